"""GA bookkeeping: random streams against libstdc++ <random>, the oracle's GA
steps against the reference's expressions and distributions, and bit-exact
agreement of the library's GA with the oracle's GA.

Pins:
  orc_mt_* / orc_canonical_float / orc_uniform_int   libstdc++ (g++ build of
      tests/std_random_probe.cpp), the standard's 10000th-output answer
  orc_ga_init     initial genes = the stream's first P V floats; the rank
                  probabilities' closed form
  orc_ga_sort     stable ascending order, rows follow (ties included)
  get_a_parent    strict cum[n] > u at an exact boundary draw; the joint
                  distribution of the two parents (chi-square, 1.2e5 pairs),
                  which also separates swapped children
  orc_ga_crossover complementary children, point in [1, V - 1] (uniform, all
                  hit), draw order parent, parent (redrawn), point
  orc_ga_mutation count, index range, index-before-value order and the
                  stream position after it, against the reference's
                  expression compiled by g++ -std=c++14 (and shown to differ
                  under -std=c++17, where the value is drawn first)
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def probe_exes(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    exes = {}
    for std in (14, 17):
        exes[std] = str(d / f"probe{std}")
        subprocess.run(["g++", "-O2", f"-std=c++{std}", os.path.join(HERE, "std_random_probe.cpp"),
                        "-o", exes[std]], check=True)
    return exes


@pytest.fixture(scope="module")
def probe_raw(probe_exes):
    def run(std, seed, schedule):
        return subprocess.run([probe_exes[std], str(seed), schedule], check=True,
                              capture_output=True, text=True).stdout.split("\n")
    return run


@pytest.fixture(scope="module")
def probe(probe_exes):
    exe = probe_exes[14]

    def run(seed, schedule):
        out = subprocess.run([exe, str(seed), schedule], check=True, capture_output=True,
                             text=True).stdout.split("\n")
        vals = []
        for line in out:
            if line.startswith("u "):
                vals.append(struct.unpack("<f", bytes.fromhex(line[2:])[::-1])[0])
            elif line.startswith("i "):
                vals.append(int(line[2:]))
        return vals
    return run


def test_canonical_float_matches_libstdcxx(probe):
    ref = probe(12345678, "f" * 3000)
    g = O.MT19937(12345678)
    got = [g.canonical() for _ in range(3000)]
    assert np.array_equal(np.array(got, np.float32), np.array(ref, np.float32))


@pytest.mark.parametrize("a,b", [(1, 8), (1, 12), (9, 1799), (13, 2599), (141, 28199),
                                 (0, 1), (5, 5)])
def test_uniform_int_matches_libstdcxx(probe, a, b):
    sched = ",".join([f"i{a}:{b}"] * 2000) + ","
    ref = probe(987654321, sched)
    g = O.MT19937(987654321)
    got = [g.uniform_int(a, b) for _ in range(2000)]
    assert got == ref


def test_interleaved_stream_matches_libstdcxx(probe):
    # the draw pattern of one crossover + one mutation: f f [f...] i, then i f pairs
    rng = np.random.default_rng(1)
    sched, ops = [], []
    for _ in range(400):
        if rng.random() < 0.5:
            sched.append("f")
            ops.append(("f",))
        else:
            a = int(rng.integers(0, 50))
            b = a + int(rng.integers(0, 5000))
            sched.append(f"i{a}:{b},")
            ops.append(("i", a, b))
    ref = probe(5489, "".join(sched))
    g = O.MT19937(5489)
    got = [g.canonical() if op[0] == "f" else g.uniform_int(op[1], op[2]) for op in ops]
    for x, y, op in zip(got, ref, ops):
        if op[0] == "f":
            assert np.float32(x) == np.float32(y)
        else:
            assert x == y


def test_mt19937_known_answer():
    # C++ standard [rand.predef]: the 10000th output of a default-seeded
    # std::mt19937 is 4123659995.
    g = O.MT19937(5489)
    for _ in range(9999):
        g.next()
    assert g.next() == 4123659995


def _untemper(y):
    """Inverse of std::mt19937's output tempering: the state word whose next
    output is y."""
    def undo_right(v, k):
        x = v
        for _ in range(32 // k + 1):
            x = v ^ (x >> k)
        return x & 0xFFFFFFFF

    def undo_left(v, k, mask):
        x = v
        for _ in range(32 // k + 1):
            x = v ^ ((x << k) & mask)
        return x & 0xFFFFFFFF
    y = undo_right(y, 18)
    y = undo_left(y, 15, 0xEFC60000)
    y = undo_left(y, 7, 0x9D2C5680)
    return undo_right(y, 11)


def _clone_rng(ga):
    m = O.MT19937(0)
    m.s = O._MT.from_buffer_copy(ga.g.rng)
    return m


def _crossover_samples(parents, pairs, V, calls, seed):
    """Run the oracle's crossover on parent rows that identify themselves
    (row n holds n + 0.5 in every gene) and read back, per child pair, the
    two parents and the crossover point."""
    P = parents + 2 * pairs
    ga = O.GA(P, parents, V, seed, generations=10, mutation_rate=0.0)
    ident = (np.arange(parents, dtype=np.float32) + 0.5)[:, None] * np.ones(V, np.float32)
    p1s, p2s, cps = [], [], []
    for _ in range(calls):
        ga.population[:parents] = ident
        O.lib().orc_ga_crossover(O.C.byref(ga.g))
        pop = ga.population
        assert np.array_equal(pop[:parents], ident)  # parent rows untouched
        c1, c2 = pop[parents::2], pop[parents + 1::2]
        a, b = c1[:, 0], c1[:, -1]
        # both parents in every child: a crossover point in [1, V - 1] and two
        # distinct parents
        assert np.all(a != b)
        cp = np.sum(c1 == a[:, None], axis=1)
        n = np.arange(V)[None, :]
        want1 = np.where(n < cp[:, None], a[:, None], b[:, None])
        want2 = np.where(n < cp[:, None], b[:, None], a[:, None])
        # child 1 = parent 1 before the point, parent 2 after; child 2 the
        # complement (src/ga.cu:173-184)
        assert np.array_equal(c1, want1) and np.array_equal(c2, want2)
        p1s.append((a - 0.5).astype(int))
        p2s.append((b - 0.5).astype(int))
        cps.append(cp)
    return np.concatenate(p1s), np.concatenate(p2s), np.concatenate(cps)


def test_cumulative_rank_probabilities_closed_form():
    # src/ga.cu:41-46: cum[n] = sum_{k<=n} (parents - k) / (parents (parents + 1) / 2),
    # numerator and denominator exact integers in float
    for parents in (2, 10, 100, 128, 2048):
        ga = O.GA(parents + 2, parents, 3, 1)
        cum = np.ctypeslib.as_array(ga.g.cumulative_probabilities, shape=(parents,))
        S = np.cumsum(np.arange(parents, 0, -1)).astype(np.float32)
        D = np.float32(parents * (parents + 1) // 2)
        assert np.array_equal(cum, S / D)
        assert cum[-1] == 1.0


def test_parent_draw_boundary_is_strict():
    # get_a_parent (src/ga.cu:200-212) takes the first n with cum[n] > u: a
    # draw u equal to cum[n] selects n + 1.  The next output of the stream is
    # set so that u = float(x) / 2^32 hits cum[n] exactly (and, for the
    # other case, lands one float below it).
    parents, V = 4, 3
    ga = O.GA(parents + 2, parents, V, 1, generations=10, mutation_rate=0.0)
    cum = np.ctypeslib.as_array(ga.g.cumulative_probabilities, shape=(parents,)).copy()
    ident = (np.arange(parents, dtype=np.float32) + 0.5)[:, None] * np.ones(V, np.float32)
    for n in range(parents - 1):
        m, e = np.frexp(np.float64(cum[n]))
        for below, expect in ((False, n + 1), (True, n)):
            u = np.nextafter(cum[n], np.float32(0)) if below else cum[n]
            x = int(np.float64(u) * 2.0 ** 32)
            assert np.float32(np.float32(x) / np.float32(2.0 ** 32)) == u
            ga.population[:parents] = ident
            ga.g.rng.idx = 0
            ga.g.rng.mt[0] = _untemper(x)
            O.lib().orc_ga_crossover(O.C.byref(ga.g))
            assert int(ga.population[parents, 0] - 0.5) == expect, (n, below)


def test_parent_draws_follow_rank_probabilities():
    # 1.2e5 child pairs: the first parent ~ w_n = (parents - n) / (parents
    # (parents + 1) / 2); the second, redrawn while equal to the first,
    # ~ w_m / (1 - w_n).  A shifted index, a reversed rank or swapped
    # children (first/second exchanged: w_m w_n / (1 - w_m)) fail this.
    from scipy import stats
    parents, V = 10, 9
    p1, p2, cp = _crossover_samples(parents, 600, V, 200, 2024)
    N = len(p1)
    w = np.arange(parents, 0, -1) / (parents * (parents + 1) / 2)
    joint = np.zeros((parents, parents))
    np.add.at(joint, (p1, p2), 1)
    expect = N * w[:, None] * w[None, :] / (1 - w[:, None])
    np.fill_diagonal(expect, 0)
    assert np.all(np.diag(joint) == 0)
    mask = expect > 0
    chi = np.sum((joint[mask] - expect[mask]) ** 2 / expect[mask])
    assert stats.chi2.sf(chi, mask.sum() - 1) > 1e-6, chi
    swapped = expect.T * (expect.sum() / expect.T.sum())
    chi_sw = np.sum((joint[mask] - swapped[mask]) ** 2 / swapped[mask])
    assert stats.chi2.sf(chi_sw, mask.sum() - 1) < 1e-12  # the test can tell
    # crossover point uniform on [1, V - 1], every value drawn
    counts = np.bincount(cp, minlength=V + 1)
    assert counts[0] == 0 and counts[V] == 0 and np.all(counts[1:V] > 0)
    chi_cp = np.sum((counts[1:V] - N / (V - 1)) ** 2 / (N / (V - 1)))
    assert stats.chi2.sf(chi_cp, V - 2) > 1e-6


def test_crossover_draw_order():
    # per child pair: parent 1, parent 2 (redrawn while equal), then the
    # crossover point (src/ga.cu:164-170), from the stream in that order
    parents, pairs, V = 6, 50, 7
    ga = O.GA(parents + 2 * pairs, parents, V, 99, generations=10, mutation_rate=0.0)
    cum = np.ctypeslib.as_array(ga.g.cumulative_probabilities, shape=(parents,)).copy()
    rng = _clone_rng(ga)
    ident = (np.arange(parents, dtype=np.float32) + 0.5)[:, None] * np.ones(V, np.float32)
    ga.population[:parents] = ident
    O.lib().orc_ga_crossover(O.C.byref(ga.g))
    pop = ga.population

    def parent():
        u = np.float32(rng.canonical())
        return int(np.argmax(cum > u))
    for m in range(pairs):
        a = parent()
        b = parent()
        while b == a:
            b = parent()
        cp = rng.uniform_int(1, V - 1)
        c1 = pop[parents + 2 * m]
        assert int(c1[0] - 0.5) == a and int(c1[-1] - 0.5) == b
        assert np.sum(c1 == c1[0]) == cp


def test_ga_initial_population_and_sort():
    P, parents, V = 40, 20, 9
    ga = O.GA(P, parents, V, 777, generations=50, mutation_rate=0.0)
    # initial genes: P * V uniform floats in row-major order (src/ga.cu:53-56)
    g = O.MT19937(777)
    assert np.array_equal(ga.population.ravel(),
                          np.array([g.canonical() for _ in range(P * V)], np.float32))
    pop0 = ga.population.copy()
    fit = np.random.default_rng(3).random(P).astype(np.float32)
    fit[[5, 9, 30]] = fit[2]  # ties keep their order (insertion sort is stable)
    ga.step(0, fit)  # rate 0: sort + crossover only
    order = np.argsort(fit, kind="stable")
    assert np.array_equal(ga.population[:parents], pop0[order[:parents]])
    assert np.array_equal(ga.fitness, np.sort(fit))


@pytest.mark.parametrize("P,parents,V,G,gen,rate,seed", [
    (30, 10, 5, 10, 2, 0.5, 42), (200, 100, 9, 1000, 0, 0.2, 12345678),
    (64, 32, 141, 100, 37, 0.2, 9), (256, 128, 13, 50, 49, 0.3, 1)])
def test_mutation_matches_reference_expression(probe_raw, P, parents, V, G, gen, rate, seed):
    # The oracle's mutation (count, index range [V, P V - 1], index drawn
    # before value, nothing else drawn) against libstdc++ running the
    # reference's expression `population[r1(rng)] = r2(rng)` as g++
    # -std=c++14 compiles it, from the same stream position (after the P V
    # initial draws).  Every cell of rows >= 1 holds a sentinel first, so
    # each mutation shows.
    ga = O.GA(P, parents, V, seed, generations=G, mutation_rate=rate)
    ga.population[1:] = 2.0
    ga.population[0] = 3.0
    O.lib().orc_ga_mutation(O.C.byref(ga.g), gen)
    m = P * V
    count = probe_raw(14, seed, f"c{m}:{G}:{gen}:{rate!r},")
    n = int(count[0].split()[1])
    assert n == int(np.round(np.float32(m) * np.float32(rate * (1 - np.float32(gen) / np.float32(G)))))
    out = probe_raw(14, seed, f"s{m}," + f"m{V}:{m - 1}," * n + "fff")
    writes = [ln.split() for ln in out if ln.startswith("m ")]
    assert len(writes) == n
    want = np.full(m, 2.0, np.float32)
    want[:V] = 3.0
    for _, k, bits in writes:
        want[int(k)] = struct.unpack(">f", bytes.fromhex(bits))[0]
    assert np.array_equal(ga.population.ravel(), want)
    # the stream continues where the reference's would
    nxt = [struct.unpack(">f", bytes.fromhex(ln[2:]))[0] for ln in out if ln.startswith("u ")]
    rng = _clone_rng(ga)
    assert [np.float32(rng.canonical()) for _ in range(3)] == [np.float32(x) for x in nxt]
    if n > 0:
        # the C++17 order (value drawn first) gives another population
        out17 = probe_raw(17, seed, f"s{m}," + f"m{V}:{m - 1}," * n)
        w17 = [ln.split() for ln in out17 if ln.startswith("m ")]
        assert w17 != writes


def test_product_ga_bit_exact_with_oracle():
    G = pytest.importorskip("gpuga_amd")
    for P, parents, V, seed in [(200, 100, 9, 12345678), (256, 128, 13, 1), (64, 32, 141, 9),
                                (50, 26, 15, 3)]:
        ga = G.GA(P, parents, V, maximum_generation=30, mutation_rate=0.2, seed=seed)
        og = O.GA(P, parents, V, seed, generations=30, mutation_rate=0.2)
        assert np.array_equal(ga.population, og.population)
        rng = np.random.default_rng(seed)
        for g in range(30):
            f = rng.random(P).astype(np.float32)
            if g % 7 == 3:
                f[rng.integers(0, P, 5)] = f[0]  # ties
            ga.step(g, f)
            og.step(g, f)
            assert np.array_equal(ga.population, og.population), (P, V, g)
