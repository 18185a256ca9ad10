"""GA bookkeeping: random streams against libstdc++ <random>, GA invariants,
and bit-exact agreement of the library's GA with the oracle's GA.

Pins: orc_canonical_float / orc_uniform_int / orc_mt_next against the C++
standard library (g++ build of tests/std_random_probe.cpp); the GA steps
(src/ga.cu:105-212) against invariants that a dropped term, a wrong index or
a swapped parent would break.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("probe") / "probe")
    subprocess.run(["g++", "-O2", "-std=c++14", os.path.join(HERE, "std_random_probe.cpp"), "-o",
                    exe], check=True)

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


def test_ga_invariants():
    P, parents, V = 40, 20, 9
    ga = O.GA(P, parents, V, 777, generations=50, mutation_rate=0.0)
    pop0 = ga.population.copy()
    fit = np.random.default_rng(3).random(P).astype(np.float32)
    ga.step(0, fit)  # rate 0: sort + crossover only
    pop = ga.population
    order = np.argsort(fit, kind="stable")
    # sort: rows 0..parents-1 are the best individuals in ascending fitness
    assert np.array_equal(pop[:parents], pop0[order[:parents]])
    assert np.array_equal(ga.fitness, np.sort(fit))
    # crossover: every child gene comes from one of the parent rows at the
    # same position, children come in complementary pairs
    for m in range(0, P - parents, 2):
        c1, c2 = pop[parents + m], pop[parents + m + 1]
        for n in range(V):
            assert (c1[n] in pop[:parents, n]) and (c2[n] in pop[:parents, n])
        same = [np.flatnonzero(np.all(pop[:parents] == c1, axis=1))]
        del same


def test_mutation_count_and_elite():
    P, parents, V, G = 30, 10, 5, 10
    ga = O.GA(P, parents, V, 42, generations=G, mutation_rate=0.5)
    fit = np.arange(P, dtype=np.float32)  # already sorted: identity permutation
    ga.fitness[:] = fit
    O.lib().orc_ga_sort(O.C.byref(ga.g))
    O.lib().orc_ga_crossover(O.C.byref(ga.g))
    before = ga.population.copy()
    O.lib().orc_ga_mutation(O.C.byref(ga.g), 2)
    after = ga.population
    changed = np.argwhere(before != after)
    # round(P*V*rate*(1-g/G)) draws, never in row 0 (the elite)
    expected = int(np.round(np.float32(P * V) * np.float32(0.5 * (1 - 2 / G))))
    assert len(changed) <= expected
    assert np.all(changed[:, 0] >= 1)


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
