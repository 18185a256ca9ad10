"""Ragged Tersoff training sets: configurations of two, three and five atoms
in large boxes (two atoms out of each other's reach, a dimer whose bond
order has zeta = 0, bent trimers, a tetrahedral cluster; the reference's
reader rejects a configuration of one atom) beside diamond cells of 8
and 64 atoms, with the small clusters among the force configurations.  The
population path (small launches with a wave per item, one-chunk force
configurations assembled by the evaluating wave, the packed last lane
block for P = 70) against the FP64 oracle: fitness and the three terms
within 1e-5 relative, as the other Tersoff parity tests.  The reference
takes any such set (src/fitness.cu:59-161, src/minimal_tersoff.cu:122-188
with its zeta cut).  The same for the rigid-ion and NN2B (FP32 tabulation)
paths on clusters of both species beside bulk cells."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def _cluster(rng, pos, L=14.0):
    pos = np.asarray(pos, float) + rng.normal(0.0, 0.03, (len(pos), 3)) + L / 2
    return (np.diag([L, L, L]), pos, np.zeros(len(pos), np.int32))


# two-species bounds (D0, alpha, r0, S [same species, mixed], n, beta, h, R1, R2)
TWO = (np.array([3.0, 1.3, 2.2, 1.8, 3.0, 1.3, 2.2, 1.8, 0.6, 0.2, -1.0, 2.7, 3.0], np.float32),
       np.array([3.2, 1.5, 2.3, 2.2, 3.2, 1.5, 2.3, 2.2, 0.9, 0.8, -0.5, 2.7, 3.0], np.float32))


def _ragged_set(seed, two=False):
    from gpuga_amd import synth
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    t = 2.35 / np.sqrt(3.0)
    tetra = [[0, 0, 0], [t, t, t], [t, -t, -t], [-t, t, -t], [-t, -t, t]]
    cells = [
        _cluster(rng, [[0, 0, 0], [5.0, 0, 0]]),                      # no neighbours
        _cluster(rng, [[0, 0, 0], [2.35, 0, 0]]),                     # dimer: zeta = 0
        _cluster(rng, [[0, 0, 0], [2.3, 0, 0], [-0.8, 2.2, 0]]),      # bent trimer
        _cluster(rng, tetra),                                         # 4 + 4 x 1 bonds
        _cluster(rng, [[0, 0, 0], [2.4, 0, 0], [1.2, 2.1, 0]]),       # triangle
        _cluster(rng, tetra[:3]),
    ]
    # diamond cells of 8 and 64 atoms after the clusters
    for reps in ((1, 1, 1), (2, 2, 2), (1, 1, 1), (2, 2, 2)):
        frac, ty = synth._supercell(synth._DIAMOND, np.zeros(8, np.int32), reps)
        L = np.array(reps) * 5.43 * (1 + rng.uniform(-0.03, 0.03, 3))
        cells.append((np.diag(L), frac * L + rng.normal(0.0, 0.1, frac.shape), ty))
    d = synth._assemble(cells, 0.5, rng)
    if two:
        # alternate species, so every cluster has like and unlike bonds
        d["type"] = (np.arange(len(d["type"])) % 2).astype(np.int32)
        d.update(potential_type=2, cutoff=3.0, weights=(0.05, 0.15, 0.8), pmin=TWO[0],
                 pmax=TWO[1], truth=(0.5 * (TWO[0] + TWO[1])).astype(np.float32))
    else:
        d.update(potential_type=1, cutoff=3.0, weights=(0.05, 0.15, 0.8),
                 pmin=synth.TERSOFF_SI_MIN, pmax=synth.TERSOFF_SI_MAX,
                 truth=synth.TERSOFF_SI_TRUTH)
    pot = dict(potential_type=d["potential_type"], cutoff=3.0, weights=d["weights"],
               pmin=d["pmin"], pmax=d["pmax"])
    d = synth.label(d, O.Case(potential=pot, arrays=d).find_force)
    return d, pot


@pytest.mark.parametrize("two,P", [(False, 70), (False, 64), (False, 5), (True, 70), (True, 5)])
def test_ragged_clusters_match_oracle(two, P):
    import gpuga_amd as G
    from oracle import oracle as O
    print_case("ragged clusters", "two species" if two else "one species", P)
    d, pot = _ragged_set(41, two)
    assert d["Nc_force"] == 5
    nn_counts = None
    genes = np.random.default_rng(43).random((P, len(d["pmin"])), dtype=np.float32)
    with G.Fitness(arrays=d) as fit:
        NN, _, _ = fit.neighbors()
        nn_counts = NN
        f, terms = fit.compute(genes, terms=True)
        again, t2 = fit.compute(genes, terms=True)
    # the set holds what it is meant to: atoms with 0, 1, 2, 3 and 4 neighbours
    assert {0, 1, 2, 4} <= set(int(x) for x in nn_counts)
    assert np.array_equal(terms, t2) and np.array_equal(f, again)
    ref, ref_terms = O.Case(potential=pot, arrays=d).fitness(genes, threads=16)
    np.testing.assert_allclose(f, ref, rtol=1e-5)
    # a term near zero (an individual whose energies, virials or forces
    # nearly equal the labels) is an FP32 difference of nearly equal numbers:
    # 1e-7 absolute floor (two-species set: 2.3e-8 on a term of 1.6e-3)
    np.testing.assert_allclose(terms, ref_terms, rtol=1e-5, atol=1e-7)


def test_ragged_clusters_per_individual_path():
    """The per-individual launches (the plain form) on the same set."""
    import gpuga_amd as G
    from oracle import oracle as O
    print_case("ragged clusters, per-individual path")
    d, pot = _ragged_set(41)
    genes = np.random.default_rng(47).random((6, 9), dtype=np.float32)
    with G.Fitness(arrays=d, backend=1) as fit:
        f, terms = fit.compute(genes, terms=True)
    ref, ref_terms = O.Case(potential=pot, arrays=d).fitness(genes, threads=16)
    np.testing.assert_allclose(terms, ref_terms, rtol=1e-5)
    np.testing.assert_allclose(f, ref, rtol=1e-5)


def _pair_ragged_set(kind, seed):
    """Rigid ion (two species, 10 A cut-off) or NN2B (n = 8, 6 A cut-off):
    clusters of two to five atoms of both species in boxes wider than twice
    the cut-off (a pair out of reach, like and unlike pairs), then bulk
    cells, the clusters among the force configurations; labels from the
    oracle at the ground-truth parameters."""
    from gpuga_amd import synth
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    L = 24.0 if kind == "ri" else 14.0
    far = 11.0 if kind == "ri" else 6.5
    lattice = [
        ([[0, 0, 0], [far, 0, 0]], [0, 1]),                           # out of reach
        ([[0, 0, 0], [2.3, 0, 0]], [0, 1]),                           # unlike pair
        ([[0, 0, 0], [2.6, 0, 0], [0, 2.6, 0]], [1, 1, 0]),
        ([[0, 0, 0], [2.2, 0.3, 0], [-0.4, 2.4, 0.5], [0.5, -0.6, 2.5]], [0, 1, 1, 0]),
        ([[0, 0, 0], [3.1, 0, 0], [0, 3.0, 0], [0, 0, 2.9], [2.0, 2.0, 2.0]], [0, 0, 1, 1, 1]),
    ]
    cells = []
    for pos, ty in lattice:
        p = np.asarray(pos, float) + rng.normal(0.0, 0.03, (len(pos), 3)) + L / 2
        cells.append((np.diag([L, L, L]), p, np.asarray(ty, np.int32)))
    if kind == "ri":
        bulk = synth.fluorite_ab2(Nc=3, seed=seed)
        lo, hi = synth.ri_bounds()
        extra = dict(potential_type=0, cutoff=10.0, weights=(0.1, 0.2, 0.7), pmin=lo, pmax=hi,
                     truth=synth.RI_TRUTH)
    else:
        bulk = synth.wurtzite_ab(Nc=3, n=8, seed=seed)
        V = 8 * 12 + 1
        extra = dict(potential_type=3, num_neurons=8, cutoff=6.0, weights=(0.4, 0.2, 0.4),
                     pmin=np.full(V, -3, np.float32), pmax=np.full(V, 3, np.float32),
                     truth=bulk["truth"])
    box = bulk["box"].reshape(-1, 3, 3)
    r = bulk["r"].reshape(3, -1)
    off = 0
    for c, n in enumerate(bulk["Na"]):
        cells.append((box[c], r[:, off:off + n].T.astype(float), bulk["type"][off:off + n]))
        off += n
    d = synth._assemble(cells, 0.5, rng)
    d.update(extra)
    pot = {k: d[k] for k in ("potential_type", "cutoff", "weights", "pmin", "pmax")}
    if kind == "nn2b":
        pot["num_neurons"] = 8
    d = synth.label(d, O.Case(potential=pot, arrays=d).find_force)
    return d, pot


@pytest.mark.parametrize("kind,P", [("ri", 70), ("ri", 5), ("nn2b", 70), ("nn2b", 5)])
def test_ragged_pair_sets_match_oracle(kind, P):
    import gpuga_amd as G
    from oracle import oracle as O
    print_case("ragged pair set", kind, P)
    d, pot = _pair_ragged_set(kind, 53)
    orc = O.Case(potential=pot, arrays=d)
    NN = orc.neighbors()[0]
    assert 0 in set(int(x) for x in NN)
    genes = np.random.default_rng(59).random((P, len(d["pmin"])), dtype=np.float32)
    ref, ref_terms = orc.fitness(genes, threads=16)
    out = {}
    for backend in (0, 1):
        with G.Fitness(arrays=d, backend=backend, nn2b_bf16=0) as fit:
            out[backend] = fit.compute(genes, terms=True)
    for backend, (f, terms) in out.items():
        fe = np.abs(f - ref) / np.abs(ref)
        te = np.abs(terms - ref_terms)
        print(f"backend {backend}: fitness rel err max {fe.max():.3e}; term abs err max "
              f"{te.max():.3e} at {np.unravel_index(te.argmax(), te.shape)}, rel "
              f"{(te / np.abs(ref_terms)).max():.3e}")
    # Fitness within 1e-5 relative.  Terms within 1e-5, except the rigid-ion
    # energy and virial terms, 5e-5: a cluster's energy and virial are sums
    # of a few pair terms in which the shifted Coulomb part nearly cancels,
    # and both paths (population and per-individual, different code) show
    # the same 2.2e-5 against the FP64 oracle, so it is the FP32 arithmetic,
    # not the schedule.
    tol = np.array([5e-5, 5e-5, 1e-5]) if kind == "ri" else np.full(3, 1e-5)
    for backend, (f, terms) in out.items():
        np.testing.assert_allclose(f, ref, rtol=1e-5, err_msg=f"backend {backend}")
        rel = np.abs(terms - ref_terms) / np.abs(ref_terms)
        assert (rel <= tol).all(), (backend, rel.max(axis=0))
