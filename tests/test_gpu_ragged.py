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
with its zeta cut)."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def _cluster(rng, pos, L=14.0):
    pos = np.asarray(pos, float) + rng.normal(0.0, 0.03, (len(pos), 3)) + L / 2
    return (np.diag([L, L, L]), pos, np.zeros(len(pos), np.int32))


def _ragged_set(seed):
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
    d.update(potential_type=1, cutoff=3.0, weights=(0.05, 0.15, 0.8),
             pmin=synth.TERSOFF_SI_MIN, pmax=synth.TERSOFF_SI_MAX,
             truth=synth.TERSOFF_SI_TRUTH)
    pot = dict(potential_type=1, cutoff=3.0, weights=d["weights"], pmin=d["pmin"],
               pmax=d["pmax"])
    d = synth.label(d, O.Case(potential=pot, arrays=d).find_force)
    return d, pot


@pytest.mark.parametrize("P", [70, 64, 5])
def test_ragged_clusters_match_oracle(P):
    import gpuga_amd as G
    from oracle import oracle as O
    print_case("ragged clusters", P)
    d, pot = _ragged_set(41)
    assert d["Nc_force"] == 5
    nn_counts = None
    genes = np.random.default_rng(43).random((P, 9), dtype=np.float32)
    with G.Fitness(arrays=d) as fit:
        NN, _, _ = fit.neighbors()
        nn_counts = NN
        f, terms = fit.compute(genes, terms=True)
        again, t2 = fit.compute(genes, terms=True)
    # the set holds what it is meant to: atoms with 0, 1, 2, 3 and 4 neighbours
    assert {0, 1, 2, 4} <= set(int(x) for x in nn_counts)
    assert np.array_equal(terms, t2) and np.array_equal(f, again)
    ref, ref_terms = O.Case(potential=pot, arrays=d).fitness(genes, threads=16)
    np.testing.assert_allclose(terms, ref_terms, rtol=1e-5)
    np.testing.assert_allclose(f, ref, rtol=1e-5)


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
