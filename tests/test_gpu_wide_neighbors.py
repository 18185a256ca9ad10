"""Tersoff training sets with more neighbours than the atom records hold (the
reference takes any neighbour count, src/minimal_tersoff.cu:147-173 and
src/neighbor.cu:51-63).

Diamond Si with a 4.6 A cut-off has 28 neighbours per atom (shells at 2.35,
3.84 and 4.50 A).  Two population paths serve it:
  * R2 at most 3.2 A (the usual bounds): bonds at or beyond the largest R2
    any individual can have contribute exact zeros (fc = fc' = 0), so the
    atom records keep the 4 nearest neighbours;
  * R2 up to 4.6 A: 28 neighbours lie inside the taper's reach, and the
    looped kernel (per-bond lane values in LDS, geometry from the bond and
    triplet tables) evaluates them.
Both against the FP64 oracle, which builds the full 28-neighbour list as
the reference does; fitness and the three terms within 1e-5 relative, with
force configurations in the set (the per-bond force exchange of the pruned
list) and several lane blocks (P = 70)."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

# one- and two-species bounds (D0, alpha, r0, S [x2], n, beta, h, R1, R2)
ONE = (np.array([2.9, 1.3, 2.2, 2.0, 0.5, 0.0, -0.8, 2.8, 3.2], np.float32),
       np.array([3.3, 1.5, 2.4, 2.0, 0.8, 0.4, -0.6, 2.8, 3.2], np.float32))
TWO = (np.array([3.0, 1.3, 2.2, 1.8, 3.0, 1.3, 2.2, 1.8, 0.6, 0.2, -1.0, 2.7, 3.0], np.float32),
       np.array([3.2, 1.5, 2.3, 2.2, 3.2, 1.5, 2.3, 2.2, 0.9, 0.8, -0.5, 2.7, 3.0], np.float32))


def _set(two, cutoff, r1, r2, seed):
    """12 diamond cells of 64 atoms (zinc blende for two species), R1 fixed,
    R2 in the range r2, labels from the oracle at the middle of the bounds."""
    from gpuga_amd import synth
    from oracle import oracle as O
    d = synth.tersoff_si(Nc=12, reps=(2, 2, 2), seed=seed, force_fraction=0.25)
    lo, hi = (TWO if two else ONE)
    lo, hi = lo.copy(), hi.copy()
    lo[-2], hi[-2] = r1, r1
    lo[-1], hi[-1] = r2
    if two:
        # the second basis of each conventional cell (the last four sites) is
        # species 1
        d["type"] = np.tile(np.array([0, 0, 0, 0, 1, 1, 1, 1], np.int32), len(d["type"]) // 8)
    d.update(potential_type=2 if two else 1, cutoff=cutoff, pmin=lo, pmax=hi,
             truth=(0.5 * (lo + hi)).astype(np.float32))
    pot = dict(potential_type=d["potential_type"], cutoff=cutoff, weights=d["weights"],
               pmin=lo, pmax=hi)
    d = synth.label(d, O.Case(potential=pot, arrays=d).find_force)
    return d, pot


def _check(d, pot, P, seed, records, max_nn=28):
    """records: the atom-record bond slots the population kernel reads (0:
    the looped kernel, which reads no records)."""
    import gpuga_amd as G
    from oracle import oracle as O
    fit = G.Fitness(arrays=d)
    try:
        assert fit.max_neighbors == max_nn
        genes = np.random.default_rng(seed).random((P, fit.V), dtype=np.float32)
        genes[0, -1] = 1.0  # R2 at its upper bound
        f, terms = fit.compute(genes, terms=True)
        again = fit.compute(genes)
        _, _, rec = fit.geometry()
    finally:
        fit.close()
    M = records
    assert len(rec) == (fit.N * (1 + M + M * M + M // 2) if M else 0)
    assert np.array_equal(f, again)
    sel = [0, 1, 63, 64, P - 1]
    ref, ref_terms = O.Case(potential=pot, arrays=d).fitness(genes[sel], threads=16)
    np.testing.assert_allclose(terms[sel], ref_terms, rtol=1e-5)
    np.testing.assert_allclose(f[sel], ref, rtol=1e-5)


@pytest.mark.parametrize("two", [False, True])
def test_pruned_list_matches_full_list_oracle(two):
    print_case("wide cutoff, pruned to R2", two)
    d, pot = _set(two, 4.6, 2.8, (3.0, 3.2), seed=21)
    _check(d, pot, P=70, seed=5, records=4)


@pytest.mark.parametrize("two", [False, True])
def test_looped_kernel_matches_oracle(two):
    print_case("wide cutoff, looped kernel", two)
    d, pot = _set(two, 4.6, 3.0, (4.2, 4.6), seed=22)
    _check(d, pot, P=70, seed=6, records=0)


def test_pruned_list_rejects_genes_above_one():
    """Pruning holds for genes in [0, 1]: an outer-radius gene above 1 is an
    error, not a silently different result."""
    import gpuga_amd as G
    print_case("wide cutoff, gene above 1")
    d, _ = _set(False, 4.6, 2.8, (3.0, 3.2), seed=23)
    fit = G.Fitness(arrays=d)
    try:
        genes = np.full((8, fit.V), 0.5, np.float32)
        genes[3, -1] = 1.5
        with pytest.raises(G.GpugaError, match="genes must lie in"):
            fit.compute(genes)
        genes[3, -1] = 1.0
        assert np.all(np.isfinite(fit.compute(genes)))
    finally:
        fit.close()


def test_taper_reached_by_some_lanes():
    """R1 from 2.2 to 2.8 A: the wave's smallest R1 falls below many nearest
    bonds (2.35 A +- displacements), so atoms take the body with the taper
    terms, others the one without them (every bond inside every lane's R1);
    both against the oracle."""
    print_case("taper reached by some lanes")
    d, pot = _set(False, 3.0, 2.8, (3.0, 3.2), seed=24)
    d["pmin"] = d["pmin"].copy()
    d["pmin"][-2] = 2.2
    pot["pmin"] = d["pmin"]
    _check(d, pot, P=70, seed=7, records=4, max_nn=4)
