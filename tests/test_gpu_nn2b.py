"""NN2B population path (tabulated network, Hermite interpolation) against the
FP64 oracle, which evaluates the reference's network per pair
(src/nn2b.cu:49-193).

* Convergence: the tabulation reconstructs p from (p, p') and p' from
  (p', p'') with cubic Hermite interpolation, so its error falls as h^4; the
  fitness error of the FP32 tabulation on the reference's nn2b case with 32,
  64 and 128 grid points must fall by at least 4x and 8x per halving of h
  (a numpy model of the same construction gives 5-14x: the p' error leaves
  its pre-asymptotic range between 32 and 64 points), and be below 1e-5 at
  the default 1024 points.
* Per-atom forces, per-configuration energies and virials of the tabulated
  path: the training set is labelled with the oracle's own outputs at one
  individual's parameters, so that individual's three fitness terms are the
  relative RMS differences between the tabulated path and the oracle, per
  atom (force) and per configuration (energy, virial)."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def _rel_err(fit, genes, ref):
    return float(np.max(np.abs(fit.compute(genes) - ref) / np.abs(ref)))


def test_tabulation_error_falls_as_h4(golden, oracle_case):
    import gpuga_amd as G
    print_case("nn2b tabulation convergence")
    case = oracle_case("nn2b")
    genes = np.random.default_rng(12).random((6, case.V), dtype=np.float32)
    ref, _ = case.fitness(genes, threads=16)
    err = {}
    for grid in (32, 64, 128, 1024):
        with G.Fitness(golden("nn2b"), nn2b_bf16=0, nn2b_grid=grid) as fit:
            err[grid] = _rel_err(fit, genes, ref)
    print("nn2b fitness rel error by grid points:", err)
    assert err[32] / err[64] > 4.0
    assert err[64] / err[128] > 8.0
    assert err[1024] < 1e-5


@pytest.mark.parametrize("bf16,tol", [(0, 2e-5), (1, 1e-4)])
def test_tabulated_forces_energies_virials_match_oracle(bf16, tol):
    import gpuga_amd as G
    from gpuga_amd import synth
    from oracle import oracle as O
    print_case("nn2b tabulated per-atom outputs", bf16)
    d = synth.wurtzite_ab(Nc=40, n=10, seed=14)
    pot = dict(potential_type=3, cutoff=d["cutoff"], weights=d["weights"], pmin=d["pmin"],
               pmax=d["pmax"], num_neurons=10)
    case = O.Case(potential=pot, arrays=d)
    genes = np.random.default_rng(15).random((3, case.V), dtype=np.float32)
    # labels: the oracle's force, energy and virial at individual 1's parameters
    d["truth"] = np.asarray(case.denormalise(genes[1]), np.float32).reshape(-1)
    assert d["truth"].shape == (case.V,)
    d = synth.label(d, case.find_force)
    with G.Fitness(arrays=d, nn2b_bf16=bf16) as fit:
        _, terms = fit.compute(genes, terms=True)
    print("nn2b tabulated vs oracle, bf16=%d: relative RMS difference of energy, virial, "
          "force = %s" % (bf16, np.array2string(terms[1], precision=3)))
    assert np.all(terms[1] < tol)
    # the other individuals are far from these labels
    assert np.all(terms[[0, 2]] > 10 * terms[1].max())


def _species_set(Nc, n, seed, ws_zero=False):
    from gpuga_amd import synth
    from oracle import oracle as O
    d = synth.wurtzite_ab(Nc=Nc, n=n, seed=seed, species=True)
    if ws_zero:
        d["truth"] = d["truth"].copy()
        d["truth"][n * (n + 4) + 1:] = 0.0
    pot = dict(potential_type=4, cutoff=d["cutoff"], weights=d["weights"], pmin=d["pmin"],
               pmax=d["pmax"], num_neurons=n)
    d = synth.label(d, O.Case(potential=pot, arrays=d).find_force)
    return d, pot


@pytest.mark.parametrize("bf16,tol", [(0, 1e-5), (1, 1e-4)])
def test_species_pair_input_matches_oracle(bf16, tol):
    """NN2B with the species-pair input (potential type 4): the first layer
    sees the pair's species code t_i + t_j; the population path tabulates the
    network once per code.  Fitness and terms of individuals of two lane
    blocks against the oracle; per-atom outputs of the per-individual path
    too."""
    import gpuga_amd as G
    from oracle import oracle as O
    print_case("nn2b species-pair input", bf16)
    n = 10
    d, pot = _species_set(40, n, 16)
    case = O.Case(potential=pot, arrays=d)
    genes = np.random.default_rng(17).random((70, case.V), dtype=np.float32)
    with G.Fitness(arrays=d, nn2b_bf16=bf16) as fit:
        assert fit.V == n * (n + 5) + 1
        f, terms = fit.compute(genes, terms=True)
        params = fit.denormalise(genes[5])
        ff, fw, fe = fit.find_force(params)
    sel = [0, 5, 63, 64, 69]
    ref, ref_terms = case.fitness(genes[sel], threads=16)
    np.testing.assert_allclose(terms[sel], ref_terms, rtol=tol)
    np.testing.assert_allclose(f[sel], ref, rtol=tol)
    of, ow, oe = case.find_force(params)
    for got, want in ((ff, of), (fw, ow), (fe, oe)):
        np.testing.assert_allclose(got, want, atol=2e-5 * np.abs(want).max())


def test_species_pair_input_with_zero_weights_is_the_reference_form():
    """Species weights of zero: type 4 gives type 3's fitness bit for bit on
    the GPU (the same network, every code's table the same)."""
    import gpuga_amd as G
    print_case("nn2b species weights zero")
    n = 10
    d4, _ = _species_set(24, n, 18, ws_zero=True)
    genes4 = np.random.default_rng(19).random((20, n * (n + 5) + 1), dtype=np.float32)
    # genes of 0.5 are weights of 0 in [-3, 3]
    genes4[:, n * (n + 4) + 1:] = 0.5
    d3 = dict(d4)
    d3.update(potential_type=3, pmin=d4["pmin"][: n * (n + 4) + 1],
              pmax=d4["pmax"][: n * (n + 4) + 1])
    out = []
    # (one fixed item size: two autotuned handles may pick different sizes,
    # which regroups the FP64 sums)
    for d, g in ((d3, genes4[:, : n * (n + 4) + 1]), (d4, genes4)):
        with G.Fitness(arrays=d, nn2b_bf16=0, chunk_cost=200) as fit:
            out.append(fit.compute(g, terms=True))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
