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
