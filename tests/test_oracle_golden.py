"""The FP64 oracle against the reference's own held outputs (golden fixtures).

* ga.out pairs: each line holds the elite's parameters (printed %g, 6
  significant digits) and its fitness.  The oracle's fitness at the printed
  parameters must agree within the change that rounding the parameters to 6
  digits can cause (first-order bound from the oracle's own gradient) plus
  the 6-digit rounding of the printed fitness and the reference's FP32
  arithmetic (2e-6 relative).
* fitness from outputs: the reference's final energy/force/virial.out,
  pushed through the oracle's fitness reduction, reproduce the last ga.out
  fitness (pins the reduction, src/fitness.cu:437-555).
* hfo2: every RI parameter is fixed (min = max), so the reference's
  energy/force/virial.out are exact known answers for the oracle's RI.
Excluded: develop/wurzite/{InAs, InP, ZnSe} (a one-species Tersoff on a
two-species set, where the reference reads an unset parameter set,
src/minimal_tersoff.cu:45-52,239-244) and GaP (outputs from another run).
"""
import os

import numpy as np
import pytest

from conftest import read_ga_out

CASES = ["si_diamond", "si_example", "gaas", "si_type2", "nn2b", "hfo2"]


def half_digit(p):
    """Half a unit in the 6th significant digit of a %g-printed value."""
    a = np.abs(p.astype(np.float64))
    out = np.zeros_like(a)
    nz = a > 0
    out[nz] = 0.5 * 10.0 ** (np.floor(np.log10(a[nz])) - 5)
    return out


def fd_bound(case, params):
    """First-order bound on |fitness(p) - fitness(p + delta)| for |delta| <=
    half a printed digit, from central differences of the oracle."""
    delta = half_digit(params)
    grad = np.zeros(len(params))
    for k in np.flatnonzero(delta > 0):
        h = max(abs(float(params[k])) * 1e-4, 1e-6)
        up = params.astype(np.float64).copy()
        dn = up.copy()
        up[k] += h
        dn[k] -= h
        fu = case.fitness_of_params(up.astype(np.float32))[0]
        fd = case.fitness_of_params(dn.astype(np.float32))[0]
        grad[k] = (fu - fd) / (float(np.float32(up[k])) - float(np.float32(dn[k])))
    return float(np.sum(np.abs(grad) * delta))


@pytest.mark.parametrize("name", ["si_diamond", "si_example", "gaas", "si_type2", "nn2b"])
def test_ga_out_pairs(golden, oracle_case, name):
    case = oracle_case(name)
    rows = read_ga_out(os.path.join(golden(name), "ga.out"))
    picks = [rows[0], rows[len(rows) // 2], rows[-1]]
    if name == "nn2b":
        picks = [rows[0], rows[-1]]
    for gen, ref, params in picks:
        fit, _ = case.fitness_of_params(params)
        bound = fd_bound(case, params) + abs(ref) * (5e-6 + 2e-6) + 1e-12
        assert abs(fit - ref) <= bound, (name, gen, fit, ref, bound)


@pytest.mark.parametrize("name", CASES)
def test_fitness_from_reference_outputs(golden, oracle_case, name):
    case = oracle_case(name)
    d = golden(name)
    N, Nc, Ncf, Nf = case.N, case.Nc, case.ts.Nc_force, case.ts.N_force
    Na = case.array("Na", Nc)
    off = case.array("Na_sum", Nc)
    fo = np.loadtxt(os.path.join(d, "force.out")).reshape(-1, 6)
    eo = np.loadtxt(os.path.join(d, "energy.out")).reshape(-1, 2)
    vo = np.loadtxt(os.path.join(d, "virial.out")).reshape(-1, 2)
    assert fo.shape[0] == Nf and eo.shape[0] == Nc - Ncf and vo.shape[0] == 6 * (Nc - Ncf)
    f = np.zeros(3 * N)
    w = np.zeros(6 * N)
    e = np.zeros(N)
    for k in range(3):
        f[k * N:k * N + Nf] = fo[:, k]
    for c in range(Ncf, Nc):
        e[off[c]] = eo[c - Ncf, 0]
        for k in range(6):
            w[k * N + off[c]] = vo[k * (Nc - Ncf) + c - Ncf, 0]
    fit, _ = case.error_terms(f, w, e)
    ref = read_ga_out(os.path.join(d, "ga.out"))[-1][1]
    tol = 3e-5 if name == "nn2b" else 2e-5  # 6 printed digits; nn2b forces are tiny
    assert abs(fit - ref) <= tol * abs(ref), (fit, ref)
    del Na


def test_rigid_ion_known_answer(golden, oracle_case):
    case = oracle_case("hfo2")
    d = golden("hfo2")
    pmin, pmax = case.bounds()
    assert np.array_equal(pmin, pmax)  # every parameter fixed in potential.in
    f, w, e = case.find_force(pmin)
    fit, _ = case.error_terms(f, w, e)
    assert abs(fit - 2.2073) <= 2.2073 * 3e-5
    N, Nc, Ncf, Nf = case.N, case.Nc, case.ts.Nc_force, case.ts.N_force
    Na = case.array("Na", Nc)
    off = case.array("Na_sum", Nc)
    eo = np.loadtxt(os.path.join(d, "energy.out"))
    fo = np.loadtxt(os.path.join(d, "force.out"))
    vo = np.loadtxt(os.path.join(d, "virial.out"))
    E = np.array([e[off[c]:off[c] + Na[c]].sum() for c in range(Ncf, Nc)])
    # the reference sums ~800 FP32 site energies; 5e-5 relative covers it
    np.testing.assert_allclose(E, eo[:, 0], rtol=5e-5)
    F = np.stack([f[k * N:k * N + Nf] for k in range(3)], 1)
    np.testing.assert_allclose(F, fo[:, :3], rtol=1e-4, atol=5e-5)
    W = np.concatenate([[w[k * N + off[c]:k * N + off[c] + Na[c]].sum() for c in range(Ncf, Nc)]
                        for k in range(6)])
    np.testing.assert_allclose(W, vo[:, 0], rtol=2e-5, atol=2e-4 * np.abs(vo[:, 0]).max())


@pytest.mark.parametrize("name", ["si_diamond", "gaas"])
def test_tersoff_elite_outputs(golden, oracle_case, name):
    """E, F, W of the final elite against the reference's *.out; the elite is
    known to 6 digits, so the allowed difference follows from the oracle's
    sensitivity to those digits (bounded generously at 50x a digit)."""
    case = oracle_case(name)
    d = golden(name)
    params = read_ga_out(os.path.join(d, "ga.out"))[-1][2]
    f, w, e = case.find_force(params)
    N, Nc, Ncf, Nf = case.N, case.Nc, case.ts.Nc_force, case.ts.N_force
    Na = case.array("Na", Nc)
    off = case.array("Na_sum", Nc)
    eo = np.loadtxt(os.path.join(d, "energy.out"))
    E = np.array([e[off[c]:off[c] + Na[c]].sum() for c in range(Ncf, Nc)])
    np.testing.assert_allclose(E, eo[:, 0], rtol=2e-4)
    fo = np.loadtxt(os.path.join(d, "force.out"))
    F = np.stack([f[k * N:k * N + Nf] for k in range(3)], 1)
    np.testing.assert_allclose(F, fo[:, :3], atol=2e-3 * np.abs(fo[:, :3]).max())


@pytest.mark.parametrize("name", ["si_diamond", "gaas", "nn2b", "hfo2"])
def test_neighbor_list_invariants_and_margin(oracle_case, name):
    """Symmetric, ascending, no self, strict cutoff; and no pair within 2e-6
    relative of rc^2.  The FP32 evaluation of d^2 (two 3-term products, three
    squares, about a dozen roundings of 2^-24 each on terms no larger than
    d^2) is accurate to ~5e-7 relative, so the list is the same whatever fused
    multiply-add contraction the reference's compiler chose."""
    case = oracle_case(name)
    NN, start, NL = case.neighbors()
    N = case.N
    r = case.array("r", 3 * N).reshape(3, N).astype(np.float64)
    h = case.array("h", 18 * case.Nc).reshape(case.Nc, 18).astype(np.float64)
    Na = case.array("Na", case.Nc)
    off = case.array("Na_sum", case.Nc)
    rc2 = float(np.float32(case.pot.cutoff) * np.float32(case.pot.cutoff))
    pairs = set()
    margin = np.inf
    for c in range(case.Nc):
        a, b = off[c], off[c] + Na[c]
        H = h[c, :9].reshape(3, 3)
        Hi = h[c, 9:].reshape(3, 3)
        for i in range(a, b):
            lst = NL[start[i]:start[i] + NN[i]]
            assert np.all(np.diff(lst) > 0) and i not in lst
            assert np.all((lst >= a) & (lst < b))
            dr = r[:, a:b] - r[:, i:i + 1]
            s = Hi @ dr
            s -= np.rint(s)
            d2 = np.sum((H @ s) ** 2, axis=0)
            d2[i - a] = np.inf
            margin = min(margin, float(np.min(np.abs(d2 - rc2)) / rc2))
            expect = np.flatnonzero(d2 < rc2) + a
            assert np.array_equal(expect, lst)
            pairs.update((i, int(j)) for j in lst)
        if len(pairs) > 200000:
            break
    assert all((j, i) in pairs for i, j in list(pairs)[:50000])
    assert margin > 2e-6, margin
