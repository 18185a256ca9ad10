"""The FP64 oracle against what the mathematics fixes.

* forces are -dE/dr (central differences of the oracle's own energy);
* for an isolated cluster the configuration virial equals sum_i r_i (x) F_i
  (component pairs xx yy zz xy yz zx as the reference defines them,
  src/minimal_tersoff.cu:386-391, src/ri.cu:167-172);
* forces sum to zero per configuration;
* closed forms: a Tersoff dimer (zeta = 0, b = 1) and equilateral trimer,
  the two parameter sets of the two-species Tersoff, rigid-ion pairs of all
  three type combinations (Buckingham + damped shifted-force Coulomb of
  Fennell & Gezelter, J. Chem. Phys. 124, 234104 (2006), + generalised Morse),
  and an NN2B dimer evaluated with numpy.  The closed-form systems sit in a
  32 A box, whose FP32 inverse (src/fitness.cu:40-57) is exact, so the
  minimum image (src/mic.cuh:23-31) returns the separation unchanged.
A dropped term, a wrong sign or index, or a swapped operand in any of the
potentials fails at least one of these.
"""
import math

import numpy as np
import pytest
from scipy.special import erfc

from oracle import oracle as O

TERSOFF_P = np.array([3.0, 1.4, 2.3, 2.0, 0.7, 0.3, -0.7, 2.8, 3.2], np.float32)
TERSOFF2_P = np.array([3.0, 1.4, 2.3, 2.0, 2.6, 1.3, 2.4, 1.9, 0.7, 0.3, -0.7, 2.8, 3.2],
                      np.float32)
RI_P = np.array([2.4, 900.0, 0.3, 5.0, 12372.16, 0.2286, 81.36, 1388.77, 0.3623, 175.0, 10.0,
                 0.3474, 1.623, 2.048, 2.0], np.float32)


def make_case(ptype, positions, types, box=32.0, cutoff=3.2, n=0, cells=None):
    """One or more configurations; energy configs only unless force configs
    are needed (the first config is always a force config)."""
    if cells is None:
        cells = [(np.asarray(positions, np.float64), np.asarray(types, np.int32),
                  np.eye(3) * box)]
    Na = np.array([len(c[1]) for c in cells], np.int32)
    r = np.concatenate([c[0] for c in cells]).T.astype(np.float32)
    typ = np.concatenate([c[1] for c in cells]).astype(np.int32)
    boxes = np.concatenate([np.asarray(c[2], np.float32).reshape(9) for c in cells])
    if len(cells) == 1:  # a second, distant copy so that Nc >= 2
        Na = np.array([Na[0], Na[0]], np.int32)
        r = np.concatenate([r, r], 1)
        typ = np.concatenate([typ, typ])
        boxes = np.concatenate([boxes, boxes])
    V = {0: 15, 1: 9, 2: 13, 3: n * (n + 4) + 1, 4: n * (n + 5) + 1}[ptype]
    pot = dict(potential_type=ptype, cutoff=cutoff, weights=(1.0, 1.0, 1.0),
               pmin=np.zeros(V, np.float32), pmax=np.ones(V, np.float32), num_neurons=n)
    arrays = dict(Na=Na, type=typ, box=boxes, r=r.reshape(-1), Nc_force=1,
                  pe_ref=np.ones(len(Na), np.float32),
                  virial_ref=np.ones(6 * len(Na), np.float32),
                  force_ref=np.ones(r.size, np.float32))
    return O.Case(potential=pot, arrays=arrays)


def energy(case, params, c=0):
    f, w, e = case.find_force(params)
    Na = case.array("Na", case.Nc)
    off = case.array("Na_sum", case.Nc)
    return e[off[c]:off[c] + Na[c]].sum()


def perturbed_positions(case, i, k, h):
    r = np.ctypeslib.as_array(case.ts.r, shape=(3 * case.N,))
    idx = k * case.N + i
    old = r[idx]
    return r, idx, old


def fd_forces(case, params, atoms, h=1e-3):
    """-dE/dr_i by central differences on the float positions (first config)."""
    r = np.ctypeslib.as_array(case.ts.r, shape=(3 * case.N,))
    out = np.zeros((len(atoms), 3))
    for a, i in enumerate(atoms):
        for k in range(3):
            idx = k * case.N + i
            old = r[idx]
            r[idx] = np.float32(old + h)
            up, xu = energy(case, params), float(r[idx])
            r[idx] = np.float32(old - h)
            dn, xd = energy(case, params), float(r[idx])
            r[idx] = old
            out[a, k] = -(up - dn) / (xu - xd)
    return out


def diamond_cluster(seed, n=10, a=5.43, disp=0.08):
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0], [.25, .25, .25],
                     [.25, .75, .75], [.75, .25, .75], [.75, .75, .25], [1, .5, .5],
                     [.5, 1, .5], [.5, .5, 1], [1, 1, 1]])[:n] * a
    return base + 12.0 + rng.normal(0, disp, base.shape)


def periodic_diamond(seed, a=5.43, disp=0.1):
    rng = np.random.default_rng(seed)
    f = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0], [.25, .25, .25],
                  [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    frac = np.concatenate([f + [i, j, k] for i in range(2) for j in range(2) for k in range(2)]) / 2
    L = 2 * a * (1 + rng.uniform(-0.03, 0.03, 3))
    return frac * L + rng.normal(0, disp, frac.shape), np.diag(L)


POTENTIALS = {
    "tersoff1": (1, TERSOFF_P, 3.2, 0),
    "tersoff2": (2, TERSOFF2_P, 3.2, 0),
    "ri": (0, RI_P, 12.0, 0),
}


@pytest.mark.parametrize("name", ["tersoff1", "tersoff2", "ri", "nn2b"])
def test_forces_are_energy_gradient(name):
    if name == "nn2b":
        n = 6
        rng = np.random.default_rng(2)
        params = rng.uniform(-1, 1, n * (n + 4) + 1).astype(np.float32)
        params[:n] = rng.uniform(0.5, 1.5, n)
        # below the 7-8 A taper, where the reference's NN2B force is the gradient
        ptype, cutoff = 3, 6.5
    else:
        ptype, params, cutoff, n = POTENTIALS[name]
    pos, H = periodic_diamond(4)
    types = np.arange(len(pos)) % 2
    case = make_case(ptype, None, None, cutoff=cutoff, n=n,
                     cells=[(pos, types, H), (pos * 1.01, types, H * 1.01)])
    f, w, e = case.find_force(params)
    atoms = [0, 5, 17, 40]
    fd = fd_forces(case, params, atoms)
    N = case.N
    got = np.array([[f[k * N + i] for k in range(3)] for i in atoms])
    scale = np.abs(f[:3 * N]).max()
    np.testing.assert_allclose(got, fd, atol=2e-4 * scale + 1e-6)


@pytest.mark.parametrize("name", ["tersoff1", "tersoff2", "ri"])
def test_cluster_virial_and_newton(name):
    ptype, params, cutoff, n = POTENTIALS[name]
    pos = diamond_cluster(1)
    types = np.arange(len(pos)) % 2
    case = make_case(ptype, pos, types, box=40.0, cutoff=cutoff)
    f, w, e = case.find_force(params)
    N, na = case.N, len(pos)
    F = np.stack([f[k * N:k * N + na] for k in range(3)], 1)
    R = np.asarray(pos)
    Rf = R.astype(np.float32).astype(np.float64)
    W = np.array([w[k * N:k * N + na].sum() for k in range(6)])
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
    expect = np.array([np.sum(Rf[:, a] * F[:, b]) for a, b in pairs])
    np.testing.assert_allclose(W, expect, atol=1e-9 * np.abs(expect).max() + 1e-9)
    np.testing.assert_allclose(F.sum(0), 0.0, atol=1e-10 * np.abs(F).max())


def fdist(d, x0=10.0):
    """The separation the oracle sees: both coordinates are stored as float."""
    return float(np.float32(x0 + d)) - float(np.float32(x0))


def tersoff_pair_closed(p, d):
    D0, al, r0, S = map(float, p[:4])
    R1, R2 = float(p[-2]), float(p[-1])
    fR = D0 / (S - 1) * math.exp(-math.sqrt(2 * S) * al * (d - r0))
    fA = S * D0 / (S - 1) * math.exp(-math.sqrt(2 / S) * al * (d - r0))
    if d < R1:
        fc = 1.0
    elif d < R2:
        fc = 0.5 * math.cos(math.pi * (d - R1) / (R2 - R1)) + 0.5
    else:
        fc = 0.0
    return fc, fR, fA


@pytest.mark.parametrize("d", [2.35, 2.95])
def test_tersoff_dimer_closed_form(d):
    case = make_case(1, [[10, 10, 10], [10 + d, 10, 10]], [0, 0])
    d = fdist(d)
    fc, fR, fA = tersoff_pair_closed(TERSOFF_P, d)
    assert energy(case, TERSOFF_P) == pytest.approx(fc * (fR - fA), rel=1e-10, abs=1e-12)


def test_tersoff_trimer_closed_form():
    d = 2.4
    pos = [[10, 10, 10], [10 + d, 10, 10], [10 + d / 2, 10 + d * math.sqrt(3) / 2, 10]]
    case = make_case(1, pos, [0, 0, 0])
    fc, fR, fA = tersoff_pair_closed(TERSOFF_P, d)
    n, beta, h = map(float, TERSOFF_P[4:7])
    zeta = fc * beta * (0.5 - h) ** 2  # the other bond at 60 degrees
    b = (1 + zeta ** n) ** (-1 / (2 * n))
    # coordinates rounded to float make the triangle equilateral to ~1e-7
    assert energy(case, TERSOFF_P) == pytest.approx(3 * fc * (fR - b * fA), rel=2e-6)


def test_tersoff_two_species_parameter_sets():
    d = 2.5
    like = make_case(2, [[10, 10, 10], [10 + d, 10, 10]], [1, 1])
    unlike = make_case(2, [[10, 10, 10], [10 + d, 10, 10]], [0, 1])
    d = fdist(d)
    p = TERSOFF2_P
    set0 = np.concatenate([p[0:4], p[8:13]])
    set1 = np.concatenate([p[4:8], p[8:13]])
    fc, fR, fA = tersoff_pair_closed(set0, d)
    assert energy(like, p) == pytest.approx(fc * (fR - fA), rel=1e-10)
    fc, fR, fA = tersoff_pair_closed(set1, d)
    assert energy(unlike, p) == pytest.approx(fc * (fR - fA), rel=1e-10)


def ri_pair_closed(p, t, d):
    """Buckingham + DSF Coulomb (alpha = 0.2 1/A, cutoff p[10]) + generalised
    Morse on unlike pairs; charges q0 and -q0/2."""
    q0 = float(p[0])
    q = [q0, -q0 / 2]
    ti, tj = [(0, 0), (0, 1), (1, 1)][t]
    qq = 14.399645 * q[ti] * q[tj]
    A, rho, C = map(float, p[1 + 3 * t:4 + 3 * t])
    rc, al = float(p[10]), 0.2
    dsf = (erfc(al * d) / d - erfc(al * rc) / rc
           + (erfc(al * rc) / rc ** 2 + 2 * al / math.sqrt(math.pi) * math.exp(-al * al * rc * rc) / rc)
           * (d - rc))
    U = A * math.exp(-d / rho) - C / d ** 6 + qq * dsf
    if t == 1:
        D0, a, r0, S = map(float, p[11:15])
        U += D0 / (S - 1) * math.exp(-math.sqrt(2 * S) * a * (d - r0))
        U -= S * D0 / (S - 1) * math.exp(-math.sqrt(2 / S) * a * (d - r0))
    return U


@pytest.mark.parametrize("types,t", [([0, 0], 0), ([0, 1], 1), ([1, 1], 2)])
def test_rigid_ion_pair_closed_form(types, t):
    d = 2.2
    case = make_case(0, [[10, 10, 10], [10 + d, 10, 10]], types, cutoff=12.0)
    d = fdist(d)
    assert energy(case, RI_P) == pytest.approx(ri_pair_closed(RI_P, t, d), rel=1e-10)


@pytest.mark.parametrize("d", [2.5, 7.4])
def test_nn2b_dimer_closed_form(d):
    n = 5
    rng = np.random.default_rng(3)
    w = rng.uniform(-1, 1, n * (n + 4) + 1).astype(np.float32)
    case = make_case(3, [[10, 10, 10], [10 + d, 10, 10]], [0, 0], cutoff=8.5, n=n)
    d = fdist(d)
    w0, b0 = w[:n].astype(float), w[n:2 * n].astype(float)
    W1 = w[2 * n:2 * n + n * n].astype(float).reshape(n, n)
    b1, w2, b2 = w[n * (n + 2):n * (n + 3)], w[n * (n + 3):n * (n + 4)], float(w[-1])
    x1 = np.tanh(w0 * d - b0)
    x2 = np.tanh(W1 @ x1 - b1.astype(float))
    p = float(w2.astype(float) @ x2) - b2
    fc = 1.0 if d < 7 else 0.5 * math.cos(float(np.float32(3.1415927)) * (d - 7)) + 0.5
    assert energy(case, w) == pytest.approx(p * fc, rel=1e-10, abs=1e-12)


def test_nn2b_taper_force_quirk():
    """Inside the 7-8 A taper the reference's NN2B force is
    (p' fc + (p fc) fc') / r times r_ij (src/nn2b.cu:145-148), not the
    gradient of p fc; the oracle keeps that definition."""
    n, d = 4, 7.3
    rng = np.random.default_rng(8)
    w = rng.uniform(-1, 1, n * (n + 4) + 1).astype(np.float32)
    case = make_case(3, [[10, 10, 10], [10 + d, 10, 10]], [0, 0], cutoff=8.5, n=n)
    d = fdist(d)
    f, _, _ = case.find_force(w)
    w0, b0 = w[:n].astype(float), w[n:2 * n].astype(float)
    W1 = w[2 * n:2 * n + n * n].astype(float).reshape(n, n)
    b1, w2 = w[n * (n + 2):n * (n + 3)].astype(float), w[n * (n + 3):n * (n + 4)].astype(float)
    x1 = np.tanh(w0 * d - b0)
    z = W1 @ x1 - b1
    x2 = np.tanh(z)
    p = float(w2 @ x2) - float(w[-1])
    dp = float(w2 @ ((1 - x2 ** 2) * (W1 @ ((1 - x1 ** 2) * w0))))
    pf = float(np.float32(3.1415927))
    fc = 0.5 * math.cos(pf * (d - 7)) + 0.5
    fcp = -0.5 * pf * math.sin(pf * (d - 7))
    f2 = (dp * fc + p * fc * fcp) / d
    N = case.N
    # atom 0 sees its neighbour at +d along x: F_0 = r_01 f2 = (d, 0, 0) f2
    assert f[0] == pytest.approx(d * f2, rel=1e-10)
    assert abs(d * f2 - d * (dp * fc + p * fcp) / d) > 1e-6  # differs from the gradient
    del N


# NN2B with the species-pair input (potential type 4, no reference
# counterpart): x1 = tanh(w0 r + ws s - b0), s = t_i + t_j, ws after the
# reference's n (n + 4) + 1 parameters.

def _nn2b_species_params(n, seed):
    rng = np.random.default_rng(seed)
    w = rng.uniform(-1, 1, n * (n + 5) + 1).astype(np.float32)
    w[:n] = rng.uniform(0.5, 1.5, n)
    return w


def test_nn2b_species_with_zero_weights_is_the_reference_form():
    """ws = 0: type 4 gives type 3's forces, virials and energies bit for bit
    on a binary wurtzite set."""
    from gpuga_amd import synth
    n = 6
    d = synth.wurtzite_ab(Nc=3, n=n, seed=9)
    w3 = _nn2b_species_params(n, 4)[: n * (n + 4) + 1]
    w4 = np.concatenate([w3, np.zeros(n, np.float32)])
    out = []
    for ptype, w in ((3, w3), (4, w4)):
        pot = dict(potential_type=ptype, cutoff=d["cutoff"], weights=d["weights"],
                   pmin=np.zeros(len(w), np.float32), pmax=np.ones(len(w), np.float32),
                   num_neurons=n)
        out.append(O.Case(potential=pot, arrays=d).find_force(w))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("types,code", [([0, 0], 0), ([0, 1], 1), ([1, 1], 2)])
def test_nn2b_species_dimer_closed_form(types, code):
    n, d = 5, 2.5
    w = _nn2b_species_params(n, 6)
    case = make_case(4, [[10, 10, 10], [10 + d, 10, 10]], types, cutoff=8.5, n=n)
    d = fdist(d)
    w0, b0 = w[:n].astype(float), w[n:2 * n].astype(float)
    W1 = w[2 * n:2 * n + n * n].astype(float).reshape(n, n)
    b1, w2 = w[n * (n + 2):n * (n + 3)].astype(float), w[n * (n + 3):n * (n + 4)].astype(float)
    ws = w[n * (n + 4) + 1:].astype(float)
    x1 = np.tanh(w0 * d + ws * code - b0)
    p = float(w2 @ np.tanh(W1 @ x1 - b1)) - float(w[n * (n + 4)])
    assert energy(case, w) == pytest.approx(p, rel=1e-10, abs=1e-12)


def test_nn2b_species_forces_are_energy_gradient():
    n = 6
    w = _nn2b_species_params(n, 7)
    pos, H = periodic_diamond(4)
    types = np.arange(len(pos)) % 2
    case = make_case(4, None, None, cutoff=6.5, n=n,
                     cells=[(pos, types, H), (pos * 1.01, types, H * 1.01)])
    f, _, _ = case.find_force(w)
    atoms = [0, 5, 17, 40]
    fd = fd_forces(case, w, atoms)
    N = case.N
    got = np.array([[f[k * N + i] for k in range(3)] for i in atoms])
    np.testing.assert_allclose(got, fd, atol=2e-4 * np.abs(f[:3 * N]).max() + 1e-6)
