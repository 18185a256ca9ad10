"""Seeded synthetic training sets for the benchmark workloads.

Each generator returns a dict of arrays in train.in order (the keys accepted
by gpuga_amd.Fitness(arrays=...)) with labels left at zero; `label()` fills
them from a ground-truth parameter set through a labelling callback
(find_force of the GPU library in the benchmark, the FP64 oracle in tests).
`write_dir()` writes potential.in / train.in / ga.in in the reference's
formats (doc/GPUGA.tex:116-193 of the reference), with float round-trip
precision.

Workloads (sizes as named in the benchmark configurations):
  tersoff_si(...)   diamond Si supercells, a = 5.43 A, +-3% axis strain,
                    N(0, 0.1 A) displacements; minimal Tersoff, 9 parameters
  fluorite_ab2(...) AB2 fluorite 2x2x2 cells (96 atoms), a = 5.1 A, +-2%
                    strain, N(0, 0.05 A); rigid-ion potential, 15 parameters
  wurtzite_ab(...)  AB wurtzite 3x3x2 cells (72 atoms); NN2B pair potential
                    (species=True: with the species pair as a network input)
"""
from __future__ import annotations

import math
import os

import numpy as np

# Minimal Tersoff Si: ground truth = the elite of the reference's own test run
# (tests/si_diamond/ga1.out, last line), bounds of tests/si_diamond/potential.in.
TERSOFF_SI_TRUTH = np.array([3.18157, 1.43165, 2.24158, 2.0, 0.663971, 0.317766, -0.677574,
                             2.8, 3.2], np.float32)
TERSOFF_SI_MIN = np.array([2.9, 1.3, 2.2, 2, 0.5, 0.0, -0.8, 2.8, 3.2], np.float32)
TERSOFF_SI_MAX = np.array([3.3, 1.5, 2.4, 2, 0.8, 0.4, -0.6, 2.8, 3.2], np.float32)
TERSOFF_NAMES = ["D0", "alpha", "r0", "S", "n", "beta", "h", "R1", "R2"]

# Rigid ion AB2: charges q0 / -q0/2 (src/ri.cu:41-43), Buckingham and Morse
# values of develop/hfo2/potential.in in the reference, q0 = 4 x 0.6.
RI_TRUTH = np.array([2.4, 0.0, 1.0, 0.0, 12372.16, 0.2286, 81.36, 1388.77, 0.3623, 175.0,
                     10.0, 0.3474, 1.623, 2.048, 2.0], np.float32)
RI_NAMES = ["Q0", "A00", "rho00", "C00", "A01", "rho01", "C01", "A11", "rho11", "C11",
            "cutoff", "D0", "alpha", "r0", "S"]
RI_FIXED = {1, 2, 3, 10, 14}


def ri_bounds(spread=0.1):
    lo = RI_TRUTH.copy()
    hi = RI_TRUTH.copy()
    for k in range(15):
        if k not in RI_FIXED:
            lo[k] = RI_TRUTH[k] * (1 - spread)
            hi[k] = RI_TRUTH[k] * (1 + spread)
    return lo.astype(np.float32), hi.astype(np.float32)


_DIAMOND = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])


def _supercell(frac, types, reps):
    n1, n2, n3 = reps
    out, t = [], []
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                out.append((frac + [i, j, k]) / [n1, n2, n3])
                t.append(types)
    return np.concatenate(out), np.concatenate(t)


def _assemble(cells, force_fraction, rng):
    """cells: list of (box3x3 rows a,b,c; cart positions [Na,3]; types[Na])."""
    Nc = len(cells)
    Nc_force = max(1, min(Nc - 1, int(round(Nc * force_fraction))))
    Na = np.array([len(c[2]) for c in cells], np.int32)
    N = int(Na.sum())
    r = np.zeros((3, N), np.float32)
    typ = np.zeros(N, np.int32)
    box = np.zeros((Nc, 9), np.float32)
    o = 0
    for c, (b, pos, t) in enumerate(cells):
        n = len(t)
        r[:, o:o + n] = pos.T
        typ[o:o + n] = t
        box[c] = np.asarray(b, np.float32).reshape(9)
        o += n
    return {"Na": Na, "type": typ, "box": box.reshape(-1), "r": r.reshape(-1),
            "Nc_force": Nc_force, "pe_ref": np.zeros(Nc, np.float32),
            "virial_ref": np.zeros(6 * Nc, np.float32), "force_ref": np.zeros(3 * N, np.float32)}


def _diamond_reps_for(target, rng, max_rep=4):
    cands = [(a, b, c) for a in range(1, max_rep + 1) for b in range(1, max_rep + 1)
             for c in range(1, max_rep + 1)]
    sizes = np.array([8 * a * b * c for a, b, c in cands])
    logd = np.abs(np.log(sizes) - math.log(target))
    best = np.flatnonzero(logd <= logd.min() + 1e-12)
    return cands[int(rng.choice(best))]


def tersoff_si(Nc=60, sizes="fixed", reps=(2, 2, 2), a=5.43, strain=0.03, disp=0.1,
               force_fraction=None, seed=1):
    """Diamond-Si training set.  sizes="fixed": every cell is `reps` of the
    conventional cell; sizes="loguniform": 8..512 atoms, log-uniform."""
    rng = np.random.default_rng(seed)
    cells = []
    for _ in range(Nc):
        if sizes == "fixed":
            rp = reps
        else:
            target = math.exp(rng.uniform(math.log(8), math.log(512)))
            rp = _diamond_reps_for(target, rng)
        frac, t = _supercell(_DIAMOND, np.zeros(8, np.int32), rp)
        L = np.array(rp) * a * (1 + rng.uniform(-strain, strain, 3))
        pos = frac * L + rng.normal(0.0, disp, frac.shape)
        cells.append((np.diag(L), pos, t))
    ff = 0.2 if force_fraction is None else force_fraction
    d = _assemble(cells, ff, rng)
    d.update(potential_type=1, cutoff=3.0, weights=(0.05, 0.15, 0.8),
             pmin=TERSOFF_SI_MIN, pmax=TERSOFF_SI_MAX, names=TERSOFF_NAMES,
             truth=TERSOFF_SI_TRUTH)
    return d


def fluorite_ab2(Nc=500, a=5.1, strain=0.02, disp=0.05, force_fraction=0.2, seed=3):
    rng = np.random.default_rng(seed)
    cat = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    an = np.array([[x, y, z] for x in (.25, .75) for y in (.25, .75) for z in (.25, .75)])
    frac = np.concatenate([cat, an])
    t = np.array([0] * 4 + [1] * 8, np.int32)
    cells = []
    for _ in range(Nc):
        f2, t2 = _supercell(frac, t, (2, 2, 2))
        L = 2 * a * (1 + rng.uniform(-strain, strain, 3))
        pos = f2 * L + rng.normal(0.0, disp, f2.shape)
        cells.append((np.diag(L), pos, t2))
    d = _assemble(cells, force_fraction, rng)
    lo, hi = ri_bounds()
    d.update(potential_type=0, cutoff=10.0, weights=(0.1, 0.2, 0.7), pmin=lo, pmax=hi,
             names=RI_NAMES, truth=RI_TRUTH)
    return d


def nn2b_truth(n, seed=7):
    """A smooth, bounded ground-truth NN2B weight set with n neurons."""
    rng = np.random.default_rng(seed)
    V = n * (n + 4) + 1
    w = rng.uniform(-1.0, 1.0, V).astype(np.float32)
    w[:n] = rng.uniform(0.5, 1.5, n)              # w0: input scale ~ 1/A
    w[n:2 * n] = rng.uniform(1.0, 5.0, n)         # b0: centres over the pair range
    w[n * (n + 3):n * (n + 4)] *= 0.5             # w2
    return w.astype(np.float32)


def wurtzite_ab(Nc=800, a=3.82, c_over_a=1.633, u=0.375, strain=0.02, disp=0.05, n=32,
                force_fraction=0.2, seed=5, species=False):
    """AB wurtzite cells for NN2B; species=True: potential type 4, whose first
    layer also takes the pair's species code t_i + t_j (n more weights)."""
    rng = np.random.default_rng(seed)
    frac = np.array([[1 / 3, 2 / 3, 0], [2 / 3, 1 / 3, .5], [1 / 3, 2 / 3, u], [2 / 3, 1 / 3, .5 + u]])
    t = np.array([0, 0, 1, 1], np.int32)
    cells = []
    for _ in range(Nc):
        f2, t2 = _supercell(frac, t, (3, 3, 2))
        s = 1 + rng.uniform(-strain, strain, 3)
        A = np.array([[3 * a, 0, 0], [-1.5 * a, 1.5 * a * math.sqrt(3), 0],
                      [0, 0, 2 * a * c_over_a]]) * s[:, None]
        pos = f2 @ A + rng.normal(0.0, disp, f2.shape)
        cells.append((A, pos, t2))
    d = _assemble(cells, force_fraction, rng)
    V = n * (n + 4) + 1 + (n if species else 0)
    truth = nn2b_truth(n)
    if species:
        # species weights: A-A, A-B, B-B pairs see first-layer inputs shifted by
        # 0, ws, 2 ws
        ws = np.random.default_rng(seed + 1000).uniform(-0.5, 0.5, n).astype(np.float32)
        truth = np.concatenate([truth, ws]).astype(np.float32)
    d.update(potential_type=4 if species else 3, num_neurons=n, cutoff=6.0,
             weights=(0.9, 0.1, 0.0),
             pmin=np.full(V, -3, np.float32), pmax=np.full(V, 3, np.float32),
             names=[f"w{k}" for k in range(V)], truth=truth)
    return d


def label(d, find_force):
    """Fill pe_ref / virial_ref / force_ref from find_force(params) -> (f, w, e)
    per-atom arrays ([3][N], [6][N], [N]) at the ground-truth parameters."""
    f, w, e = find_force(d["truth"])
    f = np.asarray(f, np.float64).reshape(3, -1)
    w = np.asarray(w, np.float64).reshape(6, -1)
    e = np.asarray(e, np.float64)
    Na = d["Na"]
    off = np.concatenate([[0], np.cumsum(Na)])
    Nc = len(Na)
    pe = np.array([e[off[c]:off[c + 1]].sum() for c in range(Nc)])
    vir = np.array([[w[k, off[c]:off[c + 1]].sum() for c in range(Nc)] for k in range(6)])
    out = dict(d)
    out["pe_ref"] = pe.astype(np.float32)
    out["virial_ref"] = vir.reshape(-1).astype(np.float32)
    out["force_ref"] = f.reshape(-1).astype(np.float32)
    return out


def random_genes(P, V, seed=11):
    return np.random.default_rng(seed).random((P, V), dtype=np.float32)


def write_dir(d, path, ga=None):
    """Write potential.in, train.in and ga.in (reference formats)."""
    os.makedirs(path, exist_ok=True)
    names = d.get("names")
    with open(os.path.join(path, "potential.in"), "w") as f:
        f.write(f"potential_type {d['potential_type']}\n")
        if d["potential_type"] in (3, 4):
            f.write(f"num_neurons {d['num_neurons']}\n")
        wf, we, wv = d["weights"]
        f.write(f"cutoff {d['cutoff']!r}\nweight_force {wf!r}\nweight_energy {we!r}\n"
                f"weight_virial {wv!r}\n")
        if d["potential_type"] < 3:
            for k, (lo, hi) in enumerate(zip(d["pmin"], d["pmax"])):
                f.write(f"{names[k] if names else 'p%d' % k} {float(lo):.9g} {float(hi):.9g}\n")
    Na = np.asarray(d["Na"])
    Nc, Ncf = len(Na), int(d["Nc_force"])
    N = int(Na.sum())
    r = np.asarray(d["r"], np.float32).reshape(3, N)
    fr = np.asarray(d["force_ref"], np.float32).reshape(3, N)
    box = np.asarray(d["box"], np.float32).reshape(Nc, 9)
    pe = np.asarray(d["pe_ref"], np.float32)
    vir = np.asarray(d["virial_ref"], np.float32).reshape(6, Nc)
    typ = np.asarray(d["type"])
    g = lambda x: "%.9g" % float(x)
    lines = [f"{Nc} {Ncf}"] + [str(int(n)) for n in Na]
    o = 0
    for c in range(Nc):
        if c >= Ncf:
            lines.append(" ".join(g(x) for x in [pe[c], *vir[:, c]]))
        lines.append(" ".join(g(x) for x in box[c]))
        for i in range(o, o + Na[c]):
            if c < Ncf:
                lines.append(f"{typ[i]} " + " ".join(g(x) for x in (*r[:, i], *fr[:, i])))
            else:
                lines.append(f"{typ[i]} " + " ".join(g(x) for x in r[:, i]))
        o += Na[c]
    with open(os.path.join(path, "train.in"), "w") as f:
        f.write("\n".join(lines) + "\n")
    if ga is not None:
        with open(os.path.join(path, "ga.in"), "w") as f:
            for k, v in ga.items():
                f.write(f"{k} {v}\n")


def served_mix(root, jobs=16, generations=20, rng=None, log=None):
    """The served mixed workload's jobs: minimal Tersoff Si, AB2 rigid ion
    and NN2B wurtzite in turn, 100-2000 configurations (log-uniform; a
    quarter of that for rigid ion, half for NN2B), populations of 64-256,
    each labelled with its ground-truth parameters and written as a job
    directory under root.  rng: the generator the sizes are drawn from
    (default: seeded 2024).  Returns [(path, kind, P, atoms)]."""
    import gpuga_amd
    if rng is None:
        rng = np.random.default_rng(2024)
    kinds = ["tersoff", "ri", "nn2b"]
    out = []
    for j in range(jobs):
        kind = kinds[j % 3]
        Nc = int(math.exp(rng.uniform(math.log(100), math.log(2000))))
        P = int(rng.choice([64, 128, 192, 256]))
        if kind == "tersoff":
            d = tersoff_si(Nc=Nc, sizes="loguniform", seed=100 + j)
        elif kind == "ri":
            d = fluorite_ab2(Nc=max(2, Nc // 4), seed=100 + j)
        else:
            d = wurtzite_ab(Nc=max(2, Nc // 2), seed=100 + j)
        lab = gpuga_amd.Fitness(arrays=d)
        d = label(d, lab.find_force)
        lab.close()
        path = os.path.join(root, f"job{j:02d}_{kind}")
        write_dir(d, path, ga={"maximum_generation": generations, "population_size": P,
                               "parent_number": P // 2, "mutation_rate": 0.2,
                               "seed": 1000 + j})
        out.append((path, kind, P, int(np.sum(d["Na"]))))
        if log:
            log(f"prepared job {j}: {kind} Nc={len(d['Na'])} N={out[-1][3]} P={P}")
    return out

