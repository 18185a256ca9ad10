"""Benchmark: GA generations of the batched minimal-Tersoff Si workload.

Workload (the batched Tersoff configuration): population 200, 1000 diamond-Si
configurations of 8-512 atoms (log-uniform), +-3% axis strain, N(0, 0.1 A)
displacements, 20% force configurations, labels from the ground-truth
parameter set (computed by this library's find_force), cutoff 3.0 A.

A step is one GA generation: the fitness of all 200 individuals over all
configurations (one persistent, graph-replayed HIP launch) followed by the
GA's sort / crossover / mutation, by default on the device as well, so the
timed generations run back to back with the population resident (--host-ga:
genes and fitness cross to the host GA every generation).  The metric is
individual x atom fitness evaluations per second.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload NAME]
                  [--no-cpu-baseline] [--plain] [--host-ga]

Extra workloads (not the headline line): tersoff_small, rigid_ion, nn2b,
nn2b_species (NN2B with the species pair as a network input), serve (16 mixed
jobs, Poisson arrivals).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_FP32_PEAK = 157.3e12  # FLOP/s, vector FP32 (vendor figure)
MI355X_HBM_PEAK = 8.0e12     # B/s
# Algorithmic work per individual x atom evaluation (4 neighbours per atom):
# pair terms ~106 FLOP x 4, triplets ~57 FLOP x 12, assembly -> ~1.1 kFLOP.
TERSOFF_FLOP_PER_UNIT = 1.1e3
# Pair potentials: the work per pair the population kernels evaluate (each
# unordered pair once in energy/virial configurations, every ordered pair in
# force configurations; distance-only factors such as erfc(a r)/r come from
# the set-up tables).  FMA = 2 FLOP; exp2 not counted.
#  RI, per pair: Buckingham + shifted-force Coulomb value and derivative
#  (14 FLOP), + energy and virial accumulation in energy configurations
#  (31 FLOP in all); a Morse term on unlike pairs adds 10.
RI_FLOP_FORCE_PAIR, RI_FLOP_ENERGY_PAIR, RI_FLOP_MORSE = 14.0, 31.0, 10.0
#  NN2B: per individual the 1-n-n-1 network with its first two derivatives
#  on every grid point of the tabulation (6 n^2 + 23 n FLOP), then per pair
#  the Hermite reconstruction of p and p' and the taper (18 FLOP) plus the
#  force (6) or energy and virial (17) accumulation.
NN2B_FLOP_FORCE_PAIR, NN2B_FLOP_ENERGY_PAIR = 24.0, 35.0
NN2B_GRID = 1024

WORKLOADS = {
    "tersoff_batched": dict(kind="tersoff", Nc=1000, sizes="loguniform", P=200, seed=2),
    "tersoff_small": dict(kind="tersoff", Nc=60, sizes="fixed", P=50, seed=1),
    "rigid_ion": dict(kind="ri", Nc=500, P=200, seed=3),
    "nn2b": dict(kind="nn2b", Nc=800, P=256, seed=5),
    # the same with the species pair as a second network input (type 4)
    "nn2b_species": dict(kind="nn2b", Nc=800, P=256, seed=5, species=True),
}


def make_workload(name):
    import gpuga_amd
    from gpuga_amd import synth
    w = WORKLOADS[name]
    if w["kind"] == "tersoff":
        d = synth.tersoff_si(Nc=w["Nc"], sizes=w["sizes"], seed=w["seed"])
    elif w["kind"] == "ri":
        d = synth.fluorite_ab2(Nc=w["Nc"], seed=w["seed"])
    else:
        d = synth.wurtzite_ab(Nc=w["Nc"], seed=w["seed"], species=w.get("species", False))
    unlabelled = gpuga_amd.Fitness(arrays=d)
    d = synth.label(d, unlabelled.find_force)
    unlabelled.close()
    return d, w


def run_serve(args):
    """Served mixed workload: 16 fitting jobs (minimal Tersoff Si, AB2 rigid
    ion, NN2B wurtzite; 100-2000 configurations; populations 64-256) arrive
    as a Poisson process at 2 jobs/s and are served concurrently with
    training sets and populations resident on the GPU, admitted by a
    device-memory budget.  Returns the measurement as a dict."""
    import tempfile

    import gpuga_amd
    from gpuga_amd import synth
    root = tempfile.mkdtemp(prefix="gpuga_serve_")
    rng = np.random.default_rng(2024)
    jobs = synth.served_mix(root, jobs=args.jobs, generations=args.generations, rng=rng,
                            log=lambda m: print(m, file=sys.stderr, flush=True))
    # arrival times continue the same stream
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, len(jobs)))
    arrivals -= arrivals[0]
    srv = gpuga_amd.Server(memory_budget_bytes=args.budget_gb * 1e9, max_active=16,
                           serve_share=1 if args.share_gpu else 0)
    t0 = time.perf_counter()
    submitted = 0
    while True:
        now = time.perf_counter() - t0
        while submitted < len(jobs) and arrivals[submitted] <= now:
            srv.submit(jobs[submitted][0])
            submitted += 1
        a, q, f = srv.poll()
        if submitted == len(jobs) and f == len(jobs):
            break
        if a == 0 and q == 0 and submitted < len(jobs):
            time.sleep(max(0.0, min(0.01, arrivals[submitted] - now)))
    wall = time.perf_counter() - t0
    lat = srv.latencies()
    adm = srv.admissions()
    srv.close()
    gens = len(lat)
    work = sum(P * N * args.generations for _, _, P, N in jobs)
    return dict(jobs=jobs, wall=wall, lat=lat, adm=adm, gens=gens, work=work)


def serve_summary(r, args):
    """The served workload's figures as they appear in the bench line."""
    lat, adm = r["lat"], r["adm"]
    return {"generations_per_s": r["gens"] / r["wall"], "generations": r["gens"],
            "wall_s": r["wall"], "jobs": len(r["jobs"]), "arrival_rate_per_s": args.rate,
            "generations_per_job": args.generations,
            "p50_generation_ms": float(np.percentile(lat, 50)),
            "p99_generation_ms": float(np.percentile(lat, 99)),
            "p50_admission_ms": float(np.percentile(adm, 50)),
            "p99_admission_ms": float(np.percentile(adm, 99)),
            "individual_atom_evals_per_s": r["work"] / r["wall"],
            "mix": [f"{k}:P{P}:N{N}" for _, k, P, N in r["jobs"]]}


def serve(args):
    r = run_serve(args)
    jobs, wall, gens = r["jobs"], r["wall"], r["gens"]
    sm = serve_summary(r, args)
    out = {
        "metric": "generations/s (served mixed workload)",
        "value": gens / wall,
        "unit": "generations/s",
        "n_gpus": 1,
        "steps": gens,
        "warmup": 0,
        "ms_per_step": 1e3 * wall / gens,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 sums)",
        "data": "synthetic (seeded Si, AB2 fluorite and AB wurtzite cells)",
        "config": dict(workload="serve", **{k: v for k, v in sm.items()
                                             if k not in ("generations_per_s", "generations")}),
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(d, w, genes, budget_s=12.0):
    """FP64 oracle on the same training set, a bounded sample of individuals:
    one thread (the reported value), then OpenMP over individuals with one
    thread per host core of this process's share (at most 16)."""
    from oracle import oracle as O
    case = O.Case(potential=dict(potential_type=d["potential_type"], cutoff=d["cutoff"],
                                 weights=d["weights"], pmin=d["pmin"], pmax=d["pmax"],
                                 num_neurons=d.get("num_neurons", 0)),
                  arrays=d)
    N = case.N
    t0 = time.perf_counter()
    case.fitness(genes[:1], threads=1)
    t1 = time.perf_counter() - t0
    n = max(1, min(len(genes), int(budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    case.fitness(genes[:n], threads=1)
    dt = time.perf_counter() - t0
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    m = max(threads, min(len(genes), threads * n))
    t0 = time.perf_counter()
    case.fitness(genes[:m], threads=threads)
    dtm = time.perf_counter() - t0
    return {"value": n * N / dt, "unit": "individual*atom evals/s", "cores": 1, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{n} individuals x {N} atoms (full training set), FP64 scalar C oracle, "
                      f"1 thread, {dt:.1f} s",
            "openmp": {"value": m * N / dtm, "cores": threads,
                       "sample": f"{m} individuals x {N} atoms, {threads} threads, {dtm:.1f} s"}}


def pair_counts(fit, d):
    """Pairs a pair-potential population kernel evaluates per individual, by
    pair type t_i + t_j: every bond of a force-configuration atom, each
    unordered pair of the other configurations."""
    NN, start, NL = fit.neighbors()
    typ = np.asarray(d["type"])
    owner = np.repeat(np.arange(fit.N), NN)
    pt = typ[owner] + typ[NL]
    force = owner < fit.N_force
    fp = np.bincount(pt[force], minlength=3)[:3]
    ep = np.bincount(pt[~force & (NL > owner)], minlength=3)[:3]
    return fp.astype(np.float64), ep.astype(np.float64)


def plain_leg(fit, ga, steps, batched_ms):
    """The per-individual-launch form of the same generation on the same GPU
    (one find_force, one reduction and one host synchronisation per
    individual, then the host GA step), timed over `steps` generations after
    an untimed one, against the batched generation time of this run."""
    import gpuga_amd
    P, V = ga.population.shape
    g = gpuga_amd.GA(P, ga.parent_number, V, maximum_generation=steps + 1, mutation_rate=0.2,
                     seed=4242)
    g.step(0, fit.compute(g.population, plain=True))
    t0 = time.perf_counter()
    for s in range(1, steps + 1):
        g.step(s, fit.compute(g.population, plain=True))
    ms = 1e3 * (time.perf_counter() - t0) / steps
    return {"ms_per_generation": ms, "generations": steps,
            "individual_atom_evals_per_s": P * fit.N / (ms * 1e-3),
            "batched_ms_per_generation": batched_ms, "speedup_batched_over_plain": ms / batched_ms}


def trace_split(fit, genes, n_force_cfg):
    """Where the population launch's wave time goes, from one generation's
    work-item timeline (gpuga_fitness_trace): per resident wave, the time in
    phase-0 items of force and of energy configurations and in
    force-assembly items (an item's time ends when the wave has fetched its
    next item, so it includes that fetch), before its first item (entry and
    first fetch) and idle after its last one, as fractions of waves x launch
    span."""
    rows = fit.trace(genes)
    c, k, a0, a1, ph, slot, t0, t1, entry, cb, ce = (
        rows[:, 0], rows[:, 1], rows[:, 3], rows[:, 4], rows[:, 5], rows[:, 6], rows[:, 7],
        rows[:, 8], rows[:, 9], rows[:, 10], rows[:, 11])
    begin = float(np.nanmin(np.where(np.isnan(entry), t0, entry)))
    end = float(t1.max())
    span = end - begin
    waves = np.unique(slot)
    dur = t1 - t0
    force = c < n_force_cfg
    cls = {"phase0_force": (ph == 0) & force, "phase0_energy": (ph == 0) & ~force,
           "phase1_assembly": ph == 1}
    total = span * len(waves)
    out = {"span_us": span / 1e3, "waves": int(len(waves)), "items": int(len(rows))}
    for name, m in cls.items():
        atoms = float(np.sum(a1[m] - a0[m]))
        # an item's time in three parts: set-up (item load, parameters, or
        # the dependency wait), compute, and the rest (reductions, next fetch)
        setup, comp, rest = cb[m] - t0[m], ce[m] - cb[m], t1[m] - ce[m]
        out[name] = {"frac": float(dur[m].sum() / total), "items": int(m.sum()),
                     "mean_item_us": float(dur[m].mean() / 1e3) if m.any() else 0.0,
                     "us_per_atom_step": float(dur[m].sum() / max(atoms, 1) / 1e3),
                     "mean_setup_us": float(np.mean(setup) / 1e3) if m.any() else 0.0,
                     "mean_compute_us": float(np.mean(comp) / 1e3) if m.any() else 0.0,
                     "mean_after_us": float(np.mean(rest) / 1e3) if m.any() else 0.0,
                     "compute_us_per_atom": float(comp.sum() / max(atoms, 1) / 1e3)}
    head = tail = 0.0
    for w in waves:
        m = slot == w
        head += float(t0[m].min()) - begin
        tail += end - float(t1[m].max())
    out["head_frac"] = head / total
    out["tail_frac"] = tail / total
    # item time against its atom count (a least-squares line per class)
    steps = a1 - a0
    for name, m in cls.items():
        if m.sum() > 2:
            A = np.stack([np.ones(m.sum()), steps[m]], 1)
            coef = np.linalg.lstsq(A, dur[m], rcond=None)[0]
            out[name]["fit_us"] = {"per_item": float(coef[0] / 1e3),
                                   "per_atom_step": float(coef[1] / 1e3)}
    last = np.argsort(t1)[-12:]
    out["last_items"] = [{"c": int(c[i]), "k": int(k[i]), "phase": int(ph[i]),
                          "atoms": int(a1[i] - a0[i]), "start_us": float((t0[i] - begin) / 1e3),
                          "compute_us": [float((cb[i] - begin) / 1e3),
                                         float((ce[i] - begin) / 1e3)],
                          "end_us": float((t1[i] - begin) / 1e3)} for i in last]
    out["compute_end_us_max"] = float((np.nanmax(ce) - begin) / 1e3)
    # when the waves that took an item entered the kernel (dispatch spread)
    ent = (entry[~np.isnan(entry)] - begin) / 1e3
    if ent.size:
        out["wave_entry_us"] = {q: float(np.percentile(ent, v)) for q, v in
                                (("p10", 10), ("p50", 50), ("p90", 90), ("max", 100))}
    # speed by dispatch round: workgroup b of the grid is resident on its CU
    # with b mod (grid / occupancy) as the slot; round = b // (workgroups per
    # round) orders the waves of one SIMD by age
    wg = (slot // 4).astype(int)
    n_wg = int(wg.max()) + 1
    rounds = 4
    per_round = max(1, n_wg // rounds)
    rnd = np.minimum(wg // per_round, rounds - 1)
    ph0 = ph == 0
    out["us_per_atom_by_round"] = [
        float(dur[ph0 & (rnd == r)].sum() / max(float(np.sum((a1 - a0)[ph0 & (rnd == r)])), 1) / 1e3)
        for r in range(rounds)]
    starts = np.sort(t0 - begin)
    out["item_start_us"] = {"p90": float(np.percentile(starts, 90) / 1e3),
                            "p99": float(np.percentile(starts, 99) / 1e3),
                            "max": float(starts[-1] / 1e3)}
    ends = np.sort([float(t1[slot == w].max()) - begin for w in waves])
    out["wave_end_us"] = {"p10": float(np.percentile(ends, 10) / 1e3),
                          "p50": float(np.percentile(ends, 50) / 1e3),
                          "p90": float(np.percentile(ends, 90) / 1e3),
                          "max": float(ends[-1] / 1e3)}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measured_traffic(workload):
    """HBM bytes per launch of the dominant kernel from the committed
    FETCH_SIZE / WRITE_SIZE profile of this workload (each counter scaled by
    what it reads for a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f).get(workload)
    except (OSError, ValueError):
        return None
    if not t:
        return None
    return 1024.0 * (t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="tersoff_batched",
                    choices=sorted(WORKLOADS) + ["serve"])
    ap.add_argument("--jobs", type=int, default=16, help="serve: number of jobs")
    ap.add_argument("--rate", type=float, default=2.0, help="serve: arrivals per second")
    ap.add_argument("--generations", type=int, default=20, help="serve: generations per job")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="serve: device memory budget")
    ap.add_argument("--share-gpu", action="store_true",
                    help="serve: each running job's launches on its share of the GPU")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain", action="store_true", help="per-individual launches")
    ap.add_argument("--host-ga", action="store_true",
                    help="GA step on the host each generation (genes and fitness cross PCIe)")
    ap.add_argument("--chunk-cost", type=int, default=0,
                    help="work-item size; 0 = autotune before timing")
    ap.add_argument("--no-serve", action="store_true",
                    help="headline line without the served-workload leg")
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="generations of the per-individual-launch leg (0: none)")
    ap.add_argument("--tail-split", type=int, default=0,
                    help="cut configurations smaller than one item into this many "
                         "pieces per item (0: the library default)")
    ap.add_argument("--trace-split", action="store_true",
                    help="add the work-item timeline split of one population launch")
    ap.add_argument("--population", type=int, default=0,
                    help="population size (0: the workload's own)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the population-sharded multi-GPU form "
              "(replicated training set, one all-gather of P fitness values per "
              "generation) is not built.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "serve":
        serve(args)
        return
    import gpuga_amd
    from gpuga_amd import synth
    d, w = make_workload(args.workload)
    fit = gpuga_amd.Fitness(arrays=d, backend=1 if args.plain else 0, chunk_cost=args.chunk_cost,
                            tail_split=args.tail_split)
    P, V, N = (args.population or w["P"]), fit.V, fit.N
    chunk = args.chunk_cost
    shape = None
    if not args.plain:
        # (tunes the shape first when --chunk-cost is 0)
        shape = fit.launch_shape(P)
        chunk = shape["chunk_cost"]
    parents = P // 2 + (P - P // 2) % 2  # the GA pairs the children: P - parents even
    ga = gpuga_amd.GA(P, parents, V, maximum_generation=args.warmup + args.steps,
                      mutation_rate=0.2, seed=12345678)

    def step(g):
        genes = ga.population
        f = fit.compute(genes, plain=args.plain)
        ga.step(g, f)
        return f

    device_ga = not args.plain and not args.host_ga
    if device_ga:
        # generations back to back on the device: population launch + GA step
        fit.run_ga(ga, 0, args.warmup)
        t0 = time.perf_counter()
        bests, _ = fit.run_ga(ga, args.warmup, args.steps)
        elapsed = time.perf_counter() - t0
        best = float(bests[-1])
        # the population launch alone, for the roofline (one untimed call
        # first: it captures the launch's graph, which the events would
        # otherwise count as idle device time)
        genes = ga.population
        fit.compute(genes)
        fit.set_timing(True)
        fit.timing(reset=True)
        for _ in range(10):
            fit.compute(genes)
        kms, launches = fit.timing(reset=True)
    else:
        for g in range(args.warmup):
            step(g)
        fit.set_timing(True)
        fit.timing(reset=True)
        t0 = time.perf_counter()
        best = None
        for g in range(args.warmup, args.warmup + args.steps):
            best = float(np.min(step(g)))
        elapsed = time.perf_counter() - t0
        kms, launches = fit.timing(reset=True)
    ms_step = 1e3 * elapsed / args.steps
    value = P * N * args.steps / elapsed
    kernel_s = kms * 1e-3 / max(launches, 1)
    # Algorithmic work per generation (the per-unit counts above, FMA = 2
    # FLOP) over the generation's kernel time.
    peak, bound = MI355X_FP32_PEAK, "valu"
    if w["kind"] == "tersoff":
        flop = TERSOFF_FLOP_PER_UNIT * P * N
        note = None
    else:
        fp, ep = pair_counts(fit, d)
        if w["kind"] == "ri":
            # the Morse term (parameter 11, D0) runs on unlike pairs when any
            # individual can have one
            morse = RI_FLOP_MORSE if (d["pmin"][11] != 0 or d["pmax"][11] != 0) else 0.0
            flop = P * (RI_FLOP_FORCE_PAIR * fp.sum() + RI_FLOP_ENERGY_PAIR * ep.sum() +
                        morse * (fp[1] + ep[1]))
            note = (f"{int(fp.sum())} ordered force-configuration pairs, {int(ep.sum())} "
                    "unordered energy-configuration pairs per individual")
        else:
            n = fit.num_neurons
            flop = P * (NN2B_FLOP_FORCE_PAIR * fp.sum() + NN2B_FLOP_ENERGY_PAIR * ep.sum() +
                        NN2B_GRID * (6.0 * n * n + 23.0 * n))
            note = (f"{int(fp.sum())} ordered force-configuration pairs, {int(ep.sum())} "
                    f"unordered energy-configuration pairs and a {NN2B_GRID}-point tabulation "
                    "per individual; time = tabulation + pair kernel")
    if args.plain:
        flop /= P  # timing covers one individual per launch group
    achieved = flop / kernel_s
    data = {"tersoff": "seeded diamond-Si cells", "ri": "seeded AB2 fluorite cells",
            "nn2b": "seeded AB wurtzite cells"}[w["kind"]]
    out = {
        "metric": "individual*atom fitness evals/s",
        "value": value,
        "unit": "individual*atom evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 sums)",
        "data": f"synthetic ({data}, labels from ground-truth parameters)",
        "config": {"workload": args.workload, "population": P, "configurations": fit.Nc,
                   "atoms": N, "bonds": fit.bonds, "parameters": V,
                   "backend": "per-individual" if args.plain else "population",
                   "ga": "device" if device_ga else "host",
                   "chunk_cost": chunk, "launch_shape": shape,
                   "s_per_generation": ms_step / 1e3,
                   "best_fitness": best},
        "roofline": {"bound": bound, "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None if args.plain else measured_traffic(args.workload),
                     "kernel_ms": kernel_s * 1e3},
    }
    if note:
        out["roofline"]["note"] = note
    if out["roofline"]["traffic"] is not None:
        out["roofline"]["traffic_source"] = (
            "rocprofv3 FETCH_SIZE + WRITE_SIZE of this workload's population kernel, "
            "each scaled by the counter's reading for a copy of known size "
            "(profiles/traffic.json; not measured in this run)")
    if args.trace_split and not args.plain:
        out["trace_split"] = trace_split(fit, ga.population, fit.Nc_force)
    if device_ga and args.plain_steps > 0:
        out["plain"] = plain_leg(fit, ga, args.plain_steps, ms_step)
    fit.close()
    if not args.no_cpu_baseline and not args.plain:
        genes = synth.random_genes(256, V, seed=99)
        out["cpu_baseline"] = cpu_baseline(d, w, genes)
    if args.workload == "tersoff_batched" and device_ga and not args.no_serve:
        r = run_serve(args)
        out["serve"] = serve_summary(r, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
