"""Serving loop: concurrent fitting jobs, each with its own resident training
set, population and stream.  A job's results (best fitness, elite, ga.out)
must be the same run alone and run together with other jobs, and the
memory budget must hold jobs back until memory frees."""
import os
import shutil
import time

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def job_dir(golden, tmp, name, gens, P, parents, seed):
    d = os.path.join(tmp, f"{name}_{seed}")
    if os.path.isdir(d):
        shutil.rmtree(d)
    shutil.copytree(golden(name), d)
    with open(os.path.join(d, "ga.in"), "w") as f:
        f.write(f"maximum_generation {gens}\npopulation_size {P}\nparent_number {parents}\n"
                f"mutation_rate 0.2\nseed {seed}\n")
    return d


JOBS = [("si_diamond", 6, 64, 32, 1), ("gaas", 5, 40, 20, 2), ("nn2b", 6, 70, 36, 3),
        ("si_example", 4, 130, 64, 4)]


def run_jobs(G, dirs, budget=64e9, max_active=8, share=0):
    s = G.Server(memory_budget_bytes=budget, max_active=max_active, nn2b_bf16=0,
                 serve_share=share)
    ids = [s.submit(d, write_outputs=True) for d in dirs]
    s.drain()
    return s, ids


@pytest.mark.parametrize("share", [0, 1])
def test_results_same_alone_and_in_the_mix(golden, tmp_path, share):
    """share = 1: in the mix each running job's launches use 1/n of the GPU
    (a different grid than alone); the results must not change."""
    import gpuga_amd as G
    print_case("serve alone vs mix", share)
    alone, mixed = {}, {}
    for name, gens, P, par, seed in JOBS:
        d = job_dir(golden, str(tmp_path / "alone"), name, gens, P, par, seed)
        s, ids = run_jobs(G, [d])
        info = s.job(ids[0], 2000)
        alone[name] = (info, open(os.path.join(d, "ga.out")).read())
        assert info["state"] == 2 and info["generations"] == gens
    dirs = [job_dir(golden, str(tmp_path / "mix"), *j) for j in JOBS]
    s, ids = run_jobs(G, dirs, share=share)
    for (name, gens, *_), jid, d in zip(JOBS, ids, dirs):
        info = s.job(jid, 2000)
        assert info["state"] == 2
        assert info["best"] == alone[name][0]["best"]
        assert np.array_equal(info["elite"], alone[name][0]["elite"])
        assert open(os.path.join(d, "ga.out")).read() == alone[name][1]
        for out in ("force.out", "energy.out", "virial.out"):
            assert os.path.getsize(os.path.join(d, out)) > 0
    lat = s.latencies()
    assert len(lat) == sum(j[1] for j in JOBS)
    assert np.all(lat > 0)


def test_memory_budget_serialises_jobs(golden, tmp_path):
    import gpuga_amd as G
    print_case("serve budget")
    dirs = [job_dir(golden, str(tmp_path), "si_diamond", 3, 64, 32, s) for s in (11, 12, 13)]
    s = G.Server(memory_budget_bytes=1.0, max_active=8)  # admits one job at a time
    ids = [s.submit(d) for d in dirs]
    # files are read on worker threads: poll until the first job is admitted
    seen = []
    for _ in range(20000):
        a, q, f = s.poll()
        seen.append(a)
        if a > 0 or f > 0:
            break
        time.sleep(0.001)
    assert a == 1 and q == 2
    while f < len(ids):
        a, q, f = s.poll()
        seen.append(a)
    assert max(seen) == 1  # never two jobs resident under the budget
    assert all(s.job(i, 9)["state"] == 2 for i in ids)


def test_failed_jobs_do_not_stop_the_others(golden, tmp_path):
    """A job whose files cannot be read fails on its worker with the
    reference's message; the other jobs are served to the end."""
    import gpuga_amd as G
    print_case("serve failure")
    good = job_dir(golden, str(tmp_path), "si_diamond", 3, 64, 32, 21)
    bad = os.path.join(str(tmp_path), "missing")
    broken = job_dir(golden, str(tmp_path), "gaas", 3, 40, 20, 22)
    with open(os.path.join(broken, "ga.in"), "w") as f:
        f.write("population_size 10\n")
    s = G.Server(memory_budget_bytes=64e9, max_active=8)
    ids = [s.submit(d) for d in (bad, good, broken)]
    s.drain()
    with pytest.raises(G.GpugaError, match="Failed to open"):
        s.job(ids[0], 9)
    assert s.job(ids[1], 9)["state"] == 2 and s.job(ids[1], 9)["generations"] == 3
    with pytest.raises(G.GpugaError, match="population_size should >= 20"):
        s.job(ids[2], 9)
    assert len(s.admissions()) == 1


def test_served_mix_alone_equals_mix(tmp_path):
    """The served mixed workload's own jobs (16: minimal Tersoff Si, AB2 rigid
    ion and NN2B wurtzite in turn, 100-2000 configurations, P 64-256), two
    generations each: every job's best fitness, elite and ga.out are the
    same run alone and submitted together with all the others."""
    import gpuga_amd as G
    from gpuga_amd import synth
    print_case("served mix, alone vs mix")
    jobs = synth.served_mix(str(tmp_path / "alone"), jobs=16, generations=2)
    assert {k for _, k, _, _ in jobs} == {"tersoff", "ri", "nn2b"}
    alone = {}
    for path, kind, P, N in jobs:
        s, ids = run_jobs(G, [path])
        info = s.job(ids[0], 2000)
        assert info["state"] == 2 and info["generations"] == 2, (path, info)
        alone[os.path.basename(path)] = (info, open(os.path.join(path, "ga.out")).read())
        s.close()
    mixed = synth.served_mix(str(tmp_path / "mix"), jobs=16, generations=2)
    s, ids = run_jobs(G, [p for p, *_ in mixed], max_active=16)
    for (path, *_), jid in zip(mixed, ids):
        info = s.job(jid, 2000)
        ref, ref_out = alone[os.path.basename(path)]
        assert info["state"] == 2
        assert info["best"] == ref["best"], path
        assert np.array_equal(info["elite"], ref["elite"]), path
        assert open(os.path.join(path, "ga.out")).read() == ref_out, path
    s.close()
