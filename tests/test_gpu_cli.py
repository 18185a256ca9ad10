"""End to end: the `gpuga < input.txt` driver and Fitness::predict against the
reference's own output files (the reference's only test, tests/run1.bat:1-9,
runs the driver and compares force.out, energy.out, virial.out and ga.out).

hfo2 fixes every parameter (min = max in potential.in), so its outputs do not
depend on the random stream: every ga.out line is the reference's line with
fitness 2.2073, and the final energy/force/virial.out are the reference's at
those parameters.  Tolerances are those the FP64 oracle meets against the
same files (tests/test_oracle_golden.py):
  energy.out  5e-5 relative (the reference sums ~800 FP32 site energies)
  force.out   1e-4 relative + 5e-5 absolute
  virial.out  2e-5 relative + 2e-4 of the largest |W|
  fitness     1e-5 relative to 2.2073 (the reference prints 6 digits)
si_diamond: predict at the last ga.out elite (known to 6 digits) against the
reference's energy.out (2e-4 relative) and force.out (2e-3 of the largest
|F|), the bounds that follow from the elite's rounding.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, print_case, read_ga_out

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "gpuga_amd", "gpuga")


def _write_ga_in(d, G, P, parents):
    with open(os.path.join(d, "ga.in"), "w") as f:
        f.write(f"maximum_generation {G}\npopulation_size {P}\nparent_number {parents}\n"
                f"mutation_rate 0.2\n")


def _run_cli(tmp_path, dirs, env=None):
    inp = tmp_path / "input.txt"
    inp.write_text(f"{len(dirs)}\n" + "".join(f"{d}\n" for d in dirs))
    with open(inp) as fi:
        r = subprocess.run([CLI], stdin=fi, capture_output=True, text=True, timeout=300,
                           env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _ga_fields(line, V):
    # "%d %g " + V x "%g " + "\n" (src/ga.cu:141-147)
    assert re.fullmatch(r"-?\d+ (\S+ ){%d}\n" % (V + 1), line), repr(line)
    return line.split()


def test_cli_hfo2_matches_reference_outputs(golden, tmp_path):
    print_case("gpuga < input.txt", "hfo2")
    src = golden("hfo2")
    d = str(tmp_path / "hfo2")
    shutil.copytree(src, d)
    for f in ("ga.out", "energy.out", "force.out", "virial.out"):
        os.remove(os.path.join(d, f))
    G, V = 25, 15
    _write_ga_in(d, G, 20, 10)
    out = _run_cli(tmp_path, [d])
    assert "Finished running gpuga." in out
    ref_line = open(os.path.join(src, "ga.out")).readline()
    ref = _ga_fields(ref_line, V)
    lines = open(os.path.join(d, "ga.out")).readlines()
    assert len(lines) == G
    for g, line in enumerate(lines):
        v = _ga_fields(line, V)
        assert int(v[0]) == g
        assert abs(float(v[1]) - 2.2073) <= 2.2073 * 1e-5, line
        assert v[2:] == ref[2:]  # the fixed parameters, as the reference prints them
    # the screen copy every 10 generations (generation + 1)
    shown = [ln for ln in out.splitlines() if re.match(r"^(10|20) ", ln)]
    assert [int(s.split()[0]) for s in shown] == [10, 20]
    # predict's files at the final elite
    for name, rtol, atol_frac in (("energy.out", 5e-5, 0.0), ("virial.out", 2e-5, 2e-4)):
        got = np.loadtxt(os.path.join(d, name))
        want = np.loadtxt(os.path.join(src, name))
        assert got.shape == want.shape
        np.testing.assert_array_equal(got[:, 1], want[:, 1])  # reference column
        np.testing.assert_allclose(got[:, 0], want[:, 0], rtol=rtol,
                                   atol=atol_frac * np.abs(want[:, 0]).max())
    got = np.loadtxt(os.path.join(d, "force.out"))
    want = np.loadtxt(os.path.join(src, "force.out"))
    assert got.shape == want.shape
    np.testing.assert_array_equal(got[:, 3:], want[:, 3:])
    np.testing.assert_allclose(got[:, :3], want[:, :3], rtol=1e-4, atol=5e-5)


def test_cli_runs_jobs_in_turn_and_reports_errors(golden, tmp_path):
    print_case("gpuga < input.txt", "two jobs + missing dir")
    dirs = []
    for name in ("hfo2", "si_diamond"):
        d = str(tmp_path / name)
        shutil.copytree(golden(name), d)
        _write_ga_in(d, 3, 20, 10)
        dirs.append(d)
    out = _run_cli(tmp_path, dirs)
    assert out.count("Run simulation for") == 2
    for d in dirs:
        assert len(open(os.path.join(d, "ga.out")).readlines()) == 3
    inp = tmp_path / "bad.txt"
    inp.write_text(f"1\n{tmp_path / 'missing'}\n")
    with open(inp) as fi:
        r = subprocess.run([CLI], stdin=fi, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "ERROR" in r.stdout


def test_cli_device_ga_equals_host_ga_loop(golden, tmp_path):
    """The driver runs generations with the device GA (gpuga_fitness_run_ga);
    its ga.out equals a host-GA loop over the same population launches."""
    import gpuga_amd as G
    print_case("gpuga < input.txt vs host loop", "si_diamond")
    d = str(tmp_path / "si")
    shutil.copytree(golden("si_diamond"), d)
    gens = 12
    _write_ga_in(d, gens, 40, 20)
    _run_cli(tmp_path, [d])
    cli_lines = open(os.path.join(d, "ga.out")).readlines()
    with G.Fitness(d) as fit:
        ga = G.GA(40, 20, fit.V, maximum_generation=gens, mutation_rate=0.2, seed=12345678)
        pmin, pmax = fit.parameters_min, fit.parameters_max
        for g in range(gens):
            s = ga.step(g, fit.compute(ga.population))
            elite = ga.population[0]
            b = (pmax - pmin).astype(np.float32)
            params = (pmin + (b * elite).astype(np.float32)).astype(np.float32)
            want = "%d %s " % (g, _g(s[0])) + "".join(_g(x) + " " for x in params) + "\n"
            assert cli_lines[g] == want, (g, cli_lines[g], want)


def _g(x):
    """C's %g of a float (printf promotes it to double)."""
    return "%g" % float(np.float32(x))


def test_predict_at_reference_elite(golden, tmp_path):
    import gpuga_amd as G
    print_case("predict", "si_diamond")
    src = golden("si_diamond")
    params = read_ga_out(os.path.join(src, "ga.out"))[-1][2]
    with G.Fitness(src) as fit:
        pmin, pmax = fit.parameters_min, fit.parameters_max
        span = np.where(pmax > pmin, pmax - pmin, 1).astype(np.float64)
        genes = np.clip((params - pmin) / span, 0, 1).astype(np.float32)
        assert np.allclose(fit.denormalise(genes), params, rtol=1e-6, atol=1e-7)
        fit.predict(str(tmp_path), genes)
    eo = np.loadtxt(os.path.join(src, "energy.out"))
    e = np.loadtxt(tmp_path / "energy.out")
    assert e.shape == eo.shape
    np.testing.assert_array_equal(e[:, 1], eo[:, 1])
    np.testing.assert_allclose(e[:, 0], eo[:, 0], rtol=2e-4)
    fo = np.loadtxt(os.path.join(src, "force.out"))
    f = np.loadtxt(tmp_path / "force.out")
    assert f.shape == fo.shape
    np.testing.assert_array_equal(f[:, 3:], fo[:, 3:])
    np.testing.assert_allclose(f[:, :3], fo[:, :3], atol=2e-3 * np.abs(fo[:, :3]).max())
    vo = np.loadtxt(os.path.join(src, "virial.out"))
    v = np.loadtxt(tmp_path / "virial.out")
    assert v.shape == vo.shape
    np.testing.assert_array_equal(v[:, 1], vo[:, 1])
    np.testing.assert_allclose(v[:, 0], vo[:, 0], atol=2e-3 * np.abs(vo[:, 0]).max())
