"""The allocation-free entry point (gpuga_fitness_compute_plain_noalloc)
against gpuga_fitness_compute_plain, through the C ABI: the same kernels in
the same order on the fixed per-GPU workspace instead of the handle's
buffers, so fitness and terms are bit-identical.  Calls in a row share the
workspace, which keeps the last call's contents: a larger training set
first, then smaller ones (and another potential), each compared with the
allocating path; a set that needs more than the workspace is refused."""
import ctypes as C

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def _noalloc(G, fit, genes):
    g = np.ascontiguousarray(genes, np.float32)
    P = g.shape[0]
    f = np.zeros(P, np.float32)
    t = np.zeros((P, 3), np.float64)
    rc = G.lib().gpuga_fitness_compute_plain_noalloc(
        fit._h, P, g.ctypes.data_as(C.POINTER(C.c_float)), f.ctypes.data_as(C.POINTER(C.c_float)),
        t.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise G.GpugaError(G.lib().gpuga_last_error().decode())
    return f, t


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_noalloc_equals_plain_on_calls_in_a_row(golden):
    import gpuga_amd as G
    from gpuga_amd import synth
    from oracle import oracle as O
    small = synth.tersoff_si(Nc=4, reps=(1, 1, 1), seed=31)
    small = synth.label(small, O.Case(potential=dict(
        potential_type=1, cutoff=small["cutoff"], weights=small["weights"], pmin=small["pmin"],
        pmax=small["pmax"]), arrays=small).find_force)
    cases = [("si_diamond", dict(input_dir=golden("si_diamond")), 12),
             ("si_diamond again", dict(input_dir=golden("si_diamond")), 7),
             ("small Tersoff set", dict(arrays=small), 5),
             ("hfo2 (rigid ion)", dict(input_dir=golden("hfo2")), 4)]
    for n, (name, src, P) in enumerate(cases):
        print_case("noalloc", name, P)
        if "input_dir" in src:
            fit = G.Fitness(src["input_dir"])
        else:
            fit = G.Fitness(arrays=src["arrays"])
        try:
            genes = np.random.default_rng(40 + n).random((P, fit.V), dtype=np.float32)
            f0, t0 = _noalloc(G, fit, genes)
            f1, t1 = fit.compute(genes, terms=True, plain=True)
            f2, t2 = _noalloc(G, fit, genes[: max(1, P // 2)])
        finally:
            fit.close()
        assert _same(f0, f1) and _same(t0, t1), name
        assert _same(f2, f1[: max(1, P // 2)]) and _same(t2, t1[: max(1, P // 2)]), name
        assert np.all(np.isfinite(f0))


def test_noalloc_refuses_a_set_beyond_the_workspace():
    import gpuga_amd as G
    from gpuga_amd import synth
    print_case("noalloc", "beyond the workspace")
    # 1200 cells of 512 atoms: 30 floats per atom of workspace, 74 MB > 64 MiB
    d = synth.tersoff_si(Nc=1200, reps=(4, 4, 4), seed=32)
    fit = G.Fitness(arrays=d)
    try:
        genes = np.full((1, fit.V), 0.5, np.float32)
        with pytest.raises(G.GpugaError, match="the fixed workspace holds"):
            _noalloc(G, fit, genes)
    finally:
        fit.close()
