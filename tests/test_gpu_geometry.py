"""Geometry tables built on the device (bond vectors and lengths, triplet
cosines and gradients, the Tersoff atom records) against the host build of
the same tables, bit for bit: the minimum image and the triplet factors are
FP64 expressions evaluated without contraction on both sides, with IEEE
correctly rounded add, multiply, divide and square root, then rounded to
float.  Cases: the golden one- and two-species Tersoff sets (orthorhombic and
hexagonal boxes), the rigid-ion and NN2B sets (bond table only), and
synthetic Tersoff sets whose cutoffs give records of 4, 8 and 16 bonds."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _check(d):
    import gpuga_amd as G
    host = G.inspect_geometry(d)
    fit = G.Fitness(d)
    try:
        dev = fit.geometry()
    finally:
        fit.close()
    for name, h, g in zip(("bond", "tri", "rec"), host, dev):
        assert _same(h, g), (name, h.shape, g.shape,
                             int(np.sum(h.view(np.uint32) != g.view(np.uint32)))
                             if h.shape == g.shape else None)
    return host


@pytest.mark.parametrize("name", ["si_diamond", "gaas", "si_type2", "hfo2", "nn2b"])
def test_golden_geometry(golden, name):
    print_case("geometry", name)
    bond, tri, rec = _check(golden(name))
    assert len(bond) > 0
    if name in ("hfo2", "nn2b"):
        assert len(rec) == 0


@pytest.mark.parametrize("cutoff,r2max,slots", [(3.0, 3.2, 4), (3.5, 3.5, 8), (3.9, 3.9, 16),
                                                (3.9, 3.2, 4)])
def test_synthetic_tersoff_geometry(tmp_path, cutoff, r2max, slots):
    """Records of 4, 8 and 16 bond slots; with R2 at most 3.2 A under a 3.9 A
    cut-off, the records hold only the bonds shorter than 3.2 A (the longer
    ones contribute exact zeros), on the device as on the host."""
    from gpuga_amd import synth
    print_case("geometry synthetic", cutoff, r2max)
    d = synth.tersoff_si(Nc=12, reps=(3, 3, 3), seed=4)
    d["cutoff"] = cutoff
    d["pmax"] = d["pmax"].copy()
    d["pmax"][-1] = r2max
    synth.write_dir(d, str(tmp_path))
    bond, tri, rec = _check(str(tmp_path))
    N = int(np.sum(d["Na"]))
    assert len(rec) == N * (1 + slots + slots * slots + slots // 2)
