"""GPU parity: the HIP library (through its C ABI) against the FP64 oracle.

Tolerances (FP32 pair/triplet arithmetic, FP64 per-configuration sums):
  fitness and its three terms     1e-5 relative     (Tersoff, RI, NN2B-FP32)
  NN2B with split-BF16 MFMA hidden layer 1e-4 relative (stated BF16 error)
  per-atom force / virial / energy 2e-5 of the case's largest magnitude
                                   (per-atom FP32 accumulation of bond terms)
"""
import os

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def G():
    import gpuga_amd
    return gpuga_amd


_FIT = {}


@pytest.fixture(autouse=True, scope="module")
def _close_cached():
    """The module's cached handles are closed when its tests are done."""
    yield
    while _FIT:
        _FIT.popitem()[1].close()


def product(G, golden, name, **opts):
    key = (name, tuple(sorted(opts.items())))
    if key not in _FIT:
        _FIT[key] = G.Fitness(golden(name), **opts)
    return _FIT[key]


def genes_for(V, P, seed):
    return np.random.default_rng(seed).random((P, V), dtype=np.float32)


@pytest.mark.parametrize("name", ["si_diamond", "gaas", "si_type2", "nn2b", "hfo2"])
def test_find_force_matches_oracle(G, golden, oracle_case, name):
    print_case("find_force", name)
    orc = oracle_case(name)
    fit = product(G, golden, name)
    pmin, pmax = orc.bounds()
    g = genes_for(orc.V, 1, 3)[0]
    params = orc.denormalise(g)
    f, w, e = fit.find_force(params)
    fo, wo, eo = orc.find_force(params)
    for got, ref, what in ((f, fo, "force"), (w, wo, "virial"), (e, eo, "energy")):
        scale = max(np.abs(ref).max(), 1e-3)
        err = np.abs(got.astype(np.float64) - ref).max() / scale
        assert err < 2e-5, f"{name} {what}: max error {err:.3e} of scale {scale:.3e}"


@pytest.mark.parametrize("name,P", [("si_diamond", 70), ("gaas", 20), ("si_type2", 20),
                                    ("nn2b", 64)])
def test_population_fitness_matches_oracle(G, golden, oracle_case, name, P):
    print_case("population", name, P)
    orc = oracle_case(name)
    fit = product(G, golden, name, nn2b_bf16=0)
    genes = genes_for(orc.V, P, 7)
    got, terms = fit.compute(genes, terms=True)
    ref, rterms = orc.fitness(genes, threads=THREADS)
    rel = np.abs(got - ref) / np.abs(ref)
    assert rel.max() < 1e-5, f"{name}: fitness rel error {rel.max():.3e}"
    trel = np.abs(terms - rterms) / np.maximum(np.abs(rterms), 1e-30)
    assert trel.max() < 1e-5, f"{name}: term rel error {trel.max():.3e}"


def test_population_rigid_ion_matches_oracle(G, golden, oracle_case):
    print_case("population", "hfo2")
    orc = oracle_case("hfo2")
    fit = product(G, golden, "hfo2")
    genes = genes_for(orc.V, 3, 9)
    got, terms = fit.compute(genes, terms=True)
    ref, rterms = orc.fitness(genes, threads=THREADS)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-5
    np.testing.assert_allclose(terms, rterms, rtol=1e-5)


def test_nn2b_bf16_error_is_stated_bound(G, golden, oracle_case):
    print_case("population", "nn2b bf16")
    orc = oracle_case("nn2b")
    fit = product(G, golden, "nn2b", nn2b_bf16=1)
    genes = genes_for(orc.V, 64, 5)
    got = fit.compute(genes)
    ref, _ = orc.fitness(genes, threads=THREADS)
    rel = np.abs(got - ref) / np.abs(ref)
    print("nn2b bf16 fitness rel error: max %.3e median %.3e" % (rel.max(), np.median(rel)))
    assert rel.max() < 1e-4


@pytest.mark.parametrize("name", ["si_diamond", "hfo2", "nn2b"])
def test_population_equals_per_individual_path(G, golden, name):
    print_case("plain-vs-population", name)
    fit = product(G, golden, name, nn2b_bf16=0)
    genes = genes_for(fit.V, 10, 13)
    a, ta = fit.compute(genes, terms=True)
    b, tb = fit.compute(genes, terms=True, plain=True)
    np.testing.assert_allclose(ta, tb, rtol=2e-5)
    np.testing.assert_allclose(a, b, rtol=2e-5)


@pytest.mark.parametrize("P", [1, 63, 64, 65, 130])
def test_population_sizes(G, golden, oracle_case, P):
    print_case("population size", P)
    orc = oracle_case("si_diamond")
    fit = product(G, golden, "si_diamond")
    genes = genes_for(orc.V, P, 100 + P)
    got = fit.compute(genes)
    sel = np.unique(np.linspace(0, P - 1, min(P, 6)).astype(int))
    ref, _ = orc.fitness(genes[sel], threads=THREADS)
    np.testing.assert_allclose(got[sel], ref, rtol=1e-5)


@pytest.mark.parametrize("name,P", [("si_diamond", 72), ("si_diamond", 65), ("si_diamond", 8),
                                    ("gaas", 76), ("si_type2", 130)])
def test_packed_lane_block_matches_unpacked(G, golden, name, P):
    """A last lane block of <= 16 individuals runs several atoms per wave;
    its fitness and terms must equal the one-atom-per-wave form (same FP32
    formulas, FP64 sums regrouped), for the individuals of that block and the
    full blocks alike."""
    print_case("packed lanes", name, P)
    packed = product(G, golden, name)
    plain = product(G, golden, name, pack_lanes=0)
    genes = genes_for(packed.V, P, 300 + P)
    a, ta = packed.compute(genes, terms=True)
    b, tb = plain.compute(genes, terms=True)
    np.testing.assert_allclose(ta, tb, rtol=1e-5)
    np.testing.assert_allclose(a, b, rtol=1e-5)
    again, tag = packed.compute(genes, terms=True)
    assert np.array_equal(ta, tag)


def test_replay_is_bit_identical(G, golden):
    print_case("graph replay")
    fit = product(G, golden, "si_diamond")
    genes = genes_for(fit.V, 200, 21)
    first, t1 = fit.compute(genes, terms=True)
    for _ in range(4):
        again, t2 = fit.compute(genes, terms=True)
        assert np.array_equal(first, again)
        assert np.array_equal(t1, t2)
    with G.Fitness(golden("si_diamond"), use_graph=0, chunk_cost=fit_chunk(fit, 200)) as eager:
        e, te = eager.compute(genes, terms=True)
    np.testing.assert_allclose(te, t1, rtol=1e-12)


def fit_chunk(fit, P):
    return fit.autotune(P)


def test_chunking_does_not_change_results(G, golden):
    print_case("chunking")
    genes = genes_for(9, 100, 33)
    with G.Fitness(golden("si_diamond"), chunk_cost=64) as small, \
            G.Fitness(golden("si_diamond"), chunk_cost=10 ** 8) as whole:
        a, ta = small.compute(genes, terms=True)
        b, tb = whole.compute(genes, terms=True)
    np.testing.assert_allclose(ta, tb, rtol=1e-12)


@pytest.mark.parametrize("name", ["si_diamond", "hfo2"])
def test_launch_shape_does_not_change_results(G, golden, name):
    """Queue count and resident workgroups per CU only change which wave
    takes which item; every fold has a fixed order, so the results are the
    same bits.  The autotuned shape, given back as options, reproduces the
    tuned handle's results bit for bit."""
    print_case("launch shape", name)
    P = 130
    with G.Fitness(golden(name), chunk_cost=200) as base:
        genes = genes_for(base.V, P, 91)
        ref, tref = base.compute(genes, terms=True)
    for opts in ({"queues": 16}, {"queues": 64}, {"queues": 3}, {"blocks_per_cu": 1},
                 {"blocks_per_cu": 2, "queues": 64}):
        with G.Fitness(golden(name), chunk_cost=200, **opts) as f:
            got, tgot = f.compute(genes, terms=True)
            assert f.launch_shape(P)["chunk_cost"] == 200
        assert np.array_equal(tgot, tref), opts
        assert np.array_equal(got, ref), opts
    with G.Fitness(golden(name)) as tuned:
        sh = tuned.launch_shape(P)
        assert sh["chunk_cost"] >= 64 and 1 <= sh["queues"] <= 64 and sh["blocks_per_cu"] >= 1
        a, ta = tuned.compute(genes, terms=True)
    with G.Fitness(golden(name), chunk_cost=sh["chunk_cost"], queues=sh["queues"],
                   blocks_per_cu=sh["blocks_per_cu"], pack_lanes=sh["packed"]) as fixed:
        b, tb = fixed.compute(genes, terms=True)
    assert np.array_equal(ta, tb)
    # (another item size regroups the FP64 sums, and packing the last lane
    # block sums its lane groups by a butterfly: the packed-lane bound)
    np.testing.assert_allclose(ta, tref, rtol=1e-5)


@pytest.mark.parametrize("name,P", [("si_diamond", 128), ("gaas", 128), ("si_diamond", 50)])
def test_lone_wave_form_same_bits(G, golden, name, P):
    """A grid of at most one workgroup per CU runs the four-bond Tersoff
    records with the lone-wave instantiation (each next atom's record staged
    through LDS; with one lane block, P = 50, the individuals' constants
    loaded during the first fetch), four workgroups per CU with the default
    one: the same arithmetic on the same values, so the same bits."""
    print_case("lone-wave form", name, P)
    genes = genes_for(9 if name == "si_diamond" else 13, P, 97)
    out = {}
    for bpc in (1, 4):
        with G.Fitness(golden(name), chunk_cost=200, blocks_per_cu=bpc) as f:
            out[bpc] = f.compute(genes, terms=True)
    assert np.array_equal(out[1][1], out[4][1])
    assert np.array_equal(out[1][0], out[4][0])


@pytest.mark.parametrize("name", ["si_diamond", "gaas", "hfo2"])
def test_cross_wave_handoff_stress(G, golden, name):
    """Tiny work items split every configuration into many chunks, so chunk
    partials, per-bond forces and error terms pass between waves on different
    XCDs many times per launch; ten launches must reproduce the single-item
    result (FP64 partial sums regrouped: 1e-12)."""
    print_case("handoff stress", name)
    with G.Fitness(golden(name), chunk_cost=10 ** 8) as whole, \
            G.Fitness(golden(name), chunk_cost=24, use_graph=1) as tiny:
        genes = genes_for(whole.V, 130, 77)
        ref_fit, ref_terms = whole.compute(genes, terms=True)
        first = None
        for rep in range(10):
            f, t = tiny.compute(genes, terms=True)
            np.testing.assert_allclose(t, ref_terms, rtol=1e-12)
            if first is None:
                first = t
            assert np.array_equal(t, first)
