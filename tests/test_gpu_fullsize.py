"""The headline workload at its full size (population 200, 1000 diamond-Si
configurations of 8-512 atoms, 120 200 atoms): the population launch
against the FP64 oracle for a sample of individuals (the oracle evaluates
one individual over the whole set in about 0.1 s), against the
per-individual launches for another sample, and bit-identical over
repeated launches and over the GA run's generations (replayed graphs).
Tolerance: fitness and terms 1e-5 relative (FP32 bond arithmetic, FP64
sums)."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def headline():
    import gpuga_amd
    from gpuga_amd import synth
    d = synth.tersoff_si(Nc=1000, sizes="loguniform", seed=2)
    lab = gpuga_amd.Fitness(arrays=d)
    d = synth.label(d, lab.find_force)
    lab.close()
    fit = gpuga_amd.Fitness(arrays=d)
    yield d, fit
    fit.close()


def test_full_size_population_matches_oracle(headline):
    from oracle import oracle as O
    d, fit = headline
    print_case("full size vs oracle", fit.N)
    assert fit.N == 120200 and fit.Nc == 1000
    genes = np.random.default_rng(77).random((200, fit.V), dtype=np.float32)
    f, terms = fit.compute(genes, terms=True)
    case = O.Case(potential=dict(potential_type=1, cutoff=3.0, weights=(0.05, 0.15, 0.8),
                                 pmin=d["pmin"], pmax=d["pmax"]), arrays=d)
    sel = [0, 63, 64, 127, 191, 192, 199]  # every lane block, the packed one included
    ref, ref_terms = case.fitness(genes[sel], threads=16)
    np.testing.assert_allclose(terms[sel], ref_terms, rtol=1e-5)
    np.testing.assert_allclose(f[sel], ref, rtol=1e-5)
    again = fit.compute(genes)
    assert np.array_equal(f, again)


def test_full_size_population_equals_per_individual_launches(headline):
    _, fit = headline
    print_case("full size vs per-individual")
    genes = np.random.default_rng(78).random((200, fit.V), dtype=np.float32)
    f, terms = fit.compute(genes, terms=True)
    sel = [5, 100, 198]
    fp, tp = fit.compute(genes[sel], terms=True, plain=True)
    np.testing.assert_allclose(tp, terms[sel], rtol=1e-5)
    np.testing.assert_allclose(fp, f[sel], rtol=1e-5)


def test_full_size_ga_run_equals_host_loop(headline):
    import gpuga_amd as G
    _, fit = headline
    print_case("full size run_ga vs host loop")
    gens = 6
    host = G.GA(200, 100, fit.V, maximum_generation=gens, mutation_rate=0.2, seed=12345678)
    dev = G.GA(200, 100, fit.V, maximum_generation=gens, mutation_rate=0.2, seed=12345678)
    best_h = [host.step(g, fit.compute(host.population))[0] for g in range(gens)]
    best_d, _ = fit.run_ga(dev, 0, gens)
    assert np.array_equal(np.array(best_h, np.float32), best_d)
    assert np.array_equal(host.population, dev.population)


def _labelled(make):
    import gpuga_amd
    from gpuga_amd import synth
    d = make()
    lab = gpuga_amd.Fitness(arrays=d)
    d = synth.label(d, lab.find_force)
    lab.close()
    return d


def _oracle_case(d):
    from oracle import oracle as O
    pot = dict(potential_type=d["potential_type"], cutoff=d["cutoff"], weights=d["weights"],
               pmin=d["pmin"], pmax=d["pmax"], num_neurons=d.get("num_neurons", 0))
    return O.Case(potential=pot, arrays=d)


def test_full_size_rigid_ion_matches_oracle():
    """The rigid-ion fluorite workload at full size (500 x 96 atoms, about 95
    neighbours, P = 200, 15 parameters): individuals of every lane block
    against the oracle, tolerance 1e-5 relative."""
    import gpuga_amd
    from gpuga_amd import synth
    print_case("full size rigid ion vs oracle")
    d = _labelled(lambda: synth.fluorite_ab2(Nc=500, seed=3))
    fit = gpuga_amd.Fitness(arrays=d)
    try:
        genes = np.random.default_rng(79).random((200, fit.V), dtype=np.float32)
        f, terms = fit.compute(genes, terms=True)
    finally:
        fit.close()
    sel = [0, 64, 130, 199]
    ref, ref_terms = _oracle_case(d).fitness(genes[sel], threads=16)
    np.testing.assert_allclose(terms[sel], ref_terms, rtol=1e-5)
    np.testing.assert_allclose(f[sel], ref, rtol=1e-5)


@pytest.mark.parametrize("bf16,tol", [(0, 1e-5), (1, 1e-4)])
def test_full_size_nn2b_matches_oracle(bf16, tol):
    """The NN2B wurtzite workload at full size (800 x 72 atoms, n = 32,
    1153 parameters, P = 256): one individual of every lane block against the
    oracle, fitness and each of the three terms (force included); FP32
    tabulation within 1e-5 relative, the split-BF16 hidden layer within its
    stated 1e-4."""
    import gpuga_amd
    from gpuga_amd import synth
    print_case("full size nn2b vs oracle", bf16)
    d = _labelled(lambda: synth.wurtzite_ab(Nc=800, seed=5))
    fit = gpuga_amd.Fitness(arrays=d, nn2b_bf16=bf16)
    try:
        genes = np.random.default_rng(80).random((256, fit.V), dtype=np.float32)
        f, terms = fit.compute(genes, terms=True)
    finally:
        fit.close()
    sel = [3, 70, 140, 255]  # lane blocks 0, 1, 2, 3
    ref, ref_terms = _oracle_case(d).fitness(genes[sel], threads=16)
    rel = np.abs(f[sel] - ref) / np.abs(ref)
    trel = np.abs(terms[sel] - ref_terms) / np.abs(ref_terms)
    print("nn2b full size, bf16=%d: fitness rel error max %.3e, terms (E, W, F) %s"
          % (bf16, rel.max(), np.array2string(trel.max(axis=0), precision=3)))
    assert rel.max() < tol
    assert trel.max() < tol
