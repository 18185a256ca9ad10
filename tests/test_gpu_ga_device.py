"""Device GA step (ga_device.hip) against the host GA (ga.cpp, itself pinned
bit-exact to the oracle and libstdc++ <random>): the same populations, sorted
fitness and random state, generation after generation.

The device step takes the random decisions the host draws for it (staged
draw records) and does the sort, crossover and mutation.  Shapes cover the
rank sort (P <= 512) and the bitonic sort (P > 512) with many equal
fitness values, the one-thread insertion sort when a NaN is present, a
population split over many workgroups (P = 256, V = 1153, the NN2B n = 32
layout: 7 rows per workgroup, 59034 mutations per early generation, many of
them on the same gene), and the largest population (4096)."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

_OPEN = []


def request_close(fit):
    _OPEN.append(fit)


@pytest.fixture(autouse=True, scope="module")
def _close_handles():
    yield
    while _OPEN:
        _OPEN.pop().close()


def fitness_seq(P, gens, seed, ties=False, nan_at=None):
    rng = np.random.default_rng(seed)
    out = []
    for g in range(gens):
        f = rng.random(P).astype(np.float32)
        if ties:
            f = np.round(f * 8) / 8  # many equal values: stability of the sort
        if nan_at is not None and g == nan_at:
            f[rng.integers(0, P, 3)] = np.nan
        out.append(f.astype(np.float32))
    return out


@pytest.mark.parametrize("P,parents,V,gens,ties,nan_at", [
    (20, 10, 9, 40, False, None),
    (200, 100, 9, 30, True, None),
    (64, 32, 141, 20, False, 3),
    (130, 64, 15, 20, True, 5),
    (256, 128, 1153, 6, False, None),
    (1000, 500, 9, 8, True, None),
    (600, 300, 15, 6, False, 2),
    (4096, 2048, 12, 3, True, None),
])
def test_device_step_equals_host_step(P, parents, V, gens, ties, nan_at):
    import gpuga_amd as G
    print_case("device GA step", P, parents, V)
    host = G.GA(P, parents, V, maximum_generation=gens, mutation_rate=0.2, seed=777)
    dev = G.GA(P, parents, V, maximum_generation=gens, mutation_rate=0.2, seed=777)
    assert np.array_equal(host.population, dev.population)
    for g, f in enumerate(fitness_seq(P, gens, 31 + P, ties, nan_at)):
        sh = host.step(g, f)
        sd = dev.step_device(g, f)
        assert np.array_equal(sh, sd, equal_nan=True), g
        assert np.array_equal(host.population, dev.population), g


@pytest.mark.parametrize("name,P,parents", [("si_diamond", 64, 32), ("hfo2", 40, 20),
                                            ("gaas", 200, 100)])
def test_run_ga_equals_host_loop(golden, name, P, parents):
    """gpuga_fitness_run_ga (population kernel + device GA step per
    generation, no host round trip) against compute + host GA step."""
    import gpuga_amd as G
    print_case("run_ga", name, P)
    fit = G.Fitness(golden(name))
    request_close(fit)
    gens = 12
    host = G.GA(P, parents, fit.V, maximum_generation=gens, mutation_rate=0.2, seed=4242)
    dev = G.GA(P, parents, fit.V, maximum_generation=gens, mutation_rate=0.2, seed=4242)
    best_h, elite_h = [], []
    for g in range(gens):
        f = fit.compute(host.population)
        s = host.step(g, f)
        best_h.append(s[0])
        elite_h.append(host.population[0].copy())  # row 0 is never crossed or mutated
    best_d, elite_d = fit.run_ga(dev, 0, gens)
    assert np.array_equal(np.array(best_h, np.float32), best_d)
    assert np.array_equal(np.array(elite_h), elite_d)
    assert np.array_equal(host.population, dev.population)
    # continuing from the returned state matches the host continuing too
    f = fit.compute(host.population)
    s = host.step(gens, f)
    b2, e2 = fit.run_ga(dev, gens, 1)
    assert b2[0] == s[0]
    assert np.array_equal(host.population, dev.population)


@pytest.mark.parametrize("batch,gens", [(0, 70), (5, 23)])
def test_run_ga_across_staging_batches(golden, batch, gens):
    """A run longer than one batch of staged draw records (32 generations
    by default; 5 forces batches that start on either population buffer)
    alternates population buffers, replays whole-batch graphs and
    single-generation graphs for the rest, and restages; it equals the host
    loop."""
    import gpuga_amd as G
    print_case("run_ga batches", "hfo2", 40, batch, gens)
    fit = G.Fitness(golden("hfo2"), ga_batch=batch)
    request_close(fit)
    host = G.GA(40, 20, fit.V, maximum_generation=gens, mutation_rate=0.2, seed=99)
    dev = G.GA(40, 20, fit.V, maximum_generation=gens, mutation_rate=0.2, seed=99)
    best_h = []
    for g in range(gens):
        s = host.step(g, fit.compute(host.population))
        best_h.append(s[0])
    best_d, _ = fit.run_ga(dev, 0, gens)
    assert np.array_equal(np.array(best_h, np.float32), best_d)
    assert np.array_equal(host.population, dev.population)


@pytest.mark.parametrize("P,parents,V,gens,ties", [(20, 10, 9, 30, False), (200, 100, 9, 20, True),
                                                   (256, 128, 1153, 4, False),
                                                   (1000, 500, 15, 5, True)])
def test_device_step_equals_oracle(P, parents, V, gens, ties):
    """The device GA step against the oracle's GA (oracle/ga.c, pinned to
    libstdc++ and the reference's expressions) directly."""
    import gpuga_amd as G
    from oracle import oracle as O
    print_case("device GA step vs oracle", P, parents, V)
    dev = G.GA(P, parents, V, maximum_generation=gens, mutation_rate=0.2, seed=31337)
    og = O.GA(P, parents, V, 31337, generations=gens, mutation_rate=0.2)
    assert np.array_equal(dev.population, og.population)
    for g, f in enumerate(fitness_seq(P, gens, 5 + P, ties)):
        s = dev.step_device(g, f)
        og.step(g, f)
        assert np.array_equal(s, og.fitness), g
        assert np.array_equal(dev.population, og.population), g


@pytest.mark.parametrize("name,P,parents", [("si_diamond", 200, 100), ("nn2b", 64, 32)])
def test_run_ga_equals_oracle_ga(golden, name, P, parents):
    """gpuga_fitness_run_ga against the oracle's GA stepped on the product's
    fitness of the oracle's population (the population launch is
    deterministic: the same genes give the same fitness bit for bit)."""
    import gpuga_amd as G
    from oracle import oracle as O
    print_case("run_ga vs oracle GA", name, P)
    fit = G.Fitness(golden(name))
    request_close(fit)
    gens = 10
    dev = G.GA(P, parents, fit.V, maximum_generation=gens, mutation_rate=0.2, seed=2718)
    og = O.GA(P, parents, fit.V, 2718, generations=gens, mutation_rate=0.2)
    best_o = []
    for g in range(gens):
        og.step(g, fit.compute(og.population))
        best_o.append(og.fitness[0])
    best_d, _ = fit.run_ga(dev, 0, gens)
    assert np.array_equal(np.array(best_o, np.float32), best_d)
    assert np.array_equal(dev.population, og.population)
