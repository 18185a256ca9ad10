"""Neighbour list built on the device (neighbor.hip; the set-up of every
Fitness) against the oracle's list, itself pinned bit for bit to the
reference fixtures (test_host.py, test_oracle_golden.py), and against the
host build: identical CSR lists on every golden case, on triclinic cells
(a sheared box folds many pairs through the minimum image), and on a
100k-atom set."""
import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import gpuga_amd
    return gpuga_amd


@pytest.mark.parametrize("name", ["si_diamond", "gaas", "si_type2", "nn2b", "hfo2"])
def test_device_list_equals_oracle(G, golden, oracle_case, name):
    print_case("device neighbour list", name)
    with G.Fitness(golden(name)) as fit:
        NN, start, NL = fit.neighbors()
    a, b, c = oracle_case(name).neighbors()
    assert np.array_equal(NN, a) and np.array_equal(start, b) and np.array_equal(NL, c)


def triclinic_set(seed, Nc=40):
    """Random atoms (about Si density) in sheared cells of 8-96 atoms."""
    from gpuga_amd import synth
    rng = np.random.default_rng(seed)
    cells = []
    for _ in range(Nc):
        n = int(rng.integers(8, 97))
        L = (n / 0.05) ** (1 / 3)
        box = np.diag(L * rng.uniform(0.9, 1.1, 3))
        box[1, 0] = L * rng.uniform(-0.45, 0.45)  # b has an a component
        box[2, 0] = L * rng.uniform(-0.45, 0.45)
        box[2, 1] = L * rng.uniform(-0.45, 0.45)
        frac = rng.random((n, 3))
        cells.append((box, frac @ box, np.zeros(n, np.int32)))
    d = synth._assemble(cells, 0.2, rng)
    d.update(potential_type=1, cutoff=3.0, weights=(0.05, 0.15, 0.8),
             pmin=synth.TERSOFF_SI_MIN, pmax=synth.TERSOFF_SI_MAX)
    return d


@pytest.mark.parametrize("seed", [1, 2])
def test_device_list_triclinic(G, seed):
    from oracle import oracle as O
    print_case("device neighbour list, triclinic", seed)
    d = triclinic_set(seed)
    with G.Fitness(arrays=d) as fit:
        NN, start, NL = fit.neighbors()
    case = O.Case(potential=dict(potential_type=1, cutoff=3.0, weights=(0.05, 0.15, 0.8),
                                 pmin=d["pmin"], pmax=d["pmax"]), arrays=d)
    a, b, c = case.neighbors()
    assert NN.sum() > 0
    assert np.array_equal(NN, a) and np.array_equal(start, b) and np.array_equal(NL, c)


def test_device_list_large_set_equals_host_build(G, tmp_path):
    from gpuga_amd import synth
    print_case("device neighbour list, large set")
    d = synth.tersoff_si(Nc=800, sizes="loguniform", seed=11)
    synth.write_dir(d, str(tmp_path), ga={"maximum_generation": 2, "population_size": 20,
                                          "parent_number": 10, "mutation_rate": 0.2})
    with G.Fitness(str(tmp_path)) as fit:
        NN, start, NL = fit.neighbors()
    a, b, c = G.inspect_neighbors(str(tmp_path))
    assert len(NN) > 90000
    assert np.array_equal(NN, a) and np.array_equal(start, b) and np.array_equal(NL, c)
