"""GPU failure path of the population kernel: a force-assembly wait that gives
up.

The host forces the timeout (wait limit < 0: the assembly items report a
timeout without waiting), so every kernel still runs to its normal end.  The
launch's results must be discarded with an error, and the same handle must
afterwards give the oracle's fitness again: the plan's counters may not keep
counts of the failed launch (a late partial-force item can count its chunk
after the assembly items returned the counter to zero).

Tolerance: fitness 1e-5 relative to the FP64 oracle, bit-identical to the
handle's own result before the failure.
"""
import os

import numpy as np
import pytest

from conftest import print_case

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.mark.parametrize("P", [70, 200])
def test_forced_wait_timeout_then_recovers(golden, oracle_case, P):
    import gpuga_amd as G
    print_case("forced wait timeout", "si_diamond", P)
    orc = oracle_case("si_diamond")
    assert orc.ts.Nc_force > 0  # the case has force configurations (assembly items)
    genes = np.random.default_rng(11).random((P, orc.V), dtype=np.float32)
    with G.Fitness(golden("si_diamond")) as fit:
        before = fit.compute(genes)
        for _ in range(2):
            fit.set_wait_limit(-1)
            with pytest.raises(G.GpugaError, match="timed out"):
                fit.compute(genes)
        fit.set_wait_limit(1 << 23)
        for _ in range(3):
            after = fit.compute(genes)
            assert np.array_equal(after, before)
        ref, _ = orc.fitness(genes[:8], threads=THREADS)
        assert np.max(np.abs(after[:8] - ref) / np.abs(ref)) < 1e-5


def test_forced_wait_timeout_in_device_ga_run(golden):
    import gpuga_amd as G
    print_case("forced wait timeout in run_ga", "si_diamond")
    with G.Fitness(golden("si_diamond")) as fit:
        P = 64
        ga = G.GA(P, 32, fit.V, maximum_generation=20, mutation_rate=0.2, seed=5)
        ref = G.GA(P, 32, fit.V, maximum_generation=20, mutation_rate=0.2, seed=5)
        pop0 = ga.population
        fit.set_wait_limit(-1)
        with pytest.raises(G.GpugaError, match="timed out"):
            fit.run_ga(ga, 0, 4)
        # the GA handle is as before the failed call (population and stream)
        assert np.array_equal(ga.population, pop0)
        fit.set_wait_limit(1 << 23)
        best, _ = fit.run_ga(ga, 0, 4)
        ref_best, _ = fit.run_ga(ref, 0, 4)
        assert np.array_equal(best, ref_best)
        assert np.array_equal(ga.population, ref.population)
