"""Continuous-batching server: staggered arrivals, admission by memory
budget, every result equal to the same request solved alone.  GPU only."""
import numpy as np
import pytest

from cbns3d_cart_mpi_amd import capi as C
from cbns3d_cart_mpi_amd import serve as S

pytestmark = pytest.mark.gpu


def test_served_results_equal_solves_alone():
    reqs = S.make_requests(12, seed=5, grids=(16, 24, 32), steps=(3, 12), rate=400.0, precision="alt")
    # a budget for about three of the largest at once, so admission has to wait
    budget = 3 * C.batch_bytes_needed(C.make_params(32))
    out = S.serve(reqs, budget=budget, chunk=2, capacity=16)
    assert out["cell_steps"] == sum(r.work for r in reqs)
    assert all(r.done >= r.admitted >= r.arrival for r in reqs)
    S.sequential_fp64(reqs, precision=None)  # each request alone, in its own precision
    for r in reqs:
        assert np.array_equal(r.result, r.alone), r


def test_sweep_runs_mixed_precisions_and_fields():
    reqs = S.make_requests(6, seed=9, grids=(16, 24), steps=(5, 5), precision="alt")
    res = S.run_sweep(reqs, budget=8 << 30)
    assert res["mcell_steps_per_s"] > 0
    assert all(np.isfinite(r.t_final) and r.t_final > 0 for r in reqs)
