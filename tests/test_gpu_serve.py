"""Continuous-batching server and parameter sweeps: staggered arrivals,
admission by memory budget, concurrent size-class lanes; every served result
checked against the same request solved alone and against the oracle.
GPU only.

Bars: FP64 requests served in the bit-exact form equal the oracle exactly;
in the tolerance-bound form every field within 1e-10 of its own maximum (a
few dozen steps; the form's bar is 1e-12 after one step and 1e-9 after 100);
FP32-flux requests within 1e-5 relative L2 per field."""
import numpy as np
import pytest

from cbns3d_cart_mpi_amd import capi as C
from cbns3d_cart_mpi_amd import serve as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def oracle_solve(r):
    """The oracle's run of request r (src/cbns3d.cpp:176-270 on one rank)."""
    p, Q = S.prepare(r)
    b = O.Block(O.make_params(r.N, Mach=p.Mach, Re=p.Re))
    if Q is None:
        b.init_tgv()
    else:
        Qg = np.zeros((r.N + 8,) * 3 + (5,))
        Qg[4:-4, 4:-4, 4:-4] = Q
        b.set_state(Qg, 0.0, mu0=0.0)
    b.run(r.steps)
    return b.Q_interior(0), b.t_cur


def rel_l2(Q, R):
    return [float(np.sqrt(np.mean((Q[..., e] - R[..., e]) ** 2) / np.mean(R[..., e] ** 2))) for e in range(5)]


def rel_linf_own(Q, R):
    return [float(np.abs(Q[..., e] - R[..., e]).max() / np.abs(R[..., e]).max()) for e in range(5)]


def test_served_results_equal_solves_alone():
    reqs = S.make_requests(12, seed=5, grids=(16, 24, 32), steps=(3, 12), rate=400.0, precision="alt")
    # a budget for about three of the largest at once, so admission has to wait
    budget = 3 * C.batch_bytes_needed(C.make_params(32))
    out = S.serve(reqs, budget=budget, chunk=2, capacity=16, fp64_form=C.CBNS_FP64)
    assert out["cell_steps"] == sum(r.work for r in reqs)
    assert all(r.done >= r.admitted >= r.arrival for r in reqs)
    assert out["failed"] == 0
    S.sequential_fp64(reqs, precision=None)  # each request alone, in its own precision
    for r in reqs:
        assert np.array_equal(r.result, r.alone), r


@pytest.mark.parametrize("fp64_form", [C.CBNS_FP64, C.CBNS_FP64_FAST])
def test_served_results_against_the_oracle(fp64_form):
    """Every served request (TGV and turbulence, FP64 and FP32-flux, grids
    16-24^3, 8-24 steps, arriving while others run) against the oracle's run
    of the same request."""
    O.set_threads(16)
    reqs = S.make_requests(10, seed=31, grids=(16, 20, 24), steps=(8, 24), rate=300.0, precision="alt")
    out = S.serve(reqs, budget=64 << 30, chunk_ms=2.0, capacity=16, fp64_form=fp64_form)
    assert out["failed"] == 0
    for r in reqs:
        R, t = oracle_solve(r)
        if r.precision == C.CBNS_FP32_FLUX:
            err = rel_l2(r.result, R)
            assert max(err) <= 1e-5, (r, err)
            assert abs(r.t_final - t) <= 1e-7 * t
        elif fp64_form == C.CBNS_FP64:
            assert r.t_final == t and np.array_equal(r.result, R), r
        else:
            err = rel_linf_own(r.result, R)
            assert max(err) <= 1e-10, (r, err)
            assert abs(r.t_final - t) <= 1e-12 * t


def test_sweep_lanes_give_per_solve_latency_and_match_the_oracle():
    """A sweep of 32^3 and 48^3 solves (TGV and turbulence, 300 steps each)
    admitted at once into one lane per size: the lanes finish at their own
    times (per-solve completion times, p50 below p99), every FP64 solve
    equals its per-instance oracle run bit for bit and every FP32-flux solve
    is within 1e-5 relative L2 of it."""
    O.set_threads(16)
    reqs = S.make_requests(8, seed=77, grids=(32, 48), steps=(300, 300), precision="alt")
    assert {r.N for r in reqs} == {32, 48} and {r.ic for r in reqs} == {"tgv", "turb"}
    res = S.run_sweep_lanes(reqs, budget=64 << 30, collect=True, fp64_form=C.CBNS_FP64)
    lat = sorted(r.done for r in reqs)
    assert lat[0] < lat[-1] and res["p50_s"] < res["p99_s"]
    for r in reqs:
        R, t = oracle_solve(r)
        if r.precision == C.CBNS_FP64:
            assert r.t_final == t and np.array_equal(r.result, R), r
        else:
            assert max(rel_l2(r.result, R)) <= 1e-5, r


def test_sweep_runs_mixed_precisions_and_fields():
    reqs = S.make_requests(6, seed=9, grids=(16, 24), steps=(5, 5), precision="alt")
    res = S.run_sweep(reqs, budget=8 << 30)
    assert res["mcell_steps_per_s"] > 0
    assert all(np.isfinite(r.t_final) and r.t_final > 0 for r in reqs)


def test_batch_reuses_the_memory_of_removed_solves():
    """A removed solve's arrays are pooled and taken by the next solve of the
    same shape (no device allocation or release between chunks), and the
    pool is released on request."""
    p = C.make_params(24)
    with C.Batch(8 << 30, capacity=4) as b:
        i = b.add(p, steps=2)
        b.run(2)
        b.remove(i)
        assert b.pooled_bytes == C.batch_bytes_needed(p)
        j = b.add(p, steps=3)
        assert b.pooled_bytes == 0
        b.run(3)
        Q = b.get_state(j)
        b.remove(j)
        b.trim()
        assert b.pooled_bytes == 0
    with C.Solver(p) as s:
        s.init_tgv()
        s.run(3)
        assert np.array_equal(Q, s.get_state())
