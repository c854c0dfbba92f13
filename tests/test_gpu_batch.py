"""Batched form (persistent kernels over a packed block table, one hipGraph
per RK3 step): every solve equals its single-solve run bitwise, whatever it
is batched with.  GPU only."""
import numpy as np
import pytest

from cbns3d_cart_mpi_amd import capi as C

pytestmark = pytest.mark.gpu
GB = 1 << 30


def single(N, steps, dims=(1, 1, 1), precision=C.CBNS_FP64, **kw):
    with C.Solver(C.make_params(N, **kw), dims, precision=precision) as s:
        s.init_tgv()
        s.run(steps)
        return s.get_state(), s.t_cur


def test_batch_of_uneven_solves_matches_reference_and_single_runs(golden):
    g16 = golden("ref_tgv16.npz")
    gu = golden("ref_uneven.npz")
    gr = golden("ref_restart16.npz")
    with C.Batch(64 * GB) as b:
        ids = {
            "tgv16": b.add(C.make_params(16), steps=10),
            "uneven": b.add(C.make_params((26, 20, 16)), dims=(3, 2, 1), steps=int(gu["nsteps"])),
            "restart": b.add(C.make_params(16), Q=gr["Q_in"], t_cur=float(gr["t_in"]), mu0=0.0,
                             steps=int(gr["nsteps"])),
            "fp32": b.add(C.make_params(24, Mach=0.3, Re=400.0), dims=(1, 2, 1), precision=C.CBNS_FP32_FLUX, steps=4),
            "odd": b.add(C.make_params((21, 16, 12), Mach=0.2, Re=800.0), steps=7),
        }
        b.run(12)
        done, live = b.poll()
        assert live == 5 and sorted(done) == sorted(ids.values())
        st = {k: b.status(v) for k, v in ids.items()}
        assert st["tgv16"]["steps_done"] == 10 and st["odd"]["steps_done"] == 7
        assert st["tgv16"]["t_cur"] == float(g16["t_n10"])
        assert np.array_equal(b.get_state(ids["tgv16"]), g16["Q_n10"])
        assert st["uneven"]["t_cur"] == float(gu["t"])
        assert np.array_equal(b.get_state(ids["uneven"]), gu["Q"])
        assert st["restart"]["t_cur"] == float(gr["t_out"])
        assert np.array_equal(b.get_state(ids["restart"]), gr["Q_out"])
        Q, t = single(24, 4, (1, 2, 1), C.CBNS_FP32_FLUX, Mach=0.3, Re=400.0)
        assert st["fp32"]["t_cur"] == t and np.array_equal(b.get_state(ids["fp32"]), Q)
        Q, t = single((21, 16, 12), 7, Mach=0.2, Re=800.0)  # odd x-width: per-cell stage tail
        assert st["odd"]["t_cur"] == t and np.array_equal(b.get_state(ids["odd"]), Q)


def test_alone_equals_in_mix():
    p = C.make_params(32, Mach=0.4, Re=200.0)
    with C.Batch(16 * GB) as alone:
        a = alone.add(p, precision=C.CBNS_FP32_FLUX, steps=6)
        alone.run(6)
        Qa = alone.get_state(a)
    with C.Batch(32 * GB) as mix:
        mix.add(C.make_params(48), steps=9)
        x = mix.add(p, precision=C.CBNS_FP32_FLUX, steps=6)
        mix.add(C.make_params((24, 32, 16)), dims=(2, 2, 1), precision=C.CBNS_FP32_FLUX, steps=3)
        mix.add(C.make_params(16, Mach=0.5), steps=20)
        mix.run(4)
        mix.run(3)  # replays of the same captured step graph
        assert np.array_equal(mix.get_state(x), Qa)


@pytest.mark.parametrize("precision", [C.CBNS_FP64, C.CBNS_FP32_FLUX])
def test_odd_plane_viscous_chunks_equal_single_runs(precision):
    """z-extents whose viscous march leaves a chunk with an odd number of
    planes (n2 + 4 = 25: chunks of 16 and 9; 29: 16 and 13), so a persistent
    workgroup's next unit refills the LDS buffer its last plane read."""
    # the last case has an odd x-width: per-cell stage tail, dt from kb_dt every step
    cases = [((20, 16, 21), (1, 1, 1), 5), ((18, 22, 25), (1, 2, 1), 4), ((16, 16, 16), (1, 1, 1), 6),
             ((21, 18, 16), (1, 1, 1), 4)]
    with C.Batch(32 * GB) as b:
        ids = [b.add(C.make_params(N, Mach=0.3, Re=500.0), dims=d, precision=precision, steps=s) for N, d, s in cases]
        b.run(6)
        for (N, d, s), i in zip(cases, ids):
            Q, t = single(N, s, d, precision, Mach=0.3, Re=500.0)
            assert b.status(i)["t_cur"] == t
            assert np.array_equal(b.get_state(i), Q), N


def test_viscous_overlap_and_grids_change_nothing():
    """The step graph's schedule (viscous kernels in line or beside the face
    sweeps, with any grid), the memory-bound kernels' grid and the band of the
    work order leave every solve's values unchanged."""
    cases = [((24, 20, 16), (1, 1, 1), C.CBNS_FP64), ((20, 16, 21), (2, 1, 1), C.CBNS_FP32_FLUX),
             ((32, 32, 32), (1, 1, 1), C.CBNS_FP32_FLUX)]
    out = []
    for ov, wg, band in ((0, 0, None), (64, 0, 7), (None, 96, 64), (300, 5, 1)):
        with C.Batch(32 * GB) as b:
            if ov is not None:
                b.set_overlap(ov)
            if band is not None:
                b.set_band(band)
            b.set_workgroups(wg)
            ids = [b.add(C.make_params(N, Mach=0.3, Re=700.0), dims=d, precision=p, steps=5) for N, d, p in cases]
            b.run(5)
            out.append([(b.status(i)["t_cur"], b.get_state(i)) for i in ids])
    for other in out[1:]:
        for (t0, q0), (t1, q1) in zip(out[0], other):
            assert t0 == t1 and np.array_equal(q0, q1)


def test_admission_by_memory_budget_and_reuse_of_slots():
    p = C.make_params(32)
    need = C.batch_bytes_needed(p)
    with C.Batch(2 * need + need // 2, capacity=8) as b:
        i0 = b.add(p, steps=2)
        i1 = b.add(p, steps=2)
        with pytest.raises(C.CbnsError) as e:
            b.add(p, steps=2)
        assert e.value.status == -3
        assert b.memory[0] == 2 * need
        b.run(3)
        Q1 = b.get_state(i1)
        b.remove(i0)
        i2 = b.add(p, steps=2)
        b.run(3)
        assert np.array_equal(b.get_state(i2), Q1)  # same inputs, same result in a reused slot


def test_finished_solves_stay_unchanged():
    with C.Batch(8 * GB) as b:
        i = b.add(C.make_params(16), steps=2)
        b.run(2)
        Q = b.get_state(i)
        b.run(5)
        assert np.array_equal(b.get_state(i), Q)
        assert b.status(i)["steps_done"] == 2 and not b.status(i)["active"]


def test_step_budgets_pause_and_resume_without_changing_results():
    p = C.make_params(16)
    with C.Batch(8 * GB) as b:
        i = b.add(p, steps=6)
        j = b.add(C.make_params(24), steps=6)
        b.set_budgets({i: 2, j: 5})
        b.run(5)
        assert b.status(i)["steps_done"] == 2 and b.status(j)["steps_done"] == 5
        b.set_budgets({i: 4, j: 1})
        b.run(4)
        assert b.status(i)["steps_done"] == 6 and b.status(j)["steps_done"] == 6
        done, _ = b.poll()
        assert sorted(done) == sorted([i, j])
        Qi, Qj = b.get_state(i), b.get_state(j)
    Q, _ = single(16, 6)
    assert np.array_equal(Qi, Q)
    Q, _ = single(24, 6)
    assert np.array_equal(Qj, Q)


def test_scaled_solve_recomputes_faces_as_single_run():
    """A Taylor-Green state scaled by 1e31 (many faces leave the range the
    face kernels' unscaled quotients are checked for and are recomputed with
    the IEEE forms) beside an ordinary solve: both equal their single runs."""
    N = (20, 16, 12)
    with C.Solver(C.make_params(N)) as s:
        s.init_tgv()
        Q0 = s.get_state() * 1e31
        s.set_state(Q0, 0.0, mu0=0.0)
        s.run(3)
        Qs, ts = s.get_state(), s.t_cur
    with C.Batch(16 * GB) as b:
        a = b.add(C.make_params(N), Q=Q0, t_cur=0.0, mu0=0.0, steps=3)
        o = b.add(C.make_params(24), steps=3)
        b.run(3)
        assert b.status(a)["t_cur"] == ts and np.array_equal(b.get_state(a), Qs)
        Q, t = single(24, 3)
        assert b.status(o)["t_cur"] == t and np.array_equal(b.get_state(o), Q)


@pytest.mark.parametrize("precision", [C.CBNS_FP64, C.CBNS_FP32_FLUX])
def test_minimum_blocks_in_batch_equal_single_runs(precision):
    """A solve of 24 minimum-size blocks (4 cells per axis) and one of 4- and
    5-cell odd blocks, batched together: each equals its single run."""
    cases = [((12, 8, 16), (3, 2, 4), 3), ((13, 9, 10), (3, 2, 2), 4)]
    with C.Batch(16 * GB) as b:
        ids = [b.add(C.make_params(N, Mach=0.3, Re=600.0), dims=d, precision=precision, steps=s)
               for N, d, s in cases]
        b.run(4)
        for (N, d, s), i in zip(cases, ids):
            Q, t = single(N, s, d, precision, Mach=0.3, Re=600.0)
            assert b.status(i)["t_cur"] == t
            assert np.array_equal(b.get_state(i), Q), N


def test_non_finite_solve_stops_alone():
    """A solve that turns non-finite in its first step (a Taylor-Green state
    scaled by 1e-200) is flagged and deactivated by the on-device dt
    bookkeeping; the ordinary solve beside it runs on and equals its single
    run."""
    N = (20, 16, 12)
    with C.Solver(C.make_params(N)) as s:
        s.init_tgv()
        Q0 = s.get_state() * 1e-200
    with C.Batch(16 * GB) as b:
        bad = b.add(C.make_params(N), Q=Q0, t_cur=0.0, mu0=0.0, steps=5)
        good = b.add(C.make_params(24), steps=5)
        b.run(5)
        st = b.status(bad)
        assert st["nonfinite"] and not st["active"] and st["steps_done"] < 5
        Q, t = single(24, 5)
        assert b.status(good)["t_cur"] == t and np.array_equal(b.get_state(good), Q)


def test_tolerance_bound_fp64_solves_in_a_mix_equal_their_single_runs():
    """CBNS_FP64_FAST solves batched with FP64 and FP32-flux solves give their
    single-solve results bit for bit (the same sweep bodies, contraction fixed
    per function), including a solve scaled by 1e50 whose every cell goes
    through the batch's cell fixup, an odd x-width and a split solve."""
    from oracle import oracle as O

    ob = O.Block(O.make_params(12))
    ob.init_tgv()
    Qs = ob.Q_interior(0) * 1e50
    cases = {
        "fast": (C.make_params(32, Mach=0.3, Re=400.0), (1, 1, 1), None, 6),
        "fast_split": (C.make_params((24, 20, 16), Mach=0.2, Re=900.0), (2, 1, 2), None, 5),
        "fast_odd": (C.make_params((21, 16, 12), Mach=0.4, Re=300.0), (1, 1, 1), None, 4),
        "fast_scaled": (C.make_params(12), (1, 1, 1), Qs, 3),
    }
    with C.Batch(32 * GB) as b:
        b.add(C.make_params(24), steps=8)
        b.add(C.make_params(16, Mach=0.5), precision=C.CBNS_FP32_FLUX, steps=8)
        ids = {k: b.add(p, dims=d, precision=C.CBNS_FP64_FAST, Q=Q, steps=n) for k, (p, d, Q, n) in cases.items()}
        b.run(8)
        for k, (p, d, Q, n) in cases.items():
            with C.Solver(p, d, precision=C.CBNS_FP64_FAST) as s:
                if Q is None:
                    s.init_tgv()
                else:
                    s.set_state(Q, 0.0, mu0=0.0)
                s.run(n)
                st = b.status(ids[k])
                assert st["steps_done"] == n and st["t_cur"] == s.t_cur, k
                got = b.get_state(ids[k])
                assert np.isfinite(got).all(), k
                assert np.array_equal(got, s.get_state()), k
