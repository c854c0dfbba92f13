"""Benchmark of the FP64 RHS + TVD-RK3 time-step path on one MI355X.

Workload: Taylor-Green vortex, 256^3 periodic cells, FP64, one on-device block
(the reference's TGV setup: gamma 1.4, Mach 0.1, Re 1600, Pr 0.71,
T_inf 255.556, CFL 0.5).  A step is one full TVD-RK3 step (dt reduction, three
RHS evaluations, three ghost fills) over the whole grid, with the state
resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--dims 1,1,1]

Prints one JSON line.  value = Mcell-steps/s = N^3 * K / wall / 1e6 over the
K timed steps.  The dominant kernel (the inviscid face fluxes: in the
headline tolerance-bound form one march that forms the x and y faces of
each stage, in the bit-exact form an x and a y sweep; the z faces are formed
by the stage tail, which fuses them with the update) is timed live with HIP
events on the solver's stream during the timed steps; its roofline is FP64
arithmetic
(algorithmic 1570 flop per face: calc_eigenvectors 161 + 2 x rec_weno5 642 +
lf_flux 125, counted literally from the reference source).  The fused z-face
and stage-tail kernel is reported beside it against both FP64 arithmetic and
HBM bandwidth.  The CPU baseline is the repository's CPU
restatement of the reference (oracle/, OpenMP) on the host's cores, timed on a
bounded sample of steps of the same 256^3 workload.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_FACE = 1570          # literal count of the reference's per-face arithmetic
FLOP_PER_CELL_STEP = 15145    # literal count per cell-step, all kernels
PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 vector (and matrix) peak
PEAK_HBM_GBS = 8000.0
# Every timed launch adds two events to the stream (about 2.5% of a step when
# all kernels are timed), so the timed region times only the roofline kernel
# and the per-kernel breakdown comes from this many further steps.
BREAKDOWN_STEPS = 10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--dims", default="1,1,1")
    ap.add_argument("--band", default="auto", help="rows per work band, or 'auto' to tune")
    ap.add_argument("--schedule", default="auto",
                    help="stage schedule: overlap | overlap:N (N viscous workgroups) | serial, or 'auto' to tune")
    ap.add_argument("--shape", default="32,32",
                    help="viscous planes per chunk and cells per run, with given --band/--schedule/--occ")
    ap.add_argument("--occ", default="auto", choices=["auto", "2", "3"],
                    help="face-kernel waves per SIMD of the pair sweeps, or 'auto' to tune")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-fma", action="store_true", help="skip the contracted FP64 form (default: skipped)",
                    default=True)
    ap.add_argument("--fma", dest="no_fma", action="store_false", help="also time the contracted FP64 form")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-turb", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-io", action="store_true")
    ap.add_argument("--io-only", action="store_true",
                    help="run only the output/diagnostics leg (for profiling it) and print it")
    ap.add_argument("--batched-steps", type=int, default=60)
    ap.add_argument("--batched-copies", type=int, default=2)
    ap.add_argument("--turb-n", type=int, default=128)
    ap.add_argument("--turb-steps", type=int, default=500)
    ap.add_argument("--sweep-steps", type=int, default=300)
    ap.add_argument("--serve-requests", type=int, default=80)
    ap.add_argument("--cpu-max-steps", type=int, default=10)
    ap.add_argument("--cpu-budget-s", type=float, default=30.0)
    return ap.parse_args()


PMC_SUMMARY = os.path.join(ROOT, "profiles", "traffic_valu_tgv256_pmc.md")


def read_traffic():
    """HBM bytes per launch of the tolerance-bound face kernel (the x+y
    march) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary."""
    fn = PMC_SUMMARY
    if not os.path.exists(fn):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(fn).read())
    return float(m.group(1)) if m else None


def read_valu_active():
    """VALU-active fraction of the tolerance-bound face kernel (the x+y
    march), from the committed rocprofv3 PMC summary (SQ_ACTIVE_INST_VALU /
    (CUs x cycles))."""
    fn = PMC_SUMMARY
    if not os.path.exists(fn):
        return None
    m = re.search(r"valu_active_face_kernel:\s*([0-9.]+)", open(fn).read())
    return float(m.group(1)) if m else None


def cpu_baseline(n, max_steps, budget_s):
    """The oracle (CPU restatement) on this host's cores, a bounded sample."""
    from oracle import oracle as O

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    # a one-GPU share of the GPU host is 16 cores (whatever the host's core
    # count), so at most 16 threads are used
    cores = max(1, min(cores, 16))
    O.set_threads(cores)
    b = O.Block(O.make_params(n))
    b.init_tgv()
    b.step()  # warm-up (page-in of ~10 GB of arrays)
    t0 = time.perf_counter()
    k = 0
    while k < max_steps:
        b.step()
        k += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    del b
    return {"value": n ** 3 * k / el / 1e6, "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
            "sample": f"{k} TGV {n}^3 FP64 steps ({el:.1f} s) after 1 warm-up step, oracle/ C restatement "
                      "(bit-exact with the reference) with OpenMP; on the 8-core build container it runs at "
                      "1.09x (1 thread) and 1.68x (8 threads) the reference binary's speed at -np 1 / -np 8 "
                      "(oracle/calibrate_cpu.py, TGV 64^3), so it overstates the reference's CPU rate"}


FIELDS = ("rho", "rho_u", "rho_v", "rho_w", "E")


def form_leg(C, N, dims, steps_total, steps_timed, Q64, precision):
    """Another arithmetic form (FP32-flux, or contracted FP64) on the same
    workload: throughput over steps_timed steps after its own short tuning,
    and its per-field error against the plain FP64 state at the same step
    count (relative L2 = RMS difference / RMS of the FP64 field; relative
    L-inf = max difference / max |FP64 field|)."""
    import numpy as np

    from cbns3d_cart_mpi_amd import autotune

    s = C.Solver(C.make_params(N), dims, precision=precision)
    s.init_tgv()
    pre = steps_total - steps_timed - BREAKDOWN_STEPS  # error taken at the FP64 leg's step count
    occs = (2,) if precision == C.CBNS_FP32_FLUX else (2, 3)
    best = None
    if pre >= 3 * (2 * len(autotune.SCHEDULES) * len(occs) + len(autotune.SHAPES)):
        best, tms = autotune.tune(s, steps=2, bands=(16, 32), schedules=autotune.SCHEDULES, occs=occs)
        pre -= 3 * len(tms)
    s.run(pre)
    s.sync()
    t0 = time.perf_counter()
    s.run(steps_timed)
    s.sync()
    el = time.perf_counter() - t0
    s.set_profiling(True)
    s.run(BREAKDOWN_STEPS)
    kt = s.kernel_times()
    Q = s.get_state()
    s.close()
    err = {}
    for e, nm in enumerate(FIELDS):
        d = Q[..., e] - Q64[..., e]
        ref = Q64[..., e]
        err[nm] = {"rel_l2": float(np.sqrt(np.mean(d * d) / np.mean(ref * ref))),
                   "rel_linf": float(np.abs(d).max() / np.abs(ref).max())}
    v = N ** 3 * steps_timed / el / 1e6
    return {"value": round(v, 2), "unit": "Mcell-steps/s", "ms_per_step": round(el / steps_timed * 1e3, 3),
            "error_vs_fp64_after_steps": steps_total, "error": err,
            "kernels_avg_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in kt.items() if v[1]},
            "autotune": {"band": best[0], "schedule": best[1], "face_waves_per_simd": best[2],
                         "visc_planes_run_cells": list(best[3])} if best else None}


def sweep_leg(steps):
    """64 independent solves (grids {32,48,64,96,128}^3, Re log-U[100,3200],
    Mach U[0.1,0.5], half Taylor-Green / half turbulence), all submitted at
    once; one lane (batch: one persistent launch per kernel per stage, one
    hipGraph per step) per grid size, the lanes running concurrently, so each
    solve has its own completion time.  FP64 (bit-exact and tolerance-bound)
    and FP32-flux; per-solve error of FP32-flux against FP64."""
    from cbns3d_cart_mpi_amd import capi as C
    from cbns3d_cart_mpi_amd import serve as S

    import numpy as np

    out = {}
    ref = {}
    worst = []
    for name, prec, form in (("fp64", "fp64", C.CBNS_FP64), ("fp64_fast", "fp64", C.CBNS_FP64_FAST),
                             ("fp32", "fp32", C.CBNS_FP64)):
        reqs = S.make_requests(64, seed=2024, steps=(steps, steps), precision=prec)
        r = S.run_sweep_lanes(reqs, collect=name != "fp64_fast", fp64_form=form)
        out[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        for q in reqs:
            if name == "fp64":
                ref[q.rid] = q.result
            elif name == "fp32":  # per-solve error of the FP32-flux form against the same solve in FP64
                Q64 = ref.pop(q.rid)
                worst.append(max(float(np.sqrt(np.mean((q.result[..., e] - Q64[..., e]) ** 2) /
                                               np.mean(Q64[..., e] ** 2))) for e in range(5)))
            q.result = None
    out["fp32_over_fp64"] = round(out["fp32"]["mcell_steps_per_s"] / out["fp64"]["mcell_steps_per_s"], 3)
    out["fp32_rel_l2_worst_field_per_solve"] = {"max": float(np.max(worst)), "median": float(np.median(worst))}
    out["config"] = (f"64 solves x {steps} steps submitted at t = 0, one lane per grid size; latency = submission "
                     "to the result on the host")
    return out


def batched_leg(N, steps, copies, plain_exact, plain_fast):
    """The batched form (persistent kernels over the packed block table, one
    hipGraph per step, band and viscous schedule from the batch autotuner) on
    `copies` Taylor-Green N^3 solves in each form: Mcell-steps/s of each
    against the single-solve FP64 paths' rates."""
    from cbns3d_cart_mpi_amd import autotune
    from cbns3d_cart_mpi_amd import capi as C

    tune_steps = 2
    bands, overlaps = (16, 32, 64), (0, 256, 512)
    tuning = (tune_steps + 1) * len(bands) * len(overlaps)
    out = {}
    for name, prec in (("fp64", C.CBNS_FP64), ("fp64_fast", C.CBNS_FP64_FAST), ("fp32", C.CBNS_FP32_FLUX)):
        with C.Batch(200 << 30, capacity=copies) as b:
            for _ in range(copies):
                b.add(C.make_params(N), precision=prec, steps=2 + tuning + steps)
            b.run(1)  # captures the step graph and takes the first step (untimed)
            best, _ = autotune.tune_batch(b, steps=tune_steps, bands=bands, overlaps=overlaps)
            b.run(1)  # re-capture on the chosen schedule (untimed)
            b.sync()
            t0 = time.perf_counter()
            b.run(steps)
            b.sync()
            el = time.perf_counter() - t0
        v = copies * N ** 3 * steps / el / 1e6
        out[name] = {"mcell_steps_per_s": round(v, 2), "ms_per_step": round(el / steps * 1e3, 3),
                     "autotune": {"band": best[0], "viscous_overlap_workgroups": best[1]}}
    out["fp32_over_plain_fp64"] = round(out["fp32"]["mcell_steps_per_s"] / plain_exact, 3)
    out["fp32_over_plain_fp64_fast"] = round(out["fp32"]["mcell_steps_per_s"] / plain_fast, 3)
    out["config"] = (f"{copies} Taylor-Green {N}^3 solves in one batch, {steps} timed steps; ratios against the "
                     "single-solve bit-exact and tolerance-bound FP64 rates of this run")
    return out


def turbulence_leg(N, steps, seed=2024):
    """Decaying turbulence (M_t 0.6, Re_lambda 100, seeded) at N^3: the plain
    FP64 path and the FP32-flux form over `steps` steps from the same state;
    throughput of each, per-field error of FP32-flux against FP64 and the
    dilatation-spectrum difference (host numpy), and the device spectra
    (cbns_spectra): their time per call, the FP64 dilatation spectrum against
    the host definition and the FP32-flux spectra against FP64's."""
    from cbns3d_cart_mpi_amd import capi as C
    from cbns3d_cart_mpi_amd.turbulence import precision_errors, turbulence_state

    import numpy as np

    from cbns3d_cart_mpi_amd.turbulence import dilatation_spectrum

    Q0, Re = turbulence_state(N, seed=seed)
    p = C.make_params(N, Mach=1.0, Re=Re, nstep_max=steps, t_end=1.0e3)
    out, Q, spec = {}, {}, {}
    spec_ms = []
    for name, prec in (("fp64", C.CBNS_FP64), ("fp64_fast", C.CBNS_FP64_FAST), ("fp32", C.CBNS_FP32_FLUX)):
        with C.Solver(p, precision=prec) as s:
            s.set_state(Q0, 0.0, mu0=0.0)
            s.run(1)  # untimed first step
            s.sync()
            t0 = time.perf_counter()
            s.run(steps - 1)
            el = time.perf_counter() - t0
            Q[name] = s.get_state()
            out[name] = {"mcell_steps_per_s": round(N ** 3 * (steps - 1) / el / 1e6, 2), "t_end": s.t_cur}
            s.spectra(N // 2 + 1)  # plans the transform
            t0 = time.perf_counter()
            spec[name] = s.spectra(N // 2 + 1)
            spec_ms.append((time.perf_counter() - t0) * 1e3)
    out["fp32_error_vs_fp64"] = precision_errors(Q["fp32"], Q["fp64"], N)
    out["fp64_fast_error_vs_fp64"] = precision_errors(Q["fp64_fast"], Q["fp64"], N)
    # the device spectra (cbns_spectra) beside the host numpy definition
    host = dilatation_spectrum(Q["fp64"], N)
    d64, d32 = spec["fp64"][2], spec["fp32"][2]
    out["device_spectra"] = {
        "ms_per_call": round(min(spec_ms), 3),
        "fp64_dilatation_vs_host_numpy_max_rel": float(np.abs(d64 - host).max() / host.max()),
        "fp32_vs_fp64_dilatation_max_rel_shell": float((np.abs(d32[1:] - d64[1:]) / d64[1:]).max()),
        "fp32_vs_fp64_ke_max_rel_shell": float((np.abs(spec["fp32"][0][1:] - spec["fp64"][0][1:])
                                                / spec["fp64"][0][1:]).max())}
    out["config"] = f"{N}^3, seed {seed}, Mach 1 (sound speed 1), M_t 0.6, Re {Re:.4g}, {steps} steps"
    return out


def io_leg(N, steps=20):
    """Output and diagnostics beside the step on TGV N^3 (tolerance-bound
    form, one block): the time of `steps` steps alone and with an
    asynchronous snapshot (cbns_snapshot_begin, after a first one that
    allocated its buffers) issued before them, whose
    device copy runs on the solver's stream and whose host copy runs on a
    copy stream beside the steps; the wait for the host copy after them; the
    synchronous copies an output takes without it (cbns_get_block_state and
    the five cbns_get_block_primitive fields); and the device velocity
    spectra (cbns_spectra) per call."""
    import numpy as np

    from cbns3d_cart_mpi_amd import capi as C

    with C.Solver(C.make_params(N, nstep_max=10 ** 6, t_end=20.0), precision=C.CBNS_FP64_FAST) as s:
        s.init_tgv()
        s.run(3)
        s.sync()

        def timed(f):
            t0 = time.perf_counter()
            r = f()
            s.sync()
            return (time.perf_counter() - t0) * 1e3, r

        # the first snapshot allocates its device and pinned host buffers
        first, _ = timed(lambda: (s.snapshot_begin(), s.snapshot_wait()))
        plain, _ = timed(lambda: s.run(steps))
        snap_state = s.get_block_state(0)  # the state the snapshot will hold
        t_snap = s.t_cur
        t0 = time.perf_counter()
        s.snapshot_begin()
        s.run(steps)
        s.sync()
        with_snap = (time.perf_counter() - t0) * 1e3
        wait, t_held = timed(s.snapshot_wait)
        same = bool(np.array_equal(s.snapshot_block_state(0), snap_state)) and t_held == t_snap
        sync_copy, _ = timed(lambda: (s.get_block_state(0), [s.get_block_primitive(0, f) for f in range(5)]))
        s.spectra()  # plans the transform
        spec = min(timed(s.spectra)[0] for _ in range(3))
    return {"grid": N, "steps": steps,
            "steps_ms": round(plain, 2), "steps_ms_with_snapshot": round(with_snap, 2),
            "snapshot_wait_ms_after_steps": round(wait, 3), "snapshot_equals_state_of_its_moment": same,
            "first_snapshot_ms_incl_allocation": round(first, 1),
            "synchronous_output_copies_ms": round(sync_copy, 2), "device_spectra_ms_per_call": round(spec, 3)}


def serve_leg(n):
    """Continuous batching of a Poisson stream (reduced size): n requests,
    grids {32..192}^3 with weights 1/N, steps U[50, 200], FP64 / FP32-flux
    alternating (FP64 requests in the tolerance-bound form), 200 GB budget,
    offered load 85% of the capacity measured on a probe of the same mix,
    three size-class lanes; against the same requests solved one at a time as
    bit-exact plain FP64, measured request by request and from the per-size
    cost model that the full 1000-request run uses.  The probe is also served
    in each arithmetic form alone (FP32-flux only, and FP64 only in each
    form): the server's capacity per form."""
    from cbns3d_cart_mpi_amd import serve as S

    import numpy as np

    grids = (32, 48, 64, 96, 128, 160, 192)
    w = [1.0 / g for g in grids]
    probe = S.make_requests(32, 7, grids, w, (50, 50))
    cap = S.run_sweep_lanes(probe, fp64_form=S.C.CBNS_FP64_FAST, lanes=S.LANES)["mcell_steps_per_s"] * 1e6
    # the same probe served in one arithmetic form at a time
    cap_form = {}
    for nm, prec, form in (("fp32_flux", "fp32", S.C.CBNS_FP64_FAST), ("fp64_tolerance_bound", "fp64", S.C.CBNS_FP64_FAST),
                           ("fp64_bit_exact", "fp64", S.C.CBNS_FP64)):
        pr = S.make_requests(32, 7, grids, w, (50, 50), precision=prec)
        cap_form[nm] = round(S.run_sweep_lanes(pr, fp64_form=form, lanes=S.LANES)["mcell_steps_per_s"], 1)
    mean_work = np.mean([r.work for r in S.make_requests(n, 2024, grids, w, (50, 200))])
    rate = 0.85 * cap / mean_work
    reqs = S.make_requests(n, 2024, grids, w, (50, 200), rate=rate)
    srv = S.serve(reqs, collect=True)
    seq = S.sequential_fp64(reqs)
    model = S.sequential_model(reqs)
    rnd = lambda d: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}
    return {"requests": n, "rate_per_s": round(rate, 3), "capacity_probe_mcell_steps_per_s": round(cap / 1e6, 1),
            "capacity_probe_by_form_mcell_steps_per_s": cap_form,
            "fp32_flux_capacity_over_sequential": round(cap_form["fp32_flux"] / seq["mcell_steps_per_s"], 3),
            "served": rnd(srv), "sequential_fp64": rnd(seq), "sequential_fp64_model": rnd(model),
            "served_over_sequential": round(srv["mcell_steps_per_s"] / seq["mcell_steps_per_s"], 3),
            "config": "reduced: steps U[50,200] instead of U[200,2000]; 80 instead of 1000 requests "
                      "(the full configuration: profiles/serving_full.md)"}


def zfused_bytes_per_cell(precision):
    """Algorithmic HBM bytes per cell of one launch of the fused z-face and
    stage-tail kernel, averaged over the three stages: the x- and y-face
    fluxes (5 each) and the viscous divergence (4) (the tolerance-bound form
    reads instead the 5 values its x+y march stored: the x- plus y-flux
    differences less the divergence), the stage input (5, read once along
    the column's march), Q^n (5, stages 2 and 3 only) and the updated state
    (5 written); the bit-exact form also writes 6 primitives."""
    operands = 5 if precision == 3 else 10 + 4
    b = 8 * (operands + 5 + 5) + 8 * 5 * 2 / 3
    return b + (48 if precision == 0 else 0)


def plain_leg(C, N, dims, precision, steps, warmup, fixed=None):
    """One FP64 form on TGV N^3: autotuned launch shapes, `warmup` untimed
    steps, then `steps` timed steps with the face sweeps timed live (HIP
    events on the solver's streams), then a per-kernel breakdown over further
    steps run serially (the viscous kernel after the sweeps, so each kernel's
    duration is its own).  Returns (line fragment, solver)."""
    from cbns3d_cart_mpi_amd import autotune

    s = C.Solver(C.make_params(N, nstep_max=10 ** 6, t_end=20.0), dims, precision=precision)
    s.init_tgv()
    s.run(warmup)
    tuned = None
    tuned_steps = 0
    if fixed is None:
        # the tolerance-bound form's face kernel (the x+y march) has one register
        # variant and reads the viscous divergence (its stages run serially)
        fast = precision == C.CBNS_FP64_FAST
        best, tms = autotune.tune(s, steps=2, bands=(16, 32, 64),
                                  schedules=("serial",) if fast else autotune.SCHEDULES,
                                  occs=(2,) if fast else (2, 3))
        tuned_steps = 3 * len(tms)
        tuned = {"band": best[0], "schedule": best[1], "face_waves_per_simd": best[2],
                 "visc_planes_run_cells": list(best[3]),
                 "ms_per_step (band/schedule/waves)": {f"{k[0]}/{k[1]}/{k[2]}": round(v * 1e3, 3)
                                                       for k, v in tms.items() if len(k) == 3},
                 "ms_per_step (visc planes/run cells)": {f"{k[3][0]}/{k[3][1]}": round(v * 1e3, 3)
                                                         for k, v in tms.items() if len(k) == 4}}
    else:  # launch shapes given on the command line (profiling runs)
        band, sched, occ, shape = fixed
        s.set_band(band)
        autotune.apply_schedule(s, sched)
        s.set_face_occupancy(occ)
        s.set_tile_shape(*shape)
        tuned = {"band": band, "schedule": sched, "face_waves_per_simd": occ, "visc_planes_run_cells": list(shape),
                 "given": True}
    sweeps = ("face_x", "face_y", "face_z")
    s.sync()
    s.set_profiling_classes(sweeps)  # live HIP-event timing of the roofline kernel over the timed steps
    t0 = time.perf_counter()
    dts = s.run(steps)  # returns after a stream synchronisation
    el = time.perf_counter() - t0
    kt_face = s.kernel_times()
    assert len(dts) == steps
    # per-kernel breakdown outside the timed region, serial schedule
    sched = tuned["schedule"] if tuned else "overlap"
    autotune.apply_schedule(s, "serial")
    s.set_profiling(True)
    s.run(BREAKDOWN_STEPS)
    kt = s.kernel_times()
    s.set_profiling(False)
    autotune.apply_schedule(s, sched)
    faces_per_launch = (N + 1) * N * N
    flop = FLOP_PER_FACE * faces_per_launch  # one axis of faces
    sweeps = tuple(k for k in sweeps if kt_face[k][1])  # the z faces of the fused forms are in the stage tail
    # the tolerance-bound form's x+y march (timed as face_y) forms two axes of faces per launch
    flop_k = {k: flop * (2 if precision == C.CBNS_FP64_FAST else 1) for k in sweeps}
    per_axis = {k: kt_face[k][0] / max(kt_face[k][1], 1) for k in sweeps}
    avg_ms = sum(kt_face[k][0] for k in sweeps) / max(sum(kt_face[k][1] for k in sweeps), 1)
    achieved = (sum(flop_k[k] * kt_face[k][1] for k in sweeps) / max(sum(kt_face[k][1] for k in sweeps), 1)
                / (avg_ms * 1e-3) / 1e12)
    serial = {k: kt[k][0] / max(kt[k][1], 1) for k in sweeps}
    value = N ** 3 * steps / el / 1e6
    zf = None
    if "face_z" not in sweeps and kt["update"][1]:  # the fused z-face and stage-tail kernel
        zms = kt["update"][0] / kt["update"][1]
        zbytes = zfused_bytes_per_cell(precision) * N ** 3
        zf = {"avg_launch_ms_serial": round(zms, 4),
              "valu": {"achieved_tflops": round(flop / (zms * 1e-3) / 1e12, 3),
                       "frac": round(flop / (zms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4),
                       "algorithmic_flop_per_launch": flop},
              "hbm": {"achieved_gbs": round(zbytes / (zms * 1e-3) / 1e9, 1),
                      "frac": round(zbytes / (zms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
                      "algorithmic_bytes_per_launch": round(zbytes)}}
    frag = {
        "value": round(value, 2), "ms_per_step": round(el / steps * 1e3, 3),
        "face_roofline": {
            "achieved": round(achieved, 3), "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
            "avg_launch_ms": round(avg_ms, 4),
            "per_axis_launch_ms_timed_steps": {k: round(v, 4) for k, v in per_axis.items()},
            "per_axis_frac_timed_steps": {k: round(flop_k[k] / (v * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)
                                          for k, v in per_axis.items()},
            "per_axis_frac_serial": {k: round(flop_k[k] / (v * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)
                                     for k, v in serial.items()},
            "algorithmic_flop_per_launch": {k: flop_k[k] for k in sweeps}},
        "kernels_avg_ms_serial": {k: round(v[0] / max(v[1], 1), 4) for k, v in kt.items() if v[1]},
        "fused_z_update": zf,
        "autotune": tuned,
        "steps_before_timed": warmup + tuned_steps,
    }
    return frag, s


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the path shards by "
              "on-device blocks but multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    import numpy as np

    from cbns3d_cart_mpi_amd import capi as C

    dims = tuple(int(x) for x in a.dims.split(","))
    N = a.n
    if a.io_only:
        print(json.dumps({"io": io_leg(N)}), flush=True)
        return
    # headline: the tolerance-bound FP64 path (within 1e-12 of the reference
    # after one step); the bit-exact plain path is measured beside it
    given = "auto" not in (a.band, a.schedule, a.occ)
    # given shapes: --occ applies to the headline form, the bit-exact form keeps 2 waves per SIMD
    shape = tuple(int(x) for x in a.shape.split(","))
    fixed = lambda occ: ((int(a.band), a.schedule, occ, shape) if given else None)
    if not given and any(x != "auto" for x in (a.band, a.schedule, a.occ)):
        sys.exit("give all of --band, --schedule and --occ, or none")
    head, s = plain_leg(C, N, dims, C.CBNS_FP64_FAST, a.steps, a.warmup, fixed(int(a.occ) if given else 3))
    ke, en = s.diagnostics()
    mem = s.memory_bytes
    s_fast = s
    exact, s = plain_leg(C, N, dims, C.CBNS_FP64, a.steps, a.warmup, fixed(2))
    # both forms' states after the same number of steps from the same start
    # (the autotuners may have run different numbers of steps)
    done_fast = head["steps_before_timed"] + a.steps + BREAKDOWN_STEPS
    done_exact = exact["steps_before_timed"] + a.steps + BREAKDOWN_STEPS
    steps_head = max(done_fast, done_exact)
    s_fast.run(steps_head - done_fast)
    s.run(steps_head - done_exact)
    Q_fast = s_fast.get_state()
    s_fast.close()
    Q64 = s.get_state()
    s.close()
    fast_err = {nm: float(np.abs(Q_fast[..., e] - Q64[..., e]).max() / np.abs(Q64[..., e]).max())
                for e, nm in enumerate(FIELDS)}
    del Q_fast
    fp32 = fma = None
    if not a.no_fma:
        fma = form_leg(C, N, dims, steps_head, a.steps, Q64, C.CBNS_FP64_FMA)
    if not a.no_fp32:
        fp32 = form_leg(C, N, dims, steps_head, a.steps, Q64, C.CBNS_FP32_FLUX)
        fp32["over_bit_exact_fp64"] = round(fp32["value"] / exact["value"], 3)
        fp32["over_tolerance_bound_fp64"] = round(fp32["value"] / head["value"], 3)
    del Q64

    value = head["value"]
    fr = head["face_roofline"]
    roof = {"bound": "valu", "achieved": fr["achieved"], "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": fr["frac"], "traffic": read_traffic(),
            "kernel": "tolerance-bound FP64 face kernel k_xyface: the x and y faces of a stage in one march "
                      "along y, which stores each cell's x+y flux differences less its viscous divergence "
                      "(timed as face_y); FP64 "
                      "VALU-issue bound; achieved = 1570 algorithmic flop per face (literal count of the "
                      "reference) x (x faces + y faces) per launch / its average launch duration in the timed "
                      "steps; the z faces are in the fused stage tail (fused_z_update)",
            "avg_launch_ms": fr["avg_launch_ms"],
            "algorithmic_flop_per_launch": 2 * FLOP_PER_FACE * (N + 1) * N * N,
            "per_axis_frac_timed_steps": fr["per_axis_frac_timed_steps"],
            "per_axis_frac_serial": fr["per_axis_frac_serial"],
            "valu_active_frac": read_valu_active(),
            "fused_z_update": head["fused_z_update"]}
    line = {
        "metric": "Mcell-steps/s TGV 256^3 FP64 vs FP32-flux (L2 err stated); sweep p50/p99 latency",
        "value": value,
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Taylor-Green initial condition, no dataset)",
        "config": {"workload": f"Taylor-Green vortex {N}^3 periodic, FP64 tolerance-bound path "
                               f"(CBNS_FP64_FAST, within 1e-12 of the reference after one step), blocks {a.dims}",
                   "grid": [N, N, N], "blocks": list(dims), "Re": 1600.0, "Mach": 0.1, "CFL": 0.5},
        "roofline": roof,
        "algorithmic_tflops_all_kernels": round(FLOP_PER_CELL_STEP * value * 1e6 / 1e12, 3),
        "kernels_avg_ms_serial": head["kernels_avg_ms_serial"],
        "autotune": head["autotune"],
        "device_bytes": mem,
        "ke_enstrophy_end": [ke, en],
        "fp64_bit_exact": {k: v for k, v in exact.items() if k != "steps_before_timed"},
        "fp64_fast_vs_bit_exact_rel_linf_own_max": {"after_steps": steps_head, "error": fast_err},
    }
    if fma:
        line["fp64_fma"] = fma
    if fp32:
        line["fp32_flux"] = fp32
    if not a.no_io:
        # before the legs that open many streams (batches, lanes): the snapshot's
        # copy stream then has a hardware queue of its own
        line["io"] = io_leg(N)
    if not a.no_batched:
        line["batched"] = batched_leg(N, a.batched_steps, a.batched_copies, exact["value"], value)
    if not a.no_turb:
        line["turbulence"] = turbulence_leg(a.turb_n, a.turb_steps)
    if not a.no_sweep:
        line["sweep"] = sweep_leg(a.sweep_steps)
    if not a.no_serve:
        line["serving"] = serve_leg(a.serve_requests)
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(N, a.cpu_max_steps, a.cpu_budget_s)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
