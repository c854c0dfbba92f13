"""Parameter-sweep solves: a batched sweep and a continuous-batching server.

    python -m cbns3d_cart_mpi_amd.serve sweep  [--n 64] [--steps 300] [--precision fp64|fp32|alt]
    python -m cbns3d_cart_mpi_amd.serve serve  [--n 128] [--steps-min 20] [--steps-max 200] [--rate R]
                                               [--baseline measured|model|none] [--fp64-form fast|exact]

Requests are drawn from a seeded stream:
  * grid N^3 from a set of sizes (sweep: {32,48,64,96,128}, uniform; serving:
    {32,48,64,96,128,160,192} with weights proportional to 1/N);
  * Re log-uniform in [100, 3200], Mach uniform in [0.1, 0.5];
  * half start from the Taylor-Green field (input Mach = the drawn Mach), half
    from the seeded decaying-turbulence field (cbns3d_cart_mpi_amd.turbulence)
    with the drawn Mach as its turbulent Mach number (input Mach = 1, so that
    the sound speed is 1);
  * FP64 or FP32-flux (serving: alternating).

Serving is continuous batching: requests arrive on a Poisson clock; each is
admitted as soon as its state fits the HBM budget, stays resident on the
device, and advances with the other live solves of its size class in the
same persistent launches (one hipGraph per step per class), the classes
running concurrently; a class is serviced (finished solves collected and
removed, waiting requests admitted, its next chunk issued) whenever its last
chunk has finished, while the other classes keep running.  FP64 requests
are served in the tolerance-bound FP64 form by default (--fp64-form exact
for the bit-exact one).  Time-to-completion is measured from arrival to the
host having the result.  The baseline processes the same requests one at a
time, first come first served, as plain (bit-exact) FP64 single solves,
measured request by request or from a per-size cost model.
"""
import argparse
import json
import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi as C
from .turbulence import turbulence_state

GB = 1 << 30


@dataclass
class Request:
    rid: int
    N: int
    steps: int
    Re: float
    mach: float
    ic: str          # "tgv" or "turb"
    precision: int
    seed: int
    arrival: float = 0.0
    admitted: float = math.nan
    done: float = math.nan
    t_final: float = math.nan
    lane: int = -1       # size class it was served in
    bid: int = -1        # instance id in that class's batch
    ran: int = 0         # steps issued to it so far
    nonfinite: bool = False  # it met a non-finite value (the result is the failed state)
    payload: object = field(default=None, repr=False)
    result: object = field(default=None, repr=False)

    @property
    def work(self):
        return self.N ** 3 * self.steps


def make_requests(n, seed=2024, grids=(32, 48, 64, 96, 128), weights=None, steps=(300, 300), rate=None,
                  precision="alt"):
    rng = np.random.default_rng(seed)
    grids = np.asarray(grids)
    w = None if weights is None else np.asarray(weights, float) / np.sum(weights)
    t = 0.0
    out = []
    for r in range(n):
        N = int(rng.choice(grids, p=w))
        st = int(rng.integers(steps[0], steps[1] + 1))
        Re = float(np.exp(rng.uniform(np.log(100.0), np.log(3200.0))))
        mach = float(rng.uniform(0.1, 0.5))
        ic = "tgv" if r % 2 == 0 else "turb"
        if precision == "alt":
            prec = C.CBNS_FP64 if (r // 2) % 2 == 0 else C.CBNS_FP32_FLUX
        else:
            prec = C.CBNS_FP32_FLUX if precision == "fp32" else C.CBNS_FP64
        if rate:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(r, N, st, Re, mach, ic, prec, seed=int(rng.integers(1 << 31)), arrival=t))
    return out


def prepare(req):
    """(params, Q or None) of a request; the turbulence field is its payload."""
    if req.ic == "tgv":
        return C.make_params(req.N, Mach=req.mach, Re=req.Re, nstep_max=req.steps), None
    Q, _ = turbulence_state(req.N, seed=req.seed, mach_t=req.mach)
    return C.make_params(req.N, Mach=1.0, Re=req.Re, nstep_max=req.steps), Q


def percentile(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else math.nan


def run_sweep(reqs, budget=200 * GB, collect=False, workgroups=None):
    """All requests admitted at time 0 into one batch and advanced together.
    With collect, each request's final state is kept in r.result; workgroups
    sets the persistent grid of the memory-bound kernels."""
    for r in reqs:
        r.payload = prepare(r)
    t_setup = time.perf_counter()
    with C.Batch(budget, capacity=max(len(reqs), 1)) as b:
        if workgroups:
            b.set_workgroups(workgroups)
        ids = {}
        for r in reqs:
            p, Q = r.payload
            ids[r.rid] = b.add(p, precision=r.precision, Q=Q, t_cur=0.0, mu0=0.0, steps=r.steps)
        b.run(1)  # captures the step graph and takes the first step (untimed)
        b.sync()
        t0 = time.perf_counter()
        steps_max = max(r.steps for r in reqs)
        b.run(steps_max)  # the remaining steps; solves stop at their own counts
        b.sync()
        el = time.perf_counter() - t0
        for r in reqs:
            st = b.status(ids[r.rid])
            assert not st["nonfinite"], r
            r.t_final = st["t_cur"]
            r.done = el
            if collect:
                r.result = b.get_state(ids[r.rid])
    work = sum(r.N ** 3 * (r.steps - 1) for r in reqs)  # timed steps only
    return {"wall_s": el, "setup_s": t0 - t_setup, "cell_steps": work, "mcell_steps_per_s": work / el / 1e6,
            "p50_s": percentile([r.done for r in reqs], 50), "p99_s": percentile([r.done for r in reqs], 99)}


LANES = ((0, 64), (65, 128), (129, 1 << 30))  # size classes by grid edge N


def _form(r, fp64_form):
    """Arithmetic a request is solved in: FP32-flux as asked, FP64 requests
    in fp64_form (CBNS_FP64: bit-exact with the reference; CBNS_FP64_FAST:
    tolerance-bound, within 1e-12 of it)."""
    return r.precision if r.precision == C.CBNS_FP32_FLUX else fp64_form


def serve(reqs, budget=200 * GB, chunk_ms=40.0, capacity=256, collect=True, lanes=LANES, max_live=None, log=None,
          fp64_form=C.CBNS_FP64_FAST, rate_guess=1.2e9, chunk=None):
    """Continuous batching over a Poisson request stream (arrival times in s).

    Solves are grouped into size-class lanes, each a device batch with its own
    stream and step graph, so a small solve never advances in lockstep with a
    large one, and the lanes run concurrently on the GPU.  The host services
    a lane only when its last chunk has finished (cbns_batch_query, no wait):
    it collects the solves that are done (their results come back to the
    host and their memory is kept for reuse), admits waiting requests of that
    lane (first come first admitted, while the shared HBM budget and the
    optional cap on resident solves allow), and issues the lane's next chunk
    of steps, sized to about chunk_ms of device time but never past the step
    at which its next solve finishes.  The other lanes keep running
    meanwhile.  Time-to-completion runs from arrival to the result on the
    host.  log, if given, is called with each finished request; chunk, if
    given, fixes the steps per chunk instead."""
    for r in reqs:
        r.payload = prepare(r)
    pending = sorted(reqs, key=lambda r: r.arrival)
    waiting = []
    done = []
    batches = [C.Batch(budget, capacity=capacity) for _ in lanes]
    live = [dict() for _ in lanes]
    used = [0]
    host = {"admit": 0.0, "issue": 0.0, "poll": 0.0, "collect": 0.0, "idle": 0.0}
    lane_of = lambda r: next(n for n, (lo, hi) in enumerate(lanes) if lo <= r.N <= hi)
    try:
        t0 = time.perf_counter()
        k = 0
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            while k < len(pending) and pending[k].arrival <= now:
                waiting.append(pending[k])
                k += 1
            serviced = False
            for li, b in enumerate(batches):
                if live[li] and b.busy():
                    continue
                # lane idle: collect its finished solves
                tc = time.perf_counter()
                if live[li]:
                    fin, _ = b.poll_status()
                    host["poll"] += time.perf_counter() - tc
                    for st in fin:
                        i = st["id"]
                        r = live[li].pop(i)
                        r.t_final = st["t_cur"]
                        r.nonfinite = st["nonfinite"]
                        if collect:
                            r.result = b.get_state(i)
                        r.done = time.perf_counter() - t0
                        b.remove(i)
                        used[0] -= C.batch_bytes_needed(r.payload[0])
                        done.append(r)
                        if log:
                            log(r)
                ta = time.perf_counter()
                host["collect"] += ta - tc  # includes the poll
                # admit this lane's waiting requests, first come first admitted
                for r in [w for w in waiting if lane_of(w) == li]:
                    p, Q = r.payload
                    need = C.batch_bytes_needed(p)
                    nlive = sum(len(x) for x in live)
                    if max_live is not None and nlive >= max_live:
                        break
                    if used[0] + need > budget:
                        if not nlive:
                            raise RuntimeError(f"request {r.rid} does not fit the budget alone")
                        break
                    if used[0] + need + sum(x.pooled_bytes for x in batches) > budget:
                        for x in batches:  # memory kept for reuse by other lanes goes first
                            if x is not b and not x.busy():
                                x.trim()
                    i = b.add(p, precision=_form(r, fp64_form), Q=Q, t_cur=0.0, mu0=0.0, steps=r.steps)
                    used[0] += need
                    r.admitted = time.perf_counter() - t0
                    r.lane, r.bid, r.ran = li, i, 0
                    live[li][i] = r
                    waiting.remove(r)
                tb = time.perf_counter()
                host["admit"] += tb - ta
                if live[li]:  # next chunk: about chunk_ms, never past the lane's next completion
                    cells = sum(r.N ** 3 for r in live[li].values())
                    left = min(r.steps - r.ran for r in live[li].values())
                    want = chunk if chunk else int(chunk_ms * 1e-3 / (cells / rate_guess + 1.5e-4))
                    nrep = max(1, min(left, want))
                    b.run(nrep)
                    for r in live[li].values():
                        r.ran += nrep
                    serviced = True
                host["issue"] += time.perf_counter() - tb
            if not serviced:
                ti = time.perf_counter()
                if not any(live) and k < len(pending):
                    time.sleep(max(0.0, pending[k].arrival - (time.perf_counter() - t0)))
                else:
                    time.sleep(2e-4)
                host["idle"] += time.perf_counter() - ti
        el = time.perf_counter() - t0
    finally:
        for b in batches:
            b.close()
    lat = [r.done - r.arrival for r in reqs]
    work = sum(r.work for r in reqs)
    return {"wall_s": el, "cell_steps": work, "mcell_steps_per_s": work / el / 1e6,
            "p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99), "mean_s": float(np.mean(lat)),
            "failed": sum(bool(getattr(r, "nonfinite", False)) for r in reqs),
            "host_s": {k: round(v, 4) for k, v in host.items()}}


def size_lanes(reqs):
    """One lane per grid size present in reqs."""
    return tuple((N, N) for N in sorted({r.N for r in reqs}))


def run_sweep_lanes(reqs, budget=200 * GB, collect=False, fp64_form=C.CBNS_FP64, chunk_ms=40.0, lanes=None):
    """A parameter sweep: every request admitted at time 0 into its lane (by
    default one lane per grid size), the lanes advancing concurrently;
    per-solve completion times from the start (small solves finish first)."""
    for r in reqs:
        r.arrival = 0.0
    res = serve(reqs, budget, chunk_ms=chunk_ms, collect=collect, fp64_form=fp64_form,
                lanes=lanes or size_lanes(reqs))
    bad = [r.rid for r in reqs if getattr(r, "nonfinite", False)]
    assert not bad, f"solves turned non-finite: {bad}"
    return res


def sequential_fp64(reqs, precision=C.CBNS_FP64, log=None):
    """The same requests solved one at a time as single solves (plain FP64
    unless told otherwise), first come first served; latencies from the
    measured service times.  Each request's result is kept in r.alone."""
    service = []
    for r in reqs:
        p, Q = r.payload if r.payload is not None else prepare(r)
        t0 = time.perf_counter()
        with C.Solver(p, precision=r.precision if precision is None else precision) as s:
            if Q is None:
                s.init_tgv()
            else:
                s.set_state(Q, 0.0, mu0=0.0)
            s.run(r.steps)
            r.alone = s.get_state()
        service.append(time.perf_counter() - t0)
        if log:
            log(r)
    clock, lat = 0.0, []
    for r, sv in sorted(zip(reqs, service), key=lambda x: x[0].arrival):
        start = max(clock, r.arrival)
        clock = start + sv
        lat.append(clock - r.arrival)
    work = sum(r.work for r in reqs)
    busy = sum(service)
    return {"busy_s": busy, "mcell_steps_per_s": work / busy / 1e6, "p50_s": percentile(lat, 50),
            "p99_s": percentile(lat, 99), "mean_s": float(np.mean(lat)), "makespan_s": clock}


def sequential_model(reqs, precision=C.CBNS_FP64, probe_steps=20, log=None):
    """The sequential baseline without running every request: for each grid
    size in the stream, one single solve measures its fixed cost (create,
    initial state, result read-back, destroy) and its time per step (over
    probe_steps steps after two untimed ones); a request's service time is
    fixed + steps x per-step, and the same first-come-first-served queue as
    sequential_fp64 gives the latencies.  The step cost of a size does not
    depend on the physics or the initial field, so the model is exact up to
    run-to-run noise (bench.py's serving leg measures both on the same
    requests)."""
    costs = {}
    for N in sorted({r.N for r in reqs}):
        p = C.make_params(N)
        t0 = time.perf_counter()
        with C.Solver(p, precision=precision) as s:
            s.init_tgv()
            s.run(2)
            s.sync()
            t1 = time.perf_counter()
            s.run(probe_steps)
            t2 = time.perf_counter()
            s.get_state()
        t3 = time.perf_counter()
        per = (t2 - t1) / probe_steps
        costs[N] = ((t1 - t0) - 2 * per + (t3 - t2), per)
        if log:
            log(f"sequential cost N={N}: fixed {costs[N][0]:.4f} s, per step {per * 1e3:.4f} ms")
    clock, lat, busy = 0.0, [], 0.0
    for r in sorted(reqs, key=lambda r: r.arrival):
        fixed, per = costs[r.N]
        sv = fixed + r.steps * per
        busy += sv
        start = max(clock, r.arrival)
        clock = start + sv
        lat.append(clock - r.arrival)
    work = sum(r.work for r in reqs)
    return {"busy_s": busy, "mcell_steps_per_s": work / busy / 1e6, "p50_s": percentile(lat, 50),
            "p99_s": percentile(lat, 99), "mean_s": float(np.mean(lat)), "makespan_s": clock,
            "per_size_cost_s": {str(N): [round(c[0], 5), round(c[1], 7)] for N, c in costs.items()}}


FORMS = {"fast": C.CBNS_FP64_FAST, "exact": C.CBNS_FP64}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["sweep", "serve"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--steps-min", type=int, default=20)
    ap.add_argument("--steps-max", type=int, default=200)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "alt"])
    ap.add_argument("--fp64-form", default="fast", choices=sorted(FORMS),
                    help="arithmetic of FP64 requests: tolerance-bound (fast) or bit-exact")
    ap.add_argument("--rate", type=float, default=None, help="arrivals per second (serve)")
    ap.add_argument("--load", type=float, default=0.85, help="offered load vs measured capacity when --rate is unset")
    ap.add_argument("--probe", type=int, default=48, help="requests of the capacity probe")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--chunk-ms", type=float, default=40.0)
    ap.add_argument("--slots", type=int, default=0, help="solves resident at once (0: as memory allows)")
    ap.add_argument("--lanes", default="sizes", choices=["sizes", "classes"],
                    help="one lane per grid size, or three size classes (N <= 64, <= 128, > 128)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--baseline", default="measured", choices=["measured", "model", "none"],
                    help="sequential plain-FP64 baseline: run every request, or the per-size cost model")
    a = ap.parse_args(argv)
    budget = int(a.budget_gb * GB)
    form = FORMS[a.fp64_form]
    if a.mode == "sweep":
        reqs = make_requests(a.n or 64, a.seed, steps=(a.steps, a.steps), precision=a.precision)
        res = run_sweep_lanes(reqs, budget, fp64_form=form, chunk_ms=a.chunk_ms)
        res["requests"] = len(reqs)
        print(json.dumps({"sweep": res}))
        return 0
    grids = (32, 48, 64, 96, 128, 160, 192)
    weights = [1.0 / g for g in grids]
    n = a.n or 128
    rate = a.rate
    cap = None
    if rate is None:
        # calibrate the offered load: the served mix's capacity, all admitted at once
        probe = make_requests(a.probe, a.seed + 1, grids, weights, (a.steps_min, a.steps_min))
        cap = run_sweep_lanes(probe, budget, fp64_form=form, chunk_ms=a.chunk_ms,
                              lanes=None if a.lanes == "sizes" else LANES)["mcell_steps_per_s"] * 1e6
        mean_work = np.mean([r.work for r in make_requests(n, a.seed, grids, weights, (a.steps_min, a.steps_max))])
        rate = a.load * cap / mean_work
        print(f"capacity probe: {cap / 1e6:.1f} Mcell-steps/s -> {rate:.4f} requests/s at load {a.load}", flush=True)
    reqs = make_requests(n, a.seed, grids, weights, (a.steps_min, a.steps_max), rate=rate)
    lanes = tuple((g, g) for g in grids) if a.lanes == "sizes" else LANES
    t_start = time.perf_counter()

    def progress(r):
        if isinstance(r, str):
            print(r, flush=True)
        else:
            print(f"request {r.rid} N={r.N} steps={r.steps} {'fp32' if r.precision else 'fp64'} "
                  f"latency {r.done - r.arrival:.3f} s (at {time.perf_counter() - t_start:.1f} s)", flush=True)

    res = {"requests": n, "rate_per_s": rate, "load": a.load, "capacity_probe_mcell_steps_per_s":
           None if cap is None else cap / 1e6, "slots": a.slots, "fp64_form": a.fp64_form,
           "steps": [a.steps_min, a.steps_max], "budget_gb": a.budget_gb,
           "lanes": a.lanes, "chunk_ms": a.chunk_ms,
           "served": serve(reqs, budget, a.chunk_ms, max_live=a.slots or None, log=progress, fp64_form=form,
                           collect=False, lanes=lanes)}
    print(json.dumps(res), flush=True)
    if a.baseline == "measured":
        res["sequential_fp64"] = sequential_fp64(reqs)
    elif a.baseline == "model":
        res["sequential_fp64_model"] = sequential_model(reqs, log=progress)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
