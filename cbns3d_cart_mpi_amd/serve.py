"""Parameter-sweep solves: a batched sweep and a continuous-batching server.

    python -m cbns3d_cart_mpi_amd.serve sweep  [--n 64] [--steps 300] [--precision fp64|fp32]
    python -m cbns3d_cart_mpi_amd.serve serve  [--n 128] [--steps-min 20] [--steps-max 200] [--rate R]

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
running concurrently; completed solves are collected and removed between
graph-replay chunks, freeing room for waiting requests.
Time-to-completion is measured from arrival to the host having the result.
The baseline processes the same requests one at a time, first come first
served, as plain FP64 single solves.
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


def serve(reqs, budget=200 * GB, chunk=4, capacity=1024, collect=True, lanes=LANES, max_live=None, log=None):
    """Continuous batching over a Poisson request stream (arrival times in s).

    Solves are grouped into size-class lanes, each a device batch with its own
    stream and step graph, so a small solve never advances in lockstep with a
    large one.  Between chunks the server admits waiting requests (first come,
    first admitted, while the shared HBM budget allows), then issues every
    lane's chunk asynchronously so the lanes run concurrently on the GPU: the
    lane holding the largest solve replays `chunk` steps and a lane whose
    largest solve has n cells replays about chunk * n_max / n steps.  Finished
    solves are collected and removed after the chunk, freeing memory.
    max_live caps the number of solves resident at once (None: memory only);
    log, if given, is called with each finished request."""
    for r in reqs:
        r.payload = prepare(r)
    pending = sorted(reqs, key=lambda r: r.arrival)
    waiting = []
    done = []
    batches = [C.Batch(budget, capacity=capacity) for _ in lanes]
    live = [dict() for _ in lanes]
    used = [0]
    host = {"admit": 0.0, "issue": 0.0, "poll": 0.0, "collect": 0.0}
    try:
        t0 = time.perf_counter()
        k = 0
        while len(done) < len(reqs):
            now = time.perf_counter() - t0
            while k < len(pending) and pending[k].arrival <= now:
                waiting.append(pending[k])
                k += 1
            ta = time.perf_counter()
            while waiting:  # first come, first admitted, while memory allows
                r = waiting[0]
                p, Q = r.payload
                need = C.batch_bytes_needed(p)
                nlive = sum(len(x) for x in live)
                if max_live is not None and nlive >= max_live:
                    break
                if used[0] + need > budget:
                    if not nlive:
                        raise RuntimeError(f"request {r.rid} does not fit the budget alone")
                    break
                li = next(n for n, (lo, hi) in enumerate(lanes) if lo <= r.N <= hi)
                i = batches[li].add(p, precision=r.precision, Q=Q, t_cur=0.0, mu0=0.0, steps=r.steps)
                used[0] += need
                r.admitted = time.perf_counter() - t0
                r.lane, r.bid, r.ran = li, i, 0
                live[li][i] = r
                waiting.pop(0)
            tb = time.perf_counter()
            host["admit"] += tb - ta
            active = [li for li in range(len(lanes)) if live[li]]
            if not active:
                if k < len(pending):
                    time.sleep(max(0.0, pending[k].arrival - (time.perf_counter() - t0)))
                continue
            n_all = max(r.N ** 3 for li in active for r in live[li].values())
            for li in active:  # asynchronous: the lanes overlap on the device
                n_lane = max(r.N ** 3 for r in live[li].values())
                left = max(r.steps - r.ran for r in live[li].values())
                nrep = max(1, min(left, (chunk * n_all) // n_lane))
                batches[li].run(nrep)
                for r in live[li].values():
                    r.ran += nrep
            tc = time.perf_counter()
            host["issue"] += tc - tb
            for li in active:
                td = time.perf_counter()
                fin, _ = batches[li].poll()
                te = time.perf_counter()
                host["poll"] += te - td
                for i in fin:
                    r = live[li].pop(i)
                    r.t_final = batches[li].status(i)["t_cur"]
                    if collect:
                        r.result = batches[li].get_state(i)
                    r.done = time.perf_counter() - t0
                    batches[li].remove(i)
                    used[0] -= C.batch_bytes_needed(r.payload[0])
                    done.append(r)
                    if log:
                        log(r)
                host["collect"] += time.perf_counter() - te
        el = time.perf_counter() - t0
    finally:
        for b in batches:
            b.close()
    lat = [r.done - r.arrival for r in reqs]
    work = sum(r.work for r in reqs)
    return {"wall_s": el, "cell_steps": work, "mcell_steps_per_s": work / el / 1e6,
            "p50_s": percentile(lat, 50), "p99_s": percentile(lat, 99), "mean_s": float(np.mean(lat)),
            "host_s": {k: round(v, 4) for k, v in host.items()}}


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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["sweep", "serve"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--steps-min", type=int, default=20)
    ap.add_argument("--steps-max", type=int, default=200)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "alt"])
    ap.add_argument("--rate", type=float, default=None, help="arrivals per second (serve)")
    ap.add_argument("--load", type=float, default=0.85, help="offered load vs batch capacity when --rate is unset")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--chunk", type=int, default=4)
    ap.add_argument("--slots", type=int, default=2, help="solves resident at once (0: as memory allows)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-baseline", action="store_true")
    a = ap.parse_args(argv)
    budget = int(a.budget_gb * GB)
    if a.mode == "sweep":
        reqs = make_requests(a.n or 64, a.seed, steps=(a.steps, a.steps), precision=a.precision)
        res = run_sweep(reqs, budget)
        res["requests"] = len(reqs)
        print(json.dumps({"sweep": res}))
        return 0
    grids = (32, 48, 64, 96, 128, 160, 192)
    weights = [1.0 / g for g in grids]
    n = a.n or 128
    rate = a.rate
    if rate is None:
        # calibrate the offered load on a short batched sweep of the same mix
        probe = make_requests(min(n, 32), a.seed + 1, grids, weights, (a.steps_min, a.steps_min))
        cap = run_sweep(probe, budget)["mcell_steps_per_s"] * 1e6
        mean_work = np.mean([r.work for r in make_requests(n, a.seed, grids, weights, (a.steps_min, a.steps_max))])
        rate = a.load * cap / mean_work
    reqs = make_requests(n, a.seed, grids, weights, (a.steps_min, a.steps_max), rate=rate)
    def progress(r):
        print(f"request {r.rid} N={r.N} steps={r.steps} done", flush=True)

    res = {"requests": n, "rate_per_s": rate, "slots": a.slots,
           "served": serve(reqs, budget, a.chunk, max_live=a.slots or None, log=progress)}
    print(json.dumps(res), flush=True)
    if not a.no_baseline:
        res["sequential_fp64"] = sequential_fp64(reqs, log=progress)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
