"""Launch-shape autotuner for a solver context.

Every knob tuned here changes only the order in which workgroups visit the
grid or which stream a kernel runs on; results are bitwise identical for
every choice (the tests check it), so the tuner may pick freely by time.

Knobs: rows per band of the stencil kernels' work order (XCD/L2 locality),
the register-allocation variant (waves per SIMD) of the FP64 face kernels,
and the stage schedule:
  "overlap"   viscous-flux kernel on a side stream beside the face sweeps;
  "serial"    viscous-flux kernel after the sweeps, on the same stream.
"""
import itertools
import time

BANDS = (4, 8, 16, 32, 64)
SCHEDULES = ("overlap", "serial")


def _time_steps(solver, steps):
    solver.sync()
    t0 = time.perf_counter()
    solver.run(steps)
    return (time.perf_counter() - t0) / steps


def apply_schedule(solver, sched):
    if sched not in SCHEDULES:
        raise ValueError(f"unknown schedule {sched!r}")
    solver.set_overlap(sched == "overlap")


def tune(solver, steps=3, bands=BANDS, schedules=SCHEDULES, occs=(2, 3)):
    """Try each (band, schedule, face occupancy) on `steps` steps of the live
    state and leave the solver on the fastest.  Returns
    (best, {choice: seconds_per_step}); the state advances by
    (steps + 1) steps per choice."""
    timings = {}
    for band, sched, occ in itertools.product(bands, schedules, occs):
        solver.set_band(band)
        apply_schedule(solver, sched)
        solver.set_face_occupancy(occ)
        _time_steps(solver, 1)  # settle after a switch
        timings[(band, sched, occ)] = _time_steps(solver, steps)
    best = min(timings, key=timings.get)
    solver.set_band(best[0])
    apply_schedule(solver, best[1])
    solver.set_face_occupancy(best[2])
    return best, timings
