"""Launch-shape autotuner for a solver context and for a batch.

Every knob tuned here changes only the order in which workgroups visit the
grid, how many persistent workgroups a launch has, or which stream a kernel
runs on; results are bitwise identical for every choice (the tests check
it), so the tuner may pick freely by time, on the live state.

Single context (tune): rows per band of the stencil kernels' work order
(XCD/L2 locality), the register-allocation variant (waves per SIMD) of the
FP64 face kernels, and the stage schedule:
  "overlap"   viscous-flux kernel on a side stream beside the face sweeps
              ("overlap:n": on n workgroups taking its tile units in turn);
  "serial"    viscous-flux kernel after the sweeps, on the same stream;
then, on the best of those, the tile shapes of the z-marching kernels
(planes per chunk of the viscous tile march, cells per run of the fused
z-face pass and the tolerance-bound form's x+y march).
Batch (tune_batch): rows per band, and the viscous kernels' schedule in the
step graph (0: in line; n: beside the sweeps with n persistent workgroups).
"""
import itertools
import time

BANDS = (4, 8, 16, 32, 64)
SCHEDULES = ("overlap", "overlap:128", "overlap:256", "serial")


def _time_steps(solver, steps):
    solver.sync()
    t0 = time.perf_counter()
    solver.run(steps)
    return (time.perf_counter() - t0) / steps


def apply_schedule(solver, sched):
    if sched not in SCHEDULES:
        raise ValueError(f"unknown schedule {sched!r}")
    over = sched.startswith("overlap")
    solver.set_overlap(over)
    solver.set_viscous_workgroups(int(sched.split(":")[1]) if ":" in sched else 0)


SERIAL_MARGIN = 0.03  # an overlapped schedule must beat the best serial one by this fraction
# (viscous planes per chunk, cells per run); (32, 32) is the library's default
SHAPES = ((16, 32), (32, 16), (32, 32), (32, 64), (64, 32), (64, 64))


def tune(solver, steps=3, bands=BANDS, schedules=SCHEDULES, occs=(2, 3), shapes=SHAPES):
    """Try each (band, schedule, face occupancy) on `steps` steps of the live
    state and keep the fastest, preferring the serial schedule unless an
    overlapped one is faster by more than SERIAL_MARGIN (within timing noise
    the kernels then run alone, so their own durations are what a profile of
    the step shows); then try each tile shape (viscous planes per chunk,
    cells per run) on that choice, the default (32, 32) as well.  Leaves the
    solver on the result.  Returns (best, {choice: seconds_per_step}) with
    best = (band, schedule, occupancy, (planes, cells)) and the timings of
    both passes (keys of 3 and 4 entries); the state advances by
    (steps + 1) steps per choice."""
    timings = {}
    solver.set_tile_shape(0, 0)
    for band, sched, occ in itertools.product(bands, schedules, occs):
        solver.set_band(band)
        apply_schedule(solver, sched)
        solver.set_face_occupancy(occ)
        _time_steps(solver, 1)  # settle after a switch
        timings[(band, sched, occ)] = _time_steps(solver, steps)
    best = min(timings, key=timings.get)
    serial = [k for k in timings if k[1] == "serial"]
    if serial and best[1] != "serial":
        best_serial = min(serial, key=timings.get)
        if timings[best_serial] <= timings[best] * (1.0 + SERIAL_MARGIN):
            best = best_serial
    solver.set_band(best[0])
    apply_schedule(solver, best[1])
    solver.set_face_occupancy(best[2])
    shape_t = {}
    for shape in dict.fromkeys(tuple(shapes) + ((32, 32),)):
        solver.set_tile_shape(*shape)
        _time_steps(solver, 1)
        shape_t[shape] = _time_steps(solver, steps)
        timings[best + (shape,)] = shape_t[shape]
    shape = min(shape_t, key=shape_t.get)
    solver.set_tile_shape(*shape)
    return best + (shape,), timings


BATCH_BANDS = (16, 32, 64, 128)
BATCH_OVERLAPS = (0, 128, 256, 384, 512)


def _time_batch(batch, steps):
    batch.sync()
    t0 = time.perf_counter()
    batch.run(steps)
    batch.sync()
    return (time.perf_counter() - t0) / steps


def tune_batch(batch, steps=2, bands=BATCH_BANDS, overlaps=BATCH_OVERLAPS):
    """Try each (band, viscous overlap) on `steps` steps of the live batch and
    leave it on the fastest.  Returns (best, {choice: seconds_per_step}).
    Every solve advances by (steps + 1) steps per choice (a choice's first
    step re-captures the step graph), so the solves need that many steps
    left; the steps taken are ordinary steps of the solves."""
    timings = {}
    for band, ov in itertools.product(bands, overlaps):
        batch.set_band(band)
        batch.set_overlap(ov)
        batch.run(1)  # re-captures the step graph
        timings[(band, ov)] = _time_batch(batch, steps)
    best = min(timings, key=timings.get)
    batch.set_band(best[0])
    batch.set_overlap(best[1])
    return best, timings
