"""Launch-shape autotuner for a solver context.

Every knob tuned here changes only the order in which workgroups visit the
grid or which stream a kernel runs on; results are bitwise identical for
every choice (the tests check it), so the tuner may pick freely by time.

Knobs: rows per band of the stencil kernels' work order (XCD/L2 locality),
the side-stream overlap of the viscous kernels with the face fluxes, and the
register-allocation variant (waves per SIMD) of the FP64 face kernel.
"""
import itertools
import time

BANDS = (4, 8, 16, 32, 64)


def _time_steps(solver, steps):
    solver.sync()
    t0 = time.perf_counter()
    solver.run(steps)
    return (time.perf_counter() - t0) / steps


def tune(solver, steps=3, bands=BANDS, overlaps=(True, False), occs=(2, 3)):
    """Try each (band, overlap, face occupancy) on `steps` steps of the live
    state and leave the solver on the fastest.  Returns
    (best, {choice: seconds_per_step}); the state advances by
    (steps + 1) steps per choice."""
    timings = {}
    for band, ov, occ in itertools.product(bands, overlaps, occs):
        solver.set_band(band)
        solver.set_overlap(ov)
        solver.set_face_occupancy(occ)
        _time_steps(solver, 1)  # settle after a switch
        timings[(band, ov, occ)] = _time_steps(solver, steps)
    best = min(timings, key=timings.get)
    solver.set_band(best[0])
    solver.set_overlap(best[1])
    solver.set_face_occupancy(best[2])
    return best, timings
