"""Host logic that runs without a GPU: the launch-shape autotuners (on
stand-in solver and batch objects with a scripted clock) and bench.py's
contract pieces (profile readers, the one-GPU guard)."""
import os
import subprocess
import sys

import pytest

from cbns3d_cart_mpi_amd import autotune

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class FakeSolver:
    """Records the knobs and advances a clock by a cost that depends on them."""

    def __init__(self, cost, shape_cost=lambda shape: 0.0):
        self.cost = cost
        self.shape_cost = shape_cost
        self.band, self.overlap, self.occ, self.shape = 16, True, 2, (32, 32)
        self.steps = 0
        self.now = 0.0

    def set_band(self, b):
        self.band = b

    def set_overlap(self, on):
        self.overlap = on

    def set_face_occupancy(self, w):
        self.occ = w

    def set_viscous_workgroups(self, n):
        self.vw = n

    def set_tile_shape(self, planes, cells):
        self.shape = (planes or 32, cells or 32)

    def sync(self):
        pass

    def run(self, n):
        self.steps += n
        self.now += n * (self.cost(self.band, self.overlap, self.occ) + self.shape_cost(self.shape))


@pytest.fixture
def clock(monkeypatch):
    """Route autotune's perf_counter to the fake object's clock."""
    box = {}
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: box["obj"].now)
    return box


def test_tune_tries_every_choice_and_keeps_the_fastest(clock):
    def cost(band, overlap, occ):
        return 1.0 + abs(band - 32) / 100 + (0.0 if overlap else 0.05) + (0.02 if occ == 3 else 0.0)

    s = FakeSolver(cost, lambda shape: 0.01 * (shape != (64, 32)))
    clock["obj"] = s
    best, tms = autotune.tune(s, steps=3)
    launch = [k for k in tms if len(k) == 3]
    assert len(launch) == len(autotune.BANDS) * len(autotune.SCHEDULES) * 2
    assert len(tms) - len(launch) == len(set(autotune.SHAPES) | {(32, 32)})
    assert best == (32, "overlap", 2, (64, 32))
    assert (s.band, s.overlap, s.occ, s.shape) == (32, True, 2, (64, 32))
    assert s.steps == 4 * len(tms)  # one settling step + `steps` timed steps per choice
    assert tms[(32, "overlap", 2)] == pytest.approx(1.01)  # the first pass runs the default shape
    assert tms[(32, "overlap", 2, (64, 32))] == pytest.approx(1.0)


def test_tune_prefers_serial_within_the_margin(clock):
    """An overlapped schedule less than SERIAL_MARGIN faster than the best
    serial one does not win; one more than that faster does."""
    for gain, want in ((0.5 * autotune.SERIAL_MARGIN, "serial"), (2.0 * autotune.SERIAL_MARGIN, "overlap")):
        def cost(band, overlap, occ, gain=gain):
            return (1.0 + abs(band - 32) / 100) * ((1.0 - gain) if overlap else 1.0)

        s = FakeSolver(cost)
        clock["obj"] = s
        best, _ = autotune.tune(s, steps=3, occs=(2,))
        # the three overlapped schedules cost the same here: any of them may win
        assert best[0] == 32 and best[1].split(":")[0] == want, (gain, best)
        assert s.overlap == (want != "serial")


def test_apply_schedule_rejects_unknown_names():
    with pytest.raises(ValueError):
        autotune.apply_schedule(FakeSolver(lambda *a: 1.0), "sideways")


class FakeBatch:
    def __init__(self):
        self.band, self.ov, self.steps, self.now = 32, 0, 0, 0.0

    def set_band(self, b):
        self.band = b

    def set_overlap(self, n):
        self.ov = n

    def sync(self):
        pass

    def run(self, n):
        self.steps += n
        self.now += n * (2.0 + abs(self.band - 64) / 1000 + abs(self.ov - 256) / 10000)


def test_tune_batch_tries_every_choice_and_keeps_the_fastest(clock):
    b = FakeBatch()
    clock["obj"] = b
    best, tms = autotune.tune_batch(b, steps=2)
    assert len(tms) == len(autotune.BATCH_BANDS) * len(autotune.BATCH_OVERLAPS)
    assert best == (64, 256) and (b.band, b.ov) == (64, 256)
    assert b.steps == 3 * len(tms)  # a re-capturing step + `steps` timed steps per choice


def test_bench_reads_the_committed_profiles():
    sys.path.insert(0, ROOT)
    import bench

    traffic = bench.read_traffic()
    valu = bench.read_valu_active()
    # one launch of the tolerance-bound FP64 face kernel (the x+y march) at
    # 256^3 moves about 1.4 GB algorithmically (the state read once, each
    # cell's x+y differences written); the VALU issue fraction of a
    # VALU-bound kernel lies well inside (0.5, 1]
    assert traffic is not None and 1.3e9 < traffic < 3e9
    assert valu is not None and 0.5 < valu <= 1.0


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_fastdiv_and_the_face_decode_restated():
    """The face sweeps divide work indices by the block's row lengths and
    plane counts with the round-up multiply-high method (FastDiv in
    csrc/cbns_kernels.h) and take bands of a power-of-two height by shifts
    (face_decode); restated here, both equal the integer divisions and the
    generic band order for every numerator below 2^31 tried."""
    import random

    def make(d):
        l = 0
        while (1 << l) < d:
            l += 1
        m = ((1 << 32) * ((1 << l) - d)) // d + 1
        assert m < (1 << 32)
        return m, l

    def fdiv(n, f):
        return (((n * f[0]) >> 32) + n) >> f[1]

    rnd = random.Random(7)
    for d in list(range(1, 600)) + [rnd.randrange(1, 1 << 20) for _ in range(500)]:
        f = make(d)
        for n in [0, 1, d - 1, d, d + 1, (1 << 31) - 1] + [rnd.randrange(0, 1 << 31) for _ in range(100)]:
            assert fdiv(n, f) == n // d, (n, d)

    def band(t, ni, nj, nk, jb):  # band_decode
        i, R = t % ni, t // ni
        full = nj // jb
        if R < full * nk * jb:
            b = R // (nk * jb)
            rem = R - b * nk * jb
            k = rem // jb
            return i, b * jb + rem - k * jb, k
        last, rem = nj - full * jb, R - full * nk * jb
        k = rem // last
        return i, full * jb + rem - k * last, k

    def fast(t, ni, nj, nk, jb):  # face_decode
        jl = jb.bit_length() - 1
        R = fdiv(t, make(ni))
        i = t - R * ni
        full = nj >> jl
        if R < full * nk * jb:
            Rb = R >> jl
            b = fdiv(Rb, make(nk))
            return i, (b << jl) + (R & (jb - 1)), Rb - b * nk
        last, rem = nj - full * jb, R - full * nk * jb
        k = rem // last
        return i, full * jb + rem - k * last, k

    for ni, nj, nk in [(257, 256, 256), (21, 16, 12), (33, 17, 9), (5, 4, 4), (129, 65, 3)]:
        for jb in (1, 2, 4, 8, 16, 32, 64, 128):
            for t in range(0, ni * nj * nk, max(1, ni * nj * nk // 3000)):
                assert band(t, ni, nj, nk, jb) == fast(t, ni, nj, nk, jb), (ni, nj, nk, jb, t)
