"""cbns_face_flux_eval_noalloc (the entry point that allocates nothing on the
device) against cbns_face_flux_eval, through the C ABI: bit-identical
fluxes for every precision and direction, on two calls in a row of which
the second is smaller (the fixed scratch keeps the first call's contents,
which must not leak into the second), plus the capacity limit.  GPU only."""
import numpy as np
import pytest

from cbns3d_cart_mpi_amd import capi as C

pytestmark = pytest.mark.gpu


def _faces(rng, n, gam1=0.4):
    """Six-cell stencils of physical states, smooth and discontinuous, with
    the primitives of the two middle cells as calc_primitive forms them."""
    rho = np.exp(rng.uniform(-0.7, 0.7, (n, 6)))
    vel = rng.standard_normal((n, 6, 3)) * rng.choice([0.01, 0.3, 1.0], size=(n, 1, 1))
    p = np.exp(rng.uniform(0.0, 4.3, (n, 6)))
    jump = rng.random(n) < 0.2
    rho[jump, 3:] *= 4.0
    cells = np.empty((n, 6, 5))
    cells[..., 0] = rho
    cells[..., 1:4] = rho[..., None] * vel
    cells[..., 4] = p / gam1 + 0.5 * rho * (vel ** 2).sum(-1)
    prim = np.empty((n, 2, 5))
    for s, c in ((0, 2), (1, 3)):
        rr = cells[:, c, 0]
        uu, vv, ww = cells[:, c, 1] / rr, cells[:, c, 2] / rr, cells[:, c, 3] / rr
        prim[:, s] = np.column_stack([rr, uu, vv, ww, gam1 * (cells[:, c, 4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww))])
    return cells, prim


@pytest.mark.parametrize("precision", [C.CBNS_FP64, C.CBNS_FP32_FLUX, C.CBNS_FP64_FMA])
def test_noalloc_face_eval_equals_the_allocating_entry_point(precision):
    rng = np.random.default_rng(7 + precision)
    for d in range(3):
        for n in (C.CBNS_NOALLOC_MAX_FACES, 1237):  # second call smaller than the first
            cells, prim = _faces(rng, n)
            want = C.face_flux_eval(d, 1.4, cells, prim, precision)
            got = C.face_flux_eval_noalloc(d, 1.4, cells, prim, precision)
            # strong jumps reconstruct negative pressures on some faces, whose
            # fluxes are NaN on both paths (as in the reference arithmetic)
            assert np.array_equal(got, want, equal_nan=True), (precision, d, n)
            assert np.isfinite(want).all(axis=1).mean() > 0.8


def test_noalloc_face_eval_rejects_more_faces_than_its_scratch_holds():
    rng = np.random.default_rng(3)
    cells, prim = _faces(rng, C.CBNS_NOALLOC_MAX_FACES + 1)
    with pytest.raises(C.CbnsError) as e:
        C.face_flux_eval_noalloc(0, 1.4, cells, prim)
    assert e.value.status == C.CBNS_ERR_CAPACITY
    assert C.face_flux_eval_noalloc(0, 1.4, cells[:0], prim[:0]).shape == (0, 5)
