"""Per-face parity of the contracted FP64 form and the FP32-flux form with the
FP64 oracle (src/Block3d.cpp:586-612: calc_eigenvectors, rec_weno5, lf_flux),
through cbns_face_flux_eval.  GPU only.

Bars, on the per-component error err = |dev - ref| / (|ref| + scale_e) with
scale_e the RMS of the oracle's component e over the set of faces:
  * contracted FP64: err <= 1e-12 everywhere;
  * FP32-flux:       every component within 1e-3 and, within the form's
                     relative tolerance 1e-5, 99.9% of the components of
                     moderate and discontinuous stencils and 99% of those of
                     extreme ones (whose finite faces number ~700).  Single
                     faces of harsh random stencils reach ~1e-4: the WENO
                     weights of a stencil near a switch between sub-stencils
                     amplify the FP32 rounding of its smoothness indicators.
The extreme set reaches smoothness indicators above 1e8 (checked here from
the oracle's characteristic variables), where the FP32 WENO weights leave
their one-reciprocal form; the discontinuous set has density ratios of 4
across the face, which take the FP64 Roe average of the FP32 form."""
import numpy as np
import pytest

from cbns3d_cart_mpi_amd import capi as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GAMMA = 1.4


def _faces(rng, n, lrho=(-0.7, 0.7), lp=(0.0, 4.3), vscale=(0.01, 0.3, 1.0), gam1=GAMMA - 1.0):
    rho = np.exp(rng.uniform(*lrho, (n, 6)))
    vel = rng.standard_normal((n, 6, 3)) * rng.choice(vscale, size=(n, 1, 1))
    p = np.exp(rng.uniform(*lp, (n, 6)))
    smooth = rng.random(n) < 0.5
    rho[smooth] = rho[smooth, :1] * (1 + 0.01 * np.arange(6))
    p[smooth] = p[smooth, :1] * (1 + 0.02 * np.sin(np.arange(6)))
    jump = rng.random(n) < 0.2
    rho[jump, 3:] *= 4.0
    cells = np.empty((n, 6, 5))
    cells[..., 0] = rho
    cells[..., 1:4] = rho[..., None] * vel
    cells[..., 4] = p / gam1 + 0.5 * rho * (vel ** 2).sum(-1)
    prim = np.empty((n, 2, 5))
    for s, c in ((0, 2), (1, 3)):  # calc_primitive of cells i-1, i (src/Block3d.cpp:369-373)
        rr = cells[:, c, 0]
        uu, vv, ww = cells[:, c, 1] / rr, cells[:, c, 2] / rr, cells[:, c, 3] / rr
        prim[:, s] = np.column_stack([rr, uu, vv, ww, gam1 * (cells[:, c, 4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww))])
    return cells, prim, jump


def _oracle(d, cells, prim):
    n = np.zeros(3)
    n[d] = 1.0
    out = np.empty((cells.shape[0], 5))
    for f in range(cells.shape[0]):
        R, L = O.calc_eigenvectors(GAMMA, prim[f, 0], prim[f, 1], n)
        sL = O.rec_weno5(cells[f, 0:5], R, L, GAMMA - 1.0)
        sR = O.rec_weno5(cells[f, [5, 4, 3, 2, 1]], R, L, GAMMA - 1.0)
        out[f] = O.lf_flux(GAMMA, sR, sL, n)
    return out


def _max_is(d, cells, prim):
    """Largest WENO smoothness indicator of each face's left-biased
    reconstruction, from the oracle's characteristic variables in FP64."""
    n = np.zeros(3)
    n[d] = 1.0
    out = np.empty(cells.shape[0])
    for f in range(cells.shape[0]):
        _, L = O.calc_eigenvectors(GAMMA, prim[f, 0], prim[f, 1], n)
        w = cells[f, 0:5] @ L.T  # [cell][characteristic]
        a, b, c, dd, e = w
        is1 = 13 / 12 * (a - 2 * b + c) ** 2 + 0.25 * (a - 4 * b + 3 * c) ** 2
        is2 = 13 / 12 * (b - 2 * c + dd) ** 2 + 0.25 * (b - dd) ** 2
        is3 = 13 / 12 * (c - 2 * dd + e) ** 2 + 0.25 * (3 * c - 4 * dd + e) ** 2
        out[f] = max(is1.max(), is2.max(), is3.max())
    return out


def _rel_err(dev, ref):
    ok = np.isfinite(ref).all(axis=1)
    scale = np.sqrt(np.mean(ref[ok] ** 2, axis=0))
    err = np.abs(dev[ok] - ref[ok]) / (np.abs(ref[ok]) + scale)
    return err, ok


@pytest.mark.parametrize("d", [0, 1, 2])
def test_fma_face_flux_within_1e12_of_oracle(d):
    rng = np.random.default_rng(300 + d)
    cells, prim, _ = _faces(rng, 2000)
    ref = _oracle(d, cells, prim)
    dev = C.face_flux_eval(d, GAMMA, cells, prim, precision=C.CBNS_FP64_FMA)
    err, ok = _rel_err(dev, ref)
    print(f"FP64-FMA face d={d}: max rel err {err.max():.3e} over {ok.sum()} faces")
    assert np.isfinite(dev[ok]).all()
    assert err.max() <= 1e-12


@pytest.mark.parametrize("d", [0, 1, 2])
def test_fp32_face_flux_within_1e5_of_oracle(d):
    rng = np.random.default_rng(400 + d)
    cells, prim, jump = _faces(rng, 3000)
    ref = _oracle(d, cells, prim)
    dev = C.face_flux_eval(d, GAMMA, cells, prim, precision=C.CBNS_FP32_FLUX)
    err, ok = _rel_err(dev, ref)
    print(f"FP32 face d={d}: rel err max {err.max():.3e} (jump faces {err[jump[ok]].max():.3e}), "
          f"99.9% {np.quantile(err, 0.999):.3e}, 99% {np.quantile(err, 0.99):.3e}, median {np.median(err):.3e} "
          f"over {ok.sum()} faces")
    assert jump[ok].sum() > 300
    assert np.isfinite(dev[ok]).all()
    assert np.quantile(err, 0.999) <= 1e-5
    assert err.max() <= 1e-3


@pytest.mark.parametrize("d", [0, 2])
def test_fp32_face_flux_extreme_stencils(d):
    """Dense, fast, high-pressure stencils: smoothness indicators far past the
    range of the one-reciprocal WENO weights."""
    rng = np.random.default_rng(500 + d)
    cells, prim, _ = _faces(rng, 1500, lrho=(4.0, 8.0), lp=(8.0, 12.0), vscale=(30.0, 300.0))
    big = _max_is(d, cells, prim)
    assert (big > 1e8).mean() > 0.3, np.percentile(big, [10, 50, 90])
    ref = _oracle(d, cells, prim)
    dev = C.face_flux_eval(d, GAMMA, cells, prim, precision=C.CBNS_FP32_FLUX)
    err, ok = _rel_err(dev, ref)
    print(f"FP32 extreme face d={d}: rel err max {err.max():.3e}, 99.9% {np.quantile(err, 0.999):.3e}, "
          f"99% {np.quantile(err, 0.99):.3e}, median {np.median(err):.3e}; IS>1e8 on {(big > 1e8).mean():.2f} "
          f"of {ok.sum()} finite faces")
    assert ok.sum() > 500
    assert np.isfinite(dev[ok]).all()
    assert np.quantile(err, 0.99) <= 1e-5
    assert err.max() <= 1e-3
