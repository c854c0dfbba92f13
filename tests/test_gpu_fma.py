"""The contracted FP64 form (CBNS_FP64_FMA) against the FP64 oracle.  GPU only.

Tolerance (the plain path's stated bar): relative L-inf error <= 1e-12 after
one step and <= 1e-9 after 100 steps of the 64^3 Taylor-Green case, per
conserved field: max |difference| / max |oracle field| for rho and E, and
for the momentum components the difference relative to the momentum scale
max |rho u_i| over the three components (rho w starts identically zero, so
its own maximum after one step is not a scale); t_cur within the same
bounds.  The form is deterministic, so its decompositions agree bit for bit."""
import numpy as np
import pytest

from cbns3d_cart_mpi_amd import capi as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu
FIELDS = ("rho", "rho_u", "rho_v", "rho_w", "E")


def rel_linf(Q, R):
    mom = np.abs(R[..., 1:4]).max()
    scale = [np.abs(R[..., 0]).max(), mom, mom, mom, np.abs(R[..., 4]).max()]
    return {nm: float(np.abs(Q[..., e] - R[..., e]).max() / scale[e]) for e, nm in enumerate(FIELDS)}


@pytest.mark.parametrize("steps,tol", [(1, 1e-12), (100, 1e-9)])
def test_fma_form_tgv64_within_plain_path_tolerance(steps, tol):
    O.set_threads(16)
    b = O.Block(O.make_params(64))
    b.init_tgv()
    b.run(steps)
    with C.Solver(C.make_params(64), precision=C.CBNS_FP64_FMA) as s:
        s.init_tgv()
        s.run(steps)
        Q = s.get_state()
        t = s.t_cur
    err = rel_linf(Q, b.Q_interior(0))
    print(f"FP64-FMA TGV 64^3 {steps} steps: rel Linf {err}, t {t!r} vs {b.t_cur!r}")
    assert abs(t - b.t_cur) <= tol * b.t_cur
    for nm, e in err.items():
        assert e <= tol, (nm, e)


def test_fma_form_decompositions_identical():
    with C.Solver(C.make_params((26, 20, 16)), precision=C.CBNS_FP64_FMA) as a, \
            C.Solver(C.make_params((26, 20, 16)), (3, 2, 1), precision=C.CBNS_FP64_FMA) as b:
        a.init_tgv()
        b.init_tgv()
        a.run(3)
        b.run(3)
        assert np.array_equal(a.get_state(), b.get_state())


def test_fma_form_is_not_offered_in_the_batch():
    with C.Batch(1 << 30, capacity=2) as bt:
        with pytest.raises(C.CbnsError):
            bt.add(C.make_params(16), precision=C.CBNS_FP64_FMA)
