"""The fused z-face and stage-tail pass of the single solve: the cells
beside z faces whose range checks fail are left un-updated by the pass and
finished by the cell fixup, from its list and, when more cells fail than the
list holds (65,536), by its pass over every cell.  Against the oracle on
scaled Taylor-Green states (the same flow: u, T and dt are unchanged).
GPU only."""
import numpy as np
import pytest

from cbns3d_cart_mpi_amd import capi as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def scaled_oracle(N, scale, steps):
    """The oracle's run of the Taylor-Green state of N scaled by `scale`
    (src/cbns3d.cpp:176-240 from a restart state with mu = 0)."""
    O.set_threads(16)
    ob = O.Block(O.make_params(N))
    ob.init_tgv()
    Q0 = ob.Q_interior(0) * scale
    o = O.Block(O.make_params(N))
    Qg = np.zeros((N[2] + 8, N[1] + 8, N[0] + 8, 5))
    Qg[4:-4, 4:-4, 4:-4] = Q0
    o.set_state(Qg, 0.0, mu0=0.0)
    o.run(steps)
    return Q0, o.Q_interior(0), o.t_cur


@pytest.mark.parametrize("N,dims", [((12, 12, 12), (1, 1, 1)), ((48, 48, 48), (1, 1, 1)), ((40, 36, 44), (2, 1, 2))])
def test_bit_exact_form_on_a_scaled_state_equals_the_oracle(N, dims):
    """A state scaled by 1e31 takes many faces out of the range the bit-exact
    form's unscaled quotients are checked for: its z-face cells go through
    the cell fixup (the whole 48^3 grid has more such cells than the list
    holds), and the result equals the oracle bit for bit."""
    Q0, R, t = scaled_oracle(N, 1e31, 2)
    with C.Solver(C.make_params(N), dims) as s:
        s.set_state(Q0, 0.0, mu0=0.0)
        s.run(2)
        assert s.t_cur == t
        assert np.array_equal(s.get_state(), R)


def test_tolerance_bound_form_past_the_cell_list():
    """A state scaled by 1e50 fails every face's checks in the
    tolerance-bound form: at 48^3 the 110,592 cells that the fused z pass
    leaves exceed the cell list, so the cell fixup checks every cell (and
    forms every cell's x, y and z faces); two steps within the form's
    one-step bar of the oracle (1e-12 of each field's maximum, rho w against
    the momentum scale)."""
    N = (48, 48, 48)
    Q0, R, t = scaled_oracle(N, 1e50, 2)
    with C.Solver(C.make_params(N), precision=C.CBNS_FP64_FAST) as s:
        s.set_state(Q0, 0.0, mu0=0.0)
        s.run(2)
        Q = s.get_state()
        assert s.t_cur == t
    assert np.isfinite(Q).all()
    for e in (0, 1, 2, 4):
        assert np.abs(Q[..., e] - R[..., e]).max() <= 1e-12 * np.abs(R[..., e]).max(), e
    assert np.abs(Q[..., 3] - R[..., 3]).max() <= 1e-12 * np.abs(R[..., 1:4]).max()


@pytest.mark.parametrize("precision", [C.CBNS_FP64, C.CBNS_FP64_FAST, C.CBNS_FP32_FLUX])
def test_short_runs_on_small_grids_equal_long_ones(precision):
    """The run length of the fused z pass (and the viscous chunk depth)
    shrink on grids with few columns; a 16^3 solve (runs of 2 cells) and the
    same flow in a 2 x 2 x 2 decomposition (blocks of 8^3) agree bit for bit
    with each other and, in the bit-exact form, with the oracle."""
    N = (16, 16, 16)
    with C.Solver(C.make_params(N), precision=precision) as a, \
            C.Solver(C.make_params(N), (2, 2, 2), precision=precision) as b:
        a.init_tgv()
        b.init_tgv()
        a.run(5)
        b.run(5)
        Qa = a.get_state()
        assert np.array_equal(Qa, b.get_state())
    if precision == C.CBNS_FP64:
        ob = O.Block(O.make_params(N))
        ob.init_tgv()
        ob.run(5)
        assert np.array_equal(Qa, ob.Q_interior(0))


@pytest.mark.parametrize("precision", [C.CBNS_FP64, C.CBNS_FP64_FAST, C.CBNS_FP32_FLUX])
def test_tile_shapes_give_identical_results(precision):
    """The autotuner's tile shapes (planes per viscous z-chunk, cells per run
    of the fused z pass and the x+y march) change only the work split:
    every shape gives the default's state bit for bit, including runs that
    do not divide the grid."""
    N = (40, 36, 44)
    states = []
    for shape in ((0, 0), (4, 2), (12, 5), (256, 256)):
        with C.Solver(C.make_params(N), (1, 2, 1), precision=precision) as s:
            s.set_tile_shape(*shape)
            s.init_tgv()
            s.run(3)
            states.append(s.get_state())
    for k in range(1, len(states)):
        assert np.array_equal(states[0], states[k]), k
