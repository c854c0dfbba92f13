"""Output and diagnostics beside the time step, through the C ABI (GPU only):
the asynchronous snapshot holds the state of its moment while the run goes
on, a resume with the mu sidecar continues bit for bit as the uninterrupted
run, and the device velocity spectra equal the host definition and the
closed forms of single modes."""
import numpy as np
import pytest

from cbns3d_cart_mpi_amd import capi as C
from cbns3d_cart_mpi_amd.turbulence import turbulence_state, velocity_spectra
from spectral_modes import check_single_mode_spectra, oblique_mode_state, single_mode_state

pytestmark = pytest.mark.gpu

FORMS = {"fp64": C.CBNS_FP64, "fp64_fast": C.CBNS_FP64_FAST, "fp32_flux": C.CBNS_FP32_FLUX}


def held(s):
    """Q and the primitives rho u v w p of every block, read synchronously."""
    return ([s.get_block_state(b) for b in range(s.num_blocks)],
            [[s.get_block_primitive(b, f) for f in range(5)] for b in range(s.num_blocks)])


@pytest.mark.parametrize("form", FORMS)
def test_snapshot_holds_the_state_of_its_moment(form):
    with C.Solver(C.make_params((24, 20, 16)), (2, 1, 1), FORMS[form]) as s:
        s.init_tgv()
        for steps_before, steps_after in ((3, 4), (2, 3)):  # the second snapshot overwrites the first
            s.run(steps_before)
            want_q, want_p = held(s)
            t = s.t_cur
            s.snapshot_begin()
            s.run(steps_after)  # the snapshot's copies run beside these steps
            assert s.snapshot_wait() == t
            for b in range(s.num_blocks):
                assert np.array_equal(s.snapshot_block_state(b), want_q[b])
                for f in range(5):
                    assert np.array_equal(s.snapshot_block_primitive(b, f), want_p[b][f]), (b, f)
            assert not np.array_equal(s.get_block_state(0), want_q[0])  # the run did go on


@pytest.mark.parametrize("form", ["fp64", "fp64_fast"])
def test_mu_sidecar_resume_is_bit_identical(form):
    N, dims, prec = (24, 20, 16), (2, 1, 1), FORMS[form]
    with C.Solver(C.make_params(N), dims, prec) as a:
        a.init_tgv()
        a.run(8)
        Q8, t8 = a.get_state(), a.t_cur
    with C.Solver(C.make_params(N), dims, prec) as b:
        b.init_tgv()
        b.run(4)
        Q4, t4 = b.get_state(), b.t_cur
        mus = [b.get_block_mu(k) for k in range(b.num_blocks)]
    assert all(np.all(m > 0) for m in mus)
    with C.Solver(C.make_params(N), dims, prec) as c:
        c.set_state(Q4, t4, mu0=0.0)
        for k, m in enumerate(mus):
            c.set_block_mu(k, m)
        c.run(4)
        assert c.t_cur == t8
        assert np.array_equal(c.get_state(), Q8)
    with C.Solver(C.make_params(N), dims, prec) as d:  # the reference's restart: mu = 0 at the first dt
        d.set_state(Q4, t4, mu0=0.0)
        d.run(4)
        assert d.t_cur != t8


def close(dev, host):
    """Device and host spectra agree shell by shell to 1e-12 of the largest
    value of their scale: the energy spectrum's for ke, the enstrophy
    spectrum's (the same k^2 weight) for the enstrophy and the dilatation,
    which on a solenoidal field is rounding noise."""
    scale = (host[0].max(), host[1].max(), host[1].max())
    for d, h, s in zip(dev, host, scale):
        assert np.abs(d - h).max() <= 1e-12 * s, np.abs(d - h).max() / s


def spectra_of(Q, dims, **kw):
    N = (Q.shape[2], Q.shape[1], Q.shape[0])
    with C.Solver(C.make_params(N, **kw), dims) as s:
        s.set_state(Q.reshape(-1, 5), 0.0)
        return s.spectra()


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 2, 1)])
def test_spectra_of_turbulence_equal_the_host_definition(dims):
    Q, Re = turbulence_state(32, seed=7)
    dev = spectra_of(Q, dims, Mach=1.0, Re=Re)
    host = velocity_spectra(Q, nshell=len(dev[0]))
    close(dev, host)
    assert dev[2].sum() < 1e-20 * dev[1].sum()  # the initial field is solenoidal


def test_spectra_of_a_running_taylor_green_flow_equal_the_host_definition():
    with C.Solver(C.make_params((24, 20, 16)), (2, 1, 1)) as s:
        s.init_tgv()
        s.run(5)
        Q = s.get_state()
        dev = s.spectra()
    host = velocity_spectra(Q, nshell=len(dev[0]))
    close(dev, host)
    assert dev[2].sum() > 1e-9 * dev[1].sum()  # compressible (M = 0.1): dilatation 1.6e-7 of the enstrophy


@pytest.mark.parametrize("shape,dims", [((16, 12, 10), (2, 1, 1)), ((15, 11, 9), (1, 1, 1))])
def test_spectra_of_white_noise_equal_the_host_definition(shape, dims):
    """A random compressible field, so that every mode carries energy, the
    Nyquist ones of an even grid included (numpy labels both a Nyquist mode
    and its conjugate partner -N/2, so the partner's k . u_hat is not the
    mirror of the held mode's): each device spectrum equals the host
    definition to 1e-12 of its own maximum."""
    rng = np.random.default_rng(5)
    N0, N1, N2 = shape
    Q = np.empty((N2, N1, N0, 5))
    Q[..., 0] = 1.0 + 0.1 * rng.random((N2, N1, N0))
    Q[..., 1:4] = Q[..., :1] * 0.1 * rng.standard_normal((N2, N1, N0, 3))
    Q[..., 4] = 2.5 + 0.5 * (Q[..., 1:4] ** 2).sum(-1) / Q[..., 0]
    dev = spectra_of(Q, dims)
    host = velocity_spectra(Q, nshell=len(dev[0]))
    for d, h in zip(dev, host):
        assert np.abs(d - h).max() <= 1e-12 * h.max(), np.abs(d - h).max() / h.max()


@pytest.mark.parametrize("make", [single_mode_state, oblique_mode_state])
def test_spectra_of_single_modes_equal_their_closed_forms(make):
    Q, exp = make(16, 12, 10)
    ke, en, dil = spectra_of(Q, (2, 1, 1))
    check_single_mode_spectra((ke, en, dil), exp, tol=1e-13)
