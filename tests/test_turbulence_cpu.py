"""Host-side diagnostics of the decaying-turbulence workload (CPU only):
the dilatation spectrum against closed forms."""
import numpy as np

from cbns3d_cart_mpi_amd.turbulence import dilatation_spectrum, precision_errors, turbulence_state


def _grid(N):
    x = -np.pi + (np.arange(N) + 0.5) * 2 * np.pi / N
    return np.meshgrid(x, x, x, indexing="ij")


def test_dilatation_spectrum_of_a_potential_mode():
    # u = a sin(m x) (rho = 1): theta = a m cos(m x), whose two Fourier
    # coefficients +-m carry (a m / 2)^2 each, so E(m) = (a m)^2 / 4
    N, a, m = 24, 0.3, 3
    Z, Y, X = _grid(N)
    Q = np.zeros((N, N, N, 5))
    Q[..., 0] = 2.0
    Q[..., 1] = 2.0 * a * np.sin(m * X)
    s = dilatation_spectrum(Q, N)
    assert abs(s[m] - (a * m) ** 2 / 4) < 1e-14
    assert np.all(np.delete(s, m) < 1e-28)


def test_dilatation_of_a_solenoidal_field_vanishes():
    Q, _ = turbulence_state(24, seed=5)
    s = dilatation_spectrum(Q, 24)
    ke = 0.5 * np.mean((Q[..., 1:4] ** 2).sum(-1))
    assert s.sum() < 1e-24 * ke


def test_precision_errors_are_zero_for_equal_states_and_scale_linearly():
    Q, _ = turbulence_state(16, seed=1)
    Z, Y, X = _grid(16)
    Q[..., 1] += 0.01 * np.sin(2 * X)  # give it some dilatation
    e0 = precision_errors(Q, Q, 16)
    assert all(v["rel_l2"] == 0.0 for v in e0.values())
    e1 = precision_errors(Q * (1 + 1e-7), Q, 16)
    for nm in ("rho", "rho_u", "rho_v", "rho_w", "E"):
        assert abs(e1[nm]["rel_l2"] - 1e-7) < 1e-12
    # u = rho u / rho is unchanged by a common factor, so theta is too
    assert e1["dilatation_spectrum"]["rel_l2"] < 1e-12
