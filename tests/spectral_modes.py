"""Velocity fields whose shell spectra (kinetic energy, enstrophy,
dilatation of cbns3d_cart_mpi_amd.turbulence.velocity_spectra and of the
device's cbns_spectra) have closed forms: a plane wave a cos(k . x) of one
velocity component has the two Fourier coefficients a/2 at +-k (amplitudes
divided by the cell count), so its shell rint(|k|) holds 0.5 x 2 x (a/2)^2 =
a^2/4 of energy, |k x e|^2 a^2/4 of enstrophy and (k . e)^2 a^2/4 of
dilatation (e the component's unit vector)."""
import numpy as np


def single_mode_state(N0, N1, N2, a=0.3, b=0.2, c=0.1, m=3, p=5, q=2):
    """rho = 1 + 0.2 sin z; u = a cos(m x) (compressive, k || u_hat),
    v = b sin(p x) (solenoidal, k = (p, 0, 0)), w = c cos(q y) (solenoidal,
    k = (0, q, 0)) on an N0 x N1 x N2 grid: the shells m, p and q hold
    ke = a^2/4, b^2/4, c^2/4, the dilatation only m^2 a^2 / 4 at m and the
    enstrophy p^2 b^2 / 4 at p and q^2 c^2 / 4 at q.  Returns (Q, expected
    (ke, enstrophy, dilatation) as {shell: value})."""
    x = [-np.pi + (np.arange(n) + 0.5) * 2 * np.pi / n for n in (N0, N1, N2)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    rho = 1.0 + 0.2 * np.sin(Z)
    Q = np.zeros((N2, N1, N0, 5))
    Q[..., 0] = rho
    Q[..., 1] = rho * a * np.cos(m * X)
    Q[..., 2] = rho * b * np.sin(p * X)
    Q[..., 3] = rho * c * np.cos(q * Y)
    Q[..., 4] = 1.0 / 0.4 + 0.5 * (Q[..., 1] ** 2 + Q[..., 2] ** 2 + Q[..., 3] ** 2) / rho
    exp = ({m: a * a / 4, p: b * b / 4, q: c * c / 4}, {p: p * p * b * b / 4, q: q * q * c * c / 4},
           {m: m * m * a * a / 4})
    return Q, exp


def check_single_mode_spectra(spectra, exp, tol=1e-14):
    """Each spectrum equals its expected {shell: value} (zero elsewhere) to
    tol times the largest expected value of all of them."""
    scale = max(v for e in exp for v in e.values())
    for s, e in zip(spectra, exp):
        want = np.zeros_like(s)
        for k, v in e.items():
            want[k] = v
        assert np.abs(s - want).max() < tol * scale, (s, want)


def oblique_mode_state(N0, N1, N2, d=(0.3, 0.2, 0.1)):
    """Three compressive modes, each with two equal velocity components
    along a diagonal k: v = w = d1 cos(y + z) (|k| = 1.41, shell 1),
    u = w = d2 cos(2x + 2z) (shell 3), u = v = d3 cos(3x + 3y) (shell 4).
    Each has k x u_hat = 0 (the two terms of one cross-product component
    cancel, so a sign slip in one shows) and k . u_hat = kappa d: shell s
    holds ke = d^2 / 2, dilatation = kappa^2 d^2 and no enstrophy."""
    x = [-np.pi + (np.arange(n) + 0.5) * 2 * np.pi / n for n in (N0, N1, N2)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    m1, m2, m3 = np.cos(Y + Z), np.cos(2 * X + 2 * Z), np.cos(3 * X + 3 * Y)
    u = d[1] * m2 + d[2] * m3
    v = d[0] * m1 + d[2] * m3
    w = d[0] * m1 + d[1] * m2
    Q = np.zeros((N2, N1, N0, 5))
    Q[..., 0] = 1.0
    Q[..., 1], Q[..., 2], Q[..., 3] = u, v, w
    Q[..., 4] = 1.0 / 0.4 + 0.5 * (u * u + v * v + w * w)
    exp = ({1: d[0] ** 2 / 2, 3: d[1] ** 2 / 2, 4: d[2] ** 2 / 2}, {},
           {1: d[0] ** 2, 3: 4 * d[1] ** 2, 4: 9 * d[2] ** 2})
    return Q, exp
