"""Seeded compressible decaying-turbulence initial state.

A random solenoidal velocity field on the periodic [-pi, pi)^3 grid with
energy spectrum E(k) ~ k^4 exp(-2 (k/k0)^2) (k0 = 4) and uniformly random
phases, scaled to a turbulent Mach number M_t = u_rms / c (u_rms over all
three components), with uniform rho = 1 and p = 1/gamma.

Non-dimensionalisation matches the solver's (reference src/Block3d.cpp:259-262,
src/Block3d.cpp:381): with the input Mach number set to 1, gamma M^2 = gamma,
so T = gamma p / rho = 1 and the sound speed is 1.  The Reynolds number is
chosen so that the Taylor-microscale Reynolds number of the initial field,
Re_lambda = u' lambda / nu with u' = u_rms / sqrt(3),
lambda = sqrt(<u^2> / <(du/dx)^2>) (spectral derivatives, averaged over the
three components) and nu = mu / (rho Re) = 1 / Re at T = 1, equals the
requested value:  Re = Re_lambda / (u' lambda).

The generator uses numpy's PCG64 stream seeded with `seed`, so the same
seed gives the same field on every host.
"""
import numpy as np


def velocity_field(N, seed, k0=4.0):
    """(u, v, w), each [z][y][x], unscaled, zero mean, solenoidal in the
    spectral sense."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(N, 1.0 / N)
    KZ, KY, KX = np.meshgrid(k1, k1, k1, indexing="ij")
    k2 = KX * KX + KY * KY + KZ * KZ
    kk = np.sqrt(k2)
    # per-mode amplitude: E(k) / (4 pi k^2) ~ k^2 exp(-2 (k/k0)^2)
    amp = kk * np.exp(-((kk / k0) ** 2))
    # no Nyquist modes: their real part would not stay divergence-free
    if N % 2 == 0:
        nyq = (np.abs(KX) == N // 2) | (np.abs(KY) == N // 2) | (np.abs(KZ) == N // 2)
        amp[nyq] = 0.0
    r = rng.standard_normal((3, N, N, N)) + 1j * rng.standard_normal((3, N, N, N))
    K = np.stack([KX, KY, KZ])
    with np.errstate(invalid="ignore", divide="ignore"):
        kr = (K * r).sum(axis=0) / np.where(k2 > 0, k2, 1.0)
    uh = (r - K * kr) * amp  # projection onto the plane normal to k
    uh[:, 0, 0, 0] = 0.0
    u = np.real(np.fft.ifftn(uh, axes=(1, 2, 3)))
    return u


def taylor_reynolds_factor(u, N):
    """u' * lambda of a velocity field on [-pi, pi)^3 (spectral derivatives)."""
    k1 = np.fft.fftfreq(N, 1.0 / N)
    KZ, KY, KX = np.meshgrid(k1, k1, k1, indexing="ij")
    Ks = (KX, KY, KZ)
    num, den = 0.0, 0.0
    for a in range(3):
        d = np.real(np.fft.ifftn(1j * Ks[a] * np.fft.fftn(u[a])))
        num += np.mean(u[a] ** 2)
        den += np.mean(d ** 2)
    u_prime = np.sqrt(num / 3.0)
    lam = np.sqrt(num / den)
    return u_prime * lam


def turbulence_state(N, seed=0, mach_t=0.6, gamma=1.4, re_lambda=100.0, k0=4.0):
    """Returns (Q, Re): Q the interior AoS state [k][j][i][eq] (N^3 x 5) and
    the Reynolds number giving Re_lambda for it (use with Mach = 1)."""
    u = velocity_field(N, seed, k0)
    urms = np.sqrt(np.mean((u * u).sum(axis=0)))
    u *= mach_t / urms  # sound speed 1 at rho = 1, p = 1/gamma
    Re = re_lambda / taylor_reynolds_factor(u, N)
    rho = 1.0
    p = 1.0 / gamma
    Q = np.empty((N, N, N, 5))
    Q[..., 0] = rho
    Q[..., 1] = rho * u[0]
    Q[..., 2] = rho * u[1]
    Q[..., 3] = rho * u[2]
    Q[..., 4] = p / (gamma - 1.0) + 0.5 * rho * (u * u).sum(axis=0)
    return Q, float(Re)


def dilatation_spectrum(Q, N):
    """Shell spectrum of the dilatation theta = div u of an interior AoS state
    (spectral derivatives on the periodic [-pi, pi)^3 grid): E(k) for integer
    shells k = 0 .. N/2, E(k) = sum over |kvec| in [k - 1/2, k + 1/2) of
    |theta_hat|^2 / 2 with theta_hat normalised by N^3."""
    k1 = np.fft.fftfreq(N, 1.0 / N)
    KZ, KY, KX = np.meshgrid(k1, k1, k1, indexing="ij")
    th = np.zeros((N, N, N), complex)
    for a, K in enumerate((KX, KY, KZ)):
        th += 1j * K * np.fft.fftn(Q[..., 1 + a] / Q[..., 0])
    e = 0.5 * np.abs(th / N ** 3) ** 2
    shell = np.rint(np.sqrt(KX * KX + KY * KY + KZ * KZ)).astype(int).ravel()
    spec = np.bincount(shell, weights=e.ravel())
    return spec[: N // 2 + 1]


def velocity_spectra(Q, nshell=None):
    """Shell spectra of the velocity of an interior AoS state Q [k][j][i][eq]
    of any shape on the periodic [-pi, pi)^3 grid (numpy, the host
    counterpart of the device's cbns_spectra): with u_hat the FFT of
    u = (rho u) / rho (and v, w) divided by the cell count and k the integer
    wavenumbers in fftfreq order, shell s = rint(|k|) holds
    ke = sum 0.5 |u_hat|^2, enstrophy = sum 0.5 |k x u_hat|^2 and
    dilatation = sum 0.5 |k . u_hat|^2.  Returns the three arrays of
    nshell shells (default: up to the largest |k|)."""
    N2, N1, N0 = Q.shape[:3]
    kk = [np.fft.fftfreq(n, 1.0 / n) for n in (N2, N1, N0)]
    KZ, KY, KX = np.meshgrid(*kk, indexing="ij")
    K = (KX, KY, KZ)
    uh = [np.fft.fftn(Q[..., 1 + a] / Q[..., 0]) / (N0 * N1 * N2) for a in range(3)]
    ke = 0.5 * sum(np.abs(u) ** 2 for u in uh)
    cross = [K[1] * uh[2] - K[2] * uh[1], K[2] * uh[0] - K[0] * uh[2], K[0] * uh[1] - K[1] * uh[0]]
    en = 0.5 * sum(np.abs(c) ** 2 for c in cross)
    dil = 0.5 * np.abs(sum(K[a] * uh[a] for a in range(3))) ** 2
    shell = np.rint(np.sqrt(KX * KX + KY * KY + KZ * KZ)).astype(int).ravel()
    if nshell is None:
        nshell = shell.max() + 1
    out = []
    for e in (ke, en, dil):
        s = np.bincount(shell, weights=e.ravel(), minlength=nshell)
        out.append(s[:nshell])
    return tuple(out)


def precision_errors(Q32, Q64, N):
    """Per-field relative L2 (RMS difference / RMS of the FP64 field) and
    relative L-inf (max difference / max |FP64 field|) of an FP32-flux state
    against the FP64 state; for rho and E, whose fields sit on a large
    uniform background, also the RMS difference relative to the RMS of the
    field's fluctuation about its mean (rel_l2_fluct); and the
    dilatation-spectrum difference: the largest relative difference over
    shells 1 .. N/2 and the relative L2 difference of the whole spectrum."""
    fields = ("rho", "rho_u", "rho_v", "rho_w", "E")
    err = {}
    for e, nm in enumerate(fields):
        d = Q32[..., e] - Q64[..., e]
        ref = Q64[..., e]
        err[nm] = {"rel_l2": float(np.sqrt(np.mean(d * d) / np.mean(ref * ref))),
                   "rel_linf": float(np.abs(d).max() / np.abs(ref).max())}
        if nm in ("rho", "E"):
            fl = np.mean((ref - ref.mean()) ** 2)
            err[nm]["rel_l2_fluct"] = float(np.sqrt(np.mean(d * d) / fl)) if fl > 0 else None
    s32, s64 = dilatation_spectrum(Q32, N), dilatation_spectrum(Q64, N)
    rel = np.abs(s32[1:] - s64[1:]) / s64[1:]
    err["dilatation_spectrum"] = {"max_rel_shell": float(rel.max()),
                                  "rel_l2": float(np.linalg.norm(s32 - s64) / np.linalg.norm(s64))}
    return err
