"""Manufactured periodic field with closed-form kinetic energy and enstrophy
under the 4th-order central stencil (src/calc_gradient.cpp:29-64).

Velocity is a sum of plane waves with integer wave vectors (0,1,1), (1,0,1)
and (1,1,0), each vorticity component a single cosine wave with two
correlated contributions, and the density carries the second harmonic of
each wave, so that the density weight and every term of the vorticity show
in the means:

    u = A2 sin(x+z) + A3 sin(x+y)
    v = B1 sin(y+z) + B3 sin(x+y)
    w = C1 sin(y+z) + C2 sin(x+z)
    rho = 1 + e1 cos 2(y+z) + e2 cos 2(x+z) + e3 cos 2(x+y),   p = 1

On a uniform periodic grid of spacing h the stencil
((1/12) f(x-2h) - (2/3) f(x-h) + (2/3) f(x+h) - (1/12) f(x+2h)) / h maps
sin(x + c) to D(h) cos(x + c) exactly, with D(h) = (8 sin h - sin 2h) / (6 h),
since f(x+h) - f(x-h) = 2 cos(x+c) sin h.  Hence

    omega_x = dw/dy - dv/dz = (C1 Dy - B1 Dz) cos(y+z)
    omega_y = du/dz - dw/dx = (A2 Dz - C2 Dx) cos(x+z)
    omega_z = dv/dx - du/dy = (B3 Dx - A3 Dy) cos(x+y)

Grid means of products of distinct waves vanish (no wave number reaches a
multiple of the grid size for N > 4), <sin^2> = <cos^2> = 1/2,
<cos 2t sin^2 t> = -1/4 and <cos 2t cos^2 t> = +1/4, so

    KE  = 1/2 sum_m P_m (1/2 - e_m/4),   P = (B1^2 + C1^2, A2^2 + C2^2, A3^2 + B3^2)
    ENS = 1/2 sum_m W_m^2 (1/2 + e_m/4), W = (C1 Dy - B1 Dz, A2 Dz - C2 Dx, B3 Dx - A3 Dy)
"""
import math

import numpy as np

AMP = dict(A2=0.3, A3=-0.2, B1=0.25, B3=0.15, C1=-0.35, C2=0.1)
RHO = (0.2, -0.15, 0.1)
GAMMA = 1.4


def _centres(n):
    h = 2.0 * math.pi / n
    return -math.pi + (np.arange(n) + 0.5) * h


def state(N, amp=AMP, rho_e=RHO, gamma=GAMMA):
    """Interior conserved state, AoS [k][j][i][eq], for grid N = (IM, JM, KM)."""
    x, y, z = _centres(N[0]), _centres(N[1]), _centres(N[2])
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    a = amp
    u = a["A2"] * np.sin(X + Z) + a["A3"] * np.sin(X + Y)
    v = a["B1"] * np.sin(Y + Z) + a["B3"] * np.sin(X + Y)
    w = a["C1"] * np.sin(Y + Z) + a["C2"] * np.sin(X + Z)
    e1, e2, e3 = rho_e
    rho = 1.0 + e1 * np.cos(2 * (Y + Z)) + e2 * np.cos(2 * (X + Z)) + e3 * np.cos(2 * (X + Y))
    p = np.ones_like(rho)
    Q = np.empty(rho.shape + (5,))
    Q[..., 0] = rho
    Q[..., 1] = rho * u
    Q[..., 2] = rho * v
    Q[..., 3] = rho * w
    Q[..., 4] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return Q


def stencil_factor(n):
    h = 2.0 * math.pi / n
    return (8.0 * math.sin(h) - math.sin(2.0 * h)) / (6.0 * h)


def closed_form(N, amp=AMP, rho_e=RHO, variant=None):
    """(KE, enstrophy) of the field on grid N.  variant names a plausible
    mistake to evaluate instead: "no_rho" (density weight dropped),
    "w_sign" (every w term of the vorticity with the wrong sign)."""
    a = amp
    Dx, Dy, Dz = (stencil_factor(n) for n in N)
    e = (0.0, 0.0, 0.0) if variant == "no_rho" else rho_e
    P = (a["B1"] ** 2 + a["C1"] ** 2, a["A2"] ** 2 + a["C2"] ** 2, a["A3"] ** 2 + a["B3"] ** 2)
    sw = -1.0 if variant == "w_sign" else 1.0
    W = (sw * a["C1"] * Dy - a["B1"] * Dz, a["A2"] * Dz - sw * a["C2"] * Dx, a["B3"] * Dx - a["A3"] * Dy)
    ke = 0.5 * sum(P[m] * (0.5 - e[m] / 4.0) for m in range(3))
    ens = 0.5 * sum(W[m] ** 2 * (0.5 + e[m] / 4.0) for m in range(3))
    return ke, ens
