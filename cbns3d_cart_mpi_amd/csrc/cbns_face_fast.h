// Per-face inviscid arithmetic of the tolerance-bound FP64 form
// (CBNS_FP64_FAST), for gfx950.
//
// The same face flux as the reference's rec_riemann_x/y/z
// (src/Block3d.cpp:586-612: calc_eigenvectors, two rec_weno5, lf_flux),
// every operation in FP64, but restructured so that one face takes about
// half the VALU instructions of the operation-for-operation form:
//   * primitives of the two cells beside the face come from their conserved
//     values (calc_primitive's expressions, src/Block3d.cpp:369-373) with one
//     refined reciprocal of rho instead of three quotients, so the face reads
//     only the state;
//   * Roe average (src/eigenvectors_roe.cpp:87-106): H = gamma/(gamma-1) p/rho
//     + |u|^2/2 without the square root and its square, one reciprocal of
//     1 + ratio for the four weighted means, and one inverse square root of
//     c^2 giving both c and 1/(2c), hence K = (gamma-1)/(2c^2) = 2(gamma-1)
//     (1/(2c))^2;
//   * characteristic projection with the reference's L rows
//     (src/eigenvectors_roe.cpp:138-169) in factored form: with
//     t = E - u.m + q^2/2 rho, a = K t and b = (qn rho - m_n)/(2c), the
//     acoustic and entropy fields are a + b, rho - 2a, a - b, and the two
//     shear fields the frame rows (-ql, l) and (-qm, m); the scaling of every
//     field is the reference's, which matters because WENO5's epsilon is not
//     scale-free;
//   * WENO5 (src/weno.cpp:25-56) with one division: the weights d_k / IS_k^2
//     times IS_1^2 IS_2^2 IS_3^2, which cancels in the normalisation, and the
//     candidates' 1/6 folded into that one reciprocal; the second
//     differences of the left and right stencils of a face are shared;
//   * back-projection with R's columns (src/eigenvectors_roe.cpp:108-136) in
//     factored form, and the LF flux (src/lf_flux.cpp:42-97) from the
//     reconstructed conserved values: rho u = m, rho E = E and rho u H =
//     (E + p) u_n are the reference's quantities without the round trip
//     through primitives;
//   * a * b + c contracts into FMAs (this header is used only by a
//     translation unit built with -ffp-contract=fast).
// Each face checks the operand ranges its shortcuts need (densities and
// pressures in [2^-300, 2^300], c^2 in [2^-600, 2^600], every WENO
// denominator below 1e300) and that its flux is finite; a face failing a
// check is recomputed with the IEEE division and square root of the
// reference-order form, so no value depends on a shortcut outside its
// range.
#pragma once
#include "cbns_face.h"
#include "cbns_kernels.h"

namespace cbns {

// 1/b within about one ulp: hardware reciprocal and two Newton steps.
__device__ __forceinline__ double rcp_nr(double b) {
  double y = __builtin_amdgcn_rcp(b);
  y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
  y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
  return y;
}

// Left and right WENO5 reconstructions at one face of one characteristic
// field f[0..5] (cells i-3 .. i+2): left from f0..f4, right from f5..f1
// (src/Block3d.cpp:586-612 passes the stencils that way).  dmax collects
// the larger normalisation denominator for the range check.
__device__ __forceinline__ void weno5_lr(const double (&f)[6], double& wl, double& wr, double& dmax) {
  const double eps = 1.0e-6, c13 = 13.0 / 12.0;
  // second differences centred on cells 1..4, shared by both stencils
  const double s1 = f[0] - 2.0 * f[1] + f[2];
  const double s2 = f[1] - 2.0 * f[2] + f[3];
  const double s3 = f[2] - 2.0 * f[3] + f[4];
  const double s4 = f[3] - 2.0 * f[4] + f[5];
  const double t1 = c13 * s1 * s1, t2 = c13 * s2 * s2, t3 = c13 * s3 * s3, t4 = c13 * s4 * s4;
  // left (Um2..Up2 = f0..f4): diff2, diff4, diff6 of src/weno.cpp:34-39
  const double e1 = f[0] - 4.0 * f[1] + 3.0 * f[2];
  const double e2 = f[1] - f[3];
  const double e3 = 3.0 * f[2] - 4.0 * f[3] + f[4];
  // right (Um2..Up2 = f5..f1)
  const double r1 = f[5] - 4.0 * f[4] + 3.0 * f[3];
  const double r2 = f[4] - f[2];
  const double r3 = 3.0 * f[3] - 4.0 * f[2] + f[1];
  const double ISl1 = t1 + 0.25 * e1 * e1 + eps, ISl2 = t2 + 0.25 * e2 * e2 + eps, ISl3 = t3 + 0.25 * e3 * e3 + eps;
  const double ISr1 = t4 + 0.25 * r1 * r1 + eps, ISr2 = t3 + 0.25 * r2 * r2 + eps, ISr3 = t2 + 0.25 * r3 * r3 + eps;
  // weights (0.1, 0.6, 0.3) / IS^2 scaled by IS1^2 IS2^2 IS3^2; times 6
  // for the candidates' common 1/6: den = 6 (0.1 B C + 0.6 A C + 0.3 A B)
  {
    const double A = ISl1 * ISl1, B = ISl2 * ISl2, C = ISl3 * ISl3;
    const double P1 = B * C, P2 = A * C, P3 = A * B;
    const double g1 = 2.0 * f[0] - 7.0 * f[1] + 11.0 * f[2];
    const double g2 = -f[1] + 5.0 * f[2] + 2.0 * f[3];
    const double g3 = 2.0 * f[2] + 5.0 * f[3] - f[4];
    const double den = 0.6 * P1 + 3.6 * P2 + 1.8 * P3;
    const double num = 0.1 * P1 * g1 + 0.6 * P2 * g2 + 0.3 * P3 * g3;
    wl = num * rcp_nr(den);
    dmax = fmax(dmax, den);
  }
  {
    const double A = ISr1 * ISr1, B = ISr2 * ISr2, C = ISr3 * ISr3;
    const double P1 = B * C, P2 = A * C, P3 = A * B;
    const double g1 = 2.0 * f[5] - 7.0 * f[4] + 11.0 * f[3];
    const double g2 = -f[4] + 5.0 * f[3] + 2.0 * f[2];
    const double g3 = 2.0 * f[3] + 5.0 * f[2] - f[1];
    const double den = 0.6 * P1 + 3.6 * P2 + 1.8 * P3;
    const double num = 0.1 * P1 * g1 + 0.6 * P2 * g2 + 0.3 * P3 * g3;
    wr = num * rcp_nr(den);
    dmax = fmax(dmax, den);
  }
}

// Conserved state -> (1/rho, u, v, w, p) by calc_primitive's expressions
// (src/Block3d.cpp:369-373) with one reciprocal.
__device__ __forceinline__ void prim_fast(const double (&a)[5], double gam1, double& ir, double (&u)[3], double& p) {
  ir = rcp_nr(a[0]);
  u[0] = a[1] * ir;
  u[1] = a[2] * ir;
  u[2] = a[3] * ir;
  p = gam1 * (a[4] - 0.5 * a[0] * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
}

// Face flux along axis D from the conserved values q[c][e] of cells
// i-3 .. i+2.  Returns false where a range check fails (flux then unset).
template <int D>
__device__ __forceinline__ bool face_flux_fast(const double (&q)[6][5], double gamma, double gam1, double gg1,
                                               double (&flux)[5]) {
  // ---- primitives of cells i-1, i and the Roe average
  double irL, irR, uL[3], uR[3], pL, pR;
  prim_fast(q[2], gam1, irL, uL, pL);
  prim_fast(q[3], gam1, irR, uR, pR);
  bool ok = in_range300(q[2][0]) && in_range300(q[3][0]) && in_range300(pL) && in_range300(pR);
  const double HL = gg1 * pL * irL + 0.5 * (uL[0] * uL[0] + uL[1] * uL[1] + uL[2] * uL[2]);
  const double HR = gg1 * pR * irR + 0.5 * (uR[0] * uR[0] + uR[1] * uR[1] + uR[2] * uR[2]);
  const double ratio = sqrt_noscale(q[3][0] * irL);
  const double inv = rcp_nr(1.0 + ratio);
  const double uu = (uL[0] + ratio * uR[0]) * inv;
  const double vv = (uL[1] + ratio * uR[1]) * inv;
  const double ww = (uL[2] + ratio * uR[2]) * inv;
  const double H = (HL + ratio * HR) * inv;
  const double q2 = uu * uu + vv * vv + ww * ww;
  const double c2 = gam1 * (H - 0.5 * q2);
  ok = ok && c2 >= 0x1p-600 && c2 <= 0x1p600;
  // c and h = 1/(2c) from one inverse square root (two coupled Newton steps)
  double c, h;
  {
    const double y = __builtin_amdgcn_rsq(c2);
    double g = c2 * y;
    h = 0.5 * y;
    double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, c2);
    g = __builtin_fma(d, h, g);
    r = __builtin_fma(-h, g, 0.5);
    h = __builtin_fma(h, r, h);
    c = g;
  }
  const double K = 2.0 * gam1 * h * h;  // (gamma - 1) / (2 c^2)
  const double hq2 = 0.5 * q2;
  const double qn = D == 0 ? uu : (D == 1 ? vv : ww);
  // frame (src/eigenvectors_roe.cpp:48-85):
  //   x: l = (-0, 0, 1), m = (0, -1, 0): ql = ww,  qm = -vv
  //   y: l = (0, 0, -1), m = (-1, 0, 0): ql = -ww, qm = -uu
  //   z: l = (0, 1, -0), m = (-1, 0, 0): ql = vv,  qm = -uu
  // ---- characteristic fields of the six cells
  double w[5][6];
#pragma unroll
  for (int cc = 0; cc < 6; cc++) {
    const double r = q[cc][0], mx = q[cc][1], my = q[cc][2], mz = q[cc][3], e = q[cc][4];
    const double mn = D == 0 ? mx : (D == 1 ? my : mz);
    const double t = e - uu * mx - vv * my - ww * mz + hq2 * r;
    const double a = K * t;
    const double b = h * (qn * r - mn);
    w[0][cc] = a + b;
    w[1][cc] = r - 2.0 * a;
    w[2][cc] = a - b;
    if (D == 0) {
      w[3][cc] = mz - ww * r;
      w[4][cc] = vv * r - my;
    } else if (D == 1) {
      w[3][cc] = ww * r - mz;
      w[4][cc] = uu * r - mx;
    } else {
      w[3][cc] = my - vv * r;
      w[4][cc] = uu * r - mx;
    }
  }
  double wl[5], wr[5], dmax = 0.0;
#pragma unroll
  for (int m = 0; m < 5; m++) weno5_lr(w[m], wl[m], wr[m], dmax);
  ok = ok && dmax < 1.0e300;
  // ---- back-projection and LF flux
  const double qc = qn * c;
  double U[2][5], F[2][5], wsd[2];
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const double* x = s == 0 ? wl : wr;
    const double sum = x[0] + x[2], dif = x[2] - x[0];
    const double rho = sum + x[1];
    double m0 = uu * rho, m1 = vv * rho, m2 = ww * rho;
    if (D == 0) {
      m0 += c * dif;
      m1 -= x[4];
      m2 += x[3];
    } else if (D == 1) {
      m1 += c * dif;
      m2 -= x[3];
      m0 -= x[4];
    } else {
      m2 += c * dif;
      m1 += x[3];
      m0 -= x[4];
    }
    const double ql_w = D == 0 ? ww * x[3] : (D == 1 ? -ww * x[3] : vv * x[3]);
    const double qm_w = D == 0 ? -vv * x[4] : -uu * x[4];
    const double E = H * sum + qc * dif + hq2 * x[1] + ql_w + qm_w;
    ok = ok && in_range300(fabs(rho));
    const double ir = rcp_nr(rho);
    const double mn = D == 0 ? m0 : (D == 1 ? m1 : m2);
    const double un = mn * ir;
    const double p = gam1 * (E - 0.5 * (m0 * m0 + m1 * m1 + m2 * m2) * ir);
    const double cs = sqrt_noscale(gamma * p * ir);
    wsd[s] = fabs(un) + cs;
    U[s][0] = rho;
    U[s][1] = m0;
    U[s][2] = m1;
    U[s][3] = m2;
    U[s][4] = E;
    F[s][0] = mn;
    F[s][1] = m0 * un + (D == 0 ? p : 0.0);
    F[s][2] = m1 * un + (D == 1 ? p : 0.0);
    F[s][3] = m2 * un + (D == 2 ? p : 0.0);
    F[s][4] = (E + p) * un;
  }
  // src/lf_flux.cpp:56-71: the right side's speed unless the left's is
  // larger.  A side with a non-finite sound speed (a negative reconstructed
  // pressure) makes the reference's flux non-finite through its H; such a
  // face fails the check below and takes the reference-order form.
  const double ws = wsd[0] > wsd[1] ? wsd[0] : wsd[1];
  double chk = wsd[0] + wsd[1];
#pragma unroll
  for (int e = 0; e < 5; e++) {
    flux[e] = 0.5 * (F[0][e] + F[1][e] - ws * (U[1][e] - U[0][e]));
    chk += flux[e];
  }
  return ok && fabs(chk) < __builtin_inf();
}

// The same face with the IEEE division and square root of the
// reference-order form (face_flux without the speculative shortcuts), for
// faces whose range checks fail.  The primitives of cells i-1, i are
// calc_primitive's (src/Block3d.cpp:369-373).
template <int D>
__device__ __forceinline__ void face_flux_fast_fallback(const double (&q)[6][5], const Phys& ph, double (&flux)[5]) {
  double pl[5], pr[5];
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const double(&a)[5] = q[2 + s];
    double* o = s == 0 ? pl : pr;
    const double rr = a[0], uu = a[1] / rr, vv = a[2] / rr, ww = a[3] / rr;
    o[0] = rr;
    o[1] = uu;
    o[2] = vv;
    o[3] = ww;
    o[4] = ph.gam1 * (a[4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
  }
  bool ok = true;
  face_flux<D, false, false>(q, pl, pr, ph.gamma, ph.gam1, ph.rgam1, ph.fgam1, flux, ok);
}

template <int D>
__device__ __forceinline__ void face_flux_fast_checked(const double (&q)[6][5], const Phys& ph, double (&flux)[5]) {
  if (!face_flux_fast<D>(q, ph.gamma, ph.gam1, ph.gg1, flux)) face_flux_fast_fallback<D>(q, ph, flux);
}

}  // namespace cbns
