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
//   * fused multiply-adds written out as such, with the compiler's own
//     contraction off in these functions, so that every translation unit
//     and every kernel that inlines them forms the same values.
// Each face checks the operand ranges its shortcuts need (densities and
// pressures in [2^-300, 2^300], c^2 in [2^-600, 2^600], every WENO
// denominator below 1e300) and that its flux is finite; a face failing a
// check is recomputed with the IEEE division and square root of the
// reference-order form, so no value depends on a shortcut outside its
// range.
#pragma once
#include "cbns_bodies.h"
#include "cbns_face.h"
#include "cbns_kernels.h"

namespace cbns {


// Left and right WENO5 reconstructions at one face of one characteristic
// field f[0..5] (cells i-3 .. i+2): left from f0..f4, right from f5..f1
// (src/Block3d.cpp:586-612 passes the stencils that way).  dmax collects
// the larger normalisation denominator for the range check.
//   IS_k as src/weno.cpp:33-42, with the second differences centred on
//   cells 1..4 shared by both stencils;
//   result = sum_k (d_k / IS_k^2) g_k / 6 / sum_k (d_k / IS_k^2), d = (0.1,
//   0.6, 0.3) and g_k the candidates' numerators (src/weno.cpp:50-54), times
//   10 IS_1^2 IS_2^2 IS_3^2 above and below: (P1 g1 + 6 P2 g2 + 3 P3 g3) /
//   (6 P1 + 36 P2 + 18 P3) with P1 = IS_2^2 IS_3^2, P2 = IS_1^2 IS_3^2,
//   P3 = IS_1^2 IS_2^2 (every constant exact).
__device__ __forceinline__ void weno5_side(double a, double b, double c, double d, double e, double IS1, double IS2,
                                           double IS3, double& w, double& dmax) {
#pragma clang fp contract(off)
  const double A = IS1 * IS1, B = IS2 * IS2, C = IS3 * IS3;
  const double P1 = B * C, P2 = A * C, P3 = A * B;
  const double g1 = __builtin_fma(2.0, a, __builtin_fma(-7.0, b, 11.0 * c));  // 2a - 7b + 11c
  const double g2 = __builtin_fma(12.0, d, __builtin_fma(30.0, c, -6.0 * b));  // 6 (-b + 5c + 2d)
  const double g3 = __builtin_fma(15.0, d, __builtin_fma(6.0, c, -3.0 * e));   // 3 (2c + 5d - e)
  const double num = __builtin_fma(P1, g1, __builtin_fma(P2, g2, P3 * g3));
  const double den = __builtin_fma(6.0, P1, __builtin_fma(36.0, P2, 18.0 * P3));
  w = num * rcp_nr(den);
  dmax = fmax(dmax, den);
}

// The smoothness indicators are formed times 4: 4 IS = (4 (13/12) s) s +
// 4 eps + e^2, with (4 (13/12) s) s + 4 eps in one fused multiply-add and
// the rest the reference-order operations scaled by a power of two, hence
// exact (no 0.25 e product); the weights' numerator and denominator then
// carry a factor 256 that cancels in their quotient (the reciprocal of a
// power-of-two multiple is that power's multiple), and dmax is 256 times
// the unscaled denominator.
__device__ __forceinline__ void weno5_lr(const double (&f)[6], double& wl, double& wr, double& dmax) {
#pragma clang fp contract(off)
  const double eps4 = 4.0 * 1.0e-6, c13x4 = 4.0 * (13.0 / 12.0);
  const double s1 = __builtin_fma(-2.0, f[1], f[0] + f[2]);
  const double s2 = __builtin_fma(-2.0, f[2], f[1] + f[3]);
  const double s3 = __builtin_fma(-2.0, f[3], f[2] + f[4]);
  const double s4 = __builtin_fma(-2.0, f[4], f[3] + f[5]);
  const double t1 = __builtin_fma(c13x4 * s1, s1, eps4), t2 = __builtin_fma(c13x4 * s2, s2, eps4);
  const double t3 = __builtin_fma(c13x4 * s3, s3, eps4), t4 = __builtin_fma(c13x4 * s4, s4, eps4);
  // first-difference terms (diff2, diff4, diff6 of src/weno.cpp:34-39)
  const double e1 = __builtin_fma(3.0, f[2], __builtin_fma(-4.0, f[1], f[0]));  // left:  f0 - 4 f1 + 3 f2
  const double e2 = f[1] - f[3];
  const double e3 = __builtin_fma(-4.0, f[3], __builtin_fma(3.0, f[2], f[4]));  //        3 f2 - 4 f3 + f4
  const double r1 = __builtin_fma(3.0, f[3], __builtin_fma(-4.0, f[4], f[5]));  // right: f5 - 4 f4 + 3 f3
  const double r2 = f[4] - f[2];
  const double r3 = __builtin_fma(-4.0, f[2], __builtin_fma(3.0, f[3], f[1]));  //        3 f3 - 4 f2 + f1
  const double ISl1 = __builtin_fma(e1, e1, t1), ISl2 = __builtin_fma(e2, e2, t2);
  const double ISl3 = __builtin_fma(e3, e3, t3);
  const double ISr1 = __builtin_fma(r1, r1, t4), ISr2 = __builtin_fma(r2, r2, t3);
  const double ISr3 = __builtin_fma(r3, r3, t2);
  weno5_side(f[0], f[1], f[2], f[3], f[4], ISl1, ISl2, ISl3, wl, dmax);
  weno5_side(f[5], f[4], f[3], f[2], f[1], ISr1, ISr2, ISr3, wr, dmax);
}

// Conserved state -> (1/rho, u, v, w, p, |u|^2) by calc_primitive's
// expressions (src/Block3d.cpp:369-373) with one reciprocal.
__device__ __forceinline__ void prim_fast(const double (&a)[5], double gam1, double& ir, double (&u)[3], double& p,
                                          double& k) {
#pragma clang fp contract(off)
  ir = rcp_nr(a[0]);
  u[0] = a[1] * ir;
  u[1] = a[2] * ir;
  u[2] = a[3] * ir;
  k = __builtin_fma(u[2], u[2], __builtin_fma(u[1], u[1], u[0] * u[0]));
  p = gam1 * __builtin_fma(-0.5 * a[0], k, a[4]);
}

// Face flux along axis D from the conserved values q[c][e] of cells
// i-3 .. i+2.  Returns false where a range check fails (flux then unset).
// Every fused multiply-add is written out (no contraction is left to the
// compiler), so the values do not depend on the surrounding code.
template <int D>
__device__ __forceinline__ bool face_flux_fast(const double (&q)[6][5], double gamma, double gam1, double gg1,
                                               double (&flux)[5]) {
#pragma clang fp contract(off)
  // ---- primitives of cells i-1, i and the Roe average
  double irL, irR, uL[3], uR[3], pL, pR, kL, kR;
  prim_fast(q[2], gam1, irL, uL, pL, kL);
  prim_fast(q[3], gam1, irR, uR, pR, kR);
  bool ok = in_range300(q[2][0]) && in_range300(q[3][0]) && in_range300(pL) && in_range300(pR);
  const double HL = __builtin_fma(gg1 * pL, irL, 0.5 * kL);
  const double HR = __builtin_fma(gg1 * pR, irR, 0.5 * kR);
  const double ratio = sqrt_noscale(q[3][0] * irL);
  const double inv = rcp_nr(1.0 + ratio);
  const double uu = __builtin_fma(ratio, uR[0], uL[0]) * inv;
  const double vv = __builtin_fma(ratio, uR[1], uL[1]) * inv;
  const double ww = __builtin_fma(ratio, uR[2], uL[2]) * inv;
  const double H = __builtin_fma(ratio, HR, HL) * inv;
  const double q2 = __builtin_fma(ww, ww, __builtin_fma(vv, vv, uu * uu));
  const double c2 = gam1 * __builtin_fma(-0.5, q2, H);
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
  const double K = (2.0 * gam1) * (h * h);  // (gamma - 1) / (2 c^2)
  const double hq2 = 0.5 * q2;
  const double qn = D == 0 ? uu : (D == 1 ? vv : ww);
  // frame (src/eigenvectors_roe.cpp:48-85):
  //   x: l = (-0, 0, 1), m = (0, -1, 0): ql = ww,  qm = -vv
  //   y: l = (0, 0, -1), m = (-1, 0, 0): ql = -ww, qm = -uu
  //   z: l = (0, 1, -0), m = (-1, 0, 0): ql = vv,  qm = -uu
  // ---- characteristic fields of the six cells: a = K t, b = (qn r - m_n)/(2c)
  double w[5][6];
#pragma unroll
  for (int cc = 0; cc < 6; cc++) {
    const double r = q[cc][0], mx = q[cc][1], my = q[cc][2], mz = q[cc][3], e = q[cc][4];
    const double mn = D == 0 ? mx : (D == 1 ? my : mz);
    double t = __builtin_fma(hq2, r, e);
    t = __builtin_fma(-uu, mx, t);
    t = __builtin_fma(-vv, my, t);
    t = __builtin_fma(-ww, mz, t);
    const double a = K * t;
    const double b = h * __builtin_fma(qn, r, -mn);
    w[0][cc] = a + b;
    w[1][cc] = __builtin_fma(-2.0, a, r);
    w[2][cc] = a - b;
    if (D == 0) {
      w[3][cc] = __builtin_fma(-ww, r, mz);
      w[4][cc] = __builtin_fma(vv, r, -my);
    } else if (D == 1) {
      w[3][cc] = __builtin_fma(ww, r, -mz);
      w[4][cc] = __builtin_fma(uu, r, -mx);
    } else {
      w[3][cc] = __builtin_fma(-vv, r, my);
      w[4][cc] = __builtin_fma(uu, r, -mx);
    }
  }
  double wl[5], wr[5], dmax = 0.0;
#pragma unroll
  for (int m = 0; m < 5; m++) weno5_lr(w[m], wl[m], wr[m], dmax);
  ok = ok && dmax < 256.0 * 1.0e300;  // every unscaled WENO denominator below 1e300
  // ---- back-projection (R's columns) and LF flux
  const double qc = qn * c;
  double U[2][5], F[2][5], wsd[2];
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const double* x = s == 0 ? wl : wr;
    const double sum = x[0] + x[2], dif = x[2] - x[0];
    const double rho = sum + x[1];
    double m0, m1, m2, lm;  // lm = ql x3 + qm x4
    if (D == 0) {
      m0 = __builtin_fma(c, dif, uu * rho);
      m1 = __builtin_fma(vv, rho, -x[4]);
      m2 = __builtin_fma(ww, rho, x[3]);
      lm = __builtin_fma(ww, x[3], -vv * x[4]);
    } else if (D == 1) {
      m1 = __builtin_fma(c, dif, vv * rho);
      m2 = __builtin_fma(ww, rho, -x[3]);
      m0 = __builtin_fma(uu, rho, -x[4]);
      lm = __builtin_fma(-ww, x[3], -uu * x[4]);
    } else {
      m2 = __builtin_fma(c, dif, ww * rho);
      m1 = __builtin_fma(vv, rho, x[3]);
      m0 = __builtin_fma(uu, rho, -x[4]);
      lm = __builtin_fma(vv, x[3], -uu * x[4]);
    }
    const double E = __builtin_fma(H, sum, __builtin_fma(qc, dif, __builtin_fma(hq2, x[1], lm)));
    ok = ok && in_range300(fabs(rho));
    const double ir = rcp_nr(rho);
    const double mn = D == 0 ? m0 : (D == 1 ? m1 : m2);
    const double un = mn * ir;
    const double mm = __builtin_fma(m2, m2, __builtin_fma(m1, m1, m0 * m0));
    const double p = gam1 * __builtin_fma(-0.5 * mm, ir, E);
    const double cs = sqrt_noscale((gamma * p) * ir);
    wsd[s] = fabs(un) + cs;
    U[s][0] = rho;
    U[s][1] = m0;
    U[s][2] = m1;
    U[s][3] = m2;
    U[s][4] = E;
    F[s][0] = mn;
    F[s][1] = D == 0 ? __builtin_fma(m0, un, p) : m0 * un;
    F[s][2] = D == 1 ? __builtin_fma(m1, un, p) : m1 * un;
    F[s][3] = D == 2 ? __builtin_fma(m2, un, p) : m2 * un;
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
    flux[e] = 0.5 * __builtin_fma(-ws, U[1][e] - U[0][e], F[0][e] + F[1][e]);
    chk += flux[e];
  }
  return ok && fabs(chk) < __builtin_inf();
}

// The same face with the IEEE division and square root of the
// reference-order form (face_flux without the speculative shortcuts, in a
// translation unit without contraction: the plain path's values), for
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

// ---- sweep bodies (shared by the single-solve and the batched kernels) ----
// A face whose checks fail is appended to fx for the fixup pass (its flux is
// then written there, not here).
__device__ __forceinline__ void fix_append(const FixList& fx, int b, int D, long long fi) {
  const unsigned s = atomicAdd(fx.count, 1u);
  if (s < fx.cap) fx.ent[s] = (unsigned long long)fi | ((unsigned long long)D << 40) | ((unsigned long long)b << 42);
}

template <int D>
__device__ __forceinline__ void load_face_cells(const BlockDesc& B, int which, int c0, double (&q)[6][5]) {
  const int sD = D == 0 ? 1 : (D == 1 ? B.g[0] : B.g[0] * B.g[1]);
  gptr<const double> __restrict__ Q = gmem<const double>(B.Q[which]);
  const long long sg = B.sg;
#pragma unroll
  for (int cc = 0; cc < 6; cc++)
#pragma unroll
    for (int e = 0; e < 5; e++) q[cc][e] = Q[e * sg + c0 + (cc - 3) * sD];
}

// One face per work item (x sweep): cells i-3 .. i+2 of face (i, j, k).
template <int D>
__device__ __forceinline__ void face_body_fast(const BlockDesc& B, int b, int t, int which, int jb, const Phys& ph,
                                               const FixList& fx) {
  const int ni = B.n[0] + (D == 0), nj = B.n[1] + (D == 1);
  int i, j, k;
  single_face_decode<D>(B, t, jb, i, j, k);
  const int fi = i + ni * (j + nj * k);
  double q[6][5], fl[5];
  load_face_cells<D>(B, which, gidx(B, i, j, k), q);
  if (face_flux_fast<D>(q, ph.gamma, ph.gam1, ph.gg1, fl)) {
    gptr<double> __restrict__ F = gmem(B.F[D]);
#pragma unroll
    for (int e = 0; e < 5; e++) F[e * B.sf[D] + fi] = fl[e];
  } else {
    fix_append(fx, b, D, fi);
  }
}

// Two faces adjacent along D per work item (y and z sweeps): faces m and
// m + 1 (m even) from the seven cells m-3 .. m+3, as face_pair_body.
template <int D>
__device__ __forceinline__ void face_pair_body_fast(const BlockDesc& B, int b, int t, int which, int jb,
                                                    const Phys& ph, const FixList& fx) {
  static_assert(D == 1 || D == 2, "pairs along y or z");
  const int ni = B.n[0], nj = B.n[1] + (D == 1), nk = B.n[2] + (D == 2);
  int i, j, k;
  pair_decode<D>(B, t, jb, i, j, k);
  const int c0 = gidx(B, i, j, k);
  const int sD = D == 1 ? B.g[0] : B.g[0] * B.g[1];
  gptr<const double> __restrict__ Q = gmem<const double>(B.Q[which]);
  const long long sg = B.sg;
  double q7[7][5];
#pragma unroll
  for (int cc = 0; cc < 7; cc++)
#pragma unroll
    for (int e = 0; e < 5; e++) q7[cc][e] = Q[e * sg + c0 + (cc - 3) * sD];
  gptr<double> __restrict__ F = gmem(B.F[D]);
  const long long sf = B.sf[D];
  const int sF = D == 1 ? ni : ni * nj;
  const int fi = i + ni * (j + nj * k);
#pragma unroll
  for (int h = 0; h < 2; h++) {
    if (h == 1 && (D == 1 ? j + 1 >= nj : k + 1 >= nk)) break;
    double qq[6][5], fl[5];
#pragma unroll
    for (int cc = 0; cc < 6; cc++)
#pragma unroll
      for (int e = 0; e < 5; e++) qq[cc][e] = q7[cc + h][e];
    if (face_flux_fast<D>(qq, ph.gamma, ph.gam1, ph.gg1, fl)) {
#pragma unroll
      for (int e = 0; e < 5; e++) F[e * sf + fi + h * sF] = fl[e];
    } else {
      fix_append(fx, b, D, fi + h * sF);
    }
  }
}

template <int D>
__device__ __forceinline__ void face_flux_fast_checked(const double (&q)[6][5], const Phys& ph, double (&flux)[5]) {
  if (!face_flux_fast<D>(q, ph.gamma, ph.gam1, ph.gg1, flux)) face_flux_fast_fallback<D>(q, ph, flux);
}

}  // namespace cbns
