// Per-face inviscid arithmetic of the plain FP64 path, for gfx950.
//
// One face = Roe frame of the two adjacent cells (src/eigenvectors_roe.cpp:32-171),
// characteristic WENO5 reconstruction of the left and right states from the
// six cells around the face (src/weno.cpp:25-152), and the local
// Lax-Friedrichs flux (src/lf_flux.cpp:32-99), as rec_riemann_x/y/z drive them
// (src/Block3d.cpp:586-612).
//
// Exactness.  The library is compiled with -ffp-contract=off, so every
// expression below rounds exactly as the reference's (x86-64 SSE2, no FMA)
// does, provided the operation order is kept.  It is, with these exact
// simplifications for axis-aligned normals:
//   * n, l, m are compile-time unit vectors; the reference's normalisation
//     divides by sqrt(1) = 1, which is exact;
//   * a product with a frame component of 1 is the operand itself; a product
//     with -1 is its negation (both exact);
//   * a term whose frame component is 0 adds a signed zero, which leaves every
//     non-zero sum unchanged; only the sign of an exactly-zero result can
//     differ, and no such zero is ever divided by, so all values agree;
//   * the six cells of a face are projected once each with the face's L and
//     shared by the left and right stencils — the reference projects the same
//     cells with the same L in the same order, once per stencil;
//   * the characteristic fields are streamed m = 0..4, accumulating R's column
//     m into both reconstructed states, which adds R[li][m] * w_m in the same
//     order m = 0..4 as the reference's back-projection loop;
//   * the LF wave speed |ws . n| reduces to the normal component, the only
//     non-zero term;
//   * the face kernels run the speculative form (SPEC, below): quotients and
//     square roots whose operand ranges each face checks take the IEEE
//     sequences without their range scaling and special-case fix-ups, which
//     cannot act on such operands; a face failing a check is recomputed with
//     the IEEE forms.
#pragma once
#include <hip/hip_runtime.h>

namespace cbns {

// x / 6.0, correctly rounded, in three operations.  y = RN(1/6) has relative
// error 2^-54, so q0 = RN(x y) is within one ulp of x/6; the remainder
// r = x - 6 q0 is then exact in one FMA, and RN(q0 + r y) is the correctly
// rounded quotient (Markstein's correction theorem).  x/6 is never a
// rounding tie (that would need a 55-bit x).  Differs from IEEE division only
// in the sign of a zero result and for infinite x, neither of which reaches
// it here.
__device__ __forceinline__ double div6(double x) {
  const double y = 1.0 / 6.0;
  const double q0 = x * y;
  const double r = __builtin_fma(-q0, 6.0, x);
  return __builtin_fma(r, y, q0);
}

// x / h for a run-time constant h with y = RN(1/h): the same correction as
// div6, valid (and then correctly rounded) when y is within 2^-54 relative of
// 1/h, which the host establishes exactly before setting `fast`.  Otherwise
// the IEEE division.  `fast` is uniform across the launch.
__device__ __forceinline__ double divc(double x, double h, double y, int fast) {
  if (fast) {
    const double q0 = x * y;
    const double r = __builtin_fma(-q0, h, x);
    return __builtin_fma(r, y, q0);
  }
  return x / h;
}

// a / b.  FMA = false: the IEEE quotient.  FMA = true (the contracted FP64
// form, and the checked operands of the plain path's speculative form): the
// same reciprocal refinement and one-FMA correction as the compiler's IEEE
// sequence without its range scaling and special-case fix-up, so the
// quotient is still correctly rounded for operands of moderate magnitude
// (nonzero, |a|, |b| between about 2^-300 and 2^300).  The contracted form
// takes sqrt_noscale for its square roots likewise.
template <bool FMA>
__device__ __forceinline__ double qdiv(double a, double b) {
  if constexpr (!FMA) {
    return a / b;
  } else {
    double y = __builtin_amdgcn_rcp(b);
    y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
    y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
    const double q = a * y;
    return __builtin_fma(__builtin_fma(-b, q, a), y, q);
  }
}

// 1/b within about one ulp: hardware reciprocal and two Newton steps (the
// tolerance-bound FP64 form; not correctly rounded).
__device__ __forceinline__ double rcp_nr(double b) {
  double y = __builtin_amdgcn_rcp(b);
  y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
  y = __builtin_fma(y, __builtin_fma(-b, y, 1.0), y);
  return y;
}

// sqrt(x) as the compiler's IEEE sequence computes it (v_rsq_f64, then two
// Newton-Raphson corrections of g ~ sqrt(x) and h ~ 1/(2 sqrt(x))), without
// its input scaling for x < 2^-767 and its select of x itself for zero and
// infinity: the same value for 2^-767 <= x < infinity.
__device__ __forceinline__ double sqrt_noscale(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = y * 0.5;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// Speculative form of the plain path (SPEC).  The compiler's IEEE sequence
// for a / b is v_div_scale x2, v_rcp_f64, two Newton steps, q = a y,
// fma(fma(-b, q, a), y, q) in v_div_fmas, then v_div_fixup.  The scaling is
// the identity and the fix-up passes the quotient through unless an operand
// or the quotient is zero, infinite, NaN or near the ends of the exponent
// range, so qdiv<true> gives the IEEE quotient bit for bit when the operands
// are moderate; likewise sqrt_noscale.  SPEC takes those forms for the
// quotients and roots whose operands it checks, and clears ok where a check
// fails; the caller recomputes such a face with the IEEE forms.  Ranges used:
//   positive operands in [2^-300, 2^300]: quotients of two lie in
//     [2^-600, 2^600], exponent gap < 768, numerator far from 2^-970;
//   WENO weights c / IS^2 with IS in [1e-6, 2^200): divisor in
//     [1e-12, 2^400], quotient in [2^-404, 2^40];
//   WENO final quotient: divisor (the weight sum) in [2^-404, 2^41],
//     numerator 0 or of magnitude in [2^-500, 2^300].  A numerator of -0
//     gives +0 where IEEE division gives -0 (the only difference).
__device__ __forceinline__ bool in_range300(double x) { return x >= 0x1p-300 && x <= 0x1p300; }

// src/weno.cpp:25-56
template <bool FMA = false, bool SPEC = false>
__device__ __forceinline__ double weno5(double Um2, double Um1, double U0, double Up1, double Up2, bool& ok) {
  const double eps = 1.0e-6;
  double diff1 = Um2 - 2.0 * Um1 + U0;
  double diff2 = Um2 - 4.0 * Um1 + 3.0 * U0;
  double diff3 = Um1 - 2.0 * U0 + Up1;
  double diff4 = Um1 - Up1;
  double diff5 = U0 - 2.0 * Up1 + Up2;
  double diff6 = 3.0 * U0 - 4.0 * Up1 + Up2;
  double IS1 = (13.0 / 12.0) * diff1 * diff1 + 0.25 * diff2 * diff2 + eps;
  double IS2 = (13.0 / 12.0) * diff3 * diff3 + 0.25 * diff4 * diff4 + eps;
  double IS3 = (13.0 / 12.0) * diff5 * diff5 + 0.25 * diff6 * diff6 + eps;
  constexpr bool F = FMA || SPEC;
  double w1 = qdiv<F>(0.1, IS1 * IS1);
  double w2 = qdiv<F>(0.6, IS2 * IS2);
  double w3 = qdiv<F>(0.3, IS3 * IS3);
  double sum = w1 + w2 + w3;
  double f1 = div6(2.0 * Um2 - 7.0 * Um1 + 11.0 * U0);
  double f2 = div6(-Um1 + 5.0 * U0 + 2.0 * Up1);
  double f3 = div6(2.0 * U0 + 5.0 * Up1 - Up2);
  const double num = w1 * f1 + w2 * f2 + w3 * f3;
  if constexpr (SPEC) {
    const double an = fabs(num);
    ok = ok && IS1 < 0x1p200 && IS2 < 0x1p200 && IS3 < 0x1p200 && an <= 0x1p300 &&
         (an >= 0x1p-500 || an == 0.0);
  }
  return qdiv<F>(num, sum);
}

// Reconstructed conserved state -> primitive (src/weno.cpp:146-150).
template <bool FMA = false>
__device__ __forceinline__ void cons_to_prim(const double (&a)[5], double gam1, double (&s)[5]) {
  double rho = a[0];
  double u = qdiv<FMA>(a[1], rho);
  double v = qdiv<FMA>(a[2], rho);
  double w = qdiv<FMA>(a[3], rho);
  s[0] = rho;
  s[1] = u;
  s[2] = v;
  s[3] = w;
  s[4] = gam1 * (a[4] - 0.5 * rho * (u * u + v * v + w * w));
}

// Face flux along axis D (0 = x, 1 = y, 2 = z).
//   q[c][e]  conserved values of cells i-3 .. i+2 along D (c = 0..5)
//   pl, pr   rho, u, v, w, p of cells i-1 and i (calc_primitive of the same Q)
// FMA = true is the contracted FP64 form: the same expressions in a
// translation unit that lets the compiler fuse a * b + c, with qdiv<true>.
template <int D, bool FMA = false, bool SPEC = false>
__device__ __forceinline__ void face_flux(const double (&q)[6][5], const double (&pl)[5],
                                          const double (&pr)[5], double gamma, double gam1,
                                          double rg1, int fg1, double (&flux)[5], bool& ok) {
  constexpr bool F = FMA || SPEC;
  const auto root = [](double x) { return F ? sqrt_noscale(x) : sqrt(x); };
  if constexpr (SPEC)
    ok = in_range300(pl[0]) && in_range300(pl[4]) && in_range300(pr[0]) && in_range300(pr[4]) &&
         gamma >= 0x1p-100 && gamma <= 0x1p100 && gam1 >= 0x1p-100 && gam1 <= 0x1p100;
  // ---- Roe average (src/eigenvectors_roe.cpp:87-106)
  double c_L = root(qdiv<F>(gamma * pl[4], pl[0]));
  double H_L = divc(c_L * c_L, gam1, rg1, fg1) + 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
  double c_R = root(qdiv<F>(gamma * pr[4], pr[0]));
  double H_R = divc(c_R * c_R, gam1, rg1, fg1) + 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
  double ratio = root(qdiv<F>(pr[0], pl[0]));
  double uu = qdiv<FMA>(pl[1] + ratio * pr[1], 1.0 + ratio);
  double vv = qdiv<FMA>(pl[2] + ratio * pr[2], 1.0 + ratio);
  double ww = qdiv<FMA>(pl[3] + ratio * pr[3], 1.0 + ratio);
  double H = qdiv<FMA>(H_L + ratio * H_R, 1.0 + ratio);
  double q2 = uu * uu + vv * vv + ww * ww;
  const double c2 = gam1 * (H - 0.5 * q2);
  if constexpr (SPEC) ok = ok && c2 >= 0x1p-600 && c2 <= 0x1p600;  // c in [2^-300, 2^300]
  double c = root(c2);
  // frame (n, l, m) per axis (src/eigenvectors_roe.cpp:48-85):
  //   x: n=(1,0,0) l=(-0,0,1) m=(0,-1,0)   qn=uu ql=ww  qm=-vv
  //   y: n=(0,1,0) l=(0,0,-1) m=(-1,0,0)   qn=vv ql=-ww qm=-uu
  //   z: n=(0,0,1) l=(0,1,-0) m=(-1,0,0)   qn=ww ql=vv  qm=-uu
  const double qn = D == 0 ? uu : (D == 1 ? vv : ww);

  // ---- L rows (src/eigenvectors_roe.cpp:138-169)
  double K = qdiv<F>(gam1, 2.0 * c * c);
  double one_2c = qdiv<F>(0.5, c);
  double L00 = 0.5 * K * q2 + one_2c * qn;
  double L10 = 1.0 - K * q2;
  double L20 = 0.5 * K * q2 - one_2c * qn;
  double Ku = K * uu, Kv = K * vv, Kw = K * ww;
  double L01 = D == 0 ? -(Ku + one_2c) : -Ku;
  double L02 = D == 1 ? -(Kv + one_2c) : -Kv;
  double L03 = D == 2 ? -(Kw + one_2c) : -Kw;
  double L21 = D == 0 ? -(Ku - one_2c) : -Ku;
  double L22 = D == 1 ? -(Kv - one_2c) : -Kv;
  double L23 = D == 2 ? -(Kw - one_2c) : -Kw;
  double L11 = 2.0 * K * uu, L12 = 2.0 * K * vv, L13 = 2.0 * K * ww, L14 = -2.0 * K;

  // ---- R entries that are not 0 or +-1 (src/eigenvectors_roe.cpp:108-136)
  double Rm0[3], Rp2[3];  // velocity rows of columns 0 and 2
  Rm0[0] = D == 0 ? uu - c : uu;
  Rm0[1] = D == 1 ? vv - c : vv;
  Rm0[2] = D == 2 ? ww - c : ww;
  Rp2[0] = D == 0 ? uu + c : uu;
  Rp2[1] = D == 1 ? vv + c : vv;
  Rp2[2] = D == 2 ? ww + c : ww;
  double R40 = H - qn * c;
  double R41 = 0.5 * q2;
  double R42 = H + qn * c;

  double aL[5], aR[5];  // reconstructed conserved states (back-projection accumulators)
#pragma unroll
  for (int m = 0; m < 5; m++) {
    double fc[6];
#pragma unroll
    for (int cc = 0; cc < 6; cc++) {
      const double f0 = q[cc][0], f1 = q[cc][1], f2 = q[cc][2], f3 = q[cc][3], f4 = q[cc][4];
      if (m == 0) {
        fc[cc] = L00 * f0 + L01 * f1 + L02 * f2 + L03 * f3 + K * f4;
      } else if (m == 1) {
        fc[cc] = L10 * f0 + L11 * f1 + L12 * f2 + L13 * f3 + L14 * f4;
      } else if (m == 2) {
        fc[cc] = L20 * f0 + L21 * f1 + L22 * f2 + L23 * f3 + K * f4;
      } else if (m == 3) {  // row (-ql, l)
        fc[cc] = D == 0 ? f3 - ww * f0 : (D == 1 ? ww * f0 - f3 : f2 - vv * f0);
      } else {  // row (-qm, m)
        fc[cc] = D == 0 ? vv * f0 - f2 : uu * f0 - f1;
      }
    }
    // left state from cells i-3..i+1, right state from cells i+2..i-2
    const double wl = weno5<FMA, SPEC>(fc[0], fc[1], fc[2], fc[3], fc[4], ok);
    const double wr = weno5<FMA, SPEC>(fc[5], fc[4], fc[3], fc[2], fc[1], ok);
    if (m == 0) {
      aL[0] = wl;              aR[0] = wr;
      aL[1] = Rm0[0] * wl;     aR[1] = Rm0[0] * wr;
      aL[2] = Rm0[1] * wl;     aR[2] = Rm0[1] * wr;
      aL[3] = Rm0[2] * wl;     aR[3] = Rm0[2] * wr;
      aL[4] = R40 * wl;        aR[4] = R40 * wr;
    } else if (m == 1) {
      aL[0] = aL[0] + wl;      aR[0] = aR[0] + wr;
      aL[1] = aL[1] + uu * wl; aR[1] = aR[1] + uu * wr;
      aL[2] = aL[2] + vv * wl; aR[2] = aR[2] + vv * wr;
      aL[3] = aL[3] + ww * wl; aR[3] = aR[3] + ww * wr;
      aL[4] = aL[4] + R41 * wl; aR[4] = aR[4] + R41 * wr;
    } else if (m == 2) {
      aL[0] = aL[0] + wl;          aR[0] = aR[0] + wr;
      aL[1] = aL[1] + Rp2[0] * wl; aR[1] = aR[1] + Rp2[0] * wr;
      aL[2] = aL[2] + Rp2[1] * wl; aR[2] = aR[2] + Rp2[1] * wr;
      aL[3] = aL[3] + Rp2[2] * wl; aR[3] = aR[3] + Rp2[2] * wr;
      aL[4] = aL[4] + R42 * wl;    aR[4] = aR[4] + R42 * wr;
    } else if (m == 3) {  // column (0, l, ql)
      if (D == 0) {
        aL[3] = aL[3] + wl;      aR[3] = aR[3] + wr;
        aL[4] = aL[4] + ww * wl; aR[4] = aR[4] + ww * wr;
      } else if (D == 1) {
        aL[3] = aL[3] - wl;      aR[3] = aR[3] - wr;
        aL[4] = aL[4] - ww * wl; aR[4] = aR[4] - ww * wr;
      } else {
        aL[2] = aL[2] + wl;      aR[2] = aR[2] + wr;
        aL[4] = aL[4] + vv * wl; aR[4] = aR[4] + vv * wr;
      }
    } else {  // column (0, m, qm)
      if (D == 0) {
        aL[2] = aL[2] - wl;      aR[2] = aR[2] - wr;
        aL[4] = aL[4] - vv * wl; aR[4] = aR[4] - vv * wr;
      } else {
        aL[1] = aL[1] - wl;      aR[1] = aR[1] - wr;
        aL[4] = aL[4] - uu * wl; aR[4] = aR[4] - uu * wr;
      }
    }
  }

  double sL[5], sR[5];
  cons_to_prim<FMA>(aL, gam1, sL);
  cons_to_prim<FMA>(aR, gam1, sR);

  // ---- local Lax-Friedrichs flux (src/lf_flux.cpp:32-99)
  if constexpr (SPEC) ok = ok && in_range300(sR[0]) && in_range300(sR[4]) && in_range300(sL[0]) && in_range300(sL[4]);
  double cR = root(qdiv<F>(gamma * sR[4], sR[0]));
  double HR = divc(cR * cR, gam1, rg1, fg1) + 0.5 * (sR[1] * sR[1] + sR[2] * sR[2] + sR[3] * sR[3]);
  double rER = divc(sR[4], gam1, rg1, fg1) + 0.5 * sR[0] * (sR[1] * sR[1] + sR[2] * sR[2] + sR[3] * sR[3]);
  double cL = root(qdiv<F>(gamma * sL[4], sL[0]));
  double HL = divc(cL * cL, gam1, rg1, fg1) + 0.5 * (sL[1] * sL[1] + sL[2] * sL[2] + sL[3] * sL[3]);
  double rEL = divc(sL[4], gam1, rg1, fg1) + 0.5 * sL[0] * (sL[1] * sL[1] + sL[2] * sL[2] + sL[3] * sL[3]);
  const double qR = sR[1 + D], qL = sL[1 + D];
  double ws = fabs(qR) + cR;
  double wsL = fabs(qL) + cL;
  if (wsL > ws) ws = wsL;

  double diss[5], fL[5], fR[5];
  diss[0] = ws * (sR[0] - sL[0]);
  diss[1] = ws * (sR[0] * sR[1] - sL[0] * sL[1]);
  diss[2] = ws * (sR[0] * sR[2] - sL[0] * sL[2]);
  diss[3] = ws * (sR[0] * sR[3] - sL[0] * sL[3]);
  diss[4] = ws * (rER - rEL);
  const double mL = sL[0] * qL, mR = sR[0] * qR;
  fL[0] = mL;
  fR[0] = mR;
#pragma unroll
  for (int a = 1; a <= 3; a++) {
    fL[a] = (a == 1 + D) ? mL * sL[a] + sL[4] : mL * sL[a];
    fR[a] = (a == 1 + D) ? mR * sR[a] + sR[4] : mR * sR[a];
  }
  fL[4] = mL * HL;
  fR[4] = mR * HR;
#pragma unroll
  for (int e = 0; e < 5; e++) flux[e] = 0.5 * (fL[e] + fR[e] - diss[e]);
}


// ---------------------------------------------------------------------------
// FP32-flux form.  The characteristic projections, WENO5 (smoothness
// indicators, weights, candidates) and the LF flux run in FP32 on deviations
// from a constant reference state Ug = (rho_g, 0, 0, 0, E_g) of the solve:
//   * WENO5 is translation-equivariant and the projection is linear, so
//     reconstructing U - Ug and adding Ug back is the same reconstruction;
//   * the flux is returned minus F(Ug) = (0, p_g n, 0), a constant that
//     cancels exactly in the flux difference of the update.
// Large backgrounds (p ~ 1/(gamma M^2), E) therefore never meet FP32
// rounding directly.  Roe averages are taken in FP32 except at
// near-vacuum or strong-gradient faces, where they are taken in FP64.
// Unlike the plain path, this form lets the compiler contract a * b + c
// into fused multiply-adds.
struct RefState {
  double rho, E, p;  // rho_g, E_g, p_g = (gamma - 1) E_g
};

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// FP32 arithmetic contracts a * b + c within each expression only (clang's
// "on"), so the fused operations are fixed by the source and do not depend
// on the kernel these functions are inlined into.
// One side of a face's WENO5 (src/weno.cpp:25-56) in FP32 from the
// candidate stencil a..e and its three smoothness indicators (here 4 IS).
// Weights c_k / IS_k^2 scaled by IS1^2 IS2^2 IS3^2 (the scale cancels in
// the normalisation): one reciprocal instead of four.  IS >= eps keeps the
// products from underflowing; while den < 256e30 (every IS below ~1e7, so
// |f_k| below ~1e4) neither den nor the numerator can overflow.  Past that
// (or for a NaN) the unscaled weights are used.
__device__ __forceinline__ float weno5f_side(float a, float b, float c, float d, float e, float IS1, float IS2,
                                             float IS3) {
#pragma clang fp contract(on)
  const float f1 = 2.0f * a - 7.0f * b + 11.0f * c;
  const float f2 = -b + 5.0f * c + 2.0f * d;
  const float f3 = 2.0f * c + 5.0f * d - e;
  const float A = IS1 * IS1, B = IS2 * IS2, C = IS3 * IS3;
  const float s1 = 0.1f * (B * C), s2 = 0.6f * (A * C), s3 = 0.3f * (A * B);
  const float den = 6.0f * (s1 + s2 + s3);
  if (den < 256.0f * 1.0e30f) return (s1 * f1 + s2 * f2 + s3 * f3) * frcp(den);
  const float w1 = 0.1f * frcp(A), w2 = 0.6f * frcp(B), w3 = 0.3f * frcp(C);
  return (w1 * f1 + w2 * f2 + w3 * f3) * frcp(6.0f * (w1 + w2 + w3));
}

// Left and right FP32 WENO5 reconstructions at one face of one
// characteristic field f[0..5] (cells i-3 .. i+2): left from f0..f4, right
// from f5..f1, the second differences centred on cells 1..4 shared by both
// sides, and the smoothness indicators formed times 4 (4 IS = (4 (13/12) s)
// s + 4 eps + e^2; the factor cancels in the weights).
__device__ __forceinline__ void weno5f_lr(const float (&f)[6], float& wl, float& wr) {
#pragma clang fp contract(on)
  const float eps4 = 4.0f * 1.0e-6f, c13x4 = 4.0f * (13.0f / 12.0f);
  const float s1 = f[0] - 2.0f * f[1] + f[2], s2 = f[1] - 2.0f * f[2] + f[3];
  const float s3 = f[2] - 2.0f * f[3] + f[4], s4 = f[3] - 2.0f * f[4] + f[5];
  const float t1 = c13x4 * s1 * s1 + eps4, t2 = c13x4 * s2 * s2 + eps4;
  const float t3 = c13x4 * s3 * s3 + eps4, t4 = c13x4 * s4 * s4 + eps4;
  const float e1 = f[0] - 4.0f * f[1] + 3.0f * f[2], e2 = f[1] - f[3], e3 = 3.0f * f[2] - 4.0f * f[3] + f[4];
  const float r1 = f[5] - 4.0f * f[4] + 3.0f * f[3], r2 = f[4] - f[2], r3 = 3.0f * f[3] - 4.0f * f[2] + f[1];
  wl = weno5f_side(f[0], f[1], f[2], f[3], f[4], e1 * e1 + t1, e2 * e2 + t2, e3 * e3 + t3);
  wr = weno5f_side(f[5], f[4], f[3], f[2], f[1], r1 * r1 + t4, r2 * r2 + t3, r3 * r3 + t2);
}

// Deviation of one cell's conserved state from the reference state: FP64
// difference, then FP32.
__device__ __forceinline__ void dev_f32(const double (&q)[5], const RefState& g, float (&d)[5]) {
  d[0] = (float)(q[0] - g.rho);
  d[1] = (float)q[1];
  d[2] = (float)q[2];
  d[3] = (float)q[3];
  d[4] = (float)(q[4] - g.E);
}

// FP32-flux face flux from the deviations d of cells i-3 .. i+2 and the
// stored (float) primitives pf, qf = rho, u, v, w, p of cells i-1 and i.
template <int D>
__device__ __forceinline__ void face_flux_f32_dev(const float (&d)[6][5], const float (&pf)[5],
                                                  const float (&qf)[5], double gamma, double gam1,
                                                  const RefState& g, float (&flux)[5]) {
#pragma clang fp contract(on)
  const float gf = (float)gamma, g1 = (float)gam1;
  // ---- Roe average
  float uu, vv, ww, H, c;
  const float rl = pf[0], rrf = qf[0];
  const float irl = frcp(rl), ratf = rrf * irl;
  const float rho_min = (float)(1.0e-3 * g.rho);
  const bool wide = !(ratf > 0.5f && ratf < 2.0f) || rl < rho_min || rrf < rho_min;
  if (wide) {
    double pl[5], pr[5];
#pragma unroll
    for (int e = 0; e < 5; e++) {
      pl[e] = (double)pf[e];
      pr[e] = (double)qf[e];
    }
    const double rr = pr[0] / pl[0];
    double c_L = sqrt(gamma * pl[4] / pl[0]);
    double H_L = c_L * c_L / gam1 + 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    double c_R = sqrt(gamma * pr[4] / pr[0]);
    double H_R = c_R * c_R / gam1 + 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    double ratio = sqrt(rr);
    double inv = 1.0 / (1.0 + ratio);
    double u_ = (pl[1] + ratio * pr[1]) * inv, v_ = (pl[2] + ratio * pr[2]) * inv;
    double w_ = (pl[3] + ratio * pr[3]) * inv, H_ = (H_L + ratio * H_R) * inv;
    uu = (float)u_;
    vv = (float)v_;
    ww = (float)w_;
    H = (float)H_;
    c = (float)sqrt(gam1 * (H_ - 0.5 * (u_ * u_ + v_ * v_ + w_ * w_)));
  } else {
    const float ul = pf[1], vl = pf[2], wl = pf[3], plf = pf[4];
    const float ur = qf[1], vr = qf[2], wr = qf[3], prf = qf[4];
    const float gg1 = gf * frcp(g1);
    const float HL = gg1 * plf * irl + 0.5f * (ul * ul + vl * vl + wl * wl);
    const float HR = gg1 * prf * frcp(rrf) + 0.5f * (ur * ur + vr * vr + wr * wr);
    const float ratio = fsqrt(ratf);
    const float inv = frcp(1.0f + ratio);
    uu = (ul + ratio * ur) * inv;
    vv = (vl + ratio * vr) * inv;
    ww = (wl + ratio * wr) * inv;
    H = (HL + ratio * HR) * inv;
    c = fsqrt(g1 * (H - 0.5f * (uu * uu + vv * vv + ww * ww)));
  }
  const float q2 = uu * uu + vv * vv + ww * ww;
  const float qn = D == 0 ? uu : (D == 1 ? vv : ww);
  const float ic = frcp(c);
  const float K = 0.5f * g1 * ic * ic;  // (gamma - 1) / (2 c^2)
  const float one_2c = 0.5f * ic;
  const float hq2 = 0.5f * q2;
  // characteristic fields of the six cells' deviations with the reference's
  // L rows (src/eigenvectors_roe.cpp:138-169) in factored form: with
  // t = E - u.m + q^2/2 rho, a = K t and b = (qn rho - m_n)/(2c), the
  // acoustic and entropy fields are a + b, rho - 2a, a - b, the shear fields
  // the frame rows
  float w[5][6];
#pragma unroll
  for (int cc = 0; cc < 6; cc++) {
    const float r = d[cc][0], mx = d[cc][1], my = d[cc][2], mz = d[cc][3], e = d[cc][4];
    const float mn = D == 0 ? mx : (D == 1 ? my : mz);
    const float t = e + hq2 * r - uu * mx - vv * my - ww * mz;
    const float a = K * t;
    const float b = one_2c * (qn * r - mn);
    w[0][cc] = a + b;
    w[1][cc] = r - 2.0f * a;
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
  float xl[5], xr[5];
#pragma unroll
  for (int m = 0; m < 5; m++) weno5f_lr(w[m], xl[m], xr[m]);
  // back-projection with R's columns (src/eigenvectors_roe.cpp:108-136) in
  // factored form: the deviations of rho, rho u, rho v, rho w, E
  const float qc = qn * c;
  float aL[5], aR[5];
#pragma unroll
  for (int sd = 0; sd < 2; sd++) {
    const float* x = sd == 0 ? xl : xr;
    float* o = sd == 0 ? aL : aR;
    const float sum = x[0] + x[2], dif = x[2] - x[0];
    const float rho = sum + x[1];
    float lm;
    if (D == 0) {
      o[1] = uu * rho + c * dif;
      o[2] = vv * rho - x[4];
      o[3] = ww * rho + x[3];
      lm = ww * x[3] - vv * x[4];
    } else if (D == 1) {
      o[2] = vv * rho + c * dif;
      o[3] = ww * rho - x[3];
      o[1] = uu * rho - x[4];
      lm = -ww * x[3] - uu * x[4];
    } else {
      o[3] = ww * rho + c * dif;
      o[2] = vv * rho + x[3];
      o[1] = uu * rho - x[4];
      lm = vv * x[3] - uu * x[4];
    }
    o[0] = rho;
    o[4] = H * sum + qc * dif + hq2 * x[1] + lm;
  }
  // ---- LF flux minus F(Ug), from the reconstructed deviations
  const float rho_g = (float)g.rho, p_g = (float)g.p, Hg = (float)(g.E + g.p);
  float F[2][5], un[2], cs[2];
  const float* a[2] = {aL, aR};
#pragma unroll
  for (int s = 0; s < 2; s++) {
    const float rho = rho_g + a[s][0];
    const float ir = frcp(rho);
    const float u = a[s][1] * ir, v = a[s][2] * ir, w = a[s][3] * ir;
    const float dp = g1 * (a[s][4] - 0.5f * (a[s][1] * u + a[s][2] * v + a[s][3] * w));
    const float vn = D == 0 ? u : (D == 1 ? v : w);
    cs[s] = fsqrt(gf * (p_g + dp) * ir);
    un[s] = vn;
    F[s][0] = a[s][1 + D];
    F[s][1] = a[s][1] * vn + (D == 0 ? dp : 0.0f);
    F[s][2] = a[s][2] * vn + (D == 1 ? dp : 0.0f);
    F[s][3] = a[s][3] * vn + (D == 2 ? dp : 0.0f);
    F[s][4] = (Hg + a[s][4] + dp) * vn;
  }
  float ws = fabsf(un[1]) + cs[1];
  const float wsL = fabsf(un[0]) + cs[0];
  if (wsL > ws) ws = wsL;
#pragma unroll
  for (int e = 0; e < 5; e++) flux[e] = 0.5f * (F[0][e] + F[1][e] - ws * (aR[e] - aL[e]));
}

template <int D>
__device__ __forceinline__ void face_flux_f32(const double (&q)[6][5], const double (&pl)[5],
                                              const double (&pr)[5], double gamma, double gam1,
                                              const RefState& g, float (&flux)[5]) {
  float d[6][5], pf[5], qf[5];
#pragma unroll
  for (int cc = 0; cc < 6; cc++) dev_f32(q[cc], g, d[cc]);
#pragma unroll
  for (int e = 0; e < 5; e++) {
    pf[e] = (float)pl[e];
    qf[e] = (float)pr[e];
  }
  face_flux_f32_dev<D>(d, pf, qf, gamma, gam1, g, flux);
}

}  // namespace cbns
