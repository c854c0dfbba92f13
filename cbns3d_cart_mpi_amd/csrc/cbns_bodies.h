// Per-work-item bodies of the time-step path, shared by the single-solve
// kernels (one launch per block set, grid.y = block) and the batched
// persistent kernels (one launch per stage over a packed table of blocks of
// many solves).  A body computes one work item t of one block; both launch
// forms therefore produce identical values for a given block.
#pragma once
#include "cbns_face.h"
#include "cbns_kernels.h"

namespace cbns {

constexpr int TPB = 256;
constexpr int NXCD = 8;  // XCDs of an MI355X; workgroups are dispatched round-robin over them

__device__ __forceinline__ int gidx(const BlockDesc& B, int i, int j, int k) {
  return (i + NG) + B.g[0] * ((j + NG) + B.g[1] * (k + NG));
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Stencil kernels want neighbouring rows on the same L2, so the linear
// workgroup id is remapped so that XCD r runs one contiguous range of logical
// tiles (a bijection on [0, total) for any total).
struct Tile {
  int x, y;  // logical (blockIdx.x, blockIdx.y)
};
__device__ __forceinline__ Tile xcd_tile() {
  const int gx = gridDim.x;
  const int total = gx * gridDim.y;
  const int lin = blockIdx.y * gx + blockIdx.x;
  const int base = total / NXCD, rem = total % NXCD;
  const int r = lin % NXCD, q = lin / NXCD;
  const int logical = r * base + (r < rem ? r : rem) + q;
  return {logical % gx, logical / gx};
}

// Work index t -> (i, j, k) of an ni x nj x nk box in band order
// (j-band of jb rows, k, j within the band, i): a contiguous run of t covers
// a band of rows over successive planes, so a run of workgroups on one XCD
// keeps the y-neighbour rows (same band) and the z-neighbour rows (same rows,
// nearby planes) of its stencils in that XCD's L2.
__device__ __forceinline__ void band_decode(int t, int ni, int nj, int nk, int jb, int& i, int& j, int& k) {
  i = t % ni;
  const int R = t / ni;
  const int full = nj / jb;
  const int Rfull = full * nk * jb;
  if (R < Rfull) {
    const int b = R / (nk * jb);
    const int rem = R - b * nk * jb;
    k = rem / jb;
    j = b * jb + (rem - k * jb);
  } else {
    const int last = nj - full * jb;
    const int rem = R - Rfull;
    k = rem / last;
    j = full * jb + (rem - k * last);
  }
}

__device__ __forceinline__ int fdiv(int n, FastDiv f) {
  return (int)((__umulhi((unsigned)n, f.m) + (unsigned)n) >> f.l);
}

// band_decode of the face sweeps: the divisions by ni and nk through the
// block's FastDiv (fni, fnk), and with jb a power of two those by jb and
// nk jb as a shift and fnk; other bands take band_decode.
__device__ __forceinline__ void face_decode(int t, int ni, FastDiv fni, int nj, int nk, FastDiv fnk, int jb, int& i,
                                            int& j, int& k) {
  if ((jb & (jb - 1)) != 0) {  // uniform: a launch argument
    band_decode(t, ni, nj, nk, jb, i, j, k);
    return;
  }
  const int jl = __builtin_ctz(jb);
  const int R = fdiv(t, fni);
  i = t - R * ni;
  const int full = nj >> jl;
  if (R < full * nk * jb) {
    // R = (band nk + k) jb + row within the band
    const int Rb = R >> jl;
    const int b = fdiv(Rb, fnk);
    k = Rb - b * nk;
    j = (b << jl) + (R & (jb - 1));
  } else {
    const int last = nj - full * jb;
    const int rem = R - full * nk * jb;
    k = rem / last;
    j = full * jb + (rem - k * last);
  }
}

// x / h as divc, or with R (the tolerance-bound form's viscous terms and
// flux differences) x times RN(1/h): within about one ulp of the quotient,
// in one operation instead of three.
template <bool R>
__device__ __forceinline__ double divr(double x, double h, double rh, int fh) {
  if constexpr (R) return x * rh;
  else return divc(x, h, rh, fh);
}

// Stencil sum (1/12) a - (2/3) b + (2/3) c - (1/12) d of the 4th-order
// central difference of the values a, b, c, d at -2, -1, +1, +2: in the
// reference's order and rounding (src/calc_gradient.cpp:40-60; the
// divergence's c0 a + c1 b - c1 c - c0 d with c1 = -2/3,
// src/Block3d.cpp:894-908, rounds identically), or with R (the
// tolerance-bound and FP32-flux forms' viscous pass) as
// (1/12)(a - d) + (2/3)(c - b) in one multiply and one fused multiply-add.
template <bool R>
__device__ __forceinline__ double st4(double a, double b, double c, double d) {
  if constexpr (R) return __builtin_fma(1.0 / 12.0, a - d, (2.0 / 3.0) * (c - b));
  else return (1.0 / 12.0) * a - 2.0 / 3.0 * b + (2.0 / 3.0) * c - 1.0 / 12.0 * d;
}

// 4th-order central derivative (src/calc_gradient.cpp:29-64): the stencil
// sum, then the division by the spacing (divr).
template <bool R = false, class P>
__device__ __forceinline__ double grad4(P __restrict__ f, int c, int s, double h, double rh, int fh) {
  return divr<R>(st4<R>(f[c - 2 * s], f[c - s], f[c + s], f[c + 2 * s]), h, rh, fh);
}

// Divergence stencil of calc_viscous_flux_contribution (src/Block3d.cpp:894-908).
template <typename VT>
__device__ __forceinline__ double cdiff4(gptr<const VT> __restrict__ f, int c, int s) {
  return st4<false>(f[c - 2 * s], f[c - s], f[c + s], f[c + 2 * s]);
}

__device__ __forceinline__ int face_count(const BlockDesc& B, int D) {
  return (B.n[0] + (D == 0)) * (B.n[1] + (D == 1)) * (B.n[2] + (D == 2));
}

// Work index t -> (i, j, k) of single face t along D (band order).
template <int D>
__device__ __forceinline__ void single_face_decode(const BlockDesc& B, int t, int jb, int& i, int& j, int& k) {
  const int ni = B.n[0] + (D == 0), nj = B.n[1] + (D == 1), nk = B.n[2] + (D == 2);
  if constexpr (D == 2) band_decode(t, ni, nj, nk, jb, i, j, k);  // single z faces: the contracted form's sweep
  else face_decode(t, ni, B.dv_row[D == 0 ? 0 : 1], nj, nk, B.dv_plane[0], jb, i, j, k);
}

// ---- inviscid face flux: rec_riemann_x/y/z (src/Block3d.cpp:555-748) ----
// Face (i,j,k) along D sits between cells i-1 and i; t < face_count(B, D).
// PT: storage type of the primitives (double, or float in the FP32-flux form).
template <int D, typename PT = double>
__device__ __forceinline__ void face_load(const BlockDesc& B, int t, int which, int jb, double (&q)[6][5],
                                          double (&pl)[5], double (&pr)[5], int& fi) {
  const int ni = B.n[0] + (D == 0), nj = B.n[1] + (D == 1);
  int i, j, k;
  single_face_decode<D>(B, t, jb, i, j, k);
  fi = i + ni * (j + nj * k);
  const int c0 = gidx(B, i, j, k);
  const int sD = D == 0 ? 1 : (D == 1 ? B.g[0] : B.g[0] * B.g[1]);
  gptr<const double> __restrict__ Q = gmem<const double>(B.Q[which]);
  const long long sg = B.sg;
#pragma unroll
  for (int cc = 0; cc < 6; cc++)
#pragma unroll
    for (int e = 0; e < 5; e++) q[cc][e] = Q[e * sg + c0 + (cc - 3) * sD];
  gptr<const PT> __restrict__ P = gmem<const PT>((const PT*)B.prim);
#pragma unroll
  for (int f = 0; f < 5; f++) {
    pl[f] = (double)P[f * sg + c0 - sD];
    pr[f] = (double)P[f * sg + c0];
  }
}

// Plain path: the face flux of the face whose cells i-3 .. i+2 along D start
// at ghosted index c0 - 3 sD, recomputed from memory with the IEEE division
// and square root (face_flux<D, false, false>) for a face whose speculative
// form's operand checks failed.
template <int D>
__device__ __forceinline__ void face_flux_ieee(const BlockDesc& B, int which, int c0, int sD, const Phys& ph,
                                            double (&fl)[5]) {
  gptr<const double> __restrict__ Q = gmem<const double>(B.Q[which]);
  gptr<const double> __restrict__ P = gmem<const double>(B.prim);
  const long long sg = B.sg;
  double q[6][5], pl[5], pr[5];
#pragma unroll
  for (int cc = 0; cc < 6; cc++)
#pragma unroll
    for (int e = 0; e < 5; e++) q[cc][e] = Q[e * sg + c0 + (cc - 3) * sD];
#pragma unroll
  for (int f = 0; f < 5; f++) {
    pl[f] = P[f * sg + c0 - sD];
    pr[f] = P[f * sg + c0];
  }
  bool ok = true;
  face_flux<D, false, false>(q, pl, pr, ph.gamma, ph.gam1, ph.rgam1, ph.fgam1, fl, ok);
}

// The plain path runs the speculative form (face_flux SPEC) and recomputes
// a face whose checks fail with the IEEE forms.
template <int D, bool FMA = false>
__device__ __forceinline__ void face_body(const BlockDesc& B, int t, int which, int jb, const Phys& ph) {
  double q[6][5], pl[5], pr[5], fl[5];
  int fi;
  face_load<D>(B, t, which, jb, q, pl, pr, fi);
  bool ok = true;
  face_flux<D, FMA, !FMA>(q, pl, pr, ph.gamma, ph.gam1, ph.rgam1, ph.fgam1, fl, ok);
  if constexpr (!FMA) {
    if (!ok) {
      int i, j, k;
      single_face_decode<D>(B, t, jb, i, j, k);
      face_flux_ieee<D>(B, which, gidx(B, i, j, k), D == 0 ? 1 : (D == 1 ? B.g[0] : B.g[0] * B.g[1]), ph, fl);
    }
  }
  gptr<double> __restrict__ F = gmem(B.F[D]);
#pragma unroll
  for (int e = 0; e < 5; e++) F[e * B.sf[D] + fi] = fl[e];
}

template <int D>
__device__ __forceinline__ void face_body_f32(const BlockDesc& B, int t, int which, int jb, const Phys& ph,
                                              const RefState& g) {
  double q[6][5], pl[5], pr[5];
  float fl[5];
  int fi;
  face_load<D, float>(B, t, which, jb, q, pl, pr, fi);
  face_flux_f32<D>(q, pl, pr, ph.gamma, ph.gam1, g, fl);
  gptr<float> __restrict__ F = gmem((float*)B.F[D]);
#pragma unroll
  for (int e = 0; e < 5; e++) F[e * B.sf[D] + fi] = fl[e];
}

// y and z sweeps (every form): two faces adjacent along D per work item,
// faces m and m + 1 with m even (m = j for D = 1, k for D = 2), from the
// seven cells m-3 .. m+3 and the primitives of m-1 .. m+1: 35 state and 15
// primitive loads for two faces instead of 60 and 20, and one index decode.
// The last face of an odd count (n + 1 faces) is the first of a pair whose
// second face does not exist; its extra cell (m + 3 <= n + 3) is still
// inside the ghosted block.  Each face's arithmetic is face_body's.
__device__ __forceinline__ int face_pair_count(const BlockDesc& B, int D) {
  const int ni = B.n[0], nj = B.n[1] + (D == 1), nk = B.n[2] + (D == 2);
  return D == 1 ? ni * ((nj + 1) / 2) * nk : ni * nj * ((nk + 1) / 2);
}

// Work index t -> (i, j, k) of the first face of a pair along D.
template <int D>
__device__ __forceinline__ void pair_decode(const BlockDesc& B, int t, int jb, int& i, int& j, int& k) {
  const int ni = B.n[0], nj = B.n[1] + (D == 1), nk = B.n[2] + (D == 2);
  if (D == 1) {
    face_decode(t, ni, B.dv_row[1], (nj + 1) / 2, nk, B.dv_plane[0], jb, i, j, k);
    j *= 2;
  } else {
    face_decode(t, ni, B.dv_row[1], nj, (nk + 1) / 2, B.dv_plane[1], jb, i, j, k);
    k *= 2;
  }
}

template <int D, bool FMA = false>
__device__ __forceinline__ void face_pair_body(const BlockDesc& B, int t, int which, int jb, const Phys& ph) {
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
  gptr<const double> __restrict__ P = gmem<const double>(B.prim);
  double p[3][5];
#pragma unroll
  for (int f = 0; f < 5; f++)
#pragma unroll
    for (int s = 0; s < 3; s++) p[s][f] = P[f * sg + c0 + (s - 1) * sD];
  gptr<double> __restrict__ F = gmem(B.F[D]);
  const long long sf = B.sf[D];
  const int sF = D == 1 ? ni : ni * nj;  // face-array stride along D
  const int fi = i + ni * (j + nj * k);
#pragma unroll
  for (int h = 0; h < 2; h++) {
    if (h == 1 && (D == 1 ? j + 1 >= nj : k + 1 >= nk)) break;
    double qq[6][5];
#pragma unroll
    for (int cc = 0; cc < 6; cc++)
#pragma unroll
      for (int e = 0; e < 5; e++) qq[cc][e] = q7[cc + h][e];
    double fl[5];
    bool ok = true;
    face_flux<D, FMA, !FMA>(qq, p[h], p[h + 1], ph.gamma, ph.gam1, ph.rgam1, ph.fgam1, fl, ok);
    if constexpr (!FMA) {
      if (!ok) face_flux_ieee<D>(B, which, c0 + h * sD, sD, ph, fl);
    }
#pragma unroll
    for (int e = 0; e < 5; e++) F[e * sf + fi + h * sF] = fl[e];
  }
}

template <int D>
__device__ __forceinline__ void face_pair_body_f32(const BlockDesc& B, int t, int which, int jb, const Phys& ph,
                                                   const RefState& g) {
  static_assert(D == 1 || D == 2, "pairs along y or z");
  const int ni = B.n[0], nj = B.n[1] + (D == 1), nk = B.n[2] + (D == 2);
  int i, j, k;
  pair_decode<D>(B, t, jb, i, j, k);
  const int c0 = gidx(B, i, j, k);
  const int sD = D == 1 ? B.g[0] : B.g[0] * B.g[1];
  gptr<const double> __restrict__ Q = gmem<const double>(B.Q[which]);
  const long long sg = B.sg;
  float d[7][5];
#pragma unroll
  for (int cc = 0; cc < 7; cc++) {
    double q[5];
#pragma unroll
    for (int e = 0; e < 5; e++) q[e] = Q[e * sg + c0 + (cc - 3) * sD];
    dev_f32(q, g, d[cc]);
  }
  gptr<const float> __restrict__ P = gmem<const float>((const float*)B.prim);
  float p[3][5];
#pragma unroll
  for (int f = 0; f < 5; f++)
#pragma unroll
    for (int s = 0; s < 3; s++) p[s][f] = P[f * sg + c0 + (s - 1) * sD];
  gptr<float> __restrict__ F = gmem((float*)B.F[D]);
  const long long sf = B.sf[D];
  const int sF = D == 1 ? ni : ni * nj;  // face-array stride along D
  const int fi = i + ni * (j + nj * k);
#pragma unroll
  for (int h = 0; h < 2; h++) {
    if (h == 1 && (D == 1 ? j + 1 >= nj : k + 1 >= nk)) break;
    float dd[6][5];
#pragma unroll
    for (int cc = 0; cc < 6; cc++)
#pragma unroll
      for (int e = 0; e < 5; e++) dd[cc][e] = d[cc + h][e];
    float fl[5];
    face_flux_f32_dev<D>(dd, p[h], p[h + 1], ph.gamma, ph.gam1, g, fl);
#pragma unroll
    for (int e = 0; e < 5; e++) F[e * sf + fi + h * sF] = fl[e];
  }
}

// One face on explicit inputs: cells [nf][6][5] (cells i-3..i+2), prim
// [nf][2][5] (rho u v w p of cells i-1, i), flux [nf][5].  MODE 0: plain
// FP64; 1: contracted FP64 (instantiated only in the translation unit built
// with FMA contraction); 2: FP32-flux, with the face's own reference state
// (rho and E means of cells i-1 and i, p = (gamma - 1) E), the primitives
// rounded to float as they are stored, and F(Ug) (p_g in the normal
// momentum) added back in FP64 so that the result is the flux itself.
template <int D, int MODE>
__device__ __forceinline__ void face_eval_body(const Phys& ph, int t, const double* __restrict__ cells,
                                               const double* __restrict__ prim, double* __restrict__ flux) {
  double q[6][5], pl[5], pr[5], fl[5];
#pragma unroll
  for (int c = 0; c < 6; c++)
#pragma unroll
    for (int e = 0; e < 5; e++) q[c][e] = cells[(long long)t * 30 + 5 * c + e];
#pragma unroll
  for (int e = 0; e < 5; e++) {
    pl[e] = prim[(long long)t * 10 + e];
    pr[e] = prim[(long long)t * 10 + 5 + e];
  }
  if constexpr (MODE == 2) {
    RefState g;
    g.rho = 0.5 * (q[2][0] + q[3][0]);
    g.E = 0.5 * (q[2][4] + q[3][4]);
    g.p = ph.gam1 * g.E;
#pragma unroll
    for (int e = 0; e < 5; e++) {
      pl[e] = (double)(float)pl[e];
      pr[e] = (double)(float)pr[e];
    }
    float f32[5];
    face_flux_f32<D>(q, pl, pr, ph.gamma, ph.gam1, g, f32);
#pragma unroll
    for (int e = 0; e < 5; e++) fl[e] = (double)f32[e] + (e == 1 + D ? g.p : 0.0);
  } else {
    bool ok = true;
    face_flux<D, MODE == 1, MODE == 0>(q, pl, pr, ph.gamma, ph.gam1, ph.rgam1, ph.fgam1, fl, ok);
    if (MODE == 0 && !ok) face_flux<D, false, false>(q, pl, pr, ph.gamma, ph.gam1, ph.rgam1, ph.fgam1, fl, ok);
  }
#pragma unroll
  for (int e = 0; e < 5; e++) flux[(long long)t * 5 + e] = fl[e];
}

// ---- viscous fluxes and their divergence ------------------------------------
// calc_gradient (src/calc_gradient.cpp:17-66), calc_viscous_terms
// (src/Block3d.cpp:750-810), calc_viscous_flux (src/Block3d.cpp:812-868) and
// the divergence of calc_viscous_flux_contribution (src/Block3d.cpp:879-916)
// in one tile march.  Only the divergence (4 fields on the interior, B.dfv)
// and mu (on the interior, for calc_dt) are written to HBM: the nine flux
// components txx tyy tzz txy txz tyz ex ey ez live in LDS and registers.
// The fluxes and the divergence are rounded to VT as stored fields would be
// (double for the FP64 forms, float for FP32-flux).
// Sutherland viscosity, stress tensor, heat flux and energy fluxes of one
// point from its gradients gr[u, v, w, T][x, y, z], temperature and velocity
// (src/Block3d.cpp:777-804, 843-862).  out = txx tyy tzz txy txz tyz ex ey ez;
// mu_out = Sutherland mu (before the 1/Re factor).
__device__ __forceinline__ void visc_terms(const Phys& ph, const double (&gr)[4][3], double T_l, double uu,
                                           double vv, double ww, double (&out)[9], double& mu_out) {
  const double u_x = gr[0][0], u_y = gr[0][1], u_z = gr[0][2];
  const double v_x = gr[1][0], v_y = gr[1][1], v_z = gr[1][2];
  const double w_x = gr[2][0], w_y = gr[2][1], w_z = gr[2][2];
  const double T_x = gr[3][0], T_y = gr[3][1], T_z = gr[3][2];
  double mu_l = (1.0 + ph.C_T_inf) / (T_l + ph.C_T_inf) * sqrt(T_l * T_l * T_l);
  mu_out = mu_l;
  mu_l *= ph.Re_inv;
  const double txx = (2.0 / 3.0) * mu_l * (2.0 * u_x - v_y - w_z);
  const double tyy = (2.0 / 3.0) * mu_l * (2.0 * v_y - u_x - w_z);
  const double tzz = (2.0 / 3.0) * mu_l * (2.0 * w_z - u_x - v_y);
  const double txy = mu_l * (u_y + v_x);
  const double txz = mu_l * (u_z + w_x);
  const double tyz = mu_l * (v_z + w_y);
  mu_l *= ph.kappa;
  const double qx = mu_l * T_x, qy = mu_l * T_y, qz = mu_l * T_z;
  out[0] = txx;
  out[1] = tyy;
  out[2] = tzz;
  out[3] = txy;
  out[4] = txz;
  out[5] = tyz;
  out[6] = uu * txx + vv * txy + ww * txz - qx;
  out[7] = uu * txy + vv * tyy + ww * tyz - qy;
  out[8] = uu * txz + vv * tyz + ww * tzz - qz;
}

// FROMQ (tolerance-bound FP64 form, whose stage tail stores no primitives):
// u, v, w, T are formed from the conserved values of Q[which] as each
// point is staged (calc_primitive's expressions, src/Block3d.cpp:369-381,
// with one refined reciprocal of rho).
__device__ __forceinline__ void prim_uvwT(const double (&q)[5], const Phys& ph, double (&o)[4]) {
#pragma clang fp contract(off)
  const double ir = rcp_nr(q[0]);
  o[0] = q[1] * ir;
  o[1] = q[2] * ir;
  o[2] = q[3] * ir;
  const double k = __builtin_fma(o[2], o[2], __builtin_fma(o[1], o[1], o[0] * o[0]));
  const double p = ph.gam1 * __builtin_fma(-0.5 * q[0], k, q[4]);
  o[3] = (ph.gM2 * p) * ir;
}

// Loads and stores at 32-bit byte offsets from a per-field base: the base
// is uniform (scalar registers) and the offset one VGPR, so an access needs
// no 64-bit address arithmetic.  Every field of a block is below 4 GB
// (MAX_BLOCK_CELLS).
template <typename T>
__device__ __forceinline__ T gld(gptr<const T> base, unsigned idx) {
  return *(gptr<const T>)((gptr<const char>)base + idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void gst(gptr<T> base, unsigned idx, T v) {
  *(gptr<T>)((gptr<char>)base + idx * (unsigned)sizeof(T)) = v;
}

// Tile march.  A workgroup owns a VX x VY (16 x 16) tile of interior columns
// and marches the planes p of one z-chunk [Z0, Z1), from Z0 - 2 to Z1 + 1.
// Each thread keeps a five-plane register queue (p - 2 .. p + 2) of u, v, w,
// T of its own column.  Per plane the workgroup
//  - stages u, v, w, T of the (VX + 8) x (VY + 8) region around the tile in
//    LDS (double-buffered) for the x- and y-gradients: each thread writes
//    its own column's plane-p value from its queue, and only the 320-point
//    ring around the tile is loaded (one ring point per thread, a second
//    for half of waves 2 and 3, the waves without side points);
//  - evaluates the fluxes at each thread's own column point, with the
//    z-gradient from its queue; the z-fluxes Gv = (txz, tyz, tzz, ez) enter a
//    five-plane queue of their own;
//  - on the chunk's planes, wave 0 also evaluates Ev = (txx, txy, txz, ex) at
//    the two columns beside each x-side of the tile (16 rows x 4 columns)
//    and wave 1 Fv = (txy, tyy, tyz, ey) at the two rows beside each y-side
//    (4 rows x 16 columns): only the six x/y-gradients and the flux
//    components that side's divergence takes, with a queue of the two
//    velocities whose z-gradients it needs (loaded by these two waves alone);
//    Ev and Fv of the tile and its sides go to LDS;
//  - forms the x- and y-parts of the divergence of plane p - 1 from the Ev
//    and Fv that the previous plane left in LDS (Ev and Fv are
//    double-buffered, so that one barrier per plane suffices), and completes
//    the divergence of plane p - 2 with the z-part from the Gv queue, summed
//    in the reference's order (x-part, + y-part, + z-part).
// The next plane's loads are issued before the current plane's arithmetic.
// Every value is formed by the same expressions as the reference's separate
// passes (grad4, visc_terms, the divergence stencil); in the tolerance-bound
// (FROMQ) and FP32-flux (float VT) forms the divisions by the spacings are
// multiplications by their reciprocals (divr).  One call = one (column tile,
// z-chunk) unit of vz planes, made by all VX * VY threads of a workgroup with
// the same tile and chunk (it synchronises the workgroup).
template <typename VT, bool FROMQ = false>
__device__ __forceinline__ void visc_tile_body(const BlockDesc& B, const Phys& ph, int tile, int chunk, int which,
                                               int vz) {
  static_assert(VX == 16 && VY == 16, "one wave of side points per side kind");
  static_assert(!FROMQ || sizeof(VT) == 8, "the state is FP64");
  // divisions by the spacings as multiplications by their reciprocals: the
  // tolerance-bound form (FROMQ) and the FP32-flux form (float storage)
  constexpr bool RX = FROMQ || sizeof(VT) == 4;
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int ntx = (n0 + VX - 1) / VX, nty = (n1 + VY - 1) / VY;
  const int Z0 = chunk * vz;
  if (tile >= ntx * nty || Z0 >= n2) return;  // uniform per workgroup
  const int Z1 = Z0 + vz < n2 ? Z0 + vz : n2;
  const int X0 = (tile % ntx) * VX, Y0 = (tile / ntx) * VY;
  const int wx = n0 - X0 < VX ? n0 - X0 : VX, wy = n1 - Y0 < VY ? n1 - Y0 : VY;  // the tile's extent
  const int tid = threadIdx.x, tx = tid % VX, ty = tid / VX;
  const bool act = tx < wx && ty < wy;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  // side point of this thread (offsets from X0, Y0): wave 0 Ev at an x-side
  // column, wave 1 Fv at a y-side row; sk: the point exists
  int sx = 0, sy = 0;
  bool sk = false;
  if (wave == 0) {
    const int r = lane >> 2, s = lane & 3;
    sx = s < 2 ? s - 2 : wx + s - 2;
    sy = r;
    sk = r < wy;
  } else if (wave == 1) {
    const int c = lane & 15, s = lane >> 4;
    sx = c;
    sy = s < 2 ? s - 2 : wy + s - 2;
    sk = c < wx;
  }
  const bool sidew = wave < 2;  // uniform
  constexpr int SX = VX + 8, SY = VY + 8, NRING = SX * SY - VX * VY;
  __shared__ VT L[2][4][SY][SX];       // u, v, w, T of the staged region, in the storage type
  __shared__ VT LE[2][4][VY][VX + 4];  // Ev of the tile and its x-sides
  __shared__ VT LF[2][4][VY + 4][VX];  // Fv of the tile and its y-sides
  const int sz = B.g[0] * B.g[1];
  const long long sg = B.sg;
  // source fields: u, v, w, T of the primitives (sharing VT), or with FROMQ
  // rho, rho u, rho v, rho w, E of the state
  constexpr int NF = FROMQ ? 5 : 4;
  const VT* src = FROMQ ? (const VT*)B.Q[which] : (const VT*)B.prim;
  gptr<const VT> Pf[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) Pf[f] = gmem<const VT>(src + (FROMQ ? f * sg : (f < 3 ? f + 1 : 5) * sg));
  // the values of one point as u, v, w, T
  const auto uvwT = [&](const VT (&x)[NF], double (&o)[4]) {
    if constexpr (FROMQ) {
      const double q[5] = {(double)x[0], (double)x[1], (double)x[2], (double)x[3], (double)x[NF - 1]};
      prim_uvwT(q, ph, o);
    } else {
#pragma unroll
      for (int f = 0; f < 4; f++) o[f] = (double)x[f];
    }
  };
  // Positions past the block's last ghost cell are clamped to it: no used
  // value comes from them (the farthest point a side point's stencil reaches
  // is (X0 + wx + 3, Y0 + wy + 3)), and every load stays inside the ghosted
  // block.
  const auto clampx = [&](int c) { return c < n0 + NG - 1 ? c : n0 + NG - 1; };
  const auto clampy = [&](int c) { return c < n1 + NG - 1 ? c : n1 + NG - 1; };
  // ring entries of this thread: LDS offset in a field plane (-1: none) and
  // column index in the ghosted block.  Entry e < 320: rows 0-3, rows
  // VY+4 .. VY+7, then columns 0-3 and VX+4 .. VX+7 of rows 4 .. VY+3.
  int rl[2];
  unsigned rg[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    int e = q == 0 ? tid : (wave >= 2 && lane < 32 ? VX * VY + (wave - 2) * 32 + lane : -1);
    if (e >= NRING) e = -1;
    int lx = 0, ly = 0;
    if (e >= 0) {
      if (e < 4 * SX) {
        ly = e / SX;
        lx = e % SX;
      } else if (e < 8 * SX) {
        const int f = e - 4 * SX;
        ly = VY + 4 + f / SX;
        lx = f % SX;
      } else {
        const int f = e - 8 * SX, side = f / (4 * VY), g = f - side * 4 * VY;
        ly = 4 + g / 4;
        lx = (g & 3) + side * (VX + 4);
      }
    }
    rl[q] = e >= 0 ? ly * SX + lx : -1;
    rg[q] = (unsigned)gidx(B, clampx(X0 - 4 + lx), clampy(Y0 - 4 + ly), 0);
  }
  // own column (clamped like the staging; the tile's own column for act threads)
  const unsigned ci = (unsigned)gidx(B, clampx(X0 + tx), clampy(Y0 + ty), 0);
  const int own_l = (ty + 4) * SX + tx + 4;
  // side column and the fields its z-gradients take: with FROMQ rho, rho a,
  // rho w, else a, w (a = u beside an x-side, v beside a y-side)
  const unsigned cs = sk ? (unsigned)gidx(B, X0 + sx, Y0 + sy, 0) : ci;
  constexpr int NH = FROMQ ? 3 : 2;
  gptr<const VT> Ph[NH];
  if constexpr (FROMQ) {
    Ph[0] = Pf[0];
    Ph[1] = Pf[wave == 1 ? 2 : 1];
    Ph[2] = Pf[3];
  } else {
    Ph[0] = Pf[wave == 1 ? 1 : 0];
    Ph[1] = Pf[2];
  }
  // the side point's two velocities (prim_uvwT's values with FROMQ)
  const auto side_vel = [&](const VT (&h)[NH], VT& va, VT& vw) {
    if constexpr (FROMQ) {
      const double ir = rcp_nr((double)h[0]);
      va = (VT)((double)h[1] * ir);
      vw = (VT)((double)h[2] * ir);
    } else {
      va = h[0];
      vw = h[1];
    }
  };
  // queues in the staging type (every value is exact in it: VT is the
  // primitives' storage type, and FROMQ forms double values)
  VT r[4][5];   // u, v, w, T of the own column, planes p - 2 .. p + 2
  VT sq[2][5];  // the side point's two velocities, planes p - 2 .. p + 2
  VT gq[4][5];  // Gv of the own column, planes p - 4 .. p
  double dq[4]; // x- plus y-part of the divergence of plane p - 2
  VT pf[2][NF], rn[NF], hn[NH];  // next-plane prefetch (raw source values)
#pragma unroll
  for (int d = 0; d < 4; d++) {
    VT x[NF];
    double o[4];
#pragma unroll
    for (int f = 0; f < NF; f++) x[f] = gld(Pf[f], ci + (unsigned)((Z0 - 4 + d) * sz));
    uvwT(x, o);
#pragma unroll
    for (int f = 0; f < 4; f++) r[f][d] = (VT)o[f];
    if (sidew) {
      VT h[NH];
#pragma unroll
      for (int f = 0; f < NH; f++) h[f] = gld(Ph[f], cs + (unsigned)((Z0 - 4 + d) * sz));
      side_vel(h, sq[0][d], sq[1][d]);
    }
  }
#pragma unroll
  for (int e = 0; e < 4; e++) {
#pragma unroll
    for (int d = 0; d < 5; d++) gq[e][d] = (VT)0;
    dq[e] = 0.0;
  }
#pragma unroll
  for (int f = 0; f < NF; f++) {
    rn[f] = gld(Pf[f], ci + (unsigned)(Z0 * sz));
#pragma unroll
    for (int q = 0; q < 2; q++)
      if (rl[q] >= 0) pf[q][f] = gld(Pf[f], rg[q] + (unsigned)((Z0 - 2) * sz));
  }
  if (sidew) {
#pragma unroll
    for (int f = 0; f < NH; f++) hn[f] = gld(Ph[f], cs + (unsigned)(Z0 * sz));
  }
  gptr<VT> DV[4];
#pragma unroll
  for (int e = 0; e < 4; e++) DV[e] = gmem((VT*)B.dfv + e * B.sn);
  const gptr<double> MU = gmem(B.mu);
  const unsigned oi = (unsigned)(X0 + tx + n0 * (Y0 + ty));  // interior offset of the own column (act)
  const unsigned pn = (unsigned)(n0 * n1);
  // Ev = (txx, txy, txz, ex), Fv = (txy, tyy, tyz, ey), Gv = (txz, tyz, tzz, ez)
  const int Ef[4] = {0, 3, 4, 6}, Ff[4] = {3, 1, 5, 7}, Gf[4] = {4, 5, 2, 8};
  for (int p = Z0 - 2; p < Z1 + 2; p++) {
    const int buf = (p - Z0) & 1;
    // stage plane p: ring entries from the prefetch, the tile from the own queue;
    // the queues take plane p + 2
#pragma unroll
    for (int q = 0; q < 2; q++) {
      if (rl[q] >= 0) {
        double o[4];
        uvwT(pf[q], o);
#pragma unroll
        for (int f = 0; f < 4; f++) (&L[buf][f][0][0])[rl[q]] = (VT)o[f];
      }
    }
    {
      double o[4];
      uvwT(rn, o);
#pragma unroll
      for (int f = 0; f < 4; f++) r[f][4] = (VT)o[f];
    }
#pragma unroll
    for (int f = 0; f < 4; f++) (&L[buf][f][0][0])[own_l] = r[f][2];
    if (sidew) side_vel(hn, sq[0][4], sq[1][4]);
    __syncthreads();
    if (p + 1 < Z1 + 2) {  // planes p + 1 (staging) and p + 3 (queues): at most n2 + 3, the last ghost plane
#pragma unroll
      for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int q = 0; q < 2; q++)
          if (rl[q] >= 0) pf[q][f] = gld(Pf[f], rg[q] + (unsigned)((p + 1) * sz));
        rn[f] = gld(Pf[f], ci + (unsigned)((p + 3) * sz));
      }
      if (sidew) {
#pragma unroll
        for (int f = 0; f < NH; f++) hn[f] = gld(Ph[f], cs + (unsigned)((p + 3) * sz));
      }
    }
    // x- and y-parts of the divergence of plane p - 1 (its Ev, Fv are in LDS)
    double dxy[4] = {0.0, 0.0, 0.0, 0.0};
    if (act && p - 1 >= Z0 && p - 1 < Z1) {
      const int b = buf ^ 1;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        double d = divr<RX>(st4<RX>(LE[b][e][ty][tx], LE[b][e][ty][tx + 1], LE[b][e][ty][tx + 3], LE[b][e][ty][tx + 4]),
                            ph.dx, ph.rdx, ph.fdx);
        d += divr<RX>(st4<RX>(LF[b][e][ty][tx], LF[b][e][ty + 1][tx], LF[b][e][ty + 3][tx], LF[b][e][ty + 4][tx]),
                      ph.dy, ph.rdy, ph.fdy);
        dxy[e] = d;
      }
    }
    const bool out = p >= Z0 && p < Z1;  // uniform: a plane of the chunk
    if (act) {
      double gr[4][3];
#pragma unroll
      for (int f = 0; f < 4; f++) {
        gr[f][0] = grad4<RX>(&L[buf][f][ty + 4][0], tx + 4, 1, ph.dx, ph.rdx, ph.fdx);
        gr[f][1] = grad4<RX>(&L[buf][f][0][tx + 4], (ty + 4) * SX, SX, ph.dy, ph.rdy, ph.fdy);
        gr[f][2] = grad4<RX>(&r[f][0], 2, 1, ph.dz, ph.rdz, ph.fdz);
      }
      double o[9], mu_l;
      visc_terms(ph, gr, r[3][2], r[0][2], r[1][2], r[2][2], o, mu_l);
#pragma unroll
      for (int e = 0; e < 4; e++) gq[e][4] = (VT)o[Gf[e]];
      if (out) {
        gst(MU, oi + pn * (unsigned)p, mu_l);
#pragma unroll
        for (int e = 0; e < 4; e++) {
          LE[buf][e][ty][tx + 2] = (VT)o[Ef[e]];
          LF[buf][e][ty + 2][tx] = (VT)o[Ff[e]];
        }
      }
    }
    if (out && sidew && sk) {
      // visc_terms' expressions for the components this side keeps; the
      // z-gradients are a_z (u_z beside an x-side, v_z beside a y-side) and w_z
      const int lx = sx + 4, ly = sy + 4;
      const double T_l = (double)L[buf][3][ly][lx], uu = (double)L[buf][0][ly][lx],
                   vv = (double)L[buf][1][ly][lx], ww = (double)L[buf][2][ly][lx];
      const double u_x = grad4<RX>(&L[buf][0][ly][0], lx, 1, ph.dx, ph.rdx, ph.fdx);
      const double u_y = grad4<RX>(&L[buf][0][0][lx], ly * SX, SX, ph.dy, ph.rdy, ph.fdy);
      const double v_x = grad4<RX>(&L[buf][1][ly][0], lx, 1, ph.dx, ph.rdx, ph.fdx);
      const double v_y = grad4<RX>(&L[buf][1][0][lx], ly * SX, SX, ph.dy, ph.rdy, ph.fdy);
      const double a_z = grad4<RX>(&sq[0][0], 2, 1, ph.dz, ph.rdz, ph.fdz);
      const double w_z = grad4<RX>(&sq[1][0], 2, 1, ph.dz, ph.rdz, ph.fdz);
      double mu_l = (1.0 + ph.C_T_inf) / (T_l + ph.C_T_inf) * sqrt(T_l * T_l * T_l);
      mu_l *= ph.Re_inv;
      const double txy = mu_l * (u_y + v_x);
      if (wave == 0) {  // Ev = (txx, txy, txz, ex)
        const double w_x = grad4<RX>(&L[buf][2][ly][0], lx, 1, ph.dx, ph.rdx, ph.fdx);
        const double T_x = grad4<RX>(&L[buf][3][ly][0], lx, 1, ph.dx, ph.rdx, ph.fdx);
        const double txx = (2.0 / 3.0) * mu_l * (2.0 * u_x - v_y - w_z);
        const double txz = mu_l * (a_z + w_x);
        mu_l *= ph.kappa;
        const double qx = mu_l * T_x;
        LE[buf][0][sy][sx + 2] = (VT)txx;
        LE[buf][1][sy][sx + 2] = (VT)txy;
        LE[buf][2][sy][sx + 2] = (VT)txz;
        LE[buf][3][sy][sx + 2] = (VT)(uu * txx + vv * txy + ww * txz - qx);
      } else {  // Fv = (txy, tyy, tyz, ey)
        const double w_y = grad4<RX>(&L[buf][2][0][lx], ly * SX, SX, ph.dy, ph.rdy, ph.fdy);
        const double T_y = grad4<RX>(&L[buf][3][0][lx], ly * SX, SX, ph.dy, ph.rdy, ph.fdy);
        const double tyy = (2.0 / 3.0) * mu_l * (2.0 * v_y - u_x - w_z);
        const double tyz = mu_l * (a_z + w_y);
        mu_l *= ph.kappa;
        const double qy = mu_l * T_y;
        LF[buf][0][sy + 2][sx] = (VT)txy;
        LF[buf][1][sy + 2][sx] = (VT)tyy;
        LF[buf][2][sy + 2][sx] = (VT)tyz;
        LF[buf][3][sy + 2][sx] = (VT)(uu * txy + vv * tyy + ww * tyz - qy);
      }
    }
    // the divergence of plane p - 2 (< Z1) is complete
    if (act && p - 2 >= Z0) {
      const unsigned o = oi + pn * (unsigned)(p - 2);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const double dz = st4<RX>(gq[e][0], gq[e][1], gq[e][3], gq[e][4]);
        gst(DV[e], o, (VT)(dq[e] + divr<RX>(dz, ph.dz, ph.rdz, ph.fdz)));
      }
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
      dq[e] = dxy[e];
#pragma unroll
      for (int d = 0; d < 4; d++) gq[e][d] = gq[e][d + 1];
    }
#pragma unroll
    for (int f = 0; f < 4; f++)
#pragma unroll
      for (int d = 0; d < 4; d++) r[f][d] = r[f][d + 1];
#pragma unroll
    for (int d = 0; d < 4; d++) {
      sq[0][d] = sq[0][d + 1];
      sq[1][d] = sq[1][d + 1];
    }
  }
}

// ---- stage tail: RK3 update + primitives ---------------------------------

// RK3 stage combination (src/Block3d.cpp:949-955) of cell c from the
// right-hand side df, Q^n = Qn, the stage input Qi, into Qo; then the cell's
// primitives into P (src/Block3d.cpp:369-381).  Qo may be Qn (stage 3).
template <typename PT>
__device__ __forceinline__ void rk3_prim(const double (&df)[NEQ], gptr<const double> Qn, gptr<const double> Qi,
                                         gptr<double> Qo, gptr<PT> __restrict__ P, long long sg, int c,
                                         double dt, int stage,
                                         const Phys& ph) {
  double qn[NEQ];
#pragma unroll
  for (int e = 0; e < NEQ; e++) {
    const long long q = e * sg + c;
    if (1 == stage) qn[e] = Qn[q] - dt * df[e];
    else if (2 == stage) qn[e] = 0.75 * Qn[q] + 0.25 * (Qi[q] - dt * df[e]);
    else qn[e] = (Qn[q] + 2.0 * (Qi[q] - dt * df[e])) / 3.0;
    Qo[q] = qn[e];
  }
  const double rr = qn[0];
  const double uu = qn[1] / rr;
  const double vv = qn[2] / rr;
  const double ww = qn[3] / rr;
  const double pp = ph.gam1 * (qn[4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
  P[c] = (PT)rr;
  P[sg + c] = (PT)uu;
  P[2 * sg + c] = (PT)vv;
  P[3 * sg + c] = (PT)ww;
  P[4 * sg + c] = (PT)pp;
  P[5 * sg + c] = (PT)(ph.gM2 * pp / rr);
}

// src/Block3d.cpp:920-962 and 359-385 for one interior cell, with the
// viscous divergence of the viscous pass (B.dfv); stage 1..3 only, in place
// on Q / Q_p as the reference.  FT is the face-array type (double, or float
// for FP32-flux), which is also the type of the primitives and of the
// divergence.
template <typename FT>
__device__ __forceinline__ void update_body(const BlockDesc& B, int t, double dt, int stage, const Phys& ph,
                                            int jb) {
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  int i, j, k;
  band_decode(t, n0, n1, n2, jb, i, j, k);
  const int ct = i + n0 * (j + n1 * k);
  const int c = gidx(B, i, j, k);
  double dfv[4];
#pragma unroll
  for (int e = 0; e < 4; e++) dfv[e] = (double)gmem<const FT>((const FT*)B.dfv)[e * B.sn + ct];
  const int iE = i + (n0 + 1) * (j + n1 * k);
  const int iF = i + n0 * (j + (n1 + 1) * k);
  const int iG = ct;
  double df[NEQ];
#pragma unroll
  for (int e = 0; e < NEQ; e++) {
    gptr<const FT> __restrict__ Ep = gmem<const FT>((const FT*)B.F[0] + e * B.sf[0]);
    gptr<const FT> __restrict__ Fp = gmem<const FT>((const FT*)B.F[1] + e * B.sf[1]);
    gptr<const FT> __restrict__ Gp = gmem<const FT>((const FT*)B.F[2] + e * B.sf[2]);
    double d = divc((double)Ep[iE + 1] - (double)Ep[iE], ph.dx, ph.rdx, ph.fdx);
    d += divc((double)Fp[iF + n0] - (double)Fp[iF], ph.dy, ph.rdy, ph.fdy);
    d += divc((double)Gp[iG + n0 * n1] - (double)Gp[iG], ph.dz, ph.rdz, ph.fdz);
    if (e > 0) d -= dfv[e - 1];
    df[e] = d;
  }
  rk3_prim<FT>(df, gmem<const double>(B.Q[0]), gmem<const double>(B.Q[1]), gmem(stage == 3 ? B.Q[0] : B.Q[1]),
               gmem((FT*)B.prim), B.sg, c, dt, stage, ph);
}

// ---- calc_dt spectral radius of one cell from its primitives and mu
// (src/Block3d.cpp:527-546).  Returns t1 + t2 + t3; *bad is set for a NaN
// or infinite (>= 1e308) result.
__device__ __forceinline__ double dt_point(double rr, double uu, double vv, double ww, double pp, double mu,
                                           const Phys& ph, int* bad) {
  const double cs = sqrt(ph.gamma * pp / rr);
  const double factor_v = 2.0 * mu * ph.C_dt_v * sqrt(uu * uu + vv * vv + ww * ww) / cs / rr;
  double t1 = fabs(uu) + divc(cs, ph.dx, ph.rdx, ph.fdx);
  t1 += divc(divc(factor_v, ph.dx, ph.rdx, ph.fdx), ph.dx, ph.rdx, ph.fdx);
  double t2 = fabs(vv) + divc(cs, ph.dy, ph.rdy, ph.fdy);
  t2 += divc(divc(factor_v, ph.dy, ph.rdy, ph.fdy), ph.dy, ph.rdy, ph.fdy);
  double t3 = fabs(ww) + divc(cs, ph.dz, ph.rdz, ph.fdz);
  t3 += divc(divc(factor_v, ph.dz, ph.rdz, ph.fdz), ph.dz, ph.rdz, ph.fdz);
  const double s = t1 + t2 + t3;
  if (!(s < 1.0e308)) *bad = 1;
  return s;
}

// ---- stage tail over x-pairs of cells ---------------------------------------
// The same arithmetic as update_body for the two cells (i, j, k), (i+1, j, k)
// with i even, so that every field of the pair is one 2-wide load or store
// (for FP32-flux, the float face arrays move 8 bytes per lane instead of 4).  Requires an even n0: every field stride and row offset is
// then even, so the pair loads are naturally aligned.
template <typename T>
using vec2 = T __attribute__((ext_vector_type(2)));

template <typename T>
__device__ __forceinline__ vec2<T> ld2(gptr<const T> p, long long o) {
  return *(gptr<const vec2<T>>)(p + o);
}

// With dts != nullptr (stage 3) it also returns in *s_out the larger calc_dt
// spectral radius of the two updated cells (a NaN never wins), with mu from
// this stage's viscous pass, which is what the next step's calc_dt reads.
template <typename FT>
__device__ __forceinline__ void update_pair_body(const BlockDesc& B, int t, double dt, int stage, const Phys& ph,
                                                 int jb, bool want_s = false, double* s_out = nullptr,
                                                 int* bad = nullptr) {
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  int ih, j, k;
  band_decode(t, n0 / 2, n1, n2, jb, ih, j, k);
  const int i = 2 * ih;
  // viscous divergence of both cells (the viscous pass's B.dfv)
  const int iG = i + n0 * (j + n1 * k);
  double dfv[2][4];
#pragma unroll
  for (int e = 0; e < 4; e++) {
    const vec2<FT> v = ld2(gmem<const FT>((const FT*)B.dfv), e * B.sn + iG);
    dfv[0][e] = (double)v.x;
    dfv[1][e] = (double)v.y;
  }
  // inviscid flux differences, RK3 and primitives
  const int iE = i + (n0 + 1) * (j + n1 * k);
  const int iF = i + n0 * (j + (n1 + 1) * k);
  const long long sg = B.sg;
  const int c = gidx(B, i, j, k);
  gptr<const double> Qn = gmem<const double>(B.Q[0]);
  gptr<const double> Qi = gmem<const double>(B.Q[1]);
  gptr<double> Qo = gmem(stage == 3 ? B.Q[0] : B.Q[1]);
  double qn[2][NEQ];
#pragma unroll
  for (int e = 0; e < NEQ; e++) {
    gptr<const FT> Ep = gmem<const FT>((const FT*)B.F[0] + e * B.sf[0]);
    gptr<const FT> Fp = gmem<const FT>((const FT*)B.F[1] + e * B.sf[1]);
    gptr<const FT> Gp = gmem<const FT>((const FT*)B.F[2] + e * B.sf[2]);
    const double e0 = (double)Ep[iE], e1 = (double)Ep[iE + 1], e2 = (double)Ep[iE + 2];
    const vec2<FT> fa = ld2(Fp, iF), fb = ld2(Fp, iF + n0);
    const vec2<FT> ga = ld2(Gp, iG), gb = ld2(Gp, iG + (long long)n0 * n1);
    double df0 = divc(e1 - e0, ph.dx, ph.rdx, ph.fdx);
    df0 += divc((double)fb.x - (double)fa.x, ph.dy, ph.rdy, ph.fdy);
    df0 += divc((double)gb.x - (double)ga.x, ph.dz, ph.rdz, ph.fdz);
    double df1 = divc(e2 - e1, ph.dx, ph.rdx, ph.fdx);
    df1 += divc((double)fb.y - (double)fa.y, ph.dy, ph.rdy, ph.fdy);
    df1 += divc((double)gb.y - (double)ga.y, ph.dz, ph.rdz, ph.fdz);
    if (e > 0) {
      df0 -= dfv[0][e - 1];
      df1 -= dfv[1][e - 1];
    }
    const long long q = e * sg + c;
    const vec2<double> a = ld2(Qn, q);
    vec2<double> r;
    if (1 == stage) {
      r.x = a.x - dt * df0;
      r.y = a.y - dt * df1;
    } else {
      const vec2<double> b = ld2(Qi, q);
      if (2 == stage) {
        r.x = 0.75 * a.x + 0.25 * (b.x - dt * df0);
        r.y = 0.75 * a.y + 0.25 * (b.y - dt * df1);
      } else {
        r.x = (a.x + 2.0 * (b.x - dt * df0)) / 3.0;
        r.y = (a.y + 2.0 * (b.y - dt * df1)) / 3.0;
      }
    }
    *(gptr<vec2<double>>)(Qo + q) = r;
    qn[0][e] = r.x;
    qn[1][e] = r.y;
  }
  gptr<FT> P = gmem((FT*)B.prim);
  vec2<FT> pr[6];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const double rr = qn[h][0];
    const double uu = qn[h][1] / rr;
    const double vv = qn[h][2] / rr;
    const double ww = qn[h][3] / rr;
    const double pp = ph.gam1 * (qn[h][4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
    pr[0][h] = (FT)rr;
    pr[1][h] = (FT)uu;
    pr[2][h] = (FT)vv;
    pr[3][h] = (FT)ww;
    pr[4][h] = (FT)pp;
    pr[5][h] = (FT)(ph.gM2 * pp / rr);
    if (want_s) {
      const double sc = dt_point(rr, uu, vv, ww, pp, gmem<const double>(B.mu)[iG + h], ph, bad);
      *s_out = (*s_out < sc) ? sc : *s_out;  // src/Block3d.cpp:546: a NaN never wins
    }
  }
#pragma unroll
  for (int f = 0; f < 6; f++) *(gptr<vec2<FT>>)(P + f * sg + c) = pr[f];
}

// ---- ghost gather of the state and its primitives ------------------------
// The g-th ghost cell (0 <= g < sg - sn) of a block as a ghosted linear index:
// first the NG z-planes below and above the interior (whole planes), then
// per interior plane the NG y-rows on each side (whole rows), then per
// interior row the NG x-cells on each side.
__device__ __forceinline__ int ghost_cell(const BlockDesc& B, int g) {
  const int g0 = B.g[0], g1 = B.g[1], n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int plane = g0 * g1;
  const int zs = 2 * NG * plane;
  if (g < zs) {
    const int p = g / plane;
    return (p < NG ? p : n2 + p) * plane + (g - p * plane);
  }
  g -= zs;
  const int yrow = 2 * NG * g0;
  if (g < n2 * yrow) {
    const int kk = g / yrow, r = g - kk * yrow;
    const int row = r / g0;
    return (kk + NG) * plane + (row < NG ? row : n1 + row) * g0 + (r - row * g0);
  }
  g -= n2 * yrow;
  const int row = g / (2 * NG), c = g - row * (2 * NG);
  const int kk = row / n1, jj = row - kk * n1;
  return (kk + NG) * plane + (jj + NG) * g0 + (c < NG ? c : n0 + c);
}

// boundary_condition (src/boundary_condition.cpp:4-231): every ghost cell
// takes its periodic image's value, gathered from the block (of the same
// solve: blocks[block0 + ...]) that owns the image.  The primitives of a
// ghost are its image's (calc_primitive is pointwise).  t < B.sg.
template <typename PT = double>
__device__ __forceinline__ void ghost_body(const BlockDesc& B, const BlockDesc* __restrict__ blocks,
                                           const Domain& dom, int block0, int t, int which, bool with_prim) {
  const int i = t % B.g[0] - NG;
  const int r = t / B.g[0];
  const int j = r % B.g[1] - NG;
  const int k = r / B.g[1] - NG;
  if (i >= 0 && i < B.n[0] && j >= 0 && j < B.n[1] && k >= 0 && k < B.n[2]) return;
  const int gl[3] = {B.o[0] + i, B.o[1] + j, B.o[2] + k};
  int bc[3], lc[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    int x = gl[a] % dom.N[a];
    x = x < 0 ? x + dom.N[a] : x;
    int bb = x / dom.base[a];
    bb = bb > dom.dims[a] - 1 ? dom.dims[a] - 1 : bb;
    bc[a] = bb;
    lc[a] = x - bb * dom.base[a];
  }
  const BlockDesc& S = blocks[block0 + (bc[0] * dom.dims[1] + bc[1]) * dom.dims[2] + bc[2]];
  const int src = gidx(S, lc[0], lc[1], lc[2]);
  gptr<const double> __restrict__ Qs = gmem<const double>(S.Q[which]);
  gptr<double> __restrict__ Qd = gmem(B.Q[which]);
#pragma unroll
  for (int e = 0; e < NEQ; e++) Qd[e * B.sg + t] = Qs[e * S.sg + src];
  if (with_prim) {
#pragma unroll
    for (int f = 0; f < 6; f++) gmem((PT*)B.prim)[f * B.sg + t] = gmem<const PT>((const PT*)S.prim)[f * S.sg + src];
  }
}

// ---- calc_dt spectral radius of one interior cell (src/Block3d.cpp:527-546)
// Returns t1 + t2 + t3; *bad is set for a NaN or infinite result.
// PT = float (FP32-flux form, float primitives): the FP64 primitives are
// formed from Q by calc_primitive's expressions instead of read.
template <typename PT = double>
__device__ __forceinline__ double dt_cell(const BlockDesc& B, int t, const Phys& ph, int* bad) {
  const int n0 = B.n[0], n1 = B.n[1];
  const int i = t % n0;
  const int r = t / n0;
  const int j = r % n1;
  const int k = r / n1;
  const int c = gidx(B, i, j, k);
  const long long sg = B.sg;
  double rr, uu, vv, ww, pp;
  if constexpr (sizeof(PT) == 8) {
    gptr<const double> __restrict__ P = gmem<const double>(B.prim);
    rr = P[c], uu = P[sg + c], vv = P[2 * sg + c], ww = P[3 * sg + c];
    pp = P[4 * sg + c];
  } else {  // src/Block3d.cpp:369-373
    gptr<const double> __restrict__ Q = gmem<const double>(B.Q[0]);
    rr = Q[c];
    uu = Q[sg + c] / rr;
    vv = Q[2 * sg + c] / rr;
    ww = Q[3 * sg + c] / rr;
    pp = ph.gam1 * (Q[4 * sg + c] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
  }
  return dt_point(rr, uu, vv, ww, pp, gmem<const double>(B.mu)[t], ph, bad);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

}  // namespace cbns
