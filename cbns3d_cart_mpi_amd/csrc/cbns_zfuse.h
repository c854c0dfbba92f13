// The z-face sweep fused with the stage tail: rec_riemann_z
// (src/Block3d.cpp:690-748) and update_rk3 with the viscous contribution and
// calc_primitive of the updated cell (src/Block3d.cpp:870-962, 359-385) in
// one pass, for the bit-exact FP64, the FP32-flux and the tolerance-bound
// FP64 forms.
//
// A thread owns one column (i, j) and a run of cells along z.  It marches
// the faces of its run with a six-cell register queue of the stage input, so
// each cell of Q[in] is loaded once per run (plus the five more its first
// face needs) and no z-flux reaches HBM; after face k it updates cell k - 1
// from faces k - 1 and k.  The operands of a cell's update (x- and y-face
// fluxes, viscous divergence, Q^n) are loaded before its face's arithmetic,
// so that the tail's HBM traffic runs under the face arithmetic, which
// bounds the z sweep on its own.  The stage input is never written (stage 1:
// Q -> S, stage 2: S -> Q_p, stage 3: Q_p -> Q, with S a third state), so
// the runs of a column need no ordering.
//
// The face of each form is its sweep kernels' arithmetic on the same
// operands: the primitives of the two cells beside a face (which the plain
// and FP32-flux sweeps read from the stored primitives) are formed here by
// calc_primitive's expressions from the queued state, which gives the
// stored values (rounded to float for FP32-flux as stored there); the
// FP32-flux face fluxes are rounded to float as its face arrays store them.
// Every value therefore equals the separate sweep and tail kernels'.
// The tolerance-bound form's x and y faces are formed by a march of the same
// shape along y (xyrun_fast), x faces across the lanes of a wave, y faces
// along each lane's column: it stores, per cell, the x- plus y-flux
// differences in the reference's order and rounding (src/Block3d.cpp:932-938),
// so no x or y face flux reaches memory and its fused z pass reads five
// values per cell instead of the x- and y-face fluxes of both sides.
// Shared by the single-solve kernels (cbns_kernels_zfuse.hip) and the
// batched ones (cbns_batch.hip).
#pragma once
#include "cbns_face_fast.h"

namespace cbns {

enum { ZF_FP64 = 0, ZF_FP32 = 1, ZF_FAST = 2 };

template <int FORM>
using zft = typename std::conditional<FORM == ZF_FP32, float, double>::type;

// Operands of the update of one cell other than its z-fluxes and its stage
// input, in their storage types.  XY (tolerance-bound form): xy holds the
// cell's x- plus y-flux differences less its viscous divergence
// (xyrun_fast), and e0 .. f1 and dv are unused.
template <typename FT>
struct ZOperands {
  FT e0[5], e1[5], f0[5], f1[5], dv[4];
  double xy[5];
  double qn[5];
};

template <typename FT, bool XY = false>
__device__ __forceinline__ void zload(const BlockDesc& B, int i, int j, int k, ZOperands<FT>& o, int stage = 2) {
  const int n0 = B.n[0], n1 = B.n[1];
  const int iE = i + (n0 + 1) * (j + n1 * k), iF = i + n0 * (j + (n1 + 1) * k), ct = i + n0 * (j + n1 * k);
  gptr<const FT> __restrict__ E = gmem<const FT>((const FT*)B.F[0]);
  gptr<const FT> __restrict__ F = gmem<const FT>((const FT*)B.F[1]);
  gptr<const FT> __restrict__ DV = gmem<const FT>((const FT*)B.dfv);
  gptr<const double> __restrict__ Qn = gmem<const double>(B.Q[0]);
  const int c = gidx(B, i, j, k);
  // stage 1 takes Q^n from the stage input: its loads point all lanes at one line
  const long long qs = stage > 1 ? B.sg : 0;
  const int qc = stage > 1 ? c : 0;
#pragma unroll
  for (int e = 0; e < 5; e++) {
    if constexpr (XY) {
      o.xy[e] = gmem<const double>(B.F[1])[e * B.sf[1] + ct];
    } else {
      o.e0[e] = E[e * B.sf[0] + iE];
      o.e1[e] = E[e * B.sf[0] + iE + 1];
      o.f0[e] = F[e * B.sf[1] + iF];
      o.f1[e] = F[e * B.sf[1] + iF + n0];
    }
    o.qn[e] = Qn[e * qs + qc];  // used by stages 2 and 3 (loaded always: no branch among the loads)
  }
  if constexpr (!XY) {
#pragma unroll
    for (int e = 0; e < 4; e++) o.dv[e] = DV[e * B.sn + ct];
  }
}

// x- plus y-flux differences of one cell from its x-face fluxes e0, e1 and
// y-face fluxes f0, f1, as update_rk3 sums them (src/Block3d.cpp:932-938);
// R (tolerance-bound form): the divisions by the spacings as divr's
// multiplications by their reciprocals.
template <bool R = false>
__device__ __forceinline__ double xy_diff(double e0, double e1, double f0, double f1, const Phys& ph) {
  double df = divr<R>(e1 - e0, ph.dx, ph.rdx, ph.fdx);
  df += divr<R>(f1 - f0, ph.dy, ph.rdy, ph.fdy);
  return df;
}

// The update of interior cell (i, j, k) with the expressions and order of
// update_pair_body: x-, y- and z-flux differences less the viscous
// divergence (the tolerance-bound form: its x+y differences already less
// the divergence, then the z differences), then the RK3 combination of Q^n
// (for stage 1 the stage input qi itself) and qi, into Q[out]; the plain
// and FP32-flux forms also store the cell's primitives.  With want_s, the calc_dt spectral radius of the
// updated cell (src/Block3d.cpp:527-546, mu of this stage's viscous pass)
// goes into *s (a NaN never wins; *bad flags a non-finite one).
template <int FORM>
__device__ __forceinline__ void zapply(const BlockDesc& B, int i, int j, int k, const ZOperands<zft<FORM>>& o,
                                       const double (&gk)[5], const double (&gk1)[5], const double (&qi)[5], int out,
                                       int stage, double dt, const Phys& ph, bool want_s, double* s, int* bad) {
  using FT = zft<FORM>;
  gptr<double> __restrict__ Qo = gmem(B.Q[out]);
  const long long sg = B.sg;
  const int c = gidx(B, i, j, k);
  double r[5];
#pragma unroll
  for (int e = 0; e < 5; e++) {
    double df = FORM == ZF_FAST ? o.xy[e]
                                : xy_diff((double)o.e0[e], (double)o.e1[e], (double)o.f0[e], (double)o.f1[e], ph);
    df += divr<FORM == ZF_FAST>(gk1[e] - gk[e], ph.dz, ph.rdz, ph.fdz);
    if (FORM != ZF_FAST && e > 0) df -= (double)o.dv[e - 1];  // the tolerance-bound xy holds it
    if (1 == stage) r[e] = qi[e] - dt * df;
    else if (2 == stage) r[e] = 0.75 * o.qn[e] + 0.25 * (qi[e] - dt * df);
    else r[e] = (o.qn[e] + 2.0 * (qi[e] - dt * df)) / 3.0;
    Qo[e * sg + c] = r[e];
  }
  if (FORM == ZF_FAST && !want_s) return;
  const double rr = r[0];
  const double uu = r[1] / rr;
  const double vv = r[2] / rr;
  const double ww = r[3] / rr;
  const double pp = ph.gam1 * (r[4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
  if constexpr (FORM != ZF_FAST) {
    gptr<FT> __restrict__ P = gmem((FT*)B.prim);
    P[c] = (FT)rr;
    P[sg + c] = (FT)uu;
    P[2 * sg + c] = (FT)vv;
    P[3 * sg + c] = (FT)ww;
    P[4 * sg + c] = (FT)pp;
    P[5 * sg + c] = (FT)(ph.gM2 * pp / rr);
  }
  if (want_s) {
    const double sc = dt_point(rr, uu, vv, ww, pp, gmem<const double>(B.mu)[i + B.n[0] * (j + B.n[1] * k)], ph, bad);
    *s = (*s < sc) ? sc : *s;
  }
}

// rho, u, v, w, p of one cell by calc_primitive's expressions
// (src/Block3d.cpp:369-373), as the plain and FP32-flux sweeps read them
// from the stored primitives (rounded to float for FP32-flux).
template <int FORM>
__device__ __forceinline__ void zprim(const double (&q)[5], const Phys& ph, double (&p)[5]) {
  const double rr = q[0];
  const double uu = q[1] / rr;
  const double vv = q[2] / rr;
  const double ww = q[3] / rr;
  const double pp = ph.gam1 * (q[4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
  const double v[5] = {rr, uu, vv, ww, pp};
#pragma unroll
  for (int f = 0; f < 5; f++) p[f] = FORM == ZF_FP32 ? (double)(float)v[f] : v[f];
}

// One thread's run: cells [k0, k0 + zm) of column (i, j) of block B (index
// b of the block array), i < n0, j < n1, k0 < n2.  The tolerance-bound and
// plain forms leave the cells beside a face whose range checks fail to cfx:
// the cell fixup recomputes such a face with the reference-order form (the
// plain form's sweeps recompute it in place; the result is the same) and
// updates the cells.  Keeping that form out of the run keeps its registers
// to the checked form's.
template <int FORM>
__device__ __forceinline__ void zrun(const BlockDesc& B, int b, int i, int j, int k0, int zm, int in, int out,
                                     int stage, double dt, const Phys& ph, const RefState& ref, bool want,
                                     double* s, int* bad, const FixList& cfx) {
  using FT = zft<FORM>;
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int k1 = k0 + zm < n2 ? k0 + zm : n2;  // cells [k0, k1), faces k0 .. k1
  const int sz = B.g[0] * B.g[1];
  const long long sg = B.sg;
  gptr<const double> __restrict__ Qin = gmem<const double>(B.Q[in]);
  const int c0 = gidx(B, i, j, 0);
  double q[6][5];  // cells k - 3 .. k + 2 of face k (planes -3 .. n2 + 2: inside the ghosts)
#pragma unroll
  for (int cc = 0; cc < 6; cc++)
#pragma unroll
    for (int e = 0; e < 5; e++) q[cc][e] = Qin[e * sg + c0 + (k0 - 3 + cc) * sz];
  double pl[5];  // primitives of cell k - 1 (FP32-flux form; the plain form forms them at each face)
  if constexpr (FORM != ZF_FAST) zprim<FORM>(q[2], ph, pl);
  double gp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  bool okp = true;
  for (int k = k0; k <= k1; k++) {
    // The other forms issue the loads of this step (the operands of cell
    // k - 1 and the next queued cell, plane k + 3 <= n2 + 3, inside the
    // ghosts) before the face, with no branch among them, so that the
    // compiler's wait for them falls after the face arithmetic; the first
    // face of a run updates no cell and loads cell k0's operands unused.
    // The plain form's heavier face leaves no registers to hold values
    // across it: it loads them after the face and forms the primitives of
    // both of its cells at each face.
    constexpr bool early = FORM != ZF_FP64;
    ZOperands<FT> o;
    if (early) zload<FT, FORM == ZF_FAST>(B, i, j, k > k0 ? k - 1 : k0, o, stage);
    double nx[5];
    if (early) {
#pragma unroll
      for (int e = 0; e < 5; e++) nx[e] = Qin[e * sg + c0 + (k + 3) * sz];
    }
    double g[5];
    bool ok = true;
    if constexpr (FORM == ZF_FAST) {
      ok = face_flux_fast<2>(q, ph.gamma, ph.gam1, ph.gg1, g);
    } else {
      double pr[5];  // primitives of cell k
      zprim<FORM>(q[3], ph, pr);
      if constexpr (FORM == ZF_FP64) zprim<FORM>(q[2], ph, pl);
      if constexpr (FORM == ZF_FP32) {
        float g32[5];
        face_flux_f32<2>(q, pl, pr, ph.gamma, ph.gam1, ref, g32);
#pragma unroll
        for (int e = 0; e < 5; e++) g[e] = (double)g32[e];
      } else {
        face_flux<2, false, true>(q, pl, pr, ph.gamma, ph.gam1, ph.rgam1, ph.fgam1, g, ok);
      }
#pragma unroll
      for (int f = 0; f < 5; f++) pl[f] = pr[f];
    }
    if (k > k0) {  // cell k - 1 is q[2]
      if (!early && ok && okp) zload<FT>(B, i, j, k - 1, o, stage);
      // a tolerance-bound cell beside a failing y face carries a NaN mark
      const bool okxy = FORM != ZF_FAST || o.xy[0] == o.xy[0];
      if (ok && okp && okxy) zapply<FORM>(B, i, j, k - 1, o, gp, g, q[2], out, stage, dt, ph, want, s, bad);
      else {
        const unsigned slot = atomicAdd(cfx.count, 1u);
        if (slot < cfx.cap)
          cfx.ent[slot] = (unsigned long long)(i + (long long)n0 * (j + (long long)n1 * (k - 1))) |
                          ((unsigned long long)b << 42);
      }
    }
    if (!early && k < k1) {
#pragma unroll
      for (int e = 0; e < 5; e++) nx[e] = Qin[e * sg + c0 + (k + 3) * sz];
    }
    if (early || k < k1) {
#pragma unroll
      for (int cc = 0; cc < 5; cc++)
#pragma unroll
        for (int e = 0; e < 5; e++) q[cc][e] = q[cc + 1][e];
#pragma unroll
      for (int e = 0; e < 5; e++) q[5][e] = nx[e];
    }
#pragma unroll
    for (int e = 0; e < 5; e++) gp[e] = g[e];
    okp = ok;
  }
}

// One face of a listed cell: the checked form, and where its checks fail
// the reference-order form with IEEE division and square root.  Returns
// whether the checked form passed.
template <int FORM>
__device__ __forceinline__ bool zface_checked(const double (&q)[6][5], const Phys& ph, bool recompute,
                                              double (&g)[5]) {
  bool ok = true;
  if constexpr (FORM == ZF_FAST) {
    ok = face_flux_fast<2>(q, ph.gamma, ph.gam1, ph.gg1, g);
    if (!ok && recompute) face_flux_fast_fallback<2>(q, ph, g);
  } else {
    double pl[5], pr[5];
    zprim<FORM>(q[2], ph, pl);
    zprim<FORM>(q[3], ph, pr);
    face_flux<2, false, true>(q, pl, pr, ph.gamma, ph.gam1, ph.rgam1, ph.fgam1, g, ok);
    if (!ok && recompute) {
      bool chk = true;
      face_flux<2, false, false>(q, pl, pr, ph.gamma, ph.gam1, ph.rgam1, ph.fgam1, g, chk);
    }
  }
  return ok;
}

// A listed cell: both of its z-faces (zface_checked), then its update.  With
// all set (more failures than the list holds), a cell is updated only if
// one of its faces fails its checks: exactly the cells the fused pass left.
// The tolerance-bound form's cells also take their x and y faces here (its
// x+y march stores no face fluxes), and a cell is left alone in the all
// pass only if its x and y faces pass as well.
template <int FORM>
__device__ __forceinline__ void zcell_fix(const BlockDesc& B, int i, int j, int k, int in, int out, int stage,
                                          double dt, const Phys& ph, bool all, bool want_s, double* s, int* bad) {
  const int sz = B.g[0] * B.g[1];
  const long long sg = B.sg;
  gptr<const double> __restrict__ Qin = gmem<const double>(B.Q[in]);
  const int c = gidx(B, i, j, k);
  // the two faces along axis D of the cell from its seven cells along D
  const auto faces = [&](auto dc, double (&g0)[5], double (&g1)[5], bool& ok0, bool& ok1) {
    constexpr int D = decltype(dc)::value;
    const int sD = D == 0 ? 1 : (D == 1 ? B.g[0] : sz);
    double q7[7][5];
#pragma unroll
    for (int cc = 0; cc < 7; cc++)
#pragma unroll
      for (int e = 0; e < 5; e++) q7[cc][e] = Qin[e * sg + c + (cc - 3) * sD];
    double qa[6][5], qb[6][5];
#pragma unroll
    for (int cc = 0; cc < 6; cc++)
#pragma unroll
      for (int e = 0; e < 5; e++) {
        qa[cc][e] = q7[cc][e];
        qb[cc][e] = q7[cc + 1][e];
      }
    if constexpr (D == 2) {
      ok0 = zface_checked<FORM>(qa, ph, true, g0);
      ok1 = zface_checked<FORM>(qb, ph, true, g1);
    } else {
      ok0 = face_flux_fast<D>(qa, ph.gamma, ph.gam1, ph.gg1, g0);
      if (!ok0) face_flux_fast_fallback<D>(qa, ph, g0);
      ok1 = face_flux_fast<D>(qb, ph.gamma, ph.gam1, ph.gg1, g1);
      if (!ok1) face_flux_fast_fallback<D>(qb, ph, g1);
    }
  };
  double g0[5], g1[5], e0[5], e1[5], f0[5], f1[5];
  bool ok0, ok1, okx0 = true, okx1 = true, oky0 = true, oky1 = true;
  faces(std::integral_constant<int, 2>(), g0, g1, ok0, ok1);
  if constexpr (FORM == ZF_FAST) {
    faces(std::integral_constant<int, 0>(), e0, e1, okx0, okx1);
    faces(std::integral_constant<int, 1>(), f0, f1, oky0, oky1);
  }
  if (all && ok0 && ok1 && okx0 && okx1 && oky0 && oky1) return;
  ZOperands<double> o;
  zload<double, FORM == ZF_FAST>(B, i, j, k, o, stage);
  if constexpr (FORM == ZF_FAST) {  // as xyrun_fast forms them
    gptr<const double> __restrict__ DV = gmem<const double>(B.dfv);
    const int ct = i + B.n[0] * (j + B.n[1] * k);
#pragma unroll
    for (int e = 0; e < 5; e++) {
      o.xy[e] = xy_diff<true>(e0[e], e1[e], f0[e], f1[e], ph);
      if (e > 0) o.xy[e] -= DV[(e - 1) * B.sn + ct];
    }
  }
  double qc[5];
#pragma unroll
  for (int e = 0; e < 5; e++) qc[e] = Qin[e * sg + c];
  zapply<FORM>(B, i, j, k, o, g0, g1, qc, out, stage, dt, ph, want_s, s, bad);
}

// 64-bit value of lane src (uniform over the wave) to every lane.
__device__ __forceinline__ double lane_value(double v, int src) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Most rows per run of the x+y march: lane l forms the right x face of row
// j0 + l for the wave's last column.
constexpr int XY_RUN_MAX = 64;

// Tolerance-bound form: the x and y faces of a stage in one march, which
// stores each interior cell's x- plus y-flux differences (xy_diff<true>)
// less its viscous divergence (B.dfv, from the viscous pass of the stage,
// which runs before the march) into B.F[1] at the cell's interior index;
// no face flux reaches memory.  The fused z pass adds the z differences:
// the reference sums x, y, z, then subtracts the divergence
// (src/Block3d.cpp:932-940); the form moves the subtraction before the z
// term, a reassociation within its tolerance that takes four loads per cell
// off the memory-heavier z pass.
// One wave owns the 64 columns i0 .. i0 + 63 at one k (lane l: i = i0 + l)
// and marches rows [j0, j0 + ym), ym <= 64: the y faces j0 .. j1 with a
// six-cell register queue of the stage input along each lane's column (as
// zrun marches z).  At row j each lane also forms the x face on the left of
// its cell (cells i - 3 .. i + 2 of the row, face_flux_fast<0>) and takes the
// face on the right from lane l + 1; the last lane's right faces (x index
// i0 + 64, one per row of the run) are formed before the march, row j0 + l
// on lane l, when that face lies in the block.  Lanes past the block's last
// column form copies of face n0 (its stencil lies in the ghosts) and store
// nothing, so every lane of the wave takes part in every step and the lane
// exchanges see defined values.  Cell j's x part divr(E(i+1) - E(i), dx) is
// formed at row j and its y part added after y face j + 1, the order and
// rounding of xy_diff<true>, then the divergence subtracted.  A cell beside any x or y face whose range
// checks fail gets a NaN in place of its differences: the fused z pass then
// leaves it to the cell fixup, which forms its faces with the
// reference-order form.
// The caller guarantees k < n2, j0 < n1 and i0 < n0, uniform over the wave.
// The x stencil of a row is loaded after the row's y face (228 VGPRs, two
// waves per SIMD; loaded before it, the queue and the stencil together
// spill).
__device__ __forceinline__ void xyrun_fast(const BlockDesc& B, int i0, int k, int j0, int ym, int which,
                                           const Phys& ph) {
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  const int n0 = B.n[0], n1 = B.n[1];
  const int i = i0 + lane;
  // the lane's queue column: its own up to n0 + 2, the last column an x
  // stencil of a stored cell reaches (ghost columns are marched, not stored)
  const int cq = i < n0 + 2 ? i : n0 + 2;
  const int j1 = j0 + ym < n1 ? j0 + ym : n1;  // cells [j0, j1), y faces j0 .. j1
  const int sy = B.g[0];
  const long long sg = B.sg;
  gptr<const double> __restrict__ Qin = gmem<const double>(B.Q[which]);
  gptr<double> __restrict__ XY = gmem(B.F[1]);
  gptr<const double> __restrict__ DV = gmem<const double>(B.dfv);
  const long long sfx = B.sf[1], sn = B.sn;
  const int c0 = gidx(B, cq, 0, k);
  // Row buffers of the x stencils, two per wave (rows alternate): position p
  // holds cell i0 - 3 + p of the row.  Each lane writes its queued cell at
  // lane + 3; lanes 0 .. 4 load and write the halo cells i0 - 3 .. i0 - 1 and
  // i0 + 64, i0 + 65 (clamped into the ghosts where only lanes past the
  // block read them).  Only the wave itself touches its buffers, and its LDS
  // accesses complete in order, so no barrier is needed.
  constexpr int XW = 72;
  __shared__ double xrow[2][ZY][NEQ][XW];
  const bool hl = lane < 5;
  const int hcell = lane < 3 ? i0 - 3 + lane : (i0 + 61 + lane < n0 + 3 ? i0 + 61 + lane : n0 + 3);
  const int hpos = lane < 3 ? lane : 64 + lane;
  const int h0 = gidx(B, hl ? hcell : cq, 0, k);
  // the last lane's right faces: face i0 + 64 of row j0 + lane (rows past
  // the run repeat its last row)
  double ex[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int okex = 1;
  if (i0 + 64 <= n0) {  // uniform over the wave
    const int jr = j0 + lane < j1 ? j0 + lane : j1 - 1;
    double xq[6][5];
    load_face_cells<0>(B, which, gidx(B, i0 + 64, jr, k), xq);
    okex = face_flux_fast<0>(xq, ph.gamma, ph.gam1, ph.gg1, ex);
  }
  double q[6][5];  // cells j - 3 .. j + 2 of y face j (rows -3 .. n1 + 2: inside the ghosts)
#pragma unroll
  for (int cc = 0; cc < 6; cc++)
#pragma unroll
    for (int e = 0; e < 5; e++) q[cc][e] = Qin[e * sg + c0 + (j0 - 3 + cc) * sy];
  {  // row j0 into buffer 0
    double h[5];
#pragma unroll
    for (int e = 0; e < 5; e++) h[e] = hl ? Qin[e * sg + h0 + j0 * sy] : 0.0;
#pragma unroll
    for (int e = 0; e < 5; e++) {
      xrow[0][w][e][lane + 3] = q[3][e];
      if (hl) xrow[0][w][e][hpos] = h[e];
    }
  }
  double gp[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, xp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  bool okp = true, okxp = true;
  for (int j = j0; j <= j1; j++) {
    const bool xrow_j = j < j1;   // uniform: the last y face of a run closes cell j1 - 1 only
    const bool next = j + 1 < j1;  // uniform: row j + 1 has x faces in this run
    const int bj = (j - j0) & 1;
    double nx[5];
#pragma unroll
    for (int e = 0; e < 5; e++) nx[e] = Qin[e * sg + c0 + (j + 3) * sy];  // row j + 3 <= n1 + 3
    double g[5];
    const bool ok = face_flux_fast<1>(q, ph.gamma, ph.gam1, ph.gg1, g);
    // the divergence of cell j - 1 (a lane past the block reads cell n0 - 1's)
    double dv[4];
    const int ct = (i < n0 ? i : n0 - 1) + n0 * ((j > j0 ? j - 1 : j0) + n1 * k);
#pragma unroll
    for (int e = 0; e < 4; e++) dv[e] = DV[e * sn + ct];
    double xd[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    bool okx = true;
    if (xrow_j) {
      // the halo of row j + 1 goes out before this row's x face
      double hn[5];
#pragma unroll
      for (int e = 0; e < 5; e++) hn[e] = (next && hl) ? Qin[e * sg + h0 + (j + 1) * sy] : 0.0;
      double xq[6][5];
#pragma unroll
      for (int cc = 0; cc < 6; cc++)
#pragma unroll
        for (int e = 0; e < 5; e++) xq[cc][e] = xrow[bj][w][e][lane + cc];
      double gx[5];
      const int okl = face_flux_fast<0>(xq, ph.gamma, ph.gam1, ph.gg1, gx);
      const int src = j - j0;
#pragma unroll
      for (int e = 0; e < 5; e++) {
        const double right = __shfl_down(gx[e], 1, 64), last = lane_value(ex[e], src);
        xd[e] = divr<true>((lane == 63 ? last : right) - gx[e], ph.dx, ph.rdx, ph.fdx);
      }
      const int okr = __shfl_down(okl, 1, 64), okl63 = __builtin_amdgcn_readlane(okex, src);
      okx = okl && (lane == 63 ? okl63 : okr);
      if (next) {  // row j + 1 (the queue's cell j + 1) into the other buffer
#pragma unroll
        for (int e = 0; e < 5; e++) {
          xrow[bj ^ 1][w][e][lane + 3] = q[4][e];
          if (hl) xrow[bj ^ 1][w][e][hpos] = hn[e];
        }
      }
    }
    if (j > j0 && i < n0) {  // cell j - 1
      const bool good = ok && okp && okxp;
#pragma unroll
      for (int e = 0; e < 5; e++) {
        double df = xp[e];
        df += divr<true>(g[e] - gp[e], ph.dy, ph.rdy, ph.fdy);
        if (e > 0) df -= dv[e - 1];
        XY[e * sfx + ct] = good ? df : __builtin_nan("");
      }
    }
#pragma unroll
    for (int cc = 0; cc < 5; cc++)
#pragma unroll
      for (int e = 0; e < 5; e++) q[cc][e] = q[cc + 1][e];
#pragma unroll
    for (int e = 0; e < 5; e++) {
      q[5][e] = nx[e];
      gp[e] = g[e];
      xp[e] = xd[e];
    }
    okp = ok;
    okxp = okx;
  }
}

}  // namespace cbns
