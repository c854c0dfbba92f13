// Face-flux sweeps of the tolerance-bound FP64 form (CBNS_FP64_FAST):
// rec_riemann_x/y/z (src/Block3d.cpp:555-748) with the restructured face
// arithmetic of cbns_face_fast.h (its fused multiply-adds written out, so
// the batched kernels that inline the same bodies form the same values).
// The sweeps read only the state (the primitives of the two cells beside a
// face are formed from it); everything else of a step is shared with the
// plain path.  The fixup pass recomputes a failing face with the plain
// path's IEEE arithmetic.
//
// A face whose range checks fail is not recomputed in the sweep: the sweep
// appends it to a FixList and a fixup pass after the sweeps recomputes it
// with the reference-order form (IEEE division and square root).  Keeping
// that form out of the sweep kernels keeps their registers to the fast
// path's own (4 waves per SIMD for single faces, 3 for pairs, against 2
// with the recompute inlined).
#include "cbns_batch.h"
#include "cbns_face_fast.h"

namespace cbns {

static inline unsigned nblocks_fast(long long count) { return (unsigned)((count + TPB - 1) / TPB); }

FixList fixlist_at(void* buf) {
  FixList f;
  f.count = (unsigned*)buf;
  f.done = (unsigned*)buf + 1;
  f.ent = (unsigned long long*)((char*)buf + 256);
  f.cap = FIX_CAP;
  return f;
}

// OCC: waves per SIMD the register allocation must allow (4 for single
// faces, 3 for pairs; both fit without spills).
template <int D, bool PAIR, int OCC>
__global__ __launch_bounds__(TPB, OCC) void k_face_flux_fast(const BlockDesc* __restrict__ blocks, int which, int jb,
                                                             Phys ph, FixList fx) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[tl.y];
  const int t = tl.x * TPB + threadIdx.x;
  if constexpr (PAIR) {
    if (t >= face_pair_count(B, D)) return;
    face_pair_body_fast<D>(B, tl.y, t, which, jb, ph, fx);
  } else {
    if (t >= face_count(B, D)) return;
    face_body_fast<D>(B, tl.y, t, which, jb, ph, fx);
  }
}

hipError_t launch_face_flux_fast(const BlockDesc* db, const BlockDesc* hb, int nb, int dir, int which,
                                 const Phys& ph, int jb, int occ, const FixList& fx, hipStream_t s) {
  long long mx = 0;
  for (int b = 0; b < nb; b++) mx = std::max(mx, face_items(hb[b], dir, true));
  const dim3 grid(nblocks_fast(mx), nb);
  // occ: waves per SIMD of the pair kernels (y, z), 2 or 3; single faces (x) fit 4
  if (dir == 0) hipLaunchKernelGGL((k_face_flux_fast<0, false, 4>), grid, dim3(TPB), 0, s, db, which, jb, ph, fx);
  else if (occ >= 3) {
    if (dir == 1) hipLaunchKernelGGL((k_face_flux_fast<1, true, 3>), grid, dim3(TPB), 0, s, db, which, jb, ph, fx);
    else hipLaunchKernelGGL((k_face_flux_fast<2, true, 3>), grid, dim3(TPB), 0, s, db, which, jb, ph, fx);
  } else {
    if (dir == 1) hipLaunchKernelGGL((k_face_flux_fast<1, true, 2>), grid, dim3(TPB), 0, s, db, which, jb, ph, fx);
    else hipLaunchKernelGGL((k_face_flux_fast<2, true, 2>), grid, dim3(TPB), 0, s, db, which, jb, ph, fx);
  }
  return hipGetLastError();
}

// ---- fixup of the faces whose checks failed --------------------------------
template <int D>
__device__ __forceinline__ void fixup_face(const BlockDesc& B, int which, long long fi, const Phys& ph) {
  const int ni = B.n[0] + (D == 0), nj = B.n[1] + (D == 1);
  const int i = (int)(fi % ni);
  const long long r = fi / ni;
  const int j = (int)(r % nj), k = (int)(r / nj);
  double q[6][5], fl[5];
  load_face_cells<D>(B, which, gidx(B, i, j, k), q);
  face_flux_fast_fallback<D>(q, ph, fl);
  gptr<double> __restrict__ F = gmem(B.F[D]);
#pragma unroll
  for (int e = 0; e < 5; e++) F[e * B.sf[D] + fi] = fl[e];
}

// Every face of sweep D of block B checked again (more failures than the
// list holds): faces that fail are recomputed, the others are left as the
// sweep wrote them.
template <int D>
__device__ __forceinline__ void fixup_all(const BlockDesc& B, int which, const Phys& ph, long long first,
                                          long long stride) {
  const long long n = face_count(B, D);
  for (long long fi = first; fi < n; fi += stride) {
    const int ni = B.n[0] + (D == 0), nj = B.n[1] + (D == 1);
    const int i = (int)(fi % ni);
    const long long r = fi / ni;
    const int j = (int)(r % nj), k = (int)(r / nj);
    double q[6][5], fl[5];
    load_face_cells<D>(B, which, gidx(B, i, j, k), q);
    if (!face_flux_fast<D>(q, ph.gamma, ph.gam1, ph.gg1, fl)) {
      face_flux_fast_fallback<D>(q, ph, fl);
      gptr<double> __restrict__ F = gmem(B.F[D]);
#pragma unroll
      for (int e = 0; e < 5; e++) F[e * B.sf[D] + fi] = fl[e];
    }
  }
}

constexpr int FIX_WG = 64;  // workgroups of the fixup pass

// BATCH: blocks of many solves, each with its own physics (insts); the
// single-solve form passes ph.  In the pass over every face, only blocks of
// stepping tolerance-bound solves are checked.
template <bool BATCH>
__global__ __launch_bounds__(TPB) void k_face_fixup(const BlockDesc* __restrict__ blocks,
                                                    const InstDesc* __restrict__ insts, int nb, int which, Phys ph,
                                                    FixList fx) {
  const unsigned n = *fx.count;  // every workgroup reads it before the last one resets it (below)
  const long long first = blockIdx.x * (long long)TPB + threadIdx.x, stride = (long long)gridDim.x * TPB;
  if (n <= fx.cap) {
    for (long long e = first; e < n; e += stride) {
      const unsigned long long v = fx.ent[e];
      const long long fi = (long long)(v & ((1ull << 40) - 1));
      const int D = (int)((v >> 40) & 3), b = (int)(v >> 42);
      const Phys& P = BATCH ? insts[blocks[b].inst].ph : ph;
      if (D == 0) fixup_face<0>(blocks[b], which, fi, P);
      else if (D == 1) fixup_face<1>(blocks[b], which, fi, P);
      else fixup_face<2>(blocks[b], which, fi, P);
    }
  } else {
    for (int b = 0; b < nb; b++) {
      const BlockDesc& B = blocks[b];
      if (BATCH && (!insts[B.inst].stepping || insts[B.inst].precision != PREC_FP64_FAST)) continue;
      const Phys& P = BATCH ? insts[B.inst].ph : ph;
      fixup_all<0>(B, which, P, first, stride);
      fixup_all<1>(B, which, P, first, stride);
      fixup_all<2>(B, which, P, first, stride);
    }
  }
  // the last workgroup to finish clears the list for the next stage
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(fx.done, 1u) == gridDim.x - 1) {
      atomicExch(fx.count, 0u);
      atomicExch(fx.done, 0u);
    }
  }
}

hipError_t launch_face_fixup(const BlockDesc* db, const BlockDesc* hb, int nb, int which, const Phys& ph,
                             const FixList& fx, hipStream_t s) {
  (void)hb;
  hipLaunchKernelGGL(k_face_fixup<false>, dim3(FIX_WG), dim3(TPB), 0, s, db, (const InstDesc*)nullptr, nb, which, ph,
                     fx);
  return hipGetLastError();
}

hipError_t batch_launch_fixup(const BlockDesc* blocks, const InstDesc* insts, int nb, int which, const FixList& fx,
                              hipStream_t s) {
  hipLaunchKernelGGL(k_face_fixup<true>, dim3(FIX_WG), dim3(TPB), 0, s, blocks, insts, nb, which, Phys{}, fx);
  return hipGetLastError();
}

// ---- z faces fused with the stage tail (single solve) ----------------------
// rec_riemann_z (src/Block3d.cpp:690-748) and update_rk3 with the viscous
// contribution (src/Block3d.cpp:870-962) in one pass.  A thread owns one
// column (i, j) and a run of ZM cells along z: it marches the faces of the
// run with a six-cell register queue of the stage input, so each cell of
// Q[in] is loaded once per run (plus the five more its first face needs)
// and no z-flux reaches HBM; after face k it updates cell k - 1 from faces
// k - 1 and k.  The operands of a cell's update are loaded before its
// face's arithmetic, so that the tail's HBM traffic runs under the face
// arithmetic.  The stage input is never written (stage 1: Q -> S, stage 2:
// S -> Q_p, stage 3: Q_p -> Q, with S a third state), so runs of one column
// need no ordering.
constexpr int ZX = 64, ZY = 4, ZM = 16;  // columns per workgroup along x, y; cells per run

// Operands of the update of one cell other than its z-fluxes and its stage
// input: its x- and y-face fluxes, the viscous divergence and Q^n.
struct ZOperands {
  double e0[5], e1[5], f0[5], f1[5], dv[4], qn[5];
};

__device__ __forceinline__ void zload(const BlockDesc& B, int i, int j, int k, int stage, ZOperands& o) {
  const int n0 = B.n[0], n1 = B.n[1];
  const int iE = i + (n0 + 1) * (j + n1 * k), iF = i + n0 * (j + (n1 + 1) * k), ct = i + n0 * (j + n1 * k);
  gptr<const double> __restrict__ E = gmem<const double>(B.F[0]);
  gptr<const double> __restrict__ F = gmem<const double>(B.F[1]);
  gptr<const double> __restrict__ DV = gmem<const double>(B.dfv);
  gptr<const double> __restrict__ Qn = gmem<const double>(B.Q[0]);
  const int c = gidx(B, i, j, k);
#pragma unroll
  for (int e = 0; e < 5; e++) {
    o.e0[e] = E[e * B.sf[0] + iE];
    o.e1[e] = E[e * B.sf[0] + iE + 1];
    o.f0[e] = F[e * B.sf[1] + iF];
    o.f1[e] = F[e * B.sf[1] + iF + n0];
    if (stage > 1) o.qn[e] = Qn[e * B.sg + c];
  }
#pragma unroll
  for (int e = 0; e < 4; e++) o.dv[e] = DV[e * B.sn + ct];
}

// The update of interior cell (i, j, k) with the expressions and order of
// update_pair_body: x-, y- and z-flux differences less the viscous
// divergence, then the RK3 combination of Q^n (for stage 1 the stage input
// qi itself) and qi, into Q[out].  With want_s, the calc_dt spectral radius
// of the updated cell (src/Block3d.cpp:527-546, mu of this stage's viscous
// pass) goes into *s (a NaN never wins; *bad flags a non-finite one).
__device__ __forceinline__ void zapply(const BlockDesc& B, int i, int j, int k, const ZOperands& o,
                                       const double (&gk)[5], const double (&gk1)[5], const double (&qi)[5], int out,
                                       int stage, double dt, const Phys& ph, bool want_s, double* s, int* bad) {
  gptr<double> __restrict__ Qo = gmem(B.Q[out]);
  const int c = gidx(B, i, j, k);
  double r[5];
#pragma unroll
  for (int e = 0; e < 5; e++) {
    double df = divc(o.e1[e] - o.e0[e], ph.dx, ph.rdx, ph.fdx);
    df += divc(o.f1[e] - o.f0[e], ph.dy, ph.rdy, ph.fdy);
    df += divc(gk1[e] - gk[e], ph.dz, ph.rdz, ph.fdz);
    if (e > 0) df -= o.dv[e - 1];
    if (1 == stage) r[e] = qi[e] - dt * df;
    else if (2 == stage) r[e] = 0.75 * o.qn[e] + 0.25 * (qi[e] - dt * df);
    else r[e] = (o.qn[e] + 2.0 * (qi[e] - dt * df)) / 3.0;
    Qo[e * B.sg + c] = r[e];
  }
  if (want_s) {
    const double rr = r[0];
    const double uu = r[1] / rr;
    const double vv = r[2] / rr;
    const double ww = r[3] / rr;
    const double pp = ph.gam1 * (r[4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
    const double sc = dt_point(rr, uu, vv, ww, pp, gmem<const double>(B.mu)[i + B.n[0] * (j + B.n[1] * k)], ph, bad);
    *s = (*s < sc) ? sc : *s;
  }
}

// Block-wide max of s into *smax (integer atomic max on the bits of a
// non-negative double; a NaN never wins) and the non-finite flag.  Made by
// every thread of the workgroup.
__device__ __forceinline__ void reduce_smax(double s, int bad, unsigned long long* smax, int* flag) {
  s = wave_max(s);
  __shared__ double red[TPB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int a = 1; a < (int)(blockDim.x / 64); a++) m = red[a] > m ? red[a] : m;
    atomicMax(smax, (unsigned long long)__double_as_longlong(m));
  }
  if (bad) atomicOr(flag, 1);
}

__global__ __launch_bounds__(ZX * ZY) void k_zface_update_fast(const BlockDesc* __restrict__ blocks, int in, int out,
                                                               int stage, double dt, Phys ph,
                                                               unsigned long long* smax, int* flag, FixList cfx) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[blockIdx.z];
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int ntx = (n0 + ZX - 1) / ZX;
  const int i = (tl.x % ntx) * ZX + (int)threadIdx.x % ZX;
  const int j = (tl.x / ntx) * ZY + (int)threadIdx.x / ZX;
  const int k0 = tl.y * ZM;
  const bool want = smax != nullptr;  // uniform
  double s = 0.0;
  int bad = 0;
  if (i < n0 && j < n1 && k0 < n2) {
    const int k1 = k0 + ZM < n2 ? k0 + ZM : n2;  // cells [k0, k1), faces k0 .. k1
    const int sz = B.g[0] * B.g[1];
    const long long sg = B.sg;
    gptr<const double> __restrict__ Qin = gmem<const double>(B.Q[in]);
    const int c0 = gidx(B, i, j, 0);
    double q[6][5];  // cells k - 3 .. k + 2 of face k (planes -3 .. n2 + 2: inside the ghosts)
#pragma unroll
    for (int cc = 0; cc < 6; cc++)
#pragma unroll
      for (int e = 0; e < 5; e++) q[cc][e] = Qin[e * sg + c0 + (k0 - 3 + cc) * sz];
    double gp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    bool okp = true;
    for (int k = k0; k <= k1; k++) {
      ZOperands o;
      if (k > k0) zload(B, i, j, k - 1, stage, o);
      double nx[5];
      if (k < k1) {
#pragma unroll
        for (int e = 0; e < 5; e++) nx[e] = Qin[e * sg + c0 + (k + 3) * sz];
      }
      double g[5];
      const bool ok = face_flux_fast<2>(q, ph.gamma, ph.gam1, ph.gg1, g);
      if (k > k0) {  // cell k - 1 is q[2]
        if (ok && okp) zapply(B, i, j, k - 1, o, gp, g, q[2], out, stage, dt, ph, want, &s, &bad);
        else {
          const unsigned slot = atomicAdd(cfx.count, 1u);
          if (slot < cfx.cap)
            cfx.ent[slot] = (unsigned long long)(i + (long long)n0 * (j + (long long)n1 * (k - 1))) |
                            ((unsigned long long)blockIdx.z << 42);
        }
      }
      if (k < k1) {
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
  if (want) reduce_smax(s, bad, smax, flag);
}

hipError_t launch_zface_update_fast(const BlockDesc* db, const BlockDesc* hb, int nb, int in, int out, int stage,
                                    double dt, const Phys& ph, unsigned long long* smax, int* flag,
                                    const FixList& cfx, hipStream_t s) {
  int tiles = 0, runs = 0;
  for (int b = 0; b < nb; b++) {
    tiles = std::max(tiles, ((hb[b].n[0] + ZX - 1) / ZX) * ((hb[b].n[1] + ZY - 1) / ZY));
    runs = std::max(runs, (hb[b].n[2] + ZM - 1) / ZM);
  }
  hipLaunchKernelGGL(k_zface_update_fast, dim3(tiles, runs, nb), dim3(ZX * ZY), 0, s, db, in, out, stage, dt, ph,
                     smax, flag, cfx);
  return hipGetLastError();
}

// A listed cell: both of its z-faces with the checked form (the fast form
// where its checks pass, the reference-order form where they fail), then its
// update.  With all set (more failures than the list holds), a cell is
// updated only if one of its faces fails its checks: exactly the cells the
// fused pass left.
__device__ __forceinline__ void zcell_fix(const BlockDesc& B, int i, int j, int k, int in, int out, int stage,
                                          double dt, const Phys& ph, bool all, bool want_s, double* s, int* bad) {
  const int sz = B.g[0] * B.g[1];
  const long long sg = B.sg;
  gptr<const double> __restrict__ Qin = gmem<const double>(B.Q[in]);
  const int c = gidx(B, i, j, k);
  double q7[7][5];
#pragma unroll
  for (int cc = 0; cc < 7; cc++)
#pragma unroll
    for (int e = 0; e < 5; e++) q7[cc][e] = Qin[e * sg + c + (cc - 3) * sz];
  double qa[6][5], qb[6][5];
#pragma unroll
  for (int cc = 0; cc < 6; cc++)
#pragma unroll
    for (int e = 0; e < 5; e++) {
      qa[cc][e] = q7[cc][e];
      qb[cc][e] = q7[cc + 1][e];
    }
  double g0[5], g1[5];
  const bool ok0 = face_flux_fast<2>(qa, ph.gamma, ph.gam1, ph.gg1, g0);
  const bool ok1 = face_flux_fast<2>(qb, ph.gamma, ph.gam1, ph.gg1, g1);
  if (all && ok0 && ok1) return;
  if (!ok0) face_flux_fast_fallback<2>(qa, ph, g0);
  if (!ok1) face_flux_fast_fallback<2>(qb, ph, g1);
  ZOperands o;
  zload(B, i, j, k, stage, o);
  zapply(B, i, j, k, o, g0, g1, q7[3], out, stage, dt, ph, want_s, s, bad);
}

__global__ __launch_bounds__(TPB) void k_zcell_fixup(const BlockDesc* __restrict__ blocks, int nb, int in, int out,
                                                     int stage, double dt, Phys ph, unsigned long long* smax,
                                                     int* flag, FixList cfx) {
  const unsigned n = *cfx.count;  // every workgroup reads it before the last one resets it (below)
  const long long first = blockIdx.x * (long long)TPB + threadIdx.x, stride = (long long)gridDim.x * TPB;
  const bool want = smax != nullptr;
  double s = 0.0;
  int bad = 0;
  if (n <= cfx.cap) {
    for (long long e = first; e < n; e += stride) {
      const unsigned long long v = cfx.ent[e];
      const long long ct = (long long)(v & ((1ull << 40) - 1));
      const BlockDesc& B = blocks[(int)(v >> 42)];
      const int i = (int)(ct % B.n[0]);
      const long long r = ct / B.n[0];
      zcell_fix(B, i, (int)(r % B.n[1]), (int)(r / B.n[1]), in, out, stage, dt, ph, false, want, &s, &bad);
    }
  } else {
    for (int b = 0; b < nb; b++) {
      const BlockDesc& B = blocks[b];
      for (long long ct = first; ct < B.sn; ct += stride) {
        const int i = (int)(ct % B.n[0]);
        const long long r = ct / B.n[0];
        zcell_fix(B, i, (int)(r % B.n[1]), (int)(r / B.n[1]), in, out, stage, dt, ph, true, want, &s, &bad);
      }
    }
  }
  if (want) reduce_smax(s, bad, smax, flag);
  // the last workgroup to finish clears the list for the next stage
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(cfx.done, 1u) == gridDim.x - 1) {
      atomicExch(cfx.count, 0u);
      atomicExch(cfx.done, 0u);
    }
  }
}

hipError_t launch_zcell_fixup(const BlockDesc* db, const BlockDesc* hb, int nb, int in, int out, int stage,
                              double dt, const Phys& ph, unsigned long long* smax, int* flag, const FixList& cfx,
                              hipStream_t s) {
  (void)hb;
  hipLaunchKernelGGL(k_zcell_fixup, dim3(FIX_WG), dim3(TPB), 0, s, db, nb, in, out, stage, dt, ph, smax, flag, cfx);
  return hipGetLastError();
}

// Tolerance-bound face flux on explicit inputs (parity tests): cells
// [nf][6][5] of cells i-3 .. i+2, flux [nf][5]; a face failing its checks
// takes the reference-order form in place.
template <int D>
__global__ __launch_bounds__(TPB) void k_face_eval_fast(Phys ph, int nf, const double* __restrict__ cells,
                                                        double* __restrict__ flux) {
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= nf) return;
  double q[6][5], fl[5];
#pragma unroll
  for (int c = 0; c < 6; c++)
#pragma unroll
    for (int e = 0; e < 5; e++) q[c][e] = cells[(long long)t * 30 + 5 * c + e];
  face_flux_fast_checked<D>(q, ph, fl);
#pragma unroll
  for (int e = 0; e < 5; e++) flux[(long long)t * 5 + e] = fl[e];
}

hipError_t launch_face_eval_fast(int dir, const Phys& ph, int nf, const double* cells, double* flux, hipStream_t s) {
  const dim3 grid(nblocks_fast(nf));
  if (dir == 0) hipLaunchKernelGGL(k_face_eval_fast<0>, grid, dim3(TPB), 0, s, ph, nf, cells, flux);
  else if (dir == 1) hipLaunchKernelGGL(k_face_eval_fast<1>, grid, dim3(TPB), 0, s, ph, nf, cells, flux);
  else hipLaunchKernelGGL(k_face_eval_fast<2>, grid, dim3(TPB), 0, s, ph, nf, cells, flux);
  return hipGetLastError();
}

}  // namespace cbns
