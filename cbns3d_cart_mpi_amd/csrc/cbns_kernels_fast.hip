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

// The faces that the method-level sweeps (cbns_rec_riemann) left to the
// list; in the pass over every face, only the face arrays in the mask dirs
// are checked.  (The fused steps list no faces: their x+y march and fused
// z pass leave failing faces' cells to the cell fixup.)
__global__ __launch_bounds__(TPB) void k_face_fixup(const BlockDesc* __restrict__ blocks, int nb, int which, Phys ph,
                                                    FixList fx, int dirs) {
  const unsigned n = *fx.count;  // every workgroup reads it before the last one resets it (below)
  const long long first = blockIdx.x * (long long)TPB + threadIdx.x, stride = (long long)gridDim.x * TPB;
  if (n <= fx.cap) {
    for (long long e = first; e < n; e += stride) {
      const unsigned long long v = fx.ent[e];
      const long long fi = (long long)(v & ((1ull << 40) - 1));
      const int D = (int)((v >> 40) & 3), b = (int)(v >> 42);
      if (D == 0) fixup_face<0>(blocks[b], which, fi, ph);
      else if (D == 1) fixup_face<1>(blocks[b], which, fi, ph);
      else fixup_face<2>(blocks[b], which, fi, ph);
    }
  } else {
    for (int b = 0; b < nb; b++) {
      const BlockDesc& B = blocks[b];
      if (dirs & 1) fixup_all<0>(B, which, ph, first, stride);
      if (dirs & 2) fixup_all<1>(B, which, ph, first, stride);
      if (dirs & 4) fixup_all<2>(B, which, ph, first, stride);
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
                             const FixList& fx, hipStream_t s, int dirs) {
  (void)hb;
  hipLaunchKernelGGL(k_face_fixup, dim3(FIX_WG), dim3(TPB), 0, s, db, nb, which, ph, fx, dirs);
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
