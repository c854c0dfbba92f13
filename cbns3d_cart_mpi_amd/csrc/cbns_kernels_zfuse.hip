// Single-solve kernels of the z-face sweep fused with the stage tail (the
// arithmetic and its layout: cbns_zfuse.h).
#include "cbns_zfuse.h"

namespace cbns {

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

template <int FORM>
__global__ __launch_bounds__(ZX * ZY, 2) void k_zface_update(const BlockDesc* __restrict__ blocks, int in, int out,
                                                          int stage, double dt, Phys ph, RefState ref,
                                                          unsigned long long* smax, int* flag, FixList cfx,
                                                          int zm) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[blockIdx.z];
  const int ntx = (B.n[0] + ZX - 1) / ZX;
  const int i = (tl.x % ntx) * ZX + (int)threadIdx.x % ZX;
  const int j = (tl.x / ntx) * ZY + (int)threadIdx.x / ZX;
  const int k0 = tl.y * zm;
  const bool want = smax != nullptr;  // uniform
  double s = 0.0;
  int bad = 0;
  if (i < B.n[0] && j < B.n[1] && k0 < B.n[2])
    zrun<FORM>(B, blockIdx.z, i, j, k0, zm, in, out, stage, dt, ph, ref, want, &s, &bad, cfx);
  if (want) reduce_smax(s, bad, smax, flag);
}

hipError_t launch_zface_update(const BlockDesc* db, const BlockDesc* hb, int nb, int precision, int in, int out,
                               int stage, double dt, const Phys& ph, const RefState& ref, unsigned long long* smax,
                               int* flag, const FixList& cfx, hipStream_t s, int zm_top) {
  int tiles = 0, n2 = 0;
  for (int b = 0; b < nb; b++) {
    tiles = std::max(tiles, ((hb[b].n[0] + ZX - 1) / ZX) * ((hb[b].n[1] + ZY - 1) / ZY));
    n2 = std::max(n2, hb[b].n[2]);
  }
  const int zm = chunk_len((long long)tiles * nb, n2, zm_top, ZM_LOW, 1024);
  const dim3 grid(tiles, (n2 + zm - 1) / zm, nb), blk(ZX * ZY);
  if (precision == PREC_FP64_FAST)
    hipLaunchKernelGGL(k_zface_update<ZF_FAST>, grid, blk, 0, s, db, in, out, stage, dt, ph, ref, smax, flag, cfx,
                       zm);
  else if (precision == PREC_FP32_FLUX)
    hipLaunchKernelGGL(k_zface_update<ZF_FP32>, grid, blk, 0, s, db, in, out, stage, dt, ph, ref, smax, flag, cfx,
                       zm);
  else if (precision == PREC_FP64)
    hipLaunchKernelGGL(k_zface_update<ZF_FP64>, grid, blk, 0, s, db, in, out, stage, dt, ph, ref, smax, flag, cfx,
                       zm);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// Tolerance-bound form: the x and y faces in one march (xyrun_fast), one
// wave per 64 columns at one k, ZY waves per workgroup, runs of ym <= 64
// cells along y.  Every condition on the way to the march is uniform over
// a wave (k from the wave index, j0 from the workgroup), so all 64 lanes
// take part in its lane exchanges.
__global__ __launch_bounds__(ZX * ZY, 2) void k_xyface(const BlockDesc* __restrict__ blocks, int which, Phys ph,
                                                       int ym) {
  static_assert(ZX == 64, "one wave per row of columns");
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[blockIdx.z];
  const int ntx = (B.n[0] + ZX - 1) / ZX;
  const int i0 = (tl.x % ntx) * ZX;
  const int k = (tl.x / ntx) * ZY + (int)threadIdx.x / ZX;
  const int j0 = tl.y * ym;
  if (i0 < B.n[0] && k < B.n[2] && j0 < B.n[1]) xyrun_fast(B, i0, k, j0, ym, which, ph);
}

hipError_t launch_xyface(const BlockDesc* db, const BlockDesc* hb, int nb, int which, const Phys& ph,
                         hipStream_t s, int ym_top) {
  int tiles = 0, n1 = 0;
  for (int b = 0; b < nb; b++) {
    tiles = std::max(tiles, ((hb[b].n[0] + ZX - 1) / ZX) * ((hb[b].n[2] + ZY - 1) / ZY));
    n1 = std::max(n1, hb[b].n[1]);
  }
  const int ym = chunk_len((long long)tiles * nb, n1, std::min(ym_top, XY_RUN_MAX), ZM_LOW, 1024);
  const dim3 grid(tiles, (n1 + ym - 1) / ym, nb), blk(ZX * ZY);
  hipLaunchKernelGGL(k_xyface, grid, blk, 0, s, db, which, ph, ym);
  return hipGetLastError();
}

constexpr int ZFIX_WG = 64;  // workgroups of the cell fixup pass

template <int FORM>
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
      zcell_fix<FORM>(B, i, (int)(r % B.n[1]), (int)(r / B.n[1]), in, out, stage, dt, ph, false, want, &s, &bad);
    }
  } else {
    for (int b = 0; b < nb; b++) {
      const BlockDesc& B = blocks[b];
      for (long long ct = first; ct < B.sn; ct += stride) {
        const int i = (int)(ct % B.n[0]);
        const long long r = ct / B.n[0];
        zcell_fix<FORM>(B, i, (int)(r % B.n[1]), (int)(r / B.n[1]), in, out, stage, dt, ph, true, want, &s, &bad);
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

hipError_t launch_zcell_fixup(const BlockDesc* db, const BlockDesc* hb, int nb, int precision, int in, int out,
                              int stage, double dt, const Phys& ph, unsigned long long* smax, int* flag,
                              const FixList& cfx, hipStream_t s) {
  (void)hb;
  if (precision == PREC_FP64_FAST)
    hipLaunchKernelGGL(k_zcell_fixup<ZF_FAST>, dim3(ZFIX_WG), dim3(TPB), 0, s, db, nb, in, out, stage, dt, ph, smax,
                       flag, cfx);
  else if (precision == PREC_FP64)
    hipLaunchKernelGGL(k_zcell_fixup<ZF_FP64>, dim3(ZFIX_WG), dim3(TPB), 0, s, db, nb, in, out, stage, dt, ph, smax,
                       flag, cfx);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace cbns
