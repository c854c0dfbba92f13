// Persistent batch kernels: each launch walks a packed tile table that mixes
// the blocks of many solves.  The grid is the number of workgroups that can
// be resident at once (batch_grids).  Tiles are handed out dynamically, as
// the hardware dispatcher hands out the workgroups of a one-shot launch:
// workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2, so XCD r serves the r-th eighth of the table (as the one-shot kernels'
// xcd_tile does) in table order through its own counter, and a workgroup
// takes the next tile when it finishes one.  The tiles in flight on an XCD
// thus form one tight window that slides through its range, and the stencil
// neighbours of a tile (the rows and planes before it) are still in that
// XCD's L2.  (A static assignment, tile q + m G to workgroup q, lets slow
// workgroups fall behind and widens the window past the L2: the z sweep
// then fetched three times its data.)  The stage tail (update, ghost fill)
// is by default launched directly instead, one workgroup per tile, with the
// same per-XCD contiguous ranges (for_tiles with ctr == nullptr).  The
// per-item arithmetic is the shared bodies', so every solve's values equal a
// single-solve run's.
#include "cbns_batch.h"
#include "cbns_bodies.h"
#include "cbns_face_fast.h"
#include "cbns_zfuse.h"

namespace cbns {

// Runs body(x) for the tiles x this workgroup takes from its XCD's range of
// [0, n).  ctr: this launch's 8 counters, zero at launch (kb_reset_counters).
// The next tile's index is fetched (device-scope atomic) while the current
// one runs.  body must not synchronise the workgroup unless every thread
// calls it alike.
static_assert(BATCH_CTR_PER_LAUNCH == NXCD, "one tile counter per XCD");

template <class F>
__device__ __forceinline__ void for_tiles(int n, int* __restrict__ ctr, F&& body) {
  if (ctr == nullptr) {  // direct launch: one tile per workgroup, no queue
    // workgroups reach the XCDs in turn; XCD r takes the r-th contiguous run
    // of tiles, as the queues do (the grid is NXCD * per workgroups)
    const int per = (n + NXCD - 1) / NXCD;
    const int x = (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
    if (x < n) body(x);
    return;
  }
  const bool split = gridDim.x >= NXCD;  // every XCD has workgroups
  const int r = split ? (int)(blockIdx.x % NXCD) : 0;
  const int parts = split ? NXCD : 1;
  const int lo = (int)((long long)n * r / parts), hi = (int)((long long)n * (r + 1) / parts);
  __shared__ int s_k;
  int* c = ctr + r * BATCH_CTR_STRIDE;  // counters sit on separate lines: no contention between XCDs
  if (threadIdx.x == 0) s_k = atomicAdd(c, 1);
  __syncthreads();
  int k = s_k;
  while (lo + k < hi) {
    int nxt = 0;
    if (threadIdx.x == 0) nxt = atomicAdd(c, 1);
    body(lo + k);
    __syncthreads();  // all done with this tile (and with s_k)
    if (threadIdx.x == 0) s_k = nxt;
    __syncthreads();
    k = s_k;
  }
}

__global__ void kb_reset_counters(int* ctr, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctr[i] = 0;
}

template <int D, int P>
__global__ __launch_bounds__(TPB, 2) void kb_face(const BlockDesc* __restrict__ blocks,
                                                  const InstDesc* __restrict__ insts, TileTable tt, int which,
                                                  int jb, int* ctr) {
  for_tiles(tt.n, ctr, [&](int x) {
    const TileRef tr = tt.tiles[x];
    const BlockDesc& B = blocks[tr.block];
    const InstDesc& I = insts[B.inst];
    if (!I.stepping) return;
    const int t = tr.tile * TPB + threadIdx.x;
    if (D > 0) {  // y, z: pairs of faces (the table counts pairs)
      if (t >= face_pair_count(B, D)) return;
      if (P == 0) face_pair_body<D == 0 ? 1 : D>(B, t, which, jb, I.ph);
      else face_pair_body_f32<D == 0 ? 1 : D>(B, t, which, jb, I.ph, I.ref);
      return;
    }
    if (t >= face_count(B, D)) return;
    if (P == 0) face_body<D>(B, t, which, jb, I.ph);
    else face_body_f32<D>(B, t, which, jb, I.ph, I.ref);
  });
}

// Tolerance-bound solves' x and y faces: one march along y (xyrun_fast),
// one (column tile (i, k), y-run) unit of one block per table entry, the
// run length chosen from the block's size as the table was built.
__global__ __launch_bounds__(TPB, 2) void kb_xyface(const BlockDesc* __restrict__ blocks,
                                                    const InstDesc* __restrict__ insts, TileTable tt, int which,
                                                    int* ctr) {
  static_assert(TPB == ZX * ZY && ZX == 64, "one wave per row of columns");
  static_assert(BATCH_ZM <= XY_RUN_MAX, "run length of the x+y march");
  for_tiles(tt.n, ctr, [&](int x) {
    const TileRef tr = tt.tiles[x];
    const BlockDesc& B = blocks[tr.block];
    const InstDesc& I = insts[B.inst];
    if (!I.stepping) return;
    const int ntx = (B.n[0] + ZX - 1) / ZX, nt = ntx * ((B.n[2] + ZY - 1) / ZY);
    const int ym = chunk_len(nt, B.n[1], BATCH_ZM, ZM_LOW, BATCH_UNITS);
    const int tile = tr.tile % nt, run = tr.tile / nt;
    const int i0 = (tile % ntx) * ZX, k = (tile / ntx) * ZY + (int)threadIdx.x / ZX;
    const int j0 = run * ym;
    // uniform over each wave: all its lanes take part in the march's lane exchanges
    if (k < B.n[2] && j0 < B.n[1]) xyrun_fast(B, i0, k, j0, ym, which, I.ph);
  });
}

// FROMQ: the tolerance-bound solves' march, which forms u, v, w, T from
// the stage input state `which` (they store no primitives)
// float storage (FP32-flux) fits 3 waves per SIMD, as the single-solve march
template <typename VT, bool FROMQ = false>
__global__ __launch_bounds__(TPB, sizeof(VT) == 4 ? 3 : 2) void kb_visc(const BlockDesc* __restrict__ blocks,
                                               const InstDesc* __restrict__ insts, TileTable tt, int jb,
                                               int which, int* ctr) {
  // visc_tile_body synchronises the workgroup: every condition before it is
  // uniform over the workgroup (one tile, one block, one solve); for_tiles'
  // barrier after each unit also protects the LDS planes that the next unit
  // refills
  for_tiles(tt.n, ctr, [&](int x) {
    const TileRef tr = tt.tiles[x];
    const BlockDesc& B = blocks[tr.block];
    const InstDesc& I = insts[B.inst];
    if (!I.stepping) return;
    // one (column tile, z-chunk) unit of the viscous tile march per entry
    const int ntiles = ((B.n[0] + VX - 1) / VX) * ((B.n[1] + VY - 1) / VY);
    const int vz = chunk_len(ntiles, B.n[2], BATCH_VZ, VISC_LOW, BATCH_UNITS);
    visc_tile_body<VT, FROMQ>(B, I.ph, tr.tile % ntiles, tr.tile / ntiles, which, vz);
  });
}

// Flush of a workgroup's running spectral-radius max for solve inst (an
// integer atomic max on the bits of a non-negative double, per wave).
__device__ __forceinline__ void flush_smax(InstDesc* insts, int inst, double acc, int bad) {
  const double m = wave_max(acc);
  if ((threadIdx.x & 63) == 0) atomicMax(&insts[inst].smax, (unsigned long long)__double_as_longlong(m));
  if (bad) atomicOr(&insts[inst].nonfinite, 1);
}

// Stage tail: the z-face sweep fused with the update (zrun, cbns_zfuse.h),
// one (column tile, z-run) unit of one block per table entry, the run length
// chosen from the block's size as the table was built.  At stage 3 it also
// reduces the next step's calc_dt spectral radius of every updated cell
// into its solve's smax (as the single-solve stage tail does), so kb_dt
// skips such solves on the next step.  The descriptors are read through the
// read-only insts (so that their uniform fields come through the scalar
// cache) and only smax and nonfinite, which no kernel of the stage reads,
// are written through insts_rw.
template <int FORM>
__global__ __launch_bounds__(TPB, 2) void kb_zface_update(const BlockDesc* __restrict__ blocks,
                                                          const InstDesc* __restrict__ insts, InstDesc* insts_rw,
                                                          TileTable tt, int in, int out, int stage, int* ctr,
                                                          FixList cfx) {
  static_assert(TPB == ZX * ZY, "one column per thread");
  int cur = -1, bad = 0;
  double acc = 0.0;
  for_tiles(tt.n, ctr, [&](int x) {
    const TileRef tr = tt.tiles[x];
    const BlockDesc& B = blocks[tr.block];
    const InstDesc& I = insts[B.inst];
    if (!I.stepping) return;
    if (stage == 3 && B.inst != cur) {  // uniform: one block per tile
      if (cur >= 0) flush_smax(insts_rw, cur, acc, bad);
      cur = B.inst;
      acc = 0.0;
      bad = 0;
    }
    const int ntx = (B.n[0] + ZX - 1) / ZX, nt = ntx * ((B.n[1] + ZY - 1) / ZY);
    const int zm = chunk_len(nt, B.n[2], BATCH_ZM, ZM_LOW, BATCH_UNITS);
    const int tile = tr.tile % nt, run = tr.tile / nt;
    const int i = (tile % ntx) * ZX + (int)threadIdx.x % ZX, j = (tile / ntx) * ZY + (int)threadIdx.x / ZX;
    const int k0 = run * zm;
    if (i < B.n[0] && j < B.n[1] && k0 < B.n[2])
      zrun<FORM>(B, tr.block, i, j, k0, zm, in, out, stage, I.dt, I.ph, I.ref, stage == 3, &acc, &bad, cfx);
  });
  if (cur >= 0) flush_smax(insts_rw, cur, acc, bad);
}

// Cells of FP64 and tolerance-bound solves that the fused pass left
// (zcell_fix), with each solve's own physics; at stage 3 each such cell's
// spectral radius goes into its solve's smax.  More failures than the list
// holds: every cell of every stepping FP64 solve is checked.  The last
// workgroup resets the list.
__device__ __forceinline__ void zfix_one(const BlockDesc& B, const InstDesc& I, InstDesc* insts_rw, int i, int j,
                                         int k, int in, int out, int stage, bool all) {
  double s = 0.0;
  int bad = 0;
  if (I.precision == PREC_FP64_FAST)
    zcell_fix<ZF_FAST>(B, i, j, k, in, out, stage, I.dt, I.ph, all, stage == 3, &s, &bad);
  else
    zcell_fix<ZF_FP64>(B, i, j, k, in, out, stage, I.dt, I.ph, all, stage == 3, &s, &bad);
  if (stage == 3) {
    atomicMax(&insts_rw[B.inst].smax, (unsigned long long)__double_as_longlong(s));
    if (bad) atomicOr(&insts_rw[B.inst].nonfinite, 1);
  }
}

constexpr int ZFIX_WG = 64;  // workgroups of the cell fixup pass

__global__ __launch_bounds__(TPB) void kb_zcell_fixup(const BlockDesc* __restrict__ blocks,
                                                      const InstDesc* __restrict__ insts, InstDesc* insts_rw, int nb,
                                                      int in, int out, int stage, FixList cfx) {
  const unsigned n = *cfx.count;  // every workgroup reads it before the last one resets it (below)
  const long long first = blockIdx.x * (long long)TPB + threadIdx.x, stride = (long long)gridDim.x * TPB;
  if (n <= cfx.cap) {
    for (long long e = first; e < n; e += stride) {
      const unsigned long long v = cfx.ent[e];
      const long long ct = (long long)(v & ((1ull << 40) - 1));
      const BlockDesc& B = blocks[(int)(v >> 42)];
      const long long r = ct / B.n[0];
      zfix_one(B, insts[B.inst], insts_rw, (int)(ct % B.n[0]), (int)(r % B.n[1]), (int)(r / B.n[1]), in, out, stage,
               false);
    }
  } else {
    for (int b = 0; b < nb; b++) {
      const BlockDesc& B = blocks[b];
      const InstDesc& I = insts[B.inst];
      if (!I.stepping || I.precision == PREC_FP32_FLUX) continue;
      for (long long ct = first; ct < B.sn; ct += stride) {
        const long long r = ct / B.n[0];
        zfix_one(B, I, insts_rw, (int)(ct % B.n[0]), (int)(r % B.n[1]), (int)(r / B.n[1]), in, out, stage, true);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(cfx.done, 1u) == gridDim.x - 1) {
      atomicExch(cfx.count, 0u);
      atomicExch(cfx.done, 0u);
    }
  }
}

__global__ __launch_bounds__(TPB) void kb_ghost(const BlockDesc* __restrict__ blocks,
                                                const InstDesc* __restrict__ insts, TileTable tt, int which,
                                                int* ctr) {
  for_tiles(tt.n, ctr, [&](int x) {
    const TileRef tr = tt.tiles[x];
    const BlockDesc& B = blocks[tr.block];
    const InstDesc& I = insts[B.inst];
    if (!I.stepping) return;
    const int g = tr.tile * TPB + threadIdx.x;
    if (g >= B.sg - B.sn) return;
    // the tolerance-bound solves' ghosts take the state alone
    if (I.precision == PREC_FP32_FLUX) ghost_body<float>(B, blocks, I.dom, I.block0, ghost_cell(B, g), which, true);
    else ghost_body<double>(B, blocks, I.dom, I.block0, ghost_cell(B, g), which, I.precision != PREC_FP64_FAST);
  });
}

// Per-solve max of the spectral radius: a workgroup keeps a running max
// for the solve it is on and flushes it (integer atomic max on the bits of a
// non-negative double) when its next tile belongs to another solve, or at
// the end.
// When no live solve needs this pass (every stepping solve took its smax
// from the last stage-3 update), each workgroup leaves before touching the
// tile queue.
__global__ __launch_bounds__(TPB) void kb_dt(const BlockDesc* __restrict__ blocks,
                                             const InstDesc* __restrict__ insts, InstDesc* insts_rw, TileTable tt,
                                             int* ctr, int ninst) {
  bool any = false;
  for (int k = 0; k < ninst; k++) any = any || (insts[k].active && insts[k].budget > 0 && !insts[k].s_fused);
  if (!any) return;  // uniform: every thread reads the same flags
  int cur = -1;
  double acc = 0.0;
  int bad = 0;
  for_tiles(tt.n, ctr, [&](int x) {
    const TileRef tr = tt.tiles[x];
    const BlockDesc& B = blocks[tr.block];
    const int inst = B.inst;
    if (!insts[inst].active || insts[inst].budget <= 0 || insts[inst].s_fused) return;
    if (inst != cur) {
      if (cur >= 0) flush_smax(insts_rw, cur, acc, bad);
      cur = inst;
      acc = 0.0;
      bad = 0;
    }
    const int t = tr.tile * TPB + threadIdx.x;
    if (t < B.n[0] * B.n[1] * B.n[2]) {
      // FP32-flux and tolerance-bound solves: calc_primitive's values formed from the state
      const int pr = insts[inst].precision;
      const double s = pr == PREC_FP32_FLUX || pr == PREC_FP64_FAST ? dt_cell<float>(B, t, insts[inst].ph, &bad)
                                                                    : dt_cell<double>(B, t, insts[inst].ph, &bad);
      acc = (acc < s) ? s : acc;  // src/Block3d.cpp:546: a NaN never wins
    }
  });
  if (cur >= 0) flush_smax(insts_rw, cur, acc, bad);
}

// The driver's per-step bookkeeping on the device (src/cbns3d.cpp:178-194):
// is_finished, dt = CFL / s_max, the clip to t_end, t_cur += dt; a solve also
// stops when it has taken its requested steps or met a non-finite value.
__global__ void kb_finalize(InstDesc* insts, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  InstDesc& I = insts[i];
  I.stepping = 0;
  I.s_fused = 0;
  if (!I.active) return;
  const unsigned long long bits = I.smax;
  I.smax = 0ull;
  if (I.nonfinite || I.steps_done >= I.steps_target || I.t_cur + 1.0e-5 > I.t_end) {
    I.active = 0;
    I.dt = 0.0;
    return;
  }
  if (I.budget <= 0) return;  // paused until the host grants more steps
  const double s_max = __longlong_as_double((long long)bits);
  double dt = I.ph.CFL / s_max;
  if (I.t_end - I.t_cur < dt) dt = I.t_end - I.t_cur;
  I.t_cur += dt;
  I.dt = dt;
  I.steps_done += 1;
  I.budget -= 1;
  I.stepping = 1;
  I.s_fused = I.pairs;  // this step's stage 3 forms the next step's smax
}

__global__ void kb_publish(const InstDesc* insts, InstStatus* status, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const InstDesc& I = insts[i];
    InstStatus st;
    st.t_cur = I.t_cur;
    st.steps_done = I.steps_done;
    st.steps_target = I.steps_target;
    st.active = I.active;
    st.nonfinite = I.nonfinite;
    status[i] = st;
  }
  __threadfence_system();
}

// Step budgets granted by the host: pairs of (solve id, steps).
__global__ void kb_set_budget(InstDesc* insts, const int2* pairs, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) insts[pairs[k].x].budget = pairs[k].y;
}

hipError_t batch_launch_face(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int dir, int which, int jb,
                             int form, int wgs, int* ctr, hipStream_t s) {
  if (tt.n == 0) return hipSuccess;
  const dim3 g(wgs), b(TPB);
  if (form == 0) {
    if (dir == 0) hipLaunchKernelGGL((kb_face<0, 0>), g, b, 0, s, blocks, insts, tt, which, jb, ctr);
    else if (dir == 1) hipLaunchKernelGGL((kb_face<1, 0>), g, b, 0, s, blocks, insts, tt, which, jb, ctr);
    else hipLaunchKernelGGL((kb_face<2, 0>), g, b, 0, s, blocks, insts, tt, which, jb, ctr);
  } else {
    if (dir == 0) hipLaunchKernelGGL((kb_face<0, 1>), g, b, 0, s, blocks, insts, tt, which, jb, ctr);
    else if (dir == 1) hipLaunchKernelGGL((kb_face<1, 1>), g, b, 0, s, blocks, insts, tt, which, jb, ctr);
    else hipLaunchKernelGGL((kb_face<2, 1>), g, b, 0, s, blocks, insts, tt, which, jb, ctr);
  }
  return hipGetLastError();
}

hipError_t batch_launch_visc(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int jb, int precision,
                             int which, int wgs, int* ctr, hipStream_t s) {
  if (tt.n == 0) return hipSuccess;
  if (precision == PREC_FP32_FLUX)
    hipLaunchKernelGGL(kb_visc<float>, dim3(wgs), dim3(TPB), 0, s, blocks, insts, tt, jb, which, ctr);
  else if (precision == PREC_FP64_FAST)
    hipLaunchKernelGGL((kb_visc<double, true>), dim3(wgs), dim3(TPB), 0, s, blocks, insts, tt, jb, which, ctr);
  else
    hipLaunchKernelGGL(kb_visc<double>, dim3(wgs), dim3(TPB), 0, s, blocks, insts, tt, jb, which, ctr);
  return hipGetLastError();
}

hipError_t batch_launch_zupdate(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int precision, int in,
                                int out, int stage, int wgs, int* ctr, const FixList& cfx, hipStream_t s) {
  if (tt.n == 0) return hipSuccess;
  if (wgs == 0) {  // direct launch
    wgs = (tt.n + NXCD - 1) / NXCD * NXCD;
    ctr = nullptr;
  }
  const dim3 g(wgs), b(TPB);
  if (precision == PREC_FP64)
    hipLaunchKernelGGL(kb_zface_update<ZF_FP64>, g, b, 0, s, blocks, insts, insts, tt, in, out, stage, ctr, cfx);
  else if (precision == PREC_FP64_FAST)
    hipLaunchKernelGGL(kb_zface_update<ZF_FAST>, g, b, 0, s, blocks, insts, insts, tt, in, out, stage, ctr, cfx);
  else
    hipLaunchKernelGGL(kb_zface_update<ZF_FP32>, g, b, 0, s, blocks, insts, insts, tt, in, out, stage, ctr, cfx);
  return hipGetLastError();
}

hipError_t batch_launch_ymarch(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int which, int wgs, int* ctr,
                               hipStream_t s) {
  if (tt.n == 0) return hipSuccess;
  if (wgs == 0) {  // direct launch
    wgs = (tt.n + NXCD - 1) / NXCD * NXCD;
    ctr = nullptr;
  }
  hipLaunchKernelGGL(kb_xyface, dim3(wgs), dim3(TPB), 0, s, blocks, insts, tt, which, ctr);
  return hipGetLastError();
}

hipError_t batch_launch_zfixup(const BlockDesc* blocks, InstDesc* insts, int nb, int in, int out, int stage,
                               const FixList& cfx, hipStream_t s) {
  hipLaunchKernelGGL(kb_zcell_fixup, dim3(ZFIX_WG), dim3(TPB), 0, s, blocks, insts, insts, nb, in, out, stage, cfx);
  return hipGetLastError();
}

hipError_t batch_launch_ghost(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int which, int wgs,
                              int* ctr, hipStream_t s) {
  if (tt.n == 0) return hipSuccess;
  if (wgs == 0) {  // direct launch
    wgs = (tt.n + NXCD - 1) / NXCD * NXCD;
    ctr = nullptr;
  }
  hipLaunchKernelGGL(kb_ghost, dim3(wgs), dim3(TPB), 0, s, blocks, insts, tt, which, ctr);
  return hipGetLastError();
}

hipError_t batch_launch_dt(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int wgs, int* ctr, int ninst,
                           hipStream_t s) {
  if (tt.n == 0) return hipSuccess;
  hipLaunchKernelGGL(kb_dt, dim3(wgs), dim3(TPB), 0, s, blocks, insts, insts, tt, ctr, ninst);
  return hipGetLastError();
}

hipError_t batch_launch_reset_counters(int* ctr, int n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(kb_reset_counters, dim3((n + 255) / 256), dim3(256), 0, s, ctr, n);
  return hipGetLastError();
}

hipError_t batch_launch_set_budget(InstDesc* insts, const int2* pairs, int n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(kb_set_budget, dim3((n + 63) / 64), dim3(64), 0, s, insts, pairs, n);
  return hipGetLastError();
}

hipError_t batch_launch_publish(const InstDesc* insts, InstStatus* status, int ninst, hipStream_t s) {
  if (ninst == 0) return hipSuccess;
  hipLaunchKernelGGL(kb_publish, dim3((ninst + 63) / 64), dim3(64), 0, s, insts, status, ninst);
  return hipGetLastError();
}

hipError_t batch_launch_finalize(InstDesc* insts, int ninst, hipStream_t s) {
  if (ninst == 0) return hipSuccess;
  hipLaunchKernelGGL(kb_finalize, dim3((ninst + 63) / 64), dim3(64), 0, s, insts, ninst);
  return hipGetLastError();
}

// Resident workgroups of each persistent kernel on the current device
// (workgroups per CU at the kernel's register and LDS use, times the CUs).
hipError_t batch_grids(BatchGrids* g) {
  int dev = 0, cus = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  g->cus = cus;
  auto res = [&](const void* f, int* out) {
    int per = 0;
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, TPB, 0);
    *out = (per > 0 ? per : 1) * cus;
  };
  res((const void*)kb_face<0, 0>, &g->face[0][0]);
  res((const void*)kb_face<1, 0>, &g->face[0][1]);
  res((const void*)kb_face<2, 0>, &g->face[0][2]);
  res((const void*)kb_face<0, 1>, &g->face[1][0]);
  res((const void*)kb_face<1, 1>, &g->face[1][1]);
  res((const void*)kb_face<2, 1>, &g->face[1][2]);
  res((const void*)kb_visc<double>, &g->visc[0]);
  res((const void*)kb_visc<float>, &g->visc[1]);
  res((const void*)kb_visc<double, true>, &g->visc[2]);
  res((const void*)kb_zface_update<ZF_FP64>, &g->update[0]);
  res((const void*)kb_zface_update<ZF_FP32>, &g->update[1]);
  res((const void*)kb_zface_update<ZF_FAST>, &g->update[2]);
  res((const void*)kb_xyface, &g->ymarch);
  res((const void*)kb_ghost, &g->ghost);
  res((const void*)kb_dt, &g->dt);
  return e;
}

}  // namespace cbns
