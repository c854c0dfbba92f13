// HIP kernels of the plain FP64 right-hand-side and time-step path (gfx950).
//
// Work mapping: grid.y selects the on-device block, grid.x * 256 threads walk
// a flat index over that block's work items with x fastest, so each wavefront
// touches 64 consecutive cells of one x-row (coalesced SoA loads).  Stencil
// neighbours along y and z are the same rows shifted by g0 or g0*g1, served
// from L1/L2.
#include <algorithm>

#include "cbns_bodies.h"

namespace cbns {

static inline unsigned nblocks_for(long long count) { return (unsigned)((count + TPB - 1) / TPB); }




// ---------------------------------------------------------------------------
// Inviscid face fluxes: rec_riemann_x/y/z (src/Block3d.cpp:555-748).
// One thread per face; face (i,j,k) along D sits between cells i-1 and i.
// OCC = minimum workgroups per CU the register allocation must allow: 2 keeps
// every value in registers (182 VGPRs), 3 trades a few spills for a third
// wave per SIMD.  Both are compiled; the launch picks one (autotuned).
template <int D, int OCC>
__global__ __launch_bounds__(TPB, OCC) void k_face_flux(const BlockDesc* __restrict__ blocks, int which, int jb,
                                                        Phys ph) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[tl.y];
  const int t = tl.x * TPB + threadIdx.x;
  if (D == 0 || OCC >= 3) {  // the pair body does not fit 3 waves' registers
    if (t >= face_count(B, D)) return;
    face_body<D>(B, t, which, jb, ph);
  } else {
    if (t >= face_pair_count(B, D)) return;
    face_pair_body<D == 0 ? 1 : D>(B, t, which, jb, ph);
  }
}

// FP32-flux form: same stencil and work items, FP32 face arrays holding
// F - F(Ug).
template <int D>
__global__ __launch_bounds__(TPB) void k_face_flux_f32(const BlockDesc* __restrict__ blocks, int which, int jb,
                                                       Phys ph, RefState g) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[tl.y];
  const int t = tl.x * TPB + threadIdx.x;
  if (D == 0) {
    if (t >= face_count(B, D)) return;
    face_body_f32<D>(B, t, which, jb, ph, g);
  } else {
    if (t >= face_pair_count(B, D)) return;
    face_pair_body_f32<D == 0 ? 1 : D>(B, t, which, jb, ph, g);
  }
}

hipError_t launch_face_flux(const BlockDesc* db, const BlockDesc* hb, int nb, int dir, int which,
                            const Phys& ph, int precision, const RefState& g, int jb, int occ, const FixList& fx,
                            hipStream_t s) {
  if (precision == PREC_FP64_FMA) return launch_face_flux_fma(db, hb, nb, dir, which, ph, jb, occ, s);
  if (precision == PREC_FP64_FAST) return launch_face_flux_fast(db, hb, nb, dir, which, ph, jb, occ, fx, s);
  // work items: pairs of faces along y and z, except the FP64 sweeps at 3
  // waves per SIMD, which take single faces
  const bool single = precision == PREC_FP64 && occ >= 3;
  long long mx = 0;
  for (int b = 0; b < nb; b++) mx = std::max(mx, face_items(hb[b], dir, !single));
  dim3 grid(nblocks_for(mx), nb);
  if (precision == PREC_FP64) {
#define CBNS_FACE(DIR, OCC) \
  hipLaunchKernelGGL((k_face_flux<DIR, OCC>), grid, dim3(TPB), 0, s, db, which, jb, ph)
    if (occ >= 3) {
      if (dir == 0) CBNS_FACE(0, 3);
      else if (dir == 1) CBNS_FACE(1, 3);
      else CBNS_FACE(2, 3);
    } else {
      if (dir == 0) CBNS_FACE(0, 2);
      else if (dir == 1) CBNS_FACE(1, 2);
      else CBNS_FACE(2, 2);
    }
#undef CBNS_FACE
  } else {
    if (dir == 0) hipLaunchKernelGGL(k_face_flux_f32<0>, grid, dim3(TPB), 0, s, db, which, jb, ph, g);
    else if (dir == 1) hipLaunchKernelGGL(k_face_flux_f32<1>, grid, dim3(TPB), 0, s, db, which, jb, ph, g);
    else hipLaunchKernelGGL(k_face_flux_f32<2>, grid, dim3(TPB), 0, s, db, which, jb, ph, g);
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// calc_primitive over the whole ghosted box (src/Block3d.cpp:359-385).
// PT: primitive storage (float in the FP32-flux form), computed in FP64.
template <typename PT>
__global__ __launch_bounds__(TPB) void k_primitive(const BlockDesc* __restrict__ blocks, int which, double gam1,
                                                   double gM2) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[tl.y];
  const long long sg = B.sg;
  const int t = tl.x * TPB + threadIdx.x;
  if (t >= sg) return;
  gptr<const double> __restrict__ Q = gmem<const double>(B.Q[which]);
  gptr<PT> __restrict__ P = gmem((PT*)B.prim);
  double rr = Q[t];
  double uu = Q[sg + t] / rr;
  double vv = Q[2 * sg + t] / rr;
  double ww = Q[3 * sg + t] / rr;
  double pp = gam1 * (Q[4 * sg + t] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
  P[t] = (PT)rr;
  P[sg + t] = (PT)uu;
  P[2 * sg + t] = (PT)vv;
  P[3 * sg + t] = (PT)ww;
  P[4 * sg + t] = (PT)pp;
  P[5 * sg + t] = (PT)(gM2 * pp / rr);
}

hipError_t launch_primitive(const BlockDesc* db, const BlockDesc* hb, int nb, int which, const Phys& ph,
                            int precision, hipStream_t s) {
  long long mx = 0;
  for (int b = 0; b < nb; b++) mx = hb[b].sg > mx ? hb[b].sg : mx;
  const dim3 grid(nblocks_for(mx), nb);
  if (precision == PREC_FP32_FLUX)
    hipLaunchKernelGGL(k_primitive<float>, grid, dim3(TPB), 0, s, db, which, ph.gam1, ph.gM2);
  else hipLaunchKernelGGL(k_primitive<double>, grid, dim3(TPB), 0, s, db, which, ph.gam1, ph.gM2);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Ghost fill: boundary_condition (src/boundary_condition.cpp:4-231) without
// MPI.  The reference's three directional exchange phases leave every ghost
// cell equal to the interior cell of its periodic image; here each ghost cell
// gathers that value directly from the block that owns it.
__global__ __launch_bounds__(TPB) void k_ghost_fill(const BlockDesc* __restrict__ blocks, Domain dom, int which) {
  const BlockDesc& B = blocks[blockIdx.y];
  const int g = blockIdx.x * TPB + threadIdx.x;
  if (g >= B.sg - B.sn) return;
  ghost_body(B, blocks, dom, 0, ghost_cell(B, g), which, false);
}

hipError_t launch_ghost_fill(const BlockDesc* db, const BlockDesc* hb, const Domain& dom, int which,
                             hipStream_t s) {
  long long mx = 0;
  for (int b = 0; b < dom.nb; b++) mx = std::max(mx, hb[b].sg - hb[b].sn);
  hipLaunchKernelGGL(k_ghost_fill, dim3(nblocks_for(mx), dom.nb), dim3(TPB), 0, s, db, dom, which);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Viscous pass.  k_visc_flux: calc_gradient (src/calc_gradient.cpp:17-66),
// calc_viscous_terms (src/Block3d.cpp:750-810), calc_viscous_flux
// (src/Block3d.cpp:812-868) and the divergence of
// calc_viscous_flux_contribution (src/Block3d.cpp:879-916) in one tile
// march, writing the divergence (B.dfv) and mu, which is kept on the
// interior for calc_dt (the reference stores it everywhere; only the
// interior is read).  Grid: (column tiles, z-chunks, blocks); see
// visc_tile_body.
template <typename VT, bool FROMQ = false>
__global__ __launch_bounds__(VX * VY) void k_visc_flux(const BlockDesc* __restrict__ blocks, Phys ph, int which,
                                                       int vz) {
  const Tile tl = xcd_tile();
  visc_tile_body<VT, FROMQ>(blocks[blockIdx.z], ph, tl.x, tl.y, which, vz);
}

static int visc_chunk_planes(const BlockDesc* hb, int nb, int top) {
  long long tiles = 0;
  int n2 = 0;
  for (int b = 0; b < nb; b++) {
    tiles = std::max(tiles, (long long)((hb[b].n[0] + VX - 1) / VX) * ((hb[b].n[1] + VY - 1) / VY));
    n2 = std::max(n2, hb[b].n[2]);
  }
  return chunk_len(tiles * nb, n2, top, VISC_LOW, 1024);
}

static dim3 visc_grid(const BlockDesc* hb, int nb, int vz) {
  int tiles = 0, chunks = 0;
  for (int b = 0; b < nb; b++) {
    const int n0 = hb[b].n[0], n1 = hb[b].n[1], n2 = hb[b].n[2];
    tiles = std::max(tiles, ((n0 + VX - 1) / VX) * ((n1 + VY - 1) / VY));
    chunks = std::max(chunks, (n2 + vz - 1) / vz);
  }
  return dim3(tiles, chunks, nb);
}

// The same march with a fixed number of workgroups, each taking units
// (block, column tile, z-chunk) in turn: beside the face sweeps on a second
// stream it then occupies only that many workgroup slots.
template <typename VT, bool FROMQ = false>
__global__ __launch_bounds__(VX * VY) void k_visc_flux_units(const BlockDesc* __restrict__ blocks, Phys ph,
                                                             int tiles, int chunks, int nb, int which, int vz) {
  const int per_block = tiles * chunks, total = per_block * nb;
  for (int u = blockIdx.x; u < total; u += gridDim.x) {  // uniform per workgroup
    const int b = u / per_block, r = u - b * per_block;
    visc_tile_body<VT, FROMQ>(blocks[b], ph, r % tiles, r / tiles, which, vz);
    __syncthreads();  // the next unit refills the LDS planes this one read
  }
}

hipError_t launch_visc_flux(const BlockDesc* db, const BlockDesc* hb, int nb, const Phys& ph, int precision,
                            int jb, hipStream_t s, int wgs, int from_state, int vz_top) {
  (void)jb;  // the tile march fixes its own order
  const int vz = visc_chunk_planes(hb, nb, vz_top);
  const dim3 grid = visc_grid(hb, nb, vz);
  const bool fromq = from_state >= 0;  // u, v, w, T formed from Q[from_state]
  const int w = fromq ? from_state : 0;
  if (wgs > 0) {
    if (precision == PREC_FP32_FLUX)
      hipLaunchKernelGGL(k_visc_flux_units<float>, dim3(wgs), dim3(VX * VY), 0, s, db, ph, (int)grid.x, (int)grid.y,
                         nb, w, vz);
    else if (fromq)
      hipLaunchKernelGGL((k_visc_flux_units<double, true>), dim3(wgs), dim3(VX * VY), 0, s, db, ph, (int)grid.x,
                         (int)grid.y, nb, w, vz);
    else
      hipLaunchKernelGGL(k_visc_flux_units<double>, dim3(wgs), dim3(VX * VY), 0, s, db, ph, (int)grid.x, (int)grid.y,
                         nb, w, vz);
    return hipGetLastError();
  }
  if (precision == PREC_FP32_FLUX)
    hipLaunchKernelGGL(k_visc_flux<float>, grid, dim3(VX * VY), 0, s, db, ph, w, vz);
  else if (fromq)
    hipLaunchKernelGGL((k_visc_flux<double, true>), grid, dim3(VX * VY), 0, s, db, ph, w, vz);
  else
    hipLaunchKernelGGL(k_visc_flux<double>, grid, dim3(VX * VY), 0, s, db, ph, w, vz);
  return hipGetLastError();
}

hipError_t launch_viscous(const BlockDesc* db, const BlockDesc* hb, int nb, const Phys& ph, int precision, int jb,
                          hipStream_t s) {
  return launch_visc_flux(db, hb, nb, ph, precision, jb, s);
}

// ---------------------------------------------------------------------------
// update_rk3 (src/Block3d.cpp:920-962) on the interior.
template <typename FT>
__global__ __launch_bounds__(TPB) void k_update(const BlockDesc* __restrict__ blocks, double dt, int stage,
                                                Phys ph, int jb) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[tl.y];
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int t = tl.x * TPB + threadIdx.x;
  if (t >= n0 * n1 * n2) return;
  int i, j, k;
  band_decode(t, n0, n1, n2, jb, i, j, k);
  const int ct = i + n0 * (j + n1 * k);
  const int c = gidx(B, i, j, k);
  const int iE = i + (n0 + 1) * (j + n1 * k);
  const int iF = i + n0 * (j + (n1 + 1) * k);
  const int iG = ct;
  gptr<double> __restrict__ Q = gmem(B.Q[0]);
  gptr<double> __restrict__ Qp = gmem(B.Q[1]);
  const long long sg = B.sg;
#pragma unroll
  for (int e = 0; e < NEQ; e++) {
    gptr<const FT> __restrict__ Ep = gmem<const FT>((const FT*)B.F[0] + e * B.sf[0]);
    gptr<const FT> __restrict__ Fp = gmem<const FT>((const FT*)B.F[1] + e * B.sf[1]);
    gptr<const FT> __restrict__ Gp = gmem<const FT>((const FT*)B.F[2] + e * B.sf[2]);
    double df = divc((double)Ep[iE + 1] - (double)Ep[iE], ph.dx, ph.rdx, ph.fdx);
    df += divc((double)Fp[iF + n0] - (double)Fp[iF], ph.dy, ph.rdy, ph.fdy);
    df += divc((double)Gp[iG + n0 * n1] - (double)Gp[iG], ph.dz, ph.rdz, ph.fdz);
    if (e > 0) df -= (double)gmem<const FT>((const FT*)B.dfv)[(e - 1) * B.sn + ct];
    const long long q = e * sg + c;
    if (1 == stage) {
      Qp[q] = Q[q] - dt * df;
    } else if (2 == stage) {
      Qp[q] = 0.75 * Q[q] + 0.25 * (Qp[q] - dt * df);
    } else if (3 == stage) {
      Q[q] = (Q[q] + 2.0 * (Qp[q] - dt * df)) / 3.0;
    }
  }
}

hipError_t launch_update(const BlockDesc* db, const BlockDesc* hb, int nb, double dt, int stage,
                         const Phys& ph, int precision, int jb, hipStream_t s) {
  long long mn = 0;
  for (int b = 0; b < nb; b++) mn = hb[b].sn > mn ? hb[b].sn : mn;
  if (precision != PREC_FP32_FLUX)
    hipLaunchKernelGGL(k_update<double>, dim3(nblocks_for(mn), nb), dim3(TPB), 0, s, db, dt, stage, ph, jb);
  else
    hipLaunchKernelGGL(k_update<float>, dim3(nblocks_for(mn), nb), dim3(TPB), 0, s, db, dt, stage, ph, jb);
  return hipGetLastError();
}



// ---------------------------------------------------------------------------
// Fused stage tail of the time-step path: the viscous divergence
// (src/Block3d.cpp:879-916), the RK3 update (src/Block3d.cpp:920-962) and the
// primitive recovery of the updated interior cell (src/Block3d.cpp:359-385),
// one pass per interior cell.  Each value is formed by the same expressions
// as the separate kernels, so the results are identical; the viscous
// divergence and the primitives never round-trip through HBM.
template <typename FT>
__global__ __launch_bounds__(TPB) void k_update_fused(const BlockDesc* __restrict__ blocks, double dt, int stage,
                                                      Phys ph, int jb) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[tl.y];
  const int t = tl.x * TPB + threadIdx.x;
  if (t >= B.n[0] * B.n[1] * B.n[2]) return;
  update_body<FT>(B, t, dt, stage, ph, jb);
}

// The same over x-pairs of cells (update_pair_body), for blocks with even n0.
// With smax != nullptr it also reduces the next step's calc_dt spectral
// radius of the updated cells into *smax (as k_dt does: integer atomic max on
// the bits, one per workgroup) and raises *flag for a non-finite one.
template <typename FT>
__global__ __launch_bounds__(TPB) void k_update_pair(const BlockDesc* __restrict__ blocks, double dt, int stage,
                                                     Phys ph, int jb, unsigned long long* smax, int* flag) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[tl.y];
  const int t = tl.x * TPB + threadIdx.x;
  const bool want = smax != nullptr;  // uniform
  double s = 0.0;
  int bad = 0;
  if (t < (B.n[0] / 2) * B.n[1] * B.n[2]) update_pair_body<FT>(B, t, dt, stage, ph, jb, want, &s, &bad);
  if (!want) return;
  s = wave_max(s);
  __shared__ double red[TPB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int a = 1; a < TPB / 64; a++) m = red[a] > m ? red[a] : m;
    atomicMax(smax, (unsigned long long)__double_as_longlong(m));
  }
  if (bad) atomicOr(flag, 1);
}

hipError_t launch_update_fused(const BlockDesc* db, const BlockDesc* hb, int nb, double dt, int stage,
                               const Phys& ph, int precision, int jb, hipStream_t s, unsigned long long* smax,
                               int* flag, bool* s_done) {
  if (s_done) *s_done = false;
  long long mn = 0;
  bool even = true;
  for (int b = 0; b < nb; b++) {
    mn = hb[b].sn > mn ? hb[b].sn : mn;
    even = even && hb[b].n[0] % 2 == 0;
  }
  if (stage < 1 || stage > 3) return hipSuccess;
  if (even) {
    const dim3 grid(nblocks_for(mn / 2), nb);
    if (precision != PREC_FP32_FLUX)
      hipLaunchKernelGGL(k_update_pair<double>, grid, dim3(TPB), 0, s, db, dt, stage, ph, jb, smax, flag);
    else
      hipLaunchKernelGGL(k_update_pair<float>, grid, dim3(TPB), 0, s, db, dt, stage, ph, jb, smax, flag);
    if (s_done) *s_done = smax != nullptr;
    return hipGetLastError();
  }
  const dim3 grid(nblocks_for(mn), nb);
  if (precision != PREC_FP32_FLUX) hipLaunchKernelGGL(k_update_fused<double>, grid, dim3(TPB), 0, s, db, dt, stage, ph, jb);
  else hipLaunchKernelGGL(k_update_fused<float>, grid, dim3(TPB), 0, s, db, dt, stage, ph, jb);
  return hipGetLastError();
}

// Ghost fill of the state and of its primitives together: the primitives of
// a ghost cell are those of its periodic image (calc_primitive is pointwise).
template <typename PT>
__global__ __launch_bounds__(TPB) void k_ghost_fill_prim(const BlockDesc* __restrict__ blocks, Domain dom, int which) {
  const BlockDesc& B = blocks[blockIdx.y];
  const int g = blockIdx.x * TPB + threadIdx.x;
  if (g >= B.sg - B.sn) return;
  ghost_body<PT>(B, blocks, dom, 0, ghost_cell(B, g), which, true);
}

hipError_t launch_ghost_fill_prim(const BlockDesc* db, const BlockDesc* hb, const Domain& dom, int which,
                                  int precision, hipStream_t s) {
  long long mx = 0;
  for (int b = 0; b < dom.nb; b++) mx = std::max(mx, hb[b].sg - hb[b].sn);
  const dim3 grid(nblocks_for(mx), dom.nb);
  if (precision == PREC_FP32_FLUX) hipLaunchKernelGGL(k_ghost_fill_prim<float>, grid, dim3(TPB), 0, s, db, dom, which);
  else hipLaunchKernelGGL(k_ghost_fill_prim<double>, grid, dim3(TPB), 0, s, db, dom, which);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// calc_dt spectral radius (src/Block3d.cpp:515-553), max over every block.
// The max is exact and order-independent, so one global max followed by one
// division equals the reference's per-rank CFL/s_max and MIN-allreduce.
// Non-negative doubles order like their bit patterns, so the cross-workgroup
// max is an integer atomic max.  A NaN never wins (the reference's ternary
// skips it) but raises the non-finite flag, as does +inf.

template <typename PT>
__global__ __launch_bounds__(TPB) void k_dt(const BlockDesc* __restrict__ blocks, Phys ph,
                                            unsigned long long* smax, int* flag) {
  const BlockDesc& B = blocks[blockIdx.y];
  const int total = B.n[0] * B.n[1] * B.n[2];
  double s_max = 0.0;
  int bad = 0;
  for (int t = blockIdx.x * TPB + threadIdx.x; t < total; t += gridDim.x * TPB) {
    const double s = dt_cell<PT>(B, t, ph, &bad);
    s_max = (s_max < s) ? s : s_max;  // src/Block3d.cpp:546: a NaN never wins
  }
  s_max = wave_max(s_max);
  __shared__ double red[TPB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = s_max;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int a = 1; a < TPB / 64; a++) m = red[a] > m ? red[a] : m;
    atomicMax(smax, (unsigned long long)__double_as_longlong(m));
  }
  if (bad) atomicOr(flag, 1);
}

hipError_t launch_dt_reduce(const BlockDesc* db, const BlockDesc* hb, int nb, const Phys& ph,
                            unsigned long long* d_smax, int* d_flag, int precision, hipStream_t s) {
  long long mn = 0;
  for (int b = 0; b < nb; b++) mn = hb[b].sn > mn ? hb[b].sn : mn;
  unsigned gx = nblocks_for(mn);
  gx = gx > 1024 ? 1024 : gx;
  // the FP32-flux form (float primitives) and the tolerance-bound form (no
  // stored primitives) form the FP64 primitives from the state (dt_cell<float>)
  if (precision == PREC_FP32_FLUX || precision == PREC_FP64_FAST)
    hipLaunchKernelGGL(k_dt<float>, dim3(gx, nb), dim3(TPB), 0, s, db, ph, d_smax, d_flag);
  else hipLaunchKernelGGL(k_dt<double>, dim3(gx, nb), dim3(TPB), 0, s, db, ph, d_smax, d_flag);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// init_Q_p and other whole-array copies between Q and Q_p (src/Block3d.cpp:387-394).
__global__ __launch_bounds__(TPB) void k_copy_state(const BlockDesc* __restrict__ blocks, int from, int to) {
  const BlockDesc& B = blocks[blockIdx.y];
  const long long n = NEQ * B.sg;
  gptr<const double> __restrict__ a = gmem<const double>(B.Q[from]);
  gptr<double> __restrict__ b = gmem(B.Q[to]);
  for (long long t = blockIdx.x * (long long)TPB + threadIdx.x; t < n; t += (long long)gridDim.x * TPB) b[t] = a[t];
}

hipError_t launch_copy_state(const BlockDesc* db, const BlockDesc* hb, int nb, int from, int to, hipStream_t s) {
  long long mx = 0;
  for (int b = 0; b < nb; b++) mx = NEQ * hb[b].sg > mx ? NEQ * hb[b].sg : mx;
  unsigned gx = nblocks_for(mx);
  gx = gx > 4096 ? 4096 : gx;
  hipLaunchKernelGGL(k_copy_state, dim3(gx, nb), dim3(TPB), 0, s, db, from, to);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Taylor-Green start: initial_condition + calc_conservative
// (src/Block3d.cpp:237-300) from per-axis tables of sin(x_c), cos(x_c),
// cos(2 x_c) evaluated on the host with the C library.
__global__ __launch_bounds__(TPB) void k_tgv(const BlockDesc* __restrict__ blocks, const double* __restrict__ tab,
                                             int ts, double gM2, double gam1, double p_inf) {
  const BlockDesc& B = blocks[blockIdx.y];
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= n0 * n1 * n2) return;
  const int i = t % n0;
  const int r = t / n0;
  const int j = r % n1;
  const int k = r / n1;
  const int gi = B.o[0] + i, gj = B.o[1] + j, gk = B.o[2] + k;
  // tab[(3*a + f) * ts + index], f: 0 sin, 1 cos, 2 cos(2x)
  const double sx = tab[0 * ts + gi], cx = tab[1 * ts + gi], c2x = tab[2 * ts + gi];
  const double sy = tab[3 * ts + gj], cy = tab[4 * ts + gj], c2y = tab[5 * ts + gj];
  const double cz = tab[7 * ts + gk], c2z = tab[8 * ts + gk];
  const double uu = sx * cy * cz;
  const double vv = -cx * sy * cz;
  const double ww = 0.0;
  const double pp = p_inf + (c2x + c2y) * (c2z + 2.0) / 16.0;
  const double rr = gM2 * pp;
  const int c = gidx(B, i, j, k);
  double* __restrict__ Q = B.Q[0];
  const long long sg = B.sg;
  Q[c] = rr;
  Q[sg + c] = rr * uu;
  Q[2 * sg + c] = rr * vv;
  Q[3 * sg + c] = rr * ww;
  Q[4 * sg + c] = pp / gam1 + 0.5 * rr * (uu * uu + vv * vv + ww * ww);
  B.mu[t] = 1.0;
}

hipError_t launch_tgv(const BlockDesc* db, const BlockDesc* hb, int nb, const double* tab, int ts,
                      const Phys& ph, double p_inf, hipStream_t s) {
  long long mn = 0;
  for (int b = 0; b < nb; b++) mn = hb[b].sn > mn ? hb[b].sn : mn;
  hipLaunchKernelGGL(k_tgv, dim3(nblocks_for(mn), nb), dim3(TPB), 0, s, db, tab, ts, ph.gM2, ph.gam1, p_inf);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Global interior AoS [k][j][i][eq] <-> per-block SoA.
__global__ __launch_bounds__(TPB) void k_aos(const BlockDesc* __restrict__ blocks, Domain dom, int which,
                                             double* __restrict__ aos, int to_blocks) {
  const BlockDesc& B = blocks[blockIdx.y];
  const int n0 = B.n[0], n1 = B.n[1], n2 = B.n[2];
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= n0 * n1 * n2) return;
  const int i = t % n0;
  const int r = t / n0;
  const int j = r % n1;
  const int k = r / n1;
  const long long gidx_ = (long long)NEQ *
      ((long long)(B.o[0] + i) + (long long)dom.N[0] * ((long long)(B.o[1] + j) + (long long)dom.N[1] * (B.o[2] + k)));
  const int c = gidx(B, i, j, k);
  double* __restrict__ Q = B.Q[which];
#pragma unroll
  for (int e = 0; e < NEQ; e++) {
    if (to_blocks) Q[e * B.sg + c] = aos[gidx_ + e];
    else aos[gidx_ + e] = Q[e * B.sg + c];
  }
}

hipError_t launch_scatter_aos(const BlockDesc* db, const BlockDesc* hb, const Domain& dom, int which,
                              const double* aos, hipStream_t s) {
  long long mn = 0;
  for (int b = 0; b < dom.nb; b++) mn = hb[b].sn > mn ? hb[b].sn : mn;
  hipLaunchKernelGGL(k_aos, dim3(nblocks_for(mn), dom.nb), dim3(TPB), 0, s, db, dom, which, (double*)aos, 1);
  return hipGetLastError();
}

hipError_t launch_gather_aos(const BlockDesc* db, const BlockDesc* hb, const Domain& dom, int which, double* aos,
                             hipStream_t s) {
  long long mn = 0;
  for (int b = 0; b < dom.nb; b++) mn = hb[b].sn > mn ? hb[b].sn : mn;
  hipLaunchKernelGGL(k_aos, dim3(nblocks_for(mn), dom.nb), dim3(TPB), 0, s, db, dom, which, aos, 0);
  return hipGetLastError();
}

__global__ __launch_bounds__(TPB) void k_fill_mu(const BlockDesc* __restrict__ blocks, double mu0) {
  const BlockDesc& B = blocks[blockIdx.y];
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t < B.sn) B.mu[t] = mu0;
}

hipError_t launch_fill_mu(const BlockDesc* db, const BlockDesc* hb, int nb, double mu0, hipStream_t s) {
  long long mn = 0;
  for (int b = 0; b < nb; b++) mn = hb[b].sn > mn ? hb[b].sn : mn;
  hipLaunchKernelGGL(k_fill_mu, dim3(nblocks_for(mn), nb), dim3(TPB), 0, s, db, mu0);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Global sums over the periodic interior in a fixed global cell order, so the
// result is bitwise independent of the block decomposition: thread t of a
// fixed grid of GSUM_WG workgroups visits global cells t, t + GSUM_WG*TPB, ...
// (each read from the block that owns it), then fixed-order wave, workgroup
// and final sums.
//   what = 0: (sum rho, sum E)           -- FP32-flux reference state
//   what = 1: (sum 0.5 rho |u|^2, sum 0.5 rho |omega|^2) with omega by the
//             4th-order central stencil   -- diagnostics
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int WHAT, typename PT = double>
__global__ __launch_bounds__(TPB) void k_global_sums(const BlockDesc* __restrict__ blocks, Domain dom, Phys ph,
                                                     double* part) {
  const long long total = (long long)dom.N[0] * dom.N[1] * dom.N[2];
  double sa = 0.0, sb = 0.0;
  for (long long t = blockIdx.x * (long long)TPB + threadIdx.x; t < total; t += (long long)gridDim.x * TPB) {
    const int gi = (int)(t % dom.N[0]);
    const long long r = t / dom.N[0];
    const int gj = (int)(r % dom.N[1]);
    const int gk = (int)(r / dom.N[1]);
    const int g3[3] = {gi, gj, gk};
    int bc[3], lc[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      int bb = g3[a] / dom.base[a];
      bb = bb > dom.dims[a] - 1 ? dom.dims[a] - 1 : bb;
      bc[a] = bb;
      lc[a] = g3[a] - bb * dom.base[a];
    }
    const BlockDesc& B = blocks[(bc[0] * dom.dims[1] + bc[1]) * dom.dims[2] + bc[2]];
    const int c = gidx(B, lc[0], lc[1], lc[2]);
    if (WHAT == 0) {
      sa += gmem<const double>(B.Q[0])[c];
      sb += gmem<const double>(B.Q[0])[4 * B.sg + c];
    } else {
      const long long sg = B.sg;
      const int sy = B.g[0], sz = B.g[0] * B.g[1];
      const PT* pb = (const PT*)B.prim;
      gptr<const PT> u = gmem<const PT>(pb + sg);
      gptr<const PT> v = gmem<const PT>(pb + 2 * sg);
      gptr<const PT> w = gmem<const PT>(pb + 3 * sg);
      const double rr = gmem<const PT>(pb)[c], uu = u[c], vv = v[c], ww = w[c];
      sa += 0.5 * rr * (uu * uu + vv * vv + ww * ww);
      const double ox = grad4(w, c, sy, ph.dy, ph.rdy, ph.fdy) - grad4(v, c, sz, ph.dz, ph.rdz, ph.fdz);
      const double oy = grad4(u, c, sz, ph.dz, ph.rdz, ph.fdz) - grad4(w, c, 1, ph.dx, ph.rdx, ph.fdx);
      const double oz = grad4(v, c, 1, ph.dx, ph.rdx, ph.fdx) - grad4(u, c, sy, ph.dy, ph.rdy, ph.fdy);
      sb += 0.5 * rr * (ox * ox + oy * oy + oz * oz);
    }
  }
  sa = wave_sum(sa);
  sb = wave_sum(sb);
  __shared__ double red[2][TPB / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = sa;
    red[1][wv] = sb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int x = 0; x < TPB / 64; x++) {
      a += red[0][x];
      b += red[1][x];
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

__global__ void k_sums_final(const double* part, int nparts, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int p = 0; p < nparts; p++) {
      a += part[2 * p];
      b += part[2 * p + 1];
    }
    out[0] = a;
    out[1] = b;
  }
}

hipError_t launch_global_sums(const BlockDesc* db, const Domain& dom, const Phys& ph, int what, double* d_part,
                              double* d_out, int precision, hipStream_t s) {
  if (what == 0) hipLaunchKernelGGL(k_global_sums<0>, dim3(GSUM_WG), dim3(TPB), 0, s, db, dom, ph, d_part);
  else if (precision == PREC_FP32_FLUX)
    hipLaunchKernelGGL((k_global_sums<1, float>), dim3(GSUM_WG), dim3(TPB), 0, s, db, dom, ph, d_part);
  else hipLaunchKernelGGL((k_global_sums<1, double>), dim3(GSUM_WG), dim3(TPB), 0, s, db, dom, ph, d_part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(64), 0, s, d_part, GSUM_WG, d_out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Face flux on explicit inputs (parity tests of the per-face arithmetic).
template <int D, bool F32>
__global__ __launch_bounds__(TPB) void k_face_eval(Phys ph, int nf, const double* __restrict__ cells,
                                                   const double* __restrict__ prim, double* __restrict__ flux) {
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= nf) return;
  face_eval_body<D, F32 ? 2 : 0>(ph, t, cells, prim, flux);
}

hipError_t launch_face_eval(int dir, const Phys& ph, int precision, int nf, const double* cells, const double* prim,
                            double* flux, hipStream_t s) {
  if (precision == PREC_FP64_FMA) return launch_face_eval_fma(dir, ph, nf, cells, prim, flux, s);
  if (precision == PREC_FP64_FAST) return launch_face_eval_fast(dir, ph, nf, cells, flux, s);
  dim3 grid(nblocks_for(nf));
#define CBNS_EVAL(DIR, F32) hipLaunchKernelGGL((k_face_eval<DIR, F32>), grid, dim3(TPB), 0, s, ph, nf, cells, prim, flux)
  if (precision == PREC_FP32_FLUX) {
    if (dir == 0) CBNS_EVAL(0, true);
    else if (dir == 1) CBNS_EVAL(1, true);
    else CBNS_EVAL(2, true);
  } else {
    if (dir == 0) CBNS_EVAL(0, false);
    else if (dir == 1) CBNS_EVAL(1, false);
    else CBNS_EVAL(2, false);
  }
#undef CBNS_EVAL
  return hipGetLastError();
}

}  // namespace cbns
