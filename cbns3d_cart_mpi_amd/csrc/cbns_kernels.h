// Device-side data layout and kernel launchers shared by the kernels and the
// C ABI.  All state lives in HBM in SoA form (one array per conserved or
// primitive field, x fastest) so that every stencil load of a wavefront is
// one contiguous run along x.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cbns {

constexpr int NEQ = 5;
constexpr int NG = 4;   // ghost layers
constexpr int MAXB = 512;
// Most ghosted cells of one block: every field of a block then stays below
// 4 GB, so kernels address it at 32-bit byte offsets from the field's base.
// Larger grids run as several on-device blocks (the decomposition is free
// and bit-exact).
constexpr long long MAX_BLOCK_CELLS = 1LL << 29;
// whether a block of n interior cells per axis fits that bound
inline bool block_fits(int n0, int n1, int n2) {
  return (long long)(n0 + 2 * NG) * (n1 + 2 * NG) * (long long)(n2 + 2 * NG) <= MAX_BLOCK_CELLS;
}

// Physical constants of one solve (src/read_input.cpp:78-90 and the static
// caches of src/Block3d.cpp:563-564, 764-767).
struct Phys {
  double gamma, gam1, gM2, CFL, C_T_inf, Re_inv, kappa, C_dt_v;
  double dx, dy, dz;
  // Correctly rounded division by a run-time constant h in three operations
  // is valid when RN(1/h) is within 2^-54 relative of 1/h (checked exactly on
  // the host, cbns_create); rh = RN(1/h), and the flag says whether it holds.
  double rdx, rdy, rdz, rgam1;
  int fdx, fdy, fdz, fgam1;
  double gg1;  // gamma / (gamma - 1), for the tolerance-bound face form
};

// Division of an n < 2^31 by a run-time constant d >= 1 in three
// instructions: q = (mulhi(n, m) + n) >> l with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1 (the round-up method of Granlund and
// Montgomery, "Division by invariant integers using multiplication", PLDI
// 1994; the add does not overflow for n < 2^31).
struct FastDiv {
  unsigned m, l;
};
inline FastDiv make_fastdiv(unsigned d) {
  unsigned l = 0;
  while ((1ull << l) < d) l++;
  FastDiv f;
  f.m = (unsigned)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  f.l = l;
  return f;
}

// One on-device block: the reference's Block3d arrays, SoA.
struct BlockDesc {
  double* Q[3];   // Q, Q_p and (tolerance-bound single solve) the stage-1 state: [5][g2][g1][g0], ghosted
  double* F[3];   // face fluxes Ep, Fp, Gp: [5][n2+dz][n1+dy][n0+dx]
  double* dfv;    // viscous divergence [4][n2][n1][n0] (float storage in the FP32-flux form)
  double* prim;   // rho, u, v, w, p, T: [6][g2][g1][g0] (float storage in the FP32-flux form)
  double* mu;     // Sutherland viscosity on the interior [n2][n1][n0]
  int n[3];       // interior cells
  int g[3];       // ghosted extent n + 8
  int o[3];       // global origin
  int inst;       // owning solve (batched form); 0 for a single solve
  long long sg;   // ghosted field stride
  long long sf[3];// face-array field strides
  long long sn;   // interior field stride
  // the face sweeps' work-index divisors: row lengths n0 + 1 (x faces) and
  // n0 (y and z face pairs), plane counts n2 (x faces, y pairs) and
  // (n2 + 2) / 2 (z pairs); set by set_face_divisors
  FastDiv dv_row[2], dv_plane[2];
};

inline void set_face_divisors(BlockDesc& B) {
  B.dv_row[0] = make_fastdiv(B.n[0] + 1);
  B.dv_row[1] = make_fastdiv(B.n[0]);
  B.dv_plane[0] = make_fastdiv(B.n[2]);
  B.dv_plane[1] = make_fastdiv((B.n[2] + 2) / 2);
}

// Partition of the periodic domain (src/cbns3d.cpp:102-126).
struct Domain {
  int N[3];     // global cells
  int dims[3];  // blocks per axis
  int base[3];  // N / dims
  int nb;
};

// Descriptor pointers are generic; these views tell the compiler that they
// address global memory, so accesses compile to global_* instead of flat_*
// instructions (no LDS aperture check, SGPR-base addressing).
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> gmem(T* p) {
  return (gptr<T>)p;
}

struct Launch {
  hipStream_t stream;
  const BlockDesc* dblocks;  // device copy
  int nb;
  int maxcount;              // per-kernel max work per block (set by launcher)
};

// Arithmetic of the face fluxes (the values of cbns_precision in the C ABI).
constexpr int PREC_FP64 = 0, PREC_FP32_FLUX = 1, PREC_FP64_FMA = 2, PREC_FP64_FAST = 3;

// Work items of a face sweep of block b: faces along x, and along y and z
// pairs of faces (face_pair_body) or, with pairs = false, single faces.
inline long long face_items(const BlockDesc& b, int dir, bool pairs = true) {
  const long long ni = b.n[0] + (dir == 0), nj = b.n[1] + (dir == 1), nk = b.n[2] + (dir == 2);
  if (!pairs || dir == 0) return ni * nj * nk;
  return dir == 1 ? ni * ((nj + 1) / 2) * nk : ni * nj * ((nk + 1) / 2);
}

// Faces of the tolerance-bound form whose range checks fail: the sweep
// appends them here (block, axis, face index) and launch_face_fixup
// recomputes them with IEEE division after the sweeps, so that the sweep
// kernels carry no fallback code (and its registers).  count and done start
// at zero; the fixup pass resets them.  More than cap failures send the fixup
// over every face of every sweep instead.
struct FixList {
  unsigned* count;
  unsigned* done;
  unsigned long long* ent;  // face index | axis << 40 | block << 42
  unsigned cap;
};
constexpr unsigned FIX_CAP = 1u << 16;
constexpr size_t FIX_BYTES = 256 + sizeof(unsigned long long) * FIX_CAP;  // counters on their own line, then entries

// kernel launchers (cbns_kernels.hip; launch_face_flux_fma in cbns_kernels_fma.hip,
// launch_face_flux_fast and launch_face_fixup in cbns_kernels_fast.hip)
hipError_t launch_face_flux_fast(const BlockDesc* db, const BlockDesc* hb, int nb, int dir, int which,
                                 const Phys& ph, int jb, int occ, const FixList& fx, hipStream_t s);
// Recompute the faces listed in fx (of the sweeps of state `which` since the
// last fixup) with the reference-order form; resets fx.  dirs: bit mask of
// the face arrays the pass over every face (more failures than fx holds)
// checks.
hipError_t launch_face_fixup(const BlockDesc* db, const BlockDesc* hb, int nb, int which, const Phys& ph,
                             const FixList& fx, hipStream_t s, int dirs = 7);
FixList fixlist_at(void* buf);
// Single solve (cbns_kernels_zfuse.hip): the z-face sweep fused with the
// stage tail, for PREC_FP64, PREC_FP32_FLUX and PREC_FP64_FAST.  Stage
// `stage` reads its input state Q[in] and writes Q[out]; Q[0] is Q^n (stages
// 2 and 3 read it pointwise).  FP64 forms: cells beside a z-face whose range
// checks fail are not updated but appended to cfx (block, cell) for
// launch_zcell_fixup, which recomputes their faces with the reference-order
// form and updates them.  With smax (stage 3), the calc_dt
// spectral radius of every updated cell is reduced into *smax (*flag set for
// a non-finite one).
struct RefState;
hipError_t launch_zface_update(const BlockDesc* db, const BlockDesc* hb, int nb, int precision, int in, int out,
                               int stage, double dt, const Phys& ph, const RefState& ref, unsigned long long* smax,
                               int* flag, const FixList& cfx, hipStream_t s, int zm_top = 32);
// Tolerance-bound form: its x and y faces in one march along y that stores
// each interior cell's x- plus y-flux differences into F[1] (no x sweep, no
// face fixup); cells beside a failing x or y face are marked for the fused
// z pass to leave to the cell fixup.  Runs are at most 64 cells long.
hipError_t launch_xyface(const BlockDesc* db, const BlockDesc* hb, int nb, int which, const Phys& ph,
                         hipStream_t s, int ym_top = 32);
hipError_t launch_zcell_fixup(const BlockDesc* db, const BlockDesc* hb, int nb, int precision, int in, int out,
                              int stage, double dt, const Phys& ph, unsigned long long* smax, int* flag,
                              const FixList& cfx, hipStream_t s);
hipError_t launch_face_eval_fast(int dir, const Phys& ph, int nfaces, const double* cells, double* flux,
                                 hipStream_t s);
hipError_t launch_face_flux_fma(const BlockDesc* db, const BlockDesc* hb, int nb, int dir, int which,
                                const Phys& ph, int jb, int occ, hipStream_t s);
struct RefState;
hipError_t launch_face_flux(const BlockDesc* db, const BlockDesc* hb, int nb, int dir, int which,
                            const Phys& ph, int precision, const RefState& g, int jb, int occ, const FixList& fx,
                            hipStream_t s);
hipError_t launch_primitive(const BlockDesc* db, const BlockDesc* hb, int nb, int which, const Phys& ph,
                            int precision, hipStream_t s);
hipError_t launch_ghost_fill(const BlockDesc* db, const BlockDesc* hb, const Domain& dom, int which,
                             hipStream_t s);
// viscous fields stored as float for PREC_FP32_FLUX, double otherwise
hipError_t launch_viscous(const BlockDesc* db, const BlockDesc* hb, int nb, const Phys& ph, int precision, int jb,
                          hipStream_t s);
// wgs > 0: a fixed number of workgroups taking the tile units in turn;
// from_state >= 0 (FP64 only): u, v, w, T formed from Q[from_state] instead
// of read from the stored primitives; vz_top: most planes per z-chunk
hipError_t launch_visc_flux(const BlockDesc* db, const BlockDesc* hb, int nb, const Phys& ph, int precision,
                            int jb, hipStream_t s, int wgs = 0, int from_state = -1, int vz_top = 32);
// smax/flag (optional): also reduce the next step's calc_dt spectral radius of
// the updated cells (only over x-pairs, for blocks of even n0; *s_done says
// whether it was done).
hipError_t launch_update_fused(const BlockDesc* db, const BlockDesc* hb, int nb, double dt, int stage,
                               const Phys& ph, int precision, int jb, hipStream_t s,
                               unsigned long long* smax = nullptr, int* flag = nullptr, bool* s_done = nullptr);
hipError_t launch_ghost_fill_prim(const BlockDesc* db, const BlockDesc* hb, const Domain& dom, int which,
                                  int precision, hipStream_t s);
hipError_t launch_update(const BlockDesc* db, const BlockDesc* hb, int nb, double dt, int stage,
                         const Phys& ph, int precision, int jb, hipStream_t s);
hipError_t launch_dt_reduce(const BlockDesc* db, const BlockDesc* hb, int nb, const Phys& ph,
                            unsigned long long* d_smax, int* d_flag, int precision, hipStream_t s);
hipError_t launch_copy_state(const BlockDesc* db, const BlockDesc* hb, int nb, int from, int to, hipStream_t s);
hipError_t launch_tgv(const BlockDesc* db, const BlockDesc* hb, int nb, const double* tab, int tabstride,
                      const Phys& ph, double p_inf, hipStream_t s);
// interior AoS (global) <-> block SoA
hipError_t launch_scatter_aos(const BlockDesc* db, const BlockDesc* hb, const Domain& dom, int which,
                              const double* aos, hipStream_t s);
hipError_t launch_gather_aos(const BlockDesc* db, const BlockDesc* hb, const Domain& dom, int which,
                             double* aos, hipStream_t s);
hipError_t launch_fill_mu(const BlockDesc* db, const BlockDesc* hb, int nb, double mu0, hipStream_t s);
// Viscous tile march: VX x VY interior columns per workgroup (VX * VY
// threads), at most VZ planes per z-chunk (fewer where a grid has too few
// column tiles to fill the device: chunk_len).
constexpr int VX = 16, VY = 16, VZ = 32;  // VZ: the launchers' default vz_top
// Planes per chunk of a march over `tiles` column tiles (of all blocks it
// covers) and n2 planes: top, halved (down to low) until there are at least
// `want` (tile, chunk) units.  The single solve asks 1024 units of its
// whole grid; the batched form 256 of each block (the other solves fill the
// device), from the same block sizes on the host and the device.
__host__ __device__ inline int chunk_len(long long tiles, int n2, int top, int low, long long want) {
  int v = top;
  while (v > low && tiles * ((n2 + v - 1) / v) < want) v /= 2;
  return v;
}
constexpr int VISC_LOW = 4;    // fewest planes per viscous chunk
// Fused z-face and stage-tail pass (cbns_zfuse.h): ZX x ZY columns per
// workgroup, one per thread, each with a run of at most ZM and at least
// ZM_LOW cells along z (chunk_len).
constexpr int ZX = 64, ZY = 4, ZM = 32, ZM_LOW = 2;  // ZM: the launchers' default run length
constexpr int BATCH_UNITS = 64;  // units per block the batched marches ask for
// most viscous planes per chunk and cells per run of the batched marches
constexpr int BATCH_VZ = 32, BATCH_ZM = 32;
constexpr int GSUM_WG = 512;  // workgroups of the fixed-order global sums
hipError_t launch_global_sums(const BlockDesc* db, const Domain& dom, const Phys& ph, int what, double* d_part,
                              double* d_out, int precision, hipStream_t s);
hipError_t launch_face_eval(int dir, const Phys& ph, int precision, int nfaces, const double* cells,
                            const double* prim, double* flux, hipStream_t s);
hipError_t launch_face_eval_fma(int dir, const Phys& ph, int nfaces, const double* cells, const double* prim,
                                double* flux, hipStream_t s);
// RN(1/h) and whether three-operation division by h is exact (see Phys).
void reciprocal_check(double h, double* rh, int* fast);

// Shell spectra of the velocity field (cbns_spectra.hip): buffers and FFT
// plan per context, made at the first call.  spectra_run writes the kinetic
// energy, enstrophy and dilatation spectra of shells 0 .. nshell - 1 to
// host_out[0..3 nshell) and synchronises s; 0 ok, 1 allocation failed, 2 HIP
// or hipFFT error, 3 a domain of more than 2^31 - 1 cells.
struct Spectra;
Spectra* spectra_create();
void spectra_destroy(Spectra* sp);
int spectra_run(Spectra* sp, const BlockDesc* dblocks, const BlockDesc* hblocks, const Domain& dom, int which,
                int nshell, double* host_out, hipStream_t s);

}  // namespace cbns
