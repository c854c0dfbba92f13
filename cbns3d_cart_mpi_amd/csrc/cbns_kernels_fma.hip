// Face-flux sweeps of the contracted FP64 form (CBNS_FP64_FMA).  This
// translation unit is compiled with -ffp-contract=fast: the compiler may fuse
// a * b + c into one FMA, and quotients use qdiv<true> (the IEEE sequence
// without range scaling).  Everything else of a step is shared with the plain
// path.  Results agree with the plain path to about one rounding per
// operation rather than bit for bit.
#include "cbns_bodies.h"

namespace cbns {

static inline unsigned nblocks_fma(long long count) { return (unsigned)((count + TPB - 1) / TPB); }

template <int D, int OCC>
__global__ __launch_bounds__(TPB, OCC) void k_face_flux_fma(const BlockDesc* __restrict__ blocks, int which, int jb,
                                                            Phys ph) {
  const Tile tl = xcd_tile();
  const BlockDesc& B = blocks[tl.y];
  const int t = tl.x * TPB + threadIdx.x;
  if (D == 0 || OCC >= 3) {  // the pair body does not fit 3 waves' registers
    if (t >= face_count(B, D)) return;
    face_body<D, true>(B, t, which, jb, ph);
  } else {
    if (t >= face_pair_count(B, D)) return;
    face_pair_body<D == 0 ? 1 : D, true>(B, t, which, jb, ph);
  }
}

hipError_t launch_face_flux_fma(const BlockDesc* db, const BlockDesc* hb, int nb, int dir, int which,
                                const Phys& ph, int jb, int occ, hipStream_t s) {
  long long mx = 0;
  for (int b = 0; b < nb; b++) {
    const long long c = face_items(hb[b], dir, occ < 3);
    mx = c > mx ? c : mx;
  }
  dim3 grid(nblocks_fma(mx), nb);
#define CBNS_FACE(DIR, OCC) \
  hipLaunchKernelGGL((k_face_flux_fma<DIR, OCC>), grid, dim3(TPB), 0, s, db, which, jb, ph)
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
  return hipGetLastError();
}

// Contracted-form face flux on explicit inputs (parity tests).
template <int D>
__global__ __launch_bounds__(TPB) void k_face_eval_fma(Phys ph, int nf, const double* __restrict__ cells,
                                                       const double* __restrict__ prim, double* __restrict__ flux) {
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= nf) return;
  face_eval_body<D, 1>(ph, t, cells, prim, flux);
}

hipError_t launch_face_eval_fma(int dir, const Phys& ph, int nf, const double* cells, const double* prim,
                                double* flux, hipStream_t s) {
  dim3 grid(nblocks_fma(nf));
  if (dir == 0) hipLaunchKernelGGL(k_face_eval_fma<0>, grid, dim3(TPB), 0, s, ph, nf, cells, prim, flux);
  else if (dir == 1) hipLaunchKernelGGL(k_face_eval_fma<1>, grid, dim3(TPB), 0, s, ph, nf, cells, prim, flux);
  else hipLaunchKernelGGL(k_face_eval_fma<2>, grid, dim3(TPB), 0, s, ph, nf, cells, prim, flux);
  return hipGetLastError();
}

}  // namespace cbns
