// C ABI of the MI355X-native cbns3d right-hand-side and time-step path.
// Owns all device memory of a context; host arrays are caller-owned and
// cross the boundary only through the explicit set/get calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/cbns3d_amd.h"
#include "cbns_face.h"
#include "cbns_kernels.h"

using namespace cbns;

struct cbns_ctx {
  cbns_params par;
  int precision;
  int device;
  Domain dom;
  Phys ph;
  std::vector<BlockDesc> hb;
  BlockDesc* db = nullptr;
  hipStream_t stream = nullptr;
  // second stream: the viscous kernels of a stage run beside its face kernels
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool overlap = true;
  int visc_wgs = 0;  // with overlap: workgroups of the side-stream viscous kernel (0 = one per unit)
  int jband = 16;  // rows per band of the stencil kernels' work order
  int face_occ = 2;  // register-allocation variant of the FP64 face kernels (not of the x+y march)
  int visc_top = VZ;  // most planes per viscous z-chunk
  int run_top = ZM;   // most cells per run of the fused z pass and the x+y march (which caps it at 64)
  unsigned long long* d_smax = nullptr;
  int* d_flag = nullptr;
  double* d_diag = nullptr;  // partials then the two sums
  int diag_parts = 0;
  std::vector<void*> allocs;
  uint64_t bytes = 0;
  bool have_state = false;
  int nonfinite = 0;
  // d_smax/d_flag hold the calc_dt reduction of the current state, made by the
  // last stage's update; cleared by every call that changes the state or mu
  bool s_ready = false;
  // The last update also copies that reduction to pinned host memory (h_dt)
  // and records ev_dt after the copy, so that the next step can issue its
  // first face sweeps, which do not depend on dt, before it waits for dt.
  struct DtWords {
    unsigned long long smax;
    int flag;
  };
  DtWords* h_dt = nullptr;
  hipEvent_t ev_dt = nullptr;
  bool dt_inflight = false;
  RefState ref{1.0, 0.0, 0.0};  // FP32-flux reference state
  FixList fix{};                  // faces of the tolerance-bound form left to the fixup pass
  FixList cfix{};                 // cells its fused z pass left to the cell fixup pass
  // the tolerance-bound form's fused steps store no primitives (its kernels
  // form them from the state): set after such a step, cleared by
  // refresh_prims before any call that reads the stored primitives
  bool prim_stale = false;
  // snapshot for file output: Q and the primitives rho u v w p of every
  // block, copied on the stream into snap_dev, then to pinned snap_host on
  // the copy stream while later work runs (ev_snap_host marks the end)
  hipStream_t copy = nullptr;
  hipEvent_t ev_snap_dev = nullptr, ev_snap_host = nullptr;
  char* snap_dev = nullptr;
  char* snap_host = nullptr;
  size_t snap_bytes = 0;
  std::vector<size_t> snap_off;  // per block: byte offset of Q; its primitives follow
  bool snap_pending = false;
  double snap_t = 0.0;
  Spectra* spec = nullptr;  // velocity spectra: FFT plan and buffers (first cbns_spectra call)
  // live kernel timing
  unsigned prof = 0;  // bit k: time the launches of kernel class k
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  double k_ms[CBNS_KCLASSES] = {0};
  int64_t k_n[CBNS_KCLASSES] = {0};
};

static hipEvent_t take_event(cbns_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Add the durations of completed bracketed launches (called after a stream sync).
static void harvest(cbns_ctx* c) {
  for (auto& p : c->pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
      c->k_ms[p.first] += ms;
      c->k_n[p.first] += 1;
    }
    c->ev_pool.push_back(p.second.first);
    c->ev_pool.push_back(p.second.second);
  }
  c->pending.clear();
}

// Launch `body` on stream s bracketed by events of class k when timing is on.
template <class F>
static hipError_t timed_on(cbns_ctx* c, hipStream_t s, int k, F body) {
  if (!(c->prof >> k & 1u)) return body();
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) return hipErrorOutOfMemory;
  hipError_t e = hipEventRecord(a, s);
  if (e == hipSuccess) e = body();
  if (e == hipSuccess) e = hipEventRecord(b, s);
  c->pending.push_back({k, {a, b}});
  return e;
}
template <class F>
static hipError_t timed(cbns_ctx* c, int k, F body) {
  return timed_on(c, c->stream, k, body);
}

#define CK(call)                                   \
  do {                                             \
    hipError_t e_ = (call);                        \
    if (e_ != hipSuccess) return CBNS_ERR_HIP;     \
  } while (0)

// y = RN(1/h); the three-operation quotient x*y corrected by one FMA
// remainder is correctly rounded for every x when |y h - 1| <= 2^-54
// (Markstein's condition: q0 = RN(x y) is then within one ulp of x/h).
// fma(y, h, -1) is exact here (|y h - 1| < 2^-52 needs < 53 bits).
void cbns::reciprocal_check(double h, double* rh, int* fast) {
  const double y = 1.0 / h;
  *rh = y;
  *fast = (std::isfinite(y) && std::fabs(std::fma(y, h, -1.0)) <= 0x1p-54) ? 1 : 0;
}

static int dev_alloc(cbns_ctx* c, void** p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) return CBNS_ERR_NOMEM;
  c->allocs.push_back(*p);
  c->bytes += bytes;
  return CBNS_OK;
}

// Recompute the stored primitives of Q where the fused steps of the
// tolerance-bound form left them stale (calc_primitive over the ghosted
// blocks, src/Block3d.cpp:359-385).
static int refresh_prims(cbns_ctx* c) {
  if (!c->prim_stale) return CBNS_OK;
  CK(launch_primitive(c->db, c->hb.data(), c->dom.nb, CBNS_Q, c->ph, c->precision, c->stream));
  c->prim_stale = false;
  return CBNS_OK;
}

extern "C" {

const char* cbns_strerror(int s) {
  switch (s) {
    case CBNS_OK: return "ok";
    case CBNS_ERR_ARG: return "invalid argument";
    case CBNS_ERR_HIP: return "HIP runtime error";
    case CBNS_ERR_NOMEM: return "device allocation failed";
    case CBNS_ERR_NODEV: return "no gfx950 device";
    case CBNS_ERR_STATE: return "call out of order";
    case CBNS_ERR_IO: return "file I/O failed";
    case CBNS_ERR_FORMAT: return "checkpoint dimensions do not match";
    case CBNS_ERR_NONFINITE: return "non-finite value in the time-step reduction";
    case CBNS_ERR_CAPACITY: return "request exceeds a fixed capacity (no-allocation scratch)";
    default: return "unknown status";
  }
}

// src/read_input.cpp:78-90
int cbns_params_derive(cbns_params* p) {
  if (!p) return CBNS_ERR_ARG;
  p->Mach2 = p->Mach * p->Mach;
  p->Re_inv = 1.0 / p->Re;
  p->Pr_inv = 1.0 / p->Pr;
  p->gam1 = p->gamma - 1.0;
  p->gM2 = p->gamma * p->Mach2;
  p->gam1_inv = 1.0 / p->gam1;
  p->p_inf = 1.0 / p->gM2;
  p->C_T_inf = 110.5 / p->T_inf;
  const double a = 4.0 / 3.0, b = p->gamma * p->Pr_inv;
  p->C_dt_v = ((a < b) ? b : a) * p->Re_inv;
  return CBNS_OK;
}

int cbns_device_count(int* n) {
  if (!n) return CBNS_ERR_ARG;
  *n = 0;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess) return CBNS_OK;
  for (int d = 0; d < cnt; d++) {
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, d) == hipSuccess && std::strncmp(pr.gcnArchName, "gfx950", 6) == 0) (*n)++;
  }
  return CBNS_OK;
}

int cbns_create(const cbns_params* p, const int dims[3], int precision, cbns_ctx** out) {
  if (!p || !dims || !out) return CBNS_ERR_ARG;
  *out = nullptr;
  if ((precision < CBNS_FP64 || precision > CBNS_FP64_FAST)) return CBNS_ERR_ARG;
  const int N[3] = {(int)p->IM, (int)p->JM, (int)p->KM};
  int largest[3];
  for (int a = 0; a < 3; a++) {
    if (dims[a] < 1 || N[a] < 1) return CBNS_ERR_ARG;
    const int base = N[a] / dims[a];
    const int last = N[a] - (dims[a] - 1) * base;
    if (base < NG || last < NG) return CBNS_ERR_ARG;  // halo must fit in one neighbour
    largest[a] = last;
  }
  if (!block_fits(largest[0], largest[1], largest[2])) return CBNS_ERR_ARG;  // split into more blocks
  const int nb = dims[0] * dims[1] * dims[2];
  if (nb > MAXB) return CBNS_ERR_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return CBNS_ERR_NODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return CBNS_ERR_NODEV;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return CBNS_ERR_NODEV;

  cbns_ctx* c = new (std::nothrow) cbns_ctx();
  if (!c) return CBNS_ERR_NOMEM;
  c->par = *p;
  c->precision = precision;
  c->device = dev;
  for (int a = 0; a < 3; a++) {
    c->dom.N[a] = N[a];
    c->dom.dims[a] = dims[a];
    c->dom.base[a] = N[a] / dims[a];
  }
  c->dom.nb = nb;
  // static caches of the reference (src/Block3d.cpp:208-219, 563-564, 764-767)
  const double two_pi = 2.0 * 3.14159265358979323846;
  c->ph.gamma = p->gamma;
  c->ph.gam1 = p->gam1;
  c->ph.gM2 = p->gM2;
  c->ph.CFL = p->CFL;
  c->ph.C_T_inf = p->C_T_inf;
  c->ph.Re_inv = p->Re_inv;
  c->ph.kappa = -p->Pr_inv * p->gam1_inv / p->Mach2;
  c->ph.C_dt_v = p->C_dt_v;
  c->ph.dx = two_pi / p->IM;
  c->ph.dy = two_pi / p->JM;
  c->ph.dz = two_pi / p->KM;
  reciprocal_check(c->ph.dx, &c->ph.rdx, &c->ph.fdx);
  reciprocal_check(c->ph.dy, &c->ph.rdy, &c->ph.fdy);
  reciprocal_check(c->ph.dz, &c->ph.rdz, &c->ph.fdz);
  reciprocal_check(c->ph.gam1, &c->ph.rgam1, &c->ph.fgam1);
  c->ph.gg1 = c->ph.gamma / c->ph.gam1;

  int st = CBNS_OK;
  c->hb.resize(nb);
  for (int bx = 0; bx < dims[0] && st == CBNS_OK; bx++)
    for (int by = 0; by < dims[1] && st == CBNS_OK; by++)
      for (int bz = 0; bz < dims[2] && st == CBNS_OK; bz++) {
        const int b = (bx * dims[1] + by) * dims[2] + bz;
        BlockDesc& B = c->hb[b];
        std::memset(&B, 0, sizeof(B));
        const int bc[3] = {bx, by, bz};
        for (int a = 0; a < 3; a++) {
          const int base = N[a] / dims[a];
          B.n[a] = (bc[a] == dims[a] - 1) ? N[a] - (dims[a] - 1) * base : base;
          B.g[a] = B.n[a] + 2 * NG;
          B.o[a] = bc[a] * base;
        }
        B.sg = (long long)B.g[0] * B.g[1] * B.g[2];
        B.sn = (long long)B.n[0] * B.n[1] * B.n[2];
        for (int d = 0; d < 3; d++)
          B.sf[d] = (long long)(B.n[0] + (d == 0)) * (B.n[1] + (d == 1)) * (B.n[2] + (d == 2));
        set_face_divisors(B);
        void* q;
        if ((st = dev_alloc(c, &q, sizeof(double) * NEQ * B.sg))) break;
        B.Q[0] = (double*)q;
        if ((st = dev_alloc(c, &q, sizeof(double) * NEQ * B.sg))) break;
        B.Q[1] = (double*)q;
        if (precision != CBNS_FP64_FMA) {  // the stage-1 state of the fused z pass's steps
          if ((st = dev_alloc(c, &q, sizeof(double) * NEQ * B.sg))) break;
          B.Q[2] = (double*)q;
        }
        for (int d = 0; d < 3 && !st; d++) {
          if ((st = dev_alloc(c, &q, sizeof(double) * NEQ * B.sf[d]))) break;
          B.F[d] = (double*)q;
        }
        if (st) break;
        if ((st = dev_alloc(c, &q, sizeof(double) * 4 * B.sn))) break;
        B.dfv = (double*)q;
        // primitives: float in the FP32-flux form, double otherwise
        const size_t vbytes = precision == CBNS_FP32_FLUX ? sizeof(float) : sizeof(double);
        if ((st = dev_alloc(c, &q, vbytes * 6 * B.sg))) break;
        B.prim = (double*)q;
        if ((st = dev_alloc(c, &q, sizeof(double) * B.sn))) break;
        B.mu = (double*)q;
      }
  void* q = nullptr;
  if (!st) st = dev_alloc(c, &q, sizeof(BlockDesc) * nb);
  c->db = (BlockDesc*)q;
  if (!st) st = dev_alloc(c, &q, 64);
  c->d_smax = (unsigned long long*)q;
  if (!st) st = dev_alloc(c, &q, 64);
  c->d_flag = (int*)q;
  c->diag_parts = GSUM_WG;
  if (!st) st = dev_alloc(c, &q, sizeof(double) * (2 * c->diag_parts + 2));
  c->d_diag = (double*)q;
  if (!st) st = dev_alloc(c, &q, FIX_BYTES);
  if (!st && hipMemset(q, 0, 256) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st) c->fix = fixlist_at(q);
  if (!st) st = dev_alloc(c, &q, FIX_BYTES);
  if (!st && hipMemset(q, 0, 256) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st) c->cfix = fixlist_at(q);
  if (!st && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st && hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st && hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st && hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st && hipEventCreateWithFlags(&c->ev_dt, hipEventDisableTiming) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st && hipHostMalloc((void**)&c->h_dt, sizeof(cbns_ctx::DtWords), hipHostMallocDefault) != hipSuccess)
    st = CBNS_ERR_NOMEM;
  if (!st && hipMemcpy(c->db, c->hb.data(), sizeof(BlockDesc) * nb, hipMemcpyHostToDevice) != hipSuccess)
    st = CBNS_ERR_HIP;
  if (st) {
    cbns_destroy(c);
    return st;
  }
  *out = c;
  return CBNS_OK;
}

int cbns_destroy(cbns_ctx* c) {
  if (!c) return CBNS_ERR_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->copy) {
    (void)hipStreamSynchronize(c->copy);
    (void)hipStreamDestroy(c->copy);
  }
  if (c->ev_snap_dev) (void)hipEventDestroy(c->ev_snap_dev);
  if (c->ev_snap_host) (void)hipEventDestroy(c->ev_snap_host);
  if (c->snap_dev) (void)hipFree(c->snap_dev);
  spectra_destroy(c->spec);
  if (c->snap_host) (void)hipHostFree(c->snap_host);
  harvest(c);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  if (c->side) (void)hipStreamSynchronize(c->side);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_dt) (void)hipEventDestroy(c->ev_dt);
  if (c->h_dt) (void)hipHostFree(c->h_dt);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return CBNS_OK;
}

int cbns_num_blocks(const cbns_ctx* c, int* nb) {
  if (!c || !nb) return CBNS_ERR_ARG;
  *nb = c->dom.nb;
  return CBNS_OK;
}

int cbns_block_info(const cbns_ctx* c, int b, int n[3], int origin[3]) {
  if (!c || b < 0 || b >= c->dom.nb) return CBNS_ERR_ARG;
  for (int a = 0; a < 3; a++) {
    if (n) n[a] = c->hb[b].n[a];
    if (origin) origin[a] = c->hb[b].o[a];
  }
  return CBNS_OK;
}

int cbns_memory_bytes(const cbns_ctx* c, uint64_t* bytes) {
  if (!c || !bytes) return CBNS_ERR_ARG;
  *bytes = c->bytes;
  return CBNS_OK;
}

int cbns_sync(cbns_ctx* c) {
  if (!c) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(c->stream));
  CK(hipStreamSynchronize(c->side));
  return CBNS_OK;
}

// ---- method-level -------------------------------------------------------

int cbns_boundary_condition(cbns_ctx* c, int which) {
  if (!c || (which != CBNS_Q && which != CBNS_QP)) return CBNS_ERR_ARG;
  c->s_ready = false;
  CK(launch_ghost_fill(c->db, c->hb.data(), c->dom, which, c->stream));
  return CBNS_OK;
}

int cbns_calc_primitive(cbns_ctx* c, int which) {
  if (!c || (which != CBNS_Q && which != CBNS_QP)) return CBNS_ERR_ARG;
  c->s_ready = false;
  c->prim_stale = false;
  CK(launch_primitive(c->db, c->hb.data(), c->dom.nb, which, c->ph, c->precision, c->stream));
  return CBNS_OK;
}

int cbns_init_Q_p(cbns_ctx* c) {
  if (!c) return CBNS_ERR_ARG;
  CK(launch_copy_state(c->db, c->hb.data(), c->dom.nb, CBNS_Q, CBNS_QP, c->stream));
  return CBNS_OK;
}

// dt = CFL / s_max from the reduction in d_smax / d_flag (src/Block3d.cpp:551).
static int read_dt(cbns_ctx* c, double* dt) {
  unsigned long long bits = 0;
  int flag = 0;
  CK(hipMemcpyAsync(&bits, c->d_smax, sizeof(bits), hipMemcpyDeviceToHost, c->stream));
  CK(hipMemcpyAsync(&flag, c->d_flag, sizeof(flag), hipMemcpyDeviceToHost, c->stream));
  CK(hipStreamSynchronize(c->stream));
  harvest(c);
  double s_max;
  std::memcpy(&s_max, &bits, sizeof(s_max));
  c->nonfinite = flag;
  *dt = c->par.CFL / s_max;  // src/Block3d.cpp:551
  c->s_ready = false;
  c->dt_inflight = false;
  return CBNS_OK;
}

int cbns_calc_dt(cbns_ctx* c, double* dt) {
  if (!c || !dt) return CBNS_ERR_ARG;
  if (!c->have_state) return CBNS_ERR_STATE;
  CK(hipMemsetAsync(c->d_smax, 0, sizeof(unsigned long long), c->stream));
  CK(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
  CK(timed(c, 7, [&] {
    return launch_dt_reduce(c->db, c->hb.data(), c->dom.nb, c->ph, c->d_smax, c->d_flag, c->precision, c->stream);
  }));
  return read_dt(c, dt);
}

int cbns_rec_riemann(cbns_ctx* c, int which, int dir) {
  if (!c || (which != CBNS_Q && which != CBNS_QP) || dir < 0 || dir > 2) return CBNS_ERR_ARG;
  CK(launch_face_flux(c->db, c->hb.data(), c->dom.nb, dir, which, c->ph, c->precision, c->ref, c->jband, c->face_occ,
                      c->fix, c->stream));
  if (c->precision == CBNS_FP64_FAST)
    CK(launch_face_fixup(c->db, c->hb.data(), c->dom.nb, which, c->ph, c->fix, c->stream, 1 << dir));
  return CBNS_OK;
}

int cbns_calc_viscous_flux_contribution(cbns_ctx* c) {
  if (!c) return CBNS_ERR_ARG;
  c->s_ready = false;
  int st = refresh_prims(c);
  if (st) return st;
  CK(launch_viscous(c->db, c->hb.data(), c->dom.nb, c->ph, c->precision, c->jband, c->stream));
  return CBNS_OK;
}

int cbns_update_rk3(cbns_ctx* c, double dt, int stage) {
  if (!c) return CBNS_ERR_ARG;
  c->s_ready = false;
  if (stage < 1 || stage > 3) return CBNS_OK;  // the reference silently does nothing
  CK(launch_update(c->db, c->hb.data(), c->dom.nb, dt, stage, c->ph, c->precision, c->jband, c->stream));
  return CBNS_OK;
}

// ---- setup ----------------------------------------------------------------

static int finish_setup(cbns_ctx* c) {
  // src/cbns3d.cpp:149-153
  int st;
  c->s_ready = false;
  if ((st = cbns_boundary_condition(c, CBNS_Q))) return st;
  if ((st = cbns_calc_primitive(c, CBNS_Q))) return st;
  if ((st = cbns_init_Q_p(c))) return st;
  CK(hipStreamSynchronize(c->stream));
  c->have_state = true;
  if (c->precision == CBNS_FP32_FLUX) {
    // reference state of the deviation form: interior means of rho and E
    double* out = c->d_diag + 2 * c->diag_parts;
    CK(launch_global_sums(c->db, c->dom, c->ph, 0, c->d_diag, out, c->precision, c->stream));
    double sums[2];
    CK(hipMemcpyAsync(sums, out, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
    CK(hipStreamSynchronize(c->stream));
    const double n = (double)c->dom.N[0] * c->dom.N[1] * c->dom.N[2];
    c->ref.rho = sums[0] / n;
    c->ref.E = sums[1] / n;
    c->ref.p = c->ph.gam1 * c->ref.E;
  }
  return CBNS_OK;
}

int cbns_init_tgv(cbns_ctx* c) {
  if (!c) return CBNS_ERR_ARG;
  const int* N = c->dom.N;
  const int ts = std::max(N[0], std::max(N[1], N[2]));
  std::vector<double> tab(9 * (size_t)ts, 0.0);
  const double tmp = 2.0 * 3.14159265358979323846;
  const double h[3] = {c->ph.dx, c->ph.dy, c->ph.dz};
  for (int a = 0; a < 3; a++)
    for (int i = 0; i < N[a]; i++) {
      // gen_mesh edge (src/Block3d.cpp:223-233) and cell centre (src/Block3d.cpp:243-251)
      const double edge = (unsigned)i * h[a] - tmp / 2.0;
      const double xc = edge + 0.5 * h[a];
      tab[(3 * a + 0) * ts + i] = std::sin(xc);
      tab[(3 * a + 1) * ts + i] = std::cos(xc);
      tab[(3 * a + 2) * ts + i] = std::cos(2.0 * xc);
    }
  double* dtab = nullptr;
  CK(hipMalloc(&dtab, sizeof(double) * tab.size()));
  int st = CBNS_OK;
  if (hipMemcpy(dtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess ||
      launch_tgv(c->db, c->hb.data(), c->dom.nb, dtab, ts, c->ph, c->par.p_inf, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    st = CBNS_ERR_HIP;
  (void)hipFree(dtab);
  if (st) return st;
  return finish_setup(c);
}

int cbns_set_state(cbns_ctx* c, const double* Q_aos, double t_cur, double mu0) {
  if (!c || !Q_aos) return CBNS_ERR_ARG;
  const size_t n = (size_t)NEQ * c->dom.N[0] * c->dom.N[1] * c->dom.N[2];
  double* d = nullptr;
  CK(hipMalloc(&d, n * sizeof(double)));
  int st = CBNS_OK;
  if (hipMemcpy(d, Q_aos, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      launch_scatter_aos(c->db, c->hb.data(), c->dom, CBNS_Q, d, c->stream) != hipSuccess ||
      launch_fill_mu(c->db, c->hb.data(), c->dom.nb, mu0, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    st = CBNS_ERR_HIP;
  (void)hipFree(d);
  if (st) return st;
  c->par.t_cur = t_cur;
  return finish_setup(c);
}

int cbns_get_state(cbns_ctx* c, int which, double* Q_aos) {
  if (!c || !Q_aos || (which != CBNS_Q && which != CBNS_QP)) return CBNS_ERR_ARG;
  const size_t n = (size_t)NEQ * c->dom.N[0] * c->dom.N[1] * c->dom.N[2];
  double* d = nullptr;
  CK(hipMalloc(&d, n * sizeof(double)));
  int st = CBNS_OK;
  if (launch_gather_aos(c->db, c->hb.data(), c->dom, which, d, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(Q_aos, d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    st = CBNS_ERR_HIP;
  (void)hipFree(d);
  return st;
}

int cbns_get_block_state(cbns_ctx* c, int b, int which, double* out) {
  if (!c || !out || b < 0 || b >= c->dom.nb || (which != CBNS_Q && which != CBNS_QP)) return CBNS_ERR_ARG;
  const BlockDesc& B = c->hb[b];
  std::vector<double> soa((size_t)NEQ * B.sg);
  CK(hipStreamSynchronize(c->stream));
  CK(hipMemcpy(soa.data(), B.Q[which], soa.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (long long t = 0; t < B.sg; t++)
    for (int e = 0; e < NEQ; e++) out[NEQ * t + e] = soa[e * B.sg + t];
  return CBNS_OK;
}

int cbns_set_block_state(cbns_ctx* c, int b, int which, const double* in) {
  if (!c || !in || b < 0 || b >= c->dom.nb || (which != CBNS_Q && which != CBNS_QP)) return CBNS_ERR_ARG;
  const BlockDesc& B = c->hb[b];
  std::vector<double> soa((size_t)NEQ * B.sg);
  for (long long t = 0; t < B.sg; t++)
    for (int e = 0; e < NEQ; e++) soa[e * B.sg + t] = in[NEQ * t + e];
  CK(hipStreamSynchronize(c->stream));
  CK(hipMemcpy(B.Q[which], soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
  c->have_state = true;
  c->s_ready = false;
  return CBNS_OK;
}

int cbns_get_block_primitive(cbns_ctx* c, int b, int f, double* out) {
  if (!c || !out || b < 0 || b >= c->dom.nb || f < 0 || f > 6) return CBNS_ERR_ARG;
  int st = refresh_prims(c);
  if (st) return st;
  const BlockDesc& B = c->hb[b];
  CK(hipStreamSynchronize(c->stream));
  if (f < 6 && c->precision == CBNS_FP32_FLUX) {  // float storage, widened
    std::vector<float> pf(B.sg);
    CK(hipMemcpy(pf.data(), (const float*)B.prim + f * B.sg, B.sg * sizeof(float), hipMemcpyDeviceToHost));
    for (long long t = 0; t < B.sg; t++) out[t] = pf[t];
    return CBNS_OK;
  }
  if (f < 6) {
    CK(hipMemcpy(out, B.prim + f * B.sg, B.sg * sizeof(double), hipMemcpyDeviceToHost));
    return CBNS_OK;
  }
  std::vector<double> mu(B.sn);
  CK(hipMemcpy(mu.data(), B.mu, B.sn * sizeof(double), hipMemcpyDeviceToHost));
  std::fill(out, out + B.sg, 0.0);
  for (int k = 0; k < B.n[2]; k++)
    for (int j = 0; j < B.n[1]; j++)
      for (int i = 0; i < B.n[0]; i++)
        out[(i + NG) + (long long)B.g[0] * ((j + NG) + (long long)B.g[1] * (k + NG))] =
            mu[i + (long long)B.n[0] * (j + (long long)B.n[1] * k)];
  return CBNS_OK;
}

// ---- asynchronous snapshot ----------------------------------------------

// bytes of one block's primitive field in the snapshot (their storage type)
static size_t snap_prim_bytes(const cbns_ctx* c) {
  return c->precision == CBNS_FP32_FLUX ? sizeof(float) : sizeof(double);
}

int cbns_snapshot_begin(cbns_ctx* c) {
  if (!c) return CBNS_ERR_ARG;
  if (!c->have_state) return CBNS_ERR_STATE;
  const int nb = c->dom.nb;
  const size_t pb = snap_prim_bytes(c);
  if (!c->snap_dev) {  // first snapshot: buffers, copy stream, events
    c->snap_off.resize(nb);
    size_t off = 0;
    for (int b = 0; b < nb; b++) {
      c->snap_off[b] = off;
      off += (size_t)c->hb[b].sg * (NEQ * sizeof(double) + 5 * pb);
      off = (off + 255) & ~(size_t)255;
    }
    if (hipMalloc(&c->snap_dev, off) != hipSuccess) {
      c->snap_dev = nullptr;
      return CBNS_ERR_NOMEM;
    }
    if (hipHostMalloc((void**)&c->snap_host, off, hipHostMallocDefault) != hipSuccess) {
      c->snap_host = nullptr;
      (void)hipFree(c->snap_dev);
      c->snap_dev = nullptr;
      return CBNS_ERR_NOMEM;
    }
    c->snap_bytes = off;
    CK(hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking));
    CK(hipEventCreateWithFlags(&c->ev_snap_dev, hipEventDisableTiming));
    CK(hipEventCreateWithFlags(&c->ev_snap_host, hipEventDisableTiming));
  }
  // the previous snapshot's host copy must end before its device copy is overwritten
  if (c->snap_pending) CK(hipStreamWaitEvent(c->stream, c->ev_snap_host, 0));
  // with overlap on, the side stream may still write mu and the viscous fields,
  // never Q or the primitives: the stream's order is enough
  int st = refresh_prims(c);
  if (st) return st;
  for (int b = 0; b < nb; b++) {
    const BlockDesc& B = c->hb[b];
    char* dst = c->snap_dev + c->snap_off[b];
    CK(hipMemcpyAsync(dst, B.Q[CBNS_Q], (size_t)B.sg * NEQ * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    CK(hipMemcpyAsync(dst + (size_t)B.sg * NEQ * sizeof(double), B.prim, (size_t)B.sg * 5 * pb,
                      hipMemcpyDeviceToDevice, c->stream));
  }
  CK(hipEventRecord(c->ev_snap_dev, c->stream));
  CK(hipStreamWaitEvent(c->copy, c->ev_snap_dev, 0));
  CK(hipMemcpyAsync(c->snap_host, c->snap_dev, c->snap_bytes, hipMemcpyDeviceToHost, c->copy));
  CK(hipEventRecord(c->ev_snap_host, c->copy));
  c->snap_pending = true;
  c->snap_t = c->par.t_cur;
  return CBNS_OK;
}

int cbns_snapshot_wait(cbns_ctx* c, double* t_cur) {
  if (!c) return CBNS_ERR_ARG;
  if (!c->snap_pending) return CBNS_ERR_STATE;
  CK(hipEventSynchronize(c->ev_snap_host));
  if (t_cur) *t_cur = c->snap_t;
  return CBNS_OK;
}

int cbns_snapshot_block_state(cbns_ctx* c, int b, double* out) {
  if (!c || !out || b < 0 || b >= c->dom.nb) return CBNS_ERR_ARG;
  int st = cbns_snapshot_wait(c, nullptr);
  if (st) return st;
  const BlockDesc& B = c->hb[b];
  const double* soa = (const double*)(c->snap_host + c->snap_off[b]);
  for (long long t = 0; t < B.sg; t++)
    for (int e = 0; e < NEQ; e++) out[NEQ * t + e] = soa[e * B.sg + t];
  return CBNS_OK;
}

int cbns_snapshot_block_primitive(cbns_ctx* c, int b, int f, double* out) {
  if (!c || !out || b < 0 || b >= c->dom.nb || f < 0 || f > 4) return CBNS_ERR_ARG;
  int st = cbns_snapshot_wait(c, nullptr);
  if (st) return st;
  const BlockDesc& B = c->hb[b];
  const char* p = c->snap_host + c->snap_off[b] + (size_t)B.sg * NEQ * sizeof(double);
  if (c->precision == CBNS_FP32_FLUX) {  // float storage, widened
    const float* pf = (const float*)p + (size_t)f * B.sg;
    for (long long t = 0; t < B.sg; t++) out[t] = pf[t];
  } else {
    std::memcpy(out, (const double*)p + (size_t)f * B.sg, (size_t)B.sg * sizeof(double));
  }
  return CBNS_OK;
}

int cbns_spectra(cbns_ctx* c, int nshell, double* ke, double* enstrophy, double* dilatation) {
  if (!c || nshell < 1) return CBNS_ERR_ARG;
  if (!c->have_state) return CBNS_ERR_STATE;
  if (!c->spec && !(c->spec = spectra_create())) return CBNS_ERR_NOMEM;
  std::vector<double> h(3 * (size_t)nshell);
  const int r = spectra_run(c->spec, c->db, c->hb.data(), c->dom, CBNS_Q, nshell, h.data(), c->stream);
  if (r == 1) return CBNS_ERR_NOMEM;
  if (r == 3) return CBNS_ERR_ARG;
  if (r) return CBNS_ERR_HIP;
  double* outs[3] = {ke, enstrophy, dilatation};
  for (int a = 0; a < 3; a++)
    if (outs[a]) std::memcpy(outs[a], h.data() + (size_t)a * nshell, sizeof(double) * nshell);
  return CBNS_OK;
}

int cbns_get_block_mu(cbns_ctx* c, int b, double* mu) {
  if (!c || !mu || b < 0 || b >= c->dom.nb) return CBNS_ERR_ARG;
  const BlockDesc& B = c->hb[b];
  CK(hipStreamSynchronize(c->stream));
  CK(hipStreamSynchronize(c->side));
  CK(hipMemcpy(mu, B.mu, B.sn * sizeof(double), hipMemcpyDeviceToHost));
  return CBNS_OK;
}

int cbns_set_block_mu(cbns_ctx* c, int b, const double* mu) {
  if (!c || !mu || b < 0 || b >= c->dom.nb) return CBNS_ERR_ARG;
  const BlockDesc& B = c->hb[b];
  CK(hipStreamSynchronize(c->stream));
  CK(hipStreamSynchronize(c->side));
  CK(hipMemcpy(B.mu, mu, B.sn * sizeof(double), hipMemcpyHostToDevice));
  c->s_ready = false;  // the next dt reads the new mu
  c->dt_inflight = false;
  return CBNS_OK;
}

// ---- fused stepping -------------------------------------------------------

int cbns_get_time(const cbns_ctx* c, double* t) {
  if (!c || !t) return CBNS_ERR_ARG;
  *t = c->par.t_cur;
  return CBNS_OK;
}

int cbns_set_time(cbns_ctx* c, double t) {
  if (!c) return CBNS_ERR_ARG;
  c->par.t_cur = t;
  return CBNS_OK;
}

// src/cbns3d.cpp:203-240 per stage: the x and y face sweeps and the
// viscous march (beside them on the side stream when overlap is on; the
// tolerance-bound form runs its viscous march first, then one march over its
// x and y faces that stores each cell's x+y flux differences less the
// viscous divergence), the z-face sweep fused with the RK3 update (and, but
// for the tolerance-bound form, the primitives) with its cell fixup, and the
// ghost gather of the state (with its primitives).  The contracted FP64 form
// keeps a separate z sweep and update.  launch_sweeps issues the kernels
// that read only the stage input, finish_stage the rest.
// Whether a form's steps run the z-face sweep fused with the stage tail
// (every form but the contracted FP64 one).
static bool zfused(const cbns_ctx* c) { return c->precision != CBNS_FP64_FMA; }

// Input and output state of a fused stage.  The fused z pass keeps the stage
// input unwritten (it updates while it sweeps): stage 1 goes from Q to a
// third state S, stage 2 from S to Q_p, stage 3 from Q_p to Q, so Q_p holds
// the reference's stage-2 state after each step.  The contracted form
// updates Q_p and Q in place as the reference does.
static void stage_io(const cbns_ctx* c, int stage, int* in, int* out) {
  if (zfused(c)) {
    *in = stage == 1 ? CBNS_Q : (stage == 2 ? 2 : CBNS_QP);
    *out = stage == 1 ? 2 : (stage == 2 ? CBNS_QP : CBNS_Q);
  } else {
    *in = stage == 1 ? CBNS_Q : CBNS_QP;
    *out = stage == 3 ? CBNS_Q : CBNS_QP;
  }
}

// Whether the viscous pass of a stage runs beside its face sweeps.  The
// tolerance-bound form's x+y march reads the divergence, so its viscous
// pass runs first, on the context's stream.
static bool side_viscous(const cbns_ctx* c) { return c->overlap && c->precision != CBNS_FP64_FAST; }

static int launch_sweeps(cbns_ctx* c, int stage) {
  int in, out;
  stage_io(c, stage, &in, &out);
  const BlockDesc* hb = c->hb.data();
  if (side_viscous(c)) {
    CK(hipEventRecord(c->ev_fork, c->stream));
    CK(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    CK(timed_on(c, c->side, 3, [&] {
      return launch_visc_flux(c->db, hb, c->dom.nb, c->ph, c->precision, c->jband, c->side, c->visc_wgs,
                              c->precision == CBNS_FP64_FAST ? in : -1, c->visc_top);
    }));
    CK(hipEventRecord(c->ev_join, c->side));
  }
  // the z faces are formed by the fused stage tail
  if (c->precision == CBNS_FP64_FAST) {
    // the viscous march from the stage input, then the x and y faces in one
    // march that stores each cell's x- plus y-flux differences less the
    // viscous divergence (timed as the y class: it carries both axes)
    CK(timed(c, 3, [&] {
      return launch_visc_flux(c->db, hb, c->dom.nb, c->ph, c->precision, c->jband, c->stream, 0, in, c->visc_top);
    }));
    CK(timed(c, 1, [&] { return launch_xyface(c->db, hb, c->dom.nb, in, c->ph, c->stream, c->run_top); }));
    return CBNS_OK;
  }
  const int nd = zfused(c) ? 2 : 3;
  for (int d = 0; d < nd; d++)
    CK(timed(c, d, [&] {
      return launch_face_flux(c->db, hb, c->dom.nb, d, in, c->ph, c->precision, c->ref, c->jband, c->face_occ,
                              c->fix, c->stream);
    }));
  return CBNS_OK;
}

static int finish_stage(cbns_ctx* c, int stage, double dt) {
  int in, out;
  stage_io(c, stage, &in, &out);
  const BlockDesc* hb = c->hb.data();
  // the tolerance-bound form stores no primitives in its fused steps: its
  // viscous march forms them from the stage input, its stage tail (fused
  // with its z faces) skips them and its ghost fill copies the state alone
  const bool lean = c->precision == CBNS_FP64_FAST;
  if (side_viscous(c)) CK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  else if (!lean)  // (the tolerance-bound form's ran before its x+y march)
    CK(timed(c, 3, [&] {
      return launch_visc_flux(c->db, hb, c->dom.nb, c->ph, c->precision, c->jband, c->stream, 0, -1, c->visc_top);
    }));
  bool s_done = false;
  if (stage == 3) {  // the next step's calc_dt reduction rides on the last update
    CK(hipMemsetAsync(c->d_smax, 0, sizeof(unsigned long long), c->stream));
    CK(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
  }
  if (zfused(c)) {
    unsigned long long* sm = stage == 3 ? c->d_smax : nullptr;
    int* fl = stage == 3 ? c->d_flag : nullptr;
    CK(timed(c, 4, [&] {
      return launch_zface_update(c->db, hb, c->dom.nb, c->precision, in, out, stage, dt, c->ph, c->ref, sm, fl,
                                 c->cfix, c->stream, c->run_top);
    }));
    if (c->precision != CBNS_FP32_FLUX)
      CK(launch_zcell_fixup(c->db, hb, c->dom.nb, c->precision, in, out, stage, dt, c->ph, sm, fl, c->cfix,
                            c->stream));
    s_done = stage == 3;
  } else {
    CK(timed(c, 4, [&] {
      return launch_update_fused(c->db, hb, c->dom.nb, dt, stage, c->ph, c->precision, c->jband, c->stream,
                                 stage == 3 ? c->d_smax : nullptr, stage == 3 ? c->d_flag : nullptr, &s_done);
    }));
  }
  if (stage == 3) {
    c->s_ready = s_done;
    c->dt_inflight = false;
    if (s_done) {
      CK(hipMemcpyAsync(&c->h_dt->smax, c->d_smax, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
      CK(hipMemcpyAsync(&c->h_dt->flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
      CK(hipEventRecord(c->ev_dt, c->stream));
      c->dt_inflight = true;
    }
  }
  CK(timed(c, 5, [&] {
    return lean ? launch_ghost_fill(c->db, hb, c->dom, out, c->stream)
                : launch_ghost_fill_prim(c->db, hb, c->dom, out, c->precision, c->stream);
  }));
  if (lean) c->prim_stale = true;
  return CBNS_OK;
}

// first_issued: stage 1's sweeps are already on the stream.
static int run_stages(cbns_ctx* c, double dt, bool first_issued) {
  for (int stage = 1; stage <= 3; stage++) {
    if (stage > 1 || !first_issued) {
      int st = launch_sweeps(c, stage);
      if (st) return st;
    }
    int st = finish_stage(c, stage, dt);
    if (st) return st;
  }
  return CBNS_OK;
}

int cbns_step(cbns_ctx* c, double* dt_out, int* finished) {
  if (!c) return CBNS_ERR_ARG;
  if (!c->have_state) return CBNS_ERR_STATE;
  if (finished) *finished = 0;
  // is_finished (src/Block3d.cpp:502-513)
  if (c->par.t_cur + 1.0e-5 > c->par.t_end) {
    if (finished) *finished = 1;
    return CBNS_OK;
  }
  double dt;
  bool first_issued = false;
  if (c->s_ready && c->dt_inflight) {
    // stage 1's sweeps read only the state: issue them, then wait for the
    // last update's dt reduction (already on its way to pinned memory)
    // while they run, instead of leaving the GPU idle for the round trip
    int st = launch_sweeps(c, 1);
    if (st) return st;
    first_issued = true;
    CK(hipEventSynchronize(c->ev_dt));
    double s_max;
    std::memcpy(&s_max, &c->h_dt->smax, sizeof(s_max));
    c->nonfinite = c->h_dt->flag;
    dt = c->par.CFL / s_max;  // src/Block3d.cpp:551
    c->s_ready = false;
    c->dt_inflight = false;
  } else {
    int st = c->s_ready ? read_dt(c, &dt) : cbns_calc_dt(c, &dt);
    if (st) return st;
  }
  if (c->nonfinite) {
    // the issued sweeps change no state, but with overlap on the side
    // stream's viscous kernel of stage 1 writes mu and the viscous fields:
    // order every later call on the context's stream after it
    if (first_issued && side_viscous(c)) CK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
    return CBNS_ERR_NONFINITE;
  }
  // src/cbns3d.cpp:191-194
  if (c->par.t_end - c->par.t_cur < dt) dt = c->par.t_end - c->par.t_cur;
  c->par.t_cur += dt;
  if (dt_out) *dt_out = dt;
  return run_stages(c, dt, first_issued);
}

int cbns_run(cbns_ctx* c, int nsteps, double* dts, int* steps_done) {
  if (!c || nsteps < 0) return CBNS_ERR_ARG;
  int n = 0;
  for (; n < nsteps; n++) {
    int fin = 0;
    double dt = 0.0;
    int st = cbns_step(c, &dt, &fin);
    if (st) {
      if (steps_done) *steps_done = n;
      return st;
    }
    if (fin) break;
    if (dts) dts[n] = dt;
  }
  if (steps_done) *steps_done = n;
  CK(hipStreamSynchronize(c->stream));
  harvest(c);
  return CBNS_OK;
}

int cbns_set_band(cbns_ctx* c, int rows) {
  if (!c || rows < 1) return CBNS_ERR_ARG;
  c->jband = rows;
  return CBNS_OK;
}

int cbns_set_face_occupancy(cbns_ctx* c, int waves) {
  if (!c || (waves != 2 && waves != 3)) return CBNS_ERR_ARG;
  c->face_occ = waves;
  return CBNS_OK;
}

int cbns_set_tile_shape(cbns_ctx* c, int visc_planes, int run_cells) {
  if (!c) return CBNS_ERR_ARG;
  if (visc_planes != 0 && (visc_planes < VISC_LOW || visc_planes > 256)) return CBNS_ERR_ARG;
  if (run_cells != 0 && (run_cells < ZM_LOW || run_cells > 256)) return CBNS_ERR_ARG;
  c->visc_top = visc_planes ? visc_planes : VZ;
  c->run_top = run_cells ? run_cells : ZM;
  return CBNS_OK;
}

int cbns_set_overlap(cbns_ctx* c, int on) {
  if (!c) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(c->stream));
  c->overlap = on != 0;
  return CBNS_OK;
}

int cbns_set_viscous_workgroups(cbns_ctx* c, int workgroups) {
  if (!c || workgroups < 0) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(c->stream));
  c->visc_wgs = workgroups;
  return CBNS_OK;
}

int cbns_set_profiling_classes(cbns_ctx* c, unsigned mask) {
  if (!c || (mask >> CBNS_KCLASSES)) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(c->stream));
  if (c->side) CK(hipStreamSynchronize(c->side));
  harvest(c);
  c->prof = mask;
  for (int k = 0; k < CBNS_KCLASSES; k++) {
    c->k_ms[k] = 0.0;
    c->k_n[k] = 0;
  }
  return CBNS_OK;
}

int cbns_set_profiling(cbns_ctx* c, int on) {
  return cbns_set_profiling_classes(c, on ? (1u << CBNS_KCLASSES) - 1 : 0u);
}

int cbns_kernel_times(cbns_ctx* c, double ms[CBNS_KCLASSES], int64_t n[CBNS_KCLASSES]) {
  if (!c || !ms || !n) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(c->stream));
  harvest(c);
  for (int k = 0; k < CBNS_KCLASSES; k++) {
    ms[k] = c->k_ms[k];
    n[k] = c->k_n[k];
  }
  return CBNS_OK;
}

int cbns_diagnostics(cbns_ctx* c, double* ke, double* ens) {
  if (!c || !ke || !ens) return CBNS_ERR_ARG;
  int st = refresh_prims(c);
  if (st) return st;
  CK(launch_global_sums(c->db, c->dom, c->ph, 1, c->d_diag, c->d_diag + 2 * c->diag_parts, c->precision,
                        c->stream));
  double out[2];
  CK(hipMemcpyAsync(out, c->d_diag + 2 * c->diag_parts, sizeof(out), hipMemcpyDeviceToHost, c->stream));
  CK(hipStreamSynchronize(c->stream));
  const double n = (double)c->dom.N[0] * c->dom.N[1] * c->dom.N[2];
  *ke = out[0] / n;
  *ens = out[1] / n;
  return CBNS_OK;
}

int cbns_face_flux_eval(int precision, int dir, double gamma, int nf, const double* cells, const double* prim,
                        double* flux) {
  if (!cells || !prim || !flux || nf < 0 || dir < 0 || dir > 2) return CBNS_ERR_ARG;
  Phys ph{};
  ph.gamma = gamma;
  ph.gam1 = gamma - 1.0;
  reciprocal_check(ph.gam1, &ph.rgam1, &ph.fgam1);
  ph.gg1 = ph.gamma / ph.gam1;
  if ((precision < CBNS_FP64 || precision > CBNS_FP64_FAST)) return CBNS_ERR_ARG;
  if (nf == 0) return CBNS_OK;
  double *dc = nullptr, *dp = nullptr, *df = nullptr;
  int st = CBNS_OK;
  if (hipMalloc(&dc, sizeof(double) * 30 * (size_t)nf) != hipSuccess ||
      hipMalloc(&dp, sizeof(double) * 10 * (size_t)nf) != hipSuccess ||
      hipMalloc(&df, sizeof(double) * 5 * (size_t)nf) != hipSuccess)
    st = CBNS_ERR_NOMEM;
  if (!st && (hipMemcpy(dc, cells, sizeof(double) * 30 * (size_t)nf, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(dp, prim, sizeof(double) * 10 * (size_t)nf, hipMemcpyHostToDevice) != hipSuccess ||
              launch_face_eval(dir, ph, precision, nf, dc, dp, df, nullptr) != hipSuccess ||
              hipDeviceSynchronize() != hipSuccess ||
              hipMemcpy(flux, df, sizeof(double) * 5 * (size_t)nf, hipMemcpyDeviceToHost) != hipSuccess))
    st = CBNS_ERR_HIP;
  if (dc) (void)hipFree(dc);
  if (dp) (void)hipFree(dp);
  if (df) (void)hipFree(df);
  return st;
}

}  // extern "C"

// ---- face flux evaluation without device allocation -----------------------
// Staging of cbns_face_flux_eval_noalloc: per face 30 cell values, 10
// primitive values and 5 flux values, for up to CBNS_NOALLOC_MAX_FACES faces.
// One copy per GPU, alive until the process ends.  In host code `scratch`
// names a host-side stand-in: its device address on each GPU comes from
// hipGetSymbolAddress (scratch_on_current_device) and only that address is
// passed to copies and kernels.
static constexpr size_t NOALLOC_CELLS = 30 * (size_t)CBNS_NOALLOC_MAX_FACES;
static constexpr size_t NOALLOC_PRIM = 10 * (size_t)CBNS_NOALLOC_MAX_FACES;
static constexpr size_t NOALLOC_FLUX = 5 * (size_t)CBNS_NOALLOC_MAX_FACES;
static __device__ __align__(256) double scratch[NOALLOC_CELLS + NOALLOC_PRIM + NOALLOC_FLUX];

static constexpr int NOALLOC_MAXDEV = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[NOALLOC_MAXDEV] = {};

// Device address of the current device's scratch, looked up once per device
// (with that device current) and kept; CBNS_ERR_NODEV on a device that is
// not a gfx950.
static int scratch_on_current_device(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAXDEV) return CBNS_ERR_NODEV;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess || std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
      return CBNS_ERR_NODEV;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CBNS_ERR_HIP;
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return CBNS_OK;
}

extern "C" int cbns_face_flux_eval_noalloc(int precision, int dir, double gamma, int nf, const double* cells,
                                           const double* prim, double* flux) {
  if (!cells || !prim || !flux || nf < 0 || dir < 0 || dir > 2) return CBNS_ERR_ARG;
  if ((precision < CBNS_FP64 || precision > CBNS_FP64_FAST)) return CBNS_ERR_ARG;
  if (nf > CBNS_NOALLOC_MAX_FACES) return CBNS_ERR_CAPACITY;
  if (nf == 0) return CBNS_OK;
  Phys ph{};
  ph.gamma = gamma;
  ph.gam1 = gamma - 1.0;
  reciprocal_check(ph.gam1, &ph.rgam1, &ph.fgam1);
  ph.gg1 = ph.gamma / ph.gam1;
  double* base = nullptr;
  int st = scratch_on_current_device(&base);
  if (st) return st;
  double* dc = base;
  double* dp = base + NOALLOC_CELLS;
  double* df = base + NOALLOC_CELLS + NOALLOC_PRIM;
  // every staged value is overwritten before it is read: the cells and
  // primitives by the copies, the fluxes by the kernel (one per face)
  CK(hipMemcpy(dc, cells, sizeof(double) * 30 * (size_t)nf, hipMemcpyHostToDevice));
  CK(hipMemcpy(dp, prim, sizeof(double) * 10 * (size_t)nf, hipMemcpyHostToDevice));
  CK(launch_face_eval(dir, ph, precision, nf, dc, dp, df, nullptr));
  CK(hipMemcpy(flux, df, sizeof(double) * 5 * (size_t)nf, hipMemcpyDeviceToHost));
  return CBNS_OK;
}
