// C ABI of the batched form: many independent solves of uneven sizes in one
// set of persistent launches per RK stage, one hipGraph per RK3 step.
//
// Composition changes (add / remove) rebuild the packed block and tile tables
// and re-capture the step graph; between changes, cbns_batch_run replays the
// graph with no host synchronisation.  Every solve's arithmetic is the
// single-solve path's (shared per-item bodies, per-solve constants, an
// order-independent dt maximum), so a solve's result does not depend on what
// it is batched with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "../../include/cbns3d_amd.h"
#include "cbns_batch.h"
#include "cbns_kernels.h"

using namespace cbns;

namespace {

struct Inst {
  bool live = false;
  cbns_params par{};
  int dims[3] = {1, 1, 1};
  int precision = 0;
  Domain dom{};
  std::vector<BlockDesc> blocks;  // block0-relative (instance-local) descriptors
  BlockDesc* dblocks = nullptr;   // instance-local device copy (setup, gather)
  std::vector<std::pair<void*, size_t>> allocs;
  uint64_t bytes = 0;
  int block0 = 0;
};

enum {
  T_FACE64 = 0, T_FACE32 = 3,
  T_VISC64 = 6, T_VISC32, T_VISCF,
  T_UPD64, T_UPD32, T_UPDF,
  T_YMF,  // the tolerance-bound solves' x+y march
  T_GHOST, T_DT, T_COUNT
};

// Kernel form of a solve: 0 FP64, 1 FP32-flux, 2 tolerance-bound FP64.
int face_form(int precision) { return precision == CBNS_FP32_FLUX ? 1 : (precision == CBNS_FP64_FAST ? 2 : 0); }
const int FACE_CLASS[2] = {T_FACE64, T_FACE32};  // the tolerance-bound solves march theirs (T_YMF)
const int VISC_CLASS[3] = {T_VISC64, T_VISC32, T_VISCF};
const int UPD_CLASS[3] = {T_UPD64, T_UPD32, T_UPDF};
const int FORM_PREC[3] = {PREC_FP64, PREC_FP32_FLUX, PREC_FP64_FAST};

}  // namespace

struct cbns_batch {
  uint64_t budget = 0, used = 0;
  int cap = 0;
  std::vector<Inst> inst;
  InstDesc* d_insts = nullptr;
  BlockDesc* d_blocks = nullptr;
  int d_blocks_cap = 0;
  TileRef* d_tiles[T_COUNT] = {nullptr};
  int tiles_cap[T_COUNT] = {0};
  TileTable tt[T_COUNT] = {};
  double* d_part = nullptr;  // global-sum partials (FP32 reference states)
  int2* d_pairs = nullptr;   // staging of step budgets
  int pairs_cap = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // viscous kernels beside the face sweeps (overlap)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int visc_side_wgs = 0;       // 0: viscous kernels in line; > 0: on the side stream with this grid
                               // (default: one workgroup per CU, the fastest measured)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool dirty = true;
  int jb = 32;
  BatchGrids grid{};  // resident workgroups of each persistent kernel
  int* d_ctr = nullptr;  // tile counters of the step graph's launches
  FixList cfix{};        // cells of FP64 solves that the fused z pass left to the cell fixup
  // Arrays of removed solves, kept for reuse by later solves of the same
  // shapes (by size): hipFree synchronises the whole device, which would stall
  // every other stream of a server between its chunks.  Released when fresh
  // memory would take used + pooled past the budget, and at destroy.
  std::map<size_t, std::vector<void*>> pool;
  uint64_t pooled = 0;
  double* d_stage = nullptr;  // grow-only staging of states and tables (add, get_state)
  size_t stage_cap = 0;
  hipEvent_t ev_run = nullptr;  // recorded after the last run's graph launches
  // status of every descriptor after the last run (host-mapped, written by
  // kb_publish at the end of each run); valid for reading once ev_run is done
  InstStatus* h_status = nullptr;
  InstStatus* d_status = nullptr;
  bool published = false;  // h_status reflects the device descriptors
  bool has_fast = false; // a live tolerance-bound solve (the step graph has its x+y march)
  bool has_f64 = false;  // a live FP64 solve of either form (the step graph has cell fixup passes)
  int wg_mem = 0;     // override of the memory-bound kernels' grid (0: resident)
  int ninst_hw = 0;  // highest live id + 1
};

#define CK(call)                                 \
  do {                                           \
    hipError_t e_ = (call);                      \
    if (e_ != hipSuccess) return CBNS_ERR_HIP;   \
  } while (0)

static Phys make_phys(const cbns_params* p) {
  Phys ph{};
  const double two_pi = 2.0 * 3.14159265358979323846;
  ph.gamma = p->gamma;
  ph.gam1 = p->gam1;
  ph.gM2 = p->gM2;
  ph.CFL = p->CFL;
  ph.C_T_inf = p->C_T_inf;
  ph.Re_inv = p->Re_inv;
  ph.kappa = -p->Pr_inv * p->gam1_inv / p->Mach2;
  ph.C_dt_v = p->C_dt_v;
  ph.dx = two_pi / p->IM;
  ph.dy = two_pi / p->JM;
  ph.dz = two_pi / p->KM;
  reciprocal_check(ph.dx, &ph.rdx, &ph.fdx);
  reciprocal_check(ph.dy, &ph.rdy, &ph.fdy);
  reciprocal_check(ph.dz, &ph.rdz, &ph.fdz);
  reciprocal_check(ph.gam1, &ph.rgam1, &ph.fgam1);
  ph.gg1 = ph.gamma / ph.gam1;
  return ph;
}

// Bytes of device memory one solve needs (its blocks' arrays).
static uint64_t solve_bytes(const cbns_params* p, const int dims[3]) {
  const int N[3] = {(int)p->IM, (int)p->JM, (int)p->KM};
  uint64_t total = 0;
  for (int bx = 0; bx < dims[0]; bx++)
    for (int by = 0; by < dims[1]; by++)
      for (int bz = 0; bz < dims[2]; bz++) {
        const int bc[3] = {bx, by, bz};
        uint64_t n[3];
        for (int a = 0; a < 3; a++) {
          const int base = N[a] / dims[a];
          n[a] = (bc[a] == dims[a] - 1) ? N[a] - (dims[a] - 1) * base : base;
        }
        const uint64_t g = (n[0] + 8) * (n[1] + 8) * (n[2] + 8);
        const uint64_t in = n[0] * n[1] * n[2];
        const uint64_t f = (n[0] + 1) * n[1] * n[2] + n[0] * (n[1] + 1) * n[2] + n[0] * n[1] * (n[2] + 1);
        total += 8 * (3 * NEQ * g + NEQ * f + 6 * g + in + 4 * in);
      }
  return total + sizeof(BlockDesc) * dims[0] * dims[1] * dims[2];
}

// A solve's arrays go to the batch's pool (see cbns_batch::pool).
// Host <-> device copies of the batch go on its own stream (then wait for
// it): a copy on the null stream can queue behind another batch's work on a
// shared hardware queue.
static hipError_t copy_on(cbns_batch* b, void* dst, const void* src, size_t n, hipMemcpyKind k) {
  hipError_t e = hipMemcpyAsync(dst, src, n, k, b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  return e;
}

static void free_inst(cbns_batch* b, Inst& I) {
  for (auto& a : I.allocs) {
    b->pool[a.second].push_back(a.first);
    b->pooled += a.second;
  }
  I.allocs.clear();
  I.blocks.clear();
  I.dblocks = nullptr;
  I.live = false;
}

static void drain_pool(cbns_batch* b) {
  for (auto& kv : b->pool)
    for (void* p : kv.second) (void)hipFree(p);
  b->pool.clear();
  b->pooled = 0;
}

// Device memory for a solve: a pooled array of the same size, else a fresh
// one (after releasing the pool if it would take the batch past its budget).
static void* batch_alloc(cbns_batch* b, size_t bytes, uint64_t need) {
  auto it = b->pool.find(bytes);
  if (it != b->pool.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    b->pooled -= bytes;
    return p;
  }
  if (b->used + b->pooled + need > b->budget) drain_pool(b);
  void* q = nullptr;
  if (hipMalloc(&q, bytes) != hipSuccess) return nullptr;
  return q;
}

// Grow-only staging buffer of at least `bytes`.
static int stage(cbns_batch* b, size_t bytes, double** out) {
  if (bytes > b->stage_cap) {
    if (b->d_stage) (void)hipFree(b->d_stage);
    b->d_stage = nullptr;
    b->stage_cap = 0;
    if (hipMalloc(&b->d_stage, bytes) != hipSuccess) return CBNS_ERR_NOMEM;
    b->stage_cap = bytes;
  }
  *out = b->d_stage;
  return CBNS_OK;
}

static int grow(void** p, int* cap, int need, size_t elem) {
  if (need <= *cap) return CBNS_OK;
  int nc = std::max(need, *cap * 2 + 16);
  void* q = nullptr;
  if (hipMalloc(&q, (size_t)nc * elem) != hipSuccess) return CBNS_ERR_NOMEM;
  if (*p) (void)hipFree(*p);
  *p = q;
  *cap = nc;
  return CBNS_OK;
}

// Rebuild the packed block table, the per-class tile tables and the step graph.
static int rebuild(cbns_batch* b) {
  CK(hipStreamSynchronize(b->stream));
  std::vector<BlockDesc> all;
  std::vector<TileRef> tab[T_COUNT];
  int hw = 0;
  for (int id = 0; id < b->cap; id++) {
    Inst& I = b->inst[id];
    if (!I.live) continue;
    hw = id + 1;
    I.block0 = (int)all.size();
    for (size_t k = 0; k < I.blocks.size(); k++) {
      BlockDesc B = I.blocks[k];
      B.inst = id;
      const int bi = (int)all.size();
      all.push_back(B);
      auto add = [&](int cls, long long count) {
        const int nt = (int)((count + 255) / 256);
        for (int t = 0; t < nt; t++) tab[cls].push_back({bi, t});
      };
      // x and y sweeps (the z faces are formed by the fused stage tail); the
      // tolerance-bound solves march their x and y faces together in
      // (column tile, y-run) units, the run lengths as the kernel forms them
      const bool fast = I.precision == CBNS_FP64_FAST;
      if (!fast)
        for (int d = 0; d < 2; d++) add(FACE_CLASS[face_form(I.precision)] + d, face_items(B, d));
      else {
        const int yt = ((B.n[0] + ZX - 1) / ZX) * ((B.n[2] + ZY - 1) / ZY);
        const int ym = chunk_len(yt, B.n[1], BATCH_ZM, ZM_LOW, BATCH_UNITS);
        for (int t = 0; t < yt * ((B.n[1] + ym - 1) / ym); t++) tab[T_YMF].push_back({bi, t});
      }
      // whole-workgroup units of the viscous march and of the fused stage
      // tail: (column tile, z-chunk), the chunk lengths as the kernels form them
      const int vt = ((B.n[0] + VX - 1) / VX) * ((B.n[1] + VY - 1) / VY);
      const int vz = chunk_len(vt, B.n[2], BATCH_VZ, VISC_LOW, BATCH_UNITS);
      for (int t = 0; t < vt * ((B.n[2] + vz - 1) / vz); t++) tab[VISC_CLASS[face_form(I.precision)]].push_back({bi, t});
      const int zt = ((B.n[0] + ZX - 1) / ZX) * ((B.n[1] + ZY - 1) / ZY);
      const int zm = chunk_len(zt, B.n[2], BATCH_ZM, ZM_LOW, BATCH_UNITS);
      for (int t = 0; t < zt * ((B.n[2] + zm - 1) / zm); t++) tab[UPD_CLASS[face_form(I.precision)]].push_back({bi, t});
      add(T_GHOST, B.sg - B.sn);  // ghost cells only
      add(T_DT, B.sn);
    }
    // the device descriptor's block0 (other fields hold live state)
    CK(copy_on(b, (char*)(b->d_insts + id) + offsetof(InstDesc, block0), &I.block0, sizeof(int),
                 hipMemcpyHostToDevice));
  }
  b->ninst_hw = hw;
  b->has_fast = b->has_f64 = false;
  for (int id = 0; id < b->cap; id++) {
    if (!b->inst[id].live) continue;
    b->has_fast = b->has_fast || b->inst[id].precision == CBNS_FP64_FAST;
    b->has_f64 = b->has_f64 || b->inst[id].precision != CBNS_FP32_FLUX;
  }
  int st;
  if ((st = grow((void**)&b->d_blocks, &b->d_blocks_cap, (int)all.size() + 1, sizeof(BlockDesc)))) return st;
  if (!all.empty())
    CK(copy_on(b, b->d_blocks, all.data(), all.size() * sizeof(BlockDesc), hipMemcpyHostToDevice));
  for (int c = 0; c < T_COUNT; c++) {
    if ((st = grow((void**)&b->d_tiles[c], &b->tiles_cap[c], (int)tab[c].size() + 1, sizeof(TileRef)))) return st;
    if (!tab[c].empty())
      CK(copy_on(b, b->d_tiles[c], tab[c].data(), tab[c].size() * sizeof(TileRef), hipMemcpyHostToDevice));
    b->tt[c] = {b->d_tiles[c], (int)tab[c].size()};
  }
  // capture one RK3 step (src/cbns3d.cpp:178-240) as a graph
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (hw == 0) {
    b->dirty = false;
    return CBNS_OK;
  }
  hipStream_t s = b->stream;
  CK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const BatchGrids& G = b->grid;
  auto mem = [&](int resident) { return b->wg_mem > 0 ? b->wg_mem : resident; };
  // the stage tail (update, ghost fill) defaults to a direct launch, one tile
  // per workgroup and no queue (grid 0), with the tiles of each XCD contiguous
  auto tail = [&]() { return b->wg_mem > 0 ? b->wg_mem : 0; };
  // every persistent launch of the step takes its own block of counters, all
  // zeroed by the first node of the graph
  constexpr int LAUNCHES = 1 + 3 * 12;  // dt; per stage 3 viscous, 4 face, 1 x+y march, 3 fused z + update, 1 ghost
  constexpr int K = BATCH_CTR_PER_LAUNCH * BATCH_CTR_STRIDE;
  static_assert(LAUNCHES * K <= BATCH_COUNTERS, "counter block");
  int used = 0;
  auto ctr = [&]() { return b->d_ctr + K * used++; };
  hipError_t e = batch_launch_reset_counters(b->d_ctr, LAUNCHES * K, s);
  if (e == hipSuccess) e = batch_launch_dt(b->d_blocks, b->d_insts, b->tt[T_DT], mem(G.dt), ctr(), hw, s);
  if (e == hipSuccess) e = batch_launch_finalize(b->d_insts, hw, s);
  for (int stage = 1; stage <= 3 && e == hipSuccess; stage++) {
    // the fused z pass keeps its stage input unwritten: Q -> S -> Q_p -> Q
    // (S: the third state), so Q_p holds the stage-2 state after each step
    const int in = stage == 1 ? CBNS_Q : (stage == 2 ? 2 : CBNS_QP);
    const int out = stage == 1 ? 2 : (stage == 2 ? CBNS_QP : CBNS_Q);
    const bool ov = b->visc_side_wgs > 0;
    hipStream_t sv = ov ? b->side : s;
    if (ov) {  // fork: the viscous kernels run beside the face sweeps
      e = hipEventRecord(b->ev_fork, s);
      if (e == hipSuccess) e = hipStreamWaitEvent(sv, b->ev_fork, 0);
    }
    for (int f = 0; f < 3 && e == hipSuccess; f++)
      e = batch_launch_visc(b->d_blocks, b->d_insts, b->tt[VISC_CLASS[f]], b->jb, FORM_PREC[f], in,
                            ov ? b->visc_side_wgs : G.visc[f], ctr(), sv);
    if (ov && e == hipSuccess) e = hipEventRecord(b->ev_join, sv);
    for (int d = 0; d < 2 && e == hipSuccess; d++)
      for (int f = 0; f < 2 && e == hipSuccess; f++)
        e = batch_launch_face(b->d_blocks, b->d_insts, b->tt[FACE_CLASS[f] + d], d, in, b->jb, f, G.face[f][d], ctr(),
                              s);
    // join before the tolerance-bound solves' x+y march (which reads their
    // viscous divergence) and the stage tail
    if (ov && e == hipSuccess) e = hipStreamWaitEvent(s, b->ev_join, 0);
    if (b->has_fast && e == hipSuccess)
      e = batch_launch_ymarch(b->d_blocks, b->d_insts, b->tt[T_YMF], in, tail(), ctr(), s);
    for (int f = 0; f < 3 && e == hipSuccess; f++)
      e = batch_launch_zupdate(b->d_blocks, b->d_insts, b->tt[UPD_CLASS[f]], FORM_PREC[f], in, out, stage, tail(),
                               ctr(), b->cfix, s);
    if (b->has_f64 && e == hipSuccess)
      e = batch_launch_zfixup(b->d_blocks, b->d_insts, (int)all.size(), in, out, stage, b->cfix, s);
    if (e == hipSuccess)
      e = batch_launch_ghost(b->d_blocks, b->d_insts, b->tt[T_GHOST], out, tail(), ctr(), s);
  }
  if (used > LAUNCHES) e = hipErrorInvalidValue;  // a launch without a zeroed counter block
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(s, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return CBNS_ERR_HIP;
  }
  b->graph = g;
  CK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->dirty = false;
  return CBNS_OK;
}

extern "C" {

int cbns_batch_create(uint64_t hbm_budget_bytes, int capacity, cbns_batch** out) {
  if (!out || capacity < 1) return CBNS_ERR_ARG;
  *out = nullptr;
  int n = 0;
  cbns_device_count(&n);
  if (n == 0) return CBNS_ERR_NODEV;
  cbns_batch* b = new (std::nothrow) cbns_batch();
  if (!b) return CBNS_ERR_NOMEM;
  b->budget = hbm_budget_bytes;
  b->cap = capacity;
  b->inst.resize(capacity);
  int st = CBNS_OK;
  if (hipMalloc(&b->d_insts, sizeof(InstDesc) * capacity) != hipSuccess ||
      hipMalloc(&b->d_part, sizeof(double) * (2 * GSUM_WG + 2)) != hipSuccess ||
      hipMalloc(&b->d_ctr, sizeof(int) * BATCH_COUNTERS) != hipSuccess)
    st = CBNS_ERR_NOMEM;
  void* cfixbuf = nullptr;
  if (!st && hipMalloc(&cfixbuf, FIX_BYTES) != hipSuccess) st = CBNS_ERR_NOMEM;
  if (!st && hipMemset(cfixbuf, 0, 256) != hipSuccess) st = CBNS_ERR_HIP;
  if (cfixbuf) b->cfix = fixlist_at(cfixbuf);
  if (!st && hipHostMalloc((void**)&b->h_status, sizeof(InstStatus) * capacity, hipHostMallocMapped) != hipSuccess)
    st = CBNS_ERR_NOMEM;
  if (!st && hipHostGetDevicePointer((void**)&b->d_status, b->h_status, 0) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st) std::memset(b->h_status, 0, sizeof(InstStatus) * capacity);
  if (!st && hipMemset(b->d_insts, 0, sizeof(InstDesc) * capacity) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st && batch_grids(&b->grid) != hipSuccess) st = CBNS_ERR_HIP;
  if (!st) b->visc_side_wgs = b->grid.cus;
  if (!st && (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
              hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&b->ev_run, hipEventDisableTiming) != hipSuccess))
    st = CBNS_ERR_HIP;
  if (st) {
    cbns_batch_destroy(b);
    return st;
  }
  *out = b;
  return CBNS_OK;
}

int cbns_batch_destroy(cbns_batch* b) {
  if (!b) return CBNS_ERR_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (Inst& I : b->inst) free_inst(b, I);
  drain_pool(b);
  if (b->d_stage) (void)hipFree(b->d_stage);
  if (b->ev_run) (void)hipEventDestroy(b->ev_run);
  if (b->h_status) (void)hipHostFree(b->h_status);
  for (int c = 0; c < T_COUNT; c++)
    if (b->d_tiles[c]) (void)hipFree(b->d_tiles[c]);
  if (b->d_blocks) (void)hipFree(b->d_blocks);
  if (b->d_insts) (void)hipFree(b->d_insts);
  if (b->d_part) (void)hipFree(b->d_part);
  if (b->d_ctr) (void)hipFree(b->d_ctr);
  if (b->cfix.count) (void)hipFree(b->cfix.count);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->d_pairs) (void)hipFree(b->d_pairs);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return CBNS_OK;
}

int cbns_batch_bytes_needed(const cbns_params* p, const int dims[3], uint64_t* bytes) {
  if (!p || !dims || !bytes) return CBNS_ERR_ARG;
  *bytes = solve_bytes(p, dims);
  return CBNS_OK;
}

int cbns_batch_memory(const cbns_batch* b, uint64_t* used, uint64_t* budget) {
  if (!b) return CBNS_ERR_ARG;
  if (used) *used = b->used;
  if (budget) *budget = b->budget;
  return CBNS_OK;
}

int cbns_batch_add(cbns_batch* b, const cbns_params* p, const int dims[3], int precision, const double* Q_aos,
                   double t_cur, double mu0, int steps, int* id_out) {
  if (!b || !p || !dims || !id_out || steps < 0) return CBNS_ERR_ARG;
  if (precision != CBNS_FP64 && precision != CBNS_FP32_FLUX && precision != CBNS_FP64_FAST) return CBNS_ERR_ARG;
  const int N[3] = {(int)p->IM, (int)p->JM, (int)p->KM};
  int largest[3];
  for (int a = 0; a < 3; a++) {
    if (dims[a] < 1 || N[a] < 1) return CBNS_ERR_ARG;
    const int base = N[a] / dims[a];
    if (base < NG || N[a] - (dims[a] - 1) * base < NG) return CBNS_ERR_ARG;
    largest[a] = N[a] - (dims[a] - 1) * base;
  }
  if (!block_fits(largest[0], largest[1], largest[2])) return CBNS_ERR_ARG;  // split into more blocks
  const uint64_t need = solve_bytes(p, dims);
  if (b->used + need > b->budget) return CBNS_ERR_NOMEM;
  int id = -1;
  for (int k = 0; k < b->cap; k++)
    if (!b->inst[k].live) {
      id = k;
      break;
    }
  if (id < 0) return CBNS_ERR_NOMEM;
  CK(hipStreamSynchronize(b->stream));
  Inst& I = b->inst[id];
  I = Inst();
  I.par = *p;
  I.precision = precision;
  std::memcpy(I.dims, dims, sizeof(I.dims));
  for (int a = 0; a < 3; a++) {
    I.dom.N[a] = N[a];
    I.dom.dims[a] = dims[a];
    I.dom.base[a] = N[a] / dims[a];
  }
  I.dom.nb = dims[0] * dims[1] * dims[2];
  auto alloc = [&](size_t bytes) -> double* {
    void* q = batch_alloc(b, bytes, need);
    if (q) I.allocs.push_back({q, bytes});
    return (double*)q;
  };
  bool ok = true;
  for (int bx = 0; bx < dims[0] && ok; bx++)
    for (int by = 0; by < dims[1] && ok; by++)
      for (int bz = 0; bz < dims[2] && ok; bz++) {
        BlockDesc B;
        std::memset(&B, 0, sizeof(B));
        const int bc[3] = {bx, by, bz};
        for (int a = 0; a < 3; a++) {
          const int base = N[a] / dims[a];
          B.n[a] = (bc[a] == dims[a] - 1) ? N[a] - (dims[a] - 1) * base : base;
          B.g[a] = B.n[a] + 2 * NG;
          B.o[a] = bc[a] * base;
        }
        B.inst = id;
        B.sg = (long long)B.g[0] * B.g[1] * B.g[2];
        B.sn = (long long)B.n[0] * B.n[1] * B.n[2];
        for (int d = 0; d < 3; d++)
          B.sf[d] = (long long)(B.n[0] + (d == 0)) * (B.n[1] + (d == 1)) * (B.n[2] + (d == 2));
        set_face_divisors(B);
        ok = (B.Q[0] = alloc(8 * NEQ * B.sg)) && (B.Q[1] = alloc(8 * NEQ * B.sg)) && (B.Q[2] = alloc(8 * NEQ * B.sg)) &&
             (B.F[0] = alloc(8 * NEQ * B.sf[0])) && (B.F[1] = alloc(8 * NEQ * B.sf[1])) &&
             (B.F[2] = alloc(8 * NEQ * B.sf[2])) && (B.prim = alloc((precision == CBNS_FP32_FLUX ? 4 : 8) * 6 * B.sg)) && (B.mu = alloc(8 * B.sn)) &&
             (B.dfv = alloc(8 * 4 * B.sn));
        I.blocks.push_back(B);
      }
  if (ok) ok = (I.dblocks = (BlockDesc*)alloc(sizeof(BlockDesc) * I.blocks.size())) != nullptr;
  if (!ok) {
    free_inst(b, I);
    return CBNS_ERR_NOMEM;
  }
  I.bytes = need;
  const Phys ph = make_phys(p);
  int st = CBNS_OK;
  hipStream_t s = b->stream;
  const int nb = (int)I.blocks.size();
  if (copy_on(b, I.dblocks, I.blocks.data(), sizeof(BlockDesc) * nb, hipMemcpyHostToDevice) != hipSuccess)
    st = CBNS_ERR_HIP;
  // initial state: Taylor-Green (Q_aos == NULL) or the given interior state
  if (!st && !Q_aos) {
    const int ts = std::max(N[0], std::max(N[1], N[2]));
    std::vector<double> tab(9 * (size_t)ts, 0.0);
    const double tmp = 2.0 * 3.14159265358979323846;
    const double h[3] = {ph.dx, ph.dy, ph.dz};
    for (int a = 0; a < 3; a++)
      for (int i = 0; i < N[a]; i++) {
        const double xc = ((unsigned)i * h[a] - tmp / 2.0) + 0.5 * h[a];
        tab[(3 * a + 0) * ts + i] = std::sin(xc);
        tab[(3 * a + 1) * ts + i] = std::cos(xc);
        tab[(3 * a + 2) * ts + i] = std::cos(2.0 * xc);
      }
    double* dtab = nullptr;
    st = stage(b, sizeof(double) * tab.size(), &dtab);
    if (!st && (copy_on(b, dtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess ||
                launch_tgv(I.dblocks, I.blocks.data(), nb, dtab, ts, ph, p->p_inf, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
      st = CBNS_ERR_HIP;
    t_cur = p->t_cur;
  } else if (!st) {
    const size_t n = (size_t)NEQ * N[0] * N[1] * N[2];
    double* d = nullptr;
    st = stage(b, n * sizeof(double), &d);
    if (!st && (copy_on(b, d, Q_aos, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                launch_scatter_aos(I.dblocks, I.blocks.data(), I.dom, CBNS_Q, d, s) != hipSuccess ||
                launch_fill_mu(I.dblocks, I.blocks.data(), nb, mu0, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
      st = CBNS_ERR_HIP;
  }
  // boundary_condition, calc_primitive, init_Q_p (src/cbns3d.cpp:149-153)
  if (!st && (launch_ghost_fill(I.dblocks, I.blocks.data(), I.dom, CBNS_Q, s) != hipSuccess ||
              launch_primitive(I.dblocks, I.blocks.data(), nb, CBNS_Q, ph, precision, s) != hipSuccess ||
              launch_copy_state(I.dblocks, I.blocks.data(), nb, CBNS_Q, CBNS_QP, s) != hipSuccess))
    st = CBNS_ERR_HIP;
  InstDesc D;
  std::memset(&D, 0, sizeof(D));
  D.ph = ph;
  D.ref = RefState{1.0, 0.0, 0.0};
  if (!st && precision == CBNS_FP32_FLUX) {
    double sums[2];
    if (launch_global_sums(I.dblocks, I.dom, ph, 0, b->d_part, b->d_part + 2 * GSUM_WG, precision, s) != hipSuccess ||
        hipMemcpyAsync(sums, b->d_part + 2 * GSUM_WG, sizeof(sums), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      st = CBNS_ERR_HIP;
    const double n = (double)N[0] * N[1] * N[2];
    D.ref.rho = sums[0] / n;
    D.ref.E = sums[1] / n;
    D.ref.p = ph.gam1 * D.ref.E;
  }
  if (st) {
    free_inst(b, I);
    return st;
  }
  D.dom = I.dom;
  D.nblk = nb;
  D.precision = precision;
  D.pairs = 1;  // every solve's stage tail reduces the next step's smax
  D.active = 1;
  D.steps_target = steps;
  D.budget = 0x7fffffff;  // unbounded until the host sets budgets
  D.t_cur = t_cur;
  D.t_end = p->t_end;
  if (copy_on(b, b->d_insts + id, &D, sizeof(D), hipMemcpyHostToDevice) != hipSuccess) {
    free_inst(b, I);
    return CBNS_ERR_HIP;
  }
  I.live = true;
  b->used += need;
  b->dirty = true;
  b->published = false;
  *id_out = id;
  return CBNS_OK;
}

int cbns_batch_remove(cbns_batch* b, int id) {
  if (!b || id < 0 || id >= b->cap || !b->inst[id].live) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(b->stream));
  // stream-ordered clear of its active flag (no synchronous copy)
  CK(hipMemsetAsync((char*)(b->d_insts + id) + offsetof(InstDesc, active), 0, sizeof(int), b->stream));
  b->used -= b->inst[id].bytes;
  free_inst(b, b->inst[id]);
  b->dirty = true;
  return CBNS_OK;
}

int cbns_batch_run(cbns_batch* b, int nsteps) {
  if (!b || nsteps < 0) return CBNS_ERR_ARG;
  if (b->dirty) {
    int st = rebuild(b);
    if (st) return st;
  }
  if (!b->exec) return CBNS_OK;
  for (int n = 0; n < nsteps; n++) CK(hipGraphLaunch(b->exec, b->stream));
  CK(batch_launch_publish(b->d_insts, b->d_status, b->ninst_hw, b->stream));
  CK(hipEventRecord(b->ev_run, b->stream));
  b->published = true;
  return CBNS_OK;
}

int cbns_batch_pooled(const cbns_batch* b, uint64_t* bytes) {
  if (!b || !bytes) return CBNS_ERR_ARG;
  *bytes = b->pooled;
  return CBNS_OK;
}

int cbns_batch_trim(cbns_batch* b) {
  if (!b) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(b->stream));
  drain_pool(b);
  return CBNS_OK;
}

int cbns_batch_query(cbns_batch* b, int* busy) {
  if (!b || !busy) return CBNS_ERR_ARG;
  const hipError_t e = hipEventQuery(b->ev_run);
  if (e == hipSuccess) *busy = 0;
  else if (e == hipErrorNotReady) *busy = 1;
  else return CBNS_ERR_HIP;
  return CBNS_OK;
}

int cbns_batch_set_budgets(cbns_batch* b, const int* ids, const int* steps, int n) {
  if (!b || n < 0 || (n > 0 && (!ids || !steps))) return CBNS_ERR_ARG;
  b->published = false;
  std::vector<int2> pr(n);
  for (int k = 0; k < n; k++) {
    if (ids[k] < 0 || ids[k] >= b->cap || !b->inst[ids[k]].live || steps[k] < 0) return CBNS_ERR_ARG;
    pr[k] = make_int2(ids[k], steps[k]);
  }
  if (n == 0) return CBNS_OK;
  CK(hipStreamSynchronize(b->stream));
  int st = grow((void**)&b->d_pairs, &b->pairs_cap, n, sizeof(int2));
  if (st) return st;
  CK(copy_on(b, b->d_pairs, pr.data(), sizeof(int2) * n, hipMemcpyHostToDevice));
  CK(batch_launch_set_budget(b->d_insts, b->d_pairs, n, b->stream));
  return CBNS_OK;
}

int cbns_batch_set_workgroups(cbns_batch* b, int wgs) {
  if (!b || wgs < 0 || wgs > 65535) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(b->stream));
  b->wg_mem = wgs;
  b->dirty = true;  // the step graph holds the grid size
  return CBNS_OK;
}

int cbns_batch_set_band(cbns_batch* b, int rows) {
  if (!b || rows < 1 || rows > 4096) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(b->stream));
  b->jb = rows;
  b->dirty = true;  // the step graph holds the band
  return CBNS_OK;
}

int cbns_batch_set_overlap(cbns_batch* b, int viscous_workgroups) {
  if (!b || viscous_workgroups < 0 || viscous_workgroups > 65535) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(b->stream));
  b->visc_side_wgs = viscous_workgroups;
  b->dirty = true;  // the step graph holds the schedule
  return CBNS_OK;
}

int cbns_batch_sync(cbns_batch* b) {
  if (!b) return CBNS_ERR_ARG;
  CK(hipStreamSynchronize(b->stream));
  return CBNS_OK;
}

int cbns_batch_status(cbns_batch* b, int id, int* active, int* steps_done, double* t_cur, int* nonfinite) {
  if (!b || id < 0 || id >= b->cap || !b->inst[id].live) return CBNS_ERR_ARG;
  InstDesc D;
  CK(hipStreamSynchronize(b->stream));
  CK(copy_on(b, &D, b->d_insts + id, sizeof(D), hipMemcpyDeviceToHost));
  if (active) *active = D.active;
  if (steps_done) *steps_done = D.steps_done;
  if (t_cur) *t_cur = D.t_cur;
  if (nonfinite) *nonfinite = D.nonfinite;
  return CBNS_OK;
}

int cbns_batch_poll(cbns_batch* b, int* ids, int max_ids, int* n_done, int* n_live) {
  return cbns_batch_poll_status(b, ids, nullptr, nullptr, nullptr, max_ids, n_done, n_live);
}

int cbns_batch_poll_status(cbns_batch* b, int* ids, double* t_cur, int* steps_done, int* nonfinite, int max_ids,
                           int* n_done, int* n_live) {
  if (!b || (!ids && max_ids > 0)) return CBNS_ERR_ARG;
  int hw = 0;  // highest live id + 1 (only those descriptors are read)
  for (int id = 0; id < b->cap; id++)
    if (b->inst[id].live) hw = id + 1;
  std::vector<InstStatus> D(hw);
  if (b->published) {  // after the run's publish: read the host copy, no device operation
    CK(hipEventSynchronize(b->ev_run));
    std::atomic_thread_fence(std::memory_order_acquire);
    const volatile InstStatus* hs = b->h_status;
    for (int id = 0; id < hw; id++)
      D[id] = InstStatus{hs[id].t_cur, hs[id].steps_done, hs[id].steps_target, hs[id].active, hs[id].nonfinite};
  } else {
    CK(hipStreamSynchronize(b->stream));
    std::vector<InstDesc> full(hw);
    if (hw) CK(copy_on(b, full.data(), b->d_insts, sizeof(InstDesc) * hw, hipMemcpyDeviceToHost));
    for (int id = 0; id < hw; id++)
      D[id] = InstStatus{full[id].t_cur, full[id].steps_done, full[id].steps_target, full[id].active,
                         full[id].nonfinite};
  }
  int nd = 0, nl = 0;
  for (int id = 0; id < hw; id++) {
    if (!b->inst[id].live) continue;
    nl++;
    const bool done = !D[id].active || D[id].steps_done >= D[id].steps_target;
    if (done) {
      if (nd < max_ids) {
        ids[nd] = id;
        if (t_cur) t_cur[nd] = D[id].t_cur;
        if (steps_done) steps_done[nd] = D[id].steps_done;
        if (nonfinite) nonfinite[nd] = D[id].nonfinite;
      }
      nd++;
    }
  }
  if (n_done) *n_done = std::min(nd, max_ids);
  if (n_live) *n_live = nl;
  return CBNS_OK;
}

int cbns_batch_get_state(cbns_batch* b, int id, double* Q_aos) {
  if (!b || !Q_aos || id < 0 || id >= b->cap || !b->inst[id].live) return CBNS_ERR_ARG;
  Inst& I = b->inst[id];
  const size_t n = (size_t)NEQ * I.dom.N[0] * I.dom.N[1] * I.dom.N[2];
  CK(hipStreamSynchronize(b->stream));
  double* d = nullptr;
  int st = stage(b, n * sizeof(double), &d);
  if (st) return st;
  if (launch_gather_aos(I.dblocks, I.blocks.data(), I.dom, CBNS_Q, d, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess ||
      copy_on(b, Q_aos, d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    st = CBNS_ERR_HIP;
  return st;
}

}  // extern "C"
