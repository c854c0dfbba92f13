// Batched form: many independent solves (instances) of uneven sizes advanced
// together.  Each stage kernel is one persistent launch over a packed table of
// work tiles drawn from every live instance's blocks; dt, t and the active
// flag of every instance live on the device, so a whole RK3 step is a fixed
// launch sequence that is captured once as a hipGraph and replayed.
#pragma once
#include "cbns_face.h"
#include "cbns_kernels.h"

namespace cbns {

struct InstDesc {
  Phys ph;
  RefState ref;
  Domain dom;
  int block0, nblk;
  int precision;
  int active;        // 1 while stepping; 0 when idle, finished or failed
  int nonfinite;     // set by the dt reduction, ends the solve
  int steps_done, steps_target;
  int budget;        // steps it may still take before the host grants more
  int stepping;      // whether the current step advances it (set by finalize)
  int pairs;         // the stage tail reduces the next step's smax (every solve: the fused z pass)
  int s_fused;       // smax was formed by the last stage-3 update (kb_dt skips it)
  double t_cur, t_end, dt;
  unsigned long long smax;  // dt reduction target, reset after each use
};

// Per-solve status published to pinned host memory after each run, so a
// server can poll a finished lane without a device copy (which could queue
// behind another lane's work on a shared hardware queue).
struct InstStatus {
  double t_cur;
  int steps_done, steps_target, active, nonfinite;
};

// One work tile: 256 consecutive work items of one block.
struct TileRef {
  int block, tile;
};

struct TileTable {
  const TileRef* tiles;
  int n;
};

// Dynamic tile queues: each persistent launch of a step graph owns one
// counter per XCD, each on its own line.
constexpr int BATCH_CTR_PER_LAUNCH = 8;  // one tile counter per XCD
constexpr int BATCH_CTR_STRIDE = 64;     // ints between counters: one 256-byte line each
constexpr int BATCH_COUNTERS = 64 * BATCH_CTR_PER_LAUNCH * BATCH_CTR_STRIDE;  // of one step graph

// Persistent grid of each kernel: its resident workgroups on the device.
struct BatchGrids {
  int face[2][3];  // [form: FP64, FP32-flux][direction]
  int visc[3], update[3], ghost, dt;  // [FP64, FP32-flux, tolerance-bound FP64]
  int ymarch;  // the tolerance-bound solves' x+y march
  int cus;  // compute units of the device
};
hipError_t batch_grids(BatchGrids* g);

// wgs: the launch's persistent workgroups; ctr: its NXCD tile counters,
// zero at launch (batch_launch_reset_counters earlier on the stream)
// form: 0 FP64, 1 FP32-flux (the tolerance-bound solves march their x and
// y faces: batch_launch_ymarch)
hipError_t batch_launch_face(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int dir, int which, int jb,
                             int form, int wgs, int* ctr, hipStream_t s);
// The tolerance-bound solves' x and y faces: the march along y
// (xyrun_fast) that stores each cell's x- plus y-flux differences less its
// viscous divergence (after their viscous pass), one (column tile, y-run)
// unit per table entry.
hipError_t batch_launch_ymarch(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int which, int wgs, int* ctr,
                               hipStream_t s);
// precision: PREC_FP64, PREC_FP32_FLUX or PREC_FP64_FAST (whose march forms
// u, v, w, T from the stage input state `which`)
hipError_t batch_launch_visc(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int jb, int precision,
                             int which, int wgs, int* ctr, hipStream_t s);
// Stage tail of the solves of one form (PREC_FP64, PREC_FP32_FLUX or
// PREC_FP64_FAST): the z-face sweep fused with the update, from Q[in] into
// Q[out]; FP64 solves' cells beside failing z-faces go to cfx.
hipError_t batch_launch_zupdate(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int precision, int in,
                                int out, int stage, int wgs, int* ctr, const FixList& cfx, hipStream_t s);
// those cells (their faces in the reference-order form, then their update); resets cfx
hipError_t batch_launch_zfixup(const BlockDesc* blocks, InstDesc* insts, int nb, int in, int out, int stage,
                               const FixList& cfx, hipStream_t s);
hipError_t batch_launch_ghost(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int which, int wgs,
                              int* ctr, hipStream_t s);
hipError_t batch_launch_dt(const BlockDesc* blocks, InstDesc* insts, TileTable tt, int wgs, int* ctr, int ninst,
                           hipStream_t s);
// zeroes n counters with a kernel (the step graph holds no memset nodes)
hipError_t batch_launch_reset_counters(int* ctr, int n, hipStream_t s);
hipError_t batch_launch_finalize(InstDesc* insts, int ninst, hipStream_t s);
// copies every descriptor's status to host-mapped memory (after a run)
hipError_t batch_launch_publish(const InstDesc* insts, InstStatus* status, int ninst, hipStream_t s);
hipError_t batch_launch_set_budget(InstDesc* insts, const int2* pairs, int n, hipStream_t s);

}  // namespace cbns
