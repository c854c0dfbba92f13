// Block3d facade over the C ABI (reference class: include/Block3d.h:14-201).
//
// Same public method names, argument meanings and call order as the
// reference, so the reference's driver sequence runs unchanged.  One
// difference of scope: the reference holds one block per MPI rank; here one
// Block3d holds the whole periodic domain as proc_info->dims on-device blocks
// (the reference's rank decomposition, src/cbns3d.cpp:102-126), all resident
// on one GPU.  State arrays live on the device; get_Q()/get_Q_p() return
// handles that the methods accept where the reference passes Q or Q_p.
#ifndef CBNS_HOST_BLOCK3D_H
#define CBNS_HOST_BLOCK3D_H

#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "process_info.h"
#include "simulation_params.h"

class Block3d {
  ProcessInfo* proc_info;
  SimulationParams* sim_pars;
  static const size_type NEQ{constant::NEQ};
  static const size_type NG{constant::NG};
  static const size_type NGV{constant::NGV};

  // whole-domain sizes
  size_type IM, JM, KM;

  value_type dx, dy, dz;

  cbns_ctx* ctx = nullptr;
  int status = 0;  // last C-ABI status
  bool mu_sidecar = false;

  // asynchronous output: writer pool, its job queue and the snapshot's state
  bool async_io = false;
  long epoch = 0;        // bumped by every call that may change Q or the primitives
  long snap_epoch = -1;  // the epoch the context's snapshot holds
  std::vector<std::thread> io_workers;
  std::deque<std::function<int()>> io_jobs;
  std::mutex io_m;
  std::condition_variable io_cv, io_idle_cv;
  int io_busy = 0;
  bool io_stop = false;
  int io_err = 0;
  void io_start(int threads);
  void io_stop_all();
  void io_drain();
  int io_take_error();
  int io_snapshot();
  void io_enqueue(std::function<int()> job);
  void touch() { epoch++; }

  // handles standing for Q and Q_p in the reference's pointer-passing calls
  value_type tag_Q = 0.0, tag_Q_p = 0.0;

  int which(const value_type* Q) const;

 public:
  Block3d() = delete;
  Block3d(const Block3d&) = delete;
  Block3d& operator=(const Block3d&) = delete;
  // num_x/y/z: the whole domain; proc_info->dims: on-device blocks per axis.
  Block3d(ProcessInfo* proc_info, SimulationParams* sim_pars, const size_type num_x, const size_type num_y,
          const size_type num_z);
  ~Block3d();

  void allocate_mem();
  void free_mem();

  // Mesh of the whole domain starting at cell (i_begin, j_begin, k_begin)
  // (src/Block3d.cpp:208-235); per-block coordinates are slices of it.
  void gen_mesh(const size_type i_begin, const size_type j_begin, const size_type k_begin);

  // Taylor-Green start (src/Block3d.cpp:237-271); the conserved variables are
  // formed on the device together with it, so calc_conservative has nothing
  // left to do.
  void initial_condition();
  void calc_conservative(value_type* Q);

  value_type* get_Q();
  value_type* get_Q_p();

  // Checkpoint I/O in the reference's per-rank format
  // (src/Block3d.cpp:318-357, 964-987).  fname is the name the reference's
  // main hands its own rank: the configured name with the rank as %03d and
  // ".dat" appended (src/cbns3d.cpp:132-134); here that rank is
  // proc_info->rank.  With one block, fname is the file.  With several,
  // block b's file is fname with those rank digits replaced by b's, i.e. the
  // file the reference's rank b would use; a name without the rank suffix is
  // refused (-1).
  int read_bin(const std::string fname);
  int output_bin(const std::string fname);

  // mu sidecar (no counterpart in the reference): with it on, output_bin also
  // writes each block's interior mu to "<checkpoint file>.mu" (t_cur, the
  // interior sizes as three size_type, then KM x JM x IM doubles).  read_bin
  // loads a block's sidecar whenever that file exists, so the resumed run's
  // first dt uses the mu of the uninterrupted run and the whole resumed run
  // is bit-identical to it; without a sidecar mu is zero at the first dt, as
  // in the reference's restart path.
  void set_mu_sidecar(bool on) { mu_sidecar = on; }

  // Asynchronous file output (no counterpart in the reference, whose writers
  // format on the calling rank's thread while the run waits).  With it on,
  // output_bin and output_vtr take a device snapshot of the state (one per
  // state: a checkpoint and a VTK output of the same step share it) and
  // return at once; `threads` writer threads write the files from the
  // snapshot's pinned host copy while the next steps run.  The files are
  // those of the synchronous writers, byte for byte.  A writer's failure is
  // returned (-1) by the next output call or by io_wait(), which waits for
  // every pending write; the destructor waits too.
  void set_async_io(bool on, int threads = 4);
  int io_wait();

  friend void boundary_condition(Block3d* blk, value_type* Q);

  void calc_primitive(const value_type* Q);
  void init_Q_p();

  // VTK XML output (src/Block3d.cpp:396-469): of this rank's block
  // (proc_info->rank, as the reference's per-rank call), or of block b.
  int output_vtr(const std::string fname);
  int output_vtr(const std::string fname, int b);
  int output_vtm(const std::string fname, const std::string vtr_fname, const size_type num_blocks);

  bool is_finished();
  value_type calc_dt();

  void rec_riemann_x(const value_type* Q);
  void rec_riemann_y(const value_type* Q);
  void rec_riemann_z(const value_type* Q);

  // calc_viscous_terms, calc_viscous_flux and the divergence run as one
  // fused device pass; the first two are kept as names for callers that
  // invoke the reference's sequence, and do their work in the third.
  void calc_viscous_terms();
  void calc_viscous_flux();
  void calc_viscous_flux_contribution();

  void update_rk3(const value_type dt, const size_type stage);

  // Fused step of the driver loop on the device (src/cbns3d.cpp:186-240).
  int step(value_type* dt_out);

  int num_blocks() const;
  int last_status() const { return status; }
  cbns_ctx* context() { return ctx; }

 private:
  std::vector<value_type> x, y, z;  // cell edges of the whole domain
};

void boundary_condition(Block3d* blk, value_type* Q);

#endif
