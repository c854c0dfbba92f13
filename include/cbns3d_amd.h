/*
 * cbns3d_amd — C ABI of the MI355X-native right-hand-side and time-step path
 * of the cbns3d_cart_mpi compressible Navier-Stokes solver.
 *
 * Plain C types only (no HIP or torch types).  Every call returns 0 on
 * success and a negative cbns_status on failure; nothing aborts inside the
 * library.  One context is bound to one device and one HIP stream; calls on
 * a context are stream-ordered and must come from one host thread at a time.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   SimulationParams ............ include/simulation_params.h:15-53
 *   read_input derivations ...... src/read_input.cpp:78-90
 *   Block3d ctor / gen_mesh ..... src/Block3d.cpp:11-33, 208-235
 *   initial_condition +
 *     calc_conservative ......... src/Block3d.cpp:237-300
 *   read_bin / output_bin ....... src/Block3d.cpp:318-357, 964-987
 *   boundary_condition .......... src/boundary_condition.cpp:4-231
 *   calc_primitive .............. src/Block3d.cpp:359-385
 *   init_Q_p .................... src/Block3d.cpp:387-394
 *   is_finished ................. src/Block3d.cpp:502-513
 *   calc_dt + MPI_Allreduce MIN . src/Block3d.cpp:515-553, src/cbns3d.cpp:186-194
 *   rec_riemann_x/y/z ........... src/Block3d.cpp:555-748
 *   calc_viscous_flux_contribution src/Block3d.cpp:750-918
 *   update_rk3 .................. src/Block3d.cpp:920-962
 *   time loop body .............. src/cbns3d.cpp:176-240
 */
#ifndef CBNS3D_AMD_H
#define CBNS3D_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CBNS_NEQ 5 /* conserved variables (include/constants.h:13) */
#define CBNS_NG 4  /* ghost layers (include/constants.h:15-17) */
#define CBNS_NGV 2 /* viscous halo (include/constants.h:16) */

typedef enum cbns_status {
  CBNS_OK = 0,
  CBNS_ERR_ARG = -1,      /* invalid argument (null pointer, bad index, bad size) */
  CBNS_ERR_HIP = -2,      /* a HIP runtime call failed */
  CBNS_ERR_NOMEM = -3,    /* device allocation failed */
  CBNS_ERR_NODEV = -4,    /* no usable gfx950 device */
  CBNS_ERR_STATE = -5,    /* call out of order (e.g. stepping before a state is set) */
  CBNS_ERR_IO = -6,       /* file could not be opened, read or written */
  CBNS_ERR_FORMAT = -7,   /* checkpoint dimensions do not match (src/Block3d.cpp:334-338) */
  CBNS_ERR_NONFINITE = -8, /* a non-finite value met the time-step reduction */
  CBNS_ERR_CAPACITY = -9   /* request larger than a fixed capacity (the no-allocation entry point's scratch) */
} cbns_status;

/* SimulationParams (include/simulation_params.h:15-53) as plain data, with the
 * checkpoint name left to the host facade.  The first group is what input.txt
 * holds (src/read_input.cpp:24-72); the second is derived by cbns_params_derive. */
typedef struct cbns_params {
  double CFL, t_cur, t_end, Pr, gamma, Mach, Re, T_inf, L;
  uint32_t IM, JM, KM, nstep_max, checkpoint_freq, num_output;
  int32_t is_restart;
  double Mach2, gM2, Re_inv, p_inf, gam1, gam1_inv, Pr_inv, C_T_inf, C_dt_v;
} cbns_params;

/* Arithmetic of the inviscid face flux. */
typedef enum cbns_precision {
  CBNS_FP64 = 0,      /* the plain path: every operation in FP64, rounded as the
                         reference's (bit-identical results) */
  CBNS_FP32_FLUX = 1, /* WENO, characteristic projections and LF flux in FP32 on
                         deviations from a reference state; state, Roe averages,
                         viscous terms, RK3 and reductions in FP64 */
  CBNS_FP64_FMA = 2,  /* the plain path with the face fluxes' a * b + c fused
                         into FMAs, quotients and square roots without range
                         scaling: FP64
                         throughout, within a few roundings of CBNS_FP64 per
                         face; single contexts only (not the batched form) */
  CBNS_FP64_FAST = 3  /* tolerance-bound FP64: the face fluxes' arithmetic
                         restructured (characteristic projections in factored
                         form, one division per WENO5 reconstruction, shared
                         reciprocals, FMA contraction) and formed from the
                         state alone; FP64 throughout, everything else as
                         CBNS_FP64; within 1e-12 relative of the reference
                         after one step */
} cbns_precision;

typedef struct cbns_ctx cbns_ctx;

/* Which state array a call addresses: the reference's Q or Q_p. */
#define CBNS_Q 0
#define CBNS_QP 1

/* src/read_input.cpp:78-90: Mach2, Re_inv, Pr_inv, gam1, gM2, gam1_inv, p_inf,
 * C_T_inf, C_dt_v from the input values. */
int cbns_params_derive(cbns_params* p);

/* Number of visible devices that this library can run on (gfx950). */
int cbns_device_count(int* n);

/* Create a context on the current device: the periodic IMxJMxKM domain split
 * into dims[0] x dims[1] x dims[2] on-device blocks exactly as the reference
 * splits it over MPI ranks (src/cbns3d.cpp:102-126: N/dims cells per block, the
 * last block takes the remainder).  Blocks are numbered like the reference's
 * Cartesian ranks (row-major, z fastest).  Every block needs >= 4 cells per
 * axis (the reference's halo depth) and at most 2^29 cells with its ghost
 * layers, (n0 + 8)(n1 + 8)(n2 + 8) <= 2^29 (about 800^3: each field of a
 * block stays below 4 GB); a larger grid takes more blocks.  CBNS_ERR_ARG
 * otherwise.  No state is set yet. */
int cbns_create(const cbns_params* p, const int dims[3], int precision, cbns_ctx** out);
int cbns_destroy(cbns_ctx* ctx);

int cbns_num_blocks(const cbns_ctx* ctx, int* nb);
/* Interior size and global origin of block b. */
int cbns_block_info(const cbns_ctx* ctx, int b, int n[3], int origin[3]);
/* Device memory held by the context, in bytes. */
int cbns_memory_bytes(const cbns_ctx* ctx, uint64_t* bytes);

/* Taylor-Green start (src/cbns3d.cpp:136-153 with is_restart = 1):
 * initial_condition + calc_conservative, boundary_condition, calc_primitive,
 * init_Q_p.  mu := 1.0 on the interior (src/Block3d.cpp:265).  The sin/cos
 * factors are evaluated on the host with the C library, as the reference does. */
int cbns_init_tgv(cbns_ctx* ctx);

/* Set Q from a host array holding the whole periodic interior, AoS
 * [k][j][i][eq] (IMxJMxKMx5 doubles), then boundary_condition, calc_primitive
 * and init_Q_p; t_cur := t_cur.  mu on the interior := mu0 (0.0 reproduces the
 * reference's restart path, where mu is never written before the first dt). */
int cbns_set_state(cbns_ctx* ctx, const double* Q_aos, double t_cur, double mu0);
/* Copy the interior of Q (which = CBNS_Q) or Q_p (CBNS_QP) to a host AoS array. */
int cbns_get_state(cbns_ctx* ctx, int which, double* Q_aos);

/* Per-block ghosted AoS arrays in the checkpoint layout (src/Block3d.cpp:964-987:
 * [k][j][i][eq] over (KM+8)(JM+8)(IM+8)), for the Block3d facade. */
int cbns_get_block_state(cbns_ctx* ctx, int b, int which, double* Q_aos_ghosted);
int cbns_set_block_state(cbns_ctx* ctx, int b, int which, const double* Q_aos_ghosted);

/* Primitive field f (0 rho, 1 u, 2 v, 3 w, 4 p, 5 T, 6 mu) of block b, ghosted
 * SoA [k][j][i] ((KM+8)(JM+8)(IM+8) doubles); for mu only the interior is kept
 * on the device and ghosts read as 0. */
int cbns_get_block_primitive(cbns_ctx* ctx, int b, int f, double* out_ghosted);

/* Interior mu of block b, [k][j][i] (KM x JM x IM doubles of the block): the
 * Sutherland viscosity the last viscous pass stored (src/Block3d.cpp:777-779),
 * which the next calc_dt reads (src/Block3d.cpp:533-546).  The reference's
 * checkpoint (src/Block3d.cpp:964-987) does not hold it, so its restart path
 * (src/Block3d.cpp:318-357) starts from an unwritten mu; saving it beside a
 * checkpoint and setting it after cbns_set_state makes the resumed run's
 * first dt, and so the whole resumed run, bit-identical to the uninterrupted
 * run.  No counterpart in the reference. */
int cbns_get_block_mu(cbns_ctx* ctx, int b, double* mu_interior);
int cbns_set_block_mu(cbns_ctx* ctx, int b, const double* mu_interior);

/* Asynchronous snapshot for file output, so that checkpoint and VTK writes
 * (src/Block3d.cpp:396-500, 964-987) overlap the following steps instead of
 * stopping them.  No counterpart in the reference, whose writers copy and
 * format on the calling rank's thread.
 * cbns_snapshot_begin enqueues on the context's stream a device copy of Q and
 * of the primitives rho u v w p of every block as they stand after the work
 * issued so far, and a copy of it to pinned host memory on a copy stream;
 * it returns at once and later steps run meanwhile.  cbns_snapshot_wait
 * blocks until the host copy is complete and gives the snapshot's t_cur.
 * cbns_snapshot_block_state and cbns_snapshot_block_primitive (f = 0 rho,
 * 1 u, 2 v, 3 w, 4 p) give block b's arrays in the layouts of
 * cbns_get_block_state and cbns_get_block_primitive; they wait for the host
 * copy themselves and may be called from another host thread while the
 * context's own thread keeps stepping.  A snapshot stays readable until the
 * next cbns_snapshot_begin, which overwrites it: the caller finishes reading
 * one snapshot before it begins the next.  The first call allocates the
 * device and pinned host buffers (one state and five primitive fields per
 * block), which the context keeps. */
int cbns_snapshot_begin(cbns_ctx* ctx);
int cbns_snapshot_wait(cbns_ctx* ctx, double* t_cur);
int cbns_snapshot_block_state(cbns_ctx* ctx, int b, double* Q_aos_ghosted);
int cbns_snapshot_block_primitive(cbns_ctx* ctx, int b, int f, double* out_ghosted);

/* ---- method-level entry points (the reference's stage order) ----------- */
int cbns_boundary_condition(cbns_ctx* ctx, int which);  /* src/boundary_condition.cpp:4-231 */
int cbns_calc_primitive(cbns_ctx* ctx, int which);      /* src/Block3d.cpp:359-385 */
int cbns_init_Q_p(cbns_ctx* ctx);                       /* src/Block3d.cpp:387-394 */
/* Global CFL/s_max over every block = the reference's local calc_dt followed by
 * MPI_Allreduce(MIN) (src/Block3d.cpp:515-553, src/cbns3d.cpp:186-189). */
int cbns_calc_dt(cbns_ctx* ctx, double* dt);
int cbns_rec_riemann(cbns_ctx* ctx, int which, int dir); /* src/Block3d.cpp:555-748 */
int cbns_calc_viscous_flux_contribution(cbns_ctx* ctx); /* src/Block3d.cpp:870-918 */
int cbns_update_rk3(cbns_ctx* ctx, double dt, int stage); /* src/Block3d.cpp:920-962 */

/* ---- fused time stepping ----------------------------------------------- */
/* One step of the driver loop (src/cbns3d.cpp:178-240): returns 1 in *finished
 * (and does nothing) when t_cur + 1e-5 > t_end, else computes dt, clips it to
 * t_end - t_cur, advances t_cur and runs the three stages.  *dt_out may be NULL. */
int cbns_step(cbns_ctx* ctx, double* dt_out, int* finished);
/* nsteps steps; dts (may be NULL) receives each step's dt.  *steps_done = the
 * steps taken before is_finished stopped the loop. */
int cbns_run(cbns_ctx* ctx, int nsteps, double* dts, int* steps_done);

int cbns_get_time(const cbns_ctx* ctx, double* t_cur);
int cbns_set_time(cbns_ctx* ctx, double t_cur);

/* Interior means of 0.5 rho |u|^2 and 0.5 rho |omega|^2 (omega by the
 * reference's 4th-order central stencil, src/calc_gradient.cpp:29-64) from the
 * current primitives of Q.  No counterpart in the reference. */
int cbns_diagnostics(cbns_ctx* ctx, double* ke, double* enstrophy);

/* Shell spectra of the velocity (u = (rho u) / rho, v, w) of Q on the periodic
 * [-pi, pi)^3 grid, formed on the device: a batched hipFFT real-to-complex
 * 3-D transform of u, v, w (amplitudes divided by the cell count), then one
 * deterministic reduction per shell.  For shells s = 0 .. nshell - 1 (the
 * modes with rint(|k|) = s, integer wavenumbers in numpy fftfreq order):
 * ke[s] = sum 0.5 |u_hat|^2, enstrophy[s] = sum 0.5 |k x u_hat|^2,
 * dilatation[s] = sum 0.5 |k . u_hat|^2.  Any output may be NULL.  The first
 * call allocates the transform buffers (3 N^3 doubles, 3 N^2 (N/2 + 1)
 * complex) and the plan, which the context keeps.  A domain of more than
 * 2^31 - 1 cells (past hipFFT's 32-bit sizes) returns CBNS_ERR_ARG.
 * Diagnostics of the decaying-turbulence workload; no counterpart in the
 * reference. */
int cbns_spectra(cbns_ctx* ctx, int nshell, double* ke, double* enstrophy, double* dilatation);

/* Synchronise the context's stream. */
int cbns_sync(cbns_ctx* ctx);

/* Run the viscous kernels of each stage on a second stream beside the face
 * fluxes (default on).  Results are identical either way.  The
 * tolerance-bound form's face kernel reads the viscous divergence, so its
 * stages run the two in turn whatever this says. */
int cbns_set_overlap(cbns_ctx* ctx, int on);
/* With overlap on: the side-stream viscous kernel runs on this many
 * workgroups that take its tile units in turn, so that it occupies fewer of
 * the slots the face sweeps use (0, the default: one workgroup per unit).
 * A tuning knob only; results are identical. */
int cbns_set_viscous_workgroups(cbns_ctx* ctx, int workgroups);
/* Rows per band of the stencil kernels' work order (default 16): each XCD
 * walks bands of this many y-rows through all z-planes.  A tuning knob only;
 * results are identical for every value. */
int cbns_set_band(cbns_ctx* ctx, int rows);
/* Register-allocation variant of the FP64 face kernels: 2 or 3 waves per
 * SIMD (default 2).  The tolerance-bound form's stepping marches its x and y
 * faces in one kernel with a single variant, which this leaves unchanged.
 * A tuning knob only; results are identical. */
int cbns_set_face_occupancy(cbns_ctx* ctx, int waves);
/* Tile shapes of the marching kernels: at most visc_planes planes per
 * z-chunk of the viscous tile march (4 .. 256, default 32) and at most
 * run_cells cells per run of the fused z-face pass and of the
 * tolerance-bound form's x+y march (2 .. 256, default 32; the x+y march
 * takes at most 64); 0 keeps the default.  Each is halved on grids with
 * too few column tiles to give 1024 work units.  A tuning knob only; results
 * are identical for every value.  No counterpart in the reference. */
int cbns_set_tile_shape(cbns_ctx* ctx, int visc_planes, int run_cells);

/* Live kernel timing: while on, every launch made by cbns_step / cbns_run is
 * bracketed by HIP events on the context's stream and its duration is added
 * to its class.  Turning it on clears the totals.  No counterpart in the
 * reference (which times only the whole loop, src/cbns3d.cpp:174,272-286). */
#define CBNS_KCLASSES 8 /* face x, face y, face z, viscous, update, ghost fill, primitive, dt */
int cbns_set_profiling(cbns_ctx* ctx, int on);
/* The same for the kernel classes whose bit is set in mask (bit k = class k
 * above; 0 = off): each timed launch adds two events to the stream, so a
 * caller that needs one class's live time can leave the others untimed. */
int cbns_set_profiling_classes(cbns_ctx* ctx, unsigned mask);
int cbns_kernel_times(cbns_ctx* ctx, double ms_total[CBNS_KCLASSES], int64_t launches[CBNS_KCLASSES]);

/* Evaluate the device face flux of one face along axis dir on host-given data:
 * cells[6][5] conserved values of cells i-3..i+2 (AoS), prim[2][5] = rho, u, v,
 * w, p of cells i-1 and i.  Same arithmetic as cbns_rec_riemann in the given
 * precision; for parity tests of the per-face path (src/Block3d.cpp:586-612).
 * nfaces faces at once.  CBNS_FP32_FLUX evaluates the deviation form about
 * the face's own reference state (rho and E means of cells i-1 and i, at
 * rest), with the primitives rounded to float as the solver stores them, and
 * returns the flux itself (F(Ug) added back in FP64). */
int cbns_face_flux_eval(int precision, int dir, double gamma, int nfaces,
                        const double* cells, const double* prim, double* flux);

/* cbns_face_flux_eval without any device allocation, for callers that cannot
 * allocate (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results, bit for bit; the inputs and outputs are staged through one fixed
 * device array per GPU (CBNS_NOALLOC_MAX_FACES faces, 2.95 MB, held until
 * the process ends) instead of buffers allocated per call.  nfaces above
 * CBNS_NOALLOC_MAX_FACES returns CBNS_ERR_CAPACITY.  Every call on a GPU
 * uses that GPU's one array: calls that overlap on one GPU (several host
 * threads) race and must be serialised by the caller.  Synchronous: returns
 * when flux holds the results. */
#define CBNS_NOALLOC_MAX_FACES 8192
int cbns_face_flux_eval_noalloc(int precision, int dir, double gamma, int nfaces,
                                const double* cells, const double* prim, double* flux);

/* ---- batched form ------------------------------------------------------
 * Many independent solves (each with its own grid, physics, precision, step
 * count and block split) advanced together: each RK stage is one persistent
 * launch per kernel over a packed table of every live solve's blocks, and one
 * RK3 step (dt reduction, dt/t bookkeeping on the device, three stages) is
 * replayed as one hipGraph.  Solves are admitted while the device memory they
 * need fits the budget.  A solve's result is bitwise the same alone or in any
 * mix.  No counterpart in the reference (one solve per MPI job there). */
typedef struct cbns_batch cbns_batch;
int cbns_batch_create(uint64_t hbm_budget_bytes, int capacity, cbns_batch** out);
int cbns_batch_destroy(cbns_batch* b);
/* Device bytes a solve of this size and split needs in a batch. */
int cbns_batch_bytes_needed(const cbns_params* p, const int dims[3], uint64_t* bytes);
int cbns_batch_memory(const cbns_batch* b, uint64_t* used, uint64_t* budget);
/* Admit a solve: Taylor-Green start when Q_aos is NULL (t from p->t_cur), else
 * the given interior AoS state at t_cur with mu := mu0; it runs `steps` steps
 * (or until t_end).  precision: CBNS_FP64, CBNS_FP32_FLUX or CBNS_FP64_FAST.
 * Block sizes are bounded as in cbns_create (CBNS_ERR_ARG).
 * CBNS_ERR_NOMEM when the budget or capacity is full.  The arrays of removed
 * solves are kept for reuse by later solves of the same shapes (releasing
 * device memory synchronises the whole device); they are released when new
 * memory would take the batch past its budget, and by cbns_batch_destroy. */
int cbns_batch_add(cbns_batch* b, const cbns_params* p, const int dims[3], int precision, const double* Q_aos,
                   double t_cur, double mu0, int steps, int* id);
int cbns_batch_remove(cbns_batch* b, int id);
/* Replay nsteps RK3 steps of every live solve (asynchronous).  A solve with a
 * step budget (below) takes at most that many of them, then pauses. */
int cbns_batch_run(cbns_batch* b, int nsteps);
/* Grant each listed solve `steps` more steps (replacing its budget).  Solves
 * start unbounded.  Lets a scheduler give small solves many cheap steps and
 * large ones few within one run. */
int cbns_batch_set_budgets(cbns_batch* b, const int* ids, const int* steps, int n);
/* Persistent workgroups per launch of the memory-bound stage kernels (stage
 * tail, ghost gather, dt).  0, the default, launches the stage tail (update,
 * ghost fill) directly, one tile per workgroup with each XCD's tiles
 * contiguous, and gives the dt pass its resident workgroups on the device.  A
 * tuning knob only; results are identical for every value.  Takes effect at
 * the next run. */
int cbns_batch_set_workgroups(cbns_batch* b, int workgroups);
/* Stage schedule of the step graph: 0 runs the viscous-flux kernels before
 * the face sweeps; n > 0 runs them beside the sweeps on a second stream with
 * n persistent workgroups (joined before the stage tail).  The default is one
 * workgroup per compute unit.  A tuning knob only; results are identical for
 * every value.  Takes effect at the next run.  No counterpart in the
 * reference. */
int cbns_batch_set_overlap(cbns_batch* b, int viscous_workgroups);
/* Rows per band in the work order of the face sweeps and the stage tail
 * (default 32), as cbns_set_band for a single context.  A tuning knob only;
 * results are identical for every value.  Takes effect at the next run.  No
 * counterpart in the reference. */
int cbns_batch_set_band(cbns_batch* b, int rows);
int cbns_batch_sync(cbns_batch* b);
/* Whether the steps issued by the last cbns_batch_run are still running
 * (*busy = 1) or done (0), without waiting: a server polls its batches with
 * it and services only the idle ones.  No counterpart in the reference. */
int cbns_batch_query(cbns_batch* b, int* busy);
/* Device bytes held for reuse from removed solves, and their release (see
 * cbns_batch_add).  No counterpart in the reference. */
int cbns_batch_pooled(const cbns_batch* b, uint64_t* bytes);
int cbns_batch_trim(cbns_batch* b);
int cbns_batch_status(cbns_batch* b, int id, int* active, int* steps_done, double* t_cur, int* nonfinite);
/* Ids of live solves that are done (took their steps, reached t_end or failed);
 * *n_live = live solves. */
int cbns_batch_poll(cbns_batch* b, int* ids, int max_ids, int* n_done, int* n_live);
/* The same, with each done solve's t_cur, steps taken and non-finite flag
 * (any of the three arrays may be NULL), read back in one copy. */
int cbns_batch_poll_status(cbns_batch* b, int* ids, double* t_cur, int* steps_done, int* nonfinite, int max_ids,
                           int* n_done, int* n_live);
int cbns_batch_get_state(cbns_batch* b, int id, double* Q_aos);

const char* cbns_strerror(int status);

#ifdef __cplusplus
}
#endif
#endif
