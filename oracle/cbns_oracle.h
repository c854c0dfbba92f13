/*
 * CPU FP64 oracle for the cbns3d_cart_mpi right-hand-side and time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a literal C restatement of the
 * reference solver, used as the parity checker for the HIP path (tests,
 * the smoke check and the CPU-baseline leg of the benchmark).  The product
 * path never calls it.
 *
 * Layouts follow the reference exactly (include/Block3d.h:169-199):
 *   Q, Q_p  AoS [k][j][i][eq] over the ghosted box (NG = 4 each side)
 *   prims   SoA [k][j][i]     over the ghosted box
 *   grads   SoA over the (N+2*NGV)^3 box, NGV = 2
 *   Ep      [k][j][i][eq] with i over IM+1 faces (Fp, Gp analogously)
 *   dfv     [k][j][i][eq-1] over the interior
 *
 * One block covers the whole periodic domain; the reference's MPI halo
 * exchange is restated as its closed form (every ghost = its periodic
 * image), which is what the three-phase exchange produces.
 */
#ifndef CBNS_ORACLE_H
#define CBNS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* SimulationParams (include/simulation_params.h:15-53) without the string. */
typedef struct or_params {
  double CFL, t_cur, t_end, Pr, gamma, Mach, Re, T_inf, L;
  int IM, JM, KM, nstep_max, checkpoint_freq, num_output, is_restart;
  /* derived by or_params_derive (src/read_input.cpp:78-90) */
  double Mach2, gM2, Re_inv, p_inf, gam1, gam1_inv, Pr_inv, C_T_inf, C_dt_v;
} or_params;

void or_params_derive(or_params* p);

/* Point functions (general face normal). */
double or_weno5(double Um2, double Um1, double U0, double Up1, double Up2);
void or_rec_weno5(const double* f1, const double* f2, const double* f3,
                  const double* f4, const double* f5, const double* R,
                  const double* L, double gam1, double* out5);
void or_calc_eigenvectors(double gamma, const double* primL, const double* primR,
                          double nx, double ny, double nz, double* R, double* L);
void or_lf_flux(double gamma, const double* primR, const double* primL,
                double nx, double ny, double nz, double* flux);

typedef struct or_block or_block;

or_block* or_create(const or_params* p);
void or_destroy(or_block* b);
int or_dims(const or_block* b, int* dims6); /* IM JM KM IM_G JM_G KM_G */

/* Setup (src/cbns3d.cpp:136-153): TGV IC + conservative, or a given state. */
void or_init_tgv(or_block* b);
/* Q_ghosted: full AoS ghosted array (checkpoint layout); mu interior := mu0
 * (0.0 reproduces the restart path's zero-filled mu). */
void or_set_state(or_block* b, const double* Q_ghosted, double t_cur, double mu0);

/* Hot-path methods, same meaning as the reference's (src/Block3d.cpp). */
void or_boundary_condition(or_block* b, int which);     /* 0 = Q, 1 = Q_p */
void or_calc_primitive(or_block* b, int which);
void or_init_Q_p(or_block* b);
double or_calc_dt(or_block* b);                          /* local CFL/s_max */
void or_rec_riemann(or_block* b, int which, int dir);   /* dir 0,1,2 */
void or_calc_viscous_flux_contribution(or_block* b);
void or_update_rk3(or_block* b, double dt, int stage);

/* One time step as the driver does it (src/cbns3d.cpp:178-240).
 * Returns 1 if finished before stepping, else 0; *dt_out = dt used. */
int or_step(or_block* b, double* dt_out);
double or_t_cur(const or_block* b);

/* Array access (pointers into the oracle's own storage). */
double* or_Q(or_block* b, int which);
double* or_prim(or_block* b, int field); /* 0 rho 1 u 2 v 3 w 4 p 5 T 6 mu */
double* or_face_flux(or_block* b, int dir);
double* or_dfv(or_block* b);

/* Interior copy of Q (AoS [k][j][i][eq], no ghosts). */
void or_get_interior(or_block* b, int which, double* out);

/* Diagnostics: interior means of 0.5 rho |u|^2 and 0.5 rho |omega|^2, with
 * omega from the reference's 4th-order central stencil (calc_gradient). */
void or_diagnostics(or_block* b, double* ke, double* enstrophy);

void or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
