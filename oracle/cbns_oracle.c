/*
 * CPU FP64 oracle — TEST INFRASTRUCTURE ONLY (the parity checker; the
 * product path never calls it).  A literal restatement of the reference's
 * right-hand-side and time-step path: every expression keeps the
 * reference's operand order and association so that, built without FMA
 * contraction, each result rounds exactly as the reference's does.
 * Each function cites the reference lines it restates.
 */
#include "cbns_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NEQ 5
#define NGI 3
#define NGV 2
#define NG 4 /* max(2*NGV, NGI), include/constants.h:15-17 */
static const double PI_ = 3.14159265358979323846; /* include/constants.h:19 */

/* src/read_input.cpp:78-90 */
void or_params_derive(or_params* p) {
  p->Mach2 = p->Mach * p->Mach;
  p->Re_inv = 1.0 / p->Re;
  p->Pr_inv = 1.0 / p->Pr;
  p->gam1 = p->gamma - 1.0;
  p->gM2 = p->gamma * p->Mach2;
  p->gam1_inv = 1.0 / p->gam1;
  p->p_inf = 1.0 / p->gM2;
  p->C_T_inf = 110.5 / p->T_inf;
  {
    double a = 4.0 / 3.0, b = p->gamma * p->Pr_inv; /* std::max(a, b) */
    p->C_dt_v = ((a < b) ? b : a) * p->Re_inv;
  }
}

/* ------------------------------------------------------------------------ */
/* Point functions                                                           */

/* src/weno.cpp:25-56 (Jiang-Shu WENO5, eps = 1e-6) */
double or_weno5(double Um2, double Um1, double U0, double Up1, double Up2) {
  const double eps = 1.0e-6;
  double diff1 = Um2 - 2.0 * Um1 + U0;
  double diff2 = Um2 - 4.0 * Um1 + 3.0 * U0;
  double diff3 = Um1 - 2.0 * U0 + Up1;
  double diff4 = Um1 - Up1;
  double diff5 = U0 - 2.0 * Up1 + Up2;
  double diff6 = 3.0 * U0 - 4.0 * Up1 + Up2;

  double IS1 = (13.0 / 12.0) * diff1 * diff1 + 0.25 * diff2 * diff2 + eps;
  double IS2 = (13.0 / 12.0) * diff3 * diff3 + 0.25 * diff4 * diff4 + eps;
  double IS3 = (13.0 / 12.0) * diff5 * diff5 + 0.25 * diff6 * diff6 + eps;

  double w1 = 0.1 / (IS1 * IS1);
  double w2 = 0.6 / (IS2 * IS2);
  double w3 = 0.3 / (IS3 * IS3);
  double sum = w1 + w2 + w3;

  double f1 = (2.0 * Um2 - 7.0 * Um1 + 11.0 * U0) / 6.0;
  double f2 = (-Um1 + 5.0 * U0 + 2.0 * Up1) / 6.0;
  double f3 = (2.0 * U0 + 5.0 * Up1 - Up2) / 6.0;
  return (w1 * f1 + w2 * f2 + w3 * f3) / sum;
}

/* src/weno.cpp:77-152.  R, L row-major 5x5; out5 = rho, u, v, w, p. */
void or_rec_weno5(const double* f1, const double* f2, const double* f3,
                  const double* f4, const double* f5, const double* R,
                  const double* L, double gam1, double* out5) {
  const double* fs[5] = {f1, f2, f3, f4, f5};
  double f_sten[NEQ][5], fc_sten[NEQ][5];
  for (int c = 0; c < 5; c++)
    for (int e = 0; e < NEQ; e++) f_sten[e][c] = fs[c][e];
  for (int li = 0; li < NEQ; li++)
    for (int lj = 0; lj < 5; lj++) {
      double ss = 0.0;
      for (int lk = 0; lk < NEQ; lk++) ss += L[5 * li + lk] * f_sten[lk][lj];
      fc_sten[li][lj] = ss;
    }
  for (int li = 0; li < NEQ; li++)
    f_sten[li][0] = or_weno5(fc_sten[li][0], fc_sten[li][1], fc_sten[li][2],
                             fc_sten[li][3], fc_sten[li][4]);
  for (int li = 0; li < NEQ; li++) {
    double ss = 0.0;
    for (int lj = 0; lj < NEQ; lj++) ss += R[5 * li + lj] * f_sten[lj][0];
    f_sten[li][1] = ss;
  }
  double rho = f_sten[0][1];
  double u = f_sten[1][1] / rho;
  double v = f_sten[2][1] / rho;
  double w = f_sten[3][1] / rho;
  out5[0] = rho;
  out5[1] = u;
  out5[2] = v;
  out5[3] = w;
  out5[4] = gam1 * (f_sten[4][1] - 0.5 * rho * (u * u + v * v + w * w));
}

/* src/eigenvectors_roe.cpp:32-171.  primX = rho, u, v, w, p. */
void or_calc_eigenvectors(double gamma, const double* primL, const double* primR,
                          double n_x, double n_y, double n_z, double* R, double* L) {
  const double gam1 = gamma - 1.0;
  const double rho_L = primL[0], u_L = primL[1], v_L = primL[2], w_L = primL[3], p_L = primL[4];
  const double rho_R = primR[0], u_R = primR[1], v_R = primR[2], w_R = primR[3], p_R = primR[4];

  double l_x = 0.0, l_y = n_z, l_z = -n_y, m_x, m_y, m_z;
  double n_abs = sqrt(l_x * l_x + l_y * l_y + l_z * l_z);
  if (n_abs > 1.0e-8) {
    m_x = -n_y * n_y - n_z * n_z;
    m_y = n_x * n_y;
    m_z = n_x * n_z;
  } else {
    l_x = -n_z;
    l_y = 0.0;
    l_z = n_x;
    m_x = n_x * n_y;
    m_y = -n_x * n_x - n_z * n_z;
    m_z = n_y * n_z;
  }
  n_abs = sqrt(n_x * n_x + n_y * n_y + n_z * n_z);
  n_x /= n_abs; n_y /= n_abs; n_z /= n_abs;
  n_abs = sqrt(l_x * l_x + l_y * l_y + l_z * l_z);
  l_x /= n_abs; l_y /= n_abs; l_z /= n_abs;
  n_abs = sqrt(m_x * m_x + m_y * m_y + m_z * m_z);
  m_x /= n_abs; m_y /= n_abs; m_z /= n_abs;

  double c_L = sqrt(gamma * p_L / rho_L);
  double H_L = c_L * c_L / gam1 + 0.5 * (u_L * u_L + v_L * v_L + w_L * w_L);
  double c_R = sqrt(gamma * p_R / rho_R);
  double H_R = c_R * c_R / gam1 + 0.5 * (u_R * u_R + v_R * v_R + w_R * w_R);

  double ratio = sqrt(rho_R / rho_L);
  double uu = (u_L + ratio * u_R) / (1.0 + ratio);
  double vv = (v_L + ratio * v_R) / (1.0 + ratio);
  double ww = (w_L + ratio * w_R) / (1.0 + ratio);
  double H = (H_L + ratio * H_R) / (1.0 + ratio);
  double q2 = uu * uu + vv * vv + ww * ww;
  double c = sqrt(gam1 * (H - 0.5 * q2));
  double qn = uu * n_x + vv * n_y + ww * n_z;
  double ql = uu * l_x + vv * l_y + ww * l_z;
  double qm = uu * m_x + vv * m_y + ww * m_z;

#define RR(a, b) R[5 * (a) + (b)]
#define LL(a, b) L[5 * (a) + (b)]
  RR(0, 0) = 1.0; RR(1, 0) = uu - c * n_x; RR(2, 0) = vv - c * n_y;
  RR(3, 0) = ww - c * n_z; RR(4, 0) = H - qn * c;
  RR(0, 1) = 1.0; RR(1, 1) = uu; RR(2, 1) = vv; RR(3, 1) = ww; RR(4, 1) = 0.5 * q2;
  RR(0, 2) = 1.0; RR(1, 2) = uu + c * n_x; RR(2, 2) = vv + c * n_y;
  RR(3, 2) = ww + c * n_z; RR(4, 2) = H + qn * c;
  RR(0, 3) = 0.0; RR(1, 3) = l_x; RR(2, 3) = l_y; RR(3, 3) = l_z; RR(4, 3) = ql;
  RR(0, 4) = 0.0; RR(1, 4) = m_x; RR(2, 4) = m_y; RR(3, 4) = m_z; RR(4, 4) = qm;

  double K_2c2 = gam1 / (2.0 * c * c);
  double one_2c = 0.5 / c;
  LL(0, 0) = 0.5 * K_2c2 * q2 + one_2c * qn;
  LL(1, 0) = 1.0 - K_2c2 * q2;
  LL(2, 0) = 0.5 * K_2c2 * q2 - one_2c * qn;
  LL(3, 0) = -ql;
  LL(4, 0) = -qm;
  LL(0, 1) = -(K_2c2 * uu + one_2c * n_x);
  LL(1, 1) = 2.0 * K_2c2 * uu;
  LL(2, 1) = -(K_2c2 * uu - one_2c * n_x);
  LL(3, 1) = l_x;
  LL(4, 1) = m_x;
  LL(0, 2) = -(K_2c2 * vv + one_2c * n_y);
  LL(1, 2) = 2.0 * K_2c2 * vv;
  LL(2, 2) = -(K_2c2 * vv - one_2c * n_y);
  LL(3, 2) = l_y;
  LL(4, 2) = m_y;
  LL(0, 3) = -(K_2c2 * ww + one_2c * n_z);
  LL(1, 3) = 2.0 * K_2c2 * ww;
  LL(2, 3) = -(K_2c2 * ww - one_2c * n_z);
  LL(3, 3) = l_z;
  LL(4, 3) = m_z;
  LL(0, 4) = K_2c2;
  LL(1, 4) = -2.0 * K_2c2;
  LL(2, 4) = K_2c2;
  LL(3, 4) = 0.0;
  LL(4, 4) = 0.0;
#undef RR
#undef LL
}

/* src/lf_flux.cpp:32-99.  Argument order as the reference: right, then left. */
void or_lf_flux(double gamma, const double* primR, const double* primL,
                double nx, double ny, double nz, double* flux) {
  const double gam1 = gamma - 1.0;
  const double rho_R = primR[0], u_R = primR[1], v_R = primR[2], w_R = primR[3], p_R = primR[4];
  const double rho_L = primL[0], u_L = primL[1], v_L = primL[2], w_L = primL[3], p_L = primL[4];

  double c_R = sqrt(gamma * p_R / rho_R);
  double H_R = c_R * c_R / gam1 + 0.5 * (u_R * u_R + v_R * v_R + w_R * w_R);
  double rE_R = p_R / gam1 + 0.5 * rho_R * (u_R * u_R + v_R * v_R + w_R * w_R);
  double c_L = sqrt(gamma * p_L / rho_L);
  double H_L = c_L * c_L / gam1 + 0.5 * (u_L * u_L + v_L * v_L + w_L * w_L);
  double rE_L = p_L / gam1 + 0.5 * rho_L * (u_L * u_L + v_L * v_L + w_L * w_L);
  double q_R = u_R * nx + v_R * ny + w_R * nz;
  double q_L = u_L * nx + v_L * ny + w_L * nz;

  double ws[3], ws_max;
  ws[0] = fabs(u_R) + c_R;
  ws_max = fabs(u_L) + c_L;
  if (ws_max > ws[0]) ws[0] = ws_max;
  ws[1] = fabs(v_R) + c_R;
  ws_max = fabs(v_L) + c_L;
  if (ws_max > ws[1]) ws[1] = ws_max;
  ws[2] = fabs(w_R) + c_R;
  ws_max = fabs(w_L) + c_L;
  if (ws_max > ws[2]) ws[2] = ws_max;
  ws_max = fabs(ws[0] * nx + ws[1] * ny + ws[2] * nz);

  double diss[NEQ], f_L[NEQ], f_R[NEQ];
  diss[0] = ws_max * (rho_R - rho_L);
  diss[1] = ws_max * (rho_R * u_R - rho_L * u_L);
  diss[2] = ws_max * (rho_R * v_R - rho_L * v_L);
  diss[3] = ws_max * (rho_R * w_R - rho_L * w_L);
  diss[4] = ws_max * (rE_R - rE_L);

  f_L[0] = rho_L * q_L;
  f_L[1] = rho_L * q_L * u_L + p_L * nx;
  f_L[2] = rho_L * q_L * v_L + p_L * ny;
  f_L[3] = rho_L * q_L * w_L + p_L * nz;
  f_L[4] = rho_L * q_L * H_L;
  f_R[0] = rho_R * q_R;
  f_R[1] = rho_R * q_R * u_R + p_R * nx;
  f_R[2] = rho_R * q_R * v_R + p_R * ny;
  f_R[3] = rho_R * q_R * w_R + p_R * nz;
  f_R[4] = rho_R * q_R * H_R;
  for (int li = 0; li < NEQ; li++) flux[li] = 0.5 * (f_L[li] + f_R[li] - diss[li]);
}

/* ------------------------------------------------------------------------ */
/* Block                                                                     */

struct or_block {
  or_params par;
  int IM, JM, KM, IM_G, JM_G, KM_G, IM_GV, JM_GV, KM_GV;
  double dx, dy, dz;
  double *x, *y, *z;
  double *rho, *u, *v, *w, *p, *T, *mu;
  double *ux, *vx, *wx, *uy, *vy, *wy, *uz, *vz, *wz, *Tx, *Ty, *Tz;
  double *txx, *tyy, *tzz, *txy, *txz, *tyz, *qx, *qy, *qz;
  double *Q, *Q_p, *Ev, *Fv, *Gv, *dfv, *Ep, *Fp, *Gp;
};

/* include/Block3d.h:169-199 */
static inline long idx_g(const or_block* b, int i, int j, int k) {
  return (NG + i) + (long)b->IM_G * ((NG + j) + (long)b->JM_G * (NG + k));
}
static inline long idx_v(const or_block* b, int i, int j, int k) {
  return (NGV + i) + (long)b->IM_GV * ((NGV + j) + (long)b->JM_GV * (NGV + k));
}
static inline long idx_Q(const or_block* b, int e, int i, int j, int k) {
  return e + NEQ * idx_g(b, i, j, k);
}
static inline long idx_Ep(const or_block* b, int e, int i, int j, int k) {
  return e + NEQ * (i + (long)(b->IM + 1) * (j + (long)b->JM * k));
}
static inline long idx_Fp(const or_block* b, int e, int i, int j, int k) {
  return e + NEQ * (i + (long)b->IM * (j + (long)(b->JM + 1) * k));
}
static inline long idx_Gp(const or_block* b, int e, int i, int j, int k) {
  return e + NEQ * (i + (long)b->IM * (j + (long)b->JM * k));
}
static inline long idx_Ev(const or_block* b, int e, int i, int j, int k) {
  return e + (NEQ - 1) * idx_v(b, i, j, k);
}
static inline long idx_dfv(const or_block* b, int e, int i, int j, int k) {
  return e + (NEQ - 1) * (i + (long)b->IM * (j + (long)b->JM * k));
}

static double* alloc0(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

void or_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* src/Block3d.cpp:11-135 and src/Block3d.cpp:208-235 (one block, origin 0) */
or_block* or_create(const or_params* p) {
  or_block* b = (or_block*)calloc(1, sizeof(or_block));
  b->par = *p;
  b->IM = p->IM; b->JM = p->JM; b->KM = p->KM;
  b->IM_G = b->IM + 2 * NG; b->JM_G = b->JM + 2 * NG; b->KM_G = b->KM + 2 * NG;
  b->IM_GV = b->IM + 2 * NGV; b->JM_GV = b->JM + 2 * NGV; b->KM_GV = b->KM + 2 * NGV;
  long ng = (long)b->IM_G * b->JM_G * b->KM_G;
  long nv = (long)b->IM_GV * b->JM_GV * b->KM_GV;
  long nn = (long)b->IM * b->JM * b->KM;
  b->x = alloc0(b->IM + 1); b->y = alloc0(b->JM + 1); b->z = alloc0(b->KM + 1);
  b->rho = alloc0(ng); b->u = alloc0(ng); b->v = alloc0(ng); b->w = alloc0(ng);
  b->p = alloc0(ng); b->T = alloc0(ng); b->mu = alloc0(ng);
  double** gv[] = {&b->ux, &b->vx, &b->wx, &b->uy, &b->vy, &b->wy, &b->uz, &b->vz,
                   &b->wz, &b->Tx, &b->Ty, &b->Tz, &b->txx, &b->tyy, &b->tzz,
                   &b->txy, &b->txz, &b->tyz, &b->qx, &b->qy, &b->qz};
  for (unsigned a = 0; a < sizeof(gv) / sizeof(gv[0]); a++) *gv[a] = alloc0(nv);
  b->Q = alloc0(NEQ * ng); b->Q_p = alloc0(NEQ * ng);
  b->Ev = alloc0((NEQ - 1) * nv); b->Fv = alloc0((NEQ - 1) * nv); b->Gv = alloc0((NEQ - 1) * nv);
  b->dfv = alloc0((NEQ - 1) * nn);
  b->Ep = alloc0((long)NEQ * (b->IM + 1) * b->JM * b->KM);
  b->Fp = alloc0((long)NEQ * b->IM * (b->JM + 1) * b->KM);
  b->Gp = alloc0((long)NEQ * b->IM * b->JM * (b->KM + 1));

  /* gen_mesh(0, 0, 0) */
  double tmp = 2.0 * PI_;
  b->dx = tmp / p->IM; b->dy = tmp / p->JM; b->dz = tmp / p->KM;
  tmp /= 2.0;
  for (int i = 0; i < b->IM + 1; i++) b->x[i] = (double)(unsigned)i * b->dx - tmp;
  for (int j = 0; j < b->JM + 1; j++) b->y[j] = (double)(unsigned)j * b->dy - tmp;
  for (int k = 0; k < b->KM + 1; k++) b->z[k] = (double)(unsigned)k * b->dz - tmp;
  return b;
}

void or_destroy(or_block* b) {
  if (!b) return;
  double* arrs[] = {b->x, b->y, b->z, b->rho, b->u, b->v, b->w, b->p, b->T, b->mu,
                    b->ux, b->vx, b->wx, b->uy, b->vy, b->wy, b->uz, b->vz, b->wz,
                    b->Tx, b->Ty, b->Tz, b->txx, b->tyy, b->tzz, b->txy, b->txz,
                    b->tyz, b->qx, b->qy, b->qz, b->Q, b->Q_p, b->Ev, b->Fv, b->Gv,
                    b->dfv, b->Ep, b->Fp, b->Gp};
  for (unsigned a = 0; a < sizeof(arrs) / sizeof(arrs[0]); a++) free(arrs[a]);
  free(b);
}

int or_dims(const or_block* b, int* d) {
  d[0] = b->IM; d[1] = b->JM; d[2] = b->KM;
  d[3] = b->IM_G; d[4] = b->JM_G; d[5] = b->KM_G;
  return 0;
}

double* or_Q(or_block* b, int which) { return which ? b->Q_p : b->Q; }
double* or_prim(or_block* b, int f) {
  double* a[] = {b->rho, b->u, b->v, b->w, b->p, b->T, b->mu};
  return (f >= 0 && f < 7) ? a[f] : 0;
}
double* or_face_flux(or_block* b, int dir) { return dir == 0 ? b->Ep : dir == 1 ? b->Fp : b->Gp; }
double* or_dfv(or_block* b) { return b->dfv; }
double or_t_cur(const or_block* b) { return b->par.t_cur; }

/* src/Block3d.cpp:237-300 (initial_condition + calc_conservative) */
void or_init_tgv(or_block* b) {
  const or_params* sp = &b->par;
  for (int k = 0; k < b->KM; k++) {
    double z_l = b->z[k] + 0.5 * b->dz;
    for (int j = 0; j < b->JM; j++) {
      double y_l = b->y[j] + 0.5 * b->dy;
      for (int i = 0; i < b->IM; i++) {
        double x_l = b->x[i] + 0.5 * b->dx;
        long id = idx_g(b, i, j, k);
        b->u[id] = sin(x_l) * cos(y_l) * cos(z_l);
        b->v[id] = -cos(x_l) * sin(y_l) * cos(z_l);
        b->w[id] = 0.0;
        b->p[id] = sp->p_inf + (cos(2.0 * x_l) + cos(2.0 * y_l)) * (cos(2.0 * z_l) + 2.0) / 16.0;
        b->rho[id] = sp->gM2 * b->p[id];
        b->T[id] = 1.0;
        b->mu[id] = 1.0;
      }
    }
  }
  for (int k = 0; k < b->KM; k++)
    for (int j = 0; j < b->JM; j++)
      for (int i = 0; i < b->IM; i++) {
        long id = idx_g(b, i, j, k), iq = idx_Q(b, 0, i, j, k);
        double rr = b->rho[id], uu = b->u[id], vv = b->v[id], ww = b->w[id], pp = b->p[id];
        b->Q[iq] = rr;
        b->Q[iq + 1] = rr * uu;
        b->Q[iq + 2] = rr * vv;
        b->Q[iq + 3] = rr * ww;
        b->Q[iq + 4] = pp / sp->gam1 + 0.5 * rr * (uu * uu + vv * vv + ww * ww);
      }
  /* src/cbns3d.cpp:149-153 */
  or_boundary_condition(b, 0);
  or_calc_primitive(b, 0);
  or_init_Q_p(b);
}

/* src/Block3d.cpp:318-357 (read_bin, after a successful read) followed by
 * src/cbns3d.cpp:149-153.  mu is not part of the checkpoint. */
void or_set_state(or_block* b, const double* Qg, double t_cur, double mu0) {
  long n = (long)NEQ * b->IM_G * b->JM_G * b->KM_G;
  memcpy(b->Q, Qg, (size_t)n * sizeof(double));
  b->par.t_cur = t_cur;
  long ng = (long)b->IM_G * b->JM_G * b->KM_G;
  for (long a = 0; a < ng; a++) b->mu[a] = mu0;
  or_boundary_condition(b, 0);
  or_calc_primitive(b, 0);
  or_init_Q_p(b);
}

/* src/boundary_condition.cpp:4-231, one rank with periodic neighbours = itself.
 * The three phases (x over interior j,k; y over all i, interior k; z over
 * all i, j) leave every ghost equal to its periodic image. */
static inline int wrap(int a, int n) { return ((a % n) + n) % n; }
void or_boundary_condition(or_block* b, int which) {
  double* Q = which ? b->Q_p : b->Q;
  /* x phase */
  for (int k = 0; k < b->KM; k++)
    for (int j = 0; j < b->JM; j++)
      for (int g = 0; g < NG; g++) {
        int il = g - NG, ir = b->IM + g;
        for (int e = 0; e < NEQ; e++) {
          Q[idx_Q(b, e, il, j, k)] = Q[idx_Q(b, e, wrap(il, b->IM), j, k)];
          Q[idx_Q(b, e, ir, j, k)] = Q[idx_Q(b, e, wrap(ir, b->IM), j, k)];
        }
      }
  /* y phase */
  for (int k = 0; k < b->KM; k++)
    for (int i = -NG; i < b->IM + NG; i++)
      for (int g = 0; g < NG; g++) {
        int jl = g - NG, jr = b->JM + g;
        for (int e = 0; e < NEQ; e++) {
          Q[idx_Q(b, e, i, jl, k)] = Q[idx_Q(b, e, i, wrap(jl, b->JM), k)];
          Q[idx_Q(b, e, i, jr, k)] = Q[idx_Q(b, e, i, wrap(jr, b->JM), k)];
        }
      }
  /* z phase */
  for (int j = -NG; j < b->JM + NG; j++)
    for (int i = -NG; i < b->IM + NG; i++)
      for (int g = 0; g < NG; g++) {
        int kl = g - NG, kr = b->KM + g;
        for (int e = 0; e < NEQ; e++) {
          Q[idx_Q(b, e, i, j, kl)] = Q[idx_Q(b, e, i, j, wrap(kl, b->KM))];
          Q[idx_Q(b, e, i, j, kr)] = Q[idx_Q(b, e, i, j, wrap(kr, b->KM))];
        }
      }
}

/* src/Block3d.cpp:359-385 */
void or_calc_primitive(or_block* b, int which) {
  const double* Q = which ? b->Q_p : b->Q;
  const double gam1 = b->par.gam1, gM2 = b->par.gM2;
  long n = (long)b->IM_G * b->JM_G * b->KM_G;
#pragma omp parallel for schedule(static)
  for (long i1 = 0; i1 < n; i1++) {
    long i2 = NEQ * i1;
    double rr = Q[i2];
    double uu = Q[i2 + 1] / rr;
    double vv = Q[i2 + 2] / rr;
    double ww = Q[i2 + 3] / rr;
    double pp = gam1 * (Q[i2 + 4] - 0.5 * rr * (uu * uu + vv * vv + ww * ww));
    b->rho[i1] = rr; b->u[i1] = uu; b->v[i1] = vv; b->w[i1] = ww; b->p[i1] = pp;
    b->T[i1] = gM2 * pp / rr;
  }
}

/* src/Block3d.cpp:387-394 */
void or_init_Q_p(or_block* b) {
  long n = (long)NEQ * b->IM_G * b->JM_G * b->KM_G;
  memcpy(b->Q_p, b->Q, (size_t)n * sizeof(double));
}

/* src/Block3d.cpp:515-553.  The ternary ignores NaN as the reference does. */
double or_calc_dt(or_block* b) {
  const or_params* sp = &b->par;
  double s_max = 0.0;
#pragma omp parallel for schedule(static) reduction(max : s_max)
  for (int k = 0; k < b->KM; k++)
    for (int j = 0; j < b->JM; j++)
      for (int i = 0; i < b->IM; i++) {
        long id = idx_g(b, i, j, k);
        double rr = b->rho[id], uu = b->u[id], vv = b->v[id], ww = b->w[id], pp = b->p[id];
        double c = sqrt(sp->gamma * pp / rr);
        double factor_v = 2.0 * b->mu[id] * sp->C_dt_v * sqrt(uu * uu + vv * vv + ww * ww) / c / rr;
        double t1 = fabs(uu) + c / b->dx;
        t1 += factor_v / b->dx / b->dx;
        double t2 = fabs(vv) + c / b->dy;
        t2 += factor_v / b->dy / b->dy;
        double t3 = fabs(ww) + c / b->dz;
        t3 += factor_v / b->dz / b->dz;
        s_max = ((s_max < t1 + t2 + t3) ? (t1 + t2 + t3) : s_max);
      }
  return sp->CFL / s_max;
}

/* src/Block3d.cpp:555-748 (rec_riemann_x/y/z), one routine for all three. */
void or_rec_riemann(or_block* b, int which, int dir) {
  const double* Q = which ? b->Q_p : b->Q;
  const double gamma = b->par.gamma, gam1 = b->par.gam1;
  const int di = dir == 0, dj = dir == 1, dk = dir == 2;
  const double nx = di ? 1.0 : 0.0, ny = dj ? 1.0 : 0.0, nz = dk ? 1.0 : 0.0;
  const int ni = b->IM + di, nj = b->JM + dj, nk = b->KM + dk;
  double* F = dir == 0 ? b->Ep : dir == 1 ? b->Fp : b->Gp;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nk; k++)
    for (int j = 0; j < nj; j++)
      for (int i = 0; i < ni; i++) {
        double R[25], L[25], pL[5], pR[5], sL[5], sR[5];
        long id2 = idx_g(b, i - di, j - dj, k - dk), id3 = idx_g(b, i, j, k);
        pL[0] = b->rho[id2]; pL[1] = b->u[id2]; pL[2] = b->v[id2]; pL[3] = b->w[id2]; pL[4] = b->p[id2];
        pR[0] = b->rho[id3]; pR[1] = b->u[id3]; pR[2] = b->v[id3]; pR[3] = b->w[id3]; pR[4] = b->p[id3];
        or_calc_eigenvectors(gamma, pL, pR, nx, ny, nz, R, L);
#define QP(o) (Q + idx_Q(b, 0, i + (o) * di, j + (o) * dj, k + (o) * dk))
        or_rec_weno5(QP(-3), QP(-2), QP(-1), QP(0), QP(1), R, L, gam1, sL);
        or_rec_weno5(QP(2), QP(1), QP(0), QP(-1), QP(-2), R, L, gam1, sR);
#undef QP
        long id1 = dir == 0 ? idx_Ep(b, 0, i, j, k) : dir == 1 ? idx_Fp(b, 0, i, j, k) : idx_Gp(b, 0, i, j, k);
        or_lf_flux(gamma, sR, sL, nx, ny, nz, F + id1);
      }
}

/* src/calc_gradient.cpp:17-66 */
static void calc_gradient(or_block* b, const double* f, double* f_x, double* f_y, double* f_z) {
#pragma omp parallel for schedule(static)
  for (int k = -NGV; k < b->KM + NGV; k++)
    for (int j = -NGV; j < b->JM + NGV; j++)
      for (int i = -NGV; i < b->IM + NGV; i++) {
        long i0 = idx_v(b, i, j, k);
        long m1 = idx_g(b, i - 1, j, k), m2 = idx_g(b, i - 2, j, k);
        long p1 = idx_g(b, i + 1, j, k), p2 = idx_g(b, i + 2, j, k);
        f_x[i0] = (1.0 / 12.0) * f[m2] - 2.0 / 3.0 * f[m1] + (2.0 / 3.0) * f[p1] - 1.0 / 12.0 * f[p2];
        m1 = idx_g(b, i, j - 1, k); m2 = idx_g(b, i, j - 2, k);
        p1 = idx_g(b, i, j + 1, k); p2 = idx_g(b, i, j + 2, k);
        f_y[i0] = (1.0 / 12.0) * f[m2] - 2.0 / 3.0 * f[m1] + (2.0 / 3.0) * f[p1] - 1.0 / 12.0 * f[p2];
        m1 = idx_g(b, i, j, k - 1); m2 = idx_g(b, i, j, k - 2);
        p1 = idx_g(b, i, j, k + 1); p2 = idx_g(b, i, j, k + 2);
        f_z[i0] = (1.0 / 12.0) * f[m2] - 2.0 / 3.0 * f[m1] + (2.0 / 3.0) * f[p1] - 1.0 / 12.0 * f[p2];
        f_x[i0] /= b->dx;
        f_y[i0] /= b->dy;
        f_z[i0] /= b->dz;
      }
}

/* src/Block3d.cpp:750-810 */
static void calc_viscous_terms(or_block* b) {
  calc_gradient(b, b->u, b->ux, b->uy, b->uz);
  calc_gradient(b, b->v, b->vx, b->vy, b->vz);
  calc_gradient(b, b->w, b->wx, b->wy, b->wz);
  calc_gradient(b, b->T, b->Tx, b->Ty, b->Tz);
  const double C_T_inf = b->par.C_T_inf, Re_inv = b->par.Re_inv;
  const double kappa = -b->par.Pr_inv * b->par.gam1_inv / b->par.Mach2;
#pragma omp parallel for schedule(static)
  for (int k = -NGV; k < b->KM + NGV; k++)
    for (int j = -NGV; j < b->JM + NGV; j++)
      for (int i = -NGV; i < b->IM + NGV; i++) {
        long i1 = idx_g(b, i, j, k), i2 = idx_v(b, i, j, k);
        double T_l = b->T[i1];
        double mu_l = (1.0 + C_T_inf) / (T_l + C_T_inf) * sqrt(T_l * T_l * T_l);
        b->mu[i1] = mu_l;
        double u_x = b->ux[i2], u_y = b->uy[i2], u_z = b->uz[i2];
        double v_x = b->vx[i2], v_y = b->vy[i2], v_z = b->vz[i2];
        double w_x = b->wx[i2], w_y = b->wy[i2], w_z = b->wz[i2];
        mu_l *= Re_inv;
        b->txx[i2] = (2.0 / 3.0) * mu_l * (2.0 * u_x - v_y - w_z);
        b->tyy[i2] = (2.0 / 3.0) * mu_l * (2.0 * v_y - u_x - w_z);
        b->tzz[i2] = (2.0 / 3.0) * mu_l * (2.0 * w_z - u_x - v_y);
        b->txy[i2] = mu_l * (u_y + v_x);
        b->txz[i2] = mu_l * (u_z + w_x);
        b->tyz[i2] = mu_l * (v_z + w_y);
        mu_l *= kappa;
        b->qx[i2] = mu_l * b->Tx[i2];
        b->qy[i2] = mu_l * b->Ty[i2];
        b->qz[i2] = mu_l * b->Tz[i2];
      }
}

/* src/Block3d.cpp:812-868 */
static void calc_viscous_flux(or_block* b) {
#pragma omp parallel for schedule(static)
  for (int k = -NGV; k < b->KM + NGV; k++)
    for (int j = -NGV; j < b->JM + NGV; j++)
      for (int i = -NGV; i < b->IM + NGV; i++) {
        long i1 = idx_g(b, i, j, k), i2 = idx_v(b, i, j, k), i3 = idx_Ev(b, 0, i, j, k);
        double uu = b->u[i1], vv = b->v[i1], ww = b->w[i1];
        double txx = b->txx[i2], tyy = b->tyy[i2], tzz = b->tzz[i2];
        double txy = b->txy[i2], txz = b->txz[i2], tyz = b->tyz[i2];
        b->Ev[i3] = txx; b->Ev[i3 + 1] = txy; b->Ev[i3 + 2] = txz;
        b->Ev[i3 + 3] = uu * txx + vv * txy + ww * txz - b->qx[i2];
        b->Fv[i3] = txy; b->Fv[i3 + 1] = tyy; b->Fv[i3 + 2] = tyz;
        b->Fv[i3 + 3] = uu * txy + vv * tyy + ww * tyz - b->qy[i2];
        b->Gv[i3] = txz; b->Gv[i3 + 1] = tyz; b->Gv[i3 + 2] = tzz;
        b->Gv[i3 + 3] = uu * txz + vv * tyz + ww * tzz - b->qz[i2];
      }
}

/* src/Block3d.cpp:870-918 */
void or_calc_viscous_flux_contribution(or_block* b) {
  calc_viscous_terms(b);
  calc_viscous_flux(b);
  const double c0 = 1.0 / 12.0, c1 = -2.0 / 3.0;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < b->KM; k++)
    for (int j = 0; j < b->JM; j++)
      for (int i = 0; i < b->IM; i++)
        for (int e = 0; e < NEQ - 1; e++) {
          double df = (c0 * b->Ev[idx_Ev(b, e, i - 2, j, k)] + c1 * b->Ev[idx_Ev(b, e, i - 1, j, k)] -
                       c1 * b->Ev[idx_Ev(b, e, i + 1, j, k)] - c0 * b->Ev[idx_Ev(b, e, i + 2, j, k)]) / b->dx;
          df += (c0 * b->Fv[idx_Ev(b, e, i, j - 2, k)] + c1 * b->Fv[idx_Ev(b, e, i, j - 1, k)] -
                 c1 * b->Fv[idx_Ev(b, e, i, j + 1, k)] - c0 * b->Fv[idx_Ev(b, e, i, j + 2, k)]) / b->dy;
          df += (c0 * b->Gv[idx_Ev(b, e, i, j, k - 2)] + c1 * b->Gv[idx_Ev(b, e, i, j, k - 1)] -
                 c1 * b->Gv[idx_Ev(b, e, i, j, k + 1)] - c0 * b->Gv[idx_Ev(b, e, i, j, k + 2)]) / b->dz;
          b->dfv[idx_dfv(b, e, i, j, k)] = df;
        }
}

/* src/Block3d.cpp:920-962 */
void or_update_rk3(or_block* b, double dt, int stage) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < b->KM; k++)
    for (int j = 0; j < b->JM; j++)
      for (int i = 0; i < b->IM; i++)
        for (int e = 0; e < NEQ; e++) {
          double df = (b->Ep[idx_Ep(b, e, i + 1, j, k)] - b->Ep[idx_Ep(b, e, i, j, k)]) / b->dx;
          df += (b->Fp[idx_Fp(b, e, i, j + 1, k)] - b->Fp[idx_Fp(b, e, i, j, k)]) / b->dy;
          df += (b->Gp[idx_Gp(b, e, i, j, k + 1)] - b->Gp[idx_Gp(b, e, i, j, k)]) / b->dz;
          if (e > 0) df -= b->dfv[idx_dfv(b, e - 1, i, j, k)];
          long iq = idx_Q(b, e, i, j, k);
          if (1 == stage) {
            b->Q_p[iq] = b->Q[iq] - dt * df;
          } else if (2 == stage) {
            b->Q_p[iq] = 0.75 * b->Q[iq] + 0.25 * (b->Q_p[iq] - dt * df);
          } else if (3 == stage) {
            b->Q[iq] = (b->Q[iq] + 2.0 * (b->Q_p[iq] - dt * df)) / 3.0;
          }
        }
}

/* src/Block3d.cpp:502-513 and src/cbns3d.cpp:178-240 */
int or_step(or_block* b, double* dt_out) {
  or_params* sp = &b->par;
  if (sp->t_cur + 1.0e-5 > sp->t_end) return 1;
  double dt = or_calc_dt(b); /* MPI_Allreduce(MIN) over one rank = identity */
  if (sp->t_end - sp->t_cur < dt) dt = sp->t_end - sp->t_cur;
  sp->t_cur += dt;
  if (dt_out) *dt_out = dt;
  for (int stage = 1; stage <= 3; stage++) {
    int in = (stage == 1) ? 0 : 1;
    or_rec_riemann(b, in, 0);
    or_rec_riemann(b, in, 1);
    or_rec_riemann(b, in, 2);
    or_calc_viscous_flux_contribution(b);
    or_update_rk3(b, dt, stage);
    int out = (stage == 3) ? 0 : 1;
    or_boundary_condition(b, out);
    or_calc_primitive(b, out);
  }
  return 0;
}

void or_get_interior(or_block* b, int which, double* out) {
  const double* Q = which ? b->Q_p : b->Q;
  long o = 0;
  for (int k = 0; k < b->KM; k++)
    for (int j = 0; j < b->JM; j++) {
      memcpy(out + o, Q + idx_Q(b, 0, 0, j, k), (size_t)NEQ * b->IM * sizeof(double));
      o += (long)NEQ * b->IM;
    }
}

/* Diagnostics (no counterpart in the reference): means over the interior of
 * 0.5 rho |u|^2 and 0.5 rho |omega|^2 from the primitives of Q, with the
 * derivatives taken by the reference's calc_gradient stencil. */
void or_diagnostics(or_block* b, double* ke, double* ens) {
  double sk = 0.0, se = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sk, se)
  for (int k = 0; k < b->KM; k++)
    for (int j = 0; j < b->JM; j++)
      for (int i = 0; i < b->IM; i++) {
        long id = idx_g(b, i, j, k);
        double rr = b->rho[id], uu = b->u[id], vv = b->v[id], ww = b->w[id];
        sk += 0.5 * rr * (uu * uu + vv * vv + ww * ww);
#define D(f, di, dj, dk, h)                                                                        \
  (((1.0 / 12.0) * f[idx_g(b, i - 2 * di, j - 2 * dj, k - 2 * dk)] -                              \
    2.0 / 3.0 * f[idx_g(b, i - di, j - dj, k - dk)] + (2.0 / 3.0) * f[idx_g(b, i + di, j + dj, k + dk)] - \
    1.0 / 12.0 * f[idx_g(b, i + 2 * di, j + 2 * dj, k + 2 * dk)]) / (h))
        double wy = D(b->w, 0, 1, 0, b->dy), vz = D(b->v, 0, 0, 1, b->dz);
        double uz = D(b->u, 0, 0, 1, b->dz), wx = D(b->w, 1, 0, 0, b->dx);
        double vx = D(b->v, 1, 0, 0, b->dx), uy = D(b->u, 0, 1, 0, b->dy);
#undef D
        double ox = wy - vz, oy = uz - wx, oz = vx - uy;
        se += 0.5 * rr * (ox * ox + oy * oy + oz * oz);
      }
  double n = (double)b->IM * b->JM * b->KM;
  *ke = sk / n;
  *ens = se / n;
}
