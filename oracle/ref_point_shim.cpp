// C-linkage shim over the reference's per-face point functions, compiled
// together with the reference's own sources (never copied) into
// oracle/_ref/libref_point.so.  Test infrastructure only: it lets the
// oracle's restatement and the device face arithmetic be compared with
// the reference's compiled arithmetic on arbitrary inputs.
//
//   weno5             src/weno.cpp:25-56
//   rec_weno5         src/weno.cpp:77-152
//   calc_eigenvectors src/eigenvectors_roe.cpp:32-171
//   lf_flux           src/lf_flux.cpp:32-99
#include "weno.h"
#include "eigenvectors_roe.h"
#include "lf_flux.h"

value_type weno5(const value_type, const value_type, const value_type,
                 const value_type, const value_type);

extern "C" {

double refpt_weno5(double a, double b, double c, double d, double e) {
  return weno5(a, b, c, d, e);
}

// f: 5 cells x 5 conserved values (cell-major, as rec_weno5's f1..f5);
// R, L: 5x5 row-major; out: rho, u, v, w, p.
void refpt_rec_weno5(const double* f, const double* R, const double* L,
                     double gam1, double* out) {
  value_type Rm[constant::NEQ][constant::NEQ], Lm[constant::NEQ][constant::NEQ];
  for (int a = 0; a < 5; a++)
    for (int b = 0; b < 5; b++) { Rm[a][b] = R[5 * a + b]; Lm[a][b] = L[5 * a + b]; }
  rec_weno5(f, f + 5, f + 10, f + 15, f + 20, Rm, Lm, gam1,
            out[0], out[1], out[2], out[3], out[4]);
}

// primL, primR: rho, u, v, w, p of the two cells; n: normal; R, L out row-major.
void refpt_calc_eigenvectors(double gamma, const double* primL,
                             const double* primR, const double* n,
                             double* R, double* L) {
  value_type Rm[constant::NEQ][constant::NEQ], Lm[constant::NEQ][constant::NEQ];
  calc_eigenvectors(gamma, primL[0], primL[1], primL[2], primL[3], primL[4],
                    primR[0], primR[1], primR[2], primR[3], primR[4],
                    n[0], n[1], n[2], Rm, Lm);
  for (int a = 0; a < 5; a++)
    for (int b = 0; b < 5; b++) { R[5 * a + b] = Rm[a][b]; L[5 * a + b] = Lm[a][b]; }
}

// Argument order as the reference: right state first, then left.
void refpt_lf_flux(double gamma, const double* primR, const double* primL,
                   const double* n, double* flux) {
  lf_flux(gamma, primR[0], primR[1], primR[2], primR[3], primR[4],
          primL[0], primL[1], primL[2], primL[3], primL[4],
          n[0], n[1], n[2], flux);
}

}
