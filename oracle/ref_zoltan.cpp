// ref_zoltan.cpp — C entry points over the reference's zoltan-application
// kernels (zoltan/{bf,penalty,hybrid}.cpp, compiled where they lie by
// oracle/Makefile into _ref/libzoltan_ref.so).  Test infrastructure only:
// tests/golden/gen_zoltan.py writes golden vectors with it.  The OpenMP
// pragmas of hybrid.cpp are inactive (no -fopenmp): serial semantics.
#include <vector>

void bf(double xA[3], double yA[3], double zA[3], double xB[3], double yB[3], double zB[3],
        double& xPA, double& yPA, double& zPA, double& xPB, double& yPB, double& zPB);
void penalty(unsigned int it, unsigned int nt, double a[3], double b[3], double c[3],
             double* d[3], double* e[3], double* f[3], double* p[3], double* q[3]);
void hybrid(int it, int nt, double a[3], double b[3], double c[3], double* d[3], double* e[3],
            double* f[3], double* p[3], double* q[3]);

extern "C" {

// P and Q keep their input values when the reference returns early.
void ref_zoltan_bf(const double* A9, const double* B9, double* P, double* Q) {
  double xa[3] = {A9[0], A9[1], A9[2]}, ya[3] = {A9[3], A9[4], A9[5]},
         za[3] = {A9[6], A9[7], A9[8]};
  double xb[3] = {B9[0], B9[1], B9[2]}, yb[3] = {B9[3], B9[4], B9[5]},
         zb[3] = {B9[6], B9[7], B9[8]};
  bf(xa, ya, za, xb, yb, zb, P[0], P[1], P[2], Q[0], Q[1], Q[2]);
}

// Arrays are vertex-major: D[k*n + i] is coordinate k of vertex D of triangle i.
void ref_zoltan_penalty(unsigned it, unsigned nt, const double* abc9, double* D, double* E,
                        double* F, double* P, double* Q, int n) {
  double a[3] = {abc9[0], abc9[1], abc9[2]}, b[3] = {abc9[3], abc9[4], abc9[5]},
         c[3] = {abc9[6], abc9[7], abc9[8]};
  double* d[3] = {D, D + n, D + 2 * n};
  double* e[3] = {E, E + n, E + 2 * n};
  double* f[3] = {F, F + n, F + 2 * n};
  double* p[3] = {P, P + n, P + 2 * n};
  double* q[3] = {Q, Q + n, Q + 2 * n};
  penalty(it, nt, a, b, c, d, e, f, p, q);
}

void ref_zoltan_hybrid(int it, int nt, const double* abc9, double* D, double* E, double* F,
                       double* P, double* Q, int n) {
  double a[3] = {abc9[0], abc9[1], abc9[2]}, b[3] = {abc9[3], abc9[4], abc9[5]},
         c[3] = {abc9[6], abc9[7], abc9[8]};
  double* d[3] = {D, D + n, D + 2 * n};
  double* e[3] = {E, E + n, E + 2 * n};
  double* f[3] = {F, F + n, F + 2 * n};
  double* p[3] = {P, P + n, P + 2 * n};
  double* q[3] = {Q, Q + n, Q + 2 * n};
  hybrid(it, nt, a, b, c, d, e, f, p, q);
}

}  // extern "C"
