/*
 * ofp_oracle.c -- CPU oracle for the vector_dist neighbour-search path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  See ofp_oracle.h for
 * the conventions and ofp_oracle_body.inc for the per-function reference
 * citations.  Build: oracle/Makefile (gcc -O3 -ffp-contract=off).
 */
#include "ofp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Monaghan cubic spline, 3-D normalisation 1/(pi h^3):
 * W = s/h^3 {1 - 1.5q^2 + 0.75q^3 (q<1); 0.25(2-q)^3 (1<=q<2); 0}. */
double ofo_w3(double r, double h) {
  const double s = 1.0 / (M_PI * h * h * h);
  double q = r / h;
  if (q < 1.0) return s * (1.0 - 1.5 * q * q + 0.75 * q * q * q);
  if (q < 2.0) {
    double t = 2.0 - q;
    return s * 0.25 * t * t * t;
  }
  return 0.0;
}

/* dW/dr = s/h^4 {-3q + 2.25q^2 (q<1); -0.75(2-q)^2 (1<=q<2); 0}. */
double ofo_dw3(double r, double h) {
  const double s = 1.0 / (M_PI * h * h * h * h);
  double q = r / h;
  if (q < 1.0) return s * (-3.0 * q + 2.25 * q * q);
  if (q < 2.0) {
    double t = 2.0 - q;
    return s * (-0.75 * t * t);
  }
  return 0.0;
}

/* Tait EOS, P = B((rho/rho0)^7 - 1), B = c0^2 rho0 / 7. */
void ofo_tait(const double* rho, size_t n, double rho0, double c0, double* prs) {
  double B = c0 * c0 * rho0 / 7.0;
  for (size_t i = 0; i < n; ++i) {
    double r = rho[i] / rho0;
    double r2 = r * r, r4 = r2 * r2;
    prs[i] = B * (r4 * r2 * r - 1.0);
  }
}

#define T float
#define S _f32
#define FLOOR_T floorf
#include "ofp_oracle_body.inc"
#undef T
#undef S
#undef FLOOR_T

#define T double
#define S _f64
#define FLOOR_T floor
#include "ofp_oracle_body.inc"
#undef T
#undef S
#undef FLOOR_T
