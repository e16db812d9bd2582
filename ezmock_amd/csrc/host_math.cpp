// host_math.cpp -- host-side set-up math of the engine (runs once per mock on
// at most ~1000 table entries; not part of the device hot path).
//
//   ez_hyp2f1       Gauss 2F1 as in the reference's math/hypergeom.c:42-194
//                   (J. Pearson's Taylor series and single-fraction
//                   recurrence, z < -1 and -1 <= z < 0 transformations).
//   ez_cspline_ypp  natural cubic spline second derivatives as in
//                   math/cspline.c:35-75 (Hornbeck 2020 tridiagonal sweep).
// The same algorithms (and operation order) are used so the growth factors
// and P(k) tables agree with the reference to the last bit.
#include <math.h>
#include <stddef.h>
#include "host_math.h"

static int series(double a, double b, double c, double z, double tol,
    int maxiter, double *res) {
  if (fabs(a) >= 50 || fabs(b) >= 50 || fabs(c) <= 1 || fabs(z) >= 0.9) {
    *res = HUGE_VAL;
    return 1;
  }
  double term = 1, sum = 1;
  int quiet = 0;
  for (int n = 0; n < maxiter; n++) {
    const double dn = n;
    term *= (a + dn) * (b + dn) * z / ((c + dn) * (dn + 1));
    if (fabs(term) < fabs(sum) * tol) {
      if (++quiet >= 3) { *res = sum; return 0; }
    }
    else quiet = 0;
    sum += term;
  }
  return 1;
}

static int fraction(double a, double b, double c, double z, double tol,
    int maxiter, double *res) {
  if (fabs(a) >= 50 || fabs(b) >= 50 || fabs(c) >= 20 || fabs(z) >= 0.9) {
    *res = HUGE_VAL;
    return 1;
  }
  double al = 0, be = 1, ga = 1, ze = 1;
  int quiet = 0;
  for (int n = 0; n < maxiter; n++) {
    const double dn = n, last = ze;
    const double f = (dn + 1) * (c + dn);
    al = (al + be) * f;
    be *= (a + dn) * (b + dn) * z;
    ga *= f;
    ze = (al + be) / ga;
    if (fabs(ze - last) < fabs(last) * tol) {
      if (++quiet >= 3) { *res = ze; return 0; }
    }
    else quiet = 0;
  }
  return 1;
}

int ez_hyp2f1(double a, double b, double c, double z, double tol, int maxiter,
    double *res) {
  if (fabs(a) >= 50 || fabs(b) >= 50) { *res = HUGE_VAL; return 1; }
  if (z < -1) {
    const double w = 1 / (1 - z);
    double f1, f2;
    if (ez_hyp2f1(a, c - b, a - b + 1, w, tol, maxiter, &f1) ||
        ez_hyp2f1(b, c - a, b - a + 1, w, tol, maxiter, &f2)) {
      *res = HUGE_VAL;
      return 1;
    }
    const double lc = lgamma(c);
    const double g1 = lc + lgamma(b - a) - lgamma(b) - lgamma(c - a);
    const double g2 = lc + lgamma(a - b) - lgamma(a) - lgamma(c - b);
    // Gamma(b - a) < 0 for -1 < b - a < -1/2: the first term changes sign.
    *res = -pow(w, a) * exp(g1) * f1 + pow(w, b) * exp(g2) * f2;
    return 0;
  }
  if (z < 0) {
    const double w = z / (z - 1);
    double f;
    if (ez_hyp2f1(a, c - b, c, w, tol, maxiter, &f)) { *res = HUGE_VAL; return 1; }
    *res = pow(1 - z, -a) * f;
    return 0;
  }
  if (z > 0.5) { *res = HUGE_VAL; return 1; }
  if (fabs(c) > 1) return series(a, b, c, z, tol, maxiter, res);
  return fraction(a, b, c, z, tol, maxiter, res);
}

void ez_cspline_ypp(const double *x, const double *y, size_t n, double *ypp) {
  if (n < 2 || !ypp) return;
  double *w = ypp + n;
  w[0] = w[n - 1] = ypp[0] = ypp[n - 1] = 0;
  double h_prev = x[1] - x[0];
  double s_prev = (y[1] - y[0]) / h_prev;
  size_t j;
  for (j = 1; j < n - 1; j++) {
    const double h = x[j + 1] - x[j];
    const double s = (y[j + 1] - y[j]) / h;
    const double diag = (h + h_prev) * 2;
    const double inv = 1 / (diag - h_prev * w[j - 1]);
    const double rhs = (s - s_prev) * 6;
    ypp[j] = (rhs - h_prev * ypp[j - 1]) * inv;
    w[j] = h * inv;
    h_prev = h;
    s_prev = s;
  }
  while (j) {
    j--;
    ypp[j] -= w[j] * ypp[j + 1];
  }
}
