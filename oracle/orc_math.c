/*
 * orc_math.c -- Gauss 2F1 and natural cubic spline of the CPU oracle
 * (test infrastructure only).
 *
 * Restates libEZmock/math/hypergeom.c:42-194 (Pearson's Taylor series and
 * single-fraction recurrence, with the z < -1 and -1 <= z < 0
 * transformations) and libEZmock/math/cspline.c:35-105 (Thomas-style
 * natural spline, Hornbeck 2020).  The same algorithms are used so that the
 * oracle reproduces the reference's floating-point results; the reference
 * files themselves are compiled into oracle/_ref to check this.
 */
#include "orc.h"
#include <math.h>

#define HYP_TOL 1e-12
#define HYP_MAXITER 10000

/* hypergeom.c:42-74: Taylor series, |a|,|b| < 50, |c| > 1, |z| < 0.9. */
static int hyp_series(double a, double b, double c, double z, double tol,
    int maxiter, double *res) {
  if (fabs(a) >= 50 || fabs(b) >= 50 || fabs(c) <= 1 || fabs(z) >= 0.9) {
    *res = HUGE_VAL; return 1;
  }
  double term = 1, sum = 1;
  int quiet = 0;
  for (int n = 0; n < maxiter; n++) {
    double dn = n;
    term *= (a + dn) * (b + dn) * z / ((c + dn) * (dn + 1));
    if (fabs(term) < fabs(sum) * tol) {
      if (++quiet >= 3) { *res = sum; return 0; }
    }
    else quiet = 0;
    sum += term;
  }
  return 1;
}

/* hypergeom.c:92-124: single-fraction recurrence, |c| < 20. */
static int hyp_fraction(double a, double b, double c, double z, double tol,
    int maxiter, double *res) {
  if (fabs(a) >= 50 || fabs(b) >= 50 || fabs(c) >= 20 || fabs(z) >= 0.9) {
    *res = HUGE_VAL; return 1;
  }
  double al = 0, be = 1, ga = 1, ze = 1;
  int quiet = 0;
  for (int n = 0; n < maxiter; n++) {
    double dn = n, last = ze;
    double f = (dn + 1) * (c + dn);
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

/* hypergeom.c:151-194. */
int orc_hyp2f1(double a, double b, double c, double z, double tol,
    int maxiter, double *res) {
  if (fabs(a) >= 50 || fabs(b) >= 50) { *res = HUGE_VAL; return 1; }
  if (z < -1) {
    double w = 1 / (1 - z), f1, f2;
    if (orc_hyp2f1(a, c - b, a - b + 1, w, tol, maxiter, &f1) ||
        orc_hyp2f1(b, c - a, b - a + 1, w, tol, maxiter, &f2)) {
      *res = HUGE_VAL; return 1;
    }
    double lc = lgamma(c);
    double g1 = lc + lgamma(b - a) - lgamma(b) - lgamma(c - a);
    double g2 = lc + lgamma(a - b) - lgamma(a) - lgamma(c - b);
    /* Gamma(b - a) < 0 for -1 < b - a < -1/2, hence the minus sign. */
    *res = -pow(w, a) * exp(g1) * f1 + pow(w, b) * exp(g2) * f2;
    return 0;
  }
  else if (z < 0) {
    double w = z / (z - 1), f;
    if (orc_hyp2f1(a, c - b, c, w, tol, maxiter, &f)) {
      *res = HUGE_VAL; return 1;
    }
    *res = pow(1 - z, -a) * f;
    return 0;
  }
  if (z > 0.5) { *res = HUGE_VAL; return 1; }
  if (fabs(c) > 1) return hyp_series(a, b, c, z, tol, maxiter, res);
  return hyp_fraction(a, b, c, z, tol, maxiter, res);
}

/* cspline.c:35-75: second derivatives of a natural cubic spline.
   ypp holds 2n doubles; the upper half is the work array. */
void orc_cspline_ypp(const double *x, const double *y, size_t n,
    double *ypp) {
  if (n < 2 || !ypp) return;
  double *w = ypp + n;
  w[0] = w[n - 1] = ypp[0] = ypp[n - 1] = 0;
  double h_prev = x[1] - x[0];
  double s_prev = (y[1] - y[0]) / h_prev;
  size_t j;
  for (j = 1; j < n - 1; j++) {
    double h = x[j + 1] - x[j];
    double s = (y[j + 1] - y[j]) / h;
    double diag = (h + h_prev) * 2;
    double inv = 1 / (diag - h_prev * w[j - 1]);
    double rhs = (s - s_prev) * 6;
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

/* cspline.c:91-105: evaluate in the interval [x[i], x[i+1]]. */
double orc_cspline_eval(const double *x, const double *y, const double *ypp,
    double xv, size_t i) {
  size_t j = i + 1;
  double h = x[j] - x[i];
  double dl = xv - x[i];
  double dr = x[j] - xv;
  double h2 = h * h;
  double lin = dl * y[j] + dr * y[i];
  double cr = (dl * dl - h2) * dl * ypp[j];
  double cl = (dr * dr - h2) * dr * ypp[i];
  return (lin + 0x1.5555555555555p-3 * (cr + cl)) / h;
}
