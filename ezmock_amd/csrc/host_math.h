// host_math.h -- host-side set-up math (see host_math.cpp).
#pragma once
#include <stddef.h>

int ez_hyp2f1(double a, double b, double c, double z, double tol, int maxiter,
    double *res);
void ez_cspline_ypp(const double *x, const double *y, size_t n, double *ypp);
