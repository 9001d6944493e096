// Device geometry and angular helpers shared by the sphere-table and
// real-space kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace pp {

// cartesian vector of fractional f (rows of lat are the lattice vectors)
__device__ __forceinline__ void d_frac_to_cart(const double *f, const double *lat, double *o) {
  o[0] = f[0] * lat[0] + f[1] * lat[3] + f[2] * lat[6];
  o[1] = f[0] * lat[1] + f[1] * lat[4] + f[2] * lat[7];
  o[2] = f[0] * lat[2] + f[1] * lat[5] + f[2] * lat[8];
}

__device__ __forceinline__ double d_fact(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; i++)
    r *= i;
  return r;
}

// associated Legendre P_l^m(x), Condon-Shortley phase (utils.c:377-386)
__device__ __forceinline__ double d_legendre(int l, int m, double x) {
  double sgn = 1.0, fac = 1.0;
  if (m < 0) {
    m = -m;
    sgn = (m & 1) ? -1.0 : 1.0;
    fac = d_fact(l - m) / d_fact(l + m);
  }
  if (m > l)
    return 0.0;
  double pmm = 1.0, s = sqrt((1.0 - x) * (1.0 + x)), odd = 1.0;
  for (int i = 1; i <= m; i++) {
    pmm *= -odd * s;
    odd += 2.0;
  }
  double res;
  if (l == m) {
    res = pmm;
  } else {
    double pm1 = x * (2 * m + 1) * pmm;
    if (l == m + 1) {
      res = pm1;
    } else {
      double pl = 0.0;
      for (int ll = m + 2; ll <= l; ll++) {
        pl = (x * (2 * ll - 1) * pm1 - (ll + m - 1) * pmm) / (ll - m);
        pmm = pm1;
        pm1 = pl;
      }
      res = pl;
    }
  }
  return sgn * fac * res;
}

}  // namespace pp
