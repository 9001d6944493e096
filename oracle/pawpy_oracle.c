/*
 * pawpy_oracle.c -- TEST INFRASTRUCTURE ONLY (see pawpy_oracle.h).
 *
 * A plain-C, FP64 restatement of the reference's CPU projection path
 * (pawpyseed 0.7.1).  The reference itself cannot be compiled in this
 * image (its sources include <mkl.h>, which is not installed), so this file
 * is the oracle; its functions are pinned in tests/ against library
 * routines (scipy splines and spherical harmonics, numpy FFTs), brute-force
 * enumerations and closed forms, and against the reference's own fixture
 * test_files/WAVECAR.  The FFT below is a project-written mixed-radix
 * transform (the reference uses MKL DFTI); or_fft_use_mkl switches the 3D
 * backward transforms to MKL's, through the FFTW3 interface that libmkl_rt
 * exports, where that library can be loaded at run time (the CPU baseline
 * of the benchmark times the reference's FFT library that way).
 */
#define _GNU_SOURCE /* setenv, dlopen */
#include "pawpy_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* Splines                                                              */
/* ------------------------------------------------------------------ */

/* Cubic spline with y'(x0) = secant slope (y1-y0)/(x1-x0) and y''(x_{n-1}) = 0,
 * the boundary conditions of utils.c:709-734.  Second derivatives M come
 * from a Thomas solve; per-interval coefficients follow utils.c:740-746:
 * value = y_i + t (b_i + t (c_i + t d_i)), t = r - x_i.  b,c,d[n-1] = 0. */
void or_spline_coeff(const double *x, const double *y, int n, double *b,
                     double *c, double *d) {
  double *diag = (double *)malloc(sizeof(double) * n);
  double *upper = (double *)malloc(sizeof(double) * n);
  double *rhs = (double *)malloc(sizeof(double) * n);
  double *M = (double *)malloc(sizeof(double) * n);
  double h0 = x[1] - x[0];
  double slope0 = (y[1] - y[0]) / h0;
  /* clamped start: 2 h0 M0 + h0 M1 = 6 ((y1-y0)/h0 - slope0) */
  diag[0] = 2.0 * h0;
  upper[0] = h0;
  rhs[0] = 6.0 * ((y[1] - y[0]) / h0 - slope0);
  for (int i = 1; i < n - 1; i++) {
    double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
    double lower = hl;
    double dg = 2.0 * (hl + hr);
    double rr = 6.0 * ((y[i + 1] - y[i]) / hr - (y[i] - y[i - 1]) / hl);
    double w = lower / diag[i - 1];
    diag[i] = dg - w * upper[i - 1];
    upper[i] = hr;
    rhs[i] = rr - w * rhs[i - 1];
  }
  /* natural end M_{n-1} = 0 */
  M[n - 1] = 0.0;
  for (int i = n - 2; i >= 0; i--)
    M[i] = (rhs[i] - upper[i] * M[i + 1]) / diag[i];
  for (int i = 0; i < n - 1; i++) {
    double h = x[i + 1] - x[i];
    b[i] = (y[i + 1] - y[i]) / h - h * (2.0 * M[i] + M[i + 1]) / 6.0;
    c[i] = 0.5 * M[i];
    d[i] = (M[i + 1] - M[i]) / (6.0 * h);
  }
  b[n - 1] = c[n - 1] = d[n - 1] = 0.0;
  free(diag);
  free(upper);
  free(rhs);
  free(M);
}

/* utils.c:751-764: exact integral of the piecewise cubic. */
double or_spline_integral(const double *x, const double *a, const double *b,
                          const double *c, const double *d, int n) {
  double s = 0.0;
  for (int i = 0; i < n - 1; i++) {
    double h = x[i + 1] - x[i];
    s += h * (a[i] + h * (0.5 * b[i] + h * (c[i] / 3.0 + 0.25 * d[i] * h)));
  }
  return s;
}

/* utils.c:461-475: zero beyond the last grid point, f[0] below the first,
 * interval index min((int)(r/rmax*n), n-2). */
double or_proj_interpolate(double r, double rmax, int n, const double *x,
                           const double *f, const double *b, const double *c,
                           const double *d) {
  if (r > x[n - 1])
    return 0.0;
  if (r < x[0])
    return f[0];
  int i = (int)(r / rmax * n);
  if (i > n - 2)
    i = n - 2;
  double t = r - x[i];
  return f[i] + t * (b[i] + t * (c[i] + t * d[i]));
}

/* ------------------------------------------------------------------ */
/* Angular functions                                                    */
/* ------------------------------------------------------------------ */

static double fact(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; i++)
    r *= i;
  return r;
}

/* Associated Legendre P_l^m(x) with the Condon-Shortley phase, the
 * convention of utils.c:377-386, by the three-term recurrence in l. */
double or_legendre(int l, int m, double x) {
  if (m < 0) {
    int mm = -m;
    double sgn = (mm % 2) ? -1.0 : 1.0;
    return sgn * fact(l - mm) / fact(l + mm) * or_legendre(l, mm, x);
  }
  if (m > l)
    return 0.0;
  /* P_m^m = (-1)^m (2m-1)!! (1-x^2)^{m/2} */
  double pmm = 1.0;
  double s = sqrt((1.0 - x) * (1.0 + x));
  double odd = 1.0;
  for (int i = 1; i <= m; i++) {
    pmm *= -odd * s;
    odd += 2.0;
  }
  if (l == m)
    return pmm;
  double pm1 = x * (2 * m + 1) * pmm;
  if (l == m + 1)
    return pm1;
  double pl = 0.0;
  for (int ll = m + 2; ll <= l; ll++) {
    pl = (x * (2 * ll - 1) * pm1 - (ll + m - 1) * pmm) / (ll - m);
    pmm = pm1;
    pm1 = pl;
  }
  return pl;
}

/* utils.c:441-449: complex Y_lm(theta, phi), physics convention. */
void or_ylm(int l, int m, double theta, double phi, double *out2) {
  double norm = sqrt((2 * l + 1) / (4.0 * OR_PI) * fact(l - m) / fact(l + m));
  double p = norm * or_legendre(l, m, cos(theta));
  out2[0] = p * cos(m * phi);
  out2[1] = p * sin(m * phi);
}

/* utils.c:547-564: radial spline value times Y_lm of the direction vec
 * (|vec| = r); phi from acos(x/rho), mirrored for y < 0, 0 on the z axis. */
void or_proj_value(int l, int m, double rmax, int n, const double *x,
                   const double *f, const double *b, const double *c,
                   const double *d, const double *vec, double r, double *out2) {
  double radial = or_proj_interpolate(r, rmax, n, x, f, b, c, d);
  double y[2];
  if (r == 0.0) {
    or_ylm(l, m, 0.0, 0.0, y);
  } else {
    double theta = acos(vec[2] / r);
    double phi = 0.0;
    if (r - fabs(vec[2]) != 0.0)
      phi = acos(vec[0] / sqrt(vec[0] * vec[0] + vec[1] * vec[1]));
    if (vec[1] < 0)
      phi = 2.0 * OR_PI - phi;
    or_ylm(l, m, theta, phi, y);
  }
  out2[0] = radial * y[0];
  out2[1] = radial * y[1];
}

/* ------------------------------------------------------------------ */
/* Geometry                                                             */
/* ------------------------------------------------------------------ */

static void frac_to_cart(const double *f, const double *lat, double *out) {
  for (int j = 0; j < 3; j++)
    out[j] = f[0] * lat[j] + f[1] * lat[3 + j] + f[2] * lat[6 + j];
}

static double norm3(const double *v) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

static void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

static double det3(const double *m) {
  return m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) +
         m[2] * (m[3] * m[7] - m[4] * m[6]);
}

/* utils.c:51-73: shortest of the 27 images of coord - center. */
void or_min_cart_path(const double *coord, const double *center,
                      const double *lattice, double *path, double *r) {
  *r = INFINITY;
  for (int i = -1; i <= 1; i++)
    for (int j = -1; j <= 1; j++)
      for (int k = -1; k <= 1; k++) {
        double fv[3] = {coord[0] + i - center[0], coord[1] + j - center[1],
                        coord[2] + k - center[2]};
        double cv[3];
        frac_to_cart(fv, lattice, cv);
        double dist = norm3(cv);
        if (dist < *r) {
          path[0] = cv[0];
          path[1] = cv[1];
          path[2] = cv[2];
          *r = dist;
        }
      }
}

static void site_box(const double *lattice, const int *fftg, double rmax,
                     int *g) {
  /* utils.c:641-646: half-width int(|a_j x a_k| rmax / V * N_i) + 1 */
  double v = det3(lattice), cr[3];
  cross3(lattice + 3, lattice + 6, cr);
  g[0] = (int)(norm3(cr) * rmax / v * fftg[0]) + 1;
  cross3(lattice + 0, lattice + 6, cr);
  g[1] = (int)(norm3(cr) * rmax / v * fftg[1]) + 1;
  cross3(lattice + 0, lattice + 3, cr);
  g[2] = (int)(norm3(cr) * rmax / v * fftg[2]) + 1;
}

int or_site_capacity(const double *lattice, const int *fftg, double rmax) {
  int g[3];
  site_box(lattice, fftg, rmax, g);
  return (2 * g[0] + 1) * (2 * g[1] + 1) * (2 * g[2] + 1);
}

int or_site_table(const double *lattice, const int *fftg, const double *coord,
                  double rmax, int n, const double *x, int nchan,
                  const int *ls, const double *f, const double *b,
                  const double *c, const double *d, int cap, int *idx,
                  double *path, double *values) {
  int g[3];
  site_box(lattice, fftg, rmax, g);
  int c0 = (int)round(coord[0] * fftg[0]);
  int c1 = (int)round(coord[1] * fftg[1]);
  int c2 = (int)round(coord[2] * fftg[2]);
  /* utils.c:651-652: strict radius (n-1)/n * rmax */
  double R0 = (n - 1) * rmax / n;
  int np = 0;
  for (int i = c0 - g[0]; i <= c0 + g[0]; i++)
    for (int j = c1 - g[1]; j <= c1 + g[1]; j++)
      for (int k = c2 - g[2]; k <= c2 + g[2]; k++) {
        double fv[3] = {(double)i / fftg[0] - coord[0],
                        (double)j / fftg[1] - coord[1],
                        (double)k / fftg[2] - coord[2]};
        double cv[3];
        frac_to_cart(fv, lattice, cv);
        if (!(norm3(cv) < R0))
          continue;
        if (np >= cap)
          return -1;
        int ii = (i % fftg[0] + fftg[0]) % fftg[0];
        int jj = (j % fftg[1] + fftg[1]) % fftg[1];
        int kk = (k % fftg[2] + fftg[2]) % fftg[2];
        idx[np] = (ii * fftg[1] + jj) * fftg[2] + kk;
        path[3 * np + 0] = cv[0];
        path[3 * np + 1] = cv[1];
        path[3 * np + 2] = cv[2];
        /* the value uses the minimum-image vector from the site to the
         * wrapped grid point (utils.c:566-576) */
        double pf[3] = {(double)ii / fftg[0], (double)jj / fftg[1],
                        (double)kk / fftg[2]};
        double mv[3], r;
        or_min_cart_path(pf, coord, lattice, mv, &r);
        int p = 0;
        for (int ch = 0; ch < nchan; ch++) {
          int l = ls[ch];
          for (int m = -l; m <= l; m++, p++) {
            double v2[2];
            or_proj_value(l, m, rmax, n, x, f + (long)ch * n, b + (long)ch * n,
                          c + (long)ch * n, d + (long)ch * n, mv, r, v2);
            values[2 * ((long)p * cap + np) + 0] = v2[0];
            values[2 * ((long)p * cap + np) + 1] = v2[1];
          }
        }
        np++;
      }
  return np;
}

/* ------------------------------------------------------------------ */
/* FFT                                                                  */
/* ------------------------------------------------------------------ */

/* One 1D plan: radices (4 first, then primes) and per-stage twiddles
 * w^(k r) = exp(+2 pi i k r / (p R)) for k < p, r < R, plus the R-th roots. */
typedef struct {
  int n, nf, f[32];
  double *tw[32];   /* [p][R] complex per stage */
  double *root[32]; /* [R] complex per stage */
} fft_plan;

static void plan_make(fft_plan *pl, int n) {
  pl->n = n;
  pl->nf = 0;
  int m = n;
  while (m % 4 == 0) {
    pl->f[pl->nf++] = 4;
    m /= 4;
  }
  for (int q = 2; m > 1 && q <= m; q++)
    while (m % q == 0) {
      pl->f[pl->nf++] = q;
      m /= q;
    }
  int p = 1;
  for (int s = 0; s < pl->nf; s++) {
    int R = pl->f[s];
    pl->tw[s] = (double *)malloc(sizeof(double) * 2 * p * R);
    pl->root[s] = (double *)malloc(sizeof(double) * 2 * R);
    for (int k = 0; k < p; k++)
      for (int r = 0; r < R; r++) {
        double ang = 2.0 * OR_PI * (double)(k * r) / (double)(p * R);
        pl->tw[s][2 * (k * R + r)] = cos(ang);
        pl->tw[s][2 * (k * R + r) + 1] = sin(ang);
      }
    for (int r = 0; r < R; r++) {
      pl->root[s][2 * r] = cos(2.0 * OR_PI * r / R);
      pl->root[s][2 * r + 1] = sin(2.0 * OR_PI * r / R);
    }
    p *= R;
  }
}

static void plan_free(fft_plan *pl) {
  for (int s = 0; s < pl->nf; s++) {
    free(pl->tw[s]);
    free(pl->root[s]);
  }
}

/* Backward transform of nb sequences stored [e][b] (b fastest, length n
 * each) in a, using b as scratch of the same size; result left in a.
 * Stockham autosort: stage (R, p): for butterfly j = (q, k), k = j mod p,
 * inputs a[j + r n/R] * w^(k r), outputs b[(j/p) p R + k + r p]. */
static void fft_batch(const fft_plan *pl, double *a, double *b, int nb) {
  int n = pl->n, p = 1;
  double *in = a, *out = b;
  double v[2 * 64];
  for (int s = 0; s < pl->nf; s++) {
    int R = pl->f[s], T = n / R;
    const double *tw = pl->tw[s], *rt = pl->root[s];
    for (int j = 0; j < T; j++) {
      int k = j % p;
      int o = (j / p) * p * R + k;
      for (int c = 0; c < nb; c++) {
        for (int r = 0; r < R; r++) {
          const double *x = in + 2 * ((long)(j + r * T) * nb + c);
          double wr = tw[2 * (k * R + r)], wi = tw[2 * (k * R + r) + 1];
          v[2 * r] = x[0] * wr - x[1] * wi;
          v[2 * r + 1] = x[0] * wi + x[1] * wr;
        }
        if (R == 4) {
          double a0r = v[0] + v[4], a0i = v[1] + v[5], a1r = v[0] - v[4], a1i = v[1] - v[5];
          double b0r = v[2] + v[6], b0i = v[3] + v[7], b1r = -(v[3] - v[7]), b1i = v[2] - v[6];
          double *y0 = out + 2 * ((long)(o + 0 * p) * nb + c), *y1 = out + 2 * ((long)(o + 1 * p) * nb + c);
          double *y2 = out + 2 * ((long)(o + 2 * p) * nb + c), *y3 = out + 2 * ((long)(o + 3 * p) * nb + c);
          y0[0] = a0r + b0r;
          y0[1] = a0i + b0i;
          y2[0] = a0r - b0r;
          y2[1] = a0i - b0i;
          y1[0] = a1r + b1r;
          y1[1] = a1i + b1i;
          y3[0] = a1r - b1r;
          y3[1] = a1i - b1i;
        } else if (R == 2) {
          double *y0 = out + 2 * ((long)o * nb + c), *y1 = out + 2 * ((long)(o + p) * nb + c);
          y0[0] = v[0] + v[2];
          y0[1] = v[1] + v[3];
          y1[0] = v[0] - v[2];
          y1[1] = v[1] - v[3];
        } else {
          for (int q = 0; q < R; q++) {
            double sr = 0.0, si = 0.0;
            for (int r = 0; r < R; r++) {
              int e = (r * q) % R;
              sr += v[2 * r] * rt[2 * e] - v[2 * r + 1] * rt[2 * e + 1];
              si += v[2 * r] * rt[2 * e + 1] + v[2 * r + 1] * rt[2 * e];
            }
            double *y = out + 2 * ((long)(o + q * p) * nb + c);
            y[0] = sr;
            y[1] = si;
          }
        }
      }
    }
    double *t = in;
    in = out;
    out = t;
    p *= R;
  }
  if (in != a)
    memcpy(a, in, sizeof(double) * 2 * (size_t)n * nb);
}

/* MKL's 3D transforms, through the FFTW3 entry points of libmkl_rt
 * (FFTW3 API: fftw_plan_dft_3d(n0, n1, n2, in, out, sign, flags), row-major
 * complex doubles, sign +1 = the unnormalised backward transform, flag
 * FFTW_ESTIMATE = 1 << 6; fftw_execute_dft reuses a plan on another array
 * of the same layout and is thread safe).  MKL runs sequentially inside:
 * the callers' OpenMP loops over (band, k) provide the parallelism.  One
 * plan for the last grid, made under a lock. */
typedef void *(*plan3d_fn)(int, int, int, double *, double *, int, unsigned);
typedef void (*exec_fn)(void *, double *, double *);
typedef void (*destroy_fn)(void *);
static plan3d_fn mkl_plan3d;
static exec_fn mkl_exec;
static destroy_fn mkl_destroy;
static int mkl_use;           /* 1: 3D transforms through MKL */
static void *mkl_plan;
static int mkl_dims[3];

int or_fft_use_mkl(int enable) {
  if (!enable) {
    mkl_use = 0;
    return 0;
  }
  if (!mkl_exec) {
    const char *env = getenv("PP_ORACLE_MKL");
    const char *cands[] = {env, "libmkl_rt.so.1", "libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1",
                           "/opt/intel/oneapi/mkl/latest/lib/libmkl_rt.so.2"};
    void *h = NULL;
    /* MKL reads its threading layer at its first call */
    setenv("MKL_THREADING_LAYER", "SEQUENTIAL", 0);
    for (unsigned i = 0; i < sizeof(cands) / sizeof(cands[0]) && !h; i++)
      if (cands[i])
        h = dlopen(cands[i], RTLD_NOW | RTLD_LOCAL);
    if (!h)
      return 0;
    mkl_plan3d = (plan3d_fn)dlsym(h, "fftw_plan_dft_3d");
    mkl_exec = (exec_fn)dlsym(h, "fftw_execute_dft");
    mkl_destroy = (destroy_fn)dlsym(h, "fftw_destroy_plan");
    if (!mkl_plan3d || !mkl_exec || !mkl_destroy) {
      mkl_exec = NULL;
      return 0;
    }
  }
  mkl_use = 1;
  return 1;
}

/* x in place through MKL; 0 when MKL is not in use */
static int mkl_ifft3d(double *x, const int *fftg) {
  if (!mkl_use)
    return 0;
  void *plan;
#pragma omp critical(or_mkl_plan)
  {
    if (!mkl_plan || mkl_dims[0] != fftg[0] || mkl_dims[1] != fftg[1] || mkl_dims[2] != fftg[2]) {
      if (mkl_plan)
        mkl_destroy(mkl_plan);
      mkl_plan = mkl_plan3d(fftg[0], fftg[1], fftg[2], x, x, 1, 1U << 6);
      mkl_dims[0] = fftg[0];
      mkl_dims[1] = fftg[1];
      mkl_dims[2] = fftg[2];
    }
    plan = mkl_plan;
  }
  if (!plan)
    return 0;
  mkl_exec(plan, x, x);
  return 1;
}

/* work: 4 * nmax * nmax doubles (two [n][nb] complex buffers) */
static void ifft3d_with(double *x, const int *fftg, fft_plan *pl, double *work) {
  if (mkl_ifft3d(x, fftg))
    return;
  int n0 = fftg[0], n1 = fftg[1], n2 = fftg[2];
  int nmax = n0 > n1 ? n0 : n1;
  nmax = nmax > n2 ? nmax : n2;
  double *A = work, *B = work + 2 * (size_t)nmax * nmax;
  /* z transforms: for each x, sequences along z, batch over y (transpose in) */
  for (int i = 0; i < n0; i++) {
    double *pl0 = x + 2 * (size_t)i * n1 * n2;
    for (int j = 0; j < n1; j++)
      for (int k = 0; k < n2; k++) {
        A[2 * ((size_t)k * n1 + j)] = pl0[2 * ((size_t)j * n2 + k)];
        A[2 * ((size_t)k * n1 + j) + 1] = pl0[2 * ((size_t)j * n2 + k) + 1];
      }
    fft_batch(&pl[2], A, B, n1);
    for (int j = 0; j < n1; j++)
      for (int k = 0; k < n2; k++) {
        pl0[2 * ((size_t)j * n2 + k)] = A[2 * ((size_t)k * n1 + j)];
        pl0[2 * ((size_t)j * n2 + k) + 1] = A[2 * ((size_t)k * n1 + j) + 1];
      }
  }
  /* y transforms: for each x the plane is already [j][k] = [e][b] */
  for (int i = 0; i < n0; i++) {
    double *pl0 = x + 2 * (size_t)i * n1 * n2;
    memcpy(A, pl0, sizeof(double) * 2 * (size_t)n1 * n2);
    fft_batch(&pl[1], A, B, n2);
    memcpy(pl0, A, sizeof(double) * 2 * (size_t)n1 * n2);
  }
  /* x transforms: for each y, [i][k] with stride n1*n2 between i */
  for (int j = 0; j < n1; j++) {
    for (int i = 0; i < n0; i++)
      memcpy(A + 2 * (size_t)i * n2, x + 2 * (((size_t)i * n1 + j) * n2), sizeof(double) * 2 * n2);
    fft_batch(&pl[0], A, B, n2);
    for (int i = 0; i < n0; i++)
      memcpy(x + 2 * (((size_t)i * n1 + j) * n2), A + 2 * (size_t)i * n2, sizeof(double) * 2 * n2);
  }
}

static double *fft_work(const int *fftg) {
  int nmax = fftg[0];
  if (fftg[1] > nmax)
    nmax = fftg[1];
  if (fftg[2] > nmax)
    nmax = fftg[2];
  return (double *)malloc(sizeof(double) * 4 * (size_t)nmax * nmax);
}

void or_ifft3d_inplace(double *x, const int *fftg) {
  fft_plan pl[3];
  for (int a = 0; a < 3; a++)
    plan_make(&pl[a], fftg[a]);
  double *work = fft_work(fftg);
  ifft3d_with(x, fftg, pl, work);
  free(work);
  for (int a = 0; a < 3; a++)
    plan_free(&pl[a]);
}

/* linalg.c:14-45: zero the grid, assign x[(G+N)%N] = C_G (z fastest),
 * backward transform, scale by V^{-1/2}. */
static void fft3d_with(double *x, const double *lattice, const int *Gs,
                       const float *Cs, int nw, const int *fftg,
                       fft_plan *pl, double *work) {
  long ng = (long)fftg[0] * fftg[1] * fftg[2];
  memset(x, 0, sizeof(double) * 2 * ng);
  for (int w = 0; w < nw; w++) {
    int g1 = (Gs[3 * w + 0] + fftg[0]) % fftg[0];
    int g2 = (Gs[3 * w + 1] + fftg[1]) % fftg[1];
    int g3 = (Gs[3 * w + 2] + fftg[2]) % fftg[2];
    long o = ((long)g1 * fftg[1] + g2) * fftg[2] + g3;
    x[2 * o] = Cs[2 * w];
    x[2 * o + 1] = Cs[2 * w + 1];
  }
  ifft3d_with(x, fftg, pl, work);
  double s = 1.0 / sqrt(det3(lattice));
  for (long i = 0; i < 2 * ng; i++)
    x[i] *= s;
}

void or_fft3d(double *x, const double *lattice, const int *Gs,
              const float *Cs_c64, int nw, const int *fftg) {
  fft_plan pl[3];
  for (int a = 0; a < 3; a++)
    plan_make(&pl[a], fftg[a]);
  double *work = fft_work(fftg);
  fft3d_with(x, lattice, Gs, Cs_c64, nw, fftg, pl, work);
  free(work);
  for (int a = 0; a < 3; a++)
    plan_free(&pl[a]);
}

/* ------------------------------------------------------------------ */
/* Sphere sums                                                          */
/* ------------------------------------------------------------------ */

static void reclat(const double *lat, double *rec) {
  double v = det3(lat);
  cross3(lat + 3, lat + 6, rec + 0);
  cross3(lat + 6, lat + 0, rec + 3);
  cross3(lat + 0, lat + 3, rec + 6);
  for (int i = 0; i < 9; i++)
    rec[i] *= 2.0 * OR_PI / v;
}

/* projector.c:223-274: xvals = x[idx] dv e^{i k.path} once per point,
 * then <p_j|psi> = sum conj(values_j) xvals (zdotc) for each projector. */
void or_onto_projector(const double *grid, const double *lattice,
                       const double *kfrac, const int *fftg, int nsites,
                       const int *site_pt, const int *site_nproj,
                       const long *val_off, const int *proj_off,
                       const int *idx, const double *path,
                       const double *values, double *proj_out) {
  double dv = det3(lattice) / fftg[0] / fftg[1] / fftg[2];
  double rec[9], kc[3];
  reclat(lattice, rec);
  frac_to_cart(kfrac, rec, kc);
  int maxn = 0;
  for (int s = 0; s < nsites; s++)
    if (site_pt[s + 1] - site_pt[s] > maxn)
      maxn = site_pt[s + 1] - site_pt[s];
  double *xv = (double *)malloc(sizeof(double) * 2 * (size_t)(maxn > 0 ? maxn : 1));
  for (int s = 0; s < nsites; s++) {
    int p0 = site_pt[s], np = site_pt[s + 1] - site_pt[s];
    for (int i = 0; i < np; i++) {
      const double *pa = path + 3 * (long)(p0 + i);
      double kr = kc[0] * pa[0] + kc[1] * pa[1] + kc[2] * pa[2];
      double cr = cos(kr), ci = sin(kr);
      long g = idx[p0 + i];
      double xr = grid[2 * g] * dv, xi = grid[2 * g + 1] * dv;
      xv[2 * i] = xr * cr - xi * ci;
      xv[2 * i + 1] = xr * ci + xi * cr;
    }
    for (int p = 0; p < site_nproj[s]; p++) {
      double sr = 0.0, si = 0.0;
      const double *val = values + 2 * (val_off[s] + (long)p * np);
      for (int i = 0; i < np; i++) {
        double vr = val[2 * i], vi = val[2 * i + 1];
        sr += vr * xv[2 * i] + vi * xv[2 * i + 1];
        si += vr * xv[2 * i + 1] - vi * xv[2 * i];
      }
      proj_out[2 * (proj_off[s] + p)] = sr;
      proj_out[2 * (proj_off[s] + p) + 1] = si;
    }
  }
  free(xv);
}

void or_setup_projections(const double *lattice, const int *fftg, int nk,
                          const double *kfracs, const int *nw,
                          const int *const *Gs, const float *C, long ldc,
                          int nband, int nsites, const int *site_pt,
                          const int *site_nproj, const long *val_off,
                          const int *proj_off, int ptot, const int *idx,
                          const double *path, const double *values,
                          int npairs, const int *pairs, double *proj_out,
                          int nthreads) {
  long ng = (long)fftg[0] * fftg[1] * fftg[2];
  int total = pairs ? npairs : nk * nband;
#ifdef _OPENMP
  if (nthreads > 0)
    omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    fft_plan pl[3];
    for (int a = 0; a < 3; a++)
      plan_make(&pl[a], fftg[a]);
    double *work = fft_work(fftg);
    double *x = (double *)malloc(sizeof(double) * 2 * ng);
#pragma omp for schedule(dynamic, 1)
    for (int w = 0; w < total; w++) {
      int k, b;
      if (pairs) {
        k = pairs[2 * w];
        b = pairs[2 * w + 1];
      } else {
        /* projector.c:589-591 ordering: kpt = w % K, band = w / K */
        k = w % nk;
        b = w / nk;
      }
      const float *cb = C + 2 * (((long)k * nband + b) * ldc);
      fft3d_with(x, lattice, Gs[k], cb, nw[k], fftg, pl, work);
      or_onto_projector(x, lattice, kfracs + 3 * k, fftg, nsites, site_pt,
                        site_nproj, val_off, proj_off, idx, path, values,
                        proj_out + 2 * (((long)k * nband + b) * ptot));
    }
    free(x);
    free(work);
    for (int a = 0; a < 3; a++)
      plan_free(&pl[a]);
  }
}

/* ------------------------------------------------------------------ */
/* Band overlaps                                                        */
/* ------------------------------------------------------------------ */

void or_pseudoprojection(double *out, const float *C_ref, const float *C_proj,
                         long ldc, int nband_ref, int nband_proj, int nwk,
                         int nspin, const int *nw, int band, int flip,
                         int fp32acc) {
  int nkt = nwk * nspin;
#pragma omp parallel for
  for (int b = 0; b < nband_ref; b++) {
    for (int k = 0; k < nkt; k++) {
      int kp = k;
      if (nspin == 2 && flip)
        kp = (k < nwk) ? k + nwk : k - nwk;
      const float *a = C_ref + 2 * (((long)kp * nband_ref + b) * ldc);
      const float *p = C_proj + 2 * (((long)k * nband_proj + band) * ldc);
      /* pseudoprojector.c:85: length of the reference band at k */
      int n = nw[k];
      if (fp32acc) {
        float sr = 0.f, si = 0.f;
        for (int g = 0; g < n; g++) {
          sr += a[2 * g] * p[2 * g] + a[2 * g + 1] * p[2 * g + 1];
          si += a[2 * g] * p[2 * g + 1] - a[2 * g + 1] * p[2 * g];
        }
        out[2 * (b * nkt + k)] = sr;
        out[2 * (b * nkt + k) + 1] = si;
      } else {
        double sr = 0.0, si = 0.0;
        for (int g = 0; g < n; g++) {
          double ar = a[2 * g], ai = a[2 * g + 1];
          double pr = p[2 * g], pi = p[2 * g + 1];
          sr += ar * pr + ai * pi;
          si += ar * pi - ai * pr;
        }
        out[2 * (b * nkt + k)] = sr;
        out[2 * (b * nkt + k) + 1] = si;
      }
    }
  }
}

/* projector.c:850-910 (O_M only): for w = b_R * nkt + k,
 * overlap[w] += sum_{s in M} sum_{i,j: l_i=l_j, m_i=m_j}
 *               conj(P_R[s1][i]) dO[n_i][n_j] P_S[s2][j]. */
void or_compensation_terms_m(double *overlap, int band, int nwk, int nspin,
                             int nbR, const double *proj_R, int ptotR,
                             const int *proj_off_R, const int *labels_R,
                             const double *proj_S, int ptotS,
                             const int *proj_off_S, int num_M, const int *M_R,
                             const int *M_S, const int *sp_nchan,
                             const int *sp_meta_off, const int *meta_n,
                             const int *meta_l, const int *meta_m,
                             const int *sp_do_off, const double *dO, int flip) {
  int nkt = nwk * nspin;
  /* number of projectors of a species from its channel table */
#pragma omp parallel for
  for (int w = 0; w < nbR * nkt; w++) {
    int k = w % nkt;
    int kR = k;
    if (nspin == 2 && flip)
      kR = (kR < nwk) ? kR + nwk : kR - nwk;
    int bR = w / nkt;
    /* proj_S holds only the projected band's rows: [nkt][ptotS] */
    const double *pr = proj_R + 2 * (((long)kR * nbR + bR) * ptotR);
    const double *ps = proj_S + 2 * ((long)k * ptotS);
    double tr = 0.0, ti = 0.0;
    for (int s = 0; s < num_M; s++) {
      int s1 = M_R[s], s2 = M_S[s];
      int e = labels_R[s1];
      int mo = sp_meta_off[e];
      int npj = proj_off_R[s1 + 1] - proj_off_R[s1];
      int nch = sp_nchan[e];
      const double *D = dO + sp_do_off[e];
      for (int i = 0; i < npj; i++) {
        double ar = pr[2 * (proj_off_R[s1] + i)];
        double ai = -pr[2 * (proj_off_R[s1] + i) + 1];
        for (int j = 0; j < npj; j++) {
          if (meta_l[mo + i] != meta_l[mo + j] ||
              meta_m[mo + i] != meta_m[mo + j])
            continue;
          double dv = D[meta_n[mo + i] * nch + meta_n[mo + j]];
          double br = ps[2 * (proj_off_S[s2] + j)];
          double bi = ps[2 * (proj_off_S[s2] + j) + 1];
          tr += dv * (ar * br - ai * bi);
          ti += dv * (ar * bi + ai * br);
        }
      }
    }
    overlap[2 * w] += tr;
    overlap[2 * w + 1] += ti;
  }
}

/* projector.c:507-558: same-l spline integrals of partial-wave products. */
void or_pwave_overlaps(const double *grid, int n, int nchan, const int *ls,
                       const double *aewaves, const double *pswaves,
                       double *psov, double *aeov, double *diov) {
  double *prod = (double *)malloc(sizeof(double) * n);
  double *b = (double *)malloc(sizeof(double) * n);
  double *c = (double *)malloc(sizeof(double) * n);
  double *d = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < nchan * nchan; i++)
    psov[i] = aeov[i] = diov[i] = 0.0;
  for (int i = 0; i < nchan; i++)
    for (int j = i; j < nchan; j++) {
      if (ls[i] != ls[j])
        continue;
      const double *ps1 = pswaves + (long)i * n, *ps2 = pswaves + (long)j * n;
      const double *ae1 = aewaves + (long)i * n, *ae2 = aewaves + (long)j * n;
      for (int t = 0; t < 3; t++) {
        for (int g = 0; g < n; g++)
          prod[g] = t == 0   ? ps1[g] * ps2[g]
                    : t == 1 ? ae1[g] * ae2[g]
                             : (ae1[g] - ps1[g]) * (ae2[g] - ps2[g]);
        or_spline_coeff(grid, prod, n, b, c, d);
        double v = or_spline_integral(grid, prod, b, c, d, n);
        double *dst = t == 0 ? psov : t == 1 ? aeov : diov;
        dst[i * nchan + j] = v;
        dst[j * nchan + i] = v;
      }
    }
  free(prod);
  free(b);
  free(c);
  free(d);
}

/* ------------------------------------------------------------------ */
/* Spherical Bessel transform and the filtered partial-wave difference  */
/* ------------------------------------------------------------------ */

#define OR_C_EV 0.262465831

/* x <- sum_j x[j] exp(+2 pi i j k / n), unnormalised (the DFTI backward
 * transform of sbt.c:128-144); direct O(n^2) sum with exact twiddles. */
static void dft_backward(double *x, int n) {
  double *y = (double *)calloc(2 * (size_t)n, sizeof(double));
  for (int k = 0; k < n; k++) {
    long double re = 0.0L, im = 0.0L;
    for (int j = 0; j < n; j++) {
      long q = ((long)j * k) % n;
      long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)q / (long double)n;
      long double cs = cosl(a), sn = sinl(a);
      re += x[2 * j] * cs - x[2 * j + 1] * sn;
      im += x[2 * j] * sn + x[2 * j + 1] * cs;
    }
    y[2 * k] = (double)re;
    y[2 * k + 1] = (double)im;
  }
  memcpy(x, y, 2 * (size_t)n * sizeof(double));
  free(y);
}

typedef struct {
  int N, lmax;
  double *ks, *rs; /* N each */
  double *mult;    /* (lmax+1) x N complex */
} sbt_desc;

/* sbt.c:24-96: log grids r_i = r0 e^{(i-N/2) drho}, k_i = kmin e^{i drho}
 * (N = 2n, kmin = sqrt((encut+enbuf) c) e^{-(N/2-1) drho}) and the
 * multiplier table M_l(t_i), t_i = i 2 pi / (N drho), of Talman's
 * algorithm (J. Talman, CPC 180 (2009) 332): the phase of the Gamma-function
 * ratio from a Stirling series with ten shift terms, then the
 * recurrence M_{l+1} = e^{-2i atan(2t/(2l+1))} M_{l-1}. */
static void sbt_setup(sbt_desc *d, double encut, double enbuf, int lmax, int n, const double *r) {
  const int N = 2 * n;
  if (lmax == 0)
    lmax = 1;
  d->N = N;
  d->lmax = lmax;
  d->ks = (double *)malloc(sizeof(double) * N);
  d->rs = (double *)malloc(sizeof(double) * N);
  d->mult = (double *)calloc(2 * (size_t)(lmax + 1) * N, sizeof(double));
  const double drho = log(r[1] / r[0]);
  const double dt = 2.0 * OR_PI / N / drho;
  const double kmin = sqrt((encut + enbuf) * OR_C_EV) * exp(-(N / 2 - 1) * drho);
  const double kappamin = log(kmin);
  for (int i = 0; i < N; i++) {
    d->ks[i] = kmin * exp(i * drho);
    d->rs[i] = r[0] * exp((i - N / 2) * drho);
  }
  const double rhomin = log(d->rs[0]);
  for (int i = 0; i < N; i++) {
    const double t = i * dt;
    const double rad = sqrt(10.5 * 10.5 + t * t);
    const double ph = atan(2.0 * t / 21.0);
    double p1 = -10.0 * ph - t * log(rad) + t + sin(ph) / (12.0 * rad) - sin(3.0 * ph) / (360.0 * pow(rad, 3)) +
                sin(5.0 * ph) / (1260.0 * pow(rad, 5)) - sin(7.0 * ph) / (1680.0 * pow(rad, 7));
    for (int j = 1; j <= 10; j++)
      p1 += atan(2.0 * t / (2 * j - 1));
    const double p2 = -atan(tanh(OR_PI * t / 2.0));
    const double p3 = (kappamin + rhomin) * t;
    const double phase = p1 + p2 + p3;
    double amp = sqrt(OR_PI / 2.0) / N;
    if (i == 0)
      amp *= 0.5;
    double *m0 = d->mult + 2 * (size_t)i;
    m0[0] = amp * cos(phase);
    m0[1] = amp * sin(phase);
    /* M_1 = e^{2i(-p2 - atan 2t)} M_0 */
    const double a1 = 2.0 * (-p2 - atan(2.0 * t));
    double *m1 = d->mult + 2 * ((size_t)N + i);
    m1[0] = cos(a1) * m0[0] - sin(a1) * m0[1];
    m1[1] = sin(a1) * m0[0] + cos(a1) * m0[1];
    for (int l = 1; l < lmax; l++) {
      const double al = -2.0 * atan(2.0 * t / (2 * l + 1));
      const double *mp = d->mult + 2 * ((size_t)(l - 1) * N + i);
      double *mn = d->mult + 2 * ((size_t)(l + 1) * N + i);
      mn[0] = cos(al) * mp[0] - sin(al) * mp[1];
      mn[1] = sin(al) * mp[0] + cos(al) * mp[1];
    }
  }
}

static void sbt_free(sbt_desc *d) {
  free(d->ks);
  free(d->rs);
  free(d->mult);
}

/* Two backward DFTs around the multiplier: in[m] -> DFT -> x M_l, upper
 * half zeroed -> DFT (sbt.c:133-144, 190-200). */
static void sbt_core(const sbt_desc *d, double *x, int l) {
  const int N = d->N;
  dft_backward(x, N);
  const double *M = d->mult + 2 * (size_t)l * N;
  for (int i = 0; i < N; i++) {
    if (i >= N / 2) {
      x[2 * i] = x[2 * i + 1] = 0.0;
    } else {
      const double re = x[2 * i] * M[2 * i] - x[2 * i + 1] * M[2 * i + 1];
      const double im = x[2 * i] * M[2 * i + 1] + x[2 * i + 1] * M[2 * i];
      x[2 * i] = re;
      x[2 * i + 1] = im;
    }
  }
  dft_backward(x, N);
}

/* sbt.c:98-153: g(k_p) = int r^2 j_l(k_p r) R(r) dr for f = r R(r)
 * sampled on the n-point log grid r (below r[0], f is continued as
 * f[0] (r/r[0])^{l+1}).  Writes ks[n] and g[n]. */
void or_sbt_forward(double encut, double enbuf, int lmax, int n, const double *r, const double *f, int l,
                    double *ks_out, double *g) {
  sbt_desc d;
  sbt_setup(&d, encut, enbuf, lmax, n, r);
  const int N = d.N;
  double *x = (double *)calloc(2 * (size_t)N, sizeof(double));
  const double C = f[0] / pow(d.rs[N / 2], l + 1);
  for (int i = 0; i < N; i++) {
    const double fi = i < N / 2 ? C * pow(d.rs[i], l + 1) : f[i - N / 2];
    x[2 * i] = sqrt(d.rs[i]) * fi;
  }
  sbt_core(&d, x, l);
  for (int p = 0; p < n; p++) {
    ks_out[p] = d.ks[p];
    g[p] = x[2 * p] * 2.0 / pow(d.ks[p], 1.5);
  }
  free(x);
  sbt_free(&d);
}

/* sbt.c:155-213: R(r_p) = (2/pi) int k^2 j_l(k r_p) g(k) dk for g on the
 * k grid of the same descriptor; r_p = rs[n + p] (the input r grid). */
void or_sbt_inverse(double encut, double enbuf, int lmax, int n, const double *r, const double *g, int l,
                    double *rs_out, double *R) {
  sbt_desc d;
  sbt_setup(&d, encut, enbuf, lmax, n, r);
  const int N = d.N;
  double *x = (double *)calloc(2 * (size_t)N, sizeof(double));
  for (int i = 0; i < N / 2; i++)
    x[2 * i] = pow(d.ks[i], 1.5) * g[i];
  sbt_core(&d, x, l);
  for (int p = 0; p < n; p++) {
    rs_out[p] = d.rs[p + N / 2];
    R[p] = x[2 * (p + N / 2)] / OR_PI * 2.0 * 2.0 / pow(d.rs[p + N / 2], 1.5);
  }
  free(x);
  sbt_free(&d);
}

/* utils.c:477-489 (wave_interpolate): cubic spline on a log grid. */
double or_wave_interpolate(double r, int n, const double *x, const double *f, const double *b, const double *c,
                           const double *d) {
  if (r > x[n - 1])
    return 0.0;
  if (r < x[0])
    return f[0];
  int i = (int)(log(r / x[0]) / log(x[1] / x[0]));
  if (i > n - 2)
    i = n - 2;
  const double t = r - x[i];
  return f[i] + t * (b[i] + t * (c[i] + t * d[i]));
}

/* projector.c:114-163 for one partial-wave difference f = r (phi - phi~)
 * on the n-point log grid wave_grid: SBT onto the k grid (encut 1e7 eV),
 * keep k < sqrt(c grid_encut), transform back, and sample the result on
 * the linear grid j wave_rmax / np (j < np) of the sphere tables
 * (sdw[0] = 0 for l > 0, sdw[0] = sdw[1] for l = 0).  Also returns the
 * k grid and the unfiltered transform (kwave, used by the off-site terms). */
void or_smooth_diffwave(int n, const double *wave_grid, const double *diffwave, int l, int lmax,
                        double grid_encut, int np, double *sdw, double *kgrid, double *kwave) {
  or_sbt_forward(1e7, 0.0, lmax, n, wave_grid, diffwave, l, kgrid, kwave);
  double *kd = (double *)calloc((size_t)n, sizeof(double));
  const double kcut = sqrt(OR_C_EV * grid_encut);
  for (int q = 0; q < n && kgrid[q] < kcut; q++)
    kd[q] = kwave[q];
  double *rs = (double *)malloc(sizeof(double) * n);
  double *sm = (double *)malloc(sizeof(double) * n);
  or_sbt_inverse(1e7, 0.0, lmax, n, wave_grid, kd, l, rs, sm);
  /* the dense grid of projector.c:62-69 (scale 1): wave_grid[0] times
   * powers of wave_grid[1] / wave_grid[0] */
  double *dg = (double *)malloc(sizeof(double) * n);
  dg[0] = wave_grid[0];
  const double fac = wave_grid[1] / wave_grid[0];
  for (int p = 1; p < n; p++)
    dg[p] = dg[p - 1] * fac;
  double *b = (double *)malloc(sizeof(double) * n);
  double *c = (double *)malloc(sizeof(double) * n);
  double *d = (double *)malloc(sizeof(double) * n);
  or_spline_coeff(dg, sm, n, b, c, d);
  const double wave_rmax = wave_grid[n - 1];
  if (l > 0)
    sdw[0] = 0.0;
  for (int p = 1; p < np; p++)
    sdw[p] = or_wave_interpolate(wave_rmax / np * p, n, dg, sm, b, c, d);
  if (l == 0)
    sdw[0] = np > 1 ? sdw[1] : 0.0;
  free(kd);
  free(rs);
  free(sm);
  free(dg);
  free(b);
  free(c);
  free(d);
}

/* projector.c:912-938 (O_R and O_S): for w = b_R * nkt + k,
 * overlap[w] += sum_{s < num_N_R} sum_i WP_S[s][i] conj(P_R[N_R[s]][i])
 *             + sum_{s < num_N_S} sum_i conj(WP_R[s][i]) P_S[N_S[s]][i].
 * P_R, WP_R: [nkt][nbR][ptot]; WP_S, P_S: the projected band's rows
 * [nkt][ptot].  WP_* site s occupies wp_off[s] .. wp_off[s+1]. */
void or_compensation_terms_rs(double *overlap, int nwk, int nspin, int nbR,
                              const double *proj_R, int ptotR,
                              const int *proj_off_R, int num_N_R,
                              const int *N_R, const double *wp_S,
                              int wp_ptot_S, const int *wp_off_S,
                              const double *wp_R, int wp_ptot_R,
                              const int *wp_off_R, const double *proj_S,
                              int ptotS, const int *proj_off_S, int num_N_S,
                              const int *N_S, int flip) {
  int nkt = nwk * nspin;
#pragma omp parallel for
  for (int w = 0; w < nbR * nkt; w++) {
    int k = w % nkt;
    int kR = k;
    if (nspin == 2 && flip)
      kR = (kR < nwk) ? kR + nwk : kR - nwk;
    int bR = w / nkt;
    double tr = 0.0, ti = 0.0;
    const double *pr = proj_R + 2 * (((long)kR * nbR + bR) * ptotR);
    const double *ws = wp_S + 2 * ((long)k * wp_ptot_S);
    for (int s = 0; s < num_N_R; s++) {
      int site = N_R[s];
      int np = wp_off_S[s + 1] - wp_off_S[s];
      for (int i = 0; i < np; i++) {
        double ar = ws[2 * (wp_off_S[s] + i)], ai = ws[2 * (wp_off_S[s] + i) + 1];
        double br = pr[2 * (proj_off_R[site] + i)], bi = -pr[2 * (proj_off_R[site] + i) + 1];
        tr += ar * br - ai * bi;
        ti += ar * bi + ai * br;
      }
    }
    const double *wr = wp_R + 2 * (((long)kR * nbR + bR) * wp_ptot_R);
    const double *ps = proj_S + 2 * ((long)k * ptotS);
    for (int s = 0; s < num_N_S; s++) {
      int site = N_S[s];
      int np = wp_off_R[s + 1] - wp_off_R[s];
      for (int i = 0; i < np; i++) {
        double ar = wr[2 * (wp_off_R[s] + i)], ai = -wr[2 * (wp_off_R[s] + i) + 1];
        double br = ps[2 * (proj_off_S[site] + i)], bi = ps[2 * (proj_off_S[site] + i) + 1];
        tr += ar * br - ai * bi;
        ti += ar * bi + ai * br;
      }
    }
    overlap[2 * w] += tr;
    overlap[2 * w + 1] += ti;
  }
}

/* ------------------------------------------------------------------ */
/* Off-site partial-wave overlaps (O_N)                                 */
/* ------------------------------------------------------------------ */

/* Wigner 3j symbol (j1 j2 j3; m1 m2 m3) for integer arguments, Racah's
 * closed form. */
double or_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0 || j3 < abs(j1 - j2) || j3 > j1 + j2 || abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3)
    return 0.0;
  double tri = fact(j1 + j2 - j3) * fact(j1 - j2 + j3) * fact(-j1 + j2 + j3) / fact(j1 + j2 + j3 + 1);
  double pre = sqrt(tri * fact(j1 + m1) * fact(j1 - m1) * fact(j2 + m2) * fact(j2 - m2) * fact(j3 + m3) *
                    fact(j3 - m3));
  int kmin = 0;
  if (j2 - j3 - m1 > kmin)
    kmin = j2 - j3 - m1;
  if (j1 - j3 + m2 > kmin)
    kmin = j1 - j3 + m2;
  int kmax = j1 + j2 - j3;
  if (j1 - m1 < kmax)
    kmax = j1 - m1;
  if (j2 + m2 < kmax)
    kmax = j2 + m2;
  double sum = 0.0;
  for (int k = kmin; k <= kmax; k++) {
    double den = fact(k) * fact(j1 + j2 - j3 - k) * fact(j1 - m1 - k) * fact(j2 + m2 - k) * fact(j3 - j2 + m1 + k) *
                 fact(j3 - j1 - m2 + k);
    sum += ((k % 2) ? -1.0 : 1.0) / den;
  }
  int ph = j1 - j2 - m3;
  return (((ph % 2) + 2) % 2 ? -1.0 : 1.0) * pre * sum;
}

/* gaunt.py:17-27: SBTFACS[l1][l2][(l3-|l1-l2|)/2][l1+m1][m2] =
 * (l1 l2 l3; 0 0 0)(l1 l2 l3; -m1 m2 m1-m2) sqrt((2l1+1)(2l2+1)(2l3+1)/4pi). */
double or_sbtfac(int l1, int l2, int l3, int m1, int m2) {
  return or_wigner3j(l1, l2, l3, 0, 0, 0) * or_wigner3j(l1, l2, l3, -m1, m2, m1 - m2) *
         sqrt((2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / 4.0 / OR_PI);
}

/* utils.c:807-827: j_l(x) by upward recurrence, 1 or 0 below 1e-5. */
static double sbf(double x, int l) {
  if (x < 10e-6)
    return l == 0 ? 1.0 : 0.0;
  double jm = sin(x) / x, j = sin(x) / (x * x) - cos(x) / x, jp = 0.0;
  if (l == 0)
    return jm;
  if (l == 1)
    return j;
  for (int ll = 1; ll < l; ll++) {
    jp = (2 * ll + 1) / x * j - jm;
    jm = j;
    j = jp;
  }
  return jp;
}

/* radial.c:116-196 (reciprocal_offsite_wave_overlap): overlap of the
 * partial-wave differences (l1, m1) at the origin and (l2, m2) at dcoord
 * (cartesian) from their transforms f1, f2 on the log k grids k1, k2
 * (spline coefficients b, c, d):
 *   sum_{L = |l1-l2|, step 2}^{l1+l2} I_L F(L) Y_L^{m1-m2}(dcoord) i^{l2+L-l1} 8 (-1)^m1,
 * I_L = int k^2 f1 f2 j_L(k R) dk by spline integration over 500 log
 * points of [max(k1[0], k2[0]), min(k1[-1], k2[-1])]; at R = 0 only
 * L = 0, l1 = l2, m1 = m2 contributes, as (2/pi) I_0.  out2 = (re, im). */
void or_offsite_overlap(const double *dcoord, const double *k1, const double *f1, const double *b1,
                        const double *c1, const double *d1, int n1, const double *k2, const double *f2,
                        const double *b2, const double *c2, const double *d2, int n2, int l1, int m1, int l2,
                        int m2, double *out2) {
  enum { NK = 500 };
  int lx = l1, ly = l2, mx = m1, my = m2;
  if (l1 < l2) {
    lx = l2;
    ly = l1;
    mx = m2;
    my = m1;
  }
  if (my < 0) {
    mx = -mx;
    my = -my;
  }
  double kmax = k1[n1 - 1] < k2[n2 - 1] ? k1[n1 - 1] : k2[n2 - 1];
  double kmin = k1[0] > k2[0] ? k1[0] : k2[0];
  double R = sqrt(dcoord[0] * dcoord[0] + dcoord[1] * dcoord[1] + dcoord[2] * dcoord[2]);
  double theta = 0.0, phi = 0.0;
  if (R < 10e-12) {
    R = 0.0;
  } else {
    theta = acos(dcoord[2] / R);
    if (R - fabs(dcoord[2]) >= 10e-12)
      phi = acos(dcoord[0] / sqrt(dcoord[0] * dcoord[0] + dcoord[1] * dcoord[1]));
    if (dcoord[1] < 0)
      phi = 2.0 * OR_PI - phi;
  }
  double kg[NK], fn[NK], bb[NK], cc[NK], dd[NK];
  double tr = 0.0, ti = 0.0;
  const double mult = ((m1 % 2) ? -8.0 : 8.0);
  for (int L = abs(l1 - l2); L <= l1 + l2; L += 2) {
    for (int q = 0; q < NK; q++) {
      kg[q] = kmin * pow(kmax / kmin, (double)q / NK);
      fn[q] = or_wave_interpolate(kg[q], n1, k1, f1, b1, c1, d1) * or_wave_interpolate(kg[q], n2, k2, f2, b2, c2, d2) *
              kg[q] * kg[q] * sbf(kg[q] * R, L);
    }
    or_spline_coeff(kg, fn, NK, bb, cc, dd);
    double I = or_spline_integral(kg, fn, bb, cc, dd, NK);
    if (R > 10e-10) {
      double y[2];
      or_ylm(L, m1 - m2, theta, phi, y);
      double f = I * or_sbtfac(lx, ly, L, mx, my) * mult;
      /* i^(l2 + L - l1) */
      int p = (((l2 + L - l1) % 4) + 4) % 4;
      double ir = p == 0 ? 1.0 : p == 2 ? -1.0 : 0.0, ii = p == 1 ? 1.0 : p == 3 ? -1.0 : 0.0;
      tr += f * (y[0] * ir - y[1] * ii);
      ti += f * (y[0] * ii + y[1] * ir);
    } else if (L == 0 && l1 == l2 && m1 == m2) {
      tr += I * 2.0 / OR_PI;
    }
  }
  out2[0] = tr;
  out2[1] = ti;
}

/* projector.c:940-959 (O_N): for w = b_R * nkt + k,
 * overlap[w] += sum_s sum_ij conj(P_R[N_RS_R[s]][i]) Ov_s[i][j]
 *                         P_S[N_RS_S[s]][j] exp(2 pi i kR . dcoord_s),
 * kR the (possibly spin-flipped) k of R in the units the reference mixes.
 * Ov: per pair s, P_R_s x P_S_s complex at ov_off[s]; kpts [nkt][3]. */
void or_compensation_terms_n(double *overlap, int nwk, int nspin, int nbR, const double *proj_R, int ptotR,
                             const int *proj_off_R, const double *proj_S, int ptotS, const int *proj_off_S,
                             int num_N_RS, const int *N_RS_R, const int *N_RS_S, const double *ov,
                             const long *ov_off, const double *dcoords, const double *kpts, int flip) {
  int nkt = nwk * nspin;
#pragma omp parallel for
  for (int w = 0; w < nbR * nkt; w++) {
    int k = w % nkt;
    int kR = k;
    if (nspin == 2 && flip)
      kR = (kR < nwk) ? kR + nwk : kR - nwk;
    int bR = w / nkt;
    const double *pr = proj_R + 2 * (((long)kR * nbR + bR) * ptotR);
    const double *ps = proj_S + 2 * ((long)k * ptotS);
    double tr = 0.0, ti = 0.0;
    for (int s = 0; s < num_N_RS; s++) {
      int s1 = N_RS_R[s], s2 = N_RS_S[s];
      int n1 = proj_off_R[s1 + 1] - proj_off_R[s1], n2 = proj_off_S[s2 + 1] - proj_off_S[s2];
      const double *dc = dcoords + 3 * s;
      double a = 2.0 * OR_PI * (kpts[3 * kR] * dc[0] + kpts[3 * kR + 1] * dc[1] + kpts[3 * kR + 2] * dc[2]);
      double er = cos(a), ei = sin(a);
      const double *O = ov + 2 * ov_off[s];
      for (int i = 0; i < n1; i++) {
        double ar = pr[2 * (proj_off_R[s1] + i)], ai = -pr[2 * (proj_off_R[s1] + i) + 1];
        for (int j = 0; j < n2; j++) {
          double orr = O[2 * (i * n2 + j)], oi = O[2 * (i * n2 + j) + 1];
          double br = ps[2 * (proj_off_S[s2] + j)], bi = ps[2 * (proj_off_S[s2] + j) + 1];
          /* conj(P_R) Ov P_S e */
          double xr = ar * orr - ai * oi, xi = ar * oi + ai * orr;
          double yr = xr * br - xi * bi, yi = xr * bi + xi * br;
          tr += yr * er - yi * ei;
          ti += yr * ei + yi * er;
        }
      }
    }
    overlap[2 * w] += tr;
    overlap[2 * w + 1] += ti;
  }
}

/* ------------------------------------------------------------------ */
/* aug_recip: plane-wave coefficients of the augmentation functions     */
/* ------------------------------------------------------------------ */

/* projector.c:276-330 + 420-453 + linalg.c:47-79 for one band:
 * x(r) = sum_s sum_p P[proj_off_P[s] + p] values_p(r) e^{-i k.path(r)} over
 * the site tables (values = p(r) as stored, not conjugated), then the
 * forward transform with scale sqrt(V)/N^3 sampled at the band's G
 * vectors: CA[w] = sqrt(V)/N^3 sum_r x(r) e^{-2 pi i G.r/N}.  P holds the
 * band's projections at the sites of the tables (site s at proj_off_P[s]).
 * CA is returned in FP64 (the reference stores it as float complex). */
void or_aug_freqs(double *CA, const double *lattice, const double *kfrac, const int *fftg, const int *Gs, int nw,
                  int nsites, const int *site_pt, const int *site_nproj, const long *val_off, const int *idx,
                  const double *path, const double *values, const double *P, const int *proj_off_P) {
  const long n = (long)fftg[0] * fftg[1] * fftg[2];
  double *x = (double *)calloc(2 * (size_t)n, sizeof(double));
  double rec[9], kc[3];
  reclat(lattice, rec);
  frac_to_cart(kfrac, rec, kc);
  for (int s = 0; s < nsites; s++) {
    int p0 = site_pt[s], np = site_pt[s + 1] - site_pt[s];
    for (int i = 0; i < np; i++) {
      const double *pa = path + 3 * (long)(p0 + i);
      double kr = -(kc[0] * pa[0] + kc[1] * pa[1] + kc[2] * pa[2]);
      double er = cos(kr), ei = sin(kr);
      double ar = 0.0, ai = 0.0;
      for (int p = 0; p < site_nproj[s]; p++) {
        const double *v = values + 2 * (val_off[s] + (long)p * np + i);
        const double *pp = P + 2 * (proj_off_P[s] + p);
        ar += pp[0] * v[0] - pp[1] * v[1];
        ai += pp[0] * v[1] + pp[1] * v[0];
      }
      long g = idx[p0 + i];
      x[2 * g] += ar * er - ai * ei;
      x[2 * g + 1] += ar * ei + ai * er;
    }
  }
  /* forward DFT = conj(backward DFT(conj x)) */
  for (long i = 0; i < n; i++)
    x[2 * i + 1] = -x[2 * i + 1];
  or_ifft3d_inplace(x, fftg);
  const double scl = sqrt(det3(lattice)) / (double)n;
  for (int w = 0; w < nw; w++) {
    int g1 = ((Gs[3 * w] % fftg[0]) + fftg[0]) % fftg[0];
    int g2 = ((Gs[3 * w + 1] % fftg[1]) + fftg[1]) % fftg[1];
    int g3 = ((Gs[3 * w + 2] % fftg[2]) + fftg[2]) % fftg[2];
    long g = ((long)g1 * fftg[1] + g2) * fftg[2] + g3;
    CA[2 * w] = x[2 * g] * scl;
    CA[2 * w + 1] = -x[2 * g + 1] * scl;
  }
  free(x);
}

/* ------------------------------------------------------------------ */
/* All-electron real-space states (density.c)                          */
/* ------------------------------------------------------------------ */

/* utils.c:522-545 (wave_value2): (f(|pos|) / |pos|) Y_lm(pos) with f the
 * log-grid spline (r below the first grid point divides by that point). */
void or_wave_value2(int n, const double *x, const double *f, const double *b,
                    const double *c, const double *d, int l, int m,
                    const double *pos, double *out2) {
  double r = norm3(pos);
  double radial = or_wave_interpolate(r, n, x, f, b, c, d);
  radial /= (r < x[0]) ? x[0] : r;
  double y[2];
  if (r == 0.0) {
    or_ylm(l, m, 0.0, 0.0, y);
  } else {
    double theta = acos(pos[2] / r);
    double phi = 0.0;
    if (r - fabs(pos[2]) != 0.0)
      phi = acos(pos[0] / sqrt(pos[0] * pos[0] + pos[1] * pos[1]));
    if (pos[1] < 0)
      phi = 2.0 * OR_PI - phi;
    or_ylm(l, m, theta, phi, y);
  }
  out2[0] = radial * y[0];
  out2[1] = radial * y[1];
}

/* density.c:232-310 (realspace_state) after its fft3d: x (c128 [N^3],
 * x slowest) holds the transformed pseudo state on entry.  Each point is
 * multiplied by e^{2 pi i k.frac} (:242-255); then every site s adds, at
 * the points of its box (site_box around round(coord N)) with
 * |r - R| < rmax_s (the last point of its wave grid; r the unwrapped box
 * point, :277-284), sum_n wave_value2(diffwave of channel n, l_n, m_n,
 * r - R) proj[proj_off[s] + n] e^{2 pi i k.(R + T)} with T the lattice
 * vector that wraps the point into the cell (:285-303).  The projectors
 * of a site run over its channels, m = -l..l within each.
 * Species e: n_e = wn[e] wave-grid points at wg + woff[e]; channels
 * chan_off[e] .. chan_off[e+1] with l = chan_l[ch] and diffwave spline
 * f, b, c, d at (long)ch * wmax (wmax >= every n_e). */
void or_realspace_state(double *x, const double *lattice, const int *fftg,
                        const double *kfrac, int nsites, const int *labels,
                        const double *coords, const int *wn, const long *woff,
                        const double *wg, const int *chan_off,
                        const int *chan_l, int wmax, const double *f,
                        const double *b, const double *c, const double *d,
                        const int *proj_off, const double *proj) {
  const int n0 = fftg[0], n1 = fftg[1], n2 = fftg[2];
  for (int i = 0; i < n0; i++)
    for (int j = 0; j < n1; j++)
      for (int k = 0; k < n2; k++) {
        double kr = kfrac[0] * i / n0 + kfrac[1] * j / n1 + kfrac[2] * k / n2;
        double ph[2] = {cos(2.0 * OR_PI * kr), sin(2.0 * OR_PI * kr)};
        double *v = x + 2 * (((long)i * n1 + j) * n2 + k);
        double re = v[0] * ph[0] - v[1] * ph[1];
        double im = v[0] * ph[1] + v[1] * ph[0];
        v[0] = re;
        v[1] = im;
      }
  for (int s = 0; s < nsites; s++) {
    const int e = labels[s];
    const int n = wn[e];
    const double *xg = wg + woff[e];
    const double rmax = xg[n - 1];
    const double *R = coords + 3 * s;
    int g[3];
    site_box(lattice, fftg, rmax, g);
    int c0 = (int)round(R[0] * n0), c1 = (int)round(R[1] * n1), c2 = (int)round(R[2] * n2);
    for (int i = c0 - g[0]; i <= c0 + g[0]; i++)
      for (int j = c1 - g[1]; j <= c1 + g[1]; j++)
        for (int k = c2 - g[2]; k <= c2 + g[2]; k++) {
          double fv[3] = {(double)i / n0 - R[0], (double)j / n1 - R[1], (double)k / n2 - R[2]};
          double cv[3];
          frac_to_cart(fv, lattice, cv);
          if (!(norm3(cv) < rmax))
            continue;
          int ii = (i % n0 + n0) % n0, jj = (j % n1 + n1) % n1, kk = (k % n2 + n2) % n2;
          /* (ii - i) / N is exact: the wrap is a whole number of cells */
          double pc[3] = {R[0] + (ii - i) / n0, R[1] + (jj - j) / n1, R[2] + (kk - k) / n2};
          double phase = kfrac[0] * pc[0] + kfrac[1] * pc[1] + kfrac[2] * pc[2];
          double e2[2] = {cos(2.0 * OR_PI * phase), sin(2.0 * OR_PI * phase)};
          double acc[2] = {0.0, 0.0};
          int p = proj_off[s];
          for (int ch = chan_off[e]; ch < chan_off[e + 1]; ch++) {
            int l = chan_l[ch];
            long o = (long)ch * wmax;
            for (int m = -l; m <= l; m++, p++) {
              double w[2];
              or_wave_value2(n, xg, f + o, b + o, c + o, d + o, l, m, cv, w);
              double pr = proj[2 * p], pi = proj[2 * p + 1];
              double t0 = w[0] * pr - w[1] * pi, t1 = w[0] * pi + w[1] * pr;
              acc[0] += t0 * e2[0] - t1 * e2[1];
              acc[1] += t0 * e2[1] + t1 * e2[0];
            }
          }
          double *v = x + 2 * (((long)ii * n1 + jj) * n2 + kk);
          v[0] += acc[0];
          v[1] += acc[1];
        }
  }
}

/* density.c:312-326 (remove_phase): x *= e^{-2 pi i k.frac}. */
void or_remove_phase(double *x, const int *fftg, const double *kfrac) {
  const int n0 = fftg[0], n1 = fftg[1], n2 = fftg[2];
  for (int i = 0; i < n0; i++)
    for (int j = 0; j < n1; j++)
      for (int k = 0; k < n2; k++) {
        double kr = kfrac[0] * i / n0 + kfrac[1] * j / n1 + kfrac[2] * k / n2;
        double ph[2] = {cos(2.0 * OR_PI * kr), -sin(2.0 * OR_PI * kr)};
        double *v = x + 2 * (((long)i * n1 + j) * n2 + k);
        double re = v[0] * ph[0] - v[1] * ph[1];
        double im = v[0] * ph[1] + v[1] * ph[0];
        v[0] = re;
        v[1] = im;
      }
}
