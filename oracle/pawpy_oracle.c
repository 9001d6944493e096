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
 * transform (the reference uses MKL DFTI).
 */
#include "pawpy_oracle.h"

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

/* work: 4 * nmax * nmax doubles (two [n][nb] complex buffers) */
static void ifft3d_with(double *x, const int *fftg, fft_plan *pl, double *work) {
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
