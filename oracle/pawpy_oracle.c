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

/* One 1D plan: factors of n (4,2,3,5,7, then any prime) and the table
 * w[j] = exp(+2 pi i j / n). */
typedef struct {
  int n, nf, f[32];
  double *w;
} fft_plan;

static void plan_make(fft_plan *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) {
    p->f[p->nf++] = 4;
    m /= 4;
  }
  for (int q = 2; m > 1 && q <= m; q++)
    while (m % q == 0) {
      p->f[p->nf++] = q;
      m /= q;
    }
  p->w = (double *)malloc(sizeof(double) * 2 * n);
  for (int j = 0; j < n; j++) {
    p->w[2 * j] = cos(2.0 * OR_PI * j / n);
    p->w[2 * j + 1] = sin(2.0 * OR_PI * j / n);
  }
}

static void plan_free(fft_plan *p) { free(p->w); }

/* Recursive decimation in time: out[k] = sum_j in[j*stride] e^{+2 pi i jk/n}.
 * level selects the radix; tmp has room for n complex values. */
static void fft_rec(const fft_plan *p, int level, int n, const double *in,
                    long stride, double *out, double *tmp) {
  if (n == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  int r = p->f[level];
  int m = n / r;
  /* sub-transforms of the r decimated sequences into out[q*m ...] */
  for (int q = 0; q < r; q++)
    fft_rec(p, level + 1, m, in + 2 * q * stride, stride * r, out + 2 * q * m,
            tmp);
  /* combine: X[k + s m] = sum_q W_n^{q k} W_r^{q s} Y_q[k] */
  long wstep = p->n / n; /* table stride for W_n */
  for (int k = 0; k < m; k++) {
    double y[2 * 64];
    for (int q = 0; q < r; q++) {
      long e = (wstep * (long)q * k) % p->n;
      double wr = p->w[2 * e], wi = p->w[2 * e + 1];
      double ar = out[2 * (q * m + k)], ai = out[2 * (q * m + k) + 1];
      y[2 * q] = ar * wr - ai * wi;
      y[2 * q + 1] = ar * wi + ai * wr;
    }
    for (int s = 0; s < r; s++) {
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < r; q++) {
        long e = (wstep * (long)m * ((long)q * s % r)) % p->n;
        double wr = p->w[2 * e], wi = p->w[2 * e + 1];
        sr += y[2 * q] * wr - y[2 * q + 1] * wi;
        si += y[2 * q] * wi + y[2 * q + 1] * wr;
      }
      tmp[2 * (k + s * m)] = sr;
      tmp[2 * (k + s * m) + 1] = si;
    }
  }
  memcpy(out, tmp, sizeof(double) * 2 * n);
}

static void fft_line(const fft_plan *p, double *line, double *work) {
  /* work: 3n complex (input copy, output, combine scratch) */
  int n = p->n;
  double *in = work, *out = work + 2 * n, *tmp = work + 4 * n;
  memcpy(in, line, sizeof(double) * 2 * n);
  fft_rec(p, 0, n, in, 1, out, tmp);
  memcpy(line, out, sizeof(double) * 2 * n);
}

static void ifft3d_with(double *x, const int *fftg, fft_plan *pl,
                        double *work) {
  int n0 = fftg[0], n1 = fftg[1], n2 = fftg[2];
  int nmax = n0 > n1 ? n0 : n1;
  nmax = nmax > n2 ? nmax : n2;
  double *line = work + 6 * nmax;
  /* z lines (contiguous) */
  for (long a = 0; a < (long)n0 * n1; a++)
    fft_line(&pl[2], x + 2 * a * n2, work);
  /* y lines */
  for (int i = 0; i < n0; i++)
    for (int k = 0; k < n2; k++) {
      for (int j = 0; j < n1; j++) {
        long o = 2 * (((long)i * n1 + j) * n2 + k);
        line[2 * j] = x[o];
        line[2 * j + 1] = x[o + 1];
      }
      fft_line(&pl[1], line, work);
      for (int j = 0; j < n1; j++) {
        long o = 2 * (((long)i * n1 + j) * n2 + k);
        x[o] = line[2 * j];
        x[o + 1] = line[2 * j + 1];
      }
    }
  /* x lines */
  for (int j = 0; j < n1; j++)
    for (int k = 0; k < n2; k++) {
      for (int i = 0; i < n0; i++) {
        long o = 2 * (((long)i * n1 + j) * n2 + k);
        line[2 * i] = x[o];
        line[2 * i + 1] = x[o + 1];
      }
      fft_line(&pl[0], line, work);
      for (int i = 0; i < n0; i++) {
        long o = 2 * (((long)i * n1 + j) * n2 + k);
        x[o] = line[2 * i];
        x[o + 1] = line[2 * i + 1];
      }
    }
}

void or_ifft3d_inplace(double *x, const int *fftg) {
  fft_plan pl[3];
  for (int a = 0; a < 3; a++)
    plan_make(&pl[a], fftg[a]);
  int nmax = fftg[0];
  if (fftg[1] > nmax)
    nmax = fftg[1];
  if (fftg[2] > nmax)
    nmax = fftg[2];
  double *work = (double *)malloc(sizeof(double) * 2 * 4 * nmax);
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
  int nmax = fftg[0];
  if (fftg[1] > nmax)
    nmax = fftg[1];
  if (fftg[2] > nmax)
    nmax = fftg[2];
  double *work = (double *)malloc(sizeof(double) * 2 * 4 * nmax);
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

/* projector.c:223-274: xvals = x[idx] dv e^{i k.path};
 * <p_j|psi> = sum conj(values_j) xvals (zdotc). */
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
  for (int s = 0; s < nsites; s++) {
    int p0 = site_pt[s], np = site_pt[s + 1] - site_pt[s];
    for (int p = 0; p < site_nproj[s]; p++) {
      double sr = 0.0, si = 0.0;
      const double *val = values + 2 * (val_off[s] + (long)p * np);
      for (int i = 0; i < np; i++) {
        const double *pa = path + 3 * (long)(p0 + i);
        double kr = kc[0] * pa[0] + kc[1] * pa[1] + kc[2] * pa[2];
        double cr = cos(kr), ci = sin(kr);
        long g = idx[p0 + i];
        double xr = grid[2 * g] * dv, xi = grid[2 * g + 1] * dv;
        double yr = xr * cr - xi * ci, yi = xr * ci + xi * cr;
        double vr = val[2 * i], vi = val[2 * i + 1];
        sr += vr * yr + vi * yi;
        si += vr * yi - vi * yr;
      }
      proj_out[2 * (proj_off[s] + p)] = sr;
      proj_out[2 * (proj_off[s] + p) + 1] = si;
    }
  }
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
  int nmax = fftg[0];
  if (fftg[1] > nmax)
    nmax = fftg[1];
  if (fftg[2] > nmax)
    nmax = fftg[2];
#ifdef _OPENMP
  if (nthreads > 0)
    omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    fft_plan pl[3];
    for (int a = 0; a < 3; a++)
      plan_make(&pl[a], fftg[a]);
    double *work = (double *)malloc(sizeof(double) * 2 * 4 * nmax);
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
