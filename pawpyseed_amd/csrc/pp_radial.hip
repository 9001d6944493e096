// Per-species radial setup on the host: the spherical Bessel transform of
// the partial-wave differences and their plane-wave-filtered form.
//
// Replaces sbt.c:24-213 (Talman's NumSBT, J. Talman, Comput. Phys.
// Commun. 180 (2009) 332) and the filter of projector.c:114-163.  The
// transform pairs are
//   g(k) = int r^2 j_l(k r) R(r) dr          (input f = r R(r)),
//   R(r) = (2/pi) int k^2 j_l(k r) g(k) dk,
// on log grids r_i = r0 e^{(i - n) drho}, k_i = kmin e^{i drho} of 2n
// points (the input grid is the upper half of r).  Each transform is two
// length-2n DFTs around the multiplier M_l(t) = phase of the Gamma-function
// ratio of the Mellin kernel, evaluated from a Stirling series.  The DFTs
// use Bluestein's chirp-z form over power-of-two FFTs, so any n works in
// O(n log n).  This is setup work, run once per species.
#include <cmath>
#include <complex>
#include <vector>

#include "pp_internal.h"

namespace pp {

using cd = std::complex<double>;
static const double kPi = 3.14159265358979323846;
static const double kCev = 0.262465831;  // 2 m_e / hbar^2 in 1/(eV A^2)

// in-place radix-2 FFT, sign +1: e^{+2 pi i jk/M}, sign -1: e^{-...}
static void fft_pow2(std::vector<cd> &a, int sign) {
  const size_t M = a.size();
  for (size_t i = 1, j = 0; i < M; i++) {
    size_t bit = M >> 1;
    for (; j & bit; bit >>= 1)
      j ^= bit;
    j ^= bit;
    if (i < j)
      std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= M; len <<= 1) {
    const double ang = sign * 2.0 * kPi / (double)len;
    for (size_t i = 0; i < M; i += len)
      for (size_t q = 0; q < len / 2; q++) {
        const cd w = std::polar(1.0, ang * (double)q);
        const cd u = a[i + q], v = a[i + q + len / 2] * w;
        a[i + q] = u + v;
        a[i + q + len / 2] = u - v;
      }
  }
}

// x_k <- sum_j x_j e^{+2 pi i jk/N} (unnormalised), any N, via
// jk = (j^2 + k^2 - (k-j)^2)/2 and one power-of-two convolution.
static void dft_backward(std::vector<cd> &x) {
  const long N = (long)x.size();
  size_t M = 1;
  while (M < (size_t)(2 * N - 1))
    M <<= 1;
  std::vector<cd> chirp(N);
  for (long j = 0; j < N; j++)
    chirp[j] = std::polar(1.0, kPi * (double)((j * j) % (2 * N)) / (double)N);
  std::vector<cd> a(M, cd(0, 0)), b(M, cd(0, 0));
  for (long j = 0; j < N; j++)
    a[j] = x[j] * chirp[j];
  b[0] = std::conj(chirp[0]);
  for (long m = 1; m < N; m++)
    b[m] = b[M - m] = std::conj(chirp[m]);
  fft_pow2(a, -1);
  fft_pow2(b, -1);
  for (size_t i = 0; i < M; i++)
    a[i] *= b[i];
  fft_pow2(a, +1);
  for (long k = 0; k < N; k++)
    x[k] = chirp[k] * a[k] / (double)M;
}

struct SbtPlan {
  int N = 0, lmax = 0;
  std::vector<double> ks, rs;
  std::vector<cd> mult;  // [(lmax+1) * N]
};

// Grids and multiplier table of sbt.c:24-96 for an n-point log grid r.
static void sbt_plan(SbtPlan &p, double encut, double enbuf, int lmax, int n, const double *r) {
  const int N = 2 * n;
  if (lmax < 1)
    lmax = 1;
  p.N = N;
  p.lmax = lmax;
  p.ks.resize(N);
  p.rs.resize(N);
  p.mult.assign((size_t)(lmax + 1) * N, cd(0, 0));
  const double drho = std::log(r[1] / r[0]);
  const double dt = 2.0 * kPi / N / drho;
  const double kmin = std::sqrt((encut + enbuf) * kCev) * std::exp(-(N / 2 - 1) * drho);
  for (int i = 0; i < N; i++) {
    p.ks[i] = kmin * std::exp(i * drho);
    p.rs[i] = r[0] * std::exp((i - N / 2) * drho);
  }
  const double shift = std::log(kmin) + std::log(p.rs[0]);  // kappa_min + rho_min
  for (int i = 0; i < N; i++) {
    const double t = i * dt;
    // arg Gamma(10.5 + i t) by Stirling's series about 10.5 + i t, then
    // shifted down to 0.5 + i t by ten atan terms
    const double mod = std::hypot(10.5, t);
    const double th = std::atan(t / 10.5);
    double arg = -10.0 * th - t * std::log(mod) + t + std::sin(th) / (12.0 * mod) -
                 std::sin(3.0 * th) / (360.0 * mod * mod * mod) +
                 std::sin(5.0 * th) / (1260.0 * std::pow(mod, 5)) - std::sin(7.0 * th) / (1680.0 * std::pow(mod, 7));
    for (int j = 1; j <= 10; j++)
      arg += std::atan(2.0 * t / (2 * j - 1));
    const double refl = -std::atan(std::tanh(kPi * t / 2.0));
    const double phase = arg + refl + shift * t;
    cd m0 = std::polar(std::sqrt(kPi / 2.0) / N, phase);
    if (i == 0)
      m0 *= 0.5;
    p.mult[i] = m0;
    p.mult[(size_t)N + i] = std::polar(1.0, 2.0 * (-refl - std::atan(2.0 * t))) * m0;
    for (int l = 1; l < lmax; l++)
      p.mult[(size_t)(l + 1) * N + i] =
          std::polar(1.0, -2.0 * std::atan(2.0 * t / (2 * l + 1))) * p.mult[(size_t)(l - 1) * N + i];
  }
}

static void sbt_apply(const SbtPlan &p, std::vector<cd> &x, int l) {
  dft_backward(x);
  const cd *M = p.mult.data() + (size_t)l * p.N;
  for (int i = 0; i < p.N; i++)
    x[i] = i < p.N / 2 ? x[i] * M[i] : cd(0, 0);
  dft_backward(x);
}

// g(k_p) for f = r R(r) on the n input points (p < n)
static void sbt_forward(const SbtPlan &p, const double *f, int l, double *g) {
  const int N = p.N, n = N / 2;
  std::vector<cd> x(N);
  const double C = f[0] / std::pow(p.rs[n], l + 1);  // r^{l+1} continuation below r[0]
  for (int i = 0; i < N; i++)
    x[i] = std::sqrt(p.rs[i]) * (i < n ? C * std::pow(p.rs[i], l + 1) : f[i - n]);
  sbt_apply(p, x, l);
  for (int q = 0; q < n; q++)
    g[q] = x[q].real() * 2.0 / std::pow(p.ks[q], 1.5);
}

// R(r_{n+q}) from g on the lower n points of the k grid
static void sbt_inverse(const SbtPlan &p, const double *g, int l, double *R) {
  const int N = p.N, n = N / 2;
  std::vector<cd> x(N, cd(0, 0));
  for (int i = 0; i < n; i++)
    x[i] = std::pow(p.ks[i], 1.5) * g[i];
  sbt_apply(p, x, l);
  for (int q = 0; q < n; q++)
    R[q] = x[q + n].real() * 4.0 / kPi / std::pow(p.rs[q + n], 1.5);
}

// Filtered partial-wave difference of projector.c:114-163 (see the
// declaration in pawpy_abi.h).
int smooth_diffwave_host(int n, const double *wave_grid, const double *diffwave, int l, int lmax, double grid_encut,
                         int np, double *sdw, double *kgrid, double *kwave) {
  SbtPlan p;
  sbt_plan(p, 1e7, 0.0, lmax, n, wave_grid);
  sbt_forward(p, diffwave, l, kwave);
  for (int q = 0; q < n; q++)
    kgrid[q] = p.ks[q];
  std::vector<double> kd(n, 0.0), sm(n), dg(n), b(n), c(n), d(n);
  const double kcut = std::sqrt(kCev * grid_encut);
  for (int q = 0; q < n && kgrid[q] < kcut; q++)
    kd[q] = kwave[q];
  sbt_inverse(p, kd.data(), l, sm.data());
  dg[0] = wave_grid[0];
  const double fac = wave_grid[1] / wave_grid[0];
  for (int q = 1; q < n; q++)
    dg[q] = dg[q - 1] * fac;
  spline_coeff_host(dg.data(), sm.data(), n, b.data(), c.data(), d.data());
  const double wave_rmax = wave_grid[n - 1];
  const double dlog = std::log(dg[1] / dg[0]);
  for (int q = 1; q < np; q++) {
    const double rr = wave_rmax / np * q;
    double v;
    if (rr > dg[n - 1]) {
      v = 0.0;
    } else if (rr < dg[0]) {
      v = sm[0];
    } else {
      int i = std::min((int)(std::log(rr / dg[0]) / dlog), n - 2);
      const double t = rr - dg[i];
      v = sm[i] + t * (b[i] + t * (c[i] + t * d[i]));
    }
    sdw[q] = v;
  }
  sdw[0] = l == 0 ? (np > 1 ? sdw[1] : 0.0) : 0.0;
  return PP_OK;
}

// ------------------------------------------------------------ off-site
static double factorial(int n) {
  double r = 1.0;
  for (int i = 2; i <= n; i++)
    r *= i;
  return r;
}

// Wigner 3j symbol for integer arguments (Racah's formula)
double wigner3j_host(int j1, int j2, int j3, int m1, int m2, int m3) {
  if (m1 + m2 + m3 != 0 || j3 < std::abs(j1 - j2) || j3 > j1 + j2 || std::abs(m1) > j1 || std::abs(m2) > j2 ||
      std::abs(m3) > j3)
    return 0.0;
  const double delta =
      factorial(j1 + j2 - j3) * factorial(j1 - j2 + j3) * factorial(j2 + j3 - j1) / factorial(j1 + j2 + j3 + 1);
  const double norm = std::sqrt(delta * factorial(j1 + m1) * factorial(j1 - m1) * factorial(j2 + m2) *
                                factorial(j2 - m2) * factorial(j3 + m3) * factorial(j3 - m3));
  const int lo = std::max(0, std::max(j2 - j3 - m1, j1 - j3 + m2));
  const int hi = std::min(j1 + j2 - j3, std::min(j1 - m1, j2 + m2));
  double s = 0.0;
  for (int t = lo; t <= hi; t++)
    s += (t & 1 ? -1.0 : 1.0) / (factorial(t) * factorial(j1 + j2 - j3 - t) * factorial(j1 - m1 - t) *
                                 factorial(j2 + m2 - t) * factorial(j3 - j2 + m1 + t) * factorial(j3 - j1 - m2 + t));
  return ((j1 - j2 - m3) & 1 ? -1.0 : 1.0) * norm * s;
}

// associated Legendre P_l^m(x) with the Condon-Shortley phase, m >= 0
static double plm(int l, int m, double x) {
  double pmm = 1.0;
  const double sx = std::sqrt((1.0 - x) * (1.0 + x));
  for (int i = 1; i <= m; i++)
    pmm *= -(2 * i - 1) * sx;
  if (l == m)
    return pmm;
  double p1 = x * (2 * m + 1) * pmm;
  for (int ll = m + 2; ll <= l; ll++) {
    const double p2 = (x * (2 * ll - 1) * p1 - (ll + m - 1) * pmm) / (ll - m);
    pmm = p1;
    p1 = p2;
  }
  return p1;
}

// Y_lm(theta, phi), Y_l,-m = (-1)^m conj(Y_lm)
static cd ylm_host(int l, int m, double theta, double phi) {
  const int am = std::abs(m);
  const double n = std::sqrt((2 * l + 1) / (4.0 * kPi) * factorial(l - am) / factorial(l + am));
  cd y = n * plm(l, am, std::cos(theta)) * std::polar(1.0, am * phi);
  if (m < 0)
    y = (am & 1 ? -1.0 : 1.0) * std::conj(y);
  return y;
}

// j_l(x) by upward recurrence (1 or 0 below 1e-5, as utils.c:807-827)
static double sph_jl(double x, int l) {
  if (x < 1e-5)
    return l == 0 ? 1.0 : 0.0;
  double a = std::sin(x) / x, b = std::sin(x) / (x * x) - std::cos(x) / x;
  if (l == 0)
    return a;
  for (int ll = 1; ll < l; ll++) {
    const double c = (2 * ll + 1) / x * b - a;
    a = b;
    b = c;
  }
  return b;
}

// cubic spline on a log grid (utils.c:477-489)
static double log_spline(double r, int n, const double *x, const double *f, const double *b, const double *c,
                         const double *d) {
  if (r > x[n - 1])
    return 0.0;
  if (r < x[0])
    return f[0];
  const int i = std::min((int)(std::log(r / x[0]) / std::log(x[1] / x[0])), n - 2);
  const double t = r - x[i];
  return f[i] + t * (b[i] + t * (c[i] + t * d[i]));
}

// Overlap int phi_1(r) conj(phi_2(r - dcoord)) d^3r of phi_i = R_i Y_{l_i m_i}
// from the transforms f_i(k) = int r^2 j_l R_i dr on log k grids, as
//   sum_L I_L 8 (-1)^m1 F(L) Y_L^{m1-m2}(dcoord) i^{l2+L-l1},
// I_L = int k^2 f1 f2 j_L(k R) dk (500-point log grid, spline integral),
// F(L) = (lx ly L; 0 0 0)(lx ly L; -mx my mx-my) sqrt((2lx+1)(2ly+1)(2L+1)/4pi)
// with (lx, mx) the larger-l function and my >= 0 (radial.c:116-196,
// gaunt.py:17-27).  R = 0 keeps only (2/pi) I_0 for equal (l, m).
cd offsite_overlap_host(const double *dcoord, const double *k1, const double *f1, const double *const *s1, int n1,
                        const double *k2, const double *f2, const double *const *s2, int n2, int l1, int m1, int l2,
                        int m2) {
  constexpr int NK = 500;
  int lx = l1, ly = l2, mx = m1, my = m2;
  if (l1 < l2) {
    std::swap(lx, ly);
    std::swap(mx, my);
  }
  if (my < 0) {
    mx = -mx;
    my = -my;
  }
  const double kmax = std::min(k1[n1 - 1], k2[n2 - 1]), kmin = std::max(k1[0], k2[0]);
  double R = std::sqrt(dcoord[0] * dcoord[0] + dcoord[1] * dcoord[1] + dcoord[2] * dcoord[2]);
  double theta = 0.0, phi = 0.0;
  if (R < 1e-11) {
    R = 0.0;
  } else {
    theta = std::acos(dcoord[2] / R);
    if (R - std::fabs(dcoord[2]) >= 1e-11)
      phi = std::acos(dcoord[0] / std::sqrt(dcoord[0] * dcoord[0] + dcoord[1] * dcoord[1]));
    if (dcoord[1] < 0)
      phi = 2.0 * kPi - phi;
  }
  std::vector<double> kg(NK), fn(NK), b(NK), c(NK), d(NK);
  cd total(0, 0);
  static const cd ipow[4] = {cd(1, 0), cd(0, 1), cd(-1, 0), cd(0, -1)};
  for (int L = std::abs(l1 - l2); L <= l1 + l2; L += 2) {
    for (int q = 0; q < NK; q++) {
      kg[q] = kmin * std::pow(kmax / kmin, (double)q / NK);
      fn[q] = log_spline(kg[q], n1, k1, f1, s1[0], s1[1], s1[2]) * log_spline(kg[q], n2, k2, f2, s2[0], s2[1], s2[2]) *
              kg[q] * kg[q] * sph_jl(kg[q] * R, L);
    }
    spline_coeff_host(kg.data(), fn.data(), NK, b.data(), c.data(), d.data());
    double I = 0.0;
    for (int q = 0; q < NK - 1; q++) {
      const double h = kg[q + 1] - kg[q];
      I += h * (fn[q] + h * (b[q] / 2 + h * (c[q] / 3 + d[q] * h / 4)));
    }
    if (R > 1e-9) {
      const double F = wigner3j_host(lx, ly, L, 0, 0, 0) * wigner3j_host(lx, ly, L, -mx, my, mx - my) *
                       std::sqrt((2 * lx + 1) * (2 * ly + 1) * (2 * L + 1) / 4.0 / kPi);
      total += I * F * (m1 & 1 ? -8.0 : 8.0) * ylm_host(L, m1 - m2, theta, phi) * ipow[((l2 + L - l1) % 4 + 4) % 4];
    } else if (L == 0 && l1 == l2 && m1 == m2) {
      total += I * 2.0 / kPi;
    }
  }
  return total;
}

}  // namespace pp

using namespace pp;

extern "C" {

int pp_sbt_forward(double encut, double enbuf, int lmax, int n, const double *r, const double *f, int l,
                   double *ks, double *g) {
  PP_REQUIRE(n >= 2 && r && f && ks && g && l >= 0 && r[0] > 0 && r[1] > r[0], PP_ERR_ARG,
             "pp_sbt_forward: bad arguments");
  SbtPlan p;
  sbt_plan(p, encut, enbuf, std::max(lmax, l), n, r);
  sbt_forward(p, f, l, g);
  for (int q = 0; q < n; q++)
    ks[q] = p.ks[q];
  return PP_OK;
}

int pp_smooth_diffwave(int n, const double *wave_grid, const double *diffwave, int l, int lmax, double grid_encut,
                       int np, double *sdw, double *kgrid, double *kwave) {
  PP_REQUIRE(n >= 2 && np >= 1 && wave_grid && diffwave && sdw && kgrid && kwave && l >= 0 && l <= std::max(lmax, 1) &&
                 wave_grid[0] > 0 && wave_grid[1] > wave_grid[0],
             PP_ERR_ARG, "pp_smooth_diffwave: bad arguments");
  return smooth_diffwave_host(n, wave_grid, diffwave, l, lmax, grid_encut, np, sdw, kgrid, kwave);
}

int pp_offsite_overlap(const double *dcoord, int n1, const double *r1, const double *f1, int n2, const double *r2,
                       const double *f2, int l1, int m1, int l2, int m2, double *out2) {
  PP_REQUIRE(dcoord && r1 && f1 && r2 && f2 && out2 && n1 >= 2 && n2 >= 2 && l1 >= 0 && l2 >= 0 && l1 <= 3 &&
                 l2 <= 3 && std::abs(m1) <= l1 && std::abs(m2) <= l2 && r1[0] > 0 && r1[1] > r1[0] && r2[0] > 0 &&
                 r2[1] > r2[0],
             PP_ERR_ARG, "pp_offsite_overlap: bad arguments");
  // pawpyc.pyx:151-190: both functions transformed at encut 1e5 eV
  std::vector<double> k1(n1), g1(n1), k2(n2), g2(n2);
  SbtPlan p1, p2;
  sbt_plan(p1, 1e5, 0.0, l1, n1, r1);
  sbt_forward(p1, f1, l1, g1.data());
  sbt_plan(p2, 1e5, 0.0, l2, n2, r2);
  sbt_forward(p2, f2, l2, g2.data());
  for (int q = 0; q < n1; q++)
    k1[q] = p1.ks[q];
  for (int q = 0; q < n2; q++)
    k2[q] = p2.ks[q];
  std::vector<double> b1(n1), c1(n1), d1(n1), b2(n2), c2(n2), d2(n2);
  spline_coeff_host(k1.data(), g1.data(), n1, b1.data(), c1.data(), d1.data());
  spline_coeff_host(k2.data(), g2.data(), n2, b2.data(), c2.data(), d2.data());
  const double *s1[3] = {b1.data(), c1.data(), d1.data()};
  const double *s2[3] = {b2.data(), c2.data(), d2.data()};
  const cd v = offsite_overlap_host(dcoord, k1.data(), g1.data(), s1, n1, k2.data(), g2.data(), s2, n2, l1, m1, l2, m2);
  out2[0] = v.real();
  out2[1] = v.imag();
  return PP_OK;
}

}  // extern "C"
