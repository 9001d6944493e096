// All-electron real-space states on the GPU (density.c:158-327).
//
// realspace_state (density.c:232-310) forms, for one band,
//   psi(r) = fft3d(C)(r) e^{2 pi i k.r}
//          + sum_s sum_n (phi_n - phi~_n)(r - R_s) <p~_n|psi~> e^{2 pi i k.(R_s + T)}
// with the partial-wave differences (utils.c:522-545: diffwave(|d|)/|d|
// times the complex Y_lm of d, spline on the log wave grid) taken at
// the unwrapped box vector d = r - R_s of every grid point with
// |d| < rmax_s (the last point of the species' wave grid), T the lattice
// vector that wraps the point into the cell.
//
// Plan (pp_rs_plan_create): per site the box points inside its sphere
// (count, host scan, ordered fill; the order of the reference's i-j-k
// loops) and a complex table of the partial-wave values per point and
// projector, independent of k and band.  Sites whose spheres can share a
// grid point get different colours, so one launch per colour adds the
// augmentation with plain read-modify-writes (deterministic); when a
// site's box wraps onto itself the plan uses atomic adds instead.
//
// Per band chunk (pp_realspace_batched; two chunks in flight on two
// streams): scatter + backward FFT into the band-fastest chunk (the
// projection path's kernels), augmentation added
// in the chunk with the phase e^{-2 pi i k.(r_unwrapped - R)} / scale so
// that the emit kernel's common factor scale e^{2 pi i k.r} restores the
// reference's terms, then either the band-major states (LDS transpose,
// with or without the Bloch phase, i.e. remove_phase) or the weighted
// density sum_b w_b |psi_b|^2 (ae_state_density / ae_chg_density,
// density.c:158-179).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "pp_geom.h"
#include "pp_internal.h"

namespace pp {

struct RsSite {
  double R[3];
  double rmax;
  int c[3], g[3];
  int nproj, nchan, n;  // projectors, channels, wave-grid points
  int chan_off;         // first channel of the species in chan_l
  int proj_off;         // first projection of the site
  long spl_off;         // species block: x[n], then f, b, c, d (n each) per channel
};

struct RsLat {
  double m[9];
};

struct RsWeights {
  double w[64];
};

// candidate c of a site's box -> unwrapped (i, j, k), k fastest (the
// reference's loop order, density.c:272-276)
__device__ __forceinline__ void rs_cand(const RsSite &s, int c, int &i, int &j, int &k) {
  const int w1 = 2 * s.g[1] + 1, w2 = 2 * s.g[2] + 1;
  k = c % w2 - s.g[2] + s.c[2];
  j = (c / w2) % w1 - s.g[1] + s.c[1];
  i = c / (w1 * w2) - s.g[0] + s.c[0];
}

// unwrapped box vector r - R in cartesian coordinates (density.c:277-281)
__device__ __forceinline__ double rs_vec(const RsSite &s, const RsLat &L, int f0, int f1, int f2, int i, int j, int k,
                                         double *v) {
  double f[3] = {(double)i / f0 - s.R[0], (double)j / f1 - s.R[1], (double)k / f2 - s.R[2]};
  d_frac_to_cart(f, L.m, v);
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

__global__ __launch_bounds__(256) void k_rs_count(const RsSite *__restrict__ geo, RsLat L, int f0, int f1, int f2,
                                                  int nblk, int *__restrict__ blkcnt) {
  const int s = blockIdx.y;
  const RsSite g = geo[s];
  const int ncand = (2 * g.g[0] + 1) * (2 * g.g[1] + 1) * (2 * g.g[2] + 1);
  const int c = blockIdx.x * 256 + threadIdx.x;
  bool in = false;
  if (c < ncand) {
    int i, j, k;
    double v[3];
    rs_cand(g, c, i, j, k);
    in = rs_vec(g, L, f0, f1, f2, i, j, k, v) < g.rmax;
  }
  const int cnt = __syncthreads_count(in);
  if (threadIdx.x == 0)
    blkcnt[(long)s * nblk + blockIdx.x] = cnt;
}

// points in candidate order: (wrapped grid index, unwrapped i, j, k)
__global__ __launch_bounds__(256) void k_rs_fill(const RsSite *__restrict__ geo, RsLat L, int f0, int f1, int f2,
                                                 int nblk, const long *__restrict__ blkoff, int4 *__restrict__ pts) {
  __shared__ int wsum[4];
  const int s = blockIdx.y;
  const RsSite g = geo[s];
  const int ncand = (2 * g.g[0] + 1) * (2 * g.g[1] + 1) * (2 * g.g[2] + 1);
  const int c = blockIdx.x * 256 + threadIdx.x;
  bool in = false;
  int i = 0, j = 0, k = 0;
  if (c < ncand) {
    double v[3];
    rs_cand(g, c, i, j, k);
    in = rs_vec(g, L, f0, f1, f2, i, j, k, v) < g.rmax;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long m = __ballot(in);
  const int pre = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0)
    wsum[w] = __popcll(m);
  __syncthreads();
  int base = 0;
  for (int q = 0; q < w; q++)
    base += wsum[q];
  if (in) {
    const int ii = (i % f0 + f0) % f0, jj = (j % f1 + f1) % f1, kk = (k % f2 + f2) % f2;
    pts[blkoff[(long)s * nblk + blockIdx.x] + base + pre] = make_int4((ii * f1 + jj) * f2 + kk, i, j, k);
  }
}

// tab[tab_off[s] + q * P + p] = wave_value2 of projector p at point q
// (utils.c:477-489, 522-545; Y_lm of utils.c:441-449)
__global__ __launch_bounds__(256) void k_rs_values(const RsSite *__restrict__ geo, RsLat L, int f0, int f1, int f2,
                                                   const int *__restrict__ site_pt, const int4 *__restrict__ pts,
                                                   const long *__restrict__ tab_off, const double *__restrict__ spl,
                                                   const int *__restrict__ chan_l, cplx<double> *__restrict__ tab) {
  const int s = blockIdx.y;
  const RsSite g = geo[s];
  const int p0 = site_pt[s], np = site_pt[s + 1] - p0;
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= np)
    return;
  const int4 pt = pts[p0 + q];
  double v[3];
  const double r = rs_vec(g, L, f0, f1, f2, pt.y, pt.z, pt.w, v);
  double theta = 0.0, phi = 0.0;
  if (r != 0.0) {
    theta = acos(v[2] / r);
    if (r - fabs(v[2]) != 0.0)
      phi = acos(v[0] / sqrt(v[0] * v[0] + v[1] * v[1]));
    if (v[1] < 0)
      phi = 2.0 * M_PI - phi;
  }
  const double ct = cos(theta);
  const int n = g.n;
  const double *x = spl + g.spl_off;
  int id = 0;
  double t = 0.0;
  const bool below = r < x[0], beyond = r > x[n - 1];
  if (!below && !beyond) {
    id = (int)(log(r / x[0]) / log(x[1] / x[0]));
    if (id > n - 2)
      id = n - 2;
    t = r - x[id];
  }
  const double rdiv = below ? x[0] : r;
  cplx<double> *out = tab + tab_off[s] + (long)q * g.nproj;
  int p = 0;
  for (int ch = 0; ch < g.nchan; ch++) {
    const int l = chan_l[g.chan_off + ch];
    const double *fc = x + (long)n * (1 + 4 * ch);
    const double *bc = fc + n, *cc = bc + n, *dc = cc + n;
    double radial = beyond ? 0.0 : below ? fc[0] : fc[id] + t * (bc[id] + t * (cc[id] + t * dc[id]));
    radial /= rdiv;
    for (int m = -l; m <= l; m++, p++) {
      const double norm = sqrt((2 * l + 1) / (4.0 * M_PI) * d_fact(l - m) / d_fact(l + m));
      const double a = norm * d_legendre(l, m, ct) * radial;
      double sn, cs;
      sincos(m * phi, &sn, &cs);
      out[p] = mk<double>(a * cs, a * sn);
    }
  }
}

// Adds the augmentation of the sites sites[blockIdx.y] to the band-fastest
// chunk: lane = band, the site's coefficients in LDS; each wavefront takes
// U points of the block's range at a time, their grid values loaded
// together before the sums so that the read-modify-writes overlap.
// proj rows: proj[b * ldp + proj_off + p] for the chunk's bands.
template <typename T, bool ATOMIC, int U>
__global__ __launch_bounds__(256) void k_rs_augment(cplx<T> *__restrict__ grid, int NB, int nbc,
                                                    const RsSite *__restrict__ geo, const int *__restrict__ sites,
                                                    const int *__restrict__ site_pt, const int4 *__restrict__ pts,
                                                    const long *__restrict__ tab_off,
                                                    const cplx<double> *__restrict__ tab,
                                                    const cplx<double> *__restrict__ proj, long ldp, double k0,
                                                    double k1, double k2, double iscl, int f0, int f1, int f2,
                                                    int ppb) {
  extern __shared__ cplx<double> coef[];  // [P][64]
  const int s = sites[blockIdx.y];
  const RsSite g = geo[s];
  const int P = g.nproj;
  const int p0 = site_pt[s], np = site_pt[s + 1] - p0;
  const int begin = blockIdx.x * ppb;
  if (begin >= np)
    return;  // uniform over the block
  const int end = min(np, begin + ppb);
  for (int t = threadIdx.x; t < P * 64; t += 256) {
    const int p = t % P, b = t / P;
    coef[p * 64 + b] = b < nbc ? proj[(long)b * ldp + g.proj_off + p] : mk<double>(0.0, 0.0);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool act = lane < nbc;
  const cplx<double> *tsite = tab + tab_off[s];
  for (int q0 = begin + wv * U; q0 < end; q0 += 4 * U) {
    int4 pt[U];
    cplx<T> v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      pt[u] = q0 + u < end ? pts[p0 + q0 + u] : make_int4(-1, 0, 0, 0);
      v[u] = mk<T>((T)0, (T)0);
    }
    if (!ATOMIC) {
#pragma unroll
      for (int u = 0; u < U; u++)
        if (pt[u].x >= 0 && act)
          v[u] = grid[(long)pt[u].x * NB + lane];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (pt[u].x < 0)
        continue;
      const cplx<double> *row = tsite + (long)(q0 + u) * P;
      double ar = 0.0, ai = 0.0;
      for (int p = 0; p < P; p++) {
        const cplx<double> w = row[p], c = coef[p * 64 + lane];
        ar += w.x * c.x - w.y * c.y;
        ai += w.x * c.y + w.y * c.x;
      }
      // e^{2 pi i k.(R + T)} e^{-2 pi i k.r} = e^{-2 pi i k.(r_unwrapped - R)}
      const double kt = k0 * ((double)pt[u].y / f0 - g.R[0]) + k1 * ((double)pt[u].z / f1 - g.R[1]) +
                        k2 * ((double)pt[u].w / f2 - g.R[2]);
      double sn, cs;
      sincospi(-2.0 * kt, &sn, &cs);
      const double vr = (ar * cs - ai * sn) * iscl, vi = (ar * sn + ai * cs) * iscl;
      if (act) {
        cplx<T> *dst = grid + (long)pt[u].x * NB + lane;
        if (ATOMIC) {
          atomicAdd(&dst->x, (T)vr);
          atomicAdd(&dst->y, (T)vi);
        } else {
          v[u].x += (T)vr;
          v[u].y += (T)vi;
          *dst = v[u];
        }
      }
    }
  }
}

// The same with the site's coefficients of the lane's band in registers
// (P <= RP) and U points of a wavefront in flight: their grid values are
// loaded together before the sums, so the read-modify-writes of one
// wavefront overlap.
template <typename T, bool ATOMIC, int RP, int U>
__global__ __launch_bounds__(256) void k_rs_augment_r(cplx<T> *__restrict__ grid, int NB, int nbc,
                                                      const RsSite *__restrict__ geo, const int *__restrict__ sites,
                                                      const int *__restrict__ site_pt, const int4 *__restrict__ pts,
                                                      const long *__restrict__ tab_off,
                                                      const cplx<double> *__restrict__ tab,
                                                      const cplx<double> *__restrict__ proj, long ldp, double k0,
                                                      double k1, double k2, double iscl, int f0, int f1, int f2,
                                                      int ppb) {
  const int s = sites[blockIdx.y];
  const RsSite g = geo[s];
  const int P = g.nproj;
  const int p0 = site_pt[s], np = site_pt[s + 1] - p0;
  const int begin = blockIdx.x * ppb;
  if (begin >= np)
    return;
  const int end = min(np, begin + ppb);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool act = lane < nbc;
  cplx<double> cf[RP];
#pragma unroll
  for (int p = 0; p < RP; p++)
    cf[p] = (p < P && act) ? proj[(long)lane * ldp + g.proj_off + p] : mk<double>(0.0, 0.0);
  const cplx<double> *tsite = tab + tab_off[s];
  for (int q0 = begin + wv * U; q0 < end; q0 += 4 * U) {
    int4 pt[U];
    cplx<T> v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      pt[u] = q0 + u < end ? pts[p0 + q0 + u] : make_int4(-1, 0, 0, 0);
      v[u] = mk<T>((T)0, (T)0);
    }
    if (!ATOMIC) {
#pragma unroll
      for (int u = 0; u < U; u++)
        if (pt[u].x >= 0 && act)
          v[u] = grid[(long)pt[u].x * NB + lane];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (pt[u].x < 0)
        continue;
      const cplx<double> *row = tsite + (long)(q0 + u) * P;
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int p = 0; p < RP; p++) {
        if (p < P) {
          const cplx<double> w = row[p];
          ar += w.x * cf[p].x - w.y * cf[p].y;
          ai += w.x * cf[p].y + w.y * cf[p].x;
        }
      }
      const double kt = k0 * ((double)pt[u].y / f0 - g.R[0]) + k1 * ((double)pt[u].z / f1 - g.R[1]) +
                        k2 * ((double)pt[u].w / f2 - g.R[2]);
      double sn, cs;
      sincospi(-2.0 * kt, &sn, &cs);
      const double vr = (ar * cs - ai * sn) * iscl, vi = (ar * sn + ai * cs) * iscl;
      if (act) {
        cplx<T> *dst = grid + (long)pt[u].x * NB + lane;
        if (ATOMIC) {
          atomicAdd(&dst->x, (T)vr);
          atomicAdd(&dst->y, (T)vi);
        } else {
          v[u].x += (T)vr;
          v[u].y += (T)vi;
          *dst = v[u];
        }
      }
    }
  }
}

// Tiled form: a work-group stages the table rows, grid indices and phase
// factors of its ppb points in LDS (one contiguous read of the rows), the
// lane's band coefficients sit in registers (P <= RP), and each wavefront
// sums U points at a time from broadcast LDS reads of the rows.
template <typename T, int RP, int U>
__global__ __launch_bounds__(256) void k_rs_augment_t(cplx<T> *__restrict__ grid, int NB, int nbc,
                                                      const RsSite *__restrict__ geo, const int *__restrict__ sites,
                                                      const int *__restrict__ site_pt, const int4 *__restrict__ pts,
                                                      const long *__restrict__ tab_off,
                                                      const cplx<double> *__restrict__ tab,
                                                      const cplx<double> *__restrict__ proj, long ldp, double k0,
                                                      double k1, double k2, double iscl, int f0, int f1, int f2,
                                                      int ppb) {
  extern __shared__ cplx<double> smem[];
  const int s = sites[blockIdx.y];
  const RsSite g = geo[s];
  const int P = g.nproj;
  const int p0 = site_pt[s], np = site_pt[s + 1] - p0;
  const int begin = blockIdx.x * ppb;
  if (begin >= np)
    return;  // uniform over the block
  const int cnt = min(ppb, np - begin);
  cplx<double> *Ts = smem;       // [cnt][P]
  cplx<double> *Ph = Ts + (long)ppb * P;  // [cnt]
  int *Ix = (int *)(Ph + ppb);   // [cnt]
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool act = lane < nbc;
  const cplx<double> *rows = tab + tab_off[s] + (long)begin * P;
  // the table tile goes from global memory straight to LDS
  // (global_load_lds_dwordx4: lane l of a wavefront writes its 16-byte
  // entry at the wave-uniform LDS address + 16 l, masked-off lanes write
  // nothing), so all of a work-group's tile loads are in flight at once and
  // hold no registers; a copy through registers waited on each load in turn
  // (the kernel's registers leave no room to batch them)
  for (int i = wv; i * 64 < cnt * P; i += 4) {
    const int t = i * 64 + lane;
    if (t < cnt * P)
      __builtin_amdgcn_global_load_lds((const void *)(rows + t),
                                       (__attribute__((address_space(3))) void *)(Ts + i * 64), 16, 0, 0);
  }
  // the site's projections of this lane's band, loaded with the tile
  cplx<double> cf[RP];
#pragma unroll
  for (int p = 0; p < RP; p++)
    cf[p] = (p < P && act) ? proj[(long)lane * ldp + g.proj_off + p] : mk<double>(0.0, 0.0);
  for (int t = threadIdx.x; t < cnt; t += 256) {
    const int4 pt = pts[p0 + begin + t];
    Ix[t] = pt.x;
    // e^{2 pi i k.(R + T)} e^{-2 pi i k.r} = e^{-2 pi i k.(r_unwrapped - R)}
    const double kt = k0 * ((double)pt.y / f0 - g.R[0]) + k1 * ((double)pt.z / f1 - g.R[1]) +
                      k2 * ((double)pt.w / f2 - g.R[2]);
    double sn, cs;
    sincospi(-2.0 * kt, &sn, &cs);
    Ph[t] = mk<double>(cs * iscl, sn * iscl);
  }
  // the direct-to-LDS loads count as outstanding vector loads: wait for all
  // of them before the barrier that publishes the tile
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int q0 = wv * U; q0 < cnt; q0 += 4 * U) {
    cplx<T> v[U];
#pragma unroll
    for (int u = 0; u < U; u++)
      v[u] = (q0 + u < cnt && act) ? grid[(long)Ix[q0 + u] * NB + lane] : mk<T>((T)0, (T)0);
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (q0 + u >= cnt)
        break;
      const cplx<double> *row = Ts + (q0 + u) * P;
      double ar = 0.0, ai = 0.0;
#pragma unroll
      for (int p = 0; p < RP; p++) {
        if (p < P) {
          const cplx<double> w = row[p];
          ar += w.x * cf[p].x - w.y * cf[p].y;
          ai += w.x * cf[p].y + w.y * cf[p].x;
        }
      }
      const cplx<double> e = Ph[q0 + u];
      if (act) {
        v[u].x += (T)(ar * e.x - ai * e.y);
        v[u].y += (T)(ar * e.y + ai * e.x);
        grid[(long)Ix[q0 + u] * NB + lane] = v[u];
      }
    }
  }
}

// out[b * ldo + pt] = scl grid[pt][b] (e^{2 pi i k.r} unless remove_phase):
// a tile of RS_TP points x NB bands through LDS.
constexpr int RS_TP = 32;
template <typename T>
__global__ __launch_bounds__(256) void k_rs_state(const cplx<T> *__restrict__ grid, int NB, int nbc, long npts,
                                                  double scl, int f0, int f1, int f2, double k0, double k1,
                                                  double k2, int phase_on, cplx<double> *__restrict__ out,
                                                  long ldo) {
  __shared__ cplx<double> tile[RS_TP * 65];
  __shared__ cplx<double> ph[RS_TP];
  const long pt0 = (long)blockIdx.x * RS_TP;
  const int npt = (int)min((long)RS_TP, npts - pt0);
  if (threadIdx.x < RS_TP) {
    cplx<double> e = mk<double>(scl, 0.0);
    const long pt = pt0 + threadIdx.x;
    if (phase_on && (int)threadIdx.x < npt) {
      const int kk = (int)(pt % f2), jj = (int)((pt / f2) % f1), ii = (int)(pt / ((long)f1 * f2));
      const double a = k0 * ii / f0 + k1 * jj / f1 + k2 * kk / f2;
      double sn, cs;
      sincospi(2.0 * a, &sn, &cs);
      e = mk<double>(scl * cs, scl * sn);
    }
    ph[threadIdx.x] = e;
  }
  const cplx<T> *src = grid + pt0 * NB;
  // all of a thread's loads in flight before the LDS writes
  cplx<T> v[RS_TP * 64 / 256];
#pragma unroll
  for (int u = 0; u < RS_TP * 64 / 256; u++) {
    const int t = threadIdx.x + u * 256;
    v[u] = t < npt * NB ? src[t] : mk<T>((T)0, (T)0);
  }
#pragma unroll
  for (int u = 0; u < RS_TP * 64 / 256; u++) {
    const int t = threadIdx.x + u * 256;
    if (t < npt * NB) {
      const int p = t / NB, b = t - p * NB;
      tile[p * 65 + b] = cvt<double>(v[u]);
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < RS_TP * nbc; t += 256) {
    const int b = t / RS_TP, p = t - b * RS_TP;
    if (p < npt)
      out[(long)b * ldo + pt0 + p] = tile[p * 65 + b] * ph[p];
  }
}

// dens[pt] += scl^2 sum_b w_b |grid[pt][b]|^2: lane = band, one point per
// wavefront step.
template <typename T>
__global__ __launch_bounds__(256) void k_rs_density(const cplx<T> *__restrict__ grid, int NB, int nbc, long npts,
                                                    double scl2, RsWeights W, int ppw,
                                                    double *__restrict__ dens) {
  const int lane = threadIdx.x & 63;
  const long w0 = ((long)blockIdx.x * 4 + (threadIdx.x >> 6)) * ppw;
  const double wl = lane < nbc ? W.w[lane] : 0.0;
  constexpr int U = 8;  // points whose loads are in flight together and whose sums share one reduction
  // The U band sums of a batch are reduced over the 64 lanes together: the
  // exchanges across lane bits 5, 4 and 3 each halve the values a lane holds
  // (it keeps the points of its half and adds the partner's partial sums of
  // them), those across bits 2, 1, 0 finish one sum; 10 exchanges for 8
  // points instead of 6 per point.  Lane l then holds the sum of point
  // q(l) = 4 b5 + 2 b4 + b3 of the batch, and lane 8 q(l) of each group of 8
  // updates that point's density.
  const int qmine = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
  const bool h5 = (lane >> 5) & 1, h4 = (lane >> 4) & 1, h3 = (lane >> 3) & 1;
  for (int q0 = 0; q0 < ppw; q0 += U) {
    cplx<T> a[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const long pt = w0 + q0 + u;
      a[u] = (q0 + u < ppw && pt < npts && lane < nbc) ? grid[pt * NB + lane] : mk<T>((T)0, (T)0);
    }
    // the density value this lane updates, read with the batch
    const long ptm = w0 + q0 + qmine;
    const bool owner = (lane & 7) == 0 && q0 + qmine < ppw && ptm < npts;
    const double dold = owner ? dens[ptm] : 0.0;
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const cplx<double> c = cvt<double>(a[u]);
      v[u] = wl * (c.x * c.x + c.y * c.y);
    }
    double v4[4], v2[2];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double send = h5 ? v[j] : v[j + 4], keep = h5 ? v[j + 4] : v[j];
      v4[j] = keep + __shfl_xor(send, 32, 64);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const double send = h4 ? v4[j] : v4[j + 2], keep = h4 ? v4[j + 2] : v4[j];
      v2[j] = keep + __shfl_xor(send, 16, 64);
    }
    double v1 = (h3 ? v2[1] : v2[0]) + __shfl_xor(h3 ? v2[0] : v2[1], 8, 64);
    v1 += __shfl_xor(v1, 4, 64);
    v1 += __shfl_xor(v1, 2, 64);
    v1 += __shfl_xor(v1, 1, 64);
    if (owner)
      dens[ptm] = dold + scl2 * v1;
  }
}

__global__ __launch_bounds__(256) void k_rs_add(double *__restrict__ d, const double *__restrict__ e, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    d[i] += e[i];
}

}  // namespace pp

using namespace pp;

struct pp_rsplan {
  pp_ctx *ctx = nullptr;
  double lattice[9];
  int fftg[3];
  int nsites = 0, ptot = 0, pmax = 0;
  long npts = 0;
  bool atomic = false;
  std::vector<int> proj_off, site_pt;
  std::vector<int> color_off, color_sites, color_maxpts;
  DevBuf d_geo, d_pts, d_tab, d_tab_off, d_site_pt, d_colors, d_spl, d_chan_l;
  DevBuf dens2;  // densities of the second stream's chunks
};

namespace {

double cross_norm(const double *a, const double *b) {
  const double c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2], c2 = a[0] * b[1] - a[1] * b[0];
  return sqrt(c0 * c0 + c1 * c1 + c2 * c2);
}

// Greedy colouring: sites s, t conflict when some lattice image brings
// them within rmax_s + rmax_t (their spheres may share a grid point).
void colour_sites(const double *lat, const std::vector<RsSite> &geo, std::vector<int> &colour) {
  const int ns = (int)geo.size();
  colour.assign(ns, 0);
  std::vector<char> used;
  for (int s = 0; s < ns; s++) {
    used.assign(ns + 1, 0);
    for (int t = 0; t < s; t++) {
      const double lim = geo[s].rmax + geo[t].rmax;
      bool hit = false;
      for (int a = -2; a <= 2 && !hit; a++)
        for (int b = -2; b <= 2 && !hit; b++)
          for (int c = -2; c <= 2 && !hit; c++) {
            const double f[3] = {geo[s].R[0] - geo[t].R[0] + a, geo[s].R[1] - geo[t].R[1] + b,
                                 geo[s].R[2] - geo[t].R[2] + c};
            double v[3];
            for (int q = 0; q < 3; q++)
              v[q] = f[0] * lat[q] + f[1] * lat[3 + q] + f[2] * lat[6 + q];
            hit = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) < lim;
          }
      if (hit)
        used[colour[t]] = 1;
    }
    int c = 0;
    while (used[c])
      c++;
    colour[s] = c;
  }
}

}  // namespace

extern "C" int pp_rs_plan_create(pp_ctx *ctx, pp_rsplan **out, const double *lattice, const int *fftg, int num_sites,
                                 const int *labels, const double *frac_coords, int num_elems, const int *nchan,
                                 const int *ls, const int *wave_n, const double *wave_grid,
                                 const double *diffwave) {
  PP_REQUIRE(ctx && out && lattice && fftg && num_sites >= 0 && num_elems > 0 && nchan && wave_n, PP_ERR_ARG,
             "pp_rs_plan_create: bad arguments");
  PP_REQUIRE(num_sites == 0 || (labels && frac_coords), PP_ERR_ARG, "pp_rs_plan_create: null site arrays");
  PP_REQUIRE(num_sites <= 65535, PP_ERR_ARG, "pp_rs_plan_create: more than 65535 sites (one grid row per site)");
  PP_REQUIRE(fftg[0] > 0 && fftg[1] > 0 && fftg[2] > 0 && (long)fftg[0] * fftg[1] * fftg[2] < (1L << 31),
             PP_ERR_ARG, "pp_rs_plan_create: bad grid");
  const double vol = det3(lattice);
  PP_REQUIRE(vol > 0, PP_ERR_ARG, "pp_rs_plan_create: the lattice must be right-handed");
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  // species blocks: x[n], then f, b, c, d per channel (host splines)
  std::vector<long> spl_off(num_elems);
  std::vector<int> chan_off(num_elems + 1, 0), nproj(num_elems, 0);
  std::vector<double> spl;
  long goff = 0, doff = 0;
  for (int e = 0; e < num_elems; e++) {
    const int n = wave_n[e];
    PP_REQUIRE(n >= 2 && nchan[e] >= 0, PP_ERR_ARG, "pp_rs_plan_create: wave grid needs 2 or more points");
    PP_REQUIRE(wave_grid[goff] > 0 && wave_grid[goff + 1] > wave_grid[goff], PP_ERR_ARG,
               "pp_rs_plan_create: the wave grid must be positive and increasing");
    chan_off[e + 1] = chan_off[e] + nchan[e];
    spl_off[e] = (long)spl.size();
    spl.insert(spl.end(), wave_grid + goff, wave_grid + goff + n);
    for (int ch = 0; ch < nchan[e]; ch++) {
      const int l = ls[chan_off[e] + ch];
      PP_REQUIRE(l >= 0 && l <= 3, PP_ERR_ARG, "pp_rs_plan_create: l must be 0..3");
      nproj[e] += 2 * l + 1;
      const double *f = diffwave + doff + (long)ch * n;
      std::vector<double> b(n), c(n), d(n);
      spline_coeff_host(wave_grid + goff, f, n, b.data(), c.data(), d.data());
      spl.insert(spl.end(), f, f + n);
      spl.insert(spl.end(), b.begin(), b.end());
      spl.insert(spl.end(), c.begin(), c.end());
      spl.insert(spl.end(), d.begin(), d.end());
    }
    goff += n;
    doff += (long)nchan[e] * n;
  }
  std::unique_ptr<pp_rsplan> hold(new pp_rsplan());
  pp_rsplan *rp = hold.get();
  rp->ctx = ctx;
  memcpy(rp->lattice, lattice, sizeof(rp->lattice));
  memcpy(rp->fftg, fftg, sizeof(rp->fftg));
  rp->nsites = num_sites;
  std::vector<RsSite> geo(num_sites);
  rp->proj_off.assign(num_sites + 1, 0);
  int nblk = 1;
  for (int s = 0; s < num_sites; s++) {
    const int e = labels[s];
    PP_REQUIRE(e >= 0 && e < num_elems, PP_ERR_ARG, "pp_rs_plan_create: label out of range");
    RsSite &g = geo[s];
    const double *x = wave_grid;
    long o = 0;
    for (int q = 0; q < e; q++)
      o += wave_n[q];
    g.rmax = x[o + wave_n[e] - 1];
    for (int q = 0; q < 3; q++) {
      g.R[q] = frac_coords[3 * s + q];
      g.c[q] = (int)std::round(g.R[q] * fftg[q]);
    }
    // density.c:266-274
    g.g[0] = (int)(cross_norm(lattice + 3, lattice + 6) * g.rmax / vol * fftg[0]) + 1;
    g.g[1] = (int)(cross_norm(lattice + 0, lattice + 6) * g.rmax / vol * fftg[1]) + 1;
    g.g[2] = (int)(cross_norm(lattice + 0, lattice + 3) * g.rmax / vol * fftg[2]) + 1;
    for (int q = 0; q < 3; q++)
      if (2 * g.g[q] + 1 > fftg[q])
        rp->atomic = true;  // the box wraps onto itself
    g.nproj = nproj[e];
    g.nchan = nchan[e];
    g.n = wave_n[e];
    g.chan_off = chan_off[e];
    g.spl_off = spl_off[e];
    g.proj_off = rp->proj_off[s];
    rp->proj_off[s + 1] = rp->proj_off[s] + nproj[e];
    rp->pmax = std::max(rp->pmax, nproj[e]);
    const long ncand = (long)(2 * g.g[0] + 1) * (2 * g.g[1] + 1) * (2 * g.g[2] + 1);
    PP_REQUIRE(ncand < (1L << 31), PP_ERR_ARG, "pp_rs_plan_create: site box too large");
    nblk = std::max(nblk, (int)((ncand + 255) / 256));
  }
  rp->ptot = rp->proj_off[num_sites];
  PP_REQUIRE(rp->pmax <= 64, PP_ERR_ARG, "pp_rs_plan_create: more than 64 projectors per site");
  int st = PP_OK;
  const int ns = std::max(1, num_sites);
  if ((st = rp->d_geo.ensure(sizeof(RsSite) * ns)) || (st = rp->d_spl.ensure(sizeof(double) * std::max<size_t>(1, spl.size()))) ||
      (st = rp->d_chan_l.ensure(sizeof(int) * std::max(1, chan_off[num_elems]))))
    return st;
  hipStream_t sm = ctx->stream;
  if (num_sites)
    PP_CHECK_HIP(hipMemcpyAsync(rp->d_geo.p, geo.data(), sizeof(RsSite) * num_sites, hipMemcpyHostToDevice, sm));
  if (!spl.empty())
    PP_CHECK_HIP(hipMemcpyAsync(rp->d_spl.p, spl.data(), sizeof(double) * spl.size(), hipMemcpyHostToDevice, sm));
  if (chan_off[num_elems])
    PP_CHECK_HIP(hipMemcpyAsync(rp->d_chan_l.p, ls, sizeof(int) * chan_off[num_elems], hipMemcpyHostToDevice, sm));
  RsLat L;
  memcpy(L.m, lattice, sizeof(L.m));
  rp->site_pt.assign(num_sites + 1, 0);
  if (num_sites) {
    DevBuf cnt, boff;
    if ((st = cnt.ensure(sizeof(int) * (size_t)num_sites * nblk)) ||
        (st = boff.ensure(sizeof(long) * (size_t)num_sites * nblk)))
      return st;
    hipLaunchKernelGGL(k_rs_count, dim3(nblk, num_sites), dim3(256), 0, sm, (const RsSite *)rp->d_geo.p, L, fftg[0],
                       fftg[1], fftg[2], nblk, (int *)cnt.p);
    PP_CHECK_HIP(hipGetLastError());
    std::vector<int> hc((size_t)num_sites * nblk);
    PP_CHECK_HIP(hipMemcpyAsync(hc.data(), cnt.p, sizeof(int) * hc.size(), hipMemcpyDeviceToHost, sm));
    PP_CHECK_HIP(hipStreamSynchronize(sm));
    std::vector<long> ho(hc.size());
    long acc = 0;
    for (int s = 0; s < num_sites; s++) {
      for (int b = 0; b < nblk; b++) {
        ho[(size_t)s * nblk + b] = acc;
        acc += hc[(size_t)s * nblk + b];
      }
      PP_REQUIRE(acc < (1L << 31), PP_ERR_ARG, "pp_rs_plan_create: more than 2^31 sphere points");
      rp->site_pt[s + 1] = (int)acc;
    }
    rp->npts = acc;
    if ((st = rp->d_pts.ensure(sizeof(int4) * std::max(1L, acc))))
      return st;
    PP_CHECK_HIP(hipMemcpyAsync(boff.p, ho.data(), sizeof(long) * ho.size(), hipMemcpyHostToDevice, sm));
    hipLaunchKernelGGL(k_rs_fill, dim3(nblk, num_sites), dim3(256), 0, sm, (const RsSite *)rp->d_geo.p, L, fftg[0],
                       fftg[1], fftg[2], nblk, (const long *)boff.p, (int4 *)rp->d_pts.p);
    PP_CHECK_HIP(hipGetLastError());
    // value table
    std::vector<long> toff(num_sites + 1, 0);
    int maxp = 1;
    for (int s = 0; s < num_sites; s++) {
      const int np = rp->site_pt[s + 1] - rp->site_pt[s];
      toff[s + 1] = toff[s] + (long)np * geo[s].nproj;
      maxp = std::max(maxp, np);
    }
    if ((st = rp->d_tab.ensure(sizeof(cplx<double>) * std::max(1L, toff[num_sites]))) ||
        (st = rp->d_tab_off.ensure(sizeof(long) * (num_sites + 1))) ||
        (st = rp->d_site_pt.ensure(sizeof(int) * (num_sites + 1))))
      return st;
    PP_CHECK_HIP(
        hipMemcpyAsync(rp->d_tab_off.p, toff.data(), sizeof(long) * toff.size(), hipMemcpyHostToDevice, sm));
    PP_CHECK_HIP(hipMemcpyAsync(rp->d_site_pt.p, rp->site_pt.data(), sizeof(int) * rp->site_pt.size(),
                                hipMemcpyHostToDevice, sm));
    hipLaunchKernelGGL(k_rs_values, dim3((maxp + 255) / 256, num_sites), dim3(256), 0, sm,
                       (const RsSite *)rp->d_geo.p, L, fftg[0], fftg[1], fftg[2], (const int *)rp->d_site_pt.p,
                       (const int4 *)rp->d_pts.p, (const long *)rp->d_tab_off.p, (const double *)rp->d_spl.p,
                       (const int *)rp->d_chan_l.p, (cplx<double> *)rp->d_tab.p);
    PP_CHECK_HIP(hipGetLastError());
    // colours (one launch per colour; a single "colour" with atomics when
    // a box wraps onto itself)
    std::vector<int> colour;
    if (rp->atomic)
      colour.assign(num_sites, 0);
    else
      colour_sites(lattice, geo, colour);
    const int nc = *std::max_element(colour.begin(), colour.end()) + 1;
    rp->color_off.assign(nc + 1, 0);
    rp->color_maxpts.assign(nc, 0);
    for (int s = 0; s < num_sites; s++)
      rp->color_off[colour[s] + 1]++;
    for (int c = 0; c < nc; c++)
      rp->color_off[c + 1] += rp->color_off[c];
    rp->color_sites.assign(num_sites, 0);
    std::vector<int> fillc(rp->color_off.begin(), rp->color_off.end() - 1);
    for (int s = 0; s < num_sites; s++) {
      rp->color_sites[fillc[colour[s]]++] = s;
      rp->color_maxpts[colour[s]] =
          std::max(rp->color_maxpts[colour[s]], rp->site_pt[s + 1] - rp->site_pt[s]);
    }
    if ((st = rp->d_colors.ensure(sizeof(int) * num_sites)))
      return st;
    PP_CHECK_HIP(hipMemcpyAsync(rp->d_colors.p, rp->color_sites.data(), sizeof(int) * num_sites,
                                hipMemcpyHostToDevice, sm));
    PP_CHECK_HIP(hipStreamSynchronize(sm));  // host staging buffers end here
  }
  *out = hold.release();
  return PP_OK;
}

extern "C" void pp_rs_plan_destroy(pp_rsplan *rp) { delete rp; }

extern "C" int pp_rs_plan_info(pp_rsplan *rp, int *total_projs, long *total_points, int *ncolors, int *atomic,
                               int *proj_off) {
  PP_REQUIRE(rp, PP_ERR_ARG, "pp_rs_plan_info: null plan");
  if (total_projs)
    *total_projs = rp->ptot;
  if (total_points)
    *total_points = rp->npts;
  if (ncolors)
    *ncolors = rp->color_off.empty() ? 0 : (int)rp->color_off.size() - 1;
  if (atomic)
    *atomic = rp->atomic ? 1 : 0;
  if (proj_off)
    memcpy(proj_off, rp->proj_off.data(), sizeof(int) * rp->proj_off.size());
  return PP_OK;
}

namespace {

// Plane-wave sphere bounds of one k-point on the plan's grid and the
// populated lines of the z and y passes (one host synchronisation), so the
// passes skip empty lines and the chunk needs no zero fill.  prune is
// false when the grid aliases a G vector or the G set does not fill its
// z-line intervals (the passes then run in full).
int rs_bounds(pp_ctx *ctx, const int *fftg, int nw, const int *G_dev, SphereBounds &bd, bool &prune) {
  const int n0 = fftg[0], n1 = fftg[1], n2 = fftg[2];
  const int nbnd = 2 * n0 * n1 + 2 * n0 + 3;
  prune = false;
  if (nw == 0 || !ctx->tune_rs_prune)
    return PP_OK;
  PP_TRY(ctx->bounds.ensure(sizeof(int) * (size_t)nbnd));
  int *bnd = (int *)ctx->bounds.p;
  PP_TRY(sphere_bounds(ctx, fftg, nw, G_dev, bnd));
  ctx->bounds_host.resize(nbnd);
  PP_CHECK_HIP(hipMemcpyAsync(ctx->bounds_host.data(), bnd, sizeof(int) * nbnd, hipMemcpyDeviceToHost, ctx->stream));
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  const int *hb = ctx->bounds_host.data();
  if (hb[nbnd - 1] || !bounds_filled(hb, fftg, nw))
    return PP_OK;  // aliasing, or a G set with holes: no pruning
  bd = SphereBounds{bnd, bnd + 2 * n0 * n1, bnd + 2 * n0 * n1 + 2 * n0, n0, n1, n2};
  const int *yb = hb + 2 * n0 * n1, *xb = yb + 2 * n0;
  ctx->lines_host.clear();
  for (int sx = xb[0]; sx <= xb[1]; sx++) {
    const int x = (sx + n0) % n0;
    for (int sy = yb[2 * x]; sy <= yb[2 * x + 1]; sy++)
      ctx->lines_host.push_back(x * n1 + (sy + n1) % n1);
  }
  const int nzl = (int)ctx->lines_host.size();
  for (int sx = xb[0]; sx <= xb[1]; sx++)
    for (int z = 0; z < n2; z++)
      ctx->lines_host.push_back(((sx + n0) % n0) * n2 + z);
  if (!ctx->lines_host.empty()) {
    PP_TRY(ctx->lines.ensure(sizeof(int) * ctx->lines_host.size()));
    // the stream is idle (synchronised above); lines_host stays in the
    // context until the next call's synchronisation
    PP_CHECK_HIP(hipMemcpyAsync(ctx->lines.p, ctx->lines_host.data(), sizeof(int) * ctx->lines_host.size(),
                                hipMemcpyHostToDevice, ctx->stream));
    bd.zl = (const int *)ctx->lines.p;
    bd.nzl = nzl;
    bd.yl = (const int *)ctx->lines.p + nzl;
    bd.nyl = (int)ctx->lines_host.size() - nzl;
  }
  prune = true;
  return PP_OK;
}

template <typename T>
int rs_chunk(pp_ctx *ctx, pp_rsplan *rp, FftPlan *pl, const double *kf, int nw, const int *gidx,
             const cplx<float> *C, long ldc, int nbc, const cplx<double> *proj, long ldp, int remove_phase,
             cplx<double> *states, long ldo, const double *wts, double *dens, const SphereBounds *bd,
             cplx<T> *grid) {
  const int NB = nbc;
  const long npts = (long)rp->fftg[0] * rp->fftg[1] * rp->fftg[2];
  PP_TRY(fft_chunk<T>(ctx, pl, grid, NB, nw, gidx, C, ldc, nbc, bd));
  const double scl = 1.0 / sqrt(det3(rp->lattice));
  hipEvent_t ev;
  if (rp->nsites && rp->ptot) {
    tbegin(ctx, K_RS_AUG, &ev);
    const int ppb = 64;
    const size_t lds = sizeof(cplx<double>) * 64 * (size_t)std::max(1, rp->pmax);
    const int nc = (int)rp->color_off.size() - 1;
    for (int c = 0; c < nc; c++) {
      const int nsc = rp->color_off[c + 1] - rp->color_off[c];
      const int gx = std::max(1, (rp->color_maxpts[c] + ppb - 1) / ppb);
      const int *sites = (const int *)rp->d_colors.p + rp->color_off[c];
#define RS_AUG_R(AT, RP)                                                                                         \
  hipLaunchKernelGGL((k_rs_augment_r<T, AT, RP, 4>), dim3(gx, nsc), dim3(256), 0, ctx->stream, grid, NB, nbc,        \
                     (const RsSite *)rp->d_geo.p, sites, (const int *)rp->d_site_pt.p, (const int4 *)rp->d_pts.p,   \
                     (const long *)rp->d_tab_off.p, (const cplx<double> *)rp->d_tab.p, proj, ldp, kf[0], kf[1],      \
                     kf[2], 1.0 / scl, rp->fftg[0], rp->fftg[1], rp->fftg[2], ppb)
      if (rp->pmax <= 18 && ctx->tune_rs_kernel == 2 && !rp->atomic) {
        const size_t tl = (sizeof(cplx<double>) * ((size_t)ppb * rp->pmax + ppb) + sizeof(int) * ppb);
        const int u = ctx->tune_rs_unroll;
        if (u >= 4)
          hipLaunchKernelGGL((k_rs_augment_t<T, 18, 4>), dim3(gx, nsc), dim3(256), tl, ctx->stream, grid, NB, nbc,
                             (const RsSite *)rp->d_geo.p, sites, (const int *)rp->d_site_pt.p,
                             (const int4 *)rp->d_pts.p, (const long *)rp->d_tab_off.p,
                             (const cplx<double> *)rp->d_tab.p, proj, ldp, kf[0], kf[1], kf[2], 1.0 / scl,
                             rp->fftg[0], rp->fftg[1], rp->fftg[2], ppb);
        else if (u == 2)
          hipLaunchKernelGGL((k_rs_augment_t<T, 18, 2>), dim3(gx, nsc), dim3(256), tl, ctx->stream, grid, NB, nbc,
                             (const RsSite *)rp->d_geo.p, sites, (const int *)rp->d_site_pt.p,
                             (const int4 *)rp->d_pts.p, (const long *)rp->d_tab_off.p,
                             (const cplx<double> *)rp->d_tab.p, proj, ldp, kf[0], kf[1], kf[2], 1.0 / scl,
                             rp->fftg[0], rp->fftg[1], rp->fftg[2], ppb);
        else
          hipLaunchKernelGGL((k_rs_augment_t<T, 18, 1>), dim3(gx, nsc), dim3(256), tl, ctx->stream, grid, NB, nbc,
                             (const RsSite *)rp->d_geo.p, sites, (const int *)rp->d_site_pt.p,
                             (const int4 *)rp->d_pts.p, (const long *)rp->d_tab_off.p,
                             (const cplx<double> *)rp->d_tab.p, proj, ldp, kf[0], kf[1], kf[2], 1.0 / scl,
                             rp->fftg[0], rp->fftg[1], rp->fftg[2], ppb);
      } else if (rp->pmax <= 18 && ctx->tune_rs_kernel == 1) {
        if (rp->atomic)
          RS_AUG_R(true, 18);
        else
          RS_AUG_R(false, 18);
      } else if (rp->pmax <= 32 && ctx->tune_rs_kernel == 1) {
        if (rp->atomic)
          RS_AUG_R(true, 32);
        else
          RS_AUG_R(false, 32);
      } else {
#define RS_AUG_L(AT, U)                                                                                          \
  hipLaunchKernelGGL((k_rs_augment<T, AT, U>), dim3(gx, nsc), dim3(256), lds, ctx->stream, grid, NB, nbc,           \
                     (const RsSite *)rp->d_geo.p, sites, (const int *)rp->d_site_pt.p, (const int4 *)rp->d_pts.p,   \
                     (const long *)rp->d_tab_off.p, (const cplx<double> *)rp->d_tab.p, proj, ldp, kf[0], kf[1],      \
                     kf[2], 1.0 / scl, rp->fftg[0], rp->fftg[1], rp->fftg[2], ppb)
        const int u = ctx->tune_rs_unroll;
        if (rp->atomic)
          RS_AUG_L(true, 1);
        else if (u >= 4)
          RS_AUG_L(false, 4);
        else if (u == 2)
          RS_AUG_L(false, 2);
        else
          RS_AUG_L(false, 1);
#undef RS_AUG_L
      }
#undef RS_AUG_R
      PP_CHECK_HIP(hipGetLastError());
    }
    tend(ctx, K_RS_AUG, ev);
  }
  tbegin(ctx, K_RS_EMIT, &ev);
  if (states) {
    hipLaunchKernelGGL(k_rs_state<T>, dim3((unsigned)((npts + RS_TP - 1) / RS_TP)), dim3(256), 0, ctx->stream, grid,
                       NB, nbc, npts, scl, rp->fftg[0], rp->fftg[1], rp->fftg[2], kf[0], kf[1], kf[2],
                       remove_phase ? 0 : 1, states, ldo);
    PP_CHECK_HIP(hipGetLastError());
  }
  if (dens) {
    RsWeights W;
    for (int b = 0; b < 64; b++)
      W.w[b] = b < nbc ? wts[b] : 0.0;
    const int ppw = 16;
    const long blocks = (npts + 4L * ppw - 1) / (4L * ppw);
    hipLaunchKernelGGL(k_rs_density<T>, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, grid, NB, nbc, npts,
                       scl * scl, W, ppw, dens);
    PP_CHECK_HIP(hipGetLastError());
  }
  tend(ctx, K_RS_EMIT, ev);
  return PP_OK;
}

}  // namespace

extern "C" int pp_realspace_batched(pp_ctx *ctx, pp_rsplan *rp, const double *kfrac, int nw, const int *G_dev,
                                    const void *C_dev, long ldc, int nband, const void *proj_dev, long ldp,
                                    int remove_phase, void *states_dev, long ldo, const double *weights,
                                    double *density_dev, int precision) {
  PP_REQUIRE(ctx && rp && kfrac && nw >= 0 && nband >= 0 && ldc >= nw, PP_ERR_ARG,
             "pp_realspace_batched: bad arguments");
  PP_REQUIRE(precision == PP_FP64 || precision == PP_FP32, PP_ERR_ARG, "pp_realspace_batched: precision");
  PP_REQUIRE(states_dev || density_dev, PP_ERR_ARG, "pp_realspace_batched: no output");
  PP_REQUIRE(!density_dev || weights, PP_ERR_ARG, "pp_realspace_batched: density needs weights");
  const long npts = (long)rp->fftg[0] * rp->fftg[1] * rp->fftg[2];
  PP_REQUIRE(!states_dev || ldo >= npts, PP_ERR_ARG, "pp_realspace_batched: ldo below the grid size");
  PP_REQUIRE(rp->ptot == 0 || (proj_dev && ldp >= rp->ptot), PP_ERR_ARG,
             "pp_realspace_batched: projections missing or ldp below the plan's projector count");
  PP_REQUIRE(nw == 0 || (G_dev && C_dev), PP_ERR_ARG, "pp_realspace_batched: null coefficients");
  if (nband == 0)
    return PP_OK;
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  FftPlan *pl = nullptr;
  PP_TRY(get_fft_plan(ctx, rp->fftg, &pl));
  const bool f64 = precision == PP_FP64;
  const int nbmax = std::min(64, nband);
  const size_t csz = f64 ? sizeof(cplx<double>) : sizeof(cplx<float>);
  PP_TRY(ctx->grid.ensure(csz * (size_t)npts * nbmax));
  PP_TRY(ctx->gidx.ensure(sizeof(int) * (size_t)std::max(1, nw)));
  PP_TRY(compute_gidx(ctx, rp->fftg, nw, G_dev, (int *)ctx->gidx.p));
  SphereBounds bd{};
  bool prune = false;
  PP_TRY(rs_bounds(ctx, rp->fftg, nw, G_dev, bd, prune));
  // chunks to run (a density-only call skips chunks whose weights are all 0)
  std::vector<int> starts;
  for (int b0 = 0; b0 < nband; b0 += 64) {
    bool any = !density_dev || states_dev;
    for (int b = 0; b < std::min(64, nband - b0) && !any; b++)
      any = weights[b0 + b] != 0.0;
    if (any)
      starts.push_back(b0);
  }
  // two chunks in flight on two streams (default priority), each with its
  // own chunk buffer; the second stream's densities go to a buffer of the
  // plan that is added to density_dev at the end
  const bool two = ctx->tune_streams > 1 && starts.size() > 1;
  hipStream_t streams[2] = {ctx->stream, ctx->stream};
  if (two) {
    PP_TRY(ctx->grid2.ensure(csz * (size_t)npts * nbmax));
    if (!ctx->aux)
      PP_CHECK_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    if (!ctx->fork)
      PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
    if (!ctx->join)
      PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
    if (density_dev)
      PP_TRY(rp->dens2.ensure(sizeof(double) * (size_t)npts));
    streams[1] = ctx->aux;
    PP_CHECK_HIP(hipEventRecord(ctx->fork, ctx->stream));
    PP_CHECK_HIP(hipStreamWaitEvent(ctx->aux, ctx->fork, 0));
    if (density_dev) {
      ctx->stream = ctx->aux;
      int zs = zero_async(ctx, rp->dens2.p, sizeof(double) * (size_t)npts);
      ctx->stream = streams[0];
      PP_TRY(zs);
    }
  }
  int st = PP_OK;
  for (size_t c = 0; c < starts.size() && st == PP_OK; c++) {
    const int b0 = starts[c];
    const int nbc = std::min(64, nband - b0);
    const int sl = two ? (int)(c & 1) : 0;
    ctx->stream = streams[sl];
    ctx->slot = sl;
    void *grid = sl ? ctx->grid2.p : ctx->grid.p;
    double *dens = density_dev ? (sl ? (double *)rp->dens2.p : density_dev) : nullptr;
    const cplx<float> *C = (const cplx<float> *)C_dev + (size_t)b0 * ldc;
    const cplx<double> *P = proj_dev ? (const cplx<double> *)proj_dev + (size_t)b0 * ldp : nullptr;
    cplx<double> *S = states_dev ? (cplx<double> *)states_dev + (size_t)b0 * ldo : nullptr;
    const double *W = density_dev ? weights + b0 : nullptr;
    if (f64)
      st = rs_chunk<double>(ctx, rp, pl, kfrac, nw, (const int *)ctx->gidx.p, C, ldc, nbc, P, ldp, remove_phase, S,
                            ldo, W, dens, prune ? &bd : nullptr, (cplx<double> *)grid);
    else
      st = rs_chunk<float>(ctx, rp, pl, kfrac, nw, (const int *)ctx->gidx.p, C, ldc, nbc, P, ldp, remove_phase, S,
                           ldo, W, dens, prune ? &bd : nullptr, (cplx<float> *)grid);
  }
  ctx->stream = streams[0];
  ctx->slot = 0;
  if (two) {
    // the caller's stream sees everything the second stream did
    (void)hipEventRecord(ctx->join, ctx->aux);
    (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
    if (st == PP_OK && density_dev) {
      hipLaunchKernelGGL(k_rs_add, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, ctx->stream, density_dev,
                         (const double *)rp->dens2.p, npts);
      PP_CHECK_HIP(hipGetLastError());
    }
  }
  return st;
}

namespace pp {

// x[pt] *= e^{-2 pi i k.r} (density.c:312-326)
__global__ __launch_bounds__(256) void k_rs_unphase(cplx<double> *__restrict__ x, long npts, int f0, int f1, int f2,
                                                    double k0, double k1, double k2) {
  const long pt = (long)blockIdx.x * 256 + threadIdx.x;
  if (pt >= npts)
    return;
  const int kk = (int)(pt % f2), jj = (int)((pt / f2) % f1), ii = (int)(pt / ((long)f1 * f2));
  double sn, cs;
  sincospi(-2.0 * (k0 * ii / f0 + k1 * jj / f1 + k2 * kk / f2), &sn, &cs);
  x[pt] = x[pt] * mk<double>(cs, sn);
}

// part[b * nparts + q] = sum over slice q of conj(R[b][r]) S[r]
__global__ __launch_bounds__(256) void k_rs_dot(const cplx<double> *__restrict__ R, long ldr,
                                                const cplx<double> *__restrict__ S, long n, int nparts,
                                                cplx<double> *__restrict__ part) {
  __shared__ double sr[256], si[256];
  const int b = blockIdx.y, q = blockIdx.x;
  const long chunk = (n + nparts - 1) / nparts;
  const long r0 = (long)q * chunk, r1 = min(n, r0 + chunk);
  double ar = 0.0, ai = 0.0;
  for (long r = r0 + threadIdx.x; r < r1; r += 256) {
    const cplx<double> a = R[(long)b * ldr + r], s = S[r];
    ar += a.x * s.x + a.y * s.y;
    ai += a.x * s.y - a.y * s.x;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sr[threadIdx.x] += sr[threadIdx.x + o];
      si[threadIdx.x] += si[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    part[(long)b * nparts + q] = mk<double>(sr[0], si[0]);
}

__global__ __launch_bounds__(64) void k_rs_dot_sum(const cplx<double> *__restrict__ part, int nparts, int nb,
                                                   double scale, cplx<double> *__restrict__ out) {
  const int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= nb)
    return;
  double ar = 0.0, ai = 0.0;
  for (int q = 0; q < nparts; q++) {
    ar += part[(long)b * nparts + q].x;
    ai += part[(long)b * nparts + q].y;
  }
  out[b] = mk<double>(scale * ar, scale * ai);
}

int rs_remove_phase(pp_ctx *ctx, cplx<double> *x, const int *fftg, const double *kf) {
  const long npts = (long)fftg[0] * fftg[1] * fftg[2];
  hipLaunchKernelGGL(k_rs_unphase, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, ctx->stream, x, npts, fftg[0],
                     fftg[1], fftg[2], kf[0], kf[1], kf[2]);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

int rs_dots(pp_ctx *ctx, const cplx<double> *R, long ldr, int nb, const cplx<double> *S, long n, double scale,
            cplx<double> *out) {
  if (nb <= 0)
    return PP_OK;
  const int nparts = (int)std::min<long>(256, std::max<long>(1, n / 4096));
  DevBuf part;
  PP_TRY(part.ensure(sizeof(cplx<double>) * (size_t)nb * nparts));
  hipLaunchKernelGGL(k_rs_dot, dim3(nparts, nb), dim3(256), 0, ctx->stream, R, ldr, S, n, nparts,
                     (cplx<double> *)part.p);
  PP_CHECK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_rs_dot_sum, dim3((nb + 63) / 64), dim3(64), 0, ctx->stream, (const cplx<double> *)part.p,
                     nparts, nb, scale, out);
  PP_CHECK_HIP(hipGetLastError());
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));  // part is released on return
  return PP_OK;
}

}  // namespace pp
