// Projector spheres: table build and the <p_i|psi_nk> sphere sums.
//
// Table build replaces projector.c:193-209 / utils.c:590-696 (serial in the
// reference): per site, every grid point of the bounding box
// +-(int(|a_j x a_k| rmax / V * N_i) + 1) around round(frac * N) with
// |r - R| < (n-1)/n * rmax.  The points of a site are listed line by line
// (j, k outer, i innermost; the reference loops i-j-k, which only changes
// the order of the floating-point sums), so the points of one site on one
// x-line (j, k) form one run of consecutive i, a segment.  Each point
// stores its wrapped grid index, its unwrapped path r - R (for the Bloch
// phase) and, for every projector j = (channel, l, m) of the site, the
// real value t_j of the projector in real-harmonic form:
//   t_(l,m)  = A_lm cos(m phi)   (m >= 0),
//   t_(l,-m) = A_lm sin(m phi)   (m > 0),
// A_lm = N_lm P_l^m(cos theta) spline_l(|r|), evaluated on the
// minimum-image vector (utils.c:547-576).
//
// Sphere sums replace projector.c:223-274 (onto_projector_helper):
// <p_j|psi> = sum_i conj(p_j(i)) psi(idx_i) dv exp(i k.path_i).  With
// s_j = sum_i t_j(i) psi(idx_i) dv exp(i k.path_i) and
// Y_l,-m = (-1)^m conj(Y_lm) (utils.c:377-386, 441-449):
//   <p_(l,m)|psi>  = s_(l,m) - i s_(l,-m),
//   <p_(l,-m)|psi> = (-1)^m (s_(l,m) + i s_(l,-m)),
// so the sums run over real table values (half the flops and bytes of a
// complex table) and the reduction kernel forms the complex projections.
// One work-group per (site, part of its points); lane = band of the chunk,
// so each point's NB bands are one contiguous load from the band-fastest
// grid while the projector values and phase are uniform across the wave.
#include <cmath>
#include <cstring>

#include "pp_geom.h"
#include "pp_sphere.h"

namespace pp {

struct Lat9 {
  double m[9];
};

// decode candidate c of a site's box into (i, j, k): i fastest, then k,
// then j; c % (2 g0 + 1) is the position of i on its line
__device__ __forceinline__ void box_ijk(const SiteGeo &s, int c, int &i, int &j, int &k) {
  int w0 = 2 * s.g[0] + 1, w2 = 2 * s.g[2] + 1;
  i = c % w0 - s.g[0] + s.c[0];
  int t = c / w0;
  k = t % w2 - s.g[2] + s.c[2];
  j = t / w2 - s.g[1] + s.c[1];
}

__device__ __forceinline__ bool inside(const SiteGeo &s, const Lat9 &L, const int *fftg, int i, int j, int k,
                                       double *cv) {
  double fv[3] = {(double)i / fftg[0] - s.coord[0], (double)j / fftg[1] - s.coord[1],
                  (double)k / fftg[2] - s.coord[2]};
  d_frac_to_cart(fv, L.m, cv);
  return sqrt(cv[0] * cv[0] + cv[1] * cv[1] + cv[2] * cv[2]) < s.R0;
}

// Candidate c of a site's box is inside its sphere.
__device__ __forceinline__ bool cand_inside(const SiteGeo &s, const Lat9 &L, const int *fftg, int c, double *cv) {
  int i, j, k;
  box_ijk(s, c, i, j, k);
  return inside(s, L, fftg, i, j, k, cv);
}

// points (counts[s]) and segments (counts[nsites + s]) of every site.  The
// candidates are visited in tiles of 256 consecutive ones; a segment starts
// at an inside candidate that is the first of its line or follows an
// outside one, judged from the same inside flags as the points.
__global__ __launch_bounds__(256) void k_site_count(const SiteGeo *__restrict__ geo, Lat9 L, int f0, int f1, int f2,
                                                    int nsites, int *__restrict__ counts) {
  const SiteGeo s = geo[blockIdx.x];
  const int fftg[3] = {f0, f1, f2};
  const int w0 = 2 * s.g[0] + 1;
  const int ncand = w0 * (2 * s.g[1] + 1) * (2 * s.g[2] + 1);
  __shared__ int flags[256], red[256], reds[256];
  __shared__ int prev_last;
  if (threadIdx.x == 0)
    prev_last = 0;
  int cnt = 0, nseg = 0;
  for (int c0 = 0; c0 < ncand; c0 += 256) {
    const int c = c0 + threadIdx.x;
    double cv[3];
    const int in = (c < ncand && cand_inside(s, L, fftg, c, cv)) ? 1 : 0;
    flags[threadIdx.x] = in;
    __syncthreads();
    const int prev = threadIdx.x > 0 ? flags[threadIdx.x - 1] : prev_last;
    cnt += in;
    nseg += (in && (c % w0 == 0 || !prev)) ? 1 : 0;
    __syncthreads();
    if (threadIdx.x == 255)
      prev_last = flags[255];
    __syncthreads();
  }
  red[threadIdx.x] = cnt;
  reds[threadIdx.x] = nseg;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[threadIdx.x] += red[threadIdx.x + o];
      reds[threadIdx.x] += reds[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counts[blockIdx.x] = red[0];
    counts[nsites + blockIdx.x] = reds[0];
  }
}

// Ordered compaction of the inside candidates and of the segment starts
// (block scans per tile of 256).  A segment is stored as (first point,
// length, wrapped i of the first point, site) and its wrapped line
// j * N2 + k; the lengths are filled in by k_seg_len.
__global__ __launch_bounds__(256) void k_site_fill(const SiteGeo *__restrict__ geo, Lat9 L, int f0, int f1, int f2,
                                                   const int *__restrict__ site_pt, const int *__restrict__ seg_off,
                                                   int *__restrict__ pidx, double *__restrict__ path,
                                                   int4 *__restrict__ seg, int *__restrict__ segline) {
  const SiteGeo s = geo[blockIdx.x];
  const int fftg[3] = {f0, f1, f2};
  const int w0 = 2 * s.g[0] + 1;
  const int ncand = w0 * (2 * s.g[1] + 1) * (2 * s.g[2] + 1);
  __shared__ int flags[256], scan[256], scans[256];
  __shared__ int running, running_s, prev_last;
  if (threadIdx.x == 0) {
    running = site_pt[blockIdx.x];
    running_s = seg_off[blockIdx.x];
    prev_last = 0;
  }
  __syncthreads();
  for (int c0 = 0; c0 < ncand; c0 += 256) {
    const int c = c0 + threadIdx.x;
    int i = 0, j = 0, k = 0;
    double cv[3] = {0, 0, 0};
    int in = 0;
    if (c < ncand) {
      box_ijk(s, c, i, j, k);
      in = inside(s, L, fftg, i, j, k, cv) ? 1 : 0;
    }
    flags[threadIdx.x] = in;
    __syncthreads();
    const int prev = threadIdx.x > 0 ? flags[threadIdx.x - 1] : prev_last;
    const int st = (in && (c % w0 == 0 || !prev)) ? 1 : 0;
    scan[threadIdx.x] = in;
    scans[threadIdx.x] = st;
    __syncthreads();
    // Hillis-Steele inclusive scans
    for (int o = 1; o < 256; o <<= 1) {
      int v = threadIdx.x >= o ? scan[threadIdx.x - o] : 0;
      int vs = threadIdx.x >= o ? scans[threadIdx.x - o] : 0;
      __syncthreads();
      scan[threadIdx.x] += v;
      scans[threadIdx.x] += vs;
      __syncthreads();
    }
    int pos = running + scan[threadIdx.x] - in;
    if (in) {
      int ii = (i % fftg[0] + fftg[0]) % fftg[0];
      int jj = (j % fftg[1] + fftg[1]) % fftg[1];
      int kk = (k % fftg[2] + fftg[2]) % fftg[2];
      pidx[pos] = (ii * fftg[1] + jj) * fftg[2] + kk;
      path[3 * (long)pos + 0] = cv[0];
      path[3 * (long)pos + 1] = cv[1];
      path[3 * (long)pos + 2] = cv[2];
      if (st) {
        const int q = running_s + scans[threadIdx.x] - 1;
        seg[q] = make_int4(pos, 0, ii, blockIdx.x);
        segline[q] = jj * fftg[2] + kk;
      }
    }
    __syncthreads();
    if (threadIdx.x == 255) {
      running += scan[255];
      running_s += scans[255];
      prev_last = flags[255];
    }
    __syncthreads();
  }
}

// Coverage masks of the x-lines: every segment sets the bits of its points.
__global__ __launch_bounds__(256) void k_xmask(const int4 *__restrict__ seg, const int *__restrict__ segline,
                                               long nseg, int n0, int mw, unsigned *__restrict__ mask) {
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  if (q >= nseg)
    return;
  const int4 sg = seg[q];
  unsigned *m = mask + (long)segline[q] * mw;
  int x = sg.z;
  for (int i = 0; i < sg.y; i++) {
    atomicOr(&m[x >> 5], 1u << (x & 31));
    if (++x == n0)
      x = 0;
  }
}

// Segment lengths: the points of segment q run to the first point of the
// next segment of its site, or to the end of the site's points.
__global__ __launch_bounds__(256) void k_seg_len(int4 *__restrict__ seg, long nseg, const int *__restrict__ seg_off,
                                                 const int *__restrict__ site_pt) {
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  if (q >= nseg)
    return;
  const int4 sg = seg[q];
  const int end = q + 1 < seg_off[sg.w + 1] ? seg[q + 1].x : site_pt[sg.w + 1];
  seg[q].y = end - sg.x;
}


template <typename T>
__global__ __launch_bounds__(256) void k_site_values(const SiteGeo *__restrict__ geo, Lat9 L, int f0, int f1, int f2,
                                                     const int *__restrict__ site_pt, const long *__restrict__ tab_off,
                                                     const int *__restrict__ pidx, const double *__restrict__ spl,
                                                     const int *__restrict__ chan_l, T *__restrict__ table) {
  const int s_id = blockIdx.y;
  const SiteGeo s = geo[s_id];
  const int p0 = site_pt[s_id], np = site_pt[s_id + 1] - p0;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= np)
    return;
  const int fftg[3] = {f0, f1, f2};
  int g = pidx[p0 + i];
  int kk = g % fftg[2];
  int jj = (g / fftg[2]) % fftg[1];
  int ii = g / (fftg[1] * fftg[2]);
  double pf[3] = {(double)ii / fftg[0], (double)jj / fftg[1], (double)kk / fftg[2]};
  // minimum image of the wrapped point (utils.c:51-73, 566-576)
  double best = INFINITY, v[3] = {0, 0, 0};
  for (int a = -1; a <= 1; a++)
    for (int b = -1; b <= 1; b++)
      for (int c = -1; c <= 1; c++) {
        double fv[3] = {pf[0] + a - s.coord[0], pf[1] + b - s.coord[1], pf[2] + c - s.coord[2]}, cv[3];
        d_frac_to_cart(fv, L.m, cv);
        double d = sqrt(cv[0] * cv[0] + cv[1] * cv[1] + cv[2] * cv[2]);
        if (d < best) {
          best = d;
          v[0] = cv[0];
          v[1] = cv[1];
          v[2] = cv[2];
        }
      }
  const double r = best;
  double theta = 0.0, phi = 0.0;
  if (r != 0.0) {
    theta = acos(v[2] / r);
    if (r - fabs(v[2]) != 0.0)
      phi = acos(v[0] / sqrt(v[0] * v[0] + v[1] * v[1]));
    if (v[1] < 0)
      phi = 2.0 * M_PI - phi;
  }
  const double ct = cos(theta);
  const int n = s.ngrid;
  const double *x = spl + s.spl_off;
  T *out = table + tab_off[s_id] + (long)i * s.nproj;
  int p = 0;
  for (int ch = 0; ch < s.nchan; ch++) {
    const int l = chan_l[s.chan_off + ch];
    const double *f = x + (long)(1 + ch) * n;
    const double *bb = x + (long)(1 + s.nchan + ch) * n;
    const double *cc = x + (long)(1 + 2 * s.nchan + ch) * n;
    const double *dd = x + (long)(1 + 3 * s.nchan + ch) * n;
    // utils.c:461-475
    double radial;
    if (r > x[n - 1]) {
      radial = 0.0;
    } else if (r < x[0]) {
      radial = f[0];
    } else {
      int id = (int)(r / s.rmax * n);
      if (id > n - 2)
        id = n - 2;
      double t = r - x[id];
      radial = f[id] + t * (bb[id] + t * (cc[id] + t * dd[id]));
    }
    for (int m = -l; m <= l; m++, p++) {
      // real form: slot m >= 0 holds A_|m| cos(|m| phi), slot -|m| holds
      // A_|m| sin(|m| phi), with A_|m| = N_l|m| P_l^|m|(cos theta) radial
      const int am = m < 0 ? -m : m;
      double norm = sqrt((2 * l + 1) / (4.0 * M_PI) * d_fact(l - am) / d_fact(l + am));
      double pl = norm * d_legendre(l, am, ct) * radial;
      double sn, cs;
      sincos(am * phi, &sn, &cs);
      out[p] = (T)(m < 0 ? pl * sn : pl * cs);
    }
  }
}

// phase[pt] = dv / sqrt(V) * exp(i kc . path)
template <typename T>
__global__ __launch_bounds__(256) void k_phase(long npts, const double *__restrict__ path, double k0, double k1,
                                               double k2, double scl, cplx<T> *__restrict__ phase) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= npts)
    return;
  double kr = k0 * path[3 * i] + k1 * path[3 * i + 1] + k2 * path[3 * i + 2];
  double sn, cs;
  sincos(kr, &sn, &cs);
  phase[i] = mk<T>((T)(scl * cs), (T)(scl * sn));
}

// One work-group (4 wavefronts) per (site, part of its points); lane b
// of every wavefront accumulates the site's projectors for band b of the
// chunk.  Tiles of TP points stage their projector values, phases and grid
// indices in LDS with coalesced loads; each wavefront then takes every
// fourth point of the tile, four points per step with their grid loads in
// flight together, and reads the projector values as LDS broadcasts.  The
// four wavefronts' sums are added in LDS at the end.
template <typename T, int PMAX, bool FULL, int U>
__device__ __forceinline__ void sphere_sum_block(const cplx<T> *__restrict__ g, int NB,
                                                 const int *__restrict__ site_pt, const int *__restrict__ site_np,
                                                 const long *__restrict__ tab_off, const int *__restrict__ pidx,
                                                 const cplx<T> *__restrict__ phase, const T *__restrict__ table,
                                                 int nparts, cplx<double> *__restrict__ partial, int streaming,
                                                 int blk) {
  constexpr int TP = 64;
  __shared__ __align__(16) T s_tab[TP * PMAX];
  __shared__ cplx<T> s_ph[TP];
  __shared__ int s_idx[TP];
  const int s = blk / nparts, part = blk % nparts;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int p0 = site_pt[s], p1 = site_pt[s + 1];
  const int np = FULL ? PMAX : site_np[s];  // FULL: every site has PMAX projectors
  const int len = p1 - p0;
  const int chunk = (len + nparts - 1) / nparts;
  const int a = p0 + part * chunk;
  const int e = min(p1, a + chunk);
  const bool active = lane < NB;
  cplx<T> acc[PMAX];
#pragma unroll
  for (int j = 0; j < PMAX; j++)
    acc[j] = mk<T>(0, 0);
  const T *tsite = table + tab_off[s];
  // a tile's table values, indices and phases: every load of a thread is
  // issued before any is written to LDS, so they are in flight together
  constexpr int SL = (TP * PMAX + 255) / 256;
  T tv[SL];
  int iv = 0;
  cplx<T> pv = mk<T>(0, 0);
  auto fetch = [&](int f0) {
    const int fn = min(TP, e - f0);
    const T *tsrc = tsite + (long)(f0 - p0) * np;
#pragma unroll
    for (int u = 0; u < SL; u++) {
      const int q = tid + u * 256;
      tv[u] = q < fn * np ? tsrc[q] : (T)0;
    }
    if (tid < fn) {
      iv = pidx[f0 + tid];
      pv = phase[f0 + tid];
    }
  };
  for (int t0 = a; t0 < e; t0 += TP) {
    const int n = min(TP, e - t0);
    fetch(t0);
#pragma unroll
    for (int u = 0; u < SL; u++) {
      const int q = tid + u * 256;
      if (q < n * np)
        s_tab[q] = tv[u];
    }
    if (tid < n) {
      s_idx[tid] = iv;
      s_ph[tid] = pv;
    }
    __syncthreads();
    int i = wave;
    for (; i + 4 * (U - 1) < n; i += 4 * U) {
      cplx<T> x[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const cplx<T> *gp = g + (long)s_idx[i + 4 * u] * NB + lane;
        x[u] = active ? (streaming ? load_nt(gp) : *gp) : mk<T>(0, 0);
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const cplx<T> xv = x[u] * s_ph[i + 4 * u];
        const T *tp = s_tab + (i + 4 * u) * np;
#pragma unroll
        for (int j = 0; j < PMAX; j++)
          if (FULL || j < np) {
            acc[j].x += tp[j] * xv.x;
            acc[j].y += tp[j] * xv.y;
          }
      }
    }
    for (; i < n; i += 4) {
      cplx<T> xv = active ? g[(long)s_idx[i] * NB + lane] : mk<T>(0, 0);
      xv = xv * s_ph[i];
      const T *tp = s_tab + i * np;
#pragma unroll
      for (int j = 0; j < PMAX; j++)
        if (FULL || j < np) {
          acc[j].x += tp[j] * xv.x;
          acc[j].y += tp[j] * xv.y;
        }
    }
    __syncthreads();
  }
  // add the four wavefronts' sums, PMAX/2 projectors at a time through the
  // table tile (TP * PMAX reals = 64 lanes x PMAX/2 complex)
  cplx<T> *red = reinterpret_cast<cplx<T> *>(s_tab);
  constexpr int JH = PMAX / 2;
  for (int w = 1; w < 4; w++) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      if (wave == w) {
#pragma unroll
        for (int j = 0; j < JH; j++)
          red[j * 64 + lane] = acc[h * JH + j];
      }
      __syncthreads();
      if (wave == 0) {
#pragma unroll
        for (int j = 0; j < JH; j++)
          acc[h * JH + j] = acc[h * JH + j] + red[j * 64 + lane];
      }
      __syncthreads();
    }
  }
  if (wave == 0 && active) {
    cplx<double> *o = partial + ((long)blk * NB + lane) * PMAX;
#pragma unroll
    for (int j = 0; j < PMAX; j++)
      if (FULL || j < np)
        o[j] = cvt<double>(acc[j]);
  }
}
template <typename T, int PMAX, bool FULL, int U>
__global__ __launch_bounds__(256) void k_sphere_sum(const cplx<T> *__restrict__ g, int NB,
                                                    const int *__restrict__ site_pt, const int *__restrict__ site_np,
                                                    const long *__restrict__ tab_off, const int *__restrict__ pidx,
                                                    const cplx<T> *__restrict__ phase, const T *__restrict__ table,
                                                    int nparts, cplx<double> *__restrict__ partial, int streaming) {
  sphere_sum_block<T, PMAX, FULL, U>(g, NB, site_pt, site_np, tab_off, pidx, phase, table, nparts, partial, streaming,
                                     blockIdx.x);
}
// The sphere sums of a batch of work items (one per blockIdx.y), each with
// its own chunk, plan and phases.
template <typename T, int PMAX, bool FULL, int U>
__global__ __launch_bounds__(256) void k_sphere_sum_items(const ChunkItem *__restrict__ items, int NB, int streaming) {
  const ChunkItem &it = items[blockIdx.y];
  const int blk = blockIdx.x;
  if (blk >= it.nsites * it.nparts)
    return;
  sphere_sum_block<T, PMAX, FULL, U>((const cplx<T> *)it.grid, NB, (const int *)it.site_pt,
                                     (const int *)it.site_np, (const long *)it.tab_off, (const int *)it.pidx,
                                     (const cplx<T> *)it.phase, (const T *)it.table, it.nparts,
                                     (cplx<double> *)it.partial, streaming, blk);
}

// Sum the parts of every (site, band, projector) in FP64 and form the
// complex projections from the real-harmonic sums: thread (s, b, j) with
// m_j >= 0 writes slot j and, for m_j > 0, its partner slot j - 2 m_j.
__device__ __forceinline__ void sphere_reduce_one(const cplx<double> *__restrict__ partial, int NB, int nbc,
                                                  int PMAX, int nparts, int nsites, const int *__restrict__ site_np,
                                                  const int *__restrict__ proj_off, const int *__restrict__ proj_m,
                                                  long ptot, cplx<double> *__restrict__ out, long t) {
  long total = (long)nsites * nbc * PMAX;
  if (t >= total)
    return;
  int j = (int)(t % PMAX);
  long r = t / PMAX;
  int b = (int)(r % nbc);
  int s = (int)(r / nbc);
  if (j >= site_np[s])
    return;
  const int m = proj_m[proj_off[s] + j];
  if (m < 0)
    return;
  cplx<double> c = mk<double>(0, 0), sn = mk<double>(0, 0);
  const cplx<double> *ps = partial + ((long)s * nparts * NB + b) * PMAX;
  // the partner slot of m > 0 (m = 0 reads its own slot again and leaves
  // the sum unused), loaded without a branch and unrolled, so that several
  // parts' loads are in flight together; the sums keep their order
  const int jo = m > 0 ? j - 2 * m : j;
#pragma unroll 4
  for (int q = 0; q < nparts; q++) {
    c = c + ps[(long)q * NB * PMAX + j];
    sn = sn + ps[(long)q * NB * PMAX + jo];
  }
  cplx<double> *o = out + (long)b * ptot + proj_off[s];
  if (m == 0) {
    o[j] = c;
  } else {
    const double sg = (m & 1) ? -1.0 : 1.0;
    o[j] = mk<double>(c.x + sn.y, c.y - sn.x);
    o[j - 2 * m] = mk<double>(sg * (c.x - sn.y), sg * (c.y + sn.x));
  }
}
__global__ __launch_bounds__(256) void k_sphere_reduce(const cplx<double> *__restrict__ partial, int NB, int nbc,
                                                       int PMAX, int nparts, int nsites,
                                                       const int *__restrict__ site_np,
                                                       const int *__restrict__ proj_off,
                                                       const int *__restrict__ proj_m, long ptot,
                                                       cplx<double> *__restrict__ out) {
  sphere_reduce_one(partial, NB, nbc, PMAX, nparts, nsites, site_np, proj_off, proj_m, ptot, out,
                    (long)blockIdx.x * 256 + threadIdx.x);
}
__global__ __launch_bounds__(256) void k_sphere_reduce_items(const ChunkItem *__restrict__ items, int NB, int PMAX) {
  const ChunkItem &it = items[blockIdx.y];
  sphere_reduce_one((const cplx<double> *)it.partial, NB, it.nbc, PMAX, it.nparts, it.nsites,
                    (const int *)it.site_np, (const int *)it.proj_off, (const int *)it.proj_m, it.ptot,
                    (cplx<double> *)it.out, (long)blockIdx.x * 256 + threadIdx.x);
}

// Coefficients of the real-harmonic table slots for the expansion
// sum_j P_j p_j(r) (projector.c:324-327): with p_(l,+-m) built from the real
// slots t_(l,m) = A cos m phi, t_(l,-m) = A sin m phi (m > 0),
//   sum_m P_m Y_lm = sum_{m>0} t_(l,m) (P_m + (-1)^m P_-m)
//                 + i t_(l,-m) (P_m - (-1)^m P_-m) + t_(l,0) P_0.
// Q[(s * PMAX + j) * NB + b] for band b of the chunk; P rows of stride
// ptotP, site s of the plan at src_off[s].
template <typename T>
__global__ __launch_bounds__(256) void k_expand_coef(int nsites, int PMAX, int NB, int nbc,
                                                     const int *__restrict__ site_np, const int *__restrict__ proj_off,
                                                     const int *__restrict__ proj_m, const int *__restrict__ src_off,
                                                     const cplx<double> *__restrict__ P, long ptotP,
                                                     cplx<T> *__restrict__ Q) {
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)nsites * PMAX * NB)
    return;
  const int b = (int)(t % NB);
  const long r = t / NB;
  const int j = (int)(r % PMAX), s = (int)(r / PMAX);
  cplx<double> q = mk<double>(0, 0);
  if (j < site_np[s] && b < nbc) {
    const int m = proj_m[proj_off[s] + j];
    const cplx<double> *row = P + (long)b * ptotP + src_off[s];
    if (m == 0) {
      q = row[j];
    } else {
      const int am = m < 0 ? -m : m;
      const int jp = m > 0 ? j : j + 2 * am, jn = jp - 2 * am;  // slots of +am and -am
      const double sg = (am & 1) ? -1.0 : 1.0;
      const cplx<double> a = row[jp], c = row[jn];
      if (m > 0)
        q = mk<double>(a.x + sg * c.x, a.y + sg * c.y);
      else  // i (P_m - (-1)^m P_-m)
        q = mk<double>(-(a.y - sg * c.y), a.x - sg * c.x);
    }
  }
  Q[t] = cvt<T>(q);
}

// grid[idx][b] += conj(e^{-i k.path} sum_j Q_j t_j) for every sphere point
// (conjugated so the backward passes give the forward transform).  Lane =
// band; spheres of different sites may share points, so the adds are
// atomic (vector global float/double atomics).
template <typename T, int PMAX>
__global__ __launch_bounds__(256) void k_sphere_expand(cplx<T> *__restrict__ g, int NB, int nbc,
                                                       const int *__restrict__ site_pt, const int *__restrict__ site_np,
                                                       const long *__restrict__ tab_off, const int *__restrict__ pidx,
                                                       const cplx<T> *__restrict__ phase, const T *__restrict__ table,
                                                       const cplx<T> *__restrict__ Q, int nparts) {
  const int s = blockIdx.x / nparts, part = blockIdx.x % nparts;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p0 = site_pt[s], len = site_pt[s + 1] - p0;
  const int np = site_np[s];
  const int chunk = (len + nparts - 1) / nparts;
  const int a = p0 + part * chunk, e = min(p0 + len, a + chunk);
  if (lane >= nbc)
    return;
  cplx<T> q[PMAX];
#pragma unroll
  for (int j = 0; j < PMAX; j++)
    q[j] = j < np ? Q[((long)s * PMAX + j) * NB + lane] : mk<T>(0, 0);
  const T *tsite = table + tab_off[s];
  for (int i = a + wave; i < e; i += 4) {
    const T *tp = tsite + (long)(i - p0) * np;
    cplx<T> v = mk<T>(0, 0);
#pragma unroll
    for (int j = 0; j < PMAX; j++)
      if (j < np) {
        v.x += q[j].x * tp[j];
        v.y += q[j].y * tp[j];
      }
    const cplx<T> x = v * phase[i];
    T *dst = reinterpret_cast<T *>(g + (long)pidx[i] * NB + lane);
    atomicAdd(dst, x.x);
    atomicAdd(dst + 1, -x.y);
  }
}

// ------------------------------------------------------------------ host
static void cross(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static double nrm(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
double det3(const double *m) {
  return m[0] * m[4] * m[8] + m[1] * m[5] * m[6] + m[2] * m[3] * m[7] - m[2] * m[4] * m[6] -
         m[1] * m[3] * m[8] - m[0] * m[5] * m[7];
}

// Host spline coefficients with the reference's boundary conditions
// (utils.c:699-749): y'(x0) = secant slope, natural end.
void spline_coeff_host(const double *x, const double *y, int n, double *b, double *c, double *d) {
  std::vector<double> diag(n), up(n), rhs(n), M(n);
  double h0 = x[1] - x[0];
  diag[0] = 2.0 * h0;
  up[0] = h0;
  rhs[0] = 0.0;  // 6((y1-y0)/h0 - slope0) with slope0 the secant slope
  for (int i = 1; i < n - 1; i++) {
    double hl = x[i] - x[i - 1], hr = x[i + 1] - x[i];
    double w = hl / diag[i - 1];
    diag[i] = 2.0 * (hl + hr) - w * up[i - 1];
    up[i] = hr;
    rhs[i] = 6.0 * ((y[i + 1] - y[i]) / hr - (y[i] - y[i - 1]) / hl) - w * rhs[i - 1];
  }
  M[n - 1] = 0.0;
  for (int i = n - 2; i >= 0; i--)
    M[i] = (rhs[i] - up[i] * M[i + 1]) / diag[i];
  for (int i = 0; i < n - 1; i++) {
    double h = x[i + 1] - x[i];
    b[i] = (y[i + 1] - y[i]) / h - h * (2.0 * M[i] + M[i + 1]) / 6.0;
    c[i] = 0.5 * M[i];
    d[i] = (M[i + 1] - M[i]) / (6.0 * h);
  }
  b[n - 1] = c[n - 1] = d[n - 1] = 0.0;
}

// line-block work lists of the fused x pass (pp_fused.hip)
int fused_plan(pp_sphere *sp, const std::vector<SiteGeo> &geo);

// Sphere radius and candidate box of a site with coord set:
// R0 = (n-1)/n rmax (utils.c:651-652, the site's own species), box
// half-widths g_i = int(|a_j x a_k| rmax / V * N_i) + 1 around
// round(frac * N).
void site_box(const double *lattice, const int *fftg, double rmax, int ngrid, SiteGeo &g) {
  const double vol = det3(lattice);
  g.rmax = rmax;
  g.ngrid = ngrid;
  g.R0 = (ngrid - 1) * rmax / ngrid;
  double cr[3];
  cross(lattice + 3, lattice + 6, cr);
  g.g[0] = (int)(nrm(cr) * rmax / vol * fftg[0]) + 1;
  cross(lattice + 0, lattice + 6, cr);
  g.g[1] = (int)(nrm(cr) * rmax / vol * fftg[1]) + 1;
  cross(lattice + 0, lattice + 3, cr);
  g.g[2] = (int)(nrm(cr) * rmax / vol * fftg[2]) + 1;
  for (int a = 0; a < 3; a++)
    g.c[a] = (int)round(g.coord[a] * fftg[a]);
}

static int sphere_build(pp_sphere *sp) {
  pp_ctx *ctx = sp->ctx;
  sp->gen++;
  const int ns = sp->nsites;
  // per-site geometry
  std::vector<SiteGeo> geo(ns);
  std::vector<long> spl_off(sp->nelems);
  long so = 0;
  for (int e = 0; e < sp->nelems; e++) {
    spl_off[e] = so;
    so += (long)(1 + 4 * sp->nchan[e]) * sp->ngrid[e];
  }
  sp->nproj_site.assign(ns, 0);
  sp->proj_off.assign(ns + 1, 0);
  sp->pmax = 0;
  for (int s = 0; s < ns; s++) {
    SiteGeo &g = geo[s];
    int e = sp->labels[s];
    for (int a = 0; a < 3; a++)
      g.coord[a] = sp->frac[3 * s + a];
    site_box(sp->lattice, sp->fftg, sp->rmax[e], sp->ngrid[e], g);
    g.elem = e;
    g.nchan = sp->nchan[e];
    int np = 0;
    for (int ch = 0; ch < g.nchan; ch++)
      np += 2 * sp->ls[sp->ls_off[e] + ch] + 1;
    g.nproj = np;
    g.spl_off = spl_off[e];
    g.chan_off = sp->ls_off[e];
    sp->nproj_site[s] = np;
    sp->proj_off[s + 1] = sp->proj_off[s] + np;
    if (np > sp->pmax)
      sp->pmax = np;
    if (s == 0 || np < sp->pmin)
      sp->pmin = np;
  }
  sp->ptot = sp->proj_off[ns];
  PP_REQUIRE(sp->pmax <= 64, PP_ERR_UNSUPPORTED, "more than 64 projectors on one site");
  std::vector<int> proj_m(std::max(1, sp->ptot));
  for (int s = 0, q = 0; s < ns; s++) {
    int e = sp->labels[s];
    for (int ch = 0; ch < sp->nchan[e]; ch++) {
      int l = sp->ls[sp->ls_off[e] + ch];
      for (int m = -l; m <= l; m++)
        proj_m[q++] = m;
    }
  }
  PP_TRY(sp->d_proj_m.ensure(sizeof(int) * proj_m.size()));
  PP_CHECK_HIP(hipMemcpyAsync(sp->d_proj_m.p, proj_m.data(), sizeof(int) * proj_m.size(), hipMemcpyHostToDevice,
                              ctx->stream));
  Lat9 L;
  memcpy(L.m, sp->lattice, sizeof(L.m));
  PP_TRY(sp->d_sitemeta.ensure(sizeof(SiteGeo) * ns));
  PP_CHECK_HIP(hipMemcpyAsync(sp->d_sitemeta.p, geo.data(), sizeof(SiteGeo) * ns, hipMemcpyHostToDevice, ctx->stream));
  PP_TRY(sp->d_counts.ensure(sizeof(int) * 2 * ns));
  hipEvent_t ev;
  tbegin(ctx, K_TABLE, &ev);
  hipLaunchKernelGGL(k_site_count, dim3(ns), dim3(256), 0, ctx->stream, (const SiteGeo *)sp->d_sitemeta.p, L,
                     sp->fftg[0], sp->fftg[1], sp->fftg[2], ns, (int *)sp->d_counts.p);
  tend(ctx, K_TABLE, ev);
  PP_CHECK_HIP(hipGetLastError());
  std::vector<int> counts(2 * ns);
  PP_CHECK_HIP(
      hipMemcpyAsync(counts.data(), sp->d_counts.p, sizeof(int) * 2 * ns, hipMemcpyDeviceToHost, ctx->stream));
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  sp->site_pt_h.assign(ns + 1, 0);
  sp->seg_off_h.assign(ns + 1, 0);
  std::vector<long> tab_off(ns);
  long to = 0;
  int maxcount = 0;
  for (int s = 0; s < ns; s++) {
    sp->site_pt_h[s + 1] = sp->site_pt_h[s] + counts[s];
    sp->seg_off_h[s + 1] = sp->seg_off_h[s] + counts[ns + s];
    tab_off[s] = to;
    to += (long)counts[s] * sp->nproj_site[s];
    if (counts[s] > maxcount)
      maxcount = counts[s];
  }
  sp->npts = sp->site_pt_h[ns];
  sp->nseg = sp->seg_off_h[ns];
  size_t csz = sp->precision == PP_FP64 ? sizeof(double) : sizeof(float);
  PP_TRY(sp->d_site_pt.ensure(sizeof(int) * (ns + 1)));
  PP_TRY(sp->d_site_np.ensure(sizeof(int) * ns));
  PP_TRY(sp->d_tab_off.ensure(sizeof(long) * ns));
  PP_TRY(sp->d_proj_off.ensure(sizeof(int) * (ns + 1)));
  PP_TRY(sp->d_pidx.ensure(sizeof(int) * std::max<long>(1, sp->npts)));
  PP_TRY(sp->d_path.ensure(sizeof(double) * 3 * std::max<long>(1, sp->npts)));
  PP_TRY(sp->d_table.ensure(csz * std::max<long>(1, to)));
  PP_TRY(sp->d_seg_off.ensure(sizeof(int) * (ns + 1)));
  PP_TRY(sp->d_seg.ensure(sizeof(int4) * std::max<long>(1, sp->nseg)));
  PP_TRY(sp->d_segline.ensure(sizeof(int) * std::max<long>(1, sp->nseg)));
  PP_CHECK_HIP(hipMemcpyAsync(sp->d_seg_off.p, sp->seg_off_h.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice,
                              ctx->stream));
  PP_CHECK_HIP(hipMemcpyAsync(sp->d_site_pt.p, sp->site_pt_h.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice,
                              ctx->stream));
  PP_CHECK_HIP(hipMemcpyAsync(sp->d_site_np.p, sp->nproj_site.data(), sizeof(int) * ns, hipMemcpyHostToDevice,
                              ctx->stream));
  PP_CHECK_HIP(hipMemcpyAsync(sp->d_tab_off.p, tab_off.data(), sizeof(long) * ns, hipMemcpyHostToDevice, ctx->stream));
  PP_CHECK_HIP(hipMemcpyAsync(sp->d_proj_off.p, sp->proj_off.data(), sizeof(int) * (ns + 1), hipMemcpyHostToDevice,
                              ctx->stream));
  tbegin(ctx, K_TABLE, &ev);
  hipLaunchKernelGGL(k_site_fill, dim3(ns), dim3(256), 0, ctx->stream, (const SiteGeo *)sp->d_sitemeta.p, L,
                     sp->fftg[0], sp->fftg[1], sp->fftg[2], (const int *)sp->d_site_pt.p,
                     (const int *)sp->d_seg_off.p, (int *)sp->d_pidx.p, (double *)sp->d_path.p, (int4 *)sp->d_seg.p,
                     (int *)sp->d_segline.p);
  if (sp->nseg > 0)
    hipLaunchKernelGGL(k_seg_len, dim3((unsigned)((sp->nseg + 255) / 256)), dim3(256), 0, ctx->stream,
                       (int4 *)sp->d_seg.p, sp->nseg, (const int *)sp->d_seg_off.p, (const int *)sp->d_site_pt.p);
  tend(ctx, K_TABLE, ev);
  PP_CHECK_HIP(hipGetLastError());
  // x-line coverage masks (the x pass writes only covered points)
  sp->xmask_words = (sp->fftg[0] + 31) / 32;
  const size_t mbytes = sizeof(unsigned) * (size_t)sp->fftg[1] * sp->fftg[2] * sp->xmask_words;
  PP_TRY(sp->d_xmask.ensure(mbytes));
  PP_TRY(zero_async(ctx, sp->d_xmask.p, mbytes));
  if (sp->nseg > 0) {
    hipLaunchKernelGGL(k_xmask, dim3((unsigned)((sp->nseg + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const int4 *)sp->d_seg.p, (const int *)sp->d_segline.p, sp->nseg, sp->fftg[0],
                       sp->xmask_words, (unsigned *)sp->d_xmask.p);
    PP_CHECK_HIP(hipGetLastError());
  }
  if (maxcount > 0) {
    dim3 grd((maxcount + 255) / 256, ns);
    tbegin(ctx, K_TABLE, &ev);
    if (sp->precision == PP_FP64)
      hipLaunchKernelGGL(k_site_values<double>, grd, dim3(256), 0, ctx->stream, (const SiteGeo *)sp->d_sitemeta.p,
                         L, sp->fftg[0], sp->fftg[1], sp->fftg[2], (const int *)sp->d_site_pt.p,
                         (const long *)sp->d_tab_off.p, (const int *)sp->d_pidx.p, (const double *)sp->d_spline.p,
                         (const int *)sp->d_chan.p, (double *)sp->d_table.p);
    else
      hipLaunchKernelGGL(k_site_values<float>, grd, dim3(256), 0, ctx->stream, (const SiteGeo *)sp->d_sitemeta.p, L,
                         sp->fftg[0], sp->fftg[1], sp->fftg[2], (const int *)sp->d_site_pt.p,
                         (const long *)sp->d_tab_off.p, (const int *)sp->d_pidx.p, (const double *)sp->d_spline.p,
                         (const int *)sp->d_chan.p, (float *)sp->d_table.p);
    tend(ctx, K_TABLE, ev);
    PP_CHECK_HIP(hipGetLastError());
  }
  // the fused x pass's work lists are built when that path is first used
  sp->geo_h = geo;
  sp->fx_built = false;
  sp->fused_ok = false;
  return ctx->tune_fused ? fused_plan(sp, geo) : PP_OK;
}

template <typename T, int PMAX>
static void launch_sum(pp_ctx *ctx, pp_sphere *sp, const cplx<T> *grid, int NB, const cplx<T> *phase, int nparts,
                       cplx<double> *partial) {
#define PP_SS(FULL, U)                                                                                            \
  hipLaunchKernelGGL((k_sphere_sum<T, PMAX, FULL, U>), dim3(sp->nsites * nparts), dim3(256), 0, ctx->stream, grid,  \
                     NB, (const int *)sp->d_site_pt.p, (const int *)sp->d_site_np.p, (const long *)sp->d_tab_off.p, \
                     (const int *)sp->d_pidx.p, phase, (const T *)sp->d_table.p, nparts, partial,             \
                     ctx->tune_sphere_nt)
  const bool u8 = ctx->tune_sphere_unroll == 8;
  if (sp->pmin == PMAX) {
    if (u8)
      PP_SS(true, 8);
    else
      PP_SS(true, 4);
  } else {
    if (u8)
      PP_SS(false, 8);
    else
      PP_SS(false, 4);
  }
#undef PP_SS
}

int sphere_parts_of(pp_ctx *ctx, pp_sphere *sp);
int sphere_pm_of(pp_sphere *sp, bool *full);

template <typename T>
int sphere_sums(pp_ctx *ctx, pp_sphere *sp, const cplx<T> *grid, int NB, int nbc, const cplx<T> *phase,
                cplx<double> *out, long ptot) {
  if (sp->nsites == 0 || nbc == 0)
    return PP_OK;
  const int nparts = sphere_parts_of(ctx, sp);
  const int PM = sphere_pm_of(sp, nullptr);
  DevBuf &pbuf = ctx->slot ? ctx->partial2 : ctx->partial;
  PP_TRY(pbuf.ensure(sizeof(cplx<double>) * (size_t)sp->nsites * nparts * NB * PM));
  cplx<double> *partial = (cplx<double> *)pbuf.p;
  hipEvent_t ev;
  tbegin(ctx, K_SPHERE_SUM, &ev);
  switch (PM) {
  case 8:
    launch_sum<T, 8>(ctx, sp, grid, NB, phase, nparts, partial);
    break;
  case 18:
    launch_sum<T, 18>(ctx, sp, grid, NB, phase, nparts, partial);
    break;
  case 32:
    launch_sum<T, 32>(ctx, sp, grid, NB, phase, nparts, partial);
    break;
  default:
    launch_sum<T, 64>(ctx, sp, grid, NB, phase, nparts, partial);
  }
  tend(ctx, K_SPHERE_SUM, ev);
  PP_CHECK_HIP(hipGetLastError());
  long total = (long)sp->nsites * nbc * PM;
  tbegin(ctx, K_SPHERE_REDUCE, &ev);
  hipLaunchKernelGGL(k_sphere_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, partial, NB,
                     nbc, PM, nparts, sp->nsites, (const int *)sp->d_site_np.p, (const int *)sp->d_proj_off.p,
                     (const int *)sp->d_proj_m.p, ptot, out);
  tend(ctx, K_SPHERE_REDUCE, ev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

// Sphere-sum parts per site of a plan (the tuning value, else about eight
// four-wave work-groups per CU).
int sphere_parts_of(pp_ctx *ctx, pp_sphere *sp) {
  int nparts = ctx->tune_sphere_parts;
  if (nparts <= 0) {
    nparts = (2048 + sp->nsites - 1) / std::max(1, sp->nsites);
    long avg = sp->npts / std::max(1, sp->nsites);
    int maxp = (int)std::max<long>(1, avg / 256);
    if (nparts > maxp)
      nparts = maxp;
    if (nparts < 1)
      nparts = 1;
  }
  return nparts;
}
// Projector slots per site the sum kernels hold for a plan (8, 18, 32, 64)
// and whether every site has exactly that many.
int sphere_pm_of(pp_sphere *sp, bool *full) {
  const int PM = sp->pmax <= 8 ? 8 : sp->pmax <= 18 ? 18 : sp->pmax <= 32 ? 32 : 64;
  if (full)
    *full = sp->pmin == PM;
  return PM;
}
// The plan's part of a work item.
void sphere_item_fields(pp_ctx *ctx, pp_sphere *sp, ChunkItem &it) {
  it.site_pt = (const PP_G int *)sp->d_site_pt.p;
  it.site_np = (const PP_G int *)sp->d_site_np.p;
  it.proj_off = (const PP_G int *)sp->d_proj_off.p;
  it.proj_m = (const PP_G int *)sp->d_proj_m.p;
  it.pidx = (const PP_G int *)sp->d_pidx.p;
  it.tab_off = (const PP_G long *)sp->d_tab_off.p;
  it.table = (const PP_G void *)sp->d_table.p;
  it.nsites = sp->nsites;
  it.nparts = sphere_parts_of(ctx, sp);
}

// Sphere sums and their reduction for a batch of work items: one launch
// each.  PM and full as sphere_pm_of for every item's plan; maxblk bounds
// nsites * nparts and maxred nsites * nbc * PM over the items.
template <typename T>
int sphere_sums_items(pp_ctx *ctx, const ChunkItem *items, int nitems, int NB, int PM, bool full, long maxblk,
                      long maxred) {
  if (nitems == 0 || maxblk == 0)
    return PP_OK;
  hipEvent_t ev;
  tbegin(ctx, K_SPHERE_SUM, &ev);
  dim3 grd((unsigned)maxblk, (unsigned)nitems);
  const bool u8 = ctx->tune_sphere_unroll == 8;
#define PP_SI(P, F, U)                                                                                           \
  hipLaunchKernelGGL((k_sphere_sum_items<T, P, F, U>), grd, dim3(256), 0, ctx->stream, items, NB, ctx->tune_sphere_nt)
#define PP_SIP(P)                                                                                                \
  do {                                                                                                           \
    if (full) {                                                                                                  \
      if (u8)                                                                                                    \
        PP_SI(P, true, 8);                                                                                       \
      else                                                                                                       \
        PP_SI(P, true, 4);                                                                                       \
    } else if (u8)                                                                                               \
      PP_SI(P, false, 8);                                                                                        \
    else                                                                                                         \
      PP_SI(P, false, 4);                                                                                        \
  } while (0)
  switch (PM) {
  case 8:
    PP_SIP(8);
    break;
  case 18:
    PP_SIP(18);
    break;
  case 32:
    PP_SIP(32);
    break;
  default:
    PP_SIP(64);
  }
#undef PP_SIP
#undef PP_SI
  tend(ctx, K_SPHERE_SUM, ev);
  PP_CHECK_HIP(hipGetLastError());
  tbegin(ctx, K_SPHERE_REDUCE, &ev);
  hipLaunchKernelGGL(k_sphere_reduce_items, dim3((unsigned)((maxred + 255) / 256), (unsigned)nitems), dim3(256), 0,
                     ctx->stream, items, NB, PM);
  tend(ctx, K_SPHERE_REDUCE, ev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}
template int sphere_sums_items<double>(pp_ctx *, const ChunkItem *, int, int, int, bool, long, long);
template int sphere_sums_items<float>(pp_ctx *, const ChunkItem *, int, int, int, bool, long, long);

// phase[pt] = dv / sqrt(V) e^{i k.path} (sphere sums) or, with expand,
// e^{-i k.path} (the augmentation functions of projector.c:276-330)
template <typename T>
int make_phase(pp_ctx *ctx, pp_sphere *sp, const double *kfrac, cplx<T> *phase, bool expand) {
  if (sp->npts == 0)
    return PP_OK;
  // reciprocal lattice with 2 pi, rows b_i (utils.h:131)
  double rec[9], v = det3(sp->lattice);
  cross(sp->lattice + 3, sp->lattice + 6, rec + 0);
  cross(sp->lattice + 6, sp->lattice + 0, rec + 3);
  cross(sp->lattice + 0, sp->lattice + 3, rec + 6);
  for (int i = 0; i < 9; i++)
    rec[i] *= 2.0 * M_PI / v;
  double kc[3];
  for (int j = 0; j < 3; j++)
    kc[j] = kfrac[0] * rec[j] + kfrac[1] * rec[3 + j] + kfrac[2] * rec[6 + j];
  double dv = v / sp->fftg[0] / sp->fftg[1] / sp->fftg[2];
  double scl = dv / sqrt(v);
  if (expand) {
    scl = 1.0;
    for (int j = 0; j < 3; j++)
      kc[j] = -kc[j];
  }
  hipEvent_t ev;
  tbegin(ctx, K_PHASE, &ev);
  hipLaunchKernelGGL(k_phase<T>, dim3((unsigned)((sp->npts + 255) / 256)), dim3(256), 0, ctx->stream, sp->npts,
                     (const double *)sp->d_path.p, kc[0], kc[1], kc[2], scl, phase);
  tend(ctx, K_PHASE, ev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

template int sphere_sums<double>(pp_ctx *, pp_sphere *, const cplx<double> *, int, int, const cplx<double> *,
                                 cplx<double> *, long);
template int sphere_sums<float>(pp_ctx *, pp_sphere *, const cplx<float> *, int, int, const cplx<float> *,
                                cplx<double> *, long);
template int make_phase<double>(pp_ctx *, pp_sphere *, const double *, cplx<double> *, bool);
template int make_phase<float>(pp_ctx *, pp_sphere *, const double *, cplx<float> *, bool);

// Accumulate conj of the augmentation functions of the plan's sites for
// a band chunk into grid (zeroed by the caller); P rows [nbc][ptotP] with
// plan site s at src_off_dev[s]; phase from make_phase(expand = true).
template <typename T>
int sphere_expand(pp_ctx *ctx, pp_sphere *sp, cplx<T> *grid, int NB, int nbc, const cplx<T> *phase,
                  const cplx<double> *P, long ptotP, const int *src_off_dev) {
  if (sp->nsites == 0 || nbc == 0)
    return PP_OK;
  const int PM = sp->pmax <= 8 ? 8 : sp->pmax <= 18 ? 18 : sp->pmax <= 32 ? 32 : 64;
  DevBuf &qb = ctx->slot ? ctx->partial2 : ctx->partial;
  const long nq = (long)sp->nsites * PM * NB;
  PP_TRY(qb.ensure(sizeof(cplx<T>) * (size_t)nq));
  cplx<T> *Q = (cplx<T> *)qb.p;
  hipEvent_t ev;
  tbegin(ctx, K_AUG_PREP, &ev);
  hipLaunchKernelGGL(k_expand_coef<T>, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, ctx->stream, sp->nsites, PM,
                     NB, nbc, (const int *)sp->d_site_np.p, (const int *)sp->d_proj_off.p, (const int *)sp->d_proj_m.p,
                     src_off_dev, P, ptotP, Q);
  tend(ctx, K_AUG_PREP, ev);
  PP_CHECK_HIP(hipGetLastError());
  long avg = sp->npts / std::max(1, sp->nsites);
  int nparts = (int)std::max<long>(1, std::min<long>((2048 + sp->nsites - 1) / sp->nsites, avg / 64));
  tbegin(ctx, K_SPHERE_SUM, &ev);
#define PP_EX(PMX)                                                                                                  \
  hipLaunchKernelGGL((k_sphere_expand<T, PMX>), dim3((unsigned)(sp->nsites * nparts)), dim3(256), 0, ctx->stream,  \
                     grid, NB, nbc, (const int *)sp->d_site_pt.p, (const int *)sp->d_site_np.p,                   \
                     (const long *)sp->d_tab_off.p, (const int *)sp->d_pidx.p, phase, (const T *)sp->d_table.p, Q, \
                     nparts)
  if (PM == 8)
    PP_EX(8);
  else if (PM == 18)
    PP_EX(18);
  else if (PM == 32)
    PP_EX(32);
  else
    PP_EX(64);
#undef PP_EX
  tend(ctx, K_SPHERE_SUM, ev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

template int sphere_expand<double>(pp_ctx *, pp_sphere *, cplx<double> *, int, int, const cplx<double> *,
                                   const cplx<double> *, long, const int *);
template int sphere_expand<float>(pp_ctx *, pp_sphere *, cplx<float> *, int, int, const cplx<float> *,
                                  const cplx<double> *, long, const int *);
int sphere_nsites(pp_sphere *sp) { return sp->nsites; }
long sphere_generation(pp_sphere *sp) { return sp->gen; }
const unsigned *sphere_xmask(pp_sphere *sp, int *words) {
  *words = sp->xmask_words;
  return (const unsigned *)sp->d_xmask.p;
}

long sphere_npts(pp_sphere *sp) { return sp->npts; }
int sphere_ptot(pp_sphere *sp) { return sp->ptot; }
int sphere_precision(pp_sphere *sp) { return sp->precision; }
const int *sphere_fftg(pp_sphere *sp) { return sp->fftg; }
const double *sphere_lattice(pp_sphere *sp) { return sp->lattice; }

}  // namespace pp

using namespace pp;

extern "C" int pp_sphere_create(pp_ctx *ctx, pp_sphere **out, const double *lattice, const int *fftg,
                                int num_sites, const int *labels, const double *frac_coords, int num_elems,
                                const int *nchan, const int *ls, const double *rmax, const int *ngrid,
                                const double *proj, int precision) {
  PP_REQUIRE(ctx && out && lattice && fftg && num_sites >= 0 && num_elems > 0, PP_ERR_ARG,
             "pp_sphere_create: bad arguments");
  PP_REQUIRE(precision == PP_FP64 || precision == PP_FP32 || precision == PP_BF16X3 ||
                 precision == PP_BF16X6 || precision == PP_FP16X3,
             PP_ERR_ARG,
             "pp_sphere_create: unknown precision");
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  pp_sphere *sp = new pp_sphere();
  sp->ctx = ctx;
  sp->precision = precision == PP_FP64 ? PP_FP64 : PP_FP32;
  memcpy(sp->lattice, lattice, sizeof(sp->lattice));
  memcpy(sp->fftg, fftg, sizeof(sp->fftg));
  sp->nsites = num_sites;
  sp->nelems = num_elems;
  sp->labels.assign(labels, labels + num_sites);
  sp->frac.assign(frac_coords, frac_coords + 3 * num_sites);
  sp->nchan.assign(nchan, nchan + num_elems);
  sp->ngrid.assign(ngrid, ngrid + num_elems);
  sp->rmax.assign(rmax, rmax + num_elems);
  for (int s = 0; s < num_sites; s++)
    if (labels[s] < 0 || labels[s] >= num_elems) {
      delete sp;
      PP_REQUIRE(false, PP_ERR_ARG, "pp_sphere_create: site label out of range");
    }
  // species spline blocks: x, f[nchan], b[nchan], c[nchan], d[nchan]
  std::vector<double> spl;
  int lo = 0;
  long po = 0;
  for (int e = 0; e < num_elems; e++) {
    sp->ls_off.push_back(lo);
    int n = ngrid[e];
    if (n < 3) {
      delete sp;
      PP_REQUIRE(false, PP_ERR_ARG, "pp_sphere_create: projector grid needs >= 3 points");
    }
    std::vector<double> x(n);
    for (int j = 0; j < n; j++)
      x[j] = rmax[e] / n * j;  // projector.c:56
    size_t base = spl.size();
    spl.resize(base + (size_t)(1 + 4 * nchan[e]) * n);
    memcpy(&spl[base], x.data(), sizeof(double) * n);
    for (int ch = 0; ch < nchan[e]; ch++) {
      sp->ls.push_back(ls[lo + ch]);
      if (ls[lo + ch] < 0 || ls[lo + ch] > 3) {
        delete sp;
        PP_REQUIRE(false, PP_ERR_UNSUPPORTED, "pp_sphere_create: l must be 0..3");
      }
      const double *f = proj + po + (long)ch * n;
      double *F = &spl[base + (size_t)(1 + ch) * n];
      double *B = &spl[base + (size_t)(1 + nchan[e] + ch) * n];
      double *C = &spl[base + (size_t)(1 + 2 * nchan[e] + ch) * n];
      double *D = &spl[base + (size_t)(1 + 3 * nchan[e] + ch) * n];
      memcpy(F, f, sizeof(double) * n);
      spline_coeff_host(x.data(), f, n, B, C, D);
    }
    lo += nchan[e];
    po += (long)nchan[e] * n;
  }
  int st = sp->d_spline.ensure(sizeof(double) * spl.size());
  if (st == PP_OK)
    st = sp->d_chan.ensure(sizeof(int) * std::max<size_t>(1, sp->ls.size()));
  if (st != PP_OK) {
    delete sp;
    return st;
  }
  PP_CHECK_HIP(hipMemcpy(sp->d_spline.p, spl.data(), sizeof(double) * spl.size(), hipMemcpyHostToDevice));
  PP_CHECK_HIP(hipMemcpy(sp->d_chan.p, sp->ls.data(), sizeof(int) * sp->ls.size(), hipMemcpyHostToDevice));
  st = sphere_build(sp);
  if (st != PP_OK) {
    delete sp;
    return st;
  }
  *out = sp;
  return PP_OK;
}

extern "C" int pp_sphere_rebuild(pp_sphere *sp, const double *frac_coords) {
  PP_REQUIRE(sp && frac_coords, PP_ERR_ARG, "pp_sphere_rebuild: bad arguments");
  sp->frac.assign(frac_coords, frac_coords + 3 * sp->nsites);
  return sphere_build(sp);
}

extern "C" void pp_sphere_destroy(pp_sphere *sp) { delete sp; }

extern "C" int pp_sphere_coverage(pp_sphere *sp, long *covered) {
  PP_REQUIRE(sp && covered, PP_ERR_ARG, "pp_sphere_coverage: null argument");
  PP_CHECK_HIP(hipSetDevice(sp->ctx->device));
  const size_t nw = (size_t)sp->fftg[1] * sp->fftg[2] * sp->xmask_words;
  std::vector<unsigned> m(nw);
  PP_CHECK_HIP(hipStreamSynchronize(sp->ctx->stream));
  if (nw)
    PP_CHECK_HIP(hipMemcpy(m.data(), sp->d_xmask.p, sizeof(unsigned) * nw, hipMemcpyDeviceToHost));
  long c = 0;
  for (unsigned w : m)
    c += __builtin_popcount(w);
  *covered = c;
  return PP_OK;
}

extern "C" int pp_sphere_info(pp_sphere *sp, int *total_projs, long *total_points, int *proj_off) {
  PP_REQUIRE(sp, PP_ERR_ARG, "pp_sphere_info: null plan");
  if (total_projs)
    *total_projs = sp->ptot;
  if (total_points)
    *total_points = sp->npts;
  if (proj_off)
    memcpy(proj_off, sp->proj_off.data(), sizeof(int) * (sp->nsites + 1));
  return PP_OK;
}
