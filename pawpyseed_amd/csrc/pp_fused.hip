// The last FFT pass (along x) fused with the sphere sums, FP32 path.
//
// Unfused, the x pass writes the whole real-space chunk to HBM and the
// sphere sums read nearly all of it back (the projector spheres of a
// packed structure cover most grid points).  Here a work-group transforms
// the x-lines of a block of FX_BY x FX_BZ lines (y, z) one at a time in
// LDS and, while a line is there, adds its points to the sums
//   s_j = sum_i t_j(i) psi(idx_i) dv exp(i k.path_i)
// of every site whose sphere reaches the block (projector.c:223-274, in
// the real-harmonic form of pp_sphere.hip).  The points of a site on one
// line are one segment of consecutive x (the sphere is convex), so the
// sums read the transformed line from LDS and the projector values and
// phases of the segment from the site's tables.  The transformed grid is
// never written.
//
// Work item: a block and up to FX_G sites; wavefront w holds the sums of
// sites w, w + 4, w + 8 in registers, its 64 lanes being 16 bands x 2
// points of a segment x 2 halves of the projector list.  Each (item, site) writes one
// partial row of FX_PMAX x NB sums; the reduction adds the rows of each
// site in FP64 and forms the complex projections.
//
// Work lists (built on the host from the site boxes at plan time): a site
// joins every block holding a line whose distance from the site is below
// R0, a superset of the lines that carry its points; the segment table
// [item][line][slot] holds the one segment of each (site, line), filled on
// the device from the segments of the point lists.
#include <algorithm>
#include <cmath>

#include "pp_sphere.h"
#include "pp_stages.h"

namespace pp {

constexpr int FX_STAGE = 4;  // staging loads per lane and site in flight

__global__ __launch_bounds__(256) void k_fill_int(int *__restrict__ p, long n, int v) {
  long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    p[i] = v;
}

// Segment q goes to the slot of its site in the items of its block.
__global__ __launch_bounds__(256) void k_seg_table(const int4 *__restrict__ seg, const int *__restrict__ segline,
                                                   long nseg, int n2, int nbz, const int *__restrict__ bitem,
                                                   const int *__restrict__ isites, int *__restrict__ segtab,
                                                   int *__restrict__ flag) {
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  if (q >= nseg)
    return;
  const int4 sg = seg[q];
  const int yz = segline[q];
  const int y = yz / n2, z = yz - y * n2;
  const int b = (y / FX_BY) * nbz + z / FX_BZ;
  const int line = (y % FX_BY) * FX_BZ + z % FX_BZ;
  for (int it = bitem[b]; it < bitem[b + 1]; it++)
    for (int s = 0; s < FX_G; s++)
      if (isites[it * FX_G + s] == sg.w) {
        const int old = atomicCAS(&segtab[((long)it * FX_BL + line) * FX_G + s], -1, (int)q);
        if (old != -1)
          atomicOr(flag, 2);
        return;
      }
  atomicOr(flag, 1);
}

// Grid: ceil(nitems / 8) * 8 * nslices work-groups; the nslices band
// slices of one item are 8 apart in blockIdx, so they run on one XCD and
// share its L2 for the item's tables.  Lane l of a wavefront: band
// b = l % NBW, point parity pq and projector half from l / NBW, so each
// lane holds FX_S x FX_PMAX / 2 complex sums.
//
// Per line: the x-line is loaded into LDS and, in the same round trip,
// every wavefront copies the table rows and phases of its own sites'
// segments on the line into its LDS staging area (maxlen points per site:
// a segment never exceeds the site box width 2 g0 + 1); the stages run;
// then each wavefront adds its segments from LDS alone.
template <int NBW, int MAXB, int RMAX, bool FULL>
__global__ __launch_bounds__(256) void k_fft_x_sphere(
    const cplx<float> *__restrict__ g, AxisPlan pl, const cplx<float> *__restrict__ tw, int twlen, long es, long s1,
    long s2, int n1, int n2, int nslices, int nitems, int maxlen, SphereBounds bd, const int4 *__restrict__ items,
    const int *__restrict__ isites, const int *__restrict__ segtab, const int4 *__restrict__ seg,
    const int *__restrict__ site_pt, const int *__restrict__ site_np, const long *__restrict__ tab_off,
    const cplx<float> *__restrict__ phase, const float *__restrict__ table, int NB, cplx<float> *__restrict__ partial) {
  static_assert(NBW == 16, "lanes are 16 bands x 2 points x 2 projector halves");
  constexpr int P = FX_PMAX, PH = FX_PMAX / 2, S = FX_S, G = FX_G;
  extern __shared__ unsigned char smem_raw[];
  __shared__ int2 s_seg[FX_BL * G];  // per (line, slot): first point, length << 16 | x of the first point
  __shared__ long s_tbase[G];        // table row of point p of the slot's site: s_tbase + p * np
  __shared__ int s_np[G];
  const int bid = blockIdx.x, xcd = bid & 7, r8 = bid >> 3;
  const int slice = r8 % nslices, item = (r8 / nslices) * 8 + xcd;
  if (item >= nitems)
    return;
  const int4 it = items[item];  // y0, z0, sites, -
  const int n = pl.n, tot = n * NBW;
  cplx<float> *buf = reinterpret_cast<cplx<float> *>(smem_raw);
  cplx<float> *twl = buf + tot;
  cplx<float> *s_ph = twl + twlen;                                     // [G][maxlen]
  float *s_tab = reinterpret_cast<float *>(s_ph + (long)G * maxlen);  // [G][maxlen][np]
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, b = lane % NBW;
  const int pq = (lane / NBW) & 1, j0 = (lane / NBW >> 1) * PH;
  stage_twiddles<float, 256>(twl, tw, twlen);
  for (int q = tid; q < FX_BL * G; q += 256) {
    const int sid = segtab[(long)item * FX_BL * G + q];
    int2 v = make_int2(0, 0);
    if (sid >= 0) {
      const int4 sg = seg[sid];
      v = make_int2(sg.x, (sg.y << 16) | sg.z);
    }
    s_seg[q] = v;
  }
  if (tid < G) {
    const int site = isites[item * G + tid];
    const int np = site >= 0 ? (FULL ? P : site_np[site]) : 0;
    s_np[tid] = np;
    s_tbase[tid] = site >= 0 ? tab_off[site] - (long)site_pt[site] * np : 0;
  }
  int lo = -(1 << 30), hi = 1 << 30;
  if (bd.xb) {
    lo = bd.xb[0];
    hi = bd.xb[1];
  }
  cplx<float> acc[S][PH];
#pragma unroll
  for (int r = 0; r < S; r++)
#pragma unroll
    for (int j = 0; j < PH; j++)
      acc[r][j] = mk<float>(0, 0);
  __syncthreads();
  for (int ln = 0; ln < FX_BL; ln++) {
    const int y = it.x + ln / FX_BZ, z = it.y + ln % FX_BZ;
    if (y >= n1 || z >= n2)
      continue;
    // lines no site of the item reaches are not transformed at all (the
    // test reads the same LDS words in every thread)
    int any = 0;
    for (int q = 0; q < G; q++)
      any |= s_seg[ln * G + q].y;
    if (!any)
      continue;
    const long base = (long)y * s1 + (long)z * s2 + (long)slice * NBW;
    for (int t = tid; t < tot; t += 256) {
      const int e = t / NBW, bb = t % NBW;
      const int se = sgn_idx(e, n);
      buf[t] = (se >= lo && se <= hi) ? g[base + (long)e * es + bb] : mk<float>(0, 0);
    }
    // this wavefront's segments: table rows and phases into its staging,
    // all loads of a batch in flight before the first store
    {
      int nfl[S], p0[S], len[S];
      const float *src[S];
#pragma unroll
      for (int r = 0; r < S; r++) {
        const int s = wave + 4 * r;
        const int2 sv = s_seg[ln * G + s];
        len[r] = sv.y >> 16;
        p0[r] = sv.x;
        nfl[r] = len[r] * s_np[s];
        src[r] = table + s_tbase[s] + (long)sv.x * s_np[s];
      }
      const int most = max(nfl[0], max(nfl[1], nfl[2]));
      for (int e0 = 0; e0 < most; e0 += 64 * FX_STAGE) {
        float v[S][FX_STAGE];
#pragma unroll
        for (int r = 0; r < S; r++)
#pragma unroll
          for (int k = 0; k < FX_STAGE; k++) {
            const int e = e0 + lane + 64 * k;
            v[r][k] = e < nfl[r] ? src[r][e] : 0.0f;
          }
#pragma unroll
        for (int r = 0; r < S; r++)
#pragma unroll
          for (int k = 0; k < FX_STAGE; k++) {
            const int e = e0 + lane + 64 * k;
            if (e < nfl[r])
              s_tab[(long)(wave + 4 * r) * maxlen * P + e] = v[r][k];
          }
      }
#pragma unroll
      for (int r = 0; r < S; r++)
        for (int e = lane; e < len[r]; e += 64)
          s_ph[(wave + 4 * r) * maxlen + e] = phase[p0[r] + e];
    }
    __syncthreads();
    stages_ip<float, NBW, MAXB, 256, false, RMAX>(buf, twl, pl, n);
#pragma unroll
    for (int r = 0; r < S; r++) {
      const int s = wave + 4 * r;
      const int2 sv = s_seg[ln * G + s];
      const int len = sv.y >> 16;
      if (len == 0)
        continue;
      const int np = FULL ? P : s_np[s];
      const float *tp = s_tab + (long)s * maxlen * P + j0;
      const cplx<float> *ph = s_ph + s * maxlen;
      int x = (sv.y & 0xffff) + pq;
      if (x >= n)
        x -= n;
      for (int i = pq; i < len; i += 2) {
        const cplx<float> v = buf[x * NBW + b] * ph[i];
        const float *t = tp + i * np;
#pragma unroll
        for (int j = 0; j < PH; j++)
          if (FULL || j0 + j < np) {
            const float tv = t[j];
            acc[r][j].x += tv * v.x;
            acc[r][j].y += tv * v.y;
          }
        x += 2;
        while (x >= n)
          x -= n;
      }
    }
    __syncthreads();  // the next line overwrites buf and the staging
  }
  // add the two point parities of every (band, projector)
#pragma unroll
  for (int r = 0; r < S; r++)
#pragma unroll
    for (int j = 0; j < PH; j++) {
      acc[r][j].x += __shfl_xor(acc[r][j].x, NBW);
      acc[r][j].y += __shfl_xor(acc[r][j].y, NBW);
    }
  if (pq == 0) {
#pragma unroll
    for (int r = 0; r < S; r++) {
      const int s = wave + 4 * r;
      if (isites[item * G + s] < 0)
        continue;
      cplx<float> *o = partial + ((long)(item * G + s) * P + j0) * NB + slice * NBW + b;
#pragma unroll
      for (int j = 0; j < PH; j++)
        o[(long)j * NB] = acc[r][j];
    }
  }
}

// Add the partial rows of every (site, projector, band) in FP64 and form
// the complex projections from the real-harmonic sums, as k_sphere_reduce.
__global__ __launch_bounds__(256) void k_fused_reduce(const cplx<float> *__restrict__ partial, int NB, int nbc,
                                                      int nsites, const int *__restrict__ site_np,
                                                      const int *__restrict__ proj_off, const int *__restrict__ proj_m,
                                                      const int *__restrict__ rows_off, const int *__restrict__ rows,
                                                      long ptot, cplx<double> *__restrict__ out) {
  constexpr int P = FX_PMAX;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)nsites * P * nbc)
    return;
  const int b = (int)(t % nbc);
  const long r = t / nbc;
  const int j = (int)(r % P), s = (int)(r / P);
  if (j >= site_np[s])
    return;
  const int m = proj_m[proj_off[s] + j];
  if (m < 0)
    return;
  cplx<double> c = mk<double>(0, 0), sn = mk<double>(0, 0);
  for (int q = rows_off[s]; q < rows_off[s + 1]; q++) {
    const cplx<float> *pr = partial + (long)rows[q] * P * NB + b;
    c = c + cvt<double>(pr[(long)j * NB]);
    if (m > 0)
      sn = sn + cvt<double>(pr[(long)(j - 2 * m) * NB]);
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

// ------------------------------------------------------------------ host
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

bool fused_lists(const double *lattice, const int *fftg, const std::vector<SiteGeo> &geo, FusedLists &fl) {
  const int ns = (int)geo.size();
  const int N1 = fftg[1], N2 = fftg[2];
  for (const SiteGeo &g : geo)
    if (2 * g.g[1] + 1 > N1 || 2 * g.g[2] + 1 > N2)
      return false;
  const int nby = (N1 + FX_BY - 1) / FX_BY, nbz = (N2 + FX_BZ - 1) / FX_BZ, nblk = nby * nbz;
  const double *a0 = lattice, *a1 = lattice + 3, *a2 = lattice + 6;
  const double a00 = dot3(a0, a0);
  // sites of every block with the number of their lines in it
  std::vector<std::vector<std::pair<int, int>>> blist(nblk);
  std::vector<int> mark(nblk, -1), widx(nblk, 0);
  for (int s = 0; s < ns; s++) {
    const SiteGeo &g = geo[s];
    const double R = g.R0 * (1.0 + 1e-9) + 1e-12;
    for (int dj = -g.g[1]; dj <= g.g[1]; dj++) {
      const int j = g.c[1] + dj;
      const double fj = (double)j / N1 - g.coord[1];
      for (int dk = -g.g[2]; dk <= g.g[2]; dk++) {
        const int k = g.c[2] + dk;
        const double fk = (double)k / N2 - g.coord[2];
        double w[3];
        for (int c = 0; c < 3; c++)
          w[c] = fj * a1[c] + fk * a2[c];
        const double wa = dot3(w, a0);
        const double d2 = dot3(w, w) - wa * wa / a00;  // squared distance of the site from the line
        if (d2 >= R * R)
          continue;
        const int jw = (j % N1 + N1) % N1, kw = (k % N2 + N2) % N2;
        const int b = (jw / FX_BY) * nbz + kw / FX_BZ;
        if (mark[b] != s) {
          mark[b] = s;
          widx[b] = (int)blist[b].size();
          blist[b].push_back({s, 0});
        }
        blist[b][widx[b]].second++;
      }
    }
  }
  // items of up to FX_G sites per block; sites in decreasing order of
  // their line counts dealt to the four wavefronts in a snake (0 1 2 3 3 2
  // 1 0 ...) so the wavefronts carry similar work
  fl.items.clear();
  fl.isites.clear();
  fl.bitem.assign(nblk + 1, 0);
  fl.rows_off.assign(ns + 1, 0);
  int nitems = 0;
  for (int b = 0; b < nblk; b++) {
    fl.bitem[b] = nitems;
    auto &L = blist[b];
    if (L.empty())
      continue;
    std::stable_sort(L.begin(), L.end(), [](const std::pair<int, int> &u, const std::pair<int, int> &v) {
      return u.second > v.second;
    });
    for (size_t c0 = 0; c0 < L.size(); c0 += FX_G) {
      const int nsl = (int)std::min<size_t>(FX_G, L.size() - c0);
      fl.items.push_back((b / nbz) * FX_BY);
      fl.items.push_back((b % nbz) * FX_BZ);
      fl.items.push_back(nsl);
      fl.items.push_back(0);
      const size_t o = fl.isites.size();
      fl.isites.resize(o + FX_G, -1);
      for (int u = 0; u < nsl; u++) {
        const int r = u / 4, pos = u % 4;
        const int wave = (r & 1) ? 3 - pos : pos;
        fl.isites[o + wave + 4 * r] = L[c0 + u].first;
        fl.rows_off[L[c0 + u].first + 1]++;
      }
      nitems++;
    }
  }
  fl.bitem[nblk] = nitems;
  // rows of every site (row = item * FX_G + slot), in item order
  for (int s = 0; s < ns; s++)
    fl.rows_off[s + 1] += fl.rows_off[s];
  fl.rows.assign(std::max(1, fl.rows_off[ns]), 0);
  std::vector<int> fill(fl.rows_off.begin(), fl.rows_off.end() - 1);
  for (int it = 0; it < nitems; it++)
    for (int s = 0; s < FX_G; s++) {
      const int site = fl.isites[(size_t)it * FX_G + s];
      if (site >= 0)
        fl.rows[fill[site]++] = it * FX_G + s;
    }
  fl.nitems = nitems;
  fl.nblk = nblk;
  fl.nbz = nbz;
  return true;
}

int fused_plan(pp_sphere *sp, const std::vector<SiteGeo> &geo) {
  pp_ctx *ctx = sp->ctx;
  const int ns = sp->nsites;
  sp->fused_ok = false;
  sp->fx_built = true;
  sp->fx_nitems = sp->fx_nrows = sp->fx_nrows_used = 0;
  if (sp->precision != PP_FP32 || sp->pmax > FX_PMAX || ns == 0 || sp->nseg == 0)
    return PP_OK;
  FusedLists fl;
  if (!fused_lists(sp->lattice, sp->fftg, geo, fl))
    return PP_OK;
  // one packed upload: items | item sites | block -> items | site -> rows
  sp->fx_items_off = 0;
  sp->fx_sites_off = (long)fl.items.size();
  sp->fx_bitem_off = sp->fx_sites_off + (long)fl.isites.size();
  sp->fx_rows_off = sp->fx_bitem_off + (long)fl.bitem.size();
  sp->fx_rows = sp->fx_rows_off + (long)fl.rows_off.size();
  sp->fx_host.clear();
  for (const std::vector<int> *v : {&fl.items, &fl.isites, &fl.bitem, &fl.rows_off, &fl.rows})
    sp->fx_host.insert(sp->fx_host.end(), v->begin(), v->end());
  PP_TRY(sp->d_fx.ensure(sizeof(int) * sp->fx_host.size()));
  PP_CHECK_HIP(hipMemcpyAsync(sp->d_fx.p, sp->fx_host.data(), sizeof(int) * sp->fx_host.size(),
                              hipMemcpyHostToDevice, ctx->stream));
  const long ntab = (long)std::max(1, fl.nitems) * FX_BL * FX_G;
  PP_TRY(sp->d_segtab.ensure(sizeof(int) * ntab));
  PP_TRY(sp->d_fxflag.ensure(sizeof(int)));
  int *dfx = (int *)sp->d_fx.p;
  hipEvent_t ev;
  tbegin(ctx, K_TABLE, &ev);
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((ntab + 255) / 256)), dim3(256), 0, ctx->stream,
                     (int *)sp->d_segtab.p, ntab, -1);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, ctx->stream, (int *)sp->d_fxflag.p, 1L, 0);
  hipLaunchKernelGGL(k_seg_table, dim3((unsigned)((sp->nseg + 255) / 256)), dim3(256), 0, ctx->stream,
                     (const int4 *)sp->d_seg.p, (const int *)sp->d_segline.p, sp->nseg, sp->fftg[2], fl.nbz,
                     dfx + sp->fx_bitem_off, dfx + sp->fx_sites_off, (int *)sp->d_segtab.p, (int *)sp->d_fxflag.p);
  tend(ctx, K_TABLE, ev);
  PP_CHECK_HIP(hipGetLastError());
  sp->fx_nitems = fl.nitems;
  sp->fx_nrows = fl.nitems * FX_G;
  sp->fx_nrows_used = fl.rows_off[ns];
  sp->fx_nblocks = fl.nblk;
  sp->fx_nbz = fl.nbz;
  // staging length per site: a segment never exceeds its box width
  sp->fx_maxlen = 1;
  for (const SiteGeo &g : geo)
    sp->fx_maxlen = std::max(sp->fx_maxlen, 2 * g.g[0] + 1);
  sp->fused_ok = true;
  return PP_OK;
}

// In-place stage plan of the x axis for 16-band tiles: 0 when the fused
// kernel cannot run it, else the butterflies per thread (2, 3, 4 or 8);
// rmax = the largest radix.
static int fused_maxb(const AxisPlan &ap, int &rmax) {
  int maxb = 0;
  rmax = 2;
  for (int s = 0; s < ap.nst; s++) {
    if (ap.radix[s] < 2 || ap.radix[s] > 5)
      return 0;
    rmax = std::max(rmax, ap.radix[s]);
    maxb = std::max(maxb, ((ap.n / ap.radix[s]) * 16 + 255) / 256);
  }
  if (maxb <= 2)
    return 2;
  if (maxb <= 3)
    return 3;
  if (maxb <= 4)
    return 4;
  return maxb <= 8 ? 8 : 0;
}

// Twiddles of the radix 2-5 stages (the composite factorization's follow
// them in the axis table and are not needed here).
static int basic_twlen(const AxisPlan &ap) { return ap.nstc > 0 ? ap.twoffc[0] : ap.twlen; }

// dynamic LDS of the fused kernel: line tile, twiddles, phase and table
// staging of FX_G sites x fx_maxlen points
static size_t fused_lds(const AxisPlan &ap, const pp_sphere *sp) {
  return ((size_t)ap.n * 16 + basic_twlen(ap) + (size_t)FX_G * sp->fx_maxlen) * sizeof(cplx<float>) +
         (size_t)FX_G * sp->fx_maxlen * FX_PMAX * sizeof(float);
}

// Work lists of the current tables, built on first use.
static int fused_ensure(pp_sphere *sp) { return sp->fx_built ? PP_OK : fused_plan(sp, sp->geo_h); }

bool fused_usable(pp_ctx *ctx, FftPlan *pl, pp_sphere *sp, int NB, bool fp32) {
  if (!ctx->tune_fused || !fp32 || NB % 16 != 0 || ctx->tune_fft_pingpong)
    return false;
  if (fused_ensure(sp) != PP_OK || !sp->fused_ok)
    return false;
  const AxisPlan &ap = pl->ax[0];
  int rmax = 0;
  return fused_maxb(ap, rmax) != 0 && fused_lds(ap, sp) <= 64 * 1024 - 2048;
}

// x pass of a chunk whose z and y passes are done, fused with the sphere
// sums and their reduction into out[b * ptot + proj_off[s] + j].
int fused_x_sums(pp_ctx *ctx, FftPlan *pl, pp_sphere *sp, const cplx<float> *grid, int NB, int nbc,
                 const SphereBounds *bounds, const cplx<float> *phase, cplx<double> *out, long ptot) {
  if (sp->nsites == 0 || nbc == 0)
    return PP_OK;
  constexpr int NBW = 16;
  const int n0 = pl->fftg[0], n1 = pl->fftg[1], n2 = pl->fftg[2];
  SphereBounds bd = {nullptr, nullptr, nullptr, n0, n1, n2};
  if (bounds)
    bd = *bounds;
  const AxisPlan &ap = pl->ax[0];
  int rmax = 0;
  const int maxb = fused_maxb(ap, rmax);
  PP_REQUIRE(maxb != 0 && NB % NBW == 0, PP_ERR_UNSUPPORTED,
             "fused x pass: unsupported x length or band chunk");
  const int nslices = NB / NBW;
  DevBuf &pbuf = ctx->slot ? ctx->partial2 : ctx->partial;
  PP_TRY(pbuf.ensure(sizeof(cplx<float>) * (size_t)std::max(1, sp->fx_nrows) * FX_PMAX * NB));
  cplx<float> *partial = (cplx<float> *)pbuf.p;
  const int *dfx = (const int *)sp->d_fx.p;
  const size_t lds = fused_lds(ap, sp);
  PP_REQUIRE(lds <= 64 * 1024 - 2048, PP_ERR_UNSUPPORTED, "fused x pass: staging does not fit in LDS");
  const int twlen = basic_twlen(ap);
  const long nwg = (long)((sp->fx_nitems + 7) / 8) * 8 * nslices;
  const long es = (long)n1 * n2 * NB, s1 = (long)n2 * NB, s2 = NB;
  const cplx<float> *tw = fft_twiddles<float>(pl, 0);
  const bool full = sp->pmin == FX_PMAX;
  hipEvent_t ev;
  tbegin(ctx, K_FFT_X_SPHERE, &ev);
#define PP_FX(MB, RM, FULL)                                                                                      \
  hipLaunchKernelGGL((k_fft_x_sphere<NBW, MB, RM, FULL>), dim3((unsigned)nwg), dim3(256), lds, ctx->stream, grid,  \
                     ap, tw, twlen, es, s1, s2, n1, n2, nslices, sp->fx_nitems, sp->fx_maxlen, bd,                  \
                     (const int4 *)(dfx + sp->fx_items_off),                                                        \
                     dfx + sp->fx_sites_off, (const int *)sp->d_segtab.p, (const int4 *)sp->d_seg.p,                \
                     (const int *)sp->d_site_pt.p, (const int *)sp->d_site_np.p, (const long *)sp->d_tab_off.p,     \
                     phase, (const float *)sp->d_table.p, NB, partial)
#define PP_FX_M(FULL)                                                                                            \
  do {                                                                                                           \
    if (rmax <= 4 && maxb <= 3) {                                                                                \
      if (maxb == 2)                                                                                             \
        PP_FX(2, 4, FULL);                                                                                       \
      else                                                                                                       \
        PP_FX(3, 4, FULL);                                                                                       \
    } else if (maxb <= 2) {                                                                                      \
      PP_FX(2, 5, FULL);                                                                                         \
    } else if (maxb <= 4) {                                                                                      \
      PP_FX(4, 5, FULL);                                                                                         \
    } else {                                                                                                     \
      PP_FX(8, 5, FULL);                                                                                         \
    }                                                                                                            \
  } while (0)
  if (full)
    PP_FX_M(true);
  else
    PP_FX_M(false);
#undef PP_FX_M
#undef PP_FX
  tend(ctx, K_FFT_X_SPHERE, ev);
  PP_CHECK_HIP(hipGetLastError());
  const long total = (long)sp->nsites * FX_PMAX * nbc;
  tbegin(ctx, K_SPHERE_REDUCE, &ev);
  hipLaunchKernelGGL(k_fused_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, partial, NB,
                     nbc, sp->nsites, (const int *)sp->d_site_np.p, (const int *)sp->d_proj_off.p,
                     (const int *)sp->d_proj_m.p, dfx + sp->fx_rows_off, dfx + sp->fx_rows, ptot, out);
  tend(ctx, K_SPHERE_REDUCE, ev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

}  // namespace pp

// Work lists of the fused x pass of a plan (diagnostic): whether the plan
// has them, their item, partial-row and segment counts, and the
// consistency flag of the segment table (0 = every segment found its slot
// exactly once).
extern "C" int pp_sphere_fused_info(pp_sphere *sp, int *usable, int *nitems, int *nrows, long *nseg, int *flag) {
  PP_REQUIRE(sp, PP_ERR_ARG, "pp_sphere_fused_info: null plan");
  PP_CHECK_HIP(hipSetDevice(sp->ctx->device));
  PP_TRY(pp::fused_ensure(sp));
  if (usable)
    *usable = sp->fused_ok ? 1 : 0;
  if (nitems)
    *nitems = sp->fx_nitems;
  if (nrows)
    *nrows = sp->fx_nrows_used;
  if (nseg)
    *nseg = sp->nseg;
  if (flag) {
    *flag = 0;
    if (sp->fused_ok) {
      PP_CHECK_HIP(hipStreamSynchronize(sp->ctx->stream));
      PP_CHECK_HIP(hipMemcpy(flag, sp->d_fxflag.p, sizeof(int), hipMemcpyDeviceToHost));
    }
  }
  return PP_OK;
}

// Work lists of the fused x pass computed on the host alone (no device
// needed), for checking them against independent point lists: sites with
// fractional coordinates frac[ns][3], sphere grids rmax[s] / ngrid[s].
// Sizes first (any output pointer may be NULL): counts[0] = items,
// counts[1] = blocks, counts[2] = total rows; then items[4 * items],
// isites[FX_G * items], bitem[blocks + 1], rows_off[ns + 1],
// rows[counts[2]].  Returns PP_ERR_UNSUPPORTED when a site box wraps onto
// itself in y or z.
extern "C" int pp_fused_lists_host(const double *lattice, const int *fftg, int ns, const double *frac,
                                   const double *rmax, const int *ngrid, int *counts, int *items, int *isites,
                                   int *bitem, int *rows_off, int *rows) {
  PP_REQUIRE(lattice && fftg && frac && rmax && ngrid && counts && ns >= 0, PP_ERR_ARG,
             "pp_fused_lists_host: bad arguments");
  std::vector<pp::SiteGeo> geo(ns);
  for (int s = 0; s < ns; s++) {
    for (int a = 0; a < 3; a++)
      geo[s].coord[a] = frac[3 * s + a];
    pp::site_box(lattice, fftg, rmax[s], ngrid[s], geo[s]);
  }
  pp::FusedLists fl;
  PP_REQUIRE(pp::fused_lists(lattice, fftg, geo, fl), PP_ERR_UNSUPPORTED,
             "pp_fused_lists_host: a site box wraps onto itself");
  counts[0] = fl.nitems;
  counts[1] = fl.nblk;
  counts[2] = fl.rows_off[ns];
  auto put = [](int *dst, const std::vector<int> &v, size_t n) {
    if (dst)
      std::copy(v.begin(), v.begin() + n, dst);
  };
  put(items, fl.items, fl.items.size());
  put(isites, fl.isites, fl.isites.size());
  put(bitem, fl.bitem, fl.bitem.size());
  put(rows_off, fl.rows_off, fl.rows_off.size());
  put(rows, fl.rows, (size_t)fl.rows_off[ns]);
  return PP_OK;
}
