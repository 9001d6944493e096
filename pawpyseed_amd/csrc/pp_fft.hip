// Plane-wave scatter and batched 3D backward FFT for gfx950.
//
// Replaces the per-band work of linalg.c:14-45 (fft3d): zero the grid,
// place C_G at ((G+N) mod N) with z fastest, backward (e^{+i}) transform.
// The 1/sqrt(V) scale of the reference is folded into the per-point phase
// of the sphere sums, so these kernels compute the unscaled transform.
//
// Layout in HBM: a chunk of NB bands is stored band-fastest,
//   grid[((x*N1 + y)*N2 + z)*NB + b],
// so every FFT pass moves runs of NB contiguous complex values per grid
// point and the sphere-sum gather reads one contiguous run per point.
// Each 1D pass loads one line x NBW bands into LDS, runs a mixed-radix
// Stockham transform (radices 4,2,3,5 and any other small prime) there,
// and writes the line back in place.
#include <cmath>

#include "pp_internal.h"
#include "pp_stages.h"

namespace pp {

// ------------------------------------------------------------------ plan
static int factor_axis(int n, AxisPlan &a) {
  a.n = n;
  a.nst = 0;
  int m = n;
  while (m % 4 == 0 && a.nst < 12) {
    a.radix[a.nst++] = 4;
    m /= 4;
  }
  for (int q = 2; m > 1 && q <= m; q++)
    while (m % q == 0) {
      if (a.nst >= 12 || q > 13)
        return PP_ERR_UNSUPPORTED;
      a.radix[a.nst++] = q;
      m /= q;
    }
  return PP_OK;
}

// Composite factorization for the in-place passes: the largest radix with
// a register-level DFT first (16, 12, 9, 8, 6, then 5, 4, 3, 2); false when
// a prime factor above 5 remains.
static bool factor_composite(int n, int *radix, int &nst) {
  static const int pref[] = {16, 12, 9, 8, 6, 5, 4, 3, 2};
  nst = 0;
  int m = n;
  while (m > 1) {
    int r = 0;
    for (int c : pref)
      if (m % c == 0) {
        r = c;
        break;
      }
    if (r == 0 || nst >= 12)
      return false;
    radix[nst++] = r;
    m /= r;
  }
  return true;
}

// twiddles exp(+2 pi i k r / (p R)), k < p, r = 1..R-1, then the roots
// exp(+2 pi i m / R) of the generic small DFT, for one stage
static void push_stage_twiddles(std::vector<double> &tw, int p, int R) {
  for (int k = 0; k < p; k++)
    for (int r = 1; r < R; r++) {
      long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)(k * r) / (long double)(p * R);
      tw.push_back((double)cosl(ang));
      tw.push_back((double)sinl(ang));
    }
  for (int m = 0; m < R; m++) {
    long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)R;
    tw.push_back((double)cosl(ang));
    tw.push_back((double)sinl(ang));
  }
}

int get_fft_plan(pp_ctx *ctx, const int *fftg, FftPlan **out) {
  long long key = ((long long)fftg[0] << 40) | ((long long)fftg[1] << 20) | fftg[2];
  auto it = ctx->fft_plans.find(key);
  if (it != ctx->fft_plans.end()) {
    *out = it->second;
    return PP_OK;
  }
  FftPlan *pl = new FftPlan();
  std::vector<double> tw;  // interleaved c128
  for (int a = 0; a < 3; a++) {
    pl->fftg[a] = fftg[a];
    AxisPlan &ap = pl->ax[a];
    if (factor_axis(fftg[a], ap) != PP_OK) {
      delete pl;
      PP_REQUIRE(false, PP_ERR_UNSUPPORTED,
                 "FFT length " + std::to_string(fftg[a]) + " has a prime factor above 13");
    }
    pl->twbase[a] = (int)(tw.size() / 2);
    int p = 1;
    for (int s = 0; s < ap.nst; s++) {
      int R = ap.radix[s];
      ap.pprev[s] = p;
      ap.pmag[s] = p > 1 ? (unsigned)(((1ull << 32) + (unsigned long long)p - 1) / (unsigned long long)p) : 0u;
      ap.twoff[s] = (int)(tw.size() / 2) - pl->twbase[a];
      push_stage_twiddles(tw, p, R);
      p *= R;
    }
    // the composite factorization, when it has fewer stages
    int rc[12], nc = 0;
    ap.nstc = 0;
    if (factor_composite(fftg[a], rc, nc) && nc < ap.nst) {
      ap.nstc = nc;
      p = 1;
      for (int s = 0; s < nc; s++) {
        ap.radixc[s] = rc[s];
        ap.pprevc[s] = p;
        ap.pmagc[s] = p > 1 ? (unsigned)(((1ull << 32) + (unsigned long long)p - 1) / (unsigned long long)p) : 0u;
        ap.twoffc[s] = (int)(tw.size() / 2) - pl->twbase[a];
        push_stage_twiddles(tw, p, rc[s]);
        p *= rc[s];
      }
    }
    ap.twlen = (int)(tw.size() / 2) - pl->twbase[a];
  }
  size_t ntw = tw.size() / 2;
  std::vector<float> tw32(tw.size());
  for (size_t i = 0; i < tw.size(); i++)
    tw32[i] = (float)tw[i];
  if (pl->tw64.ensure(ntw * 16) != PP_OK || pl->tw32.ensure(ntw * 8) != PP_OK) {
    delete pl;
    return PP_ERR_NOMEM;
  }
  PP_CHECK_HIP(hipMemcpy(pl->tw64.p, tw.data(), ntw * 16, hipMemcpyHostToDevice));
  PP_CHECK_HIP(hipMemcpy(pl->tw32.p, tw32.data(), ntw * 8, hipMemcpyHostToDevice));
  ctx->fft_plans[key] = pl;
  *out = pl;
  return PP_OK;
}

// ------------------------------------------------------------------ kernels
__global__ __launch_bounds__(256) void k_zero(uint4 *__restrict__ p, size_t n16) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t st = (size_t)gridDim.x * blockDim.x;
  uint4 z = {0u, 0u, 0u, 0u};
  for (; i < n16; i += st)
    p[i] = z;
}

__global__ __launch_bounds__(256) void k_zero_bytes(unsigned char *__restrict__ p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n)
    p[i] = 0;
}

__global__ __launch_bounds__(256) void k_gidx(int n0, int n1, int n2, int nw, const int *__restrict__ G,
                                              int *__restrict__ gidx) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nw)
    return;
  int a = ((G[3 * w] % n0) + n0) % n0;
  int b = ((G[3 * w + 1] % n1) + n1) % n1;
  int c = ((G[3 * w + 2] % n2) + n2) % n2;
  gidx[w] = (a * n1 + b) * n2 + c;
}

// Scatter a 32-coefficient x NB-band tile through LDS: coalesced reads
// along G of each band row, contiguous NB-band writes per grid point.  A
// thread's 8 coefficient loads, and then its 8 grid-index loads, are issued
// together before any is used (a loop that uses each load before issuing
// the next waits for them one after another).
template <typename T>
__device__ __forceinline__ void scatter_tile(cplx<T> *__restrict__ g, int NB, const int *__restrict__ gidx, int nw,
                                             const cplx<float> *__restrict__ C, long ldc, int nbc, int g0) {
  __shared__ cplx<T> tile[64][33];
  constexpr int EPT = 64 * 32 / 256;
  cplx<T> v[EPT];
#pragma unroll
  for (int u = 0; u < EPT; u++) {
    const int e = threadIdx.x + u * 256, b = e >> 5, gg = e & 31;
    v[u] = mk<T>(0, 0);
    if (b < nbc && g0 + gg < nw)
      v[u] = cvt<T>(C[(long)b * ldc + g0 + gg]);
  }
#pragma unroll
  for (int u = 0; u < EPT; u++) {
    const int e = threadIdx.x + u * 256;
    tile[e >> 5][e & 31] = v[u];
  }
  __syncthreads();
  const int b = threadIdx.x & 63, w = threadIdx.x >> 6;
  int gi[EPT];
#pragma unroll
  for (int u = 0; u < EPT; u++) {
    const int gg = w + 4 * u;
    gi[u] = g0 + gg < nw ? gidx[g0 + gg] : -1;
  }
#pragma unroll
  for (int u = 0; u < EPT; u++)
    if (b < nbc && b < NB && gi[u] >= 0)
      g[(long)gi[u] * NB + b] = tile[b][w + 4 * u];
}
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(cplx<T> *__restrict__ g, int NB, const int *__restrict__ gidx,
                                                 int nw, const cplx<float> *__restrict__ C, long ldc, int nbc) {
  scatter_tile<T>(g, NB, gidx, nw, C, ldc, nbc, blockIdx.x * 32);
}
// The scatter of a batch of work items (one per blockIdx.y).
template <typename T>
__global__ __launch_bounds__(256) void k_scatter_items(const ChunkItem *__restrict__ items, int NB) {
  const ChunkItem &it = items[blockIdx.y];
  const int g0 = blockIdx.x * 32;
  if (g0 >= it.nw)
    return;
  scatter_tile<T>((cplx<T> *)it.grid, NB, (const int *)it.gidx, it.nw, (const cplx<float> *)it.C, it.ldc, it.nbc,
                  g0);
}

// One Stockham stage of radix R over a line tile in LDS:
// v[r] = in[j + r*Tn] * w^(k r) (k = j mod p), small DFT, then
// out[(j/p)*p*R + k + r*p] = V[r].
template <typename T, int NBW, int R>
__device__ __forceinline__ void stage(const cplx<T> *__restrict__ in, cplx<T> *__restrict__ out,
                                      const cplx<T> *__restrict__ tws, int p, int Tn) {
  for (int t = threadIdx.x; t < Tn * NBW; t += 256) {
    const int j = t / NBW, b = t % NBW;
    const int k = j % p;
    cplx<T> v[R];
    v[0] = in[j * NBW + b];
#pragma unroll
    for (int r = 1; r < R; r++)
      v[r] = in[(j + r * Tn) * NBW + b] * tws[k * (R - 1) + (r - 1)];
    Dft<T, R>::run(v);
    const int o = (j / p) * p * R + k;
#pragma unroll
    for (int r = 0; r < R; r++)
      out[(o + r * p) * NBW + b] = v[r];
  }
}

// Any other small prime radix: each output computed directly from LDS
// (O(R^2) work, no local arrays).
template <typename T, int NBW>
__device__ void stage_generic(const cplx<T> *__restrict__ in, cplx<T> *__restrict__ out,
                              const cplx<T> *__restrict__ tws, int p, int Tn, int R) {
  const cplx<T> *roots = tws + p * (R - 1);
  for (int t = threadIdx.x; t < Tn * NBW; t += 256) {
    const int j = t / NBW, b = t % NBW;
    const int k = j % p;
    const int o = (j / p) * p * R + k;
    for (int q = 0; q < R; q++) {
      cplx<T> acc = in[j * NBW + b];
      for (int r = 1; r < R; r++) {
        cplx<T> x = in[(j + r * Tn) * NBW + b] * tws[k * (R - 1) + (r - 1)];
        acc = acc + x * roots[(r * q) % R];
      }
      out[(o + q * p) * NBW + b] = acc;
    }
  }
}

// Line l of a pass: from the populated-line list of the z / y pass when
// there is one, else l itself.
template <int AXIS>
__device__ __forceinline__ int pass_line(const SphereBounds &bd, int l) {
  const int *ll = AXIS == 2 ? bd.zl : AXIS == 1 ? bd.yl : nullptr;
  return ll ? ll[l] : l;
}

// One pass of 1D transforms along an axis.  Work-group: one line (fixed
// other two coordinates) x NBW bands.  Element e of band slot b lives at
// g[base + e*es + b].
//
// With bounds (SphereBounds, from the G vectors of the k-point) the pass
// only touches the part of the grid the plane-wave sphere can reach: the
// sphere is convex, so on every z-line its coefficients fill one signed
// interval [zlo, zhi]; the populated z-lines of a plane x lie in
// [ylo(x), yhi(x)] and the populated planes in [xlo, xhi].  Elements
// outside those intervals are zero by construction and are not loaded, and
// lines whose input is entirely outside are skipped (z and y passes), so no
// zero fill of the grid is needed.
template <typename T, int NBW, int AXIS>
__global__ __launch_bounds__(256) void k_fft_pass(cplx<T> *__restrict__ g, AxisPlan pl,
                                                  const cplx<T> *__restrict__ tw, long es, long s1, long s2,
                                                  int n2, int nslices, SphereBounds bd) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T> *buf0 = reinterpret_cast<cplx<T> *>(smem_raw);
  const int n = pl.n;
  cplx<T> *buf1 = buf0 + n * NBW;
  const int blk = blockIdx.x;
  const int slice = blk % nslices;
  const int line = pass_line<AXIS>(bd, blk / nslices);
  const int l1 = line / n2, l2 = line % n2;
  int lo = -(1 << 30), hi = 1 << 30;
  if (bd.xb) {
    if (AXIS == 2) {  // line (x, y)
      const int sx = sgn_idx(l1, bd.n0), sy = sgn_idx(l2, bd.n1);
      if (sx < bd.xb[0] || sx > bd.xb[1] || sy < bd.yb[2 * l1] || sy > bd.yb[2 * l1 + 1])
        return;
      lo = bd.zb[2 * (l1 * bd.n1 + l2)];
      hi = bd.zb[2 * (l1 * bd.n1 + l2) + 1];
    } else if (AXIS == 1) {  // line (x, z)
      const int sx = sgn_idx(l1, bd.n0);
      if (sx < bd.xb[0] || sx > bd.xb[1])
        return;
      lo = bd.yb[2 * l1];
      hi = bd.yb[2 * l1 + 1];
    } else {  // line (y, z)
      lo = bd.xb[0];
      hi = bd.xb[1];
    }
  }
  const long base = (long)l1 * s1 + (long)l2 * s2 + (long)slice * NBW;
  const int tot = n * NBW;
  for (int t = threadIdx.x; t < tot; t += 256) {
    int e = t / NBW, b = t % NBW;
    const int se = sgn_idx(e, n);
    buf0[t] = (se >= lo && se <= hi) ? g[base + (long)e * es + b] : mk<T>(0, 0);
  }
  __syncthreads();
  cplx<T> *in = buf0, *out = buf1;
  for (int s = 0; s < pl.nst; s++) {
    const int R = pl.radix[s], p = pl.pprev[s], Tn = n / R;
    const cplx<T> *tws = tw + pl.twoff[s];
    switch (R) {
    case 4:
      stage<T, NBW, 4>(in, out, tws, p, Tn);
      break;
    case 3:
      stage<T, NBW, 3>(in, out, tws, p, Tn);
      break;
    case 2:
      stage<T, NBW, 2>(in, out, tws, p, Tn);
      break;
    case 5:
      stage<T, NBW, 5>(in, out, tws, p, Tn);
      break;
    default:
      stage_generic<T, NBW>(in, out, tws, p, Tn, R);
    }
    __syncthreads();
    cplx<T> *tmp = in;
    in = out;
    out = tmp;
  }
  for (int t = threadIdx.x; t < tot; t += 256) {
    int e = t / NBW, b = t % NBW;
    g[base + (long)e * es + b] = in[t];
  }
}

// As k_fft_pass with in-place stages (radices with a register-level DFT),
// at most MAXB butterflies per thread per stage, NT threads per work-group.
//
// xcd_per > 0: work-group w runs tile (w % 8) * xcd_per + w / 8, so each
// of the 8 XCDs (which receive work-groups round-robin) walks one
// contiguous range of lines; tiles past nblk exit.
//
// streaming: 1 = the line is written back with non-temporal stores, 2 =
// also read with non-temporal loads (each pass touches its data once).
// The work of one work-group of the in-place pass: tile blk (line
// blk / nslices, band slice blk % nslices) of the chunk g.
// N > 0: the axis length and its stages are compile-time constants
// (stages_fixed); the host launches that form only when its plan has the
// same stages.
template <typename T, int NBW, int AXIS, int MAXB, int NT, bool COMP, int N = 0>
__device__ __forceinline__ void fft_pass_ip_tile(cplx<T> *__restrict__ g, const AxisPlan &pl,
                                                 const cplx<T> *__restrict__ tw, long es, long s1, long s2, int n2,
                                                 int nslices, const SphereBounds &bd, int blk, int streaming) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
  const int n = N > 0 ? N : pl.n;
  const int slice = blk % nslices;
  const int line = pass_line<AXIS>(bd, blk / nslices);
  const int l1 = line / n2, l2 = line % n2;
  int lo = -(1 << 30), hi = 1 << 30;
  if (bd.xb) {
    if (AXIS == 2) {
      const int sx = sgn_idx(l1, bd.n0), sy = sgn_idx(l2, bd.n1);
      if (sx < bd.xb[0] || sx > bd.xb[1] || sy < bd.yb[2 * l1] || sy > bd.yb[2 * l1 + 1])
        return;
      lo = bd.zb[2 * (l1 * bd.n1 + l2)];
      hi = bd.zb[2 * (l1 * bd.n1 + l2) + 1];
    } else if (AXIS == 1) {
      const int sx = sgn_idx(l1, bd.n0);
      if (sx < bd.xb[0] || sx > bd.xb[1])
        return;
      lo = bd.yb[2 * l1];
      hi = bd.yb[2 * l1 + 1];
    } else {
      lo = bd.xb[0];
      hi = bd.xb[1];
    }
  }
  const long base = (long)l1 * s1 + (long)l2 * s2 + (long)slice * NBW;
  const int tot = n * NBW;
  cplx<T> *twl = buf + tot;
  // the line's loads are issued in batches of FFT_LB per thread before any
  // is written to LDS, so that many are in flight at once (a load-then-store
  // loop waits for each load in turn); FP32 lines move two bands (16 bytes)
  // per access.  Compile-time lengths: a thread's elements are EST apart on
  // the axis, so their addresses step by a uniform stride from one base.
  constexpr int PK = (sizeof(T) == 4 && NBW % 2 == 0) ? 2 : 1;  // complex values per access
  constexpr int EST = PK * NT / NBW;
  constexpr bool FXLS = N > 0 && (PK * NT) % NBW == 0;
  constexpr int FX_MW = N > 0 ? (N + 31) / 32 : 1;  // coverage words of an x line
  unsigned mw[FX_MW];
  // compile-time lengths: the first 2 NT twiddles (all of them for the
  // compiled lengths) are read into registers ahead of the line's loads and
  // written to LDS after them, so that the two wait on one memory latency
  // rather than one after the other
  cplx<T> twv[2];
  if constexpr (FXLS) {
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const int i = threadIdx.x + r * NT;
      twv[r] = i < pl.twlen ? tw[i] : mk<T>(0, 0);
    }
    for (int i = threadIdx.x + 2 * NT; i < pl.twlen; i += NT)
      twl[i] = tw[i];
  } else
    stage_twiddles<T, NT>(twl, tw, pl.twlen);
  if constexpr (FXLS) {
    constexpr int TOTA = N * NBW / PK, EPT = (TOTA + NT - 1) / NT;
    const int t = threadIdx.x, e0 = PK * t / NBW, bb = PK * t % NBW;
    const cplx<T> *p = g + base + (long)e0 * es + bb;
    const long step = (long)EST * es;
    cplx<T> va[EPT][PK];
#pragma unroll
    for (int u = 0; u < EPT; u++) {
#pragma unroll
      for (int c = 0; c < PK; c++)
        va[u][c] = mk<T>(0, 0);
      const int e = e0 + u * EST;
      const int se = e < N - N / 2 ? e : e - N;
      if ((TOTA % NT == 0 || t + u * NT < TOTA) && se >= lo && se <= hi) {
        if constexpr (PK == 2)
          load2(reinterpret_cast<const cplx<float> *>(p + u * step), streaming > 1, va[u][0], va[u][1]);
        else
          va[u][0] = streaming > 1 ? load_nt(p + u * step) : p[u * step];
      }
    }
#pragma unroll
    for (int r = 0; r < 2; r++)
      if (threadIdx.x + r * NT < pl.twlen)
        twl[threadIdx.x + r * NT] = twv[r];
#pragma unroll
    for (int u = 0; u < EPT; u++)
      if (TOTA % NT == 0 || t + u * NT < TOTA) {
#pragma unroll
        for (int c = 0; c < PK; c++)
          buf[PK * (t + u * NT) + c] = va[u][c];
      }
    // the x line's coverage words, read here so that their loads complete
    // during the stages and the masked stores below test registers (a load
    // of the word beside each store waited for it store by store)
    if constexpr (AXIS == 0) {
      const unsigned *xm0 = bd.xm ? bd.xm + (long)line * bd.xmw : nullptr;
#pragma unroll
      for (int q = 0; q < FX_MW; q++)
        mw[q] = (xm0 && q < bd.xmw) ? xm0[q] : ~0u;
    }
  } else if constexpr (sizeof(T) == 4 && NBW % 2 == 0) {
    const int tot2 = tot / 2;
    for (int t0 = threadIdx.x; t0 < tot2; t0 += NT * FFT_LB) {
      cplx<T> va[FFT_LB], vb[FFT_LB];
#pragma unroll
      for (int u = 0; u < FFT_LB; u++) {
        const int t = t0 + u * NT;
        va[u] = vb[u] = mk<T>(0, 0);
        if (t < tot2) {
          const int e = 2 * t / NBW, b = 2 * t % NBW;
          const int se = sgn_idx(e, n);
          if (se >= lo && se <= hi)
            load2(reinterpret_cast<const cplx<float> *>(g + base + (long)e * es + b), streaming > 1, va[u], vb[u]);
        }
      }
#pragma unroll
      for (int u = 0; u < FFT_LB; u++)
        if (t0 + u * NT < tot2) {
          buf[2 * (t0 + u * NT)] = va[u];
          buf[2 * (t0 + u * NT) + 1] = vb[u];
        }
    }
  } else
  for (int t0 = threadIdx.x; t0 < tot; t0 += NT * FFT_LB) {
    cplx<T> v[FFT_LB];
#pragma unroll
    for (int u = 0; u < FFT_LB; u++) {
      const int t = t0 + u * NT;
      v[u] = mk<T>(0, 0);
      if (t < tot) {
        const int e = t / NBW, b = t % NBW;
        const int se = sgn_idx(e, n);
        if (se >= lo && se <= hi)
          v[u] = streaming > 1 ? load_nt(g + base + (long)e * es + b) : g[base + (long)e * es + b];
      }
    }
#pragma unroll
    for (int u = 0; u < FFT_LB; u++)
      if (t0 + u * NT < tot)
        buf[t0 + u * NT] = v[u];
  }
  __syncthreads();
  if constexpr (N > 0)
    stages_fixed<T, NBW, N, NT>(buf, twl);
  else
    stages_ip<T, NBW, MAXB, NT, COMP>(buf, twl, pl, n);
  const unsigned *xm = (AXIS == 0 && bd.xm) ? bd.xm + (long)line * bd.xmw : nullptr;
  if constexpr (FXLS) {
    constexpr int TOTA = N * NBW / PK, EPT = (TOTA + NT - 1) / NT;
    const int t = threadIdx.x, e0 = PK * t / NBW, bb = PK * t % NBW;
    cplx<T> *p = g + base + (long)e0 * es + bb;
    const long step = (long)EST * es;
    cplx<T> va[EPT][PK];
#pragma unroll
    for (int u = 0; u < EPT; u++)
      if (TOTA % NT == 0 || t + u * NT < TOTA) {
#pragma unroll
        for (int c = 0; c < PK; c++)
          va[u][c] = buf[PK * (t + u * NT) + c];
      }
#pragma unroll
    for (int u = 0; u < EPT; u++) {
      const int e = e0 + u * EST;
      if (!(TOTA % NT == 0 || t + u * NT < TOTA))
        continue;
      if constexpr (AXIS == 0) {
        unsigned w = mw[0];
#pragma unroll
        for (int q = 1; q < FX_MW; q++)
          w = (e >> 5) == q ? mw[q] : w;
        if (!((w >> (e & 31)) & 1u))
          continue;  // no projector sphere reads this point
      }
      if constexpr (PK == 2)
        store2(reinterpret_cast<cplx<float> *>(p + u * step), streaming != 0, va[u][0], va[u][1]);
      else if (streaming)
        store_nt(p + u * step, va[u][0]);
      else
        p[u * step] = va[u][0];
    }
    return;
  }
  if constexpr (sizeof(T) == 4 && NBW % 2 == 0) {
    const int tot2 = tot / 2;
    for (int t0 = threadIdx.x; t0 < tot2; t0 += NT * FFT_LB) {
      cplx<T> va[FFT_LB], vb[FFT_LB];
#pragma unroll
      for (int u = 0; u < FFT_LB; u++)
        if (t0 + u * NT < tot2) {
          va[u] = buf[2 * (t0 + u * NT)];
          vb[u] = buf[2 * (t0 + u * NT) + 1];
        }
#pragma unroll
      for (int u = 0; u < FFT_LB; u++) {
        const int t = t0 + u * NT;
        if (t >= tot2)
          break;
        const int e = 2 * t / NBW, b = 2 * t % NBW;
        if (xm && !((xm[e >> 5] >> (e & 31)) & 1u))
          continue;  // no projector sphere reads this point
        store2(reinterpret_cast<cplx<float> *>(g + base + (long)e * es + b), streaming != 0, va[u], vb[u]);
      }
    }
    return;
  }
  for (int t0 = threadIdx.x; t0 < tot; t0 += NT * FFT_LB) {
    cplx<T> v[FFT_LB];
#pragma unroll
    for (int u = 0; u < FFT_LB; u++)
      if (t0 + u * NT < tot)
        v[u] = buf[t0 + u * NT];
#pragma unroll
    for (int u = 0; u < FFT_LB; u++) {
      const int t = t0 + u * NT;
      if (t >= tot)
        break;
      const int e = t / NBW, b = t % NBW;
      if (xm && !((xm[e >> 5] >> (e & 31)) & 1u))
        continue;  // no projector sphere reads this point
      if (streaming)
        store_nt(g + base + (long)e * es + b, v[u]);
      else
        g[base + (long)e * es + b] = v[u];
    }
  }
}

template <typename T, int NBW, int AXIS, int MAXB, int NT, bool COMP = false>
__global__ __launch_bounds__(NT, (FftWaves<T, MAXB, COMP>::value)) void k_fft_pass_ip(cplx<T> *__restrict__ g, AxisPlan pl,
                                                     const cplx<T> *__restrict__ tw, long es, long s1, long s2,
                                                     int n2, int nslices, SphereBounds bd, int xcd_per, int nblk,
                                                     int streaming) {
  const int blk = xcd_per > 0 ? (int)(blockIdx.x % 8) * xcd_per + (int)(blockIdx.x / 8) : (int)blockIdx.x;
  if (blk >= nblk)
    return;
  fft_pass_ip_tile<T, NBW, AXIS, MAXB, NT, COMP>(g, pl, tw, es, s1, s2, n2, nslices, bd, blk, streaming);
}

// The in-place pass for an axis of compile-time length N (see
// fft_pass_ip_tile); FP32 kernels at 8 wavefronts per SIMD as
// k_fft_pass_ip's.
template <typename T, int NBW, int AXIS, int N, int NT>
__global__ __launch_bounds__(NT, (PassWaves<T, 4, N, NBW>::value)) void k_fft_pass_fx(
    cplx<T> *__restrict__ g, AxisPlan pl, const cplx<T> *__restrict__ tw, long es, long s1, long s2, int n2,
    int nslices, SphereBounds bd, int xcd_per, int nblk, int streaming) {
  const int blk = xcd_per > 0 ? (int)(blockIdx.x % 8) * xcd_per + (int)(blockIdx.x / 8) : (int)blockIdx.x;
  if (blk >= nblk)
    return;
  fft_pass_ip_tile<T, NBW, AXIS, 4, NT, false, N>(g, pl, tw, es, s1, s2, n2, nslices, bd, blk, streaming);
}

// The same pass over the chunks of a batch of work items (one per
// blockIdx.y): each item brings its own chunk buffer and the plane-wave
// bounds and line lists of its k-point; tiles past the item's lines exit.
// (7 rather than 8 waves per SIMD where the single-chunk pass takes 8: the
// descriptor's pointers and counts cost a few registers, and 64 would spill)
template <typename T, int NBW, int AXIS, int MAXB, int NT, int N = 0>
__global__ __launch_bounds__(NT, (PassWaves<T, MAXB, N, NBW>::value == 8 ? 7 : PassWaves<T, MAXB, N, NBW>::value))
void k_fft_pass_ip_items(
    const ChunkItem *__restrict__ items, AxisPlan pl, const cplx<T> *__restrict__ tw, long es, long s1, long s2,
    int n2, int nslices, int streaming) {
  const ChunkItem *it = items + blockIdx.y;
  SphereBounds bd;
  bd.zb = (const int *)it->zb;
  bd.yb = (const int *)it->yb;
  bd.xb = (const int *)it->xb;
  bd.n0 = it->n0;
  bd.n1 = it->n1;
  bd.n2 = it->n2;
  bd.zl = AXIS == 2 ? (const int *)it->zl : nullptr;
  bd.nzl = AXIS == 2 ? it->nzl : 0;
  bd.yl = AXIS == 1 ? (const int *)it->yl : nullptr;
  bd.nyl = AXIS == 1 ? it->nyl : 0;
  bd.xm = AXIS == 0 ? (const unsigned *)it->xm : nullptr;
  bd.xmw = AXIS == 0 ? it->xmw : 0;
  const long nlin = AXIS == 2 ? (bd.zl ? bd.nzl : (long)bd.n0 * bd.n1)
                              : AXIS == 1 ? (bd.yl ? bd.nyl : (long)bd.n0 * bd.n2) : (long)bd.n1 * bd.n2;
  const int blk = (int)blockIdx.x;
  if (blk >= nlin * nslices)
    return;
  fft_pass_ip_tile<T, NBW, AXIS, MAXB, NT, false, N>((cplx<T> *)it->grid, pl, tw, es, s1, s2, n2, nslices, bd, blk,
                                                     streaming);
}

// Line bounds of a pass (see k_fft_pass); false when the line is skipped.
template <int AXIS>
__device__ __forceinline__ bool line_range(const SphereBounds &bd, int l1, int l2, int &lo, int &hi) {
  lo = -(1 << 30);
  hi = 1 << 30;
  if (!bd.xb)
    return true;
  if (AXIS == 2) {
    const int sx = sgn_idx(l1, bd.n0), sy = sgn_idx(l2, bd.n1);
    if (sx < bd.xb[0] || sx > bd.xb[1] || sy < bd.yb[2 * l1] || sy > bd.yb[2 * l1 + 1])
      return false;
    lo = bd.zb[2 * (l1 * bd.n1 + l2)];
    hi = bd.zb[2 * (l1 * bd.n1 + l2) + 1];
  } else if (AXIS == 1) {
    const int sx = sgn_idx(l1, bd.n0);
    if (sx < bd.xb[0] || sx > bd.xb[1])
      return false;
    lo = bd.yb[2 * l1];
    hi = bd.yb[2 * l1 + 1];
  } else {
    lo = bd.xb[0];
    hi = bd.xb[1];
  }
  return true;
}

// Persistent, software-pipelined form of k_fft_pass_ip: a fixed grid of
// work-groups walks the lines; each thread holds its EPT elements of the
// next line in registers while the current line is transformed in LDS, so
// the global loads of line i+1 overlap the LDS stages of line i.
template <typename T, int NBW, int AXIS, int MAXB, int EPT>
__global__ __launch_bounds__(256) void k_fft_pass_pl(cplx<T> *__restrict__ g, AxisPlan pl,
                                                     const cplx<T> *__restrict__ tw, long es, long s1, long s2,
                                                     int n2, int nslices, long nlines, SphereBounds bd) {
  extern __shared__ unsigned char smem_raw[];
  cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
  const int n = pl.n;
  const int tot = n * NBW;
  cplx<T> *twl = buf + tot;
  stage_twiddles<T>(twl, tw, pl.twlen);
  cplx<T> reg[EPT];
  // find the first line of this work-group that has work
  long line = blockIdx.x;
  int lo = 0, hi = 0;
  for (; line < nlines; line += gridDim.x) {
    const long ln = pass_line<AXIS>(bd, (int)(line / nslices));
    if (line_range<AXIS>(bd, (int)(ln / n2), (int)(ln % n2), lo, hi))
      break;
  }
  auto load = [&](long L, int llo, int lhi) {
    const long ln = pass_line<AXIS>(bd, (int)(L / nslices));
    const int slice = (int)(L % nslices);
    const long base = (ln / n2) * s1 + (ln % n2) * s2 + (long)slice * NBW;
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      const int t = threadIdx.x + q * 256;
      cplx<T> v = mk<T>(0, 0);
      if (t < tot) {
        const int e = t / NBW, b = t % NBW;
        const int se = sgn_idx(e, n);
        if (se >= llo && se <= lhi)
          v = g[base + (long)e * es + b];
      }
      reg[q] = v;
    }
  };
  if (line < nlines)
    load(line, lo, hi);
  while (line < nlines) {
#pragma unroll
    for (int q = 0; q < EPT; q++) {
      const int t = threadIdx.x + q * 256;
      if (t < tot)
        buf[t] = reg[q];
    }
    __syncthreads();
    // prefetch the next line with work
    long next = line + gridDim.x;
    int nlo = 0, nhi = 0;
    for (; next < nlines; next += gridDim.x) {
      const long ln = pass_line<AXIS>(bd, (int)(next / nslices));
      if (line_range<AXIS>(bd, (int)(ln / n2), (int)(ln % n2), nlo, nhi))
        break;
    }
    if (next < nlines)
      load(next, nlo, nhi);
    stages_ip<T, NBW, MAXB, 256>(buf, twl, pl, n);
    {
      const long ln = pass_line<AXIS>(bd, (int)(line / nslices));
      const int slice = (int)(line % nslices);
      const long base = (ln / n2) * s1 + (ln % n2) * s2 + (long)slice * NBW;
      for (int t = threadIdx.x; t < tot; t += 256) {
        const int e = t / NBW, b = t % NBW;
        g[base + (long)e * es + b] = buf[t];
      }
    }
    __syncthreads();
    line = next;
  }
}

// Bounds of the plane-wave sphere of one k-point in signed grid indices.
// bnd layout: zb[2*n0*n1] | yb[2*n0] | xb[2] | alias flag.
__global__ __launch_bounds__(256) void k_bounds_init(int *__restrict__ bnd, int n) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    bnd[i] = (i == n - 1) ? 0 : ((i & 1) ? -(1 << 30) : (1 << 30));
}

__global__ __launch_bounds__(256) void k_bounds(int n0, int n1, int n2, int nw, const int *__restrict__ G,
                                                int *__restrict__ bnd) {
  const int w = blockIdx.x * 256 + threadIdx.x;
  if (w >= nw)
    return;
  int *zb = bnd, *flag = bnd + 2 * n0 * n1 + 2 * n0 + 2;
  const int g0 = G[3 * w], g1 = G[3 * w + 1], g2 = G[3 * w + 2];
  const int x = ((g0 % n0) + n0) % n0, y = ((g1 % n1) + n1) % n1, z = ((g2 % n2) + n2) % n2;
  const int sx = sgn_idx(x, n0), sy = sgn_idx(y, n1), sz = sgn_idx(z, n2);
  if (sx != g0 || sy != g1 || sz != g2)
    atomicOr(flag, 1);  // the grid aliases this G: no pruning
  atomicMin(&zb[2 * (x * n1 + y)], sz);
  atomicMax(&zb[2 * (x * n1 + y) + 1], sz);
}

// y extent of every plane x and the x extent, from the z-line bounds (one
// work-group; each thread scans the lines of its planes).
__global__ __launch_bounds__(256) void k_bounds_rows(int n0, int n1, int *__restrict__ bnd) {
  const int *zb = bnd;
  int *yb = bnd + 2 * n0 * n1, *xb = yb + 2 * n0;
  int xlo = 1 << 30, xhi = -(1 << 30);
  for (int x = threadIdx.x; x < n0; x += 256) {
    int lo = 1 << 30, hi = -(1 << 30);
    for (int y = 0; y < n1; y++)
      if (zb[2 * (x * n1 + y)] <= zb[2 * (x * n1 + y) + 1]) {
        const int sy = sgn_idx(y, n1);
        lo = min(lo, sy);
        hi = max(hi, sy);
      }
    yb[2 * x] = lo;
    yb[2 * x + 1] = hi;
    if (lo <= hi) {
      const int sx = sgn_idx(x, n0);
      xlo = min(xlo, sx);
      xhi = max(xhi, sx);
    }
  }
  __shared__ int rlo[256], rhi[256];
  rlo[threadIdx.x] = xlo;
  rhi[threadIdx.x] = xhi;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      rlo[threadIdx.x] = min(rlo[threadIdx.x], rlo[threadIdx.x + o]);
      rhi[threadIdx.x] = max(rhi[threadIdx.x], rhi[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    xb[0] = rlo[0];
    xb[1] = rhi[0];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_to_bandmajor(const cplx<T> *__restrict__ g, int NB, int nbc, long npts,
                                                      double scl, cplx<double> *__restrict__ out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts * nbc)
    return;
  long pt = i / nbc;
  int b = (int)(i % nbc);
  out[(long)b * npts + pt] = scale(cvt<double>(g[pt * NB + b]), scl);
}

// ------------------------------------------------------------------ host
int zero_async(pp_ctx *ctx, void *p, size_t bytes) {
  hipEvent_t ev;
  tbegin(ctx, K_ZERO, &ev);
  size_t n16 = bytes / 16;
  if (n16) {
    size_t blocks = (n16 + 255) / 256;
    if (blocks > 65536)
      blocks = 65536;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (uint4 *)p, n16);
  }
  size_t tail = bytes - n16 * 16;
  if (tail)
    hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(256), 0, ctx->stream, (unsigned char *)p + n16 * 16, tail);
  tend(ctx, K_ZERO, ev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

int compute_gidx(pp_ctx *ctx, const int *fftg, int nw, const int *G_dev, int *gidx_dev) {
  if (nw <= 0)
    return PP_OK;
  hipLaunchKernelGGL(k_gidx, dim3((nw + 255) / 256), dim3(256), 0, ctx->stream, fftg[0], fftg[1], fftg[2], nw,
                     G_dev, gidx_dev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

// Axis lengths with a compile-time pass (k_fft_pass_fx): the grids of the
// 64-atom (60), 216-atom (108), 512-atom (192) cells and the doubled
// 216-atom grid (216).
#define PP_FIXED_LENGTHS(X) X(60) X(108) X(192) X(216)
static bool fixed_length(int n) {
#define PP_FX_IS(NN) n == NN ||
  return PP_FIXED_LENGTHS(PP_FX_IS) false;
#undef PP_FX_IS
}
// The compile-time pass serves an axis when its length has one and the
// host plan's stages and twiddle offsets are the ones it was compiled with.
// Its results agree with the generic pass's to rounding (the compiler is
// free to contract the complex products differently in the two).
static bool fixed_plan_matches(const AxisPlan &ap) {
  if (!fixed_length(ap.n))
    return false;
  const FftFactors f = fft_factors(ap.n);
  if (f.nst != ap.nst)
    return false;
  for (int s = 0; s < f.nst; s++)
    if (f.r[s] != ap.radix[s] || f.p[s] != ap.pprev[s] || f.twoff[s] != ap.twoff[s])
      return false;
  return true;
}

template <typename T, int NBW, int AXIS>
static int launch_pass(pp_ctx *ctx, cplx<T> *grid, const AxisPlan &ap, const cplx<T> *tw, long es, long s1,
                       long s2, int n1, int n2, int NB, int cls, const SphereBounds &bd) {
  int nslices = NB / NBW;
  const long nlin = (AXIS == 2 && bd.zl) ? bd.nzl : (AXIS == 1 && bd.yl) ? bd.nyl : (long)n1 * n2;
  long nblk = nlin * nslices;
  if (nblk == 0)
    return PP_OK;
  // in-place form when every radix is 2..5 and each thread holds <= 8 butterflies
  int maxb = 0;
  bool ip = ctx->tune_fft_pingpong == 0 && NBW >= 8;
  // 512-thread work-groups for 32-band tiles keep the per-thread work of a
  // 16-band tile with 256 threads
  const int nt = (NBW >= 32 && ctx->tune_fft_threads == 512) ? 512 : 256;
  for (int s = 0; s < ap.nst; s++) {
    if (ap.radix[s] > 5)
      ip = false;
    maxb = std::max(maxb, ((ap.n / ap.radix[s]) * NBW + nt - 1) / nt);
  }
  if (maxb > 8)
    ip = false;
  // the composite factorization: fewer LDS round trips, when every stage
  // fits two butterflies per thread of 256
  bool comp = ip && ap.nstc > 0 && nt == 256 && ctx->tune_fft_persistent == 0 && ctx->tune_fft_radix == 1;
  for (int s = 0; comp && s < ap.nstc; s++)
    if (!dft_supported(ap.radixc[s]) || ((ap.n / ap.radixc[s]) * NBW + 255) / 256 > 2)
      comp = false;
  // streaming access: fft_nt 0..2 for every pass; 3 keeps the z pass's
  // stores cacheable (the y pass reads them next), streams the y pass's
  // stores and both directions of the x pass
  const int nt_mode = ctx->tune_fft_nt < 3 ? ctx->tune_fft_nt : (AXIS == 2 ? 0 : AXIS == 1 ? 1 : 2);
  // XCD-contiguous line ranges (fft_xcd = 1)
  const int xcd_per = ctx->tune_fft_xcd ? (int)((nblk + 7) / 8) : 0;
  const long nlaunch = xcd_per ? 8L * xcd_per : nblk;
  hipEvent_t ev;
  tbegin(ctx, cls, &ev);
  bool launched = false;
  const int ept = (ap.n * NBW + 255) / 256;
  if constexpr (NBW >= 8) {
    if (comp) {
      AxisPlan apc = ap;
      apc.nst = ap.nstc;
      for (int s = 0; s < ap.nstc; s++) {
        apc.radix[s] = ap.radixc[s];
        apc.pprev[s] = ap.pprevc[s];
        apc.pmag[s] = ap.pmagc[s];
        apc.twoff[s] = ap.twoffc[s];
      }
      size_t lds = ((size_t)ap.n * NBW + ap.twlen) * sizeof(cplx<T>);
      hipLaunchKernelGGL((k_fft_pass_ip<T, NBW, AXIS, 2, 256, true>), dim3((unsigned)nlaunch), dim3(256), lds,
                         ctx->stream, grid, apc, tw, es, s1, s2, n2, nslices, bd, xcd_per, (int)nblk, nt_mode);
      launched = true;
    }
  }
  if constexpr (NBW >= 8) {
    if (!launched && ip && ctx->tune_fft_persistent != 0 && ept <= 16) {
      size_t lds = ((size_t)ap.n * NBW + ap.twlen) * sizeof(cplx<T>);
      long nlines = nblk;
      long grid_wg = std::min<long>(nlines, (long)ctx->tune_fft_persistent > 1 ? ctx->tune_fft_persistent : 2048);
#define PP_PL(MB, EP)                                                                                          \
  hipLaunchKernelGGL((k_fft_pass_pl<T, NBW, AXIS, MB, EP>), dim3((unsigned)grid_wg), dim3(256), lds, ctx->stream,   \
                     grid, ap, tw, es, s1, s2, n2, nslices, nlines, bd)
      if (maxb <= 4 && ept <= 8)
        PP_PL(4, 8);
      else if (ept <= 8)
        PP_PL(8, 8);
      else if (maxb <= 4)
        PP_PL(4, 16);
      else
        PP_PL(8, 16);
#undef PP_PL
      launched = true;
    }
    // the compile-time pass for the default tile width (16 FP32, 8 FP64 bands)
    // and twice it
    if constexpr (NBW == (sizeof(T) == 4 ? 16 : 8) || NBW == (sizeof(T) == 4 ? 32 : 16)) {
      if (!launched && ip && nt == 256 && ctx->tune_fft_fixed && fixed_plan_matches(ap)) {
        size_t lds = ((size_t)ap.n * NBW + ap.twlen) * sizeof(cplx<T>);
#define PP_FX(NN)                                                                                               \
  case NN:                                                                                                      \
    hipLaunchKernelGGL((k_fft_pass_fx<T, NBW, AXIS, NN, 256>), dim3((unsigned)nlaunch), dim3(256), lds,         \
                       ctx->stream, grid, ap, tw, es, s1, s2, n2, nslices, bd, xcd_per, (int)nblk, nt_mode);    \
    launched = true;                                                                                            \
    break;
        switch (ap.n) { PP_FIXED_LENGTHS(PP_FX) default: break; }
#undef PP_FX
      }
    }
    if (!launched && ip) {
      size_t lds = ((size_t)ap.n * NBW + ap.twlen) * sizeof(cplx<T>);
#define PP_IP(MB, NT)                                                                                          \
  hipLaunchKernelGGL((k_fft_pass_ip<T, NBW, AXIS, MB, NT>), dim3((unsigned)nlaunch), dim3(NT), lds, ctx->stream,    \
                     grid, ap, tw, es, s1, s2, n2, nslices, bd, xcd_per, (int)nblk, nt_mode)
      bool done = false;
      if constexpr (NBW >= 32) {
        if (nt == 512) {
          if (maxb <= 2)
            PP_IP(2, 512);
          else if (maxb <= 4)
            PP_IP(4, 512);
          else
            PP_IP(8, 512);
          done = true;
        }
      }
      if (!done) {
        if (maxb <= 2)
          PP_IP(2, 256);
        else if (maxb <= 4)
          PP_IP(4, 256);
        else
          PP_IP(8, 256);
      }
#undef PP_IP
      launched = true;
    }
  }
  if (!launched) {
    size_t lds = (size_t)2 * ap.n * NBW * sizeof(cplx<T>);
    hipLaunchKernelGGL((k_fft_pass<T, NBW, AXIS>), dim3((unsigned)nblk), dim3(256), lds, ctx->stream, grid, ap, tw,
                       es, s1, s2, n2, nslices, bd);
  }
  tend(ctx, cls, ev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

template <typename T, int AXIS>
static int run_pass(pp_ctx *ctx, cplx<T> *grid, const AxisPlan &ap, const cplx<T> *tw, long es, long s1, long s2,
                    int n1, int n2, int NB, int cls, const SphereBounds &bd) {
  const size_t ldsmax = 64 * 1024;
  int nbw = ctx->tune_fft_tile > 0 ? ctx->tune_fft_tile : (sizeof(T) == 4 ? 16 : 8);
  while (nbw > 1 && (NB % nbw != 0 || (size_t)2 * ap.n * nbw * sizeof(cplx<T>) > ldsmax))
    nbw >>= 1;
  PP_REQUIRE((size_t)2 * ap.n * nbw * sizeof(cplx<T>) <= ldsmax, PP_ERR_UNSUPPORTED,
             "FFT line too long for LDS: " + std::to_string(ap.n));
  switch (nbw) {
  case 32:
    return launch_pass<T, 32, AXIS>(ctx, grid, ap, tw, es, s1, s2, n1, n2, NB, cls, bd);
  case 16:
    return launch_pass<T, 16, AXIS>(ctx, grid, ap, tw, es, s1, s2, n1, n2, NB, cls, bd);
  case 8:
    return launch_pass<T, 8, AXIS>(ctx, grid, ap, tw, es, s1, s2, n1, n2, NB, cls, bd);
  case 4:
    return launch_pass<T, 4, AXIS>(ctx, grid, ap, tw, es, s1, s2, n1, n2, NB, cls, bd);
  case 2:
    return launch_pass<T, 2, AXIS>(ctx, grid, ap, tw, es, s1, s2, n1, n2, NB, cls, bd);
  default:
    return launch_pass<T, 1, AXIS>(ctx, grid, ap, tw, es, s1, s2, n1, n2, NB, cls, bd);
  }
}

template <> const cplx<double> *fft_twiddles<double>(FftPlan *pl, int axis) {
  return (const cplx<double> *)pl->tw64.p + pl->twbase[axis];
}
template <> const cplx<float> *fft_twiddles<float>(FftPlan *pl, int axis) {
  return (const cplx<float> *)pl->tw32.p + pl->twbase[axis];
}
template <typename T> static const cplx<T> *tw_of(FftPlan *pl, int axis) { return fft_twiddles<T>(pl, axis); }

// One pass over a batch of work items (k_fft_pass_ip_items): the in-place
// form only; PP_ERR_UNSUPPORTED when the axis needs another form (a radix
// above 5, more than 8 butterflies per thread, the ping-pong stages).
template <typename T, int NBW, int AXIS>
static int launch_pass_items(pp_ctx *ctx, const ChunkItem *items, int nitems, long maxlines, const AxisPlan &ap,
                             const cplx<T> *tw, long es, long s1, long s2, int n2, int NB, int cls) {
  if constexpr (NBW < 8) {
    return PP_ERR_UNSUPPORTED;
  } else {
    const int nslices = NB / NBW;
    const long nblk = maxlines * nslices;
    if (nblk == 0 || nitems == 0)
      return PP_OK;
    const int nt = (NBW >= 32 && ctx->tune_fft_threads == 512) ? 512 : 256;
    int maxb = 0;
    bool ip = ctx->tune_fft_pingpong == 0;
    for (int s = 0; s < ap.nst; s++) {
      if (ap.radix[s] > 5)
        ip = false;
      maxb = std::max(maxb, ((ap.n / ap.radix[s]) * NBW + nt - 1) / nt);
    }
    if (!ip || maxb > 8)
      return PP_ERR_UNSUPPORTED;
    const int nt_mode = ctx->tune_fft_nt < 3 ? ctx->tune_fft_nt : (AXIS == 2 ? 0 : AXIS == 1 ? 1 : 2);
    const size_t lds = ((size_t)ap.n * NBW + ap.twlen) * sizeof(cplx<T>);
    dim3 grd((unsigned)nblk, (unsigned)nitems);
    hipEvent_t ev;
    tbegin(ctx, cls, &ev);
    bool done = false;
    if constexpr (NBW == (sizeof(T) == 4 ? 16 : 8) || NBW == (sizeof(T) == 4 ? 32 : 16)) {
      if (nt == 256 && ctx->tune_fft_fixed && fixed_plan_matches(ap)) {
#define PP_FXI(NN)                                                                                              \
  case NN:                                                                                                      \
    hipLaunchKernelGGL((k_fft_pass_ip_items<T, NBW, AXIS, 4, 256, NN>), grd, dim3(256), lds, ctx->stream, items, ap, \
                       tw, es, s1, s2, n2, nslices, nt_mode);                                                   \
    done = true;                                                                                                \
    break;
        switch (ap.n) { PP_FIXED_LENGTHS(PP_FXI) default: break; }
#undef PP_FXI
      }
    }
#define PP_IPI(MB, NT)                                                                                           \
  hipLaunchKernelGGL((k_fft_pass_ip_items<T, NBW, AXIS, MB, NT>), grd, dim3(NT), lds, ctx->stream, items, ap, tw, es, \
                     s1, s2, n2, nslices, nt_mode)
    if constexpr (NBW >= 32) {
      if (!done && nt == 512) {
        if (maxb <= 2)
          PP_IPI(2, 512);
        else if (maxb <= 4)
          PP_IPI(4, 512);
        else
          PP_IPI(8, 512);
        done = true;
      }
    }
    if (!done) {
      if (maxb <= 2)
        PP_IPI(2, 256);
      else if (maxb <= 4)
        PP_IPI(4, 256);
      else
        PP_IPI(8, 256);
    }
#undef PP_IPI
    tend(ctx, cls, ev);
    PP_CHECK_HIP(hipGetLastError());
    return PP_OK;
  }
}

template <typename T, int AXIS>
static int run_pass_items(pp_ctx *ctx, const ChunkItem *items, int nitems, long maxlines, const AxisPlan &ap,
                          const cplx<T> *tw, long es, long s1, long s2, int n2, int NB, int cls) {
  const size_t ldsmax = 64 * 1024;
  int nbw = ctx->tune_fft_tile > 0 ? ctx->tune_fft_tile : (sizeof(T) == 4 ? 16 : 8);
  while (nbw > 1 && (NB % nbw != 0 || (size_t)2 * ap.n * nbw * sizeof(cplx<T>) > ldsmax))
    nbw >>= 1;
  switch (nbw) {
  case 32:
    return launch_pass_items<T, 32, AXIS>(ctx, items, nitems, maxlines, ap, tw, es, s1, s2, n2, NB, cls);
  case 16:
    return launch_pass_items<T, 16, AXIS>(ctx, items, nitems, maxlines, ap, tw, es, s1, s2, n2, NB, cls);
  case 8:
    return launch_pass_items<T, 8, AXIS>(ctx, items, nitems, maxlines, ap, tw, es, s1, s2, n2, NB, cls);
  default:
    return PP_ERR_UNSUPPORTED;
  }
}

// True when the item form of the passes serves this grid and chunk width.
template <typename T> bool fft_items_supported(pp_ctx *ctx, FftPlan *pl, int NB) {
  if (ctx->tune_fft_pingpong)
    return false;
  int nbw = ctx->tune_fft_tile > 0 ? ctx->tune_fft_tile : (sizeof(T) == 4 ? 16 : 8);
  for (int a = 0; a < 3; a++) {
    const AxisPlan &ap = pl->ax[a];
    int w = nbw;
    while (w > 1 && (NB % w != 0 || (size_t)2 * ap.n * w * sizeof(cplx<T>) > 64 * 1024))
      w >>= 1;
    if (w < 8)
      return false;
    const int nt = (w >= 32 && ctx->tune_fft_threads == 512) ? 512 : 256;
    for (int s = 0; s < ap.nst; s++)
      if (ap.radix[s] > 5 || ((ap.n / ap.radix[s]) * w + nt - 1) / nt > 8)
        return false;
  }
  return true;
}

// Scatter and the three pruned passes of a batch of work items (device
// array `items` of nitems), every item's chunk pruned to its k-point's
// plane-wave sphere: one launch per kernel for the whole batch.  maxw,
// maxzl and maxyl bound the plane waves, z lines and y lines of the items.
template <typename T>
int fft_items(pp_ctx *ctx, FftPlan *pl, const ChunkItem *items, int nitems, int NB, int maxw, long maxzl,
              long maxyl) {
  const long n0 = pl->fftg[0], n1 = pl->fftg[1], n2 = pl->fftg[2];
  if (maxw > 0) {
    hipEvent_t ev;
    tbegin(ctx, K_SCATTER, &ev);
    hipLaunchKernelGGL(k_scatter_items<T>, dim3((unsigned)((maxw + 31) / 32), (unsigned)nitems), dim3(256), 0,
                       ctx->stream, items, NB);
    tend(ctx, K_SCATTER, ev);
    PP_CHECK_HIP(hipGetLastError());
  }
  PP_TRY((run_pass_items<T, 2>(ctx, items, nitems, maxzl, pl->ax[2], fft_twiddles<T>(pl, 2), NB, n1 * n2 * NB,
                               n2 * NB, (int)n1, NB, K_FFT_Z)));
  PP_TRY((run_pass_items<T, 1>(ctx, items, nitems, maxyl, pl->ax[1], fft_twiddles<T>(pl, 1), n2 * NB, n1 * n2 * NB,
                               NB, (int)n2, NB, K_FFT_Y)));
  PP_TRY((run_pass_items<T, 0>(ctx, items, nitems, n1 * n2, pl->ax[0], fft_twiddles<T>(pl, 0), n1 * n2 * NB,
                               n2 * NB, NB, (int)n2, NB, K_FFT_X)));
  return PP_OK;
}
template int fft_items<double>(pp_ctx *, FftPlan *, const ChunkItem *, int, int, int, long, long);
template int fft_items<float>(pp_ctx *, FftPlan *, const ChunkItem *, int, int, int, long, long);
template bool fft_items_supported<double>(pp_ctx *, FftPlan *, int);
template bool fft_items_supported<float>(pp_ctx *, FftPlan *, int);


template <typename T>
int fft_chunk(pp_ctx *ctx, FftPlan *pl, cplx<T> *grid, int NB, int nw, const int *gidx_dev,
              const cplx<float> *C_dev, long ldc, int nbc, const SphereBounds *bounds, bool last_pass) {
  const long n0 = pl->fftg[0], n1 = pl->fftg[1], n2 = pl->fftg[2];
  const long npts = n0 * n1 * n2;
  SphereBounds bd = {nullptr, nullptr, nullptr, (int)n0, (int)n1, (int)n2};
  if (bounds)
    bd = *bounds;
  else
    PP_TRY(zero_async(ctx, grid, (size_t)npts * NB * sizeof(cplx<T>)));
  if (nw > 0) {
    hipEvent_t ev;
    tbegin(ctx, K_SCATTER, &ev);
    hipLaunchKernelGGL(k_scatter<T>, dim3((nw + 31) / 32), dim3(256), 0, ctx->stream, grid, NB, gidx_dev, nw, C_dev,
                       ldc, nbc);
    tend(ctx, K_SCATTER, ev);
    PP_CHECK_HIP(hipGetLastError());
  }
  // z lines: (x, y); y lines: (x, z); x lines: (y, z)
  PP_TRY((run_pass<T, 2>(ctx, grid, pl->ax[2], tw_of<T>(pl, 2), NB, n1 * n2 * NB, n2 * NB, (int)n0, (int)n1, NB,
                         K_FFT_Z, bd)));
  PP_TRY((run_pass<T, 1>(ctx, grid, pl->ax[1], tw_of<T>(pl, 1), n2 * NB, n1 * n2 * NB, NB, (int)n0, (int)n2, NB,
                         K_FFT_Y, bd)));
  if (last_pass)
    PP_TRY((run_pass<T, 0>(ctx, grid, pl->ax[0], tw_of<T>(pl, 0), n1 * n2 * NB, n2 * NB, NB, (int)n1, (int)n2, NB,
                           K_FFT_X, bd)));
  return PP_OK;
}

int sphere_bounds(pp_ctx *ctx, const int *fftg, int nw, const int *G_dev, int *bnd_dev) {
  const int n = 2 * fftg[0] * fftg[1] + 2 * fftg[0] + 2 + 1;
  hipLaunchKernelGGL(k_bounds_init, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, bnd_dev, n);
  if (nw > 0)
    hipLaunchKernelGGL(k_bounds, dim3((nw + 255) / 256), dim3(256), 0, ctx->stream, fftg[0], fftg[1], fftg[2], nw,
                       G_dev, bnd_dev);
  hipLaunchKernelGGL(k_bounds_rows, dim3(1), dim3(256), 0, ctx->stream, fftg[0], fftg[1], bnd_dev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

bool bounds_filled(const int *hb, const int *fftg, long nw) {
  const long nl = (long)fftg[0] * fftg[1];
  long tot = 0;
  for (long l = 0; l < nl; l++)
    if (hb[2 * l] <= hb[2 * l + 1])
      tot += hb[2 * l + 1] - hb[2 * l] + 1;
  return tot == nw;
}

// The three unpruned passes over a full chunk (no scatter, no zero fill).
template <typename T>
int fft_passes(pp_ctx *ctx, FftPlan *pl, cplx<T> *grid, int NB) {
  const long n0 = pl->fftg[0], n1 = pl->fftg[1], n2 = pl->fftg[2];
  SphereBounds bd = {nullptr, nullptr, nullptr, (int)n0, (int)n1, (int)n2};
  PP_TRY((run_pass<T, 2>(ctx, grid, pl->ax[2], tw_of<T>(pl, 2), NB, n1 * n2 * NB, n2 * NB, (int)n0, (int)n1, NB,
                         K_FFT_Z, bd)));
  PP_TRY((run_pass<T, 1>(ctx, grid, pl->ax[1], tw_of<T>(pl, 1), n2 * NB, n1 * n2 * NB, NB, (int)n0, (int)n2, NB,
                         K_FFT_Y, bd)));
  PP_TRY((run_pass<T, 0>(ctx, grid, pl->ax[0], tw_of<T>(pl, 0), n1 * n2 * NB, n2 * NB, NB, (int)n1, (int)n2, NB,
                         K_FFT_X, bd)));
  return PP_OK;
}

// out[b * ldo + w] = scl conj(grid[gidx[w]][b]): the plane-wave coefficients
// of a forward transform done as conj(backward(conj x)).
template <typename T>
__global__ __launch_bounds__(256) void k_gather_conj(const cplx<T> *__restrict__ g, int NB, int nbc,
                                                     const int *__restrict__ gidx, int nw, double scl,
                                                     cplx<double> *__restrict__ out, long ldo) {
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)nw * nbc)
    return;
  const int b = (int)(t % nbc);
  const int w = (int)(t / nbc);
  const cplx<double> v = cvt<double>(g[(long)gidx[w] * NB + b]);
  out[(long)b * ldo + w] = mk<double>(scl * v.x, -scl * v.y);
}

template <typename T>
int gather_conj(pp_ctx *ctx, const cplx<T> *grid, int NB, int nbc, const int *gidx, int nw, double scl,
                cplx<double> *out, long ldo) {
  const long n = (long)nw * nbc;
  if (n == 0)
    return PP_OK;
  hipEvent_t ev;
  tbegin(ctx, K_AUG_PREP, &ev);
  hipLaunchKernelGGL(k_gather_conj<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, grid, NB, nbc,
                     gidx, nw, scl, out, ldo);
  tend(ctx, K_AUG_PREP, ev);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

template int fft_passes<double>(pp_ctx *, FftPlan *, cplx<double> *, int);
template int fft_passes<float>(pp_ctx *, FftPlan *, cplx<float> *, int);
template int gather_conj<double>(pp_ctx *, const cplx<double> *, int, int, const int *, int, double, cplx<double> *,
                                 long);
template int gather_conj<float>(pp_ctx *, const cplx<float> *, int, int, const int *, int, double, cplx<double> *,
                                long);

template int fft_chunk<double>(pp_ctx *, FftPlan *, cplx<double> *, int, int, const int *, const cplx<float> *,
                               long, int, const SphereBounds *, bool);
template int fft_chunk<float>(pp_ctx *, FftPlan *, cplx<float> *, int, int, const int *, const cplx<float> *, long,
                              int, const SphereBounds *, bool);

int grid_to_bandmajor(pp_ctx *ctx, int precision, const void *grid, int NB, int nbc, long npts, double scl,
                      void *out) {
  long n = npts * nbc;
  unsigned blocks = (unsigned)((n + 255) / 256);
  if (precision == PP_FP64)
    hipLaunchKernelGGL(k_to_bandmajor<double>, dim3(blocks), dim3(256), 0, ctx->stream,
                       (const cplx<double> *)grid, NB, nbc, npts, scl, (cplx<double> *)out);
  else
    hipLaunchKernelGGL(k_to_bandmajor<float>, dim3(blocks), dim3(256), 0, ctx->stream, (const cplx<float> *)grid,
                       NB, nbc, npts, scl, (cplx<double> *)out);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

}  // namespace pp
