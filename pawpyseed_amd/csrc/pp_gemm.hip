// Complex "NT" GEMM: C[m][n] (+)= sum_k conj(A[m][k]) B[n][k].
//
// The band-overlap matrix of pseudoprojector.c:63-90 (one cdotc per band
// pair) and the O_M augmentation sum of projector.c:890-910 are both this
// contraction with K contiguous in both operands.  FP64 path: FP64 matrix
// cores (k_zgemm_mfma64) or, selectable, a VALU tile GEMM (64x64 output
// tiles, 4x4 complex accumulators per thread); K tiles of 16 staged through
// LDS; split-K with a deterministic reduction when the tile grid alone
// cannot fill the 256 CUs.
#include "pp_internal.h"

namespace pp {

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_zgemm_f64(int M, int N, int kbeg_stride, int K, int nsplit,
                                                   const cplx<TA> *__restrict__ A, long lda,
                                                   const cplx<TB> *__restrict__ B, long ldb,
                                                   cplx<double> *__restrict__ C, long ldc, long sSplit) {
  __shared__ cplx<double> As[16][64 + 1];
  __shared__ cplx<double> Bs[16][64 + 1];
  const int tilesN = (N + 63) / 64;
  int split, tile;
  if (!gemm_block(((M + 63) / 64) * tilesN, nsplit, split, tile))
    return;
  const int tm = tile / tilesN, tn = tile % tilesN;
  const int k0 = split * kbeg_stride;
  const int k1 = min(K, k0 + kbeg_stride);
  C += (long)split * sSplit;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  double ar[4][4], ai[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++)
      ar[i][j] = ai[i][j] = 0.0;
  for (int kb = k0; kb < k1; kb += 16) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int e = tid + q * 256;
      int r = e >> 4, kk = e & 15;
      int gm = tm * 64 + r, gn = tn * 64 + r, gk = kb + kk;
      cplx<double> va = mk<double>(0, 0), vb = mk<double>(0, 0);
      if (gk < k1) {
        if (gm < M)
          va = cvt<double>(A[(long)gm * lda + gk]);
        if (gn < N)
          vb = cvt<double>(B[(long)gn * ldb + gk]);
      }
      As[kk][r] = va;
      Bs[kk][r] = vb;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; kk++) {
      cplx<double> a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++)
        a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; j++)
        b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          // conj(a) * b
          ar[i][j] = fma(a[i].x, b[j].x, ar[i][j]);
          ar[i][j] = fma(a[i].y, b[j].y, ar[i][j]);
          ai[i][j] = fma(a[i].x, b[j].y, ai[i][j]);
          ai[i][j] = fma(-a[i].y, b[j].x, ai[i][j]);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int gm = tm * 64 + ty + 16 * i, gn = tn * 64 + tx + 16 * j;
      if (gm < M && gn < N)
        C[(long)gm * ldc + gn] = mk<double>(ar[i][j], ai[i][j]);
    }
}

typedef double doublex4 __attribute__((ext_vector_type(4)));

// FP64 matrix-core form: conj(A) B^T as two real GEMMs over the
// interleaved (re, im) storage (K' = 2K; Im uses B'' = (b_im, -b_re)) on
// v_mfma_f64_16x16x4_f64.  Work-group: 4 wavefronts, 64x64 complex
// outputs, each wavefront a 32x32 quadrant of 2x2 16x16 tiles.  Output
// layout of a 16x16 tile: lane l, register v holds row l/16 + 4v, column
// l%16 (unlike the FP32 16x16x4 form, whose register v is row 4(l/16)+v).
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_zgemm_mfma64(int M, int N, int kchunk, int K, int nsplit,
                                                      const cplx<TA> *__restrict__ A, long lda,
                                                      const cplx<TB> *__restrict__ B, long ldb,
                                                      cplx<double> *__restrict__ C, long ldc, long sSplit) {
  __shared__ double As[64][36];
  __shared__ double Bs[64][36];
  const int tilesN = (N + 63) / 64;
  int split, tile;
  if (!gemm_block(((M + 63) / 64) * tilesN, nsplit, split, tile))
    return;
  const int tm = tile / tilesN, tn = tile % tilesN;
  const int k0 = split * kchunk;
  const int k1 = min(K, k0 + kchunk);
  C += (long)split * sSplit;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  doublex4 cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      cr[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
      ci[a][b] = doublex4{0.0, 0.0, 0.0, 0.0};
    }
  const int lrow = threadIdx.x >> 2, part = threadIdx.x & 3;
  const int gm = tm * 64 + lrow, gn = tn * 64 + lrow;
  // a k-step's 4 + 4 values of a thread are loaded together, the next
  // step's while the current step's MFMAs run
  cplx<TA> pa[4];
  cplx<TB> pb[4];
  auto fetch = [&](int kb) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int gk = kb + part * 4 + c;
      pa[c] = (gk < k1 && gm < M) ? A[(long)gm * lda + gk] : mk<TA>((TA)0, (TA)0);
      pb[c] = (gk < k1 && gn < N) ? B[(long)gn * ldb + gk] : mk<TB>((TB)0, (TB)0);
    }
  };
  if (k0 < k1)
    fetch(k0);
  for (int kb = k0; kb < k1; kb += 16) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const cplx<double> va = cvt<double>(pa[c]), vb = cvt<double>(pb[c]);
      As[lrow][part * 8 + 2 * c] = va.x;
      As[lrow][part * 8 + 2 * c + 1] = va.y;
      Bs[lrow][part * 8 + 2 * c] = vb.x;
      Bs[lrow][part * 8 + 2 * c + 1] = vb.y;
    }
    __syncthreads();
    if (kb + 16 < k1)
      fetch(kb + 16);
    const int sub = lane & 15, kq = lane >> 4;
    // B'' = (b_im, -b_re): the neighbour k ^ 1 with its sign bit flipped for
    // odd k -- one LDS read and no divergent branch
    const unsigned long long sgn = (unsigned long long)(kq & 1) << 63;
#pragma unroll
    for (int kk = 0; kk < 32; kk += 4) {
      const int kr = kk + kq;
      const double a0 = As[wm * 32 + sub][kr];
      const double a1 = As[wm * 32 + 16 + sub][kr];
#pragma unroll
      for (int nj = 0; nj < 2; nj++) {
        const int col = wn * 32 + nj * 16 + sub;
        const double b = Bs[col][kr];
        const double b2 = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, Bs[col][kr ^ 1]) ^ sgn);
        cr[0][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b, cr[0][nj], 0, 0, 0);
        ci[0][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b2, ci[0][nj], 0, 0, 0);
        cr[1][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b, cr[1][nj], 0, 0, 0);
        ci[1][nj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b2, ci[1][nj], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int mi = 0; mi < 2; mi++)
#pragma unroll
    for (int nj = 0; nj < 2; nj++) {
      const int j = tn * 64 + wn * 32 + nj * 16 + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = tm * 64 + wm * 32 + mi * 16 + (lane >> 4) + 4 * v;
        if (i < M && j < N)
          C[(long)i * ldc + j] = mk<double>(cr[mi][nj][v], ci[mi][nj][v]);
      }
    }
}

// C[m][n] (+)= sum over splits of partial[s][m][n], in FP64 (partials FP64
// or FP32)
template <typename TP>
__global__ __launch_bounds__(256) void k_splitk_reduce(int M, int N, int nsplit, const cplx<TP> *__restrict__ P,
                                                       cplx<double> *__restrict__ C, long ldc, int accumulate) {
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)M * N)
    return;
  int m = (int)(t / N), n = (int)(t % N);
  cplx<double> s = accumulate ? C[(long)m * ldc + n] : mk<double>(0, 0);
  for (int q = 0; q < nsplit; q++)
    s = s + cvt<double>(P[(long)q * M * N + t]);
  C[(long)m * ldc + n] = s;
}

int gemm_fp32_mfma(pp_ctx *ctx, int precision, int M, int N, int K, const void *A, long lda, const void *B, long ldb,
                   cplx<double> *C, long ldc, bool accumulate, int slot);
bool plane_precision(int precision);
int plane_gemm_pipeline(pp_ctx *ctx, int precision, int M, int N, int K, const void *A, long lda, long sA,
                        const void *B, long ldb, long sB, cplx<double> *C, long ldc, long sC, int nbatch,
                        const int *amap, const int *bmap, const int *klen, bool accumulate);

// Split-K policy of the overlap GEMMs.  FP64: enough K chunks that the
// 64x64 output tiles times the chunks give about 1024 work-groups (four per
// CU), chunks of at least 256 complex terms.  FP32: chunks of at most 512
// complex terms, whose FP32 sums are exact to the FP32 accumulation of 512
// terms, summed in FP64 by the reduction (a kernel that instead folds its
// FP32 sums into FP64 inside longer chunks holds 64 more registers per lane
// and runs at a third of the occupancy: 94 against 119 TFLOP/s at 216
// atoms).  (The split-plane GEMMs chunk K by their plane layout.)  A
// positive `tune` forces the count.
int splitk_count(int precision, int M, int N, int K, int tune) {
  if (tune > 0)
    return tune;
  const int tiles = std::max(1, ((M + 63) / 64) * ((N + 63) / 64));
  int n = (1024 + tiles - 1) / tiles;
  if (precision == PP_FP32)
    n = std::max(n, (K + 511) / 512);
  n = std::min(n, std::max(1, K / 256));
  return std::max(1, n);
}

// FP64 GEMM of one batch into C (nsplit == 1) or into per-chunk partials P
// ([nsplit][M][N]); kchunk a multiple of 16.  No allocation.
int launch_zgemm64(hipStream_t st, bool valu, int M, int N, int K, int atype, const void *A, long lda, int btype,
                   const void *B, long ldb, cplx<double> *dst, long dld, long sSplit, int nsplit, int kchunk) {
  const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
  const dim3 grd = gemm_grid(tiles, nsplit);
  if (!valu) {
#define PP_MF(TA, TB)                                                                                          \
  hipLaunchKernelGGL((k_zgemm_mfma64<TA, TB>), grd, dim3(256), 0, st, M, N, kchunk, K, nsplit, (const cplx<TA> *)A, lda, \
                     (const cplx<TB> *)B, ldb, dst, dld, sSplit)
    if (atype == 1 && btype == 1)
      PP_MF(float, float);
    else if (atype == 0 && btype == 0)
      PP_MF(double, double);
    else if (atype == 0 && btype == 1)
      PP_MF(double, float);
    else
      PP_MF(float, double);
#undef PP_MF
  } else {
#define PP_MV(TA, TB)                                                                                          \
  hipLaunchKernelGGL((k_zgemm_f64<TA, TB>), grd, dim3(256), 0, st, M, N, kchunk, K, nsplit, (const cplx<TA> *)A, lda, \
                     (const cplx<TB> *)B, ldb, dst, dld, sSplit)
    if (atype == 1 && btype == 1)
      PP_MV(float, float);
    else if (atype == 0 && btype == 0)
      PP_MV(double, double);
    else if (atype == 0 && btype == 1)
      PP_MV(double, float);
    else
      PP_MV(float, double);
#undef PP_MV
  }
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

int launch_splitk_reduce(hipStream_t st, int M, int N, int nsplit, const void *P, bool fp32_partials,
                         cplx<double> *C, long ldc, bool accumulate) {
  const long tot = (long)M * N;
  if (fp32_partials)
    hipLaunchKernelGGL(k_splitk_reduce<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, M, N, nsplit,
                       (const cplx<float> *)P, C, ldc, accumulate ? 1 : 0);
  else
    hipLaunchKernelGGL(k_splitk_reduce<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, M, N, nsplit,
                       (const cplx<double> *)P, C, ldc, accumulate ? 1 : 0);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

static int zgemm_one(pp_ctx *ctx, int precision, int M, int N, int K, int atype, const void *A, long lda, int btype,
                     const void *B, long ldb, cplx<double> *C, long ldc, bool accumulate, int cls, int slot) {
  if (M <= 0 || N <= 0)
    return PP_OK;
  if (precision != PP_FP64 && atype == 1 && btype == 1)
    return gemm_fp32_mfma(ctx, precision, M, N, K, A, lda, B, ldb, C, ldc, accumulate, slot);
  int nsplit = splitk_count(PP_FP64, M, N, K, ctx->tune_splitk);
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = (kchunk + 15) / 16 * 16;
  nsplit = std::max(1, (K + kchunk - 1) / std::max(1, kchunk));
  bool direct = nsplit == 1 && !accumulate;
  cplx<double> *dst = C;
  long dld = ldc, sSplit = 0;
  if (!direct) {
    PP_TRY(ctx->gemm_ws.ensure(sizeof(cplx<double>) * (size_t)nsplit * M * N));
    dst = (cplx<double> *)ctx->gemm_ws.p;
    dld = N;
    sSplit = (long)M * N;
  }
  hipEvent_t ev;
  tbegin(ctx, cls, &ev);
  int st = launch_zgemm64(ctx->stream, ctx->tune_gemm_f64_valu != 0, M, N, K, atype, A, lda, btype, B, ldb, dst, dld,
                          sSplit, nsplit, std::max(16, kchunk));
  tend(ctx, cls, ev);
  PP_TRY(st);
  if (!direct) {
    tbegin(ctx, K_GEMM_REDUCE, &ev);
    st = launch_splitk_reduce(ctx->stream, M, N, nsplit, ctx->gemm_ws.p, false, C, ldc, accumulate);
    tend(ctx, K_GEMM_REDUCE, ev);
  }
  return st;
}

int zgemm_nt(pp_ctx *ctx, int precision, int M, int N, int K, int atype, const void *A, long lda, long sA, int btype,
             const void *B, long ldb, long sB, void *C, long ldc, long sC, int nbatch, const int *amap,
             const int *bmap, const int *klen, bool accumulate) {
  size_t ea = atype == 0 ? 16 : 8, eb = btype == 0 ? 16 : 8;
  // FP32 / 3xBF16 batches: the split-K reduction of one batch runs on the
  // second stream while the next batch's GEMM runs (two partials buffers)
  const bool pipe = precision != PP_FP64 && atype == 1 && btype == 1 && nbatch > 1 && ctx->tune_streams > 1 &&
                    M > 0 && N > 0;
  if (pipe) {
    if (!ctx->aux)
      PP_CHECK_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    for (int i = 0; i < 2; i++) {
      if (!ctx->gemm_done[i])
        PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->gemm_done[i], hipEventDisableTiming));
      if (!ctx->red_done[i])
        PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->red_done[i], hipEventDisableTiming));
      PP_CHECK_HIP(hipEventRecord(ctx->red_done[i], ctx->stream));  // no reduction pending yet
    }
    if (!ctx->join)
      PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
  }
  int st = PP_OK;
  if (pipe && plane_precision(precision)) {
    st = plane_gemm_pipeline(ctx, precision, M, N, K, A, lda, sA, B, ldb, sB, (cplx<double> *)C, ldc, sC, nbatch,
                             amap, bmap, klen, accumulate);
    (void)hipEventRecord(ctx->join, ctx->aux);
    (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
    return st;
  }
  for (int b = 0; b < nbatch && st == PP_OK; b++) {
    int ia = amap ? amap[b] : b, ib = bmap ? bmap[b] : b;
    int kk = klen ? klen[b] : K;
    const char *pa = (const char *)A + (size_t)ia * sA * ea;
    const char *pb = (const char *)B + (size_t)ib * sB * eb;
    cplx<double> *pc = (cplx<double> *)C + (size_t)b * sC;
    st = zgemm_one(ctx, precision, M, N, kk, atype, pa, lda, btype, pb, ldb, pc, ldc, accumulate,
                   atype == 0 ? K_AUG_GEMM : K_GEMM, pipe ? (b & 1) : -1);
  }
  if (pipe) {
    // the caller's stream sees every reduction
    (void)hipEventRecord(ctx->join, ctx->aux);
    (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
  }
  return st;
}

}  // namespace pp
