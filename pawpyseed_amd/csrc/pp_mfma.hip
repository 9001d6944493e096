// Mixed-precision band-overlap GEMM on the CDNA4 matrix cores.
//
// Complex conj(A) B^T with c64 operands is two real GEMMs over the
// interleaved (re, im) storage taken as a real K' = 2K dimension:
//   Re = sum_k' A'[m][k'] B'[n][k'],   B' = (b_re, b_im) as stored,
//   Im = sum_k' A'[m][k'] B''[n][k'],  B'' = (b_im, -b_re),
// so no operand repacking is needed for the real part and B'' is the
// stored pair swapped with one sign flip.
//
// PP_FP32: v_mfma_f32_16x16x4_f32; FP32 sums are folded into FP64 every
// 512 complex terms inside a split-K chunk, chunks summed in FP64.
// PP_BF16X3 / PP_BF16X6: each FP32 operand is split once per call into two
// (hi, lo) or three (hi, mid, lo) bf16 planes; the 3 or 6 plane products
// that carry 16 or 24 significant bits run on v_mfma_f32_16x16x32_bf16
// with FP32 accumulation, folded into FP64 like the FP32 path.
//
// Work-group: 4 wavefronts, 64x64 complex outputs, each wavefront a 32x32
// quadrant as 2x2 16x16 MFMA tiles (real and imaginary accumulators).
#include "pp_internal.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace pp {

// MFMA 16x16 output layout: lane l, register v holds row 4*(l/16)+v,
// column l%16 of the 16x16 tile.
template <typename TO>
__device__ __forceinline__ void store_tile(cplx<TO> *C, long ldc, int M, int N, int r0, int c0, floatx4 re,
                                           floatx4 im, int lane) {
  const int j = c0 + (lane & 15);
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int i = r0 + 4 * (lane >> 4) + v;
    if (i < M && j < N)
      C[(long)i * ldc + j] = mk<TO>((TO)re[v], (TO)im[v]);
  }
}

// As store_tile, adding FP64 partial sums of earlier K blocks.
__device__ __forceinline__ void store_tile64(cplx<double> *C, long ldc, int M, int N, int r0, int c0,
                                             const double *dre, const double *dim, floatx4 re, floatx4 im,
                                             int lane) {
  const int j = c0 + (lane & 15);
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int i = r0 + 4 * (lane >> 4) + v;
    if (i < M && j < N)
      C[(long)i * ldc + j] = mk<double>(dre[v] + (double)re[v], dim[v] + (double)im[v]);
  }
}

// Load 4 consecutive complex (8 floats) of row `row` starting at complex
// column gk, zero outside [0, M) x [.., k1).
template <bool VEC>
__device__ __forceinline__ void load8(const float *__restrict__ P, long ld, int row, int nrows, int gk, int k1,
                                      float *v) {
  if (VEC && row < nrows && gk + 4 <= k1) {
    const float4 *p = reinterpret_cast<const float4 *>(P + ((long)row * ld + gk) * 2);
    float4 x = p[0], y = p[1];
    v[0] = x.x;
    v[1] = x.y;
    v[2] = x.z;
    v[3] = x.w;
    v[4] = y.x;
    v[5] = y.y;
    v[6] = y.z;
    v[7] = y.w;
  } else {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      bool ok = row < nrows && gk + c < k1;
      v[2 * c] = ok ? P[((long)row * ld + gk + c) * 2] : 0.f;
      v[2 * c + 1] = ok ? P[((long)row * ld + gk + c) * 2 + 1] : 0.f;
    }
  }
}

// FOLD: FP32 sums folded into FP64 every 512 complex terms, FP64 partial
// sums out; otherwise a chunk's FP32 sums are its (FP32) partial sums.
template <bool VEC, bool FOLD>
__global__ __launch_bounds__(256) void k_cgemm_f32(int M, int N, int kchunk, int K, int nsplit,
                                                   const float *__restrict__ A, long lda,
                                                   const float *__restrict__ B, long ldb, void *__restrict__ Cv,
                                                   long ldc, long sSplit) {
  __shared__ float As[64][36];
  __shared__ float Bs[64][36];
  const int tilesN = (N + 63) / 64;
  int split, tile;
  if (!gemm_block(((M + 63) / 64) * tilesN, nsplit, split, tile))
    return;
  const int tm = tile / tilesN, tn = tile % tilesN;
  const int k0 = split * kchunk;
  const int k1 = min(K, k0 + kchunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  floatx4 cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      cr[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
      ci[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
  double dr[2][2][4], di[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int v = 0; v < 4; v++)
        dr[a][b][v] = di[a][b][v] = 0.0;
  const int lrow = threadIdx.x >> 2, part = threadIdx.x & 3;
  int since_flush = 0;
  // the next k-step's A and B values are loaded into registers while the
  // current step's MFMAs run
  float va[8], vb[8];
  if (k0 < k1) {
    load8<VEC>(A, lda, tm * 64 + lrow, M, k0 + part * 4, k1, va);
    load8<VEC>(B, ldb, tn * 64 + lrow, N, k0 + part * 4, k1, vb);
  }
  for (int kb = k0; kb < k1; kb += 16) {
#pragma unroll
    for (int c = 0; c < 8; c++) {
      As[lrow][part * 8 + c] = va[c];
      Bs[lrow][part * 8 + c] = vb[c];
    }
    __syncthreads();
    if (kb + 16 < k1) {
      load8<VEC>(A, lda, tm * 64 + lrow, M, kb + 16 + part * 4, k1, va);
      load8<VEC>(B, ldb, tn * 64 + lrow, N, kb + 16 + part * 4, k1, vb);
    }
    const int sub = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int kk = 0; kk < 32; kk += 4) {
      const int kr = kk + kq;
      const float a0 = As[wm * 32 + sub][kr];
      const float a1 = As[wm * 32 + 16 + sub][kr];
#pragma unroll
      for (int nj = 0; nj < 2; nj++) {
        const int col = wn * 32 + nj * 16 + sub;
        const float b = Bs[col][kr];
        const float b2 = (kq & 1) ? -Bs[col][kr - 1] : Bs[col][kr + 1];
        cr[0][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b, cr[0][nj], 0, 0, 0);
        ci[0][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b2, ci[0][nj], 0, 0, 0);
        cr[1][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b, cr[1][nj], 0, 0, 0);
        ci[1][nj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b2, ci[1][nj], 0, 0, 0);
      }
    }
    __syncthreads();
    // bound the FP32 accumulation length: fold into FP64 every 512 terms
    if (FOLD && ++since_flush == 32) {
      since_flush = 0;
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
#pragma unroll
          for (int v = 0; v < 4; v++) {
            dr[a][b][v] += (double)cr[a][b][v];
            di[a][b][v] += (double)ci[a][b][v];
          }
          cr[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
          ci[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
    }
  }
#pragma unroll
  for (int mi = 0; mi < 2; mi++)
#pragma unroll
    for (int nj = 0; nj < 2; nj++) {
      const int r0 = tm * 64 + wm * 32 + mi * 16, c0 = tn * 64 + wn * 32 + nj * 16;
      if constexpr (FOLD)
        store_tile64((cplx<double> *)Cv + (long)split * sSplit, ldc, M, N, r0, c0, dr[mi][nj], di[mi][nj],
                     cr[mi][nj], ci[mi][nj], lane);
      else
        store_tile<float>((cplx<float> *)Cv + (long)split * sSplit, ldc, M, N, r0, c0, cr[mi][nj], ci[mi][nj], lane);
    }
}

// round-to-nearest-even float -> bf16 bits
__device__ __forceinline__ unsigned short bf16_bits(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ bf16x8 ld8(const unsigned short *p) {
  uint4 u = *reinterpret_cast<const uint4 *>(p);
  return __builtin_bit_cast(bf16x8, u);
}

// Split of FP32 data into NT bf16 planes, x = hi + mid (+ lo) up to the
// rounding of the last plane: hi = bf16(x), mid = bf16(x - hi),
// lo = bf16(x - hi - mid), each residual exact in FP32.  X holds rows of
// K complex values (row stride ldx complex); plane t of row r occupies
// P[t * sp + r * ldp + j] for the 2K real values j, zeros up to ldp (a
// multiple of 64).  One thread converts 8 consecutive reals.
template <int NT>
__global__ __launch_bounds__(256) void k_split_bf16(const float *__restrict__ X, long ldx, int rows, int K, long ldp,
                                                    long sp, unsigned short *__restrict__ P) {
  const long per_row = ldp / 8;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  const long r = t / per_row, j0 = (t % per_row) * 8;
  if (r >= rows)
    return;
  const float *x = X + r * 2 * ldx;
  const long nreal = 2L * K;
  float v[8];
#pragma unroll
  for (int i = 0; i < 8; i++)
    v[i] = j0 + i < nreal ? x[j0 + i] : 0.f;
  unsigned short h[NT][8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    float rem = v[i];
#pragma unroll
    for (int q = 0; q < NT; q++) {
      h[q][i] = bf16_bits(rem);
      rem -= bf16_val(h[q][i]);
    }
  }
#pragma unroll
  for (int q = 0; q < NT; q++) {
    uint4 w;
    w.x = (unsigned)h[q][0] | ((unsigned)h[q][1] << 16);
    w.y = (unsigned)h[q][2] | ((unsigned)h[q][3] << 16);
    w.z = (unsigned)h[q][4] | ((unsigned)h[q][5] << 16);
    w.w = (unsigned)h[q][6] | ((unsigned)h[q][7] << 16);
    *reinterpret_cast<uint4 *>(P + q * sp + r * ldp + j0) = w;
  }
}

// (re, im) bf16 pairs of a B row -> the pairs (im, -re) of B'': swap the
// halves of each 32-bit word and flip the sign of the new upper half.
__device__ __forceinline__ bf16x8 qswap(uint4 b) {
  uint4 q;
  q.x = ((b.x >> 16) | (b.x << 16)) ^ 0x80000000u;
  q.y = ((b.y >> 16) | (b.y << 16)) ^ 0x80000000u;
  q.z = ((b.z >> 16) | (b.z << 16)) ^ 0x80000000u;
  q.w = ((b.w >> 16) | (b.w << 16)) ^ 0x80000000u;
  return __builtin_bit_cast(bf16x8, q);
}

// Complex conj(A) B^T from NT bf16 planes of each operand (k_split_bf16),
// on v_mfma_f32_16x16x32_bf16: every product of planes (i, j) with
// i + j < NT, the smallest terms first (NT = 2: 3 products, 16 significant
// bits per operand; NT = 3: 6 products, 24 bits, FP32-like).  A tile's
// planes go global -> registers -> LDS as 16-byte accesses, the next
// k-step's loads in flight during the current step's MFMAs; B'' for the
// imaginary part is formed from B's registers (qswap).  Planes hold
// zeros past the data, so chunks need no k bounds.  FOLD: FP32 sums folded
// into FP64 every 512 complex terms.
template <int NT, bool FOLD>
__global__ __launch_bounds__(256) void k_cgemm_bf16p(int M, int N, int kchunk, int K, int nsplit,
                                                     const unsigned short *__restrict__ A, long lda, long spa,
                                                     const unsigned short *__restrict__ B, long ldb, long spb,
                                                     cplx<double> *__restrict__ C, long ldc, long sSplit) {
  __shared__ __attribute__((aligned(16))) unsigned short As[NT][64][72];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[NT][64][72];
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
  floatx4 cr[2][2], ci[2][2];
  double dr[2][2][4], di[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      cr[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
      ci[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int v = 0; v < 4; v++)
        dr[a][b][v] = di[a][b][v] = 0.0;
    }
  const int lrow = threadIdx.x >> 2, part = threadIdx.x & 3;
  const int ra = tm * 64 + lrow, rb = tn * 64 + lrow;
  const uint4 zero = {0u, 0u, 0u, 0u};
  uint4 pa[NT][2], pb[NT][2];
  auto fetch = [&](int kb) {
    const long off = 2L * kb + part * 16;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int h = 0; h < 2; h++) {
        pa[t][h] = ra < M ? *reinterpret_cast<const uint4 *>(A + t * spa + (long)ra * lda + off + h * 8) : zero;
        pb[t][h] = rb < N ? *reinterpret_cast<const uint4 *>(B + t * spb + (long)rb * ldb + off + h * 8) : zero;
      }
  };
  if (k0 < k1)
    fetch(k0);
  int since = 0;
  const int sub = lane & 15, kq = lane >> 4;
  for (int kb = k0; kb < k1; kb += 32) {
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int h = 0; h < 2; h++) {
        *reinterpret_cast<uint4 *>(&As[t][lrow][part * 16 + h * 8]) = pa[t][h];
        *reinterpret_cast<uint4 *>(&Bs[t][lrow][part * 16 + h * 8]) = pb[t][h];
      }
    __syncthreads();
    if (kb + 32 < k1)
      fetch(kb + 32);
#pragma unroll
    for (int kk = 0; kk < 64; kk += 32) {
      const int kc = kk + kq * 8;
      bf16x8 a[NT][2];
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int mi = 0; mi < 2; mi++)
          a[t][mi] = ld8(&As[t][wm * 32 + mi * 16 + sub][kc]);
#pragma unroll
      for (int nj = 0; nj < 2; nj++) {
        const int col = wn * 32 + nj * 16 + sub;
        bf16x8 b[NT], q[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const uint4 raw = *reinterpret_cast<const uint4 *>(&Bs[t][col][kc]);
          b[t] = __builtin_bit_cast(bf16x8, raw);
          q[t] = qswap(raw);
        }
#pragma unroll
        for (int mi = 0; mi < 2; mi++) {
          // plane pairs (i, j), i + j < NT, smallest terms first
#pragma unroll
          for (int d = NT - 1; d >= 0; d--)
#pragma unroll
            for (int i = d; i >= 0; i--) {
              const int j = d - i;
              cr[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][mi], b[j], cr[mi][nj], 0, 0, 0);
              ci[mi][nj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][mi], q[j], ci[mi][nj], 0, 0, 0);
            }
        }
      }
    }
    __syncthreads();
    if (FOLD && ++since == 16) {
      since = 0;
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
#pragma unroll
          for (int v = 0; v < 4; v++) {
            dr[a][b][v] += (double)cr[a][b][v];
            di[a][b][v] += (double)ci[a][b][v];
          }
          cr[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
          ci[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
    }
  }
#pragma unroll
  for (int mi = 0; mi < 2; mi++)
#pragma unroll
    for (int nj = 0; nj < 2; nj++)
      store_tile64(C, ldc, M, N, tm * 64 + wm * 32 + mi * 16, tn * 64 + wn * 32 + nj * 16, dr[mi][nj], di[mi][nj],
                   cr[mi][nj], ci[mi][nj], lane);
}

int splitk_count(int precision, int M, int N, int K, int tune);
int launch_splitk_reduce(hipStream_t st, int M, int N, int nsplit, const void *P, bool fp32_partials,
                         cplx<double> *C, long ldc, bool accumulate);

// K chunk (complex terms) of an FP32 / 3xBF16 GEMM split nsplit ways: a
// multiple of the kernel's k-step; nsplit is reduced to the chunks needed.
int cgemm_chunk(int precision, int K, int &nsplit) {
  const int kstep = (precision == PP_BF16X3 || precision == PP_BF16X6) ? 32 : 16;
  int kchunk = (K + nsplit - 1) / std::max(1, nsplit);
  kchunk = (kchunk + kstep - 1) / kstep * kstep;
  if (kchunk == 0)
    kchunk = kstep;
  nsplit = std::max(1, (K + kchunk - 1) / kchunk);
  return kchunk;
}

// The FP32 GEMM of one batch into partial sums P ([nsplit][M][ldp], split
// stride sSplit): FP32 partials for chunks of at most 512 complex terms,
// else FP64 partials with the FP32 sums folded into FP64 every 512 terms
// (fp32_partials_for says which).  No allocation.
bool fp32_partials_for(int precision, int kchunk) { return precision == PP_FP32 && kchunk <= 512; }
int launch_cgemm32(hipStream_t st, int precision, int M, int N, int K, const void *A, long lda, const void *B,
                   long ldb, void *P, long ldp, long sSplit, int nsplit, int kchunk) {
  const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
  const bool vec = (lda % 2 == 0) && (ldb % 2 == 0) && ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0);
  const dim3 grd = gemm_grid(tiles, nsplit);
  const bool fold = !fp32_partials_for(precision, kchunk);
#define PP_CF(V, F)                                                                                              \
  hipLaunchKernelGGL((k_cgemm_f32<V, F>), grd, dim3(256), 0, st, M, N, kchunk, K, nsplit, (const float *)A, lda,    \
                     (const float *)B, ldb, P, ldp, sSplit)
  if (vec && fold)
    PP_CF(true, true);
  else if (vec)
    PP_CF(true, false);
  else if (fold)
    PP_CF(false, true);
  else
    PP_CF(false, false);
#undef PP_CF
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

// Split-bf16 GEMM of one batch (PP_BF16X3: two planes per operand, 3
// products; PP_BF16X6: three planes, 6 products): both operands split once
// into the plane buffer `planes` ((M + N) * NT * ldp shorts, ldp = 2K
// rounded up to 64), then the plane GEMM into partials P as launch_cgemm32.
int launch_cgemm_bf16(hipStream_t st, int precision, int M, int N, int K, const void *A, long lda, const void *B,
                      long ldb, unsigned short *planes, cplx<double> *P, long ldp_out, long sSplit, int nsplit,
                      int kchunk) {
  const int NT = precision == PP_BF16X6 ? 3 : 2;
  const long ldp = (2L * K + 63) / 64 * 64;
  unsigned short *PA = planes, *PB = planes + (size_t)NT * M * ldp;
  const long nA = (long)M * (ldp / 8), nB = (long)N * (ldp / 8);
  if (NT == 3) {
    hipLaunchKernelGGL(k_split_bf16<3>, dim3((unsigned)((nA + 255) / 256)), dim3(256), 0, st, (const float *)A, lda, M,
                       K, ldp, (long)M * ldp, PA);
    hipLaunchKernelGGL(k_split_bf16<3>, dim3((unsigned)((nB + 255) / 256)), dim3(256), 0, st, (const float *)B, ldb, N,
                       K, ldp, (long)N * ldp, PB);
  } else {
    hipLaunchKernelGGL(k_split_bf16<2>, dim3((unsigned)((nA + 255) / 256)), dim3(256), 0, st, (const float *)A, lda, M,
                       K, ldp, (long)M * ldp, PA);
    hipLaunchKernelGGL(k_split_bf16<2>, dim3((unsigned)((nB + 255) / 256)), dim3(256), 0, st, (const float *)B, ldb, N,
                       K, ldp, (long)N * ldp, PB);
  }
  PP_CHECK_HIP(hipGetLastError());
  const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
  const dim3 grd = gemm_grid(tiles, nsplit);
  const bool fold = kchunk > 512;
#define PP_BP(NTV, F)                                                                                            \
  hipLaunchKernelGGL((k_cgemm_bf16p<NTV, F>), grd, dim3(256), 0, st, M, N, kchunk, K, nsplit, PA, ldp, (long)M * ldp, \
                     PB, ldp, (long)N * ldp, P, ldp_out, sSplit)
  if (NT == 3) {
    if (fold)
      PP_BP(3, true);
    else
      PP_BP(3, false);
  } else {
    if (fold)
      PP_BP(2, true);
    else
      PP_BP(2, false);
  }
#undef PP_BP
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

// slot < 0: partials in gemm_ws, reduction on the context's stream.
// slot 0 / 1: partials in gemm_ws / gemm_ws2 and the reduction on the
// second stream (after the GEMM's event), so that it overlaps the next
// batch's GEMM; the GEMM first waits until the slot's previous reduction
// has read its partials.  The caller joins the second stream at the end.
int gemm_fp32_mfma(pp_ctx *ctx, int precision, int M, int N, int K, const void *A, long lda, const void *B, long ldb,
                   cplx<double> *C, long ldc, bool accumulate, int slot) {
  int nsplit = splitk_count(precision, M, N, K, ctx->tune_splitk);
  const int kchunk = cgemm_chunk(precision, K, nsplit);
  DevBuf &ws = slot == 1 ? ctx->gemm_ws2 : ctx->gemm_ws;
  PP_TRY(ws.ensure(sizeof(cplx<double>) * (size_t)nsplit * M * N));
  cplx<double> *P = (cplx<double> *)ws.p;
  if (slot >= 0)
    PP_CHECK_HIP(hipStreamWaitEvent(ctx->stream, ctx->red_done[slot], 0));
  hipEvent_t ev;
  int st;
  if (precision == PP_BF16X3 || precision == PP_BF16X6) {
    const int NT = precision == PP_BF16X6 ? 3 : 2;
    const long ldp = (2L * K + 63) / 64 * 64;
    PP_TRY(ctx->bf16_ws.ensure(sizeof(unsigned short) * (size_t)NT * (M + N) * ldp));
    tbegin(ctx, K_GEMM, &ev);
    st = launch_cgemm_bf16(ctx->stream, precision, M, N, K, A, lda, B, ldb, (unsigned short *)ctx->bf16_ws.p, P,
                           (long)N, (long)M * N, nsplit, kchunk);
  } else {
    tbegin(ctx, K_GEMM, &ev);
    st = launch_cgemm32(ctx->stream, precision, M, N, K, A, lda, B, ldb, P, (long)N, (long)M * N, nsplit, kchunk);
  }
  tend(ctx, K_GEMM, ev);
  PP_TRY(st);
  hipStream_t main = ctx->stream;
  if (slot >= 0) {
    PP_CHECK_HIP(hipEventRecord(ctx->gemm_done[slot], main));
    PP_CHECK_HIP(hipStreamWaitEvent(ctx->aux, ctx->gemm_done[slot], 0));
    ctx->stream = ctx->aux;
  }
  tbegin(ctx, K_GEMM_REDUCE, &ev);
  st = launch_splitk_reduce(ctx->stream, M, N, nsplit, P, fp32_partials_for(precision, kchunk), C, ldc, accumulate);
  tend(ctx, K_GEMM_REDUCE, ev);
  ctx->stream = main;
  PP_TRY(st);
  if (slot >= 0)
    PP_CHECK_HIP(hipEventRecord(ctx->red_done[slot], ctx->aux));
  return PP_OK;
}

}  // namespace pp
