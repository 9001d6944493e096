// Mixed-precision band-overlap GEMM on the CDNA4 matrix cores.
//
// Complex conj(A) B^T with c64 operands is two real GEMMs over the
// interleaved (re, im) storage taken as a real K' = 2K dimension:
//   Re = sum_k' A'[m][k'] B'[n][k'],   B' = (b_re, b_im) as stored,
//   Im = sum_k' A'[m][k'] B''[n][k'],  B'' = (b_im, -b_re),
// so no operand repacking is needed for the real part and B'' is the
// stored pair swapped with one sign flip.
//
// PP_FP32: v_mfma_f32_16x16x4_f32 over chunks of at most 512 complex
// terms with FP32 partial sums (or FP32 sums folded into FP64 every 512
// terms inside longer chunks), chunks summed in FP64.
// PP_BF16X3 / PP_BF16X6 / PP_FP16X3: each FP32 operand is split once per
// call into two (hi, lo) or three (hi, mid, lo) 16-bit planes; the 3 or 6
// plane products run on v_mfma_f32_16x16x32_{bf16,f16} with FP32
// accumulation over chunks of 512 complex terms, chunks summed in FP64.  The fp16 split scales each chunk of a row by a power of two first,
// so that its two planes carry 22 significant bits where bf16's carry 16.
//
// FP32 work-group: 4 wavefronts, 64x64 complex outputs, each wavefront a
// 32x32 quadrant as 2x2 16x16 MFMA tiles (real and imaginary
// accumulators); split-plane work-group: 128x128, 64x64 per wavefront.
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
    // B'' = (b_im, -b_re): the neighbour k ^ 1 (k even for re, odd for im)
    // with its sign bit flipped for odd k -- one LDS read and no branch
    // (a conditional choice of the two reads compiles to a divergent branch)
    const unsigned sgn = (unsigned)(kq & 1) << 31;
#pragma unroll
    for (int kk = 0; kk < 32; kk += 4) {
      const int kr = kk + kq;
      const float a0 = As[wm * 32 + sub][kr];
      const float a1 = As[wm * 32 + 16 + sub][kr];
#pragma unroll
      for (int nj = 0; nj < 2; nj++) {
        const int col = wn * 32 + nj * 16 + sub;
        const float b = Bs[col][kr];
        const float b2 = __builtin_bit_cast(float, __builtin_bit_cast(unsigned, Bs[col][kr ^ 1]) ^ sgn);
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

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// 2^e as a float, for -126 <= e <= 127
__device__ __forceinline__ float pow2f(int e) { return __uint_as_float((unsigned)(e + 127) << 23); }

// The NT 16-bit planes of two consecutive reals x, each plane the
// round-to-nearest-even conversion of what the planes before it leave (one
// packed conversion per pair on gfx950); w[q] holds plane q of the pair,
// the first real in its low half.  bf16: the residuals are exact in FP32.
// fp16 (NT = 2): the reals are scaled by sc = 2^e first, hi = fp16(x sc)
// and lo = fp16(x sc - hi) rounded once from the exact value (v_fma_mix),
// which equals the rounding of the FP32 residual whenever x sc is exact in
// FP32 (the residual of a normal FP32 number and its fp16 rounding is).
template <int NT, bool F16>
__device__ __forceinline__ void split_pair(f32x2 x, float sc, unsigned (&w)[NT]) {
  if constexpr (F16) {
    static_assert(NT == 2, "fp16 planes come in pairs");
    const f16x2 h = __builtin_convertvector(x * sc, f16x2);
    const f16x2 l = {(_Float16)__builtin_fmaf(x[0], sc, -(float)h[0]),
                     (_Float16)__builtin_fmaf(x[1], sc, -(float)h[1])};
    w[0] = __builtin_bit_cast(unsigned, h);
    w[1] = __builtin_bit_cast(unsigned, l);
  } else {
#pragma unroll
    for (int q = 0; q < NT; q++) {
      const bf16x2 h = __builtin_convertvector(x, bf16x2);
      w[q] = __builtin_bit_cast(unsigned, h);
      if (q + 1 < NT)
        x = x - __builtin_convertvector(h, f32x2);
    }
  }
}

// Planes are split and scaled in chunks of kPlaneChunk reals (1024 complex
// terms) of a row; rows are padded with zeros to a multiple of it.
static constexpr int kPlaneChunk = 2048;

// Split of FP32 rows into NT 16-bit planes, one wavefront per (row, chunk).
// bf16 (F16 false): x = hi + mid (+ lo) up to the rounding of the last
// plane, hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid), each
// residual exact in FP32.  fp16 (F16 true, NT = 2): the chunk is first
// scaled by 2^e, e chosen so that its largest magnitude lies in
// [2^14, 2^15) (e at most 127, which only chunks below 2^-112 reach), then
// hi = fp16(y), lo = fp16(y - hi): 11 + 11 significant bits, all of them
// normal for entries within 2^-17 of the chunk's largest, and the rest off
// by at most 2^-40 of it; e goes to E[chunk][row].  The
// rows of X (M rows, stride ldx complex) and Y (N rows) are planes' rows 0..M-1
// and M..M+N-1.  Planes are chunk-major: real j of row r in plane t is
// P[t * sp + ((j / 2048) * (M + N) + r) * 2048 + j % 2048] (2K data, zeros
// up to ldp), so that the rows a GEMM work-group reads for one K chunk form
// one contiguous block (few translation entries) instead of rows a whole
// row length apart.  Each lane converts 4 groups of 8
// consecutive reals; the chunk maximum is a wavefront reduction, so the
// kernel holds no LDS and runs beside an LDS-bound kernel on the same CUs.
template <int NT, bool F16, bool PLANES>
__global__ __launch_bounds__(64) void k_split_planes(const float *__restrict__ X, long ldx, int M,
                                                     const float *__restrict__ Y, long ldy, int N, int K, long ldp,
                                                     long sp, unsigned short *__restrict__ P, int *__restrict__ E) {
  constexpr int G = kPlaneChunk / 512;  // groups of 8 reals per lane
  const int r = blockIdx.x, chunk = blockIdx.y;
  const float *x = r < M ? X + (long)r * 2 * ldx : Y + (long)(r - M) * 2 * ldy;
  const long nreal = 2L * K;
  float v[G][8];
#pragma unroll
  for (int g = 0; g < G; g++) {
    const long j0 = (long)chunk * kPlaneChunk + g * 512 + threadIdx.x * 8;
    if (j0 + 8 <= nreal && (((uintptr_t)(x + j0)) & 15) == 0) {
      const float4 p = *reinterpret_cast<const float4 *>(x + j0), q = *reinterpret_cast<const float4 *>(x + j0 + 4);
      v[g][0] = p.x;
      v[g][1] = p.y;
      v[g][2] = p.z;
      v[g][3] = p.w;
      v[g][4] = q.x;
      v[g][5] = q.y;
      v[g][6] = q.z;
      v[g][7] = q.w;
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++)
        v[g][i] = j0 + i < nreal ? x[j0 + i] : 0.f;
    }
  }
  int e = 0;
  if constexpr (F16) {
    static_assert(NT == 2, "fp16 planes come in pairs");
    float m = 0.f;
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
      for (int i = 0; i < 8; i++)
        m = fmaxf(m, fabsf(v[g][i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
      m = fmaxf(m, __shfl_xor(m, o));
    int ex = 0;
    if (m > 0.f && m <= 3.4e38f)
      (void)frexpf(m, &ex);  // m < 2^ex <= 2^128
    e = m > 0.f && m <= 3.4e38f ? min(15 - ex, 127) : 0;
    if (threadIdx.x == 0)
      E[(long)chunk * (M + N) + r] = e;
  }
  if constexpr (!PLANES)
    return;
#pragma unroll
  for (int g = 0; g < G; g++) {
    unsigned w[4][NT];
#pragma unroll
    for (int i = 0; i < 4; i++)
      split_pair<NT, F16>(f32x2{v[g][2 * i], v[g][2 * i + 1]}, pow2f(e), w[i]);
#pragma unroll
    for (int q = 0; q < NT; q++) {
      uint4 o;
      o.x = w[0][q];
      o.y = w[1][q];
      o.z = w[2][q];
      o.w = w[3][q];
      *reinterpret_cast<uint4 *>(P + q * sp + ((long)chunk * (M + N) + r) * kPlaneChunk + g * 512 +
                                 threadIdx.x * 8) = o;
    }
  }
}

// (re, im) 16-bit pairs of a B row -> the pairs (im, -re) of B'': swap the
// halves of each 32-bit word and flip the sign of the new upper half (the
// sign is bit 15 in bf16 and fp16 alike).
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 qswap(u32x4 b) { return ((b >> 16) | (b << 16)) ^ 0x80000000u; }

template <bool F16>
__device__ __forceinline__ floatx4 mfma16(u32x4 a, u32x4 b, floatx4 c) {
  if constexpr (F16) {
    typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(halfx8, a), __builtin_bit_cast(halfx8, b), c, 0,
                                                  0, 0);
  } else {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
  }
}

// Complex conj(A) B^T from NT 16-bit planes of each operand (k_split_planes)
// on v_mfma_f32_16x16x32_{bf16,f16}: every product of planes (i, j) with
// i + j < NT, the smallest terms first (NT = 2: 3 products; NT = 3: 6).
// Work-group: 4 wavefronts, 128x128 complex outputs, each wavefront a 64x64
// quadrant as 4x4 16x16 MFMA tiles (real and imaginary accumulators, 128
// registers), so that every 16-byte operand read from LDS feeds 16 or 32
// MFMAs and the global operand traffic per output is half that of a 64x64
// tile.  A k-step of KS reals per row: its planes go global -> registers
// -> LDS as 16-byte accesses into one of two LDS buffers while the other
// buffer's step is computed, one barrier per step.  WPE wavefronts per
// SIMD: 1 (up to 512 registers, KS = 64: 144 KiB of LDS) or 2 (256
// registers with the accumulators in architectural VGPRs, KS = 32: 80 KiB,
// two work-groups per CU whose waves cover each other's LDS and memory
// waits).  Planes hold zeros past
// the data, so a chunk of CH reals needs no k bounds.  The chunk's FP32
// sums are its partial sums (row stride ldc, split stride sSplit); fp16
// hi and lo 16-bit planes of 8 consecutive reals, exactly as k_split_planes
// forms them (scale sc = 2^e for fp16)
template <bool F16>
__device__ __forceinline__ void split8(f32x4 x0, f32x4 x1, float sc, u32x4 &hi, u32x4 &lo) {
  const float v[8] = {x0[0], x0[1], x0[2], x0[3], x1[0], x1[1], x1[2], x1[3]};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    unsigned w[2];
    split_pair<2, F16>(f32x2{v[2 * i], v[2 * i + 1]}, sc, w);
    hi[i] = w[0];
    lo[i] = w[1];
  }
}

// four FP32 values of a row from x + off; WH false: those at or past
// `left` read as zeros
template <bool WH>
__device__ __forceinline__ f32x4 ld4(const float *x, long off, long left) {
  if constexpr (WH) {
    return *reinterpret_cast<const f32x4 *>(x + off);
  } else {
    f32x4 r;
#pragma unroll
    for (int i = 0; i < 4; i++)
      r[i] = off + i < left ? x[off + i] : 0.f;
    return r;
  }
}

// Operands: CVT false -- NT chunk-major planes of 16-bit values
// (k_split_planes): rows of lda = ldb = 2048 reals per chunk block, block
// stride cs, plane stride sp; CVT true (two-plane
// modes) -- the FP32 rows themselves (row stride lda, ldb complex, K
// complex terms each), split into hi / lo planes on their way into LDS with
// the same arithmetic, so that no plane buffer goes through HBM (fp16: the
// exponents EA / EB of an exponent-only k_split_planes pass).
template <int NT, int KS, bool F16, int CH, int WPE, bool CVT, int TMW>
__global__ __launch_bounds__(128 * TMW) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_cgemm_planes(
    int M, int N, int nsplit, int K, const void *__restrict__ Av, long lda, const void *__restrict__ Bv, long ldb,
    long sp, long cs, const int *__restrict__ EA, const int *__restrict__ EB, long sE, cplx<float> *__restrict__ C,
    long ldc, long sSplit) {
  static_assert(!CVT || NT == 2, "in-kernel splits form two planes");
  static_assert(TMW == 2 || TMW == 4, "2 or 4 wavefronts along M");
  constexpr int LS = KS + 8;           // LDS row stride (16-bit words): 16 bytes of padding
  constexpr int HP = KS / 32;          // 16-byte loads per (row, plane) per thread
  constexpr int TM = 64 * TMW;         // output rows (A rows) of the tile; 128 columns
  constexpr int RS = 32 * TMW;         // rows one pass of the work-group's threads stages
  constexpr int RA = TM / RS, RB = 128 / RS;  // passes over A (2) and B (2 or 1)
  __shared__ __attribute__((aligned(16))) unsigned short As[2][NT][TM][LS];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[2][NT][128][LS];
  const int tilesN = (N + 127) / 128;
  int split, tile;
  if (!gemm_block(((M + TM - 1) / TM) * tilesN, nsplit, split, tile))
    return;
  const int tm = tile / tilesN, tn = tile % tilesN;
  const long k0 = (long)split * CH;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lrow = threadIdx.x >> 2, part = threadIdx.x & 3;
  // This thread's rows of the tile: lrow + h * RS of A and of B.  A row past
  // the matrix reads the last row instead: it only feeds outputs of rows
  // (columns) past the matrix, which are not stored.
  int ra[RA], rb[RB];
#pragma unroll
  for (int h = 0; h < RA; h++)
    ra[h] = min(tm * TM + h * RS + lrow, M - 1);
#pragma unroll
  for (int h = 0; h < RB; h++)
    rb[h] = min(tn * 128 + h * RS + lrow, N - 1);
  const long kp = k0 + part * (KS / 4);
  // the scale exponents of the plane chunk this K chunk lies in
  const long ec = (long)split * CH / kPlaneChunk * sE;
  const unsigned short *pa[RA], *pb[RB];
  const float *xa[RA], *xb[RB];
  float sa[RA], sb[RB];  // fp16: the rows' chunk scales 2^e
  bool whole = true;  // the chunk lies inside the 2K reals of every row
  if constexpr (CVT) {
    const float *A = (const float *)Av, *B = (const float *)Bv;
#pragma unroll
    for (int h = 0; h < RA; h++) {
      xa[h] = A + (long)ra[h] * 2 * lda + kp;
      sa[h] = F16 ? pow2f(EA[ec + ra[h]]) : 1.f;
    }
#pragma unroll
    for (int h = 0; h < RB; h++) {
      xb[h] = B + (long)rb[h] * 2 * ldb + kp;
      sb[h] = F16 ? pow2f(EB[ec + rb[h]]) : 1.f;
    }
    whole = k0 + CH <= 2L * K;
  } else {
    // chunk-major planes: rows of lda reals within the chunk's block of
    // stride cs
    const long cb = k0 / lda * cs, kc0 = k0 % lda + part * (KS / 4);
    const unsigned short *A = (const unsigned short *)Av + cb, *B = (const unsigned short *)Bv + cb;
#pragma unroll
    for (int h = 0; h < RA; h++)
      pa[h] = A + (long)ra[h] * lda + kc0;
#pragma unroll
    for (int h = 0; h < RB; h++)
      pb[h] = B + (long)rb[h] * ldb + kc0;
  }
  const long nreal_left = 2L * K - kp;  // reals of a row from this thread's first
  // global -> register staging of one k-step: [row pass][plane][16-byte piece]
  // (planes), or [row pass][16-byte piece] of FP32 values (CVT)
  u32x4 ga[RA][NT][HP], gb[RB][NT][HP];
  f32x4 fa[RA][2 * HP], fb[RB][2 * HP];
  // WH: the chunk is whole (loads need no bounds); the main loop is
  // instantiated for both cases under one uniform branch, so that neither
  // carries a branch per load
#define PP_FETCH(step, WH)                                                                                       \
  {                                                                                                              \
    const long off_ = (long)(step) * KS;                                                                         \
    if constexpr (CVT) {                                                                                         \
      _Pragma("unroll") for (int q = 0; q < 2 * HP; q++) {                                                       \
        _Pragma("unroll") for (int h = 0; h < RA; h++) fa[h][q] = ld4<WH>(xa[h], off_ + q * 4, nreal_left);      \
        _Pragma("unroll") for (int h = 0; h < RB; h++) fb[h][q] = ld4<WH>(xb[h], off_ + q * 4, nreal_left);      \
      }                                                                                                          \
    } else {                                                                                                     \
      _Pragma("unroll") for (int t = 0; t < NT; t++) _Pragma("unroll") for (int q = 0; q < HP; q++) {           \
        _Pragma("unroll") for (int h = 0; h < RA; h++) ga[h][t][q] =                                             \
            *reinterpret_cast<const u32x4 *>(pa[h] + t * sp + off_ + q * 8);                                     \
        _Pragma("unroll") for (int h = 0; h < RB; h++) gb[h][t][q] =                                             \
            *reinterpret_cast<const u32x4 *>(pb[h] + t * sp + off_ + q * 8);                                     \
      }                                                                                                          \
    }                                                                                                            \
  }
#define PP_STASH(buf)                                                                                            \
  {                                                                                                              \
    if constexpr (CVT) {                                                                                         \
      _Pragma("unroll") for (int q = 0; q < HP; q++) {                                                           \
        _Pragma("unroll") for (int h = 0; h < RA; h++)                                                           \
            split8<F16>(fa[h][2 * q], fa[h][2 * q + 1], sa[h], ga[h][0][q], ga[h][1][q]);                        \
        _Pragma("unroll") for (int h = 0; h < RB; h++)                                                           \
            split8<F16>(fb[h][2 * q], fb[h][2 * q + 1], sb[h], gb[h][0][q], gb[h][1][q]);                        \
      }                                                                                                          \
    }                                                                                                            \
    _Pragma("unroll") for (int t = 0; t < NT; t++) _Pragma("unroll") for (int q = 0; q < HP; q++) {             \
      _Pragma("unroll") for (int h = 0; h < RA; h++)                                                             \
          *reinterpret_cast<u32x4 *>(&As[buf][t][h * RS + lrow][part * (KS / 4) + q * 8]) = ga[h][t][q];         \
      _Pragma("unroll") for (int h = 0; h < RB; h++)                                                             \
          *reinterpret_cast<u32x4 *>(&Bs[buf][t][h * RS + lrow][part * (KS / 4) + q * 8]) = gb[h][t][q];         \
    }                                                                                                            \
  }
  floatx4 cr[4][4], ci[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      cr[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
      ci[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
  constexpr int nsteps = CH / KS;
  static_assert(nsteps >= 2, "at least two k-steps per chunk");
  const int sub = lane & 15, kq = lane >> 4;
  // step s: compute from LDS buffer s & 1 while step s + 1 goes from
  // registers into the other buffer and step s + 2 is loaded (the last
  // steps re-load the chunk's last step and fill a buffer no one reads)
#define PP_KLOOP(WH)                                                                                             \
  {                                                                                                              \
    PP_FETCH(0, WH);                                                                                             \
    PP_STASH(0);                                                                                                 \
    __syncthreads();                                                                                             \
    PP_FETCH(1, WH);                                                                                             \
    for (int s = 0; s < nsteps; s++) {                                                                           \
      const int buf = s & 1;                                                                                     \
      _Pragma("unroll") for (int kk = 0; kk < KS; kk += 32) {                                                    \
        const int kc = kk + kq * 8;                                                                              \
        u32x4 a[NT][4];                                                                                          \
        _Pragma("unroll") for (int t = 0; t < NT; t++) _Pragma("unroll") for (int mi = 0; mi < 4; mi++) a[t][mi] = \
            *reinterpret_cast<const u32x4 *>(&As[buf][t][wm * 64 + mi * 16 + sub][kc]);                        \
        _Pragma("unroll") for (int nj = 0; nj < 4; nj++) {                                                       \
          const int col = wn * 64 + nj * 16 + sub;                                                               \
          u32x4 b[NT], q[NT];                                                                                    \
          _Pragma("unroll") for (int t = 0; t < NT; t++) {                                                       \
            b[t] = *reinterpret_cast<const u32x4 *>(&Bs[buf][t][col][kc]);                                       \
            q[t] = qswap(b[t]);                                                                                  \
          }                                                                                                      \
          _Pragma("unroll") for (int mi = 0; mi < 4; mi++) {                                                     \
            _Pragma("unroll") for (int d = NT - 1; d >= 0; d--) _Pragma("unroll") for (int i = d; i >= 0; i--) { \
              cr[mi][nj] = mfma16<F16>(a[i][mi], b[d - i], cr[mi][nj]);                                          \
              ci[mi][nj] = mfma16<F16>(a[i][mi], q[d - i], ci[mi][nj]);                                          \
            }                                                                                                    \
          }                                                                                                      \
        }                                                                                                        \
      }                                                                                                          \
      PP_STASH(buf ^ 1);                                                                                         \
      PP_FETCH(min(s + 2, nsteps - 1), WH);                                                                      \
      __syncthreads();                                                                                           \
    }                                                                                                            \
  }
  if (whole)
    PP_KLOOP(true)
  else
    PP_KLOOP(false)
#undef PP_KLOOP
#undef PP_FETCH
#undef PP_STASH
  C += (long)split * sSplit;
#pragma unroll
  for (int mi = 0; mi < 4; mi++) {
    const int r0 = tm * TM + wm * 64 + mi * 16 + 4 * (lane >> 4);
#pragma unroll
    for (int nj = 0; nj < 4; nj++) {
      const int j = tn * 128 + wn * 64 + nj * 16 + (lane & 15);
      if (j >= N)
        continue;
      const int eb = F16 ? EB[ec + j] : 0;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int i = r0 + v;
        if (i < M) {
          float re = cr[mi][nj][v], im = ci[mi][nj][v];
          if constexpr (F16) {
            const int e = EA[ec + i] + eb;
            re = ldexpf(re, -e);
            im = ldexpf(im, -e);
          }
          C[(long)i * ldc + j] = mk<float>(re, im);
        }
      }
    }
  }
}

int splitk_count(int precision, int M, int N, int K, int tune);
int launch_splitk_reduce(hipStream_t st, int M, int N, int nsplit, const void *P, bool fp32_partials,
                         cplx<double> *C, long ldc, bool accumulate);

// K chunk (complex terms) of an FP32 GEMM split nsplit ways: a multiple
// of the kernel's k-step; nsplit is reduced to the chunks needed.
int cgemm_chunk(int precision, int K, int &nsplit) {
  const int kstep = 16;
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

bool plane_precision(int precision) {
  return precision == PP_BF16X3 || precision == PP_BF16X6 || precision == PP_FP16X3;
}

// Layout of a plane GEMM's workspace and partial sums: ldp reals per plane
// row (2K rounded up to whole plane chunks of 1024 complex terms), NT planes
// of M + N rows, then the fp16 scale exponents [plane chunk][M + N] (with
// cvt, the GEMM splits its FP32 operands itself and the workspace holds the
// exponents alone).  The GEMM accumulates K chunks of 512 complex terms in
// FP32 (two per plane chunk), one partial-sum plane each, summed in FP64:
// FP32 sums over 1024 terms of 3 products per k-step would round more than
// the FP32 GEMM's 512-term chunks do.
void plane_layout(int precision, int M, int N, int K, bool cvt, long &ldp, int &nt, int &nsplit, size_t &ws_bytes) {
  nt = precision == PP_BF16X6 ? 3 : 2;
  ldp = (2L * K + kPlaneChunk - 1) / kPlaneChunk * kPlaneChunk;
  nsplit = (int)std::max<long>(1, ldp / (kPlaneChunk / 2));
  const long nchunks = ldp / kPlaneChunk;
  ws_bytes = cvt ? 256 : sizeof(unsigned short) * (size_t)nt * (M + N) * ldp;
  if (precision == PP_FP16X3)
    ws_bytes += sizeof(int) * (size_t)nchunks * (M + N);
}

// Whether the GEMM splits its operands itself under the context's settings
// (gemm_planes 3, two-plane modes, 16-byte aligned FP32 rows).
bool plane_cvt(pp_ctx *ctx, int precision, const void *A, long lda, const void *B, long ldb) {
  return ctx->tune_gemm_planes >= 3 && (precision == PP_FP16X3 || precision == PP_BF16X3) && lda % 2 == 0 &&
         ldb % 2 == 0 && (uintptr_t)A % 16 == 0 && (uintptr_t)B % 16 == 0;
}

// Split of both operands into planes (ws as plane_layout); with cvt the
// fp16 exponents only (nothing at all for bf16).  No allocation.
int launch_split_planes(hipStream_t st, int precision, int M, int N, int K, const void *A, long lda, const void *B,
                        long ldb, void *ws, bool cvt) {
  long ldp;
  int NT, nsplit;
  size_t wsb;
  plane_layout(precision, M, N, K, cvt, ldp, NT, nsplit, wsb);
  const long nchunks = ldp / kPlaneChunk;
  PP_REQUIRE(nchunks <= 65535, PP_ERR_ARG, "plane GEMM: K too large");
  if (cvt && precision != PP_FP16X3)
    return PP_OK;
  unsigned short *PA = (unsigned short *)ws;
  const long sp = (long)(M + N) * ldp;
  int *E = cvt ? (int *)ws : (int *)(PA + (size_t)NT * sp);
  const dim3 sg((unsigned)(M + N), (unsigned)nchunks);
  if (cvt)
    hipLaunchKernelGGL((k_split_planes<2, true, false>), sg, dim3(64), 0, st, (const float *)A, lda, M,
                       (const float *)B, ldb, N, K, ldp, sp, PA, E);
  else if (precision == PP_FP16X3)
    hipLaunchKernelGGL((k_split_planes<2, true, true>), sg, dim3(64), 0, st, (const float *)A, lda, M,
                       (const float *)B, ldb, N, K, ldp, sp, PA, E);
  else if (NT == 3)
    hipLaunchKernelGGL((k_split_planes<3, false, true>), sg, dim3(64), 0, st, (const float *)A, lda, M,
                       (const float *)B, ldb, N, K, ldp, sp, PA, nullptr);
  else
    hipLaunchKernelGGL((k_split_planes<2, false, true>), sg, dim3(64), 0, st, (const float *)A, lda, M,
                       (const float *)B, ldb, N, K, ldp, sp, PA, nullptr);
  PP_CHECK_HIP(hipGetLastError());
  return PP_OK;
}

// The plane GEMM over the planes in ws (or, with cvt, over A and B
// themselves) into FP32 partials P ([nsplit][M][ldp_out], split stride
// sSplit); variant: the context's gemm_planes (1: 128x128 tiles, one
// work-group per CU; 2, 3: two per CU; 4: 256x128 tiles of 8 wavefronts,
// with cvt).  No allocation.
int launch_plane_gemm(hipStream_t st, int precision, int M, int N, int K, const void *A, long lda, const void *B,
                      long ldb, const void *ws, cplx<float> *P, long ldp_out, long sSplit, int variant,
                      bool cvt) {
  long ldp;
  int NT, nsplit;
  size_t wsb;
  plane_layout(precision, M, N, K, cvt, ldp, NT, nsplit, wsb);
  const unsigned short *PA = (const unsigned short *)ws, *PB = PA + (size_t)M * kPlaneChunk;
  const long sp = (long)(M + N) * ldp, cs = (long)(M + N) * kPlaneChunk;
  constexpr long LP = kPlaneChunk;
  const int *E = cvt ? (const int *)ws : (const int *)(PA + (size_t)NT * sp);
  const int tiles = ((M + 127) / 128) * ((N + 127) / 128);
  const dim3 grd = gemm_grid(tiles, nsplit);
  const long sE = M + N;
  constexpr int CH = kPlaneChunk / 2;
#define PP_PG(NTV, KSV, F, WPEV, CVTV, AA, LA, BB, LB, EE)                                                      \
  hipLaunchKernelGGL((k_cgemm_planes<NTV, KSV, F, CH, WPEV, CVTV, 2>), grd, dim3(256), 0, st, M, N, nsplit, K, AA, LA,  \
                     BB, LB, sp, cs, EE, EE ? EE + M : nullptr, sE, P, ldp_out, sSplit)
#define PP_PG4(F, AA, LA, BB, LB, EE)                                                                          \
  hipLaunchKernelGGL((k_cgemm_planes<2, 32, F, CH, 2, true, 4>), grd4, dim3(512), 0, st, M, N, nsplit, K, AA, LA, BB, \
                     LB, sp, cs, EE, EE ? EE + M : nullptr, sE, P, ldp_out, sSplit)
  const dim3 grd4 = gemm_grid(((M + 255) / 256) * ((N + 127) / 128), nsplit);
  const int *noE = nullptr;
  if (cvt && variant >= 4 && precision == PP_FP16X3)
    PP_PG4(true, A, lda, B, ldb, E);
  else if (cvt && variant >= 4)
    PP_PG4(false, A, lda, B, ldb, noE);
  else if (cvt && precision == PP_FP16X3)
    PP_PG(2, 32, true, 2, true, A, lda, B, ldb, E);
  else if (cvt)
    PP_PG(2, 32, false, 2, true, A, lda, B, ldb, noE);
  else if (precision == PP_FP16X3 && variant >= 2)
    PP_PG(2, 32, true, 2, false, PA, LP, PB, LP, E);
  else if (precision == PP_FP16X3)
    PP_PG(2, 64, true, 1, false, PA, LP, PB, LP, E);
  else if (NT == 2 && variant >= 2)
    PP_PG(2, 32, false, 2, false, PA, LP, PB, LP, noE);
  else if (NT == 3)
    PP_PG(3, 32, false, 1, false, PA, LP, PB, LP, noE);
  else
    PP_PG(2, 64, false, 1, false, PA, LP, PB, LP, noE);
#undef PP_PG
#undef PP_PG4
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
  const bool planes = plane_precision(precision);
  const bool cvt = planes && plane_cvt(ctx, precision, A, lda, B, ldb);
  int nsplit, kchunk = 0;
  long ldp;
  int nt;
  size_t wsb = 0;
  if (planes) {
    plane_layout(precision, M, N, K, cvt, ldp, nt, nsplit, wsb);
  } else {
    nsplit = splitk_count(precision, M, N, K, ctx->tune_splitk);
    kchunk = cgemm_chunk(precision, K, nsplit);
  }
  const bool fp32p = planes || fp32_partials_for(precision, kchunk);
  DevBuf &ws = slot == 1 ? ctx->gemm_ws2 : ctx->gemm_ws;
  PP_TRY(ws.ensure((fp32p ? sizeof(cplx<float>) : sizeof(cplx<double>)) * (size_t)nsplit * M * N));
  void *P = ws.p;
  if (slot >= 0)
    PP_CHECK_HIP(hipStreamWaitEvent(ctx->stream, ctx->red_done[slot], 0));
  hipEvent_t ev;
  int st;
  if (planes) {
    PP_TRY(ctx->bf16_ws.ensure(wsb));
    if (!cvt || precision == PP_FP16X3) {
      tbegin(ctx, K_GEMM_SPLIT, &ev);
      st = launch_split_planes(ctx->stream, precision, M, N, K, A, lda, B, ldb, ctx->bf16_ws.p, cvt);
      tend(ctx, K_GEMM_SPLIT, ev);
      PP_TRY(st);
    }
    tbegin(ctx, K_GEMM, &ev);
    st = launch_plane_gemm(ctx->stream, precision, M, N, K, A, lda, B, ldb, ctx->bf16_ws.p, (cplx<float> *)P,
                           (long)N, (long)M * N, ctx->tune_gemm_planes, cvt);
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
  st = launch_splitk_reduce(ctx->stream, M, N, nsplit, P, fp32p, C, ldc, accumulate);
  tend(ctx, K_GEMM_REDUCE, ev);
  ctx->stream = main;
  PP_TRY(st);
  if (slot >= 0)
    PP_CHECK_HIP(hipEventRecord(ctx->red_done[slot], ctx->aux));
  return PP_OK;
}

// A batch of split-plane GEMMs (one per k-point) as a two-stream pipeline:
// the operand split of batch b + 1 runs on the second stream while batch
// b's GEMM runs on the context's stream (two plane buffers), and the split-K
// reduction of batch b follows it on the second stream (two partials
// buffers).  The split holds no LDS, so its wavefronts run on the CUs the
// LDS-bound GEMM fills.  The second stream's work is ordered
//   split(0), split(1), reduce(0), split(2), reduce(1), ...
// and every buffer hand-over is an event: a plane buffer is split into
// after the GEMM two batches back read it (gemm_done), a GEMM starts after
// its split (split_done) and after the reduction two batches back read its
// partials (red_done).  The caller's stream joins the second stream at the
// end (zgemm_nt).
int plane_gemm_pipeline(pp_ctx *ctx, int precision, int M, int N, int K, const void *A, long lda, long sA,
                        const void *B, long ldb, long sB, cplx<double> *C, long ldc, long sC, int nbatch,
                        const int *amap, const int *bmap, const int *klen, bool accumulate) {
  size_t wsmax = 0, pmax = 0;
  const bool cvt = plane_cvt(ctx, precision, A, lda, B, ldb) && sA % 2 == 0 && sB % 2 == 0;
  for (int b = 0; b < nbatch; b++) {
    long ldp;
    int nt, nsplit;
    size_t wsb;
    plane_layout(precision, M, N, klen ? klen[b] : K, cvt, ldp, nt, nsplit, wsb);
    wsmax = std::max(wsmax, wsb);
    pmax = std::max(pmax, sizeof(cplx<float>) * (size_t)nsplit * M * N);
  }
  PP_TRY(ctx->bf16_ws.ensure(wsmax));
  PP_TRY(ctx->bf16_ws2.ensure(wsmax));
  PP_TRY(ctx->gemm_ws.ensure(pmax));
  PP_TRY(ctx->gemm_ws2.ensure(pmax));
  hipStream_t main = ctx->stream, aux = ctx->aux;
  for (int i = 0; i < 2; i++) {
    if (!ctx->split_done[i])
      PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->split_done[i], hipEventDisableTiming));
    PP_CHECK_HIP(hipEventRecord(ctx->gemm_done[i], main));  // no GEMM reads a plane buffer yet
  }
  void *planes[2] = {ctx->bf16_ws.p, ctx->bf16_ws2.p};
  void *parts[2] = {ctx->gemm_ws.p, ctx->gemm_ws2.p};
  auto operand = [&](int b, const void *&pa, const void *&pb, int &kk) {
    const int ia = amap ? amap[b] : b, ib = bmap ? bmap[b] : b;
    kk = klen ? klen[b] : K;
    pa = (const char *)A + (size_t)ia * sA * sizeof(cplx<float>);
    pb = (const char *)B + (size_t)ib * sB * sizeof(cplx<float>);
  };
  auto split = [&](int b) -> int {
    const int slot = b & 1;
    const void *pa, *pb;
    int kk;
    operand(b, pa, pb, kk);
    PP_CHECK_HIP(hipStreamWaitEvent(aux, ctx->gemm_done[slot], 0));
    if (cvt && precision != PP_FP16X3) {  // the GEMM splits bf16 rows itself: nothing to prepare
      PP_CHECK_HIP(hipEventRecord(ctx->split_done[slot], aux));
      return PP_OK;
    }
    hipEvent_t ev;
    ctx->stream = aux;
    tbegin(ctx, K_GEMM_SPLIT, &ev);
    const int st = launch_split_planes(aux, precision, M, N, kk, pa, lda, pb, ldb, planes[slot], cvt);
    tend(ctx, K_GEMM_SPLIT, ev);
    ctx->stream = main;
    PP_TRY(st);
    PP_CHECK_HIP(hipEventRecord(ctx->split_done[slot], aux));
    return PP_OK;
  };
  PP_TRY(split(0));
  for (int b = 0; b < nbatch; b++) {
    const int slot = b & 1;
    if (b + 1 < nbatch)
      PP_TRY(split(b + 1));
    const void *pa, *pb;
    int kk;
    operand(b, pa, pb, kk);
    long ldp;
    int nt, nsplit;
    size_t wsb;
    plane_layout(precision, M, N, kk, cvt, ldp, nt, nsplit, wsb);
    PP_CHECK_HIP(hipStreamWaitEvent(main, ctx->split_done[slot], 0));
    PP_CHECK_HIP(hipStreamWaitEvent(main, ctx->red_done[slot], 0));
    hipEvent_t ev;
    tbegin(ctx, K_GEMM, &ev);
    int st = launch_plane_gemm(main, precision, M, N, kk, pa, lda, pb, ldb, planes[slot],
                               (cplx<float> *)parts[slot], (long)N, (long)M * N, ctx->tune_gemm_planes, cvt);
    tend(ctx, K_GEMM, ev);
    PP_TRY(st);
    PP_CHECK_HIP(hipEventRecord(ctx->gemm_done[slot], main));
    PP_CHECK_HIP(hipStreamWaitEvent(aux, ctx->gemm_done[slot], 0));
    ctx->stream = aux;
    tbegin(ctx, K_GEMM_REDUCE, &ev);
    st = launch_splitk_reduce(aux, M, N, nsplit, parts[slot], true, C + (size_t)b * sC, ldc, accumulate);
    tend(ctx, K_GEMM_REDUCE, ev);
    ctx->stream = main;
    PP_TRY(st);
    PP_CHECK_HIP(hipEventRecord(ctx->red_done[slot], aux));
  }
  return PP_OK;
}

}  // namespace pp
