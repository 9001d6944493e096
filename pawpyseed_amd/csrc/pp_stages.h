// LDS-resident FFT stages shared by the FFT passes and the fused x pass
// with sphere sums.  A line tile of n elements x NBW bands lives in LDS as
// buf[e * NBW + b].
#pragma once
#include "pp_dft.h"
#include "pp_internal.h"

namespace pp {

// Signed frequency of wrapped index i on an axis of length n.
__device__ __forceinline__ int sgn_idx(int i, int n) { return i < n - n / 2 ? i : i - n; }

// In-place Stockham stage: every thread keeps the R inputs of up to MAXB
// butterflies in registers across one barrier, so a single LDS buffer
// serves as input and output (half the LDS of the ping-pong form, twice
// the work-groups per CU).
//
// Thread t of NT owns band slot b = t % NBW and butterflies j = t / NBW +
// q * (NT / NBW); j / p comes from the multiplier pmag = ceil(2^32 / p)
// (exact for j < 2^16), and the twiddles are read from the LDS copy of
// the axis table.
template <typename T, int NBW, int R, int MAXB, int NT = 256>
__device__ __forceinline__ void stage_ip(cplx<T> *__restrict__ buf, const cplx<T> *__restrict__ tws, unsigned p,
                                         unsigned pmag, unsigned Tn) {
  constexpr unsigned JS = NT / NBW;
  const unsigned b = threadIdx.x % NBW, j0 = threadIdx.x / NBW;
  cplx<T> v[MAXB][R];
  unsigned o[MAXB];
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    const unsigned j = j0 + q * JS;
    if (j < Tn) {
      const unsigned h = p > 1 ? __umulhi(j, pmag) : j;
      const unsigned k = j - h * p;
      o[q] = h * p * R + k;
      v[q][0] = buf[j * NBW + b];
#pragma unroll
      for (int r = 1; r < R; r++)
        v[q][r] = buf[(j + r * Tn) * NBW + b] * tws[k * (R - 1) + (r - 1)];
      Dft<T, R>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    const unsigned j = j0 + q * JS;
    if (j < Tn) {
#pragma unroll
      for (int r = 0; r < R; r++)
        buf[(o[q] + r * p) * NBW + b] = v[q][r];
    }
  }
  __syncthreads();
}

// The stage list of an axis of length n as the host plan factors it
// (radix 4 while it divides, then the primes in increasing order), with
// each stage's p (product of the earlier radices) and its offset in the
// axis twiddle table, which holds p (R - 1) twiddles then R roots per
// stage.
struct FftFactors {
  int nst, r[12], p[12], twoff[12];
};
constexpr FftFactors fft_factors(int n) {
  FftFactors f{};
  int m = n;
  while (m % 4 == 0 && f.nst < 12) {
    f.r[f.nst++] = 4;
    m /= 4;
  }
  for (int q = 2; m > 1 && q <= m && f.nst < 12; q++)
    while (m % q == 0 && f.nst < 12) {
      f.r[f.nst++] = q;
      m /= q;
    }
  int p = 1, off = 0;
  for (int s = 0; s < f.nst; s++) {
    f.p[s] = p;
    f.twoff[s] = off;
    off += p * (f.r[s] - 1) + f.r[s];
    p *= f.r[s];
  }
  return f;
}

// stage_ip for a stage known at compile time (radix R, p = P, TN = n / R
// butterflies per band): every LDS and twiddle offset is a constant off one
// base per butterfly, and the first stage (p = 1, all twiddles 1) skips
// its twiddle products.
template <typename T, int NBW, int R, int P, int TN, int NT>
__device__ __forceinline__ void stage_fixed(cplx<T> *__restrict__ buf, const cplx<T> *__restrict__ tws) {
  constexpr unsigned JS = NT / NBW, MB = (TN + JS - 1) / JS;
  const unsigned b = threadIdx.x % NBW, j0 = threadIdx.x / NBW;
  cplx<T> v[MB][R];
  unsigned o[MB];
#pragma unroll
  for (unsigned q = 0; q < MB; q++) {
    const unsigned j = j0 + q * JS;
    if (TN % JS == 0 || q + 1 < MB || j < (unsigned)TN) {
      const unsigned k = j % P;
      o[q] = (j / P) * P * R + k;
      v[q][0] = buf[j * NBW + b];
#pragma unroll
      for (int r = 1; r < R; r++) {
        v[q][r] = buf[(j + r * TN) * NBW + b];
        if constexpr (P > 1)
          v[q][r] = v[q][r] * tws[k * (R - 1) + (r - 1)];
      }
      Dft<T, R>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (unsigned q = 0; q < MB; q++) {
    const unsigned j = j0 + q * JS;
    if (TN % JS == 0 || q + 1 < MB || j < (unsigned)TN) {
#pragma unroll
      for (int r = 0; r < R; r++)
        buf[(o[q] + r * P) * NBW + b] = v[q][r];
    }
  }
  __syncthreads();
}

// All stages of a line of compile-time length N (stage S onwards).
template <typename T, int NBW, int N, int NT, int S = 0>
__device__ __forceinline__ void stages_fixed(cplx<T> *__restrict__ buf, const cplx<T> *__restrict__ twl) {
  constexpr FftFactors f = fft_factors(N);
  if constexpr (S < f.nst) {
    stage_fixed<T, NBW, f.r[S], f.p[S], N / f.r[S], NT>(buf, twl + f.twoff[S]);
    stages_fixed<T, NBW, N, NT, S + 1>(buf, twl);
  }
}

// Copy the twiddle/root table of one axis into LDS (visible after the
// caller's next barrier).
template <typename T, int NT = 256>
__device__ __forceinline__ void stage_twiddles(cplx<T> *__restrict__ dst, const cplx<T> *__restrict__ tw, int len) {
  for (int i = threadIdx.x; i < len; i += NT)
    dst[i] = tw[i];
}

// loads per thread issued together when a pass stages its line in LDS
constexpr int FFT_LB = 8;
// Minimum wavefronts per SIMD the in-place pass kernels are compiled for:
// 8 (at most 64 registers) for the FP32 radix-2..5 kernels with up to four
// butterflies per thread, where the extra resident work-groups keep more
// line loads in flight (x pass 8.4 -> 7.8 ms per step at 108^3; the
// 8-band-tile kernels spill a few registers); no bound for the others,
// which would spill.
template <typename T, int MAXB, bool COMP> struct FftWaves {
  static constexpr int value = (!COMP && sizeof(T) == 4 && MAXB <= 4) ? 8 : 1;
};
// The same for a pass of compile-time length N (0: the generic pass): FP32
// 16-band tiles of lines up to 192 at 8 wavefronts per SIMD; longer lines
// (up to seven butterflies per thread in a radix-2 stage of 216) and wider
// tiles at 4.
template <typename T, int MAXB, int N, int NBW> struct PassWaves {
  static constexpr int value = N == 0 ? FftWaves<T, MAXB, false>::value
                               : sizeof(T) == 4 ? (N <= 192 && NBW <= 16 ? 8 : 4) : 1;
};

// All stages of a line in LDS (radices up to RMAX; the caller guarantees
// the plan has no larger one).  Composite radices are compiled only into
// the COMP instantiation with MAXB <= 2 (the host selects the composite
// factorization only when every stage needs at most two butterflies per
// thread); keeping them out of the other instantiations keeps their
// register allocation that of radices 2-5.
template <typename T, int NBW, int MAXB, int NT, bool COMP = false, int RMAX = 5>
__device__ __forceinline__ void stages_ip(cplx<T> *__restrict__ buf, const cplx<T> *__restrict__ twl,
                                          const AxisPlan &pl, int n) {
  for (int s = 0; s < pl.nst; s++) {
    const unsigned R = pl.radix[s], p = pl.pprev[s], pm = pl.pmag[s], Tn = n / R;
    const cplx<T> *tws = twl + pl.twoff[s];
    switch (R) {
    case 2:
      stage_ip<T, NBW, 2, MAXB, NT>(buf, tws, p, pm, Tn);
      break;
    case 3:
      stage_ip<T, NBW, 3, MAXB, NT>(buf, tws, p, pm, Tn);
      break;
    case 4:
      stage_ip<T, NBW, 4, MAXB, NT>(buf, tws, p, pm, Tn);
      break;
    case 5:
      if constexpr (RMAX >= 5)
        stage_ip<T, NBW, 5, MAXB, NT>(buf, tws, p, pm, Tn);
      break;
    default:
      if constexpr (COMP && MAXB <= 2) {
        switch (R) {
        case 6:
          stage_ip<T, NBW, 6, MAXB, NT>(buf, tws, p, pm, Tn);
          break;
        case 8:
          stage_ip<T, NBW, 8, MAXB, NT>(buf, tws, p, pm, Tn);
          break;
        case 9:
          stage_ip<T, NBW, 9, MAXB, NT>(buf, tws, p, pm, Tn);
          break;
        case 12:
          stage_ip<T, NBW, 12, MAXB, NT>(buf, tws, p, pm, Tn);
          break;
        case 16:
          stage_ip<T, NBW, 16, MAXB, NT>(buf, tws, p, pm, Tn);
          break;
        }
      }
    }
  }
}

}  // namespace pp
