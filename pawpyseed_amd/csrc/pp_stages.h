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

// Copy the twiddle/root table of one axis into LDS (visible after the
// caller's next barrier).
template <typename T, int NT = 256>
__device__ __forceinline__ void stage_twiddles(cplx<T> *__restrict__ dst, const cplx<T> *__restrict__ tw, int len) {
  for (int i0 = threadIdx.x; i0 < len; i0 += 4 * NT) {
    cplx<T> v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (i0 + u * NT < len)
        v[u] = tw[i0 + u * NT];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (i0 + u * NT < len)
        dst[i0 + u * NT] = v[u];
  }
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
