// The band-overlap GEMM without device allocation: pp_overlap_gemm_noalloc.
//
// Callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc) get the
// overlap of pp_overlap_gemm with its split-K partial sums kept in a fixed
// device array of this library instead of the context's grown buffers.
// The array is a file-scope __device__ object: each GPU the library runs on
// holds its own copy (64 MiB) from module load until the process ends.
// Host code never passes the array's name to a kernel (on the host it is
// only a stand-in symbol): its device address is looked up once per device
// with hipGetSymbolAddress and that address is what the kernels receive.
//
// Every call on one GPU shares that GPU's array, and a call leaves its
// partial sums there; the reduction reads only what the same call wrote.
// Calls that overlap in time on one GPU (two streams, two threads) race on
// the array and must be serialised by the caller.
#include <mutex>

#include "pp_internal.h"

namespace pp {

int splitk_count(int precision, int M, int N, int K, int tune);
int cgemm_chunk(int precision, int K, int &nsplit);
int launch_cgemm32(hipStream_t st, int precision, int M, int N, int K, const void *A, long lda, const void *B,
                   long ldb, void *P, long ldp, long sSplit, int nsplit, int kchunk);
bool fp32_partials_for(int precision, int kchunk);
int launch_zgemm64(hipStream_t st, bool valu, int M, int N, int K, int atype, const void *A, long lda, int btype,
                   const void *B, long ldb, cplx<double> *dst, long dld, long sSplit, int nsplit, int kchunk);
int launch_splitk_reduce(hipStream_t st, int M, int N, int nsplit, const void *P, bool fp32_partials,
                         cplx<double> *C, long ldc, bool accumulate);

static constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices] = {nullptr};

// Device address of `scratch` on `device` (the current device), looked up
// on the first call there.
static int scratch_ptr(int device, cplx<double> **out) {
  PP_REQUIRE(device >= 0 && device < kMaxDevices, PP_ERR_ARG, "pp_overlap_gemm_noalloc: device ordinal too large");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    PP_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PP_REQUIRE(p != nullptr, PP_ERR_HIP, "pp_overlap_gemm_noalloc: no device address for the scratch array");
    g_scratch_dev[device] = p;
  }
  *out = (cplx<double> *)g_scratch_dev[device];
  return PP_OK;
}

}  // namespace pp

using namespace pp;

extern "C" {

size_t pp_noalloc_scratch_bytes(void) { return kScratchBytes; }

int pp_overlap_gemm_noalloc(pp_ctx *ctx, int nwk, int nspin, int flip, const int *nw, const void *CR_dev, long ldR,
                            int nbR, const void *CS_dev, long ldS, int nbS, void *out_dev, int precision) {
  PP_REQUIRE(ctx && nw && nwk > 0 && (nspin == 1 || nspin == 2) && nbR >= 0 && nbS >= 0, PP_ERR_ARG,
             "pp_overlap_gemm_noalloc: bad arguments");
  PP_REQUIRE(precision == PP_FP64 || precision == PP_FP32, PP_ERR_ARG,
             "pp_overlap_gemm_noalloc: precision must be PP_FP64 or PP_FP32");
  const int nkt = nwk * nspin;
  for (int k = 0; k < nkt; k++)
    PP_REQUIRE(nw[k] >= 0 && nw[k] <= ldR && nw[k] <= ldS, PP_ERR_ARG, "pp_overlap_gemm_noalloc: nw exceeds ld");
  if (nbR == 0 || nbS == 0)
    return PP_OK;
  PP_REQUIRE(CR_dev && CS_dev && out_dev, PP_ERR_ARG, "pp_overlap_gemm_noalloc: null buffer");
  // at least two K chunks of partial sums must fit the scratch array
  const size_t plane = sizeof(cplx<double>) * (size_t)nbR * nbS;
  PP_REQUIRE(2 * plane <= kScratchBytes, PP_ERR_ARG,
             "pp_overlap_gemm_noalloc: nbR * nbS = " + std::to_string((long)nbR * nbS) +
                 " exceeds the scratch limit of " + std::to_string(kScratchBytes / (2 * sizeof(cplx<double>))) +
                 " output entries per k-point");
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  cplx<double> *P = nullptr;
  PP_TRY(scratch_ptr(ctx->device, &P));
  const int max_split64 = (int)(kScratchBytes / plane);  // FP64 partial planes that fit
  hipStream_t st = ctx->stream;
  for (int k = 0; k < nkt; k++) {
    const int kR = (flip && nspin == 2) ? (k < nwk ? k + nwk : k - nwk) : k;
    const void *A = (const char *)CR_dev + (size_t)kR * nbR * ldR * sizeof(cplx<float>);
    const void *B = (const char *)CS_dev + (size_t)k * nbS * ldS * sizeof(cplx<float>);
    cplx<double> *C = (cplx<double> *)out_dev + (size_t)k * nbR * nbS;
    const int K = nw[k];
    // the split policy of pp_overlap_gemm where its partials fit the array
    int nsplit = splitk_count(precision, nbR, nbS, K, ctx->tune_splitk);
    int kchunk;
    bool fp32p = false;
    if (precision == PP_FP32) {
      kchunk = cgemm_chunk(precision, K, nsplit);
      fp32p = fp32_partials_for(precision, kchunk);
      if ((fp32p ? (size_t)nsplit * plane / 2 : (size_t)nsplit * plane) > kScratchBytes) {
        // fewer, longer chunks with in-chunk FP64 folds and FP64 partials
        nsplit = max_split64;
        kchunk = cgemm_chunk(precision, K, nsplit);
        fp32p = fp32_partials_for(precision, kchunk);
      }
      PP_TRY(launch_cgemm32(st, precision, nbR, nbS, K, A, ldR, B, ldS, P, (long)nbS, (long)nbR * nbS, nsplit,
                            kchunk));
    } else {
      nsplit = std::min(nsplit, max_split64);
      kchunk = (K + nsplit - 1) / nsplit;
      kchunk = std::max(16, (kchunk + 15) / 16 * 16);
      nsplit = std::max(1, (K + kchunk - 1) / kchunk);
      PP_TRY(launch_zgemm64(st, ctx->tune_gemm_f64_valu != 0, nbR, nbS, K, 1, A, ldR, 1, B, ldS, P, (long)nbS,
                            (long)nbR * nbS, nsplit, kchunk));
    }
    // the next k-point's GEMM overwrites the partials only after this
    // reduction has read them (one stream, in order)
    PP_TRY(launch_splitk_reduce(st, nbR, nbS, nsplit, P, fp32p, C, nbS, false));
  }
  return PP_OK;
}

}  // extern "C"
