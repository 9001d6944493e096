// Internal host-side interfaces shared by the library's translation units.
#pragma once
#include <functional>
#include <map>
#include <string>
#include <complex>
#include <vector>

#include "pp_common.h"

namespace pp {

// Kernel classes timed by the context's optional event timer.
enum KClass {
  K_ZERO = 0,
  K_SCATTER,
  K_FFT_Z,
  K_FFT_Y,
  K_FFT_X,
  K_PHASE,
  K_SPHERE_SUM,
  K_SPHERE_REDUCE,
  K_TABLE,
  K_GEMM,
  K_GEMM_REDUCE,
  K_AUG_PREP,
  K_AUG_GEMM,
  K_MISC,
  K_FFT_X_SPHERE,
  K_RS_AUG,
  K_RS_EMIT,
  K_GEMM_SPLIT,
  K_NCLASS
};
const char *kclass_name(int c);

// Device buffer owned by the library (grows, never shrinks).
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need);
  void release();
  ~DevBuf() { release(); }
};

// Length-n 1D FFT plan of one axis: radices and twiddle offsets.
struct AxisPlan {
  int n = 0, nst = 0;
  int radix[12];
  int pprev[12];
  int twoff[12];
  unsigned pmag[12];  // ceil(2^32 / pprev): j / pprev == umulhi(j, pmag) for j < 2^16 (pprev > 1)
  int twlen = 0;      // twiddle + root entries of this axis (both factorizations)
  // second factorization with composite radices (6, 8, 9, 12, 16): fewer
  // stages for the in-place passes; nstc = 0 when there is none
  int nstc = 0;
  int radixc[12], pprevc[12], twoffc[12];
  unsigned pmagc[12];
};

// Plane-wave sphere bounds of one k-point in signed grid indices (device
// pointers; xb == nullptr means "no pruning").
struct SphereBounds {
  const int *zb, *yb, *xb;
  int n0, n1, n2;
  // populated lines of the z and y passes (packed l1 * n2 + l2 of the
  // pass), nullptr = every line
  const int *zl = nullptr, *yl = nullptr;
  int nzl = 0, nyl = 0;
  // x pass: points of x-line y * n2 + z that some projector sphere holds,
  // bit x of word xm[line * xmw + x / 32]; only those are written
  // (nullptr = all)
  const unsigned *xm = nullptr;
  int xmw = 0;
};

// Pointers of the work-item descriptors below are typed as global memory:
// a generic pointer read from device memory would make the compiler emit
// flat accesses through it (which also wait on the LDS counter).
#define PP_G __attribute__((address_space(1)))

// One work item of a batched setup over several structures and k-points
// (pp_setup_projections_multi): a chunk of up to NB bands of one k-point
// of one structure, with everything its kernels read.  The item kernels
// take an array of these, one per blockIdx.y.
struct ChunkItem {
  PP_G void *grid;             // band-fastest chunk buffer of the item's slot
  const PP_G cplx<float> *C;   // coefficients of the chunk's first band, row stride ldc
  long ldc;
  int nbc, nw;                 // bands in the chunk, plane waves of the k-point
  const PP_G int *gidx;        // wrapped grid index of every G of the k-point
  // plane-wave sphere bounds and line lists of the k-point (SphereBounds)
  const PP_G int *zb, *yb, *xb, *zl, *yl;
  const PP_G unsigned *xm;
  int n0, n1, n2, nzl, nyl, xmw;
  const PP_G void *phase;      // Bloch phase of every sphere point of (structure, k-point)
  // sphere plan of the structure
  const PP_G int *site_pt, *site_np, *proj_off, *proj_m, *pidx;
  const PP_G long *tab_off;
  const PP_G void *table;
  int nsites, nparts;
  PP_G void *partial;          // sphere-sum partial sums of the slot
  PP_G cplx<double> *out;      // projections of the chunk's first band, row stride ptot
  long ptot;
};

struct FftPlan {
  int fftg[3] = {0, 0, 0};
  AxisPlan ax[3];
  DevBuf tw64, tw32;  // concatenated per-axis twiddle tables (c128 / c64)
  int twbase[3] = {0, 0, 0};
};

struct Timer {
  bool on = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> pool;
  double ms[K_NCLASS] = {0};
  long launches[K_NCLASS] = {0};
};

}  // namespace pp

struct pp_ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  int tune_fft_bands = 0, tune_sphere_parts = 0, tune_splitk = 0, tune_fft_tile = 0, tune_fft_prune = 0;
  int tune_fft_pingpong = 0;
  int tune_fft_persistent = 0;  // 0 off, 1 on (2048 work-groups), > 1 work-group count
  int tune_gemm_f64_valu = 0;   // 1: FP64 GEMM on the VALU instead of the matrix cores
  int tune_streams = 2;         // band chunks in flight (1 or 2 streams)
  int tune_sphere_unroll = 8;   // sphere-sum gathers in flight per wavefront (4 or 8)
  int tune_fft_threads = 256;   // work-group size of the in-place FFT passes with 32-band tiles (256 or 512)
  int tune_fft_radix = 0;       // 0: radices 2-5, 1: composite radices where they fit
  int tune_sphere_nt = 1;       // 1: sphere-sum gathers with non-temporal loads
  int tune_fft_nt = 2;          // FFT passes: 1 non-temporal stores, 2 non-temporal loads and stores
  int tune_fft_xmask = 0;        // 1: the x pass writes only points a projector sphere holds
  int tune_fft_xcd = 0;         // 1: FFT pass work-groups in XCD-contiguous line ranges
  int tune_fft_fixed = 1;       // 1: compile-time passes for the axis lengths that have one
  int tune_fused = 0;           // 1: x pass fused with the sphere sums where the plan allows
  int tune_rs_kernel = 2;       // real-space augmentation: 0 LDS coefficients, 1 registers, 2 tiled
  int tune_rs_unroll = 4;       // real-space augmentation: points of a wavefront in flight (1, 2, 4)
  int tune_rs_prune = 1;        // real-space FFT passes restricted to the plane-wave sphere's lines
  int tune_multi_items = 8;     // work items per launch of the multi-structure setup
  int tune_gemm_planes = 3;     // two-plane split GEMMs: 1, 2 work-groups per CU; 3: as 2, operands split in the GEMM
  hipStream_t aux = nullptr;    // second stream of the chunk pipeline (default priority)
  hipEvent_t fork = nullptr, join = nullptr;  // hand-over between the two streams
  int slot = 0;                 // chunk buffer set in use (0 or 1)
  pp::DevBuf grid, grid2;       // band-fastest FFT chunk, one per slot
  pp::DevBuf partial, partial2; // sphere-sum partials, one per slot
  pp::DevBuf gidx;      // wrapped grid index per G
  pp::DevBuf phase;     // per-point Bloch phase of every k
  pp::DevBuf gemm_ws;   // split-K partials
  pp::DevBuf gemm_ws2;  // second partials buffer (reduction of one batch overlaps the next GEMM)
  pp::DevBuf bf16_ws;   // 16-bit planes of the split-plane GEMMs' operands
  pp::DevBuf bf16_ws2;  // second plane buffer (the split of one batch overlaps the previous batch's GEMM)
  hipEvent_t gemm_done[2] = {nullptr, nullptr};  // GEMM of a partials slot finished
  hipEvent_t red_done[2] = {nullptr, nullptr};   // reduction of a partials slot finished
  hipEvent_t split_done[2] = {nullptr, nullptr}; // operand split into a plane buffer finished
  pp::DevBuf aug_ws;    // gathered / dO-applied projections
  // Index tables and dO coefficients of the last augmentation call, on the
  // device and (as the content key) on the host; uploads of changed tables
  // go through a pinned staging buffer.
  pp::DevBuf aug_tab;
  std::vector<int> aug_ints;
  std::vector<char> aug_coefs;
  void *aug_pinned = nullptr;
  size_t aug_pinned_bytes = 0;
  hipEvent_t aug_upload = nullptr;  // the last upload out of aug_pinned finished
  hipEvent_t aug_done = nullptr;    // the last call's kernels (readers of aug_tab / aug_ws) finished
  bool aug_pending = false;
  pp::DevBuf bounds;    // plane-wave sphere bounds per k
  pp::DevBuf lines;     // populated z / y pass lines per k
  pp::DevBuf m_items, m_grids, m_partials;  // work items, chunk grids, partial sums of the multi-structure setup
  std::vector<int> lines_host, bounds_host;
  std::map<long long, pp::FftPlan *> fft_plans;
  pp::Timer timer;
};

// A captured setup pipeline: the graph, and what it was captured against
// (the build generation of every sphere plan it reads).  The graph owns the
// buffers the host-side preparation fills from the call's G vectors and
// k-points (grid indices, Bloch phases, plane-wave bounds, line lists, work
// items), so eager calls with other inputs on the same context leave them
// alone.  A single-structure graph uses the context's chunk grids and
// partial sums, scratch whose contents no launch carries over, so only
// their addresses must stay those of the capture (bufs); a multi-structure
// graph owns those too.
struct pp_graph {
  pp_ctx *ctx = nullptr;
  std::vector<std::pair<pp_sphere *, long>> plans;
  std::vector<std::pair<void *, size_t>> bufs;
  bool multi = false;
  pp::DevBuf gidx, phase, bounds, lines;
  std::vector<int> lines_host, bounds_host;
  pp::DevBuf items, grids, partials;
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
};

namespace pp {
// timing helpers: call begin before a launch and end after it
void tbegin(pp_ctx *ctx, int cls, hipEvent_t *start);
void tend(pp_ctx *ctx, int cls, hipEvent_t start);

int get_fft_plan(pp_ctx *ctx, const int *fftg, FftPlan **out);

// Fill a device buffer with zeros (a kernel; no memset node).
int zero_async(pp_ctx *ctx, void *p, size_t bytes);

// Band-fastest grid chunk pipeline pieces (pp_fft.hip).
int compute_gidx(pp_ctx *ctx, const int *fftg, int nw, const int *G_dev, int *gidx_dev);
// Scatter and backward FFT of a band chunk; with last_pass false the x
// pass is left to the caller (the fused x pass with sphere sums).
template <typename T>
int fft_chunk(pp_ctx *ctx, FftPlan *pl, cplx<T> *grid, int NB, int nw, const int *gidx_dev,
              const cplx<float> *C_dev, long ldc, int nbc, const SphereBounds *bounds, bool last_pass = true);
// twiddle / root table of one axis in the precision of T
template <typename T> const cplx<T> *fft_twiddles(FftPlan *pl, int axis);
template <typename T> int fft_passes(pp_ctx *ctx, FftPlan *pl, cplx<T> *grid, int NB);
template <typename T>
int gather_conj(pp_ctx *ctx, const cplx<T> *grid, int NB, int nbc, const int *gidx, int nw, double scl,
                cplx<double> *out, long ldo);
// bnd_dev: 2*n0*n1 + 2*n0 + 3 ints; the last is set when the grid aliases a G.
int sphere_bounds(pp_ctx *ctx, const int *fftg, int nw, const int *G_dev, int *bnd_dev);
// True when the G vectors fill every z-line interval of the host copy hb
// of those bounds (sum of interval lengths == nw), as a cutoff sphere does;
// the pruned passes rely on it, since they load no zero fill.
bool bounds_filled(const int *hb, const int *fftg, long nw);
int grid_to_bandmajor(pp_ctx *ctx, int precision, const void *grid, int NB, int nbc, long npts,
                      double scl, void *out);
double det3(const double *m);
// Real-space states (pp_realspace.hip): x *= e^{-2 pi i k.r} on the grid
// fftg; out[b] = scale sum_r conj(R[b * ldr + r]) S[r] (synchronises).
int rs_remove_phase(pp_ctx *ctx, cplx<double> *x, const int *fftg, const double *kf);
int rs_dots(pp_ctx *ctx, const cplx<double> *R, long ldr, int nb, const cplx<double> *S, long n, double scale,
            cplx<double> *out);
void spline_coeff_host(const double *x, const double *y, int n, double *b, double *c, double *d);
int smooth_diffwave_host(int n, const double *wave_grid, const double *diffwave, int l, int lmax, double grid_encut,
                         int np, double *sdw, double *kgrid, double *kwave);
double wigner3j_host(int j1, int j2, int j3, int m1, int m2, int m3);
std::complex<double> offsite_overlap_host(const double *dcoord, const double *k1, const double *f1,
                                          const double *const *s1, int n1, const double *k2, const double *f2,
                                          const double *const *s2, int n2, int l1, int m1, int l2, int m2);

// Work-group -> (split-K chunk, output tile) of the overlap GEMMs.  The
// hardware hands work-groups round-robin to the 8 XCDs in launch order, so
// block b runs on XCD b % 8.  The (chunk, tile) items, chunk-major, are cut
// into 8 contiguous ranges and XCD x takes range x: with many chunks an XCD
// holds whole chunks, so its L2 keeps a chunk's A and B rows for all of the
// chunk's tiles; with fewer chunks than XCDs the tiles of a chunk spread
// over all of them (consecutive tiles share their A rows).  Grid:
// gemm_grid(tiles, nsplit) work-groups, of which those past the items do
// nothing.
__device__ __forceinline__ bool gemm_block(int ntiles, int nsplit, int &split, int &tile) {
  const long total = (long)ntiles * nsplit, per = (total + 7) / 8;
  const long r = (long)(blockIdx.x >> 3), i = (long)(blockIdx.x & 7) * per + r;
  if (r >= per || i >= total)
    return false;
  split = (int)(i / ntiles);
  tile = (int)(i % ntiles);
  return true;
}
inline dim3 gemm_grid(int ntiles, int nsplit) {
  return dim3((unsigned)(((long)ntiles * nsplit + 7) / 8 * 8));
}

// Autotuner (pp_api.hip): coordinate descent over the named parameters,
// timing run() on the context's stream.
struct TuneCand {
  const char *name;
  std::vector<int> values;
};
int autotune_core(pp_ctx *ctx, const std::vector<TuneCand> &cands, const std::function<int()> &run, int repeats,
                  char *report, int report_len);

// Complex GEMM out[b][m][n] (+)= sum_k conj(A[b][m][k]) B[b][n][k] (pp_gemm.hip).
// type codes: 0 = c128, 1 = c64.  Batch strides in elements; amap/bmap
// optional per-batch index remaps (host arrays of length nbatch).
int zgemm_nt(pp_ctx *ctx, int precision, int M, int N, int K, int atype, const void *A, long lda,
             long sA, int btype, const void *B, long ldb, long sB, void *C, long ldc, long sC,
             int nbatch, const int *amap, const int *bmap, const int *klen, bool accumulate);
}  // namespace pp
