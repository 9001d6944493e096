// Context, errors, timing and the batched device-resident entry points.
#include <cstring>
#include <mutex>

#include "pp_internal.h"

struct pp_sphere;

namespace pp {

static thread_local std::string g_err;
static thread_local int g_code = PP_OK;

void set_error(int code, const std::string &msg) {
  g_code = code;
  g_err = msg;
}
int last_code() { return g_code; }

const char *kclass_name(int c) {
  static const char *names[K_NCLASS] = {"zero",          "scatter",    "fft_z",       "fft_y",    "fft_x",
                                        "phase",         "sphere_sum", "sphere_reduce", "table",  "gemm",
                                        "gemm_reduce",   "aug_prep",   "aug_gemm", "misc", "fft_x_sphere"};
  return (c >= 0 && c < K_NCLASS) ? names[c] : "?";
}

int DevBuf::ensure(size_t need) {
  if (need <= bytes && p)
    return PP_OK;
  release();
  size_t alloc = need < 256 ? 256 : need;
  hipError_t e = hipMalloc(&p, alloc);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    set_error(PP_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(alloc) + "): " + hipGetErrorString(e));
    return PP_ERR_NOMEM;
  }
  bytes = alloc;
  return PP_OK;
}

void DevBuf::release() {
  if (p)
    (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

void tbegin(pp_ctx *ctx, int cls, hipEvent_t *start) {
  *start = nullptr;
  if (!ctx->timer.on)
    return;
  hipEvent_t e;
  if (!ctx->timer.pool.empty()) {
    e = ctx->timer.pool.back();
    ctx->timer.pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    return;
  }
  (void)hipEventRecord(e, ctx->stream);
  *start = e;
  (void)cls;
}

void tend(pp_ctx *ctx, int cls, hipEvent_t start) {
  if (!ctx->timer.on || !start)
    return;
  hipEvent_t e;
  if (!ctx->timer.pool.empty()) {
    e = ctx->timer.pool.back();
    ctx->timer.pool.pop_back();
  } else if (hipEventCreate(&e) != hipSuccess) {
    return;
  }
  (void)hipEventRecord(e, ctx->stream);
  ctx->timer.pending.push_back({cls, {start, e}});
}

static void timer_collect(pp_ctx *ctx) {
  if (ctx->timer.pending.empty())
    return;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &pe : ctx->timer.pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
      ctx->timer.ms[pe.first] += ms;
      ctx->timer.launches[pe.first] += 1;
    }
    ctx->timer.pool.push_back(pe.second.first);
    ctx->timer.pool.push_back(pe.second.second);
  }
  ctx->timer.pending.clear();
}

// sphere helpers implemented in pp_sphere.hip
template <typename T>
int sphere_sums(pp_ctx *ctx, pp_sphere *sp, const cplx<T> *grid, int NB, int nbc, const cplx<T> *phase,
                cplx<double> *out, long ptot);
template <typename T> int make_phase(pp_ctx *ctx, pp_sphere *sp, const double *kfrac, cplx<T> *phase, bool expand);
template <typename T>
int sphere_expand(pp_ctx *ctx, pp_sphere *sp, cplx<T> *grid, int NB, int nbc, const cplx<T> *phase,
                  const cplx<double> *P, long ptotP, const int *src_off_dev);
int sphere_nsites(pp_sphere *sp);
long sphere_npts(pp_sphere *sp);
int sphere_ptot(pp_sphere *sp);
int sphere_precision(pp_sphere *sp);
const int *sphere_fftg(pp_sphere *sp);
const double *sphere_lattice(pp_sphere *sp);
// fused x pass with sphere sums (pp_fused.hip)
bool fused_usable(pp_ctx *ctx, FftPlan *pl, pp_sphere *sp, int NB, bool fp32);
int fused_x_sums(pp_ctx *ctx, FftPlan *pl, pp_sphere *sp, const cplx<float> *grid, int NB, int nbc,
                 const SphereBounds *bounds, const cplx<float> *phase, cplx<double> *out, long ptot);

template <typename T>
static int setup_batched(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac, const int *gofs, const int *G_dev,
                         const cplx<float> *C, long ldc, int nband, cplx<double> *proj) {
  const int *fftg = sphere_fftg(sp);
  FftPlan *pl = nullptr;
  PP_TRY(get_fft_plan(ctx, fftg, &pl));
  int NB = ctx->tune_fft_bands > 0 ? ctx->tune_fft_bands : 64;
  if (NB > 64)
    NB = 64;
  if (NB > nband)
    NB = nband < 1 ? 1 : nband;
  const long npts_grid = (long)fftg[0] * fftg[1] * fftg[2];
  const long ptot = sphere_ptot(sp);
  int maxw = 0;
  for (int k = 0; k < nk; k++)
    maxw = std::max(maxw, gofs[k + 1] - gofs[k]);
  const long npts_sph = std::max<long>(1, sphere_npts(sp));
  const bool two = ctx->tune_streams > 1 && (long)nk * ((nband + NB - 1) / NB) > 1;
  // the x pass and the sphere sums as one kernel where the plan allows
  const bool fused = fused_usable(ctx, pl, sp, NB, sizeof(T) == 4);
  PP_TRY(ctx->grid.ensure(sizeof(cplx<T>) * (size_t)npts_grid * NB));
  if (two) {
    PP_TRY(ctx->grid2.ensure(sizeof(cplx<T>) * (size_t)npts_grid * NB));
    if (!ctx->aux)
      PP_CHECK_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
  }
  PP_TRY(ctx->gidx.ensure(sizeof(int) * (size_t)std::max(1, gofs[nk])));
  PP_TRY(ctx->phase.ensure(sizeof(cplx<T>) * (size_t)npts_sph * nk));
  cplx<T> *grids[2] = {(cplx<T> *)ctx->grid.p, two ? (cplx<T> *)ctx->grid2.p : (cplx<T> *)ctx->grid.p};
  cplx<T> *phase = (cplx<T> *)ctx->phase.p;
  int *gidx = (int *)ctx->gidx.p;
  // plane-wave sphere bounds of every k (one host sync for the alias flags)
  const int nbnd = 2 * fftg[0] * fftg[1] + 2 * fftg[0] + 3;
  PP_TRY(ctx->bounds.ensure(sizeof(int) * (size_t)nbnd * nk));
  int *bnd = (int *)ctx->bounds.p;
  for (int k = 0; k < nk; k++)
    PP_TRY(sphere_bounds(ctx, fftg, gofs[k + 1] - gofs[k], G_dev + 3 * (long)gofs[k], bnd + (size_t)k * nbnd));
  ctx->bounds_host.resize((size_t)nbnd * nk);
  PP_CHECK_HIP(hipMemcpyAsync(ctx->bounds_host.data(), bnd, sizeof(int) * ctx->bounds_host.size(),
                              hipMemcpyDeviceToHost, ctx->stream));
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<int> alias(nk, 1);
  // populated lines of the z pass ((x, y) in the sphere's hull) and of the
  // y pass ((x, z) with x in the slab), so no work-group is launched for an
  // empty line
  std::vector<long> zl_off(nk + 1, 0), yl_off(nk + 1, 0);
  ctx->lines_host.clear();
  const int n0 = fftg[0], n1 = fftg[1], n2 = fftg[2];
  for (int k = 0; k < nk; k++) {
    const int *hb = ctx->bounds_host.data() + (size_t)k * nbnd;
    alias[k] = hb[nbnd - 1];
    const int *yb = hb + 2 * n0 * n1, *xb = yb + 2 * n0;
    zl_off[k] = (long)ctx->lines_host.size();
    if (!alias[k])
      for (int sx = xb[0]; sx <= xb[1]; sx++) {
        const int x = (sx + n0) % n0;
        for (int sy = yb[2 * x]; sy <= yb[2 * x + 1]; sy++)
          ctx->lines_host.push_back(x * n1 + (sy + n1) % n1);
      }
    yl_off[k] = (long)ctx->lines_host.size();
    if (!alias[k])
      for (int sx = xb[0]; sx <= xb[1]; sx++)
        for (int z = 0; z < n2; z++)
          ctx->lines_host.push_back(((sx + n0) % n0) * n2 + z);
    zl_off[k + 1] = (long)ctx->lines_host.size();
  }
  int *dlines = nullptr;
  if (!ctx->lines_host.empty()) {
    PP_TRY(ctx->lines.ensure(sizeof(int) * ctx->lines_host.size()));
    dlines = (int *)ctx->lines.p;
    // the streams are idle here (synchronised above), and lines_host lives in
    // the context until the next call's synchronisation
    PP_CHECK_HIP(hipMemcpyAsync(dlines, ctx->lines_host.data(), sizeof(int) * ctx->lines_host.size(),
                                hipMemcpyHostToDevice, ctx->stream));
  }
  // grid indices and Bloch phases of every k, so the chunks of all k can be
  // in flight on both streams
  for (int k = 0; k < nk; k++) {
    PP_TRY(compute_gidx(ctx, fftg, gofs[k + 1] - gofs[k], G_dev + 3 * (long)gofs[k], gidx + gofs[k]));
    PP_TRY(make_phase<T>(ctx, sp, kfrac + 3 * k, phase + (size_t)k * npts_sph, false));
  }
  // Chunk pipeline: chunk c runs on stream c % 2 with buffer set c % 2, so
  // one chunk's sphere sums overlap the next chunk's FFT passes.
  hipStream_t streams[2] = {ctx->stream, two ? ctx->aux : ctx->stream};
  hipEvent_t fork = nullptr, join = nullptr;
  if (two) {
    PP_CHECK_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    PP_CHECK_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    PP_CHECK_HIP(hipEventRecord(fork, ctx->stream));
    PP_CHECK_HIP(hipStreamWaitEvent(ctx->aux, fork, 0));
  }
  int st = PP_OK;
  long c = 0;
  for (int k = 0; k < nk && st == PP_OK; k++) {
    const int nw = gofs[k + 1] - gofs[k];
    const int *kb = bnd + (size_t)k * nbnd;
    SphereBounds bd = {kb, kb + 2 * fftg[0] * fftg[1], kb + 2 * fftg[0] * fftg[1] + 2 * fftg[0], fftg[0], fftg[1],
                       fftg[2]};
    if (dlines && !alias[k] && ctx->tune_fft_prune != 3) {
      bd.zl = dlines + zl_off[k];
      bd.nzl = (int)(yl_off[k] - zl_off[k]);
      bd.yl = dlines + yl_off[k];
      bd.nyl = (int)(zl_off[k + 1] - yl_off[k]);
    }
    const bool prune = alias[k] == 0 && ctx->tune_fft_prune != 2;
    for (int b0 = 0; b0 < nband && st == PP_OK; b0 += NB, c++) {
      const int nbc = std::min(NB, nband - b0);
      const int s = two ? (int)(c & 1) : 0;
      ctx->stream = streams[s];
      ctx->slot = s;
      st = fft_chunk<T>(ctx, pl, grids[s], NB, nw, gidx + gofs[k], C + ((long)k * nband + b0) * ldc, ldc, nbc,
                        prune ? &bd : nullptr, !fused);
      if (st == PP_OK && fused)
        st = fused_x_sums(ctx, pl, sp, (const cplx<float> *)grids[s], NB, nbc, prune ? &bd : nullptr,
                          (const cplx<float> *)(phase + (size_t)k * npts_sph), proj + ((long)k * nband + b0) * ptot,
                          ptot);
      else if (st == PP_OK)
        st = sphere_sums<T>(ctx, sp, grids[s], NB, nbc, phase + (size_t)k * npts_sph,
                            proj + ((long)k * nband + b0) * ptot, ptot);
    }
  }
  ctx->stream = streams[0];
  ctx->slot = 0;
  if (two) {
    // the caller's stream sees everything the second stream did
    (void)hipEventRecord(join, ctx->aux);
    (void)hipStreamWaitEvent(ctx->stream, join, 0);
    (void)hipEventDestroy(fork);
    (void)hipEventDestroy(join);
  }
  return st;
}

// projector.c:420-453 for every band and k of one structure: the
// augmentation functions of the plan's sites weighted by the band's
// projections, forward-transformed (scale sqrt(V)/N^3, linalg.c:47-79) and
// sampled at the band's G vectors.  One band chunk at a time on one stream.
template <typename T>
static int aug_freqs_batched(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac, const int *gofs,
                             const int *G_dev, const cplx<double> *P, long ptotP, const int *src_off_dev, int nband,
                             cplx<double> *out, long ldo) {
  const int *fftg = sphere_fftg(sp);
  FftPlan *pl = nullptr;
  PP_TRY(get_fft_plan(ctx, fftg, &pl));
  const int NB = std::min(64, std::max(1, nband));
  const long npts_grid = (long)fftg[0] * fftg[1] * fftg[2];
  const long npts_sph = std::max<long>(1, sphere_npts(sp));
  PP_TRY(ctx->grid.ensure(sizeof(cplx<T>) * (size_t)npts_grid * NB));
  PP_TRY(ctx->gidx.ensure(sizeof(int) * (size_t)std::max(1, gofs[nk])));
  PP_TRY(ctx->phase.ensure(sizeof(cplx<T>) * (size_t)npts_sph));
  cplx<T> *grid = (cplx<T> *)ctx->grid.p;
  cplx<T> *phase = (cplx<T> *)ctx->phase.p;
  int *gidx = (int *)ctx->gidx.p;
  const double scl = sqrt(det3(sphere_lattice(sp))) / (double)npts_grid;
  ctx->slot = 0;
  for (int k = 0; k < nk; k++) {
    const int nw = gofs[k + 1] - gofs[k];
    PP_TRY(compute_gidx(ctx, fftg, nw, G_dev + 3 * (long)gofs[k], gidx));
    PP_TRY(make_phase<T>(ctx, sp, kfrac + 3 * k, phase, true));
    for (int b0 = 0; b0 < nband; b0 += NB) {
      const int nbc = std::min(NB, nband - b0);
      PP_TRY(zero_async(ctx, grid, sizeof(cplx<T>) * (size_t)npts_grid * NB));
      PP_TRY(sphere_expand<T>(ctx, sp, grid, NB, nbc, phase, P + ((long)k * nband + b0) * ptotP, ptotP,
                              src_off_dev));
      PP_TRY(fft_passes<T>(ctx, pl, grid, NB));
      PP_TRY(gather_conj<T>(ctx, grid, NB, nbc, gidx, nw, scl, out + ((long)k * nband + b0) * ldo, ldo));
    }
  }
  return PP_OK;
}

}  // namespace pp

using namespace pp;

extern "C" {

const char *pp_last_error(void) { return g_err.c_str(); }
int pp_last_status(void) { return g_code; }
const char *pp_version(void) { return "pawpyseed_amd 0.1.0 (gfx950)"; }

int pp_aug_freqs_batched(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac, const int *gofs, const int *G_dev,
                         const void *proj_dev, long ptot_proj, const int *src_off, int nband, void *out_dev, long ldo,
                         int precision) {
  PP_REQUIRE(ctx && sp && nk > 0 && kfrac && gofs && nband >= 0 && ptot_proj >= 0 && ldo >= 0, PP_ERR_ARG,
             "pp_aug_freqs_batched: bad arguments");
  PP_REQUIRE(precision == PP_FP64 || precision == PP_FP32, PP_ERR_ARG, "pp_aug_freqs_batched: precision");
  PP_REQUIRE(sphere_precision(sp) == (precision == PP_FP64 ? PP_FP64 : PP_FP32), PP_ERR_ARG,
             "pp_aug_freqs_batched: the plan's table precision differs from the call's");
  if (nband == 0)
    return PP_OK;
  const int ns = sphere_nsites(sp);
  PP_REQUIRE(ns == 0 || (src_off && proj_dev && out_dev), PP_ERR_ARG, "pp_aug_freqs_batched: null argument");
  for (int k = 0; k < nk; k++)
    PP_REQUIRE(gofs[k + 1] - gofs[k] >= 0 && gofs[k + 1] - gofs[k] <= ldo, PP_ERR_ARG,
               "pp_aug_freqs_batched: plane-wave count exceeds ldo");
  std::vector<int> poff(ns + 1);
  int ptot_plan = 0;
  PP_TRY(pp_sphere_info(sp, &ptot_plan, nullptr, poff.data()));
  for (int s = 0; s < ns; s++)
    PP_REQUIRE(src_off[s] >= 0 && src_off[s] + (poff[s + 1] - poff[s]) <= ptot_proj, PP_ERR_ARG,
               "pp_aug_freqs_batched: site offset outside the projection rows");
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  PP_TRY(ctx->aug_ws.ensure(sizeof(int) * (size_t)std::max(1, ns)));
  int *d_src = (int *)ctx->aug_ws.p;
  if (ns > 0)
    PP_CHECK_HIP(hipMemcpyAsync(d_src, src_off, sizeof(int) * ns, hipMemcpyHostToDevice, ctx->stream));
  int st;
  if (precision == PP_FP64)
    st = aug_freqs_batched<double>(ctx, sp, nk, kfrac, gofs, G_dev, (const cplx<double> *)proj_dev, ptot_proj, d_src,
                                   nband, (cplx<double> *)out_dev, ldo);
  else
    st = aug_freqs_batched<float>(ctx, sp, nk, kfrac, gofs, G_dev, (const cplx<double> *)proj_dev, ptot_proj, d_src,
                                  nband, (cplx<double> *)out_dev, ldo);
  // the host copy of src_off must outlive the async upload
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return st;
}

int pp_ctx_create(pp_ctx **out, int device) {
  PP_REQUIRE(out, PP_ERR_ARG, "pp_ctx_create: null out");
  int ndev = 0;
  PP_CHECK_HIP(hipGetDeviceCount(&ndev));
  PP_REQUIRE(device >= 0 && device < ndev, PP_ERR_ARG, "pp_ctx_create: no such device");
  PP_CHECK_HIP(hipSetDevice(device));
  pp_ctx *c = new pp_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    PP_CHECK_HIP(e);
  }
  c->stream = c->own;
  *out = c;
  return PP_OK;
}

void pp_ctx_destroy(pp_ctx *ctx) {
  if (!ctx)
    return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &kv : ctx->fft_plans)
    delete kv.second;
  for (auto &pe : ctx->timer.pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  for (auto e : ctx->timer.pool)
    (void)hipEventDestroy(e);
  if (ctx->aux) {
    (void)hipStreamSynchronize(ctx->aux);
    (void)hipStreamDestroy(ctx->aux);
  }
  if (ctx->own)
    (void)hipStreamDestroy(ctx->own);
  delete ctx;
}

int pp_ctx_set_stream(pp_ctx *ctx, void *stream) {
  PP_REQUIRE(ctx, PP_ERR_ARG, "pp_ctx_set_stream: null ctx");
  ctx->stream = stream ? (hipStream_t)stream : ctx->own;
  return PP_OK;
}

int pp_ctx_synchronize(pp_ctx *ctx) {
  PP_REQUIRE(ctx, PP_ERR_ARG, "pp_ctx_synchronize: null ctx");
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return PP_OK;
}

int pp_ctx_set_timing(pp_ctx *ctx, int enable) {
  PP_REQUIRE(ctx, PP_ERR_ARG, "pp_ctx_set_timing: null ctx");
  timer_collect(ctx);
  ctx->timer.on = enable != 0;
  return PP_OK;
}

int pp_ctx_timing_reset(pp_ctx *ctx) {
  PP_REQUIRE(ctx, PP_ERR_ARG, "pp_ctx_timing_reset: null ctx");
  timer_collect(ctx);
  for (int c = 0; c < K_NCLASS; c++) {
    ctx->timer.ms[c] = 0;
    ctx->timer.launches[c] = 0;
  }
  return PP_OK;
}

int pp_ctx_timing_read(pp_ctx *ctx, int n, const char **names, double *ms_total, long *launches) {
  PP_REQUIRE(ctx, PP_ERR_ARG, "pp_ctx_timing_read: null ctx");
  timer_collect(ctx);
  for (int c = 0; c < K_NCLASS && c < n; c++) {
    if (names)
      names[c] = kclass_name(c);
    if (ms_total)
      ms_total[c] = ctx->timer.ms[c];
    if (launches)
      launches[c] = ctx->timer.launches[c];
  }
  return K_NCLASS;
}

int pp_ctx_set_tuning(pp_ctx *ctx, int fft_bands, int fft_tile, int fft_prune, int sphere_parts, int gemm_splitk) {
  PP_REQUIRE(ctx, PP_ERR_ARG, "pp_ctx_set_tuning: null ctx");
  PP_REQUIRE(fft_bands >= 0 && fft_bands <= 64, PP_ERR_ARG, "fft_bands must be 0..64");
  PP_REQUIRE(fft_tile == 0 || fft_tile == 1 || fft_tile == 2 || fft_tile == 4 || fft_tile == 8 || fft_tile == 16 ||
                 fft_tile == 32,
             PP_ERR_ARG, "fft_tile must be 0 or a power of two <= 32");
  PP_REQUIRE(fft_prune >= 0 && fft_prune <= 2, PP_ERR_ARG, "fft_prune must be 0 (auto), 1 (on) or 2 (off)");
  PP_REQUIRE(sphere_parts >= 0 && gemm_splitk >= 0, PP_ERR_ARG, "negative tuning value");
  ctx->tune_fft_bands = fft_bands;
  ctx->tune_fft_tile = fft_tile;
  ctx->tune_fft_prune = fft_prune;
  ctx->tune_sphere_parts = sphere_parts;
  ctx->tune_splitk = gemm_splitk;
  return PP_OK;
}

int pp_ctx_set_param(pp_ctx *ctx, const char *name, int value) {
  PP_REQUIRE(ctx && name, PP_ERR_ARG, "pp_ctx_set_param: null argument");
  std::string n(name);
  if (n == "fft_pingpong")
    ctx->tune_fft_pingpong = value != 0;
  else if (n == "gemm_f64_valu")
    ctx->tune_gemm_f64_valu = value != 0;
  else if (n == "fft_persistent")
    ctx->tune_fft_persistent = value < 0 ? 0 : value;
  else if (n == "streams")
    ctx->tune_streams = value > 1 ? 2 : 1;
  else if (n == "sphere_unroll")
    ctx->tune_sphere_unroll = value >= 8 ? 8 : 4;
  else if (n == "fft_threads")
    ctx->tune_fft_threads = value >= 512 ? 512 : 256;
  else if (n == "fft_radix")
    ctx->tune_fft_radix = value != 0 ? 1 : 0;
  else if (n == "fused")
    ctx->tune_fused = value != 0 ? 1 : 0;
  else if (n == "fft_tile")
    return pp_ctx_set_tuning(ctx, ctx->tune_fft_bands, value, ctx->tune_fft_prune, ctx->tune_sphere_parts,
                             ctx->tune_splitk);
  else if (n == "fft_prune")
    return pp_ctx_set_tuning(ctx, ctx->tune_fft_bands, ctx->tune_fft_tile, value, ctx->tune_sphere_parts,
                             ctx->tune_splitk);
  else if (n == "fft_bands")
    return pp_ctx_set_tuning(ctx, value, ctx->tune_fft_tile, ctx->tune_fft_prune, ctx->tune_sphere_parts,
                             ctx->tune_splitk);
  else if (n == "sphere_parts")
    return pp_ctx_set_tuning(ctx, ctx->tune_fft_bands, ctx->tune_fft_tile, ctx->tune_fft_prune, value,
                             ctx->tune_splitk);
  else if (n == "gemm_splitk")
    return pp_ctx_set_tuning(ctx, ctx->tune_fft_bands, ctx->tune_fft_tile, ctx->tune_fft_prune,
                             ctx->tune_sphere_parts, value);
  else
    PP_REQUIRE(false, PP_ERR_ARG, "pp_ctx_set_param: unknown parameter " + n);
  return PP_OK;
}

int pp_setup_projections_batched(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac, const int *gofs,
                                 const int *G_dev, const void *C_dev, long ldc, int nband, void *proj_dev,
                                 int precision) {
  PP_REQUIRE(ctx && sp && kfrac && gofs && nk >= 0 && nband >= 0, PP_ERR_ARG,
             "pp_setup_projections_batched: bad arguments");
  PP_REQUIRE(nband == 0 || (C_dev && proj_dev && G_dev), PP_ERR_ARG, "pp_setup_projections_batched: null buffer");
  for (int k = 0; k < nk; k++)
    PP_REQUIRE(gofs[k + 1] >= gofs[k] && gofs[k + 1] - gofs[k] <= ldc, PP_ERR_ARG,
               "pp_setup_projections_batched: plane-wave count exceeds ldc");
  int want = precision == PP_FP64 ? PP_FP64 : PP_FP32;
  PP_REQUIRE(sphere_precision(sp) == want, PP_ERR_ARG,
             "pp_setup_projections_batched: sphere plan built for another precision");
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  if (nband == 0 || nk == 0)
    return PP_OK;
  if (want == PP_FP64)
    return setup_batched<double>(ctx, sp, nk, kfrac, gofs, G_dev, (const cplx<float> *)C_dev, ldc, nband,
                                 (cplx<double> *)proj_dev);
  return setup_batched<float>(ctx, sp, nk, kfrac, gofs, G_dev, (const cplx<float> *)C_dev, ldc, nband,
                              (cplx<double> *)proj_dev);
}

int pp_overlap_gemm(pp_ctx *ctx, int nwk, int nspin, int flip, const int *nw, const void *CR_dev, long ldR, int nbR,
                    const void *CS_dev, long ldS, int nbS, void *out_dev, int precision) {
  PP_REQUIRE(ctx && nw && nwk > 0 && (nspin == 1 || nspin == 2) && nbR >= 0 && nbS >= 0, PP_ERR_ARG,
             "pp_overlap_gemm: bad arguments");
  PP_REQUIRE(precision == PP_FP64 || precision == PP_FP32 || precision == PP_BF16X3, PP_ERR_ARG,
             "pp_overlap_gemm: unknown precision");
  const int nkt = nwk * nspin;
  for (int k = 0; k < nkt; k++)
    PP_REQUIRE(nw[k] >= 0 && nw[k] <= ldR && nw[k] <= ldS, PP_ERR_ARG, "pp_overlap_gemm: nw exceeds ld");
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  std::vector<int> amap(nkt);
  for (int k = 0; k < nkt; k++)
    amap[k] = (flip && nspin == 2) ? (k < nwk ? k + nwk : k - nwk) : k;
  return zgemm_nt(ctx, precision, nbR, nbS, 0, 1, CR_dev, ldR, (long)nbR * ldR, 1, CS_dev, ldS, (long)nbS * ldS,
                  out_dev, nbS, (long)nbR * nbS, nkt, amap.data(), nullptr, nw, false);
}

}  // extern "C"
