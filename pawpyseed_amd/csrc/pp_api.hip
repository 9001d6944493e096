// Context, errors, timing and the batched device-resident entry points.
#include <cstring>
#include <functional>
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
                                        "gemm_reduce",   "aug_prep",   "aug_gemm", "misc", "fft_x_sphere",
                                        "rs_augment",    "rs_emit",    "gemm_split"};
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
long sphere_generation(pp_sphere *sp);
const unsigned *sphere_xmask(pp_sphere *sp, int *words);
// fused x pass with sphere sums (pp_fused.hip)
bool fused_usable(pp_ctx *ctx, FftPlan *pl, pp_sphere *sp, int NB, bool fp32);
int fused_x_sums(pp_ctx *ctx, FftPlan *pl, pp_sphere *sp, const cplx<float> *grid, int NB, int nbc,
                 const SphereBounds *bounds, const cplx<float> *phase, cplx<double> *out, long ptot);

// Host-side preparation of a setup call: buffers, FFT plan, plane-wave
// sphere bounds of every k (one host sync for the alias flags), the
// populated-line lists, grid indices and Bloch phases.  Everything after it
// (setup_launch) is device work on the context's streams only, so it can be
// captured in a graph.
struct SetupState {
  FftPlan *pl = nullptr;
  int NB = 64, nbnd = 0;
  bool two = false, fused = false;
  long npts_sph = 1, ptot = 0;
  void *grids[2] = {nullptr, nullptr};
  void *phase = nullptr;
  int *gidx = nullptr, *bnd = nullptr, *dlines = nullptr;
  std::vector<int> alias;
  std::vector<long> zl_off, yl_off;
};

// Where a prepared setup keeps what setup_prepare computes from the call's
// G vectors and k-points: the context's buffers for eager calls, a graph's
// own buffers for a captured pipeline (so that later eager calls on the
// same context cannot change what a replay reads).
struct PrepBufs {
  DevBuf *gidx, *phase, *bounds, *lines;
  std::vector<int> *lines_host, *bounds_host;
};
static PrepBufs ctx_prep_bufs(pp_ctx *ctx) {
  return {&ctx->gidx, &ctx->phase, &ctx->bounds, &ctx->lines, &ctx->lines_host, &ctx->bounds_host};
}

template <typename T>
static int setup_prepare(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac, const int *gofs, const int *G_dev,
                         int nband, SetupState &S, const PrepBufs &pb) {
  const int *fftg = sphere_fftg(sp);
  PP_TRY(get_fft_plan(ctx, fftg, &S.pl));
  int NB = ctx->tune_fft_bands > 0 ? ctx->tune_fft_bands : 64;
  if (NB > 64)
    NB = 64;
  if (NB > nband)
    NB = nband < 1 ? 1 : nband;
  S.NB = NB;
  const long npts_grid = (long)fftg[0] * fftg[1] * fftg[2];
  S.ptot = sphere_ptot(sp);
  S.npts_sph = std::max<long>(1, sphere_npts(sp));
  S.two = ctx->tune_streams > 1 && (long)nk * ((nband + NB - 1) / NB) > 1;
  // the x pass and the sphere sums as one kernel where the plan allows
  S.fused = fused_usable(ctx, S.pl, sp, NB, sizeof(T) == 4);
  PP_TRY(ctx->grid.ensure(sizeof(cplx<T>) * (size_t)npts_grid * NB));
  if (S.two) {
    PP_TRY(ctx->grid2.ensure(sizeof(cplx<T>) * (size_t)npts_grid * NB));
    if (!ctx->aux)
      PP_CHECK_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    if (!ctx->fork)
      PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
    if (!ctx->join)
      PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
  }
  PP_TRY(pb.gidx->ensure(sizeof(int) * (size_t)std::max(1, gofs[nk])));
  PP_TRY(pb.phase->ensure(sizeof(cplx<T>) * (size_t)S.npts_sph * nk));
  S.grids[0] = ctx->grid.p;
  S.grids[1] = S.two ? ctx->grid2.p : ctx->grid.p;
  S.phase = pb.phase->p;
  S.gidx = (int *)pb.gidx->p;
  // plane-wave sphere bounds of every k (one host sync for the alias flags)
  S.nbnd = 2 * fftg[0] * fftg[1] + 2 * fftg[0] + 3;
  PP_TRY(pb.bounds->ensure(sizeof(int) * (size_t)S.nbnd * nk));
  S.bnd = (int *)pb.bounds->p;
  for (int k = 0; k < nk; k++)
    PP_TRY(sphere_bounds(ctx, fftg, gofs[k + 1] - gofs[k], G_dev + 3 * (long)gofs[k], S.bnd + (size_t)k * S.nbnd));
  pb.bounds_host[0].resize((size_t)S.nbnd * nk);
  PP_CHECK_HIP(hipMemcpyAsync(pb.bounds_host[0].data(), S.bnd, sizeof(int) * pb.bounds_host[0].size(),
                              hipMemcpyDeviceToHost, ctx->stream));
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  S.alias.assign(nk, 1);
  // populated lines of the z pass ((x, y) in the sphere's hull) and of the
  // y pass ((x, z) with x in the slab), so no work-group is launched for an
  // empty line
  S.zl_off.assign(nk + 1, 0);
  S.yl_off.assign(nk + 1, 0);
  pb.lines_host[0].clear();
  const int n0 = fftg[0], n1 = fftg[1], n2 = fftg[2];
  for (int k = 0; k < nk; k++) {
    const int *hb = pb.bounds_host[0].data() + (size_t)k * S.nbnd;
    // no pruning when the grid aliases a G or the G set leaves holes in a
    // z-line's interval (not a cutoff sphere)
    S.alias[k] = hb[S.nbnd - 1] || !bounds_filled(hb, fftg, gofs[k + 1] - gofs[k]);
    const int *yb = hb + 2 * n0 * n1, *xb = yb + 2 * n0;
    S.zl_off[k] = (long)pb.lines_host[0].size();
    if (!S.alias[k])
      for (int sx = xb[0]; sx <= xb[1]; sx++) {
        const int x = (sx + n0) % n0;
        for (int sy = yb[2 * x]; sy <= yb[2 * x + 1]; sy++)
          pb.lines_host[0].push_back(x * n1 + (sy + n1) % n1);
      }
    S.yl_off[k] = (long)pb.lines_host[0].size();
    if (!S.alias[k])
      for (int sx = xb[0]; sx <= xb[1]; sx++)
        for (int z = 0; z < n2; z++)
          pb.lines_host[0].push_back(((sx + n0) % n0) * n2 + z);
    S.zl_off[k + 1] = (long)pb.lines_host[0].size();
  }
  S.dlines = nullptr;
  if (!pb.lines_host[0].empty()) {
    PP_TRY(pb.lines->ensure(sizeof(int) * pb.lines_host[0].size()));
    S.dlines = (int *)pb.lines->p;
    // the streams are idle here (synchronised above), and lines_host lives in
    // the context until the next call's synchronisation
    PP_CHECK_HIP(hipMemcpyAsync(S.dlines, pb.lines_host[0].data(), sizeof(int) * pb.lines_host[0].size(),
                                hipMemcpyHostToDevice, ctx->stream));
  }
  // grid indices and Bloch phases of every k, so the chunks of all k can be
  // in flight on both streams
  for (int k = 0; k < nk; k++) {
    PP_TRY(compute_gidx(ctx, fftg, gofs[k + 1] - gofs[k], G_dev + 3 * (long)gofs[k], S.gidx + gofs[k]));
    PP_TRY(make_phase<T>(ctx, sp, kfrac + 3 * k, (cplx<T> *)S.phase + (size_t)k * S.npts_sph, false));
  }
  return PP_OK;
}

// The chunk pipeline of a prepared setup call.  Chunk c runs on stream
// c % 2 with buffer set c % 2, so one chunk's sphere sums overlap the next
// chunk's FFT passes; the caller's stream waits for the second stream at
// the end.
template <typename T>
static int setup_launch(pp_ctx *ctx, pp_sphere *sp, int nk, const int *gofs, const cplx<float> *C, long ldc,
                        int nband, cplx<double> *proj, const SetupState &S) {
  const int *fftg = sphere_fftg(sp);
  const int NB = S.NB;
  const long ptot = S.ptot;
  const bool two = S.two;
  hipStream_t streams[2] = {ctx->stream, two ? ctx->aux : ctx->stream};
  if (two) {
    PP_CHECK_HIP(hipEventRecord(ctx->fork, ctx->stream));
    PP_CHECK_HIP(hipStreamWaitEvent(ctx->aux, ctx->fork, 0));
  }
  int st = PP_OK;
  long c = 0;
  for (int k = 0; k < nk && st == PP_OK; k++) {
    const int nw = gofs[k + 1] - gofs[k];
    const int *kb = S.bnd + (size_t)k * S.nbnd;
    SphereBounds bd = {kb, kb + 2 * fftg[0] * fftg[1], kb + 2 * fftg[0] * fftg[1] + 2 * fftg[0], fftg[0], fftg[1],
                       fftg[2]};
    if (S.dlines && !S.alias[k] && ctx->tune_fft_prune != 3) {
      bd.zl = S.dlines + S.zl_off[k];
      bd.nzl = (int)(S.yl_off[k] - S.zl_off[k]);
      bd.yl = S.dlines + S.yl_off[k];
      bd.nyl = (int)(S.zl_off[k + 1] - S.yl_off[k]);
    }
    const bool prune = S.alias[k] == 0 && ctx->tune_fft_prune != 2;
    if (ctx->tune_fft_xmask && !S.fused)
      bd.xm = sphere_xmask(sp, &bd.xmw);
    const cplx<T> *phase_k = (const cplx<T> *)S.phase + (size_t)k * S.npts_sph;
    for (int b0 = 0; b0 < nband && st == PP_OK; b0 += NB, c++) {
      const int nbc = std::min(NB, nband - b0);
      const int s = two ? (int)(c & 1) : 0;
      ctx->stream = streams[s];
      ctx->slot = s;
      cplx<T> *grid = (cplx<T> *)S.grids[s];
      st = fft_chunk<T>(ctx, S.pl, grid, NB, nw, S.gidx + gofs[k], C + ((long)k * nband + b0) * ldc, ldc, nbc,
                        prune ? &bd : nullptr, !S.fused);
      if (st == PP_OK && S.fused)
        st = fused_x_sums(ctx, S.pl, sp, (const cplx<float> *)grid, NB, nbc, prune ? &bd : nullptr,
                          (const cplx<float> *)phase_k, proj + ((long)k * nband + b0) * ptot, ptot);
      else if (st == PP_OK)
        st = sphere_sums<T>(ctx, sp, grid, NB, nbc, phase_k, proj + ((long)k * nband + b0) * ptot, ptot);
    }
  }
  ctx->stream = streams[0];
  ctx->slot = 0;
  if (two) {
    // the caller's stream sees everything the second stream did
    (void)hipEventRecord(ctx->join, ctx->aux);
    (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
  }
  return st;
}

template <typename T>
static int setup_batched(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac, const int *gofs, const int *G_dev,
                         const cplx<float> *C, long ldc, int nband, cplx<double> *proj) {
  SetupState S;
  PP_TRY(setup_prepare<T>(ctx, sp, nk, kfrac, gofs, G_dev, nband, S, ctx_prep_bufs(ctx)));
  return setup_launch<T>(ctx, sp, nk, gofs, C, ldc, nband, proj, S);
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
  if (ctx->fork)
    (void)hipEventDestroy(ctx->fork);
  if (ctx->join)
    (void)hipEventDestroy(ctx->join);
  for (int i = 0; i < 2; i++) {
    if (ctx->gemm_done[i])
      (void)hipEventDestroy(ctx->gemm_done[i]);
    if (ctx->red_done[i])
      (void)hipEventDestroy(ctx->red_done[i]);
    if (ctx->split_done[i])
      (void)hipEventDestroy(ctx->split_done[i]);
  }
  if (ctx->aug_upload) {
    (void)hipEventSynchronize(ctx->aug_upload);
    (void)hipEventDestroy(ctx->aug_upload);
  }
  if (ctx->aug_done) {
    (void)hipEventSynchronize(ctx->aug_done);
    (void)hipEventDestroy(ctx->aug_done);
  }
  if (ctx->aug_pinned)
    (void)hipHostFree(ctx->aug_pinned);
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
  else if (n == "fft_xmask")
    ctx->tune_fft_xmask = value != 0 ? 1 : 0;
  else if (n == "fft_xcd")
    ctx->tune_fft_xcd = value != 0 ? 1 : 0;
  else if (n == "fft_fixed")
    ctx->tune_fft_fixed = value != 0 ? 1 : 0;
  else if (n == "sphere_nt")
    ctx->tune_sphere_nt = value != 0 ? 1 : 0;
  else if (n == "fft_nt")
    ctx->tune_fft_nt = value < 0 ? 0 : value > 3 ? 3 : value;
  else if (n == "fused")
    ctx->tune_fused = value != 0 ? 1 : 0;
  else if (n == "rs_kernel")
    ctx->tune_rs_kernel = value < 0 ? 0 : value > 2 ? 2 : value;
  else if (n == "rs_prune")
    ctx->tune_rs_prune = value != 0 ? 1 : 0;
  else if (n == "rs_unroll")
    ctx->tune_rs_unroll = value >= 4 ? 4 : value >= 2 ? 2 : 1;
  else if (n == "multi_items")
    ctx->tune_multi_items = value < 1 ? 1 : value > 64 ? 64 : value;
  else if (n == "gemm_planes")
    ctx->tune_gemm_planes = value >= 4 ? 4 : value == 3 ? 3 : value == 2 ? 2 : 1;
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

int pp_ctx_get_param(pp_ctx *ctx, const char *name, int *value) {
  PP_REQUIRE(ctx && name && value, PP_ERR_ARG, "pp_ctx_get_param: null argument");
  const std::string n(name);
  const std::pair<const char *, int> table[] = {
      {"fft_pingpong", ctx->tune_fft_pingpong}, {"gemm_f64_valu", ctx->tune_gemm_f64_valu},
      {"fft_persistent", ctx->tune_fft_persistent}, {"streams", ctx->tune_streams},
      {"sphere_unroll", ctx->tune_sphere_unroll}, {"fft_threads", ctx->tune_fft_threads},
      {"fft_radix", ctx->tune_fft_radix}, {"fft_xcd", ctx->tune_fft_xcd}, {"fft_fixed", ctx->tune_fft_fixed},
      {"fft_xmask", ctx->tune_fft_xmask}, {"fft_nt", ctx->tune_fft_nt},
      {"sphere_nt", ctx->tune_sphere_nt},
      {"fused", ctx->tune_fused}, {"rs_kernel", ctx->tune_rs_kernel}, {"rs_prune", ctx->tune_rs_prune},
      {"rs_unroll", ctx->tune_rs_unroll}, {"multi_items", ctx->tune_multi_items},
      {"gemm_planes", ctx->tune_gemm_planes},
      {"fft_tile", ctx->tune_fft_tile}, {"fft_prune", ctx->tune_fft_prune}, {"fft_bands", ctx->tune_fft_bands},
      {"sphere_parts", ctx->tune_sphere_parts}, {"gemm_splitk", ctx->tune_splitk}};
  for (const auto &e : table)
    if (n == e.first) {
      *value = e.second;
      return PP_OK;
    }
  PP_REQUIRE(false, PP_ERR_ARG, "pp_ctx_get_param: unknown parameter " + n);
  return PP_OK;
}

}  // extern "C"

namespace pp {
// Coordinate descent over tuning parameters: each candidate value is timed
// with HIP events around run (the caller's own inputs, on the context's
// stream), best of `repeats` after one warm-up run, and the fastest value
// of a parameter is kept before the next parameter is tried.  A candidate
// replaces the value in place only when it is faster by more than the
// run-to-run spread (2%), so that equal settings keep the library default.
// On failure the parameter being tried gets its previous value back.  A
// final run leaves the outputs of the chosen setting.  report (may be
// NULL) receives one line per candidate: "name value ms".
int autotune_core(pp_ctx *ctx, const std::vector<TuneCand> &cands, const std::function<int()> &run, int repeats,
                  char *report, int report_len) {
  PP_REQUIRE(repeats >= 1, PP_ERR_ARG, "autotune: repeats must be >= 1");
  const bool timing = ctx->timer.on;
  ctx->timer.on = false;
  hipEvent_t e0, e1;
  PP_CHECK_HIP(hipEventCreate(&e0));
  PP_CHECK_HIP(hipEventCreate(&e1));
  std::string rep;
  auto run_once = [&](float &ms) -> int {
    PP_CHECK_HIP(hipEventRecord(e0, ctx->stream));
    PP_TRY(run());
    PP_CHECK_HIP(hipEventRecord(e1, ctx->stream));
    PP_CHECK_HIP(hipEventSynchronize(e1));
    PP_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    return PP_OK;
  };
  auto trial = [&](float &best) -> int {
    float ms = 0.f;
    PP_TRY(run_once(ms));  // warm-up: buffers, plans
    best = 1e30f;
    for (int r = 0; r < repeats; r++) {
      PP_TRY(run_once(ms));
      best = std::min(best, ms);
    }
    return PP_OK;
  };
  int st = PP_OK;
  for (const TuneCand &c : cands) {
    int cur = 0;
    pp_ctx_get_param(ctx, c.name, &cur);
    int best_v = cur;
    float best_ms = 1e30f;
    for (int v : c.values) {
      if ((st = pp_ctx_set_param(ctx, c.name, v)) != PP_OK)
        break;
      float ms = 0.f;
      if ((st = trial(ms)) != PP_OK)
        break;
      rep += std::string(c.name) + " " + std::to_string(v) + " " + std::to_string(ms) + "\n";
      const bool incumbent = v == cur;
      if (best_ms == 1e30f || (incumbent ? ms < best_ms * 1.02f : ms * 1.02f < best_ms)) {
        best_ms = ms;
        best_v = v;
      }
    }
    if (st != PP_OK) {
      const std::string msg = pp_last_error();
      pp_ctx_set_param(ctx, c.name, cur);
      set_error(st, msg);
      break;
    }
    pp_ctx_set_param(ctx, c.name, best_v);
  }
  if (st == PP_OK) {
    float ms = 0.f;
    st = run_once(ms);  // outputs of the chosen setting
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->timer.on = timing;
  if (report && report_len > 0) {
    const size_t n = std::min<size_t>(rep.size(), (size_t)report_len - 1);
    memcpy(report, rep.data(), n);
    report[n] = 0;
  }
  return st;
}
}  // namespace pp

extern "C" {

// Project-written autotuner of the setup path over streams, fft_tile,
// fft_threads, sphere_unroll, sphere_parts, fft_radix, fft_xcd, fft_xmask,
// fft_nt, sphere_nt, fft_bands and (FP32) fused, on the caller's inputs.
int pp_autotune(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac, const int *gofs, const int *G_dev,
                const void *C_dev, long ldc, int nband, void *proj_dev, int precision, int repeats, char *report,
                int report_len) {
  PP_REQUIRE(ctx && sp && repeats >= 1, PP_ERR_ARG, "pp_autotune: bad arguments");
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  const bool fp32 = precision != PP_FP64;
  std::vector<TuneCand> cands = {
      {"streams", {1, 2}},
      {"fft_tile", fp32 ? std::vector<int>{16, 32} : std::vector<int>{8, 16}},
      {"fft_threads", {256, 512}},
      {"sphere_unroll", {4, 8}},
      {"sphere_parts", {0, 4, 16}},
      {"fft_radix", {0, 1}},
      {"fft_xcd", {0, 1}},
      {"fft_xmask", {0, 1}},
      {"fft_nt", {0, 1, 2, 3}},
      {"sphere_nt", {0, 1}},
      {"fft_bands", {64, 32}},
  };
  if (fp32)
    cands.push_back({"fused", {0, 1}});
  return autotune_core(
      ctx, cands,
      [&]() {
        return pp_setup_projections_batched(ctx, sp, nk, kfrac, gofs, G_dev, C_dev, ldc, nband, proj_dev, precision);
      },
      repeats, report, report_len);
}

}  // extern "C"

static std::vector<std::pair<void *, size_t>> ctx_buffers(pp_ctx *ctx) {
  std::vector<std::pair<void *, size_t>> v;
  for (const DevBuf *b : {&ctx->grid, &ctx->grid2, &ctx->partial, &ctx->partial2})
    v.push_back({b->p, b->bytes});
  return v;
}

template <typename T>
static int graph_capture(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac, const int *gofs, const int *G_dev,
                         const cplx<float> *C, long ldc, int nband, cplx<double> *proj, pp_graph *gr) {
  SetupState S;
  const PrepBufs pb = {&gr->gidx, &gr->phase, &gr->bounds, &gr->lines, &gr->lines_host, &gr->bounds_host};
  PP_TRY(setup_prepare<T>(ctx, sp, nk, kfrac, gofs, G_dev, nband, S, pb));
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  PP_CHECK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int st = setup_launch<T>(ctx, sp, nk, gofs, C, ldc, nband, proj, S);
  hipError_t e = hipStreamEndCapture(ctx->stream, &gr->g);
  PP_TRY(st);
  PP_CHECK_HIP(e);
  PP_CHECK_HIP(hipGraphInstantiate(&gr->ge, gr->g, nullptr, nullptr, 0));
  return PP_OK;
}

extern "C" {

int pp_graph_capture(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac, const int *gofs, const int *G_dev,
                     const void *C_dev, long ldc, int nband, void *proj_dev, int precision, pp_graph **out) {
  PP_REQUIRE(out, PP_ERR_ARG, "pp_graph_capture: null output");
  *out = nullptr;
  PP_REQUIRE(nk > 0 && nband > 0, PP_ERR_ARG, "pp_graph_capture: nothing to capture");
  // one eager run validates the arguments, sizes every buffer the graph
  // uses and leaves valid projections
  PP_TRY(pp_setup_projections_batched(ctx, sp, nk, kfrac, gofs, G_dev, C_dev, ldc, nband, proj_dev, precision));
  pp_graph *gr = new pp_graph();
  gr->ctx = ctx;
  gr->plans.push_back({sp, sphere_generation(sp)});
  const bool timing = ctx->timer.on;
  ctx->timer.on = false;
  int st = precision == PP_FP64
               ? graph_capture<double>(ctx, sp, nk, kfrac, gofs, G_dev, (const cplx<float> *)C_dev, ldc, nband,
                                       (cplx<double> *)proj_dev, gr)
               : graph_capture<float>(ctx, sp, nk, kfrac, gofs, G_dev, (const cplx<float> *)C_dev, ldc, nband,
                                      (cplx<double> *)proj_dev, gr);
  ctx->timer.on = timing;
  if (st != PP_OK) {
    pp_graph_destroy(gr);
    return st;
  }
  gr->bufs = ctx_buffers(ctx);
  *out = gr;
  return PP_OK;
}

int pp_graph_replay(pp_graph *gr) {
  PP_REQUIRE(gr && gr->ge, PP_ERR_ARG, "pp_graph_replay: null graph");
  for (const auto &pg : gr->plans)
    PP_REQUIRE(sphere_generation(pg.first) == pg.second, PP_ERR_ARG,
               "pp_graph_replay: a sphere plan was rebuilt since the capture");
  PP_REQUIRE(gr->multi || ctx_buffers(gr->ctx) == gr->bufs, PP_ERR_ARG,
             "pp_graph_replay: context buffers were reallocated since the capture");
  PP_CHECK_HIP(hipSetDevice(gr->ctx->device));
  PP_CHECK_HIP(hipGraphLaunch(gr->ge, gr->ctx->stream));
  return PP_OK;
}

void pp_graph_destroy(pp_graph *gr) {
  if (!gr)
    return;
  if (gr->ge)
    (void)hipGraphExecDestroy(gr->ge);
  if (gr->g)
    (void)hipGraphDestroy(gr->g);
  delete gr;
}

int pp_overlap_gemm(pp_ctx *ctx, int nwk, int nspin, int flip, const int *nw, const void *CR_dev, long ldR, int nbR,
                    const void *CS_dev, long ldS, int nbS, void *out_dev, int precision) {
  PP_REQUIRE(ctx && nw && nwk > 0 && (nspin == 1 || nspin == 2) && nbR >= 0 && nbS >= 0, PP_ERR_ARG,
             "pp_overlap_gemm: bad arguments");
  PP_REQUIRE(precision == PP_FP64 || precision == PP_FP32 || precision == PP_BF16X3 ||
                 precision == PP_BF16X6 || precision == PP_FP16X3,
             PP_ERR_ARG,
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
