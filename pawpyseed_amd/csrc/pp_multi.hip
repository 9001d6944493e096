// The setup of several structures over many k-points with uneven band
// counts: one work list of (structure, k-point, 64-band chunk) items.
//
// The reference sets up each structure on its own (setup_projections,
// projector.c:560-602, once per Wavefunction of the
// setup_multiple_projections generator, projector.py:311-382) and loops
// over (band, k) inside.  Here every item of every structure goes into one
// list; a batch of `multi_items` items runs as one launch per kernel
// (scatter, z / y / x pass, sphere sums, reduction), each work-group
// reading its item's descriptor (chunk buffer, coefficients, plane-wave
// bounds of the k-point, sphere plan and phases of the structure).  Items
// of different k-points and structures share a launch, so the uneven
// plane-wave and band counts of the items even out inside one grid.
// Batches alternate between the context's two streams with two sets of
// chunk buffers.  The host-side preparation (bounds of every k-point, one
// host synchronisation, line lists, grid indices, phases of every
// (structure, k-point), the item array) runs once per call; the launches
// after it can be captured as one HIP graph (pp_graph_capture_multi).
#include <cstring>

#include "pp_internal.h"

namespace pp {

long sphere_npts(pp_sphere *sp);
int sphere_ptot(pp_sphere *sp);
int sphere_precision(pp_sphere *sp);
const int *sphere_fftg(pp_sphere *sp);
long sphere_generation(pp_sphere *sp);
const unsigned *sphere_xmask(pp_sphere *sp, int *words);
template <typename T> int make_phase(pp_ctx *ctx, pp_sphere *sp, const double *kfrac, cplx<T> *phase, bool expand);
int sphere_pm_of(pp_sphere *sp, bool *full);
void sphere_item_fields(pp_ctx *ctx, pp_sphere *sp, ChunkItem &it);
template <typename T>
int sphere_sums_items(pp_ctx *ctx, const ChunkItem *items, int nitems, int NB, int PM, bool full, long maxblk,
                      long maxred);
template <typename T>
int fft_items(pp_ctx *ctx, FftPlan *pl, const ChunkItem *items, int nitems, int NB, int maxw, long maxzl,
              long maxyl);
template <typename T> bool fft_items_supported(pp_ctx *ctx, FftPlan *pl, int NB);

// Where a multi-structure setup keeps its prepared data: the context's
// buffers for an eager call, a graph's own for a captured one.
struct MultiBufs {
  DevBuf *gidx, *phase, *bounds, *lines, *items, *grids, *partials;
  std::vector<int> *lines_host, *bounds_host;
};

struct MultiBatch {
  int first = 0, count = 0, set = 0;
  int maxw = 0;
  long maxzl = 0, maxyl = 0, maxblk = 0, maxred = 0;
};

struct MultiState {
  FftPlan *pl = nullptr;
  int NB = 64, PM = 8;
  bool full = true, two = false;
  ChunkItem *d_items = nullptr;
  std::vector<MultiBatch> batches;
};

struct MultiArgs {
  int nstruct;
  pp_sphere *const *plans;
  int nk;
  const double *kfrac;
  const int *gofs;
  const int *G_dev;
  const int *nband;
  const void *const *C;
  const long *ldc;
  void *const *proj;
};

// Host-side preparation: validates, sizes the buffers, computes the
// plane-wave bounds of every k-point (one host synchronisation), the line
// lists, grid indices, phases of every (structure, k-point) and the work
// items, and uploads the items.  PP_ERR_UNSUPPORTED when a k-point's G set
// aliases the grid or leaves holes in a z-line (then the per-structure
// path applies) or the grid needs the non-in-place FFT stages.
template <typename T>
static int multi_prepare(pp_ctx *ctx, const MultiArgs &a, const MultiBufs &mb, MultiState &S) {
  const int *fftg = sphere_fftg(a.plans[0]);
  PP_TRY(get_fft_plan(ctx, fftg, &S.pl));
  const int n0 = fftg[0], n1 = fftg[1], n2 = fftg[2];
  const long npts_grid = (long)n0 * n1 * n2;
  int NB = ctx->tune_fft_bands > 0 ? std::min(64, ctx->tune_fft_bands) : 64;
  int maxnb = 0;
  for (int i = 0; i < a.nstruct * a.nk; i++)
    maxnb = std::max(maxnb, a.nband[i]);
  NB = std::max(1, std::min(NB, maxnb));
  S.NB = NB;
  if (!fft_items_supported<T>(ctx, S.pl, NB))
    return PP_ERR_UNSUPPORTED;
  // plans: phases offsets, projector slots
  std::vector<long> ph_off(a.nstruct * a.nk + 1, 0);
  S.PM = 8;
  S.full = true;
  for (int s = 0; s < a.nstruct; s++) {
    bool full = false;
    const int pm = sphere_pm_of(a.plans[s], &full);
    S.PM = std::max(S.PM, pm);
    for (int k = 0; k < a.nk; k++)
      ph_off[s * a.nk + k + 1] = ph_off[s * a.nk + k] + std::max<long>(1, sphere_npts(a.plans[s]));
  }
  for (int s = 0; s < a.nstruct; s++) {
    bool full = false;
    const int pm = sphere_pm_of(a.plans[s], &full);
    S.full = S.full && full && pm == S.PM;
  }
  // plane-wave bounds of every k (one host synchronisation)
  const int nbnd = 2 * n0 * n1 + 2 * n0 + 3;
  PP_TRY(mb.bounds->ensure(sizeof(int) * (size_t)nbnd * a.nk));
  int *bnd = (int *)mb.bounds->p;
  for (int k = 0; k < a.nk; k++)
    PP_TRY(sphere_bounds(ctx, fftg, a.gofs[k + 1] - a.gofs[k], a.G_dev + 3 * (long)a.gofs[k],
                         bnd + (size_t)k * nbnd));
  std::vector<int> &hb_all = *mb.bounds_host;
  hb_all.resize((size_t)nbnd * a.nk);
  PP_CHECK_HIP(hipMemcpyAsync(hb_all.data(), bnd, sizeof(int) * hb_all.size(), hipMemcpyDeviceToHost, ctx->stream));
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<int> &lh = *mb.lines_host;
  lh.clear();
  std::vector<long> zl_off(a.nk + 1, 0), yl_off(a.nk + 1, 0);
  for (int k = 0; k < a.nk; k++) {
    const int *hb = hb_all.data() + (size_t)k * nbnd;
    if (hb[nbnd - 1] || !bounds_filled(hb, fftg, a.gofs[k + 1] - a.gofs[k]))
      return PP_ERR_UNSUPPORTED;
    const int *yb = hb + 2 * n0 * n1, *xb = yb + 2 * n0;
    zl_off[k] = (long)lh.size();
    for (int sx = xb[0]; sx <= xb[1]; sx++) {
      const int x = (sx + n0) % n0;
      for (int sy = yb[2 * x]; sy <= yb[2 * x + 1]; sy++)
        lh.push_back(x * n1 + (sy + n1) % n1);
    }
    yl_off[k] = (long)lh.size();
    for (int sx = xb[0]; sx <= xb[1]; sx++)
      for (int z = 0; z < n2; z++)
        lh.push_back(((sx + n0) % n0) * n2 + z);
    zl_off[k + 1] = (long)lh.size();
  }
  PP_TRY(mb.lines->ensure(sizeof(int) * std::max<size_t>(1, lh.size())));
  int *dl = (int *)mb.lines->p;
  if (!lh.empty())
    PP_CHECK_HIP(hipMemcpyAsync(dl, lh.data(), sizeof(int) * lh.size(), hipMemcpyHostToDevice, ctx->stream));
  PP_TRY(mb.gidx->ensure(sizeof(int) * (size_t)std::max(1, a.gofs[a.nk])));
  int *gidx = (int *)mb.gidx->p;
  PP_TRY(mb.phase->ensure(sizeof(cplx<T>) * (size_t)ph_off.back()));
  cplx<T> *phase = (cplx<T> *)mb.phase->p;
  for (int k = 0; k < a.nk; k++) {
    PP_TRY(compute_gidx(ctx, fftg, a.gofs[k + 1] - a.gofs[k], a.G_dev + 3 * (long)a.gofs[k], gidx + a.gofs[k]));
    for (int s = 0; s < a.nstruct; s++)
      PP_TRY(make_phase<T>(ctx, a.plans[s], a.kfrac + 3 * k, phase + ph_off[s * a.nk + k], false));
  }
  // work items in k-major order (the items of one k-point share its
  // bounds, lines and grid indices), batches of Q items
  const int Q = std::max(1, ctx->tune_multi_items);
  std::vector<ChunkItem> items;
  for (int k = 0; k < a.nk; k++) {
    const int *kb = bnd + (size_t)k * nbnd;
    SphereBounds bd = {kb, kb + 2 * n0 * n1, kb + 2 * n0 * n1 + 2 * n0, n0, n1, n2};
    bd.zl = dl + zl_off[k];
    bd.nzl = (int)(yl_off[k] - zl_off[k]);
    bd.yl = dl + yl_off[k];
    bd.nyl = (int)(zl_off[k + 1] - yl_off[k]);
    for (int s = 0; s < a.nstruct; s++) {
      const int i = s * a.nk + k;
      SphereBounds bds = bd;
      if (ctx->tune_fft_xmask)
        bds.xm = sphere_xmask(a.plans[s], &bds.xmw);
      const long ptot = sphere_ptot(a.plans[s]);
      for (int b0 = 0; b0 < a.nband[i]; b0 += NB) {
        ChunkItem it;
        memset(&it, 0, sizeof(it));
        it.C = (const PP_G cplx<float> *)((const cplx<float> *)a.C[i] + (long)b0 * a.ldc[i]);
        it.ldc = a.ldc[i];
        it.nbc = std::min(NB, a.nband[i] - b0);
        it.nw = a.gofs[k + 1] - a.gofs[k];
        it.gidx = (const PP_G int *)(gidx + a.gofs[k]);
        it.zb = (const PP_G int *)bds.zb;
        it.yb = (const PP_G int *)bds.yb;
        it.xb = (const PP_G int *)bds.xb;
        it.zl = (const PP_G int *)bds.zl;
        it.yl = (const PP_G int *)bds.yl;
        it.xm = (const PP_G unsigned *)bds.xm;
        it.n0 = bds.n0;
        it.n1 = bds.n1;
        it.n2 = bds.n2;
        it.nzl = bds.nzl;
        it.nyl = bds.nyl;
        it.xmw = bds.xmw;
        it.phase = (const PP_G void *)(phase + ph_off[i]);
        sphere_item_fields(ctx, a.plans[s], it);
        it.out = (PP_G cplx<double> *)((cplx<double> *)a.proj[i] + (long)b0 * ptot);
        it.ptot = ptot;
        items.push_back(it);
      }
    }
  }
  const int nitems = (int)items.size();
  S.two = ctx->tune_streams > 1 && nitems > Q;
  const int nsets = S.two ? 2 : 1;
  const int slots = std::min(Q, std::max(1, nitems));
  size_t part_bytes = 0;
  for (int s = 0; s < a.nstruct; s++) {
    ChunkItem tmp;
    sphere_item_fields(ctx, a.plans[s], tmp);
    part_bytes = std::max(part_bytes, sizeof(cplx<double>) * (size_t)tmp.nsites * tmp.nparts * NB * S.PM);
  }
  part_bytes = (std::max<size_t>(part_bytes, 256) + 255) / 256 * 256;
  const size_t grid_bytes = (sizeof(cplx<T>) * (size_t)npts_grid * NB + 255) / 256 * 256;
  PP_TRY(mb.grids->ensure(grid_bytes * nsets * slots));
  PP_TRY(mb.partials->ensure(part_bytes * nsets * slots));
  for (int b = 0; b * Q < nitems; b++) {
    MultiBatch bt;
    bt.first = b * Q;
    bt.count = std::min(Q, nitems - bt.first);
    bt.set = S.two ? (b & 1) : 0;
    for (int j = 0; j < bt.count; j++) {
      ChunkItem &it = items[bt.first + j];
      const size_t slot = (size_t)bt.set * slots + j;
      it.grid = (PP_G void *)((char *)mb.grids->p + slot * grid_bytes);
      it.partial = (PP_G void *)((char *)mb.partials->p + slot * part_bytes);
      bt.maxw = std::max(bt.maxw, it.nw);
      bt.maxzl = std::max<long>(bt.maxzl, it.nzl);
      bt.maxyl = std::max<long>(bt.maxyl, it.nyl);
      bt.maxblk = std::max<long>(bt.maxblk, (long)it.nsites * it.nparts);
      bt.maxred = std::max<long>(bt.maxred, (long)it.nsites * it.nbc * S.PM);
    }
    S.batches.push_back(bt);
  }
  PP_TRY(mb.items->ensure(sizeof(ChunkItem) * std::max(1, nitems)));
  S.d_items = (ChunkItem *)mb.items->p;
  if (nitems)
    PP_CHECK_HIP(hipMemcpyAsync(S.d_items, items.data(), sizeof(ChunkItem) * nitems, hipMemcpyHostToDevice,
                                ctx->stream));
  // the host arrays must outlive the uploads; the streams are idle after
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return PP_OK;
}

// The batches of a prepared call: batch b on stream b % 2 with chunk
// buffer set b % 2; the caller's stream waits for the second stream.
template <typename T> static int multi_launch(pp_ctx *ctx, const MultiState &S) {
  const bool two = S.two;
  hipStream_t streams[2] = {ctx->stream, ctx->stream};
  if (two) {
    if (!ctx->aux)
      PP_CHECK_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    if (!ctx->fork)
      PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
    if (!ctx->join)
      PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->join, hipEventDisableTiming));
    streams[1] = ctx->aux;
    PP_CHECK_HIP(hipEventRecord(ctx->fork, ctx->stream));
    PP_CHECK_HIP(hipStreamWaitEvent(ctx->aux, ctx->fork, 0));
  }
  int st = PP_OK;
  for (const MultiBatch &bt : S.batches) {
    ctx->stream = streams[bt.set];
    const ChunkItem *it = S.d_items + bt.first;
    st = fft_items<T>(ctx, S.pl, it, bt.count, S.NB, bt.maxw, bt.maxzl, bt.maxyl);
    if (st == PP_OK)
      st = sphere_sums_items<T>(ctx, it, bt.count, S.NB, S.PM, S.full, bt.maxblk, bt.maxred);
    if (st != PP_OK)
      break;
  }
  ctx->stream = streams[0];
  if (two) {
    (void)hipEventRecord(ctx->join, ctx->aux);
    (void)hipStreamWaitEvent(ctx->stream, ctx->join, 0);
  }
  return st;
}

static int multi_check(pp_ctx *ctx, const MultiArgs &a, int precision) {
  PP_REQUIRE(ctx && a.nstruct > 0 && a.plans && a.nk > 0 && a.kfrac && a.gofs && a.nband && a.C && a.ldc && a.proj,
             PP_ERR_ARG, "pp_setup_projections_multi: bad arguments");
  PP_REQUIRE(precision == PP_FP64 || precision == PP_FP32 || precision == PP_BF16X3 ||
                 precision == PP_BF16X6 || precision == PP_FP16X3,
             PP_ERR_ARG,
             "pp_setup_projections_multi: unknown precision");
  const int want = precision == PP_FP64 ? PP_FP64 : PP_FP32;
  for (int s = 0; s < a.nstruct; s++) {
    PP_REQUIRE(a.plans[s], PP_ERR_ARG, "pp_setup_projections_multi: null plan");
    PP_REQUIRE(sphere_precision(a.plans[s]) == want, PP_ERR_ARG,
               "pp_setup_projections_multi: a plan was built for another precision");
    PP_REQUIRE(!memcmp(sphere_fftg(a.plans[s]), sphere_fftg(a.plans[0]), 3 * sizeof(int)), PP_ERR_ARG,
               "pp_setup_projections_multi: the plans' FFT grids differ");
  }
  for (int k = 0; k < a.nk; k++)
    PP_REQUIRE(a.gofs[k + 1] >= a.gofs[k], PP_ERR_ARG, "pp_setup_projections_multi: bad gofs");
  for (int s = 0; s < a.nstruct; s++)
    for (int k = 0; k < a.nk; k++) {
      const int i = s * a.nk + k;
      PP_REQUIRE(a.nband[i] >= 0, PP_ERR_ARG, "pp_setup_projections_multi: negative band count");
      PP_REQUIRE(a.nband[i] == 0 || (a.C[i] && a.proj[i] && a.G_dev), PP_ERR_ARG,
                 "pp_setup_projections_multi: null buffer");
      PP_REQUIRE(a.nband[i] == 0 || a.gofs[k + 1] - a.gofs[k] <= a.ldc[i], PP_ERR_ARG,
                 "pp_setup_projections_multi: plane-wave count exceeds ldc");
    }
  return PP_OK;
}

// Per-structure fallback (an aliasing grid, a G set with holes, FFT stages
// outside the item form): each (structure, k-point) through the
// single-structure setup.
static int multi_fallback(pp_ctx *ctx, const MultiArgs &a, int precision) {
  for (int s = 0; s < a.nstruct; s++)
    for (int k = 0; k < a.nk; k++) {
      const int i = s * a.nk + k;
      if (a.nband[i] == 0)
        continue;
      const int go[2] = {0, a.gofs[k + 1] - a.gofs[k]};
      PP_TRY(pp_setup_projections_batched(ctx, a.plans[s], 1, a.kfrac + 3 * k, go, a.G_dev + 3 * (long)a.gofs[k],
                                          a.C[i], a.ldc[i], a.nband[i], a.proj[i], precision));
    }
  return PP_OK;
}

}  // namespace pp

using namespace pp;

extern "C" {

int pp_setup_projections_multi(pp_ctx *ctx, int nstruct, pp_sphere *const *plans, int nk, const double *kfrac,
                               const int *gofs, const int *G_dev, const int *nband, const void *const *C_dev,
                               const long *ldc, void *const *proj_dev, int precision) {
  const MultiArgs a = {nstruct, plans, nk, kfrac, gofs, G_dev, nband, C_dev, ldc, proj_dev};
  PP_TRY(multi_check(ctx, a, precision));
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  const MultiBufs mb = {&ctx->gidx, &ctx->phase, &ctx->bounds, &ctx->lines, &ctx->m_items, &ctx->m_grids,
                        &ctx->m_partials, &ctx->lines_host, &ctx->bounds_host};
  MultiState S;
  int st = precision == PP_FP64 ? multi_prepare<double>(ctx, a, mb, S) : multi_prepare<float>(ctx, a, mb, S);
  if (st == PP_ERR_UNSUPPORTED)
    return multi_fallback(ctx, a, precision);
  PP_TRY(st);
  return precision == PP_FP64 ? multi_launch<double>(ctx, S) : multi_launch<float>(ctx, S);
}

int pp_graph_capture_multi(pp_ctx *ctx, int nstruct, pp_sphere *const *plans, int nk, const double *kfrac,
                           const int *gofs, const int *G_dev, const int *nband, const void *const *C_dev,
                           const long *ldc, void *const *proj_dev, int precision, pp_graph **out) {
  PP_REQUIRE(out, PP_ERR_ARG, "pp_graph_capture_multi: null output");
  *out = nullptr;
  const MultiArgs a = {nstruct, plans, nk, kfrac, gofs, G_dev, nband, C_dev, ldc, proj_dev};
  PP_TRY(multi_check(ctx, a, precision));
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  pp_graph *gr = new pp_graph();
  gr->ctx = ctx;
  gr->multi = true;
  for (int s = 0; s < nstruct; s++)
    gr->plans.push_back({plans[s], sphere_generation(plans[s])});
  const MultiBufs mb = {&gr->gidx, &gr->phase, &gr->bounds, &gr->lines, &gr->items, &gr->grids, &gr->partials,
                        &gr->lines_host, &gr->bounds_host};
  MultiState S;
  const bool timing = ctx->timer.on;
  ctx->timer.on = false;
  int st = precision == PP_FP64 ? multi_prepare<double>(ctx, a, mb, S) : multi_prepare<float>(ctx, a, mb, S);
  if (st == PP_ERR_UNSUPPORTED)
    set_error(st, "pp_graph_capture_multi: these inputs need the per-structure setup (aliasing grid, a G set "
                  "with holes, or FFT stages outside the work-item form)");
  if (st == PP_OK) {
    // one eager run leaves valid projections and warms the kernels
    st = precision == PP_FP64 ? multi_launch<double>(ctx, S) : multi_launch<float>(ctx, S);
    if (st == PP_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
      st = PP_ERR_HIP;
  }
  if (st == PP_OK) {
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      set_error(PP_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
      st = PP_ERR_HIP;
    } else {
      st = precision == PP_FP64 ? multi_launch<double>(ctx, S) : multi_launch<float>(ctx, S);
      e = hipStreamEndCapture(ctx->stream, &gr->g);
      if (st == PP_OK && e != hipSuccess) {
        set_error(PP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        st = PP_ERR_HIP;
      }
      if (st == PP_OK && (e = hipGraphInstantiate(&gr->ge, gr->g, nullptr, nullptr, 0)) != hipSuccess) {
        set_error(PP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        st = PP_ERR_HIP;
      }
    }
  }
  ctx->timer.on = timing;
  if (st != PP_OK) {
    pp_graph_destroy(gr);
    return st;
  }
  *out = gr;
  return PP_OK;
}

// Autotuner of the multi-structure setup: coordinate descent (as
// pp_autotune) over multi_items, streams, fft_tile, sphere_unroll,
// sphere_parts, fft_nt, sphere_nt and fft_xmask on the caller's inputs.
int pp_autotune_multi(pp_ctx *ctx, int nstruct, pp_sphere *const *plans, int nk, const double *kfrac,
                      const int *gofs, const int *G_dev, const int *nband, const void *const *C_dev,
                      const long *ldc, void *const *proj_dev, int precision, int repeats, char *report,
                      int report_len) {
  const MultiArgs a = {nstruct, plans, nk, kfrac, gofs, G_dev, nband, C_dev, ldc, proj_dev};
  PP_TRY(multi_check(ctx, a, precision));
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  const bool fp32 = precision != PP_FP64;
  std::vector<TuneCand> cands = {
      {"multi_items", {1, 2, 4, 8, 16}},
      {"streams", {1, 2}},
      {"fft_tile", fp32 ? std::vector<int>{16, 32} : std::vector<int>{8, 16}},
      {"sphere_unroll", {4, 8}},
      {"sphere_parts", {0, 4, 16}},
      {"fft_nt", {0, 1, 2, 3}},
      {"sphere_nt", {0, 1}},
      {"fft_xmask", {0, 1}},
  };
  return autotune_core(
      ctx, cands,
      [&]() {
        return pp_setup_projections_multi(ctx, nstruct, plans, nk, kfrac, gofs, G_dev, nband, C_dev, ldc, proj_dev,
                                          precision);
      },
      repeats, report, report_len);
}

}  // extern "C"
