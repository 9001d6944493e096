// Reference-named C entry points over the device-resident data model.
//
// pswf_t keeps the reference's pointer-tree layout (pawpyseed/core/utils.h)
// for callers that read it; the library additionally keeps each
// structure's coefficients, G vectors and projections in HBM (WfDev,
// pswf_t::dev) and answers the per-band calls of the reference
// (pseudoprojection, compensation_terms) from all-band matrices computed
// once on the GPU.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <tuple>

#include "pp_internal.h"

namespace pp {

struct PairCache {
  std::vector<pp_cdouble> m;  // [nkt][nbR][nbS]
  int nbR = 0, nbS = 0;
};

// Identities of wavefunctions for the caches other wavefunctions keep: a
// process-wide counter, never reused, so a freed pswf_t whose address comes
// back from calloc for another structure can never hit an old entry.
static std::atomic<unsigned long long> g_uid{0};
static unsigned long long next_uid() { return ++g_uid; }

struct WfDev {
  // identity of this wavefunction's coefficients (fixed once uploaded) and
  // of its current projections (renewed by every setup_projections)
  unsigned long long uid = next_uid();
  unsigned long long proj_uid = 0;
  int nkt = 0, nband = 0;
  long ldc = 0;
  std::vector<int> nw, gofs;
  DevBuf C, G;  // c64 [nkt][nband][ldc]; int32 [sum nw][3]
  bool uploaded = false;
  DevBuf proj;     // c128 [nkt][nband][ptot] (spinors: the up component)
  DevBuf proj_dn;  // spinors: the down component, same layout
  int ptot = 0, nsites = 0, proj_precision = -1;
  std::vector<int> proj_off, labels;
  std::vector<int> sp_nchan, sp_ls;
  std::vector<double> sp_dO;  // aepw - pspw per species
  // keyed by the other wavefunction's uid (overlaps) or proj_uid (terms
  // built from its projections)
  std::map<std::tuple<unsigned long long, int, int>, PairCache> overlap_cache;
  std::map<std::tuple<unsigned long long, int, unsigned long long>, PairCache> aug_cache;
  // <(phi - phi~)_smooth | psi~> at the other structure's unmatched sites
  // (overlap_setup_real): c128 [nkt][nband][wp_ptot]; wp_labels index the
  // other structure's species
  DevBuf wp;
  int wp_n = 0, wp_ptot = 0;
  unsigned long long wp_of = 0;  // uid of the wavefunction whose sites they are
  std::vector<int> wp_off, wp_labels, wp_nchan, wp_ls;
  unsigned long long wp_hash = 0;
  // off-site overlap blocks of the N_RS pairs (overlap_setup_real), kept on
  // the projected structure: conj of the partial-wave overlaps, [P_R x P_S]
  // per pair at on_off[t], and the minimum-image pair vectors
  int on_n = 0;
  unsigned long long on_of = 0;  // uid of the reference wavefunction
  std::vector<int> on_R, on_S;
  std::vector<pp_cdouble> on_ov;
  std::vector<long> on_off;
  std::vector<double> on_dc;
  // aug_recip (overlap_setup_recip): plane-wave coefficients of this
  // structure's augmentation functions at its own unmatched sites, c128
  // [nkt][nband][ca_ld]
  DevBuf ca;
  long ca_ld = 0;
  int ca_n = -1;
  unsigned long long ca_hash = 0;
  // real-space states: the partial-wave sphere plan of the last
  // (grid, labels, coords) asked for
  pp_rsplan *rs = nullptr;
  unsigned long long rs_hash = 0;
  ~WfDev() {
    if (rs)
      pp_rs_plan_destroy(rs);
  }
};

static std::mutex g_mu;
static pp_ctx *g_ctx = nullptr;
static int g_precision = PP_FP64;

static int global_ctx(pp_ctx **out) {
  if (!g_ctx) {
    int dev = 0;
    const char *e = getenv("PAWPY_AMD_DEVICE");
    if (e)
      dev = atoi(e);
    PP_TRY(pp_ctx_create(&g_ctx, dev));
  }
  *out = g_ctx;
  return PP_OK;
}

static WfDev *dev_of(pswf_t *wf) {
  if (!wf->dev)
    wf->dev = new WfDev();
  return (WfDev *)wf->dev;
}

static int ensure_upload(pp_ctx *ctx, pswf_t *wf) {
  WfDev *d = dev_of(wf);
  if (d->uploaded)
    return PP_OK;
  d->nkt = wf->nwk * wf->nspin;
  d->nband = wf->nband;
  d->nw.assign(d->nkt, 0);
  d->gofs.assign(d->nkt + 1, 0);
  int maxw = 0;
  for (int k = 0; k < d->nkt; k++) {
    d->nw[k] = wf->kpts[k]->num_waves;
    d->gofs[k + 1] = d->gofs[k] + d->nw[k];
    maxw = std::max(maxw, d->nw[k]);
  }
  d->ldc = std::max(64, (maxw + 63) / 64 * 64);
  size_t cbytes = sizeof(pp_cfloat) * (size_t)d->nkt * d->nband * d->ldc;
  PP_TRY(d->C.ensure(cbytes));
  PP_TRY(d->G.ensure(sizeof(int) * 3 * (size_t)std::max(1, d->gofs[d->nkt])));
  std::vector<pp_cfloat> stage((size_t)d->nband * d->ldc);
  for (int k = 0; k < d->nkt; k++) {
    std::fill(stage.begin(), stage.end(), pp_cfloat(0.f, 0.f));
    for (int b = 0; b < d->nband; b++)
      memcpy(&stage[(size_t)b * d->ldc], wf->kpts[k]->bands[b]->Cs, sizeof(pp_cfloat) * d->nw[k]);
    PP_CHECK_HIP(hipMemcpy((pp_cfloat *)d->C.p + (size_t)k * d->nband * d->ldc, stage.data(),
                           sizeof(pp_cfloat) * stage.size(), hipMemcpyHostToDevice));
    if (d->nw[k])
      PP_CHECK_HIP(hipMemcpy((int *)d->G.p + 3 * (size_t)d->gofs[k], wf->kpts[k]->Gs, sizeof(int) * 3 * d->nw[k],
                             hipMemcpyHostToDevice));
  }
  d->uploaded = true;
  return PP_OK;
}

static void free_projection_list(projection_t *pl, int num) {
  if (!pl)
    return;
  for (int i = 0; i < num; i++) {
    free(pl[i].ns);
    free(pl[i].ls);
    free(pl[i].ms);
    free(pl[i].overlaps);
  }
  free(pl);
}

static double **spline3(const double *x, const double *y, int n) {
  double **c = (double **)malloc(3 * sizeof(double *));
  for (int i = 0; i < 3; i++)
    c[i] = (double *)malloc(sizeof(double) * n);
  spline_coeff_host(x, y, n, c[0], c[1], c[2]);
  return c;
}

static void free_spline(double **s) {
  if (!s)
    return;
  for (int i = 0; i < 3; i++)
    free(s[i]);
  free(s);
}

// utils.c:751-764
static double spline_integral_host(const double *x, const double *a, double **s, int n) {
  double r = 0.0;
  for (int i = 0; i < n - 1; i++) {
    double h = x[i + 1] - x[i];
    r += h * (a[i] + h * (s[0][i] / 2 + h * (s[1][i] / 3 + s[2][i] * h / 4)));
  }
  return r;
}

// ------------------------------------------------------------------ WAVECAR
struct Src {
  FILE *fp = nullptr;
  const char *mem = nullptr;
  bool read(long off, void *dst, size_t n) {
    if (fp) {
      if (fseek(fp, off, SEEK_SET) != 0)
        return false;
      return fread(dst, 1, n, fp) == n;
    }
    memcpy(dst, mem + off, n);
    return true;
  }
};

static void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double mag3(const double *a) { return sqrt(dot3(a, a)); }

// Half-extents of the G box from the cutoff and the reciprocal lattice
// (the geometric bound of reader.c:55-127).
static void gbox(const double *rec, double encut, double *nb) {
  const double c = 0.262465831;
  double m[3] = {mag3(rec), mag3(rec + 3), mag3(rec + 6)};
  double g = sqrt(encut * c), v[3];
  auto sinb = [&](const double *u, const double *w) { return fabs(sin(acos(dot3(u, w) / (mag3(u) * mag3(w))))); };
  auto sinp = [&](const double *w, const double *u, const double *x) {
    cross3(u, x, v);
    return fabs(dot3(w, v) / (mag3(v) * mag3(w)));
  };
  const double *b1 = rec, *b2 = rec + 3, *b3 = rec + 6;
  double A[3] = {g / (m[0] * sinb(b1, b2)) + 1, g / (m[1] * sinb(b1, b2)) + 1, g / (m[2] * sinp(b3, b1, b2)) + 1};
  double B[3] = {g / (m[0] * sinb(b1, b3)) + 1, g / (m[1] * sinp(b2, b1, b3)) + 1, g / (m[2] * sinb(b1, b3)) + 1};
  double C[3] = {g / (m[0] * sinp(b1, b2, b3)) + 1, g / (m[1] * sinb(b2, b3)) + 1, g / (m[2] * sinb(b2, b3)) + 1};
  for (int i = 0; i < 3; i++)
    nb[i] = std::max(A[i], std::max(B[i], C[i]));
}

static pswf_t *new_pswf(int nspin, int nwk, int nband, double encut, const double *lattice) {
  pswf_t *wf = (pswf_t *)calloc(1, sizeof(pswf_t));
  wf->nspin = nspin;
  wf->nwk = nwk;
  wf->nband = nband;
  wf->encut = encut;
  wf->lattice = (double *)malloc(9 * sizeof(double));
  wf->reclattice = (double *)malloc(9 * sizeof(double));
  memcpy(wf->lattice, lattice, 9 * sizeof(double));
  double v = det3(lattice);
  cross3(lattice + 3, lattice + 6, wf->reclattice);
  cross3(lattice + 6, lattice + 0, wf->reclattice + 3);
  cross3(lattice + 0, lattice + 3, wf->reclattice + 6);
  for (int i = 0; i < 9; i++)
    wf->reclattice[i] *= 2.0 * M_PI / v;
  wf->G_bounds = (int *)calloc(6, sizeof(int));
  wf->kpts = (kpoint_t **)calloc((size_t)nwk * nspin, sizeof(kpoint_t *));
  return wf;
}

static kpoint_t *new_kpoint(int nband, int nplane) {
  kpoint_t *kp = (kpoint_t *)calloc(1, sizeof(kpoint_t));
  kp->num_bands = nband;
  kp->num_waves = nplane;
  kp->bands = (band_t **)calloc(nband, sizeof(band_t *));
  kp->k = (double *)calloc(3, sizeof(double));
  kp->Gs = (int *)malloc(sizeof(int) * 3 * (size_t)std::max(1, nplane));
  for (int b = 0; b < nband; b++) {
    band_t *bd = (band_t *)calloc(1, sizeof(band_t));
    bd->n = b;
    bd->num_waves = nplane;
    bd->Cs = (pp_cfloat *)malloc(sizeof(pp_cfloat) * (size_t)std::max(1, nplane));
    kp->bands[b] = bd;
  }
  return kp;
}

static pswf_t *read_wavecar_src(Src &src, const double *kpt_weights) {
  double head[3];
  if (!src.read(0, head, sizeof(head))) {
    set_error(PP_ERR_IO, "WAVECAR: cannot read header");
    return nullptr;
  }
  if (!(head[1] == 1.0 || head[1] == 2.0) || head[2] != 45200.0) {
    set_error(PP_ERR_IO, "WAVECAR: unsupported header (only single-precision WAVECARs)");
    return nullptr;
  }
  const int nspin = (int)head[1];
  // a record holds at least the 12 doubles of record 1 (counts, cutoff,
  // lattice) and whole doubles; the size bound keeps offsets in range
  const long nrecl = (head[0] >= 96.0 && head[0] <= (double)(1L << 40)) ? lround(head[0]) : -1;
  if (nrecl < 96 || nrecl % 8 != 0) {
    set_error(PP_ERR_IO, "WAVECAR: bad record length " + std::to_string(nrecl));
    return nullptr;
  }
  const long nrd = nrecl / 8;  // doubles per record
  std::vector<double> rec(nrd);
  if (!src.read(nrecl, rec.data(), nrecl)) {
    set_error(PP_ERR_IO, "WAVECAR: cannot read record 1");
    return nullptr;
  }
  if (!std::isfinite(rec[0]) || !std::isfinite(rec[1]) || !(rec[2] > 0.0) || !std::isfinite(rec[2])) {
    set_error(PP_ERR_IO, "WAVECAR: bad record 1");
    return nullptr;
  }
  const double nwk_d = rec[0], nband_d = rec[1];
  // each k-point record holds 4 doubles and 3 per band (energy, occupation)
  if (nwk_d < 1 || nband_d < 1 || nwk_d > 1e7 || 7 + 3 * (nband_d - 1) > (double)nrd) {
    set_error(PP_ERR_IO, "WAVECAR: k-point or band count does not fit the record length");
    return nullptr;
  }
  int nwk = (int)lround(nwk_d), nband = (int)lround(nband_d);
  double encut = rec[2];
  pswf_t *wf = new_pswf(nspin, nwk, nband, encut, &rec[3]);
  double nb[3];
  gbox(wf->reclattice, encut, nb);
  const double *b1 = wf->reclattice, *b2 = wf->reclattice + 3, *b3 = wf->reclattice + 6;
  std::vector<float> cbuf(nrecl / 4);
  for (int iwk = 0; iwk < nwk * nspin; iwk++) {
    long irec = (long)iwk * (1 + nband);
    if (!src.read((irec + 2) * nrecl, rec.data(), nrecl)) {
      set_error(PP_ERR_IO, "WAVECAR: truncated k-point record");
      free_pswf(wf);
      return nullptr;
    }
    // a band record of nplane complex64 values must fit one record
    if (!std::isfinite(rec[0]) || rec[0] < 0 || rec[0] > (double)nrd) {
      set_error(PP_ERR_IO, "WAVECAR: plane-wave count does not fit the record length");
      free_pswf(wf);
      return nullptr;
    }
    int nplane = (int)lround(rec[0]);
    kpoint_t *kp = new_kpoint(nband, nplane);
    wf->kpts[iwk] = kp;
    kp->up = (short)(iwk / nwk);
    kp->k[0] = rec[1];
    kp->k[1] = rec[2];
    kp->k[2] = rec[3];
    for (int b = 0; b < nband; b++) {
      kp->bands[b]->energy = rec[4 + 3 * b];
      kp->bands[b]->occ = rec[6 + 3 * b];
    }
    // G enumeration in VASP order (reader.c:231-271)
    int ncnt = 0;
    for (int ig3 = 0; ig3 <= 2 * nb[2]; ig3++) {
      int g3 = ig3 > nb[2] ? (int)(ig3 - 2 * nb[2] - 1) : ig3;
      for (int ig2 = 0; ig2 <= 2 * nb[1]; ig2++) {
        int g2 = ig2 > nb[1] ? (int)(ig2 - 2 * nb[1] - 1) : ig2;
        for (int ig1 = 0; ig1 <= 2 * nb[0]; ig1++) {
          int g1 = ig1 > nb[0] ? (int)(ig1 - 2 * nb[0] - 1) : ig1;
          double s[3];
          for (int j = 0; j < 3; j++)
            s[j] = (kp->k[0] + g1) * b1[j] + (kp->k[1] + g2) * b2[j] + (kp->k[2] + g3) * b3[j];
          double gt = mag3(s);
          if (gt * gt / 0.262465831 <= encut) {
            if (ncnt < nplane) {
              kp->Gs[3 * ncnt] = g1;
              kp->Gs[3 * ncnt + 1] = g2;
              kp->Gs[3 * ncnt + 2] = g3;
            }
            ncnt++;
            wf->G_bounds[0] = std::min(wf->G_bounds[0], g1);
            wf->G_bounds[1] = std::max(wf->G_bounds[1], g1);
            wf->G_bounds[2] = std::min(wf->G_bounds[2], g2);
            wf->G_bounds[3] = std::max(wf->G_bounds[3], g2);
            wf->G_bounds[4] = std::min(wf->G_bounds[4], g3);
            wf->G_bounds[5] = std::max(wf->G_bounds[5], g3);
          }
        }
      }
    }
    if (ncnt * 2 == nplane) {
      // spinor: the G list repeats for the second component (reader.c:274-281)
      wf->is_ncl = 1;
      memcpy(kp->Gs + 3 * ncnt, kp->Gs, sizeof(int) * 3 * ncnt);
    } else if (ncnt != nplane) {
      set_error(PP_ERR_IO, "WAVECAR: plane-wave count " + std::to_string(nplane) + " differs from the " +
                               std::to_string(ncnt) + " G vectors inside the cutoff");
      free_pswf(wf);
      return nullptr;
    }
    for (int b = 0; b < nband; b++) {
      irec++;
      if (!src.read((irec + 2) * nrecl, cbuf.data(), (size_t)nplane * 8)) {
        set_error(PP_ERR_IO, "WAVECAR: truncated band record");
        free_pswf(wf);
        return nullptr;
      }
      memcpy(kp->bands[b]->Cs, cbuf.data(), (size_t)nplane * 8);
    }
    kp->weight = kpt_weights ? kpt_weights[iwk % nwk] : 1.0 / nwk;
  }
  return wf;
}

}  // namespace pp

using namespace pp;

extern "C" {

int pp_set_precision(int precision) {
  PP_REQUIRE(precision == PP_FP64 || precision == PP_FP32 || precision == PP_BF16X3 ||
                 precision == PP_BF16X6 || precision == PP_FP16X3,
             PP_ERR_ARG,
             "pp_set_precision: unknown precision");
  std::lock_guard<std::mutex> lk(g_mu);
  g_precision = precision;
  return PP_OK;
}
int pp_get_precision(void) { return g_precision; }

pswf_t *read_wavefunctions(char *filename, double *kpt_weights) {
  Src s;
  s.fp = fopen(filename, "rb");
  if (!s.fp) {
    set_error(PP_ERR_IO, std::string("cannot open ") + filename);
    return nullptr;
  }
  pswf_t *wf = read_wavecar_src(s, kpt_weights);
  fclose(s.fp);
  return wf;
}

pswf_t *read_wavefunctions_from_str(char *start, double *kpt_weights) {
  if (!start) {
    set_error(PP_ERR_ARG, "read_wavefunctions_from_str: null buffer");
    return nullptr;
  }
  Src s;
  s.mem = start;
  return read_wavecar_src(s, kpt_weights);
}

pswf_t *pp_pswf_from_arrays(int nspin, int nwk, int nband, double encut, const double *lattice, const double *kpts,
                            const double *weights, const int *nplane, const int *Gs, const pp_cfloat *coeffs,
                            const double *occ, const double *energy) {
  if ((nspin != 1 && nspin != 2) || nwk <= 0 || nband <= 0 || !lattice || !kpts || !nplane || !Gs || !coeffs) {
    set_error(PP_ERR_ARG, "pp_pswf_from_arrays: bad arguments");
    return nullptr;
  }
  pswf_t *wf = new_pswf(nspin, nwk, nband, encut, lattice);
  long go = 0, co = 0;
  for (int iwk = 0; iwk < nwk * nspin; iwk++) {
    int np = nplane[iwk];
    kpoint_t *kp = new_kpoint(nband, np);
    wf->kpts[iwk] = kp;
    kp->up = (short)(iwk / nwk);
    for (int j = 0; j < 3; j++)
      kp->k[j] = kpts[3 * (iwk % nwk) + j];
    kp->weight = weights ? weights[iwk % nwk] : 1.0 / nwk;
    memcpy(kp->Gs, Gs + 3 * go, sizeof(int) * 3 * (size_t)np);
    for (int w = 0; w < np; w++) {
      int g[3] = {Gs[3 * (go + w)], Gs[3 * (go + w) + 1], Gs[3 * (go + w) + 2]};
      for (int a = 0; a < 3; a++) {
        wf->G_bounds[2 * a] = std::min(wf->G_bounds[2 * a], g[a]);
        wf->G_bounds[2 * a + 1] = std::max(wf->G_bounds[2 * a + 1], g[a]);
      }
    }
    // a G list that repeats itself is the two components of a spinor
    // (a collinear list holds distinct vectors; reader.c:274-281)
    if (np > 0 && np % 2 == 0 && !memcmp(Gs + 3 * go, Gs + 3 * (go + np / 2), sizeof(int) * 3 * (size_t)(np / 2)))
      wf->is_ncl = 1;
    go += np;
    for (int b = 0; b < nband; b++) {
      memcpy(kp->bands[b]->Cs, coeffs + co, sizeof(pp_cfloat) * (size_t)np);
      co += np;
      kp->bands[b]->occ = occ ? occ[(long)iwk * nband + b] : 0.0;
      kp->bands[b]->energy = energy ? energy[(long)iwk * nband + b] : 0.0;
    }
  }
  return wf;
}

void free_ppot_list(ppot_t *pps, int length) {
  if (!pps)
    return;
  for (int i = 0; i < length; i++) {
    ppot_t *pp = pps + i;
    for (int k = 0; pp->funcs && k < pp->num_projs; k++) {
      funcset_t *f = pp->funcs + k;
      free(f->proj);
      free(f->aewave);
      free(f->pswave);
      free(f->diffwave);
      free(f->kwave);
      free(f->smooth_diffwave);
      free_spline(f->proj_spline);
      free_spline(f->aewave_spline);
      free_spline(f->pswave_spline);
      free_spline(f->diffwave_spline);
      free_spline(f->kwave_spline);
      free_spline(f->smooth_diffwave_spline);
    }
    free(pp->funcs);
    free(pp->wave_grid);
    free(pp->kwave_grid);
    free(pp->proj_grid);
    free(pp->smooth_grid);
    free(pp->pspw_overlap_matrix);
    free(pp->aepw_overlap_matrix);
    free(pp->diff_overlap_matrix);
  }
  free(pps);
}

void free_pswf(pswf_t *wf) {
  if (!wf)
    return;
  std::lock_guard<std::mutex> lk(g_mu);
  for (int i = 0; wf->kpts && i < wf->nwk * wf->nspin; i++) {
    kpoint_t *kp = wf->kpts[i];
    if (!kp)
      continue;
    for (int b = 0; b < kp->num_bands; b++) {
      band_t *bd = kp->bands[b];
      if (!bd)
        continue;
      free(bd->Cs);
      free_projection_list(bd->projections, wf->num_sites);
      free_projection_list(bd->wave_projections, wf->wp_num);
      free_projection_list(bd->up_projections, wf->num_sites);
      free_projection_list(bd->down_projections, wf->num_sites);
      free(bd->CRs);
      free(bd->CAs);
      free(bd);
    }
    free(kp->Gs);
    free(kp->bands);
    free(kp->k);
    free(kp);
  }
  if (wf->overlaps) {
    for (int i = 0; i < wf->num_aug_overlap_sites; i++)
      free(wf->overlaps[i]);
    free(wf->overlaps);
  }
  free(wf->num_projs);
  free(wf->kpts);
  free(wf->G_bounds);
  free(wf->lattice);
  free(wf->reclattice);
  if (wf->pps)
    free_ppot_list(wf->pps, wf->num_elems);
  free(wf->dcoords);
  free(wf->fftg);
  delete (WfDev *)wf->dev;
  free(wf);
}

double *get_occs(pswf_t *wf) {
  int nk = wf->nwk * wf->nspin;
  double *o = (double *)malloc(sizeof(double) * (size_t)nk * wf->nband);
  for (int k = 0; k < nk; k++)
    for (int b = 0; b < wf->nband; b++)
      o[(long)b * nk + k] = wf->kpts[k]->bands[b]->occ;
  return o;
}
int get_nband(pswf_t *wf) { return wf->nband; }
int get_nwk(pswf_t *wf) { return wf->nwk; }
int get_nspin(pswf_t *wf) { return wf->nspin; }
int is_ncl(pswf_t *wf) { return wf->is_ncl; }
double get_encut(pswf_t *wf) { return wf->encut; }
double get_energy(pswf_t *wf, int band, int kpt, int spin) { return wf->kpts[kpt + spin * wf->nwk]->bands[band]->energy; }
double get_occ(pswf_t *wf, int band, int kpt, int spin) { return wf->kpts[kpt + spin * wf->nwk]->bands[band]->occ; }
void set_num_sites(pswf_t *wf, int nsites) { wf->num_sites = nsites; }
void free_ptr(void *ptr) { free(ptr); }

ppot_t *get_projector_list(int num_els, int *labels, int *ls, double *wave_grids, double *projectors,
                           double *aewaves, double *pswaves, double *rmaxs, double grid_encut) {
  if (num_els <= 0 || !labels || !ls || !wave_grids || !projectors || !aewaves || !pswaves || !rmaxs) {
    set_error(PP_ERR_ARG, "get_projector_list: bad arguments");
    return nullptr;
  }
  ppot_t *pps = (ppot_t *)calloc(num_els, sizeof(ppot_t));
  long wgt = 0, wt = 0, pt = 0, lnum = 0;
  for (int i = 0; i < num_els; i++) {
    ppot_t *pp = pps + i;
    pp->num_projs = labels[4 * i + 1];
    pp->rmax = rmaxs[i];
    pp->proj_gridsize = labels[4 * i + 2];
    pp->wave_gridsize = labels[4 * i + 3];
    int nw = pp->wave_gridsize, np = pp->proj_gridsize;
    pp->wave_grid = (double *)malloc(sizeof(double) * nw);
    pp->kwave_grid = (double *)calloc(nw, sizeof(double));
    memcpy(pp->wave_grid, wave_grids + wgt, sizeof(double) * nw);
    wgt += nw;
    pp->proj_grid = (double *)malloc(sizeof(double) * np);
    for (int j = 0; j < np; j++)
      pp->proj_grid[j] = pp->rmax / np * j;  // projector.c:56
    pp->wave_rmax = pp->wave_grid[nw - 1];
    pp->smooth_grid = (double *)malloc(sizeof(double) * np);
    for (int j = 0; j < np; j++)
      pp->smooth_grid[j] = pp->wave_rmax / np * j;
    pp->funcs = (funcset_t *)calloc(pp->num_projs, sizeof(funcset_t));
    for (int k = 0; k < pp->num_projs; k++) {
      funcset_t *f = pp->funcs + k;
      f->l = ls[lnum++];
      if (f->l > pp->lmax)
        pp->lmax = f->l;
      pp->total_projs += 2 * f->l + 1;
      f->proj = (double *)malloc(sizeof(double) * np);
      f->aewave = (double *)malloc(sizeof(double) * nw);
      f->pswave = (double *)malloc(sizeof(double) * nw);
      f->diffwave = (double *)malloc(sizeof(double) * nw);
      memcpy(f->proj, projectors + pt, sizeof(double) * np);
      pt += np;
      for (int j = 0; j < nw; j++) {
        f->aewave[j] = aewaves[wt + j];
        f->pswave[j] = pswaves[wt + j];
        f->diffwave[j] = aewaves[wt + j] - pswaves[wt + j];
      }
      wt += nw;
      f->proj_spline = spline3(pp->proj_grid, f->proj, np);
      f->aewave_spline = spline3(pp->wave_grid, f->aewave, nw);
      f->pswave_spline = spline3(pp->wave_grid, f->pswave, nw);
      f->diffwave_spline = spline3(pp->wave_grid, f->diffwave, nw);
    }
    // projector.c:114-163: transform of every partial-wave difference and
    // its plane-wave-filtered form on the smooth grid (the O_R/O_S tables)
    for (int k = 0; k < pp->num_projs; k++) {
      funcset_t *f = pp->funcs + k;
      f->kwave = (double *)malloc(sizeof(double) * nw);
      f->smooth_diffwave = (double *)malloc(sizeof(double) * np);
      if (smooth_diffwave_host(nw, pp->wave_grid, f->diffwave, f->l, pp->lmax, grid_encut, np, f->smooth_diffwave,
                               pp->kwave_grid, f->kwave) != PP_OK) {
        free_ppot_list(pps, i + 1);
        set_error(PP_ERR_ARG, "get_projector_list: partial-wave transform failed");
        return nullptr;
      }
      f->kwave_spline = spline3(pp->kwave_grid, f->kwave, nw);
      f->smooth_diffwave_spline = spline3(pp->smooth_grid, f->smooth_diffwave, np);
    }
    // projector.c:507-558: same-l spline integrals, symmetric
    int n = pp->num_projs;
    pp->pspw_overlap_matrix = (double *)calloc((size_t)n * n, sizeof(double));
    pp->aepw_overlap_matrix = (double *)calloc((size_t)n * n, sizeof(double));
    pp->diff_overlap_matrix = (double *)calloc((size_t)n * n, sizeof(double));
    std::vector<double> prod(nw);
    for (int a = 0; a < n; a++)
      for (int b = a; b < n; b++) {
        if (pp->funcs[a].l != pp->funcs[b].l)
          continue;
        for (int t = 0; t < 3; t++) {
          const funcset_t *fa = pp->funcs + a, *fb = pp->funcs + b;
          for (int j = 0; j < nw; j++)
            prod[j] = t == 0   ? fa->pswave[j] * fb->pswave[j]
                      : t == 1 ? fa->aewave[j] * fb->aewave[j]
                               : fa->diffwave[j] * fb->diffwave[j];
          double **s = spline3(pp->wave_grid, prod.data(), nw);
          double v = spline_integral_host(pp->wave_grid, prod.data(), s, nw);
          free_spline(s);
          double *M = t == 0 ? pp->pspw_overlap_matrix : t == 1 ? pp->aepw_overlap_matrix : pp->diff_overlap_matrix;
          M[a * n + b] = v;
          M[b * n + a] = v;
        }
      }
  }
  return pps;
}

}  // extern "C"

namespace pp {

// Species arrays of a ppot_t list for pp_sphere_create: the projectors
// (smooth = 0) or the filtered partial-wave differences on the smooth grid
// with radius wave_rmax (smooth = 1, utils.c:601-604 and :674-679).
struct SpeciesArrays {
  std::vector<int> nchan, ngrid, ls;
  std::vector<double> rmax, vals, dO, ident;
};

static void species_arrays(const ppot_t *pps, int num_elems, int smooth, SpeciesArrays &a) {
  a.nchan.assign(num_elems, 0);
  a.ngrid.assign(num_elems, 0);
  a.rmax.assign(num_elems, 0.0);
  for (int e = 0; e < num_elems; e++) {
    const ppot_t &pp = pps[e];
    a.nchan[e] = pp.num_projs;
    a.ngrid[e] = pp.proj_gridsize;
    a.rmax[e] = smooth ? pp.wave_rmax : pp.rmax;
    for (int ch = 0; ch < pp.num_projs; ch++) {
      a.ls.push_back(pp.funcs[ch].l);
      const double *v = smooth ? pp.funcs[ch].smooth_diffwave : pp.funcs[ch].proj;
      a.vals.insert(a.vals.end(), v, v + pp.proj_gridsize);
    }
    for (int i = 0; i < pp.num_projs * pp.num_projs; i++) {
      a.dO.push_back(pp.aepw_overlap_matrix[i] - pp.pspw_overlap_matrix[i]);
      a.ident.push_back(i % (pp.num_projs + 1) == 0 ? 1.0 : 0.0);
    }
  }
}

// <table_j | psi~_nk> for every band and k of wf at the given sites (grid
// and lattice of wf), into out c128 [nkt][nband][ptot].  For a spinor
// wavefunction out gets the up component and out_dn (when given) the down
// one: each half of the 2 W coefficients is projected with the first W G
// vectors (projector.c:359-392).
static int project_sites(pp_ctx *ctx, pswf_t *wf, const SpeciesArrays &a, int num_elems, int nsites,
                         const int *labels, const double *coords, int prec, DevBuf &out, int &ptot,
                         std::vector<int> &proj_off, DevBuf *out_dn = nullptr) {
  WfDev *d = dev_of(wf);
  pp_sphere *sp = nullptr;
  PP_TRY(pp_sphere_create(ctx, &sp, wf->lattice, wf->fftg, nsites, labels, coords, num_elems, a.nchan.data(),
                          a.ls.data(), a.rmax.data(), a.ngrid.data(), a.vals.data(), prec));
  ptot = 0;
  proj_off.assign(nsites + 1, 0);
  pp_sphere_info(sp, &ptot, nullptr, proj_off.data());
  int st = out.ensure(sizeof(pp_cdouble) * (size_t)d->nkt * d->nband * std::max(1, ptot));
  std::vector<double> kf(3 * (size_t)d->nkt);
  for (int k = 0; k < d->nkt; k++)
    for (int j = 0; j < 3; j++)
      kf[3 * k + j] = wf->kpts[k]->k[j];
  if (st == PP_OK && out_dn)
    st = out_dn->ensure(sizeof(pp_cdouble) * (size_t)d->nkt * d->nband * std::max(1, ptot));
  if (st == PP_OK && !wf->is_ncl) {
    st = pp_setup_projections_batched(ctx, sp, d->nkt, kf.data(), d->gofs.data(), (const int *)d->G.p, d->C.p,
                                      d->ldc, d->nband, out.p, prec);
  } else if (st == PP_OK) {
    for (int comp = 0; comp < (out_dn ? 2 : 1) && st == PP_OK; comp++) {
      DevBuf &dst = comp ? *out_dn : out;
      for (int k = 0; k < d->nkt && st == PP_OK; k++) {
        const int W = d->nw[k] / 2;
        const int gk[2] = {0, W};
        const pp_cfloat *Ck = (const pp_cfloat *)d->C.p + (size_t)k * d->nband * d->ldc + (size_t)comp * W;
        st = pp_setup_projections_batched(ctx, sp, 1, kf.data() + 3 * k, gk, (const int *)d->G.p + 3 * (size_t)d->gofs[k],
                                          Ck, d->ldc, d->nband,
                                          (pp_cdouble *)dst.p + (size_t)k * d->nband * std::max(1, ptot), prec);
      }
    }
  }
  if (st == PP_OK)
    st = pp_ctx_synchronize(ctx);
  pp_sphere_destroy(sp);
  return st;
}

// Host projection_t lists of the reference (per band and k) from a device
// table [nkt][nband][ptot], stored as band->projections, wave_projections,
// up_projections or down_projections.
enum ListKind { L_PROJ, L_WAVE, L_UP, L_DOWN };
static int fill_host_lists(pswf_t *wf, const DevBuf &tab, int ptot, int nsites, const int *labels,
                           const std::vector<int> &proj_off, const ppot_t *pps, ListKind kind) {
  WfDev *d = dev_of(wf);
  std::vector<pp_cdouble> host((size_t)d->nkt * d->nband * std::max(1, ptot));
  PP_CHECK_HIP(hipMemcpy(host.data(), tab.p, sizeof(pp_cdouble) * host.size(), hipMemcpyDeviceToHost));
  for (int k = 0; k < d->nkt; k++)
    for (int b = 0; b < d->nband; b++) {
      projection_t *pl = (projection_t *)calloc(std::max(1, nsites), sizeof(projection_t));
      const pp_cdouble *row = host.data() + ((size_t)k * d->nband + b) * ptot;
      for (int s = 0; s < nsites; s++) {
        const ppot_t &pp = pps[labels[s]];
        int P = proj_off[s + 1] - proj_off[s];
        pl[s].num_projs = pp.num_projs;
        pl[s].total_projs = P;
        pl[s].ns = (int *)malloc(sizeof(int) * P);
        pl[s].ls = (int *)malloc(sizeof(int) * P);
        pl[s].ms = (int *)malloc(sizeof(int) * P);
        pl[s].overlaps = (pp_cdouble *)malloc(sizeof(pp_cdouble) * P);
        int p = 0;
        for (int ch = 0; ch < pp.num_projs; ch++) {
          int l = pp.funcs[ch].l;
          for (int m = -l; m <= l; m++, p++) {
            pl[s].ns[p] = ch;
            pl[s].ls[p] = l;
            pl[s].ms[p] = m;
            pl[s].overlaps[p] = row[proj_off[s] + p];
          }
        }
      }
      band_t *bd = wf->kpts[k]->bands[b];
      (kind == L_WAVE ? bd->wave_projections
                      : kind == L_UP ? bd->up_projections : kind == L_DOWN ? bd->down_projections : bd->projections) = pl;
    }
  return PP_OK;
}

static unsigned long long hash_ints(unsigned long long h, const int *v, int n) {
  for (int i = 0; i < n; i++)
    h = (h ^ (unsigned)v[i]) * 1099511628211ull;
  return (h ^ (unsigned long long)n) * 1099511628211ull;
}

}  // namespace pp

extern "C" {

int setup_projections(pswf_t *wf, ppot_t *pps, int num_elems, int num_sites, int *fftg, int *labels,
                      double *coords) {
  PP_REQUIRE(wf && pps && num_elems > 0 && num_sites >= 0 && fftg && (num_sites == 0 || (labels && coords)),
             PP_ERR_ARG, "setup_projections: bad arguments");
  std::lock_guard<std::mutex> lk(g_mu);
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  for (int s = 0; s < num_sites; s++)
    PP_REQUIRE(labels[s] >= 0 && labels[s] < num_elems, PP_ERR_ARG, "setup_projections: label out of range");
  // release projections of an earlier setup
  for (int k = 0; k < wf->nwk * wf->nspin; k++)
    for (int b = 0; b < wf->nband; b++) {
      band_t *bd = wf->kpts[k]->bands[b];
      free_projection_list(bd->projections, wf->num_sites);
      free_projection_list(bd->up_projections, wf->num_sites);
      free_projection_list(bd->down_projections, wf->num_sites);
      bd->projections = bd->up_projections = bd->down_projections = nullptr;
    }
  if (wf->pps && wf->pps != pps)
    free_ppot_list(wf->pps, wf->num_elems);
  wf->pps = pps;
  wf->num_elems = num_elems;
  wf->num_sites = num_sites;
  if (!wf->fftg)
    wf->fftg = (int *)malloc(3 * sizeof(int));
  memcpy(wf->fftg, fftg, 3 * sizeof(int));
  PP_TRY(ensure_upload(ctx, wf));
  WfDev *d = dev_of(wf);
  SpeciesArrays a;
  species_arrays(pps, num_elems, 0, a);
  int prec = g_precision == PP_FP64 ? PP_FP64 : PP_FP32;
  int ptot = 0;
  std::vector<int> proj_off;
  PP_TRY(project_sites(ctx, wf, a, num_elems, num_sites, labels, coords, prec, d->proj, ptot, proj_off,
                       wf->is_ncl ? &d->proj_dn : nullptr));
  d->ptot = ptot;
  d->nsites = num_sites;
  d->proj_off = proj_off;
  d->labels.assign(labels, labels + num_sites);
  d->sp_nchan = a.nchan;
  d->sp_ls = a.ls;
  d->sp_dO = a.dO;
  d->proj_precision = prec;
  d->proj_uid = next_uid();
  d->aug_cache.clear();
  if (wf->is_ncl) {
    // onto_projector_ncl (projector.c:359-392): up and down lists
    PP_TRY(fill_host_lists(wf, d->proj, ptot, num_sites, labels, proj_off, pps, L_UP));
    return fill_host_lists(wf, d->proj_dn, ptot, num_sites, labels, proj_off, pps, L_DOWN);
  }
  return fill_host_lists(wf, d->proj, ptot, num_sites, labels, proj_off, pps, L_PROJ);
}

// projector.c:674-712 (and :797-840): off-site overlaps of the N_RS pairs
// from the transforms of the partial-wave differences (radial.c:116-196),
// kept on wf_S as the reference does.
static int setup_offsite(pswf_t *wf_R, pswf_t *wf_S, const int *labels_R, const int *labels_S, const double *coords_R,
                         const double *coords_S, const int *N_RS_R, const int *N_RS_S, int num_N_RS) {
  for (int t = 0; t < num_N_RS; t++)
    PP_REQUIRE(N_RS_R[t] >= 0 && N_RS_R[t] < wf_R->num_sites && N_RS_S[t] >= 0 && N_RS_S[t] < wf_S->num_sites,
               PP_ERR_ARG, "overlap setup: N_RS site out of range");
  if (wf_S->overlaps) {
    for (int i = 0; i < wf_S->num_aug_overlap_sites; i++)
      free(wf_S->overlaps[i]);
    free(wf_S->overlaps);
    wf_S->overlaps = nullptr;
  }
  free(wf_S->dcoords);
  wf_S->dcoords = nullptr;
  WfDev *dS = dev_of(wf_S);
  dS->on_n = num_N_RS;
  dS->on_of = dev_of(wf_R)->uid;
  dS->on_R.assign(N_RS_R, N_RS_R + num_N_RS);
  dS->on_S.assign(N_RS_S, N_RS_S + num_N_RS);
  dS->on_ov.clear();
  dS->on_off.assign(1, 0);
  dS->on_dc.assign(3 * (size_t)num_N_RS, 0.0);
  dS->aug_cache.clear();
  if (num_N_RS > 0) {
    wf_S->overlaps = (pp_cdouble **)calloc(num_N_RS, sizeof(pp_cdouble *));
    wf_S->dcoords = (double *)malloc(3 * sizeof(double) * num_N_RS);
  }
  for (int t = 0; t < num_N_RS; t++) {
    const int s1 = N_RS_R[t], s2 = N_RS_S[t];
    const ppot_t &p1 = wf_R->pps[labels_R[s1]];
    const ppot_t &p2 = wf_S->pps[labels_S[s2]];
    PP_REQUIRE(p1.funcs[0].kwave && p2.funcs[0].kwave, PP_ERR_ARG,
               "overlap_setup_real: projector list without partial-wave transforms");
    // minimum-image cartesian vector from the R site to the S site
    double best = INFINITY, dc[3] = {0, 0, 0};
    for (int a = -1; a <= 1; a++)
      for (int b = -1; b <= 1; b++)
        for (int c = -1; c <= 1; c++) {
          const double f[3] = {coords_S[3 * s2] + a - coords_R[3 * s1], coords_S[3 * s2 + 1] + b - coords_R[3 * s1 + 1],
                               coords_S[3 * s2 + 2] + c - coords_R[3 * s1 + 2]};
          double v[3];
          for (int j = 0; j < 3; j++)
            v[j] = f[0] * wf_R->lattice[j] + f[1] * wf_R->lattice[3 + j] + f[2] * wf_R->lattice[6 + j];
          const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
          if (r < best) {
            best = r;
            dc[0] = v[0];
            dc[1] = v[1];
            dc[2] = v[2];
          }
        }
    memcpy(wf_S->dcoords + 3 * t, dc, sizeof(dc));
    memcpy(dS->on_dc.data() + 3 * t, dc, sizeof(dc));
    std::vector<pp_cdouble> blk((size_t)p1.total_projs * p2.total_projs);
    int tj = 0;
    for (int j = 0; j < p1.num_projs; j++) {
      const funcset_t &f1 = p1.funcs[j];
      for (int m1 = -f1.l; m1 <= f1.l; m1++, tj++) {
        int tk = 0;
        for (int k = 0; k < p2.num_projs; k++) {
          const funcset_t &f2 = p2.funcs[k];
          for (int m2 = -f2.l; m2 <= f2.l; m2++, tk++)
            blk[(size_t)tj * p2.total_projs + tk] =
                std::conj(offsite_overlap_host(dc, p1.kwave_grid, f1.kwave, f1.kwave_spline, p1.wave_gridsize,
                                               p2.kwave_grid, f2.kwave, f2.kwave_spline, p2.wave_gridsize, f1.l, m1,
                                               f2.l, m2));
        }
      }
    }
    wf_S->overlaps[t] = (pp_cdouble *)malloc(sizeof(pp_cdouble) * blk.size());
    memcpy(wf_S->overlaps[t], blk.data(), sizeof(pp_cdouble) * blk.size());
    dS->on_ov.insert(dS->on_ov.end(), blk.begin(), blk.end());
    dS->on_off.push_back((long)dS->on_ov.size());
  }
  wf_S->num_aug_overlap_sites = num_N_RS;
  wf_R->num_aug_overlap_sites = num_N_RS;
  return PP_OK;
}

int overlap_setup_real(pswf_t *wf_R, pswf_t *wf_S, int *labels_R, int *labels_S, double *coords_R,
                       double *coords_S, int *N_R, int *N_S, int *N_RS_R, int *N_RS_S, int num_N_R, int num_N_S,
                       int num_N_RS) {
  PP_REQUIRE(wf_R && wf_S && num_N_R >= 0 && num_N_S >= 0 && num_N_RS >= 0, PP_ERR_ARG,
             "overlap_setup_real: bad arguments");
  PP_REQUIRE(num_N_RS == 0 || (N_RS_R && N_RS_S && labels_R && labels_S && coords_R && coords_S), PP_ERR_ARG,
             "overlap_setup_real: null N_RS data");
  PP_REQUIRE(!wf_R->is_ncl && !wf_S->is_ncl, PP_ERR_UNSUPPORTED,
             "overlap_setup_real: noncollinear wavefunctions cannot be projected (projector.py:74-75)");
  PP_REQUIRE(wf_R->pps && wf_S->pps && wf_R->fftg && wf_S->fftg, PP_ERR_ARG,
             "overlap_setup_real: call setup_projections on both wavefunctions first");
  PP_REQUIRE(num_N_R == 0 || (N_R && labels_R && coords_R), PP_ERR_ARG, "overlap_setup_real: null N_R data");
  PP_REQUIRE(num_N_S == 0 || (N_S && labels_S && coords_S), PP_ERR_ARG, "overlap_setup_real: null N_S data");
  for (int s = 0; s < num_N_R; s++)
    PP_REQUIRE(N_R[s] >= 0 && N_R[s] < wf_R->num_sites && labels_R[N_R[s]] >= 0 &&
                   labels_R[N_R[s]] < wf_R->num_elems,
               PP_ERR_ARG, "overlap_setup_real: N_R site out of range");
  for (int s = 0; s < num_N_S; s++)
    PP_REQUIRE(N_S[s] >= 0 && N_S[s] < wf_S->num_sites && labels_S[N_S[s]] >= 0 &&
                   labels_S[N_S[s]] < wf_S->num_elems,
               PP_ERR_ARG, "overlap_setup_real: N_S site out of range");
  for (int e = 0; e < wf_R->num_elems; e++)
    PP_REQUIRE(num_N_R == 0 || wf_R->pps[e].funcs[0].smooth_diffwave, PP_ERR_ARG,
               "overlap_setup_real: projector list without filtered partial waves");
  for (int e = 0; e < wf_S->num_elems; e++)
    PP_REQUIRE(num_N_S == 0 || wf_S->pps[e].funcs[0].smooth_diffwave, PP_ERR_ARG,
               "overlap_setup_real: projector list without filtered partial waves");
  std::lock_guard<std::mutex> lk(g_mu);
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  PP_TRY(ensure_upload(ctx, wf_R));
  PP_TRY(ensure_upload(ctx, wf_S));
  int prec = g_precision == PP_FP64 ? PP_FP64 : PP_FP32;
  // projector.c:604-672: S's bands onto R's filtered partial waves at the
  // N_R sites, and R's bands onto S's at the N_S sites
  struct Side {
    pswf_t *wf;        // whose bands are projected
    pswf_t *owner;     // whose sites and species
    int n;
    const int *sites;
    const int *labels;
    const double *coords;
  } sides[2] = {{wf_S, wf_R, num_N_R, N_R, labels_R, coords_R}, {wf_R, wf_S, num_N_S, N_S, labels_S, coords_S}};
  for (auto &sd : sides) {
    WfDev *d = dev_of(sd.wf);
    for (int k = 0; k < sd.wf->nwk * sd.wf->nspin; k++)
      for (int b = 0; b < sd.wf->nband; b++) {
        free_projection_list(sd.wf->kpts[k]->bands[b]->wave_projections, sd.wf->wp_num);
        sd.wf->kpts[k]->bands[b]->wave_projections = nullptr;
      }
    sd.wf->wp_num = sd.n;
    d->wp_n = sd.n;
    d->wp_of = dev_of(sd.owner)->uid;
    d->aug_cache.clear();
    d->wp_hash = hash_ints(1469598103934665603ull, sd.sites, sd.n);
    std::vector<int> lab(sd.n);
    std::vector<double> crd(3 * (size_t)sd.n);
    for (int s = 0; s < sd.n; s++) {
      lab[s] = sd.labels[sd.sites[s]];
      for (int j = 0; j < 3; j++)
        crd[3 * s + j] = sd.coords[3 * sd.sites[s] + j];
    }
    SpeciesArrays a;
    species_arrays(sd.owner->pps, sd.owner->num_elems, 1, a);
    d->wp_labels = lab;
    d->wp_nchan = a.nchan;
    d->wp_ls = a.ls;
    if (sd.n == 0) {
      d->wp_ptot = 0;
      d->wp_off.assign(1, 0);
      continue;
    }
    // the smooth tables live on the projected structure's grid
    PP_TRY(project_sites(ctx, sd.wf, a, sd.owner->num_elems, sd.n, lab.data(), crd.data(), prec, d->wp, d->wp_ptot,
                         d->wp_off));
    PP_TRY(fill_host_lists(sd.wf, d->wp, d->wp_ptot, sd.n, lab.data(), d->wp_off, sd.owner->pps, L_WAVE));
  }
  PP_TRY(setup_offsite(wf_R, wf_S, labels_R, labels_S, coords_R, coords_S, N_RS_R, N_RS_S, num_N_RS));
  return PP_OK;
}

int pp_get_projections(pswf_t *wf, int band, int kidx, pp_cdouble *out) {
  PP_REQUIRE(wf && wf->dev && out, PP_ERR_ARG, "pp_get_projections: no projections (call setup_projections)");
  WfDev *d = (WfDev *)wf->dev;
  PP_REQUIRE(d->proj_precision >= 0, PP_ERR_ARG, "pp_get_projections: no projections (call setup_projections)");
  PP_REQUIRE(band >= 0 && band < d->nband && kidx >= 0 && kidx < d->nkt, PP_ERR_ARG,
             "pp_get_projections: index out of range");
  PP_CHECK_HIP(hipMemcpy(out, (pp_cdouble *)d->proj.p + ((size_t)kidx * d->nband + band) * d->ptot,
                         sizeof(pp_cdouble) * d->ptot, hipMemcpyDeviceToHost));
  if (wf->is_ncl)  // spinors: the down component follows the up one
    PP_CHECK_HIP(hipMemcpy(out + d->ptot, (pp_cdouble *)d->proj_dn.p + ((size_t)kidx * d->nband + band) * d->ptot,
                           sizeof(pp_cdouble) * d->ptot, hipMemcpyDeviceToHost));
  return PP_OK;
}

int pseudoprojection(pp_cdouble *projections, pswf_t *wf_ref, pswf_t *wf_proj, int BAND_NUM, int flip_spin) {
  PP_REQUIRE(projections && wf_ref && wf_proj, PP_ERR_ARG, "pseudoprojection: null argument");
  PP_REQUIRE(BAND_NUM >= 0 && BAND_NUM < wf_proj->nband, PP_ERR_ARG, "pseudoprojection: band index out of range");
  PP_REQUIRE(wf_ref->nwk == wf_proj->nwk && wf_ref->nspin == wf_proj->nspin, PP_ERR_ARG,
             "pseudoprojection: k-point sets differ");
  std::lock_guard<std::mutex> lk(g_mu);
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  PP_TRY(ensure_upload(ctx, wf_ref));
  PP_TRY(ensure_upload(ctx, wf_proj));
  WfDev *dr = dev_of(wf_ref), *dp = dev_of(wf_proj);
  const int nkt = dr->nkt;
  int flip = (flip_spin && wf_ref->nspin == 2) ? 1 : 0;
  auto key = std::make_tuple(dr->uid, flip, g_precision);
  auto it = dp->overlap_cache.find(key);
  if (it == dp->overlap_cache.end()) {
    // pseudoprojector.c:85: length of the reference band at the unflipped k
    for (int k = 0; k < nkt; k++)
      PP_REQUIRE(dr->nw[k] <= dp->ldc, PP_ERR_ARG, "pseudoprojection: plane-wave bases differ");
    DevBuf out;
    PP_TRY(out.ensure(sizeof(pp_cdouble) * (size_t)nkt * dr->nband * dp->nband));
    PP_TRY(pp_overlap_gemm(ctx, wf_ref->nwk, wf_ref->nspin, flip, dr->nw.data(), dr->C.p, dr->ldc, dr->nband,
                           dp->C.p, dp->ldc, dp->nband, out.p, g_precision));
    PairCache pc;
    pc.nbR = dr->nband;
    pc.nbS = dp->nband;
    pc.m.resize((size_t)nkt * pc.nbR * pc.nbS);
    PP_CHECK_HIP(hipMemcpyAsync(pc.m.data(), out.p, sizeof(pp_cdouble) * pc.m.size(), hipMemcpyDeviceToHost,
                                ctx->stream));
    PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    it = dp->overlap_cache.emplace(key, std::move(pc)).first;
  }
  const PairCache &pc = it->second;
  for (int b = 0; b < pc.nbR; b++)
    for (int k = 0; k < nkt; k++)
      projections[(long)b * nkt + k] = pc.m[((size_t)k * pc.nbR + b) * pc.nbS + BAND_NUM];
  return PP_OK;
}

int compensation_terms(pp_cdouble *overlap, int BAND_NUM, pswf_t *wf_S, pswf_t *wf_R, int num_M, int num_N_R,
                       int num_N_S, int num_N_RS, int *M_R, int *M_S, int *N_R, int *N_S, int *N_RS_R, int *N_RS_S,
                       int *proj_labels, double *proj_coords, int *ref_labels, double *ref_coords, int *fft_grid,
                       int flip_spin) {
  PP_REQUIRE(wf_R && wf_S, PP_ERR_ARG, "compensation_terms: null wavefunction");
  PP_REQUIRE(!wf_R->is_ncl && !wf_S->is_ncl, PP_ERR_UNSUPPORTED,
             "compensation_terms: noncollinear wavefunctions cannot be projected (projector.py:74-75)");
  (void)proj_labels;
  (void)proj_coords;
  (void)ref_labels;
  (void)ref_coords;
  (void)fft_grid;
  PP_REQUIRE(overlap && wf_S && wf_R && num_M >= 0 && num_N_R >= 0 && num_N_S >= 0, PP_ERR_ARG,
             "compensation_terms: bad arguments");
  PP_REQUIRE(num_N_RS >= 0 && (num_N_RS == 0 || (N_RS_R && N_RS_S)), PP_ERR_ARG,
             "compensation_terms: bad N_RS lists");
  PP_REQUIRE(BAND_NUM >= 0 && BAND_NUM < wf_S->nband, PP_ERR_ARG, "compensation_terms: band index out of range");
  PP_REQUIRE(num_M == 0 || (M_R && M_S), PP_ERR_ARG, "compensation_terms: null site list");
  PP_REQUIRE(num_N_R == 0 || N_R, PP_ERR_ARG, "compensation_terms: null N_R list");
  PP_REQUIRE(num_N_S == 0 || N_S, PP_ERR_ARG, "compensation_terms: null N_S list");
  std::lock_guard<std::mutex> lk(g_mu);
  WfDev *dS = (WfDev *)wf_S->dev, *dR = (WfDev *)wf_R->dev;
  PP_REQUIRE(dS && dR && dS->proj_precision >= 0 && dR->proj_precision >= 0, PP_ERR_ARG,
             "compensation_terms: call setup_projections on both wavefunctions first");
  const int nkt = dR->nkt;
  for (int t = 0; t < num_M; t++)
    PP_REQUIRE(M_R[t] >= 0 && M_R[t] < dR->nsites && M_S[t] >= 0 && M_S[t] < dS->nsites, PP_ERR_ARG,
               "compensation_terms: site index out of range");
  for (int t = 0; t < num_N_R; t++)
    PP_REQUIRE(N_R[t] >= 0 && N_R[t] < dR->nsites, PP_ERR_ARG, "compensation_terms: N_R site out of range");
  for (int t = 0; t < num_N_S; t++)
    PP_REQUIRE(N_S[t] >= 0 && N_S[t] < dS->nsites, PP_ERR_ARG, "compensation_terms: N_S site out of range");
  // the wave projections of overlap_setup_real must be those of these lists
  const unsigned long long hR = hash_ints(1469598103934665603ull, N_R, num_N_R);
  const unsigned long long hS = hash_ints(1469598103934665603ull, N_S, num_N_S);
  PP_REQUIRE(num_N_R == 0 || (dS->wp_n == num_N_R && dS->wp_of == dR->uid && dS->wp_hash == hR), PP_ERR_ARG,
             "compensation_terms: call overlap_setup_real with this N_R list first");
  PP_REQUIRE(num_N_S == 0 || (dR->wp_n == num_N_S && dR->wp_of == dS->uid && dR->wp_hash == hS), PP_ERR_ARG,
             "compensation_terms: call overlap_setup_real with this N_S list first");
  if (num_N_RS > 0) {
    bool same = dS->on_n == num_N_RS && dS->on_of == dR->uid;
    for (int t = 0; same && t < num_N_RS; t++)
      same = dS->on_R[t] == N_RS_R[t] && dS->on_S[t] == N_RS_S[t];
    PP_REQUIRE(same, PP_ERR_ARG, "compensation_terms: call overlap_setup_real with these N_RS pairs first");
  }
  if (num_M + num_N_R + num_N_S + num_N_RS == 0)
    return PP_OK;
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  int flip = (flip_spin && wf_R->nspin == 2) ? 1 : 0;
  unsigned long long h = 1469598103934665603ull;
  for (int t = 0; t < num_M; t++) {
    h = (h ^ (unsigned)M_R[t]) * 1099511628211ull;
    h = (h ^ (unsigned)M_S[t]) * 1099511628211ull;
  }
  h ^= (unsigned long long)num_M << 32;
  h = hash_ints(h, N_R, num_N_R);
  h = hash_ints(h, N_S, num_N_S);
  if (num_N_RS > 0) {
    h = hash_ints(h, N_RS_R, num_N_RS);
    h = hash_ints(h, N_RS_S, num_N_RS);
  }
  auto key = std::make_tuple(dR->proj_uid, flip, h);
  auto it = dS->aug_cache.find(key);
  if (it == dS->aug_cache.end()) {
    DevBuf out;
    size_t n = (size_t)nkt * dR->nband * dS->nband;
    PP_TRY(out.ensure(sizeof(pp_cdouble) * n));
    PP_TRY(zero_async(ctx, out.p, sizeof(pp_cdouble) * n));
    // O_M (projector.c:890-910): dO and channel tables of R's species
    PP_TRY(pp_augmentation(ctx, wf_R->nwk, wf_R->nspin, flip, dR->proj.p, dR->nband, dR->ptot, dR->proj_off.data(),
                           dR->labels.data(), dS->proj.p, dS->nband, dS->ptot, dS->proj_off.data(),
                           dS->labels.data(), num_M, M_R, M_S, (int)dR->sp_nchan.size(), dR->sp_nchan.data(),
                           dR->sp_ls.data(), dR->sp_dO.data(), out.p));
    // O_R (:915-924) = sum_i conj(<p_i|R>) <(phi_i - phi~_i)|S> over the N_R
    // sites, and O_S (:929-938) = sum_i conj(<(phi_i - phi~_i)|R>) <p_i|S>
    // over the N_S sites: the same contraction with identity channel
    // matrices (only i = j, equal (n, l, m), contributes)
    auto ident = [](const std::vector<int> &nchan) {
      std::vector<double> I;
      for (int nc : nchan)
        for (int i = 0; i < nc * nc; i++)
          I.push_back(i % (nc + 1) == 0 ? 1.0 : 0.0);
      return I;
    };
    if (num_N_R > 0) {
      std::vector<int> seq(num_N_R);
      for (int t = 0; t < num_N_R; t++)
        seq[t] = t;
      std::vector<double> I = ident(dR->sp_nchan);
      PP_TRY(pp_augmentation(ctx, wf_R->nwk, wf_R->nspin, flip, dR->proj.p, dR->nband, dR->ptot,
                             dR->proj_off.data(), dR->labels.data(), dS->wp.p, dS->nband, dS->wp_ptot,
                             dS->wp_off.data(), dS->wp_labels.data(), num_N_R, N_R, seq.data(),
                             (int)dR->sp_nchan.size(), dR->sp_nchan.data(), dR->sp_ls.data(), I.data(), out.p));
    }
    if (num_N_S > 0) {
      std::vector<int> seq(num_N_S);
      for (int t = 0; t < num_N_S; t++)
        seq[t] = t;
      std::vector<double> I = ident(dS->sp_nchan);
      PP_TRY(pp_augmentation(ctx, wf_R->nwk, wf_R->nspin, flip, dR->wp.p, dR->nband, dR->wp_ptot,
                             dR->wp_off.data(), dR->wp_labels.data(), dS->proj.p, dS->nband, dS->ptot,
                             dS->proj_off.data(), dS->labels.data(), num_N_S, seq.data(), N_S,
                             (int)dS->sp_nchan.size(), dS->sp_nchan.data(), dS->sp_ls.data(), I.data(), out.p));
    }
    // O_N (:940-959) = sum over N_RS pairs of conj(P_R) Ov P_S e^{2 pi i kR . dcoord}
    if (num_N_RS > 0) {
      std::vector<double> kp(3 * (size_t)nkt);
      for (int k = 0; k < nkt; k++)
        for (int j = 0; j < 3; j++)
          kp[3 * k + j] = wf_R->kpts[k]->k[j];
      PP_TRY(pp_augmentation_offsite(ctx, wf_R->nwk, wf_R->nspin, flip, dR->proj.p, dR->nband, dR->ptot,
                                     dR->proj_off.data(), dS->proj.p, dS->nband, dS->ptot, dS->proj_off.data(),
                                     num_N_RS, N_RS_R, N_RS_S, dS->on_ov.data(), dS->on_off.data(),
                                     dS->on_dc.data(), kp.data(), out.p));
    }
    PairCache pc;
    pc.nbR = dR->nband;
    pc.nbS = dS->nband;
    pc.m.resize(n);
    PP_CHECK_HIP(hipMemcpyAsync(pc.m.data(), out.p, sizeof(pp_cdouble) * n, hipMemcpyDeviceToHost, ctx->stream));
    PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    it = dS->aug_cache.emplace(key, std::move(pc)).first;
  }
  const PairCache &pc = it->second;
  for (int b = 0; b < pc.nbR; b++)
    for (int k = 0; k < nkt; k++)
      overlap[(long)b * nkt + k] += pc.m[((size_t)k * pc.nbR + b) * pc.nbS + BAND_NUM];
  return PP_OK;
}

int overlap_setup_recip(pswf_t *wf_R, pswf_t *wf_S, int *labels_R, int *labels_S, double *coords_R,
                        double *coords_S, int *N_R, int *N_S, int *N_RS_R, int *N_RS_S, int num_N_R, int num_N_S,
                        int num_N_RS) {
  PP_REQUIRE(wf_R && wf_S && num_N_R >= 0 && num_N_S >= 0 && num_N_RS >= 0, PP_ERR_ARG,
             "overlap_setup_recip: bad arguments");
  PP_REQUIRE(!wf_R->is_ncl && !wf_S->is_ncl, PP_ERR_UNSUPPORTED,
             "overlap_setup_recip: noncollinear wavefunctions cannot be projected (projector.py:74-75)");
  PP_REQUIRE(wf_R->pps && wf_S->pps && wf_R->fftg && wf_S->fftg && wf_R->dev && wf_S->dev, PP_ERR_ARG,
             "overlap_setup_recip: call setup_projections on both wavefunctions first");
  PP_REQUIRE(num_N_R == 0 || (N_R && labels_R && coords_R), PP_ERR_ARG, "overlap_setup_recip: null N_R data");
  PP_REQUIRE(num_N_S == 0 || (N_S && labels_S && coords_S), PP_ERR_ARG, "overlap_setup_recip: null N_S data");
  PP_REQUIRE(num_N_RS == 0 || (N_RS_R && N_RS_S && labels_R && labels_S && coords_R && coords_S), PP_ERR_ARG,
             "overlap_setup_recip: null N_RS data");
  for (int s = 0; s < num_N_R; s++)
    PP_REQUIRE(N_R[s] >= 0 && N_R[s] < wf_R->num_sites, PP_ERR_ARG, "overlap_setup_recip: N_R site out of range");
  for (int s = 0; s < num_N_S; s++)
    PP_REQUIRE(N_S[s] >= 0 && N_S[s] < wf_S->num_sites, PP_ERR_ARG, "overlap_setup_recip: N_S site out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  const int prec = g_precision == PP_FP64 ? PP_FP64 : PP_FP32;
  // projector.c:727-790: each structure's bands expanded in its own
  // filtered partial waves at its own unmatched sites (R at N_R, S at N_S)
  struct Side {
    pswf_t *wf;
    int n;
    const int *sites;
    const int *labels;
    const double *coords;
  } sides[2] = {{wf_R, num_N_R, N_R, labels_R, coords_R}, {wf_S, num_N_S, N_S, labels_S, coords_S}};
  wf_R->wp_num = 0;
  wf_S->wp_num = 0;
  for (auto &sd : sides) {
    WfDev *d = dev_of(sd.wf);
    for (int k = 0; k < sd.wf->nwk * sd.wf->nspin; k++)
      for (int b = 0; b < sd.wf->nband; b++) {
        free_projection_list(sd.wf->kpts[k]->bands[b]->wave_projections, sd.wf->wp_num);
        sd.wf->kpts[k]->bands[b]->wave_projections = nullptr;
        free(sd.wf->kpts[k]->bands[b]->CAs);
        sd.wf->kpts[k]->bands[b]->CAs = nullptr;
      }
    d->wp_n = 0;
    d->aug_cache.clear();
    d->ca_n = sd.n;
    d->ca_hash = hash_ints(1469598103934665603ull, sd.sites, sd.n);
    if (sd.n == 0)
      continue;
    for (int e = 0; e < sd.wf->num_elems; e++)
      PP_REQUIRE(sd.wf->pps[e].funcs[0].smooth_diffwave, PP_ERR_ARG,
                 "overlap_setup_recip: projector list without filtered partial waves");
    std::vector<int> lab(sd.n), src(sd.n);
    std::vector<double> crd(3 * (size_t)sd.n);
    for (int s = 0; s < sd.n; s++) {
      lab[s] = sd.labels[sd.sites[s]];
      src[s] = d->proj_off[sd.sites[s]];
      for (int j = 0; j < 3; j++)
        crd[3 * s + j] = sd.coords[3 * sd.sites[s] + j];
    }
    SpeciesArrays a;
    species_arrays(sd.wf->pps, sd.wf->num_elems, 1, a);
    pp_sphere *sp = nullptr;
    PP_TRY(pp_sphere_create(ctx, &sp, sd.wf->lattice, sd.wf->fftg, sd.n, lab.data(), crd.data(), sd.wf->num_elems,
                            a.nchan.data(), a.ls.data(), a.rmax.data(), a.ngrid.data(), a.vals.data(), prec));
    d->ca_ld = d->ldc;
    int st = d->ca.ensure(sizeof(pp_cdouble) * (size_t)d->nkt * d->nband * d->ca_ld);
    std::vector<double> kf(3 * (size_t)d->nkt);
    for (int k = 0; k < d->nkt; k++)
      for (int j = 0; j < 3; j++)
        kf[3 * k + j] = sd.wf->kpts[k]->k[j];
    if (st == PP_OK)
      st = zero_async(ctx, d->ca.p, sizeof(pp_cdouble) * (size_t)d->nkt * d->nband * d->ca_ld);
    if (st == PP_OK)
      st = pp_aug_freqs_batched(ctx, sp, d->nkt, kf.data(), d->gofs.data(), (const int *)d->G.p, d->proj.p, d->ptot,
                                src.data(), d->nband, d->ca.p, d->ca_ld, prec);
    if (st == PP_OK)
      st = pp_ctx_synchronize(ctx);
    pp_sphere_destroy(sp);
    PP_TRY(st);
    // host copies in the reference's float complex CAs of every band
    std::vector<pp_cdouble> host((size_t)d->nkt * d->nband * d->ca_ld);
    PP_CHECK_HIP(hipMemcpy(host.data(), d->ca.p, sizeof(pp_cdouble) * host.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < d->nkt; k++)
      for (int b = 0; b < d->nband; b++) {
        const int nw = d->nw[k];
        pp_cfloat *ca = (pp_cfloat *)malloc(sizeof(pp_cfloat) * std::max(1, nw));
        const pp_cdouble *row = host.data() + ((size_t)k * d->nband + b) * d->ca_ld;
        for (int w = 0; w < nw; w++)
          ca[w] = pp_cfloat((float)row[w].real(), (float)row[w].imag());
        sd.wf->kpts[k]->bands[b]->CAs = ca;
      }
  }
  wf_R->wp_num = num_N_S;
  wf_S->wp_num = num_N_R;
  return setup_offsite(wf_R, wf_S, labels_R, labels_S, coords_R, coords_S, N_RS_R, N_RS_S, num_N_RS);
}

int compensation_terms_recip(pp_cdouble *overlap, int BAND_NUM, pswf_t *wf_S, pswf_t *wf_R, int num_M, int num_N_R,
                             int num_N_S, int num_N_RS, int *M_R, int *M_S, int *N_R, int *N_S, int *N_RS_R,
                             int *N_RS_S, int *proj_labels, double *proj_coords, int *ref_labels, double *ref_coords,
                             int *fft_grid, int flip_spin) {
  PP_REQUIRE(wf_R && wf_S, PP_ERR_ARG, "compensation_terms_recip: null wavefunction");
  PP_REQUIRE(!wf_R->is_ncl && !wf_S->is_ncl, PP_ERR_UNSUPPORTED,
             "compensation_terms_recip: noncollinear wavefunctions cannot be projected (projector.py:74-75)");
  (void)proj_labels;
  (void)proj_coords;
  (void)ref_labels;
  (void)ref_coords;
  (void)fft_grid;
  PP_REQUIRE(overlap && wf_S && wf_R && num_M >= 0 && num_N_R >= 0 && num_N_S >= 0 && num_N_RS >= 0, PP_ERR_ARG,
             "compensation_terms_recip: bad arguments");
  PP_REQUIRE(BAND_NUM >= 0 && BAND_NUM < wf_S->nband, PP_ERR_ARG,
             "compensation_terms_recip: band index out of range");
  PP_REQUIRE(num_M == 0 || (M_R && M_S), PP_ERR_ARG, "compensation_terms_recip: null site list");
  PP_REQUIRE(num_N_RS == 0 || (N_RS_R && N_RS_S), PP_ERR_ARG, "compensation_terms_recip: null N_RS lists");
  std::lock_guard<std::mutex> lk(g_mu);
  WfDev *dS = (WfDev *)wf_S->dev, *dR = (WfDev *)wf_R->dev;
  PP_REQUIRE(dS && dR && dS->proj_precision >= 0 && dR->proj_precision >= 0, PP_ERR_ARG,
             "compensation_terms_recip: call setup_projections on both wavefunctions first");
  for (int t = 0; t < num_M; t++)
    PP_REQUIRE(M_R[t] >= 0 && M_R[t] < dR->nsites && M_S[t] >= 0 && M_S[t] < dS->nsites, PP_ERR_ARG,
               "compensation_terms_recip: site index out of range");
  PP_REQUIRE(num_N_R == 0 || N_R, PP_ERR_ARG, "compensation_terms_recip: null N_R list");
  PP_REQUIRE(num_N_S == 0 || N_S, PP_ERR_ARG, "compensation_terms_recip: null N_S list");
  PP_REQUIRE((num_N_R == 0 || (dR->ca_n == num_N_R &&
                               dR->ca_hash == hash_ints(1469598103934665603ull, N_R, num_N_R))) &&
                 (num_N_S == 0 || (dS->ca_n == num_N_S &&
                                   dS->ca_hash == hash_ints(1469598103934665603ull, N_S, num_N_S))),
             PP_ERR_ARG, "compensation_terms_recip: call overlap_setup_recip with these N_R/N_S lists first");
  if (num_N_RS > 0) {
    bool same = dS->on_n == num_N_RS && dS->on_of == dR->uid;
    for (int t = 0; same && t < num_N_RS; t++)
      same = dS->on_R[t] == N_RS_R[t] && dS->on_S[t] == N_RS_S[t];
    PP_REQUIRE(same, PP_ERR_ARG, "compensation_terms_recip: call overlap_setup_recip with these N_RS pairs first");
  }
  const int nkt = dR->nkt;
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  int flip = (flip_spin && wf_R->nspin == 2) ? 1 : 0;
  unsigned long long h = 0x9e3779b97f4a7c15ull;  // distinct from the real-space keys
  h = hash_ints(h, M_R, num_M);
  h = hash_ints(h, M_S, num_M);
  h = hash_ints(h, N_R, num_N_R);
  h = hash_ints(h, N_S, num_N_S);
  h = hash_ints(h, N_RS_R, num_N_RS);
  h = hash_ints(h, N_RS_S, num_N_RS);
  auto key = std::make_tuple(dR->proj_uid, flip, h);
  auto it = dS->aug_cache.find(key);
  if (it == dS->aug_cache.end()) {
    DevBuf out;
    size_t n = (size_t)nkt * dR->nband * dS->nband;
    PP_TRY(out.ensure(sizeof(pp_cdouble) * n));
    PP_TRY(zero_async(ctx, out.p, sizeof(pp_cdouble) * n));
    std::vector<int> amap(nkt), klen(nkt);
    for (int k = 0; k < nkt; k++) {
      amap[k] = flip ? (k < wf_R->nwk ? k + wf_R->nwk : k - wf_R->nwk) : k;
      klen[k] = dR->nw[amap[k]];  // projector.c:1012, 1020: the R k-point's plane-wave count
      PP_REQUIRE(klen[k] <= dS->ldc && klen[k] <= dR->ldc, PP_ERR_ARG,
                 "compensation_terms_recip: plane-wave bases differ");
    }
    // projector.c:1008-1022: cdotc(CA_R, C_S) + cdotc(C_R, CA_S) for every pair
    if (num_N_R > 0)
      PP_TRY(zgemm_nt(ctx, PP_FP64, dR->nband, dS->nband, 0, 0, dR->ca.p, dR->ca_ld, (long)dR->nband * dR->ca_ld, 1,
                      dS->C.p, dS->ldc, (long)dS->nband * dS->ldc, out.p, dS->nband, (long)dR->nband * dS->nband,
                      nkt, amap.data(), nullptr, klen.data(), true));
    if (num_N_S > 0)
      PP_TRY(zgemm_nt(ctx, PP_FP64, dR->nband, dS->nband, 0, 1, dR->C.p, dR->ldc, (long)dR->nband * dR->ldc, 0,
                      dS->ca.p, dS->ca_ld, (long)dS->nband * dS->ca_ld, out.p, dS->nband,
                      (long)dR->nband * dS->nband, nkt, amap.data(), nullptr, klen.data(), true));
    // O_M (:1030-1050) and O_N (:1056-1071), as in compensation_terms
    PP_TRY(pp_augmentation(ctx, wf_R->nwk, wf_R->nspin, flip, dR->proj.p, dR->nband, dR->ptot, dR->proj_off.data(),
                           dR->labels.data(), dS->proj.p, dS->nband, dS->ptot, dS->proj_off.data(),
                           dS->labels.data(), num_M, M_R, M_S, (int)dR->sp_nchan.size(), dR->sp_nchan.data(),
                           dR->sp_ls.data(), dR->sp_dO.data(), out.p));
    if (num_N_RS > 0) {
      std::vector<double> kp(3 * (size_t)nkt);
      for (int k = 0; k < nkt; k++)
        for (int j = 0; j < 3; j++)
          kp[3 * k + j] = wf_R->kpts[k]->k[j];
      PP_TRY(pp_augmentation_offsite(ctx, wf_R->nwk, wf_R->nspin, flip, dR->proj.p, dR->nband, dR->ptot,
                                     dR->proj_off.data(), dS->proj.p, dS->nband, dS->ptot, dS->proj_off.data(),
                                     num_N_RS, N_RS_R, N_RS_S, dS->on_ov.data(), dS->on_off.data(),
                                     dS->on_dc.data(), kp.data(), out.p));
    }
    PairCache pc;
    pc.nbR = dR->nband;
    pc.nbS = dS->nband;
    pc.m.resize(n);
    PP_CHECK_HIP(hipMemcpyAsync(pc.m.data(), out.p, sizeof(pp_cdouble) * n, hipMemcpyDeviceToHost, ctx->stream));
    PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    it = dS->aug_cache.emplace(key, std::move(pc)).first;
  }
  const PairCache &pc = it->second;
  for (int b = 0; b < pc.nbR; b++)
    for (int k = 0; k < nkt; k++)
      overlap[(long)b * nkt + k] += pc.m[((size_t)k * pc.nbR + b) * pc.nbS + BAND_NUM];
  return PP_OK;
}

int fft3d(pp_cdouble *x, int *G_bounds, double *lattice, double *kpt, int *Gs, pp_cfloat *Cs, int num_waves,
          int *fftg) {
  (void)G_bounds;
  (void)kpt;
  PP_REQUIRE(x && lattice && fftg && num_waves >= 0 && (num_waves == 0 || (Gs && Cs)), PP_ERR_ARG,
             "fft3d: bad arguments");
  std::lock_guard<std::mutex> lk(g_mu);
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  long npts = (long)fftg[0] * fftg[1] * fftg[2];
  DevBuf dG, dC, dX;
  PP_TRY(dG.ensure(sizeof(int) * 3 * (size_t)std::max(1, num_waves)));
  PP_TRY(dC.ensure(sizeof(pp_cfloat) * (size_t)std::max(1, num_waves)));
  PP_TRY(dX.ensure(sizeof(pp_cdouble) * (size_t)npts));
  if (num_waves) {
    PP_CHECK_HIP(hipMemcpy(dG.p, Gs, sizeof(int) * 3 * num_waves, hipMemcpyHostToDevice));
    PP_CHECK_HIP(hipMemcpy(dC.p, Cs, sizeof(pp_cfloat) * num_waves, hipMemcpyHostToDevice));
  }
  PP_TRY(pp_fft3d_batched(ctx, lattice, fftg, num_waves, (const int *)dG.p, dC.p, std::max(1, num_waves), 1, dX.p,
                          g_precision == PP_FP64 ? PP_FP64 : PP_FP32));
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  PP_CHECK_HIP(hipMemcpy(x, dX.p, sizeof(pp_cdouble) * npts, hipMemcpyDeviceToHost));
  return PP_OK;
}


// ---- all-electron real-space states (density.c) ----

}  // extern "C"

namespace pp {

// The partial-wave sphere plan of wf on grid fftg for (labels, coords),
// cached on the wavefunction; its projector layout must be the one of the
// projections of setup_projections.
static int rs_plan_for(pp_ctx *ctx, pswf_t *wf, const int *fftg, const int *labels, const double *coords,
                       pp_rsplan **out) {
  WfDev *d = dev_of(wf);
  PP_REQUIRE(wf->pps && d->proj_precision >= 0, PP_ERR_ARG, "realspace: no projections (call setup_projections)");
  const int ns = wf->num_sites;
  for (int s = 0; s < ns; s++)
    PP_REQUIRE(labels[s] >= 0 && labels[s] < wf->num_elems, PP_ERR_ARG, "realspace: label out of range");
  unsigned long long h = 1469598103934665603ull;
  h = hash_ints(h, fftg, 3);
  h = hash_ints(h, labels, ns);
  h = hash_ints(h, (const int *)coords, 6 * ns);
  // the partial waves come from the projector list of the last
  // setup_projections, whose identity proj_uid carries
  h = hash_ints(h, (const int *)&d->proj_uid, (int)(sizeof(d->proj_uid) / sizeof(int)));
  if (d->rs && d->rs_hash == h) {
    *out = d->rs;
    return PP_OK;
  }
  std::vector<int> nchan, ls, wn;
  std::vector<double> wg, dw;
  for (int e = 0; e < wf->num_elems; e++) {
    const ppot_t &pp = wf->pps[e];
    nchan.push_back(pp.num_projs);
    wn.push_back(pp.wave_gridsize);
    wg.insert(wg.end(), pp.wave_grid, pp.wave_grid + pp.wave_gridsize);
    for (int ch = 0; ch < pp.num_projs; ch++) {
      ls.push_back(pp.funcs[ch].l);
      dw.insert(dw.end(), pp.funcs[ch].diffwave, pp.funcs[ch].diffwave + pp.wave_gridsize);
    }
  }
  pp_rsplan *rp = nullptr;
  PP_TRY(pp_rs_plan_create(ctx, &rp, wf->lattice, fftg, ns, labels, coords, wf->num_elems, nchan.data(), ls.data(),
                           wn.data(), wg.data(), dw.data()));
  std::vector<int> po(ns + 1, 0);
  pp_rs_plan_info(rp, nullptr, nullptr, nullptr, nullptr, po.data());
  if (po != d->proj_off) {
    pp_rs_plan_destroy(rp);
    PP_REQUIRE(false, PP_ERR_ARG, "realspace: the sites' projectors differ from those of setup_projections");
  }
  if (d->rs)
    pp_rs_plan_destroy(d->rs);
  d->rs = rp;
  d->rs_hash = h;
  *out = rp;
  return PP_OK;
}

// States (states_dev c128 [nb][N^3]) and/or density (dens_dev +=
// sum_b w_b |psi_b|^2) of bands b0 .. b0+nb-1 at k-index k, spinor
// component comp, on the device.
static int rs_run(pp_ctx *ctx, pswf_t *wf, pp_rsplan *rp, int k, int b0, int nb, int comp, int remove,
                  void *states_dev, long ldo, const double *w, double *dens_dev) {
  WfDev *d = dev_of(wf);
  const int W = wf->is_ncl ? d->nw[k] / 2 : d->nw[k];
  const pp_cfloat *C = (const pp_cfloat *)d->C.p + ((size_t)k * d->nband + b0) * d->ldc + (size_t)comp * W;
  const DevBuf &pj = comp ? d->proj_dn : d->proj;
  const pp_cdouble *P = (const pp_cdouble *)pj.p + ((size_t)k * d->nband + b0) * std::max(1, d->ptot);
  return pp_realspace_batched(ctx, rp, wf->kpts[k]->k, W, (const int *)d->G.p + 3 * (size_t)d->gofs[k], C, d->ldc,
                              nb, P, std::max(1, d->ptot), remove, states_dev, ldo, w, dens_dev,
                              g_precision == PP_FP64 ? PP_FP64 : PP_FP32);
}

static int rs_check(pswf_t *wf, const int *fftg, const int *labels, const double *coords, const char *who) {
  PP_REQUIRE(wf && fftg && (wf->num_sites == 0 || (labels && coords)), PP_ERR_ARG,
             std::string(who) + ": bad arguments");
  PP_REQUIRE(fftg[0] > 0 && fftg[1] > 0 && fftg[2] > 0, PP_ERR_ARG, std::string(who) + ": bad grid");
  return PP_OK;
}

// one state (spinors: both components) into host x
static int rs_state_host(pp_cdouble *x, int band, int k, pswf_t *wf, int *fftg, int *labels, double *coords,
                         int ncomp, const char *who) {
  PP_TRY(rs_check(wf, fftg, labels, coords, who));
  PP_REQUIRE(x, PP_ERR_ARG, std::string(who) + ": null output");
  PP_REQUIRE(band >= 0 && band < wf->nband && k >= 0 && k < wf->nwk * wf->nspin, PP_ERR_ARG,
             std::string(who) + ": band or k-point index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  pp_rsplan *rp;
  PP_TRY(rs_plan_for(ctx, wf, fftg, labels, coords, &rp));
  const long npts = (long)fftg[0] * fftg[1] * fftg[2];
  DevBuf X;
  PP_TRY(X.ensure(sizeof(pp_cdouble) * (size_t)npts * ncomp));
  for (int c = 0; c < ncomp; c++)
    PP_TRY(rs_run(ctx, wf, rp, k, band, 1, c, 0, (pp_cdouble *)X.p + (size_t)c * npts, npts, nullptr, nullptr));
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  PP_CHECK_HIP(hipMemcpy(x, X.p, sizeof(pp_cdouble) * (size_t)npts * ncomp, hipMemcpyDeviceToHost));
  return PP_OK;
}

// P += weighted densities of every k and the bands with occ > 0
static int rs_density_host(double *P, pswf_t *wf, int *fftg, int *labels, double *coords, int only_band, int only_k,
                           const char *who) {
  PP_TRY(rs_check(wf, fftg, labels, coords, who));
  PP_REQUIRE(P, PP_ERR_ARG, std::string(who) + ": null output");
  std::lock_guard<std::mutex> lk(g_mu);
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  pp_rsplan *rp;
  PP_TRY(rs_plan_for(ctx, wf, fftg, labels, coords, &rp));
  const long npts = (long)fftg[0] * fftg[1] * fftg[2];
  DevBuf D;
  PP_TRY(D.ensure(sizeof(double) * (size_t)npts));
  PP_CHECK_HIP(hipMemcpy(D.p, P, sizeof(double) * npts, hipMemcpyHostToDevice));
  const int ncomp = wf->is_ncl ? 2 : 1;
  if (only_band >= 0) {
    const double one = 1.0;
    PP_TRY(rs_run(ctx, wf, rp, only_k, only_band, 1, 0, 0, nullptr, 0, &one, (double *)D.p));
  } else {
    // density.c:158-203: weight_k occ (2 / nspin; spinors 1) for occ > 0
    const double spin_mult = wf->is_ncl ? 1.0 : 2.0 / wf->nspin;
    std::vector<double> w(wf->nband);
    for (int k = 0; k < wf->nwk * wf->nspin; k++) {
      int nb = 0;
      for (int b = 0; b < wf->nband; b++) {
        const double occ = wf->kpts[k]->bands[b]->occ;
        w[b] = occ > 0 ? wf->kpts[k]->weight * occ * spin_mult : 0.0;
        if (occ > 0)
          nb = b + 1;
      }
      for (int c = 0; c < ncomp; c++)
        PP_TRY(rs_run(ctx, wf, rp, k, 0, nb, c, 0, nullptr, 0, w.data(), (double *)D.p));
    }
  }
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  PP_CHECK_HIP(hipMemcpy(P, D.p, sizeof(double) * npts, hipMemcpyDeviceToHost));
  return PP_OK;
}

}  // namespace pp

extern "C" {

int realspace_state(pp_cdouble *x, int BAND_NUM, int KPOINT_NUM, pswf_t *wf, int *fftg, int *labels,
                    double *coords) {
  PP_REQUIRE(wf && !wf->is_ncl, PP_ERR_ARG, "realspace_state: spinor wavefunction (use ncl_realspace_state)");
  return rs_state_host(x, BAND_NUM, KPOINT_NUM, wf, fftg, labels, coords, 1, "realspace_state");
}

int ncl_realspace_state(pp_cdouble *x, int BAND_NUM, int KPOINT_NUM, pswf_t *wf, int *fftg, int *labels,
                        double *coords) {
  PP_REQUIRE(wf && wf->is_ncl, PP_ERR_ARG, "ncl_realspace_state: not a spinor wavefunction");
  return rs_state_host(x, BAND_NUM, KPOINT_NUM, wf, fftg, labels, coords, 2, "ncl_realspace_state");
}

int remove_phase(pp_cdouble *x, int KPOINT_NUM, pswf_t *wf, int *fftg) {
  PP_REQUIRE(x && wf && fftg && KPOINT_NUM >= 0 && KPOINT_NUM < wf->nwk * wf->nspin, PP_ERR_ARG,
             "remove_phase: bad arguments");
  std::lock_guard<std::mutex> lk(g_mu);
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  const long npts = (long)fftg[0] * fftg[1] * fftg[2];
  DevBuf X;
  PP_TRY(X.ensure(sizeof(pp_cdouble) * (size_t)npts));
  PP_CHECK_HIP(hipMemcpy(X.p, x, sizeof(pp_cdouble) * npts, hipMemcpyHostToDevice));
  PP_TRY(rs_remove_phase(ctx, (cplx<double> *)X.p, fftg, wf->kpts[KPOINT_NUM]->k));
  PP_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  PP_CHECK_HIP(hipMemcpy(x, X.p, sizeof(pp_cdouble) * npts, hipMemcpyDeviceToHost));
  return PP_OK;
}

int ae_state_density(double *P, int BAND_NUM, int KPOINT_NUM, pswf_t *wf, int *fftg, int *labels, double *coords) {
  PP_REQUIRE(wf && !wf->is_ncl, PP_ERR_ARG, "ae_state_density: spinor wavefunction");
  PP_REQUIRE(BAND_NUM >= 0 && BAND_NUM < wf->nband && KPOINT_NUM >= 0 && KPOINT_NUM < wf->nwk * wf->nspin,
             PP_ERR_ARG, "ae_state_density: band or k-point index out of range");
  return rs_density_host(P, wf, fftg, labels, coords, BAND_NUM, KPOINT_NUM, "ae_state_density");
}

int ae_chg_density(double *P, pswf_t *wf, int *fftg, int *labels, double *coords) {
  PP_REQUIRE(wf && !wf->is_ncl, PP_ERR_ARG, "ae_chg_density: spinor wavefunction (use ncl_ae_chg_density)");
  return rs_density_host(P, wf, fftg, labels, coords, -1, -1, "ae_chg_density");
}

int ncl_ae_chg_density(double *P, pswf_t *wf, int *fftg, int *labels, double *coords) {
  PP_REQUIRE(wf && wf->is_ncl, PP_ERR_ARG, "ncl_ae_chg_density: not a spinor wavefunction");
  return rs_density_host(P, wf, fftg, labels, coords, -1, -1, "ncl_ae_chg_density");
}

int project_realspace_state(pp_cdouble *projs, int BAND_NUM, pswf_t *wf, pswf_t *wf_R, int *fftg, int *labels,
                            double *coords, int *labels_R, double *coords_R) {
  PP_REQUIRE(projs && wf && wf_R && !wf->is_ncl && !wf_R->is_ncl, PP_ERR_ARG,
             "project_realspace_state: bad arguments");
  PP_TRY(rs_check(wf, fftg, labels, coords, "project_realspace_state"));
  PP_TRY(rs_check(wf_R, fftg, labels_R, coords_R, "project_realspace_state"));
  PP_REQUIRE(BAND_NUM >= 0 && BAND_NUM < wf->nband, PP_ERR_ARG, "project_realspace_state: band index out of range");
  PP_REQUIRE(wf_R->nband >= wf->nband && wf_R->nwk * wf_R->nspin >= wf->nwk * wf->nspin, PP_ERR_ARG,
             "project_realspace_state: the basis has fewer bands or k-points than the projected wavefunction");
  if (wf == wf_R)  // one wavefunction holds one cached plan
    for (int i = 0; i < wf->num_sites; i++)
      PP_REQUIRE(labels[i] == labels_R[i] && coords[3 * i] == coords_R[3 * i] &&
                     coords[3 * i + 1] == coords_R[3 * i + 1] && coords[3 * i + 2] == coords_R[3 * i + 2],
                 PP_ERR_ARG, "project_realspace_state: one wavefunction with two different site lists");
  std::lock_guard<std::mutex> lk(g_mu);
  pp_ctx *ctx;
  PP_TRY(global_ctx(&ctx));
  pp_rsplan *rpS, *rpR;
  PP_TRY(rs_plan_for(ctx, wf, fftg, labels, coords, &rpS));
  PP_TRY(rs_plan_for(ctx, wf_R, fftg, labels_R, coords_R, &rpR));
  const long npts = (long)fftg[0] * fftg[1] * fftg[2];
  const int nkt = wf->nwk * wf->nspin, nband = wf->nband;
  const int cb = std::min(64, nband);
  DevBuf S, R, O;
  PP_TRY(S.ensure(sizeof(pp_cdouble) * (size_t)npts));
  PP_TRY(R.ensure(sizeof(pp_cdouble) * (size_t)npts * cb));
  PP_TRY(O.ensure(sizeof(pp_cdouble) * (size_t)nband));
  std::vector<pp_cdouble> o(nband);
  const double scale = det3(wf->lattice) / npts;
  for (int k = 0; k < nkt; k++) {
    PP_TRY(rs_run(ctx, wf, rpS, k, BAND_NUM, 1, 0, 0, S.p, npts, nullptr, nullptr));
    for (int b0 = 0; b0 < nband; b0 += cb) {
      const int nb = std::min(cb, nband - b0);
      PP_TRY(rs_run(ctx, wf_R, rpR, k, b0, nb, 0, 0, R.p, npts, nullptr, nullptr));
      PP_TRY(rs_dots(ctx, (const cplx<double> *)R.p, npts, nb, (const cplx<double> *)S.p, npts, scale,
                     (cplx<double> *)O.p + b0));
    }
    PP_CHECK_HIP(hipMemcpy(o.data(), O.p, sizeof(pp_cdouble) * nband, hipMemcpyDeviceToHost));
    for (int b = 0; b < nband; b++)
      projs[(long)b * nkt + k] = o[b];
  }
  return PP_OK;
}

}  // extern "C"
