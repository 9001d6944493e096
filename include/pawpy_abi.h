/*
 * pawpy_abi.h -- C-ABI of libpawpy_amd.so, the MI355X (gfx950) projection
 * path of pawpyseed.
 *
 * Two layers:
 *  1. Reference-named entry points.  Same symbol names and argument
 *     meaning as the reference's C core, so the reference's Python layer
 *     (through its Cython externs, pawpyc_extern.pxd, or the ctypes
 *     binding in pawpyseed_amd.core) can call them unchanged.  Functions
 *     the reference declares `void` return an int status here (0 = OK);
 *     on x86-64 a caller that ignores the return value is unaffected.
 *     Errors never exit the process: the status is returned and
 *     pp_last_error() gives the message.
 *  2. Batched device-resident entry points (pp_*): plain device pointers
 *     and sizes, asynchronous on the context's stream.
 *
 * Complex data are interleaved (re, im) pairs; c64 = two floats, c128 =
 * two doubles.
 */
#ifndef PAWPY_ABI_H
#define PAWPY_ABI_H

#include <stddef.h>

#include "pawpy_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PP_OK 0
#define PP_ERR_ARG 1
#define PP_ERR_HIP 2
#define PP_ERR_NOMEM 3
#define PP_ERR_UNSUPPORTED 4
#define PP_ERR_IO 5

/* arithmetic of the hot path */
#define PP_FP64 0   /* FP64 everywhere (the plain path, the parity path) */
#define PP_FP32 1   /* FFT, sphere sums and overlap GEMM in FP32 (MFMA), FP64 reductions */
#define PP_BF16X3 2 /* as PP_FP32, overlap GEMM as 3 products of (hi, lo) bf16 planes on MFMA */
#define PP_BF16X6 3 /* as PP_FP32, overlap GEMM as 6 products of (hi, mid, lo) bf16 planes on MFMA */
#define PP_FP16X3 4 /* as PP_FP32, overlap GEMM as 3 products of (hi, lo) fp16 planes, each 1024-term chunk of a row
                       scaled by a power of two (22 significant bits per operand) */

/* Message of the last failing call on this thread ("" if none). */
const char *pp_last_error(void);
int pp_last_status(void);
/* Library version string. */
const char *pp_version(void);
/* Precision used by the reference-named entry points (default PP_FP64). */
int pp_set_precision(int precision);
int pp_get_precision(void);

/* ===================================================================== */
/* 1. Reference-named entry points                                       */
/* ===================================================================== */

/* replaces reader.c:317-323 read_wavefunctions (WAVECAR reader; G vectors
 * enumerated as reader.c:231-271). Returns NULL on error. */
pswf_t *read_wavefunctions(char *filename, double *kpt_weights);
/* replaces reader.c:325-330 read_wavefunctions_from_str (in-memory WAVECAR) */
pswf_t *read_wavefunctions_from_str(char *start, double *kpt_weights);

/* No reference counterpart: builds a pswf_t from in-memory arrays (the
 * data a WAVECAR holds).  kpts [nwk][3]; weights [nwk]; nplane [nspin*nwk];
 * Gs: concatenated int32 [sum nplane][3]; coeffs: concatenated c64
 * [sum over (spin,k) of nband*nplane] (band-major per k);
 * occ, energy: [nspin*nwk][nband].  k-index = k + spin*nwk.  A k-point
 * whose G list is one list of W vectors given twice (nplane = 2 W, as the
 * reader keeps a spinor's list, reader.c:274-281) marks the wavefunction
 * noncollinear, its coefficients being the up then the down component. */
pswf_t *pp_pswf_from_arrays(int nspin, int nwk, int nband, double encut,
                            const double *lattice, const double *kpts,
                            const double *weights, const int *nplane,
                            const int *Gs, const pp_cfloat *coeffs,
                            const double *occ, const double *energy);

/* replaces utils.c:275-301 free_pswf */
void free_pswf(pswf_t *wf);
/* replaces utils.c:322-327 free_ppot_list */
void free_ppot_list(ppot_t *pps, int length);
/* replace the accessors of utils.c:343-373 */
double *get_occs(pswf_t *wf);
int get_nband(pswf_t *wf);
int get_nwk(pswf_t *wf);
int get_nspin(pswf_t *wf);
int is_ncl(pswf_t *wf);
double get_encut(pswf_t *wf);
double get_energy(pswf_t *wf, int band, int kpt, int spin);
double get_occ(pswf_t *wf, int band, int kpt, int spin);
void set_num_sites(pswf_t *wf, int nsites);
void free_ptr(void *ptr);

/* replaces projector.c:20-171 get_projector_list: projector splines
 * (utils.c:699-749), PAW partial-wave overlap matrices
 * (projector.c:507-558), and the spherical Bessel transforms and
 * plane-wave-filtered partial-wave differences (kwave, smooth_diffwave,
 * projector.c:114-163) that the O_R/O_S and O_N terms use. */
ppot_t *get_projector_list(int num_els, int *labels, int *ls,
                           double *wave_grids, double *projectors,
                           double *aewaves, double *pswaves, double *rmaxs,
                           double grid_encut);

/* replaces projector.c:560-602 setup_projections: sphere tables
 * (utils.c:590-696), then scatter -> 3D FFT -> sphere sums for every
 * (band, k) on the GPU.  Takes ownership of pps as the reference does.
 * Fills band->projections of every band (host copies) and keeps the
 * projections resident in HBM for compensation_terms.  Spinor
 * wavefunctions (is_ncl) get band->up_projections and down_projections,
 * each component projected on its own (onto_projector_ncl,
 * projector.c:359-392). */
int setup_projections(pswf_t *wf, ppot_t *pps, int num_elems, int num_sites,
                      int *fftg, int *labels, double *coords);

/* overlap_setup_real / _recip and compensation_terms / _recip return
 * PP_ERR_UNSUPPORTED for spinor wavefunctions, which the reference's
 * Projector refuses (projector.py:74-75). */

/* replaces pseudoprojector.c:63-90 pseudoprojection:
 * projections[b*NK + k] = <C_ref[k'][b] | C_proj[k][BAND_NUM]>.
 * The first call for a (wf_ref, wf_proj, flip_spin) triple computes the
 * whole band-overlap matrix of every k as one GEMM in HBM; later calls
 * read their column from it. */
int pseudoprojection(pp_cdouble *projections, pswf_t *wf_ref, pswf_t *wf_proj,
                     int BAND_NUM, int flip_spin);

/* replaces projector.c:604-720 overlap_setup_real: S's bands projected
 * onto R's plane-wave-filtered partial-wave differences at the N_R sites
 * (on S's grid), and R's onto S's at the N_S sites, on the GPU (same
 * scatter/FFT/sphere-sum path as setup_projections, with the
 * smooth_diffwave tables of get_projector_list, radius wave_rmax); fills
 * band->wave_projections (wf_S: num_N_R sites, wf_R: num_N_S sites) and
 * keeps them in HBM.  For the N_RS pairs, the off-site partial-wave
 * overlaps (radial.c:116-196) into wf_S->overlaps / wf_S->dcoords as the
 * reference stores them.  Call setup_projections on both wavefunctions
 * first. */
int overlap_setup_real(pswf_t *wf_R, pswf_t *wf_S, int *labels_R, int *labels_S,
                       double *coords_R, double *coords_S, int *N_R, int *N_S,
                       int *N_RS_R, int *N_RS_S, int num_N_R, int num_N_S,
                       int num_N_RS);

/* replaces projector.c:727-848 overlap_setup_recip (the aug_recip method):
 * the plane-wave coefficients of each structure's augmentation functions
 * at its own unmatched sites (R's bands at N_R, S's at N_S; the CAs of
 * projector.c:420-453), computed on the GPU from the resident projections
 * (expansion on the grid, forward FFT, G sampling) and kept in HBM (FP64;
 * band->CAs gets float complex copies); plus the N_RS off-site overlaps as
 * in overlap_setup_real.  Call setup_projections on both first. */
int overlap_setup_recip(pswf_t *wf_R, pswf_t *wf_S, int *labels_R, int *labels_S,
                        double *coords_R, double *coords_S, int *N_R, int *N_S,
                        int *N_RS_R, int *N_RS_S, int num_N_R, int num_N_S,
                        int num_N_RS);

/* replaces projector.c:965-1077 compensation_terms_recip (adds into
 * overlap): <CA_R|C_S> + <C_R|CA_S> over the R k-point's plane waves, plus
 * O_M and O_N, for all band pairs at once (FP64 GEMMs), cached like
 * compensation_terms. */
int compensation_terms_recip(pp_cdouble *overlap, int BAND_NUM, pswf_t *wf_S,
                             pswf_t *wf_R, int num_M, int num_N_R, int num_N_S,
                             int num_N_RS, int *M_R, int *M_S, int *N_R, int *N_S,
                             int *N_RS_R, int *N_RS_S, int *proj_labels,
                             double *proj_coords, int *ref_labels,
                             double *ref_coords, int *fft_grid, int flip_spin);

/* replaces projector.c:850-963 compensation_terms (adds into overlap).
 * O_M, O_R, O_S and O_N are computed for all band pairs at once (FP64) and
 * cached per (wf_R, flip_spin, site lists); O_R/O_S/O_N need a preceding
 * overlap_setup_real with the same N_R/N_S/N_RS lists. */
int compensation_terms(pp_cdouble *overlap, int BAND_NUM, pswf_t *wf_S,
                       pswf_t *wf_R, int num_M, int num_N_R, int num_N_S,
                       int num_N_RS, int *M_R, int *M_S, int *N_R, int *N_S,
                       int *N_RS_R, int *N_RS_S, int *proj_labels,
                       double *proj_coords, int *ref_labels, double *ref_coords,
                       int *fft_grid, int flip_spin);

/* replaces linalg.c:14-45 fft3d for one band (host buffers). */
int fft3d(pp_cdouble *x, int *G_bounds, double *lattice, double *kpt, int *Gs,
          pp_cfloat *Cs, int num_waves, int *fftg);

/* All-electron real-space states (density.h, density.c:158-327), on the
 * GPU over the device-resident coefficients and projections of a preceding
 * setup_projections.  Each returns a status instead of void.
 * realspace_state (density.c:232-310): x c128 [fftg] (x slowest) gets the
 * state of (BAND_NUM, KPOINT_NUM) on the grid fftg with the partial-wave
 * differences of the sites (labels, coords). */
int realspace_state(pp_cdouble *x, int BAND_NUM, int KPOINT_NUM, pswf_t *wf,
                    int *fftg, int *labels, double *coords);
/* density.c:312-326: x *= e^{-2 pi i k.r} (on the GPU). */
int remove_phase(pp_cdouble *x, int KPOINT_NUM, pswf_t *wf, int *fftg);
/* density.c:28-39: P += |psi|^2 of one state. */
int ae_state_density(double *P, int BAND_NUM, int KPOINT_NUM, pswf_t *wf,
                     int *fftg, int *labels, double *coords);
/* density.c:158-179: P += sum over k and bands with occ > 0 of
 * |psi|^2 weight_k occ (2 / nspin). */
int ae_chg_density(double *P, pswf_t *wf, int *fftg, int *labels,
                   double *coords);
/* density.c:181-203 and 328-421 (spinors): x c128 [2][fftg], the up
 * component then the down one; the density sums both (factor 1). */
int ncl_realspace_state(pp_cdouble *x, int BAND_NUM, int KPOINT_NUM,
                        pswf_t *wf, int *fftg, int *labels, double *coords);
int ncl_ae_chg_density(double *P, pswf_t *wf, int *fftg, int *labels,
                       double *coords);
/* density.c:205-230: projs[b nwk nspin + k] = V/N^3 sum_r
 * conj(psi_R,b,k(r)) psi_BAND_NUM,k(r) over the bands b < wf->nband of
 * wf_R (the `realspace` projection method). */
int project_realspace_state(pp_cdouble *projs, int BAND_NUM, pswf_t *wf,
                            pswf_t *wf_R, int *fftg, int *labels,
                            double *coords, int *labels_R, double *coords_R);

/* No reference counterpart: copies projections of one (band, k-index)
 * into out (c128 [total projectors]) from the device-resident table; for
 * a spinor wavefunction out holds the up component, then the down one
 * (c128 [2][total projectors]). */
int pp_get_projections(pswf_t *wf, int band, int kidx, pp_cdouble *out);

/* ===================================================================== */
/* 2. Batched device-resident entry points                               */
/* ===================================================================== */

typedef struct pp_ctx pp_ctx;
typedef struct pp_sphere pp_sphere;
typedef struct pp_graph pp_graph;

/* Context: one HIP stream (default priority) on `device`, workspaces. */
int pp_ctx_create(pp_ctx **out, int device);
void pp_ctx_destroy(pp_ctx *ctx);
/* Use an existing stream (e.g. the caller's); NULL = the context's own. */
int pp_ctx_set_stream(pp_ctx *ctx, void *hip_stream);
int pp_ctx_synchronize(pp_ctx *ctx);
/* Per-kernel-class timing with HIP events on the context's stream.
 * pp_ctx_timing_read fills up to n entries of names/ms_total/launches and
 * returns the number of classes. */
int pp_ctx_set_timing(pp_ctx *ctx, int enable);
int pp_ctx_timing_read(pp_ctx *ctx, int n, const char **names,
                       double *ms_total, long *launches);
int pp_ctx_timing_reset(pp_ctx *ctx);
/* Tuning knobs (0 = automatic): bands per FFT chunk (<= 64), bands per
 * FFT work-group tile (power of two <= 32), FFT pruning to the plane-wave
 * sphere (1 on, 2 off), sphere-sum parts per site, GEMM split-K factor. */
int pp_ctx_set_tuning(pp_ctx *ctx, int fft_bands, int fft_tile, int fft_prune,
                      int sphere_parts, int gemm_splitk);
/* One tuning knob by name: "fft_bands", "fft_tile", "fft_prune",
 * "fft_pingpong" (1 = two-buffer LDS FFT stages instead of in-place),
 * "fft_persistent" (0 off, 1 on with 2048 work-groups, > 1 work-group
 * count: FFT passes as a fixed grid that prefetches the next line),
 * "sphere_parts", "gemm_splitk", "gemm_f64_valu" (1 = FP64 GEMM on the
 * VALU instead of the FP64 matrix cores), "streams" (1 or 2 band chunks in
 * flight, default 2), "sphere_unroll" (4 or 8 sphere-sum gathers in flight
 * per wavefront, default 8), "fft_threads" (256 or 512: work-group size of
 * the FFT passes when fft_tile is 32), "fft_radix" (0: radices 2-5,
 * default; 1: composite radices 6/8/9/12/16 where they fit, fewer LDS
 * round trips but more registers), "fused" (0, default: separate x-pass and
 * sphere-sum kernels; 1: the FP32 path runs them as one kernel where the
 * plan allows -- it writes no real-space grid but measured slower on the
 * 216-atom workload), "fft_xcd" (1: FFT pass work-groups walk
 * XCD-contiguous line ranges), "fft_fixed" (1, default: FFT passes along
 * axes of length 60, 108, 192 or 216 run kernels compiled for that length,
 * with every index and offset a constant; 0: the generic in-place passes;
 * the two agree to rounding), "fft_nt" (FFT passes: 0 cached, 1
 * non-temporal stores, 2 non-temporal loads and stores, default 2; 3:
 * cached z pass, streamed stores in the y pass, streamed x pass),
 * "sphere_nt" (1, default: sphere-sum gathers with non-temporal loads),
 * "fft_xmask" (1: the x pass of a setup writes only the grid points some
 * projector sphere holds, the only ones the sphere sums read; 0, default:
 * every point -- the masked stores measured slower on the 216-atom
 * workload although they write 31% fewer bytes), "rs_kernel" (real-space
 * augmentation: 0 a site's coefficients in LDS, 1 in registers (up to 32
 * projectors per site), 2, default: tiles of points with their table rows
 * in LDS and the coefficients in registers (up to 18 projectors per site;
 * else 0)), "rs_unroll" (1, 2 or 4,
 * default 4: sphere points of a wavefront in flight in the real-space
 * augmentation), "rs_prune" (1, default: the real-space FFT passes skip
 * lines outside the plane-wave sphere), "multi_items" (1..64, default 8:
 * work items per launch of pp_setup_projections_multi), "gemm_planes"
 * (two-plane split GEMMs, PP_FP16X3 and PP_BF16X3: 1 one work-group per
 * CU with 64-real k-steps over planes split beforehand, 2 two work-groups
 * per CU with 32-real k-steps, 3, default: as 2 with the FP32 operands
 * split inside the GEMM, where their rows are 16-byte aligned: no plane
 * buffer; 4 as 3 with 256 x 128 output tiles of 8 wavefronts; all give
 * bit-identical results). */
int pp_ctx_set_param(pp_ctx *ctx, const char *name, int value);
/* Current value of a parameter of pp_ctx_set_param (0 = the built-in
 * default for fft_tile, fft_bands, sphere_parts, gemm_splitk). */
int pp_ctx_get_param(pp_ctx *ctx, const char *name, int *value);
/* Autotuner of the setup path (no counterpart in the reference; the
 * pp_autotune of the boundary's batched entry points): coordinate descent
 * over streams, fft_tile, fft_threads, sphere_unroll, sphere_parts,
 * fft_radix, fft_xcd, fft_xmask, fft_nt, sphere_nt, fft_bands and (FP32)
 * fused, each
 * candidate timed
 * with HIP events on the caller's own inputs, best of `repeats` runs after
 * a warm-up.  The fastest values stay set on ctx, and proj_dev holds the
 * projections of a final run with them.  Arguments as
 * pp_setup_projections_batched; report (may be NULL) receives "name value
 * ms" lines, truncated to report_len. */
/* The setup path captured in a HIP graph (no counterpart in the
 * reference; the pp_graph_capture/replay of the boundary's batched entry
 * points).  Capture runs the setup once eagerly (sizing every buffer;
 * proj_dev then holds valid projections), prepares the host-side lists,
 * and records the chunk pipeline (scatter, FFT passes, sphere sums on the
 * context's two streams) as one graph.  Replay launches it on the
 * context's stream, asynchronously: it recomputes proj_dev from the
 * current contents of C_dev for the same G vectors, k-points and plan.
 * Replay refuses (PP_ERR_ARG) once the plan was rebuilt or the context's
 * work buffers were reallocated by a larger call after the capture. */
int pp_graph_capture(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac,
                     const int *gofs, const int *G_dev, const void *C_dev,
                     long ldc, int nband, void *proj_dev, int precision,
                     pp_graph **out);
int pp_graph_replay(pp_graph *graph);
void pp_graph_destroy(pp_graph *graph);
int pp_autotune(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac,
                const int *gofs, const int *G_dev, const void *C_dev, long ldc,
                int nband, void *proj_dev, int precision, int repeats,
                char *report, int report_len);

/* Sphere plan of one structure (replaces the table build of
 * projector.c:193-209 / utils.c:590-696, done on the GPU).
 * Species e: nchan[e] channels with l = ls[...] (concatenated), radius
 * rmax[e], ngrid[e] points of the linear grid x_j = j*rmax/ngrid with
 * projector values proj[...] (concatenated [nchan][ngrid] per species).
 * precision selects the table type (PP_FP64 or PP_FP32/PP_BF16X3). */
int pp_sphere_create(pp_ctx *ctx, pp_sphere **out, const double *lattice,
                     const int *fftg, int num_sites, const int *labels,
                     const double *frac_coords, int num_elems, const int *nchan,
                     const int *ls, const double *rmax, const int *ngrid,
                     const double *proj, int precision);
void pp_sphere_destroy(pp_sphere *sp);
/* total projectors, total sphere points, and per-site projector offsets
 * (proj_off[num_sites+1], may be NULL) */
int pp_sphere_info(pp_sphere *sp, int *total_projs, long *total_points,
                   int *proj_off);
/* Grid points held by at least one of the plan's projector spheres (the
 * union; the x pass of a setup writes only those).  Synchronises the
 * context's stream.  No counterpart in the reference. */
int pp_sphere_coverage(pp_sphere *sp, long *covered);
/* Rebuild the tables of an existing plan for new coordinates (same sites
 * and species): the per-structure setup step of setup_projections. */
int pp_sphere_rebuild(pp_sphere *sp, const double *frac_coords);
/* Diagnostic, no counterpart in the reference: whether the plan carries
 * (building them now if needed; they are otherwise built on the fused
 * path's first use after each rebuild) the work lists of the fused x pass
 * with sphere sums (FP32 plans with at
 * most 18 projectors per site whose boxes do not wrap onto themselves in
 * y or z), their work items, the partial rows that hold a site (one per
 * item and site), the plan's segments (runs of a site's points
 * on one x-line), and the work-list consistency flag (0 = every segment
 * has its slot; reading it synchronises the context's stream).  Any
 * pointer may be NULL. */
int pp_sphere_fused_info(pp_sphere *sp, int *usable, int *nitems, int *nrows,
                         long *nseg, int *flag);
/* Diagnostic, no counterpart in the reference, host only (no device
 * needed): the fused x-pass work lists of sites frac[ns][3] with sphere
 * grids rmax[s] / ngrid[s].  counts[0..2] = work items, line blocks, site
 * rows; then, where the pointers are not NULL, items[4 * items] (y0, z0,
 * sites, 0), isites[12 * items] (site per slot or -1), bitem[blocks + 1],
 * rows_off[ns + 1], rows[counts[2]] (item * 12 + slot).  Returns
 * PP_ERR_UNSUPPORTED when a site box wraps onto itself in y or z. */
int pp_fused_lists_host(const double *lattice, const int *fftg, int ns,
                        const double *frac, const double *rmax,
                        const int *ngrid, int *counts, int *items, int *isites,
                        int *bitem, int *rows_off, int *rows);

/* <p_i|psi_nk> for all bands and k-points of one structure:
 * kfrac [nk][3] (host); gofs [nk+1] (host) offsets into G_dev (device
 * int32 [sum W][3]); C_dev device c64 [nk][nband][ldc]; proj_dev device c128
 * [nk][nband][total_projs]. */
int pp_setup_projections_batched(pp_ctx *ctx, pp_sphere *sp, int nk,
                                 const double *kfrac, const int *gofs,
                                 const int *G_dev, const void *C_dev, long ldc,
                                 int nband, void *proj_dev, int precision);

/* Several structures sharing one plane-wave basis, set up as one work list
 * (the setup_projections of every Wavefunction that the reference's
 * setup_multiple_projections generator yields, projector.py:311-382, each
 * projector.c:560-602): nstruct plans (same FFT grid and precision); the
 * basis of nk k-points as in pp_setup_projections_batched (kfrac, gofs,
 * G_dev); per (structure s, k-point k), index i = s * nk + k, the band
 * count nband[i], coefficients C_dev[i] (device c64 [nband[i]][ldc[i]])
 * and projections proj_dev[i] (device c128 [nband[i]][total projectors of
 * plan s]).  Every (structure, k, 64-band chunk) is a work item; a batch of
 * "multi_items" items (tuning parameter, default 8) runs as one launch per
 * kernel, batches alternating over two streams.  One host synchronisation
 * (the plane-wave bounds of every k); asynchronous afterwards.  Inputs the
 * work-item kernels do not cover (a grid that aliases a G vector, a G set
 * with holes) go through the single-structure path instead. */
int pp_setup_projections_multi(pp_ctx *ctx, int nstruct, pp_sphere *const *plans, int nk,
                               const double *kfrac, const int *gofs, const int *G_dev,
                               const int *nband, const void *const *C_dev, const long *ldc,
                               void *const *proj_dev, int precision);
/* The same as one HIP graph (owning its prepared data and chunk buffers);
 * replay with pp_graph_replay, which refuses once a plan was rebuilt. */
int pp_graph_capture_multi(pp_ctx *ctx, int nstruct, pp_sphere *const *plans, int nk,
                           const double *kfrac, const int *gofs, const int *G_dev,
                           const int *nband, const void *const *C_dev, const long *ldc,
                           void *const *proj_dev, int precision, pp_graph **out);
/* Autotuner of pp_setup_projections_multi (as pp_autotune) over
 * multi_items, streams, fft_tile, sphere_unroll, sphere_parts, fft_nt,
 * sphere_nt and fft_xmask. */
int pp_autotune_multi(pp_ctx *ctx, int nstruct, pp_sphere *const *plans, int nk,
                      const double *kfrac, const int *gofs, const int *G_dev,
                      const int *nband, const void *const *C_dev, const long *ldc,
                      void *const *proj_dev, int precision, int repeats, char *report,
                      int report_len);

/* Band overlaps of every k as one complex GEMM each:
 * out[k][r][s] = sum_G conj(CR[kR(k)][r][G]) CS[k][s][G], G < nw[k],
 * kR(k) = k, or the other spin's k when flip && nspin == 2
 * (k-index = k + spin*nwk).  CR, CS device c64 with row stride ld and
 * k stride ld*nb; out device c128 [nkt][nbR][nbS]. */
int pp_overlap_gemm(pp_ctx *ctx, int nwk, int nspin, int flip, const int *nw,
                    const void *CR_dev, long ldR, int nbR, const void *CS_dev,
                    long ldS, int nbS, void *out_dev, int precision);

/* No counterpart in the reference: pp_overlap_gemm without any device
 * allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result; PP_FP64 or PP_FP32 only.  The split-K partial sums
 * live in a fixed 64 MiB device array of the library (one per GPU, held
 * until the process ends), so nbR * nbS may not exceed
 * pp_noalloc_scratch_bytes() / 32 (2,097,152) output entries per k-point
 * (PP_ERR_ARG otherwise).  The result equals pp_overlap_gemm's whenever
 * its split-K partial sums fit the array (FP64 at 512 x 512 bands), else
 * it differs by the rounding of another split (FP32: fewer, longer chunks
 * with in-chunk FP64 folds).  Runs on the context's stream only and touches
 * none of the context's buffers.  Calls that overlap in time on one GPU
 * share the array and race: serialise them. */
int pp_overlap_gemm_noalloc(pp_ctx *ctx, int nwk, int nspin, int flip, const int *nw,
                            const void *CR_dev, long ldR, int nbR, const void *CS_dev,
                            long ldS, int nbS, void *out_dev, int precision);
/* Size in bytes of the scratch array of pp_overlap_gemm_noalloc. */
size_t pp_noalloc_scratch_bytes(void);

/* O_M augmentation of every band pair (projector.c:890-910), FP64:
 * out[k][r][s] += sum_{t in M} sum_{i,j: (l,m) equal}
 *     conj(PR[kR][r][offR(M_R[t])+i]) dO_e[n_i][n_j] PS[k][s][offS(M_S[t])+j].
 * projR_dev c128 [nkt][nbR][ptotR], projS_dev c128 [nkt][nbS][ptotS];
 * proj_off_R/S (host) per-site offsets; labels_R/S (host) site species
 * (dO and n_i from the R site's species, n_j from the S site's);
 * species tables as in pp_sphere_create plus dO (concatenated nchan^2,
 * host).  out_dev c128 [nkt][nbR][nbS]. */
int pp_augmentation(pp_ctx *ctx, int nwk, int nspin, int flip,
                    const void *projR_dev, int nbR, int ptotR,
                    const int *proj_off_R, const int *labels_R,
                    const void *projS_dev, int nbS, int ptotS,
                    const int *proj_off_S, const int *labels_S, int num_M,
                    const int *M_R, const int *M_S, int num_elems,
                    const int *nchan, const int *ls, const double *dO,
                    void *out_dev);

/* The off-site O_N term of projector.c:940-959 for every band pair:
 * out[k][r][s] += sum over pairs t of
 *   sum_ij conj(P_R[kR][r][off_R(N_RS_R[t]) + i]) ov_t[i][j]
 *          P_S[k][s][off_S(N_RS_S[t]) + j] exp(2 pi i kpts[kR] . dcoords[t]),
 * kR the spin-flipped k when flip (the reference multiplies the fractional
 * k by the cartesian pair vector).  ov: host c128 blocks P_R x P_S at
 * ov_off[t] (num_pairs + 1 offsets), as overlap_setup_real computes them
 * (conj of the partial-wave overlaps, radial.c:116-196); dcoords [3 t];
 * kpts [nkt][3].  FP64 on the GPU. */
int pp_augmentation_offsite(pp_ctx *ctx, int nwk, int nspin, int flip,
                            const void *projR_dev, int nbR, int ptotR,
                            const int *proj_off_R, const void *projS_dev,
                            int nbS, int ptotS, const int *proj_off_S,
                            int num_pairs, const int *N_RS_R,
                            const int *N_RS_S, const pp_cdouble *ov,
                            const long *ov_off, const double *dcoords,
                            const double *kpts, void *out_dev);

/* Plane-wave coefficients of nband (<= 64) bands to their real-space
 * grids with the reference's 1/sqrt(V) scale (FP64 or FP32 FFT), the
 * device form of fft3d: G_dev int32 [nw][3], C_dev c64 [nband][ldc],
 * grid_out_dev c128 [nband][N0][N1][N2]. */
int pp_fft3d_batched(pp_ctx *ctx, const double *lattice, const int *fftg,
                     int nw, const int *G_dev, const void *C_dev, long ldc,
                     int nband, void *grid_out_dev, int precision);

/* Plane-wave coefficients of the augmentation functions of a plan's
 * sites for every band and k (the CAs of projector.c:420-453): per band,
 * x(r) = sum over the plan's sites s and projectors j of
 * P[k][b][src_off[s] + j] p_j(r) e^{-i k.path}, forward-transformed with
 * scale sqrt(V)/N^3 (linalg.c:47-79) and sampled at the G vectors of k.
 * proj_dev c128 [nk][nband][ptot_proj]; out_dev c128 [nk][nband][ldo]
 * (the reference keeps them as float complex).  The plan's table
 * precision selects the FFT precision. */
int pp_aug_freqs_batched(pp_ctx *ctx, pp_sphere *sp, int nk, const double *kfrac,
                         const int *gofs, const int *G_dev, const void *proj_dev,
                         long ptot_proj, const int *src_off, int nband,
                         void *out_dev, long ldo, int precision);

/* ---- per-species radial setup (host) ---- */

/* Spherical Bessel transform g(k) = int r^2 j_l(k r) R(r) dr of
 * f = r R(r) given on the n-point log grid r, on the k grid
 * k_q = sqrt((encut + enbuf) c) e^{(q - n + 1) drho} (q < n).  Replaces
 * spherical_bessel_transform_setup + wave_spherical_bessel_transform
 * (sbt.c:24-153, bound as pawpyc.spherical_bessel_transform,
 * pawpyc.pyx:130-150).  Writes ks[n] and g[n]. */
int pp_sbt_forward(double encut, double enbuf, int lmax, int n, const double *r,
                   const double *f, int l, double *ks, double *g);

/* The plane-wave-filtered partial-wave difference of one channel
 * (projector.c:114-163): diffwave = r (phi - phi~) on the n-point log
 * wave grid, transformed (encut 1e7 eV), cut at k < sqrt(c grid_encut),
 * transformed back and sampled on the linear grid j wave_rmax / np
 * (sdw[np]); kgrid[n], kwave[n] get the unfiltered transform. */
int pp_smooth_diffwave(int n, const double *wave_grid, const double *diffwave,
                       int l, int lmax, double grid_encut, int np, double *sdw,
                       double *kgrid, double *kwave);

/* Overlap of the functions f1/r1 Y_{l1 m1} at the origin and f2/r2 Y_{l2 m2}
 * at dcoord (cartesian), as int phi_1 conj(phi_2) d^3r, from their
 * spherical Bessel transforms (encut 1e5 eV) on the log grids r1[n1],
 * r2[n2]; l <= 3.  Replaces pawpyc.reciprocal_offsite_wave_overlap
 * (pawpyc.pyx:151-190) over radial.c:116-196.  out2 = (re, im). */
int pp_offsite_overlap(const double *dcoord, int n1, const double *r1,
                       const double *f1, int n2, const double *r2,
                       const double *f2, int l1, int m1, int l2, int m2,
                       double *out2);

/* ---- all-electron real-space states ---- */

typedef struct pp_rsplan pp_rsplan;
/* Partial-wave spheres of one structure on a real-space grid (the site
 * loops of realspace_state, density.c:259-309, done on the GPU).  Species
 * e: nchan[e] channels with l = ls[...] (concatenated, l <= 3), wave_n[e]
 * points of its log wave grid (wave_grid, concatenated) and the
 * partial-wave differences diffwave = r (phi - phi~) on it (projector.c:
 * 84-99; [nchan][wave_n] per species, concatenated).  A site's sphere
 * holds the box points (density.c:266-284) within the last point of its
 * wave grid; the table of (diffwave/r) Y_lm per point and projector is
 * built once.  Projectors of a site: channels in order, m = -l..l. */
int pp_rs_plan_create(pp_ctx *ctx, pp_rsplan **out, const double *lattice,
                      const int *fftg, int num_sites, const int *labels,
                      const double *frac_coords, int num_elems,
                      const int *nchan, const int *ls, const int *wave_n,
                      const double *wave_grid, const double *diffwave);
void pp_rs_plan_destroy(pp_rsplan *plan);
/* total projectors, sphere points, launch colours (sites whose spheres
 * may share a grid point never share one), whether the plan adds with
 * atomics (a site box wraps onto itself), per-site projector offsets
 * (proj_off[num_sites + 1]); pointers may be NULL. */
int pp_rs_plan_info(pp_rsplan *plan, int *total_projs, long *total_points,
                    int *ncolors, int *atomic, int *proj_off);
/* realspace_state / ae_state_density / ae_chg_density for nband bands of
 * one k-point (density.c:28-39, 158-179, 232-310): scatter + FFT of
 * C_dev (c64 [nband][ldc], G_dev int32 [nw][3]) on the plan's grid, the
 * Bloch phase (omitted with remove_phase, density.c:312-326) and the
 * partial-wave differences weighted by proj_dev (c128 [nband][ldp], the
 * plan's projector order).  Outputs (either may be NULL): states_dev c128
 * [nband][ldo] (x slowest) and density_dev f64 [N^3] += sum_b weights[b]
 * |psi_b|^2 (weights: host [nband]).  precision: PP_FP64 or PP_FP32 FFT.
 * Asynchronous on the context's stream after one host synchronisation
 * (the plane-wave bounds); with the "streams" parameter at 2 two 64-band
 * chunks are in flight on two streams, joined into the context's stream
 * before the call returns. */
int pp_realspace_batched(pp_ctx *ctx, pp_rsplan *plan, const double *kfrac,
                         int nw, const int *G_dev, const void *C_dev, long ldc,
                         int nband, const void *proj_dev, long ldp,
                         int remove_phase, void *states_dev, long ldo,
                         const double *weights, double *density_dev,
                         int precision);

#ifdef __cplusplus
}
#endif
#endif
