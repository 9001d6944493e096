/*
 * pawpy_types.h -- data model of the pawpyseed_amd C-ABI.
 *
 * The structs keep the field order and types of the reference's
 * pawpyseed/core/utils.h:30-166 (funcset_t, ppot_t, projection_t, band_t,
 * rayleigh_set_t, kpoint_t, pswf_t), because the reference's Cython
 * boundary reads fields directly (pawpyc.pyx:537 reads
 * wf_ptr.kpts[k].bands[b].occ).  pswf_t gains one trailing member, `dev`,
 * the library's device-resident state (coefficients, projections and
 * cached overlap matrices in HBM); code written against the reference
 * never sees it.
 */
#ifndef PAWPY_TYPES_H
#define PAWPY_TYPES_H

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> pp_cdouble;
typedef std::complex<float> pp_cfloat;
extern "C" {
#else
#include <complex.h>
typedef double complex pp_cdouble;
typedef float complex pp_cfloat;
#endif

typedef struct funcset {
  int l;
  double *proj;
  double **proj_spline;
  double *aewave;
  double **aewave_spline;
  double *pswave;
  double **pswave_spline;
  double *diffwave;
  double **diffwave_spline;
  double *kwave;
  double **kwave_spline;
  double *smooth_diffwave;
  double **smooth_diffwave_spline;
  double *dense_kwave;
  double **dense_kwave_spline;
} funcset_t;

typedef struct ppot {
  int num_projs;
  int total_projs;
  int lmax;
  funcset_t *funcs;
  double rmax;
  double wave_rmax;
  double *pspw_overlap_matrix;
  double *aepw_overlap_matrix;
  double *diff_overlap_matrix;
  int proj_gridsize;
  int wave_gridsize;
  int num_cart_gridpts;
  double *wave_grid;
  double *kwave_grid;
  double *proj_grid;
  double *smooth_grid;
  double *dense_kgrid;
} ppot_t;

typedef struct projection {
  int num_projs;
  int total_projs;
  int *ns;
  int *ls;
  int *ms;
  pp_cdouble *overlaps;
} projection_t;

typedef struct band {
  int n;
  int num_waves;
  double occ;
  double N;
  double energy;
  pp_cfloat *Cs;
  pp_cdouble *CRs;
  pp_cfloat *CAs;
  projection_t *projections;
  projection_t *up_projections;
  projection_t *down_projections;
  projection_t *wave_projections;
} band_t;

typedef struct rayleigh_set {
  int l;
  pp_cdouble *terms;
} rayleigh_set_t;

typedef struct kpoint {
  short int up;
  int num_waves;
  int *Gs;
  double *k;
  double weight;
  int num_bands;
  band_t **bands;
  rayleigh_set_t **expansion;
} kpoint_t;

typedef struct pswf {
  double encut;
  int num_elems;
  int *num_projs;
  int num_sites;
  ppot_t *pps;
  int *G_bounds;
  kpoint_t **kpts;
  int nspin;
  int nband;
  int nwk;
  double *lattice;
  double *reclattice;
  int *fftg;
  int is_ncl;
  int wp_num;
  int num_aug_overlap_sites;
  double *dcoords;
  pp_cdouble **overlaps;
  /* pawpyseed_amd extension: device-resident state, owned by the library */
  void *dev;
} pswf_t;

#ifdef __cplusplus
}
#endif
#endif
