/*
 * pawpy_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, FP64, OpenMP) of pawpyseed's projection path.
 * It is the checker for the HIP library and, timed on the host cores, the
 * CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it; the product library never links or calls it.
 *
 * Every routine names the reference lines (pawpyseed 0.7.1) it follows.
 * Complex numbers are interleaved (re, im) double pairs unless the name
 * says c64 (interleaved floats).
 */
#ifndef PAWPY_ORACLE_H
#define PAWPY_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- radial splines: utils.c:699-749, 751-764, 461-475 ---- */
void or_spline_coeff(const double *x, const double *y, int n, double *b,
                     double *c, double *d);
double or_spline_integral(const double *x, const double *a, const double *b,
                          const double *c, const double *d, int n);
double or_proj_interpolate(double r, double rmax, int n, const double *x,
                           const double *f, const double *b, const double *c,
                           const double *d);

/* ---- angular part: utils.c:377-386, 441-449, 547-564 ---- */
double or_legendre(int l, int m, double x);
void or_ylm(int l, int m, double theta, double phi, double *out2);
void or_proj_value(int l, int m, double rmax, int n, const double *x,
                   const double *f, const double *b, const double *c,
                   const double *d, const double *vec, double r, double *out2);

/* ---- geometry: utils.c:51-73, 149-160 ---- */
void or_min_cart_path(const double *coord, const double *center,
                      const double *lattice, double *path, double *r);

/* Sphere table of one site, utils.c:590-696 (pr0_pw1 = 0 branch).
 * Channel data: nchan radial functions f[ch*n + j] on the linear grid
 * x[j] with spline coefficients b,c,d laid out the same way; ls[ch].
 * Output (capacity cap points): idx[p], path[3p], values[P*cap + p] with
 * the projector order of the reference (channel, then m=-l..l).
 * Returns the number of points, or -1 if cap is too small. */
int or_site_table(const double *lattice, const int *fftg, const double *coord,
                  double rmax, int n, const double *x, int nchan,
                  const int *ls, const double *f, const double *b,
                  const double *c, const double *d, int cap, int *idx,
                  double *path, double *values);

/* Bounding-box capacity used by or_site_table for this site radius. */
int or_site_capacity(const double *lattice, const int *fftg, double rmax);

/* ---- FFT: linalg.c:14-45 (backward, scale 1/sqrt(V)) ---- */
/* Switch the 3D backward transforms (or_fft3d, or_ifft3d_inplace,
 * or_setup_projections) to MKL's, loaded at run time from libmkl_rt
 * (PP_ORACLE_MKL names another path); returns 1 when MKL is in use, 0 when
 * it cannot be loaded (the oracle's own transform stays).  enable = 0
 * switches back. */
int or_fft_use_mkl(int enable);
void or_fft3d(double *x, const double *lattice, const int *Gs,
              const float *Cs_c64, int nw, const int *fftg);
/* plain backward 3D DFT with no scale and no scatter, for tests */
void or_ifft3d_inplace(double *x, const int *fftg);

/* ---- sphere sums: projector.c:223-274 for every site at once ----
 * CSR tables: site_pt[s]..site_pt[s+1] points of site s; idx/path per point;
 * values at val_off[s] + p*npts_s + i; projections out at proj_off[s] + p. */
void or_onto_projector(const double *grid, const double *lattice,
                       const double *kfrac, const int *fftg, int nsites,
                       const int *site_pt, const int *site_nproj,
                       const long *val_off, const int *proj_off,
                       const int *idx, const double *path,
                       const double *values, double *proj_out);

/* projector.c:560-602 hot loop: OpenMP over (band, k) pairs, each pair one
 * fft3d + one onto_projector.  C is c64 [k][b][ldc]; proj_out c128
 * [k][b][ptot]; pairs = list of (k, b) to do (npairs; NULL = all). */
void or_setup_projections(const double *lattice, const int *fftg, int nk,
                          const double *kfracs, const int *nw,
                          const int *const *Gs, const float *C, long ldc,
                          int nband, int nsites, const int *site_pt,
                          const int *site_nproj, const long *val_off,
                          const int *proj_off, int ptot, const int *idx,
                          const double *path, const double *values,
                          int npairs, const int *pairs, double *proj_out,
                          int nthreads);

/* ---- pseudoprojection: pseudoprojector.c:63-90 ----
 * out[b*nkt + k] = sum_G conj(C_ref[k'][b][G]) C_proj[k][band][G]
 * nkt = nwk*nspin, k' = spin-flipped k when flip && nspin==2.
 * fp32acc=1 accumulates in single precision (the reference's cdotc). */
void or_pseudoprojection(double *out, const float *C_ref, const float *C_proj,
                         long ldc, int nband_ref, int nband_proj, int nwk,
                         int nspin, const int *nw, int band, int flip,
                         int fp32acc);

/* ---- compensation_terms, O_M term: projector.c:850-910 ----
 * proj_R c128 [nkt][nbR][ptotR], proj_S c128 [nkt][nbS][ptotS];
 * per-site projector metadata from site offsets and per-species tables:
 * sp_nproj[e] projectors of species e with ns/ls/ms at sp_meta_off[e];
 * dO[e] = aepw - pspw channel matrix (nchan[e]^2) at sp_do_off[e]. */
void or_compensation_terms_m(double *overlap, int band, int nwk, int nspin,
                             int nbR, const double *proj_R, int ptotR,
                             const int *proj_off_R, const int *labels_R,
                             const double *proj_S, int ptotS,
                             const int *proj_off_S, int num_M, const int *M_R,
                             const int *M_S, const int *sp_nchan,
                             const int *sp_meta_off, const int *meta_n,
                             const int *meta_l, const int *meta_m,
                             const int *sp_do_off, const double *dO, int flip);

/* ---- PAW partial-wave overlaps: projector.c:507-558 ----
 * ps/ae waves [nchan][n] on grid; outputs nchan^2 matrices. */
void or_pwave_overlaps(const double *grid, int n, int nchan, const int *ls,
                       const double *aewaves, const double *pswaves,
                       double *psov, double *aeov, double *diov);

/* projector.c:912-938: the O_R and O_S terms added into overlap
 * (c128 [nbR * nkt]) from the wave projections of overlap_setup_real. */
void or_compensation_terms_rs(double *overlap, int nwk, int nspin, int nbR,
                              const double *proj_R, int ptotR,
                              const int *proj_off_R, int num_N_R,
                              const int *N_R, const double *wp_S,
                              int wp_ptot_S, const int *wp_off_S,
                              const double *wp_R, int wp_ptot_R,
                              const int *wp_off_R, const double *proj_S,
                              int ptotS, const int *proj_off_S, int num_N_S,
                              const int *N_S, int flip);

/* ---- spherical Bessel transform: sbt.c:24-213 (Talman's NumSBT) ----
 * forward: g(k) = int r^2 j_l(k r) R(r) dr from f = r R(r) on the n-point
 * log grid r; inverse: R(r) = (2/pi) int k^2 j_l(k r) g(k) dk. */
void or_sbt_forward(double encut, double enbuf, int lmax, int n, const double *r, const double *f, int l,
                    double *ks_out, double *g);
void or_sbt_inverse(double encut, double enbuf, int lmax, int n, const double *r, const double *g, int l,
                    double *rs_out, double *R);
/* utils.c:477-489 */
double or_wave_interpolate(double r, int n, const double *x, const double *f, const double *b, const double *c,
                           const double *d);
/* projector.c:114-163: SBT-filtered partial-wave difference on the
 * np-point linear grid j wave_rmax / np (sdw[np]); kgrid[n], kwave[n]. */
void or_smooth_diffwave(int n, const double *wave_grid, const double *diffwave, int l, int lmax,
                        double grid_encut, int np, double *sdw, double *kgrid, double *kwave);

/* ---- off-site overlaps: gaunt.py:17-27, radial.c:116-196, projector.c:940-959 ---- */
double or_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);
double or_sbtfac(int l1, int l2, int l3, int m1, int m2);
void or_offsite_overlap(const double *dcoord, const double *k1, const double *f1, const double *b1,
                        const double *c1, const double *d1, int n1, const double *k2, const double *f2,
                        const double *b2, const double *c2, const double *d2, int n2, int l1, int m1, int l2,
                        int m2, double *out2);
void or_compensation_terms_n(double *overlap, int nwk, int nspin, int nbR, const double *proj_R, int ptotR,
                             const int *proj_off_R, const double *proj_S, int ptotS, const int *proj_off_S,
                             int num_N_RS, const int *N_RS_R, const int *N_RS_S, const double *ov,
                             const long *ov_off, const double *dcoords, const double *kpts, int flip);

/* ---- aug_recip: projector.c:276-330, 420-453; linalg.c:47-79 ---- */
void or_aug_freqs(double *CA, const double *lattice, const double *kfrac, const int *fftg, const int *Gs, int nw,
                  int nsites, const int *site_pt, const int *site_nproj, const long *val_off, const int *idx,
                  const double *path, const double *values, const double *P, const int *proj_off_P);

/* ---- all-electron real-space states: density.c:232-327, utils.c:522-545 ---- */
void or_wave_value2(int n, const double *x, const double *f, const double *b,
                    const double *c, const double *d, int l, int m,
                    const double *pos, double *out2);
void or_realspace_state(double *x, const double *lattice, const int *fftg,
                        const double *kfrac, int nsites, const int *labels,
                        const double *coords, const int *wn, const long *woff,
                        const double *wg, const int *chan_off,
                        const int *chan_l, int wmax, const double *f,
                        const double *b, const double *c, const double *d,
                        const int *proj_off, const double *proj);
void or_remove_phase(double *x, const int *fftg, const double *kfrac);

#ifdef __cplusplus
}
#endif
#endif
