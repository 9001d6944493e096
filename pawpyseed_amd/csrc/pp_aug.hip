// PAW augmentation terms for every band pair, FP64: the on-site O_M
// (projector.c:890-910) and the off-site O_N (projector.c:940-959).
//
// O_M[k][r][s] = sum_{t in M} sum_{i,j: (l_i,m_i) = (l_j,m_j)}
//     conj(P_R[kR][r][offR(t)+i]) dO[n_i][n_j] P_S[k][s][offS(t)+j].
// dO is block-diagonal in (l, m), so Y = dO P_S is a short sparse sum per
// row; the remaining contraction over the matched-site projector index is
// the complex GEMM conj(X_R) Y^T with X_R the gathered R projections.
// Also the device form of fft3d (linalg.c:14-45) for up to 64 bands.
#include <cstring>

#include "pp_internal.h"

namespace pp {

// X[k][b][t] = PR[kmap[k]][b][gR[t]];  Y[k][s][t] = sum_q coef[q] PS[k][s][colS[q]]
__global__ __launch_bounds__(256) void k_aug_gather(int nkt, int nbR, int KM, const int *__restrict__ kmap,
                                                    const cplx<double> *__restrict__ PR, long ptotR,
                                                    const int *__restrict__ gR, cplx<double> *__restrict__ X) {
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  long total = (long)nkt * nbR * KM;
  if (t >= total)
    return;
  int i = (int)(t % KM);
  long r = t / KM;
  int b = (int)(r % nbR);
  int k = (int)(r / nbR);
  X[t] = PR[((long)kmap[k] * nbR + b) * ptotR + gR[i]];
}

__global__ __launch_bounds__(256) void k_aug_apply(int nkt, int nbS, int KM, int nnz,
                                                   const cplx<double> *__restrict__ PS, long ptotS,
                                                   const int *__restrict__ rowptr, const int *__restrict__ colS,
                                                   const cplx<double> *__restrict__ coef,
                                                   cplx<double> *__restrict__ Y) {
  long t = (long)blockIdx.x * 256 + threadIdx.x;
  long total = (long)nkt * nbS * KM;
  if (t >= total)
    return;
  int i = (int)(t % KM);
  long r = t / KM;  // = k * nbS + s
  const int k = (int)(r / nbS);
  const cplx<double> *ps = PS + r * ptotS;
  const cplx<double> *ck = coef + (long)k * nnz;
  cplx<double> acc = mk<double>(0, 0);
  for (int q = rowptr[i]; q < rowptr[i + 1]; q++)
    acc = acc + ck[q] * ps[colS[q]];
  Y[t] = acc;
}

// out[k] += conj(X[k]) Y[k]^T over the KM gathered R projector rows gR,
// with Y[k][s][i] = sum_q coef[k][q] PS[k][s][colS[q]] (CSR rows rowptr).
static int aug_contract(pp_ctx *ctx, int nwk, int nspin, int flip, const void *projR_dev, int nbR, int ptotR,
                        const void *projS_dev, int nbS, int ptotS, const std::vector<int> &gR,
                        const std::vector<int> &rowptr, const std::vector<int> &colS,
                        const std::vector<cplx<double>> &coef, void *out_dev) {
  const int KM = (int)gR.size();
  const int nkt = nwk * nspin;
  const int nnz = (int)colS.size();
  if (KM == 0)
    return PP_OK;
  std::vector<int> kmap(nkt);
  for (int k = 0; k < nkt; k++)
    kmap[k] = (flip && nspin == 2) ? (k < nwk ? k + nwk : k - nwk) : k;
  // device tables: ints, then coefficients; X and Y in aug_ws
  std::vector<int> ints;
  ints.insert(ints.end(), gR.begin(), gR.end());
  ints.insert(ints.end(), rowptr.begin(), rowptr.end());
  ints.insert(ints.end(), colS.begin(), colS.end());
  if (colS.empty())
    ints.push_back(0);
  ints.insert(ints.end(), kmap.begin(), kmap.end());
  const size_t ibytes = (ints.size() * sizeof(int) + 255) / 256 * 256;
  const size_t craw = coef.size() * sizeof(cplx<double>);
  const size_t cbytes = (std::max<size_t>(1, coef.size()) * sizeof(cplx<double>) + 255) / 256 * 256;
  const size_t xbytes = sizeof(cplx<double>) * (size_t)nkt * nbR * KM;
  const size_t ybytes = sizeof(cplx<double>) * (size_t)nkt * nbS * KM;
  // the previous call's kernels (on whatever stream they ran) read the
  // tables and X / Y: this call's writes wait for them on the device
  if (!ctx->aug_done)
    PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->aug_done, hipEventDisableTiming));
  if (ctx->aug_pending)
    PP_CHECK_HIP(hipStreamWaitEvent(ctx->stream, ctx->aug_done, 0));
  PP_TRY(ctx->aug_ws.ensure(xbytes + ybytes));
  const bool same = ints == ctx->aug_ints && craw == ctx->aug_coefs.size() &&
                    (craw == 0 || std::memcmp(coef.data(), ctx->aug_coefs.data(), craw) == 0);
  if (!same) {
    // upload through pinned staging: no host sync, and no dependence on
    // how the runtime stages pageable copies
    PP_TRY(ctx->aug_tab.ensure(ibytes + cbytes));
    if (!ctx->aug_upload)
      PP_CHECK_HIP(hipEventCreateWithFlags(&ctx->aug_upload, hipEventDisableTiming));
    else
      PP_CHECK_HIP(hipEventSynchronize(ctx->aug_upload));  // the staging buffer is free again
    if (ctx->aug_pinned_bytes < ibytes + cbytes) {
      if (ctx->aug_pinned)
        PP_CHECK_HIP(hipHostFree(ctx->aug_pinned));
      ctx->aug_pinned = nullptr;
      ctx->aug_pinned_bytes = 0;
      PP_CHECK_HIP(hipHostMalloc(&ctx->aug_pinned, ibytes + cbytes, hipHostMallocDefault));
      ctx->aug_pinned_bytes = ibytes + cbytes;
    }
    char *h = (char *)ctx->aug_pinned;
    std::memcpy(h, ints.data(), ints.size() * sizeof(int));
    if (craw)
      std::memcpy(h + ibytes, coef.data(), craw);
    PP_CHECK_HIP(hipMemcpyAsync(ctx->aug_tab.p, h, ibytes + craw, hipMemcpyHostToDevice, ctx->stream));
    PP_CHECK_HIP(hipEventRecord(ctx->aug_upload, ctx->stream));
    ctx->aug_ints = ints;
    ctx->aug_coefs.assign((const char *)coef.data(), (const char *)coef.data() + craw);
  }
  char *base = (char *)ctx->aug_tab.p;
  int *d_gR = (int *)base;
  int *d_rowptr = d_gR + gR.size();
  int *d_colS = d_rowptr + rowptr.size();
  int *d_kmap = d_colS + std::max<size_t>(1, colS.size());
  cplx<double> *d_coef = (cplx<double> *)(base + ibytes);
  cplx<double> *X = (cplx<double> *)ctx->aug_ws.p;
  cplx<double> *Y = (cplx<double> *)((char *)ctx->aug_ws.p + xbytes);
  hipEvent_t ev;
  long tx = (long)nkt * nbR * KM, ty = (long)nkt * nbS * KM;
  tbegin(ctx, K_AUG_PREP, &ev);
  hipLaunchKernelGGL(k_aug_gather, dim3((unsigned)((tx + 255) / 256)), dim3(256), 0, ctx->stream, nkt, nbR, KM,
                     d_kmap, (const cplx<double> *)projR_dev, (long)ptotR, d_gR, X);
  hipLaunchKernelGGL(k_aug_apply, dim3((unsigned)((ty + 255) / 256)), dim3(256), 0, ctx->stream, nkt, nbS, KM, nnz,
                     (const cplx<double> *)projS_dev, (long)ptotS, d_rowptr, d_colS, d_coef, Y);
  tend(ctx, K_AUG_PREP, ev);
  PP_CHECK_HIP(hipGetLastError());
  PP_TRY(zgemm_nt(ctx, PP_FP64, nbR, nbS, KM, 0, X, KM, (long)nbR * KM, 0, Y, KM, (long)nbS * KM, out_dev, nbS,
                  (long)nbR * nbS, nkt, nullptr, nullptr, nullptr, true));
  PP_CHECK_HIP(hipEventRecord(ctx->aug_done, ctx->stream));
  ctx->aug_pending = true;
  return PP_OK;
}

}  // namespace pp

using namespace pp;

extern "C" int pp_augmentation(pp_ctx *ctx, int nwk, int nspin, int flip, const void *projR_dev, int nbR, int ptotR,
                               const int *proj_off_R, const int *labels_R, const void *projS_dev, int nbS, int ptotS,
                               const int *proj_off_S, const int *labels_S, int num_M, const int *M_R,
                               const int *M_S, int num_elems, const int *nchan, const int *ls, const double *dO,
                               void *out_dev) {
  PP_REQUIRE(ctx && nwk > 0 && (nspin == 1 || nspin == 2) && nbR >= 0 && nbS >= 0 && num_M >= 0, PP_ERR_ARG,
             "pp_augmentation: bad arguments");
  if (num_M == 0 || nbR == 0 || nbS == 0)
    return PP_OK;
  PP_REQUIRE(projR_dev && projS_dev && out_dev && proj_off_R && proj_off_S && labels_R && labels_S && M_R && M_S &&
                 nchan && ls && dO,
             PP_ERR_ARG, "pp_augmentation: null argument");
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  // species tables: projector (n, l, m) lists and dO offsets
  std::vector<int> ls_off(num_elems), do_off(num_elems);
  std::vector<std::vector<int>> pn(num_elems), pl(num_elems), pm(num_elems);
  int lo = 0, dofs = 0;
  for (int e = 0; e < num_elems; e++) {
    ls_off[e] = lo;
    do_off[e] = dofs;
    for (int ch = 0; ch < nchan[e]; ch++) {
      int l = ls[lo + ch];
      for (int m = -l; m <= l; m++) {
        pn[e].push_back(ch);
        pl[e].push_back(l);
        pm[e].push_back(m);
      }
    }
    lo += nchan[e];
    dofs += nchan[e] * nchan[e];
  }
  // matched-site projector index space and the sparse rows of dO P_S
  std::vector<int> gR, rowptr(1, 0), colS;
  std::vector<double> coef;
  for (int t = 0; t < num_M; t++) {
    int s1 = M_R[t], s2 = M_S[t];
    int eR = labels_R[s1], eS = labels_S[s2];
    PP_REQUIRE(eR >= 0 && eR < num_elems && eS >= 0 && eS < num_elems, PP_ERR_ARG,
               "pp_augmentation: species label out of range");
    int nR = (int)pn[eR].size(), nS = (int)pn[eS].size();
    PP_REQUIRE(proj_off_R[s1] + nR <= ptotR && proj_off_S[s2] + nS <= ptotS, PP_ERR_ARG,
               "pp_augmentation: projector offsets exceed the projection table");
    const double *D = dO + do_off[eR];
    int nc = nchan[eR];
    for (int i = 0; i < nR; i++) {
      gR.push_back(proj_off_R[s1] + i);
      for (int j = 0; j < nS; j++) {
        if (pl[eR][i] != pl[eS][j] || pm[eR][i] != pm[eS][j])
          continue;
        int ni = pn[eR][i], nj = pn[eS][j];
        PP_REQUIRE(nj < nc, PP_ERR_ARG, "pp_augmentation: channel index outside dO");
        colS.push_back(proj_off_S[s2] + j);
        coef.push_back(D[ni * nc + nj]);
      }
      rowptr.push_back((int)colS.size());
    }
  }
  // the same real coefficients at every k
  const int nkt = nwk * nspin;
  std::vector<cplx<double>> cf((size_t)nkt * coef.size());
  for (int k = 0; k < nkt; k++)
    for (size_t q = 0; q < coef.size(); q++)
      cf[(size_t)k * coef.size() + q] = mk<double>(coef[q], 0.0);
  return aug_contract(ctx, nwk, nspin, flip, projR_dev, nbR, ptotR, projS_dev, nbS, ptotS, gR, rowptr, colS, cf,
                      out_dev);
}

extern "C" int pp_augmentation_offsite(pp_ctx *ctx, int nwk, int nspin, int flip, const void *projR_dev, int nbR,
                                       int ptotR, const int *proj_off_R, const void *projS_dev, int nbS, int ptotS,
                                       const int *proj_off_S, int num_pairs, const int *N_RS_R, const int *N_RS_S,
                                       const pp_cdouble *ov, const long *ov_off, const double *dcoords,
                                       const double *kpts, void *out_dev) {
  PP_REQUIRE(ctx && nwk > 0 && (nspin == 1 || nspin == 2) && nbR >= 0 && nbS >= 0 && num_pairs >= 0, PP_ERR_ARG,
             "pp_augmentation_offsite: bad arguments");
  if (num_pairs == 0 || nbR == 0 || nbS == 0)
    return PP_OK;
  PP_REQUIRE(projR_dev && projS_dev && out_dev && proj_off_R && proj_off_S && N_RS_R && N_RS_S && ov && ov_off &&
                 dcoords && kpts,
             PP_ERR_ARG, "pp_augmentation_offsite: null argument");
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  const int nkt = nwk * nspin;
  std::vector<int> gR, rowptr(1, 0), colS;
  std::vector<std::pair<int, long>> src;  // (pair, entry of its overlap block) per nonzero
  for (int t = 0; t < num_pairs; t++) {
    const int s1 = N_RS_R[t], s2 = N_RS_S[t];
    const int n1 = proj_off_R[s1 + 1] - proj_off_R[s1], n2 = proj_off_S[s2 + 1] - proj_off_S[s2];
    PP_REQUIRE(n1 >= 0 && n2 >= 0 && proj_off_R[s1] + n1 <= ptotR && proj_off_S[s2] + n2 <= ptotS, PP_ERR_ARG,
               "pp_augmentation_offsite: projector offsets exceed the projection table");
    PP_REQUIRE(ov_off[t + 1] - ov_off[t] == (long)n1 * n2, PP_ERR_ARG,
               "pp_augmentation_offsite: overlap block size differs from the projector counts");
    for (int i = 0; i < n1; i++) {
      gR.push_back(proj_off_R[s1] + i);
      for (int j = 0; j < n2; j++) {
        colS.push_back(proj_off_S[s2] + j);
        src.emplace_back(t, ov_off[t] + (long)i * n2 + j);
      }
      rowptr.push_back((int)colS.size());
    }
  }
  // coef[k][q] = Ov e^{2 pi i kR . dcoord} with kR the (spin-flipped) k of R
  std::vector<cplx<double>> cf((size_t)nkt * src.size());
  for (int k = 0; k < nkt; k++) {
    const int kR = (flip && nspin == 2) ? (k < nwk ? k + nwk : k - nwk) : k;
    for (size_t q = 0; q < src.size(); q++) {
      const double *dc = dcoords + 3 * src[q].first;
      const double a = 2.0 * M_PI * (kpts[3 * kR] * dc[0] + kpts[3 * kR + 1] * dc[1] + kpts[3 * kR + 2] * dc[2]);
      const pp_cdouble o = ov[src[q].second];
      cf[(size_t)k * src.size() + q] = mk<double>(o.real(), o.imag()) * mk<double>(cos(a), sin(a));
    }
  }
  return aug_contract(ctx, nwk, nspin, flip, projR_dev, nbR, ptotR, projS_dev, nbS, ptotS, gR, rowptr, colS, cf,
                      out_dev);
}

extern "C" int pp_fft3d_batched(pp_ctx *ctx, const double *lattice, const int *fftg, int nw, const int *G_dev,
                                const void *C_dev, long ldc, int nband, void *grid_out_dev, int precision) {
  PP_REQUIRE(ctx && lattice && fftg && nw >= 0 && nband >= 0 && nband <= 64, PP_ERR_ARG,
             "pp_fft3d_batched: bad arguments (nband must be <= 64)");
  PP_REQUIRE(nw <= ldc, PP_ERR_ARG, "pp_fft3d_batched: nw exceeds ldc");
  if (nband == 0)
    return PP_OK;
  PP_CHECK_HIP(hipSetDevice(ctx->device));
  FftPlan *pl = nullptr;
  PP_TRY(get_fft_plan(ctx, fftg, &pl));
  const long npts = (long)fftg[0] * fftg[1] * fftg[2];
  int NB = nband;
  bool f64 = precision == PP_FP64;
  size_t csz = f64 ? sizeof(cplx<double>) : sizeof(cplx<float>);
  PP_TRY(ctx->grid.ensure(csz * (size_t)npts * NB));
  PP_TRY(ctx->gidx.ensure(sizeof(int) * (size_t)std::max(1, nw)));
  PP_TRY(compute_gidx(ctx, fftg, nw, G_dev, (int *)ctx->gidx.p));
  if (f64)
    PP_TRY(fft_chunk<double>(ctx, pl, (cplx<double> *)ctx->grid.p, NB, nw, (const int *)ctx->gidx.p,
                             (const cplx<float> *)C_dev, ldc, nband, nullptr));
  else
    PP_TRY(fft_chunk<float>(ctx, pl, (cplx<float> *)ctx->grid.p, NB, nw, (const int *)ctx->gidx.p,
                            (const cplx<float> *)C_dev, ldc, nband, nullptr));
  return grid_to_bandmajor(ctx, f64 ? PP_FP64 : PP_FP32, ctx->grid.p, NB, nband, npts, 1.0 / sqrt(det3(lattice)),
                           grid_out_dev);
}
