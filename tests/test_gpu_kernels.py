"""GPU parity of the HIP kernels against the CPU oracle, through the C-ABI.

Tolerances (written per test): FP64 path <= 1e-9 relative (normwise,
max|d| / max|ref|) against the FP64 oracle -- in practice ~1e-14; FP32
path <= 1e-5 for FFT / projections and <= 1e-6 for the overlap GEMM;
split-plane overlaps: bf16x3 <= 2e-5, bf16x6 and fp16x3 <= 1e-6.
"""

import numpy as np
import pytest

from helpers import device_inputs, normwise, oracle_projections
from oracle import pyoracle as orc
from pawpyseed_amd import synthetic

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pawpyseed_amd.core.batched import Context
    c = Context(0)
    yield c
    c.close()


def _fft_case(ctx, wl, nb, precision):
    from pawpyseed_amd.core import batched as B
    G = wl.ensure_gvecs()[0]
    C = wl.coefficients("R", nband=nb)[0]
    Cd, Gd, gofs, nw = device_inputs([G], [C])
    out = torch.zeros((nb,) + tuple(int(n) for n in wl.fftg), dtype=torch.complex128, device="cuda")
    B.fft3d(ctx, wl.lattice, wl.fftg, Gd, Cd[0], out, precision)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    errs = []
    for b in range(nb):
        want = orc.fft3d(wl.lattice, G, C[b], wl.fftg)
        errs.append(normwise(got[b], want))
    return max(errs)


@pytest.mark.parametrize("fftg,encut", [((20, 20, 20), 150.0), ((24, 18, 30), 120.0), ((28, 14, 21), 60.0),
                                        ((60, 60, 60), 400.0), ((26, 22, 33), 90.0), ((48, 40, 64), 300.0),
                                        ((36, 54, 32), 250.0)])
def test_fft3d_fp64(ctx, fftg, encut):
    """Grids whose axes factor into radices 2-5 only, composite radices
    (6, 8, 9, 12, 16: e.g. 108 = 12 x 9, 64 = 16 x 4, 40 = 8 x 5) and other
    primes (7, 11, 13, generic stages)."""
    wl = synthetic.tiny_config(encut=encut, fftg=fftg)
    print("fft3d fp64", fftg, flush=True)
    assert _fft_case(ctx, wl, 5, "fp64") < 1e-12


@pytest.mark.parametrize("fftg,encut", [((24, 18, 30), 120.0), ((60, 60, 60), 400.0), ((48, 40, 64), 300.0),
                                        ((36, 54, 32), 250.0)])
def test_fft3d_composite_radices(ctx, fftg, encut):
    """The opt-in composite factorization (fft_radix = 1: 12 x 2, 9 x 2,
    6 x 5, 12 x 5, 16 x 3, 8 x 5, 16 x 4, 12 x 3, 9 x 6, 16 x 2 here)."""
    wl = synthetic.tiny_config(encut=encut, fftg=fftg)
    ctx.set_param("fft_radix", 1)
    try:
        err64 = _fft_case(ctx, wl, 5, "fp64")
        err32 = _fft_case(ctx, wl, 3, "fp32")
    finally:
        ctx.set_param("fft_radix", 0)
    print("fft3d composite", fftg, err64, err32, flush=True)
    assert err64 < 1e-12 and err32 < 1e-5


def test_fft3d_fp32(ctx):
    wl = synthetic.tiny_config(encut=400.0, fftg=(60, 60, 60))
    err = _fft_case(ctx, wl, 3, "fp32")
    print("fft3d fp32 normwise error", err)
    assert err < 1e-5


def test_fft3d_reference_wavecar(ctx, wavecar):
    """tests.c:28-80 fft_check on test_files/WAVECAR (k0, band 0, 20^3)."""
    from pawpyseed_amd.core import batched as B
    G = wavecar["G_0"]
    C = wavecar["C_0"]
    Cd, Gd, gofs, nw = device_inputs([G], [C])
    out = torch.zeros((C.shape[0], 20, 20, 20), dtype=torch.complex128, device="cuda")
    B.fft3d(ctx, wavecar["lattice"], wavecar["grid_fftg"], Gd, Cd[0], out, "fp64")
    got = out.cpu().numpy()
    want = wavecar["grid_k0_b0"]
    assert np.abs(got[0] - want).max() < 1e-10 * np.abs(want).max()
    for b in range(1, C.shape[0]):
        assert normwise(got[b], orc.fft3d(wavecar["lattice"], G, C[b], wavecar["grid_fftg"])) < 1e-12


def _proj_case(ctx, wl, frac, nb, precision, labels=None):
    from pawpyseed_amd.core import batched as B
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("S", nband=nb)
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, frac, np.zeros(len(frac), np.int32) if labels is None else labels,
                        [wl.species], precision)
    Cd, Gd, gofs, nw = device_inputs(Gs, Cs)
    out = torch.zeros((len(Gs), nb, plan.ptot), dtype=torch.complex128, device="cuda")
    B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, out, precision)
    torch.cuda.synchronize()
    want, T = oracle_projections(wl, frac, Cs)
    assert plan.npts == len(T.idx)
    assert np.array_equal(plan.proj_off, T.proj_off)
    got = out.cpu().numpy()
    plan.close()
    return got, want


def test_projections_fp64_tiny_two_k(ctx):
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=8, kpts=[[0, 0, 0], [0.25, -0.25, 0.1]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    got, want = _proj_case(ctx, wl, wl.frac_S, 8, "fp64")
    err = normwise(got, want)
    print("projections fp64 tiny", err)
    assert err < 1e-9


def test_projections_fp64_ragged_band_chunk(ctx):
    """nband not a multiple of the 64-band chunk, and > 64."""
    wl = synthetic.tiny_config(encut=100.0, fftg=(20, 20, 20), nband=70, kpts=[[0.1, 0.2, 0.3]], ls=(0, 1),
                               alphas=(1.5, 3.0))
    got, want = _proj_case(ctx, wl, wl.frac_S, 70, "fp64")
    assert normwise(got, want) < 1e-9


def test_projections_two_species(ctx):
    """Sites of two elements with different projector sets (9 and 18
    projectors, different radii) in one plan; FP64 and FP32 paths."""
    from pawpyseed_amd.core import batched as B
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=8, kpts=[[0, 0, 0], [0.25, -0.25, 0.1]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    sp2 = synthetic.make_species([0, 0, 1, 1, 2, 2], [1.2, 2.5, 1.4, 2.8, 1.6, 3.0], 0.95)
    species = [wl.species, sp2]
    labels = (np.arange(len(wl.frac_S)) % 2).astype(np.int32)
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("S", nband=8)
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S, species, labels)
    want = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, Cs, T)
    Cd, Gd, gofs, nw = device_inputs(Gs, Cs)
    for precision, tol in (("fp64", 1e-9), ("fp32", 1e-5)):
        plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, labels, species, precision)
        assert np.array_equal(plan.proj_off, T.proj_off)
        out = torch.zeros((len(Gs), 8, plan.ptot), dtype=torch.complex128, device="cuda")
        B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, out, precision)
        torch.cuda.synchronize()
        err = normwise(out.cpu().numpy(), want)
        print("projections two species", precision, err, flush=True)
        plan.close()
        assert err < tol


def test_projections_fp64_cfg0(ctx):
    """64-site cell of the 400 eV / 60^3 workload, Gamma, 16 bands."""
    wl = synthetic.config0()
    got, want = _proj_case(ctx, wl, wl.frac_S, 16, "fp64")
    err = normwise(got, want)
    print("projections fp64 cfg0", err)
    assert err < 1e-9


def test_projections_nonorthogonal_cell(ctx, wavecar):
    """Hexagonal Ga4 cell of test_files (CONTCAR positions), its WAVECAR
    coefficients, synthetic l=0..2 projectors."""
    frac = np.array([[0.1560427174376784, 0.8439572825631387, 0.9166540580595424],
                     [0.3439572825623216, 0.6560427174368613, 0.4166540580595424],
                     [0.8439572825623216, 0.1560427174368613, 0.0833459419404576],
                     [0.6560427174376784, 0.3439572825631387, 0.5833459419404576]])
    sp = synthetic.make_species([0, 1, 2], [1.5, 2.0, 2.5], 1.2)
    wl = synthetic.Workload("ga4", wavecar["lattice"], frac, frac, sp, np.array([20, 20, 20], np.int32),
                            float(wavecar["encut"]), wavecar["kpts"], np.array([0.5, 0.5]), 6, np.eye(3),
                            np.arange(4), np.arange(4), gvecs=[wavecar["G_0"], wavecar["G_1"]])
    from pawpyseed_amd.core import batched as B
    Cs = [wavecar["C_0"], wavecar["C_1"]]
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, frac, np.zeros(4, np.int32), [sp], "fp64")
    Cd, Gd, gofs, nw = device_inputs(wl.gvecs, Cs)
    out = torch.zeros((2, 6, plan.ptot), dtype=torch.complex128, device="cuda")
    B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, out, "fp64")
    want, T = oracle_projections(wl, frac, Cs)
    assert normwise(out.cpu().numpy(), want) < 1e-9


def test_projections_fp32(ctx):
    wl = synthetic.config0()
    got, want = _proj_case(ctx, wl, wl.frac_S, 8, "fp32")
    err = normwise(got, want)
    print("projections fp32 cfg0 normwise error", err)
    assert err < 1e-5


def _overlap_case(ctx, precision, nbR=70, nbS=45, nspin=1, flip=False):
    from pawpyseed_amd.core import batched as B
    kp = [[0, 0, 0], [0.25, 0.25, 0.0]]
    wl = synthetic.tiny_config(encut=250.0, nband=nbR, kpts=kp)
    rng = np.random.default_rng(7)
    CR = [c for c in wl.coefficients("R", rng=rng, nband=nbR)] * nspin
    CS = [c for c in wl.coefficients("S", rng=rng, nband=nbS)] * nspin
    if nspin == 2:
        CR = CR[:2] + [c * np.exp(0.3j) for c in CR[2:]]
    nwk = len(kp)
    CRd, _, _, nw = device_inputs(wl.ensure_gvecs() * nspin, CR)
    CSd, _, _, _ = device_inputs(wl.ensure_gvecs() * nspin, CS)
    out = torch.zeros((nwk * nspin, nbR, nbS), dtype=torch.complex128, device="cuda")
    B.overlap_gemm(ctx, nwk, nspin, flip, nw, CRd, CSd, out, precision)
    got = out.cpu().numpy()
    nkt = nwk * nspin
    want = np.zeros_like(got)
    for band in range(nbS):
        res = orc.pseudoprojection(CR, CS, band, nwk, nspin, flip=flip)
        want[:, :, band] = res.reshape(nbR, nkt).T
    return got, want


def test_overlap_fp64(ctx):
    got, want = _overlap_case(ctx, "fp64")
    assert np.abs(got - want).max() < 1e-12


def test_overlap_fp64_spin_flip(ctx):
    got, want = _overlap_case(ctx, "fp64", nbR=33, nbS=20, nspin=2, flip=True)
    assert np.abs(got - want).max() < 1e-12


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-6), ("bf16x3", 2e-5), ("bf16x6", 1e-6), ("fp16x3", 1e-6)])
def test_overlap_mixed(ctx, precision, tol):
    got, want = _overlap_case(ctx, precision, nbR=130, nbS=67)
    err = normwise(got, want)
    print("overlap", precision, "normwise error", err)
    assert err < tol


@pytest.mark.parametrize("precision", ["fp32", "fp16x3", "bf16x6"])
def test_overlap_split_planes_row_scales(ctx, precision):
    """Split-plane overlaps of rows whose magnitudes span eleven decades (and
    one zero row), 150 S bands (two 128-column tiles), two plane chunks of
    K, against FP64: the fp16 split scales each 1024-term chunk of a row by
    its own power of two, so no row loses accuracy to the others'
    magnitudes.  With the overlaps divided by the row scales of both
    operands, normwise error <= 1e-6 and every R row's own normwise error
    <= 3e-6 (the FP32 GEMM, held to the same bounds, is the control); the
    zero row's overlaps are exactly zero."""
    from pawpyseed_amd.core import batched as B
    wl = synthetic.tiny_config(encut=250.0, nband=20, kpts=[[0.1, 0.2, -0.3]])
    rng = np.random.default_rng(23)
    G = wl.ensure_gvecs()[0]
    assert len(G) > 1024
    c = wl.coefficients("R", rng=rng, nband=20)[0]
    d = wl.coefficients("S", rng=rng, nband=150)[0]
    scale_r = 10.0 ** rng.uniform(-8, 3, 20)
    scale_r[5] = 0.0
    scale_s = 10.0 ** rng.uniform(-8, 3, 150)
    CR = [(c * scale_r[:, None]).astype(np.complex64)]
    CS = [(d * scale_s[:, None]).astype(np.complex64)]
    CRd, _, _, nw = device_inputs([G], CR)
    CSd, _, _, _ = device_inputs([G], CS)
    out = torch.zeros((1, 20, 150), dtype=torch.complex128, device="cuda")
    B.overlap_gemm(ctx, 1, 1, False, nw, CRd, CSd, out, precision)
    got = out.cpu().numpy()[0]
    want = np.conj(CR[0].astype(np.complex128)) @ CS[0].astype(np.complex128).T
    assert np.all(got[5] == 0)
    rows = [m for m in range(20) if m != 5]
    norm = np.outer(scale_r[rows], scale_s)
    g, w = got[rows] / norm, want[rows] / norm
    err = normwise(g, w)
    row_err = max(np.linalg.norm(g[m] - w[m]) / np.linalg.norm(w[m]) for m in range(len(rows)))
    print(f"{precision} row-scaled overlap: normwise {err:.2e}, worst row {row_err:.2e}", flush=True)
    assert err < 1e-6 and row_err < 3e-6


@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_overlap_split_planes_work_group_variants(ctx, precision):
    """The two-plane GEMM with one work-group per CU (64-real k-steps), with
    two (32-real k-steps), with two whose GEMM splits the FP32 operands
    itself (no plane buffer), and with 256 x 128 tiles of 8 wavefronts that
    split them too: every output accumulates the same products of the same
    planes in the same order, so the results are bit-identical; 300 x 67
    bands (a partial 256-row tile), two k-points with odd plane-wave counts
    (a partial last chunk), against FP64 <= 1e-6 (fp16x3) / 2e-5 (bf16x3)."""
    before = ctx.get_param("gemm_planes")
    try:
        out = {}
        for gp in (1, 2, 3, 4):
            ctx.set_param("gemm_planes", gp)
            out[gp], want = _overlap_case(ctx, precision, nbR=300, nbS=67)
    finally:
        ctx.set_param("gemm_planes", before)
    err = normwise(out[2], want)
    same = {gp: np.array_equal(out[gp], out[2]) for gp in (1, 3, 4)}
    print(precision, "gemm_planes 1 / 3 / 4 vs 2 identical:", same, "error", err, flush=True)
    assert all(same.values())
    assert err < (1e-6 if precision == "fp16x3" else 2e-5)


def test_overlap_near_orthogonal_pairs(ctx):
    """Pairs of nearly orthogonal bands: S band i = d_i + eps_i c_i with d_i
    orthogonal to R band c_i, eps_i from 1e-8 to 1, so overlap entries span
    eight decades.  FP64: elementwise relative error <= 1e-6 on every entry
    with |O| > 1e-8; FP32 / 3xBF16 are held to their normwise bounds."""
    from pawpyseed_amd.core import batched as B
    wl = synthetic.tiny_config(encut=250.0, nband=9, kpts=[[0.1, -0.2, 0.3]])
    rng = np.random.default_rng(11)
    G = wl.ensure_gvecs()[0]
    c = wl.coefficients("R", rng=rng, nband=9)[0].astype(np.complex128)
    d = wl.coefficients("S", rng=rng, nband=9)[0].astype(np.complex128)
    eps = 10.0 ** -np.arange(9.0)[::-1]  # 1e-8 ... 1
    S = np.empty_like(d)
    for i in range(9):
        di = d[i] - np.vdot(c[i], d[i]) / np.vdot(c[i], c[i]) * c[i]
        di /= np.linalg.norm(di)
        S[i] = di + eps[i] * c[i] / np.linalg.norm(c[i])
    CR, CS = [c.astype(np.complex64)], [S.astype(np.complex64)]
    CRd, _, _, nw = device_inputs([G], CR)
    CSd, _, _, _ = device_inputs([G], CS)
    want = np.stack([orc.pseudoprojection(CR, CS, b, 1, 1) for b in range(9)], axis=1)  # [bR, bS]
    diag = np.abs(np.diag(want))
    assert diag.min() < 1e-7 and diag.max() > 0.5  # the stored pairs really span the range
    for precision, tol in (("fp64", None), ("fp32", 1e-6), ("bf16x3", 2e-5), ("bf16x6", 1e-6), ("fp16x3", 1e-6)):
        out = torch.zeros((1, 9, 9), dtype=torch.complex128, device="cuda")
        B.overlap_gemm(ctx, 1, 1, False, nw, CRd, CSd, out, precision)
        got = out.cpu().numpy()[0]
        big = np.abs(want) > 1e-8
        rel = float((np.abs(got - want)[big] / np.abs(want)[big]).max())
        print("near-orthogonal overlap", precision, "max elementwise rel err (|O|>1e-8)", rel,
              "normwise", normwise(got, want), flush=True)
        if tol is None:
            assert rel < 1e-6
        else:
            assert normwise(got, want) < tol


def test_uneven_structures_onto_one_basis(ctx):
    """Three perturbed structures with different band counts onto one
    basis, at random (non-grid) k-points whose plane-wave counts differ:
    projections of every structure, band overlaps (FP64) and O_M, checked
    against the oracle's setup_projections, pseudoprojection and
    compensation_terms."""
    from pawpyseed_amd.core import batched as B
    rng = np.random.default_rng(5)
    kp = rng.uniform(-0.5, 0.5, (3, 3))
    wl = synthetic.tiny_config(encut=140.0, fftg=(24, 24, 24), nband=10, kpts=kp, ls=(0, 1, 2),
                               alphas=(1.5, 3.0, 2.0))
    Gs = wl.ensure_gvecs()
    nk = len(Gs)
    assert len({len(g) for g in Gs}) > 1
    nbR = 10
    CR = wl.coefficients("R", rng=rng, nband=nbR)
    frac_R = wl.frac_R
    planR = B.SpherePlan(ctx, wl.lattice, wl.fftg, frac_R, np.zeros(len(frac_R), np.int32), [wl.species], "fp64")
    CRd, Gd, gofs, nw = device_inputs(Gs, CR)
    projR = torch.zeros((nk, nbR, planR.ptot), dtype=torch.complex128, device="cuda")
    B.setup_projections(ctx, planR, wl.kpts, gofs, Gd, CRd, projR, "fp64")
    wantR, TR = oracle_projections(wl, frac_R, CR)
    assert normwise(projR.cpu().numpy(), wantR) < 1e-9
    sites = np.arange(len(frac_R), dtype=np.int32)
    labels = np.zeros(len(frac_R), np.int32)
    for j, nbS in enumerate((7, 12, 9)):
        cart = frac_R @ wl.lattice + rng.normal(0.0, 0.03, frac_R.shape)
        frac_S = np.mod(cart @ np.linalg.inv(wl.lattice), 1.0)
        CS = wl.coefficients("S", rng=rng, nband=nbS)
        planS = B.SpherePlan(ctx, wl.lattice, wl.fftg, frac_S, labels, [wl.species], "fp64")
        CSd, _, _, _ = device_inputs(Gs, CS)
        projS = torch.zeros((nk, nbS, planS.ptot), dtype=torch.complex128, device="cuda")
        B.setup_projections(ctx, planS, wl.kpts, gofs, Gd, CSd, projS, "fp64")
        wantS, TS = oracle_projections(wl, frac_S, CS)
        assert normwise(projS.cpu().numpy(), wantS) < 1e-9
        O = torch.zeros((nk, nbR, nbS), dtype=torch.complex128, device="cuda")
        B.overlap_gemm(ctx, nk, 1, False, nw, CRd, CSd, O, "fp64")
        B.augmentation(ctx, nk, 1, False, projR, planR.proj_off, labels, projS, planS.proj_off, labels, sites,
                       sites, [wl.species], [wl.delta_o], O)
        got = O.cpu().numpy()
        for band in range(nbS):
            res = orc.pseudoprojection(CR, CS, band, nk, 1)
            res = orc.compensation_terms_m(res, band, nk, 1, wantR, TR, labels, wantS[:, band], TS, sites, sites,
                                           [wl.delta_o], [wl.species.ls])
            want = res.reshape(nbR, nk).T
            assert normwise(got[:, :, band], want) < 1e-9, (j, band)
        print("structure", j, "nbS", nbS, "ok", flush=True)
        planS.close()
    planR.close()


def test_augmentation_matches_oracle(ctx):
    from pawpyseed_amd.core import batched as B
    rng = np.random.default_rng(3)
    sp = synthetic.make_species([0, 0, 1, 1, 2], [1.5, 3.0, 1.5, 3.0, 2.0], 1.2)
    P = sp.total_projs
    nsR, nsS, nwk, nspin, nbR, nbS = 6, 5, 2, 2, 40, 9
    nkt = nwk * nspin
    pR = rng.standard_normal((nkt, nbR, nsR * P)) + 1j * rng.standard_normal((nkt, nbR, nsR * P))
    pS = rng.standard_normal((nkt, nbS, nsS * P)) + 1j * rng.standard_normal((nkt, nbS, nsS * P))
    dO = synthetic.channel_delta_o(rng, sp.ls)
    M_R = np.array([0, 2, 3, 5], np.int32)
    M_S = np.array([1, 0, 4, 2], np.int32)
    offR = np.arange(nsR + 1, dtype=np.int32) * P
    offS = np.arange(nsS + 1, dtype=np.int32) * P
    for flip in (False, True):
        out = torch.zeros((nkt, nbR, nbS), dtype=torch.complex128, device="cuda")
        B.augmentation(ctx, nwk, nspin, flip, torch.from_numpy(pR).cuda(), offR, np.zeros(nsR, np.int32),
                       torch.from_numpy(pS).cuda(), offS, np.zeros(nsS, np.int32), M_R, M_S, [sp], [dO], out)
        got = out.cpu().numpy()

        class T:
            pass

        TR, TS = T(), T()
        TR.ptot, TR.proj_off = nsR * P, offR
        TS.ptot, TS.proj_off = nsS * P, offS
        for band in range(nbS):
            want = orc.compensation_terms_m(np.zeros(nbR * nkt, complex), band, nwk, nspin, pR, TR,
                                            np.zeros(nsR, np.int32), pS[:, band], TS, M_R, M_S, [dO], [sp.ls],
                                            flip=flip)
            np.testing.assert_allclose(got[:, :, band], want.reshape(nbR, nkt).T, atol=1e-11)


def test_augmentation_dense_delta_o_cfg0(ctx):
    """The 64-site cell of the 400 eV / 60^3 workload with a dense ΔO (all
    4 x 4 channel entries random, cross-l included) over all 64 bands: O_M
    uses only the entries between projectors of equal (l, m)
    (projector.c:897-906), so the GPU result equals the oracle's with the
    same dense matrix (1e-11) and, to rounding, the GPU result with the
    cross-l entries zeroed (1e-13)."""
    from pawpyseed_amd.core import batched as B
    wl = synthetic.config0()
    rng = np.random.default_rng(17)
    dense = synthetic.channel_delta_o(rng, wl.species.ls, dense=True)
    ls = np.asarray(wl.species.ls)
    assert np.abs(dense[ls[:, None] != ls[None, :]]).min() > 0  # really dense
    same_l = dense.copy()
    same_l[ls[:, None] != ls[None, :]] = 0.0
    Gs = wl.ensure_gvecs()
    CR = wl.coefficients("R", nband=64)
    CS = wl.coefficients("S", nband=64)
    labels = np.zeros(len(wl.frac_R), np.int32)
    proj = {}
    for name, frac, C in (("R", wl.frac_R, CR), ("S", wl.frac_S, CS)):
        plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, frac, labels, [wl.species], "fp64")
        Cd, Gd, gofs, nw = device_inputs(Gs, C)
        out = torch.zeros((1, 64, plan.ptot), dtype=torch.complex128, device="cuda")
        B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, out, "fp64")
        proj[name] = (out, plan.proj_off.copy())
        plan.close()
    (pR, offR), (pS, offS) = proj["R"], proj["S"]
    got = {}
    for key, dO in (("dense", dense), ("same_l", same_l)):
        o = torch.zeros((1, 64, 64), dtype=torch.complex128, device="cuda")
        B.augmentation(ctx, 1, 1, False, pR, offR, labels, pS, offS, labels, wl.M_R, wl.M_S, [wl.species], [dO], o)
        got[key] = o.cpu().numpy()[0]
    pRh, pSh = pR.cpu().numpy(), pS.cpu().numpy()

    class T:
        pass

    TR, TS = T(), T()
    TR.ptot, TR.proj_off = pRh.shape[2], offR
    TS.ptot, TS.proj_off = pSh.shape[2], offS
    want = np.zeros((64, 64), complex)
    for band in range(64):
        w = orc.compensation_terms_m(np.zeros(64, complex), band, 1, 1, pRh, TR, labels, pSh[:, band], TS,
                                     wl.M_R, wl.M_S, [dense], [wl.species.ls])
        want[:, band] = w
    scale = np.abs(want).max()
    e_or = np.abs(got["dense"] - want).max() / scale
    e_l = np.abs(got["dense"] - got["same_l"]).max() / scale
    print(f"cfg0 dense dO: vs oracle {e_or:.2e}, vs equal-l part {e_l:.2e}", flush=True)
    assert e_or < 1e-11 and e_l < 1e-13


def test_empty_and_invalid_inputs(ctx):
    from pawpyseed_amd.core import batched as B
    from pawpyseed_amd.core._lib import PAWpyError
    wl = synthetic.tiny_config()
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, np.zeros(len(wl.frac_S), np.int32), [wl.species])
    Cd = torch.zeros((1, 0, 64), dtype=torch.complex64, device="cuda")
    out = torch.zeros((1, 0, plan.ptot), dtype=torch.complex128, device="cuda")
    Gd = torch.zeros((1, 3), dtype=torch.int32, device="cuda")
    B.setup_projections(ctx, plan, wl.kpts, np.array([0, 0], np.int32), Gd, Cd, out)  # nband = 0: no-op
    with pytest.raises(PAWpyError):
        B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, np.full(len(wl.frac_S), 3, np.int32), [wl.species])
    with pytest.raises(PAWpyError):  # table precision differs from the call's
        C1 = torch.zeros((1, 1, 64), dtype=torch.complex64, device="cuda")
        o1 = torch.zeros((1, 1, plan.ptot), dtype=torch.complex128, device="cuda")
        B.setup_projections(ctx, plan, wl.kpts, np.array([0, 1], np.int32), Gd, C1, o1, "fp32")
    # zero plane waves: overlap is zero
    o = torch.ones((1, 3, 2), dtype=torch.complex128, device="cuda")
    B.overlap_gemm(ctx, 1, 1, False, [0], torch.zeros((1, 3, 64), dtype=torch.complex64, device="cuda"),
                   torch.zeros((1, 2, 64), dtype=torch.complex64, device="cuda"), o)
    assert torch.count_nonzero(o).item() == 0
    plan.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_xmask_writes_only_covered_points(ctx, precision):
    """The x pass writes only the grid points some projector sphere holds
    (fft_xmask): the plan's coverage equals the union of the oracle's
    sphere points, and the projections are bit-identical with the
    unmasked pass (the sphere sums read covered points only)."""
    from pawpyseed_amd.core import batched as B
    wl = synthetic.config0()
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("S", nband=64)
    labels = np.zeros(len(wl.frac_S), np.int32)
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, labels, [wl.species], precision)
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S, [wl.species], labels)
    union = len(np.unique(T.idx))
    assert plan.coverage() == union
    print("coverage", union, "of", int(np.prod(wl.fftg)), flush=True)
    Cd, Gd, gofs, nw = device_inputs(Gs, Cs)
    outs = []
    for xm in (0, 1):
        ctx.set_param("fft_xmask", xm)
        out = torch.zeros((len(Gs), 64, plan.ptot), dtype=torch.complex128, device="cuda")
        B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, out, precision)
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    ctx.set_param("fft_xmask", 1)
    assert np.array_equal(outs[0], outs[1])
    plan.close()


@pytest.mark.parametrize("wide", [False, True])
@pytest.mark.parametrize("fftg", [(108, 60, 216), (192, 108, 60)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_fixed_length_passes(ctx, fftg, precision, tol, wide):
    """The FFT passes compiled for the axis lengths 60, 108, 192 and 216
    (fft_fixed = 1, the default), with the default band tile (16 FP32, 8
    FP64) and twice it, give the projections of the generic passes
    (fft_fixed = 0) to rounding (the compiler contracts the complex
    products into FMAs differently in the two forms, so they are not
    bit-identical) and the oracle's (first 8 bands).  64 bands: a full
    chunk, so the tiles are the ones the fixed passes are compiled for."""
    from pawpyseed_amd.core import batched as B
    nb = 64
    wl = synthetic.tiny_config(encut=150.0, fftg=fftg, nband=nb, kpts=[[0.25, -0.25, 0.1]], ls=(0, 1, 2),
                               alphas=(1.5, 3.0, 1.5))
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("S", nband=nb)
    labels = np.zeros(len(wl.frac_S), np.int32)
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S, [wl.species], labels)
    want = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, [c[:8] for c in Cs], T)
    Cd, Gd, gofs, nw = device_inputs(Gs, Cs)
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, labels, [wl.species], precision)
    outs = []
    ctx.set_param("fft_tile", (32 if precision == "fp32" else 16) if wide else 0)
    for fx in (0, 1):
        ctx.set_param("fft_fixed", fx)
        out = torch.zeros((1, nb, plan.ptot), dtype=torch.complex128, device="cuda")
        B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, out, precision)
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    ctx.set_param("fft_fixed", 1)
    ctx.set_param("fft_tile", 0)
    err = normwise(outs[1][:, :8], want)
    diff = normwise(outs[1], outs[0])
    print("fixed-length passes", fftg, precision, "wide" if wide else "", err, "against generic", diff, flush=True)
    assert diff < tol / 10
    assert err < tol
    plan.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_projections_g_set_with_holes(ctx, precision, tol):
    """A plane-wave list that is not a cutoff sphere (a random subset of
    it, so z-line intervals have holes): the pruned passes, which load no
    zero fill, must not be used (the interval-fill check on the host
    bounds), and the projections still equal the oracle's."""
    from pawpyseed_amd.core import batched as B
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=6, kpts=[[0.25, -0.25, 0.1]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    G = wl.ensure_gvecs()[0]
    keep = np.sort(np.random.default_rng(3).choice(len(G), size=len(G) * 2 // 3, replace=False))
    Gs = [np.ascontiguousarray(G[keep])]
    Cs = [wl.coefficients("S", nband=6)[0][:, keep]]
    labels = np.zeros(len(wl.frac_S), np.int32)
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S, [wl.species], labels)
    want = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, Cs, T)
    Cd, Gd, gofs, nw = device_inputs(Gs, Cs)
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, labels, [wl.species], precision)
    out = torch.zeros((1, 6, plan.ptot), dtype=torch.complex128, device="cuda")
    # garbage in the work buffers first: a pruned pass would read it
    big = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=64, kpts=[[0, 0, 0]])
    Cb, Gb, gob, _ = device_inputs(big.ensure_gvecs(), big.coefficients("R", nband=64))
    junk = torch.zeros((1, 64, plan.ptot), dtype=torch.complex128, device="cuda")
    B.setup_projections(ctx, plan, big.kpts, gob, Gb, Cb, junk, precision)
    B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, out, precision)
    torch.cuda.synchronize()
    err = normwise(out.cpu().numpy(), want)
    print("projections with a holed G set", precision, err, flush=True)
    assert err < tol
    plan.close()
