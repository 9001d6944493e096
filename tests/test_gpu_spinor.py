"""Spinor projections through the batched device API at the size of the
large noncollinear workload: 512 sites, a = 21.72 A, 192^3 grid, 2
k-points, W = 275,953 plane waves per spinor component, 18 projectors per
site (the workload's 2048 bands cut to 16 here).  Each component is
projected on its own (onto_projector_ncl, projector.c:359-392); sampled
bands are checked against the FP64 oracle.

Tolerances (normwise, max|d| / max|ref|): FP64 path 1e-9, FP32 path 1e-5.
"""

import numpy as np
import pytest

from helpers import normwise
from oracle import pyoracle as orc
from pawpyseed_amd import synthetic

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NB = 16
SAMPLE = (0, 11)


@pytest.fixture(scope="module")
def ctx():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pawpyseed_amd.core.batched import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def case():
    wl = synthetic.config4(nband=NB)
    Gs = wl.ensure_gvecs()
    Cu = wl.coefficients("S", nband=NB)
    Cd = wl.coefficients("R", nband=NB)
    labels = np.zeros(len(wl.frac_S), np.int32)
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S, [wl.species], labels)
    want_up = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, [c[list(SAMPLE)] for c in Cu], T)
    want_dn = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, [c[list(SAMPLE)] for c in Cd], T)
    return wl, Gs, Cu, Cd, labels, want_up, want_dn


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_spinor_projections_large_cell(ctx, case, precision, tol):
    from pawpyseed_amd.core import batched as B
    wl, Gs, Cu, Cd, labels, want_up, want_dn = case
    nk = len(Gs)
    W = [len(g) for g in Gs]
    ldc = -(-2 * max(W) // 64) * 64
    C2 = np.zeros((nk, NB, ldc), np.complex64)
    for k in range(nk):
        C2[k, :, :W[k]] = Cu[k]
        C2[k, :, W[k]:2 * W[k]] = Cd[k]
    C2d = torch.from_numpy(C2).cuda()
    gofs = np.concatenate([[0], np.cumsum(W)]).astype(np.int32)
    Gd = torch.from_numpy(np.ascontiguousarray(np.concatenate(Gs), np.int32)).cuda()
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, labels, [wl.species], precision)
    up = torch.zeros((nk, NB, plan.ptot), dtype=torch.complex128, device="cuda")
    dn = torch.zeros_like(up)
    torch.cuda.synchronize()
    print("spinor case", precision, "sites", len(wl.frac_S), "W", W, "grid", list(wl.fftg), flush=True)
    B.setup_projections_spinor(ctx, plan, wl.kpts, gofs, Gd, C2d, up, dn, precision)
    got_up = up.cpu().numpy()[:, list(SAMPLE)]
    got_dn = dn.cpu().numpy()[:, list(SAMPLE)]
    eu, ed = normwise(got_up, want_up), normwise(got_dn, want_dn)
    print("spinor", precision, "up", eu, "down", ed, flush=True)
    assert eu < tol and ed < tol
    plan.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-6), ("bf16x6", 1e-6), ("fp16x3", 1e-6)])
def test_spinor_overlap_large_cell(ctx, case, precision, tol):
    """The band overlap of spinor bands (pseudoprojector.c:63-90 over the
    reference band's num_waves = 2 W, reader.c:274-281: the up then the down
    component) at the large cell's full plane-wave count, 2 k-points, 16 x 16
    bands of two structures, against the oracle's FP64 sums."""
    from pawpyseed_amd.core import batched as B
    wl, Gs, Cu, Cd, labels, _, _ = case
    nk = len(Gs)
    W = [len(g) for g in Gs]
    ldc = -(-2 * max(W) // 64) * 64
    rng = np.random.default_rng(23)
    spin = []
    for which in ("R", "S"):
        C2 = np.zeros((nk, NB, ldc), np.complex64)
        for k in range(nk):
            up = wl.coefficients(which, rng=rng, nband=NB)[k]
            dn = wl.coefficients(which, rng=rng, nband=NB)[k]
            both = np.concatenate([up, dn], axis=1)
            C2[k, :, : 2 * W[k]] = both / np.linalg.norm(both, axis=1, keepdims=True)
        spin.append(C2)
    CRd, CSd = (torch.from_numpy(c).cuda() for c in spin)
    nw2 = np.array([2 * w for w in W], np.int32)
    out = torch.zeros((nk, NB, NB), dtype=torch.complex128, device="cuda")
    B.overlap_gemm(ctx, nk, 1, False, nw2, CRd, CSd, out, precision)
    got = out.cpu().numpy()
    CR = [spin[0][k, :, : 2 * W[k]] for k in range(nk)]
    CS = [spin[1][k, :, : 2 * W[k]] for k in range(nk)]
    want = np.zeros_like(got)
    for b in range(NB):
        want[:, :, b] = orc.pseudoprojection(CR, CS, b, nk, 1).reshape(NB, nk).T
    # only the down halves: the up-only overlap misses them by O(1)
    up_only = np.zeros_like(got)
    for b in range(NB):
        up_only[:, :, b] = orc.pseudoprojection([c[:, : W[k]] for k, c in enumerate(CR)],
                                                [c[:, : W[k]] for k, c in enumerate(CS)], b, nk, 1).reshape(NB, nk).T
    assert normwise(up_only, want) > 0.1
    err = normwise(got, want)
    print(f"spinor overlap {precision} (2W = {nw2.tolist()}): {err:.2e}", flush=True)
    assert err < tol
