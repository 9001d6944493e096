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
