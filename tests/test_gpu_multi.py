"""The multi-structure setup (pp_setup_projections_multi): several
structures sharing one plane-wave basis, k-points with uneven plane-wave
counts and uneven band counts per (structure, k), all as one work list of
64-band chunks, against the FP64 oracle (projector.c:560-602 per structure)
and against the single-structure path; its HIP graph against eager runs.

Tolerances: FP64 1e-9 and FP32 1e-5 relative (normwise per sampled band)
against the oracle; the work-item batch size does not change the result
(bit-identical); FP64 against the single-structure path 1e-12.
"""

import numpy as np
import pytest

from helpers import normwise
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


def _basis(lattice, kpts, encuts):
    Gs = [synthetic.gvectors(lattice, k, e) for k, e in zip(kpts, encuts)]
    nw = np.array([len(g) for g in Gs], np.int32)
    gofs = np.concatenate([[0], np.cumsum(nw)]).astype(np.int32)
    Gd = torch.from_numpy(np.ascontiguousarray(np.concatenate(Gs), np.int32)).cuda()
    return Gs, nw, gofs, Gd


def _coeffs(wl, Gs, nbands_s, seed):
    """complex64 [nk, max nband, ldc] on the GPU and the per-k views."""
    nw = [len(g) for g in Gs]
    ldc = int(-(-max(nw) // 64) * 64)
    C = synthetic.device_coefficients(torch, wl, Gs, seed, int(max(nbands_s)), ldc, nbands=nbands_s)
    return C, [C[k, : nbands_s[k]] for k in range(len(Gs))]


def _small():
    """Two structures of a 16-site cell on a 24^3 grid, three k-points
    with cutoffs 150, 120 and 135 eV and band counts from 3 to 70."""
    wl = synthetic.tiny_config(ncell=1, encut=150.0, fftg=(24, 24, 24), nband=8,
                               kpts=[[0, 0, 0], [0.25, -0.25, 0.1], [0.5, 0.1, 0.3]], ls=(0, 1, 2),
                               alphas=(1.5, 3.0, 1.5))
    encuts = np.array([150.0, 120.0, 135.0])
    nbands = np.array([[70, 3, 64], [5, 66, 17]])
    return wl, encuts, nbands


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_multi_small_against_oracle_and_single(ctx, precision, tol):
    from pawpyseed_amd.core import batched as B
    wl, encuts, nbands = _small()
    Gs, nw, gofs, Gd = _basis(wl.lattice, wl.kpts, encuts)
    wl.encut = float(encuts.max())
    fracs = [wl.frac_R, wl.frac_S]
    labels = np.zeros(len(wl.frac_R), np.int32)
    plans = [B.SpherePlan(ctx, wl.lattice, wl.fftg, f, labels, [wl.species], precision) for f in fracs]
    Cfull, C, P = [], [], []
    for s in range(2):
        Cs, views = _coeffs(wl, Gs, nbands[s], 11 + s)
        Cfull.append(Cs)
        C.append(views)
        P.append([torch.zeros((int(nbands[s][k]), plans[s].ptot), dtype=torch.complex128, device="cuda")
                  for k in range(len(Gs))])
    B.setup_projections_multi(ctx, plans, wl.kpts, gofs, Gd, C, P, precision)
    got = [[p.cpu().numpy() for p in Ps] for Ps in P]
    for s in range(2):
        T = orc.site_tables(wl.lattice, wl.fftg, fracs[s], [wl.species], labels)
        for k in range(len(Gs)):
            Ch = C[s][k][:, : nw[k]].cpu().numpy()
            want = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts[k:k + 1], [Gs[k]], [Ch], T)[0]
            err = max(normwise(got[s][k][b], want[b]) for b in range(nbands[s][k]))
            print(f"multi {precision} structure {s} k {k} ({nbands[s][k]} bands, W {nw[k]}): {err:.2e}", flush=True)
            assert err < tol
            # the single-structure path on the same (structure, k)
            one = torch.zeros((1, int(nbands[s][k]), plans[s].ptot), dtype=torch.complex128, device="cuda")
            go = np.array([0, nw[k]], np.int32)
            B.setup_projections(ctx, plans[s], wl.kpts[k:k + 1], go, Gd[gofs[k]:gofs[k + 1]],
                                C[s][k].unsqueeze(0), one, precision)
            assert normwise(got[s][k], one[0].cpu().numpy()) < (1e-12 if precision == "fp64" else 1e-6)
    # the batch size only changes how items share launches
    before = ctx.get_param("multi_items")
    try:
        for q in (1, 3, 64):
            ctx.set_param("multi_items", q)
            for Ps in P:
                for p in Ps:
                    p.zero_()
            B.setup_projections_multi(ctx, plans, wl.kpts, gofs, Gd, C, P, precision)
            for s in range(2):
                for k in range(len(Gs)):
                    assert np.array_equal(P[s][k].cpu().numpy(), got[s][k]), (q, s, k)
    finally:
        ctx.set_param("multi_items", before)
    for p in plans:
        p.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_multi_216_atoms_uneven(ctx, precision, tol):
    """The batched uneven workload at a reduced k-count: the 216-site cell,
    three perturbed structures (215, 216, 216 sites), 4 random wedge
    k-points with plane-wave counts spread +-10%, 384-640 bands per
    (structure, k).  Sampled bands of every (structure, k) against the
    oracle; the captured graph replays the eager result."""
    from pawpyseed_amd.core import batched as B
    wl = synthetic.config3(nk=4)
    Gs, nw, gofs, Gd = _basis(wl.lattice, wl.kpts, wl.encuts)
    print(f"cfg3 (4 k): W {nw.tolist()}, bands {wl.nbands[1:].tolist()}", flush=True)
    plans, C, P, keep = [], [], [], []
    for s in range(1, 4):
        labels = np.zeros(len(wl.frac[s]), np.int32)
        plans.append(B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac[s], labels, [wl.species], precision))
        Cs, views = _coeffs(wl, Gs, wl.nbands[s], 300 + s)
        keep.append(Cs)
        C.append(views)
        P.append([torch.zeros((int(wl.nbands[s][k]), plans[-1].ptot), dtype=torch.complex128, device="cuda")
                  for k in range(len(Gs))])
    B.setup_projections_multi(ctx, plans, wl.kpts, gofs, Gd, C, P, precision)
    torch.cuda.synchronize()
    eager = [[p.clone() for p in Ps] for Ps in P]
    for i, s in enumerate(range(1, 4)):
        labels = np.zeros(len(wl.frac[s]), np.int32)
        T = orc.site_tables(wl.lattice, wl.fftg, wl.frac[s], [wl.species], labels)
        for k in range(len(Gs)):
            nb = int(wl.nbands[s][k])
            bands = [0, nb // 2, nb - 1]
            Ch = C[i][k][bands, : nw[k]].cpu().numpy()
            want = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts[k:k + 1], [Gs[k]], [Ch], T, nthreads=0)[0]
            got = P[i][k][bands].cpu().numpy()
            err = max(normwise(got[j], want[j]) for j in range(len(bands)))
            print(f"cfg3 {precision} structure {s} k {k}: {err:.2e}", flush=True)
            assert err < tol
    g = B.MultiGraph(ctx, plans, wl.kpts, gofs, Gd, C, P, precision)
    for rep in range(2):
        for Ps in P:
            for p in Ps:
                p.zero_()
        torch.cuda.synchronize()
        g.replay()
        for i in range(3):
            for k in range(len(Gs)):
                assert torch.equal(P[i][k], eager[i][k]), (rep, i, k)
        print(f"cfg3 graph replay {rep} matches the eager run", flush=True)
    g.close()
    for p in plans:
        p.close()
