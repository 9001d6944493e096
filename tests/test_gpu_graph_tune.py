"""The setup path as a replayed HIP graph (pp_graph_capture / replay) and
the autotuner (pp_autotune), against eager runs and the FP64 oracle.

Graph: every replay is checked against the eager result of the same
inputs, with the program's own eager work (changing the coefficients,
other setup calls) and synchronisations between replays.  Tolerances:
replay vs eager bit-identical (same kernels, same order); projections vs
oracle 1e-9 (FP64) / 1e-5 (FP32), normwise.
"""

import numpy as np
import pytest

from helpers import device_inputs, normwise
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


def _case(nband=70):
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=nband, kpts=[[0, 0, 0], [0.25, -0.25, 0.1]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("S", nband=nband)
    T = orc.site_tables(wl.lattice, wl.fftg, wl.frac_S, [wl.species], np.zeros(len(wl.frac_S), np.int32))
    want = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, Cs, T)
    return wl, Gs, Cs, want


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_graph_replay_matches_eager(ctx, precision, tol):
    from pawpyseed_amd.core import batched as B
    wl, Gs, Cs, want = _case()
    Cd, Gd, gofs, nw = device_inputs(Gs, Cs)
    labels = np.zeros(len(wl.frac_S), np.int32)
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, labels, [wl.species], precision)
    nk, nb = len(Gs), Cs[0].shape[0]
    eager = torch.zeros((nk, nb, plan.ptot), dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, eager, precision)
    e0 = eager.clone()
    assert normwise(e0.cpu().numpy(), want) < tol
    out = torch.zeros_like(eager)
    torch.cuda.synchronize()
    g = B.SetupGraph(ctx, plan, wl.kpts, gofs, Gd, Cd, out, precision)
    torch.cuda.synchronize()
    assert torch.equal(out, e0)  # the capture's own eager run
    C0 = Cd.clone()
    for rep, scale in enumerate((1.0, 2.0, 1.0, -0.5)):
        # eager work between replays: new coefficients, an eager setup into
        # another buffer, a synchronisation
        Cd.copy_(C0 * scale)
        out.zero_()
        torch.cuda.synchronize()
        B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, eager, precision)
        torch.cuda.synchronize()
        g.replay()
        print("graph replay", precision, rep, "scale", scale, flush=True)
        assert torch.equal(out, eager), f"replay {rep} differs from the eager run"
        assert normwise(out.cpu().numpy(), scale * want) < tol
    # a rebuilt plan invalidates the graph
    plan.rebuild(wl.frac_S)
    with pytest.raises(Exception):
        g.replay()
    g.close()
    plan.close()


def test_autotune_keeps_results(ctx):
    from pawpyseed_amd.core import batched as B
    wl, Gs, Cs, want = _case(nband=64)
    Cd, Gd, gofs, nw = device_inputs(Gs, Cs)
    labels = np.zeros(len(wl.frac_S), np.int32)
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, labels, [wl.species], "fp32")
    out = torch.zeros((len(Gs), 64, plan.ptot), dtype=torch.complex128, device="cuda")
    before = {n: ctx.get_param(n) for n in B.TUNED}
    try:
        chosen, trials = B.autotune(ctx, plan, wl.kpts, gofs, Gd, Cd, out, "fp32", repeats=1)
        torch.cuda.synchronize()
        print("autotune chosen", chosen, "trials", len(trials), flush=True)
        assert len(trials) >= 24
        assert {n for n, _, _ in trials} == set(B.TUNED)
        assert all(ms > 0 for _, _, ms in trials)
        for n, v in chosen.items():
            tried = [val for name, val, _ in trials if name == n]
            assert v in tried or (n == "fft_tile" and v == 0), (n, v, tried)
        assert normwise(out.cpu().numpy(), want) < 1e-5
    finally:
        for n, v in before.items():
            ctx.set_param(n, v)
        plan.close()
