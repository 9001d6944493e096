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


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("fp32", 1e-5)])
def test_graph_replay_after_other_inputs(ctx, precision, tol):
    """A replay after eager calls on the same context with other k-points,
    another plane-wave set and another plan still gives the captured
    inputs' projections: the graph keeps its own grid indices, phases,
    bounds and line lists."""
    from pawpyseed_amd.core import batched as B
    wl, Gs, Cs, want = _case(nband=40)
    Cd, Gd, gofs, nw = device_inputs(Gs, Cs)
    labels = np.zeros(len(wl.frac_S), np.int32)
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, wl.frac_S, labels, [wl.species], precision)
    out = torch.zeros((len(Gs), 40, plan.ptot), dtype=torch.complex128, device="cuda")
    g = B.SetupGraph(ctx, plan, wl.kpts, gofs, Gd, Cd, out, precision)
    torch.cuda.synchronize()
    first = out.clone()
    assert normwise(first.cpu().numpy(), want) < tol
    # other inputs on the same context: a lower cutoff (fewer plane waves),
    # other k-points, the R sites, fewer bands
    wl2 = synthetic.tiny_config(encut=110.0, fftg=(24, 24, 24), nband=24, kpts=[[0.1, 0.2, -0.3], [0.5, 0, 0]],
                                ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    Gs2 = wl2.ensure_gvecs()
    Cs2 = wl2.coefficients("R", nband=24)
    Cd2, Gd2, gofs2, _ = device_inputs(Gs2, Cs2)
    plan2 = B.SpherePlan(ctx, wl2.lattice, wl2.fftg, wl2.frac_R, labels, [wl2.species], precision)
    other = torch.zeros((len(Gs2), 24, plan2.ptot), dtype=torch.complex128, device="cuda")
    B.setup_projections(ctx, plan2, wl2.kpts, gofs2, Gd2, Cd2, other, precision)
    T2 = orc.site_tables(wl2.lattice, wl2.fftg, wl2.frac_R, [wl2.species], labels)
    want2 = orc.setup_projections(wl2.lattice, wl2.fftg, wl2.kpts, Gs2, Cs2, T2)
    assert normwise(other.cpu().numpy(), want2) < tol
    out.zero_()
    torch.cuda.synchronize()
    g.replay()
    print("replay after other inputs", precision, flush=True)
    assert torch.equal(out, first)
    g.close()
    plan.close()
    plan2.close()
