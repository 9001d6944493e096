"""The FP32 x pass fused with the sphere sums (param "fused" 1), against
the FP64 oracle and against the separate x-pass and sphere-sum kernels
(the default).

Tolerances: FP32 projections <= 1e-5 relative (normwise, max|d| /
max|ref|) against the oracle; fused and separate FP32 kernels agree to
<= 2e-6 (both round to FP32, in different summation orders).
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
    c.set_param("fused", 0)
    c.close()


def _run(ctx, wl, frac, species, labels, nb, fused):
    from pawpyseed_amd.core import batched as B
    Gs = wl.ensure_gvecs()
    Cs = wl.coefficients("S", nband=nb)
    Cd, Gd, gofs, nw = device_inputs(Gs, Cs)
    plan = B.SpherePlan(ctx, wl.lattice, wl.fftg, frac, labels, species, "fp32")
    info = plan.fused_info()
    out = torch.zeros((len(Gs), nb, plan.ptot), dtype=torch.complex128, device="cuda")
    ctx.set_param("fused", 1 if fused else 0)
    try:
        B.setup_projections(ctx, plan, wl.kpts, gofs, Gd, Cd, out, "fp32")
        torch.cuda.synchronize()
    finally:
        ctx.set_param("fused", 0)
    got = out.cpu().numpy()
    plan.close()
    return got, info, Gs, Cs


def _check(ctx, name, wl, frac, species, labels, nb, expect_usable=True):
    print("fused case", name, "fftg", list(wl.fftg), "sites", len(frac), "bands", nb, flush=True)
    fz, info, Gs, Cs = _run(ctx, wl, frac, species, labels, nb, True)
    usable, nitems, nrows, nseg, flag = info
    print("  work lists: usable", usable, "items", nitems, "rows", nrows, "segments", nseg, "flag", flag,
          flush=True)
    assert usable == expect_usable
    if name == "dense":
        assert nitems == 153
    assert flag == 0
    sep, _, _, _ = _run(ctx, wl, frac, species, labels, nb, False)
    T = orc.site_tables(wl.lattice, wl.fftg, frac, species, labels)
    want = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, Gs, Cs, T)
    e_or = normwise(fz, want)
    e_sep = normwise(fz, sep)
    print(f"  vs oracle {e_or:.2e}, vs separate kernels {e_sep:.2e}", flush=True)
    assert e_or < 1e-5
    assert e_sep < 2e-6


def test_fused_tiny_two_k(ctx):
    """8 sites, 18 projectors each, two k-points, one 16-band chunk."""
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=16, kpts=[[0, 0, 0], [0.25, -0.25, 0.1]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    _check(ctx, "tiny", wl, wl.frac_S, [wl.species], np.zeros(len(wl.frac_S), np.int32), 16)


def test_fused_ragged_chunks_edge_blocks(ctx):
    """70 bands (a 64-band chunk and a 6-band one), y and z lengths that are
    not multiples of the 4-line block, x length 30."""
    wl = synthetic.tiny_config(encut=100.0, fftg=(30, 18, 22), nband=70, kpts=[[0.1, 0.2, 0.3]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    _check(ctx, "ragged", wl, wl.frac_S, [wl.species], np.zeros(len(wl.frac_S), np.int32), 70)


def test_fused_two_species(ctx):
    """9 and 18 projectors per site, different radii (not every site has
    the full projector count)."""
    wl = synthetic.tiny_config(encut=150.0, fftg=(24, 24, 24), nband=32, kpts=[[0, 0, 0], [0.25, -0.25, 0.1]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    sp2 = synthetic.make_species([0, 1, 1], [1.2, 1.4, 2.8], 0.95)
    labels = (np.arange(len(wl.frac_S)) % 2).astype(np.int32)
    _check(ctx, "two species", wl, wl.frac_S, [wl.species, sp2], labels, 32)


def test_fused_dense_spheres_split_items(ctx):
    """64 sites with overlapping spheres (radius 2.2 A at 1.54 A bonds):
    many blocks hold more than 12 sites, so their lines are transformed by
    two work items (153 items for 100 blocks)."""
    wl = synthetic.tiny_config(ncell=2, encut=60.0, fftg=(40, 40, 40), nband=16, kpts=[[0.0, 0.5, 0.25]],
                               ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5), rmax=2.2)
    _check(ctx, "dense", wl, wl.frac_S, [wl.species], np.zeros(len(wl.frac_S), np.int32), 16)


def test_fused_cfg0(ctx):
    """The 64-site, 60^3 cell of the 400 eV workload, 64 bands."""
    wl = synthetic.config0()
    _check(ctx, "cfg0", wl, wl.frac_S, [wl.species], np.zeros(len(wl.frac_S), np.int32), 64)


def test_fused_falls_back_when_boxes_wrap(ctx):
    """Spheres wider than the cell along y and z: no fused work lists, the
    separate kernels run."""
    wl = synthetic.tiny_config(encut=40.0, fftg=(12, 12, 12), nband=16, kpts=[[0, 0, 0]], ls=(0, 1),
                               alphas=(1.5, 3.0), rmax=2.6)
    _check(ctx, "wrapping boxes", wl, wl.frac_S, [wl.species], np.zeros(len(wl.frac_S), np.int32), 16,
           expect_usable=False)
