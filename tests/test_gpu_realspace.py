"""All-electron real-space states on the GPU (density.c:158-327) against the
oracle's restatement of realspace_state (pinned in test_oracle_realspace).

* batched API (pp_rs_plan_create / pp_realspace_batched): states, states
  without the Bloch phase and weighted densities of up to 70 bands (two
  band chunks) on a cubic cell with overlapping spheres (coloured
  launches), a skewed two-species cell with sites near the faces, and a
  small cell whose site boxes wrap onto themselves (atomic adds);
* the reference's entry points through the Python mirror: get_state_realspace
  (also on another grid and with remove_phase), get_state_realspace_density,
  get_realspace_density (all bands and a band subset), the spinor variants,
  the 'realspace' Projector method and the volumetric writers.

Tolerances (normwise, max|d| / max|ref|): FP64 1e-10, FP32 FFT 1e-5.
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


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pawpyseed_amd.core import _lib
    _lib.lib().pp_set_precision(_lib.PP_FP64)
    return True


def _case(name, nband, rng):
    if name == "cubic":
        # rmax 1.3 A: neighbouring spheres overlap (bond 2.35 A)
        wl = synthetic.tiny_config(ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5), rmax=1.3, fftg=(20, 20, 20),
                                   kpts=[[0.25, -0.25, 0.1]], nband=nband)
        G = wl.ensure_gvecs()[0]
        C = wl.coefficients("S", nband=nband)[0]
        return wl.lattice, (20, 20, 20), np.asarray(wl.kpts[0]), G, C, [wl.species], \
            np.zeros(len(wl.frac_S), np.int32), np.asarray(wl.frac_S)
    if name == "skewed":
        lat = np.array([[5.0, 0, 0], [1.0, 5.5, 0], [0.5, 0.7, 6.0]])
        species = [synthetic.make_species([0, 1], [1.5, 3.0], 1.2), synthetic.make_species([2], [2.0], 0.9)]
        coords = np.array([[0.02, 0.5, 0.97], [0.51, 0.99, 0.03], [0.97, 0.01, 0.5]])
        labels = np.array([0, 1, 0], np.int32)
        G = np.unique(rng.integers(-5, 6, size=(60, 3)), axis=0).astype(np.int32)
        C = (0.1 * (rng.normal(size=(nband, len(G))) + 1j * rng.normal(size=(nband, len(G))))).astype(np.complex64)
        return lat, (20, 22, 24), np.array([0.3, 0.1, -0.2]), G, C, species, labels, coords
    # small cell: each site's box wraps onto itself
    lat = np.diag([2.4, 2.6, 2.5])
    species = [synthetic.make_species([0, 1], [1.5, 3.0], 1.4)]
    coords = np.array([[0.1, 0.2, 0.3], [0.6, 0.7, 0.8]])
    labels = np.zeros(2, np.int32)
    G = np.unique(rng.integers(-3, 4, size=(40, 3)), axis=0).astype(np.int32)
    C = (0.1 * (rng.normal(size=(nband, len(G))) + 1j * rng.normal(size=(nband, len(G))))).astype(np.complex64)
    return lat, (12, 12, 12), np.array([0.1, -0.2, 0.25]), G, C, species, labels, coords


def _box_count(lat, fftg, species, labels, coords):
    """points inside each site's sphere over its box (density.c:259-284)"""
    vol = np.linalg.det(lat)
    tot = 0
    for s, e in enumerate(labels):
        rmax = species[e].grid[-1]
        g = [int(np.linalg.norm(np.cross(lat[(a + 1) % 3], lat[(a + 2) % 3])) * rmax / vol * fftg[a]) + 1
             for a in range(3)]
        c = [int(np.round(coords[s][a] * fftg[a])) for a in range(3)]
        ax = [np.arange(c[a] - g[a], c[a] + g[a] + 1) / fftg[a] - coords[s][a] for a in range(3)]
        f = np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1).reshape(-1, 3)
        tot += int((np.linalg.norm(f @ lat, axis=1) < rmax).sum())
    return tot


@pytest.mark.parametrize("name,precision,nband,tol", [
    ("cubic", "fp64", 70, 1e-10), ("cubic", "fp32", 8, 1e-5), ("skewed", "fp64", 5, 1e-10),
    ("wrapping", "fp64", 4, 1e-10)])
def test_realspace_states_batched(ctx, name, precision, nband, tol):
    from pawpyseed_amd.core import batched as B
    rng = np.random.default_rng(7)
    lat, fftg, k, G, C, species, labels, coords = _case(name, nband, rng)
    print("realspace case", name, precision, "bands", nband, "grid", fftg, flush=True)
    W = orc.WaveData(species)
    plan = B.RealspacePlan(ctx, lat, fftg, coords, labels, species)
    assert plan.npts == _box_count(lat, fftg, species, labels, coords)
    assert plan.atomic == (name == "wrapping")
    if name == "cubic":
        assert plan.ncolors >= 2  # the diamond neighbours' spheres overlap
    P = plan.ptot
    proj = rng.normal(size=(nband, P)) + 1j * rng.normal(size=(nband, P))
    ldc = -(-len(G) // 64) * 64
    Ch = np.zeros((nband, ldc), np.complex64)
    Ch[:, :len(G)] = C
    Cd = torch.from_numpy(Ch).cuda()
    Gd = torch.from_numpy(np.ascontiguousarray(G, np.int32)).cuda()
    Pd = torch.from_numpy(proj).cuda()
    out = torch.zeros((nband,) + tuple(fftg), dtype=torch.complex128, device="cuda")
    out_np = torch.zeros_like(out)
    dens = torch.zeros(tuple(fftg), dtype=torch.float64, device="cuda")
    w = rng.uniform(0.1, 1.0, nband)
    B.realspace_states(ctx, plan, k, Gd, Cd, Pd, states_out=out, precision=precision)
    B.realspace_states(ctx, plan, k, Gd, Cd, Pd, states_out=out_np, density_out=dens, weights=w,
                       remove_phase=True, precision=precision)
    got, got_np, got_d = out.cpu().numpy(), out_np.cpu().numpy(), dens.cpu().numpy()
    want_d = np.zeros(fftg)
    for b in range(nband):
        want = orc.realspace_state(lat, fftg, k, G, C[b], W, labels, coords, proj[b])
        want_np = orc.realspace_state(lat, fftg, k, G, C[b], W, labels, coords, proj[b], remove_phase=True)
        e, e2 = normwise(got[b], want), normwise(got_np[b], want_np)
        assert e < tol and e2 < tol, (b, e, e2)
        want_d += w[b] * np.abs(want) ** 2
    assert normwise(got_d, want_d) < tol
    plan.close()


def _wf_case(nband=6, ncl=False):
    from pawpyseed_amd.core import CoreRegion, NCLWavefunction, Pseudopotential, Structure, Wavefunction
    wl = synthetic.tiny_config(ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5), rmax=1.1, fftg=(20, 20, 20), nband=nband,
                               kpts=[[0, 0, 0], [0.25, -0.25, 0.1]])
    Gs = wl.ensure_gvecs()
    struct = Structure(wl.lattice, ["C"] * len(wl.frac_S), wl.frac_S)
    cr = CoreRegion({"C": Pseudopotential.from_species(wl.species)})
    occ = np.tile(np.linspace(1.0, 0.0, nband), (len(Gs), 1))
    occ[:, -2:] = 0.0
    if ncl:
        Cu, Cd = wl.coefficients("S", nband=nband), wl.coefficients("R", nband=nband)
        C2 = [np.concatenate([u, d], axis=1) for u, d in zip(Cu, Cd)]
        wf = NCLWavefunction.from_arrays(struct, cr, wl.fftg, wl.kpts, wl.kws, Gs, C2, wl.encut, occs=occ,
                                         setup_projectors=True)
        return wl, wf, (Cu, Cd), occ
    C = wl.coefficients("S", nband=nband)
    wf = Wavefunction.from_arrays(struct, cr, wl.fftg, wl.kpts, wl.kws, Gs, C, wl.encut, occs=occ,
                                  setup_projectors=True)
    return wl, wf, C, occ


def _oracle_state(wl, wf, C, b, k, fftg, proj=None, remove_phase=False):
    W = orc.WaveData([wl.species])
    p = wf.get_projections(b, k) if proj is None else proj
    return orc.realspace_state(wl.lattice, fftg, wl.kpts[k], wl.ensure_gvecs()[k], C[k][b], W,
                               np.zeros(len(wl.frac_S), np.int32), wl.frac_S, p, remove_phase=remove_phase)


def test_wavefunction_realspace_states_and_densities(gpu):
    wl, wf, C, occ = _wf_case()
    for (b, k) in [(0, 0), (3, 1)]:
        got = wf.get_state_realspace(b, k, 0)
        assert got.shape == (20, 20, 20)
        assert normwise(got, _oracle_state(wl, wf, C, b, k, (20, 20, 20))) < 1e-10
    got = wf.get_state_realspace(2, 1, 0, dim=(24, 24, 24), remove_phase=True)
    want = _oracle_state(wl, wf, C, 2, 1, (24, 24, 24), remove_phase=True)
    assert normwise(got, want) < 1e-10
    # densities on the doubled grid of dim // 2
    d1 = wf.get_state_realspace_density(1, 1, 0, dim=(40, 40, 40))
    assert d1.shape == (40, 40, 40)
    want1 = np.abs(_oracle_state(wl, wf, C, 1, 1, (40, 40, 40))) ** 2
    assert normwise(d1, want1) < 1e-10
    dens = wf.get_realspace_density()
    want = np.zeros((40, 40, 40))
    for k in range(2):
        for b in range(wf.nband):
            if occ[k, b] > 0:
                want += wl.kws[k] * occ[k, b] * 2.0 * np.abs(_oracle_state(wl, wf, C, b, k, (40, 40, 40))) ** 2
    assert normwise(dens, want) < 1e-10
    sub = wf.get_realspace_density(bands=[0, 2])
    want = sum(wl.kws[k] * np.abs(_oracle_state(wl, wf, C, b, k, (40, 40, 40))) ** 2 for k in range(2) for b in (0, 2))
    assert normwise(sub, want) < 1e-10
    with pytest.raises(ValueError):
        wf.get_state_realspace(wf.nband, 0, 0)
    wf.free()


def test_projector_realspace_method(gpu):
    from pawpyseed_amd.core import CoreRegion, Projector, Pseudopotential, Structure, Wavefunction
    wl, wf, C, occ = _wf_case(nband=4)
    CR = wl.coefficients("R", nband=4)
    Gs = wl.ensure_gvecs()
    struct = Structure(wl.lattice, ["C"] * len(wl.frac_R), wl.frac_R)
    cr = CoreRegion({"C": Pseudopotential.from_species(wl.species)})
    basis = Wavefunction.from_arrays(struct, cr, wl.fftg, wl.kpts, wl.kws, Gs, CR, wl.encut, setup_projectors=True)
    pr = Projector(wf, basis, method="realspace")
    res = pr.single_band_projection(2)
    fine = (40, 40, 40)
    vol = abs(np.linalg.det(wl.lattice))
    W = orc.WaveData([wl.species])
    for k in range(2):
        s = _oracle_state(wl, wf, C, 2, k, fine)
        for b in range(4):
            r = orc.realspace_state(wl.lattice, fine, wl.kpts[k], Gs[k], CR[k][b], W,
                                    np.zeros(len(wl.frac_R), np.int32), wl.frac_R, basis.get_projections(b, k))
            want = vol / np.prod(fine) * np.vdot(r, s)
            assert abs(res[b * 2 + k] - want) < 1e-10 * max(1.0, abs(want)), (b, k)
    basis.free()
    wf.free()


def test_noncollinear_realspace(gpu):
    wl, wf, (Cu, Cd), occ = _wf_case(nband=4, ncl=True)
    up, dn = wf.get_state_realspace(1, 1, 0)
    p = wf.get_projections(1, 1)
    assert normwise(up, _oracle_state(wl, wf, Cu, 1, 1, (20, 20, 20), proj=p[0])) < 1e-10
    assert normwise(dn, _oracle_state(wl, wf, Cd, 1, 1, (20, 20, 20), proj=p[1])) < 1e-10
    dens = wf.get_realspace_density()
    want = np.zeros((20, 20, 20))
    for k in range(2):
        for b in range(wf.nband):
            if occ[k, b] > 0:
                pk = wf.get_projections(b, k)
                for comp, Cc in ((0, Cu), (1, Cd)):
                    want += wl.kws[k] * occ[k, b] * np.abs(_oracle_state(wl, wf, Cc, b, k, (20, 20, 20),
                                                                         proj=pk[comp])) ** 2
    assert normwise(dens, want) < 1e-10
    wf.free()


def test_volumetric_writers(gpu, tmp_path):
    import os
    wl, wf, C, occ = _wf_case(nband=3)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        res = wf.write_state_realspace(1, 0, 0, fileprefix="t")
        dens = wf.write_density_realspace("D.vasp")
    finally:
        os.chdir(cwd)
    for name, arr, dim in (("tB1K0S0_REAL.vasp", res.real, (20, 20, 20)), ("tB1K0S0_IMAG.vasp", res.imag,
                                                                          (20, 20, 20)),
                           ("D.vasp", dens, (40, 40, 40))):
        lines = open(tmp_path / name).read().split("\n")
        i = lines.index("Direct") + len(wl.frac_S) + 2
        assert [int(x) for x in lines[i].split()] == list(dim)
        vals = np.array(" ".join(lines[i + 1:]).split(), float)
        # z slowest, x fastest (density.c:461-477)
        np.testing.assert_allclose(vals, np.asarray(arr).transpose(2, 1, 0).ravel(), rtol=1e-5, atol=1e-12)
    wf.free()


@pytest.mark.parametrize("kernel,unroll", [(0, 1), (0, 4), (1, 4), (2, 1), (2, 2)])
def test_augmentation_variants_agree(ctx, kernel, unroll):
    """Every selectable augmentation kernel (rs_kernel 0: coefficients in
    LDS, 1: in registers, 2: tiled; rs_unroll points in flight) gives the
    oracle's states, FP64 and FP32 chunks, on the overlapping cubic case."""
    from pawpyseed_amd.core import batched as B
    rng = np.random.default_rng(17)
    nband = 6
    lat, fftg, k, G, C, species, labels, coords = _case("cubic", nband, rng)
    W = orc.WaveData(species)
    plan = B.RealspacePlan(ctx, lat, fftg, coords, labels, species)
    proj = rng.normal(size=(nband, plan.ptot)) + 1j * rng.normal(size=(nband, plan.ptot))
    ldc = -(-len(G) // 64) * 64
    Ch = np.zeros((nband, ldc), np.complex64)
    Ch[:, :len(G)] = C
    Cd, Gd = torch.from_numpy(Ch).cuda(), torch.from_numpy(np.ascontiguousarray(G, np.int32)).cuda()
    Pd = torch.from_numpy(proj).cuda()
    before = (ctx.get_param("rs_kernel"), ctx.get_param("rs_unroll"))
    try:
        ctx.set_param("rs_kernel", kernel)
        ctx.set_param("rs_unroll", unroll)
        for precision, tol in (("fp64", 1e-10), ("fp32", 1e-5)):
            out = torch.zeros((nband,) + tuple(fftg), dtype=torch.complex128, device="cuda")
            B.realspace_states(ctx, plan, k, Gd, Cd, Pd, states_out=out, precision=precision)
            got = out.cpu().numpy()
            for b in range(nband):
                want = orc.realspace_state(lat, fftg, k, G, C[b], W, labels, coords, proj[b])
                assert normwise(got[b], want) < tol, (precision, b)
    finally:
        ctx.set_param("rs_kernel", before[0])
        ctx.set_param("rs_unroll", before[1])
    plan.close()


def test_wrapping_boxes_fp32_atomics(ctx):
    """Atomic adds of the FP32 chunk where site boxes wrap onto themselves."""
    from pawpyseed_amd.core import batched as B
    rng = np.random.default_rng(19)
    nband = 3
    lat, fftg, k, G, C, species, labels, coords = _case("wrapping", nband, rng)
    W = orc.WaveData(species)
    plan = B.RealspacePlan(ctx, lat, fftg, coords, labels, species)
    assert plan.atomic
    proj = rng.normal(size=(nband, plan.ptot)) + 1j * rng.normal(size=(nband, plan.ptot))
    ldc = -(-len(G) // 64) * 64
    Ch = np.zeros((nband, ldc), np.complex64)
    Ch[:, :len(G)] = C
    out = torch.zeros((nband,) + tuple(fftg), dtype=torch.complex128, device="cuda")
    B.realspace_states(ctx, plan, k, torch.from_numpy(np.ascontiguousarray(G, np.int32)).cuda(),
                       torch.from_numpy(Ch).cuda(), torch.from_numpy(proj).cuda(), states_out=out, precision="fp32")
    got = out.cpu().numpy()
    for b in range(nband):
        want = orc.realspace_state(lat, fftg, k, G, C[b], W, labels, coords, proj[b])
        assert normwise(got[b], want) < 1e-5
    plan.close()


def test_full_size_216_atoms_doubled_grid(ctx):
    """The all-electron density workload's size: 216-atom cell (215 S
    sites, 1.6 A spheres) on the 216^3 grid.  Parseval for the pseudo part
    (projections 0): sum_r |psi(r)|^2 V/N^3 = sum_G |C_G|^2 to 1e-12, and
    one full state (band 0, k 1) against the oracle to 1e-10."""
    from pawpyseed_amd.core import batched as B
    wl = synthetic.config1(nband=2)
    Gs = wl.ensure_gvecs()
    k = 1
    G = np.ascontiguousarray(Gs[k], np.int32)
    C = wl.coefficients("S", nband=2)[k]
    fine = tuple(int(2 * n) for n in wl.fftg)
    labels = np.zeros(len(wl.frac_S), np.int32)
    plan = B.RealspacePlan(ctx, wl.lattice, fine, wl.frac_S, labels, [wl.species])
    print("full-size realspace: grid", fine, "points", plan.npts, "colours", plan.ncolors, flush=True)
    ldc = -(-len(G) // 64) * 64
    Ch = np.zeros((2, ldc), np.complex64)
    Ch[:, :len(G)] = C
    Cd, Gd = torch.from_numpy(Ch).cuda(), torch.from_numpy(G).cuda()
    rng = np.random.default_rng(23)
    proj = 0.05 * (rng.normal(size=(2, plan.ptot)) + 1j * rng.normal(size=(2, plan.ptot)))
    zero = torch.zeros((2, plan.ptot), dtype=torch.complex128, device="cuda")
    out = torch.zeros((2,) + fine, dtype=torch.complex128, device="cuda")
    B.realspace_states(ctx, plan, wl.kpts[k], Gd, Cd, zero, states_out=out, precision="fp64")
    vol = abs(np.linalg.det(wl.lattice))
    for b in range(2):
        norm = float((out[b].abs() ** 2).sum().item()) * vol / np.prod(fine)
        want = float(np.sum(np.abs(C[b].astype(np.complex128)) ** 2))
        assert abs(norm - want) < 1e-12 * want, (b, norm, want)
    B.realspace_states(ctx, plan, wl.kpts[k], Gd, Cd, torch.from_numpy(proj).cuda(), states_out=out,
                       precision="fp64")
    got = out[0].cpu().numpy()
    W = orc.WaveData([wl.species])
    want = orc.realspace_state(wl.lattice, fine, wl.kpts[k], G, C[0], W, labels, wl.frac_S, proj[0])
    err = normwise(got, want)
    print("full-size realspace state error", err, flush=True)
    assert err < 1e-10
    plan.close()
