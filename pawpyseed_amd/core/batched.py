"""Python wrapper of the batched device-resident entry points.

Device memory comes from torch tensors (plumbing only: every computation
runs in the HIP library); pointers and sizes go through the C-ABI.
"""

from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import (PRECISIONS, as_f64, as_i32, c_double_p, c_int_p, check, lib, ptr)


def _prec(p):
    return PRECISIONS[p] if isinstance(p, str) else int(p)


def species_arrays(species_list):
    """Flatten species (objects with ls, realprojs, rmax) into the
    pp_sphere_create arrays."""
    nchan = as_i32([len(sp.ls) for sp in species_list])
    ls = as_i32([l for sp in species_list for l in sp.ls])
    rmax = as_f64([sp.rmax for sp in species_list])
    ngrid = as_i32([len(sp.realprojs[0]) for sp in species_list])
    proj = as_f64(np.concatenate([np.asarray(p, np.float64) for sp in species_list for p in sp.realprojs]))
    return nchan, ls, rmax, ngrid, proj


class Context:
    """A pp_ctx: one HIP stream on one device plus workspaces."""

    def __init__(self, device: int = 0, stream=None):
        """stream: optional torch.cuda.Stream to launch on (its handle must
        be non-null); by default the context owns a stream and the wrapper
        functions below synchronize it before returning (sync=True)."""
        h = ct.c_void_p()
        check(lib().pp_ctx_create(ct.byref(h), int(device)), "pp_ctx_create")
        self.h = h
        self.device = device
        self.stream = None
        if stream is not None:
            if not stream.cuda_stream:
                raise ValueError("pass a non-default torch.cuda.Stream")
            self.set_stream(stream.cuda_stream)
            self.stream = stream

    def set_stream(self, stream_handle):
        """Launch on an external HIP stream (e.g. torch.cuda.current_stream().cuda_stream)."""
        check(lib().pp_ctx_set_stream(self.h, ct.c_void_p(stream_handle) if stream_handle else None),
              "pp_ctx_set_stream")

    def synchronize(self):
        check(lib().pp_ctx_synchronize(self.h), "pp_ctx_synchronize")

    def set_timing(self, on: bool):
        check(lib().pp_ctx_set_timing(self.h, int(bool(on))), "pp_ctx_set_timing")

    def timing_reset(self):
        check(lib().pp_ctx_timing_reset(self.h), "pp_ctx_timing_reset")

    def timing(self):
        n = 32
        names = (ct.c_char_p * n)()
        ms = np.zeros(n)
        cnt = np.zeros(n, np.int64)
        k = lib().pp_ctx_timing_read(self.h, n, names, ptr(ms, c_double_p), cnt.ctypes.data_as(ct.POINTER(ct.c_long)))
        return {names[i].decode(): (float(ms[i]), int(cnt[i])) for i in range(k) if cnt[i] > 0}

    def set_tuning(self, fft_bands=0, fft_tile=0, fft_prune=0, sphere_parts=0, gemm_splitk=0):
        check(lib().pp_ctx_set_tuning(self.h, fft_bands, fft_tile, fft_prune, sphere_parts, gemm_splitk),
              "pp_ctx_set_tuning")

    def set_param(self, name: str, value: int):
        check(lib().pp_ctx_set_param(self.h, name.encode(), int(value)), "pp_ctx_set_param")

    def get_param(self, name: str) -> int:
        v = ct.c_int()
        check(lib().pp_ctx_get_param(self.h, name.encode(), ct.byref(v)), "pp_ctx_get_param")
        return v.value

    def close(self):
        if getattr(self, "h", None):
            lib().pp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SpherePlan:
    """pp_sphere: sphere tables of one structure on the device."""

    def __init__(self, ctx: Context, lattice, fftg, frac_coords, labels, species_list, precision="fp64"):
        self.ctx = ctx
        self._keep = []
        lat = as_f64(lattice).ravel()
        fg = as_i32(fftg)
        fr = as_f64(frac_coords).reshape(-1, 3)
        lab = as_i32(labels)
        nchan, ls, rmax, ngrid, proj = species_arrays(species_list)
        h = ct.c_void_p()
        check(lib().pp_sphere_create(ctx.h, ct.byref(h), ptr(lat, c_double_p), ptr(fg, c_int_p), len(fr),
                                     ptr(lab, c_int_p), ptr(fr, c_double_p), len(species_list),
                                     ptr(nchan, c_int_p), ptr(ls, c_int_p), ptr(rmax, c_double_p),
                                     ptr(ngrid, c_int_p), ptr(proj, c_double_p), _prec(precision)),
              "pp_sphere_create")
        self.h = h
        self.nsites = len(fr)
        self.precision = _prec(precision)
        self._info()

    def _info(self):
        tp = ct.c_int()
        npt = ct.c_long()
        off = np.zeros(self.nsites + 1, np.int32)
        check(lib().pp_sphere_info(self.h, ct.byref(tp), ct.byref(npt), ptr(off, c_int_p)), "pp_sphere_info")
        self.ptot = tp.value
        self.npts = npt.value
        self.proj_off = off

    def coverage(self) -> int:
        """Grid points held by at least one projector sphere."""
        c = ct.c_long()
        check(lib().pp_sphere_coverage(self.h, ct.byref(c)), "pp_sphere_coverage")
        return c.value

    def fused_info(self):
        """(usable, work items, partial rows, segments, flag) of the plan's
        fused x-pass work lists; flag 0 means every segment found its slot
        once."""
        u, it, nr, fl = ct.c_int(), ct.c_int(), ct.c_int(), ct.c_int()
        ns = ct.c_long()
        check(lib().pp_sphere_fused_info(self.h, ct.byref(u), ct.byref(it), ct.byref(nr), ct.byref(ns),
                                         ct.byref(fl)), "pp_sphere_fused_info")
        return bool(u.value), it.value, nr.value, ns.value, fl.value

    def rebuild(self, frac_coords):
        fr = as_f64(frac_coords).reshape(-1, 3)
        check(lib().pp_sphere_rebuild(self.h, ptr(fr, c_double_p)), "pp_sphere_rebuild")
        self._info()

    def close(self):
        if getattr(self, "h", None):
            lib().pp_sphere_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dptr(t):
    return ct.c_void_p(t.data_ptr())


def _done(ctx, sync):
    if sync:
        ctx.synchronize()


def setup_projections(ctx: Context, plan: SpherePlan, kfrac, gofs, G_dev, C_dev, proj_out, precision="fp64", sync=True):
    """<p|psi> of every (k, band): C_dev complex64 [nk, nband, ldc] (cuda),
    G_dev int32 [sum W, 3] (cuda), proj_out complex128 [nk, nband, ptot]."""
    nk, nband, ldc = C_dev.shape
    kf = as_f64(kfrac).reshape(-1, 3)
    go = as_i32(gofs)
    assert proj_out.shape == (nk, nband, plan.ptot)
    check(lib().pp_setup_projections_batched(ctx.h, plan.h, nk, ptr(kf, c_double_p), ptr(go, c_int_p),
                                             _dptr(G_dev), _dptr(C_dev), ldc, nband, _dptr(proj_out),
                                             _prec(precision)), "pp_setup_projections_batched")
    _done(ctx, sync)


def setup_projections_spinor(ctx: Context, plan: SpherePlan, kfrac, gofs, G_dev, C2_dev, proj_up, proj_dn,
                             precision="fp64", sync=True):
    """Spinor bands (onto_projector_ncl, projector.c:359-392): C2_dev
    complex64 [nk, nband, ldc] holds the up component of band (k, b) in
    [0, W_k) of its row and the down one in [W_k, 2 W_k); both go through
    the batched setup as band sets with row stride ldc, into proj_up /
    proj_dn complex128 [nk, nband, ptot]."""
    nk, nband, ldc = C2_dev.shape
    kf = as_f64(kfrac).reshape(-1, 3)
    go = as_i32(gofs)
    assert proj_up.shape == proj_dn.shape == (nk, nband, plan.ptot)
    assert str(C2_dev.dtype) == "torch.complex64" and str(proj_up.dtype) == str(proj_dn.dtype) == "torch.complex128"
    for k in range(nk):
        W = int(go[k + 1] - go[k])
        assert 2 * W <= ldc
        gk = np.array([0, W], np.int32)
        kk = np.ascontiguousarray(kf[k])
        for comp, out in ((0, proj_up), (1, proj_dn)):
            cptr = C2_dev.data_ptr() + ((k * nband) * ldc + comp * W) * 8
            gptr = G_dev.data_ptr() + 3 * int(go[k]) * 4
            pptr = out.data_ptr() + k * nband * plan.ptot * 16
            check(lib().pp_setup_projections_batched(ctx.h, plan.h, 1, ptr(kk, c_double_p), ptr(gk, c_int_p),
                                                     ct.c_void_p(gptr), ct.c_void_p(cptr), ldc, nband,
                                                     ct.c_void_p(pptr), _prec(precision)),
                  "pp_setup_projections_batched")
    _done(ctx, sync)


class SetupGraph:
    """pp_graph: the setup path of fixed buffers captured once, replayed on
    the context's stream (recomputes proj_out from the current C_dev)."""

    def __init__(self, ctx: Context, plan: SpherePlan, kfrac, gofs, G_dev, C_dev, proj_out, precision="fp64"):
        nk, nband, ldc = C_dev.shape
        self._kf = as_f64(kfrac).reshape(-1, 3)
        self._go = as_i32(gofs)
        assert proj_out.shape == (nk, nband, plan.ptot)
        self.ctx = ctx
        self._keep = (plan, G_dev, C_dev, proj_out)
        h = ct.c_void_p()
        check(lib().pp_graph_capture(ctx.h, plan.h, nk, ptr(self._kf, c_double_p), ptr(self._go, c_int_p),
                                     _dptr(G_dev), _dptr(C_dev), ldc, nband, _dptr(proj_out), _prec(precision),
                                     ct.byref(h)), "pp_graph_capture")
        self.h = h

    def replay(self, sync=True):
        check(lib().pp_graph_replay(self.h), "pp_graph_replay")
        _done(self.ctx, sync)

    def close(self):
        if getattr(self, "h", None):
            lib().pp_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


TUNED = ("streams", "fft_tile", "fft_threads", "sphere_unroll", "sphere_parts", "fft_radix", "fft_xcd", "fft_xmask",
         "fft_nt", "sphere_nt", "fft_bands", "fused")


def autotune(ctx: Context, plan: SpherePlan, kfrac, gofs, G_dev, C_dev, proj_out, precision="fp64", repeats=2):
    """pp_autotune on these inputs: the fastest settings stay on ctx and
    proj_out holds the projections of a run with them.  Returns
    (settings, [(name, value, ms), ...])."""
    nk, nband, ldc = C_dev.shape
    kf = as_f64(kfrac).reshape(-1, 3)
    go = as_i32(gofs)
    assert proj_out.shape == (nk, nband, plan.ptot)
    rep = ct.create_string_buffer(8192)
    check(lib().pp_autotune(ctx.h, plan.h, nk, ptr(kf, c_double_p), ptr(go, c_int_p), _dptr(G_dev), _dptr(C_dev),
                            ldc, nband, _dptr(proj_out), _prec(precision), int(repeats), rep, len(rep)),
          "pp_autotune")
    trials = []
    for line in rep.value.decode().splitlines():
        n, v, ms = line.split()
        trials.append((n, int(v), float(ms)))
    return {n: ctx.get_param(n) for n in TUNED}, trials


def overlap_gemm(ctx: Context, nwk, nspin, flip, nw, CR, CS, out, precision="fp64", sync=True):
    """out[k, r, s] = <CR[kR(k), r] | CS[k, s]> over the first nw[k] plane waves."""
    nkt, nbR, ldR = CR.shape
    _, nbS, ldS = CS.shape
    nwa = as_i32(nw)
    assert out.shape == (nkt, nbR, nbS)
    check(lib().pp_overlap_gemm(ctx.h, nwk, nspin, int(bool(flip)), ptr(nwa, c_int_p), _dptr(CR), ldR, nbR,
                                _dptr(CS), ldS, nbS, _dptr(out), _prec(precision)), "pp_overlap_gemm")
    _done(ctx, sync)


def augmentation(ctx: Context, nwk, nspin, flip, projR, proj_off_R, labels_R, projS, proj_off_S, labels_S,
                 M_R, M_S, species_list, dO_list, out, sync=True):
    """out[k, r, s] += O_M of projector.c:890-910 (FP64)."""
    nkt, nbR, ptotR = projR.shape
    _, nbS, ptotS = projS.shape
    offR, offS = as_i32(proj_off_R), as_i32(proj_off_S)
    labR, labS = as_i32(labels_R), as_i32(labels_S)
    mr, ms = as_i32(M_R), as_i32(M_S)
    nchan = as_i32([len(sp.ls) for sp in species_list])
    ls = as_i32([l for sp in species_list for l in sp.ls])
    dO = as_f64(np.concatenate([np.asarray(d, np.float64).ravel() for d in dO_list]))
    check(lib().pp_augmentation(ctx.h, nwk, nspin, int(bool(flip)), _dptr(projR), nbR, ptotR, ptr(offR, c_int_p),
                                ptr(labR, c_int_p), _dptr(projS), nbS, ptotS, ptr(offS, c_int_p),
                                ptr(labS, c_int_p), len(mr), ptr(mr, c_int_p), ptr(ms, c_int_p), len(species_list),
                                ptr(nchan, c_int_p), ptr(ls, c_int_p), ptr(dO, c_double_p), _dptr(out)),
          "pp_augmentation")
    _done(ctx, sync)


def fft3d(ctx: Context, lattice, fftg, G_dev, C_dev, out, precision="fp64", sync=True):
    """Real-space grids (c128 [nband, N0, N1, N2]) of up to 64 bands."""
    nband, ldc = C_dev.shape
    lat = as_f64(lattice).ravel()
    fg = as_i32(fftg)
    check(lib().pp_fft3d_batched(ctx.h, ptr(lat, c_double_p), ptr(fg, c_int_p), G_dev.shape[0], _dptr(G_dev),
                                 _dptr(C_dev), ldc, nband, _dptr(out), _prec(precision)), "pp_fft3d_batched")
    _done(ctx, sync)


def wave_arrays(species_list, diffwaves=None):
    """Flatten species (objects with ls, grid, aewaves, pswaves) into the
    pp_rs_plan_create arrays: per species its log wave grid and the
    partial-wave differences ae - ps of every channel (projector.c:84-99),
    or the given ``diffwaves[e][ch]``."""
    nchan = as_i32([len(sp.ls) for sp in species_list])
    ls = as_i32([l for sp in species_list for l in sp.ls])
    wn = as_i32([len(sp.grid) for sp in species_list])
    wg = as_f64(np.concatenate([np.asarray(sp.grid, np.float64) for sp in species_list]))
    dw = []
    for e, sp in enumerate(species_list):
        for ch in range(len(sp.ls)):
            dw.append(np.asarray(diffwaves[e][ch], np.float64) if diffwaves is not None else
                      np.asarray(sp.aewaves[ch], np.float64) - np.asarray(sp.pswaves[ch], np.float64))
    return nchan, ls, wn, wg, as_f64(np.concatenate(dw) if dw else np.zeros(1))


class RealspacePlan:
    """pp_rsplan: partial-wave spheres of one structure on a real-space
    grid (the site loops of realspace_state, density.c:259-309)."""

    def __init__(self, ctx: Context, lattice, fftg, frac_coords, labels, species_list, diffwaves=None):
        self.ctx = ctx
        lat = as_f64(lattice).ravel()
        self.fftg = tuple(int(n) for n in fftg)
        fg = as_i32(fftg)
        fr = as_f64(frac_coords).reshape(-1, 3)
        lab = as_i32(labels)
        nchan, ls, wn, wg, dw = wave_arrays(species_list, diffwaves)
        h = ct.c_void_p()
        check(lib().pp_rs_plan_create(ctx.h, ct.byref(h), ptr(lat, c_double_p), ptr(fg, c_int_p), len(fr),
                                      ptr(lab, c_int_p), ptr(fr, c_double_p), len(species_list), ptr(nchan, c_int_p),
                                      ptr(ls, c_int_p), ptr(wn, c_int_p), ptr(wg, c_double_p), ptr(dw, c_double_p)),
              "pp_rs_plan_create")
        self.h = h
        self.nsites = len(fr)
        tp, nc, at = ct.c_int(), ct.c_int(), ct.c_int()
        npt = ct.c_long()
        off = np.zeros(self.nsites + 1, np.int32)
        check(lib().pp_rs_plan_info(h, ct.byref(tp), ct.byref(npt), ct.byref(nc), ct.byref(at), ptr(off, c_int_p)),
              "pp_rs_plan_info")
        self.ptot, self.npts, self.ncolors, self.atomic = tp.value, npt.value, nc.value, bool(at.value)
        self.proj_off = off

    @property
    def gridsize(self):
        return int(np.prod(self.fftg))

    def close(self):
        if getattr(self, "h", None):
            lib().pp_rs_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def realspace_states(ctx: Context, plan: RealspacePlan, kfrac, G_dev, C_dev, proj_dev, states_out=None,
                     density_out=None, weights=None, remove_phase=False, precision="fp64", sync=True):
    """All-electron states and/or weighted density of the bands of one
    k-point (density.c:28-39, 158-179, 232-326).  C_dev complex64 [nband,
    ldc], G_dev int32 [W, 3] (W <= ldc), proj_dev complex128 [nband, ldp]
    in the plan's projector order; states_out complex128 [nband, N0, N1,
    N2]; density_out float64 [N0, N1, N2] += sum_b weights[b] |psi_b|^2."""
    nband, ldc = C_dev.shape
    kf = as_f64(kfrac).ravel()
    ldp = proj_dev.shape[1] if proj_dev is not None else 1
    w = None
    if density_out is not None:
        w = as_f64(np.ones(nband) if weights is None else weights)
        assert len(w) == nband
    if states_out is not None:
        assert states_out.shape[0] == nband and states_out[0].numel() == plan.gridsize
    check(lib().pp_realspace_batched(ctx.h, plan.h, ptr(kf, c_double_p), G_dev.shape[0], _dptr(G_dev),
                                     _dptr(C_dev), ldc, nband,
                                     _dptr(proj_dev) if proj_dev is not None else None, ldp,
                                     1 if remove_phase else 0,
                                     _dptr(states_out) if states_out is not None else None, plan.gridsize,
                                     ptr(w, c_double_p) if w is not None else None,
                                     _dptr(density_out) if density_out is not None else None,
                                     _prec(precision)), "pp_realspace_batched")
    _done(ctx, sync)


def _multi_arrays(plans, C, proj):
    """Pointer, stride and band-count arrays of the multi-structure entry
    points: C[s][k] complex64 [nband_sk, ldc_sk] and proj[s][k] complex128
    [nband_sk, ptot_s] (cuda tensors, rows contiguous; views into larger
    tensors allowed)."""
    ns, nk = len(plans), len(C[0])
    assert len(C) == ns and len(proj) == ns and all(len(c) == nk and len(p) == nk for c, p in zip(C, proj))
    nb = np.zeros(ns * nk, np.int32)
    ld = np.zeros(ns * nk, np.int64)
    cp = (ct.c_void_p * (ns * nk))()
    pp = (ct.c_void_p * (ns * nk))()
    for s in range(ns):
        for k in range(nk):
            c, p = C[s][k], proj[s][k]
            assert str(c.dtype) == "torch.complex64" and str(p.dtype) == "torch.complex128"
            assert c.shape[0] == p.shape[0] and p.shape[1] == plans[s].ptot
            assert c.stride(1) == 1 and p.stride(1) == 1 and p.stride(0) == plans[s].ptot
            i = s * nk + k
            nb[i] = c.shape[0]
            ld[i] = c.stride(0) if c.shape[0] > 1 else max(c.shape[1], 1)
            cp[i] = c.data_ptr()
            pp[i] = p.data_ptr()
    hp = (ct.c_void_p * ns)(*[p.h.value for p in plans])
    return hp, nb, ld, cp, pp


def setup_projections_multi(ctx: Context, plans, kfrac, gofs, G_dev, C, proj, precision="fp64", sync=True):
    """<p|psi> of several structures sharing one plane-wave basis, every
    (structure, k, 64-band chunk) as one work list (pp_setup_projections_multi):
    plans[s] the sphere plans; kfrac [nk, 3], gofs [nk + 1] and G_dev int32
    [sum W, 3] the basis; C[s][k] complex64 [nband_sk, ldc] and proj[s][k]
    complex128 [nband_sk, ptot_s] per structure and k-point."""
    hp, nb, ld, cp, pp = _multi_arrays(plans, C, proj)
    kf = as_f64(kfrac).reshape(-1, 3)
    go = as_i32(gofs)
    check(lib().pp_setup_projections_multi(ctx.h, len(plans), hp, len(kf), ptr(kf, c_double_p), ptr(go, c_int_p),
                                           _dptr(G_dev), ptr(nb, c_int_p), cp,
                                           ld.ctypes.data_as(ct.POINTER(ct.c_long)), pp, _prec(precision)),
          "pp_setup_projections_multi")
    _done(ctx, sync)


class MultiGraph:
    """pp_graph_capture_multi: the multi-structure setup of fixed buffers
    captured once (after one eager run), replayed on the context's stream."""

    def __init__(self, ctx: Context, plans, kfrac, gofs, G_dev, C, proj, precision="fp64"):
        hp, nb, ld, cp, pp = _multi_arrays(plans, C, proj)
        self._kf = as_f64(kfrac).reshape(-1, 3)
        self._go = as_i32(gofs)
        self._keep = (plans, G_dev, C, proj, hp, nb, ld, cp, pp)
        self.ctx = ctx
        h = ct.c_void_p()
        check(lib().pp_graph_capture_multi(ctx.h, len(plans), hp, len(self._kf), ptr(self._kf, c_double_p),
                                           ptr(self._go, c_int_p), _dptr(G_dev), ptr(nb, c_int_p), cp,
                                           ld.ctypes.data_as(ct.POINTER(ct.c_long)), pp, _prec(precision),
                                           ct.byref(h)), "pp_graph_capture_multi")
        self.h = h

    def replay(self, sync=True):
        check(lib().pp_graph_replay(self.h), "pp_graph_replay")
        _done(self.ctx, sync)

    def close(self):
        if getattr(self, "h", None):
            lib().pp_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


TUNED_MULTI = ("multi_items", "streams", "fft_tile", "sphere_unroll", "sphere_parts", "fft_nt", "sphere_nt",
               "fft_xmask")


def autotune_multi(ctx: Context, plans, kfrac, gofs, G_dev, C, proj, precision="fp64", repeats=2):
    """pp_autotune_multi on these inputs; returns (settings, trials) as autotune."""
    hp, nb, ld, cp, pp = _multi_arrays(plans, C, proj)
    kf = as_f64(kfrac).reshape(-1, 3)
    go = as_i32(gofs)
    rep = ct.create_string_buffer(8192)
    check(lib().pp_autotune_multi(ctx.h, len(plans), hp, len(kf), ptr(kf, c_double_p), ptr(go, c_int_p),
                                  _dptr(G_dev), ptr(nb, c_int_p), cp, ld.ctypes.data_as(ct.POINTER(ct.c_long)), pp,
                                  _prec(precision), int(repeats), rep, len(rep)), "pp_autotune_multi")
    trials = []
    for line in rep.value.decode().splitlines():
        n, v, ms = line.split()
        trials.append((n, int(v), float(ms)))
    return {n: ctx.get_param(n) for n in TUNED_MULTI}, trials
