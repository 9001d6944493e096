"""Shared test helpers: synthetic inputs as device tensors and oracle results."""

import numpy as np

from oracle import pyoracle as orc


def pack_coeffs(Cs, ld_mult=64):
    """List of per-k complex64 [nb, W_k] -> padded [nk, nb, ldc] and nw."""
    nw = np.array([c.shape[1] for c in Cs], np.int32)
    ldc = int(max(ld_mult, -(-nw.max() // ld_mult) * ld_mult))
    out = np.zeros((len(Cs), Cs[0].shape[0], ldc), np.complex64)
    for k, c in enumerate(Cs):
        out[k, :, : c.shape[1]] = c
    return out, nw


def device_inputs(Gs, Cs, device="cuda"):
    import torch
    C, nw = pack_coeffs(Cs)
    gofs = np.concatenate([[0], np.cumsum([len(g) for g in Gs])]).astype(np.int32)
    G = np.ascontiguousarray(np.concatenate([np.asarray(g, np.int32) for g in Gs]))
    return torch.from_numpy(C).to(device), torch.from_numpy(G).to(device), gofs, nw


def oracle_projections(wl, frac, Cs, pairs=None, nthreads=0):
    T = orc.site_tables(wl.lattice, wl.fftg, frac, wl.species)
    P = orc.setup_projections(wl.lattice, wl.fftg, wl.kpts, wl.ensure_gvecs(), Cs, T, pairs=pairs,
                              nthreads=nthreads)
    return P, T


def normwise(a, b):
    """max |a - b| / max |b|"""
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))
