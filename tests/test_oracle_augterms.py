"""The oracle's O_R, O_S and O_N terms (projector.c:912-959) against an
independent numpy route.

The oracle restates the reference's loops one (band, k) at a time in C.
The route here is written from the equations instead: each term is a
gather of the site blocks followed by one einsum over all bands and
k-points, with the reference's conventions made explicit:

- O_R[b, k] = sum_i <(phi - phi~)_R,i | psi_S,k>  conj(<p_R,i | psi_R,b,kR>)
  (wave projection of the S band times the conjugated R projection);
- O_S[b, k] = sum_i conj(<(phi - phi~)_S,i | psi_R,b,kR>) <p_S,i | psi_S,k>;
- O_N[b, k] = sum_t e^{2 pi i kfrac[kR] . d_t} sum_ij conj(P_R[kR, b, i]) Ov_t[i, j] P_S[k, j],
  with the fractional k of the R k-point (kpt_R->k, projector.c:955);

where kR is k with the spin flipped when flip_spin and nspin == 2
(projector.c:874-882).  The tests use two spins, distinct k vectors per
k-index (so that kpts[k] and kpts[kR] differ), and check that each
plausible slip (the conjugate on the other factor, the unflipped k in the
phase, a transposed block) moves the result far beyond the tolerance.
They also check the exchange identity O_R(R, S)[r, s] = conj(O_S(S, R)[s, r]),
which ties the two terms' conjugate placements together.
"""

import types

import numpy as np
import pytest

from oracle import pyoracle as orc


def _tables(nproj):
    """Stand-in for SiteTables: the terms read only ptot and proj_off."""
    off = np.concatenate([[0], np.cumsum(nproj)]).astype(np.int32)
    return types.SimpleNamespace(proj_off=off, ptot=int(off[-1]))


def _crand(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _kr(nwk, nspin, flip):
    k = np.arange(nwk * nspin)
    if flip and nspin == 2:
        return np.where(k < nwk, k + nwk, k - nwk)
    return k


def _rows(T, sites):
    return np.concatenate([np.arange(T.proj_off[s], T.proj_off[s + 1]) for s in sites]) if len(sites) else \
        np.zeros(0, int)


def _route_rs(nwk, nspin, flip, proj_R, T_R, N_R, wp_S, WT_S, wp_R, WT_R, proj_S, T_S, N_S, conj_slip=False):
    """O_R + O_S as [nbR * nkt] (res[b * nkt + k]); wp_S, proj_S: [nkt, .]
    rows of the projected band."""
    kR = _kr(nwk, nspin, flip)
    PRg = proj_R[kR][:, :, _rows(T_R, N_R)]                   # [k, b, i] at the N_R sites, R's own rows
    WSg = wp_S[:, _rows(WT_S, np.arange(len(N_R)))]           # [k, i] S band onto R's partial waves
    if conj_slip:
        o_r = np.einsum("ki,kbi->kb", np.conj(WSg), PRg)
    else:
        o_r = np.einsum("ki,kbi->kb", WSg, np.conj(PRg))
    WRg = wp_R[kR][:, :, _rows(WT_R, np.arange(len(N_S)))]    # [k, b, i] R bands onto S's partial waves
    PSg = proj_S[:, _rows(T_S, N_S)]                          # [k, i]
    o_s = np.einsum("kbi,ki->kb", np.conj(WRg), PSg)
    return (o_r + o_s).T.ravel()


def _route_n(nwk, nspin, flip, proj_R, T_R, proj_S, T_S, N_RS_R, N_RS_S, ovs, dcoords, kpts, k_slip=False,
             transpose_slip=False):
    kR = _kr(nwk, nspin, flip)
    nkt, nb = proj_R.shape[0], proj_R.shape[1]
    out = np.zeros((nkt, nb), complex)
    for t, (s1, s2) in enumerate(zip(N_RS_R, N_RS_S)):
        PR = proj_R[kR][:, :, T_R.proj_off[s1]:T_R.proj_off[s1 + 1]]
        PS = proj_S[:, T_S.proj_off[s2]:T_S.proj_off[s2 + 1]]
        O = ovs[t] if not transpose_slip else ovs[t].T.reshape(ovs[t].shape)
        kk = kpts[np.arange(nkt)] if k_slip else kpts[kR]
        phase = np.exp(2j * np.pi * (kk @ dcoords[t]))
        out += phase[:, None] * np.einsum("kbi,ij,kj->kb", np.conj(PR), O, PS)
    return out.T.ravel()


def _case(seed=7, nwk=2, nspin=2, nbR=5):
    rng = np.random.default_rng(seed)
    nkt = nwk * nspin
    T_R = _tables([4, 9, 4, 9, 1])
    T_S = _tables([4, 9, 4, 9])
    N_R = np.array([1, 4, 2], np.int32)       # R sites without a match
    N_S = np.array([3, 0], np.int32)          # S sites without a match
    WT_S = _tables([T_R.proj_off[s + 1] - T_R.proj_off[s] for s in N_R])
    WT_R = _tables([T_S.proj_off[s + 1] - T_S.proj_off[s] for s in N_S])
    d = dict(
        proj_R=_crand(rng, (nkt, nbR, T_R.ptot)), wp_S=_crand(rng, (nkt, WT_S.ptot)),
        wp_R=_crand(rng, (nkt, nbR, WT_R.ptot)), proj_S=_crand(rng, (nkt, T_S.ptot)),
        T_R=T_R, T_S=T_S, WT_S=WT_S, WT_R=WT_R, N_R=N_R, N_S=N_S,
        kpts=rng.uniform(-0.5, 0.5, (nkt, 3)), nwk=nwk, nspin=nspin)
    N_RS_R = np.array([1, 4, 4], np.int32)
    N_RS_S = np.array([3, 3, 0], np.int32)
    d["N_RS_R"], d["N_RS_S"] = N_RS_R, N_RS_S
    d["ovs"] = [_crand(rng, (T_R.proj_off[a + 1] - T_R.proj_off[a], T_S.proj_off[b + 1] - T_S.proj_off[b]))
                for a, b in zip(N_RS_R, N_RS_S)]
    d["dcoords"] = rng.uniform(-2.0, 2.0, (len(N_RS_R), 3))
    return d


def _oracle_rs(d, flip):
    return orc.compensation_terms_rs(np.zeros(d["proj_R"].shape[1] * d["proj_R"].shape[0], complex), d["nwk"],
                                     d["nspin"], d["proj_R"], d["T_R"], d["N_R"], d["wp_S"], d["WT_S"], d["wp_R"],
                                     d["WT_R"], d["proj_S"], d["T_S"], d["N_S"], flip=flip)


def _oracle_n(d, flip):
    return orc.compensation_terms_n(np.zeros(d["proj_R"].shape[1] * d["proj_R"].shape[0], complex), d["nwk"],
                                    d["nspin"], d["proj_R"], d["T_R"], d["proj_S"], d["T_S"], d["N_RS_R"],
                                    d["N_RS_S"], d["ovs"], d["dcoords"], d["kpts"], flip=flip)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("flip", [False, True])
def test_rs_terms_match_numpy_route(flip):
    d = _case()
    got = _oracle_rs(d, flip)
    want = _route_rs(d["nwk"], d["nspin"], flip, d["proj_R"], d["T_R"], d["N_R"], d["wp_S"], d["WT_S"], d["wp_R"],
                     d["WT_R"], d["proj_S"], d["T_S"], d["N_S"])
    assert _rel(got, want) < 1e-13
    # the conjugate on the wave projection instead of the R projection
    slip = _route_rs(d["nwk"], d["nspin"], flip, d["proj_R"], d["T_R"], d["N_R"], d["wp_S"], d["WT_S"], d["wp_R"],
                     d["WT_R"], d["proj_S"], d["T_S"], d["N_S"], conj_slip=True)
    assert _rel(got, slip) > 1e-2
    # the unflipped R k-point
    if flip:
        noflip = _route_rs(d["nwk"], d["nspin"], False, d["proj_R"], d["T_R"], d["N_R"], d["wp_S"], d["WT_S"],
                           d["wp_R"], d["WT_R"], d["proj_S"], d["T_S"], d["N_S"])
        assert _rel(got, noflip) > 1e-2


@pytest.mark.parametrize("flip", [False, True])
def test_offsite_term_matches_numpy_route(flip):
    d = _case(seed=11)
    got = _oracle_n(d, flip)
    args = (d["nwk"], d["nspin"], flip, d["proj_R"], d["T_R"], d["proj_S"], d["T_S"], d["N_RS_R"], d["N_RS_S"],
            d["ovs"], d["dcoords"], d["kpts"])
    assert _rel(got, _route_n(*args)) < 1e-13
    assert _rel(got, _route_n(*args, transpose_slip=True)) > 1e-2
    if flip:
        # the phase takes the R k-point's own k vector (kpt_R->k)
        assert _rel(got, _route_n(*args, k_slip=True)) > 1e-2


def test_rs_exchange_identity():
    """Projecting S onto R and R onto S: O_R of one equals the conjugate of
    O_S of the other with the band indices exchanged."""
    rng = np.random.default_rng(3)
    nwk, nspin, nbR, nbS = 2, 1, 4, 3
    nkt = nwk * nspin
    T_R = _tables([4, 9, 9])
    N_R = np.array([2, 0], np.int32)
    WT = _tables([T_R.proj_off[s + 1] - T_R.proj_off[s] for s in N_R])
    P_R = _crand(rng, (nkt, nbR, T_R.ptot))              # <p_R | psi_R>
    W_S = _crand(rng, (nkt, nbS, WT.ptot))               # <(phi - phi~)_R | psi_S>
    none = np.zeros(0, np.int32)
    empty = _tables([])
    for s in range(nbS):
        o_r = orc.compensation_terms_rs(np.zeros(nbR * nkt, complex), nwk, nspin, P_R, T_R, N_R, W_S[:, s], WT,
                                        np.zeros((nkt, nbR, 1), complex), empty, np.zeros((nkt, 1), complex), empty,
                                        none)
        for r in range(nbR):
            # roles exchanged: R' = S (bands W_S as wave projections), S' = R
            o_s = orc.compensation_terms_rs(np.zeros(nbS * nkt, complex), nwk, nspin,
                                            np.zeros((nkt, nbS, 1), complex), empty, none,
                                            np.zeros((nkt, 1), complex), empty, W_S, WT, P_R[:, r], T_R, N_R)
            for k in range(nkt):
                assert abs(o_r[r * nkt + k] - np.conj(o_s[s * nkt + k])) < 1e-12
