"""Work lists of the fused x pass (built on the host, pp_fused_lists_host)
checked against the oracle's sphere point lists (utils.c:590-696), no GPU.

For every site and every (y, z) line holding some of its points: the
block of the line has exactly one slot for the site among its work items;
the points of the site on the line are one run of consecutive x (cyclic),
so one segment per (site, line) covers them; an item holds at most 12
sites; the partial rows of every site are its (item, slot) pairs.
"""

import ctypes as ct

import numpy as np
import pytest

from oracle import pyoracle as orc
from pawpyseed_amd import synthetic
from pawpyseed_amd.core._lib import lib

BY = BZ = 4
G = 12


def _lists(lattice, fftg, frac, rmax, ngrid):
    L = lib()
    f = L.pp_fused_lists_host
    f.restype = ct.c_int
    dp, ip = ct.POINTER(ct.c_double), ct.POINTER(ct.c_int)
    f.argtypes = [dp, ip, ct.c_int, dp, dp, ip, ip, ip, ip, ip, ip, ip]
    lat = np.ascontiguousarray(lattice, np.float64).ravel()
    fg = np.ascontiguousarray(fftg, np.int32)
    fr = np.ascontiguousarray(frac, np.float64).reshape(-1, 3)
    rm = np.ascontiguousarray(rmax, np.float64)
    ng = np.ascontiguousarray(ngrid, np.int32)
    cnt = np.zeros(3, np.int32)
    P = lambda a, t: a.ctypes.data_as(t)  # noqa: E731
    st = f(P(lat, dp), P(fg, ip), len(fr), P(fr, dp), P(rm, dp), P(ng, ip), P(cnt, ip), None, None, None, None,
           None)
    if st != 0:
        return None
    nit, nblk, nrows = (int(v) for v in cnt)
    items = np.zeros(4 * nit, np.int32)
    isites = np.zeros(G * nit, np.int32)
    bitem = np.zeros(nblk + 1, np.int32)
    roff = np.zeros(len(fr) + 1, np.int32)
    rows = np.zeros(max(1, nrows), np.int32)
    assert f(P(lat, dp), P(fg, ip), len(fr), P(fr, dp), P(rm, dp), P(ng, ip), P(cnt, ip), P(items, ip),
             P(isites, ip), P(bitem, ip), P(roff, ip), P(rows, ip)) == 0
    return items.reshape(-1, 4), isites.reshape(-1, G), bitem, roff, rows[:nrows]


def _check(lattice, fftg, frac, species, labels):
    fftg = np.asarray(fftg)
    rmax = np.array([species[l].rmax for l in labels])
    ngrid = np.array([species[l].npts for l in labels], np.int32)
    out = _lists(lattice, fftg, frac, rmax, ngrid)
    assert out is not None
    items, isites, bitem, roff, rows = out
    N0, N1, N2 = (int(n) for n in fftg)
    nbz = -(-N2 // BZ)
    assert bitem[0] == 0 and bitem[-1] == len(items) and np.all(np.diff(bitem) >= 0)
    for b in range(len(bitem) - 1):
        for it in range(bitem[b], bitem[b + 1]):
            assert items[it, 0] == (b // nbz) * BY and items[it, 1] == (b % nbz) * BZ
            used = isites[it][isites[it] >= 0]
            assert items[it, 2] == len(used) <= G
            assert len(set(used.tolist())) == len(used)
    # rows: site -> (item, slot) pairs
    for s in range(len(frac)):
        want = sorted(it * G + q for it in range(len(items)) for q in range(G) if isites[it, q] == s)
        assert sorted(rows[roff[s]:roff[s + 1]].tolist()) == want
    T = orc.site_tables(lattice, fftg, frac, species, labels)
    nseg = 0
    for s in range(len(frac)):
        idx = T.idx[T.site_pt[s]:T.site_pt[s + 1]]
        x, yz = idx // (N1 * N2), idx % (N1 * N2)
        for line in np.unique(yz):
            y, z = divmod(int(line), N2)
            b = (y // BY) * nbz + z // BZ
            slots = [(it, q) for it in range(bitem[b], bitem[b + 1]) for q in range(G) if isites[it, q] == s]
            assert len(slots) == 1, (s, y, z)
            xs = np.sort(x[yz == line])
            # one cyclic run of consecutive x
            gaps = np.count_nonzero(np.diff(xs) != 1) + (0 if len(xs) == N0 else int((xs[0] - xs[-1]) % N0 != 1))
            assert gaps <= 1 or len(xs) == N0, (s, y, z, xs)
            nseg += 1
    return len(items), nseg


@pytest.mark.parametrize("name", ["tiny", "ragged", "dense", "cfg0", "hexagonal"])
def test_fused_lists_cover_every_segment(name):
    if name == "tiny":
        wl = synthetic.tiny_config(fftg=(24, 24, 24), ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    elif name == "ragged":
        wl = synthetic.tiny_config(fftg=(30, 18, 22), ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    elif name == "dense":
        # overlapping spheres: 153 items for 100 blocks
        wl = synthetic.tiny_config(ncell=2, fftg=(40, 40, 40), ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5), rmax=2.2)
    elif name == "cfg0":
        wl = synthetic.config0()
    else:
        wl = synthetic.tiny_config(fftg=(30, 30, 48))
        a, c = 3.2, 5.2
        wl.lattice = np.array([[a, 0, 0], [-a / 2, a * np.sqrt(3) / 2, 0], [0, 0, c]])
    labels = np.zeros(len(wl.frac_S), np.int32)
    nit, nseg = _check(wl.lattice, wl.fftg, wl.frac_S, [wl.species], labels)
    print(name, "items", nit, "site lines", nseg)
    if name == "dense":
        assert nit > 100


def test_fused_lists_two_species():
    wl = synthetic.tiny_config(fftg=(24, 24, 24), ls=(0, 1, 2), alphas=(1.5, 3.0, 1.5))
    sp2 = synthetic.make_species([0, 1, 1], [1.2, 1.4, 2.8], 0.95, npts=80)
    labels = (np.arange(len(wl.frac_S)) % 2).astype(np.int32)
    _check(wl.lattice, wl.fftg, wl.frac_S, [wl.species, sp2], labels)


def test_fused_lists_216_atom_cell():
    """The 216-atom, 108^3 workload: every block fits one work item."""
    wl = synthetic.config1(nband=8)
    labels = np.zeros(len(wl.frac_S), np.int32)
    nit, nseg = _check(wl.lattice, wl.fftg, wl.frac_S, [wl.species], labels)
    assert nit == 27 * 27


def test_fused_lists_refuse_wrapping_boxes():
    wl = synthetic.tiny_config(fftg=(12, 12, 12), ls=(0, 1), rmax=2.6)
    n = len(wl.frac_S)
    assert _lists(wl.lattice, wl.fftg, wl.frac_S, np.full(n, 2.6), np.full(n, 100, np.int32)) is None
