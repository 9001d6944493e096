"""Write tests/golden/wavecar_ga4.npz from the reference's test fixture.

Reads the binary data file test_files/WAVECAR that the reference's own
tests use (Ga4, ISPIN=2, 2 k-points, 24 bands, ENCUT=320 eV; the record
layout follows pawpyseed/core/reader.c:129-315) and stores a small subset
as a fixture:

* header: lattice, encut, nspin, nwk, nband, k-points;
* per (spin, k): the plane-wave count recorded in the file, occupations,
  energies and the first NB bands' complex64 coefficients;
* the G vectors of each k enumerated by pawpyseed_amd.synthetic.gvectors;
  the script asserts their count equals the file's plane-wave count,
  which pins that enumeration against VASP's;
* the 20^3 real-space grid of (spin 0, k 0, band 0) by explicit
  plane-wave summation, psi(f) = V^{-1/2} sum_G C_G exp(2 pi i G.f), the
  brute-force check of the reference's fft_check (tests/tests.c:28-80).

A second fixture, wavecar_ncl.npz, comes from the reference's spinor file
test_files/noncollinear/WAVECAR (Ga4, 4 k-points, 24 spinor bands, 520 eV;
each record holds 2 W coefficients, the up then the down component,
reader.c:274-281): the header, and for the first NCL_K k-points the
plane-wave count, the enumerated W G vectors (the script asserts
2 W equals the file's count), and the first NB spinor bands.

Run in the build container (needs /root/reference):
    python tests/golden/make_wavecar_fixture.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from pawpyseed_amd.synthetic import gvectors  # noqa: E402

SRC = "/root/reference/test_files/WAVECAR"
SRC_NCL = "/root/reference/test_files/noncollinear/WAVECAR"
NCL_K = 2
NB = 6
NGRID = (20, 20, 20)  # NGX/NGY/NGZ of test_files/vasprun.xml


def main():
    raw = np.fromfile(SRC, dtype=np.uint8)
    head = raw[:24].view(np.float64)
    nrecl, nspin, nprec = int(round(head[0])), int(round(head[1])), int(round(head[2]))
    assert nprec == 45200, "single-precision WAVECAR expected"
    rec1 = raw[nrecl:nrecl + 12 * 8].view(np.float64)
    nwk, nband, encut = int(round(rec1[0])), int(round(rec1[1])), float(rec1[2])
    lattice = rec1[3:12].reshape(3, 3).copy()

    out = dict(lattice=lattice, encut=encut, nspin=nspin, nwk=nwk, nband=nband, nb_stored=NB)
    kpts = []
    for iwk in range(nwk * nspin):
        irec = iwk * (1 + nband)
        off = (irec + 2) * nrecl
        kr = raw[off:off + nrecl].view(np.float64)
        nplane = int(round(kr[0]))
        k = kr[1:4].copy()
        eo = kr[4:4 + 3 * nband].reshape(nband, 3)
        coeffs = []
        for b in range(NB):
            o = (irec + 1 + b + 2) * nrecl
            coeffs.append(raw[o:o + 8 * nplane].view(np.complex64).copy())
        G = gvectors(lattice, k, encut)
        assert len(G) == nplane, (len(G), nplane)
        if iwk < nwk:
            kpts.append(k)
        out[f"nplane_{iwk}"] = nplane
        out[f"G_{iwk}"] = G
        out[f"C_{iwk}"] = np.stack(coeffs)
        out[f"energy_{iwk}"] = eo[:, 0].copy()
        out[f"occ_{iwk}"] = eo[:, 2].copy()
    out["kpts"] = np.array(kpts)

    # brute-force real-space grid for (spin 0, k 0, band 0)
    G = out["G_0"].astype(np.float64)
    C = out["C_0"][0].astype(np.complex128)
    n0, n1, n2 = NGRID
    f = np.stack(np.meshgrid(np.arange(n0) / n0, np.arange(n1) / n1, np.arange(n2) / n2, indexing="ij"),
                 axis=-1).reshape(-1, 3)
    vol = abs(np.linalg.det(lattice))
    psi = np.zeros(len(f), np.complex128)
    for start in range(0, len(f), 1000):
        ph = np.exp(2j * np.pi * (f[start:start + 1000] @ G.T))
        psi[start:start + 1000] = ph @ C
    out["grid_fftg"] = np.array(NGRID, np.int32)
    out["grid_k0_b0"] = (psi / np.sqrt(vol)).reshape(NGRID)
    np.savez_compressed(os.path.join(HERE, "wavecar_ga4.npz"), **out)
    print("wrote", os.path.join(HERE, "wavecar_ga4.npz"), {k: np.shape(v) for k, v in out.items()})


def main_ncl():
    raw = np.fromfile(SRC_NCL, dtype=np.uint8)
    head = raw[:24].view(np.float64)
    nrecl, nspin, nprec = int(round(head[0])), int(round(head[1])), int(round(head[2]))
    assert nprec == 45200 and nspin == 1
    rec1 = raw[nrecl:nrecl + 12 * 8].view(np.float64)
    nwk, nband, encut = int(round(rec1[0])), int(round(rec1[1])), float(rec1[2])
    lattice = rec1[3:12].reshape(3, 3).copy()
    out = dict(lattice=lattice, encut=encut, nwk=nwk, nband=nband, nb_stored=NB, nk_stored=NCL_K)
    kpts = []
    for iwk in range(NCL_K):
        irec = iwk * (1 + nband)
        off = (irec + 2) * nrecl
        kr = raw[off:off + nrecl].view(np.float64)
        nplane = int(round(kr[0]))
        k = kr[1:4].copy()
        eo = kr[4:4 + 3 * nband].reshape(nband, 3)
        coeffs = []
        for b in range(NB):
            o = (irec + 1 + b + 2) * nrecl
            coeffs.append(raw[o:o + 8 * nplane].view(np.complex64).copy())
        G = gvectors(lattice, k, encut)
        assert 2 * len(G) == nplane, (len(G), nplane)
        kpts.append(k)
        out[f"nplane_{iwk}"] = nplane
        out[f"G_{iwk}"] = G
        out[f"C_{iwk}"] = np.stack(coeffs)
        out[f"energy_{iwk}"] = eo[:NB, 0].copy()
        out[f"occ_{iwk}"] = eo[:NB, 2].copy()
    out["kpts"] = np.array(kpts)
    np.savez_compressed(os.path.join(HERE, "wavecar_ncl.npz"), **out)
    print("wrote", os.path.join(HERE, "wavecar_ncl.npz"), {k: np.shape(v) for k, v in out.items()})


if __name__ == "__main__":
    main()
    main_ncl()
