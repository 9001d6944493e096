"""Seeded synthetic inputs for the projection path.

No VASP outputs or POTCARs can be shipped, so every workload is generated:
a diamond-lattice cell, its plane-wave basis enumerated exactly as the
reference's WAVECAR reader does (``pawpyseed/core/reader.c:55-127`` for the
box bounds, ``reader.c:231-271`` for the VASP ordering of G vectors),
Gaussian-damped random plane-wave coefficients, and analytic projector
radial functions tabulated on the reference's linear projector grid
(``pawpyseed/core/wavefunction.py:134-136``: ``x_j = j * rmax / n``).

Random streams: ``numpy.random.SeedSequence(20261202).spawn(5)``;
stream 0 geometry jitter, 1 R-coefficients, 2 S-coefficients, 3 the PAW
overlap difference matrix, 4 k-points and band counts.
"""

from __future__ import annotations

import dataclasses
import math

import numpy as np

#: hbar^2/2m_e in eV*A^2 inverted: E[eV] = |k|^2 [A^-2] / C_EV (reader.c:11)
C_EV = 0.262465831
SEED = 20261202

DIAMOND_BASIS = np.array(
    [
        [0.0, 0.0, 0.0],
        [0.0, 0.5, 0.5],
        [0.5, 0.0, 0.5],
        [0.5, 0.5, 0.0],
        [0.25, 0.25, 0.25],
        [0.25, 0.75, 0.75],
        [0.75, 0.25, 0.75],
        [0.75, 0.75, 0.25],
    ]
)


def streams(n: int = 5):
    return [np.random.Generator(np.random.PCG64(s)) for s in np.random.SeedSequence(SEED).spawn(n)]


def reciprocal_lattice(lattice: np.ndarray) -> np.ndarray:
    """Rows b_i with b_i . a_j = 2 pi delta_ij (the reference keeps 2 pi, utils.h:131)."""
    return 2.0 * np.pi * np.linalg.inv(np.asarray(lattice, dtype=np.float64)).T


def diamond_cell(ncell: int, a0: float = 5.43):
    """Cubic supercell of ``ncell^3`` conventional diamond cells."""
    lattice = np.eye(3) * (a0 * ncell)
    frac = []
    for i in range(ncell):
        for j in range(ncell):
            for k in range(ncell):
                frac.append((DIAMOND_BASIS + np.array([i, j, k])) / ncell)
    return lattice, np.concatenate(frac, axis=0)


def _box_bounds(lattice: np.ndarray, encut: float):
    """Half-extents of the G box; the same geometric bound as reader.c:55-127."""
    rec = reciprocal_lattice(lattice)
    mags = np.linalg.norm(rec, axis=1)
    gmax = math.sqrt(encut * C_EV)

    def sin_between(u, v):
        return abs(math.sin(math.acos(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v)))))

    def sin_to_plane(w, u, v):
        cr = np.cross(u, v)
        return abs(np.dot(w, cr) / (np.linalg.norm(cr) * np.linalg.norm(w)))

    b1, b2, b3 = rec
    nA = [gmax / (mags[0] * sin_between(b1, b2)) + 1,
          gmax / (mags[1] * sin_between(b1, b2)) + 1,
          gmax / (mags[2] * sin_to_plane(b3, b1, b2)) + 1]
    nB = [gmax / (mags[0] * sin_between(b1, b3)) + 1,
          gmax / (mags[1] * sin_to_plane(b2, b1, b3)) + 1,
          gmax / (mags[2] * sin_between(b1, b3)) + 1]
    nC = [gmax / (mags[0] * sin_to_plane(b1, b2, b3)) + 1,
          gmax / (mags[1] * sin_between(b2, b3)) + 1,
          gmax / (mags[2] * sin_between(b2, b3)) + 1]
    return [max(nA[i], nB[i], nC[i]) for i in range(3)]


def gvectors(lattice: np.ndarray, kpt, encut: float) -> np.ndarray:
    """Integer G vectors inside the cutoff sphere, in VASP (WAVECAR) order.

    z-index outermost, x innermost, each index running 0..nmax then the
    negative half (reader.c:231-271); kept when |k+G|^2 / C_EV <= encut.
    """
    nb = _box_bounds(lattice, encut)
    rec = reciprocal_lattice(lattice)

    def axis(nmax):
        # loop index i in [0, 2*nmax] (nmax is a double in the reference; the
        # comparisons are made against it as a double)
        top = int(math.floor(2 * nmax))
        out = []
        for i in range(top + 1):
            out.append(i if i <= nmax else int(i - 2 * nmax - 1))
        return np.array(out, dtype=np.int64)

    g1, g2, g3 = axis(nb[0]), axis(nb[1]), axis(nb[2])
    # ordering: ig3 outer, ig2 middle, ig1 inner
    G3, G2, G1 = np.meshgrid(g3, g2, g1, indexing="ij")
    G = np.stack([G1.ravel(), G2.ravel(), G3.ravel()], axis=1)
    k = np.asarray(kpt, dtype=np.float64)
    kg = (G + k[None, :]) @ rec
    e = np.einsum("ij,ij->i", kg, kg) / C_EV
    return G[e <= encut].astype(np.int32)


def gaussian_coefficients(rng, lattice, kpt, G, nband, encut, dtype=np.complex64):
    """c_G ~ CN(0,1) * exp(-|k+G|^2 / (2 (Gcut/2)^2)), each band normalised."""
    rec = reciprocal_lattice(lattice)
    kg = (G.astype(np.float64) + np.asarray(kpt)[None, :]) @ rec
    g2 = np.einsum("ij,ij->i", kg, kg)
    gcut = math.sqrt(C_EV * encut)
    env = np.exp(-g2 / (2.0 * (0.5 * gcut) ** 2))
    c = (rng.standard_normal((nband, len(G))) + 1j * rng.standard_normal((nband, len(G)))) / math.sqrt(2.0)
    c *= env[None, :]
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    return c.astype(dtype)


@dataclasses.dataclass
class Species:
    """Projector set of one element, in the reference's POTCAR-derived form.

    ``realprojs[i]`` is tabulated on ``projgrid = arange(n) * rmax / n``
    (wavefunction.py:134-136); ``ls[i]`` its angular momentum.  ``aewaves``
    and ``pswaves`` are partial waves on the logarithmic ``grid``.
    """

    ls: list
    alphas: list
    rmax: float
    npts: int
    realprojs: list
    projgrid: np.ndarray
    grid: np.ndarray
    aewaves: list
    pswaves: list

    @property
    def total_projs(self) -> int:
        return int(sum(2 * l + 1 for l in self.ls))


def smooth_cut(r, rmax):
    x = np.clip(r / rmax, 0.0, 1.0)
    return (1.0 - x * x) ** 3


def make_species(ls, alphas, rmax, npts=100, nlog=400):
    projgrid = np.arange(npts) * rmax / npts
    realprojs = [projgrid ** l * np.exp(-a * projgrid ** 2) * smooth_cut(projgrid, rmax)
                 for l, a in zip(ls, alphas)]
    # logarithmic partial-wave grid ending at the augmentation radius
    grid = rmax * np.exp(np.linspace(math.log(1e-4), 0.0, nlog))
    pswaves, aewaves = [], []
    for l, a in zip(ls, alphas):
        ps = grid ** (l + 1) * np.exp(-0.5 * a * grid ** 2)
        bump = 0.3 * grid ** (l + 1) * np.exp(-4.0 * a * grid ** 2) * smooth_cut(grid, rmax)
        pswaves.append(ps)
        aewaves.append(ps + bump)
    return Species(list(ls), list(alphas), rmax, npts, realprojs, projgrid, grid, aewaves, pswaves)


def channel_delta_o(rng, ls, scale=0.1, dense=False):
    """Symmetric channel matrix with N(0, scale) entries; cross-l entries zeroed
    unless ``dense`` (the reference only couples equal-l channels,
    projector.c:519 and :897-906)."""
    n = len(ls)
    a = rng.normal(0.0, scale, (n, n))
    a = 0.5 * (a + a.T)
    if not dense:
        ls = np.asarray(ls)
        a[ls[:, None] != ls[None, :]] = 0.0
    return a


def monkhorst_pack(n1, n2, n3):
    ks = []
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                ks.append([(2 * i - n1 + 1) / (2 * n1), (2 * j - n2 + 1) / (2 * n2), (2 * k - n3 + 1) / (2 * n3)])
    return np.array(ks, dtype=np.float64)


@dataclasses.dataclass
class Workload:
    """One reference structure R and one projected structure S sharing a basis."""

    name: str
    lattice: np.ndarray
    frac_R: np.ndarray
    frac_S: np.ndarray
    species: Species
    fftg: np.ndarray
    encut: float
    kpts: np.ndarray
    kws: np.ndarray
    nband: int
    delta_o: np.ndarray          # channel x channel (aepw - pspw)
    M_R: np.ndarray              # matched sites (index lists for compensation_terms)
    M_S: np.ndarray
    gvecs: list = None           # per k-point int32 [W,3]

    def ensure_gvecs(self):
        if self.gvecs is None:
            self.gvecs = [gvectors(self.lattice, k, self.encut) for k in self.kpts]
        return self.gvecs

    def coefficients(self, which: str, rng=None, nband=None):
        """Per k-point complex64 [nband, W_k] arrays for structure 'R' or 'S'."""
        st = streams()
        rng = rng or (st[1] if which == "R" else st[2])
        nb = nband or self.nband
        return [gaussian_coefficients(rng, self.lattice, k, G, nb, self.encut)
                for k, G in zip(self.kpts, self.ensure_gvecs())]


def config0(jitter=0.05):
    """64 diamond sites, a=10.86 A, 400 eV, Gamma only, 60^3 grid, 64 bands,
    8 projectors per site (l=0,0,1,1; alpha 1.5/3.0; rmax 1.2 A)."""
    st = streams()
    lattice, frac = diamond_cell(2)
    cart = frac @ lattice + st[0].normal(0.0, jitter, frac.shape)
    frac_S = np.mod(cart @ np.linalg.inv(lattice), 1.0)
    sp = make_species([0, 0, 1, 1], [1.5, 3.0, 1.5, 3.0], 1.2)
    dO = channel_delta_o(st[3], sp.ls)
    idx = np.arange(len(frac), dtype=np.int32)
    return Workload("cfg0", lattice, frac, frac_S, sp, np.array([60, 60, 60], np.int32), 400.0,
                    np.zeros((1, 3)), np.ones(1), 64, dO, idx, idx.copy())


def config1(nband=512):
    """216 diamond sites, a=16.29 A, 520 eV, 2x2x1 MP (4 k), 108^3 grid,
    512 bands, 18 projectors per site (l=0,1,2 x alpha 1.5/3.0; rmax 1.6 A).
    S: one vacancy (site 0 removed) and its 4 nearest neighbours moved
    0.15 A towards it; compensation site lists match the 215 S sites to
    their R counterparts by index."""
    st = streams()
    lattice, frac = diamond_cell(3)
    sp = make_species([0, 0, 1, 1, 2, 2], [1.5, 3.0, 1.5, 3.0, 1.5, 3.0], 1.6)
    cart = frac @ lattice
    vac = 0
    d = cart - cart[vac]
    d -= np.round(d @ np.linalg.inv(lattice)) @ lattice
    dist = np.linalg.norm(d, axis=1)
    nn = np.argsort(dist)[1:5]
    cart_S = cart.copy()
    for i in nn:
        cart_S[i] -= 0.15 * d[i] / dist[i]
    keep = np.array([i for i in range(len(frac)) if i != vac])
    frac_S = np.mod(cart_S[keep] @ np.linalg.inv(lattice), 1.0)
    dO = channel_delta_o(st[3], sp.ls)
    M_R = keep.astype(np.int32)
    M_S = np.arange(len(keep), dtype=np.int32)
    kpts = monkhorst_pack(2, 2, 1)
    return Workload("cfg1", lattice, frac, frac_S, sp, np.array([108, 108, 108], np.int32), 520.0,
                    kpts, np.full(len(kpts), 1.0 / len(kpts)), nband, dO, M_R, M_S)


def config4(nband=2048):
    """Large noncollinear case: 512 diamond sites, a=21.72 A, 520 eV
    (W ~ 275.9k per spinor component), 2 k-points (2x1x1 MP), 192^3 grid,
    spinor bands of two W_k-coefficient components, 18 projectors per site
    (l=0,1,2 x alpha 1.5/3.0; rmax 1.6 A).  S: seeded 0.05 A jitter of R."""
    st = streams()
    lattice, frac = diamond_cell(4)
    cart = frac @ lattice + st[0].normal(0.0, 0.05, frac.shape)
    frac_S = np.mod(cart @ np.linalg.inv(lattice), 1.0)
    sp = make_species([0, 0, 1, 1, 2, 2], [1.5, 3.0, 1.5, 3.0, 1.5, 3.0], 1.6)
    dO = channel_delta_o(st[3], sp.ls)
    idx = np.arange(len(frac), dtype=np.int32)
    kpts = monkhorst_pack(2, 1, 1)
    return Workload("cfg4", lattice, frac, frac_S, sp, np.array([192, 192, 192], np.int32), 520.0,
                    kpts, np.full(len(kpts), 1.0 / len(kpts)), nband, dO, idx, idx.copy())


def tiny_config(ncell=1, encut=150.0, fftg=(20, 20, 20), nband=8, kpts=None, ls=(0, 1), alphas=(1.5, 3.0),
                rmax=1.1, jitter=0.03):
    """Small cell for tests that the pure-Python checks finish in seconds."""
    st = streams()
    lattice, frac = diamond_cell(ncell)
    cart = frac @ lattice + st[0].normal(0.0, jitter, frac.shape)
    frac_S = np.mod(cart @ np.linalg.inv(lattice), 1.0)
    sp = make_species(list(ls), list(alphas), rmax)
    dO = channel_delta_o(st[3], sp.ls)
    kp = np.zeros((1, 3)) if kpts is None else np.asarray(kpts, np.float64)
    idx = np.arange(len(frac), dtype=np.int32)
    return Workload("tiny", lattice, frac, frac_S, sp, np.array(fftg, np.int32), encut,
                    kp, np.full(len(kp), 1.0 / len(kp)), nband, dO, idx, idx.copy())


def device_coefficients(torch, wl, Gs, seed, nband, ldc, nbands=None):
    """Gaussian-damped random coefficients (as gaussian_coefficients) drawn
    on the GPU with a seeded torch generator: complex64 [nk, nband, ldc],
    zero beyond each k-point's W_k.  ``nbands`` (optional, per k) zeroes the
    bands of k beyond nbands[k]."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(int(seed))
    rec = reciprocal_lattice(wl.lattice)
    gcut = math.sqrt(C_EV * wl.encut)
    C = torch.zeros((len(Gs), nband, ldc), dtype=torch.complex64, device="cuda")
    for k, (kp, G) in enumerate(zip(wl.kpts, Gs)):
        kg = (G.astype(np.float64) + np.asarray(kp)[None, :]) @ rec
        env = np.exp(-np.einsum("ij,ij->i", kg, kg) / (2.0 * (0.5 * gcut) ** 2)).astype(np.float32)
        envt = torch.from_numpy(env).cuda()
        nb = nband if nbands is None else int(nbands[k])
        re = torch.randn((nb, len(G)), generator=gen, device="cuda")
        im = torch.randn((nb, len(G)), generator=gen, device="cuda")
        c = torch.complex(re * envt, im * envt)
        c /= torch.linalg.vector_norm(c, dim=1, keepdim=True)
        C[k, :nb, : len(G)] = c
        del re, im, c
    return C


#: overlap magnitudes of the planted near-orthogonal band pairs
NEAR_ORTHOGONAL_EPS = (1e-8, 1e-7, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2, 1e-1, 1.0)


def plant_near_orthogonal(torch, CR, CS, nw, pairs, eps=NEAR_ORTHOGONAL_EPS):
    """Near-orthogonal band pairs of the precision workload: for each
    (k, r, s) of ``pairs`` and magnitude e of ``eps`` (cycled), S band s at k
    becomes (d + e R_r) / |d + e R_r| with d its old coefficients made
    orthogonal to R band r (Gram-Schmidt in FP64) and normalised, so that
    <R_r | S_s> = e / sqrt(1 + e^2) before the rounding to complex64.
    CS is changed in place; returns the list (k, r, s, e)."""
    out = []
    for i, (k, r, s) in enumerate(pairs):
        e = float(eps[i % len(eps)])
        W = int(nw[k])
        rv = CR[k, r, :W].to(torch.complex128)
        d = CS[k, s, :W].to(torch.complex128)
        rv = rv / torch.linalg.vector_norm(rv)
        for _ in range(2):  # twice, so the residual overlap is at FP64 rounding
            d = d - torch.vdot(rv, d) * rv
        d = d / torch.linalg.vector_norm(d)
        v = d + e * rv
        v = v / torch.linalg.vector_norm(v)
        CS[k, s, :W] = v.to(torch.complex64)
        out.append((int(k), int(r), int(s), e))
    return out


@dataclasses.dataclass
class MultiWorkload:
    """Several structures projected onto one reference over many k-points
    with uneven plane-wave and band counts.  ``encuts[k]`` is the cutoff of
    k-point k's sphere (its G list: ``gvecs[k]``), ``nbands[i][k]`` the band
    count of structure i (0 = the reference) at k."""

    name: str
    lattice: np.ndarray
    frac: list                   # [reference, structure 1, ...] fractional coordinates
    species: Species
    fftg: np.ndarray
    kpts: np.ndarray
    kws: np.ndarray
    encuts: np.ndarray
    nbands: np.ndarray           # [nstruct + 1, nk]
    delta_o: np.ndarray
    M: list                      # per structure: (M_R, M_S) matched-site lists
    gvecs: list = None

    @property
    def encut(self):
        return float(self.encuts.max())

    def ensure_gvecs(self):
        if self.gvecs is None:
            self.gvecs = [gvectors(self.lattice, k, e) for k, e in zip(self.kpts, self.encuts)]
        return self.gvecs


def wedge_kpoints(rng, nk):
    """Random k-points in the irreducible wedge of the simple-cubic
    Brillouin zone, 0 <= kz <= ky <= kx <= 1/2 (fractional)."""
    k = np.sort(rng.uniform(0.0, 0.5, (nk, 3)), axis=1)[:, ::-1]
    return np.ascontiguousarray(k)


def config3(nk=64, nband_range=(384, 640), spread=0.10):
    """Batched uneven workload: the 216-site cell (a = 16.29 A, 108^3 grid,
    18 projectors per site, rmax 1.6 A) and three perturbed structures
    projected onto the bulk reference: the vacancy of config1 (215 sites),
    a seeded 0.05 A jitter of every site, and 8 sites moved 0.1 A.  nk
    random k-points in the irreducible wedge (stream 4).  With one cutoff,
    shifting k changes the sphere's plane-wave count by well under 1% for
    this cell, so each k-point's cutoff is drawn around 520 eV as
    520 (1 + u)^(2/3), u ~ U[-spread, spread], which makes W_k vary by
    +-spread; bands per (structure, k) ~ U[nband_range] (stream 4)."""
    st = streams()
    rng = st[4]
    base = config1()
    lattice, frac_R = base.lattice, base.frac_R
    cart = frac_R @ lattice
    inv = np.linalg.inv(lattice)
    jit = np.mod((cart + st[0].normal(0.0, 0.05, cart.shape)) @ inv, 1.0)
    moved = cart.copy()
    sel = rng.choice(len(cart), 8, replace=False)
    dirs = rng.normal(size=(8, 3))
    moved[sel] += 0.1 * dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    moved = np.mod(moved @ inv, 1.0)
    kpts = wedge_kpoints(rng, nk)
    u = rng.uniform(-spread, spread, nk)
    encuts = 520.0 * (1.0 + u) ** (2.0 / 3.0)
    nbands = rng.integers(nband_range[0], nband_range[1] + 1, (4, nk))
    idx = np.arange(len(frac_R), dtype=np.int32)
    M = [(base.M_R, base.M_S), (idx, idx.copy()), (idx, idx.copy())]
    return MultiWorkload("cfg3", lattice, [frac_R, base.frac_S, jit, moved], base.species, base.fftg.copy(), kpts,
                         np.full(nk, 1.0 / nk), encuts, nbands, base.delta_o, M)
