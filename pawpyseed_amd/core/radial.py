"""Radial helpers of the reference's Cython module (pawpyc.pyx:130-190):
spherical Bessel transforms and off-site partial-wave overlaps, computed by
libpawpy_amd.so's host code."""

from __future__ import annotations

import numpy as np

from ._lib import as_f64, c_double_p, check, lib, ptr


def spherical_bessel_transform(encut, l, r, f):
    """pawpyc.pyx:130-150: (k grid, g(k) = int r^2 j_l(k r) f(r)/r dr) for
    f = r R(r) on the log grid r."""
    r, f = as_f64(r), as_f64(f)
    n = len(r)
    ks, g = np.zeros(n), np.zeros(n)
    check(lib().pp_sbt_forward(float(encut), 0.0, int(l), n, ptr(r, c_double_p), ptr(f, c_double_p), int(l),
                               ptr(ks, c_double_p), ptr(g, c_double_p)), "spherical_bessel_transform")
    return ks, g


def reciprocal_offsite_wave_overlap(dcoord, r1, f1, r2, f2, l1, m1, l2, m2):
    """pawpyc.pyx:151-190: int phi_1(r) conj(phi_2(r - dcoord)) d^3r for
    phi_i = f_i / r Y_{l_i m_i} (f_i = r R_i on the log grids r_i)."""
    d, r1, f1, r2, f2 = (as_f64(a) for a in (dcoord, r1, f1, r2, f2))
    out = np.zeros(2)
    check(lib().pp_offsite_overlap(ptr(d, c_double_p), len(r1), ptr(r1, c_double_p), ptr(f1, c_double_p), len(r2),
                                   ptr(r2, c_double_p), ptr(f2, c_double_p), int(l1), int(m1), int(l2), int(m2),
                                   ptr(out, c_double_p)), "reciprocal_offsite_wave_overlap")
    return complex(out[0], out[1])
