"""ctypes mirrors of the reference-layout structs ppot_t / funcset_t
(pawpyseed/core/utils.h:30-70), used to read or overwrite the PAW overlap
matrices of a ppot_t list returned by get_projector_list."""

import ctypes as ct

_dp = ct.POINTER(ct.c_double)
_dpp = ct.POINTER(_dp)


class FuncSet(ct.Structure):
    _fields_ = [("l", ct.c_int), ("proj", _dp), ("proj_spline", _dpp), ("aewave", _dp), ("aewave_spline", _dpp),
                ("pswave", _dp), ("pswave_spline", _dpp), ("diffwave", _dp), ("diffwave_spline", _dpp),
                ("kwave", _dp), ("kwave_spline", _dpp), ("smooth_diffwave", _dp), ("smooth_diffwave_spline", _dpp),
                ("dense_kwave", _dp), ("dense_kwave_spline", _dpp)]


class PPot(ct.Structure):
    _fields_ = [("num_projs", ct.c_int), ("total_projs", ct.c_int), ("lmax", ct.c_int),
                ("funcs", ct.POINTER(FuncSet)), ("rmax", ct.c_double), ("wave_rmax", ct.c_double),
                ("pspw_overlap_matrix", _dp), ("aepw_overlap_matrix", _dp), ("diff_overlap_matrix", _dp),
                ("proj_gridsize", ct.c_int), ("wave_gridsize", ct.c_int), ("num_cart_gridpts", ct.c_int),
                ("wave_grid", _dp), ("kwave_grid", _dp), ("proj_grid", _dp), ("smooth_grid", _dp),
                ("dense_kgrid", _dp)]


def ppot_array(ptr_value, n):
    return ct.cast(ct.c_void_p(ptr_value), ct.POINTER(PPot * n)).contents
