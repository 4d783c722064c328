"""ctypes binding of libofea_hip.so (declared in include/openfea_amd.h).

The library is built in-tree (openfea_amd/libofea_hip.so) by
``make -C openfea_amd/csrc``.  There is no fallback: every call goes to the
native library and a missing build raises at import of the first call.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libofea_hip.so")

OFEA_OK = 0
OFEA_NOT_CONVERGED = 1
OFEA_ERR_ARG = -1
OFEA_ERR_HIP = -2
OFEA_ERR_STATE = -3
OFEA_ERR_IO = -4
OFEA_ERR_LIMIT = -5

C3D4 = 0
C3D8 = 1

PCG_JACOBI = 0
PCG_BLOCK_JACOBI = 1
PCG_MIXED = 2
PCG_LEARNED = 3
PCG_MIXED32 = 4
PCG_LEARNED_MIXED = 5
PCG_TWO_LEVEL = 6
PCG_TWO_LEVEL_MIXED = 7
PCG_LEARNED_TWO_LEVEL = 8
PCG_LEARNED_TWO_LEVEL_MIXED = 9
PCG_TWO_LEVEL_MIXED32 = 10
PCG_LEARNED_TWO_LEVEL_MIXED32 = 11

FP32 = 0
FP64 = 1
BF16 = 2

POST_STRAIN = 0
POST_STRESS = 1

_d = C.c_double
_i = C.c_int
_f = C.c_float
_ll = C.c_longlong
_vp = C.c_void_p
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
_pll = C.POINTER(C.c_longlong)

# name -> (restype, argtypes); every symbol of include/openfea_amd.h
SIGNATURES = {
    "ofea_last_error": (C.c_char_p, []),
    "ofea_version": (C.c_char_p, []),
    "ofea_inp_open": (_i, [C.c_char_p, C.POINTER(_vp)]),
    "ofea_inp_sizes": (_i, [_vp, _pi, _pi, _pi, _pi]),
    "ofea_inp_get": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _pi, _pd, _pd, _pd, _pd, _pd, _pd]),
    "ofea_inp_close": (None, [_vp]),
    "ofea_build_free_index": (_i, [_i, _vp, _i, _vp]),
    "ofea_build_single_load": (_i, [_i, _vp, _vp, _i, _i, _d, _vp]),
    "ofea_reset_displacement": (_i, [_i, _vp, _vp, _vp]),
    "ofea_build_csr": (_i, [_i, _i, _vp, _vp, _vp, _pi, _pll, _vp, _vp]),
    "ofea_create": (_i, [C.POINTER(_vp), _i]),
    "ofea_destroy": (_i, [_vp]),
    "ofea_set_mesh": (_i, [_vp, _i, _vp, _i, _vp, _vp, _vp]),
    "ofea_pattern": (_i, [_vp, _pi, _pll, _vp, _vp]),
    "ofea_set_elem_E": (_i, [_vp, _vp]),
    "ofea_assemble": (_i, [_vp, _d, _d, _vp]),
    "ofea_spmv": (_i, [_vp, _vp, _vp]),
    "ofea_set_rhs": (_i, [_vp, _vp]),
    "ofea_solve": (_i, [_vp, _i, _vp, _vp, _d, _i, _pi, _pd]),
    "ofea_solve_noalloc": (_i, [_vp, _i, _vp, _vp, _d, _i, _pi, _pd]),
    "ofea_get_solution": (_i, [_vp, _vp]),
    "ofea_lumped_mass": (_i, [_vp, _d, _vp]),
    "ofea_explicit_run": (_i, [_vp, _d, _d, _vp, _vp, _vp, _vp, _d, _i, _i, _vp]),
    "ofea_explicit_analysis": (_i, [_vp, _d, _d, _vp, _vp, _vp, _vp, _d, _d, _i, _i, _i, _i, _vp, _vp, _vp, _pi,
                                    _pi]),
    "ofea_build_explicit_load": (_i, [_i, _vp, _i, _i, _d, _vp]),
    "ofea_nl_assemble": (_i, [_vp, _d, _d, _vp, _vp, _vp]),
    "ofea_nl_solve": (_i, [_vp, _d, _d, _vp, _i, _d, _i, _i, _d, _vp, _pi, _pi, _pd]),
    "ofea_nl_solve_ref": (_i, [_vp, _d, _d, _vp, _d, _i, _i, _d, _vp, _pi, _pi, _pd]),
    "ofea_learned_nparam": (_i, []),
    "ofea_learned_init": (_i, [_vp, C.c_ulonglong]),
    "ofea_learned_set_params": (_i, [_vp, _vp, _i]),
    "ofea_learned_get_params": (_i, [_vp, _vp, _i]),
    "ofea_learned_loss_grad": (_i, [_vp, _pd, _vp]),
    "ofea_learned_train": (_i, [_vp, _i, _f, _vp]),
    "ofea_learned_train_multi": (_i, [_vp, _i, _i, _f, _vp]),
    "ofea_learned_blocks": (_i, [_vp, _vp, _vp]),
    "ofea_learned_outputs": (_i, [_vp, _vp]),
    "ofea_learned_refresh": (_i, [_vp]),
    "ofea_learned_precision": (_i, [_vp, _i]),
    "ofea_coarse_apply": (_i, [_vp, _vp, _vp]),
    "ofea_coarse_aggregates": (_i, [_vp, _vp]),
    "ofea_coarse_builds": (_i, [_vp]),
    "ofea_assemble_batch": (_i, [_vp, _i, _vp, _d, _vp]),
    "ofea_solve_batch": (_i, [_vp, _i, _i, _vp, _vp, _d, _i, _vp, _vp]),
    "ofea_solve_batch_pc": (_i, [_vp, _i, _i, _vp, _vp, _d, _i, _i, _vp, _vp]),
    "ofea_autotune": (_i, [_vp, _pi, _vp]),
    "ofea_batch_tiles": (_i, [_vp, _i, _pi, _pf]),
    "ofea_timer_start": (_i, [_vp]),
    "ofea_timer_stop": (_i, [_vp, _pf]),
    "ofea_sync": (_i, [_vp]),
    "ofea_stream_copy": (_i, [_vp, _ll, _i, _pf]),
    "ofea_reset_ca_displacement": (_i, [_i, _vp, _vp, _i, _vp, _i, _vp]),
    "ofea_ca_info_open": (_i, [C.c_char_p, C.POINTER(_vp)]),
    "ofea_ca_info_sizes": (_i, [_vp, _pi, _pi, _pi]),
    "ofea_ca_info_get": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ofea_ca_info_close": (None, [_vp]),
    "ofea_write_vtk": (_i, [C.c_char_p, _i, _vp, _i, _vp, _vp, _vp, _d, _vp]),
    "ofea_ca_delta": (_i, [_vp, _d, _d, _i, _vp, _vp, _vp, _vp, _vp]),
    "ofea_ca_solve": (_i, [_vp, _i, _vp, _vp, _i, _d, _i, _vp, _pi, _pi, _vp]),
    "ofea_post_nodal": (_i, [_vp, _d, _d, _vp, _i, _vp, _vp]),
    "ofea_inp_nodal_loads": (_i, [_vp, _vp]),
    "ofea_inp_boundary": (_i, [_vp, _pi, _vp, _pi, _pi]),
    "ofea_inp_initial_velocity": (_i, [_vp, _vp]),
    "ofea_inp_materials": (_i, [_vp, _vp, _vp, _vp]),
    "ofea_sfem_setup": (_i, [_vp]),
    "ofea_sfem_assemble": (_i, [_vp, _d, _d, _vp]),
}

_LIB = None


class OfeaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"openfea_amd error {code}: {msg}")
        self.code = code


def lib():
    """Load the native library (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"native library missing: {LIB_PATH}; build it with `make -C openfea_amd/csrc` "
                "(the HIP path has no fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc, allow=(OFEA_OK,)):
    if rc not in allow:
        msg = lib().ofea_last_error().decode(errors="replace")
        raise OfeaError(rc, msg)
    return rc


def ptr(a):
    """Data pointer of a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(_vp)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
