"""ctypes binding of ``libflexmm_amd.so`` (declarations in include/flexmm_amd.h).

Loading fails loudly if the library has not been built: there is no CPU
fallback for any entry point.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libflexmm_amd.so")

ABI_VERSION = 1

# mirrors of the #defines in include/flexmm_amd.h
E_ARG, E_STATE, E_LIMIT, E_INTERNAL = -1, -2, -3, -4
KERNEL_LAPLACE = 0
KERNEL_MODIFIED_HELMHOLTZ = 1
KERNEL_STOKESLET = 2
EVAL_POTENTIAL, EVAL_GRADIENT, EVAL_STOKES, EVAL_DIPOLE, EVAL_STOKES_DIPOLE = 0, 1, 2, 3, 4
PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED_M2L = 2
M2L_SVD, M2L_DENSE = 0, 1
TREE_AUTO, TREE_GENERAL = 0, 1
TREE_ARRAYS = dict(keys=0, perm=1, sx=2, sy=3, code=4, level=5, bot=6, top=7, parent=8, child=9,
                   flags=10, colleagues=11, xmid=12, ymid=13, row=14, m2l=15, evalid=16)
EXP_ARRAYS = dict(check=0, multipole=1, local=2, tau_sorted=3)
NF_LEAF, NF_FAKE, NF_XLIST, NF_UP, NF_DOWN, NF_CHILDREN, NF_SOURCE = 1, 2, 4, 8, 16, 32, 64

# every exported symbol and its ctypes signature (restype, argtypes)
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
SIGNATURES = {
    "fmm2d_abi_version": (_int, []),
    "fmm2d_last_error": (ctypes.c_char_p, []),
    "fmm2d_plan_create": (_int, [_int, _vp, _int, _int, _int, ctypes.POINTER(_vp)]),
    "fmm2d_plan_create_cheb": (_int, [_int, _vp, _int, _int, _int, ctypes.POINTER(_vp)]),
    "fmm2d_plan_destroy": (_int, [_vp]),
    "fmm2d_tree_build": (_int, [_vp, _vp, _vp, _i64, ctypes.POINTER(ctypes.c_double), _i64, _vp]),
    "fmm2d_tree_info": (_int, [_vp, ctypes.POINTER(_i64), _i64]),
    "fmm2d_tree_export": (_int, [_vp, _int, _vp, _i64, _vp]),
    "fmm2d_operator_size": (_i64, [_int]),
    "fmm2d_set_operators": (_int, [_vp, _int, _vp, _i64, _vp]),
    "fmm2d_clear_operators": (_int, [_vp]),
    "fmm2d_upward": (_int, [_vp, _vp, _i64, _vp]),
    "fmm2d_upward_dipole": (_int, [_vp, _vp, _vp, _i64, _vp]),
    "fmm2d_downward": (_int, [_vp, _vp]),
    "fmm2d_evaluate": (_int, [_vp, _vp, _vp, _i64, _int, _int, _vp, _vp]),
    "fmm2d_evaluate_noalloc": (_int, [_vp, _vp, _vp, _i64, _int, _int, _vp, _vp]),
    "fmm2d_export_expansions": (_int, [_vp, _int, _vp, _i64, _vp]),
    "fmm2d_set_m2l_mode": (_int, [_vp, _int]),
    "fmm2d_set_tree_mode": (_int, [_vp, _int]),
    "fmm2d_set_profiling": (_int, [_vp, _int]),
    "fmm2d_phase_times": (_int, [_vp, ctypes.POINTER(ctypes.c_double), _int]),
    "fmm2d_work_counts": (_int, [_vp, ctypes.POINTER(_i64), _int]),
}
WORK_COUNTS = ("p2p_pairs", "l2t_targets", "targets", "m2l_pairs", "s2m_points", "up_rows",
               "down_boxes", "m2m_parents", "l2l_children")
PHASES = ("tree", "upward", "downward", "evaluate", "eval_kernel", "s2m", "m2l")


class FMMError(RuntimeError):
    """A non-zero return code of the C-ABI, with the library's message."""

    def __init__(self, code, message):
        super().__init__("flexmm_amd error %d: %s" % (code, message))
        self.code = code


_lock = threading.Lock()
_lib = None


def load(path=LIB_PATH):
    """Load and declare the library (once)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                "flexmm_amd HIP library not found at %s; build it with "
                "`python -m flexmm_amd.build` (there is no CPU fallback)" % path)
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        v = lib.fmm2d_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError("libflexmm_amd ABI version %d, expected %d" % (v, ABI_VERSION))
        _lib = lib
        return lib


def check(code):
    """Raise FMMError for a non-zero return code."""
    if code != 0:
        msg = _lib.fmm2d_last_error()
        raise FMMError(code, msg.decode() if msg else "")
