"""ctypes binding of libeuvl.so (include/euvl.h).

The library is required: importing the GPU entry points without it raises
immediately; there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeuvl.so")

c_int, c_int64, c_double, c_char, c_void_p = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_char, ctypes.c_void_p
P_int = ctypes.POINTER(ctypes.c_int)
P_double = ctypes.POINTER(ctypes.c_double)


class EuvlError(RuntimeError):
    pass


# status of euvl_graph_launch when the handle's workspace moved since the capture (include/euvl.h)
ERR_STALE = -6


# euvl_set_precision bits (include/euvl.h)
PREC_FP32_MATFUN, PREC_FP32_COUPLING, PREC_FP32_LU, PREC_FP32_GEMM = 1, 2, 4, 8
PREC_GEMM_4M = 16  # complex products with 4 real MFMAs, FP64 and FP32 (default: 3, the zgemm3m form)
PREC_FP32_IMAGING = 32  # Abbe and SOCS transforms in FP32 (x pass on FP32 MFMAs), compensated sums
PREC_FP64 = 0
# the precision form: mode-coupling and top-solve products and the whole imaging in FP32; the matrix
# functions of the absorber operator (the eigendecomposition's counterpart) and the multilayer cascade in FP64
PREC_MIXED = PREC_FP32_COUPLING | PREC_FP32_LU | PREC_FP32_IMAGING
# every product of the EM solve in FP32 too (the matrix functions as well): faster, larger error
PREC_FAST = PREC_FP32_MATFUN | PREC_FP32_COUPLING | PREC_FP32_LU | PREC_FP32_IMAGING


class EmParams(ctypes.Structure):
    _fields_ = [
        ("ndivx", c_int), ("ndivy", c_int), ("mx", c_int), ("my", c_int),
        ("na", c_double), ("wavelength", c_double), ("k", c_double),
        ("dx", c_double), ("dy", c_double),
        ("lsmax_x", c_int), ("lsmax_y", c_int),
        ("eabs_re", c_double), ("eabs_im", c_double), ("dabs", c_double),
        ("n_layers", c_int), ("h_layer_eps", P_double), ("h_layer_d", P_double),
    ]


class StccParams(ctypes.Structure):
    _fields_ = [
        ("k", c_double), ("kx0", c_double), ("ky0", c_double), ("z0", c_double), ("dx", c_double),
        ("dy", c_double), ("na", c_double), ("mx", c_int), ("my", c_int), ("lpmax_x", c_int), ("lpmax_y", c_int),
        ("xdiv", c_int),
    ]


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("polar", c_char), ("p", ctypes.POINTER(EmParams)), ("h_lindex", P_int), ("h_mindex", P_int),
        ("nrange", c_int), ("lmax", c_int), ("mmax", c_int), ("n_off", c_int), ("h_off", P_double),
        ("n_clips", c_int), ("kx0", c_double), ("ky0", c_double), ("z0", c_double),
        ("ncut", c_int), ("h_pupil3", P_int), ("lpmax_x", c_int), ("lpmax_y", c_int), ("h_vc", P_double),
        ("h_ab", P_double), ("n_src", c_int), ("h_src", P_int), ("z", c_double), ("nx", c_int), ("ny", c_int),
        ("weight", c_double), ("alpha_margin", c_double),
    ]


# name -> (argtypes); every entry point returns int
_SIGS = {
    "euvl_create": [ctypes.POINTER(c_void_p)],
    "euvl_destroy": [c_void_p],
    "euvl_set_stream": [c_void_p, c_void_p],
    "euvl_workspace_bytes": [c_void_p],
    "euvl_reserve_workspace": [c_void_p, c_int64],
    "euvl_release_workspace": [c_void_p],
    "euvl_gemm_stats": [c_void_p, c_int, P_double],
    "euvl_set_precision": [c_void_p, c_int],
    "euvl_autotune_gemm": [c_void_p, c_int, c_int, c_int, c_int, c_int, P_int, P_double],
    "euvl_set_fft_min_n": [c_int, c_int],
    "euvl_zgemm_batched": [c_void_p, c_int, c_int, c_int, P_double, c_void_p, c_int64, c_int64, c_void_p, c_int64,
                           c_int64, P_double, c_void_p, c_int64, c_int64, P_double, c_int],
    "euvl_zgetrf_batched": [c_void_p, c_int, c_void_p, c_int64, c_int64, c_void_p, c_int],
    "euvl_zgetrf_nopiv_batched": [c_void_p, c_int, c_void_p, c_int64, c_int64, c_void_p, c_int],
    "euvl_zgesv_nopiv_batched": [c_void_p, c_int, c_int, c_void_p, c_int64, c_int64, c_int],
    "euvl_zgetrs_batched": [c_void_p, c_int, c_int, c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_int64,
                            c_int64, c_int],
    "euvl_zcos_sinc_batched": [c_void_p, c_int, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_int, P_int],
    "euvl_mask_spectrum": [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p],
    "euvl_mask_spectrum_noalloc": [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p],
    "euvl_mask_spectrum_noalloc_max": [],
    "euvl_multilayer": [c_void_p, c_char, ctypes.POINTER(EmParams), c_int, c_void_p, c_void_p, c_void_p, c_void_p],
    "euvl_amp_s": [c_void_p, c_char, ctypes.POINTER(EmParams), P_int, P_int, c_int, c_int, c_int, c_void_p, c_int,
                   P_double, P_double, c_void_p, P_int],
    "euvl_abbe_image": [c_void_p, c_char, ctypes.POINTER(EmParams), c_void_p, c_int, c_int, c_int, P_int, c_int, P_double,
                        c_int, P_int, c_int, P_int, c_double, c_double, c_double, c_double, c_int, c_int, c_double,
                        c_void_p],
    "euvl_m3d_fit": [c_void_p, ctypes.POINTER(EmParams), c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_int,
                     P_int, P_double, P_double, c_double, c_double, c_double, c_int, P_int, c_void_p, c_void_p,
                     c_void_p],
    "euvl_linear_images": [c_void_p, ctypes.POINTER(EmParams), c_void_p, c_int, c_int, P_double, P_double, P_double,
                           c_void_p, c_void_p, c_void_p, c_int, P_int, c_int, P_double, c_double, c_double, c_double,
                           c_double, c_double, c_int, c_int, c_void_p, c_void_p],
    "euvl_tcc_kernels": [c_void_p, ctypes.POINTER(EmParams), c_int, P_int, c_int, P_double, c_double, c_double,
                         c_double, c_double, c_int, c_void_p, c_void_p, P_int],
    "euvl_socs_image": [c_void_p, ctypes.POINTER(EmParams), c_void_p, c_int, c_int, P_double, P_double, P_double,
                        c_void_p, c_void_p, c_void_p, c_int, P_int, c_void_p, c_void_p, c_int, c_int, c_int, c_int,
                        c_double, c_double, c_double, c_int, c_int, c_void_p],
    "euvl_mask_linespace": [c_void_p, ctypes.c_uint, c_int, c_int, c_int, c_int, c_int, c_void_p],
    "euvl_mask_linespace_plan": [ctypes.c_uint, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                 c_int64, ctypes.POINTER(c_int64)],
    "euvl_mask_raster": [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                         c_int, c_void_p],
    "euvl_mask_raster_ms": [c_void_p, P_double],
    "euvl_cnn_compmask": [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p],
    "euvl_cnn_conv_block": [c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_void_p],
    "euvl_cnn_linear": [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p],
    "euvl_stcc_image": [c_void_p, c_char, ctypes.POINTER(StccParams), c_void_p, P_double, P_double, P_double, c_void_p,
                        c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p],
    "euvl_plan_create": [c_void_p, ctypes.POINTER(PlanDesc), ctypes.POINTER(c_void_p)],
    "euvl_plan_destroy": [c_void_p],
    "euvl_plan_workspace_bytes": [c_void_p],
    "euvl_plan_reserve": [c_void_p, c_int, ctypes.POINTER(c_void_p)],
    "euvl_plan_run": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int],
    "euvl_plan_decisions": [c_void_p, P_int],
    "euvl_plan_set_decisions": [c_void_p, c_int],
    "euvl_graph_begin": [c_void_p],
    "euvl_graph_end": [c_void_p, ctypes.POINTER(c_void_p)],
    "euvl_graph_launch": [c_void_p, c_void_p],
    "euvl_graph_nodes": [c_void_p, ctypes.POINTER(c_int64)],
    "euvl_graph_destroy": [c_void_p],
    "euvl_augment_batch": [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                           c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
}

EXPORTS = tuple(_SIGS) + ("euvl_last_error",)

_lib = None


def load():
    """Load libeuvl.so (raises EuvlError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EuvlError(f"{LIB_PATH} not found: build it with `python -m euvlitho_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int64 if name in ("euvl_workspace_bytes", "euvl_plan_workspace_bytes",
                                                "euvl_mask_spectrum_noalloc_max") else c_int
    lib.euvl_last_error.argtypes = []
    lib.euvl_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check(status, what=""):
    if status != 0:
        msg = load().euvl_last_error().decode(errors="replace")
        raise EuvlError(f"{what} failed with status {status}: {msg}")


def ptr(t):
    """Raw device pointer of a torch tensor (must be contiguous)."""
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return c_void_p(t.data_ptr())


def arr_i(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(P_int)


def arr_d(a):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(P_double)
