"""ctypes binding of libmf6hip.so (the C ABI declared in include/mf6hip.h).

The library is built in-tree (``modflow6_amd/libmf6hip.so``) by
``__graft_entry__.build()`` / ``make -C modflow6_amd/csrc``.  There is no
fallback: if the library is missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmf6hip.so")

MF6HIP_FP64 = 0
# status codes (include/mf6hip.h)
MF6HIP_OK, MF6HIP_EINVAL, MF6HIP_EHIP, MF6HIP_EUNSUPPORTED = 0, 1, 2, 3
MF6HIP_MIXED = 1
MF6HIP_MIXED_FP16 = 2
MF6HIP_ORDER_RB = 0
PKG = {"WEL": 1, "DRN": 2, "RCH": 3, "CHD": 4}
# MF6HIP_TERM_*: the flow terms of mf6hip_gwf_budget / mf6hip_gwf_flow_term
TERM = {"STO-SS": 0, "STO-SY": 1, "WEL": 2, "DRN": 3, "RCH": 4, "CHD": 5}

# every entry point declared in include/mf6hip.h
EXPORTS = [
    "mf6hip_create", "mf6hip_solve", "mf6hip_destroy", "mf6hip_last_error",
    "mf6hip_gwf_create", "mf6hip_gwf_set_package", "mf6hip_gwf_set_package_ex",
    "mf6hip_gwf_set_hfb", "mf6hip_gwf_set_gnc",
    "mf6hip_gwf_set_x",
    "mf6hip_gwf_get_x", "mf6hip_gwf_set_x_device", "mf6hip_gwf_get_x_device",
    "mf6hip_gwf_advance", "mf6hip_gwf_formulate",
    "mf6hip_gwf_linear_solve", "mf6hip_gwf_outer_update",
    "mf6hip_gwf_residual_norms", "mf6hip_gwf_backtrack", "mf6hip_gwf_budget",
    "mf6hip_gwf_flow_term",
    "mf6hip_last_timing", "mf6hip_set_timing", "mf6hip_apply_op",
    "mf6hip_load_factor", "mf6hip_get_system",
    "mf6hip_unet_create", "mf6hip_unet_destroy", "mf6hip_unet_num_params",
    "mf6hip_unet_set_params", "mf6hip_unet_get_params", "mf6hip_unet_forward",
    "mf6hip_unet_grad", "mf6hip_unet_set_data", "mf6hip_unet_train",
    "mf6hip_attach_unet", "mf6hip_last_learned_iterations",
    "mf6hip_batch_create", "mf6hip_batch_add", "mf6hip_batch_finalize",
    "mf6hip_batch_set_system", "mf6hip_batch_autotune", "mf6hip_batch_solve",
    "mf6hip_batch_get_x", "mf6hip_batch_destroy",
    "mf6hip_xmi_set_nonlinear", "mf6hip_xmi_prepare_solve", "mf6hip_xmi_solve",
    "mf6hip_xmi_last_iteration", "mf6hip_xmi_finalize_solve",
    "mf6hip_csr_residual_norms_noalloc",
]


class ImsSettingsC(C.Structure):
    """mf6hip_ims_settings (mirrors ImsLinearSettingsType)."""
    _fields_ = [("dvclose", C.c_double), ("rclose", C.c_double),
                ("icnvgopt", C.c_int), ("iter1", C.c_int),
                ("ilinmeth", C.c_int), ("iscl", C.c_int), ("iord", C.c_int),
                ("north", C.c_int), ("relax", C.c_double),
                ("level", C.c_int), ("droptol", C.c_double),
                ("precision", C.c_int), ("ordering", C.c_int),
                ("nlay", C.c_int), ("nrow", C.c_int), ("ncol", C.c_int),
                ("inner_fp32_max", C.c_int),
                ("inner_fp32_rtol", C.c_double),
                ("kernel_variant", C.c_int),
                ("preconditioner", C.c_int),
                ("coarse_size", C.c_int)]


class CnvgOutC(C.Structure):
    _fields_ = [("capacity", C.c_int), ("count", C.c_int),
                ("dvmax", C.POINTER(C.c_double)),
                ("rmax", C.POINTER(C.c_double)),
                ("alpha", C.POINTER(C.c_double)),
                ("omega", C.POINTER(C.c_double)),
                ("locdv", C.POINTER(C.c_int)),
                ("locr", C.POINTER(C.c_int))]


class GwfModelC(C.Structure):
    _fields_ = [("nlay", C.c_int), ("nrow", C.c_int), ("ncol", C.c_int),
                ("delr", C.POINTER(C.c_double)),
                ("delc", C.POINTER(C.c_double)),
                ("top", C.POINTER(C.c_double)),
                ("bot", C.POINTER(C.c_double)),
                ("k11", C.POINTER(C.c_double)),
                ("k33", C.POINTER(C.c_double)),
                ("icelltype", C.POINTER(C.c_int)),
                ("ibound", C.POINTER(C.c_int)),
                ("inewton", C.c_int), ("inewtonur", C.c_int),
                ("ivarcv", C.c_int), ("idewatcv", C.c_int),
                ("icellavg", C.c_int), ("has_sto", C.c_int),
                ("iconvert", C.POINTER(C.c_int)),
                ("ss", C.POINTER(C.c_double)),
                ("sy", C.POINTER(C.c_double)),
                ("k22", C.POINTER(C.c_double)),
                ("angle1", C.POINTER(C.c_double)),
                ("angle2", C.POINTER(C.c_double)),
                ("angle3", C.POINTER(C.c_double)),
                ("istor_coef", C.c_int), ("iconf_ss", C.c_int), ("iorig_ss", C.c_int),
                ("ithickstrt", C.c_int), ("strt", C.POINTER(C.c_double))]


class PtcC(C.Structure):
    """mf6hip_ptc: pseudo-transient continuation control and state."""
    _fields_ = [("iptc", C.c_int), ("iallowptc", C.c_int),
                ("iptcopt", C.c_int), ("kstp", C.c_int), ("kper", C.c_int),
                ("ptcf", C.c_double), ("ptcdel0", C.c_double),
                ("ptcexp", C.c_double), ("ptcdel", C.c_double),
                ("l2norm0", C.c_double)]


class UnderRelaxC(C.Structure):
    """mf6hip_underrelax: IMS under-relaxation settings and Cooley state."""
    _fields_ = [("nonmeth", C.c_int), ("gamma", C.c_double),
                ("theta", C.c_double), ("akappa", C.c_double),
                ("amomentum", C.c_double), ("bigch", C.c_double),
                ("bigchold", C.c_double), ("relaxold", C.c_double)]


class PackageOptionsC(C.Structure):
    """mf6hip_package_options: WEL AUTO_FLOW_REDUCE, DRN AUXDEPTHNAME /
    DEV_CUBIC_SCALING, AUXMULTNAME (rejected)."""
    _fields_ = [("flowred", C.c_double), ("iflowred", C.c_int), ("idrndepth", C.c_int),
                ("icubic_scaling", C.c_int), ("has_auxmult", C.c_int),
                ("drndepth", C.POINTER(C.c_double))]


class NonlinearC(C.Structure):
    """mf6hip_nonlinear: the IMS NONLINEAR block."""
    _fields_ = [("outer_dvclose", C.c_double), ("outer_maximum", C.c_int),
                ("nonmeth", C.c_int), ("gamma", C.c_double), ("theta", C.c_double),
                ("akappa", C.c_double), ("amomentum", C.c_double),
                ("backtracking_number", C.c_int),
                ("backtracking_tolerance", C.c_double),
                ("backtracking_reduction_factor", C.c_double),
                ("backtracking_residual_limit", C.c_double),
                ("iallowptc", C.c_int), ("iptcopt", C.c_int),
                ("ptcexp", C.c_double), ("ptcdel0", C.c_double)]


class XmiIterationC(C.Structure):
    """mf6hip_xmi_iteration: what one mf6hip_xmi_solve call did."""
    _fields_ = [("kiter", C.c_int), ("inner", C.c_int), ("icnvg_inner", C.c_int),
                ("hncg", C.c_double), ("lrch", C.c_int), ("converged", C.c_int),
                ("nur", C.c_int), ("nbt", C.c_int), ("iptc", C.c_int)]


class Mf6HipError(RuntimeError):
    pass


_LIB = None


def load():
    """Load libmf6hip.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise Mf6HipError(
            f"{LIBPATH} not found: build it with __graft_entry__.build() or "
            "make -C modflow6_amd/csrc")
    L = C.CDLL(LIBPATH)
    i, d, vp = C.c_int, C.c_double, C.c_void_p
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    L.mf6hip_create.argtypes = [i, i, ip, ip, C.POINTER(ImsSettingsC),
                                C.POINTER(vp)]
    L.mf6hip_solve.argtypes = [vp, i, i, dp, dp, dp, ip, ip,
                               C.POINTER(CnvgOutC)]
    L.mf6hip_destroy.argtypes = [vp]
    L.mf6hip_last_error.argtypes = [C.c_char_p, i]
    L.mf6hip_gwf_create.argtypes = [vp, C.POINTER(GwfModelC)]
    L.mf6hip_gwf_set_package.argtypes = [vp, i, i, ip, dp, dp]
    L.mf6hip_gwf_set_package_ex.argtypes = [vp, i, i, ip, dp, dp, C.POINTER(PackageOptionsC)]
    L.mf6hip_gwf_set_hfb.argtypes = [vp, i, ip, ip, dp]
    L.mf6hip_gwf_set_gnc.argtypes = [vp, i, i, ip, ip, ip, dp, i]
    L.mf6hip_gwf_set_x.argtypes = [vp, dp]
    L.mf6hip_gwf_get_x.argtypes = [vp, dp]
    L.mf6hip_gwf_set_x_device.argtypes = [vp, vp]
    L.mf6hip_gwf_budget.argtypes = [vp, dp, dp, C.POINTER(C.c_double)]
    L.mf6hip_gwf_flow_term.argtypes = [vp, i, dp, ip]
    L.mf6hip_gwf_get_x_device.argtypes = [vp, vp]
    L.mf6hip_gwf_advance.argtypes = [vp, i, d]
    L.mf6hip_gwf_formulate.argtypes = [vp, i, i, dp, dp]
    L.mf6hip_gwf_residual_norms.argtypes = [vp, dp, dp]
    L.mf6hip_gwf_backtrack.argtypes = [vp, d, d, ip]
    L.mf6hip_gwf_linear_solve.argtypes = [vp, i, i, C.POINTER(PtcC), ip, ip,
                                          dp, ip, C.POINTER(CnvgOutC)]
    L.mf6hip_gwf_outer_update.argtypes = [vp, i, C.POINTER(UnderRelaxC), ip,
                                          dp, dp, ip]
    L.mf6hip_last_timing.argtypes = [vp, dp, dp, ip]
    L.mf6hip_set_timing.argtypes = [vp, i, i]
    L.mf6hip_apply_op.argtypes = [vp, i, dp, dp]
    L.mf6hip_load_factor.argtypes = [vp, dp]
    L.mf6hip_get_system.argtypes = [vp, dp, dp]
    fp = C.POINTER(C.c_float)
    L.mf6hip_unet_create.argtypes = [i, i, i, i, i, ip, C.POINTER(vp)]
    L.mf6hip_unet_destroy.argtypes = [vp]
    L.mf6hip_unet_num_params.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.mf6hip_unet_set_params.argtypes = [vp, fp]
    L.mf6hip_unet_get_params.argtypes = [vp, fp]
    L.mf6hip_unet_forward.argtypes = [vp, fp, fp]
    L.mf6hip_unet_grad.argtypes = [vp, fp, fp, fp, fp]
    L.mf6hip_unet_set_data.argtypes = [vp, i, fp]
    L.mf6hip_unet_train.argtypes = [vp, i, i, ip, C.c_float, fp]
    L.mf6hip_attach_unet.argtypes = [vp, vp, fp, i, i]
    L.mf6hip_last_learned_iterations.argtypes = [vp, ip]
    L.mf6hip_batch_create.argtypes = [i, C.POINTER(vp)]
    L.mf6hip_batch_add.argtypes = [vp, C.POINTER(ImsSettingsC)]
    L.mf6hip_batch_finalize.argtypes = [vp]
    L.mf6hip_batch_set_system.argtypes = [vp, i, ip, ip, i, dp, dp, dp]
    L.mf6hip_batch_autotune.argtypes = [vp, C.c_char_p, i]
    L.mf6hip_batch_solve.argtypes = [vp, ip, ip, dp]
    L.mf6hip_batch_get_x.argtypes = [vp, i, dp]
    L.mf6hip_batch_destroy.argtypes = [vp]
    L.mf6hip_xmi_set_nonlinear.argtypes = [vp, C.POINTER(NonlinearC)]
    L.mf6hip_xmi_prepare_solve.argtypes = [vp, i, i, i, d]
    L.mf6hip_xmi_solve.argtypes = [vp, ip]
    L.mf6hip_xmi_last_iteration.argtypes = [vp, C.POINTER(XmiIterationC)]
    L.mf6hip_xmi_finalize_solve.argtypes = [vp, ip, ip]
    L.mf6hip_csr_residual_norms_noalloc.argtypes = [i, i] + [vp] * 8
    for nm in EXPORTS:
        getattr(L, nm).restype = i
    _LIB = L
    return L


def last_error():
    buf = C.create_string_buffer(1024)
    load().mf6hip_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise Mf6HipError(f"{what} failed (status {rc}): {last_error()}")


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))
