"""ctypes binding of ``libxlb_biot.so`` (declared in ``include/xlb_biot.h``).

torch is imported first so that its HIP runtime (soname ``libamdhip64.so.7``) is
the one the kernel library resolves against: one runtime per process.  There is
no CPU fallback: a missing library or a missing GPU raises at first use.
"""
import ctypes
import os

import torch

XB_F64, XB_F32, XB_F64_S32, XB_F64_S16, XB_F32_S16 = 0, 1, 2, 3, 4
XB_MODE_STEP, XB_MODE_SMOOTH, XB_MODE_RESIDUAL = 0, 1, 2
XB_RULE_ZERO = 100
XB_EINVAL, XB_EUNSUPPORTED, XB_EHIP, XB_ENONFINITE = -1, -2, -3, -4

LIB_PATH = os.path.normpath(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "lib", "libxlb_biot.so"))

c_int32, c_int64, c_double, c_void_p = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class BCTable(ctypes.Structure):
    _fields_ = [("nslots", c_int32), ("cell_slot", c_void_p), ("code_pc", c_void_p), ("code_ps", c_void_p),
                ("const_pc", c_void_p), ("const_ps", c_void_p)]


class ElastArgs(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("dtype", c_int32), ("mode", c_int32),
                ("lam", c_double), ("mu", c_double), ("gamma", c_double),
                ("f_in", c_void_p), ("f_out", c_void_p), ("fpc_out", c_void_p),
                ("force_a", c_void_p), ("force_b", c_void_p), ("resid", c_void_p),
                ("ix0", c_int32), ("ix1", c_int32), ("iy0", c_int32), ("iy1", c_int32),
                ("bc", ctypes.POINTER(BCTable)), ("coarse_out", c_void_p), ("norm_partials", c_void_p),
                ("rows", c_int32), ("pad_", c_int32),
                ("div_in", c_void_p), ("div_out", c_void_p), ("nl_scale", c_double),
                ("grad_bc", ctypes.POINTER(BCTable)), ("lam_field", c_void_p), ("mu_field", c_void_p)]


class DarcyArgs(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("dtype", c_int32), ("pad_", c_int32),
                ("omega", c_double), ("omega_field", c_void_p),
                ("f_in", c_void_p), ("f_out", c_void_p), ("fpc_out", c_void_p),
                ("source_a", c_void_p), ("source_b", c_void_p), ("bc", ctypes.POINTER(BCTable)),
                ("rows", c_int32), ("pad2_", c_int32)]


class ElastAdjointArgs(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("dtype", c_int32), ("pad_", c_int32),
                ("lam", c_double), ("mu", c_double), ("lam_field", c_void_p), ("mu_field", c_void_p),
                ("abar_in", c_void_p), ("f_in", c_void_p), ("abar_out", c_void_p), ("abar_add", c_void_p),
                ("grad_lam", c_void_p), ("grad_mu", c_void_p), ("grad_force", c_void_p),
                ("bc", ctypes.POINTER(BCTable))]


class DarcyAdjointArgs(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("dtype", c_int32), ("pad_", c_int32),
                ("omega", c_double), ("omega_field", c_void_p), ("source_a", c_void_p), ("source_b", c_void_p),
                ("abar_in", c_void_p), ("f_in", c_void_p), ("abar_out", c_void_p), ("abar_add", c_void_p),
                ("grad_omega", c_void_p), ("grad_source", c_void_p), ("bc", ctypes.POINTER(BCTable))]


XB_ADJ_SMOOTH, XB_ADJ_RESIDUAL = 1, 2
# rule tables of at most this many slots are staged in LDS (xb_device.h LDS_BC_SLOTS)
LDS_BC_SLOTS = 32
ADJ_SMALL_MAX_CELLS = 65536  # xb_elast_smooth_adjoint_small


class TailLevel(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("ix0", c_int32), ("ix1", c_int32), ("iy0", c_int32),
                ("iy1", c_int32), ("bc", ctypes.POINTER(BCTable)), ("grad_bc", ctypes.POINTER(BCTable)),
                ("f0", c_void_p), ("f1", c_void_p), ("div0", c_void_p), ("div1", c_void_p), ("resid", c_void_p),
                ("scratch", c_void_p)]


class TailPlan(ctypes.Structure):
    _fields_ = [("dtype", c_int32), ("nlevels", c_int32), ("k_smooth", c_int32), ("coupled", c_int32),
                ("lam", c_double), ("mu", c_double), ("gamma", c_double), ("nl_scale", c_double),
                ("table", c_void_p), ("lds_bytes", ctypes.c_int64)]


class ElastMGAdjointArgs(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("dtype", c_int32), ("mode", c_int32),
                ("lam", c_double), ("mu", c_double), ("gamma", c_double), ("lam_field", c_void_p),
                ("mu_field", c_void_p), ("ix0", c_int32), ("ix1", c_int32), ("iy0", c_int32), ("iy1", c_int32),
                ("abar_in", c_void_p), ("abar_add", c_void_p), ("abar_out", c_void_p), ("rbar", c_void_p),
                ("force_bar", c_void_p), ("f_in", c_void_p), ("resid", c_void_p), ("grad_lam", c_void_p),
                ("grad_mu", c_void_p), ("bc", ctypes.POINTER(BCTable))]


XB_BATCH_ELAST_STEP, XB_BATCH_PROLONG_DIV, XB_BATCH_FILL, XB_BATCH_NORM, XB_BATCH_ELAST_SMOOTHN = 1, 2, 3, 4, 5


class ProlongJob(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("cnx", c_int32), ("cny", c_int32),
                ("ix0", c_int32), ("ix1", c_int32), ("iy0", c_int32), ("iy1", c_int32),
                ("lam", c_double), ("mu", c_double), ("f", c_void_p), ("coarse", c_void_p), ("div", c_void_p)]


class FillJob(ctypes.Structure):
    _fields_ = [("n", c_int64), ("out", c_void_p), ("value", c_double)]


class NormJob(ctypes.Structure):
    _fields_ = [("partials", c_void_p), ("count", c_int32), ("pad_", c_int32)]


class BatchPlan(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("dtype", c_int32), ("mode", c_int32), ("flags", c_int32),
                ("njobs", c_int32), ("nblocks", c_int32), ("table", c_void_p), ("table_bytes", c_int64),
                ("jobs_offset", c_int64), ("blocks_offset", c_int64), ("insts_offset", c_int64),
                ("jobinfo_offset", c_int64), ("work_offset", c_int64), ("use_list", c_int32),
                ("list_waves", c_int32)]


class ConvergeState(ctypes.Structure):
    _fields_ = [("r0", c_void_p), ("rlast", c_void_p), ("cycles", c_void_p), ("active", c_void_p),
                ("status", c_void_p), ("nactive", c_void_p), ("rref", c_void_p)]


# name -> (restype, argtypes); every exported symbol of include/xlb_biot.h
SIGNATURES = {
    "xb_abi_version": (c_int32, []),
    "xb_last_error": (ctypes.c_char_p, []),
    "xb_elast_step": (c_int32, [ctypes.POINTER(ElastArgs), c_void_p]),
    "xb_elast_step2": (c_int32, [ctypes.POINTER(ElastArgs), c_void_p]),
    "xb_darcy_step2": (c_int32, [ctypes.POINTER(DarcyArgs), c_void_p]),
    "xb_elast_partials_count": (c_int64, [ctypes.POINTER(ElastArgs)]),
    "xb_elast_prolong_add": (c_int32, [c_int32, c_int32, c_int32, c_void_p, c_void_p, c_int32, c_int32,
                                       c_int32, c_int32, c_int32, c_int32, c_void_p]),
    "xb_elast_prolong_add_div": (c_int32, [c_int32, c_int32, c_int32, c_double, c_double, c_void_p, c_void_p,
                                           c_int32, c_int32, c_int32, c_int32, c_int32, c_int32, c_void_p,
                                           c_void_p, c_void_p, c_void_p]),
    "xb_elast_macro": (c_int32, [c_int32, c_int32, c_int32, c_double, c_double, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "xb_elast_divergence": (c_int32, [c_int32, c_int32, c_int32, c_double, c_double, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_void_p]),
    "xb_elast_equilibrium": (c_int32, [c_int32, c_int32, c_int32, c_double, c_double, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "xb_elast_adjoint": (c_int32, [ctypes.POINTER(ElastAdjointArgs), c_void_p]),
    "xb_darcy_adjoint": (c_int32, [ctypes.POINTER(DarcyAdjointArgs), c_void_p]),
    "xb_elast_mg_adjoint": (c_int32, [ctypes.POINTER(ElastMGAdjointArgs), c_void_p]),
    "xb_darcy_stepn": (c_int32, [ctypes.POINTER(DarcyArgs), c_int32, c_void_p]),
    "xb_elast_stepn": (c_int32, [ctypes.POINTER(ElastArgs), c_int32, c_void_p]),
    "xb_elast_smoothn": (c_int32, [ctypes.POINTER(ElastArgs), c_int32, c_void_p]),
    "xb_vcycle_tail_bytes": (c_int64, [c_int32]),
    "xb_vcycle_tail_prepare": (c_int32, [c_int32, c_double, c_double, c_double, c_int32, c_double, c_int32,
                                         ctypes.POINTER(TailLevel), c_int32, c_void_p, c_int64,
                                         ctypes.POINTER(TailPlan)]),
    "xb_vcycle_tail": (c_int32, [ctypes.POINTER(TailPlan), c_void_p]),
    "xb_elast_smooth_adjoint_small": (c_int32, [ctypes.POINTER(ElastMGAdjointArgs), c_int32, c_double,
                                                ctypes.POINTER(BCTable), c_void_p, c_void_p]),
    "xb_elast_smooth_adjoint_tiles": (c_int32, [ctypes.POINTER(ElastMGAdjointArgs), c_int32, c_double,
                                                ctypes.POINTER(BCTable), ctypes.POINTER(c_void_p), c_void_p,
                                                c_void_p]),
    "xb_elast_coupling_adjoint": (c_int32, [c_int32, c_int32, c_int32, c_double, c_double, c_void_p, c_void_p,
                                            c_void_p, c_double, ctypes.POINTER(BCTable), c_void_p, c_void_p,
                                            c_void_p, c_void_p, c_void_p, c_void_p]),
    "xb_darcy_gradient_adjoint": (c_int32, [c_int32, c_int32, c_int32, c_void_p, c_double, ctypes.POINTER(BCTable),
                                            c_double, c_void_p, c_void_p]),
    "xb_elast_restrict_sum": (c_int32, [c_int32, c_int32, c_int32, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                        c_int32, c_int32, c_int32, c_void_p]),
    "xb_elast_macro_adjoint": (c_int32, [c_int32, c_int32, c_int32, c_double, c_double, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "xb_elast_divergence_adjoint": (c_int32, [c_int32, c_int32, c_int32, c_double, c_double, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p]),
    "xb_darcy_macro_adjoint": (c_int32, [c_int32, c_int32, c_int32, c_double, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "xb_darcy_step": (c_int32, [ctypes.POINTER(DarcyArgs), c_void_p]),
    "xb_darcy_macro": (c_int32, [c_int32, c_int32, c_int32, c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "xb_darcy_gradient": (c_int32, [c_int32, c_int32, c_int32, c_void_p, c_double, ctypes.POINTER(BCTable),
                                    c_void_p, c_double, c_double, c_void_p, c_void_p]),
    "xb_lincomb": (c_int32, [c_int64, c_int32, c_double, c_void_p, c_double, c_void_p, c_void_p, c_void_p]),
    "xb_fill": (c_int32, [c_int64, c_int32, c_double, c_void_p, c_void_p]),
    "xb_sqdiff_partials": (c_int32, [c_int64, c_int32, c_void_p, c_void_p, c_void_p, c_int32,
                                     ctypes.POINTER(c_int32), c_void_p]),
    "xb_norm_finalize": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p]),
    "xb_batch_table_bytes": (c_int64, [c_int32, c_int32, c_void_p, c_int32]),
    "xb_batch_prepare": (c_int32, [c_int32, c_int32, c_void_p, c_void_p, c_int32, c_void_p, c_int64,
                                   ctypes.POINTER(BatchPlan), c_void_p]),
    "xb_batch_launch": (c_int32, [ctypes.POINTER(BatchPlan), c_void_p, c_void_p]),
    "xb_batch_converge": (c_int32, [ctypes.POINTER(BatchPlan), c_int32, ctypes.POINTER(ConvergeState), c_double,
                                    c_int32, c_void_p]),
    "xb_batch_norms": (c_int32, [ctypes.POINTER(BatchPlan), c_void_p, c_void_p]),
    "xb_batch_compact": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p]),
    "xb_elast_residual_noalloc": (c_int32, [ctypes.POINTER(ElastArgs), c_void_p, c_void_p]),
}

# capacity of the static partials of xb_elast_residual_noalloc (XB_STATIC_PARTIALS)
STATIC_PARTIALS = 1 << 17

_lib = None


def load_library(path=LIB_PATH):
    """Load the kernel library (no GPU needed to load it)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError(f"HIP kernel library not built: {path} (run __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def lib():
    """The library, on a machine with a ROCm GPU (raises otherwise)."""
    if not torch.cuda.is_available():
        raise RuntimeError("xlb-biot kernels need a ROCm GPU; none is visible (there is no CPU fallback)")
    return load_library()


def check(status, what):
    if status != 0:
        msg = load_library().xb_last_error()
        msg = msg.decode() if msg else ""
        exc = ValueError if status in (XB_EINVAL, XB_EUNSUPPORTED) else RuntimeError
        raise exc(f"{what} failed ({status}): {msg}")


def ptr(t):
    return None if t is None else c_void_p(t.data_ptr())


def stream():
    return c_void_p(torch.cuda.current_stream().cuda_stream)


def fill(t, value):
    """``t[...] = value`` with the library's fill kernel (graph-capture safe)."""
    check(lib().xb_fill(t.numel(), dtype_code(t.dtype), float(value), ptr(t), stream()), "xb_fill")
    return t


def dtype_code(dtype):
    if dtype == torch.float64:
        return XB_F64
    if dtype == torch.float32:
        return XB_F32
    raise ValueError(f"unsupported dtype {dtype}")
