"""ctypes binding of the C ABI in include/eplace.h and include/eplace_gen.h.

The product path is libeplace.so (hand-written HIP kernels for gfx950).  If the
library is missing this module raises at import time: there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeplace.so")

EP_OK, EP_EINVAL, EP_EDEVICE, EP_ENAN, EP_ENOMEM, EP_ESTATE, EP_ENOTSUP, EP_EIO = range(8)
EP_PREC_FP32, EP_PREC_BF16X3, EP_PREC_FP16X3 = 0, 1, 2
EP_GEN_MESH, EP_GEN_STDCELL = 0, 1
EP_TR_NONE, EP_TR_DCT, EP_TR_IDCT, EP_TR_IDXST = -1, 0, 1, 2
EP_WA_PACKET_NOALLOC_MAX = 1 << 20
EP_WA_EXP_HW, EP_WA_EXP_FAST64 = 0, 1

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


class EpNetlist(C.Structure):
    _fields_ = [
        ("n_mov", C.c_int32), ("n_fix", C.c_int32),
        ("w", _fp), ("h", _fp), ("x", _fp), ("y", _fp),
        ("n_nets", C.c_int32), ("n_pins", C.c_int32),
        ("net_start", _ip), ("pin_cell", _ip), ("pin_dx", _fp), ("pin_dy", _fp),
        ("xl", C.c_float), ("yl", C.c_float), ("xh", C.c_float), ("yh", C.c_float),
    ]


class EpParams(C.Structure):
    _fields_ = [
        ("bins_x", C.c_int32), ("bins_y", C.c_int32), ("target_density", C.c_float),
        ("gamma_bins", C.c_float), ("gamma_schedule", C.c_int32), ("lambda_eta", C.c_float),
        ("hpwl_ref", C.c_double), ("mu_max", C.c_float), ("mu_min", C.c_float),
        ("max_backtrack", C.c_int32), ("stop_overflow", C.c_float), ("max_iters", C.c_int32),
        ("init_disp_bins", C.c_float), ("use_fillers", C.c_int32), ("filler_seed", C.c_uint64),
        ("precision", C.c_int32), ("use_graph", C.c_int32), ("learned_step", C.c_int32),
        ("fork_min_cells", C.c_int32), ("cell_order", C.c_int32),
    ]


class EpStepInfo(C.Structure):
    _fields_ = [
        ("iteration", C.c_int32), ("evals", C.c_int32), ("learned_used", C.c_int32),
        ("learned_accepted", C.c_int32), ("hpwl", C.c_double), ("overflow", C.c_double),
        ("lambda_", C.c_double), ("gamma", C.c_double), ("alpha", C.c_double), ("wa", C.c_double),
    ]


class EpRunInfo(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("evals", C.c_int64), ("learned_proposals", C.c_int32),
        ("learned_accepted", C.c_int32), ("hpwl", C.c_double), ("overflow", C.c_double),
        ("seconds", C.c_double),
    ]


class EpKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int32), ("avg_us", C.c_double),
                ("bytes", C.c_double), ("flops", C.c_double)]


class EpGenParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("n_movable", C.c_int32), ("utilisation", C.c_double),
        ("macro_area_frac", C.c_double), ("macro_min", C.c_int32), ("macro_max", C.c_int32),
        ("site_min", C.c_int32), ("site_max", C.c_int32), ("nets_per_cell", C.c_double),
        ("locality_sigma", C.c_double), ("init_sigma_frac", C.c_double),
        ("shuffle_storage", C.c_int32), ("seed", C.c_uint64),
    ]


class EpDesign(C.Structure):
    _fields_ = [("nl", EpNetlist), ("mesh_side", C.c_int32)]


# every symbol the headers under include/ declare
EXPORTS = [
    "ep_default_params", "ep_create", "ep_plan_fillers", "ep_destroy", "ep_last_error", "ep_get_sizes",
    "ep_get_cells", "ep_get_nets", "ep_set_positions", "ep_get_positions", "ep_wa_grad",
    "ep_get_wa_grad", "ep_density", "ep_get_density", "ep_field", "ep_get_field", "ep_field_error", "ep_gradient",
    "ep_cell_bins", "ep_hpwl", "ep_overflow", "ep_init_optimizer", "ep_step", "ep_run",
    "ep_set_precision", "ep_profile", "ep_take_step_records", "ep_set_step_model", "ep_wa_packet",
    "ep_transform", "ep_wa_packet_noalloc", "ep_wa_packet_exp",
    "ep_gen_default_params", "ep_gen_create", "ep_gen_free",
    "ep_mlp_default_train_params", "ep_mlp_train", "ep_mlp_forward", "ep_mlp_grad",
    "ep_batch_create", "ep_batch_destroy", "ep_batch_reset", "ep_batch_init", "ep_batch_run",
    "ep_batch_results", "ep_batch_sizes", "ep_batch_positions", "ep_batch_gradient",
    "ep_batch_state",
    "ep_lefdef_read", "ep_lefdef_read_dir", "ep_lefdef_free", "ep_lefdef_netlist", "ep_lefdef_get_info",
    "ep_lefdef_write_def", "ep_legalize",
]

MLP_IN, MLP_HID = 16, 64
MLP_NWEIGHTS = 34 + MLP_HID * MLP_IN + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID + 1


class EpMlpTrainParams(C.Structure):
    _fields_ = [("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("seed", C.c_uint64)]


class EpBatchRunInfo(C.Structure):
    _fields_ = [("instances", C.c_int32), ("finished", C.c_int32),
                ("instance_iterations", C.c_int64), ("evals", C.c_int64),
                ("launches", C.c_int32), ("seconds", C.c_double),
                ("density_cells_per_thread", C.c_int32),
                ("transform_k_step", C.c_int32)]


class EpLefdefInfo(C.Structure):
    _fields_ = [("units", C.c_int32), ("n_macros", C.c_int32), ("n_components", C.c_int32),
                ("n_io_pins", C.c_int32), ("n_nets", C.c_int32), ("n_mov", C.c_int32),
                ("n_fix", C.c_int32), ("view_nets", C.c_int32), ("view_pins", C.c_int32),
                ("die", C.c_double * 4)]


class EpLegalInfo(C.Structure):
    _fields_ = [("total_disp", C.c_double), ("max_disp", C.c_double), ("sum_sq_disp", C.c_double),
                ("rows", C.c_int32), ("segments", C.c_int32), ("seconds", C.c_double)]


class EplaceError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[ep status {code}] {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m aieplace_amd.build` "
            "(the HIP library is the only implementation; there is no fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.ep_last_error.restype = C.c_char_p
    L.ep_default_params.argtypes = [C.POINTER(EpParams)]
    L.ep_default_params.restype = None
    L.ep_create.argtypes = [C.POINTER(EpNetlist), C.POINTER(EpParams), C.POINTER(vp)]
    L.ep_plan_fillers.argtypes = [C.POINTER(EpNetlist), C.POINTER(EpParams), _ip, _fp, _fp,
                                  _fp, _fp, C.c_int32]
    L.ep_destroy.argtypes = [vp]
    L.ep_get_sizes.argtypes = [vp, _ip, _ip, _ip, _ip, _ip]
    L.ep_get_cells.argtypes = [vp, _fp, _fp, _fp, _fp]
    L.ep_get_nets.argtypes = [vp, _ip, _ip, _fp, _fp, _ip]
    L.ep_set_positions.argtypes = [vp, _fp, _fp]
    L.ep_get_positions.argtypes = [vp, _fp, _fp]
    L.ep_wa_grad.argtypes = [vp, C.c_float, _dp, _dp]
    L.ep_get_wa_grad.argtypes = [vp, _fp, _fp]
    L.ep_density.argtypes = [vp, _dp]
    L.ep_get_density.argtypes = [vp, _fp, _fp, _fp]
    L.ep_field.argtypes = [vp, C.c_int32]
    L.ep_get_field.argtypes = [vp, _fp, _fp]
    L.ep_field_error.argtypes = [vp, C.c_int32, _dp, _dp]
    L.ep_gradient.argtypes = [vp, C.c_float, C.c_float, _fp, _fp, _fp, _fp, _fp, _fp, _dp, _dp]
    L.ep_cell_bins.argtypes = [vp, C.c_int32, _ip]
    L.ep_hpwl.argtypes = [vp, _dp]
    L.ep_overflow.argtypes = [vp, _dp]
    L.ep_init_optimizer.argtypes = [vp]
    L.ep_step.argtypes = [vp, C.POINTER(EpStepInfo)]
    L.ep_run.argtypes = [vp, C.POINTER(EpRunInfo), C.POINTER(EpStepInfo), C.c_int32]
    L.ep_set_precision.argtypes = [vp, C.c_int32]
    L.ep_profile.argtypes = [vp, C.c_int32, C.POINTER(EpKernelStat), C.c_int32, _ip]
    L.ep_take_step_records.argtypes = [vp, _fp, _fp, C.c_int32, _ip]
    L.ep_set_step_model.argtypes = [vp, _fp, C.c_int32]
    L.ep_wa_packet.argtypes = [_fp, C.c_int32, C.c_float, _fp, C.c_int32]
    L.ep_wa_packet_noalloc.argtypes = [_fp, C.c_int32, C.c_float, _fp, C.c_int32]
    L.ep_wa_packet_exp.argtypes = [_fp, C.c_int32, C.c_float, _fp, C.c_int32, C.c_int32]
    L.ep_transform.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _fp, _fp, C.c_int32]
    L.ep_gen_default_params.argtypes = [C.POINTER(EpGenParams), C.c_int32]
    L.ep_gen_default_params.restype = None
    L.ep_gen_create.argtypes = [C.POINTER(EpGenParams), C.POINTER(C.POINTER(EpDesign))]
    L.ep_gen_free.argtypes = [C.POINTER(EpDesign)]
    L.ep_gen_free.restype = None
    L.ep_mlp_default_train_params.argtypes = [C.c_void_p]
    L.ep_mlp_default_train_params.restype = None
    L.ep_mlp_train.argtypes = [_fp, _fp, C.c_int32, C.c_void_p, _fp, _fp]
    L.ep_mlp_forward.argtypes = [_fp, _fp, C.c_int32, _fp]
    L.ep_mlp_grad.argtypes = [_fp, _fp, _fp, C.c_int32, _fp, _fp]
    L.ep_batch_create.argtypes = [C.POINTER(C.POINTER(EpNetlist)), _ip, C.c_int32,
                                  C.POINTER(EpParams), C.POINTER(vp)]
    L.ep_batch_destroy.argtypes = [vp]
    L.ep_batch_reset.argtypes = [vp]
    L.ep_batch_init.argtypes = [vp]
    L.ep_batch_run.argtypes = [vp, C.c_int32, C.POINTER(EpBatchRunInfo)]
    L.ep_batch_results.argtypes = [vp, _ip, C.POINTER(C.c_int64), _dp, _dp]
    L.ep_batch_sizes.argtypes = [vp, C.c_int32, _ip, _ip, _ip]
    L.ep_batch_positions.argtypes = [vp, C.c_int32, _fp, _fp]
    L.ep_batch_gradient.argtypes = [vp, C.c_int32, _fp, _fp]
    L.ep_batch_state.argtypes = [vp, C.c_int32, _dp, _dp, _dp]
    L.ep_lefdef_read.argtypes = [C.POINTER(C.c_char_p), C.c_int32, C.c_char_p, C.POINTER(vp)]
    L.ep_lefdef_read_dir.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.ep_lefdef_free.argtypes = [vp]
    L.ep_lefdef_netlist.argtypes = [vp, C.POINTER(EpNetlist)]
    L.ep_lefdef_get_info.argtypes = [vp, C.POINTER(EpLefdefInfo), C.c_char_p]
    L.ep_lefdef_write_def.argtypes = [vp, _fp, _fp, C.c_char_p]
    L.ep_legalize.argtypes = [C.POINTER(EpNetlist), _fp, _fp, C.c_float, C.c_float, C.c_int32, _fp, _fp,
                              C.POINTER(EpLegalInfo)]
    return L


lib = _load()


def check(rc: int) -> None:
    if rc != EP_OK:
        msg = lib.ep_last_error()
        raise EplaceError(rc, msg.decode() if msg else "")


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_fp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def default_params(**kw) -> EpParams:
    p = EpParams()
    lib.ep_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p
