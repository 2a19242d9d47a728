"""MI355X-native closed-shell CCSD(T) triples energy (ExaChem fused (T) path).

The product is the C-ABI library ``exachem_amd/_lib/libexachem_t.so`` (host
C++ + gfx950 HIP kernels, declared in ``include/exachem_t.h``).  This package
is a thin ctypes front end used by tests and the benchmark; it never computes
anything itself and raises if the native library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libexachem_t.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "exachem_t.h")

XT_OK, XT_EINVAL, XT_ENOMEM, XT_EHIP, XT_ESTATE, XT_ENODEV, XT_ERANGE = 0, -1, -2, -3, -4, -5, -6
PREC_FP64, PREC_MIXED = 0, 1
SCHED_DRIVER, SCHED_ANALYTIC, SCHED_LEARNED = 0, 1, 2

_lib = None


class XtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"xt error {code}: {msg}")
        self.code = code


def build(force: bool = False) -> str:
    """Compile the HIP library in-tree (hipcc, gfx950)."""
    args = ["make", "-s", "-C", os.path.join(_HERE, "csrc")]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)
    return LIB_PATH


class XtOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("schedule", C.c_int), ("use_graph", C.c_int),
                ("record_task_time", C.c_int), ("workgroups", C.c_int),
                ("progress_ms", C.c_int), ("reserved", C.c_int * 2)]


class LearnStats(C.Structure):
    _fields_ = [("n_sample", C.c_long), ("final_loss", C.c_double),
                ("fit_rmse_model", C.c_double), ("fit_rmse_flops", C.c_double),
                ("calib_ms", C.c_double), ("train_ms", C.c_double)]


MLP_IN, MLP_HIDDEN, MLP_PARAMS = 7, 64, 4737


def lib():
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    ip, dp, lp, vp = C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_long), C.c_void_p
    sz = C.c_size_t
    L.xt_last_error.restype = C.c_char_p
    L.xt_version.restype = C.c_char_p
    L.xt_device_available.restype = C.c_int
    L.xt_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.xt_destroy.argtypes = [vp]
    L.xt_set_tiling.argtypes = [vp, C.c_int, C.c_int, C.c_int, ip, C.c_int, ip, C.c_int]
    L.xt_num_tasks.argtypes = [vp, lp]
    L.xt_task_list.argtypes = [vp, lp, dp]
    L.xt_flops.argtypes = [vp, dp, dp, dp]
    L.xt_upload_closed_shell.argtypes = [vp] + [dp] * 7
    L.xt_synth_upload.argtypes = [vp, C.c_uint64, C.c_double, C.c_double]
    L.xt_synth_closed_shell.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_double,
                                        C.c_double] + [dp] * 7
    L.xt_default_opts.argtypes = [C.POINTER(XtOpts)]
    L.xt_default_opts.restype = None
    L.xt_run.argtypes = [vp, C.POINTER(XtOpts), dp, dp, C.POINTER(C.c_float)]
    L.xt_last_kernel_ms.argtypes = [vp, dp]
    L.xt_run_tasks.argtypes = [vp, C.POINTER(XtOpts), C.c_long, lp, dp, dp, C.POINTER(C.c_float)]
    L.xt_sum_task_energies.argtypes = [C.c_long, dp, dp]
    L.xt_memory_required.argtypes = [vp, C.c_long, dp]
    L.xt_upload_peak_bytes.argtypes = [vp, dp]
    L.xt_task_codes.argtypes = [vp, C.c_long, C.POINTER(C.c_uint64)]
    llp = C.POINTER(C.c_longlong)
    L.xt_queue_layout.argtypes = [vp, C.POINTER(XtOpts), C.c_long, lp, ip, ip, llp, llp, ip, ip]
    L.xt_prepare.argtypes = [vp, C.c_int]
    L.xt_set_group_shape.argtypes = [vp, C.c_int, ip]
    L.xt_get_group_shape.argtypes = [vp, C.c_int, ip]
    L.xt_autotune.argtypes = [vp, C.POINTER(XtOpts), C.c_long, ip, dp, ip]
    fp = C.POINTER(C.c_float)
    L.xt_mlp_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
    L.xt_mlp_destroy.argtypes = [vp]
    L.xt_mlp_get_params.argtypes = [vp, fp]
    L.xt_mlp_set_params.argtypes = [vp, fp]
    L.xt_mlp_loss_grad.argtypes = [vp, C.c_long, fp, fp, fp, fp]
    L.xt_mlp_train.argtypes = [vp, C.c_long, fp, fp, C.c_int, C.c_float, fp]
    L.xt_mlp_predict.argtypes = [vp, C.c_long, fp, fp]
    L.xt_task_features.argtypes = [vp, C.c_int, fp]
    L.xt_set_task_cost.argtypes = [vp, dp]
    L.xt_learn_schedule.argtypes = [vp, vp, C.POINTER(XtOpts), C.c_long, C.c_int, C.c_int,
                                    C.c_float, C.POINTER(LearnStats)]
    L.xt_upload_cholesky.argtypes = [vp, dp, dp, dp, dp, C.c_long, dp]
    L.xt_expand_amplitudes.argtypes = [vp, vp, vp, vp]
    L.xt_gather_task.argtypes = [vp, lp, vp] + [vp] * 6 + [ip] * 7
    L.xt_packed_sizes.argtypes = [vp, lp]
    L.xt_gather_last_elems.argtypes = [vp, lp]
    L.xt_fully_fused_ccsd_t_gpu.argtypes = (
        [vp, sz] + [sz] * 6 + [vp] * 6 + [ip] * 6 + [sz] * 8 + [vp] * 6 + [vp, vp])
    L.xt_fully_fused_ccsd_t_gpu_noalloc.argtypes = L.xt_fully_fused_ccsd_t_gpu.argtypes
    for name in ("xt_create", "xt_destroy", "xt_set_tiling", "xt_num_tasks", "xt_task_list",
                 "xt_flops", "xt_upload_closed_shell", "xt_synth_upload", "xt_synth_closed_shell",
                 "xt_run", "xt_last_kernel_ms", "xt_fully_fused_ccsd_t_gpu",
                 "xt_fully_fused_ccsd_t_gpu_noalloc", "xt_run_tasks",
                 "xt_sum_task_energies", "xt_memory_required", "xt_upload_peak_bytes",
                 "xt_task_codes", "xt_queue_layout", "xt_prepare", "xt_mlp_create", "xt_mlp_destroy",
                 "xt_mlp_get_params", "xt_mlp_set_params", "xt_mlp_loss_grad", "xt_mlp_train",
                 "xt_mlp_predict", "xt_task_features", "xt_set_task_cost", "xt_learn_schedule",
                 "xt_set_group_shape", "xt_get_group_shape", "xt_autotune", "xt_gather_task",
                 "xt_packed_sizes", "xt_gather_last_elems", "xt_upload_cholesky",
                 "xt_expand_amplitudes"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc):
    if rc != XT_OK:
        raise XtError(rc, lib().xt_last_error().decode())
    return rc


def header_symbols(path: str = HEADER):
    """Function names declared in the C header."""
    import re
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(xt_\w+)\s*\(", text, re.M)))


from .triples import (  # noqa: E402
    CCSDTriples, CostModel, Tiling, default_opts, fully_fused_ccsd_t_gpu, sum_task_energies,
    synth_closed_shell, uneven_tiles)

__all__ = ["CCSDTriples", "CostModel", "Tiling", "default_opts", "fully_fused_ccsd_t_gpu",
           "sum_task_energies", "synth_closed_shell", "uneven_tiles", "lib", "build", "XtError",
           "XtOpts", "LearnStats"]
