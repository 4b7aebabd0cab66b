"""ctypes binding of `libjxhip.so` (the C ABI declared in include/jxhip.h).

The library must be built (`jaxhip.build.build()`); loading fails loudly otherwise.
Launches go on PyTorch's current HIP stream; pointers come from `Tensor.data_ptr()`.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import torch

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libjxhip.so"
HOST_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libjxhost.so"

JXHIP_PROG_COLS = 8
JXHIP_TOPO_COLS = 16
JXHIP_NODE_ICOLS = 8
JXHIP_NODE_FCOLS = 4
JXHIP_NPARAM = 8
JXHIP_NSYNPARAM = 6
JXHIP_SYN_MIXED = 4

_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64


class FusedArgs(C.Structure):
    """`jxhip_fused_args` (include/jxhip.h)."""

    _fields_ = [
        ("progs", _p), ("n_progs", _i64), ("topos", _p), ("seg", _p), ("seg_level_ptr", _p),
        ("seg_child", _p),
        ("node_i", _p), ("node_f", _p), ("child_idx", _p), ("stim_col", _p),
        ("stim_scale", _p), ("rec_node", _p),
        ("node_p", _p), ("bparam", _p), ("bparam_mode", _i32),
        ("v", _p), ("gate_m", _p), ("gate_h", _p), ("gate_n", _p),
        ("stim", _p), ("stim_cols", _i64), ("stim_b", _i64),
        ("rec", _p),
        ("syn_model", _i32), ("syn_pre", _p), ("syn_ptr", _p), ("tgt_node", _p),
        ("syn_param", _p * JXHIP_NSYNPARAM),
        ("syn_scale", _p), ("syn_s", _p),
        ("sims_total", _i64), ("dt", C.c_double), ("diff", C.c_double),
        ("S", _i32), ("threads", _i32), ("items_per_thread", _i32), ("lds_bytes", _i64),
        ("ckpt_v", _p), ("ckpt_m", _p), ("ckpt_h", _p), ("ckpt_n", _p), ("ckpt_s", _p),
        ("ckpt_every", _i64), ("ckpt_stride", _i64), ("ckpt_syn_stride", _i64),
        ("max_items", _i64), ("coef_global", _i32), ("work_ax", _p),
        ("dt_v", C.c_double), ("rhs_scale", C.c_double), ("explicit_step", _i32),
        ("syn_kind", _p), ("node_p2", _p), ("gate_ext", _p), ("gate_ext_stride", C.c_int64),
        ("parts", _i32), ("coop", _i32), ("group_nodes", _i64), ("vx", _p), ("vx_stride", _i64),
        ("traj_v", _p), ("traj_steps", _i64), ("node_sym", _p), ("group_sync", _p),
    ]


class AdjointArgs(C.Structure):
    """`jxhip_adjoint_args` (include/jxhip.h)."""

    _fields_ = [
        ("g_rec", _p), ("g_param", _p * JXHIP_NPARAM), ("g_v0", _p),
        ("work_v", _p), ("work_m", _p), ("work_h", _p), ("work_n", _p),
        ("lam_m", _p), ("lam_h", _p), ("lam_n", _p), ("work_stride", _i64), ("n_steps", _i64),
        ("lam_v", _p), ("rec_first", _p), ("rec_next", _p),
        ("pre_ptr", _p), ("pre_pos", _p), ("work_s", _p), ("work_syn_stride", _i64),
        ("lam_s", _p), ("work_q", _p), ("g_syn_gs", _p), ("work_lin", _p),
        ("work_lin_stride", _i64), ("work_q_stride", _i64),
    ]


# Exported symbols and their argument types (kept in sync with include/jxhip.h).
_F = "F"  # floating pointer placeholder
_SIGS = {
    "jxhip_version": ([], C.c_int),
    "jxhip_error_string": ([C.c_int], C.c_char_p),
    "jxhip_tree_solve_levels": ([_p, _p, _p, _p, _p, _p, _i64, _i64, _p, _i32, _i32, _p, _p, _p, _p, _p, _p], C.c_int),
    "jxhip_tree_solve_regs": ([_p, _p, _p, _p, _p, _i64, _i64, _p, _i32, _i32, _p, _p, _p, _p, _p, _p], C.c_int),
    "jxhip_tree_solve_grads": ([_p, _p, _p, _p, _p, _p, _i64, _i64, _p], C.c_int),
    "jxhip_hh_step": ([_p, _i64, _p, _i64, _p, _p, _p, _p, _p, "dt", _p, _p, _p, _p, _p, _p, _i64, _p], C.c_int),
    "jxhip_synapse_step": ([_p, _p, _p, _p, _p, _p, "dt", "dt", _p, _p, _p, _p, _i64, _i64, _i64, _i64, _i32, _i32, _p], C.c_int),
    "jxhip_synapse_step_backward": ([_p, _p, _p, _p, _p, _p, "dt", "dt", _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _i64, _i64, _i64, _i32, _i32, _p], C.c_int),
    "jxhip_fused_forward": ([C.POINTER(FusedArgs), _p], C.c_int),
    "jxhip_fused_adjoint": ([C.POINTER(FusedArgs), C.POINTER(AdjointArgs), _p], C.c_int),
    "jxhip_fused_forward_mixed": ([C.POINTER(FusedArgs), _p], C.c_int),
    "jxhip_fused_adjoint_mixed": ([C.POINTER(FusedArgs), C.POINTER(AdjointArgs), _p], C.c_int),
    "jxhip_fused_forward_noalloc_f64": ([C.POINTER(FusedArgs), _i64, _p], C.c_int),
}
UNTYPED = ("jxhip_version", "jxhip_error_string", "jxhip_fused_forward_mixed", "jxhip_fused_adjoint_mixed",
           "jxhip_fused_forward_noalloc_f64")
TYPED = {k for k in _SIGS if k not in UNTYPED}


def exported_symbols():
    """All symbol names declared in include/jxhip.h."""
    names = []
    for k in _SIGS:
        if k in TYPED:
            names += [k + "_f32", k + "_f64"]
        else:
            names.append(k)
    return names


_lib: Optional[C.CDLL] = None
_funcs = {}


def load() -> C.CDLL:
    """Load libjxhip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP kernels first (jaxhip.build.build())"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, (args, res) in _SIGS.items():
        variants = [(name + "_f32", C.c_float), (name + "_f64", C.c_double)] if name in TYPED else [(name, None)]
        for full, ft in variants:
            fn = getattr(lib, full)
            fn.argtypes = [ft if a == "dt" else a for a in args]
            fn.restype = res
            _funcs[full] = fn
    _lib = lib
    return lib


def fn(name: str, dtype: Optional[torch.dtype] = None):
    load()
    if dtype is None:
        return _funcs[name]
    if dtype == torch.float32:
        return _funcs[name + "_f32"]
    if dtype == torch.float64:
        return _funcs[name + "_f64"]
    raise TypeError(f"jaxhip kernels support float32 and float64, not {dtype}")


def check(code: int, what: str = "") -> None:
    if code != 0:
        msg = _funcs["jxhip_error_string"](code)
        raise RuntimeError(f"jxhip {what} failed: {msg.decode() if msg else code} ({code})")


def stream_ptr(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: Optional[torch.Tensor]) -> int:
    return 0 if t is None else t.data_ptr()


def require_gpu(t: torch.Tensor, what: str) -> None:
    if not t.is_cuda:
        raise RuntimeError(
            f"{what}: the jaxhip kernels run on the GPU only (got a tensor on {t.device})"
        )


# Host library (include/jxhip_host.h): schedule construction, no GPU runtime.
_HOST_SIGS = {
    "jxhost_bfs_edge_hops": [_i64, _i64, _p, _p, _i64, _i64, _p, _p, _p, _p],
    "jxhost_segments": [_i64, _p, _p, _p, _p, _p, _p],
}
_host: Optional[C.CDLL] = None


def host() -> C.CDLL:
    """Load libjxhost.so (raises if it has not been built)."""
    global _host
    if _host is not None:
        return _host
    if not HOST_LIB_PATH.exists():
        raise RuntimeError(f"{HOST_LIB_PATH} is missing: build the library first (jaxhip.build.build())")
    lib = C.CDLL(str(HOST_LIB_PATH))
    for name, args in _HOST_SIGS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    _host = lib
    return lib


def host_check(code: int, what: str) -> None:
    if code != 0:
        reason = {1: "invalid argument", 2: "the graph is not connected",
                  3: "not a forest with parents before children"}.get(code, str(code))
        raise ValueError(f"{what}: {reason}")
