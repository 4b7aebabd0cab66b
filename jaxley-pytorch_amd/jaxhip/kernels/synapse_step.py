"""Fused synapse step of one synapse type in a regular layout: the drop-in for
jaxley-pytorch's `jaxley_pytorch::synapse_step` / `::synapse_step_backward` custom ops
(kernels/synapse_step.py:336-448) and their plan (`kernel_plan`, :79-103).

* `regular_layout(post)`: the number k of synapses per target when the synapses are
  sorted by target, each target holds the same k and targets strictly increase
  (syn_utils.py:13-27), else `None`.
* `kernel_plan(model, pre, post, n_nodes, device)`: int32 index tensors on the device
  (presynaptic node per synapse, target node per target, transposed presynaptic index
  for the deterministic backward), or `None` when the layout is not regular.
* `synapse_step(plan, v, s, params, area, dt, diff)`: differentiable; returns
  `(s_new, i_syn, lin, cst)` with, per simulation b and target t, `lin` = sum over the
  target's synapses of the finite-difference slope of i / area * 1e5 (both voltages
  perturbed by `diff`, modules/network.py:452-478) and `cst` = sum of (i / area * 1e5 -
  slope * v_post); `s_new` the updated states (KineticSynapse; SigmoidRateSynapse has
  none and returns `s`), `i_syn` the current of every synapse in nA.

Kernels: csrc/jxhip_synapse.hip (one thread per (simulation, target) forward; one
thread per synapse backward with the batch loop inside, presynaptic gradients summed
through the transposed index, no atomics).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import ctypes as C
import numpy as np
import torch

from .. import _lib

MODELS = {"KineticSynapse": 0, "SigmoidRateSynapse": 1}
PARAM_NAMES = {0: ("gS", "e_syn", "alpha", "beta", "v_th", "delta"), 1: ("gS", "x_offset", "slope")}


def regular_layout(post: np.ndarray) -> Optional[int]:
    post = np.asarray(post)
    if len(post) == 0:
        return None
    targets, counts = np.unique(post, return_counts=True)
    k = int(counts[0])
    if not np.all(counts == k):
        return None
    if not np.array_equal(post, np.repeat(targets, k)):
        return None
    return k


def kernel_plan(model, pre: np.ndarray, post: np.ndarray, n_nodes: int, device) -> Optional[Dict]:
    """Index plan of a synapse group (`model`: a synapse type name or 0 / 1)."""
    model = MODELS.get(getattr(model, "name", model), model)
    k = regular_layout(post)
    if k is None:
        return None
    pre = np.asarray(pre, dtype=np.int64)
    targets = np.asarray(post, dtype=np.int64)[::k]
    targets_are_nodes = len(targets) == n_nodes and np.array_equal(targets, np.arange(n_nodes))
    pre_nodes, counts = np.unique(pre, return_counts=True)
    width = int(counts.max()) if len(counts) else 1
    inv = -np.ones((len(pre_nodes), width), dtype=np.int64)
    fill = np.zeros(len(pre_nodes), dtype=np.int64)
    slot = np.searchsorted(pre_nodes, pre)
    for e, j in enumerate(slot):
        inv[j, fill[j]] = e
        fill[j] += 1
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.int32, device=device)
    return {"model": int(model), "k": k, "pre": t(pre), "targets": t(targets),
            "targets_are_nodes": bool(targets_are_nodes), "pre_nodes": t(pre_nodes),
            "inv": t(inv), "inv_width": width, "n_nodes": int(n_nodes), "n_syn": len(pre)}


def _param_ptrs(params: Sequence[torch.Tensor]):
    """Host array of the six parameter device pointers (unused slots repeat the last)."""
    arr = (C.c_void_p * 6)()
    for q in range(6):
        arr[q] = params[min(q, len(params) - 1)].data_ptr()
    return arr


def _forward_impl(v, s, area, params, pre, targets, k: int, model: int, targets_are_nodes: bool,
                  dt: float, diff: float):
    B, n_nodes = v.shape
    E = pre.shape[0]
    T = E // k
    dev, dtype = v.device, v.dtype
    v, s, area = v.contiguous(), s.contiguous(), area.contiguous()
    params = [p.contiguous() for p in params]
    s_new = torch.empty((B, E), dtype=dtype, device=dev)
    i_syn = torch.empty((B, E), dtype=dtype, device=dev)
    lin = torch.empty((B, T), dtype=dtype, device=dev)
    cst = torch.empty((B, T), dtype=dtype, device=dev)
    f = _lib.fn("jxhip_synapse_step", dtype)
    pp = _param_ptrs(params)
    _lib.check(f(v.data_ptr(), s.data_ptr(), pre.data_ptr(), targets.data_ptr(),
                 C.cast(pp, C.c_void_p), area.data_ptr(), dt, diff, s_new.data_ptr(), i_syn.data_ptr(),
                 lin.data_ptr(), cst.data_ptr(), B, n_nodes, E, k, model,
                 int(targets_are_nodes), _lib.stream_ptr(dev)), "synapse_step")
    if model == 1:
        s_new.copy_(s)
    return s_new, i_syn, lin, cst


@torch.library.custom_op("jaxhip::synapse_step", mutates_args=())
def synapse_step_op(v: torch.Tensor, s: torch.Tensor, area: torch.Tensor, params: List[torch.Tensor],
                    pre: torch.Tensor, targets: torch.Tensor, pre_nodes: torch.Tensor,
                    inv: torch.Tensor, k: int, model: int, targets_are_nodes: bool, dt: float,
                    diff: float) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Forward kernel (K3): `(s_new, i_syn, lin, cst)` of a regular synapse group. The
    plan's index tensors are operator inputs (`pre_nodes`, `inv` serve the backward), so
    a traced graph holds no tensor constants."""
    return _forward_impl(v, s, area, params, pre, targets, k, model, targets_are_nodes, dt, diff)


@synapse_step_op.register_fake
def _synapse_step_fake(v, s, area, params, pre, targets, pre_nodes, inv, k, model, targets_are_nodes,
                       dt, diff):
    B, E = v.shape[0], pre.shape[0]
    return v.new_empty((B, E)), v.new_empty((B, E)), v.new_empty((B, E // k)), v.new_empty((B, E // k))


def _backward_impl(v, s, area, params, pre, targets, pre_nodes, inv, k: int, model: int,
                   targets_are_nodes: bool, dt: float, diff: float, g_sn, g_i, g_lin, g_cst):
    B, n_nodes = v.shape
    E = pre.shape[0]
    dev, dtype = v.device, v.dtype
    g_v = torch.empty((B, n_nodes), dtype=dtype, device=dev)
    g_s = torch.zeros((B, E), dtype=dtype, device=dev)
    g_p = torch.empty((len(params), E), dtype=dtype, device=dev)
    work = torch.empty(2 * E * B, dtype=dtype, device=dev)
    gsn = None if (g_sn is None or model == 1) else g_sn.contiguous()
    gi = None if g_i is None else g_i.contiguous()
    f = _lib.fn("jxhip_synapse_step_backward", dtype)
    pp = _param_ptrs([p.contiguous() for p in params])
    _lib.check(f(v.contiguous().data_ptr(), s.contiguous().data_ptr(), pre.data_ptr(), targets.data_ptr(),
                 C.cast(pp, C.c_void_p), area.contiguous().data_ptr(), dt, diff, _lib.ptr(gsn),
                 _lib.ptr(gi), g_lin.contiguous().data_ptr(), g_cst.contiguous().data_ptr(),
                 g_v.data_ptr(), g_s.data_ptr(), g_p.data_ptr(), work.data_ptr(), pre_nodes.data_ptr(),
                 inv.data_ptr(), pre_nodes.numel(), inv.shape[1], B, n_nodes, E, k, model,
                 int(targets_are_nodes), _lib.stream_ptr(dev)), "synapse_step_backward")
    if model == 1 and g_sn is not None:
        g_s = g_s + g_sn               # s passes through unchanged
    return g_v, g_s, g_p


@torch.library.custom_op("jaxhip::synapse_step_backward", mutates_args=())
def synapse_step_backward_op(v: torch.Tensor, s: torch.Tensor, area: torch.Tensor,
                             params: List[torch.Tensor], pre: torch.Tensor, targets: torch.Tensor,
                             pre_nodes: torch.Tensor, inv: torch.Tensor, k: int, model: int,
                             targets_are_nodes: bool, dt: float, diff: float,
                             g_sn: Optional[torch.Tensor], g_i: Optional[torch.Tensor],
                             g_lin: torch.Tensor,
                             g_cst: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Backward kernels (K4-K5): `(g_v, g_s, g_params)`, g_params `(n_params, E)`."""
    return _backward_impl(v, s, area, params, pre, targets, pre_nodes, inv, k, model,
                          targets_are_nodes, dt, diff, g_sn, g_i, g_lin, g_cst)


@synapse_step_backward_op.register_fake
def _synapse_step_backward_fake(v, s, area, params, pre, targets, pre_nodes, inv, k, model,
                                targets_are_nodes, dt, diff, g_sn, g_i, g_lin, g_cst):
    B, E = v.shape[0], pre.shape[0]
    return v.new_empty(v.shape), v.new_empty((B, E)), v.new_empty((len(params), E))


def _synapse_setup(ctx, inputs, output):
    v, s, area, params, pre, targets, pre_nodes, inv, k, model, targets_are_nodes, dt, diff = inputs
    ctx.save_for_backward(v, s, area, pre, targets, pre_nodes, inv, *params)
    ctx.args = (k, model, targets_are_nodes, dt, diff)
    ctx.n_params = len(params)


def _synapse_backward(ctx, g_sn, g_i, g_lin, g_cst):
    v, s, area, pre, targets, pre_nodes, inv, *params = ctx.saved_tensors
    k, model, targets_are_nodes, dt, diff = ctx.args
    B, E = v.shape[0], pre.shape[0]
    z = lambda *shape: v.new_zeros(shape)
    g_lin = z(B, E // k) if g_lin is None else g_lin
    g_cst = z(B, E // k) if g_cst is None else g_cst
    g_v, g_s, g_p = synapse_step_backward_op(v, s, area, list(params), pre, targets, pre_nodes, inv,
                                             k, model, targets_are_nodes, dt, diff, g_sn, g_i,
                                             g_lin, g_cst)
    return (g_v, g_s, None, [g_p[q] for q in range(ctx.n_params)], None, None, None, None, None,
            None, None, None, None)


synapse_step_op.register_autograd(_synapse_backward, setup_context=_synapse_setup)


class _SynapseStep(torch.autograd.Function):
    """The same forward and backward kernels without the operator dispatch (the eager
    per-step path calls this once per step and synapse type)."""

    @staticmethod
    def forward(ctx, v, s, area, plan, dt, diff, *params):
        out = _forward_impl(v, s, area, list(params), plan["pre"], plan["targets"], plan["k"],
                            plan["model"], bool(plan["targets_are_nodes"]), dt, diff)
        ctx.save_for_backward(v, s, area, *params)
        ctx.plan, ctx.dt, ctx.diff = plan, dt, diff
        return out

    @staticmethod
    def backward(ctx, g_sn, g_i, g_lin, g_cst):
        v, s, area, *params = ctx.saved_tensors
        plan = ctx.plan
        B, E, k = v.shape[0], plan["n_syn"], plan["k"]
        g_lin = v.new_zeros((B, E // k)) if g_lin is None else g_lin
        g_cst = v.new_zeros((B, E // k)) if g_cst is None else g_cst
        g_v, g_s, g_p = _backward_impl(v, s, area, params, plan["pre"], plan["targets"], plan["pre_nodes"],
                                       plan["inv"], k, plan["model"], bool(plan["targets_are_nodes"]),
                                       ctx.dt, ctx.diff, g_sn, g_i, g_lin, g_cst)
        return (g_v, g_s, None, None, None, None) + tuple(g_p[q] for q in range(len(params)))


def synapse_step(plan: Dict, v: torch.Tensor, s: torch.Tensor, params: Sequence[torch.Tensor],
                 area: torch.Tensor, dt: float, diff: float = 1e-3):
    """One synapse step of a regular synapse group; `v` (B, n_nodes), `s` (B, E),
    `params` the model's parameters in `PARAM_NAMES` order, each (E,), `area` (E,) the
    postsynaptic compartment area (um^2). See the module docstring for the outputs."""
    if v.device.type != "cuda":
        raise RuntimeError("jaxhip.kernels.synapse_step runs on the GPU")
    names = PARAM_NAMES[plan["model"]]
    if len(params) != len(names):
        raise ValueError(f"expected parameters {names}")
    if torch.compiler.is_compiling():
        return synapse_step_op(v, s, area, list(params), plan["pre"], plan["targets"], plan["pre_nodes"],
                               plan["inv"], plan["k"], plan["model"], bool(plan["targets_are_nodes"]),
                               float(dt), float(diff))
    return _SynapseStep.apply(v, s, area, plan, float(dt), float(diff), *params)
