"""Gradients of the fused simulation: forward with checkpoints every K steps and the
hand-written reverse pass (`jxhip_fused_adjoint_*`), wrapped in an autograd Function.

Semantics: the gradient of `integrate`'s recordings with respect to the HH / Leak
parameters and the initial voltages, as autograd through the reference's step would
give (jaxley_pytorch/integrate.py:624-662 with `checkpoint_lengths=[T / K, K]`,
utils/torch_utils.py:141-198). The reverse pass recomputes each K-step segment from its
checkpoint and runs, per step and backwards: the transposed tree solve
(solver_voltage.py:143-187), the adjoint of the linearized currents, and the adjoint of
the exponential-Euler gate update through the rate functions (channels/hh.py:87-108,
solver_gate.py:8-64).
"""
from __future__ import annotations

from typing import Dict, List, Tuple

import torch

from . import _lib
from .fused import ALL_PARAMS, ForwardLaunch, from_prog, to_prog


class _FusedIntegrate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, holder, v0, *param_tensors):
        st, node_f, modes, names, gates0, stim, stim_scale, T, dt, B, cfg, K = holder["args"]
        params = {k: (m, t.detach()) for k, m, t in zip(names, modes, param_tensors)}
        run = ForwardLaunch(st, node_f, params, v0.detach(), tuple(g.detach() for g in gates0),
                            stim, stim_scale, T, dt, B, cfg, ckpt_every=K)
        run.launch()
        holder["run"] = run
        ctx.holder = holder
        ctx.modes = modes
        ctx.names = names
        return run.recordings()

    @staticmethod
    def backward(ctx, g_rec):
        holder = ctx.holder
        run: ForwardLaunch = holder["run"]
        st, K = run.st, run.ckpt_every
        P, S, n, T = run.P, run.S, st.n, run.T
        dev, dtype = run.v.device, run.dtype
        # (B, n_rec, T+1) -> program layout (P, T+1, n_rec, S)
        gr = g_rec.to(dtype)
        if P * S != run.B:
            gr = torch.cat([gr, gr.new_zeros(P * S - run.B, *gr.shape[1:])], 0)
        gr = gr.reshape(P, S, st.n_rec, T + 1).permute(0, 3, 2, 1).contiguous()
        work = [torch.empty((P, K + 1, n, S), dtype=torch.float64, device=dev)]
        work += [torch.empty((P, K + 1, n, S), dtype=dtype, device=dev) for _ in range(3)]
        lam = [torch.zeros((P, n, S), dtype=dtype, device=dev) for _ in range(3)]
        g_v0 = torch.zeros((P, n, S), dtype=dtype, device=dev)
        gpar = {}
        adj = _lib.AdjointArgs()
        adj.g_rec = gr.data_ptr()
        for j, name in enumerate(ALL_PARAMS):
            if name in ctx.names and ctx.needs_input_grad[2 + ctx.names.index(name)]:
                gpar[name] = torch.zeros((P, n, S), dtype=dtype, device=dev)
                adj.g_param[j] = gpar[name].data_ptr()
            else:
                adj.g_param[j] = 0
        adj.g_v0 = g_v0.data_ptr()
        adj.work_v, adj.work_m, adj.work_h, adj.work_n = (w.data_ptr() for w in work)
        adj.lam_m, adj.lam_h, adj.lam_n = (l.data_ptr() for l in lam)
        adj.work_stride = (K + 1) * n * S
        adj.n_steps = T
        args = run.args
        saved_lds = args.lds_bytes
        args.lds_bytes = 3 * n * S * run.node_f.element_size()
        f = _lib.fn("jxhip_fused_adjoint", dtype)
        try:
            _lib.check(f(args, adj, _lib.stream_ptr(dev)), "fused_adjoint")
        finally:
            args.lds_bytes = saved_lds
        grads: List = []
        for name, mode in zip(ctx.names, ctx.modes):
            gp = gpar.get(name)
            if gp is None:
                grads.append(None)
            elif mode == 0:
                grads.append(gp.sum(dim=(0, 2)))          # shared per node: sum over sims
            else:
                grads.append(gp.reshape(-1))              # per (node, sim): program layout
        gv = from_prog(g_v0, run.B) if ctx.needs_input_grad[1] else None
        holder.pop("run", None)
        return (None, gv, *grads)


def fused_integrate_with_adjoint(st, node_f, params: Dict[str, Tuple[int, torch.Tensor]], param_inputs,
                                 v0, gates0, stim, stim_scale, T, dt, B, cfg, K, batch, comp_to_solve):
    """Differentiable fused simulation: returns `(recordings (B, n_rec, T+1), run)`."""
    if any(g.requires_grad for g in gates0):
        raise NotImplementedError("gradients with respect to initial gate states are not supported")
    K = int(max(1, min(K, T)))
    names = [k for k in ALL_PARAMS if k in params]
    modes = [params[k][0] for k in names]
    tensors = [params[k][1] for k in names]
    holder = {"args": (st, node_f, modes, names, gates0, stim, stim_scale, T, dt, B, cfg, K)}
    recs = _FusedIntegrate.apply(holder, v0, *tensors)
    return recs, holder.get("run")
