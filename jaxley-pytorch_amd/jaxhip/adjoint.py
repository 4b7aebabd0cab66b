"""Gradients of the fused simulation: forward with checkpoints every K steps and the
hand-written reverse pass (`jxhip_fused_adjoint_*`), wrapped in an autograd Function.

With a voltage trajectory (`ForwardLaunch.traj`: the streaming forward of a cell keeps
every step's voltages when they fit the memory budget) the reverse pass recomputes only
the gates between checkpoints and reads the voltages from it.

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
from .fused import (ALL_PARAMS, BATCHABLE, ForwardLaunch, adjoint_config, adjoint_lds_bytes, adjoint_threads,
                    from_prog,
                    to_prog)


class _FusedIntegrate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, holder, v0, node_p, bparam, syn_gs):
        st, node_f, bmode, gates0, stim, stim_scale, T, dt, B, cfg, K, _ = holder["args"]
        syn_params, syn_scale, syn_s0 = holder["syn"]
        run = ForwardLaunch(st, node_f, node_p.detach(), None if bparam is None else bparam.detach(),
                            bmode, v0.detach(), tuple(g.detach() for g in gates0), stim, stim_scale,
                            T, dt, B, cfg, None if syn_params is None else [x.detach() for x in syn_params],
                            syn_scale, None if syn_s0 is None else syn_s0.detach(), ckpt_every=K)
        run.launch()
        holder["run"] = run
        ctx.holder = holder
        ctx.bmode = bmode
        return run.recordings()

    @staticmethod
    def backward(ctx, g_rec):
        holder = ctx.holder
        run: ForwardLaunch = holder["run"]
        st, K = run.st, run.ckpt_every
        P, S, n, T = run.P, run.S, st.n, run.T
        dev, dtype = run.v.device, run.dtype
        # (B, n_rec, T+1) -> the layout of the recordings (per group, part blocks)
        gr = run.recordings_to_prog(g_rec.to(dtype))
        # Recomputed trajectories per group of simulations (a split network's parts side
        # by side).
        # (The voltages come from the forward's trajectory when it kept one.)
        work = [torch.empty((P, K + 1, n, S) if run.traj is None else (1,), dtype=torch.float64, device=dev)]
        work += [torch.empty((P, K + 1, n, S), dtype=dtype, device=dev) for _ in range(3)]
        lam = [torch.zeros((P, n, S), dtype=dtype, device=dev) for _ in range(3)]
        g_v0 = torch.zeros((P, n, S), dtype=dtype, device=dev)
        need_p = ctx.needs_input_grad[2] or ctx.needs_input_grad[3]
        # Only the parameters that require gradients get an accumulator (each one is a
        # read-modify-write per item and step in the reverse pass).
        wanted = holder["args"][-1]
        gpar = [torch.zeros((P, n, S), dtype=dtype, device=dev) if need_p and k in wanted else None
                for k in ALL_PARAMS]
        adj = _lib.AdjointArgs()
        adj.g_rec = gr.data_ptr()
        for j in range(len(ALL_PARAMS)):
            adj.g_param[j] = gpar[j].data_ptr() if gpar[j] is not None else 0
        adj.g_v0 = g_v0.data_ptr()
        adj.work_v, adj.work_m, adj.work_h, adj.work_n = (w.data_ptr() for w in work)
        adj.lam_m, adj.lam_h, adj.lam_n = (l.data_ptr() for l in lam)
        adj.work_stride = (K + 1) * n * S
        lam_v = torch.empty((P, n, S), dtype=dtype, device=dev)   # used when not in VGPRs
        adj.lam_v = lam_v.data_ptr()
        adj.rec_first = st.rec_first.data_ptr()
        adj.rec_next = st.rec_next.data_ptr()
        adj.n_steps = T
        g_gs = None
        n_tgt = 0
        if st.syn is not None:
            # Synapses: recomputed states per step, state adjoints, presynaptic shares and
            # the gradient of gS per (synapse, simulation).
            n_syn, n_tgt = st.syn["n_syn"], st.tgt_max
            Pp = P * st.parts                                # programs
            syn_max = st.syn_max
            kinetic = st.syn["model"] == 0
            syn_bufs = {
                "work_s": torch.empty((Pp, K + 1, syn_max, S) if kinetic else (1,), dtype=dtype, device=dev),
                "lam_s": torch.empty((P, n_syn, S) if kinetic else (1,), dtype=dtype, device=dev),
                # Presynaptic shares by step parity (the parts of a split network write
                # each other's slots).
                "work_q": torch.empty((2 if run.coop else 1, P, n_syn, S), dtype=dtype, device=dev),
                "work_lin": torch.empty((Pp, K, n_tgt, S), dtype=dtype, device=dev),
            }
            holder["syn_bufs"] = syn_bufs
            adj.pre_ptr = st.syn["pre_ptr"].data_ptr()
            adj.pre_pos = st.syn["pre_pos"].data_ptr()
            adj.work_s = syn_bufs["work_s"].data_ptr()
            adj.work_syn_stride = (K + 1) * syn_max * S
            adj.lam_s = syn_bufs["lam_s"].data_ptr()
            adj.work_q = syn_bufs["work_q"].data_ptr()
            adj.work_q_stride = P * n_syn * S if run.coop else 0
            adj.work_lin = syn_bufs["work_lin"].data_ptr()
            adj.work_lin_stride = K * n_tgt * S
            if ctx.needs_input_grad[4]:
                g_gs = torch.zeros((P, n_syn, S), dtype=dtype, device=dev)
                adj.g_syn_gs = g_gs.data_ptr()
        args = run.args
        # The reverse pass's own launch shape at the forward's S.
        nmax = st.n_max
        acfg = adjoint_config(nmax, run.B, run.node_f.element_size(), st.sched, S=S, n_tgt=n_tgt)
        saved = (args.lds_bytes, args.threads, args.items_per_thread, args.coef_global)
        args.lds_bytes = adjoint_lds_bytes(nmax, S, run.node_f.element_size(), st.sched, acfg.gcoef, n_tgt)
        args.threads = adjoint_threads(acfg, torch.tensor([], dtype=dtype).element_size())
        args.items_per_thread, args.coef_global = acfg.ipt, int(acfg.gcoef)
        f = run.kernel_name("jxhip_fused_adjoint")
        if run.gsync is not None:
            run.gsync.zero_()
        try:
            _lib.check(f(args, adj, _lib.stream_ptr(dev)), "fused_adjoint")
        finally:
            args.lds_bytes, args.threads, args.items_per_thread, args.coef_global = saved
        run.check_group_sync()
        g_node_p = g_bparam = None
        zero = torch.zeros((P, n, S), dtype=dtype, device=dev)
        gpar = [zero if g is None else g for g in gpar]
        if ctx.needs_input_grad[2]:
            g_node_p = torch.stack([g.sum(dim=(0, 2)) for g in gpar], 1)
        if ctx.needs_input_grad[3]:
            cols = [gpar[ALL_PARAMS.index(k)] for k in BATCHABLE]
            if ctx.bmode == 1:
                g_bparam = torch.stack([c.sum(1).reshape(-1) for c in cols], -1)
            else:
                g_bparam = torch.stack(cols, -1)
        gv = from_prog(g_v0, run.B) if ctx.needs_input_grad[1] else None
        # gS is shared by the simulations: its gradient is the sum over them.
        g_syn = from_prog(g_gs, run.B).sum(0) if g_gs is not None else None
        run.traj = None
        holder.pop("run", None)
        holder.pop("syn_bufs", None)
        return None, gv, g_node_p, g_bparam, g_syn


def fused_integrate_with_adjoint(st, node_f, node_p, bparam, bmode, v0, gates0, stim, stim_scale,
                                 T, dt, B, cfg, K, grad_keys=ALL_PARAMS, syn_params=None, syn_scale=None,
                                 syn_s0=None):
    """Differentiable fused simulation: returns `(recordings (B, n_rec, T+1), run)`.
    `grad_keys`: the channel parameters whose gradients are wanted (others are zero).
    Networks (one Kinetic or SigmoidRate synapse type): `syn_params` in the kernel's
    synapse order; gradients flow to `syn_params[0]` (gS)."""
    if any(g.requires_grad for g in gates0):
        raise NotImplementedError("gradients with respect to initial gate states are not supported")
    K = int(max(1, min(K, T)))
    holder = {"args": (st, node_f, bmode, gates0, stim, stim_scale, T, dt, B, cfg, K, frozenset(grad_keys)),
              "syn": (syn_params, syn_scale, syn_s0)}
    syn_gs = syn_params[0] if syn_params is not None else None
    recs = _FusedIntegrate.apply(holder, v0, node_p, bparam, syn_gs)
    return recs, holder.get("run")
