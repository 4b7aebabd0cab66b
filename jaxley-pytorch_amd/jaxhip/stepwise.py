"""Differentiable per-step simulation on the GPU, for gradients through networks.

The fused kernels run a whole simulation in one launch, and their hand-written adjoint
covers cells (HH / Leak, no synapses) whose system fits a workgroup's LDS. Gradients
through synapses take this path instead: the reference's eager step (`_step_with_plan`,
modules/base.py:3214-3400, inside the time loop of integrate.py:624-662), one step at a
time with autograd, on the device:

* synapses: the fused synapse-step operator (`kernels.synapse_step`, the HIP kernel
  pair K3 / K4-K5 with its hand-written backward; network.py:350-478) for Kinetic and
  SigmoidRate synapses in the regular layout; the reference's per-edge tensor operations
  for IonotropicSynapse and TestSynapse (network.py:480-578);
* HH, Leak and the Pospischil Na / K / Km channels: elementwise tensor operations, as
  the reference computes them (channels/hh.py:37-108, channels/pospischil.py:39-245,
  solver_gate.py:8-87; the reference has no kernel there); cells with Pospischil
  channels take this path for gradients and for batched Pospischil parameters;
* the implicit system (solver_voltage.py:250-359) and its solve by the HIP tree-solve
  kernel with its transposed-solve backward (`kernels.tree_solve`, K1 / K2).

The time loop is checkpointed every K steps (`checkpoint_lengths[-1]`, default 100):
the backward recomputes each K-step chunk from its start, as the reference's
`nested_checkpoint_scan` does (utils/torch_utils.py:141-198). Forward-mode
differentiation (inputs with tangents of `torch.autograd.forward_ad`) runs the same steps
without checkpoints; the tree-solve operator carries its forward rule.
"""
from __future__ import annotations

from typing import Dict, List, Optional

import numpy as np
import torch
import torch.autograd.forward_ad as fwAD
from torch.utils.checkpoint import checkpoint

from .geometry import axial_conductances, comp_geometry
from .kernels.synapse_step import MODELS, PARAM_NAMES, kernel_plan, synapse_step
from .kernels.tree_solve import fused_tree_solve, make_tree_solve_kernel

DIFF = 1e-3          # mV, finite difference of the synapse linearization (network.py:29)
DEFAULT_CHUNK = 100  # steps per checkpoint when integrate() gets no checkpoint_lengths


def _save_exp(x):
    """exp(min(x, 20)) (solver_gate.py:8-11)."""
    return torch.exp(torch.clamp(x, max=20.0))


def _x_over_expm1(x, y):
    """x / expm1(min(x / y, 20)), y (1 - z/2 + z^2/12) for |z| < 1e-4 (solver_gate.py:14-30);
    the unused branch is evaluated at a safe argument so that its gradient stays finite."""
    z = torch.clamp(x / y, max=20.0)
    small = torch.abs(z) < 1e-4
    z_safe = torch.where(small, torch.ones_like(z), z)
    return torch.where(small, y * (1.0 - z / 2.0 + z * z / 12.0), x / torch.expm1(z_safe))


def _gate(x, dt, alpha, beta):
    """Exponential Euler (solver_gate.py:45-64)."""
    tau = 1.0 / (alpha + beta)
    xinf = alpha * tau
    e = _save_exp(-dt / tau)
    return x * e + xinf * (1.0 - e)


def _hh_gates(m, h, n, v, dt):
    """HH gate update at the old voltage (channels/hh.py:37-50, rates :87-108)."""
    am = 0.1 * _x_over_expm1(-(v + 40.0), 10.0)
    bm = 4.0 * _save_exp(-(v + 65.0) / 18.0)
    ah = 0.07 * _save_exp(-(v + 65.0) / 20.0)
    bh = 1.0 / (_save_exp(-(v + 35.0) / 10.0) + 1.0)
    an = 0.01 * _x_over_expm1(-(v + 55.0), 10.0)
    bn = 0.125 * _save_exp(-(v + 65.0) / 80.0)
    return _gate(m, dt, am, bm), _gate(h, dt, ah, bh), _gate(n, dt, an, bn)


_CHANNEL_PARAMS = {
    "HH": ("HH_gNa", "HH_gK", "HH_gLeak", "HH_eNa", "HH_eK", "HH_eLeak"),
    "Leak": ("Leak_gLeak", "Leak_eLeak"),
    "PNa": ("Na_gNa", "eNa", "vt"), "PK": ("K_gK", "eK", "vt"), "PKm": ("Km_gKm", "Km_taumax", "eK"),
}
_CHANNEL_GATES = {"HH": ("HH_m", "HH_h", "HH_n"), "PNa": ("Na_m", "Na_h"), "PK": ("K_n",), "PKm": ("Km_p",)}
_GATE_ORDER = ("HH", "PNa", "PK", "PKm")


def _pospischil(kid, gates, v, p, dt):
    """Gates at the old voltage, then (conductance, reversal potential) of a Pospischil
    channel (channels/pospischil.py:74-245; solver_gate.py:45-87)."""
    if kid == "PKm":
        vp = v + 35.0
        pinf = 1.0 / (1.0 + _save_exp(-0.1 * vp))
        tau = p["Km_taumax"] / (3.3 * _save_exp(0.05 * vp) + _save_exp(-0.05 * vp))
        e = _save_exp(-1.0 / tau * dt)
        q = gates[0] * e + pinf * (1.0 - e)
        return [q], p["Km_gKm"] * q, p["eK"]
    u = v - p["vt"]
    if kid == "PNa":
        am = 0.32 * _x_over_expm1(-0.25 * (u - 13.0), 1.0) / 0.25
        bm = 0.28 * _x_over_expm1(0.2 * (u - 40.0), 1.0) / 0.2
        ah = 0.128 * _save_exp(-(u - 17.0) / 18.0)
        bh = 4.0 / (_save_exp(-(u - 40.0) / 5.0) + 1.0)
        m, h = _gate(gates[0], dt, am, bm), _gate(gates[1], dt, ah, bh)
        return [m, h], p["Na_gNa"] * m**3 * h, p["eNa"]
    an = 0.032 * _x_over_expm1(-0.2 * (u - 15.0), 1.0) / 0.2
    bn = 0.5 * _save_exp(-(u - 10.0) / 40.0)
    nn = _gate(gates[0], dt, an, bn)
    return [nn], p["K_gK"] * nn**4, p["eK"]


def _explicit_edges(tables, device):
    """Edge index sets of the forward-Euler vector field: all edges, and the edges into
    branchpoints (types 3 and 4) with their sinks and sources."""
    typ = np.asarray(tables["type"])
    sinks = torch.as_tensor(np.asarray(tables["sink"], dtype=np.int64), device=device)
    sources = torch.as_tensor(np.asarray(tables["source"], dtype=np.int64), device=device)
    bp = torch.as_tensor(np.nonzero(np.isin(typ, [3, 4]))[0], device=device)
    return sinks, sources, bp


def _explicit_step(v, lin_t, cst_t, g, edges, dt):
    """v + dt f(v) with f = -lin v + const plus the axial currents of every edge, the
    branchpoints at the conductance-weighted mean of their neighbours
    (`_voltage_vectorfield`, solver_voltage.py:438-502)."""
    sinks, sources, bp = edges
    f = -lin_t * v + cst_t
    vv = v
    if len(bp):
        snk, src = sinks[bp], sources[bp]
        num = vv.index_fill(-1, snk, 0.0).index_add(-1, snk, g[:, bp] * vv[:, src])
        den = torch.ones_like(vv).index_fill(-1, snk, 0.0).index_add(-1, snk, g[:, bp])
        vv = num / den
    if len(sinks):
        f = f.index_add(-1, sinks, (vv[:, sources] - vv[:, sinks]) * g)
    return v + dt * f


class _SynGroup:
    """One synapse type of the network in the regular layout of the synapse operator."""

    def __init__(self, edge: Dict, syn_vals: Dict[str, torch.Tensor], area_node: torch.Tensor,
                 n_nodes: int, B: int, device, dtype):
        stype = edge["type"]
        self.name = stype.name
        if self.name not in MODELS:
            raise NotImplementedError(f"{self.name}: the synapse operator runs "
                                      f"{tuple(MODELS)} (gradients through networks)")
        post = np.asarray(edge["post"], dtype=np.int64)
        pre = np.asarray(edge["pre"], dtype=np.int64)
        perm = np.argsort(post, kind="stable")     # the operator takes synapses by target
        self.plan = kernel_plan(self.name, pre[perm], post[perm], n_nodes, device)
        if self.plan is None:
            raise NotImplementedError("gradients through networks need every target to hold "
                                      "the same number of synapses (regular layout)")
        self.model = self.plan["model"]
        self.perm = torch.as_tensor(perm, device=device)
        self.inv = torch.empty_like(self.perm)
        self.inv[self.perm] = torch.arange(len(perm), device=device)
        self.params = []
        for p in PARAM_NAMES[self.model]:
            x = syn_vals[f"{self.name}_{p}"].to(dtype)
            if x.ndim > 1:
                raise NotImplementedError("batched synapse parameters are not supported")
            self.params.append(x[self.perm])
        self.area = area_node[torch.as_tensor(post[perm], device=device)]
        self.targets = self.plan["targets"].long()
        E = len(perm)
        key = f"{self.name}_s"
        if self.model == 0:
            s0 = syn_vals[key].to(dtype)
            self.s0 = torch.broadcast_to(s0, (*s0.shape[:-1], E)).reshape(-1, E)[:, self.perm]
            self.s0 = self.s0.expand(B, E) if self.s0.shape[0] == 1 else self.s0
        else:
            self.s0 = torch.zeros(B, E, dtype=dtype, device=device)
        self.state_key = key if self.model == 0 else None

    def apply(self, v, s, lin, cst, dt):
        s2, i_syn, sl, sc = synapse_step(self.plan, v, s, self.params, self.area, dt, DIFF)
        self.last_current = i_syn          # nA per synapse (operator order)
        return s2, lin.index_add(1, self.targets, sl), cst.index_add(1, self.targets, -sc)


def _ionotropic_update(s, pre_v, p, dt):
    """synapses/ionotropic.py:44-66."""
    u = (pre_v - p["v_th"]) / p["delta"]
    s_inf = 1.0 / (1.0 + _save_exp(-u))
    rate = p["k_minus"] * (1.0 + torch.exp(torch.clamp(u, -20.0, 50.0)))
    e = _save_exp(-rate * dt)
    return s * e + s_inf * (1.0 - e)


def _test_update(c, pre_v, p, dt):
    """synapses/test.py:25-48 (v_th = -35, delta = 10, k_minus = 1/40)."""
    s_bar = 1.0 / (1.0 + _save_exp((-35.0 - pre_v) / 10.0))
    tau = (1.0 - s_bar) / (1.0 / 40.0)
    e = _save_exp((-1.0 / tau) * dt)
    return c * e + s_bar * (1.0 - e)


# update(state, v_pre, params, dt) and current(state, v_pre, v_post, params) in nA of the
# synapse types without an operator kernel (synapses/ionotropic.py, synapses/test.py).
def _kinetic_update(s, pre_v, p, dt):
    """synapses/kinetic.py:37-53."""
    a = p["alpha"] * torch.sigmoid((pre_v - p["v_th"]) / p["delta"])
    rate = a + p["beta"]
    s_inf = a / rate
    return s_inf + (s - s_inf) * torch.exp(-dt * rate)


_TORCH_SYNAPSES = {
    "KineticSynapse": (_kinetic_update, lambda s, vp, vq, p: p["gS"] * s * (vq - p["e_syn"])),
    "SigmoidRateSynapse": (lambda s, vp, p, dt: s,
                           lambda s, vp, vq, p: -p["gS"] * torch.sigmoid((vp - p["x_offset"]) * p["slope"])),
    "IonotropicSynapse": (_ionotropic_update, lambda s, vp, vq, p: p["gS"] * s * (vq - p["e_syn"])),
    "TestSynapse": (_test_update, lambda s, vp, vq, p: p["gC"] * s * vq),
}


class _TorchSynGroup:
    """A synapse type without an operator kernel (IonotropicSynapse, TestSynapse), or in
    a layout the operator does not take: the reference's per-edge step in tensor
    operations (network.py:350-478, 480-578): state update, current at the unperturbed and
    at the perturbed voltages (both ends moved by 1e-3 mV), slope and offset added to the
    postsynaptic nodes."""

    def __init__(self, edge: Dict, syn_vals: Dict[str, torch.Tensor], area_node: torch.Tensor,
                 B: int, device, dtype):
        stype = edge["type"]
        self.name = stype.name
        if self.name not in _TORCH_SYNAPSES:
            raise NotImplementedError(f"{self.name}: no differentiable step on the jaxhip path")
        self.update, self.current = _TORCH_SYNAPSES[self.name]
        self.pre = torch.as_tensor(np.asarray(edge["pre"], dtype=np.int64), device=device)
        self.post = torch.as_tensor(np.asarray(edge["post"], dtype=np.int64), device=device)
        E = len(self.pre)
        self.params = {}
        for p in stype.param_names:
            x = syn_vals[f"{self.name}_{p}"].to(dtype)
            if x.ndim > 1:
                raise NotImplementedError("batched synapse parameters are not supported")
            self.params[p] = x
        self.area = area_node[self.post]
        self.inv = torch.arange(E, device=device)
        if stype.state_name is None:
            self.state_key = None
            self.s0 = torch.zeros(B, E, dtype=dtype, device=device)
            return
        self.state_key = f"{self.name}_{stype.state_name}"
        s0 = syn_vals[self.state_key].to(dtype)
        self.s0 = torch.broadcast_to(s0, (*s0.shape[:-1], E)).reshape(-1, E)
        self.s0 = self.s0.expand(B, E) if self.s0.shape[0] == 1 else self.s0

    def apply(self, v, s, lin, cst, dt):
        pre_v, post_v = v[:, self.pre], v[:, self.post]
        s2 = self.update(s, pre_v, self.params, dt)
        d0 = self.current(s2, pre_v, post_v, self.params) / self.area * 1e5
        d1 = self.current(s2, pre_v + DIFF, post_v + DIFF, self.params) / self.area * 1e5
        vt = (d1 - d0) / DIFF
        ct = d0 - vt * post_v
        self.last_current = d0 * self.area / 1e5     # nA per synapse
        return s2, lin.index_add(1, self.post, vt), cst.index_add(1, self.post, -ct)


def _syn_group(edge, syn_vals, area_node, n, B, device, dtype, forward_ad=False):
    """The operator kernel for Kinetic / SigmoidRate synapses in the regular layout,
    tensor operations otherwise (and under forward-mode differentiation, for which the
    operator has no rule)."""
    name = edge["type"].name
    if name in MODELS and not forward_ad:
        post = np.asarray(edge["post"], dtype=np.int64)
        perm = np.argsort(post, kind="stable")
        if kernel_plan(name, np.asarray(edge["pre"])[perm], post[perm], n, device) is not None:
            return _SynGroup(edge, syn_vals, area_node, n, B, device, dtype)
        raise NotImplementedError("gradients through Kinetic / SigmoidRate synapses need every "
                                  "target to hold the same number of synapses (regular layout)")
    return _TorchSynGroup(edge, syn_vals, area_node, B, device, dtype)


class _Recorder:
    """The reference's recorder (`_build_recorder`, integrate.py:53-84) over the per-step
    path's states. Recordable: `v`; channel gates (NaN where the channel is absent);
    channel membrane currents `i_<channel>` (mA / cm^2 at the step's starting voltage
    with the new gates, summed per current name, zero where absent; base.py:3445-3502);
    synapse states and synapse currents `i_<synapse>` (nA, after the state update;
    network.py:462-465), indexed by the synapse's index within its type. Each state is
    gathered with one index operation; the result keeps the order of the recordings."""

    def __init__(self, module, chans, groups, comp_node, gate_keys, B, n, device, dtype):
        self.B, self.dtype, self.device = B, dtype, device
        names = [s for s, _ in module.recordings]
        idx = np.array([i for _, i in module.recordings], dtype=np.int64)
        self.n_rec = len(names)
        cur_of_kid = {ch.kernel_id: ch.current_name for ch in module.channels}
        gate_pos = {}
        gi = 0
        for kid in _GATE_ORDER:
            if kid in chans:
                for k in _CHANNEL_GATES[kid]:
                    gate_pos[k] = (gi, chans[kid][2].cpu().numpy())
                    gi += 1
        syn_state = {g.state_key: j for j, g in enumerate(groups) if g.state_key is not None}
        syn_cur = {f"i_{g.name}": j for j, g in enumerate(groups)}
        self.channel_current = {}
        self.aux_names = []
        self.getters = []
        order = []
        for name in dict.fromkeys(names):
            pos = np.nonzero(np.array(names) == name)[0]
            order.append(pos)
            ind = idx[pos]
            if name == "v":
                nodes = comp_node[torch.as_tensor(ind, device=device)]
                self.getters.append(lambda v, g, s, a, nodes=nodes: v[:, nodes])
            elif name in gate_pos:
                gi, cidx = gate_pos[name]
                col = np.full(int(module._n_comp), -1, dtype=np.int64)
                col[cidx] = np.arange(len(cidx))
                c = col[ind]
                cols = torch.as_tensor(np.maximum(c, 0), device=device)
                has = torch.as_tensor(c >= 0, device=device)
                self.getters.append(lambda v, g, s, a, gi=gi, cols=cols, has=has: torch.where(
                    has, g[gi][:, cols], torch.full_like(g[gi][:, cols], float("nan"))))
            elif name in cur_of_kid.values():
                for kid, cn in cur_of_kid.items():
                    if cn == name:
                        self.channel_current[kid] = name
                self.aux_names.append(name)
                nodes = comp_node[torch.as_tensor(ind, device=device)]
                self.getters.append(lambda v, g, s, a, name=name, nodes=nodes: self._aux(a, name, nodes))
            elif name in syn_state:
                j = syn_state[name]
                sel = groups[j].inv[torch.as_tensor(ind, device=device)]
                self.getters.append(lambda v, g, s, a, j=j, sel=sel: s[j][:, sel])
            elif name in syn_cur:
                self.aux_names.append(name)
                sel = torch.as_tensor(ind, device=device)
                self.getters.append(lambda v, g, s, a, name=name, sel=sel: self._aux(a, name, sel))
            else:
                raise KeyError(f"{name} is not a recognized state in this module.")
        self.needs_aux = bool(self.aux_names)
        self.single = len(order) == 1
        self.inverse = torch.as_tensor(np.argsort(np.concatenate(order)), device=device)

    def needs(self, name: str) -> bool:
        return name in self.aux_names

    def _aux(self, aux, name, sel):
        if aux is None or name not in aux:
            return torch.zeros(1, len(sel), dtype=self.dtype, device=self.device)
        return aux[name][:, sel]

    def __call__(self, v, gates, syn_s, aux):
        recs = []
        for f in self.getters:
            r = f(v, gates, syn_s, aux)
            recs.append(r.expand(self.B, -1) if r.shape[0] != self.B else r)
        if self.single:
            return recs[0]
        return torch.cat(recs, 1)[:, self.inverse]


def has_tangents(tensors) -> bool:
    """Whether any tensor carries a forward-mode tangent (torch.autograd.forward_ad)."""
    return any(fwAD.unpack_dual(t).tangent is not None for t in tensors if isinstance(t, torch.Tensor))


def integrate_stepwise(module, inp, dt: float, checkpoint_lengths: Optional[List[int]], device,
                       dtype, return_states: bool = False, solver: str = "bwd_euler"):
    """`integrate` through the differentiable per-step path (see the module docstring).
    Returns the recordings `(*batch, n_rec, 1 + T)`. `solver` "crank_nicolson" solves
    with dt / 2 and takes 2 x - v (base.py:3327-3366); "fwd_euler" takes v + dt f(v)
    (solver_voltage.py:412-502); gates and synapses advance by dt."""
    if solver not in ("bwd_euler", "crank_nicolson", "fwd_euler"):
        raise NotImplementedError(f"solver={solver!r}")
    cn = solver == "crank_nicolson"
    fe = solver == "fwd_euler"
    dtv = dt / 2 if cn else dt
    tables, ind = module._ensure_solver()
    n = module._n_nodes
    vals = {k: x.to(dtype) for k, x in inp.vals.items()}
    comp_node = torch.as_tensor(module._comp_node, device=device)
    geo = comp_geometry(vals)
    g_ax = axial_conductances(tables, module._comp_node, n, geo, vals["axial_resistivity"],
                              vals["capacitance"])
    batch_src = [x for x in vals.values() if x.ndim > 1]
    if inp.stim is not None and inp.stim.ndim > 2:
        batch_src.append(inp.stim[..., 0])
    if inp.all_states is not None:
        batch_src += [x for x in inp.all_states.values() if x.ndim > 1]
    batch = tuple(torch.broadcast_shapes(*[x.shape[:-1] for x in batch_src])) if batch_src else ()
    B = int(np.prod(batch)) if batch else 1

    def flat(x):  # (..., m) -> (B, m) or (1, m)
        if x.ndim <= 1:
            return x.reshape(1, -1)
        return torch.broadcast_to(x, (*batch, x.shape[-1])).reshape(B, -1)

    def node_array(comp_vals, fill):
        cv = flat(comp_vals)
        out = torch.full((cv.shape[0], n), fill, dtype=dtype, device=device)
        return out.index_copy(1, comp_node, cv)

    cm = node_array(vals["capacitance"], 1.0)
    area = node_array(geo["area"], 1.0)
    gdt = dtv * flat(g_ax)
    sinks = torch.as_tensor(np.asarray(tables["sink"], dtype=np.int64), device=device)
    # Channels (HH, Leak, Pospischil Na / K / Km) and their compartments.
    chans = {}
    for ch in module.channels:
        comps = np.nonzero(module._channel_mask[ch.name])[0]
        if not len(comps):
            continue
        cidx = torch.as_tensor(comps, device=device)
        names = _CHANNEL_PARAMS[ch.kernel_id]
        prm = {k: flat(vals[k])[:, cidx] for k in names}
        chans[ch.kernel_id] = (comp_node[cidx], prm, cidx)
    # Initial states.
    gate_keys = [k for kid in _GATE_ORDER if kid in chans for k in _CHANNEL_GATES[kid]]
    init = {k: vals[k] for k in vals if k in ["v"] + gate_keys}
    if inp.all_states is not None:
        for k, x in inp.all_states.items():
            if k in init:
                init[k] = x[..., comp_node]
    v = torch.broadcast_to(node_array(init["v"], 0.0), (B, n))
    gates = []
    for kid in _GATE_ORDER:
        if kid in chans:
            cidx = chans[kid][2]
            gates += [torch.broadcast_to(flat(init[k])[:, cidx], (B, len(cidx))) for k in _CHANNEL_GATES[kid]]
    forward_ad = has_tangents(list(inp.vals.values()) + list(inp.syn_vals.values()) +
                              ([inp.stim] if inp.stim is not None else []) +
                              (list(inp.all_states.values()) if inp.all_states is not None else []))
    groups = [_syn_group(e, inp.syn_vals, area[0], n, B, device, dtype, forward_ad)
              for e in module._edges]
    if area.shape[0] > 1 and groups:
        raise NotImplementedError("batched geometry with synapses is not supported")
    syn_s = [g.s0 for g in groups]
    # Stimulus (B or 1, n_stim, T), scaled to uA/cm^2 (cell_utils.py:275-288).
    stim = stim_nodes = None
    if inp.stim is not None:
        s = inp.stim.to(dtype)
        stim = s.reshape(-1, *s.shape[-2:]) if s.ndim > 2 else s[None]
        if stim.shape[0] > 1:
            stim = torch.broadcast_to(s, (*batch, *s.shape[-2:])).reshape(B, *s.shape[-2:])
        stim_nodes = torch.as_tensor(np.asarray(inp.stim_nodes, dtype=np.int64), device=device)
    # Solve order (solver_voltage.py:294-359) and the tree-solve kernel.
    mp = torch.as_tensor(np.asarray(ind["map_to_solve_order"], dtype=np.int64), device=device)
    inv = torch.as_tensor(np.asarray(ind["inv_map_to_solve_order"], dtype=np.int64), device=device)
    mlo = torch.as_tensor(np.asarray(ind["map_to_solve_order_lower"], dtype=np.int64), device=device)
    mup = torch.as_tensor(np.asarray(ind["map_to_solve_order_upper"], dtype=np.int64), device=device)
    kern = make_tree_solve_kernel(np.asarray(ind["parent_lookup"]), device)
    if kern is None:
        raise NotImplementedError("the solve order is not supported by the tree-solve kernel")
    internal_mask = torch.zeros(n, dtype=torch.bool, device=device)
    internal_mask[comp_node] = True
    lu = torch.broadcast_to(-gdt, (B, gdt.shape[-1]))
    lo = torch.cat([lu[:, mlo], lu.new_zeros(B, 1)], 1) if len(mlo) else lu.new_zeros(B, n + 1)
    up = torch.cat([lu[:, mup], lu.new_zeros(B, 1)], 1) if len(mup) else lu.new_zeros(B, n + 1)
    diag0 = torch.zeros(B, n, dtype=dtype, device=device).index_add(
        1, comp_node, torch.ones(B, len(comp_node), dtype=dtype, device=device))
    if len(sinks):
        diag0 = diag0.index_add(1, sinks, torch.broadcast_to(gdt, (B, gdt.shape[-1])))

    g_full = torch.broadcast_to(flat(g_ax), (B, g_ax.shape[-1]))
    edges = _explicit_edges(tables, device)

    rec = _Recorder(module, chans, groups, comp_node, gate_keys, B, n, device, dtype)

    def add_current(aux, kid, nodes, g, gv, vn):
        """The channel's membrane current g v - gv (mA / cm^2) at the step's starting
        voltage with the new gates, summed per current name (base.py:3445-3502)."""
        name = rec.channel_current.get(kid)
        if name is not None:
            prev = aux.get(name, torch.zeros(B, n, dtype=dtype, device=device))
            aux[name] = prev.index_add(1, nodes, torch.broadcast_to(g * vn - gv, (B, len(nodes))))

    def step(v, gates, syn_s, t, aux=None):
        lin = torch.zeros(B, n, dtype=dtype, device=device)
        cst = torch.zeros(B, n, dtype=dtype, device=device)
        gates = list(gates)
        gi = 0
        if "HH" in chans:
            nodes, p, _ = chans["HH"]
            vh = v[:, nodes]
            m, h, nn = _hh_gates(*gates[:3], vh, dt)
            gates[:3] = [m, h, nn]
            gi = 3
            gNa = p["HH_gNa"] * m**3 * h
            gK = p["HH_gK"] * nn**4
            g = gNa + gK + p["HH_gLeak"]
            gv = gNa * p["HH_eNa"] + gK * p["HH_eK"] + p["HH_gLeak"] * p["HH_eLeak"]
            # i = g v - (gNa eNa + gK eK + gL eL): the slope of a current linear in v is g
            # (the reference's finite difference up to round-off).
            lin = lin.index_add(1, nodes, g)
            cst = cst.index_add(1, nodes, gv)
            if aux is not None:
                add_current(aux, "HH", nodes, g, gv, vh)
        if "Leak" in chans:
            nodes, p, _ = chans["Leak"]
            gl = torch.broadcast_to(p["Leak_gLeak"], (B, len(nodes)))
            gv = torch.broadcast_to(p["Leak_gLeak"] * p["Leak_eLeak"], (B, len(nodes)))
            lin = lin.index_add(1, nodes, gl)
            cst = cst.index_add(1, nodes, gv)
            if aux is not None:
                add_current(aux, "Leak", nodes, gl, gv, v[:, nodes])
        for kid in _GATE_ORDER[1:]:
            if kid in chans:
                nodes, p, _ = chans[kid]
                k = len(_CHANNEL_GATES[kid])
                new, g, e = _pospischil(kid, gates[gi:gi + k], v[:, nodes], p, dt)
                gates[gi:gi + k] = new
                gi += k
                lin = lin.index_add(1, nodes, torch.broadcast_to(g, (B, len(nodes))))
                cst = cst.index_add(1, nodes, torch.broadcast_to(g * e, (B, len(nodes))))
                if aux is not None:
                    add_current(aux, kid, nodes, g, g * e, v[:, nodes])
        lin = lin * 1000.0
        cst = cst * 1000.0
        new_s = []
        for grp, s in zip(groups, syn_s):
            s2, lin, cst = grp.apply(v, s, lin, cst, dt)
            new_s.append(s2)
            if aux is not None and rec.needs(f"i_{grp.name}"):
                aux[f"i_{grp.name}"] = grp.last_current[:, grp.inv]
        if stim is not None:
            cur = stim[..., t]
            cst = cst.index_add(1, stim_nodes, torch.broadcast_to(cur / area[:, stim_nodes] * 1e5, (B, len(stim_nodes))))
        if fe:
            return _explicit_step(v, lin / cm, cst / cm, g_full, edges, dt), gates, new_s
        diag = diag0 + dtv * lin / cm
        rhs = torch.where(internal_mask, v + dtv * cst / cm, torch.zeros_like(v))
        dd = torch.cat([diag[:, mp], diag.new_ones(B, 1)], 1)
        ss = torch.cat([rhs[:, mp], rhs.new_zeros(B, 1)], 1)
        x = fused_tree_solve(kern, dd, lo, up, ss)[:, :-1][:, inv]
        if cn:
            x = 2 * x - v
        return x, gates, new_s

    T = inp.T
    K = int(checkpoint_lengths[-1]) if checkpoint_lengths else DEFAULT_CHUNK
    K = max(1, min(K, T))
    n_g, n_s = len(gates), len(syn_s)

    def chunk(t0, t1, v, *rest):
        gates, syn_s = list(rest[:n_g]), list(rest[n_g:])
        recs = []
        for t in range(t0, t1):
            aux = {} if rec.needs_aux else None
            v, gates, syn_s = step(v, gates, syn_s, t, aux)
            recs.append(rec(v, gates, syn_s, aux))
        return (v, *gates, *syn_s, torch.stack(recs, -1))

    init_aux = None
    if rec.needs_aux:
        init_aux = {}
        for name in rec.aux_names:
            x0 = inp.all_states.get(name) if inp.all_states is not None else None
            if x0 is not None:
                init_aux[name] = torch.broadcast_to(torch.as_tensor(x0, dtype=dtype, device=device),
                                                    (*batch, x0.shape[-1])).reshape(B, -1)
    recs = [rec(v, gates, syn_s, init_aux)[..., None]]
    state = (v, *gates, *syn_s)
    for t0 in range(0, T, K):
        t1 = min(t0 + K, T)
        if torch.is_grad_enabled() and not forward_ad:
            out = checkpoint(chunk, t0, t1, *state, use_reentrant=False)
        else:
            out = chunk(t0, t1, *state)
        state, r = out[:-1], out[-1]
        recs.append(r)
    out_rec = torch.cat(recs, -1).reshape(*batch, rec.n_rec, T + 1)
    if not return_states:
        return out_rec
    v, gates, syn_s = state[0], list(state[1:1 + n_g]), list(state[1 + n_g:])
    out = {"v": v.reshape(*batch, n)}
    gi = 0
    for kid in _GATE_ORDER:
        if kid in chans:
            for k in _CHANNEL_GATES[kid]:
                full = torch.full((B, n), float("nan"), dtype=dtype, device=device)
                out[k] = full.index_copy(1, chans[kid][0], gates[gi]).reshape(*batch, n)
                gi += 1
    for grp, s in zip(groups, syn_s):
        if grp.state_key is not None:
            out[grp.state_key] = s[:, grp.inv].reshape(*batch, -1)
    return out_rec, out
