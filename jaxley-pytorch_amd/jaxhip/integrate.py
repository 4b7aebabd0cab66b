"""`integrate` and `build_init_and_step_fn`: the drop-in for jaxley-pytorch's
`jx.integrate` (jaxley_pytorch/integrate.py:396-666) on the GPU.

Same signature and conventions: device and dtype from the first floating point input;
stimuli `(..., n_inputs, time)` padded / truncated to `int(t_max // delta_t + 1)` steps;
leading batch dimensions of parameters, `data_set` values, stimuli and states broadcast
against each other; recordings `(..., n_recordings, 1 + steps)` (the initial state and
every step); `voltage_solver` in {"jaxley.dhs", "jaxley.dhs.gpu", "jaxley.dhs.cpu",
"jaxley.stone"}; `solver` "bwd_euler", "crank_nicolson" or "fwd_euler"; gradients
through autograd.

Execution: the whole time loop is one fused HIP kernel launch (`jaxhip.fused`), with a
hand-written adjoint (checkpoints every K steps, recomputation between them) for
gradients of the channel parameters and the initial voltages. There is no CPU path:
inputs on the CPU raise.
"""
from __future__ import annotations

import math
import os
from typing import Any, Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .fused import (ALL_PARAMS, BATCHABLE, HH_PARAMS, LEAK_PARAMS, POSP_PARAMS, POSP_STATES, ForwardLaunch,
                    LaunchConfig,
                    build_structure, default_config, node_float_table, syn_has_state, to_prog)
from .geometry import GEOMETRY_KEYS, axial_conductances, comp_geometry, cylinder_geometry

_SUPPORTED_SOLVERS = ("jaxley.dhs.gpu", "jaxley.dhs.cpu", "jaxley.stone")
# Time steppers (base.py:3327-3366): backward Euler; Crank-Nicolson solves with dt / 2
# and takes v' = 2 x - v; forward Euler (solver_voltage.py:412-502) needs no solve.
_SUPPORTED_STEPS = ("bwd_euler", "crank_nicolson", "fwd_euler")


def _resolve_voltage_solver(module, voltage_solver: str) -> str:
    """integrate.py:21-27."""
    if voltage_solver == "jaxley.dhs":
        return "jaxley.dhs.gpu" if module._solver_device in ("gpu", "tpu") else "jaxley.dhs.cpu"
    return voltage_solver


def _tensors(tree):
    if isinstance(tree, torch.Tensor):
        yield tree
    elif isinstance(tree, dict):
        for v in tree.values():
            yield from _tensors(v)
    elif isinstance(tree, (list, tuple)):
        for v in tree:
            yield from _tensors(v)


def first_float_tensor(*trees) -> Optional[torch.Tensor]:
    for tree in trees:
        for t in _tensors(tree):
            if t.is_floating_point():
                return t
    return None


def params_to_pstate(params, indices):
    """utils/cell_utils.py:258-272."""
    return [{"key": list(p.keys())[0], "val": list(p.values())[0], "indices": i}
            for p, i in zip(params, indices)]


def set_at(x: torch.Tensor, idx, val) -> torch.Tensor:
    """`x.at[..., idx].set(val)` along the last dimension, negative indices wrapping
    (utils/torch_utils.py:84-101)."""
    idx = torch.as_tensor(np.asarray(idx, dtype=np.int64), device=x.device)
    if idx.numel() == 0:
        return x
    n = x.shape[-1]
    val = torch.as_tensor(val, dtype=x.dtype, device=x.device)
    nvb = max(val.ndim - idx.ndim, 0)
    batch = torch.broadcast_shapes(x.shape[:-1], val.shape[:nvb])
    val = torch.broadcast_to(val, (*batch, *idx.shape)).reshape(*batch, -1)
    x = torch.broadcast_to(x, (*batch, n))
    return x.index_copy(-1, torch.remainder(idx.reshape(-1), n), val)


class _Inputs:
    """Everything `integrate` assembles before the time loop."""


def _assemble(module, params, param_state, data_stimuli, t_max, delta_t, all_states, device, dtype,
              module_stimuli: bool = True):
    inp = _Inputs()
    pstate = params_to_pstate(params, module.indices_set_by_trainables)
    if param_state is not None:
        pstate = pstate + list(param_state)
    # Compartment parameters and states (…, n_comp).
    vals: Dict[str, torch.Tensor] = {}
    for k, col in module._node_data.items():
        vals[k] = torch.as_tensor(col, dtype=dtype, device=device)
    syn_vals: Dict[str, torch.Tensor] = {}
    for e in module._edges:
        for d in (e["params"], e["states"]):
            for k, col in d.items():
                syn_vals[k] = torch.as_tensor(col, dtype=dtype, device=device)
    for p in pstate:
        key, inds, val = p["key"], p["indices"], p["val"]
        val = torch.as_tensor(val, dtype=dtype, device=device)
        if key in vals:
            vals[key] = set_at(vals[key], inds, val[..., :, None])
            if key in ("radius", "length"):
                # Those compartments become cylinders (modules/base.py:1997-2015).
                it = torch.as_tensor(np.asarray(inds, dtype=np.int64), device=device)
                geo = cylinder_geometry(vals["length"][..., it], vals["radius"][..., it])
                for gk, gv in geo.items():
                    vals[gk] = set_at(vals[gk], inds, gv)
        elif key in syn_vals:
            syn_vals[key] = set_at(syn_vals[key], inds, val[..., :, None])
        else:
            raise KeyError(f"unknown parameter {key!r}")
    inp.vals, inp.syn_vals = vals, syn_vals
    # Stimuli (…, n_stim, time) and their nodes.
    curs, nodes = [], []
    ext = module.externals if module_stimuli else {}
    if "i" in ext:
        curs.append(ext["i"].to(device=device, dtype=dtype))
        nodes.append(module.external_inds["i"])
    if data_stimuli is not None:
        for name, c, ind in zip(*data_stimuli):
            if name != "i":
                raise NotImplementedError("only current stimuli ('i') are supported")
            curs.append(torch.as_tensor(c, dtype=dtype, device=device))
            nodes.append(np.asarray(ind, dtype=np.int64))
    if curs:
        batch = torch.broadcast_shapes(*[c.shape[:-2] for c in curs])
        curs = [torch.broadcast_to(c, (*batch, *c.shape[-2:])) for c in curs]
        stim = torch.cat(curs, dim=-2)
        stim_nodes = np.concatenate(nodes).astype(np.int64)
    else:
        stim, stim_nodes = None, np.zeros(0, np.int64)
    if t_max is not None:
        T = int(t_max // delta_t + 1)
        if stim is not None:
            if stim.shape[-1] < T:
                pad = torch.zeros((*stim.shape[:-1], T - stim.shape[-1]), dtype=dtype, device=device)
                stim = torch.cat([stim, pad], -1)
            else:
                stim = stim[..., :T]
    else:
        if stim is None:
            raise AssertionError("If no stimulus or clamp are inserted you have to specify the "
                                 "simulation duration at `jx.integrate(..., t_max=)`.")
        T = stim.shape[-1]
    inp.stim, inp.stim_nodes, inp.T = stim, stim_nodes, T
    # Initial states.
    if all_states is not None:
        inp.all_states = {k: torch.as_tensor(v, dtype=dtype, device=device) for k, v in all_states.items()}
    else:
        inp.all_states = None
    return inp


def _require_gpu(device) -> None:
    if torch.device(device).type != "cuda":
        raise RuntimeError("jaxhip.integrate runs on the GPU: pass parameters, stimuli or "
                           "states on a cuda device (or set torch.set_default_device('cuda'))")


def _batch_of(*ts) -> Tuple[int, ...]:
    shapes = [t.shape[:-1] for t in ts if isinstance(t, torch.Tensor) and t.ndim > 0]
    return tuple(torch.broadcast_shapes(*shapes)) if shapes else ()


def integrate(module, params: List[Dict[str, torch.Tensor]] = [], *, param_state=None,
              data_stimuli=None, data_clamps=None, t_max: Optional[float] = None,
              delta_t: float = 0.025, solver: str = "bwd_euler",
              voltage_solver: str = "jaxley.dhs", checkpoint_lengths: Optional[List[int]] = None,
              all_states: Optional[Dict] = None, return_states: bool = False,
              compile: bool = False):
    """Simulate `module` on the GPU; see the module docstring. `compile` is accepted
    for signature compatibility (the step is always one compiled HIP kernel)."""
    voltage_solver = _resolve_voltage_solver(module, voltage_solver)
    if voltage_solver not in _SUPPORTED_SOLVERS:
        raise NotImplementedError(f"voltage_solver={voltage_solver!r} is not part of jaxhip "
                                  f"(supported: {_SUPPORTED_SOLVERS})")
    if solver not in _SUPPORTED_STEPS:
        raise NotImplementedError(f"solver={solver!r}: supported are {_SUPPORTED_STEPS}")
    if data_clamps is not None:
        raise NotImplementedError("clamps are not part of jaxhip")
    tables, _ = module._ensure_solver()
    if voltage_solver == "jaxley.stone" and np.any(np.isin(tables["type"], [1, 2, 3, 4])):
        raise NotImplementedError("The stone solver is not implemented for branched morphologies.")
    # Device and dtype from the first floating point input, module stimuli included
    # (integrate.py:480-490).
    ref = first_float_tensor(params, param_state, data_stimuli[1] if data_stimuli else None,
                             all_states, module.externals)
    device = ref.device if ref is not None else torch.get_default_device()
    dtype = ref.dtype if ref is not None else torch.get_default_dtype()
    _require_gpu(device)
    if dtype not in (torch.float32, torch.float64):
        raise TypeError(f"dtype {dtype} is not supported (float32 / float64)")
    if not module.recordings:
        raise ValueError("No recordings are set. Please set them.")
    inp = _assemble(module, params, param_state, data_stimuli, t_max, delta_t, all_states, device, dtype)
    if _per_step_path(module, inp, solver):
        from .stepwise import integrate_stepwise

        return integrate_stepwise(module, inp, delta_t, checkpoint_lengths, device, dtype, return_states,
                                  solver=solver)
    return _integrate_fused(module, inp, delta_t, checkpoint_lengths, return_states, device, dtype,
                            solver)


_GEOMETRY_INPUTS = ("length", "radius", "axial_resistivity", "capacitance") + GEOMETRY_KEYS


def _per_step_path(module, inp, solver: str) -> bool:
    """Whether a simulation takes the differentiable per-step path (`stepwise`) instead
    of the fused whole-loop kernels. The fused forward records voltages of a geometry
    shared by the batch; its hand-written adjoint covers backward-Euler HH / Leak cells
    without synapses and differentiates the channel parameters and initial voltages. The
    per-step path takes everything else: recordings of gates, currents or synapse states;
    batched or differentiated geometry; gradients through synapses, Pospischil channels,
    Crank-Nicolson / forward-Euler steps, stimuli or initial gates and synapse states;
    batched Pospischil parameters; forward-mode tangents."""
    if {s for s, _ in module.recordings} != {"v"}:
        return True
    if any(k in inp.vals and (inp.vals[k].ndim > 1 or inp.vals[k].requires_grad) for k in _GEOMETRY_INPUTS):
        return True
    if _has_tangents(inp):
        return True
    if has_pospischil(module) and any(inp.vals[k].ndim > 1 for k in POSP_PARAMS if k in inp.vals):
        return True
    if not _needs_grad(inp):
        return False
    if solver != "bwd_euler" or has_pospischil(module):
        return True
    syn_grad = []
    if module._edges:
        # The fused adjoint covers one KineticSynapse or SigmoidRateSynapse type and the
        # gradient of its gS.
        types = {e["type"].name for e in module._edges}
        kinds = {int(e["type"].kernel_model) for e in module._edges}
        if len(types) != 1 or not kinds <= {0, 1}:
            return True
        name = types.pop()
        syn_grad = [x for k, x in inp.syn_vals.items() if k != f"{name}_gS"]
    # The fused adjoint: gradients of the HH / Leak parameters and the initial voltages.
    others = [inp.vals[k] for k in inp.vals if k not in ALL_PARAMS and k != "v"] + syn_grad
    if inp.stim is not None:
        others.append(inp.stim)
    if inp.all_states is not None:
        others += [x for k, x in inp.all_states.items() if k != "v"]
    return any(t.requires_grad for t in others)


def has_pospischil(module) -> bool:
    return any(ch.kernel_id in ("PNa", "PK", "PKm") for ch in module.channels)


def _has_tangents(inp) -> bool:
    """Forward-mode differentiation (`torch.autograd.forward_ad` dual inputs): the
    per-step path, whose tree solve has the forward rule of `_DHSSolveForwardAD`
    (solver_voltage.py:190-224)."""
    from .stepwise import has_tangents

    ts = list(inp.vals.values()) + list(inp.syn_vals.values())
    if inp.stim is not None:
        ts.append(inp.stim)
    if inp.all_states is not None:
        ts += list(inp.all_states.values())
    return has_tangents(ts)


def _needs_grad(inp) -> bool:
    if not torch.is_grad_enabled():
        return False
    ts = list(inp.vals.values()) + list(inp.syn_vals.values())
    if inp.stim is not None:
        ts.append(inp.stim)
    if inp.all_states is not None:
        ts += list(inp.all_states.values())
    return any(t.requires_grad for t in ts)


_F32_MODE = os.environ.get("JAXHIP_F32_MODE", "mixed")


def set_float32_mode(mode: str) -> None:
    """How float32 simulations run: "mixed" (default) keeps gates, rates and voltages
    in float64 and solves the tree on float32 increments; "fast" runs gates and rates
    in float32 too (voltages stay float64).

    Measured on the 4096 simulations of the batched parameter sweep (random tree of 1000
    compartments, 2000 steps) against float64 on the same float32 inputs: mixed max
    0.038 mV (3 simulations above 1e-2 mV), fast max 0.21 mV (129 above 1e-2 mV). Those
    simulations sit near a spike threshold: rounding the inputs themselves to float32
    moves the float64 trajectories by up to 0.023 mV. jaxley-pytorch's own float32
    simulation of 138 of them (those above, and 8 others) ends up to 0.86 mV (median
    0.094 mV) from its float64 one; on those the mixed mode stays within 0.022 mV
    (median 0.0002) and the fast mode within 0.19 mV (median 0.009)."""
    global _F32_MODE
    if mode not in ("mixed", "fast"):
        raise ValueError(f"unknown float32 mode {mode!r}")
    _F32_MODE = mode


class Problem:
    """A simulation prepared for the fused kernel: structure, node tables, batched
    conductances, initial states, stimulus and synapses (see `_prepare`)."""


def _structure(module, rec_comps, stim_nodes, device, B: int, lsize: int, grad: bool):
    """The module's kernel tables (cached), split into parts for a network: one part per
    cell up to one program per compute unit when B simulations leave units idle (at one
    simulation per program), more parts while the largest does not fit the LDS (of the
    forward, and of the reverse pass with `grad`). B = 0: one part."""
    from .fused import adjoint_lds_bytes, cell_sizes, lds_bytes

    ncell = len(cell_sizes(module)) if B else 1
    dev = torch.device(device)
    cus = torch.cuda.get_device_properties(dev).multi_processor_count if dev.type == "cuda" else 256
    parts = max(1, min(ncell, cus // max(B, 1))) if ncell > 1 else 1
    cache = module.__dict__.setdefault("_fused_cache", {})
    base = (module._version, tuple(rec_comps.tolist()), tuple(np.asarray(stim_nodes).tolist()), str(device))
    for k in list(cache):
        if k[:4] != base:
            del cache[k]
    while True:
        key = base + (parts,)
        st = cache.get(key)
        if st is None:
            st = build_structure(module, rec_comps, stim_nodes, device, parts)
            cache[key] = st
        n, nt = st.n_max, st.tgt_max
        fits = lds_bytes(n, nt, 1, lsize, st.sched, gcoef=True) <= 160 * 1024
        if grad:
            fits = fits and adjoint_lds_bytes(n, 1, lsize, st.sched, True, nt) <= 160 * 1024
        if fits or parts >= ncell:
            return st
        parts = min(ncell, 2 * parts)


def _prepare(module, inp, dt, device, user_dtype, solver: str = "bwd_euler", split: bool = True) -> Problem:
    """Everything the fused kernel needs, independent of the launch shape. The tree's
    coefficients and the stimulus scale use the solve's step (dt / 2 for
    Crank-Nicolson); gates and synapses advance by dt. With `split`, a network is split
    into parts (consecutive cells, one program each per group of simulations) when its
    batch leaves the device's compute units idle or a part of it is all the LDS holds."""
    cn = solver == "crank_nicolson"
    dtv = dt / 2 if cn else dt
    # Element type of gates, rates, parameters and recordings in the kernel.
    dtype = torch.float64 if (user_dtype == torch.float32 and _F32_MODE == "mixed") else user_dtype
    vals = {k: v.to(dtype) for k, v in inp.vals.items()}
    rec_states = {s for s, _ in module.recordings}
    if rec_states != {"v"}:
        raise NotImplementedError("the fused kernel records voltages ('v') only")
    rec_comps = np.array([c for _, c in module.recordings], dtype=np.int64)
    for k in _GEOMETRY_INPUTS:
        if vals[k].ndim > 1:
            raise NotImplementedError(f"batched {k!r} is not supported by the fused kernel")
        if vals[k].requires_grad:
            raise NotImplementedError(f"gradients with respect to {k!r} take the per-step path")
    # Batch: parameters, states, stimulus.
    chan_keys = [k for k in ALL_PARAMS if k in vals]
    state_keys = ["v"] + [k for k in ("HH_m", "HH_h", "HH_n") + POSP_STATES if k in vals]
    init = {k: vals[k] for k in state_keys}
    if inp.all_states is not None:
        for k in state_keys:
            if k in inp.all_states:
                x = inp.all_states[k]
                init[k] = x[..., torch.as_tensor(module._comp_node, device=device)]
    batch_t = [vals[k] for k in chan_keys] + list(init.values())
    if inp.stim is not None:
        batch_t.append(inp.stim[..., 0])
    syn_keys = list(inp.syn_vals)
    batch_t += [inp.syn_vals[k] for k in syn_keys]
    batch = _batch_of(*batch_t)
    B = int(np.prod(batch)) if batch else 1
    needs_grad = _needs_grad(inp)
    st = _structure(module, rec_comps, inp.stim_nodes, device, B if split else 0,
                    torch.tensor([], dtype=user_dtype).element_size(), needs_grad)
    tables = module._comp_edges
    geo = comp_geometry(vals)
    g = axial_conductances(tables, module._comp_node, module._n_nodes, geo, vals["axial_resistivity"], vals["capacitance"])
    cm_node = torch.full((module._n_nodes,), -1.0, dtype=dtype, device=device)
    cm_node[torch.as_tensor(module._comp_node, device=device)] = vals["capacitance"]
    node_f = node_float_table(st, g, cm_node, dtv, user_dtype)

    n = st.n
    comp_of_k = torch.as_tensor(st.comp_of_k, device=device)
    is_comp = comp_of_k >= 0
    cs_safe = torch.where(is_comp, comp_of_k, torch.zeros_like(comp_of_k))
    # Branchpoints take the value of their parent compartment (their initial voltage is
    # never read by the reference: their right-hand side is 0; here it seeds the
    # increment form with a consistent value).
    par_k = torch.as_tensor(st.parent_k, device=device)
    bp_src = torch.where(is_comp, cs_safe, cs_safe[torch.clamp(par_k, min=0)])

    def comp_to_kernel(x):  # (..., n_comp) -> (..., n) kernel order, 0 at branchpoints
        y = x[..., cs_safe]
        return torch.where(is_comp, y, torch.zeros_like(y))

    # Channel parameters: a per-node table (kernel order) and, when some conductance
    # carries batch dimensions, a batched table of the four conductances.
    def kernel_cols(x):
        return torch.nan_to_num(comp_to_kernel(x), nan=0.0)

    zero_n = torch.zeros(n, dtype=dtype, device=device)
    batched = [k for k in ALL_PARAMS if k in vals and vals[k].ndim > 1]
    for k in batched:
        if k not in BATCHABLE:
            raise NotImplementedError(f"batched {k!r} is not supported by the fused kernel")
    bmode, bparam = 0, None
    if batched:
        tabs = []
        for k in BATCHABLE:
            x = vals.get(k)
            if x is None:
                tabs.append(torch.zeros(B, n, dtype=dtype, device=device))
            else:
                tabs.append(kernel_cols(torch.broadcast_to(x, (*batch, x.shape[-1])).reshape(B, -1)))
        stack = torch.stack(tabs, -1)                                     # (B, n, 4)
        chan = torch.as_tensor(np.asarray(st.node_i.cpu())[:, 3] & 6 != 0, device=device)
        first = int(torch.nonzero(chan)[0]) if bool(chan.any()) else 0
        with torch.no_grad():
            per_sim = bool(((stack - stack[:, first:first + 1]).abs()[:, chan] == 0).all())
        bmode = 1 if per_sim else 2
        bsim = stack[:, first] if per_sim else None                       # (B, 4)

    def bparam_for(S):
        """The batched conductances in the layout of programs of S simulations."""
        if bmode == 1:
            P = -(-B // S)
            b = bsim if P * S == B else torch.cat([bsim, bsim.new_zeros(P * S - B, 4)], 0)
            return b.contiguous()
        if bmode == 2:
            return torch.stack([to_prog(stack[..., c], S) for c in range(4)], -1).contiguous()
        return None
    cols = []
    for j, k in enumerate(ALL_PARAMS):
        x = vals.get(k)
        if x is None or (bmode and k in BATCHABLE):
            cols.append(zero_n)
        else:
            cols.append(kernel_cols(x))
    node_p = torch.stack(cols, 1).contiguous()
    param_inputs = {k: vals[k] for k in chan_keys}
    # States.
    v0 = torch.broadcast_to(init["v"][..., bp_src], (*batch, n)).reshape(B, n)
    gates0 = []
    for k in ("HH_m", "HH_h", "HH_n"):
        if k in init:
            x = init[k]
            x = torch.where(torch.isnan(x), torch.zeros_like(x), x)
            gates0.append(torch.broadcast_to(comp_to_kernel(x), (*batch, n)).reshape(B, n))
        else:
            gates0.append(torch.zeros(B, n, dtype=dtype, device=device))
    # Pospischil channels: their own parameter table and gates (unbatched parameters).
    posp_kw = {}
    if has_pospischil(module):
        for k in POSP_PARAMS:
            if k in vals and (vals[k].ndim > 1 or vals[k].requires_grad):
                raise NotImplementedError(f"batched or trainable {k!r} is not supported by the fused kernel")
        node_p2 = torch.stack([kernel_cols(vals[k]) if k in vals else zero_n for k in POSP_PARAMS], 1)
        gx = []
        for k in POSP_STATES:
            if k in init:
                x = torch.nan_to_num(init[k], nan=0.0)
                gx.append(torch.broadcast_to(comp_to_kernel(x), (*batch, n)).reshape(B, n))
            else:
                gx.append(torch.zeros(B, n, dtype=dtype, device=device))
        posp_kw = {"node_p2": node_p2.contiguous(), "gates_ext0": tuple(gx)}
    # Stimulus (T, n_stim, stim_b) and per-entry scale dt * 1e5 / (area cm).
    stim = stim_scale = None
    if inp.stim is not None:
        s = inp.stim.to(dtype)
        sb = s.shape[:-2]
        if sb:
            s = torch.broadcast_to(s, (*batch, *s.shape[-2:])).reshape(B, *s.shape[-2:])
            stim = s.permute(2, 1, 0).contiguous()
        else:
            stim = s.permute(1, 0)[..., None].contiguous()
        area_node = torch.full((module._n_nodes,), 1.0, dtype=dtype, device=device)
        area_node[torch.as_tensor(module._comp_node, device=device)] = geo["area"]
        sn = torch.as_tensor(inp.stim_nodes, device=device)
        scale = dtv * 1e5 / (area_node[sn] * cm_node[sn])
        order_e = torch.as_tensor(st.stim_order, device=device)
        stim_scale = scale[order_e].contiguous()
    # Synapses.
    syn_params = syn_scale = syn_s0 = None
    if st.syn is not None:
        groups = st.syn["groups"]
        perm = torch.as_tensor(st.syn["perm"], device=device)
        # Parameter slot j holds every synapse's j-th parameter of its own type (in the
        # concatenated edge order of the types, then sorted by target).
        n_slots = max(len(g["type"].param_names) for g in groups)
        syn_params = []
        for j in range(n_slots):
            cols = []
            for g in groups:
                names = g["type"].param_names
                if j < len(names):
                    x = inp.syn_vals[f"{g['type'].name}_{names[j]}"].to(dtype)
                    if x.ndim > 1:
                        raise NotImplementedError("batched synapse parameters are not supported by the fused kernel")
                    cols.append(x)
                else:
                    cols.append(torch.zeros(g["n"], dtype=dtype, device=device))
            syn_params.append(torch.cat(cols)[perm].contiguous())
        area_node = torch.full((module._n_nodes,), 1.0, dtype=dtype, device=device)
        area_node[torch.as_tensor(module._comp_node, device=device)] = geo["area"]
        syn_scale = (1e5 / area_node[torch.as_tensor(st.syn["post_node"], device=device)]).contiguous()
        if syn_has_state(st.syn):
            cols = []
            for g in groups:
                key = _syn_state_key(g)
                if key is None:
                    cols.append(torch.zeros((B, g["n"]), dtype=dtype, device=device))
                    continue
                s0 = inp.syn_vals[key].to(dtype)
                if inp.all_states is not None and key in inp.all_states:
                    s0 = inp.all_states[key].to(dtype)
                cols.append(torch.broadcast_to(s0, (*batch, s0.shape[-1])).reshape(B, -1))
            syn_s0 = torch.cat(cols, 1)[:, perm]
    P = Problem()
    P.module, P.st, P.node_f, P.node_p, P.bmode = module, st, node_f, node_p, bmode
    P.bparam_for, P.v0, P.gates0, P.stim, P.stim_scale = bparam_for, v0, gates0, stim, stim_scale
    P.syn_params, P.syn_scale, P.syn_s0 = syn_params, syn_scale, syn_s0
    P.B, P.batch, P.T, P.dt, P.dtype, P.user_dtype = B, batch, inp.T, dt, dtype, user_dtype
    # Launch arguments of the time stepper (0 = backward Euler).
    P.step_kw = {"dt_v": dtv if cn else 0.0, "rhs_scale": 2.0 if cn else 0.0,
                 "explicit_step": int(solver == "fwd_euler"), **posp_kw}
    P.needs_grad, P.device, P.stim_requires_grad = needs_grad, device, (
        inp.stim is not None and inp.stim.requires_grad)
    return P


def _integrate_fused(module, inp, dt, checkpoint_lengths, return_states, device, user_dtype,
                     solver: str = "bwd_euler"):
    P = _prepare(module, inp, dt, device, user_dtype, solver)
    st, node_f, node_p, bmode, bparam_for = P.st, P.node_f, P.node_p, P.bmode, P.bparam_for
    v0, gates0, stim, stim_scale = P.v0, P.gates0, P.stim, P.stim_scale
    syn_params, syn_scale, syn_s0 = P.syn_params, P.syn_scale, P.syn_s0
    B, batch, dtype, needs_grad = P.B, P.batch, P.dtype, P.needs_grad
    # Launch shape (measured on a short run of the real problem when not cached).
    from .autotune import _time, fused_config

    def measure(c):
        Tm = min(inp.T, 32)
        try:
            with torch.no_grad():
                run = ForwardLaunch(st, node_f, node_p.detach(),
                                    None if bmode == 0 else bparam_for(c.S).detach(), bmode,
                                    v0.detach(), tuple(g.detach() for g in gates0), stim, stim_scale,
                                    Tm, dt, B, c, syn_params, syn_scale, syn_s0, **P.step_kw)
                return _time(run.launch, reps=2)
        except RuntimeError:
            return float("inf")

    try:
        cfg = fused_config(st, B, dtype, torch.tensor([], dtype=user_dtype).element_size(), needs_grad,
                           measure=measure)
    except NotImplementedError:
        # A cell whose system (or its reverse pass's) does not fit one workgroup's LDS:
        # the per-step path (its tree solve sweeps the levels in HBM).
        from .stepwise import integrate_stepwise

        return integrate_stepwise(module, inp, dt, checkpoint_lengths, device, user_dtype, return_states,
                                  solver=solver)
    bparam = bparam_for(cfg.S)
    if needs_grad:
        from .adjoint import fused_integrate_with_adjoint

        if P.stim_requires_grad or (st.syn is not None and st.syn["model"] not in (0, 1)):
            raise RuntimeError("the hand-written adjoint covers HH / Leak cells and networks with one "
                               "Kinetic or SigmoidRate synapse type (integrate routes other gradients "
                               "to the per-step path)")
        K = checkpoint_lengths[-1] if checkpoint_lengths else 100
        grad_keys = [k for k in ALL_PARAMS if k in inp.vals and inp.vals[k].requires_grad]
        recs, final = fused_integrate_with_adjoint(
            st, node_f, node_p, bparam, bmode, v0, gates0, stim, stim_scale, inp.T, dt, B, cfg, K, grad_keys,
            syn_params=syn_params, syn_scale=syn_scale, syn_s0=syn_s0)
    else:
        with torch.no_grad():
            run = ForwardLaunch(st, node_f, node_p.detach(), None if bparam is None else bparam.detach(),
                                bmode, v0.detach(), tuple(g.detach() for g in gates0), stim, stim_scale,
                                inp.T, dt, B, cfg, syn_params, syn_scale, syn_s0, **P.step_kw)
            run.launch()
            recs = run.recordings()
            final = run
    recs = recs.to(user_dtype).reshape(*batch, st.n_rec, inp.T + 1)
    if not return_states:
        return recs
    return recs, _final_states(module, st, final, batch, B, user_dtype, device)


def _final_states(module, st, run, batch, B, dtype, device):
    from .fused import from_prog

    mp = torch.as_tensor(st.mp, device=device)
    out = {}
    bufs = [("v", run.v), ("HH_m", run.gm), ("HH_h", run.gh), ("HH_n", run.gn)]
    if getattr(run, "gx", None) is not None:
        bufs += list(zip(POSP_STATES, run.gx))
    for name, buf in bufs:
        if name != "v" and name not in module._node_data:
            continue
        xs = from_prog(buf, B)
        x = torch.full((B, st.n), float("nan"), dtype=dtype, device=device)
        x[:, mp] = xs
        out[name] = x.reshape(*batch, st.n)
    if st.syn is not None and run.syn_s is not None:
        out.update(_syn_states_out(st, from_prog(run.syn_s, B), batch, device))
    return out


def _syn_state_key(group) -> Optional[str]:
    name = group["type"].state_name
    return None if name is None else f"{group['type'].name}_{name}"


def _syn_states_out(st, xs, batch, device, dtype=None) -> Dict[str, torch.Tensor]:
    """Per-type synapse states in edge order from the kernel's target-sorted states
    `xs` (B, n_syn)."""
    perm = torch.as_tensor(st.syn["perm"], device=device)
    inv = torch.empty_like(perm)
    inv[perm] = torch.arange(len(inv), device=device)
    xe = xs[:, inv]
    if dtype is not None:
        xe = xe.to(dtype)
    out = {}
    for g in st.syn["groups"]:
        key = _syn_state_key(g)
        if key is not None:
            out[key] = xe[:, g["offset"]:g["offset"] + g["n"]].reshape(*batch, -1)
    return out


class _StepRunner:
    """One fused step (T = 1) of a module with fixed structure, stimulus nodes and batch:
    the kernel's buffers persist between steps; every call copies the states in, runs the
    step on the current stream and copies them out (no host synchronization, so a step
    can be captured in a CUDA / HIP graph)."""

    def __init__(self, module, params, all_states, param_state, delta_t, ext_nodes, ext_batch,
                 solver: str = "bwd_euler"):
        ref = first_float_tensor(params, param_state, all_states)
        device = ref.device if ref is not None else torch.get_default_device()
        dtype = ref.dtype if ref is not None else torch.get_default_dtype()
        _require_gpu(device)
        ext_nodes = np.asarray(ext_nodes, dtype=np.int64)
        ds = None
        if len(ext_nodes):
            dummy = torch.zeros((*ext_batch, len(ext_nodes), 1), dtype=dtype, device=device)
            ds = (["i"], [dummy], [ext_nodes])
        inp = _assemble(module, params, param_state, ds, delta_t, delta_t, all_states, device, dtype,
                        module_stimuli=False)
        inp.T = 1
        if inp.stim is not None:
            inp.stim = inp.stim[..., :1]
        with torch.no_grad():
            # One program per group (a step may be captured in a graph: no cooperative
            # launch).
            P = _prepare(module, inp, delta_t, device, dtype, solver, split=False)
        from .autotune import fused_config

        cfg = fused_config(P.st, P.B, P.dtype, torch.tensor([], dtype=dtype).element_size(), False)
        self.P, self.cfg, self.device, self.dtype = P, cfg, device, dtype
        self.module, self.n_nodes = module, module._n_nodes
        self.run = ForwardLaunch(P.st, P.node_f, P.node_p, None if P.bmode == 0 else P.bparam_for(cfg.S),
                                 P.bmode, P.v0, tuple(P.gates0), P.stim, P.stim_scale, 1, delta_t, P.B, cfg,
                                 P.syn_params, P.syn_scale, P.syn_s0, **P.step_kw)
        self.mp = torch.as_tensor(P.st.mp, device=device)
        run = self.run
        bufs = [(run.gm, "HH_m"), (run.gh, "HH_h"), (run.gn, "HH_n")]
        if run.gx is not None:
            bufs += [(run.gx[q], name) for q, name in enumerate(POSP_STATES)]
        self.gate_bufs = [(b, k) for b, k in bufs if k in module._node_data]
        self.gate_names = [k for _, k in self.gate_bufs]
        comp_of_k = torch.as_tensor(P.st.comp_of_k, device=device)
        self.is_comp = comp_of_k >= 0
        if P.st.syn is not None:
            self.syn_perm = torch.as_tensor(P.st.syn["perm"], device=device)

    def step(self, all_states, externals):
        from .fused import from_prog

        P, run, B = self.P, self.run, self.P.B
        v = torch.as_tensor(all_states["v"], dtype=torch.float64, device=self.device)
        v = torch.broadcast_to(v, (*P.batch, self.n_nodes)).reshape(B, self.n_nodes)
        run.v.copy_(to_prog(v[:, self.mp], run.S))
        for buf, name in self.gate_bufs:
            if name in all_states:
                x = torch.as_tensor(all_states[name], device=self.device).to(run.dtype)
                x = torch.broadcast_to(x, (*P.batch, self.n_nodes)).reshape(B, self.n_nodes)[:, self.mp]
                x = torch.nan_to_num(x, nan=0.0)
                buf.copy_(to_prog(x, run.S))
        if run.syn_s is not None:
            groups = P.st.syn["groups"]
            keys = [_syn_state_key(g) for g in groups]
            if any(k is not None and k in all_states for k in keys):
                cols = []
                prev = from_prog(run.syn_s, B)
                inv = torch.empty_like(self.syn_perm)
                inv[self.syn_perm] = torch.arange(len(inv), device=self.device)
                prev = prev[:, inv]
                for g, key in zip(groups, keys):
                    if key is not None and key in all_states:
                        x = torch.as_tensor(all_states[key], device=self.device).to(run.dtype)
                        cols.append(torch.broadcast_to(x, (*P.batch, x.shape[-1])).reshape(B, -1))
                    else:
                        cols.append(prev[:, g["offset"]:g["offset"] + g["n"]])
                s = torch.cat(cols, 1)[:, self.syn_perm]
                run.syn_s.copy_(to_prog(s, run.S))
        if P.stim is not None:
            cur = torch.as_tensor(externals["i"], device=self.device).to(run.dtype)
            if run.stim.shape[2] == 1:
                cur = cur.reshape(-1, cur.shape[-1])[:1]
            else:
                cur = torch.broadcast_to(cur, (*P.batch, cur.shape[-1])).reshape(B, -1)
            run.stim.copy_(cur.t()[None])
        run.launch()
        out = dict(all_states)
        vk = from_prog(run.v, B)
        vn = torch.empty((B, self.n_nodes), dtype=self.dtype, device=self.device)
        vn[:, self.mp] = vk.to(self.dtype)
        out["v"] = vn.reshape(*P.batch, self.n_nodes)
        for buf, name in self.gate_bufs:
            gk = from_prog(buf, B).to(self.dtype)
            gk = torch.where(self.is_comp[None], gk, torch.full_like(gk, float("nan")))
            gn = torch.full((B, self.n_nodes), float("nan"), dtype=self.dtype, device=self.device)
            gn[:, self.mp] = gk
            if name in all_states:
                prev = torch.as_tensor(all_states[name], dtype=self.dtype, device=self.device)
                prev = torch.broadcast_to(prev, (*P.batch, self.n_nodes)).reshape(B, self.n_nodes)
                gn = torch.where(torch.isnan(prev), prev, gn)
            out[name] = gn.reshape(*P.batch, self.n_nodes)
        if run.syn_s is not None:
            out.update(_syn_states_out(P.st, from_prog(run.syn_s, B), P.batch, self.device, self.dtype))
        return out


def build_init_and_step_fn(module, voltage_solver: str = "jaxley.dhs", solver: str = "bwd_euler"):
    """`init_fn, step_fn` for step-by-step simulation (integrate.py:87-274).

    * `init_fn(params, all_states=None, param_state=None, delta_t=0.025)` returns
      `(all_states, all_params)`: the states over nodes (`v`, HH gates with NaN where
      the channel is absent, synapse states) and the parameters to pass to `step_fn`.
    * `step_fn(all_states, all_params, externals, external_inds, delta_t=0.025)` runs one
      backward-Euler (Crank-Nicolson, forward-Euler) step with the currents `externals["i"]` (`(..., n_inputs)`, nA) into
      the nodes `external_inds["i"]` and returns the new states.

    The step is one launch of the fused kernel with T = 1; its buffers are kept between
    calls (keyed by the stimulus nodes and batch shape), so a step can be captured in a
    CUDA / HIP graph. Gradients are not propagated through `step_fn` (the differentiable
    path is `integrate`)."""
    voltage_solver = _resolve_voltage_solver(module, voltage_solver)
    if voltage_solver not in _SUPPORTED_SOLVERS:
        raise NotImplementedError(f"voltage_solver={voltage_solver!r} is not part of jaxhip")
    if solver not in _SUPPORTED_STEPS:
        raise NotImplementedError(f"solver={solver!r}: supported are {_SUPPORTED_STEPS}")
    external_inds0 = module.external_inds.copy()
    cache: Dict[Any, _StepRunner] = {}

    def init_fn(params, all_states=None, param_state=None, delta_t: float = 0.025):
        ref = first_float_tensor(params, param_state, all_states, module.externals)
        device = ref.device if ref is not None else torch.get_default_device()
        dtype = ref.dtype if ref is not None else torch.get_default_dtype()
        _require_gpu(device)
        n = module._n_nodes
        comp_node = torch.as_tensor(module._comp_node, device=device)
        states = {}
        inp = _assemble(module, params, param_state, None, delta_t, delta_t, None, device, dtype,
                        module_stimuli=False)
        for k in ["v"] + [k for k in ("HH_m", "HH_h", "HH_n") + POSP_STATES if k in inp.vals]:
            x = inp.vals[k]
            out = torch.full((*x.shape[:-1], n), float("nan"), dtype=dtype, device=device)
            out[..., comp_node] = x
            if k == "v":   # branchpoints start at their parent compartment's voltage
                out = torch.nan_to_num(out, nan=float(x.reshape(-1)[0]))
            states[k] = out
        state_keys = {f"{e['type'].name}_{e['type'].state_name}" for e in module._edges
                      if e["type"].state_name is not None}
        for k, x in inp.syn_vals.items():
            if k in state_keys:
                states[k] = x
        if all_states is not None:
            states.update({k: torch.as_tensor(v, device=device) for k, v in all_states.items()})
        all_params = {"params": params, "param_state": param_state, "delta_t": delta_t}
        return states, all_params

    def step_fn(all_states, all_params, externals, external_inds=external_inds0, delta_t: float = 0.025):
        inds = np.zeros(0, np.int64)
        if "i" in externals:
            inds = np.asarray(external_inds["i"], dtype=np.int64)
        v = all_states["v"]
        batch = tuple(v.shape[:-1])
        if "i" in externals:
            batch = tuple(torch.broadcast_shapes(batch, tuple(torch.as_tensor(externals["i"]).shape[:-1])))
        key = (tuple(inds.tolist()), batch, float(delta_t), str(v.device), v.dtype, module._version,
               id(all_params))
        runner = cache.get(key)
        if runner is None:
            base = {k: torch.broadcast_to(torch.as_tensor(x), (*batch, x.shape[-1])).contiguous()
                    for k, x in all_states.items()}
            runner = _StepRunner(module, all_params["params"], base, all_params["param_state"],
                                 delta_t, inds, batch, solver)
            cache.clear()
            cache[key] = runner
        with torch.no_grad():
            return runner.step(all_states, externals)

    return init_fn, step_fn
