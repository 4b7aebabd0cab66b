"""Seeded workloads shared by the tests, the benchmark and the smoke check.

Each builder returns the jaxhip module and the matching oracle model (same morphology,
geometry, mechanisms, synapses; compartment indices agree).

* c0: one cell, root + full binary tree of depth 3 (15 branches), 4 compartments per
  branch, 10 um per branch (2.5 um per compartment), radius 1 um, HH everywhere;
  0.1 nA step into compartment 0 from 1 to 2 ms; dt 0.025 ms, 400 steps.
* c1: a seeded random tree of `nbranch` branches (each new branch on a uniformly chosen
  earlier branch), 4 compartments per branch, branch length lognormal (median 30 um,
  sigma 0.5), radius 2 um * 0.85^generation; HH with gNa, gK, gLeak drawn in +-50 % of
  the defaults per simulation; 0.2 nA step into the root from 1 to 21 ms.
* c4: ball-and-stick network (1-compartment HH soma, 8-compartment Leak dendrite),
  KineticSynapses with fixed in-degree onto one dendritic compartment per cell.
"""
from __future__ import annotations

import numpy as np
import torch

import jaxhip as jx
from oracle import jx_oracle as O

DT = 0.025


def binary_tree_parents(nbranches: int):
    """`_binary_tree_parents` of the reference's tests/conftest.py:25-31."""
    parents = [-1]
    depth = 0
    while nbranches > len(parents):
        parents = [-1] + [b // 2 for b in range(0, 2**depth - 2)]
        depth += 1
    return parents[:nbranches]


def c0_stimulus(t_steps: int = 400) -> np.ndarray:
    cur = np.zeros(t_steps)
    cur[int(1.0 / DT):int(2.0 / DT)] = 0.1
    return cur


def build_c0(device=None, dtype=torch.float64):
    parents = binary_tree_parents(15)
    cell = jx.Cell(jx.Branch(jx.Compartment(), 4), parents)
    cell.set("length", 2.5)
    cell.insert(jx.HH())
    cell.record()
    om = O.OracleModel([(parents, [4] * 15)])
    om.length = np.full(60, 2.5)
    om.hh_comps = np.arange(60)
    return cell, om


def random_tree(nbranch: int, seed: int):
    rng = np.random.default_rng(seed)
    parents = [-1] + [int(rng.integers(0, k)) for k in range(1, nbranch)]
    gen = np.zeros(nbranch, dtype=np.int64)
    for k in range(1, nbranch):
        gen[k] = gen[parents[k]] + 1
    blen = rng.lognormal(np.log(30.0), 0.5, size=nbranch)
    rad = 2.0 * 0.85 ** gen
    return parents, blen, rad


def build_tree(nbranch: int, seed: int, ncomp: int = 4):
    parents, blen, rad = random_tree(nbranch, seed)
    comp_len = np.repeat(blen / ncomp, ncomp)
    comp_rad = np.repeat(rad, ncomp)
    cell = jx.Cell(jx.Branch(jx.Compartment(), ncomp), parents)
    cell.set("length", comp_len)
    cell.set("radius", comp_rad)
    cell.insert(jx.HH())
    om = O.OracleModel([(parents, [ncomp] * nbranch)])
    om.length = comp_len
    om.radius = comp_rad
    om.hh_comps = np.arange(nbranch * ncomp)
    return cell, om


def c1_stimulus(t_steps: int) -> np.ndarray:
    cur = np.zeros(t_steps)
    cur[int(1.0 / DT):int(21.0 / DT)] = 0.2
    return cur


def c1_params(B: int, seed: int):
    rng = np.random.default_rng(seed + 1)
    scale = rng.uniform(0.5, 1.5, size=(B, 3))
    return {
        "HH_gNa": 0.12 * scale[:, 0:1],
        "HH_gK": 0.036 * scale[:, 1:2],
        "HH_gLeak": 0.0003 * scale[:, 2:3],
    }


def data_set_params(cell, params: dict, device, dtype):
    """`data_set` of per-simulation values `(B, 1)` on the whole cell."""
    ps = None
    for k, v in params.items():
        ps = cell.data_set(k, torch.as_tensor(v, dtype=dtype, device=device), ps)
    return ps


def oracle_overrides(om, params: dict):
    nc = len(om.internal)
    return {k: np.broadcast_to(v, (v.shape[0], nc)).copy() for k, v in params.items()}


def c2_specs(n_req: int, seed: int = 7):
    """Uneven-batch workload: `n_req` random trees with compartment counts log-uniform in
    [20, 2000] (4 per branch), per-request HH conductances in +-50 % of the defaults and
    t_max uniform in [10, 100] ms."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_req):
        ncomp = int(4 * round(np.exp(rng.uniform(np.log(20), np.log(2000))) / 4))
        scale = rng.uniform(0.5, 1.5, size=3)
        specs.append({
            "nbranch": max(ncomp // 4, 1),
            "seed": int(rng.integers(0, 2**31 - 1)),
            "t_max": float(rng.uniform(10.0, 100.0)),
            "params": {"HH_gNa": 0.12 * scale[0:1][None], "HH_gK": 0.036 * scale[1:2][None],
                       "HH_gLeak": 0.0003 * scale[2:3][None]},
        })
    return specs


def c2_request(spec, device, dtype, cell=None):
    """A served request for one spec (the cell built here unless given): 0.2 nA step into
    the root from 1 to 21 ms, voltage recorded at the root."""
    from jaxhip.serve import SimRequest

    if cell is None:
        cell, _ = build_tree(spec["nbranch"], spec["seed"])
        cell.branch(0).comp(0).record()
    T = int(spec["t_max"] // DT + 1)
    cur = torch.tensor(c1_stimulus(T), dtype=dtype, device=device)
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    ps = data_set_params(cell, spec["params"], device, dtype)
    return SimRequest(cell, t_max=spec["t_max"], param_state=ps, data_stimuli=ds)


def build_c4(n_cells: int, k: int, seed: int, synapse: str = "KineticSynapse"):
    """Ball-and-stick network: per cell a 1-compartment soma (radius 10 um, length 20 um,
    the area of a 10 um sphere) with HH and an 8-compartment dendrite (200 um, radius
    1 um) with Leak. Every cell receives k synapses from uniformly drawn presynaptic
    somata, all onto one uniformly drawn dendritic compartment; synapses are sorted by
    target. KineticSynapse / IonotropicSynapse: gS ~ U(0, 1e-3) uS, e_syn 0 mV (80 % of
    presynaptic cells) or -80 mV; SigmoidRateSynapse: gS ~ U(0, 1e-3), default offset and
    slope.
    Returns (network, oracle model, info)."""
    rng = np.random.default_rng(seed)
    cell = jx.Cell([jx.Branch(jx.Compartment(), 1), jx.Branch(jx.Compartment(), 8)], [-1, 0])
    length = np.array([20.0] + [25.0] * 8)
    radius = np.array([10.0] + [1.0] * 8)
    cell.set("length", length)
    cell.set("radius", radius)
    cell.branch(0).insert(jx.HH())
    cell.branch(1).insert(jx.Leak())
    net = jx.Network([cell] * n_cells)
    excit = rng.uniform(size=n_cells) < 0.8
    pre_cells = rng.integers(0, n_cells, size=(n_cells, k))
    tgt_comp = rng.integers(1, 9, size=n_cells)
    pre_comp = (pre_cells * 9).reshape(-1)
    post_comp = np.repeat(np.arange(n_cells) * 9 + tgt_comp, k)
    syn = {"KineticSynapse": jx.KineticSynapse, "SigmoidRateSynapse": jx.SigmoidRateSynapse,
           "IonotropicSynapse": jx.IonotropicSynapse}[synapse]()
    net.add_synapses(net._comp_node[pre_comp], net._comp_node[post_comp], syn)
    E = len(pre_comp)
    getattr(net, syn.name).set(f"{syn.name}_gS", rng.uniform(0.0, 1e-3, size=E))
    if synapse in ("KineticSynapse", "IonotropicSynapse"):
        e_syn = np.where(excit[pre_cells.reshape(-1)], 0.0, -80.0)
        getattr(net, syn.name).set(f"{syn.name}_e_syn", e_syn)
    edge = net._edges[0]
    params = {name: edge["params"][f"{syn.name}_{name}"].copy() for name in syn.param_names}
    om = O.OracleModel([([-1, 0], [1, 8])] * n_cells)
    om.length = np.tile(length, n_cells)
    om.radius = np.tile(radius, n_cells)
    om.hh_comps = np.arange(n_cells) * 9
    om.leak_comps = np.setdiff1d(np.arange(9 * n_cells), om.hh_comps)
    s0 = edge["states"][f"{syn.name}_s"].copy() if synapse != "SigmoidRateSynapse" else None
    om.synapses = [{"type": synapse, "pre": om.internal[pre_comp], "post": om.internal[post_comp],
                    "params": params, "s0": s0}]
    info = {"soma_comps": np.arange(n_cells) * 9, "pre_comp": pre_comp, "post_comp": post_comp}
    return net, om, info


def c4_stimulus(B: int, n_cells: int, T: int, seed: int, amp: float = 0.05, p: float = 0.01):
    """Bernoulli(p) pulses of `amp` nA per soma and step, (B, n_cells, T)."""
    rng = np.random.default_rng(seed)
    return amp * (rng.uniform(size=(B, n_cells, T)) < p)


def oracle_of_swc(cells):
    """The oracle's model of cells read by `jx.read_swc` (same parents and traced
    compartment geometry), HH in every compartment."""
    om = O.OracleModel([(np.asarray(c._cell_specs[0][0]), np.asarray(c._cell_specs[0][1])) for c in cells])
    cat = lambda k: np.concatenate([c._node_data[k] for c in cells])
    om.length[:], om.radius[:] = cat("length"), cat("radius")
    om.morph = {k: cat(k) for k in ("area", "resistive_load_in", "resistive_load_out")}
    om.hh_comps = np.arange(len(om.internal))
    return om


def oracle_of_module(mod):
    """The oracle's model of a `jaxhip` cell or network (structure, geometry, HH, Leak
    and Pospischil channels with the module's parameters and initial states)."""
    om = O.OracleModel([(np.asarray(p), np.asarray(nc)) for p, nc in mod._cell_specs])
    d = mod._node_data
    for k in ("length", "radius", "axial_resistivity", "capacitance"):
        getattr(om, k)[:] = d[k]
    om.morph = {k: d[k].copy() for k in ("area", "resistive_load_in", "resistive_load_out")}
    om.v0 = d["v"].copy()
    mask = mod._channel_mask
    if "HH" in mask:
        om.hh_comps = np.nonzero(mask["HH"])[0]
        om.hh_params = {k: np.nan_to_num(d[k]) for k in om.hh_params}
        om.gates0 = tuple(np.nan_to_num(d[k]) for k in ("HH_m", "HH_h", "HH_n"))
    if "Leak" in mask:
        om.leak_comps = np.nonzero(mask["Leak"])[0]
        om.leak_params = {k: np.nan_to_num(d[k]) for k in om.leak_params}
    for name in ("Na", "K", "Km"):
        if name in mask:
            om.posp_comps[name] = np.nonzero(mask[name])[0]
    for k in om.posp_params:
        if k in d:
            om.posp_params[k] = np.nan_to_num(d[k])
    for k in om.posp_states0:
        if k in d:
            om.posp_states0[k] = np.nan_to_num(d[k])
    return om


def pospischil_cell():
    """The reference's channel steady-state cell (tests/test_channels.py:181-208) without
    its calcium channels: two branches of four compartments, Na on branch 0, K on
    branch 1, Km at branch 1's first compartment, Leak and HH everywhere."""
    cell = jx.Cell(jx.Branch(jx.Compartment(), 4), parents=[-1, 0])
    cell.branch(0).loc(0.0).record()
    cell.branch(0).insert(jx.Na())
    cell.branch(1).insert(jx.K())
    cell.branch(1).loc(0.0).insert(jx.Km())
    cell.insert(jx.Leak())
    cell.insert(jx.HH())
    return cell
