"""Seeded workloads of the tests and the smoke check: the package's workloads
(`jaxhip.workloads`) with the matching oracle model (same morphology, geometry,
mechanisms, synapses; compartment indices agree).

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

from jaxhip.workloads import (DT, binary_tree_parents, c0_stimulus, c1_params, c1_stimulus,  # noqa: F401
                              c2_request, c2_specs, c4_stimulus, data_set_params, random_tree)
from jaxhip import workloads as W


def build_c0(device=None, dtype=torch.float64):
    cell, parents = W.c0_cell()
    om = O.OracleModel([(parents, [4] * 15)])
    om.length = np.full(60, 2.5)
    om.hh_comps = np.arange(60)
    return cell, om


def build_tree(nbranch: int, seed: int, ncomp: int = 4):
    cell, parents, comp_len, comp_rad = W.tree_cell(nbranch, seed, ncomp)
    om = O.OracleModel([(parents, [ncomp] * nbranch)])
    om.length = comp_len
    om.radius = comp_rad
    om.hh_comps = np.arange(nbranch * ncomp)
    return cell, om


def oracle_overrides(om, params: dict):
    nc = len(om.internal)
    return {k: np.broadcast_to(v, (v.shape[0], nc)).copy() for k, v in params.items()}


def build_c4(n_cells: int, k: int, seed: int, synapse: str = "KineticSynapse"):
    """`jaxhip.workloads.c4_network` and its oracle model. Returns (network, oracle
    model, info)."""
    net, info = W.c4_network(n_cells, k, seed, synapse)
    syn = info["synapse"]
    edge = net._edges[0]
    params = {name: edge["params"][f"{syn.name}_{name}"].copy() for name in syn.param_names}
    om = O.OracleModel([([-1, 0], [1, 8])] * n_cells)
    om.length = np.tile(W.C4_LENGTH, n_cells)
    om.radius = np.tile(W.C4_RADIUS, n_cells)
    om.hh_comps = np.arange(n_cells) * 9
    om.leak_comps = np.setdiff1d(np.arange(9 * n_cells), om.hh_comps)
    s0 = edge["states"][f"{syn.name}_s"].copy() if synapse != "SigmoidRateSynapse" else None
    pre_comp, post_comp = info["pre_comp"], info["post_comp"]
    om.synapses = [{"type": synapse, "pre": om.internal[pre_comp], "post": om.internal[post_comp],
                    "params": params, "s0": s0}]
    return net, om, info


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
