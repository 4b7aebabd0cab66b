"""Compartment geometry and axial conductances (once per simulation, on the device).

`cylinder_area` / `cylinder_resistive_load` (utils/morph_attributes.py:261-312, with the
half-length loads of modules/compartment.py:70-78) and `compute_axial_conductances`
(utils/morph_attributes.py:315-436) for the voltage:
  type 0 (comp -> comp): g = 1e7 / (r_a(src) + r_a(sink)) / area(sink) / cm(sink), with
      r_a(src) the out-load (forward edge) or in-load (backward edge) of the source and
      r_a(sink) the in-load (forward) or out-load (backward) of the sink, times R_a;
  types 1, 2 (branchpoint -> comp): g = 1e7 / r_a(sink) / area(sink) / cm(sink), the
      sink's in-load for `ordered` edges, else its out-load;
  types 3, 4 (comp -> branchpoint): g = 1 / (r_l(src) R_a(src)) (not / area, / cm) with
      the load of the side the mirrored type 1 / 2 edge uses.
Leading batch dimensions broadcast. Unbatched CPU inputs without autograd take a numpy
form of the same arithmetic (request preparation on the host).
"""
from __future__ import annotations

import math
from typing import Dict

import numpy as np
import torch


GEOMETRY_KEYS = ("area", "volume", "resistive_load_in", "resistive_load_out")


def cylinder_geometry(length, radius) -> Dict:
    """Area, volume and the resistive loads of the two halves of cylinders (numpy or
    torch; utils/morph_attributes.py:261-312 with l / 2, modules/compartment.py:70-78)."""
    rl = (length / 2) / radius**2 / math.pi
    return {"area": 2.0 * math.pi * radius * length, "volume": length * radius**2 * math.pi,
            "resistive_load_in": rl, "resistive_load_out": rl}


def comp_geometry(vals: Dict[str, torch.Tensor]) -> Dict[str, torch.Tensor]:
    """Per-compartment area and input / output loads of the simulation's parameters."""
    return {"area": vals["area"], "rl_in": vals["resistive_load_in"], "rl_out": vals["resistive_load_out"]}


def _host_values(*ts) -> bool:
    """Unbatched CPU tensors without autograd (a served request's preparation): the
    tables are computed with numpy, the same operations in the same order."""
    return all(t.device.type == "cpu" and t.ndim == 1 and not t.requires_grad for t in ts)


def _axial_conductances_np(tables, comp_node, n_nodes, rl_in, rl_out, area, Ra, cm) -> np.ndarray:
    typ, src, snk, ordd = tables["type"], tables["source"], tables["sink"], tables["ordered"]
    comp_of_node = -np.ones(n_nodes, dtype=np.int64)
    comp_of_node[comp_node] = np.arange(len(comp_node))
    out = np.zeros(len(typ), dtype=area.dtype)
    rl = np.stack([rl_in, rl_out])
    side = lambda o, c: rl[o, c] * Ra[c]
    e0 = np.nonzero(typ == 0)[0]
    if len(e0):
        cs, ct = comp_of_node[src[e0]], comp_of_node[snk[e0]]
        r = side(ordd[e0], cs) + side(1 - ordd[e0], ct)
        out[e0] = 1.0 / r / area[ct] / cm[ct] * 1e7
    e12 = np.nonzero((typ == 1) | (typ == 2))[0]
    if len(e12):
        ct = comp_of_node[snk[e12]]
        out[e12] = 1.0 / side(1 - ordd[e12], ct) / area[ct] / cm[ct] * 1e7
    e34 = np.nonzero((typ == 3) | (typ == 4))[0]
    if len(e34):
        cs = comp_of_node[src[e34]]
        out[e34] = 1.0 / rl[1 - ordd[e12], cs] * (1.0 / Ra[cs])
    return out


def axial_conductances(tables: Dict, comp_node: np.ndarray, n_nodes: int, geo: Dict[str, torch.Tensor],
                       axial_resistivity: torch.Tensor, capacitance: torch.Tensor) -> torch.Tensor:
    if _host_values(geo["rl_in"], geo["rl_out"], geo["area"], axial_resistivity, capacitance):
        return torch.from_numpy(_axial_conductances_np(
            tables, comp_node, n_nodes, geo["rl_in"].numpy(), geo["rl_out"].numpy(), geo["area"].numpy(),
            axial_resistivity.numpy(), capacitance.numpy()))
    typ, src, snk, ordd = tables["type"], tables["source"], tables["sink"], tables["ordered"]
    dev = axial_resistivity.device
    comp_of_node = -np.ones(n_nodes, dtype=np.int64)
    comp_of_node[comp_node] = np.arange(len(comp_node))
    area = geo["area"]
    Ra = axial_resistivity
    cm = capacitance
    batch = torch.broadcast_shapes(geo["rl_in"].shape[:-1], geo["rl_out"].shape[:-1], area.shape[:-1],
                                   Ra.shape[:-1], cm.shape[:-1])
    out = torch.zeros((*batch, len(typ)), dtype=area.dtype, device=dev)
    # Index 0: the load of the compartment's first half (in), 1: of its second (out).
    rl = torch.stack(torch.broadcast_tensors(geo["rl_in"], geo["rl_out"]), dim=-2)
    T = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int64), device=dev)
    side = lambda o, c: rl[..., T(o), T(c)] * Ra[..., T(c)]
    e0 = np.nonzero(typ == 0)[0]
    if len(e0):
        cs, ct = comp_of_node[src[e0]], comp_of_node[snk[e0]]
        r = side(ordd[e0], cs) + side(1 - ordd[e0], ct)
        val = 1.0 / r / area[..., T(ct)] / cm[..., T(ct)] * 1e7
        out = out.index_copy(-1, T(e0), torch.broadcast_to(val, (*batch, len(e0))))
    e12 = np.nonzero(np.isin(typ, [1, 2]))[0]
    if len(e12):
        ct = comp_of_node[snk[e12]]
        val = 1.0 / side(1 - ordd[e12], ct) / area[..., T(ct)] / cm[..., T(ct)] * 1e7
        out = out.index_copy(-1, T(e12), torch.broadcast_to(val, (*batch, len(e12))))
    e34 = np.nonzero(np.isin(typ, [3, 4]))[0]
    if len(e34):
        # Rows of types 3 and 4 mirror those of types 1 and 2 in the same order.
        cs = comp_of_node[src[e34]]
        val = 1.0 / rl[..., T(1 - ordd[e12]), T(cs)] * (1.0 / Ra[..., T(cs)])
        out = out.index_copy(-1, T(e34), torch.broadcast_to(val, (*batch, len(e34))))
    return out
