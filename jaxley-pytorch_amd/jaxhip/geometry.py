"""Compartment geometry and axial conductances (once per simulation, on the device).

`cylinder_area` / `cylinder_resistive_load` (utils/morph_attributes.py:261-312, with the
half-length loads of modules/compartment.py:70-78) and `compute_axial_conductances`
(utils/morph_attributes.py:315-436) for the voltage:
  type 0 (comp -> comp): g = 1e7 / (r_a(src) + r_a(sink)) / area(sink) / cm(sink), with
      r_a(src) the out-load (forward edge) or in-load (backward edge) times R_a;
  types 1, 2 (branchpoint -> comp): g = 1e7 / r_a(sink) / area(sink) / cm(sink);
  types 3, 4 (comp -> branchpoint): g = 1 / (r_l(src) R_a(src)) (not / area, / cm).
Leading batch dimensions broadcast.
"""
from __future__ import annotations

import math
from typing import Dict

import numpy as np
import torch


def comp_geometry(length: torch.Tensor, radius: torch.Tensor) -> Dict[str, torch.Tensor]:
    area = 2.0 * math.pi * radius * length
    rl = (length / 2) / radius**2 / math.pi
    return {"area": area, "rl": rl}


def axial_conductances(tables: Dict, comp_node: np.ndarray, n_nodes: int, geo: Dict[str, torch.Tensor],
                       axial_resistivity: torch.Tensor, capacitance: torch.Tensor) -> torch.Tensor:
    typ, src, snk, ordd = tables["type"], tables["source"], tables["sink"], tables["ordered"]
    dev = axial_resistivity.device
    comp_of_node = -np.ones(n_nodes, dtype=np.int64)
    comp_of_node[comp_node] = np.arange(len(comp_node))
    rl = geo["rl"]
    area = geo["area"]
    Ra = axial_resistivity
    cm = capacitance
    batch = torch.broadcast_shapes(rl.shape[:-1], area.shape[:-1], Ra.shape[:-1], cm.shape[:-1])
    out = torch.zeros((*batch, len(typ)), dtype=rl.dtype, device=dev)
    ra = rl * Ra  # in- and out-loads are equal for cylinders
    T = lambda a: torch.as_tensor(np.asarray(a, dtype=np.int64), device=dev)
    e0 = np.nonzero(typ == 0)[0]
    if len(e0):
        cs, ct = comp_of_node[src[e0]], comp_of_node[snk[e0]]
        r = ra[..., T(cs)] + ra[..., T(ct)]
        val = 1.0 / r / area[..., T(ct)] / cm[..., T(ct)] * 1e7
        out = out.index_copy(-1, T(e0), torch.broadcast_to(val, (*batch, len(e0))))
    e12 = np.nonzero(np.isin(typ, [1, 2]))[0]
    if len(e12):
        ct = comp_of_node[snk[e12]]
        val = 1.0 / ra[..., T(ct)] / area[..., T(ct)] / cm[..., T(ct)] * 1e7
        out = out.index_copy(-1, T(e12), torch.broadcast_to(val, (*batch, len(e12))))
    e34 = np.nonzero(np.isin(typ, [3, 4]))[0]
    if len(e34):
        cs = comp_of_node[src[e34]]
        val = 1.0 / rl[..., T(cs)] * (1.0 / Ra[..., T(cs)])
        out = out.index_copy(-1, T(e34), torch.broadcast_to(val, (*batch, len(e34))))
    return out
