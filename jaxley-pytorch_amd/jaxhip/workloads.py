"""Seeded synthetic workloads of the hot path (the benchmark's and the tests' inputs).

* `c0_cell`: one cell, root + full binary tree of depth 3 (15 branches), 4 compartments
  per branch, 10 um per branch (2.5 um per compartment), radius 1 um, HH everywhere;
  `c0_stimulus`: 0.1 nA from 1 to 2 ms; dt 0.025 ms, 400 steps.
* `tree_cell`: a seeded random tree of `nbranch` branches (each new branch on a uniformly
  chosen earlier branch), 4 compartments per branch, branch length lognormal (median 30
  um, sigma 0.5), radius 2 um * 0.85^generation, HH everywhere; `c1_params`: gNa, gK,
  gLeak drawn in +-50 % of the defaults per simulation; `c1_stimulus`: 0.2 nA into the
  root from 1 to 21 ms.
* `c2_specs` / `c2_request`: uneven served requests, compartment counts log-uniform in
  [20, 2000], per-request conductances and durations.
* `c4_network`: ball-and-stick cells (1-compartment HH soma, 8-compartment Leak
  dendrite), synapses with a fixed in-degree onto one dendritic compartment per cell;
  `c4_stimulus`: Bernoulli current pulses into every soma.

Every draw uses `numpy.random.default_rng(seed)`.
"""
from __future__ import annotations

import numpy as np
import torch

from .modules import Branch, Cell, Compartment, Network
from .channels import HH, Leak
from .synapses import IonotropicSynapse, KineticSynapse, SigmoidRateSynapse

DT = 0.025


def binary_tree_parents(nbranches: int):
    """Parents of a root plus a full binary tree (`_binary_tree_parents` of the
    reference's tests/conftest.py:25-31)."""
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


def c0_cell():
    parents = binary_tree_parents(15)
    cell = Cell(Branch(Compartment(), 4), parents)
    cell.set("length", 2.5)
    cell.insert(HH())
    cell.record()
    return cell, parents


def random_tree(nbranch: int, seed: int):
    """Parents, branch lengths (um) and radii (um) of the seeded random tree."""
    rng = np.random.default_rng(seed)
    parents = [-1] + [int(rng.integers(0, k)) for k in range(1, nbranch)]
    gen = np.zeros(nbranch, dtype=np.int64)
    for k in range(1, nbranch):
        gen[k] = gen[parents[k]] + 1
    blen = rng.lognormal(np.log(30.0), 0.5, size=nbranch)
    rad = 2.0 * 0.85 ** gen
    return parents, blen, rad


def tree_cell(nbranch: int, seed: int, ncomp: int = 4):
    """The seeded random tree as a cell with HH; returns (cell, parents, per-compartment
    lengths, per-compartment radii)."""
    parents, blen, rad = random_tree(nbranch, seed)
    comp_len = np.repeat(blen / ncomp, ncomp)
    comp_rad = np.repeat(rad, ncomp)
    cell = Cell(Branch(Compartment(), ncomp), parents)
    cell.set("length", comp_len)
    cell.set("radius", comp_rad)
    cell.insert(HH())
    return cell, parents, comp_len, comp_rad


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
    from .serve import SimRequest

    if cell is None:
        cell = tree_cell(spec["nbranch"], spec["seed"])[0]
        cell.branch(0).comp(0).record()
    T = int(spec["t_max"] // DT + 1)
    cur = torch.tensor(c1_stimulus(T), dtype=dtype, device=device)
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    ps = data_set_params(cell, spec["params"], device, dtype)
    return SimRequest(cell, t_max=spec["t_max"], param_state=ps, data_stimuli=ds)


C4_LENGTH = np.array([20.0] + [25.0] * 8)
C4_RADIUS = np.array([10.0] + [1.0] * 8)


def c4_network(n_cells: int, k: int, seed: int, synapse: str = "KineticSynapse"):
    """Ball-and-stick network: per cell a 1-compartment soma (radius 10 um, length 20 um,
    the area of a 10 um sphere) with HH and an 8-compartment dendrite (200 um, radius
    1 um) with Leak. Every cell receives k synapses from uniformly drawn presynaptic
    somata, all onto one uniformly drawn dendritic compartment. KineticSynapse /
    IonotropicSynapse: gS ~ U(0, 1e-3) uS, e_syn 0 mV (80 % of presynaptic cells) or
    -80 mV; SigmoidRateSynapse: gS ~ U(0, 1e-3), default offset and slope.
    Returns (network, info) with the compartments of somata, synapse ends and
    parameters."""
    rng = np.random.default_rng(seed)
    cell = Cell([Branch(Compartment(), 1), Branch(Compartment(), 8)], [-1, 0])
    cell.set("length", C4_LENGTH)
    cell.set("radius", C4_RADIUS)
    cell.branch(0).insert(HH())
    cell.branch(1).insert(Leak())
    net = Network([cell] * n_cells)
    excit = rng.uniform(size=n_cells) < 0.8
    pre_cells = rng.integers(0, n_cells, size=(n_cells, k))
    tgt_comp = rng.integers(1, 9, size=n_cells)
    pre_comp = (pre_cells * 9).reshape(-1)
    post_comp = np.repeat(np.arange(n_cells) * 9 + tgt_comp, k)
    syn = {"KineticSynapse": KineticSynapse, "SigmoidRateSynapse": SigmoidRateSynapse,
           "IonotropicSynapse": IonotropicSynapse}[synapse]()
    net.add_synapses(net._comp_node[pre_comp], net._comp_node[post_comp], syn)
    E = len(pre_comp)
    getattr(net, syn.name).set(f"{syn.name}_gS", rng.uniform(0.0, 1e-3, size=E))
    if synapse in ("KineticSynapse", "IonotropicSynapse"):
        e_syn = np.where(excit[pre_cells.reshape(-1)], 0.0, -80.0)
        getattr(net, syn.name).set(f"{syn.name}_e_syn", e_syn)
    info = {"soma_comps": np.arange(n_cells) * 9, "pre_comp": pre_comp, "post_comp": post_comp,
            "synapse": syn}
    return net, info


def c4_stimulus(B: int, n_cells: int, T: int, seed: int, amp: float = 0.05, p: float = 0.01):
    """Bernoulli(p) pulses of `amp` nA per soma and step, (B, n_cells, T)."""
    rng = np.random.default_rng(seed)
    return amp * (rng.uniform(size=(B, n_cells, T)) < p)
