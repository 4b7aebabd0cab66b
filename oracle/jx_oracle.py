"""CPU oracle: a from-scratch restatement of the per-step path of jaxley-pytorch's
`jx.integrate` (float64, PyTorch on CPU, numpy / pure Python for the integer work).

TEST INFRASTRUCTURE ONLY (see `oracle/__init__.py`): it is the checker of the HIP
path, never the thing measured or shipped.

Every function cites the reference file:line it restates (paths relative to the
reference checkout). The restatement is literal where bit-exactness matters (level
schedules, synapse ordering) and follows the reference's op order for the floating
point work (finite-difference linearization with `diff = 1e-3`, `save_exp` clamps,
the `x_over_expm1` series switch, level-by-level `index_add` elimination).

Pinning (what checks each part; tests/test_oracle.py, tests/test_host.py):
  * morphology tables, axial conductances, HH, stimulus, implicit Euler step and the
    DHS solve: the reference's golden traces (tests/jaxley_identical/*), restated as
    fixtures under tests/golden/, at 1e-8 mV;
  * DHS schedule: the BFS restatement is checked against networkx's own
    `to_undirected` / `bfs_edges` on random trees (networkx is the third-party
    dependency that fixes the reference's order) and against the hand-built ragged
    indexer of tests/test_solver.py:78-86;
  * tree solve: dense `numpy.linalg.solve` (closed form) and its transpose;
  * IonotropicSynapse + network indexer: golden `test_net` trace;
  * KineticSynapse / SigmoidRateSynapse (tests/closed_forms.py): the state update
    against the analytic solution of its linear ODE at a constant presynaptic voltage
    (400 steps) and first-order convergence to a Runge-Kutta solution with a varying
    one; the sigmoid's symmetry and midpoint; the finite-difference slope against the
    analytic derivative; whole two-cell networks through `simulate` against the
    closed-form backward-Euler recurrence of the postsynaptic membrane; the golden
    network path through IonotropicSynapse exercises the same linearization;
  * vectors written by the reference itself (tests/golden/reference_vectors.npz,
    tests/test_reference_vectors.py): the DHS indexers of a 60-compartment cell, a
    1000-compartment random tree and a 6-cell network (bit-exact), and 30 ms traces of a
    ball-and-stick cell and of 6-cell networks with KineticSynapse / SigmoidRateSynapse
    (equal bit for bit), and the reference's autograd gradients of a tree's per-branch
    HH conductances and of the Kinetic network's gS and gNa (3e-13).
"""
from __future__ import annotations

import math
from collections import deque
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

F64 = torch.float64

# --------------------------------------------------------------------------------------
# Morphology tables
# --------------------------------------------------------------------------------------


def cell_tables(parents: Sequence[int], ncomp_per_branch: Sequence[int]) -> Dict:
    """Compartment edges of one cell, in the reference's row order.

    Restates `Cell._init_comp_graph` (modules/cell.py:145-225), `Branch._init_comp_graph`
    (modules/branch.py:89-107) and `compute_children_and_parents`
    (utils/cell_utils.py:306-314, `remap_to_consecutive` :218-224).
    Edge types: 0 comp<->comp, 1 bp->parent comp, 2 bp->child comp, 3 parent comp->bp,
    4 child comp->bp.
    """
    parents = np.asarray(parents, dtype=np.int64)
    ncomp = np.asarray(ncomp_per_branch, dtype=np.int64)
    nb = len(parents)
    cs = np.concatenate([[0], np.cumsum(ncomp)]).astype(np.int64)
    n_comp = int(cs[-1])
    src: List[int] = []
    snk: List[int] = []
    ordd: List[int] = []
    typ: List[int] = []
    for b in range(nb):
        n, c = int(ncomp[b]), int(cs[b])
        src += list(range(c, c + n - 1)) + list(range(c + 1, c + n))
        snk += list(range(c + 1, c + n)) + list(range(c, c + n - 1))
        ordd += [1] * (n - 1) + [0] * (n - 1)
        typ += [0] * (2 * (n - 1))
    par = parents[1:]
    child = np.arange(1, nb)
    if len(par) > 0:
        par_u, child_bp = np.unique(par, return_inverse=True)
    else:
        par_u, child_bp = np.zeros(0, np.int64), np.zeros(0, np.int64)
    nbp = len(par_u)
    t1_src = np.arange(nbp) + n_comp
    t1_snk = cs[par_u + 1] - 1
    t2_src = child_bp + n_comp
    t2_snk = cs[child]
    src += list(t1_src) + list(t2_src) + list(t1_snk) + list(t2_snk)
    snk += list(t1_snk) + list(t2_snk) + list(t1_src) + list(t2_src)
    ordd += [0] * nbp + [1] * len(child) + [1] * nbp + [0] * len(child)
    typ += [1] * nbp + [2] * len(child) + [3] * nbp + [4] * len(child)
    n_nodes = int(max(snk) + 1) if len(snk) > 0 else 1
    return {
        "source": np.asarray(src, dtype=np.int64),
        "sink": np.asarray(snk, dtype=np.int64),
        "ordered": np.asarray(ordd, dtype=np.int64),
        "type": np.asarray(typ, dtype=np.int64),
        "n_comp": n_comp,
        "n_nodes": n_nodes,
        "internal": np.arange(n_comp, dtype=np.int64),
        "ncomp_per_branch": ncomp,
        "parents": parents,
    }


# --------------------------------------------------------------------------------------
# DHS schedule (integer work: must be bit-exact)
# --------------------------------------------------------------------------------------


def _undirected_adjacency(n_nodes: int, tables: Dict) -> Dict[int, List[int]]:
    """Neighbour order of networkx's `DiGraph.to_undirected()` applied to the graph of
    `to_graph` (modules/base.py:4418-4456): nodes in `.nodes` then `_branchpoints`
    order; directed edges of type 2, 3 and forward type 0, in row order. networkx adds
    `(u, v)` for `u` in node order and `v` in successor insertion order; a new neighbour
    is appended to both adjacency dicts, an existing one keeps its position."""
    typ, s, t = tables["type"], tables["source"], tables["sink"]
    keep = np.isin(typ, [2, 3]) | ((typ == 0) & (s < t))
    succ: Dict[int, Dict[int, None]] = {i: {} for i in range(n_nodes)}
    for u, v in zip(s[keep].tolist(), t[keep].tolist()):
        succ[u][v] = None
    adj: Dict[int, Dict[int, None]] = {i: {} for i in range(n_nodes)}
    for u in range(n_nodes):
        for v in succ[u]:
            adj[u][v] = None
            adj[v][u] = None
    return {k: list(v.keys()) for k, v in adj.items()}


def bfs_edge_hops(adj: Dict[int, List[int]], root: int, allowed_nodes_per_level: int):
    """Restates `bfs_edge_hops` (utils/solver_utils.py:187-223) over networkx's
    `generic_bfs_edges` queue order and `single_source_shortest_path_length`."""
    dist = {root: 0}
    order = []
    seen = {root}
    queue = deque([root])
    while queue:
        u = queue.popleft()
        for v in adj[u]:
            if v not in seen:
                seen.add(v)
                dist[v] = dist[u] + 1
                order.append((u, v))
                queue.append(v)
    current_depth = 0
    since = 0
    last_true = 0
    out = []
    for u, v in order:
        if dist[v] > last_true:
            current_depth += 1
            since = 0
            last_true = dist[v]
        elif since >= allowed_nodes_per_level:
            current_depth += 1
            since = 0
        since += 1
        out.append((u, v, current_depth))
    return out


def group_levels(node_order: np.ndarray) -> List[np.ndarray]:
    """`dhs_group_comps_into_levels` (utils/solver_utils.py:226-251): pandas groupby on
    the level column (sorted levels, rows in original order)."""
    node_order = np.asarray(node_order, dtype=np.int64).reshape(-1, 3)
    if len(node_order) == 0:
        return []
    out = []
    for lev in sorted(set(node_order[:, 2].tolist())):
        rows = node_order[node_order[:, 2] == lev]
        out.append(np.stack([rows[:, 0], rows[:, 1]]).T)
    return out


def dhs_indexer_cell(tables: Dict, allowed_nodes_per_level: int = 32, root: int = 0) -> Dict:
    """`Module._init_solver_jaxley_dhs_solve` (modules/base.py:1142-1266) with
    `dhs_solve_index` (utils/solver_utils.py:135-184) and `dhs_permutation_indices`
    (utils/solver_utils.py:80-132)."""
    n = tables["n_nodes"]
    adj = _undirected_adjacency(n, tables)
    hops = bfs_edge_hops(adj, root, allowed_nodes_per_level)
    solve_index = {root: 0}
    node_and_parent = [(root, -1, 0)]
    for k, (i, j, lev) in enumerate(hops):
        solve_index[j] = k + 1
        node_and_parent.append((j, i, lev))
    inv = np.array([solve_index[i] for i in sorted(solve_index)], dtype=np.int64)
    mapping = np.argsort(inv)
    node_order = node_and_parent[1:]
    # Permutation of the lower / upper values (edge k has value k).
    offdiag = np.stack([tables["source"], tables["sink"]])
    edge_data = {}
    for k, e in enumerate(offdiag.T.tolist()):
        edge_data[tuple(e)] = k
    ordered = {}
    for (i, j), k in edge_data.items():
        ordered[(solve_index[i], solve_index[j])] = k
    lowers, uppers = {}, {}
    for ij, k in ordered.items():
        if ij[0] < ij[1]:
            lowers[ij] = k
        else:
            uppers[ij] = k
    lower_sorting = np.argsort([key[1] for key in lowers.keys()])
    upper_sorting = np.argsort([key[0] for key in uppers.keys()])
    lower_vals = np.asarray(list(lowers.values()), dtype=np.int64)[lower_sorting] if lowers else np.zeros(0, np.int64)
    upper_vals = np.asarray(list(uppers.values()), dtype=np.int64)[upper_sorting] if uppers else np.zeros(0, np.int64)
    new_node_order = np.asarray(
        [[solve_index[int(a)], solve_index[int(b)], int(c)] for a, b, c in node_order],
        dtype=np.int64,
    ).reshape(-1, 3)
    lu = np.concatenate([lower_vals, upper_vals])
    grouped = group_levels(new_node_order)
    parent_lookup = -np.ones(n + 1, dtype=np.int64)
    for g in grouped:
        parent_lookup[g[:, 0]] = g[:, 1]
    return {
        "map_to_solve_order": mapping,
        "inv_map_to_solve_order": inv,
        "map_to_solve_order_lower": np.concatenate([[0], lu[: n - 1]]).astype(np.int64),
        "map_to_solve_order_upper": np.concatenate([[0], lu[n - 1 :]]).astype(np.int64),
        "node_order": new_node_order,
        "node_order_grouped": grouped,
        "parent_lookup": parent_lookup,
    }


def network_tables_and_indexer(cells: List[Dict], indexers: List[Dict], solver_device: str = "gpu"):
    """Concatenation of cells into a network: `Network._init_comp_graph`
    (modules/network.py:149-190) and `Network._init_solver_jaxley_dhs_solve`
    (modules/network.py:192-293)."""
    src, snk, ordd, typ, internal = [], [], [], [], []
    inv, mp, low, up, node_order, plook = [], [], [], [], [], []
    offset = 0
    lu_off = 0
    for tab, ind in zip(cells, indexers):
        src.append(tab["source"] + offset)
        snk.append(tab["sink"] + offset)
        ordd.append(tab["ordered"])
        typ.append(tab["type"])
        internal.append(tab["internal"] + offset)
        inv.append(ind["inv_map_to_solve_order"] + offset)
        mp.append(ind["map_to_solve_order"] + offset)
        low.append(ind["map_to_solve_order_lower"] + lu_off)
        up.append(ind["map_to_solve_order_upper"] + lu_off)
        no = np.array(ind["node_order"], copy=True).reshape(-1, 3)
        if solver_device == "cpu" or len(no) == 0:
            node_order.append(no + offset)
        else:
            no[:, :2] += offset
            node_order.append(no)
        pl = ind["parent_lookup"].copy()[:-1]
        pl[1:] += offset
        plook.append(pl)
        offset += tab["n_nodes"]
        lu_off += (tab["n_nodes"] - 1) * 2
    max_ind = max(lu_off - 1, 0)
    tables = {
        "source": np.concatenate(src),
        "sink": np.concatenate(snk),
        "ordered": np.concatenate(ordd),
        "type": np.concatenate(typ),
        "n_nodes": offset,
        "n_comp": int(sum(len(i) for i in internal)),
        "internal": np.concatenate(internal),
    }
    no_all = np.concatenate(node_order, axis=0)
    grouped = group_levels(no_all)
    indexer = {
        "inv_map_to_solve_order": np.concatenate(inv),
        "map_to_solve_order": np.concatenate(mp),
        "map_to_solve_order_lower": np.minimum(np.concatenate(low), max_ind),
        "map_to_solve_order_upper": np.minimum(np.concatenate(up), max_ind),
        "node_order": no_all,
        "node_order_grouped": grouped,
        "parent_lookup": np.concatenate(plook + [np.asarray([-1])]),
    }
    return tables, indexer


# --------------------------------------------------------------------------------------
# Geometry -> axial conductances
# --------------------------------------------------------------------------------------


def node_geometry(length, radius, n_nodes: int, internal: np.ndarray) -> Dict[str, torch.Tensor]:
    """Per-node area and resistive loads (`cylinder_area`, `cylinder_resistive_load`,
    utils/morph_attributes.py:261-312, with `l / 2` as in modules/compartment.py:70-78).
    Branchpoints hold -1, as `to_torch` writes them (modules/base.py:1042-1046)."""
    length = torch.as_tensor(length, dtype=F64)
    radius = torch.as_tensor(radius, dtype=F64)
    area_c = 2.0 * np.pi * radius * length
    rl_c = (length / 2) / radius**2 / np.pi
    batch = torch.broadcast_shapes(length.shape[:-1], radius.shape[:-1])
    out = {}
    idx = torch.as_tensor(internal)
    for key, val in (("area", area_c), ("rl", rl_c), ("length", length), ("radius", radius)):
        full = -torch.ones((*batch, n_nodes), dtype=F64)
        full = full.index_copy(-1, idx, torch.broadcast_to(val, (*batch, len(internal))))
        out[key] = full
    return out


def axial_conductances(tables: Dict, params: Dict[str, torch.Tensor]) -> torch.Tensor:
    """Voltage axial conductances per compartment edge, already divided by `cm`
    (`compute_axial_conductances`, utils/morph_attributes.py:315-436)."""
    typ, s, t, o = tables["type"], tables["source"], tables["sink"], tables["ordered"]
    Ra = params["axial_resistivity"]
    rl_in, rl_out = params["resistive_load_in"], params["resistive_load_out"]
    area, cm = params["area"], params["capacitance"]
    batch = torch.broadcast_shapes(Ra.shape[:-1], rl_in.shape[:-1], area.shape[:-1], cm.shape[:-1])
    g = torch.zeros((*batch, len(typ)), dtype=F64)
    ax = 1.0 / Ra
    comp_r_a = [rl_in / ax, rl_out / ax]          # [sink (in), source (out)]
    comp_g_a = [1.0 / rl_in * ax, 1.0 / rl_out * ax]
    c0 = np.nonzero(typ == 0)[0]
    if len(c0):
        ra = torch.stack([comp_r_a[int(oo)][..., int(ss)] for oo, ss in zip(o[c0], s[c0])], -1) + torch.stack(
            [comp_r_a[1 - int(oo)][..., int(tt)] for oo, tt in zip(o[c0], t[c0])], -1
        )
        tt = torch.as_tensor(t[c0])
        val = 1.0 / ra / area[..., tt] / cm[..., tt] * 10**7
        g = g.index_copy(-1, torch.as_tensor(c0), torch.broadcast_to(val, (*batch, len(c0))))
    c12 = np.nonzero(np.isin(typ, [1, 2]))[0]
    if len(c12):
        ra = torch.stack([comp_r_a[1 - int(oo)][..., int(tt)] for oo, tt in zip(o[c12], t[c12])], -1)
        tt = torch.as_tensor(t[c12])
        val = 1.0 / ra / area[..., tt] / cm[..., tt] * 10**7
        g = g.index_copy(-1, torch.as_tensor(c12), torch.broadcast_to(val, (*batch, len(c12))))
        o12 = o[c12]
        c34 = np.nonzero(np.isin(typ, [3, 4]))[0]
        val = torch.stack([comp_g_a[1 - int(oo)][..., int(ss)] for oo, ss in zip(o12, s[c34])], -1)
        g = g.index_copy(-1, torch.as_tensor(c34), torch.broadcast_to(val, (*batch, len(c34))))
    return g


# --------------------------------------------------------------------------------------
# Gate solvers and mechanisms
# --------------------------------------------------------------------------------------


def save_exp(x, max_value: float = 20.0):
    """solver_gate.py:8-11."""
    return torch.exp(torch.clamp(torch.as_tensor(x), max=max_value))


def x_over_expm1(x, y=1.0):
    """solver_gate.py:14-30."""
    x = torch.as_tensor(x)
    z = torch.clamp(x / y, max=20.0)
    small = torch.abs(z) < 1e-4
    z_safe = torch.where(small, torch.ones_like(z), z)
    series = y * (1.0 - z / 2.0 + z * z / 12.0)
    return torch.where(small, series, x / torch.expm1(z_safe))


def solve_gate_exponential(x, dt, alpha, beta):
    """solver_gate.py:45-64."""
    tau = 1 / (alpha + beta)
    xinf = alpha * tau
    exp_term = save_exp(-dt / tau)
    return x * exp_term + xinf * (1.0 - exp_term)


def hh_rates(v):
    """channels/hh.py:87-108."""
    am = 0.1 * x_over_expm1(-(v + 40), 10)
    bm = 4.0 * save_exp(-(v + 65) / 18)
    ah = 0.07 * save_exp(-(v + 65) / 20)
    bh = 1.0 / (save_exp(-(v + 35) / 10) + 1)
    an = 0.01 * x_over_expm1(-(v + 55), 10)
    bn = 0.125 * save_exp(-(v + 65) / 80)
    return (am, bm), (ah, bh), (an, bn)


def hh_update(m, h, n, v, dt):
    """`HH.update_states` (channels/hh.py:37-50)."""
    (am, bm), (ah, bh), (an, bn) = hh_rates(v)
    return (
        solve_gate_exponential(m, dt, am, bm),
        solve_gate_exponential(h, dt, ah, bh),
        solve_gate_exponential(n, dt, an, bn),
    )


def hh_current(m, h, n, v, p):
    """`HH.compute_current` (channels/hh.py:52-67), mA/cm^2."""
    gNa = p["HH_gNa"] * (m**3) * h
    gK = p["HH_gK"] * n**4
    return gNa * (v - p["HH_eNa"]) + gK * (v - p["HH_eK"]) + p["HH_gLeak"] * (v - p["HH_eLeak"])



# Pospischil et al. (2008) channels (channels/pospischil.py:74-250): Na (m, h) and K (n)
# with the global threshold parameter vt, Km (p) with its own time-constant scale.
def posp_na_rates(v, vt):
    """`Na.m_gate`, `Na.h_gate` (pospischil.py:120-134)."""
    va, vb = v - vt - 13.0, v - vt - 40.0
    am = 0.32 * x_over_expm1(-0.25 * va) / 0.25
    bm = 0.28 * x_over_expm1(0.2 * vb) / 0.2
    ha, hb = v - vt - 17.0, v - vt - 40.0
    ah = 0.128 * save_exp(-ha / 18.0)
    bh = 4.0 / (save_exp(-hb / 5.0) + 1.0)
    return am, bm, ah, bh


def posp_k_rates(v, vt):
    """`K.n_gate` (pospischil.py:186-193)."""
    an = 0.032 * x_over_expm1(-0.2 * (v - vt - 15.0)) / 0.2
    bn = 0.5 * save_exp(-(v - vt - 10.0) / 40.0)
    return an, bn


def posp_km_gate(v, taumax):
    """`Km.p_gate` (pospischil.py:240-245): (p_inf, tau_p)."""
    vp = v + 35.0
    return 1.0 / (1.0 + save_exp(-0.1 * vp)), taumax / (3.3 * save_exp(0.05 * vp) + save_exp(-0.05 * vp))


def gate_exponential(x, dt, alpha, beta):
    """`solve_gate_exponential` (solver_gate.py:45-64)."""
    tau = 1.0 / (alpha + beta)
    e = save_exp(-dt / tau)
    return x * e + alpha * tau * (1.0 - e)


def inf_gate_exponential(x, dt, x_inf, tau):
    """`solve_inf_gate_exponential` (solver_gate.py:67-87)."""
    e = save_exp(-1.0 / tau * dt)
    return x * e + x_inf * (1.0 - e)


POSP_PARAMS = {"Na_gNa": 50e-3, "eNa": 50.0, "vt": -60.0, "K_gK": 5e-3, "eK": -90.0,
               "Km_gKm": 0.004e-3, "Km_taumax": 4000.0}
POSP_STATES = {"Na": ("Na_m", "Na_h"), "K": ("K_n",), "Km": ("Km_p",)}


def posp_update(name, st, v, p, dt):
    """New gates of one Pospischil channel at voltage v (`update_states`)."""
    if name == "Na":
        am, bm, ah, bh = posp_na_rates(v, p["vt"])
        return {"Na_m": gate_exponential(st["Na_m"], dt, am, bm), "Na_h": gate_exponential(st["Na_h"], dt, ah, bh)}
    if name == "K":
        an, bn = posp_k_rates(v, p["vt"])
        return {"K_n": gate_exponential(st["K_n"], dt, an, bn)}
    pinf, tau = posp_km_gate(v, p["Km_taumax"])
    return {"Km_p": inf_gate_exponential(st["Km_p"], dt, pinf, tau)}


def posp_current(name, st, v, p):
    """`compute_current` in mA/cm^2 (pospischil.py:108-115, 176-182, 228-233)."""
    if name == "Na":
        return p["Na_gNa"] * st["Na_m"] ** 3 * st["Na_h"] * (v - p["eNa"])
    if name == "K":
        return p["K_gK"] * st["K_n"] ** 4 * (v - p["eK"])
    return p["Km_gKm"] * st["Km_p"] * (v - p["eK"])


def hh_init(v):
    """`HH.init_state` (channels/hh.py:69-85)."""
    (am, bm), (ah, bh), (an, bn) = hh_rates(v)
    return am / (am + bm), ah / (ah + bh), an / (an + bn)


def leak_current(v, p):
    """`Leak.compute_current` (channels/pospischil.py:65-69)."""
    return p["Leak_gLeak"] * (v - p["Leak_eLeak"])


def kinetic_update(s, pre_v, p, dt):
    """`KineticSynapse.update_states` (synapses/kinetic.py:37-53)."""
    transmitter = torch.sigmoid((pre_v - p["v_th"]) / p["delta"])
    a = p["alpha"] * transmitter
    rate = a + p["beta"]
    s_inf = a / rate
    return s_inf + (s - s_inf) * torch.exp(-dt * rate)


def kinetic_current(s, pre_v, post_v, p):
    """`KineticSynapse.compute_current` (synapses/kinetic.py:55-63), nA."""
    return p["gS"] * s * (post_v - p["e_syn"])


def sigmoid_rate_current(s, pre_v, post_v, p):
    """`SigmoidRateSynapse.compute_current` (synapses/sigmoid_rate.py:42-50), nA."""
    return -p["gS"] * torch.sigmoid((pre_v - p["x_offset"]) * p["slope"])


def ionotropic_update(s, pre_v, p, dt):
    """`IonotropicSynapse.update_states` (synapses/ionotropic.py:44-66)."""
    u = (pre_v - p["v_th"]) / p["delta"]
    s_inf = 1.0 / (1.0 + save_exp(-u))
    rate = p["k_minus"] * (1.0 + torch.exp(torch.clamp(u, -20.0, 50.0)))
    exp_term = save_exp(-rate * dt)
    return s * exp_term + s_inf * (1.0 - exp_term)


def ionotropic_current(s, pre_v, post_v, p):
    """`IonotropicSynapse.compute_current` (synapses/ionotropic.py:68-72)."""
    return p["gS"] * s * (post_v - p["e_syn"])


def test_synapse_update(c, pre_v, p, dt):
    """`TestSynapse.update_states` (synapses/test.py:25-48): v_th = -35, delta = 10,
    k_minus = 1/40."""
    s_bar = 1.0 / (1.0 + save_exp((-35.0 - pre_v) / 10.0))
    tau_s = (1.0 - s_bar) / (1.0 / 40.0)
    exp_term = save_exp((-1.0 / tau_s) * dt)
    return c * exp_term + s_bar * (1.0 - exp_term)


def test_synapse_current(c, pre_v, post_v, p):
    """`TestSynapse.compute_current` (synapses/test.py:50-59), e_syn = 0."""
    return p["gC"] * c * (post_v - 0.0)


test_synapse_update.__test__ = False
test_synapse_current.__test__ = False

SYNAPSES = {
    "KineticSynapse": (kinetic_update, kinetic_current, ("gS", "e_syn", "alpha", "beta", "v_th", "delta"), 0.0),
    "SigmoidRateSynapse": (None, sigmoid_rate_current, ("gS", "x_offset", "slope"), None),
    "IonotropicSynapse": (ionotropic_update, ionotropic_current, ("gS", "e_syn", "k_minus", "v_th", "delta"), 0.2),
    "TestSynapse": (test_synapse_update, test_synapse_current, ("gC",), 0.2),
}
SYN_DEFAULTS = {
    "KineticSynapse": {"gS": 1e-4, "e_syn": 0.0, "alpha": 1.0, "beta": 0.2, "v_th": -20.0, "delta": 2.0},
    "SigmoidRateSynapse": {"gS": 1e-4, "x_offset": -60.0, "slope": 0.25},
    "IonotropicSynapse": {"gS": 1e-4, "e_syn": 0.0, "k_minus": 0.025, "v_th": -35.0, "delta": 10.0},
    "TestSynapse": {"gC": 1e-4},
}


def synapse_linearization(kind, s, pre_v, post_v, prm, area_post, dt, diff=1e-3, update=True):
    """One synapse type's part of a step (modules/network.py:382-478): the state update
    at the old voltages (when `update` and the type has a state), then the finite
    difference linearization of i / area * 1e5 with BOTH voltages perturbed by `diff`
    (network.py:452-457). Returns `(s_new, vt, ct, i)`: d0 = vt v_post + ct is the current
    per unit area (uA / cm^2), i the current in nA (the `i_<synapse>` state,
    network.py:462-465)."""
    upd, cur, _, _ = SYNAPSES[kind]
    if update and upd is not None:
        s = upd(s, pre_v, prm, dt)
    i0 = cur(s, pre_v, post_v, prm)
    i1 = cur(s, pre_v + diff, post_v + diff, prm)
    shape = torch.broadcast_shapes(i0.shape, i1.shape, post_v.shape)
    d0 = torch.broadcast_to(i0, shape) / area_post * 100_000
    d1 = torch.broadcast_to(i1, shape) / area_post * 100_000
    vt = (d1 - d0) / diff
    ct = d0 - vt * post_v
    return s, vt, ct, torch.broadcast_to(i0, shape)


HH_DEFAULTS = {"HH_gNa": 0.12, "HH_gK": 0.036, "HH_gLeak": 0.0003, "HH_eNa": 50.0, "HH_eK": -77.0, "HH_eLeak": -54.3}
LEAK_DEFAULTS = {"Leak_gLeak": 1e-4, "Leak_eLeak": -70.0}

# --------------------------------------------------------------------------------------
# DHS tree solve
# --------------------------------------------------------------------------------------


class DHSSolverOracle:
    """`DHSSolver.raw_solve` (solver_voltage.py:18-140) on CPU: leaves-first
    `index_add_` elimination per level, then recursive doubling
    (`optimize_for_gpu=True`) or the sequential root-first back-substitution."""

    def __init__(self, indexer: Dict, optimize_for_gpu: bool = True):
        levels = [np.asarray(l, dtype=np.int64).reshape(-1, 2) for l in indexer["node_order_grouped"]]
        self.levels = [(torch.as_tensor(l[:, 0]), torch.as_tensor(l[:, 1])) for l in levels]
        self.optimize_for_gpu = optimize_for_gpu
        self.parent_lookup = torch.as_tensor(np.asarray(indexer["parent_lookup"], dtype=np.int64))
        self.has_no_parent = self.parent_lookup < 0
        self.has_no_parent[0] = True
        steps = len(self.levels)
        self.num_recursive_steps = int(np.ceil(np.log2(steps + 1))) if steps > 0 else 0

    def solve(self, d, l, u, s):
        d = d.clone()
        s = s.clone()
        for child, parent in reversed(self.levels):
            mult = u[..., child] / d[..., child]
            d = d.index_add(-1, parent, -l[..., child] * mult)
            s = s.index_add(-1, parent, -s[..., child] * mult)
        if not self.optimize_for_gpu:
            ln = l / d
            s = s / d
            for child, parent in self.levels:
                s = s.index_add(-1, child, -s[..., parent] * ln[..., child])
            return s
        l = l.masked_fill(self.has_no_parent, 0.0)
        le = -l / d
        se = s / d
        pj = self.parent_lookup
        for _ in range(self.num_recursive_steps):
            se = le * se[..., pj] + se
            le = le * le[..., pj]
            pj = pj[pj]
        return se


def dense_matrix(d, l, u, parent_lookup):
    """Dense matrix of the tree system in solve order: A[c,p]=l[c], A[p,c]=u[c]."""
    n = d.shape[-1]
    A = np.diag(np.asarray(d, dtype=np.float64))
    for c in range(n):
        p = int(parent_lookup[c])
        if p >= 0 and c != 0:
            A[c, p] += float(l[c])
            A[p, c] += float(u[c])
    return A


# --------------------------------------------------------------------------------------
# Model + integrate
# --------------------------------------------------------------------------------------


def step_current(i_delay, i_dur, i_amp, delta_t, t_max, i_offset=0.0):
    """stimulus.py:8-36 (float-exact window)."""
    window_start = int(i_delay / delta_t)
    window_end = int((i_delay + i_dur) / delta_t)
    time_steps = int(t_max // delta_t) + 2
    cur = np.zeros(time_steps) + i_offset
    cur[window_start:window_end] = i_amp
    return torch.as_tensor(cur, dtype=F64)


class OracleModel:
    """A cell or network as the reference's tables + parameters (float64, CPU).

    `cells`: list of `(parents, ncomp_per_branch)`; geometry and mechanisms are set
    through the attributes below (node-indexed arrays, -1 at branchpoints)."""

    def __init__(self, cells, allowed_nodes_per_level: int = 32, solver_device: str = "gpu"):
        tabs = [cell_tables(p, nc) for p, nc in cells]
        inds = [dhs_indexer_cell(t, allowed_nodes_per_level) for t in tabs]
        if len(cells) == 1:
            self.tables, self.indexer = tabs[0], inds[0]
        else:
            self.tables, self.indexer = network_tables_and_indexer(tabs, inds, solver_device)
        self.cell_tables = tabs
        n = self.tables["n_nodes"]
        self.n_nodes = n
        internal = self.tables["internal"]
        self.internal = internal
        nc = len(internal)
        # Default compartment params (modules/compartment.py:33-37).
        self.length = np.full(nc, 10.0)
        self.radius = np.full(nc, 1.0)
        self.axial_resistivity = np.full(nc, 5000.0)
        self.capacitance = np.full(nc, 1.0)
        self.v0 = np.full(nc, -70.0)
        self.hh_comps = np.zeros(0, np.int64)      # compartment positions (0..nc-1)
        self.leak_comps = np.zeros(0, np.int64)
        self.hh_params = {k: np.full(nc, v) for k, v in HH_DEFAULTS.items()}
        self.leak_params = {k: np.full(nc, v) for k, v in LEAK_DEFAULTS.items()}
        self.gates0 = None  # (m, h, n) over compartments, default 0.2
        # Traced geometry of a morphology read from SWC: per compartment "area",
        # "resistive_load_in", "resistive_load_out" (io/graph.py:351-359 -> module
        # nodes, io/graph.py:1058-1064); None means cylinders of `length`, `radius`.
        self.morph: Optional[Dict[str, np.ndarray]] = None
        # Pospischil channels: compartments per channel ("Na", "K", "Km"), parameters
        # and initial gates per compartment.
        self.posp_comps: Dict[str, np.ndarray] = {}
        self.posp_params = {k: np.full(nc, v) for k, v in POSP_PARAMS.items()}
        self.posp_states0 = {k: np.full(nc, 0.2) for ks in POSP_STATES.values() for k in ks}
        self.synapses: List[Dict] = []   # {"type", "pre" (node), "post" (node), params, s0}

    def node_array(self, comp_vals, fill=-1.0):
        """Compartment values -> node-indexed tensor (last dim n_nodes)."""
        comp_vals = torch.as_tensor(comp_vals, dtype=F64)
        batch = comp_vals.shape[:-1]
        out = torch.full((*batch, self.n_nodes), fill, dtype=F64)
        return out.index_copy(-1, torch.as_tensor(self.internal), comp_vals)

    def params(self, overrides: Optional[Dict] = None) -> Dict:
        overrides = overrides or {}
        get = lambda k, d: overrides.get(k, d)
        length = torch.as_tensor(get("length", self.length), dtype=F64)
        radius = torch.as_tensor(get("radius", self.radius), dtype=F64)
        geo = node_geometry(length, radius, self.n_nodes, self.internal)
        p = {
            "area": geo["area"],
            "resistive_load_in": geo["rl"],
            "resistive_load_out": geo["rl"],
        }
        if self.morph is not None:
            for k in ("area", "resistive_load_in", "resistive_load_out"):
                p[k] = self.node_array(self.morph[k])
        p.update({
            "axial_resistivity": self.node_array(get("axial_resistivity", self.axial_resistivity)),
            "capacitance": self.node_array(get("capacitance", self.capacitance)),
        })
        p["g_axial"] = axial_conductances(self.tables, p)
        for k, v in self.hh_params.items():
            p[k] = torch.as_tensor(get(k, v), dtype=F64)
        for k, v in self.leak_params.items():
            p[k] = torch.as_tensor(get(k, v), dtype=F64)
        return p


def explicit_step(v, lin_t, const_t, g_ax, tables: Dict, dt: float):
    """`step_voltage_explicit` / `_voltage_vectorfield` (solver_voltage.py:412-502):
    v' = v + dt f(v), f = -lin v + const plus the axial currents of every compartment
    edge, with each branchpoint at the conductance-weighted mean of its neighbours
    (edges of types 3 and 4 point into branchpoints)."""
    typ = np.asarray(tables["type"])
    sinks = torch.as_tensor(tables["sink"])
    sources = torch.as_tensor(tables["source"])
    n = v.shape[-1]
    batch = torch.broadcast_shapes(v.shape[:-1], lin_t.shape[:-1], const_t.shape[:-1], g_ax.shape[:-1])
    vv = torch.broadcast_to(v, (*batch, n))
    g = torch.broadcast_to(g_ax, (*batch, g_ax.shape[-1]))
    f = torch.broadcast_to(-lin_t * v + const_t, (*batch, n))
    bp = np.nonzero(np.isin(typ, [3, 4]))[0]
    if len(bp):
        ci = torch.as_tensor(bp)
        snk, src = sinks[ci], sources[ci]
        num = vv.index_fill(-1, snk, 0.0).index_add(-1, snk, g[..., ci] * vv[..., src])
        den = torch.ones_like(vv).index_fill(-1, snk, 0.0).index_add(-1, snk, g[..., ci])
        vv = num / den
    if np.isin(typ, [0, 1, 2]).any():
        f = f.index_add(-1, sinks, (vv[..., sources] - vv[..., sinks]) * g)
    return v + dt * f


def _recorder(model, record, internal):
    """`_build_recorder` (integrate.py:53-84) for `(state, index)` recordings: `v` by
    compartment; HH gates by compartment (NaN where HH is absent); `i_HH` / `i_Leak`,
    the channel's membrane current in mA / cm^2 at the step's starting voltage with the
    new gates (base.py:3445-3502; 0 where absent and before the first step); synapse
    states `<type>_s` / `<type>_c` and currents `i_<type>` in nA (network.py:462-465) by
    index within the synapse type."""
    hh_pos = {int(c): j for j, c in enumerate(model.hh_comps)}
    leak_pos = {int(c): j for j, c in enumerate(model.leak_comps)}
    syn_of = {sp["type"]: j for j, sp in enumerate(model.synapses)}

    def get(state, idx, v, m, h, nn, syn_states, currents):
        nan = torch.full(v.shape[:-1], float("nan"), dtype=F64)
        zero = torch.zeros(v.shape[:-1], dtype=F64)
        if state == "v":
            return v[..., int(internal[idx])]
        if state in ("HH_m", "HH_h", "HH_n"):
            g = {"HH_m": m, "HH_h": h, "HH_n": nn}[state]
            return torch.broadcast_to(g[..., hh_pos[idx]], zero.shape) if idx in hh_pos else nan
        if state in ("i_HH", "i_Leak"):
            pos = hh_pos if state == "i_HH" else leak_pos
            if currents is None or state not in currents or idx not in pos:
                return zero
            return torch.broadcast_to(currents[state][..., pos[idx]], zero.shape)
        if state.startswith("i_") and state[2:] in syn_of:
            if currents is None or state not in currents:
                return zero
            return torch.broadcast_to(currents[state][..., idx], zero.shape)
        kind = state.rsplit("_", 1)[0]
        if kind in syn_of:
            return torch.broadcast_to(syn_states[syn_of[kind]][..., idx], zero.shape)
        raise KeyError(state)

    def rec(v, m, h, nn, syn_states, currents=None):
        return torch.stack([get(st, int(i), v, m, h, nn, syn_states, currents) for st, i in record], -1)

    return rec


def simulate(*args, dtype=torch.float64, **kwargs):
    """`_simulate` in `dtype`: float64 (the oracle) or float32, every tensor operation in
    that type as the reference computes when its inputs are float32 (integrate.py:480-490
    takes the dtype of the first floating point input)."""
    global F64
    saved = F64
    F64 = dtype
    try:
        return _simulate(*args, **kwargs)
    finally:
        F64 = saved


def _simulate(
    model: OracleModel,
    t_max: Optional[float],
    dt: float = 0.025,
    stim: Optional[torch.Tensor] = None,
    stim_comps: Sequence[int] = (),
    rec_comps: Sequence[int] = (0,),
    overrides: Optional[Dict] = None,
    voltage_solver: str = "jaxley.dhs.gpu",
    return_all: bool = False,
    record: Optional[Sequence] = None,
    solver: str = "bwd_euler",
):
    """`jx.integrate` (integrate.py:396-666) + `_step_with_plan` (modules/base.py:3214-3400)
    for HH / Leak channels and Kinetic / SigmoidRate / Ionotropic / Test synapses, with
    `solver` "bwd_euler", "crank_nicolson" (modules/base.py:3327-3366: the implicit
    solve with dt / 2, then v' = 2 x - v at every node; channels and synapses still
    advance by dt) or "fwd_euler" (`step_voltage_explicit`, solver_voltage.py:412-502).

    `stim`: `(..., n_stim, time)` currents in nA (padded / truncated to
    `int(t_max // dt + 1)` steps as integrate.py:523-541 does). Recordings are
    `(..., n_rec, 1 + steps)`: the initial state and every step; `record` (a list of
    `(state, index)`, see `_recorder`) replaces the voltages at `rec_comps`."""
    overrides = overrides or {}
    if solver not in ("bwd_euler", "crank_nicolson", "fwd_euler"):
        raise ValueError(f"solver {solver!r}")
    dtv = dt / 2 if solver == "crank_nicolson" else dt
    p = model.params(overrides)
    n = model.n_nodes
    internal = torch.as_tensor(model.internal)
    if t_max is not None:
        T = int(t_max // dt + 1)
    else:
        T = int(torch.as_tensor(stim).shape[-1])
    if stim is not None:
        stim = torch.as_tensor(stim, dtype=F64)
        if stim.shape[-1] < T:
            pad = torch.zeros((*stim.shape[:-1], T - stim.shape[-1]), dtype=F64)
            stim = torch.cat([stim, pad], -1)
        else:
            stim = stim[..., :T]
    stim_nodes = torch.as_tensor(model.internal[np.asarray(stim_comps, dtype=np.int64)]) if len(stim_comps) else None
    rec_nodes = torch.as_tensor(model.internal[np.asarray(rec_comps, dtype=np.int64)])

    hh_nodes = torch.as_tensor(model.internal[model.hh_comps])
    leak_nodes = torch.as_tensor(model.internal[model.leak_comps])
    # HH / Leak parameters are compartment-indexed `(..., n_comp)`; gather the channel's
    # compartments (`_gather_step_params`, modules/base.py:3110-3139).
    hhp = {k: torch.as_tensor(overrides.get(k, v), dtype=F64)[..., model.hh_comps] for k, v in model.hh_params.items()}
    lkp = {k: torch.as_tensor(v, dtype=F64)[..., model.leak_comps] for k, v in model.leak_params.items()}

    v = model.node_array(overrides.get("v0", model.v0))
    nh = len(model.hh_comps)
    if model.gates0 is None:
        m = torch.full((nh,), 0.2, dtype=F64)
        h = torch.full((nh,), 0.2, dtype=F64)
        nn = torch.full((nh,), 0.2, dtype=F64)
    else:
        m, h, nn = (torch.as_tensor(g, dtype=F64)[..., model.hh_comps] for g in model.gates0)

    syn_states = [torch.as_tensor(sp["s0"], dtype=F64) if sp.get("s0") is not None else None for sp in model.synapses]
    posp = {}
    for name, comps in model.posp_comps.items():
        comps = np.asarray(comps, dtype=np.int64)
        prm = {k: torch.as_tensor(overrides.get(k, x), dtype=F64)[..., comps] for k, x in model.posp_params.items()}
        st0 = {k: torch.as_tensor(model.posp_states0[k], dtype=F64)[..., comps] for k in POSP_STATES[name]}
        posp[name] = [torch.as_tensor(model.internal[comps]), prm, st0]
    batch = torch.broadcast_shapes(
        v.shape[:-1], p["g_axial"].shape[:-1], *[t.shape[:-1] for t in hhp.values()],
        *(() if stim is None else (stim.shape[:-2],)),
    )
    v = torch.broadcast_to(v, (*batch, n)).clone()

    g_ax = p["g_axial"]
    sinks = torch.as_tensor(model.tables["sink"])
    cm = p["capacitance"]
    area = p["area"]
    dense = voltage_solver in ("jax.sparse", "torch.dense")
    dhs = None if dense else DHSSolverOracle(model.indexer, optimize_for_gpu=voltage_solver.endswith("gpu") or voltage_solver == "jaxley.stone")
    ind = model.indexer
    mp = torch.as_tensor(ind["map_to_solve_order"])
    inv = torch.as_tensor(ind["inv_map_to_solve_order"])
    mlo = torch.as_tensor(ind["map_to_solve_order_lower"])
    mup = torch.as_tensor(ind["map_to_solve_order_upper"])
    diff = 1e-3

    def syn_terms(v, syn_states, update):
        lin = torch.zeros_like(v)
        const = torch.zeros_like(v)
        new_states = []
        for sp, s in zip(model.synapses, syn_states):
            _, _, names, _ = SYNAPSES[sp["type"]]
            prm = {k: torch.as_tensor(sp["params"][k], dtype=F64) for k in names}
            pre = torch.as_tensor(sp["pre"])
            post = torch.as_tensor(sp["post"])
            s, vt, ct, i0 = synapse_linearization(sp["type"], s, v[..., pre], v[..., post], prm,
                                                  area[..., post], dt, diff, update)
            new_states.append(s)
            currents[f"i_{sp['type']}"] = i0
            lin = lin.index_add(-1, post, torch.broadcast_to(vt, (*v.shape[:-1], len(post))))
            const = const.index_add(-1, post, torch.broadcast_to(-ct, (*v.shape[:-1], len(post))))
        return lin, const, new_states

    currents: Dict[str, torch.Tensor] = {}

    def one_step(v, m, h, nn, syn_states, t):
        if stim_nodes is not None:
            cur = stim[..., t]
            i_ext = torch.zeros_like(v).index_add(-1, stim_nodes, torch.broadcast_to(cur / area[..., stim_nodes] * 100_000, (*v.shape[:-1], len(stim_nodes))))
        else:
            i_ext = 0.0
        lin = torch.zeros_like(v)
        const = torch.zeros_like(v)
        if nh:
            vh = v[..., hh_nodes]
            m, h, nn = hh_update(m, h, nn, vh, dt)
            I0 = hh_current(m, h, nn, vh, hhp)
            currents["i_HH"] = I0
            I1 = hh_current(m, h, nn, vh + diff, hhp)
            vt = (I1 - I0) / diff
            ct = I0 - vt * vh
            lin = lin.index_add(-1, hh_nodes, torch.broadcast_to(vt, (*v.shape[:-1], nh)))
            const = const.index_add(-1, hh_nodes, torch.broadcast_to(-ct, (*v.shape[:-1], nh)))
        if len(model.leak_comps):
            vl = v[..., leak_nodes]
            I0 = leak_current(vl, lkp)
            currents["i_Leak"] = I0
            I1 = leak_current(vl + diff, lkp)
            vt = (I1 - I0) / diff
            ct = I0 - vt * vl
            lin = lin.index_add(-1, leak_nodes, torch.broadcast_to(vt, (*v.shape[:-1], len(leak_nodes))))
            const = const.index_add(-1, leak_nodes, torch.broadcast_to(-ct, (*v.shape[:-1], len(leak_nodes))))
        for name, entry in posp.items():
            nodes, prm, st = entry
            vp = v[..., nodes]
            st = posp_update(name, st, vp, prm, dt)
            entry[2] = st
            I0 = posp_current(name, st, vp, prm)
            I1 = posp_current(name, st, vp + diff, prm)
            vt = (I1 - I0) / diff
            ct = I0 - vt * vp
            lin = lin.index_add(-1, nodes, torch.broadcast_to(vt, (*v.shape[:-1], len(nodes))))
            const = const.index_add(-1, nodes, torch.broadcast_to(-ct, (*v.shape[:-1], len(nodes))))
        lin = lin * 1000.0
        const = const * 1000.0
        slin, sconst, syn_states = syn_terms(v, syn_states, True)
        lin_t = (lin + slin) / cm
        const_t = (const + i_ext + sconst) / cm
        # _build_diags_and_solves (solver_voltage.py:250-291)
        gdt = dtv * g_ax
        bsh = torch.broadcast_shapes(v.shape[:-1], lin_t.shape[:-1], const_t.shape[:-1], gdt.shape[:-1])
        diag = torch.broadcast_to(dtv * lin_t, (*bsh, n)).clone()
        diag = diag.index_add(-1, internal, torch.ones((*bsh, len(internal)), dtype=F64))
        if len(sinks):
            diag = diag.index_add(-1, sinks, torch.broadcast_to(gdt, (*bsh, gdt.shape[-1])))
        vals = v[..., internal] + dtv * const_t[..., internal]
        rhs = torch.zeros((*bsh, n), dtype=F64).index_copy(-1, internal, torch.broadcast_to(vals, (*bsh, len(internal))))
        if solver == "fwd_euler":
            x = explicit_step(v, lin_t, const_t, g_ax, model.tables, dt)
        elif len(sinks) == 0:
            x = rhs / diag
        elif dense:
            A = torch.diag_embed(diag)
            srcs = torch.as_tensor(model.tables["source"])
            A = A.clone()
            A[..., sinks, srcs] += torch.broadcast_to(-gdt, (*bsh, gdt.shape[-1]))
            x = torch.linalg.solve(A, rhs[..., None])[..., 0]
        else:
            lu = torch.broadcast_to(-gdt, (*bsh, gdt.shape[-1]))
            dd = diag[..., mp]
            ss = rhs[..., mp]
            lo = lu[..., mlo]
            up = lu[..., mup]
            one = torch.ones((*bsh, 1), dtype=F64)
            zero = torch.zeros((*bsh, 1), dtype=F64)
            dd = torch.cat([dd, one], -1)
            ss = torch.cat([ss, zero], -1)
            lo = torch.cat([lo, zero], -1)
            up = torch.cat([up, zero], -1)
            x = dhs.solve(dd, lo, up, ss)[..., :-1]
            x = x[..., inv]
        if dtv != dt:
            x = 2 * x - v
        return x, m, h, nn, syn_states

    # Initial synapse currents are only states; they do not change the dynamics.
    if record is None:
        rec_fn = lambda v, m, h, nn, syn_states, currents=None: v[..., rec_nodes]
    else:
        rec_fn = _recorder(model, record, internal)
    recs = [rec_fn(v, m, h, nn, syn_states)]
    for t in range(T):
        v, m, h, nn, syn_states = one_step(v, m, h, nn, syn_states, t)
        recs.append(rec_fn(v, m, h, nn, syn_states, currents))
    out = torch.stack(recs, -1)
    if return_all:
        return out, {"v": v, "m": m, "h": h, "n": nn, "syn": syn_states}
    return out
