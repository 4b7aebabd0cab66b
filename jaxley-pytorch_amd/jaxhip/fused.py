"""Whole-simulation fused kernels: host side.

`FusedPlan` turns a module (cell, network or a pack of trees) into the descriptor
tables of `jxhip_fused_forward_*` / `jxhip_fused_adjoint_*` (include/jxhip.h):
topology in solve order, per-node integer and float tables, stimulus entries,
recordings and synapses. `run_forward` launches the forward; `FusedIntegrate` wraps
forward + hand-written adjoint in a `torch.autograd.Function`.

Layout in HBM ("program layout"): simulations are split into programs of S; every
per-node (or per-synapse) array is stored as [program][node][S], simulations fastest,
so a wavefront reading one level of nodes reads contiguous memory.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .schedule import check_forest, segments

HH_PARAMS = ("HH_gNa", "HH_gK", "HH_gLeak", "HH_eNa", "HH_eK", "HH_eLeak")
LEAK_PARAMS = ("Leak_gLeak", "Leak_eLeak")
ALL_PARAMS = HH_PARAMS + LEAK_PARAMS
# Parameters that may carry batch dimensions in the fused kernel, and their column in
# the batched table.
BATCHABLE = {"HH_gNa": 0, "HH_gK": 1, "HH_gLeak": 2, "Leak_gLeak": 3}
FLAG_COMP, FLAG_HH, FLAG_LEAK, FLAG_EXPORT = 1, 2, 4, 8
CHANNEL_FLAGS = {"HH": FLAG_HH, "Leak": FLAG_LEAK, "PNa": 16, "PK": 32, "PKm": 64}
# Pospischil channels: the columns of their parameter table and their gates, in the
# kernel's order (include/jxhip.h, JXHIP_NPARAM2 and gate_ext).
POSP_PARAMS = ("Na_gNa", "eNa", "vt", "K_gK", "eK", "Km_gKm", "Km_taumax")
POSP_STATES = ("Na_m", "Na_h", "K_n", "Km_p")
SYN_DIFF = 1e-3  # mV, the finite difference of the synapse slope (network.py:29)


def to_prog(x: torch.Tensor, S: int) -> torch.Tensor:
    """(B, n) -> (P, n, S) with B padded to P * S (zeros). Always a new tensor (the
    kernels update states in place)."""
    B, n = x.shape
    P = (B + S - 1) // S
    if P * S != B:
        x = torch.cat([x, x.new_zeros(P * S - B, n)], 0)
    out = torch.empty((P, n, S), dtype=x.dtype, device=x.device)
    out.copy_(x.reshape(P, S, n).transpose(1, 2))
    return out


def from_prog(x: torch.Tensor, B: int) -> torch.Tensor:
    """(P, n, S) -> (B, n)."""
    P, n, S = x.shape
    return x.transpose(1, 2).reshape(P * S, n)[:B]


@dataclass
class Structure:
    """Integer tables of one topology in kernel order (cached per module structure). A
    network split into parts has one topology per part (consecutive cells): every part's
    nodes are a contiguous range of the kernel order, with its own solve schedule,
    recordings and synapses (those onto its nodes); the per-part values are in the
    `part_*` arrays and in the rows of `topo_row`."""

    n: int
    L: int
    mp: np.ndarray             # kernel index -> node
    kidx_of_node: np.ndarray   # node -> kernel index
    comp_of_k: np.ndarray      # kernel index -> compartment (-1 branchpoint)
    topo_row: np.ndarray       # (parts, JXHIP_TOPO_COLS)
    seg: torch.Tensor
    seg_level_ptr: torch.Tensor
    seg_child: torch.Tensor
    node_i: torch.Tensor
    child_idx: torch.Tensor
    stim_col: torch.Tensor
    rec_node: torch.Tensor
    syn: Optional[Dict]
    n_rec: int
    rec_comps: np.ndarray
    stim_order: np.ndarray     # stimulus entry e -> stimulus column
    sinks_k: np.ndarray        # sink of every compartment edge, kernel index
    in_edges: np.ndarray       # (n, max in-degree): incoming edges of every kernel node in
                               # edge order, -1 padded (fixed summation order)
    lo_edge: np.ndarray        # edge of the lower coefficient of every kernel node
    up_edge: np.ndarray
    parent_k: np.ndarray
    sched: int = 0             # bytes of the solve schedule staged in LDS (largest part)
    rec_first: Optional[torch.Tensor] = None   # per kernel node: its first recording (reverse pass)
    rec_next: Optional[torch.Tensor] = None    # per recording: the node's next recording
    parts: int = 1
    part_nb: Optional[np.ndarray] = None       # first node of every part
    part_n: Optional[np.ndarray] = None        # nodes of every part
    part_rb: Optional[np.ndarray] = None       # first recording of every part (part-major)
    part_nrec: Optional[np.ndarray] = None
    rec_order: Optional[np.ndarray] = None     # part-major recording -> recording

    @property
    def n_max(self) -> int:
        """Nodes of the largest part (the LDS and the threads of a program)."""
        return int(self.part_n.max()) if self.part_n is not None else self.n

    @property
    def tgt_max(self) -> int:
        return int(self.topo_row[:, 7].max()) if self.syn else 0

    @property
    def syn_max(self) -> int:
        return int(self.topo_row[:, 11].max()) if self.syn else 0


def cell_sizes(module) -> np.ndarray:
    """Solve nodes of every tree of the module's forest (its cells), in solve order of
    their roots."""
    tables, indexer = module._ensure_solver()
    n = tables["n_nodes"]
    parent = np.asarray(indexer["parent_lookup"][:n], dtype=np.int64)
    root = _roots(parent)
    _, counts = np.unique(root, return_counts=True)
    return counts


def _roots(parent: np.ndarray) -> np.ndarray:
    """Root of every node of a forest whose parents precede their children."""
    root = np.arange(len(parent))
    has = parent >= 0
    root[has] = parent[has]
    while True:
        nxt = root[root]
        if np.array_equal(nxt, root):
            return root
        root = nxt


def _split(sizes: np.ndarray, parts: int) -> np.ndarray:
    """Part of every tree: consecutive trees, the parts' node counts balanced (at most
    `parts` parts, none empty: fewer when trees are uneven)."""
    cum = np.cumsum(sizes) - sizes                    # nodes before each tree
    total = float(sizes.sum())
    part = np.minimum((cum * parts / total).astype(np.int64), parts - 1)
    return np.unique(part, return_inverse=True)[1].reshape(-1)


def build_structure(module, rec_comps: np.ndarray, stim_nodes: np.ndarray, device, parts: int = 1) -> Structure:
    tables, indexer = module._ensure_solver()
    n = tables["n_nodes"]
    mp_solve = np.asarray(indexer["map_to_solve_order"], dtype=np.int64)
    inv_solve = np.asarray(indexer["inv_map_to_solve_order"], dtype=np.int64)
    parent_solve = np.asarray(indexer["parent_lookup"][:n], dtype=np.int64)
    if check_forest(parent_solve) is None:
        raise ValueError("the solve order must be a forest with parents before children")
    # Parts: consecutive trees (cells); each part's kernel order from its own schedule.
    root = _roots(parent_solve)
    roots, tree_of = np.unique(root, return_inverse=True)
    parts = max(1, min(int(parts), len(roots)))
    part_of_tree = _split(np.bincount(tree_of), parts) if parts > 1 else np.zeros(len(roots), np.int64)
    parts = int(part_of_tree.max()) + 1
    part_of_solve = part_of_tree[tree_of]
    korder_l, sgs = [], []
    for q in range(parts):
        nodes = np.nonzero(part_of_solve == q)[0]
        loc = np.full(n, -1, dtype=np.int64)
        loc[nodes] = np.arange(len(nodes))
        par = parent_solve[nodes]
        sg = segments(np.where(par >= 0, loc[np.maximum(par, 0)], -1))
        korder_l.append(nodes[sg["korder"]])
        sgs.append(sg)
    korder = np.concatenate(korder_l)                # kernel -> solve
    kidx_solve = np.empty(n, dtype=np.int64)         # solve -> kernel
    kidx_solve[korder] = np.arange(n)
    part_n = np.array([len(x) for x in korder_l], dtype=np.int64)
    part_nb = np.concatenate([[0], np.cumsum(part_n)[:-1]]).astype(np.int64)
    part_of_k = np.repeat(np.arange(parts), part_n)
    nb_k = part_nb[part_of_k]                        # first node of every node's part
    mp = mp_solve[korder]                            # kernel -> node
    kidx_of_node = kidx_solve[inv_solve]             # node -> kernel
    parent_k = np.where(parent_solve[korder] >= 0, kidx_solve[np.maximum(parent_solve[korder], 0)], -1)
    comp_of_node = -np.ones(n, dtype=np.int64)
    comp_of_node[module._comp_node] = np.arange(module._n_comp)
    comp_of_k = comp_of_node[mp]
    flags = np.where(comp_of_k >= 0, FLAG_COMP, 0).astype(np.int32)
    for ch in module.channels:
        mask = module._channel_mask[ch.name]
        has = np.zeros(n, dtype=bool)
        ok = comp_of_k >= 0
        has[ok] = mask[comp_of_k[ok]]
        flags |= np.where(has, CHANNEL_FLAGS[ch.kernel_id], 0).astype(np.int32)
    # Children of every kernel node (CSR, in increasing kernel index; a part's nodes and
    # their children are contiguous, the indices part-relative).
    ch_nodes = np.nonzero(parent_k >= 0)[0]
    srt = np.lexsort((ch_nodes, parent_k[ch_nodes]))
    ch_nodes = ch_nodes[srt]
    cnt = np.bincount(parent_k[ch_nodes], minlength=n)
    cbeg = np.zeros(n, dtype=np.int64)
    np.cumsum(cnt[:-1], out=cbeg[1:])
    part_cb = cbeg[part_nb] if n else np.zeros(parts, np.int64)
    # Stimulus entries grouped per kernel node; entry e reads stimulus column stim_order[e].
    stim_k = kidx_of_node[stim_nodes] if len(stim_nodes) else np.zeros(0, np.int64)
    stim_order = np.argsort(stim_k, kind="stable")
    scnt = np.bincount(stim_k, minlength=n) if len(stim_k) else np.zeros(n, np.int64)
    sbeg = np.zeros(n, dtype=np.int64)
    np.cumsum(scnt[:-1], out=sbeg[1:])
    node_i = np.zeros((n, _lib.JXHIP_NODE_ICOLS), dtype=np.int32)
    node_i[:, 0] = np.where(parent_k >= 0, parent_k - nb_k, -1)
    node_i[:, 1] = cbeg - part_cb[part_of_k]
    node_i[:, 2] = cbeg + cnt - part_cb[part_of_k]
    node_i[:, 3] = flags
    node_i[:, 4] = sbeg
    node_i[:, 5] = sbeg + scnt
    node_i[:, 6] = -1
    syn = _build_synapses(module, kidx_of_node, n, node_i, part_of_k, nb_k, parts)
    rec_k = kidx_of_node[module._comp_node[rec_comps]] if len(rec_comps) else np.zeros(0, np.int64)
    # Nodes whose voltage other threads read from HBM during the time loop.
    node_i[rec_k, 3] |= FLAG_EXPORT
    if syn is not None:
        node_i[syn["pre"], 3] |= FLAG_EXPORT
        node_i[syn["tgt"] + syn["tgt_nb"], 3] |= FLAG_EXPORT
    # Recordings grouped by part (part-major order; rec_order maps back).
    rec_part = part_of_k[rec_k] if len(rec_k) else np.zeros(0, np.int64)
    rec_order = np.argsort(rec_part, kind="stable")
    part_nrec = np.bincount(rec_part, minlength=parts).astype(np.int64)
    part_rb = np.concatenate([[0], np.cumsum(part_nrec)[:-1]]).astype(np.int64)
    rec_local = rec_k[rec_order] - nb_k[rec_k[rec_order]] if len(rec_k) else rec_k
    # The recordings of every node as a list (first, next; part-relative indices; the
    # reverse pass adds their gradients).
    none = np.iinfo(np.int32).max
    rec_first = np.full(n, none, dtype=np.int64)
    rec_next = np.full(max(len(rec_k), 1), none, dtype=np.int64)
    for r in range(len(rec_k) - 1, -1, -1):
        kk = rec_k[rec_order[r]]
        rb = part_rb[part_of_k[kk]]
        rec_next[r] = rec_first[kk]
        rec_first[kk] = r - rb
    # Segment tables of the parts, concatenated (part-relative node indices).
    seg = np.concatenate([sg["seg"] for sg in sgs]) if n else np.zeros((0, 4), np.int32)
    lptr = np.concatenate([sg["seg_level_ptr"] for sg in sgs])
    schild = np.concatenate([sg["seg_child"] for sg in sgs])
    nseg = np.array([len(sg["seg"]) for sg in sgs], dtype=np.int64)
    nch = np.array([len(sg["seg_child"]) for sg in sgs], dtype=np.int64)
    nlp = np.array([len(sg["seg_level_ptr"]) for sg in sgs], dtype=np.int64)
    base = lambda c: np.concatenate([[0], np.cumsum(c)[:-1]]).astype(np.int64)
    Ls = np.array([sg["L"] for sg in sgs], dtype=np.int64)
    sched = [sched_bytes(len(sg["seg"]), sg["L"], int(sg["seg"][-1, 3]) if len(sg["seg"]) else 0) for sg in sgs]
    topo_row = np.zeros((parts, _lib.JXHIP_TOPO_COLS), dtype=np.int32)
    topo_row[:, 0] = part_n
    topo_row[:, 1] = Ls
    topo_row[:, 2] = part_nb
    topo_row[:, 3] = base(nlp)
    topo_row[:, 4] = part_cb
    topo_row[:, 5] = part_nrec
    topo_row[:, 6] = part_rb
    if syn is not None:
        topo_row[:, 7] = syn["part_ntgt"]
        topo_row[:, 8] = syn["part_tgb"]
        topo_row[:, 9] = syn["part_spb"]
        topo_row[:, 10] = syn["part_synb"]
        topo_row[:, 11] = syn["part_nsyn"]
    topo_row[:, 12] = base(nseg)
    topo_row[:, 13] = base(nch)
    topo_row[:, 14] = part_nb                        # pre_ptr is indexed by the group's nodes
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.int32, device=device)
    if syn is not None:
        # The synapses grouped by presynaptic node (over the group's nodes): every
        # synapse's slot in the grouping (the reverse pass stores its presynaptic share
        # there and sums each node's contiguous slots, without atomics).
        pre_np = np.asarray(syn["pre"], dtype=np.int64)
        order = np.argsort(pre_np, kind="stable")
        pos = np.empty(len(order), dtype=np.int64)
        pos[order] = np.arange(len(order))
        syn["pre_ptr"] = t(np.concatenate([[0], np.cumsum(np.bincount(pre_np, minlength=n))]))
        syn["pre_pos"] = t(pos if len(pos) else np.zeros(1, np.int64))
        for key in ("pre", "ptr", "tgt"):
            syn[key] = t(syn[key])
        if syn["kind"] is not None:
            syn["kind"] = t(syn["kind"])
    return Structure(
        n=n, L=int(Ls.max()), mp=mp, kidx_of_node=kidx_of_node, comp_of_k=comp_of_k, topo_row=topo_row,
        seg=t(seg if len(seg) else np.zeros((1, 4), np.int32)), seg_level_ptr=t(lptr), seg_child=t(schild),
        node_i=t(node_i), child_idx=t(ch_nodes - nb_k[ch_nodes] if len(ch_nodes) else np.zeros(1, np.int64)),
        stim_col=t(stim_order if len(stim_order) else np.zeros(1, np.int64)),
        rec_node=t(rec_local if len(rec_k) else np.zeros(1, np.int64)),
        syn=syn, n_rec=len(rec_k), rec_comps=rec_comps, stim_order=stim_order,
        sinks_k=kidx_of_node[tables["sink"]],
        in_edges=_in_edges(kidx_of_node[tables["sink"]], n),
        lo_edge=np.asarray(indexer["map_to_solve_order_lower"])[korder],
        up_edge=np.asarray(indexer["map_to_solve_order_upper"])[korder],
        parent_k=parent_k,
        sched=int(max(sched)),
        rec_first=t(rec_first), rec_next=t(rec_next),
        parts=parts, part_nb=part_nb, part_n=part_n, part_rb=part_rb, part_nrec=part_nrec,
        rec_order=rec_order,
    )


# Repeated end entries of a part's jagged-diagonal pointer list (the forward's synapse
# phase walks a target's synapses with up to 8 threads, each reading up to 7 entries
# past the target's last synapse).
SYN_PTR_END = 8


def _build_synapses(module, kidx_of_node: np.ndarray, n: int, node_i: np.ndarray,
                    part_of_k: np.ndarray, nb_k: np.ndarray, parts: int) -> Optional[Dict]:
    """All synapses of the network, grouped by the part of their target, each part's in
    jagged-diagonal order: the part's targets (postsynaptic nodes) numbered by
    non-increasing in-degree (ties by node), and the list holding every target's first
    synapse, then every target's second one, and so on, so that the j-th synapse of
    target t is `ptr[j] + t` (for t below `ptr[j + 1] - ptr[j]`; a part's `ptr` ends with
    SYN_PTR_END entries equal to its synapse count, so that the kernels' threads that
    walk every 2nd to 8th synapse of a target read within it). Threads of consecutive targets then read consecutive synapses. A
    target's synapses keep the order of the types in the network and of the edges within
    a type (the order the currents are summed in). Several types give the per-synapse
    model ids of JXHIP_SYN_MIXED. Presynaptic nodes index the group's nodes, targets are
    part-relative."""
    if not module._edges:
        return None
    groups, pre_l, post_l, kind_l = [], [], [], []
    off = 0
    for e in module._edges:
        m = len(e["pre"])
        groups.append({"type": e["type"], "offset": off, "n": m})
        pre_l.append(np.asarray(e["pre"], dtype=np.int64))
        post_l.append(np.asarray(e["post"], dtype=np.int64))
        kind_l.append(np.full(m, int(e["type"].kernel_model), dtype=np.int32))
        off += m
    post_nodes = np.concatenate(post_l)
    post = kidx_of_node[post_nodes]
    pre = kidx_of_node[np.concatenate(pre_l)]
    kind = np.concatenate(kind_l)
    perm_l, ptr_l, tgt_l, tgt_g = [], [], [], []
    part_ntgt = np.zeros(parts, np.int64)
    part_nsyn = np.zeros(parts, np.int64)
    part_nptr = np.zeros(parts, np.int64)
    for q in range(parts):
        mine = np.nonzero(part_of_k[post] == q)[0]
        part_nsyn[q] = len(mine)
        if not len(mine):
            ptr_l.append(np.zeros(1 + SYN_PTR_END, np.int64))
            part_nptr[q] = 1 + SYN_PTR_END
            continue
        order = mine[np.argsort(post[mine], kind="stable")]   # grouped by target
        targets, counts = np.unique(post[order], return_counts=True)
        first = np.concatenate([[0], np.cumsum(counts)[:-1]])
        rank = np.argsort(-counts, kind="stable")              # non-increasing in-degree
        targets, counts, first = targets[rank], counts[rank], first[rank]
        per_j = np.array([int((counts > j).sum()) for j in range(int(counts.max()))], dtype=np.int64)
        ptr = np.concatenate([[0], np.cumsum(per_j), [len(mine)] * SYN_PTR_END]).astype(np.int64)
        perm = np.empty(len(mine), dtype=np.int64)
        for j, c in enumerate(per_j):
            perm[ptr[j]:ptr[j] + c] = order[first[:c] + j]
        node_i[targets, 6] = np.arange(len(targets), dtype=np.int32)
        perm_l.append(perm)
        ptr_l.append(ptr)
        tgt_l.append(targets - nb_k[targets])
        tgt_g.append(targets)
        part_ntgt[q] = len(targets)
        part_nptr[q] = len(ptr)
    perm = np.concatenate(perm_l)
    tgt_local = np.concatenate(tgt_l)
    part_tgb = np.concatenate([[0], np.cumsum(part_ntgt)[:-1]]).astype(np.int64)
    mixed = len(groups) > 1
    return {
        "model": _lib.JXHIP_SYN_MIXED if mixed else int(groups[0]["type"].kernel_model),
        "groups": groups,
        "kind": kind[perm] if mixed else None,
        "perm": perm,
        "pre": pre[perm].astype(np.int32),
        "ptr": np.concatenate(ptr_l).astype(np.int32),
        "tgt": tgt_local.astype(np.int32),
        "tgt_nb": nb_k[np.concatenate(tgt_g)] if tgt_g else np.zeros(0, np.int64),
        "post_node": post_nodes[perm],
        "n_tgt": int(part_ntgt.sum()),
        "n_syn": len(post),
        "part_ntgt": part_ntgt, "part_tgb": part_tgb,
        "part_spb": np.concatenate([[0], np.cumsum(part_nptr)[:-1]]).astype(np.int64),
        "part_synb": np.concatenate([[0], np.cumsum(part_nsyn)[:-1]]).astype(np.int64),
        "part_nsyn": part_nsyn,
    }


def syn_has_state(syn: Dict) -> bool:
    """Whether the launch keeps a state per synapse (every model but the sigmoid rate)."""
    return syn["model"] != 1


def _in_edges(sinks_k: np.ndarray, n: int) -> np.ndarray:
    """Incoming edges of every kernel node, in edge order, padded with -1."""
    sinks_k = np.asarray(sinks_k, dtype=np.int64)
    deg = np.bincount(sinks_k, minlength=n) if len(sinks_k) else np.zeros(n, np.int64)
    out = np.full((n, max(int(deg.max()) if n else 0, 1)), -1, dtype=np.int64)
    order = np.argsort(sinks_k, kind="stable")              # edges grouped by sink, in order
    ks = sinks_k[order]
    first = np.concatenate([[0], np.cumsum(deg)[:-1]])
    out[ks, np.arange(len(order)) - first[ks]] = order
    return out


def node_float_table(st: Structure, g_axial: torch.Tensor, cm_node: torch.Tensor, dt: float,
                     dtype) -> torch.Tensor:
    """lo, up, dbase, inv_cm per kernel node from the (unbatched) axial conductances
    (already / cm, per compartment edge) and the node capacitances."""
    dev = g_axial.device
    n = st.n
    if dev.type == "cpu" and g_axial.ndim == 1 and not g_axial.requires_grad and not cm_node.requires_grad:
        return torch.from_numpy(_node_float_table_np(st, g_axial.double().numpy(), cm_node.double().numpy(),
                                                     dt)).to(dtype)
    gdt = (dt * g_axial.double())
    if len(gdt):
        lo = -gdt[torch.as_tensor(st.lo_edge, device=dev)]
        up = -gdt[torch.as_tensor(st.up_edge, device=dev)]
    else:
        lo = up = torch.zeros(n, dtype=torch.float64, device=dev)
    has_par = torch.as_tensor(st.parent_k >= 0, device=dev)
    lo = torch.where(has_par, lo, torch.zeros_like(lo))
    up = torch.where(has_par, up, torch.zeros_like(up))
    dbase = torch.as_tensor((st.comp_of_k >= 0).astype(np.float64), device=dev)
    if len(gdt):
        # 1 + sum of dt g over the node's incoming edges, added in edge order (as the
        # reference's index_add on the CPU does; a scatter-add on the GPU would round in a
        # different order on every call and make simulations differ in the last bits).
        inc = torch.as_tensor(st.in_edges, device=dev)
        for j in range(inc.shape[1]):
            col = inc[:, j]
            dbase = dbase + torch.where(col >= 0, gdt[torch.clamp(col, min=0)], torch.zeros_like(dbase))
    cm = cm_node.double()[torch.as_tensor(st.mp, device=dev)]
    inv_cm = torch.where(torch.as_tensor(st.comp_of_k >= 0, device=dev), 1.0 / cm, torch.zeros_like(cm))
    return torch.stack([lo, up, dbase, inv_cm], 1).to(dtype).contiguous()


def _node_float_table_np(st: Structure, g: np.ndarray, cm_node: np.ndarray, dt: float) -> np.ndarray:
    """`node_float_table` on the host (numpy; the same operations in the same order)."""
    n = st.n
    gdt = dt * g
    has_par = st.parent_k >= 0
    if len(gdt):
        lo = np.where(has_par, -gdt[st.lo_edge], 0.0)
        up = np.where(has_par, -gdt[st.up_edge], 0.0)
    else:
        lo = up = np.zeros(n)
    dbase = (st.comp_of_k >= 0).astype(np.float64)
    if len(gdt):
        inc = st.in_edges
        for j in range(inc.shape[1]):
            col = inc[:, j]
            dbase = dbase + np.where(col >= 0, gdt[np.maximum(col, 0)], 0.0)
    cm = cm_node[st.mp]
    is_comp = st.comp_of_k >= 0
    inv_cm = np.where(is_comp, 1.0 / np.where(is_comp, cm, 1.0), 0.0)
    return np.ascontiguousarray(np.stack([lo, up, dbase, inv_cm], 1))


@dataclass
class LaunchConfig:
    S: int
    threads: int
    stream: bool = False       # streaming kernel even if the items fit the registers
    gcoef: bool = False        # streaming kernel reading the coupling coefficients from HBM / L2
    sym: bool = False          # streaming kernel solving the row-scaled symmetric system (one
                               # coefficient per node in LDS, the residual's pairs from L2)


# Register-resident items per thread with a forward kernel, by the element sizes of
# (gates, solve) and the launch bound: threads <= 1024 (128 VGPRs) or <= 512 (256 VGPRs).
REGS_IPT = {
    (4, 4): {1024: (1, 2, 3, 4, 6, 8, 10, 12), 512: (16,)},
    (8, 8): {1024: (1, 2, 3, 4), 512: (6, 8, 12)},
    (8, 4): {1024: (1, 2, 3, 4), 512: (6, 8, 12)},
}


def sched_bytes(nseg: int, L: int, ncs: int) -> int:
    """LDS bytes of a staged solve schedule (segments, level pointers, segment parents,
    child segment tops), rounded to 16."""
    return (16 * nseg + 4 * (L + 2 + nseg + ncs) + 15) & ~15


def lds_bytes(n: int, n_tgt: int, S: int, esize: int, sched: int, gcoef: bool = False,
              sym: bool = False) -> int:
    """Forward: the staged schedule, lower / upper coefficient per node (unless read from
    HBM / L2; with `sym` one symmetric coefficient per node, rounded up to 16 bytes),
    diagonal and rhs per (node, simulation), synapse terms per (target, simulation)."""
    coefs = (n * esize + 15) // 16 * 16 if sym else 0 if gcoef else 2 * n * esize
    return sched + coefs + (2 * n + 2 * n_tgt) * S * esize


def items_per_thread(n: int, S: int, threads: int, gsize: int, lsize: int) -> int:
    """Register-resident items per thread of the forward: the smallest count with a
    kernel that holds all n * S items of a program, or 0 (streaming) if none does."""
    need = -(-n * S // threads)
    sets = REGS_IPT[(gsize, lsize)]
    avail = sorted(set(sets[1024]) | (set(sets[512]) if threads <= 512 else set()))
    for c in avail:
        if c >= need:
            return c
    return 0


def config_candidates(n: int, n_tgt: int, B: int, gsize: int, lsize: int,
                      sched: int = 0, stream_only: bool = False) -> List[LaunchConfig]:
    """Launch shapes worth measuring: S simulations per program (a power of two, LDS
    permitting, at least 256 programs when the batch allows) with every thread count
    whose registers hold the items of a program, and the streaming kernel at each S
    (with the coupling coefficients in HBM / L2 when only that fits the LDS). n, n_tgt
    and sched are those of the largest part; `stream_only` for the lockstep programs
    of a split network."""
    out = []
    S = 1
    while S <= 64:
        if S > 1 and S * 256 > B:
            break
        fits = lds_bytes(n, n_tgt, S, lsize, sched) <= 160 * 1024
        fits_s = lds_bytes(n, n_tgt, S, lsize, sched, sym=True) <= 160 * 1024
        fits_g = lds_bytes(n, n_tgt, S, lsize, sched, gcoef=True) <= 160 * 1024
        if not fits_g:
            break
        if fits and not stream_only:
            for thr in (256, 512, 1024):
                if items_per_thread(n, S, thr, gsize, lsize):
                    out.append(LaunchConfig(S, thr))
        thr = int(min(1024, max(64, 64 * -(-n * S // (64 * 16)))))
        for t in sorted({thr, 1024}):
            if not fits and fits_s:
                # Only the symmetric system's coefficients fit beside the system (a
                # structure that cannot be symmetrized takes the L2 pairs instead).
                out.append(LaunchConfig(S, t, stream=True, sym=True))
            out.append(LaunchConfig(S, t, stream=True, gcoef=not fits))
        S *= 2
    if not out:
        raise NotImplementedError(f"a topology of {n} nodes does not fit the fused kernel's LDS")
    return out


def default_config(n: int, n_tgt: int, B: int, esize: int, gsize: int = 0,
                   sched: int = 0, stream_only: bool = False) -> LaunchConfig:
    """A launch shape without measurement: the largest S with a register-resident shape
    (fewest threads), else the largest streaming S."""
    gsize = gsize or esize
    cands = config_candidates(n, n_tgt, B, gsize, esize, sched, stream_only)
    regs = [c for c in cands if not c.stream]
    if regs:
        return max(regs, key=lambda c: (c.S, -c.threads))
    return max(cands, key=lambda c: c.S)


ADJOINT_THREADS_MAX = 512


def adjoint_threads(cfg: "AdjointConfig", gsize: int) -> int:
    """Threads of the reverse pass: the config's, or with the adjoints in HBM (ipt 0) the
    kernel's larger instance (more waves in flight): 768 with float64 gates, 1024 with
    float32."""
    return cfg.threads if cfg.ipt else (768 if gsize == 8 else 1024)
# Voltage adjoints per thread the reverse pass keeps in VGPRs (its kernels' instances);
# 0: in HBM, for programs with more items than that.
ADJ_IPT = (1, 2, 4, 8, 12, 16)


@dataclass
class AdjointConfig:
    S: int
    threads: int
    ipt: int          # items (voltage adjoints) per thread
    gcoef: bool       # coupling coefficients read from the node table (L2) instead of LDS


def adjoint_lds_bytes(n: int, S: int, esize: int, sched: int, gcoef: bool = False, n_tgt: int = 0) -> int:
    """LDS of one reverse-pass program: the staged schedule, the tree's two coupling
    coefficients per node (unless read from L2), the transposed system (right-hand side
    and pivots per (node, simulation)) and two synapse terms per (target, simulation)."""
    return sched + (0 if gcoef else 2 * n * esize) + (2 * n + 2 * n_tgt) * S * esize


def adjoint_config(n: int, B: int, esize: int, sched: int = 0, S: Optional[int] = None,
                   n_tgt: int = 0) -> AdjointConfig:
    """The reverse pass's launch shape: the largest S (at least 256 programs when the
    batch allows) whose transposed system fits the LDS, the coefficients in LDS when they
    fit too, up to 512 threads, and VGPRs for the thread's voltage adjoints; or the shape
    for a given S (the S of the checkpointed forward)."""
    if S is None:
        S = 1
        while S < 64 and (S * 2) * 256 <= max(B, 256):
            S *= 2
        cands = [S >> k for k in range(S.bit_length())]
    else:
        cands = [S]
    for s in cands:
        for gcoef in (False, True):
            if adjoint_lds_bytes(n, s, esize, sched, gcoef, n_tgt) > 160 * 1024:
                continue
            threads = int(min(ADJOINT_THREADS_MAX, max(64, 64 * -(-n * s // 64))))
            need = -(-n * s // threads)
            ipt = next((c for c in ADJ_IPT if c >= need), 0)
            return AdjointConfig(s, threads, ipt, gcoef)
    raise NotImplementedError(f"a topology of {n} nodes does not fit the reverse pass's LDS")


def grad_config(n: int, B: int, esize: int, sched: int = 0, n_tgt: int = 0) -> LaunchConfig:
    """A forward launch shape for the checkpointed forward when none is measured: the S
    of the reverse pass (the checkpoints are in its program layout), the streaming kernel
    (coefficients from L2 when the LDS copy does not fit)."""
    S = adjoint_config(n, B, esize, sched, n_tgt=n_tgt).S
    fits = lds_bytes(n, n_tgt, S, esize, sched) <= 160 * 1024
    sym = not fits and lds_bytes(n, n_tgt, S, esize, sched, sym=True) <= 160 * 1024
    return LaunchConfig(S, 1024, stream=True, gcoef=not fits and not sym, sym=sym)


# Share of the free device memory the voltage trajectory of a checkpointed forward may
# take (JAXHIP_TRAJ_FRACTION; 0 disables it: the reverse pass then recomputes the
# voltages between checkpoints with tree solves).
TRAJ_FRACTION = float(os.environ.get("JAXHIP_TRAJ_FRACTION", "0.5"))


def keep_trajectory(nbytes: int, device) -> bool:
    if TRAJ_FRACTION <= 0 or torch.device(device).type != "cuda":
        return False
    free, _ = torch.cuda.mem_get_info(device)
    return nbytes <= TRAJ_FRACTION * free


def symmetric_rows(st: Structure, node_f: torch.Tensor) -> Optional[torch.Tensor]:
    """(n, 2) per kernel node (s, w) of the row-scaled symmetric system, in node_f's
    dtype and device, or None when the coupling cannot be symmetrized (an edge with
    coefficients of opposite signs, or one zero and one not). Row k of the system is
    scaled by w_k > 0: w = 1 at the roots and w_child = w_parent * upper / lower of the
    child's edge (the lower coefficient couples the child's row to its parent, the upper
    the parent's row to the child), so that s_child = w_child * lower = w_parent * upper
    couples both rows. The products run root to leaf (depth by depth), every w is one
    rounding from its parent's."""
    f = node_f.detach().to("cpu", torch.float64).numpy()
    lo, up = f[:, 0], f[:, 1]
    par = st.parent_k
    has = par >= 0
    if np.any(has & ((lo == 0) != (up == 0))) or np.any(has & (lo * up < 0)):
        return None
    ratio = np.where(has & (lo != 0), up / np.where(lo != 0, lo, 1.0), 1.0)
    # Root to leaf: every pass sets the nodes whose parent is set.
    w = np.ones(len(par))
    known = ~has
    while not known.all():
        ready = ~known & known[np.maximum(par, 0)]
        w[ready] = w[par[ready]] * ratio[ready]
        known |= ready
    sym = np.stack([np.where(has, w * lo, 0.0), w], 1)
    return torch.from_numpy(sym).to(device=node_f.device, dtype=node_f.dtype)


class ForwardLaunch:
    """All device buffers and the argument block of one fused forward launch."""

    def __init__(self, st: Structure, node_f: torch.Tensor, node_p: torch.Tensor,
                 bparam: Optional[torch.Tensor], bmode: int,
                 v0: torch.Tensor, gates0: Tuple[torch.Tensor, ...], stim: Optional[torch.Tensor],
                 stim_scale: Optional[torch.Tensor], T: int, dt: float, B: int, cfg: LaunchConfig,
                 syn_params: Optional[List[torch.Tensor]] = None, syn_scale: Optional[torch.Tensor] = None,
                 syn_s0: Optional[torch.Tensor] = None, ckpt_every: int = 0,
                 dt_v: float = 0.0, rhs_scale: float = 0.0, explicit_step: int = 0,
                 node_p2: Optional[torch.Tensor] = None, gates_ext0: Optional[Tuple[torch.Tensor, ...]] = None):
        dev = v0.device
        dtype = gates0[0].dtype
        S = cfg.S
        self.st, self.cfg, self.B, self.T, self.S, self.dtype = st, cfg, B, T, S, dtype
        # P groups of S simulations; a group is one program per part of the structure.
        self.P = P = (B + S - 1) // S
        n = st.n
        parts = st.parts
        # Lockstep programs (grid-wide barrier per step): parts joined by synapses.
        self.coop = coop = parts > 1 and st.syn is not None
        self.v = to_prog(v0.double(), S)
        self.gm, self.gh, self.gn = (to_prog(g, S) for g in gates0)
        nrec = max(st.n_rec, 1)
        # Recordings per group: every part's block [T + 1][its recordings][S].
        self.rec = torch.empty((P, (T + 1) * nrec * S), dtype=dtype, device=dev)
        g_i = np.repeat(np.arange(P), parts)
        q_i = np.tile(np.arange(parts), P)
        progs = np.zeros((P * parts, _lib.JXHIP_PROG_COLS), dtype=np.int64)
        progs[:, 0] = q_i
        progs[:, 1] = np.minimum(S, B - g_i * S)
        progs[:, 3] = T
        progs[:, 4] = (g_i * n + st.part_nb[q_i]) * S
        progs[:, 5] = g_i * S
        progs[:, 6] = g_i * (T + 1) * nrec * S + (T + 1) * st.part_rb[q_i] * S
        n_syn = st.syn["n_syn"] if st.syn else 0
        if st.syn:
            progs[:, 7] = (g_i * n_syn + st.syn["part_synb"][q_i]) * S
        self.progs = torch.as_tensor(progs, device=dev)
        self.topos = torch.as_tensor(st.topo_row, device=dev)
        self.node_f = node_f
        self.node_p = node_p
        self.bparam = bparam
        self.bmode = bmode
        self.stim = stim
        self.stim_scale = stim_scale
        self.syn_params = syn_params
        self.syn_scale = syn_scale
        self.syn_s = None
        if st.syn is not None and syn_has_state(st.syn):
            self.syn_s = to_prog(syn_s0, S)
        self.ckpt_every = ckpt_every
        self.n_ckpt = 0
        if ckpt_every > 0:
            C = (T + ckpt_every - 1) // ckpt_every
            self.n_ckpt = C
            self.ckpt = [torch.empty((C,) + self.v.shape, dtype=torch.float64, device=dev)]
            self.ckpt += [torch.empty((C,) + self.gm.shape, dtype=dtype, device=dev) for _ in range(3)]
            self.ckpt_s = torch.empty((C,) + self.syn_s.shape, dtype=dtype, device=dev) if self.syn_s is not None else None
        a = _lib.FusedArgs()
        a.progs = self.progs.data_ptr()
        a.n_progs = P * parts
        a.parts = parts
        a.group_nodes = n
        a.coop = int(coop)
        self.vx = None
        self.gsync = None
        if coop:
            # Exported voltages of every part by step parity (presynaptic reads).
            self.vx = torch.empty((2,) + self.v.shape, dtype=torch.float64, device=dev)
            a.vx = self.vx.data_ptr()
            a.vx_stride = self.v.numel()
            # One barrier counter per group of programs (the parts of a group wait for
            # each other only) and an error word.
            self.gsync = torch.zeros(P + 1, dtype=torch.int32, device=dev)
            a.group_sync = self.gsync.data_ptr()
        a.topos = self.topos.data_ptr()
        a.seg = st.seg.data_ptr()
        a.seg_level_ptr = st.seg_level_ptr.data_ptr()
        a.seg_child = st.seg_child.data_ptr()
        a.node_i = st.node_i.data_ptr()
        a.node_f = node_f.data_ptr()
        a.child_idx = st.child_idx.data_ptr()
        a.stim_col = st.stim_col.data_ptr()
        a.rec_node = st.rec_node.data_ptr()
        self._dummy = torch.zeros(1, dtype=dtype, device=dev)
        a.node_p = node_p.data_ptr()
        a.bparam = bparam.data_ptr() if bparam is not None else self._dummy.data_ptr()
        a.bparam_mode = bmode if bparam is not None else 0
        a.v = self.v.data_ptr()
        a.gate_m, a.gate_h, a.gate_n = self.gm.data_ptr(), self.gh.data_ptr(), self.gn.data_ptr()
        self.gx = self.node_p2 = None
        if node_p2 is not None:
            # Pospischil gates: four arrays in the program layout of gate_m, one buffer.
            self.node_p2 = node_p2
            self.gx = torch.stack([to_prog(g, S) for g in gates_ext0]).contiguous()
            a.node_p2 = node_p2.data_ptr()
            a.gate_ext = self.gx.data_ptr()
            a.gate_ext_stride = self.gm.numel()
        if stim is not None:
            a.stim = stim.data_ptr()
            a.stim_cols = stim.shape[1]
            a.stim_b = stim.shape[2]
            a.stim_scale = stim_scale.data_ptr()
        else:
            a.stim = self._dummy.data_ptr()
            a.stim_cols = 0
            a.stim_b = 1
            a.stim_scale = self._dummy.data_ptr()
        a.rec = self.rec.data_ptr()
        if st.syn is not None:
            a.syn_model = st.syn["model"]
            a.syn_pre = st.syn["pre"].data_ptr()
            a.syn_ptr = st.syn["ptr"].data_ptr()
            a.tgt_node = st.syn["tgt"].data_ptr()
            for j in range(_lib.JXHIP_NSYNPARAM):
                a.syn_param[j] = syn_params[min(j, len(syn_params) - 1)].data_ptr()
            a.syn_scale = syn_scale.data_ptr()
            a.syn_s = self.syn_s.data_ptr() if self.syn_s is not None else self._dummy.data_ptr()
            if st.syn["kind"] is not None:
                a.syn_kind = st.syn["kind"].data_ptr()
        else:
            a.syn_model = -1
        a.sims_total = B
        a.dt = dt
        # Crank-Nicolson: solve with dt / 2 (node_f and stim_scale built with it) and a
        # doubled right-hand side; 0 = backward Euler.
        a.dt_v = dt_v
        a.rhs_scale = rhs_scale
        # Forward Euler: no tree solve (branchpoint rows only).
        a.explicit_step = explicit_step
        a.diff = SYN_DIFF
        a.S = S
        a.threads = cfg.threads
        nmax, n_tgt = st.n_max, st.tgt_max
        stream = cfg.stream or coop
        a.items_per_thread = 0 if stream else items_per_thread(
            nmax, S, cfg.threads, self.gm.element_size(), node_f.element_size())
        esz = node_f.element_size()
        fits = lds_bytes(nmax, n_tgt, S, esz, st.sched) <= 160 * 1024
        # Where the solve finds the tree's coupling coefficients: the pairs in LDS, the
        # symmetric system's coefficients in LDS (when asked for, or when only they fit
        # beside the system), or the pairs in L2.
        self.node_sym = None
        if stream and not explicit_step and (cfg.sym or (not fits and not cfg.gcoef)) and \
                lds_bytes(nmax, n_tgt, S, esz, st.sched, sym=True) <= 160 * 1024:
            self.node_sym = symmetric_rows(st, node_f)
        if self.node_sym is not None:
            a.coef_global = 2
            a.node_sym = self.node_sym.data_ptr()
            a.lds_bytes = lds_bytes(nmax, n_tgt, S, esz, st.sched, sym=True)
        else:
            a.coef_global = int(stream and (cfg.gcoef or cfg.sym or not fits))
            a.lds_bytes = lds_bytes(nmax, n_tgt, S, esz, st.sched, bool(a.coef_global))
        # Scratch of the streaming kernel: the axial part of every item's residual.
        self.work_ax = torch.empty_like(self.gm) if a.items_per_thread == 0 else None
        a.work_ax = self.work_ax.data_ptr() if self.work_ax is not None else 0
        a.max_items = nmax * S
        if ckpt_every > 0:
            a.ckpt_v, a.ckpt_m, a.ckpt_h, a.ckpt_n = (c.data_ptr() for c in self.ckpt)
            a.ckpt_s = self.ckpt_s.data_ptr() if self.ckpt_s is not None else 0
            a.ckpt_every = ckpt_every
            a.ckpt_stride = self.v.numel()
            a.ckpt_syn_stride = self.syn_s.numel() if self.syn_s is not None else 0
        # The voltage of every step (for the reverse pass: no recomputed tree solves) when
        # the trajectory fits the memory budget.
        self.traj = None
        if ckpt_every > 0 and keep_trajectory(
                (T + 1) * self.v.numel() * self.gm.element_size(), dev):
            # In the gate type: the voltages the gates, synapses and linearization use.
            self.traj = torch.empty((T + 1) * self.v.numel(), dtype=self.gm.dtype, device=dev)
            a.traj_v = self.traj.data_ptr()
            a.traj_steps = T + 1
        self.args = a

    def kernel_name(self, base: str):
        if self.dtype == torch.float64 and self.node_f.dtype == torch.float32:
            return _lib.fn(base + "_mixed")
        return _lib.fn(base, self.dtype)

    def launch(self):
        f = self.kernel_name("jxhip_fused_forward")
        if self.gsync is not None:
            self.gsync.zero_()
        _lib.check(f(self.args, _lib.stream_ptr(self.v.device)), "fused_forward")
        self.check_group_sync()

    def check_group_sync(self):
        """Raises if a group barrier of the last launch gave up waiting (waits for it; not
        while a graph is being captured, where the word is left for the caller)."""
        if self.gsync is None or torch.cuda.is_current_stream_capturing():
            return
        if int(self.gsync[-1]) != 0:
            raise RuntimeError("a group barrier of the lockstep programs gave up waiting")

    def recordings(self) -> torch.Tensor:
        """(B, n_rec, T + 1), the recordings in their order."""
        st, P, S, T1 = self.st, self.P, self.S, self.T + 1
        blocks = []
        for q in range(st.parts):
            rb, nr = int(st.part_rb[q]), int(st.part_nrec[q])
            b = self.rec[:, T1 * rb * S:T1 * (rb + nr) * S].reshape(P, T1, nr, S)
            blocks.append(b.permute(0, 3, 2, 1).reshape(P * S, nr, T1))
        r = torch.cat(blocks, 1) if len(blocks) > 1 else blocks[0]
        if st.parts > 1:
            r = r[:, torch.as_tensor(np.argsort(st.rec_order), device=r.device)]
        return r[: self.B]

    def recordings_to_prog(self, g: torch.Tensor) -> torch.Tensor:
        """(B, n_rec, T + 1) -> the layout of the recordings buffer (gradients)."""
        st, P, S, T1 = self.st, self.P, self.S, self.T + 1
        if P * S != self.B:
            g = torch.cat([g, g.new_zeros(P * S - self.B, *g.shape[1:])], 0)
        if st.parts > 1:
            g = g[:, torch.as_tensor(st.rec_order, device=g.device)]
        blocks = []
        for q in range(st.parts):
            rb, nr = int(st.part_rb[q]), int(st.part_nrec[q])
            b = g[:, rb:rb + nr].reshape(P, S, nr, T1).permute(0, 3, 2, 1)
            blocks.append(b.reshape(P, -1))
        out = torch.cat(blocks, 1) if len(blocks) > 1 else blocks[0]
        if out.shape[1] < self.rec.shape[1]:
            out = torch.cat([out, out.new_zeros(P, self.rec.shape[1] - out.shape[1])], 1)
        return out.contiguous()
