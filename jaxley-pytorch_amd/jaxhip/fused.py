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
    """Integer tables of one topology in kernel order (cached per module structure)."""

    n: int
    L: int
    mp: np.ndarray             # kernel index -> node
    kidx_of_node: np.ndarray   # node -> kernel index
    comp_of_k: np.ndarray      # kernel index -> compartment (-1 branchpoint)
    topo_row: np.ndarray
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
    sched: int = 0             # bytes of the solve schedule staged in LDS
    rec_first: Optional[torch.Tensor] = None   # per kernel node: its first recording (reverse pass)
    rec_next: Optional[torch.Tensor] = None    # per recording: the node's next recording


def build_structure(module, rec_comps: np.ndarray, stim_nodes: np.ndarray, device) -> Structure:
    tables, indexer = module._ensure_solver()
    n = tables["n_nodes"]
    mp_solve = np.asarray(indexer["map_to_solve_order"], dtype=np.int64)
    inv_solve = np.asarray(indexer["inv_map_to_solve_order"], dtype=np.int64)
    parent_solve = np.asarray(indexer["parent_lookup"][:n], dtype=np.int64)
    if check_forest(parent_solve) is None:
        raise ValueError("the solve order must be a forest with parents before children")
    sg = segments(parent_solve)
    korder = sg["korder"]                           # kernel -> solve
    kidx_solve = sg["kidx"]                         # solve -> kernel
    mp = mp_solve[korder]                           # kernel -> node
    kidx_of_node = kidx_solve[inv_solve]            # node -> kernel
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
    # Children of every kernel node (CSR, in increasing kernel index).
    ch_nodes = np.nonzero(parent_k >= 0)[0]
    srt = np.lexsort((ch_nodes, parent_k[ch_nodes]))
    ch_nodes = ch_nodes[srt]
    cnt = np.bincount(parent_k[ch_nodes], minlength=n)
    cbeg = np.zeros(n, dtype=np.int64)
    np.cumsum(cnt[:-1], out=cbeg[1:])
    # Stimulus entries grouped per kernel node; entry e reads stimulus column stim_order[e].
    stim_k = kidx_of_node[stim_nodes] if len(stim_nodes) else np.zeros(0, np.int64)
    stim_order = np.argsort(stim_k, kind="stable")
    scnt = np.bincount(stim_k, minlength=n) if len(stim_k) else np.zeros(n, np.int64)
    sbeg = np.zeros(n, dtype=np.int64)
    np.cumsum(scnt[:-1], out=sbeg[1:])
    node_i = np.zeros((n, _lib.JXHIP_NODE_ICOLS), dtype=np.int32)
    node_i[:, 0] = parent_k
    node_i[:, 1] = cbeg
    node_i[:, 2] = cbeg + cnt
    node_i[:, 3] = flags
    node_i[:, 4] = sbeg
    node_i[:, 5] = sbeg + scnt
    node_i[:, 6] = -1
    syn = _build_synapses(module, kidx_of_node, n, node_i)
    rec_k = kidx_of_node[module._comp_node[rec_comps]] if len(rec_comps) else np.zeros(0, np.int64)
    # Nodes whose voltage other threads read from HBM during the time loop.
    node_i[rec_k, 3] |= FLAG_EXPORT
    if syn is not None:
        node_i[syn["pre"], 3] |= FLAG_EXPORT
        node_i[syn["tgt"], 3] |= FLAG_EXPORT
    topo_row = np.zeros(_lib.JXHIP_TOPO_COLS, dtype=np.int32)
    topo_row[:16] = [n, sg["L"], 0, 0, 0, len(rec_k), 0, syn["n_tgt"] if syn else 0, 0, 0, 0,
                     syn["n_syn"] if syn else 0, 0, 0, 0, 0]
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.int32, device=device)
    # The recordings of every node as a list (first, next; the reverse pass adds their
    # gradients).
    none = np.iinfo(np.int32).max
    rec_first = np.full(n, none, dtype=np.int64)
    rec_next = np.full(max(len(rec_k), 1), none, dtype=np.int64)
    for r in range(len(rec_k) - 1, -1, -1):
        rec_next[r] = rec_first[rec_k[r]]
        rec_first[rec_k[r]] = r
    if syn is not None:
        # The synapses grouped by presynaptic node: every synapse's slot in the grouping
        # (the reverse pass stores its presynaptic share there and sums each node's
        # contiguous slots, without atomics).
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
        n=n, L=sg["L"], mp=mp, kidx_of_node=kidx_of_node, comp_of_k=comp_of_k, topo_row=topo_row,
        seg=t(sg["seg"]), seg_level_ptr=t(sg["seg_level_ptr"]), seg_child=t(sg["seg_child"]),
        node_i=t(node_i), child_idx=t(ch_nodes if len(ch_nodes) else np.zeros(1, np.int64)),
        stim_col=t(stim_order if len(stim_order) else np.zeros(1, np.int64)),
        rec_node=t(rec_k if len(rec_k) else np.zeros(1, np.int64)),
        syn=syn, n_rec=len(rec_k), rec_comps=rec_comps, stim_order=stim_order,
        sinks_k=kidx_of_node[tables["sink"]],
        in_edges=_in_edges(kidx_of_node[tables["sink"]], n),
        lo_edge=np.asarray(indexer["map_to_solve_order_lower"])[korder],
        up_edge=np.asarray(indexer["map_to_solve_order_upper"])[korder],
        parent_k=parent_k,
        sched=sched_bytes(len(sg["seg"]), sg["L"], int(sg["seg"][-1, 3]) if len(sg["seg"]) else 0),
        rec_first=t(rec_first), rec_next=t(rec_next),
    )


def _build_synapses(module, kidx_of_node: np.ndarray, n: int, node_i: np.ndarray) -> Optional[Dict]:
    """All synapses of the network as one list in jagged-diagonal order: the targets
    (postsynaptic nodes) numbered by non-increasing in-degree (ties by node), and the
    list holding every target's first synapse, then every target's second one, and so
    on, so that the j-th synapse of target t is `ptr[j] + t` (for t below
    `ptr[j + 1] - ptr[j]`; `ptr` ends with a repeated entry). Threads of consecutive
    targets then read consecutive synapses. A target's synapses keep the order of the
    types in the network and of the edges within a type (the order the currents are
    summed in). Several types give the per-synapse model ids of JXHIP_SYN_MIXED."""
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
    order = np.argsort(post, kind="stable")         # synapses grouped by target
    targets, counts = np.unique(post[order], return_counts=True)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    rank = np.argsort(-counts, kind="stable")        # targets by non-increasing in-degree
    targets, counts, first = targets[rank], counts[rank], first[rank]
    per_j = np.array([int((counts > j).sum()) for j in range(int(counts.max()))], dtype=np.int64)
    ptr = np.concatenate([[0], np.cumsum(per_j), [len(post)]]).astype(np.int32)
    perm = np.empty(len(post), dtype=np.int64)
    for j, cnt in enumerate(per_j):
        perm[ptr[j]:ptr[j] + cnt] = order[first[:cnt] + j]
    node_i[targets, 6] = np.arange(len(targets), dtype=np.int32)
    mixed = len(groups) > 1
    return {
        "model": _lib.JXHIP_SYN_MIXED if mixed else int(groups[0]["type"].kernel_model),
        "groups": groups,
        "kind": kind[perm] if mixed else None,
        "perm": perm,
        "pre": pre[perm].astype(np.int32),
        "ptr": ptr,
        "tgt": targets.astype(np.int32),
        "post_node": post_nodes[perm],
        "n_tgt": len(targets),
        "n_syn": len(post),
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


@dataclass
class LaunchConfig:
    S: int
    threads: int
    stream: bool = False       # streaming kernel even if the items fit the registers
    gcoef: bool = False        # streaming kernel reading the coupling coefficients from HBM / L2


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


def lds_bytes(n: int, n_tgt: int, S: int, esize: int, sched: int, gcoef: bool = False) -> int:
    """Forward: the staged schedule, lower / upper coefficient per node (unless read from
    HBM / L2), diagonal and rhs per (node, simulation), synapse terms per (target,
    simulation)."""
    return sched + (0 if gcoef else 2 * n * esize) + (2 * n + 2 * n_tgt) * S * esize


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
                      sched: int = 0) -> List[LaunchConfig]:
    """Launch shapes worth measuring: S simulations per program (a power of two, LDS
    permitting, at least 256 programs when the batch allows) with every thread count
    whose registers hold the items of a program, and the streaming kernel at each S
    (with the coupling coefficients in HBM / L2 when only that fits the LDS)."""
    out = []
    S = 1
    while S <= 64:
        if S > 1 and S * 256 > B:
            break
        fits = lds_bytes(n, n_tgt, S, lsize, sched) <= 160 * 1024
        fits_g = lds_bytes(n, n_tgt, S, lsize, sched, gcoef=True) <= 160 * 1024
        if not fits_g:
            break
        if fits:
            for thr in (256, 512, 1024):
                if items_per_thread(n, S, thr, gsize, lsize):
                    out.append(LaunchConfig(S, thr))
        thr = int(min(1024, max(64, 64 * -(-n * S // (64 * 16)))))
        for t in sorted({thr, 1024}):
            out.append(LaunchConfig(S, t, stream=True, gcoef=not fits))
        S *= 2
    if not out:
        raise NotImplementedError(f"a topology of {n} nodes does not fit the fused kernel's LDS")
    return out


def default_config(n: int, n_tgt: int, B: int, esize: int, gsize: int = 0,
                   sched: int = 0) -> LaunchConfig:
    """A launch shape without measurement: the largest S with a register-resident shape
    (fewest threads), else the largest streaming S."""
    gsize = gsize or esize
    cands = config_candidates(n, n_tgt, B, gsize, esize, sched)
    regs = [c for c in cands if not c.stream]
    if regs:
        return max(regs, key=lambda c: (c.S, -c.threads))
    return max(cands, key=lambda c: c.S)


ADJOINT_THREADS_MAX = 512
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
    return LaunchConfig(S, 1024, stream=True, gcoef=not fits)


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
        self.P = P = (B + S - 1) // S
        n = st.n
        self.v = to_prog(v0.double(), S)
        self.gm, self.gh, self.gn = (to_prog(g, S) for g in gates0)
        self.rec = torch.empty((P, T + 1, max(st.n_rec, 1), S), dtype=dtype, device=dev)
        progs = np.zeros((P, _lib.JXHIP_PROG_COLS), dtype=np.int64)
        p = np.arange(P)
        progs[:, 1] = np.minimum(S, B - p * S)
        progs[:, 3] = T
        progs[:, 4] = p * n * S
        progs[:, 5] = p * S
        progs[:, 6] = p * (T + 1) * max(st.n_rec, 1) * S
        n_syn = st.syn["n_syn"] if st.syn else 0
        progs[:, 7] = p * n_syn * S
        self.progs = torch.as_tensor(progs, device=dev)
        self.topos = torch.as_tensor(st.topo_row[None], device=dev)
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
        a.n_progs = P
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
        a.items_per_thread = 0 if cfg.stream else items_per_thread(
            n, S, cfg.threads, self.gm.element_size(), node_f.element_size())
        n_tgt = st.syn["n_tgt"] if st.syn else 0
        a.coef_global = int(cfg.stream and cfg.gcoef)
        a.lds_bytes = lds_bytes(n, n_tgt, S, node_f.element_size(), st.sched, bool(a.coef_global))
        # Scratch of the streaming kernel: the axial part of every item's residual.
        self.work_ax = torch.empty_like(self.gm) if a.items_per_thread == 0 else None
        a.work_ax = self.work_ax.data_ptr() if self.work_ax is not None else 0
        a.max_items = n * S
        if ckpt_every > 0:
            a.ckpt_v, a.ckpt_m, a.ckpt_h, a.ckpt_n = (c.data_ptr() for c in self.ckpt)
            a.ckpt_s = self.ckpt_s.data_ptr() if self.ckpt_s is not None else 0
            a.ckpt_every = ckpt_every
            a.ckpt_stride = self.v.numel()
            a.ckpt_syn_stride = self.syn_s.numel() if self.syn_s is not None else 0
        self.args = a

    def kernel_name(self, base: str):
        if self.dtype == torch.float64 and self.node_f.dtype == torch.float32:
            return _lib.fn(base + "_mixed")
        return _lib.fn(base, self.dtype)

    def launch(self):
        f = self.kernel_name("jxhip_fused_forward")
        _lib.check(f(self.args, _lib.stream_ptr(self.v.device)), "fused_forward")

    def recordings(self) -> torch.Tensor:
        """(B, n_rec, T + 1)."""
        r = self.rec[..., : self.st.n_rec, :]                 # (P, T+1, n_rec, S)
        r = r.permute(0, 3, 2, 1).reshape(self.P * self.S, self.st.n_rec, self.T + 1)
        return r[: self.B]
