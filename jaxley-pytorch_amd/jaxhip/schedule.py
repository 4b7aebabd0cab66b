"""Morphology tables, the DHS solve order and the kernel schedules (host side, numpy).

Everything here is integer work done once per structure; it must reproduce the
reference bit for bit:

* compartment edges and their row order: `Cell._init_comp_graph`
  (jaxley_pytorch/modules/cell.py:145-225), `Branch._init_comp_graph` (branch.py:89-107),
  `Network._init_comp_graph` (network.py:149-190);
* the DHS solve order: `to_graph` (modules/base.py:4364-4459) -> networkx
  `DiGraph.to_undirected` -> `bfs_edge_hops` (utils/solver_utils.py:187-223) with
  `allowed_nodes_per_level` (32 by default, modules/base.py:1175-1181), then
  `dhs_permutation_indices` (solver_utils.py:80-132) and the level grouping
  (solver_utils.py:226-251); networks concatenate cells (network.py:192-293).

The kernel schedules (`forest_schedule`, `FusedTopology`, `segments`) are this
package's own: heights, height-sorted node order and child lists (CSR) of the forest in
solve order, and the segment order of the fused kernels.

The breadth-first walk with its level hops and the segment order run in the native host
library (`libjxhost.so`, include/jxhip_host.h): they are the per-node loops of a
request's preparation.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

# ------------------------------------------------------------------------------------
# Compartment edges
# ------------------------------------------------------------------------------------


def cell_tables(parents: Sequence[int], ncomp_per_branch: Sequence[int]) -> Dict:
    """Compartment edges of a cell (types 0..4, reference row order).

    Type 0: comp <-> comp within a branch (forward edges of every branch first, then
    the backward ones, branch by branch); 1: branchpoint -> last comp of the parent
    branch; 2: branchpoint -> first comp of a child branch; 3, 4: their mirror images.
    Branchpoints are numbered after all compartments, one per distinct parent branch
    in increasing order.
    """
    parents = np.asarray(parents, dtype=np.int64)
    ncomp = np.asarray(ncomp_per_branch, dtype=np.int64)
    nb = len(parents)
    cs = np.zeros(nb + 1, dtype=np.int64)
    np.cumsum(ncomp, out=cs[1:])
    n_comp = int(cs[-1])
    fwd_s, fwd_t, bwd_s, bwd_t = [], [], [], []
    src_parts, snk_parts, ord_parts = [], [], []
    for b in range(nb):
        c, n = int(cs[b]), int(ncomp[b])
        a = np.arange(c, c + n - 1, dtype=np.int64)
        src_parts += [a, a + 1]
        snk_parts += [a + 1, a]
        ord_parts += [np.ones(n - 1, np.int64), np.zeros(n - 1, np.int64)]
    par = parents[1:]
    child = np.arange(1, nb, dtype=np.int64)
    if len(par):
        par_u, child_bp = np.unique(par, return_inverse=True)
    else:
        par_u = np.zeros(0, np.int64)
        child_bp = np.zeros(0, np.int64)
    nbp = len(par_u)
    t1s = np.arange(nbp, dtype=np.int64) + n_comp
    t1t = cs[par_u + 1] - 1
    t2s = child_bp.astype(np.int64) + n_comp
    t2t = cs[child]
    n0 = int(sum(len(x) for x in src_parts))
    src = np.concatenate(src_parts + [t1s, t2s, t1t, t2t]) if nb else np.zeros(0, np.int64)
    snk = np.concatenate(snk_parts + [t1t, t2t, t1s, t2s]) if nb else np.zeros(0, np.int64)
    ordd = np.concatenate(
        ord_parts
        + [np.zeros(nbp, np.int64), np.ones(len(child), np.int64),
           np.ones(nbp, np.int64), np.zeros(len(child), np.int64)]
    )
    typ = np.concatenate(
        [np.zeros(n0, np.int64), np.full(nbp, 1), np.full(len(child), 2), np.full(nbp, 3),
         np.full(len(child), 4)]
    ).astype(np.int64)
    n_nodes = int(snk.max()) + 1 if len(snk) else 1
    n_nodes = max(n_nodes, n_comp)
    return {
        "source": src.astype(np.int64),
        "sink": snk.astype(np.int64),
        "ordered": ordd.astype(np.int64),
        "type": typ,
        "n_comp": n_comp,
        "n_nodes": n_nodes,
        "internal": np.arange(n_comp, dtype=np.int64),
    }


def _bfs_edge_hops(tables: Dict, allowed_nodes_per_level: int, root: int = 0):
    """BFS tree edges (parent, child) over the undirected compartment graph, visiting
    neighbours in networkx's `to_undirected` adjacency order, with their levels (native:
    `jxhost_bfs_edge_hops`)."""
    from . import _lib

    n = tables["n_nodes"]
    typ, s, t = tables["type"], tables["source"], tables["sink"]
    keep = np.isin(typ, [2, 3]) | ((typ == 0) & (s < t))
    ks = np.ascontiguousarray(s[keep], dtype=np.int64)
    kt = np.ascontiguousarray(t[keep], dtype=np.int64)
    par = np.empty(max(n - 1, 1), dtype=np.int64)
    chi = np.empty_like(par)
    lev = np.empty_like(par)
    dist = np.empty(n, dtype=np.int64)
    code = _lib.host().jxhost_bfs_edge_hops(n, len(ks), ks.ctypes.data, kt.ctypes.data, root,
                                            int(allowed_nodes_per_level), par.ctypes.data, chi.ctypes.data,
                                            dist.ctypes.data, lev.ctypes.data)
    _lib.host_check(code, "the DHS solve order")
    return par[:n - 1], chi[:n - 1], lev[:n - 1]


def dhs_indexer(tables: Dict, allowed_nodes_per_level: int = 32, root: int = 0) -> Dict:
    """DHS solve order of one cell (see module docstring)."""
    n = tables["n_nodes"]
    e_par, e_chi, levels = _bfs_edge_hops(tables, allowed_nodes_per_level, root)
    inv = np.zeros(n, dtype=np.int64)        # node -> solve index
    inv[root] = 0
    inv[e_chi] = np.arange(1, n, dtype=np.int64)
    mapping = np.argsort(inv)                # solve index -> node
    # Lower / upper values: edge k carries value k. In solve order, an edge whose
    # source precedes its sink is a "lower" of the sink (child); otherwise an "upper" of
    # its source (child). Both are sorted by the child's solve index.
    ss = inv[tables["source"]]
    st = inv[tables["sink"]]
    k = np.arange(len(ss), dtype=np.int64)
    is_low = ss < st
    low_vals = k[is_low][np.argsort(st[is_low], kind="stable")]
    up_vals = k[~is_low][np.argsort(ss[~is_low], kind="stable")]
    lu = np.concatenate([low_vals, up_vals])
    node_order = np.zeros((n - 1, 3), dtype=np.int64)
    node_order[:, 0] = inv[e_chi]
    node_order[:, 1] = inv[e_par]
    node_order[:, 2] = levels
    parent_lookup = -np.ones(n + 1, dtype=np.int64)
    parent_lookup[node_order[:, 0]] = node_order[:, 1]
    return _Indexer({
        "map_to_solve_order": mapping,
        "inv_map_to_solve_order": inv,
        "map_to_solve_order_lower": np.concatenate([[0], lu[: n - 1]]).astype(np.int64),
        "map_to_solve_order_upper": np.concatenate([[0], lu[n - 1:]]).astype(np.int64),
        "node_order": node_order,
        "parent_lookup": parent_lookup,
    })


class _Indexer(dict):
    """A DHS indexer; its level grouping (`node_order_grouped`, the reference's solve
    levels) is built when first read."""

    def __missing__(self, key):
        if key == "node_order_grouped":
            self[key] = group_levels(self["node_order"])
            return self[key]
        raise KeyError(key)


def group_levels(node_order: np.ndarray) -> List[np.ndarray]:
    """Rows of `(node, parent, level)` grouped by level (sorted), rows in order."""
    node_order = np.asarray(node_order, dtype=np.int64).reshape(-1, 3)
    if len(node_order) == 0:
        return []
    lev = node_order[:, 2]
    order = np.argsort(lev, kind="stable")
    sorted_rows = node_order[order]
    uniq, starts = np.unique(sorted_rows[:, 2], return_index=True)
    ends = np.append(starts[1:], len(sorted_rows))
    return [sorted_rows[a:b, :2].copy() for a, b in zip(starts, ends)]


def network_indexer(cell_tabs: List[Dict], cell_inds: List[Dict], vectorize: bool = True):
    """Tables and DHS indexer of a network of cells (cells solved in parallel levels
    when `vectorize`, as `_solver_device == "gpu"`)."""
    src, snk, ordd, typ, internal = [], [], [], [], []
    inv, mp, low, up, node_order, plook = [], [], [], [], [], []
    off = 0
    lu_off = 0
    for tab, ind in zip(cell_tabs, cell_inds):
        src.append(tab["source"] + off)
        snk.append(tab["sink"] + off)
        ordd.append(tab["ordered"])
        typ.append(tab["type"])
        internal.append(tab["internal"] + off)
        inv.append(ind["inv_map_to_solve_order"] + off)
        mp.append(ind["map_to_solve_order"] + off)
        low.append(ind["map_to_solve_order_lower"] + lu_off)
        up.append(ind["map_to_solve_order_upper"] + lu_off)
        no = ind["node_order"].copy().reshape(-1, 3)
        if not vectorize or len(no) == 0:
            no = no + off
        else:
            no[:, :2] += off
        node_order.append(no)
        pl = ind["parent_lookup"][:-1].copy()
        pl[1:] += off
        plook.append(pl)
        off += tab["n_nodes"]
        lu_off += (tab["n_nodes"] - 1) * 2
    max_ind = max(lu_off - 1, 0)
    tables = {
        "source": np.concatenate(src),
        "sink": np.concatenate(snk),
        "ordered": np.concatenate(ordd),
        "type": np.concatenate(typ),
        "n_nodes": off,
        "n_comp": int(sum(len(i) for i in internal)),
        "internal": np.concatenate(internal),
    }
    no_all = np.concatenate(node_order, axis=0)
    indexer = _Indexer({
        "inv_map_to_solve_order": np.concatenate(inv),
        "map_to_solve_order": np.concatenate(mp),
        "map_to_solve_order_lower": np.minimum(np.concatenate(low), max_ind),
        "map_to_solve_order_upper": np.minimum(np.concatenate(up), max_ind),
        "node_order": no_all,
        "parent_lookup": np.concatenate(plook + [np.asarray([-1])]),
    })
    return tables, indexer


# ------------------------------------------------------------------------------------
# Kernel schedules
# ------------------------------------------------------------------------------------


def tree_heights(parent: np.ndarray) -> tuple:
    """Depth and height of every node of a forest whose parents precede children."""
    parent = np.asarray(parent, dtype=np.int64)
    n = len(parent)
    depth = np.zeros(n, dtype=np.int64)
    nz = np.nonzero(parent >= 0)[0]
    for i in nz:
        depth[i] = depth[parent[i]] + 1
    height = np.zeros(n, dtype=np.int64)
    for i in nz[::-1]:
        p = parent[i]
        if height[i] + 1 > height[p]:
            height[p] = height[i] + 1
    return depth, height


def check_forest(parent: np.ndarray) -> Optional[np.ndarray]:
    """Tree starts if every tree is contiguous with its root first and parents precede
    children (the requirement of `build_forest_schedule`, tree_solve.py:56-72)."""
    parent = np.asarray(parent, dtype=np.int64)
    n = len(parent)
    roots = np.nonzero(parent < 0)[0]
    if n == 0 or len(roots) == 0 or roots[0] != 0:
        return None
    sizes = np.diff(np.append(roots, n))
    tree_of = np.repeat(np.arange(len(roots)), sizes)
    idx = np.arange(n)
    nr = parent >= 0
    if not (np.all(parent[nr] >= roots[tree_of[nr]]) and np.all(parent[nr] < idx[nr])):
        return None
    return roots


def forest_schedule(parent: np.ndarray, max_block: int) -> Optional[Dict[str, np.ndarray]]:
    """Blocks of consecutive whole trees (<= max_block nodes unless one tree is larger),
    with per-block heights, height-sorted order, level pointers and child lists."""
    parent = np.asarray(parent, dtype=np.int64)
    roots = check_forest(parent)
    if roots is None:
        return None
    n = len(parent)
    ends = np.append(roots[1:], n)
    blocks = []
    bs, be = 0, 0
    for s, e in zip(roots, ends):
        if e - bs > max_block and be > bs:
            blocks.append((bs, be))
            bs = s
        be = e
    blocks.append((bs, be))
    depth, height = tree_heights(parent)
    order = np.zeros(n, dtype=np.int32)
    par_rel = -np.ones(n, dtype=np.int32)
    child_rng = np.zeros((n, 2), dtype=np.int32)
    child_idx = []
    level_ptr = []
    info = np.zeros((len(blocks), 4), dtype=np.int32)
    pos = 0
    for b, (s, e) in enumerate(blocks):
        h = height[s:e]
        o = np.argsort(h, kind="stable")
        order[s:e] = o
        p = parent[s:e]
        par_rel[s:e] = np.where(p >= 0, p - s, -1)
        # children (block-relative), grouped by parent in increasing child order
        ch = np.nonzero(p >= 0)[0]
        pp = p[ch] - s
        srt = np.argsort(pp, kind="stable")
        ch, pp = ch[srt], pp[srt]
        cnt = np.bincount(pp, minlength=e - s)
        beg = np.zeros(e - s, dtype=np.int64)
        np.cumsum(cnt[:-1], out=beg[1:])
        child_rng[s:e, 0] = beg + pos
        child_rng[s:e, 1] = beg + cnt + pos
        child_idx.append(ch.astype(np.int32))
        pos += len(ch)
        H = int(h.max())
        counts = np.bincount(h, minlength=H + 1)
        info[b] = [s, e - s, H, sum(len(x) for x in level_ptr)]
        level_ptr.append(np.concatenate([[0], np.cumsum(counts)]).astype(np.int32))
    return {
        "block_info": info,
        "level_ptr": np.concatenate(level_ptr).astype(np.int32),
        "order": order,
        "parent": par_rel,
        "child_rng": child_rng,
        "child_idx": np.concatenate(child_idx).astype(np.int32) if child_idx else np.zeros(1, np.int32),
        "max_block": int(max(e - s for s, e in blocks)),
        "num_blocks": len(blocks),
        "num_nodes": n,
        "parent_glob": np.where(parent >= 0, parent, -1).astype(np.int32),
        "height": height,
        "depth": depth,
    }


def segments(parent: np.ndarray):
    """Kernel order of a forest (parents before children): segments are maximal chains
    of single-child nodes (top first); a segment's level is 0 if its bottom node has no
    children, else 1 + the largest level of its child segments. Returns the kernel order
    (kernel index -> input index), the segments `(beg, end)` in kernel indices grouped by
    level (level pointer; longest first within a level, so that the threads of a
    wavefront walk chains of similar length), and the child-segment tops of every
    segment (native: `jxhost_segments`)."""
    from . import _lib

    parent = np.ascontiguousarray(parent, dtype=np.int64)
    n = len(parent)
    korder = np.empty(max(n, 1), dtype=np.int64)
    seg = np.empty((max(n, 1), 4), dtype=np.int32)
    lptr = np.empty(n + 2, dtype=np.int32)
    seg_child = np.empty(max(n, 1), dtype=np.int32)
    sizes = np.zeros(3, dtype=np.int64)
    code = _lib.host().jxhost_segments(n, parent.ctypes.data, korder.ctypes.data, seg.ctypes.data,
                                       lptr.ctypes.data, seg_child.ctypes.data, sizes.ctypes.data)
    _lib.host_check(code, "the segment order")
    nseg, L, nsc = (int(x) for x in sizes)
    korder = korder[:n]
    kidx = np.empty(n, dtype=np.int64)
    kidx[korder] = np.arange(n)
    return {
        "korder": korder,
        "kidx": kidx,
        "seg": seg[:nseg].copy(),
        "seg_level_ptr": lptr[:L + 2].copy(),
        "seg_child": seg_child[:nsc].copy() if nsc else np.zeros(1, np.int32),
        "L": L,
        "nseg": nseg,
    }


@dataclass
class FusedTopology:
    """One topology of the fused kernels, in solve order (node k = solve index k).

    Integer tables follow the column layout of `jxhip.h`; float tables are built per
    dtype by the caller (they depend on parameters)."""

    n: int
    H: int
    order: np.ndarray          # (n,) nodes sorted by height
    level_ptr: np.ndarray      # (H+2,)
    parent: np.ndarray         # (n,) -1 root
    child_rng: np.ndarray      # (n, 2) topology-relative into child_idx
    child_idx: np.ndarray      # (n_children,)
    height: np.ndarray
    extra: Dict = field(default_factory=dict)


def fused_topology(parent_solve: np.ndarray) -> FusedTopology:
    """Schedule of one topology (a forest in solve order) for the fused kernels."""
    sched = forest_schedule(parent_solve, max_block=1 << 30)
    if sched is None or sched["num_blocks"] != 1:
        raise ValueError("the solve order must be a forest with parents before children")
    return FusedTopology(
        n=int(sched["num_nodes"]),
        H=int(sched["block_info"][0, 2]),
        order=sched["order"].astype(np.int32),
        level_ptr=sched["level_ptr"].astype(np.int32),
        parent=np.asarray(parent_solve, dtype=np.int32),
        child_rng=sched["child_rng"].astype(np.int32),
        child_idx=sched["child_idx"].astype(np.int32),
        height=sched["height"],
    )
