"""SWC morphology ingest: `read_swc` (jaxley-pytorch io/swc.py:11-68, graph backend).

The traced points of an SWC file become a `Cell` of `ncomp` compartments per branch
with per-compartment radius, length, membrane area, volume and input / output
resistive loads computed from the traced frusta, and the groups `soma`, `axon`,
`basal`, `apical` (`customN` for other type ids). The steps, each restated from the
reference in numpy and plain dicts (the reference drives networkx and pandas):

1. `_SwcGraph` - one node per SWC row, an edge parent -> child of Euclidean length
   (min 1e-5; io/graph.py:97-188).
2. `_trace_branches` - a single-point soma becomes a 2 r long two-point section; at a
   somatic branchpoint, non-somatic neurites start at their first own point; a
   depth-first walk from the lowest-numbered tip (neighbours in increasing order) cuts
   branches at tips, branchings and type changes; branches longer than
   `max_branch_len` are cut at multiples of it; optionally at tracing interruptions
   (io/graph.py:423-711).
3. `_compartments` - each branch's traced points are split into `ncomp` pieces of
   equal path length (new points interpolated), and every piece gives radius (length
   weighted mean), area (frusta), volume and the resistive loads of its two halves
   (utils/morph_attributes.py:15-297); branchpoints join branch ends
   (io/graph.py:191-420, 714-740).
4. `_solve_order` - a depth-first walk of the compartment graph from its
   lowest-numbered tip numbers compartments and branches (the module's indices) and
   gives each branch its parent (io/graph.py:747-832, 847-871, 979-1074).

Host-side integer and geometry work, done once per morphology.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence

import numpy as np

GROUP_NAMES = {0: "undefined", 1: "soma", 2: "axon", 3: "basal", 4: "apical"}


# ------------------------------------------------------------------------------------
# Frustum geometry (utils/morph_attributes.py)
# ------------------------------------------------------------------------------------


def split_xyzr_into_equal_length_segments(xyzr: np.ndarray, ncomp: int) -> List[np.ndarray]:
    """`ncomp` pieces of equal path length; split points are linear interpolations
    and end one piece and start the next (morph_attributes.py:81-145)."""
    if len(xyzr) == 1:
        return [xyzr] * ncomp
    d = np.linalg.norm(np.diff(xyzr[:, :3], axis=0), axis=1)
    cum = np.concatenate([[0.0], np.cumsum(d)])
    targets = np.linspace(0.0, cum[-1], ncomp + 1)
    seg = np.clip(np.searchsorted(cum, targets, side="right") - 1, 0, len(xyzr) - 2)
    frac = (targets - cum[seg]) / np.where(d < 1e-14, 1e-14, d)[seg]
    cuts = xyzr[seg] + frac[:, None] * (xyzr[seg + 1] - xyzr[seg])
    out = []
    for i in range(1, ncomp + 1):
        inner = xyzr[(cum > targets[i - 1]) & (cum < targets[i])]
        out.append(np.vstack([cuts[i - 1][None], inner, cuts[i][None]]))
    return out


def _frusta(xyzr: np.ndarray):
    return np.linalg.norm(np.diff(xyzr[:, :3], axis=0), axis=1), xyzr[:, 3]


def swc_radius(lengths, radii) -> float:
    """Mean radius, each point weighted by the lengths of its adjacent segments."""
    w = np.zeros(len(lengths) + 1)
    w[1:] += lengths
    w[:-1] += lengths
    return float(np.sum(radii * (w / np.sum(w))))


def swc_area(lengths, radii) -> float:
    """Lateral area of the truncated cones between consecutive points."""
    r0, r1 = radii[:-1], radii[1:]
    return float(np.sum(np.pi * (r0 + r1) * np.sqrt((r1 - r0) ** 2 + lengths**2)))


def swc_volume(lengths, radii) -> float:
    r0, r1 = radii[:-1], radii[1:]
    return float(np.sum(np.pi / 3 * lengths * (r0**2 + r0 * r1 + r1**2)))


def swc_resistive_load(lengths, radii) -> float:
    """(1 / pi) * integral of dl / r^2 over the truncated cones (a cylinder where the
    radii agree within 1e-8)."""
    r0, r1 = np.asarray(radii[:-1], float), np.asarray(radii[1:], float)
    dr = r1 - r0
    const = np.isclose(dr, 0)
    out = np.empty_like(np.asarray(lengths, float))
    out[const] = lengths[const] / r0[const] ** 2
    v = ~const
    out[v] = lengths[v] / dr[v] * (1 / r0[v] - 1 / r1[v])
    return float(np.sum(out) / np.pi)


def morph_attrs_from_xyzr(xyzr: np.ndarray, min_radius: Optional[float], ncomp: int):
    """(radius, area, volume, load of the first half, load of the second half) of one
    compartment; a single point is a sphere shared by the branch's `ncomp`
    compartments (morph_attributes.py:15-78)."""
    radii = xyzr[:, 3]
    if len(xyzr) > 1:
        lengths, _ = _frusta(xyzr)
        radius = swc_radius(lengths, radii)
        area = swc_area(lengths, radii)
        volume = swc_volume(lengths, radii)
        loads = [swc_resistive_load(*_frusta(h)) for h in split_xyzr_into_equal_length_segments(xyzr, 2)]
    else:
        radius = float(radii.mean())
        area = 4 * np.pi * radii[0] ** 2 / ncomp
        volume = 4 / 3 * np.pi * radii[0] ** 3 / ncomp
        loads = [radii[0] / ncomp / radii[0] ** 2 / np.pi] * 2
    if min_radius is None:
        assert radius > 0.0, "Radius 0.0 in SWC file. Set `read_swc(..., min_radius=...)`."
    elif radius < min_radius or np.isnan(radius):
        radius = min_radius
    return radius, area, volume, loads[0], loads[1]


# ------------------------------------------------------------------------------------
# SWC graph and branch tracing (io/graph.py)
# ------------------------------------------------------------------------------------


class _SwcGraph:
    """Nodes in insertion order with attributes; directed edges parent -> child with
    their length; `undirected()` gives neighbour sets and lengths both ways."""

    def __init__(self):
        self.nodes: Dict[int, Dict] = {}
        self.edges: Dict[tuple, float] = {}

    def add_edge(self, i, j, l):
        self.edges[(i, j)] = l

    def has_edge(self, i, j) -> bool:
        return (i, j) in self.edges

    def undirected(self):
        adj = {n: set() for n in self.nodes}
        length = {}
        for (i, j), l in self.edges.items():
            adj[i].add(j)
            adj[j].add(i)
            length[(i, j)] = length[(j, i)] = l
        return adj, length


def _read(fname: str) -> _SwcGraph:
    rows = np.loadtxt(fname)
    rows = rows.reshape(-1, 7)
    g = _SwcGraph()
    for i, typ, x, y, z, r, p in rows:
        g.nodes[int(i)] = {"id": int(typ), "x": x, "y": y, "z": z, "r": r, "p": int(p)}
    for i, p in rows[:, [0, 6]]:
        if p != -1:
            a, b = g.nodes[int(p)], g.nodes[int(i)]
            d = np.sqrt((a["x"] - b["x"]) ** 2 + (a["y"] - b["y"]) ** 2 + (a["z"] - b["z"]) ** 2)
            g.add_edge(int(p), int(i), d if d >= 1e-5 else 1e-5)
    return g


def _dfs_edges(adj: Dict[int, set], root):
    """Depth-first tree edges from `root`, neighbours in increasing order (networkx
    `dfs_edges(G, root, sort_neighbors=sorted)`)."""
    seen = {root}
    stack = [(root, iter(sorted(adj[root])))]
    while stack:
        parent, children = stack[-1]
        for child in children:
            if child not in seen:
                seen.add(child)
                yield parent, child
                stack.append((child, iter(sorted(adj[child]))))
                break
        else:
            stack.pop()


def _first_tip(adj) -> int:
    return next(n for n in sorted(adj) if len(adj[n]) < 2)


def _trace_branches(g: _SwcGraph, max_len, ignore_interruptions: bool, relevant_type_ids):
    """Branches as arrays of `(swc_parent, swc_node, length)` rows, their type ids,
    the set of points that are not part of the neurite that starts there, node types
    ('branchpoint' or not) and the (possibly rewritten) graph (io/graph.py:423-589)."""
    ids = lambda n: g.nodes[n]["id"]
    same = lambda i, j: (ids(i) not in relevant_type_ids and ids(j) not in relevant_type_ids) or ids(i) == ids(j)
    ignore = []
    soma = [n for n in g.nodes if g.nodes[n]["id"] == 1]
    if len(soma) == 1:
        s = soma[0]
        for e in list(g.edges):
            if s in e:
                g.edges[e] = 0.0
        # The point becomes a section of length 2 r from a copy of the first node.
        g.nodes[1]["p"] = 0
        for n in g.nodes:
            g.nodes[n]["p"] += 1
        copy = dict(g.nodes[1], p=-1)
        g2 = _SwcGraph()
        for n, a in g.nodes.items():
            g2.nodes[n + 1] = a
        g2.nodes[1] = copy
        for (i, j), l in g.edges.items():
            g2.edges[(i + 1, j + 1)] = l
        g2.edges[(1, s + 1)] = 2 * g.nodes[s]["r"]
        g = g2
        ignore += [1, s + 1]
    adj, length = g.undirected()
    # Somatic branchpoints (the indices are those before any relabelling).
    for node in soma:
        if node not in adj:
            continue
        if sum(1 for m in adj[node] if g.nodes[m]["id"] == 1) > 1:
            for m in adj[node]:
                if g.nodes[m]["id"] != 1:
                    ignore.append(node)
                    length[(node, m)] = length[(m, node)] = 0.0
    branches, types, cur = [], [], []
    for i, j in _dfs_edges(adj, _first_tip(adj)):
        cur.append((i, j, length[(i, j)]))
        tid = ids(i)
        if len(adj[j]) < 2:
            if not same(i, j):
                branches += [cur[:-1], cur[-1:]]
                types += [tid, ids(j)]
            else:
                branches.append(cur)
                types.append(tid)
            cur = []
        elif len(adj[j]) > 2:
            branches.append(cur)
            types.append(tid)
            cur = []
        elif not same(i, j):
            if not g.has_edge(i, j) and g.has_edge(j, i):
                branches.append(cur)
                types.append(tid)
                cur = []
            else:
                branches.append(cur[:-1])
                types.append(tid)
                cur = [cur[-1]]
    keep = [k for k, b in enumerate(branches) if len(b) > 0]
    branches = [np.array(branches[k], dtype=float) for k in keep]
    types = [types[k] for k in keep]
    bp = set()
    if max_len:
        new_b, new_t = [], []
        for b, t in zip(branches, types):
            k = np.cumsum([length[(int(i), int(j))] for i, j, _ in b]) // max_len
            parts = [b[np.where(k == kk)[0]] for kk in np.unique(k)]
            new_b += parts
            new_t += [t] * len(parts)
            bp.update(int(p[0, 0]) for p in parts[1:])
        branches, types = new_b, new_t
    bp.update(int(b[0, 0]) for b in branches[1:])
    if not ignore_interruptions:
        cuts = [g.nodes[n]["p"] for n in g.nodes
                if g.nodes[n]["p"] > -1 and g.nodes[n]["p"] != n - 1
                and n not in bp and g.nodes[n]["p"] not in bp]
        for node in cuts:
            k = next(q for q, b in enumerate(branches) if node in b[:, :2])
            b = branches[k]
            at = np.where(b[:, :2] == node)[0][1]
            branches = branches[:k] + [b[:at], b[at:]] + branches[k + 1:]
            types = types[:k] + [types[k], types[k]] + types[k + 1:]
        bp.update(cuts)
    return g, branches, types, set(ignore), bp


def _compartments(g: _SwcGraph, branches, types, ignore, bp, ncomp: int, min_radius):
    """Compartment attributes in branch order and the compartment graph: nodes are
    compartments 0..N-1 and branchpoints labelled N, N+1, ... in the order of the
    branch that ends at them; directed edges follow the tracing (io/graph.py:191-420,
    714-740)."""
    comp = {k: [] for k in ("radius", "length", "area", "volume", "rl_in", "rl_out", "id", "x", "y", "z")}
    xyzr_of = []
    edges = []                       # (u, v), u/v = ("c", index) or ("b", swc node)
    ends = {}                        # swc branchpoint -> branch index that ends there
    for b_idx, b in enumerate(branches):
        path = np.cumsum(np.concatenate([[0.0], b[:, 2]]))
        flat = b[:, :2].reshape(-1)
        _, first = np.unique(flat, return_index=True)
        nodes = flat[np.sort(first)].astype(np.int64)
        a = [g.nodes[int(n)] for n in nodes]
        idv = np.array([x["id"] for x in a], float)
        xyzr = np.array([[x["x"], x["y"], x["z"], x["r"]] for x in a], float)
        if idv[0] != idv[1]:
            idv[0] = idv[1]
            xyzr[0, 3] = xyzr[1, 3]
        data = np.column_stack([xyzr, path])
        if int(b[0, 0]) in ignore:
            xyzr = xyzr[1:]
        pieces = split_xyzr_into_equal_length_segments(xyzr, ncomp)
        attrs = np.asarray([morph_attrs_from_xyzr(p, min_radius, ncomp) for p in pieces])
        blen = path.max()
        if blen < 1e-8:
            blen = 0.1
        clen = blen / ncomp
        locs = np.linspace(clen / 2, blen - clen / 2, ncomp)
        for q, key in enumerate(("x", "y", "z")):
            comp[key] += list(np.interp(locs, path, data[:, q]))
        comp["radius"] += list(attrs[:, 0])
        comp["area"] += list(attrs[:, 1])
        comp["volume"] += list(attrs[:, 2])
        comp["rl_in"] += list(attrs[:, 3])
        comp["rl_out"] += list(attrs[:, 4])
        comp["length"] += [clen] * ncomp
        comp["id"] += [types[b_idx]] * ncomp
        xyzr_of += pieces
        c0 = b_idx * ncomp
        edges += [(("c", c0 + q), ("c", c0 + q + 1)) for q in range(ncomp - 1)]
        pre, post = int(b[0, 0]), int(b[-1, 1])
        if pre in bp:
            edges.append((("b", pre), ("c", c0)))
        if post in bp:
            edges.append((("c", c0 + ncomp - 1), ("b", post)))
            ends[post] = b_idx
    n = len(comp["radius"])
    # Branchpoints that no branch ends at have no predecessor and no label.
    order = sorted(ends, key=lambda s: ends[s])
    label = {("b", s): n + k for k, s in enumerate(order)}
    label.update({("c", c): c for c in range(n)})
    cedges = [(label[u], label[v]) for u, v in edges if u in label and v in label]
    for key in comp:
        comp[key] = np.asarray(comp[key])
    comp["xyzr"] = xyzr_of
    return comp, n, cedges


def _solve_order(n_comp: int, cedges):
    """Depth-first numbering of the compartment graph from its lowest-numbered tip:
    new compartment index per old one, new branch index per old compartment, the
    parent branch of every new branch and which compartments were walked against the
    tracing direction (io/graph.py:747-832, 835-871, 979-1074)."""
    nodes = set(range(n_comp))
    for u, v in cedges:
        nodes.update((u, v))
    adj = {x: set() for x in nodes}
    for u, v in cedges:
        adj[u].add(v)
        adj[v].add(u)
    # Branchpoints left at tips are dropped (io/graph.py:835-844).
    for x in [x for x in nodes if x >= n_comp and len(adj[x]) <= 1]:
        for y in adj.pop(x):
            adj[y].discard(x)
    directed = set(cedges)
    root = _first_tip(adj)
    new_comp = {root: 0}
    branch = {root: 0}
    flipped = set()
    pred = {}
    ci, bi = 1, 0
    for i, j in _dfs_edges(adj, root):
        pred[j] = i
        branch[j] = bi
        if j < n_comp:
            new_comp[j] = ci
        if len(adj[j]) < 2 or j >= n_comp:
            bi += 1
        if j < n_comp:
            ci += 1
        if (i, j) not in directed and (j, i) in directed:
            flipped.update((i, j))
    if len(new_comp) != n_comp:
        raise ValueError("the SWC morphology is not one connected tree")
    # Branch edges through branchpoints and direct compartment links between branches.
    parent_of = {}
    for j, i in pred.items():
        if j >= n_comp:
            continue
        u = i
        if u >= n_comp:
            u = pred[u]
        if branch[u] != branch[j]:
            parent_of[branch[j]] = branch[u]
    n_branch = max(branch[c] for c in range(n_comp)) + 1
    parents = [-1] + [parent_of[b] for b in range(1, n_branch)]
    return new_comp, branch, parents, flipped


def read_swc(fname: str, ncomp: int, max_branch_len: Optional[float] = None,
             min_radius: Optional[float] = None, assign_groups: bool = True, backend: str = "graph",
             ignore_swc_tracing_interruptions: bool = True,
             relevant_type_ids: Optional[List[int]] = None):
    """Read an SWC file into a `Cell` (io/swc.py:11-68). Compartments are cylinders in
    length only: their area, volume and resistive loads come from the traced frusta,
    and `set("radius" | "length", ...)` turns the selected ones into cylinders."""
    from .modules import Branch, Cell

    if backend != "graph":
        raise ValueError(f"Unknown backend: {backend}. Use either `custom` or `graph`.")
    relevant_type_ids = [1, 2, 3, 4] if relevant_type_ids is None else list(relevant_type_ids)
    g = _read(fname)
    g, branches, types, ignore, bp = _trace_branches(g, max_branch_len, ignore_swc_tracing_interruptions,
                                                     relevant_type_ids)
    comp, n, cedges = _compartments(g, branches, types, ignore, bp, ncomp, min_radius)
    new_comp, branch, parents, flipped = _solve_order(n, cedges)
    order = np.empty(n, dtype=np.int64)          # old compartment at each new index
    for old, new in new_comp.items():
        order[new] = old
    nb = len(parents)
    counts = np.bincount([branch[c] for c in range(n)], minlength=nb)
    if not np.all(counts == ncomp):
        raise ValueError("branches of unequal compartment counts are not supported")
    cell = Cell([Branch(ncomp=ncomp)] * nb, parents)
    cell._set_morphology(
        radius=comp["radius"][order], length=comp["length"][order], area=comp["area"][order],
        volume=comp["volume"][order], resistive_load_in=comp["rl_in"][order],
        resistive_load_out=comp["rl_out"][order],
    )
    for key in ("x", "y", "z"):
        cell._node_data[key] = comp[key][order].astype(float)
    xyzr = [comp["xyzr"][c][::-1] if c in flipped else comp["xyzr"][c] for c in order]
    cell.xyzr = [np.concatenate(xyzr[b * ncomp:(b + 1) * ncomp]) for b in range(nb)]
    if assign_groups:
        ids = comp["id"][order].astype(np.int64)
        for t in np.unique(ids):
            cell.select(nodes=np.nonzero(ids == t)[0]).add_to_group(GROUP_NAMES.get(int(t), f"custom{int(t)}"))
    return cell
