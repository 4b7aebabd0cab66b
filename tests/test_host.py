"""Host-side logic on the CPU: product schedules vs the oracle, the C ABI exports, and
the fused kernel's descriptor tables executed by a CPU model of the kernel."""
import ctypes
import re
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

import jaxhip as jx
from jaxhip import _lib
from jaxhip import schedule as S
from oracle import jx_oracle as O

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.parametrize("seed", range(20))
def test_schedule_matches_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    nb = int(rng.integers(1, 80))
    parents = [-1] + [int(rng.integers(0, k)) for k in range(1, nb)]
    ncomp = rng.integers(1, 6, size=nb)
    for allowed in (1, 4, 32):
        a = S.dhs_indexer(S.cell_tables(parents, ncomp), allowed)
        b = O.dhs_indexer_cell(O.cell_tables(parents, ncomp), allowed)
        for k in ("map_to_solve_order", "inv_map_to_solve_order", "map_to_solve_order_lower",
                  "map_to_solve_order_upper", "parent_lookup", "node_order"):
            assert np.array_equal(a[k], b[k]), k
        assert all(np.array_equal(x, y) for x, y in zip(a["node_order_grouped"], b["node_order_grouped"]))


def test_network_schedule_matches_oracle():
    specs = [([-1, 0, 0], [2, 2, 2]), ([-1], [1]), ([-1, 0, 0, 1, 1], [4] * 5)]
    cell_tabs = [S.cell_tables(p, n) for p, n in specs]
    inds = [S.dhs_indexer(t) for t in cell_tabs]
    ta, ia = S.network_indexer(cell_tabs, inds, True)
    otabs = [O.cell_tables(p, n) for p, n in specs]
    tb, ib = O.network_tables_and_indexer(otabs, [O.dhs_indexer_cell(t) for t in otabs], "gpu")
    for k in ("source", "sink", "type", "ordered"):
        assert np.array_equal(ta[k], tb[k])
    for k in ("map_to_solve_order", "inv_map_to_solve_order", "map_to_solve_order_lower",
              "map_to_solve_order_upper", "parent_lookup"):
        assert np.array_equal(ia[k], ib[k]), k


def test_forest_schedule_invariants():
    rng = np.random.default_rng(3)
    parents = [-1] + [int(rng.integers(0, k)) for k in range(1, 50)]
    ind = S.dhs_indexer(S.cell_tables(parents, [4] * 50))
    pl = ind["parent_lookup"]
    sch = S.forest_schedule(pl, 1 << 20)
    h = sch["height"]
    for c, p in enumerate(pl):
        if p >= 0:
            assert h[p] >= h[c] + 1 and p < c
    # Every node appears exactly once in the height order; levels are contiguous.
    assert sorted(sch["order"].tolist()) == list(range(len(pl)))
    lp = sch["level_ptr"]
    for lev in range(len(lp) - 1):
        assert all(h[k] == lev for k in sch["order"][lp[lev]:lp[lev + 1]])


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "jxhip.h").read_text()
    declared = set(re.findall(r"^(?:int|const char\*)\s+(jxhip_\w+)\(", header, re.M))
    assert declared == set(_lib.exported_symbols())
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    assert _lib.fn("jxhip_version")() == 100


def test_host_library_exports_every_declared_symbol():
    """libjxhost.so (g++ only, no GPU runtime) exports what include/jxhip_host.h declares."""
    header = (ROOT / "include" / "jxhip_host.h").read_text()
    declared = set(re.findall(r"^int\s+(jxhost_\w+)\(", header, re.M))
    assert declared == set(_lib._HOST_SIGS)
    lib = ctypes.CDLL(str(_lib.HOST_LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    # No GPU runtime among its dependencies (preparation workers load it).
    assert b"libamdhip64" not in _lib.HOST_LIB_PATH.read_bytes()


def _segments_reference(parent):
    """The segment order written out directly (test side): chains of single-child nodes
    from every top, levels from the leaves, sorted by (level, -length, top)."""
    n = len(parent)
    kids = [[] for _ in range(n)]
    for c in range(n):
        if parent[c] >= 0:
            kids[parent[c]].append(c)
    tops = [t for t in range(n) if parent[t] < 0 or len(kids[parent[t]]) != 1]
    chains = []
    for t in tops:
        ch = [t]
        while len(kids[ch[-1]]) == 1:
            ch.append(kids[ch[-1]][0])
        chains.append(ch)
    of = {c: q for q, ch in enumerate(chains) for c in ch}

    def level(q):
        b = chains[q][-1]
        return 0 if len(kids[b]) < 2 else 1 + max(level(of[c]) for c in kids[b])
    lv = [level(q) for q in range(len(chains))]
    order = sorted(range(len(chains)), key=lambda q: (lv[q], -len(chains[q]), chains[q][0]))
    korder = [c for q in order for c in chains[q]]
    kidx = {c: i for i, c in enumerate(korder)}
    seg, child = [], []
    for q in order:
        b = chains[q][-1]
        tk = [kidx[c] for c in kids[b]] if len(kids[b]) >= 2 else []
        seg.append((kidx[chains[q][0]], kidx[chains[q][0]] + len(chains[q]), len(child), len(child) + len(tk)))
        child += tk
    L = max(lv) if lv else 0
    lptr = np.concatenate([[0], np.cumsum(np.bincount(np.array(lv, dtype=np.int64), minlength=L + 1))])
    return np.array(korder), np.array(seg).reshape(-1, 4), lptr, np.array(child if child else [0]), L


@pytest.mark.parametrize("seed", range(12))
def test_native_segments_match_reference(seed):
    """`jxhost_segments` (native) against the order written out in Python, on random
    forests with chains, branchings and several roots."""
    rng = np.random.default_rng(300 + seed)
    n = int(rng.integers(1, 400))
    parent = np.full(n, -1, dtype=np.int64)
    for i in range(1, n):
        if rng.random() > 0.03:
            # Mostly chains (the parent is the previous node) with random branchings.
            parent[i] = i - 1 if rng.random() < 0.7 else int(rng.integers(0, i))
    sg = S.segments(parent)
    korder, seg, lptr, child, L = _segments_reference(parent)
    assert np.array_equal(sg["korder"], korder)
    assert np.array_equal(sg["seg"], seg)
    assert np.array_equal(sg["seg_level_ptr"], lptr)
    assert np.array_equal(sg["seg_child"], child)
    assert sg["L"] == L and sg["nseg"] == len(seg)
    # Invariants: a permutation; every segment a parent chain; child tops belong to
    # segments of a lower level.
    assert sorted(sg["korder"].tolist()) == list(range(n))
    for b, e, cb, ce in sg["seg"]:
        nodes = sg["korder"][b:e]
        assert all(parent[nodes[i + 1]] == nodes[i] for i in range(len(nodes) - 1))
        for t in sg["seg_child"][cb:ce]:
            assert parent[sg["korder"][t]] == nodes[-1]


def test_native_schedule_rejects_bad_input():
    with pytest.raises(ValueError, match="forest"):
        S.segments(np.array([-1, 2, 0]))
    tab = S.cell_tables([-1, 0], [2, 2])
    bad = dict(tab)
    keep = ~((tab["type"] == 0) & (tab["source"] == 0))
    bad = {**tab, "source": tab["source"][keep], "sink": tab["sink"][keep], "type": tab["type"][keep],
           "ordered": tab["ordered"][keep]}
    with pytest.raises(ValueError, match="connected"):
        S.dhs_indexer(bad)


def test_struct_layouts_match_header():
    """ctypes mirrors of the argument blocks have the C field order of jxhip.h."""
    header = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "jxhip.h").read_text(), flags=re.S)
    body = header[header.index("typedef struct jxhip_fused_args"):header.index("} jxhip_fused_args;")]
    names = re.findall(r"\b(\w+)(?:\[\w+\])?;", body)
    assert names == [f[0] for f in _lib.FusedArgs._fields_]
    body = header[header.index("typedef struct jxhip_adjoint_args"):header.index("} jxhip_adjoint_args;")]
    names = re.findall(r"\b(\w+)(?:\[\w+\])?;", body)
    assert names == [f[0] for f in _lib.AdjointArgs._fields_]


def test_integrate_refuses_cpu_tensors():
    cell = jx.Cell(jx.Branch(jx.Compartment(), 2), [-1, 0, 0])
    cell.insert(jx.HH())
    cell.record()
    ps = cell.data_set("HH_gNa", torch.tensor([[0.1]], dtype=torch.float64), None)
    with pytest.raises(RuntimeError, match="GPU"):
        jx.integrate(cell, param_state=ps, t_max=1.0)


@pytest.fixture
def cpu_model(monkeypatch):
    """Run `integrate`'s host path on the CPU, the kernel replaced by its CPU model."""
    from kernel_model import run_forward_model
    import jaxhip.fused as F

    monkeypatch.setattr(sys.modules["jaxhip.integrate"], "_require_gpu", lambda d: None)
    monkeypatch.setattr(F.ForwardLaunch, "launch", lambda self: run_forward_model(self))
    # No launch-shape measurement on the CPU (the heuristic shape is used).
    monkeypatch.setenv("JAXHIP_AUTOTUNE", "0")


def test_host_tables_c0(cpu_model):
    from models import DT, build_c0, c0_stimulus

    cell, om = build_c0()
    cur = c0_stimulus(80)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64))
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=list(range(60)))
    assert np.abs(v.numpy() - ref.numpy()).max() < 1e-9


def test_host_tables_batched_tree(cpu_model):
    from models import DT, build_tree, c1_params, c1_stimulus, data_set_params, oracle_overrides

    B, T = 3, 60
    cell, om = build_tree(12, seed=2)
    cell.branch(0).comp(0).record()
    cell.branch(3).comp(2).record()
    prm = c1_params(B, seed=2)
    ps = data_set_params(cell, prm, torch.device("cpu"), torch.float64)
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64))
    v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT)
    rc = int(cell._cumsum_ncomp[3] + 2)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0, rc],
                     overrides=oracle_overrides(om, prm))
    assert v.shape == (B, 2, T + 1)
    assert np.abs(v.numpy() - ref.numpy()).max() < 1e-9


def test_serve_arena_allocator():
    from jaxhip.serve import _Arena

    a = _Arena(100)
    x = a.alloc(10)
    y = a.alloc(20, align=8)
    z = a.alloc(70)
    assert (x, y) == (0, 16) and z is None
    assert a.alloc(6) == 10          # the alignment gap is reused
    a.free(y, 20)
    a.free(x, 10)
    assert a.alloc(30) == 0 or a.free_list[0][0] == 0
    rng = np.random.default_rng(0)
    b, held = _Arena(1000), []
    for _ in range(2000):
        if held and rng.random() < 0.5:
            off, sz = held.pop(int(rng.integers(len(held))))
            b.free(off, sz)
        else:
            sz = int(rng.integers(1, 60))
            off = b.alloc(sz, align=int(rng.choice([1, 4])))
            if off is not None:
                held.append((off, sz))
        spans = sorted(held)
        for (o1, s1), (o2, _) in zip(spans, spans[1:]):
            assert o1 + s1 <= o2                    # no overlap
        assert all(0 <= o and o + s <= 1000 for o, s in held)
    for off, sz in held:
        b.free(off, sz)
    assert b.free_list == [(0, 1000)] and b.used == 0



def test_serve_preparation_in_worker_processes():
    """A served request prepared in a spawned worker process (the server's
    `prep_processes`: numpy arrays across the boundary) equals the same preparation in
    this process, tensor by tensor."""
    import concurrent.futures as cf
    import multiprocessing

    from jaxhip.serve import _prepare_in_worker, _prepare_payload, _to_numpy
    from models import c2_request, c2_specs

    cpu = torch.device("cpu")
    reqs = [c2_request(sp, cpu, torch.float64) for sp in c2_specs(3, seed=7)]
    args = [(r.module, r.params, r.param_state, r.data_stimuli, r.t_max, r.all_states, 0.025,
             torch.float64, "mixed", 4001) for r in reqs]
    local = [_prepare_payload(*a) for a in args]
    with cf.ProcessPoolExecutor(max_workers=2, mp_context=multiprocessing.get_context("spawn")) as pool:
        remote = list(pool.map(_prepare_in_worker, [_to_numpy(a) for a in args]))
    for p, q in zip(local, remote):
        assert (p.n, p.L, p.T, p.n_rec, p.sched) == (q.n, q.L, q.T, q.n_rec, q.sched)
        for name in ("node_i", "seg", "seg_level_ptr", "seg_child", "child_idx", "rec_node", "stim_col",
                     "node_f", "node_p", "v0", "stim", "stim_scale"):
            x, y = getattr(p, name), getattr(q, name)
            assert isinstance(x, np.ndarray) and x.dtype == y.dtype and np.array_equal(x, y), name
        for g, h in zip(p.gates0, q.gates0):
            assert np.array_equal(g, h)


def test_tree_solve_operator_fake_rule():
    """`jaxhip::tree_solve` propagates shapes and dtypes without running (what
    `torch.compile` traces with)."""
    from torch._subclasses.fake_tensor import FakeTensorMode

    from jaxhip.kernels.tree_solve import tree_solve_op

    with FakeTensorMode():
        d = torch.empty(3, 10, dtype=torch.float64)
        s = torch.empty(1, 10, dtype=torch.float64)
        x = tree_solve_op(d, d, d, s, torch.empty(10, dtype=torch.int32), 0)
        assert x.shape == (3, 10) and x.dtype == torch.float64


def test_synapse_step_operators_fake_rules():
    """`jaxhip::synapse_step` and `jaxhip::synapse_step_backward` propagate shapes and
    dtypes without running, and the forward's registered backward traces through the
    backward operator (what `torch.compile` records)."""
    from torch._subclasses.fake_tensor import FakeTensorMode

    from jaxhip.kernels.synapse_step import synapse_step_backward_op, synapse_step_op

    B, n, k, T = 2, 12, 3, 4
    E = k * T
    with FakeTensorMode():
        f = lambda *shape: torch.empty(shape, dtype=torch.float64)
        i = lambda *shape: torch.empty(shape, dtype=torch.int32)
        params = [f(E).requires_grad_() for _ in range(6)]
        v = f(B, n).requires_grad_()
        outs = synapse_step_op(v, f(B, E), f(E), params, i(E), i(T), i(5), i(5, 3), k, 0, False,
                               0.025, 1e-3)
        assert [tuple(o.shape) for o in outs] == [(B, E), (B, E), (B, T), (B, T)]
        assert all(o.dtype == torch.float64 for o in outs)
        g = synapse_step_backward_op(v.detach(), f(B, E), f(E), params, i(E), i(T), i(5), i(5, 3), k, 0,
                                     False, 0.025, 1e-3, None, f(B, E), f(B, T), f(B, T))
        assert [tuple(x.shape) for x in g] == [(B, n), (B, E), (6, E)]
        gv, *gp = torch.autograd.grad(outs[2].sum() + outs[1].sum(), [v] + params)
        assert gv.shape == (B, n) and all(x.shape == (E,) for x in gp)


def test_server_launch_shapes_respect_lds():
    """The server's launch shapes for a float64 tree of ~5.5k nodes: no register-resident
    shape (its coefficients and system exceed the LDS), the streaming kernel with the
    coupling coefficients read from L2; a tree whose streaming system exceeds the LDS
    has no shape (and is rejected at acceptance)."""
    from types import SimpleNamespace

    from jaxhip.fused import lds_bytes
    from jaxhip.serve import SimServer

    fake = SimpleNamespace(gsize=8, lsize=8, thread_order=(256, 512, 1024))
    fake._candidates = lambda n, s: SimServer._candidates(fake, n, s)
    small = SimServer._candidates(fake, 900, 4096)
    assert small and not small[0].stream
    n, sched = 5500, 16 * 1400
    assert lds_bytes(n, 0, 1, 8, sched) > 160 * 1024
    cands = SimServer._candidates(fake, n, sched)
    assert [c.stream for c in cands] == [True] and cands[0].gcoef
    assert lds_bytes(n, 0, 1, 8, sched, True) <= 160 * 1024
    assert SimServer._config(fake, n, sched) == cands[0]
    assert SimServer._fits(fake, small[0], n, sched) is False
    big = 12000
    assert SimServer._candidates(fake, big, 16 * 3000) == []


def _irregular_network(synapse: str):
    """Six ball-and-stick cells (the network workload's cells) with irregular in-degree:
    cell c receives 1 + (c mod 4) synapses from seeded presynaptic somata, spread over
    two dendritic compartments for some cells; the oracle model of the same network."""
    from jaxhip.workloads import C4_LENGTH, C4_RADIUS

    rng = np.random.default_rng(3)
    n_cells = 6
    cell = jx.Cell([jx.Branch(jx.Compartment(), 1), jx.Branch(jx.Compartment(), 8)], [-1, 0])
    cell.set("length", C4_LENGTH)
    cell.set("radius", C4_RADIUS)
    cell.branch(0).insert(jx.HH())
    cell.branch(1).insert(jx.Leak())
    net = jx.Network([cell] * n_cells)
    pre_comp, post_comp = [], []
    for c in range(n_cells):
        for j in range(1 + c % 4):
            pre_comp.append(9 * int(rng.integers(0, n_cells)))
            post_comp.append(9 * c + (3 if j % 2 == 0 or c < 3 else 6))
    pre_comp, post_comp = np.array(pre_comp), np.array(post_comp)
    syn = getattr(jx, synapse)()
    net.add_synapses(net._comp_node[pre_comp], net._comp_node[post_comp], syn)
    E = len(pre_comp)
    gs = rng.uniform(0.0, 2e-2, size=E)
    getattr(net, syn.name).set(f"{syn.name}_gS", gs)
    om = O.OracleModel([([-1, 0], [1, 8])] * n_cells)
    om.length = np.tile(C4_LENGTH, n_cells)
    om.radius = np.tile(C4_RADIUS, n_cells)
    om.hh_comps = np.arange(n_cells) * 9
    om.leak_comps = np.setdiff1d(np.arange(9 * n_cells), om.hh_comps)
    edge = net._edges[0]
    params = {name: edge["params"][f"{syn.name}_{name}"].copy() for name in syn.param_names}
    s0 = edge["states"][f"{syn.name}_s"].copy() if synapse == "KineticSynapse" else None
    om.synapses = [{"type": synapse, "pre": om.internal[pre_comp], "post": om.internal[post_comp],
                    "params": params, "s0": s0}]
    return net, om, pre_comp, post_comp


@pytest.mark.parametrize("parts", [1, 3])
def test_synapse_jagged_diagonal_order(parts):
    """The fused kernels' synapse lists: per part of the network (consecutive cells), the
    synapses onto the part's nodes with the targets by non-increasing in-degree, the j-th
    synapse of target t at ptr[j] + t, and each target's synapses in edge order;
    presynaptic nodes index the whole network, targets the part's nodes."""
    from jaxhip.fused import build_structure

    net, _, pre_comp, post_comp = _irregular_network("KineticSynapse")
    st = build_structure(net, np.array([0]), np.zeros(0, np.int64), torch.device("cpu"), parts)
    assert st.parts == parts and st.part_n.sum() == st.n
    syn = st.syn
    ptr_all, perm, tgt_all = syn["ptr"].numpy(), np.asarray(syn["perm"]), syn["tgt"].numpy()
    pre = syn["pre"].numpy()
    post_k = st.kidx_of_node[net._comp_node[post_comp]]
    pre_k = st.kidx_of_node[net._comp_node[pre_comp]]
    seen = []
    for q in range(parts):
        n_p, nb, n_tgt, tgb, spb, synb, n_syn = (int(x) for x in st.topo_row[q, [0, 2, 7, 8, 9, 10, 11]])
        ptr = ptr_all[spb:]
        deg = []
        for t in range(n_tgt):
            j, mine = 0, []
            while t < ptr[j + 1] - ptr[j]:
                mine.append(synb + ptr[j] + t)
                j += 1
            deg.append(j)
            node = nb + tgt_all[tgb + t]
            assert 0 <= tgt_all[tgb + t] < n_p
            assert [perm[m] for m in mine] == [e for e in range(len(post_k)) if post_k[e] == node]
            assert all(pre[m] == pre_k[perm[m]] for m in mine)
            seen += mine
        assert deg == sorted(deg, reverse=True) and sum(deg) == n_syn
    assert sorted(seen) == list(range(len(post_k)))
    assert sorted(perm.tolist()) == list(range(len(post_k)))


@pytest.mark.parametrize("synapse", ["KineticSynapse", "SigmoidRateSynapse"])
def test_host_tables_network(cpu_model, synapse):
    """A network with irregular in-degree through the host tables and the CPU model of
    the fused kernel (synapse phase in jagged-diagonal order) against the oracle."""
    from models import DT

    net, om, _, _ = _irregular_network(synapse)
    soma = np.arange(6) * 9
    for c in range(6):
        net.cell(c).branch(0).comp(0).record()
    T = 80
    cur = np.random.default_rng(5).uniform(size=(6, T)) < 0.2
    cur = cur * 1.5
    ds = None
    for c in range(6):
        ds = net.cell(c).branch(0).comp(0).data_stimulate(torch.tensor(cur[c], dtype=torch.float64), ds)
    v = jx.integrate(net, data_stimuli=ds, delta_t=DT)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=list(soma), rec_comps=list(soma))
    assert np.abs(v.numpy() - ref[0].numpy()).max() < 1e-9


def test_split_parts_are_never_empty():
    """Uneven cells (six one-compartment cells and one of 20 compartments) asked for seven
    parts: consecutive cells, no empty part, every node in exactly one."""
    from jaxhip.fused import build_structure

    morph = jx.Cell(jx.Branch(jx.Compartment(), 4), parents=[-1, 0, 0, 1, 1])
    net = jx.Network([jx.Cell()] * 6 + [morph])
    net.insert(jx.HH())
    st = build_structure(net, np.array([0]), np.zeros(0, np.int64), torch.device("cpu"), 7)
    assert 1 < st.parts <= 7 and (st.part_n > 0).all() and st.part_n.sum() == st.n
    assert np.array_equal(st.part_nb, np.concatenate([[0], np.cumsum(st.part_n)[:-1]]))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_host_tables_numpy_equal_torch(dtype):
    """The numpy forms of the axial conductances and the node table (unbatched CPU inputs,
    a served request's preparation) equal the tensor forms bit for bit."""
    from jaxhip import fused as F
    from jaxhip.geometry import axial_conductances
    from jaxhip.workloads import c2_specs, tree_cell

    sp = c2_specs(3, seed=5)[2]
    cell = tree_cell(sp["nbranch"], sp["seed"])[0]
    cell.branch(0).comp(0).record()
    rng = np.random.default_rng(0)
    nc = cell._n_comp
    geo = {k: torch.tensor(rng.uniform(0.5, 2.0, nc), dtype=dtype) for k in ("area", "rl_in", "rl_out")}
    ra = torch.tensor(rng.uniform(50, 5000, nc), dtype=dtype)
    cm = torch.tensor(rng.uniform(0.5, 2.0, nc), dtype=dtype)
    tables = cell._comp_edges
    g_np = axial_conductances(tables, cell._comp_node, cell._n_nodes, geo, ra, cm)
    g_t = axial_conductances(tables, cell._comp_node, cell._n_nodes, {k: v[None] for k, v in geo.items()},
                             ra[None], cm[None])[0]
    assert g_np.dtype == g_t.dtype and torch.equal(g_np, g_t)
    st = F.build_structure(cell, np.array([0]), np.array([0]), torch.device("cpu"))
    cm_node = torch.full((cell._n_nodes,), -1.0, dtype=dtype)
    cm_node[torch.as_tensor(cell._comp_node)] = cm
    a = F.node_float_table(st, g_np, cm_node, 0.025, dtype)
    b = F.node_float_table(st, g_np[None].expand(1, -1)[0].clone().requires_grad_(True), cm_node, 0.025, dtype)
    assert torch.equal(a, b.detach())


def test_serve_batched_preparation_results():
    """A pool task of several requests returns each one's payload or its own error."""
    from jaxhip.serve import _prepare_batch

    def fn(x):
        if x == 2:
            raise ValueError("bad request")
        return x * 10

    out = _prepare_batch(fn, [(1,), (2,), (3,)])
    assert [ok for ok, _ in out] == [True, False, True]
    assert out[0][1] == 10 and out[2][1] == 30 and isinstance(out[1][1], ValueError)


@pytest.mark.parametrize("nbranch,seed", [(250, 11), (40, 3)])
def test_symmetric_rows_scale_the_system_symmetric(nbranch, seed):
    """The row scales of the symmetric solve: W A is symmetric (its couplings are the
    table's s) and the scaled system has the solution of A (dense solves), on the
    1000-compartment random tree and a smaller one, branchpoints included."""
    from jaxhip import fused as F
    from jaxhip.geometry import axial_conductances
    from jaxhip.workloads import tree_cell

    cell = tree_cell(nbranch, seed)[0]
    cell.branch(0).comp(0).record()
    rng = np.random.default_rng(seed)
    nc = cell._n_comp
    geo = {k: torch.tensor(rng.uniform(0.5, 2.0, nc), dtype=torch.float64) for k in ("area", "rl_in", "rl_out")}
    ra = torch.tensor(rng.uniform(50, 5000, nc), dtype=torch.float64)
    cm = torch.tensor(rng.uniform(0.5, 2.0, nc), dtype=torch.float64)
    g = axial_conductances(cell._comp_edges, cell._comp_node, cell._n_nodes, geo, ra, cm)
    st = F.build_structure(cell, np.array([0]), np.array([0]), torch.device("cpu"))
    cm_node = torch.full((cell._n_nodes,), -1.0, dtype=torch.float64)
    cm_node[torch.as_tensor(cell._comp_node)] = cm
    nf = F.node_float_table(st, g, cm_node, 0.025, torch.float64).numpy()
    sym = F.symmetric_rows(st, torch.from_numpy(nf)).numpy()
    n, par = st.n, st.parent_k
    A = np.diag(nf[:, 2] + rng.uniform(0.1, 1.0, n) * nf[:, 3])
    ch = np.nonzero(par >= 0)[0]
    A[ch, par[ch]] = nf[ch, 0]
    A[par[ch], ch] = nf[ch, 1]
    w, s = sym[:, 1], sym[:, 0]
    assert (w > 0).all() and w[par < 0].tolist() == [1.0] * int((par < 0).sum())
    WA = w[:, None] * A
    off = WA[par[ch], ch]
    np.testing.assert_allclose(WA[ch, par[ch]], s[ch], rtol=1e-15)
    np.testing.assert_allclose(off, s[ch], rtol=4e-16)
    b = rng.normal(size=n)
    x = np.linalg.solve(A, b)
    Ws = np.diag(WA.diagonal())
    Ws[ch, par[ch]] = s[ch]
    Ws[par[ch], ch] = s[ch]
    # Both solves round differently; the system's condition (1e4 here) bounds the gap.
    np.testing.assert_allclose(np.linalg.solve(Ws, w * b), x, rtol=1e-10)


def test_symmetric_rows_refuse_couplings_of_opposite_sign():
    from jaxhip import fused as F
    from jaxhip.workloads import tree_cell

    cell = tree_cell(6, 0)[0]
    cell.branch(0).comp(0).record()
    st = F.build_structure(cell, np.array([0]), np.array([0]), torch.device("cpu"))
    nf = torch.zeros((st.n, 4), dtype=torch.float64)
    nf[:, 0] = nf[:, 1] = -0.5
    assert F.symmetric_rows(st, nf) is not None
    k = int(np.nonzero(st.parent_k >= 0)[0][3])
    nf[k, 1] = 0.5
    assert F.symmetric_rows(st, nf) is None
    nf[k, 1] = 0.0
    assert F.symmetric_rows(st, nf) is None
