"""Vectors written by jaxley-pytorch itself (tests/golden/make_reference_vectors.py,
float64 on the CPU): the DHS solve indexers of the binary-tree cell, the 250-branch
random tree and a 6-cell network, bit for bit; 30 ms traces of one ball-and-stick cell
and of 6-cell ball-and-stick networks with KineticSynapse / SigmoidRateSynapse (spiking
somata) through the oracle (CPU) and through the fused kernel and the per-step path
(GPU)."""
from pathlib import Path

import numpy as np
import pytest
import torch

import jaxhip as jx
from jaxhip.workloads import binary_tree_parents, random_tree
from oracle import jx_oracle as O

REF = np.load(Path(__file__).resolve().parent / "golden" / "reference_vectors.npz")
DT, T_MAX = float(REF["dt"]), float(REF["t_max"])
N_CELLS = 6
LENGTH = np.array([20.0] + [25.0] * 8)
RADIUS = np.array([10.0] + [1.0] * 8)
SYNAPSES = ["KineticSynapse", "SigmoidRateSynapse"]
KEYS = ("map_to_solve_order", "inv_map_to_solve_order", "map_to_solve_order_lower",
        "map_to_solve_order_upper", "node_order", "parent_lookup")


def _ball_and_stick():
    cell = jx.Cell([jx.Branch(jx.Compartment(), 1), jx.Branch(jx.Compartment(), 8)], [-1, 0])
    cell.set("length", LENGTH)
    cell.set("radius", RADIUS)
    cell.branch(0).insert(jx.HH())
    cell.branch(1).insert(jx.Leak())
    return cell


def _spec(synapse):
    key = synapse.lower()
    return {k: REF[f"{key}_{k}"] for k in ("pre", "tgt", "gS", "e_syn", "amp", "delay", "v")}


def _assert_indexer(module, prefix):
    ind = module._dhs_solve_indexer
    for k in KEYS:
        assert np.array_equal(np.asarray(ind[k]), REF[f"{prefix}_{k}"]), (prefix, k)
    grouped = ind["node_order_grouped"]
    assert [len(g) for g in grouped] == REF[f"{prefix}_grouped_sizes"].tolist()
    assert np.array_equal(np.concatenate(grouped, 0), REF[f"{prefix}_grouped"])


def test_indexer_binary_tree_cell():
    cell = jx.Cell([jx.Branch(jx.Compartment(), 4)] * 15, binary_tree_parents(15))
    _assert_indexer(cell, "c0")


def test_indexer_random_tree():
    parents, _, _ = random_tree(250, 11)
    assert np.array_equal(parents, REF["c1_parents"])
    cell = jx.Cell([jx.Branch(jx.Compartment(), 4)] * 250, parents)
    _assert_indexer(cell, "c1")


def test_indexer_network():
    net = jx.Network([_ball_and_stick() for _ in range(N_CELLS)])
    _assert_indexer(net, "net")


def _network(synapse, sp, device=None, dtype=torch.float64):
    net = jx.Network([_ball_and_stick() for _ in range(N_CELLS)])
    cls = getattr(jx, synapse)
    for post in range(N_CELLS):
        for j in range(sp["pre"].shape[1]):
            jx.connect(net.cell(int(sp["pre"][post, j])).branch(0).comp(0),
                       net.cell(post).branch(1).comp(int(sp["tgt"][post])), cls())
    getattr(net, synapse).set(f"{synapse}_gS", sp["gS"].reshape(-1))
    if synapse == "KineticSynapse":
        getattr(net, synapse).set(f"{synapse}_e_syn", sp["e_syn"].reshape(-1))
    return net


def _currents(sp):
    # step_current builds the current in the default dtype (as the reference's, which
    # wrote these vectors with float64 as the default).
    prev = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    try:
        return [jx.step_current(float(sp["delay"][c]), 15.0, float(sp["amp"][c]), DT, T_MAX)
                for c in range(N_CELLS)]
    finally:
        torch.set_default_dtype(prev)


def _oracle_network(synapse, sp):
    """The oracle's model of the network (synapse parameters of the built network)."""
    net = _network(synapse, sp)
    cls = getattr(jx, synapse)()
    om = O.OracleModel([([-1, 0], [1, 8])] * N_CELLS)
    om.length = np.tile(LENGTH, N_CELLS)
    om.radius = np.tile(RADIUS, N_CELLS)
    om.hh_comps = np.arange(N_CELLS) * 9
    om.leak_comps = np.setdiff1d(np.arange(9 * N_CELLS), om.hh_comps)
    edge = net._edges[0]
    params = {name: np.asarray(edge["params"][f"{synapse}_{name}"], dtype=np.float64).copy()
              for name in cls.param_names}
    pre_comp = (sp["pre"] * 9).reshape(-1)
    post_comp = np.repeat(np.arange(N_CELLS) * 9 + 1 + sp["tgt"], sp["pre"].shape[1])
    s0 = np.zeros(len(pre_comp)) if synapse == "KineticSynapse" else None
    om.synapses = [{"type": synapse, "pre": om.internal[pre_comp], "post": om.internal[post_comp],
                    "params": params, "s0": s0}]
    return om


@pytest.mark.parametrize("synapse", SYNAPSES)
def test_oracle_network_traces(synapse):
    """The oracle against the reference's traces (every recorded soma spikes twice)."""
    sp = _spec(synapse)
    om = _oracle_network(synapse, sp)
    stim = torch.stack(_currents(sp))[None]
    rec = list(np.arange(N_CELLS) * 9) + list(np.arange(N_CELLS) * 9 + 1 + sp["tgt"])
    ref = O.simulate(om, T_MAX, DT, stim=stim, stim_comps=list(np.arange(N_CELLS) * 9), rec_comps=rec)[0]
    err = float(np.abs(ref.numpy()[:, ::4] - sp["v"]).max())
    print(synapse, "oracle vs reference vectors, max", err)
    assert err < 1e-10   # measured 0: the oracle repeats the reference's operations


def test_oracle_single_cell_trace():
    """One ball-and-stick cell (HH soma, Leak dendrite), a step current into the soma."""
    om = O.OracleModel([([-1, 0], [1, 8])])
    om.length, om.radius = LENGTH.copy(), RADIUS.copy()
    om.hh_comps, om.leak_comps = np.array([0]), np.arange(1, 9)
    prev = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    try:
        stim = jx.step_current(1.31, 15.0, 0.1536, DT, T_MAX)[None, None]
    finally:
        torch.set_default_dtype(prev)
    ref = O.simulate(om, T_MAX, DT, stim=stim, stim_comps=[0], rec_comps=[0, 4])[0].numpy()[:, ::4]
    err = float(np.abs(ref - REF["cell_v"]).max())
    print("single cell, oracle vs reference vectors, max", err)
    assert err < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["fused", "stepwise"])
@pytest.mark.parametrize("synapse", SYNAPSES)
def test_network_traces_gpu(device, synapse, path, monkeypatch):
    """`integrate` on the GPU against the reference's traces: the fused kernel (synapse
    step fused into the voltage and gate steps) and the per-step path (synapse and
    tree-solve operators), float64, 1e-6 mV; identical spike counts."""
    import sys

    integ = sys.modules["jaxhip.integrate"]   # the module (the package exports the function)
    sp = _spec(synapse)
    net = _network(synapse, sp)
    for c, cur in enumerate(_currents(sp)):
        net.cell(c).branch(0).comp(0).stimulate(cur.to(device))
    for c in range(N_CELLS):
        net.cell(c).branch(0).comp(0).record()
    for c in range(N_CELLS):
        net.cell(c).branch(1).comp(int(sp["tgt"][c])).record()
    if path == "stepwise":
        monkeypatch.setattr(integ, "_per_step_path", lambda *a, **k: True)
    v = jx.integrate(net, delta_t=DT, t_max=T_MAX).double().cpu().numpy()
    v = v.reshape(-1, v.shape[-1])[:, ::4]
    err = float(np.abs(v - sp["v"]).max())
    print(synapse, path, "vs reference vectors, max", err)
    assert err < 1e-6
    up = lambda x: ((x[:, 1:] >= 0) & (x[:, :-1] < 0)).sum(1)
    assert np.array_equal(up(v[:N_CELLS]), up(sp["v"][:N_CELLS])) and up(sp["v"][:N_CELLS]).min() > 0


# ---- gradients of the reference's autograd ----

GT_KEYS = ("HH_gNa", "HH_gK", "HH_gLeak")


def _grad_tree_cell():
    nb = len(REF["grad_tree_parents"])
    cell = jx.Cell([jx.Branch(jx.Compartment(), 4)] * nb, REF["grad_tree_parents"])
    cell.set("length", REF["grad_tree_length"])
    cell.set("radius", REF["grad_tree_radius"])
    cell.insert(jx.HH())
    cell.branch(0).comp(0).record()
    for k in GT_KEYS:
        cell.branch("all").make_trainable(k, verbose=False)
    return cell


def test_oracle_tree_gradients():
    """Autograd through the oracle against the reference's autograd: per-branch HH
    conductances of two parameter sets, MSE of the root voltage."""
    vals = REF["grad_tree_vals"]
    parents, nb = REF["grad_tree_parents"], vals.shape[-1]
    om = O.OracleModel([(parents, [4] * nb)])
    om.length, om.radius = REF["grad_tree_length"].copy(), REF["grad_tree_radius"].copy()
    om.hh_comps = np.arange(4 * nb)
    ovals = torch.tensor(vals, requires_grad=True)
    comp = torch.repeat_interleave(ovals, 4, dim=-1)
    ov = {k: comp[:, j] for j, k in enumerate(GT_KEYS)}
    stim = torch.tensor(REF["grad_tree_stim"])[None]
    T = stim.shape[-1]
    v = O.simulate(om, None, DT, stim=stim, stim_comps=[0], rec_comps=[0], overrides=ov)
    loss = ((v[:, 0] - torch.linspace(-70, -60, T + 1, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    np.testing.assert_allclose(loss.item(), float(REF["grad_tree_loss"]), rtol=1e-12)
    rel = np.abs(ovals.grad.numpy() - REF["grad_tree_grads"]).max() / np.abs(REF["grad_tree_grads"]).max()
    print("oracle tree gradients vs the reference's autograd, max relative to the largest", rel)
    assert rel < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,rtol", [(torch.float64, 1e-6), (torch.float32, 1e-2)])
def test_tree_gradients_gpu(device, dtype, rtol):
    """The hand-written reverse pass (fused kernel, checkpoints every 50 steps) against
    the reference's autograd; float32 runs the library's default ("mixed") mode."""
    cell = _grad_tree_cell()
    vals = REF["grad_tree_vals"]
    params = [{k: torch.tensor(vals[:, j], dtype=dtype, device=device, requires_grad=True)}
              for j, k in enumerate(GT_KEYS)]
    stim = torch.tensor(REF["grad_tree_stim"], dtype=dtype, device=device)
    T = stim.shape[-1]
    ds = cell.branch(0).comp(0).data_stimulate(stim)
    v = jx.integrate(cell, params, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[T // 50, 50])
    target = torch.linspace(-70, -60, T + 1, dtype=torch.float64, device=device)
    loss = ((v[:, 0].double() - target) ** 2).mean()
    loss.backward()
    g = np.stack([list(p.values())[0].grad.double().cpu().numpy() for p in params], 1)
    ref = REF["grad_tree_grads"]
    rel = np.abs(g - ref).max() / np.abs(ref).max()
    print(dtype, "tree gradients vs the reference's autograd, max relative to the largest", rel)
    np.testing.assert_allclose(loss.item(), float(REF["grad_tree_loss"]), rtol=1e-9 if dtype == torch.float64 else 1e-3)
    assert rel < rtol
    if dtype == torch.float64:
        np.testing.assert_allclose(g, ref, rtol=1e-6, atol=1e-9 * np.abs(ref).max())


def _grad_net_inputs(device):
    sp = _spec("KineticSynapse")
    net = jx.Network([_ball_and_stick() for _ in range(N_CELLS)])
    for post in range(N_CELLS):
        for j in range(sp["pre"].shape[1]):
            jx.connect(net.cell(int(sp["pre"][post, j])).branch(0).comp(0),
                       net.cell(post).branch(1).comp(int(sp["tgt"][post])), jx.KineticSynapse())
    net.KineticSynapse.set("KineticSynapse_e_syn", sp["e_syn"].reshape(-1))
    for c, cur in enumerate(_currents(sp)):
        net.cell(c).branch(0).comp(0).stimulate(cur.to(device) if device is not None else cur)
    for c in range(N_CELLS):
        net.cell(c).branch(0).comp(0).record()
    seed, *shape = REF["grad_net_w_seed_shape"].tolist()
    w = np.random.default_rng(seed).normal(size=tuple(shape))
    return net, sp, w


def test_oracle_network_gradients():
    """Autograd through the oracle against the reference's: every gS and the somata's
    gNa of the KineticSynapse network, two simulations."""
    sp = _spec("KineticSynapse")
    om = _oracle_network("KineticSynapse", sp)
    gs = torch.tensor(sp["gS"].reshape(-1), requires_grad=True)
    om.synapses[0]["params"]["gS"] = gs
    gna = torch.tensor(REF["grad_net_gna"], requires_grad=True)
    soma = list(np.arange(N_CELLS) * 9)
    stim = torch.stack(_currents(sp))[None]
    v = O.simulate(om, T_MAX, DT, stim=stim, stim_comps=soma, rec_comps=soma,
                   overrides={"HH_gNa": gna.expand(2, len(om.internal))})
    seed, *shape = REF["grad_net_w_seed_shape"].tolist()
    w = np.random.default_rng(seed).normal(size=tuple(shape))
    (v * torch.tensor(w)).sum().backward()
    err_v = float(np.abs(v.detach().numpy()[:, :, ::4] - REF["grad_net_v"]).max())
    e1 = np.abs(gs.grad.numpy() - REF["grad_net_grad_gS"]).max() / np.abs(REF["grad_net_grad_gS"]).max()
    e2 = np.abs(gna.grad.numpy() - REF["grad_net_grad_gNa"]).max() / np.abs(REF["grad_net_grad_gNa"]).max()
    print(f"oracle network vs the reference: v {err_v:.2e} mV, gS {e1:.2e}, gNa {e2:.2e}")
    assert err_v < 1e-10 and e1 < 1e-10 and e2 < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("path", ["fused", "stepwise"])
def test_network_gradients_gpu(device, path, monkeypatch):
    """Every gS and the somata's gNa of the KineticSynapse network against the
    reference's autograd: the fused reverse pass (synapse adjoint in the kernel) and the
    per-step path."""
    import sys

    integ = sys.modules["jaxhip.integrate"]
    net, sp, w = _grad_net_inputs(device)
    gs = torch.tensor(sp["gS"].reshape(-1), dtype=torch.float64, device=device, requires_grad=True)
    gna = torch.tensor(REF["grad_net_gna"], dtype=torch.float64, device=device, requires_grad=True)
    ps = net.data_set("KineticSynapse_gS", gs, None)
    ps = net.data_set("HH_gNa", gna, ps)
    if path == "stepwise":
        monkeypatch.setattr(integ, "_per_step_path", lambda *a, **k: True)
    else:
        import jaxhip.stepwise as stepwise

        def refuse(*a, **k):
            raise AssertionError("network gradients took the per-step path")
        monkeypatch.setattr(stepwise, "integrate_stepwise", refuse)
    v = jx.integrate(net, param_state=ps, delta_t=DT, t_max=T_MAX, checkpoint_lengths=[12, 100])
    (v.cpu() * torch.tensor(w)).sum().backward()
    err_v = float(np.abs(v.detach().cpu().numpy()[:, :, ::4] - REF["grad_net_v"]).max())
    e1 = np.abs(gs.grad.cpu().numpy() - REF["grad_net_grad_gS"]).max() / np.abs(REF["grad_net_grad_gS"]).max()
    e2 = np.abs(gna.grad.cpu().numpy() - REF["grad_net_grad_gNa"]).max() / np.abs(REF["grad_net_grad_gNa"]).max()
    print(path, f"network vs the reference: v {err_v:.2e} mV, gS {e1:.2e}, gNa {e2:.2e}")
    assert err_v < 1e-6 and e1 < 1e-6 and e2 < 1e-6


# ---- the configurations' cells: 60-compartment binary tree, 1000-compartment sweep tree ----

def _c0_cell():
    cell = jx.Cell([jx.Branch(jx.Compartment(), 4)] * 15, binary_tree_parents(15))
    cell.set("length", 2.5)
    cell.insert(jx.HH())
    cell.record()
    return cell


def _c0_current():
    prev = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    try:
        return jx.step_current(1.0, 1.0, 0.1, DT, 10.0)
    finally:
        torch.set_default_dtype(prev)


def _c1_cell():
    from jaxhip.workloads import tree_cell

    cell = tree_cell(250, 11)[0]
    cell.branch(0).comp(0).record()
    return cell


def _c1_stim():
    cur = np.zeros(2000)
    cur[int(1.0 / DT):int(21.0 / DT)] = 0.2
    return cur


def test_oracle_config_traces():
    """The oracle against the reference on the 60-compartment cell (all compartments)
    and on four simulations of the sweep tree (2000 steps)."""
    from models import build_c0, build_tree, oracle_overrides

    _, om = build_c0()
    ref = O.simulate(om, 10.0, DT, stim=_c0_current()[None, None], stim_comps=[0], rec_comps=list(range(60)))
    e0 = float(np.abs(ref.numpy()[0][:, ::4] - REF["c0_v"]).max())
    _, om1 = build_tree(250, 11)
    p = REF["c1_params"]
    prm = {"HH_gNa": p[:, 0:1], "HH_gK": p[:, 1:2], "HH_gLeak": p[:, 2:3]}
    ref1 = O.simulate(om1, None, DT, stim=torch.tensor(_c1_stim())[None], stim_comps=[0], rec_comps=[0],
                      overrides=oracle_overrides(om1, prm))
    e1 = float(np.abs(ref1.numpy()[:, :, ::4] - REF["c1_v"]).max())
    print("oracle vs reference: 60-compartment cell", e0, "sweep tree", e1)
    assert e0 < 1e-10 and e1 < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [None, "8,1024,sym", "1,1024"])
def test_config_traces_gpu(device, shape, monkeypatch):
    """The fused forward against the reference: the 60-compartment cell and four
    simulations of the sweep tree (2000 steps; the library's launch shape, the sweep's
    8 simulations per program with the symmetric system, one per program), float64,
    1e-6 mV, identical spike counts."""
    from jaxhip import autotune
    from jaxhip.workloads import data_set_params

    if shape is not None:
        monkeypatch.setenv("JAXHIP_FUSED_CONFIG", shape)
    monkeypatch.setattr(autotune, "_CACHE", {})
    c0 = _c0_cell()
    c0.branch(0).comp(0).stimulate(_c0_current().to(device))
    v0 = jx.integrate(c0, delta_t=DT, t_max=10.0).cpu().numpy()[:, ::4]
    c1 = _c1_cell()
    p = REF["c1_params"]
    prm = {"HH_gNa": p[:, 0:1], "HH_gK": p[:, 1:2], "HH_gLeak": p[:, 2:3]}
    ps = data_set_params(c1, prm, device, torch.float64)
    ds = c1.branch(0).comp(0).data_stimulate(torch.tensor(_c1_stim(), dtype=torch.float64, device=device))
    v1 = jx.integrate(c1, param_state=ps, data_stimuli=ds, delta_t=DT).cpu().numpy()[:, :, ::4]
    e0 = float(np.abs(v0 - REF["c0_v"]).max())
    e1 = float(np.abs(v1 - REF["c1_v"]).max())
    print(shape, "vs reference: 60-compartment cell", e0, "sweep tree", e1)
    assert e0 < 1e-6 and e1 < 1e-6
    up = lambda x: ((x[..., 1:] >= 0) & (x[..., :-1] < 0)).sum(-1)
    assert np.array_equal(up(v1), up(REF["c1_v"])) and up(REF["c1_v"]).sum() > 0


# ---- float32: the reference's own float32 simulation of the sweep ----

def test_reference_float32_misses_1e2_itself():
    """The reference's own float32 simulation of 138 sweep simulations (the ones whose
    float32 GPU run missed 1e-2 mV in one measurement, and the first eight) against its
    float64 simulation of the same float32 inputs: the tolerance is missed by the
    reference itself on most of them (controls included), while the spike counts agree."""
    e = REF["f32_sweep_err"]
    print("reference float32 vs float64: max", e.max(), "median", np.median(e), "over 1e-2:", (e > 1e-2).sum(),
          "of", len(e), "controls", np.round(e[:8], 4).tolist())
    assert (e > 1e-2).sum() > len(e) // 2 and e[:8].max() > 1e-2
    assert np.array_equal(REF["f32_sweep_spikes32"], REF["f32_sweep_spikes64"])


@pytest.mark.gpu
def test_float32_modes_vs_reference_float32(device):
    """The library's float32 modes on those simulations against its float64 run of the
    same float32 inputs, beside the reference's own float32 error on each: the largest
    and the median error of both modes stay below the reference's float32 ones; spike
    counts equal float64's."""
    from jaxhip.integrate import set_float32_mode
    from jaxhip.workloads import c1_params

    idx = REF["f32_sweep_sims"]
    parents, blen, rad = random_tree(250, 11)
    L32 = np.repeat(blen / 4, 4).astype(np.float32).astype(np.float64)
    R32 = np.repeat(rad, 4).astype(np.float32).astype(np.float64)
    prm = {k: v[idx].astype(np.float32) for k, v in c1_params(4096, 11).items()}
    cur = np.zeros(2000, dtype=np.float32)
    cur[int(1.0 / DT):int(21.0 / DT)] = np.float32(0.2)

    def run(dtype):
        cell = jx.Cell([jx.Branch(jx.Compartment(), 4)] * 250, parents)
        cell.set("length", L32)
        cell.set("radius", R32)
        cell.insert(jx.HH())
        cell.branch(0).comp(0).record()
        ps = None
        for k, v in prm.items():
            ps = cell.data_set(k, torch.tensor(v, dtype=dtype, device=device), ps)
        ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=dtype, device=device))
        with torch.no_grad():
            return jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT).double().cpu().numpy()

    v64 = run(torch.float64)
    ref_e = REF["f32_sweep_err"]
    up = lambda v: ((v[..., 1:] >= 0) & (v[..., :-1] < 0)).sum(-1)[:, 0]
    try:
        for mode in ("mixed", "fast"):
            set_float32_mode(mode)
            e = np.abs(run(torch.float32) - v64).max(axis=(1, 2))
            print(f"{mode}: max {e.max():.4f} median {np.median(e):.5f} over 1e-2 {(e > 1e-2).sum()} of {len(e)}; "
                  f"reference float32: max {ref_e.max():.4f} median {np.median(ref_e):.5f}; "
                  f"below the reference's on {(e <= ref_e).sum()}")
            assert e.max() < ref_e.max() and np.median(e) < np.median(ref_e)
            assert np.array_equal(up(run(torch.float32)), up(v64))
    finally:
        set_float32_mode("mixed")
