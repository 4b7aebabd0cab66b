"""Networks through the fused kernel (synapse step fused into the voltage and gate
steps) against the oracle, float64 and float32: ball-and-stick cells, KineticSynapse,
SigmoidRateSynapse and IonotropicSynapse with fixed in-degree onto one dendritic
compartment per cell, Bernoulli current pulses into every soma."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from models import DT, build_c4, c4_stimulus
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu


def _run(device, synapse, dtype, n_cells=12, k=4, B=2, T=400, seed=4):
    net, om, info = build_c4(n_cells, k, seed, synapse)
    assert np.array_equal(net._comp_node[info["pre_comp"]], om.internal[info["pre_comp"]])
    soma = info["soma_comps"]
    for c in soma:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    net.cell(0).branch(1).comp(3).record()
    cur = c4_stimulus(B, n_cells, T, seed + 1)
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(
            torch.tensor(cur[:, i, :], dtype=dtype, device=device), ds)
    v = jx.integrate(net, data_stimuli=ds, delta_t=DT)
    rec = list(soma) + [3 + 1]
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur), stim_comps=list(soma), rec_comps=rec)
    return v, ref


@pytest.mark.parametrize("shape", [None, "1,1024,sym"])
@pytest.mark.parametrize("synapse", ["KineticSynapse", "SigmoidRateSynapse", "IonotropicSynapse"])
def test_network_float64(device, synapse, shape, monkeypatch):
    """Forward against the oracle; with the library's launch shape and with the
    symmetric system (one coupling coefficient per node in LDS)."""
    if shape is not None:
        from jaxhip import autotune

        monkeypatch.setenv("JAXHIP_FUSED_CONFIG", shape)
        monkeypatch.setattr(autotune, "_CACHE", {})
    v, ref = _run(device, synapse, torch.float64)
    assert v.shape == ref.shape
    err = float((v.cpu() - ref).abs().max())
    print(synapse, "fp64 max err", err)
    assert err < 1e-6


def test_network_float32(device):
    v, ref = _run(device, "KineticSynapse", torch.float32)
    err = float((v.double().cpu() - ref).abs().max())
    print("fp32 max err", err)
    assert err < 1e-2


@pytest.mark.parametrize("path", ["fused", "stepwise"])
@pytest.mark.parametrize("synapse", ["KineticSynapse", "SigmoidRateSynapse"])
def test_network_gradients_vs_oracle_autograd(device, synapse, path, monkeypatch):
    """Gradients through a network with respect to every synapse's gS and the somata's
    gNa (checkpoints every 40 steps), against autograd through the float64 CPU oracle; the
    recordings against the oracle too. "fused": the hand-written reverse pass of the fused
    kernel (synapse adjoint and presynaptic shares in the kernel); "stepwise": the
    differentiable per-step path (synapse operator with its hand-written backward, tree
    solve with its transposed solve), taken here because the stimulus requires a gradient
    too."""
    import jaxhip.stepwise as stepwise

    if path == "fused":
        def refuse(*a, **k):
            raise AssertionError("network gradients took the per-step path")
        monkeypatch.setattr(stepwise, "integrate_stepwise", refuse)
    n_cells, k, T, seed = 8, 3, 240, 6
    net, om, info = build_c4(n_cells, k, seed, synapse)
    soma = info["soma_comps"]
    for c in soma:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    cur = c4_stimulus(1, n_cells, T, seed + 1) * 4.0
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(
            torch.tensor(cur[0, i, :], dtype=torch.float64, device=device, requires_grad=path == "stepwise"),
            ds)
    name = synapse
    gs0 = np.asarray(om.synapses[0]["params"]["gS"], dtype=np.float64) * 20.0
    gs = torch.tensor(gs0, dtype=torch.float64, device=device, requires_grad=True)
    gna = torch.full((n_cells,), 0.12, dtype=torch.float64, device=device, requires_grad=True)
    ps = net.data_set(f"{name}_gS", gs, None)
    ps = net.data_set("HH_gNa", gna, ps)
    w = torch.tensor(np.random.default_rng(seed).normal(size=(n_cells, T + 1)), dtype=torch.float64)
    v = jx.integrate(net, param_state=ps, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[6, 40])
    v = v.reshape(n_cells, T + 1)
    loss = (v.cpu() * w).sum()
    loss.backward()
    # oracle
    gs_o = torch.tensor(gs0, dtype=torch.float64, requires_grad=True)
    om.synapses[0]["params"]["gS"] = gs_o
    gna_o = torch.full((len(om.internal),), 0.12, dtype=torch.float64, requires_grad=True)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur), stim_comps=list(soma), rec_comps=list(soma),
                     overrides={"HH_gNa": gna_o})
    (ref[0] * w).sum().backward()
    err_v = float((v.detach().cpu() - ref[0].detach()).abs().max())
    g1, g1o = gs.grad.cpu(), gs_o.grad
    g2, g2o = gna.grad.cpu(), gna_o.grad[torch.as_tensor(soma)]
    e1 = float((g1 - g1o).abs().max() / g1o.abs().max())
    e2 = float((g2 - g2o).abs().max() / g2o.abs().max())
    print(f"{synapse} ({path}): v err {err_v:.2e}, rel grad err gS {e1:.2e}, gNa {e2:.2e}, "
          f"|dL/dgS| max {float(g1o.abs().max()):.3e}")
    assert err_v < 1e-6
    assert float(g1o.abs().max()) > 0 and float(g2o.abs().max()) > 0
    assert e1 < 1e-6 and e2 < 1e-6


@pytest.mark.parametrize("k", [4, 20])
@pytest.mark.parametrize("synapse", ["KineticSynapse", "SigmoidRateSynapse"])
def test_network_fused_adjoint_batched(device, synapse, k, monkeypatch):
    """The fused reverse pass on two simulations with per-simulation gNa (gS shared, its
    gradient summed over the simulations) and a loss on every soma: against autograd
    through the oracle, relative 1e-6 of the largest gradient. k = 20 synapses per target
    walk several synapses per lane of a target's lane group (up to 8 lanes per target)."""
    import jaxhip.stepwise as stepwise

    def refuse(*a, **k):
        raise AssertionError("network gradients took the per-step path")
    monkeypatch.setattr(stepwise, "integrate_stepwise", refuse)
    n_cells, T, seed, B = 16, 200, 8, 2
    net, om, info = build_c4(n_cells, k, seed, synapse)
    soma = info["soma_comps"]
    for c in soma:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    cur = c4_stimulus(B, n_cells, T, seed + 1) * 4.0
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(
            torch.tensor(cur[:, i, :], dtype=torch.float64, device=device), ds)
    gs0 = np.asarray(om.synapses[0]["params"]["gS"], dtype=np.float64) * 20.0
    gna0 = 0.12 * np.random.default_rng(seed).uniform(0.8, 1.2, size=(B, 1))
    gs = torch.tensor(gs0, dtype=torch.float64, device=device, requires_grad=True)
    gna = torch.tensor(gna0, dtype=torch.float64, device=device, requires_grad=True)
    ps = net.data_set(f"{synapse}_gS", gs, None)
    ps = net.data_set("HH_gNa", gna, ps)
    w = torch.tensor(np.random.default_rng(seed + 2).normal(size=(B, n_cells, T + 1)), dtype=torch.float64)
    v = jx.integrate(net, param_state=ps, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[5, 40])
    (v.cpu() * w).sum().backward()
    gs_o = torch.tensor(gs0, dtype=torch.float64, requires_grad=True)
    om.synapses[0]["params"]["gS"] = gs_o
    gna_o = torch.tensor(gna0, dtype=torch.float64, requires_grad=True)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur), stim_comps=list(soma), rec_comps=list(soma),
                     overrides={"HH_gNa": gna_o.expand(B, len(om.internal))})
    (ref * w).sum().backward()
    err_v = float((v.detach().cpu() - ref.detach()).abs().max())
    e1 = float((gs.grad.cpu() - gs_o.grad).abs().max() / gs_o.grad.abs().max())
    e2 = float((gna.grad.cpu() - gna_o.grad).abs().max() / gna_o.grad.abs().max())
    print(f"{synapse} batched: v err {err_v:.2e}, rel grad err gS {e1:.2e}, gNa {e2:.2e}")
    assert err_v < 1e-6
    assert e1 < 1e-6 and e2 < 1e-6


def _two_cell_net(synapse):
    """Cell 0 held at -10 mV (Leak reversal, no input), cell 1 at -70 mV with Leak
    (reversal -65 mV), one synapse 0 -> 1; compartments 10 um long, radius 1 um."""
    cell = jx.Cell(jx.Branch(jx.Compartment(), 1), [-1])
    cell.insert(jx.Leak())
    net = jx.Network([cell, cell])
    net.cell(0).set("Leak_eLeak", -10.0)
    net.cell(0).set("v", -10.0)
    net.cell(1).set("Leak_eLeak", -65.0)
    net.cell(1).set("v", -70.0)
    jx.connect(net.cell(0).branch(0).comp(0), net.cell(1).branch(0).comp(0), synapse)
    net.cell(1).branch(0).comp(0).record()
    return net


def _zero_input(net, dtype, device):
    """A zero current into cell 0 (fixes the device and dtype of the simulation)."""
    return net.cell(0).branch(0).comp(0).data_stimulate(torch.zeros(401, dtype=dtype, device=device))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-9), (torch.float32, 1e-3)])
def test_two_cell_kinetic_closed_form(device, dtype, tol):
    """The fused network step with a KineticSynapse against the closed-form backward
    Euler recurrence of the postsynaptic membrane (exact synapse state at a constant
    presynaptic voltage; tests/closed_forms.py): pins the state update, current, scaling
    and implicit step independently of the oracle."""
    import closed_forms as CF

    net = _two_cell_net(jx.KineticSynapse())
    net.KineticSynapse.set("KineticSynapse_gS", 1e-3)
    net.KineticSynapse.set("KineticSynapse_s", 0.2)
    v = jx.integrate(net, t_max=10.0, delta_t=DT, data_stimuli=_zero_input(net, dtype, device))
    ref = CF.two_cell_kinetic(400, DT, -10.0, -70.0, 0.2, 1e-3, 0.0, 1e-4, -65.0, CF.cylinder_area(10.0, 1.0))
    err = float(np.abs(v[0].double().cpu().numpy() - ref).max())
    print(dtype, "kinetic closed form max err", err)
    assert err < tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-9), (torch.float32, 1e-3)])
def test_two_cell_sigmoid_rate_closed_form(device, dtype, tol):
    import closed_forms as CF

    net = _two_cell_net(jx.SigmoidRateSynapse())
    net.SigmoidRateSynapse.set("SigmoidRateSynapse_gS", 2e-3)
    net.SigmoidRateSynapse.set("SigmoidRateSynapse_x_offset", -12.0)
    v = jx.integrate(net, t_max=10.0, delta_t=DT, data_stimuli=_zero_input(net, dtype, device))
    ref = CF.two_cell_sigmoid_rate(400, DT, -10.0, -70.0, 2e-3, -12.0, 0.25, 1e-4, -65.0,
                                   CF.cylinder_area(10.0, 1.0))
    err = float(np.abs(v[0].double().cpu().numpy() - ref).max())
    print(dtype, "sigmoid-rate closed form max err", err)
    assert err < tol


def _parts_of(net):
    return max(st.parts for st in net.__dict__.get("_fused_cache", {}).values())


@pytest.mark.parametrize("n_cells, dtype, tol", [(1024, torch.float64, 1e-6), (4096, torch.float32, 1e-2)])
def test_network_split_full_size(device, n_cells, dtype, tol):
    """Networks larger than one workgroup's LDS: the 1024-cell network (10,240 solve
    nodes) in float64 and a 4096-cell one (40,960 nodes) in float32 ("mixed"), 32
    KineticSynapses per cell, split into parts that advance in lockstep; current pulses
    that make the somata spike; against the float64 oracle (float32: on the same float32
    inputs), spike counts identical."""
    k, T, B = 32, 400, 2
    net, om, info = build_c4(n_cells, k, seed=12)
    soma = info["soma_comps"]
    rec = soma[::64]
    for c in rec:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    cur = c4_stimulus(B, n_cells, T, seed=13, amp=1.5, p=0.1)
    if dtype == torch.float32:
        cur = cur.astype(np.float32).astype(np.float64)
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(
            torch.tensor(cur[:, i, :], dtype=dtype, device=device), ds)
    v = jx.integrate(net, data_stimuli=ds, delta_t=DT)
    parts = _parts_of(net)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur), stim_comps=list(soma), rec_comps=list(rec))
    err = float((v.double().cpu() - ref).abs().max())
    sc = ((ref[..., 1:] >= 0) & (ref[..., :-1] < 0)).sum(-1)
    sv = ((v.double().cpu()[..., 1:] >= 0) & (v.double().cpu()[..., :-1] < 0)).sum(-1)
    print(f"{n_cells} cells, {dtype}: {parts} parts, max err {err:.3e} mV, spikes {int(sc.sum())}")
    assert parts > 1
    assert int(sc.sum()) > 20 and torch.equal(sc, sv)
    assert err < tol
