"""Edge cases of the fused path against the oracle: an unbranched 1500-compartment cable
(one segment: the longest chain walk), stimuli shorter and longer than the simulation,
checkpoint lengths that do not divide the number of steps, a network whose cells
partly receive no synapses, and a compartment recorded twice."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from models import DT, c1_stimulus
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [None, "4,1024,sym", "2,256"])
def test_unbranched_cable(device, shape, monkeypatch):
    """One branch of 1500 compartments (lengths and radii varying along it), HH
    everywhere, B = 3 with per-simulation gNa; recorded at both ends and in the middle."""
    from jaxhip import autotune

    if shape is not None:
        monkeypatch.setenv("JAXHIP_FUSED_CONFIG", shape)
    monkeypatch.setattr(autotune, "_CACHE", {})
    n, B, T = 1500, 3, 300
    rng = np.random.default_rng(2)
    length = rng.uniform(5.0, 15.0, n)
    radius = np.linspace(2.0, 0.5, n)
    cell = jx.Cell([jx.Branch(jx.Compartment(), n)], [-1])
    cell.set("length", length)
    cell.set("radius", radius)
    cell.insert(jx.HH())
    for c in (0, n // 2, n - 1):
        cell.branch(0).comp(c).record()
    gna = np.array([[0.12], [0.08], [0.16]])
    ps = cell.data_set("HH_gNa", torch.tensor(gna, dtype=torch.float64, device=device), None)
    cur = c1_stimulus(T) * 2.0
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT).cpu().numpy()
    om = O.OracleModel([([-1], [n])])
    om.length, om.radius, om.hh_comps = length.copy(), radius.copy(), np.arange(n)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0, n // 2, n - 1],
                     overrides={"HH_gNa": np.broadcast_to(gna, (B, n)).copy()}).numpy()
    err = np.abs(v - ref).max()
    print(shape, "cable max err", err, "root range", v[:, 0].min(), v[:, 0].max())
    assert v.shape == (B, 3, T + 1)
    assert err < 1e-6 and v[:, 0].max() > 0  # the stimulated end spikes


@pytest.mark.parametrize("stim_steps", [120, 400])
def test_stimulus_padding_and_truncation(device, stim_steps):
    """A stimulus shorter than the simulation (zero-padded) and longer (truncated), with
    t_max giving the number of steps."""
    from models import build_tree

    cell, om = build_tree(10, seed=5)
    cell.branch(0).comp(0).record()
    cur = np.zeros(stim_steps)
    cur[20:100] = 0.3
    t_max = 6.0
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT, t_max=t_max).cpu().numpy()
    ref = O.simulate(om, t_max, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0]).numpy()
    print(stim_steps, "max err", np.abs(v - ref).max(), v.shape)
    # int(t_max // dt + 1) steps (integrate.py:523-541 of the reference), plus the
    # initial state.
    assert v.shape[-1] == int(t_max // DT + 1) + 1 == ref.shape[-1]
    assert np.abs(v - ref).max() < 1e-6


@pytest.mark.parametrize("lengths", [[3, 70], [7, 31], [1, 1000]])
def test_adjoint_checkpoints_not_dividing_steps(device, lengths):
    """Checkpoint lengths whose product is not the number of steps (the last segment is
    shorter; one segment longer than the run): gradients against oracle autograd."""
    from models import build_tree

    B, T, nb = 2, 205, 8
    cell, om = build_tree(nb, seed=9)
    cell.branch(0).comp(0).record()
    for k in ("HH_gNa", "HH_gK", "HH_gLeak"):
        cell.branch("all").make_trainable(k, verbose=False)
    vals = np.array([0.12, 0.036, 0.0003])[None, :, None] * np.random.default_rng(3).uniform(0.8, 1.2, (B, 3, nb))
    params = [{k: torch.tensor(vals[:, j], dtype=torch.float64, device=device, requires_grad=True)}
              for j, k in enumerate(("HH_gNa", "HH_gK", "HH_gLeak"))]
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    v = jx.integrate(cell, params, data_stimuli=ds, delta_t=DT, checkpoint_lengths=lengths)
    target = torch.linspace(-70, -60, T + 1, dtype=torch.float64)
    ((v[:, 0].cpu() - target) ** 2).mean().backward()
    g = np.stack([list(p.values())[0].grad.cpu().numpy() for p in params], 1)
    ov = torch.tensor(vals, requires_grad=True)
    comp = torch.repeat_interleave(ov, 4, dim=-1)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0],
                     overrides={"HH_gNa": comp[:, 0], "HH_gK": comp[:, 1], "HH_gLeak": comp[:, 2]})
    ((ref[:, 0] - target) ** 2).mean().backward()
    np.testing.assert_allclose(g, ov.grad.numpy(), rtol=1e-6, atol=1e-9 * np.abs(ov.grad.numpy()).max())


def test_network_cells_without_synapses_and_double_recording(device):
    """Five ball-and-stick cells, KineticSynapses onto cells 1 and 3 only (cells 0, 2
    and 4 receive none; cell 4 sends none either), the soma of cell 1 recorded twice."""
    from jaxhip.workloads import C4_LENGTH, C4_RADIUS

    cell = jx.Cell([jx.Branch(jx.Compartment(), 1), jx.Branch(jx.Compartment(), 8)], [-1, 0])
    cell.set("length", C4_LENGTH)
    cell.set("radius", C4_RADIUS)
    cell.branch(0).insert(jx.HH())
    cell.branch(1).insert(jx.Leak())
    net = jx.Network([cell] * 5)
    pairs = [(0, 1, 2), (2, 1, 2), (0, 3, 5), (1, 3, 5), (2, 3, 5)]   # (pre cell, post cell, dendrite comp)
    for pre, post, comp in pairs:
        jx.connect(net.cell(pre).branch(0).comp(0), net.cell(post).branch(1).comp(comp), jx.KineticSynapse())
    gs = np.array([4e-3, 3e-3, 5e-3, 2e-3, 4e-3])
    net.KineticSynapse.set("KineticSynapse_gS", gs)
    rec = [0, 1, 1, 2, 3, 4]
    for c in rec:
        net.cell(c).branch(0).comp(0).record()
    T = 600
    cur = np.zeros((5, T))
    cur[[0, 2], 40:440] = 0.3
    ds = None
    for c in range(5):
        ds = net.cell(c).branch(0).comp(0).data_stimulate(torch.tensor(cur[c], dtype=torch.float64, device=device), ds)
    v = jx.integrate(net, data_stimuli=ds, delta_t=DT).cpu().numpy()
    om = O.OracleModel([([-1, 0], [1, 8])] * 5)
    om.length, om.radius = np.tile(C4_LENGTH, 5), np.tile(C4_RADIUS, 5)
    om.hh_comps = np.arange(5) * 9
    om.leak_comps = np.setdiff1d(np.arange(45), om.hh_comps)
    syn = jx.KineticSynapse()
    params = {name: np.full(5, val, dtype=np.float64) for name, val in
              ((n, syn.synapse_params[f"KineticSynapse_{n}"]) for n in syn.param_names)}
    params["gS"] = gs.copy()
    pre = np.array([p[0] * 9 for p in pairs])
    post = np.array([p[1] * 9 + 1 + p[2] for p in pairs])
    om.synapses = [{"type": "KineticSynapse", "pre": om.internal[pre], "post": om.internal[post], "params": params,
                    "s0": np.zeros(5)}]
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=list(np.arange(5) * 9),
                     rec_comps=[c * 9 for c in rec]).numpy()[0]
    err = np.abs(v - ref).max()
    print("network max err", err)
    assert v.shape == (6, T + 1) and err < 1e-6
    assert np.array_equal(v[1], v[2])
    assert v[0].max() > 0 and v[4].max() < -60   # cell 0 spikes; cell 4 stays at rest
