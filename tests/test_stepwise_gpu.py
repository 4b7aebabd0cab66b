"""The per-step path on the GPU against the oracle: recordings of states other than the
voltage, batched and differentiated geometry, and gradients with respect to stimuli
and initial gate states (integrate.py:53-84, 638-645; modules/base.py:1319-1344,
1917-2021). Float64; voltages and gradients within 1e-6 (relative for gradients)."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from models import DT, build_c0, build_c4, c0_stimulus, c4_stimulus, oracle_of_module
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu


def _c0_with_stim(device, T=200, amp=1.0, requires_grad=False):
    cell, om = build_c0()
    cell.delete_recordings()
    cur = c0_stimulus(T) * amp
    x = torch.tensor(cur, dtype=torch.float64, device=device, requires_grad=requires_grad)
    ds = cell.branch(0).comp(0).data_stimulate(x)
    return cell, om, cur, x, ds


def test_record_gates_and_currents(device):
    """`record("HH_m")`, `record("i_HH")` and voltages in one run (mixed recordings keep
    their order) against the oracle's recorder."""
    cell, om, cur, _, ds = _c0_with_stim(device)
    rec = [("v", 0), ("HH_m", 5), ("i_HH", 5), ("HH_h", 59), ("v", 33), ("HH_n", 0), ("i_HH", 12)]
    for state, c in rec:
        cell.branch(c // 4).comp(c % 4).record(state)
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], record=rec)
    assert v.shape == ref.shape == (len(rec), 201)
    err = (v.cpu() - ref).abs().amax(-1)
    print("max err per recording", err.tolist())
    assert float(err.max()) < 1e-6
    assert float(ref[2].abs().max()) > 1e-3          # the current moved


def test_record_synapse_states_and_currents(device):
    net, om, info = build_c4(6, 3, seed=2)
    soma = info["soma_comps"]
    cur = c4_stimulus(1, 6, 300, seed=3, p=0.2) * 20.0
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(
            torch.tensor(cur[0, i], dtype=torch.float64, device=device), ds)
    net.cell(1).branch(0).comp(0).record()
    net.KineticSynapse.edge([0, 4, 11]).record("KineticSynapse_s")
    net.KineticSynapse.edge([4, 17]).record("i_KineticSynapse")
    v = jx.integrate(net, data_stimuli=ds, delta_t=DT)
    rec = [("v", 9), ("KineticSynapse_s", 0), ("KineticSynapse_s", 4), ("KineticSynapse_s", 11),
           ("i_KineticSynapse", 4), ("i_KineticSynapse", 17)]
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur), stim_comps=list(soma), record=rec)[0]
    assert v.shape == ref.shape == (len(rec), 301)
    err = (v.cpu() - ref).abs().amax(-1)
    print("max err per recording", err.tolist(), "max s", float(ref[1:4].abs().max()))
    assert float(err.max()) < 1e-6
    assert float(ref[1:4].abs().max()) > 1e-3


def test_batched_radius(device):
    """`data_set("radius", (B, n_comp))`: every simulation its own geometry."""
    cell, om, cur, _, ds = _c0_with_stim(device)
    cell.branch(0).comp(0).record()
    cell.branch(14).comp(3).record()
    rng = np.random.default_rng(0)
    rad = rng.uniform(0.6, 1.6, size=(3, 60))
    ps = cell.data_set("radius", torch.tensor(rad, dtype=torch.float64, device=device), None)
    v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0, 59],
                     overrides={"radius": torch.tensor(rad)})
    err = float((v.cpu() - ref).abs().max())
    print("batched radius max err", err, "spread", float(ref[:, 0].std(0).max()))
    assert v.shape == (3, 2, 201)
    assert err < 1e-6
    assert float(ref[:, 0].std(0).max()) > 0.1


def _grad_check(g, g_ref, name):
    rel = float((g.cpu() - g_ref).abs().max() / g_ref.abs().max())
    print(f"{name}: relative gradient error {rel:.2e}, |g| max {float(g_ref.abs().max()):.3e}")
    assert float(g_ref.abs().max()) > 0
    assert rel < 1e-6


def test_gradient_wrt_radius_and_length(device):
    cell, om, cur, _, ds = _c0_with_stim(device, T=160)
    cell.branch(0).comp(0).record()
    r0 = np.linspace(0.8, 1.3, 60)
    rad = torch.tensor(r0, dtype=torch.float64, device=device, requires_grad=True)
    ln = torch.tensor(np.full(60, 2.5), dtype=torch.float64, device=device, requires_grad=True)
    ps = cell.data_set("radius", rad, None)
    ps = cell.data_set("length", ln, ps)
    v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[4, 40])
    w = torch.tensor(np.random.default_rng(1).normal(size=161), dtype=torch.float64)
    (v[0].cpu() * w).sum().backward()
    rad_o = torch.tensor(r0, dtype=torch.float64, requires_grad=True)
    ln_o = torch.tensor(np.full(60, 2.5), dtype=torch.float64, requires_grad=True)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0],
                     overrides={"radius": rad_o, "length": ln_o})
    (ref[0] * w).sum().backward()
    assert float((v.detach().cpu() - ref.detach()).abs().max()) < 1e-6
    _grad_check(rad.grad, rad_o.grad, "radius")
    _grad_check(ln.grad, ln_o.grad, "length")


def test_gradient_wrt_stimulus(device):
    """d(loss) / d(stimulus current) (integrate.py:638-645 differentiates the inputs)."""
    cell, om, cur, x, ds = _c0_with_stim(device, T=160, amp=1.5, requires_grad=True)
    cell.branch(3).comp(1).record()
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT)
    w = torch.tensor(np.random.default_rng(2).normal(size=161), dtype=torch.float64)
    (v[0].cpu() * w).sum().backward()
    xo = torch.tensor(cur, dtype=torch.float64, requires_grad=True)
    ref = O.simulate(om, None, DT, stim=xo[None], stim_comps=[0], rec_comps=[13])
    (ref[0] * w).sum().backward()
    _grad_check(x.grad, xo.grad, "stimulus")


def test_gradient_wrt_initial_gates(device):
    cell, om, cur, _, ds = _c0_with_stim(device, T=120)
    cell.branch(0).comp(0).record()
    n = cell._n_nodes
    m0 = torch.full((n,), 0.1, dtype=torch.float64, device=device, requires_grad=True)
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT, all_states={"HH_m": m0})
    v[0, -1].cpu().backward()
    m0o = torch.full((60,), 0.1, dtype=torch.float64, requires_grad=True)
    om.gates0 = (m0o, torch.full((60,), 0.2, dtype=torch.float64), torch.full((60,), 0.2, dtype=torch.float64))
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0])
    ref[0, -1].backward()
    g = m0.grad[torch.as_tensor(cell._comp_node, device=device)]
    _grad_check(g, m0o.grad, "initial HH_m")


def test_cell_larger_than_lds(device):
    """A single cell whose system exceeds one workgroup's LDS in float64 (3000 branches x
    4 compartments, 12,000 compartments; cells are never split between programs): the
    forward takes the per-step path (its tree solve sweeps the levels in HBM); 200 steps
    against the oracle (1e-6 mV)."""
    from models import build_tree, c1_stimulus

    T = 200
    cell, om = build_tree(3000, seed=5)
    cell.branch(0).comp(0).record()
    cell.branch(1500).comp(2).record()
    cur = torch.tensor(c1_stimulus(T) * 5.0, dtype=torch.float64, device=device)
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT)
    rc = int(cell._cumsum_ncomp[1500] + 2)
    ref = O.simulate(om, None, DT, stim=cur.cpu()[None], stim_comps=[0], rec_comps=[0, rc]).reshape(2, T + 1)
    assert v.shape == (2, T + 1)
    err = float((v.cpu() - ref).abs().max())
    print(f"12,000-compartment cell: max err {err:.2e} mV, root range {float(ref[0].min()):.1f}.."
          f"{float(ref[0].max()):.1f} mV")
    assert err < 1e-6
