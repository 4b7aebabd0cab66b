"""`solver="crank_nicolson"` on the GPU (modules/base.py:3327-3366 of jaxley-pytorch: the
implicit solve with dt / 2, then v' = 2 x - v; channels and synapses advance by dt)
against the oracle's Crank-Nicolson, whose closed form, order of convergence and limit
are pinned in test_oracle.py.

Tolerances (voltages, mV): 1e-6 in float64, 1e-2 in float32. Gradients: 1e-6 relative
to autograd through the oracle."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from models import (DT, build_tree, c1_params, c1_stimulus, data_set_params,
                    oracle_overrides)
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu

CN = "crank_nicolson"
SHAPES = ["8,1024,gcoef", "8,1024,sym", "4,512,stream", "2,256", "1,1024"]


def spike_counts(v: np.ndarray) -> np.ndarray:
    return ((v[..., 1:] >= 0.0) & (v[..., :-1] < 0.0)).sum(-1)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-6), (torch.float32, 1e-2)])
def test_crank_nicolson_launch_shapes_match_oracle(device, dtype, tol, monkeypatch):
    """Streaming (coefficients from L2 or LDS) and register-resident kernels, batched
    per-simulation conductances, B = 5 (padding simulations in the last program)."""
    from jaxhip import autotune

    B, T, nb = 5, 300, 30
    cell, om = build_tree(nb, seed=4)
    cell.branch(0).comp(0).record()
    cell.branch(nb - 1).comp(3).record()
    prm = c1_params(B, seed=4)
    ps = data_set_params(cell, prm, device, dtype)
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=dtype, device=device))
    kw = dict(stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0, 4 * nb - 1],
              overrides=oracle_overrides(om, prm))
    ref = O.simulate(om, None, DT, solver=CN, **kw).numpy()
    ref_be = O.simulate(om, None, DT, **kw).numpy()
    assert np.abs(ref - ref_be).max() > 1e-2      # the two steppers differ on this input
    for shape in SHAPES:
        monkeypatch.setenv("JAXHIP_FUSED_CONFIG", shape)
        autotune._CACHE.clear()
        print("shape", shape, dtype, flush=True)
        v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT, solver=CN)
        v = v.double().cpu().numpy()
        err = np.abs(v - ref).max()
        print(f"crank-nicolson shape {shape} {dtype} max err {err:.3e}", flush=True)
        assert v.shape == (B, 2, T + 1)
        assert err < tol, shape
        assert np.array_equal(spike_counts(v), spike_counts(ref)), shape
    autotune._CACHE.clear()


def _simple_net():
    cell = jx.Cell(jx.Branch(jx.Compartment(), 4), parents=[-1, 0, 0])
    net = jx.Network([cell, cell])
    net.insert(jx.HH())
    jx.connect(net.cell(0).branch(0).comp(0), net.cell(1).branch(0).comp(0), jx.IonotropicSynapse())
    net.IonotropicSynapse.set("IonotropicSynapse_gS", 0.01)
    net.set("length", 50.0)
    net.set("axial_resistivity", 200.0)
    net.cell(1).branch(0).comp(3).record()
    return net


def test_crank_nicolson_network_matches_oracle(device):
    """The network of the reference's solver comparison (tests/test_solver.py:29-58):
    both steppers against the oracle in float64, and within 6 mV of each other as the
    reference asserts."""
    net = _simple_net()
    ocur = O.step_current(0.5, 1.0, 0.1, 0.025, 5.0)
    ds = net.cell(0).branch(0).comp(0).data_stimulate(ocur.to(device))
    m = O.OracleModel([([-1, 0, 0], [4, 4, 4])] * 2)
    m.hh_comps = np.arange(24)
    m.length[:] = 50.0
    m.axial_resistivity[:] = 200.0
    p = dict(O.SYN_DEFAULTS["IonotropicSynapse"], gS=0.01)
    m.synapses = [{"type": "IonotropicSynapse", "pre": np.array([0]), "post": np.array([13]),
                   "params": {k: np.array([x]) for k, x in p.items()}, "s0": np.array([0.2])}]
    out = {}
    for s in ("bwd_euler", CN):
        v = jx.integrate(net, data_stimuli=ds, delta_t=DT, solver=s).cpu().numpy()
        ref = O.simulate(m, None, DT, stim=ocur[None], stim_comps=[0], rec_comps=[15], solver=s).numpy()
        err = np.abs(v - ref).max()
        print(f"network {s} max err vs oracle {err:.3e}", flush=True)
        assert v.shape == ref.shape
        assert err < 1e-6
        out[s] = v
    assert np.abs(out["bwd_euler"] - out[CN]).max() < 6.0


def test_crank_nicolson_gradients_vs_oracle_autograd(device):
    """Under autograd Crank-Nicolson steps run through the per-step path (tree-solve
    operator and its backward kernel, checkpoints every 40 steps); gradients of a loss
    on the recorded voltage w.r.t. per-branch gNa and gK against autograd through the
    oracle."""
    nb, B, T = 8, 2, 120
    cell, om = build_tree(nb, seed=7)
    cell.branch(0).comp(0).record()
    cell.branch("all").make_trainable("HH_gNa", verbose=False)
    cell.branch("all").make_trainable("HH_gK", verbose=False)
    rng = np.random.default_rng(2)
    base = np.array([0.12, 0.036])
    vals = base[None, :, None] * rng.uniform(0.7, 1.3, size=(B, 2, nb))
    params = [{k: torch.tensor(vals[:, j], dtype=torch.float64, device=device, requires_grad=True)}
              for j, k in enumerate(("HH_gNa", "HH_gK"))]
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    v = jx.integrate(cell, params, data_stimuli=ds, delta_t=DT, solver=CN, checkpoint_lengths=[3, 40])
    target = torch.linspace(-70, -60, T + 1, dtype=torch.float64, device=device)
    loss = ((v[:, 0] - target) ** 2).mean()
    loss.backward()
    ovals = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    comp = torch.repeat_interleave(ovals, 4, dim=-1)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0],
                     overrides={"HH_gNa": comp[:, 0], "HH_gK": comp[:, 1]}, solver=CN)
    oloss = ((ref[:, 0] - target.cpu()) ** 2).mean()
    oloss.backward()
    print("crank-nicolson loss", loss.item(), oloss.item(), flush=True)
    np.testing.assert_allclose(loss.item(), oloss.item(), rtol=1e-9)
    for j, p in enumerate(params):
        g = list(p.values())[0].grad.cpu().numpy()
        np.testing.assert_allclose(g, ovals.grad[:, j].numpy(), rtol=1e-6, atol=1e-9)


def test_crank_nicolson_step_loop_matches_integrate(device):
    """`build_init_and_step_fn(solver="crank_nicolson")` stepped T times equals
    `integrate(solver="crank_nicolson")`."""
    cell, _ = build_tree(10, seed=5)
    cell.branch(0).comp(0).record()
    T = 150
    cur = torch.tensor(c1_stimulus(T), dtype=torch.float64, device=device)
    init_fn, step_fn = jx.build_init_and_step_fn(cell, solver=CN)
    gk = torch.tensor([[0.04]], dtype=torch.float64, device=device)
    states, all_params = init_fn([], param_state=cell.data_set("HH_gK", gk, None))
    rec_node = int(cell._comp_node[0])
    rec = [states["v"][..., rec_node]]
    for t in range(T):
        states = step_fn(states, all_params, {"i": cur[t:t + 1]}, {"i": np.array([rec_node])}, delta_t=DT)
        rec.append(states["v"][..., rec_node])
    v_step = torch.stack([r.reshape(-1)[0] for r in rec])
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    v_int = jx.integrate(cell, param_state=cell.data_set("HH_gK", gk, None), data_stimuli=ds,
                         delta_t=DT, solver=CN).reshape(T + 1)
    err = float((v_step - v_int).abs().max())
    print("crank-nicolson step loop vs integrate", err, flush=True)
    assert err < 1e-9
