"""The reverse pass with the forward's voltage trajectory (gates recomputed between
checkpoints, no recomputed tree solves) against the reverse pass that recomputes the
voltages, and both against autograd through the oracle."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from models import DT, build_tree, c1_stimulus
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu


def _grads(device, dtype, B, T, K, nb=12, seed=7):
    cell, om = build_tree(nb, seed=seed)
    cell.branch(0).comp(0).record()
    for k in ("HH_gNa", "HH_gK", "HH_gLeak"):
        cell.branch("all").make_trainable(k, verbose=False)
    rng = np.random.default_rng(1)
    vals = np.array([0.12, 0.036, 0.0003])[None, :, None] * rng.uniform(0.7, 1.3, size=(B, 3, nb))
    params = [{k: torch.tensor(vals[:, j], dtype=dtype, device=device, requires_grad=True)}
              for j, k in enumerate(("HH_gNa", "HH_gK", "HH_gLeak"))]
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=dtype, device=device))
    v = jx.integrate(cell, params, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[T // K, K])
    target = torch.linspace(-70, -60, T + 1, dtype=torch.float64, device=device)
    loss = ((v[:, 0].double() - target) ** 2).mean()
    loss.backward()
    g = np.stack([list(p.values())[0].grad.double().cpu().numpy() for p in params], 1)
    return loss.item(), g, (om, vals, cur, target)


@pytest.fixture
def streaming_only(monkeypatch):
    """Launch shapes of the streaming kernel only (the one that keeps a trajectory), and
    a record of the trajectory decisions."""
    import jaxhip.autotune as A
    import jaxhip.fused as F

    monkeypatch.setattr(F, "items_per_thread", lambda *a, **k: 0)
    monkeypatch.setattr(A, "_CACHE", {})
    kept = []
    orig = F.keep_trajectory

    def spy(nbytes, dev):
        r = orig(nbytes, dev)
        kept.append(r)
        return r
    monkeypatch.setattr(F, "keep_trajectory", spy)
    return kept


@pytest.mark.parametrize("B,K", [(2, 50), (20, 40)])
def test_trajectory_reverse_pass_matches_recompute_and_oracle(device, streaming_only, monkeypatch, B, K):
    import jaxhip.fused as F

    T = 200
    loss_t, g_t, (om, vals, cur, target) = _grads(device, torch.float64, B, T, K)
    assert streaming_only and all(streaming_only), "the trajectory was not kept"
    monkeypatch.setattr(F, "TRAJ_FRACTION", 0.0)
    streaming_only.clear()
    loss_r, g_r, _ = _grads(device, torch.float64, B, T, K)
    assert streaming_only and not any(streaming_only)
    rel = np.abs(g_t - g_r).max() / np.abs(g_r).max()
    print(f"B={B} K={K}: trajectory vs recompute, max gradient difference {rel:.3e} (relative to the largest)")
    assert loss_t == pytest.approx(loss_r, rel=1e-12)
    assert rel < 1e-9
    # Oracle autograd (per-branch conductances, every compartment of a branch).
    ovals = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    comp = torch.repeat_interleave(ovals, 4, dim=-1)
    ov = {"HH_gNa": comp[:, 0], "HH_gK": comp[:, 1], "HH_gLeak": comp[:, 2]}
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0], overrides=ov)
    oloss = ((ref[:, 0] - target.cpu()) ** 2).mean()
    oloss.backward()
    np.testing.assert_allclose(loss_t, oloss.item(), rtol=1e-9)
    np.testing.assert_allclose(g_t, ovals.grad.numpy(), rtol=1e-6, atol=1e-9)


def test_trajectory_reverse_pass_float32(device, streaming_only, monkeypatch):
    """float32 ("fast" mode): the trajectory path against the recomputing one."""
    import jaxhip.fused as F
    from jaxhip.integrate import set_float32_mode

    set_float32_mode("fast")
    try:
        _, g_t, _ = _grads(device, torch.float32, 16, 200, 50)
        assert streaming_only and all(streaming_only)
        monkeypatch.setattr(F, "TRAJ_FRACTION", 0.0)
        _, g_r, _ = _grads(device, torch.float32, 16, 200, 50)
    finally:
        set_float32_mode("mixed")
    rel = np.abs(g_t - g_r).max() / np.abs(g_r).max()
    print(f"float32: trajectory vs recompute, max gradient difference {rel:.3e}")
    assert rel < 1e-3


def _net_grads(device, synapse, B, T=200, K=40, n_cells=16, k=4, seed=8):
    from models import build_c4, c4_stimulus

    net, om, info = build_c4(n_cells, k, seed, synapse)
    soma = info["soma_comps"]
    for c in soma:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    cur = c4_stimulus(B, n_cells, T, seed + 1) * 4.0
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(
            torch.tensor(cur[:, i, :], dtype=torch.float64, device=device), ds)
    gs = torch.tensor(np.asarray(om.synapses[0]["params"]["gS"]) * 20.0, dtype=torch.float64, device=device,
                      requires_grad=True)
    gna = torch.full((B, 1), 0.12, dtype=torch.float64, device=device, requires_grad=True)
    ps = net.data_set(f"{synapse}_gS", gs, None)
    ps = net.data_set("HH_gNa", gna, ps)
    w = torch.tensor(np.random.default_rng(seed + 2).normal(size=(B, n_cells, T + 1)), dtype=torch.float64)
    v = jx.integrate(net, param_state=ps, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[T // K, K])
    (v.cpu() * w).sum().backward()
    return gs.grad.cpu().numpy(), gna.grad.cpu().numpy()


@pytest.mark.parametrize("synapse", ["KineticSynapse", "SigmoidRateSynapse"])
@pytest.mark.parametrize("B,K", [(2, 40), (2, 25), (300, 40)])
def test_network_trajectory_reverse_pass(device, streaming_only, monkeypatch, synapse, B, K):
    """Networks: the reverse pass from the trajectory (synapse states and gates
    recomputed per synapse / item, no tree solves) against the recomputing one, split
    into parts that run in lockstep (B = 2; K = 25: segments of odd length) and not
    split (B = 300)."""
    import jaxhip.fused as F

    a1, a2 = _net_grads(device, synapse, B, K=K)
    assert streaming_only and all(streaming_only), "the trajectory was not kept"
    monkeypatch.setattr(F, "TRAJ_FRACTION", 0.0)
    b1, b2 = _net_grads(device, synapse, B, K=K)
    e1 = np.abs(a1 - b1).max() / np.abs(b1).max()
    e2 = np.abs(a2 - b2).max() / np.abs(b2).max()
    print(f"{synapse} B={B} K={K}: trajectory vs recompute, gS {e1:.2e}, gNa {e2:.2e}")
    assert e1 < 1e-9 and e2 < 1e-9


def test_trajectory_register_resident_forward(device, monkeypatch):
    """The register-resident forward keeps the trajectory too (the launch shapes the
    autotuner picks for small batches): against the recomputing reverse pass."""
    import jaxhip.autotune as A
    import jaxhip.fused as F

    monkeypatch.setattr(A, "_CACHE", {})
    shapes = []
    orig = F.ForwardLaunch.__init__

    def spy(self, *args, **kw):
        orig(self, *args, **kw)
        shapes.append((self.args.items_per_thread, self.traj is not None))
    monkeypatch.setattr(F.ForwardLaunch, "__init__", spy)
    _, g_t, _ = _grads(device, torch.float64, 2, 200, 50)
    assert any(ipt > 0 and kept for ipt, kept in shapes), shapes
    monkeypatch.setattr(F, "TRAJ_FRACTION", 0.0)
    _, g_r, _ = _grads(device, torch.float64, 2, 200, 50)
    rel = np.abs(g_t - g_r).max() / np.abs(g_r).max()
    print(f"register-resident forward: trajectory vs recompute {rel:.3e}; shapes {shapes}")
    assert rel < 1e-9
