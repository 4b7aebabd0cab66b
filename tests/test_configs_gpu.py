"""The workload shapes of the hot path at their real sizes (not shrunken) against the
oracle: the gradient workload on the 250-branch tree (750 per-branch trainable
conductances, 2000 steps, checkpoints every 100), the 1024-cell network with 32768
KineticSynapses (in-degree 32) driven to spike, and served requests that include the
largest (~2000-compartment) trees."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from jaxhip.serve import SimServer
from models import (DT, build_c4, build_tree, c1_stimulus, c2_request, c2_specs, c4_stimulus,
                    oracle_overrides)
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu

NBRANCH, SEED = 250, 11
COND = ("HH_gNa", "HH_gK", "HH_gLeak")


def spike_counts(v):
    v = np.asarray(v)
    return ((v[..., 1:] >= 0.0) & (v[..., :-1] < 0.0)).sum(-1)


def test_gradient_workload_real_shape(device):
    """750 per-branch trainable gNa / gK / gLeak per simulation on the 250-branch tree
    (B = 2, T = 2000, checkpoints every 100 steps, float64), MSE against target traces
    from seeded ground-truth parameters: the hand-written adjoint against central finite
    differences (h = 1e-6 relative) on 20 random parameters and against autograd through
    the oracle on every parameter."""
    B, T, K = 2, 2000, 100
    cell, om = build_tree(NBRANCH, SEED)
    cell.branch(0).comp(0).record()
    for k in COND:
        cell.branch("all").make_trainable(k, verbose=False)
    rng = np.random.default_rng(21)
    base = np.array([0.12, 0.036, 0.0003])[None, :, None]
    true = base * rng.uniform(0.5, 1.5, size=(B, 3, NBRANCH))
    guess = base * rng.uniform(0.5, 1.5, size=(B, 3, NBRANCH))
    cur = torch.tensor(c1_stimulus(T), dtype=torch.float64, device=device)
    ds = cell.branch(0).comp(0).data_stimulate(cur)

    def params(vals, grad):
        return [{k: torch.tensor(vals[:, j], dtype=torch.float64, device=device, requires_grad=grad)}
                for j, k in enumerate(COND)]

    with torch.no_grad():
        target = jx.integrate(cell, params(true, False), data_stimuli=ds, delta_t=DT)[:, 0]

    def loss_of(vals, grad=False):
        ps = params(vals, grad)
        v = jx.integrate(cell, ps, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[T // K, K])
        return ((v[:, 0] - target) ** 2).mean(), ps

    loss, ps = loss_of(guess, True)
    loss.backward()
    grads = np.stack([p[k].grad.cpu().numpy() for p, k in zip(ps, COND)], 1)      # (B, 3, nb)
    # Oracle autograd on the same loss.
    ov = torch.tensor(guess, dtype=torch.float64, requires_grad=True)
    comp = torch.repeat_interleave(ov, 4, dim=-1)
    ref = O.simulate(om, None, DT, stim=cur.cpu()[None], stim_comps=[0], rec_comps=[0],
                     overrides={k: comp[:, j] for j, k in enumerate(COND)})
    oloss = ((ref[:, 0] - target.cpu()) ** 2).mean()
    oloss.backward()
    g_o = ov.grad.numpy()
    rel = np.abs(grads - g_o).max() / np.abs(g_o).max()
    print(f"loss {loss.item():.6e} vs oracle {oloss.item():.6e}; adjoint vs oracle autograd: max error "
          f"relative to the largest gradient {rel:.2e}")
    np.testing.assert_allclose(loss.item(), oloss.item(), rtol=1e-8)
    assert rel < 1e-6
    h = 1e-6
    with torch.no_grad():
        for _ in range(20):
            b, j, i = int(rng.integers(B)), int(rng.integers(3)), int(rng.integers(NBRANCH))
            step = h * guess[b, j, i]
            up, dn = guess.copy(), guess.copy()
            up[b, j, i] += step
            dn[b, j, i] -= step
            fd = (loss_of(up)[0].item() - loss_of(dn)[0].item()) / (2 * step)
            print(f"fd ({b},{j},{i}) {fd:.6e} adjoint {grads[b, j, i]:.6e}")
            assert abs(fd - grads[b, j, i]) <= 1e-4 * abs(grads[b, j, i]) + 1e-6 * np.abs(grads).max()


def test_network_real_shape_spiking(device):
    """1024 ball-and-stick cells, 32 KineticSynapses per cell (32768), one simulation of
    400 steps with current pulses strong enough to make the somata spike (the synapses
    open): the fused float32 network step ("mixed") against the float64 oracle on the
    same float32 inputs; spike counts identical."""
    from jaxhip.integrate import set_float32_mode

    n_cells, k, T = 1024, 32, 400
    net, om, info = build_c4(n_cells, k, seed=9)
    soma = info["soma_comps"]
    for c in soma[::8]:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    cur = c4_stimulus(1, n_cells, T, seed=10, amp=1.5, p=0.1).astype(np.float32)
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(
            torch.tensor(cur[0, i], device=device), ds)
    set_float32_mode("mixed")
    v = jx.integrate(net, data_stimuli=ds, delta_t=DT)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur.astype(np.float64)), stim_comps=list(soma),
                     rec_comps=list(soma[::8]))[0]
    err = (v.double().cpu() - ref).abs().amax(-1)
    sc = spike_counts(ref.numpy())
    print(f"network: max err {float(err.max()):.3e} mV, spikes per recorded soma {sc.mean():.2f}")
    assert sc.sum() > 100
    assert np.array_equal(spike_counts(v.double().cpu().numpy()), sc)
    assert float(err.max()) < 1e-2


def test_served_with_largest_trees(device):
    """Served requests with trees up to 2000 compartments (the workload's largest), run
    together, against the oracle (float64, 1e-6 mV) and against each run alone."""
    specs = c2_specs(64, seed=7)
    specs = sorted(specs, key=lambda s: -s["nbranch"])
    pick = [specs[0], specs[1], specs[20], specs[-1]]
    for s in pick:
        s["t_max"] = 8.0
    assert pick[0]["nbranch"] >= 350
    srv = SimServer(device, dtype=torch.float64, delta_t=DT, hbm_budget=512 << 20, max_steps=4001,
                    max_stim_entries=64, chunk_steps=100)
    hs = [srv.submit(c2_request(s, device, torch.float64)) for s in pick]
    srv.run_until_idle()
    srv.close()
    for s, h in zip(pick, hs):
        _, om = build_tree(s["nbranch"], s["seed"])
        T = int(s["t_max"] // DT + 1)
        ref = O.simulate(om, None, DT, stim=torch.tensor(c1_stimulus(T))[None], stim_comps=[0],
                         rec_comps=[0], overrides=oracle_overrides(om, s["params"]))
        err = float((h.result.cpu() - ref[0]).abs().max())
        print(f"{4 * s['nbranch']} compartments: served vs oracle {err:.2e}")
        assert err < 1e-6
    alone = SimServer(device, dtype=torch.float64, delta_t=DT, hbm_budget=512 << 20, max_steps=4001,
                      max_stim_entries=64, chunk_steps=100)
    h0 = alone.submit(c2_request(pick[0], device, torch.float64))
    alone.run_until_idle()
    alone.close()
    assert torch.equal(h0.result, hs[0].result)
