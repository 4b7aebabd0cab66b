"""GPU parity: the HIP path against the CPU oracle on the same seeded inputs.

Tolerances (voltages, mV): 1e-6 in float64, 1e-2 in float32 (max abs error over every
recorded sample). Tree solves: 1e-10 relative to the dense solve in float64.
"""
import numpy as np
import pytest
import torch

import jaxhip as jx
from jaxhip.kernels.tree_solve import make_tree_solve_kernel
from oracle import jx_oracle as O

from models import (DT, build_c0, build_tree, c0_stimulus, c1_params, c1_stimulus,
                    data_set_params, oracle_overrides)

pytestmark = pytest.mark.gpu


def _run_c0(device, dtype, T=400):
    cell, om = build_c0()
    cur = c0_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=dtype, device=device))
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0],
                     rec_comps=list(range(60)))
    return v.double().cpu().numpy(), ref.numpy()


def spike_counts(v: np.ndarray) -> np.ndarray:
    """Upward crossings of 0 mV along the last axis (training/spiking_network.py:509-513)."""
    return ((v[..., 1:] >= 0.0) & (v[..., :-1] < 0.0)).sum(-1)


def test_c0_float64(device):
    v, ref = _run_c0(device, torch.float64)
    assert v.shape == (60, 401)
    err = np.abs(v - ref).max()
    print("c0 fp64 max err", err)
    assert err < 1e-6
    assert np.array_equal(spike_counts(v), spike_counts(ref))


def test_c0_float32(device):
    v, ref = _run_c0(device, torch.float32)
    err = np.abs(v - ref).max()
    print("c0 fp32 max err", err)
    assert err < 1e-2
    assert np.array_equal(spike_counts(v), spike_counts(ref))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-6), (torch.float32, 1e-2)])
def test_batched_random_tree(device, dtype, tol):
    B, T, nb = 6, 600, 40
    cell, om = build_tree(nb, seed=3)
    cell.branch(0).comp(0).record()
    prm = c1_params(B, seed=3)
    ps = data_set_params(cell, prm, device, dtype)
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=dtype, device=device))
    v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0],
                     overrides=oracle_overrides(om, prm))
    err = np.abs(v.double().cpu().numpy() - ref.numpy()).max()
    print(f"tree {dtype} max err", err)
    assert v.shape == (B, 1, T + 1)
    assert err < tol
    assert np.array_equal(spike_counts(v.double().cpu().numpy()), spike_counts(ref.numpy()))
    assert spike_counts(ref.numpy()).sum() > 0


SHAPES = ["8,1024,gcoef", "8,1024,sym", "4,512,gcoef", "4,512,sym", "4,512,stream", "2,256", "1,1024", "8,512"]


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-6), (torch.float32, 1e-2)])
def test_launch_shapes_match_oracle(device, dtype, tol, monkeypatch):
    """Kernels and launch shapes of the fused forward (streaming with the coupling
    coefficients in LDS, read from L2 or the symmetric system's in LDS, register-resident;
    B = 5 leaves padding simulations in the last program) against the oracle."""
    from jaxhip import autotune

    B, T, nb = 5, 300, 30
    cell, om = build_tree(nb, seed=4)
    cell.branch(0).comp(0).record()
    cell.branch(nb - 1).comp(3).record()
    prm = c1_params(B, seed=4)
    ps = data_set_params(cell, prm, device, dtype)
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=dtype, device=device))
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0],
                     rec_comps=[0, 4 * nb - 1], overrides=oracle_overrides(om, prm)).numpy()
    for shape in SHAPES:
        monkeypatch.setenv("JAXHIP_FUSED_CONFIG", shape)
        autotune._CACHE.clear()
        print("shape", shape, dtype, flush=True)
        v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT)
        err = np.abs(v.double().cpu().numpy() - ref).max()
        print(f"shape {shape} {dtype} max err {err:.3e}", flush=True)
        assert v.shape == (B, 2, T + 1)
        assert err < tol, shape
        assert np.array_equal(spike_counts(v.double().cpu().numpy()), spike_counts(ref)), shape
    autotune._CACHE.clear()


def _dense_solve(d, l, u, s, parent):
    n = len(d)
    A = np.diag(d.astype(np.float64))
    for c in range(n):
        p = parent[c]
        if p >= 0:
            A[c, p] += l[c]
            A[p, c] += u[c]
    return np.linalg.solve(A, s)


@pytest.mark.parametrize("mode", ["registers", "levels"])
@pytest.mark.parametrize("which", ["c0", "tree"])
def test_tree_solve_vs_dense(device, mode, which):
    cell, _ = build_c0() if which == "c0" else build_tree(60, seed=5)
    idx = cell._dhs_solve_indexer
    parent = np.asarray(idx["parent_lookup"])  # includes the spurious node
    kern = make_tree_solve_kernel(parent, device, mode)
    assert kern is not None
    n = len(parent)
    rng = np.random.default_rng(0)
    rows = 5
    l = -rng.uniform(0.1, 1.0, (rows, n))
    u = -rng.uniform(0.1, 1.0, (rows, n))
    d = 4.0 + rng.uniform(0, 1, (rows, n))
    s = rng.normal(size=(rows, n))
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=device)
    x = kern.solve(t(d), t(l), t(u), t(s)).cpu().numpy()
    for r in range(rows):
        ref = _dense_solve(d[r], l[r], u[r], s[r], parent)
        np.testing.assert_allclose(x[r], ref, rtol=1e-10, atol=1e-12)
    # Gradient: lambda = A^-T g through the custom backward.
    D, L, U, Sv = (t(a).requires_grad_(True) for a in (d, l, u, s))
    w = t(rng.normal(size=(rows, n)))
    (kern.fused(D, L, U, Sv) * w).sum().backward()
    for r in range(rows):
        A = np.diag(d[r])
        for c in range(n):
            p = parent[c]
            if p >= 0:
                A[c, p] += l[r, c]
                A[p, c] += u[r, c]
        lam = np.linalg.solve(A.T, w[r].cpu().numpy())
        np.testing.assert_allclose(Sv.grad[r].cpu().numpy(), lam, rtol=1e-9, atol=1e-12)
        xr = np.linalg.solve(A, s[r])
        np.testing.assert_allclose(D.grad[r].cpu().numpy(), -lam * xr, rtol=1e-9, atol=1e-12)


def test_adjoint_matches_oracle_autograd(device):
    """Gradients of a loss on the recorded voltage w.r.t. per-branch gNa / gK / gLeak
    (fp64) against autograd through the oracle, checkpoints every 50 steps."""
    nb, B, T = 12, 2, 200
    cell, om = build_tree(nb, seed=7)
    cell.branch(0).comp(0).record()
    cell.branch("all").make_trainable("HH_gNa", verbose=False)
    cell.branch("all").make_trainable("HH_gK", verbose=False)
    cell.branch("all").make_trainable("HH_gLeak", verbose=False)
    rng = np.random.default_rng(1)
    base = np.array([0.12, 0.036, 0.0003])
    vals = base[None, :, None] * rng.uniform(0.7, 1.3, size=(B, 3, nb))
    params = [{k: torch.tensor(vals[:, j], dtype=torch.float64, device=device, requires_grad=True)}
              for j, k in enumerate(("HH_gNa", "HH_gK", "HH_gLeak"))]
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    v = jx.integrate(cell, params, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[4, 50])
    target = torch.linspace(-70, -60, T + 1, dtype=torch.float64, device=device)
    loss = ((v[:, 0] - target) ** 2).mean()
    loss.backward()
    # Oracle autograd.
    ovals = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    comp = torch.repeat_interleave(ovals, 4, dim=-1)  # (B, 3, nb*4)
    ov = {"HH_gNa": comp[:, 0], "HH_gK": comp[:, 1], "HH_gLeak": comp[:, 2]}
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0],
                     overrides=ov)
    oloss = ((ref[:, 0] - target.cpu()) ** 2).mean()
    oloss.backward()
    np.testing.assert_allclose(loss.item(), oloss.item(), rtol=1e-9)
    for j, p in enumerate(params):
        g = list(p.values())[0].grad
        np.testing.assert_allclose(g.cpu().numpy(), ovals.grad[:, j].numpy(), rtol=1e-6, atol=1e-9)


def _c3_loss(cell, params_vals, ds, target, dtype, device, K=50):
    params = [{k: torch.tensor(params_vals[:, j], dtype=dtype, device=device, requires_grad=True)}
              for j, k in enumerate(("HH_gNa", "HH_gK", "HH_gLeak"))]
    v = jx.integrate(cell, params, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[4, K])
    loss = ((v[:, 0].double() - target) ** 2).mean()
    return loss, params


def test_adjoint_vs_central_differences(device):
    """Per-branch conductance gradients (fp64 adjoint, checkpoints every 50 steps) against
    central finite differences with h = 1e-6 on 20 random parameters (the C3 check); the float32 adjoint
    against the float64 one."""
    nb, B, T = 12, 2, 200
    cell, _ = build_tree(nb, seed=9)
    cell.branch(0).comp(0).record()
    for k in ("HH_gNa", "HH_gK", "HH_gLeak"):
        cell.branch("all").make_trainable(k, verbose=False)
    rng = np.random.default_rng(2)
    vals = np.array([0.12, 0.036, 0.0003])[None, :, None] * rng.uniform(0.7, 1.3, size=(B, 3, nb))
    cur = torch.tensor(c1_stimulus(T), dtype=torch.float64, device=device)
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    target = torch.linspace(-70, -60, T + 1, dtype=torch.float64, device=device)
    loss, params = _c3_loss(cell, vals, ds, target, torch.float64, device)
    loss.backward()
    grads = np.stack([list(p.values())[0].grad.cpu().numpy() for p in params], 1)   # (B, 3, nb)
    h = 1e-6
    with torch.no_grad():
        for _ in range(20):
            b, j, i = int(rng.integers(B)), int(rng.integers(3)), int(rng.integers(nb))
            scale = vals[b, j, i]
            up, dn = vals.copy(), vals.copy()
            up[b, j, i] += h * scale
            dn[b, j, i] -= h * scale
            fd = (_c3_loss(cell, up, ds, target, torch.float64, device)[0].item()
                  - _c3_loss(cell, dn, ds, target, torch.float64, device)[0].item()) / (2 * h * scale)
            print("fd", b, j, i, fd, grads[b, j, i])
            assert abs(fd - grads[b, j, i]) <= 1e-4 * abs(grads[b, j, i]) + 1e-7
    ds32 = cell.branch(0).comp(0).data_stimulate(cur.float())
    loss32, p32 = _c3_loss(cell, vals, ds32, target, torch.float32, device)
    loss32.backward()
    g32 = np.stack([list(p.values())[0].grad.double().cpu().numpy() for p in p32], 1)
    rel = np.abs(g32 - grads).max() / np.abs(grads).max()
    print("fp32 vs fp64 gradient, max error relative to the largest gradient", rel)
    assert rel < 1e-2


@pytest.mark.parametrize("shape", ["8,1024,gcoef", "8,1024,sym", "2,256"])
def test_repeated_launches_bitwise_identical(device, shape, monkeypatch):
    """Deterministic like the reference's kernels (no atomics): the same simulation run
    three times gives bitwise identical traces, for the streaming and the register
    kernel (the node tables are summed in a fixed order on the host)."""
    from jaxhip import autotune

    B, T = 64, 600
    cell, _ = build_tree(40, seed=3)
    cell.branch(0).comp(0).record()
    prm = c1_params(B, seed=3)
    ps = data_set_params(cell, prm, device, torch.float64)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(c1_stimulus(T), dtype=torch.float64, device=device))
    monkeypatch.setenv("JAXHIP_FUSED_CONFIG", shape)
    autotune._CACHE.clear()
    runs = [jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT) for _ in range(3)]
    autotune._CACHE.clear()
    assert torch.equal(runs[0], runs[1]) and torch.equal(runs[0], runs[2])


def test_adjoint_per_simulation_conductances(device):
    """Per-simulation trainable gNa / gK / gLeak (`data_set` of shape (B, 1), the batched
    sweep's parameters: the reverse pass sums their gradients per simulation in
    registers) against autograd through the oracle, B = 19 (several programs and
    padding simulations)."""
    B, T, nb = 19, 300, 20
    cell, om = build_tree(nb, seed=13)
    cell.branch(0).comp(0).record()
    cell.branch(3).comp(2).record()
    prm = c1_params(B, seed=13)
    leaves = {k: torch.tensor(v, dtype=torch.float64, device=device, requires_grad=True) for k, v in prm.items()}
    ps = None
    for k, v in leaves.items():
        ps = cell.data_set(k, v, ps)
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[3, 100])
    w = torch.tensor(np.random.default_rng(5).normal(size=(B, 2, T + 1)), dtype=torch.float64)
    (v.cpu() * w).sum().backward()
    ol = {k: torch.tensor(x, dtype=torch.float64, requires_grad=True) for k, x in prm.items()}
    nc = 4 * nb
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0, 14],
                     overrides={k: x.expand(B, nc) for k, x in ol.items()})
    (ref * w).sum().backward()
    assert float((v.detach().cpu() - ref.detach()).abs().max()) < 1e-6
    for k in prm:
        g, g_ref = leaves[k].grad.cpu(), ol[k].grad
        rel = float((g - g_ref).abs().max() / g_ref.abs().max())
        print(k, "relative gradient error", rel)
        assert rel < 1e-8
