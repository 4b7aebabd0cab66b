"""Standalone synapse and HH step operators on the GPU against closed forms (the oracle's
restatement of the reference's synapse / channel formulas) in float64, with autograd
gradients for the synapse backward."""
import numpy as np
import pytest
import torch

from jaxhip.kernels.hh_step import hh_step
from jaxhip.kernels.synapse_step import PARAM_NAMES, kernel_plan, regular_layout, synapse_step
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu

DIFF = 1e-3


def _closed_form(model, v, s, pre, post_t, k, prm, area, dt):
    """Per synapse: state update, current (nA) and its finite-difference linearization
    (both voltages perturbed by DIFF, /area * 1e5), summed per target."""
    pre_v = v[:, pre]
    post_v = v[:, post_t.repeat_interleave(k)]
    if model == 0:
        s_new = O.kinetic_update(s, pre_v, prm, dt)
        cur = O.kinetic_current
    else:
        s_new = s
        cur = O.sigmoid_rate_current
    i0 = cur(s_new, pre_v, post_v, prm)
    i1 = cur(s_new, pre_v + DIFF, post_v + DIFF, prm)
    d0, d1 = i0 / area * 1e5, i1 / area * 1e5
    vt = (d1 - d0) / DIFF
    ct = d0 - vt * post_v
    B = v.shape[0]
    return s_new, i0, vt.reshape(B, -1, k).sum(-1), ct.reshape(B, -1, k).sum(-1)


def _problem(model, seed, dtype=torch.float64):
    rng = np.random.default_rng(seed)
    n_nodes, n_tgt, k, B = 40, 9, 4, 3
    targets = np.sort(rng.choice(n_nodes, n_tgt, replace=False))
    post = np.repeat(targets, k)
    pre = rng.integers(0, n_nodes, size=n_tgt * k)
    E = len(pre)
    if model == 0:
        vals = {"gS": rng.uniform(0, 1e-3, E), "e_syn": rng.choice([0.0, -80.0], E),
                "alpha": rng.uniform(0.5, 2, E), "beta": rng.uniform(0.1, 0.5, E),
                "v_th": rng.uniform(-40, -10, E), "delta": rng.uniform(1, 5, E)}
    else:
        vals = {"gS": rng.uniform(0, 1e-3, E), "x_offset": rng.uniform(-60, -30, E),
                "slope": rng.uniform(0.05, 0.5, E)}
    t = lambda a: torch.tensor(a, dtype=dtype)
    v = t(rng.normal(-55, 15, (B, n_nodes)))
    s = t(rng.uniform(0, 1, (B, E)))
    area = t(rng.uniform(50, 500, E))
    return pre, post, targets, k, {n: t(x) for n, x in vals.items()}, v, s, area


@pytest.mark.parametrize("model", [0, 1])
def test_synapse_step_forward_backward(device, model):
    pre, post, targets, k, prm, v, s, area = _problem(model, 10 + model)
    assert regular_layout(post) == k
    plan = kernel_plan(model, pre, post, v.shape[1], device)
    names = PARAM_NAMES[model]
    dt = 0.025
    # Closed form with autograd on the CPU.
    vc, sc = v.clone().requires_grad_(), s.clone().requires_grad_()
    pc = {n: prm[n].clone().requires_grad_() for n in names}
    ref = _closed_form(model, vc, sc, torch.as_tensor(pre), torch.as_tensor(targets), k, pc, area, dt)
    # GPU.
    vg = v.to(device).requires_grad_()
    sg = s.to(device).requires_grad_()
    pg = [prm[n].to(device).requires_grad_() for n in names]
    out = synapse_step(plan, vg, sg, pg, area.to(device), dt, DIFF)
    # s_new and the currents to round-off; the finite-difference slope and offset carry
    # the cancellation of (d1 - d0) / DIFF (relative ~1e-11).
    for name, a, b, rtol in zip(("s_new", "i_syn", "lin", "cst"), out, ref, (1e-12, 1e-12, 1e-8, 1e-8)):
        err = float((a.detach().cpu() - b.detach()).abs().max())
        print(name, "max abs err", err, "scale", float(b.abs().max()))
        assert torch.allclose(a.detach().cpu(), b.detach(), rtol=rtol, atol=1e-10 * float(b.abs().max())), name
    rng = np.random.default_rng(0)
    w = [torch.tensor(rng.normal(size=tuple(o.shape))) for o in ref]
    if model == 1:
        w[0] = torch.zeros_like(w[0])            # no state
    loss_ref = sum((wi * o).sum() for wi, o in zip(w, ref))
    gref = torch.autograd.grad(loss_ref, [vc, sc] + [pc[n] for n in names], allow_unused=True)
    loss = sum((wi.to(device) * o).sum() for wi, o in zip(w, out))
    gout = torch.autograd.grad(loss, [vg, sg] + pg, allow_unused=True)
    for a, b in zip(gout, gref):
        b = torch.zeros_like(a.cpu()) if b is None else b
        assert torch.allclose(a.cpu(), b, rtol=1e-8, atol=1e-10), float((a.cpu() - b).abs().max())


@pytest.mark.parametrize("model", [0, 1])
def test_synapse_step_under_torch_compile(device, model):
    """The synapse operators inside `torch.compile(fullgraph=True)` (aot_eager: traced
    through their fake rules and the registered backward, no code generation) give the
    eager values and gradients."""
    pre, post, targets, k, prm, v, s, area = _problem(model, 20 + model)
    plan = kernel_plan(model, pre, post, v.shape[1], device)
    names = PARAM_NAMES[model]
    ag = area.to(device)
    rng = np.random.default_rng(1)
    w = [torch.tensor(rng.normal(size=sh), device=device)
         for sh in (tuple(s.shape), tuple(s.shape), (v.shape[0], len(targets)), (v.shape[0], len(targets)))]

    def f(v, s, *p):
        return sum((wi * o).sum() for wi, o in zip(w, synapse_step(plan, v, s, list(p), ag, 0.025, DIFF)))

    outs = []
    for fn in (f, torch.compile(f, backend="aot_eager", fullgraph=True)):
        args = [x.to(device).requires_grad_() for x in [v, s] + [prm[n] for n in names]]
        val = fn(*args)
        val.backward()
        outs.append((val.detach(), [a.grad for a in args]))
    torch.testing.assert_close(outs[0][0], outs[1][0], rtol=1e-12, atol=0)
    for ga, gb in zip(outs[0][1], outs[1][1]):
        torch.testing.assert_close(ga, gb, rtol=1e-12, atol=1e-15)


def test_synapse_step_float32(device):
    pre, post, targets, k, prm, v, s, area = _problem(0, 5)
    plan = kernel_plan(0, pre, post, v.shape[1], device)
    names = PARAM_NAMES[0]
    out = synapse_step(plan, v.float().to(device), s.float().to(device),
                       [prm[n].float().to(device) for n in names], area.float().to(device), 0.025)
    ref = _closed_form(0, v, s, torch.as_tensor(pre), torch.as_tensor(targets), k, prm, area, 0.025)
    for a, b in zip(out[:2], ref[:2]):
        assert torch.allclose(a.double().cpu(), b, rtol=1e-5, atol=1e-6)


def test_hh_step(device):
    rng = np.random.default_rng(3)
    B, n_nodes = 4, 30
    inds = torch.tensor(np.sort(rng.choice(n_nodes, 20, replace=False)))
    v = torch.tensor(rng.uniform(-90, 40, (B, n_nodes)))
    v[0, inds[0]] = -40.0                      # x_over_expm1 singular points
    v[1, inds[1]] = -55.0
    m, h, n = (torch.tensor(rng.uniform(0, 1, (B, 20))) for _ in range(3))
    prm = {"HH_gNa": torch.tensor(rng.uniform(0.06, 0.18, (B, 20))), "HH_gK": torch.full((20,), 0.036),
           "HH_gLeak": torch.full((20,), 3e-4), "HH_eNa": torch.full((20,), 50.0),
           "HH_eK": torch.full((20,), -77.0), "HH_eLeak": torch.full((20,), -54.3)}
    dt = 0.025
    out = hh_step(v.to(device), inds.to(device), m.to(device), h.to(device), n.to(device),
                  [prm[k].to(device) for k in ("HH_gNa", "HH_gK", "HH_gLeak", "HH_eNa", "HH_eK", "HH_eLeak")], dt)
    vv = v[:, inds]
    m1, h1, n1 = O.hh_update(m, h, n, vv, dt)
    i0 = O.hh_current(m1, h1, n1, vv, prm)
    i1 = O.hh_current(m1, h1, n1, vv + DIFF, prm)
    lin = (i1 - i0) / DIFF                      # the reference's finite difference
    for a, b in zip(out[:4], (m1, h1, n1, i0)):
        assert torch.allclose(a.cpu(), b, rtol=1e-12, atol=1e-14)
    assert torch.allclose(out[4].cpu(), lin, rtol=1e-9, atol=1e-12)
    assert torch.allclose(out[4].cpu() * vv - out[5].cpu(), i0, rtol=1e-10, atol=1e-12)
