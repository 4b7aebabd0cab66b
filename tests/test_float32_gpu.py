"""The float32 modes of the fused kernels ("mixed": float64 gates and voltages with a
float32 tree solve, `jxhip_fused_forward_mixed`; "fast": float32 gates, rates and solve,
`jxhip_fused_forward_f32` / `jxhip_fused_adjoint_f32`) on the real workload shapes.

What float32 can reach is bounded by the problem's conditioning: a simulation near a
spike threshold moves by up to 0.023 mV (4096 simulations of the 1000-compartment sweep)
when only its inputs are rounded to float32 and everything is computed in float64. The
float32 results are therefore compared with float64 on the same float32 inputs, and a
simulation counts as well conditioned when rounding its inputs moves its float64 trace
by at most 3e-3 mV. Measured on the 4096-simulation sweep: mixed <= 8.5e-3 mV on every
well-conditioned simulation (max 0.038 mV overall), fast <= 0.12 mV (0.21 overall); spike
counts identical in both.
"""
import numpy as np
import pytest
import torch

import jaxhip as jx
from models import (DT, build_c0, c0_stimulus, c1_params, c1_stimulus, data_set_params,
                    oracle_overrides, random_tree)
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu

NBRANCH, SEED = 250, 11


@pytest.fixture(params=["mixed", "fast"])
def f32_mode(request):
    from jaxhip.integrate import set_float32_mode

    set_float32_mode(request.param)
    yield request.param
    set_float32_mode("mixed")


def spike_counts(v):
    v = np.asarray(v)
    return ((v[..., 1:] >= 0.0) & (v[..., :-1] < 0.0)).sum(-1)


def _c1_cell(round32: bool):
    """The 250-branch tree of the batched sweep; geometry optionally rounded to float32
    (the values a float32 simulation computes with)."""
    parents, blen, rad = random_tree(NBRANCH, SEED)
    length, radius = np.repeat(blen / 4, 4), np.repeat(rad, 4)
    if round32:
        length, radius = (x.astype(np.float32).astype(np.float64) for x in (length, radius))
    cell = jx.Cell(jx.Branch(jx.Compartment(), 4), parents)
    cell.set("length", length)
    cell.set("radius", radius)
    cell.insert(jx.HH())
    cell.branch(0).comp(0).record()
    om = O.OracleModel([(parents, [4] * NBRANCH)])
    om.length, om.radius = length, radius
    om.hh_comps = np.arange(4 * NBRANCH)
    return cell, om


def _run(cell, prm, dtype, device, T, amp=0.2):
    ps = data_set_params(cell, prm, device, dtype)
    cur = torch.tensor(c1_stimulus(T) * (amp / 0.2), dtype=dtype, device=device)
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    with torch.no_grad():
        return jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT)[:, 0].double().cpu()


def test_c0_full_size_float32(device, f32_mode):
    """The 60-compartment cell, 400 steps, every compartment recorded, against the oracle
    (float64, same float32 inputs)."""
    cell, om = build_c0()
    cur = c0_stimulus(400)
    cur32 = cur.astype(np.float32)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur32, device=device))
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT)
    assert v.dtype == torch.float32 and v.shape == (60, 401)
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur32.astype(np.float64))[None], stim_comps=[0],
                     rec_comps=list(range(60)))
    err = np.abs(v.double().cpu().numpy() - ref.numpy()).max(-1)
    print(f"{f32_mode}: max err {err.max():.3e} mV (per compartment median {np.median(err):.2e}), "
          f"spikes {spike_counts(ref.numpy()).sum()}")
    assert err.max() < 1e-2
    assert np.array_equal(spike_counts(v.double().cpu().numpy()), spike_counts(ref.numpy()))


def test_c1_tree_float32_vs_oracle(device, f32_mode):
    """Four simulations of the real sweep (250 branches, 1127 solve nodes, 2000 steps)
    against the oracle on the same float32 inputs."""
    B, T = 4, 2000
    prm = c1_params(B, SEED)
    prm32 = {k: x.astype(np.float32).astype(np.float64) for k, x in prm.items()}
    cell, _ = _c1_cell(False)
    v = _run(cell, prm, torch.float32, device, T)
    _, om = _c1_cell(True)
    stim = torch.tensor(c1_stimulus(T) * float(np.float32(0.2)) / 0.2)[None]
    ref = O.simulate(om, None, DT, stim=stim, stim_comps=[0], rec_comps=[0],
                     overrides=oracle_overrides(om, prm32))[:, 0]
    err = (v - ref).abs().amax(-1)
    print(f"{f32_mode}: per-simulation max err {err.tolist()}, spikes {spike_counts(ref.numpy()).tolist()}")
    assert float(err.max()) < 1e-2
    assert np.array_equal(spike_counts(v.numpy()), spike_counts(ref.numpy()))
    assert spike_counts(ref.numpy()).sum() > 0


def test_c1_sweep_float32_conditioned(device, f32_mode):
    """64 simulations of the sweep, 2000 steps: float32 against float64 on the same
    float32 inputs (the float64 kernel matches the oracle to 1e-6 mV on this tree,
    test_c1_tree_float64_vs_oracle); per-simulation error, spike counts, and the
    simulations' own sensitivity to the float32 rounding of their inputs."""
    B, T = 64, 2000
    prm = c1_params(B, SEED)
    prm32 = {k: x.astype(np.float32).astype(np.float64) for k, x in prm.items()}
    v32 = _run(_c1_cell(False)[0], prm, torch.float32, device, T)
    ref_exact = _run(_c1_cell(False)[0], prm, torch.float64, device, T)
    ref = _run(_c1_cell(True)[0], prm32, torch.float64, device, T, amp=float(np.float32(0.2)))
    err = (v32 - ref).abs().amax(-1).numpy()
    sens = (ref - ref_exact).abs().amax(-1).numpy()
    good = sens <= 3e-3
    print(f"{f32_mode}: max err {err.max():.3e} mV, well-conditioned {good.sum()}/{B} (max err there "
          f"{err[good].max():.3e}), input-rounding sensitivity max {sens.max():.3e}; "
          f"errors {np.round(err, 5).tolist()}")
    assert np.array_equal(spike_counts(v32.numpy()), spike_counts(ref.numpy()))
    assert good.sum() >= 60
    assert err[good].max() < (1e-2 if f32_mode == "mixed" else 0.15)


@pytest.mark.parametrize("shape", [None, "8,1024,sym", "8,1024,gcoef"])
def test_c1_tree_float64_vs_oracle(device, shape, monkeypatch):
    """The real sweep tree at float64: four simulations of 2000 steps within 1e-6 mV of the
    oracle, identical spike counts; with the library's launch shape and with the sweep's
    (8 simulations per program: the symmetric system's coefficients in LDS, or the pairs
    from L2; 4 padding simulations)."""
    from jaxhip import autotune
    import jaxhip.fused as F

    B, T = 4, 2000
    prm = c1_params(B, SEED)
    cell, om = _c1_cell(False)
    modes = []
    if shape is not None:
        monkeypatch.setenv("JAXHIP_FUSED_CONFIG", shape)
        monkeypatch.setattr(autotune, "_CACHE", {})
        orig = F.ForwardLaunch.__init__

        def spy(self, *args, **kw):
            orig(self, *args, **kw)
            modes.append(self.args.coef_global)
        monkeypatch.setattr(F.ForwardLaunch, "__init__", spy)
    v = _run(cell, prm, torch.float64, device, T)
    if shape is not None:
        assert modes == [2 if shape.endswith("sym") else 1], modes
    ref = O.simulate(om, None, DT, stim=torch.tensor(c1_stimulus(T))[None], stim_comps=[0], rec_comps=[0],
                     overrides=oracle_overrides(om, prm))[:, 0]
    err = float((v - ref).abs().max())
    print("c1 float64 max err", err, "spikes", spike_counts(ref.numpy()).tolist())
    assert err < 1e-6
    assert np.array_equal(spike_counts(v.numpy()), spike_counts(ref.numpy()))


def test_hh_at_singular_voltage_float32(device, f32_mode):
    """tests/test_float32_safety.py:83-93 of the reference: an HH compartment started
    exactly at -40 mV (alpha_m's removable singularity) stays finite in float32 and
    matches the oracle."""
    cell = jx.Cell(jx.Branch(jx.Compartment(), 1), [-1])
    cell.insert(jx.HH())
    cell.set("v", -40.0)
    cell.record()
    ds = cell.branch(0).comp(0).data_stimulate(torch.zeros(41, dtype=torch.float32, device=device))
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT)
    assert v.dtype == torch.float32 and torch.isfinite(v).all()
    om = O.OracleModel([([-1], [1])])
    om.hh_comps = np.array([0])
    om.v0 = np.array([-40.0])
    ref = O.simulate(om, None, DT, stim=torch.zeros(1, 41, dtype=torch.float64), stim_comps=[0])
    err = float((v.double().cpu() - ref).abs().max())
    print(f"{f32_mode}: -40 mV start, max err {err:.3e}")
    assert err < 1e-2


def test_fast_adjoint_float32(device):
    """The all-float32 reverse pass (`jxhip_fused_adjoint_f32`): gradients of a loss on the
    full-size 60-compartment cell (400 steps, per-branch trainable gNa, gK and gLeak,
    checkpoints every 100 steps) against autograd through the float64 oracle on the same
    float32 inputs (relative 1e-3 of the largest gradient) and the traces within 1e-2 mV."""
    from jaxhip.integrate import set_float32_mode

    set_float32_mode("fast")
    try:
        cell, om = build_c0()
        cur32 = c0_stimulus(400).astype(np.float32)
        ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur32, device=device))
        rng = np.random.default_rng(4)
        vals = {k: (d * rng.uniform(0.8, 1.2, 15)).astype(np.float32)
                for k, d in (("HH_gNa", 0.12), ("HH_gK", 0.036), ("HH_gLeak", 3e-4))}
        for k in vals:
            cell.branch("all").make_trainable(k, verbose=False)
        params = [{k: torch.tensor(vals[k], device=device, requires_grad=True)} for k in vals]
        v = jx.integrate(cell, params, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[4, 100])
        w = torch.tensor(rng.normal(size=(60, 401)), dtype=torch.float32, device=device)
        (v * w).sum().backward()
        leaves = {k: torch.tensor(np.repeat(vals[k].astype(np.float64), 4), requires_grad=True) for k in vals}
        ref = O.simulate(om, None, DT, stim=torch.tensor(cur32.astype(np.float64))[None], stim_comps=[0],
                         rec_comps=list(range(60)), overrides=leaves)
        (ref * w.double().cpu()).sum().backward()
        assert float((v.detach().double().cpu() - ref.detach()).abs().max()) < 1e-2
        for p, k in zip(params, vals):
            g = p[k].grad.double().cpu()
            g_ref = leaves[k].grad.reshape(15, 4).sum(1)
            rel = float((g - g_ref).abs().max() / g_ref.abs().max())
            print(f"{k}: relative gradient error {rel:.2e} (|g| max {float(g_ref.abs().max()):.3e})")
            assert rel < 1e-3
    finally:
        set_float32_mode("mixed")
