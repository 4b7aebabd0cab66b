"""Pospischil Na, K and Km channels (channels/pospischil.py:74-250 of jaxley-pytorch) in
the fused HIP kernels against the oracle: a random tree with the channels on different
branches, mixed with HH and Leak, on every launch shape, in float64 (1e-6 mV) and
float32 (1e-2 mV), with the three time steppers; the reference's steady-state property
(tests/test_channels.py:181-208, without its calcium channels) through `integrate`."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from models import DT, build_tree, c1_stimulus, oracle_of_module, pospischil_cell
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = ["8,1024,gcoef", "8,1024,sym", "4,512,stream", "2,256", "1,1024"]


def _tree():
    nb = 24
    cell, _ = build_tree(nb, seed=11)
    br = lambda ids: cell.branch(list(ids))
    br(range(0, 16)).insert(jx.Na())
    br(range(8, 24)).insert(jx.K())
    br(range(4, 12)).insert(jx.Km())
    cell.insert(jx.Leak())
    cell.branch(0).set("vt", -55.0)
    cell.branch(list(range(12, 24))).set("Km_gKm", 0.002)
    cell.set("Km_taumax", 400.0)
    rec = [0, 4 * 10 + 1, 4 * nb - 1]
    for c in rec:
        cell.select(nodes=[c]).record()
    return cell, rec


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-6), (torch.float32, 1e-2)])
def test_pospischil_launch_shapes_match_oracle(device, dtype, tol, monkeypatch):
    from jaxhip import autotune

    cell, rec = _tree()
    T = 400
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=dtype, device=device))
    ref = O.simulate(oracle_of_module(cell), None, DT, stim=torch.tensor(cur)[None], stim_comps=[0],
                     rec_comps=rec).numpy()
    # The Pospischil channels change the dynamics of this input.
    plain, _ = build_tree(24, seed=11)
    plain.insert(jx.Leak())
    base = O.simulate(oracle_of_module(plain), None, DT, stim=torch.tensor(cur)[None], stim_comps=[0],
                      rec_comps=rec).numpy()
    assert np.abs(base - ref).max() > 1.0
    for shape in SHAPES:
        monkeypatch.setenv("JAXHIP_FUSED_CONFIG", shape)
        autotune._CACHE.clear()
        v = jx.integrate(cell, data_stimuli=ds, delta_t=DT).double().cpu().numpy()
        err = np.abs(v - ref).max()
        print(f"pospischil shape {shape} {dtype} max err {err:.3e}", flush=True)
        assert v.shape == ref.shape and err < tol, shape
    autotune._CACHE.clear()


@pytest.mark.parametrize("solver", ["crank_nicolson", "fwd_euler"])
def test_pospischil_solvers_match_oracle(device, solver):
    cell, rec = _tree()
    cell.set("length", 100.0)
    cell.set("axial_resistivity", 200.0)
    T = 300
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    ref = O.simulate(oracle_of_module(cell), None, DT, stim=torch.tensor(cur)[None], stim_comps=[0],
                     rec_comps=rec, solver=solver).numpy()
    v = jx.integrate(cell, data_stimuli=ds, delta_t=DT, solver=solver).cpu().numpy()
    err = np.abs(v - ref).max()
    print(f"pospischil {solver} max err {err:.3e}", flush=True)
    assert np.isfinite(ref).all() and err < 1e-6


def test_pospischil_steady_state(device):
    """tests/test_channels.py:181-208 through the fused kernel, and the traces equal the
    oracle's."""
    cell = pospischil_cell()
    cell.set("v", -62.0)
    cell.init_states()
    zero = lambda T: cell.branch(0).comp(0).data_stimulate(torch.zeros(T, dtype=torch.float64, device=device))
    v = jx.integrate(cell, data_stimuli=zero(801), delta_t=0.025).cpu().numpy()
    ref = O.simulate(oracle_of_module(cell), None, 0.025, stim=torch.zeros(1, 801, dtype=torch.float64),
                     stim_comps=[0], rec_comps=[0]).numpy()
    assert np.abs(v - ref).max() < 1e-8
    cell.set("v", float(v[0, -1]))
    cell.init_states()
    v = jx.integrate(cell, data_stimuli=zero(401), delta_t=0.025).cpu().numpy()
    print("pospischil steady state drift", abs(v[0, 0] - v[0, -1]), flush=True)
    assert abs(v[0, 0] - v[0, -1]) < 0.02


@pytest.mark.parametrize("solver", ["bwd_euler", "crank_nicolson"])
def test_pospischil_gradients_vs_oracle_autograd(device, solver):
    """Per-branch Na gNa and K gK (trainable) through the per-step path against autograd
    through the oracle, and a batch of vt values (batched Pospischil parameters)."""
    cell, rec = _tree()
    cell.branch("all").make_trainable("Na_gNa", verbose=False)
    cell.branch("all").make_trainable("K_gK", verbose=False)
    rng = np.random.default_rng(5)
    na = 0.05 * rng.uniform(0.7, 1.3, 16)
    kk = 0.005 * rng.uniform(0.7, 1.3, 16)
    params = [{"Na_gNa": torch.tensor(na, device=device, requires_grad=True)},
              {"K_gK": torch.tensor(kk, device=device, requires_grad=True)}]
    T = 200
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    v = jx.integrate(cell, params, data_stimuli=ds, delta_t=DT, solver=solver, checkpoint_lengths=[50])
    w = torch.tensor(rng.normal(size=tuple(v.shape)), device=device)
    loss = (v * w).sum()
    loss.backward()
    ona = torch.tensor(na, requires_grad=True)
    okk = torch.tensor(kk, requires_grad=True)
    z = torch.zeros(32, dtype=torch.float64)
    ov = {"Na_gNa": torch.cat([torch.repeat_interleave(ona, 4), z]),
          "K_gK": torch.cat([z, torch.repeat_interleave(okk, 4)])}
    ref = O.simulate(oracle_of_module(cell), None, DT, stim=torch.tensor(cur)[None], stim_comps=[0],
                     rec_comps=rec, overrides=ov, solver=solver)
    oloss = (ref * w.cpu()).sum()
    oloss.backward()
    print(f"pospischil {solver} loss", loss.item(), oloss.item(), flush=True)
    np.testing.assert_allclose(loss.item(), oloss.item(), rtol=1e-9)
    np.testing.assert_allclose(params[0]["Na_gNa"].grad.cpu().numpy(), ona.grad.numpy(), rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(params[1]["K_gK"].grad.cpu().numpy(), okk.grad.numpy(), rtol=1e-6, atol=1e-9)


def test_pospischil_batched_parameters(device):
    """A batch of threshold parameters vt (the per-step path) against the oracle."""
    cell, rec = _tree()
    T = 200
    cur = c1_stimulus(T)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    vts = torch.tensor([[-60.0], [-57.0], [-63.0]], dtype=torch.float64)
    ps = cell.data_set("vt", vts.to(device), None)
    v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT).cpu().numpy()
    for b in range(3):
        om = oracle_of_module(cell)
        om.posp_params["vt"] = np.where(np.isnan(cell._node_data["vt"]), 0.0, float(vts[b, 0]))
        ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=rec).numpy()
        err = np.abs(v[b] - ref).max()
        print("pospischil batched vt", float(vts[b, 0]), "max err", err, flush=True)
        assert v[b].shape == ref.shape and err < 1e-6, b


def test_pospischil_step_loop_matches_integrate(device):
    """`build_init_and_step_fn` carries the Pospischil gates between steps: T single
    steps equal one `integrate` of T steps."""
    cell, rec = _tree()
    T = 120
    cur = torch.tensor(c1_stimulus(T), dtype=torch.float64, device=device)
    init_fn, step_fn = jx.build_init_and_step_fn(cell)
    ps = cell.data_set("HH_gK", torch.tensor([[0.04]], dtype=torch.float64, device=device), None)
    states, all_params = init_fn([], param_state=ps)
    assert "Na_m" in states and "Km_p" in states
    node = int(cell._comp_node[0])
    out = [states["v"][..., node]]
    for t in range(T):
        states = step_fn(states, all_params, {"i": cur[t:t + 1]}, {"i": np.array([node])}, delta_t=DT)
        out.append(states["v"][..., node])
    v_step = torch.stack([x.reshape(-1)[0] for x in out])
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    v_int = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT)[..., 0, :].reshape(T + 1)
    err = float((v_step - v_int).abs().max())
    print("pospischil step loop vs integrate", err, flush=True)
    assert err < 1e-9
