"""Per-step API (`build_init_and_step_fn`) on the GPU: a stepped loop equals `integrate`,
and a step captured in a CUDA / HIP graph replays to the eager result."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from models import DT, build_c0, build_tree, c0_stimulus, c1_stimulus

pytestmark = pytest.mark.gpu


def test_step_loop_matches_integrate(device):
    cell, _ = build_c0()
    T = 120
    cur = torch.tensor(c0_stimulus(T), dtype=torch.float64, device=device)
    init_fn, step_fn = jx.build_init_and_step_fn(cell)
    gna = torch.tensor([[0.11]], dtype=torch.float64, device=device)
    states, all_params = init_fn([], param_state=cell.data_set("HH_gNa", gna, None))
    comp_node = torch.as_tensor(cell._comp_node, device=device)
    stim_node = np.array([cell._comp_node[0]])
    rec = [states["v"][..., comp_node]]
    for t in range(T):
        states = step_fn(states, all_params, {"i": cur[t:t + 1]}, {"i": stim_node}, delta_t=DT)
        rec.append(states["v"][..., comp_node])
    v_step = torch.stack([r.reshape(60) for r in rec], -1)
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    ps = cell.data_set("HH_gNa", gna, None)
    v_int = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT).reshape(60, T + 1)
    err = float((v_step - v_int).abs().max())
    print("step loop vs integrate", err)
    assert err < 1e-9


def test_step_graph_replay(device):
    cell, _ = build_tree(20, seed=4)
    cell.branch(0).comp(0).record()
    init_fn, step_fn = jx.build_init_and_step_fn(cell)
    states, all_params = init_fn([], param_state=cell.data_set(
        "HH_gK", torch.tensor([[0.03], [0.04]], dtype=torch.float64, device=device), None))
    states["v"] = torch.broadcast_to(states["v"], (2, states["v"].shape[-1])).contiguous()
    for k in ("HH_m", "HH_h", "HH_n"):
        states[k] = torch.broadcast_to(states[k], (2, states[k].shape[-1])).contiguous()
    inds = {"i": np.array([cell._comp_node[0]])}
    cur = torch.tensor(c1_stimulus(60), dtype=torch.float64, device=device)
    static_in = {k: v.clone() for k, v in states.items()}
    static_i = torch.zeros(1, dtype=torch.float64, device=device)
    step_fn(static_in, all_params, {"i": static_i}, inds, delta_t=DT)     # build buffers
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        static_out = step_fn(static_in, all_params, {"i": static_i}, inds, delta_t=DT)
    eager = {k: v.clone() for k, v in states.items()}
    for t in range(40, 52):                      # inside the stimulus window
        static_i.copy_(cur[t:t + 1])
        g.replay()
        torch.cuda.synchronize()
        eager = step_fn(eager, all_params, {"i": cur[t:t + 1]}, inds, delta_t=DT)
        torch.cuda.synchronize()
        for k in ("v", "HH_m"):
            a, b = static_out[k], eager[k]
            assert torch.equal(torch.nan_to_num(a, 7.0), torch.nan_to_num(b, 7.0)), (t, k)
        for k in static_in:
            static_in[k].copy_(static_out[k])
