"""Writes `tests/golden/reference_vectors.npz`: input and output vectors of jaxley-pytorch
itself (imported from /root/reference, on the CPU, in float64) for what none of the
reference's own tests holds a golden value for:

* the DHS solve indexers (`_dhs_solve_indexer`, modules/base.py:1142-1266 and
  network.py:192-293 of the reference, `allowed_nodes_per_level` 32) of the
  60-compartment binary-tree cell, the 250-branch random tree (seed 11) and a 6-cell
  ball-and-stick network;
* float64 voltage traces of 6-cell ball-and-stick networks (1-compartment HH soma,
  8-compartment Leak dendrite) whose cells are connected by KineticSynapse or by
  SigmoidRateSynapse (3 synapses onto one dendritic compartment of every cell, ordered
  by postsynaptic cell: the reference's regular layout), with step currents into the
  somata that make the cells spike; 30 ms at dt 0.025 ms, stored every 4th step;
* the same for one ball-and-stick cell without synapses (0.1536 nA from 1.31 ms);
* float64 traces of the 60-compartment binary-tree cell (10 ms, all compartments) and
  of four simulations of the 250-branch random tree (2000 steps, the root), every 4th
  step;
* gradients of the reference's autograd: per-branch gNa / gK / gLeak of a 12-branch
  random tree (2 parameter sets, MSE against a target trace, checkpoints every 50 steps)
  and every gS and the somata's gNa of the KineticSynapse network (2 simulations,
  a weighted sum of the somatic recordings, checkpoints every 100 steps).

The inputs (structures, synapse ends, parameters, stimuli) are stored beside the
outputs, so that the tests rebuild the same models. Run in the build container only:

    python tests/golden/make_reference_vectors.py

The random structures come from `jaxhip.workloads` (numpy only); everything simulated
or indexed here is the reference's.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]
sys.path.insert(0, str(REPO / "jaxley-pytorch_amd"))
sys.path.insert(0, "/root/reference")

import jaxley_pytorch as rjx  # noqa: E402
from jaxley_pytorch.channels import HH, Leak  # noqa: E402
from jaxley_pytorch.connect import connect  # noqa: E402
from jaxley_pytorch.synapses import KineticSynapse, SigmoidRateSynapse  # noqa: E402

from jaxhip.workloads import binary_tree_parents, random_tree  # noqa: E402

torch.set_default_dtype(torch.float64)

DT = 0.025
T_MAX = 30.0
N_CELLS, K_IN = 6, 3
C4_LENGTH = np.array([20.0] + [25.0] * 8)
C4_RADIUS = np.array([10.0] + [1.0] * 8)
INDEXER_KEYS = ("map_to_solve_order", "inv_map_to_solve_order", "map_to_solve_order_lower",
                "map_to_solve_order_upper", "node_order", "parent_lookup")


def ball_and_stick():
    cell = rjx.Cell([rjx.Branch(rjx.Compartment(), 1), rjx.Branch(rjx.Compartment(), 8)], [-1, 0])
    cell.set("length", C4_LENGTH)
    cell.set("radius", C4_RADIUS)
    cell.branch(0).insert(HH())
    cell.branch(1).insert(Leak())
    return cell


def network_spec(seed: int):
    """Synapse ends (presynaptic soma, postsynaptic dendritic compartment), parameters and
    stimulus amplitudes of the 6-cell network."""
    rng = np.random.default_rng(seed)
    pre = np.stack([rng.choice([c for c in range(N_CELLS) if c != post], size=K_IN, replace=False)
                    for post in range(N_CELLS)]).astype(np.int64)          # (cells, k)
    tgt = rng.integers(0, 8, size=N_CELLS).astype(np.int64)                # dendrite comp
    gS = rng.uniform(2e-3, 6e-3, size=(N_CELLS, K_IN))
    e_syn = np.where(rng.uniform(size=(N_CELLS, K_IN)) < 0.7, 0.0, -80.0)
    amp = rng.uniform(0.15, 0.45, size=N_CELLS)                             # nA
    delay = rng.uniform(1.0, 8.0, size=N_CELLS)                             # ms
    return {"pre": pre, "tgt": tgt, "gS": gS, "e_syn": e_syn, "amp": amp, "delay": delay}


def simulate_network(synapse: str, spec: dict):
    net = rjx.Network([ball_and_stick() for _ in range(N_CELLS)])
    cls = {"KineticSynapse": KineticSynapse, "SigmoidRateSynapse": SigmoidRateSynapse}[synapse]
    for post in range(N_CELLS):
        for j in range(K_IN):
            connect(net.cell(int(spec["pre"][post, j])).branch(0).comp(0),
                    net.cell(post).branch(1).comp(int(spec["tgt"][post])), cls())
    syn = getattr(net, synapse)
    syn.set(f"{synapse}_gS", spec["gS"].reshape(-1))
    if synapse == "KineticSynapse":
        syn.set(f"{synapse}_e_syn", spec["e_syn"].reshape(-1))
    for c in range(N_CELLS):
        cur = rjx.step_current(float(spec["delay"][c]), 15.0, float(spec["amp"][c]), DT, T_MAX)
        net.cell(c).branch(0).comp(0).stimulate(cur, verbose=False)
    for c in range(N_CELLS):
        net.cell(c).branch(0).comp(0).record(verbose=False)
    for c in range(N_CELLS):
        net.cell(c).branch(1).comp(int(spec["tgt"][c])).record(verbose=False)
    with torch.no_grad():
        v = rjx.integrate(net, delta_t=DT, t_max=T_MAX)
    return net, v.numpy()


def tree_gradients(out: dict):
    """Per-branch trainable gNa / gK / gLeak of a 12-branch random tree (B = 2 parameter
    sets), MSE of the root voltage against a target trace over 200 steps, checkpoints
    every 50 steps: the loss and the gradients of the reference's autograd."""
    from jaxhip.workloads import tree_cell as hip_tree_cell

    nb, B, T, K = 12, 2, 200, 50
    _, parents, comp_len, comp_rad = hip_tree_cell(nb, 7)
    cell = rjx.Cell([rjx.Branch(rjx.Compartment(), 4)] * nb, np.asarray(parents))
    cell.set("length", comp_len)
    cell.set("radius", comp_rad)
    cell.insert(HH())
    cell.branch(0).comp(0).record(verbose=False)
    keys = ("HH_gNa", "HH_gK", "HH_gLeak")
    for k in keys:
        cell.branch("all").make_trainable(k, verbose=False)
    rng = np.random.default_rng(1)
    vals = np.array([0.12, 0.036, 0.0003])[None, :, None] * rng.uniform(0.7, 1.3, size=(B, 3, nb))
    params = [{k: torch.tensor(vals[:, j], requires_grad=True)} for j, k in enumerate(keys)]
    cur = np.zeros(T)
    cur[int(1.0 / DT):int(21.0 / DT)] = 0.2
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur))
    v = rjx.integrate(cell, params=params, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[T // K, K])
    target = torch.linspace(-70, -60, T + 1)
    loss = ((v[:, 0] - target) ** 2).mean()
    loss.backward()
    out["grad_tree_parents"] = np.asarray(parents, dtype=np.int64)
    out["grad_tree_length"] = np.asarray(comp_len, dtype=np.float64)
    out["grad_tree_radius"] = np.asarray(comp_rad, dtype=np.float64)
    out["grad_tree_vals"] = vals
    out["grad_tree_stim"] = cur
    out["grad_tree_loss"] = np.array(loss.item())
    out["grad_tree_grads"] = np.stack([p[k].grad.numpy() for p, k in zip(params, keys)], 1)


def network_gradients(spec: dict, out: dict):
    """The KineticSynapse network (B = 2): gradients of a weighted sum of the recordings
    with respect to every synapse's gS (shared by the batch) and the somata's gNa (one
    per simulation), checkpoints every 100 steps."""
    net = rjx.Network([ball_and_stick() for _ in range(N_CELLS)])
    for post in range(N_CELLS):
        for j in range(K_IN):
            connect(net.cell(int(spec["pre"][post, j])).branch(0).comp(0),
                    net.cell(post).branch(1).comp(int(spec["tgt"][post])), KineticSynapse())
    net.KineticSynapse.set("KineticSynapse_e_syn", spec["e_syn"].reshape(-1))
    for c in range(N_CELLS):
        cur = rjx.step_current(float(spec["delay"][c]), 15.0, float(spec["amp"][c]), DT, T_MAX)
        net.cell(c).branch(0).comp(0).stimulate(cur, verbose=False)
    for c in range(N_CELLS):
        net.cell(c).branch(0).comp(0).record(verbose=False)
    gs = torch.tensor(spec["gS"].reshape(-1), requires_grad=True)
    gna = torch.tensor([[0.12], [0.1]], requires_grad=True)
    ps = net.data_set("KineticSynapse_gS", gs, None)
    ps = net.data_set("HH_gNa", gna, ps)
    v = rjx.integrate(net, param_state=ps, delta_t=DT, t_max=T_MAX, checkpoint_lengths=[12, 100])
    # The weights: numpy's seeded generator (the tests draw the same ones).
    w = np.random.default_rng(5).normal(size=tuple(v.shape))
    (v * torch.tensor(w)).sum().backward()
    out["grad_net_w_seed_shape"] = np.array([5, *v.shape], dtype=np.int64)
    out["grad_net_gna"] = gna.detach().numpy()
    out["grad_net_grad_gS"] = gs.grad.numpy()
    out["grad_net_grad_gNa"] = gna.grad.numpy()
    out["grad_net_v"] = v.detach().numpy()[:, :, ::4]


def config_traces(out: dict):
    """The 60-compartment binary-tree cell (HH, 2.5 um compartments, 0.1 nA into
    compartment 0 from 1 to 2 ms, 10 ms, every compartment recorded) and four
    simulations of the 250-branch random tree of the batched sweep (per-simulation gNa,
    gK, gLeak; 0.2 nA into the root from 1 to 21 ms; 2000 steps; root recorded)."""
    from jaxhip.workloads import c1_params, tree_cell as hip_tree_cell

    c0 = rjx.Cell([rjx.Branch(rjx.Compartment(), 4)] * 15, np.asarray(binary_tree_parents(15)))
    c0.set("length", 2.5)
    c0.insert(HH())
    c0.record(verbose=False)
    c0.branch(0).comp(0).stimulate(rjx.step_current(1.0, 1.0, 0.1, DT, 10.0), verbose=False)
    with torch.no_grad():
        out["c0_v"] = rjx.integrate(c0, delta_t=DT, t_max=10.0).numpy()[:, ::4]
    _, parents, comp_len, comp_rad = hip_tree_cell(250, 11)
    c1 = rjx.Cell([rjx.Branch(rjx.Compartment(), 4)] * 250, np.asarray(parents))
    c1.set("length", comp_len)
    c1.set("radius", comp_rad)
    c1.insert(HH())
    c1.branch(0).comp(0).record(verbose=False)
    prm = c1_params(4, 11)
    ps = None
    for k, v in prm.items():
        ps = c1.data_set(k, torch.tensor(v), ps)
    cur = np.zeros(2000)
    cur[int(1.0 / DT):int(21.0 / DT)] = 0.2
    ds = c1.branch(0).comp(0).data_stimulate(torch.tensor(cur))
    with torch.no_grad():
        v = rjx.integrate(c1, param_state=ps, data_stimuli=ds, delta_t=DT)
    out["c1_v"] = v.numpy()[:, :, ::4]
    out["c1_params"] = np.concatenate([prm[k] for k in ("HH_gNa", "HH_gK", "HH_gLeak")], 1)
    print("c1 spikes per simulation", [int(((x[0, 1:] >= 0) & (x[0, :-1] < 0)).sum()) for x in v.numpy()])


def indexer_arrays(module, prefix: str, out: dict):
    ind = module._dhs_solve_indexer
    for k in INDEXER_KEYS:
        out[f"{prefix}_{k}"] = np.asarray(ind[k], dtype=np.int64)
    grouped = ind["node_order_grouped"]
    out[f"{prefix}_grouped"] = np.concatenate(grouped, 0).astype(np.int64) if len(grouped) else \
        np.zeros((0, 2), dtype=np.int64)
    out[f"{prefix}_grouped_sizes"] = np.array([len(g) for g in grouped], dtype=np.int64)


def main():
    out = {}
    # Binary-tree cell: root + full binary tree of depth 3, 4 compartments per branch.
    c0 = rjx.Cell([rjx.Branch(rjx.Compartment(), 4)] * 15, np.asarray(binary_tree_parents(15)))
    indexer_arrays(c0, "c0", out)
    # The 250-branch random tree.
    parents, _, _ = random_tree(250, 11)
    c1 = rjx.Cell([rjx.Branch(rjx.Compartment(), 4)] * 250, np.asarray(parents))
    out["c1_parents"] = np.asarray(parents, dtype=np.int64)
    indexer_arrays(c1, "c1", out)
    # One ball-and-stick cell, a step current into the soma (no synapses).
    cell = ball_and_stick()
    cell.branch(0).comp(0).stimulate(rjx.step_current(1.31, 15.0, 0.1536, DT, T_MAX), verbose=False)
    cell.branch(0).comp(0).record(verbose=False)
    cell.branch(1).comp(3).record(verbose=False)
    with torch.no_grad():
        out["cell_v"] = rjx.integrate(cell, delta_t=DT, t_max=T_MAX).numpy()[:, ::4]
    for synapse, seed in (("KineticSynapse", 21), ("SigmoidRateSynapse", 22)):
        spec = network_spec(seed)
        net, v = simulate_network(synapse, spec)
        key = synapse.lower()
        for k, x in spec.items():
            out[f"{key}_{k}"] = x
        out[f"{key}_v"] = v[:, ::4]     # every 4th step (0.1 ms), the fixture stays small
        print(synapse, v.shape, "spikes per recorded soma",
              [int(((v[i, 1:] >= 0) & (v[i, :-1] < 0)).sum()) for i in range(N_CELLS)])
        if synapse == "KineticSynapse":
            indexer_arrays(net, "net", out)
            network_gradients(spec, out)
    tree_gradients(out)
    config_traces(out)
    out["dt"] = np.array(DT)
    out["t_max"] = np.array(T_MAX)
    np.savez_compressed(HERE / "reference_vectors.npz", **out)
    print("wrote", HERE / "reference_vectors.npz", {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
