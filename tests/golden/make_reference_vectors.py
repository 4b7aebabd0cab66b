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
* the same for one ball-and-stick cell without synapses (0.1536 nA from 1.31 ms).

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
    out["dt"] = np.array(DT)
    out["t_max"] = np.array(T_MAX)
    np.savez_compressed(HERE / "reference_vectors.npz", **out)
    print("wrote", HERE / "reference_vectors.npz", {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
