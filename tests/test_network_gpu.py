"""Networks through the fused kernel (synapse step fused into the voltage and gate
steps) against the oracle, float64 and float32: ball-and-stick cells, KineticSynapse
and SigmoidRateSynapse with fixed in-degree onto one dendritic compartment per cell,
Bernoulli current pulses into every soma."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from models import DT, build_c4, c4_stimulus
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu


def _run(device, synapse, dtype, n_cells=12, k=4, B=2, T=400, seed=4):
    net, om, info = build_c4(n_cells, k, seed, synapse)
    assert np.array_equal(net._comp_node[info["pre_comp"]], om.internal[info["pre_comp"]])
    soma = info["soma_comps"]
    for c in soma:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    net.cell(0).branch(1).comp(3).record()
    cur = c4_stimulus(B, n_cells, T, seed + 1)
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(
            torch.tensor(cur[:, i, :], dtype=dtype, device=device), ds)
    v = jx.integrate(net, data_stimuli=ds, delta_t=DT)
    rec = list(soma) + [3 + 1]
    ref = O.simulate(om, None, DT, stim=torch.tensor(cur), stim_comps=list(soma), rec_comps=rec)
    return v, ref


@pytest.mark.parametrize("synapse", ["KineticSynapse", "SigmoidRateSynapse"])
def test_network_float64(device, synapse):
    v, ref = _run(device, synapse, torch.float64)
    assert v.shape == ref.shape
    err = float((v.cpu() - ref).abs().max())
    print(synapse, "fp64 max err", err)
    assert err < 1e-6


def test_network_float32(device):
    v, ref = _run(device, "KineticSynapse", torch.float32)
    err = float((v.double().cpu() - ref).abs().max())
    print("fp32 max err", err)
    assert err < 1e-2
