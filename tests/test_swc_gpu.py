"""Cells read from SWC files (`jx.read_swc`) on the product path (`jaxhip.integrate`
through the fused HIP kernels, float64) against the reference's golden traces
(tests/jaxley_identical/test_swc.py of jaxley-pytorch: a 0.2 nA step from 0.5 to
1.5 ms into soma.branch(0).loc(1.0), every 20th step, tolerance 1e-8 mV), a network of
two such cells with an IonotropicSynapse, launch shapes against the oracle, and
gradients through the fused adjoint against autograd through the oracle (rtol 1e-6)."""
import json
import math
import os

import numpy as np
import pytest
import torch

import jaxhip as jx
from models import oracle_of_swc
from oracle import jx_oracle as O
from test_swc import GOLDEN, swc, swc_net_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def float64_default():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def _cur(device):
    return jx.step_current(0.5, 1.0, 0.2, 0.025, 5.0).to(device)


def _err(v, key):
    err = float(np.abs(v[:, ::20] - np.asarray(GOLDEN[key]["values"])).max())
    print(key, "max err vs reference golden", err, flush=True)
    return err


@pytest.mark.parametrize("file", ["morph_ca1_n120_single_point_soma", "morph_ca1_n120"])
def test_swc_cell_golden(device, file):
    cell = jx.read_swc(swc(file + ".swc"), ncomp=2, max_branch_len=300.0)
    _ = cell.soma
    cell.insert(jx.HH())
    cell.soma.branch(0).loc(1.0).record()
    ds = cell.soma.branch(0).loc(1.0).data_stimulate(_cur(device))
    v = jx.integrate(cell, data_stimuli=ds, delta_t=0.025).cpu().numpy()
    assert v.shape == (1, 202)
    assert _err(v, "swc_cell_" + file[6:]) <= 1e-8


@pytest.mark.parametrize("file", ["morph_ca1_n120", "morph_ca1_n120_250"])
def test_swc_net_golden(device, file):
    """tests/jaxley_identical/test_swc.py:83-190, built the way the reference's test
    builds it."""
    cell1 = jx.read_swc(swc(file + ".swc"), ncomp=2, max_branch_len=300.0)
    cell2 = jx.read_swc(swc(file + ".swc"), ncomp=2, max_branch_len=300.0)
    net = jx.Network([cell1, cell2])
    jx.connect(net.cell(0).soma.branch(0).loc(1.0), net.cell(1).soma.branch(0).loc(1.0),
               jx.IonotropicSynapse())
    net.insert(jx.HH())
    post = net.cell(1).soma.branch(0).comp(1).nodes
    area = 2 * math.pi * post["length"].item() * post["radius"].item()
    net.set("IonotropicSynapse_gS", 0.5 / (100_000.0 / area))
    ds = None
    for i in range(2):
        net.cell(i).soma.branch(0).loc(1.0).record()
        ds = net.cell(i).soma.branch(0).loc(1.0).data_stimulate(_cur(device), ds)
    v = jx.integrate(net, data_stimuli=ds, delta_t=0.025).cpu().numpy()
    assert v.shape == (2, 202)
    assert _err(v, "swc_net_" + file[6:]) <= 1e-8


def test_swc_launch_shapes_match_oracle(device, monkeypatch):
    """Streaming (LDS / L2 coefficients) and register-resident kernels on a traced
    morphology with batched conductances, against the oracle."""
    from jaxhip import autotune

    cell = jx.read_swc(swc("morph_ca1_n120_250.swc"), ncomp=4)
    cell.insert(jx.HH())
    comp = int(cell.soma.branch(0).loc(1.0)._comps[0])
    cell.soma.branch(0).loc(1.0).record()
    cell.branch(5).comp(3).record()
    rec = [comp, int(cell.branch(5).comp(3)._comps[0])]
    B, nc = 3, cell._n_comp
    gna = 0.12 * np.random.default_rng(1).uniform(0.8, 1.2, (B, nc))
    ps = cell.data_set("HH_gNa", torch.tensor(gna, device=device), None)
    cur = _cur(device)
    ds = cell.soma.branch(0).loc(1.0).data_stimulate(cur)
    ref = O.simulate(oracle_of_swc([cell]), None, 0.025, stim=cur.cpu()[None], stim_comps=[comp],
                     rec_comps=rec, overrides={"HH_gNa": torch.tensor(gna)}).numpy()
    for shape in ["8,1024,gcoef", "8,1024,sym", "4,512,stream", "2,256", "1,1024"]:
        monkeypatch.setenv("JAXHIP_FUSED_CONFIG", shape)
        autotune._CACHE.clear()
        v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=0.025).cpu().numpy()
        err = np.abs(v - ref).max()
        print("swc shape", shape, "max err vs oracle", err, flush=True)
        assert v.shape == ref.shape and err < 1e-6, shape
    autotune._CACHE.clear()


def test_swc_gradients_vs_oracle_autograd(device):
    """Per-branch gNa of a traced cell through the fused adjoint."""
    cell = jx.read_swc(swc("morph_ca1_n120_250.swc"), ncomp=2)
    cell.insert(jx.HH())
    comp = int(cell.soma.branch(0).loc(1.0)._comps[0])
    cell.soma.branch(0).loc(1.0).record()
    cell.branch("all").make_trainable("HH_gNa", verbose=False)
    nb = len(cell.comb_parents)
    vals = 0.12 * np.random.default_rng(3).uniform(0.7, 1.3, nb)
    params = [{"HH_gNa": torch.tensor(vals, device=device, requires_grad=True)}]
    cur = _cur(device)
    v = jx.integrate(cell, params, data_stimuli=cell.soma.branch(0).loc(1.0).data_stimulate(cur),
                     delta_t=0.025)
    target = torch.linspace(-70, -60, v.shape[-1], device=device)
    loss = ((v[0] - target) ** 2).mean()
    loss.backward()
    ov = torch.tensor(vals, requires_grad=True)
    ref = O.simulate(oracle_of_swc([cell]), None, 0.025, stim=cur.cpu()[None], stim_comps=[comp],
                     rec_comps=[comp], overrides={"HH_gNa": torch.repeat_interleave(ov, 2)})
    oloss = ((ref[0] - target.cpu()) ** 2).mean()
    oloss.backward()
    print("swc loss", loss.item(), oloss.item(), flush=True)
    np.testing.assert_allclose(loss.item(), oloss.item(), rtol=1e-9)
    np.testing.assert_allclose(params[0]["HH_gNa"].grad.cpu().numpy(), ov.grad.numpy(), rtol=1e-6, atol=1e-9)


def test_single_branch_swc_no_nan(device):
    """tests/test_swc.py:161-169 (without set_ncomp): no NaN in a 1 ms simulation."""
    cell = jx.read_swc(swc("morph_single_branch.swc"), ncomp=4)
    cell.insert(jx.HH())
    cell.branch(0).comp(0).record()
    ds = cell.branch(0).comp(0).data_stimulate(torch.zeros(41, dtype=torch.float64, device=device))
    v = jx.integrate(cell, data_stimuli=ds, delta_t=0.025)
    assert not torch.isnan(v).any()
