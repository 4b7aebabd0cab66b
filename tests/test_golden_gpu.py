"""The product path (`jaxhip.integrate` through the fused HIP kernel, float64) against the
reference's own golden voltage traces (tests/jaxley_identical/test_basic_modules.py and
test_radius_and_length.py of jaxley-pytorch, frozen as tests/golden/reference_traces.json:
every 20th step of a 0.02 nA step current from 0.5 to 1.5 ms, dt = 0.025 ms, 201 steps).
The reference checks these traces at 1e-8 mV; so does this test."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

import jaxhip as jx

pytestmark = pytest.mark.gpu

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_traces.json").read_text())


@pytest.fixture(autouse=True)
def float64_default():
    """The reference's tests run with float64 as the default dtype (tests/conftest.py:17),
    which `step_current` follows (stimulus.py:8-36)."""
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def _run(module, stim_view, rec_view, device):
    cur = jx.step_current(0.5, 1.0, 0.02, 0.025, 5.0).to(device)
    assert cur.dtype == torch.float64
    rec_view.record()
    v = jx.integrate(module, data_stimuli=stim_view.data_stimulate(cur), delta_t=0.025)
    return v.double().cpu().numpy()


def _check(v, key):
    gold = np.asarray(GOLD[key]["values"])
    err = np.abs(v[..., ::20] - gold).max()
    print(key, "max err vs reference golden", err)
    assert v.shape[-1] == 202          # the initial state and 201 steps
    assert err <= 1e-8, (key, err)


def test_golden_compartment(device):
    comp = jx.Compartment()
    comp.insert(jx.HH())
    _check(_run(comp, comp, comp, device)[0], "compartment")


def test_golden_branch(device):
    branch = jx.Branch(jx.Compartment(), 2)
    branch.insert(jx.HH())
    _check(_run(branch, branch.loc(0.0), branch.loc(0.0), device)[0], "branch2")


def test_golden_cell(device):
    cell = jx.Cell(jx.Branch(jx.Compartment(), 2), parents=[-1, 0, 0])
    cell.insert(jx.HH())
    _check(_run(cell, cell.branch(1).loc(0.0), cell.branch(1).loc(0.0), device)[0], "cell_3_2")


def test_golden_radius_and_length(device):
    np.random.seed(1)
    comp = jx.Compartment()
    comp.insert(jx.HH())
    comp.set("length", 5 * np.random.rand(1))
    comp.set("radius", np.random.rand(1))
    _check(_run(comp, comp, comp, device)[0], "radius_and_length_compartment")


def test_golden_network_ionotropic(device):
    """Two SimpleCell(3, 2) cells, an IonotropicSynapse from cell 0 (branch 0, loc 0) to
    cell 1 (branch 0, loc 0), gS = 0.5 / (1e5 / (2 pi 10)), both cells stimulated and
    recorded at branch 1, loc 0 (test_basic_modules.py:336-409)."""
    cell = jx.Cell(jx.Branch(jx.Compartment(), 2), parents=[-1, 0, 0])
    net = jx.Network([cell, cell])
    net.insert(jx.HH())
    jx.connect(net.cell(0).branch(0).loc(0.0), net.cell(1).branch(0).loc(0.0), jx.IonotropicSynapse())
    net.IonotropicSynapse.set("IonotropicSynapse_gS", 0.5 / (100000.0 / (2 * np.pi * 10.0 * 1.0)))
    cur = jx.step_current(0.5, 1.0, 0.02, 0.025, 5.0).to(device)
    ds = None
    for c in (0, 1):
        net.cell(c).branch(1).loc(0.0).record()
        ds = net.cell(c).branch(1).loc(0.0).data_stimulate(cur, ds)
    v = jx.integrate(net, data_stimuli=ds, delta_t=0.025).double().cpu().numpy()
    assert v.shape == (2, 202)
    _check(v, "net_2x_cell_3_2")


@pytest.mark.parametrize("solver,voltage_solver", [("fwd_euler", "jaxley.dhs"), ("bwd_euler", "jaxley.stone")])
def test_golden_branch_uneven_radiuses(device, solver, voltage_solver):
    """test_basic_modules.py:124-197: an 8-compartment branch, r_a 500, lengths
    linspace(20, 300, 8), radii linspace(1, 5, 8), 2.0 nA from 0.5 to 1.5 ms into
    loc(1.0), recorded at loc(0.0); every 50th step against the reference at 1e-5 mV."""
    branch = jx.Branch(jx.Compartment(), 8)
    branch.set("axial_resistivity", 500.0)
    branch.set("length", np.linspace(20, 300, 8))
    branch.set("radius", np.linspace(1, 5, 8))
    branch.insert(jx.HH())
    branch.loc(0.0).record()
    cur = jx.step_current(0.5, 1.0, 2.0, 0.025, 10.0).to(device)
    v = jx.integrate(branch, data_stimuli=branch.loc(1.0).data_stimulate(cur), delta_t=0.025,
                     solver=solver, voltage_solver=voltage_solver).double().cpu().numpy()[0]
    gold = np.asarray(GOLD[f"branch8_uneven_radiuses_{solver}"]["values"])
    err = np.abs(v[::50] - gold).max()
    print(f"branch8_uneven_radiuses {solver} max err vs reference golden {err:.3e}")
    assert err <= 1e-5


@pytest.mark.parametrize("which", ["complex", "mixed"])
def test_golden_net7(device, which):
    """test_complex_net / test_mixed_net (test_basic_modules.py:412-574): seven cells
    (SimpleCell(5, 4); or six single-compartment cells and one SimpleCell(5, 4)) with
    HH, IonotropicSynapse and TestSynapse in one network (the fused kernel's per-synapse
    models), fully_connect with random postsynaptic compartments after np.random.seed(0),
    per-edge conductances, module stimuli into three (six) cells; every 40th step
    against the reference at 1e-8 mV."""
    morph = jx.Cell(jx.Branch(jx.Compartment(), 4), parents=[-1, 0, 0, 1, 1])
    cells = [morph] * 7 if which == "complex" else [jx.Cell()] * 6 + [morph]
    net = jx.Network(cells)
    net.insert(jx.HH())
    np.random.seed(0)
    for pre, post in (([0, 1, 2], [3, 4, 5]), ([3, 4, 5], 6)):
        jx.fully_connect(net.cell(pre), net.cell(post), jx.IonotropicSynapse(), random_post_comp=True)
        jx.fully_connect(net.cell(pre), net.cell(post), jx.TestSynapse(), random_post_comp=True)
    f = 100_000.0 / (2 * np.pi * 10.0)
    net.set("IonotropicSynapse_gS", 0.44 / f)
    net.set("TestSynapse_gC", 0.62 / f)
    net.IonotropicSynapse.edge([0, 2, 4]).set("IonotropicSynapse_gS", 0.32 / f)
    net.TestSynapse.edge([0, 3, 5]).set("TestSynapse_gC", 0.24 / f)
    cur = jx.step_current(0.5, 0.5, 0.1, 0.025, 10.0).to(device)
    for i in range(3 if which == "complex" else 6):
        net.cell(i).branch(0).loc(0.0).stimulate(cur)
    net.cell(6).branch(0).loc(0.0).record()
    v = jx.integrate(net, delta_t=0.025).double().cpu().numpy()
    gold = np.asarray(GOLD[f"net7_{which}"]["values"])
    err = np.abs(v[..., ::40] - gold).max()
    print(f"net7_{which} max err vs reference golden {err:.3e}")
    assert err <= 1e-8, err


def test_golden_network_gradient(device):
    """tests/jaxley_identical/test_grad.py:24-105 through the product path: the complex
    network, `make_trainable` on the module, on cells, on a synapse type and on two
    synapses, `get_parameters`, loss = sum of every 40th recorded voltage; the gradient
    runs through the per-step path (the HIP tree-solve operator and its backward kernel,
    synapse steps as tensor operations). Loss within 1e-8, gradients within 1e-3."""
    morph = jx.Cell(jx.Branch(jx.Compartment(), 4), parents=[-1, 0, 0, 1, 1])
    net = jx.Network([morph] * 7)
    net.insert(jx.HH())
    np.random.seed(0)
    for pre, post in (([0, 1, 2], [3, 4, 5]), ([3, 4, 5], 6)):
        jx.fully_connect(net.cell(pre), net.cell(post), jx.IonotropicSynapse(), random_post_comp=True)
        jx.fully_connect(net.cell(pre), net.cell(post), jx.TestSynapse(), random_post_comp=True)
    f = 100_000.0 / (2 * np.pi * 10.0)
    net.set("IonotropicSynapse_gS", 0.44 / f)
    net.set("TestSynapse_gC", 0.62 / f)
    net.IonotropicSynapse.edge([0, 2, 4]).set("IonotropicSynapse_gS", 0.32 / f)
    net.TestSynapse.edge([0, 3, 5]).set("TestSynapse_gC", 0.24 / f)
    cur = jx.step_current(0.5, 0.5, 0.1, 0.025, 10.0).to(device)
    for i in range(3):
        net.cell(i).branch(0).loc(0.0).stimulate(cur)
    net.cell(6).branch(0).loc(0.0).record()
    net.make_trainable("HH_gNa")
    net.cell([0, 1, 4]).make_trainable("HH_gK")
    net.cell("all").make_trainable("HH_gLeak")
    net.IonotropicSynapse.make_trainable("IonotropicSynapse_gS")
    net.TestSynapse.edge([0, 2]).make_trainable("TestSynapse_gC")
    params = [{k: v.to(device).requires_grad_()} for p in net.get_parameters() for k, v in p.items()]
    loss = jx.integrate(net, params=params, voltage_solver="jaxley.dhs.cpu")[0, ::40].sum()
    loss.backward()
    print(f"network gradient golden: loss {loss.item():.10f}")
    assert abs(loss.item() - (-592.26920963)) <= 1e-8
    gold = [[-1420.82493379], [36.76796492, 14.50837022, 127.31361882],
            [-1511.83061117, -657.06543355, -276.59064858, -348.48049676, -10613.6754718,
             -1240.62600873, -268475.23225046], [-149134.15237682], [-75.34438431, -163.50398586]]
    for p, g in zip(params, gold):
        (key, val), = p.items()
        err = np.abs(val.grad.cpu().numpy() - np.asarray(g)).max()
        print(f"  {key}: max grad err {err:.3e}")
        assert err <= 1e-3, key


def _rl_geometry(seed, nb):
    np.random.seed(seed)
    r1 = 5 * np.random.rand(2 * nb)
    r2 = np.random.rand(2 * nb)
    return [np.flip(r1[2 * b:2 * b + 2]) for b in range(nb)], [np.flip(r2[2 * b:2 * b + 2]) for b in range(nb)]


def test_golden_radius_and_length_branch_cell(device):
    """test_radius_and_length.py:63-146: a 2-compartment branch and a SimpleCell(3, 2)
    with random (flipped) lengths and radii per branch."""
    branch = jx.Branch(jx.Compartment(), 2)
    (l,), (r,) = _rl_geometry(1, 1)
    branch.set("length", l)
    branch.set("radius", r)
    branch.insert(jx.HH())
    _check(_run(branch, branch.loc(0.0), branch.loc(0.0), device)[0], "radius_and_length_branch")
    cell = jx.Cell(jx.Branch(jx.Compartment(), 2), parents=[-1, 0, 0])
    ls, rs = _rl_geometry(1, 3)
    for b in range(3):
        cell.branch(b).set("length", ls[b])
        cell.branch(b).set("radius", rs[b])
    cell.insert(jx.HH())
    _check(_run(cell, cell.branch(1).loc(0.0), cell.branch(1).loc(0.0), device)[0], "radius_and_length_cell")


def test_golden_radius_and_length_net(device):
    """test_radius_and_length.py:149-227: two SimpleCell(3, 2) with their own random
    geometry, an IonotropicSynapse from cell 0 to cell 1 (branch 0, loc 0) with gS =
    0.5 / (1e5 / area of the postsynaptic compartment)."""
    cell = jx.Cell(jx.Branch(jx.Compartment(), 2), parents=[-1, 0, 0])
    net = jx.Network([cell, cell])
    for c, seed in ((0, 1), (1, 2)):
        ls, rs = _rl_geometry(seed, 3)
        for b in range(3):
            net.cell(c).branch(b).set("length", ls[b])
            net.cell(c).branch(b).set("radius", rs[b])
    jx.connect(net.cell(0).branch(0).loc(0.0), net.cell(1).branch(0).loc(0.0), jx.IonotropicSynapse())
    net.insert(jx.HH())
    ls, rs = _rl_geometry(2, 3)
    area = 2 * np.pi * ls[0][0] * rs[0][0]
    net.set("IonotropicSynapse_gS", 0.5 / (100_000.0 / area))
    cur = jx.step_current(0.5, 1.0, 0.02, 0.025, 5.0).to(device)
    ds = None
    for c in (0, 1):
        net.cell(c).branch(1).loc(0.0).record()
        ds = net.cell(c).branch(1).loc(0.0).data_stimulate(cur, ds)
    v = jx.integrate(net, data_stimuli=ds, delta_t=0.025).double().cpu().numpy()
    _check(v, "radius_and_length_net")
