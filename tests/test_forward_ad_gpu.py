"""Forward-mode differentiation (`torch.autograd.forward_ad`) on the GPU: the tree-solve
operator's forward rule (`_DHSSolveForwardAD`, solver_voltage.py:190-224 of
jaxley-pytorch: dx = A^{-1}(db - dA x)) against the forward mode of a dense solve, and
`integrate` with dual inputs (per-step path) against the oracle's forward mode, whose
tangents agree with its reverse mode (checked here on the same run).

Tolerances: 1e-9 relative for the operator, 1e-7 relative (of the largest tangent) for
simulations."""
import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fwAD

import jaxhip as jx
from jaxhip.kernels.tree_solve import make_tree_solve_kernel
from models import DT, build_c4, build_tree, c1_stimulus, c4_stimulus
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["registers", "levels"])
def test_tree_solve_forward_rule_vs_dense(device, mode):
    cell, _ = build_tree(30, seed=5)
    parent = np.asarray(cell._dhs_solve_indexer["parent_lookup"])
    kern = make_tree_solve_kernel(parent, device, mode)
    n = len(parent)
    rng = np.random.default_rng(3)
    rows = 3
    vals = [-rng.uniform(0.1, 1.0, (rows, n)), -rng.uniform(0.1, 1.0, (rows, n)),
            4.0 + rng.uniform(0, 1, (rows, n)), rng.normal(size=(rows, n))]
    tans = [rng.normal(size=(rows, n)) for _ in range(4)]
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=device)
    l, u, d, s = vals
    with fwAD.dual_level():
        D, L, U, Sv = (fwAD.make_dual(t(a), t(b)) for a, b in zip((d, l, u, s), (tans[2], tans[0], tans[1], tans[3])))
        x = kern.fused(D, L, U, Sv)
        dx = fwAD.unpack_dual(x).tangent.cpu().numpy()
    for r in range(rows):
        def dense(dd, ll, uu):
            A = torch.diag(dd)
            has = torch.as_tensor(parent >= 0)
            c = torch.nonzero(has).flatten()
            p = torch.as_tensor(parent)[c]
            A = A.index_put((c, p), ll[c], accumulate=True)
            return A.index_put((p, c), uu[c], accumulate=True)
        with fwAD.dual_level():
            cd = lambda a, b: fwAD.make_dual(torch.tensor(a[r]), torch.tensor(b[r]))
            A = dense(cd(d, tans[2]), cd(l, tans[0]), cd(u, tans[1]))
            xr = torch.linalg.solve(A, cd(s, tans[3]))
            ref = fwAD.unpack_dual(xr).tangent.numpy()
        np.testing.assert_allclose(dx[r], ref, rtol=1e-9, atol=1e-11)


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_cell_forward_mode_vs_oracle(device):
    """Tangent of the recorded voltages along a direction in the per-compartment gNa."""
    nb, T = 8, 150
    cell, om = build_tree(nb, seed=7)
    cell.branch(0).comp(0).record()
    cell.branch(nb - 1).comp(3).record()
    cur = c1_stimulus(T)
    nc = 4 * nb
    gna = torch.full((nc,), 0.12, dtype=torch.float64)
    direction = torch.linspace(0.5, 1.5, nc, dtype=torch.float64)
    ds = cell.branch(0).comp(0).data_stimulate(torch.tensor(cur, dtype=torch.float64, device=device))
    with fwAD.dual_level():
        g = fwAD.make_dual(gna.to(device), direction.to(device))
        v = jx.integrate(cell, param_state=cell.data_set("HH_gNa", g, None), data_stimuli=ds, delta_t=DT)
        tan = fwAD.unpack_dual(v).tangent.cpu().numpy()
        prim = fwAD.unpack_dual(v).primal.cpu().numpy()
    with fwAD.dual_level():
        og = fwAD.make_dual(gna, direction)
        ref = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0, nc - 1],
                         overrides={"HH_gNa": og})
        ref_t = fwAD.unpack_dual(ref).tangent.numpy()
        ref_p = fwAD.unpack_dual(ref).primal.numpy()
    # The oracle's forward mode agrees with its reverse mode.
    g2 = gna.clone().requires_grad_()
    r2 = O.simulate(om, None, DT, stim=torch.tensor(cur)[None], stim_comps=[0], rec_comps=[0, nc - 1],
                    overrides={"HH_gNa": g2})
    w = torch.tensor(np.random.default_rng(0).normal(size=r2.shape))
    (r2 * w).sum().backward()
    assert abs((ref_t * w.numpy()).sum() - (g2.grad * direction).sum().item()) <= 1e-9 * np.abs(ref_t).sum()
    print(f"cell forward mode: primal err {np.abs(prim - ref_p).max():.3e}, tangent rel err {_rel(tan, ref_t):.3e}")
    assert np.abs(prim - ref_p).max() < 1e-6
    assert _rel(tan, ref_t) < 1e-7


def test_network_forward_mode_vs_oracle(device):
    """Kinetic-synapse network: tangent of the somatic voltages along a direction in the
    synaptic conductances gS."""
    n_cells, k, T = 6, 3, 200
    net, om, info = build_c4(n_cells, k, seed=4)
    soma = info["soma_comps"]
    for c in soma:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    cur = c4_stimulus(1, n_cells, T, seed=5)[0] * 8
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(
            torch.tensor(cur[i], dtype=torch.float64, device=device), ds)
    E = n_cells * k
    gs = torch.as_tensor(om.synapses[0]["params"]["gS"], dtype=torch.float64)
    direction = torch.linspace(-1.0, 1.0, E, dtype=torch.float64) * gs.abs().max()
    with fwAD.dual_level():
        g = fwAD.make_dual(gs.to(device), direction.to(device))
        v = jx.integrate(net, param_state=net.data_set("KineticSynapse_gS", g, None), data_stimuli=ds,
                         delta_t=DT)
        tan = fwAD.unpack_dual(v).tangent.cpu().numpy()
    with fwAD.dual_level():
        sp = om.synapses[0]
        sp["params"] = dict(sp["params"], gS=fwAD.make_dual(gs, direction))
        ref = O.simulate(om, None, DT, stim=torch.tensor(cur), stim_comps=list(soma), rec_comps=list(soma))
        ref_t = fwAD.unpack_dual(ref).tangent.numpy()
    print(f"network forward mode: tangent rel err {_rel(tan, ref_t):.3e}, max |tangent| {np.abs(ref_t).max():.3e}")
    assert np.abs(ref_t).max() > 0
    assert _rel(tan, ref_t) < 1e-7


def test_tree_solve_under_torch_compile(device):
    """The solve operator inside `torch.compile(fullgraph=True)` (aot_eager: traced
    through its fake rule and registered backward, no code generation) gives the eager
    values and gradients."""
    cell, _ = build_tree(20, seed=5)
    parent = np.asarray(cell._dhs_solve_indexer["parent_lookup"])
    kern = make_tree_solve_kernel(parent, device)
    n = len(parent)
    rng = np.random.default_rng(4)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=device)
    d0, l0, u0, s0 = (t(4.0 + rng.uniform(0, 1, (2, n))), t(-rng.uniform(0.1, 1, (2, n))),
                      t(-rng.uniform(0.1, 1, (2, n))), t(rng.normal(size=(2, n))))
    w = t(rng.normal(size=(2, n)))

    def f(d, l, u, s):
        return (kern.fused(d, l, u, s) * w).sum()

    outs = []
    for fn in (f, torch.compile(f, backend="aot_eager", fullgraph=True)):
        args = [x.clone().requires_grad_() for x in (d0, l0, u0, s0)]
        val = fn(*args)
        val.backward()
        outs.append((val.detach(), [a.grad for a in args]))
    torch.testing.assert_close(outs[0][0], outs[1][0], rtol=1e-12, atol=0)
    for ga, gb in zip(outs[0][1], outs[1][1]):
        torch.testing.assert_close(ga, gb, rtol=1e-12, atol=1e-15)
