"""The oracle pinned against the reference's golden traces, closed forms and networkx."""
import json
import math
from pathlib import Path

import networkx as nx
import numpy as np
import pytest
import torch

from oracle import jx_oracle as O

import closed_forms as CF

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_traces.json").read_text())
CUR = O.step_current(0.5, 1.0, 0.02, 0.025, 5.0)


def _check(v, key, tol=1e-8):
    gold = np.asarray(GOLD[key]["values"])
    err = np.abs(v[..., ::20].numpy() - gold).max()
    assert err <= tol, (key, err)


def test_step_current_window():
    # stimulus.py:8-36: int(t_max // dt) + 2 samples (5.0 // 0.025 == 199.0 in binary
    # floating point), on for [int(0.5 / dt), int(1.5 / dt)) = [20, 60).
    assert CUR.shape == (201,)
    assert np.nonzero(CUR.numpy())[0].tolist() == list(range(20, 60))


def test_golden_compartment():
    m = O.OracleModel([([-1], [1])])
    m.hh_comps = np.array([0])
    v = O.simulate(m, None, 0.025, stim=CUR[None], stim_comps=[0], rec_comps=[0])
    _check(v[0], "compartment")


def test_golden_branch():
    m = O.OracleModel([([-1], [2])])
    m.hh_comps = np.arange(2)
    v = O.simulate(m, None, 0.025, stim=CUR[None], stim_comps=[0], rec_comps=[0])
    _check(v[0], "branch2")


@pytest.mark.parametrize("solver", ["jaxley.dhs.gpu", "jaxley.dhs.cpu", "jax.sparse"])
def test_golden_cell(solver):
    m = O.OracleModel([([-1, 0, 0], [2, 2, 2])])
    m.hh_comps = np.arange(6)
    v = O.simulate(m, None, 0.025, stim=CUR[None], stim_comps=[2], rec_comps=[2], voltage_solver=solver)
    _check(v[0], "cell_3_2")


def test_golden_radius_and_length():
    np.random.seed(1)
    m = O.OracleModel([([-1], [1])])
    m.hh_comps = np.array([0])
    m.length = 5 * np.random.rand(1)
    m.radius = np.random.rand(1)
    v = O.simulate(m, None, 0.025, stim=CUR[None], stim_comps=[0], rec_comps=[0])
    _check(v[0], "radius_and_length_compartment")


def test_golden_network_ionotropic():
    m = O.OracleModel([([-1, 0, 0], [2, 2, 2])] * 2)
    m.hh_comps = np.arange(12)
    gs = 0.5 / (100000.0 / (2 * np.pi * 10.0 * 1.0))
    p = dict(O.SYN_DEFAULTS["IonotropicSynapse"], gS=gs)
    # cell 1's first compartment is node 7 (6 compartments + 1 branchpoint in cell 0).
    m.synapses = [{"type": "IonotropicSynapse", "pre": np.array([0]), "post": np.array([7]),
                   "params": {k: np.array([v]) for k, v in p.items()}, "s0": np.array([0.2])}]
    v = O.simulate(m, None, 0.025, stim=torch.stack([CUR, CUR]), stim_comps=[2, 8], rec_comps=[2, 8])
    _check(v, "net_2x_cell_3_2")


def test_ragged_indexer_solve_vs_dense():
    """The hand-built ragged indexer of the reference's tests/test_solver.py:78-104."""
    ind = {"node_order_grouped": [np.array([[1, 0], [2, 0]]), np.array([[3, 1]])],
           "parent_lookup": np.array([-1, 0, 0, 1, -1])}
    d = torch.tensor([4.0, 5.0, 6.0, 7.0, 1.0], dtype=torch.float64)
    l = torch.tensor([0.0, -0.4, -0.3, -0.2, 0.0], dtype=torch.float64)
    u = torch.tensor([0.0, -0.5, -0.1, -0.6, 0.0], dtype=torch.float64)
    b = torch.tensor([1.0, 2.0, 3.0, 4.0, 0.0], dtype=torch.float64)
    A = O.dense_matrix(d, l, u, ind["parent_lookup"])
    ref = np.linalg.solve(A, b.numpy())
    for gpu in (False, True):
        x = O.DHSSolverOracle(ind, optimize_for_gpu=gpu).solve(d, l, u, b)
        np.testing.assert_allclose(x.numpy(), ref, rtol=1e-12)


def _nx_order(tables, allowed):
    """The reference's own construction with networkx (to_graph edges -> to_undirected
    -> bfs_edges) for comparison with the oracle's restatement."""
    g = nx.DiGraph()
    g.add_nodes_from(range(tables["n_nodes"]))
    typ, s, t = tables["type"], tables["source"], tables["sink"]
    keep = np.isin(typ, [2, 3]) | ((typ == 0) & (s < t))
    g.add_edges_from(np.stack([s[keep], t[keep]], 1).tolist())
    ug = g.to_undirected()
    dist = nx.single_source_shortest_path_length(ug, 0)
    out, depth, since, last = [], 0, 0, 0
    for u_, v_ in nx.bfs_edges(ug, 0):
        if dist[v_] > last:
            depth, since, last = depth + 1, 0, dist[v_]
        elif since >= allowed:
            depth, since = depth + 1, 0
        since += 1
        out.append((u_, v_, depth))
    return out


@pytest.mark.parametrize("seed", range(12))
def test_bfs_order_matches_networkx(seed):
    rng = np.random.default_rng(seed)
    nb = int(rng.integers(1, 60))
    parents = [-1] + [int(rng.integers(0, k)) for k in range(1, nb)]
    ncomp = rng.integers(1, 6, size=nb)
    tab = O.cell_tables(parents, ncomp)
    for allowed in (1, 4, 32):
        ours = O.bfs_edge_hops(O._undirected_adjacency(tab["n_nodes"], tab), 0, allowed)
        assert ours == _nx_order(tab, allowed)


def test_random_tree_solve_vs_dense():
    rng = np.random.default_rng(0)
    parents = [-1] + [int(rng.integers(0, k)) for k in range(1, 30)]
    tab = O.cell_tables(parents, [3] * 30)
    ind = O.dhs_indexer_cell(tab, 32)
    n = tab["n_nodes"] + 1
    pl = ind["parent_lookup"]
    d = torch.tensor(4 + rng.uniform(size=n))
    l = torch.tensor(-rng.uniform(size=n))
    u = torch.tensor(-rng.uniform(size=n))
    b = torch.tensor(rng.normal(size=n))
    for gpu in (False, True):
        x = O.DHSSolverOracle(ind, optimize_for_gpu=gpu).solve(d, l, u, b)
        np.testing.assert_allclose(x.numpy(), np.linalg.solve(O.dense_matrix(d, l, u, pl), b.numpy()), rtol=1e-10)


KIN = {k: torch.tensor(v, dtype=torch.float64) for k, v in O.SYN_DEFAULTS["KineticSynapse"].items()}
SIG = {k: torch.tensor(v, dtype=torch.float64) for k, v in O.SYN_DEFAULTS["SigmoidRateSynapse"].items()}


@pytest.mark.parametrize("v_pre", [-70.0, -21.0, -10.0, 30.0])
def test_kinetic_exact_at_constant_presynaptic_voltage(v_pre):
    """Exponential Euler solves the linear ODE of a KineticSynapse exactly while v_pre is
    constant: 400 steps land on the analytic s(t) at t = 400 dt (a wrong rate, sign or
    steady state breaks it)."""
    dt, s = 0.025, torch.tensor(0.05, dtype=torch.float64)
    for _ in range(400):
        s = O.kinetic_update(s, torch.tensor(v_pre, dtype=torch.float64), KIN, dt)
    np.testing.assert_allclose(float(s), CF.kinetic_exact(0.05, v_pre, 400 * dt), rtol=1e-12)


def test_kinetic_fixed_point():
    v = torch.tensor(-12.0, dtype=torch.float64)
    a = 1.0 * CF.logistic((-12.0 + 20.0) / 2.0)
    s_inf = torch.tensor(a / (a + 0.2), dtype=torch.float64)
    np.testing.assert_allclose(float(O.kinetic_update(s_inf, v, KIN, 0.025)), float(s_inf), rtol=1e-14)


def test_kinetic_first_order_in_time():
    """With a time-varying v_pre (held at its value at the start of each step, as the
    reference's step does) the state converges to the ODE's solution (classical RK4 at
    dt / 1000) at first order: halving dt halves the error."""
    vp = lambda t: -20.0 + 15.0 * math.sin(2.0 * t)
    t_end = 4.0
    errs = []
    for dt in (0.05, 0.025, 0.0125):
        s = torch.tensor(0.1, dtype=torch.float64)
        for i in range(int(round(t_end / dt))):
            s = O.kinetic_update(s, torch.tensor(vp(i * dt), dtype=torch.float64), KIN, dt)
        errs.append(abs(float(s) - CF.kinetic_rk4(0.1, vp, t_end, dt / 1000)))
    assert errs[0] > 1e-5
    for a, b in zip(errs, errs[1:]):
        assert 1.8 < a / b < 2.2, errs


def test_sigmoid_rate_linearization():
    """SigmoidRateSynapse: at v_pre = x_offset the current is -gS / 2 (logistic(0) = 1/2);
    currents mirrored about x_offset sum to -gS (logistic(-x) = 1 - logistic(x)); the
    finite-difference slope (both voltages perturbed, network.py:452-457) equals the
    derivative -gS slope s (1 - s) 1e5 / area to O(diff); nothing depends on v_post."""
    gS, area = 2e-3, 150.0
    prm = dict(SIG, gS=torch.tensor(gS, dtype=torch.float64))
    x0 = float(SIG["x_offset"])
    i_mid = O.sigmoid_rate_current(None, torch.tensor(x0, dtype=torch.float64), None, prm)
    np.testing.assert_allclose(float(i_mid), -gS / 2, rtol=1e-15)
    for u in (0.3, 7.0, 25.0):
        up = O.sigmoid_rate_current(None, torch.tensor(x0 + u, dtype=torch.float64), None, prm)
        dn = O.sigmoid_rate_current(None, torch.tensor(x0 - u, dtype=torch.float64), None, prm)
        np.testing.assert_allclose(float(up + dn), -gS, rtol=1e-14)
    for v_pre in (-75.0, -60.0, -52.0):
        vts = []
        for v_post in (-80.0, -30.0):
            _, vt, ct, _ = O.synapse_linearization("SigmoidRateSynapse", None, torch.tensor(v_pre, dtype=torch.float64),
                                                torch.tensor(v_post, dtype=torch.float64), prm,
                                                torch.tensor(area, dtype=torch.float64), 0.025)
            vts.append(float(vt))
            d0 = -gS * CF.logistic((v_pre - x0) * 0.25) * 1e5 / area
            np.testing.assert_allclose(float(vt) * v_post + float(ct), d0, rtol=1e-12)
        sg = CF.logistic((v_pre - x0) * 0.25)
        deriv = -gS * 0.25 * sg * (1 - sg) * 1e5 / area
        assert vts[0] == vts[1]
        assert abs(vts[0] - deriv) <= 1e-3 * abs(deriv) + 1e-12


def test_kinetic_linearization_is_the_conductance():
    """KineticSynapse: i = gS s (v_post - e_syn) is linear in v_post and independent of
    v_pre, so the finite difference returns the conductance gS s 1e5 / area exactly (up
    to round-off) and the state advances before the current is taken."""
    gS, area, e_syn = 5e-4, 80.0, -80.0
    prm = dict(KIN, gS=torch.tensor(gS, dtype=torch.float64), e_syn=torch.tensor(e_syn, dtype=torch.float64))
    s0, v_pre, v_post = 0.3, -5.0, -61.0
    s, vt, ct, _ = O.synapse_linearization("KineticSynapse", torch.tensor(s0, dtype=torch.float64),
                                        torch.tensor(v_pre, dtype=torch.float64),
                                        torch.tensor(v_post, dtype=torch.float64), prm,
                                        torch.tensor(area, dtype=torch.float64), 0.025)
    s_new = CF.kinetic_exact(s0, v_pre, 0.025)
    np.testing.assert_allclose(float(s), s_new, rtol=1e-13)
    np.testing.assert_allclose(float(vt), gS * s_new * 1e5 / area, rtol=1e-9)
    np.testing.assert_allclose(float(ct), -gS * s_new * 1e5 / area * e_syn, rtol=1e-9)


def _two_cells(kind, prm, s0=None):
    """Two one-compartment cells (length 10 um, radius 1 um), Leak in both, cell 0 at
    -10 mV (its Leak reversal), a synapse 0 -> 1."""
    om = O.OracleModel([([-1], [1])] * 2)
    om.leak_comps = np.array([0, 1])
    om.leak_params = {"Leak_gLeak": np.array([1e-4, 1e-4]), "Leak_eLeak": np.array([-10.0, -65.0])}
    om.v0 = np.array([-10.0, -70.0])
    om.synapses = [{"type": kind, "pre": om.internal[[0]], "post": om.internal[[1]],
                    "params": {k: np.array([v]) for k, v in prm.items()},
                    "s0": None if s0 is None else np.array([s0])}]
    return om


def test_kinetic_network_closed_form():
    """The oracle's network step with a KineticSynapse (state update, current, the
    finite-difference linearization, the 1e5 / area scaling and the implicit step)
    against the closed-form recurrence of the postsynaptic membrane."""
    prm = dict(O.SYN_DEFAULTS["KineticSynapse"], gS=1e-3, e_syn=0.0)
    om = _two_cells("KineticSynapse", prm, s0=0.2)
    v = O.simulate(om, 10.0, 0.025, rec_comps=[0, 1])
    ref = CF.two_cell_kinetic(400, 0.025, -10.0, -70.0, 0.2, 1e-3, 0.0, 1e-4, -65.0,
                              CF.cylinder_area(10.0, 1.0))
    np.testing.assert_allclose(v[1].numpy(), ref, rtol=0, atol=1e-10)
    np.testing.assert_allclose(v[0].numpy(), -10.0, rtol=0, atol=1e-12)
    assert ref[-1] > -30.0          # the synapse moved the cell far from rest


def test_sigmoid_rate_network_closed_form():
    prm = dict(O.SYN_DEFAULTS["SigmoidRateSynapse"], gS=2e-3, x_offset=-12.0)
    om = _two_cells("SigmoidRateSynapse", prm)
    v = O.simulate(om, 10.0, 0.025, rec_comps=[1])
    ref = CF.two_cell_sigmoid_rate(400, 0.025, -10.0, -70.0, 2e-3, -12.0, 0.25, 1e-4, -65.0,
                                   CF.cylinder_area(10.0, 1.0))
    np.testing.assert_allclose(v[0].numpy(), ref, rtol=0, atol=1e-10)
    assert ref[-1] > -60.0          # gS > 0 depolarizes (synapses/sigmoid_rate.py:15-17)


def test_x_over_expm1_removable_singularity():
    """tests/test_float32_safety.py:25-73: alpha_m(-40) = 1, alpha_n(-55) = 0.1."""
    (am, _), _, (an, _) = O.hh_rates(torch.tensor([-40.0, -55.0], dtype=torch.float64))
    np.testing.assert_allclose(float(am[0]), 1.0, rtol=1e-12)
    np.testing.assert_allclose(float(an[1]), 0.1, rtol=1e-12)


# Crank-Nicolson (modules/base.py:3327-3366): the implicit solve with dt / 2, then
# v' = 2 x - v.

@pytest.mark.parametrize("solver", ["bwd_euler", "crank_nicolson", "fwd_euler"])
def test_time_stepper_closed_form(solver):
    """A passive compartment, dv/dt = -(v - e) / tau: backward Euler multiplies v - e by
    1 / (1 + a) per step, Crank-Nicolson by (1 - a / 2) / (1 + a / 2) and forward Euler
    by 1 - a, a = dt / tau."""
    m = O.OracleModel([([-1], [1])])
    m.leak_comps = np.array([0])
    m.leak_params["Leak_gLeak"][:] = 1e-3      # tau = cm / (1000 gLeak) = 1 ms
    m.v0[:] = -50.0
    v = O.simulate(m, 10.0, 0.025, solver=solver)[0].numpy()
    a = 0.025 * 1000 * 1e-3
    R = {"bwd_euler": 1 / (1 + a), "crank_nicolson": (1 - a / 2) / (1 + a / 2),
         "fwd_euler": 1 - a}[solver]
    exact = -70.0 + 20.0 * R ** np.arange(len(v))
    assert np.abs(v - exact).max() < 1e-10


def _passive_cell(solver, dt):
    m = O.OracleModel([([-1, 0, 0], [2, 2, 2])])
    m.leak_comps = np.arange(6)
    m.leak_params["Leak_gLeak"][:] = 1e-3
    m.length[:] = 50.0
    m.axial_resistivity[:] = 200.0
    cur = O.step_current(0.5, 1.0, 0.05, dt, 5.0)
    v = O.simulate(m, 5.0, dt, stim=torch.as_tensor(cur)[None], stim_comps=[2], rec_comps=[0, 5],
                   solver=solver).numpy()
    return np.stack([v[:, int(round(t / dt))] for t in (1.0, 2.0, 4.0)], -1)


def test_crank_nicolson_second_order_on_a_branched_cell():
    """On a passive branched cell (branchpoint included) Crank-Nicolson converges with
    order 2, backward and forward Euler (branchpoints at the conductance-weighted mean of
    their neighbours) with order 1, all to the same solution: halving dt divides the
    error by about 4 and 2, and backward Euler extrapolated by Richardson
    (2 v(h/2) - v(h)) meets the finest Crank-Nicolson run."""
    fine = _passive_cell("crank_nicolson", 0.025 / 16)
    # Forward Euler is stable on this cell for dt <= 0.05 and asymptotic below it.
    for solver, lo, hi, dts in (("bwd_euler", 1.7, 2.3, (0.1, 0.05, 0.025)),
                                ("crank_nicolson", 3.3, 4.6, (0.1, 0.05, 0.025)),
                                ("fwd_euler", 1.8, 2.2, (0.025, 0.0125, 0.00625))):
        errs = [np.abs(_passive_cell(solver, dt) - fine).max() for dt in dts]
        for e0, e1 in zip(errs, errs[1:]):
            assert lo < e0 / e1 < hi, (solver, errs)
    h = 0.025 / 8
    rich = 2 * _passive_cell("bwd_euler", h / 2) - _passive_cell("bwd_euler", h)
    assert np.abs(rich - fine).max() < 1e-4


def test_reference_solver_comparison_network():
    """tests/test_solver.py:29-58 of the reference: two SimpleCell(3, 4) with HH, an
    IonotropicSynapse from cell 0 (branch 0, comp 0) to cell 1 (branch 0, comp 0), gS
    0.01, length 50 um, r_a 200; backward Euler and Crank-Nicolson stay within 6 mV."""
    m = O.OracleModel([([-1, 0, 0], [4, 4, 4])] * 2)
    m.hh_comps = np.arange(24)
    m.length[:] = 50.0
    m.axial_resistivity[:] = 200.0
    p = dict(O.SYN_DEFAULTS["IonotropicSynapse"], gS=0.01)
    # cell 1's first compartment is node 13 (12 compartments + 1 branchpoint in cell 0).
    m.synapses = [{"type": "IonotropicSynapse", "pre": np.array([0]), "post": np.array([13]),
                   "params": {k: np.array([v]) for k, v in p.items()}, "s0": np.array([0.2])}]
    cur = O.step_current(0.5, 1.0, 0.1, 0.025, 5.0)
    out = {s: O.simulate(m, None, 0.025, stim=cur[None], stim_comps=[0], rec_comps=[15], solver=s).numpy()
           for s in ("bwd_euler", "crank_nicolson")}
    d = np.abs(out["bwd_euler"] - out["crank_nicolson"]).max()
    assert 0 < d < 6.0


def _branch8_uneven(solver, voltage_solver="jaxley.dhs.gpu"):
    m = O.OracleModel([([-1], [8])])
    m.hh_comps = np.arange(8)
    m.axial_resistivity[:] = 500.0
    m.length = np.linspace(20, 300, 8)
    m.radius = np.linspace(1, 5, 8)
    cur = O.step_current(0.5, 1.0, 2.0, 0.025, 10.0)
    return O.simulate(m, None, 0.025, stim=cur[None], stim_comps=[7], rec_comps=[0], solver=solver,
                      voltage_solver=voltage_solver)[0]


@pytest.mark.parametrize("solver,voltage_solver", [("fwd_euler", "jaxley.dhs"), ("bwd_euler", "jaxley.stone")])
def test_golden_branch_uneven_radiuses(solver, voltage_solver):
    """test_basic_modules.py:124-197: every 50th step, 1e-5 mV."""
    v = _branch8_uneven(solver, voltage_solver)
    gold = np.asarray(GOLD[f"branch8_uneven_radiuses_{solver}"]["values"])
    assert np.abs(v[::50].numpy() - gold).max() <= 1e-5


def _net7_oracle(cells):
    """test_complex_net / test_mixed_net (test_basic_modules.py:412-574): cells 0-2 ->
    3-5 and 3-5 -> 6 with IonotropicSynapse and TestSynapse by fully_connect with random
    postsynaptic compartments (np.random.seed(0); pandas' groupby-sample draws, per
    postsynaptic cell in ascending order, three compartments with replacement)."""
    m = O.OracleModel(cells)
    nc = [int(np.sum(nb)) for _, nb in cells]
    first = np.concatenate([[0], np.cumsum(nc)])
    m.hh_comps = np.arange(first[-1])
    np.random.seed(0)

    def fully(pre_cells, post_cells):
        drawn = [first[c] + np.random.choice(nc[c], len(pre_cells), replace=True) for c in post_cells]
        post = np.reshape(np.concatenate(drawn), (len(pre_cells), len(post_cells))).T.flatten()
        pre = np.repeat(first[pre_cells], len(post_cells))
        return pre, post

    groups = {"IonotropicSynapse": [], "TestSynapse": []}
    for pre_cells, post_cells in (([0, 1, 2], [3, 4, 5]), ([3, 4, 5], [6])):
        for name in ("IonotropicSynapse", "TestSynapse"):
            groups[name].append(fully(np.array(pre_cells), np.array(post_cells)))
    f = 100_000.0 / (2 * np.pi * 10.0)
    vals = {"IonotropicSynapse": ("gS", 0.44 / f, [0, 2, 4], 0.32 / f),
            "TestSynapse": ("gC", 0.62 / f, [0, 3, 5], 0.24 / f)}
    m.synapses = []
    for name, parts in groups.items():
        pre = np.concatenate([p for p, _ in parts])
        post = np.concatenate([q for _, q in parts])
        key, base, edges, val = vals[name]
        p = {k: np.full(len(pre), v) for k, v in O.SYN_DEFAULTS[name].items()}
        p[key][:] = base
        p[key][edges] = val
        m.synapses.append({"type": name, "pre": m.internal[pre], "post": m.internal[post],
                           "params": p, "s0": np.full(len(pre), 0.2)})
    return m, first


@pytest.mark.parametrize("which", ["complex", "mixed"])
def test_golden_net7(which):
    morph = ([-1, 0, 0, 1, 1], [4] * 5)
    cells = [morph] * 7 if which == "complex" else [([-1], [1])] * 6 + [morph]
    m, first = _net7_oracle(cells)
    cur = O.step_current(0.5, 0.5, 0.1, 0.025, 10.0)
    n_stim = 3 if which == "complex" else 6
    v = O.simulate(m, None, 0.025, stim=torch.stack([cur] * n_stim), stim_comps=list(first[:n_stim]),
                   rec_comps=[first[6]])
    gold = np.asarray(GOLD[f"net7_{which}"]["values"])
    err = np.abs(v[..., ::40].numpy() - gold).max()
    assert err <= 1e-8, err


def test_golden_network_gradient():
    """tests/jaxley_identical/test_grad.py:24-105: the complex network (seven
    SimpleCell(5, 4)), loss = sum of every 40th recorded voltage of cell 6, gradients
    with respect to a shared gNa, gK of cells 0, 1, 4, gLeak per cell, one shared
    IonotropicSynapse gS (initial value the mean over its synapses) and the TestSynapse gC
    of edges 0 and 2; loss within 1e-8, gradients within 1e-3 (absolute)."""
    m, first = _net7_oracle([([-1, 0, 0, 1, 1], [4] * 5)] * 7)
    nc = 20
    f = 100_000.0 / (2 * np.pi * 10.0)
    g_na = torch.tensor([0.12], dtype=torch.float64, requires_grad=True)
    g_k = torch.tensor([0.036] * 3, dtype=torch.float64, requires_grad=True)
    g_leak = torch.tensor([0.0003] * 7, dtype=torch.float64, requires_grad=True)
    iono, test = m.synapses
    g_s = torch.tensor([float(np.mean(iono["params"]["gS"]))], dtype=torch.float64, requires_grad=True)
    g_c = torch.tensor([0.24 / f, 0.62 / f], dtype=torch.float64, requires_grad=True)
    gk_cell = torch.full((7,), 0.036, dtype=torch.float64).index_put((torch.tensor([0, 1, 4]),), g_k)
    iono["params"] = dict(iono["params"], gS=g_s.expand(len(iono["pre"])))
    gc = torch.as_tensor(test["params"]["gC"], dtype=torch.float64)
    test["params"] = dict(test["params"], gC=gc.index_put((torch.tensor([0, 2]),), g_c))
    over = {"HH_gNa": g_na.expand(7 * nc), "HH_gK": torch.repeat_interleave(gk_cell, nc),
            "HH_gLeak": torch.repeat_interleave(g_leak, nc)}
    cur = O.step_current(0.5, 0.5, 0.1, 0.025, 10.0)
    v = O.simulate(m, None, 0.025, stim=torch.stack([cur] * 3), stim_comps=list(first[:3]),
                   rec_comps=[first[6]], overrides=over, voltage_solver="jaxley.dhs.cpu")
    loss = v[0, ::40].sum()
    loss.backward()
    assert abs(loss.item() - (-592.26920963)) <= 1e-8
    gold = {"HH_gNa": [-1420.82493379], "HH_gK": [36.76796492, 14.50837022, 127.31361882],
            "HH_gLeak": [-1511.83061117, -657.06543355, -276.59064858, -348.48049676,
                         -10613.6754718, -1240.62600873, -268475.23225046],
            "IonotropicSynapse_gS": [-149134.15237682], "TestSynapse_gC": [-75.34438431, -163.50398586]}
    for key, p in (("HH_gNa", g_na), ("HH_gK", g_k), ("HH_gLeak", g_leak), ("IonotropicSynapse_gS", g_s),
                   ("TestSynapse_gC", g_c)):
        assert np.abs(p.grad.numpy() - np.asarray(gold[key])).max() <= 1e-3, (key, p.grad)


def _rl_geometry(seed, nb):
    """np.random.seed(seed); 2 compartments per branch; per branch the flipped pair of
    5 * rand lengths and rand radii (test_radius_and_length.py:63-227)."""
    np.random.seed(seed)
    r1 = 5 * np.random.rand(2 * nb)
    r2 = np.random.rand(2 * nb)
    length = np.concatenate([np.flip(r1[2 * b:2 * b + 2]) for b in range(nb)])
    radius = np.concatenate([np.flip(r2[2 * b:2 * b + 2]) for b in range(nb)])
    return length, radius


def test_golden_radius_and_length_branch_cell():
    for key, cells, rec in (("radius_and_length_branch", [([-1], [2])], 0),
                            ("radius_and_length_cell", [([-1, 0, 0], [2, 2, 2])], 2)):
        nb = len(cells[0][0])
        m = O.OracleModel(cells)
        m.hh_comps = np.arange(2 * nb)
        m.length, m.radius = _rl_geometry(1, nb)
        v = O.simulate(m, None, 0.025, stim=CUR[None], stim_comps=[rec], rec_comps=[rec])
        _check(v, key)


def test_golden_radius_and_length_net():
    m = O.OracleModel([([-1, 0, 0], [2, 2, 2])] * 2)
    m.hh_comps = np.arange(12)
    l0, r0 = _rl_geometry(1, 3)
    l1, r1 = _rl_geometry(2, 3)
    m.length, m.radius = np.concatenate([l0, l1]), np.concatenate([r0, r1])
    area_post = 2 * np.pi * m.length[6] * m.radius[6]
    p = dict(O.SYN_DEFAULTS["IonotropicSynapse"], gS=0.5 / (100_000.0 / area_post))
    m.synapses = [{"type": "IonotropicSynapse", "pre": m.internal[[0]], "post": m.internal[[6]],
                   "params": {k: np.array([v]) for k, v in p.items()}, "s0": np.array([0.2])}]
    v = O.simulate(m, None, 0.025, stim=torch.stack([CUR, CUR]), stim_comps=[2, 8], rec_comps=[2, 8])
    _check(v, "radius_and_length_net")


def test_recorded_gates_and_currents_are_consistent():
    """Recordings of gates and the HH membrane current (base.py:3445-3502: mA / cm^2 at
    the step's starting voltage with the new gates) obey the backward-Euler step of a
    single HH compartment: v' - v = -dt 1000 (i + g (v' - v)) / cm, g the total HH
    conductance from the recorded gates (a current taken at v', in uA, or with the old
    gates breaks it)."""
    om = O.OracleModel([([-1], [1])])
    om.hh_comps = np.array([0])
    rec = [("v", 0), ("HH_m", 0), ("HH_h", 0), ("HH_n", 0), ("i_HH", 0)]
    cur = O.step_current(0.5, 2.0, 0.05, 0.025, 10.0)[None]
    out = O.simulate(om, 10.0, 0.025, stim=cur * 0, stim_comps=[0], record=rec).numpy()
    v, m, h, n, i = out
    dt = 0.025
    g = 0.12 * m[1:] ** 3 * h[1:] + 0.036 * n[1:] ** 4 + 0.0003
    dv = v[1:] - v[:-1]
    np.testing.assert_allclose(dv, -dt * 1000.0 * (i[1:] + g * dv), rtol=0, atol=1e-11)
    assert i[0] == 0.0 and np.abs(i[1:]).max() > 1e-4
    # The initial gates are the reference's defaults (hh.py:30-34).
    np.testing.assert_allclose([m[0], h[0], n[0]], 0.2)
