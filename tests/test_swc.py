"""SWC ingest (`jx.read_swc`, io/swc.py:11-68 and io/graph.py of jaxley-pytorch) on the
CPU: the structure the reference's own tests assert (tests/test_swc.py:62-158), closed
forms of the traced frustum geometry, invariance of the geometry under the number of
compartments, and the oracle on the cells read from the reference's SWC files against
the reference's golden traces (tests/jaxley_identical/test_swc.py, tolerance 1e-8).

The SWC files are copies of the reference's tests/swc_files (data fixtures)."""
import json
import math
import os

import numpy as np
import pytest
import torch

import jaxhip as jx
from models import oracle_of_swc
from oracle import jx_oracle as O

HERE = os.path.dirname(__file__)
SWC = os.path.join(HERE, "golden", "swc")
GOLDEN = json.load(open(os.path.join(HERE, "golden", "reference_traces.json")))


def swc(name):
    return os.path.join(SWC, name)


def test_dummy_compartment_length():
    """tests/test_swc.py:62-70: a soma traced by two points with a neurite at each end."""
    cell = jx.read_swc(swc("morph_soma_both_ends.swc"), ncomp=1)
    assert cell.comb_parents.tolist() == [-1, 0, 1]
    assert cell.nodes.length.to_numpy().tolist() == [2.2, 1.0, 2.6]


@pytest.mark.parametrize("file,want", [
    ("morph_3_types.swc", {"soma": 1, "axon": 1, "basal": 1}),
    ("morph_3_types_single_point_soma.swc", {"soma": 1, "axon": 1, "basal": 1}),
    ("morph_ca1_n120.swc", {"soma": 2, "basal": 101, "apical": 53}),
    ("morph_l5pc_with_axon.swc", {"soma": 1, "axon": 128, "basal": 66, "apical": 129}),
])
def test_swc_types(file, want):
    """tests/test_swc.py:110-158 (first pass): compartments per group with one
    compartment per branch and branches cut at tracing interruptions."""
    cell = jx.read_swc(swc(file), ncomp=1, ignore_swc_tracing_interruptions=False)
    for key, n in want.items():
        assert len(getattr(cell, key).nodes) == n, key


def test_variable_radius_closed_forms():
    """A 120 um branch: a cone from r = 15 to 0.2 over 40 um, then a 0.2 um cylinder.
    Area = lateral area of the frusta; the load of a cone is l / (pi r0 r1)."""
    r0, r1, L = 15.0, 0.2, 40.0
    cone = math.pi * (r0 + r1) * math.sqrt((r0 - r1) ** 2 + L**2)
    cyl = lambda l: 2 * math.pi * r1 * l
    one = jx.read_swc(swc("morph_variable_radiuses_within_branch.swc"), ncomp=1).nodes
    np.testing.assert_allclose(one.area.to_numpy(), [cone + cyl(80.0)], rtol=1e-13)
    np.testing.assert_allclose(one.resistive_load_in.to_numpy(),
                               [(L / (r0 * r1) + 20.0 / r1**2) / math.pi], rtol=1e-13)
    np.testing.assert_allclose(one.resistive_load_out.to_numpy(), [60.0 / r1**2 / math.pi], rtol=1e-13)
    np.testing.assert_allclose(one.length.to_numpy(), [120.0])
    three = jx.read_swc(swc("morph_variable_radiuses_within_branch.swc"), ncomp=3).nodes
    rmid = (r0 + r1) / 2
    np.testing.assert_allclose(three.area.to_numpy(), [cone, cyl(40.0), cyl(40.0)], rtol=1e-13)
    np.testing.assert_allclose(three.radius.to_numpy(), [rmid, r1, r1], rtol=1e-13)
    np.testing.assert_allclose(three.resistive_load_in.to_numpy()[0], 20.0 / (r0 * rmid) / math.pi, rtol=1e-13)
    np.testing.assert_allclose(three.resistive_load_out.to_numpy()[0], 20.0 / (rmid * r1) / math.pi, rtol=1e-13)


@pytest.mark.parametrize("file", ["morph_ca1_n120.swc", "morph_ca1_n120_single_point_soma.swc",
                                  "morph_l5pc_with_axon.swc"])
def test_geometry_invariant_under_ncomp(file):
    """Splitting branches into more compartments only inserts interpolated points:
    per branch, membrane area, resistive load and length sum to the same totals."""
    tot = {}
    for ncomp in (1, 4):
        cell = jx.read_swc(swc(file), ncomp=ncomp, max_branch_len=300.0)
        df = cell.nodes
        b = df.global_branch_index.to_numpy()
        nb = b.max() + 1
        assert len(df) == nb * ncomp
        par = cell.comb_parents
        assert par[0] == -1 and np.all(par[1:] < np.arange(1, nb)) and np.all(par[1:] >= 0)
        load = df.resistive_load_in.to_numpy() + df.resistive_load_out.to_numpy()
        tot[ncomp] = [np.bincount(b, w) for w in (df.area.to_numpy(), load, df.length.to_numpy())]
    for a, c in zip(tot[1], tot[4]):
        np.testing.assert_allclose(a, c, rtol=1e-9)


def test_set_radius_makes_cylinders():
    """`set("radius")` turns the selected compartments into cylinders
    (modules/base.py:1316-1344); the others keep their traced geometry."""
    cell = jx.read_swc(swc("morph_ca1_n120_250.swc"), ncomp=2)
    before = cell.nodes.area.to_numpy().copy()
    cell.branch(1).set("radius", 2.0)
    df = cell.nodes
    sel = df.global_branch_index.to_numpy() == 1
    np.testing.assert_allclose(df.area.to_numpy()[sel], 2 * math.pi * 2.0 * df.length.to_numpy()[sel])
    np.testing.assert_array_equal(df.area.to_numpy()[~sel], before[~sel])


def _stim():
    return O.step_current(0.5, 1.0, 0.2, 0.025, 5.0)


@pytest.mark.parametrize("file", ["morph_ca1_n120_single_point_soma", "morph_ca1_n120"])
def test_oracle_swc_cell_golden(file):
    """tests/jaxley_identical/test_swc.py:23-80."""
    cell = jx.read_swc(swc(file + ".swc"), ncomp=2, max_branch_len=300.0)
    comp = int(cell.soma.branch(0).loc(1.0)._comps[0])
    v = O.simulate(oracle_of_swc([cell]), None, 0.025, stim=_stim()[None], stim_comps=[comp],
                   rec_comps=[comp]).numpy()
    golden = np.asarray(GOLDEN["swc_cell_" + file[6:]]["values"])
    assert np.abs(v[:, ::20] - golden).max() <= 1e-8


def swc_net_oracle(cell):
    """Two copies of `cell`, an IonotropicSynapse between their soma.branch(0).loc(1.0)
    compartments with gS = 0.5 / (1e5 / (2 pi r l)) of the postsynaptic one
    (tests/jaxley_identical/test_swc.py:83-130)."""
    om = oracle_of_swc([cell, cell])
    comp = int(cell.soma.branch(0).loc(1.0)._comps[0])
    n1 = cell._n_comp
    post = int(cell.soma.branch(0).comp(1)._comps[0])
    area = 2 * math.pi * cell._node_data["length"][post] * cell._node_data["radius"][post]
    p = dict(O.SYN_DEFAULTS["IonotropicSynapse"], gS=0.5 / (100_000.0 / area))
    om.synapses = [{"type": "IonotropicSynapse", "pre": np.array([om.internal[comp]]),
                    "post": np.array([om.internal[n1 + comp]]),
                    "params": {k: np.array([x]) for k, x in p.items()}, "s0": np.array([0.2])}]
    return om, [comp, n1 + comp]


@pytest.mark.parametrize("file", ["morph_ca1_n120", "morph_ca1_n120_250"])
def test_oracle_swc_net_golden(file):
    cell = jx.read_swc(swc(file + ".swc"), ncomp=2, max_branch_len=300.0)
    om, comps = swc_net_oracle(cell)
    v = O.simulate(om, None, 0.025, stim=torch.stack([_stim(), _stim()]), stim_comps=comps,
                   rec_comps=comps).numpy()
    golden = np.asarray(GOLDEN["swc_net_" + file[6:]]["values"])
    assert np.abs(v[:, ::20] - golden).max() <= 1e-8


@pytest.mark.parametrize("n_cells", [1, 2])
def test_axial_conductances_match_oracle(n_cells):
    """The package's axial conductances of traced cells (in and out loads differ) equal
    the oracle's, edge by edge, for a cell and a network of two copies."""
    from jaxhip.geometry import axial_conductances, comp_geometry

    cells = [jx.read_swc(swc("morph_ca1_n120.swc"), ncomp=2, max_branch_len=300.0) for _ in range(n_cells)]
    mod = cells[0] if n_cells == 1 else jx.Network(cells)
    vals = {k: torch.as_tensor(v) for k, v in mod._node_data.items()}
    t = mod._comp_edges
    g = axial_conductances(t, mod._comp_node, mod._n_nodes, comp_geometry(vals), vals["axial_resistivity"],
                           vals["capacitance"]).numpy()
    om = oracle_of_swc(cells)
    assert np.array_equal(mod._comp_node, om.internal)
    assert all(np.array_equal(np.asarray(t[k]), np.asarray(om.tables[k])) for k in ("source", "sink", "type", "ordered"))
    np.testing.assert_allclose(g, om.params()["g_axial"].numpy(), rtol=1e-13)


def test_single_branch_with_repeated_point():
    """tests/test_swc.py:161-169's morphology: one branch whose last two points coincide
    (a zero-length segment): finite, positive geometry for any compartment count."""
    for ncomp in (1, 3, 4):
        df = jx.read_swc(swc("morph_single_branch.swc"), ncomp=ncomp).nodes
        assert len(df) == ncomp
        for k in ("area", "resistive_load_in", "resistive_load_out", "radius", "length"):
            x = df[k].to_numpy()
            assert np.all(np.isfinite(x)) and np.all(x > 0), (ncomp, k)
        np.testing.assert_allclose(df.radius.to_numpy(), 8.119, rtol=1e-12)
        np.testing.assert_allclose(df.area.sum(), 2 * math.pi * 8.119 * 2.0, rtol=1e-9)


def test_read_swc_options():
    """`min_radius` clips compartment radii (io/graph.py:405-411, morph_attributes.py:
    63-75), `assign_groups=False` creates no groups, `relevant_type_ids` merges neurites
    of ignored types into one branch (io/graph.py:36-55), and an unknown backend raises."""
    base = jx.read_swc(swc("morph_ca1_n120_250.swc"), ncomp=2)
    r = base.nodes.radius.to_numpy()
    clipped = jx.read_swc(swc("morph_ca1_n120_250.swc"), ncomp=2, min_radius=1.0).nodes.radius.to_numpy()
    np.testing.assert_array_equal(clipped, np.maximum(r, 1.0))
    assert np.any(r < 1.0)
    assert jx.read_swc(swc("morph_ca1_n120_250.swc"), ncomp=2, assign_groups=False).groups == {}
    # An unbranched line of basal (3), soma (1), basal points: three branches, or one when
    # no type is relevant.
    assert len(jx.read_swc(swc("morph_soma_both_ends.swc"), ncomp=1).comb_parents) == 3
    one = jx.read_swc(swc("morph_soma_both_ends.swc"), ncomp=1, relevant_type_ids=[])
    assert len(one.comb_parents) == 1
    np.testing.assert_allclose(one.nodes.length.to_numpy(), [2.2 + 1.0 + 2.6])
    with pytest.raises(ValueError):
        jx.read_swc(swc("morph_3_types.swc"), ncomp=1, backend="custom")
