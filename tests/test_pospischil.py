"""Pospischil et al. (2008) Na, K and Km channels (channels/pospischil.py:74-250 of
jaxley-pytorch) on the CPU: the oracle's rates against the published forms of the paper
(written here independently, as rational functions of exp instead of the reference's
x / expm1 helper), the fixed points that `init_states` sets, and the reference's
steady-state property (tests/test_channels.py:181-208: a cell started at its fixed
point stays within 0.02 mV over 10 ms) on the oracle."""
import math

import numpy as np
import pytest
import torch

import jaxhip as jx
from models import oracle_of_module, pospischil_cell
from oracle import jx_oracle as O

VOLTS = [-95.3, -80.0, -71.7, -63.2, -50.1, -31.4, -7.9, 12.5, 38.0]


@pytest.mark.parametrize("vt", [-60.0, -52.5])
def test_rates_match_published_forms(vt):
    for v in VOLTS:
        x = v - vt
        am = 0.32 * (x - 13) / (1 - math.exp(-(x - 13) / 4))
        bm = 0.28 * (x - 40) / (math.exp((x - 40) / 5) - 1)
        ah = 0.128 * math.exp(-(x - 17) / 18)
        bh = 4 / (1 + math.exp(-(x - 40) / 5))
        an = 0.032 * (x - 15) / (1 - math.exp(-(x - 15) / 5))
        bn = 0.5 * math.exp(-(x - 10) / 40)
        got = [float(r) for r in O.posp_na_rates(torch.tensor(v, dtype=torch.float64), vt)]
        got += [float(r) for r in O.posp_k_rates(torch.tensor(v, dtype=torch.float64), vt)]
        np.testing.assert_allclose(got, [am, bm, ah, bh, an, bn], rtol=1e-12)
        pinf, tau = O.posp_km_gate(torch.tensor(v, dtype=torch.float64), 4000.0)
        np.testing.assert_allclose(float(pinf), 1 / (1 + math.exp(-(v + 35) / 10)), rtol=1e-13)
        np.testing.assert_allclose(float(tau), 4000.0 / (3.3 * math.exp((v + 35) / 20) + math.exp(-(v + 35) / 20)),
                                   rtol=1e-13)


def test_gate_updates_closed_form():
    """Exponential Euler is exact for a gate at constant voltage: x(t) = x_inf + (x0 -
    x_inf) exp(-t / tau), for Na m and Km p over 40 steps."""
    v, dt = torch.tensor(-55.0, dtype=torch.float64), 0.025
    am, bm, _, _ = O.posp_na_rates(v, -60.0)
    pinf, tau = O.posp_km_gate(v, 4000.0)
    m, p = torch.tensor(0.3, dtype=torch.float64), torch.tensor(0.9, dtype=torch.float64)
    for _ in range(40):
        m = O.gate_exponential(m, dt, am, bm)
        p = O.inf_gate_exponential(p, dt, pinf, tau)
    minf, tm = am / (am + bm), 1 / (am + bm)
    np.testing.assert_allclose(float(m), float(minf + (0.3 - minf) * torch.exp(-40 * dt / tm)), rtol=1e-12)
    np.testing.assert_allclose(float(p), float(pinf + (0.9 - pinf) * torch.exp(-40 * dt / tau)), rtol=1e-12)


def test_init_states_fixed_points():
    cell = pospischil_cell()
    cell.set("v", -62.0)
    cell.init_states()
    d = cell._node_data
    v = torch.tensor(-62.0, dtype=torch.float64)
    am, bm, ah, bh = (float(x) for x in O.posp_na_rates(v, -60.0))
    an, bn = (float(x) for x in O.posp_k_rates(v, -60.0))
    np.testing.assert_allclose(d["Na_m"][:4], am / (am + bm), rtol=1e-12)
    np.testing.assert_allclose(d["Na_h"][:4], ah / (ah + bh), rtol=1e-12)
    np.testing.assert_allclose(d["K_n"][4:], an / (an + bn), rtol=1e-12)
    pinf, tau = (float(x) for x in O.posp_km_gate(v, 4000.0))
    np.testing.assert_allclose(d["Km_p"][4], pinf / (pinf + tau), rtol=1e-12)
    assert np.isnan(d["Na_m"][4:]).all() and np.isnan(d["Km_p"][5:]).all()


def test_oracle_steady_state():
    """tests/test_channels.py:181-208 on the oracle (without CaT / CaL)."""
    cell = pospischil_cell()
    cell.set("v", -62.0)
    cell.init_states()
    v = O.simulate(oracle_of_module(cell), 20.0, 0.025, rec_comps=[0]).numpy()
    cell.set("v", float(v[0, -1]))
    cell.init_states()
    v = O.simulate(oracle_of_module(cell), 10.0, 0.025, rec_comps=[0]).numpy()
    assert abs(v[0, 0] - v[0, -1]) < 0.02
