"""Closed forms that pin the synapse models independently of the oracle's code.

Each function is written from the mathematics (numpy scalars, the logistic function
through tanh), not from the oracle's or the kernels' expressions:

* KineticSynapse (synapses/kinetic.py:37-63): ds/dt = alpha T (1 - s) - beta s with the
  transmitter T = logistic((v_pre - v_th) / delta). At a constant v_pre it is linear with
  constant coefficients: s(t) = s_inf + (s0 - s_inf) exp(-(alpha T + beta) t),
  s_inf = alpha T / (alpha T + beta).
* SigmoidRateSynapse (synapses/sigmoid_rate.py:42-50): i = -gS logistic((v_pre -
  x_offset) slope) nA, no state.
* A two-cell network: cell 0 (presynaptic) is one passive compartment held at its Leak
  reversal (no input: its voltage never moves), cell 1 one passive compartment with Leak
  and the synapse from cell 0. Backward Euler of cell 1's membrane equation with the
  synaptic current per unit area (nA / um^2 -> uA / cm^2: x 1e5) is a scalar recurrence.
"""
from __future__ import annotations

import math

import numpy as np


def logistic(x: float) -> float:
    return 0.5 * (1.0 + math.tanh(0.5 * x))


def kinetic_exact(s0: float, v_pre: float, t: float, alpha=1.0, beta=0.2, v_th=-20.0, delta=2.0) -> float:
    a = alpha * logistic((v_pre - v_th) / delta)
    s_inf = a / (a + beta)
    return s_inf + (s0 - s_inf) * math.exp(-(a + beta) * t)


def kinetic_rk4(s0: float, v_pre_of_t, t_end: float, h: float, alpha=1.0, beta=0.2, v_th=-20.0,
                delta=2.0) -> float:
    """Classical Runge-Kutta of ds/dt = alpha T(v_pre(t)) (1 - s) - beta s."""
    f = lambda t, s: alpha * logistic((v_pre_of_t(t) - v_th) / delta) * (1 - s) - beta * s
    s, t = s0, 0.0
    for _ in range(int(round(t_end / h))):
        k1 = f(t, s)
        k2 = f(t + h / 2, s + h / 2 * k1)
        k3 = f(t + h / 2, s + h / 2 * k2)
        k4 = f(t + h, s + h * k3)
        s += h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        t += h
    return s


def cylinder_area(length: float, radius: float) -> float:
    """Lateral area of a cylinder, um^2."""
    return 2.0 * math.pi * radius * length


def two_cell_kinetic(T: int, dt: float, v_pre: float, v0: float, s0: float, gS: float, e_syn: float,
                     gL: float, eL: float, area: float, cm: float = 1.0, alpha=1.0, beta=0.2, v_th=-20.0,
                     delta=2.0) -> np.ndarray:
    """Voltages of cell 1 over T backward-Euler steps (T + 1 values). The state advances
    first (exponential Euler, exact at a constant v_pre); the current gS s (v - e_syn)
    is linear in v, so the implicit step is exact in v."""
    a = alpha * logistic((v_pre - v_th) / delta)
    r = a + beta
    s_inf = a / r
    decay = math.exp(-r * dt)
    s, v = s0, v0
    out = [v]
    for _ in range(T):
        s = s_inf + (s - s_inf) * decay
        g = gS * s * 1e5 / area          # mS / cm^2 (uS / um^2 * 1e5)
        gl = 1000.0 * gL                 # S / cm^2 -> mS / cm^2
        v = (v + dt / cm * (gl * eL + g * e_syn)) / (1.0 + dt / cm * (gl + g))
        out.append(v)
    return np.array(out)


def two_cell_sigmoid_rate(T: int, dt: float, v_pre: float, v0: float, gS: float, x_offset: float,
                          slope: float, gL: float, eL: float, area: float, cm: float = 1.0,
                          diff: float = 1e-3) -> np.ndarray:
    """Voltages of cell 1 over T backward-Euler steps with the reference's linearization
    of the synaptic current: a finite difference that perturbs the presynaptic and the
    postsynaptic voltage together (modules/network.py:452-457). The current does not
    depend on v_post, so the slope comes from v_pre alone:
      d0 = -gS logistic((v_pre - x_offset) slope) 1e5 / area,
      vt = (d(v_pre + diff) - d0) / diff, and the step solves
      v' = v + dt / cm (gL (eL - v') - d0 - vt (v' - v))."""
    d = lambda vp: -gS * logistic((vp - x_offset) * slope) * 1e5 / area
    d0 = d(v_pre)
    vt = (d(v_pre + diff) - d0) / diff
    gl = 1000.0 * gL
    v = v0
    out = [v]
    for _ in range(T):
        v = (v + dt / cm * (gl * eL - d0 + vt * v)) / (1.0 + dt / cm * (gl + vt))
        out.append(v)
    return np.array(out)
