"""Synapses on the hot path: `KineticSynapse` (synapses/kinetic.py:12-63),
`SigmoidRateSynapse` (synapses/sigmoid_rate.py:12-50), `IonotropicSynapse`
(synapses/ionotropic.py:13-80) and `TestSynapse` (synapses/test.py:13-59) of
jaxley-pytorch; the last two in the fused forward only. A network may mix types.

As in the reference, a synapse is a point process (conductances in uS, currents in nA);
its arithmetic lives in the HIP kernels (csrc/jxhip_synapse.hip, jxhip_fused.hip).
"""
from typing import Optional


class Synapse:
    kernel_model = None
    param_names = ()
    state_name: Optional[str] = None   # the synapse's state variable (without the prefix)
    synapse_params: dict = {}
    synapse_states: dict = {}

    def __init__(self, name: Optional[str] = None):
        self._name = name if name else self.__class__.__name__

    @property
    def name(self) -> str:
        return self._name


class KineticSynapse(Synapse):
    """First-order kinetic synapse (Destexhe, Mainen & Sejnowski 1994)."""

    kernel_model = 0
    param_names = ("gS", "e_syn", "alpha", "beta", "v_th", "delta")
    state_name = "s"

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        p = self._name
        self.synapse_params = {
            f"{p}_gS": 1e-4, f"{p}_e_syn": 0.0, f"{p}_alpha": 1.0, f"{p}_beta": 0.2,
            f"{p}_v_th": -20.0, f"{p}_delta": 2.0,
        }
        self.synapse_states = {f"{p}_s": 0.0}


class SigmoidRateSynapse(Synapse):
    """Current-based rate synapse, current = -gS * sigmoid((v_pre - x_offset) * slope)."""

    kernel_model = 1
    param_names = ("gS", "x_offset", "slope")

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        p = self._name
        self.synapse_params = {f"{p}_gS": 1e-4, f"{p}_x_offset": -60.0, f"{p}_slope": 0.25}
        self.synapse_states = {}


class IonotropicSynapse(Synapse):
    """Generic ionotropic synapse (Abbott & Marder 1998): s relaxes to
    1 / (1 + exp(-(v_pre - v_th) / delta)) at rate k_minus (1 + exp(u)); current
    gS s (v_post - e_syn)."""

    kernel_model = 2
    param_names = ("gS", "e_syn", "k_minus", "v_th", "delta")
    state_name = "s"

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        p = self._name
        self.synapse_params = {
            f"{p}_gS": 1e-4, f"{p}_e_syn": 0.0, f"{p}_k_minus": 0.025, f"{p}_v_th": -35.0,
            f"{p}_delta": 10.0,
        }
        self.synapse_states = {f"{p}_s": 0.2}


class TestSynapse(Synapse):
    """The reference's test synapse: c relaxes to s_bar = 1 / (1 + exp((v_th - v_pre) /
    delta)) with time constant (1 - s_bar) / k_minus (v_th = -35 mV, delta = 10 mV,
    k_minus = 1/40 per ms); current gC c v_post."""

    __test__ = False   # not a test class for pytest
    kernel_model = 3
    param_names = ("gC",)
    state_name = "c"

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        p = self._name
        self.synapse_params = {f"{p}_gC": 1e-4}
        self.synapse_states = {f"{p}_c": 0.2}


__all__ = ["Synapse", "KineticSynapse", "SigmoidRateSynapse", "IonotropicSynapse", "TestSynapse"]
