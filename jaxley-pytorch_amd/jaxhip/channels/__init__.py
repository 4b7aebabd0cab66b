"""Ion channels on the hot path: `HH` (channels/hh.py:14-108) and `Leak`
(channels/pospischil.py:39-71) of jaxley-pytorch.

A channel here is a description (parameter and state names with defaults, a kernel id);
its arithmetic lives in the HIP kernels (csrc/jxhip_common.h, jxhip_fused.hip,
jxhip_channels.hip).
"""
from typing import Optional


class Channel:
    """Base class (channels/channel.py): name, parameters and states with defaults."""

    kernel_id = None
    channel_params: dict = {}
    channel_states: dict = {}
    current_name: str = ""

    def __init__(self, name: Optional[str] = None):
        self._name = name if name else self.__class__.__name__
        self.current_is_in_mA_per_cm2 = True

    @property
    def name(self) -> str:
        return self._name


class HH(Channel):
    """Hodgkin-Huxley channel (Sterratt et al.); defaults as channels/hh.py:17-35."""

    kernel_id = "HH"

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        if self._name != "HH":
            raise NotImplementedError("renamed HH channels are not supported by the kernels")
        p = self._name
        self.channel_params = {
            f"{p}_gNa": 0.12, f"{p}_gK": 0.036, f"{p}_gLeak": 0.0003,
            f"{p}_eNa": 50.0, f"{p}_eK": -77.0, f"{p}_eLeak": -54.3,
        }
        self.channel_states = {f"{p}_m": 0.2, f"{p}_h": 0.2, f"{p}_n": 0.2}
        self.current_name = "i_HH"


class Leak(Channel):
    """Leak current (Pospischil et al. 2008); defaults as channels/pospischil.py:39-50."""

    kernel_id = "Leak"

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        if self._name != "Leak":
            raise NotImplementedError("renamed Leak channels are not supported by the kernels")
        p = self._name
        self.channel_params = {f"{p}_gLeak": 1e-4, f"{p}_eLeak": -70.0}
        self.channel_states = {}
        self.current_name = f"i_{p}"


__all__ = ["Channel", "HH", "Leak"]
