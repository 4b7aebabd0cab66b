"""Ion channels on the hot path: `HH` (channels/hh.py:14-108) and the Pospischil et
al. (2008) `Leak`, `Na`, `K`, `Km` (channels/pospischil.py:39-250) of jaxley-pytorch.

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


class _Pospischil(Channel):
    """Pospischil channels keep their reference names: the kernels know their
    parameters (`vt`, `eNa`, `eK` are global, shared by the channels that use them)."""

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        if self._name != self.__class__.__name__:
            raise NotImplementedError("renamed Pospischil channels are not supported by the kernels")


class Na(_Pospischil):
    """Sodium channel (pospischil.py:74-134)."""

    kernel_id = "PNa"

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        self.channel_params = {"Na_gNa": 50e-3, "eNa": 50.0, "vt": -60.0}
        self.channel_states = {"Na_m": 0.2, "Na_h": 0.2}
        self.current_name = "i_Na"


class K(_Pospischil):
    """Potassium channel (pospischil.py:137-193)."""

    kernel_id = "PK"

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        self.channel_params = {"K_gK": 5e-3, "eK": -90.0, "vt": -60.0}
        self.channel_states = {"K_n": 0.2}
        self.current_name = "i_K"


class Km(_Pospischil):
    """Slow M potassium channel (pospischil.py:196-245)."""

    kernel_id = "PKm"

    def __init__(self, name: Optional[str] = None):
        super().__init__(name)
        self.channel_params = {"Km_gKm": 0.004e-3, "Km_taumax": 4000.0, "eK": -90.0}
        self.channel_states = {"Km_p": 0.2}
        self.current_name = "i_K"


__all__ = ["Channel", "HH", "Leak", "Na", "K", "Km"]
