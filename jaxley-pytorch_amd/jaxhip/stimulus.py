"""Step currents (jaxley_pytorch/stimulus.py:8-71).

The window arithmetic is the reference's and must stay float-exact: the current is on
for the steps `[int(i_delay / dt), int((i_delay + i_dur) / dt))` of a trace of
`int(t_max // dt) + 2` samples (so `t_max = 5.0, dt = 0.025` gives 201 samples, since
5.0 // 0.025 == 199.0 in binary floating point). Both public functions are built on one
helper that makes the window as a 0/1 profile and blends amplitude and offset with it,
which keeps the result differentiable with respect to the amplitude.
"""
import torch


def _window(i_delay: float, i_dur: float, delta_t: float, t_max: float, dtype, device) -> torch.Tensor:
    """0/1 profile of the stimulation window, `int(t_max // delta_t) + 2` samples."""
    n = int(t_max // delta_t) + 2
    on = torch.arange(n, device=device)
    first, stop = int(i_delay / delta_t), int((i_delay + i_dur) / delta_t)
    return ((on >= first) & (on < stop)).to(dtype)


def _blend(amp: torch.Tensor, window: torch.Tensor, offset: float) -> torch.Tensor:
    """amp inside the window, offset outside (amp's leading dimensions kept)."""
    amp = amp[..., None]
    return amp * window + (1 - window) * offset


def step_current(i_delay: float, i_dur: float, i_amp, delta_t: float, t_max: float,
                 i_offset: float = 0.0) -> torch.Tensor:
    """One step current in nA (default dtype; differentiable in a tensor `i_amp`)."""
    amp = torch.as_tensor(i_amp, dtype=torch.get_default_dtype())
    return _blend(amp, _window(i_delay, i_dur, delta_t, t_max, amp.dtype, amp.device), i_offset)


def datapoint_to_step_currents(i_delay: float, i_dur: float, i_amp, delta_t: float,
                               t_max: float, i_offset: float = 0.0) -> torch.Tensor:
    """One step current per amplitude: `(..., dim, time_steps)`; the amplitudes' dtype
    when they are a floating point tensor, else the default dtype."""
    floating = isinstance(i_amp, torch.Tensor) and i_amp.is_floating_point()
    amp = torch.as_tensor(i_amp, dtype=i_amp.dtype if floating else torch.get_default_dtype())
    return _blend(amp, _window(i_delay, i_dur, delta_t, t_max, amp.dtype, amp.device), i_offset)
