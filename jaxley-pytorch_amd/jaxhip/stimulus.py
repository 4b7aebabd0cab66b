"""Step currents (jaxley_pytorch/stimulus.py:8-71): the same float-exact window math."""
import torch


def step_current(i_delay: float, i_dur: float, i_amp, delta_t: float, t_max: float,
                 i_offset: float = 0.0) -> torch.Tensor:
    """Step current in nA, length `int(t_max // delta_t) + 2`; on for steps
    `[int(i_delay / dt), int((i_delay + i_dur) / dt))`."""
    i_amp = torch.as_tensor(i_amp, dtype=torch.get_default_dtype())
    start = int(i_delay / delta_t)
    end = int((i_delay + i_dur) / delta_t)
    steps = int(t_max // delta_t) + 2
    current = torch.zeros((steps,), dtype=i_amp.dtype, device=i_amp.device) + i_offset
    mask = torch.zeros((steps,), dtype=torch.bool, device=i_amp.device)
    mask[start:end] = True
    return torch.where(mask, i_amp, current)


def datapoint_to_step_currents(i_delay: float, i_dur: float, i_amp, delta_t: float,
                               t_max: float, i_offset: float = 0.0) -> torch.Tensor:
    """One step current per amplitude: `(..., dim, time_steps)`."""
    if isinstance(i_amp, torch.Tensor) and i_amp.is_floating_point():
        dtype = i_amp.dtype
    else:
        dtype = torch.get_default_dtype()
    i_amp = torch.as_tensor(i_amp, dtype=dtype)
    start = int(i_delay / delta_t)
    end = int((i_delay + i_dur) / delta_t)
    steps = int(t_max // delta_t) + 2
    current = torch.zeros((steps,), dtype=dtype, device=i_amp.device) + i_offset
    mask = torch.zeros((steps,), dtype=torch.bool, device=i_amp.device)
    mask[start:end] = True
    return torch.where(mask, i_amp[..., None], current)
