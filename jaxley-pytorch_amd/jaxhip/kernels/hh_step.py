"""One HH channel step on the GPU: the eager `_step_channels` + linearization of one HH
channel (modules/base.py:3414-3542 with channels/hh.py:37-67), as a single kernel
(csrc/jxhip_channels.hip).

`hh_step(v, inds, m, h, n, params, dt)`: `v` (B, n_nodes) voltages, `inds` (n_ch,) the
nodes carrying the channel, gates `m, h, n` (B, n_ch), `params` the six HH parameters
(gNa, gK, gLeak in S/cm^2, eNa, eK, eLeak in mV), each (n_ch,) or (B, n_ch). Returns
`(m', h', n', i, g, ge)`: gates after the exponential-Euler update at the old voltage,
the current i = sum g_x (v - e_x) (mA/cm^2) with the new gates, its slope g = di/dv
(the total conductance; the reference's finite difference of a current linear in v) and
ge = sum g_x e_x, so that i = g v - ge. No autograd (the differentiable path is the
fused simulation of `jaxhip.integrate`).
"""
from __future__ import annotations

from typing import Sequence

import ctypes as C
import torch

from .. import _lib


def hh_step(v: torch.Tensor, inds: torch.Tensor, m: torch.Tensor, h: torch.Tensor, n: torch.Tensor,
            params: Sequence[torch.Tensor], dt: float):
    if v.device.type != "cuda":
        raise RuntimeError("jaxhip.kernels.hh_step runs on the GPU")
    B, n_nodes = v.shape
    n_ch = int(inds.numel())
    dev, dtype = v.device, v.dtype
    v = v.contiguous()
    inds = inds.to(device=dev, dtype=torch.int32).contiguous()
    m, h, n = (x.to(dtype).expand(B, n_ch).contiguous() for x in (m, h, n))
    ps, strides = [], (C.c_int64 * 6)()
    ptrs = (C.c_void_p * 6)()
    for q, p in enumerate(params):
        p = p.to(device=dev, dtype=dtype)
        if p.ndim == 2 and p.shape[0] > 1:
            p = p.expand(B, n_ch).contiguous()
            strides[q] = n_ch
        else:
            p = p.reshape(-1).expand(n_ch).contiguous()
            strides[q] = 0
        ps.append(p)
        ptrs[q] = p.data_ptr()
    outs = [torch.empty((B, n_ch), dtype=dtype, device=dev) for _ in range(6)]
    f = _lib.fn("jxhip_hh_step", dtype)
    _lib.check(f(v.data_ptr(), n_nodes, inds.data_ptr(), n_ch, m.data_ptr(), h.data_ptr(),
                 n.data_ptr(), C.cast(ptrs, C.c_void_p), C.cast(strides, C.c_void_p), float(dt),
                 *(o.data_ptr() for o in outs), B, _lib.stream_ptr(dev)), "hh_step")
    return tuple(outs)
