"""`jxhip_fused_forward_noalloc_f64` (the allocation-free forward: the streaming kernel's
scratch in a library-held device array) through the C ABI against
`jxhip_fused_forward_f64` with a caller scratch: bitwise equal recordings, voltages and
gates on two calls in a row, the second with a smaller batch, for two streaming launch
shapes; sizes beyond the array are refused before anything runs."""
import sys

import pytest
import torch

import jaxhip  # noqa: F401
from jaxhip import _lib
from jaxhip.fused import ForwardLaunch, LaunchConfig
from models import DT, build_tree, c1_params, c1_stimulus, data_set_params

pytestmark = pytest.mark.gpu

MAX_ELEMS = (64 << 20) // 8          # the library's scratch: 64 MiB of float64
HIP_ERROR_INVALID_VALUE = 1


def _problem(device, B, T, nb=60):
    integ = sys.modules["jaxhip.integrate"]
    cell, _ = build_tree(nb, 5)
    cell.branch(0).comp(0).record()
    ps = data_set_params(cell, c1_params(B, 5), device, torch.float64)
    cur = torch.tensor(c1_stimulus(T), dtype=torch.float64, device=device)
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    inp = integ._assemble(cell, [], ps, ds, None, DT, None, device, torch.float64)
    with torch.no_grad():
        return integ._prepare(cell, inp, DT, device, torch.float64)


def _launch(P, cfg):
    return ForwardLaunch(P.st, P.node_f, P.node_p, P.bparam_for(cfg.S) if P.bmode else None, P.bmode,
                         P.v0, tuple(P.gates0), P.stim, P.stim_scale, P.T, P.dt, P.B, cfg, **P.step_kw)


def test_noalloc_forward_matches_forward(device):
    fn = _lib.fn("jxhip_fused_forward_noalloc_f64")
    shapes = (LaunchConfig(8, 1024, stream=True, gcoef=True), LaunchConfig(4, 512, stream=True))
    for B in (64, 24):                   # the second call with a smaller input
        P = _problem(device, B, 300)
        for cfg in shapes:
            print(f"B={B} shape={cfg}", flush=True)
            ref = _launch(P, cfg)
            ref.launch()
            run = _launch(P, cfg)
            run.args.work_ax = None      # the library's own scratch
            _lib.check(fn(run.args, run.v.numel(), _lib.stream_ptr(device)), "fused_forward_noalloc")
            torch.cuda.synchronize(device)
            assert torch.equal(run.recordings(), ref.recordings())
            for a, b in ((run.v, ref.v), (run.gm, ref.gm), (run.gh, ref.gh), (run.gn, ref.gn)):
                assert torch.equal(a, b)
            assert float(ref.recordings().std()) > 1.0       # the cells spiked


def test_noalloc_forward_refuses_oversize(device):
    fn = _lib.fn("jxhip_fused_forward_noalloc_f64")
    P = _problem(device, 8, 20)
    run = _launch(P, LaunchConfig(8, 1024, stream=True, gcoef=True))
    v0 = run.v.clone()
    run.args.work_ax = None
    assert fn(run.args, MAX_ELEMS + 1, _lib.stream_ptr(device)) == HIP_ERROR_INVALID_VALUE
    run.args.work_ax = run.work_ax.data_ptr()           # a caller scratch is refused too
    assert fn(run.args, run.v.numel(), _lib.stream_ptr(device)) == HIP_ERROR_INVALID_VALUE
    torch.cuda.synchronize(device)
    assert torch.equal(run.v, v0)                         # nothing ran
