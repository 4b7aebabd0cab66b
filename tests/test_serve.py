"""Serving loop (jaxhip.serve) on the GPU: results alone vs in a mix, vs `integrate` and
vs the oracle; admission under a tight HBM budget."""
import numpy as np
import pytest
import torch

import jaxhip as jx
from jaxhip.serve import SimServer
from models import DT, build_tree, c1_stimulus, c2_request, c2_specs, oracle_overrides
from oracle import jx_oracle as O

pytestmark = pytest.mark.gpu


def _serve(specs, device, dtype=torch.float64, **kw):
    srv = SimServer(device, dtype=dtype, delta_t=DT, hbm_budget=kw.pop("hbm_budget", 256 << 20),
                    max_steps=4001, max_stim_entries=64, **kw)
    hs = [srv.submit(c2_request(s, device, dtype)) for s in specs]
    srv.run_until_idle()
    assert all(h.done for h in hs)
    return srv, [h.result.cpu() for h in hs]


def _small_specs(n, seed=3):
    specs = c2_specs(n, seed)
    for s in specs:                      # keep the oracle and the test fast
        s["nbranch"] = min(s["nbranch"], 30)
        s["t_max"] = min(s["t_max"], 12.0)
    return specs


def test_served_alone_vs_mix_vs_integrate(device):
    specs = _small_specs(6)
    _, mix = _serve(specs, device, chunk_steps=64)
    for i in (0, 3):
        _, alone = _serve([specs[i]], device, chunk_steps=64)
        assert torch.equal(alone[0], mix[i])
    for i, s in enumerate(specs[:3]):
        req = c2_request(s, device, torch.float64)
        v = jx.integrate(req.module, param_state=req.param_state, data_stimuli=req.data_stimuli,
                         t_max=s["t_max"], delta_t=DT)
        assert v.shape[-1] == mix[i].shape[-1]
        assert float((v[0].cpu() - mix[i]).abs().max()) < 1e-8


def test_served_vs_oracle(device):
    specs = _small_specs(3, seed=5)
    _, res = _serve(specs, device, chunk_steps=50)
    for s, r in zip(specs, res):
        _, om = build_tree(s["nbranch"], s["seed"])
        T = int(s["t_max"] // DT + 1)
        ref = O.simulate(om, None, DT, stim=torch.tensor(c1_stimulus(T))[None], stim_comps=[0],
                         rec_comps=[0], overrides=oracle_overrides(om, s["params"]))
        print("served vs oracle", float((r - ref[0]).abs().max()))
        assert float((r - ref[0]).abs().max()) < 1e-6


def test_admission_under_budget(device):
    specs = _small_specs(12, seed=9)
    srv, res = _serve(specs, device, chunk_steps=40, hbm_budget=(4001 * 64 * 8) + (3 << 20))
    assert srv.nodes_a.used == 0 and srv.rec_a.used == 0 and not srv.live
    _, ref = _serve(specs, device, chunk_steps=40)
    for a, b in zip(res, ref):
        assert torch.equal(a, b)


def test_served_float32(device):
    specs = _small_specs(4, seed=11)
    _, r32 = _serve(specs, device, dtype=torch.float32, chunk_steps=64)
    _, r64 = _serve(specs, device, chunk_steps=64)
    for a, b in zip(r32, r64):
        assert float((a.double() - b).abs().max()) < 1e-2



def test_served_with_preparation_processes(device):
    """Requests prepared in worker processes (`prep_processes`) give the same results as
    requests prepared in a thread (request tensors on the host)."""
    specs = _small_specs(4, seed=5)
    cpu = torch.device("cpu")

    def run(**kw):
        srv = SimServer(device, dtype=torch.float64, delta_t=DT, hbm_budget=256 << 20, max_steps=4001,
                        max_stim_entries=64, chunk_steps=64, **kw)
        hs = [srv.submit(c2_request(s, cpu, torch.float64)) for s in specs]
        srv.run_until_idle()
        srv.close()
        return [h.result.cpu() for h in hs]

    for a, b in zip(run(prep_threads=1), run(prep_processes=2)):
        assert torch.equal(a, b)


def test_served_swc_cells(device):
    """Requests of cells read from SWC files (traced geometry: input and output loads
    differ) served together equal `integrate` of each."""
    from jaxhip.serve import SimRequest
    from test_swc import swc

    srv = SimServer(device, dtype=torch.float64, delta_t=DT, hbm_budget=256 << 20, max_steps=4001,
                    max_stim_entries=64, chunk_steps=64)
    reqs = []
    for f, nc, t_max in (("morph_ca1_n120_250.swc", 2, 8.0), ("morph_ca1_n120.swc", 2, 5.0),
                         ("morph_ca1_n120_single_point_soma.swc", 1, 6.0)):
        cell = jx.read_swc(swc(f), ncomp=nc, max_branch_len=300.0)
        cell.insert(jx.HH())
        cell.soma.branch(0).loc(1.0).record()
        T = int(t_max // DT + 1)
        cur = torch.tensor(c1_stimulus(T), dtype=torch.float64, device=device) * 2
        ds = cell.soma.branch(0).loc(1.0).data_stimulate(cur)
        reqs.append((cell, ds, t_max))
    hs = [srv.submit(SimRequest(c, t_max=t, data_stimuli=ds)) for c, ds, t in reqs]
    srv.run_until_idle()
    for h, (c, ds, t) in zip(hs, reqs):
        v = jx.integrate(c, data_stimuli=ds, t_max=t, delta_t=DT)
        err = float((v[0].cpu() - h.result.cpu()).abs().max())
        print("served swc vs integrate", err, flush=True)
        assert v.shape[-1] == h.result.shape[-1] and err < 1e-8
