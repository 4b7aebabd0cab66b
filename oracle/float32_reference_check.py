"""TEST INFRASTRUCTURE: the reference's float32 arithmetic (the oracle run in float32,
every tensor operation in float32 as the reference computes with float32 inputs) against
float64 on the same float32 inputs, for the first 64 simulations of the batched sweep
(random tree of 250 branches, 2000 steps): the error a float32 simulation by the
reference's own arithmetic carries, beside which the float32 modes of the HIP path are
measured (tests/test_float32_gpu.py). Run from the repository root:
    python oracle/float32_reference_check.py      (about 2 minutes on 8 cores)
"""
import sys
import time
from pathlib import Path

_ROOT = Path(__file__).resolve().parents[1]
sys.path[:0] = [str(_ROOT), str(_ROOT / "jaxley-pytorch_amd")]
import numpy as np, torch
torch.set_num_threads(8)
from oracle import jx_oracle as O
from jaxhip.workloads import random_tree, c1_params, c1_stimulus, DT
B, T = 64, 2000
parents, blen, rad = random_tree(250, 11)
L = np.repeat(blen / 4, 4).astype(np.float32).astype(np.float64)
R = np.repeat(rad, 4).astype(np.float32).astype(np.float64)
om = O.OracleModel([(parents, [4] * 250)])
om.length, om.radius, om.hh_comps = L, R, np.arange(1000)
prm = {k: v.astype(np.float32).astype(np.float64) for k, v in c1_params(B, 11).items()}
ov = {k: np.broadcast_to(v, (B, 1000)).copy() for k, v in prm.items()}
stim = torch.tensor(c1_stimulus(T) * float(np.float32(0.2)) / 0.2)[None]
t0 = time.time()
r64 = O.simulate(om, None, DT, stim=stim, stim_comps=[0], rec_comps=[0], overrides=ov)[:, 0]
print("f64", time.time() - t0, flush=True)
r32 = O.simulate(om, None, DT, stim=stim.float(), stim_comps=[0], rec_comps=[0], overrides=ov, dtype=torch.float32)[:, 0]
print("f32", time.time() - t0, flush=True)
e = (r32.double() - r64).abs().amax(-1).numpy()
sp = lambda v: ((v[..., 1:] >= 0) & (v[..., :-1] < 0)).sum(-1)
print("reference float32 arithmetic vs float64: max", e.max(), "median", np.median(e), "n>1e-2", (e > 1e-2).sum(),
      "spike diffs", int((sp(r32.double()) != sp(r64)).sum()))
print(np.round(e, 4).tolist())
