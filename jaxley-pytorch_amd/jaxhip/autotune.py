"""Project autotuner: picks kernel variants and launch shapes by measuring them.

* `choose_tree_kernel`: register vs level variant of the standalone tree solve (the
  reference estimates the work instead, kernels/tree_solve.py:496-509); both are timed on
  a synthetic diagonally dominant system of the real forest.
* `fused_config`: simulations per program (S), threads per workgroup and kernel
  (register-resident or streaming) of the fused forward; candidates are timed on a short
  run of the real structure and batch (a few steps).

Results are cached per (structure, batch size, dtype, device) for the process.
`JAXHIP_AUTOTUNE=0` disables measurement (heuristic choice).
"""
from __future__ import annotations

import hashlib
import os
from typing import Dict, Tuple

import numpy as np
import torch

_CACHE: Dict[Tuple, object] = {}


def _enabled() -> bool:
    return os.environ.get("JAXHIP_AUTOTUNE", "1") != "0"


def _time(fn, reps: int = 3) -> float:
    fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / reps


def choose_tree_kernel(parent, device, forest, levels):
    from .kernels.tree_solve import LevelTreeSolveKernel, TreeSolveKernel

    key = ("tree", hashlib.sha1(np.asarray(parent, dtype=np.int64).tobytes()).hexdigest(), str(device))
    if key in _CACHE:
        variant = _CACHE[key]
    elif not _enabled():
        # Work estimate as a fallback: the register kernel touches the whole block per level.
        reg = int(np.sum(forest["block_info"][:, 2])) * 64 * forest["npl"]
        lev = int(np.sum(levels["block_info"][:, 1])) * 4
        variant = "registers" if reg <= lev else "levels"
        _CACHE[key] = variant
    else:
        kr = TreeSolveKernel(forest, device)
        kl = LevelTreeSolveKernel(levels, device)
        n = kr.num_nodes
        rows = 256
        g = torch.Generator(device="cpu").manual_seed(0)
        l = -torch.rand(rows, n, generator=g).to(device)
        u = -torch.rand(rows, n, generator=g).to(device)
        d = 4.0 + torch.rand(rows, n, generator=g).to(device)
        s = torch.randn(rows, n, generator=g).to(device)
        tr = _time(lambda: kr.solve(d, l, u, s))
        tl = _time(lambda: kl.solve(d, l, u, s))
        variant = "registers" if tr <= tl else "levels"
        _CACHE[key] = variant
        return kr if variant == "registers" else kl
    return TreeSolveKernel(forest, device) if variant == "registers" else LevelTreeSolveKernel(levels, device)


def _key(st, B, dtype, lsize, grad):
    n_tgt = st.syn["n_tgt"] if st.syn else 0
    return ("fused", st.n, st.L, n_tgt, st.parts, int(st.topo_row.sum()), B, str(dtype), lsize, bool(grad))


def parse_config(text: str):
    """"S,threads[,stream|gcoef|sym]" (the JAXHIP_FUSED_CONFIG format)."""
    from .fused import LaunchConfig

    parts = [x.strip() for x in text.split(",")]
    S, thr = int(parts[0]), int(parts[1])
    flag = parts[2] if len(parts) > 2 else ""
    return LaunchConfig(S, thr, stream=flag in ("stream", "gcoef", "sym"), gcoef=flag == "gcoef",
                        sym=flag == "sym")


def fused_config(st, B: int, dtype, lsize: int = 0, grad: bool = False, measure=None):
    """Launch shape of the fused forward for a structure and batch size. With `grad` the
    shape is shared with the reverse pass and obeys its limits. `measure(cfg)` returns
    the time of a short run with that shape; when given (and JAXHIP_AUTOTUNE != 0) the
    fastest candidate is chosen, else a heuristic one."""
    from .fused import config_candidates, default_config, grad_config

    # The largest part of a split network sizes the LDS and the threads of a program.
    n, n_tgt, sched = st.n_max, st.tgt_max, st.sched
    stream_only = st.parts > 1 and st.syn is not None
    esize = torch.tensor([], dtype=dtype).element_size()
    key = _key(st, B, dtype, lsize, grad)
    if key in _CACHE:
        return _CACHE[key]
    lsz = lsize if lsize else esize
    env = os.environ.get("JAXHIP_FUSED_CONFIG")
    if grad:
        # The checkpointed forward runs at the reverse pass's S (the checkpoints are in its
        # program layout); among the forward shapes with that S the fastest is measured.
        from .fused import adjoint_config

        S = adjoint_config(n, B, lsz, sched, n_tgt=n_tgt).S
        cands = [c for c in config_candidates(n, n_tgt, B, esize, lsz, sched, stream_only) if c.S == S]
        if measure is not None and _enabled() and len(cands) > 1:
            times = [(measure(c), i) for i, c in enumerate(cands)]
            cfg = cands[min(times)[1]]
        else:
            cfg = grad_config(n, B, lsz, sched, n_tgt)
    elif env:
        cfg = parse_config(env)
    elif measure is not None and _enabled():
        cands = config_candidates(n, n_tgt, B, esize, lsz, sched, stream_only)
        times = [(measure(c), i) for i, c in enumerate(cands)]
        cfg = cands[min(times)[1]]
    else:
        cfg = default_config(n, n_tgt, B, lsz, esize, sched, stream_only)
    _CACHE[key] = cfg
    return cfg


def set_fused_config(st, B: int, dtype, cfg, lsize: int = 0, grad: bool = False) -> None:
    """Record a measured launch shape (used by the serving loop and the benchmark)."""
    _CACHE[_key(st, B, dtype, lsize, grad)] = cfg
