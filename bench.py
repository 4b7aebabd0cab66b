"""Benchmark: compartment-steps/s of the fused HH simulation on one MI355X.

Workload (one "step" of this benchmark = one pass of the hot path over one batch): the
batched parameter sweep of 4096 simulations on one seeded random tree of 250 branches x
4 compartments (1000 compartments, 1249 nodes), HH everywhere with gNa / gK / gLeak
drawn per simulation in +-50 % of the defaults, 0.2 nA step into the root from 1 to
21 ms, dt = 0.025 ms, 2000 time steps, float32, voltage recorded at the root. Inputs
(states, parameters, stimulus) are resident in HBM before the timed region.

Also reported (not the headline value):
  plain   - one simulation per launch (B = 1), the same kernel, a few simulations;
  fwd_adj - forward + hand-written adjoint (checkpoints every 100 steps) of the same
            batch (gradients of an MSE on the root voltage w.r.t. gNa, gK, gLeak);
  error   - float32 vs float64 on the GPU (all simulations) and vs the CPU oracle on the
            sample of the cpu_baseline leg;
  roofline of the fused forward kernel: 32 algorithmic bytes per compartment-step
            (read + write v, m, h, n in float32) / measured kernel time, vs 8 TB/s.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np
import torch

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "jaxley-pytorch_amd", ROOT / "tests"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

import jaxhip as jx  # noqa: E402
from models import DT, build_tree, c1_params, c1_stimulus  # noqa: E402

NBRANCH, SEED = 250, 11
HBM_PEAK = 8.0e12
BYTES_PER_COMP_STEP = 32


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup(B: int, T: int, dtype, device):
    cell, om = build_tree(NBRANCH, SEED)
    cell.branch(0).comp(0).record()
    prm = c1_params(B, SEED)
    ps = None
    for k, v in prm.items():
        ps = cell.data_set(k, torch.as_tensor(v, dtype=dtype, device=device), ps)
    cur = torch.tensor(c1_stimulus(T), dtype=dtype, device=device)
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    return cell, om, ps, ds, prm


def timed(fn, steps, warmup, device):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize(device)
    return (time.perf_counter() - t0) / steps


def kernel_time_ms(fn, device, reps=2):
    """Average duration of the fused forward launch measured with HIP events on the
    stream the kernel is launched on (PyTorch's current stream)."""
    import jaxhip.fused as F

    times = []
    orig = F.ForwardLaunch.launch

    def launch(self):
        s = torch.cuda.current_stream(device)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(s)
        orig(self)
        b.record(s)
        times.append((a, b))

    F.ForwardLaunch.launch = launch
    try:
        for _ in range(reps):
            fn()
        torch.cuda.synchronize(device)
    finally:
        F.ForwardLaunch.launch = orig
    return float(np.mean([a.elapsed_time(b) for a, b in times]))


def cpu_baseline(prm, T_sample, B_sample, threads):
    """The CPU oracle (a from-scratch restatement of the reference's level-parallel CPU
    step, float64, PyTorch on the host) on a bounded sample of the same workload."""
    from oracle import jx_oracle as O

    torch.set_num_threads(threads)
    _, om = build_tree(NBRANCH, SEED)
    nc = len(om.internal)
    ov = {k: np.broadcast_to(v[:B_sample], (B_sample, nc)).copy() for k, v in prm.items()}
    cur = torch.tensor(c1_stimulus(T_sample))[None]
    t0 = time.perf_counter()
    ref = O.simulate(om, None, DT, stim=cur, stim_comps=[0], rec_comps=[0], overrides=ov)
    el = time.perf_counter() - t0
    return ref, el, nc * T_sample * B_sample / el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--tsteps", type=int, default=2000)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--cpu-sims", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (simulations are "
              "independent; a multi-GPU run would shard the batch with no exchange)", file=sys.stderr)
        sys.exit(2)
    device = torch.device("cuda:0")
    B, T = args.batch, args.tsteps
    dtype = torch.float32
    cell, om, ps, ds, prm = setup(B, T, dtype, device)
    n_comp = cell._n_comp

    def run():
        with torch.no_grad():
            return jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT)

    log(f"[bench] fused forward B={B} T={T} comps={n_comp}")
    sec = timed(run, args.steps, args.warmup, device)
    comp_steps = n_comp * T * B
    value = comp_steps / sec
    ktime = kernel_time_ms(run, device)
    achieved = BYTES_PER_COMP_STEP * comp_steps / (ktime * 1e-3)
    log(f"[bench] {sec*1e3:.1f} ms per batch, kernel {ktime:.1f} ms, {value:.3e} comp-steps/s")
    out = {
        "metric": "compartment-steps/s at <=1e-2 mV vs FP64 (served vs plain); p50/p99 latency",
        "value": value,
        "unit": "compartment-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random tree, seeded per-simulation HH parameters)",
        "config": {"workload": f"batched HH parameter sweep: B={B}, random tree {NBRANCH} branches x 4 "
                               f"compartments ({n_comp} compartments), T={T} steps, dt={DT} ms",
                   "batch": B, "compartments": n_comp, "time_steps": T},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "fused_forward_kernel", "kernel_ms": ktime,
                     "bytes_per_comp_step": BYTES_PER_COMP_STEP},
    }
    rec32 = run()
    from jaxhip.integrate import _F32_MODE, set_float32_mode

    out["f32_mode"] = _F32_MODE
    if not args.no_extras:
        # Float64 on the GPU, all simulations.
        cell64, _, ps64, ds64, _ = setup(B, T, torch.float64, device)
        with torch.no_grad():
            rec64 = jx.integrate(cell64, param_state=ps64, data_stimuli=ds64, delta_t=DT)
        err = (rec32.double() - rec64).abs().amax(dim=(1, 2))
        out["max_err_f32_vs_f64_gpu_mV"] = float(err.max())
        out["frac_sims_within_1e-2_mV"] = float((err <= 1e-2).double().mean())
        # The other float32 mode, for comparison.
        other = "fast" if _F32_MODE == "mixed" else "mixed"
        set_float32_mode(other)
        so = timed(run, 1, 1, device)
        ro = run()
        set_float32_mode(_F32_MODE)
        erro = (ro.double() - rec64).abs().amax(dim=(1, 2))
        out[f"{other}_f32"] = {"value": comp_steps / so, "max_err_vs_f64_gpu_mV": float(erro.max()),
                               "frac_sims_within_1e-2_mV": float((erro <= 1e-2).double().mean())}
        log(f"[bench] {other} f32: {comp_steps / so:.3e} comp-steps/s, max err {float(erro.max()):.3e} mV")
        # Plain: one simulation per launch.
        n_plain = 4
        cp, _, psp, dsp, _ = setup(1, T, dtype, device)

        def run_plain():
            with torch.no_grad():
                for i in range(n_plain):
                    pss = None
                    for k, v in prm.items():
                        pss = cp.data_set(k, torch.as_tensor(v[i:i + 1], dtype=dtype, device=device), pss)
                    jx.integrate(cp, param_state=pss, data_stimuli=dsp, delta_t=DT)

        sp = timed(run_plain, 1, 1, device)
        plain = n_comp * T * n_plain / sp
        out["plain_comp_steps_per_s"] = plain
        out["batched_over_plain"] = value / plain
        log(f"[bench] plain {plain:.3e} comp-steps/s ({value / plain:.0f}x)")
        # Forward + adjoint.
        gp = {k: torch.as_tensor(v, dtype=dtype, device=device).requires_grad_(True) for k, v in prm.items()}
        target = torch.full((T + 1,), -65.0, dtype=dtype, device=device)

        def run_grad():
            pss = None
            for k, v in gp.items():
                pss = cell.data_set(k, v, pss)
            v = jx.integrate(cell, param_state=pss, data_stimuli=ds, delta_t=DT,
                             checkpoint_lengths=[T // 100, 100])
            loss = ((v[:, 0] - target) ** 2).mean()
            loss.backward()

        sg = timed(run_grad, 1, 1, device)
        out["fwd_adj_comp_steps_per_s"] = n_comp * T * B / sg
        out["fwd_adj_ms"] = sg * 1e3
        log(f"[bench] forward+adjoint {sg*1e3:.1f} ms")
    # CPU baseline on a bounded sample (rank 0, N = 1).
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(16, ncpu))
    Bs, Ts = args.cpu_sims, T
    log(f"[bench] cpu baseline: oracle on {Bs} simulations x {Ts} steps, {threads} threads")
    ref, el, cps = cpu_baseline(prm, Ts, Bs, threads)
    out["cpu_baseline"] = {"value": cps, "unit": "compartment-steps/s", "cores": threads, "kind": "port",
                           "sample": f"{Bs} simulations x {Ts} steps x {n_comp} compartments of the same "
                                     f"workload, float64, {el:.1f} s; CPU {platform.processor() or platform.machine()}"}
    err = float((rec32[:Bs].double().cpu() - ref).abs().max())
    out["max_err_f32_vs_oracle_mV"] = err
    log(f"[bench] cpu {cps:.3e} comp-steps/s; f32 vs oracle {err:.3e} mV")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
