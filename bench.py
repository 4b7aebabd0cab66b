"""Benchmark: compartment-steps/s of the fused HH simulation on one MI355X.

Headline (one "step" of this benchmark = one pass of the hot path over one batch): the
batched parameter sweep of 4096 simulations on one seeded random tree of 250 branches x
4 compartments (1000 compartments, 1127 solve nodes), HH everywhere with gNa / gK /
gLeak drawn per simulation in +-50 % of the defaults, 0.2 nA step into the root from 1
to 21 ms, dt = 0.025 ms, 2000 time steps, voltage recorded at the root. The headline
runs in the arithmetic that meets <= 1e-2 mV against float64 on every simulation
(float64: the float32 modes are reported beside it with their measured errors). Inputs
(states, parameters, stimulus) are resident in HBM before the timed region; the launch
shape is chosen by the autotuner (measured) before timing.

Also reported (not the headline value):
  f32 modes - "mixed" (float64 gates / voltages, float32 solve) and "fast" (float32
              gates and solve): throughput, max error vs float64 on the GPU (all
              simulations) and the fraction of simulations within 1e-2 mV;
  plain     - one simulation per launch (B = 1), the same kernel, a few simulations;
  fwd_adj   - forward + hand-written adjoint of the gradient workload: the same tree,
              B = 256, per-branch trainable gNa / gK / gLeak (750 per simulation), MSE
              to target traces over 2000 steps, checkpoints every 100 steps, float32;
  fwd_adj_sweep - forward + adjoint of the headline sweep itself (B = 4096, per-simulation
              trainable conductances, float64);
  network   - 1024 ball-and-stick cells with 32768 KineticSynapses, B = 64, 4000 steps;
              network_fwd_adj: forward + adjoint with every gS trainable;
  swc       - the CA1 morphology of the reference's SWC tests (4 compartments per
              branch), B = 2048, T = 2000, float64;
  served    - 512 requests of random trees (20-2000 compartments, t_max 10-100 ms)
              arriving as a Poisson stream, continuous batching: p50 / p99 latency and
              throughput; served_burst: the same requests all at time 0 (capacity);
  roofline  - of the fused forward kernel: 64 algorithmic bytes per compartment-step in
              float64 (read + write v, m, h, n) / measured kernel time, vs 8 TB/s;
  cpu_baseline - the CPU oracle (float64 PyTorch on the host) on a bounded sample.
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
for _p in (ROOT, ROOT / "jaxley-pytorch_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

import jaxhip as jx  # noqa: E402
from jaxhip.workloads import (DT, c1_params, c1_stimulus, c2_request, c2_specs, c4_network,  # noqa: E402
                              c4_stimulus, random_tree, tree_cell)

NBRANCH, SEED = 250, 11
HBM_PEAK = 8.0e12
BYTES_PER_COMP_STEP = {"f64": 64, "f32": 32}
# HBM bytes per launch of the headline kernel, measured with rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE (scaled by a 256 MiB device copy: fetch x 2.0, write x 1.0) for the headline
# launch shape and workload (profiles/fused_stream_f64_c1_b4096_t2000.md); reported as
# roofline.traffic, with that source, when the autotuner picks that shape. Counters are
# not collected inside this run (a counter pass is a separate rocprofv3 run).
TRAFFIC_SOURCE = "profiles/fused_stream_f64_c1_b4096_t2000.md"
MEASURED_TRAFFIC = {
    ("float64", 4096, 2000, "S=8, threads=1024, stream=True, gcoef=False, sym=True"): 374.2e9 + 371.7e9,
    ("float64", 4096, 2000, "S=8, threads=1024, stream=True, gcoef=True, sym=False"): 376.4e9 + 372.3e9,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup(B: int, T: int, dtype, device):
    cell = tree_cell(NBRANCH, SEED)[0]
    cell.branch(0).comp(0).record()
    prm = c1_params(B, SEED)
    ps = None
    for k, v in prm.items():
        ps = cell.data_set(k, torch.as_tensor(v, dtype=dtype, device=device), ps)
    cur = torch.tensor(c1_stimulus(T), dtype=dtype, device=device)
    ds = cell.branch(0).comp(0).data_stimulate(cur)
    return cell, ps, ds, prm


def timed(fn, steps, warmup, device):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize(device)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize(device)
    return (time.perf_counter() - t0) / steps


def copy_bandwidth(device, nbytes=1 << 30, reps=5) -> float:
    """Measured device-to-device copy bandwidth (bytes read + written per second) of a
    1 GiB buffer, beside the 8 TB/s peak."""
    x = torch.ones(nbytes // 4, dtype=torch.float32, device=device)
    y = torch.empty_like(x)
    y.copy_(x)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        y.copy_(x)
    b.record()
    torch.cuda.synchronize(device)
    sec = a.elapsed_time(b) * 1e-3 / reps
    del x, y
    return 2 * nbytes / sec


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


def spikes(v: torch.Tensor) -> torch.Tensor:
    """Upward crossings of 0 mV per simulation and recording."""
    return ((v[..., 1:] >= 0.0) & (v[..., :-1] < 0.0)).sum(-1)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def physical_cores() -> int:
    """Physical cores available to this process (hyperthreads not counted)."""
    try:
        import psutil

        n = psutil.cpu_count(logical=False)
    except Exception:
        n = None
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    logical = os.cpu_count() or avail
    if n:
        n = max(1, n * avail // max(logical, 1))
    return max(1, min(n or avail, avail))


def cpu_baseline(threads: int, samples=((64, 40), (4096, 2)), dtypes=(torch.float64, torch.float32)):
    """The CPU oracle (this build's from-scratch restatement of the reference's
    level-parallel CPU step, PyTorch on the host) on bounded samples of the headline
    workload: for each batch B and dtype, the median of 3 timed runs of `steps` steps after
    one warm-up run; compartment-steps/s."""
    from oracle import jx_oracle as O

    torch.set_num_threads(threads)
    parents, blen, rad = random_tree(NBRANCH, SEED)
    om = O.OracleModel([(parents, [4] * NBRANCH)])
    om.length, om.radius = np.repeat(blen / 4, 4), np.repeat(rad, 4)
    om.hh_comps = np.arange(4 * NBRANCH)
    nc = len(om.internal)
    out = {}
    for B, steps in samples:
        prm = c1_params(B, SEED)
        ov = {k: np.broadcast_to(v, (B, nc)).copy() for k, v in prm.items()}
        for dt_ in dtypes:
            cur = torch.tensor(c1_stimulus(steps), dtype=dt_)[None]
            run = lambda: O.simulate(om, None, DT, stim=cur, stim_comps=[0], rec_comps=[0], overrides=ov,
                                     dtype=dt_)
            run()
            times = []
            for _ in range(3):
                t0 = time.perf_counter()
                run()
                times.append(time.perf_counter() - t0)
            sec = float(np.median(times))
            key = f"B{B}_{str(dt_).replace('torch.float', 'f')}"
            out[key] = {"comp_steps_per_s": nc * steps * B / sec, "s_per_run": sec, "steps": steps, "batch": B}
            log(f"[bench] cpu {key}: {nc * steps * B / sec:.3e} comp-steps/s ({sec:.2f} s per run)")
    return out


def cpu_reference_sample(prm, T_sample, B_sample, threads):
    """The CPU oracle's float64 traces of the first simulations (checks the GPU result)."""
    from oracle import jx_oracle as O

    torch.set_num_threads(threads)
    parents, blen, rad = random_tree(NBRANCH, SEED)
    om = O.OracleModel([(parents, [4] * NBRANCH)])
    om.length, om.radius = np.repeat(blen / 4, 4), np.repeat(rad, 4)
    om.hh_comps = np.arange(4 * NBRANCH)
    nc = len(om.internal)
    ov = {k: np.broadcast_to(v[:B_sample], (B_sample, nc)).copy() for k, v in prm.items()}
    cur = torch.tensor(c1_stimulus(T_sample))[None]
    return O.simulate(om, None, DT, stim=cur, stim_comps=[0], rec_comps=[0], overrides=ov)


def bench_fwd_adj_sweep(device, B, T, K=100, dtype=torch.float64, mode=None):
    """Forward + hand-written adjoint of the batched sweep (B simulations, per-simulation
    trainable gNa, gK, gLeak): MSE of the root voltage against target traces simulated
    from the seeded sweep parameters, checkpoints every K steps; float64, or float32 in
    the given mode of `set_float32_mode`."""
    from jaxhip.integrate import set_float32_mode

    if mode is not None:
        set_float32_mode(mode)
    cell, _, ds, prm = setup(B, T, dtype, device)
    rng = np.random.default_rng(12)
    guess = {k: v * rng.uniform(0.9, 1.1, size=v.shape) for k, v in prm.items()}
    with torch.no_grad():
        ps_t = None
        for k, v in prm.items():
            ps_t = cell.data_set(k, torch.as_tensor(v, dtype=dtype, device=device), ps_t)
        target = jx.integrate(cell, param_state=ps_t, data_stimuli=ds, delta_t=DT)[:, 0]
    leaves = {k: torch.tensor(v, dtype=dtype, device=device, requires_grad=True)
              for k, v in guess.items()}

    def run():
        ps = None
        for k, v in leaves.items():
            ps = cell.data_set(k, v, ps)
        v = jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT,
                         checkpoint_lengths=[T // K, K])
        loss = ((v[:, 0] - target) ** 2).mean()
        loss.backward()

    sec = timed(run, 1, 1, device)
    if mode is not None:
        set_float32_mode("mixed")
    n_comp = cell._n_comp
    cs = n_comp * T * B / sec
    bpc = 224 if dtype == torch.float64 or mode == "mixed" else 112
    name = "f64" if dtype == torch.float64 else f"f32 ({mode})"
    res = {"comp_steps_per_s": cs, "ms": sec * 1e3, "batch": B, "checkpoint_every": K, "dtype": name,
           "trainable_per_sim": 3, "bytes_per_comp_step": bpc,
           "frac_of_hbm_roofline": bpc * cs / HBM_PEAK}
    log(f"[bench] forward+adjoint sweep B={B} {name}: {sec*1e3:.1f} ms, {cs:.3e} comp-steps/s")
    return res


def bench_network(device, n_cells=1024, k=32, B=64, T=4000, dtype=torch.float32):
    """The ball-and-stick network (1-compartment HH soma + 8-compartment Leak dendrite per
    cell, KineticSynapses with in-degree k onto one dendritic compartment, Bernoulli(0.01)
    0.05 nA pulses into every soma), B simulations, T steps, somata recorded; float32
    (the float32 mode of integrate). One program holds the whole network (10240 solve
    nodes): in float64 its system exceeds the LDS of a workgroup."""
    net, info = c4_network(n_cells, k, seed=9)
    soma = info["soma_comps"]
    for c in soma:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    cur = torch.tensor(c4_stimulus(B, n_cells, T, seed=10), dtype=dtype, device=device)
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(cur[:, i, :], ds)

    def run():
        with torch.no_grad():
            return jx.integrate(net, data_stimuli=ds, delta_t=DT)

    sec = timed(run, 1, 1, device)
    n_comp = net._n_comp
    v = run()
    res = {"comp_steps_per_s": n_comp * T * B / sec, "ms": sec * 1e3, "batch": B, "cells": n_cells,
           "synapses": n_cells * k, "compartments": n_comp, "time_steps": T,
           "dtype": str(dtype).replace("torch.", ""),
           "mean_spikes_per_soma": float(spikes(v.double()).float().mean()),
           # Algorithmic bytes per compartment-step of the float32 network forward: per
           # cell and step 8 B x (9 voltages + 3 gates + 32 synapse states) over its 9
           # compartments (presynaptic voltages come from L2, not counted).
           "bytes_per_comp_step": 39.1,
           "frac_of_hbm_roofline": 39.1 * n_comp * T * B / sec / HBM_PEAK}
    log(f"[bench] network {n_cells} cells x {k} synapses, B={B}, T={T}: {sec*1e3:.1f} ms, "
        f"{res['comp_steps_per_s']:.3e} comp-steps/s")
    return res


def bench_swc(device, ncomp=4, B=2048, T=2000, dtype=torch.float64):
    """A traced morphology: the CA1 pyramidal cell of the reference's SWC tests
    (tests/golden/swc/morph_ca1_n120.swc, read by jx.read_swc with `ncomp` compartments per
    branch), HH everywhere, per-simulation gNa (B values), a 0.2 nA step into the soma,
    T steps, the soma recorded; float64."""
    cell = jx.read_swc(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "swc",
                                    "morph_ca1_n120.swc"), ncomp=ncomp)
    cell.insert(jx.HH())
    cell.soma.branch(0).loc(1.0).record()
    gna = torch.tensor(np.random.default_rng(3).uniform(0.09, 0.15, (B, 1)), dtype=dtype, device=device)
    ps = cell.data_set("HH_gNa", gna, None)
    cur = jx.step_current(1.0, 2.0, 0.2, DT, (T - 1) * DT).to(device=device, dtype=dtype)
    ds = cell.soma.branch(0).loc(1.0).data_stimulate(cur)

    def run():
        with torch.no_grad():
            return jx.integrate(cell, param_state=ps, data_stimuli=ds, delta_t=DT)

    sec = timed(run, 1, 1, device)
    n_comp = cell._n_comp
    T = int(run().shape[-1]) - 1
    res = {"comp_steps_per_s": n_comp * T * B / sec, "ms": sec * 1e3, "batch": B, "compartments": n_comp,
           "branches": len(cell.comb_parents), "time_steps": T, "dtype": str(dtype).replace("torch.", ""),
           "morphology": "morph_ca1_n120.swc"}
    log(f"[bench] swc CA1 {n_comp} compartments, B={B}, T={T}: {sec*1e3:.1f} ms, "
        f"{res['comp_steps_per_s']:.3e} comp-steps/s")
    return res


def bench_network_fwd_adj(device, n_cells=1024, k=32, B=64, T=4000, K=100, dtype=torch.float32):
    """Forward + adjoint of the network workload with every synapse's gS trainable (shared
    by the B simulations): a loss on the somatic voltages, checkpoints every K steps,
    through the fused kernels (checkpointed forward, then the hand-written reverse pass
    with the synapse adjoint and the presynaptic shares in the kernel); the network is
    split into parts that advance in lockstep, so the B = 64 simulations fill the
    device's compute units."""
    net, info = c4_network(n_cells, k, seed=9)
    soma = info["soma_comps"]
    for c in soma:
        net.cell(int(c) // 9).branch(0).comp(0).record()
    cur = torch.tensor(c4_stimulus(B, n_cells, T, seed=10), dtype=dtype, device=device)
    ds = None
    for i, c in enumerate(soma):
        ds = net.cell(int(c) // 9).branch(0).comp(0).data_stimulate(cur[:, i, :], ds)
    gs = torch.tensor(net._edges[0]["params"]["KineticSynapse_gS"], dtype=dtype, device=device,
                      requires_grad=True)

    def run():
        gs.grad = None
        ps = net.data_set("KineticSynapse_gS", gs, None)
        v = jx.integrate(net, param_state=ps, data_stimuli=ds, delta_t=DT, checkpoint_lengths=[T // K, K])
        (v[..., -1] + 65.0).square().mean().backward()

    sec = timed(run, 1, 1, device)
    n_comp = net._n_comp
    res = {"comp_steps_per_s": n_comp * T * B / sec, "ms": sec * 1e3, "batch": B, "cells": n_cells,
           "time_steps": T, "trainable": n_cells * k, "checkpoint_every": K,
           "dtype": str(dtype).replace("torch.", ""),
           "path": "fused checkpointed forward + fused reverse pass (synapse adjoint in the kernel)"}
    log(f"[bench] network forward+adjoint B={B}, T={T}: {sec*1e3:.1f} ms, {res['comp_steps_per_s']:.3e} comp-steps/s")
    return res


def _served_requests(n_req, dtype, seed=7):
    """The clients' request objects (morphology, parameters, stimulus): built before any
    clock starts; the server's work begins when a request is submitted."""
    specs = c2_specs(n_req, seed=seed)
    cells, reqs = [], []
    cpu = torch.device("cpu")          # request tensors live on the host until admission
    for s in specs:
        cell = tree_cell(s["nbranch"], s["seed"])[0]
        cell.branch(0).comp(0).record()
        cells.append(cell)
        reqs.append(c2_request(s, cpu, dtype, cell))
    return cells, reqs


def _run_served(srv, cells, reqs, rate):
    """Poisson arrivals of the requests at `rate` per second (None: all at time 0) into
    an open server; latency counts from the arrival time."""
    n_req = len(reqs)
    rng = np.random.default_rng(8)
    arrivals = (np.zeros(n_req) if rate is None
                else np.cumsum(rng.exponential(1.0 / rate, size=n_req)))
    srv.handles.clear()
    launches0 = srv.launches
    handles = []
    t0 = time.perf_counter()
    i = 0
    while i < n_req or srv.outstanding():
        now = time.perf_counter() - t0
        while i < n_req and arrivals[i] <= now:
            h = srv.submit(reqs[i])
            h.t_submit = t0 + arrivals[i]
            handles.append(h)
            i += 1
        if i == n_req:
            srv.run_until_idle()              # every request has arrived
            break
        srv.poll(block=False)
        if len(srv.inflight) >= srv.max_inflight:
            srv.poll(block=True)              # waits on the GPU (releases the interpreter)
        else:
            # Yield to the preparation workers until the next arrival or event.
            nxt = arrivals[i] - (time.perf_counter() - t0)
            time.sleep(max(5e-5, min(5e-4, nxt)))
    makespan = time.perf_counter() - t0
    lat = np.array([h.latency for h in handles])
    comp_steps = sum(len(c._comp_node) * h.T for c, h in zip(cells, handles))
    return {"requests": n_req, "arrival_rate_per_s": rate if rate is not None else "all at t=0",
            "chunk_steps": srv.chunk_steps,
            "p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
            "makespan_s": makespan, "comp_steps_per_s": comp_steps / makespan,
            "launches": srv.launches - launches0, "prep_order": srv.prep_order}


def bench_served(device, n_req, legs, chunk=200, prep_processes=14):
    """The served legs `(name, rate, dtype, prep_order)`: one long-running server per
    dtype (its preparation workers warmed by an untimed burst of other requests), the
    legs one after the other on it."""
    from jaxhip.serve import SimServer

    import gc

    out = {}
    for dtype in dict.fromkeys(d for _, _, d, _ in legs):
        # Memory of the earlier legs returned (host objects, cached device blocks) before
        # the server starts.
        gc.collect()
        torch.cuda.empty_cache()
        cells, reqs = _served_requests(n_req, dtype)
        wcells, wreqs = _served_requests(128, dtype, seed=99)
        srv = SimServer(device, dtype=dtype, delta_t=DT, hbm_budget=4 << 30, chunk_steps=chunk,
                        max_steps=4001, max_stim_entries=1024, prep_processes=prep_processes)
        try:
            _run_served(srv, wcells, wreqs, None)
            for name, rate, d, order in legs:
                if d != dtype:
                    continue
                srv.prep_order = order
                res = _run_served(srv, cells, reqs, rate)
                res.update(dtype=str(dtype).replace("torch.", ""), prep_processes=prep_processes)
                out[name] = res
        finally:
            srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--tsteps", type=int, default=2000)
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--only", default=None,
                    help="comma-separated legs to run beside the headline: f32, plain, fwd_adj, "
                         "fwd_adj_sweep (all three dtypes), fwd_adj_sweep_f32, fwd_adj_sweep_f32_mixed, "
                         "network, swc, network_fwd_adj, served, cpu (default: all)")
    ap.add_argument("--cpu-sims", type=int, default=4)
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-rate-high", type=float, default=2000.0,
                    help="arrival rate of the saturated served leg (requests per second)")
    ap.add_argument("--prep-processes", type=int, default=14,
                    help="request preparation workers of the served runs (0: one thread)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (simulations are "
              "independent; a multi-GPU run would shard the batch with no exchange)", file=sys.stderr)
        sys.exit(2)
    device = torch.device("cuda:0")
    B, T = args.batch, args.tsteps
    only = None if args.only is None else set(args.only.split(","))

    def want(leg):
        return not args.no_extras and (only is None or leg in only)
    from jaxhip.integrate import set_float32_mode

    _integ = sys.modules["jaxhip.integrate"]

    # ---- headline: float64 (meets <= 1e-2 mV vs float64 on every simulation) ----
    cell64, ps64, ds64, prm = setup(B, T, torch.float64, device)
    n_comp = cell64._n_comp

    def run64():
        with torch.no_grad():
            return jx.integrate(cell64, param_state=ps64, data_stimuli=ds64, delta_t=DT)

    log(f"[bench] fused forward B={B} T={T} comps={n_comp}, float64")
    sec = timed(run64, args.steps, args.warmup, device)
    comp_steps = n_comp * T * B
    value = comp_steps / sec
    ktime = kernel_time_ms(run64, device)
    from jaxhip.autotune import _CACHE

    cfg = [v for k, v in _CACHE.items() if k[0] == "fused" and k[6] == B and "float64" in k[7]]
    bpc = BYTES_PER_COMP_STEP["f64"]
    traffic = MEASURED_TRAFFIC.get(("float64", B, T, str(cfg[0])[13:-1] if cfg else ""))
    achieved = bpc * comp_steps / (ktime * 1e-3)
    log(f"[bench] {sec*1e3:.1f} ms per batch, kernel {ktime:.1f} ms, {value:.3e} comp-steps/s, cfg {cfg}")
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
        "dtype": "f64",
        "data": "synthetic (seeded random tree, seeded per-simulation HH parameters)",
        "config": {"workload": f"batched HH parameter sweep: B={B}, random tree {NBRANCH} branches x 4 "
                               f"compartments ({n_comp} compartments), T={T} steps, dt={DT} ms",
                   "batch": B, "compartments": n_comp, "time_steps": T,
                   "launch": str(cfg[0]) if cfg else None},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "copy_GBps": copy_bandwidth(device) / 1e9,
                     "frac": achieved / HBM_PEAK,
                     "traffic": traffic,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "traffic_source": TRAFFIC_SOURCE if traffic else None,
                     "traffic_GBps": traffic / (ktime * 1e-3) / 1e9 if traffic else None,
                     "kernel": "fused_stream_kernel / fused_regs_kernel (float64)", "kernel_ms": ktime,
                     "bytes_per_comp_step": bpc},
    }
    rec64 = run64()
    # The first simulations of the sweep against vectors the reference itself computed
    # for the same inputs (its first four parameter sets, float64 on the CPU).
    vec = Path(__file__).resolve().parent / "tests" / "golden" / "reference_vectors.npz"
    if vec.exists() and B >= 4 and T == 2000:
        rv = np.load(vec)["c1_v"]
        out["max_err_f64_vs_reference_vectors_mV"] = float(
            np.abs(rec64[:4].cpu().numpy()[:, :, ::4] - rv).max())
        log(f"[bench] float64 GPU vs the reference's own vectors (4 simulations): "
            f"{out['max_err_f64_vs_reference_vectors_mV']:.3e} mV")
    if want("f32"):
        # ---- float32 modes vs float64 on the GPU (all simulations) ----
        cell32, ps32, ds32, _ = setup(B, T, torch.float32, device)
        # float64 on the float32 inputs (geometry, conductances, stimulus rounded): the
        # float32 modes' reference, and the simulations' own sensitivity to that rounding.
        parents, blen, rad = random_tree(NBRANCH, SEED)
        cr = jx.Cell(jx.Branch(jx.Compartment(), 4), parents)
        cr.set("length", np.repeat(blen / 4, 4).astype(np.float32).astype(np.float64))
        cr.set("radius", np.repeat(rad, 4).astype(np.float32).astype(np.float64))
        cr.insert(jx.HH())
        cr.branch(0).comp(0).record()
        psr = None
        for k, v in prm.items():
            psr = cr.data_set(k, torch.as_tensor(v.astype(np.float32).astype(np.float64), device=device), psr)
        curr = torch.tensor(c1_stimulus(T) * float(np.float32(0.2)) / 0.2, dtype=torch.float64, device=device)
        with torch.no_grad():
            rec64r = jx.integrate(cr, param_state=psr, data_stimuli=cr.branch(0).comp(0).data_stimulate(curr),
                                  delta_t=DT)
        sens = (rec64r - rec64).abs().amax(dim=(1, 2))
        good = sens <= 3e-3
        out["f32_input_rounding"] = {"max_shift_of_f64_mV": float(sens.max()),
                                     "well_conditioned_sims": int(good.sum()),
                                     "criterion": "float64 trace moves <= 3e-3 mV when the inputs are rounded to float32"}

        def run32():
            with torch.no_grad():
                return jx.integrate(cell32, param_state=ps32, data_stimuli=ds32, delta_t=DT)

        saved_mode = _integ._F32_MODE
        for mode in ("mixed", "fast"):
            set_float32_mode(mode)
            s32 = timed(run32, 2, 1, device)
            k32 = kernel_time_ms(run32, device)
            r32 = run32()
            err = (r32.double() - rec64).abs().amax(dim=(1, 2))
            err_r = (r32.double() - rec64r).abs().amax(dim=(1, 2))
            spikes_equal = bool((spikes(r32.double()) == spikes(rec64)).all())
            out[f"f32_{mode}"] = {"value": comp_steps / s32, "kernel_ms": k32,
                                  "max_err_vs_f64_gpu_mV": float(err.max()),
                                  "frac_sims_within_1e-2_mV": float((err <= 1e-2).double().mean()),
                                  "max_err_vs_f64_same_inputs_mV": float(err_r.max()),
                                  "max_err_well_conditioned_mV": float(err_r[good].max()),
                                  "spike_counts_identical": spikes_equal}
            log(f"[bench] f32 {mode}: {comp_steps / s32:.3e} comp-steps/s, max err {float(err.max()):.3e} mV, "
                f"{float((err <= 1e-2).double().mean()):.4f} within 1e-2")
        set_float32_mode(saved_mode)
    if want("plain"):
        # ---- plain: one simulation per launch ----
        n_plain = 4
        cp, _, dsp, _ = setup(1, T, torch.float64, device)

        def run_plain():
            with torch.no_grad():
                for i in range(n_plain):
                    pss = None
                    for k, v in prm.items():
                        pss = cp.data_set(k, torch.as_tensor(v[i:i + 1], dtype=torch.float64, device=device), pss)
                    jx.integrate(cp, param_state=pss, data_stimuli=dsp, delta_t=DT)

        sp = timed(run_plain, 1, 1, device)
        plain = n_comp * T * n_plain / sp
        out["plain_comp_steps_per_s"] = plain
        out["batched_over_plain"] = value / plain
        log(f"[bench] plain {plain:.3e} comp-steps/s ({value / plain:.0f}x)")
    if want("fwd_adj"):
        # ---- forward + adjoint: per-branch trainable conductances, B = 256 ----
        Bg = 256
        cg, _, dsg, _ = setup(1, T, torch.float32, device)
        for k in ("HH_gNa", "HH_gK", "HH_gLeak"):
            cg.branch("all").make_trainable(k, verbose=False)
        rng = np.random.default_rng(5)
        base = np.array([0.12, 0.036, 0.0003])
        true = base[None, :, None] * rng.uniform(0.5, 1.5, size=(Bg, 3, NBRANCH))
        guess = base[None, :, None] * rng.uniform(0.5, 1.5, size=(Bg, 3, NBRANCH))

        def params_of(vals, grad):
            return [{k: torch.tensor(vals[:, j], dtype=torch.float32, device=device, requires_grad=grad)}
                    for j, k in enumerate(("HH_gNa", "HH_gK", "HH_gLeak"))]

        with torch.no_grad():
            target = jx.integrate(cg, params_of(true, False), data_stimuli=dsg, delta_t=DT)[:, 0]
        gp = params_of(guess, True)

        def run_grad():
            v = jx.integrate(cg, gp, data_stimuli=dsg, delta_t=DT, checkpoint_lengths=[T // 100, 100])
            loss = ((v[:, 0] - target) ** 2).mean()
            loss.backward()

        sg = timed(run_grad, 1, 1, device)
        out["fwd_adj"] = {"comp_steps_per_s": n_comp * T * Bg / sg, "ms": sg * 1e3, "batch": Bg,
                          "trainable_per_sim": 3 * NBRANCH, "checkpoint_every": 100, "dtype": "f32 (mixed)",
                          "bytes_per_comp_step": 112,
                          "frac_of_hbm_roofline": 112 * n_comp * T * Bg / sg / HBM_PEAK}
        log(f"[bench] forward+adjoint B={Bg}: {sg*1e3:.1f} ms, {n_comp * T * Bg / sg:.3e} comp-steps/s")
    # ---- forward + adjoint of the batched sweep itself: B = 4096, float64 and float32 ----
    if want("fwd_adj_sweep"):
        out["fwd_adj_sweep"] = bench_fwd_adj_sweep(device, B, T)
    if want("fwd_adj_sweep") or want("fwd_adj_sweep_f32"):
        out["fwd_adj_sweep_f32"] = bench_fwd_adj_sweep(device, B, T, dtype=torch.float32, mode="fast")
    if want("fwd_adj_sweep") or want("fwd_adj_sweep_f32_mixed"):
        out["fwd_adj_sweep_f32_mixed"] = bench_fwd_adj_sweep(device, B, T, dtype=torch.float32, mode="mixed")
    if want("network"):
        # ---- network: 1024 ball-and-stick cells, 32768 KineticSynapses, B = 64 ----
        out["network"] = bench_network(device)
    if want("swc"):
        # ---- a traced morphology (SWC): the CA1 cell, B = 2048, float64 ----
        out["swc"] = bench_swc(device)
    if want("network_fwd_adj"):
        out["network_fwd_adj"] = bench_network_fwd_adj(device)
    if want("served"):
        # ---- served: continuous batching of uneven requests ----
        # Poisson arrivals at the given rate and at a rate above the server's capacity (the
        # server, not the arrivals, bounds latency there), then the same requests all at
        # time 0 (capacity; preparation of the longest simulations first), in float64 and
        # float32.
        if args.serve_requests > 0:
            ncpu_s = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
            procs = max(0, min(args.prep_processes, ncpu_s - 2))
            legs = (("served", args.serve_rate, torch.float64, "fifo"),
                    ("served_saturated", args.serve_rate_high, torch.float64, "fifo"),
                    ("served_burst", None, torch.float64, "longest_first"),
                    ("served_burst_f32", None, torch.float32, "longest_first"))
            res = bench_served(device, args.serve_requests, legs, 200, procs)
            for key, _, _, _ in legs:
                sv = res[key]
                if "plain_comp_steps_per_s" in out:
                    sv["over_plain"] = sv["comp_steps_per_s"] / out["plain_comp_steps_per_s"]
                out[key] = sv
                log(f"[bench] {key} {sv['requests']} requests at {sv['arrival_rate_per_s']}/s ({sv['dtype']}): "
                    f"p50 {sv['p50_latency_ms']:.1f} ms, p99 {sv['p99_latency_ms']:.1f} ms, "
                    f"{sv['comp_steps_per_s']:.3e} comp-steps/s")
    if only is not None and "cpu" not in only:
        print(json.dumps(out), flush=True)
        return
    # ---- CPU baseline on bounded samples (rank 0, N = 1) ----
    threads = physical_cores()
    log(f"[bench] cpu baseline: the oracle on {threads} threads (physical cores)")
    cpu = cpu_baseline(threads)
    head = cpu["B4096_f64"]
    out["cpu_baseline"] = {"value": head["comp_steps_per_s"], "unit": "compartment-steps/s", "cores": threads,
                           "kind": "port",
                           "sample": f"the headline workload (random tree, {n_comp} compartments, per-simulation "
                                     f"HH conductances) at B=4096 for {head['steps']} steps, float64, median of 3 "
                                     f"runs after 1 warm-up; CPU {cpu_model()}",
                           "runs": cpu}
    Bs = args.cpu_sims
    ref = cpu_reference_sample(prm, T, Bs, threads)
    err = float((rec64[:Bs].cpu() - ref).abs().max())
    out["max_err_f64_vs_oracle_mV"] = err
    out["spike_counts_f64_vs_oracle_identical"] = bool((spikes(rec64[:Bs].cpu()) == spikes(ref)).all())
    log(f"[bench] float64 GPU vs oracle on {Bs} simulations: {err:.3e} mV")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
