"""Continuous-batching simulation server on one GPU.

Simulations of different morphologies, durations and parameters arrive as requests
(`SimRequest`: a cell with its recordings, `t_max`, parameters and stimuli exactly as
`jaxhip.integrate` takes them, batch size 1). The server

* prepares each request on arrival, on the host (schedule and node tables,
  `integrate._prepare`; request tensors are best given on the CPU), in a background
  thread (`prep_threads`) or in worker processes (`prep_processes`), so that admissions
  and launches go on meanwhile,
* admits it when a slot and room in its HBM arenas are free (`hbm_budget` bytes for
  node tables, states, schedules, stimuli and recordings; first fit),
* keeps the state of every admitted simulation resident in HBM between launches,
* advances all admitted simulations together, one persistent fused-kernel launch per
  chunk of `chunk_steps` time steps: every simulation is one program (its own topology
  row), so trees of uneven size and simulations at different times run in the same
  launch, and a request that arrives while a chunk runs joins the next chunk,
* retires a simulation when the chunk holding its last step has completed, copies its
  recordings out and frees its arena space for the next requests.

The launch shape of a chunk (threads per program, register-resident or streaming
kernel) follows the largest admitted tree; `autotune=True` measures the candidates on
copies of the live state the first time a size class is seen and keeps the fastest.

Every simulation is computed by its own workgroup with the arithmetic of
`jaxhip.integrate`, so its result does not depend on what else is in the launch (the
tests check alone vs in a mix, and against `integrate`).
"""
from __future__ import annotations

import concurrent.futures as cf
import importlib
import multiprocessing
import time
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional

import numpy as np
import torch

from . import _lib
from .fused import LaunchConfig, items_per_thread, lds_bytes
from .integrate import _assemble, _prepare, _require_gpu

_LDS_MAX = 160 * 1024   # bytes of LDS one workgroup may use on gfx950


class _Arena:
    """First-fit allocator over [0, capacity) in units of one element."""

    def __init__(self, capacity: int):
        self.capacity = int(capacity)
        self.free_list = [(0, self.capacity)]
        self.used = 0

    def alloc(self, size: int, align: int = 1) -> Optional[int]:
        size = max(int(size), 1)
        for idx, (off, ln) in enumerate(self.free_list):
            start = -(-off // align) * align
            if start + size <= off + ln:
                rest = []
                if start > off:
                    rest.append((off, start - off))
                if start + size < off + ln:
                    rest.append((start + size, off + ln - start - size))
                self.free_list[idx:idx + 1] = rest
                self.used += size
                return start
        return None

    def free(self, off: int, size: int) -> None:
        size = max(int(size), 1)
        self.used -= size
        self.free_list.append((off, size))
        self.free_list.sort()
        merged = []
        for o, ln in self.free_list:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free_list = merged


@dataclass
class SimRequest:
    """One simulation (batch size 1), with `integrate`'s arguments."""

    module: Any
    t_max: Optional[float] = None
    param_state: Any = None
    data_stimuli: Any = None
    all_states: Any = None
    params: List = field(default_factory=list)
    tag: Any = None


class SimHandle:
    """State of a submitted request; `result` is `(n_recordings, 1 + steps)` once done."""

    def __init__(self, rid: int, request: SimRequest):
        self.id = rid
        self.request = request
        self.t_submit = time.perf_counter()
        self.t_admit: Optional[float] = None
        self.t_done: Optional[float] = None
        self.result: Optional[torch.Tensor] = None
        self.steps_done = 0
        self.T = 0

    @property
    def done(self) -> bool:
        return self.result is not None

    @property
    def latency(self) -> Optional[float]:
        return None if self.t_done is None else self.t_done - self.t_submit


class _Prepared:
    """A request prepared on the host: schedule, node tables (channel parameters with the
    request's values merged in), initial state and stimulus, as CPU tensors."""


def _prepare_payload(module, params, param_state, data_stimuli, t_max, all_states, dt: float,
                     user_dtype, f32_mode: str, max_steps: int) -> _Prepared:
    """`integrate._assemble` + `_prepare` for one served request (batch size 1)."""
    integ = importlib.import_module("jaxhip.integrate")
    if integ.has_pospischil(module):
        raise NotImplementedError("served requests carry HH / Leak channels only")
    integ._F32_MODE = f32_mode
    cpu = torch.device("cpu")
    inp = _assemble(module, params, param_state, data_stimuli, t_max, dt, all_states, cpu, user_dtype)
    with torch.no_grad():
        P = _prepare(module, inp, dt, cpu, user_dtype, split=False)
    if P.B != 1:
        raise ValueError("a served request is one simulation (batch size 1)")
    if P.st.syn is not None:
        raise NotImplementedError("the server runs cells without synapses")
    if P.T + 1 > max_steps:
        raise ValueError(f"{P.T} steps exceed the server's max_steps={max_steps}")
    if P.stim is not None and P.stim.shape[2] != 1:
        raise ValueError("stimulus with batch dimensions in a served request")
    st = P.st
    q = _Prepared()
    q.n, q.L, q.T, q.n_rec, q.sched = st.n, st.L, P.T, st.n_rec, st.sched
    q.node_i, q.seg, q.seg_level_ptr, q.seg_child = st.node_i, st.seg, st.seg_level_ptr, st.seg_child
    q.child_idx, q.rec_node, q.stim_col = st.child_idx, st.rec_node, st.stim_col
    q.node_f = P.node_f
    node_p = P.node_p.clone()
    if P.bmode:
        bp = P.bparam_for(1)                       # (1, 4) or (1, n, 1, 4)
        flags = st.node_i[:, 3]
        for c, j in enumerate((0, 1, 2, 6)):
            val = bp[0, c] if P.bmode == 1 else bp.reshape(st.n, 4)[:, c]
            mask = (flags & (2 if j < 6 else 4)) != 0
            node_p[:, j] = torch.where(mask, val.to(node_p.dtype) * torch.ones_like(node_p[:, j]), node_p[:, j])
    q.node_p = node_p
    q.v0 = P.v0[0].double()
    q.gates0 = [g[0] for g in P.gates0]
    q.stim = None if P.stim is None else P.stim[:, :, 0].contiguous()
    q.stim_scale = P.stim_scale
    return q


def _to_numpy(x):
    """Tensors in nested lists / tuples / dicts as numpy arrays (they cross a process
    boundary as plain bytes; torch tensors would go through shared-memory handles)."""
    if isinstance(x, torch.Tensor):
        return ("__tensor__", x.detach().cpu().numpy())
    if isinstance(x, dict):
        return {k: _to_numpy(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return type(x)(_to_numpy(v) for v in x)
    return x


def _to_torch(x):
    if isinstance(x, tuple) and len(x) == 2 and isinstance(x[0], str) and x[0] == "__tensor__":
        return torch.from_numpy(x[1])
    if isinstance(x, dict):
        return {k: _to_torch(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return type(x)(_to_torch(v) for v in x)
    return x


def _prepare_in_worker(args):
    """Worker-process side of `_prepare_payload`: numpy in, numpy out."""
    torch.set_num_threads(1)
    q = _prepare_payload(*_to_torch(args))
    q.__dict__ = _to_numpy(q.__dict__)
    return q


class _Live:
    """An admitted simulation: its prepared problem and arena placements."""

    def __init__(self, handle: SimHandle, P: _Prepared):
        self.handle, self.P = handle, P
        self.n, self.T, self.n_rec = P.n, P.T, P.n_rec
        self.nseg = int(P.seg.shape[0])
        self.ncs = int(P.seg_child.shape[0]) if self.nseg else 0
        self.nchild = int(P.child_idx.shape[0])
        self.n_stim = 0 if P.stim is None else int(P.stim_scale.shape[0])
        self.aux_size = 4 * self.nseg + (P.L + 2) + self.ncs + self.nchild + max(self.n_rec, 1)


class SimServer:
    """See the module docstring. `dtype` float32 (mode of `integrate.set_float32_mode`)
    or float64; `delta_t` is shared by all requests; `max_steps` bounds a request's
    duration (the stimulus table holds `max_steps` rows). `prep_processes` > 0 prepares
    requests in that many worker processes (preparation is Python-bound, so threads do
    not add throughput; tensors cross the process boundary as numpy arrays; the workers
    are started with "spawn", so a script that creates the server does its work under
    `if __name__ == "__main__":`); otherwise `prep_threads` > 0 prepares requests in that
    many background threads (0: inside `submit`); with more than one thread, requests
    must not share a module object (its schedule cache is built on first use). `close()`
    ends the workers."""

    def __init__(self, device=None, dtype=torch.float32, delta_t: float = 0.025,
                 hbm_budget: int = 2 << 30, max_active: int = 1024, chunk_steps: int = 200,
                 max_steps: int = 4001, max_stim_entries: int = 4096, autotune: bool = False,
                 thread_order=(256, 512, 1024), prep_threads: int = 1, prep_processes: int = 0):
        device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        _require_gpu(device)
        self.device, self.user_dtype, self.dt = device, dtype, float(delta_t)
        mode = importlib.import_module("jaxhip.integrate")._F32_MODE
        self.gdtype = torch.float64 if (dtype == torch.float64 or mode == "mixed") else torch.float32
        self.ldtype = dtype
        self.chunk_steps, self.max_steps, self.autotune = int(chunk_steps), int(max_steps), autotune
        self.thread_order = tuple(thread_order)
        gs, ls = torch.tensor([], dtype=self.gdtype).element_size(), torch.tensor([], dtype=self.ldtype).element_size()
        self.gsize, self.lsize = gs, ls
        # Arena capacities from the budget: per node 8 int32 + 4 solve + 8 parameters +
        # float64 voltage + 3 gates + the axial scratch; schedules ~8 int32 per node;
        # recordings get a third.
        per_node = 32 + 4 * ls + 8 * gs + 8 + 4 * gs + 32
        stim_bytes = max_steps * max_stim_entries * gs + max_stim_entries * (4 + gs)
        rest = max(int(hbm_budget) - stim_bytes, 1 << 20)
        self.node_cap = int(rest * 2 // 3 // per_node)
        self.rec_cap = int(rest // 3 // gs)
        dev = device
        self.node_i = torch.zeros((self.node_cap, _lib.JXHIP_NODE_ICOLS), dtype=torch.int32, device=dev)
        self.node_f = torch.zeros((self.node_cap, _lib.JXHIP_NODE_FCOLS), dtype=self.ldtype, device=dev)
        self.node_p = torch.zeros((self.node_cap, _lib.JXHIP_NPARAM), dtype=self.gdtype, device=dev)
        self.v = torch.zeros(self.node_cap, dtype=torch.float64, device=dev)
        self.gm, self.gh, self.gn = (torch.zeros(self.node_cap, dtype=self.gdtype, device=dev) for _ in range(3))
        # Scratch of the streaming kernel (axial part of every item's residual), same places.
        self.work_ax = torch.zeros(self.node_cap, dtype=self.gdtype, device=dev)
        self.aux = torch.zeros(8 * self.node_cap, dtype=torch.int32, device=dev)
        self.rec = torch.zeros(self.rec_cap, dtype=self.gdtype, device=dev)
        self.stim = torch.zeros((max_steps, max_stim_entries, 1), dtype=self.gdtype, device=dev)
        self.stim_col = torch.zeros(max_stim_entries, dtype=torch.int32, device=dev)
        self.stim_scale = torch.zeros(max_stim_entries, dtype=self.gdtype, device=dev)
        self.topos = torch.zeros((max_active, _lib.JXHIP_TOPO_COLS), dtype=torch.int32, device=dev)
        self._dummy = torch.zeros(1, dtype=self.gdtype, device=dev)
        self.nodes_a, self.aux_a = _Arena(self.node_cap), _Arena(self.aux.numel())
        self.rec_a, self.stim_a = _Arena(self.rec_cap), _Arena(max_stim_entries)
        self.slots = list(range(max_active))[::-1]
        self.queue: List[SimHandle] = []
        self.preparing: List = []     # (handle, future) in submission order
        self._procs = prep_processes > 0
        if self._procs:
            # Worker processes started fresh ("spawn": nothing of this process's GPU
            # context is inherited); they import jaxhip and never touch the GPU.
            self._pool = cf.ProcessPoolExecutor(max_workers=prep_processes,
                                                mp_context=multiprocessing.get_context("spawn"))
        else:
            self._pool = cf.ThreadPoolExecutor(max_workers=prep_threads) if prep_threads > 0 else None
        self.live: Dict[int, _Live] = {}
        self.inflight: List = []      # (event, [live finishing], pinned progs)
        self.handles: List[SimHandle] = []
        self._next_id = 0
        self._cfg_cache: Dict[int, LaunchConfig] = {}
        self.stream = torch.cuda.current_stream(dev)
        self.up_stream = torch.cuda.Stream(dev)
        self.launches = 0
        self.steps_run = 0

    # ---------------------------------------------------------------- requests
    def submit(self, request: SimRequest) -> SimHandle:
        """Registers a request and starts its preparation; returns its handle at once."""
        h = SimHandle(self._next_id, request)
        self._next_id += 1
        self.handles.append(h)
        # Prepared on the host (no GPU work, so a submission never waits for a running
        # chunk); the tables go to the arenas at admission by asynchronous copies.
        r = h.request
        args = (r.module, r.params, r.param_state, r.data_stimuli, r.t_max, r.all_states, self.dt,
                self.user_dtype, importlib.import_module("jaxhip.integrate")._F32_MODE, self.max_steps)
        if self._pool is None:
            self._accept(h, _prepare_payload(*args))
            self.queue.append(h)
        elif self._procs:
            self.preparing.append((h, self._pool.submit(_prepare_in_worker, _to_numpy(args))))
        else:
            self.preparing.append((h, self._pool.submit(_prepare_payload, *args)))
        return h

    def _accept(self, h: SimHandle, P: _Prepared) -> None:
        if self._procs:
            P.__dict__ = _to_torch(P.__dict__)
        L = _Live(h, P)
        if (L.n > self.node_cap or L.aux_size + 4 > self.aux.numel()
                or (L.T + 1) * max(L.n_rec, 1) > self.rec_cap or L.n_stim > self.stim_col.numel()):
            raise ValueError("the request does not fit the server's HBM budget")
        if not self._candidates(L.n, P.sched):
            raise ValueError(f"the request's tree ({L.n} nodes) does not fit a workgroup's LDS")
        h.T = P.T
        h._P = P

    def _collect_prepared(self, wait: bool = False) -> None:
        """Moves prepared requests (in submission order) to the admission queue; a failed
        preparation raises here. `wait`: block for the oldest one if none is ready."""
        while self.preparing:
            h, fut = self.preparing[0]
            if not fut.done():
                if not wait:
                    return
                fut.result()
            self._accept(h, fut.result())     # re-raises a preparation error
            self.preparing.pop(0)
            self.queue.append(h)
            wait = False

    def close(self) -> None:
        """Ends the preparation workers (threads or processes)."""
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None

    def outstanding(self) -> int:
        """Requests not finished: being prepared, queued for admission or running."""
        return len(self.preparing) + len(self.queue) + len(self.live)

    def _try_admit(self, h: SimHandle) -> bool:
        P = h._P
        L = _Live(h, P)
        if not self.slots:
            return False
        allocs = []
        for arena, size, align in ((self.nodes_a, L.n, 1), (self.aux_a, L.aux_size, 4),
                                   (self.rec_a, (L.T + 1) * max(L.n_rec, 1), 1),
                                   (self.stim_a, max(L.n_stim, 1), 1)):
            off = arena.alloc(size, align)
            if off is None:
                for a, o, sz in allocs:
                    a.free(o, sz)
                return False
            allocs.append((arena, off, size))
        L.allocs = allocs
        L.node_off, L.aux_off, L.rec_off, L.stim_off = (o for _, o, _ in allocs)
        L.slot = self.slots.pop()
        self._upload(L)
        h.t_admit = time.perf_counter()
        self.live[h.id] = L
        return True

    def _upload(self, L: _Live) -> None:
        """Copies the request's tables and initial state into its arena places: pinned
        host buffers, asynchronous copies on the upload stream (the compute stream waits
        for it before the next chunk)."""
        P = L.P
        n, no = L.n, L.node_off
        pins = []

        def put(dst, src):
            src = src.contiguous()
            if src.device.type == "cpu":
                src = src.pin_memory()
                pins.append(src)
            dst.copy_(src, non_blocking=True)

        ni = P.node_i.clone()
        ni[:, 4:6] += L.stim_off
        node_p = P.node_p
        a = L.aux_off
        sb = a // 4
        parts = [P.seg.reshape(-1), P.seg_level_ptr[:P.L + 2], P.seg_child[:L.ncs], P.child_idx,
                 P.rec_node[:max(L.n_rec, 1)]]
        offs = np.cumsum([a] + [int(x.numel()) for x in parts])
        lb, scb, cb, rb = (int(x) for x in offs[1:5])
        row = torch.tensor([n, P.L, no, lb, cb, L.n_rec, rb, 0, 0, 0, 0, 0, sb, scb, 0, 0],
                           dtype=torch.int32)
        with torch.cuda.stream(self.up_stream):
            put(self.node_i[no:no + n], ni)
            put(self.node_f[no:no + n], P.node_f.to(self.ldtype))
            put(self.node_p[no:no + n], node_p.to(self.gdtype))
            put(self.v[no:no + n], P.v0)
            for buf, g in zip((self.gm, self.gh, self.gn), P.gates0):
                put(buf[no:no + n], g.to(self.gdtype))
            put(self.aux[a:int(offs[-1])], torch.cat([x.to(torch.int32) for x in parts]))
            put(self.topos[L.slot], row)
            if L.n_stim:
                e0 = L.stim_off
                put(self.stim_col[e0:e0 + L.n_stim], P.stim_col[:L.n_stim] + e0)
                put(self.stim_scale[e0:e0 + L.n_stim], P.stim_scale.to(self.gdtype))
                put(self.stim[:P.T, e0:e0 + P.stim.shape[1], 0], P.stim.to(self.gdtype))
        L.pins = pins

    def _retire(self, L: _Live) -> None:
        h = L.handle
        r = self.rec[L.rec_off:L.rec_off + (L.T + 1) * max(L.n_rec, 1)]
        h.result = r.reshape(L.T + 1, max(L.n_rec, 1))[:, :L.n_rec].t().to(self.user_dtype).clone()
        h.t_done = time.perf_counter()
        L.pins = None
        for arena, off, size in L.allocs:
            arena.free(off, size)
        self.slots.append(L.slot)
        del self.live[h.id]
        h._P = None

    # ---------------------------------------------------------------- launches
    def _candidates(self, n_max: int, sched_max: int) -> List[LaunchConfig]:
        """Launch shapes that can run a chunk whose largest tree has `n_max` nodes and a
        staged schedule of `sched_max` bytes: register-resident shapes whose registers
        hold the tree and whose LDS (schedule, coefficients, system) fits, then the
        streaming kernel (coupling coefficients in LDS, or read from L2 when only that
        fits)."""
        out = []
        regs_lds = lds_bytes(n_max, 0, 1, self.lsize, sched_max)
        for thr in self.thread_order:
            if items_per_thread(n_max, 1, thr, self.gsize, self.lsize) and regs_lds <= _LDS_MAX:
                out.append(LaunchConfig(1, thr))
        g = regs_lds > _LDS_MAX
        if lds_bytes(n_max, 0, 1, self.lsize, sched_max, g) <= _LDS_MAX:
            out.append(LaunchConfig(1, 1024, stream=True, gcoef=g))
        return out

    def _fits(self, cfg: LaunchConfig, n_max: int, sched_max: int) -> bool:
        return any(c == cfg for c in self._candidates(n_max, sched_max))

    def _config(self, n_max: int, sched_max: int) -> LaunchConfig:
        # A program is one simulation: the first thread count of `thread_order` whose
        # registers hold the largest tree (small workgroups let several programs share a
        # CU, which hides the latency of each program's tree solve), else streaming.
        cands = self._candidates(n_max, sched_max)
        if not cands:
            raise ValueError(f"a tree of {n_max} nodes does not fit a workgroup's LDS")
        return cands[0]

    def _args(self, progs_dev, n_progs, cfg, n_max, sched_max):
        a = _lib.FusedArgs()
        a.progs, a.n_progs, a.topos = progs_dev.data_ptr(), n_progs, self.topos.data_ptr()
        base = self.aux.data_ptr()
        a.seg = a.seg_level_ptr = a.seg_child = a.child_idx = a.rec_node = base
        a.node_i, a.node_f, a.node_p = self.node_i.data_ptr(), self.node_f.data_ptr(), self.node_p.data_ptr()
        a.bparam, a.bparam_mode = self._dummy.data_ptr(), 0
        a.v, a.gate_m, a.gate_h, a.gate_n = (t.data_ptr() for t in (self.v, self.gm, self.gh, self.gn))
        a.stim, a.stim_cols, a.stim_b = self.stim.data_ptr(), self.stim.shape[1], 1
        a.stim_col, a.stim_scale = self.stim_col.data_ptr(), self.stim_scale.data_ptr()
        a.rec = self.rec.data_ptr()
        a.syn_model = -1
        a.sims_total = len(self.slots) + len(self.live)
        a.dt, a.diff, a.S = self.dt, 1e-3, 1
        a.threads = cfg.threads
        a.items_per_thread = 0 if cfg.stream else items_per_thread(n_max, 1, cfg.threads, self.gsize, self.lsize)
        a.coef_global = int(cfg.stream and cfg.gcoef)
        a.lds_bytes = lds_bytes(n_max, 0, 1, self.lsize, sched_max, bool(a.coef_global))
        a.work_ax = self.work_ax.data_ptr()
        a.max_items = n_max
        return a

    def _kernel(self):
        if self.gdtype == torch.float64 and self.ldtype == torch.float32:
            return _lib.fn("jxhip_fused_forward_mixed")
        return _lib.fn("jxhip_fused_forward", self.ldtype)

    def _launch_chunk(self) -> None:
        lives = list(self.live.values())
        progs = np.zeros((len(lives), _lib.JXHIP_PROG_COLS), dtype=np.int64)
        finishing = []
        for r, L in enumerate(lives):
            t0 = L.handle.steps_done
            t1 = min(t0 + self.chunk_steps, L.T)
            progs[r] = [L.slot, 1, t0, t1, L.node_off, L.slot, L.rec_off + t0 * max(L.n_rec, 1), 0]
            if t1 >= L.T:
                finishing.append(L)
        n_max = max(L.n for L in lives)
        sched_max = max(L.P.sched for L in lives)
        # Shapes are cached per size class; a cached shape is used only if it still fits
        # (a larger tree of the same class may need the streaming kernel).
        key = n_max.bit_length()
        cfg = self._cfg_cache.get(key)
        if cfg is None or not self._fits(cfg, n_max, sched_max):
            cfg = self._tune(n_max, sched_max, progs) if self.autotune else self._config(n_max, sched_max)
            self._cfg_cache[key] = cfg
        self.stream.wait_stream(self.up_stream)
        host = torch.from_numpy(progs).pin_memory()
        dev = host.to(self.device, non_blocking=True)
        a = self._args(dev, len(lives), cfg, n_max, sched_max)
        _lib.check(self._kernel()(a, _lib.stream_ptr(self.device)), "fused_forward (served chunk)")
        # Only a launched chunk advances its simulations.
        for r, L in enumerate(lives):
            L.handle.steps_done = int(progs[r, 3])
        ev = torch.cuda.Event()
        ev.record(self.stream)
        self.inflight.append((ev, finishing, host, dev))
        self.launches += 1
        self.steps_run += int((progs[:, 3] - progs[:, 2]).sum())
        # Finishing simulations leave the live set only when their chunk has completed;
        # they are not launched again (steps_done == T).

    def _tune(self, n_max, sched_max, progs) -> LaunchConfig:
        """Measure the candidate shapes on copies of the live state (a few steps)."""
        cands = self._candidates(n_max, sched_max)
        if not cands:
            raise ValueError(f"a tree of {n_max} nodes does not fit a workgroup's LDS")
        saved = [t.clone() for t in (self.v, self.gm, self.gh, self.gn, self.rec)]
        p = progs.copy()
        p[:, 3] = np.minimum(p[:, 2] + 16, p[:, 3])
        dev = torch.from_numpy(p).to(self.device)
        best, best_t = cands[0], float("inf")
        for c in cands:
            a = self._args(dev, len(p), c, n_max, sched_max)
            f = self._kernel()
            st, en = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            st.record(self.stream)
            code = f(a, _lib.stream_ptr(self.device))
            en.record(self.stream)
            torch.cuda.synchronize(self.device)
            if code == 0 and st.elapsed_time(en) < best_t:
                best, best_t = c, st.elapsed_time(en)
        for t, s in zip((self.v, self.gm, self.gh, self.gn, self.rec), saved):
            t.copy_(s)
        return best

    # ---------------------------------------------------------------- driving
    def poll(self, block: bool = False) -> int:
        """Retire completed chunks, admit queued requests, launch the next chunk.
        Returns the number of simulations still queued or running."""
        while self.inflight and (block or self.inflight[0][0].query()):
            ev, finishing, _, _ = self.inflight.pop(0)
            ev.synchronize()
            for L in finishing:
                self._retire(L)
            block = False
        # With nothing running or queued, a blocking poll waits for a preparation.
        self._collect_prepared(wait=block and not self.queue and not self.live)
        pending = []
        for h in self.queue:
            if not self._try_admit(h):
                pending.append(h)
        self.queue = pending
        runnable = [L for L in self.live.values() if L.handle.steps_done < L.T]
        if runnable and len(self.inflight) < 2:
            saved = self.live
            self.live = {k: L for k, L in saved.items() if L.handle.steps_done < L.T}
            try:
                self._launch_chunk()
            finally:
                self.live = saved
        return self.outstanding()

    def run_until_idle(self) -> None:
        while self.poll(block=bool(self.inflight) or bool(self.preparing)):
            pass
        while self.inflight:
            self.poll(block=True)

    def latency_stats(self) -> Dict[str, float]:
        lat = np.array([h.latency for h in self.handles if h.latency is not None])
        if not len(lat):
            return {}
        return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "mean_s": float(lat.mean()), "n": int(len(lat))}
