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

import bisect
import concurrent.futures as cf
import heapq
import importlib
import multiprocessing
import sys
import time
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Tuple

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
        fl = self.free_list
        j = bisect.bisect_left(fl, (off, size))
        # Merge with the free ranges just before and after it.
        if j > 0 and fl[j - 1][0] + fl[j - 1][1] == off:
            j -= 1
            off, size = fl[j][0], fl[j][1] + size
            del fl[j]
        if j < len(fl) and off + size == fl[j][0]:
            size += fl[j][1]
            del fl[j]
        fl.insert(j, (off, size))


class _Staging:
    """Pinned host buffers and their device copies, one pair per element type, reused for
    every upload (grown when a batch needs more). `put` waits until the previous upload
    has left the host buffers, packs the arrays, starts one host-to-device copy per type
    on `stream` and returns device views of the arrays in the same structure."""

    def __init__(self, device, dtypes: Dict[str, torch.dtype]):
        self.device, self.dtypes = device, dtypes
        self.host: Dict[str, torch.Tensor] = {}
        self.dev: Dict[str, torch.Tensor] = {}
        self.done: Optional[torch.cuda.Event] = None

    def put(self, groups: Dict[str, List[np.ndarray]], stream) -> Dict[str, List[torch.Tensor]]:
        if self.done is not None:
            self.done.synchronize()
        out: Dict[str, List[torch.Tensor]] = {}
        with torch.cuda.stream(stream):
            for kind, arrays in groups.items():
                total = sum(int(x.size) for x in arrays)
                cap = self.host[kind].numel() if kind in self.host else 0
                if total > cap:
                    # Earlier scatters on `stream` may still read the old device buffer.
                    stream.synchronize()
                    cap = max(total, 2 * cap, 1 << 16)
                    self.host[kind] = torch.empty(cap, dtype=self.dtypes[kind], pin_memory=True)
                    self.dev[kind] = torch.empty(cap, dtype=self.dtypes[kind], device=self.device)
                hbuf, views, off = self.host[kind].numpy(), [], 0
                for x in arrays:
                    hbuf[off:off + x.size] = x.reshape(-1)
                    views.append((off, x.size))
                    off += x.size
                d = self.dev[kind]
                d[:total].copy_(self.host[kind][:total], non_blocking=True)
                out[kind] = [d[o:o + sz] for o, sz in views]
            self.done = torch.cuda.Event()
            self.done.record(stream)
        return out


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
        self.t_prepared: Optional[float] = None
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
    request's values merged in), initial state and stimulus. Every integer array lies in
    one int32 array (`ints`) and every floating array in one float64 array (`reals`),
    back to back, so that a prepared request crosses a process boundary as two buffers;
    the arrays are read as named views (`layout`: name -> (kind, offset, shape))."""

    _INTS = ("node_i", "seg", "seg_level_ptr", "seg_child", "child_idx", "rec_node", "stim_col")
    _REALS = ("node_f", "node_p", "v0", "gate_m", "gate_h", "gate_n", "stim_scale", "stim")

    def pack(self, arrays: Dict[str, Any]) -> None:
        layout, parts, off = {}, {"i": [], "r": []}, {"i": 0, "r": 0}
        for name in self._INTS + self._REALS:
            x = arrays.get(name)
            if x is None:
                continue
            kind = "i" if name in self._INTS else "r"
            x = _np(x, np.int32 if kind == "i" else np.float64)
            layout[name] = (kind, off[kind], x.shape, x.size)
            parts[kind].append(x.reshape(-1))
            off[kind] += x.size
        self.layout = layout
        self.ints = np.concatenate(parts["i"]) if parts["i"] else np.zeros(0, np.int32)
        self.reals = np.concatenate(parts["r"]) if parts["r"] else np.zeros(0, np.float64)

    def __getattr__(self, name):
        lay = self.__dict__.get("layout")
        if lay is None or name.startswith("__"):
            raise AttributeError(name)
        if name == "gates0":
            return [getattr(self, k) for k in ("gate_m", "gate_h", "gate_n")]
        if name not in lay:
            if name in self._INTS + self._REALS:
                return None
            raise AttributeError(name)
        kind, off, shape, size = lay[name]
        buf = self.ints if kind == "i" else self.reals
        return buf[off:off + size].reshape(shape)


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
    node_p = P.node_p.clone()
    if P.bmode:
        bp = P.bparam_for(1)                       # (1, 4) or (1, n, 1, 4)
        flags = st.node_i[:, 3]
        for c, j in enumerate((0, 1, 2, 6)):
            val = bp[0, c] if P.bmode == 1 else bp.reshape(st.n, 4)[:, c]
            mask = (flags & (2 if j < 6 else 4)) != 0
            node_p[:, j] = torch.where(mask, val.to(node_p.dtype) * torch.ones_like(node_p[:, j]), node_p[:, j])
    nseg = int(st.seg.shape[0]) if st.n else 0
    arrays = {"node_i": st.node_i, "seg": st.seg, "seg_level_ptr": st.seg_level_ptr[:st.L + 2],
              "seg_child": st.seg_child, "child_idx": st.child_idx, "rec_node": st.rec_node,
              "node_f": P.node_f, "node_p": node_p, "v0": P.v0[0],
              "gate_m": P.gates0[0][0], "gate_h": P.gates0[1][0], "gate_n": P.gates0[2][0]}
    if nseg == 0:
        arrays["seg"] = np.zeros((0, 4), np.int32)
    if P.stim is not None:
        stim = P.stim[:, :, 0]
        if stim.shape[1] > len(P.stim_scale):
            raise ValueError("more stimulus columns than stimulus entries")
        arrays.update(stim=stim, stim_scale=P.stim_scale, stim_col=st.stim_col[:len(P.stim_scale)])
    q.pack(arrays)
    return q


def _request_length(r: SimRequest, dt: float) -> Tuple[int, int]:
    """(time steps, compartments) of a request (its duration from `t_max`, else from its
    stimulus)."""
    if r.t_max is not None:
        T = int(r.t_max // dt + 1)
    elif r.data_stimuli is not None and len(r.data_stimuli[1]):
        T = int(np.asarray(r.data_stimuli[1][0].shape)[-1])
    else:
        T = 1
    return T, int(r.module._n_comp)


def _np(x, dtype=None):
    """A contiguous numpy copy that owns its memory (CPU tensors and arrays alike: a
    view of a tensor's storage would dangle once torch moves that storage, as pickling a
    module with cached tensors into a worker process does)."""
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.array(x, dtype=dtype, order="C", copy=True)


def _to_numpy(x):
    """Tensors in nested lists / tuples / dicts as numpy arrays (they cross a process
    boundary as plain bytes; torch tensors would go through shared-memory handles)."""
    if isinstance(x, torch.Tensor):
        if x.device.type != "cpu" or x.requires_grad:
            x = x.detach().cpu()
        return ("__tensor__", x.numpy())
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
    return _prepare_payload(*_to_torch(args))      # plain numpy arrays already


def _prepare_batch(fn, batch):
    """Several requests in one pool task: per request (True, payload) or (False, the
    exception its preparation raised)."""
    out = []
    for args in batch:
        try:
            out.append((True, fn(*args)))
        except Exception as e:  # noqa: BLE001 - re-raised by the server for that request
            out.append((False, e))
    return out


def _prepare_batch_in_worker(batch):
    return _prepare_batch(_prepare_in_worker, [(a,) for a in batch])


class _Done:
    """A finished pool task's remaining results (a future that is already done)."""

    def __init__(self, results):
        self._results = results

    def done(self) -> bool:
        return True

    def result(self):
        return self._results


class _Live:
    """An admitted simulation: its prepared problem and arena placements."""

    def __init__(self, handle: SimHandle, P: _Prepared):
        self.handle, self.P = handle, P
        self.n, self.T, self.n_rec = P.n, P.T, P.n_rec
        lay = P.layout
        self.nseg = int(lay["seg"][2][0])
        self.ncs = int(lay["seg_child"][2][0]) if self.nseg else 0
        self.nchild = int(lay["child_idx"][2][0])
        self.n_stim = int(lay["stim_scale"][2][0]) if "stim" in lay else 0
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
    must not share a module object (its schedule cache is built on first use).
    `prep_order` ("fifo" or "longest_first") is the order in which submitted requests
    are handed to the preparation workers. With worker processes the interpreter's switch
    interval is set to `switch_interval` (None: left alone) while the server is open, so
    that the pool's hand-off threads keep the workers busy. `close()` ends the workers
    and restores the switch interval."""

    def __init__(self, device=None, dtype=torch.float32, delta_t: float = 0.025,
                 hbm_budget: int = 2 << 30, max_active: int = 1024, chunk_steps: int = 200,
                 max_steps: int = 4001, max_stim_entries: int = 4096, autotune: bool = False,
                 thread_order=(256, 512, 1024), prep_threads: int = 1, prep_processes: int = 0,
                 prep_order: str = "fifo", switch_interval: Optional[float] = 1e-3):
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
        self._s_run = np.zeros(max_active, dtype=bool)
        self._s_t0, self._s_T, self._s_node_off, self._s_rec_off, self._s_nrec1, self._s_n, self._s_sched = (
            np.zeros(max_active, dtype=np.int64) for _ in range(7))
        self._s_live: List[Optional[_Live]] = [None] * max_active
        self.max_inflight = 3
        self.queue: List[SimHandle] = []
        self.preparing: List = []     # (handles, future of their payloads) in dispatch order
        self._procs = prep_processes > 0
        self._switch = None
        if self._procs and switch_interval is not None:
            # The pool's hand-off threads need the interpreter often: with the default 5 ms
            # switch interval a busy serving loop starves them and the workers sit idle.
            self._switch = sys.getswitchinterval()
            sys.setswitchinterval(switch_interval)
        if self._procs:
            # Worker processes started fresh ("spawn": nothing of this process's GPU
            # context is inherited); they import jaxhip and never touch the GPU.
            self._pool = cf.ProcessPoolExecutor(max_workers=prep_processes,
                                                mp_context=multiprocessing.get_context("spawn"))
        else:
            self._pool = cf.ThreadPoolExecutor(max_workers=prep_threads) if prep_threads > 0 else None
        if prep_order not in ("fifo", "longest_first"):
            raise ValueError(f"prep_order must be 'fifo' or 'longest_first', not {prep_order!r}")
        self.prep_order = prep_order
        # Submitted requests not yet handed to a preparation worker: at most a few per
        # worker are queued in the pool, the rest wait here, so that `prep_order` decides.
        self.waiting: List = []       # heap of (priority, id, handle, args)
        self._prep_depth = 2 * max(prep_processes, prep_threads, 1)
        self.live: Dict[int, _Live] = {}
        self.inflight: List = []      # (event, [live finishing], pinned progs)
        self.handles: List[SimHandle] = []
        self._next_id = 0
        self._cfg_cache: Dict[int, LaunchConfig] = {}
        self.stream = torch.cuda.current_stream(dev)
        self.up_stream = torch.cuda.Stream(dev)
        self._stage = _Staging(dev, {"i64": torch.int64, "i32": torch.int32, "l": self.ldtype,
                                     "g": self.gdtype, "d": torch.float64})
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
        else:
            if self.prep_order == "longest_first":
                T, nc = _request_length(r, self.dt)
                key = (-T, -nc)
            else:
                key = (0, 0)
            heapq.heappush(self.waiting, (key, h.id, h, args))
            self._dispatch()
        return h

    def _dispatch(self) -> None:
        """Hands waiting requests to the preparation pool while it holds fewer than a few
        per worker: in submission order ("fifo") or the most time steps first, larger
        trees first among equal durations ("longest_first": a simulation's steps run
        chunk after chunk, so a burst's long simulations start first, which shortens the
        time to finish all of them)."""
        while self.waiting and len(self.preparing) < self._prep_depth:
            # With many requests waiting (a burst) small ones go to a worker together, up
            # to 8 and about 2000 compartments per task (fewer hand-offs through the pool
            # per request); a large tree goes alone, so that it starts early.
            nb = max(1, min(8, len(self.waiting) // self._prep_depth))
            items = [heapq.heappop(self.waiting)]
            ncomp = items[0][2].request.module._n_comp
            while len(items) < nb and self.waiting and ncomp + self.waiting[0][2].request.module._n_comp <= 2048:
                items.append(heapq.heappop(self.waiting))
                ncomp += items[-1][2].request.module._n_comp
            hs = [h for _, _, h, _ in items]
            if self._procs:
                fut = self._pool.submit(_prepare_batch_in_worker, [_to_numpy(args) for _, _, _, args in items])
            else:
                fut = self._pool.submit(_prepare_batch, _prepare_payload, [args for _, _, _, args in items])
            self.preparing.append((hs, fut))

    def _accept(self, h: SimHandle, P: _Prepared) -> None:
        L = _Live(h, P)
        if (L.n > self.node_cap or L.aux_size + 4 > self.aux.numel()
                or (L.T + 1) * max(L.n_rec, 1) > self.rec_cap or L.n_stim > self.stim_col.numel()):
            raise ValueError("the request does not fit the server's HBM budget")
        if not self._candidates(L.n, P.sched):
            raise ValueError(f"the request's tree ({L.n} nodes) does not fit a workgroup's LDS")
        h.T = P.T
        h._P = P
        h.t_prepared = time.perf_counter()

    def _collect_prepared(self, wait: bool = False, limit: Optional[int] = None) -> None:
        """Moves prepared requests (in the order they were handed to the pool, at most
        `limit`) to the admission queue and refills the pool; a failed preparation raises
        here. `wait`: block for the oldest one if none is ready."""
        done = 0
        while self.preparing and (limit is None or done < limit):
            hs, fut = self.preparing[0]
            if not fut.done():
                if not wait:
                    break
                fut.result()
            self.preparing.pop(0)
            results = fut.result()
            for j, (h, (ok, P)) in enumerate(zip(hs, results)):
                try:
                    if not ok:
                        raise P
                    self._accept(h, P)
                except Exception:
                    # The failed request is dropped; the rest of the task's requests go
                    # back to the head of the line.
                    if j + 1 < len(hs):
                        self.preparing.insert(0, (hs[j + 1:], _Done(results[j + 1:])))
                    raise
                self.queue.append(h)
                done += 1
            wait = False
        self._dispatch()

    def close(self) -> None:
        """Ends the preparation workers (threads or processes) and restores the
        interpreter's switch interval."""
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None
        if self._switch is not None:
            sys.setswitchinterval(self._switch)
            self._switch = None

    def outstanding(self) -> int:
        """Requests not finished: waiting for or in preparation, queued for admission or
        running."""
        return (len(self.waiting) + sum(len(hs) for hs, _ in self.preparing) + len(self.queue)
                + len(self.live))

    def _try_admit(self, h: SimHandle) -> Optional[_Live]:
        """Places an admitted request in the arenas (its data is uploaded with the other
        requests admitted in the same poll, `_upload_batch`)."""
        P = h._P
        L = _Live(h, P)
        if not self.slots:
            return None
        allocs = []
        for arena, size, align in ((self.nodes_a, L.n, 1), (self.aux_a, L.aux_size, 4),
                                   (self.rec_a, (L.T + 1) * max(L.n_rec, 1), 1),
                                   (self.stim_a, max(L.n_stim, 1), 1)):
            off = arena.alloc(size, align)
            if off is None:
                for a, o, sz in allocs:
                    a.free(o, sz)
                return None
            allocs.append((arena, off, size))
        L.allocs = allocs
        L.node_off, L.aux_off, L.rec_off, L.stim_off = (o for _, o, _ in allocs)
        L.slot = sl = self.slots.pop()
        h.t_admit = time.perf_counter()
        self.live[h.id] = L
        # Per-slot launch fields (a chunk's program table is built from them vectorized).
        self._s_run[sl] = L.T > 0
        self._s_t0[sl] = 0
        self._s_T[sl] = L.T
        self._s_node_off[sl] = L.node_off
        self._s_rec_off[sl] = L.rec_off
        self._s_nrec1[sl] = max(L.n_rec, 1)
        self._s_n[sl] = L.n
        self._s_sched[sl] = P.sched
        self._s_live[sl] = L
        return L

    def _upload_batch(self, lives: List[_Live]) -> None:
        """Copies the tables and initial states of the requests admitted in one poll into
        their arena places: every array is packed on the host into one pinned staging
        buffer per element type (reused from poll to poll), copied to the device in one
        transfer per type and scattered to the arena places by index (a handful of
        operations per poll, however many requests); all on the upload stream, which the
        compute stream waits for before the next chunk."""
        if not lives:
            return
        m = len(lives)
        # Per request: arena places, sizes and where each array lies in its payload.
        f = np.zeros((13, m), dtype=np.int64)
        for r, L in enumerate(lives):
            lay, P = L.P.layout, L.P
            st = lay.get("stim")
            f[:, r] = (L.node_off, L.n, L.aux_off, L.aux_size, L.stim_off, L.n_stim,
                       st[2][1] if st else 0, st[2][0] if st else 0, lay["seg"][1],
                       lay["stim_col"][1] if st else 0, lay["node_f"][1], lay["stim_scale"][1] if st else 0,
                       lay["stim"][1] if st else 0)
        node_off, n, aux_off, aux_size, stim_off, n_stim, ncol, Tst, seg0, scol0, nf0, ss0, stv0 = f
        ilen = np.array([L.P.ints.size for L in lives], dtype=np.int64)
        rlen = np.array([L.P.reals.size for L in lives], dtype=np.int64)
        ib = np.concatenate([[0], np.cumsum(ilen)[:-1]])
        rb = np.concatenate([[0], np.cumsum(rlen)[:-1]])
        I = np.concatenate([L.P.ints for L in lives])
        R = np.concatenate([L.P.reals for L in lives])

        def ranges(starts, lens):
            """Concatenated aranges [starts[r], starts[r] + lens[r])."""
            tot = int(lens.sum())
            return np.arange(tot, dtype=np.int64) - np.repeat(np.cumsum(lens) - lens - starts, lens)

        # Node rows: node_i (its stimulus range moved to the request's entries), node_f,
        # node_p, v0 and the three gates lie in this order in each payload (float arrays
        # one after another from node_f's offset).
        ni = I[ranges(ib, 8 * n)].reshape(-1, _lib.JXHIP_NODE_ICOLS)
        ni[:, 4:6] += np.repeat(stim_off, n)[:, None].astype(np.int32)
        nf = R[ranges(rb + nf0, 4 * n)]
        npar = R[ranges(rb + nf0 + 4 * n, 8 * n)]
        v0 = R[ranges(rb + nf0 + 12 * n, n)]
        gm, gh, gn = (R[ranges(rb + nf0 + (13 + q) * n, n)] for q in range(3))
        nidx = ranges(node_off, n)
        # Schedule tables, child lists and recorded nodes: one block per request.
        aidx = ranges(aux_off, aux_size)
        aval = I[ranges(ib + seg0, aux_size)]
        lb = aux_off + 4 * np.array([L.nseg for L in lives], dtype=np.int64)
        Ls = np.array([L.P.L for L in lives], dtype=np.int64)
        scb = lb + Ls + 2
        cb = scb + np.array([L.ncs for L in lives], dtype=np.int64)
        rbase = cb + np.array([L.nchild for L in lives], dtype=np.int64)
        topo = np.zeros((m, _lib.JXHIP_TOPO_COLS), dtype=np.int64)
        topo[:, 0], topo[:, 1], topo[:, 2], topo[:, 3], topo[:, 4] = n, Ls, node_off, lb, cb
        topo[:, 5] = [L.n_rec for L in lives]
        topo[:, 6], topo[:, 12], topo[:, 13] = rbase, aux_off // 4, scb
        slots = np.array([L.slot for L in lives], dtype=np.int64)
        i64 = [nidx, aidx, slots]
        i32 = [ni.reshape(-1), aval, topo.astype(np.int32).reshape(-1)]
        gvals = [npar, gm, gh, gn]
        has_stim = bool(n_stim.any())
        if has_stim:
            # Stimulus entries and the (T, columns) stimulus of each request at its
            # entries' columns of the shared table.
            i64.append(ranges(stim_off, n_stim))
            i32.append((I[ranges(ib + scol0, n_stim)] + np.repeat(stim_off, n_stim)).astype(np.int32))
            gvals.append(R[ranges(rb + ss0, n_stim)])
            cnt = Tst * ncol
            k = ranges(np.zeros(m, dtype=np.int64), cnt)
            c_rep = np.repeat(np.maximum(ncol, 1), cnt)
            i64.append(k // c_rep * self.stim.shape[1] + np.repeat(stim_off, cnt) + k % c_rep)
            gvals.append(R[ranges(rb + stv0, cnt)])
        dev = self._stage.put({"i64": i64, "i32": i32, "l": [nf], "g": gvals, "d": [v0]}, self.up_stream)
        with torch.cuda.stream(self.up_stream):
            nix, aix, six = dev["i64"][:3]
            self.node_i.index_copy_(0, nix, dev["i32"][0].view(-1, self.node_i.shape[1]))
            self.node_f.index_copy_(0, nix, dev["l"][0].view(-1, self.node_f.shape[1]))
            self.node_p.index_copy_(0, nix, dev["g"][0].view(-1, self.node_p.shape[1]))
            self.v.index_copy_(0, nix, dev["d"][0])
            for buf, x in zip((self.gm, self.gh, self.gn), dev["g"][1:4]):
                buf.index_copy_(0, nix, x)
            self.aux.index_copy_(0, aix, dev["i32"][1])
            self.topos.index_copy_(0, six, dev["i32"][2].view(-1, self.topos.shape[1]))
            if has_stim:
                eix, stx = dev["i64"][3:5]
                self.stim_col.index_copy_(0, eix, dev["i32"][3])
                self.stim_scale.index_copy_(0, eix, dev["g"][4])
                self.stim.view(-1).index_copy_(0, stx, dev["g"][5])

    def _retire_all(self, lives: List[_Live]) -> None:
        """Results of simulations whose last chunk has completed: their recordings are
        gathered from the arena in one copy (views of it per request), then their places
        are freed."""
        if not lives:
            return
        lens = np.array([(L.T + 1) * max(L.n_rec, 1) for L in lives], dtype=np.int64)
        offs = np.array([L.rec_off for L in lives], dtype=np.int64)
        idx = np.arange(int(lens.sum()), dtype=np.int64) - np.repeat(np.cumsum(lens) - lens - offs, lens)
        flat = self.rec[torch.from_numpy(idx).to(self.device, non_blocking=True)].to(self.user_dtype)
        pos = 0
        for L, ln in zip(lives, lens.tolist()):
            L.handle.result = flat[pos:pos + ln].view(L.T + 1, max(L.n_rec, 1))[:, :L.n_rec].t()
            pos += ln
            self._retire(L)

    def _retire(self, L: _Live) -> None:
        h = L.handle
        h.t_done = time.perf_counter()
        for arena, off, size in L.allocs:
            arena.free(off, size)
        h.steps_done = L.T
        self.slots.append(L.slot)
        self._s_run[L.slot] = False
        self._s_live[L.slot] = None
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
        """One launch advancing every admitted simulation with steps left by up to
        `chunk_steps` steps (program table built from the per-slot fields)."""
        act = np.flatnonzero(self._s_run)
        t0 = self._s_t0[act]
        T = self._s_T[act]
        t1 = np.minimum(t0 + self.chunk_steps, T)
        progs = np.zeros((len(act), _lib.JXHIP_PROG_COLS), dtype=np.int64)
        progs[:, 0] = act
        progs[:, 1] = 1
        progs[:, 2] = t0
        progs[:, 3] = t1
        progs[:, 4] = self._s_node_off[act]
        progs[:, 5] = act
        progs[:, 6] = self._s_rec_off[act] + t0 * self._s_nrec1[act]
        fin = act[t1 >= T]
        finishing = [self._s_live[sl] for sl in fin]
        n_max = int(self._s_n[act].max())
        sched_max = int(self._s_sched[act].max())
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
        a = self._args(dev, len(act), cfg, n_max, sched_max)
        _lib.check(self._kernel()(a, _lib.stream_ptr(self.device)), "fused_forward (served chunk)")
        # Only a launched chunk advances its simulations; finishing ones leave the live
        # set when their chunk has completed and are not launched again.
        self._s_t0[act] = t1
        self._s_run[fin] = False
        ev = torch.cuda.Event()
        ev.record(self.stream)
        self.inflight.append((ev, finishing, host, dev))
        self.launches += 1
        self.steps_run += int((t1 - t0).sum())

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
    def poll(self, block: bool = False, max_collect: int = 64) -> int:
        """Retire completed chunks, keep up to `max_inflight` chunks queued on the GPU,
        collect prepared requests (at most `max_collect` per call, so that launches stay
        regular), admit and upload them, launch again. Returns the number of simulations
        still waiting, in preparation, queued or running."""
        while self.inflight and (block or self.inflight[0][0].query()):
            ev, finishing, _, _ = self.inflight.pop(0)
            ev.synchronize()
            self._retire_all(finishing)
            block = False
        # The GPU first: the simulations already admitted get their next chunk before the
        # host work of new admissions.
        self._maybe_launch()
        # With nothing running or queued, a blocking poll waits for a preparation.
        self._collect_prepared(wait=block and not self.queue and not self.live, limit=max_collect)
        pending, admitted = [], []
        for h in self.queue:
            L = self._try_admit(h)
            if L is None:
                pending.append(h)
            else:
                admitted.append(L)
        self.queue = pending
        self._upload_batch(admitted)
        self._maybe_launch()
        return self.outstanding()

    def _maybe_launch(self) -> None:
        if len(self.inflight) < self.max_inflight and self._s_run.any():
            self._launch_chunk()

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
