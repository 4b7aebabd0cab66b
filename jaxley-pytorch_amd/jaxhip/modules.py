"""Neuron models: `Compartment`, `Branch`, `Cell`, `Network` and their views.

The same construction and configuration surface as jaxley-pytorch's modules
(jaxley_pytorch/modules/{compartment,branch,cell,network,base}.py) for what the hot
path needs: morphology, `insert`, `set`, `data_set`, `make_trainable`,
`get_parameters`, `record`, `stimulate`, `data_stimulate`, `init_states`, views by
`cell` / `branch` / `comp` / `loc`, and synapses (`connect`). Tables are numpy arrays
(one row per compartment); the compartment-edge tables and the DHS solve order are
built by `jaxhip.schedule` exactly as the reference builds them.
"""
from __future__ import annotations

import copy
from typing import Any, Dict, List, Optional, Sequence, Union

import numpy as np
import torch

from . import schedule as sch
from .geometry import cylinder_geometry
from .channels import HH, Channel, K, Km, Leak, Na
from .synapses import IonotropicSynapse, KineticSynapse, SigmoidRateSynapse, Synapse, TestSynapse

COMP_PARAMS = {"length": 10.0, "radius": 1.0, "axial_resistivity": 5000.0, "capacitance": 1.0}
COMP_STATES = {"v": -70.0}


def _is_all(x) -> bool:
    return isinstance(x, str) and x == "all"


def _as_list(idx, n: int) -> np.ndarray:
    if _is_all(idx):
        return np.arange(n)
    arr = np.atleast_1d(np.asarray(idx, dtype=np.int64))
    return np.where(arr < 0, arr + n, arr)


class View:
    """A selection of compartments of a module (`base`)."""

    def __init__(self, base: "Module", comps: np.ndarray, scope: str, groups: np.ndarray):
        self.base = base
        self._comps = np.asarray(comps, dtype=np.int64)
        self._scope = scope
        self._groups = np.asarray(groups, dtype=np.int64)

    # -- selection ------------------------------------------------------------------
    def _sub(self, comps, scope, groups):
        return View(self.base, comps, scope, groups)

    def cell(self, idx: Any) -> "View":
        b = self.base
        cells = np.unique(b._comp_cell[self._comps])
        sel = cells[_as_list(idx, len(cells))] if not _is_all(idx) else cells
        mask = np.isin(b._comp_cell[self._comps], sel)
        comps = self._comps[mask]
        return self._sub(comps, "cell", b._comp_cell[comps])

    def branch(self, idx: Any) -> "View":
        b = self.base
        comps = self._comps
        cells = b._comp_cell[comps]
        # Branch indices are local to the cell if one cell is in view, else global.
        if len(np.unique(cells)) == 1 and b._is_network:
            local = b._comp_branch[comps] - b._cell_branch_off[cells[0]]
            branches = np.unique(local)
            sel = branches[_as_list(idx, len(branches))] if not _is_all(idx) else branches
            mask = np.isin(local, sel)
        else:
            branches = np.unique(b._comp_branch[comps])
            sel = branches[_as_list(idx, len(branches))] if not _is_all(idx) else branches
            mask = np.isin(b._comp_branch[comps], sel)
        comps = comps[mask]
        return self._sub(comps, "branch", b._comp_branch[comps])

    def comp(self, idx: Any) -> "View":
        b = self.base
        comps = self._comps
        branches = np.unique(b._comp_branch[comps])
        if len(branches) == 1:
            local = comps - b._cumsum_ncomp[branches[0]]
            sel = _as_list(idx, b._ncomp_per_branch[branches[0]])
            comps = comps[np.isin(local, sel)]
        else:
            comps = comps[_as_list(idx, len(comps))]
        return self._sub(comps, "comp", comps)

    def loc(self, at: Any) -> "View":
        """Compartment at a location in [0, 1] of every branch in view
        (modules/base.py:704-740)."""
        b = self.base
        out = []
        for br in np.unique(b._comp_branch[self._comps]):
            ncomp = int(b._ncomp_per_branch[br])
            locs = np.linspace(0, 1, ncomp) if _is_all(at) else np.atleast_1d(np.asarray(at, dtype=float))
            edges = np.linspace(0, 1 + 1e-10, ncomp + 1)
            out.append(np.digitize(locs, edges) - 1 + b._cumsum_ncomp[br])
        comps = np.concatenate(out).astype(np.int64)
        return self._sub(comps, "comp", comps)

    def select(self, nodes: Any = None) -> "View":
        """Compartments by global index (`View.select(nodes=...)`, modules/base.py)."""
        comps = self._comps if nodes is None else np.asarray(nodes, dtype=np.int64).ravel()
        return self._sub(comps, "comp", comps)

    def add_to_group(self, group_name: str):
        """Add the compartments in view to a named group (modules/base.py)."""
        b = self.base
        prev = b.groups.get(group_name, np.zeros(0, np.int64))
        b.groups[group_name] = np.union1d(prev, self._comps)

    def __getattr__(self, name):
        # `cell.soma` -> the compartments in view that belong to group `soma`.
        if name.startswith("_") or name == "base":
            raise AttributeError(name)
        base = self.__dict__.get("base")
        groups = base.__dict__.get("groups", {}) if base is not None else {}
        if name in groups:
            comps = self._comps[np.isin(self._comps, groups[name])]
            return self._sub(comps, self._scope, np.zeros(len(comps), dtype=np.int64))
        raise AttributeError(name)

    # -- configuration ----------------------------------------------------------------
    def insert(self, channel: Channel):
        self.base._insert(channel, self._comps)
        return self

    def set(self, key: str, val):
        self.base._set(key, val, self._comps)

    def data_set(self, key: str, val, param_state: Optional[List[Dict]]):
        """Set a parameter for a simulation without modifying the module; the value is
        broadcast to `(..., num_compartments_in_view)` (modules/base.py:1351-1408)."""
        b = self.base
        if key not in b._node_data and not b._is_synapse_key(key):
            raise KeyError("Key not recognized.")
        if b._is_synapse_key(key):
            idx = np.arange(b._num_edges_of_key(key)).reshape(-1, 1)
        else:
            valid = ~np.isnan(b._node_data[key][self._comps])
            idx = self._comps[valid].reshape(-1, 1)
        if not isinstance(val, torch.Tensor):
            val = torch.as_tensor(np.asarray(val), dtype=torch.get_default_dtype())
        num = idx.shape[0]
        val = torch.broadcast_to(val, (*val.shape[:-1], num)) if val.ndim > 0 else val.expand(num)
        entry = [{"indices": idx, "key": key, "val": val}]
        return (param_state + entry) if param_state is not None else entry

    def make_trainable(self, key: str, init_val: Optional[Union[float, list]] = None, verbose: bool = True):
        """One trainable parameter per group of the view (per cell, branch or compartment
        depending on the last selection; modules/base.py:1625-1750)."""
        b = self.base
        if key not in b._node_data:
            raise KeyError(f"Key '{key}' not found in nodes")
        vals = b._node_data[key][self._comps]
        valid = ~np.isnan(vals)
        comps, groups = self._comps[valid], self._groups[valid]
        if len(comps) == 0:
            raise ValueError("No settable parameters found in the selected compartments.")
        uniq = np.unique(groups)
        per = [comps[groups == g] for g in uniq]
        width = max(len(p) for p in per)
        inds = np.stack([np.pad(p, (0, width - len(p)), constant_values=-1) for p in per])
        if init_val is None:
            new = np.array([np.mean(b._node_data[key][p]) for p in per])
        elif isinstance(init_val, float):
            new = np.full(len(per), init_val)
        else:
            new = np.asarray(init_val, dtype=float)
            assert len(new) == len(per)
        b.trainable_params.append({key: torch.as_tensor(new, dtype=torch.get_default_dtype())})
        b.indices_set_by_trainables.append(inds)
        b.num_trainable_params += len(per)
        if verbose:
            print(f"Number of newly added trainable parameters: {len(per)}. "
                  f"Total number of trainable parameters: {b.num_trainable_params}")

    def record(self, state: str = "v", verbose: bool = False):
        b = self.base
        for c in self._comps:
            b.recordings.append((state, int(c)))

    def stimulate(self, current: torch.Tensor, verbose: bool = False):
        b = self.base
        current = torch.as_tensor(current)
        nodes = b._comp_node[self._comps]
        cur = torch.broadcast_to(current, (len(nodes), current.shape[-1]))
        b._stimuli.append(cur)
        b._stim_nodes.append(nodes)

    def data_stimulate(self, current, data_stimuli=None, verbose: bool = False):
        """Stimulus for one simulation without modifying the module
        (returns `(["i", ...], [currents], [node indices])`)."""
        b = self.base
        current = torch.as_tensor(current)
        nodes = b._comp_node[self._comps]
        cur = current[..., None, :] if current.ndim >= 1 else current
        cur = torch.broadcast_to(cur, (*current.shape[:-1], len(nodes), current.shape[-1]))
        if data_stimuli is None:
            return (["i"], [cur], [nodes])
        names, curs, inds = data_stimuli
        return (names + ["i"], curs + [cur], inds + [nodes])

    @property
    def nodes(self):
        import pandas as pd

        b = self.base
        df = pd.DataFrame({k: v[self._comps] for k, v in b._node_data.items()})
        df["global_comp_index"] = self._comps
        df["global_branch_index"] = b._comp_branch[self._comps]
        df["global_cell_index"] = b._comp_cell[self._comps]
        for ch in b.channels:
            df[ch.name] = b._channel_mask[ch.name][self._comps]
        df.index = b._comp_node[self._comps]
        return df


class Module(View):
    """Base of all modules; a view of all its compartments."""

    _is_network = False

    def __init__(self, cells: Sequence[tuple], vectorize_cells: bool = True):
        self._cell_specs = [(np.asarray(p, dtype=np.int64), np.asarray(nc, dtype=np.int64)) for p, nc in cells]
        self._vectorize = vectorize_cells
        self._solver_device = "gpu"
        self._init_layout()
        self._allowed_nodes_per_level = 32
        self._node_data: Dict[str, np.ndarray] = {
            k: np.full(self._n_comp, v, dtype=float) for k, v in {**COMP_PARAMS, **COMP_STATES}.items()
        }
        geo = cylinder_geometry(self._node_data["length"], self._node_data["radius"])
        self._node_data.update({k: np.array(v, dtype=float) for k, v in geo.items()})  # own buffers
        self.groups: Dict[str, np.ndarray] = {}
        self._channel_mask: Dict[str, np.ndarray] = {}
        self.channels: List[Channel] = []
        self.synapses: List[Synapse] = []
        self._edges: List[Dict] = []
        self.recordings: List[tuple] = []
        self._stimuli: List[torch.Tensor] = []
        self._stim_nodes: List[np.ndarray] = []
        self.trainable_params: List[Dict] = []
        self.indices_set_by_trainables: List[np.ndarray] = []
        self.num_trainable_params = 0
        self._version = 0

    # Attributes derived from the cell specs (rebuilt by `_init_layout`) and caches: not
    # part of a pickled module.
    _DERIVED = ("_cell_tabs", "_cell_node_off", "_comp_node", "_n_comp", "_n_nodes", "_ncomp_per_branch",
                "_cumsum_ncomp", "_cell_branch_off", "_comp_branch", "_comp_cell", "_tables", "_indexer",
                "_fused_cache", "base", "_comps", "_groups")

    def __getstate__(self):
        """A compact pickled form (request preparation in worker processes pickles every
        served module): the cell specs and the user-set data, the per-compartment
        columns packed into one array; the derived layout is rebuilt on unpickling."""
        d = {k: v for k, v in self.__dict__.items() if k not in self._DERIVED}
        cols = d["_node_data"]
        keys = list(cols)
        if keys and all(isinstance(cols[k], np.ndarray) and cols[k].dtype == np.float64
                        and cols[k].shape == (self._n_comp,) for k in keys):
            d["_node_data"] = ("__packed__", keys, np.stack([cols[k] for k in keys]))
        d["__scope_groups__"] = None if self._groups is None else (self._groups if self._groups.any() else None)
        return d

    def __setstate__(self, d):
        cols = d.pop("_node_data")
        groups = d.pop("__scope_groups__", None)
        self.__dict__.update(d)
        self._init_layout(reset_view=False)
        if isinstance(cols, tuple) and len(cols) == 3 and cols[0] == "__packed__":
            _, keys, arr = cols
            cols = {k: arr[i].copy() for i, k in enumerate(keys)}
        self._node_data = cols
        self.base = self
        self._comps = np.arange(self._n_comp)
        self._groups = np.zeros(self._n_comp, dtype=np.int64) if groups is None else groups

    def _init_layout(self, reset_view: bool = True):
        """Compartment / node layout of the cell specs (tables, offsets, index maps)."""
        tabs = [sch.cell_tables(p, nc) for p, nc in self._cell_specs]
        node_off = np.concatenate([[0], np.cumsum([t["n_nodes"] for t in tabs])]).astype(np.int64)
        self._cell_node_off = node_off
        self._comp_node = np.concatenate([t["internal"] + node_off[i] for i, t in enumerate(tabs)])
        self._n_comp = len(self._comp_node)
        self._n_nodes = int(node_off[-1])
        self._ncomp_per_branch = np.concatenate([nc for _, nc in self._cell_specs])
        self._cumsum_ncomp = np.concatenate([[0], np.cumsum(self._ncomp_per_branch)]).astype(np.int64)
        nbranch = np.array([len(p) for p, _ in self._cell_specs])
        self._cell_branch_off = np.concatenate([[0], np.cumsum(nbranch)]).astype(np.int64)
        self._comp_branch = np.repeat(np.arange(len(self._ncomp_per_branch)), self._ncomp_per_branch)
        self._comp_cell = np.repeat(np.arange(len(self._cell_specs)), [int(nc.sum()) for _, nc in self._cell_specs])
        self._cell_tabs = tabs
        self._tables = None
        self._indexer = None
        if reset_view:
            comps = np.arange(self._n_comp)
            View.__init__(self, self, comps, "module", np.zeros(self._n_comp, dtype=np.int64))

    # -- structure ----------------------------------------------------------------------
    @property
    def initialized(self) -> bool:
        return True

    def _ensure_solver(self):
        if self._indexer is None:
            inds = [sch.dhs_indexer(t, self._allowed_nodes_per_level) for t in self._cell_tabs]
            if len(self._cell_tabs) == 1:
                self._tables, self._indexer = self._cell_tabs[0], inds[0]
            else:
                self._tables, self._indexer = sch.network_indexer(self._cell_tabs, inds, self._vectorize)
        return self._tables, self._indexer

    def _init_solver_jaxley_dhs_solve(self, allowed_nodes_per_level: Optional[int] = None, root: int = 0):
        """Rebuild the DHS solve order (modules/base.py:1142-1266)."""
        if allowed_nodes_per_level is not None:
            self._allowed_nodes_per_level = allowed_nodes_per_level
            self._solver_device = "cpu" if allowed_nodes_per_level == 1 else "gpu"
        self._indexer = None
        self._version += 1
        self._ensure_solver()

    @property
    def _dhs_solve_indexer(self):
        return self._ensure_solver()[1]

    @property
    def _comp_edges(self):
        return self._ensure_solver()[0]

    # -- mechanisms ---------------------------------------------------------------------
    def _insert(self, channel: Channel, comps: np.ndarray):
        if not isinstance(channel, (HH, Leak, Na, K, Km)):
            raise NotImplementedError(f"channel {type(channel).__name__} is not on the jaxhip path")
        name = channel.name
        if name not in self._channel_mask:
            self.channels.append(channel)
            self._channel_mask[name] = np.zeros(self._n_comp, dtype=bool)
            for k in list(channel.channel_params) + list(channel.channel_states):
                if k not in self._node_data:
                    self._node_data[k] = np.full(self._n_comp, np.nan)
        self._channel_mask[name][comps] = True
        for k, v in {**channel.channel_params, **channel.channel_states}.items():
            col = self._node_data[k]
            fresh = np.isnan(col[comps])
            col[comps[fresh]] = v
        self._version += 1

    def _is_synapse_key(self, key: str) -> bool:
        return any(key in e["params"] or key in e["states"] for e in self._edges)

    def _num_edges_of_key(self, key: str) -> int:
        for e in self._edges:
            if key in e["params"] or key in e["states"]:
                return len(e["pre"])
        return 0

    def _set(self, key: str, val, comps: np.ndarray):
        if isinstance(val, torch.Tensor):
            val = val.detach().cpu().numpy()
        if key in self._node_data:
            col = self._node_data[key]
            valid = ~np.isnan(col[comps])
            col[comps[valid]] = np.broadcast_to(np.asarray(val, dtype=float), comps.shape)[valid]
            if key in ("radius", "length"):
                # Set compartments become cylinders (modules/base.py:1316-1344).
                rows = comps[valid]
                geo = cylinder_geometry(self._node_data["length"][rows], self._node_data["radius"][rows])
                for k, x in geo.items():
                    self._node_data[k][rows] = x
        elif self._is_synapse_key(key):
            for e in self._edges:
                for d in (e["params"], e["states"]):
                    if key in d:
                        d[key][:] = np.broadcast_to(np.asarray(val, dtype=float), d[key].shape)
        else:
            raise KeyError(f"Key '{key}' not found in nodes or edges")
        self._version += 1

    def add_synapses(self, pre_nodes, post_nodes, synapse_type: Synapse):
        """Append synapses between node indices (`Network._append_multiple_synapses`,
        modules/network.py:736-790): one row per (pre, post) pair, parameters and states
        at the synapse type's defaults."""
        if not isinstance(synapse_type, (KineticSynapse, SigmoidRateSynapse, IonotropicSynapse, TestSynapse)):
            raise NotImplementedError(f"synapse {type(synapse_type).__name__} is not on the jaxhip path")
        pre = np.asarray(pre_nodes, dtype=np.int64).ravel()
        post = np.asarray(post_nodes, dtype=np.int64).ravel()
        assert len(pre) == len(post)
        for e in self._edges:
            if e["type"].name == synapse_type.name:
                e["pre"] = np.concatenate([e["pre"], pre])
                e["post"] = np.concatenate([e["post"], post])
                for d, defaults in ((e["params"], synapse_type.synapse_params), (e["states"], synapse_type.synapse_states)):
                    for k, v in defaults.items():
                        d[k] = np.concatenate([d[k], np.full(len(pre), v)])
                break
        else:
            self.synapses.append(synapse_type)
            self._edges.append({
                "type": synapse_type,
                "pre": pre,
                "post": post,
                "params": {k: np.full(len(pre), v) for k, v in synapse_type.synapse_params.items()},
                "states": {k: np.full(len(pre), v) for k, v in synapse_type.synapse_states.items()},
            })
        self._version += 1

    # -- state and parameters -----------------------------------------------------------
    def init_states(self, delta_t: float = 0.025):
        """Channels in their steady state at the current voltages (modules/base.py:2419-2474;
        HH: x = alpha / (alpha + beta), channels/hh.py:69-85)."""
        for ch in self.channels:
            if ch.kernel_id == "HH":
                m = self._channel_mask[ch.name]
                v = self._node_data["v"][m]
                am = 0.1 * _xoe(-(v + 40), 10)
                bm = 4.0 * np.exp(np.minimum(-(v + 65) / 18, 20))
                ah = 0.07 * np.exp(np.minimum(-(v + 65) / 20, 20))
                bh = 1.0 / (np.exp(np.minimum(-(v + 35) / 10, 20)) + 1)
                an = 0.01 * _xoe(-(v + 55), 10)
                bn = 0.125 * np.exp(np.minimum(-(v + 65) / 80, 20))
                self._node_data["HH_m"][m] = am / (am + bm)
                self._node_data["HH_h"][m] = ah / (ah + bh)
                self._node_data["HH_n"][m] = an / (an + bn)
            elif ch.kernel_id in ("PNa", "PK", "PKm"):
                # Fixed points of the gates (pospischil.py:117-125, 184-189, 235-238;
                # Km takes p_inf / (p_inf + tau_p), as the reference does).
                m = self._channel_mask[ch.name]
                v = self._node_data["v"][m]
                d = self._node_data
                if ch.kernel_id == "PKm":
                    vp = v + 35.0
                    pinf = 1.0 / (1.0 + np.exp(np.minimum(-0.1 * vp, 20)))
                    tau = d["Km_taumax"][m] / (3.3 * np.exp(np.minimum(0.05 * vp, 20)) + np.exp(np.minimum(-0.05 * vp, 20)))
                    d["Km_p"][m] = pinf / (pinf + tau)
                    continue
                u = v - d["vt"][m]
                if ch.kernel_id == "PNa":
                    am = 0.32 * _xoe(-0.25 * (u - 13.0), 1.0) / 0.25
                    bm = 0.28 * _xoe(0.2 * (u - 40.0), 1.0) / 0.2
                    ah = 0.128 * np.exp(np.minimum(-(u - 17.0) / 18.0, 20))
                    bh = 4.0 / (np.exp(np.minimum(-(u - 40.0) / 5.0, 20)) + 1.0)
                    d["Na_m"][m] = am / (am + bm)
                    d["Na_h"][m] = ah / (ah + bh)
                else:
                    an = 0.032 * _xoe(-0.2 * (u - 15.0), 1.0) / 0.2
                    bn = 0.5 * np.exp(np.minimum(-(u - 10.0) / 40.0, 20))
                    d["K_n"][m] = an / (an + bn)
        self._version += 1

    def get_parameters(self) -> List[Dict[str, torch.Tensor]]:
        return self.trainable_params

    def delete_recordings(self):
        self.recordings = []

    def delete_stimuli(self):
        self._stimuli = []
        self._stim_nodes = []

    @property
    def externals(self) -> Dict[str, torch.Tensor]:
        if not self._stimuli:
            return {}
        return {"i": torch.cat([torch.as_tensor(s) for s in self._stimuli], dim=-2)}

    @property
    def external_inds(self) -> Dict[str, np.ndarray]:
        if not self._stim_nodes:
            return {}
        return {"i": np.concatenate(self._stim_nodes)}

    @property
    def edges(self):
        import pandas as pd

        rows = []
        for e in self._edges:
            df = pd.DataFrame({"pre_index": e["pre"], "post_index": e["post"], **e["params"], **e["states"]})
            df["type"] = e["type"].name
            rows.append(df)
        return pd.concat(rows, ignore_index=True) if rows else pd.DataFrame()

    def __getattr__(self, name):
        # `net.KineticSynapse` -> a setter view over all synapses of that type.
        if name.startswith("_"):
            raise AttributeError(name)
        for e in self.__dict__.get("_edges", []):
            if e["type"].name == name:
                return _EdgeView(self, e)
        return View.__getattr__(self, name)

    @property
    def comb_parents(self) -> np.ndarray:
        """Parent branch of every branch, cells offset by their first branch
        (modules/base.py `comb_parents`)."""
        out = []
        for c, (p, _) in enumerate(self._cell_specs):
            off = self._cell_branch_off[c]
            out.append(np.where(p < 0, -1, p + off))
        return np.concatenate(out)

    def _set_morphology(self, **cols):
        """Per-compartment geometry as read from a morphology (no cylinder rule)."""
        for k, v in cols.items():
            self._node_data[k] = np.asarray(v, dtype=float).copy()
        self._version += 1


class _EdgeView:
    """The synapses of one type (`net.IonotropicSynapse`), or a subset of them by their
    index within the type (`.edge([0, 2])`, as `View.edge`, modules/base.py)."""

    def __init__(self, base, edges, idx=None):
        self.base, self.e = base, edges
        self._per_edge = idx is not None
        self.idx = np.arange(len(edges["pre"])) if idx is None else np.asarray(idx, dtype=np.int64)

    def edge(self, idx) -> "_EdgeView":
        sel = self.idx if isinstance(idx, str) and idx == "all" else self.idx[np.atleast_1d(idx)]
        return _EdgeView(self.base, self.e, sel)

    def set(self, key, val):
        for d in (self.e["params"], self.e["states"]):
            if key in d:
                d[key][self.idx] = np.broadcast_to(np.asarray(val, dtype=float), self.idx.shape)
                self.base._version += 1
                return
        raise KeyError(key)

    def record(self, state: str, verbose: bool = False):
        """Record a synapse state or current (`i_<synapse>`) of the synapses in view, by
        their index within the type (modules/base.py:2544-2563)."""
        stype = self.e["type"]
        known = {f"{stype.name}_{stype.state_name}"} if stype.state_name else set()
        if state not in known | {f"i_{stype.name}"}:
            raise KeyError(f"{state} is not a recognized state in this module.")
        for i in self.idx:
            self.base.recordings.append((state, int(i)))

    def make_trainable(self, key: str, init_val: Optional[Union[float, list]] = None, verbose: bool = True):
        """One trainable parameter for all synapses of the type view, or one per synapse of
        an `.edge(...)` view (modules/base.py:1625-1750); the initial value is the mean
        of the current values of the synapses it controls."""
        col = self.e["params"].get(key, self.e["states"].get(key))
        if col is None:
            raise KeyError(f"Key '{key}' not found in edges")
        inds = self.idx[:, None] if self._per_edge else self.idx[None, :]
        if init_val is None:
            new = np.array([np.mean(col[r]) for r in inds])
        elif isinstance(init_val, float):
            new = np.full(len(inds), init_val)
        else:
            new = np.asarray(init_val, dtype=float)
            assert len(new) == len(inds)
        b = self.base
        b.trainable_params.append({key: torch.as_tensor(new, dtype=torch.get_default_dtype())})
        b.indices_set_by_trainables.append(inds)
        b.num_trainable_params += len(inds)
        if verbose:
            print(f"Number of newly added trainable parameters: {len(inds)}. "
                  f"Total number of trainable parameters: {b.num_trainable_params}")


def _xoe(x, y):
    z = np.minimum(x / y, 20.0)
    small = np.abs(z) < 1e-4
    zs = np.where(small, 1.0, z)
    return np.where(small, y * (1 - z / 2 + z * z / 12), x / np.expm1(zs))


class Compartment(Module):
    """A single compartment (modules/compartment.py)."""

    def __init__(self):
        super().__init__([([-1], [1])])
        self._scope = "comp"


class Branch(Module):
    """A branch of `ncomp` compartments (modules/branch.py)."""

    def __init__(self, compartments: Optional[Union[Compartment, List[Compartment]]] = None,
                 ncomp: Optional[int] = None):
        if isinstance(compartments, list):
            n = len(compartments)
            src = compartments
        else:
            n = 1 if ncomp is None else int(ncomp)
            src = [compartments] * n if compartments is not None else []
        super().__init__([([-1], [n])])
        self._inherit_from([(c, np.array([i])) for i, c in enumerate(src) if c is not None])


class Cell(Module):
    """A cell of branches (modules/cell.py); `parents[0] == -1`."""

    def __init__(self, branches: Optional[Union[Branch, List[Branch]]] = None,
                 parents: Optional[List[int]] = None):
        parents = [-1] if parents is None else list(parents)
        if branches is None:
            blist = [Branch()] * len(parents)
        elif isinstance(branches, Branch):
            blist = [branches] * len(parents)
        else:
            blist = list(branches)
        assert len(blist) == len(parents)
        super().__init__([(parents, [b._n_comp for b in blist])])
        off = np.concatenate([[0], np.cumsum([b._n_comp for b in blist])])
        self._inherit_from([(b, np.arange(off[i], off[i + 1])) for i, b in enumerate(blist)])


class Network(Module):
    """A network of cells connected by synapses (modules/network.py)."""

    _is_network = True

    def __init__(self, cells: List[Cell], vectorize_cells: Optional[bool] = None):
        specs = []
        for c in cells:
            specs += c._cell_specs
        super().__init__(specs, vectorize_cells=True if vectorize_cells is None else vectorize_cells)
        off = np.concatenate([[0], np.cumsum([c._n_comp for c in cells])])
        self._inherit_from([(c, np.arange(off[i], off[i + 1])) for i, c in enumerate(cells)])


def _inherit_from(self, parts):
    """Copy parameters, states and channels of constituent modules."""
    for mod, comps in parts:
        for ch in mod.channels:
            self._insert(ch, comps[mod._channel_mask[ch.name]])
        for k, col in mod._node_data.items():
            if k not in self._node_data:
                self._node_data[k] = np.full(self._n_comp, np.nan)
            self._node_data[k][comps] = col
        for name, members in mod.groups.items():
            prev = self.groups.get(name, np.zeros(0, np.int64))
            self.groups[name] = np.union1d(prev, comps[members])
        if getattr(mod, "xyzr", None) is not None:
            self.xyzr = list(getattr(self, "xyzr", None) or []) + list(mod.xyzr)


Module._inherit_from = _inherit_from


def fully_connect(pre_cell_view: View, post_cell_view: View, synapse_type: Synapse,
                  random_post_comp: bool = False):
    """One synapse from branch 0, compartment 0 of every presynaptic cell to every
    postsynaptic cell (connect.py:80-150): pre-major order, onto branch 0, compartment 0
    of the postsynaptic cell, or with `random_post_comp` onto compartments drawn from
    numpy's global random state exactly as the reference draws them (pandas'
    groupby-sample: for each postsynaptic cell in ascending order, `num_pre` draws with
    replacement among the cell's compartments in the view)."""
    base = pre_cell_view.base
    assert post_cell_view.base is base and isinstance(base, Network), "pre and post must be in the same network"
    pre_cells = np.unique(base._comp_cell[pre_cell_view._comps])
    post_cells = np.unique(base._comp_cell[post_cell_view._comps])
    num_pre, num_post = len(pre_cells), len(post_cells)
    pre_comps = np.array([base.cell(int(c)).branch(0).comp(0)._comps[0] for c in pre_cells], dtype=np.int64)
    pre_comps = np.repeat(pre_comps, num_post)
    if random_post_comp:
        drawn = []
        for c in post_cells:
            comps = post_cell_view._comps[base._comp_cell[post_cell_view._comps] == c]
            drawn.append(comps[np.random.choice(len(comps), num_pre, replace=True)])
        post_comps = np.reshape(np.concatenate(drawn), (num_pre, num_post)).T.flatten()
    else:
        first = np.array([base.cell(int(c)).branch(0).comp(0)._comps[0] for c in post_cells], dtype=np.int64)
        post_comps = first[np.tile(np.arange(num_post), num_pre)]
    base.add_synapses(base._comp_node[pre_comps], base._comp_node[post_comps], synapse_type)


def connect(pre: View, post: View, synapse_type: Synapse):
    """Connect compartments pairwise with synapses (connect.py:22-75)."""
    base = pre.base
    assert post.base is base and isinstance(base, Network), "pre and post must be in the same network"
    pre_nodes = base._comp_node[pre._comps]
    post_nodes = base._comp_node[post._comps]
    base.add_synapses(pre_nodes, post_nodes, synapse_type)
