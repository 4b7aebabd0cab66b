"""Tree (Hines / DHS) solve on the GPU: the drop-in for jaxley-pytorch's
`jaxley_pytorch.kernels.tree_solve` (kernels/tree_solve.py:378-635).

Same surface: `make_tree_solve_kernel(parent, device, mode)` returns a kernel object with
`.solve(d, l, u, s)` (no autograd) and `.fused(d, l, u, s)` (differentiable; the backward
is the transposed solve, tree_solve.py:575-587), or `None` if the node order does not
satisfy the kernel's requirements. `mode` (default: env var `JAXHIP_TREE_KERNEL`, also
read under the reference's name `JAXLEY_PYTORCH_TREE_KERNEL`) is `registers`, `levels`
or `auto`; `auto` asks the autotuner (measured timings, cached per structure).

Both variants are hand-written HIP kernels (csrc/jxhip_tree_solve.hip):
* `registers` (`TreeSolveKernel`): one wavefront per (block of trees, row), nodes held in
  VGPRs, children's terms exchanged through LDS, recursive-doubling back-substitution;
  blocks up to 2048 nodes.
* `levels` (`LevelTreeSolveKernel`): one workgroup per (block, row), diagonal and rhs in
  LDS, one barrier per height level; any tree size.
"""
from __future__ import annotations

import os
import weakref
from typing import Dict, Optional

import numpy as np
import torch
import torch.autograd.forward_ad as fwAD

from .. import _lib
from ..schedule import forest_schedule

# Largest block of the register variant (64 lanes x 32 nodes per lane).
MAX_BLOCK = 2048
LEVEL_BLOCK_NODES = 1024
_NPLS = (1, 2, 4, 8, 16, 32)


def build_forest_schedule(parent: np.ndarray, max_block: int = MAX_BLOCK) -> Optional[Dict]:
    """Schedule of the register variant (`None` if a tree exceeds `max_block` nodes or the
    order is unsupported; tree_solve.py:45-138)."""
    sched = forest_schedule(parent, max_block)
    if sched is None or sched["max_block"] > max_block:
        return None
    npl = next(p for p in _NPLS if 64 * p >= sched["max_block"])
    sched["npl"] = npl
    return sched


def build_level_schedule(parent: np.ndarray, block_nodes: int = LEVEL_BLOCK_NODES) -> Optional[Dict]:
    """Schedule of the level variant (blocks of about `block_nodes` nodes; a single tree
    may be larger; tree_solve.py:142-224)."""
    return forest_schedule(parent, block_nodes)


class _KernelBase:
    variant = ""

    def __init__(self, sched: Dict, device: torch.device):
        self.device = torch.device(device)
        self.num_nodes = int(sched["num_nodes"])
        self.num_blocks = int(sched["num_blocks"])
        self.max_block = int(sched["max_block"])
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.int32, device=self.device)
        self.block_info = t(sched["block_info"])
        self.level_ptr = t(sched["level_ptr"])
        self.order = t(sched["order"])
        self.parent = t(sched["parent"])
        self.child_rng = t(sched["child_rng"])
        self.child_idx = t(sched["child_idx"])
        self.parent_glob = t(sched["parent_glob"])
        self.sched = sched
        # Handle of this kernel for the `jaxhip::tree_solve` operator (an operator's
        # arguments are tensors and numbers).
        global _NEXT_HANDLE
        self.handle = _NEXT_HANDLE
        _NEXT_HANDLE += 1
        _KERNELS[self.handle] = self

    def _prep(self, *xs):
        ref = xs[0]
        _lib.require_gpu(ref, "tree solve")
        n = ref.shape[-1]
        if n != self.num_nodes:
            raise ValueError(f"expected {self.num_nodes} nodes, got {n}")
        batch = torch.broadcast_shapes(*[x.shape[:-1] for x in xs])
        flat = [torch.broadcast_to(x, (*batch, n)).reshape(-1, n).contiguous() for x in xs]
        dtype = ref.dtype
        return batch, n, [f.to(dtype) for f in flat]

    def solve(self, diags, lowers, uppers, solves):
        """Solve without autograd; inputs `(..., num_nodes)` (broadcast)."""
        batch, n, (d, l, u, s) = self._prep(diags, lowers, uppers, solves)
        x = torch.empty_like(s)
        self._launch(d, l, u, s, x)
        return x.reshape(*batch, n)

    def fused(self, diags, lowers, uppers, solves):
        """Differentiable solve (backward: transposed solve + elementwise gradients).
        Under `torch.compile` it is the `jaxhip::tree_solve` operator (fake tensor rule,
        registered backward); eagerly it is the autograd function with the same
        backward and the forward rule of `_DHSSolveForwardAD` (solver_voltage.py:190-224),
        which skips the operator dispatch (tens of microseconds per call on the per-step
        path)."""
        if torch.compiler.is_compiling():
            return tree_solve_op(diags, lowers, uppers, solves, self.parent_glob, self.handle)
        return _TreeSolve.apply(diags, lowers, uppers, solves, self)

    def grads(self, lam, x, need_d=True, need_l=True, need_u=True):
        rows = lam.reshape(-1, self.num_nodes).contiguous()
        xx = x.reshape(-1, self.num_nodes).contiguous()
        g_d = torch.empty_like(rows) if need_d else None
        g_l = torch.empty_like(rows) if need_l else None
        g_u = torch.empty_like(rows) if need_u else None
        f = _lib.fn("jxhip_tree_solve_grads", rows.dtype)
        _lib.check(
            f(_lib.ptr(rows), _lib.ptr(xx), _lib.ptr(self.parent_glob), _lib.ptr(g_d),
              _lib.ptr(g_l), _lib.ptr(g_u), rows.shape[0], self.num_nodes,
              _lib.stream_ptr(rows.device)),
            "tree_solve_grads",
        )
        shp = lam.shape
        return tuple(None if g is None else g.reshape(shp) for g in (g_d, g_l, g_u))


class TreeSolveKernel(_KernelBase):
    """Register (wave-resident) variant."""

    variant = "registers"

    def __init__(self, sched: Dict, device: torch.device):
        super().__init__(sched, device)
        self.npl = int(sched["npl"])

    def _launch(self, d, l, u, s, x):
        f = _lib.fn("jxhip_tree_solve_regs", d.dtype)
        _lib.check(
            f(d.data_ptr(), l.data_ptr(), u.data_ptr(), s.data_ptr(), x.data_ptr(), d.shape[0],
              self.num_nodes, _lib.ptr(self.block_info), self.num_blocks, self.npl,
              _lib.ptr(self.level_ptr), _lib.ptr(self.order), _lib.ptr(self.parent),
              _lib.ptr(self.child_rng), _lib.ptr(self.child_idx), _lib.stream_ptr(d.device)),
            "tree_solve_regs",
        )


class LevelTreeSolveKernel(_KernelBase):
    """Level-sweeping variant."""

    variant = "levels"

    def _launch(self, d, l, u, s, x):
        lds = 2 * d.element_size() * self.max_block
        work = torch.empty_like(d) if lds > 64 * 1024 else None
        f = _lib.fn("jxhip_tree_solve_levels", d.dtype)
        _lib.check(
            f(d.data_ptr(), l.data_ptr(), u.data_ptr(), s.data_ptr(), x.data_ptr(), _lib.ptr(work),
              d.shape[0], self.num_nodes, _lib.ptr(self.block_info), self.num_blocks,
              self.max_block, _lib.ptr(self.level_ptr), _lib.ptr(self.order),
              _lib.ptr(self.parent), _lib.ptr(self.child_rng), _lib.ptr(self.child_idx),
              _lib.stream_ptr(d.device)),
            "tree_solve_levels",
        )


class _TreeSolve(torch.autograd.Function):
    """x = A^{-1} b; backward lam = A^{-T} g (l and u swapped), g_d = -lam x,
    g_u[c] = -lam[p] x[c], g_l[c] = -lam[c] x[p], g_b = lam (solver_voltage.py:143-187).
    The backward re-enters `fused`, so it is differentiable itself."""

    @staticmethod
    def forward(ctx, diags, lowers, uppers, solves, kernel):
        x = kernel.solve(diags, lowers, uppers, solves)
        ctx.save_for_backward(diags, lowers, uppers, x)
        ctx.save_for_forward(diags, lowers, uppers, x)
        ctx.kernel = kernel
        ctx.shapes = [t.shape for t in (diags, lowers, uppers, solves)]
        return x

    @staticmethod
    def jvp(ctx, t_d, t_l, t_u, t_s, _):
        """Forward mode (`_DHSSolveForwardAD`, solver_voltage.py:190-224): the tangent is
        a solve with the same matrix, dx = A^{-1} (db - dA x), where row c holds lowers[c]
        in its parent's column and the parent's row holds uppers[c] in column c."""
        diags, lowers, uppers, x = ctx.saved_tensors
        kernel = ctx.kernel
        par = kernel.parent_glob.long()
        has = par >= 0
        ps = torch.where(has, par, torch.zeros_like(par))
        rhs = t_s if t_s is not None else torch.zeros_like(x)
        if t_d is not None:
            rhs = rhs - t_d * x
        if t_l is not None:
            rhs = rhs - torch.where(has, t_l * x[..., ps], torch.zeros_like(x))
        if t_u is not None:
            ch = torch.nonzero(has).flatten()
            rhs = torch.broadcast_to(rhs, torch.broadcast_shapes(rhs.shape, t_u.shape, x.shape))
            rhs = rhs.index_add(-1, par[ch], -(t_u * x)[..., ch])
        return kernel.solve(diags, lowers, uppers, rhs)

    @staticmethod
    def backward(ctx, g):
        diags, lowers, uppers, x = ctx.saved_tensors
        kernel = ctx.kernel
        lam = kernel.fused(diags, uppers, lowers, g)
        if torch.is_grad_enabled():
            # Double backward: express the gradients with differentiable torch ops.
            par = kernel.parent_glob.long()
            has = par >= 0
            ps = torch.where(has, par, torch.zeros_like(par))
            g_d = -lam * x
            g_u = torch.where(has, -lam[..., ps] * x, torch.zeros_like(x))
            g_l = torch.where(has, -lam * x[..., ps], torch.zeros_like(x))
        else:
            g_d, g_l, g_u = kernel.grads(lam, x, *ctx.needs_input_grad[:3])
        out = []
        for gi, shp, need in zip((g_d, g_l, g_u, lam), ctx.shapes, ctx.needs_input_grad[:4]):
            if not need or gi is None:
                out.append(None)
                continue
            # Sum over broadcast dimensions.
            while gi.dim() > len(shp):
                gi = gi.sum(0)
            for ax, (a, b) in enumerate(zip(gi.shape, shp)):
                if b == 1 and a != 1:
                    gi = gi.sum(ax, keepdim=True)
            out.append(gi)
        return (*out, None)


_KERNELS: "weakref.WeakValueDictionary[int, _KernelBase]" = weakref.WeakValueDictionary()
_NEXT_HANDLE = 0


@torch.library.custom_op("jaxhip::tree_solve", mutates_args=())
def tree_solve_op(diags: torch.Tensor, lowers: torch.Tensor, uppers: torch.Tensor,
                  solves: torch.Tensor, parent: torch.Tensor, handle: int) -> torch.Tensor:
    """x = A^{-1} b on the GPU by the kernel registered under `handle` (inputs broadcast
    to `(..., num_nodes)`); `parent` (int32, -1 at roots) is the kernel's node parent,
    used by the backward."""
    return _KERNELS[handle].solve(diags, lowers, uppers, solves)


@tree_solve_op.register_fake
def _tree_solve_fake(diags, lowers, uppers, solves, parent, handle):
    shape = torch.broadcast_shapes(diags.shape, lowers.shape, uppers.shape, solves.shape)
    return diags.new_empty(shape)


def _tree_solve_setup(ctx, inputs, output):
    diags, lowers, uppers, solves, parent, handle = inputs
    ctx.save_for_backward(diags, lowers, uppers, output, parent)
    ctx.kernel = _KERNELS[handle]           # keeps the kernel alive with the graph
    ctx.handle = handle
    ctx.shapes = [t.shape for t in (diags, lowers, uppers, solves)]


def _sum_to(g, shape):
    while g.dim() > len(shape):
        g = g.sum(0)
    for ax, (a, b) in enumerate(zip(g.shape, shape)):
        if b == 1 and a != 1:
            g = g.sum(ax, keepdim=True)
    return g


def _tree_solve_backward(ctx, g):
    """lam = A^{-T} g (the same operator with lower and upper swapped), g_d = -lam x,
    g_u[c] = -lam[p] x[c], g_l[c] = -lam[c] x[p], g_b = lam (solver_voltage.py:143-187);
    tensor operations only, so the backward is differentiable and traceable."""
    diags, lowers, uppers, x, parent = ctx.saved_tensors
    lam = tree_solve_op(diags, uppers, lowers, g, parent, ctx.handle)
    par = parent.long()
    has = par >= 0
    ps = torch.where(has, par, torch.zeros_like(par))
    grads = (-lam * x, torch.where(has, -lam * x[..., ps], torch.zeros_like(x)),
             torch.where(has, -lam[..., ps] * x, torch.zeros_like(x)), lam)
    return (*(_sum_to(gi, shp) for gi, shp in zip(grads, ctx.shapes)), None, None)


tree_solve_op.register_autograd(_tree_solve_backward, setup_context=_tree_solve_setup)


def _env_mode() -> str:
    return os.environ.get("JAXHIP_TREE_KERNEL", os.environ.get("JAXLEY_PYTORCH_TREE_KERNEL", "auto"))


def make_tree_solve_kernel(parent: np.ndarray, device: torch.device, mode: Optional[str] = None):
    """The GPU solver for the forest given by `parent` (-1 = root), or `None` if the
    order is not supported (tree_solve.py:485-513). `auto` asks the autotuner."""
    mode = mode or _env_mode()
    if mode not in ("auto", "registers", "levels"):
        raise ValueError(f"unknown tree kernel mode {mode!r}")
    forest = build_forest_schedule(parent) if mode in ("auto", "registers") else None
    levels = build_level_schedule(parent) if mode in ("auto", "levels") else None
    if mode == "registers":
        return TreeSolveKernel(forest, device) if forest is not None else None
    if mode == "levels":
        return LevelTreeSolveKernel(levels, device) if levels is not None else None
    if forest is None:
        return LevelTreeSolveKernel(levels, device) if levels is not None else None
    if levels is None:
        return TreeSolveKernel(forest, device)
    from ..autotune import choose_tree_kernel

    return choose_tree_kernel(parent, device, forest, levels)


def fused_tree_solve(kernel, diags, lowers, uppers, solves):
    """Differentiable tree solve (inputs of the same shape)."""
    return kernel.fused(diags, lowers, uppers, solves)
