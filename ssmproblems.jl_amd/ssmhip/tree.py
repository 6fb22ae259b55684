"""Sparse path storage of a device bootstrap filter (GF/ext/CUDAExt.jl:146-298).

``ParallelParticleTree(state, capacity)`` roots a tree at the filter's current particles;
``insert()`` appends the particles after the filter's last step (ParallelAncestorCallback,
CUDAExt.jl:273-298); ``ancestry()`` returns every leaf's path x_0 .. x_depth as an array
(N, depth + 1, Dx), the trajectories ``get_ancestry`` builds (GF/src/containers.jl:592-611).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .filters import PostInit, PostUpdate


class ParallelParticleTree:
    def __init__(self, state, capacity=None):
        self.ctx, self.state = state.ctx, state
        self.N, self.Dx = state.N, state.Dx
        cap = int(capacity) if capacity is not None else 4 * self.N
        p = C.c_void_p()
        self.ctx.check(self.ctx.lib.ssm_bf_tree_create(state.ptr, cap, C.byref(p)))
        self.ptr = p

    def insert(self):
        self.ctx.check(self.ctx.lib.ssm_bf_tree_insert(self.ptr))

    def info(self):
        cap, free, depth = C.c_int64(), C.c_int64(), C.c_int32()
        self.ctx.check(self.ctx.lib.ssm_bf_tree_info(self.ptr, C.byref(cap), C.byref(free),
                                                     C.byref(depth)))
        return {"capacity": cap.value, "free": free.value, "depth": depth.value,
                "used": cap.value - free.value}

    def ancestry(self):
        depth = self.info()["depth"]
        out = np.zeros((self.N, depth + 1, self.Dx))
        self.ctx.check(self.ctx.lib.ssm_bf_tree_ancestry(self.ptr, out.ctypes.data_as(_lib._dp)))
        return out

    def path(self, i):
        """Path of particle i, (depth + 1, Dx) (get_ancestry(tree, i, T), CUDAExt.jl:248-262)."""
        depth = self.info()["depth"]
        out = np.zeros((depth + 1, self.Dx))
        self.ctx.check(self.ctx.lib.ssm_bf_tree_path(self.ptr, int(i),
                                                     out.ctypes.data_as(_lib._dp)))
        return out

    def export(self):
        """(parents, offspring, leaves): the node pool's parent links (-1 at roots) and
        offspring counts, and each particle's leaf slot, all 0-based int32."""
        cap = self.info()["capacity"]
        par = np.zeros(cap, dtype=np.int32)
        off = np.zeros(cap, dtype=np.int32)
        leaves = np.zeros(self.N, dtype=np.int32)
        self.ctx.check(self.ctx.lib.ssm_bf_tree_export(
            self.ptr, par.ctypes.data_as(_lib._ip), off.ctypes.data_as(_lib._ip),
            leaves.ctypes.data_as(_lib._ip)))
        return par, off, leaves

    def close(self):
        if getattr(self, "ptr", None):
            self.ctx.lib.ssm_bf_tree_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParallelAncestorCallback:
    """callback for ``ssmhip.filter``: roots the tree at PostInit and inserts each
    generation at PostUpdate (CUDAExt.jl:273-298)."""

    def __init__(self, capacity=None):
        self.capacity = capacity
        self.tree = None

    def __call__(self, model, algo, *args):
        event = args[-1]
        state = args[-3] if event == PostUpdate else args[0]
        if event == PostInit:
            self.tree = ParallelParticleTree(state, self.capacity)
        elif event == PostUpdate:
            self.tree.insert()
