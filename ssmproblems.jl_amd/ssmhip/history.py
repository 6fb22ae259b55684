"""Dense particle history of a device bootstrap filter (GF/src/containers.jl:328-418).

``DenseParticleContainer(state)`` stores the filter's current particles as time 0;
``push()`` appends the generation after the filter's last step: states, normalised FP64
log-weights and ancestors (push!, containers.jl:364-374).  ``get_ancestry(i)`` walks the
ancestors back from the last generation (containers.jl:406-418); ``particle(t, i)`` is
``Particle(c, t, i)`` (containers.jl:396-404) with 0-based i and ancestors.
``DenseAncestorCallback`` fills one during ``ssmhip.filter`` (GF/src/callbacks.jl:53-80).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .filters import PostInit, PostUpdate


class DenseParticleContainer:
    def __init__(self, state, capacity=64):
        self.ctx, self.state = state.ctx, state
        self.N, self.Dx = state.N, state.Dx
        p = C.c_void_p()
        self.ctx.check(self.ctx.lib.ssm_bf_history_create(state.ptr, int(capacity),
                                                          C.byref(p)))
        self.ptr = p

    def push(self):
        self.ctx.check(self.ctx.lib.ssm_bf_history_push(self.ptr))

    def record(self, on=True):
        """While on, every step of the filter (per-step or a graph-replayed run) appends its
        generation on the device; do not push those steps by hand as well."""
        self.ctx.check(self.ctx.lib.ssm_bf_history_record(self.ptr, int(bool(on))))

    def backward_sample(self, u, dynamics=None):
        """Backward simulation over the stored generations with uniforms u[0..T] (u[t] for
        time t): (indices (T + 1,), trajectory (T + 1, Dx)).  dynamics: None (the filter's
        homogeneous model) or T PackedLG models, dynamics[t] the parameters of step t + 1
        (csmc.jl:424 weights time t with the transition of step t + 1)."""
        T = len(self)
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(T + 1))
        idx = np.zeros(T + 1, dtype=np.int32)
        traj = np.zeros((T + 1, self.Dx))
        if dynamics is None:
            self.ctx.check(self.ctx.lib.ssm_bf_history_backward_sample(
                self.ptr, u.ctypes.data_as(_lib._dp), idx.ctypes.data_as(_lib._ip),
                traj.ctypes.data_as(_lib._dp)))
            return idx, traj
        if len(dynamics) != T:
            raise ValueError(f"{len(dynamics)} dynamics for {T} stored steps")
        arr = (_lib.LGModelC * T)(*[d.struct for d in dynamics])
        self.ctx.check(self.ctx.lib.ssm_bf_history_backward_sample_tv(
            self.ptr, arr, u.ctypes.data_as(_lib._dp), idx.ctypes.data_as(_lib._ip),
            traj.ctypes.data_as(_lib._dp)))
        return idx, traj

    def reset(self):
        """Empties the container and stores the filter's current particles as time 0."""
        self.ctx.check(self.ctx.lib.ssm_bf_history_reset(self.ptr))

    def __len__(self):
        T = C.c_int32()
        self.ctx.check(self.ctx.lib.ssm_bf_history_info(self.ptr, C.byref(T), None))
        return T.value

    def ancestry(self, i0=0, count=None):
        """Paths of particles i0 .. i0 + count - 1 as an array (count, T + 1, Dx)."""
        count = self.N - i0 if count is None else int(count)
        out = np.zeros((count, len(self) + 1, self.Dx))
        self.ctx.check(self.ctx.lib.ssm_bf_history_ancestry(self.ptr, int(i0), count,
                                                            out.ctypes.data_as(_lib._dp)))
        return out

    def get_ancestry(self, i):
        """Trajectory of particle i: (T + 1, Dx), row 0 the time-0 state."""
        return self.ancestry(i, 1)[0]

    def generation(self, t):
        """(x (N, Dx), logw (N,), ancestors (N,)) of time t >= 1; time 0 gives states only."""
        x = np.zeros((self.N, self.Dx))
        if t == 0:
            self.ctx.check(self.ctx.lib.ssm_bf_history_generation(
                self.ptr, 0, x.ctypes.data_as(_lib._dp), None, None))
            return x, None, None
        lw = np.zeros(self.N)
        anc = np.zeros(self.N, dtype=np.int32)
        self.ctx.check(self.ctx.lib.ssm_bf_history_generation(
            self.ptr, int(t), x.ctypes.data_as(_lib._dp), lw.ctypes.data_as(_lib._dp),
            anc.ctypes.data_as(_lib._ip)))
        return x, lw, anc

    def particle(self, t, i):
        """(state, log-weight, ancestor) of particle i at time t >= 1."""
        x, lw, anc = self.generation(t)
        return x[i], lw[i], int(anc[i])

    def close(self):
        if getattr(self, "ptr", None):
            self.ctx.lib.ssm_bf_history_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DenseAncestorCallback:
    """callback for ``ssmhip.filter``: stores time 0 at PostInit and every generation at
    PostUpdate (callbacks.jl:53-80).  On a homogeneous model ``filter`` keeps its single
    graph-replayed run and the history records the generations on the device."""

    records_on_device = True

    def __init__(self, capacity=64):
        self.capacity = capacity
        self.container = None

    def __call__(self, model, algo, *args):
        event = args[-1]
        state = args[-3] if event == PostUpdate else args[0]
        if event == PostInit:
            self.container = DenseParticleContainer(state, self.capacity)
        elif event == PostUpdate:
            self.container.push()
