"""Filtering server: continuous batching of whole filter runs on the device.

A request is one ``filter(model, algo, ys)`` call (GF/src/GeneralisedFilters.jl:58-84) with
a ``BootstrapFilter`` or a ``KalmanFilter`` on a linear-Gaussian model -- the likelihood
evaluations an outer parameter sampler issues.  ``Server.submit`` hands the model
parameters (read through the unchanged calc_* protocol) and the observations to the
library, which admits requests by its HBM budget and runs every admitted filter
concurrently on the device; ``wait``/``poll`` return the log-evidence and the latency from
submit to completion.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import RequestC, default_context
from .filters import BootstrapFilter, _proposal, rng_to_seed
from .kalman import KalmanFilter
from .models import PackedLG, StateSpaceModel

QUEUED, RUNNING, DONE, FAILED, CANCELLED = 0, 1, 2, 3, 4
REQ_BF, REQ_KF = 0, 1
STATS_FIELDS = ("completed", "failed", "cancelled", "bf_batched_steps", "kf_groups",
                "bytes_in_use", "peak_bytes", "peak_active_bf", "arena_bytes", "queued",
                "bf_chunks", "bf_graph_captures", "bf_capture_ms", "bf_chunk_ms")


@dataclass
class Result:
    state: int
    ll: float
    latency_ms: float
    status: int


class Server:
    """ssm_server: a device arena of ``hbm_budget_bytes`` and at most ``max_active``
    concurrent bootstrap filters."""

    def __init__(self, hbm_budget_bytes=1 << 30, max_active=256, ctx=None):
        self.ctx = ctx or default_context()
        p = C.c_void_p()
        self.ctx.check(self.ctx.lib.ssm_serve_create(self.ctx.ptr, int(hbm_budget_bytes),
                                                     int(max_active), C.byref(p)))
        self.ptr = p
        self.lib = self.ctx.lib

    def _submit(self, kind, packed: PackedLG, ys, N=0, threshold=1.0, resampler=0, seed=0):
        ys = np.ascontiguousarray(np.asarray(ys, dtype=np.float64).reshape(-1, packed.Dy))
        q = RequestC(kind, packed.struct, int(N), ys.shape[0], ys.ctypes.data_as(_lib._dp),
                     float(threshold), int(resampler), int(seed))
        rid = C.c_int64()
        self.ctx.check(self.lib.ssm_serve_submit(self.ptr, C.byref(q), C.byref(rid)))
        return rid.value  # the library copied the parameters and observations

    def submit(self, rng, model: StateSpaceModel, algo, observations, seed=None):
        """Queues filter(rng, model, algo, observations); returns the request id.  For a
        BootstrapFilter the rng is consumed once to key the device noise (or pass seed)."""
        packed = PackedLG.from_model(model, 1)
        if isinstance(algo, KalmanFilter):
            return self._submit(REQ_KF, packed, observations)
        if isinstance(algo, BootstrapFilter):
            if _proposal(algo) is not None:
                raise _lib.ArgumentError(_lib.SSM_ERR_ARG,
                                         "the server runs the bootstrap proposal only")
            if algo.precision != "f32":
                raise _lib.ArgumentError(_lib.SSM_ERR_ARG, "the server runs FP32 particles")
            s = rng_to_seed(rng) if seed is None else int(seed)
            return self._submit(REQ_BF, packed, observations, N=algo.N,
                                threshold=algo.threshold, resampler=algo.resampler.kind,
                                seed=s)
        raise _lib.ArgumentError(_lib.SSM_ERR_ARG, f"unsupported algorithm {type(algo)}")

    def submit_params(self, kind, params, ys, N=0, threshold=1.0, resampler=0, seed=0):
        """Low-level submit from (mu0, Sigma0, A, b, Q, H, c, R)."""
        return self._submit(kind, PackedLG(*params), ys, N, threshold, resampler, seed)

    def poll(self, rid) -> Result:
        st, ll, lat, status = C.c_int32(), C.c_double(), C.c_double(), C.c_int32()
        self.ctx.check(self.lib.ssm_serve_poll(self.ptr, int(rid), C.byref(st), C.byref(ll),
                                               C.byref(lat), C.byref(status)))
        return Result(st.value, ll.value, lat.value, status.value)

    def wait(self, rid, timeout_ms=-1.0) -> Result:
        st, ll, lat, status = C.c_int32(), C.c_double(), C.c_double(), C.c_int32()
        self.ctx.check(self.lib.ssm_serve_wait(self.ptr, int(rid), float(timeout_ms),
                                               C.byref(st), C.byref(ll), C.byref(lat),
                                               C.byref(status)))
        return Result(st.value, ll.value, lat.value, status.value)

    def cancel(self, rid):
        self.ctx.check(self.lib.ssm_serve_cancel(self.ptr, int(rid)))

    def release(self, rid):
        self.ctx.check(self.lib.ssm_serve_release(self.ptr, int(rid)))

    def stats(self):
        out = (C.c_double * len(STATS_FIELDS))()
        self.ctx.check(self.lib.ssm_serve_stats(self.ptr, out, len(STATS_FIELDS)))
        return dict(zip(STATS_FIELDS, list(out)))

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.ssm_serve_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
