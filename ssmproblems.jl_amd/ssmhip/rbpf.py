"""RBPF(BF(N), KalmanFilter()) on the device (GF/src/algorithms/rbpf.jl:8-84).

The hierarchical model (GF/src/models/hierarchical.jl:4-45) has a scalar outer state and a
linear-Gaussian inner state whose drift depends on the previous outer state (prev_outer)
and whose observation noise is exp(rs * x_t) R0 (new_outer); see ssm_rb_model in
include/ssmhip.h.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import RBModelC, default_context
from .filters import BootstrapFilter, _proposal, rng_to_seed
from .kalman import KalmanFilter


@dataclass
class RBPF:
    """RBPF(pf, af) -- rbpf.jl:8-14."""
    pf: BootstrapFilter
    af: KalmanFilter = field(default_factory=KalmanFilter)


class RBModel:
    def __init__(self, om0, os0, oa, ob, oq, mu0, Sigma0, A, b, Cv, Q, H, c, R0, rs):
        f = lambda M: np.ascontiguousarray(np.atleast_2d(np.asarray(M, dtype=np.float64)).flatten(order="F"))
        v = lambda x: np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
        self.D = np.atleast_2d(A).shape[0]
        self.Dy = np.atleast_2d(H).shape[0]
        self._bufs = [v(mu0), f(Sigma0), f(A), v(b), v(Cv), f(Q), f(H), v(c), f(R0)]
        self.struct = RBModelC(float(om0), float(os0), float(oa), float(ob), float(oq), self.D,
                               self.Dy, float(rs), *[b_.ctypes.data_as(_lib._dp) for b_ in self._bufs])


class DeviceRBParticles:
    def __init__(self, ctx, model: RBModel, N: int, seed: int):
        self.ctx, self.model, self.N, self.seed = ctx, model, N, seed
        p = C.c_void_p()
        ctx.check(ctx.lib.ssm_rbpf_create(ctx.ptr, C.byref(model.struct), N, seed, C.byref(p)))
        self.ptr = p

    def initialise(self, z=None):
        zp = None
        if z is not None:
            z = np.ascontiguousarray(np.asarray(z, dtype=np.float64).reshape(self.N))
            zp = z.ctypes.data_as(_lib._dp)
        self.ctx.check(self.ctx.lib.ssm_rbpf_initialise(self.ptr, zp))

    def step(self, t, y, threshold, kind, u=None, z=None):
        y = np.ascontiguousarray(np.asarray(y, dtype=np.float64).reshape(self.model.Dy))
        ll = C.c_double()
        did = C.c_int32()
        up = zp = None
        if u is not None:
            u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
            up = u.ctypes.data_as(_lib._dp)
        if z is not None:
            z = np.ascontiguousarray(np.asarray(z, dtype=np.float64).reshape(self.N))
            zp = z.ctypes.data_as(_lib._dp)
        self.ctx.check(self.ctx.lib.ssm_rbpf_step(self.ptr, t, y.ctypes.data_as(_lib._dp),
                                                  float(threshold), kind, up, zp, C.byref(ll),
                                                  C.byref(did)))
        self.last_resampled = bool(did.value)
        return ll.value

    def run(self, ys, threshold, kind, t0=1, return_steps=False):
        ys = np.ascontiguousarray(np.asarray(ys, dtype=np.float64).reshape(-1, self.model.Dy))
        T = ys.shape[0]
        tot = C.c_double()
        steps = np.zeros(T) if return_steps else None
        self.ctx.check(self.ctx.lib.ssm_rbpf_run(
            self.ptr, t0, T, ys.ctypes.data_as(_lib._dp), float(threshold), kind, C.byref(tot),
            steps.ctypes.data_as(_lib._dp) if steps is not None else None))
        return (tot.value, steps) if return_steps else tot.value

    def export(self):
        N, D = self.N, self.model.D
        xo, mu, Sig = np.zeros(N), np.zeros((N, D)), np.zeros((N, D, D))
        lw, anc, w = np.zeros(N), np.zeros(N, dtype=np.int32), np.zeros(N)
        self.ctx.check(self.ctx.lib.ssm_rbpf_export(
            self.ptr, xo.ctypes.data_as(_lib._dp), mu.ctypes.data_as(_lib._dp),
            Sig.ctypes.data_as(_lib._dp), lw.ctypes.data_as(_lib._dp),
            anc.ctypes.data_as(_lib._ip), w.ctypes.data_as(_lib._dp)))
        return xo, mu, Sig, lw, anc, w

    def last_ms(self):
        ms = C.c_double()
        self.ctx.check(self.ctx.lib.ssm_rbpf_last_timing(self.ptr, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "ptr", None):
            self.ctx.lib.ssm_rbpf_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rbpf_filter(rng, model: RBModel, algo: RBPF, observations, ctx=None, seed=None):
    """filter(rng, hier_model, RBPF(BF(N), KF()), ys) -> (state, log_evidence)."""
    ctx = ctx or default_context()
    if _proposal(algo.pf) is not None:
        raise TypeError("the device RBPF runs the bootstrap proposal for the outer state")
    st = DeviceRBParticles(ctx, model, algo.pf.N, rng_to_seed(rng) if seed is None else seed)
    st.initialise()
    ll = st.run(observations, algo.pf.threshold, algo.pf.resampler.kind)
    return st, ll
