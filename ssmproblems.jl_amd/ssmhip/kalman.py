"""KalmanFilter on the device, single and batched (GF/src/algorithms/kalman.jl:16-79).

``filter(rng, model, KalmanFilter(), ys)`` (through ``ssmhip.filter``) runs one filter;
``KFBatch`` holds many filters of uneven size resident on the device and runs them all in
one call, returning each filter's log-likelihood (and optionally its final mean and
covariance).  Parameters come from the unchanged calc_* protocol of each model.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import default_context
from .models import StateSpaceModel, lg_params


@dataclass
class KalmanFilter:
    """KalmanFilter(; jitter=nothing) -- kalman.jl:16-21 (``filter`` runs without jitter,
    the reference default; ``kf_loglikelihood`` takes one)."""
    jitter: float = None


KF = KalmanFilter


def pack_lg(mu0, Sigma0, A, b, Q, H, c, R):
    """One filter's parameters, column-major FP64, in the C-ABI packing order."""
    A = np.atleast_2d(np.asarray(A, dtype=np.float64))
    H = np.atleast_2d(np.asarray(H, dtype=np.float64))
    Dx, Dy = A.shape[0], H.shape[0]
    f = lambda M, r, c_: np.asarray(M, dtype=np.float64).reshape(r, c_).flatten(order="F")
    return np.concatenate([np.asarray(mu0, dtype=np.float64).reshape(Dx), f(Sigma0, Dx, Dx),
                           f(A, Dx, Dx), np.asarray(b, dtype=np.float64).reshape(Dx),
                           f(Q, Dx, Dx), f(H, Dy, Dx), np.asarray(c, dtype=np.float64).reshape(Dy),
                           f(R, Dy, Dy)]), Dx, Dy


class KFBatch:
    """B device-resident Kalman filters; params: list of (mu0, Sigma0, A, b, Q, H, c, R),
    ys: list of (T x Dy_f) observation arrays (same T for all)."""

    def __init__(self, params, ys, ctx=None):
        self.ctx = ctx or default_context()
        packs = [pack_lg(*p) for p in params]
        self.B = len(packs)
        self.Dx = np.array([p[1] for p in packs], dtype=np.int32)
        self.Dy = np.array([p[2] for p in packs], dtype=np.int32)
        sizes = np.array([p[0].size for p in packs], dtype=np.int64)
        self.param_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.params = np.ascontiguousarray(np.concatenate([p[0] for p in packs]))
        ys = [np.asarray(y, dtype=np.float64).reshape(-1, int(d)) for y, d in zip(ys, self.Dy)]
        self.T = ys[0].shape[0]
        if any(y.shape[0] != self.T for y in ys):
            raise ValueError("all filters of a batch share T")
        ysz = np.array([y.size for y in ys], dtype=np.int64)
        self.y_off = np.concatenate([[0], np.cumsum(ysz)[:-1]]).astype(np.int64)
        self.ys = np.ascontiguousarray(np.concatenate([y.reshape(-1) for y in ys]))
        self._create()

    @classmethod
    def from_packed(cls, Dx, Dy, params, param_off, ys, y_off, T, ctx=None):
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        self.B = len(Dx)
        self.Dx = np.ascontiguousarray(Dx, dtype=np.int32)
        self.Dy = np.ascontiguousarray(Dy, dtype=np.int32)
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.param_off = np.ascontiguousarray(param_off, dtype=np.int64)
        self.ys = np.ascontiguousarray(ys, dtype=np.float64)
        self.y_off = np.ascontiguousarray(y_off, dtype=np.int64)
        self.T = int(T)
        self._create()
        return self

    def _create(self):
        p = C.c_void_p()
        i64 = C.POINTER(C.c_int64)
        self.ctx.check(self.ctx.lib.ssm_kf_batch_create(
            self.ctx.ptr, self.B, self.Dx.ctypes.data_as(_lib._ip),
            self.Dy.ctypes.data_as(_lib._ip), self.params.ctypes.data_as(_lib._dp),
            self.param_off.ctypes.data_as(i64), self.T, self.ys.ctypes.data_as(_lib._dp),
            self.y_off.ctypes.data_as(i64), C.byref(p)))
        self.ptr = p

    def run(self, return_state=False):
        ll = np.zeros(self.B)
        st = np.zeros(self.B, dtype=np.int32)
        mu = Sig = None
        if return_state:
            mu = np.zeros(int(self.Dx.sum()))
            Sig = np.zeros(int((self.Dx.astype(np.int64) ** 2).sum()))
        self.ctx.check(self.ctx.lib.ssm_kf_batch_run(
            self.ptr, ll.ctypes.data_as(_lib._dp), st.ctypes.data_as(_lib._ip),
            mu.ctypes.data_as(_lib._dp) if mu is not None else None,
            Sig.ctypes.data_as(_lib._dp) if Sig is not None else None))
        if not return_state:
            return ll
        mus, sigs, mo, co = [], [], 0, 0
        for dx in self.Dx:
            dx = int(dx)
            mus.append(mu[mo:mo + dx])
            sigs.append(Sig[co:co + dx * dx].reshape(dx, dx, order="F"))
            mo += dx
            co += dx * dx
        return ll, mus, sigs

    def last_ms(self):
        ms = C.c_double()
        self.ctx.check(self.ctx.lib.ssm_kf_batch_last_timing(self.ptr, C.byref(ms)))
        return ms.value

    CLASSES = ("Dx<=16", "Dx<=32", "Dx<=64")
    CANDIDATES = ((1, 2, 4), (2, 4, 8), (4, 8))

    def tune(self, T_tune=8):
        """Times each workgroup shape of each size class on this batch (T_tune steps) and
        keeps the fastest for the rest of the process.  Returns {class: {"waves": chosen,
        "ms": {waves: time}}}."""
        nw = (C.c_int32 * 3)()
        ms = (C.c_double * 9)()
        self.ctx.check(self.ctx.lib.ssm_kf_batch_tune(self.ptr, int(T_tune), nw, ms))
        out = {}
        for c, name in enumerate(self.CLASSES):
            out[name] = {"waves": nw[c],
                         "ms": {w: ms[3 * c + k] for k, w in enumerate(self.CANDIDATES[c])}}
        return out

    def close(self):
        if getattr(self, "ptr", None):
            self.ctx.lib.ssm_kf_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class MvNormalState:
    """The filtered distribution MvNormal(mu, Sigma) returned by filter(..., KF(), ...)."""
    mu: np.ndarray
    Sigma: np.ndarray


def kalman_filter(model: StateSpaceModel, observations, ctx=None):
    """filter(model, KalmanFilter(), ys) for a homogeneous LG model on the device."""
    ys = np.asarray(observations, dtype=np.float64)
    p = lg_params(model, 1)
    kb = KFBatch([p], [ys.reshape(ys.shape[0], -1)], ctx=ctx)
    ll, mus, sigs = kb.run(return_state=True)
    kb.close()
    return MvNormalState(mus[0], sigs[0]), float(ll[0])


def _colmajor_blocks(M, r, c_, lead):
    """(lead..., r, c) row-major blocks -> flat FP64, each block column-major."""
    M = np.asarray(M, dtype=np.float64).reshape(*lead, r, c_)
    return np.ascontiguousarray(np.swapaxes(M, -1, -2)).reshape(-1)


def kf_loglikelihood_batch(mu0, Sigma0, A, b, Q, H, c, R, ys, jitter=None, ctx=None,
                           return_status=False):
    """kf_loglikelihood for B filters of one shape in one device call
    (GF/src/integrations/logdensity.jl:149-161).

    mu0 (B, Dx), Sigma0 (B, Dx, Dx), ys (B, T, Dy).  The step parameters are either per step,
    A (B, T, Dx, Dx), b (B, T, Dx), Q (B, T, Dx, Dx), H (B, T, Dy, Dx), c (B, T, Dy),
    R (B, T, Dy, Dy) -- As[t] of the reference -- or one set per filter without the T axis;
    a mix is broadcast along T.  Returns the B log-likelihoods (and the status codes)."""
    ctx = ctx or default_context()
    B, Dx, Dy, T, tv, shapes, mu0c, S0, flat, yv = _kfll_args(mu0, Sigma0, A, b, Q, H, c, R, ys)
    ll = np.zeros(B)
    st = np.zeros(B, dtype=np.int32)
    rc = ctx.lib.ssm_kf_loglikelihood(
        ctx.ptr, B, Dx, Dy, T, int(tv), mu0c.ctypes.data_as(_lib._dp), S0.ctypes.data_as(_lib._dp),
        *(flat[k].ctypes.data_as(_lib._dp) for k in ("A", "b", "Q", "H", "c", "R")),
        yv.ctypes.data_as(_lib._dp), float(jitter or 0.0), ll.ctypes.data_as(_lib._dp),
        st.ctypes.data_as(_lib._ip))
    if rc and not return_status:
        ctx.check(rc)
    return (ll, st) if return_status else ll


class DeviceKFInputs:
    """The inputs of kf_loglikelihood_batch uploaded once to device memory (the same flat
    layouts), for a sampler that keeps them resident: ``loglikelihood()`` calls
    ssm_kf_loglikelihood_device, which copies nothing from the host."""

    def __init__(self, mu0, Sigma0, A, b, Q, H, c, R, ys, ctx=None):
        self.ctx = ctx or default_context()
        (self.B, self.Dx, self.Dy, self.T, self.tv, _, mu0c, S0, flat,
         yv) = _kfll_args(mu0, Sigma0, A, b, Q, H, c, R, ys)
        arrs = [mu0c, S0] + [flat[k] for k in ("A", "b", "Q", "H", "c", "R")] + [yv]
        self.bufs = []
        for a in arrs:
            d = _lib.DeviceBuffer(self.ctx, max(8, a.nbytes))
            d.upload(a)
            self.bufs.append(d)

    def loglikelihood(self, jitter=None, return_status=False):
        ll = np.zeros(self.B)
        st = np.zeros(self.B, dtype=np.int32)
        rc = self.ctx.lib.ssm_kf_loglikelihood_device(
            self.ctx.ptr, self.B, self.Dx, self.Dy, self.T, int(self.tv),
            *(d.ptr for d in self.bufs), float(jitter or 0.0), ll.ctypes.data_as(_lib._dp),
            st.ctypes.data_as(_lib._ip))
        if rc and not return_status:
            self.ctx.check(rc)
        return (ll, st) if return_status else ll


def _kfll_args(mu0, Sigma0, A, b, Q, H, c, R, ys):
    """Shapes, the time-varying flag and the flat column-major inputs of the C-ABI."""
    mu0 = np.asarray(mu0, dtype=np.float64)
    B, Dx = mu0.shape
    ys = np.asarray(ys, dtype=np.float64)
    T = ys.shape[1]
    Dy = ys.reshape(B, T, -1).shape[2]
    shapes = {"A": (Dx, Dx), "b": (Dx,), "Q": (Dx, Dx), "H": (Dy, Dx), "c": (Dy,), "R": (Dy, Dy)}
    arrs = {k: np.asarray(v, dtype=np.float64) for k, v in
            dict(A=A, b=b, Q=Q, H=H, c=c, R=R).items()}
    tv = any(arrs[k].ndim == len(shapes[k]) + 2 for k in shapes)
    flat = {}
    for k, shp in shapes.items():
        a = arrs[k]
        if tv and a.ndim == len(shp) + 1:
            a = np.broadcast_to(a[:, None], (B, T) + shp)
        lead = (B, T) if tv else (B,)
        if a.shape != lead + shp:
            raise ValueError(f"{k}: shape {a.shape}, expected {lead + shp}")
        flat[k] = (_colmajor_blocks(a, *shp, lead) if len(shp) == 2
                   else np.ascontiguousarray(a.reshape(-1)))
    return (B, Dx, Dy, T, tv, shapes, np.ascontiguousarray(mu0.reshape(-1)),
            _colmajor_blocks(Sigma0, Dx, Dx, (B,)), flat,
            np.ascontiguousarray(ys.reshape(-1)))


def kf_loglikelihood_grad_batch(mu0, Sigma0, A, b, Q, H, c, R, ys, jitter=None, ctx=None):
    """kf_loglikelihood and its reverse rule (kalman_rrule.jl:22-107) for B filters up to
    Dx = 64, Dy = 32, in FP64: returns (ll (B,), grads) with grads["mu0"] (B, Dx), grads["Sigma0"] (B, Dx, Dx) and
    grads[k] for k in A, b, Q, H, c, R shaped like the inputs (per step, or summed over the
    steps for one parameter set per filter)."""
    ctx = ctx or default_context()
    B, Dx, Dy, T, tv, shapes, mu0c, S0, flat, yv = _kfll_args(mu0, Sigma0, A, b, Q, H, c, R, ys)
    lead = (B, T) if tv else (B,)
    out = {"mu0": np.zeros(B * Dx), "Sigma0": np.zeros(B * Dx * Dx)}
    for k, shp in shapes.items():
        out[k] = np.zeros(int(np.prod(lead + shp)))
    ll = np.zeros(B)
    st = np.zeros(B, dtype=np.int32)
    ctx.check(ctx.lib.ssm_kf_loglikelihood_grad(
        ctx.ptr, B, Dx, Dy, T, int(tv), mu0c.ctypes.data_as(_lib._dp), S0.ctypes.data_as(_lib._dp),
        *(flat[k].ctypes.data_as(_lib._dp) for k in ("A", "b", "Q", "H", "c", "R")),
        yv.ctypes.data_as(_lib._dp), float(jitter or 0.0), ll.ctypes.data_as(_lib._dp),
        st.ctypes.data_as(_lib._ip),
        *(out[k].ctypes.data_as(_lib._dp) for k in ("mu0", "Sigma0", "A", "b", "Q", "H", "c", "R"))))
    unflat = lambda v, r, cc, ld: np.swapaxes(v.reshape(*ld, cc, r), -1, -2)  # column-major blocks
    grads = {"mu0": out["mu0"].reshape(B, Dx), "Sigma0": unflat(out["Sigma0"], Dx, Dx, (B,))}
    for k, shp in shapes.items():
        grads[k] = unflat(out[k], *shp, lead) if len(shp) == 2 else out[k].reshape(lead + shp)
    return ll, grads


class SRKalmanFilter:
    """SRKalmanFilter(; jitter=nothing) -- GF/src/algorithms/srkf.jl:93-176: the covariance
    carried as its upper Cholesky factor, propagated by QR factorisations."""

    def __init__(self, jitter=None):
        self.jitter = jitter


def SRKF():
    return SRKalmanFilter()


def srkf_filter_batch(mu0, Sigma0, A, b, Q, H, c, R, ys, jitter=None, ctx=None):
    """filter(model, SRKF(), ys) for B filters of one shape up to Dx = 64, Dy = 32 in one call,
    arguments as kf_loglikelihood_batch.  Returns (ll (B,), mu (B, Dx), U (B, Dx, Dx)), U the
    final filtered covariance's upper Cholesky factor (Sigma = U' U)."""
    ctx = ctx or default_context()
    B, Dx, Dy, T, tv, shapes, mu0c, S0, flat, yv = _kfll_args(mu0, Sigma0, A, b, Q, H, c, R, ys)
    ll = np.zeros(B)
    st = np.zeros(B, dtype=np.int32)
    mu = np.zeros(B * Dx)
    U = np.zeros(B * Dx * Dx)
    ctx.check(ctx.lib.ssm_srkf_filter(
        ctx.ptr, B, Dx, Dy, T, int(tv), mu0c.ctypes.data_as(_lib._dp), S0.ctypes.data_as(_lib._dp),
        *(flat[k].ctypes.data_as(_lib._dp) for k in ("A", "b", "Q", "H", "c", "R")),
        yv.ctypes.data_as(_lib._dp), float(jitter or 0.0), ll.ctypes.data_as(_lib._dp),
        st.ctypes.data_as(_lib._ip), mu.ctypes.data_as(_lib._dp), U.ctypes.data_as(_lib._dp)))
    return ll, mu.reshape(B, Dx), np.swapaxes(U.reshape(B, Dx, Dx), -1, -2)


def kf_loglikelihood(mu0, Sigma0, As, bs, Qs, Hs, cs, Rs, ys, jitter=None, ctx=None):
    """kf_loglikelihood(mu0, Sigma0, As, bs, Qs, Hs, cs, Rs, ys, jitter) of one filter: the
    sequences hold one parameter per step (logdensity.jl:149-161)."""
    ys = np.asarray(ys, dtype=np.float64)
    T = ys.shape[0]
    ys = ys.reshape(1, T, -1)
    st = lambda seq: np.asarray(seq, dtype=np.float64)[None]
    return float(kf_loglikelihood_batch(np.asarray(mu0, dtype=np.float64)[None], st(Sigma0),
                                        st(As), st(bs), st(Qs), st(Hs), st(cs), st(Rs), ys,
                                        jitter=jitter, ctx=ctx)[0])
