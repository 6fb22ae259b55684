"""Filters and the filter/step driver, mirroring GeneralisedFilters' host API.

``filter(rng, model, algo, observations)`` keeps the reference's loop and log-evidence
accumulation (GF/src/GeneralisedFilters.jl:58-84): initialise, then step t = 1..T.  For a
BootstrapFilter on a linear-Gaussian model the particle state lives on the device; each
``step`` is one call into the library (resample if the ESS rule says so, predict, update,
marginalise), and with no callback a homogeneous model runs all T steps in one call whose
steps replay from HIP graphs.  Model parameters are read through the unchanged calc_params
protocol (GF/src/models/linear_gaussian.jl:44-63) at the same points as the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import SSM_F32, SSM_F64, default_context
from .models import PackedLG, StateSpaceModel, lg_params
from .proposals import LatentProposal
from .resamplers import AbstractResampler, ESSResampler, Systematic, draw_uniforms

# Callback triggers (GF/src/callbacks.jl:16-44).
PostInit, PostResample, PostPredict, PostUpdate = "PostInit", "PostResample", "PostPredict", "PostUpdate"


def rng_to_seed(rng: np.random.Generator) -> int:
    """Consumes one UInt64 from the host rng to key the device Philox streams."""
    return int(rng.integers(0, 2**63 - 1, dtype=np.int64))


@dataclass
class BootstrapFilter:
    """BF(N; threshold=1.0, resampler=Systematic()) -- particles.jl:92-108, :197-201."""
    N: int
    threshold: float = 1.0
    resampler: AbstractResampler = field(default_factory=Systematic)
    precision: str = "f32"

    @property
    def ess_resampler(self):
        return ESSResampler(self.threshold, self.resampler)


BF = BootstrapFilter


class ParticleFilter(BootstrapFilter):
    """ParticleFilter(N, proposal; threshold=1.0, resampler=Systematic()) -- particles.jl:
    92-108.  A linear-Gaussian proposal (ssmhip.proposals) makes it a guided filter; the
    LatentProposal is the bootstrap filter."""

    def __init__(self, N, proposal, threshold=1.0, resampler=None, precision="f32"):
        super().__init__(N, threshold, resampler if resampler is not None else Systematic(),
                         precision)
        self.proposal = proposal

    @property
    def guided(self):
        return not isinstance(self.proposal, LatentProposal)


PF = ParticleFilter


class PredictiveStatistic:
    """The representative state of a look-ahead score (particles.jl:264-266)."""


class MeanPredictive(PredictiveStatistic):
    """The transition mean (particles.jl:346-352)."""


class ModePredictive(PredictiveStatistic):
    """The transition mode (particles.jl:354-360): the mean for Gaussian transitions."""


class AuxiliaryParticleFilter(BootstrapFilter):
    """AuxiliaryParticleFilter(pf, MeanPredictive()) -- particles.jl:272-281, :316-344 with
    RepresentativeStateLookAhead: each step first scores the particles with the observation
    density at their transition mean and resamples with the AuxiliaryResampler
    (resamplers.jl:92-143).  pf is a BootstrapFilter; its N, threshold, resampler and
    precision are the filter's."""

    def __init__(self, pf: BootstrapFilter, pp: PredictiveStatistic = None):
        if _proposal(pf) is not None:
            raise TypeError("the device auxiliary filter wraps a bootstrap filter")
        pp = MeanPredictive() if pp is None else pp
        if not isinstance(pp, (MeanPredictive, ModePredictive)):
            raise TypeError(f"look-ahead {type(pp).__name__} is not supported on the device "
                            "(MeanPredictive, ModePredictive)")
        super().__init__(pf.N, pf.threshold, pf.resampler, pf.precision)
        self.pf = pf
        self.weight_strategy = pp


APF = AuxiliaryParticleFilter


def _proposal(algo):
    """The filter's proposal if it is not the latent dynamics."""
    prop = getattr(algo, "proposal", None)
    return None if prop is None or isinstance(prop, LatentProposal) else prop


def _homogeneous(model, algo):
    prop = _proposal(algo)
    return model.is_homogeneous and (prop is None or prop.is_homogeneous)


class DeviceParticles:
    """Device-resident ParticleDistribution (containers.jl:125-141) of a BF on an LG model."""

    def __init__(self, ctx, model: StateSpaceModel, algo: BootstrapFilter, seed: int):
        self.ctx = ctx
        self.algo = algo
        self.N = algo.N
        self.packed = PackedLG.from_model(model, 1)
        self.Dx, self.Dy = self.packed.Dx, self.packed.Dy
        prec = SSM_F32 if algo.precision == "f32" else SSM_F64
        p = C.c_void_p()
        ctx.check(ctx.lib.ssm_bf_create(ctx.ptr, self.packed.ref(), self.N, prec, seed,
                                        C.byref(p)))
        self.ptr = p
        self.seed = seed
        if _proposal(algo) is not None:
            self.set_proposal(_proposal(algo).linear_form(1))
        if isinstance(algo, AuxiliaryParticleFilter):
            ctx.check(ctx.lib.ssm_bf_set_auxiliary(self.ptr, _lib.SSM_APF_MEAN))

    def set_proposal(self, form):
        """Uploads a linear-Gaussian proposal (F, g0, K, P); None: the latent dynamics."""
        if form is None:
            self.ctx.check(self.ctx.lib.ssm_bf_set_proposal(self.ptr, None, None, None, None))
            return
        F, g0, K, P = form
        cm = lambda M, r, c: np.ascontiguousarray(
            np.asarray(M, dtype=np.float64).reshape(r, c).T).reshape(-1)
        Fv, Pv = cm(F, self.Dx, self.Dx), cm(P, self.Dx, self.Dx)
        gv = np.ascontiguousarray(np.asarray(g0, dtype=np.float64).reshape(self.Dx))
        Kv = None if K is None else cm(K, self.Dx, self.Dy)
        self.ctx.check(self.ctx.lib.ssm_bf_set_proposal(
            self.ptr, Fv.ctypes.data_as(_lib._dp), gv.ctypes.data_as(_lib._dp),
            None if Kv is None else Kv.ctypes.data_as(_lib._dp), Pv.ctypes.data_as(_lib._dp)))

    def set_params(self, packed: PackedLG):
        self.packed = packed
        self.ctx.check(self.ctx.lib.ssm_bf_set_model(self.ptr, packed.ref()))

    def export(self):
        """(x [N x Dx], normalised logw [N], ancestors [N], resample weights [N])."""
        x = np.zeros((self.N, self.Dx))
        lw = np.zeros(self.N)
        anc = np.zeros(self.N, dtype=np.int32)
        w = np.zeros(self.N)
        self.ctx.check(self.ctx.lib.ssm_bf_export(
            self.ptr, x.ctypes.data_as(_lib._dp), lw.ctypes.data_as(_lib._dp),
            anc.ctypes.data_as(_lib._ip), w.ctypes.data_as(_lib._dp)))
        return x, lw, anc, w

    def import_state(self, x, logw):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(self.N, self.Dx))
        lw = np.ascontiguousarray(np.asarray(logw, dtype=np.float64))
        self.ctx.check(self.ctx.lib.ssm_bf_import(self.ptr, x.ctypes.data_as(_lib._dp),
                                                  lw.ctypes.data_as(_lib._dp)))

    def set_reference(self, x_ref, mode=_lib.SSM_CSMC_PLAIN):
        """Pins particle 0 to the reference trajectory x_ref ((T + 1) x Dx, row t = time t)
        for conditional SMC; None returns to the unconditional filter."""
        if x_ref is None:
            self.ctx.check(self.ctx.lib.ssm_bf_set_reference(self.ptr, 0, None, 0))
            self.ref_T = None
            return
        xr = np.ascontiguousarray(np.asarray(x_ref, dtype=np.float64).reshape(-1, self.Dx))
        self.ctx.check(self.ctx.lib.ssm_bf_set_reference(self.ptr, xr.shape[0] - 1,
                                                         xr.ctypes.data_as(_lib._dp), int(mode)))
        self.ref_T = xr.shape[0] - 1

    def sample_index(self, u):
        """0-based index of the first particle whose cumulative normalised weight reaches u."""
        idx = C.c_int32()
        self.ctx.check(self.ctx.lib.ssm_bf_sample_index(self.ptr, float(u), C.byref(idx)))
        return idx.value

    @property
    def particles(self):
        return self.export()[0]

    def weights(self):
        lw = self.export()[1]
        w = np.exp(lw - lw.max())
        return w / w.sum()

    def close(self):
        if self.ptr:
            self.ctx.lib.ssm_bf_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def initialise(rng, model: StateSpaceModel, algo: BootstrapFilter, z=None, ctx=None, seed=None):
    """initialise(rng, prior, BF) (particles.jl:36-50) on the device."""
    if not model.is_linear_gaussian:
        raise TypeError("the device BootstrapFilter needs a linear-Gaussian model")
    ctx = ctx or default_context()
    seed = rng_to_seed(rng) if seed is None else seed
    st = DeviceParticles(ctx, model, algo, seed)
    zp = None
    if z is not None:
        z = np.ascontiguousarray(np.asarray(z, dtype=np.float64).reshape(algo.N, st.Dx))
        zp = z.ctypes.data_as(_lib._dp)
    ctx.check(ctx.lib.ssm_bf_initialise(st.ptr, zp))
    return st


def step(rng, model: StateSpaceModel, algo: BootstrapFilter, t: int, state: DeviceParticles,
         observation, callback=None, u=None, z=None):
    """step(rng, model, BF, t, state, y) -> (state, ll_increment), particles.jl:147-162.

    ``u``/``z`` override the device's own Philox draws (parity tests); otherwise the
    draws are keyed by (state.seed, t) and do not consume ``rng``."""
    if not model.is_homogeneous:
        state.set_params(PackedLG.from_model(model, t))
    prop = _proposal(algo)
    if prop is not None and not prop.is_homogeneous:  # calc_params of the proposal at t
        state.set_proposal(prop.linear_form(t))
    y = np.ascontiguousarray(np.asarray(observation, dtype=np.float64).reshape(state.Dy))
    ll = C.c_double()
    did = C.c_int32()
    up = zp = None
    if u is not None:
        u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
        up = u.ctypes.data_as(_lib._dp)
    if z is not None:
        z = np.ascontiguousarray(np.asarray(z, dtype=np.float64).reshape(state.N, state.Dx))
        zp = z.ctypes.data_as(_lib._dp)
    state.ctx.check(state.ctx.lib.ssm_bf_step(state.ptr, t, y.ctypes.data_as(_lib._dp),
                                              algo.threshold, algo.resampler.kind, up, zp,
                                              C.byref(ll), C.byref(did)))
    state.last_resampled = bool(did.value)
    if callback is not None:
        callback(model, algo, t, state, observation, PostUpdate)
    return state, ll.value


def run(state: DeviceParticles, algo: BootstrapFilter, observations, t0=1, return_steps=False):
    """All steps t0..t0+T-1 of a homogeneous model in one library call (HIP graphs)."""
    ys = np.ascontiguousarray(np.asarray(observations, dtype=np.float64).reshape(-1, state.Dy))
    T = ys.shape[0]
    tot = C.c_double()
    steps = np.zeros(T) if return_steps else None
    state.ctx.check(state.ctx.lib.ssm_bf_run(
        state.ptr, t0, T, ys.ctypes.data_as(_lib._dp), algo.threshold, algo.resampler.kind,
        C.byref(tot), steps.ctypes.data_as(_lib._dp) if steps is not None else None))
    return (tot.value, steps) if return_steps else tot.value


def last_timing(state: DeviceParticles):
    ms = C.c_double()
    per = C.c_double()
    state.ctx.check(state.ctx.lib.ssm_bf_last_timing(state.ptr, C.byref(ms), C.byref(per)))
    return ms.value, per.value


def filter(rng, model: StateSpaceModel, algo, observations, callback=None, ctx=None):
    """filter(rng, model, algo, observations) -> (state, log_evidence).

    GeneralisedFilters.jl:58-84: initialise, callback(PostInit), then step 1..T summing
    the increments.  A homogeneous model without a callback takes the single-call path."""
    if isinstance(algo, BootstrapFilter):
        state = initialise(rng, model, algo, ctx=ctx)
        if callback is not None:
            callback(model, algo, state, observations, PostInit)
        ys = np.asarray(observations, dtype=np.float64).reshape(len(observations), -1)
        if callback is None and _homogeneous(model, algo):
            return state, run(state, algo, ys)
        container = getattr(callback, "container", None)
        if getattr(callback, "records_on_device", False) and _homogeneous(model, algo):
            # a dense history takes every generation inside the run's own graphs
            container.record(True)
            try:
                return state, run(state, algo, ys)
            finally:
                container.record(False)
        state, ll = step(rng, model, algo, 1, state, ys[0], callback=callback)
        for t in range(2, ys.shape[0] + 1):
            state, inc = step(rng, model, algo, t, state, ys[t - 1], callback=callback)
            ll += inc
        return state, ll
    from .kalman import KalmanFilter, MvNormalState, SRKalmanFilter, kalman_filter
    from .kalman import srkf_filter_batch
    if isinstance(algo, SRKalmanFilter):  # srkf.jl:93-176; calc_params at every step
        if not model.is_linear_gaussian:
            raise TypeError("the device SRKF needs a linear-Gaussian model")
        ys = np.asarray(observations, dtype=np.float64).reshape(len(observations), -1)
        T = ys.shape[0]
        steps = [lg_params(model, t) for t in range(1, T + 1)]
        mu0, Sigma0 = steps[0][0], np.atleast_2d(steps[0][1])
        Dx, Dy = mu0.shape[0], ys.shape[1]
        shp = (None, None, (Dx, Dx), (Dx,), (Dx, Dx), (Dy, Dx), (Dy,), (Dy, Dy))
        per = lambda k: np.stack([np.reshape(q[k], shp[k]) for q in steps])[None]
        ll, mu, U = srkf_filter_batch(mu0[None], Sigma0[None], per(2), per(3), per(4), per(5),
                                      per(6), per(7), ys[None], jitter=algo.jitter, ctx=ctx)
        st = MvNormalState(mu[0], U[0].T @ U[0])
        st.U = U[0]
        return st, float(ll[0])
    if isinstance(algo, KalmanFilter):
        if not model.is_homogeneous:
            raise TypeError("the device KalmanFilter path needs a homogeneous LG model")
        return kalman_filter(model, observations, ctx=ctx)
    raise TypeError(f"unsupported filter {type(algo).__name__}")
