"""FP64 CPU oracle for the GeneralisedFilters filtering hot path -- TEST INFRASTRUCTURE.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may
import this module, and only as the checker (or the timed CPU baseline), never as the thing
measured or shipped.  The product library (``libssmhip.so``) does not link or call it.

It wraps ``liboracle.so`` (built from ``oracle.c`` by ``oracle/Makefile``) and adds a numpy
restatement of the closed-form joint Gaussian used by the reference's Kalman tests.

Parity pinning of each oracle function (what checks it beyond retyping its formula):

* ``philox4x32_10``: the Random123 known-answer vectors (Salmon et al. SC'11).
* ``systematic`` (resamplers.jl:216-238), ``systematic_murray`` (CUDAExt.jl:49-58),
  ``stratified`` (CUDAExt.jl:68-78), ``multinomial`` (CUDAExt.jl:31-39): brute force in
  exact rational arithmetic on small inputs (tests/golden/gen_resample_fixtures.py), the
  reference's known-answer vectors for offspring<->ancestors (test/components/resamplers.jl:
  76-92), closed special cases (uniform weights -> identity, one-hot -> constant) and the
  reference's statistical test (resampled mean ~ weighted mean, resamplers.jl:5-34).
* ``kf_filter`` / ``kf_update`` (kalman.jl:42-79, kalman_gradient.jl:12-29): the closed-form
  joint Gaussian (``joint_gaussian_loglik``, restating GFTest/models/linear_gaussian.jl:171-215,
  as test/algorithms/kalman.jl:5-43 does) and the MVN conditional for T=1.
* ``kf_loglikelihood`` (logdensity.jl:149-161): the joint Gaussian with per-step parameters;
  with jitter j, the jitter-free joint Gaussian whose Q_t (t >= 2) is Q_t + j A_t A_t'
  (A (S + j I) A' + Q = A S A' + (Q + j A A')); constant parameters equal ``kf_filter``
  (the reference's kf_loglikelihood == filter(KF) test, logdensity.jl:93-132).
* ``bf_step`` (particles.jl:36-162, containers.jl:140-173): special cases (H = 0 makes every
  particle's weight equal, so ll_t is the exact N(y; c, R) density; Q -> 0 makes propagation
  deterministic; uniform weights give ESS = N and an identity systematic resample; the baseline
  after a resample is log N) and the reference's statistical test against the exact KF
  log-likelihood (test/algorithms/particles.jl:5-27).
* ``rbpf_step`` (rbpf.jl:16-84): RBPF log-likelihood against the KF of the joint model built
  like GFTest's dummy LG model (test/algorithms/rbpf.jl:5-34), and the degenerate outer-noise-0
  case where every particle is identical and the RBPF equals a KF with R_t = exp(x_t) R0.
* ``PTree`` (ParallelParticleTree, CUDAExt.jl:146-262): paths against a brute-force walk of
  the dense ancestor history (DenseParticleContainer's get_ancestry, containers.jl:406-418),
  nodes in use against the number of distinct (time, particle) pairs on the surviving lines,
  and a hand-worked insert with pool growth (tests/test_oracle_tree.py).
* ``csmc_step`` / ``as_logweights`` / ``categorical`` / ``csmc_sweep`` (csmc.jl:239-306,
  ancestor_sampling.jl:64-130, resamplers.jl:51-65): the ancestor-sampling log-weights
  against scipy's multivariate normal density, the categorical draw against a binary search
  of the cumulative weights, the step against ``bf_step`` on every particle but the pinned
  one, and the reference's statistical test of both sweeps (test/algorithms/csmc.jl:5-95):
  the CSMC chain's mean of x_2 against the exact smoothing mean from the joint Gaussian.
* ``guided_step`` / ``optimal_proposal`` / ``overdispersed_proposal`` (particles.jl:92-195,
  GFTest/proposals.jl:11-49): the proposal N(A x + b, Q) reproduces ``bf_step`` (the C
  restatement) in states, weights, ancestors and log-likelihood; with the optimal proposal
  every particle's weight increment equals the closed-form predictive density
  N(y; H (A x + b) + c, H Q H' + R), whatever its draw; with the overdispersed proposal the
  increment log f - log q equals (D/2) log k - (k - 1) |z|^2 / 2; and the reference's
  guided-filter test (test/algorithms/particles.jl:30-50) against the KF.
* ``srkf_filter`` (srkf.jl:93-176): log-likelihood, filtered mean and U' U against the
  Joseph-form KF (``kf_step``), which is the same filter in exact arithmetic, over several
  shapes with per-step parameters; U upper triangular with a positive diagonal; with
  jitter j at T = 1, U - j I equals the Cholesky factor of the KF's filtered covariance.
* ``logsumexp``: parity unpinned beyond closed cases (LogExpFunctions' exact rounding is
  third-party; only tolerances are claimed).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

SYSTEMATIC, SYSTEMATIC_MURRAY, STRATIFIED, MULTINOMIAL = 0, 1, 2, 3
STREAM_PROPAGATE, STREAM_RESAMPLE, STREAM_INIT, STREAM_OUTER, STREAM_ANCESTOR = 0, 1, 2, 3, 4


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)


class LGModel(C.Structure):
    _fields_ = [("Dx", C.c_int32), ("Dy", C.c_int32)] + [
        (n, _dp) for n in ("mu0", "Sigma0", "A", "b", "Q", "H", "c", "R")
    ]


class RBModel(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("om0", "os0", "oa", "ob", "oq")] + [
        ("D", C.c_int32), ("Dy", C.c_int32), ("rs", C.c_double)] + [
        (n, _dp) for n in ("mu0", "Sigma0", "A", "b", "C", "Q", "H", "c", "R0")]


def _declare(L):
    L.orc_philox4x32_10.argtypes = [_up, _up, _up]
    L.orc_philox_normals.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, C.c_int,
                                     C.c_int, _dp]
    L.orc_philox_uniforms.argtypes = [C.c_uint64, C.c_uint32, C.c_int64, _dp]
    L.orc_resample.argtypes = [C.c_int, C.c_int64, _dp, _dp, _ip]
    L.orc_resample.restype = C.c_int
    L.orc_offspring_to_ancestors.argtypes = [C.c_int64, _ip, _ip]
    L.orc_ancestors_to_offspring.argtypes = [C.c_int64, _ip, _ip]
    L.orc_logsumexp.argtypes = [C.c_int64, _dp]
    L.orc_logsumexp.restype = C.c_double
    L.orc_bf_initialise.argtypes = [C.POINTER(LGModel), C.c_int64, _dp, _dp]
    L.orc_bf_initialise.restype = C.c_int
    L.orc_bf_step.argtypes = [C.POINTER(LGModel), C.c_int64, _dp, _dp, C.POINTER(C.c_int),
                              C.c_double, C.c_int, _dp, _dp, _dp, _dp, _ip, _dp,
                              C.POINTER(C.c_int), _dp]
    L.orc_bf_step.restype = C.c_int
    L.orc_apf_step.argtypes = [C.POINTER(LGModel), C.c_int64, _dp, _dp, C.POINTER(C.c_int),
                               C.c_double, C.c_int, _dp, _dp, _dp, _dp, _ip, _dp,
                               C.POINTER(C.c_int), _dp]
    L.orc_apf_step.restype = C.c_int
    L.orc_csmc_step.argtypes = [C.POINTER(LGModel), C.c_int64, _dp, _dp, C.POINTER(C.c_int),
                                C.c_double, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int32, _ip,
                                _dp, C.POINTER(C.c_int), _dp]
    L.orc_csmc_step.restype = C.c_int
    L.orc_as_logweights.argtypes = [C.POINTER(LGModel), C.c_int64, _dp, _dp, C.c_int, _dp, _dp]
    L.orc_as_logweights.restype = C.c_int
    L.orc_kf_predict.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
    L.orc_kf_update.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double, _dp, _dp,
                                _dp]
    L.orc_kf_update.restype = C.c_int
    L.orc_kf_filter.argtypes = [C.POINTER(LGModel), C.c_int, _dp, C.c_double, _dp, _dp, _dp,
                                _dp]
    L.orc_kf_filter.restype = C.c_int
    L.orc_ptree_create.argtypes = [C.c_int64, C.c_int, C.c_int64, _dp]
    L.orc_ptree_create.restype = C.c_void_p
    L.orc_ptree_free.argtypes = [C.c_void_p]
    L.orc_ptree_insert.argtypes = [C.c_void_p, _dp, _ip]
    L.orc_ptree_ancestry.argtypes = [C.c_void_p, _dp]
    L.orc_ptree_size.argtypes = [C.c_void_p]
    L.orc_ptree_size.restype = C.c_int64
    L.orc_ptree_export.argtypes = [C.c_void_p, _ip, _ip, _ip]
    L.orc_rbpf_initialise.argtypes = [C.POINTER(RBModel), C.c_int64, _dp, _dp, _dp, _dp]
    L.orc_rbpf_step.argtypes = [C.POINTER(RBModel), C.c_int64, _dp, _dp, _dp, _dp,
                                C.POINTER(C.c_int), C.c_double, C.c_int, _dp, _dp, _dp, _dp,
                                _ip, _dp, C.POINTER(C.c_int), _dp]
    L.orc_rbpf_step.restype = C.c_int


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _colmajor(M):
    """Column-major FP64 buffer of a 2-D array (flattened Fortran order)."""
    return np.ascontiguousarray(np.asarray(M, dtype=np.float64).flatten(order="F"))


# --------------------------------------------------------------------------------------
# Philox
# --------------------------------------------------------------------------------------

def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(c.ctypes.data_as(_up), k.ctypes.data_as(_up),
                            out.ctypes.data_as(_up))
    return out


def philox_normals(seed, stream, t, N, D, f64_layout=False):
    """N x D standard normals exactly as the device draws them for (seed, stream, t)."""
    z = np.zeros((N, D), dtype=np.float64)
    lib().orc_philox_normals(C.c_uint64(seed), stream, t, N, D, int(f64_layout), _d(z))
    return z


def philox_uniforms(seed, t, n):
    u = np.zeros(n, dtype=np.float64)
    lib().orc_philox_uniforms(C.c_uint64(seed), t, n, _d(u))
    return u


def stream_uniform(seed, stream, t, k=0):
    """The [0, 1) uniform the device draws for (seed, stream, step t, index k): Philox
    counter (k, t, stream, 0), key (low, high word of seed), top 53 bits of words 0-1."""
    w = philox4x32_10([k, t, stream, 0], [seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF])
    v = ((int(w[0]) << 32) | int(w[1])) >> 11
    return v * (1.0 / 9007199254740992.0)


# --------------------------------------------------------------------------------------
# Resamplers
# --------------------------------------------------------------------------------------

def resample(kind, w, u):
    w = _f64(w)
    u = _f64(np.atleast_1d(u))
    n = w.shape[0]
    anc = np.zeros(n, dtype=np.int32)
    rc = lib().orc_resample(kind, n, _d(w), _d(u), _i(anc))
    if rc:
        raise ValueError(f"orc_resample rc={rc}")
    return anc


def offspring_to_ancestors(O):
    O = np.ascontiguousarray(O, dtype=np.int32)
    anc = np.zeros_like(O)
    lib().orc_offspring_to_ancestors(O.shape[0], _i(O), _i(anc))
    return anc


def ancestors_to_offspring(anc):
    anc = np.ascontiguousarray(anc, dtype=np.int32)
    off = np.zeros_like(anc)
    lib().orc_ancestors_to_offspring(anc.shape[0], _i(anc), _i(off))
    return off


def logsumexp(x):
    x = _f64(x)
    return lib().orc_logsumexp(x.shape[0], _d(x))


# --------------------------------------------------------------------------------------
# Linear-Gaussian models
# --------------------------------------------------------------------------------------

class LG:
    """Homogeneous LG SSM parameters (linear_gaussian.jl:96-131) as numpy arrays."""

    def __init__(self, mu0, Sigma0, A, b, Q, H, c, R):
        self.mu0 = _f64(mu0)
        self.Sigma0 = np.asarray(Sigma0, dtype=np.float64)
        self.A = np.asarray(A, dtype=np.float64)
        self.b = _f64(b)
        self.Q = np.asarray(Q, dtype=np.float64)
        self.H = np.asarray(H, dtype=np.float64)
        self.c = _f64(c)
        self.R = np.asarray(R, dtype=np.float64)
        self.Dx = self.A.shape[0]
        self.Dy = self.H.shape[0]

    def cstruct(self):
        keep = [_colmajor(M) for M in (self.Sigma0, self.A, self.Q, self.H, self.R)]
        m = LGModel(self.Dx, self.Dy, _d(self.mu0), _d(keep[0]), _d(keep[1]), _d(self.b),
                    _d(keep[2]), _d(keep[3]), _d(self.c), _d(keep[4]))
        return m, keep


def bf_initialise(model: LG, z):
    z = _f64(z)
    N = z.shape[0]
    x = np.zeros((N, model.Dx))
    m, keep = model.cstruct()
    rc = lib().orc_bf_initialise(C.byref(m), N, _d(z), _d(x))
    if rc:
        raise ValueError(f"orc_bf_initialise rc={rc}")
    return x


class BFState:
    """AoS oracle particle state: x (N x D), logw (N), typeless flag."""

    def __init__(self, x, logw=None, typeless=True):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        N = self.x.shape[0]
        self.logw = np.zeros(N) if logw is None else _f64(logw).copy()
        self.typeless = typeless


def bf_step(model: LG, state: BFState, y, threshold, kind, u, z, w_override=None):
    """One reference BF step in place; returns (anc, ll_inc, did_resample, w_used)."""
    N = state.x.shape[0]
    anc = np.zeros(N, dtype=np.int32)
    ll = C.c_double()
    did = C.c_int()
    typ = C.c_int(int(state.typeless))
    w_out = np.zeros(N)
    m, keep = model.cstruct()
    u = _f64(np.atleast_1d(u))
    z = _f64(z)
    y = _f64(np.atleast_1d(y))
    wo = None if w_override is None else _f64(w_override)
    rc = lib().orc_bf_step(C.byref(m), N, _d(state.x), _d(state.logw), C.byref(typ),
                           float(threshold), kind, _d(u), _d(z), _d(y), _d(wo), _i(anc),
                           C.byref(ll), C.byref(did), _d(w_out))
    if rc:
        raise ValueError(f"orc_bf_step rc={rc}")
    state.typeless = bool(typ.value)
    return anc, ll.value, bool(did.value), w_out


def apf_step(model: LG, state: BFState, y, threshold, kind, u, z, w_override=None):
    """One auxiliary-particle-filter step (MeanPredictive look-ahead) in place
    (orc_apf_step: particles.jl:248-344, resamplers.jl:92-143); returns
    (anc, ll_inc, did_resample, auxiliary weights used)."""
    N = state.x.shape[0]
    anc = np.zeros(N, dtype=np.int32)
    ll = C.c_double()
    did = C.c_int()
    typ = C.c_int(int(state.typeless))
    w_out = np.zeros(N)
    m, keep = model.cstruct()
    u = _f64(np.atleast_1d(u))
    z = _f64(z)
    y = _f64(np.atleast_1d(y))
    wo = None if w_override is None else _f64(w_override)
    rc = lib().orc_apf_step(C.byref(m), N, _d(state.x), _d(state.logw), C.byref(typ),
                            float(threshold), kind, _d(u), _d(z), _d(y), _d(wo), _i(anc),
                            C.byref(ll), C.byref(did), _d(w_out))
    if rc:
        raise ValueError(f"orc_apf_step rc={rc}")
    state.typeless = bool(typ.value)
    return anc, ll.value, bool(did.value), w_out


def csmc_step(model: LG, state: BFState, y, threshold, kind, u, z, x_ref_t, ref_anc,
              w_override=None):
    """One conditional-SMC step in place (orc_csmc_step): particle 0 takes ancestor ref_anc
    when the step resamples and moves to x_ref_t.  Returns (anc, ll_inc, did_resample,
    w_used)."""
    N = state.x.shape[0]
    anc = np.zeros(N, dtype=np.int32)
    ll = C.c_double()
    did = C.c_int()
    typ = C.c_int(int(state.typeless))
    w_out = np.zeros(N)
    m, keep = model.cstruct()
    u = _f64(np.atleast_1d(u))
    z = _f64(z)
    y = _f64(np.atleast_1d(y))
    xr = _f64(x_ref_t)
    wo = None if w_override is None else _f64(w_override)
    rc = lib().orc_csmc_step(C.byref(m), N, _d(state.x), _d(state.logw), C.byref(typ),
                             float(threshold), kind, _d(u), _d(z), _d(y), _d(wo), _d(xr),
                             int(ref_anc), _i(anc), C.byref(ll), C.byref(did), _d(w_out))
    if rc:
        raise ValueError(f"orc_csmc_step rc={rc}")
    state.typeless = bool(typ.value)
    return anc, ll.value, bool(did.value), w_out


def as_logweights(model: LG, state: BFState, x_ref):
    """Ancestor-sampling log-weights log w_i + log N(x_ref; A x_i + b, Q) (orc_as_logweights)."""
    N = state.x.shape[0]
    out = np.zeros(N)
    m, keep = model.cstruct()
    xr = _f64(x_ref)
    rc = lib().orc_as_logweights(C.byref(m), N, _d(state.x), _d(state.logw),
                                 int(state.typeless), _d(xr), _d(out))
    if rc:
        raise ValueError(f"orc_as_logweights rc={rc}")
    return out


def categorical(logw, u):
    """StatsBase.sample(Weights(softmax(logw))) with the uniform u: the multinomial rule
    (first j whose cumulative weight reaches u W_n) for one draw."""
    lw = _f64(logw)
    w = np.exp(lw - lw.max())
    return int(resample(MULTINOMIAL, w, np.full(lw.shape[0], float(u)))[0])


def backward_sample(model: LG, xs, logws, u, dynamics=None):
    """CSMC-BS's backward pass (csmc.jl:415-444): xs[t] (N, Dx) for t = 0..K, logws[t-1] the
    normalised log-weights of generation t, u[t] the uniform of time t.  Index K ~ the last
    weights; index t ~ softmax(log w_t + log N(x*_{t+1}; A x_t + b, Q)), the time-0 weights
    uniform.  dynamics (time-varying models): dynamics[t] is the model of step t + 1, whose
    transition weights time t (ancestor_weight(..., t + 1, ...) at csmc.jl:424); None uses
    `model` at every time.  Returns (indices (K+1,), trajectory (K+1, Dx))."""
    K = len(xs) - 1
    idx = np.zeros(K + 1, dtype=np.int64)
    idx[K] = categorical(logws[K - 1], u[K])
    for t in range(K - 1, -1, -1):
        st = BFState(np.asarray(xs[t], dtype=np.float64),
                     None if t == 0 else logws[t - 1], typeless=(t == 0))
        m_t = model if dynamics is None else dynamics[t]
        idx[t] = categorical(as_logweights(m_t, st, np.asarray(xs[t + 1])[idx[t + 1]]), u[t])
    traj = np.array([np.asarray(xs[t])[idx[t]] for t in range(K + 1)])
    return idx, traj


def csmc_sweep(model: LG, ys, N, x_ref, ancestor_sampling, threshold, kind, rng,
               backward=False):
    """One CSMC sweep (csmc.jl:239-306, :392-446 with backward=True) with numpy draws:
    returns (trajectory (K+1, Dx), log-likelihood).  x_ref None runs the unconditional
    filter (the first iteration)."""
    ys = np.asarray(ys, dtype=np.float64).reshape(len(ys), -1)
    K, D = ys.shape[0], model.Dx
    st = BFState(bf_initialise(model, rng.standard_normal((N, D))))
    if x_ref is not None:
        st.x[0] = x_ref[0]
    xs, ancs, lws = [st.x.copy()], [], []
    ll = 0.0
    for t in range(1, K + 1):
        u = rng.random(1 if kind in (SYSTEMATIC, SYSTEMATIC_MURRAY) else N)
        z = rng.standard_normal((N, D))
        if x_ref is None:
            anc, inc, _, _ = bf_step(model, st, ys[t - 1], threshold, kind, u, z)
        else:
            thr, a0 = threshold, 0
            if ancestor_sampling:
                a0 = categorical(as_logweights(model, st, x_ref[t]), rng.random())
                thr = np.inf
            anc, inc, _, _ = csmc_step(model, st, ys[t - 1], thr, kind, u, z, x_ref[t], a0)
        ll += inc
        xs.append(st.x.copy())
        ancs.append(anc)
        lws.append(st.logw.copy())
    if backward:
        u = rng.random(K + 1)[::-1]  # the last time's draw first
        return backward_sample(model, xs, lws, u)[1], ll
    j = categorical(st.logw, rng.random())
    return dense_ancestry(xs, ancs)[0][j], ll


def kf_filter(model: LG, ys, jitter=0.0):
    ys = np.ascontiguousarray(np.asarray(ys, dtype=np.float64).reshape(-1, model.Dy))
    T = ys.shape[0]
    mu = np.zeros(model.Dx)
    Sig = np.zeros(model.Dx * model.Dx)
    ll = C.c_double()
    lls = np.zeros(T)
    m, keep = model.cstruct()
    rc = lib().orc_kf_filter(C.byref(m), T, _d(ys), float(jitter), _d(mu), _d(Sig),
                             C.byref(ll), _d(lls))
    if rc:
        raise ValueError(f"orc_kf_filter rc={rc}")
    return mu, Sig.reshape(model.Dx, model.Dx, order="F"), ll.value, lls


def kf_step(Dx, Dy, mu, Sigma, A, b, Q, H, c, R, y, jitter=0.0):
    """One KF predict+update (kalman.jl:42-79); returns (mu, Sigma, ll)."""
    mu = _f64(mu).copy()
    S = _colmajor(Sigma)
    A_, Q_, H_, R_ = (_colmajor(M) for M in (A, Q, H, R))
    lib().orc_kf_predict(Dx, _d(A_), _d(_f64(b)), _d(Q_), _d(mu), _d(S))
    ll = C.c_double()
    rc = lib().orc_kf_update(Dx, Dy, _d(H_), _d(_f64(c)), _d(R_), _d(_f64(y)), float(jitter),
                             _d(mu), _d(S), C.byref(ll))
    if rc:
        raise ValueError(f"orc_kf_update rc={rc}")
    return mu, S.reshape(Dx, Dx, order="F"), ll.value


def simulate_lg(model: LG, T, rng):
    """Forward simulation (SSMProblems forward_simulation.jl:14-21) with numpy noise."""
    L0 = np.linalg.cholesky(model.Sigma0)
    LQ = np.linalg.cholesky(model.Q)
    LR = np.linalg.cholesky(model.R)
    x = model.mu0 + L0 @ rng.standard_normal(model.Dx)
    x0 = x.copy()
    xs, ys = [], []
    for _ in range(T):
        x = model.A @ x + model.b + LQ @ rng.standard_normal(model.Dx)
        xs.append(x.copy())
    for t in range(T):
        ys.append(model.H @ xs[t] + model.c + LR @ rng.standard_normal(model.Dy))
    return x0, np.array(xs), np.array(ys)


def joint_gaussian(model: LG, T):
    """Joint Gaussian of Z = [X0..XT, Y1..YT]; restates GFTest/models/linear_gaussian.jl:171-215."""
    rep = lambda M: [M] * T
    return joint_gaussian_tv(model.mu0, model.Sigma0, rep(model.A), rep(model.b), rep(model.Q),
                             rep(model.H), rep(model.c), rep(model.R))


def joint_gaussian_tv(mu0, Sigma0, As, bs, Qs, Hs, cs, Rs):
    """The same joint Gaussian with step-t parameters As[t-1] ... (t = 1..T)."""
    T = len(As)
    Dx, Dy = np.asarray(As[0]).shape[0], np.atleast_2d(Hs[0]).shape[0]
    n = Dx + T * (Dx + Dy)
    P = np.zeros((n, n))
    mu_e = np.zeros(n)
    S_e = np.zeros((n, n))
    mu_e[:Dx] = mu0
    S_e[:Dx, :Dx] = Sigma0
    for t in range(1, T + 1):
        iA, jA = t * Dx, (t - 1) * Dx
        P[iA:iA + Dx, jA:jA + Dx] = As[t - 1]
        iH, jH = Dx * (T + 1) + (t - 1) * Dy, Dx * t
        P[iH:iH + Dy, jH:jH + Dx] = Hs[t - 1]
        mu_e[iA:iA + Dx] = bs[t - 1]
        mu_e[iH:iH + Dy] = cs[t - 1]
        S_e[iA:iA + Dx, iA:iA + Dx] = Qs[t - 1]
        S_e[iH:iH + Dy, iH:iH + Dy] = Rs[t - 1]
    IP = np.eye(n) - P
    mu_Z = np.linalg.solve(IP, mu_e)
    Sig_Z = np.linalg.solve(IP, np.linalg.solve(IP, S_e).T).T
    return mu_Z, Sig_Z


def kf_loglikelihood_grad(mu0, Sigma0, As, bs, Qs, Hs, cs, Rs, ys, jitter=0.0):
    """The log-likelihood and its gradients w.r.t. mu0, Sigma0 and every step's A, b, Q, H,
    c, R: the reference's reverse rule (GF/src/integrations/kalman_rrule.jl:22-107) with the
    backward recursion of GF/src/algorithms/kalman_gradient.jl:103-263 (Parellier et al.),
    in FP64 numpy.  The forward pass is the Joseph-form filter with Symmetric(:U) and
    jitter (kalman_gradient.jl:12-29).  Returns (ll, grads) with grads[name] a list over
    steps (or one array for mu0 and Sigma0), in the log-likelihood convention."""
    T = len(As)
    ys = np.asarray(ys, dtype=np.float64).reshape(T, -1)
    mu = _f64(mu0).copy()
    S_ = np.asarray(Sigma0, dtype=np.float64).copy()
    ll = 0.0
    caches = []
    for t in range(T):
        A, b, Q = (np.asarray(v, dtype=np.float64) for v in (As[t], bs[t], Qs[t]))
        H, c, R = (np.atleast_2d(np.asarray(v, dtype=np.float64)) if k != 1 else
                   np.asarray(v, dtype=np.float64) for k, v in enumerate((Hs[t], cs[t], Rs[t])))
        mu_prev, S_prev = mu.copy(), S_.copy()
        mu_pred = A @ mu + b                                   # kalman.jl:42-49
        S_pred = A @ S_ @ A.T + Q
        z = ys[t] - H @ mu_pred - c                            # kalman_gradient.jl:12-29
        Sy = H @ S_pred @ H.T + R
        K = np.linalg.solve(Sy, H @ S_pred).T
        IKH = np.eye(len(mu)) - K @ H
        Sf = IKH @ S_pred @ IKH.T + K @ R @ K.T
        if jitter:
            Sf = Sf + jitter * np.eye(len(mu))
        Sf = np.triu(Sf) + np.triu(Sf, 1).T
        mu_f = mu_pred + K @ z
        L = np.linalg.cholesky(Sy)
        w = np.linalg.solve(L, z)
        ll += -0.5 * (len(z) * np.log(2 * np.pi) + 2 * np.log(np.diag(L)).sum() + w @ w)
        caches.append(dict(mu_prev=mu_prev, S_prev=S_prev, mu_pred=mu_pred, S_pred=S_pred,
                           mu_f=mu_f, z=z, S=Sy, K=K, IKH=IKH, A=A, H=H, R=R))
        mu, S_ = mu_f, Sf
    # backward pass (kalman_rrule.jl:63-103), NLL convention, negated at the end
    dmu = np.zeros_like(mu)
    dS = np.zeros_like(S_)
    g = {k: [None] * T for k in ("A", "b", "Q", "H", "c", "R")}
    for t in range(T - 1, -1, -1):
        cc = caches[t]
        z, S, K, IKH, H, R, A = cc["z"], cc["S"], cc["K"], cc["IKH"], cc["H"], cc["R"], cc["A"]
        Siz = np.linalg.solve(S, z)
        Sinv = np.linalg.inv(S)
        # gradient_c, kalman_gradient.jl:256-263
        g["c"][t] = -Siz - K.T @ dmu
        # gradient_H, kalman_gradient.jl:216-246
        P_f = IKH @ cc["S_pred"]
        dH = Sinv @ H @ cc["S_pred"] - np.outer(Siz, cc["mu_pred"]) \
            - np.outer(Siz, Siz) @ H @ cc["S_pred"]
        dH = dH + np.outer(Siz, dmu) @ P_f - np.outer(K.T @ dmu, cc["mu_f"])
        dH = dH - 2 * K.T @ dS @ P_f
        g["H"][t] = dH
        # gradient_R, kalman_gradient.jl:160-174
        cross = 0.5 * (np.outer(K.T @ dmu, Siz) + np.outer(Siz, K.T @ dmu))
        g["R"][t] = K.T @ dS @ K - cross + 0.5 * (Sinv - np.outer(Siz, Siz))
        # backward_gradient_update, kalman_gradient.jl:103-124
        dmu_pred = IKH.T @ dmu - H.T @ Siz
        Riz = np.linalg.solve(R, z)
        cross2 = 0.5 * (np.outer(dmu, H.T @ Riz) + np.outer(H.T @ Riz, dmu))
        dS_pred = IKH.T @ (dS + cross2) @ IKH \
            + 0.5 * (H.T @ Sinv @ H - H.T @ np.outer(Siz, Siz) @ H)
        # dynamics parameters, kalman_gradient.jl:196-208, :150-152
        g["b"][t] = dmu_pred
        g["A"][t] = np.outer(dmu_pred, cc["mu_prev"]) + 2 * dS_pred @ A @ cc["S_prev"]
        g["Q"][t] = dS_pred
        # backward_gradient_predict, kalman_gradient.jl:139-143
        dmu = A.T @ dmu_pred
        dS = A.T @ dS_pred @ A
    grads = {k: [-v for v in vals] for k, vals in g.items()}
    grads["mu0"] = -dmu
    grads["Sigma0"] = -dS
    return ll, grads


def kf_loglikelihood(mu0, Sigma0, As, bs, Qs, Hs, cs, Rs, ys, jitter=0.0):
    """kf_loglikelihood (GF/src/integrations/logdensity.jl:149-161): kalman_predict then the
    Joseph-form update with jitter, step parameters As[t] ..., summed increments."""
    mu = _f64(mu0)
    S = np.asarray(Sigma0, dtype=np.float64)
    ys = np.asarray(ys, dtype=np.float64).reshape(len(As), -1)
    Dx, Dy = mu.shape[0], ys.shape[1]
    ll = 0.0
    for t in range(len(As)):
        mu, S, inc = kf_step(Dx, Dy, mu, S, As[t], bs[t], Qs[t], Hs[t], cs[t], Rs[t], ys[t],
                             jitter)
        ll += inc
    return ll


def optimal_proposal(model: LG):
    """(F, g0, K, P) of GFTest's OptimalProposal (GF/src/GFTest/proposals.jl:11-31): the
    predicted state N(A x + b, Q) conditioned on y by the jitter-free Joseph-form Kalman
    update (kalman_gradient.jl:12-29), as an affine map x, y -> mean with covariance P."""
    A, b, Q = (np.atleast_2d(model.A), np.atleast_1d(model.b), np.atleast_2d(model.Q))
    H, c, R = (np.atleast_2d(model.H), np.atleast_1d(model.c), np.atleast_2d(model.R))
    S = H @ Q @ H.T + R
    K = np.linalg.solve(S, H @ Q).T                            # Q H' S^-1
    IKH = np.eye(len(b)) - K @ H
    P = IKH @ Q @ IKH.T + K @ R @ K.T                          # Joseph form
    P = np.triu(P) + np.triu(P, 1).T                           # Symmetric(:U)
    return IKH @ A, IKH @ b - K @ c, K, P


def overdispersed_proposal(model: LG, k):
    """(F, g0, K, P) of GFTest's OverdispersedProposal (proposals.jl:33-49): N(A x + b, k Q)."""
    return (np.atleast_2d(model.A), np.atleast_1d(model.b), None,
            k * np.atleast_2d(model.Q))


def _mvn_logpdf_rows(X, M, S):
    """log N(X[i]; M[i], S) for every row, by S's Cholesky factor."""
    L = np.linalg.cholesky(0.5 * (S + S.T))
    Rz = np.linalg.solve(L, (X - M).T)
    D = S.shape[0]
    return -0.5 * (D * np.log(2 * np.pi) + 2 * np.log(np.diag(L)).sum() + (Rz * Rz).sum(0))


def guided_step(model: LG, state: BFState, y, threshold, kind, u, z, prop, w_override=None):
    """One step of ParticleFilter(N, proposal) in place (particles.jl:147-162): maybe_resample
    as ``bf_step``, then ``propogate`` (particles.jl:175-195) draws x' = F x + g0 + K y + L_P z
    from the proposal and adds log f(x' | x) - log q(x' | x, y), then the observation's
    log-density, and normalises the log-weights (marginalise!).  prop = (F, g0, K or None, P).  Returns (anc, ll_inc, did_resample, w_used)."""
    N, D = state.x.shape
    y = np.atleast_1d(_f64(y))
    if state.typeless:
        w = np.full(N, 1.0 / N)
    else:
        w = np.exp(state.logw - logsumexp(state.logw))
    did = threshold * N >= 1.0 / np.cumsum(w * w)[-1]  # summed left to right, as bf_step
    if did:
        anc = resample(kind, w if w_override is None else w_override, u)
        state.x = state.x[anc].copy()
        state.logw = np.zeros(N)
        baseline = logsumexp(state.logw)
    else:
        anc = np.arange(N, dtype=np.int32)
        if state.typeless:
            baseline = np.log(N)
            state.logw = np.zeros(N)
        else:
            baseline = logsumexp(state.logw)
    state.typeless = False
    F, g0, K, P = prop
    A, b, Q = (np.atleast_2d(model.A), np.atleast_1d(model.b), np.atleast_2d(model.Q))
    H, c, R = (np.atleast_2d(model.H), np.atleast_1d(model.c), np.atleast_2d(model.R))
    g = np.atleast_1d(g0) + (0.0 if K is None else np.atleast_2d(K) @ y)
    Mq = state.x @ np.atleast_2d(F).T + g
    xn = Mq + _f64(z).reshape(N, D) @ np.linalg.cholesky(np.atleast_2d(P)).T
    log_p = _mvn_logpdf_rows(xn, state.x @ A.T + b, Q)          # logdensity(dyn, x, x')
    log_q = _mvn_logpdf_rows(xn, Mq, np.atleast_2d(P))          # logdensity(proposal, ...)
    state.x = xn
    state.logw = state.logw + (log_p - log_q)
    state.logw = state.logw + _mvn_logpdf_rows(np.broadcast_to(y, (N, len(y))), xn @ H.T + c, R)
    lse = logsumexp(state.logw)                                  # marginalise!
    state.logw = state.logw - lse
    return anc, float(lse - baseline), bool(did), w


def _qr_r_positive(M):
    """R of a QR factorisation with its rows made to have a nonnegative diagonal
    (_correct_cholesky_sign, srkf.jl:101)."""
    R = np.linalg.qr(M, mode="r")
    return np.sign(np.diag(R))[:, None] * R


def srkf_filter(mu0, Sigma0, As, bs, Qs, Hs, cs, Rs, ys, jitter=None):
    """filter(model, SRKF(jitter), ys) with per-step parameters (srkf.jl:93-176) in FP64
    numpy.  Sigma = U' U with U upper triangular.  Returns (ll, mu, U) after the last step."""
    T = len(As)
    ys = np.asarray(ys, dtype=np.float64).reshape(T, -1)
    mu = _f64(mu0).copy()
    U = np.linalg.cholesky(np.atleast_2d(_f64(Sigma0))).T
    ll = 0.0
    for t in range(T):
        A, Q = np.atleast_2d(_f64(As[t])), np.atleast_2d(_f64(Qs[t]))
        H, R = np.atleast_2d(_f64(Hs[t])), np.atleast_2d(_f64(Rs[t]))
        b, c = np.atleast_1d(_f64(bs[t])), np.atleast_1d(_f64(cs[t]))
        Dx, Dy = A.shape[0], H.shape[0]
        mu = A @ mu + b                                             # srkf_predict
        U = _qr_r_positive(np.vstack([np.linalg.cholesky(Q).T, U @ A.T]))[:Dx]
        M = np.zeros((Dy + Dx, Dy + Dx))                            # _srkf_form_update_matrix
        M[:Dy, :Dy] = np.linalg.cholesky(R).T
        M[Dy:, :Dy] = U @ H.T
        M[Dy:, Dy:] = U
        Rm = _qr_r_positive(M)
        U_S, PHt, U = Rm[:Dy, :Dy], Rm[:Dy, Dy:].T, Rm[Dy:, Dy:]
        z = ys[t] - H @ mu - c
        w = np.linalg.solve(U_S.T, z)
        mu = mu + PHt @ w
        ll += -0.5 * (w @ w + 2 * np.log(np.diag(U_S)).sum() + Dy * np.log(2 * np.pi))
        if jitter is not None:
            U = U + jitter * np.eye(Dx)
    return ll, mu, U


def mvn_logpdf(y, mu, S):
    S = 0.5 * (S + S.T)
    L = np.linalg.cholesky(S)
    r = np.linalg.solve(L, y - mu)
    return -0.5 * (len(y) * np.log(2 * np.pi) + 2 * np.log(np.diag(L)).sum() + r @ r)


def joint_gaussian_loglik(model: LG, ys):
    """Exact log p(y_1:T) from the joint Gaussian (the reference KF test's oracle)."""
    ys = np.asarray(ys, dtype=np.float64).reshape(-1, model.Dy)
    T = ys.shape[0]
    mu_Z, Sig_Z = joint_gaussian(model, T)
    iy = slice(model.Dx * (T + 1), None)
    return mvn_logpdf(ys.reshape(-1), mu_Z[iy], Sig_Z[iy, iy])


def joint_posterior(model: LG, ys):
    """Mean (T+1, Dx) and covariance of X0..XT given y_1:T, by Gaussian conditioning of the
    joint Gaussian (the smoothing distribution the reference's CSMC tests compare with)."""
    ys = np.asarray(ys, dtype=np.float64).reshape(-1, model.Dy)
    T = ys.shape[0]
    mu_Z, Sig_Z = joint_gaussian(model, T)
    nx = model.Dx * (T + 1)
    Sxx, Sxy, Syy = Sig_Z[:nx, :nx], Sig_Z[:nx, nx:], Sig_Z[nx:, nx:]
    K = np.linalg.solve(Syy, Sxy.T).T
    mean = mu_Z[:nx] + K @ (ys.reshape(-1) - mu_Z[nx:])
    cov = Sxx - K @ Sxy.T
    return mean.reshape(T + 1, model.Dx), cov


# --------------------------------------------------------------------------------------
# RBPF
# --------------------------------------------------------------------------------------

class RB:
    """Hierarchical model: scalar outer AR(1) + D-dim inner LG (see oracle.c orc_rb_model)."""

    def __init__(self, om0, os0, oa, ob, oq, mu0, Sigma0, A, b, Cv, Q, H, c, R0, rs):
        self.om0, self.os0, self.oa, self.ob, self.oq = map(float, (om0, os0, oa, ob, oq))
        self.mu0, self.b, self.C, self.c = _f64(mu0), _f64(b), _f64(Cv), _f64(c)
        self.Sigma0, self.A, self.Q, self.H, self.R0 = (
            np.asarray(M, dtype=np.float64) for M in (Sigma0, A, Q, H, R0))
        self.D = self.A.shape[0]
        self.Dy = self.H.shape[0]
        self.rs = float(rs)

    def cstruct(self):
        keep = [_colmajor(M) for M in (self.Sigma0, self.A, self.Q, self.H, self.R0)]
        m = RBModel(self.om0, self.os0, self.oa, self.ob, self.oq, self.D, self.Dy, self.rs,
                    _d(self.mu0), _d(keep[0]), _d(keep[1]), _d(self.b), _d(self.C),
                    _d(keep[2]), _d(keep[3]), _d(self.c), _d(keep[4]))
        return m, keep


class RBState:
    def __init__(self, xo, mu, Sigma, logw=None, typeless=True):
        self.xo = _f64(xo).copy()
        N = self.xo.shape[0]
        self.mu = np.ascontiguousarray(mu, dtype=np.float64).copy()      # N x D
        self.Sigma = np.ascontiguousarray(Sigma, dtype=np.float64).copy()  # N x D x D (col-major per particle)
        self.logw = np.zeros(N) if logw is None else _f64(logw).copy()
        self.typeless = typeless


def rbpf_initialise(model: RB, zo):
    zo = _f64(zo)
    N = zo.shape[0]
    xo = np.zeros(N)
    mu = np.zeros((N, model.D))
    Sig = np.zeros((N, model.D, model.D))
    m, keep = model.cstruct()
    lib().orc_rbpf_initialise(C.byref(m), N, _d(zo), _d(xo), _d(mu), _d(Sig))
    return RBState(xo, mu, Sig)


def rbpf_step(model: RB, st: RBState, y, threshold, kind, u, zo, w_override=None):
    N = st.xo.shape[0]
    anc = np.zeros(N, dtype=np.int32)
    ll = C.c_double()
    did = C.c_int()
    typ = C.c_int(int(st.typeless))
    w_out = np.zeros(N)
    m, keep = model.cstruct()
    u = _f64(np.atleast_1d(u))
    wo = None if w_override is None else _f64(w_override)
    rc = lib().orc_rbpf_step(C.byref(m), N, _d(st.xo), _d(st.mu), _d(st.Sigma), _d(st.logw),
                             C.byref(typ), float(threshold), kind, _d(u), _d(_f64(zo)),
                             _d(_f64(np.atleast_1d(y))), _d(wo), _i(anc), C.byref(ll),
                             C.byref(did), _d(w_out))
    if rc:
        raise ValueError(f"orc_rbpf_step rc={rc}")
    st.typeless = bool(typ.value)
    return anc, ll.value, bool(did.value), w_out


class PTree:
    """Sequential ParallelParticleTree (CUDAExt.jl:146-262): roots at x0 (N, D); ``insert``
    appends a generation x (N, D) with 0-based ancestors; ``ancestry`` -> (N, depth+1, D)."""

    def __init__(self, x0, capacity):
        x0 = _f64(x0)
        self.N, self.D = x0.shape
        self.depth = 0
        self.ptr = lib().orc_ptree_create(self.N, self.D, int(capacity), _d(x0))
        if not self.ptr:
            raise ValueError("capacity below N")

    def insert(self, x, anc):
        x = _f64(x).reshape(self.N, self.D)
        a = np.ascontiguousarray(np.asarray(anc, dtype=np.int32))
        grew = lib().orc_ptree_insert(self.ptr, _d(x), _i(a))
        self.depth += 1
        return grew

    def ancestry(self):
        out = np.zeros((self.N, self.depth + 1, self.D))
        lib().orc_ptree_ancestry(self.ptr, _d(out))
        return out

    def export(self):
        M = lib().orc_ptree_size(self.ptr)
        par = np.zeros(M, dtype=np.int32)
        off = np.zeros(M, dtype=np.int32)
        leaves = np.zeros(self.N, dtype=np.int32)
        lib().orc_ptree_export(self.ptr, _i(par), _i(off), _i(leaves))
        return par, off, leaves

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_ptree_free(self.ptr)
            self.ptr = None


def dense_ancestry(xs, ancs):
    """Paths by walking the stored ancestors back from the last generation, as
    DenseParticleContainer's get_ancestry (containers.jl:406-418).  xs[t]: (N, D) states,
    t = 0..T; ancs[t]: 0-based ancestors of generation t+1.  Returns the paths (N, T+1, D)
    and the number of distinct (time, particle) pairs on the surviving lines."""
    T = len(ancs)
    N, D = np.asarray(xs[0]).shape
    idx = np.arange(N)
    paths = np.zeros((N, T + 1, D))
    distinct = 0
    for t in range(T, -1, -1):
        paths[:, t, :] = np.asarray(xs[t])[idx]
        distinct += len(np.unique(idx))
        if t > 0:
            idx = np.asarray(ancs[t - 1])[idx]
    return paths, distinct


# ---- the all-cores CPU baseline (oracle/bf_omp.c) ----------------------------------------
class _BfOmpModel(C.Structure):
    _fields_ = [("D", C.c_int), ("Dy", C.c_int), ("A", C.c_double * 256), ("b", C.c_double * 16),
                ("LQ", C.c_double * 256), ("H", C.c_double * 256), ("c", C.c_double * 16),
                ("LRi", C.c_double * 256), ("cst", C.c_double)]


_omp = None


def _omp_lib():
    global _omp
    if _omp is None:
        _omp = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbfomp.so"))
        _omp.bfomp_run.argtypes = [C.POINTER(_BfOmpModel), C.c_int64, C.c_int, _dp, C.c_double,
                                   C.c_uint64, C.c_int, _dp, _dp, C.c_int, _dp]
        _omp.bfomp_run.restype = C.c_int
    return _omp


def bfomp_run(model: LG, x0, ys, threshold, seed, nthreads):
    """T systematic BF steps from the states x0 (N, Dx) with OpenMP over nthreads threads,
    FP64 SoA (bf_omp.c, the all-cores CPU baseline).  Returns (ll, resampled steps, x)."""
    x0 = np.asarray(x0, dtype=np.float64)
    N, D = x0.shape
    ys = np.ascontiguousarray(np.asarray(ys, dtype=np.float64).reshape(len(ys), -1))
    Dy = ys.shape[1]
    m = _BfOmpModel()
    m.D, m.Dy = D, Dy
    fill = lambda dst, M: [dst.__setitem__(k, float(v)) for k, v in
                           enumerate(np.asarray(M, dtype=np.float64).flatten(order="F"))]
    fill(m.A, model.A)
    fill(m.b, model.b)
    fill(m.LQ, np.linalg.cholesky(model.Q))
    fill(m.H, model.H)
    fill(m.c, model.c)
    LR = np.linalg.cholesky(model.R)
    fill(m.LRi, np.linalg.inv(LR))
    m.cst = -0.5 * (Dy * np.log(2 * np.pi) + 2 * np.log(np.diag(LR)).sum())
    x = np.ascontiguousarray(x0.T.copy())  # SoA planes
    lw = np.zeros(N)
    ll = C.c_double()
    n_rs = _omp_lib().bfomp_run(C.byref(m), N, ys.shape[0], _d(ys), float(threshold),
                                int(seed), int(nthreads), _d(x), _d(lw), 1, C.byref(ll))
    return ll.value, n_rs, x.T.copy()
