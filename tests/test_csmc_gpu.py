"""Device conditional SMC (GF/src/algorithms/csmc.jl) against the FP64 oracle.

Teacher-forced per step as the bootstrap-filter tests are: the oracle's CSMC step starts
from the state the device exported, with the same noise, uniforms and resampler weights,
and with the device's choice of the reference's ancestor.  That choice is checked on its
own: 0 for NoRefreshment, and for AncestorSampling the oracle's categorical draw from its
ancestor-sampling weights with the device's Philox uniform (bit-exact in FP64; in FP32 a
different index is accepted only when the weight between the two choices is below 1e-5 of
the total, the FP32 rounding of the weights).  Tolerances are the bootstrap filter's.
The sweeps are checked against the exact smoothing distribution as the reference's own
CSMC tests are (test/algorithms/csmc.jl:5-95).
"""
import numpy as np
import pytest

import ssmhip
from conftest import config2_model, random_lg, to_ssmhip_model
from test_bf_gpu import TOL

pytestmark = pytest.mark.gpu

DEV_SEED = 0x1234ABCD5678


def _between(w, a, b):
    lo, hi = sorted((int(a), int(b)))
    return w[lo + 1:hi].sum() if hi > lo else 0.0


def _teacher_forced_csmc(ctx, orc, m, N, prec, mode, thresholds, seed=1):
    rng = np.random.default_rng(seed)
    model = to_ssmhip_model(m)
    algo = ssmhip.BF(N, threshold=1.0, resampler=ssmhip.Systematic(), precision=prec)
    K = len(thresholds)
    _, _, ys = orc.simulate_lg(m, K, rng)
    x_ref = rng.standard_normal((K + 1, m.Dx))
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=DEV_SEED)
    st.set_reference(x_ref, mode)
    z0 = np.ascontiguousarray(rng.standard_normal((N, m.Dx)))
    ctx.check(ctx.lib.ssm_bf_initialise(st.ptr, z0.ctypes.data_as(ssmhip._lib._dp)))
    cast = np.float32 if prec == "f32" else np.float64
    x_gpu, lw_gpu, _, _ = st.export()
    # initialise with ref_state: particle 0 is the reference's time-0 state
    np.testing.assert_array_equal(x_gpu[0], x_ref[0].astype(cast).astype(np.float64))
    np.testing.assert_allclose(x_gpu[1:], orc.bf_initialise(m, z0)[1:], rtol=2e-6, atol=2e-6)
    typeless = True
    tol = TOL[prec]
    ancestor = mode == ssmhip._lib.SSM_CSMC_ANCESTOR
    for t, thr in enumerate(thresholds, start=1):
        ost = orc.BFState(x_gpu.copy(), lw_gpu.copy(), typeless)
        if ancestor:
            thr = np.inf  # ancestor sampling resamples at every step
            lw_as = orc.as_logweights(m, ost, x_ref[t])
            a_o = orc.categorical(lw_as, orc.stream_uniform(DEV_SEED, orc.STREAM_ANCESTOR, t))
        algo.threshold = thr
        u = rng.random(1)
        z = rng.standard_normal((N, m.Dx))
        _, ll = ssmhip.step(rng, model, algo, t, st, ys[t - 1], u=u, z=z)
        x_new, lw_new, anc_gpu, w_gpu = st.export()
        did = st.last_resampled
        if ancestor:
            assert did, f"t={t}: ancestor sampling must resample"
            if anc_gpu[0] != a_o:
                w_as = np.exp(lw_as - lw_as.max())
                assert prec == "f32" and _between(w_as, anc_gpu[0], a_o) <= 1e-5 * w_as.sum(), \
                    f"t={t}: reference ancestor {anc_gpu[0]} vs oracle {a_o}"
        elif did:
            assert anc_gpu[0] == 0, f"t={t}: reference ancestor {anc_gpu[0]}"
        ref_anc = int(anc_gpu[0]) if did else 0
        anc_o, ll_o, did_o, _ = orc.csmc_step(m, ost, ys[t - 1], thr, 0, u, z, x_ref[t], ref_anc,
                                              w_override=w_gpu if did else None)
        assert did == did_o, f"t={t} resample decision"
        if did:  # (the exported ancestors are those of the last resample)
            np.testing.assert_array_equal(anc_gpu, anc_o)
        np.testing.assert_array_equal(x_new[0], x_ref[t].astype(cast).astype(np.float64))
        np.testing.assert_allclose(x_new, ost.x, rtol=tol["x"][0], atol=tol["x"][1])
        live = ost.logw > -30.0
        a_lw = max(tol["lw"][0], 2e-5 * m.Dx * m.Dy) if prec == "f32" else tol["lw"][0]
        r_lw = max(tol["lw"][2], 4e-6 * m.Dx * m.Dy) if prec == "f32" else tol["lw"][2]
        np.testing.assert_allclose(lw_new[live], ost.logw[live], atol=a_lw, rtol=r_lw)
        r_ll = max(1e-7, 4e-6 * m.Dx * m.Dy) if prec == "f32" else 1e-12
        assert ll == pytest.approx(ll_o, abs=tol["ll"], rel=r_ll), f"t={t}"
        x_gpu, lw_gpu, typeless = x_new, lw_new, False
    return st


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_teacher_forced_csmc_config2(ctx, orc, mode, prec):
    # D = 4, Dy = 2: the specialised step kernel; thresholds mix kept and resampled steps
    _teacher_forced_csmc(ctx, orc, config2_model(), 3000, prec, mode,
                         [1.0, 0.0, 0.5, 1.0, 0.0, 0.8, 1.0, 1.0])


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("Dx,Dy,N", [(1, 1, 10), (3, 2, 700), (7, 5, 5000)])
def test_teacher_forced_csmc_generic_dims(ctx, orc, mode, Dx, Dy, N):
    m = random_lg(np.random.default_rng(Dx * 10 + Dy), Dx, Dy)
    _teacher_forced_csmc(ctx, orc, m, N, "f64", mode, [1.0, 0.6, 1.0, 0.0, 1.0])


def test_teacher_forced_csmc_many_tiles(ctx, orc):
    # 300000 particles: many tiles in the ancestor draw's fold and tile search
    _teacher_forced_csmc(ctx, orc, config2_model(), 300000, "f64", 1, [1.0, 1.0, 1.0])


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_sample_index_matches_oracle(ctx, orc, prec):
    m = config2_model()
    model = to_ssmhip_model(m)
    rng = np.random.default_rng(7)
    N = 70000
    algo = ssmhip.BF(N, threshold=0.5, precision="f64")
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=5)
    # after initialise the weights are the lazy uniform ones: u in the middle of
    # [j / N, (j + 1) / N) draws j
    for j in (0, 1, 12345, N - 1):
        assert st.sample_index((j + 0.5) / N) == j
    assert st.sample_index(0.0) == 0
    _, _, ys = orc.simulate_lg(m, 3, rng)
    for t in range(1, 4):
        ssmhip.step(rng, model, algo, t, st, ys[t - 1])
    lw = st.export()[1]
    w = np.exp(lw - lw.max())
    for u in np.concatenate([[0.0, 0.5, 1.0 - 1e-12], rng.random(30)]):
        got, want = st.sample_index(u), orc.categorical(lw, u)
        if got != want:  # a rounding tie of the two cumulative sums
            assert prec == "f32" and _between(w, got, want) <= 1e-6 * w.sum(), (u, got, want)
    with pytest.raises(ssmhip.SSMError):
        st.sample_index(1.0)


@pytest.mark.parametrize("mode", [0, 1])
def test_run_equals_steps_with_reference(ctx, orc, mode):
    # the graph-replayed run (chunks of 16 steps) and the per-step calls launch the same
    # kernels with the same Philox keys: identical increments and states
    m = config2_model()
    model = to_ssmhip_model(m)
    T = 37
    rng = np.random.default_rng(3)
    ys = rng.standard_normal((T, 2))
    x_ref = rng.standard_normal((T + 1, 4))
    algo = ssmhip.BF(20000, threshold=0.7, precision="f32")
    out = []
    for path in ("run", "steps"):
        st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=99)
        st.set_reference(x_ref, mode)
        ctx.check(ctx.lib.ssm_bf_initialise(st.ptr, None))
        if path == "run":
            tot, incs = ssmhip.run(st, algo, ys, return_steps=True)
        else:
            incs = np.array([ssmhip.step(rng, model, algo, t, st, ys[t - 1])[1]
                             for t in range(1, T + 1)])
        out.append((incs, st.export()[:3]))
    assert np.array_equal(out[0][0], out[1][0])
    for fa, fb in zip(out[0][1], out[1][1]):
        assert np.array_equal(fa, fb)
    assert np.array_equal(out[0][1][0][0], x_ref[T].astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("refresh", ["plain", "ancestor", "backward"])
def test_csmc_chain_matches_smoothing_mean(ctx, orc, refresh):
    # the reference's CSMC tests (Dx = Dy = 1, K = 10, N = 10, threshold 0.6): the chain's
    # mean of every x_t against the exact smoothing mean (the reference checks t = 2 only)
    m = orc.LG([0.2], [[1.0]], [[0.9]], [0.1], [[0.5]], [[1.2]], [-0.1], [[0.4]])
    ys = np.random.default_rng(11).standard_normal((10, 1)) + 0.5
    mean, cov = orc.joint_posterior(m, ys)
    model = ssmhip.CSMCModel(to_ssmhip_model(m), ys)
    refreshment = {"plain": ssmhip.NoRefreshment(), "ancestor": ssmhip.AncestorSampling(),
                   "backward": ssmhip.BackwardSimulation()}[refresh]
    csmc = ssmhip.ConditionalSMC(ssmhip.BF(10, threshold=0.6), refreshment)
    rng = np.random.default_rng(12)
    state = None
    trajs, lls = [], []
    for i in range(200 + 6000):
        _, state = ssmhip.mcmc_step(rng, model, csmc, state, ctx=ctx)
        if i >= 200:
            trajs.append(state.trajectory.as_array()[:, 0])
            lls.append(state.log_likelihood)
    trajs = np.array(trajs)
    assert trajs.shape == (6000, 11)
    bm = trajs.reshape(40, -1, 11).mean(axis=1)
    se = bm.std(axis=0, ddof=1) / np.sqrt(40)
    sd = np.sqrt(np.diag(cov))
    assert np.all(se < 0.05 * sd), (se, sd)
    z = (trajs.mean(axis=0) - mean[:, 0]) / se
    assert np.all(np.abs(z) < 5), z
    # 1/Z-hat is unbiased for 1/Z
    L = -np.array(lls)
    e = np.exp(L - L.max())
    be = e.reshape(40, -1).mean(axis=1)
    rel_se = be.std(ddof=1) / np.sqrt(40) / e.mean()
    ll_exact = orc.joint_gaussian_loglik(m, ys)
    assert rel_se < 0.1
    assert abs(L.max() + np.log(e.mean()) + ll_exact) < 5 * rel_se


@pytest.mark.parametrize("Dx,Dy,N,K", [(4, 2, 20000, 12), (3, 2, 777, 9), (1, 1, 10, 10)])
def test_backward_sample_matches_oracle(ctx, orc, Dx, Dy, N, K):
    # CSMC-BS's backward pass over the device history (FP64) against the oracle's on the
    # exported generations with the same uniforms: the same indices, and the trajectory is
    # the stored states at those indices
    rng = np.random.default_rng(40 + Dx)
    m = config2_model() if Dx == 4 else random_lg(rng, Dx, Dy)
    if Dx != 4:
        m.A = 0.9 * m.A / max(abs(np.linalg.eigvals(m.A)))
    model = to_ssmhip_model(m)
    algo = ssmhip.BF(N, threshold=0.6, precision="f64")
    _, _, ys = orc.simulate_lg(m, K, rng)
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=9)
    st.set_reference(rng.standard_normal((K + 1, Dx)), 0)
    ctx.check(ctx.lib.ssm_bf_initialise(st.ptr, None))
    h = ssmhip.DenseParticleContainer(st, capacity=4)
    h.record(True)
    ssmhip.run(st, algo, ys)
    assert len(h) == K
    xs = [h.generation(t)[0] for t in range(K + 1)]
    lws = [h.generation(t)[1] for t in range(1, K + 1)]
    for trial in range(3):
        u = rng.random(K + 1)
        idx, traj = h.backward_sample(u)
        idx_o, traj_o = orc.backward_sample(m, xs, lws, u)
        np.testing.assert_array_equal(idx, idx_o)
        np.testing.assert_array_equal(traj, traj_o)
    with pytest.raises(ssmhip.SSMError):
        h.backward_sample(np.full(K + 1, 1.0))


def test_reference_errors(ctx):
    m = config2_model()
    model = to_ssmhip_model(m)
    rng = np.random.default_rng(0)
    algo = ssmhip.BF(64)
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=1)
    with pytest.raises(ssmhip.SSMError):
        st.set_reference(np.zeros((3, 4)), 7)  # unknown refreshment
    st.set_reference(np.zeros((3, 4)), 0)
    ctx.check(ctx.lib.ssm_bf_initialise(st.ptr, None))
    for t in (1, 2):
        ssmhip.step(rng, model, algo, t, st, np.zeros(2))
    with pytest.raises(ssmhip.SSMError):  # past the reference's last state
        ssmhip.step(rng, model, algo, 3, st, np.zeros(2))
    with pytest.raises(ssmhip.SSMError):
        ssmhip.run(st, algo, np.zeros((5, 2)), t0=1)
    st.set_reference(None)  # back to the unconditional filter
    ssmhip.step(rng, model, algo, 3, st, np.zeros(2))


class _TimeVaryingDynamics(ssmhip.models.LinearGaussianLatentDynamics):
    """A_t, b_t and Q_t that change every step (calc_params(dyn, t))."""

    def __init__(self, A, b, Q):
        self.A0, self.b0, self.Q0 = A, b, Q

    def calc_A(self, step, **kw):
        return self.A0 * (1.0 + 0.3 * np.sin(step))

    def calc_b(self, step, **kw):
        return self.b0 + 0.2 * step

    def calc_Q(self, step, **kw):
        return self.Q0 * (1.0 + 0.5 * (step % 3))


def test_backward_sample_time_varying_matches_oracle(ctx, orc):
    # Backward simulation on time-varying dynamics: time t is weighted with the dynamics of
    # step t + 1 (csmc.jl:424), against the oracle given the same per-step models.
    rng = np.random.default_rng(77)
    Dx, Dy, N, K = 3, 2, 3000, 8
    m = random_lg(rng, Dx, Dy)
    m.A = 0.9 * m.A / max(abs(np.linalg.eigvals(m.A)))
    dyn = _TimeVaryingDynamics(m.A, m.b, m.Q)
    obs = ssmhip.models.HomogeneousLinearGaussianObservationProcess(m.H, m.c, m.R)
    model = ssmhip.models.StateSpaceModel(ssmhip.models.HomogeneousGaussianPrior(m.mu0, m.Sigma0),
                                          dyn, obs)
    assert not model.is_homogeneous
    ys = rng.standard_normal((K, Dy))
    algo = ssmhip.BF(N, threshold=0.5, precision="f64")
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=5)
    h = ssmhip.DenseParticleContainer(st, capacity=4)
    h.record(True)
    for t in range(1, K + 1):
        ssmhip.step(rng, model, algo, t, st, ys[t - 1])
    h.record(False)
    xs = [h.generation(t)[0] for t in range(K + 1)]
    lws = [h.generation(t)[1] for t in range(1, K + 1)]
    packed = [ssmhip.models.PackedLG.from_model(model, t + 1) for t in range(K)]
    dyn_o = [orc.LG(m.mu0, m.Sigma0, dyn.calc_A(t + 1), dyn.calc_b(t + 1), dyn.calc_Q(t + 1),
                    m.H, m.c, m.R) for t in range(K)]
    for trial in range(3):
        u = rng.random(K + 1)
        idx, traj = h.backward_sample(u, packed)
        idx_o, traj_o = orc.backward_sample(m, xs, lws, u, dynamics=dyn_o)
        np.testing.assert_array_equal(idx, idx_o)
        np.testing.assert_array_equal(traj, traj_o)
    with pytest.raises(ValueError):
        h.backward_sample(rng.random(K + 1), packed[:-1])


def test_csmc_backward_simulation_time_varying_runs(ctx):
    # ConditionalSMC with BackwardSimulation on a non-homogeneous model: the sweep draws its
    # trajectory with the per-step dynamics (no error, a path of the right shape)
    rng = np.random.default_rng(8)
    m = config2_model()
    dyn = _TimeVaryingDynamics(m.A, m.b, m.Q)
    model = ssmhip.models.StateSpaceModel(
        ssmhip.models.HomogeneousGaussianPrior(m.mu0, m.Sigma0), dyn,
        ssmhip.models.HomogeneousLinearGaussianObservationProcess(m.H, m.c, m.R))
    ys = rng.standard_normal((6, 2))
    csmc = ssmhip.ConditionalSMC(ssmhip.BF(512, threshold=0.5, precision="f64"),
                                 ssmhip.BackwardSimulation())
    traj, ll = ssmhip.csmc_sample(rng, model, csmc, ys, None, ctx=ctx)
    assert np.isfinite(ll)
    traj2, ll2 = ssmhip.csmc_sample(rng, model, csmc, ys, traj, ctx=ctx)
    assert np.isfinite(ll2) and traj2.as_array().shape == (7, 4)
