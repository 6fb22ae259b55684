"""Device guided particle filter (ParticleFilter with a proposal, GF/src/algorithms/
particles.jl:92-195) against the FP64 oracle's ``guided_step``.

Teacher-forced per step as the bootstrap filter's tests: the oracle starts from the state the
device exported, with the same noise and uniforms, and its resampler takes the weights the
device consumed, so ancestors match bit for bit.  Tolerances: FP32 path, states 3e-5
relative and absolute, live log-weights (oracle > -30) within 5e-4 + 5e-4 |logw|,
log-likelihood increments 2e-4 -- the bootstrap filter's plus the FP32 rounding of the two
extra quadratic forms of log f - log q; FP64 path 1e-9."""
import numpy as np
import pytest

import ssmhip
from conftest import config2_model, random_lg, to_ssmhip_model

pytestmark = pytest.mark.gpu

TOL = {"f32": dict(x=3e-5, lw=5e-4, ll=2e-4), "f64": dict(x=1e-10, lw=1e-9, ll=1e-9)}


def _proposals(orc, m, model):
    return {"optimal": (ssmhip.OptimalProposal(model.dyn, model.obs), orc.optimal_proposal(m)),
            "overdispersed": (ssmhip.OverdispersedProposal(model.dyn, 1.5),
                              orc.overdispersed_proposal(m, 1.5))}


def _teacher_forced(ctx, orc, m, N, prec, which, thresholds, seed=1):
    rng = np.random.default_rng(seed)
    model = to_ssmhip_model(m)
    prop, oprop = _proposals(orc, m, model)[which]
    algo = ssmhip.ParticleFilter(N, prop, threshold=1.0, precision=prec)
    st = ssmhip.initialise(rng, model, algo, z=rng.standard_normal((N, m.Dx)), ctx=ctx)
    x_gpu, lw_gpu, _, _ = st.export()
    typeless = True
    tol = TOL[prec]
    _, _, ys = orc.simulate_lg(m, len(thresholds), rng)
    for t, thr in enumerate(thresholds, start=1):
        algo.threshold = thr
        u = rng.random(1)
        z = rng.standard_normal((N, m.Dx))
        ost = orc.BFState(x_gpu.copy(), lw_gpu.copy(), typeless)
        _, ll = ssmhip.step(rng, model, algo, t, st, ys[t - 1], u=u, z=z)
        x_new, lw_new, anc_gpu, w_gpu = st.export()
        did_gpu = st.last_resampled
        anc_o, ll_o, did_o, _ = orc.guided_step(m, ost, ys[t - 1], thr, 0, u, z, oprop,
                                                w_override=w_gpu if did_gpu else None)
        assert did_gpu == did_o, f"t={t} resample decision"
        np.testing.assert_array_equal(anc_gpu, anc_o)
        np.testing.assert_allclose(x_new, ost.x, rtol=tol["x"], atol=tol["x"], err_msg=f"t={t}")
        live = ost.logw > -30.0
        np.testing.assert_allclose(lw_new[live], ost.logw[live], rtol=tol["lw"], atol=tol["lw"],
                                   err_msg=f"t={t}")
        assert ll == pytest.approx(ll_o, abs=tol["ll"]), f"t={t}"
        x_gpu, lw_gpu = x_new, lw_new
        typeless = False


@pytest.mark.parametrize("which", ["optimal", "overdispersed"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_teacher_forced_config2_model(ctx, orc, which, prec):
    _teacher_forced(ctx, orc, config2_model(), 4096, prec, which, [1.0, 0.0, 0.5, 1.0, 0.0])


@pytest.mark.parametrize("Dx,Dy", [(1, 1), (3, 2), (6, 4)])
def test_teacher_forced_other_dims(ctx, orc, Dx, Dy):
    m = random_lg(np.random.default_rng(Dx + 10 * Dy), Dx, Dy)
    for which in ("optimal", "overdispersed"):
        _teacher_forced(ctx, orc, m, 1000, "f64", which, [1.0, 0.0, 1.0])


def test_run_equals_steps(ctx, orc):
    # the homogeneous guided filter's graph-replayed run equals its step-by-step path
    m = config2_model()
    model = to_ssmhip_model(m)
    T = 25
    rng = np.random.default_rng(3)
    ys = rng.standard_normal((T, 2))
    algo = ssmhip.PF(50000, ssmhip.OptimalProposal(model.dyn, model.obs), threshold=0.7)
    a = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=99)
    tot, steps = ssmhip.run(a, algo, ys, return_steps=True)
    b = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=99)
    incs = [ssmhip.step(rng, model, algo, t, b, ys[t - 1])[1] for t in range(1, T + 1)]
    assert np.array_equal(np.array(incs), steps)
    for fa, fb in zip(a.export()[:3], b.export()[:3]):
        assert np.array_equal(fa, fb)
    # the bootstrap filter with the same seed gives another estimate: the proposal is in use
    c = ssmhip.initialise(rng, model, ssmhip.BF(50000, threshold=0.7), ctx=ctx, seed=99)
    assert ssmhip.run(c, ssmhip.BF(50000, threshold=0.7), ys) != tot


def test_guided_filter_vs_kalman(ctx, orc):
    # test/algorithms/particles.jl:30-50 on the device: random 1-D model, four observations,
    # the optimal proposal, N = 10^6, resampling at every other step; the weighted mean and
    # the log-likelihood against the KF
    rng = np.random.default_rng(1234)
    m = random_lg(rng, 1, 1)
    _, _, ys = orc.simulate_lg(m, 4, rng)
    mu_kf, _, ll_kf, _ = orc.kf_filter(m, ys)
    model = to_ssmhip_model(m)
    algo = ssmhip.PF(10**6, ssmhip.OptimalProposal(model.dyn, model.obs), precision="f64")
    st = ssmhip.initialise(rng, model, algo, ctx=ctx)
    ll = 0.0
    for t in range(1, 5):
        algo.threshold = 1.0 if t % 2 == 1 else 0.0
        st, inc = ssmhip.step(rng, model, algo, t, st, ys[t - 1])
        ll += inc
    x, lw, _, _ = st.export()
    w = np.exp(lw - lw.max())
    w /= w.sum()
    assert ll == pytest.approx(ll_kf, abs=3e-3)
    assert (w * x[:, 0]).sum() == pytest.approx(mu_kf[0], rel=3e-3)


def test_time_varying_proposal(ctx, orc):
    # a proposal whose linear form changes with the step (an overdispersed proposal over
    # step-dependent dynamics) is re-uploaded at every step
    m = config2_model()
    model = to_ssmhip_model(m)

    class Dyn(ssmhip.models.LinearGaussianLatentDynamics):
        def calc_A(self, step, **kw):
            return (1.0 - 0.05 * step) * m.A

        def calc_b(self, step, **kw):
            return m.b

        def calc_Q(self, step, **kw):
            return m.Q

    tv = ssmhip.StateSpaceModel(model.prior, Dyn(), model.obs)
    prop = ssmhip.OverdispersedProposal(tv.dyn, 2.0)
    algo = ssmhip.PF(2000, prop, threshold=0.0, precision="f64")  # no resampling
    rng = np.random.default_rng(4)
    st = ssmhip.initialise(rng, tv, algo, z=rng.standard_normal((2000, 4)), ctx=ctx)
    x, lw, _, _ = st.export()
    ost = orc.BFState(x.copy(), lw.copy(), True)
    for t in (1, 2, 3):
        z, y = rng.standard_normal((2000, 4)), rng.standard_normal(2)
        _, ll = ssmhip.step(rng, tv, algo, t, st, y, z=z)
        mt = orc.LG(m.mu0, m.Sigma0, (1.0 - 0.05 * t) * m.A, m.b, m.Q, m.H, m.c, m.R)
        _, ll_o, did, _ = orc.guided_step(mt, ost, y, 0.0, 0, [0.5], z,
                                          orc.overdispersed_proposal(mt, 2.0))
        assert not did and not st.last_resampled
        assert ll == pytest.approx(ll_o, abs=1e-9), f"t={t}"
        np.testing.assert_allclose(st.export()[0], ost.x, rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(st.export()[1], ost.logw, rtol=1e-9, atol=1e-9)


def test_errors(ctx, orc):
    m = config2_model()
    model = to_ssmhip_model(m)
    bad = ssmhip.LinearGaussianProposal(m.A, m.b, -np.eye(4))
    with pytest.raises(ssmhip.PosDefException):
        ssmhip.initialise(np.random.default_rng(0), model, ssmhip.PF(16, bad), ctx=ctx)
    st = ssmhip.initialise(np.random.default_rng(0), model,
                           ssmhip.PF(16, ssmhip.OverdispersedProposal(model.dyn, 2.0)), ctx=ctx)
    with pytest.raises(ssmhip.ArgumentError):  # conditional SMC runs the bootstrap proposal
        st.set_reference(np.zeros((3, 4)))
    st.set_proposal(None)  # back to the bootstrap filter, then a reference is accepted
    st.set_reference(np.zeros((3, 4)))
    with pytest.raises(ssmhip.ArgumentError):
        st.set_proposal((m.A, m.b, None, m.Q))
    # paths that run only the bootstrap proposal refuse a guided filter instead of
    # dropping its proposal
    pf = ssmhip.PF(64, ssmhip.OverdispersedProposal(model.dyn, 1.5))
    with ssmhip.Server(hbm_budget_bytes=64 << 20, max_active=4, ctx=ctx) as sv:
        with pytest.raises(ssmhip.ArgumentError):
            sv.submit(np.random.default_rng(0), model, pf, np.zeros((3, 2)))
    with pytest.raises(TypeError):
        ssmhip.rbpf_filter(np.random.default_rng(0), None, ssmhip.RBPF(pf), np.zeros((3, 1)),
                           ctx=ctx)
    with pytest.raises(TypeError):
        ssmhip.CSMCWorkspace(model, pf, 3, ctx=ctx)
