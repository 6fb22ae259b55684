"""Device BootstrapFilter against the FP64 oracle.

Teacher-forced per step: the oracle starts from the state the device exported, uses the
same noise and uniforms, and hands its resampler the weights the device consumed, so
ancestors must match bit for bit while states, log-weights and log-likelihood increments
must match within the tolerances stated below (FP32 state arithmetic vs the FP64 oracle).
"""
import numpy as np
import pytest

import ssmhip
from conftest import config1_model, config2_model, random_lg, to_ssmhip_model

pytestmark = pytest.mark.gpu

# FP32 path tolerances (stated): states rtol 2e-5 / atol 2e-5; normalised log-weights of
# particles that carry weight (oracle value > -30) within a + r |logw| with
# a = max(2e-4, 2e-5 Dx Dy), r = max(2e-4, 4e-6 Dx Dy) -- the FP32 state error carried
# through the gradient of the observation quadratic form (and into the normaliser), which
# grows with Dx Dy for dense H -- the others (weights
# below e^-30) rtol 3e-4; log-likelihood increments atol 1e-4 plus 4e-6 Dx Dy relative.
# FP64 path: 1e-9.
TOL = {"f32": dict(x=(2e-5, 2e-5), lw=(2e-4, 3e-4, 2e-4), ll=1e-4, w=2e-5),
       "f64": dict(x=(1e-10, 1e-10), lw=(1e-9, 1e-12, 1e-12), ll=1e-9, w=1e-10)}
RES = {0: ssmhip.Systematic(), 1: ssmhip.Systematic(form="gpu"), 2: ssmhip.Stratified(),
       3: ssmhip.Multinomial()}


def _teacher_forced(ctx, orc, m, N, prec, kind, thresholds, seed=1, anc_flips=0.0):
    # anc_flips: the fraction of ancestors allowed to differ from the oracle's sequential
    # cumsum, whose rounding error grows like sqrt(N) ulp: where a stratum boundary falls
    # between the two cumulative sums the neighbouring particle is chosen.  Such slots are
    # left out of the state comparison.
    rng = np.random.default_rng(seed)
    model = to_ssmhip_model(m)
    algo = ssmhip.BF(N, threshold=1.0, resampler=RES[kind], precision=prec)
    z0 = rng.standard_normal((N, m.Dx))
    st = ssmhip.initialise(rng, model, algo, z=z0, ctx=ctx)
    x_gpu, lw_gpu, _, _ = st.export()
    np.testing.assert_allclose(x_gpu, orc.bf_initialise(m, z0), rtol=2e-6, atol=2e-6)
    typeless = True
    tol = TOL[prec]
    _, _, ys = orc.simulate_lg(m, len(thresholds), rng)
    for t, thr in enumerate(thresholds, start=1):
        algo.threshold = thr
        u = rng.random(N) if kind >= 2 else rng.random(1)
        z = rng.standard_normal((N, m.Dx))
        ost = orc.BFState(x_gpu.copy(), lw_gpu.copy(), typeless)
        _, ll = ssmhip.step(rng, model, algo, t, st, ys[t - 1], u=u, z=z)
        x_new, lw_new, anc_gpu, w_gpu = st.export()
        did_gpu = st.last_resampled
        # oracle from the same starting state; resampler fed the device's weights
        anc_o, ll_o, did_o, w_o = orc.bf_step(m, ost, ys[t - 1], thr, kind, u, z,
                                              w_override=w_gpu if did_gpu else None)
        assert did_gpu == did_o, f"t={t} resample decision"
        same = np.ones(N, dtype=bool)
        if did_gpu:
            np.testing.assert_allclose(w_gpu, w_o, rtol=tol["w"], atol=1e-12)
            same = anc_gpu == anc_o
            flips = np.flatnonzero(~same)
            assert len(flips) <= anc_flips * N, f"t={t} {len(flips)} ancestors differ"
            # a flip moves along the cumulative order by less than the rounding of the sums:
            # the weight strictly between the two choices is below 1e-9 of the total
            for s_ in flips:
                lo_, hi_ = sorted((anc_gpu[s_], anc_o[s_]))
                between = w_gpu[lo_ + 1:hi_].sum()
                assert between <= 1e-9 * w_gpu.sum(), f"t={t} slot {s_}: {lo_}..{hi_} {between}"
        np.testing.assert_allclose(x_new[same], ost.x[same], rtol=tol["x"][0], atol=tol["x"][1])
        live = (ost.logw > -30.0) & same
        r_lw = max(tol["lw"][2], 4e-6 * m.Dx * m.Dy) if prec == "f32" else tol["lw"][2]
        a_lw = max(tol["lw"][0], 2e-5 * m.Dx * m.Dy) if prec == "f32" else tol["lw"][0]
        np.testing.assert_allclose(lw_new[live], ost.logw[live], atol=a_lw, rtol=r_lw)
        # weights below e^-30 of the live ones: FP32 cancellation in the residual
        # y - Hx - c dominates, growing with Dx * Dy like the live tolerance
        r_dead = max(tol["lw"][1], 5e-6 * m.Dx * m.Dy) if prec == "f32" else tol["lw"][1]
        dead = ~live & same
        np.testing.assert_allclose(lw_new[dead], ost.logw[dead], rtol=r_dead)
        r_ll = max(1e-7, 4e-6 * m.Dx * m.Dy) if prec == "f32" else 1e-12
        assert ll == pytest.approx(ll_o, abs=tol["ll"], rel=r_ll), f"t={t}"
        x_gpu, lw_gpu = x_new, lw_new
        typeless = False


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_teacher_forced_config2_model(ctx, orc, kind, prec):
    # D = 4, Dy = 2; N not a multiple of the tile; resample, keep, resample, keep(0.5 rule).
    _teacher_forced(ctx, orc, config2_model(), 4096 + 37, prec, kind, [1.0, 0.0, 1.0, 0.5, 0.5])


@pytest.mark.parametrize("Dx,Dy", [(1, 1), (3, 2), (7, 5), (16, 16)])
def test_teacher_forced_generic_dims(ctx, orc, Dx, Dy):
    # GFTest-style random model with its dynamics rescaled to spectral radius 0.9 (the
    # unscaled rand(Dx, Dx) matrix grows the state ~3x per step, which FP32 cannot follow).
    rng = np.random.default_rng(Dx * 31 + Dy)
    m = random_lg(rng, Dx, Dy)
    m.A = 0.9 * m.A / max(abs(np.linalg.eigvals(m.A)))
    _teacher_forced(ctx, orc, m, 3001, "f32", 0, [1.0, 0.0, 1.0])
    _teacher_forced(ctx, orc, m, 3001, "f64", 2, [1.0, 0.5, 1.0])


def test_teacher_forced_many_tiles(ctx, orc):
    # 8197 tiles: the finish streams each thread's chunk of tiles from memory in groups
    # (more than 1024 tiles); resample, keep (baseline sums), then the 0.5 rule.
    _teacher_forced(ctx, orc, config2_model(), (1 << 23) + 4321, "f32", 0, [1.0, 0.0, 0.5],
                    anc_flips=1e-5)


def test_tiny_particle_counts(ctx, orc):
    m = config2_model()
    for N in (1, 2, 3):
        _teacher_forced(ctx, orc, m, N, "f64", 0, [1.0, 1.0, 0.0])


def test_philox_noise_matches_oracle(ctx, orc):
    # Without overrides the device draws Philox4x32-10 noise keyed by (seed, t, i); the
    # oracle's restatement of the same draws must reproduce the step.
    m = config2_model()
    N = 5000
    rng = np.random.default_rng(0)
    model = to_ssmhip_model(m)
    algo = ssmhip.BF(N, threshold=1.0, resampler=ssmhip.Systematic(), precision="f32")
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=123456789)
    x0, _, _, _ = st.export()
    z0 = orc.philox_normals(123456789, 2, 0, N, 4)
    np.testing.assert_allclose(x0, orc.bf_initialise(m, z0), rtol=2e-5, atol=2e-5)
    y = np.array([0.3, -0.2])
    ost = orc.BFState(x0.copy(), None, True)
    _, ll = ssmhip.step(rng, model, algo, 1, st, y)
    x1, lw1, anc, w = st.export()
    z1 = orc.philox_normals(123456789, 0, 1, N, 4)
    u1 = orc.philox_uniforms(123456789, 1, 1)
    anc_o, ll_o, did_o, _ = orc.bf_step(m, ost, y, 1.0, 0, u1, z1, w_override=w)
    assert np.array_equal(anc, anc_o)
    np.testing.assert_allclose(x1, ost.x, rtol=5e-5, atol=5e-5)
    assert ll == pytest.approx(ll_o, abs=1e-4)


@pytest.mark.parametrize("kind,N,prec", [(0, 20000, "f32"), (2, 20000, "f32"),
                                         (3, 20000, "f32"), (0, 300000, "f32"),
                                         (2, 5000, "f64"), (3, 1 << 20, "f32"),
                                         (1, 70001, "f32"), (0, 1 << 20, "f32"),
                                         (0, (1 << 20) + 1, "f32"), (3, 3000, "f64")])
def test_run_equals_steps(ctx, kind, N, prec):
    # The graph-replayed run (up to 1024 tiles: the finish folded into the next step's
    # resampler) and the step-by-step path (resampler, step kernel, finish) run the same
    # instructions on the same counters: identical log-likelihood increments, bit for bit.
    m = config2_model()
    model = to_ssmhip_model(m)
    T = 37
    rng = np.random.default_rng(3)
    ys = rng.standard_normal((T, 2))
    algo = ssmhip.BF(N, threshold=0.7, resampler=RES[kind], precision=prec)
    a = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=99)
    tot, steps = ssmhip.run(a, algo, ys, return_steps=True)
    b = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=99)
    incs = [ssmhip.step(rng, model, algo, t, b, ys[t - 1])[1] for t in range(1, T + 1)]
    assert np.array_equal(np.array(incs), steps)
    assert tot == pytest.approx(sum(incs), rel=1e-12)
    for fa, fb in zip(a.export()[:3], b.export()[:3]):  # states, log-weights, ancestors
        assert np.array_equal(fa, fb)


@pytest.mark.parametrize("kind,N,Dx,Dy,prec,obs_scale", [
    (1, 30000, 4, 2, "f32", 1.0),      # GPU systematic form, register model
    (0, 4099, 1, 1, "f32", 1.0),       # 1-D model, ragged last tile
    (2, 10000, 3, 2, "f32", 1.0),      # generic-shape step
    (0, 6000, 3, 3, "f64", 1.0),       # FP64 generic-shape step
    (0, 50000, 2, 2, "f32", 1e-4),     # concentrated weights: heavy ancestors, long ranges
    (2, 50000, 2, 2, "f32", 1e-4),
])
def test_run_equals_steps_shapes(ctx, kind, N, Dx, Dy, prec, obs_scale):
    # The graph-replayed run against the step-by-step path over the kernel shapes, both
    # precisions and concentrated weights (heavy ancestors): bit for bit.
    rng = np.random.default_rng(N + Dx)
    m = random_lg(rng, Dx, Dy, obs_noise_scale=obs_scale)
    m.A = 0.9 * m.A / max(abs(np.linalg.eigvals(m.A)))
    model = to_ssmhip_model(m)
    T = 21
    ys = rng.standard_normal((T, Dy))
    algo = ssmhip.BF(N, threshold=0.8, resampler=RES[kind], precision=prec)
    a = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=5)
    tot, steps = ssmhip.run(a, algo, ys, return_steps=True)
    b = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=5)
    incs = [ssmhip.step(rng, model, algo, t, b, ys[t - 1])[1] for t in range(1, T + 1)]
    assert np.array_equal(np.array(incs), steps)
    for fa, fb in zip(a.export()[:3], b.export()[:3]):
        assert np.array_equal(fa, fb)
    # a run that continues from a state an ssm_bf_step call left
    ys2 = rng.standard_normal((18, Dy))
    ssmhip.step(rng, model, algo, T + 1, a, ys2[0])
    ssmhip.step(rng, model, algo, T + 1, b, ys2[0])
    _, s2 = ssmhip.run(a, algo, ys2[1:], t0=T + 2, return_steps=True)
    incs2 = [ssmhip.step(rng, model, algo, t, b, ys2[t - T - 1])[1]
             for t in range(T + 2, T + 19)]
    assert np.array_equal(np.array(incs2), s2)
    for fa, fb in zip(a.export()[:3], b.export()[:3]):
        assert np.array_equal(fa, fb)


def test_config1_fp64_statistical(ctx, orc):
    # 1-D LG model, T = 100, BF(1024; threshold 0.5) systematic, FP64 on the device: over
    # 30 seeds the mean log-likelihood is within 3 standard errors of the exact KF value.
    m = config1_model()
    rng = np.random.default_rng(1234)
    _, _, ys = orc.simulate_lg(m, 100, rng)
    _, _, ll_kf, _ = orc.kf_filter(m, ys)
    model = to_ssmhip_model(m)
    algo = ssmhip.BF(1024, threshold=0.5, resampler=ssmhip.Systematic(), precision="f64")
    lls = []
    for rep in range(30):
        _, ll = ssmhip.filter(np.random.default_rng(rep), model, algo, ys, ctx=ctx)
        lls.append(ll)
    lls = np.array(lls)
    se = lls.std(ddof=1) / np.sqrt(len(lls))
    print(f"KF {ll_kf:.4f}  PF mean {lls.mean():.4f} +- {se:.4f}")
    assert abs(lls.mean() - ll_kf) < 3 * se + 0.02


def test_config2_large_n_vs_kalman(ctx, orc):
    # D = 4 model at N = 2^20, FP32, T = 200: the PF log-likelihood is within 1e-3 relative of
    # the FP64 Kalman oracle (Monte Carlo error dominates; FP32 error is far below it).
    m = config2_model()
    rng = np.random.default_rng(1)
    _, _, ys = orc.simulate_lg(m, 200, rng)
    _, _, ll_kf, _ = orc.kf_filter(m, ys)
    model = to_ssmhip_model(m)
    for R in (ssmhip.Systematic(), ssmhip.Stratified(), ssmhip.Multinomial()):
        algo = ssmhip.BF(1 << 20, threshold=1.0, resampler=R, precision="f32")
        _, ll = ssmhip.filter(np.random.default_rng(5), model, algo, ys, ctx=ctx)
        print(f"{R!r}: PF {ll:.5f} KF {ll_kf:.5f} rel {abs(ll - ll_kf) / abs(ll_kf):.2e}")
        assert abs(ll - ll_kf) <= 1e-3 * abs(ll_kf)


def test_statistical_alternating_vs_kf(ctx, orc):
    # test/algorithms/particles.jl:5-27 on the device: random 1-D model, T = 4, N = 10^6,
    # resample / skip alternating (threshold 1 then 0): ll within 1e-3 of the KF.
    rng = np.random.default_rng(1234)
    m = random_lg(rng, 1, 1)
    _, _, ys = orc.simulate_lg(m, 4, rng)
    mu_kf, _, ll_kf, _ = orc.kf_filter(m, ys)
    model = to_ssmhip_model(m)
    algo = ssmhip.BF(10**6, threshold=1.0, precision="f32")
    st = ssmhip.initialise(rng, model, algo, ctx=ctx)
    ll = 0.0
    for t in range(1, 5):
        algo.threshold = 1.0 if t % 2 == 1 else 0.0
        st, inc = ssmhip.step(rng, model, algo, t, st, ys[t - 1])
        ll += inc
    x, lw, _, _ = st.export()
    w = np.exp(lw - lw.max())
    w /= w.sum()
    assert ll == pytest.approx(ll_kf, abs=1e-3)
    assert (w * x[:, 0]).sum() == pytest.approx(mu_kf[0], rel=1e-3)


def test_errors(ctx):
    m = config2_model()
    model = to_ssmhip_model(m)
    bad = ssmhip.create_homogeneous_linear_gaussian_model(m.mu0, m.Sigma0, m.A, m.b,
                                                          -np.eye(4), m.H, m.c, m.R)
    with pytest.raises(ssmhip.PosDefException):
        ssmhip.initialise(np.random.default_rng(0), bad, ssmhip.BF(16), ctx=ctx)
    st = ssmhip.DeviceParticles(ctx, model, ssmhip.BF(16), 1)
    with pytest.raises(ssmhip.SSMError):
        ssmhip.step(np.random.default_rng(0), model, ssmhip.BF(16), 1, st, np.zeros(2))
