"""The device auxiliary particle filter (AuxiliaryParticleFilter(BF, MeanPredictive()),
GF/src/algorithms/particles.jl:316-344) against the FP64 oracle, teacher-forced per step as
the bootstrap filter's tests are: the oracle starts from the exported state with the same
noise and uniforms and is handed the look-ahead weights the device consumed, so ancestors
match bit for bit while states, log-weights and increments match within the bootstrap
filter's stated tolerances (the first-stage correction adds an FP32 log-sum-exp of the
look-ahead weights, within them)."""
import numpy as np
import pytest

import ssmhip
from conftest import config2_model, random_lg, to_ssmhip_model
from test_bf_gpu import RES, TOL

pytestmark = pytest.mark.gpu


def _teacher_forced_apf(ctx, orc, m, N, prec, kind, thresholds, seed=3):
    rng = np.random.default_rng(seed)
    model = to_ssmhip_model(m)
    algo = ssmhip.APF(ssmhip.BF(N, threshold=1.0, resampler=RES[kind], precision=prec),
                      ssmhip.MeanPredictive())
    z0 = rng.standard_normal((N, m.Dx))
    st = ssmhip.initialise(rng, model, algo, z=z0, ctx=ctx)
    x_gpu, lw_gpu, _, _ = st.export()
    typeless = True
    tol = TOL[prec]
    _, _, ys = orc.simulate_lg(m, len(thresholds), rng)
    n_rs = 0
    for t, thr in enumerate(thresholds, start=1):
        algo.threshold = thr
        u = rng.random(N) if kind >= 2 else rng.random(1)
        z = rng.standard_normal((N, m.Dx))
        ost = orc.BFState(x_gpu.copy(), lw_gpu.copy(), typeless)
        _, ll = ssmhip.step(rng, model, algo, t, st, ys[t - 1], u=u, z=z)
        x_new, lw_new, anc_gpu, w_gpu = st.export()
        did = st.last_resampled
        anc_o, ll_o, did_o, w_o = orc.apf_step(m, ost, ys[t - 1], thr, kind, u, z,
                                               w_override=w_gpu if did else None)
        assert did == did_o, f"t={t} resample decision"
        n_rs += did
        if did:
            np.testing.assert_allclose(w_gpu, w_o, rtol=50 * tol["w"], atol=1e-12)
            assert np.array_equal(anc_gpu, anc_o), f"t={t} ancestors"
        np.testing.assert_allclose(x_new, ost.x, rtol=tol["x"][0], atol=tol["x"][1])
        live = ost.logw > -30.0
        a_lw = max(tol["lw"][0], 2e-5 * m.Dx * m.Dy) if prec == "f32" else tol["lw"][0]
        r_lw = max(tol["lw"][2], 4e-6 * m.Dx * m.Dy) if prec == "f32" else tol["lw"][2]
        np.testing.assert_allclose(lw_new[live], ost.logw[live], atol=a_lw, rtol=r_lw)
        r_ll = max(1e-7, 4e-6 * m.Dx * m.Dy) if prec == "f32" else 1e-12
        assert ll == pytest.approx(ll_o, abs=tol["ll"], rel=r_ll), f"t={t}"
        x_gpu, lw_gpu = x_new, lw_new
        typeless = False
    return n_rs


@pytest.mark.parametrize("kind", [0, 2, 3])
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_teacher_forced_config2_model(ctx, orc, kind, prec):
    # D = 4, Dy = 2, N off the tile size: resample (from the lazy uniform weights), keep,
    # resample, resample, keep
    n = _teacher_forced_apf(ctx, orc, config2_model(), 4096 + 37, prec, kind,
                            [1.0, 0.0, 1.0, 1.0, 0.0])
    assert n == 3


@pytest.mark.parametrize("Dx,Dy", [(1, 1), (3, 2), (6, 4)])
def test_teacher_forced_generic_dims(ctx, orc, Dx, Dy):
    rng = np.random.default_rng(Dx * 13 + Dy)
    m = random_lg(rng, Dx, Dy)
    m.A = 0.9 * m.A / max(abs(np.linalg.eigvals(m.A)))
    _teacher_forced_apf(ctx, orc, m, 3001, "f32", 0, [0.0, 1.0, 1.0])
    _teacher_forced_apf(ctx, orc, m, 3001, "f64", 2, [1.0, 0.0, 1.0])


@pytest.mark.parametrize("N,prec", [(20000, "f32"), (1 << 20, "f32"), (3000, "f64")])
def test_run_equals_steps(ctx, N, prec):
    # graph-replayed run (look-ahead, its fold, resampler, step, finish per step) against the
    # step-by-step path: identical increments, states and weights, bit for bit
    m = config2_model()
    model = to_ssmhip_model(m)
    T = 37
    rng = np.random.default_rng(4)
    ys = rng.standard_normal((T, 2))
    algo = ssmhip.APF(ssmhip.BF(N, threshold=0.7, precision=prec))
    a = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=21)
    tot, steps = ssmhip.run(a, algo, ys, return_steps=True)
    b = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=21)
    incs = [ssmhip.step(rng, model, algo, t, b, ys[t - 1])[1] for t in range(1, T + 1)]
    assert np.array_equal(np.array(incs), steps)
    for fa, fb in zip(a.export()[:3], b.export()[:3]):
        assert np.array_equal(fa, fb)


def test_abf_statistical_vs_kf(ctx, orc):
    # test/algorithms/particles.jl:57-82 on the device: 1-D random LG model, T = 4,
    # ESSResampler(0.8) and, to exercise the first stage every step, threshold 1.0;
    # N = 2^20 FP32: the log-likelihood within 1e-3 of the exact KF, the filtered mean
    # within rtol 1e-2.
    rng = np.random.default_rng(1234)
    m = random_lg(rng, 1, 1)
    _, _, ys = orc.simulate_lg(m, 4, rng)
    mu_kf, _, ll_kf, _ = orc.kf_filter(m, ys)
    model = to_ssmhip_model(m)
    for thr in (0.8, 1.0):
        algo = ssmhip.APF(ssmhip.BF(1 << 20, threshold=thr), ssmhip.MeanPredictive())
        st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=77)
        ll = ssmhip.run(st, algo, ys)
        x, lw, _, _ = st.export()
        w = np.exp(lw - lw.max())
        w /= w.sum()
        assert ll == pytest.approx(ll_kf, abs=1e-3), thr
        assert (w * x[:, 0]).sum() == pytest.approx(mu_kf[0], rel=1e-2), thr


def test_refusals(ctx):
    m = config2_model()
    model = to_ssmhip_model(m)
    rng = np.random.default_rng(0)
    st = ssmhip.initialise(rng, model, ssmhip.APF(ssmhip.BF(256)), ctx=ctx, seed=1)
    with pytest.raises(ssmhip.SSMError):
        st.set_reference(np.zeros((5, 4)))
    with pytest.raises(TypeError):
        ssmhip.APF(ssmhip.BF(256), object())
