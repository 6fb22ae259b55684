"""Device RBPF against the FP64 oracle.

Teacher-forced per step (same starting state, uniforms and outer noise; the oracle's
resampler gets the device's weights): ancestors bit-exact; outer states, inner means and
covariances, log-weights and increments within the FP32 tolerances stated below.
"""
import math

import numpy as np
import pytest

import ssmhip
from conftest import rand_cov
from test_oracle import dummy_hierarchical

pytestmark = pytest.mark.gpu

# stated FP32 tolerances (inner Kalman products in FP32, S factorised in FP64)
TOL_X = 2e-5        # outer state, rtol/atol
TOL_MU = 6e-4       # inner mean, relative to max(1, |mu|)
TOL_SIG = 4e-4      # inner covariance, relative to its largest entry
TOL_LW = (5e-4, 1e-4)  # live log-weights: atol + max(rtol, 4e-6 D Dy) * |logw|
TOL_LL = 3e-4       # plus 4e-6 D Dy relative


def rb_pair(orc, om0, os0, oa, ob, oq, mu0, S0, A, b, Cv, Q, H, c, R0, rs):
    args = (om0, os0, oa, ob, oq, mu0, S0, A, b, Cv, Q, H, c, R0, rs)
    return orc.RB(*args), ssmhip.RBModel(*args)


def config4_pair(orc, D=32, Dy=8, seed=1234):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((D, D))
    Qm, Rm = np.linalg.qr(G)
    A = 0.9 * Qm * np.sign(np.diag(Rm))
    H = rng.standard_normal((Dy, D))
    oq = 0.3 ** 2
    return rb_pair(orc, 0.0, oq / (1 - 0.81), 0.9, 0.0, oq, np.zeros(D), np.eye(D), A,
                   np.zeros(D), np.zeros(D), 0.1 * np.eye(D), H, np.zeros(Dy), 0.2 * np.eye(Dy),
                   1.0)


def _teacher_forced(ctx, orc, om, dm, N, kind, thresholds, ys, seed=2):
    rng = np.random.default_rng(seed)
    st = ssmhip.DeviceRBParticles(ctx, dm, N, 77)
    z0 = rng.standard_normal(N)
    st.initialise(z0)
    ost = orc.rbpf_initialise(om, z0)
    xo, mu, Sig, lw, _, _ = st.export()
    np.testing.assert_allclose(xo, ost.xo, rtol=TOL_X, atol=TOL_X)
    typeless = True
    for t, thr in enumerate(thresholds, start=1):
        u = rng.random(N) if kind >= 2 else rng.random(1)
        z = rng.standard_normal(N)
        o = orc.RBState(xo, mu, Sig, lw, typeless)
        ll = st.step(t, ys[t - 1], thr, kind, u=u, z=z)
        xo, mu, Sig, lw, anc, w = st.export()
        anc_o, ll_o, did_o, w_o = orc.rbpf_step(om, o, ys[t - 1], thr, kind, u, z,
                                                w_override=w if st.last_resampled else None)
        assert st.last_resampled == did_o
        if did_o:
            np.testing.assert_allclose(w, w_o, rtol=2e-5, atol=1e-12)
            assert np.array_equal(anc, anc_o), f"t={t}"
        np.testing.assert_allclose(xo, o.xo, rtol=TOL_X, atol=TOL_X)
        np.testing.assert_allclose(mu, o.mu, atol=TOL_MU * max(1.0, np.abs(o.mu).max()), rtol=0)
        np.testing.assert_allclose(Sig, o.Sigma, atol=TOL_SIG * np.abs(o.Sigma).max(), rtol=0)
        live = o.logw > -30
        r_lw = max(TOL_LW[1], 4e-6 * om.D * om.Dy)
        np.testing.assert_allclose(lw[live], o.logw[live], atol=TOL_LW[0], rtol=r_lw)
        assert ll == pytest.approx(ll_o, abs=TOL_LL, rel=4e-6 * om.D * om.Dy), f"t={t}"
        typeless = False


@pytest.mark.parametrize("kind", [0, 2, 3])
def test_teacher_forced_config4_model(ctx, orc, kind):
    # D = 32, Dy = 8, R_t = exp(x_t) 0.2 I; N not a multiple of the block's waves.
    om, dm = config4_pair(orc)
    ys = np.random.default_rng(9).standard_normal((4, 8))
    _teacher_forced(ctx, orc, om, dm, 1030, kind, [1.0, 0.0, 1.0, 0.5], ys)


@pytest.mark.parametrize("D,Dy", [(1, 1), (5, 3), (17, 16)])
def test_teacher_forced_generic_dims(ctx, orc, D, Dy):
    rng = np.random.default_rng(D * 13 + Dy)
    A = rng.random((D, D))
    A = 0.9 * A / max(abs(np.linalg.eigvals(A)))
    om, dm = rb_pair(orc, 0.1, 0.2, 0.8, 0.05, 0.1, rng.random(D), rand_cov(rng, D), A,
                     rng.random(D), rng.random(D), rand_cov(rng, D, 0.1), rng.random((Dy, D)),
                     rng.random(Dy), rand_cov(rng, Dy), 0.5)
    ys = rng.standard_normal((3, Dy))
    _teacher_forced(ctx, orc, om, dm, 700, 0, [1.0, 0.0, 1.0], ys)


def test_rbpf_statistical_vs_joint_kf(ctx, orc):
    # test/algorithms/rbpf.jl:5-34 on the device: the hierarchical split of a joint LG model;
    # RBPF log-likelihood within 1e-3 of the joint KF (N = 2^18, T = 4, alternating).
    rng = np.random.default_rng(1234)
    full, rb = dummy_hierarchical(rng)
    _, _, ys = orc.simulate_lg(full, 4, rng)
    _, _, ll_kf, _ = orc.kf_filter(full, ys)
    dm = ssmhip.RBModel(rb.om0, rb.os0, rb.oa, rb.ob, rb.oq, rb.mu0, rb.Sigma0, rb.A, rb.b, rb.C,
                        rb.Q, rb.H, rb.c, rb.R0, rb.rs)
    N = 1 << 18
    st = ssmhip.DeviceRBParticles(ctx, dm, N, 5)
    st.initialise()
    ll = 0.0
    for t in range(1, 5):
        ll += st.step(t, ys[t - 1], 1.0 if t % 2 == 1 else 0.0, 0)
    assert ll == pytest.approx(ll_kf, abs=2e-3)


def test_degenerate_outer_equals_kf(ctx, orc):
    # Outer noise 0 and a point-mass outer prior: every particle is identical and the RBPF
    # log-likelihood equals the KF with R_t = exp(x_t) R0 (FP64 oracle), within FP32 error.
    rng = np.random.default_rng(9)
    D, Dy, T = 8, 4, 10
    A = 0.8 * np.eye(D)
    Q = rand_cov(rng, D, 0.1)
    H = rng.standard_normal((Dy, D))
    R0 = rand_cov(rng, Dy, 0.5)
    dm = ssmhip.RBModel(0.3, 0.0, 0.9, 0.05, 0.0, np.zeros(D), np.eye(D), A, np.zeros(D),
                        np.zeros(D), Q, H, np.zeros(Dy), R0, 1.0)
    ys = rng.standard_normal((T, Dy))
    st = ssmhip.DeviceRBParticles(ctx, dm, 256, 3)
    st.initialise()
    ll = st.run(ys, 0.5, 0)
    x, mu, S = 0.3, np.zeros(D), np.eye(D)
    ll_kf = 0.0
    for t in range(T):
        x = 0.9 * x + 0.05
        mu, S, l = orc.kf_step(D, Dy, mu, S, A, np.zeros(D), Q, H, np.zeros(Dy),
                               math.exp(x) * R0, ys[t])
        ll_kf += l
    assert ll == pytest.approx(ll_kf, rel=1e-5, abs=1e-4)


def test_run_equals_steps(ctx, orc):
    om, dm = config4_pair(orc, D=32, Dy=8)
    ys = np.random.default_rng(4).standard_normal((12, 8))
    a = ssmhip.DeviceRBParticles(ctx, dm, 2048, 11)
    a.initialise()
    tot, steps = a.run(ys, 0.5, 0, return_steps=True)
    b = ssmhip.DeviceRBParticles(ctx, dm, 2048, 11)
    b.initialise()
    incs = [b.step(t, ys[t - 1], 0.5, 0) for t in range(1, 13)]
    assert np.array_equal(np.array(incs), steps)
    assert np.array_equal(a.export()[2], b.export()[2])
