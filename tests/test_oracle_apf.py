"""The oracle's auxiliary particle filter step (oracle.apf_step, a restatement of
GF/src/algorithms/particles.jl:248-344 and GF/src/resamplers.jl:92-143) pinned three ways:
the reference's own ABF test against the exact Kalman filter, the reduction to the bootstrap
filter when the look-ahead score does not depend on the particle, and the textbook
two-stage estimator of the evidence computed independently with scipy."""
import numpy as np
import pytest
from scipy.stats import multivariate_normal

import oracle
from conftest import random_lg


@pytest.mark.parametrize("threshold", [0.8, 1.0])
def test_abf_statistical_vs_kf(threshold):
    # test/algorithms/particles.jl:57-82: 1-D random LG model, T = 4, BF(10^6) with
    # ESSResampler(0.8) wrapped in AuxiliaryParticleFilter(bf, MeanPredictive()); the
    # log-likelihood within 1e-3 of the KF's and the filtered mean within rtol 1e-2.  At
    # threshold 1.0 every step resamples, so the first-stage correction enters each increment.
    rng = np.random.default_rng(1234)
    m = random_lg(rng, 1, 1)
    _, _, ys = oracle.simulate_lg(m, 4, rng)
    mu_kf, _, ll_kf, _ = oracle.kf_filter(m, ys)
    N = 10**6
    st = oracle.BFState(oracle.bf_initialise(m, rng.standard_normal((N, 1))))
    ll, n_rs = 0.0, 0
    for t in range(4):
        _, inc, did, _ = oracle.apf_step(m, st, ys[t], threshold, 0, rng.random(1),
                                         rng.standard_normal((N, 1)))
        ll += inc
        n_rs += did
    if threshold == 1.0:
        assert n_rs == 4
    w = np.exp(st.logw - st.logw.max())
    w /= w.sum()
    assert ll == pytest.approx(ll_kf, abs=1e-3)
    assert (w * st.x[:, 0]).sum() == pytest.approx(mu_kf[0], rel=1e-2)


@pytest.mark.parametrize("threshold", [1.0, 0.0])
def test_lookahead_independent_of_the_particle_is_the_bootstrap_filter(threshold):
    # A = 0: the transition mean b, and so log eta, is the same for every particle; then the
    # auxiliary weights are the filter's weights, every resampled particle starts from the
    # same -log eta and the first-stage correction cancels it: the bootstrap filter's step.
    rng = np.random.default_rng(7)
    m = random_lg(rng, 3, 2)
    m.A = np.zeros((3, 3))
    N = 5000
    x0 = rng.standard_normal((N, 3))
    lw0 = np.log(rng.dirichlet(np.ones(N)))
    a_ = oracle.BFState(x0.copy(), lw0.copy(), False)
    b_ = oracle.BFState(x0.copy(), lw0.copy(), False)
    y = rng.standard_normal(2)
    u = rng.random(1)
    z = rng.standard_normal((N, 3))
    anc_b, ll_b, did_b, w_b = oracle.bf_step(m, b_, y, threshold, 0, u, z)
    anc_a, ll_a, did_a, w_a = oracle.apf_step(m, a_, y, threshold, 0, u, z, w_override=w_b)
    assert did_a == did_b == (threshold == 1.0)
    np.testing.assert_allclose(w_a, w_b, rtol=1e-12, atol=1e-300)
    assert np.array_equal(anc_a, anc_b)
    np.testing.assert_allclose(a_.x, b_.x, rtol=1e-14)
    np.testing.assert_allclose(a_.logw, b_.logw, rtol=1e-10, atol=1e-10)
    assert ll_a == pytest.approx(ll_b, rel=1e-12)


def test_increment_is_the_two_stage_estimator():
    # Pitt & Shephard's auxiliary estimator of p(y_t | y_1:t-1), from first principles:
    # [sum_i w_i eta_i] * [1/N sum_j g(y | x'_j) / eta_{a_j}], with eta_i the observation
    # density at the transition mean A x_i + b and g at the propagated state.
    rng = np.random.default_rng(11)
    m = random_lg(rng, 2, 2)
    m.A = 0.8 * m.A / max(abs(np.linalg.eigvals(m.A)))
    N = 7
    x0 = rng.standard_normal((N, 2))
    w0 = rng.dirichlet(np.ones(N))
    st = oracle.BFState(x0.copy(), np.log(w0), False)
    y = rng.standard_normal(2)
    u = rng.random(1)
    z = rng.standard_normal((N, 2))
    anc, ll, did, wa = oracle.apf_step(m, st, y, 1.0, 0, u, z)
    assert did
    mvn = lambda mean: multivariate_normal(mean, m.R).pdf(y)
    eta = np.array([mvn(m.H @ (m.A @ x + m.b) + m.c) for x in x0])
    np.testing.assert_allclose(wa, w0 * eta / (w0 * eta).sum(), rtol=1e-12)
    LQ = np.linalg.cholesky(m.Q)
    xp = np.array([m.A @ x0[a] + m.b + LQ @ z[i] for i, a in enumerate(anc)])
    np.testing.assert_allclose(st.x, xp, rtol=1e-12)
    g = np.array([mvn(m.H @ x + m.c) for x in xp])
    want = np.log((w0 * eta).sum()) + np.log(np.mean(g / eta[anc]))
    assert ll == pytest.approx(want, rel=1e-12)
