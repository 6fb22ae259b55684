"""The oracle's guided particle filter step (GF/src/algorithms/particles.jl:147-195) and the
reference's test proposals (GF/src/GFTest/proposals.jl): the latent proposal reduces it to
the C oracle's bootstrap step; the optimal proposal's weight increment is the closed-form
predictive density of every particle; the overdispersed proposal's log f - log q is a
closed form in the draw; and the reference's guided-filter test against the KF
(test/algorithms/particles.jl:30-50) at a CPU-sized particle count."""
import numpy as np
import pytest

import oracle
from conftest import rand_cov


def _model(rng, D, Dy):
    A = 0.8 * np.linalg.qr(rng.standard_normal((D, D)))[0]
    return oracle.LG(rng.standard_normal(D), rand_cov(rng, D), A, 0.1 * rng.standard_normal(D),
                     rand_cov(rng, D, 0.5), rng.standard_normal((Dy, D)),
                     0.1 * rng.standard_normal(Dy), rand_cov(rng, Dy, 0.3))


def _state(rng, N, D):
    lw = rng.standard_normal(N)
    return rng.standard_normal((N, D)), lw - np.log(np.exp(lw).sum())


def _lse(v):
    return v.max() + np.log(np.exp(v - v.max()).sum())


def _check_increment(st, lw, want, inc):
    # the step adds `want` to the normalised log-weights lw, then normalises again; the
    # log-likelihood increment is the log of the weighted mean of exp(want)
    np.testing.assert_allclose(st.logw, lw + want - _lse(lw + want), rtol=1e-10, atol=1e-10)
    assert inc == pytest.approx(_lse(lw + want) - _lse(lw), abs=1e-10)


@pytest.mark.parametrize("threshold,typeless,N", [(0.0, False, 64), (1.0, False, 64),
                                                  (0.0, True, 64), (1.0, True, 2000)])
def test_latent_proposal_equals_bootstrap_step(threshold, typeless, N):
    # N = 2000 with uniform weights: the ESS is N up to the rounding of the sum of squares,
    # so the resample decision at threshold 1 depends on summing as the reference does
    rng = np.random.default_rng(1)
    m = _model(rng, 3, 2)
    x, lw = _state(rng, N, 3)
    u, z, y = rng.random(1), rng.standard_normal((N, 3)), rng.standard_normal(2)
    a = oracle.BFState(x.copy(), lw, typeless)
    b = oracle.BFState(x.copy(), lw, typeless)
    anc_a, ll_a, did_a, w_a = oracle.bf_step(m, a, y, threshold, oracle.SYSTEMATIC, u, z)
    prop = (m.A, m.b, None, m.Q)
    anc_b, ll_b, did_b, _ = oracle.guided_step(m, b, y, threshold, oracle.SYSTEMATIC, u, z, prop,
                                               w_override=w_a)
    assert did_a == did_b
    if N == 64:  # 1/64 squares and sums exactly, so the decision is threshold's alone
        assert did_a == (threshold == 1.0)
    np.testing.assert_array_equal(anc_a, anc_b)
    np.testing.assert_allclose(b.x, a.x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(b.logw, a.logw, rtol=1e-12, atol=1e-12)
    assert ll_b == pytest.approx(ll_a, abs=1e-12)


@pytest.mark.parametrize("D,Dy", [(1, 1), (3, 2), (4, 2)])
def test_optimal_proposal_increment_is_the_predictive_density(D, Dy):
    # with q = p(x' | x, y) the increment log f + log g - log q = log p(y | x) for any x'
    rng = np.random.default_rng(D + 10 * Dy)
    m = _model(rng, D, Dy)
    N = 50
    x, lw = _state(rng, N, D)
    z, y = 2.0 * rng.standard_normal((N, D)), rng.standard_normal(Dy)
    st = oracle.BFState(x.copy(), lw, False)
    _, inc, _, _ = oracle.guided_step(m, st, y, 0.0, oracle.SYSTEMATIC, [0.5], z,
                                      oracle.optimal_proposal(m))
    S = m.H @ m.Q @ m.H.T + m.R
    want = np.array([oracle.mvn_logpdf(y, m.H @ (m.A @ xi + m.b) + m.c, S) for xi in x])
    _check_increment(st, lw, want, inc)
    # and the proposal's mean and covariance are those of the conditional p(x' | x, y)
    F, g0, K, P = oracle.optimal_proposal(m)
    Pc = np.linalg.inv(np.linalg.inv(m.Q) + m.H.T @ np.linalg.inv(m.R) @ m.H)
    np.testing.assert_allclose(P, Pc, rtol=1e-10, atol=1e-12)
    xi = x[0]
    mc = Pc @ (np.linalg.solve(m.Q, m.A @ xi + m.b) + m.H.T @ np.linalg.solve(m.R, y - m.c))
    np.testing.assert_allclose(F @ xi + g0 + K @ y, mc, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("k", [1.5, 4.0])
def test_overdispersed_increment_closed_form(k):
    # q = N(A x + b, k Q), x' = A x + b + sqrt(k) L_Q z:
    # log f(x' | x) - log q(x' | x) = (D / 2) log k - (k - 1) |z|^2 / 2
    rng = np.random.default_rng(7)
    D, Dy = 3, 2
    m = _model(rng, D, Dy)
    m0 = oracle.LG(m.mu0, m.Sigma0, m.A, m.b, m.Q, np.zeros((Dy, D)), m.c, m.R)  # g = const
    N = 40
    x, lw = _state(rng, N, D)
    z, y = rng.standard_normal((N, D)), rng.standard_normal(Dy)
    st = oracle.BFState(x.copy(), lw, False)
    _, inc, _, _ = oracle.guided_step(m0, st, y, 0.0, oracle.SYSTEMATIC, [0.5], z,
                                      oracle.overdispersed_proposal(m0, k))
    lg = oracle.mvn_logpdf(y, m0.c, m0.R)
    want = 0.5 * D * np.log(k) - 0.5 * (k - 1) * (z * z).sum(1) + lg
    _check_increment(st, lw, want, inc)


def test_guided_filter_against_kalman():
    # the reference's guided-filter test (1-D model, four observations, OptimalProposal)
    # with N = 50000 instead of 10^6; the tolerances allow the Monte Carlo error of the
    # prior draws (about 0.01 at this N, for the bootstrap filter as well)
    m = oracle.LG([0.2], [[1.0]], [[0.9]], [0.1], [[0.5]], [[1.2]], [-0.1], [[0.4]])
    rng = np.random.default_rng(5)
    _, _, ys = oracle.simulate_lg(m, 4, rng)
    mu_kf, _, ll_kf, _ = oracle.kf_filter(m, ys)
    prop = oracle.optimal_proposal(m)
    N = 50000
    st = oracle.BFState(oracle.bf_initialise(m, rng.standard_normal((N, 1))))
    ll = 0.0
    for t in range(4):
        _, inc, _, _ = oracle.guided_step(m, st, ys[t], 0.5, oracle.SYSTEMATIC, rng.random(1),
                                          rng.standard_normal((N, 1)), prop)
        ll += inc
    w = np.exp(st.logw - st.logw.max())
    w /= w.sum()
    assert abs(np.sum(w * st.x[:, 0]) - mu_kf[0]) < 0.02
    assert ll == pytest.approx(ll_kf, abs=0.03)
