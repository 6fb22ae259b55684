"""Particle Gibbs (GF/src/integrations/particle_gibbs.jl) over the device CSMC: the host
trajectory log-density against the joint Gaussian (CPU), and the parameter chain against
the exact posterior of the reference's own test model (GPU)."""
import numpy as np
import pytest
from scipy.stats import norm

import oracle
import ssmhip


def _model(b, a=0.8, q2=0.1, r2=0.5, s02=1.0):
    # the reference's PG test model (test/integrations/particle_gibbs.jl:5-40)
    return ssmhip.create_homogeneous_linear_gaussian_model(
        [0.0], [[s02]], [[a]], [b], [[q2]], [[1.0]], [0.0], [[r2]])


def test_trajectory_logdensity_is_the_joint_density():
    # logdensity.jl:19-29: log p(x_0:T, y_1:T) of an LG model equals the joint Gaussian's
    # density at the stacked vector [x_0..x_T, y_1..y_T]
    rng = np.random.default_rng(2)
    m = oracle.LG(rng.standard_normal(2), np.eye(2) * 0.7, 0.9 * np.eye(2) + 0.05,
                  rng.standard_normal(2), np.eye(2) * 0.3, rng.standard_normal((1, 2)),
                  [0.2], [[0.4]])
    model = ssmhip.create_homogeneous_linear_gaussian_model(m.mu0, m.Sigma0, m.A, m.b, m.Q,
                                                            m.H, m.c, m.R)
    T = 6
    xs = rng.standard_normal((T + 1, 2))
    ys = rng.standard_normal((T, 1))
    mu, S = oracle.joint_gaussian(m, T)
    want = oracle.mvn_logpdf(np.concatenate([xs.reshape(-1), ys.reshape(-1)]), mu, S)
    got = ssmhip.trajectory_logdensity(model, xs, ys)
    assert got == pytest.approx(want, rel=1e-10, abs=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("refresh", ["plain", "ancestor"])
def test_pg_chain_matches_exact_posterior(ctx, refresh):
    rng = np.random.default_rng(1234)
    m_true = oracle.LG([0.0], [[1.0]], [[0.8]], [1.0], [[0.1]], [[1.0]], [0.0], [[0.5]])
    _, _, ys = oracle.simulate_lg(m_true, 5, rng)
    prior = norm(0.0, 2.0)
    # exact posterior of b on a grid: prior times the Kalman likelihood
    grid = np.linspace(-8, 10, 3601)
    lp = prior.logpdf(grid) + np.array([
        oracle.kf_filter(oracle.LG([0.0], [[1.0]], [[0.8]], [g], [[0.1]], [[1.0]], [0.0],
                                   [[0.5]]), ys)[2] for g in grid])
    w = np.exp(lp - lp.max())
    w /= w.sum()
    mean, sd = (grid * w).sum(), np.sqrt(((grid - (grid * w).sum()) ** 2 * w).sum())
    refreshment = ssmhip.AncestorSampling() if refresh == "ancestor" else ssmhip.NoRefreshment()
    pg = ssmhip.ParticleGibbs(ssmhip.ConditionalSMC(ssmhip.BF(10, threshold=0.6), refreshment),
                              ssmhip.RandomWalkMH(scale=0.8, steps=3))
    model = ssmhip.ParticleGibbsModel(prior, ssmhip.ParameterisedSSM(lambda th: _model(th[0]), ys))
    thetas, trans, _ = ssmhip.pg_sample(rng, model, pg, 4200, ctx=ctx)
    th = thetas[200:, 0]
    bm = th.reshape(40, -1).mean(axis=1)
    se = bm.std(ddof=1) / np.sqrt(40)
    assert se < 0.1 * sd, (se, sd)
    assert abs(th.mean() - mean) < 5 * se, (th.mean(), mean, se)
    assert abs(th.std() - sd) < 0.15 * sd, (th.std(), sd)
    assert 0.05 < np.mean([t.stat["acceptance_rate"] for t in trans]) < 0.95
