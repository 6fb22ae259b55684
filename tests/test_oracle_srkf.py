"""The oracle's square-root Kalman filter (GF/src/algorithms/srkf.jl:93-176) against the
Joseph-form KF of the C oracle, the same filter in exact arithmetic: log-likelihood,
filtered mean and covariance U' U; the factor's shape; and the jitter rule (added to the
factor's diagonal, srkf.jl:143-146)."""
import numpy as np
import pytest

import oracle
from test_oracle_kfll import _tv_params


def _kf(pars, ys):
    mu0, S0, As, bs, Qs, Hs, cs, Rs = pars
    mu, S = np.asarray(mu0, float), np.asarray(S0, float)
    Dx, Dy = len(mu), np.atleast_2d(Hs[0]).shape[0]
    ll = 0.0
    for t in range(len(As)):
        mu, S, inc = oracle.kf_step(Dx, Dy, mu, S, As[t], bs[t], Qs[t], Hs[t], cs[t], Rs[t],
                                    ys[t])
        ll += inc
    return ll, mu, S


@pytest.mark.parametrize("Dx,Dy,T", [(1, 1, 12), (2, 2, 10), (4, 3, 8), (3, 1, 9), (5, 4, 6)])
def test_srkf_equals_kalman(Dx, Dy, T):
    rng = np.random.default_rng(Dx * 7 + Dy)
    pars = _tv_params(rng, Dx, Dy, T)
    ys = rng.standard_normal((T, Dy))
    ll, mu, U = oracle.srkf_filter(*pars, ys)
    ll_k, mu_k, S_k = _kf(pars, ys)
    assert ll == pytest.approx(ll_k, rel=1e-11)
    np.testing.assert_allclose(mu, mu_k, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(U.T @ U, S_k, rtol=1e-10, atol=1e-12)
    assert np.allclose(U, np.triu(U)) and np.all(np.diag(U) > 0)


def test_srkf_jitter_on_the_factor():
    rng = np.random.default_rng(3)
    pars = _tv_params(rng, 3, 2, 1)
    ys = rng.standard_normal((1, 2))
    _, _, U = oracle.srkf_filter(*pars, ys, jitter=0.05)
    _, _, S_k = _kf(pars, ys)
    np.testing.assert_allclose(U - 0.05 * np.eye(3), np.linalg.cholesky(S_k).T, rtol=1e-10,
                               atol=1e-12)
