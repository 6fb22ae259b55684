"""Device square-root Kalman filter (GF/src/algorithms/srkf.jl:93-176) against the FP64
oracle's restatement, itself pinned to the Joseph-form KF: both FP64, so log-likelihoods,
final means and factors within 1e-9 relative (1e-11 absolute) up to 4 x 4 (one filter per
lane, Givens rotations); beyond (a workgroup per filter, Householder QR, up to Dx = 64,
Dy = 32) within 1e-8 relative plus 1e-10 of the array's largest entry.  Without jitter the
device SRKF also equals the oracle's Joseph-form KF on the same filters."""
import numpy as np
import pytest

import ssmhip
from conftest import config2_model, to_ssmhip_model
from test_kfll_gpu import _batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Dx,Dy,T,B,jit", [(1, 1, 20, 64, None), (2, 2, 12, 300, None),
                                           (4, 3, 8, 129, None), (3, 1, 10, 70, 0.01),
                                           (4, 4, 6, 65, 0.02), (2, 4, 7, 33, None)])
def test_srkf_matches_oracle(ctx, orc, Dx, Dy, T, B, jit):
    rng = np.random.default_rng(10 * Dx + Dy)
    pars, ys, arrs = _batch(rng, B, Dx, Dy, T)
    ll, mu, U = ssmhip.srkf_filter_batch(*arrs, ys, jitter=jit, ctx=ctx)
    tol = dict(rtol=1e-9, atol=1e-11)
    for f in range(0, B, max(1, B // 10)):
        ll_o, mu_o, U_o = orc.srkf_filter(*pars[f], ys[f], jitter=jit)
        assert ll[f] == pytest.approx(ll_o, rel=1e-9)
        np.testing.assert_allclose(mu[f], mu_o, **tol)
        np.testing.assert_allclose(U[f], U_o, **tol)
    assert np.allclose(U, np.triu(U))


def test_constant_parameters_equal_tiled(ctx):
    rng = np.random.default_rng(2)
    _, ys, arrs = _batch(rng, 20, 3, 2, 9)
    const = [arrs[0], arrs[1]] + [a[:, 0] for a in arrs[2:]]
    tiled = [arrs[0], arrs[1]] + [np.repeat(a[:, :1], 9, axis=1) for a in arrs[2:]]
    a = ssmhip.srkf_filter_batch(*const, ys, ctx=ctx)
    b = ssmhip.srkf_filter_batch(*tiled, ys, ctx=ctx)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_filter_srkf_equals_kalman(ctx, orc):
    # filter(rng, model, SRKF(), ys) against the oracle's KF on the D = 4 model
    m = config2_model()
    model = to_ssmhip_model(m)
    _, _, ys = orc.simulate_lg(m, 50, np.random.default_rng(6))
    st, ll = ssmhip.filter(np.random.default_rng(0), model, ssmhip.SRKF(), ys, ctx=ctx)
    mu_k, S_k, ll_k, _ = orc.kf_filter(m, ys)
    assert ll == pytest.approx(ll_k, rel=1e-10)
    np.testing.assert_allclose(st.mu, mu_k, rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(st.Sigma, S_k, rtol=1e-9, atol=1e-11)


def _joseph_kf(orc, pars, ys):
    mu0, S0, As, bs, Qs, Hs, cs, Rs = pars
    mu, S = np.array(mu0, dtype=float), np.array(S0, dtype=float)
    ll = 0.0
    for t in range(len(As)):
        mu, S, inc = orc.kf_step(len(mu), len(cs[t]), mu, S, As[t], bs[t], Qs[t], Hs[t], cs[t],
                                 Rs[t], ys[t])
        ll += inc
    return ll, mu, S


@pytest.mark.parametrize("Dx,Dy,T,B,jit", [(5, 2, 6, 9, None), (8, 4, 6, 7, None),
                                           (16, 8, 5, 4, 0.01), (33, 17, 4, 3, None),
                                           (64, 32, 3, 2, None), (3, 7, 5, 4, None)])
def test_workgroup_srkf_matches_oracle_and_kalman(ctx, orc, Dx, Dy, T, B, jit):
    rng = np.random.default_rng(3 * Dx + Dy)
    pars, ys, arrs = _batch(rng, B, Dx, Dy, T)
    ll, mu, U = ssmhip.srkf_filter_batch(*arrs, ys, jitter=jit, ctx=ctx)
    assert np.allclose(U, np.triu(U))
    for f in range(B):
        ll_o, mu_o, U_o = orc.srkf_filter(*pars[f], ys[f], jitter=jit)
        assert ll[f] == pytest.approx(ll_o, rel=1e-10)
        np.testing.assert_allclose(mu[f], mu_o, rtol=1e-8, atol=1e-10 * np.abs(mu_o).max())
        np.testing.assert_allclose(U[f], U_o, rtol=1e-8, atol=1e-10 * np.abs(U_o).max())
        if jit is None:  # the Joseph-form KF on the same filter
            ll_k, mu_k, S_k = _joseph_kf(orc, pars[f], ys[f])
            assert ll[f] == pytest.approx(ll_k, rel=1e-9)
            np.testing.assert_allclose(mu[f], mu_k, rtol=1e-7, atol=1e-9 * np.abs(mu_k).max())
            np.testing.assert_allclose(U[f].T @ U[f], S_k, rtol=1e-7,
                                       atol=1e-9 * np.abs(S_k).max())


def test_srkf_refuses_large_filters(ctx):
    rng = np.random.default_rng(4)
    _, ys, arrs = _batch(rng, 2, 65, 2, 2)
    with pytest.raises(ssmhip.ArgumentError):
        ssmhip.srkf_filter_batch(*arrs, ys, ctx=ctx)
    _, ys, arrs = _batch(rng, 2, 4, 33, 2)
    with pytest.raises(ssmhip.ArgumentError):
        ssmhip.srkf_filter_batch(*arrs, ys, ctx=ctx)
