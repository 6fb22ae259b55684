"""Device square-root Kalman filter (GF/src/algorithms/srkf.jl:93-176) against the FP64
oracle's restatement, itself pinned to the Joseph-form KF: both FP64, so log-likelihoods,
final means and factors within 1e-9 relative (1e-11 absolute)."""
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


def test_srkf_refuses_large_filters(ctx):
    rng = np.random.default_rng(4)
    _, ys, arrs = _batch(rng, 4, 5, 2, 3)
    with pytest.raises(ssmhip.ArgumentError):
        ssmhip.srkf_filter_batch(*arrs, ys, ctx=ctx)
