"""Device gradients of kf_loglikelihood (GF/src/integrations/kalman_rrule.jl:22-107) against
the FP64 oracle's restatement, itself pinned to finite differences: both are FP64, so up to
4 x 4 (one filter per lane) log-likelihoods and every gradient entry within 1e-9 relative
(1e-10 absolute); beyond (one workgroup per filter, up to Dx = 64, Dy = 32) within 1e-8
relative plus 1e-9 of the array's largest entry, the sums running over up to 64 terms in
another order than numpy's."""
import numpy as np
import pytest

import ssmhip
from test_kfll_gpu import _batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Dx,Dy,T,B,jit", [(2, 2, 12, 300, 0.0), (1, 1, 20, 64, 0.0),
                                           (4, 3, 8, 129, 0.0), (3, 1, 10, 70, 0.01),
                                           (4, 4, 6, 65, 0.0)])
def test_gradients_match_oracle(ctx, orc, Dx, Dy, T, B, jit):
    rng = np.random.default_rng(100 * Dx + Dy)
    pars, ys, arrs = _batch(rng, B, Dx, Dy, T)
    ll, g = ssmhip.kf_loglikelihood_grad_batch(*arrs, ys, jitter=jit, ctx=ctx)
    assert g["A"].shape == (B, T, Dx, Dx) and g["R"].shape == (B, T, Dy, Dy)
    for f in range(0, B, max(1, B // 12)):
        ll_o, go = orc.kf_loglikelihood_grad(*pars[f], ys[f], jitter=jit)
        assert ll[f] == pytest.approx(ll_o, rel=1e-9)
        tol = dict(rtol=1e-9, atol=1e-10)
        np.testing.assert_allclose(g["mu0"][f], go["mu0"], **tol)
        np.testing.assert_allclose(g["Sigma0"][f], go["Sigma0"], **tol)
        for k in ("A", "b", "Q", "H", "c", "R"):
            np.testing.assert_allclose(g[k][f], np.array(go[k]).reshape(g[k][f].shape), **tol,
                                       err_msg=k)
    # the plain entry point gives the same log-likelihoods
    np.testing.assert_allclose(ssmhip.kf_loglikelihood_batch(*arrs, ys, jitter=jit, ctx=ctx), ll,
                               rtol=1e-12)


def test_constant_parameters_sum_the_steps(ctx, orc):
    # one parameter set per filter: its gradient is the sum over the steps of the per-step
    # gradients (fill(A, T) in the reference's rrule, then summed)
    rng = np.random.default_rng(3)
    pars, ys, arrs = _batch(rng, 40, 2, 2, 9)
    const = [arrs[0], arrs[1]] + [a[:, 0] for a in arrs[2:]]
    ll, g = ssmhip.kf_loglikelihood_grad_batch(*const, ys, ctx=ctx)
    tiled = [arrs[0], arrs[1]] + [np.repeat(a[:, :1], 9, axis=1) for a in arrs[2:]]
    ll_t, g_t = ssmhip.kf_loglikelihood_grad_batch(*tiled, ys, ctx=ctx)
    np.testing.assert_allclose(ll, ll_t, rtol=1e-13)
    for k in ("A", "b", "Q", "H", "c", "R"):
        np.testing.assert_allclose(g[k], g_t[k].sum(axis=1), rtol=1e-11, atol=1e-12, err_msg=k)


@pytest.mark.parametrize("Dx,Dy,T,B,jit", [(5, 2, 6, 9, 0.0), (8, 4, 6, 7, 0.0),
                                           (16, 8, 5, 3, 0.01), (33, 17, 4, 2, 0.0),
                                           (64, 32, 3, 2, 0.0), (3, 7, 5, 4, 0.0)])
def test_workgroup_gradients_match_oracle(ctx, orc, Dx, Dy, T, B, jit):
    rng = np.random.default_rng(7 * Dx + Dy)
    pars, ys, arrs = _batch(rng, B, Dx, Dy, T)
    ll, g = ssmhip.kf_loglikelihood_grad_batch(*arrs, ys, jitter=jit, ctx=ctx)
    assert g["A"].shape == (B, T, Dx, Dx) and g["H"].shape == (B, T, Dy, Dx)
    for f in range(B):
        ll_o, go = orc.kf_loglikelihood_grad(*pars[f], ys[f], jitter=jit)
        assert ll[f] == pytest.approx(ll_o, rel=1e-10)
        for k in ("mu0", "Sigma0", "A", "b", "Q", "H", "c", "R"):
            want = np.array(go[k]).reshape(g[k][f].shape)
            np.testing.assert_allclose(g[k][f], want, rtol=1e-8,
                                       atol=1e-9 * np.abs(want).max(), err_msg=k)
    # the FP32 matrix-core kf_loglikelihood agrees to its own tolerance
    np.testing.assert_allclose(ssmhip.kf_loglikelihood_batch(*arrs, ys, jitter=jit, ctx=ctx),
                               ll, rtol=1e-4)


def test_workgroup_constant_parameters_sum_the_steps(ctx):
    rng = np.random.default_rng(8)
    pars, ys, arrs = _batch(rng, 3, 6, 5, 7)
    const = [arrs[0], arrs[1]] + [a[:, 0] for a in arrs[2:]]
    ll, g = ssmhip.kf_loglikelihood_grad_batch(*const, ys, ctx=ctx)
    tiled = [arrs[0], arrs[1]] + [np.repeat(a[:, :1], 7, axis=1) for a in arrs[2:]]
    ll_t, g_t = ssmhip.kf_loglikelihood_grad_batch(*tiled, ys, ctx=ctx)
    np.testing.assert_array_equal(ll, ll_t)
    for k in ("A", "b", "Q", "H", "c", "R"):
        np.testing.assert_allclose(g[k], g_t[k].sum(axis=1), rtol=1e-11,
                                   atol=1e-12 * np.abs(g[k]).max(), err_msg=k)


def test_gradient_refuses_large_filters(ctx):
    rng = np.random.default_rng(4)
    _, ys, arrs = _batch(rng, 2, 65, 2, 2)
    with pytest.raises(ssmhip.ArgumentError):
        ssmhip.kf_loglikelihood_grad_batch(*arrs, ys, ctx=ctx)
    _, ys, arrs = _batch(rng, 2, 4, 33, 2)
    with pytest.raises(ssmhip.ArgumentError):
        ssmhip.kf_loglikelihood_grad_batch(*arrs, ys, ctx=ctx)
