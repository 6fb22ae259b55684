"""Device Kalman filters (single and batched, uneven sizes) against the FP64 oracle.

The device computes products in FP32 and factorises S in FP64; stated tolerances:
log-likelihood relative error <= 1e-4 against the FP64 oracle (the target bar), filtered
means / covariances rtol 2e-3 relative to their scale (FP32 storage of Sigma over T steps).
"""
import numpy as np
import pytest

import ssmhip
from conftest import random_lg, to_ssmhip_model
from ssmhip import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Dy", [2, 3, 4])
def test_kf_matches_mvn_conditional(ctx, orc, Dy):
    # test/algorithms/kalman.jl:5-43 on the device: Dx = 3, T = 1.
    rng = np.random.default_rng(1234 + Dy)
    Dx = 3
    m = random_lg(rng, Dx, Dy)
    _, _, ys = orc.simulate_lg(m, 1, rng)
    st, ll = ssmhip.filter(rng, to_ssmhip_model(m), ssmhip.KF(), ys, ctx=ctx)
    mu_Z, S_Z = orc.joint_gaussian(m, 1)
    ix = slice(Dx, 2 * Dx)
    iy = slice(2 * Dx, 2 * Dx + Dy)
    y = ys[0]
    mu_X1 = mu_Z[ix] + S_Z[ix, iy] @ np.linalg.solve(S_Z[iy, iy], y - mu_Z[iy])
    S_X1 = S_Z[ix, ix] - S_Z[ix, iy] @ np.linalg.solve(S_Z[iy, iy], S_Z[iy, ix])
    np.testing.assert_allclose(st.mu, mu_X1, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(st.Sigma, S_X1, rtol=1e-4, atol=1e-5)
    assert ll == pytest.approx(orc.mvn_logpdf(y, mu_Z[iy], S_Z[iy, iy]), rel=1e-4)


@pytest.mark.parametrize("Dx,Dy,T", [(2, 2, 5), (3, 1, 8), (5, 3, 6)])
def test_kf_loglik_joint_closed_form(ctx, orc, Dx, Dy, T):
    rng = np.random.default_rng(7 * Dx + Dy)
    m = random_lg(rng, Dx, Dy)
    _, _, ys = orc.simulate_lg(m, T, rng)
    _, ll = ssmhip.filter(rng, to_ssmhip_model(m), ssmhip.KF(), ys, ctx=ctx)
    assert ll == pytest.approx(orc.joint_gaussian_loglik(m, ys), rel=1e-4)


def _check_batch(ctx, orc, w, rel=1e-4, check_state=True, sample=None):
    kb = ssmhip.KFBatch.from_packed(w["Dx"], w["Dy"], w["params"], w["param_off"], w["ys"],
                                    w["y_off"], w["T"], ctx=ctx)
    ll, mus, sigs = kb.run(return_state=True)
    kb.close()
    worst = 0.0
    idx = range(w["B"]) if sample is None else sample
    for f in idx:
        p, ys = W.unpack_lg(w, f)
        mu_o, S_o, ll_o, _ = orc.kf_filter(orc.LG(*p), ys)
        err = abs(ll[f] - ll_o) / abs(ll_o)
        worst = max(worst, err)
        assert err <= rel, (f, int(w["Dx"][f]), int(w["Dy"][f]), ll[f], ll_o)
        if check_state:
            scale = np.abs(S_o).max()
            np.testing.assert_allclose(sigs[f], S_o, atol=2e-3 * scale, rtol=0)
            np.testing.assert_allclose(mus[f], mu_o, atol=2e-3 * max(1.0, np.abs(mu_o).max()),
                                       rtol=0)
    return worst


def test_batch_every_size_class(ctx, orc):
    # All Dx in 1..64 (Dy = Dx // 2, at least 1) across the three size classes, T = 50.
    w = W.batched_lg(192, 50, seed=11, dx_lo=1, dx_hi=64)
    worst = _check_batch(ctx, orc, w)
    print(f"worst ll rel err {worst:.2e}")


def test_batch_config3_shape(ctx, orc):
    # The batched-KF benchmark configuration at reduced B: Dx ~ U{2..64}, Dy = Dx/2,
    # A spectral radius 0.9, Wishart Q and R, T = 200; log-likelihood error <= 1e-4 relative.
    w = W.batched_lg(1024, 200, seed=3)
    sample = list(range(0, 1024, 7))
    worst = _check_batch(ctx, orc, w, check_state=False, sample=sample)
    print(f"worst ll rel err {worst:.2e}")


def test_batch_independence_of_mix(ctx, orc):
    # A filter's result does not depend on what else is in the batch (one workgroup per
    # filter, fixed arithmetic order): bit-identical alone and in a mixed batch.
    w = W.batched_lg(300, 30, seed=5)
    kb = ssmhip.KFBatch.from_packed(w["Dx"], w["Dy"], w["params"], w["param_off"], w["ys"],
                                    w["y_off"], w["T"], ctx=ctx)
    ll_all = kb.run()
    for f in (0, 150, 299):
        p, ys = W.unpack_lg(w, f)
        one = ssmhip.KFBatch([p], [ys], ctx=ctx).run()
        assert one[0] == ll_all[f]


def test_non_pd_reported(ctx):
    p = (np.zeros(2), np.eye(2), 0.5 * np.eye(2), np.zeros(2), 0.1 * np.eye(2),
         np.ones((1, 2)), np.zeros(1), -10.0 * np.eye(1))
    kb = ssmhip.KFBatch([p], [np.zeros((3, 1))], ctx=ctx)
    with pytest.raises(ssmhip.PosDefException):
        kb.run()


def test_autotune_keeps_results(ctx, orc):
    # every workgroup shape runs the same per-tile and per-column arithmetic, so the tuned
    # kernels reproduce the default kernels' log-likelihoods bit for bit
    w = W.batched_lg(400, 20, seed=21, dx_lo=1, dx_hi=64)
    kb = ssmhip.KFBatch.from_packed(w["Dx"], w["Dy"], w["params"], w["param_off"], w["ys"],
                                    w["y_off"], w["T"], ctx=ctx)
    before = kb.run()
    tuned = kb.tune(T_tune=4)
    after = kb.run()
    kb.close()
    for name, v in tuned.items():
        assert v["waves"] in v["ms"], (name, v)
    assert np.array_equal(before, after)
