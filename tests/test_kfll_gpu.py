"""Device kf_loglikelihood (GF/src/integrations/logdensity.jl:149-161) against the FP64
oracle: per-step parameters, jitter, one parameter set per filter, and the reference's
kf_loglikelihood == filter(KF) identity.  Tolerance as for the device Kalman filters: the
log-likelihood within 1e-4 relative (FP32 products, FP64 factorisation of S); filters up to
4 x 4 run one per lane entirely in FP64, within 1e-9."""
import numpy as np
import pytest

import ssmhip
from conftest import random_lg, to_ssmhip_model
from test_oracle_kfll import _tv_params

pytestmark = pytest.mark.gpu


def _batch(rng, B, Dx, Dy, T):
    pars = [_tv_params(rng, Dx, Dy, T) for _ in range(B)]
    ys = rng.standard_normal((B, T, Dy))
    st = lambda k: np.array([p[k] for p in pars])
    return pars, ys, [st(k) for k in range(8)]


@pytest.mark.parametrize("Dx,Dy,T,B", [(1, 1, 20, 300), (2, 2, 30, 1000), (4, 4, 15, 257),
                                       (3, 1, 10, 70), (4, 2, 12, 64), (5, 3, 9, 33),
                                       (20, 10, 8, 40), (48, 24, 6, 16)])
def test_time_varying_matches_oracle(ctx, orc, Dx, Dy, T, B):
    rng = np.random.default_rng(Dx + 100 * Dy)
    pars, ys, arrs = _batch(rng, B, Dx, Dy, T)
    ll = ssmhip.kf_loglikelihood_batch(*arrs, ys, ctx=ctx)
    want = np.array([orc.kf_loglikelihood(*p, y) for p, y in zip(pars, ys)])
    np.testing.assert_allclose(ll, want, rtol=_rtol(Dx, Dy))


def _rtol(Dx, Dy):
    # filters up to 4 x 4 run one per lane in FP64; larger ones in the FP32 matrix-core kernel
    return 1e-9 if Dx <= 4 and Dy <= 4 else 1e-4


@pytest.mark.parametrize("Dx,Dy", [(6, 3), (2, 2)])
def test_jitter_matches_oracle(ctx, orc, Dx, Dy):
    rng = np.random.default_rng(9)
    pars, ys, arrs = _batch(rng, 32, Dx, Dy, 10)
    ll = ssmhip.kf_loglikelihood_batch(*arrs, ys, jitter=0.02, ctx=ctx)
    want = np.array([orc.kf_loglikelihood(*p, y, jitter=0.02) for p, y in zip(pars, ys)])
    np.testing.assert_allclose(ll, want, rtol=_rtol(Dx, Dy))
    assert np.all(np.abs(ll - ssmhip.kf_loglikelihood_batch(*arrs, ys, ctx=ctx)) > 1e-3)


def test_constant_parameters_equal_kf_batch(ctx):
    # one parameter set per filter takes the same kernel body as the packed batch: the
    # same FP32 values in LDS, so the same log-likelihoods bit for bit
    rng = np.random.default_rng(4)
    ms = [random_lg(rng, 5, 3) for _ in range(24)]
    T = 9
    ys = rng.standard_normal((24, T, 3))
    kb = ssmhip.KFBatch([(m.mu0, m.Sigma0, m.A, m.b, m.Q, m.H, m.c, m.R) for m in ms], list(ys),
                        ctx=ctx)
    ll_b = kb.run()
    kb.close()
    st = lambda k: np.array([getattr(m, k) for m in ms])
    ll = ssmhip.kf_loglikelihood_batch(st("mu0"), st("Sigma0"), st("A"), st("b"), st("Q"),
                                       st("H"), st("c"), st("R"), ys, ctx=ctx)
    assert np.array_equal(ll, ll_b)


def test_reference_identity_with_filter(ctx, orc):
    # test/integrations/logdensity.jl:93-132: kf_loglikelihood with fill(A, T) ... equals
    # filter(model, KF(), ys) (here two kernels: one filter per lane for Dx, Dy <= 4 against
    # the workgroup kernel of the batch, so to FP32 rounding)
    rng = np.random.default_rng(1234)
    m = random_lg(rng, 2, 2)
    _, _, ys = orc.simulate_lg(m, 5, rng)
    rep = lambda M: [M] * 5
    ll = ssmhip.kf_loglikelihood(m.mu0, m.Sigma0, rep(m.A), rep(m.b), rep(m.Q), rep(m.H),
                                 rep(m.c), rep(m.R), ys, ctx=ctx)
    _, ll_f = ssmhip.filter(rng, to_ssmhip_model(m), ssmhip.KF(), ys, ctx=ctx)
    assert ll == pytest.approx(ll_f, rel=1e-5)
    assert ll == pytest.approx(orc.kf_filter(m, ys)[2], rel=1e-4)


@pytest.mark.parametrize("Dx,Dy", [(3, 2), (6, 5)])
def test_not_positive_definite_reports_status(ctx, Dx, Dy):
    rng = np.random.default_rng(5)
    _, ys, arrs = _batch(rng, 8, Dx, Dy, 4)
    arrs[7] = arrs[7].copy()
    arrs[7][3] = -np.eye(Dy)  # filter 3: R negative definite at every step
    ll, st = ssmhip.kf_loglikelihood_batch(*arrs, ys, ctx=ctx, return_status=True)
    assert st[3] == ssmhip._lib.SSM_ERR_NOT_PD and np.isnan(ll[3])
    assert np.all(st[np.arange(8) != 3] == 0) and np.all(np.isfinite(np.delete(ll, 3)))
    with pytest.raises(ssmhip.PosDefException):
        ssmhip.kf_loglikelihood_batch(*arrs, ys, ctx=ctx)


@pytest.mark.parametrize("Dx,Dy,T,B,tv", [(2, 2, 30, 1000, True), (3, 1, 10, 70, False),
                                          (5, 3, 9, 33, True)])
def test_device_resident_inputs_equal_host_inputs(ctx, Dx, Dy, T, B, tv):
    # ssm_kf_loglikelihood_device on inputs uploaded once equals the host-array entry point
    # bit for bit (the same kernels on the same filter-minor or caller-layout data), call
    # after call
    rng = np.random.default_rng(5 * Dx + T)
    pars, ys, arrs = _batch(rng, B, Dx, Dy, T)
    if not tv:  # one parameter set per filter: drop the step axis
        arrs = arrs[:2] + [a[:, 0] for a in arrs[2:]]
    want = ssmhip.kf_loglikelihood_batch(*arrs, ys, ctx=ctx)
    dev = ssmhip.DeviceKFInputs(*arrs, ys, ctx=ctx)
    for _ in range(2):
        assert np.array_equal(dev.loglikelihood(), want)
