"""The oracle's kf_loglikelihood (GF/src/integrations/logdensity.jl:149-161) against the
joint Gaussian with per-step parameters, with and without jitter, and against kf_filter for
constant parameters (the reference's kf_loglikelihood == filter(KF) test,
test/integrations/logdensity.jl:93-132)."""
import numpy as np
import pytest

import oracle
from conftest import rand_cov, random_lg


def _tv_params(rng, Dx, Dy, T):
    As = [0.9 * np.linalg.qr(rng.standard_normal((Dx, Dx)))[0] for _ in range(T)]
    bs = [0.1 * rng.standard_normal(Dx) for _ in range(T)]
    Qs = [rand_cov(rng, Dx, 0.3) for _ in range(T)]
    Hs = [rng.standard_normal((Dy, Dx)) for _ in range(T)]
    cs = [0.1 * rng.standard_normal(Dy) for _ in range(T)]
    Rs = [rand_cov(rng, Dy, 0.5) for _ in range(T)]
    return rng.standard_normal(Dx), rand_cov(rng, Dx), As, bs, Qs, Hs, cs, Rs


def _joint_ll(mu0, S0, As, bs, Qs, Hs, cs, Rs, ys):
    mu_Z, Sig_Z = oracle.joint_gaussian_tv(mu0, S0, As, bs, Qs, Hs, cs, Rs)
    Dx = len(mu0)
    iy = slice(Dx * (len(As) + 1), None)
    return oracle.mvn_logpdf(np.asarray(ys).reshape(-1), mu_Z[iy], Sig_Z[iy, iy])


@pytest.mark.parametrize("Dx,Dy,T", [(1, 1, 6), (3, 2, 5), (4, 3, 7)])
def test_time_varying_against_joint_gaussian(Dx, Dy, T):
    rng = np.random.default_rng(10 * Dx + Dy)
    par = _tv_params(rng, Dx, Dy, T)
    ys = rng.standard_normal((T, Dy))
    ll = oracle.kf_loglikelihood(*par, ys)
    assert ll == pytest.approx(_joint_ll(*par, ys), rel=1e-10, abs=1e-10)


def test_jitter_is_extra_process_noise():
    # A (S + j I) A' + Q = A S A' + (Q + j A A'): jitter j after each update equals the
    # jitter-free filter with Q_t + j A_t A_t' from the second step on
    rng = np.random.default_rng(3)
    Dx, Dy, T, j = 3, 2, 6, 0.05
    mu0, S0, As, bs, Qs, Hs, cs, Rs = _tv_params(rng, Dx, Dy, T)
    ys = rng.standard_normal((T, Dy))
    ll = oracle.kf_loglikelihood(mu0, S0, As, bs, Qs, Hs, cs, Rs, ys, jitter=j)
    Qj = [Qs[0]] + [Qs[t] + j * As[t] @ As[t].T for t in range(1, T)]
    assert ll == pytest.approx(_joint_ll(mu0, S0, As, bs, Qj, Hs, cs, Rs, ys), rel=1e-10)
    assert abs(ll - oracle.kf_loglikelihood(mu0, S0, As, bs, Qs, Hs, cs, Rs, ys)) > 1e-3


def test_constant_parameters_equal_kf_filter():
    rng = np.random.default_rng(1234)
    m = random_lg(rng, 2, 2)
    _, _, ys = oracle.simulate_lg(m, 5, rng)
    rep = lambda M: [M] * 5
    ll = oracle.kf_loglikelihood(m.mu0, m.Sigma0, rep(m.A), rep(m.b), rep(m.Q), rep(m.H),
                                 rep(m.c), rep(m.R), ys)
    assert ll == pytest.approx(oracle.kf_filter(m, ys)[2], rel=1e-13)
