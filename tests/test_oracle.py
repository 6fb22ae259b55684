"""CPU checks that pin the FP64 oracle (oracle/oracle.c) before it is trusted as a checker.

Each check is an independent fact the oracle must reproduce: published known-answer
vectors, exact-rational fixtures, closed forms, invariants, and the reference's own
statistical tests.  None of them retypes the oracle's formula.
"""
import math

import numpy as np
import pytest

from conftest import config1_model, rand_cov, random_lg


# ---- Philox4x32-10: Random123 known-answer vectors (Salmon et al. SC'11) ---------------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(orc, ctr, key, expect):
    assert list(orc.philox4x32_10(ctr, key)) == expect


def test_philox_normals_are_standard(orc):
    z = orc.philox_normals(42, 0, 7, 200000, 4)
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.01
    # Kolmogorov-Smirnov against N(0,1) on one component.
    from scipy import stats
    assert stats.kstest(z[:, 1], "norm").pvalue > 1e-3
    z64 = orc.philox_normals(42, 0, 7, 100000, 3, f64_layout=True)
    assert abs(z64.mean()) < 0.01 and abs(z64.var() - 1.0) < 0.02


def test_philox_uniforms_range(orc):
    u = orc.philox_uniforms(5, 3, 100000)
    assert u.min() >= 0.0 and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


# ---- resamplers -----------------------------------------------------------------------
def test_resample_fixtures_exact(orc, fixtures):
    for c in fixtures["cases"]:
        a = orc.resample(c["kind_id"], np.array(c["w"]), np.array(c["u"]))
        assert list(a) == c["ancestors"], (c["kind"], c["tag"])


def test_reference_known_answers(orc, fixtures):
    k = fixtures["known"]
    o = k["offspring_to_ancestors"]
    assert list(orc.offspring_to_ancestors(o["offspring"])) == o["ancestors"]
    a = k["ancestors_to_offspring"]
    assert list(orc.ancestors_to_offspring(a["ancestors"])) == a["offspring"]


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_uniform_weights_identity(orc, kind):
    # With w = 1/n (n a power of two) every systematic/stratified scheme keeps each particle.
    rng = np.random.default_rng(0)
    n = 1024
    u = rng.random(n)
    a = orc.resample(kind, np.full(n, 1.0 / n), u)
    assert np.array_equal(a, np.arange(n))


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_one_hot(orc, kind):
    n = 257
    w = np.zeros(n)
    w[123] = 1.0
    a = orc.resample(kind, w, np.random.default_rng(1).random(n))
    assert np.all(a == 123)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_invariants_and_statistics(orc, kind):
    # test/components/resamplers.jl:5-34: weights prop. to N(1,1)/N(0,1) density ratio;
    # the resampled mean matches the weighted mean (rtol 0.1).
    rng = np.random.default_rng(1234)
    n = 10**6
    xs = rng.standard_normal(n)
    ws = np.exp(xs - 0.5)
    ws /= ws.sum()
    u = rng.random(n)
    a = orc.resample(kind, ws, u)
    assert a.min() >= 0 and a.max() < n
    if kind != 3:
        assert np.all(np.diff(a) >= 0)
    off = orc.ancestors_to_offspring(a)
    assert off.sum() == n
    mu0 = (ws * xs).sum()
    mu1 = xs[a].mean()
    assert abs(mu1 - mu0) <= 0.1 * abs(mu0)


# ---- Kalman filter against the closed-form joint Gaussian ------------------------------
@pytest.mark.parametrize("Dy", [2, 3, 4])
def test_kf_matches_mvn_conditional(orc, Dy):
    # test/algorithms/kalman.jl:5-43: Dx = 3, T = 1; filtered moments equal the MVN
    # conditional of X1 given Y1, and ll equals log N(y; mu_Y, Sigma_Y).
    rng = np.random.default_rng(1234 + Dy)
    Dx = 3
    m = random_lg(rng, Dx, Dy)
    _, _, ys = orc.simulate_lg(m, 1, rng)
    mu, Sig, ll, _ = orc.kf_filter(m, ys)
    mu_Z, S_Z = orc.joint_gaussian(m, 1)
    ix = slice(Dx, 2 * Dx)
    iy = slice(2 * Dx, 2 * Dx + Dy)
    y = ys[0]
    mu_X1 = mu_Z[ix] + S_Z[ix, iy] @ np.linalg.solve(S_Z[iy, iy], y - mu_Z[iy])
    S_X1 = S_Z[ix, ix] - S_Z[ix, iy] @ np.linalg.solve(S_Z[iy, iy], S_Z[iy, ix])
    np.testing.assert_allclose(mu, mu_X1, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(Sig, S_X1, rtol=1e-9, atol=1e-12)
    assert ll == pytest.approx(orc.mvn_logpdf(y, mu_Z[iy], S_Z[iy, iy]), rel=1e-10)


@pytest.mark.parametrize("Dx,Dy,T", [(2, 2, 5), (3, 1, 8), (5, 3, 6)])
def test_kf_loglik_joint(orc, Dx, Dy, T):
    rng = np.random.default_rng(7 * Dx + Dy)
    m = random_lg(rng, Dx, Dy)
    _, _, ys = orc.simulate_lg(m, T, rng)
    _, _, ll, lls = orc.kf_filter(m, ys)
    assert ll == pytest.approx(orc.joint_gaussian_loglik(m, ys), rel=1e-9)
    assert ll == pytest.approx(lls.sum(), rel=1e-12)


def test_kf_not_pd_is_reported(orc):
    m = orc.LG([0.0], [[1.0]], [[1.0]], [0.0], [[1.0]], [[1.0]], [0.0], [[-5.0]])
    with pytest.raises(ValueError):
        orc.kf_filter(m, np.array([[0.0]]))


# ---- bootstrap filter -------------------------------------------------------------------
def test_bf_h_zero_gives_exact_ll(orc):
    # With H = 0 every particle gets the same weight log N(y; c, R), so ll_t is exact.
    rng = np.random.default_rng(3)
    m = orc.LG(np.zeros(2), np.eye(2), 0.9 * np.eye(2), np.zeros(2), 0.1 * np.eye(2),
               np.zeros((1, 2)), [0.5], [[0.3]])
    N = 512
    st = orc.BFState(orc.bf_initialise(m, rng.standard_normal((N, 2))))
    y = np.array([1.1])
    for t, thr in enumerate([1.0, 0.0, 0.5]):
        _, ll, _, _ = orc.bf_step(m, st, y, thr, 0, rng.random(1),
                                  rng.standard_normal((N, 2)))
        exact = -0.5 * (math.log(2 * math.pi * 0.3) + (1.1 - 0.5) ** 2 / 0.3)
        assert ll == pytest.approx(exact, abs=1e-12)


def test_bf_baseline_after_resample_and_zero_noise(orc):
    # Q -> 0 makes propagation deterministic: x' = A x[a] + b exactly.
    rng = np.random.default_rng(4)
    m = orc.LG(np.zeros(1), [[1.0]], [[0.7]], [0.2], [[1e-300]], [[1.0]], [0.0], [[0.5]])
    N = 64
    x0 = orc.bf_initialise(m, rng.standard_normal((N, 1)))
    st = orc.BFState(x0.copy())
    _, ll1, did, w = orc.bf_step(m, st, [0.3], 1.0, 0, [0.5], np.zeros((N, 1)))
    assert did and np.allclose(w, 1.0 / N)
    # uniform weights with u0 = 0.5 -> identity ancestors
    np.testing.assert_allclose(st.x[:, 0], 0.7 * x0[:, 0] + 0.2, rtol=1e-15)
    # ll_1 = log mean_i N(y; x_i', R)  (baseline log N after the resample)
    g = -0.5 * (np.log(2 * np.pi * 0.5) + (0.3 - st.x[:, 0]) ** 2 / 0.5)
    assert ll1 == pytest.approx(np.log(np.mean(np.exp(g))), rel=1e-12)


def test_bf_statistical_vs_kf(orc):
    # test/algorithms/particles.jl:5-27: 1-D LG model, T = 4, N = 10^6, alternating
    # resample / no resample; ll within 1e-3 of the exact KF and the filtered mean rtol 1e-3.
    rng = np.random.default_rng(1234)
    m = random_lg(rng, 1, 1)
    _, _, ys = orc.simulate_lg(m, 4, rng)
    mu_kf, _, ll_kf, _ = orc.kf_filter(m, ys)
    N = 10**6
    st = orc.BFState(orc.bf_initialise(m, rng.standard_normal((N, 1))))
    ll = 0.0
    for t in range(4):
        thr = 1.0 if t % 2 == 0 else 0.0
        _, inc, did, _ = orc.bf_step(m, st, ys[t], thr, 0, rng.random(1),
                                     rng.standard_normal((N, 1)))
        assert did == (t % 2 == 0)
        ll += inc
    w = np.exp(st.logw - st.logw.max())
    w /= w.sum()
    assert ll == pytest.approx(ll_kf, abs=1e-3)
    assert (w * st.x[:, 0]).sum() == pytest.approx(mu_kf[0], rel=1e-3)


def test_bf_config1_loglik_within_mc_error(orc):
    # 1-D LG model, T = 100, BF(1024; threshold = 0.5) systematic, FP64: over 20 seeds the
    # mean PF log-likelihood is within 3 standard errors of the exact KF value.
    m = config1_model()
    rng = np.random.default_rng(1234)
    _, _, ys = orc.simulate_lg(m, 100, rng)
    _, _, ll_kf, _ = orc.kf_filter(m, ys)
    N = 1024
    lls = []
    for rep in range(20):
        r = np.random.default_rng(100 + rep)
        st = orc.BFState(orc.bf_initialise(m, r.standard_normal((N, 1))))
        tot = 0.0
        for t in range(100):
            _, inc, _, _ = orc.bf_step(m, st, ys[t], 0.5, 0, r.random(1),
                                       r.standard_normal((N, 1)))
            tot += inc
        lls.append(tot)
    lls = np.array(lls)
    se = lls.std(ddof=1) / np.sqrt(len(lls))
    assert abs(lls.mean() - ll_kf) < 3 * se + 0.05


# ---- RBPF -------------------------------------------------------------------------------
def dummy_hierarchical(rng, obs_noise_scale=1.0):
    """GFTest/models/dummy_linear_gaussian.jl:48-154 with D_outer = D_inner = Dy = 1:
    returns the joint LG model and the matching hierarchical (RB) model."""
    import oracle
    mu0 = rng.random(2)
    S0 = np.zeros((2, 2))
    S0[0, 0] = rand_cov(rng, 1)[0, 0]
    S0[1, 1] = rand_cov(rng, 1)[0, 0]
    A = np.zeros((2, 2))
    A[0, 0] = rng.random()
    A[1, 0] = rng.random()
    A[1, 1] = rng.random()
    b = rng.random(2)
    Q = np.diag([rand_cov(rng, 1, 0.1)[0, 0], rand_cov(rng, 1, 0.1)[0, 0]])
    H = np.array([[0.0, rng.random()]])
    c = rng.random(1)
    R = rand_cov(rng, 1, obs_noise_scale)
    full = oracle.LG(mu0, S0, A, b, Q, H, c, R)
    rb = oracle.RB(mu0[0], S0[0, 0], A[0, 0], b[0], Q[0, 0], [mu0[1]], [[S0[1, 1]]], [[A[1, 1]]],
                   [b[1]], [A[1, 0]], [[Q[1, 1]]], [[H[0, 1]]], c, R, 0.0)
    return full, rb


def test_rbpf_statistical_vs_joint_kf(orc):
    # test/algorithms/rbpf.jl:5-34: RBPF on the hierarchical split of a joint LG model
    # matches the joint KF log-likelihood within 1e-3 (T = 4, alternating resampling).
    rng = np.random.default_rng(1234)
    full, rb = dummy_hierarchical(rng)
    _, _, ys = orc.simulate_lg(full, 4, rng)
    _, _, ll_kf, _ = orc.kf_filter(full, ys)
    N = 200000
    st = orc.rbpf_initialise(rb, rng.standard_normal(N))
    ll = 0.0
    for t in range(4):
        thr = 1.0 if t % 2 == 0 else 0.0
        _, inc, _, _ = orc.rbpf_step(rb, st, ys[t], thr, 0, rng.random(1), rng.standard_normal(N))
        ll += inc
    assert ll == pytest.approx(ll_kf, abs=3e-3)


def test_rbpf_degenerate_outer_equals_kf(orc):
    # Outer noise 0 and a point-mass outer prior: every particle carries the same x_t, so
    # the RBPF log-likelihood equals a KF whose observation noise is exp(x_t) R0 exactly.
    rng = np.random.default_rng(9)
    D, Dy, T = 3, 2, 6
    A = 0.8 * np.eye(D)
    Q = rand_cov(rng, D, 0.1)
    H = rng.standard_normal((Dy, D))
    R0 = rand_cov(rng, Dy, 0.5)
    rb = orc.RB(0.3, 0.0, 0.9, 0.05, 0.0, np.zeros(D), np.eye(D), A, np.zeros(D), np.zeros(D), Q,
                H, np.zeros(Dy), R0, 1.0)
    ys = rng.standard_normal((T, Dy))
    N = 16
    st = orc.rbpf_initialise(rb, np.zeros(N))
    ll = 0.0
    x = 0.3
    mu, S = np.zeros(D), np.eye(D)
    ll_kf = 0.0
    for t in range(T):
        _, inc, _, _ = orc.rbpf_step(rb, st, ys[t], 0.5, 0, [0.5], np.zeros(N))
        ll += inc
        x = 0.9 * x + 0.05
        mu, S, l = orc.kf_step(D, Dy, mu, S, A, np.zeros(D), Q, H, np.zeros(Dy),
                               math.exp(x) * R0, ys[t])
        ll_kf += l
    assert ll == pytest.approx(ll_kf, rel=1e-12)


def test_all_cores_baseline_tracks_the_kalman_filter(orc):
    # The OpenMP CPU baseline (bf_omp.c) is the same bootstrap filter: on the 1-D config-1
    # model its log-likelihood averaged over seeds is within 3 standard errors of the exact
    # KF value, and it resamples at threshold 1.0 on every step.
    m = config1_model()
    rng = np.random.default_rng(99)
    _, _, ys = orc.simulate_lg(m, 50, rng)
    _, _, ll_kf, _ = orc.kf_filter(m, ys)
    lls = []
    for rep in range(10):
        x0 = orc.bf_initialise(m, np.random.default_rng(rep).standard_normal((4096, 1)))
        ll, n_rs, x = orc.bfomp_run(m, x0, ys, 1.0, 1000 + rep, 4)
        assert n_rs == 50 and np.all(np.isfinite(x))
        lls.append(ll)
    se = np.std(lls, ddof=1) / np.sqrt(len(lls))
    assert abs(np.mean(lls) - ll_kf) < 3 * se + 0.05
