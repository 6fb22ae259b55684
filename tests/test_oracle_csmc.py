"""The oracle's conditional SMC (GF/src/algorithms/csmc.jl:239-306): the ancestor-sampling
weights against scipy's density, the single categorical draw against a binary search, the
pinned step against the plain step, and the reference's statistical test of the CSMC and
CSMC-AS chains (test/algorithms/csmc.jl:5-95) against the exact smoothing mean."""
import numpy as np
import pytest
from scipy.stats import multivariate_normal

import oracle
from conftest import rand_cov


def _model(rng, D, Dy):
    A = 0.8 * np.linalg.qr(rng.standard_normal((D, D)))[0]
    return oracle.LG(rng.standard_normal(D), rand_cov(rng, D), A, 0.1 * rng.standard_normal(D),
                     rand_cov(rng, D, 0.5), rng.standard_normal((Dy, D)),
                     0.1 * rng.standard_normal(Dy), rand_cov(rng, Dy, 0.3))


def test_as_logweights_against_scipy():
    rng = np.random.default_rng(3)
    m = _model(rng, 3, 2)
    N = 64
    lw = rng.standard_normal(N)
    lw -= np.log(np.exp(lw).sum())
    st = oracle.BFState(rng.standard_normal((N, 3)), lw, typeless=False)
    x_ref = rng.standard_normal(3)
    got = oracle.as_logweights(m, st, x_ref)
    want = np.array([lw[i] + multivariate_normal.logpdf(x_ref, m.A @ st.x[i] + m.b, m.Q)
                     for i in range(N)])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11)
    # the lazy uniform weights of initialise contribute 0
    st.typeless = True
    np.testing.assert_allclose(oracle.as_logweights(m, st, x_ref), want - lw, atol=1e-11)


def test_categorical_is_a_binary_search():
    rng = np.random.default_rng(4)
    for n in (1, 2, 7, 300):
        lw = 3.0 * rng.standard_normal(n)
        w = np.exp(lw - lw.max())
        W = np.cumsum(w)
        for u in np.concatenate([[0.0, 1.0 - 2.0**-53], rng.random(20)]):
            want = min(int(np.searchsorted(W, u * W[-1], side="left")), n - 1)
            assert oracle.categorical(lw, u) == want
    # one-hot weights draw the hot particle for every u > 0 (u = 0 gives the first index
    # with cumulative weight >= 0, as StatsBase's linear search does)
    lw = np.full(9, -np.inf)
    lw[5] = 0.0
    assert [oracle.categorical(lw, u) for u in (2.0**-50, 0.3, 0.999, 0.0)] == [5, 5, 5, 0]


@pytest.mark.parametrize("threshold", [0.0, 1.0])
def test_pinned_step_equals_plain_step_off_the_reference(threshold):
    rng = np.random.default_rng(5)
    m = _model(rng, 2, 2)
    N = 40
    x0 = rng.standard_normal((N, 2))
    lw = rng.standard_normal(N)
    lw -= np.log(np.exp(lw).sum())
    u, z, y = rng.random(1), rng.standard_normal((N, 2)), rng.standard_normal(2)
    x_ref_t = np.array([0.25, -1.5])
    a = oracle.BFState(x0.copy(), lw, typeless=False)
    b = oracle.BFState(x0.copy(), lw, typeless=False)
    anc_a, _, did_a, _ = oracle.bf_step(m, a, y, threshold, oracle.SYSTEMATIC, u, z)
    anc_b, _, did_b, _ = oracle.csmc_step(m, b, y, threshold, oracle.SYSTEMATIC, u, z, x_ref_t, 7)
    assert did_a == did_b == (threshold == 1.0)
    np.testing.assert_array_equal(b.x[0], x_ref_t)
    np.testing.assert_array_equal(b.x[1:], a.x[1:])
    assert anc_b[0] == (7 if did_b else 0)
    np.testing.assert_array_equal(anc_b[1:], anc_a[1:])
    # the unnormalised log-weights agree off the reference: a constant offset after the
    # normalisation by different totals
    d = b.logw[1:] - a.logw[1:]
    np.testing.assert_allclose(d, d[0], atol=1e-12)


def _chain(m, ys, refresh, sweeps, burn, seed, N=10):
    rng = np.random.default_rng(seed)
    asamp, back = refresh == "ancestor", refresh == "backward"
    traj, _ = oracle.csmc_sweep(m, ys, N, None, asamp, 0.6, oracle.SYSTEMATIC, rng, back)
    trajs, lls = [], []
    for i in range(burn + sweeps):
        traj, ll = oracle.csmc_sweep(m, ys, N, traj, asamp, 0.6, oracle.SYSTEMATIC, rng, back)
        if i >= burn:
            trajs.append(traj[:, 0])
            lls.append(ll)
    return np.array(trajs), np.array(lls)


def batch_se(v, nbatch=40):
    """Batch-means standard error of the mean of a chain (along axis 0)."""
    v = np.asarray(v)
    bm = v[: len(v) // nbatch * nbatch].reshape(nbatch, -1, *v.shape[1:]).mean(axis=1)
    return bm.std(axis=0, ddof=1) / np.sqrt(nbatch)


@pytest.mark.parametrize("refresh", ["plain", "ancestor", "backward"])
def test_csmc_chain_matches_smoothing_mean(refresh):
    # the reference's setting (Dx = Dy = 1, K = 10, N = 10, threshold 0.6) with a fixed
    # model; the reference checks x_2 only, every time is checked here (dropping the
    # transition density from the ancestor weights biases x_0, x_1, x_3 and x_9 by 7-37 SE
    # while leaving x_2 within 2 SE)
    m = oracle.LG([0.2], [[1.0]], [[0.9]], [0.1], [[0.5]], [[1.2]], [-0.1], [[0.4]])
    ys = np.random.default_rng(11).standard_normal((10, 1)) + 0.5
    mean, cov = oracle.joint_posterior(m, ys)
    trajs, lls = _chain(m, ys, refresh, sweeps=6000, burn=200, seed=12)
    se = batch_se(trajs)
    sd = np.sqrt(np.diag(cov))
    assert np.all(se < 0.05 * sd), (se, sd)  # the chain mixes well enough to judge
    z = (trajs.mean(axis=0) - mean[:, 0]) / se
    assert np.all(np.abs(z) < 5), z
    # 1/Z-hat is unbiased for 1/Z (Elements of SMC, sec. 5.2; csmc.jl test)
    ll_exact = oracle.joint_gaussian_loglik(m, ys)
    L = -lls
    e = np.exp(L - L.max())
    rel_se = batch_se(e) / e.mean()  # ~ the SE of the log estimate
    lr = L.max() + np.log(e.mean())
    assert rel_se < 0.1, rel_se
    assert abs(lr + ll_exact) < 5 * rel_se, (lr, -ll_exact, rel_se)


def test_backward_sample_exact_distribution():
    # N = 3 particles, K = 2.  Index 2 ~ w_2; index 1 | i2 ~ softmax(lw_1 + log f(x2*[i2] | x1));
    # index 0 | i1 ~ softmax(log f(x1*[i1] | x0)).  Each conditional is checked by sweeping its
    # uniform over a fine grid with the later draws held at the middle of their intervals
    rng = np.random.default_rng(8)
    m = _model(rng, 2, 1)
    xs = [rng.standard_normal((3, 2)) for _ in range(3)]
    lws = []
    for _ in range(2):
        lw = rng.standard_normal(3)
        lws.append(lw - np.log(np.exp(lw).sum()))

    def softmax(v):
        e = np.exp(v - v.max())
        return e / e.sum()

    def logf(x_next, x):
        return multivariate_normal.logpdf(x_next, m.A @ x + m.b, m.Q)

    def mid(w, k):
        c = np.concatenate([[0.0], np.cumsum(w)])
        return 0.5 * (c[k] + c[k + 1])

    grid = (np.arange(400) + 0.5) / 400
    w2 = softmax(lws[1])
    for i2 in range(3):
        w1 = softmax(np.array([lws[0][i] + logf(xs[2][i2], xs[1][i]) for i in range(3)]))
        got = [oracle.backward_sample(m, xs, lws, [0.5, u1, mid(w2, i2)])[0] for u1 in grid]
        assert all(g[2] == i2 for g in got)
        freq = np.bincount([g[1] for g in got], minlength=3) / len(grid)
        np.testing.assert_allclose(freq, w1, atol=1.01 / len(grid))
        for i1 in range(3):
            w0 = softmax(np.array([logf(xs[1][i1], xs[0][i]) for i in range(3)]))
            got = [oracle.backward_sample(m, xs, lws, [u0, mid(w1, i1), mid(w2, i2)])
                   for u0 in grid]
            assert all(g[0][1] == i1 and g[0][2] == i2 for g in got)
            freq = np.bincount([g[0][0] for g in got], minlength=3) / len(grid)
            np.testing.assert_allclose(freq, w0, atol=1.01 / len(grid))
            for idx, traj in got[:5]:
                np.testing.assert_array_equal(traj, [xs[t][idx[t]] for t in range(3)])


def test_backward_sample_weights_time_t_with_the_dynamics_of_step_t_plus_1():
    # csmc.jl:424 weights time t with ancestor_weight(..., t + 1, ...).  Two particles at
    # time 0 (x = 0 and x = 10), both at 10 afterwards.  Step 1's dynamics (A = 1, tiny Q)
    # only reach x* = 10 from x = 10; step 2's (A = 0) treat both particles alike.  With the
    # right step the time-0 draw is particle 1 whatever the uniform; with step 2's dynamics
    # (or the homogeneous model below) u = 0.01 would pick particle 0.
    def lg(a):
        return oracle.LG([0.0], [[1.0]], [[a]], [0.0], [[1e-2]], [[1.0]], [0.0], [[1.0]])
    xs = [np.array([[0.0], [10.0]]), np.array([[10.0], [10.0]]), np.array([[10.0], [10.0]])]
    lws = [np.log([0.5, 0.5]), np.log([0.5, 0.5])]
    u = np.array([0.01, 0.3, 0.7])
    idx, _ = oracle.backward_sample(lg(0.0), xs, lws, u, dynamics=[lg(1.0), lg(0.0)])
    assert idx[0] == 1
    idx_h, _ = oracle.backward_sample(lg(0.0), xs, lws, u)
    assert idx_h[0] == 0
    # constant dynamics reduce to the homogeneous pass
    rng = np.random.default_rng(3)
    m = lg(0.9)
    xs = [rng.standard_normal((50, 1)) for _ in range(5)]
    lws = [np.log(rng.dirichlet(np.ones(50))) for _ in range(4)]
    u = rng.random(5)
    a, ta = oracle.backward_sample(m, xs, lws, u)
    b, tb = oracle.backward_sample(m, xs, lws, u, dynamics=[m] * 4)
    assert np.array_equal(a, b) and np.array_equal(ta, tb)
