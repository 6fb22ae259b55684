"""The oracle's Kalman log-likelihood gradients (GF/src/integrations/kalman_rrule.jl:22-107,
GF/src/algorithms/kalman_gradient.jl:103-263) against central finite differences of the
oracle's kf_loglikelihood, entry by entry (symmetric perturbations for Sigma0, Q and R)."""
import numpy as np
import pytest

import oracle
from test_oracle_kfll import _tv_params


def _fd(f, X, h=1e-6, sym=False):
    G = np.zeros_like(X)
    for idx in np.ndindex(X.shape):
        if sym and idx[0] > idx[1]:
            continue
        E = np.zeros_like(X)
        E[idx] = h
        if sym and idx[0] != idx[1]:
            E[idx[::-1]] = h
        d = (f(X + E) - f(X - E)) / (2 * h)
        if sym and idx[0] != idx[1]:
            d *= 0.5  # the symmetric gradient's entry: d f / d(e_ij + e_ji) = 2 G_ij
            G[idx[::-1]] = d
        G[idx] = d
    return G


@pytest.mark.parametrize("Dx,Dy,T,jit", [(2, 2, 4, 0.0), (3, 1, 3, 0.0), (2, 1, 3, 0.01)])
def test_gradients_match_finite_differences(Dx, Dy, T, jit):
    rng = np.random.default_rng(7 * Dx + Dy)
    mu0, S0, As, bs, Qs, Hs, cs, Rs = _tv_params(rng, Dx, Dy, T)
    ys = rng.standard_normal((T, Dy))
    args = [mu0, S0, As, bs, Qs, Hs, cs, Rs]
    ll, g = oracle.kf_loglikelihood_grad(*args, ys, jitter=jit)
    assert ll == pytest.approx(oracle.kf_loglikelihood(*args, ys, jitter=jit), rel=1e-12)

    def with_(pos, t, X):
        a = [list(v) if isinstance(v, list) else v for v in args]
        if t is None:
            a[pos] = X
        else:
            a[pos][t] = X
        return oracle.kf_loglikelihood(*a, ys, jitter=jit)

    tol = dict(rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(g["mu0"], _fd(lambda X: with_(0, None, X), mu0), **tol)
    np.testing.assert_allclose(g["Sigma0"], _fd(lambda X: with_(1, None, X), S0, sym=True),
                               **tol)
    for name, pos, sym in (("A", 2, False), ("b", 3, False), ("Q", 4, True), ("H", 5, False),
                           ("c", 6, False), ("R", 7, True)):
        for t in range(T):
            fd = _fd(lambda X: with_(pos, t, X), np.asarray(args[pos][t], dtype=np.float64), sym=sym)
            np.testing.assert_allclose(g[name][t], fd, **tol, err_msg=f"{name}[{t}]")
