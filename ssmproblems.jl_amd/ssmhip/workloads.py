"""Seeded synthetic workloads of the benchmark configurations (numpy, vectorised by size).

batched_lg(B, ...)  B independent linear-Gaussian SSMs: Dx ~ U{dx_lo..dx_hi}, Dy = Dx // 2
                    (at least 1), A = rho * (seeded orthogonal), Q and R seeded Wishart
                    (nu = dim + 2, scale I / nu), H ~ N(0, 1), mu0 = 0, Sigma0 = I, with T
                    observations simulated from each model.  Returns the packed arrays of
                    the ssm_kf_batch_create C-ABI.
"""
from __future__ import annotations

import numpy as np


def _orthogonal(rng, n, d):
    G = rng.standard_normal((n, d, d))
    Qm, Rm = np.linalg.qr(G)
    s = np.sign(np.diagonal(Rm, axis1=1, axis2=2))
    s[s == 0] = 1.0
    return Qm * s[:, None, :]


def _wishart(rng, n, d):
    nu = d + 2
    X = rng.standard_normal((n, d, nu))
    return X @ np.swapaxes(X, 1, 2) / nu


def _simulate_torch(A, Wn, x0):
    """x_t = A x_{t-1} + Wn[:, :, t] for all t, on the GPU through torch (data generation
    only; the filters never use torch)."""
    import torch
    dev = torch.device("cuda")
    At = torch.from_numpy(A).to(dev)
    Wt = torch.from_numpy(Wn).to(dev)
    x = torch.from_numpy(x0).to(dev)
    X = torch.empty_like(Wt)
    for t in range(Wn.shape[2]):
        x = torch.bmm(At, x) + Wt[:, :, t:t + 1]
        X[:, :, t:t + 1] = x
    return X.cpu().numpy()


def batched_lg(B, T, seed=1234, dx_lo=2, dx_hi=64, rho=0.9, dy_of=lambda d: max(1, d // 2),
               use_gpu=False):
    """Filters are stored in increasing-Dx order (a permutation of independent draws), so
    each size group packs contiguously.  use_gpu runs the state recursion of the data
    simulation through torch on the GPU (faster for large B; same arithmetic up to
    rounding)."""
    rng = np.random.default_rng(seed)
    Dx = np.sort(rng.integers(dx_lo, dx_hi + 1, size=B)).astype(np.int32)
    Dy = np.array([dy_of(int(d)) for d in Dx], dtype=np.int32)
    psize = 3 * Dx.astype(np.int64) ** 2 + 2 * Dx + Dy.astype(np.int64) * Dx + Dy + Dy.astype(np.int64) ** 2
    param_off = np.concatenate([[0], np.cumsum(psize)[:-1]]).astype(np.int64)
    ysize = T * Dy.astype(np.int64)
    y_off = np.concatenate([[0], np.cumsum(ysize)[:-1]]).astype(np.int64)
    params = np.empty(int(psize.sum()))
    ys = np.empty(int(ysize.sum()))
    for d in np.unique(Dx):
        idx = np.nonzero(Dx == d)[0]
        n, f0 = len(idx), int(idx[0])
        dy = int(Dy[f0])
        A = rho * _orthogonal(rng, n, d)
        Q = _wishart(rng, n, d)
        R = _wishart(rng, n, dy)
        H = rng.standard_normal((n, dy, d))
        LQ = np.linalg.cholesky(Q)
        LR = np.linalg.cholesky(R)
        # column-major packing: mu0, Sigma0, A, b, Q, H, c, R
        cm = lambda M: np.swapaxes(M, 1, 2).reshape(n, -1)
        packed = np.concatenate([np.zeros((n, d)), cm(np.broadcast_to(np.eye(d), (n, d, d))),
                                 cm(A), np.zeros((n, d)), cm(Q), cm(H), np.zeros((n, dy)), cm(R)],
                                axis=1)
        o = int(param_off[f0])
        params[o:o + packed.size] = packed.reshape(-1)
        # x0 ~ N(0, I); x_t = A x_{t-1} + LQ e_t; y_t = H x_t + LR v_t
        Wn = LQ @ rng.standard_normal((n, d, T))
        x = rng.standard_normal((n, d, 1))
        if use_gpu:
            X = _simulate_torch(A, Wn, x)
        else:
            X = np.empty((n, d, T))
            for t in range(T):
                x = A @ x + Wn[:, :, t:t + 1]
                X[:, :, t:t + 1] = x
        Y = H @ X + LR @ rng.standard_normal((n, dy, T))      # (n, dy, T)
        o = int(y_off[f0])
        ys[o:o + n * T * dy] = np.swapaxes(Y, 1, 2).reshape(-1)
    return dict(B=B, T=T, Dx=Dx, Dy=Dy, params=params, param_off=param_off, ys=ys, y_off=y_off)


def unpack_lg(w, f):
    """(mu0, Sigma0, A, b, Q, H, c, R) of filter f of a batched_lg workload."""
    dx, dy = int(w["Dx"][f]), int(w["Dy"][f])
    p = w["params"][w["param_off"][f]:]
    out, o = [], 0
    for shape in ((dx,), (dx, dx), (dx, dx), (dx,), (dx, dx), (dy, dx), (dy,), (dy, dy)):
        k = int(np.prod(shape))
        v = p[o:o + k]
        out.append(v.reshape(shape, order="F") if len(shape) == 2 else v.copy())
        o += k
    ys = w["ys"][w["y_off"][f]:w["y_off"][f] + w["T"] * dy].reshape(w["T"], dy)
    return out, ys
