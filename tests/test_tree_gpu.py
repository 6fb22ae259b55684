"""Device path storage (ParallelParticleTree, GF/ext/CUDAExt.jl:146-298) against the dense
history (DenseParticleContainer, GF/src/containers.jl:293-418).

Every leaf's path read from the sparse tree must equal, bit for bit, the path obtained by
walking the stored ancestors of the dense history back from the last generation
(get_ancestry, containers.jl:405-418).  After the last insert the tree holds exactly the
ancestral lines of the current particles: the nodes in use equal the number of distinct
(time, particle) pairs on those lines.
"""
import numpy as np
import pytest

import ssmhip
from conftest import random_lg, to_ssmhip_model

pytestmark = pytest.mark.gpu


def _dense_paths(xs, ancs):
    """xs[t]: (N, Dx) states at t = 0..T; ancs[t]: ancestors of generation t+1 (t = 0..T-1)."""
    T = len(ancs)
    N = xs[0].shape[0]
    idx = np.arange(N)
    paths = np.zeros((N, T + 1, xs[0].shape[1]))
    distinct = 0
    for t in range(T, -1, -1):
        paths[:, t, :] = xs[t][idx]
        distinct += len(np.unique(idx))
        if t > 0:
            idx = ancs[t - 1][idx]
    return paths, distinct


@pytest.mark.parametrize("prec,N,Dx,Dy,T,thr,kind", [("f32", 5000, 3, 2, 30, 0.6, 0),
                                                     ("f64", 3001, 2, 1, 25, 1.0, 2),
                                                     ("f32", 70000, 4, 2, 12, 0.5, 3)])
def test_tree_paths_equal_dense_history(ctx, prec, N, Dx, Dy, T, thr, kind):
    rng = np.random.default_rng(N + Dx)
    m = random_lg(rng, Dx, Dy)
    m.A = 0.9 * m.A / max(abs(np.linalg.eigvals(m.A)))
    model = to_ssmhip_model(m)
    res = {0: ssmhip.Systematic(), 2: ssmhip.Stratified(), 3: ssmhip.Multinomial()}[kind]
    algo = ssmhip.BF(N, threshold=thr, resampler=res, precision=prec)
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=11)
    # capacity N: the pool has to grow several times
    tree = ssmhip.ParallelParticleTree(st, capacity=N)
    xs, ancs = [st.export()[0]], []
    ys = rng.standard_normal((T, Dy))
    for t in range(1, T + 1):
        ssmhip.step(rng, model, algo, t, st, ys[t - 1])
        x, _, anc, _ = st.export()
        ancs.append(anc.astype(np.int64) if st.last_resampled else np.arange(N))
        xs.append(x)
        tree.insert()
    info = tree.info()
    paths = tree.ancestry()
    want, distinct = _dense_paths(xs, ancs)
    assert info["depth"] == T
    assert info["capacity"] > N  # it grew
    assert np.array_equal(paths, want)
    assert info["used"] == distinct, info
    tree.close()


def test_callback_with_filter(ctx):
    rng = np.random.default_rng(4)
    m = random_lg(rng, 2, 2)
    m.A = 0.8 * m.A / max(abs(np.linalg.eigvals(m.A)))
    model = to_ssmhip_model(m)
    algo = ssmhip.BF(2048, threshold=0.8, precision="f32")
    ys = rng.standard_normal((15, 2))
    cb = ssmhip.ParallelAncestorCallback(capacity=4096)
    state, ll = ssmhip.filter(rng, model, algo, ys, callback=cb, ctx=ctx)
    assert np.isfinite(ll)
    paths = cb.tree.ancestry()
    assert paths.shape == (2048, 16, 2)
    # the last point of every path is the particle itself
    assert np.array_equal(paths[:, -1, :], state.export()[0])
    cb.tree.close()


def test_bad_capacity(ctx):
    rng = np.random.default_rng(0)
    m = random_lg(rng, 2, 1)
    algo = ssmhip.BF(1000, precision="f32")
    st = ssmhip.initialise(rng, to_ssmhip_model(m), algo, ctx=ctx, seed=1)
    with pytest.raises(ssmhip.ArgumentError):
        ssmhip.ParallelParticleTree(st, capacity=999)
