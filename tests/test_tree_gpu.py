"""Device path storage (ParallelParticleTree, GF/ext/CUDAExt.jl:146-298) against the
oracle's sequential tree and the dense history (DenseParticleContainer,
GF/src/containers.jl:328-418).

The device tree's node pool (parent links, offspring counts, leaf slots) must equal the
oracle tree's bit for bit after every insert, given the same ancestors.  Every leaf's path
must equal the path obtained by walking the stored ancestors of the dense history back from
the last generation (get_ancestry, containers.jl:406-418), and the nodes in use must equal
the number of distinct (time, particle) pairs on those lines.
"""
import numpy as np
import pytest

import oracle
import ssmhip
from conftest import random_lg, to_ssmhip_model

pytestmark = pytest.mark.gpu


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
    otree = oracle.PTree(xs[0], capacity=N)
    ys = rng.standard_normal((T, Dy))
    for t in range(1, T + 1):
        ssmhip.step(rng, model, algo, t, st, ys[t - 1])
        x, _, anc, _ = st.export()
        ancs.append(anc.astype(np.int64) if st.last_resampled else np.arange(N))
        xs.append(x)
        tree.insert()
        otree.insert(x, ancs[-1])
        if t % 5 == 0 or t == T:
            for got, want in zip(tree.export(), otree.export()):
                assert np.array_equal(got, want), f"pool differs after insert {t}"
    info = tree.info()
    paths = tree.ancestry()
    want, distinct = oracle.dense_ancestry(xs, ancs)
    assert info["depth"] == T
    assert info["capacity"] > N  # it grew
    assert np.array_equal(paths, want)
    assert np.array_equal(paths, otree.ancestry())
    assert info["used"] == distinct, info
    for i in (0, N // 2, N - 1):
        assert np.array_equal(tree.path(i), want[i])
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
    tree = ssmhip.ParallelParticleTree(st, capacity=1000)
    with pytest.raises(ssmhip.ArgumentError):
        tree.path(1000)
    assert np.array_equal(tree.path(7)[0], st.export()[0][7])
    tree.close()


def test_one_tree_per_filter_and_lifetimes(ctx):
    rng = np.random.default_rng(3)
    m = random_lg(rng, 2, 1)
    model = to_ssmhip_model(m)
    algo = ssmhip.BF(4096, threshold=1.0, precision="f32")
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=5)
    tree = ssmhip.ParallelParticleTree(st, capacity=8192)
    with pytest.raises(ssmhip.SSMError) as e:
        ssmhip.ParallelParticleTree(st, capacity=8192)
    assert e.value.code == ssmhip._lib.SSM_ERR_STATE
    ssmhip.step(rng, model, algo, 1, st, np.zeros(1))
    tree.insert()
    tree.close()
    # a new tree can be attached once the old one is gone; the filter outliving it
    tree = ssmhip.ParallelParticleTree(st, capacity=8192)
    st.close()
    with pytest.raises(ssmhip.SSMError) as e:
        tree.insert()
    assert e.value.code == ssmhip._lib.SSM_ERR_STATE
    tree.close()


def test_steps_from_run_use_counted_children(ctx):
    """After ssm_bf_run (graph replay, no resampler counts) the insert counts the children
    itself; the pool must still equal the oracle's."""
    rng = np.random.default_rng(8)
    m = random_lg(rng, 3, 2)
    m.A = 0.9 * m.A / max(abs(np.linalg.eigvals(m.A)))
    model = to_ssmhip_model(m)
    N = 20000
    algo = ssmhip.BF(N, threshold=1.0, precision="f32")
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=9)
    tree = ssmhip.ParallelParticleTree(st, capacity=N)
    otree = oracle.PTree(st.export()[0], capacity=N)
    ys = rng.standard_normal((6, 2))
    for t in range(1, 7):
        if t % 2:
            ssmhip.step(rng, model, algo, t, st, ys[t - 1])
        else:
            ssmhip.run(st, algo, ys[t - 1:t], t0=t)
        x, _, anc, _ = st.export()
        tree.insert()
        otree.insert(x, anc.astype(np.int64))
    for got, want in zip(tree.export(), otree.export()):
        assert np.array_equal(got, want)
    assert np.array_equal(tree.ancestry(), otree.ancestry())
    tree.close()
