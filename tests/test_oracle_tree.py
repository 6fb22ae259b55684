"""The oracle's path storage (sequential ParallelParticleTree, GF/ext/CUDAExt.jl:146-262)
against a hand-worked insert sequence and against the brute-force walk of the dense
ancestor history (DenseParticleContainer's get_ancestry, GF/src/containers.jl:406-418)."""
import numpy as np
import pytest

import oracle


def _used(par, off, N):
    # nodes in use after an insert: the N leaves (count 0) and every node with children
    return int(np.sum(off != 0)) + N


def test_hand_worked_inserts():
    # N = 3 roots in a pool of 3: the first insert has one free slot and must double it
    t = oracle.PTree(np.array([[0.0], [1.0], [2.0]]), capacity=3)
    assert t.insert(np.array([[10.0], [11.0], [12.0]]), [0, 0, 2]) == 1
    par, off, leaves = t.export()
    # root 1 is childless and freed; the new particles take the free slots 1, 3, 4
    assert par.tolist() == [-1, 0, -1, 0, 2, -1]
    assert off.tolist() == [2, 0, 1, 0, 0, 0]
    assert leaves.tolist() == [1, 3, 4]
    assert t.insert(np.array([[20.0], [21.0], [22.0]]), [2, 2, 2]) == 0
    par, off, leaves = t.export()
    # slots 1 and 3 are childless; freeing them frees root 0; free slots are 0, 1, 3, 5
    assert par.tolist() == [4, 4, -1, 4, 2, -1]
    assert off.tolist() == [0, 0, 1, 0, 3, 0]
    assert leaves.tolist() == [0, 1, 3]
    assert t.ancestry()[:, :, 0].tolist() == [[2, 12, 20], [2, 12, 21], [2, 12, 22]]


@pytest.mark.parametrize("N,D,T,conc", [(200, 2, 40, 0.3), (57, 1, 25, 3.0), (1000, 3, 12, 1.0)])
def test_paths_equal_dense_walk(N, D, T, conc):
    rng = np.random.default_rng(N + T)
    xs = [rng.standard_normal((N, D))]
    ancs = []
    tree = oracle.PTree(xs[0], capacity=N)
    for t in range(T):
        if rng.random() < 0.3:  # a step that kept its particles
            a = np.arange(N)
        else:  # ancestors drawn from uneven weights, sorted as systematic output is
            w = rng.gamma(conc, size=N)
            a = np.sort(rng.choice(N, size=N, p=w / w.sum()))
        x = rng.standard_normal((N, D))
        tree.insert(x, a)
        ancs.append(a)
        xs.append(x)
        want, distinct = oracle.dense_ancestry(xs, ancs)
        assert np.array_equal(tree.ancestry(), want)
        par, off, leaves = tree.export()
        assert _used(par, off, N) == distinct
        assert len(np.unique(leaves)) == N


def test_identity_ancestors_keep_every_node():
    N, T = 16, 9
    rng = np.random.default_rng(1)
    xs = [rng.standard_normal((N, 1))]
    tree = oracle.PTree(xs[0], capacity=N)
    grew = 0
    for _ in range(T):
        xs.append(rng.standard_normal((N, 1)))
        grew += tree.insert(xs[-1], np.arange(N))
    par, off, _ = tree.export()
    assert _used(par, off, N) == N * (T + 1)
    assert len(par) == 16 * N and grew == 4  # 1 -> 2 -> 4 -> 8 -> 16 times N
    assert np.array_equal(tree.ancestry()[:, :, 0], np.concatenate(xs, axis=1))


def test_capacity_below_n():
    with pytest.raises(ValueError):
        oracle.PTree(np.zeros((5, 1)), capacity=4)
