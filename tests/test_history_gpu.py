"""Device dense history (DenseParticleContainer, GF/src/containers.jl:328-418) against the
per-step exports of the filter and the oracle's walk of those exports.

Every stored generation must equal, bit for bit, what ssm_bf_export reported right after
that step (states, normalised log-weights, ancestors with the identity after a step that
kept its particles), and every path must equal the oracle's dense walk
(get_ancestry, containers.jl:406-418) and the sparse tree's path of the same particle.
"""
import numpy as np
import pytest

import oracle
import ssmhip
from conftest import random_lg, to_ssmhip_model

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,N,Dx,Dy,T,thr,kind", [("f32", 5000, 3, 2, 30, 0.6, 0),
                                                     ("f64", 3001, 2, 1, 21, 1.0, 3),
                                                     ("f32", 40000, 4, 2, 9, 0.5, 2)])
def test_history_equals_exports_and_tree(ctx, prec, N, Dx, Dy, T, thr, kind):
    rng = np.random.default_rng(N + T)
    m = random_lg(rng, Dx, Dy)
    m.A = 0.9 * m.A / max(abs(np.linalg.eigvals(m.A)))
    model = to_ssmhip_model(m)
    res = {0: ssmhip.Systematic(), 2: ssmhip.Stratified(), 3: ssmhip.Multinomial()}[kind]
    algo = ssmhip.BF(N, threshold=thr, resampler=res, precision=prec)
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=3)
    hist = ssmhip.DenseParticleContainer(st, capacity=4)  # grows several times
    tree = ssmhip.ParallelParticleTree(st)
    xs, ancs, lws = [st.export()[0]], [], []
    ys = rng.standard_normal((T, Dy))
    for t in range(1, T + 1):
        ssmhip.step(rng, model, algo, t, st, ys[t - 1])
        x, lw, anc, _ = st.export()
        ancs.append(anc.astype(np.int64) if st.last_resampled else np.arange(N))
        xs.append(x)
        lws.append(lw)
        hist.push()
        tree.insert()
    assert len(hist) == T
    assert np.array_equal(hist.generation(0)[0], xs[0])
    for t in (1, T // 2, T):
        x, lw, anc = hist.generation(t)
        assert np.array_equal(x, xs[t])
        assert np.array_equal(lw, lws[t - 1])
        assert np.array_equal(anc, ancs[t - 1])
    want, _ = oracle.dense_ancestry(xs, ancs)
    assert np.array_equal(hist.ancestry(), want)
    assert np.array_equal(hist.ancestry(), tree.ancestry())
    assert np.array_equal(hist.ancestry(N // 3, 17), want[N // 3:N // 3 + 17])
    i = N - 1
    assert np.array_equal(hist.get_ancestry(i), want[i])
    s, w, a = hist.particle(T, i)
    assert np.array_equal(s, xs[T][i]) and w == lws[-1][i] and a == ancs[-1][i]
    tree.close()
    hist.close()


def test_callback_and_errors(ctx):
    rng = np.random.default_rng(2)
    m = random_lg(rng, 2, 2)
    m.A = 0.8 * m.A / max(abs(np.linalg.eigvals(m.A)))
    model = to_ssmhip_model(m)
    algo = ssmhip.BF(1024, threshold=0.7, precision="f32")
    ys = rng.standard_normal((12, 2))
    cb = ssmhip.DenseAncestorCallback(capacity=2)
    state, ll = ssmhip.filter(rng, model, algo, ys, callback=cb, ctx=ctx)
    h = cb.container
    assert len(h) == 12 and np.isfinite(ll)
    assert np.array_equal(h.ancestry()[:, -1, :], state.export()[0])
    with pytest.raises(ssmhip.ArgumentError):
        h.generation(13)
    with pytest.raises(ssmhip.ArgumentError):
        h.ancestry(1000, 25)
    with pytest.raises(ssmhip.ArgumentError):
        ssmhip.DenseParticleContainer(state, capacity=0)
    # two histories on one filter are fine; the filter may go first
    h2 = ssmhip.DenseParticleContainer(state)
    state.close()
    with pytest.raises(ssmhip.SSMError) as e:
        h2.push()
    assert e.value.code == ssmhip._lib.SSM_ERR_STATE
    assert h.ancestry().shape == (1024, 13, 2)  # the stored history stays readable
    h2.close()
    h.close()


@pytest.mark.parametrize("prec,kind,thr", [("f32", 0, 0.7), ("f64", 2, 1.0), ("f32", 3, 0.5)])
def test_recording_during_graph_run_equals_pushes(ctx, prec, kind, thr):
    # DenseAncestorCallback inside filter(...): a recording history filled by the filter's own
    # steps, here 37 graph-replayed steps (two 16-step graphs and 5 single steps) in two runs,
    # equals the manual pushes after per-step calls with the same Philox key, bit for bit
    rng = np.random.default_rng(17)
    m = random_lg(rng, 3, 2)
    m.A = 0.9 * m.A / max(abs(np.linalg.eigvals(m.A)))
    model = to_ssmhip_model(m)
    res = {0: ssmhip.Systematic(), 2: ssmhip.Stratified(), 3: ssmhip.Multinomial()}[kind]
    algo = ssmhip.BF(6000, threshold=thr, resampler=res, precision=prec)
    ys = rng.standard_normal((37, 2))
    a = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=21)
    ha = ssmhip.DenseParticleContainer(a, capacity=2)  # grows during the first run
    ha.record(True)
    ll1 = ssmhip.run(a, algo, ys[:20], t0=1)
    ll2 = ssmhip.run(a, algo, ys[20:], t0=21)
    assert len(ha) == 37
    b = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=21)
    hb = ssmhip.DenseParticleContainer(b, capacity=64)
    incs = []
    for t in range(1, 38):
        incs.append(ssmhip.step(rng, model, algo, t, b, ys[t - 1])[1])
        hb.push()
    assert ll1 + ll2 == pytest.approx(sum(incs), rel=1e-12)
    for t in (0, 1, 16, 17, 20, 21, 37):
        for ga, gb in zip(ha.generation(t), hb.generation(t)):
            if ga is not None:
                assert np.array_equal(ga, gb), t
    assert np.array_equal(ha.ancestry(0, 50), hb.ancestry(0, 50))
    # per-step calls record too; turning recording off stops it
    ssmhip.step(rng, model, algo, 38, a, ys[0])
    assert len(ha) == 38
    ha.record(False)
    ssmhip.step(rng, model, algo, 39, a, ys[1])
    assert len(ha) == 38
