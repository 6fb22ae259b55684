"""Device resamplers against the exact fixtures and the FP64 oracle (bit-exact ancestors)."""
import ctypes as C

import numpy as np
import pytest

import ssmhip
from ssmhip import _lib

pytestmark = pytest.mark.gpu
KINDS = [0, 1, 2, 3]


def test_fixtures_bit_exact(ctx, fixtures):
    for c in fixtures["cases"]:
        R = {0: ssmhip.Systematic(), 1: ssmhip.Systematic(form="gpu"), 2: ssmhip.Stratified(),
             3: ssmhip.Multinomial()}[c["kind_id"]]
        a = ssmhip.sample_ancestors(None, R, np.array(c["w"]), u=np.array(c["u"]), ctx=ctx)
        assert list(a) == c["ancestors"], (c["kind"], c["tag"], c["n"])


def test_reference_known_answers(ctx, fixtures):
    k = fixtures["known"]
    o = k["offspring_to_ancestors"]
    assert list(ssmhip.offspring_to_ancestors(o["offspring"], ctx=ctx)) == o["ancestors"]
    a = k["ancestors_to_offspring"]
    assert list(ssmhip.ancestors_to_offspring(a["ancestors"], ctx=ctx)) == a["offspring"]


def _resampler(kind):
    return {0: ssmhip.Systematic(), 1: ssmhip.Systematic(form="gpu"), 2: ssmhip.Stratified(),
            3: ssmhip.Multinomial()}[kind]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [1, 2, 5, 1000, 2047, 2048, 2049, 4099, 65536 + 7, 1 << 20])
def test_random_weights_match_oracle(ctx, orc, kind, n):
    rng = np.random.default_rng(n * 7 + kind)
    w = rng.random(n) ** 3
    w /= w.sum()
    u = rng.random(n)
    a_gpu = ssmhip.sample_ancestors(None, _resampler(kind), w, u=u if kind >= 2 else u[:1], ctx=ctx)
    a_orc = orc.resample(kind, w, u)
    mism = int((a_gpu != a_orc).sum())
    print(f"kind={kind} n={n} mismatches={mism}")
    assert mism == 0


@pytest.mark.parametrize("kind", KINDS)
def test_philox_uniforms_match_oracle(ctx, orc, kind):
    n, seed, step = 100003, 987654321, 17
    rng = np.random.default_rng(5)
    w = rng.random(n)
    w /= w.sum()
    dw = _lib.DeviceBuffer(ctx, 8 * n)
    da = _lib.DeviceBuffer(ctx, 4 * n)
    dw.upload(w)
    ctx.check(ctx.lib.ssm_sample_ancestors(ctx.ptr, kind, n, dw.ptr, 1, None, seed, step, da.ptr))
    a = da.download(np.zeros(n, dtype=np.int32))
    u = orc.philox_uniforms(seed, step, n)
    assert np.array_equal(a, orc.resample(kind, w, u))


@pytest.mark.parametrize("kind", KINDS)
def test_f32_weights(ctx, orc, kind):
    n = 30011
    rng = np.random.default_rng(11)
    w32 = (rng.random(n) ** 2).astype(np.float32)
    u = rng.random(n)
    dw = _lib.DeviceBuffer(ctx, 4 * n)
    du = _lib.DeviceBuffer(ctx, 8 * n)
    da = _lib.DeviceBuffer(ctx, 4 * n)
    dw.upload(w32)
    du.upload(u)
    ctx.check(ctx.lib.ssm_sample_ancestors(ctx.ptr, kind, n, dw.ptr, 0, du.ptr, 0, 0, da.ptr))
    a = da.download(np.zeros(n, dtype=np.int32))
    # the CPU systematic definition assumes normalised weights; feed the same values
    assert np.array_equal(a, orc.resample(kind, w32.astype(np.float64), u))


@pytest.mark.parametrize("kind", KINDS)
def test_large_invariants_and_statistics(ctx, kind):
    # test/components/resamplers.jl:38-74 (the reference's never-run GPU tests): resampled
    # mean ~ weighted mean; plus invariants at n = 2^22.
    n = 1 << 22
    rng = np.random.default_rng(1234)
    xs = rng.standard_normal(n)
    ws = np.exp(xs - 0.5)
    ws /= ws.sum()
    a = ssmhip.sample_ancestors(rng, _resampler(kind), ws, ctx=ctx)
    assert a.min() >= 0 and a.max() < n
    if kind != 3:
        assert np.all(np.diff(a) >= 0)
    off = ssmhip.ancestors_to_offspring(a, ctx=ctx)
    assert off.sum() == n
    mu0 = (ws * xs).sum()
    assert abs(xs[a].mean() - mu0) <= 0.1 * abs(mu0)


@pytest.mark.parametrize("kind", KINDS)
def test_degenerate_one_hot_large(ctx, kind):
    # One particle carries all the weight: one tile fills every output slot.
    n = (1 << 20) + 3
    w = np.zeros(n)
    w[n // 3] = 1.0
    a = ssmhip.sample_ancestors(np.random.default_rng(2), _resampler(kind), w, ctx=ctx)
    assert np.all(a == n // 3)


def test_bad_arguments_raise(ctx):
    with pytest.raises(ssmhip.ArgumentError):
        ssmhip.sample_ancestors(None, ssmhip.Systematic(), np.zeros(0), u=np.zeros(1), ctx=ctx)
