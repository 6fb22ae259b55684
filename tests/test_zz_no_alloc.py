"""The resampler entry point that allocates nothing (ssm_sample_ancestors_noalloc) against
ssm_sample_ancestors, through the C-ABI: two calls in a row, the second with a smaller
input, every resampler kind and both weight types, bit for bit; sizes past its maximum are
refused."""
import numpy as np
import pytest

from ssmhip import _lib

pytestmark = pytest.mark.gpu
MAXN = 1 << 20


def _run(ctx, fn, kind, w, u, dtype, seed, step):
    n = w.size
    es = 8 if dtype == _lib.SSM_F64 else 4
    dw = _lib.DeviceBuffer(ctx, es * n)
    dw.upload(w.astype(np.float64 if dtype == _lib.SSM_F64 else np.float32))
    da = _lib.DeviceBuffer(ctx, 4 * n)
    du = None
    if u is not None:
        du = _lib.DeviceBuffer(ctx, 8 * u.size)
        du.upload(u)
    ctx.check(fn(ctx.ptr, kind, n, dw.ptr, dtype, du.ptr if du else None, seed, step, da.ptr))
    ctx.synchronize()
    return da.download(np.zeros(n, dtype=np.int32))


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("dtype", [_lib.SSM_F64, _lib.SSM_F32])
def test_noalloc_equals_allocating_entry_point(ctx, kind, dtype):
    L = ctx.lib
    rng = np.random.default_rng(11 + kind + 7 * dtype)
    for n, explicit_u in ((MAXN, False), (70001, True)):  # the second call is smaller
        w = rng.random(n) ** 3
        w /= w.sum()
        u = rng.random(n if kind >= 2 else 1) if explicit_u else None
        a_ref = _run(ctx, L.ssm_sample_ancestors, kind, w, u, dtype, 42 + n, 3)
        a_new = _run(ctx, L.ssm_sample_ancestors_noalloc, kind, w, u, dtype, 42 + n, 3)
        assert np.array_equal(a_new, a_ref), (kind, dtype, n)
        assert a_new.min() >= 0 and a_new.max() < n


def test_noalloc_refuses_sizes_past_its_maximum(ctx):
    L = ctx.lib
    n = MAXN + 1
    dw = _lib.DeviceBuffer(ctx, 8 * n)
    da = _lib.DeviceBuffer(ctx, 4 * n)
    rc = L.ssm_sample_ancestors_noalloc(ctx.ptr, 0, n, dw.ptr, _lib.SSM_F64, None, 1, 1, da.ptr)
    assert rc == _lib.SSM_ERR_ARG
    assert b"maximum" in L.ssm_last_error(ctx.ptr)
