"""Resampler types and sample_ancestors, mirroring GF/src/resamplers.jl.

The types carry no state; they select the device algorithm:

* ``Systematic()``                resamplers.jl:214-238 (CPU definition, the default)
* ``Systematic(form="gpu")``      CUDAExt.jl:41-58 (cumulative-offspring definition)
* ``Stratified()``                CUDAExt.jl:60-78
* ``Multinomial()``               CUDAExt.jl:31-39 (inverse CDF)
* ``ESSResampler(threshold, r)``  resamplers.jl:171-183 (resample iff threshold*N >= ESS)

``sample_ancestors(rng, resampler, weights)`` draws its uniforms from ``rng`` on the host
(so a caller holding the uniforms can reproduce the result) and runs the device kernels;
ancestors are returned 0-based.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (SSM_MULTINOMIAL, SSM_STRATIFIED, SSM_SYSTEMATIC, SSM_SYSTEMATIC_MURRAY,
                   default_context)


class AbstractResampler:
    kind: int = -1
    n_uniforms_per_particle = True


class Systematic(AbstractResampler):
    def __init__(self, form: str = "cpu"):
        if form not in ("cpu", "gpu"):
            raise ValueError("form must be 'cpu' or 'gpu'")
        self.form = form
        self.kind = SSM_SYSTEMATIC if form == "cpu" else SSM_SYSTEMATIC_MURRAY
        self.n_uniforms_per_particle = False

    def __repr__(self):
        return f"Systematic(form={self.form!r})"


class Stratified(AbstractResampler):
    kind = SSM_STRATIFIED


class Multinomial(AbstractResampler):
    kind = SSM_MULTINOMIAL


class ESSResampler:
    """Conditional resampler: resample iff threshold * N >= ESS (resamplers.jl:179-183)."""

    def __init__(self, threshold: float, resampler: AbstractResampler = None):
        self.threshold = float(threshold)
        self.resampler = resampler if resampler is not None else Systematic()


def draw_uniforms(rng: np.random.Generator, resampler: AbstractResampler, n: int):
    """The uniforms a resampler consumes: u0 for systematic, n otherwise."""
    if resampler.n_uniforms_per_particle:
        return rng.random(n)
    return np.array([rng.random()])


def sample_ancestors(rng, resampler: AbstractResampler, weights, u=None, ctx=None):
    """sample_ancestors(rng, resampler, weights) -> 0-based int32 ancestors (device kernels)."""
    ctx = ctx or default_context()
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64))
    n = w.shape[0]
    if u is None:
        u = draw_uniforms(rng, resampler, n)
    u = np.ascontiguousarray(np.asarray(u, dtype=np.float64))
    anc = np.zeros(n, dtype=np.int32)
    ctx.check(ctx.lib.ssm_sample_ancestors_host(
        ctx.ptr, resampler.kind, n, w.ctypes.data_as(_lib._dp), u.ctypes.data_as(_lib._dp),
        anc.ctypes.data_as(_lib._ip)))
    return anc


def offspring_to_ancestors(offspring, ctx=None):
    """CUDAExt.jl:105-117 on the device: cumulative offspring counts -> ancestors (0-based)."""
    ctx = ctx or default_context()
    o = np.ascontiguousarray(np.asarray(offspring, dtype=np.int32))
    n = o.shape[0]
    din = _lib.DeviceBuffer(ctx, 4 * n)
    dout = _lib.DeviceBuffer(ctx, 4 * n)
    din.upload(o)
    ctx.check(ctx.lib.ssm_offspring_to_ancestors(ctx.ptr, n, din.ptr, dout.ptr))
    return dout.download(np.zeros(n, dtype=np.int32))


def ancestors_to_offspring(ancestors, ctx=None):
    """CUDAExt.jl:130-142 on the device: ancestors (0-based) -> offspring counts."""
    ctx = ctx or default_context()
    a = np.ascontiguousarray(np.asarray(ancestors, dtype=np.int32))
    n = a.shape[0]
    din = _lib.DeviceBuffer(ctx, 4 * n)
    dout = _lib.DeviceBuffer(ctx, 4 * n)
    din.upload(a)
    ctx.check(ctx.lib.ssm_ancestors_to_offspring(ctx.ptr, n, din.ptr, dout.ptr))
    return dout.download(np.zeros(n, dtype=np.int32))
