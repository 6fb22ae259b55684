"""Model definitions mirroring the reference's interface (host side, unchanged semantics).

SSMProblems/src/SSMProblems.jl:25-272 defines StatePrior / LatentDynamics /
ObservationProcess / StateSpaceModel; GF/src/models/linear_gaussian.jl:31-131 defines the
calc_* parameter protocol and the homogeneous linear-Gaussian types; and
GF/src/models/hierarchical.jl:4-45 the outer/inner split used by the RBPF.  The device
library never sees these objects: the filters read their parameters through calc_params
each step, exactly where the reference does, and pass plain arrays across the C-ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import LGModelC, _dp


class StatePrior:
    pass


class LatentDynamics:
    pass


class ObservationProcess:
    pass


class GaussianPrior(StatePrior):
    def calc_mu0(self, **kw):
        raise NotImplementedError

    def calc_Sigma0(self, **kw):
        raise NotImplementedError

    def calc_initial(self, **kw):
        return self.calc_mu0(**kw), self.calc_Sigma0(**kw)


class LinearGaussianLatentDynamics(LatentDynamics):
    """calc_params(dyn, step) -> (A, b, Q), linear_gaussian.jl:44-50."""

    def calc_params(self, step, **kw):
        return self.calc_A(step, **kw), self.calc_b(step, **kw), self.calc_Q(step, **kw)


class LinearGaussianObservationProcess(ObservationProcess):
    """calc_params(obs, step) -> (H, c, R), linear_gaussian.jl:57-63."""

    def calc_params(self, step, **kw):
        return self.calc_H(step, **kw), self.calc_c(step, **kw), self.calc_R(step, **kw)


def _arr(a):
    return np.asarray(a, dtype=np.float64)


class HomogeneousGaussianPrior(GaussianPrior):
    def __init__(self, mu0, Sigma0):
        self.mu0, self.Sigma0 = _arr(mu0).reshape(-1), np.atleast_2d(_arr(Sigma0))

    def calc_mu0(self, **kw):
        return self.mu0

    def calc_Sigma0(self, **kw):
        return self.Sigma0


class HomogeneousLinearGaussianLatentDynamics(LinearGaussianLatentDynamics):
    def __init__(self, A, b, Q):
        self.A, self.b, self.Q = np.atleast_2d(_arr(A)), _arr(b).reshape(-1), np.atleast_2d(_arr(Q))

    def calc_A(self, step, **kw):
        return self.A

    def calc_b(self, step, **kw):
        return self.b

    def calc_Q(self, step, **kw):
        return self.Q


class HomogeneousLinearGaussianObservationProcess(LinearGaussianObservationProcess):
    def __init__(self, H, c, R):
        self.H, self.c, self.R = np.atleast_2d(_arr(H)), _arr(c).reshape(-1), np.atleast_2d(_arr(R))

    def calc_H(self, step, **kw):
        return self.H

    def calc_c(self, step, **kw):
        return self.c

    def calc_R(self, step, **kw):
        return self.R


class StateSpaceModel:
    """SSMProblems.jl:264-272: prior, dyn, obs."""

    def __init__(self, prior, dyn, obs):
        self.prior, self.dyn, self.obs = prior, dyn, obs

    @property
    def is_linear_gaussian(self):
        return (isinstance(self.prior, GaussianPrior)
                and isinstance(self.dyn, LinearGaussianLatentDynamics)
                and isinstance(self.obs, LinearGaussianObservationProcess))

    @property
    def is_homogeneous(self):
        return (isinstance(self.dyn, HomogeneousLinearGaussianLatentDynamics)
                and isinstance(self.obs, HomogeneousLinearGaussianObservationProcess))


def create_homogeneous_linear_gaussian_model(mu0, Sigma0, A, b, Q, H, c, R):
    """linear_gaussian.jl:125-131."""
    return StateSpaceModel(HomogeneousGaussianPrior(mu0, Sigma0),
                           HomogeneousLinearGaussianLatentDynamics(A, b, Q),
                           HomogeneousLinearGaussianObservationProcess(H, c, R))


def lg_params(model: StateSpaceModel, step: int, **kw):
    """(mu0, Sigma0, A, b, Q, H, c, R) at `step` through the calc_* protocol."""
    mu0, Sigma0 = model.prior.calc_initial(**kw)
    A, b, Q = model.dyn.calc_params(step, **kw)
    H, c, R = model.obs.calc_params(step, **kw)
    return tuple(np.atleast_1d(_arr(v)) for v in (mu0, Sigma0, A, b, Q, H, c, R))


class PackedLG:
    """Column-major FP64 buffers + the ssm_lg_model struct that points at them."""

    def __init__(self, mu0, Sigma0, A, b, Q, H, c, R):
        A = np.atleast_2d(A)
        H = np.atleast_2d(H)
        Dx, Dy = A.shape[0], H.shape[0]
        f = lambda M, r, c_: np.ascontiguousarray(np.asarray(M, dtype=np.float64).reshape(r, c_).flatten(order="F"))
        self._bufs = [np.ascontiguousarray(np.asarray(mu0, dtype=np.float64).reshape(Dx)),
                      f(Sigma0, Dx, Dx), f(A, Dx, Dx),
                      np.ascontiguousarray(np.asarray(b, dtype=np.float64).reshape(Dx)),
                      f(Q, Dx, Dx), f(H, Dy, Dx),
                      np.ascontiguousarray(np.asarray(c, dtype=np.float64).reshape(Dy)),
                      f(R, Dy, Dy)]
        self.Dx, self.Dy = Dx, Dy
        self.struct = LGModelC(Dx, Dy, *[b_.ctypes.data_as(_dp) for b_ in self._bufs])

    @classmethod
    def from_model(cls, model: StateSpaceModel, step: int = 1, **kw):
        return cls(*lg_params(model, step, **kw))

    def ref(self):
        return C.byref(self.struct)

    def same_as(self, other) -> bool:
        """True when both hold the same dimensions and parameter values."""
        return (other is not None and self.Dx == other.Dx and self.Dy == other.Dy and
                all(np.array_equal(a, b) for a, b in zip(self._bufs, other._bufs)))


# ---- hierarchical models (hierarchical.jl:4-45) -----------------------------------------

class HierarchicalSSM:
    """Outer prior/dynamics plus an inner model whose parameters may depend on the outer
    state through the prev_outer / new_outer keyword arguments (hierarchical.jl:36-45)."""

    def __init__(self, outer_prior, outer_dyn, inner_prior, inner_dyn, obs):
        self.outer_prior, self.outer_dyn = outer_prior, outer_dyn
        self.inner_model = StateSpaceModel(inner_prior, inner_dyn, obs)
