"""Proposals of the guided particle filter (GF/src/algorithms/particles.jl:92-108, :175-201).

``ParticleFilter(N, proposal)`` draws each particle's next state from the proposal
q(x' | x, y) instead of the latent dynamics f(x' | x) and adds log f(x' | x) -
log q(x' | x, y) to its weight before the observation's density (``propogate``,
particles.jl:175-195).  The device runs proposals of the linear-Gaussian form
q = N(F x + g0 + K y, P); ``linear_form(step)`` gives (F, g0, K, P) at a step through the
calc_params protocol.  Both of the reference's test proposals (GF/src/GFTest/proposals.jl)
have this form.
"""
from __future__ import annotations

import numpy as np

from .models import (HomogeneousLinearGaussianLatentDynamics,
                     HomogeneousLinearGaussianObservationProcess, _arr)


class AbstractProposal:
    is_homogeneous = True

    def linear_form(self, step: int, **kw):
        raise NotImplementedError


class LatentProposal(AbstractProposal):
    """The bootstrap filter's proposal, the latent dynamics (particles.jl:197-201)."""


def _mat(v):
    return np.atleast_2d(np.asarray(_arr(v), dtype=np.float64))


def _vec(v):
    return np.atleast_1d(np.asarray(_arr(v), dtype=np.float64))


class LinearGaussianProposal(AbstractProposal):
    """q(x' | x, y) = N(F x + g0 + K y, P) with fixed F, g0, K (None: 0) and P."""

    def __init__(self, F, g0, P, K=None):
        self.F, self.g0, self.P = _mat(F), _vec(g0), _mat(P)
        self.K = None if K is None else _mat(K)

    def linear_form(self, step: int, **kw):
        return self.F, self.g0, self.K, self.P


class OptimalProposal(AbstractProposal):
    """p(x_t | x_{t-1}, y_t) of a linear-Gaussian model (GF/src/GFTest/proposals.jl:11-31):
    the predicted state N(A x + b, Q) updated with y_t by the jitter-free Kalman update
    (kalman_gradient.jl:12-29).  With the gain K = Q H' (H Q H' + R)^-1 its mean is
    (I - K H)(A x + b) + K (y - c) and its covariance the Joseph form
    (I - K H) Q (I - K H)' + K R K', upper triangle mirrored (Symmetric(:U))."""

    def __init__(self, dyn, obs):
        self.dyn, self.obs = dyn, obs

    @property
    def is_homogeneous(self):
        return (isinstance(self.dyn, HomogeneousLinearGaussianLatentDynamics)
                and isinstance(self.obs, HomogeneousLinearGaussianObservationProcess))

    def linear_form(self, step: int, **kw):
        A, b, Q = self.dyn.calc_params(step, **kw)
        H, c, R = self.obs.calc_params(step, **kw)
        A, Q, H, R = _mat(A), _mat(Q), _mat(H), _mat(R)
        b, c = _vec(b), _vec(c)
        S = H @ Q @ H.T + R
        K = np.linalg.solve(S, H @ Q).T
        IKH = np.eye(A.shape[0]) - K @ H
        P = IKH @ Q @ IKH.T + K @ R @ K.T
        P = np.triu(P) + np.triu(P, 1).T
        return IKH @ A, IKH @ b - K @ c, K, P


class OverdispersedProposal(AbstractProposal):
    """The latent dynamics with the covariance inflated by k: N(A x + b, k Q)
    (GF/src/GFTest/proposals.jl:33-49)."""

    def __init__(self, dyn, k: float):
        self.dyn, self.k = dyn, float(k)

    @property
    def is_homogeneous(self):
        return isinstance(self.dyn, HomogeneousLinearGaussianLatentDynamics)

    def linear_form(self, step: int, **kw):
        A, b, Q = self.dyn.calc_params(step, **kw)
        return _mat(A), _vec(b), None, self.k * _mat(Q)
