"""Particle Gibbs over the device conditional SMC (GF/src/integrations/particle_gibbs.jl,
GF/src/integrations/logdensity.jl:19-29, :163-242).

``ParticleGibbs(csmc, param)`` alternates a parameter update on
``SSMParameterLogDensity(prior, ParameterisedSSM(build, ys), trajectory)`` (log p(theta) +
log p(x_0:T, y_1:T | theta), ``trajectory_logdensity``) with one CSMC sweep of the model
built from the new parameters, conditioned on the current trajectory.  The sweeps run on the
device (``csmc_sample``, one workspace for the whole chain; a new parameter draw only
re-uploads the model).  The parameter update is host work, as in the reference, and takes
any object with ``step(rng, logdensity, theta, state) -> (theta, state, stats)``;
``RandomWalkMH`` is provided (the reference's NUTS needs an AD backend this package does not
have).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np

from .csmc import ConditionalSMC, CSMCWorkspace, ReferenceTrajectory, csmc_sample
from .models import StateSpaceModel, lg_params


def _mvn_logpdf(x, mu, S):
    L = np.linalg.cholesky(np.asarray(S, dtype=np.float64))
    r = np.linalg.solve(L, np.asarray(x, dtype=np.float64) - np.asarray(mu, dtype=np.float64))
    return -0.5 * (r.size * np.log(2 * np.pi) + 2 * np.log(np.diag(L)).sum() + r @ r)


def trajectory_logdensity(model: StateSpaceModel, trajectory, observations) -> float:
    """log p(x_0) + sum_t log p(x_t | x_{t-1}) + log p(y_t | x_t) of a linear-Gaussian model
    (logdensity.jl:19-29), parameters read through calc_params at each step."""
    xs = trajectory.as_array() if isinstance(trajectory, ReferenceTrajectory) else trajectory
    xs = np.asarray(xs, dtype=np.float64)
    ys = np.asarray(observations, dtype=np.float64).reshape(len(observations), -1)
    mu0, S0, *_ = lg_params(model, 1)
    ll = _mvn_logpdf(xs[0], mu0, S0)
    for t in range(1, ys.shape[0] + 1):
        _, _, A, b, Q, H, c, R = lg_params(model, t)
        ll += _mvn_logpdf(xs[t], A @ xs[t - 1] + b, Q)
        ll += _mvn_logpdf(ys[t - 1], H @ xs[t] + c, R)
    return float(ll)


@dataclass
class ParameterisedSSM:
    """ParameterisedSSM(build, observations): theta -> model (logdensity.jl:163-186)."""
    build: Callable[[np.ndarray], StateSpaceModel]
    observations: Any


@dataclass
class SSMParameterLogDensity:
    """log p(theta) + trajectory_logdensity(build(theta), trajectory, y)
    (logdensity.jl:188-242); ``prior`` has ``logpdf(theta)`` (e.g. a scipy frozen
    distribution)."""
    prior: Any
    param_model: ParameterisedSSM
    trajectory: Any

    def __call__(self, theta) -> float:
        lp = float(np.sum(self.prior.logpdf(theta)))
        if not np.isfinite(lp):
            return -np.inf
        model = self.param_model.build(np.asarray(theta, dtype=np.float64))
        return lp + trajectory_logdensity(model, self.trajectory, self.param_model.observations)

    logdensity = __call__


@dataclass
class RandomWalkMH:
    """Gaussian random-walk Metropolis-Hastings on theta: ``steps`` proposals per call with
    standard deviation ``scale``."""
    scale: Any = 0.1
    steps: int = 1

    def step(self, rng, logdensity, theta, state=None):
        theta = np.asarray(theta, dtype=np.float64)
        lp = logdensity(theta)
        acc = 0
        for _ in range(self.steps):
            prop = theta + np.asarray(self.scale) * rng.standard_normal(theta.shape)
            lq = logdensity(prop)
            if np.log(rng.random()) < lq - lp:
                theta, lp = prop, lq
                acc += 1
        return theta, None, {"acceptance_rate": acc / self.steps, "log_density": lp}


@dataclass
class ParticleGibbs:
    """ParticleGibbs(csmc, param) (particle_gibbs.jl:7-40)."""
    csmc: ConditionalSMC
    param: Any


@dataclass
class ParticleGibbsModel:
    """ParticleGibbsModel(prior, param_model) (particle_gibbs.jl:42-60)."""
    prior: Any
    param_model: ParameterisedSSM


@dataclass
class ParticleGibbsState:
    theta: np.ndarray
    trajectory: ReferenceTrajectory
    param_state: Any = None
    workspace: CSMCWorkspace | None = field(default=None, repr=False, compare=False)


@dataclass
class ParticleGibbsTransition:
    theta: np.ndarray
    stat: dict


def pg_step(rng, model: ParticleGibbsModel, pg: ParticleGibbs, state=None, initial_params=None,
            ctx=None):
    """AbstractMCMC.step for ParticleGibbs (particle_gibbs.jl:132-199): returns
    (transition, state).  The first step draws theta from the prior (or takes
    initial_params) and runs an unconditional sweep before its parameter update."""
    ys = model.param_model.observations
    K = len(ys)
    if state is None:
        theta = (np.asarray(initial_params, dtype=np.float64) if initial_params is not None
                 else np.atleast_1d(np.asarray(model.prior.rvs(random_state=rng),
                                               dtype=np.float64)))
        ssm = model.param_model.build(theta)
        ws = CSMCWorkspace(ssm, pg.csmc.pf, K, ctx)
        traj, _ = csmc_sample(rng, ssm, pg.csmc, ys, None, workspace=ws)
        param_state = None
    else:
        theta, traj, param_state, ws = state.theta, state.trajectory, state.param_state, \
            state.workspace
    ld = SSMParameterLogDensity(model.prior, model.param_model, traj)
    theta, param_state, stats = pg.param.step(rng, ld, theta, param_state)
    ssm = model.param_model.build(theta)
    traj, _ = csmc_sample(rng, ssm, pg.csmc, ys, traj, workspace=ws)
    return (ParticleGibbsTransition(theta, stats),
            ParticleGibbsState(theta, traj, param_state, ws))


def pg_sample(rng, model: ParticleGibbsModel, pg: ParticleGibbs, n: int, initial_params=None,
              ctx=None):
    """AbstractMCMC.sample: n transitions; returns (thetas (n, d), transitions, state)."""
    state, out = None, []
    for i in range(n):
        tr, state = pg_step(rng, model, pg, state, initial_params if i == 0 else None, ctx)
        out.append(tr)
    return np.array([t.theta for t in out]), out, state
