"""Conditional SMC sweeps on the device bootstrap filter (GF/src/algorithms/csmc.jl).

``csmc_sample(rng, model, ConditionalSMC(BF(N), refreshment), observations, ref_traj)``
runs one sweep and returns ``(ReferenceTrajectory, log_likelihood)`` as ``_csmc_sample``
does (csmc.jl:239-306).  The particles, their history and every draw stay on the device:
particle 0 is pinned to the reference (initialise and each step), its ancestor is particle
0 (NoRefreshment) or drawn on the device from the ancestor-sampling weights
(AncestorSampling), each generation goes into a device DenseParticleContainer that the
filter's steps fill themselves (a homogeneous model runs the whole sweep as one
graph-replayed call), and the returned trajectory is the path of one particle drawn from
the final weights (_sample_trajectory, csmc.jl:111-117).  ``mcmc_step`` is the AbstractMCMC
interface (csmc.jl:450-462).  BackwardSimulation draws the trajectory backwards through the device
history with the transition density (csmc.jl:392-446).  The model must be
linear-Gaussian; the Rao-Blackwellised CSMC is not provided.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import default_context
from .filters import BootstrapFilter, DeviceParticles, _proposal, rng_to_seed, run, step
from .history import DenseParticleContainer
from .models import PackedLG, StateSpaceModel


class NoRefreshment:
    """Vanilla CSMC: the reference keeps its own ancestry (csmc.jl:13-19)."""
    mode = _lib.SSM_CSMC_PLAIN


class AncestorSampling:
    """CSMC-AS / PGAS: the reference's ancestor is redrawn every step (csmc.jl:21-29)."""
    mode = _lib.SSM_CSMC_ANCESTOR


class BackwardSimulation:
    """CSMC-BS: a plain conditional forward pass, then the trajectory drawn backwards with
    the transition density (csmc.jl:31-41)."""
    mode = _lib.SSM_CSMC_PLAIN


@dataclass
class ConditionalSMC:
    """ConditionalSMC(pf, refreshment=NoRefreshment()) -- csmc.jl:43-67."""
    pf: BootstrapFilter
    refreshment: object = field(default_factory=NoRefreshment)


class ReferenceTrajectory:
    """A trajectory indexed from 0: traj[0] is the time-0 state, traj[t] the state at t
    (containers.jl:254-292)."""

    def __init__(self, x0, xs):
        self.x0 = np.asarray(x0, dtype=np.float64)
        self.xs = np.asarray(xs, dtype=np.float64)

    def __len__(self):
        return self.xs.shape[0] + 1

    def __getitem__(self, t):
        return self.x0 if t == 0 else self.xs[t - 1]

    def as_array(self):
        """(T + 1, Dx), row t the state at time t."""
        return np.concatenate([self.x0[None, :], self.xs], axis=0)

    def __eq__(self, other):
        return (isinstance(other, ReferenceTrajectory) and np.array_equal(self.x0, other.x0)
                and np.array_equal(self.xs, other.xs))


@dataclass
class CSMCModel:
    """The model wrapper of the AbstractMCMC interface (csmc.jl:83-95)."""
    ssm: StateSpaceModel
    observations: object


class CSMCWorkspace:
    """The device filter and particle history one chain reuses across sweeps."""

    def __init__(self, model: StateSpaceModel, pf: BootstrapFilter, K: int, ctx=None):
        if not model.is_linear_gaussian:
            raise TypeError("device conditional SMC needs a linear-Gaussian model")
        if _proposal(pf) is not None:
            raise TypeError("device conditional SMC runs the bootstrap proposal")
        self.ctx = ctx or default_context()
        self.state = DeviceParticles(self.ctx, model, pf, 0)
        self.ctx.check(self.ctx.lib.ssm_bf_initialise(self.state.ptr, None))
        self.history = DenseParticleContainer(self.state, capacity=max(int(K), 1))
        self.history.record(True)  # every step appends its generation on the device
        self.N, self.Dx = pf.N, self.state.Dx

    def close(self):
        self.history.close()
        self.state.close()


@dataclass
class CSMCState:
    """State of the sampler: the current reference trajectory (csmc.jl:69-81)."""
    trajectory: ReferenceTrajectory
    log_likelihood: float = float("nan")
    workspace: CSMCWorkspace | None = field(default=None, repr=False, compare=False)


def csmc_sample(rng, model: StateSpaceModel, csmc: ConditionalSMC, observations, ref_traj,
                workspace: CSMCWorkspace | None = None, ctx=None):
    """One sweep: (ReferenceTrajectory, log_likelihood).  ref_traj None runs the first,
    unconditional sweep.  Consumes one UInt64 of rng to key the sweep's device draws and
    one uniform for the trajectory draw."""
    ys = np.asarray(observations, dtype=np.float64)
    K = ys.shape[0]
    ys = ys.reshape(K, -1)
    ws = workspace or CSMCWorkspace(model, csmc.pf, K, ctx)
    st = ws.state
    lib = ws.ctx.lib
    ws.ctx.check(lib.ssm_bf_set_seed(st.ptr, rng_to_seed(rng)))
    ancestor_sampling = isinstance(csmc.refreshment, AncestorSampling)
    pf = csmc.pf
    if ancestor_sampling:
        # resample at every step (csmc.jl:288 calls resample, not maybe_resample)
        pf = dataclasses.replace(pf, threshold=float("inf"))
    if ref_traj is None:
        st.set_reference(None)
    else:
        xr = ref_traj.as_array() if isinstance(ref_traj, ReferenceTrajectory) else ref_traj
        xr = np.asarray(xr, dtype=np.float64).reshape(-1, ws.Dx)
        if xr.shape[0] != K + 1:
            raise ValueError(f"reference has {xr.shape[0]} states, need {K + 1}")
        st.set_reference(xr, csmc.refreshment.mode)
    packed = PackedLG.from_model(model, 1)
    if not packed.same_as(st.packed):  # a new model (particle Gibbs' parameter draw)
        st.set_params(packed)
    ws.ctx.check(lib.ssm_bf_initialise(st.ptr, None))
    ws.history.reset()
    if model.is_homogeneous:  # all K steps in one call, replayed from HIP graphs
        ll = run(st, pf, ys)
    else:  # calc_params per step
        ll = 0.0
        for t in range(1, K + 1):
            _, inc = step(rng, model, pf, t, st, ys[t - 1])
            ll += inc
    if isinstance(csmc.refreshment, BackwardSimulation):
        # the last time's draw first, then t = K-1 .. 0, as csmc.jl:417-444 consumes rng
        u = rng.random(K + 1)[::-1]
        # time t is weighted with the dynamics of step t + 1 (csmc.jl:424)
        dyn = None if model.is_homogeneous else [PackedLG.from_model(model, t + 1)
                                                 for t in range(K)]
        _, path = ws.history.backward_sample(u, dyn)
        return ReferenceTrajectory(path[0], path[1:]), ll
    idx = st.sample_index(float(rng.random()))
    path = ws.history.get_ancestry(idx)
    return ReferenceTrajectory(path[0], path[1:]), ll


def mcmc_step(rng, model: CSMCModel, csmc: ConditionalSMC, state: CSMCState | None = None,
              ctx=None):
    """AbstractMCMC.step (csmc.jl:450-462): one sweep conditioned on state's trajectory
    (none for the first step).  Returns (sample, state); both are the new CSMCState."""
    ws = state.workspace if state is not None else None
    if ws is None:
        ws = CSMCWorkspace(model.ssm, csmc.pf, len(model.observations), ctx)
    ref = state.trajectory if state is not None else None
    traj, ll = csmc_sample(rng, model.ssm, csmc, model.observations, ref, workspace=ws)
    new = CSMCState(traj, ll, ws)
    return new, new
