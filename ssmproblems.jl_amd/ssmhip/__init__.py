"""ssmhip -- MI355X (gfx950) filtering hot path of GeneralisedFilters behind a C-ABI.

The Python modules mirror the reference's host API (models, resamplers, filters) over
libssmhip.so; the numerics all run in the library's HIP kernels.
"""
from ._lib import (SSMError, ArgumentError, PosDefException, AdmissionError, Context,
                   default_context, load, header_symbols, LIB_PATH)
from .models import (StateSpaceModel, HomogeneousGaussianPrior,
                     HomogeneousLinearGaussianLatentDynamics,
                     HomogeneousLinearGaussianObservationProcess,
                     create_homogeneous_linear_gaussian_model, PackedLG, HierarchicalSSM)
from .resamplers import (Systematic, Stratified, Multinomial, ESSResampler, sample_ancestors,
                         offspring_to_ancestors, ancestors_to_offspring)
from .kalman import (KalmanFilter, KF, KFBatch, MvNormalState, pack_lg, kf_loglikelihood,
                     kf_loglikelihood_batch, kf_loglikelihood_grad_batch, SRKalmanFilter,
                     SRKF, srkf_filter_batch, DeviceKFInputs)
from .rbpf import RBPF, RBModel, DeviceRBParticles, rbpf_filter
from .filters import (BootstrapFilter, BF, ParticleFilter, PF, DeviceParticles, initialise,
                      step, run, filter, rng_to_seed, last_timing, AuxiliaryParticleFilter,
                      APF, MeanPredictive, ModePredictive, PredictiveStatistic)
from .proposals import (AbstractProposal, LatentProposal, LinearGaussianProposal,
                        OptimalProposal, OverdispersedProposal)
from .serve import Server
from .tree import ParallelParticleTree, ParallelAncestorCallback
from .history import DenseParticleContainer, DenseAncestorCallback
from .csmc import (ConditionalSMC, NoRefreshment, AncestorSampling, BackwardSimulation,
                   ReferenceTrajectory,
                   CSMCModel, CSMCState, CSMCWorkspace, csmc_sample, mcmc_step)
from .particle_gibbs import (ParticleGibbs, ParticleGibbsModel, ParticleGibbsState,
                             ParticleGibbsTransition, ParameterisedSSM, SSMParameterLogDensity,
                             RandomWalkMH, trajectory_logdensity, pg_step, pg_sample)

__all__ = [n for n in dir() if not n.startswith("_")]
