"""Shared pytest setup: markers, import paths, the oracle and a device context."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "ssmproblems.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def ctx():
    import ssmhip
    return ssmhip.default_context()


@pytest.fixture(scope="session")
def fixtures():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "resample_fixtures.json")) as f:
        return json.load(f)


def rand_cov(rng, d, scale=1.0, var_range=(0.8, 1.2)):
    """GFTest/utils.jl:14-36 restated: random correlation matrix scaled to given variances."""
    A = rng.random((d, d))
    S = A @ A.T
    Di = np.diag(1.0 / np.sqrt(np.diag(S)))
    S = Di @ S @ Di
    v = (rng.random(d) * (var_range[1] - var_range[0]) + var_range[0]) * scale
    Ds = np.diag(np.sqrt(v))
    S = Ds @ S @ Ds
    return (S + S.T) / 2


def random_lg(rng, Dx, Dy, process_noise_scale=0.1, obs_noise_scale=1.0):
    """GFTest/models/linear_gaussian.jl:6-38 restated (create_linear_gaussian_model)."""
    import oracle
    return oracle.LG(rng.random(Dx), rand_cov(rng, Dx), rng.random((Dx, Dx)), rng.random(Dx),
                     rand_cov(rng, Dx, process_noise_scale), rng.random((Dy, Dx)),
                     rng.random(Dy), rand_cov(rng, Dy, obs_noise_scale))


def config2_model(seed=1234):
    """D=4 LG SSM: A = 0.95 * seeded orthogonal, Q = 0.1 I, H seeded 2x4, R = 0.2 I, prior N(0, I)."""
    import oracle
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((4, 4))
    Qm, Rm = np.linalg.qr(G)
    Qm = Qm * np.sign(np.diag(Rm))
    A = 0.95 * Qm
    H = rng.standard_normal((2, 4))
    return oracle.LG(np.zeros(4), np.eye(4), A, np.zeros(4), 0.1 * np.eye(4), H, np.zeros(2),
                     0.2 * np.eye(2))


def config1_model():
    """1-D LG SSM: prior N(0,1); x_t = x_{t-1} + N(0, 0.1^2); y_t = x_t + N(0, 0.2^2)."""
    import oracle
    return oracle.LG([0.0], [[1.0]], [[1.0]], [0.0], [[0.01]], [[1.0]], [0.0], [[0.04]])


def to_ssmhip_model(m):
    import ssmhip
    return ssmhip.create_homogeneous_linear_gaussian_model(m.mu0, m.Sigma0, m.A, m.b, m.Q, m.H,
                                                           m.c, m.R)
