"""The filtering server: continuous batching, HBM-budget admission, results alone vs in a mix.

Bootstrap-filter requests run the single-filter per-tile code inside batched launches, so a
served log-likelihood must equal the plain per-filter path's (ssm_bf_run with the same
Philox seed) bit for bit, whatever else is in the batch.  Kalman requests are checked
against the FP64 oracle (1e-4 relative, the batched-KF tolerance) and alone vs in a mix
(bit for bit: one workgroup runs each filter with a fixed order of operations).
"""
import threading
import time

import numpy as np
import pytest

import ssmhip
from ssmhip import serve as S

pytestmark = pytest.mark.gpu

RES = {0: ssmhip.Systematic(), 2: ssmhip.Stratified(), 3: ssmhip.Multinomial()}


def lg_model(rng, Dx, Dy, rho=0.9):
    """Seeded stable LG model: A = rho x orthogonal, SPD Q and R, prior N(0, I)."""
    G = rng.standard_normal((Dx, Dx))
    Qm, Rm = np.linalg.qr(G)
    A = rho * Qm * np.sign(np.diag(Rm))
    Wq = rng.standard_normal((Dx, Dx + 2))
    Wr = rng.standard_normal((Dy, Dy + 2))
    Q = 0.1 * Wq @ Wq.T / (Dx + 2) + 0.01 * np.eye(Dx)
    R = 0.5 * Wr @ Wr.T / (Dy + 2) + 0.05 * np.eye(Dy)
    H = rng.standard_normal((Dy, Dx))
    return (np.zeros(Dx), np.eye(Dx), A, 0.1 * rng.standard_normal(Dx), Q, H,
            0.1 * rng.standard_normal(Dy), R)


def simulate(rng, p, T):
    mu0, S0, A, b, Q, H, c, R = p
    LQ, LR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    x = mu0 + np.linalg.cholesky(S0) @ rng.standard_normal(len(mu0))
    ys = np.zeros((T, H.shape[0]))
    for t in range(T):
        x = A @ x + b + LQ @ rng.standard_normal(len(x))
        ys[t] = H @ x + c + LR @ rng.standard_normal(H.shape[0])
    return ys


def plain_bf_ll(ctx, p, ys, N, threshold, kind, seed):
    model = ssmhip.create_homogeneous_linear_gaussian_model(*p)
    algo = ssmhip.BF(N, threshold=threshold, resampler=RES[kind], precision="f32")
    st = ssmhip.initialise(None, model, algo, ctx=ctx, seed=seed)
    return ssmhip.run(st, algo, ys)


CASES = [  # (Dx, Dy, N, T, threshold, kind)
    (1, 1, 256, 40, 0.5, 0),
    (4, 2, 3000, 50, 1.0, 0),
    (4, 2, 70000, 30, 0.5, 2),
    (3, 2, 5000, 45, 1.0, 3),
    (7, 5, 1024, 35, 0.5, 0),
    (16, 8, 2049, 25, 1.0, 2),
    # many tiles a fold thread (196 and 256 tiles): the served and plain finishes must
    # fold in the same order
    (4, 2, 200000, 30, 1.0, 0),
    (2, 1, 262144, 20, 0.5, 3),
]


def test_served_bf_equals_plain_path(ctx):
    rng = np.random.default_rng(1)
    reqs = []
    for (Dx, Dy, N, T, thr, kind) in CASES:
        p = lg_model(rng, Dx, Dy)
        ys = simulate(rng, p, T)
        reqs.append((p, ys, N, thr, kind, int(rng.integers(1, 2**62))))
    with S.Server(hbm_budget_bytes=1 << 30, max_active=64, ctx=ctx) as sv:
        ids = [sv.submit_params(S.REQ_BF, p, ys, N, thr, kind, seed)
               for (p, ys, N, thr, kind, seed) in reqs]
        got = [sv.wait(i, 60000.0) for i in ids]
    for (p, ys, N, thr, kind, seed), r, case in zip(reqs, got, CASES):
        assert r.state == S.DONE, (case, r)
        ref = plain_bf_ll(ctx, p, ys, N, thr, kind, seed)
        print(case, r.ll, ref)
        assert r.ll == ref, case


def _mix(rng, n, T_range=(20, 80)):
    out = []
    for _ in range(n):
        kind = rng.choice([S.REQ_BF, S.REQ_BF, S.REQ_KF])
        Dx = int(rng.integers(1, 17))
        Dy = int(rng.integers(1, min(Dx, 8) + 1))
        p = lg_model(rng, Dx, Dy)
        ys = simulate(rng, p, int(rng.integers(*T_range)))
        N = int(np.exp(rng.uniform(np.log(256), np.log(20000))))
        out.append((kind, p, ys, N, float(rng.choice([0.5, 1.0])), int(rng.choice([0, 2, 3])),
                    int(rng.integers(1, 2**62))))
    return out


def test_same_alone_and_in_the_mix(ctx):
    rng = np.random.default_rng(7)
    pb = lg_model(rng, 4, 2)
    probe_bf = (S.REQ_BF, pb, simulate(rng, pb, 60), 4096, 0.5, 0, 12345)
    pk = lg_model(rng, 9, 4)
    probe_kf = (S.REQ_KF, pk, simulate(rng, pk, 70), 0, 1.0, 0, 0)
    with S.Server(hbm_budget_bytes=512 << 20, max_active=32, ctx=ctx) as sv:
        alone = [sv.wait(sv.submit_params(k, p, ys, N, thr, kd, sd), 60000.0)
                 for (k, p, ys, N, thr, kd, sd) in (probe_bf, probe_kf)]
    assert all(a.state == S.DONE for a in alone)
    others = _mix(rng, 30)
    with S.Server(hbm_budget_bytes=512 << 20, max_active=32, ctx=ctx) as sv:
        ids = []
        # submit in two waves so the probes arrive while earlier requests are running
        for q in others[:15]:
            ids.append(sv.submit_params(q[0], *q[1:]))
        time.sleep(0.005)
        pid = [sv.submit_params(k, p, ys, N, thr, kd, sd)
               for (k, p, ys, N, thr, kd, sd) in (probe_bf, probe_kf)]
        for q in others[15:]:
            ids.append(sv.submit_params(q[0], *q[1:]))
        mixed = [sv.wait(i, 60000.0) for i in pid]
        rest = [sv.wait(i, 60000.0) for i in ids]
        st = sv.stats()
    assert all(r.state == S.DONE for r in rest)
    assert [m.ll for m in mixed] == [a.ll for a in alone]
    assert st["completed"] == len(others) + 2
    print("stats", st)


def test_kf_requests_match_oracle(ctx, orc):
    rng = np.random.default_rng(3)
    reqs = []
    for Dx, Dy, T in [(1, 1, 50), (5, 2, 80), (16, 8, 120), (33, 16, 40), (64, 32, 30)]:
        p = lg_model(rng, Dx, Dy)
        reqs.append((p, simulate(rng, p, T)))
    with S.Server(hbm_budget_bytes=256 << 20, max_active=8, ctx=ctx) as sv:
        ids = [sv.submit_params(S.REQ_KF, p, ys) for p, ys in reqs]
        got = [sv.wait(i, 60000.0) for i in ids]
    for (p, ys), r in zip(reqs, got):
        assert r.state == S.DONE and r.status == 0
        ref = orc.kf_filter(orc.LG(*p), ys)[2]
        assert r.ll == pytest.approx(ref, rel=1e-4), (p[2].shape, r.ll, ref)


def test_admission_by_budget(ctx):
    rng = np.random.default_rng(11)
    p = lg_model(rng, 8, 4)
    ys = simulate(rng, p, 40)
    budget = 8 << 20
    with S.Server(hbm_budget_bytes=budget, max_active=16, ctx=ctx) as sv:
        # 262144 particles x 8 dims x 2 planes x 4 B alone exceed 8 MiB: refused at submit
        with pytest.raises(ssmhip.AdmissionError):
            sv.submit_params(S.REQ_BF, p, ys, 1 << 18, 1.0, 0, 1)
        # each of these needs ~1.3 MiB: they cannot all be resident at once
        ids = [sv.submit_params(S.REQ_BF, p, ys, 20000, 1.0, 0, 100 + k) for k in range(12)]
        got = [sv.wait(i, 60000.0) for i in ids]
        st = sv.stats()
    assert all(r.state == S.DONE for r in got)
    assert st["peak_bytes"] <= st["arena_bytes"] <= budget
    assert st["peak_active_bf"] < 12
    assert all(np.isfinite(r.ll) for r in got)


def test_cancel(ctx):
    rng = np.random.default_rng(5)
    p = lg_model(rng, 4, 2)
    ys_long = simulate(rng, p, 2000)
    ys = simulate(rng, p, 30)
    with S.Server(hbm_budget_bytes=64 << 20, max_active=1, ctx=ctx) as sv:
        a = sv.submit_params(S.REQ_BF, p, ys_long, 100000, 1.0, 0, 1)
        b = sv.submit_params(S.REQ_BF, p, ys, 1000, 1.0, 0, 2)   # queued behind a
        sv.cancel(b)
        sv.cancel(a)
        ra, rb = sv.wait(a, 60000.0), sv.wait(b, 60000.0)
        c = sv.submit_params(S.REQ_BF, p, ys, 1000, 1.0, 0, 3)
        rc = sv.wait(c, 60000.0)
    assert rb.state == S.CANCELLED
    assert ra.state in (S.CANCELLED, S.DONE)
    assert rc.state == S.DONE and np.isfinite(rc.ll)


def test_submit_from_threads_while_running(ctx):
    rng = np.random.default_rng(9)
    qs = _mix(rng, 24)
    with S.Server(hbm_budget_bytes=256 << 20, max_active=16, ctx=ctx) as sv:
        ids = [None] * len(qs)

        def worker(lo, hi):
            for k in range(lo, hi):
                ids[k] = sv.submit_params(qs[k][0], *qs[k][1:])
                time.sleep(0.001)

        th = [threading.Thread(target=worker, args=(k * 8, k * 8 + 8)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        got = [sv.wait(i, 60000.0) for i in ids]
    assert all(r.state == S.DONE for r in got)
    assert all(r.latency_ms > 0 for r in got)
