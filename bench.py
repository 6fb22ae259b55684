"""Bootstrap-filter throughput on one MI355X.

Workload: the D=4 linear-Gaussian SSM (A = 0.95 * seeded orthogonal, Q = 0.1 I, H seeded
2x4 Gaussian, R = 0.2 I, prior N(0, I)), T = 1000 synthetic observations simulated from
seed 1234, BootstrapFilter with N = 2^20 FP32 particles (SoA), systematic resampling at
the reference's default threshold 1.0 (resample every step).  One bench step is one full
filter(model, BF, ys) call: initialise plus T = 1000 predict/update/resample steps.

Prints one JSON line: particle-steps/s (whole job), the roofline of the dominant kernel
(the fused predict/update step kernel, timed live with HIP events), the log-likelihood
error against an exact FP64 Kalman filter, and a CPU baseline (the FP64 oracle, one core,
on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ssmproblems.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
RESAMPLERS = {"systematic": 0, "systematic_gpu": 1, "stratified": 2, "multinomial": 3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_model(seed=1234):
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((4, 4))
    Qm, Rm = np.linalg.qr(G)
    A = 0.95 * Qm * np.sign(np.diag(Rm))
    H = rng.standard_normal((2, 4))
    return dict(mu0=np.zeros(4), Sigma0=np.eye(4), A=A, b=np.zeros(4), Q=0.1 * np.eye(4), H=H,
                c=np.zeros(2), R=0.2 * np.eye(2))


def simulate(p, T, seed):
    rng = np.random.default_rng(seed)
    L0, LQ, LR = (np.linalg.cholesky(p[k]) for k in ("Sigma0", "Q", "R"))
    x = p["mu0"] + L0 @ rng.standard_normal(4)
    ys = np.zeros((T, 2))
    for t in range(T):
        x = p["A"] @ x + p["b"] + LQ @ rng.standard_normal(4)
        ys[t] = p["H"] @ x + p["c"] + LR @ rng.standard_normal(2)
    return ys


def kalman_loglik(p, ys):
    """Exact log p(y_1:T) by an FP64 Kalman filter (numpy; independent of the library)."""
    mu, S = p["mu0"].copy(), p["Sigma0"].copy()
    ll = 0.0
    for y in ys:
        mu = p["A"] @ mu + p["b"]
        S = p["A"] @ S @ p["A"].T + p["Q"]
        z = y - p["H"] @ mu - p["c"]
        Sy = p["H"] @ S @ p["H"].T + p["R"]
        L = np.linalg.cholesky(Sy)
        v = np.linalg.solve(L, z)
        ll += -0.5 * (len(y) * np.log(2 * np.pi) + 2 * np.log(np.diag(L)).sum() + v @ v)
        K = np.linalg.solve(Sy, p["H"] @ S).T
        IKH = np.eye(len(mu)) - K @ p["H"]
        mu = mu + K @ z
        S = IKH @ S @ IKH.T + K @ p["R"] @ K.T
    return ll


def cpu_baseline(p, ys, N, kind, threshold, budget_s=12.0):
    """The FP64 oracle (single thread) on the same model and particle count, for as many
    steps as fit in ~budget_s seconds; returns particle-steps/s and the sample size."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    m = oracle.LG(p["mu0"], p["Sigma0"], p["A"], p["b"], p["Q"], p["H"], p["c"], p["R"])
    rng = np.random.default_rng(7)
    st = oracle.BFState(oracle.bf_initialise(m, rng.standard_normal((N, 4))))
    steps = 0
    t0 = time.perf_counter()
    while True:
        u = rng.random(N if kind >= 2 else 1)
        z = rng.standard_normal((N, 4))
        oracle.bf_step(m, st, ys[steps % len(ys)], threshold, kind, u, z)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= len(ys):
            break
    return N * steps / el, steps, el


def cpu_info():
    """(logical CPUs, model name) of this host, from os and /proc/cpuinfo."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return os.cpu_count() or 1, name


def cpu_baseline_all_cores(p, ys, N, threshold, budget_s=10.0):
    """The all-cores CPU port of the same filter (oracle/bf_omp.c: FP64 SoA, OpenMP over
    particles, systematic resampling by blocks) on this host's CPU share (OMP_NUM_THREADS,
    else every logical CPU): runs growing step counts until one takes ~budget_s / 2.
    Returns (particle-steps/s, threads, steps, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    m = oracle.LG(p["mu0"], p["Sigma0"], p["A"], p["b"], p["Q"], p["H"], p["c"], p["R"])
    nproc, _ = cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", nproc) or nproc)
    x0 = oracle.bf_initialise(m, np.random.default_rng(5).standard_normal((N, 4)))
    T = 2
    while True:
        t0 = time.perf_counter()
        oracle.bfomp_run(m, x0, ys[:T], threshold, 17, threads)
        el = time.perf_counter() - t0
        if el > budget_s / 2 or T >= len(ys):
            return N * T / el, threads, T, el
        T = min(len(ys), max(T + 1, int(T * min(8.0, budget_s / 2 / max(el, 1e-3)))))


def copy_calibration(nbytes=256 << 20, reps=5):
    """A plain device-to-device copy of known size: the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters (each launch reads and writes nbytes)."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": reps}))


def traffic_from_profile(N, resampler, threshold):
    """HBM bytes per launch of the step kernel from the committed counter summary for this
    workload (two rocprofv3 --pmc passes, FETCH_SIZE then WRITE_SIZE, scaled by the copy
    calibration), or None when no summary covers it."""
    path = os.path.join(ROOT, "profiles", "k_bf_step_lg4_traffic.txt")
    if not os.path.exists(path):
        return None
    want = f"N={N} resampler={resampler} threshold={threshold}"
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch") and want in line:
                return float(line.split(":", 1)[1].split()[0])
    return None


def kf_flops(Dx, Dy):
    """Algorithmic flops of one Kalman predict + Joseph update, each product counted as 2mnk
    with full outputs: A Sigma A^T 4Dx^3, (I-KH) Sigma (I-KH)^T 4Dx^3, H Sigma H^T
    2Dx^2Dy + 2DxDy^2, Sigma H^T 2Dx^2Dy, the two solves for K 2DxDy^2, K H 2Dx^2Dy,
    K R K^T 2DxDy^2 + 2Dx^2Dy."""
    Dx = np.asarray(Dx, dtype=np.float64)
    Dy = np.asarray(Dy, dtype=np.float64)
    return 8 * Dx ** 3 + 8 * Dx ** 2 * Dy + 6 * Dx * Dy ** 2


def kf_loglik_general(mu0, S0, A, b, Q, H, c, R, ys):
    """Exact FP64 Kalman log-likelihood with the Joseph update (numpy)."""
    mu, S = mu0.copy(), S0.copy()
    ll = 0.0
    for y in ys:
        mu = A @ mu + b
        S = A @ S @ A.T + Q
        z = y - H @ mu - c
        Sy = H @ S @ H.T + R
        L = np.linalg.cholesky(Sy)
        v = np.linalg.solve(L, z)
        ll += -0.5 * (len(y) * np.log(2 * np.pi) + 2 * np.log(np.diag(L)).sum() + v @ v)
        K = np.linalg.solve(Sy, H @ S).T
        IKH = np.eye(len(mu)) - K @ H
        mu = mu + K @ z
        S = IKH @ S @ IKH.T + K @ R @ K.T
        S = np.triu(S) + np.triu(S, 1).T
    return ll


def bench_kf(args, T=200, budget_s=10.0):
    """Batched Kalman filters: B LG models, Dx ~ U{2..64}, Dy = Dx // 2, A = 0.9 x seeded
    orthogonal, Q and R seeded Wishart, T = 200.  One bench step = one run of all B."""
    import ssmhip
    from ssmhip import workloads as W
    ctx = ssmhip.default_context()
    name, cus = ctx.device_info()
    log(f"device {name} CUs={cus}; generating {args.B} models")
    w = W.batched_lg(args.B, T, seed=1234)
    kb = ssmhip.KFBatch.from_packed(w["Dx"], w["Dy"], w["params"], w["param_off"], w["ys"],
                                    w["y_off"], T, ctx=ctx)
    tuned = kb.tune(T_tune=8)
    log(f"autotune (waves per workgroup): {tuned}")
    for _ in range(args.warmup):
        ll = kb.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    dev_ms = []
    for _ in range(args.steps):
        ll = kb.run()
        dev_ms.append(kb.last_ms())
    el = time.perf_counter() - t0
    flops = float((kf_flops(w["Dx"], w["Dy"]) * T).sum())
    ms_dev = float(np.mean(dev_ms))
    achieved = flops / (ms_dev * 1e-3) / 1e12
    # accuracy against an exact FP64 filter on a sample spanning the sizes
    rng = np.random.default_rng(5)
    sample = np.unique(np.concatenate([rng.choice(args.B, 48, replace=False),
                                       [0, args.B - 1]]))
    errs = []
    for f in sample:
        prm, ys_f = W.unpack_lg(w, int(f))
        ref = kf_loglik_general(*prm, ys_f)
        errs.append(abs(ll[f] - ref) / abs(ref))
    log(f"{args.steps} runs in {el:.3f}s ({ms_dev:.1f} ms device per run); "
        f"{achieved:.2f} TFLOP/s algorithmic; max ll rel err {max(errs):.2e}")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        done, t1 = 0, time.perf_counter()
        order = rng.permutation(args.B)
        while time.perf_counter() - t1 < budget_s and done < args.B:
            prm, ys_f = W.unpack_lg(w, int(order[done]))
            oracle.kf_filter(oracle.LG(*prm), ys_f)
            done += 1
        secs = time.perf_counter() - t1
        cpu = {"value": done / secs, "unit": "filters/s", "cores": 1, "kind": "port",
               "sample": f"FP64 oracle kf_filter (oracle/oracle.c, single thread) on {done} "
                         f"randomly chosen filters of the same batch, T={T}, {secs:.1f}s"}
        log(f"cpu baseline {done / secs:.1f} filters/s")
    out = {
        "metric": "filters/s at stated loglik err (batched Kalman filters)",
        "value": args.steps * args.B / el,
        "unit": "filters/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 inverse of S)",
        "data": "synthetic (seeded LG models and simulated observations)",
        "config": {"workload": f"batched KalmanFilter, B={args.B}, Dx~U{{2..64}}, Dy=Dx//2, "
                               f"T={T}, Joseph update", "B": args.B, "T": T},
        "loglik_max_rel_err": float(max(errs)), "loglik_err_sample": int(len(sample)),
        "autotune": {k: {"waves": v["waves"], "ms": {str(a): b for a, b in v["ms"].items()}}
                     for k, v in tuned.items()},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s",
                     "frac": achieved / 157.3, "traffic": None,
                     "kernel": "k_kf_batch (3 size classes, one run)",
                     "flops_per_run": flops, "ms_per_run_device": ms_dev},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def rbpf_model(D=32, Dy=8, seed=1234):
    """Outer AR(1) x_t = 0.9 x_{t-1} + N(0, 0.3^2), stationary start; inner D-dim LG with
    A = 0.9 x seeded orthogonal, Q = 0.1 I, H seeded Dy x D, R_t = exp(x_t) 0.2 I."""
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((D, D))
    Qm, Rm = np.linalg.qr(G)
    A = 0.9 * Qm * np.sign(np.diag(Rm))
    H = rng.standard_normal((Dy, D))
    oq = 0.3 ** 2
    return dict(om0=0.0, os0=oq / (1 - 0.81), oa=0.9, ob=0.0, oq=oq, mu0=np.zeros(D),
                Sigma0=np.eye(D), A=A, b=np.zeros(D), Cv=np.zeros(D), Q=0.1 * np.eye(D), H=H,
                c=np.zeros(Dy), R0=0.2 * np.eye(Dy), rs=1.0)


RB_KEYS = ("om0", "os0", "oa", "ob", "oq", "mu0", "Sigma0", "A", "b", "Cv", "Q", "H", "c",
           "R0", "rs")


def rbpf_simulate(m, T, seed):
    rng = np.random.default_rng(seed)
    D, Dy = m["A"].shape[0], m["H"].shape[0]
    xo = m["om0"] + np.sqrt(m["os0"]) * rng.standard_normal()
    z = m["mu0"] + np.linalg.cholesky(m["Sigma0"]) @ rng.standard_normal(D)
    LQ = np.linalg.cholesky(m["Q"])
    ys = np.zeros((T, Dy))
    for t in range(T):
        xprev = xo
        xo = m["oa"] * xo + m["ob"] + np.sqrt(m["oq"]) * rng.standard_normal()
        z = m["A"] @ z + m["b"] + m["Cv"] * xprev + LQ @ rng.standard_normal(D)
        R = np.exp(m["rs"] * xo) * m["R0"]
        ys[t] = m["H"] @ z + m["c"] + np.linalg.cholesky(R) @ rng.standard_normal(Dy)
    return ys


def bench_rbpf(args, N=16384, T=500, threshold=0.5, budget_s=10.0):
    """RBPF: N particles, each an outer scalar and an inner 32-dim Kalman filter (FP32
    32x32 covariance), T = 500, systematic resampling at ESS < N/2.  One bench step = one
    full filter run."""
    import ssmhip
    ctx = ssmhip.default_context()
    name, cus = ctx.device_info()
    log(f"device {name} CUs={cus}")
    m = rbpf_model()
    D, Dy = m["A"].shape[0], m["H"].shape[0]
    ys = rbpf_simulate(m, T, 99)
    dm = ssmhip.RBModel(*(m[k] for k in RB_KEYS))
    st = ssmhip.DeviceRBParticles(ctx, dm, N, 2024)
    lls = []

    def one():
        st.initialise()
        return st.run(ys, threshold, 0)

    for _ in range(args.warmup):
        one()
    ctx.synchronize()
    t0 = time.perf_counter()
    dev = []
    for _ in range(args.steps):
        lls.append(one())
        dev.append(st.last_ms())
    el = time.perf_counter() - t0
    value = args.steps * N * T / el
    ms_dev = float(np.mean(dev))
    fl = float(kf_flops(D, Dy))
    achieved = fl * N * T / (ms_dev * 1e-3) / 1e12
    log(f"{args.steps} filters in {el:.3f}s: {value:.3e} particle-steps/s; device "
        f"{ms_dev:.1f} ms per run; {achieved:.2f} TFLOP/s algorithmic; ll {lls[-1]:.3f}")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        om = oracle.RB(*(m[k] for k in RB_KEYS))
        rng = np.random.default_rng(3)
        ost = oracle.rbpf_initialise(om, rng.standard_normal(N))
        steps, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < budget_s and steps < T:
            oracle.rbpf_step(om, ost, ys[steps], threshold, 0, rng.random(1),
                             rng.standard_normal(N))
            steps += 1
        secs = time.perf_counter() - t1
        cpu = {"value": N * steps / secs, "unit": "particle-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"FP64 oracle rbpf_step (oracle/oracle.c, single thread) on the same "
                         f"model with N={N}: {steps} steps in {secs:.1f}s"}
        log(f"cpu baseline {N * steps / secs:.3e} particle-steps/s ({steps} steps)")
    out = {
        "metric": "particle-steps/s (RBPF, per-particle Kalman on MFMA)",
        "value": value, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 inverse of S)",
        "data": "synthetic (seeded hierarchical model and simulated observations)",
        "config": {"workload": f"RBPF, outer AR(1), inner LG D={D} Dy={Dy}, R_t = exp(x_t) R0, "
                               f"N={N}, T={T}, systematic at threshold {threshold}",
                   "N": N, "T": T, "D": D, "Dy": Dy},
        "loglik": lls[-1],
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 157.3, "unit": "TFLOP/s",
                     "frac": achieved / 157.3, "traffic": None,
                     "kernel": "k_rbpf_step + partials + finish + resampler (whole run)",
                     "flops_per_particle_step": fl, "ms_per_run_device": ms_dev},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def serve_requests(M, seed=1234):
    """The request stream: each request is a C1-type 1-D BF (the random-walk model,
    threshold 0.5), a C2-type LG BF or a KF on a seeded LG model with Dx ~ U{1..16},
    Dy = max(1, Dx // 2); N log-uniform in [256, 262144]; T ~ U[50, 2000]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(M):
        cls = int(rng.integers(0, 3))
        T = int(rng.integers(50, 2001))
        N = int(np.exp(rng.uniform(np.log(256), np.log(262144))))
        if cls == 0:
            p = dict(mu0=np.zeros(1), Sigma0=np.eye(1), A=np.eye(1), b=np.zeros(1),
                     Q=np.array([[0.01]]), H=np.eye(1), c=np.zeros(1), R=np.array([[0.04]]))
        else:
            dx = int(rng.integers(1, 17))
            dy = max(1, dx // 2)
            G = rng.standard_normal((dx, dx))
            Qm, Rm = np.linalg.qr(G)
            A = 0.9 * Qm * np.sign(np.diag(Rm))
            Wq = rng.standard_normal((dx, dx + 2))
            Wr = rng.standard_normal((dy, dy + 2))
            p = dict(mu0=np.zeros(dx), Sigma0=np.eye(dx), A=A, b=np.zeros(dx),
                     Q=Wq @ Wq.T / (dx + 2), H=rng.standard_normal((dy, dx)), c=np.zeros(dy),
                     R=Wr @ Wr.T / (dy + 2))
        dx, dy = p["A"].shape[0], p["H"].shape[0]
        L0, LQ, LR = (np.linalg.cholesky(p[k]) for k in ("Sigma0", "Q", "R"))
        x = p["mu0"] + L0 @ rng.standard_normal(dx)
        E = rng.standard_normal((T, dx)) @ LQ.T
        V = rng.standard_normal((T, dy)) @ LR.T
        ys = np.zeros((T, dy))
        for t in range(T):
            x = p["A"] @ x + p["b"] + E[t]
            ys[t] = p["H"] @ x + p["c"] + V[t]
        kind = 1 if cls == 2 else 0
        out.append(dict(kind=kind, cls=cls, params=tuple(p[k] for k in (
            "mu0", "Sigma0", "A", "b", "Q", "H", "c", "R")), ys=ys, N=N,
            threshold=0.5 if cls == 0 else 1.0, seed=int(rng.integers(1, 2**62))))
    return out


def plain_run(ctx, q, timings=None):
    """The plain per-filter GPU path: one filter object per request, created, run from
    captured graphs, read back and destroyed.  With `timings`, a bootstrap filter is then
    re-initialised and run again on the same object (graphs already captured), and
    (wall s of the whole first call, wall s of the second run alone) is appended: the
    second is the plain path's cost without allocation, graph capture or instantiation."""
    import ssmhip
    t0 = time.perf_counter()
    model = ssmhip.create_homogeneous_linear_gaussian_model(*q["params"])
    if q["kind"] == 1:
        ll = ssmhip.filter(None, model, ssmhip.KalmanFilter(), q["ys"], ctx=ctx)[1]
        if timings is not None:
            el = time.perf_counter() - t0
            timings.append((el, el))
        return ll
    algo = ssmhip.BF(q["N"], threshold=q["threshold"], precision="f32")
    st = ssmhip.initialise(None, model, algo, ctx=ctx, seed=q["seed"])
    ll = ssmhip.run(st, algo, q["ys"])
    if timings is not None:
        t1 = time.perf_counter()
        ctx.check(ctx.lib.ssm_bf_initialise(st.ptr, None))
        t2 = time.perf_counter()
        ssmhip.run(st, algo, q["ys"])
        t3 = time.perf_counter()
        timings.append((t1 - t0, t3 - t2))
    st.close()
    return ll


def bench_serve(args, budget=16 << 30, max_active=256, budget_s=10.0):
    import ssmhip
    from ssmhip import serve as S
    ctx = ssmhip.default_context()
    name, cus = ctx.device_info()
    M = args.requests
    log(f"device {name} CUs={cus}; generating {M} requests")
    reqs = serve_requests(M)
    K = max(8, M // 4)

    def submit(sv, q):
        return sv.submit_params(S.REQ_KF if q["kind"] == 1 else S.REQ_BF, q["params"], q["ys"],
                                q["N"], q["threshold"], 0, q["seed"])

    if args.serve_saturation_only:
        runs = []
        for _ in range(2):
            with S.Server(budget, max_active, ctx=ctx) as sv:
                t0 = time.perf_counter()
                ids = [submit(sv, q) for q in reqs]
                res = [sv.wait(i) for i in ids]
                el = time.perf_counter() - t0
                st = sv.stats()
            assert all(r.state == S.DONE for r in res), "served requests failed"
            runs.append({"filters_per_s": M / el, "bf_chunk_ms": st["bf_chunk_ms"]})
            log(f"saturation: {M} in {el:.2f}s = {M / el:.1f} filters/s, "
                f"chunks {st['bf_chunk_ms']:.1f} ms")
        print(json.dumps({"workload": "serve saturation runs", "requests": M, "runs": runs}),
              flush=True)
        return
    # warm-up: every code path once (kernels, graphs, server threads)
    plain_run(ctx, reqs[0])
    with S.Server(budget, max_active, ctx=ctx) as sv:
        for q in reqs[:4]:
            sv.wait(submit(sv, q))
    # plain per-filter path on the first K requests: each call whole (create, capture,
    # run, destroy), and the bootstrap filters' second run on the same object alone
    tims = []
    ll_plain = [plain_run(ctx, q, tims) for q in reqs[:K]]
    el_plain = sum(a for a, _ in tims)
    el_plain_run = sum(b for _, b in tims)
    plain_fps = K / el_plain
    plain_run_fps = K / el_plain_run
    log(f"plain path: {K} filters in {el_plain:.2f}s = {plain_fps:.1f} filters/s; "
        f"reused filters (run only) {el_plain_run:.2f}s = {plain_run_fps:.1f} filters/s")
    # served, saturation: the same K requests, then all M, submitted at once
    with S.Server(budget, max_active, ctx=ctx) as sv:
        t0 = time.perf_counter()
        ids = [submit(sv, q) for q in reqs[:K]]
        res_k = [sv.wait(i) for i in ids]
        el_k = time.perf_counter() - t0
        t0 = time.perf_counter()
        ids = [submit(sv, q) for q in reqs]
        res = [sv.wait(i) for i in ids]
        el_all = time.perf_counter() - t0
        st_sat = sv.stats()
    assert all(r.state == S.DONE for r in res_k + res), "served requests failed"
    served_fps_k = K / el_k
    served_fps = M / el_all
    log(f"served: same {K} in {el_k:.2f}s = {served_fps_k:.1f} filters/s "
        f"({served_fps_k / plain_fps:.1f}x the plain path, {served_fps_k / plain_run_fps:.1f}x "
        f"its run alone); all {M} in {el_all:.2f}s = {served_fps:.1f} filters/s")
    bf_exact = sum(1 for q, a, b in zip(reqs[:K], res_k, ll_plain) if q["kind"] == 0 and a.ll == b)
    n_bf = sum(1 for q in reqs[:K] if q["kind"] == 0)
    kf_err = [abs(r.ll - kf_loglik_general(*q["params"], q["ys"])) / abs(r.ll)
              for q, r in zip(reqs, res) if q["kind"] == 1][:64]
    # latency under Poisson arrivals at fractions of the measured saturation rate
    lat = {}
    rng = np.random.default_rng(99)
    for frac in (0.25, 0.5, 1.0, 2.0):
        rate = frac * served_fps
        gaps = rng.exponential(1.0 / rate, size=M)
        with S.Server(budget, max_active, ctx=ctx) as sv:
            ids = []
            t0 = time.perf_counter()
            due = t0
            for q, g in zip(reqs, gaps):
                due += g
                dt = due - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids.append(submit(sv, q))
            rr = [sv.wait(i) for i in ids]
            el = time.perf_counter() - t0
        L = np.array([r.latency_ms for r in rr])
        lat[f"{frac}x"] = {"arrival_rate": rate, "p50_ms": float(np.percentile(L, 50)),
                           "p99_ms": float(np.percentile(L, 99)), "filters_per_s": M / el}
        log(f"poisson {frac}x ({rate:.0f}/s): p50 {np.percentile(L, 50):.1f} ms, "
            f"p99 {np.percentile(L, 99):.1f} ms, {M / el:.1f} filters/s")
    cpu = None
    if not args.no_cpu:
        cpu = serve_cpu_baseline(reqs, budget_s)
        log(f"cpu baseline {cpu['value']:.3f} filters/s")
    out = {
        "metric": "filters/s + p50/p99 at stated loglik err (served request stream)",
        "value": served_fps, "unit": "filters/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": el_all * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 weight sums, FP64 inverse of S)",
        "data": "synthetic (seeded models and simulated observations)",
        "config": {"workload": f"{M} requests: 1-D BF / LG BF (Dx 1..16) / KF, N log-uniform "
                               f"[256, 262144], T ~ U[50, 2000]; budget {budget >> 30} GiB, "
                               f"{max_active} slots", "requests": M},
        "plain_filters_per_s": plain_fps,
        "plain_run_only_filters_per_s": plain_run_fps,
        "plain_setup_s": el_plain - el_plain_run,
        "served_filters_per_s_same_requests": served_fps_k,
        "speedup_vs_plain": served_fps_k / plain_fps,
        "speedup_vs_plain_run_only": served_fps_k / plain_run_fps,
        "bf_served_equal_to_plain": f"{bf_exact}/{n_bf}",
        "kf_loglik_max_rel_err": float(max(kf_err)) if kf_err else None,
        "latency": lat,
        "server_stats": st_sat,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def serve_cpu_baseline(reqs, budget_s):
    """FP64 oracle, one core: requests in stream order, each run for as many steps as fit
    its share of the budget and its time extrapolated to the full horizon."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rng = np.random.default_rng(3)
    total, n, t_start = 0.0, 0, time.perf_counter()
    for q in reqs:
        if time.perf_counter() - t_start > budget_s:
            break
        m = oracle.LG(*q["params"])
        T = q["ys"].shape[0]
        t0 = time.perf_counter()
        if q["kind"] == 1:
            oracle.kf_filter(m, q["ys"])
            total += time.perf_counter() - t0
        else:
            dx = q["params"][2].shape[0]
            st = oracle.BFState(oracle.bf_initialise(m, rng.standard_normal((q["N"], dx))))
            steps = 0
            while steps < T and time.perf_counter() - t0 < budget_s / 8:
                oracle.bf_step(m, st, q["ys"][steps], q["threshold"], 0, rng.random(1),
                               rng.standard_normal((q["N"], dx)))
                steps += 1
            total += (time.perf_counter() - t0) * T / steps
        n += 1
    return {"value": n / total, "unit": "filters/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (oracle/oracle.c, single thread) on the first {n} requests "
                      f"of the stream; BF runs cut short are extrapolated to their full T"}


def bench_tree(args, T=200, budget_s=10.0):
    """Path storage: the headline bootstrap filter (D=4 LG, N particles, systematic at
    threshold 1.0) with its particles appended to the device ancestry tree after every step
    (ParallelAncestorCallback).  One bench step = T filter steps, each followed by an
    insert; the insert cost is the difference to the same steps without the tree."""
    import ssmhip
    ctx = ssmhip.default_context()
    name, cus = ctx.device_info()
    log(f"device {name} CUs={cus}")
    p = make_model()
    N = args.N
    ys = simulate(p, T, 1234)
    model = ssmhip.create_homogeneous_linear_gaussian_model(*(p[k] for k in (
        "mu0", "Sigma0", "A", "b", "Q", "H", "c", "R")))
    algo = ssmhip.BF(N, threshold=1.0, resampler=ssmhip.Systematic(), precision="f32")
    rng = np.random.default_rng(2024)
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=2024)

    def run(mode):
        ctx.check(ctx.lib.ssm_bf_initialise(st.ptr, None))
        store = None
        if mode == "tree":
            store = ssmhip.ParallelParticleTree(st)
        elif mode == "history":
            store = ssmhip.DenseParticleContainer(st, capacity=T)
        add = None if store is None else (store.insert if mode == "tree" else store.push)
        t_ins = 0.0
        ctx.synchronize()
        t0 = time.perf_counter()
        for t in range(1, T + 1):
            ssmhip.step(rng, model, algo, t, st, ys[t - 1])
            if add is not None:
                t1 = time.perf_counter()
                add()
                ctx.synchronize()
                t_ins += time.perf_counter() - t1
        ctx.synchronize()
        el = time.perf_counter() - t0
        info = store.info() if mode == "tree" else None
        if store is not None:
            store.close()
        return el, t_ins, info

    for _ in range(args.warmup):
        for mode in ("none", "tree", "history"):
            run(mode)
    base, full, ins, push, info = [], [], [], [], None
    for _ in range(args.steps):
        base.append(run("none")[0])
        el, t_ins, info = run("tree")
        full.append(el)
        ins.append(t_ins)
        push.append(run("history")[1])
    us_insert = float(np.mean(ins)) / T * 1e6
    us_push = float(np.mean(push)) / T * 1e6
    us_step = float(np.mean(base)) / T * 1e6
    value = N / (us_insert * 1e-6)
    M = info["capacity"]
    # algorithmic bytes of one insert: per particle 4 (clear counts) + 20 (ancestor, leaf
    # gather, parent write, count atomic read+write) + 12 (leaf, count, offspring scatter)
    # + 4 (prune test) + 4 (new-leaf write) + 48 (new leaf, parent, state 16 read; parent,
    # leaf, state 16 written) = 92 B, plus two reads of the M offspring counts
    bytes_insert = 92.0 * N + 8.0 * M
    achieved = bytes_insert / (us_insert * 1e-6) / 1e9
    # dense history push: state 16 B read + written, log-weight 4 B read + 8 B written,
    # ancestor 4 B read + written
    bytes_push = 52.0 * N
    log(f"step {us_step:.1f} us, insert {us_insert:.1f} us ({achieved:.0f} GB/s); pool "
        f"{M} nodes, {info['used']} used, depth {info['depth']}; dense history push "
        f"{us_push:.1f} us ({bytes_push / (us_push * 1e-6) / 1e9:.0f} GB/s)")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        rng2 = np.random.default_rng(5)
        x = rng2.standard_normal((N, 4))
        ot = oracle.PTree(x, 4 * N)
        n, secs = 0, 0.0
        while secs < budget_s and n < T:
            a = np.sort(rng2.integers(0, N, size=N)).astype(np.int32)
            t1 = time.perf_counter()
            ot.insert(x, a)
            secs += time.perf_counter() - t1
            n += 1
        cpu = {"value": N * n / secs, "unit": "particle-inserts/s", "cores": 1, "kind": "port",
               "sample": f"oracle PTree insert (oracle/oracle.c, sequential ParallelParticleTree, "
                         f"single thread), N={N}, D=4, sorted uniform ancestors: {n} inserts "
                         f"in {secs:.1f}s"}
        log(f"cpu baseline {N * n / secs:.3e} particle-inserts/s ({n} inserts)")
    out = {
        "metric": "particle-inserts/s (device path storage after each filter step)",
        "value": value, "unit": "particle-inserts/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": float(np.mean(full)) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded LG simulation, Philox noise on device)",
        "config": {"workload": f"bootstrap filter D=4 Dy=2 LG, N={N}, T={T}, systematic at "
                               f"threshold 1.0, ancestry tree insert after every step; one "
                               f"step = T filter steps with inserts",
                   "N": N, "T": T},
        "us_per_filter_step": us_step, "us_per_insert": us_insert,
        "pool_nodes": M, "nodes_used": info["used"],
        "dense_history": {"us_per_push": us_push, "bytes_per_push": bytes_push,
                          "GB_per_s": bytes_push / (us_push * 1e-6) / 1e9,
                          "note": "host-timed incl. launch and a stream sync"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "ssm_bf_tree_insert (4 launches: link+prune, free counts, scan, assign; host-timed incl. launch)",
                     "bytes_per_insert": bytes_insert},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_csmc(args, budget_s=10.0):
    """Particle Gibbs sweeps: conditional SMC with ancestor sampling (CSMC-AS) on the
    headline D=4 LG model, N particles, K steps, systematic resampling at every step, the
    device particle history recorded by the steps themselves and one trajectory drawn per
    sweep.  A sweep is one graph-replayed run of K steps plus the draw and the path
    read-back.  One bench step = one sweep conditioned on the previous sweep's trajectory."""
    import ssmhip
    ctx = ssmhip.default_context()
    name, cus = ctx.device_info()
    log(f"device {name} CUs={cus}")
    p = make_model()
    N, K = args.N, args.K
    ys = simulate(p, K, 1234)
    prm = [p[k] for k in ("mu0", "Sigma0", "A", "b", "Q", "H", "c", "R")]
    model = ssmhip.create_homogeneous_linear_gaussian_model(*prm)
    csmc = ssmhip.ConditionalSMC(ssmhip.BF(N, threshold=1.0, precision="f32"),
                                 ssmhip.AncestorSampling())
    ws = ssmhip.CSMCWorkspace(model, csmc.pf, K, ctx)
    rng = np.random.default_rng(77)
    traj, _ = ssmhip.csmc_sample(rng, model, csmc, ys, None, workspace=ws)
    for _ in range(args.warmup):
        traj, _ = ssmhip.csmc_sample(rng, model, csmc, ys, traj, workspace=ws)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        traj, ll = ssmhip.csmc_sample(rng, model, csmc, ys, traj, workspace=ws)
    ctx.synchronize()
    el = time.perf_counter() - t0
    sweeps = args.steps / el
    # algorithmic bytes per particle and step (FP32): the step kernel 40 (state read and
    # written 32, log-weight 4, ancestor 4), the ancestor-sampling weights 24 (state 16,
    # log-weight 4, weight 4) and their tile partials 4, the resampler 8 (log-weight read,
    # ancestor written), the history push 52
    bpp = 40 + 24 + 4 + 8 + 52
    achieved = bpp * N * K * sweeps / 1e9
    log(f"{args.steps} sweeps in {el:.3f}s: {sweeps:.2f} sweeps/s, "
        f"{sweeps * N * K:.3e} particle-steps/s, {achieved:.0f} GB/s algorithmic")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        om = oracle.LG(*prm)
        orng = np.random.default_rng(5)
        Nc = min(N, 16384)
        xr = traj.as_array()
        n, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < budget_s:
            xr, _ = oracle.csmc_sweep(om, ys, Nc, xr, True, 1.0, oracle.SYSTEMATIC, orng)
            n += 1
        secs = time.perf_counter() - t1
        cpu = {"value": n * Nc * K / secs, "unit": "particle-steps/s", "cores": 1,
               "kind": "port",
               "sample": f"oracle csmc_sweep (oracle/oracle.c steps, single thread), CSMC-AS, "
                         f"N={Nc}, K={K}: {n} sweeps in {secs:.1f}s"}
        log(f"cpu baseline {n * Nc * K / secs:.3e} particle-steps/s ({n} sweeps at N={Nc})")
    out = {
        "metric": "CSMC-AS sweeps/s (particle Gibbs trajectory updates)",
        "value": sweeps, "unit": "sweeps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded LG simulation, Philox noise on device)",
        "config": {"workload": f"conditional SMC with ancestor sampling, D=4 Dy=2 LG, N={N}, "
                               f"K={K}, systematic every step, dense device history; one "
                               f"step = one sweep", "N": N, "K": K},
        "particle_steps_per_s": sweeps * N * K,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "one sweep (host-timed: one graph-replayed run of K steps, the "
                               "trajectory draw and its read-back)",
                     "bytes_per_particle_step": bpp},
        "cpu_baseline": cpu,
    }
    ws.close()
    print(json.dumps(out), flush=True)


def cpu_baseline_guided(p, ys, N, kind, threshold, which, budget_s=10.0):
    """The oracle's guided step (FP64 numpy) on a bounded sample: up to 2^17 particles,
    steps until the budget is spent.  Returns (particle-steps/s, steps, seconds, N)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    m = oracle.LG(*(p[k] for k in ("mu0", "Sigma0", "A", "b", "Q", "H", "c", "R")))
    prop = (oracle.optimal_proposal(m) if which == "optimal"
            else oracle.overdispersed_proposal(m, 1.5))
    n = min(N, 1 << 17)
    rng = np.random.default_rng(5)
    st = oracle.BFState(oracle.bf_initialise(m, rng.standard_normal((n, 4))))
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < len(ys):
        oracle.guided_step(m, st, ys[steps], threshold, kind, rng.random(n if kind >= 2 else 1),
                           rng.standard_normal((n, 4)), prop)
        steps += 1
    secs = time.perf_counter() - t0
    return steps * n / secs, steps, secs, n


def bench_kfll(args, T=100, budget_s=10.0):
    """kf_loglikelihood as a batched likelihood service: B filters of one shape with
    per-step parameters (As[t], bs[t], ... of logdensity.jl:149-161), seeded stable
    dynamics.  One bench step = one call evaluating all B likelihoods (with --grad, also
    their gradients w.r.t. every parameter, the reverse rule of kalman_rrule.jl:22-107).
    The default shape, Dx = Dy = 2, is the reference's own kf_loglikelihood test model size."""
    import ssmhip
    ctx = ssmhip.default_context()
    name, cus = ctx.device_info()
    B, Dx, Dy = args.B, args.Dx, args.Dy
    log(f"device {name} CUs={cus}; B={B} Dx={Dx} Dy={Dy} T={T}")
    rng = np.random.default_rng(31)
    qr = np.linalg.qr(rng.standard_normal((B, T, Dx, Dx)))[0]
    A = 0.9 * qr
    b = 0.1 * rng.standard_normal((B, T, Dx))
    G = rng.standard_normal((B, T, Dx, Dx)) / np.sqrt(Dx)
    Q = 0.2 * (G @ np.swapaxes(G, -1, -2)) + 0.05 * np.eye(Dx)
    H = rng.standard_normal((B, T, Dy, Dx))
    c = 0.1 * rng.standard_normal((B, T, Dy))
    Gr = rng.standard_normal((B, T, Dy, Dy)) / np.sqrt(Dy)
    R = 0.3 * (Gr @ np.swapaxes(Gr, -1, -2)) + 0.1 * np.eye(Dy)
    mu0 = rng.standard_normal((B, Dx))
    S0 = np.broadcast_to(np.eye(Dx), (B, Dx, Dx)).copy()
    ys = rng.standard_normal((B, T, Dy))
    args_ = (mu0, S0, A, b, Q, H, c, R, ys)
    grad, srkf = args.grad, args.srkf
    if (grad or srkf) and (Dx > 4 or Dy > 4):
        raise SystemExit("--grad and --srkf cover filters up to Dx = Dy = 4")

    def call():
        if grad:
            return ssmhip.kf_loglikelihood_grad_batch(*args_, ctx=ctx)
        if srkf:  # filter(model, SRKF(), ys) per filter: the square-root KF kernel
            return ssmhip.srkf_filter_batch(*args_, ctx=ctx)[0], None
        return ssmhip.kf_loglikelihood_batch(*args_, ctx=ctx), None

    for _ in range(args.warmup):
        ll, g = call()
    t0 = time.perf_counter()
    dev = []
    for _ in range(args.steps):
        ll, g = call()
        ms = ssmhip._lib.C.c_double()
        ctx.check(ctx.lib.ssm_kf_loglikelihood_last_timing(ctx.ptr, ssmhip._lib.C.byref(ms)))
        dev.append(ms.value)
    el = time.perf_counter() - t0
    ms_dev = float(np.mean(dev))
    # the same call on inputs kept in device memory (ssm_kf_loglikelihood_device): the
    # wall time a sampler that uploads its parameters once would see
    wall_dev = None
    if not grad and not srkf:
        dres = ssmhip.DeviceKFInputs(*args_, ctx=ctx)
        dres.loglikelihood()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            dres.loglikelihood()
        wall_dev = (time.perf_counter() - t1) / args.steps * 1e3
        del dres
    # algorithmic HBM bytes per filter and step: its parameter block (FP64) and observation;
    # with --grad also the gradient block written (the parameter block without the
    # observation) -- the forward cache the reverse pass re-reads is not counted
    psz = (2 * Dx * Dx + Dx + Dy * Dx + Dy + Dy * Dy + Dy) * 8
    if grad:
        psz += psz - 8 * Dy
    bytes_run = float(B) * T * psz
    achieved = bytes_run / (ms_dev * 1e-3) / 1e9
    flops = float(kf_flops(Dx, Dy)) * T * B
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    errs = []
    for f in range(0, B, max(1, B // 16)):
        ref = (oracle.srkf_filter if srkf else oracle.kf_loglikelihood)(
            mu0[f], S0[f], A[f], b[f], Q[f], H[f], c[f], R[f], ys[f])
        ref = ref[0] if srkf else ref
        errs.append(abs(ll[f] - ref) / abs(ref))
    gerr = None
    if grad:  # the gradients of a few filters against the oracle's reverse rule
        gerr = 0.0
        for f in range(0, B, max(1, B // 4)):
            _, go = oracle.kf_loglikelihood_grad(mu0[f], S0[f], A[f], b[f], Q[f], H[f], c[f],
                                                 R[f], ys[f])
            for k in ("A", "b", "Q", "H", "c", "R"):
                want = np.array(go[k]).reshape(g[k][f].shape)
                gerr = max(gerr, float(np.max(np.abs(g[k][f] - want)) /
                                       max(np.max(np.abs(want)), 1e-300)))
    log(f"{args.steps} calls in {el:.3f}s; device {ms_dev:.3f} ms per call "
        f"({B / (ms_dev * 1e-3):.3e} evals/s, {achieved:.0f} GB/s, "
        f"{flops / (ms_dev * 1e-3) / 1e12:.2f} TFLOP/s); wall incl. host packing and copies "
        f"{el / args.steps * 1e3:.1f} ms; wall with device-resident inputs "
        f"{wall_dev if wall_dev is not None else float('nan'):.2f} ms; max rel err {max(errs):.2e}"
        + (f", gradient max rel err {gerr:.2e}" if grad else ""))
    cpu = None
    if not args.no_cpu:
        n, t1 = 0, time.perf_counter()
        fn = (oracle.kf_loglikelihood_grad if grad else oracle.srkf_filter if srkf
              else oracle.kf_loglikelihood)
        while time.perf_counter() - t1 < budget_s and n < B:
            fn(mu0[n], S0[n], A[n], b[n], Q[n], H[n], c[n], R[n], ys[n])
            n += 1
        secs = time.perf_counter() - t1
        what = ("oracle kf_loglikelihood_grad (FP64 numpy forward and reverse pass)" if grad
                else "oracle srkf_filter (FP64 numpy QR per step)" if srkf
                else "oracle kf_loglikelihood (oracle/oracle.c predict/update per step)")
        cpu = {"value": n / secs, "unit": "likelihood evals/s", "cores": 1, "kind": "port",
               "sample": f"{what}, single thread, on the first {n} filters of the batch, "
                         f"{secs:.1f}s"}
        log(f"cpu baseline {n / secs:.1f} evals/s")
    out = {
        "metric": ("kf_loglikelihood gradient evals/s (batched, per-step parameters)" if grad
                   else "square-root KF filter evals/s (batched, per-step parameters)" if srkf
                   else "kf_loglikelihood evals/s (batched, per-step parameters)"),
        "value": B / (ms_dev * 1e-3), "unit": "likelihood evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_dev,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if Dx <= 4 and Dy <= 4 else "f32 (FP64 inverse of S)",
        "data": "synthetic (seeded stable time-varying LG parameters and observations)",
        "config": {"workload": ("filter(model, SRKF(), ys)" if srkf else
                                f"kf_loglikelihood{' and gradient' if grad else ''}") + f", B={B}, "
                               f"Dx={Dx}, Dy={Dy}, T={T}, per-step parameters, inputs resident "
                               f"on the device", "B": B, "T": T},
        "wall_ms_per_call_incl_pcie": el / args.steps * 1e3,
        "wall_ms_per_call_device_inputs": wall_dev,
        "loglik_max_rel_err": float(max(errs)),
        "grad_max_rel_err": gerr,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": ("k_kf_tiny_grad (one filter per lane, FP64, forward cache of "
                                "80 doubles per step)" if grad else
                                "k_srkf_tiny (one filter per lane, FP64, Givens QR)" if srkf else
                                "k_kf_tiny (one filter per lane, FP64)" if Dx <= 4 and Dy <= 4
                                else "k_kf_batch<..., true>"),
                     "bytes_per_call": bytes_run},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def report_apf(args, value, ms_per_step, lls, ll_kf):
    """The auxiliary particle filter's line: whole-filter throughput and the whole step's
    HBM fraction at 72 B per particle-step when every step resamples (look-ahead: state 16 +
    log-weight 4 read, look-ahead weight 4 written; resampler: 4 read, ancestor 4 written;
    step kernel 40)."""
    step_s = ms_per_step * 1e-3 / args.T
    bpps = 72.0 if args.threshold >= 1.0 else None
    whole = None
    if bpps is not None:
        wa = bpps * args.N / step_s / 1e9
        whole = {"bytes_per_particle_step": bpps, "achieved": wa, "unit": "GB/s",
                 "frac": wa / HBM_PEAK_GBS}
    out = {
        "metric": "auxiliary-filter particle-steps/s, % HBM roofline", "value": value,
        "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded LG simulation, Philox noise on device)",
        "config": {"workload": f"auxiliary particle filter (MeanPredictive), D=4 Dy=2 LG SSM, "
                               f"N={args.N}, T={args.T}, {args.resampler} resampling, "
                               f"threshold {args.threshold}, FP32 SoA; one step = one filter() run",
                   "N": args.N, "T": args.T},
        "loglik": lls[-1], "loglik_kalman": ll_kf,
        "loglik_rel_err": abs(lls[-1] - ll_kf) / abs(ll_kf),
        "roofline": {"bound": "hbm", "achieved": whole["achieved"] if whole else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": whole["frac"] if whole else None, "traffic": None,
                     "kernel": "whole step (k_apf_weights, fold, k_scan_offspring, k_bf_step, "
                               "finish)", "step_total_us": step_s * 1e6},
        "cpu_baseline": None,
    }
    log(f"APF: {value:.4e} particle-steps/s, {step_s * 1e6:.2f} us per step, "
        f"ll={lls[-1]:.4f} kf={ll_kf:.4f}")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--N", type=int, default=1 << 20)
    ap.add_argument("--T", type=int, default=1000)
    ap.add_argument("--resampler", default="systematic", choices=sorted(RESAMPLERS))
    ap.add_argument("--threshold", type=float, default=1.0)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--reps", type=int, default=200,
                    help="back-to-back launches per kernel for the per-launch timing")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (counter calibration)")
    ap.add_argument("--workload", default="bf",
                    choices=["bf", "kf", "rbpf", "serve", "tree", "csmc", "kfll"],
                    help="bf: bootstrap filter (default); kf: batched Kalman filters; "
                         "rbpf: Rao-Blackwellised PF; serve: request stream through the "
                         "server against the plain per-filter path; tree: ancestry path "
                         "storage after every filter step; csmc: CSMC-AS sweeps; kfll: "
                         "batched kf_loglikelihood with per-step parameters")
    ap.add_argument("--requests", type=int, default=256, help="serve: requests per phase")
    ap.add_argument("--serve-saturation-only", action="store_true",
                    help="serve: only two saturation runs of the server (no plain path, no "
                         "arrival phases), for kernel profiles of the batched lane")
    ap.add_argument("--B", type=int, default=None,
                    help="kf, kfll: number of filters (default 65536)")
    ap.add_argument("--K", type=int, default=100, help="csmc: steps per sweep")
    ap.add_argument("--Dx", type=int, default=2, help="kfll: state dimension")
    ap.add_argument("--Dy", type=int, default=2, help="kfll: observation dimension")
    ap.add_argument("--proposal", default="none", choices=["none", "optimal", "overdispersed"],
                    help="bf: a guided filter with GFTest's OptimalProposal or "
                         "OverdispersedProposal(k = 1.5) instead of the bootstrap filter")
    ap.add_argument("--apf", action="store_true",
                    help="bf: the auxiliary particle filter (MeanPredictive look-ahead)")
    ap.add_argument("--srkf", action="store_true",
                    help="kfll: the square-root Kalman filter (final means and factors)")
    ap.add_argument("--grad", action="store_true",
                    help="kfll: the log-likelihoods and their parameter gradients")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the filtering "
              "path runs on a single MI355X)",
              file=sys.stderr)
        sys.exit(2)
    if args.B is None:
        args.B = 65536
    if args.workload == "kf":
        return bench_kf(args)
    if args.workload == "rbpf":
        return bench_rbpf(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "tree":
        return bench_tree(args)
    if args.workload == "csmc":
        return bench_csmc(args)
    if args.workload == "kfll":
        return bench_kfll(args)

    import ssmhip
    ctx = ssmhip.default_context()
    name, cus = ctx.device_info()
    log(f"device {name} CUs={cus}")
    p = make_model()
    ys = simulate(p, args.T, 1234)
    ll_kf = kalman_loglik(p, ys)
    model = ssmhip.create_homogeneous_linear_gaussian_model(*(p[k] for k in (
        "mu0", "Sigma0", "A", "b", "Q", "H", "c", "R")))
    kind = RESAMPLERS[args.resampler]
    R = {0: ssmhip.Systematic(), 1: ssmhip.Systematic(form="gpu"), 2: ssmhip.Stratified(),
         3: ssmhip.Multinomial()}[kind]
    guided = args.proposal != "none"
    if guided:  # ParticleFilter(N, proposal) with GFTest's proposals, fused into the step
        prop = (ssmhip.OptimalProposal(model.dyn, model.obs) if args.proposal == "optimal"
                else ssmhip.OverdispersedProposal(model.dyn, 1.5))
        algo = ssmhip.PF(args.N, prop, threshold=args.threshold, resampler=R, precision="f32")
    else:
        algo = ssmhip.BF(args.N, threshold=args.threshold, resampler=R, precision="f32")
    if args.apf:  # AuxiliaryParticleFilter(BF, MeanPredictive())
        algo = ssmhip.APF(algo, ssmhip.MeanPredictive())
    rng = np.random.default_rng(2024)
    st = ssmhip.initialise(rng, model, algo, ctx=ctx, seed=2024)
    lib = ctx.lib

    def one_filter():
        ctx.check(lib.ssm_bf_initialise(st.ptr, None))
        return ssmhip.run(st, algo, ys)

    for _ in range(args.warmup):
        one_filter()
    ctx.synchronize()
    lls = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lls.append(one_filter())
    ctx.synchronize()
    el = time.perf_counter() - t0
    value = args.steps * args.T * args.N / el
    ms_per_step = el / args.steps * 1e3
    log(f"{args.steps} filters in {el:.3f}s: {value:.4e} particle-steps/s; "
        f"ll={lls[-1]:.4f} kf={ll_kf:.4f}")

    # Per-kernel timing with HIP events (plain launches, same kernels and stream).
    import ctypes as C
    ctx.check(lib.ssm_bf_initialise(st.ptr, None))
    if args.apf:
        return report_apf(args, value, ms_per_step, lls, ll_kf)
    ms = (C.c_double * 6)()
    llp = C.c_double()
    TP = min(args.profile_steps, args.T)
    yp = np.ascontiguousarray(ys[:TP])
    ctx.check(lib.ssm_bf_profile(st.ptr, 1, TP, yp.ctypes.data_as(C.POINTER(C.c_double)),
                                 args.threshold, kind, args.reps, ms, C.byref(llp)))
    # per-launch durations from back-to-back launches (what a profiler reports per
    # dispatch); the in-sequence figures also carry the gap an event pair adds
    ms_rs, ms_st, ms_fin = ms[3], ms[4], ms[5]
    log(f"in-sequence (event pair per kernel): resample {ms[0] * 1e3:.2f} us, step "
        f"{ms[1] * 1e3:.2f} us, finish {ms[2] * 1e3:.2f} us")
    # Algorithmic bytes per particle and launch: step kernel 40 B
    # (ancestor 4 + gathered state 16 + new state 16 + log-weight 4 when resampling;
    # state 16 + log-weight 4 read and 16 + 4 written otherwise); resampler 8 B
    # (log-weight read 4, ancestor write 4).
    step_bytes = 40.0 * args.N
    resample_bytes = 8.0 * args.N
    nb = (args.N + 1023) // 1024
    finish_bytes = 64.0 * nb  # per tile: the 5 FP64 partials read, 3 FP64 outputs written
    achieved = step_bytes / (ms_st * 1e-3) / 1e9
    log(f"step kernel {ms_st * 1e3:.2f} us ({achieved:.0f} GB/s); resampler "
        f"{ms_rs * 1e3:.2f} us ({resample_bytes / (ms_rs * 1e-3) / 1e9:.0f} GB/s); "
        f"finish {ms_fin * 1e3:.2f} us")
    # the dominant kernel by measured share of the step (per-launch durations); the roofline
    # below is that kernel's, with the whole step's and the resampler's beside it
    kern = {"k_bf_step": (step_bytes, ms_st), "k_scan_offspring": (resample_bytes, ms_rs),
            "k_pf_finish": (finish_bytes, ms_fin)}
    tot_ms = ms_st + ms_rs + ms_fin
    shares = {k: v[1] / tot_ms for k, v in kern.items()}
    dominant = max(shares, key=shares.get)
    dom_bytes, dom_ms = kern[dominant]
    dom_achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    step_total_s = ms_per_step * 1e-3 / args.T
    # whole step: 40 B per particle, plus 28 B on steps that resample (every step at a
    # threshold >= 1); at lower thresholds the resampled fraction is not reported here
    whole_bpps = 68.0 if args.threshold >= 1.0 else None
    whole = None
    if whole_bpps is not None:
        wa = whole_bpps * args.N / step_total_s / 1e9
        whole = {"bytes_per_particle_step": whole_bpps, "achieved": wa, "unit": "GB/s",
                 "frac": wa / HBM_PEAK_GBS, "us_per_step": step_total_s * 1e6}
        log(f"whole step {step_total_s * 1e6:.2f} us: {wa:.0f} GB/s = "
            f"{100 * wa / HBM_PEAK_GBS:.1f}% of HBM peak (68 B per particle-step)")

    cpu = None
    if not args.no_cpu and guided:
        v, steps, secs, n_s = cpu_baseline_guided(p, ys, args.N, kind, args.threshold,
                                                  args.proposal)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"FP64 oracle guided_step (numpy, single thread) on the same model "
                         f"and proposal with N={n_s}: {steps} steps in {secs:.1f}s"}
        log(f"cpu baseline {v:.3e} particle-steps/s ({steps} steps)")
    elif not args.no_cpu:
        v1, steps1, secs1 = cpu_baseline(p, ys, args.N, kind, args.threshold, budget_s=8.0)
        log(f"cpu baseline, faithful (1 thread): {v1:.3e} particle-steps/s ({steps1} steps)")
        nproc, model_name = cpu_info()
        faithful = {"value": v1, "cores": 1,
                    "sample": f"FP64 oracle (oracle/oracle.c, single thread, AoS, the "
                              f"reference's loop order) with N={args.N}: {steps1} steps in "
                              f"{secs1:.1f}s incl. noise draws"}
        if kind == 0:  # the all-cores port implements the systematic resampler
            v, thr, steps, secs = cpu_baseline_all_cores(p, ys, args.N, args.threshold)
            cpu = {"value": v, "unit": "particle-steps/s", "cores": thr, "kind": "port",
                   "sample": f"all-cores FP64 port (oracle/bf_omp.c, OpenMP over {thr} threads, "
                             f"SoA, Philox noise in the loop) with N={args.N}: {steps} steps in "
                             f"{secs:.1f}s", "nproc": nproc, "cpu_model": model_name,
                   "faithful_single_thread": faithful}
            log(f"cpu baseline, all cores ({thr} threads of {nproc}, {model_name}): "
                f"{v:.3e} particle-steps/s ({steps} steps)")
        else:
            cpu = dict(faithful, unit="particle-steps/s", kind="port", nproc=nproc,
                       cpu_model=model_name)

    ll_err = abs(lls[-1] - ll_kf) / abs(ll_kf)
    out = {
        "metric": ("guided-filter particle-steps/s, % HBM roofline" if guided else
                   "particle-steps/s, % HBM roofline; filters/s + p50/p99 at stated loglik err"),
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded LG simulation, Philox noise on device)",
        "config": {
            "workload": (f"guided filter ({args.proposal} proposal), " if guided else
                         "bootstrap filter, ") + f"D=4 Dy=2 LG SSM, N={args.N}, T={args.T}, "
                        f"{args.resampler} resampling, threshold {args.threshold}, FP32 SoA; "
                        "one step = one full filter() run",
            "N": args.N, "T": args.T, "Dx": 4, "Dy": 2, "resampler": args.resampler,
            "threshold": args.threshold, "parallelism": "single GPU",
        },
        "filters_per_s": args.steps / el,
        "loglik": lls[-1],
        "loglik_kalman": ll_kf,
        "loglik_rel_err": ll_err,
        "roofline": {"bound": "hbm", "achieved": dom_achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom_achieved / HBM_PEAK_GBS,
                     "traffic": (None if guided or dominant != "k_bf_step" else
                                 traffic_from_profile(args.N, args.resampler, args.threshold)),
                     "kernel": (("k_bf_step<float,4,2,false,false,true> (guided)" if guided
                                 else "k_bf_step<float,4,2>") if dominant == "k_bf_step"
                                else dominant),
                     "dominant_by_share": dominant,
                     "kernel_share": shares,
                     "bytes_per_launch": dom_bytes, "us_per_launch": dom_ms * 1e3,
                     "step_kernel": {"bytes_per_launch": step_bytes, "us_per_launch": ms_st * 1e3,
                                     "achieved": achieved, "frac": achieved / HBM_PEAK_GBS,
                                     "us_per_launch_in_sequence": ms[1] * 1e3},
                     "resampler": {"bytes_per_launch": resample_bytes,
                                   "us_per_launch": ms_rs * 1e3,
                                   "achieved": resample_bytes / (ms_rs * 1e-3) / 1e9,
                                   "frac": resample_bytes / (ms_rs * 1e-3) / 1e9 / HBM_PEAK_GBS},
                     "finish_us_per_launch": ms_fin * 1e3,
                     "whole_step": whole,
                     "step_total_us": step_total_s * 1e6},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
