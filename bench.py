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
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the filtering "
              "path runs on a single MI355X; see DESIGN.md for how it would shard)",
              file=sys.stderr)
        sys.exit(2)

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
    algo = ssmhip.BF(args.N, threshold=args.threshold, resampler=R, precision="f32")
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
    achieved = step_bytes / (ms_st * 1e-3) / 1e9
    log(f"step kernel {ms_st * 1e3:.2f} us ({achieved:.0f} GB/s); resampler "
        f"{ms_rs * 1e3:.2f} us ({resample_bytes / (ms_rs * 1e-3) / 1e9:.0f} GB/s); "
        f"finish {ms_fin * 1e3:.2f} us")

    cpu = None
    if not args.no_cpu:
        v, steps, secs = cpu_baseline(p, ys, args.N, kind, args.threshold)
        cpu = {"value": v, "unit": "particle-steps/s", "cores": 1, "kind": "port",
               "sample": f"FP64 oracle (oracle/oracle.c, single thread, AoS) on the same model "
                         f"with N={args.N}: {steps} steps in {secs:.1f}s incl. noise draws"}
        log(f"cpu baseline {v:.3e} particle-steps/s ({steps} steps)")

    ll_err = abs(lls[-1] - ll_kf) / abs(ll_kf)
    out = {
        "metric": "particle-steps/s, % HBM roofline; filters/s + p50/p99 at stated loglik err",
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
            "workload": f"bootstrap filter, D=4 Dy=2 LG SSM, N={args.N}, T={args.T}, "
                        f"{args.resampler} resampling, threshold {args.threshold}, FP32 SoA; "
                        "one step = one full filter() run",
            "N": args.N, "T": args.T, "Dx": 4, "Dy": 2, "resampler": args.resampler,
            "threshold": args.threshold, "parallelism": "single GPU",
        },
        "filters_per_s": args.steps / el,
        "loglik": lls[-1],
        "loglik_kalman": ll_kf,
        "loglik_rel_err": ll_err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(args.N, args.resampler, args.threshold),
                     "kernel": "k_bf_step<float,4,2>", "bytes_per_launch": step_bytes,
                     "us_per_launch": ms_st * 1e3,
                     "us_per_launch_in_sequence": ms[1] * 1e3,
                     "resample_us_per_launch": ms_rs * 1e3, "finish_us_per_launch": ms_fin * 1e3,
                     "step_total_us": ms_per_step * 1e3 / args.T},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
