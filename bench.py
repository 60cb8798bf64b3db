"""Throughput of the MI355X 3-D Laplace FMM on one GPU.

One step = one full evaluation (tree build, traversal, P2M/M2M, M2L, P2P,
L2L/L2P) of N = 2^20 bodies uniform in [0,1)^3 with q ~ U[-1,1), P = 10,
theta = 0.4, ncrit = 64, starting from bodies resident in HBM and ending with
p and F resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--workload plain]

Other workloads (one JSON line each):
  --workload learned    N = 2^22: 90% Plummer (a = 0.05) + 10% uniform,
                        q ~ U[-1,1); the learned per-pair M2L order (MLP
                        trained on the device from a separate 2^20-body set
                        of the same distribution) against the fixed order
                        12 it falls back to, both timed here
  --workload batched    256 independent problems (N log-uniform in 1e3..1e5,
                        uniform / Plummer, P in {6, 8, 12}) re-evaluated by
                        one hipGraph launch per step
  --workload periodic   N = 2^20 in the unit periodic cube, q = +1/-1
                        alternating, P = 10, images = 3, dipole correction;
                        checked against the FP64 Ewald sum (alpha = 10,
                        ksize = 11, cutoff 0.5) on sampled targets
  --workload copy       a 256 MiB device copy: the reference kernel that
                        scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters
                        to bytes

The roofline of the dominant kernel divides its algorithmic work by its
average duration over the timed steps, read from the HIP events the library
records around the launch on its own stream.  HBM traffic per launch, when
profiles/traffic.json holds it for the workload, comes from rocprofv3 PMC
runs scaled by the copy calibration.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D Laplace FMM particles/s, N=2^20, P=10, L2 err≤1e-5; learned vs fixed order"
PEAK = {"fp32": 157.3, "bf16": 2500.0, "hbm": 8000.0}
P2P_FLOP = 20  # the conventional count per pair interaction (potential + force, rsqrt as one)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_run(n: int, P: int, ncrit: int, theta: float, threads: int, reps: int, ntargets: int) -> dict | None:
    """One child process: `reps` evaluations of the same N-body set on
    `threads` OpenMP threads by the reference's own 3d/ code (oracle/_ref,
    built from the reference sources; the FP64 oracle port when it is
    absent).  Returns the per-run stage times and the sampled targets of the
    last run (tree order) for the error check."""
    code = f"""
import json, sys, time
sys.path.insert(0, {ROOT!r})
import numpy as np, oracle
from minimal_amd.inputs import uniform_cube, sample_targets
b0 = uniform_cube({n}, 0, "sym")
runs = []
for _ in range({reps}):
    b = b0.copy()
    if oracle.reference_available():
        r = oracle.Reference(); r.set({P}, {ncrit}, {theta}); t = r.fmm(b); kind = "reference"
    else:
        o = oracle.Oracle(); t = o.fmm(b, {P}, {ncrit}, {theta}, full=False); kind = "port"
    runs.append([float(x) for x in t])
idx = sample_targets({n}, {ntargets})
tg = b[idx]
print(json.dumps({{"kind": kind, "runs": runs, "X": tg["X"].tolist(), "q": tg["q"].tolist(),
                  "p": tg["p"].tolist(), "F": tg["F"].tolist()}}))
"""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    try:
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=900)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0:
        log("cpu baseline failed:", r.stderr[-2000:])
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(n: int, P: int, ncrit: int, theta: float, threads: int, direct_error=None) -> dict | None:
    """The reference's CPU evaluation timed on this host, as planned for the
    baseline: the reference's 3d/ lazy path as written (FP64, M2L truncated
    at degree j+n <= P-1), the median of 3 evaluations of the full workload
    on `threads` threads (the host's CPU share), and one evaluation on one
    thread of an N/8 sample (a bounded single-core figure).  The time is the
    same interval as the device's: tree + upward + lists + M2L/P2P +
    downward.  direct_error(X, q, p, F) -> (err_p, err_F) checks the
    sampled targets against the FP64 direct sum."""
    d = _cpu_run(n, P, ncrit, theta, threads, 3, 1000)
    if d is None:
        return None
    evals = sorted(sum(t) for t in d["runs"])
    med = evals[len(evals) // 2]
    stage = d["runs"][[sum(t) for t in d["runs"]].index(med)]
    n1 = n // 8
    d1 = _cpu_run(n1, P, ncrit, theta, 1, 1, 10)
    one = None
    if d1 is not None:
        t1 = sum(d1["runs"][0])
        one = {"value": n1 / t1, "unit": "particles/s", "cores": 1,
               "sample": f"N={n1} (N/8) of the same distribution, one evaluation on one thread: {t1:.3f} s"}
    err = None
    if direct_error is not None:
        ep, eF = direct_error(np.array(d["X"]), np.array(d["q"]), np.array(d["p"]), np.array(d["F"]))
        err = {"p": ep, "F": eF, "targets": len(d["q"]), "vs": "FP64 direct sum"}
    what = ("reference 3d/ lazy traversal, FP64, M2L truncated at j+n <= P-1 as written"
            if d["kind"] == "reference" else "FP64 oracle port, M2L truncated at j+n <= P-1 as the reference")
    return {"value": n / med, "unit": "particles/s", "cores": threads, "kind": d["kind"],
            "sample": f"N={n} uniform [0,1)^3, q~U[-1,1), P={P}, theta={theta}, ncrit={ncrit}: median of 3 "
                      f"evaluations (tree+upward+lists+M2L/P2P+downward) on {threads} threads, {med:.3f} s ({what})",
            "runs_s": evals, "stage_s": stage, "m2l_truncation": "reference (as written)", "error": err,
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "one_thread": one}


def upload(dev, host, torch) -> None:
    """Copy a host body array into a device buffer and wait for the copy:
    the library's streams do not wait on torch's stream, so device data must
    be complete before a library call reads it."""
    dev.copy_(torch.from_numpy(host.view(np.uint8).copy()).cuda())
    torch.cuda.synchronize()


def to_device(host, torch):
    """A new device buffer holding a host body array, complete on return."""
    d = torch.from_numpy(host.view(np.uint8).copy()).cuda()
    torch.cuda.synchronize()
    return d


def timed(f, step, steps: int, warmup: int, torch) -> tuple[float, dict]:
    """Warm up, then time `steps` calls of step() between device
    synchronisations (ms per step).  Then the same number of steps again,
    each followed by a read of the per-stage durations (HIP events recorded
    by the library on its streams, which waits for both streams): their
    means are the stage breakdown, kept out of the timed region so that its
    extra per-step wait does not count."""
    for _ in range(warmup):
        step()
    f.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    f.synchronize()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    acc = {}
    for _ in range(steps):
        step()
        st = f.stage_times()
        for k, v in st.items():
            acc[k] = acc.get(k, 0.0) + v
    return ms, {k: v / steps for k, v in acc.items()}


def accuracy(f, dev, host, n, ntargets, torch, fmm_mod):
    """One evaluation of a fresh copy, then the FP64 direct sum on the
    stride-sampled targets (3d/fmm.cxx:57-61).  Returns (err_p, err_F)."""
    from minimal_amd.inputs import BODY_DTYPE, rel_l2, sample_targets
    upload(dev, host, torch)
    ptr = dev.data_ptr()
    f.evaluate(ptr, n=n, on_device=True)
    f.synchronize()
    res = dev.cpu().numpy().view(BODY_DTYPE)
    idx = sample_targets(n, ntargets)
    tg = res[idx].copy()
    tg["p"] = 0
    tg["F"] = 0
    tdev = to_device(tg, torch)
    f.direct(tdev.data_ptr(), ptr, nt=len(tg), ns=n, on_device=True)
    ref = tdev.cpu().numpy().view(BODY_DTYPE)
    return rel_l2(res["p"][idx], ref["p"]), rel_l2(res["F"][idx], ref["F"]), len(idx)


def direct_error(f, ptr, n, X, q, p, F, torch):
    """Relative L2 of given target values against the device FP64 direct
    sum over the n bodies at ptr (targets X, q; values p, F)."""
    from minimal_amd.inputs import BODY_DTYPE, rel_l2
    tg = np.zeros(len(q), BODY_DTYPE)
    tg["X"], tg["q"] = X, q
    tdev = to_device(tg, torch)
    f.direct(tdev.data_ptr(), ptr, nt=len(tg), ns=n, on_device=True)
    ref = tdev.cpu().numpy().view(BODY_DTYPE)
    return rel_l2(p, ref["p"]), rel_l2(F, ref["F"])


def traffic_for(workload: str, kernel: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        e = t["workloads"][workload][kernel]
        return float(e["fetch_bytes"]) + float(e["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


def roofline(workload: str, kern_ms: dict, work: dict, P: int, m2l_prec: str, order_hist=None) -> dict:
    """Roofline of the longer of the two hot kernels.  P2P: 20 flop per pair
    interaction against the FP32 vector peak (it is VALU-bound: no MFMA
    form, no HBM bound at ~250 MB per launch).  M2L: the keyed GEMM's MFMA
    work, 2 (P^2)^2 flop per M2L pair (3x for the three BF16 products of the
    split-precision form) against the MFMA peak of that precision."""
    p2p = {"kernel": "p2p", "bound": "valu", "ms": kern_ms["p2p"],
           "algorithmic": f"{P2P_FLOP} flop x {work['p2p_interactions']} pair interactions per launch",
           "achieved": P2P_FLOP * work["p2p_interactions"] / (kern_ms["p2p"] * 1e-3) / 1e12,
           "peak": PEAK["fp32"], "unit": "TFLOP/s"}
    p2p["frac"] = p2p["achieved"] / p2p["peak"]
    p2p["traffic"] = traffic_for(workload, "p2p")
    flop_pair = 2.0 * P ** 4
    if order_hist is not None:  # per-pair orders of a learned evaluation: slot-weighted mean of p^4
        tot = sum(int(c) for c in order_hist)
        flop_pair = 2.0 * sum(int(c) * p ** 4 for p, c in enumerate(order_hist)) / max(tot, 1)
    mult = 3.0 if m2l_prec == "bf16x3" else 1.0
    m2l = {"kernel": "m2l_gemm", "bound": "mfma", "ms": kern_ms["m2l_kernel"],
           "algorithmic": f"{mult:g} x {flop_pair:.4g} flop x {work['m2l_pairs']} M2L pairs per launch",
           "achieved": mult * flop_pair * work["m2l_pairs"] / (kern_ms["m2l_kernel"] * 1e-3) / 1e12,
           "peak": PEAK["bf16"] if m2l_prec == "bf16x3" else PEAK["fp32"], "unit": "TFLOP/s"}
    m2l["frac"] = m2l["achieved"] / m2l["peak"]
    m2l["traffic"] = traffic_for(workload, "m2l_gemm")
    dom, other = (p2p, m2l) if p2p["ms"] >= m2l["ms"] else (m2l, p2p)
    out = dict(dom)
    out["other"] = other
    return out


def copy_workload(args, torch):
    """256 MiB device-to-device copy, three times: the known-size kernel
    whose FETCH_SIZE / WRITE_SIZE readings scale the counters to bytes."""
    x = torch.ones(64 << 20, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy", "bytes_read_per_launch": x.numel() * 4,
                      "bytes_written_per_launch": y.numel() * 4, "launches": 3}), flush=True)


def plain_learned(args, fmm, torch, dev, host, ptr, n, trunc, prec, p2p, ms_fixed):
    """Learned against fixed order on the headline bodies: an order MLP
    trained on the device (pair errors of a P = 14 context on another
    uniform 2^20 set, 16,384 pairs, 400 Adam epochs; the training time is
    reported, not timed in the steps) picks an order <= P per GEMM slot at
    log10 tolerance --plain-learned-tol; same steps, same accuracy check."""
    from minimal_amd.inputs import uniform_cube
    t0 = time.perf_counter()
    tr = fmm.FMM(P=14, ncrit=args.ncrit, theta=args.theta, m2l_precision=prec)
    tr.build_tree(uniform_cube(1 << 20, 11, "sym"))
    tr.upward()
    tr.horizontal()
    loss = tr.train_order_model(nsamples=16384, epochs=400, lr=3e-3)
    weights = tr.order_model()
    train_s = time.perf_counter() - t0
    del tr
    f = fmm.FMM(P=args.P, ncrit=args.ncrit, theta=args.theta, m2l_truncation=trunc, m2l_precision=prec,
                p2p_precision=p2p, order_policy=fmm.ORDER_LEARNED)
    f.set_order_model(weights, args.plain_learned_tol)
    upload(dev, host, torch)
    ms, kern = timed(f, lambda: f.evaluate(ptr, n=n, on_device=True), args.steps, args.warmup, torch)
    hist = f.order_stats()
    err_p, err_F, nt = accuracy(f, dev, host, n, args.targets, torch, fmm)
    log(f"learned order (tol {args.plain_learned_tol}): {ms:.3f} ms/step, err p {err_p:.3e} F {err_F:.3e}")
    return {"value": n / (ms * 1e-3), "ms_per_step": ms, "speedup_vs_fixed": ms_fixed / ms,
            "error": {"p": err_p, "F": err_F, "targets": int(nt), "vs": "FP64 direct sum"},
            "log10_tolerance": args.plain_learned_tol, "orders": f"4..{args.P}",
            "gemm_slot_orders": {str(p): int(c) for p, c in enumerate(hist) if c}, "stages_ms": kern,
            "training": {"seconds": train_s, "loss": loss, "pairs": 16384, "epochs": 400,
                         "set": "another uniform 2^20 set, P=14"}}


def plain_workload(args, fmm, torch):
    from minimal_amd.inputs import uniform_cube
    prec = {"bf16x3": fmm.PREC_BF16X3, "fp32": fmm.PREC_FP32, "fp64": fmm.PREC_FP64}[args.m2l]
    trunc = fmm.M2L_FULL if args.truncation == "full" else fmm.M2L_REFERENCE
    n = args.n or (1 << 20)
    host = uniform_cube(n, 0, "sym")
    dev = to_device(host, torch)
    ptr = dev.data_ptr()
    p2p = {"fp32": fmm.PREC_FP32, "fp32_newton": fmm.PREC_FP32_NEWTON}[args.p2p]
    f = fmm.FMM(P=args.P, ncrit=args.ncrit, theta=args.theta, m2l_truncation=trunc, m2l_precision=prec,
                p2p_precision=p2p)
    log(f"bench: N={n} P={args.P} m2l={args.m2l} truncation={args.truncation} p2p={args.p2p}")
    # Cold evaluation: a fresh context on bodies in their original (unsorted)
    # order -- device buffers grown, M2L operators generated, GEMM shape
    # autotuned -- timed once around the call.
    f.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f.evaluate(ptr, n=n, on_device=True)
    f.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    log(f"cold evaluation: {cold_ms:.2f} ms")
    ms, kern = timed(f, lambda: f.evaluate(ptr, n=n, on_device=True), args.steps, args.warmup, torch)
    # Warm context, unsorted input: each step starts from a fresh copy of the
    # bodies in their original order (the copy is outside the timing).
    uns = []
    for _ in range(max(3, min(args.steps, 10))):
        upload(dev, host, torch)
        t0 = time.perf_counter()
        f.evaluate(ptr, n=n, on_device=True)
        f.synchronize()
        uns.append((time.perf_counter() - t0) * 1e3)
    unsorted_ms = sorted(uns)[len(uns) // 2]
    log(f"unsorted input: {unsorted_ms:.3f} ms/step (median of {len(uns)})")
    work = f.work_counts()
    log("stages ms (mean of timed steps):", json.dumps(kern), "work:", json.dumps(work))
    err_p, err_F, nt = accuracy(f, dev, host, n, args.targets, torch, fmm)
    log(f"rel L2 vs FP64 direct ({nt} targets): p {err_p:.3e} F {err_F:.3e}")
    wl = "plain_n2p20" if n == 1 << 20 and args.P == 10 else "plain"
    # The timed steps run P2P on a second stream beside the upward pass and
    # the keyed GEMM, so a kernel's event span there includes time the other
    # stream holds the CUs.  The same steps with everything on one stream
    # give each kernel's own duration (the one rocprofv3 reports): the
    # roofline uses those, the overlapped spans are kept beside them.
    ro = roofline(wl, kern, work, args.P, args.m2l)
    os.environ["EXAFMM_SERIAL_STREAMS"] = "1"
    try:
        fi = fmm.FMM(P=args.P, ncrit=args.ncrit, theta=args.theta, m2l_truncation=trunc, m2l_precision=prec,
                     p2p_precision=p2p)
    finally:
        del os.environ["EXAFMM_SERIAL_STREAMS"]
    upload(dev, host, torch)
    ms_i, kern_i = timed(fi, lambda: fi.evaluate(ptr, n=n, on_device=True), args.steps, args.warmup, torch)
    roof = roofline(wl, kern_i, work, args.P, args.m2l)
    roof["kernel_ms_from"] = "one-stream pass (each kernel alone on the CUs)"
    roof["one_stream"] = {"ms_per_step": ms_i, "stages_ms": kern_i}
    roof["two_streams"] = {"ms_per_step": ms, "stages_ms": kern,
                           ro["kernel"]: {k: ro[k] for k in ("ms", "achieved", "frac")},
                           ro["other"]["kernel"]: {k: ro["other"][k] for k in ("ms", "achieved", "frac")}}
    log(f"one stream: {ms_i:.3f} ms/step, stages {json.dumps(kern_i)}")
    # The C ABI also takes bodies in host memory (copied in and out over
    # PCIe inside each call): the same step on a host copy, reported beside
    # the device-resident value, never as it.
    hb = host.copy()
    ms_h, _ = timed(f, lambda: f.evaluate(hb), 3, 1, torch)
    log(f"host buffers: {ms_h:.3f} ms/step")
    # The other P2P arithmetic on the same bodies (hardware 1/sqrt and plain
    # FP32 sums when the headline runs the refined one, or the reverse):
    # same steps, same accuracy check, reported beside the headline.
    alt_name = "fp32" if args.p2p == "fp32_newton" else "fp32_newton"
    alt_p2p = fmm.PREC_FP32 if args.p2p == "fp32_newton" else fmm.PREC_FP32_NEWTON
    fa = fmm.FMM(P=args.P, ncrit=args.ncrit, theta=args.theta, m2l_truncation=trunc, m2l_precision=prec,
                 p2p_precision=alt_p2p)
    upload(dev, host, torch)
    ms_a, kern_a = timed(fa, lambda: fa.evaluate(ptr, n=n, on_device=True), args.steps, args.warmup, torch)
    ea_p, ea_F, _ = accuracy(fa, dev, host, n, args.targets, torch, fmm)
    log(f"p2p {alt_name}: {ms_a:.3f} ms/step, p2p {kern_a['p2p']:.3f} ms, err p {ea_p:.3e} F {ea_F:.3e}")
    alt = {"p2p_precision": alt_name, "value": n / (ms_a * 1e-3), "ms_per_step": ms_a, "stages_ms": kern_a,
           "error": {"p": ea_p, "F": ea_F, "targets": int(nt), "vs": "FP64 direct sum"}}
    del fa
    lvf = None
    if args.with_learned:
        lvf = plain_learned(args, fmm, torch, dev, host, ptr, n, trunc, prec, p2p, ms)
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(n, args.P, args.ncrit, args.theta, threads,
                           lambda X, q, p, F: direct_error(f, ptr, n, X, q, p, F, torch))
    out = {
        "metric": METRIC, "value": n / (ms * 1e-3), "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"plain 3D Laplace FMM, N={n} uniform [0,1)^3, q~U[-1,1), P={args.P}, "
                               f"theta={args.theta}, ncrit={args.ncrit}",
                   "m2l_truncation": args.truncation, "m2l_precision": args.m2l, "p2p_precision": args.p2p,
                   "order_policy": "fixed"},
        "error": {"p": err_p, "F": err_F, "targets": int(nt), "vs": "FP64 direct sum"},
        "stages_ms": kern, "work": work, "roofline": roof, "cpu_baseline": cpu,
        "host_buffers": {"ms_per_step": ms_h, "value": n / (ms_h * 1e-3),
                         "note": "bodies in pageable host memory, copied in and out over PCIe inside each call"},
        "cold": {"ms": cold_ms, "note": "first evaluation of a fresh context on unsorted bodies: device buffers "
                                        "grown, M2L operators generated, GEMM shape autotuned"},
        "unsorted_input": {"ms_per_step": unsorted_ms, "value": n / (unsorted_ms * 1e-3),
                           "note": "warm context, each step on a fresh copy of the bodies in their original order "
                                   "(median, timed per step with a synchronisation)"},
        "other_p2p_arithmetic": alt,
        "learned_vs_fixed": lvf,
    }
    print(json.dumps(out), flush=True)


def learned_workload(args, fmm, torch):
    """Learned per-pair order against fixed order 12 on the same bodies.
    The model is trained on the device (exafmm_order_model_train: pair
    errors from the FP64 truncated expansions vs the exact pair sum, then
    Adam on the MLP) from a separate body set of the same distribution; the
    training time is reported, not timed in the steps."""
    from minimal_amd.inputs import plummer_background
    n = args.n or (1 << 22)
    prec = {"bf16x3": fmm.PREC_BF16X3, "fp32": fmm.PREC_FP32}[args.m2l]
    t0 = time.perf_counter()
    tr = fmm.FMM(P=14, ncrit=args.ncrit, theta=args.theta, m2l_precision=prec)
    tr.build_tree(plummer_background(1 << 20, 5))
    tr.upward()
    tr.horizontal()
    loss = tr.train_order_model(nsamples=16384, epochs=400, lr=3e-3)
    weights = tr.order_model()
    train_s = time.perf_counter() - t0
    del tr
    log(f"trained order model in {train_s:.2f} s, loss {loss:.3f}")
    host = plummer_background(n, 6)
    dev = to_device(host, torch)
    ptr = dev.data_ptr()
    res = {}
    for name in ("fixed", "learned"):
        if name == "fixed":
            f = fmm.FMM(P=args.P_fixed, ncrit=args.ncrit, theta=args.theta, m2l_precision=prec)
        else:
            f = fmm.FMM(P=args.P_learned, ncrit=args.ncrit, theta=args.theta, m2l_precision=prec,
                        order_policy=fmm.ORDER_LEARNED)
            f.set_order_model(weights, args.tol)
        upload(dev, host, torch)
        ms, kern = timed(f, lambda: f.evaluate(ptr, n=n, on_device=True), args.steps, args.warmup, torch)
        hist = f.order_stats()
        work = f.work_counts()
        err_p, err_F, nt = accuracy(f, dev, host, n, args.targets, torch, fmm)
        log(f"{name}: {ms:.3f} ms/step, err p {err_p:.3e} F {err_F:.3e}, stages {json.dumps(kern)}")
        res[name] = {"value": n / (ms * 1e-3), "ms_per_step": ms, "error": {"p": err_p, "F": err_F},
                     "stages_ms": kern, "work": work,
                     "gemm_slot_orders": {str(p): int(c) for p, c in enumerate(hist) if c},
                     "roofline": roofline("learned", kern, work, args.P_fixed, args.m2l,
                                          hist if name == "learned" else None)}
        del f
    L, F = res["learned"], res["fixed"]
    out = {
        "metric": METRIC, "value": L["value"], "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": L["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"learned per-pair M2L order, N={n}: 90% Plummer (a=0.05, r<=0.5) + 10% uniform, "
                               f"q~U[-1,1), learned orders 4..{args.P_learned} (expansions held to order "
                               f"{args.P_learned}) against fixed order {args.P_fixed}, theta={args.theta}, "
                               f"ncrit={args.ncrit}",
                   "m2l_precision": args.m2l, "p2p_precision": "fp32", "order_policy": "learned",
                   "model": "MLP 4-32-32-11, ReLU, BF16 weights, FP32 accumulate", "log10_tolerance": args.tol},
        "error": {"p": L["error"]["p"], "F": L["error"]["F"], "targets": args.targets, "vs": "FP64 direct sum"},
        "fixed_order": F, "learned": L,
        "speedup_vs_fixed": L["value"] / F["value"],
        "training": {"seconds": train_s, "loss": loss, "pairs": 16384, "epochs": 400,
                     "set": "2^20 bodies of the same distribution, another seed, P=14"},
        "roofline": L["roofline"], "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def batched_workload(args, fmm, torch):
    """256 independent problems (N log-uniform in [10^3, 10^5], even-indexed
    uniform cubes, odd-indexed Plummer spheres, q = U[0,1)/N, P drawn from {6, 8, 12},
    theta = 0.4, ncrit = 64), bodies resident in one device buffer.  One step
    is one exafmm_evaluate_batch call with EXAFMM_BATCH_FOREST: the problems of
    each order are gathered and evaluated as one forest -- tree builds,
    traversals and every pass of all problems, with one launch per pass --
    and copied back.  The first (cold) call, with context creation and the
    GEMM autotuner, is reported beside the warm steps; so is the grouped
    hipGraph that re-evaluates the charge-dependent passes of per-problem
    contexts whose trees were built beforehand (positions frozen)."""
    from minimal_amd.inputs import BODY_DTYPE, batched_problems, rel_l2, sample_targets
    probs = batched_problems(256, 3)
    ns = [len(b) for b, _ in probs]
    offs = np.concatenate([[0], np.cumsum(ns)])
    N = int(offs[-1])
    host = np.concatenate([b for b, _ in probs])
    dev = to_device(host, torch)
    base = dev.data_ptr()
    counts = {P: sum(1 for _, q in probs if q == P) for P in (6, 8, 12)}
    log(f"bench batched: {len(probs)} problems, N total {N}, P counts {counts}")
    spec = [(base + int(o) * 64, nb, P) for o, nb, (_, P) in zip(offs[:-1], ns, probs)]
    fb = fmm.FMM(ncrit=args.ncrit, theta=args.theta)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fb.evaluate_batch(spec, forest=True)
    torch.cuda.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(args.warmup):
        fb.evaluate_batch(spec, forest=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fb.evaluate_batch(spec, forest=True)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    log(f"forest batch: cold {cold_ms:.1f} ms, step {ms:.3f} ms ({N / (ms * 1e-3):.3e} particles/s)")
    # accuracy of one call on fresh bodies (p and F are accumulated, so the
    # timed calls' sums are not checked) on 24 problems (both kinds, every
    # P), FP64 direct sum on sampled targets
    upload(dev, host, torch)
    fb.evaluate_batch(spec, forest=True)
    torch.cuda.synchronize()
    res = dev.cpu().numpy().view(BODY_DTYPE)
    errs = {}
    chk = fmm.FMM(P=6)
    for i in range(0, 256, 11):  # both kinds (even uniform, odd Plummer), every order
        b = res[offs[i]:offs[i + 1]]
        idx = sample_targets(len(b), 100)
        ref = b[idx].copy()
        ref["p"] = 0
        ref["F"] = 0
        chk.direct(ref, b.copy())
        P = probs[i][1]
        e = errs.setdefault(P, [0.0, 0.0, 0])
        e[0] = max(e[0], rel_l2(b["p"][idx], ref["p"]))
        e[1] = max(e[1], rel_l2(b["F"][idx], ref["F"]))
        e[2] += 1
    log("max rel L2 per order:", errs)
    # the grouped graph over per-problem contexts (charges only; trees and
    # lists built once, outside the timed replays)
    upload(dev, host, torch)
    t0 = time.perf_counter()
    ctxs = []
    for (b, P), o in zip(probs, offs[:-1]):
        f = fmm.FMM(P=P, ncrit=args.ncrit, theta=args.theta)
        f.evaluate(base + int(o) * 64, n=len(b), on_device=True)
        ctxs.append(f)
    for f in ctxs:
        f.synchronize()
    plan_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = fmm.Graph(ctxs, streams=args.streams, grouped=True)
    create_s = time.perf_counter() - t0
    for _ in range(args.warmup):
        g.launch()
    g.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.launch()
    g.synchronize()
    torch.cuda.synchronize()
    ms_g = (time.perf_counter() - t0) * 1e3 / args.steps
    nodes = g.nodes()
    log(f"grouped graph: plan {plan_s:.2f} s, create {create_s:.2f} s ({nodes} nodes), replay {ms_g:.3f} ms")
    g.close()
    out = {
        "metric": "batched 3D Laplace FMM particles/s, 256 problems, tree and lists included",
        "value": N / (ms * 1e-3), "unit": "particles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"batched FMM: 256 problems, N log-uniform in [1e3,1e5] (total {N}), even-indexed "
                               f"uniform cubes, odd-indexed Plummer (a=0.1, r<=1), q=U[0,1)/N, P in {{6,8,12}} "
                               f"({counts}), theta={args.theta}, ncrit={args.ncrit}; one step = one "
                               f"exafmm_evaluate_batch(EXAFMM_BATCH_FOREST) call: per order one forest (tree "
                               f"build, traversal, every pass, one launch per pass over its problems)",
                   "m2l_precision": "bf16x3", "p2p_precision": "fp32"},
        "cold_ms": cold_ms,
        "error": {"max_rel_l2_by_P": {str(k): {"p": v[0], "F": v[1], "problems": v[2]} for k, v in errs.items()},
                  "targets_per_problem": 100, "vs": "FP64 direct sum"},
        "grouped_graph_charges_only": {"ms_per_step": ms_g, "value": N / (ms_g * 1e-3), "nodes": nodes,
                                       "plan_s": plan_s, "create_s": create_s},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def periodic_workload(args, fmm, torch):
    """N bodies uniform in [0,1)^3 with cycle 1, charges +1/-1 alternating
    (net neutral), P = 10, theta = 0.4, images = 3 (3dp/fmm.cxx with the
    sizes of the periodic configuration).  One step = evaluate + dipole
    correction (3dp/fmm.cxx:46-75).  Accuracy: forces and potentials of
    sampled targets against the device FP64 Ewald sum over all bodies.  The
    dipole term is applied pointwise for the potential comparison
    (phi -= 4 pi/(3 V) d.x, the form whose energy equals the reference's
    per-body correction) and the one global constant by which the two
    summation conventions differ is removed."""
    from minimal_amd.inputs import BODY_DTYPE, rel_l2, sample_targets, uniform_cube
    n = args.n or (1 << 20)
    host = uniform_cube(n, 0, "sym")
    host["q"] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    cycle = 1.0
    prec = {"bf16x3": fmm.PREC_BF16X3, "fp32": fmm.PREC_FP32, "fp64": fmm.PREC_FP64}[args.m2l]
    dev = to_device(host, torch)
    ptr = dev.data_ptr()
    f = fmm.FMM(P=args.P, ncrit=args.ncrit, theta=args.theta, m2l_truncation=fmm.M2L_FULL, m2l_precision=prec,
                p2p_precision=fmm.PREC_FP32, images=args.images, cycle=cycle)
    log(f"bench periodic: N={n} P={args.P} images={args.images} m2l={args.m2l}")

    def step():
        f.evaluate(ptr, n=n, on_device=True)
        f.dipole_correction()

    ms, kern = timed(f, step, args.steps, args.warmup, torch)
    work = f.work_counts()
    # accuracy: fresh evaluation without the per-body correction
    upload(dev, host, torch)
    f.evaluate(ptr, n=n, on_device=True)
    f.synchronize()
    res = dev.cpu().numpy().view(BODY_DTYPE)
    idx = sample_targets(n, args.targets)
    tg = res[idx].copy()
    tg["p"] = 0
    tg["F"] = 0
    tdev = to_device(tg, torch)
    ksize, alpha, sigma, cutoff = 11, 10.0, 0.25 / np.pi, 0.5
    f.ewald(tdev.data_ptr(), ptr, ksize, alpha, sigma, cutoff, nt=len(tg), ns=n, on_device=True)
    ew = tdev.cpu().numpy().view(BODY_DTYPE)
    dip = (res["X"] * res["q"][:, None]).sum(0)
    coef = 4 * np.pi / (3 * cycle ** 3)
    p = res["p"][idx] - coef * (res["X"][idx] @ dip)
    p -= np.mean(p - ew["p"])
    F = res["F"][idx] - coef * dip
    err_p = rel_l2(p, ew["p"])
    err_F = rel_l2(F, ew["F"])
    log(f"rel L2 vs FP64 Ewald ({len(idx)} targets): p {err_p:.3e} F {err_F:.3e}")
    out = {
        "metric": "periodic 3D Laplace FMM particles/s, N=2^20, P=10, images=3",
        "value": n / (ms * 1e-3), "unit": "particles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"periodic 3D Laplace FMM, N={n} uniform in the unit periodic cube, q=+1/-1 "
                               f"alternating, P={args.P}, theta={args.theta}, ncrit={args.ncrit}, "
                               f"images={args.images}, dipole correction",
                   "m2l_truncation": "full", "m2l_precision": args.m2l, "p2p_precision": "fp32"},
        "error": {"p": err_p, "F": err_F, "targets": int(len(idx)),
                  "vs": f"FP64 Ewald sum, alpha={alpha}, ksize={ksize}, cutoff={cutoff}"},
        "stages_ms": kern, "work": work, "roofline": roofline("periodic", kern, work, args.P, args.m2l),
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=0, help="bodies (default: the workload's size)")
    ap.add_argument("--P", type=int, default=10)
    ap.add_argument("--P-fixed", dest="P_fixed", type=int, default=12, help="learned workload: the fixed order compared")
    ap.add_argument("--P-learned", dest="P_learned", type=int, default=13,
                    help="learned workload: highest learned order (the learned context's P)")
    ap.add_argument("--tol", type=float, default=-7.1,
                    help="learned workload: log10 of the allowed M2L error per pair, relative to the potential "
                         "scale sqrt((sum q)^2 + sum q^2) / R0")
    ap.add_argument("--ncrit", type=int, default=64)
    ap.add_argument("--theta", type=float, default=0.4)
    ap.add_argument("--m2l", choices=["bf16x3", "fp32", "fp64"], default="bf16x3")
    ap.add_argument("--with-learned", action="store_true",
                    help="plain workload: add the learned-order comparison on the headline bodies (at P = 10 the "
                         "orders have little to cut; the learned workload is where it pays)")
    ap.add_argument("--no-learned", action="store_true", help="(accepted for older command lines; now the default)")
    ap.add_argument("--plain-learned-tol", type=float, default=-7.0,
                    help="plain workload: log10 tolerance of the learned-order comparison")
    ap.add_argument("--p2p", choices=["fp32", "fp32_newton"], default="fp32_newton",
                    help="plain workload: P2P precision (fp32_newton, the stated arithmetic of the N=2^20 "
                         "workload: Newton-refined 1/sqrt and compensated sums; fp32: hardware 1/sqrt, plain sums)")
    ap.add_argument("--truncation", choices=["full", "reference"], default="full")
    ap.add_argument("--targets", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--workload", choices=["plain", "learned", "batched", "periodic", "copy"], default="plain")
    ap.add_argument("--streams", type=int, default=8, help="batched workload: graph capture streams")
    ap.add_argument("--images", type=int, default=3)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build evaluates on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    if args.workload == "copy":
        copy_workload(args, torch)
        return
    from minimal_amd import fmm
    {"plain": plain_workload, "learned": learned_workload, "batched": batched_workload,
     "periodic": periodic_workload}[args.workload](
        args, fmm, torch)


if __name__ == "__main__":
    main()
