"""Throughput of the MI355X 3-D Laplace FMM on one GPU.

One step = one full evaluation (tree build, traversal, P2M/M2M, M2L, P2P,
L2L/L2P) of N = 2^20 bodies uniform in [0,1)^3 with q ~ U[-1,1), P = 10,
theta = 0.4, ncrit = 64, starting from bodies resident in HBM and ending with
p and F resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--workload plain]

Other workloads (one JSON line each):
  --workload periodic   N = 2^20 in the unit periodic cube, q = +1/-1
                        alternating, P = 10, images = 3, dipole correction;
                        checked against the FP64 Ewald sum (alpha = 10,
                        ksize = 11, cutoff 0.5) on sampled targets
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n: int, P: int, ncrit: int, theta: float, threads: int) -> dict | None:
    """The reference's own CPU evaluation (oracle/_ref, built from the
    reference sources) -- or the FP64 oracle port when that is absent --
    timed on this host in a child process with a fixed thread count."""
    code = f"""
import json, sys, time
sys.path.insert(0, {ROOT!r})
import numpy as np, oracle
from minimal_amd.inputs import uniform_cube
b = uniform_cube({n}, 0, "sym")
if oracle.reference_available():
    r = oracle.Reference(); r.set({P}, {ncrit}, {theta}); t = r.fmm(b); kind = "reference"
else:
    o = oracle.Oracle(); t = o.fmm(b, {P}, {ncrit}, {theta}, full=False); kind = "port"
print(json.dumps({{"kind": kind, "times": [float(x) for x in t]}}))
"""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
    try:
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=600)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0:
        log("cpu baseline failed:", r.stderr[-2000:])
        return None
    d = json.loads(r.stdout.strip().splitlines()[-1])
    t_eval = sum(d["times"])
    return {"value": n / t_eval, "unit": "particles/s", "cores": threads, "kind": d["kind"],
            "sample": f"full config: N={n} uniform [0,1)^3, q~U[-1,1), P={P}, theta={theta}, ncrit={ncrit}; "
                      f"tree+upward+lists+M2L/P2P+downward in {t_eval:.3f} s "
                      f"({'reference 3d/ lazy traversal, FP64, as written' if d['kind'] == 'reference' else 'FP64 oracle port'})",
            "stage_s": d["times"]}


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
    n = args.n
    host = uniform_cube(n, 0, "sym")
    host["q"] = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    cycle = 1.0
    prec = {"bf16x3": fmm.PREC_BF16X3, "fp32": fmm.PREC_FP32, "fp64": fmm.PREC_FP64}[args.m2l]
    dev = torch.from_numpy(host.view(np.uint8).copy()).cuda()
    ptr = dev.data_ptr()
    f = fmm.FMM(P=args.P, ncrit=args.ncrit, theta=args.theta, m2l_truncation=fmm.M2L_FULL, m2l_precision=prec,
                p2p_precision=fmm.PREC_FP32, images=args.images, cycle=cycle)
    log(f"bench periodic: N={n} P={args.P} images={args.images} m2l={args.m2l}")
    for _ in range(args.warmup):
        f.evaluate(ptr, n=n, on_device=True)
        f.dipole_correction()
    f.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f.evaluate(ptr, n=n, on_device=True)
        f.dipole_correction()
    f.synchronize()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    stages = f.stage_times()
    work = f.work_counts()
    # accuracy: fresh evaluation without the per-body correction
    dev.copy_(torch.from_numpy(host.view(np.uint8).copy()).cuda())
    f.evaluate(ptr, n=n, on_device=True)
    f.synchronize()
    res = dev.cpu().numpy().view(BODY_DTYPE)
    idx = sample_targets(n, args.targets)
    tg = res[idx].copy()
    tg["p"] = 0
    tg["F"] = 0
    tdev = torch.from_numpy(tg.view(np.uint8).copy()).cuda()
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
        "stages_ms": stages, "work": work, "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--P", type=int, default=10)
    ap.add_argument("--ncrit", type=int, default=64)
    ap.add_argument("--theta", type=float, default=0.4)
    ap.add_argument("--m2l", choices=["bf16x3", "fp32", "fp64"], default="bf16x3")
    ap.add_argument("--truncation", choices=["full", "reference"], default="full")
    ap.add_argument("--targets", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    ap.add_argument("--workload", choices=["plain", "periodic"], default="plain")
    ap.add_argument("--images", type=int, default=3)
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build evaluates on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from minimal_amd import fmm
    from minimal_amd.inputs import BODY_DTYPE, rel_l2, sample_targets, uniform_cube

    if args.workload == "periodic":
        periodic_workload(args, fmm, torch)
        return

    prec = {"bf16x3": fmm.PREC_BF16X3, "fp32": fmm.PREC_FP32, "fp64": fmm.PREC_FP64}[args.m2l]
    trunc = fmm.M2L_FULL if args.truncation == "full" else fmm.M2L_REFERENCE
    n = args.n
    host = uniform_cube(n, 0, "sym")
    dev = torch.from_numpy(host.view(np.uint8).copy()).cuda()
    ptr = dev.data_ptr()
    f = fmm.FMM(P=args.P, ncrit=args.ncrit, theta=args.theta, m2l_truncation=trunc, m2l_precision=prec,
                p2p_precision=fmm.PREC_FP32)
    log(f"bench: N={n} P={args.P} m2l={args.m2l} truncation={args.truncation}")
    for _ in range(args.warmup):
        f.evaluate(ptr, n=n, on_device=True)
    f.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f.evaluate(ptr, n=n, on_device=True)
    f.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    stages = f.stage_times()
    work = f.work_counts()
    log("stages ms:", json.dumps(stages), "work:", json.dumps(work))

    # accuracy: fresh copy, one evaluation, FP64 direct sum on sampled targets
    dev.copy_(torch.from_numpy(host.view(np.uint8).copy()).cuda())
    f.evaluate(ptr, n=n, on_device=True)
    f.synchronize()
    res = dev.cpu().numpy().view(BODY_DTYPE)
    idx = sample_targets(n, args.targets)
    tg = res[idx].copy()
    tg["p"] = 0
    tg["F"] = 0
    tdev = torch.from_numpy(tg.view(np.uint8).copy()).cuda()
    f.direct(tdev.data_ptr(), ptr, nt=len(tg), ns=n, on_device=True)
    ref = tdev.cpu().numpy().view(BODY_DTYPE)
    err_p = rel_l2(res["p"][idx], ref["p"])
    err_F = rel_l2(res["F"][idx], ref["F"])
    log(f"rel L2 vs FP64 direct ({len(idx)} targets): p {err_p:.3e} F {err_F:.3e}")

    # roofline of the dominant kernel
    p2p_flop = 20.0 * work["p2p_interactions"]
    m2l_flop = 2.0 * args.P ** 4 * work["m2l_pairs"]
    m2l_ms = stages["m2l_kernel"] if args.m2l != "fp64" else stages["m2l"]
    if stages["p2p"] >= m2l_ms:
        ach = p2p_flop / (stages["p2p"] * 1e-3) / 1e12
        roof = {"kernel": "p2p", "bound": "valu", "achieved": ach, "peak": PEAK["fp32"], "unit": "TFLOP/s",
                "frac": ach / PEAK["fp32"], "traffic": None,
                "algorithmic": f"20 flop x {work['p2p_interactions']} interactions per launch"}
    else:
        if args.m2l == "bf16x3":
            ach = 3 * m2l_flop / (m2l_ms * 1e-3) / 1e12
            pk = PEAK["bf16"]
        else:
            ach = m2l_flop / (m2l_ms * 1e-3) / 1e12
            pk = PEAK["fp32"]
        roof = {"kernel": "m2l", "bound": "mfma", "achieved": ach, "peak": pk, "unit": "TFLOP/s", "frac": ach / pk,
                "traffic": None, "algorithmic": f"2 P^4 flop x {work['m2l_pairs']} pairs per launch"}

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(n, args.P, args.ncrit, args.theta, threads)

    out = {
        "metric": METRIC, "value": n / (ms * 1e-3), "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"plain 3D Laplace FMM, N={n} uniform [0,1)^3, q~U[-1,1), P={args.P}, "
                               f"theta={args.theta}, ncrit={args.ncrit}",
                   "m2l_truncation": args.truncation, "m2l_precision": args.m2l, "p2p_precision": "fp32",
                   "order_policy": "fixed"},
        "error": {"p": err_p, "F": err_F, "targets": int(len(idx)), "vs": "FP64 direct sum"},
        "stages_ms": stages, "work": work, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
