"""Benchmark: exact k-NN queries/sec on one MI355X.

Workload (the metric's configuration): n = 4,194,304 points, d = 128, x = A z + 0.01 e
with z ~ N(0, I_16) (the ``lowrank16_n4M_d128`` generator), m = 131,072 queries from
the same model, 2,048 representatives, k = 10.  One step = one exact k-query of all m
queries with inputs and outputs resident in HBM (build excluded, timed separately).
The timed path is the BF16 MFMA filter + exact FP32 re-rank; the plain FP32 exact RBC
is timed once on the same index for the speed-up, and the CPU baseline (the oracle's
FP32 brute force, the reference's CPU arithmetic) on a bounded query sample.

Prints one JSON line (rank 0).  ``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "randomballcover-gpu_amd")

MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16
HBM_PEAK_GBS = 8000.0


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lowrank16_n4M_d128")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--m", type=int, default=None)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp32"])
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 comparison step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(args.copy_calibration << 20, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel()}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build",
              file=sys.stderr)
        sys.exit(2)

    import torch

    rbc = _load("rbc_ctypes", os.path.join(PKG, "rbc_ctypes.py"))
    datagen = _load("rbc_datagen", os.path.join(PKG, "datagen.py"))
    rbc.lib()
    dev = torch.device("cuda:0")

    t0 = time.time()
    x, q, num_reps, k = datagen.workload(args.workload, args.n, args.m)
    n, d = x.shape
    m = q.shape[0]
    gen_s = time.time() - t0
    print(f"[bench] data {args.workload}: n={n} d={d} m={m} reps={num_reps} k={k} "
          f"(generated in {gen_s:.1f}s)", file=sys.stderr, flush=True)

    xd = torch.from_numpy(x).to(dev)
    qd = torch.from_numpy(q).to(dev)
    ids = torch.empty((m, k), dtype=torch.int32, device=dev)
    ds = torch.empty((m, k), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    t0 = time.time()
    idx = rbc.ExactRBC(xd, num_reps, seed=0, device=True)
    torch.cuda.synchronize()
    build_s = time.time() - t0
    info = idx.info()
    print(f"[bench] build {build_s:.2f}s, lists {info['min_list']}..{info['max_list']}",
          file=sys.stderr, flush=True)

    prec = rbc.PREC_BF16 if args.precision == "bf16" else rbc.PREC_FP32
    for _ in range(args.warmup):
        idx.kquery_device(qd, k, ids, ds, precision=prec)
    torch.cuda.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)
    stats = []
    t0 = time.time()
    for _ in range(args.steps):
        stats.append(idx.kquery_device(qd, k, ids, ds, precision=prec))
    torch.cuda.synchronize()
    wall = time.time() - t0
    ms_step = 1e3 * wall / args.steps
    qps = m * args.steps / wall
    st = stats[-1]
    print(f"[bench] {args.precision}: {ms_step:.2f} ms/step, {qps:,.0f} q/s; stages "
          f"qreps {st['ms_qreps']:.2f} probe {st['ms_probe']:.2f} plan {st['ms_plan']:.2f} "
          f"scan {st['ms_scan']:.2f} select {st['ms_select']:.2f} ms", file=sys.stderr, flush=True)

    # Roofline of the dominant kernel (the list scan): algorithmic FLOPs and bytes per
    # launch over its hipEvent-timed duration.
    scan_ms = float(np.mean([s["ms_scan"] for s in stats]))
    flops = 2.0 * d * st["pairs_scan"]
    tflops = flops / (scan_ms * 1e-3) / 1e12
    gbs = st["bytes_scan"] / (scan_ms * 1e-3) / 1e9
    if prec == rbc.PREC_BF16:
        roof = {"bound": "mfma", "achieved": round(tflops, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4),
                "traffic": None, "kernel": "bf16_scan_kernel",
                "flops_per_launch": flops, "alg_bytes_per_launch": st["bytes_scan"],
                "alg_GBps": round(gbs, 1), "kernel_ms": round(scan_ms, 3)}
    else:
        peak = 157.3
        roof = {"bound": "valu", "achieved": round(tflops * 1.5, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(tflops * 1.5 / peak, 4), "traffic": None, "kernel": "tile_kernel<EpiEmit>",
                "kernel_ms": round(scan_ms, 3)}

    extra = {}
    if not args.no_fp32 and prec == rbc.PREC_BF16:
        torch.cuda.synchronize()
        t0 = time.time()
        ids32 = torch.empty_like(ids)
        ds32 = torch.empty_like(ds)
        s32 = idx.kquery_device(qd, k, ids32, ds32, precision=rbc.PREC_FP32)
        torch.cuda.synchronize()
        w32 = time.time() - t0
        extra["fp32_plain_qps"] = round(m / w32, 1)
        extra["speedup_vs_fp32_plain"] = round(qps / (m / w32), 2)
        extra["fp32_matches_bf16"] = bool(torch.equal(ids32, ids) and torch.equal(ds32, ds))
        print(f"[bench] fp32 plain: {w32*1e3:.1f} ms ({m/w32:,.0f} q/s), scan {s32['ms_scan']:.1f} ms",
              file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # CPU baseline leg (and its checker)

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ns = min(args.cpu_sample, m)
        t0 = time.time()
        oi32, od32 = oracle.brute_k_f32(x, q[:ns], k, 0, threads)
        cs = time.time() - t0
        cpu = {"value": round(ns / cs, 2), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"first {ns} of {m} queries, FP32 brute force over all {n} points "
                         f"(bruteK arithmetic, brute.cu:65-88), {cs:.1f}s"}
        # Checker on the same sample: FP64 oracle vs the GPU result.
        oi, od = oracle.brute_k_f64(x, q[:ns], k, 0, threads)
        gi = ids[:ns].cpu().numpy().astype(np.uint32)
        gd = ds[:ns].cpu().numpy()
        recall = float(np.mean([len(set(gi[i]) & set(oi[i])) / k for i in range(ns)]))
        rel = float(np.max(np.abs(gd.astype(np.float64) - od) / np.maximum(od, 1e-300)))
        extra.update({"recall_at_k_sample": recall, "ids_exact_sample": bool((gi == oi).all()),
                      "max_rel_err_vs_fp64_sample": rel})
        print(f"[bench] cpu baseline {ns/cs:.2f} q/s on {threads} threads; recall {recall}, "
              f"max rel err {rel:.2e}", file=sys.stderr, flush=True)

    line = {
        "metric": "exact k-NN queries/sec, n=4M d=128 k=10 recall=1.0 (BF16+learned vs FP32 plain)",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if prec == rbc.PREC_BF16 else "fp32",
        "data": "synthetic",
        "config": {"workload": args.workload, "n": n, "d": d, "m": m, "k": k, "num_reps": num_reps,
                   "precision": args.precision, "build_s": round(build_s, 2),
                   "pairs_scan_frac_of_brute": round((st["pairs_scan"] + st["pairs_probe"]) / (n * m), 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_ms": {s: round(st[f"ms_{s}"], 3) for s in ("qreps", "probe", "plan", "scan", "select")},
        **extra,
    }
    print(json.dumps(line), flush=True)
    idx.close()


if __name__ == "__main__":
    main()
