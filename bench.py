"""Benchmark: exact k-NN queries/sec on one MI355X.

Workload (the metric's configuration): n = 4,194,304 points, d = 128, x = A z + 0.01 e
with z ~ N(0, I_16) (the ``lowrank16_n4M_d128`` generator), m = 131,072 queries from
the same model, 2,048 representatives, k = 10.  One step = one exact k-query of all m
queries with inputs and outputs resident in HBM (build and proposal training excluded,
timed separately).  The timed path is the BF16 MFMA filter + exact FP32 re-rank with the
learned list proposal (MLP d -> 256 -> 256 -> R, FP16, trained on the GPU with soft
targets: the owner lists of the k exact nearest neighbours of 2^20 training queries, found
by the exact search itself).  Beside it, once each on the same index and queries: the same search without the
proposal, and the plain FP32 exact RBC.

``roofline`` is for the dominant kernel, the list scan: FLOPs the scan evaluates per
scanned (query, point) pair (64 on the 32-coordinate sketches, 2 ldb on full BF16 rows;
pairs counted on the device) over its hipEvent-timed duration on the query stream.  ``traffic`` is its HBM bytes per launch from two
``rocprofv3 --pmc`` passes (FETCH_SIZE, WRITE_SIZE) of a child run of the same query,
each scaled by what the counter reads for a 1 GiB device copy.

``query_roofline`` is the whole step against the sum of its stages' algorithmic bounds
(q -> reps FLOPs at the FP32 peak, proposal and probe and scan FLOPs at the BF16/FP16 MFMA
peak, re-rank gathers and query/result bytes at the HBM peak).  ``measured_peaks`` are
the device's own BF16 MFMA and FP32 FMA rates (``rbc_measure_peaks``) and the
bandwidth of a 4 GiB device copy, next to the published figures used for ``frac``.

``cpu_baseline`` is the CPU exact RBC (the same index's lists, triangle and Voronoi-plane
pruning, AVX2 + OpenMP on the host's cores) on a bounded query sample; ``cpu_other`` has
the reference's single-thread bruteCPU (k = 1) and an all-core FP32 brute-force k-NN.

``--workload batched256`` times the batched form instead: 256 independent databases
answered by one hipGraph-replayed persistent launch.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import csv
import glob
import importlib.util
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "randomballcover-gpu_amd")

MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
SCAN_KERNELS = ("sk_scan_kernel", "bf16_scan_kernel")  # sketch scan, full-dimension BF16 scan


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_calibration(mib):
    """One device-to-device copy of ``mib`` MiB (reference kernel for the HBM counters)."""
    import torch
    a = torch.empty(mib << 20, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    a.fill_(1)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel()}))


def measured_peaks(rbc, torch, mib=4096, reps=5):
    """The device's BF16 MFMA and FP32 FMA rates and the bandwidth of a device copy."""
    peaks = rbc.measure_peaks()
    a = torch.empty(mib << 20, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    a.fill_(1)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(reps):
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del a, b
    torch.cuda.empty_cache()
    peaks["hbm_copy_GBps"] = 2.0 * (mib << 20) / (best * 1e-3) / 1e9
    return {k: round(v, 1) for k, v in peaks.items()}


def query_roofline(st, m, d, k, num_reps, ms_step, mlp_dims):
    """Whole-step roofline: the step's time against the sum of its stages' bounds."""
    mfma = MFMA_BF16_PEAK_TFLOPS * 1e12
    bound = {
        "qreps": 2.0 * d * m * num_reps / (FP32_PEAK_TFLOPS * 1e12),
        "proposal": (2.0 * m * sum(a * b for a, b in zip(mlp_dims[:-1], mlp_dims[1:])) / mfma
                     if mlp_dims else 0.0),
        "probe": 2.0 * d * st["pairs_probe"] / mfma,
        "scan": st["flops_scan"] * st["pairs_scan"] / mfma,
        "rerank": 4.0 * d * st["pairs_rerank"] / (HBM_PEAK_GBS * 1e9),
        "io": (4.0 * d * m + 8.0 * k * m) / (HBM_PEAK_GBS * 1e9),
    }
    total_ms = 1e3 * sum(bound.values())
    return {"bound_ms": round(total_ms, 3), "ms_per_step": round(ms_step, 3),
            "frac": round(total_ms / ms_step, 4),
            "stage_bound_ms": {kk: round(1e3 * v, 4) for kk, v in bound.items()}}


BATCHED_WORKLOAD = "batched256"


def bench_batched(args, rbc, datagen, torch, peaks):
    """The batched form: 256 independent databases (n log-uniform in [2^10, 2^18], d in
    {32, 96}, m log-uniform in [1, 4096], ceil(sqrt(n)) reps, k = 16), all answered by one
    persistent launch captured in a hipGraph.  The first call autotunes the query tile and
    captures the graph; a step is one replay over all 256 query sets, inputs and outputs in
    HBM."""
    ns, dims, ms = datagen.batched_shapes()
    xs = [datagen.uniform(int(ns[i]), int(dims[i]), 100 + i) for i in range(len(ns))]
    qs = [datagen.uniform(int(ms[i]), int(dims[i]), 10000 + i) for i in range(len(ns))]
    reps = [int(np.ceil(np.sqrt(v))) for v in ns]
    k = 16
    t0 = time.time()
    b = rbc.BatchedRBC(xs, reps)
    build_s = time.time() - t0
    dev = torch.device("cuda:0")
    qd = [torch.from_numpy(q).to(dev) for q in qs]
    ids = [torch.empty((q.shape[0], k), dtype=torch.int32, device=dev) for q in qs]
    ds = [torch.empty((q.shape[0], k), dtype=torch.float32, device=dev) for q in qs]
    first = b.kquery_device(qd, k, ids, ds)
    ref = [(i.clone(), d.clone()) for i, d in zip(ids, ds)]
    for _ in range(args.warmup):
        b.kquery_device(qd, k, ids, ds)
    torch.cuda.synchronize()
    t0 = time.time()
    stats = [b.kquery_device(qd, k, ids, ds) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall = time.time() - t0
    m_total = int(ms.sum())
    qps = m_total * args.steps / wall
    same = all(torch.equal(i, ri) and torch.equal(d, rd) for (i, d), (ri, rd) in zip(zip(ids, ds), ref))
    st = stats[-1]
    log(f"{BATCHED_WORKLOAD}: {len(ns)} databases, {int(ns.sum())} points, {m_total} queries, k={k}; "
        f"build {build_s:.1f}s; first call {first['ms_graph']:.2f} ms (autotune {first['tune_ms']}, "
        f"{first['tile_queries']} queries per tile, {first['blocks_per_cu']} workgroups per CU)")
    log(f"graph replay: {1e3 * wall / args.steps:.3f} ms/step ({st['ms_graph']:.3f} ms on the GPU), "
        f"{qps:,.0f} q/s, replayed {bool(st['graph_replayed'])}, same results as the first call: {same}")
    # Roofline of the persistent kernel: the FLOPs it evaluates (2 d per (query, point)
    # and per (query, rep) distance, counted on the device) over the graph's hipEvent time
    # (which also holds the re-rank kernel).  For L2 (every instance here) the wave kernel
    # takes its distances on split-BF16 MFMA, three products per distance product, so the
    # peak for these FLOPs is a third of the dense BF16 rate.
    gflop = st["flops"] / 1e9
    ach = gflop / st["ms_graph"]  # TFLOP/s
    split_peak = MFMA_BF16_PEAK_TFLOPS / 3.0
    roof = {"bound": "mfma", "achieved": round(ach, 2), "peak": round(split_peak, 1), "unit": "TFLOP/s",
            "frac": round(ach / split_peak, 4), "traffic": None,
            "kernel": "batched_wave_kernel + batched_select (one graph replay)",
            "flops_per_launch": st["flops"], "pairs_per_launch": st["pairs_scan"],
            "pairs_frac_of_brute": round(st["pairs_scan"] / float((ns.astype(np.float64) * ms).sum()), 4),
            "peak_note": "dense BF16 MFMA rate / 3 (split-BF16: hi.hi + hi.lo + lo.hi per product)",
            "frac_of_measured_mfma": round(ach / (peaks["bf16_mfma_tflops"] / 3.0), 4)}
    log(f"batched kernel: {st['steps']:.4e} 64-point steps over all tiles, longest tile "
        f"{st['longest_tile_ms']:.2f} ms")
    roof["steps"] = st["steps"]
    roof["longest_tile_ms"] = round(st["longest_tile_ms"], 3)
    log(f"batched kernel: {st['pairs_scan']:.4e} pairs, {gflop:.1f} GFLOP in {st['ms_graph']:.2f} ms = "
        f"{ach:.1f} TFLOP/s ({ach / split_peak:.1%} of the split-BF16 MFMA peak)")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # CPU baseline leg

        threads = min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                      len(os.sched_getaffinity(0)))
        # A sample proportional to each instance's batch: its first ceil(m_i / frac) queries,
        # through the CPU exact RBC over the instance's own lists.
        frac = args.batched_cpu_frac
        t_cpu, nq, agree = 0.0, 0, 0
        for i in range(len(ns)):
            s = int(np.ceil(int(ms[i]) / frac))
            L = b.lists(i)
            c = oracle.CpuRBC(xs[i], L["rep_ids"], L["perm"], L["offsets"], L["radii"], True, threads)
            t0 = time.time()
            ci, _, _ = c.knn(qs[i][:s], k, threads)
            t_cpu += time.time() - t0
            nq += s
            agree += int((ci == ids[i][:s].cpu().numpy().astype(np.uint32)).all(1).sum())
        cpu = {"value": round(nq / t_cpu, 2), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"first ceil(m_i/{frac}) queries of every instance ({nq} queries), CPU exact "
                         f"RBC (triangle + Voronoi-plane pruning, AVX2 + OpenMP) over each instance's "
                         f"lists, {t_cpu:.1f}s of search",
               "ids_equal_gpu_sample": round(agree / nq, 6)}
        log(f"cpu exact RBC {nq / t_cpu:.1f} q/s on {threads} threads over {nq} queries ({t_cpu:.1f}s)")
    line = {
        "metric": "exact k-NN queries/sec, 256 databases in one batched launch (k=16)",
        "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * wall / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": BATCHED_WORKLOAD, "instances": len(ns), "n_total": int(ns.sum()),
                   "m_total": m_total, "k": k, "tile_queries": int(st["tile_queries"]),
                   "build_s": round(build_s, 2)},
        "roofline": roof, "cpu_baseline": cpu,
        "graph_ms": round(st["ms_graph"], 3), "graph_replayed": bool(st["graph_replayed"]),
        # L2 instances run the wave kernel (16-query tiles), whose tuner tries a quarter, half
        # and all of the workgroups per CU that fit; the tile kernel's tries 16-, 32- and
        # 64-query tiles.
        "autotune_ms": ({"quarter_of_resident_workgroups": round(first["tune_ms"][0], 3),
                         "half": round(first["tune_ms"][1], 3), "all": round(first["tune_ms"][2], 3)}
                        if first["blocks_per_cu"] > 1 else
                        {"16": round(first["tune_ms"][0], 3), "32": round(first["tune_ms"][1], 3),
                         "64": round(first["tune_ms"][2], 3)}),
        "blocks_per_cu": int(first["blocks_per_cu"]),
        "uncertified_queries": int(st["uncertified"]), "replays_match_first_call": same,
        "measured_peaks": peaks,
    }
    print(json.dumps(line), flush=True)
    b.close()


def train_proposal(rbc, datagen, idx, workload, d, num_reps, k, count, epochs, seed=0):
    """Trains the list proposal on (query, owner lists of its exact k NNs) pairs produced by
    the exact search itself (soft targets: uniform over the k lists)."""
    import torch
    qt = datagen.training_queries(workload, count)
    qtd = torch.from_numpy(qt).to("cuda:0")
    nnd = torch.empty((count, k), dtype=torch.int32, device="cuda:0")
    dnd = torch.empty((count, k), dtype=torch.float32, device="cuda:0")
    idx.kquery_device(qtd, k, nnd, dnd, precision=rbc.PREC_BF16)
    nn = nnd.cpu().numpy().view(np.uint32)
    owner = idx.lists()["owner"]
    if (nn >= owner.shape[0]).any():
        raise RuntimeError(f"exact search returned {(nn >= owner.shape[0]).sum()} invalid ids")
    labels = owner[nn]
    net = rbc.ProposalMLP(d, num_reps, seed=seed)
    losses = net.train(qt, labels, epochs=epochs, batch=1024, lr=1e-3, seed=seed)
    # Held-out quality: how many of a query's k-NN lists the network's top 8 lists hold.
    top = np.argsort(-net.logits(qt[:4096]), 1)[:, :8]
    cover = float(np.mean([np.isin(labels[i], top[i]).mean() for i in range(4096)]))
    return net, losses, cover


def build_index(args, rbc, datagen, torch):
    x, q, num_reps, k = datagen.workload(args.workload, args.n, args.m)
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(x).to(dev)
    qd = torch.from_numpy(q).to(dev)
    torch.cuda.synchronize()
    t0 = time.time()
    idx = rbc.ExactRBC(xd, num_reps, seed=0, device=True)
    torch.cuda.synchronize()
    return x, q, num_reps, k, xd, qd, idx, time.time() - t0


def pmc_child(args):
    """Runs under rocprofv3 --pmc: the measured query (with the same proposal weights),
    twice, so that the last scan launch has a warm candidate-capacity estimate."""
    import torch
    rbc = _load("rbc_ctypes", os.path.join(PKG, "rbc_ctypes.py"))
    datagen = _load("rbc_datagen", os.path.join(PKG, "datagen.py"))
    x, q, num_reps, k, xd, qd, idx, _ = build_index(args, rbc, datagen, torch)
    net = None
    if args.mlp_params:
        net = rbc.ProposalMLP(x.shape[1], num_reps, seed=0)
        net.set_params(np.load(args.mlp_params))
        idx.set_proposal(net)
    ids = torch.empty((q.shape[0], k), dtype=torch.int32, device="cuda:0")
    ds = torch.empty((q.shape[0], k), dtype=torch.float32, device="cuda:0")
    for _ in range(2):
        idx.kquery_device(qd, k, ids, ds, precision=rbc.PREC_BF16)
    torch.cuda.synchronize()
    idx.close()


def _pmc_pass(counter, child_args, outdir, timeout):
    """One rocprofv3 counter pass of a child bench process; returns the CSV rows."""
    prof = shutil.which("rocprofv3")
    d = os.path.join(outdir, counter.lower())
    cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__)] + child_args
    env = dict(os.environ, TMPDIR=tempfile.gettempdir())
    r = subprocess.run(cmd, cwd=tempfile.gettempdir(), env=env, stdout=subprocess.DEVNULL,
                       stderr=subprocess.PIPE, timeout=timeout)
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"rocprofv3 --pmc {counter}: no counter file (rc {r.returncode}): "
                           f"{r.stderr.decode(errors='replace')[-400:]}")
    with open(files[0]) as f:
        return [row for row in csv.DictReader(f) if row["Counter_Name"] == counter]


def measure_traffic(args, mlp_params, timeout=240):
    """HBM bytes per launch of the scan kernel: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) of its last launch, each scaled by the counter's reading for a 1 GiB copy."""
    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    out = tempfile.mkdtemp(prefix="rbc_pmc_")
    child = ["--pmc-child", "--workload", args.workload]
    if args.n:
        child += ["--n", str(args.n)]
    if args.m:
        child += ["--m", str(args.m)]
    if mlp_params:
        child += ["--mlp-params", mlp_params]
    res = {}
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            cal = _pmc_pass(counter, ["--copy-calibration", "1024"], os.path.join(out, "cal"), timeout)
            cal_kib = max(float(r["Counter_Value"]) for r in cal)
            factor = (1024.0 * 1024.0) / cal_kib  # 1 GiB = 1,048,576 KiB moved each way
            rows = [r for r in _pmc_pass(counter, child, out, timeout)
                    if any(kn in r["Kernel_Name"] for kn in SCAN_KERNELS)]
            if not rows:
                return None, f"no scan kernel dispatch in the {counter} pass"
            last = max(rows, key=lambda r: int(r["Dispatch_Id"]))
            res[counter] = {"kib": float(last["Counter_Value"]), "copy_factor": round(factor, 4)}
            log(f"pmc {counter}: {float(last['Counter_Value']):.0f} KiB x {factor:.3f} (copy calibration)")
    except subprocess.TimeoutExpired:
        return None, "rocprofv3 pass timed out"
    finally:
        shutil.rmtree(out, ignore_errors=True)
    tb = sum(v["kib"] * v["copy_factor"] * 1024.0 for v in res.values())
    return tb, res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="lowrank16_n4M_d128")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--m", type=int, default=None)
    ap.add_argument("--precision", default="bf16", choices=["bf16", "fp32"])
    ap.add_argument("--proposal", default="mlp", choices=["mlp", "none"],
                    help="learned list proposal in the timed path (BF16 only)")
    ap.add_argument("--train-queries", type=int, default=1 << 20)
    ap.add_argument("--train-epochs", type=int, default=3)
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 comparison step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM-traffic passes")
    ap.add_argument("--cpu-sample", type=int, default=8192,
                    help="queries timed for the CPU exact RBC baseline")
    ap.add_argument("--cpu-sample-brute", type=int, default=256)
    ap.add_argument("--cpu-sample-1thread", type=int, default=8)
    ap.add_argument("--batched-cpu-frac", type=int, default=32,
                    help="batched CPU baseline: every instance's first ceil(m_i / this) queries")
    ap.add_argument("--check-sample", type=int, default=256)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--pmc-child", action="store_true",
                    help="run only the measured query twice (for counter collection under rocprofv3)")
    ap.add_argument("--mlp-params", default=None,
                    help="with --pmc-child: proposal weights (.npy, as written by --save-mlp)")
    ap.add_argument("--save-mlp", default=None, metavar="NPY",
                    help="write the trained proposal's weights here (for a --pmc-child counter run)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              f"(queries shard across GPUs with no exchange; see the sharding design)",
              file=sys.stderr)
        sys.exit(2)
    if args.pmc_child:
        pmc_child(args)
        return

    import torch

    rbc = _load("rbc_ctypes", os.path.join(PKG, "rbc_ctypes.py"))
    datagen = _load("rbc_datagen", os.path.join(PKG, "datagen.py"))
    rbc.lib()
    peaks = measured_peaks(rbc, torch)
    log(f"measured peaks: BF16 MFMA {peaks['bf16_mfma_tflops']:,.0f} TFLOP/s, FP32 FMA "
        f"{peaks['fp32_fma_tflops']:,.1f} TFLOP/s, device copy {peaks['hbm_copy_GBps']:,.0f} GB/s")

    if args.workload == BATCHED_WORKLOAD:
        bench_batched(args, rbc, datagen, torch, peaks)
        return

    t0 = time.time()
    x, q, num_reps, k, xd, qd, idx, build_s = build_index(args, rbc, datagen, torch)
    n, d = x.shape
    m = q.shape[0]
    info = idx.info()
    log(f"{args.workload}: n={n} d={d} m={m} reps={num_reps} k={k}; data+build {time.time()-t0:.1f}s "
        f"(build {build_s:.2f}s), lists {info['min_list']}..{info['max_list']}")
    prec = rbc.PREC_BF16 if args.precision == "bf16" else rbc.PREC_FP32
    dev = torch.device("cuda:0")
    ids = torch.empty((m, k), dtype=torch.int32, device=dev)
    ds = torch.empty((m, k), dtype=torch.float32, device=dev)

    extra = {}
    net = None
    mlp_params = None
    use_mlp = args.proposal == "mlp" and prec == rbc.PREC_BF16

    def timed_arm():
        """W untimed warm-up steps, then K timed steps bracketed by synchronisations (no
        statistics requested: no per-stage events, nothing copied back), then K more steps,
        not timed here, whose statistics give the stage times and pair counts."""
        for _ in range(args.warmup):
            idx.kquery_device(qd, k, ids, ds, precision=prec, stats=False)
        torch.cuda.synchronize()
        t0 = time.time()
        for _ in range(args.steps):
            idx.kquery_device(qd, k, ids, ds, precision=prec, stats=False)
        torch.cuda.synchronize()
        wall = time.time() - t0
        sts = [idx.kquery_device(qd, k, ids, ds, precision=prec) for _ in range(args.steps)]
        return wall, sts

    if use_mlp:
        t0 = time.time()
        net, losses, cover = train_proposal(rbc, datagen, idx, args.workload, d, num_reps, k,
                                            args.train_queries, args.train_epochs)
        train_s = time.time() - t0
        log(f"proposal trained on {args.train_queries} queries' {k}-NN lists in {train_s:.1f}s: loss "
            f"{losses[0]:.3f} -> {losses[-1]:.3f}; its top 8 lists hold {100 * cover:.1f} % of the "
            f"training queries' {k}-NN lists")
        extra["proposal"] = {"train_queries": args.train_queries, "labels_per_query": k,
                             "epochs": args.train_epochs, "train_s": round(train_s, 2),
                             "loss_first": round(float(losses[0]), 4), "loss_last": round(float(losses[-1]), 4),
                             "top8_list_cover_train": round(cover, 4), "params": net.num_params,
                             "mass": 0.9}
        if args.save_mlp:
            np.save(args.save_mlp, net.get_params())
        # The same search without the proposal, timed exactly like the headline arm.
        w_plain, s_plain_all = timed_arm()
        s_plain = s_plain_all[-1]
        ref_ids, ref_ds = ids.clone(), ds.clone()
        idx.set_proposal(net)

    wall, stats = timed_arm()
    log("timed steps done")
    ms_step = 1e3 * wall / args.steps
    qps = m * args.steps / wall
    st = stats[-1]
    log(f"{args.precision}{'+mlp' if use_mlp else ''}: {ms_step:.2f} ms/step, {qps:,.0f} q/s; stages "
        f"qreps {st['ms_qreps']:.2f} probe {st['ms_probe']:.2f} plan {st['ms_plan']:.2f} "
        f"scan {st['ms_scan']:.2f} select {st['ms_select']:.2f} ms")
    # The same query with host buffers (queries in, results out over PCIe); not `value`.
    t0 = time.time()
    idx.kquery(q, k, precision=prec)
    w_host = time.time() - t0
    extra["pcie_inclusive_qps"] = round(m / w_host, 1)
    log(f"host-buffer query (PCIe-inclusive): {1e3*w_host:.1f} ms ({m/w_host:,.0f} q/s)")
    if use_mlp:
        same = bool(torch.equal(ref_ids, ids) and torch.equal(ref_ds, ds))
        gpu_plain = [s["ms_total"] for s in s_plain_all]
        gpu_with = [s["ms_total"] for s in stats]
        extra["without_proposal"] = {
            "qps": round(m * args.steps / w_plain, 1), "ms_per_step": round(1e3 * w_plain / args.steps, 3),
            "gpu_ms_min_max": [round(min(gpu_plain), 3), round(max(gpu_plain), 3)],
            "pairs_scan": s_plain["pairs_scan"], "pairs_probe": s_plain["pairs_probe"],
            "stages_ms": {s_: round(s_plain[f"ms_{s_}"], 3) for s_ in ("qreps", "probe", "plan", "scan", "select")},
            "same_results": same}
        extra["with_proposal"] = {"gpu_ms_min_max": [round(min(gpu_with), 3), round(max(gpu_with), 3)],
                                  "pairs_scan": st["pairs_scan"], "pairs_probe": st["pairs_probe"]}
        extra["learned_gain"] = round(qps / (m * args.steps / w_plain), 4)
        log(f"without proposal: {1e3*w_plain/args.steps:.2f} ms/step ({m*args.steps/w_plain:,.0f} q/s, GPU "
            f"{min(gpu_plain):.2f}..{max(gpu_plain):.2f} ms); with it {ms_step:.2f} ms/step (GPU "
            f"{min(gpu_with):.2f}..{max(gpu_with):.2f} ms): gain x{extra['learned_gain']:.4f}; pairs probed "
            f"{s_plain['pairs_probe']:.4e} -> {st['pairs_probe']:.4e}, scanned {s_plain['pairs_scan']:.4e} -> "
            f"{st['pairs_scan']:.4e}; identical results: {same}")

    # The same search with the full-dimension BF16 scan instead of the sketch scan, timed
    # like the headline arm (results must be identical).
    if prec == rbc.PREC_BF16 and st["sketch"]:
        ref_ids2, ref_ds2 = ids.clone(), ds.clone()
        idx.set_sketch(rbc.SKETCH_OFF)
        w_ns, s_ns_all = timed_arm()
        idx.set_sketch(rbc.SKETCH_AUTO)
        s_ns = s_ns_all[-1]
        extra["without_sketch"] = {
            "qps": round(m * args.steps / w_ns, 1), "ms_per_step": round(1e3 * w_ns / args.steps, 3),
            "scan_ms": round(float(np.mean([s_["ms_scan"] for s_ in s_ns_all])), 3),
            "scan_tflops": round(s_ns["flops_scan"] * s_ns["pairs_scan"] /
                                 (float(np.mean([s_["ms_scan"] for s_ in s_ns_all])) * 1e-3) / 1e12, 2),
            "pairs_scan": s_ns["pairs_scan"], "pairs_probe": s_ns["pairs_probe"], "proposal": use_mlp,
            "same_results": bool(torch.equal(ref_ids2, ids) and torch.equal(ref_ds2, ds))}
        extra["sketch_gain"] = round(qps / (m * args.steps / w_ns), 4)
        extra["sketch"] = idx.sketch_info()
        log(f"without sketch: {1e3*w_ns/args.steps:.2f} ms/step ({m*args.steps/w_ns:,.0f} q/s), scan "
            f"{extra['without_sketch']['scan_ms']:.2f} ms; gain x{extra['sketch_gain']:.3f}; identical: "
            f"{extra['without_sketch']['same_results']}")
        # Leave the headline results in ids/ds for the checks below.
        ids.copy_(ref_ids2)
        ds.copy_(ref_ds2)

    # Roofline of the dominant kernel (the list scan): algorithmic FLOPs per launch over
    # its hipEvent-timed duration.
    scan_ms = float(np.mean([s["ms_scan"] for s in stats]))
    flops = st["flops_scan"] * st["pairs_scan"]
    tflops = flops / (scan_ms * 1e-3) / 1e12
    if prec == rbc.PREC_BF16:
        roof = {"bound": "mfma", "achieved": round(tflops, 2), "peak": MFMA_BF16_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4),
                "traffic": None, "kernel": "sk_scan_kernel" if st["sketch"] else "bf16_scan_kernel<4>",
                "flops_per_pair": st["flops_scan"],
                "flops_per_launch": flops, "pairs_per_launch": st["pairs_scan"],
                "alg_bytes_per_launch": st["bytes_scan"],
                "alg_GBps": round(st["bytes_scan"] / (scan_ms * 1e-3) / 1e9, 1),
                "kernel_ms": round(scan_ms, 3),
                "frac_of_measured_peak": round(tflops / peaks["bf16_mfma_tflops"], 4)}
    else:
        roof = {"bound": "valu", "achieved": round(tflops, 2), "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "traffic": None,
                "kernel": "tile_kernel<EpiEmit>", "kernel_ms": round(scan_ms, 3)}

    qroof = None
    if prec == rbc.PREC_BF16:
        qroof = query_roofline(st, m, d, k, num_reps, ms_step, net.dims if use_mlp else None)
        log(f"whole step: {qroof['bound_ms']:.2f} ms of stage bounds in {ms_step:.2f} ms "
            f"({100 * qroof['frac']:.1f} %)")

    if not args.no_fp32 and prec == rbc.PREC_BF16:
        idx.set_proposal(None)
        ids32 = torch.empty_like(ids)
        ds32 = torch.empty_like(ds)
        idx.kquery_device(qd, k, ids32, ds32, precision=rbc.PREC_FP32)  # warm-up, as for BF16
        torch.cuda.synchronize()
        t0 = time.time()
        s32 = idx.kquery_device(qd, k, ids32, ds32, precision=rbc.PREC_FP32)
        torch.cuda.synchronize()
        w32 = time.time() - t0
        extra["fp32_plain_qps"] = round(m / w32, 1)
        extra["speedup_vs_fp32_plain"] = round(qps / (m / w32), 2)
        extra["fp32_matches_bf16"] = bool(torch.equal(ids32, ids) and torch.equal(ds32, ds))
        # The plain FP32 path plans differently (a one-list FP32 probe and planes against the
        # nearest rep only), so it scans more pairs; its scan kernel's own rate, and the
        # per-pair ratio of the two scans, isolate the precision.
        f32_tflops = 2.0 * d * s32["pairs_scan"] / (s32["ms_scan"] * 1e-3) / 1e12
        extra["fp32_plain"] = {
            "ms": round(1e3 * w32, 3), "pairs_scan": s32["pairs_scan"], "pairs_probe": s32["pairs_probe"],
            "scan_ms": round(s32["ms_scan"], 3), "scan_tflops": round(f32_tflops, 2),
            "scan_frac_of_fp32_peak": round(f32_tflops / FP32_PEAK_TFLOPS, 4),
            "stages_ms": {s_: round(s32[f"ms_{s_}"], 3) for s_ in ("qreps", "probe", "plan", "scan", "select")},
            "scan_speedup_per_pair_bf16": round((s32["ms_scan"] / s32["pairs_scan"]) /
                                                (scan_ms / st["pairs_scan"]), 2)}
        log(f"fp32 plain: {w32*1e3:.1f} ms ({m/w32:,.0f} q/s), scan {s32['ms_scan']:.1f} ms over "
            f"{s32['pairs_scan']:.4e} pairs ({f32_tflops:.1f} TFLOP/s); per pair the BF16 scan is "
            f"x{extra['fp32_plain']['scan_speedup_per_pair_bf16']} faster")
        if net is not None:
            idx.set_proposal(net)

    if prec == rbc.PREC_BF16 and not args.no_traffic:
        tmp = None
        if net is not None:
            fd, tmp = tempfile.mkstemp(suffix=".npy", prefix="rbc_mlp_")
            os.close(fd)
            np.save(tmp, net.get_params())
        try:
            tb, detail = measure_traffic(args, tmp)
        finally:
            if tmp:
                os.unlink(tmp)
        if tb is not None:
            roof["traffic"] = tb
            roof["traffic_counters"] = detail
            roof["traffic_over_alg"] = round(tb / st["bytes_scan"], 3)
            log(f"scan HBM traffic {tb/1e9:.2f} GB/launch ({tb/st['bytes_scan']:.2f}x algorithmic)")
        else:
            roof["traffic_note"] = detail
            log(f"traffic not measured: {detail}")

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle  # CPU baseline leg (and its checker)

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = min(threads, len(os.sched_getaffinity(0)))
        # CPU exact RBC over the GPU index's own lists (same reps, radii and Voronoi cells).
        L = idx.lists()
        t0 = time.time()
        crbc = oracle.CpuRBC(x, L["rep_ids"], L["perm"], L["offsets"], L["radii"], True, threads)
        cpu_build = time.time() - t0
        ns = min(args.cpu_sample, m)
        t0 = time.time()
        ci, _, cpairs = crbc.knn(q[:ns], k, threads)
        cs = time.time() - t0
        gi_s = ids[:ns].cpu().numpy().astype(np.uint32)
        cpu = {"value": round(ns / cs, 2), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"first {ns} of {m} queries, CPU exact RBC (triangle + Voronoi-plane "
                         f"pruning, AVX2 + OpenMP) over the same index's lists, {cs:.1f}s; "
                         f"{cpairs / ns / n:.4f} of the points scanned per query; layout "
                         f"{cpu_build:.1f}s (not timed)",
               "ids_equal_gpu_sample": float((ci == gi_s).all(1).mean())}
        log(f"cpu exact RBC {ns/cs:.2f} q/s on {threads} threads ({cs:.1f}s), "
            f"{cpairs/ns/n:.4f} of points per query")
        del crbc
        # The reference's bruteCPU (brute.cu:91-112): FP32 1-NN, one thread.
        n1 = max(1, min(args.cpu_sample_1thread, m))
        t0 = time.time()
        oracle.brute_cpu(x, q[:n1], 0, 1)
        c1 = time.time() - t0
        # FP32 brute-force k-NN (kernels.cu distance arithmetic), all cores.
        nb = min(args.cpu_sample_brute, m)
        t0 = time.time()
        oracle.brute_k_f32(x, q[:nb], k, 0, threads)
        cb = time.time() - t0
        extra["cpu_other"] = {
            "bruteCPU_1nn_1thread": {"value": round(n1 / c1, 3), "unit": "queries/s", "cores": 1,
                                     "sample": f"first {n1} queries, k=1, {c1:.1f}s"},
            "brute_fp32_knn_allcores": {"value": round(nb / cb, 2), "unit": "queries/s",
                                        "cores": threads, "sample": f"first {nb} queries, {cb:.1f}s"},
        }
        log(f"bruteCPU 1 thread {n1/c1:.3f} q/s; FP32 brute k-NN {nb/cb:.2f} q/s on {threads} threads")
        # Checker on a sample: FP64 oracle vs the GPU result.
        nc = min(args.check_sample, m)
        oi, od = oracle.brute_k_f64(x, q[:nc], k, 0, threads)
        gi = ids[:nc].cpu().numpy().astype(np.uint32)
        gd = ds[:nc].cpu().numpy()
        recall = float(np.mean([len(set(gi[i]) & set(oi[i])) / k for i in range(nc)]))
        rel = float(np.max(np.abs(gd.astype(np.float64) - od) / np.maximum(od, 1e-300)))
        extra.update({"check_sample": nc, "recall_at_k_sample": recall,
                      "ids_exact_sample": bool((gi == oi).all()), "max_rel_err_vs_fp64_sample": rel})
        log(f"FP64 check on {nc} queries: recall {recall}, ids exact {bool((gi == oi).all())}, "
            f"max rel err {rel:.2e}")

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
                   "precision": args.precision, "proposal": "mlp" if use_mlp else "none",
                   "build_s": round(build_s, 2),
                   "pairs_scan_frac_of_brute": round((st["pairs_scan"] + st["pairs_probe"]) / (n * m), 4)},
        "roofline": roof,
        "cpu_baseline": cpu,
        "query_roofline": qroof,
        "measured_peaks": peaks,
        "stages_ms": {s: round(st[f"ms_{s}"], 3) for s in ("qreps", "probe", "plan", "scan", "select")},
        **extra,
    }
    print(json.dumps(line), flush=True)
    idx.close()


if __name__ == "__main__":
    main()
