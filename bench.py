"""k-NN graph build throughput on one MI355X.

Default workload: 1,000,000 x 128 float32 (integers 0-218, 256-centre
mixture, ~35% zeros), 16 trees, nn = 10, leaves <= 512, BF16 MFMA candidate
Gram (FP32 accumulate, k' = 24) + exact FP32 re-rank. A step is one whole
graph build (projection, splits, leaf grouping, leaf k-NN, merge) with X
already resident in HBM. Also reported:
  * every stage's time with its algorithmic work and fraction of the MI355X
    peak (HBM 8 TB/s; dense MFMA BF16 2.5 / FP8 5.0 PFLOP/s; FP32 157.3
    TFLOP/s), and sum(work / peak) / build time;
  * the end-to-end time with X copied from pinned host memory and the lists
    copied back;
  * recall@10 (and @nn, and the reference-style nn - 1 edges) against exact
    FP64 distances: every query for the 60000 x 784 and 1M x 128 workloads,
    10,000 sampled queries with a 95% interval otherwise; the largest and 99th
    percentile relative error of the returned distances against FP64;
  * the plain FP32 build of the same input (its lists must equal the BF16 /
    FP8 ones bit for bit) and the roofline of its leaf k-NN kernel;
  * the oracle (CPU restatement, OpenMP) on the same workload, median of three
    runs where one takes under 25 s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c3|c4]
                    [--precision bf16|fp8|fp32] [--no-cpu] [--no-fp32] [--no-recall]
                    [--no-autotune | --item-plan JSON] [--density D]
"""
import argparse
import json
import math
import os
import sys
import time

# the CPU baseline's OpenMP threads bound to their cores (read when the
# OpenMP runtime loads, so before torch and the oracle are imported)
os.environ.setdefault("OMP_PROC_BIND", "true")
# CPUs this process may run on, read before any OpenMP runtime binds the main
# thread to a single core (after that, the main thread's mask lists one CPU)
START_AFFINITY = (len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                  else (os.cpu_count() or 1))

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "kNN-graph build points/sec at recall@10≥0.9; BF16/FP8+rerank vs FP32, 1 GPU"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(name):
    from drpt_amd import datagen
    if name == "c1":
        return dict(desc="60000x784 uint8 MNIST-like (10 blob prototypes, ~19% non-zero), T=8, "
                         "nn=10, reference local stage", gen=lambda n: datagen.mnist_like(n),
                    n=60000, ntrees=8, nn=10, leaf_max=0, kprime=0, mode="reference",
                    cpu_n=60000)
    if name == "c2":
        return dict(desc="1000000x128 float32 integer 256-centre mixture, T=16, nn=10, "
                         "leaf<=512, k'=24", gen=lambda n: datagen.int_mixture(n), n=1_000_000,
                    ntrees=16, nn=10, leaf_max=512, kprime=24, mode="leafmax", cpu_n=1_000_000)
    if name == "c3":
        return dict(desc="1000000x960 float32 1000-centre Gaussian mixture, T=12, nn=10, "
                         "leaf<=2048", gen=lambda n: datagen.gauss_mixture(n), n=1_000_000,
                    ntrees=12, nn=10, leaf_max=2048, kprime=24, mode="leafmax", cpu_n=20_000)
    if name == "c4":
        return dict(desc="10000000x96 float32 unit-L2 4096-centre mixture, T=32, nn=32, "
                         "leaf<=1024, FP8 k'=64", gen=lambda n: datagen.sphere_mixture(n),
                    n=10_000_000, ntrees=32, nn=32, leaf_max=1024, kprime=64, mode="leafmax",
                    cpu_n=50_000)
    raise SystemExit(f"unknown config {name}")


def make_params(w, n, d, dtype_u8, precision, seed=1):
    import drpt_amd
    if w["mode"] == "reference":
        p = drpt_amd.derive_params(n, d, w["ntrees"], w["nn"])
        p.precision = drpt_amd.PRECISIONS[precision]
        p.kprime = w["kprime"]
        p.seed = seed
    else:
        p = drpt_amd.leafmax_params(n, d, w["ntrees"], w["nn"], w["leaf_max"], precision,
                                    w["kprime"], seed)
    p.dtype = drpt_amd.DRPT_U8 if dtype_u8 else drpt_amd.DRPT_F32
    return p


def _exact_ok_fp32(X_dev):
    """Integer-valued rows whose squared norms stay below 2^23: every squared
    distance, Gram entry and partial sum is an integer below 2^24, so FP32
    matrix products give the exact values in any summation order."""
    import torch
    if X_dev.dtype == torch.uint8:
        Xf = X_dev.float()
    else:
        Xf = X_dev
        if not bool(torch.all(Xf == torch.round(Xf))):
            return False
    return float((Xf.double() ** 2).sum(1).max()) < 2.0 ** 23


def exact_knn_kth(X_dev, queries, K, budget_bytes=4 << 30):
    """The exact K-th smallest non-self squared distance of every query row
    (FP32 products when exact for this data, FP64 otherwise), queries in
    chunks whose distance block stays within budget_bytes."""
    import torch
    exact32 = _exact_ok_fp32(X_dev)
    dt = torch.float32 if exact32 else torch.float64
    chunk = max(16, min(2048, budget_bytes // (X_dev.shape[0] * (4 if exact32 else 8))))
    Xc = X_dev.to(dt)
    x2 = (Xc * Xc).sum(1)
    out = torch.empty(len(queries), dtype=torch.float64, device=X_dev.device)
    for s in range(0, len(queries), chunk):
        qq = queries[s:s + chunk]
        d = x2[qq][:, None] + x2[None, :] - 2.0 * (Xc[qq] @ Xc.T)
        d[torch.arange(len(qq), device=d.device), qq] = float("inf")
        out[s:s + len(qq)] = torch.topk(d, K, dim=1, largest=False).values[:, K - 1].double()
        del d
    return out, exact32


def recall_report(X_dev, idx, dist, K, nn, exact_all, nq_sample=10000, seed=0):
    """Tie-aware recall against exact FP64 distances: a returned neighbour j of
    query q is a hit when D(q, j) <= the exact K-th non-self distance of q.
    exact_all: every point is a query; else nq_sample seeded uniform queries
    with a 95% confidence interval (normal approximation over per-query
    recall). Also: recall@10 on the first 10 of an nn-list, the reference-style
    figure (nn - 1 written edges), and the max / p99 relative error of the
    returned FP32 distances against FP64 over the sampled queries."""
    import torch
    n = X_dev.shape[0]
    dev = X_dev.device
    if exact_all:
        q = torch.arange(n, device=dev)
    else:
        rng = np.random.default_rng(seed)
        q = torch.from_numpy(rng.choice(n, size=min(nq_sample, n), replace=False)).to(dev)
    X64 = X_dev.double()
    res = {}
    for k in sorted({10, K, K - 1} - {0}):
        if k > K:
            continue
        kth, ex32 = exact_knn_kth(X_dev, q, k)
        got = idx[q][:, :k].long()
        valid = got >= 0
        hits = torch.zeros(len(q), dtype=torch.float64, device=dev)
        for s in range(0, len(q), 65536):
            qq = q[s:s + 65536]
            g = got[s:s + 65536]
            diff = X64[qq][:, None, :] - X64[g.clamp(min=0)]
            d64 = (diff * diff).sum(2)
            hits[s:s + 65536] = ((d64 <= kth[s:s + 65536, None]) & valid[s:s + 65536]).sum(1).double()
        per_q = hits / k
        m = float(per_q.mean())
        ci = 0.0 if exact_all else 1.96 * float(per_q.std()) / math.sqrt(len(q))
        res[k] = (m, ci, bool(ex32))
    # distance error against FP64 on (up to) nq_sample queries
    qs = q[:nq_sample]
    got = idx[qs].long()
    valid = got >= 0
    diff = X64[qs][:, None, :] - X64[got.clamp(min=0)]
    d64 = (diff * diff).sum(2)
    rel = ((dist[qs].double() - d64).abs() / d64.clamp(min=1e-300))[valid]
    err_max = float(rel.max()) if rel.numel() else 0.0
    err_p99 = float(torch.quantile(rel.float(), 0.99)) if rel.numel() else 0.0
    return res, err_max, err_p99, len(q)


def source_hash(*names):
    """sha256 prefix of the kernel sources a measured profile belongs to."""
    import hashlib
    h = hashlib.sha256()
    for nme in names:
        with open(os.path.join(ROOT, "drpt_amd", "csrc", nme), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def pmc_traffic(config, prec, kernel):
    """HBM bytes per launch of the dominant kernel, from the FETCH_SIZE / WRITE_SIZE
    counter runs summarised under profiles/ (scaled by the copy-probe factors), or None
    when no such measurement exists for this configuration or the kernel sources have
    changed since it was taken."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("config") != config or t.get("precision") != prec or t.get("kernel") != kernel:
        return None
    if t.get("sources_sha16") != source_hash(*t.get("sources", [])):
        return None
    return t["traffic_bytes"]


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(w, seed, repeats=3, budget_s=25.0):
    """The oracle (C++ / OpenMP restatement of the reference build) on the
    workload at full size, on every CPU this process may run on
    (OMP_PROC_BIND=true), median of `repeats` runs (one when a run takes more
    than budget_s)."""
    import oracle
    n = w["cpu_n"]
    X = w["gen"](n)
    log(f"[bench] cpu baseline input {X.shape} ready")
    # the CPU share of this process: OMP_NUM_THREADS where the host sets it
    # (the affinity mask can list more CPUs than the process may use)
    aff = START_AFFINITY
    threads = int(os.environ.get("OMP_NUM_THREADS") or aff)
    p = make_params(w, n, X.shape[1], X.dtype == np.uint8, "fp32", seed)
    times = []
    for r in range(repeats):
        t = time.perf_counter()
        oracle.build_graph(X, p.ntrees, p.depth, p.nn,
                           local_mode=0 if w["mode"] == "reference" else 1, leaf_max=p.leaf_max,
                           seed=seed, density=p.density, nthreads=threads)
        times.append(time.perf_counter() - t)
        log(f"[bench] cpu baseline run {r + 1}: {times[-1]:.2f} s")
        if times[-1] > budget_s:
            break
    el = float(np.median(times))
    return {"value": n / el, "unit": "points/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_info(), "nproc": os.cpu_count(), "affinity_cpus": aff,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "runs_s": [round(x, 3) for x in times],
            "sample": f"{n} points of the same generator and parameters (T={p.ntrees}, nn={p.nn}, "
                      f"depth {p.depth}), oracle build_graph, median of {len(times)} run(s): {el:.2f} s"}


HBM_PEAK = 8.0e12   # bytes/s
MFMA_BF16 = 2.5e15  # flop/s, dense
MFMA_FP8 = 5.0e15
MFMA_I8 = 5.0e15    # op/s, dense
FP32_PEAK = 157.3e12


def phase_rooflines(p, d, L, xbytes, tm, ph, prec):
    """Algorithmic work of each stage of one build (bytes for HBM-bound stages,
    flops for the Gram passes), its rate over the measured time and its
    fraction of the MI355X peak; plus sum(W / peak) / T_build."""
    n, T, K, D = p.n, p.ntrees, p.nn, p.depth
    gram = tm["gram_flops"]
    out = {}

    def add(name, ms, work, peak, unit):
        if ms is None or ms <= 0:
            return
        rate = work / (ms / 1e3)
        out[name] = {"ms": ms, "work": work, "unit": unit, "achieved": rate, "peak": peak,
                     "frac": rate / peak, "ideal_ms": work / peak * 1e3}

    add("project", tm["project_ms"], n * d * xbytes + 4.0 * n * L, HBM_PEAK, "B")
    add("split", tm["split_ms"], 12.0 * n * T * (D - 1), HBM_PEAK, "B")
    add("group", ph["group_ms"], 12.0 * n * T, HBM_PEAK, "B")
    if prec == "fp32":
        add("leaf_knn_exact", ph["pool_ms"], 1.5 * gram, FP32_PEAK, "flop")
        add("merge", ph["pool_merge_ms"], n * (8.0 * K * T + 8.0 * K), HBM_PEAK, "B")
    else:
        q = 1 if prec == "fp8" else 2
        first = max(1, int(ph["first_trees"]))
        kp = p.kprime if p.kprime else min(64, 2 * K + 4)
        # integer data in [0, 255]: the LDS-resident passes' int8 Gram (and the
        # quantisation kernel's int8 copy)
        i8 = bool(ph.get("int8_passes"))
        add("prep", ph["prep_ms"], n * d * (xbytes + 2 + (1 if prec == "fp8" or i8 else 0)) + 8.0 * n
            + 12.0 * T * n * (2 if prec == "fp8" else 1), HBM_PEAK, "B")
        gpeak, gunit = (MFMA_I8, "op") if i8 else (MFMA_FP8 if prec == "fp8" else MFMA_BF16, "flop")
        add("candidate_pass", ph["cand_ms"], gram * first / T, gpeak, gunit)
        if ph["rerank"] < 0:
            # first bounds from approximate distances: the keys, one (|x~|^2, e) gather per slot
            add("first_bound", ph["known_ms"], n * first * kp * 16.0, HBM_PEAK, "B")
        else:
            add("known_merge", ph["known_ms"],
                n * (ph["rerank"] * d * xbytes + first * kp * 8.0 + 8.0 * K), HBM_PEAK, "B")
        add("bound_gather", ph["bound_ms"], 12.0 * T * n, HBM_PEAK, "B")
        add("pool_pass", ph["pool_ms"], gram, gpeak, gunit)
        add("pool_merge", ph["pool_merge_ms"], n * (16.0 * K * T + 4.0 * K), HBM_PEAK, "B")
    ideal = sum(v["ideal_ms"] for v in out.values())
    return out, ideal / tm["total_ms"] if tm["total_ms"] > 0 else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--precision", default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-recall", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--density", type=float, default=None,
                    help="projection density (an integer >= 1 gives the reference's dense "
                         "local projection, -density on the reference command line)")
    ap.add_argument("--locality", type=int, default=0,
                    help="1: the row merges read X in tree 0's leaf order (-locality 1); 0: X "
                         "itself (default: measured faster)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip drpt_autotune (work-item layout per leaf-size bucket) before warm-up")
    ap.add_argument("--item-plan", default=None, metavar="JSON",
                    help='use this plan instead of tuning, e.g. \'{"rows_per_item": [1073741824, '
                         '1073741824, 64, 64], "ring": [2, 2, 1, 1], "tree_group": 6}\' (as '
                         'autotune reports it)')
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (scale for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    import drpt_amd

    if a.copy_probe > 0:
        src = torch.ones(a.copy_probe << 18, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 4, "copies": 3}))
        return

    w = workload(a.config)
    prec = a.precision or {"c1": "fp32", "c4": "fp8"}.get(a.config, "bf16")
    log(f"[bench] generating {w['desc']}")
    X = w["gen"](w["n"])
    n, d = X.shape
    Xd = torch.from_numpy(X).to("cuda:0")
    ctx = drpt_amd.Context(0)
    p = make_params(w, n, d, X.dtype == np.uint8, prec, a.seed)
    if a.density is not None:
        p.density = a.density
    p.locality = a.locality
    out = (torch.empty((n, p.nn), dtype=torch.int32, device="cuda:0"),
           torch.empty((n, p.nn), dtype=torch.float32, device="cuda:0"))

    def run(params, steps, warmup, label):
        for i in range(warmup):
            ctx.build_graph(params, Xd, out)
            log(f"[bench] {label} warmup {i + 1}/{warmup} done")
        torch.cuda.synchronize()
        tms, phs = [], []
        t0 = time.perf_counter()
        for i in range(steps):
            _, _, tm = ctx.build_graph(params, Xd, out)
            tms.append(tm.as_dict())
            phs.append(ctx.last_phases())
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        log(f"[bench] {label}: {steps} steps in {el:.3f} s")
        return el / steps * 1e3, tms, phs

    plan = None
    if prec != "fp32" and a.item_plan:
        plan = json.loads(a.item_plan)
        ctx.set_item_plan(p, plan)
    elif prec != "fp32" and not a.no_autotune:
        t = time.perf_counter()
        plan = ctx.autotune(p, Xd)  # one-time set-up, kept by the context
        log(f"[bench] autotune {plan} in {time.perf_counter() - t:.2f} s")
    ms, tms, phs = run(p, a.steps, a.warmup, prec)
    idx_main = out[0].clone()
    dist_main = out[1].clone()
    tm_last = tms[-1]
    phases = {k: float(np.mean([t[k] for t in tms])) for k in
              ("project_ms", "split_ms", "group_ms", "knn_ms", "merge_ms", "total_ms",
               "knn_kernel_ms")}
    stages = {k: float(np.mean([q[k] for q in phs])) for k in phs[-1]}
    xbytes = 1 if X.dtype == np.uint8 else 4
    L = p.ntrees * p.depth
    tmean = dict(tm_last)
    tmean.update(phases)
    rl_phases, sum_frac = phase_rooflines(p, d, L, xbytes, tmean, stages, prec)
    # integer data in [0, 255]: the candidate and pool Grams ran on int8 rows
    int8 = prec != "fp32" and bool(stages.get("int8_passes"))
    res = {"metric": METRIC, "value": n / (ms / 1e3), "unit": "points/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int8" if int8 else prec,
           "data": "synthetic (seeded generator, see drpt_amd/datagen.py)",
           "config": {"workload": w["desc"], "n": n, "d": d, "ntrees": p.ntrees, "nn": p.nn,
                      "depth": p.depth, "leaf_max": p.leaf_max, "precision": prec,
                      "kprime": p.kprime if prec != "fp32" else 0,
                      "local_mode": w["mode"], "density": p.density, "locality": p.locality},
           "phases_ms": phases, "stages_ms": stages, "phase_rooflines": rl_phases,
           "sum_phase_fraction": sum_frac,
           "n_leaves": tm_last["n_leaves"], "max_leaf": tm_last["max_leaf"],
           "fallback_rows": tm_last["n_fallback_rows"], "item_plan": plan}

    # end to end: X from pinned host memory, lists back to the host
    Xh = torch.from_numpy(X).pin_memory()
    oi_h = torch.empty((n, p.nn), dtype=torch.int32).pin_memory()
    od_h = torch.empty((n, p.nn), dtype=torch.float32).pin_memory()
    Xe = torch.empty_like(Xd)
    e2e = []
    for _ in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        Xe.copy_(Xh, non_blocking=True)
        ctx.build_graph(p, Xe, out)
        oi_h.copy_(out[0], non_blocking=True)
        od_h.copy_(out[1], non_blocking=True)
        torch.cuda.synchronize()
        e2e.append((time.perf_counter() - t) * 1e3)
    res["end_to_end_ms"] = float(np.median(e2e))
    res["end_to_end_points_per_s"] = n / (res["end_to_end_ms"] / 1e3)
    del Xe

    if not a.no_recall:
        log("[bench] recall")
        exact_all = a.config in ("c1", "c2")
        t = time.perf_counter()
        rc, emax, ep99, nq = recall_report(Xd, idx_main, dist_main, p.nn, p.nn, exact_all)
        log(f"[bench] recall over {nq} queries in {time.perf_counter() - t:.1f} s")
        r10 = rc.get(10, rc[min(rc)])
        res["recall_at_10"] = r10[0]
        res["recall_at_10_ci95"] = r10[1]
        res["recall_at_nn"] = rc[p.nn][0]
        res["recall_ref_style_nn_minus_1"] = rc[p.nn - 1][0] if p.nn - 1 in rc else None
        res["recall_queries"] = nq
        res["recall_exact_all_queries"] = exact_all
        res["recall_at_10_meets_0_90"] = bool(r10[0] >= 0.90)
        res["dist_max_rel_err_vs_fp64"] = emax
        res["dist_p99_rel_err_vs_fp64"] = ep99

    if prec != "fp32" and not a.no_fp32:
        p32 = make_params(w, n, d, X.dtype == np.uint8, "fp32", a.seed)
        if a.density is not None:
            p32.density = a.density
        ms32, tms32, phs32 = run(p32, max(1, min(3, a.steps)), 1, "fp32")
        res["fp32_points_per_s"] = n / (ms32 / 1e3)
        res["fp32_ms_per_step"] = ms32
        res["speedup_vs_fp32"] = ms32 / ms
        res["lists_identical_to_fp32"] = bool(torch.equal(out[0], idx_main) and torch.equal(
            out[1].view(torch.int32), dist_main.view(torch.int32)))
        k32 = float(np.mean([q["pool_ms"] for q in phs32]))
        fl32 = 1.5 * tms32[-1]["gram_flops"]
        res["fp32_leaf_knn_roofline"] = {
            "kernel": "k_knn_exact", "bound": "vector FP32", "ms": k32, "flops": fl32,
            "achieved": fl32 / (k32 / 1e3) / 1e12, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
            "frac": fl32 / (k32 / 1e3) / FP32_PEAK,
            "frac_of_nonpacked_issue": fl32 / (k32 / 1e3) / (FP32_PEAK / 2)}

    kms = phases["knn_kernel_ms"]
    flops = tm_last["gram_flops"]
    if prec == "fp32":
        # exact FP32 kernel on the vector ALU: 3 flops (sub, fma) per dimension per pair
        fl = 1.5 * flops
        kernel = "k_knn_exact"
        peak = FP32_PEAK
    else:
        # the pooled filter passes: the BF16 Gram of every leaf of every tree, once,
        # over groups of trees: flops and durations summed over the launches (the
        # leaf-resident kernel and, for leaves larger than its LDS, the streaming one)
        fl = flops
        kernel = "k_pool_leaf + k_knn_lowp (pool mode)"
        peak = MFMA_I8 if int8 else MFMA_BF16
        res["pool_launches"] = stages.get("tree_groups")
    achieved = fl / (kms / 1e3)
    res["roofline"] = {"bound": "mfma" if prec != "fp32" else "fp32", "kernel": kernel,
                       "achieved": achieved / 1e12, "peak": peak / 1e12,
                       "unit": "TOP/s" if int8 else "TFLOP/s",
                       "frac": achieved / peak, "traffic": pmc_traffic(a.config, prec, kernel),
                       "flops_per_launch": fl, "kernel_ms": kms}
    if not a.no_cpu:
        log("[bench] cpu baseline (oracle)")
        res["cpu_baseline"] = cpu_baseline(w, a.seed, a.cpu_repeats)
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
