"""k-NN graph build throughput on one MI355X.

Default workload: 1,000,000 x 128 float32 (integers 0-218, 256-centre
mixture, ~35% zeros), 16 trees, nn = 10, leaves <= 512, BF16 MFMA candidate
Gram (FP32 accumulate, k' = 24) + exact FP32 re-rank. A step is one whole
graph build (projection, splits, leaf grouping, leaf k-NN, merge) with X
already resident in HBM. Also reported: the plain FP32 build of the same
input (its lists must equal the BF16 ones bit for bit), recall@10 against an
FP64 brute force on sampled queries, the roofline of the leaf k-NN kernel and
the oracle (CPU restatement, OpenMP) on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c3|c4]
                    [--precision bf16|fp8|fp32] [--no-cpu] [--no-fp32] [--no-autotune | --item-plan JSON]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "kNN-graph build points/sec at recall@10≥0.9; BF16/FP8+rerank vs FP32, 1 GPU"
PEAK = {"bf16": 2500.0, "fp32": 157.3}  # dense MFMA TFLOP/s, MI355X


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
                    ntrees=16, nn=10, leaf_max=512, kprime=24, mode="leafmax", cpu_n=200_000)
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


def recall_fp64(X_dev, idx, K, nq=2000, seed=0, dist=None):
    """Tie-aware recall@K of idx (n x K, non-self) against FP64 brute force; with dist,
    also the largest relative error of the returned distances against the FP64
    squared distance of the same pairs."""
    import torch
    n = X_dev.shape[0]
    rng = np.random.default_rng(seed)
    q = torch.from_numpy(rng.choice(n, size=min(nq, n), replace=False)).to(X_dev.device)
    X64 = X_dev.double()
    x2 = (X64 * X64).sum(1)
    hits = hits_ref = 0
    err = 0.0
    for s in range(0, len(q), 256):
        qq = q[s:s + 256]
        d = x2[qq][:, None] + x2[None, :] - 2.0 * (X64[qq] @ X64.T)
        d[torch.arange(len(qq)), qq] = float("inf")
        kth = torch.topk(d, K, dim=1, largest=False).values[:, K - 1]
        got = idx[qq].long()
        valid = got >= 0
        dg = torch.gather(d, 1, got.clamp(min=0))
        hits += int(((dg <= kth[:, None]) & valid).sum())
        kth9 = torch.topk(d, K - 1, dim=1, largest=False).values[:, K - 2]
        hits_ref += int(((dg[:, :K - 1] <= kth9[:, None]) & valid[:, :K - 1]).sum())
        if dist is not None:
            diff = X64[qq][:, None, :] - X64[got.clamp(min=0)]
            d64 = (diff * diff).sum(2)
            got_d = dist[qq].double()
            rel = ((got_d - d64).abs() / d64.clamp(min=1e-300))[valid]
            if rel.numel():
                err = max(err, float(rel.max()))
    if dist is not None:
        return hits / (len(q) * K), hits_ref / (len(q) * (K - 1)), err
    return hits / (len(q) * K), hits_ref / (len(q) * (K - 1))


def pmc_traffic(config, prec, kernel):
    """HBM bytes per launch of the dominant kernel, from the FETCH_SIZE / WRITE_SIZE
    counter runs summarised under profiles/ (scaled by the copy-probe factors), or None
    when no such measurement exists for this configuration."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("config") != config or t.get("precision") != prec or t.get("kernel") != kernel:
        return None
    return t["traffic_bytes"]


def cpu_baseline(w, seed):
    import oracle
    import drpt_amd
    n = w["cpu_n"]
    X = w["gen"](n)
    threads = min(os.cpu_count() or 1, 16)
    p = make_params(w, n, X.shape[1], X.dtype == np.uint8, "fp32", seed)
    t = time.perf_counter()
    oracle.build_graph(X, p.ntrees, p.depth, p.nn,
                       local_mode=0 if w["mode"] == "reference" else 1, leaf_max=p.leaf_max,
                       seed=seed, density=p.density, nthreads=threads)
    el = time.perf_counter() - t
    return {"value": n / el, "unit": "points/s", "cores": threads, "kind": "port",
            "sample": f"{n} points of the same generator and parameters (T={p.ntrees}, "
                      f"nn={p.nn}, depth {p.depth}), oracle build_graph, {el:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--precision", default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip drpt_autotune (work-item layout per leaf-size bucket) before warm-up")
    ap.add_argument("--item-plan", default=None, metavar="JSON",
                    help='use this plan instead of tuning, e.g. \'{"rows_per_item": [1073741824, '
                         '1073741824, 64, 64], "ring": [1, 1, 1, 1], "tree_group": 6}\' (as '
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
    out = (torch.empty((n, p.nn), dtype=torch.int32, device="cuda:0"),
           torch.empty((n, p.nn), dtype=torch.float32, device="cuda:0"))

    def run(params, steps, warmup, label):
        for i in range(warmup):
            ctx.build_graph(params, Xd, out)
            log(f"[bench] {label} warmup {i + 1}/{warmup} done")
        torch.cuda.synchronize()
        tms = []
        t0 = time.perf_counter()
        for i in range(steps):
            _, _, tm = ctx.build_graph(params, Xd, out)
            tms.append(tm.as_dict())
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        log(f"[bench] {label}: {steps} steps in {el:.3f} s")
        return el / steps * 1e3, tms

    plan = None
    if prec != "fp32" and a.item_plan:
        plan = json.loads(a.item_plan)
        ctx.set_item_plan(p, plan)
    elif prec != "fp32" and not a.no_autotune:
        t = time.perf_counter()
        plan = ctx.autotune(p, Xd)  # one-time set-up, kept by the context
        log(f"[bench] autotune {plan} in {time.perf_counter() - t:.2f} s")
    ms, tms = run(p, a.steps, a.warmup, prec)
    idx_main = out[0].clone()
    dist_main = out[1].clone()
    tm_last = tms[-1]
    phases = {k: float(np.mean([t[k] for t in tms])) for k in
              ("project_ms", "split_ms", "group_ms", "knn_ms", "merge_ms", "total_ms",
               "knn_kernel_ms")}
    res = {"metric": METRIC, "value": n / (ms / 1e3), "unit": "points/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": prec,
           "data": "synthetic (seeded generator, see drpt_amd/datagen.py)",
           "config": {"workload": w["desc"], "n": n, "d": d, "ntrees": p.ntrees, "nn": p.nn,
                      "depth": p.depth, "leaf_max": p.leaf_max, "precision": prec,
                      "kprime": p.kprime if prec != "fp32" else 0,
                      "local_mode": w["mode"]},
           "phases_ms": phases, "n_leaves": tm_last["n_leaves"], "max_leaf": tm_last["max_leaf"],
           "fallback_rows": tm_last["n_fallback_rows"], "item_plan": plan}

    log("[bench] recall")
    r, rref, derr = recall_fp64(Xd, idx_main, p.nn, dist=dist_main)
    res["recall_at_10"] = r
    res["recall_ref_style_nn_minus_1"] = rref
    res["dist_max_rel_err_vs_fp64"] = derr

    if prec != "fp32" and not a.no_fp32:
        p32 = make_params(w, n, d, X.dtype == np.uint8, "fp32", a.seed)
        ms32, _ = run(p32, max(1, min(3, a.steps)), 1, "fp32")
        res["fp32_points_per_s"] = n / (ms32 / 1e3)
        res["speedup_vs_fp32"] = ms32 / ms
        res["lists_identical_to_fp32"] = bool(torch.equal(out[0], idx_main) and torch.equal(
            out[1].view(torch.int32), dist_main.view(torch.int32)))

    kms = phases["knn_kernel_ms"]
    flops = tm_last["gram_flops"]
    if prec == "fp32":
        # exact FP32 kernel on the vector ALU: 3 flops (sub, fma) per dimension per pair
        fl = 1.5 * flops
        kernel = "k_knn_exact"
    else:
        # the pooled filter passes: the BF16 Gram of every leaf of every tree, once
        # (MFMA; FP8 builds use FP8 for the candidate pass only), over groups of
        # trees: flops and durations summed over the group launches
        fl = flops
        t = (p.kprime + 3) // 4
        t = next(s for s in (3, 4, 6, 8, 12, 16) if t <= s)
        rings = set(plan["ring"]) if plan else {1}
        path = ("ring" if rings == {1} else "regs" if rings == {0} else "mixed") \
            if (2 * d) % 16 == 0 and 2 * d <= 256 else "regs"
        kernel = f"k_knn_lowp<float, ElBf16, {t}, kModePool, {path}>"
        gsz = p.tree_group or (plan or {}).get("tree_group", 0) or \
            max(1, min(p.ntrees, max(4, 160 // p.nn)))
        res["pool_launches"] = -(-p.ntrees // gsz)
    peak = PEAK["fp32" if prec == "fp32" else "bf16"]
    achieved = fl / (kms / 1e3) / 1e12
    res["roofline"] = {"bound": "mfma", "kernel": kernel, "achieved": achieved,
                       "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                       "traffic": pmc_traffic(a.config, prec, kernel),
                       "flops_per_launch": fl, "kernel_ms": kms}
    if not a.no_cpu:
        log("[bench] cpu baseline (oracle)")
        res["cpu_baseline"] = cpu_baseline(w, a.seed)
        res["speedup_vs_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    ctx.close()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
