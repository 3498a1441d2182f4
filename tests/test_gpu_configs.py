"""Parity at each benchmark workload's own parameters and work-item plans,
against the oracle, bit for bit, at sizes the oracle finishes in seconds.

These run the code paths the benchmark runs: the reference local stage on
the full 60000 x 784 MNIST-like input; whole-leaf and row-block work items;
one and several pooled tree groups; re-rank depths below 2 nn; leaves of
1-2K points at d = 960; nn = 32 with 1024-slot pools and FP8 candidates.
The reference leaf k-NN these lists follow is MDRPT::calculate_nns
(cpp/src/drpt/algo/mdrpt.cpp:123-212) with the cross-tree merge of :180-198.
"""
import sys

import numpy as np
import pytest

import drpt_amd
from drpt_amd import datagen

pytestmark = pytest.mark.gpu

WHOLE = 1 << 30


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _same(oi, od, ri, rd, what):
    oi, od = oi.cpu().numpy(), od.cpu().numpy()
    bad = np.flatnonzero((oi != ri).any(1) | (_bits(od) != _bits(rd)).any(1))
    assert bad.size == 0, f"{what}: {bad.size} rows differ from the oracle (first {bad[:5]})"


def _oracle(orc, X, p):
    ri, rd, _, _ = orc.build_graph(X, p.ntrees, p.depth, p.nn, local_mode=p.local_mode,
                                   leaf_max=p.leaf_max, seed=p.seed, density=p.density)
    return ri, rd


def test_c1_mnist_like_full_reference_local_stage(ctx, orc):
    """60000 x 784 uint8, T = 8, nn = 10, the reference parameter rule and
    local stage (main.cpp:141-145, mdrpt.cpp:321-400), FP32."""
    _log("[c1] 60000 x 784 u8, T=8, reference local stage")
    X = datagen.mnist_like(60000)
    p = drpt_amd.derive_params(60000, 784, 8, 10)
    p.dtype = drpt_amd.DRPT_U8
    p.seed = 1
    oi, od, tm = ctx.build_graph(p, _dev(X))
    ri, rd = _oracle(orc, X, p)
    _same(oi, od, ri, rd, "c1")


def test_c2_int_mixture_tuned_plans(ctx, orc):
    """100000 x 128 integer mixture, T = 16, nn = 10, leaf <= 512, BF16
    k' = 24: the plan the benchmark's autotuner chose at 1M points (whole
    leaves resident in LDS, one 16-tree group, 4 first trees, re-rank 20),
    the same with first bounds from approximate distances alone (re-rank -1,
    one and four tree groups), the same with the streaming LDS tile ring, a
    row-block plan with four groups and a shallow re-rank, the autotuner's
    own plan for this input, and the FP32 build."""
    _log("[c2] 100000 x 128 int mixture, T=16, leaf<=512, bf16 k'=24")
    n = 100_000
    X = datagen.int_mixture(n)
    Xd = _dev(X)
    p = drpt_amd.leafmax_params(n, 128, 16, 10, leaf_max=512, precision="bf16", kprime=24, seed=1)
    ri, rd = _oracle(orc, X, p)
    plans = [
        {"rows_per_item": [WHOLE] * 4, "ring": [2] * 4, "tree_group": 16, "first_trees": 4,
         "rerank": 20},
        {"rows_per_item": [WHOLE] * 4, "ring": [2] * 4, "tree_group": 16, "first_trees": 4,
         "rerank": -1},
        {"rows_per_item": [WHOLE] * 4, "ring": [2] * 4, "tree_group": 4, "first_trees": 2,
         "rerank": -1},
        {"rows_per_item": [WHOLE] * 4, "ring": [1] * 4, "tree_group": 16, "first_trees": 4,
         "rerank": 20},
        {"rows_per_item": [WHOLE, 64, 64, 128], "ring": [1, 0, 1, 1], "tree_group": 4,
         "first_trees": 2, "rerank": 15},
    ]
    for plan in plans:
        _log(f"  plan {plan}")
        ctx.set_item_plan(p, plan)
        oi, od, tm = ctx.build_graph(p, Xd)
        _same(oi, od, ri, rd, f"c2 plan {plan}")
    tuned = ctx.autotune(p, Xd)
    _log(f"  autotuned {tuned}")
    oi, od, tm = ctx.build_graph(p, Xd)
    _same(oi, od, ri, rd, f"c2 autotuned {tuned}")
    p32 = drpt_amd.leafmax_params(n, 128, 16, 10, leaf_max=512, precision="fp32", seed=1)
    oi, od, _ = ctx.build_graph(p32, Xd)
    _same(oi, od, ri, rd, "c2 fp32")


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_leaf_resident_and_streaming_leaves(ctx, orc, prec):
    """60000 x 128 integer mixture, T = 8, leaves up to 1024: leaves the
    LDS-resident passes stage (at d = 128 up to 608 rows for the candidate
    pass and 544 for the pool pass) and larger ones,
    whose work items are gathered for the streaming kernel, in one build."""
    _log(f"[leaf-split] 60000 x 128 int mixture, T=8, leaf<=1024, {prec}")
    n = 60_000
    X = datagen.int_mixture(n)
    p = drpt_amd.leafmax_params(n, 128, 8, 10, leaf_max=1024, precision=prec, kprime=24, seed=3)
    ctx.set_item_plan(p, {"rows_per_item": [WHOLE] * 4, "ring": [2] * 4, "tree_group": 8,
                          "first_trees": 4, "rerank": 20})
    oi, od, tm = ctx.build_graph(p, _dev(X))
    _log(f"  largest leaf {tm.max_leaf}, leaves {tm.n_leaves}")
    assert tm.max_leaf > 608
    ri, rd = _oracle(orc, X, p)
    _same(oi, od, ri, rd, f"leaf-split {prec}")


def test_c3_gauss960_leaves_to_2048(ctx, orc):
    """20000 x 960 Gaussian mixture, T = 12, nn = 10, leaves up to 2048 (so
    they reach 1-2K points), BF16 with FP64 norms, the row-block plan
    (64-row items for 257-1024, 128-row items beyond) and 6-tree groups."""
    _log("[c3] 20000 x 960 gauss mixture, T=12, leaf<=2048, bf16")
    n = 20_000
    X = datagen.gauss_mixture(n)
    p = drpt_amd.leafmax_params(n, 960, 12, 10, leaf_max=2048, precision="bf16", kprime=24,
                                seed=1)
    ctx.set_item_plan(p, {"rows_per_item": [WHOLE, WHOLE, 64, 128], "ring": [1] * 4,
                          "tree_group": 6, "first_trees": 0, "rerank": 0})
    oi, od, tm = ctx.build_graph(p, _dev(X))
    _log(f"  largest leaf {tm.max_leaf}, exact rows {tm.n_fallback_rows}")
    assert tm.max_leaf > 1024
    ri, rd = _oracle(orc, X, p)
    _same(oi, od, ri, rd, "c3")


@pytest.mark.parametrize("centres,rerank", [(4096, 48), (20, 48), (4096, -1), (20, -1)])
def test_c4_sphere_nn32_fp8(ctx, orc, centres, rerank):
    """50000 x 96 unit-sphere mixture, T = 32, nn = 32, leaf <= 1024, FP8
    k' = 64, 8-tree groups, re-rank 48 or first bounds from the FP8 distances
    alone (-1); with the workload's 4096 centres, and with 20 (the 10M-point
    workload's points per centre: dense clusters, many pool survivors per row)."""
    _log(f"[c4] 50000 x 96 sphere ({centres} centres), T=32, nn=32, fp8 k'=64, rerank {rerank}")
    n = 50_000
    X = datagen.sphere_mixture(n, centres=centres)
    p = drpt_amd.leafmax_params(n, 96, 32, 32, leaf_max=1024, precision="fp8", kprime=64, seed=1)
    ctx.set_item_plan(p, {"rows_per_item": [WHOLE] * 4, "ring": [1] * 4, "tree_group": 8,
                          "first_trees": 0, "rerank": rerank})
    oi, od, tm = ctx.build_graph(p, _dev(X))
    _log(f"  exact rows {tm.n_fallback_rows}")
    ri, rd = _oracle(orc, X, p)
    _same(oi, od, ri, rd, f"c4 ({centres} centres)")


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_locality_copy_same_lists(ctx, orc, prec):
    """-locality 1: the row merges read X in tree 0's leaf order (a permuted
    copy and each point's position in it); the lists equal the oracle's and
    those of the build without it, for float and uint8 input."""
    _log(f"[locality] 30000 x 128 int mixture, T=8, leaf<=512, {prec}")
    n = 30_000
    X = datagen.int_mixture(n)
    for u8 in (False, True):
        Xi = X.astype(np.uint8) if u8 else X
        p = drpt_amd.leafmax_params(n, 128, 8, 10, leaf_max=512, precision=prec, kprime=24, seed=5)
        p.dtype = drpt_amd.DRPT_U8 if u8 else drpt_amd.DRPT_F32
        p.locality = 1
        oi, od, _ = ctx.build_graph(p, _dev(Xi))
        ri, rd = _oracle(orc, Xi, p)
        _same(oi, od, ri, rd, f"locality {prec} u8={u8}")
