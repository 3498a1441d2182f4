"""GPU parity: the HIP path through the C ABI against the oracle, bit for bit
(integer tree assignments, candidate indices and FP32 distances), on seeded
inputs the oracle finishes in seconds."""
import sys

import numpy as np
import pytest

import drpt_amd
from drpt_amd import datagen

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


CASES = {
    "mnist_u8_3000": lambda: datagen.mnist_like(3000),
    "int128_f32_4000": lambda: datagen.int_mixture(4000),
    "int128_u8_4000": lambda: datagen.int_mixture(4000).astype(np.uint8),
    # even integers up to 436: exact in bf16 but not u8 values
    "int128x2_f32_4000": lambda: datagen.int_mixture(4000) * 2,
    # u8 values but one (256, still exact in bf16)
    "int128_one256_f32_4000": lambda: _with_value(datagen.int_mixture(4000), 123, 5, 256.0),
    "gauss128_f32_3000": lambda: (np.random.default_rng(8).standard_normal((3000, 128)) * 40.0
                                  + 100.0).astype(np.float32),
    "gauss_d37_1500": lambda: np.random.default_rng(5).standard_normal((1500, 37)).astype(np.float32),
    "gauss960_800": lambda: datagen.gauss_mixture(800),
    # rows scaled by 10^U(-3, 3), signs mixed, one zero row: per-row fp8 scales
    "scaled_d64_2000": lambda: _scaled_rows(2000, 64),
}


def _with_value(X, i, j, v):
    X[i, j] = v
    return X


def _scaled_rows(n, d):
    rng = np.random.default_rng(11)
    base = rng.standard_normal((n, d)).astype(np.float32)
    base *= (10.0 ** rng.uniform(-3, 3, size=(n, 1))).astype(np.float32)
    base[7] = 0.0
    return base


@pytest.mark.parametrize("case,L", [(c, 24) for c in CASES] +
                         [("mnist_u8_3000", 160), ("gauss960_800", 100), ("gauss960_800", 240),
                          ("gauss960_800", 600)])
def test_project_bitexact(ctx, orc, case, L):
    """X.R bit for bit; L = 160 / 100 shrink the row tile beside the staged R,
    L = 240 / 600 leave R in global memory."""
    _log(f"[project] {case} L={L}")
    X = CASES[case]()
    n, d = X.shape
    R = orc.build_R(d, L, 1.0 / np.sqrt(d), 17)
    P0, cm0 = orc.project(X, R)
    P, cm = ctx.project(_dev(X), R)
    assert np.array_equal(_bits(P.cpu().numpy()), _bits(P0))
    assert np.array_equal(cm.cpu().numpy().view(np.uint32), cm0)


@pytest.mark.parametrize("case,mode,leaf_max", [("mnist_u8_3000", 0, 0), ("int128_f32_4000", 0, 0),
                                                ("int128_f32_4000", 1, 64),
                                                ("gauss_d37_1500", 1, 20)])
def test_grow_trees_bitexact(ctx, orc, case, mode, leaf_max):
    _log(f"[grow] {case} mode={mode}")
    X = CASES[case]()
    n, d = X.shape
    T = 3
    D = orc.global_depth(orc.tree_depth(n)) if mode == 0 else 9
    R = orc.build_R(d, T * D, 1.0 / np.sqrt(d), 29)
    P0, cm0 = orc.project(X, R)
    heap = ctx.grow_trees(_dev(P0), _dev(cm0.view(np.int32)), T, D, mode, leaf_max)
    heap = heap.cpu().numpy().view(np.uint32)
    for t in range(T):
        ref = orc.grow_tree(P0, cm0, D, t, mode, leaf_max, 0)
        assert np.array_equal(heap[t], ref), f"tree {t}: {(heap[t] != ref).sum()} points differ"


@pytest.mark.parametrize("mode,D,leaf_max", [(0, 14, 0), (1, 15, 3), (1, 15, 64), (1, 18, 2)])
def test_grow_trees_deep_levels_bitexact(ctx, orc, mode, D, leaf_max):
    """Levels with 1024 - 8192 nodes per tree: per-node sums and histograms in
    large LDS arrays, then global atomics; with leaf_max 64 the deepest levels
    hold few splitting nodes among many leaves; depth 18 with leaf_max 2 takes
    heap indices beyond 2^16. The trees equal the oracle's."""
    _log(f"[grow deep] mode={mode} D={D}")
    X = datagen.int_mixture(40000)
    n, d = X.shape
    T = 2
    R = orc.build_R(d, T * D, 1.0 / np.sqrt(d), 31)
    P0, cm0 = orc.project(X, R)
    heap = ctx.grow_trees(_dev(P0), _dev(cm0.view(np.int32)), T, D, mode, leaf_max)
    heap = heap.cpu().numpy().view(np.uint32)
    for t in range(T):
        ref = orc.grow_tree(P0, cm0, D, t, mode, leaf_max, 0)
        assert np.array_equal(heap[t], ref), f"tree {t}: {(heap[t] != ref).sum()} points differ"


@pytest.mark.parametrize("n,T,D,leaf_max,trees", [
    (1000000, 16, 15, 512, (0, 7, 15)),  # the 1M x 128 build's split
    (300000, 8, 17, 64, (0, 5)),         # depth 17, levels of up to 32768 nodes
])
def test_grow_trees_full_scale(ctx, orc, n, T, D, leaf_max, trees):
    """Leaf-max trees at the size of the 1M x 128 workload: every tree's
    assignment equals the oracle's on the trees listed (the others are
    grown in the same launches)."""
    _log(f"[grow full] n={n} T={T} D={D} leaf<={leaf_max}")
    X = datagen.int_mixture(n)
    d = X.shape[1]
    R = orc.build_R(d, T * D, 1.0 / np.sqrt(d), 43)
    P, cm = ctx.project(_dev(X), R)
    heap = ctx.grow_trees(P, cm, T, D, 1, leaf_max).cpu().numpy().view(np.uint32)
    Ph = P.cpu().numpy()
    cmh = cm.cpu().numpy().view(np.uint32)
    for t in trees:
        ref = orc.grow_tree(Ph[t * D:(t + 1) * D].copy(), cmh[t * D:(t + 1) * D].copy(), D, 0, 1,
                            leaf_max, 0)
        assert np.array_equal(heap[t], ref), f"tree {t}: {(heap[t] != ref).sum()} points differ"


def test_split_graph_replay(ctx, orc):
    """The split levels run as a captured hipGraph: replays into the same buffers
    (with eager work and a clobbered output between them), a recapture for a
    new size and a recapture back all equal the oracle, and equal a context
    that launches eagerly (DRPT_GRAPHS=0)."""
    import os
    import torch
    _log("[graph] split replay")
    X = CASES["int128_f32_4000"]()
    n, d = X.shape
    T, D = 3, 9
    R = orc.build_R(d, T * D, 1.0 / np.sqrt(d), 29)
    P0, cm0 = orc.project(X, R)
    refs = {m: np.stack([orc.grow_tree(P0[:, :m].copy(), cm0, D, t, 1, 64, 0) for t in range(T)])
            for m in (n, 3000)}
    Pd, cmd = _dev(P0), _dev(cm0.view(np.int32))
    Pd3 = _dev(P0[:, :3000])
    heap = torch.empty((T, n), dtype=torch.int32, device="cuda:0")
    heap3 = torch.empty((T, 3000), dtype=torch.int32, device="cuda:0")
    for it, (P, out, m) in enumerate([(Pd, heap, n), (Pd, heap, n), (Pd, heap, n), (Pd3, heap3, 3000),
                                      (Pd, heap, n), (Pd, heap, n)]):
        out.fill_(-7)
        ctx.project(_dev(X[:64]), R)  # eager work between replays
        torch.cuda.synchronize()
        got = ctx.grow_trees(P, cmd, T, D, 1, 64, out=out).cpu().numpy().view(np.uint32)
        assert np.array_equal(got, refs[m]), f"call {it} (n={m}) differs from the oracle"
    os.environ["DRPT_GRAPHS"] = "0"
    try:
        eager = drpt_amd.Context(0)
    finally:
        del os.environ["DRPT_GRAPHS"]
    with eager:
        got = eager.grow_trees(Pd, cmd, T, D, 1, 64).cpu().numpy().view(np.uint32)
    assert np.array_equal(got, refs[n])


def _groups(heap_t):
    perm = np.argsort(heap_t, kind="stable").astype(np.int32)
    h = heap_t[perm]
    starts = np.flatnonzero(np.r_[True, h[1:] != h[:-1]])
    return perm, np.r_[starts, len(h)].astype(np.int64)


@pytest.mark.parametrize("case,K", [("int128_f32_4000", 10), ("gauss_d37_1500", 7),
                                    ("mnist_u8_3000", 16), ("gauss960_800", 32)])
def test_tree_knn_fp32_bitexact(ctx, orc, case, K):
    _log(f"[tree_knn] {case} K={K}")
    X = CASES[case]()
    n, d = X.shape
    p = drpt_amd.leafmax_params(n, d, 2, K, leaf_max=150, precision="fp32", seed=3)
    p.dtype = drpt_amd.DRPT_U8 if X.dtype == np.uint8 else drpt_amd.DRPT_F32
    R = orc.build_R(d, p.ntrees * p.depth, p.density, p.seed)
    P0, cm0 = orc.project(X, R)
    heaps = np.stack([orc.grow_tree(P0, cm0, p.depth, t, 1, p.leaf_max, 0) for t in range(2)])
    ci, cd, tm = ctx.tree_knn(p, _dev(X), _dev(heaps.view(np.int32)))
    ci, cd = ci.cpu().numpy(), cd.cpu().numpy()
    for t in range(2):
        perm, offs = _groups(heaps[t])
        ri, rd = orc.leaf_knn(X, perm, offs, K)
        assert np.array_equal(ci[t], ri)
        assert np.array_equal(_bits(cd[t]), _bits(rd))


LOWP_TREE_CASES = [("int128_f32_4000", 10, 150, 0), ("int128_f32_4000", 10, 700, 12),
                   ("gauss_d37_1500", 7, 150, 0), ("mnist_u8_3000", 16, 300, 0),
                   ("gauss960_800", 32, 150, 0)]


@pytest.mark.parametrize("prec,case,K,leaf_max,kprime",
                         [("bf16",) + c for c in LOWP_TREE_CASES] +
                         [("fp8",) + c for c in LOWP_TREE_CASES] +
                         [("fp8", "scaled_d64_2000", 10, 200, 0), ("fp8", "mnist_u8_3000", 10, 300, 64)])
def test_tree_knn_lowp_bitexact(ctx, orc, prec, case, K, leaf_max, kprime):
    """Per-tree BF16 / FP8 lists (bound pass + filter pass, exact re-rank)
    equal the oracle's exact per-leaf lists bit for bit."""
    _log(f"[tree_knn {prec}] {case} K={K} leaf_max={leaf_max} kprime={kprime}")
    X = CASES[case]()
    n, d = X.shape
    p = drpt_amd.leafmax_params(n, d, 2, K, leaf_max=leaf_max, precision=prec, kprime=kprime,
                                seed=3)
    p.dtype = drpt_amd.DRPT_U8 if X.dtype == np.uint8 else drpt_amd.DRPT_F32
    R = orc.build_R(d, p.ntrees * p.depth, p.density, p.seed)
    P0, cm0 = orc.project(X, R)
    heaps = np.stack([orc.grow_tree(P0, cm0, p.depth, t, 1, p.leaf_max, 0) for t in range(2)])
    ci, cd, tm = ctx.tree_knn(p, _dev(X), _dev(heaps.view(np.int32)))
    ci, cd = ci.cpu().numpy(), cd.cpu().numpy()
    _log(f"  fallback rows {tm.n_fallback_rows} of {2 * n}, largest leaf {tm.max_leaf}")
    sizes = np.concatenate([np.unique(heaps[t], return_counts=True)[1] for t in range(2)])
    assert (tm.n_leaves, tm.max_leaf) == (len(sizes), sizes.max())
    assert tm.gram_flops == 2.0 * d * float((sizes.astype(np.int64) ** 2).sum())
    for t in range(2):
        perm, offs = _groups(heaps[t])
        ri, rd = orc.leaf_knn(X, perm, offs, K)
        assert np.array_equal(ci[t], ri)
        assert np.array_equal(_bits(cd[t]), _bits(rd))


def test_tree_knn_many_heap_slots(ctx, orc):
    """Depth 18 with five trees: 2^18 heap slots per tree, so the leaf offsets
    and work items are scanned over 1280 chunks of 1024 slots (two windows of
    the chunk-total scan); lists and leaf statistics equal the oracle's."""
    _log("[tree_knn bf16] int128_f32_4000 depth 18, T=5")
    X = CASES["int128_f32_4000"]()
    n, d = X.shape
    T, K = 5, 10
    p = drpt_amd.leafmax_params(n, d, T, K, leaf_max=64, precision="bf16", seed=5, depth=18)
    p.dtype = drpt_amd.DRPT_F32
    R = orc.build_R(d, T * p.depth, p.density, p.seed)
    P0, cm0 = orc.project(X, R)
    heaps = np.stack([orc.grow_tree(P0, cm0, p.depth, t, 1, p.leaf_max, 0) for t in range(T)])
    ci, cd, tm = ctx.tree_knn(p, _dev(X), _dev(heaps.view(np.int32)))
    ci, cd = ci.cpu().numpy(), cd.cpu().numpy()
    sizes = np.concatenate([np.unique(heaps[t], return_counts=True)[1] for t in range(T)])
    assert (tm.n_leaves, tm.max_leaf) == (len(sizes), sizes.max())
    for t in range(T):
        perm, offs = _groups(heaps[t])
        ri, rd = orc.leaf_knn(X, perm, offs, K)
        assert np.array_equal(ci[t], ri)
        assert np.array_equal(_bits(cd[t]), _bits(rd))


LOWP_BUILD_CASES = [("int128_f32_4000", 4), ("int128_f32_4000", 7), ("gauss_d37_1500", 6),
                    ("mnist_u8_3000", 6), ("gauss960_800", 5)]


@pytest.mark.parametrize("prec,case,T", [("bf16",) + c for c in LOWP_BUILD_CASES] +
                         [("fp8",) + c for c in LOWP_BUILD_CASES] +
                         [("fp8", "scaled_d64_2000", 6), ("bf16", "int128x2_f32_4000", 6),
                          ("bf16", "int128_one256_f32_4000", 6), ("bf16", "gauss128_f32_3000", 6)])
def test_lowp_lists_equal_fp32(ctx, prec, case, T):
    """Whole build: BF16 / FP8 (trees beyond the first four pruned against the
    merged bound) returns the FP32 build's lists. The pool merge's exact
    distances read the u8 copy (int128: values 0-218), the bf16 copy (x2: even
    values to 436; one256: one value 256) or X (gauss128)."""
    _log(f"[{prec}==fp32] {case} T={T}")
    X = CASES[case]()
    n, d = X.shape
    out = {}
    for pr in ("fp32", prec):
        p = drpt_amd.leafmax_params(n, d, T, 10, leaf_max=200, precision=pr, seed=5)
        p.dtype = drpt_amd.DRPT_U8 if X.dtype == np.uint8 else drpt_amd.DRPT_F32
        oi, od, tm = ctx.build_graph(p, _dev(X))
        out[pr] = (oi.cpu().numpy(), od.cpu().numpy(), tm)
    assert np.array_equal(out["fp32"][0], out[prec][0])
    assert np.array_equal(_bits(out["fp32"][1]), _bits(out[prec][1]))
    tm = out[prec][2]
    _log(f"  fallback rows {tm.n_fallback_rows} of {n * T}")
    if prec == "bf16":  # fp8's wider error bound may send many rows to the exact path
        assert tm.n_fallback_rows <= n * T // 100


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
@pytest.mark.parametrize("copies,nbase,T,overflow,d", [(20, 20, 8, False, 16), (400, 5, 3, True, 16),
                                                       (400, 5, 6, True, 16), (40, 10, 8, False, 32),
                                                       (400, 5, 3, True, 32)])
def test_lowp_duplicates_overflow(ctx, orc, prec, copies, nbase, T, overflow, d):
    """Each vector repeated: ties at distance 0 fill the candidate pools (20
    copies over 8 trees: the same duplicates in every tree's pool segment,
    merged once; 400 copies: segments overflow and those rows are redone
    exactly over every tree); the lists still equal the oracle's. d = 32
    (integer values: the int8 pool pass, pools of (id, distance) pairs): 40
    copies over 8 trees put 312 entries in every pool, past the 256 the
    two-rows-per-wave merge takes, so those rows go to the full-width merge."""
    _log(f"[edge] {prec} duplicates copies={copies} T={T} d={d}")
    rng = np.random.default_rng(4)
    base = rng.integers(0, 50, size=(nbase, d)).astype(np.float32)
    X = np.repeat(base, copies, axis=0)
    n, d = X.shape
    p = drpt_amd.leafmax_params(n, d, T, 10, leaf_max=300, precision=prec, seed=8)
    oi, od, tm = ctx.build_graph(p, _dev(X))
    ri, rd, _, _ = orc.build_graph(X, T, p.depth, 10, local_mode=1, leaf_max=300, seed=8,
                                   density=p.density)
    _log(f"  overflow rows {tm.n_fallback_rows}")
    assert (tm.n_fallback_rows > 0) == overflow
    assert np.array_equal(oi.cpu().numpy(), ri)
    assert np.array_equal(_bits(od.cpu().numpy()), _bits(rd))


def test_merge_bitexact(ctx, orc):
    _log("[merge]")
    rng = np.random.default_rng(3)
    T, n, K = 6, 500, 10
    ci = np.full((T, n, K), -1, np.int32)
    cd = np.full((T, n, K), np.inf, np.float32)
    base = rng.integers(0, 50, size=(n, n)).astype(np.float32)
    base = np.minimum(base, base.T)
    for t in range(T):
        for p in range(n):
            m = int(rng.integers(0, K + 1))
            q = rng.choice(np.delete(np.arange(n), p), size=m, replace=False)
            order = sorted(zip(base[p, q], q))
            for k, (dv, qq) in enumerate(order):
                ci[t, p, k], cd[t, p, k] = qq, dv
    oi, od = ctx.merge(_dev(ci), _dev(cd))
    ri, rd = orc.merge(ci, cd)
    assert np.array_equal(oi.cpu().numpy(), ri)
    assert np.array_equal(_bits(od.cpu().numpy()), _bits(rd))


@pytest.mark.parametrize("case,mode", [("mnist_u8_3000", 0), ("int128_f32_4000", 1),
                                       ("gauss_d37_1500", 1)])
def test_build_graph_vs_oracle(ctx, orc, case, mode):
    _log(f"[build_graph] {case} mode={mode}")
    X = CASES[case]()
    n, d = X.shape
    T, nn = 4, 10
    if mode == 0:
        p = drpt_amd.derive_params(n, d, T, nn)
        p.seed = 11
        depth, leaf_max = p.depth, 0
    else:
        p = drpt_amd.leafmax_params(n, d, T, nn, leaf_max=120, seed=11)
        depth, leaf_max = p.depth, p.leaf_max
    p.dtype = drpt_amd.DRPT_U8 if X.dtype == np.uint8 else drpt_amd.DRPT_F32
    oi, od, tm = ctx.build_graph(p, _dev(X))
    ri, rd, _, _ = orc.build_graph(X, T, depth, nn, local_mode=mode, leaf_max=leaf_max, seed=11,
                                   density=p.density)
    assert np.array_equal(oi.cpu().numpy(), ri)
    assert np.array_equal(_bits(od.cpu().numpy()), _bits(rd))


@pytest.mark.parametrize("case,T,prec", [("mnist_u8_3000", 4, "fp32"), ("int128_f32_4000", 3, "fp32"),
                                         ("mnist_u8_3000", 5, "bf16"), ("gauss_d37_1500", 4, "fp8")])
def test_build_graph_real_local_tree_vs_oracle(ctx, orc, case, T, prec):
    """-density 1 (an integer density, main.cpp:69-72): dense R and a dense
    local matrix per tree, so the reference local tree splits on real keys
    (drpt_local.cpp:121-161); the GPU's segmented (value, index) sorts give the
    oracle's local leaves and lists bit for bit."""
    _log(f"[build_graph real local] {case} T={T} {prec}")
    X = CASES[case]()
    n, d = X.shape
    p = drpt_amd.derive_params(n, d, T, 10)
    p.seed = 13
    p.density = 1.0
    p.precision = drpt_amd.PRECISIONS[prec]
    p.dtype = drpt_amd.DRPT_U8 if X.dtype == np.uint8 else drpt_amd.DRPT_F32
    oi, od, tm = ctx.build_graph(p, _dev(X))
    ri, rd, _, _ = orc.build_graph(X, T, p.depth, 10, local_mode=0, seed=13, density=1.0)
    assert np.array_equal(oi.cpu().numpy(), ri)
    assert np.array_equal(_bits(od.cpu().numpy()), _bits(rd))
    r0, _, _, _ = orc.build_graph(X, T, p.depth, 10, local_mode=0, seed=13,
                                  density=1.0 / np.sqrt(d))
    assert not np.array_equal(ri, r0)  # the local tree really changed


def test_duplicates_and_tiny_leaves(ctx, orc):
    _log("[edge] duplicates / tiny leaves")
    rng = np.random.default_rng(9)
    base = rng.integers(0, 4, size=(60, 8)).astype(np.float32)
    X = np.repeat(base, 5, axis=0)  # every vector five times: ties everywhere
    n, d = X.shape
    p = drpt_amd.leafmax_params(n, d, 3, 12, leaf_max=9, seed=2)
    oi, od, tm = ctx.build_graph(p, _dev(X))
    ri, rd, _, _ = orc.build_graph(X, 3, p.depth, 12, local_mode=1, leaf_max=9, seed=2,
                                   density=p.density)
    assert np.array_equal(oi.cpu().numpy(), ri)
    assert np.array_equal(_bits(od.cpu().numpy()), _bits(rd))


@pytest.mark.parametrize("prec,case,leaf_max,T", [("bf16", "int128_f32_4000", 700, 5),
                                                  ("fp8", "gauss_d37_1500", 100, 4),
                                                  ("bf16", "mnist_u8_3000", 300, 4)])
def test_autotune_and_item_plans(ctx, prec, case, leaf_max, T):
    """drpt_autotune returns a valid plan; builds with the tuned plan and with
    forced plans (leaves cut into 64-row items, register pipeline, mixed
    pipelines across buckets: one launch per bucket; 2 / 3 trees per pooled
    pass; first bounds from approximate distances) equal the default build
    bit for bit."""
    _log(f"[autotune] {prec} {case} leaf_max={leaf_max}")
    X = CASES[case]()
    n, d = X.shape
    p = drpt_amd.leafmax_params(n, d, T, 10, leaf_max=leaf_max, precision=prec, seed=6)
    p.dtype = drpt_amd.DRPT_U8 if X.dtype == np.uint8 else drpt_amd.DRPT_F32
    Xd = _dev(X)
    ref = drpt_amd.Context(0)
    with ref:
        ri, rd, _ = ref.build_graph(p, Xd)
        ri, rd = ri.cpu().numpy(), rd.cpu().numpy()
    with drpt_amd.Context(0) as c:
        plan = c.autotune(p, Xd)
        _log(f"  plan {plan}")
        assert all(r in (64, 128, 256, 1 << 30) for r in plan["rows_per_item"])
        assert all(g in (0, 1, 2) for g in plan["ring"])
        assert 1 <= plan["tree_group"] <= T
        assert plan["first_trees"] in (1, 2, 4)
        assert 10 <= plan["rerank"] <= 20 or plan["rerank"] == -1
        forced = [plan,
                  {"rows_per_item": [64] * 4, "ring": [0] * 4, "tree_group": 2, "first_trees": 1,
                   "rerank": 10},
                  {"rows_per_item": [64, 128, 256, 64], "ring": [1, 0, 1, 0], "tree_group": 3,
                   "first_trees": 2},
                  {"rows_per_item": [1 << 30] * 4, "ring": [0, 1, 0, 1], "tree_group": 0},
                  {"rows_per_item": [1 << 30, 64, 1 << 30, 128], "ring": [2, 2, 1, 2],
                   "tree_group": 2},
                  {"rows_per_item": [1 << 30] * 4, "ring": [2] * 4, "tree_group": 2,
                   "rerank": -1}]
        for pl in forced:
            c.set_item_plan(p, pl)
            oi, od, _ = c.build_graph(p, Xd)
            assert np.array_equal(oi.cpu().numpy(), ri), f"plan {pl}"
            assert np.array_equal(_bits(od.cpu().numpy()), _bits(rd)), f"plan {pl}"


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_lowp_k32_dense_clusters_equal_fp32(ctx, prec):
    """nn = 32 on dense clusters (about 4000 points per centre, as the 10M x 96
    workload): many survivors per row and tile, so per-thread lists are flushed
    between tiles (512-slot pools, known ids left to the row merge, several
    tree groups); the lists still equal the FP32 build's."""
    _log(f"[k32] {prec} dense clusters")
    n = 200_000
    X = datagen.sphere_mixture(n, centres=50)
    Xd = _dev(X)
    out = {}
    for pr in ("fp32", prec):
        p = drpt_amd.leafmax_params(n, 96, 8, 32, leaf_max=1024, precision=pr,
                                    kprime=64 if pr != "fp32" else 0, seed=2)
        oi, od, tm = ctx.build_graph(p, Xd)
        out[pr] = (oi.cpu().numpy(), od.cpu().numpy(), tm)
    _log(f"  overflow rows {out[prec][2].n_fallback_rows}")
    assert np.array_equal(out["fp32"][0], out[prec][0])
    assert np.array_equal(_bits(out["fp32"][1]), _bits(out[prec][1]))


DRPT_ERR_NONFINITE = -6  # include/drpt_hip.h


@pytest.mark.parametrize("bad", [np.nan, np.inf])
@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_build_nonfinite_input_refused(ctx, prec, bad):
    """A non-finite value in X makes the projection non-finite: the build
    returns DRPT_ERR_NONFINITE (the splits have run on the device meanwhile),
    and the next build of finite data on the same context is unaffected."""
    _log(f"[nonfinite] {prec} {bad}")
    X = datagen.int_mixture(3000)
    n, d = X.shape
    p = drpt_amd.leafmax_params(n, d, 4, 10, leaf_max=200, precision=prec, seed=5)
    Xb = X.copy()
    Xb[1234, 7] = bad
    with pytest.raises(drpt_amd.DrptError) as ei:
        ctx.build_graph(p, _dev(Xb))
    assert ei.value.code == DRPT_ERR_NONFINITE
    assert "non-finite" in str(ei.value)
    oi, od, _ = ctx.build_graph(p, _dev(X))
    q = drpt_amd.leafmax_params(n, d, 4, 10, leaf_max=200, precision=prec, seed=5)
    with drpt_amd.Context(0) as fresh:
        ri, rd, _ = fresh.build_graph(q, _dev(X))
    assert np.array_equal(oi.cpu().numpy(), ri.cpu().numpy())
    assert np.array_equal(_bits(od.cpu().numpy()), _bits(rd.cpu().numpy()))


@pytest.mark.parametrize("prec,K", [("bf16", 10), ("bf16", 24)])
def test_lowp_leaves_beyond_stage_equal_fp32(ctx, prec, K):
    """Leaves larger than the 2048 points a block stages: their rows take the
    exact row path (leafmates sorted in 2048-point rounds); the lists still
    equal the FP32 build's."""
    _log(f"[big leaves] {prec} K={K}")
    X = datagen.int_mixture(9000)
    n, d = X.shape
    out = {}
    for pr in ("fp32", prec):
        p = drpt_amd.leafmax_params(n, d, 3, K, leaf_max=4500, precision=pr, seed=9)
        oi, od, tm = ctx.build_graph(p, _dev(X))
        out[pr] = (oi.cpu().numpy(), od.cpu().numpy(), tm)
    _log(f"  largest leaf {out[prec][2].max_leaf}, exact rows {out[prec][2].n_fallback_rows}")
    assert out[prec][2].max_leaf > 2048
    assert np.array_equal(out["fp32"][0], out[prec][0])
    assert np.array_equal(_bits(out["fp32"][1]), _bits(out[prec][1]))
