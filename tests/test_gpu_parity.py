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
    "gauss_d37_1500": lambda: np.random.default_rng(5).standard_normal((1500, 37)).astype(np.float32),
    "gauss960_800": lambda: datagen.gauss_mixture(800),
}


@pytest.mark.parametrize("case", list(CASES))
def test_project_bitexact(ctx, orc, case):
    _log(f"[project] {case}")
    X = CASES[case]()
    n, d = X.shape
    R = orc.build_R(d, 24, 1.0 / np.sqrt(d), 17)
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


@pytest.mark.parametrize("case,K,leaf_max,kprime", [("int128_f32_4000", 10, 150, 0),
                                                    ("int128_f32_4000", 10, 700, 12),
                                                    ("gauss_d37_1500", 7, 150, 0),
                                                    ("mnist_u8_3000", 16, 300, 0),
                                                    ("gauss960_800", 32, 150, 0)])
def test_tree_knn_bf16_bitexact(ctx, orc, case, K, leaf_max, kprime):
    """Per-tree BF16 lists (bound pass + filter pass, exact re-rank) equal the
    oracle's exact per-leaf lists bit for bit."""
    _log(f"[tree_knn bf16] {case} K={K} leaf_max={leaf_max} kprime={kprime}")
    X = CASES[case]()
    n, d = X.shape
    p = drpt_amd.leafmax_params(n, d, 2, K, leaf_max=leaf_max, precision="bf16", kprime=kprime,
                                seed=3)
    p.dtype = drpt_amd.DRPT_U8 if X.dtype == np.uint8 else drpt_amd.DRPT_F32
    R = orc.build_R(d, p.ntrees * p.depth, p.density, p.seed)
    P0, cm0 = orc.project(X, R)
    heaps = np.stack([orc.grow_tree(P0, cm0, p.depth, t, 1, p.leaf_max, 0) for t in range(2)])
    ci, cd, tm = ctx.tree_knn(p, _dev(X), _dev(heaps.view(np.int32)))
    ci, cd = ci.cpu().numpy(), cd.cpu().numpy()
    _log(f"  fallback rows {tm.n_fallback_rows} of {2 * n}, largest leaf {tm.max_leaf}")
    for t in range(2):
        perm, offs = _groups(heaps[t])
        ri, rd = orc.leaf_knn(X, perm, offs, K)
        assert np.array_equal(ci[t], ri)
        assert np.array_equal(_bits(cd[t]), _bits(rd))


@pytest.mark.parametrize("case,T", [("int128_f32_4000", 4), ("int128_f32_4000", 7),
                                    ("gauss_d37_1500", 6), ("mnist_u8_3000", 6),
                                    ("gauss960_800", 5)])
def test_bf16_lists_equal_fp32(ctx, case, T):
    """Whole build: BF16 (trees beyond the first four pruned against the merged
    bound) returns the FP32 build's lists."""
    _log(f"[bf16==fp32] {case} T={T}")
    X = CASES[case]()
    n, d = X.shape
    out = {}
    for prec in ("fp32", "bf16"):
        p = drpt_amd.leafmax_params(n, d, T, 10, leaf_max=200, precision=prec, seed=5)
        p.dtype = drpt_amd.DRPT_U8 if X.dtype == np.uint8 else drpt_amd.DRPT_F32
        oi, od, tm = ctx.build_graph(p, _dev(X))
        out[prec] = (oi.cpu().numpy(), od.cpu().numpy(), tm)
    assert np.array_equal(out["fp32"][0], out["bf16"][0])
    assert np.array_equal(_bits(out["fp32"][1]), _bits(out["bf16"][1]))
    tm = out["bf16"][2]
    _log(f"  fallback rows {tm.n_fallback_rows} of {n * T}")
    assert tm.n_fallback_rows <= n * T // 100


@pytest.mark.parametrize("copies,nbase,T,overflow", [(20, 20, 8, False), (400, 5, 3, True),
                                                     (400, 5, 6, True)])
def test_bf16_duplicates_overflow(ctx, orc, copies, nbase, T, overflow):
    """Each vector repeated: ties at distance 0 fill the candidate pools (20
    copies over 8 trees: the same duplicates in every tree's pool segment,
    merged once; 400 copies: segments overflow and those rows are redone
    exactly over every tree); the lists still equal the oracle's."""
    _log(f"[edge] bf16 duplicates copies={copies} T={T}")
    rng = np.random.default_rng(4)
    base = rng.integers(0, 50, size=(nbase, 16)).astype(np.float32)
    X = np.repeat(base, copies, axis=0)
    n, d = X.shape
    p = drpt_amd.leafmax_params(n, d, T, 10, leaf_max=300, precision="bf16", seed=8)
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
