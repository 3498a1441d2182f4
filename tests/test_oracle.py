"""CPU tests pinning the oracle (tests/refpy.py restatements, exact arithmetic,
brute force and known answers). No GPU needed."""
import math
from fractions import Fraction

import numpy as np
import pytest

import refpy


def _bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(both_nan | (a.view(np.uint32) == b.view(np.uint32))))


# ---- R (math_operations.cpp:48-76) -----------------------------------------
def test_mt19937_known_answer():
    g = refpy.StdMt19937(5489)
    for _ in range(9999):
        g()
    assert g() == 4123659995  # C++ standard [rand.predef]


@pytest.mark.parametrize("d,L,density,seed", [(20, 12, 0.3, 77), (784, 3, 1 / 28, 5),
                                              (16, 8, 1.0, 123456789), (9, 7, 0.0, 4)])
def test_build_R_matches_libstdcxx_restatement(orc, d, L, density, seed):
    R1 = refpy.build_R(d, L, density, seed)
    R2 = orc.build_R(d, L, density, seed)
    assert _bits_equal(R1, R2)


def test_R_density_statistics(orc):
    R = orc.build_R(784, 56, 1 / 28, 99)
    frac = (R != 0).mean()
    assert abs(frac - 1 / 28) < 0.005
    nz = R[R != 0]
    assert abs(nz.mean()) < 0.1 and abs(nz.std() - 1.0) < 0.1


# ---- projection (math_operations.cpp:15-46) ---------------------------------
def test_project_order_fixed_double_sum(orc):
    rng = np.random.default_rng(1)
    X = rng.integers(0, 256, size=(50, 40)).astype(np.uint8)
    R = orc.build_R(40, 6, 0.3, 11)
    P, cm = orc.project(X, R)
    for c in range(6):
        nz = np.nonzero(R[:, c])[0]
        for i in range(50):
            acc = 0.0
            for j in nz:
                acc = acc + float(X[i, j]) * float(R[j, c])
            assert np.float32(acc).view(np.uint32) == P[c, i].view(np.uint32)
    assert np.array_equal(cm, np.abs(P).max(axis=1).view(np.uint32))


def test_project_close_to_exact_rational(orc):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((30, 24)).astype(np.float32)
    R = orc.build_R(24, 5, 0.5, 3)
    P, _ = orc.project(X, R)
    for c in range(5):
        for i in range(30):
            ex = sum(Fraction(float(X[i, j])) * Fraction(float(R[j, c])) for j in range(24))
            ref = np.float32(float(ex))
            assert abs(float(P[c, i]) - float(ref)) <= abs(np.spacing(ref))


def test_project_linear_in_X(orc):
    rng = np.random.default_rng(3)
    A = rng.integers(0, 100, size=(64, 32)).astype(np.float32)
    B = rng.integers(0, 100, size=(64, 32)).astype(np.float32)
    R = orc.build_R(32, 8, 0.25, 8)
    PA, _ = orc.project(A, R)
    PB, _ = orc.project(B, R)
    PC, _ = orc.project(A + B, R)
    assert np.allclose(PA + PB, PC, rtol=1e-6, atol=1e-4)


# ---- node split (math_operations.cpp:121-383) --------------------------------
def _random_nodes(seed, count):
    rng = np.random.default_rng(seed)
    for trial in range(count):
        n = int(rng.integers(2, 80))
        v = (rng.standard_normal(n) * rng.uniform(0.1, 100) + rng.uniform(-50, 50)).astype(np.float32)
        if trial % 7 == 0:
            v[:] = v[0]
        if trial % 11 == 0:
            v = np.round(v).astype(np.float32)
        if trial % 13 == 0:
            v = np.abs(v)
        yield v


@pytest.mark.parametrize("stat_mode", [0, 1])
def test_node_split_matches_restatement(orc, stat_mode):
    for v in _random_nodes(7 + stat_mode, 250):
        cm = int(np.abs(v).max().view(np.uint32))
        kq1, kq2 = refpy.quant_exponents(cm, 5000)
        a = refpy.node_split(v, stat_mode, kq1, kq2)
        b = orc.node_split(v, stat_mode, kq1, kq2)
        for k in ("median", "mu", "var"):
            assert _bits_equal(a[k], b[k]), (k, v)
        assert _bits_equal(a["E"], b["E"])
        assert np.array_equal(a["H"], b["H"])


def test_exact_mode_statistics_close_to_rational(orc):
    for v in _random_nodes(21, 100):
        n = len(v)
        cm = int(np.abs(v).max().view(np.uint32))
        kq1, kq2 = orc.quant_exponents(cm, 5000)
        r = orc.node_split(v, 0, kq1, kq2)
        mean = sum(Fraction(float(x)) for x in v) / n
        assert abs(float(r["mu"]) - float(mean)) <= 2 * abs(np.spacing(np.float32(float(mean)))) + 1e-30
        var = sum((Fraction(float(x)) - Fraction(float(r["mu"]))) ** 2 for x in v) / (n - 1)
        assert abs(float(r["var"]) - float(var)) <= 1e-5 * float(var) + 1e-25


def test_quant_exponents_headroom(orc):
    for N in (2, 1000, 60000, 1_000_000, 10_000_000):
        for m in (1e-20, 0.5, 3.0, 1.5e4, 7e30):
            bits = int(np.float32(m).view(np.uint32))
            kq1, kq2 = orc.quant_exponents(bits, N)
            assert (kq1, kq2) == refpy.quant_exponents(bits, N)
            # N terms of magnitude <= 2^shift fit in int64
            assert N * (abs(m) / 2.0 ** kq1) < 2 ** 63
            assert N * ((2 * m) ** 2 / 2.0 ** kq2) < 2 ** 63


def test_histogram_both_tails_in_slot0(orc):
    rng = np.random.default_rng(4)
    v = np.concatenate([rng.standard_normal(400), [-12.0, 12.0, 13.0]]).astype(np.float32)
    r = orc.node_split(v, 1)
    E, H = r["E"], r["H"]
    assert H.sum() == len(v)
    below = int((v <= E[0]).sum())
    above = int((v > E[13]).sum())
    assert below >= 1 and above >= 2
    assert H[0] == below + above  # both tails share slot 0
    assert H[14] == 0 and H[15] == 0


def test_single_point_node_goes_right(orc):
    r = orc.node_split(np.array([5.0], np.float32), 1)
    assert np.isnan(r["median"])  # var = 0/0, v <= NaN is false -> right child


def test_constant_node_median_uses_zero_left_edge(orc):
    # sigma = 0: every value falls in slot 0 and E[-1] := 0, so median = 0 + f*(mu - 0)
    v = np.full(10, 8.0, np.float32)
    r = orc.node_split(v, 1)
    assert r["H"][0] == 10
    assert float(r["median"]) == pytest.approx(8.0 * (5 - 0) / 10)


# ---- tree growth (drpt_global.cpp:153-352) ------------------------------------
@pytest.mark.parametrize("mode,stat_mode,leaf_max", [(0, 0, 0), (0, 1, 0), (1, 0, 40), (1, 1, 25)])
def test_grow_tree_matches_restatement(orc, mode, stat_mode, leaf_max):
    rng = np.random.default_rng(5)
    X = rng.integers(0, 50, size=(300, 16)).astype(np.float32)
    D = 6
    R = orc.build_R(16, 2 * D, 0.5, 19)
    P, cm = orc.project(X, R)
    for t in range(2):
        a = refpy.grow_tree(P, cm, D, t, mode, leaf_max, stat_mode)
        b = orc.grow_tree(P, cm, D, t, mode, leaf_max, stat_mode)
        assert np.array_equal(a, b)


def test_grow_tree_invariants(orc):
    rng = np.random.default_rng(6)
    X = rng.standard_normal((5000, 24)).astype(np.float32)
    D = 8
    R = orc.build_R(24, 3 * D, 0.2, 23)
    P, cm = orc.project(X, R)
    for t in range(3):
        h = orc.grow_tree(P, cm, D, t, 0, 0, 0)
        assert h.min() >= 2 ** (D - 1) and h.max() < 2 ** D
        sizes = np.bincount(h - 2 ** (D - 1), minlength=2 ** (D - 1))
        assert sizes.sum() == 5000
        # approximate-median splits keep leaves within a factor of the mean
        assert sizes.min() > 0 and sizes.max() < 4 * 5000 / 2 ** (D - 1)
        hl = orc.grow_tree(P, cm, D, t, 1, 100, 0)
        for leaf in np.unique(hl):
            depth = int(leaf).bit_length() - 1
            size = int((hl == leaf).sum())
            assert size <= 100 or depth == D - 1


# ---- local stage (drpt_local.cpp:49-180) --------------------------------------
def _halving(m, ld):
    if ld == 0:
        return [m]
    return _halving(m - m // 2, ld - 1) + _halving(m // 2, ld - 1)


@pytest.mark.parametrize("m,ld", [(1, 0), (8, 3), (117, 3), (938, 3), (1000, 4), (4097, 5)])
def test_local_sizes_and_permutation(orc, m, ld):
    assert list(orc.local_sizes(m, ld)) == _halving(m, ld)
    p = orc.local_perm(m, ld)
    assert sorted(p.tolist()) == list(range(m))
    assert np.array_equal(p, orc.local_perm(m, ld))


def test_local_perm_known_answer(orc):
    # a 8-point leaf split 3 levels deep on all-equal keys, as the reference's
    # DRPTLocal arranges it with GCC 11.4's libstdc++
    assert orc.local_perm(8, 3).tolist() == [6, 5, 4, 7, 0, 3, 2, 1]
    assert refpy.local_positions(8, 3) == [6, 5, 4, 7, 0, 3, 2, 1]


def _local_cases():
    rng = np.random.default_rng(14)
    ms = list(range(1, 400)) + sorted(rng.integers(400, 5001, size=60).tolist()) + [938, 1953, 4883]
    for m in ms:
        for ld in range(1, 6):
            if (1 << ld) <= m:
                yield m, ld


def test_local_perm_matches_introselect_restatement(orc):
    """orc_local_perm (the libstdc++ call) against refpy's step-by-step
    restatement of libstdc++'s introselect, for every leaf size below 400 and
    a sample up to 5000, local depths 1-5."""
    for m, ld in _local_cases():
        assert orc.local_perm(m, ld).tolist() == refpy.local_positions(m, ld), (m, ld)


def test_nth_element_restatement_selects():
    """The restatement is a selection on real keys too: the nth element is in
    place, smaller ones before it, larger after (a plausible slip in the
    partition or the median-of-three breaks this)."""
    rng = np.random.default_rng(15)
    for _ in range(400):
        m = int(rng.integers(1, 300))
        v = rng.integers(0, 25, size=m).tolist()
        a = list(range(m))
        k = int(rng.integers(0, m))
        refpy.nth_element(a, 0, k, m, lambda x, y: v[x] < v[y])
        assert sorted(a) == list(range(m))
        assert v[a[k]] == sorted(v)[k]
        assert all(v[a[i]] <= v[a[k]] for i in range(k))
        assert all(v[a[i]] >= v[a[k]] for i in range(k + 1, m))


@pytest.mark.parametrize("m,ld,levels", [(7, 2, 3), (64, 3, 3), (117, 3, 4), (500, 4, 4),
                                         (1000, 5, 5)])
def test_local_split_real_keys_matches_reference_recursion(orc, m, ld, levels):
    """Real-key local tree (-density >= 1): the oracle's leaves equal the sets
    the reference recursion (drpt_local.cpp:121-161: nth_element at size/2,
    left = [begin, end - size/2), left first) gives with the restated
    introselect, keys ordered by (value, point index); values drawn from a few
    integers so that ties are frequent."""
    rng = np.random.default_rng(m + ld)
    n = m + 50
    LP = rng.integers(0, levels, size=(ld, n)).astype(np.float32)
    pts = np.sort(rng.choice(n, size=m, replace=False)).astype(np.int32)
    got = orc.local_split(LP, pts, ld)
    a = pts.tolist()

    def rec(b, e, depth):
        if depth == ld:
            return
        size = e - b
        less = lambda x, y: (LP[depth, x], x) < (LP[depth, y], y)
        refpy.nth_element(a, b, b + size // 2, e, less)
        mid = e - size // 2
        rec(b, mid, depth + 1)
        rec(mid, e, depth + 1)

    rec(0, m, 0)
    sizes = orc.local_sizes(m, ld)
    o = 0
    for sz in sizes:
        assert sorted(got[o:o + sz].tolist()) == sorted(a[o:o + sz])
        o += sz
    # brute force: every left child holds the ceil(s/2) smallest (value, index) keys
    def check(seg, depth):
        if depth == ld:
            return
        keys = sorted((float(LP[depth, p]), int(p)) for p in seg)
        left = {p for _, p in keys[:len(seg) - len(seg) // 2]}
        first = len(seg) - len(seg) // 2
        assert set(seg[:first]) == left
        check(seg[:first], depth + 1)
        check(seg[first:], depth + 1)

    check(got.tolist(), 0)


def test_build_graph_real_local_tree(orc):
    """-density 1: dense R and dense local C; the build's lists equal the top-K
    of each point's local-leafmate union (brute force over the leaves the
    oracle reports), and differ from the tie-permutation build."""
    rng = np.random.default_rng(16)
    X = rng.integers(0, 40, size=(3000, 12)).astype(np.uint8)
    D = orc.global_depth(orc.tree_depth(3000))
    oi, od, _, _ = orc.build_graph(X, 3, D, 5, local_mode=0, seed=4, density=1.0)
    oi0, _, _, _ = orc.build_graph(X, 3, D, 5, local_mode=0, seed=4, density=1.0 / np.sqrt(12))
    assert (oi >= 0).all() and (oi != np.arange(3000)[:, None]).all()
    assert np.all(np.diff(od, axis=1) >= 0)
    for p in range(0, 3000, 97):
        d2 = sorted(_exact_int_d2(X, p, q) for q in oi[p])
        assert d2 == sorted(int(v) for v in od[p])
    assert not np.array_equal(oi, oi0)


def test_local_C_is_zero_for_truncated_density(orc):
    # (int)(1/sqrt(d)) == 0: the local projection keeps an entry only for a zero draw
    assert orc.local_C_nonzeros(784, 3, 0.0, 12345) == 0


# ---- leaf k-NN and merge (mdrpt.cpp:123-212) -----------------------------------
def _exact_int_d2(X, i, j):
    diff = X[i].astype(np.int64) - X[j].astype(np.int64)
    return int((diff * diff).sum())


def test_leaf_knn_bruteforce_integer(orc):
    rng = np.random.default_rng(8)
    X = rng.integers(0, 30, size=(200, 12)).astype(np.uint8)
    perm = rng.permutation(200).astype(np.int32)
    offs = np.array([0, 3, 60, 61, 140, 200], np.int64)
    K = 7
    ci, cd = orc.leaf_knn(X, perm, offs, K)
    for lf in range(len(offs) - 1):
        pts = perm[offs[lf]:offs[lf + 1]]
        for p in pts:
            cand = sorted((_exact_int_d2(X, p, q), int(q)) for q in pts if q != p)[:K]
            got = [(int(dv), int(i)) for dv, i in zip(cd[p], ci[p]) if i >= 0]
            assert got == cand
            assert np.all(ci[p][len(cand):] == -1)


def test_dist_fma_chain_exact_rational(orc):
    rng = np.random.default_rng(9)
    X = rng.standard_normal((6, 40)).astype(np.float32)
    for i in range(6):
        for j in range(6):
            acc = np.float32(0.0)
            for k in range(40):
                t = np.float32(X[i, k] - X[j, k])
                acc = np.float32(float(Fraction(float(t)) ** 2 + Fraction(float(acc))))
            assert orc.dist(X, i, j).view(np.uint32) == acc.view(np.uint32)


def test_merge_is_topk_of_union(orc):
    rng = np.random.default_rng(10)
    T, n, K = 5, 40, 6
    ci = np.full((T, n, K), -1, np.int32)
    cd = np.full((T, n, K), np.inf, np.float32)
    dist = {}
    for t in range(T):
        for p in range(n):
            others = [q for q in rng.choice(n, size=K + 2, replace=False) if q != p][:K - t % 2]
            lst = []
            for q in others:
                key = (min(p, q), max(p, q))
                dist.setdefault(key, float(rng.integers(0, 20)))
                lst.append((dist[key], int(q)))
            lst.sort()
            for k, (dv, q) in enumerate(lst):
                ci[t, p, k], cd[t, p, k] = q, dv
    oi, od = orc.merge(ci, cd)
    for p in range(n):
        union = {(float(cd[t, p, k]), int(ci[t, p, k])) for t in range(T) for k in range(K)
                 if ci[t, p, k] >= 0}
        want = sorted(union)[:K]
        got = [(float(dv), int(i)) for dv, i in zip(od[p], oi[p]) if i >= 0]
        assert got == want


def test_bruteforce_f64_integer_exact(orc):
    rng = np.random.default_rng(11)
    X = rng.integers(0, 256, size=(300, 20)).astype(np.uint8)
    q = np.array([0, 5, 299], np.int64)
    oi, od = orc.bruteforce_f64(X, q, 10)
    for a, qq in enumerate(q):
        ref = sorted((_exact_int_d2(X, qq, j), j) for j in range(300) if j != qq)[:10]
        assert [(int(d), int(i)) for d, i in zip(od[a], oi[a])] == ref


# ---- parameters (main.cpp:141-145, mdrpt.cpp:52, :324) --------------------------
@pytest.mark.parametrize("n,depth,dg", [(60000, 13, 7), (1_000_000, 17, 10), (10_000_000, 21, 12)])
def test_parameter_derivation(orc, n, depth, dg):
    assert orc.tree_depth(n) == depth
    assert orc.global_depth(depth, 0.6) == dg


def test_local_depth_rule(orc):
    assert orc.local_depth(500, 10, 2) == 3      # int(8.97 - 5.32)
    assert orc.local_depth(79, 10, 2) == 0       # throws in the reference (ld <= 0)
    assert orc.local_depth(256, 32, 2) == 1


# ---- whole build ------------------------------------------------------------
def test_build_graph_equals_topk_of_leafmate_union(orc):
    rng = np.random.default_rng(12)
    X = rng.integers(0, 40, size=(3000, 10)).astype(np.uint8)
    T, nn = 4, 5
    D = orc.global_depth(orc.tree_depth(3000))
    oi, od, tm, leaf = orc.build_graph(X, T, D, nn, local_mode=1, leaf_max=60, seed=3,
                                       want_leaves=True)
    for p in rng.choice(3000, size=40, replace=False):
        union = set()
        for t in range(T):
            union |= set(np.nonzero(leaf[t] == leaf[t, p])[0].tolist())
        union.discard(int(p))
        want = sorted((_exact_int_d2(X, p, q), q) for q in union)[:nn]
        got = [(int(d), int(i)) for d, i in zip(od[p], oi[p]) if i >= 0]
        assert got == want


def test_build_graph_reference_local_mode_runs(orc):
    rng = np.random.default_rng(13)
    X = rng.integers(0, 256, size=(4000, 32)).astype(np.uint8)
    D = orc.global_depth(orc.tree_depth(4000))
    oi, od, tm, _ = orc.build_graph(X, 3, D, 10, local_mode=0, seed=9)
    assert (oi >= 0).all()
    assert (oi != np.arange(4000)[:, None]).all()
    assert np.all(np.diff(od, axis=1) >= 0)
