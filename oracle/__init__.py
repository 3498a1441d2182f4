"""ORACLE — test infrastructure only (not part of the product).

ctypes bindings to ``liboracle_drpt.so``, the CPU restatement of the
HipGraph/DRPT k-NN graph construction path (see ``drpt_oracle.cpp`` for the
reference file:line each function follows and what pins it). Only tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this
module: as the checker, and as the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_drpt.so")
_lib = None


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32, i64, f32, u32 = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64,
                                  ctypes.c_float, ctypes.c_uint32)
        L.orc_build_R.argtypes = [i32, i32, f32, u32, vp]
        L.orc_project.argtypes = [vp, i32, i64, i32, vp, i32, vp, vp]
        L.orc_quant_exponents.argtypes = [u32, i64, vp, vp]
        L.orc_node_split.argtypes = [vp, i64, i32, i32, i32, vp, vp, vp, vp]
        L.orc_node_split.restype = f32
        L.orc_grow_tree.argtypes = [vp, i64, i32, i32, i32, i32, i32, vp, vp, vp]
        L.orc_local_perm.argtypes = [i32, i32, vp]
        L.orc_local_sizes.argtypes = [i32, i32, vp]
        L.orc_local_split.argtypes = [vp, i64, i32, vp, i32]
        L.orc_local_C_nonzeros.argtypes = [i32, i32, f32, u32]
        L.orc_dist.argtypes = [vp, i32, i32, i64, i64]
        L.orc_dist.restype = f32
        L.orc_leaf_knn.argtypes = [vp, i32, i64, i32, vp, vp, i64, i32, vp, vp]
        L.orc_merge.argtypes = [i64, i32, i32, vp, vp, vp, vp]
        L.orc_bruteforce_f64.argtypes = [vp, i32, i64, i32, vp, i64, i32, vp, vp]
        L.orc_tree_depth.argtypes = [i64]
        L.orc_global_depth.argtypes = [i32, ctypes.c_double]
        L.orc_local_depth.argtypes = [i64, i32, i32]
        L.orc_build_graph.argtypes = [vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Params(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("d", ctypes.c_int), ("x_is_u8", ctypes.c_int),
                ("ntrees", ctypes.c_int), ("depth", ctypes.c_int), ("nn", ctypes.c_int),
                ("local_mode", ctypes.c_int), ("leaf_max", ctypes.c_int),
                ("local_offset", ctypes.c_int), ("stat_mode", ctypes.c_int),
                ("seed", ctypes.c_uint32), ("density", ctypes.c_float),
                ("nthreads", ctypes.c_int), ("pad", ctypes.c_int)]


def build_R(d, L, density, seed):
    R = np.zeros((d, L), np.float32)
    lib().orc_build_R(d, L, density, seed, _p(R))
    return R


def project(X, R):
    X = np.ascontiguousarray(X)
    n, d = X.shape
    L = R.shape[1]
    P = np.zeros((L, n), np.float32)
    cm = np.zeros(L, np.uint32)
    lib().orc_project(_p(X), int(X.dtype == np.uint8), n, d, _p(np.ascontiguousarray(R)), L,
                      _p(P), _p(cm))
    return P, cm


def quant_exponents(colmax_bits, N):
    a = ctypes.c_int()
    b = ctypes.c_int()
    lib().orc_quant_exponents(int(colmax_bits), int(N), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def node_split(v, stat_mode=0, kq1=0, kq2=0):
    v = np.ascontiguousarray(v, np.float32)
    mu = np.zeros(1, np.float32)
    var = np.zeros(1, np.float32)
    E = np.zeros(16, np.float32)
    H = np.zeros(16, np.uint32)
    med = lib().orc_node_split(_p(v), len(v), stat_mode, kq1, kq2, _p(mu), _p(var), _p(E), _p(H))
    return dict(median=np.float32(med), mu=mu[0], var=var[0], E=E, H=H)


def grow_tree(P, colmax, D, t, mode=0, leaf_max=0, stat_mode=0, want_medians=False):
    L, n = P.shape
    node = np.zeros(n, np.uint32)
    med = np.full(1 << (D - 1), np.nan, np.float32) if want_medians else None
    rc = lib().orc_grow_tree(_p(P), n, D, t, mode, leaf_max, stat_mode, _p(colmax), _p(node),
                             _p(med) if want_medians else None)
    assert rc == 0
    return (node, med) if want_medians else node


def local_perm(m, ld):
    out = np.zeros(m, np.int32)
    lib().orc_local_perm(m, ld, _p(out))
    return out


def local_sizes(m, ld):
    out = np.zeros(1 << ld, np.int32)
    lib().orc_local_sizes(m, ld, _p(out))
    return out


def local_split(LP, pts, ld):
    """Real-key local tree of one global leaf: LP [ld][n] column-major, pts the
    leaf's point ids; returns them rearranged into local leaves."""
    LP = np.ascontiguousarray(LP, np.float32)
    out = np.ascontiguousarray(pts, np.int32).copy()
    lib().orc_local_split(_p(LP), LP.shape[1], ld, _p(out), len(out))
    return out


def local_C_nonzeros(d, ld, density_int, seed):
    return lib().orc_local_C_nonzeros(d, ld, density_int, seed)


def dist(X, i, j):
    X = np.ascontiguousarray(X)
    return np.float32(lib().orc_dist(_p(X), int(X.dtype == np.uint8), X.shape[1], i, j))


def leaf_knn(X, perm, offs, K):
    X = np.ascontiguousarray(X)
    n, d = X.shape
    perm = np.ascontiguousarray(perm, np.int32)
    offs = np.ascontiguousarray(offs, np.int64)
    ci = np.full((n, K), -1, np.int32)
    cd = np.full((n, K), np.inf, np.float32)
    lib().orc_leaf_knn(_p(X), int(X.dtype == np.uint8), n, d, _p(perm), _p(offs), len(offs) - 1,
                       K, _p(ci), _p(cd))
    return ci, cd


def merge(cand_idx, cand_dist):
    T, n, K = cand_idx.shape
    oi = np.zeros((n, K), np.int32)
    od = np.zeros((n, K), np.float32)
    lib().orc_merge(n, T, K, _p(np.ascontiguousarray(cand_idx, np.int32)),
                    _p(np.ascontiguousarray(cand_dist, np.float32)), _p(oi), _p(od))
    return oi, od


def bruteforce_f64(X, queries, K):
    X = np.ascontiguousarray(X)
    n, d = X.shape
    q = np.ascontiguousarray(queries, np.int64)
    oi = np.zeros((len(q), K), np.int32)
    od = np.zeros((len(q), K), np.float64)
    lib().orc_bruteforce_f64(_p(X), int(X.dtype == np.uint8), n, d, _p(q), len(q), K, _p(oi),
                             _p(od))
    return oi, od


def tree_depth(n):
    return lib().orc_tree_depth(n)


def global_depth(tree_depth_, ratio=0.6):
    return lib().orc_global_depth(tree_depth_, ratio)


def local_depth(gmin, nn, offset=2):
    return lib().orc_local_depth(gmin, nn, offset)


def build_graph(X, ntrees, depth, nn, local_mode, leaf_max=0, local_offset=2, stat_mode=0,
                seed=1, density=None, nthreads=0, want_leaves=False):
    """Whole build. Returns (idx[n,K], dist[n,K], timings[5], leaf_heap[T,n] or None)."""
    X = np.ascontiguousarray(X)
    n, d = X.shape
    if density is None:
        density = 1.0 / np.sqrt(d)
    p = Params(n=n, d=d, x_is_u8=int(X.dtype == np.uint8), ntrees=ntrees, depth=depth, nn=nn,
               local_mode=local_mode, leaf_max=leaf_max, local_offset=local_offset,
               stat_mode=stat_mode, seed=seed, density=density, nthreads=nthreads, pad=0)
    oi = np.zeros((n, nn), np.int32)
    od = np.zeros((n, nn), np.float32)
    tm = np.zeros(5, np.float64)
    lh = np.zeros((ntrees, n), np.uint32) if want_leaves else None
    rc = lib().orc_build_graph(ctypes.byref(p), _p(X), _p(oi), _p(od), _p(tm),
                               _p(lh) if want_leaves else None)
    if rc != 0:
        raise RuntimeError(f"oracle build_graph failed rc={rc}")
    return oi, od, tm, lh
