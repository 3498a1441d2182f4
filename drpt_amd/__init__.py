"""drpt_amd — MI355X-native DRPT k-NN graph construction.

The compute path is libdrpt_hip.so (hand-written gfx950 HIP kernels behind the
C ABI in include/drpt_hip.h); this module is the Python host side: device
buffers come from torch (plumbing only) and every compute call goes through
the C ABI. There is no CPU fallback: without the built library the calls raise.

Host interface mirroring the reference's MDRPT class
(cpp/src/drpt/algo/mdrpt.hpp:101-110): ``MDRPT(...)``, ``grow_trees(...)``,
``gather_nns(nn)``.
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import (DRPT_BF16, DRPT_F32, DRPT_FP32, DRPT_FP8, DRPT_LOCAL_LEAFMAX,
                   DRPT_LOCAL_REFERENCE, DRPT_TREE_FIXED, DRPT_TREE_LEAFMAX, DRPT_U8, DrptError,
                   Params, Phases, Timings)

__all__ = [
    "Context", "Params", "Timings", "DrptError", "derive_params", "build_projection",
    "MDRPT", "DRPT_U8", "DRPT_F32", "DRPT_FP32", "DRPT_BF16", "DRPT_FP8",
    "DRPT_LOCAL_REFERENCE", "DRPT_LOCAL_LEAFMAX", "DRPT_TREE_FIXED", "DRPT_TREE_LEAFMAX",
]

PRECISIONS = {"fp32": DRPT_FP32, "bf16": DRPT_BF16, "fp8": DRPT_FP8}


def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def derive_params(n, d, ntrees, nn, tree_depth_ratio=0.6):
    """Reference parameter derivation (main.cpp:129-145, mdrpt.cpp:52)."""
    p = Params()
    _lib.check(_lib.load().drpt_derive_params(n, d, ntrees, nn, tree_depth_ratio, ctypes.byref(p)))
    return p


def leafmax_params(n, d, ntrees, nn, leaf_max, precision="fp32", kprime=0, seed=1, depth=None,
                   tree_group=0):
    """This build's leaf-size-capped configuration: nodes larger than leaf_max
    are split with the reference split rule; depth = projection columns per tree;
    tree_group = trees per pooled filter pass of BF16/FP8 builds (0 = automatic)."""
    p = derive_params(n, d, ntrees, nn)
    p.local_mode = DRPT_LOCAL_LEAFMAX
    p.leaf_max = leaf_max
    if depth is None:
        depth = max(p.depth, int(math.ceil(math.log2(max(2, n / leaf_max)))) + 4)
    p.depth = depth
    p.precision = PRECISIONS[precision] if isinstance(precision, str) else precision
    p.kprime = kprime
    p.seed = seed
    p.tree_group = tree_group
    return p


def build_projection(d, L, density, seed):
    """Host R[d][L] (MathOp::build_sparse_local_random_matrix, math_operations.cpp:48-76)."""
    R = np.zeros((d, L), np.float32)
    _lib.check(_lib.load().drpt_build_projection(d, L, density, seed,
                                                 R.ctypes.data_as(ctypes.c_void_p)))
    return R


class Context:
    """One GPU, one stream, one device workspace (drpt_create / drpt_destroy)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        self.h = ctypes.c_void_p()
        _lib.check(self.lib.drpt_create(ctypes.byref(self.h), device))
        self.device = device

    def close(self):
        if self.h:
            self.lib.drpt_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        _lib.check(self.lib.drpt_sync(self.h))

    # ---- stages (device tensors in, device tensors out) --------------------
    def project(self, X, R):
        """P[L][n] = (X R)^T and per-column max |P| bits."""
        torch = _torch()
        n, d = X.shape
        L = R.shape[1]
        dtype = DRPT_U8 if X.dtype == torch.uint8 else DRPT_F32
        Rh = np.ascontiguousarray(R, np.float32)
        P = torch.empty((L, n), dtype=torch.float32, device=X.device)
        cm = torch.empty((L,), dtype=torch.int32, device=X.device)
        torch.cuda.synchronize(X.device)
        _lib.check(self.lib.drpt_project(self.h, _ptr(X), dtype, n, d,
                                         Rh.ctypes.data_as(ctypes.c_void_p), L, _ptr(P), _ptr(cm)))
        self.sync()
        return P, cm

    def grow_trees(self, P, colmax, ntrees, depth, tree_mode=DRPT_TREE_FIXED, leaf_max=0,
                   out=None):
        """Leaf heap index of every point in every tree, [ntrees][n] (into `out` when
        given: the same buffer lets the context replay its captured split graph)."""
        torch = _torch()
        L, n = P.shape
        heap = out if out is not None else torch.empty((ntrees, n), dtype=torch.int32,
                                                       device=P.device)
        if heap.shape != (ntrees, n) or heap.dtype != torch.int32 or not heap.is_contiguous():
            raise ValueError("out must be a contiguous int32 tensor of shape (ntrees, n)")
        torch.cuda.synchronize(P.device)
        _lib.check(self.lib.drpt_grow_trees(self.h, _ptr(P), _ptr(colmax), n, ntrees, depth,
                                            tree_mode, leaf_max, _ptr(heap)))
        self.sync()
        return heap

    def grow_trees_noalloc(self, P, colmax, ntrees, depth, tree_mode=DRPT_TREE_FIXED, leaf_max=0,
                           out=None):
        """drpt_grow_trees_noalloc: grow_trees with the split state in the
        library's fixed device array (no device allocation by the library)."""
        torch = _torch()
        L, n = P.shape
        heap = out if out is not None else torch.empty((ntrees, n), dtype=torch.int32,
                                                       device=P.device)
        if heap.shape != (ntrees, n) or heap.dtype != torch.int32 or not heap.is_contiguous():
            raise ValueError("out must be a contiguous int32 tensor of shape (ntrees, n)")
        torch.cuda.synchronize(P.device)
        _lib.check(self.lib.drpt_grow_trees_noalloc(self.h, _ptr(P), _ptr(colmax), n, ntrees,
                                                    depth, tree_mode, leaf_max, _ptr(heap)))
        self.sync()
        return heap

    def tree_knn(self, params, X, heap):
        torch = _torch()
        T, n, K = params.ntrees, params.n, params.nn
        ci = torch.empty((T, n, K), dtype=torch.int32, device=X.device)
        cd = torch.empty((T, n, K), dtype=torch.float32, device=X.device)
        tm = Timings()
        torch.cuda.synchronize(X.device)
        _lib.check(self.lib.drpt_tree_knn(self.h, ctypes.byref(params), _ptr(X), _ptr(heap),
                                          _ptr(ci), _ptr(cd), ctypes.byref(tm)))
        self.sync()
        return ci, cd, tm

    def merge(self, cand_idx, cand_dist):
        torch = _torch()
        T, n, K = cand_idx.shape
        oi = torch.empty((n, K), dtype=torch.int32, device=cand_idx.device)
        od = torch.empty((n, K), dtype=torch.float32, device=cand_idx.device)
        torch.cuda.synchronize(cand_idx.device)
        _lib.check(self.lib.drpt_merge(self.h, n, T, K, _ptr(cand_idx), _ptr(cand_dist), _ptr(oi),
                                       _ptr(od)))
        self.sync()
        return oi, od

    def build_graph(self, params, X, out=None):
        """One-shot build. Returns (idx[n,nn] int32, dist[n,nn] f32, Timings)."""
        torch = _torch()
        n, K = params.n, params.nn
        if out is None:
            oi = torch.empty((n, K), dtype=torch.int32, device=X.device)
            od = torch.empty((n, K), dtype=torch.float32, device=X.device)
        else:
            oi, od = out
        tm = Timings()
        torch.cuda.synchronize(X.device)
        _lib.check(self.lib.drpt_build_graph(self.h, ctypes.byref(params), _ptr(X), _ptr(oi),
                                             _ptr(od), ctypes.byref(tm)))
        return oi, od, tm

    def last_phases(self):
        """drpt_last_phases: stage times of the last build_graph."""
        ph = Phases()
        _lib.check(self.lib.drpt_last_phases(self.h, ctypes.byref(ph)))
        return ph.as_dict()

    def autotune(self, params, X):
        """drpt_autotune: the fastest work-item layout per leaf-size bucket
        (sizes <= 64, <= 256, <= 1024, larger) for this BF16/FP8 workload, kept
        by the context for later builds of the same shape. Returns
        {"rows_per_item": [4], "ring": [4], "tree_group": int, "first_trees": int,
        "rerank": int}."""
        torch = _torch()
        plan = (ctypes.c_int32 * 11)()
        torch.cuda.synchronize(X.device)
        _lib.check(self.lib.drpt_autotune(self.h, ctypes.byref(params), _ptr(X),
                                          ctypes.cast(plan, ctypes.c_void_p)))
        return {"rows_per_item": list(plan[:4]), "ring": list(plan[4:8]), "tree_group": plan[8],
                "first_trees": plan[9], "rerank": plan[10]}

    def set_item_plan(self, params, plan):
        """drpt_set_item_plan: restore a plan returned by autotune()."""
        arr = (ctypes.c_int32 * 11)(*(list(plan["rows_per_item"]) + list(plan["ring"]) +
                                      [plan.get("tree_group", 0), plan.get("first_trees", 0),
                                       plan.get("rerank", 0)]))
        _lib.check(self.lib.drpt_set_item_plan(self.h, ctypes.byref(params),
                                               ctypes.cast(arr, ctypes.c_void_p)))


class MDRPT:
    """Host mirror of the reference's MDRPT (mdrpt.hpp:101-110) on one GPU.

    MDRPT(ntrees, tree_depth, tree_depth_ratio, local_tree_offset,
          total_data_set_size, local_data_set_size, dimension)
    grow_trees(original_data, density, use_locality_optimization, nn)
    gather_nns(nn) -> {src_index: [(index, distance), ...]}  (self excluded,
        nn entries, ordered by (exact FP32 squared distance, index))
    Options beyond the reference: seed, precision, kprime, leaf_max, device.
    """

    def __init__(self, ntrees, tree_depth, tree_depth_ratio=0.6, local_tree_offset=2,
                 total_data_set_size=0, local_data_set_size=0, dimension=0, rank=0, world_size=1,
                 input_path="", output_path="", seed=1, precision="fp32", kprime=0, leaf_max=0,
                 device=0):
        if world_size != 1 or rank != 0:
            raise DrptError(_lib.DRPT_ERR_UNSUPPORTED, "one process drives one GPU (world_size 1)")
        self.ntrees = ntrees
        self.tree_depth = tree_depth
        self.ratio = tree_depth_ratio
        self.offset = local_tree_offset
        self.n = total_data_set_size
        self.d = dimension
        self.seed, self.precision, self.kprime, self.leaf_max = seed, precision, kprime, leaf_max
        self.ctx = Context(device)
        self._X = None
        self._density = None
        self._locality = 0
        self._result = None

    def grow_trees(self, original_data, density, use_locality_optimization=False, nn=10):
        torch = _torch()
        X = original_data
        if not torch.is_tensor(X):
            X = torch.as_tensor(np.asarray(X, dtype=np.float32))
        if not X.is_cuda:
            X = X.to(f"cuda:{self.ctx.device}")
        self._X = X.contiguous()
        self._density = float(density)
        self._locality = 1 if use_locality_optimization else 0
        self.n, self.d = self._X.shape

    def gather_nns(self, nn):
        p = derive_params(self.n, self.d, self.ntrees, nn, self.ratio)
        p.depth = int(self.tree_depth * self.ratio)
        p.dtype = DRPT_U8 if self._X.dtype == _torch().uint8 else DRPT_F32
        p.local_offset = self.offset
        p.density = self._density
        p.seed = self.seed
        p.precision = PRECISIONS[self.precision]
        p.kprime = self.kprime
        p.locality = self._locality
        if self.leaf_max:
            p.local_mode = DRPT_LOCAL_LEAFMAX
            p.leaf_max = self.leaf_max
        oi, od, tm = self.ctx.build_graph(p, self._X)
        self.timings = tm
        oi, od = oi.cpu().numpy(), od.cpu().numpy()
        out = {}
        for s in range(self.n):
            out[s] = [(int(i), float(dv)) for i, dv in zip(oi[s], od[s]) if i >= 0]
        return out
