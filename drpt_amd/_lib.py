"""ctypes binding of libdrpt_hip.so (the C ABI declared in include/drpt_hip.h).

The library is built in-tree by drpt_amd.build; there is no fallback: if it is
missing, importing the compute entry points raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdrpt_hip.so")

DRPT_OK = 0
DRPT_ERR_ARG = -1
DRPT_ERR_HIP = -2
DRPT_ERR_NOMEM = -3
DRPT_ERR_DEPTH = -4
DRPT_ERR_UNSUPPORTED = -5
DRPT_ERR_NONFINITE = -6

DRPT_U8, DRPT_F32 = 0, 1
DRPT_FP32, DRPT_BF16, DRPT_FP8 = 0, 1, 2
DRPT_LOCAL_REFERENCE, DRPT_LOCAL_LEAFMAX = 0, 1
DRPT_TREE_FIXED, DRPT_TREE_LEAFMAX = 0, 1

# every symbol include/drpt_hip.h declares
EXPORTS = [
    "drpt_last_error", "drpt_version", "drpt_create", "drpt_destroy", "drpt_sync",
    "drpt_get_stream", "drpt_derive_params", "drpt_build_projection", "drpt_project",
    "drpt_grow_trees", "drpt_tree_knn", "drpt_merge", "drpt_build_graph", "drpt_autotune",
    "drpt_set_item_plan", "drpt_grow_trees_noalloc", "drpt_last_phases",
]


class DrptError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"drpt error {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64), ("d", ctypes.c_int32), ("dtype", ctypes.c_int32),
        ("ntrees", ctypes.c_int32), ("depth", ctypes.c_int32), ("nn", ctypes.c_int32),
        ("local_mode", ctypes.c_int32), ("leaf_max", ctypes.c_int32),
        ("local_offset", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("kprime", ctypes.c_int32), ("seed", ctypes.c_uint32), ("density", ctypes.c_float),
        ("tree_group", ctypes.c_int32), ("locality", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 3),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class Timings(ctypes.Structure):
    _fields_ = [
        ("project_ms", ctypes.c_double), ("split_ms", ctypes.c_double),
        ("group_ms", ctypes.c_double), ("knn_ms", ctypes.c_double),
        ("merge_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
        ("gram_flops", ctypes.c_double), ("knn_kernel_ms", ctypes.c_double),
        ("n_leaves", ctypes.c_int64), ("n_fallback_rows", ctypes.c_int64),
        ("max_leaf", ctypes.c_int64), ("reserved", ctypes.c_int64 * 4),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class Phases(ctypes.Structure):
    _fields_ = [
        ("group_ms", ctypes.c_double), ("prep_ms", ctypes.c_double), ("cand_ms", ctypes.c_double),
        ("known_ms", ctypes.c_double), ("pool_ms", ctypes.c_double),
        ("pool_merge_ms", ctypes.c_double), ("bound_ms", ctypes.c_double),
        ("exact_ms", ctypes.c_double), ("tree_groups", ctypes.c_int64),
        ("first_trees", ctypes.c_int64), ("pool_cap", ctypes.c_int64), ("rerank", ctypes.c_int64),
        ("pool_entries", ctypes.c_int64), ("int8_passes", ctypes.c_int64),
        ("reserved", ctypes.c_int64 * 2),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


_lib = None


def load():
    """Load libdrpt_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m drpt_amd.build` first")
    # torch ships its own HIP / HSA runtime libraries. Loaded after them, this
    # library shares torch's runtime; loaded before them, its HIP calls found no
    # device once torch had initialised the GPU. So torch goes first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, u32, f32, dbl = (ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint32,
                                   ctypes.c_float, ctypes.c_double)
    L.drpt_last_error.restype = ctypes.c_char_p
    L.drpt_last_error.argtypes = []
    L.drpt_version.restype = i32
    L.drpt_create.argtypes = [ctypes.POINTER(vp), i32]
    L.drpt_destroy.argtypes = [vp]
    L.drpt_sync.argtypes = [vp]
    L.drpt_get_stream.argtypes = [vp]
    L.drpt_get_stream.restype = vp
    L.drpt_derive_params.argtypes = [i64, i32, i32, i32, dbl, ctypes.POINTER(Params)]
    L.drpt_build_projection.argtypes = [i32, i32, f32, u32, vp]
    L.drpt_project.argtypes = [vp, vp, i32, i64, i32, vp, i32, vp, vp]
    L.drpt_grow_trees.argtypes = [vp, vp, vp, i64, i32, i32, i32, i32, vp]
    L.drpt_grow_trees_noalloc.argtypes = [vp, vp, vp, i64, i32, i32, i32, i32, vp]
    L.drpt_tree_knn.argtypes = [vp, ctypes.POINTER(Params), vp, vp, vp, vp, ctypes.POINTER(Timings)]
    L.drpt_merge.argtypes = [vp, i64, i32, i32, vp, vp, vp, vp]
    L.drpt_build_graph.argtypes = [vp, ctypes.POINTER(Params), vp, vp, vp, ctypes.POINTER(Timings)]
    L.drpt_autotune.argtypes = [vp, ctypes.POINTER(Params), vp, vp]
    L.drpt_set_item_plan.argtypes = [vp, ctypes.POINTER(Params), vp]
    L.drpt_last_phases.argtypes = [vp, ctypes.POINTER(Phases)]
    for name in EXPORTS:
        fn = getattr(L, name)
        if fn.restype is ctypes.c_int or name in ("drpt_create", "drpt_destroy", "drpt_sync"):
            fn.restype = i32
    _lib = L
    return L


def check(rc):
    if rc != DRPT_OK:
        msg = load().drpt_last_error().decode(errors="replace")
        raise DrptError(rc, msg)
    return rc
