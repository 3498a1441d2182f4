"""Independent Python restatements used to pin the oracle (test helper).

Written from the reference text, not from the oracle's C++:
  * R generation: std::mt19937 raw outputs (numpy's legacy MT19937 has the
    standard init_genrand seeding), libstdc++ generate_canonical<float,24>
    and normal_distribution<float> (Marsaglia polar method) with glibc's
    logf/sqrtf called through ctypes;
  * the node split of cpp/src/drpt/math/math_operations.cpp:121-383 with the
    canonical fixed-point sums computed in exact rational arithmetic;
  * the reference tree growth (drpt_global.cpp:228-352).
"""
import ctypes
import math
from fractions import Fraction

import numpy as np

f32 = np.float32
_libm = ctypes.CDLL("libm.so.6")
_libm.logf.restype = ctypes.c_float
_libm.logf.argtypes = [ctypes.c_float]
_libm.sqrtf.restype = ctypes.c_float
_libm.sqrtf.argtypes = [ctypes.c_float]


class StdMt19937:
    """Raw 32-bit outputs of std::mt19937(seed)."""

    def __init__(self, seed):
        self.rs = np.random.RandomState(seed & 0xFFFFFFFF)
        self.buf = []

    def __call__(self):
        if not self.buf:
            self.buf = list(self.rs.randint(0, 2 ** 32, size=4096, dtype=np.uint64)[::-1])
        return int(self.buf.pop())


def canonical_float(g):
    """libstdc++ generate_canonical<float, 24>(mt19937): one draw, float arithmetic."""
    s = f32(g())                 # (g() - min) converted to float
    r = f32(4294967296.0)        # range 2^32 as float
    ret = f32(s / r)
    if ret >= f32(1.0):
        ret = np.nextafter(f32(1.0), f32(0.0))
    return ret


class StdNormalFloat:
    """libstdc++ normal_distribution<float>(0, 1), polar method with a saved value."""

    def __init__(self):
        self.saved = None

    def __call__(self, g):
        if self.saved is not None:
            v, self.saved = self.saved, None
            return v
        while True:
            x = f32(float(f32(2.0) * canonical_float(g)) - 1.0)
            y = f32(float(f32(2.0) * canonical_float(g)) - 1.0)
            r2 = f32(f32(x * x) + f32(y * y))
            if not (r2 > 1.0 or r2 == 0.0):
                break
        lg = f32(_libm.logf(float(r2)))
        mult = f32(_libm.sqrtf(float(f32(f32(f32(-2.0) * lg) / r2))))
        self.saved = f32(x * mult)
        return f32(y * mult)


def build_R(d, L, density, seed):
    g = StdMt19937(seed)
    nrm = StdNormalFloat()
    R = np.zeros((d, L), np.float32)
    dens = f32(density)
    for j in range(d):
        for i in range(L):
            u = canonical_float(g)      # uniform_real_distribution<float>(0,1): 0 + 1*u
            R[j, i] = 0.0 if u > dens else nrm(g)
    return R


# ---------------------------------------------------------------------------
def quant_exponents(colmax_bits, N):
    E = (int(colmax_bits) >> 23) & 0xFF
    e = -127 if E == 0 else E - 127
    shift = 63 - int(N).bit_length()
    return e + 1 - shift, 2 * (e + 2) - shift


def _q(v, kq):
    """round-half-even of v / 2^kq, exactly."""
    return round(Fraction(float(v)) / (Fraction(2) ** kq))


def node_split(v, stat_mode, kq1=0, kq2=0):
    """Median of one node (math_operations.cpp:121-383); returns dict."""
    v = [f32(x) for x in v]
    n = len(v)
    if stat_mode == 1:                        # sequential float32 (reference, 1 thread)
        s = f32(0.0)
        for x in v:
            s = f32(s + x)
        mu = f32(s / f32(n)) if n else f32(np.nan)
        q = f32(0.0)
        for x in v:
            dx = f32(x - mu)
            q = f32(q + f32(dx * dx))
        var = f32(q / f32(n - 1)) if n != 1 else f32(np.nan)
    else:                                     # canonical fixed-point sums, exact rationals
        S = sum(_q(x, kq1) for x in v)
        mu = f32(math.ldexp(float(S), kq1) / float(n)) if n else f32(np.nan)
        S2 = 0
        for x in v:
            w = f32(x - mu)
            S2 += _q(f32(w * w), kq2)
        den = float(n - 1)
        var = f32(math.ldexp(float(S2), kq2) / den) if den != 0 else f32(np.nan)
    sigma = f32(_libm.sqrtf(float(var)))
    s1 = f32(float(sigma) / 30.0)
    s2 = f32(float(sigma) / 6.0)
    s3 = f32(float(sigma) / 2.0)
    E = [f32(0.0)] * 16
    for j in range(1, 5):
        r = f32(f32(j) * s1)
        E[6 + j] = f32(mu + r)
        E[7 - j] = f32(mu - r)
    for j in range(1, 3):
        r = f32(sigma + f32(f32(j) * s2))
        E[10 + j] = f32(mu + r)
        E[3 - j] = f32(mu - r)
    r = f32(f32(f32(2.0) * sigma) + s3)
    E[13] = f32(mu + r)
    E[0] = f32(mu - r)
    H = [0] * 16
    for x in v:
        hit = False
        for j in range(1, 16):
            if E[j - 1] < x and E[j] >= x:
                H[j] += 1
                hit = True
        if not hit:
            H[0] += 1
    target = f32((50 * n) // 100)
    cf = f32(0.0)
    sel = 15
    for k in range(16):
        cf = f32(float(cf) + float(H[k]))
        if cf >= target:
            sel = k
            break
    count = float(H[sel])
    lo = f32(0.0) if sel == 0 else E[sel - 1]
    width = f32(E[sel] - lo)
    with np.errstate(all="ignore"):
        frac = (float(n // 2) - (float(cf) - count)) / count if count != 0 else float("nan")
    med = f32(float(lo) + frac * float(width))
    return dict(median=med, mu=mu, var=var, E=np.array(E, np.float32), H=np.array(H))


def grow_tree(P, colmax, D, t, mode, leaf_max, stat_mode):
    """Heap index of each point's leaf (drpt_global.cpp:153-352; leaf-max variant)."""
    L, n = P.shape
    cur = [list(range(n))]
    out = np.zeros(n, np.uint32)
    for l in range(D - 1):
        col = t * D + l
        kq1, kq2 = quant_exponents(colmax[col], n)
        nxt = [[] for _ in range(2 * len(cur))]
        for i, pts in enumerate(cur):
            heap = (1 << l) + i
            if not pts:
                continue
            if mode == 1 and len(pts) <= leaf_max:
                out[pts] = heap
                continue
            vals = P[col, pts]
            med = node_split(vals, stat_mode, kq1, kq2)["median"]
            for p, x in zip(pts, vals):
                (nxt[2 * i] if x <= med else nxt[2 * i + 1]).append(p)
        cur = nxt
    for i, pts in enumerate(cur):
        out[pts] = (1 << (D - 1)) + i
    return out


# ---- libstdc++ std::nth_element (bits/stl_algo.h, GCC 4.9 - 13) -------------
# The reference local tree (cpp/src/drpt/algo/drpt_local.cpp:121-161) calls
# std::nth_element(begin, begin + n/2, end, less-by-projected-value). In its
# default configuration every projected value is 0 (mdrpt.cpp:348-352 receive
# an int density of 0), so the comparator is always false and the leaves are
# whatever positions libstdc++'s introselect leaves there. This restates that
# introselect step by step on a Python list, with the comparator as a function
# of two elements: __introselect with depth limit 2*floor(log2 n), the
# median-of-three pivot moved to the front (__move_median_to_first), the
# unguarded Hoare partition, __heap_select for an exhausted depth limit and
# the final insertion sort of a range of at most 3.
def _lg(n):
    return n.bit_length() - 1


def _move_median_to_first(a, result, i, j, k, less):
    if less(a[i], a[j]):
        if less(a[j], a[k]):
            a[result], a[j] = a[j], a[result]
        elif less(a[i], a[k]):
            a[result], a[k] = a[k], a[result]
        else:
            a[result], a[i] = a[i], a[result]
    elif less(a[i], a[k]):
        a[result], a[i] = a[i], a[result]
    elif less(a[j], a[k]):
        a[result], a[k] = a[k], a[result]
    else:
        a[result], a[j] = a[j], a[result]


def _unguarded_partition(a, first, last, pivot, less):
    while True:
        while less(a[first], a[pivot]):
            first += 1
        last -= 1
        while less(a[pivot], a[last]):
            last -= 1
        if not first < last:
            return first
        a[first], a[last] = a[last], a[first]
        first += 1


def _adjust_heap(a, base, hole, length, value, less):
    top = hole
    child = hole
    while child < (length - 1) // 2:
        child = 2 * (child + 1)
        if less(a[base + child], a[base + child - 1]):
            child -= 1
        a[base + hole] = a[base + child]
        hole = child
    if length % 2 == 0 and child == (length - 2) // 2:
        child = 2 * (child + 1)
        a[base + hole] = a[base + child - 1]
        hole = child - 1
    parent = (hole - 1) // 2
    while hole > top and less(a[base + parent], value):
        a[base + hole] = a[base + parent]
        hole = parent
        parent = (hole - 1) // 2
    a[base + hole] = value


def _make_heap(a, first, last, less):
    length = last - first
    if length < 2:
        return
    parent = (length - 2) // 2
    while True:
        _adjust_heap(a, first, parent, length, a[first + parent], less)
        if parent == 0:
            return
        parent -= 1


def _heap_select(a, first, middle, last, less):
    _make_heap(a, first, middle, less)
    for i in range(middle, last):
        if less(a[i], a[first]):
            value = a[i]
            a[i] = a[first]
            _adjust_heap(a, first, 0, middle - first, value, less)


def _insertion_sort(a, first, last, less):
    for i in range(first + 1, last):
        v = a[i]
        if less(v, a[first]):
            a[first + 1:i + 1] = a[first:i]
            a[first] = v
        else:
            j = i
            while less(v, a[j - 1]):
                a[j] = a[j - 1]
                j -= 1
            a[j] = v


def nth_element(a, first, nth, last, less):
    """std::nth_element(a+first, a+nth, a+last, less) on list a, in place."""
    if first == last or nth == last:
        return
    depth = 2 * _lg(last - first)
    while last - first > 3:
        if depth == 0:
            _heap_select(a, first, nth + 1, last, less)
            a[first], a[nth] = a[nth], a[first]
            return
        depth -= 1
        mid = first + (last - first) // 2
        _move_median_to_first(a, first, first + 1, mid, last - 1, less)
        cut = _unguarded_partition(a, first + 1, last, first, less)
        if cut <= nth:
            first = cut
        else:
            last = cut
    _insertion_sort(a, first, last, less)


def local_positions(m, ld, less=lambda x, y: False):
    """Positions of a reference local tree (drpt_local.cpp:121-161): at each
    level nth_element(begin, begin + size/2, end), then the left child is
    [begin, end - size/2) and the right child [end - size/2, end)."""
    a = list(range(m))

    def rec(b, e, depth):
        if depth == ld:
            return
        size = e - b
        nth_element(a, b, b + size // 2, e, less)
        mid = e - size // 2
        rec(b, mid, depth + 1)
        rec(mid, e, depth + 1)

    rec(0, m, 0)
    return a
