"""Argument limits and edge cases of the C ABI on the GPU: work-item plans
below 64 rows per item are refused, BF16 / FP8 builds refuse nn > 32 while
the FP32 build serves nn up to 64, the exact row path gives the same lists
when its rows are processed in small batches, and inputs whose squared
norms fall below the FP32 normal range keep BF16 / FP8 lists equal to the
FP32 ones."""
import os
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


def test_item_plan_below_64_rows_refused(ctx):
    _log("[limits] rows per item < 64")
    p = drpt_amd.leafmax_params(1_000_000, 128, 16, 10, leaf_max=512, precision="bf16")
    with pytest.raises(drpt_amd.DrptError) as e:
        ctx.set_item_plan(p, {"rows_per_item": [32, 64, 64, 64], "ring": [1] * 4})
    assert e.value.code == drpt_amd._lib.DRPT_ERR_ARG and ">= 64" in str(e.value)
    ctx.set_item_plan(p, {"rows_per_item": [64, 128, 256, 1 << 30], "ring": [1] * 4})


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_lowp_nn_above_32_refused(ctx, prec):
    _log(f"[limits] {prec} nn=48")
    X = datagen.int_mixture(2000)
    p = drpt_amd.leafmax_params(2000, 128, 2, 48, leaf_max=200, precision=prec)
    with pytest.raises(drpt_amd.DrptError) as e:
        ctx.build_graph(p, _dev(X))
    assert e.value.code == drpt_amd._lib.DRPT_ERR_UNSUPPORTED and "nn > 32" in str(e.value)


def test_fp32_nn_48_vs_oracle(ctx, orc):
    _log("[limits] fp32 nn=48")
    X = datagen.int_mixture(3000)
    p = drpt_amd.leafmax_params(3000, 128, 3, 48, leaf_max=200, seed=4)
    oi, od, _ = ctx.build_graph(p, _dev(X))
    ri, rd, _, _ = orc.build_graph(X, 3, p.depth, 48, local_mode=1, leaf_max=200, seed=4,
                                   density=p.density)
    assert np.array_equal(oi.cpu().numpy(), ri)
    assert np.array_equal(_bits(od.cpu().numpy()), _bits(rd))


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_exact_rows_in_small_batches(orc, prec):
    """400 copies of each vector: pools overflow and those rows take the exact
    row path; with DRPT_ROW_EXACT_BATCH=7 they go 7 rows per batch."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _log(f"[limits] {prec} exact rows in batches of 7")
    rng = np.random.default_rng(4)
    base = rng.integers(0, 50, size=(5, 16)).astype(np.float32)
    X = np.repeat(base, 400, axis=0)
    n, d = X.shape
    p = drpt_amd.leafmax_params(n, d, 4, 10, leaf_max=300, precision=prec, seed=8)
    os.environ["DRPT_ROW_EXACT_BATCH"] = "7"
    try:
        c = drpt_amd.Context(0)
    finally:
        del os.environ["DRPT_ROW_EXACT_BATCH"]
    with c:
        oi, od, tm = c.build_graph(p, _dev(X))
    _log(f"  exact rows {tm.n_fallback_rows}")
    assert tm.n_fallback_rows > 7
    ri, rd, _, _ = orc.build_graph(X, 4, p.depth, 10, local_mode=1, leaf_max=300, seed=8,
                                   density=p.density)
    assert np.array_equal(oi.cpu().numpy(), ri)
    assert np.array_equal(_bits(od.cpu().numpy()), _bits(rd))


@pytest.mark.parametrize("prec", ["bf16", "fp8"])
def test_tiny_magnitudes_equal_fp32(ctx, orc, prec):
    """Rows scaled to about 1e-20 (squared norms near 1e-38, FP32 subnormal
    products in the Gram): the BF16 / FP8 lists equal the FP32 and oracle lists."""
    _log(f"[limits] {prec} rows scaled to 1e-20")
    rng = np.random.default_rng(17)
    X = (rng.standard_normal((2000, 48)) * 1e-20).astype(np.float32)
    n, d = X.shape
    out = {}
    for pr in ("fp32", prec):
        p = drpt_amd.leafmax_params(n, d, 4, 10, leaf_max=150, precision=pr, seed=3)
        oi, od, tm = ctx.build_graph(p, _dev(X))
        out[pr] = (oi.cpu().numpy(), od.cpu().numpy())
    ri, rd, _, _ = orc.build_graph(X, 4, p.depth, 10, local_mode=1, leaf_max=150, seed=3,
                                   density=p.density)
    assert np.array_equal(out["fp32"][0], ri)
    assert np.array_equal(_bits(out["fp32"][1]), _bits(rd))
    assert np.array_equal(out[prec][0], ri)
    assert np.array_equal(_bits(out[prec][1]), _bits(rd))
