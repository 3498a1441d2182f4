"""drpt_grow_trees_noalloc (the split state in the library's fixed device
array, no device allocation by the library) returns drpt_grow_trees' trees
bit for bit, on two calls in a row, the second with a smaller input."""
import sys

import numpy as np
import pytest

import drpt_amd
from drpt_amd import datagen

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_noalloc_grow_trees_equals_grow_trees(ctx, orc):
    calls = [  # (points, trees, depth, tree mode, leaf_max)
        (40000, 3, 12, drpt_amd.DRPT_TREE_LEAFMAX, 64),
        (5000, 2, 9, drpt_amd.DRPT_TREE_FIXED, 0),
    ]
    X = datagen.int_mixture(40000)
    for k, (n, T, D, mode, leaf_max) in enumerate(calls):
        print(f"[noalloc] call {k}: n={n} T={T} D={D} mode={mode}", flush=True)
        sys.stdout.flush()
        Xk = X[:n]
        R = orc.build_R(Xk.shape[1], T * D, 1.0 / np.sqrt(Xk.shape[1]), 41 + k)
        P, cm = ctx.project(_dev(Xk), R)
        want = ctx.grow_trees(P, cm, T, D, mode, leaf_max).cpu().numpy()
        got = ctx.grow_trees_noalloc(P, cm, T, D, mode, leaf_max).cpu().numpy()
        assert np.array_equal(got, want), f"call {k}: {(got != want).sum()} heap entries differ"
        ref = orc.grow_tree(P.cpu().numpy(), cm.cpu().numpy().view(np.uint32), D, 0, mode,
                            leaf_max, 0)
        assert np.array_equal(got[0].view(np.uint32), ref)


def test_noalloc_rejects_sizes_beyond_scratch(ctx):
    import torch
    n = 4096
    P = torch.zeros((64 * 20, n), dtype=torch.float32, device="cuda:0")
    cm = torch.zeros((64 * 20,), dtype=torch.int32, device="cuda:0")
    # 64 trees x 2^18 nodes at the deepest split level: far beyond 196608 nodes
    with pytest.raises(drpt_amd.DrptError) as e:
        ctx.grow_trees_noalloc(P, cm, 64, 20, drpt_amd.DRPT_TREE_LEAFMAX, 8)
    assert e.value.code == drpt_amd._lib.DRPT_ERR_ARG
    assert "196608" in str(e.value)
