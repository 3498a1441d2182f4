"""CPU tests of the C ABI: the library loads, exports what include/drpt_hip.h
declares, and its host-only entry points agree with the oracle. No compute
calls reach a GPU here."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(drpt_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "drpt_build_graph" in names and "drpt_project" in names
    from drpt_amd import _lib
    assert set(names) == set(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    from drpt_amd import _lib
    lib = _lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (drpt_[a-z0-9_]+)", out))
    for name in _declared():
        assert name in exported, name
        assert hasattr(lib, name)


def test_version():
    from drpt_amd import _lib
    assert _lib.load().drpt_version() >= 100


@pytest.mark.parametrize("n,d,depth", [(60000, 784, 7), (1_000_000, 128, 10),
                                       (1_000_000, 960, 10), (10_000_000, 96, 12)])
def test_derive_params_matches_reference_rule(orc, n, d, depth):
    import drpt_amd
    p = drpt_amd.derive_params(n, d, 8, 10)
    assert p.depth == depth == orc.global_depth(orc.tree_depth(n))
    assert p.local_offset == 2 and p.local_mode == drpt_amd.DRPT_LOCAL_REFERENCE
    assert np.float32(p.density) == np.float32(1.0 / np.sqrt(d))


@pytest.mark.parametrize("d,L,density,seed", [(784, 56, 1 / 28, 5), (128, 160, 0.0884, 1),
                                              (10, 4, 3.0, 2)])
def test_build_projection_equals_oracle(orc, d, L, density, seed):
    import drpt_amd
    R1 = drpt_amd.build_projection(d, L, density, seed)
    R2 = orc.build_R(d, L, density, seed)
    assert np.array_equal(R1.view(np.uint32), R2.view(np.uint32))


def test_invalid_arguments_return_codes():
    from drpt_amd import _lib
    lib = _lib.load()
    assert lib.drpt_build_projection(0, 4, 0.5, 1, None) == _lib.DRPT_ERR_ARG
    assert b"invalid" in lib.drpt_last_error()
    assert lib.drpt_derive_params(1, 3, 1, 1, 0.6, None) == _lib.DRPT_ERR_ARG
    assert lib.drpt_sync(None) == _lib.DRPT_ERR_ARG
    assert lib.drpt_create(None, 0) == _lib.DRPT_ERR_ARG


def test_no_oracle_in_product():
    """The product never links or imports the oracle."""
    for path in glob.glob(os.path.join(ROOT, "drpt_amd", "**", "*"), recursive=True):
        if path.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
            src = open(path).read()
            assert "oracle" not in src.replace("the oracle", ""), path
