"""The drpt driver: argument validation (CPU, exits before any GPU call) and,
on a GPU, an end-to-end run whose edge list equals the oracle's lists written
in the reference's format (file_writer.cpp:99-150)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "drpt_amd", "drpt")


def _run(args, **kw):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300, **kw)


@pytest.mark.parametrize("args,msg", [
    ([], "Valid input path needed!..."),
    (["-input", "x"], "Valid out path needed!..."),
    (["-input", "x", "-output", "o"], "Dataset size should be greater than 0"),
    (["-input", "x", "-output", "o", "-data-set-size", "10"], "Dimension size should be greater than 0"),
    (["-input", "x", "-output", "o", "-data-set-size", "10", "-dimension", "4"],
     "Nearest neighbours size should be greater than 0"),
    (["-input", "x", "-output", "o", "-data-set-size", "10", "-dimension", "4", "-nn", "3",
      "-data-file-format", "7"], "Unsupported data file format"),
])
def test_cli_validation_messages(args, msg):
    r = _run(args)
    assert r.returncode == 1
    assert msg in r.stdout


def test_cli_missing_input_file(tmp_path):
    r = _run(["-input", str(tmp_path / "none.idx3"), "-output", str(tmp_path) + "/",
              "-data-set-size", "10", "-dimension", "4", "-nn", "3"])
    assert r.returncode == 1
    assert "cannot open" in r.stdout


def expected_edges(idx, nn_write):
    lines = []
    for s, row in enumerate(idx):
        w = 0
        for dst in row:
            if dst < 0 or w >= nn_write:
                break
            if dst == s:
                continue
            lines.append(f"{s + 1} {dst + 1}\n")
            w += 1
    return "".join(lines)


@pytest.mark.gpu
def test_cli_end_to_end_matches_oracle(tmp_path, orc):
    from drpt_amd import datagen
    X = datagen.mnist_like(3000)
    path = tmp_path / "mnist.idx3-ubyte"
    datagen.write_idx3_ubyte(str(path), X, 28, 28)
    out = str(tmp_path) + "/"
    print("[cli] mnist 3000", flush=True)
    r = _run(["-input", str(path), "-output", out, "-data-set-size", "3000", "-dimension", "784",
              "-ntrees", "4", "-nn", "10", "-seed", "21"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Index building (s)" in r.stdout
    D = orc.global_depth(orc.tree_depth(3000))
    ri, _, _, _ = orc.build_graph(X, 4, D, 10, local_mode=0, seed=21)
    got = open(out + "results.txt").read()
    assert got == expected_edges(ri, 9)


@pytest.mark.gpu
def test_cli_fbin_bf16_autotune_matches_oracle(tmp_path, orc):
    """fbin input, leaf-max trees, BF16 candidates with the autotuner and a fixed
    tree group: the edge list equals the oracle's (the same lists as FP32)."""
    from drpt_amd import datagen
    X = datagen.int_mixture(4000)
    path = tmp_path / "int.fbin"
    datagen.write_fbin(str(path), X)
    out = str(tmp_path) + "/"
    print("[cli] fbin bf16 autotune", flush=True)
    r = _run(["-input", str(path), "-output", out, "-data-set-size", "4000", "-dimension", "128",
              "-ntrees", "6", "-nn", "10", "-seed", "5", "-data-file-format", "1",
              "-bytes-for-data-type", "4", "-data-starting-index", "8", "-precision", "bf16",
              "-leaf-max", "300", "-tree-group", "3", "-autotune", "1"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Work-item plan" in r.stdout
    import drpt_amd
    p = drpt_amd.leafmax_params(4000, 128, 6, 10, leaf_max=300, seed=5)  # the driver's depth rule
    ri, _, _, _ = orc.build_graph(X, 6, p.depth, 10, local_mode=1, leaf_max=300, seed=5,
                                  density=p.density)
    got = open(out + "results.txt").read()
    assert got == expected_edges(ri, 9)


@pytest.mark.gpu
def test_cli_full_mnist_like_matches_oracle(tmp_path, orc):
    """The reference configuration at full size: 60000 x 784 uint8 idx3, T = 8,
    nn = 10 (the reference's tree depth and local stage): results.txt equals
    the oracle's lists in the reference's edge-list format, byte for byte."""
    from drpt_amd import datagen
    X = datagen.mnist_like(60000)
    path = tmp_path / "mnist.idx3-ubyte"
    datagen.write_idx3_ubyte(str(path), X, 28, 28)
    out = str(tmp_path) + "/"
    print("[cli] mnist 60000 full", flush=True)
    r = _run(["-input", str(path), "-output", out, "-data-set-size", "60000", "-dimension", "784",
              "-ntrees", "8", "-nn", "10", "-seed", "3"])
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip().splitlines()[-1], flush=True)
    D = orc.global_depth(orc.tree_depth(60000))
    ri, _, _, _ = orc.build_graph(X, 8, D, 10, local_mode=0, seed=3)
    got = open(out + "results.txt").read()
    assert got == expected_edges(ri, 9)


@pytest.mark.gpu
def test_cli_integer_density_real_local_tree(tmp_path, orc):
    """-density 1 (parsed as an integer, as main.cpp:69-72 does): dense R and a
    dense local matrix, the real-key local tree; results.txt equals the
    oracle's lists."""
    from drpt_amd import datagen
    X = datagen.mnist_like(3000)
    path = tmp_path / "mnist.idx3-ubyte"
    datagen.write_idx3_ubyte(str(path), X, 28, 28)
    out = str(tmp_path) + "/"
    print("[cli] mnist 3000 -density 1", flush=True)
    r = _run(["-input", str(path), "-output", out, "-data-set-size", "3000", "-dimension", "784",
              "-ntrees", "5", "-nn", "10", "-seed", "23", "-density", "1"])
    assert r.returncode == 0, r.stdout + r.stderr
    D = orc.global_depth(orc.tree_depth(3000))
    ri, _, _, _ = orc.build_graph(X, 5, D, 10, local_mode=0, seed=23, density=1.0)
    got = open(out + "results.txt").read()
    assert got == expected_edges(ri, 9)
