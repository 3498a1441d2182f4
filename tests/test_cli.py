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
