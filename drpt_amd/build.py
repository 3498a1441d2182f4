"""Build libdrpt_hip.so (gfx950) and the drpt command-line driver in-tree.

Device code: hipcc --offload-arch=gfx950, no FP contraction (the split
arithmetic is specified operation by operation) and no packed-FP32
instructions. Host pieces whose values are defined by libstdc++ (the
projection matrix, the reference local-tree permutations) are compiled with
g++, as the reference is.

Usage: python -m drpt_amd.build   (or drpt_amd.build.build())
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdrpt_hip.so")
CLI = os.path.join(HERE, "drpt")
ARCH = os.environ.get("DRPT_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-pass-failed",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall"]
CLI_FLAGS = ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-pthread"]

HIP_SOURCES = ["api.hip", "util.hip", "project.hip", "split.hip", "group.hip", "leafknn.hip",
               "leafpool.hip", "merge.hip"]
HOST_SOURCES = ["host_ref.cpp"]
CLI_SOURCES = ["drpt_main.cpp", "io.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _obj(src):
    return os.path.join(OUT, os.path.splitext(src)[0] + ".o")


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append(["hipcc", *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in HOST_SOURCES:
        jobs.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in CLI_SOURCES:
        jobs.append(["g++", *HOST_FLAGS, *CLI_FLAGS, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    workers = min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(workers) as ex:
        for r in ex.map(_run, jobs):
            msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                            if "packed-fp32-ops" not in l and l.strip())
            if verbose and msg:
                print(msg, file=sys.stderr)
    lib_objs = [_obj(s) for s in HIP_SOURCES + HOST_SOURCES]
    tmp = LIB + ".tmp"
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-o", tmp, *lib_objs])
    os.replace(tmp, LIB)
    cli_objs = [_obj(s) for s in CLI_SOURCES]
    _run(["g++", "-pthread", "-o", CLI, *cli_objs, f"-L{HERE}", "-ldrpt_hip", "-Wl,-rpath,$ORIGIN",
          "-L/opt/rocm/lib", "-lamdhip64"])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
