"""Build recipe for the native libraries of openkhuzdulgraph_amd.

Outputs (in-tree, git-ignored, shipped with the tree):
  openkhuzdulgraph_amd/lib/libkz_hip.so   the C ABI of include/kz.h
                                          (host graph layer + gfx950 kernels)

Host C++ is compiled with g++ (OpenMP), HIP sources with hipcc for gfx950
only, then linked by hipcc.  Run `python -m openkhuzdulgraph_amd.build`.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
OBJ_DIR = PKG / "_build"
INCLUDE = ROOT / "include"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"

HOST_SOURCES = ["kz_graph_host.cpp"]
HIP_SOURCES = ["kz_runtime.hip", "kz_clique.hip", "kz_plan.hip", "kz_count.hip", "kz_setops.hip", "kz_workload.hip",
               "kz_census.hip", "kz_server.hip",
               "kz_gpugen.hip", "kz_fsm.hip", "kz_paged.hip"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-value",  # best-effort hipFree on teardown paths
    # packed FP32 instructions are kept out of every device build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + proc.stdout)
    # surface real warnings, skip the host pass's note about the device feature
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line and line.strip()
    )
    if out:
        sys.stderr.write(out + "\n")


def build(verbose: bool = True) -> Path:
    """Compile every source and link libkz_hip.so; returns its path."""
    OBJ_DIR.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        obj = OBJ_DIR / (src + ".o")
        jobs.append((["g++", *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)], obj))
    for src in HIP_SOURCES:
        obj = OBJ_DIR / (src + ".o")
        jobs.append(([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)], obj))
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        futs = [ex.submit(_run, cmd) for cmd, _ in jobs]
        for f in futs:
            f.result()
    out = LIB_DIR / "libkz_hip.so"
    tmp = LIB_DIR / "libkz_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for _, o in jobs], "-lgomp"])
    os.replace(tmp, out)
    # the Khuzdul application API on top of it, and the reference's apps
    # compiled unchanged against it where the reference tree is present
    khz = PKG / "khuzdul"
    subprocess.run(["make", "-s", "-B", "-C", str(khz)], check=True)
    subprocess.run(["make", "-s", "-B", "-j4", "-C", str(khz), "examples"], check=True)
    if Path("/root/reference/distributed_applications").is_dir():
        subprocess.run(["make", "-s", "-j8", "-C", str(khz), "apps"], check=True)
    if verbose:
        print(f"built {out} and {LIB_DIR / 'libkhuzdul_amd.so'}")
    return out


def clean() -> None:
    shutil.rmtree(OBJ_DIR, ignore_errors=True)


if __name__ == "__main__":
    build()
