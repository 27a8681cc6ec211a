"""Build recipe: the gfx950 library, the cumf_sgd executable and the CPU oracle.

Produces, in-tree:
  cumf_sgd_amd/libcumf_sgd_amd.so   HIP kernels + C ABI + host pipeline
  cumf_sgd_amd/cumf_sgd             command-line trainer (singleGPU/main.cpp's CLI)
  cumf_sgd_amd/cumf_predict         GPU evaluator (test/mf-predict's CLI)
  oracle/liboracle.so, oracle/_ref/ oracle only (tests and benchmark baseline)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libcumf_sgd_amd.so")
CLI = os.path.join(PKG, "cumf_sgd")
PREDICT = os.path.join(PKG, "cumf_predict")
ARCH = os.environ.get("CUMF_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    # packed FP32 ops stay off (the compiler's host pass warns that it does
    # not know the feature; that is harmless)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++14", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out:
        sys.stderr.write(out + "\n")


def build_library() -> str:
    os.makedirs(BUILD, exist_ok=True)
    csrc = os.path.join(PKG, "csrc")
    host = os.path.join(PKG, "host")
    hip_units = ["sgd_kernels", "sgd_update_f32", "sgd_update_f16", "sgd_update_bf16", "sgd_update_e4m3", "sgd_update_bf16sr", "capi",
                 "batch", "loader", "rank", "noalloc", "order"]
    jobs = [[HIPCC, *HIP_FLAGS, "-c", os.path.join(csrc, f"{u}.hip"), "-o", os.path.join(BUILD, f"{u}.o")]
            for u in hip_units]
    jobs.append(["g++", *HOST_FLAGS, "-c", os.path.join(host, "host.cpp"), "-o", os.path.join(BUILD, "host.o")])
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp,
          *[os.path.join(BUILD, f"{u}.o") for u in hip_units], os.path.join(BUILD, "host.o")])
    os.replace(tmp, LIB)
    for exe, src in ((CLI, "main.cpp"), (PREDICT, "predict.cpp")):
        _run(["g++", *HOST_FLAGS, "-o", exe, os.path.join(host, src), f"-L{PKG}", "-lcumf_sgd_amd",
              "-Wl,-rpath,$ORIGIN"])
    return LIB


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all() -> None:
    build_library()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print(LIB)
