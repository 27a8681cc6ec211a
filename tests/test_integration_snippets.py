"""The drop-in bodies that INTEGRATION.md gives for the reference's
sgd_train (singleGPU/sgd.h:141) and sgd_update_k128 (singleGPU/sgd_kernel.h:17)
compile and link against this library's header and shared object, with
stand-ins for the reference's types laid out as singleGPU/sgd.h:26-138
declares them (Parameter, mf_node, mf_problem, mf_model; the CUDA `half`
pointers only need a complete type).  No GPU is needed: nothing is run."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "cumf_sgd_amd")

# The reference's types as singleGPU/sgd.h:26-138 lays them out.
MIRROR = r"""
#pragma once
#include <cstdio>
typedef int SGDIndex;
typedef float SGDRate;
struct half { unsigned short bits; };
struct Parameter {
    int k, num_workers, u_grid, v_grid, x_grid, y_grid, ux, vy, num_iters, gpu;
    SGDRate lambda_p, lambda_q, lrate;
    float alpha, beta;
    bool do_nmf, quiet, copy_data;
};
struct mf_node { SGDIndex u, v; SGDRate rate; };
struct mf_problem {
    SGDIndex m, n;
    long long nnz;
    int u_grid, v_grid, x_grid, y_grid, ux, vy;
    long long u_seg, v_seg;
    mf_node* R;
    mf_node** R2D;
    long long* gridSize;
    long long maxGridSize;
    mf_node* gpuR;
    int cur_u_id, cur_v_id;
    mf_node* gpuRptrs[2];
    int cur_global_x_id[2], cur_global_y_id[2];
};
struct mf_model {
    int fun;
    SGDIndex m, n;
    int u_grid, v_grid, x_grid, y_grid, ux, vy;
    long long u_seg, v_seg;
    int k;
    float b;
    float* floatp;
    float* floatq;
    short* halfp;
    short* halfq;
    half* gpuHalfp;
    half* gpuHalfq;
    int cur_u_id, cur_v_id;
    half* gpuHalfPptrs[2];
    half* gpuHalfQptrs[2];
    int cur_global_x_id[2], cur_global_y_id[2];
};
static_assert(sizeof(mf_node) == 12, "mf_node is the 12-byte record");
"""


def _snippets():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```cpp\n(.*?)```", text, re.S)
    return {name: next(b for b in blocks if re.search(rf"\b{name}\(", b) and "{" in b)
            for name in ("sgd_train", "sgd_update_k128")}


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
@pytest.mark.parametrize("name", ["sgd_train", "sgd_update_k128"])
def test_integration_body_compiles_and_links(tmp_path, name):
    lib = os.path.join(LIB_DIR, "libcumf_sgd_amd.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    body = _snippets()[name]
    (tmp_path / "sgd_mirror.h").write_text(MIRROR)
    src = tmp_path / f"{name}.cpp"
    src.write_text('#include "sgd_mirror.h"\n' + body +
                   "\nint main() { return 0; }\n")
    exe = tmp_path / name
    cmd = ["g++", "-std=c++14", "-Wall", "-Werror", "-Wno-unused-variable", f"-I{tmp_path}",
           f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe), f"-L{LIB_DIR}", "-lcumf_sgd_amd",
           f"-Wl,-rpath,{LIB_DIR}"]
    out = subprocess.run(cmd, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
