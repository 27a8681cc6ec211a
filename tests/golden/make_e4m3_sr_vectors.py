"""Known-answer vectors for FP8 E4M3 stochastic rounding, taken from gfx950's
v_cvt_sr_fp8_f32 itself (the conversion the update kernel stores with).

Builds a one-kernel probe with hipcc, feeds it (f, r16 << 16) pairs (after
the kernel's clamp of |f| to 448, NaN kept) and writes f, r16 and the output
byte to the given .npz path.  Needs a gfx950 GPU; the vectors it wrote are
committed beside it as e4m3_sr_hw.npz.

    python tests/golden/make_e4m3_sr_vectors.py tests/golden/e4m3_sr_hw.npz
"""
import ctypes
import os
import subprocess
import sys
import tempfile

import numpy as np

SRC = r"""
#include <hip/hip_runtime.h>
#include <stdint.h>
__global__ void sr_vectors(const float* f, const uint32_t* r16, uint32_t* out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = f[i];
    const float c = __builtin_fminf(__builtin_fmaxf(x, -448.0f), 448.0f);
    const float v = __builtin_isnan(x) ? x : c;
    out[i] = (uint32_t)__builtin_amdgcn_cvt_sr_fp8_f32(v, (int)(r16[i] << 16), 0, 0) & 0xffu;
}
extern "C" int run(const float* f, const uint32_t* r16, uint32_t* out, int n)
{
    float* df; uint32_t* dr; uint32_t* dout;
    if (hipMalloc(&df, 4 * (size_t)n) || hipMalloc(&dr, 4 * (size_t)n) || hipMalloc(&dout, 4 * (size_t)n)) return 1;
    hipMemcpy(df, f, 4 * (size_t)n, hipMemcpyHostToDevice);
    hipMemcpy(dr, r16, 4 * (size_t)n, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(sr_vectors, dim3((n + 255) / 256), dim3(256), 0, 0, df, dr, dout, n);
    int rc = hipDeviceSynchronize() != hipSuccess ? 2 : 0;
    hipMemcpy(out, dout, 4 * (size_t)n, hipMemcpyDeviceToHost);
    hipFree(df); hipFree(dr); hipFree(dout);
    return rc;
}
"""


def cases():
    rng = np.random.default_rng(20261203)
    f, r = [], []
    # subnormal range: lo*2^-9 + 2^-(9+k) against the complement of its 20-bit fraction
    for lo in (0, 1, 3, 6, 7):
        for kk in range(1, 25):
            v = np.float64(lo) * 2.0 ** -9 + 2.0 ** -(9 + kk)
            fv = np.float32(v)
            if np.float64(fv) != v:
                continue
            frac20 = int(np.floor((v * 2 ** 9 - lo) * 2 ** 20))
            need = (1 << 20) - frac20            # r16 << 4 must reach this
            for d in (-16, -1, 0, 1, 16):
                x = ((need + d) + 15) >> 4
                if 0 <= x < (1 << 16):
                    f += [fv, -fv]
                    r += [x, x]
    # normal range: random values, words at and around the threshold
    for _ in range(2000):
        v = np.float32(np.exp2(rng.uniform(-6, 8.8)))
        frac20 = int(v.view(np.uint32)) & 0xFFFFF
        x = ((1 << 20) - frac20 + 15) >> 4
        for y in (x - 1, x):
            if 0 <= y < (1 << 16):
                f.append(v); r.append(y)
    # random values over the whole range, both signs, random words
    n = 16384
    mag = np.exp2(rng.uniform(-16, 9.2, n)).astype(np.float32)
    sgn = np.where(rng.random(n) < 0.5, -1, 1).astype(np.float32)
    f += list(mag * sgn)
    r += list(rng.integers(0, 1 << 16, n))
    specials = [0.0, -0.0, np.inf, -np.inf, 448.0, -448.0, 464.0, 480.0, 1e6, -1e6, 447.9, 2.0 ** -6, 2.0 ** -9,
                2.0 ** -10, 1.5 * 2.0 ** -10, 1e-30, -1e-30]
    for v in specials:
        for y in (0, 1, 0x8000, 0xFFFF):
            f.append(np.float32(v)); r.append(y)
    for bits in (0x7FC00000, 0xFFC00000, 0x7F800001, 0xFFFFFFFF):
        for y in (0, 0xFFFF):
            f.append(np.array([bits], np.uint32).view(np.float32)[0]); r.append(y)
    return np.array(f, np.float32), np.array(r, np.uint32)


def main(path):
    f, r = cases()
    with tempfile.TemporaryDirectory() as d:
        src, so = os.path.join(d, "p.hip"), os.path.join(d, "p.so")
        open(src, "w").write(SRC)
        subprocess.run(["hipcc", "--offload-arch=gfx950", "-O2", "-fPIC", "-shared", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-o", so, src], check=True)
        lib = ctypes.CDLL(so)
        out = np.zeros(len(f), np.uint32)
        rc = lib.run(f.ctypes.data_as(ctypes.c_void_p), r.ctypes.data_as(ctypes.c_void_p),
                     out.ctypes.data_as(ctypes.c_void_p), len(f))
        if rc:
            raise SystemExit(f"probe failed: {rc}")
    np.savez_compressed(path, f=f, r16=r.astype(np.uint16), byte=out.astype(np.uint8))
    print(f"wrote {len(f)} vectors to {path}")


if __name__ == "__main__":
    main(sys.argv[1])
