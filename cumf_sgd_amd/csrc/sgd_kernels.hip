// gfx950 kernels of the SGD hot path.
//
//  launch_sgd          dispatch of the batch-Hogwild update (sgd_update.h)
//  init_factors_kernel random factor init (random_init, singleGPU/sgd.cu:497-514)
//  fp8_rescale_kernel  between-epoch power-of-two row rescale (FP8 only)
//  export_kernel       dequantize * mult + inverse permutation gather
//                      (transform_half, sgd_kernel.cu:29-38; scale_model and
//                      shuffle_model, sgd.cu:593-635)
//  eval_kernel         test RMSE and the other mf-predict criteria (mf_predict +
//                      calc_rmse/mae/gkl/logloss/accuracy, test/mf.cpp:3794-3912)
//  split/remap/scale/moments/partition kernels: on-device preprocessing
//                      (sgd.cu:78-134, 411-487)
#include "kernels.h"
#include "sgd_update.h"

#include <hip/hip_runtime.h>

namespace cumf {

bool sgd_shape_supported(int k, int prec, int G)
{
    if (G != 64 && G != 32 && G != 16 && G != 8) return false;
    if (k != 32 && k != 64 && k != 128) return false;
    if (k % G) return false;
    int E = k / G;
    int eb = prec == kF32 ? 4 : (prec == kE4M3 ? 1 : 2);
    if (E * eb < 2 || E * eb > 32) return false;
    if (prec == kE4M3 && E % 4) return false;
    return true;
}

int sgd_default_group(int k, int prec)
{
    // 16 bytes of a row per lane (one 128-bit load and store per lane per
    // row, several samples per wave): within 2% of the fastest shape found by
    // timing every shape on the medium Netflix-shaped workload, per format
    // and k.
    const int eb = prec == kF32 ? 4 : (prec == kE4M3 ? 1 : 2);
    int G = k * eb / 16;
    if (G > 64) G = 64;
    if (G < 8) G = 8;
    return G;
}

hipError_t launch_sgd(const SgdLaunch& a, int blocks, int k, int prec, int G, bool coh, int D, hipStream_t s)
{
    switch (prec) {
    case kF32: return launch_sgd_prec<kF32>(a, blocks, k, G, coh, D, s);
    case kF16: return launch_sgd_prec<kF16>(a, blocks, k, G, coh, D, s);
    case kBF16: return launch_sgd_prec<kBF16>(a, blocks, k, G, coh, D, s);
    case kE4M3: return launch_sgd_prec<kE4M3>(a, blocks, k, G, coh, D, s);
    case kBF16SR: return launch_sgd_prec<kBF16SR>(a, blocks, k, G, coh, D, s);
    }
    return hipErrorInvalidValue;
}

size_t sgd_tile_lds_bytes(int k, int prec, int G, int rows, int waves)
{
    const size_t eb = prec == kF32 ? 4 : (prec == kE4M3 ? 1 : 2);
    const size_t stage = 64 * (size_t)((kSgdTileMaxBatch + G - 1) / G);   // words per array and wave
    return (size_t)rows * k * eb + (size_t)waves * 3 * stage * 4;
}

// gfx950: 160 KB of LDS per CU, all of it addressable by one workgroup
size_t sgd_tile_lds_limit() { return 160 * 1024; }

hipError_t launch_sgd_tile(const SgdLaunch& a, int blocks, int k, int prec, int G, int D, int rows, int waves,
                           hipStream_t s)
{
    if (a.batch_len > kSgdTileMaxBatch) return hipErrorInvalidValue;
    if (sgd_tile_lds_bytes(k, prec, G, rows, waves) > sgd_tile_lds_limit()) return hipErrorInvalidValue;
    // dynamic LDS: the tile (the kernel's record staging is static)
    const size_t eb = prec == kF32 ? 4 : (prec == kE4M3 ? 1 : 2);
    const size_t lds = (size_t)rows * k * eb;
    switch (prec) {
    case kF32: return launch_tile_prec<kF32>(a, blocks, k, G, D, waves, lds, s);
    case kF16: return launch_tile_prec<kF16>(a, blocks, k, G, D, waves, lds, s);
    case kBF16: return launch_tile_prec<kBF16>(a, blocks, k, G, D, waves, lds, s);
    case kE4M3: return launch_tile_prec<kE4M3>(a, blocks, k, G, D, waves, lds, s);
    case kBF16SR: return launch_tile_prec<kBF16SR>(a, blocks, k, G, D, waves, lds, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_sgd_batched(const SgdLaunch& a, int blocks, int prec, int D, hipStream_t s)
{
    switch (prec) {
    case kF32: return launch_batched_for<kF32>(a, blocks, D, s);
    case kF16: return launch_batched_for<kF16>(a, blocks, D, s);
    case kBF16: return launch_batched_for<kBF16>(a, blocks, D, s);
    case kE4M3: return launch_batched_for<kE4M3>(a, blocks, D, s);
    case kBF16SR: return launch_batched_for<kBF16SR>(a, blocks, D, s);
    }
    return hipErrorInvalidValue;
}

int sgd_batch_group(int k, int prec)
{
    switch (prec) {
    case kF32: return batch_group<kF32>(k);
    case kF16: return batch_group<kF16>(k);
    case kBF16: return batch_group<kBF16>(k);
    case kBF16SR: return batch_group<kBF16SR>(k);
    default: return batch_group<kE4M3>(k);
    }
}

int sgd_batch_budget_group(int k, int prec)
{
    return prec == kE4M3 ? batch_budget_group<kE4M3>(k) : batch_budget_group<kF16>(k);
}

// ---------------------------------------------------------------------------
// Device-side epoch counter.
// ---------------------------------------------------------------------------
__global__ void set_int_kernel(int32_t* p, int32_t value) { *p = value; }
__global__ void advance_int_kernel(int32_t* p) { *p += 1; }

hipError_t launch_set_int(int32_t* p, int32_t value, hipStream_t st)
{
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, p, value);
    return hipGetLastError();
}

hipError_t launch_advance_int(int32_t* p, hipStream_t st)
{
    hipLaunchKernelGGL(advance_int_kernel, dim3(1), dim3(1), 0, st, p);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Factor init: element e of row r is U(0,1] * s from word (e & 3) of
// Philox(ctr = {e >> 2, r_lo, r_hi, stream}).  One wave per row.  FP8 rows get
// the power-of-two scale for kFp8InitTarget and RNE quantization.
// ---------------------------------------------------------------------------
__device__ inline float wave_max(float x)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) x = fmaxf(x, __shfl_xor(x, d, 64));
    return x;
}

template <int PREC>
__global__ __launch_bounds__(256) void init_factors_kernel(void* F, float* scales, int64_t rows, int k,
                                                           float s, Key key, uint32_t stream)
{
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < rows; r += nwaves) {
        float vals[4];
        float amax = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int e = lane + 64 * c;
            vals[c] = 0.0f;
            if (e < k) {
                U4 w = philox(U4{(uint32_t)(e >> 2), (uint32_t)r, (uint32_t)((uint64_t)r >> 32), stream}, key.k0, key.k1);
                const uint32_t word = (e & 3) == 0 ? w.x : (e & 3) == 1 ? w.y : (e & 3) == 2 ? w.z : w.w;
                vals[c] = u01_of(word) * s;
                amax = fmaxf(amax, vals[c]);
            }
        }
        float inv = 1.0f;
        if constexpr (PREC == kE4M3) {
            amax = wave_max(amax);
            const float sc = pow2_scale(amax, kFp8InitTarget);
            inv = 1.0f / sc;
            if (lane == 0) scales[r] = sc;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int e = lane + 64 * c;
            if (e >= k) continue;
            const int64_t i = r * k + e;
            if constexpr (PREC == kF32) ((float*)F)[i] = vals[c];
            else if constexpr (PREC == kF16) ((uint16_t*)F)[i] = (uint16_t)f32_to_f16(vals[c]);
            else if constexpr (PREC == kBF16) ((uint16_t*)F)[i] = (uint16_t)f32_to_bf16(vals[c]);
            else ((uint8_t*)F)[i] = (uint8_t)f32_to_e4m3_rne(vals[c] * inv);
        }
    }
}

hipError_t launch_init(void* F, float* scales, int64_t rows, int k, int prec, float s, uint64_t seed,
                       uint32_t problem, uint32_t stream_tag, hipStream_t st)
{
    if (prec == kBF16SR) prec = kBF16;   // same storage; rounding differs only in the update
    if (rows <= 0) return hipSuccess;
    int64_t blocks64 = (rows + 3) / 4;
    int blocks = (int)(blocks64 > 65536 ? 65536 : blocks64);
    Key key = make_key(seed, problem);
    switch (prec) {
    case kF32: hipLaunchKernelGGL(init_factors_kernel<kF32>, dim3(blocks), dim3(256), 0, st, F, scales, rows, k, s, key, stream_tag); break;
    case kF16: hipLaunchKernelGGL(init_factors_kernel<kF16>, dim3(blocks), dim3(256), 0, st, F, scales, rows, k, s, key, stream_tag); break;
    case kBF16: hipLaunchKernelGGL(init_factors_kernel<kBF16>, dim3(blocks), dim3(256), 0, st, F, scales, rows, k, s, key, stream_tag); break;
    case kE4M3: hipLaunchKernelGGL(init_factors_kernel<kE4M3>, dim3(blocks), dim3(256), 0, st, F, scales, rows, k, s, key, stream_tag); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP8 between-epoch rescale.  One wave per row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void fp8_rescale_kernel(uint8_t* F, float* scales, int64_t rows, int k)
{
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    for (int64_t r = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < rows; r += nwaves) {
        const float old = scales[r];
        float vals[4];
        float amax = 0.0f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int e = lane + 64 * c;
            vals[c] = 0.0f;
            if (e < k) {
                vals[c] = e4m3_to_f32(F[r * k + e]) * old;
                amax = fmaxf(amax, fabsf(vals[c]));
            }
        }
        amax = wave_max(amax);
        if (!(amax > 0.0f)) continue;
        const float sc = pow2_scale(amax, kFp8EpochTarget);
        const float inv = 1.0f / sc;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int e = lane + 64 * c;
            if (e < k) F[r * k + e] = (uint8_t)f32_to_e4m3_rne(vals[c] * inv);
        }
        if (lane == 0) scales[r] = sc;
    }
}

hipError_t launch_fp8_rescale(uint8_t* F, float* scales, int64_t rows, int k, hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    int64_t blocks64 = (rows + 3) / 4;
    int blocks = (int)(blocks64 > 65536 ? 65536 : blocks64);
    hipLaunchKernelGGL(fp8_rescale_kernel, dim3(blocks), dim3(256), 0, st, F, scales, rows, k);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Widening helpers for the export and RMSE kernels.
// ---------------------------------------------------------------------------
template <int PREC>
__device__ inline float widen_at(const void* F, const float* scales, int64_t row, int k, int e)
{
    const int64_t i = row * k + e;
    if constexpr (PREC == kF32) return ((const float*)F)[i];
    else if constexpr (PREC == kF16) return f16_to_f32(((const uint16_t*)F)[i]);
    else if constexpr (PREC == kBF16) return bf16_to_f32(((const uint16_t*)F)[i]);
    else return e4m3_to_f32(((const uint8_t*)F)[i]) * scales[row];
}

template <int PREC>
__global__ __launch_bounds__(256) void export_kernel(const void* F, const float* scales, int64_t rows, int k,
                                                     float mult, const int32_t* map, float* out)
{
    const int64_t total = rows * k;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / k;
        const int e = (int)(i - r * k);
        const int64_t src = map ? (int64_t)map[r] : r;
        out[i] = __fmul_rn(widen_at<PREC>(F, scales, src, k, e), mult);
    }
}

hipError_t launch_export(const void* F, const float* scales, int64_t rows, int k, int prec, float mult,
                         const int32_t* map, float* out, hipStream_t st)
{
    if (prec == kBF16SR) prec = kBF16;   // same storage; rounding differs only in the update
    if (rows <= 0) return hipSuccess;
    int64_t total = rows * k;
    int blocks = (int)((total + 255) / 256 > 65536 ? 65536 : (total + 255) / 256);
    switch (prec) {
    case kF32: hipLaunchKernelGGL(export_kernel<kF32>, dim3(blocks), dim3(256), 0, st, F, scales, rows, k, mult, map, out); break;
    case kF16: hipLaunchKernelGGL(export_kernel<kF16>, dim3(blocks), dim3(256), 0, st, F, scales, rows, k, mult, map, out); break;
    case kBF16: hipLaunchKernelGGL(export_kernel<kBF16>, dim3(blocks), dim3(256), 0, st, F, scales, rows, k, mult, map, out); break;
    case kE4M3: hipLaunchKernelGGL(export_kernel<kE4M3>, dim3(blocks), dim3(256), 0, st, F, scales, rows, k, mult, map, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Test evaluation: one thread per test rating, mf_predict's sequential FP32
// dot of the exported values with separate multiply and add (std::inner_product,
// test/mf.cpp:3794-3812, b for out-of-range ids or a NaN dot), then the
// criterion's per-rating term with the reference's arithmetic; FP64 sums,
// reduced per block and then in block order.
//   CRIT 0 RMSE     float e = r - z; e*e            (calc_rmse, mf.cpp:3814-3839)
//        1 MAE      fabsf(r - z)                    (calc_mae, mf.cpp:3841-3855)
//        2 GKL      float r*logf(r/z) - r + z       (calc_gkl, mf.cpp:3857-3872)
//        5 LOGLOSS  log(1.0 + double(expf(-+z)))    (calc_logloss, mf.cpp:3874-3892)
//        6 ACC      r > 0 ? z > 0 : z < 0           (calc_accuracy, mf.cpp:3894-3912)
// ---------------------------------------------------------------------------
constexpr int kRmseThreads = 256;

template <int CRIT>
__device__ inline double eval_term(float r, float z)
{
    if constexpr (CRIT == 0) {
        const float e = __fsub_rn(r, z);
        return (double)__fmul_rn(e, e);
    } else if constexpr (CRIT == 1) {
        return (double)fabsf(__fsub_rn(r, z));
    } else if constexpr (CRIT == 2) {
        const float t = __fmul_rn(r, logf(__fdiv_rn(r, z)));
        return (double)__fadd_rn(__fsub_rn(t, r), z);
    } else if constexpr (CRIT == 5) {
        return r > 0.0f ? log(1.0 + (double)expf(-z)) : log(1.0 + (double)expf(z));
    } else {
        return r > 0.0f ? (z > 0.0f ? 1.0 : 0.0) : (z < 0.0f ? 1.0 : 0.0);
    }
}

template <int PREC, int CRIT>
__global__ __launch_bounds__(kRmseThreads) void eval_kernel(const int32_t* tu, const int32_t* tv, const float* tr,
                                                            int64_t nnz, const void* P, const float* ps,
                                                            const void* Q, const float* qs, int32_t m, int32_t n,
                                                            int k, float mult, float b, double* partials)
{
    __shared__ double red[kRmseThreads];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t u = tu[i], v = tv[i];
        float z;
        if (u < 0 || u >= m || v < 0 || v >= n) {
            z = b;
        } else {
            z = 0.0f;
            for (int e = 0; e < k; ++e) {
                const float pe = __fmul_rn(widen_at<PREC>(P, ps, u, k, e), mult);
                const float qe = __fmul_rn(widen_at<PREC>(Q, qs, v, k, e), mult);
                z = __fadd_rn(z, __fmul_rn(pe, qe));
            }
            if (__builtin_isnan(z)) z = b;
        }
        acc += eval_term<CRIT>(tr[i], z);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kRmseThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Sum n doubles in a fixed order with one block (tree over a strided split).
__global__ __launch_bounds__(256) void sum_doubles_kernel(const double* in, int64_t n, double* out)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) acc += in[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

int rmse_blocks(int64_t nnz)
{
    int64_t b = (nnz + kRmseThreads - 1) / kRmseThreads;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return (int)b;
}

// ---------------------------------------------------------------------------
// Ranking criteria (calc_mpr_auc, test/mf.cpp:3914-4034): one block per row
// that has test positives.  Every column of the row is scored with the same
// sequential FP32 dot as eval_kernel (b outside the model); each (negative
// column, positive) pair counts towards MPR when the positive scores <= the
// negative and towards AUC otherwise.  Positives are scored into LDS in
// chunks; a column is a positive when binary search finds it in the row's
// sorted positive columns.  Writes u_mpr / (n - pos) and
// u_auc / (n - pos) / pos per row.
// ---------------------------------------------------------------------------
constexpr int kRankChunk = 2048;

template <int PREC>
__device__ inline float rank_score(const void* P, const float* ps, const void* Q, const float* qs, int32_t m,
                                   int32_t n, int k, float mult, float b, int64_t u, int64_t v)
{
    if (u >= m || v >= n) return b;
    float z = 0.0f;
    for (int e = 0; e < k; ++e) {
        const float pe = __fmul_rn(widen_at<PREC>(P, ps, u, k, e), mult);
        const float qe = __fmul_rn(widen_at<PREC>(Q, qs, v, k, e), mult);
        z = __fadd_rn(z, __fmul_rn(pe, qe));
    }
    return __builtin_isnan(z) ? b : z;
}

template <int PREC>
__global__ __launch_bounds__(256) void rank_kernel(const int32_t* rows, const int64_t* pbeg, const int32_t* pcols,
                                                   int32_t nrows, int64_t ncols, const void* P, const float* ps,
                                                   const void* Q, const float* qs, int32_t m, int32_t n, int k,
                                                   float mult, float b, int transpose, double* out_mpr,
                                                   double* out_auc)
{
    __shared__ float s_pos[kRankChunk];
    __shared__ unsigned long long red[2][256];
    for (int32_t ri = blockIdx.x; ri < nrows; ri += gridDim.x) {
        const int64_t row = rows[ri], beg = pbeg[ri], cnt = pbeg[ri + 1] - beg;
        unsigned long long c_mpr = 0, c_auc = 0;
        for (int64_t c0 = 0; c0 < cnt; c0 += kRankChunk) {
            const int cc = (int)(cnt - c0 < kRankChunk ? cnt - c0 : kRankChunk);
            __syncthreads();
            for (int t = threadIdx.x; t < cc; t += blockDim.x) {
                const int64_t col = pcols[beg + c0 + t];
                s_pos[t] = transpose ? rank_score<PREC>(P, ps, Q, qs, m, n, k, mult, b, col, row)
                                     : rank_score<PREC>(P, ps, Q, qs, m, n, k, mult, b, row, col);
            }
            __syncthreads();
            for (int64_t j = threadIdx.x; j < ncols; j += blockDim.x) {
                int64_t lo = 0, hi = cnt;   // is j one of the row's positives?
                while (lo < hi) {
                    const int64_t mid = (lo + hi) >> 1;
                    if (pcols[beg + mid] < j) lo = mid + 1;
                    else hi = mid;
                }
                if (lo < cnt && pcols[beg + lo] == j) continue;
                const float sj = transpose ? rank_score<PREC>(P, ps, Q, qs, m, n, k, mult, b, j, row)
                                           : rank_score<PREC>(P, ps, Q, qs, m, n, k, mult, b, row, j);
                for (int t = 0; t < cc; ++t) {
                    if (s_pos[t] <= sj) ++c_mpr;
                    else ++c_auc;
                }
            }
        }
        red[0][threadIdx.x] = c_mpr;
        red[1][threadIdx.x] = c_auc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                red[0][threadIdx.x] += red[0][threadIdx.x + w];
                red[1][threadIdx.x] += red[1][threadIdx.x + w];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const double neg = (double)(ncols - cnt);
            out_mpr[ri] = (double)red[0][0] / neg;
            out_auc[ri] = (double)red[1][0] / neg / (double)cnt;
        }
        __syncthreads();
    }
}

hipError_t launch_rank(const int32_t* rows, const int64_t* pbeg, const int32_t* pcols, int32_t nrows, int64_t ncols,
                       const void* P, const float* ps, const void* Q, const float* qs, int32_t m, int32_t n, int k,
                       int prec, float mult, float b, int transpose, double* out_mpr, double* out_auc, hipStream_t st)
{
    if (nrows <= 0) return hipSuccess;
    if (prec == kBF16SR) prec = kBF16;
    const int blocks = nrows < 8192 ? nrows : 8192;
    switch (prec) {
    case kF32: hipLaunchKernelGGL(rank_kernel<kF32>, dim3(blocks), dim3(256), 0, st, rows, pbeg, pcols, nrows, ncols, P, ps, Q, qs, m, n, k, mult, b, transpose, out_mpr, out_auc); break;
    case kF16: hipLaunchKernelGGL(rank_kernel<kF16>, dim3(blocks), dim3(256), 0, st, rows, pbeg, pcols, nrows, ncols, P, ps, Q, qs, m, n, k, mult, b, transpose, out_mpr, out_auc); break;
    case kBF16: hipLaunchKernelGGL(rank_kernel<kBF16>, dim3(blocks), dim3(256), 0, st, rows, pbeg, pcols, nrows, ncols, P, ps, Q, qs, m, n, k, mult, b, transpose, out_mpr, out_auc); break;
    case kE4M3: hipLaunchKernelGGL(rank_kernel<kE4M3>, dim3(blocks), dim3(256), 0, st, rows, pbeg, pcols, nrows, ncols, P, ps, Q, qs, m, n, k, mult, b, transpose, out_mpr, out_auc); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

bool eval_supported(int criterion)
{
    return criterion == 0 || criterion == 1 || criterion == 2 || criterion == 5 || criterion == 6 ||
           (criterion >= 10 && criterion <= 13);
}

template <int PREC>
static hipError_t launch_eval_prec(int crit, int blocks, const int32_t* tu, const int32_t* tv, const float* tr,
                                   int64_t nnz, const void* P, const float* ps, const void* Q, const float* qs,
                                   int32_t m, int32_t n, int k, float mult, float b, double* partials, hipStream_t st)
{
#define CUMF_EVAL(C)                                                                                          \
    case C:                                                                                                   \
        hipLaunchKernelGGL((eval_kernel<PREC, C>), dim3(blocks), dim3(kRmseThreads), 0, st, tu, tv, tr, nnz, P, ps, Q, \
                           qs, m, n, k, mult, b, partials);                                                    \
        break;
    switch (crit) {
        CUMF_EVAL(0)
        CUMF_EVAL(1)
        CUMF_EVAL(2)
        CUMF_EVAL(5)
        CUMF_EVAL(6)
    default: return hipErrorInvalidValue;
    }
#undef CUMF_EVAL
    return hipGetLastError();
}

hipError_t launch_eval(int criterion, const int32_t* tu, const int32_t* tv, const float* tr, int64_t nnz,
                       const void* P, const float* ps, const void* Q, const float* qs, int32_t m, int32_t n, int k,
                       int prec, float mult, float b, double* partials, double* out, hipStream_t st)
{
    if (prec == kBF16SR) prec = kBF16;   // same storage; rounding differs only in the update
    const int blocks = rmse_blocks(nnz);
    hipError_t e;
    switch (prec) {
    case kF32: e = launch_eval_prec<kF32>(criterion, blocks, tu, tv, tr, nnz, P, ps, Q, qs, m, n, k, mult, b, partials, st); break;
    case kF16: e = launch_eval_prec<kF16>(criterion, blocks, tu, tv, tr, nnz, P, ps, Q, qs, m, n, k, mult, b, partials, st); break;
    case kBF16: e = launch_eval_prec<kBF16>(criterion, blocks, tu, tv, tr, nnz, P, ps, Q, qs, m, n, k, mult, b, partials, st); break;
    case kE4M3: e = launch_eval_prec<kE4M3>(criterion, blocks, tu, tv, tr, nnz, P, ps, Q, qs, m, n, k, mult, b, partials, st); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sum_doubles_kernel, dim3(1), dim3(256), 0, st, partials, (int64_t)blocks, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Preprocessing kernels.
// ---------------------------------------------------------------------------
static int grid_for(int64_t n, int threads)
{
    int64_t b = (n + threads - 1) / threads;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (int)b;
}

// AoS 12-byte records -> three arrays.
__global__ void split_kernel(const int32_t* aos, int64_t nnz, int32_t* u, int32_t* v, float* r)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        u[i] = aos[3 * i];
        v[i] = aos[3 * i + 1];
        r[i] = __int_as_float(aos[3 * i + 2]);
    }
}

hipError_t launch_split(const void* aos, int64_t nnz, int32_t* u, int32_t* v, float* r, hipStream_t st)
{
    if (nnz <= 0) return hipSuccess;
    hipLaunchKernelGGL(split_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, st, (const int32_t*)aos, nnz, u, v, r);
    return hipGetLastError();
}

// Range check of ids: counts records with u outside [0,m) or v outside [0,n).
__global__ void check_ids_kernel(const int32_t* u, const int32_t* v, int64_t nnz, int32_t m, int32_t n,
                                 unsigned long long* bad)
{
    unsigned long long c = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        c += (u[i] < 0 || u[i] >= m || v[i] < 0 || v[i] >= n) ? 1ull : 0ull;
    if (c) atomicAdd(bad, c);
}

hipError_t launch_check_ids(const int32_t* u, const int32_t* v, int64_t nnz, int32_t m, int32_t n,
                            unsigned long long* bad, hipStream_t st)
{
    if (nnz <= 0) return hipSuccess;
    hipLaunchKernelGGL(check_ids_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, st, u, v, nnz, m, n, bad);
    return hipGetLastError();
}

// u <- pmap[u], v <- qmap[v]; ids outside [0,m) x [0,n) are left alone.
__global__ void remap_kernel(int32_t* u, int32_t* v, int64_t nnz, const int32_t* pmap, const int32_t* qmap,
                             int32_t m, int32_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = u[i], c = v[i];
        if (a >= 0 && a < m) u[i] = pmap[a];
        if (c >= 0 && c < n) v[i] = qmap[c];
    }
}

hipError_t launch_remap(int32_t* u, int32_t* v, int64_t nnz, const int32_t* pmap, const int32_t* qmap, int32_t m,
                        int32_t n, hipStream_t st)
{
    if (nnz <= 0) return hipSuccess;
    hipLaunchKernelGGL(remap_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, st, u, v, nnz, pmap, qmap, m, n);
    return hipGetLastError();
}

__global__ void scale_kernel(float* r, int64_t nnz, float s)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        r[i] = __fmul_rn(r[i], s);
}

hipError_t launch_scale(float* r, int64_t nnz, float s, hipStream_t st)
{
    if (nnz <= 0) return hipSuccess;
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, st, r, nnz, s);
    return hipGetLastError();
}

// Sum r and r*r in double: per-block partials (fixed order), then one block.
__global__ __launch_bounds__(256) void moments_kernel(const float* r, int64_t nnz, double* partials)
{
    __shared__ double s1[256], s2[256];
    double a = 0.0, b = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const double x = (double)r[i];
        a += x;
        b += x * x;
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s1[0];
        partials[gridDim.x + blockIdx.x] = s2[0];
    }
}

int moments_blocks(int64_t nnz) { return grid_for(nnz, 256) > 2048 ? 2048 : grid_for(nnz, 256); }

hipError_t launch_moments(const float* r, int64_t nnz, double* partials, double* out2, hipStream_t st)
{
    const int blocks = moments_blocks(nnz);
    hipLaunchKernelGGL(moments_kernel, dim3(blocks), dim3(256), 0, st, r, nnz, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sum_doubles_kernel, dim3(1), dim3(256), 0, st, partials, (int64_t)blocks, out2);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sum_doubles_kernel, dim3(1), dim3(256), 0, st, partials + blocks, (int64_t)blocks, out2 + 1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable counting-sort partition by block id useg[u]*vy + vseg[v].
// A wave owns a contiguous chunk and walks it 64 records at a time; a
// record's rank among equal ids inside its group of 64 comes from a readlane
// scan, so the order inside every block is the input order.
//   pass 1: per-chunk histogram hist[chunk*nb + b]
//   pass 2: per-block column scan -> write offsets off[chunk*nb + b]
//   pass 3: scatter
// ---------------------------------------------------------------------------
__device__ inline int block_of(int32_t u, int32_t v, const int32_t* useg, const int32_t* vseg, int vy)
{
    return useg[u] * vy + vseg[v];
}

__global__ __launch_bounds__(64) void part_count_kernel(const int32_t* u, const int32_t* v, int64_t nnz, int64_t chunk,
                                                        const int32_t* useg, const int32_t* vseg, int vy, int nb,
                                                        int32_t* hist)
{
    extern __shared__ __align__(16) char dyn_smem[];
    int32_t* cnt = (int32_t*)dyn_smem;
    for (int b = threadIdx.x; b < nb; b += 64) cnt[b] = 0;
    __syncthreads();
    const int64_t c = blockIdx.x;
    const int64_t beg = c * chunk, end = beg + chunk < nnz ? beg + chunk : nnz;
    for (int64_t i = beg + threadIdx.x; i < end; i += 64) atomicAdd(&cnt[block_of(u[i], v[i], useg, vseg, vy)], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += 64) hist[c * nb + b] = cnt[b];
}

// One thread per block id: exclusive scan of its column over chunks, plus the
// block base (sum of all earlier blocks' totals, given in base[]).
__global__ void part_scan_kernel(const int32_t* hist, int64_t nchunks, int nb, const int64_t* base, int64_t* off)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int64_t run = base[b];
    for (int64_t c = 0; c < nchunks; ++c) {
        off[c * nb + b] = run;
        run += hist[c * nb + b];
    }
}

// Column totals: tot[b] = sum over chunks of hist[c*nb+b].
__global__ void part_total_kernel(const int32_t* hist, int64_t nchunks, int nb, int64_t* tot)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int64_t s = 0;
    for (int64_t c = 0; c < nchunks; ++c) s += hist[c * nb + b];
    tot[b] = s;
}

__global__ __launch_bounds__(64) void part_scatter_kernel(const int32_t* u, const int32_t* v, const float* r,
                                                          int64_t nnz, int64_t chunk, const int32_t* useg,
                                                          const int32_t* vseg, int vy, int nb, const int64_t* off,
                                                          int32_t* uo, int32_t* vo, float* ro)
{
    extern __shared__ __align__(16) char dyn_smem[];
    int64_t* pos = (int64_t*)dyn_smem;
    const int64_t c = blockIdx.x;
    for (int b = threadIdx.x; b < nb; b += 64) pos[b] = off[c * nb + b];
    __syncthreads();
    const int64_t beg = c * chunk, end = beg + chunk < nnz ? beg + chunk : nnz;
    const int lane = threadIdx.x;
    for (int64_t g = beg; g < end; g += 64) {
        const int64_t i = g + lane;
        const bool valid = i < end;
        const int bid = valid ? block_of(u[i], v[i], useg, vseg, vy) : -1;
        int rank = 0;
        const int cnt = (int)(end - g < 64 ? end - g : 64);
        for (int j = 0; j < cnt; ++j) {
            const int bj = __builtin_amdgcn_readlane(bid, j);
            rank += (j < lane && bj == bid) ? 1 : 0;
        }
        int64_t dst = 0;
        if (valid) dst = pos[bid] + rank;
        __syncthreads();
        if (valid) {
            uo[dst] = u[i];
            vo[dst] = v[i];
            ro[dst] = r[i];
            atomicAdd((unsigned long long*)&pos[bid], 1ull);
        }
        __syncthreads();
    }
}

int64_t part_chunk(int64_t nnz)
{
    // about 2048 chunks, at least 4096 records each
    int64_t c = (nnz + 2047) / 2048;
    if (c < 4096) c = 4096;
    return c;
}

hipError_t launch_partition(const int32_t* u, const int32_t* v, const float* r, int64_t nnz, const int32_t* useg,
                            const int32_t* vseg, int vy, int nb, int32_t* hist, int64_t* tot, int64_t* base,
                            int64_t* off, int32_t* uo, int32_t* vo, float* ro, hipStream_t st,
                            int64_t* host_tot)
{
    const int64_t chunk = part_chunk(nnz);
    const int64_t nchunks = (nnz + chunk - 1) / chunk;
    hipLaunchKernelGGL(part_count_kernel, dim3((unsigned)nchunks), dim3(64), nb * sizeof(int32_t), st, u, v, nnz,
                       chunk, useg, vseg, vy, nb, hist);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(part_total_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, hist, nchunks, nb, tot);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(host_tot, tot, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    // block bases on the host (nb is small)
    int64_t* hb = host_tot + nb;
    hb[0] = 0;
    for (int b = 1; b < nb; ++b) hb[b] = hb[b - 1] + host_tot[b - 1];
    if ((e = hipMemcpyAsync(base, hb, sizeof(int64_t) * nb, hipMemcpyHostToDevice, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(part_scan_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, hist, nchunks, nb, base, off);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(part_scatter_kernel, dim3((unsigned)nchunks), dim3(64), nb * sizeof(int64_t), st, u, v, r, nnz,
                       chunk, useg, vseg, vy, nb, off, uo, vo, ro);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

int64_t part_hist_entries(int64_t nnz, int nb)
{
    const int64_t chunk = part_chunk(nnz);
    return ((nnz + chunk - 1) / chunk) * (int64_t)nb;
}

// Segment lookup tables: seg[i] for i < rows given ascending boundaries
// bnd[0..ns] (bnd[0] = 0, bnd[ns] = rows).
__global__ void seg_table_kernel(int32_t* seg, int64_t rows, const int64_t* bnd, int ns)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = ns - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (bnd[mid] <= i) lo = mid; else hi = mid - 1;
        }
        seg[i] = lo;
    }
}

hipError_t launch_seg_table(int32_t* seg, int64_t rows, const int64_t* bnd, int ns, hipStream_t st)
{
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(seg_table_kernel, dim3(grid_for(rows, 256)), dim3(256), 0, st, seg, rows, bnd, ns);
    return hipGetLastError();
}

// Row/column degree histograms (for nnz-balanced stratum boundaries).
__global__ void degree_kernel(const int32_t* ids, int64_t nnz, int32_t* deg)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&deg[ids[i]], 1);
}

// Degrees of the ids in [lo, lo + kDegreeBins) counted in LDS first: a few
// hot ids (Zipf items) would serialise on global atomics.
constexpr int kDegreeBins = 16384;

__global__ __launch_bounds__(256) void degree_lds_kernel(const int32_t* ids, int64_t nnz, int32_t lo, int32_t* deg)
{
    __shared__ int32_t h[kDegreeBins];
    for (int b = threadIdx.x; b < kDegreeBins; b += blockDim.x) h[b] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t d = (uint32_t)(ids[i] - lo);
        if (d < (uint32_t)kDegreeBins) atomicAdd(&h[d], 1);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kDegreeBins; b += blockDim.x)
        if (h[b]) atomicAdd(&deg[lo + b], h[b]);
}

hipError_t launch_degree(const int32_t* ids, int64_t nnz, int32_t* deg, int32_t id_range, hipStream_t st)
{
    if (nnz <= 0) return hipSuccess;
    if (id_range <= 8 * kDegreeBins) {
        const int blocks = grid_for(nnz, 256) < 2048 ? grid_for(nnz, 256) : 2048;
        for (int32_t lo = 0; lo < id_range; lo += kDegreeBins) {
            hipLaunchKernelGGL(degree_lds_kernel, dim3(blocks), dim3(256), 0, st, ids, nnz, lo, deg);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    hipLaunchKernelGGL(degree_kernel, dim3(grid_for(nnz, 256)), dim3(256), 0, st, ids, nnz, deg);
    return hipGetLastError();
}

}  // namespace cumf
