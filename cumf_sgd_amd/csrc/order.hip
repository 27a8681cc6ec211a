// Record order inside the blocks of a partition: a stable sort of every
// block's records by column (or row) id, so that the records of one column
// follow each other.  The update kernel then finds consecutive samples of a
// worker on the same row and keeps that row in registers instead of storing
// and reloading it.  Keys are (block, id) pairs sorted with rocPRIM's stable
// radix sort; the records are then gathered in that order.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "kernels.h"

namespace cumf {

namespace {

template <typename K>
__global__ void block_key_kernel(const int32_t* __restrict__ ids, int64_t nnz, const int64_t* __restrict__ off, int nb,
                                 uint64_t stride, K* __restrict__ key, uint32_t* __restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    // block b holds records [off[b], off[b + 1]): the last b with off[b] <= i
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    key[i] = (K)((uint64_t)lo * stride + (uint64_t)(uint32_t)ids[i]);
    idx[i] = (uint32_t)i;
}

__global__ void gather_records_kernel(const uint32_t* __restrict__ idx, int64_t nnz, const int32_t* __restrict__ u,
                                      const int32_t* __restrict__ v, const float* __restrict__ r, int32_t* __restrict__ uo,
                                      int32_t* __restrict__ vo, float* __restrict__ ro)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const uint32_t j = idx[i];
    uo[i] = u[j];
    vo[i] = v[j];
    ro[i] = r[j];
}

unsigned bits_for(uint64_t x)
{
    unsigned b = 0;
    while (b < 64 && (x >> b)) ++b;
    return b;
}

// Tile keys: record i of block b (prefix offsets off) whose row u is in tile
// t = tile_of[u] of its row band gets the key (b * T + t) * stride + v, and
// counts[b * T + t] counts the records of each (block, tile).
template <typename K>
__global__ void tile_key_kernel(const int32_t* __restrict__ u, const int32_t* __restrict__ v, int64_t nnz,
                                const int64_t* __restrict__ off, int nb, const int32_t* __restrict__ tile_of, int T,
                                uint64_t stride, K* __restrict__ key, uint32_t* __restrict__ idx,
                                int32_t* __restrict__ counts)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const uint64_t bt = (uint64_t)lo * (uint64_t)T + (uint64_t)tile_of[u[i]];
    key[i] = (K)(bt * stride + (uint64_t)(uint32_t)v[i]);
    idx[i] = (uint32_t)i;
    atomicAdd(counts + bt, 1);
}

template <typename K>
hipError_t sort_pairs(int64_t nnz, unsigned end_bit, void* keys, void* keys_out, uint32_t* idx, uint32_t* idx_out,
                      void** temp, size_t* temp_bytes, hipStream_t st)
{
    hipError_t e;
    size_t need = 0;
    e = rocprim::radix_sort_pairs(nullptr, need, (K*)keys, (K*)keys_out, idx, idx_out, (size_t)nnz, 0u, end_bit, st);
    if (e != hipSuccess) return e;
    if (need > *temp_bytes) {
        if (*temp) (void)hipFree(*temp);
        *temp = nullptr;
        *temp_bytes = 0;
        if ((e = hipMalloc(temp, need)) != hipSuccess) return e;
        *temp_bytes = need;
    }
    return rocprim::radix_sort_pairs(*temp, need, (K*)keys, (K*)keys_out, idx, idx_out, (size_t)nnz, 0u, end_bit, st);
}

template <typename K>
hipError_t sort_keys(int64_t nnz, const int32_t* ids, const int64_t* off, int nb, uint64_t stride, void* keys,
                     void* keys_out, uint32_t* idx, uint32_t* idx_out, void** temp, size_t* temp_bytes, hipStream_t st)
{
    const unsigned end_bit = bits_for((uint64_t)nb * stride);
    const int threads = 256;
    const int64_t grid = (nnz + threads - 1) / threads;
    hipLaunchKernelGGL((block_key_kernel<K>), dim3((unsigned)grid), dim3(threads), 0, st, ids, nnz, off, nb, stride,
                       (K*)keys, idx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return sort_pairs<K>(nnz, end_bit, keys, keys_out, idx, idx_out, temp, temp_bytes, st);
}

template <typename K>
hipError_t sort_tile_keys(int64_t nnz, const int32_t* u, const int32_t* v, const int64_t* off, int nb,
                          const int32_t* tile_of, int T, uint64_t stride, void* keys, void* keys_out, uint32_t* idx,
                          uint32_t* idx_out, int32_t* counts, void** temp, size_t* temp_bytes, hipStream_t st)
{
    const unsigned end_bit = bits_for((uint64_t)nb * (uint64_t)T * stride);
    const int threads = 256;
    const int64_t grid = (nnz + threads - 1) / threads;
    hipLaunchKernelGGL((tile_key_kernel<K>), dim3((unsigned)grid), dim3(threads), 0, st, u, v, nnz, off, nb, tile_of, T,
                       stride, (K*)keys, idx, counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return sort_pairs<K>(nnz, end_bit, keys, keys_out, idx, idx_out, temp, temp_bytes, st);
}

}  // namespace

size_t order_key_bytes(int nb, uint64_t stride) { return (uint64_t)nb * stride >> 32 ? 8 : 4; }

hipError_t launch_order_blocks(int64_t nnz, const int32_t* ids, const int64_t* off, int nb, uint64_t stride, void* keys,
                               void* keys_out, uint32_t* idx, uint32_t* idx_out, const int32_t* u, const int32_t* v,
                               const float* r, int32_t* uo, int32_t* vo, float* ro, void** temp, size_t* temp_bytes,
                               hipStream_t st)
{
    if (nnz <= 0) return hipSuccess;
    hipError_t e = order_key_bytes(nb, stride) == 8
                       ? sort_keys<uint64_t>(nnz, ids, off, nb, stride, keys, keys_out, idx, idx_out, temp, temp_bytes, st)
                       : sort_keys<uint32_t>(nnz, ids, off, nb, stride, keys, keys_out, idx, idx_out, temp, temp_bytes, st);
    if (e != hipSuccess) return e;
    const int threads = 256;
    const int64_t grid = (nnz + threads - 1) / threads;
    hipLaunchKernelGGL(gather_records_kernel, dim3((unsigned)grid), dim3(threads), 0, st, idx_out, nnz, u, v, r, uo, vo,
                       ro);
    return hipGetLastError();
}

}  // namespace cumf

namespace cumf {

hipError_t launch_order_tiles(int64_t nnz, const int32_t* u, const int32_t* v, const float* r, const int64_t* off,
                              int nb, const int32_t* tile_of, int T, uint64_t stride, void* keys, void* keys_out,
                              uint32_t* idx, uint32_t* idx_out, int32_t* counts, int32_t* uo, int32_t* vo, float* ro,
                              void** temp, size_t* temp_bytes, hipStream_t st)
{
    if (nnz <= 0) return hipSuccess;
    hipError_t e = order_key_bytes(nb * T, stride) == 8
                       ? sort_tile_keys<uint64_t>(nnz, u, v, off, nb, tile_of, T, stride, keys, keys_out, idx, idx_out,
                                                  counts, temp, temp_bytes, st)
                       : sort_tile_keys<uint32_t>(nnz, u, v, off, nb, tile_of, T, stride, keys, keys_out, idx, idx_out,
                                                  counts, temp, temp_bytes, st);
    if (e != hipSuccess) return e;
    const int threads = 256;
    const int64_t grid = (nnz + threads - 1) / threads;
    hipLaunchKernelGGL(gather_records_kernel, dim3((unsigned)grid), dim3(threads), 0, st, idx_out, nnz, u, v, r, uo, vo,
                       ro);
    return hipGetLastError();
}

}  // namespace cumf
