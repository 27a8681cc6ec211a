// Launch interface between the C-ABI layer and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.h"

namespace cumf {

// One rating block trained by a contiguous range of workers in a launch.
struct SgdTask {
    const int32_t* u;         // block records (three arrays), global ids
    const int32_t* v;
    const float* r;
    int64_t nnz;
    void* P;                  // factor storage, row-major [rows][k]
    void* Q;
    float* ps;                // FP8 row scales (else unused)
    float* qs;
    const float* lr;          // learning-rate table indexed by epoch
    float lambda_p, lambda_q;
    int32_t worker_begin;     // workers [worker_begin, worker_begin + worker_count)
    int32_t worker_count;
    int32_t update_count;     // windows per worker per epoch
    uint32_t problem;         // Philox key word (batched problems)
    uint32_t tag;             // block id folded into the window counter
    int32_t k;                // factor length (batched launches dispatch on it)
    int32_t wave_count;       // batched launches: waves owned by the task
    int32_t window_begin;     // the task's first window index in the epoch (its windows
                              // are window_begin + [0, update_count))
    int32_t row_begin;        // tile launches: the task's P rows [row_begin, row_begin + row_count),
    int32_t row_count;        // held in the workgroup's LDS for the launch
    int32_t batch_len;        // records per window; tile launches read it per task (<= the
                              // launch's batch_len), other launches use the launch's
};

struct SgdLaunch {
    const SgdTask* tasks;     // device array
    int32_t num_tasks;
    int32_t epoch_begin, epoch_end;
    int32_t batch_len;        // <= 128
    uint64_t seed;
    int32_t num_xcd;          // > 1: workers numbered XCD-major over blockIdx % num_xcd
    int32_t task_off32;       // batched launches: 1 when every task's factors are under 4 GB
                              // (rows then addressed as the task's base + a 32-bit offset)
    const int32_t* epoch_dev; // non-null: the launch's first epoch is *epoch_dev (graph replay)
    // non-null: every task's factors, each matrix under 4 GB; the kernel then
    // addresses rows as this launch-wide base plus a 32-bit offset (one
    // instruction per address instead of three)
    void* P;
    void* Q;
    float* ps;
    float* qs;
};

constexpr int kSgdWavesPerBlock = 4;
constexpr int kSgdMaxBatch = 128;

bool sgd_shape_supported(int k, int prec, int G);
int sgd_default_group(int k, int prec);
// D: software-pipeline depth (2, 4 or 8; must divide the batch length)
hipError_t launch_sgd(const SgdLaunch& a, int blocks, int k, int prec, int G, bool coh, int D, hipStream_t s);
// Tile launch: one workgroup of `waves` waves (4, 8 or 16) per task (a tile:
// the ratings of row_count consecutive P rows within one column range), the
// tile's P rows held in LDS for the launch.  Lane group: the default one.
// Workgroups go XCD-major: task (x * (blocks / num_xcd) + j) runs on XCD x.
constexpr int kSgdTileWaves = 8;       // default workgroup
constexpr int kSgdTileMaxWaves = 16;
// Largest P tile a workgroup holds (bytes of LDS), and the default one: two
// workgroups (16 waves) per CU with the record staging of 32-record windows.
constexpr int kSgdTileMaxBytes = 128 * 1024;
constexpr int kSgdTileDefaultBytes = 64 * 1024;
// Longest window of a tile launch (records): each wave stages this many per
// worker, in a fixed number of loads.
constexpr int kSgdTileMaxBatch = 32;
// LDS bytes of a tile launch: the tile, then each wave's staged records.
size_t sgd_tile_lds_bytes(int k, int prec, int G, int rows, int waves);
size_t sgd_tile_lds_limit();
hipError_t launch_sgd_tile(const SgdLaunch& a, int blocks, int k, int prec, int G, int D, int rows, int waves,
                           hipStream_t s);
// Batched launch: tasks of mixed k (32, 64, 128), 16 bytes of a row per lane (8 to 64 lanes).
hipError_t launch_sgd_batched(const SgdLaunch& a, int blocks, int prec, int D, hipStream_t s);
int sgd_batch_group(int k, int prec);
// Lanes per worker in which a batch's wave budget is counted (k/2; FP8 k/4).
int sgd_batch_budget_group(int k, int prec);

// *p = value, and *p += 1 (device-side epoch counter of a captured epoch loop)
hipError_t launch_set_int(int32_t* p, int32_t value, hipStream_t st);
hipError_t launch_advance_int(int32_t* p, hipStream_t st);

hipError_t launch_init(void* F, float* scales, int64_t rows, int k, int prec, float s, uint64_t seed,
                       uint32_t problem, uint32_t stream_tag, hipStream_t st);
hipError_t launch_fp8_rescale(uint8_t* F, float* scales, int64_t rows, int k, hipStream_t st);
hipError_t launch_export(const void* F, const float* scales, int64_t rows, int k, int prec, float mult,
                         const int32_t* map, float* out, hipStream_t st);
int rmse_blocks(int64_t nnz);
bool eval_supported(int criterion);
// The ranking criteria end to end (host grouping + rank kernel): MPR and AUC.
int rank_eval(hipStream_t st, const int32_t* tu, const int32_t* tv, const float* tr, int64_t tnnz, const void* P,
              const float* ps, const void* Q, const float* qs, int32_t m, int32_t n, int k, int prec, float mult,
              float b, bool transpose, double* mpr, double* auc);
// calc_mpr_auc per listed row: rows[i] with sorted positive columns
// pcols[pbeg[i] .. pbeg[i+1]) over ncols columns; writes u_mpr/(ncols-pos)
// and u_auc/(ncols-pos)/pos per row.
hipError_t launch_rank(const int32_t* rows, const int64_t* pbeg, const int32_t* pcols, int32_t nrows, int64_t ncols,
                       const void* P, const float* ps, const void* Q, const float* qs, int32_t m, int32_t n, int k,
                       int prec, float mult, float b, int transpose, double* out_mpr, double* out_auc, hipStream_t st);
// criterion: 0 RMSE (sum of squares), 1 MAE, 2 GKL, 5 LOGLOSS, 6 ACC (sums of terms)
hipError_t launch_eval(int criterion, const int32_t* tu, const int32_t* tv, const float* tr, int64_t nnz,
                       const void* P, const float* ps, const void* Q, const float* qs, int32_t m, int32_t n, int k,
                       int prec, float mult, float b, double* partials, double* out, hipStream_t st);
hipError_t launch_split(const void* aos, int64_t nnz, int32_t* u, int32_t* v, float* r, hipStream_t st);
hipError_t launch_check_ids(const int32_t* u, const int32_t* v, int64_t nnz, int32_t m, int32_t n,
                            unsigned long long* bad, hipStream_t st);
hipError_t launch_remap(int32_t* u, int32_t* v, int64_t nnz, const int32_t* pmap, const int32_t* qmap, int32_t m,
                        int32_t n, hipStream_t st);
hipError_t launch_scale(float* r, int64_t nnz, float s, hipStream_t st);
int moments_blocks(int64_t nnz);
hipError_t launch_moments(const float* r, int64_t nnz, double* partials, double* out2, hipStream_t st);
int64_t part_hist_entries(int64_t nnz, int nb);
hipError_t launch_partition(const int32_t* u, const int32_t* v, const float* r, int64_t nnz, const int32_t* useg,
                            const int32_t* vseg, int vy, int nb, int32_t* hist, int64_t* tot, int64_t* base,
                            int64_t* off, int32_t* uo, int32_t* vo, float* ro, hipStream_t st, int64_t* host_tot);
hipError_t launch_seg_table(int32_t* seg, int64_t rows, const int64_t* bnd, int ns, hipStream_t st);
hipError_t launch_degree(const int32_t* ids, int64_t nnz, int32_t* deg, int32_t id_range, hipStream_t st);
// Stable sort of each block's records (blocks = prefix offsets off[0..nb]) by
// ids (< stride), gathered into uo/vo/ro.  keys/keys_out hold
// order_key_bytes(nb, stride) bytes per record, idx/idx_out 4; temp is grown
// (hipMalloc) as the sort needs.
size_t order_key_bytes(int nb, uint64_t stride);
hipError_t launch_order_blocks(int64_t nnz, const int32_t* ids, const int64_t* off, int nb, uint64_t stride, void* keys,
                               void* keys_out, uint32_t* idx, uint32_t* idx_out, const int32_t* u, const int32_t* v,
                               const float* r, int32_t* uo, int32_t* vo, float* ro, void** temp, size_t* temp_bytes,
                               hipStream_t st);
// Stable sort of each block's records by (tile of the row, column), the
// tile of row u being tile_of[u] (< T) within its row band; counts (nb * T
// entries, zeroed by the caller) receives the records per (block, tile).
// keys/keys_out hold order_key_bytes(nb * T, stride) bytes per record.
hipError_t launch_order_tiles(int64_t nnz, const int32_t* u, const int32_t* v, const float* r, const int64_t* off,
                              int nb, const int32_t* tile_of, int T, uint64_t stride, void* keys, void* keys_out,
                              uint32_t* idx, uint32_t* idx_out, int32_t* counts, int32_t* uo, int32_t* vo, float* ro,
                              void** temp, size_t* temp_bytes, hipStream_t st);

}  // namespace cumf
