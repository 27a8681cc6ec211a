// Launch-shape resolution, the learning-rate table and the graph-key hash,
// shared by the context's training entry points (capi.hip), the
// allocation-free one (noalloc.hip) and batched training (batch.hip).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/cumf_sgd_amd.h"

namespace cumf {

// The resolved shape of an update launch: G lanes per sample, S = 64/G
// samples per wave, L records per window, W workers, `blocks` workgroups,
// D row pairs in flight per worker.
struct Shape {
    int G, S, L, W, blocks, D;
};

// What the shape depends on besides the options: the factor length and
// format, the device's CU count and the rating count.
struct ShapeInput {
    int k, prec, cu_count;
    int64_t nnz;
};

// The default worker count for waves_per_cu waves per CU: the fewest workers
// that keep the same windows per worker (update_count = ceil(nnz / (W * L)),
// singleGPU/sgd_kernel.cu:203), in whole workgroups.
int default_workers(const ShapeInput& in, int S, int L, int waves_per_cu);

// Resolve opts' zero fields to the defaults and check the shape; the grid is
// rounded up to a multiple of xcd_multiple workgroups.  Returns CUMF_OK or
// CUMF_ERR_ARG (with the last-error text set).
int resolve_shape(const ShapeInput& in, const cumf_train_opts* o, Shape* sh, int xcd_multiple);

// Hogwild and gridded launches whose factors (P and Q together) fit in the
// XCDs' L2s run on one XCD.  Device memory is not kept coherent between the
// eight XCD L2s within a kernel: each L2 keeps its own copy of a row it has
// touched until the kernel ends, so on factors that small the workers of
// different XCDs would train diverging copies of the hot rows for a whole
// launch.  Larger factors are evicted long before that and train on all XCDs.
constexpr int kNumXcd = 8;
constexpr int64_t kOneXcdFactorBytes = 32ll << 20;   // 8 XCDs x 4 MB of L2
bool confine_to_one_xcd(int schedule, int coherent, int64_t m, int64_t n, int k, int prec);

// The default schedule counts an axis as skewed when its busiest row holds at
// least 1/kSkewedShareInv of all ratings.
constexpr int64_t kSkewedShareInv = 100;
// Stratum passes per epoch that the default schedule gives factors small
// enough for one XCD's launch placement.
constexpr int kSmallFactorRounds = 4;
// A P larger than this (four times MI355X's 256 MB Infinity Cache) is read
// from HBM: the default schedule then groups each block's records by row.
constexpr int64_t kHbmResidentP = 1ll << 30;

// Tile launches: the default lane group (32 bytes of a row per lane where
// supported), workers per tile, strata and passes of the default schedule.
int tile_default_group(int k, int prec);
constexpr int kTileWorkers = 32;
// What one window of a tile costs beyond its records, in sample steps of a
// worker: the window's start (a Philox draw), its record staging and the
// wait that drains the worker's memory operations before the next window.
// Measured on MI355X (99M ratings, k = 128): FP32 tiles, bound by the L2,
// train an epoch fastest from about 16 up (5.10 ms against 5.44 ms at 1),
// FP16 tiles, bound by VALU issue, the same from 1 to 64.
constexpr int kTileWindowSteps = 16;
constexpr int kTileStrata = 8;
// The learning-rate table grows in steps of this many epochs, so that an
// epoch loop driven one epoch per call keeps the table where it is.
constexpr int kLrChunk = 1024;
constexpr int kTileFp32Rounds = 2;

// lr[i] = float(alpha / (1 + beta * i^1.5) + lrate) in double for i < n
// (singleGPU/sgd_kernel.cu:137-154).
void lr_table(const cumf_train_opts* o, int n, float* out);

// FNV-1a over n bytes, continuing from h: the key of a captured epoch graph
// (start from 1469598103934665603).
inline uint64_t fnv1a(const void* p, size_t n, uint64_t h)
{
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

}  // namespace cumf
