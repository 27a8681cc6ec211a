/*
 * cumf_sgd_amd: MI355X (gfx950) matrix-factorization SGD, C ABI.
 *
 * This header is the drop-in boundary.  It replaces the reference's host/GPU
 * split, `void sgd_update_k128(Parameter, mf_model*, mf_problem*, float)`
 * (singleGPU/sgd_kernel.h:17, called once from singleGPU/sgd.cu:682), and the
 * library-level pipeline `sgd_train` (singleGPU/sgd.h:141) around it.
 *
 * Conventions (all entry points):
 *  - every function returns CUMF_OK (0) or a positive CUMF_ERR_* code and
 *    never calls exit(); cumf_last_error() returns the message of the last
 *    failure on the calling thread;
 *  - host memory belongs to the caller; device memory belongs to the context;
 *  - pointers passed as "data" may be host or device pointers (the copy uses
 *    unified addressing);
 *  - one context per host thread; calls on one context are not reentrant.
 */
#ifndef CUMF_SGD_AMD_H
#define CUMF_SGD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUMF_ABI_VERSION 3

enum cumf_status {
    CUMF_OK = 0,
    CUMF_ERR_ARG = 1,     /* invalid argument */
    CUMF_ERR_HIP = 2,     /* HIP runtime error */
    CUMF_ERR_STATE = 3,   /* call out of order (e.g. train before ratings) */
    CUMF_ERR_NOMEM = 4,   /* allocation failed */
    CUMF_ERR_IO = 5       /* file I/O failed */
};

/* Factor storage format.  Arithmetic is always FP32. */
enum cumf_precision {
    CUMF_FP32 = 0,
    CUMF_FP16 = 1,        /* the reference's native format (singleGPU/sgd.h:124-133) */
    CUMF_BF16 = 2,
    CUMF_FP8_E4M3 = 3,    /* OCP E4M3 + one power-of-two FP32 scale per row */
    CUMF_BF16_SR = 4      /* BF16 whose update write-back rounds stochastically */
};

/* Which rating blocks run in one launch. */
enum cumf_schedule {
    CUMF_SCHED_AUTO = -1,    /* cumf_sgd_train / cumf_sgd_train_file options only: a 1x1 grid
                                trains cumf_default_schedule's choice, a larger grid GRID */
    CUMF_SCHED_HOGWILD = 0,  /* all workers on the whole matrix (reference 1x1 mode); factors of
                                at most 32 MB (P and Q) train on one XCD, whose L2 then holds
                                the only L2 copy of every row */
    CUMF_SCHED_SERIAL = 1,   /* one worker, deterministic window order */
    CUMF_SCHED_GRID = 2,     /* blocks one after another, all workers each (u x v / x x y) */
    CUMF_SCHED_STRATUM = 3,  /* s x s blocks, a conflict-free diagonal per launch */
    CUMF_SCHED_TILE = 4      /* the stratum diagonals of cumf_partition_tiles' blocks; each block's
                                tiles (a few hundred consecutive rows of P) are trained by one
                                workgroup each, the tile's P rows held in that workgroup's LDS for
                                the launch */
};

/* One rating: mf_node, singleGPU/sgd.h:79-83 (12 bytes). */
typedef struct cumf_node {
    int32_t u;
    int32_t v;
    float rate;
} cumf_node;

/* Training options.  Field meanings follow Parameter (singleGPU/sgd.h:26-46). */
typedef struct cumf_train_opts {
    int32_t k;                 /* factor length: 32, 64 or 128 */
    int32_t num_workers;       /* Hogwild workers (-s); 0 = 128 waves per CU (FP8 or stratum: 512), rounded for least over-sampling */
    int32_t batch_len;         /* records per window, even, <= 128: update_vector_size (sgd_kernel.cu:198); 0 = 128.
                                  The tile schedule takes it as the longest window, at most 32 (0 = 32):
                                  each tile gets the even length up to it whose windows fill its
                                  workgroup's waves best */
    int32_t schedule;          /* enum cumf_schedule */
    int32_t lanes_per_sample;  /* lanes sharing one update: 8..64; 0 = 16 B of a row per lane (tile
                                  schedule: 32 B where k and the format allow it) */
    int32_t coherent;          /* 1: factor loads/stores coherent across XCDs (device scope) */
    int32_t pipeline_depth;    /* row pairs in flight per worker: 2, 4 or 8; 0 = 2 */
    int32_t use_graph;         /* 1: run the epoch loop as a captured HIP graph, replayed per epoch */
    float lambda_p;            /* already divided by the rating scale (sgd.cu:673) */
    float lambda_q;
    float alpha;               /* lr_t = alpha/(1+beta*t^1.5) + lrate (sgd_kernel.cu:144-149) */
    float beta;
    float lrate;
    uint64_t seed;             /* Philox key for windows and rounding */
    int32_t stratum_rounds;    /* stratum and tile schedules: passes over the s diagonals per epoch, each
                                  running 1/rounds of every block's windows; 0 = 1 */
    int32_t tiles_in_flight;   /* tile schedule: tiles one launch trains per XCD (a diagonal's other tiles
                                  in later launches); 0 = all of a diagonal's tiles in one launch.  For
                                  the tile schedule num_workers counts the workers of one tile, at most 16
                                  waves' worth (0 = 32); the tile's workgroup has 4, 8 or 16 waves */
} cumf_train_opts;

typedef struct cumf_ctx cumf_ctx;

/* ---- context ------------------------------------------------------------ */
int cumf_abi_version(void);
const char* cumf_last_error(void);
int cumf_create(int device, cumf_ctx** out);
int cumf_destroy(cumf_ctx* ctx);
int cumf_sync(cumf_ctx* ctx);

/* ---- ratings ------------------------------------------------------------ */
/* Training ratings as 12-byte records (the reference's R array,
 * singleGPU/sgd.cu:239-265) or as three arrays.  m and n are the factor row
 * counts; every id must lie in [0, m) x [0, n). */
int cumf_set_ratings(cumf_ctx* ctx, const cumf_node* R, int64_t nnz, int32_t m, int32_t n);
int cumf_set_ratings_soa(cumf_ctx* ctx, const int32_t* u, const int32_t* v, const float* r,
                         int64_t nnz, int32_t m, int32_t n);
/* read_problem (singleGPU/sgd.cu:190-335) fused with the upload: reads the
 * file, or the shards path0..path79 when any exist, straight into the
 * context's device arrays through a ring of pinned buffers and reader
 * threads (no full host copy).  A file's trailing partial record is dropped;
 * m and n are the largest ids + 1 (at least 1 for a single file).  Negative
 * ids are an error; a missing or unreadable file is CUMF_ERR_IO. */
int cumf_read_ratings(cumf_ctx* ctx, const char* path, int64_t* nnz, int32_t* m, int32_t* n);
/* The same for the test ratings (any ids; those outside the model predict b). */
int cumf_read_test(cumf_ctx* ctx, const char* path, int64_t* nnz);

/* Out-of-core ratings for the stratum schedule (the reference's streamed
 * u x v x x x y grid, singleGPU/sgd_kernel.cu:373-690, rebuilt for HBM-sized
 * blocks): R (host, ids already permuted and values scaled) is split into
 * s x s blocks exactly as cumf_partition_strata splits device ratings (even
 * s, 2..64) and kept in pinned host memory; cumf_train with the stratum
 * schedule then copies each diagonal into one of two device staging buffers
 * on a second stream while the previous diagonal trains.  Device memory
 * holds two diagonals instead of all ratings.  offsets (s*s + 1 entries, may
 * be NULL) receives the block bounds.  Moments, remap, scaling and the other
 * partitions are not available on streamed ratings; any later set/read/alloc
 * of ratings replaces them. */
int cumf_stream_ratings(cumf_ctx* ctx, const cumf_node* R, int64_t nnz, int32_t m, int32_t n, int32_t s,
                        int64_t* offsets);

/* Test ratings for cumf_test_rmse (ids outside [0,m) x [0,n) predict b,
 * test/mf.cpp:3796-3797). */
int cumf_set_test(cumf_ctx* ctx, const cumf_node* R, int64_t nnz);
int cumf_set_test_soa(cumf_ctx* ctx, const int32_t* u, const int32_t* v, const float* r, int64_t nnz);

/* Zero-copy loading of very large rating sets: allocate the context's three
 * rating arrays (nnz entries each) and return their device addresses; the
 * caller fills them on the device (any stream), then commits, which waits for
 * the device and checks every id like cumf_set_ratings_soa. */
int cumf_alloc_ratings(cumf_ctx* ctx, int64_t nnz, int32_t m, int32_t n, int32_t** u, int32_t** v, float** r);
int cumf_commit_ratings(cumf_ctx* ctx);

/* Copies of the ratings as trained (3 arrays of nnz), e.g. for a test. */
int cumf_get_ratings(cumf_ctx* ctx, int32_t* u, int32_t* v, float* r);

/* ---- on-device preprocessing (singleGPU/sgd.cu:78-134) ------------------ */
/* sum r and sum r^2 in double (collect_data, sgd.cu:78-96). */
int cumf_rating_moments(cumf_ctx* ctx, double* sum, double* sum_sq);
/* u <- p_map[u], v <- q_map[v] on training and test ratings
 * (shuffle_problem, sgd.cu:126-134).  Test ids out of range are kept. */
int cumf_remap(cumf_ctx* ctx, const int32_t* p_map, const int32_t* q_map);
/* r <- r * scale on training ratings (scale_problem, sgd.cu:98-107). */
int cumf_scale_ratings(cumf_ctx* ctx, float scale);
/* Partition the training ratings into ux x vy blocks of the reference's
 * shape, block (u/u_seg)*vy + v/v_seg (grid_problem, sgd.cu:411-487), with a
 * stable counting sort.  offsets (ux*vy+1 entries) may be NULL. */
int cumf_partition(cumf_ctx* ctx, int32_t ux, int32_t vy, int64_t* offsets);
/* Partition into s x s blocks of s contiguous row ranges and s contiguous
 * column ranges (for CUMF_SCHED_STRATUM), stable like cumf_partition.  For
 * s <= 8 (one block per XCD and diagonal) each axis is cut so that its
 * largest range holds as few ratings as any contiguous cut allows, each range
 * closed within that count nearest the remaining ratings' even share; for
 * s > 8 a range ends at the first row reaching the next 1/s of the
 * ratings. */
int cumf_partition_strata(cumf_ctx* ctx, int32_t s, int64_t* offsets);
/* Record order inside every block of the current partition: a stable sort
 * of each block's records by column id (CUMF_ORDER_COLUMNS) or row id
 * (CUMF_ORDER_ROWS), ties kept in their order.  No counterpart in the
 * reference (its blocks keep the order grid_problem's cycle-swap leaves,
 * singleGPU/sgd.cu:465-482).  With the records of a column adjacent, a
 * window's consecutive samples mostly share that column's row, which the
 * update kernel then keeps in registers instead of storing and reloading
 * (99M-rating Netflix shape, 24 x 24 strata: FP16 1.52e10 -> 2.03e10
 * updates/s, test RMSE unchanged).  Up to 2^32 - 1 ratings; not for
 * streamed ratings. */
enum cumf_order { CUMF_ORDER_ROWS = 1, CUMF_ORDER_COLUMNS = 2 };
int cumf_order_blocks(cumf_ctx* ctx, int32_t by);
/* Partition for CUMF_SCHED_TILE: the s x s blocks of cumf_partition_strata
 * (s = 0: 8, one row band per XCD and diagonal),
 * each block's records then grouped by tile and, inside a tile, ordered by
 * column (stable).  A tile is a range of consecutive rows of one row band
 * (at most tile_rows rows, cut where the band's ratings would otherwise be
 * unevenly spread over its tiles); tile_rows = 0 takes cumf_tile_rows(k,
 * precision) for the factors the context holds (FP16 at k = 128 before
 * they are initialised).  No counterpart in the reference (its blocks are
 * trained by all workers from memory, singleGPU/sgd_kernel.cu:249-372).  A
 * tile's P rows fit one workgroup's LDS, so the tile trains without sending
 * its P rows through the L2.  Up to 2^32 - 1 ratings; not for streamed
 * ratings.  offsets (s*s + 1 entries, may be NULL) receives the block
 * bounds; cumf_get_tiles returns the tiles. */
int cumf_partition_tiles(cumf_ctx* ctx, int32_t s, int32_t tile_rows, int64_t* offsets);
/* Default rows per tile for factors of length k in a format: 128 * 128 / k
 * for FP32, 96 * 128 / k for the narrower formats, at most 64 KB of rows
 * (measured fastest on a Netflix-shaped problem). */
int32_t cumf_tile_rows(int32_t k, int32_t precision);
/* The tiles of cumf_partition_tiles, block by block (block-major, then tile
 * order within the block): *count receives their number; when the arrays
 * are non-NULL, tile t holds records [rec_off[t], rec_off[t+1]) (count + 1
 * entries) and rows [row_begin[t], row_begin[t] + row_count[t]). */
int cumf_get_tiles(cumf_ctx* ctx, int64_t* count, int64_t* rec_off, int32_t* row_begin, int32_t* row_count);
/* Declare that the uploaded ratings are already grouped into ux x vy blocks
 * with the given prefix offsets (as the host cycle-swap leaves them). */
int cumf_set_blocks(cumf_ctx* ctx, int32_t ux, int32_t vy, const int64_t* offsets);

/* ---- factors ------------------------------------------------------------ */
/* Allocate P (m x k) and Q (n x k) and fill every element with
 * U(0,1] * sqrt(1/k) (random_init, sgd.cu:497-546) in the given format. */
int cumf_init_factors(cumf_ctx* ctx, int32_t k, int32_t precision, uint64_t seed);
/* Raw storage in/out: which = 0 for P, 1 for Q.  data holds rows*k elements
 * of the storage format; scales (FP8 only, else NULL) holds rows floats. */
int cumf_set_factors(cumf_ctx* ctx, int32_t which, const void* data, const float* scales);
int cumf_get_factors(cumf_ctx* ctx, int32_t which, void* data, float* scales);
/* FP32 export: out[i] = widen(F[map ? map[i] : i]) * mult for i < rows
 * (transform_half + scale_model + shuffle_model, sgd_kernel.cu:29-68,
 * sgd.cu:593-635).  out holds rows*k floats. */
int cumf_export_factors(cumf_ctx* ctx, int32_t which, float mult, const int32_t* map, float* out);

/* ---- training ----------------------------------------------------------- */
/* Run epochs [epoch_begin, epoch_end) under opts->schedule.  For
 * CUMF_SCHED_GRID, block_order lists the blocks of each epoch in the order
 * they run (order_len entries, reused every epoch when order_len equals the
 * block count, or epoch-major when it is (epoch_end-epoch_begin) times that);
 * NULL runs blocks in index order. */
int cumf_train(cumf_ctx* ctx, const cumf_train_opts* opts, int32_t epoch_begin, int32_t epoch_end,
               const int32_t* block_order, int64_t order_len);
/* Test RMSE of the current factors: predictions use widen(F) * mult (the
 * exported model), then test/mf.cpp:3794-3839 exactly: sequential FP32 dot,
 * b for out-of-range ids or nan, FP64 sum of FP32 squared errors. */
int cumf_test_rmse(cumf_ctx* ctx, float mult, float b, double* rmse);

/* Evaluation criteria of the reference's mf-predict -e (test/mf.h:30-40,
 * test/mf-predict.cpp:129-182), replacing calc_rmse/calc_mae/calc_gkl/
 * calc_logloss/calc_accuracy (test/mf.cpp:3814-3912) and calc_mpr_auc
 * (test/mf.cpp:3914-4044) on the same predictions as cumf_test_rmse.  The
 * ranking criteria score every column of each row (user, or item for the
 * column-wise ones) that has test positives (r > 0); they need distinct
 * (row, column) positives. */
enum cumf_eval {
    CUMF_EVAL_RMSE = 0, CUMF_EVAL_MAE = 1, CUMF_EVAL_GKL = 2, CUMF_EVAL_LOGLOSS = 5, CUMF_EVAL_ACC = 6,
    CUMF_EVAL_ROW_MPR = 10, CUMF_EVAL_COL_MPR = 11, CUMF_EVAL_ROW_AUC = 12, CUMF_EVAL_COL_AUC = 13
};
int cumf_test_eval(cumf_ctx* ctx, int32_t criterion, float mult, float b, double* value);

/* Load a trained FP32 model (any k up to 4096, as mf_load_model reads it,
 * test/mf.cpp:3706-3777) for evaluation: sets m, n, k and FP32 factors.
 * With ratings present, m and n must match them; training still needs k in
 * {32, 64, 128}. */
int cumf_set_model(cumf_ctx* ctx, int32_t m, int32_t n, int32_t k, const float* P, const float* Q);

/* The schedule a 1x1 grid trains by default (cumf_sgd_train and
 * cumf_sgd_train_file with no options or CUMF_SCHED_AUTO, the command line
 * without --schedule, cumf_sgd_update_k128), chosen from the context's
 * ratings for factors of length k in the given format:
 *  - P and Q together at most 32 MB: CUMF_SCHED_STRATUM with 4 passes over
 *    the diagonals per epoch;
 *  - a row and a column that each hold at least 1% of the ratings:
 *    CUMF_SCHED_HOGWILD (one diagonal at a time would pull both towards one
 *    range of the other axis per sub-epoch), its one block's records ordered
 *    by column for the round-to-nearest formats (FP32, FP16, BF16);
 *  - otherwise CUMF_SCHED_STRATUM, one pass, each block's records ordered
 *    by column (cumf_order_blocks, CUMF_ORDER_COLUMNS), or by row when P
 *    takes more than 1 GB (its rows then come from HBM), when there are at
 *    most 2^32 - 1 ratings;
 *  - except, for the round-to-nearest formats (FP32, FP16, BF16) where that
 *    would be strata ordered by column: CUMF_SCHED_TILE on the partition of
 *    cumf_partition_tiles(ctx, 0, 0, ...) (*order 0), with two passes over
 *    the diagonals per epoch for FP32, one for FP16 and BF16.
 * *order receives CUMF_ORDER_COLUMNS, CUMF_ORDER_ROWS or 0 (keep the
 * records' order).
 * It replaces the reference's fixed choice of Hogwild for a 1x1 grid
 * (singleGPU/sgd_kernel.cu:212-248), made for one coherent L2; MI355X has
 * eight XCDs with an L2 each. */
int cumf_default_schedule(cumf_ctx* ctx, int32_t k, int32_t precision, int32_t* schedule,
                          int32_t* stratum_rounds, int32_t* order);

/* Pick the launch shape: time one epoch of every supported combination of
 * lanes_per_sample {64, 32, 16, 8}, pipeline_depth {2, 4, 8}, batch_len
 * {128, 64, 32} and workers for {32, 64, 128, 256, 512} waves per CU (a field
 * that is non-zero in base stays fixed) for base's schedule on this
 * context's problem, write the one with the shortest epoch into *best (other
 * fields copied from base) and its epoch's kernel time in ms into *best_ms,
 * then restore the factors the trials trained.  Kernel-time totals and
 * last_launch are not changed.  A serial schedule returns base. */
int cumf_autotune(cumf_ctx* ctx, const cumf_train_opts* base, cumf_train_opts* best, double* best_ms);

/* Time spent in training kernels since the last reset (HIP events on the
 * context's stream), the number of training launches, and the number of
 * rating updates they performed. */
int cumf_kernel_time(cumf_ctx* ctx, double* ms, int64_t* launches, int64_t* updates);
int cumf_reset_kernel_time(cumf_ctx* ctx);

/* Resolved launch shape of the last cumf_train call. */
typedef struct cumf_launch_info {
    int32_t num_workers;
    int32_t lanes_per_sample;
    int32_t waves_per_block;
    int32_t blocks;            /* workgroups per launch (stratum: of the largest launch) */
    int32_t batch_len;
    int32_t pipeline_depth;
    int64_t updates_per_epoch;
} cumf_launch_info;
int cumf_last_launch(cumf_ctx* ctx, cumf_launch_info* info);

/* ---- batched problems ----------------------------------------------------
 * Many independent factorizations of uneven size and mixed k trained
 * together, one launch per epoch (no counterpart in the reference, which
 * trains one problem per process).  Each problem's random streams are keyed
 * by (seed, problem index). */
typedef struct cumf_batch cumf_batch;
int cumf_batch_create(int device, int32_t precision, cumf_batch** out);
int cumf_batch_destroy(cumf_batch* batch);
/* Add a problem: training ratings (ids in [0,m) x [0,n)), optional test
 * ratings, k in {32, 64, 128}, lambdas already divided by the problem's
 * rating scale.  Factors are initialised as cumf_init_factors does. */
int cumf_batch_add(cumf_batch* batch, const int32_t* u, const int32_t* v, const float* r, int64_t nnz,
                   int32_t m, int32_t n, int32_t k, float lambda_p, float lambda_q, const int32_t* tu,
                   const int32_t* tv, const float* tr, int64_t test_nnz, uint64_t seed, int32_t* index);
/* Stable sort of every added problem's training records by row id
 * (CUMF_ORDER_ROWS) or column id (CUMF_ORDER_COLUMNS), as cumf_order_blocks
 * does for one block: a worker's consecutive samples then mostly share a
 * row, which the update kernel keeps in registers.  Up to 2^32 - 1 ratings
 * per problem. */
int cumf_batch_order(cumf_batch* batch, int32_t by);
/* Epochs [epoch_begin, epoch_end) of every problem.  opts: schedule
 * (CUMF_SCHED_HOGWILD, or CUMF_SCHED_SERIAL = one worker per problem),
 * num_workers = total waves (0 = 64 per CU, shared in proportion to each
 * problem's bytes of work), alpha, beta, lrate, seed, batch_len,
 * pipeline_depth, use_graph (1: the epoch's launches are captured once as a
 * HIP graph, re-captured when the problems or the shape change, and replayed
 * per epoch); k and the lambdas come from each problem. */
int cumf_batch_train(cumf_batch* batch, const cumf_train_opts* opts, int32_t epoch_begin, int32_t epoch_end);
/* The batch's launch shape: times one epoch for total waves in
 * {4, 8, 16, 32, 64} per CU, pipeline depth {2, 4, 8} and, when
 * base->batch_len is 0, window length {128, 64, 32}, writes the fastest
 * into *best (other fields from base) and its epoch time (ms) into *best_ms,
 * and puts every problem's factors back.  A serial schedule returns base. */
int cumf_batch_autotune(cumf_batch* batch, const cumf_train_opts* base, cumf_train_opts* best, double* best_ms);
int cumf_batch_test_rmse(cumf_batch* batch, int32_t index, float mult, float b, double* rmse);
int cumf_batch_test_eval(cumf_batch* batch, int32_t index, int32_t criterion, float mult, float b, double* value);
int cumf_batch_get_factors(cumf_batch* batch, int32_t index, int32_t which, void* data, float* scales);
int cumf_batch_kernel_time(cumf_batch* batch, double* ms, int64_t* launches, int64_t* updates);

/* ---- reference-seam entry points ---------------------------------------- */
/* Mirror of Parameter (singleGPU/sgd.h:26-46). */
typedef struct cumf_param {
    int32_t k, num_workers, u_grid, v_grid, x_grid, y_grid, ux, vy, num_iters, gpu;
    float lambda_p, lambda_q, lrate, alpha, beta;
} cumf_param;

/* Drop-in for sgd_update_k128 (singleGPU/sgd_kernel.h:17,
 * singleGPU/sgd_kernel.cu:125-720): the caller passes the grid-ordered,
 * scaled, id-localised ratings (R2D[ux*vy] block pointers with gridSize
 * counts), the FP16 factor bits halfp (ux*u_seg*k) and halfq (vy*v_seg*k),
 * updated in place, and receives FP32 copies in floatp/floatq.  Schedule:
 * 1x1 -> conflict-free 24 x 24 strata over the whole matrix (the
 * reference's 1x1 Hogwild would keep per-XCD L2 copies of rows for a whole
 * epoch); otherwise blocks in the reference's order, driven by the C
 * library rand() exactly as the reference does (sgd_kernel.cu:253-267,
 * 399-451).  seed is the Philox key. */
int cumf_sgd_update_k128(const cumf_param* para, int32_t m, int32_t n, int64_t u_seg, int64_t v_seg,
                         const cumf_node* const* R2D, const int64_t* grid_size,
                         uint16_t* halfp, uint16_t* halfq, float* floatp, float* floatq,
                         uint64_t seed);

/* Drop-in for the sgd_train pipeline (singleGPU/sgd.cu:638-690) plus
 * save-ready output: statistics, id permutation (srand(seed) +
 * std::random_shuffle, sgd.cu:39-60), grid, scaling, init, training, rescale
 * and inverse permutation.  R is read only.  opts (may be NULL) overrides
 * the schedule (the stratum schedule takes s from a square u x v grid, else
 * 24), lanes_per_sample, coherent, batch_len and stratum_rounds; its k, lambdas,
 * alpha, beta, lrate and seed are taken from para and seed.  With a 1x1 grid
 * everything runs on the device, under the stratum schedule unless opts
 * says otherwise; with more blocks the id shuffle and the cycle-swap
 * partition run on the host first so that block contents and order are the
 * reference's.  P_out (m*k) and Q_out (n*k) receive the final model in original id order,
 * *b_out the model's b (ave/std_dev, sgd.cu:679).  precision selects the
 * factor format (CUMF_FP16 is the reference's).  test (may be NULL, test_nnz
 * records, original ids) is scored after every epoch into rmse_out[epoch]
 * (may be NULL). */
int cumf_sgd_train(const cumf_param* para, const cumf_train_opts* opts, const cumf_node* R,
                   int64_t nnz, int32_t m, int32_t n, int32_t precision, uint64_t seed,
                   const cumf_node* test, int64_t test_nnz, float* P_out, float* Q_out,
                   float* b_out, double* rmse_out);

/* sgd_train for a 1x1 grid straight from the files: read_problem's inputs
 * streamed into HBM (cumf_read_ratings / cumf_read_test), collect_data on
 * device FP64 sums, the same maps, scaling, training and export as
 * cumf_sgd_train.  *P_out and *Q_out (m x k, n x k, original ids, times
 * sqrt(scale)) are malloc'd; free with cumf_free_host.  rmse_out, when not
 * NULL and test_path is given, receives num_iters test RMSEs. */
int cumf_sgd_train_file(const cumf_param* para, const cumf_train_opts* opts, const char* train_path,
                        const char* test_path, int32_t precision, uint64_t seed, float** P_out, float** Q_out,
                        int32_t* m, int32_t* n, int64_t* nnz, float* b_out, double* rmse_out);

/* ---- allocation-free training -------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): epochs [epoch_begin, epoch_end) of opts->schedule
 * (CUMF_SCHED_HOGWILD or CUMF_SCHED_SERIAL; use_graph 0) over ratings u, v, r
 * (nnz, ids in [0,m) x [0,n), not checked) and factors P (m x opts->k), Q
 * (n x opts->k) in the given format (FP8: with row scales ps, qs), all in
 * the caller's device memory on `device`, updated in place.  Same kernels,
 * shapes and results as cumf_train on a context holding the same ratings and
 * factors.  The launch's task record and learning-rate table live in a fixed
 * 16 KiB device array of the library (one per GPU, kept until the process
 * ends), hence epoch_end <= CUMF_NOALLOC_MAX_EPOCHS; the call allocates
 * nothing, queues its work on `stream` (a hipStream_t; NULL = the null
 * stream) and returns.  Calls that overlap on one GPU share that array and
 * race. */
#define CUMF_NOALLOC_MAX_EPOCHS 4028
int cumf_train_noalloc(int device, void* stream, const cumf_train_opts* opts, int32_t precision,
                       const int32_t* u, const int32_t* v, const float* r, int64_t nnz, int32_t m, int32_t n,
                       void* P, void* Q, float* ps, float* qs, int32_t epoch_begin, int32_t epoch_end);

/* ---- host helpers (the reference's host pipeline, bit-exact) ------------- */
/* gen_random_map (singleGPU/sgd.cu:39-53) with srand(seed) in place of
 * srand(time(NULL)). */
int cumf_gen_random_map(uint32_t seed, int32_t size, int32_t* map);
/* collect_data (singleGPU/sgd.cu:78-96): serial double sums. */
int cumf_collect_data(const cumf_node* R, int64_t nnz, float* ave, float* std_dev);
/* grid_problem (singleGPU/sgd.cu:411-487): in-place cycle-swap partition. */
int cumf_grid_problem(cumf_node* R, int64_t nnz, int32_t m, int32_t n, int32_t ux, int32_t vy,
                      int64_t* grid_size, int64_t* offsets, int64_t* u_seg, int64_t* v_seg);
/* read_problem (singleGPU/sgd.cu:190-335): one file, or shards path0..path79
 * read in parallel.  *R_out is malloc'd (free with cumf_free_host). */
int cumf_read_problem(const char* path, cumf_node** R_out, int64_t* nnz, int32_t* m, int32_t* n);
void cumf_free_host(void* p);
/* mf_load_model (test/mf.cpp:3706-3777): the 20-byte header {fun, m, n, k, b}
 * and the P and Q rows; *P and *Q are malloc'd (free with cumf_free_host).
 * A missing or short file is CUMF_ERR_IO (the reference exits). */
int cumf_load_model(const char* path, int32_t* m, int32_t* n, int32_t* k, float* b, float** P, float** Q);
/* save_model (singleGPU/main.cpp:231-273): 20-byte header + P rows + Q rows. */
int cumf_save_model(const char* path, int32_t m, int32_t n, int32_t k, float b,
                    const float* P, const float* Q);

#ifdef __cplusplus
}
#endif

#endif /* CUMF_SGD_AMD_H */
