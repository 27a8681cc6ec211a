// C ABI for batched training: many independent matrix factorizations of
// uneven size and mixed factor length (32, 64, 128) trained together, one
// launch per epoch (sgd_batched_kernel).  No reference counterpart: the
// reference trains one problem per process.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstring>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/cumf_sgd_amd.h"
#include "errors.h"
#include "kernels.h"
#include "train_plan.h"

using namespace cumf;

namespace {

#define BHIP_TRY(expr)                                                                               \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return fail(CUMF_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                       \
    } while (0)

struct Problem {
    int64_t nnz = 0, tnnz = 0;
    int32_t m = 0, n = 0, k = 0;
    float lp = 0, lq = 0;
    int32_t* u = nullptr;
    int32_t* v = nullptr;
    float* r = nullptr;
    int32_t* tu = nullptr;
    int32_t* tv = nullptr;
    float* tr = nullptr;
    void* P = nullptr;
    void* Q = nullptr;
    float* ps = nullptr;
    float* qs = nullptr;
};

int eb_of(int prec) { return prec == CUMF_FP32 ? 4 : (prec == CUMF_FP8_E4M3 ? 1 : 2); }

int dev_alloc(void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) return CUMF_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(CUMF_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    return CUMF_OK;
}

void dev_free(void* p)
{
    if (p) (void)hipFree(p);
}

}  // namespace

struct cumf_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;
    int prec = CUMF_BF16;
    std::vector<Problem> probs;
    float* lr = nullptr;
    int lr_n = 0;
    // rates the table holds: its first lr_valid entries, for (alpha, beta, lrate) = lr_key
    int lr_valid = 0;
    float lr_key[3] = {0, 0, 0};
    SgdTask* tasks = nullptr;
    size_t tasks_n = 0;
    double* partials = nullptr;
    size_t partials_n = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double kernel_ms = 0;
    int64_t launches = 0, updates = 0;
    // use_graph: the epoch (update launch, FP8 rescales, counter advance)
    // captured once per plan and replayed; kernels read the epoch here
    int32_t* epoch_dev = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_key = 0;
};

namespace {

int check_batch(cumf_batch* b)
{
    if (!b) return fail(CUMF_ERR_ARG, "null batch");
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return fail(CUMF_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    return CUMF_OK;
}

int problem_of(cumf_batch* b, int32_t index, Problem** p)
{
    if (index < 0 || index >= (int32_t)b->probs.size()) return fail(CUMF_ERR_ARG, "problem %d out of range", index);
    *p = &b->probs[(size_t)index];
    return CUMF_OK;
}

void free_problem(Problem& p)
{
    dev_free(p.u); dev_free(p.v); dev_free(p.r);
    dev_free(p.tu); dev_free(p.tv); dev_free(p.tr);
    dev_free(p.P); dev_free(p.Q); dev_free(p.ps); dev_free(p.qs);
    p = Problem{};
}

constexpr int kBatchXcds = 8;   // MI355X: 8 XCDs of 32 CUs

// XCD-aware placement of the problems' waves.  The kernel numbers waves
// XCD-major, so XCD x runs waves [x*per_xcd, (x+1)*per_xcd).  Each XCD has its
// own L2, which device memory does not keep coherent with the others inside
// a kernel: a problem whose workers sit on two XCDs trains two copies of its
// hot rows, and one of them is lost at write-back.  So every problem that is
// at most an XCD's fair share of waves goes whole onto the least-loaded XCD
// (longest first; the per-XCD size grows from the mean until everything
// fits, and never past the mean plus the largest such problem, the longest-
// first bound); a larger problem is laid across consecutive XCDs
// first (its rows are spread over many more ratings, and it has to use more
// than one XCD to finish in time).  On return the tasks are sorted by
// worker_begin and *per_xcd is the waves per XCD, a multiple of a block.
void place_on_xcds(std::vector<SgdTask>& tasks, int nx, int64_t* per_xcd)
{
    int64_t total = 0, biggest_small = 0;
    for (const SgdTask& t : tasks) total += t.wave_count;
    const int64_t fair = (total + nx - 1) / nx;
    for (const SgdTask& t : tasks)
        if (t.wave_count <= fair) biggest_small = std::max<int64_t>(biggest_small, t.wave_count);
    std::vector<size_t> order(tasks.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return tasks[a].wave_count > tasks[b].wave_count; });
    int64_t cap = fair;
    for (;;) {
        cap = (cap + kSgdWavesPerBlock - 1) / kSgdWavesPerBlock * kSgdWavesPerBlock;
        std::vector<int64_t> load(nx, 0);
        int64_t cursor = 0;   // end of the problems laid across XCDs
        bool ok = true;
        for (size_t i : order) {
            SgdTask& t = tasks[i];
            if (t.wave_count > fair) {
                t.worker_begin = (int32_t)cursor;
                cursor += t.wave_count;
                continue;
            }
            for (int x = 0; x < nx; ++x) load[x] = std::max(load[x], std::min(cap, std::max<int64_t>(0, cursor - x * cap)));
            int best = -1;
            for (int x = 0; x < nx; ++x)
                if (load[x] + t.wave_count <= cap && (best < 0 || load[x] < load[best])) best = x;
            if (best < 0) {
                ok = false;
                break;
            }
            t.worker_begin = (int32_t)(best * cap + load[best]);
            load[best] += t.wave_count;
        }
        if (ok && cursor <= (int64_t)nx * cap) break;
        cap += std::max<int64_t>(cap / 32, kSgdWavesPerBlock);
    }
    std::stable_sort(tasks.begin(), tasks.end(),
                     [](const SgdTask& a, const SgdTask& b) { return a.worker_begin < b.worker_begin; });
    *per_xcd = cap;
}

}  // namespace

extern "C" {

int cumf_batch_create(int device, int32_t precision, cumf_batch** out)
{
    if (!out) return fail(CUMF_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (precision < CUMF_FP32 || precision > CUMF_BF16_SR) return fail(CUMF_ERR_ARG, "bad precision %d", precision);
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(CUMF_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    device = (device < 0 ? 0 : device) % count;
    BHIP_TRY(hipSetDevice(device));
    cumf_batch* b = new cumf_batch;
    b->device = device;
    b->prec = precision;
    if (hipStreamCreate(&b->stream) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
        hipEventCreate(&b->ev1) != hipSuccess) {
        cumf_batch_destroy(b);
        return fail(CUMF_ERR_HIP, "stream/event creation failed");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        b->cu_count = prop.multiProcessorCount;
    *out = b;
    return CUMF_OK;
}

int cumf_batch_destroy(cumf_batch* b)
{
    if (!b) return CUMF_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (Problem& p : b->probs) free_problem(p);
    dev_free(b->lr);
    dev_free(b->tasks);
    dev_free(b->partials);
    dev_free(b->epoch_dev);
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CUMF_OK;
}

int cumf_batch_add(cumf_batch* b, const int32_t* u, const int32_t* v, const float* r, int64_t nnz, int32_t m,
                   int32_t n, int32_t k, float lambda_p, float lambda_q, const int32_t* tu, const int32_t* tv,
                   const float* tr, int64_t test_nnz, uint64_t seed, int32_t* index)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (nnz < 0 || m <= 0 || n <= 0 || (nnz > 0 && (!u || !v || !r)) || test_nnz < 0 ||
        (test_nnz > 0 && (!tu || !tv || !tr)))
        return fail(CUMF_ERR_ARG, "bad problem arguments");
    if (k != 32 && k != 64 && k != 128) return fail(CUMF_ERR_ARG, "k=%d unsupported (32, 64 or 128)", k);
    if (b->probs.size() >= (1u << 24)) return fail(CUMF_ERR_ARG, "too many problems");
    const uint32_t idx = (uint32_t)b->probs.size();
    Problem p;
    p.nnz = nnz; p.tnnz = test_nnz; p.m = m; p.n = n; p.k = k; p.lp = lambda_p; p.lq = lambda_q;
    const size_t eb = (size_t)eb_of(b->prec);
    bool ok = !dev_alloc((void**)&p.u, 4 * (size_t)nnz) && !dev_alloc((void**)&p.v, 4 * (size_t)nnz) &&
              !dev_alloc((void**)&p.r, 4 * (size_t)nnz) && !dev_alloc((void**)&p.tu, 4 * (size_t)test_nnz) &&
              !dev_alloc((void**)&p.tv, 4 * (size_t)test_nnz) && !dev_alloc((void**)&p.tr, 4 * (size_t)test_nnz) &&
              !dev_alloc(&p.P, (size_t)m * k * eb) && !dev_alloc(&p.Q, (size_t)n * k * eb);
    if (ok && b->prec == CUMF_FP8_E4M3)
        ok = !dev_alloc((void**)&p.ps, 4 * (size_t)m) && !dev_alloc((void**)&p.qs, 4 * (size_t)n);
    if (!ok) {
        free_problem(p);
        return CUMF_ERR_NOMEM;
    }
    auto cleanup = [&](int code) {
        free_problem(p);
        return code;
    };
    hipError_t e = hipSuccess;
    if (nnz) {
        if (e == hipSuccess) e = hipMemcpyAsync(p.u, u, 4 * (size_t)nnz, hipMemcpyDefault, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(p.v, v, 4 * (size_t)nnz, hipMemcpyDefault, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(p.r, r, 4 * (size_t)nnz, hipMemcpyDefault, b->stream);
    }
    if (test_nnz) {
        if (e == hipSuccess) e = hipMemcpyAsync(p.tu, tu, 4 * (size_t)test_nnz, hipMemcpyDefault, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(p.tv, tv, 4 * (size_t)test_nnz, hipMemcpyDefault, b->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(p.tr, tr, 4 * (size_t)test_nnz, hipMemcpyDefault, b->stream);
    }
    if (e != hipSuccess) return cleanup(fail(CUMF_ERR_HIP, "upload: %s", hipGetErrorString(e)));
    // every training id must be a row of its factor matrices
    unsigned long long* bad = nullptr;
    if (dev_alloc((void**)&bad, sizeof(unsigned long long))) return cleanup(CUMF_ERR_NOMEM);
    unsigned long long hb = 0;
    e = hipMemcpyAsync(bad, &hb, sizeof(hb), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = launch_check_ids(p.u, p.v, nnz, m, n, bad, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&hb, bad, sizeof(hb), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    dev_free(bad);
    if (e != hipSuccess) return cleanup(fail(CUMF_ERR_HIP, "id check: %s", hipGetErrorString(e)));
    if (hb) return cleanup(fail(CUMF_ERR_ARG, "%llu ratings have ids outside [0,%d) x [0,%d)", hb, m, n));
    const float s = (float)std::sqrt(1.0 / k);
    e = launch_init(p.P, p.ps, m, k, b->prec, s, seed, idx, kStreamInitP, b->stream);
    if (e == hipSuccess) e = launch_init(p.Q, p.qs, n, k, b->prec, s, seed, idx, kStreamInitQ, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return cleanup(fail(CUMF_ERR_HIP, "init: %s", hipGetErrorString(e)));
    b->probs.push_back(p);
    if (index) *index = (int32_t)idx;
    return CUMF_OK;
}

int cumf_batch_order(cumf_batch* b, int32_t by)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (by != CUMF_ORDER_ROWS && by != CUMF_ORDER_COLUMNS) return fail(CUMF_ERR_ARG, "bad order %d", by);
    int64_t most = 0;
    for (const Problem& p : b->probs) most = std::max(most, p.nnz);
    if (most > (int64_t)UINT32_MAX)
        return fail(CUMF_ERR_ARG, "%lld ratings in one problem: ordering supports up to 2^32 - 1", (long long)most);
    if (most < 2) return CUMF_OK;
    // one block per problem, [0, nnz): keys are the ids themselves (< 2^31)
    struct Scratch {
        void* p[5] = {};
        ~Scratch()
        {
            for (void* q : p) dev_free(q);
        }
    } s;
    void*& off = s.p[0];
    void*& keys = s.p[1];
    void*& keys_out = s.p[2];
    void*& idx = s.p[3];
    void*& idx_out = s.p[4];
    struct Temp {
        void* p = nullptr;
        size_t bytes = 0;
        ~Temp() { dev_free(p); }
    } temp;
    if ((rc = dev_alloc(&off, 2 * sizeof(int64_t))) || (rc = dev_alloc(&keys, 4 * (size_t)most)) ||
        (rc = dev_alloc(&keys_out, 4 * (size_t)most)) || (rc = dev_alloc(&idx, 4 * (size_t)most)) ||
        (rc = dev_alloc(&idx_out, 4 * (size_t)most)))
        return rc;
    for (Problem& p : b->probs) {
        if (p.nnz < 2) continue;
        const int64_t hoff[2] = {0, p.nnz};
        const uint64_t stride = (uint64_t)(by == CUMF_ORDER_COLUMNS ? p.n : p.m);
        int32_t *uo = nullptr, *vo = nullptr;
        float* ro = nullptr;
        if (dev_alloc((void**)&uo, 4 * (size_t)p.nnz) || dev_alloc((void**)&vo, 4 * (size_t)p.nnz) ||
            dev_alloc((void**)&ro, 4 * (size_t)p.nnz)) {
            dev_free(uo); dev_free(vo); dev_free(ro);
            return CUMF_ERR_NOMEM;
        }
        hipError_t e = hipMemcpyAsync(off, hoff, sizeof(hoff), hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess)
            e = launch_order_blocks(p.nnz, by == CUMF_ORDER_COLUMNS ? p.v : p.u, (const int64_t*)off, 1, stride, keys,
                                    keys_out, (uint32_t*)idx, (uint32_t*)idx_out, p.u, p.v, p.r, uo, vo, ro, &temp.p,
                                    &temp.bytes, b->stream);
        // hoff and the old arrays stay live until the stream has drained
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) {
            dev_free(uo); dev_free(vo); dev_free(ro);
            return fail(CUMF_ERR_HIP, "order: %s", hipGetErrorString(e));
        }
        std::swap(p.u, uo);
        std::swap(p.v, vo);
        std::swap(p.r, ro);
        dev_free(uo); dev_free(vo); dev_free(ro);
    }
    return CUMF_OK;
}

int cumf_batch_train(cumf_batch* b, const cumf_train_opts* o, int32_t e0, int32_t e1)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!o) return fail(CUMF_ERR_ARG, "null options");
    if (e0 < 0 || e1 < e0 || e1 > 1 << 20) return fail(CUMF_ERR_ARG, "bad epoch range");
    if (o->schedule != CUMF_SCHED_HOGWILD && o->schedule != CUMF_SCHED_SERIAL)
        return fail(CUMF_ERR_ARG, "batched training supports the hogwild and serial schedules");
    if (o->coherent) return fail(CUMF_ERR_ARG, "batched training has no coherent mode");
    if (o->use_graph != 0 && o->use_graph != 1) return fail(CUMF_ERR_ARG, "use_graph must be 0 or 1");
    const int L = o->batch_len ? o->batch_len : 128;
    if (L < 2 || L > kSgdMaxBatch || L % 2) return fail(CUMF_ERR_ARG, "batch_len must be even and <= %d", kSgdMaxBatch);
    int D = o->pipeline_depth ? o->pipeline_depth : 4;
    if (D != 2 && D != 4 && D != 8) return fail(CUMF_ERR_ARG, "pipeline_depth must be 2, 4 or 8");
    while (L % D) D >>= 1;
    if (b->probs.empty() || e1 == e0) return CUMF_OK;

    // learning-rate table, singleGPU/sgd_kernel.cu:137-154, grown in steps of
    // kLrChunk epochs and written only when the call needs more epochs or
    // other rates (its address is in the tasks and in the graph key)
    const float lr_key[3] = {o->alpha, o->beta, o->lrate};
    if (!b->lr || b->lr_valid < e1 + 4 || std::memcmp(lr_key, b->lr_key, sizeof(lr_key)) != 0) {
        std::vector<float> lr((size_t)(e1 + 4 + kLrChunk - 1) / kLrChunk * kLrChunk);
        lr_table(o, (int)lr.size(), lr.data());
        b->lr_valid = 0;
        if ((int)lr.size() > b->lr_n) {
            dev_free(b->lr);
            b->lr = nullptr;
            b->lr_n = 0;
            if ((rc = dev_alloc((void**)&b->lr, sizeof(float) * lr.size()))) return rc;
            b->lr_n = (int)lr.size();
        }
        BHIP_TRY(hipMemcpyAsync(b->lr, lr.data(), sizeof(float) * lr.size(), hipMemcpyHostToDevice, b->stream));
        b->lr_valid = (int)lr.size();
        std::memcpy(b->lr_key, lr_key, sizeof(lr_key));
    }

    // waves per problem in proportion to its bytes of work; no more workers
    // than windows
    const bool serial = o->schedule == CUMF_SCHED_SERIAL;
    // default 64 waves per CU: the batch autotuner's pick on the 256-problem configuration
    const int64_t total_waves = o->num_workers > 0 ? o->num_workers : (int64_t)b->cu_count * 64;
    double cost_sum = 0;
    for (const Problem& p : b->probs) cost_sum += (double)p.nnz * (12.0 + 4.0 * p.k * eb_of(b->prec));
    std::vector<SgdTask> tasks;
    int64_t wave = 0, upd = 0;
    for (size_t i = 0; i < b->probs.size(); ++i) {
        const Problem& p = b->probs[i];
        if (p.nnz == 0) continue;
        // workers from the wave budget at k/2 lanes per worker (k/4 for
        // FP8), packed 16 bytes of a row per lane into the waves that run
        const int G = sgd_batch_group(p.k, b->prec), S = 64 / G;
        const int Sb = 64 / sgd_batch_budget_group(p.k, b->prec);
        const double cost = (double)p.nnz * (12.0 + 4.0 * p.k * eb_of(b->prec));
        int64_t waves = serial ? 1 : std::max<int64_t>(1, (int64_t)std::llround(total_waves * cost / cost_sum));
        // at least 16 windows per worker and epoch: a small problem runs on
        // few workers (near-serial, like its serial solve) instead of many
        // concurrent ones on a handful of rows
        int64_t workers = serial ? 1 : std::max<int64_t>(1, std::min<int64_t>(waves * Sb, p.nnz / (16LL * L)));
        waves = (workers + S - 1) / S;
        SgdTask t{};
        t.u = p.u; t.v = p.v; t.r = p.r; t.nnz = p.nnz;
        t.P = p.P; t.Q = p.Q; t.ps = p.ps; t.qs = p.qs; t.lr = b->lr;
        t.lambda_p = p.lp; t.lambda_q = p.lq;
        t.worker_begin = 0;   // placed below
        t.wave_count = (int32_t)waves;
        t.worker_count = (int32_t)workers;
        t.update_count = (int32_t)((p.nnz + workers * L - 1) / (workers * L));
        t.batch_len = L;
        t.problem = (uint32_t)i;
        t.tag = 0;
        t.k = p.k;
        tasks.push_back(t);
        wave += waves;
        upd += workers * t.update_count * L;
    }
    if (tasks.empty()) return CUMF_OK;
    if (wave > (int64_t)1 << 30) return fail(CUMF_ERR_ARG, "too many waves");
    int64_t per_xcd = 0;
    place_on_xcds(tasks, kBatchXcds, &per_xcd);
    wave = per_xcd * kBatchXcds;
    if (tasks.size() > b->tasks_n) {
        dev_free(b->tasks);
        b->tasks = nullptr;
        b->tasks_n = 0;
        if ((rc = dev_alloc((void**)&b->tasks, sizeof(SgdTask) * tasks.size()))) return rc;
        b->tasks_n = tasks.size();
    }
    BHIP_TRY(hipMemcpyAsync(b->tasks, tasks.data(), sizeof(SgdTask) * tasks.size(), hipMemcpyHostToDevice, b->stream));
    const int blocks = (int)((wave + kSgdWavesPerBlock - 1) / kSgdWavesPerBlock);
    // every problem's factors under 4 GB: rows addressed as the (wave-uniform)
    // task base plus a 32-bit offset
    int32_t off32 = 1;
    for (const Problem& p : b->probs)
        if ((uint64_t)p.m * p.k * eb_of(b->prec) > kBufRecords || (uint64_t)p.n * p.k * eb_of(b->prec) > kBufRecords)
            off32 = 0;
    // one epoch: the update launch, the FP8 rescales, and with epoch_dev the
    // device epoch counter's advance (the form captured into a graph)
    auto queue_epoch = [&](int ep, int32_t* epoch_dev) -> hipError_t {
        SgdLaunch a{};
        a.tasks = b->tasks;
        a.num_tasks = (int32_t)tasks.size();
        a.epoch_begin = ep;
        a.epoch_end = ep + 1;
        a.batch_len = L;
        a.seed = o->seed;
        a.num_xcd = kBatchXcds;
        a.task_off32 = off32;
        a.epoch_dev = epoch_dev;
        hipError_t e = launch_sgd_batched(a, blocks, b->prec, D, b->stream);
        if (e != hipSuccess) return e;
        if (b->prec == CUMF_FP8_E4M3)
            for (const Problem& p : b->probs) {
                if ((e = launch_fp8_rescale((uint8_t*)p.P, p.ps, p.m, p.k, b->stream)) != hipSuccess) return e;
                if ((e = launch_fp8_rescale((uint8_t*)p.Q, p.qs, p.n, p.k, b->stream)) != hipSuccess) return e;
            }
        return epoch_dev ? launch_advance_int(epoch_dev, b->stream) : hipSuccess;
    };
    if (o->use_graph) {
        // The key covers what the captured launches hold by value: the task
        // table's contents and address, the counter's and the learning-rate
        // table's addresses, the launch shape, and every problem's factors
        // and row counts (the FP8 rescales walk them, empty problems too).
        if (!b->epoch_dev && (rc = dev_alloc((void**)&b->epoch_dev, sizeof(int32_t)))) return rc;
        uint64_t key = fnv1a(tasks.data(), sizeof(SgdTask) * tasks.size(), 1469598103934665603ull);
        const int64_t shape[7] = {blocks, L, D, b->prec, off32, (int64_t)o->seed, (int64_t)b->probs.size()};
        key = fnv1a(shape, sizeof(shape), key);
        const void* addr[3] = {b->tasks, b->epoch_dev, b->lr};
        key = fnv1a(addr, sizeof(addr), key);
        for (const Problem& p : b->probs) {
            const void* f[4] = {p.P, p.Q, p.ps, p.qs};
            const int32_t d[3] = {p.m, p.n, p.k};
            key = fnv1a(d, sizeof(d), fnv1a(f, sizeof(f), key));
        }
        if (!b->graph_exec || b->graph_key != key) {
            if (b->graph_exec) {
                (void)hipGraphExecDestroy(b->graph_exec);
                b->graph_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            BHIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            const hipError_t qe = queue_epoch(e0, b->epoch_dev);
            const hipError_t ce = hipStreamEndCapture(b->stream, &g);
            if (qe != hipSuccess || ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(CUMF_ERR_HIP, "graph capture failed: %s",
                            hipGetErrorString(qe != hipSuccess ? qe : ce));
            }
            const hipError_t ie = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                b->graph_exec = nullptr;
                return fail(CUMF_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
            }
            b->graph_key = key;
        }
        BHIP_TRY(launch_set_int(b->epoch_dev, e0, b->stream));
    }
    BHIP_TRY(hipEventRecord(b->ev0, b->stream));
    for (int ep = e0; ep < e1; ++ep) {
        if (o->use_graph) BHIP_TRY(hipGraphLaunch(b->graph_exec, b->stream));
        else BHIP_TRY(queue_epoch(ep, nullptr));
        b->launches += 1;
        b->updates += upd;
    }
    BHIP_TRY(hipEventRecord(b->ev1, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    float ms = 0;
    BHIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->kernel_ms += ms;
    return CUMF_OK;
}

int cumf_batch_autotune(cumf_batch* b, const cumf_train_opts* base, cumf_train_opts* best, double* best_ms)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (!base || !best) return fail(CUMF_ERR_ARG, "null options");
    if (base->schedule == CUMF_SCHED_SERIAL || b->probs.empty()) {
        *best = *base;
        if (best_ms) *best_ms = 0;
        return CUMF_OK;
    }
    // keep every problem's factors (and FP8 scales) to put them back
    const size_t eb = (size_t)eb_of(b->prec);
    std::vector<void*> keep;
    struct Free {
        std::vector<void*>* v;
        ~Free() { for (void* p : *v) dev_free(p); }
    } guard{&keep};
    auto save = [&](const void* src, size_t bytes) -> int {
        void* d = nullptr;
        int r = dev_alloc(&d, bytes);
        if (r) return r;
        keep.push_back(d);
        if (bytes) BHIP_TRY(hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToDevice, b->stream));
        return CUMF_OK;
    };
    for (const Problem& p : b->probs) {
        if ((rc = save(p.P, (size_t)p.m * p.k * eb)) || (rc = save(p.Q, (size_t)p.n * p.k * eb))) return rc;
        if (b->prec == CUMF_FP8_E4M3 && ((rc = save(p.ps, 4 * (size_t)p.m)) || (rc = save(p.qs, 4 * (size_t)p.n))))
            return rc;
    }
    const double keep_ms = b->kernel_ms;
    const int64_t keep_l = b->launches, keep_u = b->updates;
    const int waves_per_cu[5] = {4, 8, 16, 32, 64};
    const int depths[3] = {2, 4, 8};
    // window lengths only when the caller leaves batch_len to the library
    std::vector<int> lens{base->batch_len};
    if (base->batch_len == 0) lens = {128, 64, 32};
    double best_t = 1e300;
    cumf_train_opts bo = *base;
    bool warm = false;
    for (int len : lens)
        for (int wpc : waves_per_cu)
            for (int D : depths) {
                cumf_train_opts o = *base;
                o.batch_len = len;
                o.num_workers = b->cu_count * wpc;
                o.pipeline_depth = D;
                if (!warm) {
                    if ((rc = cumf_batch_train(b, &o, 0, 1))) return rc;
                    warm = true;
                }
                const double before = b->kernel_ms;
                if ((rc = cumf_batch_train(b, &o, 0, 1))) return rc;
                const double t = b->kernel_ms - before;
                if (t < best_t) {
                    best_t = t;
                    bo = o;
                }
            }
    size_t j = 0;
    for (const Problem& p : b->probs) {
        BHIP_TRY(hipMemcpyAsync(p.P, keep[j++], (size_t)p.m * p.k * eb, hipMemcpyDeviceToDevice, b->stream));
        BHIP_TRY(hipMemcpyAsync(p.Q, keep[j++], (size_t)p.n * p.k * eb, hipMemcpyDeviceToDevice, b->stream));
        if (b->prec == CUMF_FP8_E4M3) {
            BHIP_TRY(hipMemcpyAsync(p.ps, keep[j++], 4 * (size_t)p.m, hipMemcpyDeviceToDevice, b->stream));
            BHIP_TRY(hipMemcpyAsync(p.qs, keep[j++], 4 * (size_t)p.n, hipMemcpyDeviceToDevice, b->stream));
        }
    }
    BHIP_TRY(hipStreamSynchronize(b->stream));
    b->kernel_ms = keep_ms;
    b->launches = keep_l;
    b->updates = keep_u;
    *best = bo;
    if (best_ms) *best_ms = best_t;
    return CUMF_OK;
}

int cumf_batch_test_eval(cumf_batch* b, int32_t index, int32_t criterion, float mult, float bias, double* value)
{
    int rc = check_batch(b);
    if (rc) return rc;
    Problem* p;
    if ((rc = problem_of(b, index, &p))) return rc;
    if (!eval_supported(criterion))
        return fail(CUMF_ERR_ARG, "criterion %d not supported (0, 1, 2, 5, 6, 10-13)", criterion);
    if (!value) return fail(CUMF_ERR_ARG, "null output");
    if (criterion >= 10) {   // row/column MPR (10, 11) and AUC (12, 13)
        double mpr = 0, auc = 0;
        if ((rc = rank_eval(b->stream, p->tu, p->tv, p->tr, p->tnnz, p->P, p->ps, p->Q, p->qs, p->m, p->n, p->k,
                            b->prec, mult, bias, criterion == 11 || criterion == 13, &mpr, &auc)))
            return rc;
        *value = criterion <= 11 ? mpr : auc;
        return CUMF_OK;
    }
    if (p->tnnz == 0) {
        *value = 0.0;
        return CUMF_OK;
    }
    const size_t need = (size_t)rmse_blocks(p->tnnz) + 1;
    if (need > b->partials_n) {
        dev_free(b->partials);
        b->partials = nullptr;
        b->partials_n = 0;
        if ((rc = dev_alloc((void**)&b->partials, sizeof(double) * need))) return rc;
        b->partials_n = need;
    }
    double* out = b->partials + need - 1;
    BHIP_TRY(launch_eval(criterion, p->tu, p->tv, p->tr, p->tnnz, p->P, p->ps, p->Q, p->qs, p->m, p->n, p->k, b->prec,
                         mult, bias, b->partials, out, b->stream));
    double h = 0;
    BHIP_TRY(hipMemcpyAsync(&h, out, sizeof(h), hipMemcpyDeviceToHost, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    *value = criterion == 0 ? std::sqrt(h / (double)p->tnnz) : h / (double)p->tnnz;
    return CUMF_OK;
}

int cumf_batch_test_rmse(cumf_batch* b, int32_t index, float mult, float bias, double* rmse)
{
    return cumf_batch_test_eval(b, index, 0, mult, bias, rmse);
}

int cumf_batch_get_factors(cumf_batch* b, int32_t index, int32_t which, void* data, float* scales)
{
    int rc = check_batch(b);
    if (rc) return rc;
    Problem* p;
    if ((rc = problem_of(b, index, &p))) return rc;
    if (which != 0 && which != 1) return fail(CUMF_ERR_ARG, "which must be 0 (P) or 1 (Q)");
    const int64_t rows = which ? p->n : p->m;
    if (data)
        BHIP_TRY(hipMemcpyAsync(data, which ? p->Q : p->P, (size_t)rows * p->k * eb_of(b->prec), hipMemcpyDefault,
                                b->stream));
    if (scales && b->prec == CUMF_FP8_E4M3)
        BHIP_TRY(hipMemcpyAsync(scales, which ? p->qs : p->ps, sizeof(float) * rows, hipMemcpyDefault, b->stream));
    BHIP_TRY(hipStreamSynchronize(b->stream));
    return CUMF_OK;
}

int cumf_batch_kernel_time(cumf_batch* b, double* ms, int64_t* launches, int64_t* updates)
{
    int rc = check_batch(b);
    if (rc) return rc;
    if (ms) *ms = b->kernel_ms;
    if (launches) *launches = b->launches;
    if (updates) *updates = b->updates;
    return CUMF_OK;
}

}  // extern "C"
