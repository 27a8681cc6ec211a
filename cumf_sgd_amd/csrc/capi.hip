// C ABI of the device path (include/cumf_sgd_amd.h): a context owns the device
// copies of the ratings and factors and drives the gfx950 kernels on its own
// stream.  Replaces the device half of sgd_update_k128
// (singleGPU/sgd_kernel.cu:125-720) and the device kernels of sgd.cu.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <climits>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cumf_sgd_amd.h"
#include "errors.h"
#include "loader.h"
#include "kernels.h"
#include "train_plan.h"

using namespace cumf;

namespace {
thread_local std::string g_last_error;
}

int cumf::fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

namespace {

#define HIP_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return fail(CUMF_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                      \
    } while (0)

template <typename T>
int ensure(T** p, size_t* cap, size_t count)
{
    if (*cap >= count && *p) return CUMF_OK;
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
        *cap = 0;
    }
    if (count == 0) return CUMF_OK;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(CUMF_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    }
    *cap = count;
    return CUMF_OK;
}

template <typename T>
void release(T** p, size_t* cap)
{
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
}

// Device scratch that lives for one call.
struct DevTmp {
    void* p = nullptr;
    ~DevTmp() { if (p) (void)hipFree(p); }
};

int prec_bytes(int prec) { return prec == CUMF_FP32 ? 4 : (prec == CUMF_FP8_E4M3 ? 1 : 2); }

}  // namespace

// One launch of the update kernel: a slice of the uploaded task table.
struct LaunchRec {
    int64_t task_off;
    int32_t task_cnt;
    int32_t num_xcd;
    int32_t blocks;   // workgroups of the launch; 0 = the shape's
};

struct cumf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;

    int32_t* u = nullptr; size_t u_cap = 0;
    int32_t* v = nullptr; size_t v_cap = 0;
    float* r = nullptr; size_t r_cap = 0;
    int64_t nnz = 0;
    int32_t m = 0, n = 0;
    bool have_ratings = false;

    int32_t* tu = nullptr; size_t tu_cap = 0;
    int32_t* tv = nullptr; size_t tv_cap = 0;
    float* tr = nullptr; size_t tr_cap = 0;
    int64_t tnnz = 0;
    bool have_test = false;

    int32_t ux = 1, vy = 1;
    std::vector<int64_t> offsets{0, 0};
    // tiles of cumf_partition_tiles, block-major: tile t holds records
    // [tile_rec[t], tile_rec[t+1]) and P rows [tile_row0[t], + tile_rows[t]);
    // block b's tiles are [block_tile[b], block_tile[b+1]).  Empty when the
    // current partition is not a tile partition.
    std::vector<int64_t> tile_rec;
    std::vector<int32_t> tile_row0, tile_rows, block_tile;
    int32_t tile_rows_max = 0;
    // bumped whenever the ratings, their order, the blocks or the tiles change
    uint64_t data_gen = 0;
    // the task table on the device (tasks) was planned under plan_key: its
    // host copy, one epoch's launches and updates, and the streamed offsets
    bool plan_valid = false;
    uint64_t plan_key = 0;
    std::vector<SgdTask> plan_tasks;
    std::vector<LaunchRec> plan_launches;
    int64_t plan_upd = 0;
    std::vector<int64_t> plan_soff;

    int32_t k = 0, prec = -1;
    void* P = nullptr; size_t P_cap = 0;
    void* Q = nullptr; size_t Q_cap = 0;
    float* ps = nullptr; size_t ps_cap = 0;
    float* qs = nullptr; size_t qs_cap = 0;
    bool have_factors = false;

    float* lr = nullptr; size_t lr_cap = 0;
    // rates the table holds: its first lr_valid entries, for (alpha, beta, lrate) = lr_key
    size_t lr_valid = 0;
    float lr_key[3] = {0, 0, 0};
    SgdTask* tasks = nullptr; size_t tasks_cap = 0;
    int32_t* epoch_dev = nullptr; size_t epoch_dev_cap = 0;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_key = 0;
    double* partials = nullptr; size_t partials_cap = 0;

    // streamed ratings (cumf_stream_ratings): block-ordered SoA in pinned host
    // memory, two device staging buffers of one stratum diagonal each, filled
    // on a copy stream while the other diagonal trains
    bool streamed = false;
    int32_t* hu = nullptr;
    int32_t* hv = nullptr;
    float* hr = nullptr;
    int32_t* su[2] = {nullptr, nullptr};
    int32_t* sv[2] = {nullptr, nullptr};
    float* sr[2] = {nullptr, nullptr};
    hipStream_t copy_stream = nullptr;
    hipEvent_t copy_done[2] = {nullptr, nullptr}, kern_done[2] = {nullptr, nullptr};

    std::vector<hipEvent_t> ev;   // start/stop pairs of training launches
    size_t ev_used = 0;
    double kernel_ms = 0.0;
    int64_t launches = 0, updates = 0;
    cumf_launch_info last{};
};

namespace {

int check_ctx(cumf_ctx* c)
{
    if (!c) return fail(CUMF_ERR_ARG, "null context");
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return fail(CUMF_ERR_HIP, "hipSetDevice(%d): %s", c->device, hipGetErrorString(e));
    return CUMF_OK;
}

// Copy `bytes` from a host or device pointer into device memory.
int copy_in(void* dst, const void* src, size_t bytes, hipStream_t s)
{
    if (bytes == 0) return CUMF_OK;
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
    return CUMF_OK;
}

int upload_soa(cumf_ctx* c, const int32_t* u, const int32_t* v, const float* r, int64_t nnz, bool test)
{
    int32_t** du = test ? &c->tu : &c->u;
    int32_t** dv = test ? &c->tv : &c->v;
    float** dr = test ? &c->tr : &c->r;
    size_t* cu = test ? &c->tu_cap : &c->u_cap;
    size_t* cv = test ? &c->tv_cap : &c->v_cap;
    size_t* cr = test ? &c->tr_cap : &c->r_cap;
    int rc;
    if ((rc = ensure(du, cu, (size_t)nnz)) || (rc = ensure(dv, cv, (size_t)nnz)) || (rc = ensure(dr, cr, (size_t)nnz)))
        return rc;
    if ((rc = copy_in(*du, u, sizeof(int32_t) * nnz, c->stream))) return rc;
    if ((rc = copy_in(*dv, v, sizeof(int32_t) * nnz, c->stream))) return rc;
    if ((rc = copy_in(*dr, r, sizeof(float) * nnz, c->stream))) return rc;
    return CUMF_OK;
}

int upload_aos(cumf_ctx* c, const cumf_node* R, int64_t nnz, bool test)
{
    int32_t** du = test ? &c->tu : &c->u;
    int32_t** dv = test ? &c->tv : &c->v;
    float** dr = test ? &c->tr : &c->r;
    size_t* cu = test ? &c->tu_cap : &c->u_cap;
    size_t* cv = test ? &c->tv_cap : &c->v_cap;
    size_t* cr = test ? &c->tr_cap : &c->r_cap;
    int rc;
    if ((rc = ensure(du, cu, (size_t)nnz)) || (rc = ensure(dv, cv, (size_t)nnz)) || (rc = ensure(dr, cr, (size_t)nnz)))
        return rc;
    if (nnz == 0) return CUMF_OK;
    DevTmp tmp;
    HIP_TRY(hipMalloc(&tmp.p, sizeof(cumf_node) * (size_t)nnz));
    if ((rc = copy_in(tmp.p, R, sizeof(cumf_node) * (size_t)nnz, c->stream))) return rc;
    HIP_TRY(launch_split(tmp.p, nnz, *du, *dv, *dr, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CUMF_OK;
}

int validate_ids(cumf_ctx* c)
{
    DevTmp bad;
    HIP_TRY(hipMalloc(&bad.p, sizeof(unsigned long long)));
    HIP_TRY(hipMemsetAsync(bad.p, 0, sizeof(unsigned long long), c->stream));
    HIP_TRY(launch_check_ids(c->u, c->v, c->nnz, c->m, c->n, (unsigned long long*)bad.p, c->stream));
    unsigned long long hb = 0;
    HIP_TRY(hipMemcpyAsync(&hb, bad.p, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (hb) return fail(CUMF_ERR_ARG, "%llu training ratings have ids outside [0,%d) x [0,%d)", hb, c->m, c->n);
    return CUMF_OK;
}

void drop_streamed(cumf_ctx* c)
{
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : {(void*)c->hu, (void*)c->hv, (void*)c->hr})
        if (p) (void)hipHostFree(p);
    c->hu = c->hv = nullptr;
    c->hr = nullptr;
    for (int b = 0; b < 2; ++b) {
        for (void* p : {(void*)c->su[b], (void*)c->sv[b], (void*)c->sr[b]})
            if (p) (void)hipFree(p);
        c->su[b] = c->sv[b] = nullptr;
        c->sr[b] = nullptr;
        if (c->copy_done[b]) (void)hipEventDestroy(c->copy_done[b]);
        if (c->kern_done[b]) (void)hipEventDestroy(c->kern_done[b]);
        c->copy_done[b] = c->kern_done[b] = nullptr;
    }
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    c->copy_stream = nullptr;
    c->streamed = false;
}

// The tile list describes the records' order and block bounds: any change
// of either drops it.
void drop_tiles(cumf_ctx* c)
{
    ++c->data_gen;
    c->tile_rec.clear();
    c->tile_row0.clear();
    c->tile_rows.clear();
    c->block_tile.clear();
    c->tile_rows_max = 0;
}

int after_ratings(cumf_ctx* c, int64_t nnz, int32_t m, int32_t n)
{
    drop_streamed(c);
    drop_tiles(c);
    c->nnz = nnz;
    c->m = m;
    c->n = n;
    c->ux = c->vy = 1;
    c->offsets.assign({0, nnz});
    int rc = validate_ids(c);
    c->have_ratings = rc == CUMF_OK;
    return rc;
}

hipEvent_t next_event(cumf_ctx* c)
{
    if (c->ev_used == c->ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev.push_back(e);
    }
    return c->ev[c->ev_used++];
}

// Sum the start/stop event pairs recorded since the last harvest.
int harvest_events(cumf_ctx* c)
{
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]));
        c->kernel_ms += ms;
    }
    c->ev_used = 0;
    return CUMF_OK;
}

}  // namespace

extern "C" {

int cumf_abi_version(void) { return CUMF_ABI_VERSION; }

const char* cumf_last_error(void) { return g_last_error.c_str(); }

int cumf_create(int device, cumf_ctx** out)
{
    if (!out) return fail(CUMF_ERR_ARG, "null output pointer");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(CUMF_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0) device = 0;
    device %= count;   // the reference selects gpu % deviceCount (singleGPU/main.cpp:292-294)
    HIP_TRY(hipSetDevice(device));
    cumf_ctx* c = new cumf_ctx;
    c->device = device;
    // a blocking stream: work the caller queued on the null stream (e.g. a
    // framework producing the ratings on the device) completes before ours
    hipError_t se = hipStreamCreate(&c->stream);
    if (se != hipSuccess) {
        delete c;
        return fail(CUMF_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(se));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->cu_count = prop.multiProcessorCount;
    *out = c;
    return CUMF_OK;
}

int cumf_destroy(cumf_ctx* c)
{
    if (!c) return CUMF_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(&c->u, &c->u_cap); release(&c->v, &c->v_cap); release(&c->r, &c->r_cap);
    release(&c->tu, &c->tu_cap); release(&c->tv, &c->tv_cap); release(&c->tr, &c->tr_cap);
    release(&c->P, &c->P_cap); release(&c->Q, &c->Q_cap);
    release(&c->ps, &c->ps_cap); release(&c->qs, &c->qs_cap);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    drop_streamed(c);
    release(&c->lr, &c->lr_cap); release(&c->tasks, &c->tasks_cap); release(&c->partials, &c->partials_cap);
    release(&c->epoch_dev, &c->epoch_dev_cap);
    for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CUMF_OK;
}

int cumf_sync(cumf_ctx* c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CUMF_OK;
}

int cumf_set_ratings(cumf_ctx* c, const cumf_node* R, int64_t nnz, int32_t m, int32_t n)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (nnz < 0 || m <= 0 || n <= 0 || (nnz > 0 && !R)) return fail(CUMF_ERR_ARG, "bad ratings (nnz=%lld m=%d n=%d)", (long long)nnz, m, n);
    c->have_ratings = false;
    if ((rc = upload_aos(c, R, nnz, false))) return rc;
    return after_ratings(c, nnz, m, n);
}

int cumf_set_ratings_soa(cumf_ctx* c, const int32_t* u, const int32_t* v, const float* r, int64_t nnz, int32_t m,
                         int32_t n)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (nnz < 0 || m <= 0 || n <= 0 || (nnz > 0 && (!u || !v || !r)))
        return fail(CUMF_ERR_ARG, "bad ratings (nnz=%lld m=%d n=%d)", (long long)nnz, m, n);
    c->have_ratings = false;
    if ((rc = upload_soa(c, u, v, r, nnz, false))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return after_ratings(c, nnz, m, n);
}

int cumf_alloc_ratings(cumf_ctx* c, int64_t nnz, int32_t m, int32_t n, int32_t** u, int32_t** v, float** r)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (nnz < 0 || m <= 0 || n <= 0 || !u || !v || !r) return fail(CUMF_ERR_ARG, "bad arguments");
    drop_streamed(c);
    c->have_ratings = false;
    if ((rc = ensure(&c->u, &c->u_cap, (size_t)nnz)) || (rc = ensure(&c->v, &c->v_cap, (size_t)nnz)) ||
        (rc = ensure(&c->r, &c->r_cap, (size_t)nnz)))
        return rc;
    c->nnz = nnz;
    c->m = m;
    c->n = n;
    *u = c->u;
    *v = c->v;
    *r = c->r;
    return CUMF_OK;
}

int cumf_commit_ratings(cumf_ctx* c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->u && c->nnz) return fail(CUMF_ERR_STATE, "no rating buffers allocated");
    HIP_TRY(hipDeviceSynchronize());   // the caller may have filled them on any stream
    return after_ratings(c, c->nnz, c->m, c->n);
}

int cumf_read_ratings(cumf_ctx* c, const char* path, int64_t* nnz_out, int32_t* m_out, int32_t* n_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    std::vector<std::string> files;
    std::vector<int64_t> counts;
    bool sharded = false;
    if ((rc = triplet_files(path, &files, &counts, &sharded))) return rc;
    int64_t nnz = 0;
    for (int64_t x : counts) nnz += x;
    c->have_ratings = false;
    if ((rc = ensure(&c->u, &c->u_cap, (size_t)nnz)) || (rc = ensure(&c->v, &c->v_cap, (size_t)nnz)) ||
        (rc = ensure(&c->r, &c->r_cap, (size_t)nnz)))
        return rc;
    int32_t mm[3];
    if ((rc = stream_triplets(files, counts, c->u, c->v, c->r, c->stream, mm))) return rc;
    if (nnz > 0 && mm[2] < 0) return fail(CUMF_ERR_ARG, "%s holds negative ids", path);
    // m, n = largest id + 1; a single file starts them at 1 (singleGPU/sgd.cu:202-203)
    const int32_t m = std::max(sharded ? -1 : 1, mm[0] + 1), n = std::max(sharded ? -1 : 1, mm[1] + 1);
    if ((rc = after_ratings(c, nnz, m, n))) return rc;
    if (nnz_out) *nnz_out = nnz;
    if (m_out) *m_out = m;
    if (n_out) *n_out = n;
    return CUMF_OK;
}

int cumf_read_test(cumf_ctx* c, const char* path, int64_t* nnz_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    std::vector<std::string> files;
    std::vector<int64_t> counts;
    bool sharded = false;
    if ((rc = triplet_files(path, &files, &counts, &sharded))) return rc;
    int64_t nnz = 0;
    for (int64_t x : counts) nnz += x;
    c->have_test = false;
    if ((rc = ensure(&c->tu, &c->tu_cap, (size_t)nnz)) || (rc = ensure(&c->tv, &c->tv_cap, (size_t)nnz)) ||
        (rc = ensure(&c->tr, &c->tr_cap, (size_t)nnz)))
        return rc;
    int32_t mm[3];
    if ((rc = stream_triplets(files, counts, c->tu, c->tv, c->tr, c->stream, mm))) return rc;
    c->tnnz = nnz;
    c->have_test = true;
    if (nnz_out) *nnz_out = nnz;
    return CUMF_OK;
}

int cumf_set_test(cumf_ctx* c, const cumf_node* R, int64_t nnz)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (nnz < 0 || (nnz > 0 && !R)) return fail(CUMF_ERR_ARG, "bad test ratings");
    if ((rc = upload_aos(c, R, nnz, true))) return rc;
    c->tnnz = nnz;
    c->have_test = true;
    return CUMF_OK;
}

int cumf_set_test_soa(cumf_ctx* c, const int32_t* u, const int32_t* v, const float* r, int64_t nnz)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (nnz < 0 || (nnz > 0 && (!u || !v || !r))) return fail(CUMF_ERR_ARG, "bad test ratings");
    if ((rc = upload_soa(c, u, v, r, nnz, true))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->tnnz = nnz;
    c->have_test = true;
    return CUMF_OK;
}

int cumf_get_ratings(cumf_ctx* c, int32_t* u, int32_t* v, float* r)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    if (c->nnz == 0) return CUMF_OK;
    if (c->streamed) {
        if (u) HIP_TRY(hipMemcpy(u, c->hu, sizeof(int32_t) * c->nnz, hipMemcpyDefault));
        if (v) HIP_TRY(hipMemcpy(v, c->hv, sizeof(int32_t) * c->nnz, hipMemcpyDefault));
        if (r) HIP_TRY(hipMemcpy(r, c->hr, sizeof(float) * c->nnz, hipMemcpyDefault));
        return CUMF_OK;
    }
    if (u) HIP_TRY(hipMemcpyAsync(u, c->u, sizeof(int32_t) * c->nnz, hipMemcpyDefault, c->stream));
    if (v) HIP_TRY(hipMemcpyAsync(v, c->v, sizeof(int32_t) * c->nnz, hipMemcpyDefault, c->stream));
    if (r) HIP_TRY(hipMemcpyAsync(r, c->r, sizeof(float) * c->nnz, hipMemcpyDefault, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CUMF_OK;
}

int cumf_rating_moments(cumf_ctx* c, double* sum, double* sum_sq)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->streamed) return fail(CUMF_ERR_STATE, "streamed ratings stay in host memory in their stratum blocks");
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    const size_t need = 2 * (size_t)moments_blocks(c->nnz) + 2;
    if ((rc = ensure(&c->partials, &c->partials_cap, need))) return rc;
    double* out2 = c->partials + need - 2;
    HIP_TRY(launch_moments(c->r, c->nnz, c->partials, out2, c->stream));
    double h[2];
    HIP_TRY(hipMemcpyAsync(h, out2, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (sum) *sum = h[0];
    if (sum_sq) *sum_sq = h[1];
    return CUMF_OK;
}

int cumf_remap(cumf_ctx* c, const int32_t* p_map, const int32_t* q_map)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->streamed) return fail(CUMF_ERR_STATE, "streamed ratings stay in host memory in their stratum blocks");
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    if (!p_map || !q_map) return fail(CUMF_ERR_ARG, "null map");
    DevTmp dp, dq;
    HIP_TRY(hipMalloc(&dp.p, sizeof(int32_t) * (size_t)c->m));
    HIP_TRY(hipMalloc(&dq.p, sizeof(int32_t) * (size_t)c->n));
    if ((rc = copy_in(dp.p, p_map, sizeof(int32_t) * c->m, c->stream))) return rc;
    if ((rc = copy_in(dq.p, q_map, sizeof(int32_t) * c->n, c->stream))) return rc;
    // maps must be permutations of [0,m) / [0,n): check the range on the host copy
    std::vector<int32_t> hp(c->m), hq(c->n);
    HIP_TRY(hipMemcpyAsync(hp.data(), dp.p, sizeof(int32_t) * c->m, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(hq.data(), dq.p, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    {
        std::vector<char> seen(std::max(c->m, c->n), 0);
        for (int32_t x : hp) {
            if (x < 0 || x >= c->m || seen[x]) return fail(CUMF_ERR_ARG, "p_map is not a permutation of [0,%d)", c->m);
            seen[x] = 1;
        }
        std::fill(seen.begin(), seen.end(), 0);
        for (int32_t x : hq) {
            if (x < 0 || x >= c->n || seen[x]) return fail(CUMF_ERR_ARG, "q_map is not a permutation of [0,%d)", c->n);
            seen[x] = 1;
        }
    }
    HIP_TRY(launch_remap(c->u, c->v, c->nnz, (const int32_t*)dp.p, (const int32_t*)dq.p, c->m, c->n, c->stream));
    if (c->have_test)
        HIP_TRY(launch_remap(c->tu, c->tv, c->tnnz, (const int32_t*)dp.p, (const int32_t*)dq.p, c->m, c->n, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->ux = c->vy = 1;
    c->offsets.assign({0, c->nnz});
    drop_tiles(c);
    return CUMF_OK;
}

int cumf_scale_ratings(cumf_ctx* c, float scale)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->streamed) return fail(CUMF_ERR_STATE, "streamed ratings stay in host memory in their stratum blocks");
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    HIP_TRY(launch_scale(c->r, c->nnz, scale, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CUMF_OK;
}

}  // extern "C"

namespace {

// Partition with per-axis segment boundaries bu[0..ux], bv[0..vy].
int partition_with(cumf_ctx* c, const std::vector<int64_t>& bu, const std::vector<int64_t>& bv, int64_t* offsets)
{
    const int ux = (int)bu.size() - 1, vy = (int)bv.size() - 1;
    const int nb = ux * vy;
    if (nb < 1 || nb > 4096) return fail(CUMF_ERR_ARG, "grid of %d blocks unsupported (1..4096)", nb);
    drop_tiles(c);
    const int64_t nnz = c->nnz;
    DevTmp useg, vseg, bnd, hist, tot, base, off, uo, vo, ro;
    HIP_TRY(hipMalloc(&useg.p, sizeof(int32_t) * (size_t)c->m));
    HIP_TRY(hipMalloc(&vseg.p, sizeof(int32_t) * (size_t)c->n));
    HIP_TRY(hipMalloc(&bnd.p, sizeof(int64_t) * (size_t)(ux + vy + 2)));
    HIP_TRY(hipMemcpyAsync(bnd.p, bu.data(), sizeof(int64_t) * (ux + 1), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync((int64_t*)bnd.p + ux + 1, bv.data(), sizeof(int64_t) * (vy + 1), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(launch_seg_table((int32_t*)useg.p, c->m, (const int64_t*)bnd.p, ux, c->stream));
    HIP_TRY(launch_seg_table((int32_t*)vseg.p, c->n, (const int64_t*)bnd.p + ux + 1, vy, c->stream));
    std::vector<int64_t> host(2 * (size_t)nb);
    if (nnz > 0) {
        const int64_t he = part_hist_entries(nnz, nb);
        HIP_TRY(hipMalloc(&hist.p, sizeof(int32_t) * (size_t)he));
        HIP_TRY(hipMalloc(&off.p, sizeof(int64_t) * (size_t)he));
        HIP_TRY(hipMalloc(&tot.p, sizeof(int64_t) * (size_t)nb));
        HIP_TRY(hipMalloc(&base.p, sizeof(int64_t) * (size_t)nb));
        HIP_TRY(hipMalloc(&uo.p, sizeof(int32_t) * (size_t)nnz));
        HIP_TRY(hipMalloc(&vo.p, sizeof(int32_t) * (size_t)nnz));
        HIP_TRY(hipMalloc(&ro.p, sizeof(float) * (size_t)nnz));
        HIP_TRY(launch_partition(c->u, c->v, c->r, nnz, (const int32_t*)useg.p, (const int32_t*)vseg.p, vy, nb,
                                 (int32_t*)hist.p, (int64_t*)tot.p, (int64_t*)base.p, (int64_t*)off.p,
                                 (int32_t*)uo.p, (int32_t*)vo.p, (float*)ro.p, c->stream, host.data()));
        // adopt the partitioned arrays
        std::swap(*(void**)&c->u, uo.p);
        std::swap(*(void**)&c->v, vo.p);
        std::swap(*(void**)&c->r, ro.p);
        c->u_cap = c->v_cap = c->r_cap = (size_t)nnz;
    } else {
        std::fill(host.begin(), host.end(), 0);
    }
    c->ux = ux;
    c->vy = vy;
    c->offsets.assign((size_t)nb + 1, 0);
    for (int b = 0; b < nb; ++b) c->offsets[b + 1] = c->offsets[b] + host[b];
    if (offsets) std::copy(c->offsets.begin(), c->offsets.end(), offsets);
    return CUMF_OK;
}

// Boundaries splitting `rows` ids into s contiguous ranges, from the degree
// histogram.  A stratum launch runs each diagonal's blocks on the XCDs,
// largest first onto the least-loaded one, so what a partition costs is the
// most loaded XCD of each diagonal, and every diagonal holds one block of
// every column range.
// - s <= kNumXcd (each XCD trains one block per diagonal, the tile
//   schedule's 8 x 8): the largest range sets every launch.  M is the least
//   capacity for which ranges filled in id order, one closed before the row
//   that would take it past M, fit in s ranges (a binary search; M is at
//   least the largest row and nnz / s); each range is then closed at the
//   row boundary nearer to the remaining ratings' even share, unless the
//   rows after it would then not fit in the ranges left at M, and always
//   before the row that would take it past M, so no range exceeds M.  A cut at
//   the first row reaching the next 1/s of the ratings left the 99M
//   workload's item range holding its most popular item (5.8% of all
//   ratings) at 1.2 times the mean, and every launch that much longer.
// - s > kNumXcd (several blocks per XCD): a range ends at the first row
//   reaching the next 1/s of the ratings.  A range overfull by a popular
//   row is followed by a short one, and the largest-first placement pairs
//   them on the XCDs (at s = 24 the most loaded XCD stays within 1% of the
//   mean, where the capped cut above, whose ranges cannot go below the most
//   popular item's 1.4 shares, leaves it at 1-3% above).
std::vector<int64_t> balanced_bounds(const std::vector<int32_t>& deg, int s, int64_t nnz)
{
    const int64_t rows = (int64_t)deg.size();
    std::vector<int64_t> b(s + 1, rows);
    b[0] = 0;
    if (s > kNumXcd) {
        int64_t run = 0;
        int next = 1;
        for (int64_t i = 0; i < rows && next < s; ++i) {
            run += deg[(size_t)i];
            while (next < s && run * s >= nnz * next) b[next++] = i + 1;
        }
        for (int j = 1; j <= s; ++j) b[j] = std::max(b[j], b[j - 1]);
        return b;
    }
    int64_t most = 0;
    for (int32_t d : deg) most = std::max<int64_t>(most, d);
    auto ranges = [&](int64_t M) {
        int64_t n = 1, run = 0;
        for (int32_t d : deg) {
            if (run + d > M) {
                ++n;
                run = 0;
            }
            run += d;
        }
        return n;
    };
    int64_t lo = std::max(most, (nnz + s - 1) / s), hi = std::max(lo, nnz);
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (ranges(mid) <= s) hi = mid;
        else lo = mid + 1;
    }
    // need[i]: the ranges of at most M that rows [i, rows) take when filled
    // in id order (the fewest any cut of them allows), through end[i], the
    // row after the first such range (two pointers)
    std::vector<int32_t> need((size_t)rows + 1, 0), end((size_t)rows, 0);
    {
        int64_t j = 0, sum = 0;
        for (int64_t i = 0; i < rows; ++i) {
            if (j < i) j = i, sum = 0;
            while (j < rows && sum + deg[(size_t)j] <= lo) sum += deg[(size_t)j++];
            end[(size_t)i] = (int32_t)j;
            sum -= deg[(size_t)i];
        }
        for (int64_t i = rows - 1; i >= 0; --i) need[(size_t)i] = 1 + need[(size_t)end[(size_t)i]];
    }
    int next = 1;
    int64_t run = 0, left = nnz;
    for (int64_t i = 0; i < rows; ++i) {
        const int64_t d = deg[(size_t)i];
        if (next < s && run > 0) {
            // close before row i when it would pass M, or when the range is
            // nearer the even share without it, (left / k - run) < (run + d -
            // left / k) in integers, and rows [i, rows) still fit in the
            // k - 1 ranges left (else a later range would pass M)
            const int64_t k = s - next + 1;
            if (run + d > lo || (2 * run * k + d * k > 2 * left && need[(size_t)i] <= k - 1)) {
                b[next++] = i;
                left -= run;
                run = 0;
            }
        }
        run += d;
    }
    return b;
}

}  // namespace

extern "C" {

int cumf_partition(cumf_ctx* c, int32_t ux, int32_t vy, int64_t* offsets)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->streamed) return fail(CUMF_ERR_STATE, "streamed ratings stay in host memory in their stratum blocks");
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    if (ux < 1 || vy < 1) return fail(CUMF_ERR_ARG, "bad grid %d x %d", ux, vy);
    // segment lengths of grid_problem (singleGPU/sgd.cu:419-423)
    const int64_t useg = ux == 1 ? c->m : (int64_t)std::ceil((double)c->m / ux);
    const int64_t vseg = vy == 1 ? c->n : (int64_t)std::ceil((double)c->n / vy);
    std::vector<int64_t> bu(ux + 1), bv(vy + 1);
    for (int i = 0; i <= ux; ++i) bu[i] = std::min<int64_t>((int64_t)i * useg, c->m);
    for (int i = 0; i <= vy; ++i) bv[i] = std::min<int64_t>((int64_t)i * vseg, c->n);
    bu[ux] = c->m;
    bv[vy] = c->n;
    return partition_with(c, bu, bv, offsets);
}

}  // extern "C"

namespace {

// Ratings per row of P (hu) and of Q (hv), counted on the device.
int rating_degrees(cumf_ctx* c, std::vector<int32_t>& hu, std::vector<int32_t>& hv)
{
    DevTmp du, dv;
    HIP_TRY(hipMalloc(&du.p, sizeof(int32_t) * (size_t)std::max(1, c->m)));
    HIP_TRY(hipMalloc(&dv.p, sizeof(int32_t) * (size_t)std::max(1, c->n)));
    HIP_TRY(hipMemsetAsync(du.p, 0, sizeof(int32_t) * (size_t)c->m, c->stream));
    HIP_TRY(hipMemsetAsync(dv.p, 0, sizeof(int32_t) * (size_t)c->n, c->stream));
    HIP_TRY(launch_degree(c->u, c->nnz, (int32_t*)du.p, c->m, c->stream));
    HIP_TRY(launch_degree(c->v, c->nnz, (int32_t*)dv.p, c->n, c->stream));
    hu.assign((size_t)c->m, 0);
    hv.assign((size_t)c->n, 0);
    HIP_TRY(hipMemcpyAsync(hu.data(), du.p, sizeof(int32_t) * c->m, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(hv.data(), dv.p, sizeof(int32_t) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CUMF_OK;
}

}  // namespace

extern "C" {

int cumf_partition_strata(cumf_ctx* c, int32_t s, int64_t* offsets)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->streamed) return fail(CUMF_ERR_STATE, "streamed ratings stay in host memory in their stratum blocks");
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    if (s < 1 || s > 64) return fail(CUMF_ERR_ARG, "strata %d out of range 1..64", s);
    std::vector<int32_t> hu, hv;
    if ((rc = rating_degrees(c, hu, hv))) return rc;
    return partition_with(c, balanced_bounds(hu, s, c->nnz), balanced_bounds(hv, s, c->nnz), offsets);
}

int cumf_order_blocks(cumf_ctx* c, int32_t by)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->streamed) return fail(CUMF_ERR_STATE, "streamed ratings stay in host memory in their stratum blocks");
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    if (by != CUMF_ORDER_ROWS && by != CUMF_ORDER_COLUMNS) return fail(CUMF_ERR_ARG, "bad order %d", by);
    const int64_t nnz = c->nnz;
    if (nnz > (int64_t)UINT32_MAX) return fail(CUMF_ERR_ARG, "%lld ratings: ordering supports up to 2^32 - 1",
                                                (long long)nnz);
    if (nnz == 0) return CUMF_OK;
    const int nb = (int)c->offsets.size() - 1;
    const uint64_t stride = (uint64_t)std::max(1, by == CUMF_ORDER_COLUMNS ? c->n : c->m);
    const size_t kb = order_key_bytes(nb, stride);
    DevTmp dof, keys, keys_out, idx, idx_out, uo, vo, ro, temp;
    size_t temp_bytes = 0;
    HIP_TRY(hipMalloc(&dof.p, sizeof(int64_t) * (size_t)(nb + 1)));
    HIP_TRY(hipMalloc(&keys.p, kb * (size_t)nnz));
    HIP_TRY(hipMalloc(&keys_out.p, kb * (size_t)nnz));
    HIP_TRY(hipMalloc(&idx.p, sizeof(uint32_t) * (size_t)nnz));
    HIP_TRY(hipMalloc(&idx_out.p, sizeof(uint32_t) * (size_t)nnz));
    HIP_TRY(hipMalloc(&uo.p, sizeof(int32_t) * (size_t)nnz));
    HIP_TRY(hipMalloc(&vo.p, sizeof(int32_t) * (size_t)nnz));
    HIP_TRY(hipMalloc(&ro.p, sizeof(float) * (size_t)nnz));
    HIP_TRY(hipMemcpyAsync(dof.p, c->offsets.data(), sizeof(int64_t) * (size_t)(nb + 1), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(launch_order_blocks(nnz, by == CUMF_ORDER_COLUMNS ? c->v : c->u, (const int64_t*)dof.p, nb, stride,
                                keys.p, keys_out.p, (uint32_t*)idx.p, (uint32_t*)idx_out.p, c->u, c->v, c->r,
                                (int32_t*)uo.p, (int32_t*)vo.p, (float*)ro.p, &temp.p, &temp_bytes, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::swap(*(void**)&c->u, uo.p);
    std::swap(*(void**)&c->v, vo.p);
    std::swap(*(void**)&c->r, ro.p);
    c->u_cap = c->v_cap = c->r_cap = (size_t)nnz;
    drop_tiles(c);
    return CUMF_OK;
}

int32_t cumf_tile_rows(int32_t k, int32_t precision)
{
    if (k < 1 || precision < CUMF_FP32 || precision > CUMF_BF16_SR) return 0;
    // At k = 128: 128 FP32 rows (64 KB, two 8-wave workgroups per CU), 96
    // rows of the narrower formats (24 KB at FP16: four 4-wave workgroups per
    // CU, which the registers of those waves allow); more rows at smaller k,
    // at most 64 KB.  Measured fastest on the 99M workload (FP16: 64-128 rows
    // within 4%, 96 best).
    const int64_t base = precision == CUMF_FP32 ? 128 : 96;
    const int64_t rows = base * 128 / k, cap = kSgdTileDefaultBytes / ((int64_t)k * prec_bytes(precision));
    return (int32_t)std::max<int64_t>(1, std::min(rows, cap));
}

int cumf_partition_tiles(cumf_ctx* c, int32_t s, int32_t tile_rows, int64_t* offsets)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->streamed) return fail(CUMF_ERR_STATE, "streamed ratings stay in host memory in their stratum blocks");
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    if (s == 0) s = kTileStrata;
    if (s < 1 || s > 64) return fail(CUMF_ERR_ARG, "strata %d out of range 1..64", s);
    if (tile_rows < 0) return fail(CUMF_ERR_ARG, "tile_rows=%d must be >= 0", tile_rows);
    if (c->nnz > (int64_t)UINT32_MAX)
        return fail(CUMF_ERR_ARG, "%lld ratings: tiles support up to 2^32 - 1", (long long)c->nnz);
    if (tile_rows == 0) tile_rows = c->have_factors ? cumf_tile_rows(c->k, c->prec) : cumf_tile_rows(128, CUMF_FP16);
    std::vector<int32_t> hu, hv;
    if ((rc = rating_degrees(c, hu, hv))) return rc;
    const std::vector<int64_t> bu = balanced_bounds(hu, s, c->nnz);
    if ((rc = partition_with(c, bu, balanced_bounds(hv, s, c->nnz), offsets))) return rc;

    // Tiles of each row band: consecutive rows, a tile closed when it holds
    // tile_rows rows or its share of the band's ratings (the band's count
    // over ceil(rows / tile_rows)), so that a band's tiles carry about equal
    // work where the row cap allows.
    std::vector<int32_t> tile_of((size_t)std::max(1, c->m), 0);
    std::vector<std::vector<std::pair<int32_t, int32_t>>> band_tiles((size_t)s);   // (first row, rows)
    int T = 1;
    for (int i = 0; i < s; ++i) {
        const int64_t r0 = bu[(size_t)i], r1 = bu[(size_t)i + 1];
        if (r1 <= r0) continue;
        int64_t band_nnz = 0;
        for (int64_t x = r0; x < r1; ++x) band_nnz += hu[(size_t)x];
        const int64_t nt = (r1 - r0 + tile_rows - 1) / tile_rows;
        const int64_t share = std::max<int64_t>(1, (band_nnz + nt - 1) / nt);
        int64_t first = r0, z = 0;
        for (int64_t x = r0; x < r1; ++x) {
            tile_of[(size_t)x] = (int32_t)band_tiles[(size_t)i].size();
            z += hu[(size_t)x];
            if (x + 1 - first == tile_rows || z >= share || x + 1 == r1) {
                band_tiles[(size_t)i].emplace_back((int32_t)first, (int32_t)(x + 1 - first));
                first = x + 1;
                z = 0;
            }
        }
        T = std::max(T, (int)band_tiles[(size_t)i].size());
    }
    const int nb = s * s;
    const int64_t nnz = c->nnz;
    if (nnz > 0) {
        const uint64_t stride = (uint64_t)std::max(1, c->n);
        const size_t kb = order_key_bytes(nb * T, stride);
        DevTmp dof, dtile, cnt, keys, keys_out, idx, idx_out, uo, vo, ro, temp;
        size_t temp_bytes = 0;
        HIP_TRY(hipMalloc(&dof.p, sizeof(int64_t) * (size_t)(nb + 1)));
        HIP_TRY(hipMalloc(&dtile.p, sizeof(int32_t) * tile_of.size()));
        HIP_TRY(hipMalloc(&cnt.p, sizeof(int32_t) * (size_t)nb * T));
        HIP_TRY(hipMalloc(&keys.p, kb * (size_t)nnz));
        HIP_TRY(hipMalloc(&keys_out.p, kb * (size_t)nnz));
        HIP_TRY(hipMalloc(&idx.p, sizeof(uint32_t) * (size_t)nnz));
        HIP_TRY(hipMalloc(&idx_out.p, sizeof(uint32_t) * (size_t)nnz));
        HIP_TRY(hipMalloc(&uo.p, sizeof(int32_t) * (size_t)nnz));
        HIP_TRY(hipMalloc(&vo.p, sizeof(int32_t) * (size_t)nnz));
        HIP_TRY(hipMalloc(&ro.p, sizeof(float) * (size_t)nnz));
        HIP_TRY(hipMemcpyAsync(dof.p, c->offsets.data(), sizeof(int64_t) * (size_t)(nb + 1), hipMemcpyHostToDevice,
                               c->stream));
        HIP_TRY(hipMemcpyAsync(dtile.p, tile_of.data(), sizeof(int32_t) * tile_of.size(), hipMemcpyHostToDevice,
                               c->stream));
        HIP_TRY(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (size_t)nb * T, c->stream));
        HIP_TRY(launch_order_tiles(nnz, c->u, c->v, c->r, (const int64_t*)dof.p, nb, (const int32_t*)dtile.p, T, stride,
                                   keys.p, keys_out.p, (uint32_t*)idx.p, (uint32_t*)idx_out.p, (int32_t*)cnt.p,
                                   (int32_t*)uo.p, (int32_t*)vo.p, (float*)ro.p, &temp.p, &temp_bytes, c->stream));
        std::vector<int32_t> counts((size_t)nb * T);
        HIP_TRY(hipMemcpyAsync(counts.data(), cnt.p, sizeof(int32_t) * counts.size(), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        std::swap(*(void**)&c->u, uo.p);
        std::swap(*(void**)&c->v, vo.p);
        std::swap(*(void**)&c->r, ro.p);
        c->u_cap = c->v_cap = c->r_cap = (size_t)nnz;
        // the non-empty tiles, block by block (window draws are tagged with
        // a tile's index in this list)
        c->block_tile.assign((size_t)nb + 1, 0);
        c->tile_rec.assign(1, 0);
        for (int b = 0; b < nb; ++b) {
            const int i = b / s;
            int64_t at = c->offsets[(size_t)b];
            for (size_t t = 0; t < band_tiles[(size_t)i].size(); ++t) {
                const int32_t cn = counts[(size_t)b * T + t];
                if (cn == 0) continue;
                at += cn;
                c->tile_rec.push_back(at);
                c->tile_row0.push_back(band_tiles[(size_t)i][t].first);
                c->tile_rows.push_back(band_tiles[(size_t)i][t].second);
            }
            if (at != c->offsets[(size_t)b + 1]) {
                drop_tiles(c);
                return fail(CUMF_ERR_STATE, "tile counts of block %d do not add up to its ratings", b);
            }
            c->block_tile[(size_t)b + 1] = (int32_t)c->tile_row0.size();
        }
        if (c->tile_row0.size() >= (1u << 24)) {
            drop_tiles(c);
            return fail(CUMF_ERR_ARG, "%zu tiles: at most 2^24 - 1 (raise tile_rows)", c->tile_row0.size());
        }
    } else {
        c->block_tile.assign((size_t)nb + 1, 0);
        c->tile_rec.assign(1, 0);
    }
    c->tile_rows_max = tile_rows;
    ++c->data_gen;
    return CUMF_OK;
}

int cumf_get_tiles(cumf_ctx* c, int64_t* count, int64_t* rec_off, int32_t* row_begin, int32_t* row_count)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!count) return fail(CUMF_ERR_ARG, "null count");
    if (c->block_tile.empty()) return fail(CUMF_ERR_STATE, "no tile partition (cumf_partition_tiles)");
    const size_t nt = c->tile_row0.size();
    *count = (int64_t)nt;
    if (rec_off) std::copy(c->tile_rec.begin(), c->tile_rec.end(), rec_off);
    if (row_begin) std::copy(c->tile_row0.begin(), c->tile_row0.end(), row_begin);
    if (row_count) std::copy(c->tile_rows.begin(), c->tile_rows.end(), row_count);
    return CUMF_OK;
}

int cumf_default_schedule(cumf_ctx* c, int32_t k, int32_t precision, int32_t* schedule, int32_t* stratum_rounds,
                          int32_t* order)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!schedule || !stratum_rounds || !order) return fail(CUMF_ERR_ARG, "null output");
    *order = 0;
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    if (k != 32 && k != 64 && k != 128) return fail(CUMF_ERR_ARG, "k=%d unsupported (32, 64 or 128)", k);
    if (precision < CUMF_FP32 || precision > CUMF_BF16_SR) return fail(CUMF_ERR_ARG, "bad precision %d", precision);
    *stratum_rounds = 1;
    // streamed ratings train only as strata, one pass
    if (c->streamed) {
        *schedule = CUMF_SCHED_STRATUM;
        return CUMF_OK;
    }
    // factors that fit in the L2s: conflict-free strata (no row is cached
    // by two XCDs within a launch) with four passes over the diagonals per
    // epoch, so that a row meets every range of the other axis four times
    // per epoch; a small problem's launches are short, the passes cost little
    if (confine_to_one_xcd(CUMF_SCHED_HOGWILD, 0, c->m, c->n, k, precision)) {
        *schedule = CUMF_SCHED_STRATUM;
        *stratum_rounds = kSmallFactorRounds;
        return CUMF_OK;
    }
    // a row and a column that each carry a large share of the ratings: one
    // stratum diagonal at a time pulls both towards one range of the other
    // axis per sub-epoch, while Hogwild interleaves them (and on factors this
    // large the per-XCD copies of a row are evicted long before a launch
    // ends); otherwise the conflict-free strata in one pass
    std::vector<int32_t> hu, hv;
    if ((rc = rating_degrees(c, hu, hv))) return rc;
    const int64_t mu = hu.empty() ? 0 : *std::max_element(hu.begin(), hu.end());
    const int64_t mv = hv.empty() ? 0 : *std::max_element(hv.begin(), hv.end());
    const bool both_skewed = mu * kSkewedShareInv >= c->nnz && mv * kSkewedShareInv >= c->nnz;
    *schedule = both_skewed ? CUMF_SCHED_HOGWILD : CUMF_SCHED_STRATUM;
    // records grouped by column, so that consecutive samples of a worker
    // share the column's row (kept in registers, not stored and reloaded):
    // always for the one-pass strata; for Hogwild on doubly skewed data with
    // round-to-nearest rows (FP32, FP16, BF16: within half of a
    // 0.005-standard-deviation bar there, at twice the speed), not with
    // stochastically rounded ones, whose write-back noise on the hot rows
    // takes the grouped run to the bar
    const bool stochastic = precision == CUMF_BF16_SR || precision == CUMF_FP8_E4M3;
    if (c->nnz <= (int64_t)UINT32_MAX && (*schedule == CUMF_SCHED_STRATUM || !stochastic)) *order = CUMF_ORDER_COLUMNS;
    // strata over a P too large for the caches (its rows then come from HBM
    // at random): records grouped by row instead, so that P's rows are the
    // ones kept in registers
    if (*order && *schedule == CUMF_SCHED_STRATUM && (int64_t)c->m * k * prec_bytes(precision) > kHbmResidentP)
        *order = CUMF_ORDER_ROWS;
    // Where that is strata ordered by column, the round-to-nearest formats
    // train row tiles instead: a tile keeps its P rows in LDS, so the L2 sees
    // only the item rows and the records (99M Netflix shape: FP32 9.2e9 ->
    // 19.1e9 updates/s, FP16 21.5e9 -> 25.3e9).  FP32 makes two passes over
    // the diagonals per epoch (one pass ends at the 0.005 bar there, two
    // 0.003 above the serialized oracle).  The stochastically rounded formats
    // spend more VALU work per update and keep the strata (BF16-SR 15.8e9
    // against 12.9e9 in tiles).
    if (*schedule == CUMF_SCHED_STRATUM && *order == CUMF_ORDER_COLUMNS && !stochastic) {
        *schedule = CUMF_SCHED_TILE;
        *stratum_rounds = precision == CUMF_FP32 ? kTileFp32Rounds : 1;
        *order = 0;
    }
    return CUMF_OK;
}

int cumf_stream_ratings(cumf_ctx* c, const cumf_node* R, int64_t nnz, int32_t m, int32_t n, int32_t s,
                        int64_t* offsets)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (nnz < 1 || m <= 0 || n <= 0 || !R) return fail(CUMF_ERR_ARG, "bad ratings (nnz=%lld m=%d n=%d)", (long long)nnz, m, n);
    if (s < 2 || s > 64 || s % 2) return fail(CUMF_ERR_ARG, "streamed strata need an even s in 2..64 (got %d)", s);
    drop_streamed(c);
    c->have_ratings = false;
    // the same blocks as cumf_partition_strata: nnz-balanced id ranges from
    // the degree histograms, then a stable counting sort by block
    std::vector<int32_t> du((size_t)m, 0), dv((size_t)n, 0);
    for (int64_t i = 0; i < nnz; ++i) {
        const cumf_node& x = R[i];
        if (x.u < 0 || x.u >= m || x.v < 0 || x.v >= n)
            return fail(CUMF_ERR_ARG, "rating %lld has ids (%d, %d) outside [0,%d) x [0,%d)", (long long)i, x.u, x.v, m, n);
        ++du[(size_t)x.u];
        ++dv[(size_t)x.v];
    }
    const std::vector<int64_t> bu = balanced_bounds(du, s, nnz), bv = balanced_bounds(dv, s, nnz);
    std::vector<uint8_t> useg((size_t)m), vseg((size_t)n);
    for (int i = 0; i < s; ++i) {
        for (int64_t x = bu[i]; x < bu[i + 1]; ++x) useg[(size_t)x] = (uint8_t)i;
        for (int64_t x = bv[i]; x < bv[i + 1]; ++x) vseg[(size_t)x] = (uint8_t)i;
    }
    const int nb = s * s;
    std::vector<int64_t> off((size_t)nb + 1, 0);
    for (int64_t i = 0; i < nnz; ++i) ++off[(size_t)(useg[(size_t)R[i].u] * s + vseg[(size_t)R[i].v]) + 1];
    for (int b = 0; b < nb; ++b) off[b + 1] += off[b];
    struct Pinned {
        void* p = nullptr;
        ~Pinned() { if (p) (void)hipHostFree(p); }
    } pu, pv, pr;
    HIP_TRY(hipHostMalloc(&pu.p, sizeof(int32_t) * (size_t)nnz, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&pv.p, sizeof(int32_t) * (size_t)nnz, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&pr.p, sizeof(float) * (size_t)nnz, hipHostMallocDefault));
    {
        std::vector<int64_t> pos(off.begin(), off.end() - 1);
        int32_t* hu = (int32_t*)pu.p;
        int32_t* hv = (int32_t*)pv.p;
        float* hr = (float*)pr.p;
        for (int64_t i = 0; i < nnz; ++i) {
            const cumf_node& x = R[i];
            const int64_t d = pos[(size_t)(useg[(size_t)x.u] * s + vseg[(size_t)x.v])]++;
            hu[d] = x.u;
            hv[d] = x.v;
            hr[d] = x.rate;
        }
    }
    // staging: the largest diagonal
    int64_t cap = 1;
    for (int d = 0; d < s; ++d) {
        int64_t t = 0;
        for (int i = 0; i < s; ++i) {
            const int blk = i * s + (i + d) % s;
            t += off[(size_t)blk + 1] - off[(size_t)blk];
        }
        cap = std::max(cap, t);
    }
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(hipMalloc(&c->su[b], sizeof(int32_t) * (size_t)cap));
        HIP_TRY(hipMalloc(&c->sv[b], sizeof(int32_t) * (size_t)cap));
        HIP_TRY(hipMalloc(&c->sr[b], sizeof(float) * (size_t)cap));
        HIP_TRY(hipEventCreateWithFlags(&c->copy_done[b], hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&c->kern_done[b], hipEventDisableTiming));
    }
    HIP_TRY(hipStreamCreate(&c->copy_stream));
    c->hu = (int32_t*)pu.p;
    c->hv = (int32_t*)pv.p;
    c->hr = (float*)pr.p;
    pu.p = pv.p = pr.p = nullptr;
    // the device copies of earlier ratings are not needed any more
    release(&c->u, &c->u_cap);
    release(&c->v, &c->v_cap);
    release(&c->r, &c->r_cap);
    c->nnz = nnz;
    c->m = m;
    c->n = n;
    c->ux = c->vy = s;
    c->offsets = off;
    drop_tiles(c);
    c->streamed = true;
    c->have_ratings = true;
    if (offsets) std::copy(off.begin(), off.end(), offsets);
    return CUMF_OK;
}

int cumf_set_blocks(cumf_ctx* c, int32_t ux, int32_t vy, const int64_t* offsets)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->streamed) return fail(CUMF_ERR_STATE, "streamed ratings stay in host memory in their stratum blocks");
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    if (ux < 1 || vy < 1 || (int64_t)ux * vy > 4096 || !offsets) return fail(CUMF_ERR_ARG, "bad block grid");
    const int nb = ux * vy;
    if (offsets[0] != 0 || offsets[nb] != c->nnz) return fail(CUMF_ERR_ARG, "block offsets must span [0, nnz]");
    for (int b = 0; b < nb; ++b)
        if (offsets[b + 1] < offsets[b]) return fail(CUMF_ERR_ARG, "block offsets must be ascending");
    c->ux = ux;
    c->vy = vy;
    c->offsets.assign(offsets, offsets + nb + 1);
    drop_tiles(c);
    return CUMF_OK;
}

int cumf_init_factors(cumf_ctx* c, int32_t k, int32_t precision, uint64_t seed)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "set ratings (m, n) before factors");
    if (k != 32 && k != 64 && k != 128) return fail(CUMF_ERR_ARG, "k=%d unsupported (32, 64 or 128)", k);
    if (precision < CUMF_FP32 || precision > CUMF_BF16_SR) return fail(CUMF_ERR_ARG, "bad precision %d", precision);
    const size_t eb = (size_t)prec_bytes(precision);
    char** P = (char**)&c->P;
    char** Q = (char**)&c->Q;
    if ((rc = ensure(P, &c->P_cap, (size_t)c->m * k * eb))) return rc;
    if ((rc = ensure(Q, &c->Q_cap, (size_t)c->n * k * eb))) return rc;
    if (precision == CUMF_FP8_E4M3) {
        if ((rc = ensure(&c->ps, &c->ps_cap, (size_t)c->m))) return rc;
        if ((rc = ensure(&c->qs, &c->qs_cap, (size_t)c->n))) return rc;
    }
    c->k = k;
    c->prec = precision;
    const float s = (float)std::sqrt(1.0 / k);   // singleGPU/sgd.cu:519
    HIP_TRY(launch_init(c->P, c->ps, c->m, k, precision, s, seed, 0, kStreamInitP, c->stream));
    HIP_TRY(launch_init(c->Q, c->qs, c->n, k, precision, s, seed, 0, kStreamInitQ, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->have_factors = true;
    return CUMF_OK;
}

int cumf_set_model(cumf_ctx* c, int32_t m, int32_t n, int32_t k, const float* P, const float* Q)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (m < 0 || n < 0 || k < 1 || k > 4096 || (m && !P) || (n && !Q)) return fail(CUMF_ERR_ARG, "bad model arguments");
    if (c->have_ratings && (c->m != m || c->n != n))
        return fail(CUMF_ERR_ARG, "model is %d x %d but the ratings are %d x %d", m, n, c->m, c->n);
    char** Pp = (char**)&c->P;
    char** Qp = (char**)&c->Q;
    if ((rc = ensure(Pp, &c->P_cap, (size_t)m * k * 4))) return rc;
    if ((rc = ensure(Qp, &c->Q_cap, (size_t)n * k * 4))) return rc;
    c->m = m;
    c->n = n;
    c->k = k;
    c->prec = CUMF_FP32;
    if (m && (rc = copy_in(c->P, P, (size_t)m * k * 4, c->stream))) return rc;
    if (n && (rc = copy_in(c->Q, Q, (size_t)n * k * 4, c->stream))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->have_factors = true;
    return CUMF_OK;
}

static int factor_ref(cumf_ctx* c, int32_t which, void** F, float** S, int64_t* rows)
{
    if (!c->have_factors) return fail(CUMF_ERR_STATE, "factors not initialised");
    if (which != 0 && which != 1) return fail(CUMF_ERR_ARG, "which must be 0 (P) or 1 (Q)");
    *F = which ? c->Q : c->P;
    *S = which ? c->qs : c->ps;
    *rows = which ? c->n : c->m;
    return CUMF_OK;
}

int cumf_set_factors(cumf_ctx* c, int32_t which, const void* data, const float* scales)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    void* F; float* S; int64_t rows;
    if ((rc = factor_ref(c, which, &F, &S, &rows))) return rc;
    if (!data) return fail(CUMF_ERR_ARG, "null data");
    if ((rc = copy_in(F, data, (size_t)rows * c->k * prec_bytes(c->prec), c->stream))) return rc;
    if (c->prec == CUMF_FP8_E4M3) {
        if (!scales) return fail(CUMF_ERR_ARG, "FP8 factors need row scales");
        if ((rc = copy_in(S, scales, sizeof(float) * rows, c->stream))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CUMF_OK;
}

int cumf_get_factors(cumf_ctx* c, int32_t which, void* data, float* scales)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    void* F; float* S; int64_t rows;
    if ((rc = factor_ref(c, which, &F, &S, &rows))) return rc;
    if (data) HIP_TRY(hipMemcpyAsync(data, F, (size_t)rows * c->k * prec_bytes(c->prec), hipMemcpyDefault, c->stream));
    if (scales && c->prec == CUMF_FP8_E4M3) HIP_TRY(hipMemcpyAsync(scales, S, sizeof(float) * rows, hipMemcpyDefault, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CUMF_OK;
}

int cumf_export_factors(cumf_ctx* c, int32_t which, float mult, const int32_t* map, float* out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    void* F; float* S; int64_t rows;
    if ((rc = factor_ref(c, which, &F, &S, &rows))) return rc;
    if (!out) return fail(CUMF_ERR_ARG, "null output");
    DevTmp dmap, dout;
    const int32_t* mp = nullptr;
    if (map) {
        std::vector<int32_t> hm(rows);
        HIP_TRY(hipMemcpy(hm.data(), map, sizeof(int32_t) * rows, hipMemcpyDefault));
        for (int32_t x : hm)
            if (x < 0 || x >= rows) return fail(CUMF_ERR_ARG, "map entry %d outside [0,%lld)", x, (long long)rows);
        HIP_TRY(hipMalloc(&dmap.p, sizeof(int32_t) * rows));
        HIP_TRY(hipMemcpyAsync(dmap.p, hm.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice, c->stream));
        mp = (const int32_t*)dmap.p;
    }
    HIP_TRY(hipMalloc(&dout.p, sizeof(float) * rows * c->k));
    HIP_TRY(launch_export(F, S, rows, c->k, c->prec, mult, mp, (float*)dout.p, c->stream));
    HIP_TRY(hipMemcpyAsync(out, dout.p, sizeof(float) * rows * c->k, hipMemcpyDefault, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CUMF_OK;
}

int cumf_test_eval(cumf_ctx* c, int32_t criterion, float mult, float b, double* value)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!eval_supported(criterion))
        return fail(CUMF_ERR_ARG, "criterion %d not supported (0, 1, 2, 5, 6, 10-13)", criterion);
    if (!c->have_factors) return fail(CUMF_ERR_STATE, "factors not initialised");
    if (!c->have_test) return fail(CUMF_ERR_STATE, "no test ratings");
    if (!value) return fail(CUMF_ERR_ARG, "null output");
    if (criterion >= 10) {   // row/column MPR (10, 11) and AUC (12, 13)
        double mpr = 0, auc = 0;
        if ((rc = rank_eval(c->stream, c->tu, c->tv, c->tr, c->tnnz, c->P, c->ps, c->Q, c->qs, c->m, c->n, c->k,
                            c->prec, mult, b, criterion == 11 || criterion == 13, &mpr, &auc)))
            return rc;
        *value = criterion <= 11 ? mpr : auc;
        return CUMF_OK;
    }
    if (c->tnnz == 0) {
        *value = 0.0;   // test/mf.cpp:3818-3819
        return CUMF_OK;
    }
    const size_t need = (size_t)rmse_blocks(c->tnnz) + 1;
    if ((rc = ensure(&c->partials, &c->partials_cap, need))) return rc;
    double* out = c->partials + need - 1;
    HIP_TRY(launch_eval(criterion, c->tu, c->tv, c->tr, c->tnnz, c->P, c->ps, c->Q, c->qs, c->m, c->n, c->k, c->prec,
                        mult, b, c->partials, out, c->stream));
    double h = 0.0;
    HIP_TRY(hipMemcpyAsync(&h, out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *value = criterion == 0 ? std::sqrt(h / (double)c->tnnz) : h / (double)c->tnnz;
    return CUMF_OK;
}

int cumf_test_rmse(cumf_ctx* c, float mult, float b, double* rmse)
{
    return cumf_test_eval(c, 0, mult, b, rmse);
}

int cumf_kernel_time(cumf_ctx* c, double* ms, int64_t* launches, int64_t* updates)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = harvest_events(c))) return rc;
    if (ms) *ms = c->kernel_ms;
    if (launches) *launches = c->launches;
    if (updates) *updates = c->updates;
    return CUMF_OK;
}

int cumf_reset_kernel_time(cumf_ctx* c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if ((rc = harvest_events(c))) return rc;
    c->kernel_ms = 0.0;
    c->launches = 0;
    c->updates = 0;
    return CUMF_OK;
}

int cumf_last_launch(cumf_ctx* c, cumf_launch_info* info)
{
    if (!c || !info) return fail(CUMF_ERR_ARG, "null argument");
    *info = c->last;
    return CUMF_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Training.
// ---------------------------------------------------------------------------
namespace {

// (Shape and ShapeInput: train_plan.h)

}  // namespace

// waves_per_cu waves per CU (the default 128, 512 for FP8, measured
// fastest: short per-worker chains, little tail), then the
// fewest workers that keep the same windows per worker, so an epoch
// over-samples the ratings as little as possible
constexpr int kDefaultWavesPerCu = 128;
constexpr int kDefaultWavesPerCuFp8 = 512;   // FP8 and stratum launches: more, shorter workers

int cumf::default_workers(const ShapeInput& in, int S, int L, int waves_per_cu)
{
    const int64_t w0 = (int64_t)in.cu_count * waves_per_cu * S;
    const int64_t nnz = std::max<int64_t>(1, in.nnz);
    const int64_t uc = std::max<int64_t>(1, (nnz + w0 * L - 1) / (w0 * L));
    int64_t w = (nnz + uc * L - 1) / (uc * L);
    const int64_t gran = (int64_t)kSgdWavesPerBlock * S;
    w = std::max<int64_t>(gran, (w + gran - 1) / gran * gran);
    return (int)std::min<int64_t>(w, 1 << 24);
}

int cumf::resolve_shape(const ShapeInput& in, const cumf_train_opts* o, Shape* sh, int xcd_multiple)
{
    sh->G = o->lanes_per_sample ? o->lanes_per_sample
                                : (o->schedule == CUMF_SCHED_TILE ? tile_default_group(in.k, in.prec)
                                                                  : sgd_default_group(in.k, in.prec));
    if (!sgd_shape_supported(in.k, in.prec, sh->G))
        return fail(CUMF_ERR_ARG, "lanes_per_sample=%d unsupported for k=%d precision=%d", sh->G, in.k, in.prec);
    if (o->coherent && (in.k / sh->G) * prec_bytes(in.prec) < 4)
        return fail(CUMF_ERR_ARG, "coherent mode needs >= 4 bytes per lane");
    sh->S = 64 / sh->G;
    sh->L = o->batch_len ? o->batch_len : 128;
    if (sh->L < 1 || sh->L > kSgdMaxBatch) return fail(CUMF_ERR_ARG, "batch_len=%d out of range 1..%d", sh->L, kSgdMaxBatch);
    if (sh->L % 2) return fail(CUMF_ERR_ARG, "batch_len=%d must be even", sh->L);
    // two row pairs in flight per worker measured fastest or equal for every
    // format (fewer registers, three waves per SIMD)
    int D = o->pipeline_depth ? o->pipeline_depth : 2;
    if (D != 2 && D != 4 && D != 8) return fail(CUMF_ERR_ARG, "pipeline_depth=%d not 2, 4 or 8", D);
    while (sh->L % D) D >>= 1;   // the depth must divide the window length
    sh->D = D;
    int W;
    if (o->schedule == CUMF_SCHED_SERIAL) {
        W = 1;
    } else if (o->num_workers > 0) {
        W = o->num_workers;
    } else {
        // stratum launches are short: one window per worker (the per-block
        // cap) measured fastest; FP8 likewise
        const bool one_window = o->schedule == CUMF_SCHED_STRATUM || in.prec == CUMF_FP8_E4M3;
        W = default_workers(in, sh->S, sh->L, one_window ? kDefaultWavesPerCuFp8 : kDefaultWavesPerCu);
    }
    const int per_block = kSgdWavesPerBlock * sh->S;
    int blocks = (W + per_block - 1) / per_block;
    if (xcd_multiple > 1) blocks = (blocks + xcd_multiple - 1) / xcd_multiple * xcd_multiple;
    sh->W = W;
    sh->blocks = blocks;
    return CUMF_OK;
}

int cumf::tile_default_group(int k, int prec)
{
    // 32 bytes of a row per lane where that is a supported shape: a tile
    // launch is bound by its VALU issue, and the per-sample work outside the
    // arithmetic (records, compares, addresses) then serves twice the
    // elements (99M workload: FP32 16.1e9 -> 19.0e9 updates/s, FP16
    // 20.8e9 -> 21.4e9)
    const int g0 = sgd_default_group(k, prec);
    return sgd_shape_supported(k, prec, g0 / 2) ? g0 / 2 : g0;
}

bool cumf::confine_to_one_xcd(int schedule, int coherent, int64_t m, int64_t n, int k, int prec)
{
    if (schedule != CUMF_SCHED_HOGWILD && schedule != CUMF_SCHED_GRID) return false;
    if (coherent) return false;   // coherent accesses already see one copy of every row
    return (m + n) * k * prec_bytes(prec) <= kOneXcdFactorBytes;
}

void cumf::lr_table(const cumf_train_opts* o, int n, float* out)
{
    // singleGPU/sgd_kernel.cu:137-154: lr[i] = float(alpha/(1+beta*i^1.5) + lrate)
    const float alpha = o->alpha, beta = o->beta, lrate = o->lrate;
    for (int i = 0; i < n; ++i) {
        double t = alpha / (1 + beta * std::pow(i, 1.5)) + lrate;
        out[i] = (float)t;
    }
}

namespace {

// A launch confined to one XCD sizes its default worker count for that XCD's
// CUs.
ShapeInput shape_input(const cumf_ctx* c, const cumf_train_opts* o)
{
    const bool one = confine_to_one_xcd(o->schedule, o->coherent, c->m, c->n, c->k, c->prec);
    return ShapeInput{c->k, c->prec, one ? std::max(1, c->cu_count / kNumXcd) : c->cu_count, c->nnz};
}

// The table is written only when the call needs more epochs than it holds or
// other rates.  Its address is part of the plan key and of every captured
// task, and growing it by one epoch per call would free device memory (a
// device-wide synchronisation), re-plan, re-upload the tasks and re-capture
// the epoch graph on every call of a one-epoch-per-call loop.
int upload_lr(cumf_ctx* c, const cumf_train_opts* o, int n)
{
    const float key[3] = {o->alpha, o->beta, o->lrate};
    if (c->lr && c->lr_valid >= (size_t)n && std::memcmp(key, c->lr_key, sizeof(key)) == 0) return CUMF_OK;
    const size_t want = ((size_t)n + kLrChunk - 1) / kLrChunk * kLrChunk;
    std::vector<float> h(want);
    lr_table(o, (int)want, h.data());
    c->lr_valid = 0;
    int rc = ensure(&c->lr, &c->lr_cap, want);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->lr, h.data(), sizeof(float) * want, hipMemcpyHostToDevice, c->stream));
    c->lr_valid = want;
    std::memcpy(c->lr_key, key, sizeof(key));
    return CUMF_OK;
}

SgdTask base_task(cumf_ctx* c, const cumf_train_opts* o, int64_t beg, int64_t end, int W, int L, uint32_t tag)
{
    SgdTask t{};
    t.u = c->u + beg;
    t.v = c->v + beg;
    t.r = c->r + beg;
    t.nnz = end - beg;
    t.P = c->P;
    t.Q = c->Q;
    t.ps = c->ps;
    t.qs = c->qs;
    t.lr = c->lr;
    t.lambda_p = o->lambda_p;
    t.lambda_q = o->lambda_q;
    t.worker_begin = 0;
    t.worker_count = W;
    // update_count = ceil(size / (W * L)) (singleGPU/sgd_kernel.cu:203)
    t.update_count = (int32_t)((t.nnz + (int64_t)W * L - 1) / ((int64_t)W * L));
    t.batch_len = L;
    t.problem = 0;
    t.tag = tag;
    return t;
}

// Window plan of one tile of n records: window length len (a multiple of D,
// at most Lmax) and windows = ceil(n / len) windows per epoch, pass r of R
// running windows [windows r / R, windows (r + 1) / R) on w_r workers of u_r
// windows each (u_r the fewest windows per worker that W workers allow, w_r
// the fewest workers that keep them).  The tile's workgroup issues whole
// waves of S slots, so a tile of, say, 26 workers of 8-slot waves would run
// 6 idle slots; len is the length whose passes issue the fewest slot-steps,
// a window costing kTileWindowSteps steps beyond its records, ties going to
// fewer samples, then to the longer window.
struct TileWindows {
    int len;
    int64_t windows;
};

TileWindows tile_windows(int64_t n, int W, int S, int Lmax, int D, int R)
{
    TileWindows best{Lmax, (n + Lmax - 1) / Lmax};
    int64_t best_cost = -1, best_samples = 0;
    for (int L = Lmax; L >= D; L -= D) {
        const int64_t N = (n + L - 1) / L;
        int64_t cost = 0, samples = 0;
        for (int r = 0; r < R; ++r) {
            const int64_t nr = N * (r + 1) / R - N * r / R;
            if (nr == 0) continue;
            const int64_t u = (nr + W - 1) / W, w = (nr + u - 1) / u;
            cost += (w + S - 1) / S * S * u * (L + kTileWindowSteps);
            samples += w * u * L;
        }
        if (best_cost < 0 || cost < best_cost || (cost == best_cost && samples < best_samples)) {
            best = TileWindows{L, N};
            best_cost = cost;
            best_samples = samples;
        }
    }
    return best;
}

// Waves of a tile workgroup holding `workers` workers of S per wave: 4, 8
// or 16.
int tile_waves(int workers, int S)
{
    const int w = (workers + S - 1) / S;
    return w <= 4 ? 4 : (w <= 8 ? 8 : 16);
}

// The launches of one epoch, appended to `tasks`.
int plan_epoch(cumf_ctx* c, const cumf_train_opts* o, const Shape& sh, int num_xcd, int ep, int e0,
               const int32_t* block_order, int64_t order_len, std::vector<SgdTask>& tasks,
               std::vector<LaunchRec>& launches, int64_t* upd_epoch)
{
    // a launch confined to XCD 0: the grid spans all XCDs (workgroups go
    // round-robin over them), workers are numbered XCD-major, and the
    // workgroups of the other XCDs find no task and exit
    const bool one_xcd = confine_to_one_xcd(o->schedule, o->coherent, c->m, c->n, c->k, c->prec);
    const int spread = one_xcd ? kNumXcd : 1;
    const int32_t spread_blocks = one_xcd ? kNumXcd * sh.blocks : 0;
    const int nb = c->ux * c->vy;
    *upd_epoch = 0;
    auto push = [&](std::vector<SgdTask>& t, int nx, int32_t blocks = 0) {
        if (t.empty()) return;
        launches.push_back(LaunchRec{(int64_t)tasks.size(), (int32_t)t.size(), nx, blocks});
        for (const SgdTask& x : t) *upd_epoch += (int64_t)x.worker_count * x.update_count * x.batch_len;
        tasks.insert(tasks.end(), t.begin(), t.end());
    };
    if (o->schedule == CUMF_SCHED_HOGWILD || o->schedule == CUMF_SCHED_SERIAL) {
        std::vector<SgdTask> t{base_task(c, o, 0, c->nnz, sh.W, sh.L, 0)};
        push(t, spread, spread_blocks);
    } else if (o->schedule == CUMF_SCHED_GRID) {
        for (int i = 0; i < nb; ++i) {
            int blk = i;
            if (block_order) blk = block_order[order_len == nb ? i : (int64_t)(ep - e0) * nb + i];
            const int64_t beg = c->offsets[blk], end = c->offsets[blk + 1];
            if (end == beg) continue;
            std::vector<SgdTask> t{base_task(c, o, beg, end, sh.W, sh.L, (uint32_t)blk)};
            push(t, spread, spread_blocks);
        }
    } else if (o->schedule == CUMF_SCHED_TILE) {
        // tiles: sub-epoch d trains the tiles of blocks (i, (i + d) mod s).
        // A block's tiles stay on one XCD (they share the block's Q rows
        // through its L2); the blocks of a diagonal go to XCDs largest first
        // onto the least-loaded one, and each XCD's tiles run largest first
        // (workgroups are dispatched to an XCD in order).  A tile gets one
        // workgroup of at most per_tile workers and its own window length
        // (tile_windows).
        const int s = c->ux;
        const int R = std::max(1, o->stratum_rounds);
        const int per_tile = sh.W;   // workers per tile (cumf_train resolved it)
        auto block_size = [&](int blk) { return c->offsets[(size_t)blk + 1] - c->offsets[(size_t)blk]; };
        struct TileEntry {
            SgdTask t;
            int64_t windows;   // the tile's windows per epoch
        };
        // diag[d][x]: the tiles XCD x trains in sub-epoch d, in launch order
        std::vector<std::vector<std::vector<TileEntry>>> diag((size_t)s,
                                                              std::vector<std::vector<TileEntry>>((size_t)num_xcd));
        for (int d = 0; d < s; ++d) {
            std::vector<int> order((size_t)s);
            for (int i = 0; i < s; ++i) order[(size_t)i] = i * s + (i + d) % s;
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return block_size(a) > block_size(b); });
            std::vector<int64_t> load((size_t)num_xcd, 0);
            for (int blk : order) {
                if (block_size(blk) == 0) continue;
                const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                load[(size_t)x] += block_size(blk);
                for (int32_t t = c->block_tile[(size_t)blk]; t < c->block_tile[(size_t)blk + 1]; ++t) {
                    const int64_t beg = c->tile_rec[(size_t)t], end = c->tile_rec[(size_t)t + 1];
                    const TileWindows tw = tile_windows(end - beg, per_tile, sh.S, sh.L, sh.D, R);
                    SgdTask tk = base_task(c, o, beg, end, 1, tw.len, (uint32_t)t);
                    tk.row_begin = c->tile_row0[(size_t)t];
                    tk.row_count = c->tile_rows[(size_t)t];
                    diag[(size_t)d][(size_t)x].push_back(TileEntry{tk, tw.windows});
                }
            }
            for (auto& l : diag[(size_t)d])
                std::stable_sort(l.begin(), l.end(), [](const TileEntry& a, const TileEntry& b) { return a.t.nnz > b.t.nnz; });
        }
        // A launch holds up to F tiles per XCD (F = tiles_in_flight, 0: all):
        // task x * per + j is XCD x's j-th tile, the rest padding (nnz 0).
        auto emit = [&](const std::vector<std::vector<SgdTask>>& lists) {
            const int F = o->tiles_in_flight;
            size_t most = 0;
            for (const auto& l : lists) most = std::max(most, l.size());
            for (size_t c0 = 0; c0 < most; c0 += F > 0 ? (size_t)F : most) {
                const size_t c1 = F > 0 ? std::min(most, c0 + (size_t)F) : most;
                size_t per = 0;
                for (const auto& l : lists) per = std::max(per, std::min(l.size(), c1) - std::min(l.size(), c0));
                if (per == 0) continue;
                std::vector<SgdTask> t((size_t)num_xcd * per, SgdTask{});
                for (int x = 0; x < num_xcd; ++x)
                    for (size_t j = c0; j < std::min(lists[(size_t)x].size(), c1); ++j)
                        t[(size_t)x * per + (j - c0)] = lists[(size_t)x][j];
                push(t, num_xcd, (int32_t)(num_xcd * per));
            }
        };
        // pass rr runs windows [N*rr/R, N*(rr+1)/R) of a tile's N: the fewest
        // windows per worker that per_tile workers allow, then the fewest
        // workers that keep them; a worker's window indices continue from
        // pass to pass
        std::vector<std::vector<std::vector<int32_t>>> next_window((size_t)s);
        for (int d = 0; d < s; ++d) {
            next_window[(size_t)d].resize((size_t)num_xcd);
            for (int x = 0; x < num_xcd; ++x) next_window[(size_t)d][(size_t)x].assign(diag[(size_t)d][(size_t)x].size(), 0);
        }
        for (int rr = 0; rr < R; ++rr)
            for (int d = 0; d < s; ++d) {
                std::vector<std::vector<SgdTask>> lists((size_t)num_xcd);
                for (int x = 0; x < num_xcd; ++x)
                    for (size_t j = 0; j < diag[(size_t)d][(size_t)x].size(); ++j) {
                        const TileEntry& te = diag[(size_t)d][(size_t)x][j];
                        const int64_t n_pass = te.windows * (rr + 1) / R - te.windows * rr / R;
                        if (n_pass == 0) continue;
                        SgdTask q = te.t;
                        q.update_count = (int32_t)((n_pass + per_tile - 1) / per_tile);
                        q.worker_count = (int32_t)((n_pass + q.update_count - 1) / q.update_count);
                        q.window_begin = next_window[(size_t)d][(size_t)x][j];
                        next_window[(size_t)d][(size_t)x][j] += q.update_count;
                        lists[(size_t)x].push_back(q);
                    }
                emit(lists);
            }
    } else {
        // stratum: sub-epoch d trains blocks (i, (i + d) mod s), i = 0..s-1,
        // which share no P and no Q rows.  Each block runs on one XCD; the
        // blocks of a diagonal go to XCDs largest first onto the least-loaded
        // one (a launch lasts as long as its most loaded XCD).  With R rounds
        // the epoch passes over the s diagonals R times, each pass running
        // 1/R of every block's windows (window indices continue across the
        // passes), so that a row meets every column range R times per epoch.
        const int s = c->ux;
        const int R = std::max(1, o->stratum_rounds);
        const int per_xcd = sh.W / num_xcd;
        struct Placed {
            SgdTask t;
            int x;   // XCD
        };
        std::vector<std::vector<Placed>> diag((size_t)s);
        for (int d = 0; d < s; ++d) {
            std::vector<int> order((size_t)s);
            for (int i = 0; i < s; ++i) order[(size_t)i] = i * s + (i + d) % s;
            auto size_of = [&](int blk) { return c->offsets[(size_t)blk + 1] - c->offsets[(size_t)blk]; };
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return size_of(a) > size_of(b); });
            std::vector<std::vector<int>> assigned((size_t)num_xcd);
            std::vector<int64_t> load((size_t)num_xcd, 0);
            for (int blk : order) {
                const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                assigned[(size_t)x].push_back(blk);
                load[(size_t)x] += size_of(blk);
            }
            for (int x = 0; x < num_xcd; ++x) {
                std::vector<int> mine;   // the XCD's non-empty blocks, largest first
                for (int blk : assigned[(size_t)x])
                    if (size_of(blk) > 0) mine.push_back(blk);
                const int64_t tot = load[(size_t)x];
                // every non-empty block gets at least one worker, also when
                // the XCD holds more blocks than its share of the workers
                const int budget = std::max(per_xcd, (int)mine.size());
                int left = budget;
                for (size_t j = 0; j < mine.size(); ++j) {
                    const int blk = mine[j];
                    const int64_t sz = size_of(blk);
                    int w = j + 1 == mine.size() ? left
                                                  : (int)std::max<int64_t>(1, tot ? (int64_t)budget * sz / tot : 1);
                    // leave one worker for each block still to place
                    w = std::max(1, std::min(w, left - (int)(mine.size() - 1 - j)));
                    // no more workers than windows: a small block gets one
                    // window per worker instead of many redundant ones
                    w = (int)std::min<int64_t>(w, (sz + sh.L - 1) / sh.L);
                    left -= w;
                    diag[(size_t)d].push_back(
                        Placed{base_task(c, o, c->offsets[blk], c->offsets[blk + 1], w, sh.L, (uint32_t)blk), x});
                }
            }
            // each non-empty block of the diagonal is trained exactly once
            size_t nonempty = 0;
            for (int blk : order) nonempty += size_of(blk) > 0;
            if (diag[(size_t)d].size() != nonempty)
                return fail(CUMF_ERR_STATE, "stratum plan placed %zu of %zu blocks of diagonal %d",
                            diag[(size_t)d].size(), nonempty, d);
        }
        // A launch's grid holds just its workers: XCD x numbers its workers
        // [x * P, x * P + its count), P the largest count rounded up to whole
        // workgroups (the kernel numbers workers XCD-major over the grid).
        const int gran = kSgdWavesPerBlock * sh.S;
        auto emit = [&](const std::vector<Placed>& pl) {
            std::vector<int64_t> off((size_t)num_xcd, 0);
            for (const Placed& q : pl) off[(size_t)q.x] += q.t.worker_count;
            const int64_t most = std::max<int64_t>(1, *std::max_element(off.begin(), off.end()));
            const int64_t P = (most + gran - 1) / gran * gran;
            std::fill(off.begin(), off.end(), 0);
            std::vector<SgdTask> t;
            for (const Placed& q : pl) {
                SgdTask tk = q.t;
                tk.worker_begin = (int32_t)(q.x * P + off[(size_t)q.x]);
                off[(size_t)q.x] += tk.worker_count;
                t.push_back(tk);
            }
            push(t, num_xcd, (int32_t)(num_xcd * (P / gran)));
        };
        // pass rr runs N*(rr+1)/R - N*rr/R (rounded up to whole windows per
        // worker) of a block's N windows on at most its workers; a worker's
        // window indices continue from pass to pass
        std::vector<std::vector<int32_t>> next_window((size_t)s);
        for (int d = 0; d < s; ++d) next_window[(size_t)d].assign(diag[(size_t)d].size(), 0);
        for (int rr = 0; rr < R; ++rr)
            for (int d = 0; d < s; ++d) {
                std::vector<Placed> pl;
                for (size_t j = 0; j < diag[(size_t)d].size(); ++j) {
                    const Placed& full = diag[(size_t)d][j];
                    Placed q = full;
                    if (R > 1) {
                        const int64_t n_all = (int64_t)full.t.worker_count * full.t.update_count;
                        const int64_t n_pass = n_all * (rr + 1) / R - n_all * rr / R;
                        if (n_pass == 0) continue;
                        // fewest windows per worker, then fewest workers:
                        // w * uc - n_pass < uc
                        q.t.update_count = (int32_t)((n_pass + full.t.worker_count - 1) / full.t.worker_count);
                        q.t.worker_count = (int32_t)((n_pass + q.t.update_count - 1) / q.t.update_count);
                        q.t.window_begin = next_window[(size_t)d][j];
                        next_window[(size_t)d][j] += q.t.update_count;
                    }
                    pl.push_back(q);
                }
                if (!pl.empty()) emit(pl);
            }
    }
    return CUMF_OK;
}

// Queue one epoch's launches (and the FP8 rescale) on the context stream.
// epoch_dev non-null: the kernels read the epoch from device memory and the
// counter is advanced at the end (the form captured into a graph).
int queue_one(cumf_ctx* c, const cumf_train_opts* o, const Shape& sh, const LaunchRec& l, int ep, int32_t* epoch_dev,
              bool timed)
{
    {
        SgdLaunch a{};
        // factors under 4 GB each: rows addressed as launch-wide base + 32-bit offset
        const uint64_t eb = (uint64_t)prec_bytes(c->prec);
        if ((uint64_t)c->m * c->k * eb <= kBufRecords && (uint64_t)c->n * c->k * eb <= kBufRecords) {
            a.P = c->P;
            a.Q = c->Q;
            a.ps = c->ps;
            a.qs = c->qs;
        }
        a.tasks = c->tasks + l.task_off;
        a.num_tasks = l.task_cnt;
        a.epoch_begin = ep;
        a.epoch_end = ep + 1;
        a.batch_len = sh.L;
        a.seed = o->seed;
        a.num_xcd = l.num_xcd;
        a.epoch_dev = epoch_dev;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        if (timed) {
            ev0 = next_event(c);
            ev1 = next_event(c);
            if (!ev0 || !ev1) return fail(CUMF_ERR_HIP, "hipEventCreate failed");
            HIP_TRY(hipEventRecord(ev0, c->stream));
        }
        if (o->schedule == CUMF_SCHED_TILE)
            HIP_TRY(launch_sgd_tile(a, l.blocks, c->k, c->prec, sh.G, sh.D, c->tile_rows_max, tile_waves(sh.W, sh.S),
                                    c->stream));
        else
            HIP_TRY(launch_sgd(a, l.blocks ? l.blocks : sh.blocks, c->k, c->prec, sh.G, o->coherent != 0, sh.D,
                               c->stream));
        if (timed) HIP_TRY(hipEventRecord(ev1, c->stream));
    }
    return CUMF_OK;
}

// After an epoch's launches: the FP8 rescale and, for graph replays, the
// device epoch counter.
int queue_epoch_tail(cumf_ctx* c, int32_t* epoch_dev)
{
    if (c->prec == CUMF_FP8_E4M3) {
        HIP_TRY(launch_fp8_rescale((uint8_t*)c->P, c->ps, c->m, c->k, c->stream));
        HIP_TRY(launch_fp8_rescale((uint8_t*)c->Q, c->qs, c->n, c->k, c->stream));
    }
    if (epoch_dev) HIP_TRY(launch_advance_int(epoch_dev, c->stream));
    return CUMF_OK;
}

int queue_epoch(cumf_ctx* c, const cumf_train_opts* o, const Shape& sh, const std::vector<LaunchRec>& ls, int ep,
                int32_t* epoch_dev, bool timed)
{
    int rc;
    for (const LaunchRec& l : ls)
        if ((rc = queue_one(c, o, sh, l, ep, epoch_dev, timed))) return rc;
    return queue_epoch_tail(c, epoch_dev);
}

// Streamed ratings: each stratum diagonal is copied from pinned host memory
// into staging buffer (d & 1) on the copy stream while the previous diagonal
// trains; a diagonal's launch waits for its copy, a copy waits for the
// launch that last read its buffer.  soff[blk] is the block's place in its
// diagonal's staging buffer; the tasks already point there.
int train_streamed(cumf_ctx* c, const cumf_train_opts* o, const Shape& sh, const std::vector<LaunchRec>& ls,
                   const std::vector<SgdTask>& tasks, const std::vector<int64_t>& soff, int e0, int e1)
{
    const int s = c->ux;
    auto diag_of = [&](const LaunchRec& l) {
        const int blk = (int)tasks[(size_t)l.task_off].tag;
        return (blk % s - blk / s + s) % s;
    };
    auto copy_diag = [&](int d) -> int {
        const int b = d & 1;
        HIP_TRY(hipStreamWaitEvent(c->copy_stream, c->kern_done[b], 0));
        for (int i = 0; i < s; ++i) {
            const int blk = i * s + (i + d) % s;
            const int64_t beg = c->offsets[(size_t)blk], sz = c->offsets[(size_t)blk + 1] - beg;
            if (sz == 0) continue;
            const int64_t at = soff[(size_t)blk];
            HIP_TRY(hipMemcpyAsync(c->su[b] + at, c->hu + beg, sizeof(int32_t) * sz, hipMemcpyHostToDevice, c->copy_stream));
            HIP_TRY(hipMemcpyAsync(c->sv[b] + at, c->hv + beg, sizeof(int32_t) * sz, hipMemcpyHostToDevice, c->copy_stream));
            HIP_TRY(hipMemcpyAsync(c->sr[b] + at, c->hr + beg, sizeof(float) * sz, hipMemcpyHostToDevice, c->copy_stream));
        }
        HIP_TRY(hipEventRecord(c->copy_done[b], c->copy_stream));
        return CUMF_OK;
    };
    int rc;
    // both buffers are free at the start
    HIP_TRY(hipEventRecord(c->kern_done[0], c->stream));
    HIP_TRY(hipEventRecord(c->kern_done[1], c->stream));
    if (ls.empty()) return CUMF_OK;
    if ((rc = copy_diag(diag_of(ls[0])))) return rc;
    for (int ep = e0; ep < e1; ++ep) {
        for (size_t li = 0; li < ls.size(); ++li) {
            const int b = diag_of(ls[li]) & 1;
            HIP_TRY(hipStreamWaitEvent(c->stream, c->copy_done[b], 0));
            if ((rc = queue_one(c, o, sh, ls[li], ep, nullptr, true))) return rc;
            HIP_TRY(hipEventRecord(c->kern_done[b], c->stream));
            if (li + 1 < ls.size()) {
                if ((rc = copy_diag(diag_of(ls[li + 1])))) return rc;
            } else if (ep + 1 < e1) {
                if ((rc = copy_diag(diag_of(ls[0])))) return rc;
            }
            const LaunchRec& l = ls[li];
            for (int32_t t = 0; t < l.task_cnt; ++t) {
                const SgdTask& x = tasks[(size_t)(l.task_off + t)];
                c->updates += (int64_t)x.worker_count * x.update_count * x.batch_len;
            }
        }
        if ((rc = queue_epoch_tail(c, nullptr))) return rc;
        c->launches += (int64_t)ls.size();
    }
    HIP_TRY(hipStreamSynchronize(c->copy_stream));
    return CUMF_OK;
}

}  // namespace

extern "C" int cumf_train(cumf_ctx* c, const cumf_train_opts* o, int32_t e0, int32_t e1, const int32_t* block_order,
                          int64_t order_len)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!o) return fail(CUMF_ERR_ARG, "null options");
    if (!c->have_ratings) return fail(CUMF_ERR_STATE, "no ratings");
    if (!c->have_factors) return fail(CUMF_ERR_STATE, "factors not initialised");
    if (o->k != c->k) return fail(CUMF_ERR_ARG, "opts k=%d but factors have k=%d", o->k, c->k);
    if (e0 < 0 || e1 < e0 || e1 > 1 << 20) return fail(CUMF_ERR_ARG, "bad epoch range [%d,%d)", e0, e1);
    if (o->schedule < CUMF_SCHED_HOGWILD || o->schedule > CUMF_SCHED_TILE) return fail(CUMF_ERR_ARG, "bad schedule");
    if (o->tiles_in_flight < 0) return fail(CUMF_ERR_ARG, "tiles_in_flight=%d must be >= 0", o->tiles_in_flight);
    if (o->use_graph != 0 && o->use_graph != 1) return fail(CUMF_ERR_ARG, "use_graph must be 0 or 1");
    if (o->stratum_rounds < 0 || o->stratum_rounds > 64)
        return fail(CUMF_ERR_ARG, "stratum_rounds=%d out of range 0..64", o->stratum_rounds);
    if (c->nnz == 0 || e1 == e0) return CUMF_OK;

    const int nb = c->ux * c->vy;
    const bool stratum = o->schedule == CUMF_SCHED_STRATUM;
    const bool tiled = o->schedule == CUMF_SCHED_TILE;
    if (stratum && (c->ux != c->vy || c->ux < 2))
        return fail(CUMF_ERR_STATE, "stratum schedule needs an s x s partition (have %d x %d)", c->ux, c->vy);
    if (tiled && c->block_tile.empty())
        return fail(CUMF_ERR_STATE, "tile schedule needs the partition of cumf_partition_tiles");
    const int num_xcd = stratum || tiled ? 8 : 1;
    Shape sh;
    if ((rc = resolve_shape(shape_input(c, o), o, &sh, num_xcd))) return rc;
    if (stratum) sh.W = sh.blocks * kSgdWavesPerBlock * sh.S;   // every wave of every XCD works
    if (tiled) {
        // one workgroup per tile: 16 or 32 bytes of a row per lane, no
        // coherent mode, the tile and the staged records in LDS
        const int g0 = sgd_default_group(c->k, c->prec);
        if (sh.G != g0 && !(sh.G == g0 / 2 && sgd_shape_supported(c->k, c->prec, g0 / 2)))
            return fail(CUMF_ERR_ARG, "tile schedule runs lanes_per_sample=%d or %d (k=%d precision=%d)", g0, g0 / 2,
                        c->k, c->prec);
        if (o->coherent) return fail(CUMF_ERR_ARG, "tile schedule has no coherent mode");
        if (sh.D != 2) return fail(CUMF_ERR_ARG, "tile schedule runs pipeline_depth 2");
        const int most = kSgdTileMaxWaves * sh.S;
        if (o->num_workers > most)
            return fail(CUMF_ERR_ARG, "tile schedule: num_workers=%d per tile exceeds %d", o->num_workers, most);
        sh.W = o->num_workers > 0 ? o->num_workers : std::min(kTileWorkers, most);   // workers per tile
        // windows of at most kSgdTileMaxBatch records (the longest by default)
        if (o->batch_len == 0) sh.L = kSgdTileMaxBatch;
        if (sh.L > kSgdTileMaxBatch)
            return fail(CUMF_ERR_ARG, "tile schedule: batch_len=%d exceeds %d", sh.L, kSgdTileMaxBatch);
        const size_t lds = sgd_tile_lds_bytes(c->k, c->prec, sh.G, c->tile_rows_max, tile_waves(sh.W, sh.S));
        if (lds > sgd_tile_lds_limit())
            return fail(CUMF_ERR_ARG, "tiles of %d rows at k=%d precision=%d need %zu bytes of LDS with the staged "
                        "records (at most %zu): partition with fewer tile rows", c->tile_rows_max, c->k, c->prec, lds,
                        sgd_tile_lds_limit());
    }
    if ((rc = upload_lr(c, o, e1 + 4))) return rc;

    c->last.num_workers = sh.W;
    c->last.lanes_per_sample = sh.G;
    c->last.waves_per_block = kSgdWavesPerBlock;
    c->last.blocks = sh.blocks;
    c->last.batch_len = sh.L;
    c->last.pipeline_depth = sh.D;
    c->last.updates_per_epoch = 0;

    const bool per_epoch_order = o->schedule == CUMF_SCHED_GRID && block_order && order_len != nb;
    if (o->schedule == CUMF_SCHED_GRID && block_order) {
        if (order_len != nb && order_len != (int64_t)nb * (e1 - e0))
            return fail(CUMF_ERR_ARG, "block_order has %lld entries (need %d or %lld)", (long long)order_len, nb,
                        (long long)nb * (e1 - e0));
        for (int64_t i = 0; i < order_len; ++i)
            if (block_order[i] < 0 || block_order[i] >= nb) return fail(CUMF_ERR_ARG, "block_order entry out of range");
    }

    // Plan every epoch's launches and upload the task table once.  A call
    // whose plan key (the ratings' generation and addresses, the factor
    // storage, the learning-rate table's address, the options and the
    // resolved shape) equals the last upload's reuses that table as it is:
    // an epoch loop driven one epoch per call neither re-plans nor uploads.
    const bool cacheable = !per_epoch_order;
    uint64_t pkey = 1469598103934665603ull;
    {
        const uint64_t gen[2] = {c->data_gen, (uint64_t)c->nnz};
        pkey = fnv1a(gen, sizeof(gen), pkey);
        const int32_t f[20] = {o->schedule, o->num_workers, o->batch_len, o->lanes_per_sample, o->coherent,
                               o->pipeline_depth, o->stratum_rounds, o->tiles_in_flight, sh.G, sh.S, sh.L, sh.W,
                               sh.blocks, sh.D, num_xcd, c->k, c->prec, c->m, c->n, c->tile_rows_max};
        pkey = fnv1a(f, sizeof(f), pkey);
        const float lam[2] = {o->lambda_p, o->lambda_q};
        pkey = fnv1a(lam, sizeof(lam), pkey);
        const void* addr[14] = {c->u,     c->v,     c->r,     c->P,     c->Q,     c->ps,    c->qs,
                                c->lr,    c->su[0], c->su[1], c->sv[0], c->sv[1], c->sr[0], c->sr[1]};
        pkey = fnv1a(addr, sizeof(addr), pkey);
    }
    std::vector<SgdTask> fresh;
    std::vector<std::vector<LaunchRec>> plan;
    std::vector<int64_t> soff;
    int64_t upd = 0;
    const bool reuse = cacheable && c->plan_valid && c->plan_key == pkey;
    if (reuse) {
        plan.push_back(c->plan_launches);
        upd = c->plan_upd;
        soff = c->plan_soff;
    } else {
        c->plan_valid = false;
        const int plans = per_epoch_order ? e1 - e0 : 1;
        for (int i = 0; i < plans; ++i) {
            plan.emplace_back();
            if ((rc = plan_epoch(c, o, sh, num_xcd, e0 + i, e0, block_order, order_len, fresh, plan.back(), &upd)))
                return rc;
        }
        if (c->streamed && !fresh.empty()) {
            if (!stratum) return fail(CUMF_ERR_STATE, "streamed ratings train with the stratum schedule");
            if (o->stratum_rounds > 1)
                return fail(CUMF_ERR_ARG, "streamed ratings make one pass over the diagonals per epoch (stratum_rounds 1)");
            // point every task at its block's place in its diagonal's staging buffer
            const int s = c->ux;
            soff.assign((size_t)s * s, 0);
            for (int d = 0; d < s; ++d) {
                int64_t at = 0;
                for (int i = 0; i < s; ++i) {
                    const int blk = i * s + (i + d) % s;
                    soff[(size_t)blk] = at;
                    at += c->offsets[(size_t)blk + 1] - c->offsets[(size_t)blk];
                }
            }
            for (SgdTask& t : fresh) {
                const int blk = (int)t.tag;
                const int b = ((blk % s - blk / s + s) % s) & 1;
                t.u = c->su[b] + soff[(size_t)blk];
                t.v = c->sv[b] + soff[(size_t)blk];
                t.r = c->sr[b] + soff[(size_t)blk];
            }
        }
        if (!fresh.empty()) {
            if ((rc = ensure(&c->tasks, &c->tasks_cap, fresh.size()))) return rc;
            HIP_TRY(hipMemcpyAsync(c->tasks, fresh.data(), sizeof(SgdTask) * fresh.size(), hipMemcpyHostToDevice,
                                   c->stream));
        }
        if (cacheable) {
            c->plan_tasks = std::move(fresh);
            c->plan_launches = plan[0];
            c->plan_upd = upd;
            c->plan_soff = soff;
            c->plan_key = pkey;
            c->plan_valid = true;
        }
    }
    const std::vector<SgdTask>& tasks = cacheable ? c->plan_tasks : fresh;
    c->last.updates_per_epoch = upd;
    // stratum and tile launches carry their own grids: report the largest
    int32_t most_blocks = 0;
    for (const LaunchRec& l : plan[0]) most_blocks = std::max(most_blocks, l.blocks);
    if (most_blocks > 0) c->last.blocks = most_blocks;
    if (tasks.empty()) return CUMF_OK;
    int64_t launches_per_epoch = 0;
    for (const LaunchRec& l : plan[0]) (void)l, ++launches_per_epoch;

    if (c->streamed) {
        if ((rc = train_streamed(c, o, sh, plan[0], tasks, soff, e0, e1))) return rc;
    } else if (o->use_graph && !per_epoch_order) {
        // the epoch loop as a graph: captured once per plan, replayed per epoch.
        // The key covers everything the captured launches hold by value: the
        // plan key (which fixes the task table's contents, the launch shapes,
        // the factor format, length and addresses and the row counts the FP8
        // rescale walks), the seed, and the addresses of the task table (a
        // larger plan reallocates it) and of the epoch counter.
        if ((rc = ensure(&c->epoch_dev, &c->epoch_dev_cap, 1))) return rc;
        uint64_t key = fnv1a(&o->seed, sizeof(o->seed), pkey);
        const void* addr[2] = {c->tasks, c->epoch_dev};
        key = fnv1a(addr, sizeof(addr), key);
        if (!c->graph_exec || c->graph_key != key) {
            if (c->graph_exec) {
                (void)hipGraphExecDestroy(c->graph_exec);
                c->graph_exec = nullptr;
            }
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int qrc = queue_epoch(c, o, sh, plan[0], e0, c->epoch_dev, false);
            hipError_t ce = hipStreamEndCapture(c->stream, &g);
            if (qrc) {
                if (g) (void)hipGraphDestroy(g);
                return qrc;
            }
            if (ce != hipSuccess) return fail(CUMF_ERR_HIP, "graph capture failed: %s", hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                c->graph_exec = nullptr;
                return fail(CUMF_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
            }
            c->graph_key = key;
        }
        HIP_TRY(launch_set_int(c->epoch_dev, e0, c->stream));
        for (int ep = e0; ep < e1; ++ep) {
            hipEvent_t ev0 = next_event(c), ev1 = next_event(c);
            if (!ev0 || !ev1) return fail(CUMF_ERR_HIP, "hipEventCreate failed");
            HIP_TRY(hipEventRecord(ev0, c->stream));
            HIP_TRY(hipGraphLaunch(c->graph_exec, c->stream));
            HIP_TRY(hipEventRecord(ev1, c->stream));
            c->launches += launches_per_epoch;
            c->updates += upd;
        }
    } else {
        for (int ep = e0; ep < e1; ++ep) {
            const std::vector<LaunchRec>& ls = plan[per_epoch_order ? ep - e0 : 0];
            if ((rc = queue_epoch(c, o, sh, ls, ep, nullptr, true))) return rc;
            for (const LaunchRec& l : ls)
                for (int32_t t = 0; t < l.task_cnt; ++t) {
                    const SgdTask& x = tasks[(size_t)(l.task_off + t)];
                    c->updates += (int64_t)x.worker_count * x.update_count * x.batch_len;
                }
            c->launches += (int64_t)ls.size();
        }
    }
    return harvest_events(c);
}

// ---------------------------------------------------------------------------
// Autotuner: time one epoch of every candidate launch shape (lane group,
// pipeline depth, waves per CU) on the context's own problem, keep the
// fastest, and restore the factors the trials trained.
// ---------------------------------------------------------------------------
extern "C" int cumf_autotune(cumf_ctx* c, const cumf_train_opts* base, cumf_train_opts* best, double* best_ms)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!base || !best) return fail(CUMF_ERR_ARG, "null options");
    if (!c->have_factors) return fail(CUMF_ERR_STATE, "factors not initialised");
    if (base->schedule == CUMF_SCHED_SERIAL) {
        *best = *base;
        if (best_ms) *best_ms = 0;
        return CUMF_OK;
    }
    const size_t eb = (size_t)prec_bytes(c->prec);
    const size_t pb = (size_t)c->m * c->k * eb, qb = (size_t)c->n * c->k * eb;
    DevTmp sp, sq, sps, sqs;
    HIP_TRY(hipMalloc(&sp.p, pb));
    HIP_TRY(hipMalloc(&sq.p, qb));
    HIP_TRY(hipMemcpyAsync(sp.p, c->P, pb, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(sq.p, c->Q, qb, hipMemcpyDeviceToDevice, c->stream));
    if (c->prec == CUMF_FP8_E4M3) {
        HIP_TRY(hipMalloc(&sps.p, 4 * (size_t)c->m));
        HIP_TRY(hipMalloc(&sqs.p, 4 * (size_t)c->n));
        HIP_TRY(hipMemcpyAsync(sps.p, c->ps, 4 * (size_t)c->m, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(sqs.p, c->qs, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, c->stream));
    }
    // keep the caller's timing totals out of the trials
    if ((rc = harvest_events(c))) return rc;
    const double keep_ms = c->kernel_ms;
    const int64_t keep_l = c->launches, keep_u = c->updates;
    const cumf_launch_info keep_last = c->last;

    // a field the caller set (non-zero) stays fixed; the zero ones are searched
    std::vector<int> groups{64, 32, 16, 8}, depths{2, 4, 8}, waves_per_cu{32, 64, 128, 256, 512}, lens{128, 64, 32};
    const bool tiled = base->schedule == CUMF_SCHED_TILE;
    std::vector<int> tile_workers{0};   // tile schedule: waves of workers per tile
    if (tiled) {
        // tile launches: the default lane group, depth 2, workgroups of 4, 8
        // or 16 waves (workers per tile) unless the caller fixed them
        const int g0 = sgd_default_group(c->k, c->prec);
        groups = {g0};
        if (sgd_shape_supported(c->k, c->prec, g0 / 2)) groups.push_back(g0 / 2);
        depths = {2};
        lens = {kSgdTileMaxBatch};
        if (!base->num_workers) tile_workers = {4, 8, 16};
    }
    if (base->lanes_per_sample) groups = {base->lanes_per_sample};
    if (base->pipeline_depth) depths = {base->pipeline_depth};
    if (base->batch_len) lens = {base->batch_len};
    double best_t = 1e300;
    cumf_train_opts bo = *base;
    bool warm = false, any = false;
    for (int L : lens)
        for (int G : groups) {
            if (!sgd_shape_supported(c->k, c->prec, G)) continue;
            if (base->coherent && (c->k / G) * (int)eb < 4) continue;
            for (int D : depths) {
                if (L % D) continue;
                const size_t nw = base->num_workers ? 1 : (tiled ? tile_workers.size() : waves_per_cu.size());
                for (size_t wi = 0; wi < nw; ++wi) {
                    cumf_train_opts o = *base;
                    o.lanes_per_sample = G;
                    o.pipeline_depth = D;
                    o.batch_len = L;
                    if (base->num_workers) o.num_workers = base->num_workers;
                    else if (tiled) o.num_workers = tile_workers[wi] * (64 / G);
                    else o.num_workers = default_workers(shape_input(c, base), 64 / G, L, waves_per_cu[wi]);
                    if (tiled && sgd_tile_lds_bytes(c->k, c->prec, G, c->tile_rows_max,
                                                    tile_waves(o.num_workers, 64 / G)) > sgd_tile_lds_limit())
                        continue;
                    o.use_graph = 0;
                    if (!warm) {   // first launch of the code object is not timed
                        if ((rc = cumf_train(c, &o, 0, 1, nullptr, 0))) return rc;
                        warm = true;
                    }
                    c->kernel_ms = 0;
                    if ((rc = cumf_train(c, &o, 0, 1, nullptr, 0))) return rc;
                    // the epoch's time, over-sampling included: that is what an
                    // epoch costs the caller
                    const double t = c->kernel_ms;
                    any = true;
                    if (t < best_t) {
                        best_t = t;
                        bo = o;
                    }
                }
            }
        }
    if (!any) return fail(CUMF_ERR_ARG, "no supported launch shape for k=%d precision=%d with the fixed fields", c->k, c->prec);
    HIP_TRY(hipMemcpyAsync(c->P, sp.p, pb, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->Q, sq.p, qb, hipMemcpyDeviceToDevice, c->stream));
    if (c->prec == CUMF_FP8_E4M3) {
        HIP_TRY(hipMemcpyAsync(c->ps, sps.p, 4 * (size_t)c->m, hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->qs, sqs.p, 4 * (size_t)c->n, hipMemcpyDeviceToDevice, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->kernel_ms = keep_ms;
    c->launches = keep_l;
    c->updates = keep_u;
    c->last = keep_last;
    bo.use_graph = base->use_graph;
    *best = bo;
    if (best_ms) *best_ms = best_t;
    return CUMF_OK;
}
