// cumf_train_noalloc: epochs of the Hogwild or serialized schedule over
// ratings and factors in the caller's device memory, without allocating.
// cumf_train keeps its launch's task record and learning-rate table in
// context memory it allocates; here they live in `scratch`, a fixed array of
// this library's device memory (one copy per GPU, held until the process
// ends), written through kernel arguments so that nothing on the host has to
// outlive the call.  The update and rescale kernels are cumf_train's own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/cumf_sgd_amd.h"
#include "errors.h"
#include "kernels.h"
#include "train_plan.h"

using namespace cumf;

namespace {

constexpr size_t kScratchBytes = 16384;
constexpr size_t kTaskBytes = 256;   // slot of the one task record, then the learning-rate table
constexpr int kMaxLr = (int)((kScratchBytes - kTaskBytes) / sizeof(float));
static_assert(sizeof(SgdTask) <= kTaskBytes, "the task record must fit its slot");
static_assert(kMaxLr == CUMF_NOALLOC_MAX_EPOCHS + 4, "the header's epoch limit must match the array");

// The device array.  In host code `scratch` names a host-side stand-in, not a
// device address: host code uses only the per-device address that
// hipGetSymbolAddress returns (g_scratch).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_scratch[kMaxDevices];

// Words staged through a kernel's arguments: the runtime copies them at the
// launch, so the host data may go out of scope right after.
constexpr int kPutWords = 256;
struct PutWords {
    uint32_t w[kPutWords];
    uint32_t* dst;
    int32_t count;
};

__global__ void put_words_kernel(PutWords a)
{
    for (int i = threadIdx.x; i < a.count; i += blockDim.x) a.dst[i] = a.w[i];
}

hipError_t put_words(uint32_t* dst, const void* src, size_t words, hipStream_t st)
{
    const uint32_t* s = (const uint32_t*)src;
    for (size_t at = 0; at < words; at += kPutWords) {
        PutWords a{};
        a.dst = dst + at;
        a.count = (int32_t)(words - at < (size_t)kPutWords ? words - at : (size_t)kPutWords);
        for (int i = 0; i < a.count; ++i) a.w[i] = s[at + (size_t)i];
        hipLaunchKernelGGL(put_words_kernel, dim3(1), dim3(256), 0, st, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// The device address of `scratch` on `device`, looked up once per device with
// that device current.
int scratch_on(int device, unsigned char** out)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(CUMF_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch[device] = (unsigned char*)ptr;
    }
    *out = g_scratch[device];
    return CUMF_OK;
}

int train_noalloc(int device, hipStream_t st, const cumf_train_opts* o, int32_t prec, const int32_t* u,
                  const int32_t* v, const float* r, int64_t nnz, int32_t m, int32_t n, void* P, void* Q, float* ps,
                  float* qs, int32_t e0, int32_t e1)
{
    unsigned char* base = nullptr;
    int rc = scratch_on(device, &base);
    if (rc) return rc;
    int cu = 0;
    hipError_t e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return fail(CUMF_ERR_HIP, "hipDeviceGetAttribute: %s", hipGetErrorString(e));
    // small factors train on one XCD, as cumf_train places them, with the
    // default worker count of that XCD's CUs
    const bool one_xcd = confine_to_one_xcd(o->schedule, o->coherent, m, n, o->k, prec);
    Shape sh;
    if ((rc = resolve_shape(ShapeInput{o->k, prec, one_xcd ? std::max(1, cu / kNumXcd) : cu, nnz}, o, &sh, 1)))
        return rc;

    // the task record and table cumf_train uploads for the same call
    // (base_task and upload_lr in capi.hip)
    float lr[kMaxLr];
    const int nlr = e1 + 4;
    lr_table(o, nlr, lr);
    SgdTask t{};
    t.u = u;
    t.v = v;
    t.r = r;
    t.nnz = nnz;
    t.P = P;
    t.Q = Q;
    t.ps = ps;
    t.qs = qs;
    t.lr = (const float*)(base + kTaskBytes);
    t.lambda_p = o->lambda_p;
    t.lambda_q = o->lambda_q;
    t.worker_begin = 0;
    t.worker_count = sh.W;
    // update_count = ceil(size / (W * L)) (singleGPU/sgd_kernel.cu:203)
    t.update_count = (int32_t)((nnz + (int64_t)sh.W * sh.L - 1) / ((int64_t)sh.W * sh.L));
    t.batch_len = sh.L;
    static_assert(sizeof(SgdTask) % 4 == 0, "the task record is staged as words");
    if ((e = put_words((uint32_t*)base, &t, sizeof(SgdTask) / 4, st)) != hipSuccess ||
        (e = put_words((uint32_t*)(base + kTaskBytes), lr, (size_t)nlr, st)) != hipSuccess)
        return fail(CUMF_ERR_HIP, "staging the task record: %s", hipGetErrorString(e));

    SgdLaunch a{};
    // factors under 4 GB each: rows addressed as launch-wide base + 32-bit offset
    const uint64_t eb = prec == CUMF_FP32 ? 4 : (prec == CUMF_FP8_E4M3 ? 1 : 2);
    if ((uint64_t)m * o->k * eb <= kBufRecords && (uint64_t)n * o->k * eb <= kBufRecords) {
        a.P = P;
        a.Q = Q;
        a.ps = ps;
        a.qs = qs;
    }
    a.tasks = (const SgdTask*)base;
    a.num_tasks = 1;
    a.batch_len = sh.L;
    a.seed = o->seed;
    a.num_xcd = one_xcd ? kNumXcd : 1;
    const int grid = one_xcd ? kNumXcd * sh.blocks : sh.blocks;
    a.epoch_dev = nullptr;
    for (int ep = e0; ep < e1; ++ep) {
        a.epoch_begin = ep;
        a.epoch_end = ep + 1;
        if ((e = launch_sgd(a, grid, o->k, prec, sh.G, o->coherent != 0, sh.D, st)) != hipSuccess)
            return fail(CUMF_ERR_HIP, "update launch: %s", hipGetErrorString(e));
        if (prec == CUMF_FP8_E4M3) {
            if ((e = launch_fp8_rescale((uint8_t*)P, ps, m, o->k, st)) != hipSuccess ||
                (e = launch_fp8_rescale((uint8_t*)Q, qs, n, o->k, st)) != hipSuccess)
                return fail(CUMF_ERR_HIP, "FP8 rescale launch: %s", hipGetErrorString(e));
        }
    }
    return CUMF_OK;
}

}  // namespace

extern "C" int cumf_train_noalloc(int device, void* stream, const cumf_train_opts* o, int32_t precision,
                                  const int32_t* u, const int32_t* v, const float* r, int64_t nnz, int32_t m,
                                  int32_t n, void* P, void* Q, float* ps, float* qs, int32_t epoch_begin,
                                  int32_t epoch_end)
{
    if (!o) return fail(CUMF_ERR_ARG, "null options");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess) return fail(CUMF_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
    if (device < 0 || device >= count || device >= kMaxDevices)
        return fail(CUMF_ERR_ARG, "device %d out of range (%d devices)", device, count);
    if (o->k != 32 && o->k != 64 && o->k != 128) return fail(CUMF_ERR_ARG, "k=%d unsupported (32, 64 or 128)", o->k);
    if (precision < CUMF_FP32 || precision > CUMF_BF16_SR) return fail(CUMF_ERR_ARG, "bad precision %d", precision);
    if (o->schedule != CUMF_SCHED_HOGWILD && o->schedule != CUMF_SCHED_SERIAL)
        return fail(CUMF_ERR_ARG, "schedule %d: the allocation-free entry point runs HOGWILD or SERIAL", o->schedule);
    if (o->use_graph != 0) return fail(CUMF_ERR_ARG, "use_graph must be 0 (capture the call itself instead)");
    if (nnz < 0 || m < 0 || n < 0) return fail(CUMF_ERR_ARG, "negative size");
    if (epoch_begin < 0 || epoch_end < epoch_begin)
        return fail(CUMF_ERR_ARG, "bad epoch range [%d,%d)", epoch_begin, epoch_end);
    if (epoch_end > CUMF_NOALLOC_MAX_EPOCHS)
        return fail(CUMF_ERR_ARG, "epoch_end=%d beyond the allocation-free limit of %d epochs (its learning-rate "
                    "table lives in a fixed %zu-byte device array)", epoch_end, CUMF_NOALLOC_MAX_EPOCHS, kScratchBytes);
    if (nnz == 0 || epoch_end == epoch_begin) return CUMF_OK;
    if (m <= 0 || n <= 0 || !u || !v || !r || !P || !Q)
        return fail(CUMF_ERR_ARG, "ratings and factors need m, n > 0 and non-null pointers");
    if (precision == CUMF_FP8_E4M3 && (!ps || !qs)) return fail(CUMF_ERR_ARG, "FP8 factors need their row scales");

    int old = 0;
    if ((e = hipGetDevice(&old)) != hipSuccess) return fail(CUMF_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
    if ((e = hipSetDevice(device)) != hipSuccess) return fail(CUMF_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    const int rc = train_noalloc(device, (hipStream_t)stream, o, precision, u, v, r, nnz, m, n, P, Q, ps, qs,
                                 epoch_begin, epoch_end);
    (void)hipSetDevice(old);
    return rc;
}
