// Streaming triplet reader: the reference's read_problem (singleGPU/sgd.cu:
// 190-335: one file, or shards path0..path79 read by one pthread each, into
// pinned host memory) fused with the upload, so a 37 GB input goes from the
// files to the resident device arrays without a full host copy.
//
// Files are cut into chunks of whole 12-byte records.  Reader threads pread
// chunks into a small ring of pinned buffers; the calling thread copies each
// filled buffer to a device staging buffer and splits it into the u/v/r
// arrays at the chunk's place, tracking the largest ids and the smallest id
// on the device, then recycles the buffer when its copy event completes.
// A file's trailing partial record is dropped, as the reference's size / 12.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <climits>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cumf_sgd_amd.h"
#include "errors.h"
#include "loader.h"

namespace cumf {

namespace {

constexpr int64_t kChunkRecords = 4 << 20;   // 48 MiB per chunk
constexpr int kBuffers = 4;

__device__ inline int wave_max_i(int x)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) x = max(x, __shfl_xor(x, d, 64));
    return x;
}
__device__ inline int wave_min_i(int x)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) x = min(x, __shfl_xor(x, d, 64));
    return x;
}

// AoS records -> u/v/r, with the largest u, the largest v and the smallest
// id of the chunk folded into mm[0..2].
__global__ __launch_bounds__(256) void split_ids_kernel(const int32_t* aos, int64_t n, int32_t* u, int32_t* v,
                                                        float* r, int32_t* mm)
{
    int mu = -1, mv = -1, mn = INT_MAX;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = aos[3 * i], b = aos[3 * i + 1];
        u[i] = a;
        v[i] = b;
        r[i] = __int_as_float(aos[3 * i + 2]);
        mu = max(mu, a);
        mv = max(mv, b);
        mn = min(mn, min(a, b));
    }
    mu = wave_max_i(mu);
    mv = wave_max_i(mv);
    mn = wave_min_i(mn);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&mm[0], mu);
        atomicMax(&mm[1], mv);
        atomicMin(&mm[2], mn);
    }
}

struct Chunk {
    int file;
    int64_t first;   // first record in the file
    int64_t count;
    int64_t dst;     // first record in the output
};

bool pread_all(int fd, char* p, int64_t bytes, int64_t off)
{
    while (bytes > 0) {
        const ssize_t got = ::pread(fd, p, (size_t)std::min<int64_t>(bytes, 1 << 30), (off_t)off);
        if (got <= 0) return false;
        p += got;
        bytes -= got;
        off += got;
    }
    return true;
}

struct HostBuf {
    void* p = nullptr;
    ~HostBuf() { if (p) (void)hipHostFree(p); }
};
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};
struct Event {
    hipEvent_t e = nullptr;
    ~Event() { if (e) (void)hipEventDestroy(e); }
};

}  // namespace

int triplet_files(const char* path, std::vector<std::string>* files, std::vector<int64_t>* counts, bool* sharded)
{
    if (!path || !*path) return fail(CUMF_ERR_ARG, "empty path");
    files->clear();
    counts->clear();
    for (int i = 0; i < 80; ++i) {   // shards path0 .. path79 (singleGPU/sgd.cu:207-217)
        std::string s = std::string(path) + std::to_string(i);
        if (::access(s.c_str(), F_OK) != -1) files->push_back(s);
    }
    *sharded = !files->empty();
    if (!*sharded) files->push_back(path);
    for (const std::string& f : *files) {
        struct stat st;
        if (::stat(f.c_str(), &st) != 0) return fail(CUMF_ERR_IO, "file %s open failed", f.c_str());
        counts->push_back((int64_t)st.st_size / 12);
    }
    return CUMF_OK;
}

int stream_triplets(const std::vector<std::string>& files, const std::vector<int64_t>& counts, int32_t* u, int32_t* v,
                    float* r, hipStream_t st, int32_t mm_out[3])
{
    mm_out[0] = mm_out[1] = -1;
    mm_out[2] = INT_MAX;
    std::vector<Chunk> chunks;
    int64_t base = 0;
    for (size_t f = 0; f < files.size(); ++f) {
        for (int64_t s = 0; s < counts[f]; s += kChunkRecords)
            chunks.push_back(Chunk{(int)f, s, std::min(kChunkRecords, counts[f] - s), base + s});
        base += counts[f];
    }
    if (chunks.empty()) return CUMF_OK;

    const int nbuf = (int)std::min<size_t>(kBuffers, chunks.size());
    const size_t bytes = (size_t)kChunkRecords * 12;
    HostBuf host[kBuffers];
    DevBuf stage[kBuffers];
    Event ev[kBuffers];
    DevBuf mm;
    for (int i = 0; i < nbuf; ++i) {
        if (hipHostMalloc(&host[i].p, bytes, hipHostMallocDefault) != hipSuccess)
            return fail(CUMF_ERR_NOMEM, "cannot allocate %zu bytes of pinned memory", bytes);
        if (hipMalloc(&stage[i].p, bytes) != hipSuccess) return fail(CUMF_ERR_NOMEM, "cannot allocate staging memory");
        if (hipEventCreateWithFlags(&ev[i].e, hipEventDisableTiming) != hipSuccess)
            return fail(CUMF_ERR_HIP, "hipEventCreate failed");
    }
    if (hipMalloc(&mm.p, sizeof(int32_t) * 3) != hipSuccess) return fail(CUMF_ERR_NOMEM, "cannot allocate id range");
    if (hipMemcpyAsync(mm.p, mm_out, sizeof(int32_t) * 3, hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(CUMF_ERR_HIP, "id range upload failed");

    std::mutex mu;
    std::condition_variable cv_read, cv_main;
    std::vector<int> free_bufs;
    for (int i = 0; i < nbuf; ++i) free_bufs.push_back(i);
    std::deque<std::pair<int, int>> filled;   // (buffer, chunk)
    size_t next = 0;
    bool stop = false;
    std::string read_error;

    auto reader = [&]() {
        std::vector<int> fds(files.size(), -1);
        for (;;) {
            int buf, ci;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (stop || next >= chunks.size()) break;
                ci = (int)next++;
                cv_read.wait(lk, [&] { return stop || !free_bufs.empty(); });
                if (stop) break;
                buf = free_bufs.back();
                free_bufs.pop_back();
            }
            const Chunk& c = chunks[(size_t)ci];
            bool ok = true;
            if (fds[(size_t)c.file] < 0) fds[(size_t)c.file] = ::open(files[(size_t)c.file].c_str(), O_RDONLY);
            ok = fds[(size_t)c.file] >= 0 && pread_all(fds[(size_t)c.file], (char*)host[buf].p, c.count * 12, c.first * 12);
            std::lock_guard<std::mutex> lk(mu);
            if (!ok) {
                if (read_error.empty()) read_error = "short read of " + files[(size_t)c.file];
                stop = true;
                cv_read.notify_all();
            } else {
                filled.emplace_back(buf, ci);
            }
            cv_main.notify_one();
        }
        for (int fd : fds)
            if (fd >= 0) ::close(fd);
    };
    const int nthreads = (int)std::min<size_t>(std::min<size_t>(8, chunks.size()), std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int i = 0; i < nthreads; ++i) th.emplace_back(reader);

    std::deque<int> inflight;
    size_t done = 0;
    hipError_t herr = hipSuccess;
    auto release = [&](int buf) {
        std::lock_guard<std::mutex> lk(mu);
        free_bufs.push_back(buf);
        cv_read.notify_one();
    };
    while (done < chunks.size() && herr == hipSuccess) {
        int buf = -1, ci = -1;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_main.wait(lk, [&] { return stop || !filled.empty() || !inflight.empty(); });
            if (stop) break;
            if (!filled.empty()) {
                buf = filled.front().first;
                ci = filled.front().second;
                filled.pop_front();
            }
        }
        if (buf < 0) {   // nothing read yet: retire the oldest copy to free its buffer
            const int old = inflight.front();
            inflight.pop_front();
            if ((herr = hipEventSynchronize(ev[old].e)) != hipSuccess) break;
            release(old);
            continue;
        }
        const Chunk& c = chunks[(size_t)ci];
        if ((herr = hipMemcpyAsync(stage[buf].p, host[buf].p, (size_t)c.count * 12, hipMemcpyHostToDevice, st)) != hipSuccess)
            break;
        int64_t blocks = (c.count + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(split_ids_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (const int32_t*)stage[buf].p,
                           c.count, u + c.dst, v + c.dst, r + c.dst, (int32_t*)mm.p);
        if ((herr = hipGetLastError()) != hipSuccess) break;
        if ((herr = hipEventRecord(ev[buf].e, st)) != hipSuccess) break;
        inflight.push_back(buf);
        ++done;
        // recycle copies that have finished without waiting
        while (!inflight.empty() && hipEventQuery(ev[inflight.front()].e) == hipSuccess) {
            release(inflight.front());
            inflight.pop_front();
        }
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        if (done < chunks.size()) stop = true;
        cv_read.notify_all();
    }
    for (std::thread& t : th) t.join();
    const hipError_t serr = hipStreamSynchronize(st);
    if (!read_error.empty()) return fail(CUMF_ERR_IO, "%s", read_error.c_str());
    if (herr != hipSuccess) return fail(CUMF_ERR_HIP, "upload failed: %s", hipGetErrorString(herr));
    if (serr != hipSuccess) return fail(CUMF_ERR_HIP, "upload failed: %s", hipGetErrorString(serr));
    if (hipMemcpy(mm_out, mm.p, sizeof(int32_t) * 3, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(CUMF_ERR_HIP, "id range download failed");
    return CUMF_OK;
}

}  // namespace cumf
