// Host side of the ranking criteria (calc_mpr_auc, test/mf.cpp:3914-4034):
// the test positives (r > 0) grouped by row with sorted columns, the rows
// that have positives and at least one negative column, one kernel, and the
// per-row terms summed in row order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/cumf_sgd_amd.h"
#include "errors.h"
#include "kernels.h"

namespace cumf {

namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

int rank_eval(hipStream_t st, const int32_t* tu, const int32_t* tv, const float* tr, int64_t tnnz, const void* P,
              const float* ps, const void* Q, const float* qs, int32_t m, int32_t n, int k, int prec, float mult,
              float b, bool transpose, double* mpr, double* auc)
{
#define RTRY(expr)                                                                                     \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(CUMF_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
    } while (0)
    std::vector<int32_t> hu((size_t)tnnz), hv((size_t)tnnz);
    std::vector<float> hr((size_t)tnnz);
    if (tnnz) {
        RTRY(hipMemcpyAsync(hu.data(), tu, sizeof(int32_t) * tnnz, hipMemcpyDeviceToHost, st));
        RTRY(hipMemcpyAsync(hv.data(), tv, sizeof(int32_t) * tnnz, hipMemcpyDeviceToHost, st));
        RTRY(hipMemcpyAsync(hr.data(), tr, sizeof(float) * tnnz, hipMemcpyDeviceToHost, st));
        RTRY(hipStreamSynchronize(st));
    }
    // m, n of the evaluation: max(test extent, model extent) (mf.cpp:3919-3932)
    int64_t pm = 0, pn = 0;
    for (int64_t i = 0; i < tnnz; ++i) {
        pm = std::max<int64_t>(pm, (int64_t)hu[(size_t)i] + 1);
        pn = std::max<int64_t>(pn, (int64_t)hv[(size_t)i] + 1);
    }
    const int64_t nrow_all = transpose ? std::max<int64_t>(pn, n) : std::max<int64_t>(pm, m);
    const int64_t ncols = transpose ? std::max<int64_t>(pm, m) : std::max<int64_t>(pn, n);
    if (ncols > INT32_MAX) return fail(CUMF_ERR_ARG, "too many columns for the ranking criteria");
    std::vector<std::pair<int32_t, int32_t>> pos;
    for (int64_t i = 0; i < tnnz; ++i) {
        if (!(hr[(size_t)i] > 0)) continue;
        const int32_t r = transpose ? hv[(size_t)i] : hu[(size_t)i], c = transpose ? hu[(size_t)i] : hv[(size_t)i];
        if (r < 0 || c < 0) return fail(CUMF_ERR_ARG, "ranking criteria need non-negative test ids");
        pos.emplace_back(r, c);
    }
    std::sort(pos.begin(), pos.end());
    if (std::adjacent_find(pos.begin(), pos.end()) != pos.end())
        return fail(CUMF_ERR_ARG, "ranking criteria need distinct (row, column) test positives");
    std::vector<int32_t> rows, cols;
    std::vector<int64_t> beg{0};
    for (size_t i = 0; i < pos.size();) {
        size_t j = i;
        while (j < pos.size() && pos[j].first == pos[i].first) ++j;
        const int64_t cnt = (int64_t)(j - i);
        if (ncols - cnt >= 1) {
            rows.push_back(pos[i].first);
            for (size_t t = i; t < j; ++t) cols.push_back(pos[t].second);
            beg.push_back((int64_t)cols.size());
        }
        i = j;
    }
    (void)nrow_all;
    const int32_t nrows = (int32_t)rows.size();
    if (nrows == 0) {
        *mpr = *auc = __builtin_nan("");
        return CUMF_OK;
    }
    DevBuf drows, dbeg, dcols, dmpr, dauc;
    RTRY(hipMalloc(&drows.p, sizeof(int32_t) * rows.size()));
    RTRY(hipMalloc(&dbeg.p, sizeof(int64_t) * beg.size()));
    RTRY(hipMalloc(&dcols.p, sizeof(int32_t) * cols.size()));
    RTRY(hipMalloc(&dmpr.p, sizeof(double) * rows.size()));
    RTRY(hipMalloc(&dauc.p, sizeof(double) * rows.size()));
    RTRY(hipMemcpyAsync(drows.p, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice, st));
    RTRY(hipMemcpyAsync(dbeg.p, beg.data(), sizeof(int64_t) * beg.size(), hipMemcpyHostToDevice, st));
    RTRY(hipMemcpyAsync(dcols.p, cols.data(), sizeof(int32_t) * cols.size(), hipMemcpyHostToDevice, st));
    RTRY(launch_rank((const int32_t*)drows.p, (const int64_t*)dbeg.p, (const int32_t*)dcols.p, nrows, ncols, P, ps, Q,
                     qs, m, n, k, prec, mult, b, transpose ? 1 : 0, (double*)dmpr.p, (double*)dauc.p, st));
    std::vector<double> hm(rows.size()), ha(rows.size());
    RTRY(hipMemcpyAsync(hm.data(), dmpr.p, sizeof(double) * rows.size(), hipMemcpyDeviceToHost, st));
    RTRY(hipMemcpyAsync(ha.data(), dauc.p, sizeof(double) * rows.size(), hipMemcpyDeviceToHost, st));
    RTRY(hipStreamSynchronize(st));
    double sm = 0, sa = 0;
    for (size_t i = 0; i < rows.size(); ++i) {
        sm += hm[i];
        sa += ha[i];
    }
    *mpr = sm / (double)cols.size();   // all_u_mpr / total_pos
    *auc = sa / (double)rows.size();   // all_u_auc / total_m
    return CUMF_OK;
#undef RTRY
}

}  // namespace cumf
