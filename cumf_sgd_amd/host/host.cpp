// Host half of the pipeline: the reference's host code around the device path
// (singleGPU/sgd.cu:39-134, 137-335, 411-487, 593-690; singleGPU/main.cpp:231-273)
// restated in C++ and exported through the C ABI.  Compiled as C++14 because
// std::random_shuffle (removed in C++17) is what the reference's id permutation
// calls, and calling the same library routine keeps the permutation bit-exact.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "../../include/cumf_sgd_amd.h"
#include "../csrc/errors.h"

using cumf::fail;

namespace {

struct CtxGuard {
    cumf_ctx* c = nullptr;
    ~CtxGuard() { if (c) cumf_destroy(c); }
};

// Propagate an error from a C-ABI call (its message is already recorded).
#define CUMF_TRY(expr)                 \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != CUMF_OK) return rc_; \
    } while (0)

// Block processing order of the reference's gridded modes, drawn with the C
// library rand() in the same sequence of std::random_shuffle calls
// (singleGPU/sgd_kernel.cu:253-267 for u x v, :388-451 for u x v x x x y).
// Returns epoch-major block ids of the ux x vy grid.
std::vector<int32_t> reference_block_order(const cumf_param& p, int epochs)
{
    std::vector<int32_t> order;
    const int vy = p.v_grid * p.y_grid;
    if (p.x_grid * p.y_grid == 1) {
        std::vector<int> uid(p.u_grid), vid(p.v_grid);
        for (int i = 0; i < p.u_grid; ++i) uid[i] = i;
        for (int i = 0; i < p.v_grid; ++i) vid[i] = i;
        for (int it = 0; it < epochs; ++it) {
            std::random_shuffle(uid.begin(), uid.end());
            for (int ui = 0; ui < p.u_grid; ++ui) {
                std::random_shuffle(vid.begin(), vid.end());
                for (int vi = 0; vi < p.v_grid; ++vi) order.push_back(uid[ui] * p.v_grid + vid[vi]);
            }
        }
        return order;
    }
    std::vector<int> uid(p.u_grid), vid(p.v_grid), xid(p.x_grid), yid(p.y_grid);
    std::vector<int> uvid(p.u_grid * p.v_grid), xyid(p.x_grid * p.y_grid);
    for (int i = 0; i < p.u_grid; ++i) uid[i] = i;
    for (int i = 0; i < p.v_grid; ++i) vid[i] = i;
    for (int i = 0; i < p.x_grid; ++i) xid[i] = i;
    for (int i = 0; i < p.y_grid; ++i) yid[i] = i;
    for (int i = 0; i < p.u_grid * p.v_grid; ++i) uvid[i] = i;
    for (int i = 0; i < p.x_grid * p.y_grid; ++i) xyid[i] = i;
    for (int it = 0; it < epochs; ++it) {
        std::random_shuffle(uvid.begin(), uvid.end());
        std::random_shuffle(uid.begin(), uid.end());
        for (int ui = 0; ui < p.u_grid; ++ui) {
            // sgd_kernel.cu:416 shuffles the empty range [begin, begin): no draw
            for (int vi = 0; vi < p.v_grid; ++vi) {
                const int cu = uvid[ui * p.v_grid + vi] / p.v_grid;
                const int cv = uvid[ui * p.v_grid + vi] % p.v_grid;
                std::random_shuffle(xid.begin(), xid.end());
                std::random_shuffle(xyid.begin(), xyid.end());
                for (int lx = 0; lx < p.x_grid; ++lx) std::random_shuffle(yid.begin(), yid.end());
                // sub-blocks run in local-id order (the lists are indexed by
                // the local id, sgd_kernel.cu:440-454)
                for (int local = 0; local < p.x_grid * p.y_grid; ++local) {
                    const int gx = cu * p.x_grid + local / p.y_grid;
                    const int gy = cv * p.y_grid + local % p.y_grid;
                    order.push_back(gx * vy + gy);
                }
            }
        }
    }
    return order;
}

cumf_train_opts opts_from(const cumf_param& p, const cumf_train_opts* extra, float scale, uint64_t seed)
{
    cumf_train_opts o;
    std::memset(&o, 0, sizeof(o));
    if (extra) o = *extra;
    o.k = p.k;
    if (!extra || extra->num_workers == 0) o.num_workers = p.num_workers;
    o.lambda_p = p.lambda_p / scale;   // singleGPU/sgd.cu:673-674
    o.lambda_q = p.lambda_q / scale;
    o.alpha = p.alpha;
    o.beta = p.beta;
    o.lrate = p.lrate;
    o.seed = seed;
    return o;
}

int train_loaded(cumf_ctx* ctx, const cumf_param& p, const cumf_train_opts* extra, const std::vector<int32_t>& order,
                 float ave, float std_dev, int32_t precision, uint64_t seed, bool eval, const int32_t* pmap,
                 const int32_t* qmap, float* P_out, float* Q_out, float* b_out, double* rmse_out);

}  // namespace

extern "C" {

int cumf_gen_random_map(uint32_t seed, int32_t size, int32_t* map)
{
    if (size < 0 || (size > 0 && !map)) return fail(CUMF_ERR_ARG, "bad map arguments");
    // singleGPU/sgd.cu:39-53 with srand(seed) in place of srand(time(NULL))
    std::srand(seed);
    std::vector<int32_t> v((size_t)size);
    for (int32_t i = 0; i < size; ++i) v[i] = i;
    std::random_shuffle(v.begin(), v.end());
    std::copy(v.begin(), v.end(), map);
    return CUMF_OK;
}

int cumf_collect_data(const cumf_node* R, int64_t nnz, float* ave, float* std_dev)
{
    if (nnz <= 0 || !R || !ave || !std_dev) return fail(CUMF_ERR_ARG, "collect_data needs nnz > 0");
    double ex = 0, ex2 = 0;
    for (int64_t i = 0; i < nnz; ++i) {
        const float r = R[i].rate;
        ex += (double)r;
        ex2 += (double)r * r;
    }
    ex = ex / (double)nnz;
    ex2 = ex2 / (double)nnz;
    *ave = (float)ex;
    *std_dev = (float)std::sqrt(ex2 - ex * ex);
    return CUMF_OK;
}

int cumf_grid_problem(cumf_node* R, int64_t nnz, int32_t m, int32_t n, int32_t ux, int32_t vy, int64_t* grid_size,
                      int64_t* offsets, int64_t* u_seg_out, int64_t* v_seg_out)
{
    if (ux < 1 || vy < 1 || m < 1 || n < 1 || nnz < 0 || (nnz > 0 && !R) || !grid_size || !offsets)
        return fail(CUMF_ERR_ARG, "bad grid_problem arguments");
    const int64_t u_seg = ux == 1 ? m : (int64_t)std::ceil((double)m / ux);
    const int64_t v_seg = vy == 1 ? n : (int64_t)std::ceil((double)n / vy);
    const int nb = ux * vy;
    auto grid_id = [&](const cumf_node& x) { return (int)((x.u / u_seg) * vy + x.v / v_seg); };
    for (int64_t i = 0; i < nnz; ++i)
        if (R[i].u < 0 || R[i].u >= m || R[i].v < 0 || R[i].v >= n)
            return fail(CUMF_ERR_ARG, "record %lld has ids outside the grid", (long long)i);
    std::fill(grid_size, grid_size + nb, 0);
    for (int64_t i = 0; i < nnz; ++i) grid_size[grid_id(R[i])]++;
    offsets[0] = 0;
    for (int g = 0; g < nb; ++g) offsets[g + 1] = offsets[g] + grid_size[g];
    std::vector<int64_t> pivots(offsets, offsets + nb);
    for (int g = 0; g < nb; ++g) {
        for (int64_t p = pivots[g]; p != offsets[g + 1];) {
            const int cg = grid_id(R[p]);
            if (cg == g) {
                ++p;
                continue;
            }
            std::swap(R[p], R[pivots[cg]]);
            pivots[cg]++;
        }
    }
    if (u_seg_out) *u_seg_out = u_seg;
    if (v_seg_out) *v_seg_out = v_seg;
    return CUMF_OK;
}

void cumf_free_host(void* p) { std::free(p); }

int cumf_read_problem(const char* path, cumf_node** R_out, int64_t* nnz_out, int32_t* m_out, int32_t* n_out)
{
    if (!path || !*path || !R_out || !nnz_out || !m_out || !n_out) return fail(CUMF_ERR_ARG, "bad read_problem arguments");
    *R_out = nullptr;
    // shards path0 .. path79 (singleGPU/sgd.cu:207-217)
    std::vector<std::string> shards;
    for (int i = 0; i < 80; ++i) {
        std::string s = std::string(path) + std::to_string(i);
        if (access(s.c_str(), F_OK) != -1) shards.push_back(s);
    }
    const bool sharded = !shards.empty();
    if (!sharded) shards.push_back(path);
    std::vector<int64_t> size(shards.size()), offset(shards.size() + 1, 0);
    for (size_t i = 0; i < shards.size(); ++i) {
        FILE* f = std::fopen(shards[i].c_str(), "rb");
        if (!f) return fail(CUMF_ERR_IO, "file %s open failed", shards[i].c_str());
        std::fseek(f, 0L, SEEK_END);
        size[i] = std::ftell(f) / 12;   // a trailing partial record is dropped
        std::fclose(f);
        offset[i + 1] = offset[i] + size[i];   // offset[0] = 0 (the reference leaves it unset)
    }
    const int64_t nnz = offset.back();
    cumf_node* R = (cumf_node*)std::malloc(sizeof(cumf_node) * (size_t)(nnz > 0 ? nnz : 1));
    if (!R) return fail(CUMF_ERR_NOMEM, "cannot allocate %lld records", (long long)nnz);
    std::vector<int32_t> max_m(shards.size(), -1), max_n(shards.size(), -1);
    std::vector<int> ok(shards.size(), 1);
    auto read_one = [&](size_t i) {
        FILE* f = std::fopen(shards[i].c_str(), "rb");
        if (!f) { ok[i] = 0; return; }
        std::vector<cumf_node> buf(1 << 16);
        int64_t done = 0;
        while (done < size[i]) {
            const size_t want = (size_t)std::min<int64_t>((int64_t)buf.size(), size[i] - done);
            const size_t got = std::fread(buf.data(), sizeof(cumf_node), want, f);
            for (size_t j = 0; j < got; ++j) {
                max_m[i] = std::max(max_m[i], buf[j].u + 1);
                max_n[i] = std::max(max_n[i], buf[j].v + 1);
            }
            std::memcpy(R + offset[i] + done, buf.data(), got * sizeof(cumf_node));
            done += (int64_t)got;
            if (got < want) break;
        }
        if (done != size[i]) ok[i] = 0;
        std::fclose(f);
    };
    if (shards.size() == 1) {
        read_one(0);
    } else {
        std::vector<std::thread> th;
        for (size_t i = 0; i < shards.size(); ++i) th.emplace_back(read_one, i);
        for (auto& t : th) t.join();
    }
    for (size_t i = 0; i < shards.size(); ++i)
        if (!ok[i]) {
            std::free(R);
            return fail(CUMF_ERR_IO, "short read of %s", shards[i].c_str());
        }
    // single file: m, n start at 1 (sgd.cu:202-203); shards: max over shards
    int32_t m = sharded ? -1 : 1, n = sharded ? -1 : 1;
    for (size_t i = 0; i < shards.size(); ++i) {
        m = std::max(m, max_m[i]);
        n = std::max(n, max_n[i]);
    }
    *R_out = R;
    *nnz_out = nnz;
    *m_out = m;
    *n_out = n;
    return CUMF_OK;
}

int cumf_save_model(const char* path, int32_t m, int32_t n, int32_t k, float b, const float* P, const float* Q)
{
    if (!path || m < 0 || n < 0 || k <= 0 || (m && !P) || (n && !Q)) return fail(CUMF_ERR_ARG, "bad save_model arguments");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(CUMF_ERR_IO, "save model failed: cannot open %s", path);
    const int32_t hdr[4] = {0, m, n, k};   // singleGPU/main.cpp:247-252
    bool ok = std::fwrite(hdr, sizeof(int32_t), 4, f) == 4;
    ok = ok && std::fwrite(&b, sizeof(float), 1, f) == 1;
    ok = ok && std::fwrite(P, sizeof(float), (size_t)m * k, f) == (size_t)m * k;
    ok = ok && std::fwrite(Q, sizeof(float), (size_t)n * k, f) == (size_t)n * k;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(CUMF_ERR_IO, "short write to %s", path);
    return CUMF_OK;
}

int cumf_load_model(const char* path, int32_t* m, int32_t* n, int32_t* k, float* b, float** P, float** Q)
{
    if (!path || !m || !n || !k || !b || !P || !Q) return fail(CUMF_ERR_ARG, "null argument");
    *P = *Q = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(CUMF_ERR_IO, "%s open failed", path);
    int32_t hdr[4];
    float bb = 0.0f;
    bool ok = std::fread(hdr, sizeof(int32_t), 4, f) == 4 && std::fread(&bb, sizeof(float), 1, f) == 1;
    if (!ok || hdr[1] < 0 || hdr[2] < 0 || hdr[3] < 1) {
        std::fclose(f);
        return fail(CUMF_ERR_IO, "%s: bad model header", path);
    }
    const size_t np = (size_t)hdr[1] * hdr[3], nq = (size_t)hdr[2] * hdr[3];
    float* p = (float*)std::malloc(sizeof(float) * (np ? np : 1));
    float* q = (float*)std::malloc(sizeof(float) * (nq ? nq : 1));
    if (!p || !q) {
        std::free(p);
        std::free(q);
        std::fclose(f);
        return fail(CUMF_ERR_NOMEM, "model of %zu floats does not fit in host memory", np + nq);
    }
    ok = std::fread(p, sizeof(float), np, f) == np && std::fread(q, sizeof(float), nq, f) == nq;
    std::fclose(f);
    if (!ok) {
        std::free(p);
        std::free(q);
        return fail(CUMF_ERR_IO, "%s: file shorter than its %d x %d x %d model", path, hdr[1], hdr[2], hdr[3]);
    }
    *m = hdr[1];
    *n = hdr[2];
    *k = hdr[3];
    *b = bb;
    *P = p;
    *Q = q;
    return CUMF_OK;
}

int cumf_sgd_update_k128(const cumf_param* para, int32_t m, int32_t n, int64_t u_seg, int64_t v_seg,
                         const cumf_node* const* R2D, const int64_t* grid_size, uint16_t* halfp, uint16_t* halfq,
                         float* floatp, float* floatq, uint64_t seed)
{
    if (!para || !R2D || !grid_size || !halfp || !halfq || !floatp || !floatq)
        return fail(CUMF_ERR_ARG, "null argument");
    const cumf_param& p = *para;
    const int ux = p.ux, vy = p.vy, nb = ux * vy;
    if (ux < 1 || vy < 1 || u_seg < 1 || v_seg < 1 || (int64_t)ux * u_seg < m || (int64_t)vy * v_seg < n)
        return fail(CUMF_ERR_ARG, "grid does not cover the matrix");
    // globalize the block-local ids (scale_problem localises them, sgd.cu:111-122)
    std::vector<int64_t> off((size_t)nb + 1, 0);
    for (int b = 0; b < nb; ++b) off[b + 1] = off[b] + grid_size[b];
    std::vector<cumf_node> R((size_t)off[nb]);
    for (int b = 0; b < nb; ++b) {
        const int64_t du = nb > 1 ? (int64_t)(b / vy) * u_seg : 0, dv = nb > 1 ? (int64_t)(b % vy) * v_seg : 0;
        for (int64_t i = 0; i < grid_size[b]; ++i) {
            cumf_node x = R2D[b][i];
            x.u = (int32_t)(x.u + du);
            x.v = (int32_t)(x.v + dv);
            R[(size_t)(off[b] + i)] = x;
        }
    }
    const int32_t rows_p = (int32_t)(ux * u_seg), rows_q = (int32_t)(vy * v_seg);
    CtxGuard g;
    CUMF_TRY(cumf_create(p.gpu, &g.c));
    CUMF_TRY(cumf_set_ratings(g.c, R.data(), (int64_t)R.size(), rows_p, rows_q));
    if (nb > 1) CUMF_TRY(cumf_set_blocks(g.c, ux, vy, off.data()));
    CUMF_TRY(cumf_init_factors(g.c, p.k, CUMF_FP16, seed));
    CUMF_TRY(cumf_set_factors(g.c, 0, halfp, nullptr));
    CUMF_TRY(cumf_set_factors(g.c, 1, halfq, nullptr));
    cumf_train_opts o;
    std::memset(&o, 0, sizeof(o));
    o.k = p.k;
    o.num_workers = p.num_workers;
    o.lambda_p = p.lambda_p;   // already divided by scale by the caller
    o.lambda_q = p.lambda_q;
    o.alpha = p.alpha;
    o.beta = p.beta;
    o.lrate = p.lrate;
    o.seed = seed;
    if (p.u_grid * p.v_grid == 1) {
        // the 1x1 grid trains the library's default schedule for these
        // ratings: conflict-free 24 x 24 strata, or Hogwild on data skewed
        // on both axes
        int32_t order = 0;
        CUMF_TRY(cumf_default_schedule(g.c, p.k, CUMF_FP16, &o.schedule, &o.stratum_rounds, &order));
        if (o.schedule == CUMF_SCHED_STRATUM) CUMF_TRY(cumf_partition_strata(g.c, 24, nullptr));
        else if (o.schedule == CUMF_SCHED_TILE) CUMF_TRY(cumf_partition_tiles(g.c, 0, 0, nullptr));
        else if (order) CUMF_TRY(cumf_partition(g.c, 1, 1, nullptr));
        if (order) CUMF_TRY(cumf_order_blocks(g.c, order));
        CUMF_TRY(cumf_train(g.c, &o, 0, p.num_iters, nullptr, 0));
    } else {
        o.schedule = CUMF_SCHED_GRID;
        std::vector<int32_t> order = reference_block_order(p, p.num_iters);
        CUMF_TRY(cumf_train(g.c, &o, 0, p.num_iters, order.data(), (int64_t)order.size()));
    }
    CUMF_TRY(cumf_get_factors(g.c, 0, halfp, nullptr));
    CUMF_TRY(cumf_get_factors(g.c, 1, halfq, nullptr));
    CUMF_TRY(cumf_export_factors(g.c, 0, 1.0f, nullptr, floatp));
    CUMF_TRY(cumf_export_factors(g.c, 1, 1.0f, nullptr, floatq));
    return CUMF_OK;
}

int cumf_sgd_train(const cumf_param* para, const cumf_train_opts* extra, const cumf_node* R, int64_t nnz, int32_t m,
                   int32_t n, int32_t precision, uint64_t seed, const cumf_node* test, int64_t test_nnz,
                   float* P_out, float* Q_out, float* b_out, double* rmse_out)
{
    if (!para || !R || nnz <= 0 || m < 1 || n < 1 || !P_out || !Q_out || !b_out)
        return fail(CUMF_ERR_ARG, "bad sgd_train arguments");
    const cumf_param& p = *para;
    // collect_data + scale (singleGPU/sgd.cu:644-649)
    float ave = 0.0f, std_dev = 0.0f;
    CUMF_TRY(cumf_collect_data(R, nnz, &ave, &std_dev));
    // id permutations (sgd.cu:656-659): both maps restart srand(seed)
    std::vector<int32_t> pmap((size_t)m), qmap((size_t)n);
    CUMF_TRY(cumf_gen_random_map((uint32_t)seed, m, pmap.data()));
    CUMF_TRY(cumf_gen_random_map((uint32_t)seed, n, qmap.data()));

    const int ux = p.ux, vy = p.vy;
    CtxGuard g;
    CUMF_TRY(cumf_create(p.gpu, &g.c));
    std::vector<int32_t> order;
    if (ux * vy == 1) {
        CUMF_TRY(cumf_set_ratings(g.c, R, nnz, m, n));
        if (test && test_nnz > 0) CUMF_TRY(cumf_set_test(g.c, test, test_nnz));
        CUMF_TRY(cumf_remap(g.c, pmap.data(), qmap.data()));
    } else {
        // host shuffle + cycle-swap grid so block contents and order are the
        // reference's (shuffle_problem sgd.cu:126-134, grid_problem :411-487)
        std::vector<cumf_node> H(R, R + nnz);
        for (int64_t i = 0; i < nnz; ++i) {
            if (H[i].u < 0 || H[i].u >= m || H[i].v < 0 || H[i].v >= n) return fail(CUMF_ERR_ARG, "rating id out of range");
            H[i].u = pmap[H[i].u];
            H[i].v = qmap[H[i].v];
        }
        std::vector<int64_t> gs((size_t)ux * vy), off((size_t)ux * vy + 1);
        int64_t us, vs;
        CUMF_TRY(cumf_grid_problem(H.data(), nnz, m, n, ux, vy, gs.data(), off.data(), &us, &vs));
        CUMF_TRY(cumf_set_ratings(g.c, H.data(), nnz, m, n));
        CUMF_TRY(cumf_set_blocks(g.c, ux, vy, off.data()));
        if (test && test_nnz > 0) {
            std::vector<cumf_node> T(test, test + test_nnz);
            for (auto& x : T) {
                if (x.u >= 0 && x.u < m) x.u = pmap[x.u];
                if (x.v >= 0 && x.v < n) x.v = qmap[x.v];
            }
            CUMF_TRY(cumf_set_test(g.c, T.data(), test_nnz));
        }
        order = reference_block_order(p, p.num_iters);
    }
    return train_loaded(g.c, p, extra, order, ave, std_dev, precision, seed, test && test_nnz > 0 && rmse_out,
                        pmap.data(), qmap.data(), P_out, Q_out, b_out, rmse_out);
}

int cumf_sgd_train_file(const cumf_param* para, const cumf_train_opts* extra, const char* train_path,
                        const char* test_path, int32_t precision, uint64_t seed, float** P_out, float** Q_out,
                        int32_t* m_out, int32_t* n_out, int64_t* nnz_out, float* b_out, double* rmse_out)
{
    if (!para || !train_path || !P_out || !Q_out || !m_out || !n_out || !b_out)
        return fail(CUMF_ERR_ARG, "bad sgd_train_file arguments");
    *P_out = *Q_out = nullptr;
    const cumf_param& p = *para;
    if (p.ux * p.vy != 1) return fail(CUMF_ERR_ARG, "the file pipeline trains the 1x1 grid only (use cumf_sgd_train)");
    for (const char* f : {train_path, test_path}) {   // file errors before device errors, as read_problem
        if (!f || !*f) continue;
        bool found = access(f, F_OK) != -1;
        for (int i = 0; i < 80 && !found; ++i) found = access((std::string(f) + std::to_string(i)).c_str(), F_OK) != -1;
        if (!found) return fail(CUMF_ERR_IO, "file %s open failed", f);
    }
    CtxGuard g;
    CUMF_TRY(cumf_create(p.gpu, &g.c));
    int64_t nnz = 0;
    int32_t m = 0, n = 0;
    CUMF_TRY(cumf_read_ratings(g.c, train_path, &nnz, &m, &n));
    if (nnz <= 0) return fail(CUMF_ERR_ARG, "%s holds no ratings", train_path);
    const bool eval = test_path && *test_path && rmse_out;
    if (test_path && *test_path) CUMF_TRY(cumf_read_test(g.c, test_path, nullptr));
    // collect_data (sgd.cu:78-96) on the device sums (FP64, another order)
    double s1 = 0, s2 = 0;
    CUMF_TRY(cumf_rating_moments(g.c, &s1, &s2));
    const double ex = s1 / (double)nnz, ex2 = s2 / (double)nnz;
    const float ave = (float)ex, std_dev = (float)std::sqrt(ex2 - ex * ex);
    std::vector<int32_t> pmap((size_t)m), qmap((size_t)n);
    CUMF_TRY(cumf_gen_random_map((uint32_t)seed, m, pmap.data()));
    CUMF_TRY(cumf_gen_random_map((uint32_t)seed, n, qmap.data()));
    CUMF_TRY(cumf_remap(g.c, pmap.data(), qmap.data()));
    float* P = (float*)std::malloc(sizeof(float) * (size_t)m * p.k);
    float* Q = (float*)std::malloc(sizeof(float) * (size_t)n * p.k);
    if (!P || !Q) {
        std::free(P);
        std::free(Q);
        return fail(CUMF_ERR_NOMEM, "model does not fit in host memory");
    }
    const std::vector<int32_t> no_order;
    int rc = train_loaded(g.c, p, extra, no_order, ave, std_dev, precision, seed, eval, pmap.data(), qmap.data(), P,
                          Q, b_out, rmse_out);
    if (rc != CUMF_OK) {
        std::free(P);
        std::free(Q);
        return rc;
    }
    *P_out = P;
    *Q_out = Q;
    *m_out = m;
    *n_out = n;
    if (nnz_out) *nnz_out = nnz;
    return CUMF_OK;
}

}  // extern "C"

namespace {

// Scale, init, train, export: the part of sgd_train (singleGPU/sgd.cu:670-690)
// after the ratings are on the device in their permuted ids.
int train_loaded(cumf_ctx* ctx, const cumf_param& p, const cumf_train_opts* extra, const std::vector<int32_t>& order,
                 float ave, float std_dev, int32_t precision, uint64_t seed, bool eval, const int32_t* pmap,
                 const int32_t* qmap, float* P_out, float* Q_out, float* b_out, double* rmse_out)
{
    const int ux = p.ux, vy = p.vy;
    const float scale = std::max(1e-4f, std_dev);
    CUMF_TRY(cumf_scale_ratings(ctx, (float)(1.0 / scale)));   // sgd.cu:672
    CUMF_TRY(cumf_init_factors(ctx, p.k, precision, seed));
    const float b = ave / std_dev;   // init_model(tr, k, ave/std_dev), sgd.cu:679
    const float fs = std::sqrt(scale);   // scale_model, sgd.cu:598

    cumf_train_opts o = opts_from(p, extra, scale, seed);
    // A 1x1 grid trains the library's default schedule for these ratings
    // (conflict-free strata, or Hogwild on data skewed on both axes);
    // gridded runs keep the reference's block order.
    int32_t block_order = 0;   // record order inside the stratum blocks (0: as given)
    if (!extra || extra->schedule == CUMF_SCHED_AUTO) {
        if (ux * vy == 1) {
            int32_t rounds = 1;
            CUMF_TRY(cumf_default_schedule(ctx, p.k, precision, &o.schedule, &rounds, &block_order));
            if (!extra || extra->stratum_rounds == 0) o.stratum_rounds = rounds;
        } else {
            o.schedule = CUMF_SCHED_GRID;
        }
    }
    // each epoch's launches replay as one captured graph (gridded runs with a
    // new block order per epoch launch one by one)
    if (!extra) o.use_graph = 1;
    if (o.schedule == CUMF_SCHED_STRATUM) {
        // a square grid gives s; otherwise 24 x 24 blocks, the measured
        // fastest for 16-bit factors on a Netflix-shaped problem
        const int s = ux == vy && ux > 1 ? ux : 24;
        CUMF_TRY(cumf_partition_strata(ctx, s, nullptr));
        if (block_order) CUMF_TRY(cumf_order_blocks(ctx, block_order));
    } else if (o.schedule == CUMF_SCHED_TILE) {
        // row tiles of 8 x 8 strata (a square grid gives s)
        CUMF_TRY(cumf_partition_tiles(ctx, ux == vy && ux > 1 ? ux : 0, 0, nullptr));
    } else if (block_order && o.schedule == CUMF_SCHED_HOGWILD) {
        // Hogwild's one block: all records grouped by column
        CUMF_TRY(cumf_partition(ctx, 1, 1, nullptr));
        CUMF_TRY(cumf_order_blocks(ctx, block_order));
    }
    for (int ep = 0; ep < p.num_iters; ++ep) {
        const int32_t* ord = nullptr;
        int64_t olen = 0;
        if (o.schedule == CUMF_SCHED_GRID && !order.empty()) {
            ord = order.data() + (size_t)ep * ux * vy;
            olen = (int64_t)ux * vy;
        }
        CUMF_TRY(cumf_train(ctx, &o, ep, ep + 1, ord, olen));
        if (eval) CUMF_TRY(cumf_test_rmse(ctx, fs, b, &rmse_out[ep]));
    }
    // export in original id order: row u of the model is shuffled row pmap[u]
    // (scale_model + shuffle_model, sgd.cu:593-635)
    CUMF_TRY(cumf_export_factors(ctx, 0, fs, pmap, P_out));
    CUMF_TRY(cumf_export_factors(ctx, 1, fs, qmap, Q_out));
    *b_out = b;
    return CUMF_OK;
}

}  // namespace
