// The batch-Hogwild update kernel template, instantiated per factor format in
// sgd_update_<format>.hip (one translation unit each, compiled in parallel).
#pragma once

#include <type_traits>

#include "kernels.h"

namespace cumf {

// ---------------------------------------------------------------------------
// Batch-Hogwild update.
//
// A wave of 64 lanes holds S = 64/G workers ("slots"); the G lanes of a slot
// share one rating update, lane l of the slot owning the E = K/G consecutive
// elements [lE, lE+E) of the P and Q rows, so a row is one coalesced
// K*bytes gather.  Per window, a slot's L records (u, v, r) are staged in
// LDS (loaded straight into LDS with global_load_lds, or in tile launches
// through registers; the next window's records load during the current
// window's last round), and the slot walks them in order as a D-deep
// software pipeline: the rows of sample j + D are requested right after
// sample j is written back, so D row pairs are in flight per slot.  A prefetched row may predate the write-back
// of one of the D - 1 samples before it; those written rows are kept in a
// register ring and forwarded when the ids match, so one slot is exactly a
// serial SGD over its window sequence.
//
// Arithmetic (shared with the CPU oracle): lane partial = p0*q0 then fma over
// the lane's other elements; butterfly sum over the slot (lane l + lane l^d,
// d = 1, 2, 4, ...); e = r - dot; with cp = fma(-lr, lp, 1), cq likewise and
// le = lr*e: p' = fma(le, q, p*cp) and q' = fma(le, p, q*cq) with the old p;
// store rounds p', q' to the storage format.  FP8 rows are used in their own
// units, their power-of-two scales entering once per sample (see the FP8
// write-back below).
// ---------------------------------------------------------------------------
constexpr int kWavesPerBlock = kSgdWavesPerBlock;
constexpr int kMaxBatch = kSgdMaxBatch;

template <int CTRL>
__device__ inline float dpp(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, true));
}

// Butterfly sum over aligned groups of G lanes.  Every lane of a group ends
// with the same value: the balanced pairwise tree over the group in lane
// order.  Steps 1, 2 use quad permutes, 4 and 8 the half-row / row mirrors
// (lane i reads lane 7-i / 15-i, whose partial sums equal those of lane
// i^4 / i^8 at that point), 16 a swizzle, and the two rows-pairs of a
// 64-lane group are combined from lanes 0, 16, 32, 48.
template <int G>
__device__ inline float slot_sum(float x)
{
    if constexpr (G >= 2) x += dpp<0xB1>(x);    // quad_perm [1,0,3,2]
    if constexpr (G >= 4) x += dpp<0x4E>(x);    // quad_perm [2,3,0,1]
    if constexpr (G >= 8) x += dpp<0x141>(x);   // row_half_mirror
    if constexpr (G >= 16) x += dpp<0x140>(x);  // row_mirror
    if constexpr (G == 32) x += __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(x), 0x401F));
    if constexpr (G == 64) {
        const float a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 0)) +
                        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 16));
        const float b = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 32)) +
                        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 48));
        x = a + b;
    }
    return x;
}

template <int NW>
struct Pending {
    Words<NW> p, q;
    int u, v;
    float r, ps, qs;
    uint32_t pa, qa;   // the lane's offset of row u in the tile (LDS) and of row v in Q (buffer)
};

// One worker (a slot of G lanes) running its windows of task T.  OFF32: the
// factors are at the wave-uniform bases FP, FQ (row scales FPs, FQs), under
// 4 GB each, and rows are addressed as base plus a 32-bit offset; otherwise
// at T's, with 64-bit offsets.  PT: the task's P rows [T.row_begin,
// T.row_begin + T.row_count) are in LDS at Pt (row-major, the workgroup's
// tile), and P rows are read and written there; every record of the task
// has its row in that range.
template <int K, int G, int PREC, bool COH, int D, bool OFF32 = false, bool PT = false>
__device__ __forceinline__ void sgd_worker(const SgdLaunch& a, const SgdTask& T, uint32_t lw, int slot, int ls,
                                           int32_t* su, int32_t* sv, float* sr, void* FP = nullptr,
                                           void* FQ = nullptr, const float* FPs = nullptr,
                                           const float* FQs = nullptr, lchar* Pt = nullptr)
{
    constexpr int E = K / G;
    constexpr int EB = PrecBytes<PREC>::value;
    constexpr int TB = E * EB;
    constexpr int NW = (TB + 3) / 4;
    constexpr bool FP8 = PREC == kE4M3;
    constexpr bool SR = FP8 || PREC == kBF16SR;   // stochastic rounding on the write-back
    constexpr bool PAIR = (PREC == kF16 || PREC == kBF16) && E % 2 == 0;   // packed two-element write-back
    constexpr bool MIX = PAIR && PREC == kF16;   // FP16 operands read in place (v_fma_mix_f32)
    constexpr int RPL = (kMaxBatch + G - 1) / G;   // staged records per lane
    static_assert(K % G == 0 && 64 % G == 0, "lane group must divide k and 64");
    static_assert(!FP8 || E % 4 == 0, "FP8 rows need 4 elements per lane");

    if (T.nnz <= 0 || T.update_count <= 0) return;
    const Key key = make_key(a.seed, T.problem);
    // window length: a tile's own, else the launch's (host guarantees
    // 1 <= L <= kMaxBatch, L % D == 0, and a tile's L at most the launch's)
    const int L = PT ? T.batch_len : a.batch_len;

    gchar* Pb = to_global((char*)T.P) + (int64_t)ls * (TB == 32 ? 16 : TB);
    gchar* Qb = to_global((char*)T.Q) + (int64_t)ls * (TB == 32 ? 16 : TB);
    gchar* Pl = to_global((char*)FP);
    gchar* Ql = to_global((char*)FQ);
    gcint* Ru = to_global(T.u);
    gcint* Rv = to_global(T.v);
    gcfloat* Rr = to_global(T.r);
    gcfloat* Ps = to_global((const float*)(OFF32 ? FPs : T.ps));
    gcfloat* Qs = to_global((const float*)(OFF32 ? FQs : T.qs));
    gcfloat* LR = to_global(T.lr);
    constexpr int64_t kRow = (int64_t)K * EB;
    // Lane l's elements: [lE, lE + E) for chunks up to 16 bytes; a 32-byte
    // chunk is two pieces of E/2 elements, [l E/2, (l+1) E/2) in each half of
    // the row, so that every 16-byte load or store instruction of the lane
    // group covers one half row contiguously (as many 64-byte L2 requests as
    // the half row needs, not twice as many).  Element i of the lane is
    // piece i / (E/2), element i % (E/2) of it.
    constexpr int SPLIT = (int)(kRow / 2);
    const uint32_t lane_off = (uint32_t)(ls * (TB == 32 ? 16 : TB));
    // the lane's part of row u of P / v of Q
    auto prow = [&](int u) -> gchar* {
        if constexpr (OFF32) return Pl + ((uint32_t)u * (uint32_t)kRow + lane_off);
        else return Pb + (int64_t)u * kRow;
    };
    auto qrow = [&](int v) -> gchar* {
        if constexpr (OFF32) return Ql + ((uint32_t)v * (uint32_t)kRow + lane_off);
        else return Qb + (int64_t)v * kRow;
    };
    const int row0 = T.row_begin;
    // Factors under 4 GB (OFF32): rows go through buffer resources, and a
    // load or store that is skipped (a row kept in registers) is issued with
    // an out-of-range offset instead of being branched around, so that the
    // wait for a sample's rows counts the same loads and stores in flight on
    // every path and stays partial.
    constexpr bool BUF = OFF32 && !COH;
    const auto Pr = factor_rsrc(FP);
    const auto Qr = factor_rsrc(FQ);
    auto boff = [&](int id) -> uint32_t { return (uint32_t)id * (uint32_t)kRow + lane_off; };
    const int64_t nnz = T.nnz;

    // The records of one window go from memory straight into LDS (global_load_lds,
    // no registers): su/sv/sr are the wave's arrays, and load c of lane
    // (slot, ls) fetches record j = ls + G*c of its slot into word 64*c + lane,
    // so record j of a slot sits at (j / G) * 64 + slot * G + j % G.
    // Tile launches load the records of a window into registers (kTileRows
    // per array and lane, every lane active: a lane past the window's end
    // loads the window's first record into a slot that is never read) and
    // write them to LDS when the window starts.  Ordinary loads, unlike
    // loads into LDS, let the compiler count the vector memory operations in
    // flight across them, so the waits for a sample's rows stay partial
    // while the next window's records load during the last round.
    constexpr int kTileRows = PT ? (kSgdTileMaxBatch + G - 1) / G : 1;
    int32_t nu[kTileRows], nv[kTileRows];
    float nr[kTileRows];
    auto fetch_window = [&](int ep, int b) {
        const int64_t start = window_start(key, T.tag, lw, (uint32_t)ep, (uint32_t)(T.window_begin + b), nnz);
        if constexpr (PT) {
#pragma unroll
            for (int c = 0; c < kTileRows; ++c) {
                const int j = ls + G * c;
                int64_t idx = start + (j < L ? j : 0);
                while (idx >= nnz) idx -= nnz;
                nu[c] = Ru[idx];
                nv[c] = Rv[idx];
                nr[c] = Rr[idx];
            }
            return;
        }
        // reads of the previous window's records have returned before they
        // are overwritten
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
#pragma unroll
        for (int c = 0; c < RPL; ++c) {
            const int j = ls + G * c;
            if (j < L) {
                int64_t idx = start + j;
                while (idx >= nnz) idx -= nnz;
                __builtin_amdgcn_global_load_lds(Ru + idx, to_lds(su + 64 * c), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(Rv + idx, to_lds(sv + 64 * c), 4, 0, 0);
                __builtin_amdgcn_global_load_lds(Rr + idx, to_lds(sr + 64 * c), 4, 0, 0);
            }
        }
    };
    auto stage_window = [&]() {
        if constexpr (PT) {
            // the wave's LDS operations run in order: the previous window's
            // reads of these words come before the writes
            const int lane = slot * G + ls;
#pragma unroll
            for (int c = 0; c < kTileRows; ++c) {
                su[64 * c + lane] = nu[c];
                sv[64 * c + lane] = nv[c];
                sr[64 * c + lane] = nr[c];
            }
        } else {
            __builtin_amdgcn_s_waitcnt(0);   // the window's records are in LDS
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    const int rbase = slot * G;
    // Request sample j's rows.  A row of the sample before it (prev_u,
    // prev_v: -1 at a window's start) is not loaded: that sample's written
    // row is forwarded from the register ring instead (below).  P rows in
    // LDS (PT) are read every sample; the ring covers one written since.
    auto request = [&](Pending<NW>& pd, int j, int prev_u, int prev_v) {
        const int at = (j / G) * 64 + rbase + j % G;
        pd.u = su[at];
        pd.v = sv[at];
        pd.r = sr[at];
        // tile launches keep the row offsets for the write-back
        if constexpr (PT) {
            pd.pa = (uint32_t)(pd.u - row0) * (uint32_t)kRow + lane_off;
            pd.p = load_lds<TB, SPLIT>(Pt + pd.pa);
        } else if constexpr (BUF) {
            pd.p = load_buf<TB, SPLIT>(Pr, pd.u != prev_u ? boff(pd.u) : kBufSkip);
        } else if (pd.u != prev_u) {
            pd.p = load_raw<TB, COH, SPLIT>(prow(pd.u));
        }
        if constexpr (PT && BUF) {
            pd.qa = boff(pd.v);
            pd.q = load_buf<TB, SPLIT>(Qr, pd.v != prev_v ? pd.qa : kBufSkip);
        } else if constexpr (BUF) {
            pd.q = load_buf<TB, SPLIT>(Qr, pd.v != prev_v ? boff(pd.v) : kBufSkip);
        } else if (pd.v != prev_v) {
            pd.q = load_raw<TB, COH, SPLIT>(qrow(pd.v));
        }
        if constexpr (FP8) {
            pd.ps = Ps[pd.u];
            pd.qs = Qs[pd.v];
        }
    };

    const int ep0 = a.epoch_dev ? *to_global(a.epoch_dev) : a.epoch_begin;
    const int nwin = (a.epoch_end - a.epoch_begin) * T.update_count;
    if (nwin <= 0) return;
    // the task's pointers are in registers before the first window's loads
    // (else each guarded load below waits for them again)
    __builtin_amdgcn_s_waitcnt(0);
    fetch_window(ep0, 0);
    // tile launches: the first window's records are in registers before the
    // window loop, so that the wait for a later window's records counts only
    // what the loop issued after them (the last round's stores stay in flight)
    if constexpr (PT) __builtin_amdgcn_s_waitcnt(0);

    float lr = 0.0f, cp = 0.0f, cq = 0.0f;
    for (int w = 0; w < nwin; ++w) {
        const int ep = ep0 + w / T.update_count;
        const int b = w % T.update_count;
        // tile launches read the rate when the epoch changes (no load that
        // the window's first wait would drain)
        if (!PT || b == 0) {
            lr = LR[ep];
            // p + lr (e q - lp p) = p (1 - lr lp) + (lr e) q: one multiply
            // and one fma per element
            cp = __builtin_fmaf(-lr, T.lambda_p, 1.0f);
            cq = __builtin_fmaf(-lr, T.lambda_q, 1.0f);
        }
        U4 srk = U4{0u, 0u, 0u, 0u};
        if constexpr (SR)
            srk = philox(U4{(uint32_t)(T.window_begin + b), (uint32_t)ep, lw, kStreamRound | (T.tag << 8)}, key.k0,
                         key.k1);
        stage_window();
        const bool more_windows = w + 1 < nwin;
        auto fetch_next = [&]() {
            if constexpr (PT) {
                // every window issues the same loads: after the last one a
                // valid window that is never read
                const int wn = more_windows ? w + 1 : 0;
                fetch_window(ep0 + wn / T.update_count, wn % T.update_count);
            } else if (more_windows) {
                fetch_window(ep0 + (w + 1) / T.update_count, (w + 1) % T.update_count);
            }
        };

        Pending<NW> pend[D];
        // ring of the last D written rows (slot d holds the sample j with j % D == d)
        Words<NW> wp[D], wq[D];
        int wu[D], wv[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            wu[d] = -1;
            wv[d] = -1;
            request(pend[d], d, d ? pend[d - 1].u : -1, d ? pend[d - 1].v : -1);
            // keep the requests in consumption order, so the wait for the
            // oldest one is a partial wait (the loop header merges this state)
            __builtin_amdgcn_sched_barrier(0);
        }

        // D samples per round.  The first and the last round are peeled: the
        // loop then runs in one steady state (every sample stores its rows and
        // requests those of sample j + D), so that the compiler's wait for a
        // sample's rows stays partial, with the D - 1 younger requests and
        // their stores still in flight.
        auto round = [&](int base, auto more) __attribute__((always_inline)) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int j = base + d;
                const int uc = pend[d].u, vc = pend[d].v;
                const float rc = pend[d].r;
                float psc = 1.0f, qsc = 1.0f;
                if constexpr (FP8) { psc = pend[d].ps; qsc = pend[d].qs; }
                // forward rows written after this sample's rows were requested
                // (samples j-D+1 .. j-1), oldest first so the newest wins; the
                // wave skips this when none of its slots has such a row.  P
                // rows in LDS (PT) are read again instead: the wave's LDS
                // operations run in order, so the read returns every write
                // before it (records ordered by column repeat P rows seldom:
                // the read stays a rare branch, and no P ring is kept)
                if constexpr (PT) {
                    bool hit_p = false, hit_q = false;
#pragma unroll
                    for (int back = D - 1; back >= 1; --back) {
                        const int s = (d - back + D) % D;
                        hit_p |= wu[s] == uc;
                        hit_q |= wv[s] == vc;
                    }
                    if (__builtin_amdgcn_ballot_w64(hit_p)) pend[d].p = load_lds<TB, SPLIT>(Pt + pend[d].pa);
                    if (__builtin_amdgcn_ballot_w64(hit_q)) {
#pragma unroll
                        for (int back = D - 1; back >= 1; --back) {
                            const int s = (d - back + D) % D;
                            if (wv[s] == vc) pend[d].q = wq[s];
                        }
                    }
                } else {
                    bool hit = false;
#pragma unroll
                    for (int back = D - 1; back >= 1; --back) {
                        const int s = (d - back + D) % D;
                        hit |= (wu[s] == uc) | (wv[s] == vc);
                    }
                    if (__builtin_amdgcn_ballot_w64(hit)) {
#pragma unroll
                        for (int back = D - 1; back >= 1; --back) {
                            const int s = (d - back + D) % D;
                            if (wu[s] == uc) pend[d].p = wp[s];
                            if (wv[s] == vc) pend[d].q = wq[s];
                        }
                    }
                }
                const Words<NW> pw = pend[d].p, qw = pend[d].q;

                // FP8 rows stay in their scaled units (row value = byte value
                // times the row's power-of-two scale): scaling by a power of
                // two commutes with every rounding below, so the dot product
                // takes the scales once and each update once, per sample
                float p[E], q[E];
                if constexpr (FP8) {
#pragma unroll
                    for (int i = 0; i < NW; ++i) {
                        const f32x2 pl = __builtin_amdgcn_cvt_pk_f32_fp8((int)pw.w[i], false);
                        const f32x2 ph = __builtin_amdgcn_cvt_pk_f32_fp8((int)pw.w[i], true);
                        const f32x2 ql = __builtin_amdgcn_cvt_pk_f32_fp8((int)qw.w[i], false);
                        const f32x2 qh = __builtin_amdgcn_cvt_pk_f32_fp8((int)qw.w[i], true);
                        p[4 * i] = pl.x; p[4 * i + 1] = pl.y; p[4 * i + 2] = ph.x; p[4 * i + 3] = ph.y;
                        q[4 * i] = ql.x; q[4 * i + 1] = ql.y; q[4 * i + 2] = qh.x; q[4 * i + 3] = qh.y;
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < E; ++e) {
                        p[e] = unpack<PREC>(pw, e);
                        q[e] = unpack<PREC>(qw, e);
                    }
                }
                float part;
                if constexpr (MIX) part = mul_f16_f16<false>(pw.w[0], qw.w[0]);
                else part = p[0] * q[0];
#pragma unroll
                for (int e = 1; e < E; ++e) part = __builtin_fmaf(p[e], q[e], part);
                float dot = slot_sum<G>(part);
                if constexpr (FP8) dot *= psc * qsc;
                const float err = rc - dot;
                const float lerr = lr * err;
                // multiplier of q in p' and of p in q' (FP8: in the rows' units)
                float lq_p = lerr, lp_q = lerr;
                if constexpr (FP8) {
                    // the scales are normal powers of two: v_rcp_f32 is exact
                    lq_p = lerr * (qsc * __builtin_amdgcn_rcpf(psc));
                    lp_q = lerr * (psc * __builtin_amdgcn_rcpf(qsc));
                }

                Words<NW> po, qo;
                if constexpr (PAIR) {
                    // two elements per word: FP16 rows are read as halves by
                    // v_fma_mix_f32; one packed conversion writes both back
#pragma unroll
                    for (int i = 0; i < NW; ++i) {
                        float p0, p1, q0, q1;
                        if constexpr (MIX) {
                            p0 = mul_f16_f32<false>(pw.w[i], cp);
                            p1 = mul_f16_f32<true>(pw.w[i], cp);
                            q0 = mul_f16_f32<false>(qw.w[i], cq);
                            q1 = mul_f16_f32<true>(qw.w[i], cq);
                        } else {
                            p0 = p[2 * i] * cp;
                            p1 = p[2 * i + 1] * cp;
                            q0 = q[2 * i] * cq;
                            q1 = q[2 * i + 1] * cq;
                        }
                        p0 = __builtin_fmaf(lerr, q[2 * i], p0);
                        p1 = __builtin_fmaf(lerr, q[2 * i + 1], p1);
                        q0 = __builtin_fmaf(lerr, p[2 * i], q0);
                        q1 = __builtin_fmaf(lerr, p[2 * i + 1], q1);
                        if constexpr (MIX) {
                            po.w[i] = f32x2_to_f16x2(p0, p1);
                            qo.w[i] = f32x2_to_f16x2(q0, q1);
                        } else {
                            po.w[i] = f32x2_to_bf16x2(p0, p1);
                            qo.w[i] = f32x2_to_bf16x2(q0, q1);
                        }
                    }
                } else if constexpr (FP8) {
                    // Element e of the lane's P row rounds with the 16-bit half
                    // (e & 1) of word e >> 1 of the lane's random words
                    // (sr_lane_seed, sr_next), element e of its Q row with the
                    // other half, each as the top half of the instruction's
                    // word; values clamp to +-448 (one v_med3_f32).
                    uint32_t w = sr_lane_seed(srk.x, (uint32_t)j, (uint32_t)ls);
#pragma unroll
                    for (int c = 0; c < NW; ++c) {
                        uint32_t op = 0u, oq = 0u;
#pragma unroll
                        for (int b = 0; b < 4; ++b) {
                            const int e = 4 * c + b;
                            if (e >= 2 && (e & 1) == 0) w = sr_next(w);
                            const float pn = __builtin_fmaf(lq_p, q[e], p[e] * cp);
                            const float qn = __builtin_fmaf(lp_q, p[e], q[e] * cq);
                            const uint32_t lo = w << 16, hi = w & 0xffff0000u;
                            const uint32_t rp = (b & 1) ? hi : lo;
                            const uint32_t rq = (b & 1) ? lo : hi;
                            op = e4m3_sr_clamped(pn, rp, op, b);
                            oq = e4m3_sr_clamped(qn, rq, oq, b);
                        }
                        po.w[c] = op;
                        qo.w[c] = oq;
                    }
                    // A NaN can only come from a non-finite multiplier (the rows
                    // and the dot product are then finite); such a sample writes
                    // NaN (0xff) to every element of both rows.
                    if (__builtin_amdgcn_ballot_w64(!(__builtin_isfinite(lq_p) && __builtin_isfinite(lp_q)))) {
                        if (!(__builtin_isfinite(lq_p) && __builtin_isfinite(lp_q))) {
#pragma unroll
                            for (int c = 0; c < NW; ++c) po.w[c] = qo.w[c] = 0xffffffffu;
                        }
                    }
                } else {
                uint32_t wr = 0u;   // BF16-SR: the lane's random words, shared by P and Q as for FP8
                if constexpr (SR) wr = sr_lane_seed(srk.x, (uint32_t)j, (uint32_t)ls);
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const float pn = __builtin_fmaf(lq_p, q[e], p[e] * cp);
                    const float qn = __builtin_fmaf(lp_q, p[e], q[e] * cq);
                    uint32_t pb = 0u, qb = 0u;
                    if constexpr (PREC == kF32) {
                        pb = __float_as_uint(pn);
                        qb = __float_as_uint(qn);
                    } else if constexpr (PREC == kF16) {
                        pb = f32_to_f16(pn);
                        qb = f32_to_f16(qn);
                    } else if constexpr (PREC == kBF16) {
                        pb = f32_to_bf16(pn);
                        qb = f32_to_bf16(qn);
                    } else {
                        // element e of the lane's P row takes the 16-bit half
                        // (e & 1) of the lane's random word e >> 1, element e of
                        // its Q row the other half
                        if (e >= 2 && (e & 1) == 0) wr = sr_next(wr);
                        if constexpr (PREC == kBF16SR && E % 2 == 0) {
                            // the instruction writes half (e & 1) of the word directly,
                            // and reads only the top half of its random word, so the
                            // word itself stands for its own top half
                            const uint32_t lo = wr << 16, hi = wr;
                            const uint32_t rp = (e & 1) ? hi : lo;
                            const uint32_t rq = (e & 1) ? lo : hi;
                            const uint32_t op = (e & 1) ? po.w[e >> 1] : 0u, oq = (e & 1) ? qo.w[e >> 1] : 0u;
                            if (e & 1) {
                                po.w[e >> 1] = bf16_sr_insert<true>(pn, rp, op);
                                qo.w[e >> 1] = bf16_sr_insert<true>(qn, rq, oq);
                            } else {
                                po.w[e >> 1] = bf16_sr_insert<false>(pn, rp, op);
                                qo.w[e >> 1] = bf16_sr_insert<false>(qn, rq, oq);
                            }
                        } else {
                            static_assert(PREC == kBF16SR, "FP8 has its own write-back above");
                            pb = f32_to_bf16_sr(pn, (e & 1) ? (wr >> 16) : wr);
                            qb = f32_to_bf16_sr(qn, (e & 1) ? wr : (wr >> 16));
                        }
                    }
                    if constexpr (!(PREC == kBF16SR && E % 2 == 0)) {
                        pack<PREC>(po, e, pb);
                        pack<PREC>(qo, e, qb);
                    }
                }
                }
                // a row that the next sample of this worker (in pend[(d + 1) % D],
                // requested already) rewrites is not stored: that sample takes
                // it from the ring and stores its own result (P rows in LDS
                // are stored every sample)
                const bool has_next = decltype(more)::value || d + 1 < D;
                const Pending<NW>& nx = pend[(d + 1) % D];
                const bool keep_p = has_next && nx.u == uc, keep_q = has_next && nx.v == vc;
                if constexpr (PT) store_lds<TB, SPLIT>(Pt + pend[d].pa, po);
                else if constexpr (BUF) store_buf<TB, SPLIT>(Pr, keep_p ? kBufSkip : boff(uc), po);
                else if (!keep_p) store_raw<TB, COH, SPLIT>(prow(uc), po);
                if constexpr (PT && BUF) store_buf<TB, SPLIT>(Qr, keep_q ? kBufSkip : pend[d].qa, qo);
                else if constexpr (BUF) store_buf<TB, SPLIT>(Qr, keep_q ? kBufSkip : boff(vc), qo);
                else if (!keep_q) store_raw<TB, COH, SPLIT>(qrow(vc), qo);
                if constexpr (!PT) wp[d] = po;
                wq[d] = qo;
                wu[d] = uc;
                wv[d] = vc;
                if constexpr (decltype(more)::value) request(pend[d], j + D, pend[(d + D - 1) % D].u, pend[(d + D - 1) % D].v);
            }
        };
        if (L > D) {
            round(0, std::true_type{});
            for (int base = D; base < L - D; base += D) round(base, std::true_type{});
        }
        // every record of the window has been read: the next window's
        // records load during the last round
        fetch_next();
        round(L - D, std::false_type{});
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

template <int K, int G, int PREC, bool COH, int D, bool OFF32>
__global__ __launch_bounds__(kWavesPerBlock * 64) void sgd_update_kernel(SgdLaunch a)
{
    constexpr int S = 64 / G;
    __shared__ int32_t s_u[kWavesPerBlock][S * kMaxBatch];
    __shared__ int32_t s_v[kWavesPerBlock][S * kMaxBatch];
    __shared__ float s_r[kWavesPerBlock][S * kMaxBatch];

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int slot = lane / G;
    const int ls = lane % G;

    int64_t blk = blockIdx.x;
    if (a.num_xcd > 1) blk = (int64_t)(blockIdx.x % a.num_xcd) * (gridDim.x / a.num_xcd) + blockIdx.x / a.num_xcd;
    const int worker = (int)((blk * kWavesPerBlock + wave) * S + slot);

    int t = 0;
    while (t < a.num_tasks && worker >= a.tasks[t].worker_begin + a.tasks[t].worker_count) ++t;
    if (t >= a.num_tasks || worker < a.tasks[t].worker_begin) return;
    const SgdTask T = a.tasks[t];
    sgd_worker<K, G, PREC, COH, D, OFF32>(a, T, (uint32_t)(worker - T.worker_begin), slot, ls, s_u[wave], s_v[wave],
                                          s_r[wave], a.P, a.Q, a.ps, a.qs);
}

// Tile launch: one workgroup per task, the task a tile (the ratings of the
// P rows [row_begin, row_begin + row_count) within one column range, no P
// row shared with another task of the launch).  The workgroup copies its P
// rows into LDS, its workers train the tile there (Q rows in memory, shared
// through the XCD's L2 with the other tiles of the same column range), and
// the rows go back to memory when every worker is done.  Per update, the L2
// then sees the Q row (when the previous sample's differs) and the record,
// instead of a P row read and written as well.
// Dynamic LDS: the tile (row_count * K * bytes); static: per wave the
// staged records of a window (64 * ceil(kSgdTileMaxBatch / G) words per
// array).
template <int K, int G, int PREC, int D, bool OFF32, int NWAVES>
__global__ __launch_bounds__(NWAVES * 64) void sgd_tile_kernel(SgdLaunch a)
{
    extern __shared__ __attribute__((aligned(16))) char lds_dyn[];
    constexpr int S = 64 / G;
    constexpr int64_t kRow = (int64_t)K * PrecBytes<PREC>::value;
    static_assert(kRow % 16 == 0, "rows are copied in 16-byte words");
    constexpr int kThreads = NWAVES * 64;

    // workgroups go round-robin over the XCDs: task x * per + j is the j-th
    // workgroup dispatched to XCD x
    const int per = (int)gridDim.x / a.num_xcd;
    const int ti = (int)(blockIdx.x % a.num_xcd) * per + (int)(blockIdx.x / a.num_xcd);
    if (ti >= a.num_tasks) return;
    const SgdTask T = a.tasks[ti];
    if (T.nnz <= 0 || T.row_count <= 0) return;

    // the staged records of each wave (a static array beside the dynamic
    // tile; kSgdTileMaxBatch records per worker)
    constexpr int kStageWords = 64 * ((kSgdTileMaxBatch + G - 1) / G);
    __shared__ int32_t s_stage[NWAVES][3][kStageWords];
    lchar* tile = (lchar*)lds_dyn;
    const int64_t bytes = (int64_t)T.row_count * kRow;
    const gchar* src = to_global((const char*)T.P) + (int64_t)T.row_begin * kRow;
    for (int64_t i = (int64_t)threadIdx.x * 16; i < bytes; i += kThreads * 16)
        *(__attribute__((address_space(3))) u32x4*)(tile + i) = *(const __attribute__((address_space(1))) u32x4*)(src + i);
    __syncthreads();

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int slot = lane / G;
    const int ls = lane % G;
    const int worker = wave * S + slot;
    if (worker < T.worker_count)
        sgd_worker<K, G, PREC, false, D, OFF32, true>(a, T, (uint32_t)worker, slot, ls, s_stage[wave][0],
                                                      s_stage[wave][1], (float*)s_stage[wave][2], a.P, a.Q, a.ps,
                                                      a.qs, tile);
    __syncthreads();

    gchar* dst = to_global((char*)T.P) + (int64_t)T.row_begin * kRow;
    for (int64_t i = (int64_t)threadIdx.x * 16; i < bytes; i += kThreads * 16)
        *(__attribute__((address_space(1))) u32x4*)(dst + i) = *(const __attribute__((address_space(3))) u32x4*)(tile + i);
}

template <int K, int G, int PREC, int D, int NWAVES>
inline hipError_t launch_tile_t(const SgdLaunch& a, int blocks, size_t lds, hipStream_t s)
{
    // dynamic LDS above the default limit must be allowed per kernel
    if (a.P) {
        hipError_t e = hipFuncSetAttribute((const void*)sgd_tile_kernel<K, G, PREC, D, true, NWAVES>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((sgd_tile_kernel<K, G, PREC, D, true, NWAVES>), dim3(blocks), dim3(NWAVES * 64), lds, s, a);
    } else {
        hipError_t e = hipFuncSetAttribute((const void*)sgd_tile_kernel<K, G, PREC, D, false, NWAVES>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((sgd_tile_kernel<K, G, PREC, D, false, NWAVES>), dim3(blocks), dim3(NWAVES * 64), lds, s,
                           a);
    }
    return hipGetLastError();
}

// Tile launches use the default lane group (16 bytes of a row per lane, at
// least 8 lanes) or half of it (32 bytes per lane), depth 2 and workgroups
// of 4, 8 or 16 waves.
template <int K, int PREC, int G>
inline hipError_t launch_tile_w(const SgdLaunch& a, int blocks, int waves, size_t lds, hipStream_t s)
{
    constexpr int EB = PrecBytes<PREC>::value;
    if constexpr (G >= 8 && K % G == 0 && (K / G) * EB <= 32 && (K / G) * EB >= 2 &&
                  (PREC != kE4M3 || (K / G) % 4 == 0)) {
        switch (waves) {
        case 4: return launch_tile_t<K, G, PREC, 2, 4>(a, blocks, lds, s);
        case 8: return launch_tile_t<K, G, PREC, 2, 8>(a, blocks, lds, s);
        case 16: return launch_tile_t<K, G, PREC, 2, 16>(a, blocks, lds, s);
        default: return hipErrorInvalidValue;
        }
    }
    return hipErrorInvalidValue;
}

template <int K, int PREC>
inline hipError_t launch_tile_k(const SgdLaunch& a, int blocks, int G, int D, int waves, size_t lds, hipStream_t s)
{
    constexpr int EB = PrecBytes<PREC>::value;
    constexpr int G0 = (K * EB / 16) > 64 ? 64 : ((K * EB / 16) < 8 ? 8 : K * EB / 16);
    if (D != 2) return hipErrorInvalidValue;
    if (G == G0) return launch_tile_w<K, PREC, G0>(a, blocks, waves, lds, s);
    if (G == G0 / 2) return launch_tile_w<K, PREC, G0 / 2>(a, blocks, waves, lds, s);
    return hipErrorInvalidValue;
}

// A pointer every lane of the wave holds (tasks own whole waves in batched
// launches), moved to scalar registers so that rows can be addressed as
// scalar base plus vector offset.
template <typename T>
__device__ inline T* wave_uniform(T* p)
{
    const uint64_t x = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return (T*)(((uint64_t)hi << 32) | lo);
}

// Lane group of a factor length in batched launches (one per k): 16 bytes
// of a row per lane, 8 to 64 lanes, as in single-problem launches.
template <int PREC>
__host__ __device__ constexpr int batch_group(int k)
{
    return k * PrecBytes<PREC>::value / 16 > 64 ? 64 : (k * PrecBytes<PREC>::value / 16 < 8 ? 8 : k * PrecBytes<PREC>::value / 16);
}
// The batch's worker budget is counted in waves of k/2 lanes per worker
// (k/4 for FP8), the unit its waves-per-CU defaults and autotuner were
// measured in: a problem keeps that many workers, packed 16 bytes of a row
// per lane into fewer waves.
template <int PREC>
__host__ __device__ constexpr int batch_budget_group(int k)
{
    return PREC == kE4M3 ? k / 4 : (k / 2 > 64 ? 64 : k / 2);
}

// Many independent problems in one launch.  Tasks own whole waves: task t
// holds waves [worker_begin, worker_begin + wave_count) (worker_begin and
// wave_count counted in waves here, tasks sorted by worker_begin, gaps
// allowed) and its first worker_count workers are active; each wave splits
// into 64/G slots for the task's k.
template <int PREC, bool COH, int D, bool OFF32>
__global__ __launch_bounds__(kWavesPerBlock * 64) void sgd_batched_kernel(SgdLaunch a)
{
    constexpr int SMAX = 64 / batch_group<PREC>(32);
    __shared__ int32_t s_u[kWavesPerBlock][SMAX * kMaxBatch];
    __shared__ int32_t s_v[kWavesPerBlock][SMAX * kMaxBatch];
    __shared__ float s_r[kWavesPerBlock][SMAX * kMaxBatch];

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    // waves numbered XCD-major (workgroups go round-robin over the XCDs), so
    // the host can keep each problem's waves on one XCD and its L2
    int blk = blockIdx.x;
    if (a.num_xcd > 1) blk = (int)(blockIdx.x % a.num_xcd) * (int)(gridDim.x / a.num_xcd) + (int)(blockIdx.x / a.num_xcd);
    const int gw = blk * kWavesPerBlock + wave;
    int t = 0;
    while (t < a.num_tasks && gw >= a.tasks[t].worker_begin + a.tasks[t].wave_count) ++t;
    if (t >= a.num_tasks || gw < a.tasks[t].worker_begin) return;
    const SgdTask T = a.tasks[t];
    const int wl = gw - T.worker_begin;
#define CUMF_BATCH_K(KK)                                                                                 \
    case KK: {                                                                                           \
        constexpr int G = batch_group<PREC>(KK);                                                         \
        const int slot = lane / G, ls = lane % G;                                                        \
        const uint32_t lw = (uint32_t)(wl * (64 / G) + slot);                                            \
        if (lw >= (uint32_t)T.worker_count) return;                                                      \
        sgd_worker<KK, G, PREC, COH, D, OFF32>(a, T, lw, slot, ls, s_u[wave], s_v[wave], s_r[wave],      \
                                               wave_uniform(T.P), wave_uniform(T.Q),                     \
                                               wave_uniform(T.ps), wave_uniform(T.qs));                  \
        return;                                                                                          \
    }
    switch (T.k) {
        CUMF_BATCH_K(32)
        CUMF_BATCH_K(64)
        CUMF_BATCH_K(128)
    default:
        return;
    }
#undef CUMF_BATCH_K
}

template <int PREC, bool COH, int D>
inline hipError_t launch_batched_t(const SgdLaunch& a, int blocks, hipStream_t s)
{
    if (a.task_off32)
        hipLaunchKernelGGL((sgd_batched_kernel<PREC, COH, D, true>), dim3(blocks), dim3(kWavesPerBlock * 64), 0, s, a);
    else
        hipLaunchKernelGGL((sgd_batched_kernel<PREC, COH, D, false>), dim3(blocks), dim3(kWavesPerBlock * 64), 0, s, a);
    return hipGetLastError();
}

template <int PREC>
inline hipError_t launch_batched_prec(const SgdLaunch& a, int blocks, int D, hipStream_t s)
{
    switch (D) {
    case 2: return launch_batched_t<PREC, false, 2>(a, blocks, s);
    case 4: return launch_batched_t<PREC, false, 4>(a, blocks, s);
    case 8: return launch_batched_t<PREC, false, 8>(a, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

template <int K, int G, int PREC, bool COH, int D>
inline hipError_t launch_sgd_t(const SgdLaunch& a, int blocks, hipStream_t s)
{
    if (a.P)
        hipLaunchKernelGGL((sgd_update_kernel<K, G, PREC, COH, D, true>), dim3(blocks), dim3(kWavesPerBlock * 64), 0,
                           s, a);
    else
        hipLaunchKernelGGL((sgd_update_kernel<K, G, PREC, COH, D, false>), dim3(blocks), dim3(kWavesPerBlock * 64), 0,
                           s, a);
    return hipGetLastError();
}

template <int K, int G, int PREC, bool COH>
inline hipError_t launch_sgd_d(const SgdLaunch& a, int blocks, int D, hipStream_t s)
{
    switch (D) {
    case 2: return launch_sgd_t<K, G, PREC, COH, 2>(a, blocks, s);
    case 4: return launch_sgd_t<K, G, PREC, COH, 4>(a, blocks, s);
    case 8: return launch_sgd_t<K, G, PREC, COH, 8>(a, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

template <int K, int G, int PREC>
inline hipError_t launch_sgd_c(const SgdLaunch& a, int blocks, bool coh, int D, hipStream_t s)
{
    constexpr int TB = (K / G) * PrecBytes<PREC>::value;
    if constexpr (TB == 2) {
        if (coh) return hipErrorInvalidValue;
        return launch_sgd_d<K, G, PREC, false>(a, blocks, D, s);
    } else {
        return coh ? launch_sgd_d<K, G, PREC, true>(a, blocks, D, s) : launch_sgd_d<K, G, PREC, false>(a, blocks, D, s);
    }
}

template <int K, int PREC>
inline hipError_t launch_sgd_g(const SgdLaunch& a, int blocks, int G, bool coh, int D, hipStream_t s)
{
    constexpr int EB = PrecBytes<PREC>::value;
    // lane groups whose per-lane chunk is 2..32 bytes (4..32 for FP8)
    switch (G) {
#define CUMF_G(GG)                                                                                     \
    case GG:                                                                                           \
        if constexpr (K % GG == 0 && (K / GG) * EB <= 32 && (K / GG) * EB >= 2 &&                      \
                      (PREC != kE4M3 || (K / GG) % 4 == 0))                                            \
            return launch_sgd_c<K, GG, PREC>(a, blocks, coh, D, s);                                    \
        break;
        CUMF_G(64)
        CUMF_G(32)
        CUMF_G(16)
        CUMF_G(8)
#undef CUMF_G
    default:
        break;
    }
    return hipErrorInvalidValue;
}


// Launch for one factor format; defined in sgd_update_<format>.hip.
template <int PREC>
hipError_t launch_sgd_prec(const SgdLaunch& a, int blocks, int k, int G, bool coh, int D, hipStream_t s);
template <int PREC>
hipError_t launch_batched_for(const SgdLaunch& a, int blocks, int D, hipStream_t s);
template <int PREC>
hipError_t launch_tile_prec(const SgdLaunch& a, int blocks, int k, int G, int D, int waves, size_t lds,
                            hipStream_t s);

}  // namespace cumf
