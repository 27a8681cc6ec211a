/*
 * CPU oracle for the cuMF_SGD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * C++ host, the Python package) links, loads or calls this file.  Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark use it, and
 * only as the reference to check against.
 *
 * Each function restates one piece of the reference's algorithm (cited as
 * file:line under the reference tree) or one piece of this project's own
 * definitions that have no reference counterpart (Philox streams, the
 * lane-group dot-product order, FP8 row scaling).  The pins that check each
 * function independently of this file live in tests/test_oracle_*.py.
 *
 * Built with -ffp-contract=off so every fused multiply-add is an explicit
 * fmaf()/fma() call, exactly where the GPU kernels use one.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11).  Counter-based; replaces the reference's clock- */
/* seeded curand XORWOW states (singleGPU/sgd_kernel.cu:22-26).              */
/* ------------------------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

EXPORT void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                              uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream tags carried in the low byte of counter word 3.  Same values as the
 * kernels. */
#define STREAM_INIT_P 1u
#define STREAM_INIT_Q 2u
#define STREAM_WINDOW 3u
#define STREAM_ROUND  4u

/* Key of every stream: the 64-bit seed, with the problem index of a batched
 * launch folded into the high word (problem 0 = a single problem). */
static inline void key_of(uint64_t seed, uint32_t problem, uint32_t key[2])
{
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32) ^ problem;
}

/* Start of Hogwild window b of worker w in epoch t: a 64-bit draw mapped onto
 * [0, nnz) by a 64x64->128 multiply-high.  The reference draws
 * (long long)(curand_uniform()*nnz) % nnz in FP32
 * (singleGPU/sgd_k128_kernel_hogwild_warp32.h:41-48), which quantizes starts
 * once nnz >= 2^24; the integer draw keeps every record reachable. */
EXPORT int64_t orc_window_start(uint64_t seed, uint32_t problem, uint32_t tag, uint32_t worker,
                                uint32_t epoch, uint32_t batch, int64_t nnz)
{
    /* tag: rating-block id of a gridded schedule (0 for the whole matrix) */
    uint32_t ctr[4] = {batch, epoch, worker, STREAM_WINDOW | (tag << 8)};
    uint32_t key[2], out[4];
    key_of(seed, problem, key);
    orc_philox4x32_10(ctr, key, out);
    uint64_t x = ((uint64_t)out[1] << 32) | out[0];
    return (int64_t)(((unsigned __int128)x * (uint64_t)nnz) >> 64);
}

/* 32 random bits -> float in (0, 1], exact: ((x >> 8) + 1) * 2^-24.
 * Same support as curand_uniform, used by random_init (singleGPU/sgd.cu:512). */
static inline float u01_of(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 0x1.0p-24f;
}

/* ------------------------------------------------------------------------ */
/* Number formats.  FP16: IEEE binary16, round-to-nearest-even, as           */
/* __float2half/__half2float in singleGPU/sgd_k128_kernel_hogwild_warp32.h   */
/* :62-99.  BF16: bfloat16 RNE.  FP8: OCP E4M3 (e4m3fn, max 448, no inf).    */
/* ------------------------------------------------------------------------ */
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

EXPORT float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = (uint32_t)(h >> 15) << 31;
    uint32_t exp = (h >> 10) & 0x1f;
    uint32_t man = h & 0x3ff;
    if (exp == 0x1f)
        return bitsf(sign | 0x7f800000u | (man << 13));
    if (exp == 0) {
        /* subnormal: man * 2^-24, exact in float */
        float v = (float)man * 0x1.0p-24f;
        return sign ? -v : v;
    }
    return bitsf(sign | ((exp + 112u) << 23) | (man << 13));
}

EXPORT uint16_t orc_f32_to_f16(float f)
{
    uint32_t x = fbits(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u)                       /* inf or nan */
        return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u)                       /* rounds to >= 65520: inf */
        return (uint16_t)(sign | 0x7c00u);
    if (ax < 0x38800000u) {                      /* below 2^-14: subnormal half */
        /* value / 2^-24 rounded to nearest even integer; rintf uses RNE */
        float v = bitsf(ax) * 0x1.0p24f;
        uint32_t q = (uint32_t)rintf(v);
        return (uint16_t)(sign | q);             /* q == 0x400 is the min normal */
    }
    /* normal: rebias exponent, round 13 dropped bits to nearest even */
    uint32_t r = ax + 0xfffu + ((ax >> 13) & 1u);
    r -= (112u << 23);
    return (uint16_t)(sign | (r >> 13));
}

EXPORT float orc_bf16_to_f32(uint16_t b) { return bitsf((uint32_t)b << 16); }

EXPORT uint16_t orc_f32_to_bf16(float f)
{
    uint32_t x = fbits(f);
    if ((x & 0x7fffffffu) > 0x7f800000u)
        return (uint16_t)((x >> 16) | 0x40u);    /* quiet nan */
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}

/* BF16 with stochastic rounding by 16 random bits r16: adding r16 to the 16
 * bits that truncation drops carries into the kept part with probability
 * equal to their value / 2^16, so |f| rounds up to the next BF16 with that
 * probability (a carry out of the largest finite value gives inf).  NaN stays
 * a quiet NaN. */
EXPORT uint16_t orc_f32_to_bf16_sr(float f, uint32_t r16)
{
    uint32_t x = fbits(f);
    if ((x & 0x7fffffffu) > 0x7f800000u)
        return (uint16_t)((x >> 16) | 0x40u);
    return (uint16_t)((x + (r16 & 0xffffu)) >> 16);
}

EXPORT float orc_e4m3_to_f32(uint8_t b)
{
    uint32_t sign = b >> 7;
    uint32_t exp = (b >> 3) & 0xf;
    uint32_t man = b & 0x7;
    float v;
    if (exp == 0xf && man == 0x7)
        return bitsf(0x7fc00000u | (sign << 31));
    if (exp == 0)
        v = (float)man * 0x1.0p-9f;
    else
        v = bitsf(((exp + 120u) << 23) | (man << 20));
    return sign ? -v : v;
}

/* |a| already < 448-region handled by caller; encode a float that holds at
 * most 3 mantissa bits and lies in [2^-6, 448] as an E4M3 byte. */
static inline uint8_t e4m3_pack_normal(uint32_t ax)
{
    uint32_t e = (ax >> 23) - 120u;
    return (uint8_t)((e << 3) | ((ax >> 20) & 7u));
}

/* Round to nearest even, saturating to +-448 (finite), nan -> 0x7f. */
EXPORT uint8_t orc_f32_to_e4m3_rne(float f)
{
    uint32_t x = fbits(f);
    uint8_t sign = (uint8_t)((x >> 24) & 0x80u);
    uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint8_t)(sign | 0x7f);
    if (ax >= 0x43e00000u) return (uint8_t)(sign | 0x7e);   /* >= 448 */
    if (ax < 0x3c800000u) {                                 /* < 2^-6 */
        uint32_t q = (uint32_t)rintf(bitsf(ax) * 0x1.0p9f);  /* 0..8 */
        return (uint8_t)(sign | q);
    }
    uint32_t r = ax + 0x7ffffu + ((ax >> 20) & 1u);
    r &= 0xfff00000u;
    if (r > 0x43e00000u) r = 0x43e00000u;                   /* saturate */
    return (uint8_t)(sign | e4m3_pack_normal(r));
}

/* Stochastic rounding with 16 random bits `r16`: the rule of gfx950's
 * v_cvt_sr_fp8_f32 given the random word r16 << 16, after clamping |f| to
 * 448.  The dropped part of |f|, in units of the E4M3 quantum (2^(e-3) for
 * normal values, 2^-9 below 2^-6), keeps its top 20 fraction bits (frac20);
 * |f| rounds up when frac20 + (r16 << 4) >= 2^20, i.e. with probability equal
 * to the dropped fraction to 2^-16.  NaN gives 0xff.  (The rule was read off
 * the instruction by an in-bounds probe over 2.3M inputs; the committed
 * vectors tests/golden/e4m3_sr_hw.npz hold its outputs.) */
EXPORT uint8_t orc_f32_to_e4m3_sr(float f, uint32_t r16)
{
    uint32_t x = fbits(f);
    uint8_t sign = (uint8_t)((x >> 24) & 0x80u);
    uint32_t ax = x & 0x7fffffffu;
    r16 &= 0xffffu;
    if (ax > 0x7f800000u) return 0xffu;
    if (ax >= 0x43e00000u) return (uint8_t)(sign | 0x7e);
    if (ax < 0x3c800000u) {
        float t = bitsf(ax) * 0x1.0p9f;         /* exact, in [0, 8) */
        float lo = floorf(t);
        uint32_t frac20 = (uint32_t)floorf((t - lo) * 0x1.0p20f);
        uint32_t q = (uint32_t)lo;
        if (frac20 + (r16 << 4) >= (1u << 20)) q += 1u;
        return (uint8_t)(sign | q);
    }
    uint32_t r = ax + (r16 << 4);
    r &= 0xfff00000u;
    if (r > 0x43e00000u) r = 0x43e00000u;
    return (uint8_t)(sign | e4m3_pack_normal(r));
}

/* ------------------------------------------------------------------------ */
/* Factor storage.  prec: 0 FP32, 1 FP16, 2 BF16, 3 FP8-E4M3 + per-row scale, */
/* 4 BF16 with stochastic rounding on the update's write-back.              */
/* ------------------------------------------------------------------------ */
enum { PREC_F32 = 0, PREC_F16 = 1, PREC_BF16 = 2, PREC_E4M3 = 3, PREC_BF16_SR = 4 };

/* FP8 row scales are powers of two.  A row is (re)scaled so that its largest
 * magnitude maps to at most FP8_INIT_TARGET (at init) or FP8_EPOCH_TARGET
 * (between epochs), leaving headroom below 448 for growth within an epoch. */
#define FP8_INIT_TARGET  14.0f
#define FP8_EPOCH_TARGET 56.0f

static inline size_t prec_bytes(int prec) { return prec == PREC_F32 ? 4 : (prec == PREC_E4M3 ? 1 : 2); }

static inline float load_elem(const void* base, int prec, int64_t i)
{
    switch (prec) {
    case PREC_F32:  return ((const float*)base)[i];
    case PREC_F16:  return orc_f16_to_f32(((const uint16_t*)base)[i]);
    case PREC_BF16:
    case PREC_BF16_SR: return orc_bf16_to_f32(((const uint16_t*)base)[i]);
    default:        return orc_e4m3_to_f32(((const uint8_t*)base)[i]);
    }
}

/* Round x to the storage format and write it; returns the value that was
 * stored, widened back to FP32 (what the next read of that element sees). */
static inline float store_elem(void* base, int prec, int64_t i, float x, float inv_scale,
                               float scale, uint32_t rnd)
{
    switch (prec) {
    case PREC_F32:
        ((float*)base)[i] = x;
        return x;
    case PREC_F16: {
        uint16_t h = orc_f32_to_f16(x);
        ((uint16_t*)base)[i] = h;
        return orc_f16_to_f32(h);
    }
    case PREC_BF16: {
        uint16_t h = orc_f32_to_bf16(x);
        ((uint16_t*)base)[i] = h;
        return orc_bf16_to_f32(h);
    }
    case PREC_BF16_SR: {
        uint16_t h = orc_f32_to_bf16_sr(x, rnd);
        ((uint16_t*)base)[i] = h;
        return orc_bf16_to_f32(h);
    }
    default: {
        uint8_t b = orc_f32_to_e4m3_sr(x * inv_scale, rnd);
        ((uint8_t*)base)[i] = b;
        return orc_e4m3_to_f32(b) * scale;
    }
    }
}

/* Smallest power of two s with amax <= target * s (s = 1 when amax is 0 or
 * not finite).  With amax = fa*2^ea and target = ft*2^et (fa, ft in
 * [0.5, 1)), that is s = 2^(ea - et + (fa > ft)), exponent clamped to
 * [-126, 126]. */
EXPORT float orc_pow2_scale(float amax, float target)
{
    if (!(amax > 0.0f) || isinf(amax)) return 1.0f;
    int ea, et;
    float fa = frexpf(amax, &ea);
    float ft = frexpf(target, &et);
    int n = ea - et + (fa > ft ? 1 : 0);
    /* kept in [2^-126, 2^126], so that s and 1/s are normal */
    return ldexpf(1.0f, n < -126 ? -126 : (n > 126 ? 126 : n));
}

/* ------------------------------------------------------------------------ */
/* Factor initialisation.  Reference: every element of every row gets        */
/* half(U(0,1] * sqrt(1/k)) (singleGPU/sgd.cu:497-546, scale at :519).       */
/* Here element e of row r draws word (e & 3) of Philox(ctr = {e >> 2, r_lo, */
/* r_hi, stream}, key = seed).                                               */
/* ------------------------------------------------------------------------ */
EXPORT void orc_init_factors(void* out, float* row_scale, int64_t rows, int k, int prec,
                             uint64_t seed, uint32_t problem, int which /*0=P,1=Q*/)
{
    float s = (float)sqrt(1.0 / k);
    uint32_t key[2];
    key_of(seed, problem, key);
    uint32_t stream = (which == 0 ? STREAM_INIT_P : STREAM_INIT_Q);
    float* tmp = (float*)malloc(sizeof(float) * (size_t)k);
    for (int64_t r = 0; r < rows; ++r) {
        float amax = 0.0f;
        for (int e = 0; e < k; e += 4) {
            uint32_t ctr[4] = {(uint32_t)(e >> 2), (uint32_t)r, (uint32_t)((uint64_t)r >> 32), stream};
            uint32_t w[4];
            orc_philox4x32_10(ctr, key, w);
            for (int j = 0; j < 4 && e + j < k; ++j) {
                float v = u01_of(w[j]) * s;
                tmp[e + j] = v;
                if (v > amax) amax = v;
            }
        }
        float sc = 1.0f, inv = 1.0f;
        if (prec == PREC_E4M3) {
            sc = orc_pow2_scale(amax, FP8_INIT_TARGET);
            inv = 1.0f / sc;
            row_scale[r] = sc;
        }
        for (int e = 0; e < k; ++e) {
            int64_t i = r * (int64_t)k + e;
            switch (prec) {
            case PREC_F32:  ((float*)out)[i] = tmp[e]; break;
            case PREC_F16:  ((uint16_t*)out)[i] = orc_f32_to_f16(tmp[e]); break;
            case PREC_BF16:
            case PREC_BF16_SR: ((uint16_t*)out)[i] = orc_f32_to_bf16(tmp[e]); break;
            default:        ((uint8_t*)out)[i] = orc_f32_to_e4m3_rne(tmp[e] * inv); break;
            }
        }
    }
    free(tmp);
}

/* Between-epoch FP8 rescale: widen each row, pick a new power-of-two scale
 * for FP8_EPOCH_TARGET, requantize with RNE (exact when the scale shrinks). */
EXPORT void orc_fp8_rescale(uint8_t* rows_q, float* row_scale, int64_t rows, int k)
{
    for (int64_t r = 0; r < rows; ++r) {
        float amax = 0.0f;
        for (int e = 0; e < k; ++e) {
            float v = fabsf(orc_e4m3_to_f32(rows_q[r * k + e]) * row_scale[r]);
            if (v > amax) amax = v;
        }
        if (!(amax > 0.0f)) continue;
        float sc = orc_pow2_scale(amax, FP8_EPOCH_TARGET);
        float inv = 1.0f / sc;
        for (int e = 0; e < k; ++e) {
            float v = orc_e4m3_to_f32(rows_q[r * k + e]) * row_scale[r];
            rows_q[r * k + e] = orc_f32_to_e4m3_rne(v * inv);
        }
        row_scale[r] = sc;
    }
}

/* ------------------------------------------------------------------------ */
/* Learning-rate schedule: lr[i] = float(alpha/(1 + beta*i^1.5) + lrate),    */
/* evaluated in double (singleGPU/sgd_kernel.cu:137-149).                    */
/* ------------------------------------------------------------------------ */
EXPORT void orc_lr_table(float alpha, float beta, float lrate, int n, float* out)
{
    for (int i = 0; i < n; ++i) {
        double t = alpha / (1 + beta * pow(i, 1.5)) + lrate;
        out[i] = (float)t;
    }
}

/* ------------------------------------------------------------------------ */
/* Host preprocessing (singleGPU/sgd.cu).                                    */
/* ------------------------------------------------------------------------ */
typedef struct { int32_t u, v; float r; } orc_node;   /* mf_node, singleGPU/sgd.h:79-83 */

/* collect_data, singleGPU/sgd.cu:78-96: serial double sums in record order. */
EXPORT void orc_collect_data(const orc_node* R, int64_t nnz, float* ave, float* std_dev)
{
    double ex = 0, ex2 = 0;
    for (int64_t i = 0; i < nnz; ++i) {
        float r = R[i].r;
        ex += (double)r;
        ex2 += (double)r * r;
    }
    ex = ex / (double)nnz;
    ex2 = ex2 / (double)nnz;
    *ave = (float)ex;
    *std_dev = (float)sqrt(ex2 - ex * ex);
}

/* gen_random_map, singleGPU/sgd.cu:39-53: iota, then std::random_shuffle
 * driven by glibc rand() after srand(seed).  libstdc++'s random_shuffle swaps
 * element i with element rand() % (i + 1) for i = 1 .. size-1. */
EXPORT void orc_gen_random_map(uint32_t seed, int32_t size, int32_t* map)
{
    srand(seed);
    for (int32_t i = 0; i < size; ++i) map[i] = i;
    for (int32_t i = 1; i < size; ++i) {
        int32_t j = rand() % (i + 1);
        if (j != i) { int32_t t = map[i]; map[i] = map[j]; map[j] = t; }
    }
}

/* gen_inv_map, singleGPU/sgd.cu:55-60. */
EXPORT void orc_gen_inv_map(const int32_t* map, int32_t size, int32_t* inv)
{
    for (int32_t i = 0; i < size; ++i) inv[map[i]] = i;
}

/* shuffle_problem, singleGPU/sgd.cu:126-134. */
EXPORT void orc_shuffle_problem(orc_node* R, int64_t nnz, const int32_t* p_map, const int32_t* q_map)
{
    for (int64_t i = 0; i < nnz; ++i) {
        R[i].u = p_map[R[i].u];
        R[i].v = q_map[R[i].v];
    }
}

/* Segment lengths, singleGPU/sgd.cu:419-423. */
EXPORT void orc_grid_segments(int32_t m, int32_t n, int ux, int vy, int64_t* u_seg, int64_t* v_seg)
{
    *u_seg = (ux == 1) ? m : (int64_t)ceil((double)m / ux);
    *v_seg = (vy == 1) ? n : (int64_t)ceil((double)n / vy);
}

/* grid_problem, singleGPU/sgd.cu:411-487: count records per block
 * (u/u_seg)*vy + v/v_seg, prefix offsets, then the in-place cycle-swap
 * partition.  grid_size has ux*vy entries, offsets ux*vy+1. */
EXPORT void orc_grid_problem(orc_node* R, int64_t nnz, int32_t m, int32_t n, int ux, int vy,
                             int64_t* grid_size, int64_t* offsets)
{
    int64_t u_seg, v_seg;
    orc_grid_segments(m, n, ux, vy, &u_seg, &v_seg);
    int nb = ux * vy;
    for (int g = 0; g < nb; ++g) grid_size[g] = 0;
    for (int64_t i = 0; i < nnz; ++i)
        grid_size[(R[i].u / u_seg) * vy + R[i].v / v_seg]++;
    offsets[0] = 0;
    for (int g = 0; g < nb; ++g) offsets[g + 1] = offsets[g] + grid_size[g];
    int64_t* pivots = (int64_t*)malloc(sizeof(int64_t) * (size_t)nb);
    for (int g = 0; g < nb; ++g) pivots[g] = offsets[g];
    for (int g = 0; g < nb; ++g) {
        for (int64_t p = pivots[g]; p != offsets[g + 1];) {
            int cg = (int)((R[p].u / u_seg) * vy + R[p].v / v_seg);
            if (cg == g) { ++p; continue; }
            int64_t nx = pivots[cg];
            orc_node t = R[p]; R[p] = R[nx]; R[nx] = t;
            pivots[cg]++;
        }
    }
    free(pivots);
}

/* scale_problem, singleGPU/sgd.cu:98-124: r *= scale (the caller passes
 * float(1.0/std)); with more than one block, ids are localised by repeated
 * subtraction of the segment length. */
EXPORT void orc_scale_problem(orc_node* R, int64_t nnz, float scale, int ux, int vy,
                              int64_t u_seg, int64_t v_seg)
{
    for (int64_t i = 0; i < nnz; ++i) {
        R[i].r = R[i].r * scale;
        if (ux * vy != 1) {
            int64_t tu = R[i].u;
            while (tu >= u_seg) tu -= u_seg;
            R[i].u = (int32_t)tu;
            int64_t tv = R[i].v;
            while (tv >= v_seg) tv -= v_seg;
            R[i].v = (int32_t)tv;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* The SGD step.  Update rule from singleGPU/sgd_k128_kernel_hogwild_warp32.h */
/* :85-99: e = r - p.q; p += lr(e q - lp p); q += lr(e p_old - lq q).        */
/* Arithmetic order (this project's kernel contract):                        */
/*  - the k elements are split over G lanes, lane l owning the E = k/G       */
/*    consecutive elements [lE, lE+E) (when E elements take 32 bytes: the    */
/*    E/2 elements [l E/2, (l+1) E/2) of each half of the row, lane_elem);   */
/*  - lane partial: a = p0*q0, then a = fma(p_j, q_j, a) for j = 1..E-1;     */
/*  - the G partials are summed by a pairwise butterfly (lane l adds lane    */
/*    l^1, then l^2, l^4, ...), i.e. a balanced tree in lane order;          */
/*  - e = r - dot;                                                           */
/*  - cp = fma(-lr, lp, 1), cq = fma(-lr, lq, 1), le = lr*e;                */
/*    p' = fma(le, q, p*cp), q' = fma(le, p, q*cq) (the same step: p (1 -    */
/*    lr lp) + lr e q).                                                      */
/* ------------------------------------------------------------------------ */
/* Element i of lane l.  A lane holding at most 16 bytes of a row owns
 * [lE, lE+E); a lane holding 32 bytes (two 16-byte pieces, half = E/2
 * elements each) owns [l*half, (l+1)*half) of each half of the row, piece
 * c = i / half, so that each of the kernel's two 16-byte accesses covers one
 * half row contiguously.  half = 0 selects the first layout. */
static inline int lane_elem(int k, int G, int half, int l, int i)
{
    if (!half) return l * (k / G) + i;
    return (i / half) * (k / 2) + l * half + i % half;
}

/* half for factors of length k in a storage format with G lanes. */
static inline int lane_half(int k, int G, int prec)
{
    int E = k / G;
    return E * (int)prec_bytes(prec) == 32 ? E / 2 : 0;
}

static float tree_dot_layout(const float* p, const float* q, int k, int G, int half, float* lanes)
{
    int E = k / G;
    for (int l = 0; l < G; ++l) {
        int e0 = lane_elem(k, G, half, l, 0);
        float a = p[e0] * q[e0];
        for (int j = 1; j < E; ++j) {
            int e = lane_elem(k, G, half, l, j);
            a = fmaf(p[e], q[e], a);
        }
        lanes[l] = a;
    }
    for (int d = 1; d < G; d <<= 1)
        for (int l = 0; l < G; l += 2 * d) lanes[l] = lanes[l] + lanes[l + d];
    return lanes[0];
}

static float tree_dot(const float* p, const float* q, int k, int G, float* lanes)
{
    return tree_dot_layout(p, q, k, G, 0, lanes);
}

EXPORT float orc_tree_dot(const float* p, const float* q, int k, int G)
{
    float lanes[256];
    return tree_dot(p, q, k, G, lanes);
}

/* The dot product in the two-piece layout of 32-byte lanes (half elements
 * per piece). */
EXPORT float orc_tree_dot_split(const float* p, const float* q, int k, int G, int half)
{
    float lanes[256];
    return tree_dot_layout(p, q, k, G, half, lanes);
}

/* Lane-local index of global element e in that layout (for the pins). */
EXPORT int orc_lane_elem(int k, int G, int half, int l, int i) { return lane_elem(k, G, half, l, i); }

typedef struct {
    int k;
    int G;              /* lanes per sample (dot-product order) */
    int prec;
    int batch_len;      /* update_vector_size, singleGPU/sgd_kernel.cu:198 */
    int workers;
    uint32_t problem;
    uint64_t seed;
    float lambda_p, lambda_q;
} orc_sgd_cfg;

/* Random bits for stochastic rounding of the FP8 and BF16-SR write-back of
 * sample j (0..L-1) of window `batch`: one Philox block per window,
 * K = Philox(ctr = {batch, epoch, worker, STREAM_ROUND}).  Lane l of the lane
 * group (elements [lE, lE+E), E = k/G) draws the seed
 * W_0 = lowbias32(K[0] ^ (j*0x9E3779B9 + l*0x85EBCA77)) and
 * W_{i+1} = sr_next(W_i); element i of its P row uses the 16-bit half (i & 1)
 * of W_{i>>1} and element i of its Q row the other half.  Sixteen bits per
 * element keep late-epoch updates far below one FP8 step unbiased. */
static inline uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

/* xorshift32 (Marsaglia 2003) plus a Weyl increment. */
static inline uint32_t sr_next(uint32_t x)
{
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x + 0x6D2B79F5u;
}

/* The 16 random bits of element e (0..k-1) of row `which` (0 P, 1 Q); the
 * lane and its element index i follow the layout (half, as lane_elem). */
static inline uint32_t sr_bits(const uint32_t K[4], uint32_t j, int which, int e, int k, int G, int half)
{
    int E = k / G, lane = e / E, i = e % E;
    if (half) {
        int r = e % (k / 2);
        lane = r / half;
        i = (e / (k / 2)) * half + r % half;
    }
    uint32_t w = lowbias32(K[0] ^ (j * 0x9E3779B9u + (uint32_t)lane * 0x85EBCA77u));
    for (int s = 0; s < (i >> 1); ++s) w = sr_next(w);
    return ((i & 1) ^ (which != 0)) ? (w >> 16) : (w & 0xffffu);
}

/* FP8 rows with power-of-two row scales sp, sq: the bytes' values are used
 * in their own units (row value = byte value * scale) and the scales enter
 * once per sample.  Scaling by a power of two commutes with every rounding
 * (away from FP32 underflow and overflow), so this is the FP32 step on the
 * widened rows: dot = (sp*sq) * tree(p8 . q8); with le = lr*e the stored
 * values are p8' = fma(le*(sq*(1/sp)), q8, p8*cp) and
 * q8' = fma(le*(sp*(1/sq)), p8, q8*cq), clamped to +-448 and rounded
 * stochastically with sr_bits (16 random bits per element).
 * A non-finite multiplier (the only source of a NaN) writes 0xff to every
 * element of both rows. */
static void sgd_update_one_e4m3(const orc_sgd_cfg* c, void* P, void* Q, float sp, float sq, int64_t bp,
                                int64_t bq, float r, float lr, const uint32_t K[4], uint32_t sj, float* p,
                                float* q, float* lanes)
{
    int k = c->k;
    uint8_t* P8 = (uint8_t*)P;
    uint8_t* Q8 = (uint8_t*)Q;
    for (int j = 0; j < k; ++j) {
        p[j] = orc_e4m3_to_f32(P8[bp + j]);
        q[j] = orc_e4m3_to_f32(Q8[bq + j]);
    }
    const int half = lane_half(k, c->G, c->prec);
    float dot = tree_dot_layout(p, q, k, c->G, half, lanes) * (sp * sq);
    float e = r - dot;
    float cp = fmaf(-lr, c->lambda_p, 1.0f), cq = fmaf(-lr, c->lambda_q, 1.0f);
    float le = lr * e;
    float lqp = le * (sq * (1.0f / sp)), lpq = le * (sp * (1.0f / sq));
    int nan_rows = !(isfinite(lqp) && isfinite(lpq));
    for (int j = 0; j < k; ++j) {
        float pn = fmaf(lqp, q[j], p[j] * cp);
        float qn = fmaf(lpq, p[j], q[j] * cq);
        P8[bp + j] = nan_rows ? 0xffu : orc_f32_to_e4m3_sr(pn, sr_bits(K, sj, 0, j, k, c->G, half));
        Q8[bq + j] = nan_rows ? 0xffu : orc_f32_to_e4m3_sr(qn, sr_bits(K, sj, 1, j, k, c->G, half));
    }
}

/* One SGD update of (u, v, r) on FP32 math with storage rounding. */
static void sgd_update_one(const orc_sgd_cfg* c, void* P, void* Q, float* ps, float* qs,
                           int32_t u, int32_t v, float r, float lr, const uint32_t K[4], uint32_t sj,
                           float* p, float* q, float* lanes)
{
    int k = c->k;
    int64_t bp = (int64_t)u * k, bq = (int64_t)v * k;
    if (c->prec == PREC_E4M3) {
        sgd_update_one_e4m3(c, P, Q, ps[u], qs[v], bp, bq, r, lr, K, sj, p, q, lanes);
        return;
    }
    float sp = 1.0f, sq = 1.0f, ip = 1.0f, iq = 1.0f;
    for (int j = 0; j < k; ++j) {
        p[j] = load_elem(P, c->prec, bp + j) * sp;
        q[j] = load_elem(Q, c->prec, bq + j) * sq;
    }
    const int half = lane_half(k, c->G, c->prec);
    float dot = tree_dot_layout(p, q, k, c->G, half, lanes);
    float e = r - dot;
    /* p + lr (e q - lp p), evaluated as p (1 - lr lp) + (lr e) q */
    float cp = fmaf(-lr, c->lambda_p, 1.0f), cq = fmaf(-lr, c->lambda_q, 1.0f);
    float le = lr * e;
    for (int j = 0; j < k; ++j) {
        float pn = fmaf(le, q[j], p[j] * cp);
        float qn = fmaf(le, p[j], q[j] * cq);
        uint32_t rp = 0, rq = 0;
        if (c->prec == PREC_E4M3 || c->prec == PREC_BF16_SR) {
            rp = sr_bits(K, sj, 0, j, k, c->G, half);
            rq = sr_bits(K, sj, 1, j, k, c->G, half);
        }
        store_elem(P, c->prec, bp + j, pn, ip, sp, rp);
        store_elem(Q, c->prec, bq + j, qn, iq, sq, rq);
    }
}

/* Serialized single-worker epochs: worker 0 draws ceil(nnz/L) windows per
 * epoch (the reference's update_count with one worker,
 * singleGPU/sgd_kernel.cu:197-208) and walks each window of L consecutive
 * records modulo nnz in order (h:50-56).  Factors are updated in place.
 * R is given as SoA arrays (u, v, r) of the block being trained. */
EXPORT int orc_sgd_serial(const orc_sgd_cfg* c, const int32_t* ru, const int32_t* rv,
                          const float* rr, int64_t nnz, void* P, void* Q, float* ps, float* qs,
                          const float* lr_table, int epoch_begin, int epoch_end)
{
    if (nnz <= 0) return 0;
    int L = c->batch_len;
    int64_t nb = (nnz + (int64_t)L - 1) / L;
    float* p = (float*)malloc(sizeof(float) * (size_t)c->k);
    float* q = (float*)malloc(sizeof(float) * (size_t)c->k);
    float lanes[256];
    for (int ep = epoch_begin; ep < epoch_end; ++ep) {
        float lr = lr_table[ep];
        for (int64_t b = 0; b < nb; ++b) {
            int64_t start = orc_window_start(c->seed, c->problem, 0, 0, (uint32_t)ep, (uint32_t)b, nnz);
            uint32_t K[4] = {0, 0, 0, 0};
            if (c->prec == PREC_E4M3 || c->prec == PREC_BF16_SR) {
                uint32_t ctr[4] = {(uint32_t)b, (uint32_t)ep, 0u, STREAM_ROUND};
                uint32_t key[2];
                key_of(c->seed, c->problem, key);
                orc_philox4x32_10(ctr, key, K);
            }
            for (int i = 0; i < L; ++i) {
                int64_t off = (start + i) % nnz;
                sgd_update_one(c, P, Q, ps, qs, ru[off], rv[off], rr[off], lr, K, (uint32_t)i, p, q, lanes);
            }
        }
    }
    free(p);
    free(q);
    return 0;
}

/* Windows [window_begin, window_begin + windows) of worker 0 of one rating
 * block in epoch `epoch`: the block's task in a gridded or stratum launch
 * (singleGPU/sgd_kernel.cu:249-372 trains one block per launch), whose
 * window draws and rounding words carry the block id as `tag`.  Each window
 * is L consecutive records of the block modulo its size (h:50-56).  With at
 * most one worker per block the stratum schedule's epoch is exactly: the
 * diagonals in order, and within a diagonal each block's windows in turn
 * (its blocks share no P or Q rows, so their order does not matter).
 * Factors are updated in place. */
EXPORT int orc_sgd_block_windows(const orc_sgd_cfg* c, uint32_t tag, const int32_t* ru, const int32_t* rv,
                                 const float* rr, int64_t nnz, void* P, void* Q, float* ps, float* qs,
                                 const float* lr_table, int epoch, int window_begin, int windows)
{
    if (nnz <= 0) return 0;
    float* p = (float*)malloc(sizeof(float) * (size_t)c->k);
    float* q = (float*)malloc(sizeof(float) * (size_t)c->k);
    float lanes[256];
    const float lr = lr_table[epoch];
    for (int w = 0; w < windows; ++w) {
        const uint32_t b = (uint32_t)(window_begin + w);
        int64_t start = orc_window_start(c->seed, c->problem, tag, 0, (uint32_t)epoch, b, nnz);
        uint32_t K[4] = {0, 0, 0, 0};
        if (c->prec == PREC_E4M3 || c->prec == PREC_BF16_SR) {
            uint32_t ctr[4] = {b, (uint32_t)epoch, 0u, STREAM_ROUND | (tag << 8)};
            uint32_t key[2];
            key_of(c->seed, c->problem, key);
            orc_philox4x32_10(ctr, key, K);
        }
        for (int i = 0; i < c->batch_len; ++i) {
            int64_t off = (start + i) % nnz;
            sgd_update_one(c, P, Q, ps, qs, ru[off], rv[off], rr[off], lr, K, (uint32_t)i, p, q, lanes);
        }
    }
    free(p);
    free(q);
    return 0;
}

/* FP64 serialized oracle: same windows, everything in double, sequential dot. */
EXPORT int orc_sgd_serial_f64(int k, int batch_len, uint64_t seed, uint32_t problem,
                              double lambda_p, double lambda_q, const int32_t* ru,
                              const int32_t* rv, const float* rr, int64_t nnz, double* P,
                              double* Q, const float* lr_table, int epoch_begin, int epoch_end)
{
    if (nnz <= 0) return 0;
    int64_t nb = (nnz + (int64_t)batch_len - 1) / batch_len;
    for (int ep = epoch_begin; ep < epoch_end; ++ep) {
        double lr = lr_table[ep];
        for (int64_t b = 0; b < nb; ++b) {
            int64_t start = orc_window_start(seed, problem, 0, 0, (uint32_t)ep, (uint32_t)b, nnz);
            for (int i = 0; i < batch_len; ++i) {
                int64_t off = (start + i) % nnz;
                double* p = P + (int64_t)ru[off] * k;
                double* q = Q + (int64_t)rv[off] * k;
                double dot = 0;
                for (int j = 0; j < k; ++j) dot += p[j] * q[j];
                double e = (double)rr[off] - dot;
                for (int j = 0; j < k; ++j) {
                    double pj = p[j], qj = q[j];
                    p[j] = pj + lr * (e * qj - lambda_p * pj);
                    q[j] = qj + lr * (e * pj - lambda_q * qj);
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Model post-processing and evaluation.                                     */
/* ------------------------------------------------------------------------ */

/* Dequantize a factor matrix to FP32. */
EXPORT void orc_widen(const void* src, const float* row_scale, int64_t rows, int k, int prec, float* dst)
{
    for (int64_t r = 0; r < rows; ++r)
        for (int e = 0; e < k; ++e) {
            float v = load_elem(src, prec, r * k + e);
            if (prec == PREC_E4M3) v *= row_scale[r];
            dst[r * k + e] = v;
        }
}

/* scale_model, singleGPU/sgd.cu:593-606: multiply the first rows*k floats by
 * float(sqrt(scale)). */
EXPORT void orc_scale_model(float* vec, int64_t rows, int k, float scale)
{
    float fs = sqrtf(scale);
    for (int64_t i = 0; i < rows * (int64_t)k; ++i) vec[i] = vec[i] * fs;
}

/* shuffle_model's inv_shuffle1, singleGPU/sgd.cu:611-631: in-place cycle
 * following on the inverse map (which it destroys). */
EXPORT void orc_inv_shuffle(float* vec, int32_t* map, int32_t size, int k)
{
    for (int32_t pivot = 0; pivot < size;) {
        if (pivot == map[pivot]) { ++pivot; continue; }
        int32_t nx = map[pivot];
        for (int d = 0; d < k; ++d) {
            float t = vec[(int64_t)pivot * k + d];
            vec[(int64_t)pivot * k + d] = vec[(int64_t)nx * k + d];
            vec[(int64_t)nx * k + d] = t;
        }
        map[pivot] = map[nx];
        map[nx] = nx;
    }
}

/* mf_predict + calc_rmse, test/mf.cpp:3794-3839: sequential float dot with
 * separate multiply and add (std::inner_product), b for out-of-range ids or a
 * nan dot, float e, double sum of float e*e. */
EXPORT double orc_rmse(const int32_t* tu, const int32_t* tv, const float* tr, int64_t nnz,
                       const float* P, const float* Q, int32_t m, int32_t n, int k, float b)
{
    if (nnz == 0) return 0;
    double loss = 0;
    for (int64_t i = 0; i < nnz; ++i) {
        float z;
        int32_t u = tu[i], v = tv[i];
        if (u < 0 || u >= m || v < 0 || v >= n) {
            z = b;
        } else {
            const float* p = P + (int64_t)u * k;
            const float* q = Q + (int64_t)v * k;
            z = 0.0f;
            for (int j = 0; j < k; ++j) {
                float t = p[j] * q[j];
                z = z + t;
            }
            if (isnan(z)) z = b;
        }
        float e = tr[i] - z;
        float e2 = e * e;
        loss += e2;
    }
    return sqrt(loss / nnz);
}

/* mf_predict (test/mf.cpp:3794-3812) for one test rating. */
static float predict_one(const float* P, const float* Q, int32_t m, int32_t n, int k, float b, int32_t u, int32_t v)
{
    if (u < 0 || u >= m || v < 0 || v >= n) return b;
    const float* p = P + (int64_t)u * k;
    const float* q = Q + (int64_t)v * k;
    float z = 0.0f;
    for (int j = 0; j < k; ++j) {
        float t = p[j] * q[j];
        z = z + t;
    }
    return isnan(z) ? b : z;
}

/* The other evaluation criteria of mf-predict -e (test/mf.h:30-40,
 * test/mf-predict.cpp:129-162), each a double sum of a per-rating term whose
 * arithmetic is the reference's:
 *   1 MAE      abs(float(r - z))                          (mf.cpp:3841-3855)
 *   2 GKL      float: r*logf(r/z) - r + z                 (mf.cpp:3857-3872)
 *   5 LOGLOSS  log(1.0 + double(expf(-+z))), by r > 0     (mf.cpp:3874-3892)
 *   6 ACC      r > 0 ? z > 0 : z < 0                      (mf.cpp:3894-3912)
 * and the mean over nnz; 0 = RMSE as orc_rmse.  Returns NAN for others. */
EXPORT double orc_eval(int criterion, const int32_t* tu, const int32_t* tv, const float* tr, int64_t nnz,
                       const float* P, const float* Q, int32_t m, int32_t n, int k, float b)
{
    if (criterion == 0) return orc_rmse(tu, tv, tr, nnz, P, Q, m, n, k, b);
    if (criterion != 1 && criterion != 2 && criterion != 5 && criterion != 6) return NAN;
    if (nnz == 0) return 0;
    double acc = 0;
    for (int64_t i = 0; i < nnz; ++i) {
        const float z = predict_one(P, Q, m, n, k, b, tu[i], tv[i]);
        const float r = tr[i];
        switch (criterion) {
        case 1: {
            float e = r - z;
            acc += fabsf(e);
            break;
        }
        case 2: {
            float q = r / z;
            float l = logf(q);
            float t = r * l;
            t = t - r;
            t = t + z;
            acc += t;
            break;
        }
        case 5:
            acc += r > 0 ? log(1.0 + (double)expf(-z)) : log(1.0 + (double)expf(z));
            break;
        default:
            acc += r > 0 ? (z > 0 ? 1 : 0) : (z < 0 ? 1 : 0);
            break;
        }
    }
    return acc / nnz;
}

/* calc_mpr_auc (test/mf.cpp:3914-4034): per row (user, or item when
 * transposed) with test positives (r > 0), every column of the row is scored
 * with mf_predict; each (negative column, positive) pair counts towards MPR
 * when the positive scores <= the negative and towards AUC otherwise.
 * Restated as the reference computes it: rows sorted by (row, col), the
 * row's positives swapped to the front and sorted by score, a binary search
 * per negative.  m, n = max(test extent, model extent).  Test pairs must be
 * distinct (the reference's swap assumes it). */
static int cmp_pair(const void* a, const void* b)
{
    const int64_t* x = (const int64_t*)a;
    const int64_t* y = (const int64_t*)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}
static int cmp_score(const void* a, const void* b)
{
    float x = *(const float*)a, y = *(const float*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}
EXPORT void orc_mpr_auc(const int32_t* tu, const int32_t* tv, const float* tr, int64_t nnz, const float* P,
                        const float* Q, int32_t model_m, int32_t model_n, int k, float b, int transpose,
                        double* mpr_out, double* auc_out)
{
    int64_t pm = 0, pn = 0;   /* read_problem's extents: largest id + 1 */
    for (int64_t i = 0; i < nnz; ++i) {
        if (tu[i] + 1 > pm) pm = tu[i] + 1;
        if (tv[i] + 1 > pn) pn = tv[i] + 1;
    }
    int64_t m = transpose ? (pn > model_n ? pn : model_n) : (pm > model_m ? pm : model_m);
    int64_t n = transpose ? (pm > model_m ? pm : model_m) : (pn > model_n ? pn : model_n);
    /* (row, col, r-as-bits) sorted by (row, col) */
    int64_t* rec = (int64_t*)malloc(sizeof(int64_t) * 3 * (size_t)(nnz ? nnz : 1));
    for (int64_t i = 0; i < nnz; ++i) {
        rec[3 * i] = transpose ? tv[i] : tu[i];
        rec[3 * i + 1] = transpose ? tu[i] : tv[i];
        uint32_t bits;
        memcpy(&bits, &tr[i], 4);
        rec[3 * i + 2] = bits;
    }
    qsort(rec, (size_t)nnz, sizeof(int64_t) * 3, cmp_pair);
    int64_t* cnt = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));
    for (int64_t i = 0; i < nnz; ++i) cnt[rec[3 * i] + 1]++;
    for (int64_t i = 1; i <= m; ++i) cnt[i] += cnt[i - 1];
    float* score = (float*)malloc(sizeof(float) * (size_t)(n ? n : 1));
    int64_t* col = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    char* is_pos = (char*)calloc((size_t)(n ? n : 1), 1);
    float* ps = (float*)malloc(sizeof(float) * (size_t)(n ? n : 1));
    int64_t total_m = 0, total_pos = 0;
    double all_mpr = 0, all_auc = 0;
    for (int64_t i = 0; i < m; ++i) {
        if (cnt[i + 1] - cnt[i] < 1) continue;
        for (int64_t j = 0; j < n; ++j) {
            int32_t uu = (int32_t)(transpose ? j : i), vv = (int32_t)(transpose ? i : j);
            score[j] = predict_one(P, Q, model_m, model_n, k, b, uu, vv);
            col[j] = j;
            is_pos[j] = 0;
        }
        int64_t pos = 0;
        for (int64_t t = cnt[i]; t < cnt[i + 1]; ++t) {
            uint32_t bits = (uint32_t)rec[3 * t + 2];
            float r;
            memcpy(&r, &bits, 4);
            if (r <= 0) continue;
            is_pos[rec[3 * t + 1]] = 1;
            ps[pos++] = score[rec[3 * t + 1]];
        }
        if (n - pos < 1 || pos < 1) continue;
        total_m++;
        total_pos += pos;
        qsort(ps, (size_t)pos, sizeof(float), cmp_score);
        double u_mpr = 0, u_auc = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (is_pos[j]) continue;
            float s = score[j];
            if (ps[pos - 1] <= s) {
                u_mpr += (double)pos;
                continue;
            }
            int64_t left = 0, right = pos - 1;
            while (left < right) {
                int64_t mid = (left + right) / 2;
                if (ps[mid] > s) right = mid;
                else left = mid + 1;
            }
            u_mpr += (double)left;
            u_auc += (double)(pos - left);
        }
        all_mpr += u_mpr / (double)(n - pos);
        all_auc += u_auc / (double)(n - pos) / (double)pos;
    }
    *mpr_out = total_pos ? all_mpr / (double)total_pos : NAN;
    *auc_out = total_m ? all_auc / (double)total_m : NAN;
    free(rec); free(cnt); free(score); free(col); free(is_pos); free(ps);
}

/* save_model, singleGPU/main.cpp:231-273: header {int 0, int m, int n, int k,
 * float b}, then m rows of P and n rows of Q as float32. */
EXPORT int orc_save_model(const char* path, int32_t m, int32_t n, int32_t k, float b,
                          const float* P, const float* Q)
{
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    int32_t hdr[4] = {0, m, n, k};
    fwrite(hdr, sizeof(int32_t), 4, f);
    fwrite(&b, sizeof(float), 1, f);
    fwrite(P, sizeof(float), (size_t)m * k, f);
    fwrite(Q, sizeof(float), (size_t)n * k, f);
    fclose(f);
    return 0;
}
