// Device helpers for the gfx950 SGD kernels: Philox4x32-10, factor storage
// formats (FP32 / FP16 / BF16 / OCP FP8-E4M3) and lane-group row I/O.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cumf {

// kBF16SR: BF16 storage whose update write-back rounds stochastically
enum Prec : int { kF32 = 0, kF16 = 1, kBF16 = 2, kE4M3 = 3, kBF16SR = 4 };

template <int PREC> struct PrecBytes { static constexpr int value = PREC == kF32 ? 4 : (PREC == kE4M3 ? 1 : 2); };

// Stream tags (low byte of Philox counter word 3).
constexpr uint32_t kStreamInitP = 1u;
constexpr uint32_t kStreamInitQ = 2u;
constexpr uint32_t kStreamWindow = 3u;
constexpr uint32_t kStreamRound = 4u;

// FP8 row-scale targets: largest |x| of a row maps to at most this value
// (headroom below 448 for growth between rescales).
constexpr float kFp8InitTarget = 14.0f;
constexpr float kFp8EpochTarget = 56.0f;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ inline U4 philox(U4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
#ifdef __HIP_DEVICE_COMPILE__
        uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
#else
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
#endif
        U4 n;
        n.x = hi1 ^ c.y ^ k0;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k1;
        n.w = lo0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

struct Key { uint32_t k0, k1; };

__host__ __device__ inline Key make_key(uint64_t seed, uint32_t problem)
{
    return Key{(uint32_t)seed, (uint32_t)(seed >> 32) ^ problem};
}

// Window start in [0, nnz): high 64 bits of a 64x64 product.
__device__ inline int64_t window_start(Key key, uint32_t tag, uint32_t worker, uint32_t epoch,
                                       uint32_t batch, int64_t nnz)
{
    U4 r = philox(U4{batch, epoch, worker, kStreamWindow | (tag << 8)}, key.k0, key.k1);
    uint64_t x = ((uint64_t)r.y << 32) | r.x;
    return (int64_t)__umul64hi(x, (uint64_t)nnz);
}

__host__ __device__ inline float u01_of(uint32_t x)
{
    return (float)((x >> 8) + 1u) * 0x1.0p-24f;
}

// ---------------------------------------------------------------------------
// Formats.  Conversions are spelled out in integer arithmetic so that they are
// bit-for-bit the same as the CPU definitions.
// ---------------------------------------------------------------------------
__device__ inline float f16_to_f32(uint32_t h)
{
    _Float16 x = __builtin_bit_cast(_Float16, (uint16_t)h);
    return (float)x;
}
__device__ inline uint32_t f32_to_f16(float f)
{
    // The value is first rounded to FP32 and then to FP16, as __float2half of
    // an FP32 expression does in the reference.  The empty asm keeps the
    // compiler from fusing the producing fma into one fma-to-f16 rounding
    // (v_fma_mixlo_f16), which rounds once and can differ in the last bit.
    asm("" : "+v"(f));
    _Float16 x = (_Float16)f;   // v_cvt_f16_f32, round to nearest even
    return __builtin_bit_cast(uint16_t, x);
}
// a * b rounded to FP32, a the FP16 in half HA of word aw, b an FP32 value.
// One v_fma_mix_f32 with a -0.0 addend: x + (-0) == x for every x, signed
// zeros included, so this is the FP32 product of the widened operands
// without a separate conversion.
template <bool HA>
__device__ inline float mul_f16_f32(uint32_t aw, float b)
{
    float d;
    const float nz = -0.0f;
    if constexpr (HA)
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(d) : "v"(aw), "v"(b), "s"(nz));
    else
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(d) : "v"(aw), "v"(b), "s"(nz));
    return d;
}
// The same for two FP16 operands, both in half H of their words.
template <bool H>
__device__ inline float mul_f16_f16(uint32_t aw, uint32_t bw)
{
    float d;
    const float nz = -0.0f;
    if constexpr (H)
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,1,0] op_sel_hi:[1,1,0]" : "=v"(d) : "v"(aw), "v"(bw), "s"(nz));
    else
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,1,0]" : "=v"(d) : "v"(aw), "v"(bw), "s"(nz));
    return d;
}
// Two FP32 values rounded to FP16 (nearest even) in one word, lo = a,
// hi = b (v_cvt_pk_f16_f32).  The values are first rounded to FP32 by their
// producers: the empty asm keeps the compiler from fusing a producing fma
// into one fma-to-f16 rounding.
__device__ inline uint32_t f32x2_to_f16x2(float a, float b)
{
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    asm("" : "+v"(a), "+v"(b));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{a, b}, h2));
}
__device__ inline float bf16_to_f32(uint32_t b) { return __uint_as_float(b << 16); }
// Round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ inline uint32_t f32_to_bf16(float f)
{
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}
// Two values in one word, lo = a, hi = b: one v_cvt_pk_bf16_f32.
__device__ inline uint32_t f32x2_to_bf16x2(float a, float b)
{
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f2{a, b}, b2));
}
// Stochastic rounding with 16 random bits r16 (low half): adding them to the
// 16 bits that truncation drops carries into the kept part with probability
// equal to their value / 2^16.  NaN stays a quiet NaN.
__device__ inline uint32_t f32_to_bf16_sr(float f, uint32_t r16)
{
    const uint32_t x = __float_as_uint(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (x >> 16) | 0x40u;
    return (x + (r16 & 0xffffu)) >> 16;
}
// The same rounding by v_cvt_sr_bf16_f32, which adds the top half of its
// random word (rword = r16 << 16; the low half is not read: over 4.2M
// inputs in both halves no result depended on it) and quiets NaN the same
// way, written into half HI of the packed pair `old`.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
template <bool HI>
__device__ inline uint32_t bf16_sr_insert(float f, uint32_t rword, uint32_t old)
{
    const bf16x2_t o = __builtin_bit_cast(bf16x2_t, old);
    return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_bf16_f32(o, f, rword, HI));
}
__device__ inline float e4m3_to_f32(uint32_t b)
{
    uint32_t sign = (b >> 7) & 1u, exp = (b >> 3) & 0xfu, man = b & 7u;
    float v;
    if (exp == 0xfu && man == 7u) return __uint_as_float(0x7fc00000u | (sign << 31));
    if (exp == 0u)
        v = (float)man * 0x1.0p-9f;
    else
        v = __uint_as_float(((exp + 120u) << 23) | (man << 20));
    return sign ? -v : v;
}
__device__ inline uint32_t e4m3_pack_normal(uint32_t ax)
{
    return (((ax >> 23) - 120u) << 3) | ((ax >> 20) & 7u);
}
__device__ inline uint32_t f32_to_e4m3_rne(float f)
{
    uint32_t x = __float_as_uint(f);
    uint32_t sign = (x >> 24) & 0x80u, ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return sign | 0x7fu;
    if (ax >= 0x43e00000u) return sign | 0x7eu;
    if (ax < 0x3c800000u) return sign | (uint32_t)rintf(__uint_as_float(ax) * 0x1.0p9f);
    uint32_t r = (ax + 0x7ffffu + ((ax >> 20) & 1u)) & 0xfff00000u;
    r = r > 0x43e00000u ? 0x43e00000u : r;
    return sign | e4m3_pack_normal(r);
}
// Stochastic rounding into byte sel (0..3, a constant after unrolling) of
// `old`, with v_cvt_sr_fp8_f32.  The instruction adds the top 20 bits of its
// random word into the 20-bit fraction of the dropped part (in units of the
// E4M3 quantum, 2^(e-3) for normal values, 2^-9 for subnormal ones) and
// truncates; rword carries 16 random bits in its top half, so |f| rounds up
// when floor(frac * 2^20) + (rword >> 12) >= 2^20.  The instruction turns
// values above 448 into NaN, so f is clamped to +-448 first, by one
// v_med3_f32 (the caller keeps NaN away: a median would make it a bound).
__device__ inline uint32_t e4m3_sr_clamped(float f, uint32_t rword, uint32_t old, int sel)
{
    const float v = __builtin_amdgcn_fmed3f(f, -448.0f, 448.0f);
    switch (sel) {
    case 0: return (uint32_t)__builtin_amdgcn_cvt_sr_fp8_f32(v, (int)rword, (int)old, 0);
    case 1: return (uint32_t)__builtin_amdgcn_cvt_sr_fp8_f32(v, (int)rword, (int)old, 1);
    case 2: return (uint32_t)__builtin_amdgcn_cvt_sr_fp8_f32(v, (int)rword, (int)old, 2);
    default: return (uint32_t)__builtin_amdgcn_cvt_sr_fp8_f32(v, (int)rword, (int)old, 3);
    }
}

// Smallest power of two s with amax <= target * s, kept in [2^-126, 2^126]:
// s and 1/s are then normal, and v_rcp_f32 returns 1/s exactly (an in-bounds
// probe compared it with 1.0f / s on every power of two).
__device__ inline float pow2_scale(float amax, float target)
{
    if (!(amax > 0.0f) || __builtin_isinf(amax)) return 1.0f;
    int ea, et;
    float fa = frexpf(amax, &ea);
    float ft = frexpf(target, &et);
    const int n = ea - et + (fa > ft ? 1 : 0);
    return ldexpf(1.0f, n < -126 ? -126 : (n > 126 ? 126 : n));
}

// ---------------------------------------------------------------------------
// Row I/O: a lane owns E consecutive elements of a row (TB = E * bytes).
// COH = true uses device-scope relaxed atomics so every access is coherent
// across XCDs (bypasses the per-XCD L2 copies).
// ---------------------------------------------------------------------------
template <int NW> struct Words { uint32_t w[NW]; };
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Global-address-space pointers: the kernels take their buffers from a task
// table in memory, so the compiler cannot infer the address space and would
// emit flat accesses, which also count against the LDS wait counter.
typedef __attribute__((address_space(1))) char gchar;
typedef __attribute__((address_space(1))) const int32_t gcint;
typedef __attribute__((address_space(1))) const float gcfloat;
template <typename T>
__device__ inline __attribute__((address_space(3))) void* to_lds(T* p)
{
    return (__attribute__((address_space(3))) void*)p;
}
// s_waitcnt immediate (gfx9 encoding) that waits for LDS reads only
constexpr int kWaitLgkm0 = 0xC07F;
template <typename T>
__device__ inline __attribute__((address_space(1))) T* to_global(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

template <int TB, bool COH, int SPLIT = 16>
__device__ inline Words<(TB + 3) / 4> load_raw(const gchar* p)
{
    Words<(TB + 3) / 4> r;
    if constexpr (TB == 32) {
        const Words<4> a = load_raw<16, COH>(p), b = load_raw<16, COH>(p + SPLIT);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            r.w[i] = a.w[i];
            r.w[4 + i] = b.w[i];
        }
    } else if constexpr (TB == 2) {
        static_assert(!COH, "2-byte coherent loads are not supported");
        r.w[0] = *(const __attribute__((address_space(1))) uint16_t*)p;
    } else if constexpr (TB == 4) {
        if constexpr (COH)
            r.w[0] = __hip_atomic_load((__attribute__((address_space(1))) uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            r.w[0] = *(const __attribute__((address_space(1))) uint32_t*)p;
    } else if constexpr (TB == 8) {
        uint64_t x;
        if constexpr (COH)
            x = __hip_atomic_load((__attribute__((address_space(1))) uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            x = *(const __attribute__((address_space(1))) uint64_t*)p;
        r.w[0] = (uint32_t)x;
        r.w[1] = (uint32_t)(x >> 32);
    } else {
        static_assert(TB == 16, "row chunk must be 2, 4, 8, 16 or 32 bytes");
        if constexpr (COH) {
            uint64_t a = __hip_atomic_load((__attribute__((address_space(1))) uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint64_t b = __hip_atomic_load((__attribute__((address_space(1))) uint64_t*)p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            r.w[0] = (uint32_t)a; r.w[1] = (uint32_t)(a >> 32);
            r.w[2] = (uint32_t)b; r.w[3] = (uint32_t)(b >> 32);
        } else {
            u32x4 x = *(const __attribute__((address_space(1))) u32x4*)p;
            r.w[0] = x.x; r.w[1] = x.y; r.w[2] = x.z; r.w[3] = x.w;
        }
    }
    return r;
}

template <int TB, bool COH, int SPLIT = 16>
__device__ inline void store_raw(gchar* p, const Words<(TB + 3) / 4>& r)
{
    if constexpr (TB == 32) {
        Words<4> a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a.w[i] = r.w[i];
            b.w[i] = r.w[4 + i];
        }
        store_raw<16, COH>(p, a);
        store_raw<16, COH>(p + SPLIT, b);
    } else if constexpr (TB == 2) {
        static_assert(!COH, "2-byte coherent stores are not supported");
        *(__attribute__((address_space(1))) uint16_t*)p = (uint16_t)r.w[0];
    } else if constexpr (TB == 4) {
        if constexpr (COH)
            __hip_atomic_store((__attribute__((address_space(1))) uint32_t*)p, r.w[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            *(__attribute__((address_space(1))) uint32_t*)p = r.w[0];
    } else if constexpr (TB == 8) {
        uint64_t x = ((uint64_t)r.w[1] << 32) | r.w[0];
        if constexpr (COH)
            __hip_atomic_store((__attribute__((address_space(1))) uint64_t*)p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else
            *(__attribute__((address_space(1))) uint64_t*)p = x;
    } else {
        if constexpr (COH) {
            __hip_atomic_store((__attribute__((address_space(1))) uint64_t*)p, ((uint64_t)r.w[1] << 32) | r.w[0], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store((__attribute__((address_space(1))) uint64_t*)p + 1, ((uint64_t)r.w[3] << 32) | r.w[2], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            *(__attribute__((address_space(1))) u32x4*)p = u32x4{r.w[0], r.w[1], r.w[2], r.w[3]};
        }
    }
}

// The same row chunks through a buffer resource (raw, stride 0) over a
// factor matrix of at most kBufRecords bytes.  A lane whose offset is
// kBufSkip lies out of range: its load returns 0 and its store is dropped,
// with no memory access.  Every lane's instruction is issued either way, so
// the number of loads and stores in flight stays the same on every path and
// the waits for a row can stay partial.
constexpr uint32_t kBufRecords = 0xFFFFFF00u;
constexpr uint32_t kBufSkip = 0xFFFFFF80u;   // out of range for every chunk of a 32-byte access
constexpr int kBufFlags = 0x00020000;   // gfx9 raw buffer: 32-bit data format, no swizzle

__device__ inline __amdgpu_buffer_rsrc_t factor_rsrc(void* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)kBufRecords, kBufFlags);
}

// A 32-byte chunk is two 16-byte pieces SPLIT bytes apart (half a row: the
// lane group's first pieces cover the row's first half contiguously, one
// instruction per half); smaller chunks are one piece.
template <int TB, int SPLIT = 16>
__device__ inline Words<(TB + 3) / 4> load_buf(__amdgpu_buffer_rsrc_t r, uint32_t off)
{
    Words<(TB + 3) / 4> w;
    if constexpr (TB == 32) {
        // the second piece at the scalar offset SPLIT (a skipped lane's
        // offset stays out of range with it)
        const u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        const u32x4 y = __builtin_amdgcn_raw_buffer_load_b128(r, off, SPLIT, 0);
        w.w[0] = x.x; w.w[1] = x.y; w.w[2] = x.z; w.w[3] = x.w;
        w.w[4] = y.x; w.w[5] = y.y; w.w[6] = y.z; w.w[7] = y.w;
    } else if constexpr (TB == 2) {
        w.w[0] = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
    } else if constexpr (TB == 4) {
        w.w[0] = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
    } else if constexpr (TB == 8) {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 x = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        w.w[0] = x.x;
        w.w[1] = x.y;
    } else {
        static_assert(TB == 16, "row chunk must be 2, 4, 8, 16 or 32 bytes");
        const u32x4 x = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
        w.w[0] = x.x; w.w[1] = x.y; w.w[2] = x.z; w.w[3] = x.w;
    }
    return w;
}

template <int TB, int SPLIT = 16>
__device__ inline void store_buf(__amdgpu_buffer_rsrc_t r, uint32_t off, const Words<(TB + 3) / 4>& w)
{
    if constexpr (TB == 32) {
        __builtin_amdgcn_raw_buffer_store_b128(u32x4{w.w[0], w.w[1], w.w[2], w.w[3]}, r, off, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b128(u32x4{w.w[4], w.w[5], w.w[6], w.w[7]}, r, off, SPLIT, 0);
    } else if constexpr (TB == 2) {
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)w.w[0], r, off, 0, 0);
    } else if constexpr (TB == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(w.w[0], r, off, 0, 0);
    } else if constexpr (TB == 8) {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(u32x2{w.w[0], w.w[1]}, r, off, 0, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b128(u32x4{w.w[0], w.w[1], w.w[2], w.w[3]}, r, off, 0, 0);
    }
}

// The same row chunks in LDS (a tile of P rows held by one workgroup).
typedef __attribute__((address_space(3))) char lchar;

template <int TB, int SPLIT = 16>
__device__ inline Words<(TB + 3) / 4> load_lds(const lchar* p)
{
    Words<(TB + 3) / 4> r;
    if constexpr (TB == 32) {
        const u32x4 x = *(const __attribute__((address_space(3))) u32x4*)p;
        const u32x4 y = *(const __attribute__((address_space(3))) u32x4*)(p + SPLIT);
        r.w[0] = x.x; r.w[1] = x.y; r.w[2] = x.z; r.w[3] = x.w;
        r.w[4] = y.x; r.w[5] = y.y; r.w[6] = y.z; r.w[7] = y.w;
    } else if constexpr (TB == 2) {
        r.w[0] = *(const __attribute__((address_space(3))) uint16_t*)p;
    } else if constexpr (TB == 4) {
        r.w[0] = *(const __attribute__((address_space(3))) uint32_t*)p;
    } else if constexpr (TB == 8) {
        const uint64_t x = *(const __attribute__((address_space(3))) uint64_t*)p;
        r.w[0] = (uint32_t)x;
        r.w[1] = (uint32_t)(x >> 32);
    } else {
        static_assert(TB == 16, "row chunk must be 2, 4, 8, 16 or 32 bytes");
        const u32x4 x = *(const __attribute__((address_space(3))) u32x4*)p;
        r.w[0] = x.x; r.w[1] = x.y; r.w[2] = x.z; r.w[3] = x.w;
    }
    return r;
}

template <int TB, int SPLIT = 16>
__device__ inline void store_lds(lchar* p, const Words<(TB + 3) / 4>& r)
{
    if constexpr (TB == 32) {
        *(__attribute__((address_space(3))) u32x4*)p = u32x4{r.w[0], r.w[1], r.w[2], r.w[3]};
        *(__attribute__((address_space(3))) u32x4*)(p + SPLIT) = u32x4{r.w[4], r.w[5], r.w[6], r.w[7]};
    } else if constexpr (TB == 2) {
        *(__attribute__((address_space(3))) uint16_t*)p = (uint16_t)r.w[0];
    } else if constexpr (TB == 4) {
        *(__attribute__((address_space(3))) uint32_t*)p = r.w[0];
    } else if constexpr (TB == 8) {
        *(__attribute__((address_space(3))) uint64_t*)p = ((uint64_t)r.w[1] << 32) | r.w[0];
    } else {
        *(__attribute__((address_space(3))) u32x4*)p = u32x4{r.w[0], r.w[1], r.w[2], r.w[3]};
    }
}

// Element i of a packed chunk, widened to FP32 (row scale applied by caller).
// OCP E4M3 byte `sel` of a word, widened exactly (v_cvt_f32_fp8 on gfx950).
template <int SEL>
__device__ inline float e4m3_byte_to_f32(uint32_t w)
{
    return __builtin_amdgcn_cvt_f32_fp8((int)w, SEL);
}

template <int PREC, int NW>
__device__ inline float unpack(const Words<NW>& r, int i)
{
    if constexpr (PREC == kF32) return __uint_as_float(r.w[i]);
    else if constexpr (PREC == kF16) return f16_to_f32((r.w[i >> 1] >> ((i & 1) * 16)) & 0xffffu);
    else if constexpr (PREC == kBF16 || PREC == kBF16SR) return bf16_to_f32((r.w[i >> 1] >> ((i & 1) * 16)) & 0xffffu);
    else {
        switch (i & 3) {
        case 0: return e4m3_byte_to_f32<0>(r.w[i >> 2]);
        case 1: return e4m3_byte_to_f32<1>(r.w[i >> 2]);
        case 2: return e4m3_byte_to_f32<2>(r.w[i >> 2]);
        default: return e4m3_byte_to_f32<3>(r.w[i >> 2]);
        }
    }
}

// lowbias32 (C. Wellons' hash prospector), a bijective 32-bit mixer.
__host__ __device__ inline uint32_t lowbias32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Random words for stochastic rounding: lane l of the worker's lane group
// draws, for sample j of the window, the seed lowbias32(K ^ (j*0x9E3779B9 +
// l*0x85EBCA77)) (K: Philox word 0 of the window) and steps it with sr_next
// for its further words: one multiply-based hash per sample and lane, then
// three shift-xors and an add per word.  Each word serves two elements of the
// P row and, with its halves swapped, the same two elements of the Q row.
__host__ __device__ inline uint32_t sr_lane_seed(uint32_t key, uint32_t j, uint32_t lane)
{
    return lowbias32(key ^ (j * 0x9E3779B9u + lane * 0x85EBCA77u));
}
// xorshift32 (Marsaglia, 2003) plus a Weyl increment, which also keeps the
// sequence off the all-zero fixed point.
__host__ __device__ inline uint32_t sr_next(uint32_t x)
{
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x + 0x6D2B79F5u;
}

template <int PREC, int NW>
__device__ inline void pack(Words<NW>& r, int i, uint32_t bits)
{
    if constexpr (PREC == kF32) r.w[i] = bits;
    else if constexpr (PREC == kF16 || PREC == kBF16 || PREC == kBF16SR) {
        if ((i & 1) == 0) r.w[i >> 1] = bits;
        else r.w[i >> 1] |= bits << 16;
    } else {
        if ((i & 3) == 0) r.w[i >> 2] = bits;
        else r.w[i >> 2] |= bits << ((i & 3) * 8);
    }
}

}  // namespace cumf
