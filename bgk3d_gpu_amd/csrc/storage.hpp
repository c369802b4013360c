// Population storage types and their conversion to the collision kind.
//
// float  -> float   (reference SINGLE)
// double -> double  (reference DOUBLE_P)
// f16    -> float   (_Float16 storage of f - w, FP32 collision)
// bf16   -> float   (bfloat16 storage of f - w, FP32 collision)
//
// Stored values are the reference's shifted populations f_i - w_i
// (SRC/col_MC.F90:139-144), which keeps them near zero where low-precision
// formats have the most resolution.  Low-precision stores round to nearest
// even, or stochastically (unbiased) with a counter-based hash so that a run
// is reproducible and independent of scheduling.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bgk {

struct bf16_t {
    uint16_t u;
};
typedef _Float16 f16_t;

enum StorageKind : int { ST_F32 = 0, ST_F64 = 1, ST_F16 = 2, ST_BF16 = 3 };

template <typename S>
struct Storage;

template <>
struct Storage<float> {
    typedef float compute;
    static constexpr int kind = ST_F32;
    __device__ __forceinline__ static float load(const float* p) { return *p; }
};
template <>
struct Storage<double> {
    typedef double compute;
    static constexpr int kind = ST_F64;
    __device__ __forceinline__ static double load(const double* p) { return *p; }
};
template <>
struct Storage<f16_t> {
    typedef float compute;
    static constexpr int kind = ST_F16;
    __device__ __forceinline__ static float load(const f16_t* p) { return (float)*p; }
};
template <>
struct Storage<bf16_t> {
    typedef float compute;
    static constexpr int kind = ST_BF16;
    __device__ __forceinline__ static float load(const bf16_t* p)
    {
        return __uint_as_float((uint32_t)p->u << 16);
    }
};

__host__ __device__ inline uint32_t hash32(uint32_t x)
{
    // integer finaliser (xor-shift-multiply), full avalanche
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Round a float to the storage type.  SR = 0: nearest-even; SR = 1:
// stochastic with the 32-bit random word r (only its high bits are used,
// which is where the per-store LCG stream below is strongest).
__device__ __forceinline__ uint32_t lcg_next(uint32_t r) { return r * 747796405u + 2891336453u; }
template <typename S, int SR>
struct Rounder;

template <int SR>
struct Rounder<float, SR> {
    __device__ __forceinline__ static float cvt(float v, uint32_t) { return v; }
};
template <int SR>
struct Rounder<double, SR> {
    __device__ __forceinline__ static double cvt(double v, uint32_t) { return v; }
};
template <>
struct Rounder<f16_t, 0> {
    __device__ __forceinline__ static f16_t cvt(float v, uint32_t) { return (f16_t)v; }
};
template <>
struct Rounder<f16_t, 1> {
    // Unbiased stochastic rounding to FP16 without branches: add a uniform
    // fraction of the FP16 spacing at |v| (2^(e-10) in the normal range,
    // 2^-24 below it; 13 random bits, so the sum is exact in FP32 for the
    // normal range) away from zero, then convert with round-toward-zero
    // (v_cvt_pkrtz_f16_f32), which truncates to the spacing in the normal
    // and the subnormal range alike.  |v| >= 2^15 converts to nearest.
    __device__ __forceinline__ static f16_t cvt(float v, uint32_t r)
    {
        const uint32_t e = __float_as_uint(v) & 0x7f800000u;
        const float ulp = __uint_as_float((e > 0x38800000u ? e : 0x38800000u) - 0x05000000u);
        const float du = (float)(r >> 19) * 1.220703125e-4f * ulp;
        const float t = v + __builtin_copysignf(du, v);
        const f16_t sr = __builtin_amdgcn_cvt_pkrtz(t, t)[0];
        return e >= 0x47000000u ? (f16_t)v : sr;
    }
};
template <>
struct Rounder<bf16_t, 0> {
    __device__ __forceinline__ static bf16_t cvt(float v, uint32_t)
    {
        uint32_t u = __float_as_uint(v);
        if ((u & 0x7f800000U) == 0x7f800000U) return bf16_t{(uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0))};
        u += 0x7fffU + ((u >> 16) & 1U);
        return bf16_t{(uint16_t)(u >> 16)};
    }
};
template <>
struct Rounder<bf16_t, 1> {
    __device__ __forceinline__ static bf16_t cvt(float v, uint32_t r)
    {
        uint32_t u = __float_as_uint(v);
        if ((u & 0x7f800000U) == 0x7f800000U) return bf16_t{(uint16_t)(u >> 16)};
        u += (r >> 16);  // add a uniform fraction of one bf16 ulp (top bits of r), then truncate
        return bf16_t{(uint16_t)(u >> 16)};
    }
};

// Stores (plain and non-temporal); bf16_t is stored through its bits.
template <typename S>
__device__ __forceinline__ void store_value(S* p, S v) { *p = v; }
template <>
__device__ __forceinline__ void store_value<bf16_t>(bf16_t* p, bf16_t v) { reinterpret_cast<uint16_t*>(p)[0] = v.u; }
template <typename S>
__device__ __forceinline__ void store_nt(S* p, S v) { __builtin_nontemporal_store(v, p); }
template <>
__device__ __forceinline__ void store_nt<bf16_t>(bf16_t* p, bf16_t v)
{
    __builtin_nontemporal_store(v.u, reinterpret_cast<uint16_t*>(p));
}

}  // namespace bgk
