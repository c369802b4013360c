// Population storage types and their conversion to the collision kind.
//
// float   -> float   (reference SINGLE)
// double  -> double  (reference DOUBLE_P)
// f16     -> float   (_Float16 storage of f - w, FP32 collision, lean form)
// bf16    -> float   (bfloat16 storage of f - w, FP32 collision, lean form)
// mix32_t -> double  (FP32 storage, FP64 collision: the reference's MIXED2
//                     build, SRC/precision.F90:38-43)
// f16r_t  -> float   (FP16 storage, FP32 collision in the reference's
//                     expression order: its MIXED1 build, :30-36)
//
// Stored values are the reference's shifted populations f_i - w_i
// (SRC/col_MC.F90:139-144), which keeps them near zero where low-precision
// formats have the most resolution.  Low-precision stores round to nearest
// even, or stochastically (unbiased) with counter-based random words so that
// a run is reproducible and independent of scheduling.  load_nt is the
// streaming form used by the time step (each value is read once per step).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bgk {

struct bf16_t {
    uint16_t u;
};
typedef _Float16 f16_t;
// Types of their own for the mixed forms, so the kernels select the
// collision kind and form at compile time; the bits stored are a float and
// a _Float16.
struct mix32_t {
    float v;
};
struct f16r_t {
    _Float16 v;
};

enum StorageKind : int { ST_F32 = 0, ST_F64 = 1, ST_F16 = 2, ST_BF16 = 3, ST_MIX32 = 4, ST_F16R = 5 };

// Every lattice lives in device (global) memory.  Loads and stores go
// through a pointer in that address space, so they compile to global memory
// instructions even where the pointer itself was read from memory (the
// serving kernel's per-request tables), which the compiler would otherwise
// treat as generic (flat) addresses.
#define BGK_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const BGK_GLOBAL T* gmem(const T* p)
{
    return (const BGK_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ BGK_GLOBAL T* gmem(T* p)
{
    return (BGK_GLOBAL T*)p;
}

// load / load_nt: a stored value in the collision kind (plain or streaming
// read); widen: the same for a value in a register.  lean: the 16-bit lean
// collide (bgk_collide_lean) instead of the reference-order one.
template <typename S>
struct Storage;

template <>
struct Storage<float> {
    typedef float compute;
    static constexpr int kind = ST_F32;
    static constexpr bool lean = false;
    __device__ __forceinline__ static float load(const float* p) { return *gmem(p); }
    __device__ __forceinline__ static float load_nt(const float* p) { return __builtin_nontemporal_load(gmem(p)); }
    __device__ __forceinline__ static float widen(float v) { return v; }
};
template <>
struct Storage<double> {
    typedef double compute;
    static constexpr int kind = ST_F64;
    static constexpr bool lean = false;
    __device__ __forceinline__ static double load(const double* p) { return *gmem(p); }
    __device__ __forceinline__ static double load_nt(const double* p) { return __builtin_nontemporal_load(gmem(p)); }
    __device__ __forceinline__ static double widen(double v) { return v; }
};
template <>
struct Storage<f16_t> {
    typedef float compute;
    static constexpr int kind = ST_F16;
    static constexpr bool lean = true;
    __device__ __forceinline__ static float load(const f16_t* p) { return (float)*gmem(p); }
    __device__ __forceinline__ static float load_nt(const f16_t* p) { return (float)__builtin_nontemporal_load(gmem(p)); }
    __device__ __forceinline__ static float widen(f16_t v) { return (float)v; }
    __device__ __forceinline__ static f16_t raw(const f16_t* p) { return *gmem(p); }
    __device__ __forceinline__ static f16_t raw_nt(const f16_t* p) { return __builtin_nontemporal_load(gmem(p)); }
};
template <>
struct Storage<bf16_t> {
    typedef float compute;
    static constexpr int kind = ST_BF16;
    static constexpr bool lean = true;
    __device__ __forceinline__ static float load(const bf16_t* p)
    {
        return __uint_as_float((uint32_t)gmem(p)->u << 16);
    }
    __device__ __forceinline__ static float load_nt(const bf16_t* p)
    {
        return __uint_as_float((uint32_t)__builtin_nontemporal_load(&gmem(p)->u) << 16);
    }
    __device__ __forceinline__ static float widen(bf16_t v) { return __uint_as_float((uint32_t)v.u << 16); }
    __device__ __forceinline__ static bf16_t raw(const bf16_t* p) { return bf16_t{gmem(p)->u}; }
    __device__ __forceinline__ static bf16_t raw_nt(const bf16_t* p)
    {
        return bf16_t{__builtin_nontemporal_load(&gmem(p)->u)};
    }
};
template <>
struct Storage<mix32_t> {
    typedef double compute;
    static constexpr int kind = ST_MIX32;
    static constexpr bool lean = false;
    __device__ __forceinline__ static double load(const mix32_t* p) { return (double)gmem(p)->v; }
    __device__ __forceinline__ static double load_nt(const mix32_t* p)
    {
        return (double)__builtin_nontemporal_load(&gmem(p)->v);
    }
    __device__ __forceinline__ static double widen(mix32_t v) { return (double)v.v; }
};
template <>
struct Storage<f16r_t> {
    typedef float compute;
    static constexpr int kind = ST_F16R;
    static constexpr bool lean = false;
    __device__ __forceinline__ static float load(const f16r_t* p) { return (float)gmem(p)->v; }
    __device__ __forceinline__ static float load_nt(const f16r_t* p)
    {
        return (float)__builtin_nontemporal_load(&gmem(p)->v);
    }
    __device__ __forceinline__ static float widen(f16r_t v) { return (float)v.v; }
    __device__ __forceinline__ static f16r_t raw(const f16r_t* p) { return f16r_t{gmem(p)->v}; }
    __device__ __forceinline__ static f16r_t raw_nt(const f16r_t* p)
    {
        return f16r_t{__builtin_nontemporal_load(&gmem(p)->v)};
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

// Random word of store d (0..18) of a cell whose hash is r0: a Weyl sequence
// with step 2^32 / golden ratio.  Every word is uniform when r0 is, and the
// high bits of one cell's 19 words are spread evenly over [0, 2^32): one
// integer add per store (a multiplicative generator costs a quarter-rate
// 32-bit multiply per store).
__device__ __forceinline__ uint32_t sr_word(uint32_t r0, int d) { return r0 + (uint32_t)(d + 1) * 0x9e3779b9u; }

// Round a value of the collision kind to the storage type.  SR = 0:
// nearest-even; SR = 1: stochastic with the 32-bit random word r.
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
// Stochastic rounding uses gfx950's conversions with a random operand
// (v_cvt_sr_f16_f32, v_cvt_sr_bf16_f32): one instruction per store.  The
// conversion draws on the high bits of the word (with a word below 2^16 it
// never rounds up; with k << 16, k uniform, it rounds 1 + 0.72 ulp up in
// 72.0 % of cases for FP16 and BF16).
typedef _Float16 sr_half2 __attribute__((ext_vector_type(2)));
typedef __bf16 sr_bf162 __attribute__((ext_vector_type(2)));

template <>
struct Rounder<f16_t, 1> {
    __device__ __forceinline__ static f16_t cvt(float v, uint32_t r)
    {
        const sr_half2 z = {0, 0};
        return __builtin_amdgcn_cvt_sr_f16_f32(z, v, r, false)[0];
    }
};
template <>
struct Rounder<bf16_t, 0> {
    // nearest-even through the hardware conversion (v_cvt_pk_bf16_f32)
    __device__ __forceinline__ static bf16_t cvt(float v, uint32_t)
    {
        return bf16_t{__builtin_bit_cast(uint16_t, (__bf16)v)};
    }
};
template <>
struct Rounder<bf16_t, 1> {
    __device__ __forceinline__ static bf16_t cvt(float v, uint32_t r)
    {
        const sr_bf162 z = {0, 0};
        const __bf16 h = __builtin_amdgcn_cvt_sr_bf16_f32(z, v, r, false)[0];
        return bf16_t{__builtin_bit_cast(uint16_t, h)};
    }
};

// FP64 to FP32 with nearest-even rounding (the reference's assignment of a
// double expression to a single-precision array element).
template <int SR>
struct Rounder<mix32_t, SR> {
    __device__ __forceinline__ static mix32_t cvt(double v, uint32_t) { return mix32_t{(float)v}; }
};
template <>
struct Rounder<f16r_t, 0> {
    __device__ __forceinline__ static f16r_t cvt(float v, uint32_t) { return f16r_t{(f16_t)v}; }
};
template <>
struct Rounder<f16r_t, 1> {
    __device__ __forceinline__ static f16r_t cvt(float v, uint32_t r) { return f16r_t{Rounder<f16_t, 1>::cvt(v, r)}; }
};

// Stores (plain and non-temporal); bf16_t and the mixed types are stored
// through their member.
template <typename S>
__device__ __forceinline__ void store_value(S* p, S v) { *gmem(p) = v; }
template <>
__device__ __forceinline__ void store_value<bf16_t>(bf16_t* p, bf16_t v) { gmem(p)->u = v.u; }
template <typename S>
__device__ __forceinline__ void store_nt(S* p, S v) { __builtin_nontemporal_store(v, gmem(p)); }
template <>
__device__ __forceinline__ void store_nt<bf16_t>(bf16_t* p, bf16_t v)
{
    __builtin_nontemporal_store(v.u, &gmem(p)->u);
}
template <>
__device__ __forceinline__ void store_value<mix32_t>(mix32_t* p, mix32_t v) { gmem(p)->v = v.v; }
template <>
__device__ __forceinline__ void store_nt<mix32_t>(mix32_t* p, mix32_t v)
{
    __builtin_nontemporal_store(v.v, &gmem(p)->v);
}
template <>
__device__ __forceinline__ void store_value<f16r_t>(f16r_t* p, f16r_t v) { gmem(p)->v = v.v; }
template <>
__device__ __forceinline__ void store_nt<f16r_t>(f16r_t* p, f16r_t v)
{
    __builtin_nontemporal_store(v.v, &gmem(p)->v);
}

}  // namespace bgk
