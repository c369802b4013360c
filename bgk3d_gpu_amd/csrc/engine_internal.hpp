// Internals shared by the handle API (engine.hip) and the serving layer
// (serving.hip): the handle, its layout, the derived collision parameters
// and the arguments of the fused kernel for one step.
#pragma once

#include "../../include/bgk3d_engine.h"

#include <cstdint>
#include <string>
#include <vector>

#include "collide.hpp"
#include "common.hpp"
#include "d3q19.hpp"
#include "reduce.hpp"
#include "storage.hpp"
#include "v1_kernels.hpp"

struct bgk_handle {
    bgk_config cfg{};
    size_t esize = 4;
    int A = 32;  // elements per 128-byte line
    // Layout of a lattice: the population planes of one z level lie next to
    // each other, population d of cell (i,j,k) at
    //   mem[buf] + base + d*plane + i + nx*j + nxy*k,
    // plane = nx*ny, z stride nxy = 19*plane (lattice 1 holds 18 populations
    // and leaves the 19th slot unused).  Rows are padded to whole 128-byte
    // lines and base puts i = 1 on a line start.  A wave's 19 loads and 19
    // stores then fall in two windows of 19 planes instead of 37 arrays
    // scattered over the whole region: fewer DRAM pages and translation
    // entries in flight per step.
    int nx = 0, ny = 0, nz = 0;
    long plane = 0, nxy = 0, lstride = 0, base = 0;
    // compact layout: rows of l + 2 cells (collide_compact), else rows padded
    // to whole lines (collide_fused)
    bool compact = false;
    void* mem[2] = {nullptr, nullptr};
    size_t bytes = 0;
    int cur = 0;
    hipStream_t stream = nullptr;
    double omega = 0, fgrad = 0, lidforce = 0, u_inflow = 0, cte_s = 0;
    // obstacle
    bool has_obs = false;
    int ic = 0, jc = 0, kc = 0, radius = 0;
    int ob[6] = {0, 0, 0, 0, 0, 0};  // link-mask box origin (i,j,k) and extent
    uint32_t* obits = nullptr;
    int mk[6] = {0, 0, 0, 0, 0, 0};  // solid-mask box origin and extent (draglift)
    int8_t* omask = nullptr;
    double* scratch = nullptr;
    double* hscratch = nullptr;
    long long steps = 0;
    bool borrowed = false;  // lattice memory, scratch and stream belong to a server
    std::vector<hipEvent_t> events;

    template <typename S>
    S* arr(int buf, int d) const
    {
        return reinterpret_cast<S*>(mem[buf]) + (long)d * plane + base;
    }
    template <typename S>
    S* cur_a(int d) const
    {
        return d == 18 ? arr<S>(0, 18) : arr<S>(cur, d);
    }
    template <typename S>
    bgk::Lattice19<S> current() const
    {
        bgk::Lattice19<S> L;
        for (int d = 0; d < 19; ++d) L.p[d] = cur_a<S>(d);
        return L;
    }
    bgk::Geometry geom() const
    {
        bgk::Geometry g;
        g.l = cfg.l;
        g.m = cfg.m;
        g.n = cfg.n;
        g.nx = nx;
        g.nxy = nxy;
        return g;
    }
};

namespace bgk {
namespace eng {

// Record an error message for bgk_last_error() and return `code`.
int fail(int code, const char* fmt, ...);

#define ENG_HIP(call)                                                                                    \
    do {                                                                                                 \
        hipError_t _e = (call);                                                                          \
        if (_e != hipSuccess) return ::bgk::eng::fail(BGK_EHIP, "%s: %s", #call, hipGetErrorString(_e)); \
    } while (0)

inline size_t storage_bytes(int kind)
{
    switch (kind) {
        case BGK_F32: return 4;
        case BGK_F64: return 8;
        default: return 2;
    }
}

inline long round_up(long v, long a) { return (v + a - 1) / a * a; }

// Row layout of the automatic choice (layout = 0).  Measured kernel times,
// compact/padded (cavities, Taylor-Green and channels alone, one handle):
//  * FP32 and FP64: 0.94-0.99 from 3.5 % row waste up (n = 100 ... 250),
//    even at 2 %; at zero waste 1.09-1.14 at 64^3, 1.00-1.04 from 96^3 to
//    224^3, 0.94-0.99 from 256^3 up (384^3, 1024 x 384^2, 512 x 512 x 256).
//  * 16-bit storage: 0.86-0.99 from 10 % waste up, even at 6-8 %, 1.04-1.07
//    at 2-4 % (n = 184, 188, 250) and 1.06-1.13 at zero waste (256^3), 0.97
//    at 384^3.
inline bool prefer_compact(const bgk_config& c, double waste)
{
    const bool wide = storage_bytes(c.storage) >= 4;
    if (waste > (wide ? 0.03 : 0.09)) return true;
    return wide && (double)c.l * (double)c.m * (double)c.n >= (double)(1 << 24);
}

// Calls f with a value of the storage type the configuration selects: the
// storage kind, or for the mixed forms the type that carries their
// collision kind (mix32_t, f16r_t).
template <typename F>
int by_storage(const bgk_config& c, F&& f)
{
    if (c.mixed == 2 && c.storage == BGK_F32) return f(mix32_t{});
    if (c.mixed == 1 && c.storage == BGK_F16) return f(f16r_t{});
    switch (c.storage) {
        case BGK_F32: return f(float{});
        case BGK_F64: return f(double{});
        case BGK_F16: return f(f16_t{});
        case BGK_BF16: return f(bf16_t{});
    }
    return fail(BGK_EINVAL, "unknown storage type %d", c.storage);
}

// The collision runs in FP64 (DOUBLE_P storage, or the MIXED2 form).
inline bool compute_double(const bgk_config& c) { return c.storage == BGK_F64 || c.mixed == 2; }

// Validate `c` and fill the layout and the derived parameters of `h`
// (no device call, no allocation).
int handle_setup(bgk_handle& h, const bgk_config& c);

// Device bytes of the two lattices of a handle set up by handle_setup.
inline size_t lattice_bytes(const bgk_handle& h) { return (size_t)h.lstride * 2 * h.esize; }

// Point mem[0] and mem[1] (the two lattices) into one region of
// lattice_bytes(h) bytes.
inline void place_lattices(bgk_handle& h, void* region)
{
    h.mem[0] = region;
    h.mem[1] = static_cast<char*>(region) + (size_t)h.lstride * h.esize;
    h.bytes = lattice_bytes(h);
}

// Stochastic-rounding seed of the step that follows `steps` completed steps.
__host__ __device__ inline uint32_t step_seed(uint32_t cfg_seed, long long steps)
{
    return hash32(cfg_seed ^ hash32((uint32_t)steps * 0x9e3779b9U + 1u));
}

// Arguments of the fused kernel for the step that reads lattice `cur`.
template <typename S>
FusedArgs<S> fused_args(const bgk_handle& h, int cur)
{
    typedef typename Storage<S>::compute C;
    FusedArgs<S> p{};
    for (int d = 0; d < 19; ++d) p.a[d] = d == 18 ? h.arr<S>(0, 18) : h.arr<S>(cur, d);
    for (int d = 0; d < 18; ++d) p.b[d] = h.arr<S>(1 - cur, d);
    p.l = h.cfg.l;
    p.m = h.cfg.m;
    p.n = h.cfg.n;
    p.nx = h.nx;
    p.nxy = h.nxy;
    p.omega = (C)h.omega;
    p.fgrad = (C)h.fgrad;
    p.fx = p.fy = p.fz = C(0);
    p.lidforce = (C)h.lidforce;
    p.cte_s = (C)h.cte_s;
    p.cst = CollideConsts<C>::shifted();
    p.obits = h.obits;
    p.oi0 = h.ob[0];
    p.oj0 = h.ob[1];
    p.ok0 = h.ob[2];
    p.oni = h.ob[3];
    p.onj = h.ob[4];
    p.onk = h.ob[5];
    p.align = 0;
    // padded rows start on a 128-byte line; with the x boundary folded into
    // the pull the cells after i = l up to that line's end are never read.
    // Compact rows: the positions the waves of a plane cover.
    if (h.compact)
        p.lend = (int)round_up((long)h.nx * h.cfg.m, h.A);
    else
        p.lend = h.cfg.boundary == BGK_BC_INFLOW ? 0 : (int)round_up(h.cfg.l, h.A);
    p.seed = step_seed(h.cfg.seed, h.steps);
    p.pstride = h.plane;
    p.uin = (C)h.u_inflow;
    return p;
}

}  // namespace eng
}  // namespace bgk
