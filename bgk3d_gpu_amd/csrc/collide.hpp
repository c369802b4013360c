// Fused pull-stream + BGK collide with the boundary conditions folded into
// the pull -- the one hot kernel behind both C-ABIs.
//
// One step of SRC/col_MC.F90:126-536 for every interior cell.  Population d
// of cell x is pulled from x - c_d.  When that source lies outside the
// interior, the value the reference's boundary routine would have left in
// the halo is produced directly:
//   AX_HALO      read the halo as it is (the caller filled it: reference
//                semantics of bcond_*_gpu, and the engine's inflow/outflow
//                faces);
//   AX_PERIODIC  wrap the coordinate (SRC/bcond.periodic.F90:105-252);
//   AX_WALL      half-way bounce-back: the cell's own opposite population
//                (SRC/bcond.driven.F90:105-247, SRC/bcond.channel.F90:255-267);
//   AX_LID       as AX_WALL, and a pull across the top wall z = n+1 adds
//                +force to direction 2 and -force to direction 11
//                (SRC/bcond.driven.F90:243-247; the z face is written last
//                there, so its value wins on shared edges).
// A wall on any crossed axis takes precedence over wrapping, as in the
// reference, where the wall faces are written after the periodic ones.
// Obstacles: bit d of the cell's link mask set means x - c_d is solid and
// the pull returns the own opposite population -- the value the reference's
// bcond_obs stores in the solid cell (SRC/bcond.obs.F90:103-121).
//
// Layout: population d is an array (0:nx-1, 0:ny-1, 0:nz-1), element
// (i,j,k) at i + nx*j + nxy*k.  A wave of 64 lanes covers 64 consecutive
// cells of one row (j,k).  If `align` is non-zero the row's first wave starts
// on the `align`-element boundary at or before i = 1 (computed from the
// output address), so that every wave's loads and stores of the same row
// hit whole 128-byte lines whatever layout the caller chose; the engine's own
// layout puts i = 1 on that boundary, making the shift zero.  Stores and the
// loads of the directions with c_x = 0 are non-temporal (streaming: read or
// written once per step); the +-x shifted loads keep the default policy, so
// the line two neighbouring waves of a row share is found in L2 by the
// second.
#pragma once

#include <type_traits>
#include <utility>

#include "common.hpp"
#include "d3q19.hpp"
#include "storage.hpp"

namespace bgk {

enum AxisBC : int { AX_HALO = 0, AX_PERIODIC = 1, AX_WALL = 2, AX_LID = 3 };

template <typename S>
struct FusedArgs {
    typedef typename Storage<S>::compute C;
    S* a[19];  // current lattice (read), a[18] updated in place
    S* b[18];  // next lattice, directions 1..18
    int l, m, n, nx;
    long nxy;
    C omega, fgrad, fx, fy, fz, lidforce;
    C cte_s;         // Smagorinsky constant (LES variants)
    CollideConsts<C> cst;
    // obstacle link masks over the box [oi0, oi0+oni) x ... (bit d: x - c_d solid)
    const uint32_t* obits;
    int oi0, oj0, ok0, oni, onj, onk;
    int align;       // 0: layout already aligned; else elements per 128-byte line
    int lend;        // > l: lanes l < i <= lend store zeros (see fused_rows); 0: off
    uint32_t seed;   // stochastic rounding stream (step-dependent)
    // Distance in elements between populations d and d+1 of one lattice
    // (library-owned layout; 0 for the reference ABI, which never uses it).
    long pstride;
    // Non-finite flag (serving): set to 1 when a cell's new rest population
    // is not finite -- a non-finite density or velocity anywhere in the
    // cell's inputs makes it so.  Null: no check.
    uint32_t* bad;
    // Inflow velocity: compact rows under the inflow boundary write both x
    // faces from the halo lanes (compact_cells).
    C uin;
};

// Raise the non-finite flag of p (if any) when the rest population v is not
// finite.  Only lanes that see one store, so a healthy step stores nothing.
template <typename S>
__device__ __forceinline__ void flag_nonfinite(const FusedArgs<S>& p, typename Storage<S>::compute v)
{
    if (p.bad && !__builtin_isfinite(v)) *gmem(p.bad) = 1u;
}

// Shifted equilibria of the inflow boundary's x faces
// (SRC/bcond.inflow.F90:131-181, rho = 1, cte1 = 0).  Inflow face x = 0,
// directions 1..5: u = (u_inflow, 0, 0), the same for every (j, k).
template <typename S>
__device__ __forceinline__ void inflow_face(typename Storage<S>::compute u_inflow, S (&v)[5])
{
    typedef typename Storage<S>::compute R;
    const R p1 = W<R>::p1, p2 = W<R>::p2, rf = R(3), qf = R(1.5), three = R(3), crho = R(1), cte1 = R(0);
    const R xj = u_inflow, yj = R(0), zj = R(0);
    const R cvsq = xj * xj + yj * yj + zj * zj;
    const R cx01 = rf * (xj - yj) + qf * (three * (xj - yj) * (xj - yj) - cvsq);
    const R cx02 = rf * (xj - zj) + qf * (three * (xj - zj) * (xj - zj) - cvsq);
    const R cx03 = rf * (xj + yj) + qf * (three * (xj + yj) * (xj + yj) - cvsq);
    const R cx04 = rf * (xj + zj) + qf * (three * (xj + zj) * (xj + zj) - cvsq);
    const R cx05 = rf * (xj) + qf * (three * (xj) * (xj)-cvsq);
    v[0] = Rounder<S, 0>::cvt(crho * p2 * (cte1 + cx01), 0u);
    v[1] = Rounder<S, 0>::cvt(crho * p2 * (cte1 + cx02), 0u);
    v[2] = Rounder<S, 0>::cvt(crho * p2 * (cte1 + cx03), 0u);
    v[3] = Rounder<S, 0>::cvt(crho * p2 * (cte1 + cx04), 0u);
    v[4] = Rounder<S, 0>::cvt(crho * p1 * (cte1 + cx05), 0u);
}

// Outflow face x = l+1, directions 10..14: rho = 1 and the momentum of cell
// (l, j, k), whose stored populations 1..18 are f[0..17].
template <typename S>
__device__ __forceinline__ void outflow_face(const typename Storage<S>::compute (&f)[19], S (&v)[5])
{
    typedef typename Storage<S>::compute R;
    const R p1 = W<R>::p1, p2 = W<R>::p2, rf = R(3), qf = R(1.5), three = R(3), crho = R(1), rhoinv = R(1),
            cte1 = R(0);
    const R xj = +(f[0] + f[1] + f[2] + f[3] + f[4] - f[9] - f[10] - f[11] - f[12] - f[13]) * rhoinv;
    const R yj = +(f[2] + f[6] + f[7] + f[8] + f[11] - f[0] - f[9] - f[15] - f[16] - f[17]) * rhoinv;
    const R zj = +(f[3] + f[5] + f[6] + f[12] + f[17] - f[1] - f[8] - f[10] - f[14] - f[15]) * rhoinv;
    const R cvsq = xj * xj + yj * yj + zj * zj;
    const R cx10 = rf * (-xj - yj) + qf * (three * (xj + yj) * (xj + yj) - cvsq);
    const R cx11 = rf * (-xj - zj) + qf * (three * (xj + zj) * (xj + zj) - cvsq);
    const R cx12 = rf * (-xj + yj) + qf * (three * (xj - yj) * (xj - yj) - cvsq);
    const R cx13 = rf * (-xj + zj) + qf * (three * (xj - zj) * (xj - zj) - cvsq);
    const R cx14 = rf * (-xj) + qf * (three * (xj) * (xj)-cvsq);
    v[0] = Rounder<S, 0>::cvt(crho * p2 * (cte1 + cx10), 0u);
    v[1] = Rounder<S, 0>::cvt(crho * p2 * (cte1 + cx11), 0u);
    v[2] = Rounder<S, 0>::cvt(crho * p2 * (cte1 + cx12), 0u);
    v[3] = Rounder<S, 0>::cvt(crho * p2 * (cte1 + cx13), 0u);
    v[4] = Rounder<S, 0>::cvt(crho * p1 * (cte1 + cx14), 0u);
}

// Whether the compact collide writes both x faces of the inflow boundary
// itself: the forms that hold all 19 new populations before storing them
// (every form but the streaming FP64 collide).  That one leaves the faces
// to the step's outflow kernel.
template <typename S, int FMODE>
constexpr bool kFillsInflowFaces = !(std::is_same<typename Storage<S>::compute, double>::value && (FMODE & 4) == 0);

// Base of population d of the lattice read (A) or written (B) this step.
// CMP (compact): population d < 18 of a lattice at a[0] + d*pstride or
// b[0] + d*pstride (a[18] is its own pointer), so 3 pointers and a stride
// live in scalar registers instead of 37 pointers.  Used by the 16-bit
// forms without a per-cell body force: there it frees 16 scalar registers
// (8 waves per SIMD instead of 7) and removes the per-lane 64-bit store
// addresses, 4-9 % faster; with the body force the compiler then held
// 84-102 vector registers, so those keep the pointer table.
template <bool CMP, typename S>
__device__ __forceinline__ const S* pop_a(const FusedArgs<S>& p, int d)
{
    if constexpr (CMP)
        return d == 18 ? p.a[18] : p.a[0] + (long)d * p.pstride;
    else
        return p.a[d];
}
template <bool CMP, typename S>
__device__ __forceinline__ S* pop_b(const FusedArgs<S>& p, int d)
{
    if constexpr (CMP)
        return d == 18 ? p.a[18] : p.b[0] + (long)d * p.pstride;
    else
        return d < 18 ? p.b[d] : p.a[18];
}
template <typename S, int FMODE>
inline constexpr bool kCompact = sizeof(S) == 2 && (FMODE & 3) == 0;


template <int AX>
__device__ __forceinline__ void axis_fix(int cd, int s, int lo_hi_n, int& shift, bool& bounce, bool& top)
{
    // s = coordinate of the source along this axis, valid range 1..lo_hi_n
    (void)cd;
    if (AX == AX_HALO) return;
    const bool below = s < 1, above = s > lo_hi_n;
    if (!(below || above)) return;
    if (AX == AX_PERIODIC) {
        shift = below ? +1 : -1;
    } else {
        bounce = true;
        if (AX == AX_LID && above) top = true;
    }
}

// Calls f(std::integral_constant<int, d>) for d = 0..18: the direction is a
// compile-time constant in the body, so per-direction choices (the load's
// cache policy below) cannot be merged across directions by the compiler.
template <typename F, int... D>
__device__ __forceinline__ void for_each_dir_(F&& f, std::integer_sequence<int, D...>)
{
    (f(std::integral_constant<int, D>{}), ...);
}
template <typename F>
__device__ __forceinline__ void for_each_dir(F&& f)
{
    for_each_dir_(f, std::make_integer_sequence<int, 19>{});
}

// Load of one pulled population.  Directions with c_x = 0 read each of
// their lines exactly once in a step (streaming: non-temporal); the +-x
// shifted directions share a line between neighbouring waves of a row, so
// they use the default policy and the second wave finds it in L2.
template <typename S, int CX>
__device__ __forceinline__ typename Storage<S>::compute pull_load(const S* ptr)
{
    if constexpr (CX == 0)
        return Storage<S>::load_nt(ptr);
    else
        return Storage<S>::load(ptr);
}
// The same load without the widening (16-bit storage).
template <typename S, int CX>
__device__ __forceinline__ S pull_raw(const S* ptr)
{
    if constexpr (CX == 0)
        return Storage<S>::raw_nt(ptr);
    else
        return Storage<S>::raw(ptr);
}

// The moving lid: a pull across the top wall adds +force to direction 2
// and -force to direction 11 (d = 1 and 10 here).  The reference stores that
// sum in the halo (SRC/bcond.driven.F90:243-247), so it is rounded to the
// storage type before the collision reads it: exact for FP32 and FP64 with
// their own arithmetic, the reference's rounding for the mixed forms.  The
// lean 16-bit forms keep the FP32 sum: rounded to BF16 it moves the lid's
// momentum input by up to half a unit (0.2 %), and an LES cavity's velocity
// error against FP64 rose past 5 %.
template <typename S, int d>
__device__ __forceinline__ typename Storage<S>::compute lid_value(typename Storage<S>::compute v, bool top,
                                                                  typename Storage<S>::compute force)
{
    if constexpr (d == 1 || d == 10) {
        if (top) {
            const typename Storage<S>::compute w = d == 1 ? v + force : v - force;
            if constexpr (Storage<S>::lean)
                return w;
            else
                return Storage<S>::widen(Rounder<S, 0>::cvt(w, 0u));
        }
    }
    return v;
}

// EDGE = the row (j,k) touches a y or z boundary; otherwise only the x
// direction can leave the interior and the j/k checks are compiled out.
//
// Addressing: j and k are the same for a whole wave, so the row's offset and
// everything the y and z boundaries decide (wrap, wall) are scalar.  For the
// 16-bit forms, whose kernels are bound by instruction issue, each lane then
// adds only its own byte offset i*sizeof(S) (plus an x wrap): every load and
// store is a scalar base + one 32-bit lane offset (saddr form) and no 64-bit
// address arithmetic runs per lane, except for the directions where a lane
// may bounce off an x wall or a solid cell, where each lane picks its own
// address.  The 32/64-bit forms are bound by HBM and keep per-lane 64-bit
// addresses, which leaves the scalar registers free for higher occupancy.
// NT0: the c_x = 0 directions load non-temporally (the default; right when
// the y and z shifts keep every wave on whole lines, as in the engine's
// layout).  Off when rows or planes do not start on a line: a y or z shifted
// line is then shared by two neighbouring waves, and a streaming load would
// fetch it twice.
template <typename S, int BX, int BY, int BZ, bool OBST, bool EDGE, bool CMP, bool NT0 = true>
__device__ __forceinline__ void pull_all(const FusedArgs<S>& p, int i, int j, int k, long row,
                                         typename Storage<S>::compute (&x)[19])
{
    typedef typename Storage<S>::compute C;
    constexpr long ES = (long)sizeof(S);
    const long rowb = row * ES;                        // scalar: byte offset of cell (0, j, k)
    const uint32_t vi = (uint32_t)i * (uint32_t)ES;    // per lane
    uint32_t ob = 0;
    if (OBST) {
        const unsigned ri = (unsigned)(i - p.oi0), rj = (unsigned)(j - p.oj0), rk = (unsigned)(k - p.ok0);
        if (ri < (unsigned)p.oni && rj < (unsigned)p.onj && rk < (unsigned)p.onk)
            ob = p.obits[(long)ri + (long)p.oni * ((long)rj + (long)p.onj * rk)];
    }
    [[maybe_unused]] S raw[19];      // 16-bit storage: the loaded values before widening
    [[maybe_unused]] bool tops[19];  // 16-bit storage: pulls across the moving lid
    for_each_dir([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        constexpr int cx = cx_of(d), cy = cy_of(d), cz = cz_of(d);
        long ds = -((long)p.nx * cy + p.nxy * cz);  // scalar element shift (y, z)
        int dx = -cx;                                // per-lane element shift (x)
        bool bounce = false, wall = false, top = false;
        if (cx != 0) {
            int sh = 0;
            axis_fix<BX>(cx, i - cx, p.l, sh, bounce, top);
            if (BX == AX_PERIODIC) dx += sh * p.l;
        }
        if (EDGE && cy != 0) {
            int sh = 0;
            axis_fix<BY>(cy, j - cy, p.m, sh, wall, top);
            if (BY == AX_PERIODIC) ds += (long)sh * p.m * p.nx;
        }
        if (EDGE && cz != 0) {
            int sh = 0;
            axis_fix<BZ>(cz, k - cz, p.n, sh, wall, top);
            if (BZ == AX_PERIODIC) ds += (long)sh * p.n * p.nxy;
        }
        if (OBST && d < 18) bounce = bounce || ((ob >> d) & 1u);
        if constexpr (sizeof(S) >= 4) {
            // 32/64-bit storage is bound by HBM, not by issue: per-lane 64-bit
            // addresses keep the scalar registers free (higher occupancy)
            const long s = row + i;
            const long src = s + ds + dx;
            const bool b = wall || bounce;
            C v = pull_load<S, NT0 ? cx : 1>(b ? (p.a[opp_of(d)] + s) : (p.a[d] + src));
            if (BZ == AX_LID && EDGE) v = lid_value<S, d>(v, top, p.lidforce);
            x[d] = v;
        } else {
            // 16-bit storage: every raw load is issued here and widened after
            // the scheduling barrier below, so the 19 loads are in flight
            // together (left to itself the scheduler has reused one register
            // for all of them and waited for each load in turn, which doubled
            // the time of some boundary forms)
            const char* own = reinterpret_cast<const char*>(pop_a<CMP>(p, opp_of(d))) + rowb;
            if (wall) {  // y or z wall: the whole wave bounces
                raw[d] = pull_raw<S, NT0 ? cx : 1>(reinterpret_cast<const S*>(own + (uint64_t)vi));
            } else {
                const char* base = reinterpret_cast<const char*>(pop_a<CMP>(p, d)) + rowb + ds * ES;
                const S* nptr = reinterpret_cast<const S*>(base + (uint64_t)(vi + (uint32_t)dx * (uint32_t)ES));
                if ((BX == AX_WALL && cx != 0) || (OBST && d < 18)) {
                    // a lane may bounce: pick its address, one load
                    const S* optr = reinterpret_cast<const S*>(own + (uint64_t)vi);
                    raw[d] = pull_raw<S, NT0 ? cx : 1>(bounce ? optr : nptr);
                } else {
                    raw[d] = pull_raw<S, NT0 ? cx : 1>(nptr);
                }
            }
            tops[d] = top;
        }
    });
    if constexpr (sizeof(S) == 2) {
        __builtin_amdgcn_sched_barrier(0);
        for_each_dir([&](auto dc) {
            constexpr int d = decltype(dc)::value;
            C v = Storage<S>::widen(raw[d]);
            if (BZ == AX_LID && EDGE) v = lid_value<S, d>(v, tops[d], p.lidforce);
            x[d] = v;
        });
    }
}

// Work mapping.  wpr = waves needed per (possibly shifted) row.  A block
// holds whole rows whenever it can, so the cache line two neighbouring
// 64-cell segments share for the +-x pulls is fetched once, by one CU:
//   4 <= wpr <= 8   one row per block, wpr waves;
//   wpr > 8         4 consecutive segments of one row per block;
//   wpr < 4         4/wpr rows per block.
struct RowMap {
    int wpr;   // waves per row
    int wpb;   // waves per block
    int rpb;   // rows per block
    int bpr;   // blocks per row
};

inline RowMap row_map(int span)
{
    RowMap r;
    r.wpr = ceil_div(span, 64);
    if (r.wpr > 8) {
        r.wpb = 4;
        r.rpb = 1;
        r.bpr = ceil_div(r.wpr, 4);
    } else if (r.wpr >= 4) {
        r.wpb = r.wpr;
        r.rpb = 1;
        r.bpr = 1;
    } else {
        r.rpb = 4 / r.wpr;
        r.wpb = r.rpb * r.wpr;
        r.bpr = 1;
    }
    return r;
}

// Block-capped variant for launches whose block size is fixed at
// 64 * maxw threads (the batched serving kernel): rows longer than maxw
// waves are split into segments of maxw waves, shorter rows are packed
// maxw / wpr to a block (a block may then hold idle waves).
inline RowMap row_map_capped(int span, int maxw)
{
    RowMap r;
    r.wpr = ceil_div(span, 64);
    if (r.wpr > maxw) {
        r.wpb = maxw;
        r.rpb = 1;
        r.bpr = ceil_div(r.wpr, maxw);
    } else {
        r.rpb = maxw / r.wpr;
        r.wpb = r.rpb * r.wpr;
        r.bpr = 1;
    }
    return r;
}

// Collision and stores of one cell whose pulled populations are x.  `cell`
// keys the stochastic-rounding stream: the cell's index in the reference's
// (0:l+1, 0:m+1, 0:n+1) layout, so the stream does not depend on the
// wave mapping or the row padding.  A lane with live = false (a padding or
// halo position that only completes a 128-byte line) stores zeros through
// the same store instructions, so a wave that covers both kinds of lanes
// issues its 19 stores once instead of branching into a second set.
// FILL (compact rows, inflow boundary; kFillsInflowFaces): a halo lane with
// face = 1 (x = 0) stores the inflow face's values in directions 1..5, with
// face = 2 (x = l+1, its pull ran on cell l) the outflow face's values from
// cell l's new populations as cell l stores them (its `cell` key is cell
// l's, so stochastic rounding draws the same words) in directions 10..14,
// zeros elsewhere -- what the outflow kernel would write before the next
// collide.
template <typename S, int FMODE, int SR, bool CMP, bool FILL = false>
__device__ __forceinline__ void collide_store(const FusedArgs<S>& p, typename Storage<S>::compute (&x)[19], int i,
                                              long row, uint32_t seed, uint32_t cell, bool live = true, int face = 0)
{
    typedef typename Storage<S>::compute C;
    constexpr long ES = (long)sizeof(S);
    const uint32_t vi = (uint32_t)i * (uint32_t)ES;
    // FMODE bits 0-1: body force (0 uniform, 1/2/3 fgrad*rho along x/y/z);
    // bit 2: Smagorinsky LES relaxation
    constexpr bool LES = (FMODE & 4) != 0;
    C fx, fy, fz;
    force_terms<C, (FMODE & 3)>(density_ref(x, p.cst.cte1), p.fgrad, p.fx, p.fy, p.fz, fx, fy, fz);
    if constexpr (std::is_same<C, double>::value && !LES) {
        // FP64 arithmetic: reference-order, each value stored as soon as it
        // is formed (77 vector registers instead of 86: 6 waves per SIMD
        // without spilling).  The FP32 form keeps the array collide: with
        // its stores spread through the arithmetic, rows that end in a
        // partly interior line (n = 134 ... 182) ran 18-24 % slower.
        static_assert(!FILL, "the streaming collide does not fill the inflow faces");
        bgk_collide_stream(x, p.omega, fx, fy, fz, p.cst, [&](auto dc, C v) {
            constexpr int d = decltype(dc)::value;
            S* dst;
            if constexpr (d < 18)
                dst = p.b[d];
            else
                dst = p.a[18];
            const C w = live ? v : C(0);
            store_nt(dst + (row + i), Rounder<S, 0>::cvt(w, 0u));
            if constexpr (d == 18) flag_nonfinite(p, w);
        });
        return;
    }
    C o[19];
    if constexpr (Storage<S>::lean)
        bgk_collide_lean<C, LES>(x, p.omega, fx, fy, fz, p.cst, o, p.cte_s);
    else
        bgk_collide_ref<C, LES>(x, p.omega, fx, fy, fz, p.cst, o, p.cte_s);

    // stochastic rounding: one hash per cell and step, then a Weyl sequence
    // over the cell's stores (deterministic: independent of scheduling)
    uint32_t r = 0;
    if (SR) r = hash32(seed ^ cell);
    [[maybe_unused]] S hv[5];
    if constexpr (FILL) {
        if (face == 1) {
            inflow_face<S>(p.uin, hv);
        } else if (face == 2) {
            C f[19];
#pragma unroll
            for (int d = 0; d < 19; ++d) f[d] = Storage<S>::widen(Rounder<S, SR>::cvt(o[d], SR ? sr_word(r, d) : 0u));
            outflow_face<S>(f, hv);
        }
    }
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        S* dst;
        if constexpr (sizeof(S) >= 4)
            dst = (d < 18 ? p.b[d] : p.a[18]) + (row + i);
        else
            dst = reinterpret_cast<S*>(reinterpret_cast<char*>(pop_b<CMP>(p, d)) + row * ES + (uint64_t)vi);
        S v = Rounder<S, SR>::cvt(live ? o[d] : C(0), SR ? sr_word(r, d) : 0u);
        if constexpr (FILL) {
            if (d < 5 && face == 1) v = hv[d];
            if (d >= 9 && d < 14 && face == 2) v = hv[d - 9];
        }
        store_nt(dst, v);
    }
    flag_nonfinite(p, live ? o[18] : C(0));
}

// Index of cell (i, j, k) in the reference's (0:l+1, 0:m+1, 0:n+1) layout.
template <typename S>
__device__ __forceinline__ uint32_t ref_cell(const FusedArgs<S>& p, int i, int j, int k)
{
    return (uint32_t)i + (uint32_t)(p.l + 2) * ((uint32_t)j + (uint32_t)(p.m + 2) * (uint32_t)k);
}

// The work of one block (bx, by, bz) of the launch geometry `rm` describes:
// pull, collide, store.  `seed` is the step's stochastic-rounding seed.
// PADZ: rows may end in lanes l < i <= lend that store zeros (below); with
// PADZ false the launcher guarantees lend <= l, and the 16-bit forms, which
// are sensitive to instruction issue, skip the per-store selects (their
// 256^3 collide took 228 instead of 214 us with them).
template <typename S, int BX, int BY, int BZ, bool OBST, int FMODE, int SR, bool PADZ = true>
__device__ __forceinline__ void fused_rows(const FusedArgs<S>& p, const RowMap& rm, int bx, int by, int bz,
                                            uint32_t seed)
{
    typedef typename Storage<S>::compute C;
    constexpr bool CMP = kCompact<S, FMODE>;
    // the 16-bit forms declare the wave index scalar, so that the row offset
    // and the boundary decisions stay in scalar registers
    const int wave = sizeof(S) == 2 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wave >= rm.wpb) return;
    int seg, j;
    if (rm.rpb == 1) {
        seg = bx * rm.wpb + wave;
        j = 1 + by;
    } else {
        seg = wave % rm.wpr;
        j = 1 + by * rm.rpb + wave / rm.wpr;
    }
    const int k = 1 + bz;
    if (j > p.m) return;
    int shift = 0;
    if (p.align) {
        const long first = (long)(reinterpret_cast<uintptr_t>(p.b[0]) / sizeof(S)) + 1 + (long)p.nx * j +
                           p.nxy * (long)k;
        shift = (int)(first & (long)(p.align - 1));
    }
    const int i = 1 - shift + seg * 64 + lane;
    if (i < 1) return;
    const long row = (long)p.nx * j + p.nxy * (long)k;  // scalar
    // Whole-line stores: when the row's last 128-byte line is only partly
    // interior, the lanes past i = l fill it with zeros (row padding and the
    // next row's i = 0 halo cell, which a folded x boundary never reads), so
    // HBM sees full-line writes instead of partial ones.  Only with a layout
    // whose rows start on a line (lend set by the engine).  Those lanes run
    // the pull on cell l (in bounds, lines the wave reads anyway) and store
    // through the wave's own store instructions.
    if (i > p.l && (!PADZ || i > p.lend)) return;
    const bool live = !PADZ || i <= p.l;
    const int il = live ? i : p.l;

    C x[19];
    const bool edge = (BY != AX_HALO && (j == 1 || j == p.m)) || (BZ != AX_HALO && (k == 1 || k == p.n));
    if (edge)
        pull_all<S, BX, BY, BZ, OBST, true, CMP>(p, il, j, k, row, x);
    else
        pull_all<S, BX, BY, BZ, OBST, false, CMP>(p, il, j, k, row, x);
    collide_store<S, FMODE, SR, CMP>(p, x, i, row, seed, ref_cell(p, il, j, k), live);
}

// The compact layout of the handle API: rows of l + 2 cells without padding
// (nx = l + 2), each population plane padded to whole 128-byte lines and
// placed so that cell (0, 1, k) starts a line.  The waves of plane k cover
// its rows j = 1..m as one range of positions r = i + nx (j - 1), 64
// consecutive cells per wave, line-aligned, up to the line end after
// (l+1, m) (p.lend positions): lanes on interior cells pull, collide and
// store; lanes on the x halo cells and on the cells of row m+1 store zeros,
// so every line is written whole.  A folded boundary never reads a halo
// cell.  The inflow boundary reads both x halos: the forms of
// kFillsInflowFaces store the faces' values from those lanes instead of
// zeros (the lane on x = l+1 has run the collide of cell l), the others
// leave them to the step's outflow kernel.  The padded layout spends whole
// lines on a row's last few cells (a 129-cell FP32 row spans five lines for
// 4.03 lines of data); this one wastes 2 cells per row, with rows that start
// anywhere in a line (y-shifted pulls then span three lines per wave, the
// third shared with the neighbouring wave through L2).  bz = k - 1, bx =
// position block of 256 (blocks of 512 positions, which keep more of the
// x-periodic wraps within one block, ran 420-438 us and of 1024 433 us,
// against 411 us, at 256^3 FP32).
template <typename S, int BX, int BY, int BZ, bool OBST, int FMODE, int SR>
__device__ __forceinline__ void compact_cells(const FusedArgs<S>& p, int bx, int bz, uint32_t seed)
{
    typedef typename Storage<S>::compute C;
    constexpr bool CMP = kCompact<S, FMODE>;
    const int r = bx * 256 + (int)threadIdx.x;
    if (r >= p.lend) return;
    const int k = 1 + bz;
    const int j = 1 + r / p.nx;
    const int i = r - (j - 1) * p.nx;
    const long row = (long)p.nx * j + p.nxy * (long)k;
    // halo positions (i = 0, i = l + 1, row m + 1) store zeros through the
    // wave's own store instructions after a pull on the nearest interior
    // cell (collide_store, live = false)
    const bool live = j <= p.m && i >= 1 && i <= p.l;
    const int il = i < 1 ? 1 : (i > p.l ? p.l : i);
    const int jl = j > p.m ? p.m : j;
    const long rowl = (long)p.nx * jl + p.nxy * (long)k;
    // Every wave runs the y/z boundary checks: a wave vote that skipped them
    // in waves with no lane on a boundary row or plane (98 % at 256^3) ran
    // slower (FP32 425-427 against 417-422 us, MIXED2 443 against 418-422).
    C x[19];
    pull_all<S, BX, BY, BZ, OBST, true, CMP>(p, il, jl, k, rowl, x);
    constexpr bool FILL = BX == AX_HALO && kFillsInflowFaces<S, FMODE>;
    const int face = (FILL && j <= p.m) ? (i == 0 ? 1 : (i == p.l + 1 ? 2 : 0)) : 0;
    collide_store<S, FMODE, SR, CMP, FILL>(p, x, i, row, seed, ref_cell(p, il, jl, k), live, face);
}

// Waves per SIMD the compiler is asked to fit (1: no request).  FP64
// without LES (streaming collide, 77 vector registers) is asked for 6; the
// other forms are left to the compiler: FP32 reaches 8 by itself, the LES
// forms need every equilibrium before relaxing, and the 16-bit forms ran
// slower when pushed from 7 to 8 waves.
template <typename S, int FMODE>
struct FusedWaves {
    static constexpr bool LES = (FMODE & 4) != 0;
    static constexpr int value = (!LES && std::is_same<typename Storage<S>::compute, double>::value) ? 6 : 1;
};

template <typename S, int BX, int BY, int BZ, bool OBST, int FMODE, int SR, bool PADZ = true>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(FusedWaves<S, FMODE>::value))) void collide_fused(
    FusedArgs<S> p, RowMap rm)
{
    fused_rows<S, BX, BY, BZ, OBST, FMODE, SR, PADZ>(p, rm, (int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z,
                                                      p.seed);
}

// Dynamic LDS the launchers reserve per collide_compact block.  The kernel
// uses none; 32 KiB per 256-thread block holds the FP32 form to 5 waves per
// SIMD (a CU has 160 KiB) instead of 8: fewer streams in flight over rows
// that do not start on a line ran 8 % faster at 256^3 (463 against 502 us)
// and 4-5 % at n = 129 ... 192; 6 waves (26 KiB) took 426-428 against
// 420-423 us, 4 waves (40 KiB) the same as 5.  MIXED2 (FP32 storage, FP64
// collide, 6 waves by register use) likewise: 425 against 436 us at 256^3.
// FP64 storage (883 against 871-878 us, 40 KiB no better) and the 16-bit
// forms (259 against 245 us) ran slower so capped and keep their natural
// occupancy.
template <typename S>
constexpr unsigned compact_lds_cap()
{
    return (std::is_same<S, float>::value || std::is_same<S, mix32_t>::value) ? (32u << 10) : 0u;
}

template <typename S, int BX, int BY, int BZ, bool OBST, int FMODE, int SR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FusedWaves<S, FMODE>::value))) void collide_compact(
    FusedArgs<S> p)
{
    compact_cells<S, BX, BY, BZ, OBST, FMODE, SR>(p, (int)blockIdx.x, (int)blockIdx.y, p.seed);
}

// Reference-layout form (the caller's arrays, every boundary already in the
// halo; FP32/FP64).  The waves of plane k cover its interior rows j = 1..m as
// ONE range of cells, from (1, 1, k) to (l, m, k): 64 consecutive cells per
// wave, the first wave starting on the 128-byte line of the output that
// holds (1, 1, k).  A wave may span the end of one row and the start of the
// next; its lanes on halo or padding cells (i = 0, i > l) neither load nor
// store.  Row-by-row waves need ceil((l + shift) / 64) waves per row when
// rows do not start on a line (nx = l + 2: five waves, the fifth holding
// 0-31 cells, for l = 256); this form needs (nx*(m-1) + l + shift) / 64 per
// plane.  blockIdx.y = k - 1; the plane's cells must be indexable in 32 bits
// (nx * ny < 2^31, checked by the launcher).  NT0 as for pull_all: set only
// when nx and nx*ny are multiples of a line.
// WL (whole lines): the lanes on the halo and padding cells between two
// interior rows (i = 0, i > l inside the range) store zeros after a pull on
// the nearest interior cell of their row, so that every line of the range is
// written whole.  Only where the caller lets the collide overwrite those
// cells of its output arrays: no routine reads them before the next boundary
// call rewrites every halo value a pull uses.  A line the kernel writes only
// in part costs the L2 a read of that line from HBM first: 484 against
// 441 us per FP32 launch at 256^3.
template <typename S, int FMODE, bool NT0, bool WL = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(FusedWaves<S, FMODE>::value))) void collide_flat(
    FusedArgs<S> p)
{
    static_assert(sizeof(S) >= 4, "the reference layout holds FP32 or FP64 populations");
    typedef typename Storage<S>::compute C;
    constexpr bool LES = (FMODE & 4) != 0;
    const int k = 1 + (int)blockIdx.y;
    const long plane = p.nxy * (long)k;
    const int lo = p.nx + 1;             // (1, 1) within the plane
    const int hi = p.nx * p.m + p.l;     // (l, m)
    const long b0 = (long)(reinterpret_cast<uintptr_t>(p.b[0]) / sizeof(S));
    const int start = lo - (int)((b0 + plane + lo) & (long)(p.align - 1));
    const int r = start + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r < lo || r > hi) return;
    const int j = (int)((unsigned)r / (unsigned)p.nx);
    const int i = r - j * p.nx;
    const bool live = i >= 1 && i <= p.l;
    if (!WL && !live) return;
    const int il = live ? i : (i < 1 ? 1 : p.l);
    const long row = plane + (long)(r - i);

    C x[19];
    pull_all<S, AX_HALO, AX_HALO, AX_HALO, false, false, false, NT0>(p, il, j, k, row, x);
    C fx, fy, fz;
    force_terms<C, (FMODE & 3)>(density_ref(x, p.cst.cte1), p.fgrad, p.fx, p.fy, p.fz, fx, fy, fz);
    if constexpr (sizeof(S) == 8 && !LES) {
        bgk_collide_stream(x, p.omega, fx, fy, fz, p.cst, [&](auto dc, C v) {
            constexpr int d = decltype(dc)::value;
            store_nt((d < 18 ? p.b[d] : p.a[18]) + (row + i), live ? v : C(0));
        });
    } else {
        C o[19];
        bgk_collide_ref<C, LES>(x, p.omega, fx, fy, fz, p.cst, o, p.cte_s);
#pragma unroll
        for (int d = 0; d < 19; ++d) store_nt((d < 18 ? p.b[d] : p.a[18]) + (row + i), live ? o[d] : C(0));
    }
}

// Launch geometry shared by both ABIs; `span` is the row length the waves
// must cover (l, or l plus the largest alignment shift).
inline dim3 fused_grid(const RowMap& rm, int m, int n)
{
    return dim3((unsigned)rm.bpr, (unsigned)ceil_div(m, rm.rpb), (unsigned)n);
}
inline dim3 fused_block(const RowMap& rm) { return dim3((unsigned)(64 * rm.wpb)); }

}  // namespace bgk
