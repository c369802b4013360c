// D3Q19 lattice and the BGK collision, as device functions.
//
// Directions use the reference numbering 1..19 (SRC/bgk3d.F90:29-48,
// SRC/alloca.F90:29-148); index d here is direction d+1.  Population d at
// interior cell x is pulled from x - c_d (SRC/col_MC.F90:275-293).
//
// bgk_collide_ref() evaluates the shifted-storage BGK step of
// SRC/col_MC.F90:295-519 term by term in the same order.  Built with
// -ffp-contract=off, every product and sum is rounded exactly where the
// reference's x86-64 build rounds it, so results are bit-identical to the
// reference's CPU path in both float and double.  The kernel is bound by HBM
// bandwidth (about 2 flop per byte), so forgoing FMA costs no time.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace bgk {

// c_x, c_y, c_z of directions 1..19
__host__ __device__ constexpr int cx_of(int d)
{
    constexpr int v[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
    return v[d];
}
__host__ __device__ constexpr int cy_of(int d)
{
    constexpr int v[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};
    return v[d];
}
__host__ __device__ constexpr int cz_of(int d)
{
    constexpr int v[19] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0};
    return v[d];
}
// opposite direction (1<->12, 2<->13, 3<->10, 4<->11, 5<->14, 6<->15,
// 7<->16, 8<->17, 9<->18; SRC/bcond.obs.F90:104-121)
__host__ __device__ constexpr int opp_of(int d)
{
    constexpr int v[19] = {11, 12, 9, 10, 13, 14, 15, 16, 17, 2, 3, 0, 1, 4, 5, 6, 7, 8, 18};
    return v[d];
}

// Weights as the reference computes them: double-precision parameters
// rounded to the working kind (SRC/storage.F90:58-66, 142-151).
template <typename R>
struct W {
    static constexpr R p0 = static_cast<R>(1.0 / 3.0);
    static constexpr R p1 = static_cast<R>(1.0 / 18.0);
    static constexpr R p2 = static_cast<R>(1.0 / (3.0 * 3.0 * (1.0 + 3.0)));
};

// The scalar constants the reference passes to its native collision
// (SRC/col_MC.F90:217-229): cte1 = 1, cte0 = 0 for shifted storage (the
// default), cte1 = 0, cte0 = 1 for NOSHIFT; weights; rf = 3, qf = 1.5,
// tre = 3.
template <typename R>
struct CollideConsts {
    R cte0, cte1, p0, p1, p2, rf, qf, tre;
    __host__ __device__ static CollideConsts shifted()
    {
        return {R(0), R(1), W<R>::p0, W<R>::p1, W<R>::p2, R(3), R(1.5), R(3)};
    }
};

__device__ __forceinline__ float dsqrt(float v) { return __builtin_sqrtf(v); }
__device__ __forceinline__ double dsqrt(double v) { return __builtin_sqrt(v); }

// Smagorinsky relaxation rate of one cell (SRC/col_MC.F90:381-466, LES
// builds): the non-equilibrium momentum flux P from n = x - e, then
// Ts = 1/(2 omega1) + sqrt(18 cteS^2 |P| + (1/omega1)^2)/2 and omega = 1/Ts,
// in the reference's order (the 2.0 is a default-real literal, exact).
template <typename R>
__device__ __forceinline__ R les_omega(const R (&x)[19], const R (&e)[19], R omega1, R cte_s)
{
    R nn[18];
#pragma unroll
    for (int d = 0; d < 18; ++d) nn[d] = x[d] - e[d];
    const R pxx = nn[0] + nn[1] + nn[2] + nn[3] + nn[4] + nn[9] + nn[10] + nn[11] + nn[12] + nn[13];
    const R pyy = nn[0] + nn[2] + nn[6] + nn[7] + nn[8] + nn[9] + nn[11] + nn[15] + nn[16] + nn[17];
    const R pzz = nn[1] + nn[3] + nn[5] + nn[6] + nn[8] + nn[10] + nn[12] + nn[14] + nn[15] + nn[17];
    const R pxz = -nn[1] + nn[3] + nn[10] - nn[12];
    const R pxy = -nn[0] + nn[2] + nn[9] - nn[11];
    const R pyz = +nn[6] - nn[8] + nn[15] - nn[17];
    const R two = R(2.0f);
    const R ptot = dsqrt(pxx * pxx + pyy * pyy + pzz * pzz + (two * pxy * pxy) + (two * pxz * pxz) +
                         (two * pyz * pyz));
    const R inv = R(1) / omega1;
    const R ts = R(1) / (R(2) * omega1) + dsqrt(R(18) * (cte_s * cte_s) * ptot + inv * inv) / R(2);
    return R(1) / ts;
}

// Shifted-storage BGK collision of one cell.
//   x[0..18]  pulled (post-streaming) populations f_i - w_i
//   omega     relaxation rate 2/(6 nu + 1) (with LES: the molecular rate
//             omega1, replaced per cell by les_omega)
//   fx,fy,fz  the force terms added to the populations with c = +-1
// out[0..18] = x - omega (x - e) + F, with the reference's term order.
template <typename R, bool LES = false>
__device__ __forceinline__ void bgk_collide_ref(const R (&x)[19], R omega, R fx, R fy, R fz,
                                                const CollideConsts<R>& c, R (&o)[19], R cte_s = R(0))
{
    const R uno = R(1), cte1 = c.cte1, cte0 = c.cte0;
    const R rf = c.rf, qf = c.qf, tre = c.tre;
    const R p0 = c.p0, p1 = c.p1, p2 = c.p2;

    const R rho = x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9] + x[10] + x[11] +
                  x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18] + cte1;
    const R rhoinv = uno / rho;
    R vx = (x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
    R vy = (x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
    R vz = (x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;

    const R vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
    const R vsq = vx2 + vy2 + vz2;
    R vxpy = vx + vy;
    const R qxpy = cte0 + qf * (tre * vxpy * vxpy - vsq);
    R vxmy = vx - vy;
    const R qxmy = cte0 + qf * (tre * vxmy * vxmy - vsq);
    R vxpz = vx + vz;
    const R qxpz = cte0 + qf * (tre * vxpz * vxpz - vsq);
    R vxmz = vx - vz;
    const R qxmz = cte0 + qf * (tre * vxmz * vxmz - vsq);
    R vypz = vy + vz;
    const R qypz = cte0 + qf * (tre * vypz * vypz - vsq);
    R vymz = vy - vz;
    const R qymz = cte0 + qf * (tre * vymz * vymz - vsq);
    const R qx = cte0 + qf * (tre * vx2 - vsq);
    const R qy = cte0 + qf * (tre * vy2 - vsq);
    const R qz = cte0 + qf * (tre * vz2 - vsq);
    const R q0 = cte0 + qf * (-vsq);

    vx = rf * vx;
    vy = rf * vy;
    vz = rf * vz;
    vxpy = rf * vxpy;
    vxmy = rf * vxmy;
    vxpz = rf * vxpz;
    vxmz = rf * vxmz;
    vypz = rf * vypz;
    vymz = rf * vymz;

    const R rp0 = rho * p0, rp1 = rho * p1, rp2 = rho * p2;
    const R s2 = cte1 * (rp2 - p2), s1 = cte1 * (rp1 - p1), s0 = cte1 * (rp0 - p0);

    R e[19];
    e[0] = rp2 * (vxmy + qxmy) + s2;
    e[1] = rp2 * (vxmz + qxmz) + s2;
    e[2] = rp2 * (vxpy + qxpy) + s2;
    e[3] = rp2 * (vxpz + qxpz) + s2;
    e[4] = rp1 * (vx + qx) + s1;
    e[5] = rp1 * (vz + qz) + s1;
    e[6] = rp2 * (vypz + qypz) + s2;
    e[7] = rp1 * (vy + qy) + s1;
    e[8] = rp2 * (vymz + qymz) + s2;
    e[9] = rp2 * (-vxpy + qxpy) + s2;
    e[10] = rp2 * (-vxpz + qxpz) + s2;
    e[11] = rp2 * (-vxmy + qxmy) + s2;
    e[12] = rp2 * (-vxmz + qxmz) + s2;
    e[13] = rp1 * (-vx + qx) + s1;
    e[14] = rp1 * (-vz + qz) + s1;
    e[15] = rp2 * (-vypz + qypz) + s2;
    e[16] = rp1 * (-vy + qy) + s1;
    e[17] = rp2 * (-vymz + qymz) + s2;
    e[18] = rp0 * (q0) + s0;

    if constexpr (LES) omega = les_omega(x, e, omega, cte_s);

    o[0] = x[0] - omega * (x[0] - e[0]) + fx - fy;
    o[1] = x[1] - omega * (x[1] - e[1]) + fx - fz;
    o[2] = x[2] - omega * (x[2] - e[2]) + fx + fy;
    o[3] = x[3] - omega * (x[3] - e[3]) + fx + fz;
    o[4] = x[4] - omega * (x[4] - e[4]) + fx;
    o[5] = x[5] - omega * (x[5] - e[5]) + fz;
    o[6] = x[6] - omega * (x[6] - e[6]) + fy + fz;
    o[7] = x[7] - omega * (x[7] - e[7]) + fy;
    o[8] = x[8] - omega * (x[8] - e[8]) + fy - fz;
    o[9] = x[9] - omega * (x[9] - e[9]) - fx - fy;
    o[10] = x[10] - omega * (x[10] - e[10]) - fx - fz;
    o[11] = x[11] - omega * (x[11] - e[11]) - fx + fy;
    o[12] = x[12] - omega * (x[12] - e[12]) - fx + fz;
    o[13] = x[13] - omega * (x[13] - e[13]) - fx;
    o[14] = x[14] - omega * (x[14] - e[14]) - fz;
    o[15] = x[15] - omega * (x[15] - e[15]) - fy - fz;
    o[16] = x[16] - omega * (x[16] - e[16]) - fy;
    o[17] = x[17] - omega * (x[17] - e[17]) - fy + fz;
    o[18] = x[18] - omega * (x[18] - e[18]);
}

// The same BGK step as bgk_collide_ref, handing each post-collision value to
// sink(std::integral_constant<int, d>{}, value) as soon as it is formed.
// Every value is computed with exactly the expression of bgk_collide_ref
// (same operands, same order, same roundings), so the results are identical;
// only the order in which the 19 values are produced differs: pairs of
// opposite directions share their quadratic term, and a population's
// register is free once its value has been emitted.  Without LES (whose
// relaxation rate needs every equilibrium first) this keeps far fewer values
// live than the array form.
template <int D>
using DirC = std::integral_constant<int, D>;

template <typename R, typename Sink>
__device__ __forceinline__ void bgk_collide_stream(const R (&x)[19], R omega, R fx, R fy, R fz,
                                                   const CollideConsts<R>& c, Sink&& sink)
{
    const R uno = R(1), cte1 = c.cte1, cte0 = c.cte0;
    const R rf = c.rf, qf = c.qf, tre = c.tre;
    const R p0 = c.p0, p1 = c.p1, p2 = c.p2;

    const R rho = x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9] + x[10] + x[11] +
                  x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18] + cte1;
    const R rhoinv = uno / rho;
    const R vx = (x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
    const R vy = (x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
    const R vz = (x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;
    const R vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
    const R vsq = vx2 + vy2 + vz2;
    const R rp0 = rho * p0, rp1 = rho * p1, rp2 = rho * p2;
    const R s2 = cte1 * (rp2 - p2), s1 = cte1 * (rp1 - p1), s0 = cte1 * (rp0 - p0);

    // diagonal pair along v = (c_a +- c_b).u: q = cte0 + qf (tre v v - vsq), scaled rf v
    auto qd = [&](R v) { return cte0 + qf * (tre * v * v - vsq); };
    // axis pair: q = cte0 + qf (tre v^2 - vsq) with v^2 the squared component
    auto qa = [&](R v2) { return cte0 + qf * (tre * v2 - vsq); };
    auto rel = [&](R xd, R e) { return xd - omega * (xd - e); };

    {  // x - y: directions 1 and 12
        const R v = vx - vy, q = qd(v), r = rf * v;
        sink(DirC<0>{}, rel(x[0], rp2 * (r + q) + s2) + fx - fy);
        sink(DirC<11>{}, rel(x[11], rp2 * (-r + q) + s2) - fx + fy);
    }
    {  // x - z: directions 2 and 13
        const R v = vx - vz, q = qd(v), r = rf * v;
        sink(DirC<1>{}, rel(x[1], rp2 * (r + q) + s2) + fx - fz);
        sink(DirC<12>{}, rel(x[12], rp2 * (-r + q) + s2) - fx + fz);
    }
    {  // x + y: directions 3 and 10
        const R v = vx + vy, q = qd(v), r = rf * v;
        sink(DirC<2>{}, rel(x[2], rp2 * (r + q) + s2) + fx + fy);
        sink(DirC<9>{}, rel(x[9], rp2 * (-r + q) + s2) - fx - fy);
    }
    {  // x + z: directions 4 and 11
        const R v = vx + vz, q = qd(v), r = rf * v;
        sink(DirC<3>{}, rel(x[3], rp2 * (r + q) + s2) + fx + fz);
        sink(DirC<10>{}, rel(x[10], rp2 * (-r + q) + s2) - fx - fz);
    }
    {  // x: directions 5 and 14
        const R q = qa(vx2), r = rf * vx;
        sink(DirC<4>{}, rel(x[4], rp1 * (r + q) + s1) + fx);
        sink(DirC<13>{}, rel(x[13], rp1 * (-r + q) + s1) - fx);
    }
    {  // z: directions 6 and 15
        const R q = qa(vz2), r = rf * vz;
        sink(DirC<5>{}, rel(x[5], rp1 * (r + q) + s1) + fz);
        sink(DirC<14>{}, rel(x[14], rp1 * (-r + q) + s1) - fz);
    }
    {  // y + z: directions 7 and 16
        const R v = vy + vz, q = qd(v), r = rf * v;
        sink(DirC<6>{}, rel(x[6], rp2 * (r + q) + s2) + fy + fz);
        sink(DirC<15>{}, rel(x[15], rp2 * (-r + q) + s2) - fy - fz);
    }
    {  // y: directions 8 and 17
        const R q = qa(vy2), r = rf * vy;
        sink(DirC<7>{}, rel(x[7], rp1 * (r + q) + s1) + fy);
        sink(DirC<16>{}, rel(x[16], rp1 * (-r + q) + s1) - fy);
    }
    {  // y - z: directions 9 and 18
        const R v = vy - vz, q = qd(v), r = rf * v;
        sink(DirC<8>{}, rel(x[8], rp2 * (r + q) + s2) + fy - fz);
        sink(DirC<17>{}, rel(x[17], rp2 * (-r + q) + s2) - fy + fz);
    }
    // rest: direction 19
    sink(DirC<18>{}, rel(x[18], rp0 * (cte0 + qf * (-vsq)) + s0));
}

// The same collision with FMA contraction allowed, for the 16-bit storage
// forms: their results are not bit-comparable with the reference anyway
// (rounding to 16 bits dominates), and fused multiply-adds cut the VALU work
// that becomes visible when each cell moves only 76 bytes.
template <typename R, bool LES = false>
__device__ __forceinline__ void bgk_collide_fast(const R (&x)[19], R omega, R fx, R fy, R fz,
                                                 const CollideConsts<R>& c, R (&o)[19], R cte_s = R(0))
{
#pragma clang fp contract(fast)
    bgk_collide_ref<R, LES>(x, omega, fx, fy, fz, c, o, cte_s);
}

// The same BGK step in a lean form for the 16-bit storage forms, whose
// kernels are bound by VALU issue rather than HBM.  In shifted storage the
// density deviation drho = rho - 1 is the plain sum of the stored values, so
// the equilibrium is formed without the cancellation of w rho - w:
//   e_d = w drho + w rho (4.5 (c.u)^2 - 1.5 u^2)  +/-  3 w rho (c.u),
// the +/- pair shared by opposite directions; the relaxation is one fused
// multiply-add per direction and 1/rho the hardware reciprocal.  Same
// mathematics as bgk_collide_ref (shifted storage, cte1 = 1), about half the
// operations, and FP32 errors far below the 16-bit store.
template <typename R, bool LES = false>
__device__ __forceinline__ void bgk_collide_lean(const R (&x)[19], R omega, R fx, R fy, R fz,
                                                 const CollideConsts<R>& c, R (&o)[19], R cte_s = R(0))
{
#pragma clang fp contract(fast)
    R drho = x[0];
#pragma unroll
    for (int d = 1; d < 19; ++d) drho += x[d];
    const R rho = R(1) + drho;
    const R jx = (x[0] + x[1] + x[2] + x[3] + x[4]) - (x[9] + x[10] + x[11] + x[12] + x[13]);
    const R jy = (x[2] + x[6] + x[7] + x[8] + x[11]) - (x[0] + x[9] + x[15] + x[16] + x[17]);
    const R jz = (x[3] + x[5] + x[6] + x[12] + x[17]) - (x[1] + x[8] + x[10] + x[14] + x[15]);
    const R rinv = __builtin_amdgcn_rcpf((float)rho);
    const R ux = jx * rinv, uy = jy * rinv, uz = jz * rinv;
    const R u15 = R(1.5) * (ux * ux + uy * uy + uz * uz);
    const R wr[3] = {c.p0 * rho, c.p1 * rho, c.p2 * rho};
    const R wd[3] = {c.p0 * drho, c.p1 * drho, c.p2 * drho};
    R e[19];
#pragma unroll
    for (int d = 0; d < 9; ++d) {
        const int cx = cx_of(d), cy = cy_of(d), cz = cz_of(d);
        const int k = cx * cx + cy * cy + cz * cz;  // 1 axis, 2 diagonal
        const R cu = (cx ? (cx > 0 ? ux : -ux) : R(0)) + (cy ? (cy > 0 ? uy : -uy) : R(0)) +
                     (cz ? (cz > 0 ? uz : -uz) : R(0));
        const R sym = wd[k] + wr[k] * (R(4.5) * cu * cu - u15);
        const R anti = R(3) * wr[k] * cu;
        e[d] = sym + anti;
        e[opp_of(d)] = sym - anti;
    }
    e[18] = wd[0] - wr[0] * u15;
    if constexpr (LES) omega = les_omega(x, e, omega, cte_s);
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        R f = R(0);
        if (cx_of(d)) f = cx_of(d) > 0 ? f + fx : f - fx;
        if (cy_of(d)) f = cy_of(d) > 0 ? f + fy : f - fy;
        if (cz_of(d)) f = cz_of(d) > 0 ? f + fz : f - fz;
        o[d] = x[d] + omega * (e[d] - x[d]) + f;
    }
}

// Force terms of one cell: mode 0 uniform (fx,fy,fz as given), 1/2/3 the
// CHANNEL body force fgrad*rho along x/y/z (SRC/col_MC.F90:472-488).
template <typename R, int FMODE>
__device__ __forceinline__ void force_terms(R rho, R fgrad, R ux, R uy, R uz, R& fx, R& fy, R& fz)
{
    if constexpr (FMODE == 0) {
        fx = ux;
        fy = uy;
        fz = uz;
    } else if constexpr (FMODE == 1) {
        fx = fgrad * rho;
        fy = R(0);
        fz = R(0);
    } else if constexpr (FMODE == 2) {
        fx = R(0);
        fy = fgrad * rho;
        fz = R(0);
    } else {
        fx = R(0);
        fy = R(0);
        fz = fgrad * rho;
    }
}

// rho exactly as the collision computes it (needed by the body force).
template <typename R>
__device__ __forceinline__ R density_ref(const R (&x)[19], R cte1)
{
    return x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9] + x[10] + x[11] + x[12] +
           x[13] + x[14] + x[15] + x[16] + x[17] + x[18] + cte1;
}

}  // namespace bgk
