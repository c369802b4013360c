// Device-side diagnostics: diagno, varm, draglift and prof_k.
//
// The reference runs these on the host after copying the whole lattice back
// every `icheck` steps (SRC/diagnostic.F90:87-113, SRC/diagno.F90:73-132).
// Here they run where the data lives.  Per-cell arithmetic keeps the
// reference's kind and expression order; sums accumulate in FP64.  Every
// reduction is two-pass and deterministic: a fixed grid writes one partial
// per block (wave64 shuffles, then LDS across the block's waves), and a
// single block folds the partials in index order, so repeated calls give the
// same bits.
#pragma once

#include "common.hpp"
#include "d3q19.hpp"
#include "storage.hpp"
#include "v1_kernels.hpp"

namespace bgk {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;  // partials per reduction pass

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum NV doubles over the block; thread 0 receives the totals.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double part[kRedThreads / 64][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) part[wave][q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = part[0][q];
            for (int w = 1; w < kRedThreads / 64; ++w) s += part[w][q];
            v[q] = s;
        }
    }
    __syncthreads();
}

// Second pass: one block folds nblk partials (layout [blk][NV]) in order.
template <int NV>
__global__ __launch_bounds__(kRedThreads) void fold_partials(const double* __restrict__ part, int nblk,
                                                             double* __restrict__ out)
{
    double v[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kRedThreads) {
#pragma unroll
        for (int q = 0; q < NV; ++q) v[q] += part[(long)b * NV + q];
    }
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) out[q] = v[q];
    }
}

// diagno (SRC/diagno.F90:89-131): per interior cell rho, u, v, w and
// 0.5|u|^2 in the storage kind; partial sums {u, v, w, rho, ke} in FP64.
// Rows (j,k) are distributed over blocks; lanes walk i.
template <typename S>
__global__ __launch_bounds__(kRedThreads) void diagno_partial(Lattice19<S> a, Geometry g,
                                                              typename Storage<S>::compute cte1,
                                                              double* __restrict__ part)
{
    typedef typename Storage<S>::compute R;
    double v[5] = {0, 0, 0, 0, 0};
    const long rows = (long)g.m * g.n;
    for (long r = blockIdx.x; r < rows; r += gridDim.x) {
        const int j = 1 + (int)(r % g.m), k = 1 + (int)(r / g.m);
        for (int i = 1 + threadIdx.x; i <= g.l; i += kRedThreads) {
            const long s = gidx(g, i, j, k);
            R x[19];
#pragma unroll
            for (int d = 0; d < 19; ++d) x[d] = Storage<S>::load(a.p[d] + s);
            const R rho = density_ref(x, cte1);
            const R rhoinv = R(1) / rho;
            const R xj = (x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
            const R yj = (x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
            const R zj = (x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;
            v[0] += (double)xj;
            v[1] += (double)yj;
            v[2] += (double)zj;
            v[3] += (double)rho;
            v[4] += (double)(R(0.5) * (xj * xj + yj * yj + zj * zj));
        }
    }
    block_sum<5>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 5; ++q) part[(long)blockIdx.x * 5 + q] = v[q];
    }
}

// varm (SRC/varm.F90:56-92): velocity sums per z-plane k (block = plane)
// -> out[3*(k-1) + {0,1,2}] = {u, w(y-vel), v(z-vel)}.
template <typename S>
__global__ __launch_bounds__(kRedThreads) void varm_planes_k(Lattice19<S> a, Geometry g,
                                                             typename Storage<S>::compute cte1,
                                                             double* __restrict__ out)
{
    typedef typename Storage<S>::compute R;
    const int k = 1 + blockIdx.x;
    double v[3] = {0, 0, 0};
    const long cells = (long)g.l * g.m;
    for (long c = threadIdx.x; c < cells; c += kRedThreads) {
        const int i = 1 + (int)(c % g.l), j = 1 + (int)(c / g.l);
        const long s = gidx(g, i, j, k);
        R x[19];
#pragma unroll
        for (int d = 0; d < 19; ++d) x[d] = Storage<S>::load(a.p[d] + s);
        const R rhoinv = R(1) / density_ref(x, cte1);
        v[0] += (double)((x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv);
        v[1] += (double)((x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv);
        v[2] += (double)((x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv);
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        out[3 * (k - 1) + 0] = v[0];
        out[3 * (k - 1) + 1] = v[1];
        out[3 * (k - 1) + 2] = v[2];
    }
}

// varm density sums per y-plane j (block = plane) -> out[j-1].
template <typename S>
__global__ __launch_bounds__(kRedThreads) void varm_planes_j(Lattice19<S> a, Geometry g,
                                                             typename Storage<S>::compute cte1,
                                                             double* __restrict__ out)
{
    typedef typename Storage<S>::compute R;
    const int j = 1 + blockIdx.x;
    double v[1] = {0};
    const long cells = (long)g.l * g.n;
    for (long c = threadIdx.x; c < cells; c += kRedThreads) {
        const int i = 1 + (int)(c % g.l), k = 1 + (int)(c / g.l);
        const long s = gidx(g, i, j, k);
        R x[19];
#pragma unroll
        for (int d = 0; d < 19; ++d) x[d] = Storage<S>::load(a.p[d] + s);
        v[0] += (double)density_ref(x, cte1);
    }
    block_sum<1>(v);
    if (threadIdx.x == 0) out[j - 1] = v[0];
}

// draglift (SRC/draglift.F90:77-176): momentum exchange over fluid cells of
// the box, F = sum_d 2 c_d a_d(x) obs(x + c_d); terms in the storage kind as
// the reference forms them, accumulation in FP64.
template <typename S>
__global__ __launch_bounds__(kRedThreads) void draglift_partial(Lattice19<S> a, Geometry g,
                                                                const int* __restrict__ obs, int is, int ie,
                                                                int js, int je, int ks, int ke,
                                                                double* __restrict__ part)
{
    typedef typename Storage<S>::compute R;
    double v[3] = {0, 0, 0};
    const int ni = ie - is + 1, nj = je - js + 1, nk = ke - ks + 1;
    const long cells = (long)ni * nj * nk;
    auto O = [&](int i, int j, int k) -> int {
        return obs[(long)(i - 1) + (long)g.l * (j - 1) + (long)g.l * g.m * (k - 1)];
    };
    for (long c = (long)blockIdx.x * kRedThreads + threadIdx.x; c < cells; c += (long)gridDim.x * kRedThreads) {
        const int i = is + (int)(c % ni), j = js + (int)((c / ni) % nj), k = ks + (int)(c / ((long)ni * nj));
        if (O(i, j, k) != 0) continue;
        const long s = gidx(g, i, j, k);
        R fxs = R(0), fys = R(0), fzs = R(0);
        bool first = true;
#pragma unroll
        for (int d = 0; d < 18; ++d) {
            const R ad = Storage<S>::load(a.p[d] + s);
            const R ob = (R)O(i + cx_of(d), j + cy_of(d), k + cz_of(d));
            const R tx = R(2.0) * (R)cx_of(d) * ad * ob;
            const R ty = R(2.0) * (R)cy_of(d) * ad * ob;
            const R tz = R(2.0) * (R)cz_of(d) * ad * ob;
            if (first) {
                fxs = tx;
                fys = ty;
                fzs = tz;
                first = false;
            } else {
                fxs = fxs + tx;
                fys = fys + ty;
                fzs = fzs + tz;
            }
        }
        v[0] += (double)fxs;
        v[1] += (double)fys;
        v[2] += (double)fzs;
    }
    block_sum<3>(v);
    if (threadIdx.x == 0) {
        part[(long)blockIdx.x * 3 + 0] = v[0];
        part[(long)blockIdx.x * 3 + 1] = v[1];
        part[(long)blockIdx.x * 3 + 2] = v[2];
    }
}

// Velocity/density along a grid line (SRC/prof_i.F90:47-78,
// SRC/prof_j.F90:47-77, SRC/prof_k.F90:47-79) or at one point
// (SRC/probe.F90:44-75; same sums, same divisions): axis 0 runs i at
// (j, k) = (c1, c2), axis 1 runs j at (i, k), axis 2 runs k at (i, j), over
// first .. first+count-1.  out[4*t + {0,1,2,3}] = u, w (y-velocity),
// v (z-velocity), den, formed in the collision kind.
template <typename S>
__global__ void prof_line_kernel(Lattice19<S> a, Geometry g, int axis, int c1, int c2, int first, int count,
                                 typename Storage<S>::compute cte1, typename Storage<S>::compute* __restrict__ out)
{
    typedef typename Storage<S>::compute R;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int q = first + t;
    const long s = axis == 0 ? gidx(g, q, c1, c2) : (axis == 1 ? gidx(g, c1, q, c2) : gidx(g, c1, c2, q));
    R x[19];
#pragma unroll
    for (int d = 0; d < 19; ++d) x[d] = Storage<S>::load(a.p[d] + s);
    const R den = density_ref(x, cte1);
    out[4 * t + 0] = (x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) / den;
    out[4 * t + 1] = (x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) / den;
    out[4 * t + 2] = (x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) / den;
    out[4 * t + 3] = den;
}

// Interior macroscopic fields (rho, ux, uy, uz) in FP64, i fastest; the
// moments are formed in the collision kind with the reference's order.
template <typename S>
__global__ void fields_kernel(Lattice19<S> a, Geometry g, typename Storage<S>::compute cte1,
                              double* __restrict__ rho, double* __restrict__ ux, double* __restrict__ uy,
                              double* __restrict__ uz)
{
    typedef typename Storage<S>::compute R;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long cells = (long)g.l * g.m * g.n;
    if (t >= cells) return;
    const int i = 1 + (int)(t % g.l), j = 1 + (int)((t / g.l) % g.m), k = 1 + (int)(t / ((long)g.l * g.m));
    const long s = gidx(g, i, j, k);
    R x[19];
#pragma unroll
    for (int d = 0; d < 19; ++d) x[d] = Storage<S>::load(a.p[d] + s);
    const R r = density_ref(x, cte1);
    const R rinv = R(1) / r;
    if (rho) rho[t] = (double)r;
    if (ux) ux[t] = (double)((x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rinv);
    if (uy) uy[t] = (double)((x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rinv);
    if (uz) uz[t] = (double)((x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rinv);
}

// Interior of one population array <-> FP64 buffer (i fastest).
template <typename S>
__global__ void get_interior(const S* __restrict__ a, Geometry g, double* __restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.l * g.m * g.n) return;
    const int i = 1 + (int)(t % g.l), j = 1 + (int)((t / g.l) % g.m), k = 1 + (int)(t / ((long)g.l * g.m));
    out[t] = (double)Storage<S>::load(a + gidx(g, i, j, k));
}

template <typename S>
__global__ void set_interior(S* __restrict__ a, Geometry g, const double* __restrict__ in)
{
    typedef typename Storage<S>::compute R;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.l * g.m * g.n) return;
    const int i = 1 + (int)(t % g.l), j = 1 + (int)((t / g.l) % g.m), k = 1 + (int)(t / ((long)g.l * g.m));
    store_value(a + gidx(g, i, j, k), Rounder<S, 0>::cvt((R)in[t], 0u));
}

}  // namespace bgk
