// Device side of the reference's output writers: the FP32 values its VTK
// plane and vorticity writers compute, in the reference's mixed-kind
// arithmetic, so that the files the host writes match the reference's byte
// for byte.
#pragma once

#include "common.hpp"
#include "d3q19.hpp"
#include "storage.hpp"
#include "v1_kernels.hpp"

namespace bgk {

// Velocity planes (SRC/vtk_xy.binary.F90:76-97, SRC/vtk_xz.binary.F90:78-96,
// SRC/vtk_yz.binary.F90:78-96): u, w (y), v (z) and den are sums of the
// stored values in the working kind assigned to real(sp), then divided in
// FP32.  plane 0: xy at k = idx, points (i fastest, then j); 1: xz at
// j = idx (i, then k); 2: yz at i = idx (j, then k).  out[3*t + {0,1,2}] =
// u/den, w/den, v/den.
template <typename S>
__global__ void vtk_plane_kernel(Lattice19<S> a, Geometry g, int plane, int idx, float* __restrict__ out)
{
    typedef typename Storage<S>::compute R;
    const int n1 = plane == 2 ? g.m : g.l;
    const int n2 = plane == 0 ? g.m : g.n;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n1 * n2) return;
    const int p = 1 + (int)(t % n1), q = 1 + (int)(t / n1);
    const long s = plane == 0 ? gidx(g, p, q, idx) : (plane == 1 ? gidx(g, p, idx, q) : gidx(g, idx, p, q));
    R x[19];
#pragma unroll
    for (int d = 0; d < 19; ++d) x[d] = Storage<S>::load(a.p[d] + s);
    const float u = (float)(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]);
    const float w = (float)(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]);
    const float v = (float)(x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]);
    // cte1 is real(sp) 1.0, added to the working-kind sum
    const float den = (float)density_ref(x, R(1.0f));
    out[3 * t + 0] = u / den;
    out[3 * t + 1] = w / den;
    out[3 * t + 2] = v / den;
}

// Velocity of one cell as SRC/vtk_om.xy.binary.F90:69-98 forms it: the 19
// values rounded to real(sp), rho summed in FP32, rhoinv = uno/rho with
// uno of the working kind, momenta times rhoinv in FP32.
template <typename S>
__device__ __forceinline__ void velocity_sp(const Lattice19<S>& a, long s, float& uu, float& vv, float& ww)
{
    typedef typename Storage<S>::compute R;
    float x[19];
#pragma unroll
    for (int d = 0; d < 19; ++d) x[d] = (float)Storage<S>::load(a.p[d] + s);
    float rho = x[0];
#pragma unroll
    for (int d = 1; d < 19; ++d) rho = rho + x[d];
    rho = rho + 1.0f;
    const float rhoinv = (float)(R(1) / (R)rho);
    uu = (x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
    vv = (x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
    ww = (x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;
}

// Vorticity on the plane k = k0 at i = 2..l-1, j = 2..m-1 (i fastest)
// (SRC/vtk_om.xy.binary.F90:120-144): central differences 0.5*(f(+1) -
// f(-1)) of the FP32 velocities; out[3*t + {0,1,2}] = dvdz - dwdy,
// dwdx - dudz, dudy - dvdx with u, v, w the x, y, z velocities -- the
// reference's sign convention, kept.
template <typename S>
__global__ void vorticity_xy_kernel(Lattice19<S> a, Geometry g, int k0, float* __restrict__ out)
{
    const int ni = g.l - 2, nj = g.m - 2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ni * nj) return;
    const int i = 2 + (int)(t % ni), j = 2 + (int)(t / ni);
    float uxp, vxp, wxp, uxm, vxm, wxm, uyp, vyp, wyp, uym, vym, wym, uzp, vzp, wzp, uzm, vzm, wzm;
    velocity_sp(a, gidx(g, i + 1, j, k0), uxp, vxp, wxp);
    velocity_sp(a, gidx(g, i - 1, j, k0), uxm, vxm, wxm);
    velocity_sp(a, gidx(g, i, j + 1, k0), uyp, vyp, wyp);
    velocity_sp(a, gidx(g, i, j - 1, k0), uym, vym, wym);
    velocity_sp(a, gidx(g, i, j, k0 + 1), uzp, vzp, wzp);
    velocity_sp(a, gidx(g, i, j, k0 - 1), uzm, vzm, wzm);
    (void)uxp, (void)uxm, (void)vyp, (void)vym, (void)wzp, (void)wzm;
    const float dudy = 0.5f * (uyp - uym), dudz = 0.5f * (uzp - uzm);
    const float dvdx = 0.5f * (vxp - vxm), dvdz = 0.5f * (vzp - vzm);
    const float dwdx = 0.5f * (wxp - wxm), dwdy = 0.5f * (wyp - wym);
    out[3 * t + 0] = dvdz - dwdy;
    out[3 * t + 1] = dwdx - dudz;
    out[3 * t + 2] = dudy - dvdx;
}

// Halo-inclusive copy of one population array (0..l+1, 0..m+1, 0..n+1, i
// fastest) between the device layout and a dense buffer, in the storage type.
template <typename S>
__global__ void pack_halo(const S* __restrict__ src, Geometry g, S* __restrict__ dst)
{
    const long nx2 = g.l + 2, ny2 = g.m + 2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx2 * ny2 * (g.n + 2)) return;
    const int i = (int)(t % nx2), j = (int)((t / nx2) % ny2), k = (int)(t / (nx2 * ny2));
    dst[t] = src[gidx(g, i, j, k)];
}

template <typename S>
__global__ void unpack_halo(const S* __restrict__ src, Geometry g, S* __restrict__ dst)
{
    const long nx2 = g.l + 2, ny2 = g.m + 2;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nx2 * ny2 * (g.n + 2)) return;
    const int i = (int)(t % nx2), j = (int)((t / nx2) % ny2), k = (int)(t / (nx2 * ny2));
    dst[gidx(g, i, j, k)] = src[t];
}

}  // namespace bgk
