// Device form of the initial state f = f_eq (SRC/init.F90:31-209): the host
// evaluates the sin/cos factors of every coordinate once (l + m + n + 6
// triples, tg_axis_table), the kernel forms each cell's 19 shifted
// populations from them with feq_cell, the function init_feq_host uses, so
// the device state is bit-identical to the host's and to the reference's.
// A 256^3 lattice needs 780 factor triples instead of 17 million per-cell
// sin/cos evaluations and host-to-device copies of the whole lattice.
#pragma once

#include "common.hpp"
#include "host_init.hpp"
#include "storage.hpp"

namespace bgk {

template <typename S>
struct InitArgs {
    typedef typename Storage<S>::compute C;
    S* p[19];               // cell (i,j,k) of population d at p[d] + i + nx*j + zstride*k; null: skipped
    const TgAxis<C>* tab;   // tg_axis_table(l, m, n)
    int l, m, n, nx;
    long zstride;
    int mode;
    C u00;
};

// One thread per cell of (0:l+1, 0:m+1, 0:n+1), i fastest.
template <typename S>
__global__ void init_feq_kernel(InitArgs<S> a)
{
    typedef typename Storage<S>::compute C;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ni = a.l + 2, nj = a.m + 2;
    if (t >= (long)ni * nj * (a.n + 2)) return;
    const int i = (int)(t % ni);
    const long r = t / ni;
    const int j = (int)(r % nj), k = (int)(r / nj);
    C v[19];
    feq_cell<C>(a.mode, a.u00, a.tab[i], a.tab[ni + j], a.tab[ni + nj + k], v);
    const long id = (long)i + (long)a.nx * j + a.zstride * k;
#pragma unroll
    for (int d = 0; d < 19; ++d)
        if (a.p[d]) store_value(a.p[d] + id, Rounder<S, 0>::cvt(v[d], 0u));
}

// Fill the populations p[d] (null entries skipped) over (0:l+1, 0:m+1,
// 0:n+1) on `stream` and wait for it.  The factor table lives in a
// temporary device buffer for the duration of the call.
template <typename S>
hipError_t init_feq_device(S* const* p, int l, int m, int n, int nx, long zstride, int mode,
                           typename Storage<S>::compute u00, hipStream_t stream)
{
    typedef typename Storage<S>::compute C;
    const std::vector<TgAxis<C>> tab = tg_axis_table<C>(l, m, n);
    const size_t bytes = tab.size() * sizeof(TgAxis<C>);
    TgAxis<C>* dtab = nullptr;
    hipError_t e = hipMalloc(&dtab, bytes);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(dtab, tab.data(), bytes, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) {
        InitArgs<S> a{};
        for (int d = 0; d < 19; ++d) a.p[d] = p[d];
        a.tab = dtab;
        a.l = l;
        a.m = m;
        a.n = n;
        a.nx = nx;
        a.zstride = zstride;
        a.mode = mode;
        a.u00 = u00;
        const long cells = (long)(l + 2) * (m + 2) * (n + 2);
        hipLaunchKernelGGL(init_feq_kernel<S>, dim3((unsigned)ceil_div(cells, 256L)), dim3(256), 0, stream, a);
        e = hipGetLastError();
    }
    const hipError_t w = hipStreamSynchronize(stream);
    (void)hipFree(dtab);
    return e != hipSuccess ? e : w;
}

}  // namespace bgk
