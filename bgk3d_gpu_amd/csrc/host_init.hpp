// Host-side initial state f = f_eq (shifted storage), shared by the
// reference-compatible ABI and the handle API.
//
// SRC/init.F90:31-209 with cte1 = 0 (shifted storage, :71-75; with the flag
// kInitUnshifted in `mode`, cte1 = 1: the populations themselves, as the
// reference's -DNOSHIFT builds store them):
//   mode 0: rest flow -> every population is 0;
//   mode 1: 3-D Taylor-Green, u = u00 sin x cos y cos z, v = -u00 cos x
//           sin y cos z, w = 0, crho = 1 + u00^2/48 (cos 2x + cos 2y)
//           (cos 2z + 2) in 2 pi units (:142-148);
//   mode 2: mode 1 with the density perturbation kept: the shifted form
//           f - w = w (rho - 1) + w rho c, where the reference stores only
//           w rho c (no reference counterpart).
// The sin/cos factors run on the host with the C library's functions of the
// working kind, like the reference; the per-cell arithmetic (feq_cell) runs
// on the host (init_feq_host) or on the device (init_feq_kernel, from a
// per-axis table of the factors) with the same bits, so either state is
// bit-identical to the reference's.
// Mixed-kind details kept: pi is a default-real literal (:63); the
// coordinates are formed in double from (real(i,mykind)-0.5d0)/real(l,mykind)
// and rounded to the working kind (:115-123).
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

#include "d3q19.hpp"

namespace bgk {

// Flag of the init mode: unshifted storage (f instead of f - w), the
// reference's -DNOSHIFT convention (SRC/init.F90:71-75).  Modes 0 and 1 only.
constexpr int kInitUnshifted = 8;

// Trigonometric factors of one coordinate as the reference forms them
// (SRC/init.F90:115-123, 142-148): the coordinate (i - 1/2)/l rounded to the
// working kind, then sin 2 pi x, cos 2 pi x and cos 4 pi x with the C
// library's functions of that kind.  Always evaluated on the host.
template <typename R>
struct TgAxis {
    R s2, c2, c4;
};

template <typename R>
inline TgAxis<R> tg_axis(int i, int l)
{
    const R due = 2, qua = 4;
    const R pi = (R)3.141592653589793238462643383279f;
    const R x = (R)(((double)(R)i - 0.5) / (double)(R)l);
    return {std::sin(due * pi * x), std::cos(due * pi * x), std::cos(qua * pi * x)};
}

// Shifted f_eq (f - w) of the 19 populations of one cell from the factors of
// its three coordinates, in the reference's expression order.  The same
// function runs on the host (init_feq_host) and on the device
// (init_feq_kernel), so both give the same bits: only +, -, * and one
// correctly rounded division, no contraction.
template <typename R>
__host__ __device__ inline void feq_cell(int mode, R u00, const TgAxis<R>& X, const TgAxis<R>& Y, const TgAxis<R>& Z,
                                         R (&h)[19])
{
    const R uno = 1, due = 2, tre = 3, qua = 4, zero = 0, rf = 3, qf = R(1.5);
    const R cte1 = (mode & kInitUnshifted) ? uno : zero;
    mode &= ~kInitUnshifted;
    const R w[19] = {W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p1, W<R>::p1, W<R>::p2,
                     W<R>::p1, W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p1,
                     W<R>::p1, W<R>::p2, W<R>::p1, W<R>::p2, W<R>::p0};
    R xj = zero, yj = zero, zj = zero, crho = uno, drho = zero;
    if (mode == 1 || mode == 2) {
        xj = u00 * X.s2 * Y.c2 * Z.c2;
        yj = -u00 * X.c2 * Y.s2 * Z.c2;
        drho = ((u00 * u00) / (qua * qua * tre)) * (X.c4 + Y.c4) * (Z.c4 + due);
        crho = uno + drho;
    }
    const R cvsq = xj * xj + yj * yj + zj * zj;
    // c.u per direction in the reference's grouping (SRC/init.F90:159-177)
    const R lin[19] = {xj - yj, xj - zj, xj + yj, xj + zj, xj, zj, yj + zj, yj, yj - zj, -xj - yj,
                       -xj - zj, -xj + yj, -xj + zj, -xj, -zj, -yj - zj, -yj, -yj + zj, zero};
    const R sq[19] = {xj - yj, xj - zj, xj + yj, xj + zj, xj, zj, yj + zj, yj, yj - zj, xj + yj,
                      xj + zj, xj - yj, xj - zj, xj, zj, yj + zj, yj, yj - zj, zero};
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        const R c = rf * lin[d] + qf * (tre * sq[d] * sq[d] - cvsq);
        h[d] = mode == 2 ? w[d] * (drho + crho * c) : crho * w[d] * (cte1 + c);
    }
}

// Planes k0..k1 (inclusive, 0 <= k0 <= k1 <= n+1) are written, plane k at
// h[d] + (k - k0)*nx*ny; the whole lattice is k0 = 0, k1 = n+1.
template <typename R>
int init_feq_host(R** h, int l, int m, int n, int nx, int ny, int mode, R u00, int k0 = 0, int k1 = -1)
{
    if (mode < 0 || (mode & ~kInitUnshifted) > 2 || mode == (2 | kInitUnshifted)) return -1;
    if (k1 < 0) k1 = n + 1;
    if (k0 < 0 || k1 > n + 1 || k0 > k1) return -1;
    const size_t nxy = (size_t)nx * (size_t)ny;
    for (int k = k0; k <= k1; ++k) {
        const TgAxis<R> Z = tg_axis<R>(k, n);
        for (int j = 0; j <= m + 1; ++j) {
            const TgAxis<R> Y = tg_axis<R>(j, m);
            for (int i = 0; i <= l + 1; ++i) {
                R v[19];
                feq_cell<R>(mode, u00, tg_axis<R>(i, l), Y, Z, v);
                const size_t id = (size_t)i + (size_t)nx * j + nxy * (size_t)(k - k0);
                for (int d = 0; d < 19; ++d) h[d][id] = v[d];
            }
        }
    }
    return 0;
}

// Factors of every coordinate of a lattice, for init_feq_kernel: i = 0..l+1,
// then j = 0..m+1, then k = 0..n+1.
template <typename R>
std::vector<TgAxis<R>> tg_axis_table(int l, int m, int n)
{
    std::vector<TgAxis<R>> t;
    t.reserve((size_t)(l + m + n + 6));
    for (int i = 0; i <= l + 1; ++i) t.push_back(tg_axis<R>(i, l));
    for (int j = 0; j <= m + 1; ++j) t.push_back(tg_axis<R>(j, m));
    for (int k = 0; k <= n + 1; ++k) t.push_back(tg_axis<R>(k, n));
    return t;
}

}  // namespace bgk
