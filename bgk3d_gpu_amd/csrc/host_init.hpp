// Host-side initial state f = f_eq (shifted storage), shared by the
// reference-compatible ABI and the handle API.
//
// SRC/init.F90:31-209 with cte1 = 0 (shifted storage, :71-75):
//   mode 0: rest flow -> every population is 0;
//   mode 1: 3-D Taylor-Green, u = u00 sin x cos y cos z, v = -u00 cos x
//           sin y cos z, w = 0, crho = 1 + u00^2/48 (cos 2x + cos 2y)
//           (cos 2z + 2) in 2 pi units (:142-148);
//   mode 2: mode 1 with the density perturbation kept: the shifted form
//           f - w = w (rho - 1) + w rho c, where the reference stores only
//           w rho c (no reference counterpart).
// Runs on the host with the C library's sin/cos of the working kind, like
// the reference, so the state is bit-identical to the reference's.
// Mixed-kind details kept: pi is a default-real literal (:63); the
// coordinates are formed in double from (real(i,mykind)-0.5d0)/real(l,mykind)
// and rounded to the working kind (:115-123).
#pragma once

#include <cmath>
#include <cstddef>

#include "d3q19.hpp"

namespace bgk {

// Planes k0..k1 (inclusive, 0 <= k0 <= k1 <= n+1) are written, plane k at
// h[d] + (k - k0)*nx*ny; the whole lattice is k0 = 0, k1 = n+1.
template <typename R>
int init_feq_host(R** h, int l, int m, int n, int nx, int ny, int mode, R u00, int k0 = 0, int k1 = -1)
{
    if (mode < 0 || mode > 2) return -1;
    if (k1 < 0) k1 = n + 1;
    if (k0 < 0 || k1 > n + 1 || k0 > k1) return -1;
    const size_t nxy = (size_t)nx * (size_t)ny;
    const R uno = 1, due = 2, tre = 3, qua = 4, zero = 0, rf = 3, qf = R(1.5), cte1 = 0;
    const R pi = (R)3.141592653589793238462643383279f;
    const R w[19] = {W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p1, W<R>::p1, W<R>::p2,
                     W<R>::p1, W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p2, W<R>::p1,
                     W<R>::p1, W<R>::p2, W<R>::p1, W<R>::p2, W<R>::p0};
    for (int k = k0; k <= k1; ++k) {
        const R z = (R)(((double)(R)k - 0.5) / (double)(R)n);
        for (int j = 0; j <= m + 1; ++j) {
            const R y = (R)(((double)(R)j - 0.5) / (double)(R)m);
            for (int i = 0; i <= l + 1; ++i) {
                const R x = (R)(((double)(R)i - 0.5) / (double)(R)l);
                R xj = zero, yj = zero, zj = zero, crho = uno, drho = zero;
                if (mode == 1 || mode == 2) {
                    xj = u00 * std::sin(due * pi * x) * std::cos(due * pi * y) * std::cos(due * pi * z);
                    yj = -u00 * std::cos(due * pi * x) * std::sin(due * pi * y) * std::cos(due * pi * z);
                    drho = ((u00 * u00) / (qua * qua * tre)) * (std::cos(qua * pi * x) + std::cos(qua * pi * y)) *
                           (std::cos(qua * pi * z) + due);
                    crho = uno + drho;
                }
                const R cvsq = xj * xj + yj * yj + zj * zj;
                // c.u per direction in the reference's grouping (SRC/init.F90:159-177)
                const R lin[19] = {xj - yj, xj - zj, xj + yj, xj + zj, xj, zj, yj + zj, yj, yj - zj, -xj - yj,
                                   -xj - zj, -xj + yj, -xj + zj, -xj, -zj, -yj - zj, -yj, -yj + zj, zero};
                const R sq[19] = {xj - yj, xj - zj, xj + yj, xj + zj, xj, zj, yj + zj, yj, yj - zj, xj + yj,
                                  xj + zj, xj - yj, xj - zj, xj, zj, yj + zj, yj, yj - zj, zero};
                const size_t id = (size_t)i + (size_t)nx * j + nxy * (size_t)(k - k0);
                for (int d = 0; d < 19; ++d) {
                    const R c = rf * lin[d] + qf * (tre * sq[d] * sq[d] - cvsq);
                    h[d][id] = mode == 2 ? w[d] * (drho + crho * c) : crho * w[d] * (cte1 + c);
                }
            }
        }
    }
    return 0;
}

}  // namespace bgk
