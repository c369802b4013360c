// gfx950 kernels behind the reference-compatible C-ABI (v1).
//
// Memory is owned by the caller and laid out as in the reference
// (SRC/alloca.F90:153-191): 19 (plus 18) separate column-major arrays of
// shape (0:l+1+ipad, 0:m+1+jpad, 0:n+1), element (i,j,k) at i + nx*j + nxy*k.
//
// The fused collide lives in collide.hpp.  The boundary kernels here fill
// the halo exactly as the reference's routines do: they cover only the
// surface and keep the reference's face order, since edge cells rely on the
// last writer.
#pragma once

#include "common.hpp"
#include "d3q19.hpp"

namespace bgk {

template <typename R>
struct Lattice19 {
    R* p[19];
};

struct Geometry {
    int l, m, n;
    int nx;      // leading dimension size(a,1)
    long nxy;    // nx * size(a,2)
};

__device__ __forceinline__ long gidx(const Geometry& g, int i, int j, int k)
{
    return (long)i + (long)g.nx * (long)j + g.nxy * (long)k;
}

// -------------------------------------------------------------------------
// Boundary faces.  A(d) is population d (1-based, as in the Fortran).
// -------------------------------------------------------------------------
#define BGK_A(d, i, j, k) a.p[(d)-1][gidx(g, (i), (j), (k))]
// bcond_driven x faces at (j, k), j=1..m, k=1..n (SRC/bcond.driven.F90:105-119).
// The y and z faces are written in the same launch (driven_faces), so the
// writes they overwrite (the same population of the same halo cell) are
// skipped here; the reference's face order x, y, z makes the later value the
// one kept (SRC/bcond.driven.F90:105-247).
template <typename R>
__device__ __forceinline__ void driven_x_cell(Lattice19<R>& a, const Geometry& g, int j, int k)
{
    const int l = g.l, l1 = g.l + 1;
    const bool y_lo = j == 1, y_hi = j == g.m;  // the y faces rewrite these
    const bool z_lo = k == 1, z_hi = k == g.n;  // the z faces rewrite these
    if (!y_lo) BGK_A(12, l1, j - 1, k) = BGK_A(1, l, j, k);
    if (!z_lo) BGK_A(13, l1, j, k - 1) = BGK_A(2, l, j, k);
    if (!y_hi) BGK_A(10, l1, j + 1, k) = BGK_A(3, l, j, k);
    if (!z_hi) BGK_A(11, l1, j, k + 1) = BGK_A(4, l, j, k);
    BGK_A(14, l1, j, k) = BGK_A(5, l, j, k);
    if (!y_lo) BGK_A(3, 0, j - 1, k) = BGK_A(10, 1, j, k);
    if (!z_lo) BGK_A(4, 0, j, k - 1) = BGK_A(11, 1, j, k);
    if (!y_hi) BGK_A(1, 0, j + 1, k) = BGK_A(12, 1, j, k);
    if (!z_hi) BGK_A(2, 0, j, k + 1) = BGK_A(13, 1, j, k);
    BGK_A(5, 0, j, k) = BGK_A(14, 1, j, k);
}

// bcond_driven y faces at (i, k), i=1..l, k=1..n (SRC/bcond.driven.F90:169-183);
// the z faces follow in the same launch, so their cells are skipped as in
// driven_x_cell.
template <typename R>
__device__ __forceinline__ void driven_y_cell(Lattice19<R>& a, const Geometry& g, int i, int k)
{
    const int m = g.m, m1 = g.m + 1;
    const bool z_lo = k == 1, z_hi = k == g.n;
    BGK_A(8, i, 0, k) = BGK_A(17, i, 1, k);
    BGK_A(12, i + 1, 0, k) = BGK_A(1, i, 1, k);
    BGK_A(3, i - 1, 0, k) = BGK_A(10, i, 1, k);
    if (!z_lo) BGK_A(7, i, 0, k - 1) = BGK_A(16, i, 1, k);
    if (!z_hi) BGK_A(9, i, 0, k + 1) = BGK_A(18, i, 1, k);
    BGK_A(10, i + 1, m1, k) = BGK_A(3, i, m, k);
    if (!z_hi) BGK_A(16, i, m1, k + 1) = BGK_A(7, i, m, k);
    BGK_A(17, i, m1, k) = BGK_A(8, i, m, k);
    if (!z_lo) BGK_A(18, i, m1, k - 1) = BGK_A(9, i, m, k);
    BGK_A(1, i - 1, m1, k) = BGK_A(12, i, m, k);
}

// bcond_driven z faces + lid, i=1..l, j=1..m (SRC/bcond.driven.F90:233-247);
// `lid` = 1 moves the top wall (driven cavity), 0 makes it a plain no-slip
// wall (the channel walls of SRC/bcond.channel.F90:244-267).
template <typename R>
__device__ __forceinline__ void walls_z_cell(Lattice19<R>& a, const Geometry& g, int i, int j, R force, int lid)
{
    const int n = g.n, n1 = g.n + 1;
    BGK_A(6, i, j, 0) = BGK_A(15, i, j, 1);
    BGK_A(4, i - 1, j, 0) = BGK_A(11, i, j, 1);
    BGK_A(7, i, j - 1, 0) = BGK_A(16, i, j, 1);
    BGK_A(13, i + 1, j, 0) = BGK_A(2, i, j, 1);
    BGK_A(18, i, j + 1, 0) = BGK_A(9, i, j, 1);
    BGK_A(15, i, j, n1) = BGK_A(6, i, j, n);
    if (lid) {
        BGK_A(2, i - 1, j, n1) = BGK_A(13, i, j, n) + force;
    } else {
        BGK_A(2, i - 1, j, n1) = BGK_A(13, i, j, n);
    }
    BGK_A(9, i, j - 1, n1) = BGK_A(18, i, j, n);
    if (lid) {
        BGK_A(11, i + 1, j, n1) = BGK_A(4, i, j, n) - force;
    } else {
        BGK_A(11, i + 1, j, n1) = BGK_A(4, i, j, n);
    }
    BGK_A(16, i, j + 1, n1) = BGK_A(7, i, j, n);
}

// bcond_channel's y faces and z walls in one launch (SRC/bcond.channel.F90,
// after the x faces): threads [0, (l+2)(n+2)) copy the y faces over
// i = 0..l+1, k = 0..n+1 as periodic_y does, the last l*m threads write the
// no-slip walls at z = 0, n+1 (walls_z_cell, lid off).  No cell the y copy reads is
// written by the walls, and the four populations both write at the same
// cell (7 and 18 at k = 0, 9 and 16 at k = n+1, for 1 <= i <= l) keep the
// walls' value, written last in the reference: the y copy skips them there.
template <typename R>
__global__ void channel_yz(Lattice19<R> a, Geometry g)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ni = g.l + 2;
    const long ny_f = (long)ni * (g.n + 2), nz_f = (long)g.l * g.m;
    if (t < ny_f) {
        const int i = (int)(t % ni), k = (int)(t / ni), m = g.m, m1 = g.m + 1;
        const bool wall_col = i >= 1 && i <= g.l;
        const bool lo = wall_col && k == 0, hi = wall_col && k == g.n + 1;
        BGK_A(3, i, 0, k) = BGK_A(3, i, m, k);
        if (!lo) BGK_A(7, i, 0, k) = BGK_A(7, i, m, k);
        BGK_A(8, i, 0, k) = BGK_A(8, i, m, k);
        if (!hi) BGK_A(9, i, 0, k) = BGK_A(9, i, m, k);
        BGK_A(12, i, 0, k) = BGK_A(12, i, m, k);
        BGK_A(1, i, m1, k) = BGK_A(1, i, 1, k);
        BGK_A(10, i, m1, k) = BGK_A(10, i, 1, k);
        if (!hi) BGK_A(16, i, m1, k) = BGK_A(16, i, 1, k);
        BGK_A(17, i, m1, k) = BGK_A(17, i, 1, k);
        if (!lo) BGK_A(18, i, m1, k) = BGK_A(18, i, 1, k);
    } else if (t < ny_f + nz_f) {
        const long u = t - ny_f;
        walls_z_cell(a, g, 1 + (int)(u % g.l), 1 + (int)(u / g.l), R(0), 0);
    }
}

// bcond_driven in one launch: threads [0, m*n) take the x faces, the next
// l*n the y faces, the last l*m the z faces with the lid.  Every value read
// is an interior cell, which no face writes, so the faces interact only
// where two of them write the same population of the same halo cell; there
// the earlier face (in the reference's order x, y, z) skips its write and the
// result equals the three sequential launches.
template <typename R>
__global__ void driven_faces(Lattice19<R> a, Geometry g, R force)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nx_f = (long)g.m * g.n, ny_f = (long)g.l * g.n, nz_f = (long)g.l * g.m;
    if (t < nx_f) {
        driven_x_cell<R>(a, g, 1 + (int)(t % g.m), 1 + (int)(t / g.m));
    } else if (t < nx_f + ny_f) {
        const long u = t - nx_f;
        driven_y_cell<R>(a, g, 1 + (int)(u % g.l), 1 + (int)(u / g.l));
    } else if (t < nx_f + ny_f + nz_f) {
        const long u = t - nx_f - ny_f;
        walls_z_cell(a, g, 1 + (int)(u % g.l), 1 + (int)(u / g.l), force, 1);
    }
}

// periodic x face over j=0..m+1, k=0..n+1 (SRC/bcond.periodic.F90:105-119),
// one thread per (population, row): q = 0..4 copies direction q+1 from
// i = l to i = 0, q = 5..9 direction q+5 from i = 1 to i = l+1.  Each value
// lands in a line of its own, so the face is bound by line transactions,
// not bytes; a thread per row (ten values each) left one wave per SIMD.
// Rows are counted in whole waves per population (x_face_rows), so q is the
// same across a wave and the population pointer is a scalar.
__host__ __device__ inline long x_face_rows(const Geometry& g) { return ((long)(g.m + 2) * (g.n + 2) + 63) / 64 * 64; }

template <typename R>
__global__ void periodic_x(Lattice19<R> a, Geometry g)
{
    const long rows = (long)(g.m + 2) * (g.n + 2), padded = x_face_rows(g);
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int q = __builtin_amdgcn_readfirstlane((int)(t / padded));
    const long row = t - (long)q * padded;
    if (q >= 10 || row >= rows) return;
    const int nj = g.m + 2;
    const int j = (int)(row % nj), k = (int)(row / nj);
    if (q < 5)
        BGK_A(q + 1, 0, j, k) = BGK_A(q + 1, g.l, j, k);
    else
        BGK_A(q + 5, g.l + 1, j, k) = BGK_A(q + 5, 1, j, k);
}

// periodic y face over i=0..l+1, k=0..n+1 (SRC/bcond.periodic.F90:171-185)
template <typename R>
__global__ void periodic_y(Lattice19<R> a, Geometry g)
{
    const int ni = g.l + 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * (g.n + 2)) return;
    const int i = t % ni, k = t / ni, m = g.m, m1 = g.m + 1;
    BGK_A(3, i, 0, k) = BGK_A(3, i, m, k);
    BGK_A(7, i, 0, k) = BGK_A(7, i, m, k);
    BGK_A(8, i, 0, k) = BGK_A(8, i, m, k);
    BGK_A(9, i, 0, k) = BGK_A(9, i, m, k);
    BGK_A(12, i, 0, k) = BGK_A(12, i, m, k);
    BGK_A(1, i, m1, k) = BGK_A(1, i, 1, k);
    BGK_A(10, i, m1, k) = BGK_A(10, i, 1, k);
    BGK_A(16, i, m1, k) = BGK_A(16, i, 1, k);
    BGK_A(17, i, m1, k) = BGK_A(17, i, 1, k);
    BGK_A(18, i, m1, k) = BGK_A(18, i, 1, k);
}

// periodic z face over i=0..l+1, j=0..m+1 (SRC/bcond.periodic.F90:238-252)
template <typename R>
__global__ void periodic_z(Lattice19<R> a, Geometry g)
{
    const int ni = g.l + 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * (g.m + 2)) return;
    const int i = t % ni, j = t / ni, n = g.n, n1 = g.n + 1;
    BGK_A(4, i, j, 0) = BGK_A(4, i, j, n);
    BGK_A(6, i, j, 0) = BGK_A(6, i, j, n);
    BGK_A(7, i, j, 0) = BGK_A(7, i, j, n);
    BGK_A(13, i, j, 0) = BGK_A(13, i, j, n);
    BGK_A(18, i, j, 0) = BGK_A(18, i, j, n);
    BGK_A(2, i, j, n1) = BGK_A(2, i, j, 1);
    BGK_A(9, i, j, n1) = BGK_A(9, i, j, 1);
    BGK_A(11, i, j, n1) = BGK_A(11, i, j, 1);
    BGK_A(15, i, j, n1) = BGK_A(15, i, j, 1);
    BGK_A(16, i, j, n1) = BGK_A(16, i, j, 1);
}

// inflow/outflow x faces over j=0..m+1, k=0..n+1
// (SRC/bcond.inflow.F90:131-181).  Note the reference writes crho*p2*(...)
// = (crho*p2)*(...) and forms c.u terms as rf*(-xj-yj) etc.; each line below
// keeps that grouping.
template <typename R>
__global__ void inflow_x(Lattice19<R> a, Geometry g, R u_inflow, R rf, R qf, R p1, R p2, R cte1)
{
    const int nj = g.m + 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nj * (g.n + 2)) return;
    const int j = t % nj, k = t / nj, l = g.l, l1 = g.l + 1;
    // `3.0` is a literal in the reference expression, not the tre argument
    const R three = R(3), crho = R(1), rhoinv = R(1);
    R f[19];
#pragma unroll
    for (int d = 0; d < 19; ++d) f[d] = a.p[d][gidx(g, l, j, k)];
    R xj = +(f[0] + f[1] + f[2] + f[3] + f[4] - f[9] - f[10] - f[11] - f[12] - f[13]) * rhoinv;
    R yj = +(f[2] + f[6] + f[7] + f[8] + f[11] - f[0] - f[9] - f[15] - f[16] - f[17]) * rhoinv;
    R zj = +(f[3] + f[5] + f[6] + f[12] + f[17] - f[1] - f[8] - f[10] - f[14] - f[15]) * rhoinv;
    R cvsq = xj * xj + yj * yj + zj * zj;
    const R cx10 = rf * (-xj - yj) + qf * (three * (xj + yj) * (xj + yj) - cvsq);
    const R cx11 = rf * (-xj - zj) + qf * (three * (xj + zj) * (xj + zj) - cvsq);
    const R cx12 = rf * (-xj + yj) + qf * (three * (xj - yj) * (xj - yj) - cvsq);
    const R cx13 = rf * (-xj + zj) + qf * (three * (xj - zj) * (xj - zj) - cvsq);
    const R cx14 = rf * (-xj) + qf * (three * (xj) * (xj)-cvsq);
    BGK_A(10, l1, j, k) = crho * p2 * (cte1 + cx10);
    BGK_A(11, l1, j, k) = crho * p2 * (cte1 + cx11);
    BGK_A(12, l1, j, k) = crho * p2 * (cte1 + cx12);
    BGK_A(13, l1, j, k) = crho * p2 * (cte1 + cx13);
    BGK_A(14, l1, j, k) = crho * p1 * (cte1 + cx14);
    xj = u_inflow;
    yj = R(0);
    zj = R(0);
    cvsq = xj * xj + yj * yj + zj * zj;
    const R cx01 = rf * (xj - yj) + qf * (three * (xj - yj) * (xj - yj) - cvsq);
    const R cx02 = rf * (xj - zj) + qf * (three * (xj - zj) * (xj - zj) - cvsq);
    const R cx03 = rf * (xj + yj) + qf * (three * (xj + yj) * (xj + yj) - cvsq);
    const R cx04 = rf * (xj + zj) + qf * (three * (xj + zj) * (xj + zj) - cvsq);
    const R cx05 = rf * (xj) + qf * (three * (xj) * (xj)-cvsq);
    BGK_A(1, 0, j, k) = crho * p2 * (cte1 + cx01);
    BGK_A(2, 0, j, k) = crho * p2 * (cte1 + cx02);
    BGK_A(3, 0, j, k) = crho * p2 * (cte1 + cx03);
    BGK_A(4, 0, j, k) = crho * p2 * (cte1 + cx04);
    BGK_A(5, 0, j, k) = crho * p1 * (cte1 + cx05);
}

// Obstacle bounce-back (SRC/bcond.obs.F90:101-122): for every solid cell S
// of the box, a_d(S) = a_opp(d)(S + c_d), which is what the fluid cell
// X = S + c_d pulls as population d next (half-way bounce-back of its own
// opposite population).  Only links whose neighbour S + c_d is fluid are
// written, so every read comes from a fluid cell this kernel never writes:
// the result is deterministic and fluid cells see exactly the reference's
// values.  (The reference's own
// GPU kernel and CPU loop also read solid neighbours they overwrite in the
// same pass; those values only ever reach other solid cells.)
template <typename R>
__global__ void obstacle_bb(Lattice19<R> a, Geometry g, const int* __restrict__ obs, int imin, int imax, int jmin,
                            int jmax, int kmin, int kmax)
{
    const int ni = imax - imin + 1, nj = jmax - jmin + 1, nk = kmax - kmin + 1;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ni * nj * nk) return;
    const int i = imin + (int)(t % ni);
    const int j = jmin + (int)((t / ni) % nj);
    const int k = kmin + (int)(t / ((long)ni * nj));
    auto solid = [&](int ii, int jj, int kk) -> bool {
        if (ii < 1 || ii > g.l || jj < 1 || jj > g.m || kk < 1 || kk > g.n) return false;
        return obs[(long)(ii - 1) + (long)g.l * (jj - 1) + (long)g.l * g.m * (kk - 1)] == 1;
    };
    if (!solid(i, j, k)) return;
    const long c = gidx(g, i, j, k);
#pragma unroll
    for (int d = 0; d < 18; ++d) {
        // reference: a_d(i,j,k) = a_opp(d)(i + c_d)
        const int ii = i + cx_of(d), jj = j + cy_of(d), kk = k + cz_of(d);
        if (solid(ii, jj, kk)) continue;
        a.p[d][c] = a.p[opp_of(d)][gidx(g, ii, jj, kk)];
    }
}

#undef BGK_A

}  // namespace bgk
