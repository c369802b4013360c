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

// ---- periodic faces in one launch -------------------------------------
// The reference fills the periodic faces one after the other, each over the
// whole face including the halo rows of the other axes: x over j = 0..m+1,
// k = 0..n+1, then y over i = 0..l+1, k = 0..n+1, then z over the full
// (i, j) face (SRC/bcond.periodic.F90:105-252).  A face writes population d
// at a halo cell when d enters the interior across it (c = +1 on the low
// side, c = -1 on the high side) and copies it from the opposite interior
// layer, reading values earlier faces may have just written.  Followed
// through, the last value of population d at halo cell h is the old value at
// h wrapped along every axis on which h lies in the halo and d enters, and
// it is written only if there is such an axis.  Each (cell, population) is
// therefore written by exactly one thread here -- the one of the last face
// in the order x, y, z that writes it -- and every value read lies, on each
// axis, either inside or on a halo side the population does not enter, so
// no thread of the launch writes it: one launch gives the three sequential
// fills' result, edges and corners included.

// -1: low halo (coordinate 0), +1: high halo (max + 1), 0: inside
__device__ __forceinline__ int halo_side(int c, int max) { return c == 0 ? -1 : (c == max + 1 ? 1 : 0); }
// population with velocity component c enters the interior across that side
__device__ __forceinline__ bool enters(int side, int c) { return (side < 0 && c > 0) || (side > 0 && c < 0); }
__device__ __forceinline__ int wrap_to(int side, int max) { return side < 0 ? max : 1; }

// The populations entering across each face, 0-based (direction - 1):
// x low / high, y low / high, z low / high.
__device__ constexpr int kEnterXlo[5] = {0, 1, 2, 3, 4}, kEnterXhi[5] = {9, 10, 11, 12, 13};
__device__ constexpr int kEnterYlo[5] = {2, 6, 7, 8, 11}, kEnterYhi[5] = {0, 9, 15, 16, 17};
__device__ constexpr int kEnterZlo[5] = {3, 5, 6, 12, 17}, kEnterZhi[5] = {1, 8, 10, 14, 15};

// Rows of the x face counted in whole waves per population, so that the
// population index q is the same across a wave.
__host__ __device__ inline long x_face_rows(const Geometry& g) { return ((long)(g.m + 2) * (g.n + 2) + 63) / 64 * 64; }

// x face, one thread per (population, row): q = 0..4 sets direction q+1 at
// i = 0 from i = l, q = 5..9 direction q+5 at i = l+1 from i = 1.  Each value
// lands in a line of its own, so this part is bound by line transactions.
// WALLS: the channel's no-slip walls at z = 0, n+1 follow the faces and
// overwrite population d at (i - c_x, j - c_y) of a wall plane for every
// interior column (i, j) (walls_z_cell); periodic z otherwise.
template <typename R, bool WALLS>
__device__ __forceinline__ void x_face_part(Lattice19<R>& a, const Geometry& g, long t)
{
    const long rows = (long)(g.m + 2) * (g.n + 2), padded = x_face_rows(g);
    const int q = __builtin_amdgcn_readfirstlane((int)(t / padded));
    const long row = t - (long)q * padded;
    if (q >= 10 || row >= rows) return;
    const int nj = g.m + 2;
    const int j = (int)(row % nj), k = (int)(row / nj);
    const int d = q < 5 ? kEnterXlo[q] : kEnterXhi[q - 5];
    const int sj = halo_side(j, g.m), sk = halo_side(k, g.n);
    if (enters(sj, cy_of(d))) return;  // the y face writes it later
    if (enters(sk, cz_of(d))) {
        if (!WALLS) return;                                   // the z face writes it later
        if (j + cy_of(d) >= 1 && j + cy_of(d) <= g.m) return;  // a wall writes it later
    }
    const int it = q < 5 ? 0 : g.l + 1, is = q < 5 ? g.l : 1;
    a.p[d][gidx(g, it, j, k)] = a.p[d][gidx(g, is, j, k)];
}

// y face at (i, k), i = 0..l+1, k = 0..n+1: directions entering at j = 0
// (from j = m) and at j = m+1 (from j = 1), the source wrapped in x where
// the cell is an x halo cell the direction enters and the x face is periodic
// (XP); without it (bcond_inflow) the x halo already holds its final values
// and is copied as it is.
template <typename R, bool WALLS, bool XP = true>
__device__ __forceinline__ void y_face_part(Lattice19<R>& a, const Geometry& g, int i, int k)
{
    const int si = halo_side(i, g.l), sk = halo_side(k, g.n);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int jt = h == 0 ? 0 : g.m + 1, js = h == 0 ? g.m : 1;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const int d = h == 0 ? kEnterYlo[q] : kEnterYhi[q];
            if (enters(sk, cz_of(d))) {
                if (!WALLS) continue;                                 // the z face writes it later
                if (i + cx_of(d) >= 1 && i + cx_of(d) <= g.l) continue;  // a wall writes it later
            }
            const int is = XP && enters(si, cx_of(d)) ? wrap_to(si, g.l) : i;
            a.p[d][gidx(g, i, jt, k)] = a.p[d][gidx(g, is, js, k)];
        }
    }
}

// z face at (i, j) over the full face (periodic only), sources wrapped in x
// (if XP) and y as above.
template <typename R, bool XP = true>
__device__ __forceinline__ void z_face_part(Lattice19<R>& a, const Geometry& g, int i, int j)
{
    const int si = halo_side(i, g.l), sj = halo_side(j, g.m);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int kt = h == 0 ? 0 : g.n + 1, ks = h == 0 ? g.n : 1;
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const int d = h == 0 ? kEnterZlo[q] : kEnterZhi[q];
            const int is = XP && enters(si, cx_of(d)) ? wrap_to(si, g.l) : i;
            const int js = enters(sj, cy_of(d)) ? wrap_to(sj, g.m) : j;
            a.p[d][gidx(g, i, j, kt)] = a.p[d][gidx(g, is, js, ks)];
        }
    }
}

// bcond_periodic in one launch: threads [0, 10*x_face_rows) the x face,
// the next (l+2)(n+2) the y face, the last (l+2)(m+2) the z face.  X = false
// leaves the x face out and copies the x halo as it stands (bcond_inflow,
// whose x faces are computed, not copied, and filled by inflow_x before this
// launch; the y and z faces of SRC/bcond.inflow.F90 then copy them along).
template <typename R, bool X>
__global__ void periodic_faces(Lattice19<R> a, Geometry g)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (X) {
        const long nxf = 10 * x_face_rows(g);
        if (t < nxf) {
            x_face_part<R, false>(a, g, t);
            return;
        }
        t -= nxf;
    }
    const int ni = g.l + 2;
    const long nyf = (long)ni * (g.n + 2), nzf = (long)ni * (g.m + 2);
    if (t < nyf) {
        y_face_part<R, false, X>(a, g, (int)(t % ni), (int)(t / ni));
    } else if (t < nyf + nzf) {
        const long u = t - nyf;
        z_face_part<R, X>(a, g, (int)(u % ni), (int)(u / ni));
    }
}
template <typename R, bool X>
inline long periodic_faces_threads(const Geometry& g)
{
    return (X ? 10 * x_face_rows(g) : 0) + (long)(g.l + 2) * (g.n + 2) + (long)(g.l + 2) * (g.m + 2);
}

// bcond_channel in one launch (SRC/bcond.channel.F90: periodic x and y
// faces, then the no-slip walls at z = 0, n+1 over the interior columns,
// with the offload bounds 1..l x 1..m): the x face, the y face and the
// walls (walls_z_cell, lid off), the faces skipping what a later face or a
// wall overwrites.  The walls read only interior cells, and no face reads a
// cell a wall writes (a face's source lies on the k of its target, and a
// wall writes a population at a wall-plane cell only where that population
// enters across the wall, which the faces' sources never are).
template <typename R>
__global__ void channel_faces(Lattice19<R> a, Geometry g)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nxf = 10 * x_face_rows(g);
    if (t < nxf) {
        x_face_part<R, true>(a, g, t);
        return;
    }
    t -= nxf;
    const int ni = g.l + 2;
    const long nyf = (long)ni * (g.n + 2), nzf = (long)g.l * g.m;
    if (t < nyf) {
        y_face_part<R, true>(a, g, (int)(t % ni), (int)(t / ni));
    } else if (t < nyf + nzf) {
        const long u = t - nyf;
        walls_z_cell(a, g, 1 + (int)(u % g.l), 1 + (int)(u / g.l), R(0), 0);
    }
}
inline long channel_faces_threads(const Geometry& g)
{
    return 10 * x_face_rows(g) + (long)(g.l + 2) * (g.n + 2) + (long)g.l * g.m;
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
// opposite population).
//
// The reference's CPU build runs this as one loop over the box, k slowest
// and i fastest, in place.  When S + c_d = N is itself a solid cell of the
// box that the loop visited before S, N's population opp(d) has already been
// replaced by a_d(N + c_opp(d)) = a_d(S), so the loop leaves a_d(S) as it
// was; in every other case it stores the old a_opp(d)(N).  Each thread
// applies exactly that rule, which gives the serial loop's result in one
// race-free pass: a thread reads a_opp(d)(N) only when N is not an earlier
// solid cell, and N's thread then skips writing that population (S is
// earlier for N), while a_d(S) is written only where N's thread does not
// read it (N is fluid, outside the box, or later than S).  Solid cells
// therefore end with the reference's values too, and so do the means over
// all cells (SRC/diagno.F90:88-130).
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
    // a solid cell the loop over the box visits (the box lies in 1..l etc.)
    auto visited_solid = [&](int ii, int jj, int kk) -> bool {
        if (ii < imin || ii > imax || jj < jmin || jj > jmax || kk < kmin || kk > kmax) return false;
        return obs[(long)(ii - 1) + (long)g.l * (jj - 1) + (long)g.l * g.m * (kk - 1)] == 1;
    };
    if (!visited_solid(i, j, k)) return;
    const long c = gidx(g, i, j, k);
#pragma unroll
    for (int d = 0; d < 18; ++d) {
        // reference: a_d(i,j,k) = a_opp(d)(i + c_d)
        const int dx = cx_of(d), dy = cy_of(d), dz = cz_of(d);
        // the neighbour precedes (i,j,k) in the loop order (k, then j, then i)
        const bool earlier = dz < 0 || (dz == 0 && (dy < 0 || (dy == 0 && dx < 0)));
        if (earlier && visited_solid(i + dx, j + dy, k + dz)) continue;
        a.p[d][c] = a.p[opp_of(d)][gidx(g, i + dx, j + dy, k + dz)];
    }
}

#undef BGK_A

}  // namespace bgk
