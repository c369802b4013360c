// Reference-compatible C-ABI (v1) for the D3Q19 BGK step on gfx950.
//
// Built twice: real_t = float (libbgk3d_gpu_sp.so) and, with -DBGK_DOUBLE,
// real_t = double (libbgk3d_gpu_dp.so), matching the reference's SINGLE and
// DOUBLE_P builds (SRC/col_MC_gpu.hpp:78-82, SRC/bcond_gpu.hpp:47-51).  The
// symbols and argument lists are those the reference's Fortran binds
// (SRC/col_MC.F90:101-123, SRC/bcond_gpu_mod.F90:30-77), so either library
// links in place of the reference's col_MC_hip.o + bcond_hip.o.
//
// Ownership: all lattice memory belongs to the caller; these entry points
// never allocate it, never free it and never keep its pointers across calls
// (the caller swaps a and b every step).  The diagnostics keep a small cached
// device scratch (partial sums) and pinned host buffer.
#include "collide.hpp"
#include "common.hpp"
#include "d3q19.hpp"
#include "reduce.hpp"
#include "v1_kernels.hpp"
#include "host_init.hpp"
#include "init_kernels.hpp"

#include "../../include/bgk3d_gpu.h"

#include <algorithm>
#include <cstdlib>
#include <cmath>

#ifdef BGK_DOUBLE
using real_t = double;
#else
using real_t = float;
#endif

namespace bgk {
hipStream_t& abi_stream()
{
    static hipStream_t s = nullptr;  // null stream, as in the reference
    return s;
}
}  // namespace bgk

using namespace bgk;

// Device scratch of diagno_gpu_noalloc: the partial sums of at most
// kRedBlocks blocks and the folded totals, 5 doubles each (41 KB).  Each GPU
// holds its own copy of this array, at its own address, from the library's
// load until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (noalloc_scratch_ptr),
// never through the array's name.
constexpr size_t kNoAllocDoubles = (size_t)(kRedBlocks + 1) * 5;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

Lattice19<real_t> lat_from(real_t** a)
{
    Lattice19<real_t> L;
    for (int d = 0; d < 19; ++d) L.p[d] = a[d];
    return L;
}

Geometry geom(int l, int m, int n, int nx, int ny)
{
    Geometry g;
    g.l = l;
    g.m = m;
    g.n = n;
    g.nx = nx;
    g.nxy = (long)nx * (long)ny;
    return g;
}

constexpr int kFaceThreads = 256;

// Host-side check of the caller's arguments before any kernel is launched on
// them: positive sizes, rows and planes that hold the halo (nx >= l+2,
// ny >= m+2), and non-null population pointers (the first `na` of `a`, the
// first `nb` of `b`).  The reference performs no such check and faults on
// the device; here the call is refused with a message on stderr (the
// reference's way of reporting launch problems, SRC/col_MC_gpu.hpp:40-47).
bool shape_ok(const char* who, int l, int m, int n, int nx, int ny, const void* const* a, int na,
              const void* const* b = nullptr, int nb = 0)
{
    const char* why = nullptr;
    if (l <= 0 || m <= 0 || n <= 0) why = "sizes must be positive";
    else if (nx < l + 2 || ny < m + 2) why = "nx, ny must hold the halo (nx >= l+2, ny >= m+2)";
    else if ((na > 0 && !a) || (nb > 0 && !b)) why = "null pointer table";
    for (int d = 0; !why && d < na; ++d)
        if (!a[d]) why = "null population pointer";
    for (int d = 0; !why && d < nb; ++d)
        if (!b[d]) why = "null population pointer";
    if (why) {
        std::fprintf(stderr, "bgk3d_gpu: %s(l=%d m=%d n=%d nx=%d ny=%d): %s; nothing launched\n", who, l, m, n, nx,
                     ny, why);
        std::fflush(stderr);
        return false;
    }
    return true;
}
bool shape_ok(const char* who, int l, int m, int n, int nx, int ny, real_t* const* a)
{
    return shape_ok(who, l, m, n, nx, ny, reinterpret_cast<const void* const*>(a), 19);
}

template <typename K, typename... Args>
void launch_face(K kern, long count, const char* name, Args... args)
{
    if (count <= 0) return;
    hipLaunchKernelGGL(kern, dim3(ceil_div(count, kFaceThreads)), dim3(kFaceThreads), 0, abi_stream(), args...);
    report_launch(name);
}

// collide_flat uses no LDS; reserving this much per 256-thread block holds
// it to 5 blocks (5 waves per SIMD) of a CU's 160 KiB.  Over the caller's 37
// separately strided arrays fewer streams in flight ran faster: 496-498
// against 510-511 us per FP32 launch at 256^3 (8 waves per SIMD); 6 waves
// per SIMD gained half of that, 4 the same as 5.
constexpr unsigned kFlatLdsCap = 32u << 10;

// BGK3D_WHOLE_LINES=1 in the environment lets col_mc_gpu store zeros into
// the halo and padding cells between the interior rows of the arrays it
// writes (b01..b18 and a19), so that its flat mapping writes whole 128-byte
// lines (collide_flat, WL).  The reference's routines never read those values:
// each bcond_*_gpu call rewrites every halo value the next collide pulls, and
// the diagnostics read interior cells.  Only a raw dump of whole arrays
// (halo included) sees the difference.  Unset: the caller's halo and padding
// are left untouched, as the reference's col_MC leaves them.
bool whole_lines()
{
    const char* v = std::getenv("BGK3D_WHOLE_LINES");
    return v && v[0] == '1' && v[1] == '\0';
}

// The fused kernel in halo mode: every pull reads the caller's halo, which
// the bcond_*_gpu entry points filled.  Waves are aligned to 128-byte lines
// from the output addresses, whatever padding (ipad) the caller chose.
template <int FM>
void launch_col(const Lattice19<real_t>& A, const Lattice19<real_t>& B, const Geometry& g, real_t omega,
                real_t fgrad, real_t ux, real_t uy, real_t uz, const CollideConsts<real_t>& c,
                real_t cte_s = real_t(0))
{
    if (g.l <= 0 || g.m <= 0 || g.n <= 0) return;
    FusedArgs<real_t> p{};
    for (int d = 0; d < 19; ++d) p.a[d] = A.p[d];
    for (int d = 0; d < 18; ++d) p.b[d] = B.p[d];
    p.l = g.l;
    p.m = g.m;
    p.n = g.n;
    p.nx = g.nx;
    p.nxy = g.nxy;
    p.omega = omega;
    p.fgrad = fgrad;
    p.fx = ux;
    p.fy = uy;
    p.fz = uz;
    p.lidforce = 0;
    p.cte_s = cte_s;
    p.cst = c;
    p.obits = nullptr;
    p.align = (int)(128 / sizeof(real_t));
    p.lend = 0;  // the caller's halo and padding are left untouched
    p.seed = 0;
    // Row by row: ceil(span / 64) waves per row, span = l plus the rows'
    // alignment shift (exact when every row shares one alignment, else the
    // largest).  Flat: the plane's interior rows as one line-aligned range,
    // (nx*(m-1) + l + shift) / 64 waves per plane.  The mapping with fewer
    // waves is launched: flat when a row's last wave would be mostly idle
    // (l = 256, nx = 258: 4.03 waves per row instead of 5), row by row when
    // rows fill their waves (l = 62 from a line start) or carry wide padding.
    int span = g.l + p.align - 1;
    if (g.nx % p.align == 0 && g.nxy % p.align == 0) {
        const long first = (long)(reinterpret_cast<uintptr_t>(p.b[0]) / sizeof(real_t)) + 1;
        span = g.l + (int)(first & (p.align - 1));
    }
    const RowMap rm = row_map(span);
    const long row_waves = (long)g.m * rm.wpr;
    const long flat_cells = (long)g.nx * (g.m - 1) + g.l + p.align - 1;
    if (g.nxy < (1L << 31) && ceil_div(flat_cells, 64L) < row_waves) {
        const dim3 grid((unsigned)ceil_div(flat_cells, 256L), (unsigned)g.n);
        const bool nt0 = g.nx % p.align == 0 && g.nxy % p.align == 0, wl = whole_lines();
        if (nt0 && wl)
            hipLaunchKernelGGL((collide_flat<real_t, FM, true, true>), grid, dim3(256), kFlatLdsCap, abi_stream(), p);
        else if (nt0)
            hipLaunchKernelGGL((collide_flat<real_t, FM, true>), grid, dim3(256), kFlatLdsCap, abi_stream(), p);
        else if (wl)
            hipLaunchKernelGGL((collide_flat<real_t, FM, false, true>), grid, dim3(256), kFlatLdsCap, abi_stream(), p);
        else
            hipLaunchKernelGGL((collide_flat<real_t, FM, false>), grid, dim3(256), kFlatLdsCap, abi_stream(), p);
        report_launch("collide_flat");
        return;
    }
    hipLaunchKernelGGL((collide_fused<real_t, AX_HALO, AX_HALO, AX_HALO, false, FM, 0>), fused_grid(rm, g.m, g.n),
                       fused_block(rm), 0, abi_stream(), p, rm);
    report_launch("collide_fused");
}

// Cached scratch for the reductions (per device; single host thread, as in
// the reference's driver).
struct Scratch {
    double* dev = nullptr;
    double* host = nullptr;
    size_t cap = 0;
};

Scratch& scratch_for_device()
{
    static Scratch s[64];
    int dev = 0;
    (void)hipGetDevice(&dev);
    return s[dev & 63];
}

double* cached_scratch(size_t doubles, double** host)
{
    Scratch& s = scratch_for_device();
    if (s.cap < doubles) {
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
        s.dev = nullptr;
        s.host = nullptr;
        s.cap = 0;
        if (hipMalloc(&s.dev, doubles * sizeof(double)) != hipSuccess) return nullptr;
        if (hipHostMalloc(&s.host, doubles * sizeof(double), hipHostMallocDefault) != hipSuccess) return nullptr;
        s.cap = doubles;
    }
    *host = s.host;
    return s.dev;
}

// The device address of `scratch` on the current device, looked up once per
// device (first call there, with that device current) and kept; null if the
// lookup fails.  Single host thread, as the reference's driver.
double* noalloc_scratch_ptr()
{
    static double* ptr[64] = {};
    static bool looked_up[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    if (!looked_up[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            std::fprintf(stderr, "bgk3d_gpu: diagno_gpu_noalloc: hipGetSymbolAddress failed: %s\n",
                         hipGetErrorString(e));
            return nullptr;
        }
        ptr[dev] = static_cast<double*>(p);
        looked_up[dev] = true;
    }
    return ptr[dev];
}

int fetch(const double* dev, double* host, size_t n)
{
    BGK_HIP_CHECK(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, abi_stream()));
    BGK_HIP_CHECK(hipStreamSynchronize(abi_stream()));
    return 0;
}

}  // namespace

extern "C" {

void col_mc_gpu(const real_t* a01, const real_t* a02, const real_t* a03, const real_t* a04, const real_t* a05,
                const real_t* a06, const real_t* a07, const real_t* a08, const real_t* a09, const real_t* a10,
                const real_t* a11, const real_t* a12, const real_t* a13, const real_t* a14, const real_t* a15,
                const real_t* a16, const real_t* a17, const real_t* a18, const real_t* a19, real_t* b01,
                real_t* b02, real_t* b03, real_t* b04, real_t* b05, real_t* b06, real_t* b07, real_t* b08,
                real_t* b09, real_t* b10, real_t* b11, real_t* b12, real_t* b13, real_t* b14, real_t* b15,
                real_t* b16, real_t* b17, real_t* b18, int l, int m, int n, int nx, int ny, real_t omega,
                real_t cte0, real_t cte1, real_t p0, real_t p1, real_t p2, real_t rf, real_t qf, real_t tre,
                real_t forcex, real_t forcey, real_t forcez)
{
    const real_t* in[19] = {a01, a02, a03, a04, a05, a06, a07, a08, a09, a10,
                            a11, a12, a13, a14, a15, a16, a17, a18, a19};
    real_t* out[18] = {b01, b02, b03, b04, b05, b06, b07, b08, b09, b10, b11, b12, b13, b14, b15, b16, b17, b18};
    if (!shape_ok("col_mc_gpu", l, m, n, nx, ny, reinterpret_cast<const void* const*>(in), 19,
                  reinterpret_cast<const void* const*>(out), 18))
        return;
    Lattice19<real_t> A, B;
    for (int d = 0; d < 19; ++d) A.p[d] = const_cast<real_t*>(in[d]);
    for (int d = 0; d < 18; ++d) B.p[d] = out[d];
    B.p[18] = A.p[18];  // direction 19 collides in place
    CollideConsts<real_t> c{cte0, cte1, p0, p1, p2, rf, qf, tre};
    launch_col<0>(A, B, geom(l, m, n, nx, ny), omega, real_t(0), forcex, forcey, forcez, c);
}

void col_mc_force_gpu(real_t** a, real_t** b, int l, int m, int n, int nx, int ny, real_t omega, real_t fgrad,
                      int axis)
{
    if (!shape_ok("col_mc_force_gpu", l, m, n, nx, ny, reinterpret_cast<const void* const*>(a), 19,
                  reinterpret_cast<const void* const*>(b), 18))
        return;
    Lattice19<real_t> A = lat_from(a), B = lat_from(b);
    B.p[18] = A.p[18];
    const Geometry g = geom(l, m, n, nx, ny);
    const auto c = CollideConsts<real_t>::shifted();
    switch (axis) {
        case 1: launch_col<1>(A, B, g, omega, fgrad, 0, 0, 0, c); break;
        case 2: launch_col<2>(A, B, g, omega, fgrad, 0, 0, 0, c); break;
        case 3: launch_col<3>(A, B, g, omega, fgrad, 0, 0, 0, c); break;
        default: launch_col<0>(A, B, g, omega, 0, 0, 0, 0, c); break;
    }
}

void col_mc_les_gpu(real_t** a, real_t** b, int l, int m, int n, int nx, int ny, real_t omega1, real_t cte_s,
                    real_t fgrad, int axis)
{
    if (!shape_ok("col_mc_les_gpu", l, m, n, nx, ny, reinterpret_cast<const void* const*>(a), 19,
                  reinterpret_cast<const void* const*>(b), 18))
        return;
    Lattice19<real_t> A = lat_from(a), B = lat_from(b);
    B.p[18] = A.p[18];
    const Geometry g = geom(l, m, n, nx, ny);
    const auto c = CollideConsts<real_t>::shifted();
    switch (axis) {
        case 1: launch_col<5>(A, B, g, omega1, fgrad, 0, 0, 0, c, cte_s); break;
        case 2: launch_col<6>(A, B, g, omega1, fgrad, 0, 0, 0, c, cte_s); break;
        case 3: launch_col<7>(A, B, g, omega1, fgrad, 0, 0, 0, c, cte_s); break;
        default: launch_col<4>(A, B, g, omega1, 0, 0, 0, 0, c, cte_s); break;
    }
}

void bcond_periodic_gpu(real_t** a, int l, int m, int n, int nx, int ny)
{
    if (!shape_ok("bcond_periodic_gpu", l, m, n, nx, ny, a)) return;
    const auto A = lat_from(a);
    const auto g = geom(l, m, n, nx, ny);
    launch_face(periodic_faces<real_t, true>, periodic_faces_threads<real_t, true>(g), "periodic_faces", A, g);
}

void bcond_channel_gpu(real_t** a, int l, int m, int n, int nx, int ny)
{
    if (!shape_ok("bcond_channel_gpu", l, m, n, nx, ny, a)) return;
    const auto A = lat_from(a);
    const auto g = geom(l, m, n, nx, ny);
    launch_face(channel_faces<real_t>, channel_faces_threads(g), "channel_faces", A, g);
}

void bcond_driven_gpu(real_t** a, int l, int m, int n, int nx, int ny, real_t force)
{
    if (!shape_ok("bcond_driven_gpu", l, m, n, nx, ny, a)) return;
    const auto A = lat_from(a);
    const auto g = geom(l, m, n, nx, ny);
    launch_face(driven_faces<real_t>, (long)m * n + (long)l * n + (long)l * m, "driven_faces", A, g, force);
}

void bcond_inflow_gpu(real_t** a, int l, int m, int n, int nx, int ny, real_t u_inflow, real_t rf, real_t qf,
                      real_t tre, real_t p1, real_t p2, real_t cte1)
{
    (void)tre;
    if (!shape_ok("bcond_inflow_gpu", l, m, n, nx, ny, a)) return;
    const auto A = lat_from(a);
    const auto g = geom(l, m, n, nx, ny);
    launch_face(inflow_x<real_t>, (long)(m + 2) * (n + 2), "inflow_x", A, g, u_inflow, rf, qf, p1, p2, cte1);
    launch_face(periodic_faces<real_t, false>, periodic_faces_threads<real_t, false>(g), "periodic_yz", A, g);
}

void bcond_obs_gpu(real_t** a, const int* obs, int l, int m, int n, int nx, int ny, int imin, int imax, int jmin,
                   int jmax, int kmin, int kmax, int obs_nx, int obs_ny)
{
    if (!shape_ok("bcond_obs_gpu", l, m, n, nx, ny, a)) return;
    if (!obs || obs_nx != l || obs_ny != m) {
        std::fprintf(stderr, "bgk3d_gpu: bcond_obs_gpu expects the (1:l,1:m,1:n) mask (got %d x %d)\n", obs_nx,
                     obs_ny);
        return;
    }
    if (imax < imin || jmax < jmin || kmax < kmin) return;  // no solid cell
    if (imin < 1 || imax > l || jmin < 1 || jmax > m || kmin < 1 || kmax > n) {
        std::fprintf(stderr, "bgk3d_gpu: bcond_obs_gpu box [%d:%d,%d:%d,%d:%d] leaves (1:l,1:m,1:n)\n", imin,
                     imax, jmin, jmax, kmin, kmax);
        return;
    }
    const auto A = lat_from(a);
    const auto g = geom(l, m, n, nx, ny);
    launch_face(obstacle_bb<real_t>, (long)(imax - imin + 1) * (jmax - jmin + 1) * (kmax - kmin + 1),
                "obstacle_bb", A, g, obs, imin, imax, jmin, jmax, kmin, kmax);
}

void gpu_device_sync(void)
{
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        std::fprintf(stderr, "bgk3d_gpu: device sync failed: %s\n", hipGetErrorString(e));
        std::fflush(stderr);
        std::abort();
    }
}

int bgk_set_stream(void* stream)
{
    abi_stream() = reinterpret_cast<hipStream_t>(stream);
    return 0;
}

int diagno_gpu(real_t** a, int l, int m, int n, int nx, int ny, real_t cte1, double* out)
{
    if (!shape_ok("diagno_gpu", l, m, n, nx, ny, a) || !out) return -1;
    double* host = nullptr;
    double* dev = cached_scratch((size_t)(kRedBlocks + 1) * 5, &host);
    if (!dev) return -2;
    const long rows = (long)m * n;
    const int nb = (int)(rows < kRedBlocks ? rows : kRedBlocks);
    hipLaunchKernelGGL(diagno_partial<real_t>, dim3(nb), dim3(kRedThreads), 0, abi_stream(), lat_from(a),
                       geom(l, m, n, nx, ny), cte1, dev + 5);
    report_launch("diagno_partial");
    hipLaunchKernelGGL(fold_partials<5>, dim3(1), dim3(kRedThreads), 0, abi_stream(), dev + 5, nb, dev);
    report_launch("fold_partials");
    if (int rc = fetch(dev, host, 5)) return rc;
    // (x/float(l))/float(m)/float(n), SRC/diagno.F90:134-138
    const double fl = (double)(float)l, fm = (double)(float)m, fn = (double)(float)n;
    for (int q = 0; q < 5; ++q) out[q] = (host[q] / fl) / fm / fn;
    return 0;
}

int diagno_gpu_noalloc(real_t** a, int l, int m, int n, int nx, int ny, real_t cte1, double* out)
{
    if (!shape_ok("diagno_gpu_noalloc", l, m, n, nx, ny, a) || !out) return -1;
    const long rows = (long)m * n;
    const int nb = (int)(rows < kRedBlocks ? rows : kRedBlocks);
    // the partials of nb blocks after the 5 totals
    if ((size_t)(nb + 1) * 5 > kNoAllocDoubles) {
        std::fprintf(stderr, "bgk3d_gpu: diagno_gpu_noalloc: %d partials exceed the static scratch of %zu doubles\n",
                     nb, kNoAllocDoubles);
        return -1;
    }
    double* dev = noalloc_scratch_ptr();
    if (!dev) return -2;
    // every partial and total is written before it is read: nothing to clear
    hipLaunchKernelGGL(diagno_partial<real_t>, dim3(nb), dim3(kRedThreads), 0, abi_stream(), lat_from(a),
                       geom(l, m, n, nx, ny), cte1, dev + 5);
    report_launch("diagno_partial");
    hipLaunchKernelGGL(fold_partials<5>, dim3(1), dim3(kRedThreads), 0, abi_stream(), dev + 5, nb, dev);
    report_launch("fold_partials");
    double host[5];
    BGK_HIP_CHECK(hipMemcpyAsync(host, dev, sizeof host, hipMemcpyDeviceToHost, abi_stream()));
    BGK_HIP_CHECK(hipStreamSynchronize(abi_stream()));
    const double fl = (double)(float)l, fm = (double)(float)m, fn = (double)(float)n;
    for (int q = 0; q < 5; ++q) out[q] = (host[q] / fl) / fm / fn;
    return 0;
}

int varm_gpu(real_t** a, int l, int m, int n, int nx, int ny, real_t cte1, double* u_k, double* w_k, double* v_k,
             double* den_j)
{
    if (!shape_ok("varm_gpu", l, m, n, nx, ny, a) || !u_k || !w_k || !v_k || !den_j) return -1;
    double* host = nullptr;
    double* dev = cached_scratch((size_t)3 * n + m, &host);
    if (!dev) return -2;
    const auto A = lat_from(a);
    const auto g = geom(l, m, n, nx, ny);
    hipLaunchKernelGGL(varm_planes_k<real_t>, dim3(n), dim3(kRedThreads), 0, abi_stream(), A, g, cte1, dev);
    report_launch("varm_planes_k");
    hipLaunchKernelGGL(varm_planes_j<real_t>, dim3(m), dim3(kRedThreads), 0, abi_stream(), A, g, cte1,
                       dev + 3 * n);
    report_launch("varm_planes_j");
    if (int rc = fetch(dev, host, (size_t)3 * n + m)) return rc;
    // rvol = 1.0/(float(l)) in default real (SRC/varm.F90:94)
    const double rvol = (double)(real_t)(1.0f / (float)l);
    for (int k = 0; k < n; ++k) {
        u_k[k] = host[3 * k + 0] * rvol;
        w_k[k] = host[3 * k + 1] * rvol;
        v_k[k] = host[3 * k + 2] * rvol;
    }
    for (int j = 0; j < m; ++j) den_j[j] = host[3 * n + j] * rvol;
    return 0;
}

int draglift_gpu(real_t** a, const int* obs, int l, int m, int n, int nx, int ny, int icoord, int jcoord,
                 int kcoord, int radius, real_t u_inflow, double* out)
{
    if (!shape_ok("draglift_gpu", l, m, n, nx, ny, a) || !obs || !out) return -1;
    // bounding box of SRC/draglift.F90:54-62 (delta = 3)
    const int delta = 3;
    const int is = icoord - radius - delta, ie = icoord + radius + delta;
    const int js = jcoord - radius - delta, je = jcoord + radius + delta;
    const int ks = kcoord - radius - delta, ke = kcoord + radius + delta;
    // every obs(x + c) read must stay inside (1:l,1:m,1:n)
    if (is < 2 || js < 2 || ks < 2 || ie > l - 1 || je > m - 1 || ke > n - 1) {
        std::fprintf(stderr, "bgk3d_gpu: draglift box [%d:%d,%d:%d,%d:%d] leaves the mask\n", is, ie, js, je, ks, ke);
        return -1;
    }
    double* host = nullptr;
    double* dev = cached_scratch((size_t)(kRedBlocks + 1) * 3, &host);
    if (!dev) return -2;
    const long cells = (long)(ie - is + 1) * (je - js + 1) * (ke - ks + 1);
    const int nb = (int)((cells + kRedThreads - 1) / kRedThreads < kRedBlocks ? (cells + kRedThreads - 1) / kRedThreads
                                                                                : kRedBlocks);
    hipLaunchKernelGGL(draglift_partial<real_t>, dim3(nb), dim3(kRedThreads), 0, abi_stream(), lat_from(a),
                       geom(l, m, n, nx, ny), obs, is, ie, js, je, ks, ke, dev + 3);
    report_launch("draglift_partial");
    hipLaunchKernelGGL(fold_partials<3>, dim3(1), dim3(kRedThreads), 0, abi_stream(), dev + 3, nb, dev);
    report_launch("fold_partials");
    if (int rc = fetch(dev, host, 3)) return rc;
    // norm = uno/(u_inflow*u_inflow*radius) in the storage kind (:178)
    const real_t norm = real_t(1) / (u_inflow * u_inflow * (real_t)radius);
    for (int q = 0; q < 3; ++q) out[q] = host[q] * (double)norm;
    return 0;
}

// Line/point profile shared by prof_i/j/k_gpu and probe_gpu (axis 0: i at
// (c1, c2) = (j, k); 1: j at (i, k); 2: k at (i, j)); 4 values per point.
static int line_profile(real_t** a, int l, int m, int n, int nx, int ny, int axis, int c1, int c2, int first,
                        int count, real_t cte1, real_t* out)
{
    const int len = axis == 0 ? l : (axis == 1 ? m : n);
    const int lim1 = axis == 0 ? m : l, lim2 = axis == 2 ? m : n;
    if (!shape_ok("line profile", l, m, n, nx, ny, a) || !out) return -1;
    if (count <= 0 || first < 1 || first + count - 1 > len) return -1;
    if (c1 < 0 || c1 > lim1 + 1 || c2 < 0 || c2 > lim2 + 1) return -1;
    double* host = nullptr;
    const size_t bytes = (size_t)4 * count * sizeof(real_t);
    double* dev = cached_scratch((bytes + sizeof(double) - 1) / sizeof(double), &host);
    if (!dev) return -2;
    real_t* d = reinterpret_cast<real_t*>(dev);
    hipLaunchKernelGGL(prof_line_kernel<real_t>, dim3(ceil_div(count, 64)), dim3(64), 0, abi_stream(), lat_from(a),
                       geom(l, m, n, nx, ny), axis, c1, c2, first, count, cte1, d);
    report_launch("prof_line_kernel");
    BGK_HIP_CHECK(hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, abi_stream()));
    BGK_HIP_CHECK(hipStreamSynchronize(abi_stream()));
    return 0;
}

int prof_i_gpu(real_t** a, int l, int m, int n, int nx, int ny, int jcoord, int kcoord, real_t cte1, real_t* out)
{
    return line_profile(a, l, m, n, nx, ny, 0, jcoord, kcoord, 1, l, cte1, out);
}

int prof_j_gpu(real_t** a, int l, int m, int n, int nx, int ny, int icoord, int kcoord, real_t cte1, real_t* out)
{
    return line_profile(a, l, m, n, nx, ny, 1, icoord, kcoord, 1, m, cte1, out);
}

int prof_k_gpu(real_t** a, int l, int m, int n, int nx, int ny, int icoord, int jcoord, real_t cte1, real_t* out)
{
    return line_profile(a, l, m, n, nx, ny, 2, icoord, jcoord, 1, n, cte1, out);
}

int probe_gpu(real_t** a, int l, int m, int n, int nx, int ny, int i0, int j0, int k0, real_t cte1, real_t* out)
{
    return line_profile(a, l, m, n, nx, ny, 0, j0, k0, i0, 1, cte1, out);
}

// Initial populations (host-side, as the reference's init), see host_init.hpp.
int bgk_init_host(real_t** h, int l, int m, int n, int nx, int ny, int mode, real_t u00)
{
    return bgk::init_feq_host<real_t>(h, l, m, n, nx, ny, mode, u00);
}

int bgk_init_gpu(real_t** a, int l, int m, int n, int nx, int ny, int mode, real_t u00)
{
    if (!shape_ok("bgk_init_gpu", l, m, n, nx, ny, a)) return -1;
    if (mode < 0 || (mode & ~kInitUnshifted) > 2 || mode == (2 | kInitUnshifted)) {
        std::fprintf(stderr,
                     "bgk3d_gpu: bgk_init_gpu: init mode %d not supported (0, 1, 2; 0 or 1 plus %d for "
                     "unshifted storage)\n",
                     mode, kInitUnshifted);
        return -1;
    }
    const hipError_t e = bgk::init_feq_device<real_t>(a, l, m, n, nx, (long)nx * ny, mode, u00, abi_stream());
    if (e != hipSuccess) {
        std::fprintf(stderr, "bgk3d_gpu: bgk_init_gpu: %s\n", hipGetErrorString(e));
        return -2;
    }
    return 0;
}

// ---- device memory plumbing for hosts that have no other HIP runtime ----
int bgk_dev_malloc(void** ptr, size_t bytes)
{
    BGK_HIP_CHECK(hipMalloc(ptr, bytes));
    return 0;
}
int bgk_dev_free(void* ptr)
{
    BGK_HIP_CHECK(hipFree(ptr));
    return 0;
}
int bgk_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    BGK_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}
int bgk_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    BGK_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}
int bgk_device_count(int* n)
{
    BGK_HIP_CHECK(hipGetDeviceCount(n));
    return 0;
}

// ---- measured stream bandwidth: copy of n floats (n*8 bytes moved) -------
typedef float v4f __attribute__((ext_vector_type(4)));
// non-temporal loads and stores, the same memory instructions as the fused
// collide, so its counters calibrate the collide's.  One 16-byte element per
// thread over the whole buffer: on gfx950 this runs about 30 % faster than a
// grid-stride loop over a capped grid.
__global__ void stream_copy_kernel(const v4f* __restrict__ src, v4f* __restrict__ dst, long n4)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}
int bgk_stream_copy(const void* src, void* dst, size_t n_floats)
{
    if (n_floats % 4) return -1;
    const long n4 = (long)(n_floats / 4);
    if ((n4 + 255) / 256 > 0x7fffffffL) return -1;
    const int blocks = (int)((n4 + 255) / 256);
    hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, abi_stream(),
                       reinterpret_cast<const v4f*>(src), reinterpret_cast<v4f*>(dst), n4);
    report_launch("stream_copy_kernel");
    return 0;
}

// ---- timing on the ABI stream (events bracket work queued there) --------
int bgk_event_create(void** ev)
{
    hipEvent_t e;
    BGK_HIP_CHECK(hipEventCreate(&e));
    *ev = e;
    return 0;
}
int bgk_event_record(void* ev)
{
    BGK_HIP_CHECK(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), abi_stream()));
    return 0;
}
int bgk_event_elapsed_ms(void* start, void* stop, float* ms)
{
    BGK_HIP_CHECK(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
    BGK_HIP_CHECK(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)));
    return 0;
}
int bgk_event_destroy(void* ev)
{
    BGK_HIP_CHECK(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)));
    return 0;
}

}  // extern "C"
