/*
 * CPU oracle for the D3Q19 BGK time step -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
 * load this library, and only as the checker.  The product library never
 * links or calls it.  See bgk_oracle_impl.h for what each function restates
 * and how it is pinned.
 *
 * Build: gcc -O2 -fPIC -shared -ffp-contract=off -fno-fast-math -fopenmp
 * (oracle/Makefile).  -ffp-contract=off keeps a*b+c as two roundings, as
 * the reference's x86-64 build does.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#define SINf32 sinf
#define COSf32 cosf
#define SQRTf32 sqrtf
#define SINf64 sin
#define COSf64 cos
#define SQRTf64 sqrt

/* Storage convention of the build: 0 shifted (f - w, the default), 1
 * unshifted (the reference's -DNOSHIFT builds).  The reference sets cte1
 * per routine: init and bcond_inflow use cte1 = 1 unshifted / 0 shifted
 * (SRC/init.F90:71-75, SRC/bcond.inflow.F90:68-72); col_MC, diagno, varm,
 * prof_i/j/k and probe use cte1 = 0 unshifted / 1 shifted, cte0 = 1 - cte1
 * (SRC/col_MC.F90:139-144, SRC/diagno.F90:55-59, SRC/varm.F90:44-48,
 * SRC/prof_i.F90:42-46, SRC/probe.F90:43-47).  Set with oracle_set_noshift. */
static int g_noshift = 0;
void oracle_set_noshift(int v) { g_noshift = v ? 1 : 0; }
int oracle_get_noshift(void) { return g_noshift; }

#define R float
#define SFX f32
#include "bgk_oracle_impl.h"
#undef R
#undef SFX

#define R double
#define SFX f64
#include "bgk_oracle_impl.h"
#undef R
#undef SFX

/* build_obs: SRC/build.obs.F90:34-97.  Hollow sphere shell
 * (R-2)^2 < d^2 < (R+2)^2 centred at (2l/5, m/2, n/2), R = 32 if m > 256
 * else ly/8 (ly = m).  d2 and the radii are kind-R values; every term is an
 * integer below 2^24 so float and double give the same mask.  obs is
 * (1:l,1:m,1:n), i fastest.  box = {imin, imax, jmin, jmax, kmin, kmax};
 * returns the number of solid cells; *radius receives R. */
long oracle_build_obs(int *obs, int l, int m, int n, int *box, int *radius)
{
    const int ic = 2 * l / 5, jc = m / 2, kc = n / 2;
    const int rad = (m > 256) ? 32 : m / 8;
    const double Rr = (double)rad;
    const double r2a = (Rr - 2) * (Rr - 2), r2b = (Rr + 2) * (Rr + 2);
    int imin = l, jmin = m, kmin = n, imax = 0, jmax = 0, kmax = 0;
    long nobs = 0;
    for (int k = 1; k <= n; ++k)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= l; ++i) {
                const double d2 = (double)((ic - i) * (ic - i) + (jc - j) * (jc - j) + (kc - k) * (kc - k));
                const size_t o = (size_t)(i - 1) + (size_t)l * (size_t)(j - 1) + (size_t)l * (size_t)m * (size_t)(k - 1);
                obs[o] = 0;
                if (d2 > r2a && d2 < r2b) {
                    obs[o] = 1;
                    ++nobs;
                    if (i < imin) imin = i;
                    if (j < jmin) jmin = j;
                    if (k < kmin) kmin = k;
                    if (i > imax) imax = i;
                    if (j > jmax) jmax = j;
                    if (k > kmax) kmax = k;
                }
            }
    box[0] = imin;
    box[1] = imax;
    box[2] = jmin;
    box[3] = jmax;
    box[4] = kmin;
    box[5] = kmax;
    *radius = rad;
    return nobs;
}

/*
 * Time loop of SRC/bgk3d.F90:169-196 for the FUSED build:
 *   boundaries (SRC/boundaries.F90:31-47) -> col_MC -> swap a<->b (1..18).
 * bc: 0 driven cavity, 1 periodic, 2 channel, 3 inflow(+obstacle if obs).
 * The swap is done on the pointer arrays the caller passes, so on return
 * a[] holds the current lattice, exactly as the Fortran pointers do.
 */
#define RUN_IMPL(R, SFX)                                                                               \
    void oracle_run_##SFX(R **a, R **b, int l, int m, int n, int nx, int ny, int bc, R omega, R u00,    \
                         int force_mode, R fgrad, R u_inflow, const int *obs, const int *box, int steps, \
                         R cteS, int nthreads)                                                         \
    {                                                                                                  \
        for (int t = 0; t < steps; ++t) {                                                              \
            if (bc == 0) oracle_bcond_driven_##SFX(a, l, m, n, nx, ny, u00);                           \
            else if (bc == 1) oracle_bcond_periodic_##SFX(a, l, m, n, nx, ny);                         \
            else if (bc == 2) oracle_bcond_channel_##SFX(a, l, m, n, nx, ny);                          \
            else oracle_bcond_inflow_##SFX(a, l, m, n, nx, ny, u_inflow);                              \
            if (obs)                                                                                   \
                oracle_bcond_obs_##SFX(a, obs, l, m, n, nx, ny, box[0], box[1], box[2], box[3], box[4],  \
                                       box[5]);                                                        \
            oracle_col_mc_##SFX(a, b, l, m, n, nx, ny, omega, force_mode, fgrad, (R)0, (R)0, (R)0, cteS, \
                                nthreads);                                                             \
            for (int d = 0; d < 18; ++d) {                                                             \
                R *tmp = a[d];                                                                         \
                a[d] = b[d];                                                                           \
                b[d] = tmp;                                                                            \
            }                                                                                          \
        }                                                                                              \
    }
RUN_IMPL(float, f32)
RUN_IMPL(double, f64)
