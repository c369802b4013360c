/*
 * bgk3d_gpu.h -- C-ABI of the gfx950 D3Q19 BGK time step (reference-compatible, v1).
 *
 * Two shared libraries export these symbols:
 *   libbgk3d_gpu_sp.so   real_t = float   (reference SINGLE build)
 *   libbgk3d_gpu_dp.so   real_t = double  (reference DOUBLE_P build; define
 *                                          BGK_DOUBLE before including)
 * Either replaces col_MC_hip.o + bcond_hip.o on the reference's `hip` link
 * line (SRC/Makefile:504-505).
 *
 * Conventions shared by every entry point (as in the reference):
 *  - lattice arrays are caller-owned device memory, Fortran column-major
 *    (0:l+1+ipad, 0:m+1+jpad, 0:n+1); nx = size(a,1), ny = size(a,2);
 *    element (i,j,k) at offset i + nx*j + nx*ny*k (SRC/alloca.F90:153-191);
 *  - `real_t** a` is a HOST array of the 19 device base pointers a01..a19
 *    (the Fortran pa(1:19), SRC/bcond_gpu_mod.F90:84-117);
 *  - work is queued on one in-order stream (the null stream unless
 *    bgk_set_stream() chose another); no entry point synchronises except
 *    gpu_device_sync() and the diagnostics, which return host values;
 *  - launch failures are printed to stderr and execution continues; an
 *    asynchronous fault aborts in gpu_device_sync() (SRC/bcond_gpu.hpp:548-571).
 */
#ifndef BGK3D_GPU_H
#define BGK3D_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef BGK_DOUBLE
typedef double bgk_real_t;
#else
typedef float bgk_real_t;
#endif

/* ---------------------------------------------------------------------------
 * Reference symbols (same names and argument lists).
 * ------------------------------------------------------------------------- */

/* Fused pull-stream + BGK collide of all interior cells: reads a01..a19,
 * writes b01..b18 and direction 19 in place into a19; the caller then swaps
 * a<->b (1..18).  cte0/cte1 select shifted (0/1) or NOSHIFT (1/0) storage;
 * forcex/y/z are uniform force terms.  Halo and padding cells are left
 * untouched, unless BGK3D_WHOLE_LINES=1 is set in the environment: then the
 * halo and padding cells between interior rows of the arrays written
 * receive zeros (values no routine reads: each bcond_*_gpu call rewrites
 * every halo value a pull uses), so that whole 128-byte lines are written
 * (about 8 % faster at 256^3).
 * Replaces: col_mc_gpu, SRC/col_MC_hip.cpp:16-56; bound at SRC/col_MC.F90:101-123,
 * called at SRC/col_MC.F90:217-229. */
void col_mc_gpu(const bgk_real_t* a01, const bgk_real_t* a02, const bgk_real_t* a03, const bgk_real_t* a04,
                const bgk_real_t* a05, const bgk_real_t* a06, const bgk_real_t* a07, const bgk_real_t* a08,
                const bgk_real_t* a09, const bgk_real_t* a10, const bgk_real_t* a11, const bgk_real_t* a12,
                const bgk_real_t* a13, const bgk_real_t* a14, const bgk_real_t* a15, const bgk_real_t* a16,
                const bgk_real_t* a17, const bgk_real_t* a18, const bgk_real_t* a19, bgk_real_t* b01,
                bgk_real_t* b02, bgk_real_t* b03, bgk_real_t* b04, bgk_real_t* b05, bgk_real_t* b06,
                bgk_real_t* b07, bgk_real_t* b08, bgk_real_t* b09, bgk_real_t* b10, bgk_real_t* b11,
                bgk_real_t* b12, bgk_real_t* b13, bgk_real_t* b14, bgk_real_t* b15, bgk_real_t* b16,
                bgk_real_t* b17, bgk_real_t* b18, int l, int m, int n, int nx, int ny, bgk_real_t omega,
                bgk_real_t cte0, bgk_real_t cte1, bgk_real_t p0, bgk_real_t p1, bgk_real_t p2, bgk_real_t rf,
                bgk_real_t qf, bgk_real_t tre, bgk_real_t forcex, bgk_real_t forcey, bgk_real_t forcez);

/* Periodic halo fill, faces x -> y -> z over full faces incl. halo rows.
 * Replaces: bcond_periodic_gpu, SRC/bcond_gpu.hpp:499-504; bound at
 * SRC/bcond_gpu_mod.F90:31-36; CPU semantics SRC/bcond.periodic.F90:67-266. */
void bcond_periodic_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny);

/* Channel: x, y periodic; no-slip bounce-back walls at z = 1 and z = n.
 * Replaces: bcond_channel_gpu, SRC/bcond_gpu.hpp:506-511; bound at
 * SRC/bcond_gpu_mod.F90:38-43; semantics SRC/bcond.channel.F90:78-281 with
 * the interior wall bounds of :244-251. */
void bcond_channel_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny);

/* Lid-driven cavity: half-way bounce-back on all faces, lid at z = n moving
 * +x with force = u00/6 (z face last, so its value wins on shared edges).
 * Replaces: bcond_driven_gpu, SRC/bcond_gpu.hpp:513-519; bound at
 * SRC/bcond_gpu_mod.F90:45-51; semantics SRC/bcond.driven.F90:61-262. */
void bcond_driven_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny, bgk_real_t force);

/* Inflow at x = 0 (equilibrium, rho = 1, u = (u_inflow,0,0)), zero-gradient
 * equilibrium outflow at x = l+1, y and z periodic.  cte1 = 0 selects shifted
 * storage (inverted convention of the boundary routines).
 * Replaces: bcond_inflow_gpu, SRC/bcond_gpu.hpp:521-531; bound at
 * SRC/bcond_gpu_mod.F90:53-60; semantics SRC/bcond.inflow.F90:91-328. */
void bcond_inflow_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny, bgk_real_t u_inflow, bgk_real_t rf,
                      bgk_real_t qf, bgk_real_t tre, bgk_real_t p1, bgk_real_t p2, bgk_real_t cte1);

/* Link-wise bounce-back on solid cells (obs == 1, int32 mask (1:l,1:m,1:n))
 * inside the box; obs_nx/obs_ny must equal l/m.  Only links whose neighbour
 * is fluid are written (deterministic; fluid results equal the reference's).
 * Replaces: bcond_obs_gpu, SRC/bcond_gpu.hpp:533-544; bound at
 * SRC/bcond_gpu_mod.F90:63-72; semantics SRC/bcond.obs.F90:101-122. */
void bcond_obs_gpu(bgk_real_t** a, const int* obs, int l, int m, int n, int nx, int ny, int imin, int imax, int jmin,
                   int jmax, int kmin, int kmax, int obs_nx, int obs_ny);

/* Full device barrier; aborts on an asynchronous fault.
 * Replaces: gpu_device_sync, SRC/bcond_gpu.hpp:548-571; bound at
 * SRC/bcond_gpu_mod.F90:75-76. */
void gpu_device_sync(void);

/* ---------------------------------------------------------------------------
 * New surface in the same style (the reference runs these on the host).
 * All return 0 on success, non-zero on error.
 * ------------------------------------------------------------------------- */

/* Fused step with the CHANNEL body force forcex|y|z = fgrad*rho per cell
 * (axis 1|2|3); axis 0 = no force.  `b` is the host array of b01..b18 (b[18]
 * is ignored: direction 19 is updated in place in a19).
 * The reference computes this only on its CPU path (SRC/col_MC.F90:472-488);
 * its native path passes zero forces (SRC/col_MC.F90:147-149). */
void col_mc_force_gpu(bgk_real_t** a, bgk_real_t** b, int l, int m, int n, int nx, int ny, bgk_real_t omega,
                      bgk_real_t fgrad, int axis);

/* col_mc_force_gpu with the Smagorinsky LES relaxation of the reference's
 * LES builds: per cell omega = 1/Ts, Ts = 1/(2 omega1) + sqrt(18 cteS^2 |P| +
 * (1/omega1)^2)/2 from the non-equilibrium momentum flux P (cteS: namelist
 * cteS, default 0.1); axis 0 = no body force.  The reference's native path
 * has no LES.  Replaces SRC/col_MC.F90:381-466 (with -DLES). */
void col_mc_les_gpu(bgk_real_t** a, bgk_real_t** b, int l, int m, int n, int nx, int ny, bgk_real_t omega1,
                    bgk_real_t cte_s, bgk_real_t fgrad, int axis);

/* Global means over interior cells: out = {<u>, <v>, <w>, <rho>, <0.5|u|^2>}
 * (diagno.dat column order), FP64 accumulation; cte1 = 1 for shifted storage.
 * Replaces the host loop of SRC/diagno.F90:86-138 (and its full D2H copy,
 * SRC/diagnostic.F90:91-94). */
int diagno_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny, bgk_real_t cte1, double* out);

/* diagno_gpu without any allocation, for callers that cannot allocate
 * memory (code between graph captures, tight serving loops, hosts that
 * forbid hipMalloc): no counterpart in the reference.  Same arguments,
 * kernels and result as diagno_gpu; its partial sums live in a static
 * device array of (1024 + 1) * 5 doubles per GPU (41 KB, held until the
 * process ends) instead of the cached scratch and pinned buffer diagno_gpu
 * allocates on first use.  The result comes back through a synchronous copy
 * into `out`.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (other host threads, other streams) race.  Returns 0, -1 for a
 * bad argument or a reduction larger than the array (cannot happen for
 * valid sizes: at most 1024 partials), -2 if the array's address cannot be
 * obtained. */
int diagno_gpu_noalloc(bgk_real_t** a, int l, int m, int n, int nx, int ny, bgk_real_t cte1, double* out);

/* Mean profiles as varm forms them: x-, y- and z-velocity sums per z-plane k
 * (u_k, w_k, v_k, length n) and density sums per y-plane j (den_j, length
 * m), each times 1/l.  Replaces SRC/varm.F90:56-99. */
int varm_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny, bgk_real_t cte1, double* u_k, double* w_k,
             double* v_k, double* den_j);

/* Momentum-exchange force on the obstacle over fluid cells of the box
 * centre +- (radius + 3); out = {Fx, Fy, Fz} / (u_inflow^2 * radius), FP64
 * accumulation.  The reference fixes the centre at (2l/5, m/2, n/2).
 * Replaces SRC/draglift.F90:48-179. */
int draglift_gpu(bgk_real_t** a, const int* obs, int l, int m, int n, int nx, int ny, int icoord, int jcoord,
                 int kcoord, int radius, bgk_real_t u_inflow, double* out);

/* Velocity/density line along k at (icoord, jcoord): out[4*(k-1)+{0,1,2,3}] =
 * {u, w (y-vel), v (z-vel), rho}, host buffer of 4*n values.
 * Replaces SRC/prof_k.F90:47-79. */
int prof_k_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny, int icoord, int jcoord, bgk_real_t cte1,
               bgk_real_t* out);

/* The same along i at (jcoord, kcoord), 4*l values.
 * Replaces SRC/prof_i.F90:47-78. */
int prof_i_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny, int jcoord, int kcoord, bgk_real_t cte1,
               bgk_real_t* out);

/* The same along j at (icoord, kcoord), 4*m values.
 * Replaces SRC/prof_j.F90:47-77. */
int prof_j_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny, int icoord, int kcoord, bgk_real_t cte1,
               bgk_real_t* out);

/* One cell: out = {u, v, w, rho} (x, y, z velocity and density), the values
 * the reference's probe prints.  Replaces SRC/probe.F90:44-75. */
int probe_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny, int i0, int j0, int k0, bgk_real_t cte1,
              bgk_real_t* out);

/* Initial state f = f_eq into HOST arrays h[0..18] (all cells incl. halo),
 * shifted storage: mode 0 rest, mode 1 3-D Taylor-Green with amplitude u00,
 * mode 2 the same with the density perturbation kept (the reference's shifted
 * form drops it; no reference counterpart).  Mode 0 or 1 plus
 * BGK_INIT_UNSHIFTED stores f itself, as the reference's -DNOSHIFT builds
 * (SRC/init.F90:71-75); such a lattice is stepped with col_mc_gpu's
 * cte0 = 1, cte1 = 0, bcond_inflow_gpu's cte1 = 1 and read by the
 * diagnostics with cte1 = 0.
 * Host-side like the reference's init, bit-identical to it.
 * Replaces SRC/init.F90:31-209 (PERIODIC + TGV3D branch for mode 1). */
#define BGK_INIT_UNSHIFTED 8
int bgk_init_host(bgk_real_t** h, int l, int m, int n, int nx, int ny, int mode, bgk_real_t u00);

/* The same state written directly into the DEVICE arrays a[0..18] (all
 * cells incl. halo; padding untouched), bit-identical to bgk_init_host: the
 * host evaluates sin/cos once per coordinate, a kernel forms every cell.
 * Waits for completion; 0 on success, -1 bad argument, -2 HIP error.
 * Replaces the host init plus the full-lattice upload the reference does
 * (SRC/init.F90:31-209, then the OpenMP target data map, SRC/bgk3d.F90:163). */
int bgk_init_gpu(bgk_real_t** a, int l, int m, int n, int nx, int ny, int mode, bgk_real_t u00);

/* Select the in-order stream (a hipStream_t) all entry points queue on;
 * NULL restores the null stream.  No counterpart in the reference. */
int bgk_set_stream(void* stream);

/* Device-memory plumbing for hosts without their own HIP runtime binding
 * (the reference uses OpenMP `target data` maps for this,
 * SRC/bgk3d.F90:162-167). */
int bgk_dev_malloc(void** ptr, size_t bytes);
int bgk_dev_free(void* ptr);
int bgk_memcpy_h2d(void* dst, const void* src, size_t bytes);
int bgk_memcpy_d2h(void* dst, const void* src, size_t bytes);
int bgk_device_count(int* n);

/* Plain device copy of n_floats floats (n_floats % 4 == 0) on the ABI
 * stream: the measured-stream-bandwidth probe (8 bytes moved per float). */
int bgk_stream_copy(const void* src, void* dst, size_t n_floats);

/* Event timing on the stream selected by bgk_set_stream (hipEvent_t handles). */
int bgk_event_create(void** ev);
int bgk_event_record(void* ev);
int bgk_event_elapsed_ms(void* start, void* stop, float* ms);
int bgk_event_destroy(void* ev);

#ifdef __cplusplus
}
#endif

#endif /* BGK3D_GPU_H */
