/*
 * bgk3d_engine.h -- handle API of the gfx950 D3Q19 BGK step (v2).
 *
 * libbgk3d_engine.so owns the lattices, so it can lay them out for the
 * hardware (rows padded so the first interior cell of every row starts a
 * 128-byte line), fold every boundary condition into the pull (no halo
 * kernels per step), store populations in FP32, FP64, FP16 or BF16 (shifted
 * f - w, collision in FP32 or FP64), and advance many independent solves
 * together, one launch per step for all of them (bgks_*).
 *
 * The reference has no such API: its driver owns the arrays and calls the
 * per-routine symbols of bgk3d_gpu.h (SRC/bgk3d.F90:169-196).  Each field
 * below names the reference input it stands for.  All functions return 0 on
 * success and a negative BGK_E* code on error (bgk_last_error() gives text).
 */
#ifndef BGK3D_ENGINE_H
#define BGK3D_ENGINE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum bgk_boundary {
    BGK_BC_DRIVEN = 0,   /* lid-driven cavity, SRC/bcond.driven.F90 (LDC build) */
    BGK_BC_PERIODIC = 1, /* fully periodic, SRC/bcond.periodic.F90 (TGV/HIT builds) */
    BGK_BC_CHANNEL = 2,  /* x,y periodic, walls z=1,n, body force, SRC/bcond.channel.F90 (POF) */
    BGK_BC_INFLOW = 3    /* inflow x=0, outflow x=l+1, y,z periodic, SRC/bcond.inflow.F90 (VKS) */
};

enum bgk_storage { BGK_F32 = 0, BGK_F64 = 1, BGK_F16 = 2, BGK_BF16 = 3 };

enum bgk_error {
    BGK_OK = 0,
    BGK_EINVAL = -1,   /* bad argument */
    BGK_ENOMEM = -2,   /* device allocation failed */
    BGK_EHIP = -3,     /* HIP runtime error */
    BGK_EBUSY = -4     /* serving queue full / slot unavailable */
};

typedef struct bgk_config {
    int l, m, n;          /* interior size (namelist lx, ly, lz; SRC/input.F90:44-48) */
    int boundary;         /* enum bgk_boundary (compile-time case flag in the reference) */
    int storage;          /* enum bgk_storage (SINGLE / DOUBLE_P; F16/BF16 are new) */
    int rounding;         /* 0 nearest-even, 1 stochastic (F16/BF16 stores only) */
    double svisc;         /* viscosity nu (namelist svisc); omega = 2/(6 nu + 1) */
    double u0;            /* lid velocity, or volume-force scale for the channel (namelist u0) */
    int init;             /* 0 rest, 1 3-D Taylor-Green (SRC/init.F90), 2 the same with
                             the density perturbation kept (not in the reference) */
    int obstacle;         /* 0 none, 1 hollow sphere shell (SRC/build.obs.F90), 2 solid sphere */
    int obs_ic, obs_jc, obs_kc, obs_radius; /* sphere centre/radius; < 0 = reference defaults */
    double u_inflow;      /* inflow velocity; < 0 = reference value (default-real 0.1) */
    int force_axis;       /* channel body-force axis 1|2|3 (FORCING_Y/Z select 2/3) */
    int device;           /* HIP device ordinal */
    unsigned seed;        /* stochastic-rounding stream */
    int les;              /* 1: Smagorinsky LES relaxation (the reference's -DLES builds,
                             SRC/col_MC.F90:381-466) */
    double cte_s;         /* Smagorinsky constant (namelist cteS); < 0 = reference default
                             (default-real 0.1, SRC/input.F90:59) */
    int wall_axis;        /* channel walls: 3 at z = 1, n (the reference, SRC/bcond.channel.F90:
                             255-267), 2 at y = 1, m (z then periodic; force_axis 1 or 3) */
    int mixed;            /* the reference's mixed-precision builds (SRC/precision.F90:30-43):
                             0 none; 1 MIXED1: storage BGK_F16, the collision in FP32 in the
                             reference's expression order (the plain F16 form uses a leaner
                             FP32 arithmetic); 2 MIXED2: storage BGK_F32, collision in FP64 */
    int layout;           /* 0 choose, 1 rows padded to 128-byte lines, 2 compact rows of l + 2
                             cells.  The default takes compact rows
                             where they measured faster: when padding would leave more than 3 %
                             of a row unused in its last line (9 % with 16-bit storage; a line
                             holds 32 values in FP32, 16 in FP64, 64 in 16-bit storage), and
                             with FP32/FP64 storage for lattices of 2^24 cells or more */
} bgk_config;

typedef struct bgk_handle bgk_handle;

/* Fill `cfg` with the reference's defaults for a driven cavity of l x m x n. */
void bgk_config_default(bgk_config* cfg, int l, int m, int n);

int bgk_create(const bgk_config* cfg, bgk_handle** out);
int bgk_destroy(bgk_handle* h);

/* Queue nsteps time steps (boundaries folded in + fused collide + swap) on
 * the handle's stream; returns without waiting. */
int bgk_step(bgk_handle* h, int nsteps);
int bgk_sync(bgk_handle* h);

/* Run nsteps and time them: total wall time of the steps (hipEvents on the
 * handle's stream) and the mean duration of the fused collide kernel
 * (events bracketing each launch). */
int bgk_time_steps(bgk_handle* h, int nsteps, double* total_ms, double* collide_avg_ms);

/* Global means {<u>, <v>, <w>, <rho>, <0.5|u|^2>} (SRC/diagno.F90), FP64 sums. */
int bgk_diagno(bgk_handle* h, double out[5]);

/* Momentum-exchange force on the obstacle / (u_inflow^2 radius)
 * (SRC/draglift.F90), FP64 sums; needs an obstacle. */
int bgk_draglift(bgk_handle* h, double out[3]);

/* Interior macroscopic fields, i fastest then j then k (l*m*n values each);
 * any pointer may be NULL. */
int bgk_fields(bgk_handle* h, double* rho, double* ux, double* uy, double* uz);

/* Velocity/density line along k at (icoord, jcoord), as the reference's
 * prof_k prints it (SRC/prof_k.F90:47-79): out[4*(k-1)+{0,1,2,3}] =
 * {u, w (y-velocity), v (z-velocity), rho}, computed in the collision kind. */
int bgk_prof_k(bgk_handle* h, int icoord, int jcoord, double* out);

/* The same along any axis: axis 0 runs i at (j, k) = (c1, c2)
 * (SRC/prof_i.F90:47-78), axis 1 runs j at (i, k) (SRC/prof_j.F90:47-77),
 * axis 2 runs k at (i, j) (SRC/prof_k.F90); 4 values per point. */
int bgk_profile(bgk_handle* h, int axis, int c1, int c2, double* out);

/* One cell: {u, v, w, rho} (x, y, z velocity, density) as the reference's
 * probe prints them (SRC/probe.F90:44-75). */
int bgk_probe(bgk_handle* h, int i0, int j0, int k0, double out[4]);

/* Restart file in the reference's format (SRC/save_raw.F90:27-69): Fortran
 * sequential unformatted records (4-byte length markers), the time step,
 * then a01..a19 over (0:l+1, 0:m+1, 0:n+1), i fastest, in the storage type.
 * The halo cells hold what the handle holds there (the library never reads
 * them back: boundaries are applied in the fused step). */
int bgk_save_raw(bgk_handle* h, const char* path, int itime);

/* Read such a file (SRC/restore_raw.F90:27-64): the grid and storage type
 * must match; the handle's step count becomes the file's time step. */
int bgk_restore_raw(bgk_handle* h, const char* path, int* itime);

/* Velocity file (SRC/save_vel.F90): a record with the time step, then one
 * record of (u, v, w) per interior cell, i fastest, in the collision kind. */
int bgk_save_vel(bgk_handle* h, const char* path, int itime);

/* Legacy VTK file with an ASCII header and big-endian FP32 vectors, as the
 * reference writes them: kind 0 velocity on the xy plane k = index
 * (SRC/vtk_xy.binary.F90), 1 xz plane j = index (SRC/vtk_xz.binary.F90),
 * 2 yz plane i = index (SRC/vtk_yz.binary.F90), 3 vorticity on the xy plane
 * k = index (SRC/vtk_om.xy.binary.F90; the reference's sign convention). */
int bgk_write_vtk(bgk_handle* h, int kind, int index, const char* path);

/* Interior values of population d (0..18, direction d+1), as double. */
int bgk_get_population(bgk_handle* h, int d, double* out);
/* Overwrite interior values of population d (rounded to the storage type). */
int bgk_set_population(bgk_handle* h, int d, const double* in);

/* Sizes: [0] nx, [1] ny, [2] nz, [3] bytes of device memory held,
 * [4] storage bytes per population value, [5] steps taken. */
int bgk_info(bgk_handle* h, long long out[6]);

/* Text of the last error on this thread. */
const char* bgk_last_error(void);

/* ---------------------------------------------------------------------------
 * Continuous-batching server: many independent lid-driven cavity solves
 * (each the LDC build of the reference, SRC/bcond.driven.F90, rest start)
 * advanced together on one GPU.  The reference's analogue is a shell loop
 * that runs one program per case (UTIL/run_sweep.sh:55-63, 128-169).
 *
 * Requests are admitted while the device arena reserved by bgks_create (the
 * HBM budget) can hold their two lattices; the requests scheduled for a
 * chunk advance sub-batch by sub-batch (at most 512 MB of lattices of one
 * storage, rounding and row layout), one kernel launch per sub-batch and
 * step, on one stream; requests join and leave between chunks of at
 * most chunk_steps steps, cut shorter by a chunk-time target if one is set
 * (see enum bgks_policy).  A request whose populations turn non-finite (the
 * step kernel checks every cell) is retired as unstable at the end of that
 * chunk.  A request's populations are bit-identical to those of a handle
 * (bgk_create with the same n, svisc, u0, storage, rounding and seed) after
 * the same number of steps, whatever else is running.  Not thread-safe: one
 * caller drives a server; at most 4096 requests are admitted at once.
 * ------------------------------------------------------------------------- */
typedef struct bgks_server bgks_server;

typedef struct bgks_request {
    int n;               /* cavity n x n x n (namelist lx = ly = lz) */
    double svisc;        /* viscosity; Re = u0 n / svisc (SRC/outdat.F90:46) */
    double u0;           /* lid velocity (namelist u0) */
    int storage;         /* enum bgk_storage */
    int rounding;        /* 0 nearest-even, 1 stochastic (F16/BF16 only) */
    long long steps;     /* time steps to run (namelist itfin) */
    unsigned seed;       /* stochastic-rounding stream */
    int keep;            /* 1: keep the lattice after completion, for
                            bgks_get_population, until bgks_release */
} bgks_request;

enum bgks_state { BGKS_QUEUED = 0, BGKS_ACTIVE = 1, BGKS_DONE = 2 };

enum bgks_policy {
    BGKS_ALL_ACTIVE = 0,     /* admit in arrival order; every admitted request advances each chunk */
    BGKS_SHORTEST_FIRST = 1  /* admit and advance the requests with the least work (cells x steps)
                                left first, enough of them to keep min_cells_per_step in flight;
                                a request passed over 8 chunks in a row goes first (aging) */
};

/* Reserve `budget_bytes` of device memory for request lattices
 * (policy BGKS_ALL_ACTIVE). */
int bgks_create(int device, size_t budget_bytes, int chunk_steps, bgks_server** out);
int bgks_destroy(bgks_server* s);

/* Choose the scheduling policy (takes effect at the next bgks_advance). */
int bgks_set_policy(bgks_server* s, int policy, long long min_cells_per_step);

/* Cut chunks so that each takes about target_ms milliseconds (estimated from
 * the previous chunk's time per cell update); 0 (the default) runs
 * chunk_steps steps.  Requests join between chunks, so this bounds how long
 * an arrival waits to start. */
int bgks_set_chunk_time(bgks_server* s, double target_ms);

/* The shortest-first selection the server makes each chunk, for hosts that
 * schedule requests themselves: of `count` requests with remaining work
 * (cells x steps), cells and chunks passed over in a row, every one passed
 * over max_skips or more times, then the others by least remaining work
 * (ties in index order) until min_cells_per_step cells are selected (at least
 * one request).  selected[] receives the indices, *n_selected their number. */
int bgks_schedule(const long long* remaining, const long long* cells, const int* skipped, int count,
                  long long min_cells_per_step, int max_skips, int* selected, int* n_selected);

/* Device bytes a request holds while admitted (both lattices, padded). */
int bgks_request_bytes(const bgks_request* r, size_t* bytes);

/* Queue a request; fails with BGK_ENOMEM if it can never fit the budget. */
int bgks_submit(bgks_server* s, const bgks_request* r, long long* id);

/* Admit what fits, run one chunk over all active requests, wait for it,
 * retire the finished ones (device diagnostics, lattice freed unless keep)
 * and admit again.  The ids finished by this call are written to done_ids
 * (at most max_done of them); *n_done receives their number. */
int bgks_advance(bgks_server* s, long long* done_ids, int max_done, int* n_done);

/* enum bgks_state and steps completed. */
int bgks_status(bgks_server* s, long long id, int* state, long long* steps_done);

/* A finished request: diagnostics means as bgk_diagno, the times (seconds
 * on the bgks_clock scale) of submission, admission and completion, and
 * whether the solution became non-finite (unstable; such a request is
 * retired when that is detected, bgks_status giving the steps it ran). */
int bgks_result(bgks_server* s, long long id, double diag[5], double times[3], int* unstable);

/* Interior values of population d of a request that holds its lattice. */
int bgks_get_population(bgks_server* s, long long id, int d, double* out);

/* Forget a finished or queued request and free its lattice. */
int bgks_release(bgks_server* s, long long id);

/* [0] queued, [1] active, [2] finished and not released, [3] arena bytes in
 * use, [4] arena bytes, [5] kernel launches, [6] cell updates, [7] chunks. */
int bgks_stats(bgks_server* s, long long out[8]);

/* Monotonic clock (seconds) used for the request times. */
double bgks_clock(void);

#ifdef __cplusplus
}
#endif

#endif /* BGK3D_ENGINE_H */
