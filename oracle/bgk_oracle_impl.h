/*
 * CPU oracle for the D3Q19 BGK time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a plain-C restatement of the reference's Fortran CPU path
 * (the `#else ... do concurrent` branches, FUSED build, shifted storage).
 * It exists so that tests, the smoke check and the CPU-baseline leg of the
 * benchmark have something to compare the HIP kernels against.  Nothing in
 * the product library links, loads or calls it.
 *
 * It is included twice by bgk_oracle.c, once with R=float (suffix _f32) and
 * once with R=double (suffix _f64), mirroring the reference's SINGLE and
 * DOUBLE_P builds (SRC/precision.F90:1-49: mystorage == mykind).
 *
 * Expression order follows the Fortran source term by term so that the
 * result is bit-identical to an amdflang -O3 build of the reference on
 * x86-64 (no FMA contraction is possible there, and this file is compiled
 * with -ffp-contract=off).  Mixed-kind literals of the Fortran source (a
 * default-real `6.0`, a `0.5d0`) are reproduced with explicit conversions.
 *
 * Array layout (SRC/alloca.F90:153-191): each of the 19 populations is a
 * column-major array (0:l+1+ipad, 0:m+1+jpad, 0:n+1) -> element (i,j,k) at
 * i + nx*j + nx*ny*k with nx = l+2+ipad, ny = m+2+jpad.  `a[d]` is the base
 * pointer of population d+1.
 *
 * Pinning (each function's check lives in tests/test_oracle_*.py):
 *   - every function is compared bit-for-bit with the compiled reference
 *     (oracle/_ref, built from the reference's own Fortran sources) on
 *     small grids, and against committed fixtures produced from it;
 *   - collide: mass/momentum conservation, f_eq fixed point, omega=0 pure
 *     streaming permutation, closed-form shifted equilibrium in float64;
 *   - periodic: translation invariance (L steps of pure streaming return the
 *     initial state bitwise);
 *   - reductions: closed form on uniform states, numpy brute force.
 */
#ifndef R
#error "define R (float or double) and SFX before including"
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

#define IDX(i, j, k) ((size_t)(i) + (size_t)nx * (size_t)(j) + nxy * (size_t)(k))

/* D3Q19 weights, as computed in SRC/storage.F90:58-66 (-DPGI branch:
 * double-precision parameters rounded to the storage kind). */
static const R FN(P0_) = (R)(1.0 / 3.0);
static const R FN(P1_) = (R)(1.0 / 18.0);
static const R FN(P2_) = (R)(1.0 / (3.0 * 3.0 * (1.0 + 3.0)));

/* ------------------------------------------------------------------------ */
/* hencol: SRC/hencol.F90:31 and :52                                         */
/* omega = (uno+uno)/((6.0*svisc)+1.0) ; fgrad = 4.0*u0*svisc/(5.0*m*m)      */
/* The literals 6.0, 1.0, 4.0, 5.0 are default-real (single) constants that  */
/* are exact in either kind, so all arithmetic happens in kind R.            */
/* ------------------------------------------------------------------------ */
void FN(oracle_hencol_)(R svisc, R u0, int m, R *omega, R *fgrad)
{
    const R uno = (R)1;
    *omega = (uno + uno) / (((R)6.0 * svisc) + (R)1.0);
    *fgrad = (R)4.0 * u0 * svisc / ((R)5.0 * (R)m * (R)m);
}

/* ------------------------------------------------------------------------ */
/* init: SRC/init.F90:31-209; cte1 = 0 shifted, 1 unshifted (:71-75).      */
/* mode 0: rest flow (all populations 0).                                    */
/* mode 1: 3-D Taylor-Green (PERIODIC + TGV3D, :114-148).                     */
/* mode 2: mode 1 with the density perturbation kept (the reference's        */
/*         shifted form stores crho*w*c and drops w*(crho-1); no reference   */
/*         counterpart, pinned by its moments in the tests).                 */
/* Every cell including the halo (0..l1, 0..m1, 0..n1) is written; the       */
/* ipad/jpad padding is left untouched (it keeps the huge() poison).         */
/* ------------------------------------------------------------------------ */
void FN(oracle_init_)(R **a, int l, int m, int n, int nx, int ny, int mode, R u00)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const R uno = (R)1, due = (R)2, tre = (R)3, qua = (R)4, zero = (R)0;
    const R rf = (R)3, qf = (R)1.5, cte1 = g_noshift ? uno : zero;
    const R p0 = FN(P0_), p1 = FN(P1_), p2 = FN(P2_);
    /* pi is a real(mykind) parameter whose value is a default-real (single)
     * literal (SRC/init.F90:63), so the DP build sees float-rounded pi. */
    const R pi = (R)3.141592653589793238462643383279f;
    for (int k = 0; k <= n + 1; ++k) {
        /* z = (real(k,mykind)-0.5d0)/real(nn,mykind): evaluated in double,
         * stored in kind R (SRC/init.F90:115) */
        const R z = (R)(((double)(R)k - 0.5) / (double)(R)n);
        for (int j = 0; j <= m + 1; ++j) {
            const R y = (R)(((double)(R)j - 0.5) / (double)(R)m);
            for (int i = 0; i <= l + 1; ++i) {
                const R x = (R)(((double)(R)i - 0.5) / (double)(R)l);
                R xj = zero, yj = zero, zj = zero, crho = uno, drho = zero;
                if (mode == 1 || mode == 2) {
                    /* SRC/init.F90:142-148 */
                    xj = u00 * CAT(SIN, SFX)(due * pi * x) * CAT(COS, SFX)(due * pi * y) *
                         CAT(COS, SFX)(due * pi * z);
                    yj = -u00 * CAT(COS, SFX)(due * pi * x) * CAT(SIN, SFX)(due * pi * y) *
                         CAT(COS, SFX)(due * pi * z);
                    zj = zero;
                    drho = ((u00 * u00) / (qua * qua * tre)) *
                           (CAT(COS, SFX)(qua * pi * x) + CAT(COS, SFX)(qua * pi * y)) *
                           (CAT(COS, SFX)(qua * pi * z) + due);
                    crho = uno + drho;
                }
                const R cvsq = xj * xj + yj * yj + zj * zj;
                /* SRC/init.F90:159-177 */
                R c[19];
                c[0] = rf * (xj - yj) + qf * (tre * (xj - yj) * (xj - yj) - cvsq);
                c[1] = rf * (xj - zj) + qf * (tre * (xj - zj) * (xj - zj) - cvsq);
                c[2] = rf * (xj + yj) + qf * (tre * (xj + yj) * (xj + yj) - cvsq);
                c[3] = rf * (xj + zj) + qf * (tre * (xj + zj) * (xj + zj) - cvsq);
                c[4] = rf * (xj) + qf * (tre * (xj) * (xj)-cvsq);
                c[5] = rf * (zj) + qf * (tre * (zj) * (zj)-cvsq);
                c[6] = rf * (yj + zj) + qf * (tre * (yj + zj) * (yj + zj) - cvsq);
                c[7] = rf * (yj) + qf * (tre * (yj) * (yj)-cvsq);
                c[8] = rf * (yj - zj) + qf * (tre * (yj - zj) * (yj - zj) - cvsq);
                c[9] = rf * (-xj - yj) + qf * (tre * (xj + yj) * (xj + yj) - cvsq);
                c[10] = rf * (-xj - zj) + qf * (tre * (xj + zj) * (xj + zj) - cvsq);
                c[11] = rf * (-xj + yj) + qf * (tre * (xj - yj) * (xj - yj) - cvsq);
                c[12] = rf * (-xj + zj) + qf * (tre * (xj - zj) * (xj - zj) - cvsq);
                c[13] = rf * (-xj) + qf * (tre * (xj) * (xj)-cvsq);
                c[14] = rf * (-zj) + qf * (tre * (zj) * (zj)-cvsq);
                c[15] = rf * (-yj - zj) + qf * (tre * (yj + zj) * (yj + zj) - cvsq);
                c[16] = rf * (-yj) + qf * (tre * (yj) * (yj)-cvsq);
                c[17] = rf * (-yj + zj) + qf * (tre * (yj - zj) * (yj - zj) - cvsq);
                /* rf*(0.0)+qf*(tre*(0.0)*(0.0)-cvsq), 0.0 default real */
                c[18] = rf * zero + qf * (tre * zero * zero - cvsq);
                static const int wclass[19] = {2, 2, 2, 2, 1, 1, 2, 1, 2, 2, 2, 2, 2, 1, 1, 2, 1, 2, 0};
                const size_t id = IDX(i, j, k);
                for (int d = 0; d < 19; ++d) {
                    const R w = wclass[d] == 0 ? p0 : (wclass[d] == 1 ? p1 : p2);
                    /* SRC/init.F90:179-197: crho*pX*(cte1+cxNN).  Mode 2 keeps
                     * the density perturbation the shifted form drops:
                     * f - w = w (rho - 1) + w rho c = w (drho + crho c). */
                    a[d][id] = mode == 2 ? w * (drho + crho * c[d]) : crho * w * (cte1 + c[d]);
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* col_MC: SRC/col_MC.F90:126-536 (do concurrent branch) for one step.       */
/* Reads a[0..18], writes b[0..17] and a[18] in place (interior only).       */
/* The caller performs the a<->b swap of populations 1..18                  */
/* (SRC/col_MC.F90:539-596).                                                 */
/* force_mode 0: uniform forcex/forcey/forcez (the values passed; 0 in the   */
/*               reference's non-CHANNEL builds, :147-149);                  */
/* force_mode 1: CHANNEL forcing forcex = fgrad*rho (:482);                  */
/* force_mode 2: FORCING_Y forcey = fgrad*rho (:475);                        */
/* force_mode 3: FORCING_Z forcez = fgrad*rho (:480);                        */
/* force_mode | 4: LES builds, per-cell Smagorinsky rate from the            */
/*               non-equilibrium flux with constant cteS (:381-466).         */
/* k-planes are independent, so the outer loop may run threaded; the result  */
/* does not depend on the thread count.                                      */
/* ------------------------------------------------------------------------ */
void FN(oracle_col_mc_)(R **a, R **b, int l, int m, int n, int nx, int ny, R omega1, int force_mode,
                        R fgrad, R fx_u, R fy_u, R fz_u, R cteS, int nthreads)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const R uno = (R)1, zero = (R)0;
    const R cte1 = g_noshift ? zero : uno, cte0 = uno - cte1; /* :139-144 */
    const R rf = (R)3, qf = (R)1.5, tre = (R)3;
    const R p0 = FN(P0_), p1 = FN(P1_), p2 = FN(P2_);
    R *const a01 = a[0], *const a02 = a[1], *const a03 = a[2], *const a04 = a[3], *const a05 = a[4];
    R *const a06 = a[5], *const a07 = a[6], *const a08 = a[7], *const a09 = a[8], *const a10 = a[9];
    R *const a11 = a[10], *const a12 = a[11], *const a13 = a[12], *const a14 = a[13], *const a15 = a[14];
    R *const a16 = a[15], *const a17 = a[16], *const a18 = a[17], *const a19 = a[18];
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1) schedule(static) collapse(2)
#else
    (void)nthreads;
#endif
    for (int k = 1; k <= n; ++k)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= l; ++i) {
                /* pull, :275-293 */
                const R x01 = a01[IDX(i - 1, j + 1, k)];
                const R x02 = a02[IDX(i - 1, j, k + 1)];
                const R x03 = a03[IDX(i - 1, j - 1, k)];
                const R x04 = a04[IDX(i - 1, j, k - 1)];
                const R x05 = a05[IDX(i - 1, j, k)];
                const R x06 = a06[IDX(i, j, k - 1)];
                const R x07 = a07[IDX(i, j - 1, k - 1)];
                const R x08 = a08[IDX(i, j - 1, k)];
                const R x09 = a09[IDX(i, j - 1, k + 1)];
                const R x10 = a10[IDX(i + 1, j + 1, k)];
                const R x11 = a11[IDX(i + 1, j, k + 1)];
                const R x12 = a12[IDX(i + 1, j - 1, k)];
                const R x13 = a13[IDX(i + 1, j, k - 1)];
                const R x14 = a14[IDX(i + 1, j, k)];
                const R x15 = a15[IDX(i, j, k + 1)];
                const R x16 = a16[IDX(i, j + 1, k + 1)];
                const R x17 = a17[IDX(i, j + 1, k)];
                const R x18 = a18[IDX(i, j + 1, k - 1)];
                const R x19 = a19[IDX(i, j, k)];
                /* moments, :295-303 */
                const R rho = x01 + x02 + x03 + x04 + x05 + x06 + x07 + x08 + x09 + x10 + x11 + x12 +
                              x13 + x14 + x15 + x16 + x17 + x18 + x19 + cte1;
                const R rhoinv = uno / rho;
                R vx = (x01 + x02 + x03 + x04 + x05 - x10 - x11 - x12 - x13 - x14) * rhoinv;
                R vy = (x03 + x07 + x08 + x09 + x12 - x01 - x10 - x16 - x17 - x18) * rhoinv;
                R vz = (x04 + x06 + x07 + x13 + x18 - x02 - x09 - x11 - x15 - x16) * rhoinv;
                /* quadratic terms, :321-342 */
                const R vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
                const R vsq = vx2 + vy2 + vz2;
                R vxpy = vx + vy;
                const R qxpy = cte0 + qf * (tre * vxpy * vxpy - vsq);
                R vxmy = vx - vy;
                const R qxmy = cte0 + qf * (tre * vxmy * vxmy - vsq);
                R vxpz = vx + vz;
                const R qxpz = cte0 + qf * (tre * vxpz * vxpz - vsq);
                R vxmz = vx - vz;
                const R qxmz = cte0 + qf * (tre * vxmz * vxmz - vsq);
                R vypz = vy + vz;
                const R qypz = cte0 + qf * (tre * vypz * vypz - vsq);
                R vymz = vy - vz;
                const R qymz = cte0 + qf * (tre * vymz * vymz - vsq);
                const R qx = cte0 + qf * (tre * vx2 - vsq);
                const R qy = cte0 + qf * (tre * vy2 - vsq);
                const R qz = cte0 + qf * (tre * vz2 - vsq);
                const R q0 = cte0 + qf * (-vsq);
                /* linear terms, :345-353 */
                vx = rf * vx;
                vy = rf * vy;
                vz = rf * vz;
                vxpy = rf * vxpy;
                vxmy = rf * vxmy;
                vxpz = rf * vxpz;
                vxmz = rf * vxmz;
                vypz = rf * vypz;
                vymz = rf * vymz;
                /* :356-358 */
                const R rp0 = rho * p0, rp1 = rho * p1, rp2 = rho * p2;
                /* equilibrium, :361-379 */
                const R e01 = rp2 * (+vxmy + qxmy) + cte1 * (rp2 - p2);
                const R e02 = rp2 * (+vxmz + qxmz) + cte1 * (rp2 - p2);
                const R e03 = rp2 * (+vxpy + qxpy) + cte1 * (rp2 - p2);
                const R e04 = rp2 * (+vxpz + qxpz) + cte1 * (rp2 - p2);
                const R e05 = rp1 * (+vx + qx) + cte1 * (rp1 - p1);
                const R e06 = rp1 * (+vz + qz) + cte1 * (rp1 - p1);
                const R e07 = rp2 * (+vypz + qypz) + cte1 * (rp2 - p2);
                const R e08 = rp1 * (+vy + qy) + cte1 * (rp1 - p1);
                const R e09 = rp2 * (+vymz + qymz) + cte1 * (rp2 - p2);
                const R e10 = rp2 * (-vxpy + qxpy) + cte1 * (rp2 - p2);
                const R e11 = rp2 * (-vxpz + qxpz) + cte1 * (rp2 - p2);
                const R e12 = rp2 * (-vxmy + qxmy) + cte1 * (rp2 - p2);
                const R e13 = rp2 * (-vxmz + qxmz) + cte1 * (rp2 - p2);
                const R e14 = rp1 * (-vx + qx) + cte1 * (rp1 - p1);
                const R e15 = rp1 * (-vz + qz) + cte1 * (rp1 - p1);
                const R e16 = rp2 * (-vypz + qypz) + cte1 * (rp2 - p2);
                const R e17 = rp1 * (-vy + qy) + cte1 * (rp1 - p1);
                const R e18 = rp2 * (-vymz + qymz) + cte1 * (rp2 - p2);
                const R e19 = rp0 * (+q0) + cte1 * (rp0 - p0);
                R omega = omega1;
                if (force_mode & 4) {
                    /* LES, :381-466: non-equilibrium parts, momentum flux,
                     * Smagorinsky relaxation time */
                    const R n01 = x01 - e01, n02 = x02 - e02, n03 = x03 - e03, n04 = x04 - e04;
                    const R n05 = x05 - e05, n06 = x06 - e06, n07 = x07 - e07, n08 = x08 - e08;
                    const R n09 = x09 - e09, n10 = x10 - e10, n11 = x11 - e11, n12 = x12 - e12;
                    const R n13 = x13 - e13, n14 = x14 - e14, n15 = x15 - e15, n16 = x16 - e16;
                    const R n17 = x17 - e17, n18 = x18 - e18;
                    const R Pxx = n01 + n02 + n03 + n04 + n05 + n10 + n11 + n12 + n13 + n14;
                    const R Pyy = n01 + n03 + n07 + n08 + n09 + n10 + n12 + n16 + n17 + n18;
                    const R Pzz = n02 + n04 + n06 + n07 + n09 + n11 + n13 + n15 + n16 + n18;
                    const R Pxz = -n02 + n04 + n11 - n13;
                    const R Pxy = -n01 + n03 + n10 - n12;
                    const R Pyz = +n07 - n09 + n16 - n18;
                    const R Pyx = Pxy, Pzx = Pxz, Pzy = Pyz;
                    /* 2.0 is a default-real literal (exact in both kinds) */
                    const R Ptotal = CAT(SQRT, SFX)((Pxx * Pxx) + (Pyy * Pyy) + (Pzz * Pzz) + ((R)2.0f * Pxy * Pyx) +
                                                    ((R)2.0f * Pxz * Pzx) + ((R)2.0f * Pyz * Pzy));
                    const R Ts = (R)1 / ((R)2 * omega1) +
                                 CAT(SQRT, SFX)((R)18 * (cteS * cteS) * Ptotal + ((R)1 / omega1) * ((R)1 / omega1)) /
                                     (R)2;
                    omega = (R)1 / Ts;
                }
                /* forcing, :472-488 */
                R forcex = fx_u, forcey = fy_u, forcez = fz_u;
                if ((force_mode & 3) == 1) {
                    forcex = fgrad * rho;
                    forcey = zero;
                    forcez = zero;
                } else if ((force_mode & 3) == 2) {
                    forcex = zero;
                    forcey = fgrad * rho;
                    forcez = zero;
                } else if ((force_mode & 3) == 3) {
                    forcex = zero;
                    forcey = zero;
                    forcez = fgrad * rho;
                }
                /* relax + write, :501-519 */
                const size_t id = IDX(i, j, k);
                b[0][id] = x01 - omega * (x01 - e01) + forcex - forcey;
                b[1][id] = x02 - omega * (x02 - e02) + forcex - forcez;
                b[2][id] = x03 - omega * (x03 - e03) + forcex + forcey;
                b[3][id] = x04 - omega * (x04 - e04) + forcex + forcez;
                b[4][id] = x05 - omega * (x05 - e05) + forcex;
                b[5][id] = x06 - omega * (x06 - e06) + forcez;
                b[6][id] = x07 - omega * (x07 - e07) + forcey + forcez;
                b[7][id] = x08 - omega * (x08 - e08) + forcey;
                b[8][id] = x09 - omega * (x09 - e09) + forcey - forcez;
                b[9][id] = x10 - omega * (x10 - e10) - forcex - forcey;
                b[10][id] = x11 - omega * (x11 - e11) - forcex - forcez;
                b[11][id] = x12 - omega * (x12 - e12) - forcex + forcey;
                b[12][id] = x13 - omega * (x13 - e13) - forcex + forcez;
                b[13][id] = x14 - omega * (x14 - e14) - forcex;
                b[14][id] = x15 - omega * (x15 - e15) - forcez;
                b[15][id] = x16 - omega * (x16 - e16) - forcey - forcez;
                b[16][id] = x17 - omega * (x17 - e17) - forcey;
                b[17][id] = x18 - omega * (x18 - e18) - forcey + forcez;
                a19[id] = x19 - omega * (x19 - e19);
            }
}

/* ------------------------------------------------------------------------ */
/* bcond_driven: SRC/bcond.driven.F90:61-262, faces in order x, y, z.        */
/* ------------------------------------------------------------------------ */
void FN(oracle_bcond_driven_)(R **a, int l, int m, int n, int nx, int ny, R u00)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const int l1 = l + 1, m1 = m + 1, n1 = n + 1;
    /* force = u00/(6.0): 6.0 is a default-real literal, exact in kind R */
    const R force = u00 / (R)6.0;
#define A(d) a[(d)-1]
    for (int k = 1; k <= n; ++k) /* :105-119 */
        for (int j = 1; j <= m; ++j) {
            A(12)[IDX(l1, j - 1, k)] = A(1)[IDX(l, j, k)];
            A(13)[IDX(l1, j, k - 1)] = A(2)[IDX(l, j, k)];
            A(10)[IDX(l1, j + 1, k)] = A(3)[IDX(l, j, k)];
            A(11)[IDX(l1, j, k + 1)] = A(4)[IDX(l, j, k)];
            A(14)[IDX(l1, j, k)] = A(5)[IDX(l, j, k)];
            A(3)[IDX(0, j - 1, k)] = A(10)[IDX(1, j, k)];
            A(4)[IDX(0, j, k - 1)] = A(11)[IDX(1, j, k)];
            A(1)[IDX(0, j + 1, k)] = A(12)[IDX(1, j, k)];
            A(2)[IDX(0, j, k + 1)] = A(13)[IDX(1, j, k)];
            A(5)[IDX(0, j, k)] = A(14)[IDX(1, j, k)];
        }
    for (int k = 1; k <= n; ++k) /* :169-183 */
        for (int i = 1; i <= l; ++i) {
            A(8)[IDX(i, 0, k)] = A(17)[IDX(i, 1, k)];
            A(12)[IDX(i + 1, 0, k)] = A(1)[IDX(i, 1, k)];
            A(3)[IDX(i - 1, 0, k)] = A(10)[IDX(i, 1, k)];
            A(7)[IDX(i, 0, k - 1)] = A(16)[IDX(i, 1, k)];
            A(9)[IDX(i, 0, k + 1)] = A(18)[IDX(i, 1, k)];
            A(10)[IDX(i + 1, m1, k)] = A(3)[IDX(i, m, k)];
            A(16)[IDX(i, m1, k + 1)] = A(7)[IDX(i, m, k)];
            A(17)[IDX(i, m1, k)] = A(8)[IDX(i, m, k)];
            A(18)[IDX(i, m1, k - 1)] = A(9)[IDX(i, m, k)];
            A(1)[IDX(i - 1, m1, k)] = A(12)[IDX(i, m, k)];
        }
    for (int j = 1; j <= m; ++j) /* :233-247 */
        for (int i = 1; i <= l; ++i) {
            A(6)[IDX(i, j, 0)] = A(15)[IDX(i, j, 1)];
            A(4)[IDX(i - 1, j, 0)] = A(11)[IDX(i, j, 1)];
            A(7)[IDX(i, j - 1, 0)] = A(16)[IDX(i, j, 1)];
            A(13)[IDX(i + 1, j, 0)] = A(2)[IDX(i, j, 1)];
            A(18)[IDX(i, j + 1, 0)] = A(9)[IDX(i, j, 1)];
            A(15)[IDX(i, j, n1)] = A(6)[IDX(i, j, n)];
            A(2)[IDX(i - 1, j, n1)] = A(13)[IDX(i, j, n)] + force;
            A(9)[IDX(i, j - 1, n1)] = A(18)[IDX(i, j, n)];
            A(11)[IDX(i + 1, j, n1)] = A(4)[IDX(i, j, n)] - force;
            A(16)[IDX(i, j + 1, n1)] = A(7)[IDX(i, j, n)];
        }
}

/* periodic x face (SRC/bcond.periodic.F90:105-119, also channel :120-135) */
static void FN(per_x_)(R **a, int l, int m, int n, int nx, int ny)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const int l1 = l + 1;
    for (int k = 0; k <= n + 1; ++k)
        for (int j = 0; j <= m + 1; ++j) {
            A(1)[IDX(0, j, k)] = A(1)[IDX(l, j, k)];
            A(2)[IDX(0, j, k)] = A(2)[IDX(l, j, k)];
            A(3)[IDX(0, j, k)] = A(3)[IDX(l, j, k)];
            A(4)[IDX(0, j, k)] = A(4)[IDX(l, j, k)];
            A(5)[IDX(0, j, k)] = A(5)[IDX(l, j, k)];
            A(10)[IDX(l1, j, k)] = A(10)[IDX(1, j, k)];
            A(11)[IDX(l1, j, k)] = A(11)[IDX(1, j, k)];
            A(12)[IDX(l1, j, k)] = A(12)[IDX(1, j, k)];
            A(13)[IDX(l1, j, k)] = A(13)[IDX(1, j, k)];
            A(14)[IDX(l1, j, k)] = A(14)[IDX(1, j, k)];
        }
}

/* periodic y face (SRC/bcond.periodic.F90:171-185, channel :187-201,
 * inflow :233-247) */
static void FN(per_y_)(R **a, int l, int m, int n, int nx, int ny)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const int m1 = m + 1;
    for (int k = 0; k <= n + 1; ++k)
        for (int i = 0; i <= l + 1; ++i) {
            A(3)[IDX(i, 0, k)] = A(3)[IDX(i, m, k)];
            A(7)[IDX(i, 0, k)] = A(7)[IDX(i, m, k)];
            A(8)[IDX(i, 0, k)] = A(8)[IDX(i, m, k)];
            A(9)[IDX(i, 0, k)] = A(9)[IDX(i, m, k)];
            A(12)[IDX(i, 0, k)] = A(12)[IDX(i, m, k)];
            A(1)[IDX(i, m1, k)] = A(1)[IDX(i, 1, k)];
            A(10)[IDX(i, m1, k)] = A(10)[IDX(i, 1, k)];
            A(16)[IDX(i, m1, k)] = A(16)[IDX(i, 1, k)];
            A(17)[IDX(i, m1, k)] = A(17)[IDX(i, 1, k)];
            A(18)[IDX(i, m1, k)] = A(18)[IDX(i, 1, k)];
        }
}

/* periodic z face (SRC/bcond.periodic.F90:238-252, inflow :299-313) */
static void FN(per_z_)(R **a, int l, int m, int n, int nx, int ny)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const int n1 = n + 1;
    for (int j = 0; j <= m + 1; ++j)
        for (int i = 0; i <= l + 1; ++i) {
            A(4)[IDX(i, j, 0)] = A(4)[IDX(i, j, n)];
            A(6)[IDX(i, j, 0)] = A(6)[IDX(i, j, n)];
            A(7)[IDX(i, j, 0)] = A(7)[IDX(i, j, n)];
            A(13)[IDX(i, j, 0)] = A(13)[IDX(i, j, n)];
            A(18)[IDX(i, j, 0)] = A(18)[IDX(i, j, n)];
            A(2)[IDX(i, j, n1)] = A(2)[IDX(i, j, 1)];
            A(9)[IDX(i, j, n1)] = A(9)[IDX(i, j, 1)];
            A(11)[IDX(i, j, n1)] = A(11)[IDX(i, j, 1)];
            A(15)[IDX(i, j, n1)] = A(15)[IDX(i, j, 1)];
            A(16)[IDX(i, j, n1)] = A(16)[IDX(i, j, 1)];
        }
}

/* bcond_periodic: SRC/bcond.periodic.F90:67-266, faces x -> y -> z */
/* the y and z faces alone (what bcond_inflow runs after its x faces) */
void FN(oracle_faces_yz_)(R **a, int l, int m, int n, int nx, int ny)
{
    FN(per_y_)(a, l, m, n, nx, ny);
    FN(per_z_)(a, l, m, n, nx, ny);
}

void FN(oracle_bcond_periodic_)(R **a, int l, int m, int n, int nx, int ny)
{
    FN(per_x_)(a, l, m, n, nx, ny);
    FN(per_y_)(a, l, m, n, nx, ny);
    FN(per_z_)(a, l, m, n, nx, ny);
}

/* bcond_channel: SRC/bcond.channel.F90:78-281.  x and y periodic, no-slip
 * walls at z = 1 and z = n.  The wall loop uses the offload bounds
 * (i = 1..l, j = 1..m, :244-251); the reference's do-concurrent branch
 * (:253) runs over 0..l+1 / 0..m+1 and writes out of bounds (i-1 = -1),
 * so that branch cannot be the definition -- see the design notes. */
void FN(oracle_bcond_channel_)(R **a, int l, int m, int n, int nx, int ny)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const int n1 = n + 1;
    FN(per_x_)(a, l, m, n, nx, ny);
    FN(per_y_)(a, l, m, n, nx, ny);
    for (int j = 1; j <= m; ++j)
        for (int i = 1; i <= l; ++i) {
            A(6)[IDX(i, j, 0)] = A(15)[IDX(i, j, 1)];
            A(4)[IDX(i - 1, j, 0)] = A(11)[IDX(i, j, 1)];
            A(7)[IDX(i, j - 1, 0)] = A(16)[IDX(i, j, 1)];
            A(13)[IDX(i + 1, j, 0)] = A(2)[IDX(i, j, 1)];
            A(18)[IDX(i, j + 1, 0)] = A(9)[IDX(i, j, 1)];
            A(15)[IDX(i, j, n1)] = A(6)[IDX(i, j, n)];
            A(2)[IDX(i - 1, j, n1)] = A(13)[IDX(i, j, n)];
            A(9)[IDX(i, j - 1, n1)] = A(18)[IDX(i, j, n)];
            A(11)[IDX(i + 1, j, n1)] = A(4)[IDX(i, j, n)];
            A(16)[IDX(i, j + 1, n1)] = A(7)[IDX(i, j, n)];
        }
}

/* bcond_inflow: SRC/bcond.inflow.F90:91-328.  x=l+1 outflow equilibrium
 * (rho = 1, velocity = momentum at x = l, :131-162), x=0 inflow equilibrium
 * (rho = 1, u = (u_inflow,0,0), :165-181), then y and z periodic.
 * cte1 = 0 shifted, 1 unshifted (:68-72).  `3.0` is a default-real literal. */
void FN(oracle_bcond_inflow_)(R **a, int l, int m, int n, int nx, int ny, R u_inflow)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const int l1 = l + 1;
    const R uno = (R)1, zero = (R)0, rf = (R)3, qf = (R)1.5, three = (R)3.0, cte1 = g_noshift ? uno : zero;
    const R p1 = FN(P1_), p2 = FN(P2_);
    for (int k = 0; k <= n + 1; ++k)
        for (int j = 0; j <= m + 1; ++j) {
            const R crho = uno, rhoinv = uno;
            const size_t s = IDX(l, j, k);
            R xj = +(A(1)[s] + A(2)[s] + A(3)[s] + A(4)[s] + A(5)[s] - A(10)[s] - A(11)[s] - A(12)[s] -
                     A(13)[s] - A(14)[s]) *
                   rhoinv;
            R yj = +(A(3)[s] + A(7)[s] + A(8)[s] + A(9)[s] + A(12)[s] - A(1)[s] - A(10)[s] - A(16)[s] -
                     A(17)[s] - A(18)[s]) *
                   rhoinv;
            R zj = +(A(4)[s] + A(6)[s] + A(7)[s] + A(13)[s] + A(18)[s] - A(2)[s] - A(9)[s] - A(11)[s] -
                     A(15)[s] - A(16)[s]) *
                   rhoinv;
            R cvsq = xj * xj + yj * yj + zj * zj;
            const R cx10 = rf * (-xj - yj) + qf * (three * (xj + yj) * (xj + yj) - cvsq);
            const R cx11 = rf * (-xj - zj) + qf * (three * (xj + zj) * (xj + zj) - cvsq);
            const R cx12 = rf * (-xj + yj) + qf * (three * (xj - yj) * (xj - yj) - cvsq);
            const R cx13 = rf * (-xj + zj) + qf * (three * (xj - zj) * (xj - zj) - cvsq);
            const R cx14 = rf * (-xj) + qf * (three * (xj) * (xj)-cvsq);
            const size_t o = IDX(l1, j, k);
            A(10)[o] = crho * p2 * (cte1 + cx10);
            A(11)[o] = crho * p2 * (cte1 + cx11);
            A(12)[o] = crho * p2 * (cte1 + cx12);
            A(13)[o] = crho * p2 * (cte1 + cx13);
            A(14)[o] = crho * p1 * (cte1 + cx14);
            xj = u_inflow;
            yj = zero;
            zj = zero;
            cvsq = xj * xj + yj * yj + zj * zj;
            const R cx01 = rf * (xj - yj) + qf * (three * (xj - yj) * (xj - yj) - cvsq);
            const R cx02 = rf * (xj - zj) + qf * (three * (xj - zj) * (xj - zj) - cvsq);
            const R cx03 = rf * (xj + yj) + qf * (three * (xj + yj) * (xj + yj) - cvsq);
            const R cx04 = rf * (xj + zj) + qf * (three * (xj + zj) * (xj + zj) - cvsq);
            const R cx05 = rf * (xj) + qf * (three * (xj) * (xj)-cvsq);
            const size_t r = IDX(0, j, k);
            A(1)[r] = crho * p2 * (cte1 + cx01);
            A(2)[r] = crho * p2 * (cte1 + cx02);
            A(3)[r] = crho * p2 * (cte1 + cx03);
            A(4)[r] = crho * p2 * (cte1 + cx04);
            A(5)[r] = crho * p1 * (cte1 + cx05);
        }
    FN(per_y_)(a, l, m, n, nx, ny);
    FN(per_z_)(a, l, m, n, nx, ny);
}

/* bcond_obs: SRC/bcond.obs.F90:101-122, executed sequentially with i
 * fastest and k slowest, the order a serial build of the do-concurrent loop
 * runs in.  Reads of neighbouring solid cells therefore see values already
 * overwritten in this pass when the neighbour precedes the cell. */
void FN(oracle_bcond_obs_)(R **a, const int *obs, int l, int m, int n, int nx, int ny, int imin, int imax,
                           int jmin, int jmax, int kmin, int kmax)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    (void)n;
    for (int k = kmin; k <= kmax; ++k)
        for (int j = jmin; j <= jmax; ++j)
            for (int i = imin; i <= imax; ++i) {
                const size_t o = (size_t)(i - 1) + (size_t)l * (size_t)(j - 1) + (size_t)l * (size_t)m * (size_t)(k - 1);
                if (obs[o] != 1) continue;
                const size_t c = IDX(i, j, k);
                A(1)[c] = A(12)[IDX(i + 1, j - 1, k)];
                A(2)[c] = A(13)[IDX(i + 1, j, k - 1)];
                A(3)[c] = A(10)[IDX(i + 1, j + 1, k)];
                A(4)[c] = A(11)[IDX(i + 1, j, k + 1)];
                A(5)[c] = A(14)[IDX(i + 1, j, k)];
                A(6)[c] = A(15)[IDX(i, j, k + 1)];
                A(7)[c] = A(16)[IDX(i, j + 1, k + 1)];
                A(8)[c] = A(17)[IDX(i, j + 1, k)];
                A(9)[c] = A(18)[IDX(i, j + 1, k - 1)];
                A(10)[c] = A(3)[IDX(i - 1, j - 1, k)];
                A(11)[c] = A(4)[IDX(i - 1, j, k - 1)];
                A(12)[c] = A(1)[IDX(i - 1, j + 1, k)];
                A(13)[c] = A(2)[IDX(i - 1, j, k + 1)];
                A(14)[c] = A(5)[IDX(i - 1, j, k)];
                A(15)[c] = A(6)[IDX(i, j, k - 1)];
                A(16)[c] = A(7)[IDX(i, j - 1, k - 1)];
                A(17)[c] = A(8)[IDX(i, j - 1, k)];
                A(18)[c] = A(9)[IDX(i, j - 1, k + 1)];
            }
}

/* diagno: SRC/diagno.F90:86-138.  Means of rho, u, v, w and 0.5|u|^2 over
 * interior cells, accumulated in double in (k, j, i) order.  out[5] =
 * {<u>, <v>, <w>, <rho>, <ke>} in the order diagno.dat writes them (:145). */
void FN(oracle_diagno_)(R **a, int l, int m, int n, int nx, int ny, double *out)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const R uno = (R)1, cte1 = g_noshift ? (R)0 : uno;
    double rtot = 0, xtot = 0, ytot = 0, ztot = 0, stot = 0;
    for (int k = 1; k <= n; ++k)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= l; ++i) {
                const size_t s = IDX(i, j, k);
                const R x01 = A(1)[s], x02 = A(2)[s], x03 = A(3)[s], x04 = A(4)[s], x05 = A(5)[s];
                const R x06 = A(6)[s], x07 = A(7)[s], x08 = A(8)[s], x09 = A(9)[s], x10 = A(10)[s];
                const R x11 = A(11)[s], x12 = A(12)[s], x13 = A(13)[s], x14 = A(14)[s], x15 = A(15)[s];
                const R x16 = A(16)[s], x17 = A(17)[s], x18 = A(18)[s], x19 = A(19)[s];
                const R rho = x01 + x02 + x03 + x04 + x05 + x06 + x07 + x08 + x09 + x10 + x11 + x12 +
                              x13 + x14 + x15 + x16 + x17 + x18 + x19 + cte1;
                const R rhoinv = uno / rho;
                const R xj = (x01 + x02 + x03 + x04 + x05 - x10 - x11 - x12 - x13 - x14) * rhoinv;
                const R yj = (x03 + x07 + x08 + x09 + x12 - x01 - x10 - x16 - x17 - x18) * rhoinv;
                const R zj = (x04 + x06 + x07 + x13 + x18 - x02 - x09 - x11 - x15 - x16) * rhoinv;
                rtot = rtot + (double)rho;
                xtot = xtot + (double)xj;
                ytot = ytot + (double)yj;
                ztot = ztot + (double)zj;
                /* 0.5*(...) : default-real 0.5 times a kind-R value */
                stot = stot + (double)((R)0.5 * (xj * xj + yj * yj + zj * zj));
            }
    /* (x/float(l))/float(m)/float(n) */
    const double fl = (double)(float)l, fm = (double)(float)m, fn = (double)(float)n;
    out[0] = (xtot / fl) / fm / fn;
    out[1] = (ytot / fl) / fm / fn;
    out[2] = (ztot / fl) / fm / fn;
    out[3] = (rtot / fl) / fm / fn;
    out[4] = (stot / fl) / fm / fn;
}

/* varm: SRC/varm.F90:56-99, with its indexing: the u (x-velocity), w
 * (y-velocity) and v (z-velocity) sums are binned by k, the density sum by
 * j; all accumulate in kind R in (k, j, i) order.  The routine prints row
 * j = 1..m as (u(j), w(j), den(j)) * (1/l); this function returns the raw
 * per-bin sums times rvol: u_k[n], w_k[n], v_k[n], den_j[m].  (The
 * reference never zeroes v; here it starts at zero.) */
void FN(oracle_varm_)(R **a, int l, int m, int n, int nx, int ny, R *u_k, R *w_k, R *v_k, R *den_j)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const R uno = (R)1, zero = (R)0, cte1 = g_noshift ? zero : uno;
    for (int k = 0; k < n; ++k) u_k[k] = w_k[k] = v_k[k] = zero;
    for (int j = 0; j < m; ++j) den_j[j] = zero;
    for (int k = 1; k <= n; ++k)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= l; ++i) {
                const size_t s = IDX(i, j, k);
                const R rho = +A(1)[s] + A(2)[s] + A(3)[s] + A(4)[s] + A(5)[s] + A(6)[s] + A(7)[s] + A(8)[s] +
                              A(9)[s] + A(10)[s] + A(11)[s] + A(12)[s] + A(13)[s] + A(14)[s] + A(15)[s] +
                              A(16)[s] + A(17)[s] + A(18)[s] + A(19)[s] + cte1;
                const R rhoinv = uno / rho;
                den_j[j - 1] = den_j[j - 1] + rho;
                u_k[k - 1] = u_k[k - 1] + (A(1)[s] + A(2)[s] + A(3)[s] + A(4)[s] + A(5)[s] - A(10)[s] -
                                           A(11)[s] - A(12)[s] - A(13)[s] - A(14)[s]) *
                                              rhoinv;
                w_k[k - 1] = w_k[k - 1] + (A(3)[s] + A(7)[s] + A(8)[s] + A(9)[s] + A(12)[s] - A(1)[s] -
                                           A(10)[s] - A(16)[s] - A(17)[s] - A(18)[s]) *
                                              rhoinv;
                v_k[k - 1] = v_k[k - 1] + (A(4)[s] + A(6)[s] + A(7)[s] + A(13)[s] + A(18)[s] - A(2)[s] -
                                           A(9)[s] - A(11)[s] - A(15)[s] - A(16)[s]) *
                                              rhoinv;
            }
    /* rvol = 1.0/(float(l)) : default real */
    const R rvol = (R)(1.0f / (float)l);
    for (int k = 0; k < n; ++k) {
        u_k[k] = u_k[k] * rvol;
        w_k[k] = w_k[k] * rvol;
        v_k[k] = v_k[k] * rvol;
    }
    for (int j = 0; j < m; ++j) den_j[j] = den_j[j] * rvol;
}

/* Line profiles (SRC/prof_i.F90:47-78, SRC/prof_j.F90:47-77,
 * SRC/prof_k.F90:47-79) and the point probe (SRC/probe.F90:44-75): the
 * three routines form, per cell, den = a01+...+a19+cte1 and the momentum
 * sums in the same order and divide each by den; probe does the same at one
 * point (xj/rho, yj/rho, zj/rho, rho).  Line along axis 0 (i = 1..l at
 * (c1, c2) = (j, k)), 1 (j = 1..m at (i, k)) or 2 (k = 1..n at (i, j)).
 * out[4*(t-1) + {0,1,2,3}] = u, w (y-velocity), v (z-velocity), den. */
static void FN(line_point_)(R **a, size_t s, R *o)
{
    const R uno = (R)1, cte1 = g_noshift ? (R)0 : uno;
    const R den = +A(1)[s] + A(2)[s] + A(3)[s] + A(4)[s] + A(5)[s] + A(6)[s] + A(7)[s] + A(8)[s] + A(9)[s] +
                  A(10)[s] + A(11)[s] + A(12)[s] + A(13)[s] + A(14)[s] + A(15)[s] + A(16)[s] + A(17)[s] +
                  A(18)[s] + A(19)[s] + cte1;
    const R u = +(A(1)[s] + A(2)[s] + A(3)[s] + A(4)[s] + A(5)[s] - A(10)[s] - A(11)[s] - A(12)[s] -
                  A(13)[s] - A(14)[s]) /
                den;
    const R w = +(A(3)[s] + A(7)[s] + A(8)[s] + A(9)[s] + A(12)[s] - A(1)[s] - A(10)[s] - A(16)[s] -
                  A(17)[s] - A(18)[s]) /
                den;
    const R v = +(A(4)[s] + A(6)[s] + A(7)[s] + A(13)[s] + A(18)[s] - A(2)[s] - A(9)[s] - A(11)[s] -
                  A(15)[s] - A(16)[s]) /
                den;
    o[0] = u;
    o[1] = w;
    o[2] = v;
    o[3] = den;
}

void FN(oracle_prof_line_)(R **a, int l, int m, int n, int nx, int ny, int axis, int c1, int c2, R *out)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    const int len = axis == 0 ? l : (axis == 1 ? m : n);
    for (int t = 1; t <= len; ++t) {
        const size_t s = axis == 0 ? IDX(t, c1, c2) : (axis == 1 ? IDX(c1, t, c2) : IDX(c1, c2, t));
        FN(line_point_)(a, s, out + 4 * (t - 1));
    }
}

void FN(oracle_prof_k_)(R **a, int l, int m, int n, int nx, int ny, int ic, int jc, R *out)
{
    FN(oracle_prof_line_)(a, l, m, n, nx, ny, 2, ic, jc, out);
}

void FN(oracle_probe_)(R **a, int nx, int ny, int i0, int j0, int k0, R *out)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    FN(line_point_)(a, IDX(i0, j0, k0), out);
}

/* draglift: SRC/draglift.F90:48-179.  Momentum exchange over fluid cells of
 * the sphere's bounding box, accumulated in kind R in (k, j, i) order, one
 * pass per component as in the reference.  out = {Fx, Fy, Fz} * norm with
 * norm = 1/(u_inflow^2 * radius).  obs is (1:l,1:m,1:n); the box must keep
 * every obs(x + c) read inside the array. */
void FN(oracle_draglift_)(R **a, const int *obs, int l, int m, int n, int nx, int ny, int radius,
                          R u_inflow, R *out)
{
    const size_t nxy = (size_t)nx * (size_t)ny;
    static const int cx[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
    static const int cy[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};
    static const int cz[19] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0};
    const int icoord = 2 * l / 5, jcoord = m / 2, kcoord = n / 2, delta = 3;
    const int is = icoord - radius - delta, ie = icoord + radius + delta;
    const int js = jcoord - radius - delta, je = jcoord + radius + delta;
    const int ks = kcoord - radius - delta, ke = kcoord + radius + delta;
#define OBS(i, j, k) obs[(size_t)((i)-1) + (size_t)l * (size_t)((j)-1) + (size_t)l * (size_t)m * (size_t)((k)-1)]
    for (int comp = 0; comp < 3; ++comp) {
        const int *cc = comp == 0 ? cx : (comp == 1 ? cy : cz);
        R tot = (R)0;
        for (int k = ks; k <= ke; ++k)
            for (int j = js; j <= je; ++j)
                for (int i = is; i <= ie; ++i) {
                    if (OBS(i, j, k) != 0) continue;
                    const size_t s = IDX(i, j, k);
                    R f[18];
                    for (int d = 0; d < 18; ++d)
                        /* 2.0*c(d)*a(d)(i,j,k)*obs(i+icx,j+icy,k+icz) */
                        f[d] = (R)2.0 * (R)cc[d] * a[d][s] * (R)OBS(i + cx[d], j + cy[d], k + cz[d]);
                    R sum = f[0];
                    for (int d = 1; d < 18; ++d) sum = sum + f[d];
                    tot = tot + sum;
                }
        out[comp] = tot;
    }
#undef OBS
    const R norm = (R)1 / (u_inflow * u_inflow * (R)radius);
    out[0] = out[0] * norm;
    out[1] = out[1] * norm;
    out[2] = out[2] * norm;
}

#undef A
#undef IDX
#undef FN
#undef CAT
#undef CAT_
