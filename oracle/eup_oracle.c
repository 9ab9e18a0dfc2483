/*
 * CPU oracle for the D3Q19 BGK hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of the reference's serial time step, used by tests/, by the smoke
 * check and by bench.py's CPU-baseline leg as the checker.  It is never linked into, called by or
 * shipped with the HIP library.  Compile with -ffp-contract=off: every expression below is written
 * in the reference's evaluation order so that the result is bit-identical to the reference built
 * with the same flag.
 *
 * State layout: one contiguous array of 19 populations, q-major; population q (1-based) occupies
 * a[(q-1)*np .. q*np-1] with np = (l+2)(m+2)(n+2) and site (i,j,k), i in 0..l+1 etc., at offset
 * i + (l+2)*(j + (m+2)*k) -- the reference's ghosted Fortran arrays a01..a19 laid end to end
 * (CODE/SRC/alloca.F90:94-133).  obs is int32 (l,m,n), i fastest, 1 = solid.
 *
 * Functions and the reference code they restate:
 *   orc_parity_steps   CODE/SRC/bgk3d.F90:114-138 (order), bcond.obs.F90:47-77, bcond.bc.F90:127-205
 *                      (outflow + border fix), :260-321 (inflow), :385-409/:472-496 (z periodic),
 *                      :535-560/:600-625 (y periodic incl. corners), movef.F90:43-68, col.F90:94-229
 *   orc_flux_x/_y      CODE/SRC/flux_X.F90:50-301, flux_Y.F90:51-282 (SERIAL offsets = 0)
 *   orc_probe          CODE/SRC/probe_global.F90:41-77
 *   orc_diagno         CODE/SRC/diagno.F90:53-131
 *   orc_physics_steps  the same collision and boundary arithmetic with y/z periodicity applied by
 *                      index wrap (ghosts fresh before bounce-back), solid sites not collided.
 *   orc_mem_force      momentum-exchange force on the solid (no reference counterpart).
 *   orc_grey_*         grey-lattice (partial bounce-back) step, force, objective and its discrete
 *                      adjoint with checkpoint-recompute (no reference counterpart).
 *
 * Parity pins (see tests/test_oracle_*.py): bit-equality with the reference executable built by
 * oracle/Makefile.ref on seeded inputs; closed forms (equilibrium moments, rest / uniform-flow
 * fixed points, index-coded streaming permutation, single-solid-site bounce-back).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* weights as the reference computes them (quad precision, then rounded once to double;
 * CODE/SRC/storage.F90:55-82).  1/3, 1/18 and 1/36 are the correctly rounded doubles. */
static const double P0 = 1.0 / 3.0;
static const double P1 = 1.0 / 18.0;
static const double P2 = 1.0 / 36.0;
static const double RF = 3.0;
static const double QF = 1.5;

static const int CX[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
static const int CY[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};
static const int CZ[19] = {0, -1, 0, 1, 0, 1, 1, 0, -1, 0, -1, 0, 1, 0, -1, -1, 0, 1, 0};
static const int OPP[19] = {11, 12, 9, 10, 13, 14, 15, 16, 17, 2, 3, 0, 1, 4, 5, 6, 7, 8, 18};

typedef struct {
    int l, m, n;
    long lp, mp, np; /* l+2, m+2, total sites */
} geom_t;

static geom_t mkgeom(int l, int m, int n)
{
    geom_t g;
    g.l = l; g.m = m; g.n = n;
    g.lp = l + 2; g.mp = m + 2;
    g.np = (long)(l + 2) * (m + 2) * (n + 2);
    return g;
}

#define IDX(g, i, j, k) ((long)(i) + (g).lp * ((long)(j) + (g).mp * (long)(k)))
#define A(q, i, j, k) a[(long)((q) - 1) * g.np + IDX(g, i, j, k)]
#define B(q, i, j, k) b[(long)((q) - 1) * g.np + IDX(g, i, j, k)]
#define OBS(i, j, k) obs[((long)(i) - 1) + (long)g.l * (((long)(j) - 1) + (long)g.m * ((long)(k) - 1))]

/* ------------------------------------------------------------------------------------------ */
/* BGK collision of one site: x[19] pre-collision -> out[19]; reference col.F90:127-229       */
/* ------------------------------------------------------------------------------------------ */
static void collide_site(const double *x, double *out, double omega, double fgrad)
{
    const double omega1 = 1.0 - omega;
    double rho = x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7]
               + x[8] + x[9] + x[10] + x[11] + x[12] + x[13] + x[14] + x[15]
               + x[16] + x[17] + x[18];
    double rhoinv = 1.0 / rho;
    double vx = (x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
    double vy = (x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
    double vz = (x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;
    double vx2 = vx * vx, vy2 = vy * vy, vz2 = vz * vz;
    double vsq = vx2 + vy2 + vz2;
    double vxpy = vx + vy;
    double qxpy = 1.0 + QF * (3.0 * vxpy * vxpy - vsq);
    double vxmy = vx - vy;
    double qxmy = 1.0 + QF * (3.0 * vxmy * vxmy - vsq);
    double vxpz = vx + vz;
    double qxpz = 1.0 + QF * (3.0 * vxpz * vxpz - vsq);
    double vxmz = vx - vz;
    double qxmz = 1.0 + QF * (3.0 * vxmz * vxmz - vsq);
    double vypz = vy + vz;
    double qypz = 1.0 + QF * (3.0 * vypz * vypz - vsq);
    double vymz = vy - vz;
    double qymz = 1.0 + QF * (3.0 * vymz * vymz - vsq);
    double qx = 1.0 + QF * (3.0 * vx2 - vsq);
    double qy = 1.0 + QF * (3.0 * vy2 - vsq);
    double qz = 1.0 + QF * (3.0 * vz2 - vsq);
    vx = RF * vx; vy = RF * vy; vz = RF * vz;
    vxpy = RF * vxpy; vxmy = RF * vxmy; vxpz = RF * vxpz;
    vxmz = RF * vxmz; vypz = RF * vypz; vymz = RF * vymz;
    double rp1 = rho * P1;
    double rp2 = rho * P2;
    /* FORCING_X (default): force along x only (col.F90:173-187) */
    double fx = fgrad * rho;
    const double fy = 0.0, fz = 0.0;
    double e[19];
    e[0] = rp2 * (vxmy + qxmy);
    e[1] = rp2 * (vxmz + qxmz);
    e[2] = rp2 * (vxpy + qxpy);
    e[3] = rp2 * (vxpz + qxpz);
    e[4] = rp1 * (vx + qx);
    e[5] = rp1 * (vz + qz);
    e[6] = rp2 * (vypz + qypz);
    e[7] = rp1 * (vy + qy);
    e[8] = rp2 * (vymz + qymz);
    e[9] = rp2 * (-vxpy + qxpy);
    e[10] = rp2 * (-vxpz + qxpz);
    e[11] = rp2 * (-vxmy + qxmy);
    e[12] = rp2 * (-vxmz + qxmz);
    e[13] = rp1 * (-vx + qx);
    e[14] = rp1 * (-vz + qz);
    e[15] = rp2 * (-vypz + qypz);
    e[16] = rp1 * (-vy + qy);
    e[17] = rp2 * (-vymz + qymz);
    e[18] = rho * P0 * (1.0 - QF * vsq);
    out[0] = omega1 * x[0] + (omega * e[0] + fx - fy);
    out[1] = omega1 * x[1] + (omega * e[1] + fx - fz);
    out[2] = omega1 * x[2] + (omega * e[2] + fx + fy);
    out[3] = omega1 * x[3] + (omega * e[3] + fx + fz);
    out[4] = omega1 * x[4] + (omega * e[4] + fx);
    out[5] = omega1 * x[5] + (omega * e[5] + fz);
    out[6] = omega1 * x[6] + (omega * e[6] + fy + fz);
    out[7] = omega1 * x[7] + (omega * e[7] + fy);
    out[8] = omega1 * x[8] + (omega * e[8] + fy - fz);
    out[9] = omega1 * x[9] + (omega * e[9] - fx - fy);
    out[10] = omega1 * x[10] + (omega * e[10] - fx - fz);
    out[11] = omega1 * x[11] + (omega * e[11] - fx + fy);
    out[12] = omega1 * x[12] + (omega * e[12] - fx + fz);
    out[13] = omega1 * x[13] + (omega * e[13] - fx);
    out[14] = omega1 * x[14] + (omega * e[14] - fz);
    out[15] = omega1 * x[15] + (omega * e[15] - fy - fz);
    out[16] = omega1 * x[16] + (omega * e[16] - fy);
    out[17] = omega1 * x[17] + (omega * e[17] - fy + fz);
    out[18] = omega1 * x[18] + (omega * e[18]);
}

/* outlet ghost values for q = 10..14 from the (not rho-divided) momentum of an interior site;
 * reference bcond.bc.F90:138-164 */
static void outlet_eq(const double *x, double *o /* o[0..4] = q10..q14 */)
{
    double xj = +(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]);
    double yj = +(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]);
    double zj = +(x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]);
    double cvsq = xj * xj + yj * yj + zj * zj;
    double c10 = RF * (-xj - yj) + QF * (3.0 * (xj + yj) * (xj + yj) - cvsq);
    double c11 = RF * (-xj - zj) + QF * (3.0 * (xj + zj) * (xj + zj) - cvsq);
    double c12 = RF * (-xj + yj) + QF * (3.0 * (xj - yj) * (xj - yj) - cvsq);
    double c13 = RF * (-xj + zj) + QF * (3.0 * (xj - zj) * (xj - zj) - cvsq);
    double c14 = RF * (-xj) + QF * (3.0 * (xj) * (xj)-cvsq);
    o[0] = P2 * (1.0 + c10);
    o[1] = P2 * (1.0 + c11);
    o[2] = P2 * (1.0 + c12);
    o[3] = P2 * (1.0 + c13);
    o[4] = P1 * (1.0 + c14);
}

/* inlet ghost values for q = 1..5 at u = (u_in, 0, 0); reference bcond.bc.F90:260-321 */
static void inlet_eq(double u_in, double *o /* o[0..4] = q1..q5 */)
{
    double xj = u_in, yj = 0.0, zj = 0.0;
    double cvsq = xj * xj + yj * yj + zj * zj;
    double c1 = RF * (xj - yj) + QF * (3.0 * (xj - yj) * (xj - yj) - cvsq);
    double c2 = RF * (xj - zj) + QF * (3.0 * (xj - zj) * (xj - zj) - cvsq);
    double c3 = RF * (xj + yj) + QF * (3.0 * (xj + yj) * (xj + yj) - cvsq);
    double c4 = RF * (xj + zj) + QF * (3.0 * (xj + zj) * (xj + zj) - cvsq);
    double c5 = RF * (xj) + QF * (3.0 * (xj) * (xj)-cvsq);
    o[0] = P2 * (1.0 + c1);
    o[1] = P2 * (1.0 + c2);
    o[2] = P2 * (1.0 + c3);
    o[3] = P2 * (1.0 + c4);
    o[4] = P1 * (1.0 + c5);
}

static void gather(const double *a, const geom_t g, long s, double *x)
{
    for (int q = 0; q < 19; ++q) x[q] = a[q * g.np + s];
}

/* ------------------------------------------------------------------------------------------ */
/* parity mode: the serial reference step, bit for bit                                        */
/* ------------------------------------------------------------------------------------------ */
static void bcond_obs(double *a, const int32_t *obs, geom_t g, const int bb[6])
{
    for (int k = bb[4]; k <= bb[5]; ++k)
        for (int j = bb[2]; j <= bb[3]; ++j)
            for (int i = bb[0]; i <= bb[1]; ++i) {
                if (OBS(i, j, k) != 1) continue;
                /* sequential in-place order as bcond.obs.F90:56-73 */
                for (int q = 0; q < 18; ++q) {
                    int p = OPP[q];
                    A(q + 1, i, j, k) = A(p + 1, i + CX[q], j + CY[q], k + CZ[q]);
                }
            }
}

static void bcond_bc(double *a, geom_t g, double u_in, int outflow, int inflow)
{
    const int l = g.l, m = g.m, n = g.n, l1 = l + 1, m1 = m + 1, n1 = n + 1;
    double x[19], o[5];
    if (outflow) {
        for (int k = 1; k <= n; ++k)
            for (int j = 1; j <= m; ++j) {
                gather(a, g, IDX(g, l, j, k), x);
                outlet_eq(x, o);
                for (int t = 0; t < 5; ++t) A(10 + t, l1, j, k) = o[t];
            }
        /* border fix (bcond.bc.F90:171-204) */
        for (int j = 1; j <= m; ++j)
            for (int t = 0; t < 5; ++t) {
                A(10 + t, l1, j, 0) = A(10 + t, l1, j, 1);
                A(10 + t, l1, j, n1) = A(10 + t, l1, j, n);
            }
        for (int k = 1; k <= n; ++k)
            for (int t = 0; t < 5; ++t) {
                A(10 + t, l1, 0, k) = A(10 + t, l1, 1, k);
                A(10 + t, l1, m1, k) = A(10 + t, l1, m, k);
            }
    }
    if (inflow) {
        inlet_eq(u_in, o);
        for (int k = 0; k <= n1; ++k)
            for (int j = 0; j <= m1; ++j)
                for (int t = 0; t < 5; ++t) A(1 + t, 0, j, k) = o[t];
    }
    /* up (z = n+1) periodic, interior range */
    for (int j = 1; j <= m; ++j)
        for (int i = 1; i <= l; ++i) {
            A(2, i, j, n1) = A(2, i, j, 1);
            A(9, i, j, n1) = A(9, i, j, 1);
            A(11, i, j, n1) = A(11, i, j, 1);
            A(15, i, j, n1) = A(15, i, j, 1);
            A(16, i, j, n1) = A(16, i, j, 1);
        }
    /* down (z = 0) periodic */
    for (int j = 1; j <= m; ++j)
        for (int i = 1; i <= l; ++i) {
            A(4, i, j, 0) = A(4, i, j, n);
            A(6, i, j, 0) = A(6, i, j, n);
            A(7, i, j, 0) = A(7, i, j, n);
            A(13, i, j, 0) = A(13, i, j, n);
            A(18, i, j, 0) = A(18, i, j, n);
        }
    /* left (y = 0) periodic with the SERIAL corner copy */
    for (int k = 1; k <= n; ++k) {
        A(3, 0, 0, k) = A(3, l, m, k);
        for (int i = 1; i <= l; ++i) {
            A(3, i, 0, k) = A(3, i, m, k);
            A(7, i, 0, k) = A(7, i, m, k);
            A(8, i, 0, k) = A(8, i, m, k);
            A(9, i, 0, k) = A(9, i, m, k);
            A(12, i, 0, k) = A(12, i, m, k);
        }
    }
    /* right (y = m+1) periodic with the SERIAL corner copy */
    for (int k = 1; k <= n; ++k) {
        A(10, l1, m1, k) = A(10, 1, 1, k);
        for (int i = 1; i <= l; ++i) {
            A(1, i, m1, k) = A(1, i, 1, k);
            A(10, i, m1, k) = A(10, i, 1, k);
            A(16, i, m1, k) = A(16, i, 1, k);
            A(17, i, m1, k) = A(17, i, 1, k);
            A(18, i, m1, k) = A(18, i, 1, k);
        }
    }
}

static void movef(const double *a, double *b, geom_t g)
{
    for (int k = 1; k <= g.n; ++k)
        for (int j = 1; j <= g.m; ++j)
            for (int i = 1; i <= g.l; ++i)
                for (int q = 0; q < 18; ++q)
                    B(q + 1, i, j, k) = A(q + 1, i - CX[q], j - CY[q], k - CZ[q]);
}

static void col(double *a, const double *b, geom_t g, double omega, double fgrad)
{
    double x[19], out[19];
    for (int k = 1; k <= g.n; ++k)
        for (int j = 1; j <= g.m; ++j)
            for (int i = 1; i <= g.l; ++i) {
                long s = IDX(g, i, j, k);
                for (int q = 0; q < 18; ++q) x[q] = b[q * g.np + s];
                x[18] = a[18 * g.np + s];
                collide_site(x, out, omega, fgrad);
                for (int q = 0; q < 19; ++q) a[q * g.np + s] = out[q];
            }
}

static void compute_bbox(const int32_t *obs, geom_t g, int bb[6])
{
    /* initial values as build.obs.F90:73-79 */
    bb[0] = g.l; bb[1] = 0; bb[2] = g.m; bb[3] = 0; bb[4] = g.n; bb[5] = 0;
    for (int k = 1; k <= g.n; ++k)
        for (int j = 1; j <= g.m; ++j)
            for (int i = 1; i <= g.l; ++i)
                if (OBS(i, j, k) == 1) {
                    if (i < bb[0]) bb[0] = i;
                    if (i > bb[1]) bb[1] = i;
                    if (j < bb[2]) bb[2] = j;
                    if (j > bb[3]) bb[3] = j;
                    if (k < bb[4]) bb[4] = k;
                    if (k > bb[5]) bb[5] = k;
                }
}

/* a: state (ghosted, 19*np); b: scratch of the same size; returns 0 */
int orc_parity_steps(double *a, double *b, const int32_t *obs, int l, int m, int n, double omega,
                     double fgrad, double u_inflow, int nsteps)
{
    geom_t g = mkgeom(l, m, n);
    int bb[6];
    compute_bbox(obs, g, bb);
    for (int t = 0; t < nsteps; ++t) {
        bcond_obs(a, obs, g, bb);
        bcond_bc(a, g, u_inflow, 1, 1);
        movef(a, b, g);
        col(a, b, g, omega, fgrad);
    }
    return 0;
}

/* individual phases, for operator-level tests */
void orc_phase_bcond_obs(double *a, const int32_t *obs, int l, int m, int n)
{
    geom_t g = mkgeom(l, m, n);
    int bb[6];
    compute_bbox(obs, g, bb);
    bcond_obs(a, obs, g, bb);
}
void orc_phase_bcond_bc(double *a, int l, int m, int n, double u_inflow)
{
    bcond_bc(a, mkgeom(l, m, n), u_inflow, 1, 1);
}
void orc_phase_movef(const double *a, double *b, int l, int m, int n)
{
    movef(a, b, mkgeom(l, m, n));
}
void orc_phase_col(double *a, const double *b, int l, int m, int n, double omega, double fgrad)
{
    col(a, b, mkgeom(l, m, n), omega, fgrad);
}
void orc_collide_site(const double *x, double *out, double omega, double fgrad)
{
    collide_site(x, out, omega, fgrad);
}

/* ------------------------------------------------------------------------------------------ */
/* physics mode: periodic y/z by index wrap, inlet/outlet as above, bounce-back from the site's */
/* own opposite population, solid sites not updated                                            */
/* ------------------------------------------------------------------------------------------ */
static inline int wrap(int v, int n) { return v < 1 ? v + n : (v > n ? v - n : v); }

/* a: state (interior used), b: scratch; inflow/outflow: 1 = inlet at x=0 / outlet at x=l+1,
 * 0 = periodic in x. */
int orc_physics_steps(double *a, double *b, const int32_t *obs, int l, int m, int n, double omega,
                      double fgrad, double u_inflow, int x_open, int nsteps)
{
    geom_t g = mkgeom(l, m, n);
    double inl[5], x[19], out[19], y[19], o[5];
    inlet_eq(u_inflow, inl);
    for (int t = 0; t < nsteps; ++t) {
        for (int k = 1; k <= n; ++k)
            for (int j = 1; j <= m; ++j)
                for (int i = 1; i <= l; ++i) {
                    long s = IDX(g, i, j, k);
                    if (OBS(i, j, k) == 1) {
                        for (int q = 0; q < 19; ++q) b[q * g.np + s] = a[q * g.np + s];
                        continue;
                    }
                    for (int q = 0; q < 18; ++q) {
                        int si = i - CX[q], sj = wrap(j - CY[q], m), sk = wrap(k - CZ[q], n);
                        if (si < 1 || si > l) {
                            if (!x_open) {
                                si = wrap(si, l);
                            } else if (si == 0) {
                                x[q] = inl[q];              /* q = 1..5 */
                                continue;
                            } else {
                                gather(a, g, IDX(g, l, sj, sk), y);
                                outlet_eq(y, o);
                                x[q] = o[q - 9];            /* q = 10..14 */
                                continue;
                            }
                        }
                        if (OBS(si, sj, sk) == 1)
                            x[q] = a[OPP[q] * g.np + s];
                        else
                            x[q] = a[q * g.np + IDX(g, si, sj, sk)];
                    }
                    x[18] = a[18 * g.np + s];
                    collide_site(x, out, omega, fgrad);
                    for (int q = 0; q < 19; ++q) b[q * g.np + s] = out[q];
                }
        for (int q = 0; q < 19; ++q)
            for (int k = 1; k <= n; ++k)
                for (int j = 1; j <= m; ++j)
                    memcpy(&a[q * g.np + IDX(g, 1, j, k)], &b[q * g.np + IDX(g, 1, j, k)],
                           sizeof(double) * (size_t)l);
    }
    return 0;
}

/* momentum-exchange force on the solid from the post-collision state (physics semantics):
 * F = sum over fluid sites x and directions p with x + c_p solid of 2 c_p f_p(x). */
void orc_mem_force(const double *a, const int32_t *obs, int l, int m, int n, int x_open, double F[3])
{
    geom_t g = mkgeom(l, m, n);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int k = 1; k <= n; ++k)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= l; ++i) {
                if (OBS(i, j, k) == 1) continue;
                long s = IDX(g, i, j, k);
                for (int p = 0; p < 18; ++p) {
                    int ti = i + CX[p], tj = wrap(j + CY[p], m), tk = wrap(k + CZ[p], n);
                    if (ti < 1 || ti > l) {
                        if (x_open) continue;
                        ti = wrap(ti, l);
                    }
                    if (OBS(ti, tj, tk) != 1) continue;
                    double f2 = 2.0 * a[p * g.np + s];
                    fx += CX[p] * f2;
                    fy += CY[p] * f2;
                    fz += CZ[p] * f2;
                }
            }
    F[0] = fx; F[1] = fy; F[2] = fz;
}

/* ------------------------------------------------------------------------------------------ */
/* diagnostics                                                                                 */
/* ------------------------------------------------------------------------------------------ */
static const double CXD[19] = {1, 1, 1, 1, 1, 0, 0, 0, 0, -1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
static const double CYD[19] = {-1, 0, 1, 0, 0, 0, 1, 1, 1, -1, 0, 1, 0, 0, 0, -1, -1, -1, 0};

static void site_fneq(const double *x, double *fn, double *rho_out)
{
    /* flux_X.F90:173-242 */
    double rho = +x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9] + x[10]
                 + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18];
    double rhoinv = 1.0 / rho;
    double xj = +(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
    double yj = +(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
    double zj = +(x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;
    double cvsq = xj * xj + yj * yj + zj * zj;
    double fe[18];
    fe[0] = rho * P2 * (1.0 + RF * (xj - yj) + QF * (3.0 * (xj - yj) * (xj - yj) - cvsq));
    fe[1] = rho * P2 * (1.0 + RF * (xj - zj) + QF * (3.0 * (xj - zj) * (xj - zj) - cvsq));
    fe[2] = rho * P2 * (1.0 + RF * (xj + yj) + QF * (3.0 * (xj + yj) * (xj + yj) - cvsq));
    fe[3] = rho * P2 * (1.0 + RF * (xj + zj) + QF * (3.0 * (xj + zj) * (xj + zj) - cvsq));
    fe[4] = rho * P1 * (1.0 + RF * (xj) + QF * (3.0 * (xj) * (xj)-cvsq));
    fe[5] = rho * P1 * (1.0 + RF * (zj) + QF * (3.0 * (zj) * (zj)-cvsq));
    fe[6] = rho * P2 * (1.0 + RF * (yj + zj) + QF * (3.0 * (yj + zj) * (yj + zj) - cvsq));
    fe[7] = rho * P1 * (1.0 + RF * (yj) + QF * (3.0 * (yj) * (yj)-cvsq));
    fe[8] = rho * P2 * (1.0 + RF * (yj - zj) + QF * (3.0 * (yj - zj) * (yj - zj) - cvsq));
    fe[9] = rho * P2 * (1.0 + RF * (-xj - yj) + QF * (3.0 * (xj + yj) * (xj + yj) - cvsq));
    fe[10] = rho * P2 * (1.0 + RF * (-xj - zj) + QF * (3.0 * (xj + zj) * (xj + zj) - cvsq));
    fe[11] = rho * P2 * (1.0 + RF * (-xj + yj) + QF * (3.0 * (xj - yj) * (xj - yj) - cvsq));
    fe[12] = rho * P2 * (1.0 + RF * (-xj + zj) + QF * (3.0 * (xj - zj) * (xj - zj) - cvsq));
    fe[13] = rho * P1 * (1.0 + RF * (-xj) + QF * (3.0 * (xj) * (xj)-cvsq));
    fe[14] = rho * P1 * (1.0 + RF * (-zj) + QF * (3.0 * (zj) * (zj)-cvsq));
    fe[15] = rho * P2 * (1.0 + RF * (-yj - zj) + QF * (3.0 * (yj + zj) * (yj + zj) - cvsq));
    fe[16] = rho * P1 * (1.0 + RF * (-yj) + QF * (3.0 * (yj) * (yj)-cvsq));
    fe[17] = rho * P2 * (1.0 + RF * (-yj + zj) + QF * (3.0 * (yj - zj) * (yj - zj) - cvsq));
    for (int q = 0; q < 18; ++q) fn[q] = x[q] - fe[q];
    *rho_out = rho;
}

/* SERIAL index clipping of flux_X/flux_Y for one axis: returns 0 = ok, 1 = return early */
typedef struct { int lo, hi; double n1, n2; int ret; } clip_t;

/* flux_X.F90:50-301.  box = {xstart, xstop, ystart, ystop, zstart, zstop}; result -> *flux */
void orc_flux_x(const double *a, int l, int m, int n, double omega, const int box[6], double *flux)
{
    geom_t g = mkgeom(l, m, n);
    int istart = box[0], istop = box[1], jstart = box[2], jstop = box[3];
    int kstart = box[4], kstop = box[5];
    int i1 = 0, i2 = 0, j1 = 0, j2 = 0, k1 = 0, k2 = 0;
    double norm1 = 1.0, norm2 = 1.0, fl[2] = {0.0, 0.0};
    if (istart > 0 && istop <= l) { i1 = istart; i2 = istop; }
    if (istart > 0 && istop > l) { i1 = istart; i2 = l; norm2 = 0.0; }
    if (istart <= 0 && istop <= l) { i1 = 1; norm1 = 0.0; i2 = istop; }
    if (istart < 0 && istop > l) return;
    if (istart > l || istop < 0) return;
    if (jstart > 0 && jstop <= m) { j1 = jstart; j2 = jstop; }
    if (jstart > 0 && jstop > m) { j1 = jstart; j2 = m; }
    if (jstart <= 0 && jstop <= m) { j1 = 1; j2 = jstop; }
    if (jstart < 0 && jstop > m) { j1 = 1; j2 = m; }
    if (jstart > m || jstop < 0) return;
    if (kstart > 0 && kstop <= n) { k1 = kstart; k2 = kstop; }
    if (kstart > 0 && kstop > n) { k1 = kstart; k2 = n; }
    if (kstart <= 0 && kstop <= n) { k1 = 1; k2 = kstop; }
    if (kstart < 0 && kstop > n) { k1 = 1; k2 = n; }
    if (kstart > n || kstop < 0) return;
    int face[2] = {i1, i2};
    const double c = (1 - 0.5 * omega);
    double x[19], fn[18], rho;
    for (int ii = 0; ii < 2; ++ii) {
        int i0 = face[ii];
        for (int k = k1; k <= k2; ++k)
            for (int j = j1; j <= j2; ++j) {
                gather(a, g, IDX(g, i0, j, k), x);
                site_fneq(x, fn, &rho);
                double sxx = -rho / RF;
                for (int q = 0; q < 18; ++q) sxx = sxx - c * CXD[q] * CXD[q] * fn[q];
                fl[ii] = fl[ii] + sxx;
            }
    }
    *flux = -fl[0] * norm1 + fl[1] * norm2;
}

/* flux_Y.F90:51-282 */
void orc_flux_y(const double *a, int l, int m, int n, double omega, const int box[6], double *flux)
{
    geom_t g = mkgeom(l, m, n);
    int istart = box[0], istop = box[1], jstart = box[2], jstop = box[3];
    int kstart = box[4], kstop = box[5];
    int i1 = 0, i2 = 0, j1 = 0, j2 = 0, k1 = 0, k2 = 0;
    double norm1 = 1.0, norm2 = 1.0, fl[2] = {0.0, 0.0};
    if (jstart > 0 && jstop <= m) { j1 = jstart; j2 = jstop; }
    if (jstart > 0 && jstop > m) { j1 = jstart; j2 = m; norm2 = 0.0; }
    if (jstart <= 0 && jstop <= m) { j1 = 1; norm1 = 0.0; j2 = jstop; }
    if (jstart < 0 && jstop > m) return;
    if (jstart > m || jstop < 0) return;
    if (istart > 0 && istop <= l) { i1 = istart; i2 = istop; }
    if (istart > 0 && istop > l) { i1 = istart; i2 = l; }
    if (istart <= 0 && istop <= l) { i1 = 1; i2 = istop; }
    if (istart < 0 && istop > l) { i1 = 1; i2 = l; }
    if (istart > l || istop < 0) return;
    if (kstart > 0 && kstop <= n) { k1 = kstart; k2 = kstop; }
    if (kstart > 0 && kstop > n) { k1 = kstart; k2 = n; }
    if (kstart <= 0 && kstop <= n) { k1 = 1; k2 = kstop; }
    if (kstart < 0 && kstop > n) { k1 = 1; k2 = n; }
    if (kstart > n || kstop < 0) return;
    int face[2] = {j1, j2};
    const double c = (1 - 0.5 * omega);
    double x[19], fn[18], rho;
    for (int ii = 0; ii < 2; ++ii) {
        int j0 = face[ii];
        for (int k = k1; k <= k2; ++k)
            for (int i = i1; i <= i2; ++i) {
                gather(a, g, IDX(g, i, j0, k), x);
                site_fneq(x, fn, &rho);
                double sxy = 0.0;
                for (int q = 0; q < 18; ++q) sxy = sxy - c * CXD[q] * CYD[q] * fn[q];
                fl[ii] = fl[ii] + sxy;
            }
    }
    *flux = +fl[0] * norm1 + fl[1] * norm2;
}

/* probe_global.F90:41-77: out = {u, v, w, rho} at interior site (i, j, k); returns 1 if the
 * site is outside the lattice (the reference then writes nothing) */
int orc_probe(const double *a, int l, int m, int n, int i, int j, int k, double out[4])
{
    geom_t g = mkgeom(l, m, n);
    if (!(k > 0 && k <= n && j > 0 && j <= m && i > 0 && i <= l)) return 1;
    double x[19];
    gather(a, g, IDX(g, i, j, k), x);
    double rho = x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9] + x[10]
                 + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18];
    double xj = (x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]);
    double yj = (x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]);
    double zj = (x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]);
    out[0] = xj / rho;
    out[1] = yj / rho;
    out[2] = zj / rho;
    out[3] = rho;
    return 0;
}

/* diagno.F90:53-131 (SERIAL): out = {<rho>, <u>, <v>, <w>, <u^2>} */
void orc_diagno(const double *a, const int32_t *obs, int l, int m, int n, double out[5])
{
    geom_t g = mkgeom(l, m, n);
    double rdv = 1. / ((double)l * (double)m * (double)n);
    double rtot = 0., xtot = 0., ytot = 0., ztot = 0., stot = 0.;
    double x[19];
    for (int k = 1; k <= n; ++k)
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= l; ++i) {
                double rho, xj, yj, zj;
                if (OBS(i, j, k) == 1) {
                    rho = 1.0; xj = 0.0; yj = 0.0; zj = 0.0;
                } else {
                    gather(a, g, IDX(g, i, j, k), x);
                    rho = +x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9]
                          + x[10] + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18];
                    double rhoinv = 1.0 / rho;
                    xj = +(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]) * rhoinv;
                    yj = +(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]) * rhoinv;
                    zj = +(x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]) * rhoinv;
                }
                rtot = rtot + rho;
                xtot = xtot + xj;
                ytot = ytot + yj;
                ztot = ztot + zj;
                stot = stot + xj * xj + yj * yj + zj * zj;
            }
    out[0] = rtot * rdv;
    out[1] = xtot * rdv;
    out[2] = ytot * rdv;
    out[3] = ztot * rdv;
    out[4] = stot * rdv;
}

/* ------------------------------------------------------------------------------------------ */
/* binary VTK output fields (reference vtk_*.binary.F90)                                       */
/* ------------------------------------------------------------------------------------------ */

/* velocity of one site as the binary VTK writers compute it (vtk_xy.binary.F90:93-115,
 * vtk_3d.binary.F90:80-100): the momenta and the density summed in FP64 in the written order,
 * each rounded to real(sp), then divided in single precision; o = {u_x, u_y, u_z} */
static void vtk_site(const double *a, geom_t g, int i, int j, int k, float o[3])
{
    double x[19];
    gather(a, g, IDX(g, i, j, k), x);
    const float u = (float)(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12] - x[13]);
    const float w = (float)(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16] - x[17]);
    const float v = (float)(x[3] + x[5] + x[6] + x[12] + x[17] - x[1] - x[8] - x[10] - x[14] - x[15]);
    const float den = (float)(x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9]
                              + x[10] + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18]);
    o[0] = u / den;
    o[1] = w / den;
    o[2] = v / den;
}

/* plane 0: yz plane x = index (vtk_yz.binary.F90:90-118, loops k, j); 1: xz plane y = index
 * (vtk_xz.binary.F90, loops k, i); 2: xy plane z = index (vtk_xy.binary.F90, loops j, i);
 * 3: the whole lattice (vtk_3d.binary.F90:78-104, loops k, j, i).  out: 3 floats per site in
 * loop order (innermost fastest). */
void orc_vtk_velocity(const double *a, int l, int m, int n, int plane, int index, float *out)
{
    geom_t g = mkgeom(l, m, n);
    long t = 0;
    if (plane == 0) {
        for (int k = 1; k <= n; ++k)
            for (int j = 1; j <= m; ++j, ++t) vtk_site(a, g, index, j, k, out + 3 * t);
    } else if (plane == 1) {
        for (int k = 1; k <= n; ++k)
            for (int i = 1; i <= l; ++i, ++t) vtk_site(a, g, i, index, k, out + 3 * t);
    } else if (plane == 2) {
        for (int j = 1; j <= m; ++j)
            for (int i = 1; i <= l; ++i, ++t) vtk_site(a, g, i, j, index, out + 3 * t);
    } else {
        for (int k = 1; k <= n; ++k)
            for (int j = 1; j <= m; ++j)
                for (int i = 1; i <= l; ++i, ++t) vtk_site(a, g, i, j, k, out + 3 * t);
    }
}

/* vorticity of the xy plane z = k0 as vtk_om.binary.F90:88-131 computes it, for i = 2..l-1
 * (fastest), j = 2..m-1.  The writer's expressions are kept as written: the densities are
 * rounded to real(sp); in each momentum only the last population is divided by the density
 * (operator precedence), the rest of the sum is not; the difference is formed in single
 * precision. */
void orc_vtk_vorticity(const double *a, int l, int m, int n, int k0, float *out)
{
    geom_t g = mkgeom(l, m, n);
    long t = 0;
    double x[19];
    for (int j = 2; j <= m - 1; ++j)
        for (int i = 2; i <= l - 1; ++i, ++t) {
            float d[4];
            const int di[4] = {0, 0, -1, 1}, dj[4] = {1, -1, 0, 0}; /* up, down, left, right */
            for (int s = 0; s < 4; ++s) {
                gather(a, g, IDX(g, i + di[s], j + dj[s], k0), x);
                d[s] = (float)(x[0] + x[1] + x[2] + x[3] + x[4] + x[5] + x[6] + x[7] + x[8] + x[9]
                               + x[10] + x[11] + x[12] + x[13] + x[14] + x[15] + x[16] + x[17] + x[18]);
            }
            float mom[4];
            for (int s = 0; s < 2; ++s) { /* up1, down1: x momentum */
                gather(a, g, IDX(g, i + di[s], j + dj[s], k0), x);
                mom[s] = (float)(x[0] + x[1] + x[2] + x[3] + x[4] - x[9] - x[10] - x[11] - x[12]
                                 - x[13] / (double)d[s]);
            }
            for (int s = 2; s < 4; ++s) { /* left1, right1: y momentum */
                gather(a, g, IDX(g, i + di[s], j + dj[s], k0), x);
                mom[s] = (float)(x[2] + x[6] + x[7] + x[8] + x[11] - x[0] - x[9] - x[15] - x[16]
                                 - x[17] / (double)d[s]);
            }
            out[t] = 0.5f * (mom[0] - mom[1]) - 0.5f * (mom[3] - mom[2]);
        }
}

/* ------------------------------------------------------------------------------------------ */
/* grey lattice (partial bounce-back) and its discrete adjoint -- no reference counterpart     */
/*                                                                                            */
/* Link-wise partial bounce-back: the link between x and its source y = x - c_q carries the   */
/* fraction sigma = 1 - (1 - s(x))(1 - s(y)) (s = 0 on ghost sites); the pulled population is */
/*     f_q(x) = (1 - sigma) src_q + sigma a_qbar(x),                                          */
/* src_q = a_q(y) inside the lattice, the inlet or outlet equilibrium across the x faces.     */
/* sigma is symmetric in the link, so mass is conserved on every link, and a binary s gives   */
/* exactly the halfway bounce-back of the physics step at every fluid site.                   */
/* Force on the grey body: F = sum over interior links (x, q) of sigma c_q (a_q(y) - a_qbar(x)) */
/* (each undirected link is met twice, once from each end).                                   */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    int l, m, n;
    double omega, u_in;
    long N;
    int xopen; /* 1: inlet at x = 0, outlet at x = l+1; 0: periodic x */
} grey_t;

#define SIDX(G, i, j, k) (((long)(i) - 1) + (long)(G).l * (((long)(j) - 1) + (long)(G).m * ((long)(k) - 1)))

/* pulled populations of site (i,j,k) (1-based, interior) from the dense interior state a[19][N];
 * sig[q] and kind[q] (0 interior source, 1 inlet, 2 outlet) describe each link */
static void grey_pull(const double *a, const double *s, grey_t G, int i, int j, int k, double *f,
                      double *sig, int *kind, double *srcv, long *srcidx)
{
    double inl[5];
    inlet_eq(G.u_in, inl);
    const long xs = SIDX(G, i, j, k);
    const double sx = s[xs];
    for (int q = 0; q < 18; ++q) {
        int si = i - CX[q], sj = wrap(j - CY[q], G.m), sk = wrap(k - CZ[q], G.n);
        double src, sy;
        if (!G.xopen) si = wrap(si, G.l);
        if (si < 1) {
            src = inl[q];
            sy = 0.0;
            kind[q] = 1;
            srcidx[q] = -1;
        } else if (si > G.l) {
            double y[19], o[5];
            long ys = SIDX(G, G.l, sj, sk);
            for (int p = 0; p < 19; ++p) y[p] = a[p * G.N + ys];
            outlet_eq(y, o);
            src = o[q - 9];
            sy = 0.0;
            kind[q] = 2;
            srcidx[q] = ys;
        } else {
            long ys = SIDX(G, si, sj, sk);
            src = a[q * G.N + ys];
            sy = s[ys];
            kind[q] = 0;
            srcidx[q] = ys;
        }
        double sg = 1.0 - (1.0 - sx) * (1.0 - sy);
        sig[q] = sg;
        srcv[q] = src;
        f[q] = (1.0 - sg) * src + sg * a[OPP[q] * G.N + xs];
    }
    f[18] = a[18 * G.N + xs];
}

static void grey_step(const double *a, double *b, const double *s, grey_t G)
{
    double f[19], sig[18], srcv[18];
    int kind[18];
    long sidx[18];
    for (int k = 1; k <= G.n; ++k)
        for (int j = 1; j <= G.m; ++j)
            for (int i = 1; i <= G.l; ++i) {
                grey_pull(a, s, G, i, j, k, f, sig, kind, srcv, sidx);
                collide_site(f, f, G.omega, 0.0);
                long xs = SIDX(G, i, j, k);
                for (int q = 0; q < 19; ++q) b[q * G.N + xs] = f[q];
            }
}

static void grey_force(const double *a, const double *s, grey_t G, double F[3])
{
    double f[19], sig[18], srcv[18];
    int kind[18];
    long sidx[18];
    double fx = 0, fy = 0, fz = 0;
    for (int k = 1; k <= G.n; ++k)
        for (int j = 1; j <= G.m; ++j)
            for (int i = 1; i <= G.l; ++i) {
                grey_pull(a, s, G, i, j, k, f, sig, kind, srcv, sidx);
                long xs = SIDX(G, i, j, k);
                for (int q = 0; q < 18; ++q) {
                    if (kind[q] != 0) continue;
                    double d = sig[q] * (srcv[q] - a[OPP[q] * G.N + xs]);
                    fx += CX[q] * d;
                    fy += CY[q] * d;
                    fz += CZ[q] * d;
                }
            }
    F[0] = fx; F[1] = fy; F[2] = fz;
}

/* (dC/df)^T lam for the BGK collision at pre-collision f (no body force) */
static void collide_adj(const double *f, const double *lam, double omega, double *out)
{
    double rho = 0, jx = 0, jy = 0, jz = 0;
    for (int q = 0; q < 19; ++q) {
        rho += f[q];
        jx += CX[q] * f[q];
        jy += CY[q] * f[q];
        jz += CZ[q] * f[q];
    }
    double ux = jx / rho, uy = jy / rho, uz = jz / rho, usq = ux * ux + uy * uy + uz * uz;
    double A = 0, Bx = 0, By = 0, Bz = 0;
    for (int q = 0; q < 19; ++q) {
        double w = q == 18 ? P0 : ((CX[q] * CX[q] + CY[q] * CY[q] + CZ[q] * CZ[q]) == 1 ? P1 : P2);
        double cu = CX[q] * ux + CY[q] * uy + CZ[q] * uz;
        A += lam[q] * w * (1.0 - 4.5 * cu * cu + 1.5 * usq);
        double g = lam[q] * w * (3.0 + 9.0 * cu);
        Bx += g * CX[q] - lam[q] * w * 3.0 * ux;
        By += g * CY[q] - lam[q] * w * 3.0 * uy;
        Bz += g * CZ[q] - lam[q] * w * 3.0 * uz;
    }
    for (int p = 0; p < 19; ++p)
        out[p] = (1.0 - omega) * lam[p] + omega * (A + Bx * CX[p] + By * CY[p] + Bz * CZ[p]);
}

/* d OUT_q / d a_p of the outlet equilibrium (momentum J of the source site, rho fixed to 1) */
static double outlet_grad(const double *y, int q, int p)
{
    double Jx = 0, Jy = 0, Jz = 0;
    for (int r = 0; r < 19; ++r) {
        Jx += CX[r] * y[r];
        Jy += CY[r] * y[r];
        Jz += CZ[r] * y[r];
    }
    double w = ((CX[q] * CX[q] + CY[q] * CY[q] + CZ[q] * CZ[q]) == 1) ? P1 : P2;
    double cJ = CX[q] * Jx + CY[q] * Jy + CZ[q] * Jz;
    double cc = CX[q] * CX[p] + CY[q] * CY[p] + CZ[q] * CZ[p];
    double Jc = Jx * CX[p] + Jy * CY[p] + Jz * CZ[p];
    return w * (3.0 * cc + 9.0 * cJ * cc - 3.0 * Jc);
}

/* one backward step: lam_out = (d a^{t+1} / d a^t)^T lam_in (+ wobj dF/da), gs += ..., at a = a^t */
static void grey_step_adj(const double *a, const double *s, grey_t G, const double *lam_in,
                          double *lam_out, double *lin, double wobj, double *gs)
{
    double f[19], sig[18], srcv[18], lam[19], tmp[19];
    int kind[18];
    long sidx[18];
    memset(lam_out, 0, sizeof(double) * 19 * G.N);
    for (int k = 1; k <= G.n; ++k)
        for (int j = 1; j <= G.m; ++j)
            for (int i = 1; i <= G.l; ++i) {
                long xs = SIDX(G, i, j, k);
                grey_pull(a, s, G, i, j, k, f, sig, kind, srcv, sidx);
                for (int q = 0; q < 19; ++q) lam[q] = lam_in[q * G.N + xs];
                collide_adj(f, lam, G.omega, tmp);
                for (int q = 0; q < 19; ++q) lin[q * G.N + xs] = tmp[q];
            }
    for (int k = 1; k <= G.n; ++k)
        for (int j = 1; j <= G.m; ++j)
            for (int i = 1; i <= G.l; ++i) {
                long xs = SIDX(G, i, j, k);
                grey_pull(a, s, G, i, j, k, f, sig, kind, srcv, sidx);
                const double sx = s[xs];
                for (int q = 0; q < 18; ++q) {
                    const double L = lin[q * G.N + xs];
                    const double sg = sig[q];
                    const double abar = a[OPP[q] * G.N + xs];
                    lam_out[OPP[q] * G.N + xs] += sg * L;
                    if (kind[q] == 0) {
                        long ys = sidx[q];
                        lam_out[q * G.N + ys] += (1.0 - sg) * L;
                        double d = abar - srcv[q];
                        gs[xs] += L * d * (1.0 - s[ys]);
                        gs[ys] += L * d * (1.0 - sx);
                        if (wobj != 0.0) {
                            double c = wobj * CX[q];
                            lam_out[q * G.N + ys] += c * sg;
                            lam_out[OPP[q] * G.N + xs] -= c * sg;
                            gs[xs] += c * (srcv[q] - abar) * (1.0 - s[ys]);
                            gs[ys] += c * (srcv[q] - abar) * (1.0 - sx);
                        }
                    } else {
                        gs[xs] += L * (abar - srcv[q]);
                        if (kind[q] == 2) {
                            double y[19];
                            long ys = sidx[q];
                            for (int p = 0; p < 19; ++p) y[p] = a[p * G.N + ys];
                            for (int p = 0; p < 19; ++p)
                                lam_out[p * G.N + ys] += (1.0 - sg) * L * outlet_grad(y, q, p);
                        }
                    }
                }
                lam_out[18 * G.N + xs] += lin[18 * G.N + xs];
            }
}

/* objective-only adjoint seed at a^t: lam = wobj dF/da, gs += wobj dF/ds */
static void grey_force_adj(const double *a, const double *s, grey_t G, double wobj, double *lam,
                           double *gs)
{
    double f[19], sig[18], srcv[18];
    int kind[18];
    long sidx[18];
    for (int k = 1; k <= G.n; ++k)
        for (int j = 1; j <= G.m; ++j)
            for (int i = 1; i <= G.l; ++i) {
                long xs = SIDX(G, i, j, k);
                grey_pull(a, s, G, i, j, k, f, sig, kind, srcv, sidx);
                for (int q = 0; q < 18; ++q) {
                    if (kind[q] != 0) continue;
                    long ys = sidx[q];
                    double c = wobj * CX[q];
                    double abar = a[OPP[q] * G.N + xs];
                    lam[q * G.N + ys] += c * sig[q];
                    lam[OPP[q] * G.N + xs] -= c * sig[q];
                    gs[xs] += c * (srcv[q] - abar) * (1.0 - s[ys]);
                    gs[ys] += c * (srcv[q] - abar) * (1.0 - s[xs]);
                }
            }
}

/* Dense-interior API: a[19*N] (N = l*m*n, i fastest), s[N]. */
int orc_grey_steps(double *a, const double *s, int l, int m, int n, double omega, double u_in,
                   int xopen, int nsteps)
{
    grey_t G = {l, m, n, omega, u_in, (long)l * m * n, xopen};
    double *b = (double *)malloc(sizeof(double) * 19 * G.N);
    if (!b) return -1;
    for (int t = 0; t < nsteps; ++t) {
        grey_step(a, b, s, G);
        memcpy(a, b, sizeof(double) * 19 * G.N);
    }
    free(b);
    return 0;
}

void orc_grey_force(const double *a, const double *s, int l, int m, int n, double u_in, int xopen,
                    double F[3])
{
    grey_t G = {l, m, n, 0.0, u_in, (long)l * m * n, xopen};
    grey_force(a, s, G, F);
}

/* J = (1/(t1-t0+1)) sum_{t=t0..t1} F_x(a^t) / norm starting from a0 (a^0); gs (may be NULL)
 * receives dJ/ds.  Forward states are checkpointed every K steps and recomputed per segment. */
int orc_grey_objective(const double *a0, const double *s, int l, int m, int n, double omega,
                       double u_in, int xopen, int t0, int t1, int K, double norm, double *J,
                       double *gs)
{
    grey_t G = {l, m, n, omega, u_in, (long)l * m * n, xopen};
    const size_t S = sizeof(double) * 19 * G.N;
    if (t0 < 1 || t1 < t0 || K < 1) return -1;
    const double wobj = 1.0 / ((double)(t1 - t0 + 1) * norm);
    int nck = t1 / K + 1;
    double *ck = (double *)malloc(S * nck);
    double *cur = (double *)malloc(S), *nxt = (double *)malloc(S);
    if (!ck || !cur || !nxt) return -1;
    memcpy(cur, a0, S);
    double acc = 0.0, F[3];
    for (int t = 0; t <= t1; ++t) {
        if (t % K == 0) memcpy(ck + (size_t)(t / K) * 19 * G.N, cur, S);
        if (t >= t0) {
            grey_force(cur, s, G, F);
            acc += F[0];
        }
        if (t == t1) break;
        grey_step(cur, nxt, s, G);
        double *sw = cur; cur = nxt; nxt = sw;
    }
    *J = acc * wobj;
    if (gs) {
        memset(gs, 0, sizeof(double) * G.N);
        double *seg = (double *)malloc(S * (size_t)(K < t1 ? K : t1)); /* a segment holds <= min(K, t1) states */
        double *lam = (double *)calloc(19 * G.N, sizeof(double));
        double *lam2 = (double *)malloc(S), *lin = (double *)malloc(S);
        if (!seg || !lam || !lam2 || !lin) {
            free(seg); free(lam); free(lam2); free(lin); free(ck); free(cur); free(nxt);
            return -1;
        }
        /* seed at t1 */
        {
            int c = t1 / K;
            memcpy(cur, ck + (size_t)c * 19 * G.N, S);
            for (int t = c * K; t < t1; ++t) {
                grey_step(cur, nxt, s, G);
                double *sw = cur; cur = nxt; nxt = sw;
            }
            grey_force_adj(cur, s, G, wobj, lam, gs);
        }
        for (int c = (t1 - 1) / K; c >= 0; --c) {
            int tb = c * K, te = (c + 1) * K < t1 ? (c + 1) * K : t1; /* states tb..te-1 */
            memcpy(seg, ck + (size_t)c * 19 * G.N, S);
            for (int t = tb + 1; t < te; ++t)
                grey_step(seg + (size_t)(t - 1 - tb) * 19 * G.N, seg + (size_t)(t - tb) * 19 * G.N, s, G);
            for (int t = te - 1; t >= tb; --t) {
                const double *at = seg + (size_t)(t - tb) * 19 * G.N;
                grey_step_adj(at, s, G, lam, lam2, lin, t >= t0 ? wobj : 0.0, gs);
                double *sw = lam; lam = lam2; lam2 = sw;
            }
        }
        free(seg); free(lam); free(lam2); free(lin);
    }
    free(ck); free(cur); free(nxt);
    return 0;
}

/* operator-level adjoint pieces for transpose tests */
void orc_collide_adj(const double *f, const double *lam, double omega, double *out)
{
    collide_adj(f, lam, omega, out);
}
void orc_grey_step_adj(const double *a, const double *s, int l, int m, int n, double omega,
                       double u_in, int xopen, const double *lam_in, double *lam_out, double wobj,
                       double *gs)
{
    grey_t G = {l, m, n, omega, u_in, (long)l * m * n, xopen};
    double *lin = (double *)malloc(sizeof(double) * 19 * G.N);
    grey_step_adj(a, s, G, lam_in, lam_out, lin, wobj, gs);
    free(lin);
}
