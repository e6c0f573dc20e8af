/*
 * hbi_oracle.c -- CPU FP64 restatement of HBI's earthquake-cycle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP library in
 * hbi_amd/ is checked against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * What it restates (reference = amtseismo/hbi, Fortran 90):
 *   - physical constants                         m_const.f90:1-11
 *   - mesh generators 2dp/2dvs, 3dph, STL frames  main_LH.f90:1701-1718, 1781-1835, 1940-1990
 *   - 2-D antiplane kernel with image            m_HACApK_calc_entry_ij.f90:55-57, 685-695
 *   - Okada rectangular kernel                   m_HACApK_calc_entry_ij.f90:1065-1145, okada.f90:10-597
 *   - TD full-space kernel                       m_HACApK_calc_entry_ij.f90:802-851, TDstressFS.f90:45-403, 560-627, 1012-1019
 *   - 2-D / 2.5-D / planar kernels 2dp, 3dp,     m_HACApK_calc_entry_ij.f90:92-166, 300-339, 412-470, 648-661,
 *     2dph (2dnh), 2dn, 25d                      697-750, 934-1064
 *   - TD half-space kernel                       m_HACApK_calc_entry_ij.f90:852-898, TDstressFS.f90:7-44, 236-293, 405-558,
 *                                                629-1011 (free-surface correction: strains as complex-step
 *                                                derivatives of the published harmonic displacement field)
 *   - geometric cluster tree + bounding boxes    m_HACApK_base.f90:3191-3333
 *   - leaf partition (np=1 lattice) + sort       m_HACApK_base.f90:2372-2460, 2937-2995, 3114-3166
 *   - ACA+ and dense fill                        m_HACApK_base.f90:1638-1946, 1949-2043; HACApK_lib.f90:31-116
 *   - H-matrix-vector product                    m_HACApK_solve.f90:220-337, 2108-2164
 *   - friction RHS (aging, slip laws)            main_LH.f90:2053-2244
 *   - Cash-Karp RK45 + elementary controller     main_LH.f90:2270-2370, 2472-2554
 *   - post-step update and monitors              main_LH.f90:1208-1336
 *
 * Arithmetic is written in the reference's evaluation order and compiled
 * without FP contraction, so results on x86-64 follow the reference to the
 * last bit except where libm or Fortran runtime intrinsics differ.
 *
 * Parity pinning of each part is documented in the tests that use it:
 *   kernels      -> golden vectors from the compiled reference kernels
 *                   (tests/golden/kernel_entries_*.npz), closed forms, symmetries
 *   partition    -> exact recount against invariants (cover, disjointness,
 *                   admissibility) and brute force on small meshes
 *   ACA+ / H     -> ||H x - A x|| / ||A x|| against the dense matrix
 *   matvec       -> dense product, linearity
 *   RK45         -> order tests on y' = lambda y, tableau identities
 *   friction     -> steady-state invariants, finite differences
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OG_PI (4.0 * atan(1.0))
static const double og_pois = 0.25;
static const double og_rigid = 32.04;
static const double og_vs = 3.464;

/* ------------------------------------------------------------------------- */
/* geometry                                                                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    int kind;   /* 1: 2dvs antiplane, 2: okada half-space (3dph/3dhr), 3: okada full-space (3dnr), 4: TD full-space (3dnt),
                   5: TD half-space (3dht), 6: 2dp, 7: 3dp, 8: 2dph/2dnh, 9: 2dn, 10: 25d,
                   11: 2dpfz, 12: 2dn_vector, 13: 2dna */
    int n;
    int vcomp;  /* 0: shear traction ('s'), 1: normal traction ('n') */
    int md;     /* TD slip mode: 0 's' (receiver rake), 1 'st', 2 'dp', 3 'open' */
    double w;   /* okada normalisation length (st_bemv%w) */
    const double *xcol, *ycol, *zcol;
    const double *xel, *xer;
    const double *ang, *angd, *dsl, *dsd, *rake;
    const double *xs1, *xs2, *xs3, *ys1, *ys2, *ys3, *zs1, *zs2, *zs3;
    const double *ev11, *ev12, *ev13, *ev21, *ev22, *ev23, *ev31, *ev32, *ev33;
    const double *yel, *yer, *xs4, *zs4, *ds;
} og_geom;

/* coordinate2dp, main_LH.f90:1701-1718 */
void og_mesh_2dp(int n, double ds0, double *xel, double *xer, double *xcol)
{
    for (int i = 1; i <= n; i++) {
        xel[i - 1] = (i - 1) * ds0;
        xer[i - 1] = i * ds0;
        xcol[i - 1] = 0.5 * (xel[i - 1] + xer[i - 1]);
    }
}

/* coordinate3ddip, main_LH.f90:1781-1835 */
void og_mesh_3ddip(int imax, int jmax, double ds0, double dipangle,
                   double *xcol, double *ycol, double *zcol, double *ang, double *angd)
{
    double pi = OG_PI;
    double *yr = (double *)malloc(sizeof(double) * (jmax + 1));
    double *zr = (double *)malloc(sizeof(double) * (jmax + 1));
    yr[0] = 0.0; zr[0] = 0.0;
    for (int j = 1; j <= jmax; j++) {
        yr[j] = yr[j - 1] - ds0 * cos(dipangle * pi / 180.0);
        zr[j] = zr[j - 1] - ds0 * sin(dipangle * pi / 180.0);
    }
    int k = 0;
    for (int i = 1; i <= imax; i++)
        for (int j = 1; j <= jmax; j++) {
            xcol[k] = ((double)(i - imax / 2) - 0.5) * ds0;
            ycol[k] = 0.5 * (yr[j - 1] + yr[j]);
            zcol[k] = 0.5 * (zr[j - 1] + zr[j]);
            angd[k] = atan2(zr[j - 1] - zr[j], yr[j - 1] - yr[j]);
            ang[k] = 0.0;
            k++;
        }
    free(yr); free(zr);
}

/* evcalc, main_LH.f90:1940-1990: local frames and areas of triangles */
void og_evcalc(int n, const double *xs1, const double *xs2, const double *xs3,
               const double *ys1, const double *ys2, const double *ys3,
               const double *zs1, const double *zs2, const double *zs3,
               double *ev11, double *ev12, double *ev13, double *ev21, double *ev22, double *ev23,
               double *ev31, double *ev32, double *ev33, double *ds)
{
    for (int k = 0; k < n; k++) {
        double vba0 = xs2[k] - xs1[k], vba1 = ys2[k] - ys1[k], vba2 = zs2[k] - zs1[k];
        double vca0 = xs3[k] - xs1[k], vca1 = ys3[k] - ys1[k], vca2 = zs3[k] - zs1[k];
        ev31[k] = vba1 * vca2 - vba2 * vca1;
        ev32[k] = vba2 * vca0 - vba0 * vca2;
        ev33[k] = vba0 * vca1 - vba1 * vca0;
        double rr = sqrt(ev31[k] * ev31[k] + ev32[k] * ev32[k] + ev33[k] * ev33[k]);
        ev31[k] = ev31[k] / rr; ev32[k] = ev32[k] / rr; ev33[k] = ev33[k] / rr;
        if (fabs(ev33[k]) < 1.0) {
            ev11[k] = -ev32[k]; ev12[k] = ev31[k]; ev13[k] = 0.0;
            rr = sqrt(ev11[k] * ev11[k] + ev12[k] * ev12[k]);
            ev11[k] = ev11[k] / rr; ev12[k] = ev12[k] / rr;
        } else {
            ev11[k] = 1.0; ev12[k] = 0.0; ev13[k] = 0.0;
        }
        ev21[k] = ev32[k] * ev13[k] - ev33[k] * ev12[k];
        ev22[k] = ev33[k] * ev11[k] - ev31[k] * ev13[k];
        ev23[k] = ev31[k] * ev12[k] - ev32[k] * ev11[k];
        double t1 = vba0 * vba0 + vba1 * vba1 + vba2 * vba2;
        double t2 = vca0 * vca0 + vca1 * vca1 + vca2 * vca2;
        double t3 = vba0 * vca0 + vba1 * vca1 + vba2 * vca2;
        ds[k] = 0.5 * sqrt(t1 * t2 - t3 * t3);
    }
}

/* ------------------------------------------------------------------------- */
/* 2-D antiplane kernel, m_HACApK_calc_entry_ij.f90:685-695 and 55-57        */
/* ------------------------------------------------------------------------- */
static double og_antiplane(double xc, double xl, double xr)
{
    double factor = og_rigid / (2.0 * OG_PI);
    return factor * (1.0 / (xc - xr) - 1.0 / (xc - xl));
}

/* ------------------------------------------------------------------------- */
/* Okada rectangular dislocation, okada.f90                                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    double xi2, et2, q2, r, r2, r3, r5, yp, d, tt, alx, ale, x11, y11, x32, y32, ey, ez, fy, fz, gy, gz;
} okc_t;

/* station geometry constants, the block repeated at okada.f90:260-290 */
static void okc_set(okc_t *c, double xi, double et, double q, double sd, double cd)
{
    c->xi2 = xi * xi;
    c->et2 = et * et;
    c->q2 = q * q;
    c->r2 = c->xi2 + c->et2 + c->q2;
    c->r = sqrt(c->r2);
    c->r3 = c->r * c->r2;
    c->r5 = c->r3 * c->r2;
    c->yp = et * cd + q * sd;
    c->d = et * sd - q * cd;
    c->tt = atan(xi * et / (q * c->r));
    double rxi = c->r + xi;
    c->alx = log(rxi);
    c->x11 = 1.0 / (c->r * rxi);
    c->x32 = (c->r + rxi) * c->x11 * c->x11 / c->r;
    double ret = c->r + et;
    c->ale = log(ret);
    c->y11 = 1.0 / (c->r * ret);
    c->y32 = (c->r + ret) * c->y11 * c->y11 / c->r;
    c->ey = sd / c->r - c->yp * q / c->r3;
    c->ez = cd / c->r + c->d * q / c->r3;
    c->fy = c->d / c->r3 + c->xi2 * c->y32 * sd;
    c->fz = c->yp / c->r3 + c->xi2 * c->y32 * cd;
    c->gy = 2.0 * c->x11 * sd - c->yp * q * c->x32;
    c->gz = 2.0 * c->x11 * cd + c->d * q * c->x32;
}

/* part A, okada.f90:227-332 */
static void ok_ua(double xi, double et, double q, double sd, double cd, double disl1, double disl2,
                  double alpha, double *dua)
{
    const double pi2 = 8.0 * atan(1.0);
    double alp1 = (1.0 - alpha) / 2.0, alp2 = alpha / 2.0;
    okc_t c; okc_set(&c, xi, et, q, sd, cd);
    double r = c.r, r3 = c.r3, q2 = c.q2, xi2 = c.xi2, d = c.d, yp = c.yp;
    double xy = xi * c.y11, qx = q * c.x11, qy = q * c.y11;
    double du[12];
    du[0] = c.tt / 2.0 + alp2 * xi * qy;
    du[1] = alp2 * q / r;
    du[2] = alp1 * c.ale - alp2 * q * qy;
    du[3] = -alp1 * qy - alp2 * xi2 * q * c.y32;
    du[4] = -alp2 * xi * q / r3;
    du[5] = alp1 * xy + alp2 * xi * q2 * c.y32;
    du[6] = alp1 * xy * sd + alp2 * xi * c.fy + d / 2.0 * c.x11;
    du[7] = alp2 * c.ey;
    du[8] = alp1 * (cd / r + qy * sd) - alp2 * q * c.fy;
    du[9] = alp1 * xy * cd + alp2 * xi * c.fz + yp / 2.0 * c.x11;
    du[10] = alp2 * c.ez;
    du[11] = -alp1 * (sd / r - qy * cd) - alp2 * q * c.fz;
    for (int i = 0; i < 12; i++) dua[i] = 0.0 + disl1 / pi2 * du[i];
    du[0] = alp2 * q / r;
    du[1] = c.tt / 2.0 + alp2 * et * qx;
    du[2] = alp1 * c.alx - alp2 * q * qx;
    du[3] = -alp2 * xi * q / r3;
    du[4] = -qy / 2.0 - alp2 * et * q / r3;
    du[5] = alp1 / r + alp2 * q2 / r3;
    du[6] = alp2 * c.ey;
    du[7] = alp1 * d * c.x11 + xy / 2.0 * sd + alp2 * et * c.gy;
    du[8] = alp1 * yp * c.x11 - alp2 * q * c.gy;
    du[9] = alp2 * c.ez;
    du[10] = alp1 * yp * c.x11 + xy / 2.0 * cd + alp2 * et * c.gz;
    du[11] = -alp1 * d * c.x11 - alp2 * q * c.gz;
    for (int i = 0; i < 12; i++) dua[i] = dua[i] + disl2 / pi2 * du[i];
}

/* part B, okada.f90:333-475 */
static void ok_ub(double xi, double et, double q, double sd, double cd, double disl1, double disl2,
                  double alpha, double *dub)
{
    const double pi2 = 8.0 * atan(1.0);
    double alp3 = (1.0 - alpha) / alpha;
    okc_t c; okc_set(&c, xi, et, q, sd, cd);
    double r = c.r, r3 = c.r3, q2 = c.q2, xi2 = c.xi2, d = c.d, yp = c.yp;
    double rd = r + d;
    double d11 = 1.0 / (r * rd);
    double aj2 = xi * yp / rd * d11;
    double aj5 = -(d + yp * yp / rd) * d11;
    double ai3, ai4, ak1, ak3, aj3, aj6;
    if (cd != 0.0) {
        if (xi == 0.0) {
            ai4 = 0.0;
        } else {
            double xs = sqrt(xi2 + q2);
            ai4 = 1.0 / cd / cd * (xi / rd * sd * cd + 2.0 * atan((et * (xs + q * cd) + xs * (r + xs) * sd) / (xi * (r + xs) * cd)));
        }
        ai3 = (yp * cd / rd - c.ale + sd * log(rd)) / cd / cd;
        ak1 = xi * (d11 - c.y11 * sd) / cd;
        ak3 = (q * c.y11 - yp * d11) / cd;
        aj3 = (ak1 - aj2 * sd) / cd;
        aj6 = (ak3 - aj5 * sd) / cd;
    } else {
        double rd2 = rd * rd;
        ai3 = (et / rd + yp * q / rd2 - c.ale) / 2.0;
        ai4 = xi * yp / rd2 / 2.0;
        ak1 = xi * q / rd * d11;
        ak3 = sd / rd * (xi2 * d11 - 1.0);
        aj3 = -xi / rd2 * (q2 * d11 - 0.5);
        aj6 = -yp / rd2 * (xi2 * d11 - 0.5);
    }
    double xy = xi * c.y11;
    double ai1 = -xi / rd * cd - ai4 * sd;
    double ai2 = log(rd) + ai3 * sd;
    double ak2 = 1.0 / r + ak3 * sd;
    double ak4 = xy * cd - ak1 * sd;
    double aj1 = aj5 * cd - aj6 * sd;
    double aj4 = -xy - aj2 * cd + aj3 * sd;
    double qx = q * c.x11, qy = q * c.y11;
    double du[12];
    du[0] = -xi * qy - c.tt - alp3 * ai1 * sd;
    du[1] = -q / r + alp3 * yp / rd * sd;
    du[2] = q * qy - alp3 * ai2 * sd;
    du[3] = xi2 * q * c.y32 - alp3 * aj1 * sd;
    du[4] = xi * q / r3 - alp3 * aj2 * sd;
    du[5] = -xi * q2 * c.y32 - alp3 * aj3 * sd;
    du[6] = -xi * c.fy - d * c.x11 + alp3 * (xy + aj4) * sd;
    du[7] = -c.ey + alp3 * (1.0 / r + aj5) * sd;
    du[8] = q * c.fy - alp3 * (qy - aj6) * sd;
    du[9] = -xi * c.fz - yp * c.x11 + alp3 * ak1 * sd;
    du[10] = -c.ez + alp3 * yp * d11 * sd;
    du[11] = q * c.fz + alp3 * ak2 * sd;
    for (int i = 0; i < 12; i++) dub[i] = 0.0 + disl1 / pi2 * du[i];
    du[0] = -q / r + alp3 * ai3 * sd * cd;
    du[1] = -et * qx - c.tt - alp3 * xi / rd * sd * cd;
    du[2] = q * qx + alp3 * ai4 * sd * cd;
    du[3] = xi * q / r3 + alp3 * aj4 * sd * cd;
    du[4] = et * q / r3 + qy + alp3 * aj5 * sd * cd;
    du[5] = -q2 / r3 + alp3 * aj6 * sd * cd;
    du[6] = -c.ey + alp3 * aj1 * sd * cd;
    du[7] = -et * c.gy - xy * sd + alp3 * aj2 * sd * cd;
    du[8] = q * c.gy + alp3 * aj3 * sd * cd;
    du[9] = -c.ez - alp3 * ak3 * sd * cd;
    du[10] = -et * c.gz - xy * cd - alp3 * xi * d11 * sd * cd;
    du[11] = q * c.gz - alp3 * ak4 * sd * cd;
    for (int i = 0; i < 12; i++) dub[i] = dub[i] + disl2 / pi2 * du[i];
}

/* part C, okada.f90:476-597 */
static void ok_uc(double xi, double et, double q, double z, double sd, double cd, double disl1,
                  double disl2, double alpha, double *duc)
{
    const double pi2 = 8.0 * atan(1.0);
    double alp4 = 1.0 - alpha, alp5 = alpha;
    okc_t c; okc_set(&c, xi, et, q, sd, cd);
    double r = c.r, r2 = c.r2, r3 = c.r3, r5 = c.r5, q2 = c.q2, xi2 = c.xi2, et2 = c.et2, d = c.d, yp = c.yp;
    double x11 = c.x11, y11 = c.y11, x32 = c.x32, y32 = c.y32;
    double cc = d + z;
    double x53 = (8.0 * r2 + 9.0 * r * xi + 3.0 * xi2) * x11 * x11 * x11 / r2;
    double y53 = (8.0 * r2 + 9.0 * r * et + 3.0 * et2) * y11 * y11 * y11 / r2;
    double h = q * cd - z;
    double z32 = sd / r3 - h * y32;
    double z53 = 3.0 * sd / r5 - h * y53;
    double y0 = y11 - xi2 * y32;
    double z0 = z32 - xi2 * z53;
    double ppy = cd / r3 + q * y32 * sd;
    double ppz = sd / r3 - q * y32 * cd;
    double qq = z * y32 + z32 + z0;
    double qqy = 3.0 * cc * d / r5 - qq * sd;
    double qqz = 3.0 * cc * yp / r5 - qq * cd + q * y32;
    double xy = xi * y11, qy = q * y11;
    double qr = 3.0 * q / r5;
    double cdr = (cc + d) / r3;
    double yy0 = yp / r3 - y0 * cd;
    double du[12];
    du[0] = alp4 * xy * cd - alp5 * xi * q * z32;
    du[1] = alp4 * (cd / r + 2.0 * qy * sd) - alp5 * cc * q / r3;
    du[2] = alp4 * qy * cd - alp5 * (cc * et / r3 - z * y11 + xi2 * z32);
    du[3] = alp4 * y0 * cd - alp5 * q * z0;
    du[4] = -alp4 * xi * (cd / r3 + 2.0 * q * y32 * sd) + alp5 * cc * xi * qr;
    du[5] = -alp4 * xi * q * y32 * cd + alp5 * xi * (3.0 * cc * et / r5 - qq);
    du[6] = -alp4 * xi * ppy * cd - alp5 * xi * qqy;
    du[7] = alp4 * 2.0 * (d / r3 - y0 * sd) * sd - yp / r3 * cd - alp5 * (cdr * sd - et / r3 - cc * yp * qr);
    du[8] = -alp4 * q / r3 + yy0 * sd + alp5 * (cdr * cd + cc * d * qr - (y0 * cd + q * z0) * sd);
    du[9] = alp4 * xi * ppz * cd - alp5 * xi * qqz;
    du[10] = alp4 * 2.0 * (yp / r3 - y0 * cd) * sd + d / r3 * cd - alp5 * (cdr * cd + cc * d * qr);
    du[11] = yy0 * cd - alp5 * (cdr * sd - cc * yp * qr - y0 * sd * sd + q * z0 * cd);
    for (int i = 0; i < 12; i++) duc[i] = 0.0 + disl1 / pi2 * du[i];
    du[0] = alp4 * cd / r - qy * sd - alp5 * cc * q / r3;
    du[1] = alp4 * yp * x11 - alp5 * cc * et * q * x32;
    du[2] = -d * x11 - xy * sd - alp5 * cc * (x11 - q2 * x32);
    du[3] = -alp4 * xi / r3 * cd + alp5 * cc * xi * qr + xi * q * y32 * sd;
    du[4] = -alp4 * yp / r3 + alp5 * cc * et * qr;
    du[5] = d / r3 - y0 * sd + alp5 * cc / r3 * (1.0 - 3.0 * q2 / r2);
    du[6] = -alp4 * et / r3 + y0 * sd * sd - alp5 * (cdr * sd - cc * yp * qr);
    du[7] = alp4 * (x11 - yp * yp * x32) - alp5 * cc * ((d + 2.0 * q * cd) * x32 - yp * et * q * x53);
    du[8] = xi * ppy * sd + yp * d * x32 + alp5 * cc * ((yp + 2.0 * q * sd) * x32 - yp * q2 * x53);
    du[9] = -q / r3 + y0 * sd * cd - alp5 * (cdr * cd + cc * d * qr);
    du[10] = alp4 * yp * d * x32 - alp5 * cc * ((yp - 2.0 * q * sd) * x32 + d * et * q * x53);
    du[11] = -xi * ppz * sd + x11 - d * d * x32 - alp5 * cc * ((d - 2.0 * q * cd) * x32 - d * q2 * x53);
    for (int i = 0; i < 12; i++) duc[i] = duc[i] + disl2 / pi2 * du[i];
}

/* okada(), okada.f90:10-226: displacement gradients of a finite rectangle,
   strike/dip slip only (no tensile terms in the reference). */
void og_dc3d(double alpha, double x, double y, double z, double depth, double dip,
             double al1, double al2, double aw1, double aw2, double disl1, double disl2,
             int fullspace, double *u)
{
    const double pi2 = 8.0 * atan(1.0), epso = 1e-6;
    double p18 = pi2 / 360.0;
    double sd = sin(dip * p18), cd = cos(dip * p18);
    if (fabs(cd) < epso) {
        cd = 0.0;
        if (sd > 0.0) sd = 1.0;
        if (sd < 0.0) sd = -1.0;
    }
    double xi[2] = { x - al1, x - al2 };
    double d = depth + z;
    double p = y * cd + d * sd;
    double q = y * sd - d * cd;
    double et[2] = { p - aw1, p - aw2 };
    double dua[12], dub[12], duc[12], du[12];
    for (int i = 0; i < 12; i++) u[i] = 0.0;
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++) {
            ok_ua(xi[j], et[k], q, sd, cd, disl1, disl2, alpha, dua);
            du[0] = -dua[0];
            du[1] = -dua[1] * cd + dua[2] * sd;
            du[2] = -dua[1] * sd - dua[2] * cd;
            du[3] = -dua[3];
            du[4] = -dua[4] * cd + dua[5] * sd;
            du[5] = -dua[4] * sd - dua[5] * cd;
            du[6] = -dua[6];
            du[7] = -dua[7] * cd + dua[8] * sd;
            du[8] = -dua[7] * sd - dua[8] * cd;
            du[9] = dua[9];
            du[10] = dua[10] * cd - dua[11] * sd;
            du[11] = dua[10] * sd + dua[11] * cd;
            for (int i = 0; i < 12; i++) u[i] = ((j + k) != 1) ? u[i] + du[i] : u[i] - du[i];
        }
    if (fullspace) return;
    d = depth - z;
    p = y * cd + d * sd;
    q = y * sd - d * cd;
    et[0] = p - aw1;
    et[1] = p - aw2;
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++) {
            ok_ua(xi[j], et[k], q, sd, cd, disl1, disl2, alpha, dua);
            ok_ub(xi[j], et[k], q, sd, cd, disl1, disl2, alpha, dub);
            ok_uc(xi[j], et[k], q, z, sd, cd, disl1, disl2, alpha, duc);
            for (int b = 0; b < 12; b += 3) {
                du[b] = dua[b] + dub[b] + z * duc[b];
                du[b + 1] = (dua[b + 1] + dub[b + 1] + z * duc[b + 1]) * cd - (dua[b + 2] + dub[b + 2] + z * duc[b + 2]) * sd;
                du[b + 2] = (dua[b + 1] + dub[b + 1] - z * duc[b + 1]) * sd + (dua[b + 2] + dub[b + 2] - z * duc[b + 2]) * cd;
            }
            du[9] = du[9] + duc[0];
            du[10] = du[10] + duc[1] * cd - duc[2] * sd;
            du[11] = du[11] - duc[1] * sd - duc[2] * cd;
            for (int i = 0; i < 12; i++) u[i] = ((j + k) != 1) ? u[i] + du[i] : u[i] - du[i];
        }
}

/* TensTrans, TDstressFS.f90:295-314; A[r][c] is the Fortran A(r+1,c+1) */
static void og_tenstrans(const double *t, const double A[3][3], double *o)
{
    double Txx = t[0], Tyy = t[1], Tzz = t[2], Txy = t[3], Txz = t[4], Tyz = t[5];
    double a11 = A[0][0], a12 = A[0][1], a13 = A[0][2];
    double a21 = A[1][0], a22 = A[1][1], a23 = A[1][2];
    double a31 = A[2][0], a32 = A[2][1], a33 = A[2][2];
    o[0] = a11 * a11 * Txx + 2 * a11 * a21 * Txy + 2 * a11 * a31 * Txz + 2 * a21 * a31 * Tyz + a21 * a21 * Tyy + a31 * a31 * Tzz;
    o[1] = a12 * a12 * Txx + 2 * a12 * a22 * Txy + 2 * a12 * a32 * Txz + 2 * a22 * a32 * Tyz + a22 * a22 * Tyy + a32 * a32 * Tzz;
    o[2] = a13 * a13 * Txx + 2 * a13 * a23 * Txy + 2 * a13 * a33 * Txz + 2 * a23 * a33 * Tyz + a23 * a23 * Tyy + a33 * a33 * Tzz;
    o[3] = a11 * a12 * Txx + (a11 * a22 + a21 * a12) * Txy + (a11 * a32 + a12 * a31) * Txz + (a32 * a21 + a31 * a22) * Tyz + a22 * a21 * Tyy + a31 * a32 * Tzz;
    o[4] = a11 * a13 * Txx + (a11 * a23 + a13 * a21) * Txy + (a11 * a33 + a31 * a13) * Txz + (a33 * a21 + a31 * a23) * Tyz + a23 * a21 * Tyy + a31 * a33 * Tzz;
    o[5] = a12 * a13 * Txx + (a13 * a22 + a12 * a23) * Txy + (a13 * a32 + a12 * a33) * Txz + (a23 * a32 + a33 * a22) * Tyz + a22 * a23 * Tyy + a32 * a33 * Tzz;
}

/* okada_ij, m_HACApK_calc_entry_ij.f90:1065-1145 (i receiver, j source, 0-based) */
static double og_okada_ij(const og_geom *g, int i, int j, int fullspace)
{
    double pi = OG_PI;
    double alpha = (1.0 + (0.5 / og_pois - 1.0)) / (1.0 + 2.0 * (0.5 / og_pois - 1.0));
    double dx = cos(g->ang[j]) * (g->xcol[i] - g->xcol[j]) + sin(g->ang[j]) * (g->ycol[i] - g->ycol[j]);
    double dy = -sin(g->ang[j]) * (g->xcol[i] - g->xcol[j]) + cos(g->ang[j]) * (g->ycol[i] - g->ycol[j]);
    double w = g->w;
    double z = g->zcol[i];
    double depth = -g->zcol[j];
    double dip = g->angd[j] * 180.0 / pi;
    double rs = g->dsl[j] / w, rd = g->dsd[j] / w;
    double u[12];
    og_dc3d(alpha, dx / w, dy / w, z / w, depth / w, dip, -0.5 * rs, 0.5 * rs, -0.5 * rd, 0.5 * rd,
            -cos(g->rake[j]), -sin(g->rake[j]), fullspace, u);
    double uxx = u[3], uyx = u[4], uzx = u[5], uxy = u[6], uyy = u[7], uzy = u[8], uxz = u[9], uyz = u[10], uzz = u[11];
    double exx = uxx, eyy = uyy, ezz = uzz;
    double exy = 0.5 * (uxy + uyx), eyz = 0.5 * (uyz + uzy), ezx = 0.5 * (uzx + uxz);
    double R = og_rigid;
    double s[6];
    s[0] = R * (exx + eyy + ezz) + 2 * R * exx;
    s[1] = R * (exx + eyy + ezz) + 2 * R * eyy;
    s[2] = R * (exx + eyy + ezz) + 2 * R * ezz;
    s[3] = 2 * R * exy;
    s[4] = 2 * R * ezx;
    s[5] = 2 * R * eyz;
    double rotang = g->ang[j] - g->ang[i];
    double dpang = g->angd[i];
    double A[3][3];
    A[0][0] = cos(rotang);               A[1][0] = -sin(rotang);              A[2][0] = 0.0;
    A[0][1] = sin(rotang) * cos(dpang);  A[1][1] = cos(rotang) * cos(dpang);  A[2][1] = sin(dpang);
    A[0][2] = sin(rotang) * sin(dpang);  A[1][2] = cos(rotang) * sin(dpang);  A[2][2] = -cos(dpang);
    double p[6];
    og_tenstrans(s, A, p);
    if (g->vcomp == 1) return p[2];
    return p[4] * cos(g->rake[i]) + p[5] * sin(g->rake[i]);
}

/* ------------------------------------------------------------------------- */
/* TD full-space stress (Nikkhoo & Walter 2015), TDstressFS.f90              */
/* ------------------------------------------------------------------------- */
static void og_cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - b[1] * a[2];
    c[1] = a[2] * b[0] - b[2] * a[0];
    c[2] = a[0] * b[1] - b[0] * a[1];
}
static double og_dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* AngDisStrain, TDstressFS.f90:560-627 */
static void og_angdisstrain(double x, double y, double z, double alpha, double bx, double by, double bz,
                            double nu, double *e)
{
    const double pi = OG_PI;
    double sinA = sin(alpha), cosA = cos(alpha);
    double eta = y * cosA - z * sinA;
    double zeta = y * sinA + z * cosA;
    double x2 = x * x, y2 = y * y, z2 = z * z;
    double r2 = x2 + y2 + z2;
    double r = sqrt(r2);
    double r3 = r * r2;
    double rz = r * (r - z);
    double r2z2 = r2 * ((r - z) * (r - z));
    double r3z = r3 * (r - z);
    double W = zeta - r;
    double W2 = W * W;
    double Wr = W * r;
    double W2r = W2 * r;
    double Wr3 = W * r3;
    double W2r2 = W2 * r2;
    double C = (r * cosA - z) / Wr;
    double S = (r * sinA - y) / Wr;
    double rFi_rx = (eta / r / (r - zeta) - y / r / (r - z)) / 4 / pi;
    double rFi_ry = (x / r / (r - z) - cosA * x / r / (r - zeta)) / 4 / pi;
    double rFi_rz = (sinA * x / r / (r - zeta)) / 4 / pi;
    double Exx = bx * (rFi_rx) + bx / 8 / pi / (1 - nu) * (eta / Wr + eta * x2 / W2r2 - eta * x2 / Wr3 + y / rz -
        x2 * y / r2z2 - x2 * y / r3z) - by * x / 8 / pi / (1 - nu) * (((2 * nu + 1) / Wr + x2 / W2r2 - x2 / Wr3) * cosA +
        (2 * nu + 1) / rz - x2 / r2z2 - x2 / r3z) + bz * x * sinA / 8 / pi / (1 - nu) * ((2 * nu + 1) / Wr + x2 / W2r2 - x2 / Wr3);
    double Eyy = by * (rFi_ry) + bx / 8 / pi / (1 - nu) * ((1 / Wr + S * S - y2 / Wr3) * eta + (2 * nu + 1) * y / rz - y * y * y / r2z2 -
        y * y * y / r3z - 2 * nu * cosA * S) - by * x / 8 / pi / (1 - nu) * (1 / rz - y2 / r2z2 - y2 / r3z +
        (1 / Wr + S * S - y2 / Wr3) * cosA) + bz * x * sinA / 8 / pi / (1 - nu) * (1 / Wr + S * S - y2 / Wr3);
    double Ezz = bz * (rFi_rz) + bx / 8 / pi / (1 - nu) * (eta / W / r + eta * C * C - eta * z2 / Wr3 + y * z / r3 +
        2 * nu * sinA * C) - by * x / 8 / pi / (1 - nu) * ((1 / Wr + C * C - z2 / Wr3) * cosA + z / r3) +
        bz * x * sinA / 8 / pi / (1 - nu) * (1 / Wr + C * C - z2 / Wr3);
    double Exy = bx * (rFi_ry) / 2 + by * (rFi_rx) / 2 - bx / 8 / pi / (1 - nu) * (x * y2 / r2z2 - nu * x / rz +
        x * y2 / r3z - nu * x * cosA / Wr + eta * x * S / Wr + eta * x * y / Wr3) +
        by / 8 / pi / (1 - nu) * (x2 * y / r2z2 - nu * y / rz + x2 * y / r3z + nu * cosA * S +
        x2 * y * cosA / Wr3 + x2 * cosA * S / Wr) - bz * sinA / 8 / pi / (1 - nu) * (nu * S + x2 * S / Wr + x2 * y / Wr3);
    double Exz = bx * (rFi_rz) / 2 + bz * (rFi_rx) / 2 - bx / 8 / pi / (1 - nu) * (-x * y / r3 + nu * x * sinA / Wr + eta * x * C / Wr +
        eta * x * z / Wr3) + by / 8 / pi / (1 - nu) * (-x2 / r3 + nu / r + nu * cosA * C + x2 * z * cosA / Wr3 +
        x2 * cosA * C / Wr) - bz * sinA / 8 / pi / (1 - nu) * (nu * C + x2 * C / Wr + x2 * z / Wr3);
    double Eyz = by * (rFi_rz) / 2 + bz * (rFi_ry) / 2 + bx / 8 / pi / (1 - nu) * (y2 / r3 - nu / r - nu * cosA * C +
        nu * sinA * S + eta * sinA * cosA / W2 - eta * (y * cosA + z * sinA) / W2r + eta * y * z / W2r2 - eta * y * z / Wr3) -
        by * x / 8 / pi / (1 - nu) * (y / r3 + sinA * (cosA * cosA) / W2 - cosA * (y * cosA + z * sinA) /
        W2r + y * z * cosA / W2r2 - y * z * cosA / Wr3) - bz * x * sinA / 8 / pi / (1 - nu) *
        (y * z / Wr3 - sinA * cosA / W2 + (y * cosA + z * sinA) / W2r - y * z / W2r2);
    (void)W2r;
    e[0] = Exx; e[1] = Eyy; e[2] = Ezz; e[3] = Exy; e[4] = Exz; e[5] = Eyz;
}

/* TDSetupD, TDstressFS.f90:352-403 */
static void og_tdsetupd(double x, double y, double z, double alpha, double bx, double by, double bz, double nu,
                        const double *tv, const double *sv, double *e)
{
    const double pi = OG_PI;
    double A11 = sv[2], A12 = -sv[1], A21 = sv[1], A22 = sv[2];
    double y1 = A11 * (y - tv[1]) + A12 * (z - tv[2]);
    double z1 = A21 * (y - tv[1]) + A22 * (z - tv[2]);
    double by1 = A11 * by + A12 * bz;
    double bz1 = A21 * by + A22 * bz;
    double ea[6];
    og_angdisstrain(x, y1, z1, -pi + alpha, bx, by1, bz1, nu, ea);
    double B[3][3] = { { 1.0, 0.0, 0.0 }, { 0.0, A11, A12 }, { 0.0, A21, A22 } };
    og_tenstrans(ea, B, e);
}

/* trimodefinder, TDstressFS.f90:329-350 */
static int og_trimode(double x, double y, double z, const double *p1, const double *p2, const double *p3)
{
    double den = (p2[1] - p3[1]) * (p1[0] - p3[0]) + (p3[0] - p2[0]) * (p1[1] - p3[1]);
    double a = ((p2[1] - p3[1]) * (x - p3[0]) + (p3[0] - p2[0]) * (y - p3[1])) / den;
    double b = ((p3[1] - p1[1]) * (x - p3[0]) + (p1[0] - p3[0]) * (y - p3[1])) / den;
    double c = 1.0 - a - b;
    int m = 1;
    if ((a <= 0.0) && (b > c) && (c > a)) m = -1;
    if ((b <= 0.0) && (c > a) && (a > b)) m = -1;
    if ((c <= 0.0) && (a > b) && (b > c)) m = -1;
    if ((a == 0.0) && (b >= 0.0) && (c >= 0.0)) m = 0;
    if ((a >= 0.0) && (b == 0.0) && (c >= 0.0)) m = 0;
    if ((a >= 0.0) && (b >= 0.0) && (c == 0.0)) m = 0;
    if ((m == 0) && (z != 0.0)) m = 1;
    return m;
}

/* TDstressFS, TDstressFS.f90:45-235: stress (Sxx,Syy,Szz,Sxy,Sxz,Syz) */
void og_tdstressfs(double X, double Y, double Z, const double *P1, const double *P2, const double *P3,
                   double Ss, double Ds, double Ts, double mu, double lambda, double *S)
{
    double nu = 1 / (1 + lambda / mu) / 2;
    double bx = Ts, by = Ss, bz = Ds;
    double P2mP1[3], P3mP1[3], Vnorm[3], Vstrike[3], Vdip[3];
    for (int k = 0; k < 3; k++) { P2mP1[k] = P2[k] - P1[k]; P3mP1[k] = P3[k] - P1[k]; }
    og_cross(P2mP1, P3mP1, Vnorm);
    double nn = sqrt(og_dot3(Vnorm, Vnorm));
    for (int k = 0; k < 3; k++) Vnorm[k] = Vnorm[k] / nn;
    const double eY[3] = { 0.0, 1.0, 0.0 }, eZ[3] = { 0.0, 0.0, 1.0 };
    og_cross(eZ, Vnorm, Vstrike);
    nn = sqrt(og_dot3(Vstrike, Vstrike));
    for (int k = 0; k < 3; k++) Vstrike[k] = Vstrike[k] / nn;
    if (og_dot3(Vstrike, Vstrike) == 0.0)
        for (int k = 0; k < 3; k++) Vstrike[k] = eY[k] * Vnorm[2];
    nn = sqrt(og_dot3(Vstrike, Vstrike));
    for (int k = 0; k < 3; k++) Vstrike[k] = Vstrike[k] / nn;
    og_cross(Vnorm, Vstrike, Vdip);
    double At[3][3];
    for (int k = 0; k < 3; k++) { At[0][k] = Vnorm[k]; At[1][k] = Vstrike[k]; At[2][k] = Vdip[k]; }
#define OG_CT(a0, a1, a2, o) do { (o)[0] = At[0][0] * (a0) + At[0][1] * (a1) + At[0][2] * (a2); \
        (o)[1] = At[1][0] * (a0) + At[1][1] * (a1) + At[1][2] * (a2); \
        (o)[2] = At[2][0] * (a0) + At[2][1] * (a1) + At[2][2] * (a2); } while (0)
    double xyz[3], p_1[3], p_2[3] = { 0.0, 0.0, 0.0 }, p_3[3];
    OG_CT(X - P2[0], Y - P2[1], Z - P2[2], xyz);
    OG_CT(P1[0] - P2[0], P1[1] - P2[1], P1[2] - P2[2], p_1);
    OG_CT(P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2], p_3);
#undef OG_CT
    double x_ = xyz[0], y_ = xyz[1], z_ = xyz[2];
    double e12[3], e13[3], e23[3];
    for (int k = 0; k < 3; k++) { e12[k] = p_2[k] - p_1[k]; e13[k] = p_3[k] - p_1[k]; e23[k] = p_3[k] - p_2[k]; }
    nn = sqrt(og_dot3(e12, e12)); for (int k = 0; k < 3; k++) e12[k] = e12[k] / nn;
    nn = sqrt(og_dot3(e13, e13)); for (int k = 0; k < 3; k++) e13[k] = e13[k] / nn;
    nn = sqrt(og_dot3(e23, e23)); for (int k = 0; k < 3; k++) e23[k] = e23[k] / nn;
    double me12[3] = { -e12[0], -e12[1], -e12[2] }, me13[3] = { -e13[0], -e13[1], -e13[2] }, me23[3] = { -e23[0], -e23[1], -e23[2] };
    double A = acos(og_dot3(e12, e13));
    double B = acos(og_dot3(me12, e23));
    double C = acos(og_dot3(e23, e13));
    int mode = og_trimode(y_, z_, x_, p_1 + 1, p_2 + 1, p_3 + 1);
    double e1[6], e2[6], e3[6], e[6];
    if (mode == 1) {
        og_tdsetupd(x_, y_, z_, A, bx, by, bz, nu, p_1, me13, e1);
        og_tdsetupd(x_, y_, z_, B, bx, by, bz, nu, p_2, e12, e2);
        og_tdsetupd(x_, y_, z_, C, bx, by, bz, nu, p_3, e23, e3);
    } else {
        og_tdsetupd(x_, y_, z_, A, bx, by, bz, nu, p_1, e13, e1);
        og_tdsetupd(x_, y_, z_, B, bx, by, bz, nu, p_2, me12, e2);
        og_tdsetupd(x_, y_, z_, C, bx, by, bz, nu, p_3, me23, e3);
    }
    for (int k = 0; k < 6; k++) e[k] = e1[k] + e2[k] + e3[k];
    if (mode == 0)
        for (int k = 0; k < 6; k++) e[k] = NAN;
    double E[6];
    og_tenstrans(e, At, E);
    S[0] = 2 * mu * E[0] + lambda * (E[0] + E[1] + E[2]);
    S[1] = 2 * mu * E[1] + lambda * (E[0] + E[1] + E[2]);
    S[2] = 2 * mu * E[2] + lambda * (E[0] + E[1] + E[2]);
    S[3] = 2 * mu * E[3];
    S[4] = 2 * mu * E[4];
    S[5] = 2 * mu * E[5];
}

/* ---- half-space triangles ------------------------------------------------
 * The free-surface correction of an angular dislocation in a half-space
 * (Nikkhoo & Walter 2015, GJI 201:1117-1139; the strain routine is
 * AngDisStrainFSC, TDstressFS.f90:629-1011) is the gradient of a harmonic
 * displacement field.  Here the displacement is written out and its gradient
 * taken by the complex-step method: u(y + i h e_k) = u(y) + i h du/dy_k + O(h^2)
 * for the real-analytic terms (products, quotients, sqrt, log of positive
 * arguments), so Im(u)/h is the derivative to rounding with h far below it.
 * The Burgers function enters only through its gradient, supplied in closed
 * form (its value carries a branch cut, its gradient does not).
 */
typedef double complex og_c;

static void og_fsc_disp(og_c y1, og_c y2, og_c y3, og_c fib, double beta, double b1, double b2, double b3,
                        double nu, double a, og_c *u)
{
    const double pi = OG_PI;
    double sinB = sin(beta), cosB = cos(beta), cotB = cosB / sinB, cot2 = cotB * cotB;
    double N1 = 1 - 2 * nu, f = 1 / (4 * pi * (1 - nu));
    og_c y3b = y3 + 2 * a;
    og_c z1b = y1 * cosB + y3b * sinB;
    og_c z3b = -y1 * sinB + y3b * cosB;
    og_c rb = csqrt(y1 * y1 + y2 * y2 + y3b * y3b);
    og_c ar = a / rb, ar2 = a / (rb * rb);
    og_c p6 = rb + y3b;            /* r + y3bar */
    og_c p7 = rb + z3b;            /* r + z3bar */
    og_c h = y3b - a;              /* y3 + a */
    og_c q = rb * cosB + y3b;
    og_c lg6 = clog(p6), lg7 = clog(p7);
    og_c u1 = 0, u2 = 0, u3 = 0;
    if (b1 != 0.0) {
        u1 += b1 * f * (-2 * (1 - nu) * N1 * fib * cot2 + N1 * y2 / p6 * ((N1 - ar) * cotB - y1 / p6 * (nu + ar))
              + N1 * y2 * cosB * cotB / p7 * (cosB + ar) + a * y2 * h * cotB / (rb * rb * rb)
              + y2 * h / (rb * p6) * (-N1 * cotB + y1 / p6 * (2 * nu + ar) + a * y1 / (rb * rb))
              + y2 * h / (rb * p7) * (cosB / p7 * (q * (N1 * cosB - ar) * cotB + 2 * (1 - nu) * (rb * sinB - y1) * cosB)
                                      - a * y3b * cosB * cotB / (rb * rb)));
        u2 += b1 * f * (N1 * ((2 * (1 - nu) * cot2 - nu) * lg6 - (2 * (1 - nu) * cot2 + N1) * cosB * lg7)
              - N1 / p6 * (y1 * cotB * (N1 - ar) + nu * y3b - a + y2 * y2 / p6 * (nu + ar))
              - N1 * z1b * cotB / p7 * (cosB + ar) - a * y1 * h * cotB / (rb * rb * rb)
              + h / p6 * (-2 * nu + (N1 * y1 * cotB - a) / rb + y2 * y2 / (rb * p6) * (2 * nu + ar) + a * y2 * y2 / (rb * rb * rb))
              + h / p7 * (cosB * cosB - (N1 * z1b * cotB + a * cosB) / rb + a * y3b * z1b * cotB / (rb * rb * rb)
                          - (y2 * y2 * cosB * cosB - a * z1b * cotB / rb * q) / (rb * p7)));
        u3 += b1 * f * (2 * (1 - nu) * (N1 * fib * cotB + y2 / p6 * (2 * nu + ar) - y2 * cosB / p7 * (cosB + ar))
              + y2 * h / rb * (2 * nu / p6 + ar2) + y2 * h * cosB / (rb * p7) * (N1 - q / p7 * (cosB + ar) - a * y3b / (rb * rb)));
    }
    if (b2 != 0.0) {
        u1 += b2 * f * (N1 * ((2 * (1 - nu) * cot2 + nu) * lg6 - (2 * (1 - nu) * cot2 + 1) * cosB * lg7)
              + N1 / p6 * (-N1 * y1 * cotB + nu * y3b - a + a * y1 * cotB / rb + y1 * y1 / p6 * (nu + ar))
              - N1 * cotB / p7 * (z1b * cosB - a * (rb * sinB - y1) / (rb * cosB)) - a * y1 * h * cotB / (rb * rb * rb)
              + h / p6 * (2 * nu + (N1 * y1 * cotB + a) / rb - y1 * y1 / (rb * p6) * (2 * nu + ar) - a * y1 * y1 / (rb * rb * rb))
              + h * cotB / p7 * (-cosB * sinB + a * y1 * y3b / (rb * rb * rb * cosB)
                                 + (rb * sinB - y1) / rb * (2 * (1 - nu) * cosB - q / p7 * (1 + a / (rb * cosB)))));
        u2 += b2 * f * (2 * (1 - nu) * N1 * fib * cot2 + N1 * y2 / p6 * (-(N1 - ar) * cotB + y1 / p6 * (nu + ar))
              - N1 * y2 * cotB / p7 * (1 + a / (rb * cosB)) - a * y2 * h * cotB / (rb * rb * rb)
              + y2 * h / (rb * p6) * (N1 * cotB - 2 * nu * y1 / p6 - a * y1 / rb * (1 / rb + 1 / p6))
              + y2 * h * cotB / (rb * p7) * (-2 * (1 - nu) * cosB + q / p7 * (1 + a / (rb * cosB)) + a * y3b / (rb * rb * cosB)));
        u3 += b2 * f * (-2 * (1 - nu) * N1 * cotB * (lg6 - cosB * lg7) - 2 * (1 - nu) * y1 / p6 * (2 * nu + ar)
              + 2 * (1 - nu) * z1b / p7 * (cosB + ar) + h / rb * (N1 * cotB - 2 * nu * y1 / p6 - a * y1 / (rb * rb))
              - h / p7 * (cosB * sinB + q * cotB / rb * (2 * (1 - nu) * cosB - q / p7)
                          + ar * (sinB - y3b * z1b / (rb * rb) - z1b * q / (rb * p7))));
    }
    if (b3 != 0.0) {
        u1 += b3 * f * (N1 * (y2 / p6 * (1 + ar) - y2 * cosB / p7 * (cosB + ar)) - y2 * h / rb * (ar2 + 1 / p6)
              + y2 * h * cosB / (rb * p7) * (q / p7 * (cosB + ar) + a * y3b / (rb * rb)));
        u2 += b3 * f * (N1 * (-sinB * lg7 - y1 / p6 * (1 + ar) + z1b / p7 * (cosB + ar)) + y1 * h / rb * (ar2 + 1 / p6)
              - h / p7 * (sinB * (cosB - ar) + z1b / rb * (1 + a * y3b / (rb * rb))
                          - (y2 * y2 * cosB * sinB - a * z1b / rb * q) / (rb * p7)));
        u3 += b3 * f * (2 * (1 - nu) * fib + 2 * (1 - nu) * y2 * sinB / p7 * (cosB + ar)
              + y2 * h * sinB / (rb * p7) * (1 + q / p7 * (cosB + ar) + a * y3b / (rb * rb)));
    }
    u[0] = u1; u[1] = u2; u[2] = u3;
}

/* free-surface correction strains v11 v22 v33 v12 v13 v23 of an angular
 * dislocation (the quantity of AngDisStrainFSC, TDstressFS.f90:629-1011) */
void og_angfsc(double y1, double y2, double y3, double beta, double b1, double b2, double b3, double nu, double a,
               double *v)
{
    const double hstep = 1e-60;
    double y3b = y3 + 2 * a;
    double rb = sqrt(y1 * y1 + y2 * y2 + y3b * y3b);
    double c = cos(beta / 2) / sin(beta / 2);
    double D = y1 - (rb + y3b) * c, den = D * D + y2 * y2;
    /* gradient of 2 atan(-y2 / (y1 - (r + y3bar) cot(beta/2))) */
    double gf[3] = { 2 * y2 * (1 - c * y1 / rb) / den, -2 * (D + c * y2 * y2 / rb) / den,
                     -2 * y2 * c * (y3b / rb + 1) / den };
    double du[3][3];   /* du[m][k] = d u_m / d y_k */
    for (int k = 0; k < 3; k++) {
        og_c yy[3] = { y1, y2, y3 };
        yy[k] += I * hstep;
        og_c u[3];
        og_fsc_disp(yy[0], yy[1], yy[2], I * hstep * gf[k], beta, b1, b2, b3, nu, a, u);
        for (int m = 0; m < 3; m++) du[m][k] = cimag(u[m]) / hstep;
    }
    v[0] = du[0][0];
    v[1] = du[1][1];
    v[2] = du[2][2];
    v[3] = (du[0][1] + du[1][0]) / 2;
    v[4] = (du[0][2] + du[2][0]) / 2;
    v[5] = (du[1][2] + du[2][1]) / 2;
}

/* AngSetupFSC_S, TDstressFS.f90:405-558: free-surface correction stress of
 * the side PA -> PB (zero for a vertical side) */
static void og_angsetup_fsc(double X, double Y, double Z, double bX, double bY, double bZ, const double *PA,
                            const double *PB, double mu, double lambda, double *S)
{
    const double pi = OG_PI;
    double nu = lambda / (lambda + mu) / 2;
    double sv[3] = { PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2] };
    double beta = acos(-sv[2] / sqrt(og_dot3(sv, sv)));
    for (int k = 0; k < 6; k++) S[k] = 0.0;
    if (fabs(beta) < 1e-3 || fabs(pi - beta) < 1e-3) return;
    double e1[3] = { sv[0], sv[1], 0.0 };
    double nn = sqrt(og_dot3(e1, e1));
    for (int k = 0; k < 3; k++) e1[k] = e1[k] / nn;
    const double e3[3] = { 0.0, 0.0, -1.0 };
    double e2[3];
    og_cross(e3, e1, e2);
    double A[3][3];   /* rows: the angular-dislocation axes */
    for (int k = 0; k < 3; k++) { A[0][k] = e1[k]; A[1][k] = e2[k]; A[2][k] = e3[k]; }
    double ya[3], yab[3], bb[3];
    double rel[3] = { X - PA[0], Y - PA[1], Z - PA[2] }, bv[3] = { bX, bY, bZ };
    for (int r = 0; r < 3; r++) {
        ya[r] = A[r][0] * rel[0] + A[r][1] * rel[1] + A[r][2] * rel[2];
        yab[r] = A[r][0] * sv[0] + A[r][1] * sv[1] + A[r][2] * sv[2];
        bb[r] = A[r][0] * bv[0] + A[r][1] * bv[1] + A[r][2] * bv[2];
    }
    /* configuration I or II (the artefact-free one for the point) */
    double ang = (beta * ya[0] >= 0.0) ? -pi + beta : beta;
    double vA[6], vB[6], v[6], E[6];
    og_angfsc(ya[0], ya[1], ya[2], ang, bb[0], bb[1], bb[2], nu, -PA[2], vA);
    og_angfsc(ya[0] - yab[0], ya[1] - yab[1], ya[2] - yab[2], ang, bb[0], bb[1], bb[2], nu, -PB[2], vB);
    for (int k = 0; k < 6; k++) v[k] = vB[k] - vA[k];
    og_tenstrans(v, A, E);
    double tr = E[0] + E[1] + E[2];
    S[0] = 2 * mu * E[0] + lambda * tr;
    S[1] = 2 * mu * E[1] + lambda * tr;
    S[2] = 2 * mu * E[2] + lambda * tr;
    S[3] = 2 * mu * E[3];
    S[4] = 2 * mu * E[4];
    S[5] = 2 * mu * E[5];
}

/* TDstress_HarFunc, TDstressFS.f90:236-293: the harmonic part, one term per side */
void og_td_harfunc(double X, double Y, double Z, const double *P1, const double *P2, const double *P3, double Ss,
                   double Ds, double Ts, double mu, double lambda, double *S)
{
    double a[3], b[3], vn[3], vs[3], vd[3];
    for (int k = 0; k < 3; k++) { a[k] = P2[k] - P1[k]; b[k] = P3[k] - P1[k]; }
    og_cross(a, b, vn);
    double nn = sqrt(og_dot3(vn, vn));
    for (int k = 0; k < 3; k++) vn[k] = vn[k] / nn;
    const double eY[3] = { 0.0, 1.0, 0.0 }, eZ[3] = { 0.0, 0.0, 1.0 };
    og_cross(eZ, vn, vs);
    /* normalised before the zero test, as there: a horizontal triangle gives NaN */
    nn = sqrt(og_dot3(vs, vs));
    for (int k = 0; k < 3; k++) vs[k] = vs[k] / nn;
    if (og_dot3(vs, vs) == 0.0)
        for (int k = 0; k < 3; k++) vs[k] = eY[k] * vn[2];
    nn = sqrt(og_dot3(vs, vs));
    for (int k = 0; k < 3; k++) vs[k] = vs[k] / nn;
    og_cross(vn, vs, vd);
    /* slip (tensile, strike, dip) in the triangle frame -> global */
    double bg[3];
    for (int k = 0; k < 3; k++) bg[k] = vn[k] * Ts + vs[k] * Ss + vd[k] * Ds;
    double s1[6], s2[6], s3[6];
    og_angsetup_fsc(X, Y, Z, bg[0], bg[1], bg[2], P1, P2, mu, lambda, s1);
    og_angsetup_fsc(X, Y, Z, bg[0], bg[1], bg[2], P2, P3, mu, lambda, s2);
    og_angsetup_fsc(X, Y, Z, bg[0], bg[1], bg[2], P3, P1, mu, lambda, s3);
    for (int k = 0; k < 6; k++) S[k] = s1[k] + s2[k] + s3[k];
}

/* TDstressHS, TDstressFS.f90:7-44: main dislocation + harmonic part + image */
void og_tdstresshs(double X, double Y, double Z, const double *P1, const double *P2, const double *P3,
                   double Ss, double Ds, double Ts, double mu, double lambda, double *S)
{
    double Sm[6], Sh[6], Si[6];
    og_tdstressfs(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lambda, Sm);
    og_td_harfunc(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lambda, Sh);
    double Q1[3] = { P1[0], P1[1], -P1[2] }, Q2[3] = { P2[0], P2[1], -P2[2] }, Q3[3] = { P3[0], P3[1], -P3[2] };
    og_tdstressfs(X, Y, Z, Q1, Q2, Q3, Ss, Ds, Ts, mu, lambda, Si);
    for (int k = 0; k < 6; k++) S[k] = Sm[k] + Si[k] + Sh[k];
}

/* matel3dn_ij / matel3dh_ij, m_HACApK_calc_entry_ij.f90:802-898: the
 * full-space kernel takes the receiver's rake for mode 's', the half-space
 * kernel the source's */
static double og_tri_ij(const og_geom *g, int i, int j)
{
    double lambda = 2 * og_pois / (1 - 2 * og_pois) * og_rigid;
    double P1[3] = { g->xs1[j], g->ys1[j], g->zs1[j] };
    double P2[3] = { g->xs2[j], g->ys2[j], g->zs2[j] };
    double P3[3] = { g->xs3[j], g->ys3[j], g->zs3[j] };
    double ss = 0.0, dd = 0.0, tt = 0.0;
    switch (g->md) {
    case 1: ss = 1.0; break;
    case 2: dd = 1.0; break;
    case 3: tt = 1.0; break;
    default: {
        int r = g->kind == 5 ? j : i;
        ss = cos(g->rake[r]); dd = sin(g->rake[r]); break;
    }
    }
    double S[6];
    if (g->kind == 5)
        og_tdstresshs(g->xcol[i], g->ycol[i], g->zcol[i], P1, P2, P3, ss, dd, tt, og_rigid, lambda, S);
    else
        og_tdstressfs(g->xcol[i], g->ycol[i], g->zcol[i], P1, P2, P3, ss, dd, tt, og_rigid, lambda, S);
    double A[3][3] = { { g->ev11[i], g->ev21[i], g->ev31[i] },
                       { g->ev12[i], g->ev22[i], g->ev32[i] },
                       { g->ev13[i], g->ev23[i], g->ev33[i] } };
    double p[6];
    og_tenstrans(S, A, p);
    if (g->vcomp == 1) return -p[2];
    return p[4] * cos(g->rake[i]) + p[5] * sin(g->rake[i]);
}

/* ---- 2-D, 2.5-D and planar 3-D kernels ------------------------------------ */

/* matels2dp_ij, m_HACApK_calc_entry_ij.f90:648-661 */
static double og_2dp_ij(const og_geom *g, int i, int j)
{
    double factor = og_rigid / (2.0 * OG_PI * (1.0 - og_pois));
    return factor * (1.0 / (g->xcol[i] - g->xer[j]) - 1.0 / (g->xcol[i] - g->xel[j]));
}

/* matel3dp_ij + ret3dp, m_HACApK_calc_entry_ij.f90:412-421, 440-470 */
static double og_3dp_ij(const og_geom *g, int i, int j)
{
    double dx[4] = { g->xcol[i] - g->xs1[j], g->xcol[i] - g->xs2[j], g->xcol[i] - g->xs3[j], g->xcol[i] - g->xs4[j] };
    double dz[4] = { g->zcol[i] - g->zs1[j], g->zcol[i] - g->zs2[j], g->zcol[i] - g->zs3[j], g->zcol[i] - g->zs4[j] };
    double dr[4];
    for (int k = 0; k < 4; k++) dr[k] = sqrt(dx[k] * dx[k] + dz[k] * dz[k]);
    double factor = og_rigid / (4.0 * OG_PI), alpha = 1 / (2 - 2 * og_pois);
    double t1 = 2.0 * (1 - alpha) * (dr[3] / (dx[3] * dz[3]) + dr[1] / (dx[1] * dz[1]) - dr[2] / (dx[2] * dz[2]) -
                                     dr[0] / (dx[0] * dz[0]));
    if (g->md == 0 && g->vcomp == 0) {
        double t2 = -(1 - 2 * alpha) / dx[0] * (dr[3] / dz[3] + dz[3] / dr[3] - dr[0] / dz[0] - dz[0] / dr[0]);
        double t3 = -(1 - 2 * alpha) / dx[1] * (dr[1] / dz[1] + dz[1] / dr[1] - dr[2] / dz[2] - dz[2] / dr[2]);
        return factor * (t1 + t2 + t3);
    }
    if (g->md == 3 && g->vcomp == 1) return factor * t1;
    return 0.0;
}

/* D2dip2, m_HACApK_calc_entry_ij.f90:143-166: edge dislocation (s1, s2) at
 * (xs, ys) in the half-space y > 0, stress at (X, Y) */
static void og_d2dip2(double X, double Y, double xs, double ys, double s1, double s2, double *sxx, double *syy,
                      double *sxy)
{
    double a = X - xs, b = Y - ys, c = Y + ys;
    double r1 = sqrt(a * a + b * b), r2 = sqrt(a * a + c * c);
    double r14 = r1 * r1 * r1 * r1, r24 = r2 * r2 * r2 * r2, r26 = r24 * r2 * r2;
    double coef = og_rigid / 2.0 / OG_PI / (1.0 - og_pois);
    double a2 = a * a, b2 = b * b, c2 = c * c, c3 = c2 * c, a4 = a2 * a2;
    double sxx1 = s1 * coef * (b * (b2 + 3 * a2) / r14 - c * (c2 + 3 * a2) / r24 +
                               2 * ys / r26 * (6 * Y * c * a2 - b * c3 - a4));
    double sxx2 = s2 * coef * (a * (b2 - a2) / r14 - a * (c2 - a2) / r24 +
                               4 * ys * a / r26 * ((2 * ys - Y) * c2 + (3 * Y + 2 * ys) * a2));
    double syy1 = s1 * coef * (b * (b2 - a2) / r14 - c * (c2 - a2) / r24 -
                               2 * ys / r26 * (6 * Y * c * a2 - (3 * Y + ys) * c3 + a4));
    double syy2 = -s2 * coef * (a * (3 * b2 + a2) / r14 - a * (3 * c2 + a2) / r24 - 4 * ys * a * Y / r26 * (3 * c2 - a2));
    double sxy1 = s1 * coef * (a * (b2 - a2) / r14 - a * (c2 - a2) / r24 + 4 * ys * a * Y / r26 * (3 * c2 - a2));
    double sxy2 = s2 * coef * (b * (b2 - a2) / r14 - c * (c2 - a2) / r24 +
                               2 * ys / r26 * (6 * Y * c * a2 - a4 + (ys - Y) * c3));
    *sxx = sxx1 + sxx2;
    *syy = syy1 + syy2;
    *sxy = sxy1 + sxy2;
}

/* matel2dph2_ij, m_HACApK_calc_entry_ij.f90:92-116 */
static double og_2dph_ij(const og_geom *g, int i, int j)
{
    double s1 = cos(g->ang[j]), s2 = sin(g->ang[j]);
    double p[3], q[3];
    og_d2dip2(g->xcol[i], g->ycol[i], g->xel[j], g->yel[j], s1, s2, &p[0], &p[1], &p[2]);
    og_d2dip2(g->xcol[i], g->ycol[i], g->xer[j], g->yer[j], s1, s2, &q[0], &q[1], &q[2]);
    double sxx = p[0] - q[0], syy = p[1] - q[1], sxy = p[2] - q[2];
    if (g->vcomp == 1) return -0.5 * (sxx + syy) + 0.5 * (sxx - syy) * cos(2 * g->ang[i]) + sxy * sin(2 * g->ang[i]);
    return -0.5 * (sxx - syy) * sin(2 * g->ang[i]) + sxy * cos(2 * g->ang[i]);
}

/* inte11s .. inte22o, m_HACApK_calc_entry_ij.f90:697-750: the in-plane
 * segment integrals at one end (k = 0: 11, 1: 12, 2: 22), shear or opening */
static double og_inte(int k, int open, double x1, double x2)
{
    double r = sqrt(x1 * x1 + x2 * x2), pa = og_vs / 6.0, c = 2 * og_vs / OG_PI * (1 - pa * pa);
    double r4 = r * r * r * r;
    if (!open) {
        if (k == 0) return -c * x2 * (3 * x1 * x1 + x2 * x2) / r4;
        if (k == 1) return c * x1 * (x1 * x1 - x2 * x2) / r4;
        return c * x2 * (x1 * x1 - x2 * x2) / r4;
    }
    if (k == 0) return c * x1 * (x1 * x1 - x2 * x2) / r4;
    if (k == 1) return c * x2 * (x1 * x1 - x2 * x2) / r4;
    return c * x1 * (3 * x1 * x1 + x2 * x2) / r4;
}

/* tensor2d_ij, m_HACApK_calc_entry_ij.f90:300-339 */
static double og_2dn_ij(const og_geom *g, int i, int j)
{
    double cj = cos(g->ang[j]), sj = sin(g->ang[j]);
    double ax = g->xcol[i] - g->xel[j], ay = g->ycol[i] - g->yel[j];
    double bx = g->xcol[i] - g->xer[j], by = g->ycol[i] - g->yer[j];
    double xp = cj * ax + sj * ay, xm = cj * bx + sj * by;
    double yp = -sj * ax + cj * ay, ym = -sj * bx + cj * by;
    int open = g->md == 3;
    double k[3];
    for (int c = 0; c < 3; c++) k[c] = -0.5 * og_rigid / og_vs * (og_inte(c, open, xp, yp) - og_inte(c, open, xm, ym));
    double sin2 = sin(-2 * g->ang[j]), cos2 = cos(-2 * g->ang[j]);
    double sxx = 0.5 * (k[0] + k[2]) + 0.5 * (k[0] - k[2]) * cos2 + k[1] * sin2;
    double sxy = -0.5 * (k[0] - k[2]) * sin2 + k[1] * cos2;
    double syy = 0.5 * (k[0] + k[2]) - 0.5 * (k[0] - k[2]) * cos2 - k[1] * sin2;
    if (g->vcomp == 1)
        return -(0.5 * (sxx + syy) - 0.5 * (sxx - syy) * cos(2 * g->ang[i]) - sxy * sin(2 * g->ang[i]));
    return 0.5 * (sxx - syy) * sin(-2 * g->ang[i]) + sxy * cos(-2 * g->ang[i]);
}

/* matelrec_ij + dc3d25 + ua25, m_HACApK_calc_entry_ij.f90:934-1064 */
static double og_25d_ij(const og_geom *g, int i, int j)
{
    double alpha = (1.0 + (0.5 / og_pois - 1.0)) / (1.0 + 2.0 * (0.5 / og_pois - 1.0));
    double st = g->ang[j];
    double x = cos(st) * (g->xcol[i] - g->xcol[j]) + sin(st) * (g->ycol[i] - g->ycol[j]);
    double q = -sin(st) * (g->xcol[i] - g->xcol[j]) + cos(st) * (g->ycol[i] - g->ycol[j]);
    double xi[2] = { x + 0.5 * g->ds[j], x - 0.5 * g->ds[j] }, et[2] = { 0.5 * g->w, -0.5 * g->w };
    double u4 = 0, u5 = 0, u7 = 0, u8 = 0, u12 = 0;
    for (int k = 0; k < 2; k++)
        for (int m = 0; m < 2; m++) {
            double X = xi[m], E = et[k];
            double r = sqrt(X * X + E * E + q * q);
            double x11 = 1.0 / (r * (r + X)), y11 = 1.0 / (r * (r + E));
            double y32 = (r + (r + E)) * y11 * y11 / r, r3 = r * r * r;
            double d4 = -(1.0 - alpha) / 2.0 * q * y11 - alpha / 2 * X * X * q * y32;
            double d6 = (1.0 - alpha) / 2.0 * X * y11 + alpha / 2 * X * q * q * y32;
            double d7 = (1.0 - alpha) / 2.0 * X * y11 + alpha / 2 * X * (E / r3 + X * X * y32) + E / 2.0 * x11;
            double d9 = (1.0 - alpha) / 2.0 * (q * y11) - alpha / 2 * q * (E / r3 + X * X * y32);
            double d11 = alpha / 2 * (E * q / r3);
            double s = (m + k == 1) ? -1.0 : 1.0, tp = OG_PI * 2;
            u4 += s * (-(d4 / tp));
            u5 += s * (d6 / tp);
            u7 += s * (-(d7 / tp));
            u8 += s * (d9 / tp);
            u12 += s * (d11 / tp);
        }
    double exx = -u4, eyy = -u8, ezz = -u12, exy = -0.5 * (u7 + u5);
    double sxx = og_rigid * (exx + eyy + ezz) + 2 * og_rigid * exx;
    double syy = og_rigid * (exx + eyy + ezz) + 2 * og_rigid * eyy;
    double sxy = 2 * og_rigid * exy;
    double d = 2 * (g->ang[i] - g->ang[j]);
    if (g->vcomp == 1) return -0.5 * (sxx + syy) + 0.5 * (sxx - syy) * cos(d) + sxy * sin(d);
    return -0.5 * (sxx - syy) * sin(d) + sxy * cos(d);
}

/* matels2dpFZ_ij, m_HACApK_calc_entry_ij.f90:663-683: the antiplane segment
   plus ten image pairs at a = k w (the numerator dx^3 - 3 a^2 as written) */
static double og_2dpfz_ij(const og_geom *g, int i, int j)
{
    double d1 = g->xcol[i] - g->xer[j], d2 = g->xcol[i] - g->xel[j];
    double v = 1.0 / d1 - 1.0 / d2;
    for (int k = 1; k <= 10; k++) {
        double a = (double)k * g->w;
        double p1 = a * a + d1 * d1, p2 = a * a + d2 * d2;
        double t1 = 2 * d1 * (d1 * d1 * d1 - 3 * (a * a)) / (p1 * p1 * p1);
        double t2 = 2 * d2 * (d2 * d2 * d2 - 3 * (a * a)) / (p2 * p2 * p2);
        v = v + t1 - t2;
    }
    return v * (og_rigid / (2.0 * OG_PI));
}

/* matel2dn_ij + kern, m_HACApK_calc_entry_ij.f90:472-562: traction vectors of
   the segment's end points, shear (vcomp 0) or normal (vcomp 1) */
static double og_2dnv_ij(const og_geom *g, int i, int j)
{
    double factor = og_rigid / (2.0 * OG_PI * (1.0 - og_pois));
    double nx = -sin(g->ang[i]), ny = cos(g->ang[i]);
    double I1[2], I2[2];
    for (int e = 0; e < 2; e++) {
        double dx = g->xcol[i] - (e ? g->xer[j] : g->xel[j]);
        double dy = g->ycol[i] - (e ? g->yer[j] : g->yel[j]);
        double r = sqrt(dx * dx + dy * dy), gx = dx / r, gy = dy / r;
        double c2 = gy * gy - gx * gx, q = ny * ny - nx * nx;
        if (g->vcomp == 0) {
            double t = 4 * nx * ny * gx * gy + q * c2;
            I1[e] = -(gy / r * t);
            I2[e] = -(gx / r * t);
        } else {
            double t = 2 * nx * ny * c2 - 2 * q * gx * gy;
            I1[e] = gx / r + gy / r * t;
            I2[e] = gy / r + gx / r * t;
        }
    }
    double sx = -sin(g->ang[j]), sy = cos(g->ang[j]);
    return factor * (sx * (-I1[1] + I1[0]) + sy * (I2[1] - I2[0]));
}

/* tensor2d3_ij + inte31s / inte32s, m_HACApK_calc_entry_ij.f90:378-393, 751-765 */
static double og_2dna_ij(const og_geom *g, int i, int j)
{
    double c = cos(g->ang[j]), s = sin(g->ang[j]);
    double ex[2] = { g->xcol[i] - g->xel[j], g->xcol[i] - g->xer[j] };
    double ey[2] = { g->ycol[i] - g->yel[j], g->ycol[i] - g->yer[j] };
    double k31[2], k32[2];
    for (int e = 0; e < 2; e++) {
        double x1 = c * ex[e] + s * ey[e], x2 = -s * ex[e] + c * ey[e];
        double r = sqrt(x1 * x1 + x2 * x2);
        k31[e] = -(og_vs * x2 / OG_PI / (r * r));
        k32[e] = og_vs / OG_PI * x1 / (r * r);
    }
    double f = -0.5 * og_rigid / og_vs;
    return sin(g->ang[i]) * (f * (k31[0] - k31[1])) + cos(g->ang[i]) * (f * (k32[0] - k32[1]));
}

/* HACApK_entry_ij, m_HACApK_calc_entry_ij.f90:24-90 (0-based, mesh order) */
double og_entry(const og_geom *g, int i, int j)
{
    switch (g->kind) {
    case 1:
        return og_antiplane(g->xcol[i], g->xel[j], g->xer[j]) - og_antiplane(g->xcol[i], -g->xel[j], -g->xer[j]);
    case 2: return og_okada_ij(g, i, j, 0);
    case 3: return og_okada_ij(g, i, j, 1);
    case 4:
    case 5: return og_tri_ij(g, i, j);
    case 6: return og_2dp_ij(g, i, j);
    case 7: return og_3dp_ij(g, i, j);
    case 8: return og_2dph_ij(g, i, j);
    case 9: return og_2dn_ij(g, i, j);
    case 10: return og_25d_ij(g, i, j);
    case 11: return og_2dpfz_ij(g, i, j);
    case 12: return og_2dnv_ij(g, i, j);
    case 13: return og_2dna_ij(g, i, j);
    }
    return NAN;
}

void og_entries(const og_geom *g, int np, const int *ii, const int *jj, double *out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int p = 0; p < np; p++) out[p] = og_entry(g, ii[p], jj[p]);
}

/* ------------------------------------------------------------------------- */
/* cluster tree and partition                                                */
/* ------------------------------------------------------------------------- */
typedef struct og_clt {
    int nstrt, nsize, nnson;
    struct og_clt *son[2];
    double bmin[3], bmax[3], zwdth;
} og_clt;

/* HACApK_generate_cbitree, m_HACApK_base.f90:3272-3333 (lod is the 0-based
   slice owned by this cluster, nsrt the 0-based start) */
static og_clt *og_cbitree(const double *coord, int n, int *lod, int nsrt, int nd, int minsz)
{
    og_clt *c = (og_clt *)calloc(1, sizeof(og_clt));
    c->nstrt = nsrt; c->nsize = nd;
    if (nd <= minsz) { c->nnson = 0; return c; }
    double zlmin[3], zlmax[3];
    for (int id = 0; id < 3; id++) {
        zlmin[id] = coord[(size_t)id * n + lod[0]]; zlmax[id] = zlmin[id];
        for (int il = 1; il < nd; il++) {
            double zg = coord[(size_t)id * n + lod[il]];
            if (zg < zlmin[id]) zlmin[id] = zg;
            else if (zlmax[id] < zg) zlmax[id] = zg;
        }
    }
    double zdiff = zlmax[0] - zlmin[0];
    int ncut = 0;
    for (int id = 0; id < 3; id++) {
        double zid = zlmax[id] - zlmin[id];
        if (zid > zdiff) { zdiff = zid; ncut = id; }
    }
    double zlmid = (zlmax[ncut] + zlmin[ncut]) / 2;
    const double *zc = coord + (size_t)ncut * n;
    /* 1-based Hoare-style sweep of the reference, kept with 1-based counters */
    int nl = 1, nr = nd;
    while (nl < nr) {
        while (nl < nd && zc[lod[nl - 1]] <= zlmid) nl++;
        while (nr >= 1 && zc[lod[nr - 1]] > zlmid) nr--;
        if (nl < nr) { int t = lod[nl - 1]; lod[nl - 1] = lod[nr - 1]; lod[nr - 1] = t; }
    }
    c->nnson = 2;
    c->son[0] = og_cbitree(coord, n, lod, nsrt, nl - 1, minsz);
    c->son[1] = og_cbitree(coord, n, lod + (nl - 1), nsrt + nl - 1, nd - nl + 1, minsz);
    return c;
}

/* HACApK_bndbox, m_HACApK_base.f90:3191-3229 */
static void og_bndbox(og_clt *c, const double *coord, int n, const int *lod)
{
    if (c->nnson == 0) {
        for (int id = 0; id < 3; id++) {
            c->bmin[id] = coord[(size_t)id * n + lod[c->nstrt]];
            c->bmax[id] = c->bmin[id];
        }
        for (int id = 0; id < 3; id++)
            for (int il = 1; il < c->nsize; il++) {
                double z = coord[(size_t)id * n + lod[c->nstrt + il]];
                if (z < c->bmin[id]) c->bmin[id] = z;
                if (c->bmax[id] < z) c->bmax[id] = z;
            }
    } else {
        for (int s = 0; s < c->nnson; s++) og_bndbox(c->son[s], coord, n, lod);
        for (int id = 0; id < 3; id++) { c->bmin[id] = c->son[0]->bmin[id]; c->bmax[id] = c->son[0]->bmax[id]; }
        for (int s = 1; s < c->nnson; s++)
            for (int id = 0; id < 3; id++) {
                if (c->son[s]->bmin[id] < c->bmin[id]) c->bmin[id] = c->son[s]->bmin[id];
                if (c->bmax[id] < c->son[s]->bmax[id]) c->bmax[id] = c->son[s]->bmax[id];
            }
    }
    double zw = (c->bmax[0] - c->bmin[0]) * (c->bmax[0] - c->bmin[0]);
    for (int id = 1; id < 3; id++) zw = zw + (c->bmax[id] - c->bmin[id]) * (c->bmax[id] - c->bmin[id]);
    c->zwdth = sqrt(zw);
}

static void og_free_clt(og_clt *c)
{
    if (!c) return;
    for (int s = 0; s < c->nnson; s++) og_free_clt(c->son[s]);
    free(c);
}

typedef struct {
    int ltmtx;  /* 1: low rank, 2: dense */
    int nstrtl, ndl, nstrtt, ndt, kt;  /* 0-based starts in cluster order */
    double *a1, *a2;  /* LR: a1(ndt,kt) col-major, a2(ndl,kt); dense: a1(ndt,ndl) */
} og_leaf;

typedef struct { og_leaf *v; int n, cap; } og_leafvec;

static void og_push(og_leafvec *lv, int ltmtx, int nstrtl, int ndl, int nstrtt, int ndt)
{
    if (lv->n == lv->cap) {
        lv->cap = lv->cap ? 2 * lv->cap : 1024;
        lv->v = (og_leaf *)realloc(lv->v, sizeof(og_leaf) * lv->cap);
    }
    og_leaf *f = &lv->v[lv->n++];
    memset(f, 0, sizeof(*f));
    f->ltmtx = ltmtx; f->nstrtl = nstrtl; f->ndl = ndl; f->nstrtt = nstrtt; f->ndt = ndt;
}

/* HACApK_generate_leafmtx, m_HACApK_base.f90:2937-2995 (strong admissibility,
   param(52)=1; nlmax = n; depth limit param(53)=100) */
static void og_leafmtx(og_leafvec *lv, const og_clt *l, const og_clt *t, int leaf, double eta,
                       int nlmax, int ndpth)
{
    int ndl = l->nsize, ndt = t->nsize, nstrtl = l->nstrt, nstrtt = t->nstrt;
    int nleaf = leaf + 1;
    ndpth = ndpth + 1;
    const int mdpth = 100;
    double zs = 0.0;
    for (int id = 0; id < 3; id++) {
        if (l->bmax[id] < t->bmin[id]) zs = zs + (t->bmin[id] - l->bmax[id]) * (t->bmin[id] - l->bmax[id]);
        else if (t->bmax[id] < l->bmin[id]) zs = zs + (l->bmin[id] - t->bmax[id]) * (l->bmin[id] - t->bmax[id]);
    }
    double zdistlt = sqrt(zs);
    if ((l->zwdth <= eta * zdistlt || t->zwdth <= eta * zdistlt) &&
        (ndl >= nleaf && ndt >= nleaf && ndl <= nlmax && ndt <= nlmax)) {
        if ((nstrtl + ndl) != nstrtt && (nstrtt + ndt) != nstrtl) {
            og_push(lv, 1, nstrtl, ndl, nstrtt, ndt);
            return;
        }
    }
    if (ndpth == mdpth || l->nnson == 0 || t->nnson == 0 || ndl <= nleaf || ndt <= nleaf) {
        og_push(lv, 2, nstrtl, ndl, nstrtt, ndt);
        return;
    }
    for (int il = 0; il < l->nnson; il++)
        for (int it = 0; it < t->nnson; it++) og_leafmtx(lv, l->son[il], t->son[it], leaf, eta, nlmax, ndpth);
}

static int og_leafcmp(const void *a, const void *b)
{
    const og_leaf *x = (const og_leaf *)a, *y = (const og_leaf *)b;
    if (x->nstrtl != y->nstrtl) return x->nstrtl < y->nstrtl ? -1 : 1;
    if (x->nstrtt != y->nstrtt) return x->nstrtt < y->nstrtt ? -1 : 1;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* ACA+ (m_HACApK_base.f90:1638-1946, HACApK_lib.f90:31-116)                 */
/* ------------------------------------------------------------------------- */
static double og_unrm(int nd, const double *za)
{
    double zscl = 0.0, zzz = 1.0;
    for (int il = 0; il < nd; il++) {
        double zza = fabs(za[il]);
        if (zza < 1.0e-30) continue;
        if (zscl < zza) {
            zzz = 1.0 + zzz * (zscl / zza) * (zscl / zza);
            zscl = zza;
        } else {
            zzz = zzz + (zza / zscl) * (zza / zscl);
        }
    }
    return zscl * sqrt(zzz);
}
static int og_maxabs(const double *za, int nd, double *zz)
{
    int il = 0; *zz = fabs(za[0]);
    for (int it = 1; it < nd; it++) if (fabs(za[it]) > *zz) { il = it; *zz = fabs(za[it]); }
    return il;
}
static int og_minabs(const double *za, int nd, double *zz)
{
    int il = 0; *zz = fabs(za[0]);
    for (int it = 1; it < nd; it++) if (fabs(za[it]) < *zz) { il = it; *zz = fabs(za[it]); }
    return il;
}

typedef struct {
    const og_geom *g; const int *lod; int nstrtl, nstrtt, ndl, ndt;
} og_blk;

/* HACApK_calc_vec: ld=1 residual column ip (length ndl), ld=0 residual row ip (length ndt).
   za is the factor whose length matches the vector, zb the other one; both
   column-major with leading dimension = their length. */
static void og_calc_vec(const og_blk *b, const double *za, const double *zb, int ndp, int ndo, int k, int ip,
                        double *vec, const int *lmsk, int ld)
{
    for (int ii = 0; ii < ndp; ii++) {
        if (lmsk[ii] == 0) {
            int ill, itt;
            if (ld == 0) { ill = ip + b->nstrtl; itt = ii + b->nstrtt; }
            else { ill = ii + b->nstrtl; itt = ip + b->nstrtt; }
            vec[ii] = og_entry(b->g, b->lod[ill], b->lod[itt]);
        }
    }
    if (k == 0) return;
    for (int ii = 0; ii < ndp; ii++) {
        double zau = 0.0;
        for (int m = 0; m < k; m++) zau = zau + za[(size_t)m * ndp + ii] * zb[(size_t)m * ndo + ip];
        vec[ii] = vec[ii] - zau;
    }
    for (int ii = 0; ii < ndp; ii++) if (lmsk[ii] == 1) vec[ii] = 0.0;
}

/* HACApK_acaplus: returns kt; zaa(ndl,kmax), zab(ndt,kmax) zero-initialised by the caller */
int og_acaplus(const og_blk *b, double *zaa, double *zab, int kmax, double eps, double znrmmat, double pACA_EPS)
{
    int ndl = b->ndl, ndt = b->ndt;
    const double za_ACA_EPS = (double)1.0e-30f;
    const double tiny_piv = (double)1.0e-20f;
    double znrm = znrmmat * (double)sqrtf((float)ndl * (float)ndt);
    double ACA_EPS = pACA_EPS * znrm;
    int ntries = (ndl > ndt ? ndl : ndt) + 1, ntries_row = 6, ntries_col = 6;
    int *lrow_msk = (int *)calloc(ndl, sizeof(int));
    int *lcol_msk = (int *)calloc(ndt, sizeof(int));
    double *pa_ref = (double *)calloc(ndl, sizeof(double));
    double *pb_ref = (double *)calloc(ndt, sizeof(double));
    int k = 0;
    int j_ref = 0;
    og_calc_vec(b, zaa, zab, ndl, ndt, k, j_ref, pa_ref, lrow_msk, 1);
    double colnorm = og_unrm(ndl, pa_ref);
    double rownorm;
    int i_ref = og_minabs(pa_ref, ndl, &rownorm);
    og_calc_vec(b, zab, zaa, ndt, ndl, k, i_ref, pb_ref, lcol_msk, 0);
    rownorm = og_unrm(ndt, pb_ref);
    double apxnorm = 0.0;
    int lstop = 0;
    int i = 0, j = 0;
    while (k < kmax && (ntries_row > 0 || ntries_col > 0) && ntries > 0) {
        ntries--;
        double *pcol = zaa + (size_t)k * ndl, *prow = zab + (size_t)k * ndt;
        double col_maxval, row_maxval, zinvmax;
        i = og_maxabs(pa_ref, ndl, &col_maxval);
        j = og_maxabs(pb_ref, ndt, &row_maxval);
        if (row_maxval > col_maxval) {
            if (j != j_ref) og_calc_vec(b, zaa, zab, ndl, ndt, k, j, pcol, lrow_msk, 1);
            else memcpy(pcol, pa_ref, sizeof(double) * ndl);
            i = og_maxabs(pcol, ndl, &col_maxval);
            if (col_maxval < ACA_EPS && k >= 1) {
                lstop = 1;
            } else {
                og_calc_vec(b, zab, zaa, ndt, ndl, k, i, prow, lcol_msk, 0);
                if (fabs(pcol[i]) > tiny_piv) zinvmax = 1.0 / pcol[i];
                else { k = k - 1 > 0 ? k - 1 : 0; break; }
                for (int ii = 0; ii < ndl; ii++) pcol[ii] = pcol[ii] * zinvmax;
            }
        } else {
            if (i != i_ref) og_calc_vec(b, zab, zaa, ndt, ndl, k, i, prow, lcol_msk, 0);
            else memcpy(prow, pb_ref, sizeof(double) * ndt);
            j = og_maxabs(prow, ndt, &row_maxval);
            if (row_maxval < ACA_EPS && k >= 1) {
                lstop = 1;
            } else {
                og_calc_vec(b, zaa, zab, ndl, ndt, k, j, pcol, lrow_msk, 1);
                if (fabs(prow[j]) > tiny_piv) zinvmax = 1.0 / prow[j];
                else { k = k - 1 > 0 ? k - 1 : 0; break; }
                for (int jj = 0; jj < ndt; jj++) prow[jj] = prow[jj] * zinvmax;
            }
        }
        lrow_msk[i] = 1; lcol_msk[j] = 1;
        if (i != i_ref) {
            zinvmax = -pcol[i_ref];
            for (int jj = 0; jj < ndt; jj++) pb_ref[jj] = pb_ref[jj] + prow[jj] * zinvmax;
            rownorm = og_unrm(ndt, pb_ref);
        }
        if (i == i_ref || rownorm < ACA_EPS) {
            if (i == i_ref) ntries_row++;
            if (ntries_row > 0) {
                rownorm = 0.0; i = i_ref;
                int stop = (i_ref - 1 + ndl) % ndl;
                while (i != stop && rownorm < za_ACA_EPS && ntries_row > 0) {
                    if (lrow_msk[i] == 0) {
                        og_calc_vec(b, zab, zaa, ndt, ndl, k + 1, i, pb_ref, lcol_msk, 0);
                        rownorm = og_unrm(ndt, pb_ref);
                        if (rownorm < ACA_EPS) lrow_msk[i] = 1;
                        ntries_row--;
                    } else {
                        rownorm = 0.0;
                    }
                    i = (i + 1) % ndl;
                }
                i_ref = (i - 1 + ndl) % ndl;
            }
        }
        if (j != j_ref) {
            zinvmax = -prow[j_ref];
            for (int ii = 0; ii < ndl; ii++) pa_ref[ii] = pa_ref[ii] + pcol[ii] * zinvmax;
            colnorm = og_unrm(ndl, pa_ref);
        }
        if (j == j_ref || colnorm < ACA_EPS) {
            if (j == j_ref) ntries_col++;
            if (ntries_col > 0) {
                colnorm = 0.0; j = j_ref;
                int stop = (j_ref - 1 + ndt) % ndt;
                while (j != stop && colnorm < za_ACA_EPS && ntries_col > 0) {
                    if (lcol_msk[j] == 0) {
                        og_calc_vec(b, zaa, zab, ndl, ndt, k + 1, j, pa_ref, lrow_msk, 1);
                        colnorm = og_unrm(ndl, pa_ref);
                        if (colnorm < ACA_EPS) lcol_msk[j] = 1;
                        ntries_col--;
                    } else {
                        colnorm = 0.0;
                    }
                    j = (j + 1) % ndt;
                }
                j_ref = (j - 1 + ndt) % ndt;
            }
        }
        if (colnorm < ACA_EPS && rownorm < ACA_EPS && k >= 1) { lstop = 1; k = k + 1; }
        if (lstop == 0) {
            double blknorm = og_unrm(ndl, pcol) * og_unrm(ndt, prow);
            if (k == 0) {
                apxnorm = znrm;
            } else if (blknorm < apxnorm * eps && rownorm < apxnorm * eps && colnorm < apxnorm * eps && k >= 1) {
                lstop = 1;
            }
        }
        if (lstop == 1 && k >= 1) break;
        k = k + 1;
    }
    free(lrow_msk); free(lcol_msk); free(pa_ref); free(pb_ref);
    return k;
}

/* ------------------------------------------------------------------------- */
/* H-matrix build (HACApK_generate + construct_LH at np=1)                   */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n, nlf;
    int *lod;  /* cluster position -> mesh element (0-based) */
    og_leaf *lf;
    og_geom g;
} og_hmat;

og_hmat *og_hmat_build(const og_geom *g, int leaf, double eta, double eps_h, int kmax, int nthreads)
{
    int n = g->n;
    og_hmat *h = (og_hmat *)calloc(1, sizeof(og_hmat));
    h->n = n; h->g = *g;
    h->lod = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) h->lod[i] = i;
    double *coord = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        coord[i] = g->xcol[i];
        coord[(size_t)n + i] = g->ycol ? g->ycol[i] : 0.0;
        coord[2 * (size_t)n + i] = g->zcol ? g->zcol[i] : 0.0;
    }
    og_clt *root = og_cbitree(coord, n, h->lod, 0, n, leaf);
    og_bndbox(root, coord, n, h->lod);
    og_leafvec lv = { 0 };
    if (root->nnson == 0 || n <= leaf + 1) {
        og_push(&lv, 2, 0, n, 0, n);
    } else {
        for (int il = 0; il < root->nnson; il++)
            for (int it = 0; it < root->nnson; it++) og_leafmtx(&lv, root->son[il], root->son[it], leaf, eta, n, 1);
    }
    qsort(lv.v, lv.n, sizeof(og_leaf), og_leafcmp);
    og_free_clt(root);
    free(coord);
    h->nlf = lv.n; h->lf = lv.v;
    double znrmmat = (double)(1.0f / (float)n);
    double pACA_EPS = (double)1.0e-6f * eps_h;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
    for (int ip = 0; ip < h->nlf; ip++) {
        og_leaf *f = &h->lf[ip];
        og_blk b = { &h->g, h->lod, f->nstrtl, f->nstrtt, f->ndl, f->ndt };
        if (f->ltmtx == 1) {
            double *zaa = (double *)calloc((size_t)f->ndl * kmax, sizeof(double));
            double *zab = (double *)calloc((size_t)f->ndt * kmax, sizeof(double));
            int kt = og_acaplus(&b, zaa, zab, kmax, eps_h, znrmmat, pACA_EPS);
            f->kt = kt;
            f->a1 = (double *)malloc(sizeof(double) * ((size_t)f->ndt * kt + 1));
            f->a2 = (double *)malloc(sizeof(double) * ((size_t)f->ndl * kt + 1));
            memcpy(f->a1, zab, sizeof(double) * (size_t)f->ndt * kt);
            memcpy(f->a2, zaa, sizeof(double) * (size_t)f->ndl * kt);
            free(zaa); free(zab);
        } else {
            f->a1 = (double *)malloc(sizeof(double) * (size_t)f->ndt * f->ndl);
            for (int il = 0; il < f->ndl; il++)
                for (int it = 0; it < f->ndt; it++)
                    f->a1[(size_t)il * f->ndt + it] = og_entry(&h->g, h->lod[il + f->nstrtl], h->lod[it + f->nstrtt]);
        }
    }
    return h;
}

/* Build the partition only (no entries): for partition tests on large meshes. */
og_hmat *og_hmat_partition(const og_geom *g, int leaf, double eta)
{
    int n = g->n;
    og_hmat *h = (og_hmat *)calloc(1, sizeof(og_hmat));
    h->n = n; h->g = *g;
    h->lod = (int *)malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) h->lod[i] = i;
    double *coord = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        coord[i] = g->xcol[i];
        coord[(size_t)n + i] = g->ycol ? g->ycol[i] : 0.0;
        coord[2 * (size_t)n + i] = g->zcol ? g->zcol[i] : 0.0;
    }
    og_clt *root = og_cbitree(coord, n, h->lod, 0, n, leaf);
    og_bndbox(root, coord, n, h->lod);
    og_leafvec lv = { 0 };
    if (root->nnson == 0 || n <= leaf + 1) og_push(&lv, 2, 0, n, 0, n);
    else
        for (int il = 0; il < root->nnson; il++)
            for (int it = 0; it < root->nnson; it++) og_leafmtx(&lv, root->son[il], root->son[it], leaf, eta, n, 1);
    qsort(lv.v, lv.n, sizeof(og_leaf), og_leafcmp);
    og_free_clt(root);
    free(coord);
    h->nlf = lv.n; h->lf = lv.v;
    return h;
}

int og_hmat_nleaves(const og_hmat *h) { return h->nlf; }
void og_hmat_lod(const og_hmat *h, int *out) { memcpy(out, h->lod, sizeof(int) * h->n); }
/* per leaf: ltmtx, nstrtl, ndl, nstrtt, ndt, kt */
void og_hmat_leaves(const og_hmat *h, int *out)
{
    for (int i = 0; i < h->nlf; i++) {
        const og_leaf *f = &h->lf[i];
        out[6 * i + 0] = f->ltmtx; out[6 * i + 1] = f->nstrtl; out[6 * i + 2] = f->ndl;
        out[6 * i + 3] = f->nstrtt; out[6 * i + 4] = f->ndt; out[6 * i + 5] = f->kt;
    }
}
/* copy factors of leaf ip: LR -> a1 (ndt*kt) then a2 (ndl*kt); dense -> a1 (ndt*ndl) */
void og_hmat_leaf_data(const og_hmat *h, int ip, double *out)
{
    const og_leaf *f = &h->lf[ip];
    if (f->ltmtx == 1) {
        memcpy(out, f->a1, sizeof(double) * (size_t)f->ndt * f->kt);
        memcpy(out + (size_t)f->ndt * f->kt, f->a2, sizeof(double) * (size_t)f->ndl * f->kt);
    } else {
        memcpy(out, f->a1, sizeof(double) * (size_t)f->ndt * f->ndl);
    }
}

void og_hmat_free(og_hmat *h)
{
    if (!h) return;
    for (int i = 0; i < h->nlf; i++) { free(h->lf[i].a1); free(h->lf[i].a2); }
    free(h->lf); free(h->lod); free(h);
}

/* ------------------------------------------------------------------------- */
/* H-matrix-vector product: HACApK_adot_body_lfmtx_hyp_contiguous at np=1     */
/* (cluster-order x and y). Threads take contiguous leaf ranges balanced by   */
/* stored entries (setcutthread_lattice, m_HACApK_use.f90:387-431) and merge  */
/* private partial vectors in thread order, so a fixed thread count gives a   */
/* reproducible result; nthreads=1 is the leaf-order serial product.          */
/* ------------------------------------------------------------------------- */
static void og_leaf_apply(const og_leaf *f, const double *zu, double *zaut, double *zbut)
{
    if (f->ltmtx == 1) {
        for (int il = 0; il < f->kt; il++) {
            double s = 0.0;
            const double *a1 = f->a1 + (size_t)il * f->ndt;
            for (int it = 0; it < f->ndt; it++) s = s + a1[it] * zu[it + f->nstrtt];
            zbut[il] = s;
        }
        for (int il = 0; il < f->kt; il++) {
            const double *a2 = f->a2 + (size_t)il * f->ndl;
            for (int it = 0; it < f->ndl; it++) zaut[it + f->nstrtl] = zaut[it + f->nstrtl] + a2[it] * zbut[il];
        }
    } else {
        for (int il = 0; il < f->ndl; il++) {
            const double *row = f->a1 + (size_t)il * f->ndt;
            double acc = zaut[il + f->nstrtl];
            for (int it = 0; it < f->ndt; it++) acc = acc + row[it] * zu[it + f->nstrtt];
            zaut[il + f->nstrtl] = acc;
        }
    }
}

void og_hmat_matvec(const og_hmat *h, const double *x, double *y, int nthreads)
{
    int n = h->n;
    int ktmax = 1;
    for (int i = 0; i < h->nlf; i++) if (h->lf[i].kt > ktmax) ktmax = h->lf[i].kt;
    if (nthreads <= 1) {
        double *zbut = (double *)malloc(sizeof(double) * ktmax);
        for (int i = 0; i < n; i++) y[i] = 0.0;
        double *zaut = (double *)calloc(n, sizeof(double));
        for (int ip = 0; ip < h->nlf; ip++) og_leaf_apply(&h->lf[ip], x, zaut, zbut);
        for (int i = 0; i < n; i++) y[i] = y[i] + zaut[i];
        free(zaut); free(zbut);
        return;
    }
    /* memory-balanced contiguous ranges */
    long long mem = 0;
    for (int ip = 0; ip < h->nlf; ip++) {
        const og_leaf *f = &h->lf[ip];
        mem += f->ltmtx == 1 ? (long long)(f->ndl + f->ndt) * f->kt : (long long)f->ndl * f->ndt;
    }
    int *lthr = (int *)malloc(sizeof(int) * (nthreads + 1));
    lthr[0] = 0;
    for (int t = 1; t <= nthreads; t++) lthr[t] = h->nlf;
    long long per = mem / nthreads, acc = 0;
    int ith = 1;
    for (int ip = 0; ip < h->nlf && ith < nthreads; ip++) {
        const og_leaf *f = &h->lf[ip];
        acc += f->ltmtx == 1 ? (long long)(f->ndl + f->ndt) * f->kt : (long long)f->ndl * f->ndt;
        if (acc > per * ith) { lthr[ith] = ip + 1; ith++; }
    }
    double *part = (double *)calloc((size_t)n * nthreads, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        double *zbut = (double *)malloc(sizeof(double) * ktmax);
        double *zaut = part + (size_t)t * n;
        for (int ip = lthr[t]; ip < lthr[t + 1]; ip++) og_leaf_apply(&h->lf[ip], x, zaut, zbut);
        free(zbut);
    }
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int t = 0; t < nthreads; t++) s = s + part[(size_t)t * n + i];
        y[i] = s;
    }
    free(part); free(lthr);
}

/* dense matrix in cluster order (rows i, cols j): A[i*n+j] = K(lod[i], lod[j]) */
void og_dense_cluster(const og_hmat *h, double *A)
{
    int n = h->n;
#pragma omp parallel for schedule(dynamic, 8)
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) A[(size_t)i * n + j] = og_entry(&h->g, h->lod[i], h->lod[j]);
}

/* ------------------------------------------------------------------------- */
/* friction + RK45 (main_LH.f90:2053-2554)                                   */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n;
    int law;          /* 0 aging, 1 slip */
    int backslip, limitsigma;
    double vref, minsig, maxsig, dtmax, hdiv;   /* hdiv: ds0 for okada problems (matvec divided by it), else 0 */
    const double *a, *b, *dc, *f0, *vpl, *taudot, *sigdot;   /* cluster order */
    const og_hmat *hs;   /* shear matrix */
    const og_hmat *hn;   /* normal matrix or NULL (sigmaconst) */
    const og_hmat *hs2, *hn2;   /* opening -> shear / normal matrices, or NULL (no opening) */
    int nthreads;
    double *vel, *xv, *hx, *hy;   /* workspaces */
    double *xn, *hx2, *hy2;       /* opening rate and its products */
    long long nmatvec;
} og_sim;

og_sim *og_sim_create(const og_hmat *hs, const og_hmat *hn, int law, int backslip, int limitsigma, double vref,
                      double minsig, double maxsig, double dtmax, double hdiv,
                      const double *a, const double *b, const double *dc, const double *f0,
                      const double *vpl, const double *taudot, const double *sigdot, int nthreads)
{
    og_sim *s = (og_sim *)calloc(1, sizeof(og_sim));
    int n = hs->n;
    s->n = n; s->law = law; s->backslip = backslip; s->limitsigma = limitsigma;
    s->vref = vref; s->minsig = minsig; s->maxsig = maxsig; s->dtmax = dtmax; s->hdiv = hdiv;
    s->a = a; s->b = b; s->dc = dc; s->f0 = f0; s->vpl = vpl; s->taudot = taudot; s->sigdot = sigdot;
    s->hs = hs; s->hn = hn; s->nthreads = nthreads;
    s->vel = (double *)malloc(sizeof(double) * n);
    s->xv = (double *)malloc(sizeof(double) * n);
    s->hx = (double *)malloc(sizeof(double) * n);
    s->hy = (double *)malloc(sizeof(double) * n);
    return s;
}
void og_sim_free(og_sim *s)
{
    if (!s) return;
    free(s->vel); free(s->xv); free(s->hx); free(s->hy);
    free(s->xn); free(s->hx2); free(s->hy2); free(s);
}
/* opening (main_LH.f90:705-716): the matrices of a tensile dislocation ('open')
   onto shear and normal traction; needs the normal matrix */
int og_sim_set_opening(og_sim *s, const og_hmat *hs2, const og_hmat *hn2)
{
    if (!s->hn || !hs2 || !hn2 || hs2->n != s->n || hn2->n != s->n) return -1;
    s->hs2 = hs2; s->hn2 = hn2;
    s->xn = (double *)malloc(sizeof(double) * s->n);
    s->hx2 = (double *)malloc(sizeof(double) * s->n);
    s->hy2 = (double *)malloc(sizeof(double) * s->n);
    return 0;
}
long long og_sim_nmatvec(const og_sim *s) { return s->nmatvec; }

/* deriv, main_LH.f90:2165-2244 */
void og_deriv(int law, int limitsigma, double vref, double minsig, double maxsig,
              double sum_gs, double sum_gn, double psi, double tau, double sig, double vel,
              double a, double b, double dc, double f0, double *dpsi, double *dtau, double *dsig)
{
    const double rigid = og_rigid, vs = og_vs;
    double dsigdt = sum_gn;
    if (limitsigma) {
        if (sig < minsig) dsigdt = 0;
        if (sig > maxsig) dsigdt = 0;
    }
    double dpsidt;
    if (law == 0) {
        if (b == 0) dpsidt = 0.0;
        else dpsidt = b / dc * vref * exp((f0 - psi) / b) - b * fabs(vel) / dc;
    } else {
        double fss = f0 + (a - b) * log(fabs(vel) / vref);
        dpsidt = -fabs(vel) / dc * (fabs(tau) / sig - fss);
    }
    double dvdtau = 2 * vref * exp(-psi / a) * cosh(tau / sig / a) / (a * sig);
    double dvdsig = -2 * vref * exp(-psi / a) * cosh(tau / sig / a) * tau / (a * (sig * sig));
    double dvdpsi = -vel / a;
    double dtaudt = sum_gs - 0.5 * rigid / vs * (dvdpsi * dpsidt + dvdsig * dsigdt);
    dtaudt = dtaudt / (1.0 + 0.5 * rigid / vs * dvdtau);
    *dpsi = dpsidt; *dtau = dtaudt; *dsig = dsigdt;
}

/* derivs, main_LH.f90:2053-2163 (sigmaconst handled by hn == NULL) */
void og_derivs(og_sim *s, const double *y, double *dydx)
{
    int n = s->n;
#pragma omp parallel for num_threads(s->nthreads > 0 ? s->nthreads : 1)
    for (int i = 0; i < n; i++) {
        double psi = y[3 * i], tau = y[3 * i + 1], sig = y[3 * i + 2];
        s->vel[i] = 2 * s->vref * exp(-psi / s->a[i]) * sinh(tau / sig / s->a[i]);
        s->xv[i] = s->backslip ? s->vel[i] - s->vpl[i] : s->vel[i];
        /* opening rate, main_LH.f90:2096-2100 */
        if (s->hs2) s->xn[i] = fmax(0.0, s->minsig - sig) * fabs(s->vel[i]);
    }
    og_hmat_matvec(s->hs, s->xv, s->hx, s->nthreads);
    s->nmatvec++;
    if (s->hn) { og_hmat_matvec(s->hn, s->xv, s->hy, s->nthreads); s->nmatvec++; }
    if (s->hs2) {
        og_hmat_matvec(s->hs2, s->xn, s->hx2, s->nthreads);
        og_hmat_matvec(s->hn2, s->xn, s->hy2, s->nthreads);
        s->nmatvec += 2;
    }
#pragma omp parallel for num_threads(s->nthreads > 0 ? s->nthreads : 1)
    for (int i = 0; i < n; i++) {
        double gs = s->hdiv != 0.0 ? s->hx[i] / s->hdiv : s->hx[i];
        double gn = 0.0;
        if (s->hn) gn = s->hdiv != 0.0 ? s->hy[i] / s->hdiv : s->hy[i];
        if (s->hs2) {   /* main_LH.f90:2129-2143 */
            gs = gs + (s->hdiv != 0.0 ? s->hx2[i] / s->hdiv : s->hx2[i]);
            gn = gn + (s->hdiv != 0.0 ? s->hy2[i] / s->hdiv : s->hy2[i]);
        }
        gn = gn + s->sigdot[i];
        gs = gs + s->taudot[i];
        og_deriv(s->law, s->limitsigma, s->vref, s->minsig, s->maxsig, gs, gn, y[3 * i], y[3 * i + 1], y[3 * i + 2],
                 s->vel[i], s->a[i], s->b[i], s->dc[i], s->f0[i], &dydx[3 * i], &dydx[3 * i + 1], &dydx[3 * i + 2]);
    }
}

/* Cash-Karp tableau of main_LH.f90:2490-2496: the b_ij, c_i and dc_i there are
   default-real (single precision) literals assigned to real(8) parameters,
   so they carry single-precision rounding; A2..A6 and B21 are double. */
typedef struct { double a2, a3, a4, a5, a6, b21, b31, b32, b41, b42, b43, b51, b52, b53, b54,
                 b61, b62, b63, b64, b65, c1, c3, c4, c6, dc1, dc3, dc4, dc5, dc6; } og_tab;
static og_tab og_tableau(void)
{
    og_tab t;
    t.a2 = .2; t.a3 = .3; t.a4 = .6; t.a5 = 1.0; t.a6 = .875; t.b21 = .2;
    t.b31 = (double)(3.f / 40.f); t.b32 = (double)(9.f / 40.f); t.b41 = (double).3f; t.b42 = (double)-.9f;
    t.b43 = (double)1.2f; t.b51 = (double)(-11.f / 54.f); t.b52 = (double)2.5f; t.b53 = (double)(-70.f / 27.f);
    t.b54 = (double)(35.f / 27.f); t.b61 = (double)(1631.f / 55296.f); t.b62 = (double)(175.f / 512.f);
    t.b63 = (double)(575.f / 13824.f); t.b64 = (double)(44275.f / 110592.f); t.b65 = (double)(253.f / 4096.f);
    t.c1 = (double)(37.f / 378.f); t.c3 = (double)(250.f / 621.f); t.c4 = (double)(125.f / 594.f); t.c6 = (double)(512.f / 1771.f);
    t.dc1 = t.c1 - (double)(2825.f / 27648.f); t.dc3 = t.c3 - (double)(18575.f / 48384.f);
    t.dc4 = t.c4 - (double)(13525.f / 55296.f); t.dc5 = (double)(-277.f / 14336.f); t.dc6 = t.c6 - (double).25f;
    return t;
}
void og_tableau_out(double *o)
{
    og_tab t = og_tableau();
    memcpy(o, &t, sizeof(t));
}

/* rkck, main_LH.f90:2472-2554 */
void og_rkck(og_sim *s, const double *y, double h, double *yout, double *yerr)
{
    int m = 3 * s->n;
    og_tab T = og_tableau();
    double *ak1 = (double *)malloc(sizeof(double) * m * 7);
    double *ak2 = ak1 + m, *ak3 = ak2 + m, *ak4 = ak3 + m, *ak5 = ak4 + m, *ak6 = ak5 + m, *yt = ak6 + m;
    og_derivs(s, y, ak1);
    for (int i = 0; i < m; i++) yt[i] = y[i] + T.b21 * h * ak1[i];
    og_derivs(s, yt, ak2);
    for (int i = 0; i < m; i++) yt[i] = y[i] + h * (T.b31 * ak1[i] + T.b32 * ak2[i]);
    og_derivs(s, yt, ak3);
    for (int i = 0; i < m; i++) yt[i] = y[i] + h * (T.b41 * ak1[i] + T.b42 * ak2[i] + T.b43 * ak3[i]);
    og_derivs(s, yt, ak4);
    for (int i = 0; i < m; i++) yt[i] = y[i] + h * (T.b51 * ak1[i] + T.b52 * ak2[i] + T.b53 * ak3[i] + T.b54 * ak4[i]);
    og_derivs(s, yt, ak5);
    for (int i = 0; i < m; i++)
        yt[i] = y[i] + h * (T.b61 * ak1[i] + T.b62 * ak2[i] + T.b63 * ak3[i] + T.b64 * ak4[i] + T.b65 * ak5[i]);
    og_derivs(s, yt, ak6);
    for (int i = 0; i < m; i++) yout[i] = y[i] + h * (T.c1 * ak1[i] + T.c3 * ak3[i] + T.c4 * ak4[i] + T.c6 * ak6[i]);
    for (int i = 0; i < m; i++)
        yerr[i] = h * (T.dc1 * ak1[i] + T.dc3 * ak3[i] + T.dc4 * ak4[i] + T.dc5 * ak5[i] + T.dc6 * ak6[i]);
    free(ak1);
}

/* errmax of rkqs, main_LH.f90:2309-2313, including the overwrite-with-10 guard */
double og_errmax(int m, const double *ytemp, const double *yerr, double eps)
{
    double errmax = 0.0;
    for (int i = 0; i < m; i++) {
        double e = fabs(yerr[i] / ytemp[i]) / eps;
        if (e > errmax) errmax = e;
        if (ytemp[i] - 1 == ytemp[i]) errmax = 10.0;
    }
    return errmax;
}

/* Step-size decision of rkqs, main_LH.f90:2335-2357, for one attempt of
   step h at time x with error ratio errmax:
     accept iff tiny < errmax < 1 -> *hnew = min(2h, SAFETY h errmax^PGROW, dtmax)
                                     (the next step's trial size), returns 1;
     reject, errmax > tiny        -> *hnew = max(h/2, SAFETY h errmax^PSHRNK), returns 0;
     reject, errmax <= tiny       -> *hnew = h/2, returns 0;
     reject and x + hnew - x < 1e-15 (step underflow; the reference stops)
                                  -> returns -1.
   SAFETY, PGROW, PSHRNK and tiny are single-precision literals stored in
   real(8) (main_LH.f90:2283). */
int og_rkqs_control(double h, double errmax, double x, double dtmax, double *hnew)
{
    const double SAFETY = (double)0.9f, PGROW = (double)-0.2f, PSHRNK = (double)-0.25f, tiny = (double)1e-20f;
    if (errmax < 1.0 && errmax > tiny) {
        double hn = 2 * h, hg = SAFETY * h * pow(errmax, PGROW);
        hn = hn < hg ? hn : hg;
        hn = hn < dtmax ? hn : dtmax;
        *hnew = hn;
        return 1;
    }
    if (errmax > tiny) {
        double h1 = 0.5 * h, h2 = SAFETY * h * pow(errmax, PSHRNK);
        h = h1 > h2 ? h1 : h2;
    } else {
        h = 0.5 * h;
    }
    *hnew = h;
    double xnew = x + h;
    if (xnew - x < 1.e-15) return -1;
    return 0;
}

/* rkqs, main_LH.f90:2270-2370.  Returns 0, or -1 when the step underflows
   (the reference stops there). */
int og_rkqs(og_sim *s, double *y, double *x, double htry, double eps, double *hdid, double *hnext,
            double *errmax_out, int *nrjct)
{
    int m = 3 * s->n;
    double *yerr = (double *)malloc(sizeof(double) * m * 2), *ytemp = yerr + m;
    double h = htry, errmax, hnew;
    *nrjct = 0;
    for (;;) {
        og_rkck(s, y, h, ytemp, yerr);
        errmax = og_errmax(m, ytemp, yerr, eps);
        int st = og_rkqs_control(h, errmax, *x, s->dtmax, &hnew);
        if (st == 1) break;
        (*nrjct)++;
        h = hnew;
        if (st < 0) { free(yerr); *errmax_out = errmax; return -1; }
    }
    *hnext = hnew;
    *hdid = h;
    *x = *x + h;
    memcpy(y, ytemp, sizeof(double) * m);
    *errmax_out = errmax;
    free(yerr);
    return 0;
}

/* Initial psi from (tau, sigma, V): main_LH.f90:1004 */
void og_init_state(int n, const double *a, const double *tau, const double *sigma, const double *vel, double vref,
                   double *y)
{
    for (int i = 0; i < n; i++) {
        double psi = a[i] * log(2 * vref / vel[i] * sinh(tau[i] / sigma[i] / a[i]));
        y[3 * i] = psi; y[3 * i + 1] = tau[i]; y[3 * i + 2] = sigma[i];
    }
}

/* Post-step update of main_LH.f90:1266-1287 after an accepted step dtdid
   (viscous flow and opening off): slip advances by the trapezoid rule over
   the old and the new slip rate, half a step each (1271 with the old V,
   1284 with the new one); V is recomputed from the new state
   (V = 2 vref e^(-psi/a) sinh(tau/(sigma a))).  Returns max |V| (1290);
   *meanslip = sum |slip| ds / sum ds, summed in element order (1316-1324). */
double og_post_step(int n, const double *y, const double *a, double vref, double dtdid, const double *ds,
                    double *slip, double *vel, double *meanslip)
{
    double mvel = 0.0, ms = 0.0, sg = 0.0;
    for (int i = 0; i < n; i++) {
        double psi = y[3 * i], tau = y[3 * i + 1], sig = y[3 * i + 2];
        slip[i] = slip[i] + vel[i] * dtdid * 0.5;
        vel[i] = 2 * vref * exp(-psi / a[i]) * sinh(tau / sig / a[i]);
        slip[i] = slip[i] + vel[i] * dtdid * 0.5;
        if (fabs(vel[i]) > mvel) mvel = fabs(vel[i]);
        ms = ms + fabs(slip[i]) * ds[i];
        sg = sg + ds[i];
    }
    *meanslip = ms / sg;
    return mvel;
}

/* The time loop of main_LH.f90:1219-1336 for nsteps accepted steps.
   slip, vel: in/out (cluster order).  ds: element areas (cluster order).
   hist (optional, nsteps x 6): t, dtdid, errmax, nrjct, log10(max V), mean slip.
   vhist, shist (optional, nsteps x nrec): V and slip at the recorded cluster
   indices rec[].
   Returns the number of steps done (< nsteps when the step underflowed or V
   left [velmin, velmax]). */
int og_run(og_sim *s, double *y, double *x, double *dtnxt, double eps_r, int nsteps,
           double *slip, double *vel, const double *ds, double velmin, double velmax,
           double *hist, int nrec, const int *rec, double *vhist, double *shist)
{
    int n = s->n;
    for (int k = 0; k < nsteps; k++) {
        double dtdid, hnext, errmax;
        int nrjct;
        if (og_rkqs(s, y, x, *dtnxt, eps_r, &dtdid, &hnext, &errmax, &nrjct) != 0) return k;
        *dtnxt = hnext;
        double meanslip;
        double mvel = og_post_step(n, y, s->a, s->vref, dtdid, ds, slip, vel, &meanslip);
        if (hist) {
            double *hrow = hist + (size_t)6 * k;
            hrow[0] = *x; hrow[1] = dtdid; hrow[2] = errmax; hrow[3] = nrjct; hrow[4] = log10(mvel); hrow[5] = meanslip;
        }
        if (vhist) for (int r = 0; r < nrec; r++) vhist[(size_t)k * nrec + r] = vel[rec[r]];
        if (shist) for (int r = 0; r < nrec; r++) shist[(size_t)k * nrec + r] = slip[rec[r]];
        if (mvel > velmax || mvel < velmin) return k + 1;
    }
    return nsteps;
}

/* The same loop on a given step grid: accepted step k has size hs[k] (one
   rkck per step, taken whatever its errmax; the error estimate is recorded),
   followed by the post-step update.  A replay of another run's step grid, so
   that two runs can be compared at identical times; the rkqs decision is
   not involved.  hist, vhist, shist as og_run. */
void og_run_grid(og_sim *s, double *y, double *x, const double *hs, int nsteps, double eps_r, double *slip,
                 double *vel, const double *ds, double *hist, int nrec, const int *rec, double *vhist, double *shist)
{
    int n = s->n, m = 3 * n;
    double *yerr = (double *)malloc(sizeof(double) * m * 2), *ytemp = yerr + m;
    for (int k = 0; k < nsteps; k++) {
        double h = hs[k];
        og_rkck(s, y, h, ytemp, yerr);
        double errmax = og_errmax(m, ytemp, yerr, eps_r);
        memcpy(y, ytemp, sizeof(double) * m);
        *x = *x + h;
        double meanslip;
        double mvel = og_post_step(n, y, s->a, s->vref, h, ds, slip, vel, &meanslip);
        if (hist) {
            double *hrow = hist + (size_t)6 * k;
            hrow[0] = *x; hrow[1] = h; hrow[2] = errmax; hrow[3] = 0; hrow[4] = log10(mvel); hrow[5] = meanslip;
        }
        if (vhist) for (int r = 0; r < nrec; r++) vhist[(size_t)k * nrec + r] = vel[rec[r]];
        if (shist) for (int r = 0; r < nrec; r++) shist[(size_t)k * nrec + r] = slip[rec[r]];
    }
    free(yerr);
}
