// entries.hpp -- FP64 matrix-entry kernels evaluated on the GPU.
//
// K(i,j): traction at receiver i from unit slip on source element j, as in
// HACApK_entry_ij (m_HACApK_calc_entry_ij.f90:24-90):
//   2dvs   antiplane with free-surface image   m_HACApK_calc_entry_ij.f90:55-57, 685-695
//   3dph   Okada (1992) rectangle, half-space  m_HACApK_calc_entry_ij.f90:1065-1145, okada.f90
//   3dnr   the same without the image part
//   3dnt   Nikkhoo & Walter (2015) triangle     m_HACApK_calc_entry_ij.f90:802-851, TDstressFS.f90
//   3dht   the same in a half-space             m_HACApK_calc_entry_ij.f90:852-898, TDstressFS.f90:7-44,
//          (main + image + free-surface correction; the correction's strains
//          are forward-mode derivatives of its displacement field, see fsc_strain)
//   2dp    in-plane segment, full space         m_HACApK_calc_entry_ij.f90:648-661
//   3dp    planar quadrilateral, full space     m_HACApK_calc_entry_ij.f90:412-421, 440-470
//   2dph   in-plane dipping segment, half-space m_HACApK_calc_entry_ij.f90:92-116, 143-166 (also '2dnh')
//   2dn    in-plane nonplanar, full space       m_HACApK_calc_entry_ij.f90:300-339, 697-750
//   25d    2.5-D rectangles                     m_HACApK_calc_entry_ij.f90:934-1064
//   2dpfz  antiplane segment + 10 image pairs   m_HACApK_calc_entry_ij.f90:663-683
//   2dn_vector in-plane nonplanar (end tractions) m_HACApK_calc_entry_ij.f90:472-562
//   2dna   antiplane nonplanar, full space      m_HACApK_calc_entry_ij.f90:378-393, 751-765
// The closed-form expressions are those of the published solutions written
// term by term in the reference's evaluation order; FP contraction is off in
// every function here so that the device result tracks the FP64 CPU result
// (remaining differences come from the device libm).  The reference has no
// singular-case guards (okada.f90:73-92 are commented out); neither do we,
// so IEEE +-Inf / NaN propagate exactly as there.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace hbi {

constexpr double kRigid = 32.04;   // m_const.f90:6
constexpr double kVs = 3.464;      // m_const.f90:7
constexpr double kPois = 0.25;     // m_const.f90:3

struct DevGeom {
    int kind, n, vcomp, md;
    double w;
    const double *xcol, *ycol, *zcol, *xel, *xer, *ang, *angd, *dsl, *dsd, *rake;
    const double *xs1, *xs2, *xs3, *ys1, *ys2, *ys3, *zs1, *zs2, *zs3;
    const double *ev11, *ev12, *ev13, *ev21, *ev22, *ev23, *ev31, *ev32, *ev33;
    const double *yel, *yer, *xs4, *zs4, *ds;
};

__device__ __forceinline__ double dpi() { return 4.0 * atan(1.0); }

// ---------------------------------------------------------------- Okada ----
struct OkGeo {
    double xi2, et2, q2, r, r2, r3, r5, yp, d, tt, alx, ale, x11, y11, x32, y32, ey, ez, fy, fz, gy, gz;
};

__device__ __noinline__ void ok_geo(OkGeo &c, double xi, double et, double q, double sd, double cd)
{
#pragma clang fp contract(off)
    c.xi2 = xi * xi;
    c.et2 = et * et;
    c.q2 = q * q;
    c.r2 = c.xi2 + c.et2 + c.q2;
    c.r = sqrt(c.r2);
    c.r3 = c.r * c.r2;
    c.r5 = c.r3 * c.r2;
    c.yp = et * cd + q * sd;
    c.d = et * sd - q * cd;
    c.tt = atan(xi * et / (q * c.r));
    const double rxi = c.r + xi;
    c.alx = log(rxi);
    c.x11 = 1.0 / (c.r * rxi);
    c.x32 = (c.r + rxi) * c.x11 * c.x11 / c.r;
    const double ret = c.r + et;
    c.ale = log(ret);
    c.y11 = 1.0 / (c.r * ret);
    c.y32 = (c.r + ret) * c.y11 * c.y11 / c.r;
    c.ey = sd / c.r - c.yp * q / c.r3;
    c.ez = cd / c.r + c.d * q / c.r3;
    c.fy = c.d / c.r3 + c.xi2 * c.y32 * sd;
    c.fz = c.yp / c.r3 + c.xi2 * c.y32 * cd;
    c.gy = 2.0 * c.x11 * sd - c.yp * q * c.x32;
    c.gz = 2.0 * c.x11 * cd + c.d * q * c.x32;
}

// strike (s1) and dip (s2) terms are combined as out = s1/2pi * du_strike + s2/2pi * du_dip
__device__ __noinline__ void ok_part_a(const OkGeo &c, double xi, double et, double q, double sd, double cd,
                                       double s1, double s2, double alpha, double *o)
{
#pragma clang fp contract(off)
    const double pi2 = 8.0 * atan(1.0);
    const double alp1 = (1.0 - alpha) / 2.0, alp2 = alpha / 2.0;
    const double r = c.r, r3 = c.r3, q2 = c.q2, xi2 = c.xi2, d = c.d, yp = c.yp;
    const double xy = xi * c.y11, qx = q * c.x11, qy = q * c.y11;
    const double f1 = s1 / pi2, f2 = s2 / pi2;
    o[0] = (0.0 + f1 * (c.tt / 2.0 + alp2 * xi * qy)) + f2 * (alp2 * q / r);
    o[1] = (0.0 + f1 * (alp2 * q / r)) + f2 * (c.tt / 2.0 + alp2 * et * qx);
    o[2] = (0.0 + f1 * (alp1 * c.ale - alp2 * q * qy)) + f2 * (alp1 * c.alx - alp2 * q * qx);
    o[3] = (0.0 + f1 * (-alp1 * qy - alp2 * xi2 * q * c.y32)) + f2 * (-alp2 * xi * q / r3);
    o[4] = (0.0 + f1 * (-alp2 * xi * q / r3)) + f2 * (-qy / 2.0 - alp2 * et * q / r3);
    o[5] = (0.0 + f1 * (alp1 * xy + alp2 * xi * q2 * c.y32)) + f2 * (alp1 / r + alp2 * q2 / r3);
    o[6] = (0.0 + f1 * (alp1 * xy * sd + alp2 * xi * c.fy + d / 2.0 * c.x11)) + f2 * (alp2 * c.ey);
    o[7] = (0.0 + f1 * (alp2 * c.ey)) + f2 * (alp1 * d * c.x11 + xy / 2.0 * sd + alp2 * et * c.gy);
    o[8] = (0.0 + f1 * (alp1 * (cd / r + qy * sd) - alp2 * q * c.fy)) + f2 * (alp1 * yp * c.x11 - alp2 * q * c.gy);
    o[9] = (0.0 + f1 * (alp1 * xy * cd + alp2 * xi * c.fz + yp / 2.0 * c.x11)) + f2 * (alp2 * c.ez);
    o[10] = (0.0 + f1 * (alp2 * c.ez)) + f2 * (alp1 * yp * c.x11 + xy / 2.0 * cd + alp2 * et * c.gz);
    o[11] = (0.0 + f1 * (-alp1 * (sd / r - qy * cd) - alp2 * q * c.fz)) + f2 * (-alp1 * d * c.x11 - alp2 * q * c.gz);
}

__device__ __noinline__ void ok_part_b(const OkGeo &c, double xi, double et, double q, double sd, double cd,
                                       double s1, double s2, double alpha, double *o)
{
#pragma clang fp contract(off)
    const double pi2 = 8.0 * atan(1.0);
    const double alp3 = (1.0 - alpha) / alpha;
    const double r = c.r, r3 = c.r3, q2 = c.q2, xi2 = c.xi2, d = c.d, yp = c.yp;
    const double rd = r + d;
    const double d11 = 1.0 / (r * rd);
    const double aj2 = xi * yp / rd * d11;
    const double aj5 = -(d + yp * yp / rd) * d11;
    double ai3, ai4, ak1, ak3, aj3, aj6;
    if (cd != 0.0) {
        if (xi == 0.0) {
            ai4 = 0.0;
        } else {
            const double xs = sqrt(xi2 + q2);
            ai4 = 1.0 / cd / cd * (xi / rd * sd * cd + 2.0 * atan((et * (xs + q * cd) + xs * (r + xs) * sd) / (xi * (r + xs) * cd)));
        }
        ai3 = (yp * cd / rd - c.ale + sd * log(rd)) / cd / cd;
        ak1 = xi * (d11 - c.y11 * sd) / cd;
        ak3 = (q * c.y11 - yp * d11) / cd;
        aj3 = (ak1 - aj2 * sd) / cd;
        aj6 = (ak3 - aj5 * sd) / cd;
    } else {
        const double rd2 = rd * rd;
        ai3 = (et / rd + yp * q / rd2 - c.ale) / 2.0;
        ai4 = xi * yp / rd2 / 2.0;
        ak1 = xi * q / rd * d11;
        ak3 = sd / rd * (xi2 * d11 - 1.0);
        aj3 = -xi / rd2 * (q2 * d11 - 0.5);
        aj6 = -yp / rd2 * (xi2 * d11 - 0.5);
    }
    const double xy = xi * c.y11;
    const double ai1 = -xi / rd * cd - ai4 * sd;
    const double ai2 = log(rd) + ai3 * sd;
    const double ak2 = 1.0 / r + ak3 * sd;
    const double ak4 = xy * cd - ak1 * sd;
    const double aj1 = aj5 * cd - aj6 * sd;
    const double aj4 = -xy - aj2 * cd + aj3 * sd;
    const double qx = q * c.x11, qy = q * c.y11;
    const double f1 = s1 / pi2, f2 = s2 / pi2;
    o[0] = (0.0 + f1 * (-xi * qy - c.tt - alp3 * ai1 * sd)) + f2 * (-q / r + alp3 * ai3 * sd * cd);
    o[1] = (0.0 + f1 * (-q / r + alp3 * yp / rd * sd)) + f2 * (-et * qx - c.tt - alp3 * xi / rd * sd * cd);
    o[2] = (0.0 + f1 * (q * qy - alp3 * ai2 * sd)) + f2 * (q * qx + alp3 * ai4 * sd * cd);
    o[3] = (0.0 + f1 * (xi2 * q * c.y32 - alp3 * aj1 * sd)) + f2 * (xi * q / r3 + alp3 * aj4 * sd * cd);
    o[4] = (0.0 + f1 * (xi * q / r3 - alp3 * aj2 * sd)) + f2 * (et * q / r3 + qy + alp3 * aj5 * sd * cd);
    o[5] = (0.0 + f1 * (-xi * q2 * c.y32 - alp3 * aj3 * sd)) + f2 * (-q2 / r3 + alp3 * aj6 * sd * cd);
    o[6] = (0.0 + f1 * (-xi * c.fy - d * c.x11 + alp3 * (xy + aj4) * sd)) + f2 * (-c.ey + alp3 * aj1 * sd * cd);
    o[7] = (0.0 + f1 * (-c.ey + alp3 * (1.0 / r + aj5) * sd)) + f2 * (-et * c.gy - xy * sd + alp3 * aj2 * sd * cd);
    o[8] = (0.0 + f1 * (q * c.fy - alp3 * (qy - aj6) * sd)) + f2 * (q * c.gy + alp3 * aj3 * sd * cd);
    o[9] = (0.0 + f1 * (-xi * c.fz - yp * c.x11 + alp3 * ak1 * sd)) + f2 * (-c.ez - alp3 * ak3 * sd * cd);
    o[10] = (0.0 + f1 * (-c.ez + alp3 * yp * d11 * sd)) + f2 * (-et * c.gz - xy * cd - alp3 * xi * d11 * sd * cd);
    o[11] = (0.0 + f1 * (q * c.fz + alp3 * ak2 * sd)) + f2 * (q * c.gz - alp3 * ak4 * sd * cd);
}

__device__ __noinline__ void ok_part_c(const OkGeo &c, double xi, double et, double q, double z, double sd, double cd,
                                       double s1, double s2, double alpha, double *o)
{
#pragma clang fp contract(off)
    const double pi2 = 8.0 * atan(1.0);
    const double alp4 = 1.0 - alpha, alp5 = alpha;
    const double r = c.r, r2 = c.r2, r3 = c.r3, r5 = c.r5, q2 = c.q2, xi2 = c.xi2, et2 = c.et2, d = c.d, yp = c.yp;
    const double x11 = c.x11, y11 = c.y11, x32 = c.x32, y32 = c.y32;
    const double cc = d + z;
    const double x53 = (8.0 * r2 + 9.0 * r * xi + 3.0 * xi2) * x11 * x11 * x11 / r2;
    const double y53 = (8.0 * r2 + 9.0 * r * et + 3.0 * et2) * y11 * y11 * y11 / r2;
    const double h = q * cd - z;
    const double z32 = sd / r3 - h * y32;
    const double z53 = 3.0 * sd / r5 - h * y53;
    const double y0 = y11 - xi2 * y32;
    const double z0 = z32 - xi2 * z53;
    const double ppy = cd / r3 + q * y32 * sd;
    const double ppz = sd / r3 - q * y32 * cd;
    const double qq = z * y32 + z32 + z0;
    const double qqy = 3.0 * cc * d / r5 - qq * sd;
    const double qqz = 3.0 * cc * yp / r5 - qq * cd + q * y32;
    const double xy = xi * y11, qy = q * y11;
    const double qr = 3.0 * q / r5;
    const double cdr = (cc + d) / r3;
    const double yy0 = yp / r3 - y0 * cd;
    const double f1 = s1 / pi2, f2 = s2 / pi2;
    o[0] = (0.0 + f1 * (alp4 * xy * cd - alp5 * xi * q * z32)) + f2 * (alp4 * cd / r - qy * sd - alp5 * cc * q / r3);
    o[1] = (0.0 + f1 * (alp4 * (cd / r + 2.0 * qy * sd) - alp5 * cc * q / r3)) + f2 * (alp4 * yp * x11 - alp5 * cc * et * q * x32);
    o[2] = (0.0 + f1 * (alp4 * qy * cd - alp5 * (cc * et / r3 - z * y11 + xi2 * z32))) +
           f2 * (-d * x11 - xy * sd - alp5 * cc * (x11 - q2 * x32));
    o[3] = (0.0 + f1 * (alp4 * y0 * cd - alp5 * q * z0)) + f2 * (-alp4 * xi / r3 * cd + alp5 * cc * xi * qr + xi * q * y32 * sd);
    o[4] = (0.0 + f1 * (-alp4 * xi * (cd / r3 + 2.0 * q * y32 * sd) + alp5 * cc * xi * qr)) + f2 * (-alp4 * yp / r3 + alp5 * cc * et * qr);
    o[5] = (0.0 + f1 * (-alp4 * xi * q * y32 * cd + alp5 * xi * (3.0 * cc * et / r5 - qq))) +
           f2 * (d / r3 - y0 * sd + alp5 * cc / r3 * (1.0 - 3.0 * q2 / r2));
    o[6] = (0.0 + f1 * (-alp4 * xi * ppy * cd - alp5 * xi * qqy)) + f2 * (-alp4 * et / r3 + y0 * sd * sd - alp5 * (cdr * sd - cc * yp * qr));
    o[7] = (0.0 + f1 * (alp4 * 2.0 * (d / r3 - y0 * sd) * sd - yp / r3 * cd - alp5 * (cdr * sd - et / r3 - cc * yp * qr))) +
           f2 * (alp4 * (x11 - yp * yp * x32) - alp5 * cc * ((d + 2.0 * q * cd) * x32 - yp * et * q * x53));
    o[8] = (0.0 + f1 * (-alp4 * q / r3 + yy0 * sd + alp5 * (cdr * cd + cc * d * qr - (y0 * cd + q * z0) * sd))) +
           f2 * (xi * ppy * sd + yp * d * x32 + alp5 * cc * ((yp + 2.0 * q * sd) * x32 - yp * q2 * x53));
    o[9] = (0.0 + f1 * (alp4 * xi * ppz * cd - alp5 * xi * qqz)) + f2 * (-q / r3 + y0 * sd * cd - alp5 * (cdr * cd + cc * d * qr));
    o[10] = (0.0 + f1 * (alp4 * 2.0 * (yp / r3 - y0 * cd) * sd + d / r3 * cd - alp5 * (cdr * cd + cc * d * qr))) +
            f2 * (alp4 * yp * d * x32 - alp5 * cc * ((yp - 2.0 * q * sd) * x32 + d * et * q * x53));
    o[11] = (0.0 + f1 * (yy0 * cd - alp5 * (cdr * sd - cc * yp * qr - y0 * sd * sd + q * z0 * cd))) +
            f2 * (-xi * ppz * sd + x11 - d * d * x32 - alp5 * cc * ((d - 2.0 * q * cd) * x32 - d * q2 * x53));
}

// displacement gradients u[3..11] = (uxx,uyx,uzx,uxy,uyy,uzy,uxz,uyz,uzz) of a
// finite rectangle with strike slip s1 and dip slip s2 (okada.f90:10-226)
__device__ __noinline__ void ok_dc3d(double alpha, double x, double y, double z, double depth, double dip,
                                     double al1, double al2, double aw1, double aw2, double s1, double s2,
                                     bool fullspace, double *u)
{
#pragma clang fp contract(off)
    const double pi2 = 8.0 * atan(1.0), epso = 1e-6;
    const double p18 = pi2 / 360.0;
    double sd = sin(dip * p18), cd = cos(dip * p18);
    if (fabs(cd) < epso) {
        cd = 0.0;
        if (sd > 0.0) sd = 1.0;
        if (sd < 0.0) sd = -1.0;
    }
    const double xi[2] = {x - al1, x - al2};
    double d = depth + z;
    double p = y * cd + d * sd;
    double q = y * sd - d * cd;
    double et[2] = {p - aw1, p - aw2};
    double da[12], db[12], dcv[12], du[12];
    for (int i = 0; i < 12; i++) u[i] = 0.0;
    OkGeo c;
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++) {
            ok_geo(c, xi[j], et[k], q, sd, cd);
            ok_part_a(c, xi[j], et[k], q, sd, cd, s1, s2, alpha, da);
            for (int b = 0; b < 12; b += 3) {
                du[b] = -da[b];
                du[b + 1] = -da[b + 1] * cd + da[b + 2] * sd;
                du[b + 2] = -da[b + 1] * sd - da[b + 2] * cd;
            }
            du[9] = da[9];
            du[10] = da[10] * cd - da[11] * sd;
            du[11] = da[10] * sd + da[11] * cd;
            const bool add = (j + k) != 1;
            for (int i = 0; i < 12; i++) u[i] = add ? u[i] + du[i] : u[i] - du[i];
        }
    if (fullspace) return;
    d = depth - z;
    p = y * cd + d * sd;
    q = y * sd - d * cd;
    et[0] = p - aw1;
    et[1] = p - aw2;
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++) {
            ok_geo(c, xi[j], et[k], q, sd, cd);
            ok_part_a(c, xi[j], et[k], q, sd, cd, s1, s2, alpha, da);
            ok_part_b(c, xi[j], et[k], q, sd, cd, s1, s2, alpha, db);
            ok_part_c(c, xi[j], et[k], q, z, sd, cd, s1, s2, alpha, dcv);
            for (int b = 0; b < 12; b += 3) {
                du[b] = da[b] + db[b] + z * dcv[b];
                du[b + 1] = (da[b + 1] + db[b + 1] + z * dcv[b + 1]) * cd - (da[b + 2] + db[b + 2] + z * dcv[b + 2]) * sd;
                du[b + 2] = (da[b + 1] + db[b + 1] - z * dcv[b + 1]) * sd + (da[b + 2] + db[b + 2] - z * dcv[b + 2]) * cd;
            }
            du[9] = du[9] + dcv[0];
            du[10] = du[10] + dcv[1] * cd - dcv[2] * sd;
            du[11] = du[11] - dcv[1] * sd - dcv[2] * cd;
            const bool add = (j + k) != 1;
            for (int i = 0; i < 12; i++) u[i] = add ? u[i] + du[i] : u[i] - du[i];
        }
}

// tensor rotation T2 = A^T T1 A of TensTrans (TDstressFS.f90:295-314); A[r][c] = A(r+1,c+1)
__device__ __forceinline__ void tens_trans(const double *t, const double A[3][3], double *o)
{
#pragma clang fp contract(off)
    const double Txx = t[0], Tyy = t[1], Tzz = t[2], Txy = t[3], Txz = t[4], Tyz = t[5];
    const double a11 = A[0][0], a12 = A[0][1], a13 = A[0][2];
    const double a21 = A[1][0], a22 = A[1][1], a23 = A[1][2];
    const double a31 = A[2][0], a32 = A[2][1], a33 = A[2][2];
    o[0] = a11 * a11 * Txx + 2 * a11 * a21 * Txy + 2 * a11 * a31 * Txz + 2 * a21 * a31 * Tyz + a21 * a21 * Tyy + a31 * a31 * Tzz;
    o[1] = a12 * a12 * Txx + 2 * a12 * a22 * Txy + 2 * a12 * a32 * Txz + 2 * a22 * a32 * Tyz + a22 * a22 * Tyy + a32 * a32 * Tzz;
    o[2] = a13 * a13 * Txx + 2 * a13 * a23 * Txy + 2 * a13 * a33 * Txz + 2 * a23 * a33 * Tyz + a23 * a23 * Tyy + a33 * a33 * Tzz;
    o[3] = a11 * a12 * Txx + (a11 * a22 + a21 * a12) * Txy + (a11 * a32 + a12 * a31) * Txz + (a32 * a21 + a31 * a22) * Tyz + a22 * a21 * Tyy + a31 * a32 * Tzz;
    o[4] = a11 * a13 * Txx + (a11 * a23 + a13 * a21) * Txy + (a11 * a33 + a31 * a13) * Txz + (a33 * a21 + a31 * a23) * Tyz + a23 * a21 * Tyy + a31 * a33 * Tzz;
    o[5] = a12 * a13 * Txx + (a13 * a22 + a12 * a23) * Txy + (a13 * a32 + a12 * a33) * Txz + (a23 * a32 + a33 * a22) * Tyz + a22 * a23 * Tyy + a32 * a33 * Tzz;
}

__device__ __noinline__ double okada_entry(const DevGeom &g, int i, int j, bool fullspace)
{
#pragma clang fp contract(off)
    const double pi = dpi();
    const double alpha = (1.0 + (0.5 / kPois - 1.0)) / (1.0 + 2.0 * (0.5 / kPois - 1.0));
    const double angj = g.ang[j];
    const double dxg = g.xcol[i] - g.xcol[j], dyg = g.ycol[i] - g.ycol[j];
    const double dx = cos(angj) * dxg + sin(angj) * dyg;
    const double dy = -sin(angj) * dxg + cos(angj) * dyg;
    const double w = g.w;
    const double z = g.zcol[i];
    const double depth = -g.zcol[j];
    const double dip = g.angd[j] * 180.0 / pi;
    const double rs = g.dsl[j] / w, rd = g.dsd[j] / w;
    double u[12];
    ok_dc3d(alpha, dx / w, dy / w, z / w, depth / w, dip, -0.5 * rs, 0.5 * rs, -0.5 * rd, 0.5 * rd,
            -cos(g.rake[j]), -sin(g.rake[j]), fullspace, u);
    const double exx = u[3], eyy = u[7], ezz = u[11];
    const double exy = 0.5 * (u[6] + u[4]), eyz = 0.5 * (u[10] + u[8]), ezx = 0.5 * (u[5] + u[9]);
    const double R = kRigid;
    double s[6];
    s[0] = R * (exx + eyy + ezz) + 2 * R * exx;
    s[1] = R * (exx + eyy + ezz) + 2 * R * eyy;
    s[2] = R * (exx + eyy + ezz) + 2 * R * ezz;
    s[3] = 2 * R * exy;
    s[4] = 2 * R * ezx;
    s[5] = 2 * R * eyz;
    const double rot = angj - g.ang[i];
    const double dp = g.angd[i];
    const double A[3][3] = {{cos(rot), sin(rot) * cos(dp), sin(rot) * sin(dp)},
                            {-sin(rot), cos(rot) * cos(dp), cos(rot) * sin(dp)},
                            {0.0, sin(dp), -cos(dp)}};
    double p[6];
    tens_trans(s, A, p);
    if (g.vcomp == 1) return p[2];
    return p[4] * cos(g.rake[i]) + p[5] * sin(g.rake[i]);
}

// ------------------------------------------------------------- triangles ----
__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - b[1] * a[2];
    c[1] = a[2] * b[0] - b[2] * a[0];
    c[2] = a[0] * b[1] - b[0] * a[1];
}

// strain of an angular dislocation (Nikkhoo & Walter 2015), TDstressFS.f90:560-627
__device__ __noinline__ void ang_dis_strain(double x, double y, double z, double alpha, double bx, double by, double bz,
                                            double nu, double *e)
{
#pragma clang fp contract(off)
    const double pi = dpi();
    const double sinA = sin(alpha), cosA = cos(alpha);
    const double eta = y * cosA - z * sinA;
    const double zeta = y * sinA + z * cosA;
    const double x2 = x * x, y2 = y * y, z2 = z * z;
    const double r2 = x2 + y2 + z2;
    const double r = sqrt(r2);
    const double r3 = r * r2;
    const double rz = r * (r - z);
    const double r2z2 = r2 * ((r - z) * (r - z));
    const double r3z = r3 * (r - z);
    const double W = zeta - r;
    const double W2 = W * W;
    const double Wr = W * r;
    const double W2r = W2 * r;
    const double Wr3 = W * r3;
    const double W2r2 = W2 * r2;
    const double C = (r * cosA - z) / Wr;
    const double S = (r * sinA - y) / Wr;
    const double rFi_rx = (eta / r / (r - zeta) - y / r / (r - z)) / 4 / pi;
    const double rFi_ry = (x / r / (r - z) - cosA * x / r / (r - zeta)) / 4 / pi;
    const double rFi_rz = (sinA * x / r / (r - zeta)) / 4 / pi;
    e[0] = bx * (rFi_rx) + bx / 8 / pi / (1 - nu) * (eta / Wr + eta * x2 / W2r2 - eta * x2 / Wr3 + y / rz -
        x2 * y / r2z2 - x2 * y / r3z) - by * x / 8 / pi / (1 - nu) * (((2 * nu + 1) / Wr + x2 / W2r2 - x2 / Wr3) * cosA +
        (2 * nu + 1) / rz - x2 / r2z2 - x2 / r3z) + bz * x * sinA / 8 / pi / (1 - nu) * ((2 * nu + 1) / Wr + x2 / W2r2 - x2 / Wr3);
    e[1] = by * (rFi_ry) + bx / 8 / pi / (1 - nu) * ((1 / Wr + S * S - y2 / Wr3) * eta + (2 * nu + 1) * y / rz - y * y * y / r2z2 -
        y * y * y / r3z - 2 * nu * cosA * S) - by * x / 8 / pi / (1 - nu) * (1 / rz - y2 / r2z2 - y2 / r3z +
        (1 / Wr + S * S - y2 / Wr3) * cosA) + bz * x * sinA / 8 / pi / (1 - nu) * (1 / Wr + S * S - y2 / Wr3);
    e[2] = bz * (rFi_rz) + bx / 8 / pi / (1 - nu) * (eta / W / r + eta * C * C - eta * z2 / Wr3 + y * z / r3 +
        2 * nu * sinA * C) - by * x / 8 / pi / (1 - nu) * ((1 / Wr + C * C - z2 / Wr3) * cosA + z / r3) +
        bz * x * sinA / 8 / pi / (1 - nu) * (1 / Wr + C * C - z2 / Wr3);
    e[3] = bx * (rFi_ry) / 2 + by * (rFi_rx) / 2 - bx / 8 / pi / (1 - nu) * (x * y2 / r2z2 - nu * x / rz +
        x * y2 / r3z - nu * x * cosA / Wr + eta * x * S / Wr + eta * x * y / Wr3) +
        by / 8 / pi / (1 - nu) * (x2 * y / r2z2 - nu * y / rz + x2 * y / r3z + nu * cosA * S +
        x2 * y * cosA / Wr3 + x2 * cosA * S / Wr) - bz * sinA / 8 / pi / (1 - nu) * (nu * S + x2 * S / Wr + x2 * y / Wr3);
    e[4] = bx * (rFi_rz) / 2 + bz * (rFi_rx) / 2 - bx / 8 / pi / (1 - nu) * (-x * y / r3 + nu * x * sinA / Wr + eta * x * C / Wr +
        eta * x * z / Wr3) + by / 8 / pi / (1 - nu) * (-x2 / r3 + nu / r + nu * cosA * C + x2 * z * cosA / Wr3 +
        x2 * cosA * C / Wr) - bz * sinA / 8 / pi / (1 - nu) * (nu * C + x2 * C / Wr + x2 * z / Wr3);
    e[5] = by * (rFi_rz) / 2 + bz * (rFi_ry) / 2 + bx / 8 / pi / (1 - nu) * (y2 / r3 - nu / r - nu * cosA * C +
        nu * sinA * S + eta * sinA * cosA / W2 - eta * (y * cosA + z * sinA) / W2r + eta * y * z / W2r2 - eta * y * z / Wr3) -
        by * x / 8 / pi / (1 - nu) * (y / r3 + sinA * (cosA * cosA) / W2 - cosA * (y * cosA + z * sinA) /
        W2r + y * z * cosA / W2r2 - y * z * cosA / Wr3) - bz * x * sinA / 8 / pi / (1 - nu) *
        (y * z / Wr3 - sinA * cosA / W2 + (y * cosA + z * sinA) / W2r - y * z / W2r2);
}

// TDSetupD, TDstressFS.f90:352-403
__device__ __noinline__ void td_setup_d(double x, double y, double z, double alpha, double bx, double by, double bz,
                                        double nu, const double *tv, const double *sv, double *e)
{
#pragma clang fp contract(off)
    const double pi = dpi();
    const double A11 = sv[2], A12 = -sv[1], A21 = sv[1], A22 = sv[2];
    const double y1 = A11 * (y - tv[1]) + A12 * (z - tv[2]);
    const double z1 = A21 * (y - tv[1]) + A22 * (z - tv[2]);
    const double by1 = A11 * by + A12 * bz;
    const double bz1 = A21 * by + A22 * bz;
    double ea[6];
    ang_dis_strain(x, y1, z1, -pi + alpha, bx, by1, bz1, nu, ea);
    const double B[3][3] = {{1.0, 0.0, 0.0}, {0.0, A11, A12}, {0.0, A21, A22}};
    tens_trans(ea, B, e);
}

__device__ __forceinline__ int tri_mode(double x, double y, double z, const double *p1, const double *p2, const double *p3)
{
#pragma clang fp contract(off)
    const double den = (p2[1] - p3[1]) * (p1[0] - p3[0]) + (p3[0] - p2[0]) * (p1[1] - p3[1]);
    const double a = ((p2[1] - p3[1]) * (x - p3[0]) + (p3[0] - p2[0]) * (y - p3[1])) / den;
    const double b = ((p3[1] - p1[1]) * (x - p3[0]) + (p1[0] - p3[0]) * (y - p3[1])) / den;
    const double c = 1.0 - a - b;
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

// full-space stress of a triangular dislocation, TDstressFS.f90:45-235
__device__ __noinline__ void td_stress_fs(double X, double Y, double Z, const double *P1, const double *P2, const double *P3,
                                          double Ss, double Ds, double Ts, double mu, double lambda, double *Sout)
{
#pragma clang fp contract(off)
    const double nu = 1 / (1 + lambda / mu) / 2;
    const double bx = Ts, by = Ss, bz = Ds;
    double a[3], b[3], vn[3], vs[3], vd[3];
    for (int k = 0; k < 3; k++) { a[k] = P2[k] - P1[k]; b[k] = P3[k] - P1[k]; }
    cross3(a, b, vn);
    double nn = sqrt(dot3(vn, vn));
    for (int k = 0; k < 3; k++) vn[k] = vn[k] / nn;
    const double eY[3] = {0.0, 1.0, 0.0}, eZ[3] = {0.0, 0.0, 1.0};
    cross3(eZ, vn, vs);
    nn = sqrt(dot3(vs, vs));
    for (int k = 0; k < 3; k++) vs[k] = vs[k] / nn;
    if (dot3(vs, vs) == 0.0)
        for (int k = 0; k < 3; k++) vs[k] = eY[k] * vn[2];
    nn = sqrt(dot3(vs, vs));
    for (int k = 0; k < 3; k++) vs[k] = vs[k] / nn;
    cross3(vn, vs, vd);
    double At[3][3];
    for (int k = 0; k < 3; k++) { At[0][k] = vn[k]; At[1][k] = vs[k]; At[2][k] = vd[k]; }
    auto ct = [&](double a0, double a1, double a2, double *o) {
        o[0] = At[0][0] * a0 + At[0][1] * a1 + At[0][2] * a2;
        o[1] = At[1][0] * a0 + At[1][1] * a1 + At[1][2] * a2;
        o[2] = At[2][0] * a0 + At[2][1] * a1 + At[2][2] * a2;
    };
    double xyz[3], p1[3], p2[3] = {0.0, 0.0, 0.0}, p3[3];
    ct(X - P2[0], Y - P2[1], Z - P2[2], xyz);
    ct(P1[0] - P2[0], P1[1] - P2[1], P1[2] - P2[2], p1);
    ct(P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2], p3);
    double e12[3], e13[3], e23[3];
    for (int k = 0; k < 3; k++) { e12[k] = p2[k] - p1[k]; e13[k] = p3[k] - p1[k]; e23[k] = p3[k] - p2[k]; }
    nn = sqrt(dot3(e12, e12)); for (int k = 0; k < 3; k++) e12[k] = e12[k] / nn;
    nn = sqrt(dot3(e13, e13)); for (int k = 0; k < 3; k++) e13[k] = e13[k] / nn;
    nn = sqrt(dot3(e23, e23)); for (int k = 0; k < 3; k++) e23[k] = e23[k] / nn;
    const double m12[3] = {-e12[0], -e12[1], -e12[2]}, m13[3] = {-e13[0], -e13[1], -e13[2]}, m23[3] = {-e23[0], -e23[1], -e23[2]};
    const double A = acos(dot3(e12, e13));
    const double B = acos(dot3(m12, e23));
    const double C = acos(dot3(e23, e13));
    const int mode = tri_mode(xyz[1], xyz[2], xyz[0], p1 + 1, p2 + 1, p3 + 1);
    double e1[6], e2[6], e3[6], e[6];
    if (mode == 1) {
        td_setup_d(xyz[0], xyz[1], xyz[2], A, bx, by, bz, nu, p1, m13, e1);
        td_setup_d(xyz[0], xyz[1], xyz[2], B, bx, by, bz, nu, p2, e12, e2);
        td_setup_d(xyz[0], xyz[1], xyz[2], C, bx, by, bz, nu, p3, e23, e3);
    } else {
        td_setup_d(xyz[0], xyz[1], xyz[2], A, bx, by, bz, nu, p1, e13, e1);
        td_setup_d(xyz[0], xyz[1], xyz[2], B, bx, by, bz, nu, p2, m12, e2);
        td_setup_d(xyz[0], xyz[1], xyz[2], C, bx, by, bz, nu, p3, m23, e3);
    }
    for (int k = 0; k < 6; k++) e[k] = e1[k] + e2[k] + e3[k];
    if (mode == 0)
        for (int k = 0; k < 6; k++) e[k] = __builtin_nan("");
    double E[6];
    tens_trans(e, At, E);
    Sout[0] = 2 * mu * E[0] + lambda * (E[0] + E[1] + E[2]);
    Sout[1] = 2 * mu * E[1] + lambda * (E[0] + E[1] + E[2]);
    Sout[2] = 2 * mu * E[2] + lambda * (E[0] + E[1] + E[2]);
    Sout[3] = 2 * mu * E[3];
    Sout[4] = 2 * mu * E[4];
    Sout[5] = 2 * mu * E[5];
}

// ------------------------------------------------- half-space triangles ----
// Value and gradient with respect to the three coordinates of the point.
struct Dual {
    double v, d[3];
};
__device__ __forceinline__ Dual dvar(double v, int k)
{
    Dual r{v, {0.0, 0.0, 0.0}};
    r.d[k] = 1.0;
    return r;
}
__device__ __forceinline__ Dual operator+(Dual a, Dual b)
{
#pragma clang fp contract(off)
    return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__device__ __forceinline__ Dual operator-(Dual a, Dual b)
{
#pragma clang fp contract(off)
    return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, {-a.d[0], -a.d[1], -a.d[2]}}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b)
{
#pragma clang fp contract(off)
    return {a.v * b.v, {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
__device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
#pragma clang fp contract(off)
    const double q = a.v / b.v;
    return {q, {(a.d[0] - q * b.d[0]) / b.v, (a.d[1] - q * b.d[1]) / b.v, (a.d[2] - q * b.d[2]) / b.v}};
}
__device__ __forceinline__ Dual operator+(Dual a, double c) { return {a.v + c, {a.d[0], a.d[1], a.d[2]}}; }
__device__ __forceinline__ Dual operator+(double c, Dual a) { return a + c; }
__device__ __forceinline__ Dual operator-(Dual a, double c) { return {a.v - c, {a.d[0], a.d[1], a.d[2]}}; }
__device__ __forceinline__ Dual operator-(double c, Dual a) { return {c - a.v, {-a.d[0], -a.d[1], -a.d[2]}}; }
__device__ __forceinline__ Dual operator*(Dual a, double c) { return {a.v * c, {a.d[0] * c, a.d[1] * c, a.d[2] * c}}; }
__device__ __forceinline__ Dual operator*(double c, Dual a) { return a * c; }
__device__ __forceinline__ Dual operator/(Dual a, double c) { return {a.v / c, {a.d[0] / c, a.d[1] / c, a.d[2] / c}}; }
__device__ __forceinline__ Dual operator/(double c, Dual a)
{
#pragma clang fp contract(off)
    const double q = c / a.v;
    return {q, {-q * a.d[0] / a.v, -q * a.d[1] / a.v, -q * a.d[2] / a.v}};
}
__device__ __forceinline__ Dual dsqrt(Dual a)
{
#pragma clang fp contract(off)
    const double r = sqrt(a.v);
    return {r, {a.d[0] / (2 * r), a.d[1] / (2 * r), a.d[2] / (2 * r)}};
}
__device__ __forceinline__ Dual dlog(Dual a) { return {log(a.v), {a.d[0] / a.v, a.d[1] / a.v, a.d[2] / a.v}}; }

// Displacement of the free-surface correction of an angular dislocation at
// depth a with angle beta (Nikkhoo & Walter 2015, GJI 201:1117-1139, the
// harmonic half of the half-space solution), one Burgers component at a time;
// fib carries the gradient of the Burgers function (its value never reaches
// a strain).
__device__ __noinline__ void fsc_disp(Dual y1, Dual y2, Dual y3, Dual fib, double beta, double b1, double b2,
                                      double b3, double nu, double a, Dual *u)
{
#pragma clang fp contract(off)
    const double pi = dpi();
    const double sinB = sin(beta), cosB = cos(beta), cotB = cosB / sinB, cot2 = cotB * cotB;
    const double N1 = 1 - 2 * nu, f = 1 / (4 * pi * (1 - nu));
    const Dual y3b = y3 + 2 * a;
    const Dual z1b = y1 * cosB + y3b * sinB;
    const Dual z3b = -y1 * sinB + y3b * cosB;
    const Dual rb = dsqrt(y1 * y1 + y2 * y2 + y3b * y3b);
    const Dual rr = rb * rb, r3 = rr * rb;
    const Dual ar = a / rb, ar2 = a / rr;
    const Dual p6 = rb + y3b, p7 = rb + z3b;
    const Dual h = y3b - a;
    const Dual q = rb * cosB + y3b;
    Dual u1{0.0, {0.0, 0.0, 0.0}}, u2 = u1, u3 = u1;
    if (b1 != 0.0) {
        u1 = b1 * f * (-2 * (1 - nu) * N1 * fib * cot2 + N1 * y2 / p6 * ((N1 - ar) * cotB - y1 / p6 * (nu + ar))
             + N1 * y2 * cosB * cotB / p7 * (cosB + ar) + a * y2 * h * cotB / r3
             + y2 * h / (rb * p6) * (-N1 * cotB + y1 / p6 * (2 * nu + ar) + a * y1 / rr)
             + y2 * h / (rb * p7) * (cosB / p7 * (q * (N1 * cosB - ar) * cotB + 2 * (1 - nu) * (rb * sinB - y1) * cosB)
                                     - a * y3b * cosB * cotB / rr));
        u2 = b1 * f * (N1 * ((2 * (1 - nu) * cot2 - nu) * dlog(p6) - (2 * (1 - nu) * cot2 + N1) * cosB * dlog(p7))
             - N1 / p6 * (y1 * cotB * (N1 - ar) + nu * y3b - a + y2 * y2 / p6 * (nu + ar))
             - N1 * z1b * cotB / p7 * (cosB + ar) - a * y1 * h * cotB / r3
             + h / p6 * (-2 * nu + (N1 * y1 * cotB - a) / rb + y2 * y2 / (rb * p6) * (2 * nu + ar) + a * y2 * y2 / r3)
             + h / p7 * (cosB * cosB - (N1 * z1b * cotB + a * cosB) / rb + a * y3b * z1b * cotB / r3
                         - (y2 * y2 * cosB * cosB - a * z1b * cotB / rb * q) / (rb * p7)));
        u3 = b1 * f * (2 * (1 - nu) * (N1 * fib * cotB + y2 / p6 * (2 * nu + ar) - y2 * cosB / p7 * (cosB + ar))
             + y2 * h / rb * (2 * nu / p6 + ar2) + y2 * h * cosB / (rb * p7) * (N1 - q / p7 * (cosB + ar) - a * y3b / rr));
    }
    if (b2 != 0.0) {
        u1 = u1 + b2 * f * (N1 * ((2 * (1 - nu) * cot2 + nu) * dlog(p6) - (2 * (1 - nu) * cot2 + 1) * cosB * dlog(p7))
             + N1 / p6 * (-N1 * y1 * cotB + nu * y3b - a + a * y1 * cotB / rb + y1 * y1 / p6 * (nu + ar))
             - N1 * cotB / p7 * (z1b * cosB - a * (rb * sinB - y1) / (rb * cosB)) - a * y1 * h * cotB / r3
             + h / p6 * (2 * nu + (N1 * y1 * cotB + a) / rb - y1 * y1 / (rb * p6) * (2 * nu + ar) - a * y1 * y1 / r3)
             + h * cotB / p7 * (-cosB * sinB + a * y1 * y3b / (r3 * cosB)
                                + (rb * sinB - y1) / rb * (2 * (1 - nu) * cosB - q / p7 * (1 + a / (rb * cosB)))));
        u2 = u2 + b2 * f * (2 * (1 - nu) * N1 * fib * cot2 + N1 * y2 / p6 * (-(N1 - ar) * cotB + y1 / p6 * (nu + ar))
             - N1 * y2 * cotB / p7 * (1 + a / (rb * cosB)) - a * y2 * h * cotB / r3
             + y2 * h / (rb * p6) * (N1 * cotB - 2 * nu * y1 / p6 - a * y1 / rb * (1 / rb + 1 / p6))
             + y2 * h * cotB / (rb * p7) * (-2 * (1 - nu) * cosB + q / p7 * (1 + a / (rb * cosB)) + a * y3b / (rr * cosB)));
        u3 = u3 + b2 * f * (-2 * (1 - nu) * N1 * cotB * (dlog(p6) - cosB * dlog(p7)) - 2 * (1 - nu) * y1 / p6 * (2 * nu + ar)
             + 2 * (1 - nu) * z1b / p7 * (cosB + ar) + h / rb * (N1 * cotB - 2 * nu * y1 / p6 - a * y1 / rr)
             - h / p7 * (cosB * sinB + q * cotB / rb * (2 * (1 - nu) * cosB - q / p7)
                         + ar * (sinB - y3b * z1b / rr - z1b * q / (rb * p7))));
    }
    if (b3 != 0.0) {
        u1 = u1 + b3 * f * (N1 * (y2 / p6 * (1 + ar) - y2 * cosB / p7 * (cosB + ar)) - y2 * h / rb * (ar2 + 1 / p6)
             + y2 * h * cosB / (rb * p7) * (q / p7 * (cosB + ar) + a * y3b / rr));
        u2 = u2 + b3 * f * (N1 * (-sinB * dlog(p7) - y1 / p6 * (1 + ar) + z1b / p7 * (cosB + ar)) + y1 * h / rb * (ar2 + 1 / p6)
             - h / p7 * (sinB * (cosB - ar) + z1b / rb * (1 + a * y3b / rr)
                         - (y2 * y2 * cosB * sinB - a * z1b / rb * q) / (rb * p7)));
        u3 = u3 + b3 * f * (2 * (1 - nu) * fib + 2 * (1 - nu) * y2 * sinB / p7 * (cosB + ar)
             + y2 * h * sinB / (rb * p7) * (1 + q / p7 * (cosB + ar) + a * y3b / rr));
    }
    u[0] = u1; u[1] = u2; u[2] = u3;
}

// free-surface correction strains v11 v22 v33 v12 v13 v23 of an angular
// dislocation (the quantity of AngDisStrainFSC, TDstressFS.f90:629-1011)
__device__ __noinline__ void fsc_strain(double y1, double y2, double y3, double beta, double b1, double b2, double b3,
                                        double nu, double a, double *v)
{
#pragma clang fp contract(off)
    const double y3b = y3 + 2 * a;
    const double rb = sqrt(y1 * y1 + y2 * y2 + y3b * y3b);
    const double c = cos(beta / 2) / sin(beta / 2);
    const double D = y1 - (rb + y3b) * c, den = D * D + y2 * y2;
    // gradient of 2 atan(-y2 / (y1 - (r + y3bar) cot(beta/2)))
    const Dual fib{0.0, {2 * y2 * (1 - c * y1 / rb) / den, -2 * (D + c * y2 * y2 / rb) / den,
                         -2 * y2 * c * (y3b / rb + 1) / den}};
    Dual u[3];
    fsc_disp(dvar(y1, 0), dvar(y2, 1), dvar(y3, 2), fib, beta, b1, b2, b3, nu, a, u);
    v[0] = u[0].d[0];
    v[1] = u[1].d[1];
    v[2] = u[2].d[2];
    v[3] = (u[0].d[1] + u[1].d[0]) / 2;
    v[4] = (u[0].d[2] + u[2].d[0]) / 2;
    v[5] = (u[1].d[2] + u[2].d[1]) / 2;
}

// AngSetupFSC_S, TDstressFS.f90:405-558: the correction stress of side PA->PB
// (zero for a vertical side)
__device__ __noinline__ void fsc_side(double X, double Y, double Z, const double *bv, const double *PA, const double *PB,
                                      double mu, double lambda, double *S)
{
#pragma clang fp contract(off)
    const double pi = dpi();
    const double nu = lambda / (lambda + mu) / 2;
    const double sv[3] = {PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2]};
    const double beta = acos(-sv[2] / sqrt(dot3(sv, sv)));
    for (int k = 0; k < 6; k++) S[k] = 0.0;
    if (fabs(beta) < 1e-3 || fabs(pi - beta) < 1e-3) return;
    double e1[3] = {sv[0], sv[1], 0.0};
    const double nn = sqrt(dot3(e1, e1));
    for (int k = 0; k < 3; k++) e1[k] = e1[k] / nn;
    const double e3[3] = {0.0, 0.0, -1.0};
    double e2[3];
    cross3(e3, e1, e2);
    double A[3][3];
    for (int k = 0; k < 3; k++) { A[0][k] = e1[k]; A[1][k] = e2[k]; A[2][k] = e3[k]; }
    const double rel[3] = {X - PA[0], Y - PA[1], Z - PA[2]};
    double ya[3], yab[3], bb[3];
    for (int r = 0; r < 3; r++) {
        ya[r] = A[r][0] * rel[0] + A[r][1] * rel[1] + A[r][2] * rel[2];
        yab[r] = A[r][0] * sv[0] + A[r][1] * sv[1] + A[r][2] * sv[2];
        bb[r] = A[r][0] * bv[0] + A[r][1] * bv[1] + A[r][2] * bv[2];
    }
    const double ang = (beta * ya[0] >= 0.0) ? -pi + beta : beta;
    double vA[6], vB[6], v[6], E[6];
    fsc_strain(ya[0], ya[1], ya[2], ang, bb[0], bb[1], bb[2], nu, -PA[2], vA);
    fsc_strain(ya[0] - yab[0], ya[1] - yab[1], ya[2] - yab[2], ang, bb[0], bb[1], bb[2], nu, -PB[2], vB);
    for (int k = 0; k < 6; k++) v[k] = vB[k] - vA[k];
    tens_trans(v, A, E);
    const double tr = E[0] + E[1] + E[2];
    S[0] = 2 * mu * E[0] + lambda * tr;
    S[1] = 2 * mu * E[1] + lambda * tr;
    S[2] = 2 * mu * E[2] + lambda * tr;
    S[3] = 2 * mu * E[3];
    S[4] = 2 * mu * E[4];
    S[5] = 2 * mu * E[5];
}

// TDstressHS + TDstress_HarFunc, TDstressFS.f90:7-44, 236-293
__device__ __noinline__ void td_stress_hs(double X, double Y, double Z, const double *P1, const double *P2, const double *P3,
                                          double Ss, double Ds, double Ts, double mu, double lambda, double *S)
{
#pragma clang fp contract(off)
    double Sm[6], Si[6], s1[6], s2[6], s3[6];
    td_stress_fs(X, Y, Z, P1, P2, P3, Ss, Ds, Ts, mu, lambda, Sm);
    double a[3], b[3], vn[3], vs[3], vd[3];
    for (int k = 0; k < 3; k++) { a[k] = P2[k] - P1[k]; b[k] = P3[k] - P1[k]; }
    cross3(a, b, vn);
    double nn = sqrt(dot3(vn, vn));
    for (int k = 0; k < 3; k++) vn[k] = vn[k] / nn;
    const double eY[3] = {0.0, 1.0, 0.0}, eZ[3] = {0.0, 0.0, 1.0};
    cross3(eZ, vn, vs);
    // normalised before the zero test, as there: a horizontal triangle gives NaN
    nn = sqrt(dot3(vs, vs));
    for (int k = 0; k < 3; k++) vs[k] = vs[k] / nn;
    if (dot3(vs, vs) == 0.0)
        for (int k = 0; k < 3; k++) vs[k] = eY[k] * vn[2];
    nn = sqrt(dot3(vs, vs));
    for (int k = 0; k < 3; k++) vs[k] = vs[k] / nn;
    cross3(vn, vs, vd);
    double bg[3];
    for (int k = 0; k < 3; k++) bg[k] = vn[k] * Ts + vs[k] * Ss + vd[k] * Ds;
    fsc_side(X, Y, Z, bg, P1, P2, mu, lambda, s1);
    fsc_side(X, Y, Z, bg, P2, P3, mu, lambda, s2);
    fsc_side(X, Y, Z, bg, P3, P1, mu, lambda, s3);
    const double Q1[3] = {P1[0], P1[1], -P1[2]}, Q2[3] = {P2[0], P2[1], -P2[2]}, Q3[3] = {P3[0], P3[1], -P3[2]};
    td_stress_fs(X, Y, Z, Q1, Q2, Q3, Ss, Ds, Ts, mu, lambda, Si);
    for (int k = 0; k < 6; k++) S[k] = Sm[k] + Si[k] + (s1[k] + s2[k] + s3[k]);
}

// matel3dn_ij / matel3dh_ij (m_HACApK_calc_entry_ij.f90:802-898): mode 's'
// takes the receiver's rake in the full space, the source's in the half-space
__device__ __noinline__ double tri_entry(const DevGeom &g, int i, int j)
{
#pragma clang fp contract(off)
    const double lambda = 2 * kPois / (1 - 2 * kPois) * kRigid;
    const double P1[3] = {g.xs1[j], g.ys1[j], g.zs1[j]};
    const double P2[3] = {g.xs2[j], g.ys2[j], g.zs2[j]};
    const double P3[3] = {g.xs3[j], g.ys3[j], g.zs3[j]};
    const bool half = g.kind == 5;
    double ss = 0.0, dd = 0.0, tt = 0.0;
    if (g.md == 1) ss = 1.0;
    else if (g.md == 2) dd = 1.0;
    else if (g.md == 3) tt = 1.0;
    else { const int r = half ? j : i; ss = cos(g.rake[r]); dd = sin(g.rake[r]); }
    double S[6];
    if (half) td_stress_hs(g.xcol[i], g.ycol[i], g.zcol[i], P1, P2, P3, ss, dd, tt, kRigid, lambda, S);
    else td_stress_fs(g.xcol[i], g.ycol[i], g.zcol[i], P1, P2, P3, ss, dd, tt, kRigid, lambda, S);
    const double A[3][3] = {{g.ev11[i], g.ev21[i], g.ev31[i]},
                            {g.ev12[i], g.ev22[i], g.ev32[i]},
                            {g.ev13[i], g.ev23[i], g.ev33[i]}};
    double p[6];
    tens_trans(S, A, p);
    if (g.vcomp == 1) return -p[2];
    return p[4] * cos(g.rake[i]) + p[5] * sin(g.rake[i]);
}

__device__ __forceinline__ double antiplane(double xc, double xl, double xr)
{
#pragma clang fp contract(off)
    const double factor = kRigid / (2.0 * dpi());
    return factor * (1.0 / (xc - xr) - 1.0 / (xc - xl));
}

// ------------------------------------------------------------ 2-D, 2.5-D ----
// planar in-plane shear crack element, full space (matels2dp_ij)
__device__ __forceinline__ double inplane(double xc, double xl, double xr)
{
#pragma clang fp contract(off)
    const double factor = kRigid / (2.0 * dpi() * (1.0 - kPois));
    return factor * (1.0 / (xc - xr) - 1.0 / (xc - xl));
}

// planar antiplane segment with ten image pairs at k w, k = 1..10
// (matels2dpFZ_ij, m_HACApK_calc_entry_ij.f90:663-683); the image term's
// numerator dx^3 - 3 a^2 is the reference's as written
__device__ __forceinline__ double antiplane_fz(double xc, double xl, double xr, double w)
{
#pragma clang fp contract(off)
    const double dx1 = xc - xr, dx2 = xc - xl;
    double s = 1.0 / dx1 - 1.0 / dx2;
    for (int k = 1; k <= 10; k++) {
        const double a = k * w, a2 = a * a;
        const double t1 = a2 + dx1 * dx1, t2 = a2 + dx2 * dx2;
        s = s + 2 * dx1 * (dx1 * dx1 * dx1 - 3 * a2) / (t1 * t1 * t1) -
            2 * dx2 * (dx2 * dx2 * dx2 - 3 * a2) / (t2 * t2 * t2);
    }
    return s * (kRigid / (2.0 * dpi()));
}

// nonplanar in-plane segment, full space, from the traction vectors of its two
// end points (matel2dn_ij + kern, m_HACApK_calc_entry_ij.f90:472-562):
// shear (vcomp 0, 's') or normal (vcomp 1, 'n') traction on the receiver
__device__ __noinline__ double seg2dv_entry(const DevGeom &g, int i, int j)
{
#pragma clang fp contract(off)
    const double factor = kRigid / (2.0 * dpi() * (1.0 - kPois));
    const double nx = -sin(g.ang[i]), ny = cos(g.ang[i]);
    const double xe[2] = {g.xel[j], g.xer[j]}, ye[2] = {g.yel[j], g.yer[j]};
    double I1[2], I2[2];
    for (int e = 0; e < 2; e++) {
        const double dx = g.xcol[i] - xe[e], dy = g.ycol[i] - ye[e];
        const double r = sqrt(dx * dx + dy * dy);
        const double gx = dx / r, gy = dy / r;
        if (g.vcomp == 0) {
            const double t = 4 * nx * ny * gx * gy + (ny * ny - nx * nx) * (gy * gy - gx * gx);
            I1[e] = -(gy / r * t);
            I2[e] = -(gx / r * t);
        } else {
            const double t = 2 * nx * ny * (gy * gy - gx * gx) - 2 * (ny * ny - nx * nx) * gx * gy;
            I1[e] = gx / r + gy / r * t;
            I2[e] = gy / r + gx / r * t;
        }
    }
    const double sx = -sin(g.ang[j]), sy = cos(g.ang[j]);
    return factor * (sx * (-I1[1] + I1[0]) + sy * (I2[1] - I2[0]));
}

// nonplanar antiplane segment, full space (tensor2d3_ij + inte31s / inte32s,
// m_HACApK_calc_entry_ij.f90:378-393, 751-765): the out-of-plane stress pair
// of the segment's frame projected on the receiver's normal
__device__ __noinline__ double seg2da_entry(const DevGeom &g, int i, int j)
{
#pragma clang fp contract(off)
    const double pi = dpi();
    const double cj = cos(g.ang[j]), sj = sin(g.ang[j]);
    const double xp = cj * (g.xcol[i] - g.xel[j]) + sj * (g.ycol[i] - g.yel[j]);
    const double xm = cj * (g.xcol[i] - g.xer[j]) + sj * (g.ycol[i] - g.yer[j]);
    const double yp = -sj * (g.xcol[i] - g.xel[j]) + cj * (g.ycol[i] - g.yel[j]);
    const double ym = -sj * (g.xcol[i] - g.xer[j]) + cj * (g.ycol[i] - g.yer[j]);
    auto i31 = [&](double x1, double x2) {
        const double r = sqrt(x1 * x1 + x2 * x2);
        return -(kVs * x2 / pi / (r * r));
    };
    auto i32 = [&](double x1, double x2) {
        const double r = sqrt(x1 * x1 + x2 * x2);
        return kVs / pi * x1 / (r * r);
    };
    const double k31 = -0.5 * kRigid / kVs * (i31(xp, yp) - i31(xm, ym));
    const double k32 = -0.5 * kRigid / kVs * (i32(xp, yp) - i32(xm, ym));
    return sin(g.ang[i]) * k31 + cos(g.ang[i]) * k32;
}

// planar rectangle in a full space seen in its own plane (ret3dp): shear for
// slip modes 's'/'o', normal for 'open'; the other combinations are zero
__device__ __noinline__ double quad_entry(const DevGeom &g, int i, int j)
{
#pragma clang fp contract(off)
    const double xc = g.xcol[i], zc = g.zcol[i];
    const double dx1 = xc - g.xs1[j], dx2 = xc - g.xs2[j], dx3 = xc - g.xs3[j], dx4 = xc - g.xs4[j];
    const double dz1 = zc - g.zs1[j], dz2 = zc - g.zs2[j], dz3 = zc - g.zs3[j], dz4 = zc - g.zs4[j];
    const bool shear = g.md == 0 && g.vcomp == 0, open = g.md == 3 && g.vcomp == 1;
    if (!shear && !open) return 0.0;
    const double factor = kRigid / (4.0 * dpi());
    const double r1 = sqrt(dx1 * dx1 + dz1 * dz1), r2 = sqrt(dx2 * dx2 + dz2 * dz2);
    const double r3 = sqrt(dx3 * dx3 + dz3 * dz3), r4 = sqrt(dx4 * dx4 + dz4 * dz4);
    const double alpha = 1 / (2 - 2 * kPois);
    // corner sum of r / (dx dz): the part both modes share
    const double corners = 2.0 * (1 - alpha) * (r4 / (dx4 * dz4) + r2 / (dx2 * dz2) - r3 / (dx3 * dz3) - r1 / (dx1 * dz1));
    if (open) return factor * corners;
    const double e1 = -(1 - 2 * alpha) / dx1 * (r4 / dz4 + dz4 / r4 - r1 / dz1 - dz1 / r1);
    const double e2 = -(1 - 2 * alpha) / dx2 * (r2 / dz2 + dz2 / r2 - r3 / dz3 - dz3 / r3);
    return factor * (corners + e1 + e2);
}

// stress (sxx, syy, sxy) at (X, Y) of an edge dislocation at (xs, ys) with
// Burgers vector (s1, s2) in a half-space y > 0 with a free surface y = 0
// (D2dip2, m_HACApK_calc_entry_ij.f90:143-166)
__device__ __noinline__ void edge_dislocation_hs(double X, double Y, double xs, double ys, double s1, double s2,
                                                 double *o)
{
#pragma clang fp contract(off)
    const double dx = X - xs, ym = Y - ys, yp = Y + ys;
    const double dx2 = dx * dx, ym2 = ym * ym, yp2 = yp * yp;
    const double r1s = dx2 + ym2, r2s = dx2 + yp2;        // r1^2, r2^2
    const double r14 = r1s * r1s, r24 = r2s * r2s, r26 = r24 * r2s;
    const double coef = kRigid / 2.0 / dpi() / (1.0 - kPois);
    const double q = 6 * Y * yp * dx2;
    const double sxx1 = s1 * coef * (ym * (ym2 + 3 * dx2) / r14 - yp * (yp2 + 3 * dx2) / r24 +
                                     2 * ys / r26 * (q - ym * (yp2 * yp) - dx2 * dx2));
    const double sxx2 = s2 * coef * (dx * (ym2 - dx2) / r14 - dx * (yp2 - dx2) / r24 +
                                     4 * ys * dx / r26 * ((2 * ys - Y) * yp2 + (3 * Y + 2 * ys) * dx2));
    const double syy1 = s1 * coef * (ym * (ym2 - dx2) / r14 - yp * (yp2 - dx2) / r24 -
                                     2 * ys / r26 * (q - (3 * Y + ys) * (yp2 * yp) + dx2 * dx2));
    const double syy2 = -s2 * coef * (dx * (3 * ym2 + dx2) / r14 - dx * (3 * yp2 + dx2) / r24 -
                                      4 * ys * dx * Y / r26 * (3 * yp2 - dx2));
    const double sxy1 = s1 * coef * (dx * (ym2 - dx2) / r14 - dx * (yp2 - dx2) / r24 +
                                     4 * ys * dx * Y / r26 * (3 * yp2 - dx2));
    const double sxy2 = s2 * coef * (ym * (ym2 - dx2) / r14 - yp * (yp2 - dx2) / r24 +
                                     2 * ys / r26 * (q - dx2 * dx2 + (ys - Y) * (yp2 * yp)));
    o[0] = sxx1 + sxx2;
    o[1] = syy1 + syy2;
    o[2] = sxy1 + sxy2;
}

// dipping segment in a half-space: the two end dislocations (matel2dph2_ij)
__device__ __noinline__ double dip2d_entry(const DevGeom &g, int i, int j)
{
#pragma clang fp contract(off)
    const double s1 = cos(g.ang[j]), s2 = sin(g.ang[j]);
    double p[3], q[3];
    edge_dislocation_hs(g.xcol[i], g.ycol[i], g.xel[j], g.yel[j], s1, s2, p);
    edge_dislocation_hs(g.xcol[i], g.ycol[i], g.xer[j], g.yer[j], s1, s2, q);
    const double sxx = p[0] - q[0], syy = p[1] - q[1], sxy = p[2] - q[2];
    const double c2 = cos(2 * g.ang[i]), s2i = sin(2 * g.ang[i]);
    if (g.vcomp == 1) return -0.5 * (sxx + syy) + 0.5 * (sxx - syy) * c2 + sxy * s2i;
    return -0.5 * (sxx - syy) * s2i + sxy * c2;
}

// nonplanar in-plane segment, full space (tensor2d_ij): the segment's kernel
// in its own frame from the end-point integrals (inte11/12/22 s|o), rotated
// to global axes and projected on the receiver
__device__ __noinline__ double seg2d_entry(const DevGeom &g, int i, int j)
{
#pragma clang fp contract(off)
    const double pi = dpi();
    const double cj = cos(g.ang[j]), sj = sin(g.ang[j]);
    const double xp = cj * (g.xcol[i] - g.xel[j]) + sj * (g.ycol[i] - g.yel[j]);
    const double xm = cj * (g.xcol[i] - g.xer[j]) + sj * (g.ycol[i] - g.yer[j]);
    const double yp = -sj * (g.xcol[i] - g.xel[j]) + cj * (g.ycol[i] - g.yel[j]);
    const double ym = -sj * (g.xcol[i] - g.xer[j]) + cj * (g.ycol[i] - g.yer[j]);
    const double pa = kVs / 6.0;                        // vs / vp, m_const.f90:7-8
    const double c = 2 * kVs / pi * (1 - pa * pa);
    const bool open = g.md == 3;
    // the three integrals at one end point (x1, x2), shear or opening
    auto ends = [&](double x1, double x2, double *t) {
        const double r = sqrt(x1 * x1 + x2 * x2);
        const double r4 = (r * r) * (r * r);
        if (!open) {
            t[0] = -c * x2 * (3 * x1 * x1 + x2 * x2) / r4;
            t[1] = c * x1 * (x1 * x1 - x2 * x2) / r4;
            t[2] = c * x2 * (x1 * x1 - x2 * x2) / r4;
        } else {
            t[0] = c * x1 * (x1 * x1 - x2 * x2) / r4;
            t[1] = c * x2 * (x1 * x1 - x2 * x2) / r4;
            t[2] = c * x1 * (3 * x1 * x1 + x2 * x2) / r4;
        }
    };
    double tp[3], tm[3];
    ends(xp, yp, tp);
    ends(xm, ym, tm);
    const double k11 = -0.5 * kRigid / kVs * (tp[0] - tm[0]);
    const double k12 = -0.5 * kRigid / kVs * (tp[1] - tm[1]);
    const double k22 = -0.5 * kRigid / kVs * (tp[2] - tm[2]);
    const double sin2 = sin(-2 * g.ang[j]), cos2 = cos(-2 * g.ang[j]);
    const double sxx = 0.5 * (k11 + k22) + 0.5 * (k11 - k22) * cos2 + k12 * sin2;
    const double sxy = -0.5 * (k11 - k22) * sin2 + k12 * cos2;
    const double syy = 0.5 * (k11 + k22) - 0.5 * (k11 - k22) * cos2 - k12 * sin2;
    if (g.vcomp == 1)
        return -(0.5 * (sxx + syy) - 0.5 * (sxx - syy) * cos(2 * g.ang[i]) - sxy * sin(2 * g.ang[i]));
    return 0.5 * (sxx - syy) * sin(-2 * g.ang[i]) + sxy * cos(-2 * g.ang[i]);
}

// 2.5-D: a rectangle of length ds along strike and width w, slip along
// strike, stresses in its plane (matelrec_ij, dc3d25, ua25)
__device__ __noinline__ double rect25_entry(const DevGeom &g, int i, int j)
{
#pragma clang fp contract(off)
    const double pi = dpi();
    const double alpha = (1.0 + (0.5 / kPois - 1.0)) / (1.0 + 2.0 * (0.5 / kPois - 1.0));
    const double st = g.ang[j];
    const double x = cos(st) * (g.xcol[i] - g.xcol[j]) + sin(st) * (g.ycol[i] - g.ycol[j]);
    const double q = -sin(st) * (g.xcol[i] - g.xcol[j]) + cos(st) * (g.ycol[i] - g.ycol[j]);
    const double dl = g.ds[j], dw = g.w;
    const double xi[2] = {x + 0.5 * dl, x - 0.5 * dl};
    const double et[2] = {0.5 * dw, -0.5 * dw};
    // displacement gradients ux,x uy,x ux,y uy,y uz,z summed over the corners
    double uxx = 0.0, uyx = 0.0, uxy = 0.0, uyy = 0.0, uzz = 0.0;
    for (int k = 0; k < 2; k++)
        for (int jj = 0; jj < 2; jj++) {
            const double X = xi[jj], E = et[k];
            const double r = sqrt(X * X + E * E + q * q);
            const double y11 = 1.0 / (r * (r + E));
            const double x11 = 1.0 / (r * (r + X));
            const double y32 = (r + (r + E)) * y11 * y11 / r;
            const double r3 = r * r * r;
            const double d4 = (-(1.0 - alpha) / 2.0 * q * y11 - alpha / 2 * (X * X) * q * y32) / (pi * 2);
            const double d6 = ((1.0 - alpha) / 2.0 * X * y11 + alpha / 2 * X * (q * q) * y32) / (pi * 2);
            const double d7 = ((1.0 - alpha) / 2.0 * X * y11 + alpha / 2 * X * (E / r3 + (X * X) * y32) + E / 2.0 * x11) /
                              (pi * 2);
            const double d9 = ((1.0 - alpha) / 2.0 * (q * y11) - alpha / 2 * q * (E / r3 + (X * X) * y32)) / (pi * 2);
            const double d11 = (alpha / 2 * (E * q / r3)) / (pi * 2);
            const double sg = (jj + k == 1) ? -1.0 : 1.0;
            uxx = uxx + sg * -d4;
            uyx = uyx + sg * d6;
            uxy = uxy + sg * -d7;
            uyy = uyy + sg * d9;
            uzz = uzz + sg * d11;
        }
    const double exx = -uxx, eyy = -uyy, ezz = -uzz, exy = -0.5 * (uxy + uyx);
    const double sxx = kRigid * (exx + eyy + ezz) + 2 * kRigid * exx;
    const double syy = kRigid * (exx + eyy + ezz) + 2 * kRigid * eyy;
    const double sxy = 2 * kRigid * exy;
    const double d = 2 * (g.ang[i] - g.ang[j]);
    if (g.vcomp == 1) return -0.5 * (sxx + syy) + 0.5 * (sxx - syy) * cos(d) + sxy * sin(d);
    return -0.5 * (sxx - syy) * sin(d) + sxy * cos(d);
}

// HACApK_entry_ij: receiver i, source j (mesh order, 0-based)
__device__ __forceinline__ double entry(const DevGeom &g, int i, int j)
{
    switch (g.kind) {
    case 1: return antiplane(g.xcol[i], g.xel[j], g.xer[j]) - antiplane(g.xcol[i], -g.xel[j], -g.xer[j]);
    case 2: return okada_entry(g, i, j, false);
    case 3: return okada_entry(g, i, j, true);
    case 4:
    case 5: return tri_entry(g, i, j);
    case 6: return inplane(g.xcol[i], g.xel[j], g.xer[j]);
    case 7: return quad_entry(g, i, j);
    case 8: return dip2d_entry(g, i, j);
    case 9: return seg2d_entry(g, i, j);
    case 10: return rect25_entry(g, i, j);
    case 11: return antiplane_fz(g.xcol[i], g.xel[j], g.xer[j], g.w);
    case 12: return seg2dv_entry(g, i, j);
    case 13: return seg2da_entry(g, i, j);
    }
    return __builtin_nan("");
}

}  // namespace hbi
