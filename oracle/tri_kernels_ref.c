/*
 * ORACLE -- test infrastructure only.  Not shipped, not measured as the
 * product.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load this file's library.
 *
 * Plain-C restatement, line for line, of the P0 closed-form pair integrals of
 * the mollified Kelvin kernel in mbem's src/mbem/kernels/tri_kernels.py,
 * including its numerical guards (1e-300 inside the log, the 1e-30 / 1e-60
 * floors, degenerate-frame and zero-length-edge skips) and its operation
 * order in the cancelling terms.  Every function names the reference lines it
 * follows.  Build with -ffp-contract=off so no fused multiply-add changes the
 * rounding of the reference's expression order (numba compiles without
 * contraction by default).
 *
 * Layout conventions (tri_kernels.py:40-41): matrix out[3f+i, 3s+j] is row
 * block f = field point, column block s = source triangle, row-major with
 * 3*ns columns; basis stacks are (nb, 3nf, 3ns) contiguous.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---- edge antiderivatives: tri_kernels.py:54-76 and :606-619 ---------- */
static inline double J1(double u, double rho2) { return log(u + sqrt(u * u + rho2) + 1e-300); }
static inline double J3(double u, double rho2)
{
    if (rho2 < 1e-60) return 0.0;
    return u / (rho2 * sqrt(u * u + rho2));
}
static inline double K1(double u, double rho2) { return sqrt(u * u + rho2); }
static inline double K3(double u, double rho2)
{
    double R = sqrt(u * u + rho2);
    if (R < 1e-30) R = 1e-30;
    return -1.0 / R;
}
static inline double J5(double u, double rho2)
{
    if (rho2 < 1e-60) return 0.0;
    double R = sqrt(u * u + rho2);
    return u * (3.0 * rho2 + 2.0 * u * u) / (3.0 * rho2 * rho2 * R * R * R);
}
static inline double K5(double u, double rho2)
{
    double R = sqrt(u * u + rho2);
    if (R < 1e-30) return 0.0;
    return -1.0 / (3.0 * R * R * R);
}

/* x ** n for a small non-negative integer n (repeated multiplication). */
static inline double ipow(double x, int n)
{
    double r = 1.0;
    for (int k = 0; k < n; ++k) r *= x;
    return r;
}

/* van Oosterom signed solid angle: tri_kernels.py:79-96 */
static double solid_angle(const double *v1, const double *v2, const double *v3,
                          double ox, double oy, double oz)
{
    double r1x = v1[0] - ox, r1y = v1[1] - oy, r1z = v1[2] - oz;
    double r2x = v2[0] - ox, r2y = v2[1] - oy, r2z = v2[2] - oz;
    double r3x = v3[0] - ox, r3y = v3[1] - oy, r3z = v3[2] - oz;
    double R1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z);
    double R2 = sqrt(r2x * r2x + r2y * r2y + r2z * r2z);
    double R3 = sqrt(r3x * r3x + r3y * r3y + r3z * r3z);
    double cx = r2y * r3z - r2z * r3y;
    double cy = r2z * r3x - r2x * r3z;
    double cz = r2x * r3y - r2y * r3x;
    double numer = r1x * cx + r1y * cy + r1z * cz;
    double d12 = r1x * r2x + r1y * r2y + r1z * r2z;
    double d23 = r2x * r3x + r2y * r3y + r2z * r3z;
    double d13 = r1x * r3x + r1y * r3y + r1z * r3z;
    double denom = R1 * R2 * R3 + R3 * d12 + R1 * d23 + R2 * d13;
    return 2.0 * atan2(numer, denom);
}

/* edge integral of xi1^a xi2^b / R^m, a,b <= 2: tri_kernels.py:99-129 */
static double edge_int(int a, int b, double dp, double nx, double ny, double tx, double ty,
                       double dJ, double dK, double dL)
{
    double total = 0.0;
    for (int i = 0; i <= a; ++i) {
        for (int j = 0; j <= b; ++j) {
            double ca = (a == 2 && i == 1) ? 2.0 : 1.0;
            double cb = (b == 2 && j == 1) ? 2.0 : 1.0;
            double coeff = ca * cb * ipow(dp, a + b - i - j) * ipow(nx, a - i) * ipow(tx, i)
                           * ipow(ny, b - j) * ipow(ty, j);
            int k = i + j;
            double delta = (k == 0) ? dJ : (k == 1 ? dK : dL);
            total += coeff * delta;
        }
    }
    return total;
}

/* binomial C(n,k), n,k <= 3: tri_kernels.py:622-629 */
static inline double binom3(int n, int k)
{
    if (k <= 0 || k >= n) return 1.0;
    if (n == 2) return 2.0;
    return 3.0;
}

/* edge integral with a+b <= 3: tri_kernels.py:632-652 */
static double edge_int3(int a, int b, double dp, double nx, double ny, double tx, double ty,
                        double dJ, double dK, double dL, double dM)
{
    double total = 0.0;
    for (int i = 0; i <= a; ++i) {
        for (int j = 0; j <= b; ++j) {
            double coeff = binom3(a, i) * binom3(b, j) * ipow(dp, a + b - i - j) * ipow(nx, a - i)
                           * ipow(tx, i) * ipow(ny, b - j) * ipow(ty, j);
            int k = i + j;
            double delta = (k == 0) ? dJ : (k == 1 ? dK : (k == 2 ? dL : dM));
            total += coeff * delta;
        }
    }
    return total;
}

/* in-plane frame of a source triangle: tri_kernels.py:136-153 */
static int tri_frame(const double *tv, double *ex, double *ey, double *nhat)
{
    double e1x = tv[3] - tv[0], e1y = tv[4] - tv[1], e1z = tv[5] - tv[2];
    double e2x = tv[6] - tv[0], e2y = tv[7] - tv[1], e2z = tv[8] - tv[2];
    double nx = e1y * e2z - e1z * e2y;
    double ny = e1z * e2x - e1x * e2z;
    double nz = e1x * e2y - e1y * e2x;
    double area2 = sqrt(nx * nx + ny * ny + nz * nz);
    if (area2 < 1e-30) return 0;
    nhat[0] = nx / area2; nhat[1] = ny / area2; nhat[2] = nz / area2;
    double e1n = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
    ex[0] = e1x / e1n; ex[1] = e1y / e1n; ex[2] = e1z / e1n;
    ey[0] = nhat[1] * ex[2] - nhat[2] * ex[1];
    ey[1] = nhat[2] * ex[0] - nhat[0] * ex[2];
    ey[2] = nhat[0] * ex[1] - nhat[1] * ex[0];
    return 1;
}

/* per-pair setup shared by every pair kernel: obs projection, the lifted
 * point's solid angle (I3 = -Omega/h) and the projected vertices
 * (tri_kernels.py:164-182, :257-273, :663-679). */
typedef struct {
    double z, h2, h, I3;
    double px[3], py[3];
} pair_geom;

static void pair_setup(const double *tv, const double *ex, const double *ey, const double *nhat,
                       const double *obs, double eps, pair_geom *g)
{
    double dx = obs[0] - tv[0], dy = obs[1] - tv[1], dz = obs[2] - tv[2];
    double z = dx * nhat[0] + dy * nhat[1] + dz * nhat[2];
    double h2 = z * z + eps * eps;
    double h = sqrt(h2);
    double opx = obs[0] - z * nhat[0];
    double opy = obs[1] - z * nhat[1];
    double opz = obs[2] - z * nhat[2];
    for (int k = 0; k < 3; ++k) {
        double wx = tv[3 * k] - opx, wy = tv[3 * k + 1] - opy, wz = tv[3 * k + 2] - opz;
        g->px[k] = wx * ex[0] + wy * ex[1] + wz * ex[2];
        g->py[k] = wx * ey[0] + wy * ey[1] + wz * ey[2];
    }
    double oex = opx + h * nhat[0], oey = opy + h * nhat[1], oez = opz + h * nhat[2];
    double Omega = solid_angle(tv, tv + 3, tv + 6, oex, oey, oez);
    g->I3 = -Omega / h;
    g->z = z; g->h2 = h2; g->h = h;
}

/* U-kernel basis blocks out[b][i][j], b = 0..2: tri_kernels.py:160-243 */
static void u_basis_pair(const double *tv, const double *ex, const double *ey, const double *nhat,
                         const double *obs, double eps, double *out)
{
    pair_geom g;
    pair_setup(tv, ex, ey, nhat, obs, eps, &g);
    double z = g.z, h2 = g.h2, I3 = g.I3;
    double BD001 = 0.0, BN1_001 = 0.0, BN2_001 = 0.0, BN1_101 = 0.0, BN2_101 = 0.0;
    double BN1_011 = 0.0, BN2_011 = 0.0;
    for (int e = 0; e < 3; ++e) {
        double ax = g.px[e], ay = g.py[e];
        double bx = g.px[(e + 1) % 3], by = g.py[(e + 1) % 3];
        double evx = bx - ax, evy = by - ay;
        double L = sqrt(evx * evx + evy * evy);
        if (L <= 1e-30) continue;
        double tx = evx / L, ty = evy / L;
        double nx = ty, ny = -tx;
        double dp = ax * nx + ay * ny;
        double ua = ax * tx + ay * ty;
        double ub = bx * tx + by * ty;
        double rho2 = dp * dp + h2;
        double dJ1 = J1(ub, rho2) - J1(ua, rho2);
        double dK1 = K1(ub, rho2) - K1(ua, rho2);
        double v00 = edge_int(0, 0, dp, nx, ny, tx, ty, dJ1, dK1, 0.0);
        double v10 = edge_int(1, 0, dp, nx, ny, tx, ty, dJ1, dK1, 0.0);
        double v01 = edge_int(0, 1, dp, nx, ny, tx, ty, dJ1, dK1, 0.0);
        BD001 += dp * v00;
        BN1_001 += nx * v00; BN2_001 += ny * v00;
        BN1_101 += nx * v10; BN2_101 += ny * v10;
        BN1_011 += nx * v01; BN2_011 += ny * v01;
    }
    (void)BN2_101;
    double I1 = BD001 - h2 * I3;
    double m10 = -BN1_001, m01 = -BN2_001;
    double m20 = I1 - BN1_101, m11 = -BN1_011, m02 = I1 - BN2_011;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            out[0 * 9 + i * 3 + j] = 0.0;
            out[1 * 9 + i * 3 + j] = 0.0;
            out[2 * 9 + i * 3 + j] = 0.0;
        }
        out[0 * 9 + i * 3 + i] = I1;
        out[1 * 9 + i * 3 + i] = eps * eps * I3;
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            out[2 * 9 + i * 3 + j] = ex[i] * ex[j] * m20 + (ex[i] * ey[j] + ey[i] * ex[j]) * m11
                                     + ey[i] * ey[j] * m02
                                     - z * (ex[i] * nhat[j] + nhat[i] * ex[j]) * m10
                                     - z * (ey[i] * nhat[j] + nhat[i] * ey[j]) * m01
                                     + z * z * nhat[i] * nhat[j] * I3;
}

/* N (lam slot) / R (mu slot) contractions: tri_kernels.py:430-444 */
static void contract_NR(const double *P, const double *n, double *out, int idxN, int idxR)
{
    for (int i = 0; i < 3; ++i) {
        double tr = P[i * 9 + 0] + P[i * 9 + 4] + P[i * 9 + 8];
        for (int j = 0; j < 3; ++j) {
            out[idxN * 9 + i * 3 + j] = n[j] * tr;
            out[idxR * 9 + i * 3 + j] = P[i * 9 + j * 3 + 0] * n[0] + P[i * 9 + j * 3 + 1] * n[1]
                                        + P[i * 9 + j * 3 + 2] * n[2] + n[0] * P[i * 9 + 0 * 3 + j]
                                        + n[1] * P[i * 9 + 1 * 3 + j] + n[2] * P[i * 9 + 2 * 3 + j];
        }
    }
}

/* T-kernel basis blocks out[b][i][j], b = 0..5: tri_kernels.py:250-427 */
static void t_basis_pair(const double *tv, const double *ex, const double *ey, const double *nhat,
                         const double *nsrc, const double *obs, double eps, double *out)
{
    pair_geom g;
    pair_setup(tv, ex, ey, nhat, obs, eps, &g);
    double z = g.z, h2 = g.h2, I3 = g.I3;
    double BN1_001 = 0.0, BN2_001 = 0.0, BD003 = 0.0, BN1_003 = 0.0, BN2_003 = 0.0;
    double BN1_103 = 0.0, BN1_013 = 0.0, BN2_013 = 0.0, BN1_203 = 0.0, BN1_113 = 0.0;
    double BN1_023 = 0.0, BN2_023 = 0.0;
    for (int e = 0; e < 3; ++e) {
        double ax = g.px[e], ay = g.py[e];
        double bx = g.px[(e + 1) % 3], by = g.py[(e + 1) % 3];
        double evx = bx - ax, evy = by - ay;
        double L = sqrt(evx * evx + evy * evy);
        if (L <= 1e-30) continue;
        double tx = evx / L, ty = evy / L;
        double nx = ty, ny = -tx;
        double dp = ax * nx + ay * ny;
        double ua = ax * tx + ay * ty;
        double ub = bx * tx + by * ty;
        double rho2 = dp * dp + h2;
        double dJ1 = J1(ub, rho2) - J1(ua, rho2);
        double dK1 = K1(ub, rho2) - K1(ua, rho2);
        double dJ3 = J3(ub, rho2) - J3(ua, rho2);
        double dK3 = K3(ub, rho2) - K3(ua, rho2);
        double dL3 = dJ1 - rho2 * dJ3;
        double v00m1 = edge_int(0, 0, dp, nx, ny, tx, ty, dJ1, dK1, 0.0);
        double v00 = edge_int(0, 0, dp, nx, ny, tx, ty, dJ3, dK3, dL3);
        double v10 = edge_int(1, 0, dp, nx, ny, tx, ty, dJ3, dK3, dL3);
        double v01 = edge_int(0, 1, dp, nx, ny, tx, ty, dJ3, dK3, dL3);
        double v20 = edge_int(2, 0, dp, nx, ny, tx, ty, dJ3, dK3, dL3);
        double v11 = edge_int(1, 1, dp, nx, ny, tx, ty, dJ3, dK3, dL3);
        double v02 = edge_int(0, 2, dp, nx, ny, tx, ty, dJ3, dK3, dL3);
        BN1_001 += nx * v00m1; BN2_001 += ny * v00m1;
        BD003 += dp * v00;
        BN1_003 += nx * v00; BN2_003 += ny * v00;
        BN1_103 += nx * v10;
        BN1_013 += nx * v01; BN2_013 += ny * v01;
        BN1_203 += nx * v20;
        BN1_113 += nx * v11;
        BN1_023 += nx * v02; BN2_023 += ny * v02;
    }
    double I5 = (BD003 + I3) / (3.0 * h2);
    double m10_3 = -BN1_001, m01_3 = -BN2_001;
    double m10_5 = -BN1_003 / 3.0, m01_5 = -BN2_003 / 3.0;
    double m20_5 = (I3 - BN1_103) / 3.0;
    double m11_5 = -BN1_013 / 3.0;
    double m02_5 = (I3 - BN2_013) / 3.0;
    double m30_5 = (2.0 * m10_3 - BN1_203) / 3.0;
    double m21_5 = (m01_3 - BN1_113) / 3.0;
    double m12_5 = -BN1_023 / 3.0;
    double m03_5 = (2.0 * m01_3 - BN2_023) / 3.0;
    double V3[3], V5[3];
    for (int i = 0; i < 3; ++i) {
        V3[i] = -ex[i] * m10_3 - ey[i] * m01_3 + nhat[i] * z * I3;
        V5[i] = -ex[i] * m10_5 - ey[i] * m01_5 + nhat[i] * z * I5;
    }
    /* T3[5] by frame expansion: tri_kernels.py:350-394 */
    double T35[27];
    for (int q = 0; q < 27; ++q) T35[q] = 0.0;
    const double *basv[3] = {ex, ey, nhat};
    for (int c1 = 0; c1 < 3; ++c1)
        for (int c2 = 0; c2 < 3; ++c2)
            for (int c3 = 0; c3 < 3; ++c3) {
                int a = (c1 == 0) + (c2 == 0) + (c3 == 0);
                int b = (c1 == 1) + (c2 == 1) + (c3 == 1);
                int nz = 3 - a - b;
                double mom;
                if (a == 3) mom = m30_5;
                else if (a == 2 && b == 1) mom = m21_5;
                else if (a == 2 && b == 0) mom = m20_5;
                else if (a == 1 && b == 2) mom = m12_5;
                else if (a == 1 && b == 1) mom = m11_5;
                else if (a == 1 && b == 0) mom = m10_5;
                else if (a == 0 && b == 3) mom = m03_5;
                else if (a == 0 && b == 2) mom = m02_5;
                else if (a == 0 && b == 1) mom = m01_5;
                else mom = I5;
                double sign = ((a + b) % 2 == 1) ? -1.0 : 1.0;
                double coeff = sign * ipow(z, nz) * mom;
                if (coeff == 0.0) continue;
                const double *b1 = basv[c1], *b2 = basv[c2], *b3 = basv[c3];
                for (int i = 0; i < 3; ++i) {
                    double ci = coeff * b1[i];
                    for (int j = 0; j < 3; ++j) {
                        double cij = ci * b2[j];
                        for (int k = 0; k < 3; ++k) T35[i * 9 + j * 3 + k] += cij * b3[k];
                    }
                }
            }
    double e2 = eps * eps;
    double P[27];
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k)
            for (int m = 0; m < 3; ++m) P[i * 9 + k * 3 + m] = (i == k) ? V3[m] : 0.0;
    contract_NR(P, nsrc, out, 0, 3);
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k)
            for (int m = 0; m < 3; ++m) {
                double val = -3.0 * T35[i * 9 + k * 3 + m];
                if (i == m) val += V3[k];
                if (k == m) val += V3[i];
                P[i * 9 + k * 3 + m] = val;
            }
    contract_NR(P, nsrc, out, 1, 4);
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k)
            for (int m = 0; m < 3; ++m) P[i * 9 + k * 3 + m] = (i == k) ? e2 * V5[m] : 0.0;
    contract_NR(P, nsrc, out, 2, 5);
}

/* 2D area moments for the stress kernels: tri_kernels.py:655-769.
 * M3 is 3x3, M5 4x4, M7 5x5 row-major; returns I3, I5, I7, z. */
static void stress_2d_moments(const double *tv, const double *ex, const double *ey,
                              const double *nhat, const double *obs, double eps, double *M3,
                              double *M5, double *M7, int need7, double *oI3, double *oI5,
                              double *oI7, double *oz)
{
    pair_geom g;
    pair_setup(tv, ex, ey, nhat, obs, eps, &g);
    double h2 = g.h2, I3 = g.I3;
    double BN1m1 = 0.0, BN2m1 = 0.0, BD003 = 0.0, BD005 = 0.0;
    double bn1_3[9] = {0}, bn2_3[9] = {0}, bn1_5[16] = {0}, bn2_5[16] = {0};
    for (int e = 0; e < 3; ++e) {
        double ax = g.px[e], ay = g.py[e];
        double bx = g.px[(e + 1) % 3], by = g.py[(e + 1) % 3];
        double evx = bx - ax, evy = by - ay;
        double L = sqrt(evx * evx + evy * evy);
        if (L <= 1e-30) continue;
        double tx = evx / L, ty = evy / L;
        double nx = ty, ny = -tx;
        double dp = ax * nx + ay * ny;
        double ua = ax * tx + ay * ty;
        double ub = bx * tx + by * ty;
        double rho2 = dp * dp + h2;
        double dJ1 = J1(ub, rho2) - J1(ua, rho2);
        double dK1 = K1(ub, rho2) - K1(ua, rho2);
        double dJ3 = J3(ub, rho2) - J3(ua, rho2);
        double dK3 = K3(ub, rho2) - K3(ua, rho2);
        double dL3 = dJ1 - rho2 * dJ3;
        double v = edge_int3(0, 0, dp, nx, ny, tx, ty, dJ1, dK1, 0.0, 0.0);
        BN1m1 += nx * v; BN2m1 += ny * v;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3 - a; ++b) {
                v = edge_int3(a, b, dp, nx, ny, tx, ty, dJ3, dK3, dL3, 0.0);
                bn1_3[a * 3 + b] += nx * v;
                bn2_3[a * 3 + b] += ny * v;
                if (a == 0 && b == 0) BD003 += dp * v;
            }
        if (need7) {
            double dJ5 = J5(ub, rho2) - J5(ua, rho2);
            double dK5 = K5(ub, rho2) - K5(ua, rho2);
            double dL5 = dJ3 - rho2 * dJ5;
            double dM5 = dK3 - rho2 * dK5;
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4 - a; ++b) {
                    v = edge_int3(a, b, dp, nx, ny, tx, ty, dJ5, dK5, dL5, dM5);
                    bn1_5[a * 4 + b] += nx * v;
                    bn2_5[a * 4 + b] += ny * v;
                    if (a == 0 && b == 0) BD005 += dp * v;
                }
        }
    }
    double I5 = (h2 > 1e-60) ? (BD003 + I3) / (3.0 * h2) : 0.0;
#define A3(a, b) M3[(a) * 3 + (b)]
#define A5(a, b) M5[(a) * 4 + (b)]
#define A7(a, b) M7[(a) * 5 + (b)]
    A3(0, 0) = I3;
    A3(1, 0) = -BN1m1;
    A3(0, 1) = -BN2m1;
    A5(0, 0) = I5;
    A5(1, 0) = -bn1_3[0 * 3 + 0] / 3.0;
    A5(0, 1) = -bn2_3[0 * 3 + 0] / 3.0;
    A5(2, 0) = (I3 - bn1_3[1 * 3 + 0]) / 3.0;
    A5(1, 1) = -bn1_3[0 * 3 + 1] / 3.0;
    A5(0, 2) = (I3 - bn2_3[0 * 3 + 1]) / 3.0;
    A5(3, 0) = (2.0 * A3(1, 0) - bn1_3[2 * 3 + 0]) / 3.0;
    A5(2, 1) = (A3(0, 1) - bn1_3[1 * 3 + 1]) / 3.0;
    A5(1, 2) = -bn1_3[0 * 3 + 2] / 3.0;
    A5(0, 3) = (2.0 * A3(0, 1) - bn2_3[0 * 3 + 2]) / 3.0;
    double I7 = 0.0;
    if (need7) {
        I7 = (h2 > 1e-60) ? (BD005 + 3.0 * I5) / (5.0 * h2) : 0.0;
        A7(0, 0) = I7;
        A7(1, 0) = -bn1_5[0 * 4 + 0] / 5.0;
        A7(0, 1) = -bn2_5[0 * 4 + 0] / 5.0;
        A7(2, 0) = (I5 - bn1_5[1 * 4 + 0]) / 5.0;
        A7(1, 1) = -bn1_5[0 * 4 + 1] / 5.0;
        A7(0, 2) = (I5 - bn2_5[0 * 4 + 1]) / 5.0;
        A7(3, 0) = (2.0 * A5(1, 0) - bn1_5[2 * 4 + 0]) / 5.0;
        A7(2, 1) = (A5(0, 1) - bn1_5[1 * 4 + 1]) / 5.0;
        A7(1, 2) = -bn1_5[0 * 4 + 2] / 5.0;
        A7(0, 3) = (2.0 * A5(0, 1) - bn2_5[0 * 4 + 2]) / 5.0;
        A7(4, 0) = (3.0 * A5(2, 0) - bn1_5[3 * 4 + 0]) / 5.0;
        A7(3, 1) = (2.0 * A5(1, 1) - bn1_5[2 * 4 + 1]) / 5.0;
        A7(2, 2) = (A5(0, 2) - bn1_5[1 * 4 + 2]) / 5.0;
        A7(1, 3) = -bn1_5[0 * 4 + 3] / 5.0;
        A7(0, 4) = (3.0 * A5(0, 2) - bn2_5[0 * 4 + 3]) / 5.0;
    }
#undef A3
#undef A5
#undef A7
    *oI3 = I3; *oI5 = I5; *oI7 = I7; *oz = g.z;
}

/* T2[n][i,j] from 2D moments M (row stride ms): tri_kernels.py:772-784 */
static void build_T2(const double *M, int ms, const double *ex, const double *ey,
                     const double *nhat, double z, double *out)
{
#define MM(a, b) M[(a) * ms + (b)]
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            out[i * 3 + j] = ex[i] * ex[j] * MM(2, 0) + (ex[i] * ey[j] + ey[i] * ex[j]) * MM(1, 1)
                             + ey[i] * ey[j] * MM(0, 2)
                             - z * (ex[i] * nhat[j] + nhat[i] * ex[j]) * MM(1, 0)
                             - z * (ey[i] * nhat[j] + nhat[i] * ey[j]) * MM(0, 1)
                             + z * z * nhat[i] * nhat[j] * MM(0, 0);
#undef MM
}

/* T3[i,j,k] = int d_i d_j d_k / R^5: tri_kernels.py:787-820 */
static void build_T3(const double *M, int ms, const double *bas, double z, double *out)
{
    for (int q = 0; q < 27; ++q) out[q] = 0.0;
    for (int c1 = 0; c1 < 3; ++c1)
        for (int c2 = 0; c2 < 3; ++c2)
            for (int c3 = 0; c3 < 3; ++c3) {
                int a = 0, b = 0;
                if (c1 == 0) a += 1; else if (c1 == 1) b += 1;
                if (c2 == 0) a += 1; else if (c2 == 1) b += 1;
                if (c3 == 0) a += 1; else if (c3 == 1) b += 1;
                int nz = 3 - a - b;
                double sign = ((a + b) % 2 == 1) ? -1.0 : 1.0;
                double coeff = sign * ipow(z, nz) * M[a * ms + b];
                if (coeff == 0.0) continue;
                for (int i = 0; i < 3; ++i) {
                    double ci = coeff * bas[c1 * 3 + i];
                    for (int j = 0; j < 3; ++j) {
                        double cij = ci * bas[c2 * 3 + j];
                        for (int k = 0; k < 3; ++k) out[i * 9 + j * 3 + k] += cij * bas[c3 * 3 + k];
                    }
                }
            }
}

/* T4[i,j,k,l] = int d_i d_j d_k d_l / R^7: tri_kernels.py:823-853 */
static void build_T4(const double *M, int ms, const double *bas, double z, double *out)
{
    for (int q = 0; q < 81; ++q) out[q] = 0.0;
    for (int c1 = 0; c1 < 3; ++c1)
        for (int c2 = 0; c2 < 3; ++c2)
            for (int c3 = 0; c3 < 3; ++c3)
                for (int c4 = 0; c4 < 3; ++c4) {
                    int cs[4] = {c1, c2, c3, c4};
                    int a = 0, b = 0;
                    for (int q = 0; q < 4; ++q) {
                        if (cs[q] == 0) a += 1; else if (cs[q] == 1) b += 1;
                    }
                    int nz = 4 - a - b;
                    double sign = ((a + b) % 2 == 1) ? -1.0 : 1.0;
                    double coeff = sign * ipow(z, nz) * M[a * ms + b];
                    if (coeff == 0.0) continue;
                    for (int i = 0; i < 3; ++i) {
                        double ci = coeff * bas[c1 * 3 + i];
                        for (int j = 0; j < 3; ++j) {
                            double cij = ci * bas[c2 * 3 + j];
                            for (int k = 0; k < 3; ++k) {
                                double cijk = cij * bas[c3 * 3 + k];
                                for (int l = 0; l < 3; ++l)
                                    out[i * 27 + j * 9 + k * 3 + l] += cijk * bas[c4 * 3 + l];
                            }
                        }
                    }
                }
}

/* force -> stress S[i,j,k]: tri_kernels.py:856-897 */
static void kelvin_stress_pair(const double *tv, const double *ex, const double *ey,
                               const double *nhat, const double *obs, double eps, double mu,
                               double nu, double lam, double *S)
{
    double M3[9] = {0}, M5[16] = {0}, M7[25] = {0};
    double I3, I5, I7, z;
    stress_2d_moments(tv, ex, ey, nhat, obs, eps, M3, M5, M7, 0, &I3, &I5, &I7, &z);
    double V3[3], V5[3];
    for (int i = 0; i < 3; ++i) {
        V3[i] = -ex[i] * M3[1 * 3 + 0] - ey[i] * M3[0 * 3 + 1] + z * nhat[i] * M3[0];
        V5[i] = -ex[i] * M5[1 * 4 + 0] - ey[i] * M5[0 * 4 + 1] + z * nhat[i] * M5[0];
    }
    double bas[9] = {ex[0], ex[1], ex[2], ey[0], ey[1], ey[2], nhat[0], nhat[1], nhat[2]};
    double T35[27];
    build_T3(M5, 4, bas, z, T35);
    double C1 = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
    double c34 = 3.0 - 4.0 * nu;
    double cbdg = 6.0 * (1.0 - nu) * eps * eps;
    double G1[27];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double dij = (i == j) ? 1.0 : 0.0;
            for (int m = 0; m < 3; ++m) {
                double dim = (i == m) ? 1.0 : 0.0;
                double djm = (j == m) ? 1.0 : 0.0;
                G1[i * 9 + j * 3 + m] = C1 * (-c34 * dij * V3[m] + dim * V3[j] + djm * V3[i]
                                              - 3.0 * T35[i * 9 + j * 3 + m] - cbdg * dij * V5[m]);
            }
        }
    for (int k = 0; k < 3; ++k) {
        double trace_k = G1[0 * 9 + k * 3 + 0] + G1[1 * 9 + k * 3 + 1] + G1[2 * 9 + k * 3 + 2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double dij = (i == j) ? 1.0 : 0.0;
                S[i * 9 + j * 3 + k] = lam * dij * trace_k + mu * G1[i * 9 + k * 3 + j]
                                       + mu * G1[j * 9 + k * 3 + i];
            }
    }
}

/* slip -> TOTAL stress H[m,n,k]: tri_kernels.py:900-964 */
static void dd_stress_pair(const double *tv, const double *ex, const double *ey,
                           const double *nhat, const double *nrm, const double *obs, double eps,
                           double mu, double nu, double lam, double *H)
{
    double M3[9] = {0}, M5[16] = {0}, M7[25] = {0};
    double I3, I5, I7, z;
    stress_2d_moments(tv, ex, ey, nhat, obs, eps, M3, M5, M7, 1, &I3, &I5, &I7, &z);
    double bas[9] = {ex[0], ex[1], ex[2], ey[0], ey[1], ey[2], nhat[0], nhat[1], nhat[2]};
    double T2_5[9], T2_7[9], T4_7[81], ID2G[81], B[27];
    build_T2(M5, 4, ex, ey, nhat, z, T2_5);
    build_T2(M7, 5, ex, ey, nhat, z, T2_7);
    build_T4(M7, 5, bas, z, T4_7);
    double C1 = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
    double c34 = 3.0 - 4.0 * nu;
    double cblob = 2.0 * (1.0 - nu) * eps * eps;
    for (int r = 0; r < 3; ++r)
        for (int p = 0; p < 3; ++p) {
            double drp = (r == p) ? 1.0 : 0.0;
            for (int s = 0; s < 3; ++s) {
                double drs = (r == s) ? 1.0 : 0.0;
                double dps = (p == s) ? 1.0 : 0.0;
                for (int q = 0; q < 3; ++q) {
                    double dsq = (s == q) ? 1.0 : 0.0;
                    double dpq = (p == q) ? 1.0 : 0.0;
                    double drq = (r == q) ? 1.0 : 0.0;
                    double val = -c34 * drp * dsq * I3 + drs * dpq * I3 + dps * drq * I3
                                 + c34 * drp * 3.0 * T2_5[s * 3 + q] - drs * 3.0 * T2_5[p * 3 + q]
                                 - dps * 3.0 * T2_5[r * 3 + q] - 3.0 * drq * T2_5[p * 3 + s]
                                 - 3.0 * dpq * T2_5[r * 3 + s] - 3.0 * dsq * T2_5[r * 3 + p]
                                 + 15.0 * T4_7[r * 27 + p * 9 + s * 3 + q]
                                 + cblob * drp * (-3.0 * dsq * I5 + 15.0 * T2_7[s * 3 + q]);
                    ID2G[r * 27 + p * 9 + s * 3 + q] = C1 * val;
                }
            }
        }
    for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) {
            double tr = 0.0;
            for (int p = 0; p < 3; ++p) tr += ID2G[r * 27 + p * 9 + s * 3 + p];
            for (int k = 0; k < 3; ++k) {
                double nq = 0.0, npd = 0.0;
                for (int q = 0; q < 3; ++q) nq += nrm[q] * ID2G[r * 27 + k * 9 + s * 3 + q];
                for (int p = 0; p < 3; ++p) npd += nrm[p] * ID2G[r * 27 + p * 9 + s * 3 + k];
                B[r * 9 + s * 3 + k] = lam * nrm[k] * tr + mu * nq + mu * npd;
            }
        }
    for (int k = 0; k < 3; ++k) {
        double trB = B[0 * 9 + 0 * 3 + k] + B[1 * 9 + 1 * 3 + k] + B[2 * 9 + 2 * 3 + k];
        for (int m = 0; m < 3; ++m)
            for (int n = 0; n < 3; ++n) {
                double dmn = (m == n) ? 1.0 : 0.0;
                H[m * 9 + n * 3 + k] = -(lam * dmn * trB + mu * (B[m * 9 + n * 3 + k] + B[n * 9 + m * 3 + k]));
            }
    }
}

/* exact finite-triangle eigenstress H*[m,n,k]: tri_kernels.py:1194-1214 */
static void eigenstress_pair(const double *tv, const double *ex, const double *ey,
                             const double *nhat, const double *nrm, const double *obs, double eps,
                             double mu, double lam, double *H)
{
    if (eps <= 0.0) {
        for (int q = 0; q < 27; ++q) H[q] = 0.0;
        return;
    }
    double M3[9] = {0}, M5[16] = {0}, M7[25] = {0};
    double I3, I5, I7, z;
    stress_2d_moments(tv, ex, ey, nhat, obs, eps, M3, M5, M7, 1, &I3, &I5, &I7, &z);
    double phi = (15.0 * eps * eps * eps * eps / (8.0 * M_PI)) * I7;
    for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
            double dmn = (m == n) ? 1.0 : 0.0;
            for (int k = 0; k < 3; ++k) {
                double dmk = (m == k) ? 1.0 : 0.0;
                double dnk = (n == k) ? 1.0 : 0.0;
                H[m * 9 + n * 3 + k] = phi * (lam * dmn * nrm[k] + mu * (dmk * nrm[n] + dnk * nrm[m]));
            }
        }
}

/* =================== exported drivers ================================= */

/* per-source frames ex, ey, nhat (3 each) and OK flag: tri_kernels.py:1052-1065 */
void orc_frames(const double *tri_verts, int64_t ns, double *frames, int32_t *ok)
{
    #pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < ns; ++s) {
        double *f = frames + 9 * s;
        ok[s] = tri_frame(tri_verts + 9 * s, f, f + 3, f + 6);
        if (!ok[s]) for (int q = 0; q < 9; ++q) f[q] = 0.0;
    }
}

/* u_basis_matrices: tri_kernels.py:451-469, out (3, 3nf, 3ns) */
void orc_u_basis_matrices(const double *x_field, int64_t nf, const double *tri_verts, int64_t ns,
                          const double *eps_arr, double *out)
{
    int64_t ncol = 3 * ns, nrow = 3 * nf;
    memset(out, 0, sizeof(double) * 3 * nrow * ncol);
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t s = 0; s < ns; ++s) {
        double ex[3], ey[3], nh[3], blk[27];
        const double *tv = tri_verts + 9 * s;
        if (!tri_frame(tv, ex, ey, nh)) continue;
        for (int64_t f = 0; f < nf; ++f) {
            u_basis_pair(tv, ex, ey, nh, x_field + 3 * f, eps_arr[s], blk);
            for (int b = 0; b < 3; ++b)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        out[(int64_t)b * nrow * ncol + (3 * f + i) * ncol + 3 * s + j] = blk[b * 9 + i * 3 + j];
        }
    }
}

/* t_basis_matrices: tri_kernels.py:472-492, out (6, 3nf, 3ns) */
void orc_t_basis_matrices(const double *x_field, int64_t nf, const double *tri_verts,
                          const double *normals, int64_t ns, const double *eps_arr, double *out)
{
    int64_t ncol = 3 * ns, nrow = 3 * nf;
    memset(out, 0, sizeof(double) * 6 * nrow * ncol);
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t s = 0; s < ns; ++s) {
        double ex[3], ey[3], nh[3], blk[54];
        const double *tv = tri_verts + 9 * s;
        if (!tri_frame(tv, ex, ey, nh)) continue;
        for (int64_t f = 0; f < nf; ++f) {
            t_basis_pair(tv, ex, ey, nh, normals + 3 * s, x_field + 3 * f, eps_arr[s], blk);
            for (int b = 0; b < 6; ++b)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        out[(int64_t)b * nrow * ncol + (3 * f + i) * ncol + 3 * s + j] = blk[b * 9 + i * 3 + j];
        }
    }
}

/* u_matrix_direct: tri_kernels.py:546-565, out (3nf, 3ns) */
void orc_u_matrix_direct(const double *x_field, int64_t nf, const double *tri_verts, int64_t ns,
                         const double *eps_arr, double g1, double g2, double g3, double *out)
{
    int64_t ncol = 3 * ns;
    memset(out, 0, sizeof(double) * 3 * nf * ncol);
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t s = 0; s < ns; ++s) {
        double ex[3], ey[3], nh[3], blk[27];
        const double *tv = tri_verts + 9 * s;
        if (!tri_frame(tv, ex, ey, nh)) continue;
        for (int64_t f = 0; f < nf; ++f) {
            u_basis_pair(tv, ex, ey, nh, x_field + 3 * f, eps_arr[s], blk);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    out[(3 * f + i) * ncol + 3 * s + j] = g1 * blk[0 * 9 + i * 3 + j]
                                                          + g2 * blk[1 * 9 + i * 3 + j]
                                                          + g3 * blk[2 * 9 + i * 3 + j];
        }
    }
}

/* t_matrix_direct: tri_kernels.py:568-591, out (3nf, 3ns); c = c1..c6 */
void orc_t_matrix_direct(const double *x_field, int64_t nf, const double *tri_verts,
                         const double *normals, int64_t ns, const double *eps_arr,
                         const double *c, double *out)
{
    int64_t ncol = 3 * ns;
    memset(out, 0, sizeof(double) * 3 * nf * ncol);
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t s = 0; s < ns; ++s) {
        double ex[3], ey[3], nh[3], blk[54];
        const double *tv = tri_verts + 9 * s;
        if (!tri_frame(tv, ex, ey, nh)) continue;
        for (int64_t f = 0; f < nf; ++f) {
            t_basis_pair(tv, ex, ey, nh, normals + 3 * s, x_field + 3 * f, eps_arr[s], blk);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    out[(3 * f + i) * ncol + 3 * s + j] =
                        c[0] * blk[0 * 9 + i * 3 + j] + c[1] * blk[1 * 9 + i * 3 + j]
                        + c[2] * blk[2 * 9 + i * 3 + j] + c[3] * blk[3 * 9 + i * 3 + j]
                        + c[4] * blk[4 * 9 + i * 3 + j] + c[5] * blk[5 * 9 + i * 3 + j];
        }
    }
}

/* u_disp_contract: tri_kernels.py:976-1009, out (nf, 3) */
void orc_u_disp_contract(const double *x_field, int64_t nf, const double *tri_verts, int64_t ns,
                         const double *eps_arr, const double *density, double g1, double g2,
                         double g3, double *out)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t f = 0; f < nf; ++f) {
        double blk[27], ex[3], ey[3], nh[3];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int64_t s = 0; s < ns; ++s) {
            const double *tv = tri_verts + 9 * s;
            if (!tri_frame(tv, ex, ey, nh)) continue;
            const double *d = density + 3 * s;
            if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
            u_basis_pair(tv, ex, ey, nh, x_field + 3 * f, eps_arr[s], blk);
            for (int j = 0; j < 3; ++j) {
                double dj = d[j];
                if (dj == 0.0) continue;
                a0 += (g1 * blk[0 * 9 + 0 * 3 + j] + g2 * blk[1 * 9 + 0 * 3 + j] + g3 * blk[2 * 9 + 0 * 3 + j]) * dj;
                a1 += (g1 * blk[0 * 9 + 1 * 3 + j] + g2 * blk[1 * 9 + 1 * 3 + j] + g3 * blk[2 * 9 + 1 * 3 + j]) * dj;
                a2 += (g1 * blk[0 * 9 + 2 * 3 + j] + g2 * blk[1 * 9 + 2 * 3 + j] + g3 * blk[2 * 9 + 2 * 3 + j]) * dj;
            }
        }
        out[3 * f] = a0; out[3 * f + 1] = a1; out[3 * f + 2] = a2;
    }
}

/* t_disp_contract: tri_kernels.py:1012-1049, out (nf, 3) */
void orc_t_disp_contract(const double *x_field, int64_t nf, const double *tri_verts,
                         const double *normals, int64_t ns, const double *eps_arr,
                         const double *density, const double *c, double *out)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t f = 0; f < nf; ++f) {
        double blk[54], ex[3], ey[3], nh[3];
        double acc[3] = {0.0, 0.0, 0.0};
        for (int64_t s = 0; s < ns; ++s) {
            const double *tv = tri_verts + 9 * s;
            if (!tri_frame(tv, ex, ey, nh)) continue;
            const double *d = density + 3 * s;
            if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
            t_basis_pair(tv, ex, ey, nh, normals + 3 * s, x_field + 3 * f, eps_arr[s], blk);
            for (int j = 0; j < 3; ++j) {
                double dj = d[j];
                if (dj == 0.0) continue;
                for (int i = 0; i < 3; ++i)
                    acc[i] += (c[0] * blk[0 * 9 + i * 3 + j] + c[1] * blk[1 * 9 + i * 3 + j]
                               + c[2] * blk[2 * 9 + i * 3 + j] + c[3] * blk[3 * 9 + i * 3 + j]
                               + c[4] * blk[4 * 9 + i * 3 + j] + c[5] * blk[5 * 9 + i * 3 + j]) * dj;
            }
        }
        out[3 * f] = acc[0]; out[3 * f + 1] = acc[1]; out[3 * f + 2] = acc[2];
    }
}

/* accumulate one 3x3x3 pair tensor T[m,n,k] contracted with the density
 * exactly as tri_kernels.py:1096-1104 does */
static inline void acc_stress(double *a, const double *T, const double *d)
{
    for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n)
            a[m * 3 + n] += T[m * 9 + n * 3 + 0] * d[0] + T[m * 9 + n * 3 + 1] * d[1]
                            + T[m * 9 + n * 3 + 2] * d[2];
}

/* kelvin_stress_contract: tri_kernels.py:1068-1108, out (nf, 3, 3) */
void orc_kelvin_stress_contract(const double *x_field, int64_t nf, const double *tri_verts,
                                int64_t ns, const double *eps_arr, const double *density,
                                double mu, double lam, double *out)
{
    double nu = lam / (2.0 * (lam + mu));
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t f = 0; f < nf; ++f) {
        double S[27], ex[3], ey[3], nh[3], a[9] = {0};
        for (int64_t s = 0; s < ns; ++s) {
            const double *tv = tri_verts + 9 * s;
            if (!tri_frame(tv, ex, ey, nh)) continue;
            const double *d = density + 3 * s;
            if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
            kelvin_stress_pair(tv, ex, ey, nh, x_field + 3 * f, eps_arr[s], mu, nu, lam, S);
            acc_stress(a, S, d);
        }
        for (int q = 0; q < 9; ++q) out[9 * f + q] = a[q];
    }
}

/* dd_stress_contract: tri_kernels.py:1111-1154, out (nf, 3, 3) */
void orc_dd_stress_contract(const double *x_field, int64_t nf, const double *tri_verts,
                            const double *normals, int64_t ns, const double *eps_arr,
                            const double *density, double mu, double lam, double *out)
{
    double nu = lam / (2.0 * (lam + mu));
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t f = 0; f < nf; ++f) {
        double H[27], ex[3], ey[3], nh[3], a[9] = {0};
        for (int64_t s = 0; s < ns; ++s) {
            const double *tv = tri_verts + 9 * s;
            if (!tri_frame(tv, ex, ey, nh)) continue;
            const double *d = density + 3 * s;
            if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
            dd_stress_pair(tv, ex, ey, nh, normals + 3 * s, x_field + 3 * f, eps_arr[s], mu, nu, lam, H);
            acc_stress(a, H, d);
        }
        for (int q = 0; q < 9; ++q) out[9 * f + q] = a[q];
    }
}

/* eigenstress_contract: tri_kernels.py:1217-1260, out (nf, 3, 3) */
void orc_eigenstress_contract(const double *x_field, int64_t nf, const double *tri_verts,
                              const double *normals, int64_t ns, const double *eps_arr,
                              const double *density, double mu, double lam, double *out)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t f = 0; f < nf; ++f) {
        double H[27], ex[3], ey[3], nh[3], a[9] = {0};
        for (int64_t s = 0; s < ns; ++s) {
            const double *tv = tri_verts + 9 * s;
            if (!tri_frame(tv, ex, ey, nh)) continue;
            const double *d = density + 3 * s;
            if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
            eigenstress_pair(tv, ex, ey, nh, normals + 3 * s, x_field + 3 * f, eps_arr[s], mu, lam, H);
            acc_stress(a, H, d);
        }
        for (int q = 0; q < 9; ++q) out[9 * f + q] = a[q];
    }
}

/* moment probe for tests: I3, I5, I7 and z of one pair (tri_kernels.py:655-769) */
void orc_moments(const double *tv, const double *obs, double eps, double *out4)
{
    double ex[3], ey[3], nh[3], M3[9] = {0}, M5[16] = {0}, M7[25] = {0};
    if (!tri_frame(tv, ex, ey, nh)) { out4[0] = out4[1] = out4[2] = out4[3] = 0.0; return; }
    stress_2d_moments(tv, ex, ey, nh, obs, eps, M3, M5, M7, 1, out4, out4 + 1, out4 + 2, out4 + 3);
}
