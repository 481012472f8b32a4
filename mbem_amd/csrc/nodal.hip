// Lagrange P0/P1/P2 nodal pair kernels on gfx950: the nodal density's
// displacement matrices and basis stacks, its displacement and stress
// contractions, and the traction rows of a field point set (the
// FAULT_TRACTION rows), for src/mbem/kernels/tri_nodal.py.
//
// Per (field point, source triangle) pair the weighted moment tables
//     W_k[n][a][b] = int_T N_k xi_1^a xi_2^b / R_eps^n dS,   n = 1, 3, 5, 7,
// come from the divergence-theorem recursion over cancellation-free edge
// primitives (closed form, small-|u| and large-|u| series) within D* L of the
// element's centroid, and from a collapsed product Gauss rule beyond
// (12 x 12 points up to 40 L, 8 x 8 past it); each node's kernel is then the
// P0 contraction of W_k.  The steps and their branch points are the
// reference's (tri_nodal.py:262-709): they decide which terms cancel and so
// the digits the result keeps.  This file is built without FP contraction,
// so those cancellation-sensitive differences round as written.
//
// Mappings.  Matrices: one wave per field point, lane = source, each lane
// writing its 3 K contiguous columns of the point's three rows (row-major,
// leading dimension ld).  Contractions and traction rows: lane = field
// point, the source loop uniform across the wave (a source's nodal frame is
// read by every lane at once), each lane owning its rows, so the traction
// rows' column scatter to shared nodes needs no atomics.
//
// The per-pair tables live in each lane's private memory (about 1,300
// doubles at order 2); nothing here is on the dense path's critical loop.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/mbem_hip.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace {

constexpr int kMaxEdge = 5;      // u^k edge primitives, k <= 5
constexpr int kMaxTable = 6;     // monomial degree of the moment table
constexpr int kGeo = 30;         // nodal frame record (29) + ok flag
constexpr int kSer = 80;         // series chains: 2 n_terms + 2 + kMaxEdge + 3 <= 80

struct NSpec {
    int order, K, q_split, n_terms, nq;
    int degn[5];      // table degree, n = -1, 1, 3, 5, 7 (or -1)
    int ekmax[4];     // edge-primitive degree, m = -1, 1, 3, 5 (or -1)
    int need0[4];     // degree the kernels read, n = 1, 3, 5, 7 (or -1)
    double dstar, dstar_far, small_u, series_u;
    const double* binom;      // 4 x n_terms
    const double* lattice;    // K x 3 (exponents as doubles)
    const double* q_lam;      // nq x 3
    const double* q_w;        // nq
    const double* q_N;        // nq x K
};

__device__ double ipow_n(double x, int n)
{
    double out = 1.0;
    if (n >= 0) {
        for (int i = 0; i < n; ++i) out = out * x;
    } else {
        const double xi = 1.0 / x;
        for (int i = 0; i < -n; ++i) out = out * xi;
    }
    return out;
}

__device__ __forceinline__ double binom_c(int a, int i)
{
    double c = 1.0;
    for (int t = 0; t < i; ++t) c = c * (a - t) / (t + 1);
    return c;
}

// large-|u| series, 0 < ua < ub, rho^2 / ua^2 <= 1/4 (tri_nodal.py:262-300):
// sum_j C(-m/2, j) x^j [t^q / q]_1^r with r^q - 1 from one-sign chains
__device__ void series_large(int m, int kmax, double ua, double ub, double rho2, const double* bm,
                             int nt, double* Pm)
{
    double T[kSer];
    const double x = rho2 / (ua * ua);
    const double r = ub / ua;
    const double log_r = log1p((ub - ua) / ua);
    const double up = (ub - ua) / ua;
    const double dn = -(ub - ua) / ub;
    const double rinv = 1.0 / r;
    const int off = 2 * nt + 2;
    T[off] = 0.0;
    const int qmax = kmax - m + 1;
    for (int q = 1; q <= qmax; ++q) T[off + q] = T[off + q - 1] * r + up;
    const int qmin = 1 - m - 2 * (nt - 1);
    for (int q = -1; q >= qmin; --q) T[off + q] = T[off + q + 1] * rinv + dn;
    double base = ipow_n(ua, 1 - m);
    for (int k = 0; k <= kmax; ++k) {
        double out = 0.0, xj = 1.0;
        for (int j = 0; j < nt; ++j) {
            const int q = k - m + 1 - 2 * j;
            const double term = (q == 0) ? log_r : T[off + q] / q;
            out = out + bm[j] * xj * term;
            xj = xj * x;
        }
        Pm[k] = base * out;
        base = base * ua;
    }
}

// small-|u| series, max |u| <= rho / 2 (tri_nodal.py:303-343)
__device__ void series_small(int m, int kmax, double ua, double ub, double rho2, const double* bm,
                             int nt, double* Pm)
{
    double YA[kSer], YB[kSer], E[kSer];
    const double rho = sqrt(rho2);
    const double ya = ua / rho, yb = ub / rho;
    const int pmax = kmax + 2 * (nt - 1) + 1;
    YA[0] = 1.0;
    YB[0] = 1.0;
    for (int p = 1; p <= pmax; ++p) {
        YA[p] = YA[p - 1] * ya;
        YB[p] = YB[p - 1] * yb;
    }
    bool close = false;
    if (ua * ub > 0.0) {
        const double ratio = ub / ua;
        close = (ratio < 2.0) && (ratio > 0.5);
        if (close) {
            const double up = (ub - ua) / ua;
            E[0] = 0.0;
            for (int p = 1; p <= pmax; ++p) E[p] = E[p - 1] * ratio + up;
        }
    }
    double scale = ipow_n(rho, 1 - m);
    for (int k = 0; k <= kmax; ++k) {
        double out = 0.0;
        for (int j = 0; j < nt; ++j) {
            const int pw = k + 2 * j + 1;
            const double dpow = close ? YA[pw] * E[pw] : YB[pw] - YA[pw];
            out = out + bm[j] * dpow / pw;
        }
        Pm[k] = scale * out;
        scale = scale * rho;
    }
}

// closed-form differences with their conjugate forms (tri_nodal.py:346-383)
__device__ void closed_form(double ua, double ub, double rho2, const int* ekmax,
                            double P[4][kMaxEdge + 1])
{
    const double Ra = sqrt(ua * ua + rho2);
    const double Rb = sqrt(ub * ub + rho2);
    const double dR = (ub - ua) * (ub + ua) / (Ra + Rb);
    double dJ1;
    if (ua >= 0.0)
        dJ1 = log1p(((ub - ua) + dR) / (ua + Ra));
    else if (ub <= 0.0)
        dJ1 = log1p(((ub - ua) - dR) / (Rb - ub));
    else
        dJ1 = log(((ub + Rb) * (Ra - ua)) / rho2);
    double dJ3;
    if (ua * ub > 0.0)
        dJ3 = ((ub - ua) * (ub + ua)) / (Ra * Rb * (ub * Ra + ua * Rb));
    else
        dJ3 = (ub / Rb - ua / Ra) / rho2;
    const double duRm = ub * pow(Rb, -3.0) - ua * pow(Ra, -3.0);
    const double dJ5 = (duRm + 2.0 * dJ3) / (3.0 * rho2);
    const double duR = 0.5 * (ub - ua) * (Ra + Rb) + 0.5 * (ua + ub) * dR;
    const double dJm1 = 0.5 * (duR + rho2 * dJ1);
    const double RaRb = Ra * Rb;
    const double S3 = (Rb * Rb + Ra * Rb) + Ra * Ra;
    const double dK[4] = {dR * S3 / 3.0, dR, dR / RaRb, dR * S3 / (3.0 * pow(RaRb, 3.0))};
    const double dJ[4] = {dJm1, dJ1, dJ3, dJ5};
    for (int mi = 0; mi < 4; ++mi) {
        if (ekmax[mi] < 0) continue;
        P[mi][0] = dJ[mi];
        if (ekmax[mi] >= 1) P[mi][1] = dK[mi];
    }
    for (int k = 2; k <= kMaxEdge; ++k)
        for (int mi = 1; mi < 4; ++mi)
            if (ekmax[mi] >= k) P[mi][k] = P[mi - 1][k - 2] - rho2 * P[mi][k - 2];
}

__device__ void edge_primitives(double ua, double ub, double rho2, const NSpec& sp,
                                double P[4][kMaxEdge + 1])
{
    const double rho = sqrt(rho2);
    const double aua = fabs(ua), aub = fabs(ub);
    const double umax = fmax(aua, aub), umin = fmin(aua, aub);
    const int nt = sp.n_terms;
    if (umax <= sp.small_u * rho) {
        for (int mi = 0; mi < 4; ++mi)
            if (sp.ekmax[mi] >= 0)
                series_small(2 * mi - 1, sp.ekmax[mi], ua, ub, rho2, sp.binom + mi * nt, nt, P[mi]);
    } else if ((ua * ub > 0.0) && umin >= sp.series_u * rho) {
        const bool negative = ua < 0.0;
        const double a = negative ? -ub : ua, b = negative ? -ua : ub;
        for (int mi = 0; mi < 4; ++mi)
            if (sp.ekmax[mi] >= 0) {
                series_large(2 * mi - 1, sp.ekmax[mi], a, b, rho2, sp.binom + mi * nt, nt, P[mi]);
                if (negative)
                    for (int k = 1; k <= sp.ekmax[mi]; k += 2) P[mi][k] = -P[mi][k];
            }
    } else {
        closed_form(ua, ub, rho2, sp.ekmax, P);
    }
}

__device__ double solid_angle_v(const double* v, double ox, double oy, double oz)
{
    const double r1x = v[0] - ox, r1y = v[1] - oy, r1z = v[2] - oz;
    const double r2x = v[3] - ox, r2y = v[4] - oy, r2z = v[5] - oz;
    const double r3x = v[6] - ox, r3y = v[7] - oy, r3z = v[8] - oz;
    const double R1 = sqrt(r1x * r1x + r1y * r1y + r1z * r1z);
    const double R2 = sqrt(r2x * r2x + r2y * r2y + r2z * r2z);
    const double R3 = sqrt(r3x * r3x + r3y * r3y + r3z * r3z);
    const double cx = r2y * r3z - r2z * r3y;
    const double cy = r2z * r3x - r2x * r3z;
    const double cz = r2x * r3y - r2y * r3x;
    const double numer = r1x * cx + r1y * cy + r1z * cz;
    const double d12 = r1x * r2x + r1y * r2y + r1z * r2z;
    const double d23 = r2x * r3x + r2y * r3y + r2z * r3z;
    const double d13 = r1x * r3x + r1y * r3y + r1z * r3z;
    const double denom = R1 * R2 * R3 + R3 * d12 + R1 * d23 + R2 * d13;
    return 2.0 * atan2(numer, denom);
}

// nodal frame: ex ey nhat centroid p2d(3 x 2) abc(3 x 3) L area2 ok
// (tri_nodal.py:418-456; the frame itself as tri_kernels.py:136-153)
__global__ __launch_bounds__(256) void k_nodal_frames(const double* __restrict__ tri_verts, long ns,
                                                      double* __restrict__ geo_all)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const double* tv = tri_verts + 9 * s;
    double* geo = geo_all + kGeo * s;
    for (int q = 0; q < kGeo; ++q) geo[q] = 0.0;
    const double e1x = tv[3] - tv[0], e1y = tv[4] - tv[1], e1z = tv[5] - tv[2];
    const double e2x = tv[6] - tv[0], e2y = tv[7] - tv[1], e2z = tv[8] - tv[2];
    const double nx = e1y * e2z - e1z * e2y;
    const double ny = e1z * e2x - e1x * e2z;
    const double nz = e1x * e2y - e1y * e2x;
    const double area2 = sqrt(nx * nx + ny * ny + nz * nz);
    if (area2 < 1e-30) return;
    double ex[3], ey[3], nh[3];
    nh[0] = nx / area2; nh[1] = ny / area2; nh[2] = nz / area2;
    const double e1n = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
    ex[0] = e1x / e1n; ex[1] = e1y / e1n; ex[2] = e1z / e1n;
    ey[0] = nh[1] * ex[2] - nh[2] * ex[1];
    ey[1] = nh[2] * ex[0] - nh[0] * ex[2];
    ey[2] = nh[0] * ex[1] - nh[1] * ex[0];
    for (int c = 0; c < 3; ++c) {
        geo[c] = ex[c];
        geo[3 + c] = ey[c];
        geo[6 + c] = nh[c];
        geo[9 + c] = ((tv[c] + tv[3 + c]) + tv[6 + c]) / 3.0;
    }
    for (int k = 0; k < 3; ++k) {
        const double dx = tv[3 * k] - geo[9], dy = tv[3 * k + 1] - geo[10], dz = tv[3 * k + 2] - geo[11];
        geo[12 + 2 * k] = dx * ex[0] + dy * ex[1] + dz * ex[2];
        geo[13 + 2 * k] = dx * ey[0] + dy * ey[1] + dz * ey[2];
    }
    double L = 0.0;
    for (int k = 0; k < 3; ++k) {
        const int k2 = (k + 1) % 3;
        const double dx = tv[3 * k2] - tv[3 * k], dy = tv[3 * k2 + 1] - tv[3 * k + 1];
        const double dz = tv[3 * k2 + 2] - tv[3 * k + 2];
        L = fmax(L, sqrt(dx * dx + dy * dy + dz * dz));
    }
    geo[27] = L;
    geo[28] = area2;
    const double den = ((geo[14] - geo[12]) * (geo[17] - geo[13]) - (geo[15] - geo[13]) * (geo[16] - geo[12]));
    for (int k = 0; k < 3; ++k) {
        const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        const double x1 = geo[12 + 2 * k1], y1 = geo[13 + 2 * k1];
        const double x2 = geo[12 + 2 * k2], y2 = geo[13 + 2 * k2];
        geo[18 + 3 * k] = (x1 * y2 - y1 * x2) / den;
        geo[19 + 3 * k] = (y1 - y2) / den;
        geo[20 + 3 * k] = (x2 - x1) / den;
    }
    geo[29] = 1.0;
}

// C[k][a][b]: N_k as a polynomial in (xi_1, xi_2) about the foot of the field
// point, from the barycentric product formula (tri_nodal.py:473-516)
template <int K>
__device__ void shape_coeffs(const NSpec& sp, const double* geo, double X1, double X2,
                             double C[K][3][3])
{
    const int order = sp.order, D = order + 1;
    for (int k = 0; k < K; ++k)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) C[k][a][b] = 0.0;
    if (order == 0) {
        C[0][0][0] = 1.0;
        return;
    }
    for (int k = 0; k < K; ++k) {
        double poly[3][3] = {{0.0}}, fac[3][3] = {{0.0}}, out[3][3];
        poly[0][0] = 1.0;
        for (int i = 0; i < 3; ++i) {
            const double l0 = (geo[18 + 3 * i] + X1 * geo[19 + 3 * i]) + X2 * geo[20 + 3 * i];
            const int pw = (int)sp.lattice[3 * k + i];
            for (int j = 0; j < pw; ++j) {
                fac[0][0] = (order * l0) / (j + 1) - (double)j / (j + 1);
                fac[1][0] = (order * geo[19 + 3 * i]) / (j + 1);
                fac[0][1] = (order * geo[20 + 3 * i]) / (j + 1);
                for (int a = 0; a < D; ++a)
                    for (int b = 0; b < D; ++b) out[a][b] = 0.0;
                for (int a = 0; a < D; ++a)
                    for (int b = 0; b < D - a; ++b) {
                        const double pab = poly[a][b];
                        if (pab == 0.0) continue;
                        for (int c = 0; c < D - a; ++c)
                            for (int d = 0; d < D - a - b - c; ++d) out[a + c][b + d] += pab * fac[c][d];
                    }
                for (int a = 0; a < D; ++a)
                    for (int b = 0; b < D; ++b) poly[a][b] = out[a][b];
            }
        }
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) C[k][a][b] = poly[a][b];
    }
}

// M[ni][a][b] = int_T xi_1^a xi_2^b / R^n, n = 2 ni - 1 (tri_nodal.py:519-620)
__device__ void moment_table(const double* tv, const double* geo, double X1, double X2, double h2,
                             const NSpec& sp, double M[5][kMaxTable + 1][kMaxTable + 1])
{
    int kall = 0;
    for (int mi = 0; mi < 4; ++mi) kall = max(kall, sp.ekmax[mi]);
    const double h = sqrt(h2);
    const double ox = ((geo[9] + X1 * geo[0]) + X2 * geo[3]) + h * geo[6];
    const double oy = ((geo[10] + X1 * geo[1]) + X2 * geo[4]) + h * geo[7];
    const double oz = ((geo[11] + X1 * geo[2]) + X2 * geo[5]) + h * geo[8];
    const double I3 = -solid_angle_v(tv, ox, oy, oz) / h;
    double BD[4] = {0.0, 0.0, 0.0, 0.0};
    double BN1[4][kMaxEdge + 1][kMaxEdge + 1], BN2[4][kMaxEdge + 1][kMaxEdge + 1];
    double P[4][kMaxEdge + 1];
    for (int mi = 0; mi < 4; ++mi)
        for (int a = 0; a <= kMaxEdge; ++a) {
            P[mi][a] = 0.0;
            for (int b = 0; b <= kMaxEdge; ++b) {
                BN1[mi][a][b] = 0.0;
                BN2[mi][a][b] = 0.0;
            }
        }
    double dpP[kMaxEdge + 1], c1P[kMaxEdge + 1], s1P[kMaxEdge + 1], c2P[kMaxEdge + 1], s2P[kMaxEdge + 1];
    dpP[0] = 1.0; c1P[0] = 1.0; s1P[0] = 1.0; c2P[0] = 1.0; s2P[0] = 1.0;
    for (int e = 0; e < 3; ++e) {
        const int e2 = (e + 1) % 3;
        const double pax = geo[12 + 2 * e] - X1, pay = geo[13 + 2 * e] - X2;
        const double pbx = geo[12 + 2 * e2] - X1, pby = geo[13 + 2 * e2] - X2;
        const double evx = pbx - pax, evy = pby - pay;
        const double Le = sqrt(evx * evx + evy * evy);
        const double tx = evx / Le, ty = evy / Le;
        const double c1 = ty, c2 = -tx;
        const double dp = pax * c1 + pay * c2;
        const double ua = pax * tx + pay * ty;
        const double ub = pbx * tx + pby * ty;
        const double rho2 = dp * dp + h2;
        edge_primitives(ua, ub, rho2, sp, P);
        for (int q = 1; q <= kall; ++q) {
            dpP[q] = dpP[q - 1] * dp;
            c1P[q] = c1P[q - 1] * c1;
            s1P[q] = s1P[q - 1] * tx;
            c2P[q] = c2P[q - 1] * c2;
            s2P[q] = s2P[q - 1] * ty;
        }
        for (int mi = 0; mi < 4; ++mi) {
            const int kmax = sp.ekmax[mi];
            if (kmax < 0) continue;
            BD[mi] = BD[mi] + dp * P[mi][0];
            for (int a = 0; a <= kmax; ++a)
                for (int b = 0; b <= kmax - a; ++b) {
                    double val = 0.0;
                    for (int i = 0; i <= a; ++i)
                        for (int j = 0; j <= b; ++j) {
                            const int pw = a + b - i - j;
                            double coeff = (binom_c(a, i) * binom_c(b, j)) * dpP[pw];
                            coeff = coeff * c1P[a - i];
                            coeff = coeff * s1P[i];
                            coeff = coeff * c2P[b - j];
                            coeff = coeff * s2P[j];
                            val = val + coeff * P[mi][i + j];
                        }
                    BN1[mi][a][b] = BN1[mi][a][b] + c1 * val;
                    BN2[mi][a][b] = BN2[mi][a][b] + c2 * val;
                }
        }
    }
    const int* degn = sp.degn;
    M[2][0][0] = I3;
    if (degn[3] >= 0) M[3][0][0] = (BD[2] + I3) / (3.0 * h2);
    if (degn[4] >= 0) M[4][0][0] = (BD[3] + 3.0 * M[3][0][0]) / (5.0 * h2);
    if (degn[1] >= 0) M[1][0][0] = (BD[1] - (1.0 * h2) * M[2][0][0]) / 1.0;
    if (degn[0] >= 0) M[0][0][0] = (BD[0] - (-1.0 * h2) * M[1][0][0]) / 3.0;
    for (int ni = 1; ni < 5; ++ni) {
        const int d = degn[ni];
        if (d < 1) continue;
        const int n = 2 * ni - 1, mi = ni - 1;
        for (int q = 1; q <= d; ++q)
            for (int a = q; a >= 0; --a) {
                const int b = q - a;
                double val;
                if (a >= 1) {
                    val = -BN1[mi][a - 1][b];
                    if (a >= 2) val = val + (a - 1) * M[mi][a - 2][b];
                } else {
                    val = -BN2[mi][0][b - 1];
                    if (b >= 2) val = val + (b - 1) * M[mi][0][b - 2];
                }
                M[ni][a][b] = val / (n - 2);
            }
    }
}

// the weighted tables W[n][k][a][b] of one pair; returns the signed height
// (tri_nodal.py:623-709)
template <int K>
__device__ double pair_tables(const double* tv, const double* geo, double ox, double oy, double oz,
                              double eps, const NSpec& sp, double W[4][K][5][5])
{
    const double dx = ox - geo[9], dy = oy - geo[10], dz = oz - geo[11];
    const double z = dx * geo[6] + dy * geo[7] + dz * geo[8];
    const double X1 = dx * geo[0] + dy * geo[1] + dz * geo[2];
    const double X2 = dx * geo[3] + dy * geo[4] + dz * geo[5];
    const double h2 = z * z + eps * eps;
    const double D = sqrt((dx * dx + dy * dy + dz * dz) + eps * eps);
    const double L = geo[27];
    if (D <= sp.dstar * L) {
        double C[K][3][3];
        double M[5][kMaxTable + 1][kMaxTable + 1];
        shape_coeffs<K>(sp, geo, X1, X2, C);
        moment_table(tv, geo, X1, X2, h2, sp, M);
        for (int ni2 = 0; ni2 < 4; ++ni2) {
            const int d = sp.need0[ni2];
            if (d < 0) continue;
            const int ni = ni2 + 1;
            for (int k = 0; k < K; ++k)
                for (int a = 0; a <= d; ++a)
                    for (int b = 0; b <= d - a; ++b) {
                        double acc = 0.0;
                        for (int ap = 0; ap <= sp.order; ++ap)
                            for (int bp = 0; bp <= sp.order - ap; ++bp) acc += C[k][ap][bp] * M[ni][a + ap][b + bp];
                        W[ni2][k][a][b] = acc;
                    }
        }
        return z;
    }
    int q0 = 0, q1 = sp.q_split;
    if (D > sp.dstar_far * L) {
        q0 = sp.q_split;
        q1 = sp.nq;
    }
    const double area2 = geo[28];
    int dmax = 0;
    for (int ni2 = 0; ni2 < 4; ++ni2) {
        const int d = sp.need0[ni2];
        dmax = max(dmax, d);
        if (d < 0) continue;
        for (int k = 0; k < K; ++k)
            for (int a = 0; a <= d; ++a)
                for (int b = 0; b <= d - a; ++b) W[ni2][k][a][b] = 0.0;
    }
    double x1p[kMaxEdge + 1], x2p[kMaxEdge + 1];
    x1p[0] = 1.0;
    x2p[0] = 1.0;
    for (int q = q0; q < q1; ++q) {
        const double l1 = sp.q_lam[3 * q], l2 = sp.q_lam[3 * q + 1], l3 = sp.q_lam[3 * q + 2];
        const double eta1 = l1 * geo[12] + l2 * geo[14] + l3 * geo[16];
        const double eta2 = l1 * geo[13] + l2 * geo[15] + l3 * geo[17];
        const double xi1 = eta1 - X1, xi2 = eta2 - X2;
        const double R2 = xi1 * xi1 + xi2 * xi2 + h2;
        const double wq = sp.q_w[q] * area2;
        const double r2i = 1.0 / R2;
        double Rn = sqrt(r2i);
        for (int a = 1; a <= dmax; ++a) {
            x1p[a] = x1p[a - 1] * xi1;
            x2p[a] = x2p[a - 1] * xi2;
        }
        for (int ni2 = 0; ni2 < 4; ++ni2) {
            const int d = sp.need0[ni2];
            if (d >= 0)
                for (int a = 0; a <= d; ++a)
                    for (int b = 0; b <= d - a; ++b) {
                        const double f = wq * x1p[a] * x2p[b] * Rn;
                        for (int k = 0; k < K; ++k) W[ni2][k][a][b] += f * sp.q_N[q * K + k];
                    }
            Rn = Rn * r2i;
        }
    }
    return z;
}

// ---- the P0 tensor builds on one node's table (tri_kernels.py:787-853) -----
// 2D moment a, b of W (5 x 5), M[a][b]
__device__ __forceinline__ void build_T2(const double* M, const double* bas, double z, double* out)
{
    const double* ex = bas;
    const double* ey = bas + 3;
    const double* nh = bas + 6;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            out[i * 3 + j] = ex[i] * ex[j] * M[2 * 5 + 0] + (ex[i] * ey[j] + ey[i] * ex[j]) * M[1 * 5 + 1]
                             + ey[i] * ey[j] * M[0 * 5 + 2]
                             - z * (ex[i] * nh[j] + nh[i] * ex[j]) * M[1 * 5 + 0]
                             - z * (ey[i] * nh[j] + nh[i] * ey[j]) * M[0 * 5 + 1]
                             + z * z * nh[i] * nh[j] * M[0];
}

__device__ void build_T3(const double* M, const double* bas, double z, double* out)
{
    for (int q = 0; q < 27; ++q) out[q] = 0.0;
    for (int c1 = 0; c1 < 3; ++c1)
        for (int c2 = 0; c2 < 3; ++c2)
            for (int c3 = 0; c3 < 3; ++c3) {
                const int a = (c1 == 0) + (c2 == 0) + (c3 == 0);
                const int b = (c1 == 1) + (c2 == 1) + (c3 == 1);
                const int nz = 3 - a - b;
                const double sign = ((a + b) % 2 == 1) ? -1.0 : 1.0;
                double zp = 1.0;
                for (int t = 0; t < nz; ++t) zp *= z;
                const double coeff = sign * zp * M[a * 5 + b];
                if (coeff == 0.0) continue;
                for (int i = 0; i < 3; ++i) {
                    const double ci = coeff * bas[c1 * 3 + i];
                    for (int j = 0; j < 3; ++j) {
                        const double cij = ci * bas[c2 * 3 + j];
                        for (int k = 0; k < 3; ++k) out[i * 9 + j * 3 + k] += cij * bas[c3 * 3 + k];
                    }
                }
            }
}

__device__ void build_T4(const double* M, const double* bas, double z, double* out)
{
    for (int q = 0; q < 81; ++q) out[q] = 0.0;
    for (int c1 = 0; c1 < 3; ++c1)
        for (int c2 = 0; c2 < 3; ++c2)
            for (int c3 = 0; c3 < 3; ++c3)
                for (int c4 = 0; c4 < 3; ++c4) {
                    const int a = (c1 == 0) + (c2 == 0) + (c3 == 0) + (c4 == 0);
                    const int b = (c1 == 1) + (c2 == 1) + (c3 == 1) + (c4 == 1);
                    const int nz = 4 - a - b;
                    const double sign = ((a + b) % 2 == 1) ? -1.0 : 1.0;
                    double zp = 1.0;
                    for (int t = 0; t < nz; ++t) zp *= z;
                    const double coeff = sign * zp * M[a * 5 + b];
                    if (coeff == 0.0) continue;
                    for (int i = 0; i < 3; ++i) {
                        const double ci = coeff * bas[c1 * 3 + i];
                        for (int j = 0; j < 3; ++j) {
                            const double cij = ci * bas[c2 * 3 + j];
                            for (int k = 0; k < 3; ++k) {
                                const double cijk = cij * bas[c3 * 3 + k];
                                for (int l = 0; l < 3; ++l) out[i * 27 + j * 9 + k * 3 + l] += cijk * bas[c4 * 3 + l];
                            }
                        }
                    }
                }
}

// out[iN] = n_j tr(P_i), out[iR] = P_ijk n_k + n_k P_ikj (tri_kernels.py:230-247)
__device__ __forceinline__ void contract_NR(const double* P, const double* n, double* out, int iN, int iR)
{
    for (int i = 0; i < 3; ++i) {
        const double tr = P[i * 9 + 0] + P[i * 9 + 4] + P[i * 9 + 8];
        for (int j = 0; j < 3; ++j) {
            out[iN * 9 + i * 3 + j] = n[j] * tr;
            out[iR * 9 + i * 3 + j] = P[i * 9 + j * 3 + 0] * n[0] + P[i * 9 + j * 3 + 1] * n[1]
                                      + P[i * 9 + j * 3 + 2] * n[2] + n[0] * P[i * 9 + 0 * 3 + j]
                                      + n[1] * P[i * 9 + 1 * 3 + j] + n[2] * P[i * 9 + 2 * 3 + j];
        }
    }
}

// ---- per-node kernels (tri_nodal.py:732-893) -------------------------------
template <int K>
__device__ void u_blocks(double W[4][K][5][5], int k, double z, const double* bas, double eps,
                         double* out /* 27 */)
{
    const double I1 = W[0][k][0][0], I3 = W[1][k][0][0];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            out[i * 3 + j] = (i == j) ? I1 : 0.0;
            out[9 + i * 3 + j] = (i == j) ? eps * eps * I3 : 0.0;
        }
    build_T2(&W[1][k][0][0], bas, z, out + 18);
}

template <int K>
__device__ void t_blocks(double W[4][K][5][5], int k, double z, const double* bas, const double* nsrc,
                         double eps, double* out /* 54 */)
{
    const double* W3 = &W[1][k][0][0];
    const double* W5 = &W[2][k][0][0];
    double V3[3], V5[3], T35[27], Pt[27];
    for (int m = 0; m < 3; ++m) {
        V3[m] = -bas[m] * W3[5] - bas[3 + m] * W3[1] + z * bas[6 + m] * W3[0];
        V5[m] = -bas[m] * W5[5] - bas[3 + m] * W5[1] + z * bas[6 + m] * W5[0];
    }
    build_T3(W5, bas, z, T35);
    const double e2 = eps * eps;
    for (int i = 0; i < 3; ++i)
        for (int kk = 0; kk < 3; ++kk)
            for (int m = 0; m < 3; ++m) Pt[i * 9 + kk * 3 + m] = (i == kk) ? V3[m] : 0.0;
    contract_NR(Pt, nsrc, out, 0, 3);
    for (int i = 0; i < 3; ++i)
        for (int kk = 0; kk < 3; ++kk)
            for (int m = 0; m < 3; ++m) {
                double val = -3.0 * T35[i * 9 + kk * 3 + m];
                if (i == m) val += V3[kk];
                if (kk == m) val += V3[i];
                Pt[i * 9 + kk * 3 + m] = val;
            }
    contract_NR(Pt, nsrc, out, 1, 4);
    for (int i = 0; i < 3; ++i)
        for (int kk = 0; kk < 3; ++kk)
            for (int m = 0; m < 3; ++m) Pt[i * 9 + kk * 3 + m] = (i == kk) ? e2 * V5[m] : 0.0;
    contract_NR(Pt, nsrc, out, 2, 5);
}

template <int K>
__device__ void dd_stress(double W[4][K][5][5], int k, double z, const double* bas, const double* nrm,
                          double eps, double mu, double nu, double lam, double* H /* [m][n][j] */)
{
    const double I3 = W[1][k][0][0], I5 = W[2][k][0][0];
    double T2_5[9], T2_7[9], T4_7[81], ID2G[81], B[27];
    build_T2(&W[2][k][0][0], bas, z, T2_5);
    build_T2(&W[3][k][0][0], bas, z, T2_7);
    build_T4(&W[3][k][0][0], bas, z, T4_7);
    const double C1 = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
    const double c34 = 3.0 - 4.0 * nu;
    const double cblob = 2.0 * (1.0 - nu) * eps * eps;
    for (int r = 0; r < 3; ++r)
        for (int p = 0; p < 3; ++p) {
            const double drp = r == p ? 1.0 : 0.0;
            for (int s = 0; s < 3; ++s) {
                const double drs = r == s ? 1.0 : 0.0, dps = p == s ? 1.0 : 0.0;
                for (int q = 0; q < 3; ++q) {
                    const double dsq = s == q ? 1.0 : 0.0, dpq = p == q ? 1.0 : 0.0, drq = r == q ? 1.0 : 0.0;
                    const double val = (-c34 * drp * dsq * I3 + drs * dpq * I3 + dps * drq * I3
                                        + c34 * drp * 3.0 * T2_5[s * 3 + q] - drs * 3.0 * T2_5[p * 3 + q]
                                        - dps * 3.0 * T2_5[r * 3 + q] - 3.0 * drq * T2_5[p * 3 + s]
                                        - 3.0 * dpq * T2_5[r * 3 + s] - 3.0 * dsq * T2_5[r * 3 + p]
                                        + 15.0 * T4_7[r * 27 + p * 9 + s * 3 + q]
                                        + cblob * drp * (-3.0 * dsq * I5 + 15.0 * T2_7[s * 3 + q]));
                    ID2G[r * 27 + p * 9 + s * 3 + q] = C1 * val;
                }
            }
        }
    for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) {
            double tr = 0.0;
            for (int p = 0; p < 3; ++p) tr += ID2G[r * 27 + p * 9 + s * 3 + p];
            for (int kk = 0; kk < 3; ++kk) {
                double nq = 0.0, npd = 0.0;
                for (int q = 0; q < 3; ++q) nq += nrm[q] * ID2G[r * 27 + kk * 9 + s * 3 + q];
                for (int p = 0; p < 3; ++p) npd += nrm[p] * ID2G[r * 27 + p * 9 + s * 3 + kk];
                B[r * 9 + s * 3 + kk] = lam * nrm[kk] * tr + mu * nq + mu * npd;
            }
        }
    for (int kk = 0; kk < 3; ++kk) {
        const double trB = B[0 * 9 + 0 * 3 + kk] + B[1 * 9 + 1 * 3 + kk] + B[2 * 9 + 2 * 3 + kk];
        for (int m = 0; m < 3; ++m)
            for (int n = 0; n < 3; ++n) {
                const double dmn = m == n ? 1.0 : 0.0;
                H[m * 9 + n * 3 + kk] = -(lam * dmn * trB + mu * (B[m * 9 + n * 3 + kk] + B[n * 9 + m * 3 + kk]));
            }
    }
}

template <int K>
__device__ void kelvin_stress(double W[4][K][5][5], int k, double z, const double* bas, double eps,
                              double mu, double nu, double lam, double* S)
{
    const double* W3 = &W[1][k][0][0];
    const double* W5 = &W[2][k][0][0];
    double V3[3], V5[3], T35[27], G1[27];
    for (int m = 0; m < 3; ++m) {
        V3[m] = -bas[m] * W3[5] - bas[3 + m] * W3[1] + z * bas[6 + m] * W3[0];
        V5[m] = -bas[m] * W5[5] - bas[3 + m] * W5[1] + z * bas[6 + m] * W5[0];
    }
    build_T3(W5, bas, z, T35);
    const double C1 = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
    const double c34 = 3.0 - 4.0 * nu;
    const double cbdg = 6.0 * (1.0 - nu) * eps * eps;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double dij = i == j ? 1.0 : 0.0;
            for (int m = 0; m < 3; ++m) {
                const double dim = i == m ? 1.0 : 0.0, djm = j == m ? 1.0 : 0.0;
                G1[i * 9 + j * 3 + m] = C1 * (-c34 * dij * V3[m] + dim * V3[j] + djm * V3[i]
                                              - 3.0 * T35[i * 9 + j * 3 + m] - cbdg * dij * V5[m]);
            }
        }
    for (int c = 0; c < 3; ++c) {
        const double trace_c = G1[0 * 9 + c * 3 + 0] + G1[1 * 9 + c * 3 + 1] + G1[2 * 9 + c * 3 + 2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double dij = i == j ? 1.0 : 0.0;
                S[i * 9 + j * 3 + c] = lam * dij * trace_c + mu * G1[i * 9 + c * 3 + j] + mu * G1[j * 9 + c * 3 + i];
            }
    }
}

template <int K>
__device__ void eigen_stress(double W[4][K][5][5], int k, const double* nrm, double eps, double mu,
                             double lam, double* H)
{
    const double phi = (15.0 * eps * eps * eps * eps / (8.0 * M_PI)) * W[3][k][0][0];
    for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
            const double dmn = m == n ? 1.0 : 0.0;
            for (int j = 0; j < 3; ++j) {
                const double dmj = m == j ? 1.0 : 0.0, dnj = n == j ? 1.0 : 0.0;
                H[m * 9 + n * 3 + j] = phi * (lam * dmn * nrm[j] + mu * (dmj * nrm[n] + dnj * nrm[m]));
            }
        }
}

// ---- kernels ---------------------------------------------------------------
struct Coef6 {
    double c[6];
};

// KIND 0: U (force -> displacement, 3 basis blocks), 1: T (slip ->
// displacement, 6 blocks).  BASIS: every block into its own (3 nf, ld)
// matrix (block stride bstride); otherwise sum_b c_b block_b.
template <int KIND, int K, bool BASIS>
__global__ __launch_bounds__(256) void k_nodal_matrix(const double* __restrict__ xf, long nf,
                                                      const double* __restrict__ tri_verts,
                                                      const double* __restrict__ normals,
                                                      const double* __restrict__ eps_arr, long ns,
                                                      const double* __restrict__ geo_all, NSpec sp,
                                                      Coef6 cf, double* __restrict__ out, long ld,
                                                      long bstride)
{
    constexpr int NB = KIND == 0 ? 3 : 6;
    const int lane = threadIdx.x & 63;
    const long f = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (f >= nf) return;
    const double ox = xf[3 * f], oy = xf[3 * f + 1], oz = xf[3 * f + 2];
    for (long s0 = 0; s0 < ns; s0 += 64) {
        const long s = s0 + lane;
        if (s >= ns) continue;
        const double* geo = geo_all + kGeo * s;
        double W[4][K][5][5];
        double blk[NB * 9];
        const bool ok = geo[29] != 0.0;
        double z = 0.0;
        if (ok) z = pair_tables<K>(tri_verts + 9 * s, geo, ox, oy, oz, eps_arr[s], sp, W);
        for (int k = 0; k < K; ++k) {
            if (ok) {
                if (KIND == 0) u_blocks<K>(W, k, z, geo, eps_arr[s], blk);
                else t_blocks<K>(W, k, z, geo, normals + 3 * s, eps_arr[s], blk);
            } else {
                for (int q = 0; q < NB * 9; ++q) blk[q] = 0.0;
            }
            const long col = 3 * (K * s + k);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    double* row = out + (3 * f + i) * ld + col + j;
                    if (BASIS) {
                        for (int b = 0; b < NB; ++b) row[b * bstride] = blk[b * 9 + i * 3 + j];
                    } else {
                        double acc;
                        if (KIND == 0) {
                            acc = (cf.c[0] * blk[i * 3 + j] + cf.c[1] * blk[9 + i * 3 + j]) + cf.c[2] * blk[18 + i * 3 + j];
                        } else {
                            acc = 0.0;
                            for (int b = 0; b < 6; ++b) acc += cf.c[b] * blk[b * 9 + i * 3 + j];
                        }
                        *row = acc;
                    }
                }
        }
    }
}

__device__ __forceinline__ bool block_zero(const double* dens, int K, long s)
{
    for (int q = 0; q < 3 * K; ++q)
        if (dens[3 * K * s + q] != 0.0) return false;
    return true;
}

// KIND 0: U displacement, 1: T displacement (out nf x 3); 2: force ->
// stress, 3: slip -> total stress, 4: eigenstress (out nf x 9)
template <int KIND, int K>
__global__ __launch_bounds__(256) void k_nodal_contract(const double* __restrict__ xf, long nf,
                                                        const double* __restrict__ tri_verts,
                                                        const double* __restrict__ normals,
                                                        const double* __restrict__ eps_arr, long ns,
                                                        const double* __restrict__ geo_all, NSpec sp,
                                                        const double* __restrict__ dens, Coef6 cf,
                                                        double mu, double nu, double lam,
                                                        double* __restrict__ out)
{
    constexpr int NO = KIND <= 1 ? 3 : 9;
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long fc = f < nf ? f : nf - 1;
    const double ox = xf[3 * fc], oy = xf[3 * fc + 1], oz = xf[3 * fc + 2];
    double acc[NO];
    for (int q = 0; q < NO; ++q) acc[q] = 0.0;
    for (long s = 0; s < ns; ++s) {
        const double* geo = geo_all + kGeo * s;
        if (geo[29] == 0.0 || block_zero(dens, K, s)) continue;
        const double eps = eps_arr[s];
        double W[4][K][5][5];
        const double z = pair_tables<K>(tri_verts + 9 * s, geo, ox, oy, oz, eps, sp, W);
        for (int k = 0; k < K; ++k) {
            const double* d = dens + 3 * (K * s + k);
            if (KIND <= 1) {
                double blk[54];
                if (KIND == 0) u_blocks<K>(W, k, z, geo, eps, blk);
                else t_blocks<K>(W, k, z, geo, normals + 3 * s, eps, blk);
                for (int j = 0; j < 3; ++j) {
                    const double dj = d[j];
                    if (dj == 0.0) continue;
                    for (int i = 0; i < 3; ++i) {
                        double v;
                        if (KIND == 0) {
                            v = cf.c[0] * blk[i * 3 + j] + cf.c[1] * blk[9 + i * 3 + j] + cf.c[2] * blk[18 + i * 3 + j];
                        } else {
                            v = 0.0;
                            for (int b = 0; b < 6; ++b) v += cf.c[b] * blk[b * 9 + i * 3 + j];
                        }
                        acc[i] += v * dj;
                    }
                }
            } else {
                double H[27];
                if (KIND == 2) kelvin_stress<K>(W, k, z, geo, eps, mu, nu, lam, H);
                else if (KIND == 3) dd_stress<K>(W, k, z, geo, normals + 3 * s, eps, mu, nu, lam, H);
                else eigen_stress<K>(W, k, normals + 3 * s, eps, mu, lam, H);
                for (int m = 0; m < 3; ++m)
                    for (int n = 0; n < 3; ++n)
                        acc[m * 3 + n] += H[m * 9 + n * 3 + 0] * d[0] + H[m * 9 + n * 3 + 1] * d[1] + H[m * 9 + n * 3 + 2] * d[2];
            }
        }
    }
    if (f >= nf) return;
    for (int q = 0; q < NO; ++q) out[NO * f + q] = acc[q];
}

// traction rows (3 nf x 3 n_cols, leading dimension ld, accumulated):
// WHICH 0 slip -> elastic (total minus eigenstress), 1 force, 2 eigenstress,
// 3 slip -> total (tri_nodal.py:1175-1227)
template <int WHICH, int K>
__global__ __launch_bounds__(256) void k_nodal_traction(const double* __restrict__ xf,
                                                        const double* __restrict__ nfield, long nf,
                                                        const double* __restrict__ tri_verts,
                                                        const double* __restrict__ normals,
                                                        const double* __restrict__ eps_arr, long ns,
                                                        const double* __restrict__ geo_all, NSpec sp,
                                                        const int64_t* __restrict__ ids, double mu,
                                                        double nu, double lam, double* __restrict__ out,
                                                        long ld)
{
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long fc = f < nf ? f : nf - 1;
    const double ox = xf[3 * fc], oy = xf[3 * fc + 1], oz = xf[3 * fc + 2];
    const double n0 = nfield[3 * fc], n1 = nfield[3 * fc + 1], n2 = nfield[3 * fc + 2];
    const bool live = f < nf;
    for (long s = 0; s < ns; ++s) {
        const double* geo = geo_all + kGeo * s;
        if (geo[29] == 0.0) continue;
        const double eps = eps_arr[s];
        double W[4][K][5][5];
        const double z = pair_tables<K>(tri_verts + 9 * s, geo, ox, oy, oz, eps, sp, W);
        for (int k = 0; k < K; ++k) {
            double H[27];
            if (WHICH == 0 || WHICH == 3) dd_stress<K>(W, k, z, geo, normals + 3 * s, eps, mu, nu, lam, H);
            else if (WHICH == 1) kelvin_stress<K>(W, k, z, geo, eps, mu, nu, lam, H);
            else eigen_stress<K>(W, k, normals + 3 * s, eps, mu, lam, H);
            if (WHICH == 0) {
                double HE[27];
                eigen_stress<K>(W, k, normals + 3 * s, eps, mu, lam, HE);
                for (int q = 0; q < 27; ++q) H[q] -= HE[q];
            }
            if (!live) continue;
            const long col = 3 * ids[K * s + k];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    out[(3 * f + i) * ld + col + j] += n0 * H[i * 9 + 0 * 3 + j] + n1 * H[i * 9 + 1 * 3 + j]
                                                       + n2 * H[i * 9 + 2 * 3 + j];
        }
    }
}

inline int launch_status() { return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP; }

// spec from its host integer record and one device buffer of tables:
// ispec = {order, K, q_split, n_terms, nq, degn[5], ekmax[4], need0[4]};
// tables = binom (4 n_terms) | lattice (3 K) | q_lam (3 nq) | q_w (nq) | q_N (K nq)
int make_spec(const int64_t* ispec, const double* fspec, const double* tables, NSpec& sp)
{
    if (!ispec || !fspec || !tables) return MBEM_EINVAL;
    sp.order = (int)ispec[0];
    sp.K = (int)ispec[1];
    sp.q_split = (int)ispec[2];
    sp.n_terms = (int)ispec[3];
    sp.nq = (int)ispec[4];
    if (sp.order < 0 || sp.order > 2 || sp.K != (sp.order + 1) * (sp.order + 2) / 2) return MBEM_EINVAL;
    if (sp.n_terms < 1 || 2 * sp.n_terms + 2 + kMaxEdge + 3 > kSer) return MBEM_EINVAL;
    if (sp.q_split < 0 || sp.q_split > sp.nq) return MBEM_EINVAL;
    for (int i = 0; i < 5; ++i) {
        sp.degn[i] = (int)ispec[5 + i];
        if (sp.degn[i] > kMaxTable) return MBEM_EINVAL;
    }
    for (int i = 0; i < 4; ++i) {
        sp.ekmax[i] = (int)ispec[10 + i];
        if (sp.ekmax[i] > kMaxEdge) return MBEM_EINVAL;
    }
    for (int i = 0; i < 4; ++i) {
        sp.need0[i] = (int)ispec[14 + i];
        if (sp.need0[i] > 4) return MBEM_EINVAL;
    }
    if (sp.degn[0] >= 1) return MBEM_EINVAL;
    sp.dstar = fspec[0];
    sp.dstar_far = fspec[1];
    sp.small_u = fspec[2];
    sp.series_u = fspec[3];
    sp.binom = tables;
    sp.lattice = sp.binom + 4 * sp.n_terms;
    sp.q_lam = sp.lattice + 3 * sp.K;
    sp.q_w = sp.q_lam + 3 * sp.nq;
    sp.q_N = sp.q_w + sp.nq;
    return MBEM_OK;
}

}  // namespace

extern "C" {

int64_t mbem_nodal_geo_doubles(int64_t ns) { return ns < 0 ? 0 : kGeo * ns; }

int64_t mbem_nodal_table_doubles(int64_t order, int64_t n_terms, int64_t nq)
{
    const int64_t K = (order + 1) * (order + 2) / 2;
    return 4 * n_terms + 3 * K + 3 * nq + nq + K * nq;
}

int mbem_nodal_frames(const double* tri_verts, int64_t ns, double* geo, hipStream_t stream)
{
    if (ns < 0 || (ns > 0 && (!tri_verts || !geo))) return MBEM_EINVAL;
    if (ns == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_nodal_frames, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, stream,
                       tri_verts, (long)ns, geo);
    return launch_status();
}

#define MBEM_BY_K(K_, CALL) \
    switch (K_) {           \
        case 1: { constexpr int KK = 1; CALL; break; } \
        case 3: { constexpr int KK = 3; CALL; break; } \
        case 6: { constexpr int KK = 6; CALL; break; } \
        default: return MBEM_EINVAL; \
    }

int mbem_nodal_matrix(int kind, const double* x_field, int64_t nf, const double* tri_verts,
                      const double* normals, const double* eps, int64_t ns, const double* geo,
                      const int64_t* ispec, const double* fspec, const double* tables,
                      const double* coeffs, double* out, int64_t ld, int64_t bstride,
                      hipStream_t stream)
{
    NSpec sp;
    if (make_spec(ispec, fspec, tables, sp) != MBEM_OK) return MBEM_EINVAL;
    if ((kind != 0 && kind != 1) || nf < 0 || ns < 0) return MBEM_EINVAL;
    if (nf == 0 || ns == 0) return MBEM_OK;
    if (!x_field || !tri_verts || !eps || !geo || !out || (kind == 1 && !normals)) return MBEM_EINVAL;
    if (ld < 3 * sp.K * ns) return MBEM_EINVAL;
    const bool basis = coeffs == nullptr;
    if (basis && bstride < 3 * nf * ld) return MBEM_EINVAL;
    Coef6 cf{};
    if (!basis)
        for (int q = 0; q < (kind == 0 ? 3 : 6); ++q) cf.c[q] = coeffs[q];
    dim3 grid((unsigned)((nf + 3) / 4)), block(256);
#define MBEM_MAT(KIND, B) hipLaunchKernelGGL((k_nodal_matrix<KIND, KK, B>), grid, block, 0, stream, \
        x_field, (long)nf, tri_verts, normals, eps, (long)ns, geo, sp, cf, out, (long)ld, (long)bstride)
    if (kind == 0 && basis) { MBEM_BY_K(sp.K, MBEM_MAT(0, true)) }
    else if (kind == 0) { MBEM_BY_K(sp.K, MBEM_MAT(0, false)) }
    else if (basis) { MBEM_BY_K(sp.K, MBEM_MAT(1, true)) }
    else { MBEM_BY_K(sp.K, MBEM_MAT(1, false)) }
#undef MBEM_MAT
    return launch_status();
}

int mbem_nodal_contract(int kind, const double* x_field, int64_t nf, const double* tri_verts,
                        const double* normals, const double* eps, int64_t ns, const double* geo,
                        const int64_t* ispec, const double* fspec, const double* tables,
                        const double* density, const double* coeffs, double mu, double nu,
                        double lam, double* out, hipStream_t stream)
{
    NSpec sp;
    if (make_spec(ispec, fspec, tables, sp) != MBEM_OK) return MBEM_EINVAL;
    if (kind < 0 || kind > 4 || nf < 0 || ns < 0) return MBEM_EINVAL;
    if (nf == 0) return MBEM_OK;
    if (!x_field || !out) return MBEM_EINVAL;
    if (ns > 0 && (!tri_verts || !eps || !geo || !density || (kind != 0 && kind != 2 && !normals)))
        return MBEM_EINVAL;
    if (kind <= 1 && !coeffs) return MBEM_EINVAL;
    Coef6 cf{};
    if (kind <= 1)
        for (int q = 0; q < (kind == 0 ? 3 : 6); ++q) cf.c[q] = coeffs[q];
    dim3 grid((unsigned)((nf + 255) / 256)), block(256);
#define MBEM_CON(KIND) hipLaunchKernelGGL((k_nodal_contract<KIND, KK>), grid, block, 0, stream, \
        x_field, (long)nf, tri_verts, normals, eps, (long)ns, geo, sp, density, cf, mu, nu, lam, out)
    switch (kind) {
        case 0: { MBEM_BY_K(sp.K, MBEM_CON(0)) } break;
        case 1: { MBEM_BY_K(sp.K, MBEM_CON(1)) } break;
        case 2: { MBEM_BY_K(sp.K, MBEM_CON(2)) } break;
        case 3: { MBEM_BY_K(sp.K, MBEM_CON(3)) } break;
        default: { MBEM_BY_K(sp.K, MBEM_CON(4)) } break;
    }
#undef MBEM_CON
    return launch_status();
}

int mbem_nodal_traction(int which, const double* x_field, const double* n_field, int64_t nf,
                        const double* tri_verts, const double* normals, const double* eps,
                        int64_t ns, const double* geo, const int64_t* ispec, const double* fspec,
                        const double* tables, const int64_t* ids, int64_t n_cols, double mu,
                        double nu, double lam, double* out, int64_t ld, hipStream_t stream)
{
    NSpec sp;
    if (make_spec(ispec, fspec, tables, sp) != MBEM_OK) return MBEM_EINVAL;
    if (which < 0 || which > 3 || nf < 0 || ns < 0 || n_cols < 0) return MBEM_EINVAL;
    if (nf == 0 || ns == 0) return MBEM_OK;
    if (!x_field || !n_field || !tri_verts || !normals || !eps || !geo || !ids || !out) return MBEM_EINVAL;
    if (ld < 3 * n_cols) return MBEM_EINVAL;
    dim3 grid((unsigned)((nf + 255) / 256)), block(256);
#define MBEM_TR(W) hipLaunchKernelGGL((k_nodal_traction<W, KK>), grid, block, 0, stream, x_field, \
        n_field, (long)nf, tri_verts, normals, eps, (long)ns, geo, sp, ids, mu, nu, lam, out, (long)ld)
    switch (which) {
        case 0: { MBEM_BY_K(sp.K, MBEM_TR(0)) } break;
        case 1: { MBEM_BY_K(sp.K, MBEM_TR(1)) } break;
        case 2: { MBEM_BY_K(sp.K, MBEM_TR(2)) } break;
        default: { MBEM_BY_K(sp.K, MBEM_TR(3)) } break;
    }
#undef MBEM_TR
    return launch_status();
}

#undef MBEM_BY_K

}  // extern "C"
