/*
 * ORACLE -- test infrastructure only.  Not shipped, not measured as the
 * product.  Only tests/ may load this file's library.
 *
 * Plain-C restatement, step for step, of the Lagrange P0/P1/P2 nodal pair
 * machinery of mbem's src/mbem/kernels/tri_nodal.py: the per-source nodal
 * frame (:418-456), the cancellation-free edge primitives with their
 * small-|u| and large-|u| series and the closed form (:248-415), the
 * shape-function coefficients about the foot point (:473-516), the moment
 * table of the divergence-theorem recursion (:519-620), the weighted tables
 * (:623-638), the far-field product Gauss tables (:641-680), the pair driver
 * (:683-709) and the per-node kernels (:732-893), with the matrix, contraction
 * and traction-row drivers (:896-1227).  Expression order follows the
 * reference's (including pow where it uses **); build with
 * -ffp-contract=off.  The spec (degrees, series coefficients, Gauss rules,
 * node lattice) is built in oracle/tri_nodal.py and passed in.
 *
 * The P0 helpers (tri_frame, solid_angle, build_T2/T3/T4, contract_NR) are
 * the ones of the P0 restatement, included below.
 */
#include "tri_kernels_ref.c"

#include <stdlib.h>

#define MAX_EDGE_DEG 5
#define MAX_TABLE_DEG 6
#define GEO_LEN 29
#define MAXK 6
#define MAX_SER 128

typedef struct {
    int order, K, q_split;
    int degn[5];     /* table degree for n = -1, 1, 3, 5, 7 (or -1) */
    int ekmax[4];    /* edge-primitive degree for m = -1, 1, 3, 5 (or -1) */
    int need0[4];    /* kernel degree at order 0 for n = 1, 3, 5, 7 (or -1) */
    double dstar, dstar_far, small_u, series_u;
    int n_terms;
    const double *binom;     /* 4 x n_terms */
    const int64_t *lattice;  /* K x 3 */
    int nq;
    const double *q_lam, *q_w, *q_N;   /* nq x 3, nq, nq x K */
} nspec;

static double COMB[MAX_EDGE_DEG + 1][MAX_EDGE_DEG + 1];

static void init_comb(void)
{
    for (int a = 0; a <= MAX_EDGE_DEG; ++a)
        for (int i = 0; i <= MAX_EDGE_DEG; ++i) {
            double c = 0.0;
            if (i <= a) {
                c = 1.0;
                for (int t = 0; t < i; ++t) c = c * (a - t) / (t + 1);
            }
            COMB[a][i] = c;
        }
}

static void unpack_spec(nspec *sp, const int64_t *ispec, const double *fspec,
                        const int64_t *lattice, const double *binom, int64_t n_terms,
                        const double *q_lam, const double *q_w, const double *q_N, int64_t nq)
{
    sp->order = (int)ispec[0];
    sp->K = (int)ispec[1];
    sp->q_split = (int)ispec[2];
    for (int i = 0; i < 5; ++i) sp->degn[i] = (int)ispec[3 + i];
    for (int i = 0; i < 4; ++i) sp->ekmax[i] = (int)ispec[8 + i];
    for (int i = 0; i < 4; ++i) sp->need0[i] = (int)ispec[12 + i];
    sp->dstar = fspec[0];
    sp->dstar_far = fspec[1];
    sp->small_u = fspec[2];
    sp->series_u = fspec[3];
    sp->n_terms = (int)n_terms;
    sp->binom = binom;
    sp->lattice = lattice;
    sp->nq = (int)nq;
    sp->q_lam = q_lam;
    sp->q_w = q_w;
    sp->q_N = q_N;
    init_comb();
}

/* x ** n for a small int n by multiplication: tri_nodal.py:248-259 */
static double ipow_s(double x, int n)
{
    double out = 1.0;
    if (n >= 0) {
        for (int i = 0; i < n; ++i) out = out * x;
    } else {
        double xi = 1.0 / x;
        for (int i = 0; i < -n; ++i) out = out * xi;
    }
    return out;
}

#define SER_OFF(nt) (2 * (nt) + 2)

/* large-|u| series, 0 < ua < ub: tri_nodal.py:262-300 */
static void series_large_row(int m, int kmax, double ua, double ub, double rho2, const double *bm,
                             int n_terms, double P[4][MAX_EDGE_DEG + 1], int mi)
{
    double T[MAX_SER];
    double x = rho2 / (ua * ua);
    double r = ub / ua;
    double log_r = log1p((ub - ua) / ua);
    double up = (ub - ua) / ua;
    double dn = -(ub - ua) / ub;
    double rinv = 1.0 / r;
    int off = SER_OFF(n_terms);
    T[off] = 0.0;
    int qmax = kmax - m + 1;
    for (int q = 1; q <= qmax; ++q) T[off + q] = T[off + q - 1] * r + up;
    int qmin = 1 - m - 2 * (n_terms - 1);
    for (int q = -1; q >= qmin; --q) T[off + q] = T[off + q + 1] * rinv + dn;
    double base = ipow_s(ua, 1 - m);
    for (int k = 0; k <= kmax; ++k) {
        double out = 0.0, xj = 1.0;
        for (int j = 0; j < n_terms; ++j) {
            int q = k - m + 1 - 2 * j;
            double term = (q == 0) ? log_r : T[off + q] / q;
            out = out + bm[j] * xj * term;
            xj = xj * x;
        }
        P[mi][k] = base * out;
        base = base * ua;
    }
}

/* small-|u| series, max |u| <= rho / 2: tri_nodal.py:303-343 */
static void series_small_row(int m, int kmax, double ua, double ub, double rho2, const double *bm,
                             int n_terms, double P[4][MAX_EDGE_DEG + 1], int mi)
{
    double YA[MAX_SER], YB[MAX_SER], E[MAX_SER];
    double rho = sqrt(rho2);
    double ya = ua / rho, yb = ub / rho;
    int pmax = kmax + 2 * (n_terms - 1) + 1;
    YA[0] = 1.0;
    YB[0] = 1.0;
    for (int p = 1; p <= pmax; ++p) {
        YA[p] = YA[p - 1] * ya;
        YB[p] = YB[p - 1] * yb;
    }
    int close = 0;
    if (ua * ub > 0.0) {
        double ratio = ub / ua;
        close = (ratio < 2.0) && (ratio > 0.5);
        if (close) {
            double up = (ub - ua) / ua;
            E[0] = 0.0;
            for (int p = 1; p <= pmax; ++p) E[p] = E[p - 1] * ratio + up;
        }
    }
    double scale = ipow_s(rho, 1 - m);
    for (int k = 0; k <= kmax; ++k) {
        double out = 0.0;
        for (int j = 0; j < n_terms; ++j) {
            int pw = k + 2 * j + 1;
            double dpow = close ? YA[pw] * E[pw] : YB[pw] - YA[pw];
            out = out + bm[j] * dpow / pw;
        }
        P[mi][k] = scale * out;
        scale = scale * rho;
    }
}

/* conjugate closed-form differences: tri_nodal.py:346-383 */
static void closed_form(double ua, double ub, double rho2, const int *ekmax,
                        double P[4][MAX_EDGE_DEG + 1])
{
    double Ra = sqrt(ua * ua + rho2);
    double Rb = sqrt(ub * ub + rho2);
    double dR = (ub - ua) * (ub + ua) / (Ra + Rb);
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
    double duRm = ub * pow(Rb, -3.0) - ua * pow(Ra, -3.0);
    double dJ5 = (duRm + 2.0 * dJ3) / (3.0 * rho2);
    double duR = 0.5 * (ub - ua) * (Ra + Rb) + 0.5 * (ua + ub) * dR;
    double dJm1 = 0.5 * (duR + rho2 * dJ1);
    double RaRb = Ra * Rb;
    double S3 = (Rb * Rb + Ra * Rb) + Ra * Ra;
    double dK[4] = {dR * S3 / 3.0, dR, dR / RaRb, dR * S3 / (3.0 * pow(RaRb, 3.0))};
    double dJ[4] = {dJm1, dJ1, dJ3, dJ5};
    for (int mi = 0; mi < 4; ++mi) {
        int kmax = ekmax[mi];
        if (kmax < 0) continue;
        P[mi][0] = dJ[mi];
        if (kmax >= 1) P[mi][1] = dK[mi];
    }
    for (int k = 2; k <= MAX_EDGE_DEG; ++k)
        for (int mi = 1; mi < 4; ++mi)
            if (ekmax[mi] >= k) P[mi][k] = P[mi - 1][k - 2] - rho2 * P[mi][k - 2];
}

/* branch between the two series and the closed form: tri_nodal.py:386-415 */
static void edge_primitives(double ua, double ub, double rho2, const nspec *sp,
                            double P[4][MAX_EDGE_DEG + 1])
{
    double rho = sqrt(rho2);
    double aua = fabs(ua), aub = fabs(ub);
    double umax = aua > aub ? aua : aub;
    double umin = aua < aub ? aua : aub;
    int nt = sp->n_terms;
    if (umax <= sp->small_u * rho) {
        for (int mi = 0; mi < 4; ++mi)
            if (sp->ekmax[mi] >= 0)
                series_small_row(2 * mi - 1, sp->ekmax[mi], ua, ub, rho2, sp->binom + mi * nt, nt,
                                 P, mi);
    } else if ((ua * ub > 0.0) && umin >= sp->series_u * rho) {
        int negative = ua < 0.0;
        double a = negative ? -ub : ua;
        double b = negative ? -ua : ub;
        for (int mi = 0; mi < 4; ++mi)
            if (sp->ekmax[mi] >= 0) {
                series_large_row(2 * mi - 1, sp->ekmax[mi], a, b, rho2, sp->binom + mi * nt, nt,
                                 P, mi);
                if (negative)
                    for (int k = 1; k <= sp->ekmax[mi]; k += 2) P[mi][k] = -P[mi][k];
            }
    } else {
        closed_form(ua, ub, rho2, sp->ekmax, P);
    }
}

/* geo: ex ey nhat centroid p2d(3 x 2) abc(3 x 3) L area2: tri_nodal.py:418-456 */
static int nodal_frame(const double *tv, double *geo)
{
    double *ex = geo, *ey = geo + 3, *nhat = geo + 6;
    if (!tri_frame(tv, ex, ey, nhat)) return 0;
    for (int c = 0; c < 3; ++c) geo[9 + c] = ((tv[c] + tv[3 + c]) + tv[6 + c]) / 3.0;
    for (int k = 0; k < 3; ++k) {
        double dx = tv[3 * k] - geo[9], dy = tv[3 * k + 1] - geo[10], dz = tv[3 * k + 2] - geo[11];
        geo[12 + 2 * k] = dx * ex[0] + dy * ex[1] + dz * ex[2];
        geo[13 + 2 * k] = dx * ey[0] + dy * ey[1] + dz * ey[2];
    }
    double L = 0.0;
    for (int k = 0; k < 3; ++k) {
        int k2 = (k + 1) % 3;
        double dx = tv[3 * k2] - tv[3 * k], dy = tv[3 * k2 + 1] - tv[3 * k + 1];
        double dz = tv[3 * k2 + 2] - tv[3 * k + 2];
        double l = sqrt(dx * dx + dy * dy + dz * dz);
        L = L > l ? L : l;
    }
    geo[27] = L;
    double e1x = tv[3] - tv[0], e1y = tv[4] - tv[1], e1z = tv[5] - tv[2];
    double e2x = tv[6] - tv[0], e2y = tv[7] - tv[1], e2z = tv[8] - tv[2];
    double nx = e1y * e2z - e1z * e2y;
    double ny = e1z * e2x - e1x * e2z;
    double nz = e1x * e2y - e1y * e2x;
    geo[28] = sqrt(nx * nx + ny * ny + nz * nz);
    double den = ((geo[14] - geo[12]) * (geo[17] - geo[13]) - (geo[15] - geo[13]) * (geo[16] - geo[12]));
    for (int k = 0; k < 3; ++k) {
        int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        double x1 = geo[12 + 2 * k1], y1 = geo[13 + 2 * k1];
        double x2 = geo[12 + 2 * k2], y2 = geo[13 + 2 * k2];
        geo[18 + 3 * k] = (x1 * y2 - y1 * x2) / den;
        geo[19 + 3 * k] = (y1 - y2) / den;
        geo[20 + 3 * k] = (x2 - x1) / den;
    }
    return 1;
}

/* C[k][a][b]: N_k(xi) = sum c xi_1^a xi_2^b about the foot X: tri_nodal.py:473-516 */
static void shape_coeffs(const nspec *sp, const double *geo, double X1, double X2,
                         double C[MAXK][3][3])
{
    int K = sp->K, order = sp->order, D = order + 1;
    for (int k = 0; k < K; ++k)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) C[k][a][b] = 0.0;
    if (order == 0) {
        C[0][0][0] = 1.0;
        return;
    }
    double poly[3][3], fac[3][3], out[3][3];
    for (int k = 0; k < K; ++k) {
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) {
                poly[a][b] = 0.0;
                fac[a][b] = 0.0;
            }
        poly[0][0] = 1.0;
        for (int i = 0; i < 3; ++i) {
            double l0 = (geo[18 + 3 * i] + X1 * geo[19 + 3 * i]) + X2 * geo[20 + 3 * i];
            for (int j = 0; j < (int)sp->lattice[3 * k + i]; ++j) {
                fac[0][0] = (order * l0) / (j + 1) - (double)j / (j + 1);
                fac[1][0] = (order * geo[19 + 3 * i]) / (j + 1);
                fac[0][1] = (order * geo[20 + 3 * i]) / (j + 1);
                for (int a = 0; a < D; ++a)
                    for (int b = 0; b < D; ++b) out[a][b] = 0.0;
                for (int a = 0; a < D; ++a)
                    for (int b = 0; b < D - a; ++b) {
                        double pab = poly[a][b];
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

/* M[ni][a][b] = int_T xi_1^a xi_2^b / R^n dS, n = -1, 1, 3, 5, 7:
 * tri_nodal.py:519-620 */
static void moment_table(const double *tv, const double *geo, double X1, double X2, double h2,
                         const nspec *sp, double M[5][MAX_TABLE_DEG + 1][MAX_TABLE_DEG + 1])
{
    const int *ekmax = sp->ekmax;
    int kall = 0;
    for (int mi = 0; mi < 4; ++mi) kall = kall > ekmax[mi] ? kall : ekmax[mi];
    double h = sqrt(h2);
    double ox = ((geo[9] + X1 * geo[0]) + X2 * geo[3]) + h * geo[6];
    double oy = ((geo[10] + X1 * geo[1]) + X2 * geo[4]) + h * geo[7];
    double oz = ((geo[11] + X1 * geo[2]) + X2 * geo[5]) + h * geo[8];
    double I3 = -solid_angle(tv, tv + 3, tv + 6, ox, oy, oz) / h;
    double BD[4], BN1[4][MAX_EDGE_DEG + 1][MAX_EDGE_DEG + 1], BN2[4][MAX_EDGE_DEG + 1][MAX_EDGE_DEG + 1];
    double P[4][MAX_EDGE_DEG + 1];
    memset(P, 0, sizeof(P));
    for (int mi = 0; mi < 4; ++mi) {
        BD[mi] = 0.0;
        for (int a = 0; a <= MAX_EDGE_DEG; ++a)
            for (int b = 0; b <= MAX_EDGE_DEG; ++b) {
                BN1[mi][a][b] = 0.0;
                BN2[mi][a][b] = 0.0;
            }
    }
    double dpP[MAX_EDGE_DEG + 1], c1P[MAX_EDGE_DEG + 1], s1P[MAX_EDGE_DEG + 1];
    double c2P[MAX_EDGE_DEG + 1], s2P[MAX_EDGE_DEG + 1];
    dpP[0] = 1.0; c1P[0] = 1.0; s1P[0] = 1.0; c2P[0] = 1.0; s2P[0] = 1.0;
    for (int e = 0; e < 3; ++e) {
        int e2 = (e + 1) % 3;
        double pax = geo[12 + 2 * e] - X1, pay = geo[13 + 2 * e] - X2;
        double pbx = geo[12 + 2 * e2] - X1, pby = geo[13 + 2 * e2] - X2;
        double evx = pbx - pax, evy = pby - pay;
        double Le = sqrt(evx * evx + evy * evy);
        double tx = evx / Le, ty = evy / Le;
        double c1 = ty, c2 = -tx;
        double dp = pax * c1 + pay * c2;
        double ua = pax * tx + pay * ty;
        double ub = pbx * tx + pby * ty;
        double rho2 = dp * dp + h2;
        edge_primitives(ua, ub, rho2, sp, P);
        for (int q = 1; q <= kall; ++q) {
            dpP[q] = dpP[q - 1] * dp;
            c1P[q] = c1P[q - 1] * c1;
            s1P[q] = s1P[q - 1] * tx;
            c2P[q] = c2P[q - 1] * c2;
            s2P[q] = s2P[q - 1] * ty;
        }
        for (int mi = 0; mi < 4; ++mi) {
            int kmax = ekmax[mi];
            if (kmax < 0) continue;
            BD[mi] = BD[mi] + dp * P[mi][0];
            for (int a = 0; a <= kmax; ++a)
                for (int b = 0; b <= kmax - a; ++b) {
                    double val = 0.0;
                    for (int i = 0; i <= a; ++i)
                        for (int j = 0; j <= b; ++j) {
                            int pw = a + b - i - j;
                            double coeff = (COMB[a][i] * COMB[b][j]) * dpP[pw];
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
    const int *degn = sp->degn;
    M[2][0][0] = I3;
    if (degn[3] >= 0) M[3][0][0] = (BD[2] + I3) / (3.0 * h2);
    if (degn[4] >= 0) M[4][0][0] = (BD[3] + 3.0 * M[3][0][0]) / (5.0 * h2);
    if (degn[1] >= 0) M[1][0][0] = (BD[1] - (1.0 * h2) * M[2][0][0]) / 1.0;
    if (degn[0] >= 0) M[0][0][0] = (BD[0] - (-1.0 * h2) * M[1][0][0]) / 3.0;
    for (int ni = 1; ni < 5; ++ni) {
        int d = degn[ni];
        if (d < 1) continue;
        int n = 2 * ni - 1, mi = ni - 1;
        for (int q = 1; q <= d; ++q)
            for (int a = q; a >= 0; --a) {
                int b = q - a;
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

/* W[ni2][k][a][b] = sum C[k][a'][b'] M[ni2 + 1][a + a'][b + b']: tri_nodal.py:623-638 */
static void weighted_tables(const nspec *sp, double C[MAXK][3][3],
                            double M[5][MAX_TABLE_DEG + 1][MAX_TABLE_DEG + 1],
                            double W[4][MAXK][5][5])
{
    for (int ni2 = 0; ni2 < 4; ++ni2) {
        int d = sp->need0[ni2];
        if (d < 0) continue;
        int ni = ni2 + 1;
        for (int k = 0; k < sp->K; ++k)
            for (int a = 0; a <= d; ++a)
                for (int b = 0; b <= d - a; ++b) {
                    double acc = 0.0;
                    for (int ap = 0; ap <= sp->order; ++ap)
                        for (int bp = 0; bp <= sp->order - ap; ++bp)
                            acc += C[k][ap][bp] * M[ni][a + ap][b + bp];
                    W[ni2][k][a][b] = acc;
                }
    }
}

/* Gauss rule [q0, q1) of N_k xi_1^a xi_2^b / R^n: tri_nodal.py:641-680 */
static void far_tables(const nspec *sp, const double *geo, double X1, double X2, double h2, int q0,
                       int q1, double W[4][MAXK][5][5])
{
    double area2 = geo[28];
    int dmax = 0;
    for (int ni2 = 0; ni2 < 4; ++ni2) {
        int d = sp->need0[ni2];
        dmax = dmax > d ? dmax : d;
        if (d < 0) continue;
        for (int k = 0; k < sp->K; ++k)
            for (int a = 0; a <= d; ++a)
                for (int b = 0; b <= d - a; ++b) W[ni2][k][a][b] = 0.0;
    }
    double x1p[MAX_EDGE_DEG + 1], x2p[MAX_EDGE_DEG + 1];
    x1p[0] = 1.0;
    x2p[0] = 1.0;
    for (int q = q0; q < q1; ++q) {
        double l1 = sp->q_lam[3 * q], l2 = sp->q_lam[3 * q + 1], l3 = sp->q_lam[3 * q + 2];
        double eta1 = l1 * geo[12] + l2 * geo[14] + l3 * geo[16];
        double eta2 = l1 * geo[13] + l2 * geo[15] + l3 * geo[17];
        double xi1 = eta1 - X1, xi2 = eta2 - X2;
        double R2 = xi1 * xi1 + xi2 * xi2 + h2;
        double wq = sp->q_w[q] * area2;
        double r2i = 1.0 / R2;
        double Rn = sqrt(r2i);
        for (int a = 1; a <= dmax; ++a) {
            x1p[a] = x1p[a - 1] * xi1;
            x2p[a] = x2p[a - 1] * xi2;
        }
        for (int ni2 = 0; ni2 < 4; ++ni2) {
            int d = sp->need0[ni2];
            if (d >= 0)
                for (int a = 0; a <= d; ++a)
                    for (int b = 0; b <= d - a; ++b) {
                        double f = wq * x1p[a] * x2p[b] * Rn;
                        for (int k = 0; k < sp->K; ++k) W[ni2][k][a][b] += f * sp->q_N[q * sp->K + k];
                    }
            Rn = Rn * r2i;
        }
    }
}

/* the weighted tables of one pair; returns the signed height z: tri_nodal.py:683-709 */
static double pair_tables(const double *tv, const double *geo, const double *obs, double eps,
                          const nspec *sp, double W[4][MAXK][5][5])
{
    double dx = obs[0] - geo[9], dy = obs[1] - geo[10], dz = obs[2] - geo[11];
    double z = dx * geo[6] + dy * geo[7] + dz * geo[8];
    double X1 = dx * geo[0] + dy * geo[1] + dz * geo[2];
    double X2 = dx * geo[3] + dy * geo[4] + dz * geo[5];
    double h2 = z * z + eps * eps;
    double D = sqrt((dx * dx + dy * dy + dz * dz) + eps * eps);
    double L = geo[27];
    if (D <= sp->dstar * L) {
        double C[MAXK][3][3];
        double M[5][MAX_TABLE_DEG + 1][MAX_TABLE_DEG + 1];
        memset(M, 0, sizeof(M));
        shape_coeffs(sp, geo, X1, X2, C);
        moment_table(tv, geo, X1, X2, h2, sp, M);
        weighted_tables(sp, C, M, W);
    } else if (D > sp->dstar_far * L) {
        far_tables(sp, geo, X1, X2, h2, sp->q_split, sp->nq, W);
    } else {
        far_tables(sp, geo, X1, X2, h2, 0, sp->q_split, W);
    }
    return z;
}

/* ---- per-node kernels: tri_nodal.py:732-893 ------------------------- */
static void t_blocks_node(double W[4][MAXK][5][5], int k, double z, const double *bas,
                          const double *nsrc, double eps, double *out /* 6 x 9 */)
{
    const double *W3 = &W[1][k][0][0], *W5 = &W[2][k][0][0];
    const double *ex = bas, *ey = bas + 3, *nhat = bas + 6;
    double V3[3], V5[3], T35[27], Pt[27];
    for (int m = 0; m < 3; ++m) {
        V3[m] = -ex[m] * W3[1 * 5 + 0] - ey[m] * W3[0 * 5 + 1] + z * nhat[m] * W3[0];
        V5[m] = -ex[m] * W5[1 * 5 + 0] - ey[m] * W5[0 * 5 + 1] + z * nhat[m] * W5[0];
    }
    build_T3(W5, 5, bas, z, T35);
    double e2 = eps * eps;
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

static void u_blocks_node(double W[4][MAXK][5][5], int k, double z, const double *bas, double eps,
                          double *out /* 3 x 9 */)
{
    double I1 = W[0][k][0][0], I3 = W[1][k][0][0];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            out[0 * 9 + i * 3 + j] = 0.0;
            out[1 * 9 + i * 3 + j] = 0.0;
        }
        out[0 * 9 + i * 3 + i] = I1;
        out[1 * 9 + i * 3 + i] = eps * eps * I3;
    }
    build_T2(&W[1][k][0][0], 5, bas, bas + 3, bas + 6, z, out + 18);
}

static void dd_stress_node(double W[4][MAXK][5][5], int k, double z, const double *bas,
                           const double *nrm, double eps, double mu, double nu, double lam,
                           double *H /* 27: H[m][n][j] */)
{
    double I3 = W[1][k][0][0], I5 = W[2][k][0][0];
    double T2_5[9], T2_7[9], T4_7[81], ID2G[81], B[27];
    build_T2(&W[2][k][0][0], 5, bas, bas + 3, bas + 6, z, T2_5);
    build_T2(&W[3][k][0][0], 5, bas, bas + 3, bas + 6, z, T2_7);
    build_T4(&W[3][k][0][0], 5, bas, z, T4_7);
    double C1 = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
    double c34 = 3.0 - 4.0 * nu;
    double cblob = 2.0 * (1.0 - nu) * eps * eps;
    for (int r = 0; r < 3; ++r)
        for (int p = 0; p < 3; ++p) {
            double drp = r == p ? 1.0 : 0.0;
            for (int s = 0; s < 3; ++s) {
                double drs = r == s ? 1.0 : 0.0, dps = p == s ? 1.0 : 0.0;
                for (int q = 0; q < 3; ++q) {
                    double dsq = s == q ? 1.0 : 0.0, dpq = p == q ? 1.0 : 0.0, drq = r == q ? 1.0 : 0.0;
                    double val = (-c34 * drp * dsq * I3 + drs * dpq * I3 + dps * drq * I3
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
        double trB = B[0 * 9 + 0 * 3 + kk] + B[1 * 9 + 1 * 3 + kk] + B[2 * 9 + 2 * 3 + kk];
        for (int m = 0; m < 3; ++m)
            for (int n = 0; n < 3; ++n) {
                double dmn = m == n ? 1.0 : 0.0;
                H[m * 9 + n * 3 + kk] = -(lam * dmn * trB + mu * (B[m * 9 + n * 3 + kk] + B[n * 9 + m * 3 + kk]));
            }
    }
}

static void kelvin_stress_node(double W[4][MAXK][5][5], int k, double z, const double *bas,
                               double eps, double mu, double nu, double lam, double *S /* 27 */)
{
    const double *W3 = &W[1][k][0][0], *W5 = &W[2][k][0][0];
    const double *ex = bas, *ey = bas + 3, *nhat = bas + 6;
    double V3[3], V5[3], T35[27], G1[27];
    for (int m = 0; m < 3; ++m) {
        V3[m] = -ex[m] * W3[1 * 5 + 0] - ey[m] * W3[0 * 5 + 1] + z * nhat[m] * W3[0];
        V5[m] = -ex[m] * W5[1 * 5 + 0] - ey[m] * W5[0 * 5 + 1] + z * nhat[m] * W5[0];
    }
    build_T3(W5, 5, bas, z, T35);
    double C1 = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
    double c34 = 3.0 - 4.0 * nu;
    double cbdg = 6.0 * (1.0 - nu) * eps * eps;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double dij = i == j ? 1.0 : 0.0;
            for (int m = 0; m < 3; ++m) {
                double dim = i == m ? 1.0 : 0.0, djm = j == m ? 1.0 : 0.0;
                G1[i * 9 + j * 3 + m] = C1 * (-c34 * dij * V3[m] + dim * V3[j] + djm * V3[i]
                                              - 3.0 * T35[i * 9 + j * 3 + m] - cbdg * dij * V5[m]);
            }
        }
    for (int c = 0; c < 3; ++c) {
        double trace_c = G1[0 * 9 + c * 3 + 0] + G1[1 * 9 + c * 3 + 1] + G1[2 * 9 + c * 3 + 2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double dij = i == j ? 1.0 : 0.0;
                S[i * 9 + j * 3 + c] = lam * dij * trace_c + mu * G1[i * 9 + c * 3 + j] + mu * G1[j * 9 + c * 3 + i];
            }
    }
}

static void eigen_node(double W[4][MAXK][5][5], int k, const double *nrm, double eps, double mu,
                       double lam, double *H)
{
    double phi = (15.0 * eps * eps * eps * eps / (8.0 * M_PI)) * W[3][k][0][0];
    for (int m = 0; m < 3; ++m)
        for (int n = 0; n < 3; ++n) {
            double dmn = m == n ? 1.0 : 0.0;
            for (int j = 0; j < 3; ++j) {
                double dmj = m == j ? 1.0 : 0.0, dnj = n == j ? 1.0 : 0.0;
                H[m * 9 + n * 3 + j] = phi * (lam * dmn * nrm[j] + mu * (dmj * nrm[n] + dnj * nrm[m]));
            }
        }
}

#define SPEC_ARGS const int64_t *ispec, const double *fspec, const int64_t *lattice, \
    const double *binom, int64_t n_terms, const double *q_lam, const double *q_w,     \
    const double *q_N, int64_t nq
#define SPEC_UNPACK nspec sp; unpack_spec(&sp, ispec, fspec, lattice, binom, n_terms, q_lam, q_w, q_N, nq)

/* ---- drivers: tri_nodal.py:896-1227 ---------------------------------- */

/* U matrices: basis stack (3, 3 nf, 3 K ns) when coeffs is NULL, else the
 * direct matrix (3 nf, 3 K ns) with u_coeffs applied (:929-955, :988-1043) */
void ornd_u_matrices(const double *x_field, int64_t nf, const double *tri_verts, int64_t ns,
                     const double *eps_arr, SPEC_ARGS, const double *coeffs, double *out)
{
    SPEC_UNPACK;
    int K = sp.K;
    int64_t ncol = 3 * K * ns, nrow = 3 * nf;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t s = 0; s < ns; ++s) {
        double geo[GEO_LEN], W[4][MAXK][5][5], blk[27];
        const double *tv = tri_verts + 9 * s;
        if (!nodal_frame(tv, geo)) continue;
        for (int64_t f = 0; f < nf; ++f) {
            double z = pair_tables(tv, geo, x_field + 3 * f, eps_arr[s], &sp, W);
            for (int k = 0; k < K; ++k) {
                u_blocks_node(W, k, z, geo, eps_arr[s], blk);
                int64_t col = 3 * (K * s + k);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        if (coeffs)
                            out[(3 * f + i) * ncol + col + j] = (coeffs[0] * blk[i * 3 + j] + coeffs[1] * blk[9 + i * 3 + j])
                                                                + coeffs[2] * blk[18 + i * 3 + j];
                        else
                            for (int b = 0; b < 3; ++b)
                                out[(b * nrow + 3 * f + i) * ncol + col + j] = blk[b * 9 + i * 3 + j];
                    }
            }
        }
    }
}

/* T matrices: basis stack (6, 3 nf, 3 K ns) or the direct matrix (:896-926) */
void ornd_t_matrices(const double *x_field, int64_t nf, const double *tri_verts,
                     const double *normals, int64_t ns, const double *eps_arr, SPEC_ARGS,
                     const double *coeffs, double *out)
{
    SPEC_UNPACK;
    int K = sp.K;
    int64_t ncol = 3 * K * ns, nrow = 3 * nf;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t s = 0; s < ns; ++s) {
        double geo[GEO_LEN], W[4][MAXK][5][5], blk[54];
        const double *tv = tri_verts + 9 * s;
        if (!nodal_frame(tv, geo)) continue;
        for (int64_t f = 0; f < nf; ++f) {
            double z = pair_tables(tv, geo, x_field + 3 * f, eps_arr[s], &sp, W);
            for (int k = 0; k < K; ++k) {
                t_blocks_node(W, k, z, geo, normals + 3 * s, eps_arr[s], blk);
                int64_t col = 3 * (K * s + k);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        if (coeffs) {
                            double acc = 0.0;
                            for (int b = 0; b < 6; ++b) acc += coeffs[b] * blk[b * 9 + i * 3 + j];
                            out[(3 * f + i) * ncol + col + j] = acc;
                        } else {
                            for (int b = 0; b < 6; ++b)
                                out[(b * nrow + 3 * f + i) * ncol + col + j] = blk[b * 9 + i * 3 + j];
                        }
                    }
            }
        }
    }
}

static int block_is_zero(const double *density, int K, int64_t s)
{
    for (int k = 0; k < K; ++k)
        for (int j = 0; j < 3; ++j)
            if (density[(K * s + k) * 3 + j] != 0.0) return 0;
    return 1;
}

/* displacement (nf, 3) of a nodal density (K ns, 3): kind 0 = U (force,
 * coeffs g1..g3), 1 = T (slip, coeffs c1..c6): tri_nodal.py:1055-1128 */
void ornd_disp_contract(const double *x_field, int64_t nf, const double *tri_verts,
                        const double *normals, int64_t ns, const double *eps_arr, SPEC_ARGS,
                        const double *density, const double *coeffs, int kind, double *out)
{
    SPEC_UNPACK;
    int K = sp.K;
    double *GEO = (double *)malloc(sizeof(double) * GEO_LEN * (ns > 0 ? ns : 1));
    int *OK = (int *)malloc(sizeof(int) * (ns > 0 ? ns : 1));
    for (int64_t s = 0; s < ns; ++s) OK[s] = nodal_frame(tri_verts + 9 * s, GEO + GEO_LEN * s);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t f = 0; f < nf; ++f) {
        double W[4][MAXK][5][5], blk[54];
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int64_t s = 0; s < ns; ++s) {
            if (!OK[s] || block_is_zero(density, K, s)) continue;
            const double *geo = GEO + GEO_LEN * s;
            double z = pair_tables(tri_verts + 9 * s, geo, x_field + 3 * f, eps_arr[s], &sp, W);
            for (int k = 0; k < K; ++k) {
                if (kind == 1)
                    t_blocks_node(W, k, z, geo, normals + 3 * s, eps_arr[s], blk);
                else
                    u_blocks_node(W, k, z, geo, eps_arr[s], blk);
                for (int j = 0; j < 3; ++j) {
                    double dj = density[(K * s + k) * 3 + j];
                    if (dj == 0.0) continue;
                    if (kind == 1) {
                        for (int i = 0; i < 3; ++i) {
                            double v = 0.0;
                            for (int b = 0; b < 6; ++b) v += coeffs[b] * blk[b * 9 + i * 3 + j];
                            if (i == 0) a0 += v * dj;
                            else if (i == 1) a1 += v * dj;
                            else a2 += v * dj;
                        }
                    } else {
                        a0 += (coeffs[0] * blk[0 * 3 + j] + coeffs[1] * blk[9 + 0 * 3 + j] + coeffs[2] * blk[18 + 0 * 3 + j]) * dj;
                        a1 += (coeffs[0] * blk[1 * 3 + j] + coeffs[1] * blk[9 + 1 * 3 + j] + coeffs[2] * blk[18 + 1 * 3 + j]) * dj;
                        a2 += (coeffs[0] * blk[2 * 3 + j] + coeffs[1] * blk[9 + 2 * 3 + j] + coeffs[2] * blk[18 + 2 * 3 + j]) * dj;
                    }
                }
            }
        }
        out[3 * f] = a0;
        out[3 * f + 1] = a1;
        out[3 * f + 2] = a2;
    }
    free(GEO);
    free(OK);
}

/* stress (nf, 3, 3) of a nodal density: which 0 = slip -> total stress,
 * 1 = force -> stress, 2 = eigenstress: tri_nodal.py:1131-1172 */
void ornd_stress_contract(const double *x_field, int64_t nf, const double *tri_verts,
                          const double *normals, int64_t ns, const double *eps_arr, SPEC_ARGS,
                          const double *density, double mu, double nu, double lam, int which,
                          double *sig)
{
    SPEC_UNPACK;
    int K = sp.K;
    double *GEO = (double *)malloc(sizeof(double) * GEO_LEN * (ns > 0 ? ns : 1));
    int *OK = (int *)malloc(sizeof(int) * (ns > 0 ? ns : 1));
    for (int64_t s = 0; s < ns; ++s) OK[s] = nodal_frame(tri_verts + 9 * s, GEO + GEO_LEN * s);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t f = 0; f < nf; ++f) {
        double W[4][MAXK][5][5], H[27], acc[9] = {0};
        for (int64_t s = 0; s < ns; ++s) {
            if (!OK[s] || block_is_zero(density, K, s)) continue;
            const double *geo = GEO + GEO_LEN * s;
            double eps = eps_arr[s];
            double z = pair_tables(tri_verts + 9 * s, geo, x_field + 3 * f, eps, &sp, W);
            for (int k = 0; k < K; ++k) {
                if (which == 0)
                    dd_stress_node(W, k, z, geo, normals + 3 * s, eps, mu, nu, lam, H);
                else if (which == 1)
                    kelvin_stress_node(W, k, z, geo, eps, mu, nu, lam, H);
                else
                    eigen_node(W, k, normals + 3 * s, eps, mu, lam, H);
                const double *d = density + (K * s + k) * 3;
                for (int m = 0; m < 3; ++m)
                    for (int n = 0; n < 3; ++n)
                        acc[m * 3 + n] += H[m * 9 + n * 3 + 0] * d[0] + H[m * 9 + n * 3 + 1] * d[1] + H[m * 9 + n * 3 + 2] * d[2];
            }
        }
        for (int q = 0; q < 9; ++q) sig[9 * f + q] = acc[q];
    }
    free(GEO);
    free(OK);
}

/* traction rows (3 nf, 3 n_cols), columns scattered by ids (ns, K):
 * which 0 = slip -> elastic (total minus eigenstress), 1 = force,
 * 2 = eigenstress alone, 3 = slip -> total: tri_nodal.py:1175-1227 */
void ornd_traction_matrix(const double *x_field, const double *n_field, int64_t nf,
                          const double *tri_verts, const double *normals, int64_t ns,
                          const double *eps_arr, SPEC_ARGS, const int64_t *ids, int64_t n_cols,
                          double mu, double nu, double lam, int which, double *out)
{
    SPEC_UNPACK;
    int K = sp.K;
    double *GEO = (double *)malloc(sizeof(double) * GEO_LEN * (ns > 0 ? ns : 1));
    int *OK = (int *)malloc(sizeof(int) * (ns > 0 ? ns : 1));
    for (int64_t s = 0; s < ns; ++s) OK[s] = nodal_frame(tri_verts + 9 * s, GEO + GEO_LEN * s);
    int64_t ncol = 3 * n_cols;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t f = 0; f < nf; ++f) {
        double W[4][MAXK][5][5], H[27], HE[27];
        const double *obs = x_field + 3 * f;
        double nf0 = n_field[3 * f], nf1 = n_field[3 * f + 1], nf2 = n_field[3 * f + 2];
        for (int64_t s = 0; s < ns; ++s) {
            if (!OK[s]) continue;
            const double *geo = GEO + GEO_LEN * s;
            double eps = eps_arr[s];
            double z = pair_tables(tri_verts + 9 * s, geo, obs, eps, &sp, W);
            for (int k = 0; k < K; ++k) {
                if (which == 0) {
                    dd_stress_node(W, k, z, geo, normals + 3 * s, eps, mu, nu, lam, H);
                    eigen_node(W, k, normals + 3 * s, eps, mu, lam, HE);
                    for (int q = 0; q < 27; ++q) H[q] -= HE[q];
                } else if (which == 1) {
                    kelvin_stress_node(W, k, z, geo, eps, mu, nu, lam, H);
                } else if (which == 2) {
                    eigen_node(W, k, normals + 3 * s, eps, mu, lam, H);
                } else {
                    dd_stress_node(W, k, z, geo, normals + 3 * s, eps, mu, nu, lam, H);
                }
                int64_t col = 3 * ids[s * K + k];
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j)
                        out[(3 * f + i) * ncol + col + j] += nf0 * H[i * 9 + 0 * 3 + j] + nf1 * H[i * 9 + 1 * 3 + j]
                                                             + nf2 * H[i * 9 + 2 * 3 + j];
            }
        }
    }
    free(GEO);
    free(OK);
}

/* the weighted tables of one pair (tests: W (4, K, 5, 5) and z) */
double ornd_pair_tables(const double *tv, const double *obs, double eps, SPEC_ARGS, double *Wout)
{
    SPEC_UNPACK;
    double geo[GEO_LEN], W[4][MAXK][5][5];
    memset(W, 0, sizeof(W));
    if (!nodal_frame(tv, geo)) return NAN;
    double z = pair_tables(tv, geo, obs, eps, &sp, W);
    for (int a = 0; a < 4; ++a)
        for (int k = 0; k < sp.K; ++k)
            for (int i = 0; i < 25; ++i) Wout[(a * sp.K + k) * 25 + i] = (&W[a][k][0][0])[i];
    return z;
}
