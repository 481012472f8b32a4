// Closed-form P0 pair integrals of the mollified Kelvin kernel, gfx950 FP64.
//
// One (observation point, source triangle) pair per call.  The edge
// antiderivatives, the van Oosterom solid angle and the 2D moment recursion
// follow mbem's src/mbem/kernels/tri_kernels.py in formula and guard order
// (:54-129, :160-221, :250-336, :606-769) because those are the cancelling
// terms.  What is built FROM the moments is restated in the source
// triangle's own orthonormal frame (ex, ey, nhat): there the displacement
// offset d = x - y has components (-xi1, -xi2, z), so every moment tensor
// (T2, T3, T4 of tri_kernels.py:772-853) is a signed, z-scaled 2D moment with
// no frame expansion at all, and each result is rotated back once.  The
// basis contractions N[P] / R[P] (tri_kernels.py:430-444) and the stress
// chains (:856-964) are written out in closed form over the unique tensor
// components, which the tensor symmetries make exact identities.
#pragma once
#include <hip/hip_runtime.h>

namespace mbem {

// ---- source record (structure of arrays, one column per source) ----------
// fields: v0 v1 v2 (9), ex (3), ey (3), nhat (3), normal (3), eps, ok, pad
enum SrcField : int {
    F_V0 = 0, F_V1 = 3, F_V2 = 6, F_EX = 9, F_EY = 12, F_NH = 15, F_N = 18,
    F_EPS = 21, F_OK = 22, NSF = 24
};

struct Src {
    double v[9];
    double ex[3], ey[3], nh[3], n[3];
    double eps;
    bool ok;
};

__device__ __forceinline__ void load_src(const double* __restrict__ src, long ld, long s, Src& S)
{
#pragma unroll
    for (int k = 0; k < 9; ++k) S.v[k] = src[(F_V0 + k) * ld + s];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        S.ex[k] = src[(F_EX + k) * ld + s];
        S.ey[k] = src[(F_EY + k) * ld + s];
        S.nh[k] = src[(F_NH + k) * ld + s];
        S.n[k] = src[(F_N + k) * ld + s];
    }
    S.eps = src[F_EPS * ld + s];
    S.ok = src[F_OK * ld + s] != 0.0;
}

// ---- reciprocal and square root -------------------------------------------
// The pair integrals divide by, and take roots of, positive lengths, heights
// and their powers, far from the denormal and overflow ranges, so the IEEE
// special-case paths of '/' and sqrt (div_scale / div_fmas / div_fixup, the
// range scaling of sqrt) are left out: the v_rcp_f64 / v_rsq_f64 seeds are
// refined by Newton steps (each squares the relative error), which leaves
// about one ulp.
__device__ __forceinline__ double frcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
// sq = sqrt(x) and is = 1/sqrt(x), x > 0: one Goldschmidt step on g = x y,
// h = y / 2 from the seed y, then one correction of each
__device__ __forceinline__ void frsq(double x, double& sq, double& is)
{
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    r = fma(-g, h, 0.5);
    h = fma(h, r, h);
    sq = g;
    is = 2.0 * h;
}
__device__ __forceinline__ double fsqrt(double x)
{
    double sq, is;
    frsq(x, sq, is);
    return sq;
}
constexpr double kThird = 1.0 / 3.0;
constexpr double kFifth = 0.2;

// ---- per-pair geometry: tri_kernels.py:163-182 ------------------------------
struct Geom {
    double z, h2, ih2, I3;
    double px[3], py[3];
};

__device__ __forceinline__ double solid_angle(const double* v, double ox, double oy, double oz)
{
    // van Oosterom: tri_kernels.py:79-96
    double r1x = v[0] - ox, r1y = v[1] - oy, r1z = v[2] - oz;
    double r2x = v[3] - ox, r2y = v[4] - oy, r2z = v[5] - oz;
    double r3x = v[6] - ox, r3y = v[7] - oy, r3z = v[8] - oz;
    // the lifted point is h >= eps > 0 off the plane: every R > 0 (the floor
    // only keeps a zero-width call finite)
    double R1 = fsqrt(fmax(r1x * r1x + r1y * r1y + r1z * r1z, 1e-300));
    double R2 = fsqrt(fmax(r2x * r2x + r2y * r2y + r2z * r2z, 1e-300));
    double R3 = fsqrt(fmax(r3x * r3x + r3y * r3y + r3z * r3z, 1e-300));
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

__device__ __forceinline__ void geom(const Src& S, double ox, double oy, double oz, Geom& g)
{
    double dx = ox - S.v[0], dy = oy - S.v[1], dz = oz - S.v[2];
    double z = dx * S.nh[0] + dy * S.nh[1] + dz * S.nh[2];
    double h2 = z * z + S.eps * S.eps;
    double h, ih;
    if (h2 > 0.0) {
        frsq(h2, h, ih);
    } else {
        h = 0.0;
        ih = 1.0 / h2;
    }
    double opx = ox - z * S.nh[0];
    double opy = oy - z * S.nh[1];
    double opz = oz - z * S.nh[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double wx = S.v[3 * k] - opx, wy = S.v[3 * k + 1] - opy, wz = S.v[3 * k + 2] - opz;
        g.px[k] = wx * S.ex[0] + wy * S.ex[1] + wz * S.ex[2];
        g.py[k] = wx * S.ey[0] + wy * S.ey[1] + wz * S.ey[2];
    }
    double Omega = solid_angle(S.v, opx + h * S.nh[0], opy + h * S.nh[1], opz + h * S.nh[2]);
    g.I3 = -Omega * ih;
    g.z = z;
    g.h2 = h2;
    g.ih2 = ih * ih;
}

// One edge in the edge-local frame (tri_kernels.py:190-201): outward normal
// (nx, ny), tangent (tx, ty), foot distance dp, endpoint abscissae ua, ub.
struct Edge {
    double nx, ny, tx, ty, dp, ua, ub, rho2;
    bool live;
};

__device__ __forceinline__ void edge(const Geom& g, int e, Edge& E)
{
    int e1 = (e == 2) ? 0 : e + 1;
    double ax = g.px[e], ay = g.py[e], bx = g.px[e1], by = g.py[e1];
    double evx = bx - ax, evy = by - ay;
    double L2 = evx * evx + evy * evy;
    E.live = L2 > 1e-60;                 // L > 1e-30
    double L, iL;
    frsq(E.live ? L2 : 1.0, L, iL);
    E.tx = E.live ? evx * iL : evx;
    E.ty = E.live ? evy * iL : evy;
    E.nx = E.ty;
    E.ny = -E.tx;
    E.dp = ax * E.nx + ay * E.ny;
    E.ua = ax * E.tx + ay * E.ty;
    E.ub = bx * E.tx + by * E.ty;
    E.rho2 = E.dp * E.dp + g.h2;
}

// ---- edge antiderivatives: tri_kernels.py:54-76, :606-619 ----------------
// One edge's endpoint terms (tri_kernels.py:54-64, same guards):
// R = sqrt(u^2 + rho2) and 1/R at both ends; J1(b) - J1(a) as one log of the
// ratio of the two arguments; J3(b) - J3(a) = (ub/Rb - ua/Ra) / rho2 (0 when
// rho2 < 1e-60); K3(b) - K3(a) with K3 = -1/max(R, 1e-30).
struct Ends {
    double Ra, Rb, iRa, iRb, irho2, dJ1, dJ3, dK3;
};

__device__ __forceinline__ void ends(const Edge& E, Ends& P)
{
    frsq(fmax(E.ua * E.ua + E.rho2, 1e-300), P.Ra, P.iRa);
    frsq(fmax(E.ub * E.ub + E.rho2, 1e-300), P.Rb, P.iRb);
    P.dJ1 = log((E.ub + P.Rb + 1e-300) * frcp(E.ua + P.Ra + 1e-300));
    P.irho2 = (E.rho2 < 1e-60) ? 0.0 : frcp(E.rho2);
    P.dJ3 = (E.ub * P.iRb - E.ua * P.iRa) * P.irho2;
    P.dK3 = fmin(P.iRa, 1e30) - fmin(P.iRb, 1e30);
}

// J5(b) - J5(a), J5 = u (3 rho2 + 2 u^2) / (3 rho2^2 R^3), and K5(b) - K5(a),
// K5 = -1 / (3 R^3) (0 for R < 1e-30): tri_kernels.py:606-619
__device__ __forceinline__ void ends5(const Edge& E, const Ends& P, double& dJ5, double& dK5)
{
    double ia3 = P.iRa * P.iRa * P.iRa, ib3 = P.iRb * P.iRb * P.iRb;
    double c = kThird * P.irho2 * P.irho2;
    dJ5 = c * (E.ub * (3.0 * E.rho2 + 2.0 * E.ub * E.ub) * ib3
               - E.ua * (3.0 * E.rho2 + 2.0 * E.ua * E.ua) * ia3);
    double ka = (P.iRa > 1e30) ? 0.0 : kThird * ia3;
    double kb = (P.iRb > 1e30) ? 0.0 : kThird * ib3;
    dK5 = ka - kb;
}

// Edge integrals of xi1^a xi2^b / R^m (tri_kernels.py:99-129, :632-652)
// written out per (a, b); dJ, dK, dL, dM are the u^0..u^3 differences.
struct EI {
    double dp, nx, ny, tx, ty;
    __device__ __forceinline__ double e00(double dJ) const { return dJ; }
    __device__ __forceinline__ double e10(double dJ, double dK) const { return dp * nx * dJ + tx * dK; }
    __device__ __forceinline__ double e01(double dJ, double dK) const { return dp * ny * dJ + ty * dK; }
    __device__ __forceinline__ double e20(double dJ, double dK, double dL) const
    {
        return dp * dp * nx * nx * dJ + 2.0 * dp * nx * tx * dK + tx * tx * dL;
    }
    __device__ __forceinline__ double e11(double dJ, double dK, double dL) const
    {
        return dp * dp * nx * ny * dJ + dp * nx * ty * dK + tx * dp * ny * dK + tx * ty * dL;
    }
    __device__ __forceinline__ double e02(double dJ, double dK, double dL) const
    {
        return dp * dp * ny * ny * dJ + 2.0 * dp * ny * ty * dK + ty * ty * dL;
    }
    __device__ __forceinline__ double e30(double dJ, double dK, double dL, double dM) const
    {
        double a = dp * nx;
        return a * a * a * dJ + 3.0 * a * a * tx * dK + 3.0 * a * tx * tx * dL + tx * tx * tx * dM;
    }
    __device__ __forceinline__ double e21(double dJ, double dK, double dL, double dM) const
    {
        // (a + tx u)^2 (b + ty u), a = dp nx, b = dp ny
        double a = dp * nx, b = dp * ny;
        return a * a * b * dJ + (a * a * ty + 2.0 * a * tx * b) * dK
               + (2.0 * a * tx * ty + tx * tx * b) * dL + tx * tx * ty * dM;
    }
    __device__ __forceinline__ double e12(double dJ, double dK, double dL, double dM) const
    {
        double a = dp * nx, b = dp * ny;
        return a * b * b * dJ + (tx * b * b + 2.0 * a * b * ty) * dK
               + (2.0 * tx * b * ty + a * ty * ty) * dL + tx * ty * ty * dM;
    }
    __device__ __forceinline__ double e03(double dJ, double dK, double dL, double dM) const
    {
        double b = dp * ny;
        return b * b * b * dJ + 3.0 * b * b * ty * dK + 3.0 * b * ty * ty * dL + ty * ty * ty * dM;
    }
};

// ---- U kernel (force -> displacement): tri_kernels.py:160-243 -------------
// Local-frame results: I1, I3 and the n=3 moments; U = g1 I1 d + g2 eps^2 I3 d
// + g3 T2[3], T2 local = [[m20, m11, -z m10], [m11, m02, -z m01], [., ., z^2 I3]].
struct UMom {
    double I1, I3, z, m10, m01, m20, m11, m02;
};

__device__ __forceinline__ void u_moments(const Src& S, double ox, double oy, double oz, UMom& M)
{
    Geom g;
    geom(S, ox, oy, oz, g);
    double BD001 = 0.0, BN1_001 = 0.0, BN2_001 = 0.0, BN1_101 = 0.0, BN1_011 = 0.0, BN2_011 = 0.0;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        Edge E;
        edge(g, e, E);
        if (!E.live) continue;
        Ends P;
        ends(E, P);
        double dJ1 = P.dJ1;
        double dK1 = P.Rb - P.Ra;
        EI ei{E.dp, E.nx, E.ny, E.tx, E.ty};
        double v00 = ei.e00(dJ1);
        double v10 = ei.e10(dJ1, dK1);
        double v01 = ei.e01(dJ1, dK1);
        BD001 += E.dp * v00;
        BN1_001 += E.nx * v00; BN2_001 += E.ny * v00;
        BN1_101 += E.nx * v10;
        BN1_011 += E.nx * v01; BN2_011 += E.ny * v01;
    }
    M.I3 = g.I3;
    M.z = g.z;
    M.I1 = BD001 - g.h2 * g.I3;
    M.m10 = -BN1_001;
    M.m01 = -BN2_001;
    M.m20 = M.I1 - BN1_101;
    M.m11 = -BN1_011;
    M.m02 = M.I1 - BN2_011;
}

// rotate a symmetric local tensor L (l00 l01 l02 l11 l12 l22) to the global
// frame: out_ij = sum_ab Q_ia L_ab Q_jb with Q = [ex | ey | nhat]
__device__ __forceinline__ void rot_sym(const Src& S, const double* L, double* out)
{
    double Lf[9] = {L[0], L[1], L[2], L[1], L[3], L[4], L[2], L[4], L[5]};
    double W[9];  // W_ib = sum_a Q_ia L_ab
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int b = 0; b < 3; ++b) W[i * 3 + b] = S.ex[i] * Lf[0 * 3 + b] + S.ey[i] * Lf[1 * 3 + b] + S.nh[i] * Lf[2 * 3 + b];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out[i * 3 + j] = W[i * 3 + 0] * S.ex[j] + W[i * 3 + 1] * S.ey[j] + W[i * 3 + 2] * S.nh[j];
}

// U basis blocks (3 x 3x3, tri_kernels.py:223-243) into b[27]
__device__ __forceinline__ void u_basis(const Src& S, const UMom& M, double* b)
{
    double T2[6] = {M.m20, M.m11, -M.z * M.m10, M.m02, -M.z * M.m01, M.z * M.z * M.I3};
    double e2I3 = S.eps * S.eps * M.I3;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            b[0 * 9 + i * 3 + j] = (i == j) ? M.I1 : 0.0;
            b[1 * 9 + i * 3 + j] = (i == j) ? e2I3 : 0.0;
        }
    rot_sym(S, T2, b + 18);
}

// combined U block g1 B1 + g2 B2 + g3 B3 (tri_kernels.py:546-565) into u[9]
__device__ __forceinline__ void u_direct(const Src& S, const UMom& M, double g1, double g2, double g3, double* u)
{
    double T2[6] = {g3 * M.m20, g3 * M.m11, -g3 * M.z * M.m10, g3 * M.m02, -g3 * M.z * M.m01,
                    g3 * M.z * M.z * M.I3};
    double dg = g1 * M.I1 + g2 * S.eps * S.eps * M.I3;
    rot_sym(S, T2, u);
    u[0] += dg; u[4] += dg; u[8] += dg;
}

// ---- T kernel (slip -> displacement): tri_kernels.py:250-427 --------------
// Local-frame moments: V3, V5 (first moments of d / R^3, d / R^5), t_i =
// sum_m T3[5]_imm and Tn_ij = sum_m T3[5]_ijm n_m, with n the source normal.
struct TMom {
    double V3[3], V5[3], t[3], Tn[6];  // local frame; Tn packed 00 01 02 11 12 22
    double nl[3];                      // the source normal in the local frame
};

__device__ __forceinline__ void t_moments(const Src& S, double ox, double oy, double oz, TMom& T)
{
    Geom g;
    geom(S, ox, oy, oz, g);
    double BN1_001 = 0.0, BN2_001 = 0.0, BD003 = 0.0, BN1_003 = 0.0, BN2_003 = 0.0;
    double BN1_103 = 0.0, BN1_013 = 0.0, BN2_013 = 0.0, BN1_203 = 0.0, BN1_113 = 0.0;
    double BN1_023 = 0.0, BN2_023 = 0.0;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        Edge E;
        edge(g, e, E);
        if (!E.live) continue;
        Ends P;
        ends(E, P);
        double dJ1 = P.dJ1, dJ3 = P.dJ3, dK3 = P.dK3;
        double dL3 = dJ1 - E.rho2 * dJ3;
        EI ei{E.dp, E.nx, E.ny, E.tx, E.ty};
        double v00m1 = dJ1;
        double v00 = dJ3;
        double v10 = ei.e10(dJ3, dK3);
        double v01 = ei.e01(dJ3, dK3);
        double v20 = ei.e20(dJ3, dK3, dL3);
        double v11 = ei.e11(dJ3, dK3, dL3);
        double v02 = ei.e02(dJ3, dK3, dL3);
        BN1_001 += E.nx * v00m1; BN2_001 += E.ny * v00m1;
        BD003 += E.dp * v00;
        BN1_003 += E.nx * v00; BN2_003 += E.ny * v00;
        BN1_103 += E.nx * v10;
        BN1_013 += E.nx * v01; BN2_013 += E.ny * v01;
        BN1_203 += E.nx * v20;
        BN1_113 += E.nx * v11;
        BN1_023 += E.nx * v02; BN2_023 += E.ny * v02;
    }
    double I3 = g.I3, z = g.z;
    double I5 = (BD003 + I3) * kThird * g.ih2;
    double m10_3 = -BN1_001, m01_3 = -BN2_001;
    double m10 = -BN1_003 * kThird, m01 = -BN2_003 * kThird;
    double m20 = (I3 - BN1_103) * kThird;
    double m11 = -BN1_013 * kThird;
    double m02 = (I3 - BN2_013) * kThird;
    double m30 = (2.0 * m10_3 - BN1_203) * kThird;
    double m21 = (m01_3 - BN1_113) * kThird;
    double m12 = -BN1_023 * kThird;
    double m03 = (2.0 * m01_3 - BN2_023) * kThird;
    // local d = (-xi1, -xi2, z)
    T.V3[0] = -m10_3; T.V3[1] = -m01_3; T.V3[2] = z * I3;
    T.V5[0] = -m10;   T.V5[1] = -m01;   T.V5[2] = z * I5;
    double z2 = z * z;
    // t_i = int d_i |d|^2 / R^5, |d|^2 = xi1^2 + xi2^2 + z^2
    T.t[0] = -(m30 + m12 + z2 * m10);
    T.t[1] = -(m21 + m03 + z2 * m01);
    T.t[2] = z * (m20 + m02 + z2 * I5);
    // local normal components
    double a = S.ex[0] * S.n[0] + S.ex[1] * S.n[1] + S.ex[2] * S.n[2];
    double b = S.ey[0] * S.n[0] + S.ey[1] * S.n[1] + S.ey[2] * S.n[2];
    double c = S.nh[0] * S.n[0] + S.nh[1] * S.n[1] + S.nh[2] * S.n[2];
    T.nl[0] = a; T.nl[1] = b; T.nl[2] = c;
    // local T3[5] components T_ijk = int d_i d_j d_k / R^5 with d = (-xi1, -xi2, z)
    double T000 = -m30, T001 = -m21, T011 = -m12, T111 = -m03;
    double T002 = z * m20, T012 = z * m11, T112 = z * m02;
    double T022 = -z2 * m10, T122 = -z2 * m01, T222 = z2 * z * I5;
    // Tn_ij = sum_k T_ijk nl_k
    T.Tn[0] = T000 * a + T001 * b + T002 * c;
    T.Tn[1] = T001 * a + T011 * b + T012 * c;
    T.Tn[2] = T002 * a + T012 * b + T022 * c;
    T.Tn[3] = T011 * a + T111 * b + T112 * c;
    T.Tn[4] = T012 * a + T112 * b + T122 * c;
    T.Tn[5] = T022 * a + T122 * b + T222 * c;
}

// general (non-symmetric) local 3x3 -> global: out = Q L Q^T
__device__ __forceinline__ void rot_gen(const Src& S, const double* L, double* out)
{
    double W[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int b = 0; b < 3; ++b) W[i * 3 + b] = S.ex[i] * L[0 * 3 + b] + S.ey[i] * L[1 * 3 + b] + S.nh[i] * L[2 * 3 + b];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) out[i * 3 + j] = W[i * 3 + 0] * S.ex[j] + W[i * 3 + 1] * S.ey[j] + W[i * 3 + 2] * S.nh[j];
}

__device__ __forceinline__ double tn_at(const double* Tn, int i, int j)
{
    // packed symmetric 00 01 02 11 12 22
    int lo = i < j ? i : j, hi = i < j ? j : i;
    return (lo == 0) ? Tn[hi] : (lo == 1 ? Tn[2 + hi] : Tn[5]);
}

// the six T basis blocks [N1 N2 N3 R1 R2 R3] (tri_kernels.py:396-427), local:
//   N1 = n_j V3_i, N2 = n_j (4 V3_i - 3 t_i), N3 = n_j eps^2 V5_i,
//   R1 = d_ij nV3 + n_i V3_j, R2 = n_i V3_j + 2 n_j V3_i + d_ij nV3 - 6 Tn_ij,
//   R3 = eps^2 (d_ij nV5 + n_i V5_j)
// combined with coefficients c[0..5] into the local 3x3 L.
__device__ __forceinline__ void t_combine_local(const TMom& T, double eps, const double* c, double* L)
{
    const double* n = T.nl;
    double e2 = eps * eps;
    double nV3 = n[0] * T.V3[0] + n[1] * T.V3[1] + n[2] * T.V3[2];
    double nV5 = n[0] * T.V5[0] + n[1] * T.V5[1] + n[2] * T.V5[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double trN = c[0] * T.V3[i] + c[1] * (4.0 * T.V3[i] - 3.0 * T.t[i]) + c[2] * e2 * T.V5[i];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double dij = (i == j) ? 1.0 : 0.0;
            double R1 = dij * nV3 + n[i] * T.V3[j];
            double R2 = n[i] * T.V3[j] + 2.0 * n[j] * T.V3[i] + dij * nV3 - 6.0 * tn_at(T.Tn, i, j);
            double R3 = e2 * (dij * nV5 + n[i] * T.V5[j]);
            L[i * 3 + j] = n[j] * trN + c[3] * R1 + c[4] * R2 + c[5] * R3;
        }
    }
}

__device__ __forceinline__ void t_direct(const Src& S, const TMom& T, const double* c, double* u)
{
    double L[9];
    t_combine_local(T, S.eps, c, L);
    rot_gen(S, L, u);
}

// ---- stress kernels: tri_kernels.py:655-769, :856-964, :1194-1214 ----------
// Local-frame n=3/5/7 moments up to the orders the stress chains consume.
struct SMom {
    double z, I3, I5, I7;
    double m3_10, m3_01;                                   // n = 3
    double m5[10];  // n = 5: 00 10 01 20 11 02 30 21 12 03
    double m7[15];  // n = 7: 00 10 01 20 11 02 30 21 12 03 40 31 22 13 04
};

template <bool NEED7>
__device__ __forceinline__ void s_moments(const Src& S, double ox, double oy, double oz, SMom& M)
{
    Geom g;
    geom(S, ox, oy, oz, g);
    double BN1m1 = 0.0, BN2m1 = 0.0, BD003 = 0.0, BD005 = 0.0;
    // n = 3 edge sums needed: bn1_3 (00 10 01 20 11 02), bn2_3 (00 01 02)
    double b1_00 = 0, b1_10 = 0, b1_01 = 0, b1_20 = 0, b1_11 = 0, b1_02 = 0;
    double b2_00 = 0, b2_01 = 0, b2_02 = 0;
    // n = 5 edge sums: bn1_5 (00 10 01 20 11 02 30 21 12 03), bn2_5 (00 01 02 03)
    double c1_00 = 0, c1_10 = 0, c1_01 = 0, c1_20 = 0, c1_11 = 0, c1_02 = 0, c1_30 = 0, c1_21 = 0,
           c1_12 = 0, c1_03 = 0;
    double c2_00 = 0, c2_01 = 0, c2_02 = 0, c2_03 = 0;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        Edge E;
        edge(g, e, E);
        if (!E.live) continue;
        Ends P;
        ends(E, P);
        double dJ1 = P.dJ1, dJ3 = P.dJ3, dK3 = P.dK3;
        double dL3 = dJ1 - E.rho2 * dJ3;
        EI ei{E.dp, E.nx, E.ny, E.tx, E.ty};
        BN1m1 += E.nx * dJ1; BN2m1 += E.ny * dJ1;
        double v = dJ3;
        b1_00 += E.nx * v; b2_00 += E.ny * v; BD003 += E.dp * v;
        v = ei.e01(dJ3, dK3); b1_01 += E.nx * v; b2_01 += E.ny * v;
        v = ei.e02(dJ3, dK3, dL3); b1_02 += E.nx * v; b2_02 += E.ny * v;
        v = ei.e10(dJ3, dK3); b1_10 += E.nx * v;
        v = ei.e11(dJ3, dK3, dL3); b1_11 += E.nx * v;
        v = ei.e20(dJ3, dK3, dL3); b1_20 += E.nx * v;
        if (NEED7) {
            double dJ5, dK5;
            ends5(E, P, dJ5, dK5);
            double dL5 = dJ3 - E.rho2 * dJ5;
            double dM5 = dK3 - E.rho2 * dK5;
            v = dJ5; c1_00 += E.nx * v; c2_00 += E.ny * v; BD005 += E.dp * v;
            v = ei.e01(dJ5, dK5); c1_01 += E.nx * v; c2_01 += E.ny * v;
            v = ei.e02(dJ5, dK5, dL5); c1_02 += E.nx * v; c2_02 += E.ny * v;
            v = ei.e03(dJ5, dK5, dL5, dM5); c1_03 += E.nx * v; c2_03 += E.ny * v;
            v = ei.e10(dJ5, dK5); c1_10 += E.nx * v;
            v = ei.e11(dJ5, dK5, dL5); c1_11 += E.nx * v;
            v = ei.e12(dJ5, dK5, dL5, dM5); c1_12 += E.nx * v;
            v = ei.e20(dJ5, dK5, dL5); c1_20 += E.nx * v;
            v = ei.e21(dJ5, dK5, dL5, dM5); c1_21 += E.nx * v;
            v = ei.e30(dJ5, dK5, dL5, dM5); c1_30 += E.nx * v;
        }
    }
    double h2 = g.h2, I3 = g.I3;
    double I5 = (h2 > 1e-60) ? (BD003 + I3) * kThird * g.ih2 : 0.0;
    M.z = g.z;
    M.I3 = I3;
    M.I5 = I5;
    M.m3_10 = -BN1m1;
    M.m3_01 = -BN2m1;
    double* m5 = M.m5;
    m5[0] = I5;
    m5[1] = -b1_00 * kThird;
    m5[2] = -b2_00 * kThird;
    m5[3] = (I3 - b1_10) * kThird;
    m5[4] = -b1_01 * kThird;
    m5[5] = (I3 - b2_01) * kThird;
    m5[6] = (2.0 * M.m3_10 - b1_20) * kThird;
    m5[7] = (M.m3_01 - b1_11) * kThird;
    m5[8] = -b1_02 * kThird;
    m5[9] = (2.0 * M.m3_01 - b2_02) * kThird;
    M.I7 = 0.0;
    if (NEED7) {
        double I7 = (h2 > 1e-60) ? (BD005 + 3.0 * I5) * kFifth * g.ih2 : 0.0;
        M.I7 = I7;
        double* m7 = M.m7;
        m7[0] = I7;
        m7[1] = -c1_00 * kFifth;
        m7[2] = -c2_00 * kFifth;
        m7[3] = (I5 - c1_10) * kFifth;
        m7[4] = -c1_01 * kFifth;
        m7[5] = (I5 - c2_01) * kFifth;
        m7[6] = (2.0 * m5[1] - c1_20) * kFifth;
        m7[7] = (m5[2] - c1_11) * kFifth;
        m7[8] = -c1_02 * kFifth;
        m7[9] = (2.0 * m5[2] - c2_02) * kFifth;
        m7[10] = (3.0 * m5[3] - c1_30) * kFifth;
        m7[11] = (2.0 * m5[4] - c1_21) * kFifth;
        m7[12] = (m5[5] - c1_12) * kFifth;
        m7[13] = -c1_03 * kFifth;
        m7[14] = (3.0 * m5[5] - c2_03) * kFifth;
    }
}

// local symmetric 2-tensor int d_i d_j / R^n from moments (00 10 01 20 11 02)
__device__ __forceinline__ void loc_T2(const double* m, double z, double* T)
{
    // packed 00 01 02 11 12 22 with d = (-xi1, -xi2, z)
    T[0] = m[3];
    T[1] = m[4];
    T[2] = -z * m[1];
    T[3] = m[5];
    T[4] = -z * m[2];
    T[5] = z * z * m[0];
}

__device__ __forceinline__ double sym2(const double* T, int i, int j)
{
    int lo = i < j ? i : j, hi = i < j ? j : i;
    return (lo == 0) ? T[hi] : (lo == 1 ? T[2 + hi] : T[5]);
}

// local rank-4 T4[7]_ijkl = int d_i d_j d_k d_l / R^7: the value depends only
// on how many indices are 0 (a), 1 (b) and 2 (c = 4 - a - b)
__device__ __forceinline__ double t4_local(const double* m7, double z, int a, int b)
{
    // moment M(a, b) index in the packed n=7 list
    int idx;
    int s = a + b;
    if (s == 0) idx = 0;
    else if (s == 1) idx = (a == 1) ? 1 : 2;
    else if (s == 2) idx = 3 + (2 - a);
    else if (s == 3) idx = 6 + (3 - a);
    else idx = 10 + (4 - a);
    int c = 4 - s;
    double zc = (c == 0) ? 1.0 : (c == 1 ? z : (c == 2 ? z * z : (c == 3 ? z * z * z : z * z * z * z)));
    double sg = (s & 1) ? -1.0 : 1.0;
    return sg * zc * m7[idx];
}

// Kelvin force -> stress, local frame: S[i,j,k] = lam d_ij tr_k + mu (G1_ikj + G1_jki)
// with G1_ijm = C1(-c34 d_ij V3_m + d_im V3_j + d_jm V3_i - 3 T3_ijm - cbdg d_ij V5_m)
// (tri_kernels.py:856-897), contracted with the local force f: sig_ij = S_ijk f_k.
__device__ __forceinline__ void kelvin_local(const SMom& M, double eps, double mu, double nu, double lam,
                                             const double* f, double* sig /*local 6 packed*/)
{
    double z = M.z;
    double V3[3] = {-M.m3_10, -M.m3_01, z * M.I3};
    double V5[3] = {-M.m5[1], -M.m5[2], z * M.m5[0]};
    const double* m = M.m5;
    double z2 = z * z;
    // local T3[5] components
    double T000 = -m[6], T001 = -m[7], T011 = -m[8], T111 = -m[9];
    double T002 = z * m[3], T012 = z * m[4], T112 = z * m[5];
    double T022 = -z2 * m[1], T122 = -z2 * m[2], T222 = z2 * z * m[0];
    // Tf_ij = sum_k T_ijk f_k
    double Tf00 = T000 * f[0] + T001 * f[1] + T002 * f[2];
    double Tf01 = T001 * f[0] + T011 * f[1] + T012 * f[2];
    double Tf02 = T002 * f[0] + T012 * f[1] + T022 * f[2];
    double Tf11 = T011 * f[0] + T111 * f[1] + T112 * f[2];
    double Tf12 = T012 * f[0] + T112 * f[1] + T122 * f[2];
    double Tf22 = T022 * f[0] + T122 * f[1] + T222 * f[2];
    double C1 = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
    double c34 = 3.0 - 4.0 * nu;
    double cbdg = 6.0 * (1.0 - nu) * eps * eps;
    // Gf_ij := sum_k G1_ikj f_k = C1(-c34 f_i V3_j + d_ij (V3.f) + f_j V3_i - 3 Tf_ij - cbdg f_i V5_j)
    double V3f = V3[0] * f[0] + V3[1] * f[1] + V3[2] * f[2];
    // trace_k f_k = sum_k sum_m G1_mkm f_k = C1(-c34 V3.f + V3.f + 3 V3.f - 3 t.f - cbdg V5.f)
    // with t_k = sum_m T3_mkm
    double tf = (T000 + T011 + T022) * f[0] + (T001 + T111 + T122) * f[1] + (T002 + T112 + T222) * f[2];
    double V5f = V5[0] * f[0] + V5[1] * f[1] + V5[2] * f[2];
    double trf = C1 * ((4.0 - c34) * V3f - 3.0 * tf - cbdg * V5f);
    double Tf[6] = {Tf00, Tf01, Tf02, Tf11, Tf12, Tf22};
    int q = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            double dij = (i == j) ? 1.0 : 0.0;
            // G1_ikj f_k + G1_jki f_k: the two non-symmetric parts add symmetric
            double Gij = -c34 * f[i] * V3[j] + dij * V3f + f[j] * V3[i] - 3.0 * Tf[q] - cbdg * f[i] * V5[j];
            double Gji = -c34 * f[j] * V3[i] + dij * V3f + f[i] * V3[j] - 3.0 * Tf[q] - cbdg * f[j] * V5[i];
            sig[q] = lam * dij * trf + mu * C1 * (Gij + Gji);
            ++q;
        }
}

// Dislocation slip -> total stress, local frame (tri_kernels.py:900-964):
// the ID2G chain contracted in closed form over the unique components.
__device__ __forceinline__ void dd_local(const SMom& M, double eps, double mu, double nu, double lam,
                                         const double* nl, const double* bk, double* sig /*6 packed*/)
{
    double z = M.z;
    double T5[6], T27[6];
    loc_T2(M.m5, z, T5);
    loc_T2(M.m7, z, T27);
    double I3 = M.I3, I5 = M.I5;
    double C1 = 1.0 / (16.0 * M_PI * mu * (1.0 - nu));
    double c34 = 3.0 - 4.0 * nu;
    double cblob = 2.0 * (1.0 - nu) * eps * eps;
    double trT5 = T5[0] + T5[3] + T5[5];
    // W_rs = sum_p T4_rspp and T7n_abc = sum_q T4_abcq n_q (local)
    double T4[3][3][3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    int a = (i == 0) + (j == 0) + (k == 0) + (l == 0);
                    int b = (i == 1) + (j == 1) + (k == 1) + (l == 1);
                    T4[i][j][k][l] = t4_local(M.m7, z, a, b);
                }
    double T5n[3], T27n[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T5n[i] = sym2(T5, i, 0) * nl[0] + sym2(T5, i, 1) * nl[1] + sym2(T5, i, 2) * nl[2];
        T27n[i] = sym2(T27, i, 0) * nl[0] + sym2(T27, i, 1) * nl[1] + sym2(T27, i, 2) * nl[2];
    }
    // B[r,s,k] b_k contracted with the slip: Bb_rs = sum_k B[r,s,k] b_k
    //   B = C1 [lam n_k tr_rs + mu (nq_rsk + npd_rsk)]
    double nb = nl[0] * bk[0] + nl[1] * bk[1] + nl[2] * bk[2];
    double T5b[3], T27b[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        T5b[i] = sym2(T5, i, 0) * bk[0] + sym2(T5, i, 1) * bk[1] + sym2(T5, i, 2) * bk[2];
        T27b[i] = sym2(T27, i, 0) * bk[0] + sym2(T27, i, 1) * bk[1] + sym2(T27, i, 2) * bk[2];
    }
    double T5nb = T5n[0] * bk[0] + T5n[1] * bk[1] + T5n[2] * bk[2];
    double Bb[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            double drs = (r == s) ? 1.0 : 0.0;
            double W = T4[r][s][0][0] + T4[r][s][1][1] + T4[r][s][2][2];
            double tr = drs * I3 * (4.0 - c34) - 3.0 * drs * trT5 + (3.0 * c34 - 18.0) * sym2(T5, r, s)
                        + 15.0 * W + cblob * (15.0 * sym2(T27, r, s) - 3.0 * drs * I5);
            // T7nb_rs = sum_k (T7 n)_rsk b_k = sum_kq T4_rskq n_q b_k (symmetric in r,s)
            double T7nb = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int qq = 0; qq < 3; ++qq) T7nb += T4[r][s][k][qq] * nl[qq] * bk[k];
            // sum_k nq_rsk b_k
            double nqb = -c34 * bk[r] * nl[s] * I3 + drs * nb * I3 + bk[s] * nl[r] * I3
                         + 3.0 * c34 * bk[r] * T5n[s] - 3.0 * drs * T5nb - 3.0 * bk[s] * T5n[r]
                         - 3.0 * nl[r] * T5b[s] - 3.0 * nb * sym2(T5, r, s) - 3.0 * nl[s] * T5b[r]
                         + 15.0 * T7nb + cblob * bk[r] * (-3.0 * nl[s] * I5 + 15.0 * T27n[s]);
            // sum_k npd_rsk b_k
            double npb = -c34 * nl[r] * bk[s] * I3 + drs * nb * I3 + nl[s] * bk[r] * I3
                         + 3.0 * c34 * nl[r] * T5b[s] - 3.0 * drs * T5nb - 3.0 * nl[s] * T5b[r]
                         - 3.0 * bk[r] * T5n[s] - 3.0 * nb * sym2(T5, r, s) - 3.0 * bk[s] * T5n[r]
                         + 15.0 * T7nb + cblob * nl[r] * (-3.0 * bk[s] * I5 + 15.0 * T27b[s]);
            Bb[r][s] = C1 * (lam * nb * tr + mu * (nqb + npb));
        }
    double trB = Bb[0][0] + Bb[1][1] + Bb[2][2];
    int q = 0;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int n = m; n < 3; ++n) {
            double dmn = (m == n) ? 1.0 : 0.0;
            sig[q++] = -(lam * dmn * trB + mu * (Bb[m][n] + Bb[n][m]));
        }
}

// Eigenstress +C:eps* (tri_kernels.py:1194-1214), global frame: the
// triangle-integrated blob Phi = 15 eps^4 / (8 pi) I7.
__device__ __forceinline__ void eigen_global(double I7, double eps, double mu, double lam, const double* n,
                                             const double* bk, double* sig /*6 packed*/)
{
    double phi = (15.0 * eps * eps * eps * eps / (8.0 * M_PI)) * I7;
    double nb = n[0] * bk[0] + n[1] * bk[1] + n[2] * bk[2];
    int q = 0;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int nn = m; nn < 3; ++nn) {
            double dmn = (m == nn) ? 1.0 : 0.0;
            sig[q++] = phi * (lam * dmn * nb + mu * (bk[m] * n[nn] + bk[nn] * n[m]));
        }
}

// rotate a packed symmetric local stress to the global frame and add it
__device__ __forceinline__ void add_rot_sym(const Src& S, const double* L, double* acc /*9*/)
{
    double G[9];
    rot_sym(S, L, G);
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] += G[q];
}

__device__ __forceinline__ void to_local(const Src& S, const double* v, double* l)
{
    l[0] = S.ex[0] * v[0] + S.ex[1] * v[1] + S.ex[2] * v[2];
    l[1] = S.ey[0] * v[0] + S.ey[1] * v[1] + S.ey[2] * v[2];
    l[2] = S.nh[0] * v[0] + S.nh[1] * v[1] + S.nh[2] * v[2];
}

}  // namespace mbem
