// ORACLE tool -- test/measurement infrastructure only.
//
// Algorithmic FLOPs per (observation point, source triangle) pair of each
// pair kernel, counted by running the oracle's own restatement of mbem's
// tri_kernels.py (tri_kernels_ref.c) with every double replaced by a
// counting type.  Counting rule: add, sub, mul, neg = 1; div, sqrt, log,
// atan2 = 1 each; comparisons and conversions = 0; the material combine of
// the *_matrix_direct drivers is included.  Prints one JSON object with the
// per-pair averages over a fixed sample of pairs (on, near and far from the
// element), per kernel and per category.
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace cnt {
unsigned long long add = 0, mul = 0, dv = 0, sq = 0, lg = 0, at = 0;
}

struct cdouble {
    double v;
    cdouble() : v(0.0) {}
    cdouble(double x) : v(x) {}
    cdouble(int x) : v(x) {}
    explicit operator double() const { return v; }
};
static inline cdouble operator+(cdouble a, cdouble b) { ++cnt::add; return cdouble(a.v + b.v); }
static inline cdouble operator-(cdouble a, cdouble b) { ++cnt::add; return cdouble(a.v - b.v); }
static inline cdouble operator*(cdouble a, cdouble b) { ++cnt::mul; return cdouble(a.v * b.v); }
static inline cdouble operator/(cdouble a, cdouble b) { ++cnt::dv; return cdouble(a.v / b.v); }
static inline cdouble operator-(cdouble a) { ++cnt::add; return cdouble(-a.v); }
static inline cdouble& operator+=(cdouble& a, cdouble b) { a = a + b; return a; }
static inline cdouble& operator-=(cdouble& a, cdouble b) { a = a - b; return a; }
static inline cdouble& operator*=(cdouble& a, cdouble b) { a = a * b; return a; }
static inline cdouble& operator/=(cdouble& a, cdouble b) { a = a / b; return a; }
static inline bool operator<(cdouble a, cdouble b) { return a.v < b.v; }
static inline bool operator>(cdouble a, cdouble b) { return a.v > b.v; }
static inline bool operator<=(cdouble a, cdouble b) { return a.v <= b.v; }
static inline bool operator>=(cdouble a, cdouble b) { return a.v >= b.v; }
static inline bool operator==(cdouble a, cdouble b) { return a.v == b.v; }
static inline bool operator!=(cdouble a, cdouble b) { return a.v != b.v; }
static inline cdouble csqrt(cdouble a) { ++cnt::sq; return cdouble(std::sqrt(a.v)); }
static inline cdouble clog(cdouble a) { ++cnt::lg; return cdouble(std::log(a.v)); }
static inline cdouble catan2(cdouble a, cdouble b) { ++cnt::at; return cdouble(std::atan2(a.v, b.v)); }

#define double cdouble
#define sqrt csqrt
#define log clog
#define atan2 catan2
#define memset(p, c, n) do { } while (0)
#define int64_t long
#define orc_frames orcx_frames
#define orc_u_basis_matrices orcx_u_basis_matrices
#define orc_t_basis_matrices orcx_t_basis_matrices
#define orc_u_matrix_direct orcx_u_matrix_direct
#define orc_t_matrix_direct orcx_t_matrix_direct
#define orc_u_disp_contract orcx_u_disp_contract
#define orc_t_disp_contract orcx_t_disp_contract
#define orc_kelvin_stress_contract orcx_kelvin_stress_contract
#define orc_dd_stress_contract orcx_dd_stress_contract
#define orc_eigenstress_contract orcx_eigenstress_contract
#define orc_moments orcx_moments
#include "tri_kernels_ref.c"
#undef double
#undef sqrt
#undef log
#undef atan2

static unsigned long long total() { return cnt::add + cnt::mul + cnt::dv + cnt::sq + cnt::lg + cnt::at; }
static void reset() { cnt::add = cnt::mul = cnt::dv = cnt::sq = cnt::lg = cnt::at = 0; }

int main()
{
    // a generic source triangle and 64 observation points on, near and far
    const double T0[9] = {0.0, 0.0, 0.0, 2.0, 0.3, 0.1, 0.4, 1.7, -0.2};
    cdouble tv[9], ex[3], ey[3], nh[3], nrm[3];
    for (int k = 0; k < 9; ++k) tv[k] = T0[k];
    tri_frame(tv, ex, ey, nh);
    for (int k = 0; k < 3; ++k) nrm[k] = nh[k];
    const int NP = 64;
    double obs[NP][3];
    unsigned s = 12345u;
    for (int p = 0; p < NP; ++p) {
        for (int k = 0; k < 3; ++k) {
            s = s * 1103515245u + 12345u;
            double r = ((s >> 8) & 0xffff) / 65535.0 - 0.5;
            double scale = (p < 16) ? 0.5 : (p < 32 ? 3.0 : 40.0);
            obs[p][k] = 0.8 + scale * r;
        }
        if (p % 8 == 0) { obs[p][0] = 0.8; obs[p][1] = 0.67; obs[p][2] = -0.033; }
    }
    const double eps = 0.4, mu = 30.0, lam = 30.0, nu = 0.25;
    struct Res { const char* name; double flops, add, mul, dv, sq, lg, at; } res[5];
    const char* names[5] = {"u_pair", "t_pair", "kelvin_stress_pair", "dd_stress_pair", "eigenstress_pair"};
    for (int kind = 0; kind < 5; ++kind) {
        reset();
        for (int p = 0; p < NP; ++p) {
            cdouble o[3] = {obs[p][0], obs[p][1], obs[p][2]};
            cdouble out[81], comb;
            if (kind == 0) {
                u_basis_pair(tv, ex, ey, nh, o, eps, out);
                cdouble g1 = 0.1, g2 = 0.2, g3 = 0.3;   // the in-loop material combine
                for (int q = 0; q < 9; ++q) comb = g1 * out[q] + g2 * out[9 + q] + g3 * out[18 + q];
            } else if (kind == 1) {
                t_basis_pair(tv, ex, ey, nh, nrm, o, eps, out);
                cdouble c[6] = {0.1, 0.2, 0.3, 0.4, 0.5, 0.6};
                for (int q = 0; q < 9; ++q)
                    comb = c[0] * out[q] + c[1] * out[9 + q] + c[2] * out[18 + q] + c[3] * out[27 + q]
                           + c[4] * out[36 + q] + c[5] * out[45 + q];
            } else if (kind == 2) {
                kelvin_stress_pair(tv, ex, ey, nh, o, eps, mu, nu, lam, out);
            } else if (kind == 3) {
                dd_stress_pair(tv, ex, ey, nh, nrm, o, eps, mu, nu, lam, out);
            } else {
                eigenstress_pair(tv, ex, ey, nh, nrm, o, eps, mu, lam, out);
            }
            (void)comb;
        }
        res[kind] = {names[kind], (double)total() / NP, (double)cnt::add / NP, (double)cnt::mul / NP,
                     (double)cnt::dv / NP, (double)cnt::sq / NP, (double)cnt::lg / NP, (double)cnt::at / NP};
    }
    std::printf("{\"rule\": \"add/sub/mul/neg=1, div/sqrt/log/atan2=1, per pair, mean over %d pairs\"", NP);
    for (auto& r : res)
        std::printf(", \"%s\": {\"flops\": %.1f, \"addsub\": %.1f, \"mul\": %.1f, \"div\": %.1f, "
                    "\"sqrt\": %.1f, \"log\": %.1f, \"atan2\": %.1f}",
                    r.name, r.flops, r.add, r.mul, r.dv, r.sq, r.lg, r.at);
    std::printf("}\n");
    return 0;
}
