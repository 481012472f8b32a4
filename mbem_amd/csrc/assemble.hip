// Pair-kernel drivers for gfx950: source preparation, dense U/T block
// assembly (optionally fused with the calibration row sums), basis stacks, and
// the matrix-free contractions behind the displacement and stress readouts.
//
// Assembly mapping (one 64-lane wavefront per tile of R observation rows):
// lane l owns source s0 + l, keeps that source's record in registers while it
// evaluates the R rows, and the wave writes each output row segment
// [3(s0..s0+63), +3) contiguously -- row-major, leading dimension ld, so a
// block lands directly inside the global matrix.  Contraction mapping: one
// lane per observation point, the source loop wave-uniform (scalar loads of
// the source record, uniform skips of degenerate / zero-density sources), the
// 3- or 9-component result written once.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "pair.hpp"
#include "../../include/mbem_hip.h"

using namespace mbem;

namespace {

struct Coef {
    double c[6];
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// ---- per-source record: tri_kernels.py:136-153 (_tri_frame) ---------------
__global__ __launch_bounds__(256) void k_prepare_src(const double* __restrict__ tv,
                                                     const double* __restrict__ nrm,
                                                     const double* __restrict__ eps, long ns,
                                                     double* __restrict__ src, long ld)
{
    // no FMA contraction here: the degenerate test (area2 < 1e-30) relies on
    // the cross product of parallel edges being exactly zero, as it is in the
    // reference's uncontracted arithmetic
#pragma clang fp contract(off)
    long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    double v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) v[k] = tv[9 * s + k];
    double e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
    double e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
    double nx = e1y * e2z - e1z * e2y;
    double ny = e1z * e2x - e1x * e2z;
    double nz = e1x * e2y - e1y * e2x;
    double area2 = sqrt(nx * nx + ny * ny + nz * nz);
    bool ok = !(area2 < 1e-30);
    double nh[3] = {0, 0, 0}, ex[3] = {0, 0, 0}, ey[3] = {0, 0, 0};
    if (ok) {
        nh[0] = nx / area2; nh[1] = ny / area2; nh[2] = nz / area2;
        double e1n = sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
        ex[0] = e1x / e1n; ex[1] = e1y / e1n; ex[2] = e1z / e1n;
        ey[0] = nh[1] * ex[2] - nh[2] * ex[1];
        ey[1] = nh[2] * ex[0] - nh[0] * ex[2];
        ey[2] = nh[0] * ex[1] - nh[1] * ex[0];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) src[(F_V0 + k) * ld + s] = v[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        src[(F_EX + k) * ld + s] = ex[k];
        src[(F_EY + k) * ld + s] = ey[k];
        src[(F_NH + k) * ld + s] = nh[k];
        src[(F_N + k) * ld + s] = nrm[3 * s + k];
    }
    src[F_EPS * ld + s] = eps[s];
    src[F_OK * ld + s] = ok ? 1.0 : 0.0;
    src[(F_OK + 1) * ld + s] = 0.0;
}

// ---- dense block assembly --------------------------------------------------
// KIND 0: U kernel (coefficients g1..g3), KIND 1: T kernel (c1..c6).
// out[(3f+i)*ld + 3s+j] = scale * K_ij + beta * out (STORE); rowsum[9f+ij] +=
// rs_scale * sum_s K_ij (ROWSUM, the calibration row sum of
// backends/dense.py:497-499 / model/equations.py:512-537, fused).
// One wave per R collocation rows, lane = source.  R = 1 is used: measured at
// N = 20,000 the assembly takes 35 ms at R = 1 (2-3 waves per SIMD, no AGPR
// spill) against 44 ms at R = 4 (occupancy 1), the source record's reuse
// across rows being worth less than the latency hiding.
template <int KIND, int R, bool STORE, bool ROWSUM>
__global__ __launch_bounds__(256) void k_block(const double* __restrict__ xf, long nf,
                                               const double* __restrict__ src, long lds, long ns,
                                               Coef cf, double* __restrict__ out, long ldo,
                                               double scale, double beta,
                                               double* __restrict__ rowsum, double rs_scale)
{
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long f0 = wave * R;
    if (f0 >= nf) return;
    double ox[R], oy[R], oz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        long f = (f0 + r < nf) ? f0 + r : nf - 1;
        ox[r] = xf[3 * f]; oy[r] = xf[3 * f + 1]; oz[r] = xf[3 * f + 2];
    }
    double rs[ROWSUM ? R : 1][9];
    if (ROWSUM) {
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int q = 0; q < 9; ++q) rs[r][q] = 0.0;
    }
    for (long s0 = 0; s0 < ns; s0 += 64) {
        const long s = s0 + lane;
        const bool valid = s < ns;
        Src S;
        if (valid) load_src(src, lds, s, S);
        else S.ok = false;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (f0 + r >= nf) break;
            double u[9];
            if (S.ok) {
                if (KIND == 0) {
                    UMom M;
                    u_moments(S, ox[r], oy[r], oz[r], M);
                    u_direct(S, M, cf.c[0], cf.c[1], cf.c[2], u);
                } else {
                    TMom T;
                    t_moments(S, ox[r], oy[r], oz[r], T);
                    t_direct(S, T, cf.c, u);
                }
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) u[q] = 0.0;
            }
            if (STORE && valid) {
                double* o = out + (3 * (f0 + r)) * ldo + 3 * s;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        double* p = o + i * ldo + j;
                        *p = (beta == 0.0) ? scale * u[i * 3 + j] : scale * u[i * 3 + j] + beta * (*p);
                    }
            }
            if (ROWSUM) {
#pragma unroll
                for (int q = 0; q < 9; ++q) rs[r][q] += u[q];
            }
        }
    }
    if (ROWSUM) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (f0 + r >= nf) break;
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                double t = wave_sum(rs[r][q]);
                if (lane == 0) rowsum[9 * (f0 + r) + q] += rs_scale * t;
            }
        }
    }
}

// ---- basis stacks: tri_kernels.py:451-492 ----------------------------------
template <int KIND>
__global__ __launch_bounds__(256) void k_basis(const double* __restrict__ xf, long nf,
                                               const double* __restrict__ src, long lds, long ns,
                                               double* __restrict__ out, long ldo, long stack)
{
    const int lane = threadIdx.x & 63;
    const long f = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (f >= nf) return;
    const double ox = xf[3 * f], oy = xf[3 * f + 1], oz = xf[3 * f + 2];
    for (long s0 = 0; s0 < ns; s0 += 64) {
        const long s = s0 + lane;
        if (s >= ns) continue;
        Src S;
        load_src(src, lds, s, S);
        constexpr int NB = KIND == 0 ? 3 : 6;
        double b[NB * 9];
        if (S.ok) {
            if (KIND == 0) {
                UMom M;
                u_moments(S, ox, oy, oz, M);
                u_basis(S, M, b);
            } else {
                TMom T;
                t_moments(S, ox, oy, oz, T);
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    double c[6] = {0, 0, 0, 0, 0, 0};
                    c[k] = 1.0;
                    t_direct(S, T, c, b + 9 * k);
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < NB * 9; ++q) b[q] = 0.0;
        }
#pragma unroll
        for (int k = 0; k < NB; ++k)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) out[k * stack + (3 * f + i) * ldo + 3 * s + j] = b[k * 9 + i * 3 + j];
    }
}

// ---- matrix-free contractions: tri_kernels.py:976-1049, :1068-1154, :1217-1260
// KIND 0: U displacement, 1: T displacement, 2: Kelvin stress, 3: dislocation
// (total) stress, 4: eigenstress.  out is (nf, 3) or (nf, 3, 3); out = scale *
// result + beta * out.  Zero-density, degenerate and (eigenstress) eps <= 0
// sources are skipped.

// one source's contribution at one point: acc += K(x, source) d
template <int KIND>
__device__ __forceinline__ void contract_pair(const Src& S, const double* d, double ox, double oy,
                                              double oz, const Coef& cf, double mu, double nu,
                                              double lam, double* acc)
{
    if (KIND == 0) {
        UMom M;
        u_moments(S, ox, oy, oz, M);
        double u[9];
        u_direct(S, M, cf.c[0], cf.c[1], cf.c[2], u);
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[i] += u[i * 3 + 0] * d[0] + u[i * 3 + 1] * d[1] + u[i * 3 + 2] * d[2];
    } else if (KIND == 1) {
        TMom T;
        t_moments(S, ox, oy, oz, T);
        double L[9], dl[3], ul[3];
        t_combine_local(T, S.eps, cf.c, L);
        to_local(S, d, dl);
#pragma unroll
        for (int i = 0; i < 3; ++i) ul[i] = L[i * 3 + 0] * dl[0] + L[i * 3 + 1] * dl[1] + L[i * 3 + 2] * dl[2];
#pragma unroll
        for (int i = 0; i < 3; ++i) acc[i] += S.ex[i] * ul[0] + S.ey[i] * ul[1] + S.nh[i] * ul[2];
    } else if (KIND == 2) {
        SMom M;
        s_moments<false>(S, ox, oy, oz, M);
        double fl[3], sl[6], g[9];
        to_local(S, d, fl);
        kelvin_local(M, S.eps, mu, nu, lam, fl, sl);
        rot_sym(S, sl, g);
        acc[0] += g[0]; acc[1] += g[1]; acc[2] += g[2]; acc[3] += g[4]; acc[4] += g[5]; acc[5] += g[8];
    } else if (KIND == 3) {
        SMom M;
        s_moments<true>(S, ox, oy, oz, M);
        double bl[3], nl[3], sl[6], g[9];
        to_local(S, d, bl);
        to_local(S, S.n, nl);
        dd_local(M, S.eps, mu, nu, lam, nl, bl, sl);
        rot_sym(S, sl, g);
        acc[0] += g[0]; acc[1] += g[1]; acc[2] += g[2]; acc[3] += g[4]; acc[4] += g[5]; acc[5] += g[8];
    } else {
        SMom M;
        s_moments<true>(S, ox, oy, oz, M);
        double sg[6];
        eigen_global(M.I7, S.eps, mu, lam, S.n, d, sg);
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[q] += sg[q];
    }
}

template <int NO>
__device__ __forceinline__ void contract_store(const double* acc, long f, double* out, double scale,
                                               double beta)
{
    if (NO == 3) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double* p = out + 3 * f + i;
            *p = (beta == 0.0) ? scale * acc[i] : scale * acc[i] + beta * (*p);
        }
    } else {
        // packed 00 01 02 11 12 22 -> full 3x3
        const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            double* p = out + 9 * f + q;
            *p = (beta == 0.0) ? scale * acc[map[q]] : scale * acc[map[q]] + beta * (*p);
        }
    }
}

// lane = point, the source loop wave-uniform (scalar loads of each source
// record): for many points
template <int KIND>
__global__ __launch_bounds__(256) void k_contract(const double* __restrict__ xf, long nf,
                                                  const double* __restrict__ src, long lds, long ns,
                                                  const double* __restrict__ dens, Coef cf,
                                                  double mu, double lam, double* __restrict__ out,
                                                  double scale, double beta)
{
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = f < nf;
    const long fc = live ? f : nf - 1;
    const double ox = xf[3 * fc], oy = xf[3 * fc + 1], oz = xf[3 * fc + 2];
    const double nu = lam / (2.0 * (lam + mu));
    constexpr int NO = (KIND <= 1) ? 3 : 6;
    double acc[NO];
#pragma unroll
    for (int q = 0; q < NO; ++q) acc[q] = 0.0;
    for (long s = 0; s < ns; ++s) {
        const double d[3] = {dens[3 * s], dens[3 * s + 1], dens[3 * s + 2]};
        if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
        Src S;
        load_src(src, lds, s, S);
        if (!S.ok) continue;
        if (KIND == 4 && !(S.eps > 0.0)) continue;
        contract_pair<KIND>(S, d, ox, oy, oz, cf, mu, nu, lam, acc);
    }
    if (!live) return;
    contract_store<NO>(acc, f, out, scale, beta);
}

// one wave per point, lane = source, a wave reduction at the end: for few
// points (a right-hand side over one patch's collocation points, a grid
// readout), where a lane per point leaves most of the device idle
template <int KIND>
__global__ __launch_bounds__(256) void k_contract_rows(const double* __restrict__ xf, long nf,
                                                       const double* __restrict__ src, long lds,
                                                       long ns, const double* __restrict__ dens,
                                                       Coef cf, double mu, double lam,
                                                       double* __restrict__ out, double scale,
                                                       double beta)
{
    const int lane = threadIdx.x & 63;
    const long f = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (f >= nf) return;
    const double ox = xf[3 * f], oy = xf[3 * f + 1], oz = xf[3 * f + 2];
    const double nu = lam / (2.0 * (lam + mu));
    constexpr int NO = (KIND <= 1) ? 3 : 6;
    double acc[NO];
#pragma unroll
    for (int q = 0; q < NO; ++q) acc[q] = 0.0;
    for (long s = lane; s < ns; s += 64) {
        const double d[3] = {dens[3 * s], dens[3 * s + 1], dens[3 * s + 2]};
        if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
        Src S;
        load_src(src, lds, s, S);
        if (!S.ok) continue;
        if (KIND == 4 && !(S.eps > 0.0)) continue;
        contract_pair<KIND>(S, d, ox, oy, oz, cf, mu, nu, lam, acc);
    }
#pragma unroll
    for (int q = 0; q < NO; ++q) acc[q] = wave_sum(acc[q]);
    if (lane == 0) contract_store<NO>(acc, f, out, scale, beta);
}

// eigenstress on a near list (evaluate.py:298-326): point p sums the source
// chunks listed for it in CSR order (pt_ptr / pt_chunk), each chunk a
// contiguous run [ch_ptr[c], ch_ptr[c+1]) of source indices ch_src.
__global__ __launch_bounds__(256) void k_eigen_near(const double* __restrict__ xf, long nf,
                                                    const double* __restrict__ src, long lds,
                                                    const double* __restrict__ dens,
                                                    const int64_t* __restrict__ pt_ptr,
                                                    const int64_t* __restrict__ pt_chunk,
                                                    const int64_t* __restrict__ ch_ptr,
                                                    const int64_t* __restrict__ ch_src, double mu,
                                                    double lam, double* __restrict__ out,
                                                    double scale, double beta)
{
    const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const double ox = xf[3 * f], oy = xf[3 * f + 1], oz = xf[3 * f + 2];
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (long k = pt_ptr[f]; k < pt_ptr[f + 1]; ++k) {
        const long c = pt_chunk[k];
        double part[6] = {0, 0, 0, 0, 0, 0};
        for (long q = ch_ptr[c]; q < ch_ptr[c + 1]; ++q) {
            const long s = ch_src[q];
            double d[3] = {dens[3 * s], dens[3 * s + 1], dens[3 * s + 2]};
            if (d[0] == 0.0 && d[1] == 0.0 && d[2] == 0.0) continue;
            Src S;
            load_src(src, lds, s, S);
            if (!S.ok || !(S.eps > 0.0)) continue;
            SMom M;
            s_moments<true>(S, ox, oy, oz, M);
            double sg[6];
            eigen_global(M.I7, S.eps, mu, lam, S.n, d, sg);
#pragma unroll
            for (int u = 0; u < 6; ++u) part[u] += sg[u];
        }
#pragma unroll
        for (int u = 0; u < 6; ++u) acc[u] += part[u];
    }
    const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        double* p = out + 9 * f + q;
        *p = (beta == 0.0) ? scale * acc[map[q]] : scale * acc[map[q]] + beta * (*p);
    }
}

// free term on a diagonal block: A[3c+a, 3c+b] += scale * D[c, a, b]
// (model/equations.py:566-582 at P0)
__global__ __launch_bounds__(256) void k_add_block_diag(double* __restrict__ A, long lda, long nc,
                                                        const double* __restrict__ D, double scale)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 9 * nc) return;
    long c = t / 9;
    int ab = (int)(t % 9), a = ab / 3, b = ab % 3;
    A[(3 * c + a) * lda + 3 * c + b] += scale * D[t];
}

inline int launch_status()
{
    return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP;
}

template <int KIND, int R>
int launch_block(const double* xf, long nf, const double* src, long lds, long ns, const Coef& cf,
                 double* out, long ldo, double scale, double beta, double* rowsum, double rs_scale,
                 int store, hipStream_t st)
{
    long waves = (nf + R - 1) / R;
    dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    if (store && rowsum)
        hipLaunchKernelGGL((k_block<KIND, R, true, true>), grid, block, 0, st, xf, nf, src, lds, ns, cf, out, ldo, scale, beta, rowsum, rs_scale);
    else if (store)
        hipLaunchKernelGGL((k_block<KIND, R, true, false>), grid, block, 0, st, xf, nf, src, lds, ns, cf, out, ldo, scale, beta, rowsum, rs_scale);
    else if (rowsum)
        hipLaunchKernelGGL((k_block<KIND, R, false, true>), grid, block, 0, st, xf, nf, src, lds, ns, cf, out, ldo, scale, beta, rowsum, rs_scale);
    else
        return MBEM_OK;
    return launch_status();
}

constexpr long kRowsBelow = 64 * 1024;

template <int KIND>
int launch_contract(const double* xf, long nf, const double* src, long lds, long ns,
                    const double* dens, const Coef& cf, double mu, double lam, double* out,
                    double scale, double beta, hipStream_t st)
{
    if (nf == 0) return MBEM_OK;
    // a lane per point gives nf / 64 waves: below one wave per SIMD (256 CUs
    // x 4), one wave per point instead (each wave then reads every source
    // record, which is cheap only while the points are few)
    if (nf < kRowsBelow) {
        dim3 grid((unsigned)((nf + 3) / 4)), block(256);
        hipLaunchKernelGGL((k_contract_rows<KIND>), grid, block, 0, st, xf, nf, src, lds, ns, dens, cf, mu, lam, out, scale, beta);
    } else {
        dim3 grid((unsigned)((nf + 255) / 256)), block(256);
        hipLaunchKernelGGL((k_contract<KIND>), grid, block, 0, st, xf, nf, src, lds, ns, dens, cf, mu, lam, out, scale, beta);
    }
    return launch_status();
}

bool bad_src(const double* src, long lds, long ns) { return ns < 0 || (ns > 0 && (!src || lds < ns)); }

}  // namespace

extern "C" {

int mbem_src_fields(void) { return NSF; }

int mbem_prepare_sources(const double* tri_verts, const double* normals, const double* eps,
                         int64_t ns, double* src, int64_t ld_src, hipStream_t stream)
{
    if (ns < 0 || ld_src < ns || (ns > 0 && (!tri_verts || !normals || !eps || !src))) return MBEM_EINVAL;
    if (ns == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_prepare_src, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, stream, tri_verts, normals, eps, ns, src, ld_src);
    return launch_status();
}

int mbem_u_block(const double* x_field, int64_t nf, const double* src, int64_t ld_src, int64_t ns,
                 const double* g, double* out, int64_t ld_out, double scale, double beta,
                 hipStream_t stream)
{
    if (nf < 0 || bad_src(src, ld_src, ns) || !g || ld_out < 3 * ns) return MBEM_EINVAL;
    if (nf == 0 || ns == 0) return MBEM_OK;
    if (!x_field || !out) return MBEM_EINVAL;
    Coef cf{{g[0], g[1], g[2], 0.0, 0.0, 0.0}};
    return launch_block<0, 1>(x_field, nf, src, ld_src, ns, cf, out, ld_out, scale, beta, nullptr, 0.0, 1, stream);
}

int mbem_t_block(const double* x_field, int64_t nf, const double* src, int64_t ld_src, int64_t ns,
                 const double* c, double* out, int64_t ld_out, double scale, double beta,
                 double* rowsum, double rs_scale, int store, hipStream_t stream)
{
    if (nf < 0 || bad_src(src, ld_src, ns) || !c) return MBEM_EINVAL;
    if (store && ld_out < 3 * ns) return MBEM_EINVAL;
    if (nf == 0 || ns == 0) return MBEM_OK;
    if (!x_field || (store && !out) || (!store && !rowsum)) return MBEM_EINVAL;
    Coef cf{{c[0], c[1], c[2], c[3], c[4], c[5]}};
    return launch_block<1, 1>(x_field, nf, src, ld_src, ns, cf, out, ld_out, scale, beta, rowsum, rs_scale, store, stream);
}

int mbem_u_basis(const double* x_field, int64_t nf, const double* src, int64_t ld_src, int64_t ns,
                 double* out, int64_t ld_out, int64_t stack_stride, hipStream_t stream)
{
    if (nf < 0 || bad_src(src, ld_src, ns) || ld_out < 3 * ns) return MBEM_EINVAL;
    if (nf == 0 || ns == 0) return MBEM_OK;
    if (!x_field || !out || stack_stride < 3 * nf * ld_out) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_basis<0>, dim3((unsigned)((nf + 3) / 4)), dim3(256), 0, stream, x_field, nf, src, ld_src, ns, out, ld_out, stack_stride);
    return launch_status();
}

int mbem_t_basis(const double* x_field, int64_t nf, const double* src, int64_t ld_src, int64_t ns,
                 double* out, int64_t ld_out, int64_t stack_stride, hipStream_t stream)
{
    if (nf < 0 || bad_src(src, ld_src, ns) || ld_out < 3 * ns) return MBEM_EINVAL;
    if (nf == 0 || ns == 0) return MBEM_OK;
    if (!x_field || !out || stack_stride < 3 * nf * ld_out) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_basis<1>, dim3((unsigned)((nf + 3) / 4)), dim3(256), 0, stream, x_field, nf, src, ld_src, ns, out, ld_out, stack_stride);
    return launch_status();
}

int mbem_u_disp_contract(const double* x_field, int64_t nf, const double* src, int64_t ld_src,
                         int64_t ns, const double* density, const double* g, double* out,
                         double scale, double beta, hipStream_t stream)
{
    if (nf < 0 || bad_src(src, ld_src, ns) || !g || (nf > 0 && (!x_field || !out)) || (ns > 0 && !density)) return MBEM_EINVAL;
    Coef cf{{g[0], g[1], g[2], 0.0, 0.0, 0.0}};
    return launch_contract<0>(x_field, nf, src, ld_src, ns, density, cf, 0.0, 0.0, out, scale, beta, stream);
}

int mbem_t_disp_contract(const double* x_field, int64_t nf, const double* src, int64_t ld_src,
                         int64_t ns, const double* density, const double* c, double* out,
                         double scale, double beta, hipStream_t stream)
{
    if (nf < 0 || bad_src(src, ld_src, ns) || !c || (nf > 0 && (!x_field || !out)) || (ns > 0 && !density)) return MBEM_EINVAL;
    Coef cf{{c[0], c[1], c[2], c[3], c[4], c[5]}};
    return launch_contract<1>(x_field, nf, src, ld_src, ns, density, cf, 0.0, 0.0, out, scale, beta, stream);
}

static int stress_contract(int kind, const double* x_field, int64_t nf, const double* src,
                           int64_t ld_src, int64_t ns, const double* density, double mu, double lam,
                           double* out, double scale, double beta, hipStream_t stream)
{
    if (nf < 0 || bad_src(src, ld_src, ns) || (nf > 0 && (!x_field || !out)) || (ns > 0 && !density)) return MBEM_EINVAL;
    if (!(mu > 0.0) || !(lam + mu != 0.0)) return MBEM_EINVAL;
    Coef cf{{0, 0, 0, 0, 0, 0}};
    if (kind == 2) return launch_contract<2>(x_field, nf, src, ld_src, ns, density, cf, mu, lam, out, scale, beta, stream);
    if (kind == 3) return launch_contract<3>(x_field, nf, src, ld_src, ns, density, cf, mu, lam, out, scale, beta, stream);
    return launch_contract<4>(x_field, nf, src, ld_src, ns, density, cf, mu, lam, out, scale, beta, stream);
}

int mbem_kelvin_stress_contract(const double* x_field, int64_t nf, const double* src, int64_t ld_src,
                                int64_t ns, const double* density, double mu, double lam,
                                double* out, double scale, double beta, hipStream_t stream)
{
    return stress_contract(2, x_field, nf, src, ld_src, ns, density, mu, lam, out, scale, beta, stream);
}

int mbem_dd_stress_contract(const double* x_field, int64_t nf, const double* src, int64_t ld_src,
                            int64_t ns, const double* density, double mu, double lam, double* out,
                            double scale, double beta, hipStream_t stream)
{
    return stress_contract(3, x_field, nf, src, ld_src, ns, density, mu, lam, out, scale, beta, stream);
}

int mbem_eigenstress_contract(const double* x_field, int64_t nf, const double* src, int64_t ld_src,
                              int64_t ns, const double* density, double mu, double lam,
                              double* out, double scale, double beta, hipStream_t stream)
{
    return stress_contract(4, x_field, nf, src, ld_src, ns, density, mu, lam, out, scale, beta, stream);
}

int mbem_eigenstress_near(const double* x_field, int64_t nf, const double* src, int64_t ld_src,
                          int64_t ns, const double* density, const int64_t* pt_ptr,
                          const int64_t* pt_chunk, const int64_t* ch_ptr, const int64_t* ch_src,
                          double mu, double lam, double* out, double scale, double beta,
                          hipStream_t stream)
{
    if (nf < 0 || bad_src(src, ld_src, ns)) return MBEM_EINVAL;
    if (nf == 0) return MBEM_OK;
    if (!x_field || !out || !pt_ptr || !pt_chunk || !ch_ptr || !ch_src || !density) return MBEM_EINVAL;
    if (!(mu > 0.0)) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_eigen_near, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, stream, x_field, nf, src, ld_src, density, pt_ptr, pt_chunk, ch_ptr, ch_src, mu, lam, out, scale, beta);
    return launch_status();
}

int mbem_add_block_diag(double* A, int64_t lda, int64_t nc, const double* D, double scale,
                        hipStream_t stream)
{
    if (nc < 0 || (nc > 0 && (!A || !D || lda < 3 * nc))) return MBEM_EINVAL;
    if (nc == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_add_block_diag, dim3((unsigned)((9 * nc + 255) / 256)), dim3(256), 0, stream, A, lda, nc, D, scale);
    return launch_status();
}

// the same operations and order as numpy's vertices[triangles], np.cross,
// sqrt(add.reduce(c * c)) and mean over the three vertices (host FP64, no
// contraction into FMAs)
int mbem_tri_geometry(const double* verts, int64_t nv, const int64_t* tris, int64_t nt,
                      double* tri_verts, double* normals, double* centroids)
{
#pragma clang fp contract(off)
    if (nv < 0 || nt < 0 || (nt > 0 && (!verts || !tris || !tri_verts || !normals || !centroids)))
        return MBEM_EINVAL;
    for (int64_t t = 0; t < 3 * nt; ++t)
        if (tris[t] < 0 || tris[t] >= nv) return MBEM_EINVAL;
    for (int64_t t = 0; t < nt; ++t) {
        const double* v0 = verts + 3 * tris[3 * t];
        const double* v1 = verts + 3 * tris[3 * t + 1];
        const double* v2 = verts + 3 * tris[3 * t + 2];
        double* tv = tri_verts + 9 * t;
        for (int k = 0; k < 3; ++k) {
            tv[k] = v0[k];
            tv[3 + k] = v1[k];
            tv[6 + k] = v2[k];
        }
        const double a0 = v1[0] - v0[0], a1 = v1[1] - v0[1], a2 = v1[2] - v0[2];
        const double b0 = v2[0] - v0[0], b1 = v2[1] - v0[1], b2 = v2[2] - v0[2];
        const double c0 = a1 * b2 - a2 * b1;
        const double c1 = a2 * b0 - a0 * b2;
        const double c2 = a0 * b1 - a1 * b0;
        const double len = std::sqrt((c0 * c0 + c1 * c1) + c2 * c2);
        normals[3 * t] = c0 / len;
        normals[3 * t + 1] = c1 / len;
        normals[3 * t + 2] = c2 / len;
        for (int k = 0; k < 3; ++k) centroids[3 * t + k] = ((v0[k] + v1[k]) + v2[k]) / 3.0;
    }
    return MBEM_OK;
}

}  // extern "C"
