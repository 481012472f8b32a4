// Batched Laplace-state solves on gfx950: many right-hand sides, each with
// its own Laplace parameter, through one pair of geometry operators.
//
// A relaxing region's operator at Laplace parameter s is
//     A(s) = (A0 + g(s) A1) D(s),   g = 1/(1 - nu(s)),
// with D(s) = 1 on u-slot columns and 1/mu(s) on t-slot columns.  With
// y = D x every system is B(g_c) y_c = b_c, B(g) = A0 + g A1, and a block of
// K columns (K = states x right-hand sides) is advanced together:
//
//   * k_mgemm<..., DUAL>: Y = alpha (A0 X + A1 X diag(g)) + beta Y on the
//     FP64 matrix cores (v_mfma_f64_16x16x4f64).  One accumulator set per
//     output tile: the LDS holds the A0 and A1 row tiles and both X and
//     X diag(g), so [A0 | A1] [X ; X g] is one GEMM of depth 2n.
//   * k_mgemm<..., TABLE>: the block-Jacobi apply as a grouped GEMM over a
//     work table (chunk inverse x the chunk's rows of a column range).
//   * the Arnoldi process of K independent FGMRES runs in lockstep: block
//     multi-dots, multi-axpys, Givens rotations and the triangular solve on
//     the device, so a whole restart cycle is a fixed kernel sequence that
//     can be captured in a HIP graph.
//
// Block vectors are row-major n x K (ld = K): column c is system c.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <stdint.h>

#include "../../include/mbem_hip.h"

namespace {

rocblas_handle g_blas = nullptr;

typedef double dbl2 __attribute__((ext_vector_type(2)));
typedef double dbl4 __attribute__((ext_vector_type(4)));

inline int launch_status() { return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP; }

// one work item of the grouped GEMM: C[m x ncols] = alpha Ainv[m x m] X[m x ncols]
// (+ beta C); offsets in elements from the base pointers of the launch.
struct Work {
    long a_off;      // chunk inverse (row-major m x m, even leading dimension a_ld)
    long a_ld;
    long m;          // chunk size (rows = inner dimension)
    long xy_off;     // first element of the chunk's rows and column range in X and Y
    long ncols;      // width of the column range
    long tile_r;     // row tile index inside the chunk
    long tile_c;     // column tile index inside the range
    long pad0;
};

// v_mfma_f64_16x16x4f64 operand layout (wave64), as measured on gfx950:
// lane l holds A[i = l % 16][k = l / 16] and B[k = l / 16][j = l % 16], and
// element v of its double4 result is D[i = l / 16 + 4 v][j = l % 16].
//
// Block tile BM x BN over WM x WN waves, depth slices of BK through a double
// buffer in LDS: the next slice's global loads are issued before the MFMAs
// of the current one and stored into the other buffer after them, one
// barrier per slice.  (The double buffer and eight waves per 128 x 128 tile
// measured 62 TFLOP/s at n = 31,104, K = 256, against 50 TFLOP/s single
// buffered.)
template <int BM, int BN, int BK, int WM, int WN, bool DUAL, bool TABLE>
__global__ __launch_bounds__(WM * WN * 64) void k_mgemm(
    const double* __restrict__ A0, const double* __restrict__ A1, long lda, long M, long Kd,
    const double* __restrict__ X, long ldx, long N, const double* __restrict__ g,
    double* __restrict__ Y, long ldy, double alpha, double beta, const Work* __restrict__ work)
{
    constexpr int NT = WM * WN * 64;
    constexpr int TM = BM / WM / 16;          // 16x16 tiles per wave (rows)
    constexpr int TN = BN / WN / 16;          // 16x16 tiles per wave (columns)
    constexpr int PADA = BM + 2;
    constexpr int PADX = BN + 2;
    static_assert(TM >= 1 && TN >= 1, "tile shape");
    static_assert(BK % 4 == 0, "BK");
    constexpr int NA2 = BM * BK / 2, NX2 = BK * BN / 2;      // dbl2 elements per tile
    constexpr int LA = (NA2 + NT - 1) / NT, LX = (NX2 + NT - 1) / NT;

    __shared__ double sA0[2][BK * PADA];
    __shared__ double sA1[2][DUAL ? BK * PADA : 1];
    __shared__ double sX[2][BK * PADX];
    __shared__ double sXG[2][DUAL ? BK * PADX : 1];

    long row0, col0;
    const double* Ab0 = A0;
    const double* Ab1 = A1;
    const double* Xb = X;
    double* Yb = Y;
    long Mr = M, Kr = Kd, Nr = N, ldar = lda;
    if (TABLE) {
        const Work w = work[blockIdx.x];
        Ab0 = A0 + w.a_off;
        ldar = w.a_ld;
        Mr = w.m;
        Kr = w.m;
        Nr = w.ncols;
        Xb = X + w.xy_off;
        Yb = Y + w.xy_off;
        row0 = w.tile_r * BM;
        col0 = w.tile_c * BN;
    } else {
        row0 = (long)blockIdx.x * BM;
        col0 = (long)blockIdx.y * BN;
    }

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int wrow = wm * (BM / WM), wcol = wn * (BN / WN);

    dbl4 acc[TM][TN];
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) acc[a][b] = dbl4{0.0, 0.0, 0.0, 0.0};

    dbl2 ra0[LA], ra1[DUAL ? LA : 1], rx[LX];

    auto load_tiles = [&](long k0) {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            const int e = tid + q * NT;
            const int r = e / (BK / 2), kk = (e % (BK / 2)) * 2;
            const long gr = row0 + r, gk = k0 + kk;
            dbl2 v0 = {0.0, 0.0}, v1 = {0.0, 0.0};
            if (e < NA2 && gr < Mr) {
                if (gk + 1 < Kr) {
                    v0 = *reinterpret_cast<const dbl2*>(Ab0 + gr * ldar + gk);
                    if (DUAL) v1 = *reinterpret_cast<const dbl2*>(Ab1 + gr * ldar + gk);
                } else if (gk < Kr) {
                    v0.x = Ab0[gr * ldar + gk];
                    if (DUAL) v1.x = Ab1[gr * ldar + gk];
                }
            }
            ra0[q] = v0;
            if (DUAL) ra1[q] = v1;
        }
#pragma unroll
        for (int q = 0; q < LX; ++q) {
            const int e = tid + q * NT;
            const int k = e / (BN / 2), cc = (e % (BN / 2)) * 2;
            const long gk = k0 + k, gc = col0 + cc;
            dbl2 v = {0.0, 0.0};
            if (e < NX2 && gk < Kr) {
                if (gc + 1 < Nr) v = *reinterpret_cast<const dbl2*>(Xb + gk * ldx + gc);
                else if (gc < Nr) v.x = Xb[gk * ldx + gc];
            }
            rx[q] = v;
        }
    };
    auto store_tiles = [&](int buf) {
#pragma unroll
        for (int q = 0; q < LA; ++q) {
            const int e = tid + q * NT;
            if (e >= NA2) continue;
            const int r = e / (BK / 2), kk = (e % (BK / 2)) * 2;
            sA0[buf][kk * PADA + r] = ra0[q].x;
            sA0[buf][(kk + 1) * PADA + r] = ra0[q].y;
            if (DUAL) {
                sA1[buf][kk * PADA + r] = ra1[q].x;
                sA1[buf][(kk + 1) * PADA + r] = ra1[q].y;
            }
        }
#pragma unroll
        for (int q = 0; q < LX; ++q) {
            const int e = tid + q * NT;
            if (e >= NX2) continue;
            const int k = e / (BN / 2), cc = (e % (BN / 2)) * 2;
            sX[buf][k * PADX + cc] = rx[q].x;
            sX[buf][k * PADX + cc + 1] = rx[q].y;
            if (DUAL) {
                const long gc = col0 + cc;
                const double g0 = (gc < Nr) ? g[gc] : 0.0;
                const double g1 = (gc + 1 < Nr) ? g[gc + 1] : 0.0;
                sXG[buf][k * PADX + cc] = rx[q].x * g0;
                sXG[buf][k * PADX + cc + 1] = rx[q].y * g1;
            }
        }
    };

    const int fi = lane & 15, fk = lane >> 4;
    load_tiles(0);
    store_tiles(0);
    __syncthreads();
    int cur = 0;
    for (long k0 = 0; k0 < Kr; k0 += BK) {
        const bool more = k0 + BK < Kr;
        if (more) load_tiles(k0 + BK);          // next slice in flight during the MFMAs
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double a0[TM], a1[DUAL ? TM : 1], b0[TN], b1[DUAL ? TN : 1];
#pragma unroll
            for (int a = 0; a < TM; ++a) {
                a0[a] = sA0[cur][(kk + fk) * PADA + wrow + a * 16 + fi];
                if (DUAL) a1[a] = sA1[cur][(kk + fk) * PADA + wrow + a * 16 + fi];
            }
#pragma unroll
            for (int b = 0; b < TN; ++b) {
                b0[b] = sX[cur][(kk + fk) * PADX + wcol + b * 16 + fi];
                if (DUAL) b1[b] = sXG[cur][(kk + fk) * PADX + wcol + b * 16 + fi];
            }
#pragma unroll
            for (int a = 0; a < TM; ++a)
#pragma unroll
                for (int b = 0; b < TN; ++b) {
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[a], b0[b], acc[a][b], 0, 0, 0);
                    if (DUAL)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[a], b1[b], acc[a][b], 0, 0, 0);
                }
        }
        if (more) store_tiles(cur ^ 1);         // the other buffer: last read two slices ago
        __syncthreads();
        cur ^= 1;
    }
    // write back: element v of tile (a, b) is row wrow + 16a + lane / 16 + 4 v,
    // column wcol + 16b + lane % 16
#pragma unroll
    for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TN; ++b) {
            const long gc = col0 + wcol + b * 16 + fi;
            if (gc >= Nr) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long gr = row0 + wrow + a * 16 + fk + 4 * v;
                if (gr >= Mr) continue;
                double* p = Yb + gr * ldy + gc;
                const double val = alpha * acc[a][b][v];
                *p = (beta == 0.0) ? val : val + beta * (*p);
            }
        }
}

// ---------------------------------------------------------------------------
// Block-vector kernels (row-major n x K blocks, ld = K)

// partial[(blockIdx.y * 4 + wave) * cnt * K + i * K + c] =
//   sum over the wave's rows r of V_{i0+i}[r, c] * W[r, c],  i < cnt <= 8
__global__ __launch_bounds__(256) void k_bdot(const double* __restrict__ V, long bstride, long cnt,
                                              const double* __restrict__ W, long n, long K,
                                              long rows_per_wave, double* __restrict__ partial)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long c = (long)blockIdx.x * 64 + lane;
    const long slot = (long)blockIdx.y * 4 + wave;
    const long r0 = slot * rows_per_wave;
    const long r1 = (r0 + rows_per_wave < n) ? r0 + rows_per_wave : n;
    double acc[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = 0.0;
    if (c < K) {
        for (long r = r0; r < r1; ++r) {
            const double w = W[r * K + c];
#pragma unroll
            for (int i = 0; i < 8; ++i)
                if (i < cnt) acc[i] += V[i * bstride + r * K + c] * w;
        }
        for (int i = 0; i < cnt; ++i) partial[(slot * cnt + i) * K + c] = acc[i];
    }
}

// out[c * ldo + o0 + i] (= or +=) sum over slots of partial[slot][i][c]
__global__ __launch_bounds__(256) void k_breduce(const double* __restrict__ partial, long nslots,
                                                 long cnt, long K, double* __restrict__ out,
                                                 long ldo, long o0, int accumulate,
                                                 double* __restrict__ copy)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cnt * K) return;
    const long i = t / K, c = t % K;
    double s = 0.0;
    for (long q = 0; q < nslots; ++q) s += partial[(q * cnt + i) * K + c];
    double* p = out + c * ldo + o0 + i;
    *p = accumulate ? *p + s : s;
    if (copy) copy[c * cnt + i] = s;
}

// Y[r, c] += sign * sum_{i < cnt} coef[c * ldc + i] * B_i[r, c]; cnt from
// cnt_dev[0] when given (the Arnoldi steps done this cycle)
__global__ __launch_bounds__(256) void k_baxpy(const double* __restrict__ B, long bstride, long cnt,
                                               const int* __restrict__ cnt_dev,
                                               const double* __restrict__ coef, long ldc,
                                               double sign, double* __restrict__ Y, long n, long K)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * K) return;
    const long c = t % K;
    const long m = cnt_dev ? (long)cnt_dev[0] : cnt;
    double s = 0.0;
    for (long i = 0; i < m; ++i) s += coef[c * ldc + i] * B[i * bstride + t];
    Y[t] += sign * s;
}

// Y[r, c] = X[r, c] * scale[c]
__global__ __launch_bounds__(256) void k_bscale(const double* __restrict__ X,
                                                const double* __restrict__ scale,
                                                double* __restrict__ Y, long n, long K)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * K) return;
    Y[t] = X[t] * scale[t % K];
}

// Y = X (block copy)
__global__ __launch_bounds__(256) void k_bcopy(const double* __restrict__ X, double* __restrict__ Y,
                                               long nk)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nk) Y[t] = X[t];
}

// Cycle start, one thread per column: beta_c = sqrt(sumsq_c); scale_c =
// 1/beta_c (0 when beta_c == 0); gvec_c = [beta_c, 0, ...]; jdone = 0.
__global__ void k_fg_start(long K, long m, const double* __restrict__ sumsq,
                           double* __restrict__ beta, double* __restrict__ scale,
                           double* __restrict__ gvec, int* __restrict__ jdone)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) jdone[0] = 0;
    if (c >= K) return;
    const double b = sqrt(sumsq[c]);
    beta[c] = b;
    scale[c] = (b > 0.0) ? 1.0 / b : 0.0;
    for (long i = 0; i <= m; ++i) gvec[c * (m + 1) + i] = 0.0;
    gvec[c * (m + 1)] = b;
}

// Arnoldi step j, one thread per column.  H[c] is (m+1) x m row-major; its
// column j becomes the CGS2 projections hp1 + hp2 (each K x (m+1)) and
// H[j+1, j] = |w| (sumsq_c = |w|^2 after orthogonalisation).  Then the next
// basis scale, the rotations, the updated gvec and the residual estimate
// |gvec[j+1]|.
__global__ void k_fg_givens(long K, long m, long j, const double* __restrict__ hp1,
                            const double* __restrict__ hp2, double* __restrict__ H,
                            double* __restrict__ cs, double* __restrict__ sn,
                            double* __restrict__ gvec, const double* __restrict__ sumsq,
                            double* __restrict__ scale, double* __restrict__ res,
                            int* __restrict__ jdone)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) jdone[0] = (int)(j + 1);
    if (c >= K) return;
    double* h = H + c * (m + 1) * m;
    double* cc = cs + c * m;
    double* ss = sn + c * m;
    double* gv = gvec + c * (m + 1);
    for (long i = 0; i <= j; ++i) h[i * m + j] = hp1[c * (m + 1) + i] + hp2[c * (m + 1) + i];
    const double nrm = sqrt(sumsq[c]);
    h[(j + 1) * m + j] = nrm;
    // breakdown guard: a zero (or denormal) new direction ends this column's
    // Krylov space; its next basis vector is set to zero
    scale[c] = (nrm > 1e-300) ? 1.0 / nrm : 0.0;
    for (long i = 0; i < j; ++i) {
        const double a = h[i * m + j], b = h[(i + 1) * m + j];
        h[i * m + j] = cc[i] * a + ss[i] * b;
        h[(i + 1) * m + j] = -ss[i] * a + cc[i] * b;
    }
    const double a = h[j * m + j], b = h[(j + 1) * m + j];
    const double d = sqrt(a * a + b * b);
    double cj = 1.0, sj = 0.0;
    if (d > 0.0) {
        cj = a / d;
        sj = b / d;
    }
    cc[j] = cj;
    ss[j] = sj;
    h[j * m + j] = d;
    h[(j + 1) * m + j] = 0.0;
    const double g0 = gv[j];
    gv[j] = cj * g0;
    gv[j + 1] = -sj * g0;
    res[c] = fabs(gv[j + 1]);
}

// y_c = R_c^{-1} gvec_c over the jdone leading steps (upper-triangular R in H)
__global__ void k_fg_solve(long K, long m, const double* __restrict__ H,
                           const double* __restrict__ gvec, const int* __restrict__ jdone,
                           double* __restrict__ y)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const long k = jdone[0];
    const double* h = H + c * (m + 1) * m;
    const double* gv = gvec + c * (m + 1);
    double* yc = y + c * m;
    for (long i = k - 1; i >= 0; --i) {
        double s = gv[i];
        for (long l = i + 1; l < k; ++l) s -= h[i * m + l] * yc[l];
        const double d = h[i * m + i];
        yc[i] = (fabs(d) > 0.0) ? s / d : 0.0;
    }
    for (long i = k; i < m; ++i) yc[i] = 0.0;
}

// fill kernel (used in captured graphs in place of hipMemsetAsync)
__global__ __launch_bounds__(256) void k_fill(double* __restrict__ y, long n, double v)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = v;
}

// out[i, j] = A0[idx[i], idx[j]] + g A1[idx[i], idx[j]]: a chunk of B(g)
__global__ __launch_bounds__(256) void k_gather_comb(const double* __restrict__ A0,
                                                     const double* __restrict__ A1, long lda,
                                                     const int64_t* __restrict__ idx, long m,
                                                     double g, double* __restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * m) return;
    const long i = t / m, j = t % m;
    const long src = idx[i] * lda + idx[j];
    out[t] = A0[src] + g * A1[src];
}

// padded chunk of B(g): out (mp x mp) = [A0 + g A1 on idx x idx, 0; 0, I]
__global__ __launch_bounds__(256) void k_gather_comb_pad(const double* __restrict__ A0,
                                                         const double* __restrict__ A1, long lda,
                                                         const int64_t* __restrict__ idx, long m,
                                                         long mp, double g, double* __restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= mp * mp) return;
    const long i = t / mp, j = t % mp;
    if (i < m && j < m) {
        const long src = idx[i] * lda + idx[j];
        out[t] = A0[src] + g * A1[src];
    } else {
        out[t] = (i == j) ? 1.0 : 0.0;
    }
}

// out[i, j] = A[idx[i], idx[j]] for a full permutation (m = n)
__global__ __launch_bounds__(256) void k_permute(const double* __restrict__ A, long lda,
                                                 const int64_t* __restrict__ idx, long n,
                                                 double* __restrict__ out, long ldo)
{
    for (long i = blockIdx.y; i < n; i += gridDim.y)
        for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x)
            out[i * ldo + j] = A[idx[i] * lda + idx[j]];
}

// ---------------------------------------------------------------------------
// tile configurations of the GEMM (the autotuner's candidates)

struct Cfg { int bm, bn, bk, wm, wn; };
constexpr Cfg kCfgs[] = {
    {64, 64, 8, 2, 2},
    {64, 128, 8, 2, 2},
    {128, 64, 8, 2, 2},
    {128, 128, 8, 4, 2},
    {128, 128, 8, 2, 4},
    {64, 128, 8, 2, 4},
    {128, 64, 8, 4, 2},
    {128, 128, 16, 2, 2},
    {256, 128, 8, 4, 2},
};
constexpr int kNCfg = sizeof(kCfgs) / sizeof(kCfgs[0]);

// static LDS of a k_mgemm instance (double-buffered tiles, rows padded by 2)
constexpr long mgemm_lds(int BM, int BN, int BK, bool DUAL)
{
    return 8L * 2 * BK * ((BM + 2) * (DUAL ? 2 : 1) + (BN + 2) * (DUAL ? 2 : 1));
}
constexpr long kLdsBytes = 160L * 1024;

template <int BM, int BN, int BK, int WM, int WN, bool DUAL, bool TABLE>
int launch_cfg(dim3 grid, hipStream_t st, const double* A0, const double* A1, long lda, long M,
               long Kd, const double* X, long ldx, long N, const double* g, double* Y, long ldy,
               double alpha, double beta, const Work* work)
{
    if constexpr (mgemm_lds(BM, BN, BK, DUAL) > kLdsBytes) {
        return MBEM_EINVAL;        // this tile does not fit the LDS in this form
    } else {
        hipLaunchKernelGGL((k_mgemm<BM, BN, BK, WM, WN, DUAL, TABLE>), grid, dim3(WM * WN * 64), 0,
                           st, A0, A1, lda, M, Kd, X, ldx, N, g, Y, ldy, alpha, beta, work);
        return launch_status();
    }
}

template <bool DUAL, bool TABLE>
int dispatch(int cfg, dim3 grid, hipStream_t st, const double* A0, const double* A1, long lda,
             long M, long Kd, const double* X, long ldx, long N, const double* g, double* Y,
             long ldy, double alpha, double beta, const Work* work)
{
#define MBEM_CFG(i, BM, BN, BK, WM, WN) \
    case i: return launch_cfg<BM, BN, BK, WM, WN, DUAL, TABLE>(grid, st, A0, A1, lda, M, Kd, X, ldx, N, g, Y, ldy, alpha, beta, work);
    switch (cfg) {
        MBEM_CFG(0, 64, 64, 8, 2, 2)
        MBEM_CFG(1, 64, 128, 8, 2, 2)
        MBEM_CFG(2, 128, 64, 8, 2, 2)
        MBEM_CFG(3, 128, 128, 8, 4, 2)
        MBEM_CFG(4, 128, 128, 8, 2, 4)
        MBEM_CFG(5, 64, 128, 8, 2, 4)
        MBEM_CFG(6, 128, 64, 8, 4, 2)
        MBEM_CFG(7, 128, 128, 16, 2, 2)
        MBEM_CFG(8, 256, 128, 8, 4, 2)
        default: return MBEM_EINVAL;
    }
#undef MBEM_CFG
}

// whether configuration cfg has a DUAL (two-operator) form within the LDS
template <bool DUAL>
bool cfg_fits(int cfg)
{
    const Cfg c = kCfgs[cfg];
    return mgemm_lds(c.bm, c.bn, c.bk, DUAL) <= kLdsBytes;
}

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// out[i] = sum_s part[s * len + i]: the split-K partial Gram matrices summed
// in a fixed order (one thread per entry)
__global__ __launch_bounds__(256) void k_sum_slices(const double* __restrict__ part, long len,
                                                    long S, double* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double acc = 0.0;
    for (long s = 0; s < S; ++s) acc += part[s * len + i];
    out[i] = acc;
}

// The coarse space's Gram matrix G (K x K, symmetric) scaled to unit
// diagonal and regularised: d_i = sqrt(G_ii) (1 for an all-zero column,
// whose diagonal becomes 1), G_ij <- G_ij / (d_i d_j) + delta [i = j].
// One thread per entry; d is written by the diagonal's threads first (a
// separate launch), so every entry reads finished d values.
__global__ __launch_bounds__(256) void k_gram_diag(const double* __restrict__ G, long K,
                                                   double* __restrict__ d)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K) return;
    const double g = G[i * K + i];
    d[i] = g > 0.0 ? sqrt(g) : 0.0;
}

__global__ __launch_bounds__(256) void k_gram_scale(double* __restrict__ G, long K,
                                                    const double* __restrict__ d, double delta)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K * K) return;
    const long i = e / K, j = e - i * K;
    const double di = d[i] > 0.0 ? d[i] : 1.0, dj = d[j] > 0.0 ? d[j] : 1.0;
    double v = G[e] / (di * dj);
    if (i == j) v = (d[i] > 0.0 ? v : 1.0) + delta;
    G[e] = v;
}

// From M, whose column-major lower triangle holds Linv = L^-1 (the row-major
// view: M[j * K + i] = Linv(i, j), i >= j), form the row-major
// Tt = Linv D^-1 and T = Tt^T, D = diag(d) (1 for a zero column).
__global__ __launch_bounds__(256) void k_coarse_tfactors(const double* __restrict__ M, long K,
                                                         const double* __restrict__ d,
                                                         double* __restrict__ Tt,
                                                         double* __restrict__ T)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K * K) return;
    const long i = e / K, j = e - i * K;
    const double dj = d[j] > 0.0 ? d[j] : 1.0;
    const double v = (i >= j) ? M[j * K + i] / dj : 0.0;
    Tt[i * K + j] = v;
    T[j * K + i] = v;
}

}  // namespace

extern "C" {

int mbem_mgemm_configs(void) { return kNCfg; }

int mbem_mgemm_config(int cfg, int32_t* out5)
{
    if (cfg < 0 || cfg >= kNCfg || !out5) return MBEM_EINVAL;
    out5[0] = kCfgs[cfg].bm;
    out5[1] = kCfgs[cfg].bn;
    out5[2] = kCfgs[cfg].bk;
    out5[3] = kCfgs[cfg].wm;
    out5[4] = kCfgs[cfg].wn;
    return MBEM_OK;
}

int mbem_mgemm_config_fits(int cfg, int dual)
{
    if (cfg < 0 || cfg >= kNCfg) return 0;
    return dual ? cfg_fits<true>(cfg) : cfg_fits<false>(cfg);
}

int mbem_gemm_rm(const double* A, int64_t lda, int64_t m, int64_t kd, const double* X, int64_t ldx,
                 int64_t ncols, double* Y, int64_t ldy, double alpha, double beta, int cfg,
                 hipStream_t stream)
{
    if (m < 0 || kd < 0 || ncols < 0 || cfg < 0 || cfg >= kNCfg) return MBEM_EINVAL;
    if (m == 0 || ncols == 0) return MBEM_OK;
    if (!A || !X || !Y || lda < kd || ldx < ncols || ldy < ncols) return MBEM_EINVAL;
    if ((lda & 1) || (ldx & 1) || !aligned16(A) || !aligned16(X)) return MBEM_EINVAL;
    const Cfg c = kCfgs[cfg];
    const long gx = (m + c.bm - 1) / c.bm, gy = (ncols + c.bn - 1) / c.bn;
    if (gx > INT32_MAX || gy > 65535) return MBEM_EINVAL;
    return dispatch<false, false>(cfg, dim3((unsigned)gx, (unsigned)gy), stream, A, nullptr, lda, m,
                                  kd, X, ldx, ncols, nullptr, Y, ldy, alpha, beta, nullptr);
}

int mbem_gemm_g(const double* A0, const double* A1, int64_t lda, int64_t m, int64_t kd,
                const double* X, int64_t ldx, int64_t ncols, const double* g, double* Y,
                int64_t ldy, double alpha, double beta, int cfg, hipStream_t stream)
{
    if (m < 0 || kd < 0 || ncols < 0 || cfg < 0 || cfg >= kNCfg) return MBEM_EINVAL;
    if (m == 0 || ncols == 0) return MBEM_OK;
    if (!A0 || !A1 || !X || !g || !Y || lda < kd || ldx < ncols || ldy < ncols) return MBEM_EINVAL;
    // 16-byte vector loads: even leading dimensions and aligned bases
    if ((lda & 1) || (ldx & 1) || !aligned16(A0) || !aligned16(A1) || !aligned16(X)) return MBEM_EINVAL;
    const Cfg c = kCfgs[cfg];
    const long gx = (m + c.bm - 1) / c.bm, gy = (ncols + c.bn - 1) / c.bn;
    if (gx > INT32_MAX || gy > 65535) return MBEM_EINVAL;
    return dispatch<true, false>(cfg, dim3((unsigned)gx, (unsigned)gy), stream, A0, A1, lda, m, kd,
                                 X, ldx, ncols, g, Y, ldy, alpha, beta, nullptr);
}

int mbem_gemm_grouped(const double* inv, const double* X, int64_t ldx, double* Y, int64_t ldy,
                      const void* work, int64_t nwork, int cfg, hipStream_t stream)
{
    if (nwork < 0 || cfg < 0 || cfg >= kNCfg) return MBEM_EINVAL;
    if (nwork == 0) return MBEM_OK;
    if (!inv || !X || !Y || !work || (ldx & 1) || !aligned16(X) || !aligned16(inv)) return MBEM_EINVAL;
    if (nwork > INT32_MAX) return MBEM_EINVAL;
    return dispatch<false, true>(cfg, dim3((unsigned)nwork), stream, inv, nullptr, 0, 0, 0, X, ldx,
                                 0, nullptr, Y, ldy, 1.0, 0.0,
                                 reinterpret_cast<const Work*>(work));
}

int mbem_work_item_bytes(void) { return (int)sizeof(Work); }

int mbem_dgemm_rowmajor(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                        hipStream_t stream)
{
    if (m < 0 || n < 0 || k < 0 || (m > 0 && n > 0 && (!A || !B || !C)) || lda < k || ldb < n || ldc < n)
        return MBEM_EINVAL;
    if (m == 0 || n == 0) return MBEM_OK;
    if (!g_blas && rocblas_create_handle(&g_blas) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_blas, stream) != rocblas_status_success) return MBEM_EHIP;
    // row-major C (m x n) = A (m x k) B (k x n)  <=>  column-major C^T = B^T A^T;
    // the 64-bit interface, since A (m x lda) passes 2^31 elements at n ~ 33,000
    const rocblas_status st = rocblas_dgemm_64(g_blas, rocblas_operation_none, rocblas_operation_none,
                                               n, m, k, &alpha, B, ldb, A, lda, &beta, C, ldc);
    return st == rocblas_status_success ? launch_status() : MBEM_EHIP;
}

int mbem_dgemm_abt(int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                   const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                   hipStream_t stream)
{
    if (m < 0 || n < 0 || k < 0 || (m > 0 && n > 0 && (!A || !B || !C)) || lda < k || ldb < k || ldc < n)
        return MBEM_EINVAL;
    if (m == 0 || n == 0) return MBEM_OK;
    if (!g_blas && rocblas_create_handle(&g_blas) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_blas, stream) != rocblas_status_success) return MBEM_EHIP;
    // row-major C (m x n) = A (m x k) B^T with B (n x k) row-major  <=>
    // column-major C^T = B A^T, B read transposed from its column-major view
    const rocblas_status st = rocblas_dgemm_64(g_blas, rocblas_operation_transpose, rocblas_operation_none,
                                               n, m, k, &alpha, B, ldb, A, lda, &beta, C, ldc);
    return st == rocblas_status_success ? launch_status() : MBEM_EHIP;
}

int64_t mbem_gram_work(int64_t k, int64_t n)
{
    if (k <= 0 || n <= 0) return 1;
    const int64_t S = n / 1024 > 0 ? (n / 1024 < 128 ? n / 1024 : 128) : 1;
    return S * k * k;
}

int mbem_gram(const double* X, int64_t ldx, int64_t k, int64_t n, double* G, double* work,
              hipStream_t stream)
{
    if (k < 0 || n < 0 || ldx < n || (k > 0 && (!X || !G || !work))) return MBEM_EINVAL;
    if (k == 0) return MBEM_OK;
    if (!g_blas && rocblas_create_handle(&g_blas) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_blas, stream) != rocblas_status_success) return MBEM_EHIP;
    const double one = 1.0, zero = 0.0;
    // X (k x n row-major) is column-major (n x k) with leading dimension ldx:
    // G = X X^T is that matrix's (k x k) Gram.  Split the n-long inner
    // dimension into S equal slices, one strided-batched DGEMM over them
    // (each slice a k x k partial), summed; the remainder columns are added
    // by one more GEMM with beta = 1.
    const int64_t S = n / 1024 > 0 ? (n / 1024 < 128 ? n / 1024 : 128) : 1;
    const int64_t L = n / S;
    rocblas_status st = rocblas_dgemm_strided_batched_64(
        g_blas, rocblas_operation_transpose, rocblas_operation_none, k, k, L, &one, X, ldx, L, X, ldx,
        L, &zero, work, k, k * k, S);
    if (st != rocblas_status_success) return MBEM_EHIP;
    const long len = (long)(k * k);
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, stream, work,
                       len, (long)S, G);
    if (S * L < n) {
        st = rocblas_dgemm_64(g_blas, rocblas_operation_transpose, rocblas_operation_none, k, k,
                              n - S * L, &one, X + S * L, ldx, X + S * L, ldx, &one, G, k);
        if (st != rocblas_status_success) return MBEM_EHIP;
    }
    return launch_status();
}

int mbem_coarse_factor(double* G, int64_t K, double delta, double* d, int32_t* info, double* Tt,
                       double* T, hipStream_t stream)
{
    if (K < 0 || K > INT32_MAX || (K > 0 && (!G || !d || !info || !Tt || !T)) || !(delta > 0.0))
        return MBEM_EINVAL;
    if (K == 0) return MBEM_OK;
    if (!g_blas && rocblas_create_handle(&g_blas) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_blas, stream) != rocblas_status_success) return MBEM_EHIP;
    const long KK = (long)(K * K);
    hipLaunchKernelGGL(k_gram_diag, dim3((unsigned)((K + 255) / 256)), dim3(256), 0, stream, G, (long)K, d);
    hipLaunchKernelGGL(k_gram_scale, dim3((unsigned)((KK + 255) / 256)), dim3(256), 0, stream, G, (long)K,
                       d, delta);
    // G symmetric: its row-major and column-major views are the same matrix
    if (rocsolver_dpotrf(g_blas, rocblas_fill_lower, (rocblas_int)K, G, (rocblas_int)K, info) !=
        rocblas_status_success)
        return MBEM_EHIP;
    if (rocsolver_dtrtri(g_blas, rocblas_fill_lower, rocblas_diagonal_non_unit, (rocblas_int)K, G,
                         (rocblas_int)K, info + 1) != rocblas_status_success)
        return MBEM_EHIP;
    hipLaunchKernelGGL(k_coarse_tfactors, dim3((unsigned)((KK + 255) / 256)), dim3(256), 0, stream, G,
                       (long)K, d, Tt, T);
    return launch_status();
}

int mbem_bdot(const double* V, int64_t bstride, int64_t cnt, const double* W, int64_t n, int64_t K,
              double* partial, int64_t nslots, hipStream_t stream)
{
    if (cnt < 1 || cnt > 8 || n < 1 || K < 1 || nslots < 4 || (nslots & 3) || !V || !W || !partial)
        return MBEM_EINVAL;
    const long rpw = (n + nslots - 1) / nslots;
    dim3 grid((unsigned)((K + 63) / 64), (unsigned)(nslots / 4));
    hipLaunchKernelGGL(k_bdot, grid, dim3(256), 0, stream, V, bstride, cnt, W, n, K, rpw, partial);
    return launch_status();
}

int mbem_breduce(const double* partial, int64_t nslots, int64_t cnt, int64_t K, double* out,
                 int64_t ldo, int64_t o0, int accumulate, double* copy, hipStream_t stream)
{
    if (cnt < 1 || K < 1 || nslots < 1 || !partial || !out) return MBEM_EINVAL;
    const long tot = cnt * K;
    hipLaunchKernelGGL(k_breduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, partial,
                       nslots, cnt, K, out, ldo, o0, accumulate, copy);
    return launch_status();
}

int mbem_baxpy(const double* B, int64_t bstride, int64_t cnt, const int* cnt_dev,
               const double* coef, int64_t ldc, double sign, double* Y, int64_t n, int64_t K,
               hipStream_t stream)
{
    if (n < 0 || K < 1 || cnt < 0 || !B || !coef || !Y) return MBEM_EINVAL;
    const long tot = n * K;
    if (tot == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_baxpy, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, B, bstride,
                       cnt, cnt_dev, coef, ldc, sign, Y, n, K);
    return launch_status();
}

int mbem_bscale(const double* X, const double* scale, double* Y, int64_t n, int64_t K,
                hipStream_t stream)
{
    if (n < 0 || K < 1 || !X || !scale || !Y) return MBEM_EINVAL;
    const long tot = n * K;
    if (tot == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_bscale, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, X, scale,
                       Y, n, K);
    return launch_status();
}

int mbem_bcopy(const double* X, double* Y, int64_t nk, hipStream_t stream)
{
    if (nk < 0 || (nk > 0 && (!X || !Y))) return MBEM_EINVAL;
    if (nk == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_bcopy, dim3((unsigned)((nk + 255) / 256)), dim3(256), 0, stream, X, Y, nk);
    return launch_status();
}

int mbem_fill(double* y, int64_t n, double v, hipStream_t stream)
{
    if (n < 0 || (n > 0 && !y)) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, y, n, v);
    return launch_status();
}

int mbem_fg_start(int64_t K, int64_t m, const double* sumsq, double* beta, double* scale,
                  double* gvec, int* jdone, hipStream_t stream)
{
    if (K < 1 || m < 1 || !sumsq || !beta || !scale || !gvec || !jdone) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_fg_start, dim3((unsigned)((K + 63) / 64)), dim3(64), 0, stream, K, m, sumsq,
                       beta, scale, gvec, jdone);
    return launch_status();
}

int mbem_fg_givens(int64_t K, int64_t m, int64_t j, const double* hp1, const double* hp2, double* H,
                   double* cs, double* sn, double* gvec, const double* sumsq, double* scale,
                   double* res, int* jdone, hipStream_t stream)
{
    if (K < 1 || m < 1 || j < 0 || j >= m || !hp1 || !hp2 || !H || !cs || !sn || !gvec || !sumsq ||
        !scale || !res || !jdone)
        return MBEM_EINVAL;
    hipLaunchKernelGGL(k_fg_givens, dim3((unsigned)((K + 63) / 64)), dim3(64), 0, stream, K, m, j, hp1,
                       hp2, H, cs, sn, gvec, sumsq, scale, res, jdone);
    return launch_status();
}

int mbem_fg_solve(int64_t K, int64_t m, const double* H, const double* gvec, const int* jdone,
                  double* y, hipStream_t stream)
{
    if (K < 1 || m < 1 || !H || !gvec || !jdone || !y) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_fg_solve, dim3((unsigned)((K + 63) / 64)), dim3(64), 0, stream, K, m, H, gvec,
                       jdone, y);
    return launch_status();
}

int mbem_gather_comb(const double* A0, const double* A1, int64_t lda, const int64_t* idx, int64_t m,
                     double g, double* out, hipStream_t stream)
{
    if (m < 0 || (m > 0 && (!A0 || !A1 || !idx || !out))) return MBEM_EINVAL;
    if (m == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_gather_comb, dim3((unsigned)((m * m + 255) / 256)), dim3(256), 0, stream, A0,
                       A1, lda, idx, m, g, out);
    return launch_status();
}

int mbem_gather_comb_pad(const double* A0, const double* A1, int64_t lda, const int64_t* idx,
                         int64_t m, int64_t mp, double g, double* out, hipStream_t stream)
{
    if (m < 0 || mp < m || (mp > 0 && (!out || (m > 0 && (!A0 || !A1 || !idx))))) return MBEM_EINVAL;
    if (mp == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_gather_comb_pad, dim3((unsigned)((mp * mp + 255) / 256)), dim3(256), 0, stream,
                       A0, A1, lda, idx, m, mp, g, out);
    return launch_status();
}

int mbem_permute(const double* A, int64_t lda, const int64_t* idx, int64_t n, double* out,
                 int64_t ldo, hipStream_t stream)
{
    if (n < 0 || (n > 0 && (!A || !idx || !out || ldo < n))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    if (n > INT32_MAX) return MBEM_EINVAL;
    long gx = (n + 255) / 256;
    if (gx > 64) gx = 64;
    const long gy = (n < 65535) ? n : 65535;
    hipLaunchKernelGGL(k_permute, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream, A, lda, idx, n,
                       out, ldo);
    return launch_status();
}

// ---- HIP graphs: capture a fixed kernel sequence on a stream, replay it ----

int mbem_graph_begin(hipStream_t stream)
{
    if (!stream) return MBEM_EINVAL;     // the legacy null stream cannot be captured
    return hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) == hipSuccess ? MBEM_OK
                                                                                          : MBEM_EHIP;
}

int mbem_graph_end(hipStream_t stream, void** exec_out)
{
    if (!stream || !exec_out) return MBEM_EINVAL;
    hipGraph_t graph = nullptr;
    if (hipStreamEndCapture(stream, &graph) != hipSuccess || !graph) return MBEM_EHIP;
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return MBEM_EHIP;
    *exec_out = exec;
    return MBEM_OK;
}

int mbem_graph_launch(void* exec, hipStream_t stream)
{
    if (!exec) return MBEM_EINVAL;
    return hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(exec), stream) == hipSuccess ? MBEM_OK
                                                                                        : MBEM_EHIP;
}

int mbem_graph_destroy(void* exec)
{
    if (!exec) return MBEM_OK;
    return hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(exec)) == hipSuccess ? MBEM_OK
                                                                                      : MBEM_EHIP;
}

int mbem_memcpy_async(void* dst, const void* src, int64_t bytes, hipStream_t stream)
{
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return MBEM_EINVAL;
    if (bytes == 0) return MBEM_OK;
    return hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault, stream) == hipSuccess ? MBEM_OK
                                                                                           : MBEM_EHIP;
}

}  // extern "C"
