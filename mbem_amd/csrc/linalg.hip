// FP64 linear algebra of the dense solve on gfx950:
//  * y = alpha A x + beta y, A row-major (the FGMRES matvec and the
//    Gauss-Seidel lower couplings): HBM-bound, one wavefront per R rows,
//    16-byte non-temporal loads of A, x reused from L2 across the R rows;
//  * split-K multi-dot h = V w and multi-axpy w += alpha V^T h, the two passes
//    of classical Gram-Schmidt (run twice) in the Arnoldi step;
//  * block-Jacobi: chunk gather + LAPACK getrf/getri (rocSOLVER) at build,
//    and one launch applying every chunk's explicit inverse with the
//    gather/scatter of its DOF list fused in.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <stdint.h>

#include <mutex>

#include "../../include/mbem_hip.h"

namespace {

typedef double dbl2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

inline int launch_status() { return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP; }

// y[r] = alpha * sum_j A[r, j] x[j] + beta * y[r]; VEC: 16-byte aligned rows
// and x.  One wave per R rows, each lane two 16-byte loads per row per step
// (columns 128 apart).  In the VEC form x goes through LDS 1,024 columns at a
// time, double-buffered: the workgroup's 256 threads load the next tile (two
// 16-byte loads each) while its four waves use the current one, then one
// barrier; each lane keeps the column order of the plain loop, so the sums
// are the same bit for bit (measured 6.4 -> 6.8 TB/s at n = 30,000, 6.6 ->
// 7.0 at 90,000, equal at 60,000).  No wave returns early: waves past the
// last row use clamped rows and store nothing, since every thread must reach
// the barriers.  Requires 256 threads per block.
template <int R, bool VEC>
__global__ __launch_bounds__(256) void k_gemv(const double* __restrict__ A, long lda, long m,
                                              long n, const double* __restrict__ x,
                                              double* __restrict__ y, double alpha, double beta)
{
    constexpr long kTile = 1024;
    __shared__ dbl2 xs[VEC ? 2 : 1][VEC ? kTile / 2 : 1];
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long r0 = wave * R;
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
    const double* rows[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rows[r] = A + ((r0 + r < m) ? r0 + r : m - 1) * lda;
    if (VEC) {
        const long n2 = n & ~1L;
        const long ntile = n2 / kTile;
        if (ntile > 0) {
            const dbl2* xp = reinterpret_cast<const dbl2*>(x);
            xs[0][threadIdx.x * 2] = xp[threadIdx.x * 2];
            xs[0][threadIdx.x * 2 + 1] = xp[threadIdx.x * 2 + 1];
            __syncthreads();
        }
        for (long k = 0; k < ntile; ++k) {
            const bool more = k + 1 < ntile;
            dbl2 xn0, xn1;
            if (more) {
                const dbl2* xp = reinterpret_cast<const dbl2*>(x + (k + 1) * kTile);
                xn0 = xp[threadIdx.x * 2];
                xn1 = xp[threadIdx.x * 2 + 1];
            }
            const dbl2* xt = xs[k & 1];
#pragma unroll
            for (int s = 0; s < (int)(kTile / 256); ++s) {
                const long j = k * kTile + 256 * s + 2 * lane;
                const dbl2 xa = xt[128 * s + lane], xb = xt[128 * s + 64 + lane];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    dbl2 a = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(rows[r] + j));
                    dbl2 b = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(rows[r] + j + 128));
                    acc[r] += a.x * xa.x + a.y * xa.y + b.x * xb.x + b.y * xb.y;
                }
            }
            if (more) {
                xs[(k + 1) & 1][threadIdx.x * 2] = xn0;
                xs[(k + 1) & 1][threadIdx.x * 2 + 1] = xn1;
                __syncthreads();
            }
        }
        long j = ntile * kTile + 2 * lane;
        for (; j + 128 < n2; j += 256) {
            dbl2 xa = *reinterpret_cast<const dbl2*>(x + j);
            dbl2 xb = *reinterpret_cast<const dbl2*>(x + j + 128);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                dbl2 a = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(rows[r] + j));
                dbl2 b = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(rows[r] + j + 128));
                acc[r] += a.x * xa.x + a.y * xa.y + b.x * xb.x + b.y * xb.y;
            }
        }
        for (; j < n2; j += 128) {
            dbl2 xa = *reinterpret_cast<const dbl2*>(x + j);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                dbl2 a = __builtin_nontemporal_load(reinterpret_cast<const dbl2*>(rows[r] + j));
                acc[r] += a.x * xa.x + a.y * xa.y;
            }
        }
        if ((n & 1) && lane == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += rows[r][n - 1] * x[n - 1];
        }
    } else {
        for (long j = lane; j < n; j += 64) {
            double xv = x[j];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += rows[r][j] * xv;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double t = wave_sum(acc[r]);
        if (lane == 0 && r0 + r < m) {
            double* p = y + r0 + r;
            *p = (beta == 0.0) ? alpha * t : alpha * t + beta * (*p);
        }
    }
}

// split-K partial dots: part[chunk * k + row] = sum_{j in chunk} V[row, j] w[j]
// part[c, row] = sum over columns of chunk c (kDotChunk wide) of V[row, j] w[j]:
// a wave per (row, chunk), 16-byte loads when rows and w are 16-byte aligned
constexpr long kDotChunk = 512;

__global__ __launch_bounds__(256) void k_multidot_part(const double* __restrict__ V, long ldv,
                                                       long k, long n, const double* __restrict__ w,
                                                       double* __restrict__ part, int vec)
{
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long c = blockIdx.y;
    if (row >= k) return;
    const long j0 = c * kDotChunk;
    const double* v = V + row * ldv;
    double a0 = 0.0, a1 = 0.0;
    if (vec && j0 + kDotChunk <= n) {
#pragma unroll
        for (int s = 0; s < (int)(kDotChunk / 128); ++s) {
            const long j = j0 + s * 128 + 2 * lane;
            const dbl2 vv = *reinterpret_cast<const dbl2*>(v + j);
            const dbl2 ww = *reinterpret_cast<const dbl2*>(w + j);
            a0 += vv.x * ww.x;
            a1 += vv.y * ww.y;
        }
    } else {
        const long j1 = (j0 + kDotChunk < n) ? j0 + kDotChunk : n;
        for (long j = j0 + lane; j < j1; j += 64) a0 += v[j] * w[j];
    }
    const double acc = wave_sum(a0 + a1);
    if (lane == 0) part[c * k + row] = acc;
}

// h[row] = sum_c part[c, row]: a wave per row, lanes over the parts
__global__ __launch_bounds__(256) void k_reduce_parts(const double* __restrict__ part, long k,
                                                      long nparts, double* __restrict__ h)
{
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= k) return;
    double s = 0.0;
    for (long c = lane; c < nparts; c += 64) s += part[c * k + row];
    s = wave_sum(s);
    if (lane == 0) h[row] = s;
}

// w[j] += alpha * sum_row V[row, j] h[row]: a workgroup per 64 columns, its
// four waves taking every fourth row, the partial sums combined in LDS in a
// fixed order
__global__ __launch_bounds__(256) void k_multiaxpy(const double* __restrict__ V, long ldv, long k,
                                                   long n, const double* __restrict__ h,
                                                   double* __restrict__ w, double alpha)
{
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long j = (long)blockIdx.x * 64 + lane;
    double s = 0.0;
    if (j < n) {
#pragma unroll 4
        for (long row = wv; row < k; row += 4) s += V[row * ldv + j] * h[row];
    }
    red[wv][lane] = s;
    __syncthreads();
    if (wv == 0 && j < n) w[j] += alpha * ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]));
}

__global__ __launch_bounds__(256) void k_axpby(long n, double a, const double* __restrict__ x,
                                               double b, double* __restrict__ y)
{
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    y[j] = (b == 0.0) ? a * x[j] : a * x[j] + b * y[j];
}

// w /= sqrt(*sumsq) when that norm exceeds 1e-300; the squared norm is read
// on the device, so the host need not see it before the next launch
__global__ __launch_bounds__(256) void k_normalize(long n, const double* __restrict__ sumsq,
                                                   double* __restrict__ w)
{
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double nrm = sqrt(fmax(*sumsq, 0.0));
    if (nrm > 1e-300) w[j] = w[j] / nrm;
}

// part[p, c] = sum over rows r of chunk p of |A[r, c]| (a thread per column,
// the wave's columns contiguous), then colsum[c] = sum_p part[p, c]: the
// column sums in a fixed order, so the 1-norm is reproducible
constexpr int kNormChunks = 64;

__global__ __launch_bounds__(256) void k_colsum_abs_part(const double* __restrict__ A, long lda,
                                                         long m, long n,
                                                         double* __restrict__ part)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const long rows = (m + kNormChunks - 1) / kNormChunks;
    const long r0 = blockIdx.y * rows;
    const long r1 = (r0 + rows < m) ? r0 + rows : m;
    double s = 0.0;
    for (long r = r0; r < r1; ++r) s += fabs(A[r * lda + c]);
    part[blockIdx.y * n + c] = s;
}

__global__ __launch_bounds__(256) void k_colsum_reduce(const double* __restrict__ part, long n,
                                                       double* __restrict__ colsum)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    double s = 0.0;
    for (int p = 0; p < kNormChunks; ++p) s += part[p * n + c];
    colsum[c] = s;
}

// chunk gather: M[i, j] = A[idx[i], idx[j]]
__global__ __launch_bounds__(256) void k_gather_block(const double* __restrict__ A, long lda,
                                                      const int64_t* __restrict__ idx, long m,
                                                      double* __restrict__ M)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * m) return;
    long i = t / m, j = t % m;
    M[t] = A[idx[i] * lda + idx[j]];
}

// padded chunk for a batched inverse: M (mp x mp) = [A[idx, idx] 0; 0 I],
// in FP64 or FP32 (F32 = true); the inverse's leading m x m block is the
// chunk's inverse (no pivot ever leaves the leading block: the identity
// rows are zero in its columns)
template <bool F32>
__global__ __launch_bounds__(256) void k_gather_pad(const double* __restrict__ A, long lda,
                                                    const int64_t* __restrict__ idx, long m,
                                                    long mp, void* __restrict__ Mv)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= mp * mp) return;
    long i = t / mp, j = t % mp;
    double v = (i < m && j < m) ? A[idx[i] * lda + idx[j]] : (i == j ? 1.0 : 0.0);
    if (F32) reinterpret_cast<float*>(Mv)[t] = (float)v;
    else reinterpret_cast<double*>(Mv)[t] = v;
}

// dst (rows x cols, ld ldd) = src (ld lds), elements of esz bytes (4 or 8)
template <typename T>
__global__ __launch_bounds__(256) void k_copy2d(const T* __restrict__ src, long lds, T* __restrict__ dst,
                                                long ldd, long rows, long cols)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= rows * cols) return;
    long i = t / cols, j = t % cols;
    dst[i * ldd + j] = src[i * lds + j];
}

// block-Jacobi apply: wave w of the launch takes rows [row0, row0 + R) of
// chunk wt[w] (table rows: chunk id, first row); z[idx[i]] = sum_j
// inv[i, j] r[idx[j]] for each row i of that chunk.
template <int R>
__global__ __launch_bounds__(256) void k_bj_apply(const int64_t* __restrict__ wtab, long nwaves,
                                                  const int64_t* __restrict__ cptr,
                                                  const int64_t* __restrict__ idx,
                                                  const int64_t* __restrict__ iptr,
                                                  const double* __restrict__ inv,
                                                  const double* __restrict__ r,
                                                  double* __restrict__ z)
{
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= nwaves) return;
    const long c = wtab[2 * wave], row0 = wtab[2 * wave + 1];
    const long i0 = cptr[c], m = cptr[c + 1] - i0;
    const int64_t* id = idx + i0;
    const double* B = inv + iptr[c];
    double acc[R];
#pragma unroll
    for (int q = 0; q < R; ++q) acc[q] = 0.0;
    for (long j = lane; j < m; j += 64) {
        double rv = r[id[j]];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            long i = row0 + q;
            if (i < m) acc[q] += __builtin_nontemporal_load(B + i * m + j) * rv;
        }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
        double t = wave_sum(acc[q]);
        long i = row0 + q;
        if (lane == 0 && i < m) z[id[i]] = t;
    }
}


// ---- blocked triangular solves on the LU factors ---------------------------
// The row-major buffer that mbem_lu_factor factors is A^T in rocSOLVER's
// column-major view: A^T = P L U with L (unit lower) and U stored column-major
// (element (i, j) at LU[i + j ld]).  A x = b is U^T L^T P^T x = b: a forward
// solve with U^T and a backward one with L^T, both read row-wise (element
// (i, j) of U^T at LU[j + i ld]), then x = P z.  A^T x = b is P L U x = b:
// z = P^T b, then L forward and U backward, read column-wise.  Each solve
// walks kTB-row diagonal blocks; the diagonal blocks' inverses are formed
// once per factorization, so one step is one launch: every workgroup forms
// x_K = inv(T_KK) r_K itself and updates its own rows beyond K.
constexpr int kTB = 64;

// inverse of diagonal block blockIdx.x of L (LOWER, unit) or U, row-major
// per block: inv[(k kTB + i) kTB + j]; a partial last block is padded with
// the identity.  Thread t forms column t by substitution, in FP64 whatever
// the factors' type T (FP64, or FP32 factors refined by the caller).
template <typename T, bool LOWER>
__global__ __launch_bounds__(kTB) void k_tri_inv_diag(const T* __restrict__ LU, long ld, long n,
                                                      double* __restrict__ inv)
{
    __shared__ double Tb[kTB][kTB + 1];
    __shared__ double Y[kTB][kTB + 1];
    const int t = threadIdx.x;
    const long k0 = (long)blockIdx.x * kTB;
    const int nb = (n - k0 < kTB) ? (int)(n - k0) : kTB;
    for (int i = 0; i < kTB; ++i)
        Tb[i][t] = (i < nb && t < nb) ? (double)LU[(k0 + i) + (k0 + t) * ld] : (i == t ? 1.0 : 0.0);
    __syncthreads();
    if (LOWER) {
        for (int i = 0; i < kTB; ++i) {
            double s = (i == t) ? 1.0 : 0.0;
            for (int j = 0; j < i; ++j) s -= Tb[i][j] * Y[j][t];
            Y[i][t] = s;
        }
    } else {
        for (int i = kTB - 1; i >= 0; --i) {
            double s = (i == t) ? 1.0 : 0.0;
            for (int j = i + 1; j < kTB; ++j) s -= Tb[i][j] * Y[j][t];
            Y[i][t] = s / Tb[i][i];
        }
    }
    for (int i = 0; i < kTB; ++i) inv[((long)blockIdx.x * kTB + i) * kTB + t] = Y[i][t];
}

// one step over diagonal block k of T x = r.  ROWWISE: T(i, j) = LU[j + i ld]
// (U^T or L^T) and inv(T_KK) is the stored block inverse transposed; else
// T(i, j) = LU[i + j ld].  LOWER: the rows after K are updated (forward),
// else the rows before K (backward).  Workgroup 0 writes x_K; the updated
// rows are each owned by one workgroup.
template <typename T, bool LOWER, bool ROWWISE>
__global__ __launch_bounds__(256) void k_trsv_step(const T* __restrict__ LU, long ld, long n,
                                                   const double* __restrict__ inv, long k,
                                                   double* __restrict__ r, double* __restrict__ x)
{
    __shared__ double xk[kTB];
    const int t = threadIdx.x;
    const long k0 = k * kTB;
    const int nb = (n - k0 < kTB) ? (int)(n - k0) : kTB;
    {
        // x_K[i] = sum_j inv(T_KK)[i, j] r[k0 + j]: 4 lanes per row, 16 j each
        const int i = t >> 2, q = t & 3;
        const double* blk = inv + k * kTB * kTB;
        double s = 0.0;
#pragma unroll 4
        for (int jj = 0; jj < 16; ++jj) {
            const int j = q * 16 + jj;
            const double rj = (j < nb) ? r[k0 + j] : 0.0;
            s += (ROWWISE ? blk[j * kTB + i] : blk[i * kTB + j]) * rj;
        }
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if (q == 0) xk[i] = s;
    }
    __syncthreads();
    if (blockIdx.x == 0 && t < nb) x[k0 + t] = xk[t];
    // rows this workgroup updates
    const long lo = LOWER ? k0 + nb : 0;
    const long hi = LOWER ? n : k0;
    if (!ROWWISE) {
        const long i = lo + (long)blockIdx.x * 256 + t;
        if (i >= hi) return;
        const T* col = LU + i + k0 * ld;
        double s = 0.0;
        for (int j = 0; j < nb; ++j) s += (double)col[j * ld] * xk[j];
        r[i] -= s;
    } else {
        // 16 lanes per row, 4 consecutive entries a lane; 64 rows a workgroup
        const int g = t >> 4, l = t & 15;
        for (int pass = 0; pass < 4; ++pass) {
            const long i = lo + (long)blockIdx.x * 64 + pass * 16 + g;
            double s = 0.0;
            if (i < hi) {
                const T* row = LU + i * ld + k0;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int j = 4 * l + e;
                    if (j < nb) s += (double)row[j] * xk[j];
                }
            }
#pragma unroll
            for (int off = 8; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
            if (l == 0 && i < hi) r[i] -= s;
        }
    }
}

// out[i] = v[perm[i]] (GATHER) or out[perm[i]] = v[i]
template <bool GATHER>
__global__ __launch_bounds__(256) void k_permute_vec(const double* __restrict__ v,
                                                     const int64_t* __restrict__ perm, long n,
                                                     double* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (GATHER) out[i] = v[perm[i]];
    else out[perm[i]] = v[i];
}

template <typename T, bool LOWER, bool ROWWISE>
int trsv_blocked(const T* LU, long n, const double* inv, double* r, double* x, hipStream_t st)
{
    const long nblk = (n + kTB - 1) / kTB;
    for (long q = 0; q < nblk; ++q) {
        const long k = LOWER ? q : nblk - 1 - q;
        const long k0 = k * kTB;
        const long nb = (n - k0 < kTB) ? n - k0 : kTB;
        const long rows = LOWER ? n - k0 - nb : k0;
        const long per = ROWWISE ? 64 : 256;
        const long g = rows > 0 ? (rows + per - 1) / per : 1;
        hipLaunchKernelGGL((k_trsv_step<T, LOWER, ROWWISE>), dim3((unsigned)g), dim3(256), 0, st, LU, n, n, inv, k, r, x);
    }
    return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP;
}

template <typename T>
int lu_prepare_impl(const T* LU, int64_t n, double* inv_l, double* inv_u, hipStream_t stream)
{
    if (n < 0 || (n > 0 && (!LU || !inv_l || !inv_u))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    const long nblk = (n + kTB - 1) / kTB;
    hipLaunchKernelGGL((k_tri_inv_diag<T, true>), dim3((unsigned)nblk), dim3(kTB), 0, stream, LU, n, n, inv_l);
    hipLaunchKernelGGL((k_tri_inv_diag<T, false>), dim3((unsigned)nblk), dim3(kTB), 0, stream, LU, n, n, inv_u);
    return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP;
}

template <typename T>
int lu_solve_blocked_impl(const T* LU, int64_t n, const int64_t* perm, const double* inv_l,
                          const double* inv_u, const double* b, double* x, double* work,
                          int adjoint, hipStream_t stream)
{
    if (n < 0 || (n > 0 && (!LU || !perm || !inv_l || !inv_u || !b || !x || !work))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    double* r = work;
    double* y = work + n;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    int rc;
    if (!adjoint) {
        // U^T w = b, L^T z = w, x = P z
        if (hipMemcpyAsync(r, b, n * sizeof(double), hipMemcpyDeviceToDevice, stream) != hipSuccess)
            return MBEM_EHIP;
        if ((rc = trsv_blocked<T, true, true>(LU, n, inv_u, r, y, stream)) != MBEM_OK) return rc;
        if ((rc = trsv_blocked<T, false, true>(LU, n, inv_l, y, r, stream)) != MBEM_OK) return rc;
        hipLaunchKernelGGL((k_permute_vec<false>), grid, block, 0, stream, r, perm, n, x);
    } else {
        // z = P^T b, L y = z, U x = y
        hipLaunchKernelGGL((k_permute_vec<true>), grid, block, 0, stream, b, perm, n, r);
        if ((rc = trsv_blocked<T, true, false>(LU, n, inv_l, r, y, stream)) != MBEM_OK) return rc;
        if ((rc = trsv_blocked<T, false, false>(LU, n, inv_u, y, x, stream)) != MBEM_OK) return rc;
    }
    return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP;
}

rocblas_handle g_handle = nullptr;

}  // namespace

extern "C" {

int mbem_gemv(const double* A, int64_t lda, int64_t m, int64_t n, const double* x, double* y,
              double alpha, double beta, hipStream_t stream)
{
    if (m < 0 || n < 0 || (m > 0 && (!A || !y || lda < n)) || (n > 0 && !x)) return MBEM_EINVAL;
    if (m == 0) return MBEM_OK;
    constexpr int R = 4;
    long waves = (m + R - 1) / R;
    dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    bool vec = ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && ((lda & 1) == 0)
               && ((reinterpret_cast<uintptr_t>(x) & 15) == 0);
    if (vec)
        hipLaunchKernelGGL((k_gemv<R, true>), grid, block, 0, stream, A, lda, m, n, x, y, alpha, beta);
    else
        hipLaunchKernelGGL((k_gemv<R, false>), grid, block, 0, stream, A, lda, m, n, x, y, alpha, beta);
    return launch_status();
}

int64_t mbem_multidot_work(int64_t k, int64_t n)
{
    int64_t parts = (n + kDotChunk - 1) / kDotChunk;
    if (parts < 1) parts = 1;
    return parts * (k > 0 ? k : 1);
}

int mbem_multidot(const double* V, int64_t ldv, int64_t k, int64_t n, const double* w, double* h,
                  double* work, hipStream_t stream)
{
    if (k < 0 || n < 0 || (k > 0 && (!V || !h || !work || ldv < n)) || (n > 0 && !w)) return MBEM_EINVAL;
    if (k == 0) return MBEM_OK;
    int64_t parts = (n + kDotChunk - 1) / kDotChunk;
    if (parts == 0) parts = 1;
    if (parts > 65535) return MBEM_EINVAL;      // grid.y
    const int vec = (ldv % 2 == 0) && (reinterpret_cast<uintptr_t>(V) % 16 == 0) &&
                    (reinterpret_cast<uintptr_t>(w) % 16 == 0);
    dim3 grid((unsigned)((k + 3) / 4), (unsigned)parts), block(256);
    hipLaunchKernelGGL(k_multidot_part, grid, block, 0, stream, V, ldv, k, n, w, work, vec);
    hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((k + 3) / 4)), dim3(256), 0, stream, work, k, parts, h);
    return launch_status();
}

// Scratch of mbem_multidot_noalloc: one call's partial sums, parts x k
// doubles (parts = ceil(n / kDotChunk)), in a fixed device array so that the
// call allocates nothing.  Each GPU holds its own copy until the process
// ends; its address on a device is read once per device (hipGetSymbolAddress
// with that device current) and kept: the array's name in host code is not a
// device address.  Every entry a call reads is written by the same call
// before it is read, so nothing needs clearing between calls; calls that
// overlap on one GPU (two streams, or two host threads) race on it.
static __device__ __align__(256) double scratch[MBEM_NOALLOC_SCRATCH_DOUBLES];

constexpr int kMaxDevices = 64;

static double* scratch_on_current_device()
{
    static std::mutex mu;
    static double* addr[kMaxDevices] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        addr[dev] = static_cast<double*>(p);
    }
    return addr[dev];
}

int mbem_multidot_noalloc(const double* V, int64_t ldv, int64_t k, int64_t n, const double* w,
                          double* h, hipStream_t stream)
{
    if (k < 0 || n < 0 || (k > 0 && (!V || !h || ldv < n)) || (n > 0 && !w)) return MBEM_EINVAL;
    if (k == 0) return MBEM_OK;
    int64_t parts = (n + kDotChunk - 1) / kDotChunk;
    if (parts == 0) parts = 1;
    if (parts > 65535) return MBEM_EINVAL;      // grid.y
    if (parts > MBEM_NOALLOC_SCRATCH_DOUBLES / k) return MBEM_ESCRATCH;
    double* work = scratch_on_current_device();
    if (!work) return MBEM_EHIP;
    const int vec = (ldv % 2 == 0) && (reinterpret_cast<uintptr_t>(V) % 16 == 0) &&
                    (reinterpret_cast<uintptr_t>(w) % 16 == 0);
    dim3 grid((unsigned)((k + 3) / 4), (unsigned)parts), block(256);
    hipLaunchKernelGGL(k_multidot_part, grid, block, 0, stream, V, ldv, k, n, w, work, vec);
    hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((k + 3) / 4)), dim3(256), 0, stream, work, k, parts, h);
    return launch_status();
}

int mbem_multiaxpy(const double* V, int64_t ldv, int64_t k, int64_t n, const double* h, double* w,
                   double alpha, hipStream_t stream)
{
    if (k < 0 || n < 0 || (k > 0 && n > 0 && (!V || !h || !w || ldv < n))) return MBEM_EINVAL;
    if (k == 0 || n == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_multiaxpy, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, stream, V, ldv, k, n, h, w, alpha);
    return launch_status();
}

int mbem_axpby(int64_t n, double a, const double* x, double b, double* y, hipStream_t stream)
{
    if (n < 0 || (n > 0 && (!x || !y))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n, a, x, b, y);
    return launch_status();
}

int mbem_normalize(int64_t n, const double* sumsq, double* w, hipStream_t stream)
{
    if (n < 0 || (n > 0 && (!sumsq || !w))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_normalize, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, n, sumsq, w);
    return launch_status();
}

int mbem_gather_block(const double* A, int64_t lda, const int64_t* idx, int64_t m, double* M,
                      hipStream_t stream)
{
    if (m < 0 || (m > 0 && (!A || !idx || !M))) return MBEM_EINVAL;
    if (m == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_gather_block, dim3((unsigned)((m * m + 255) / 256)), dim3(256), 0, stream, A, lda, idx, m, M);
    return launch_status();
}

int mbem_gather_block_pad(const double* A, int64_t lda, const int64_t* idx, int64_t m, int64_t mp,
                          void* M, int f32, hipStream_t stream)
{
    if (m < 0 || mp < m || (mp > 0 && (!M || (m > 0 && (!A || !idx))))) return MBEM_EINVAL;
    if (mp == 0) return MBEM_OK;
    const unsigned nb = (unsigned)((mp * mp + 255) / 256);
    if (f32) hipLaunchKernelGGL(k_gather_pad<true>, dim3(nb), dim3(256), 0, stream, A, lda, idx, m, mp, M);
    else hipLaunchKernelGGL(k_gather_pad<false>, dim3(nb), dim3(256), 0, stream, A, lda, idx, m, mp, M);
    return launch_status();
}

int mbem_copy2d(const void* src, int64_t lds, void* dst, int64_t ldd, int64_t rows, int64_t cols,
                int esz, hipStream_t stream)
{
    if (rows < 0 || cols < 0 || (esz != 4 && esz != 8) || (rows > 0 && cols > 0 && (!src || !dst)))
        return MBEM_EINVAL;
    if (rows == 0 || cols == 0) return MBEM_OK;
    const unsigned nb = (unsigned)((rows * cols + 255) / 256);
    if (esz == 8)
        hipLaunchKernelGGL(k_copy2d<double>, dim3(nb), dim3(256), 0, stream, (const double*)src, lds,
                           (double*)dst, ldd, rows, cols);
    else
        hipLaunchKernelGGL(k_copy2d<float>, dim3(nb), dim3(256), 0, stream, (const float*)src, lds,
                           (float*)dst, ldd, rows, cols);
    return launch_status();
}

// In-place inverse of an m x m matrix (row-major or column-major alike: the
// inverse of the transpose is the transpose of the inverse).  ipiv: m int32
// on the device, info: 1 int32 on the device.
int mbem_invert(double* M, int64_t m, int32_t* ipiv, int32_t* info, hipStream_t stream)
{
    // rocblas_int interface: the m x m array must stay below 2^31 elements
    if (m < 0 || (m > 0 && (!M || !ipiv || !info)) || m * m > INT32_MAX) return MBEM_EINVAL;
    if (m == 0) return MBEM_OK;
    if (!g_handle && rocblas_create_handle(&g_handle) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_handle, stream) != rocblas_status_success) return MBEM_EHIP;
    rocblas_int mm = (rocblas_int)m;
    if (rocsolver_dgetrf(g_handle, mm, mm, M, mm, ipiv, info) != rocblas_status_success) return MBEM_EHIP;
    if (rocsolver_dgetri(g_handle, mm, M, mm, ipiv, info) != rocblas_status_success) return MBEM_EHIP;
    return launch_status();
}

// In-place inverses of `batch` m x m matrices stored back to back (stride
// m*m): one strided-batched getrf + getri, so the per-column panel kernels
// run once for the whole batch.  ipiv: batch*m int32, info: batch int32.
int mbem_invert_batched(double* M, int64_t m, int64_t batch, int32_t* ipiv, int32_t* info,
                        hipStream_t stream)
{
    // rocblas_int sizes per matrix (m * m < 2^31 elements); the batch stride is 64-bit
    if (m < 0 || batch < 0 || (m > 0 && batch > 0 && (!M || !ipiv || !info)) || m * m > INT32_MAX
        || batch > INT32_MAX)
        return MBEM_EINVAL;
    if (m == 0 || batch == 0) return MBEM_OK;
    if (!g_handle && rocblas_create_handle(&g_handle) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_handle, stream) != rocblas_status_success) return MBEM_EHIP;
    rocblas_int mm = (rocblas_int)m;
    rocblas_stride sA = (rocblas_stride)(m * m), sP = (rocblas_stride)m;
    if (rocsolver_dgetrf_strided_batched(g_handle, mm, mm, M, mm, sA, ipiv, sP, info,
                                         (rocblas_int)batch) != rocblas_status_success)
        return MBEM_EHIP;
    if (rocsolver_dgetri_strided_batched(g_handle, mm, M, mm, sA, ipiv, sP, info,
                                         (rocblas_int)batch) != rocblas_status_success)
        return MBEM_EHIP;
    return launch_status();
}

int mbem_invert_batched_oop(void* M, int64_t m, int64_t batch, int32_t* ipiv, int32_t* info,
                            void* C, int32_t f32, hipStream_t stream)
{
    // getrf in place, then the inverse into C by solving A C = I with the
    // factors (getri_outofplace): a few large triangular solves per batch
    // instead of getri's long sequence of small blocked steps
    if (m < 0 || batch < 0 || (m > 0 && batch > 0 && (!M || !ipiv || !info || !C)) ||
        m * m > INT32_MAX || batch > INT32_MAX || M == C)
        return MBEM_EINVAL;
    if (m == 0 || batch == 0) return MBEM_OK;
    if (!g_handle && rocblas_create_handle(&g_handle) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_handle, stream) != rocblas_status_success) return MBEM_EHIP;
    rocblas_int mm = (rocblas_int)m, nb = (rocblas_int)batch;
    rocblas_stride sA = (rocblas_stride)(m * m), sP = (rocblas_stride)m;
    rocblas_status st;
    if (f32) {
        float* A = static_cast<float*>(M);
        st = rocsolver_sgetrf_strided_batched(g_handle, mm, mm, A, mm, sA, ipiv, sP, info, nb);
        if (st == rocblas_status_success)
            st = rocsolver_sgetri_outofplace_strided_batched(g_handle, mm, A, mm, sA, ipiv, sP,
                                                             static_cast<float*>(C), mm, sA, info, nb);
    } else {
        double* A = static_cast<double*>(M);
        st = rocsolver_dgetrf_strided_batched(g_handle, mm, mm, A, mm, sA, ipiv, sP, info, nb);
        if (st == rocblas_status_success)
            st = rocsolver_dgetri_outofplace_strided_batched(g_handle, mm, A, mm, sA, ipiv, sP,
                                                             static_cast<double*>(C), mm, sA, info, nb);
    }
    if (st != rocblas_status_success) return MBEM_EHIP;
    return launch_status();
}

// LU factorization of a row-major n x n matrix in place (LAPACK getrf of its
// column-major view, i.e. of A^T), and the matching solve A x = b (getrs with
// transpose on that factor).  ipiv: n int32, info: 1 int32, both on device.
int mbem_lu_factor(double* A, int64_t n, int64_t* ipiv, int64_t* info, hipStream_t stream)
{
    if (n < 0 || (n > 0 && (!A || !ipiv || !info))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    if (!g_handle && rocblas_create_handle(&g_handle) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_handle, stream) != rocblas_status_success) return MBEM_EHIP;
    // the 64-bit interface: at the dense path's sizes n^2 exceeds 2^31 elements,
    // past what the rocblas_int interface indexes
    if (rocsolver_dgetrf_64(g_handle, n, n, A, n, ipiv, info) != rocblas_status_success) return MBEM_EHIP;
    return launch_status();
}

int mbem_lu_factor_f32(float* A, int64_t n, int64_t* ipiv, int64_t* info, hipStream_t stream)
{
    if (n < 0 || (n > 0 && (!A || !ipiv || !info))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    if (!g_handle && rocblas_create_handle(&g_handle) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_handle, stream) != rocblas_status_success) return MBEM_EHIP;
    if (rocsolver_sgetrf_64(g_handle, n, n, A, n, ipiv, info) != rocblas_status_success) return MBEM_EHIP;
    return launch_status();
}

int mbem_lu_solve(const double* LU, int64_t n, const int64_t* ipiv, double* b, int64_t nrhs,
                  hipStream_t stream)
{
    if (n < 0 || nrhs < 0 || (n > 0 && (!LU || !ipiv || !b))) return MBEM_EINVAL;
    if (n == 0 || nrhs == 0) return MBEM_OK;
    if (!g_handle && rocblas_create_handle(&g_handle) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_handle, stream) != rocblas_status_success) return MBEM_EHIP;
    if (rocsolver_dgetrs_64(g_handle, rocblas_operation_transpose, n, nrhs, const_cast<double*>(LU), n,
                            const_cast<int64_t*>(ipiv), b, n) != rocblas_status_success)
        return MBEM_EHIP;
    return launch_status();
}

int mbem_lu_solve_adjoint(const double* LU, int64_t n, const int64_t* ipiv, double* b,
                          int64_t nrhs, hipStream_t stream)
{
    if (n < 0 || nrhs < 0 || (n > 0 && (!LU || !ipiv || !b))) return MBEM_EINVAL;
    if (n == 0 || nrhs == 0) return MBEM_OK;
    if (!g_handle && rocblas_create_handle(&g_handle) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_handle, stream) != rocblas_status_success) return MBEM_EHIP;
    // the row-major factors are those of A^T in rocSOLVER's column-major view,
    // so A^T x = b is the untransposed solve
    if (rocsolver_dgetrs_64(g_handle, rocblas_operation_none, n, nrhs, const_cast<double*>(LU), n,
                            const_cast<int64_t*>(ipiv), b, n) != rocblas_status_success)
        return MBEM_EHIP;
    return launch_status();
}

int64_t mbem_colsum_work(int64_t n) { return n <= 0 ? 0 : kNormChunks * n; }

int mbem_colsum_abs(const double* A, int64_t lda, int64_t m, int64_t n, double* work,
                    double* colsum, hipStream_t stream)
{
    if (m < 0 || n < 0 || (n > 0 && (!A || !work || !colsum || lda < n))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    dim3 grid((unsigned)((n + 255) / 256), kNormChunks), block(256);
    hipLaunchKernelGGL(k_colsum_abs_part, grid, block, 0, stream, A, lda, m, n, work);
    hipLaunchKernelGGL(k_colsum_reduce, dim3((unsigned)((n + 255) / 256)), block, 0, stream, work, n, colsum);
    return launch_status();
}

int64_t mbem_lu_inv_doubles(int64_t n) { return n <= 0 ? 0 : ((n + kTB - 1) / kTB) * kTB * kTB; }

int mbem_lu_prepare(const double* LU, int64_t n, double* inv_l, double* inv_u, hipStream_t stream)
{
    const int rc = lu_prepare_impl(LU, n, inv_l, inv_u, stream);
    return rc != MBEM_OK ? rc : launch_status();
}

int mbem_lu_prepare_f32(const float* LU, int64_t n, double* inv_l, double* inv_u, hipStream_t stream)
{
    const int rc = lu_prepare_impl(LU, n, inv_l, inv_u, stream);
    return rc != MBEM_OK ? rc : launch_status();
}

int mbem_lu_solve_blocked(const double* LU, int64_t n, const int64_t* perm, const double* inv_l,
                          const double* inv_u, const double* b, double* x, double* work,
                          int adjoint, hipStream_t stream)
{
    const int rc = lu_solve_blocked_impl(LU, n, perm, inv_l, inv_u, b, x, work, adjoint, stream);
    return rc != MBEM_OK ? rc : launch_status();
}

int mbem_lu_solve_blocked_f32(const float* LU, int64_t n, const int64_t* perm, const double* inv_l,
                              const double* inv_u, const double* b, double* x, double* work,
                              int adjoint, hipStream_t stream)
{
    const int rc = lu_solve_blocked_impl(LU, n, perm, inv_l, inv_u, b, x, work, adjoint, stream);
    return rc != MBEM_OK ? rc : launch_status();
}

int mbem_bj_apply(const int64_t* wave_table, int64_t nwaves, const int64_t* chunk_ptr,
                  const int64_t* idx, const int64_t* inv_ptr, const double* inv, const double* r,
                  double* z, hipStream_t stream)
{
    if (nwaves < 0 || (nwaves > 0 && (!wave_table || !chunk_ptr || !idx || !inv_ptr || !inv || !r || !z))) return MBEM_EINVAL;
    if (nwaves == 0) return MBEM_OK;
    dim3 grid((unsigned)((nwaves + 3) / 4)), block(256);
    hipLaunchKernelGGL((k_bj_apply<MBEM_BJ_ROWS_PER_WAVE>), grid, block, 0, stream, wave_table, nwaves, chunk_ptr, idx, inv_ptr, inv, r, z);
    return launch_status();
}

}  // extern "C"
