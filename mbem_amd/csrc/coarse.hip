// Coarse space of the two-level block-Jacobi preconditioner on gfx950.
//
// W is n x K with K = G q: column g q + j lives on the DOFs of group g (one
// block-Jacobi chunk's DOFs of one slot kind) and holds there the j-th of q
// shape functions (constants, rotations, linears of the group's geometry).
// The DOFs are listed group by group in ``perm`` (gptr[g] .. gptr[g+1]) and
// Wp[t q + j] is the j-th value at the t-th listed DOF.  The kernels:
//   * the coarse image Ct = (A W)^T (K x n, row-major): one pass over A, a
//     workgroup per 4 R rows, the group's listed DOFs staged through LDS
//     for all of them (below);
//   * the same image of the mixed operator (int8 far field plus the FP64
//     near field's block-CSR: 8x fewer bytes than the FP64 pass);
//   * the expansion x[perm[t]] += alpha sum_j Wp[t q + j] u[grp[t] q + j],
//     a thread per listed DOF (perm is a permutation: no two threads write
//     the same entry).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mbem_hip.h"

namespace {

inline int launch_status() { return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP; }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Ct[(g q + j) ldc + r] = sum_{t in group g} A[r, perm[t]] Wp[t q + j].
// A workgroup of four waves takes 4 R consecutive rows, a wave R of them.
// The group's listed positions go through LDS 256 at a time (their column
// and their q shape values, loaded once for the workgroup's 4 R rows rather
// than once per wave); each wave then reads its R rows at those columns
// (ascending within a group, so the gathered reads walk each row forward in
// runs), keeps q partial sums per row and reduces them across the wave once
// per group.  Rows past m are clamped to m - 1 and not stored; every thread
// reaches every barrier (the only exit is the whole workgroup's).  ROW loads
// one matrix entry as a double: FP64 rows, or int8 offset-binary codes
// times their row's FP32 scale per 64 columns.
constexpr int kTile = 256;

template <int R, int Q, class ROW>
__device__ __forceinline__ void coarse_image_wg(const ROW& row, long m,
                                                const int64_t* __restrict__ perm,
                                                const double* __restrict__ Wp,
                                                const int64_t* __restrict__ gptr, long G,
                                                double* __restrict__ Ct, long ldc)
{
    __shared__ long s_col[kTile];
    __shared__ double s_w[kTile * Q];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long wg0 = (long)blockIdx.x * 4 * R;
    if (wg0 >= m) return;                       // the whole workgroup
    const long r0 = wg0 + (long)wave * R;
    long rr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rr[r] = (r0 + r < m) ? r0 + r : m - 1;
    for (long g = 0; g < G; ++g) {
        const long t0 = gptr[g], t1 = gptr[g + 1];
        double acc[R][Q];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int j = 0; j < Q; ++j) acc[r][j] = 0.0;
        for (long tb = t0; tb < t1; tb += kTile) {
            const long nt = (t1 - tb < kTile) ? t1 - tb : kTile;
            __syncthreads();                    // the last tile's readers are done
            if (threadIdx.x < nt) s_col[threadIdx.x] = perm[tb + threadIdx.x];
            for (long e = threadIdx.x; e < nt * Q; e += blockDim.x) s_w[e] = Wp[tb * Q + e];
            __syncthreads();
            for (long u = lane; u < nt; u += 64) {
                const long col = s_col[u];
                double a[R];
#pragma unroll
                for (int r = 0; r < R; ++r) a[r] = row(rr[r], col);
#pragma unroll
                for (int j = 0; j < Q; ++j) {
                    const double w = s_w[u * Q + j];
#pragma unroll
                    for (int r = 0; r < R; ++r) acc[r][j] += a[r] * w;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int j = 0; j < Q; ++j) {
                const double v = wave_sum(acc[r][j]);
                if (lane == 0 && r0 + r < m) Ct[(g * Q + j) * ldc + r0 + r] = v;
            }
        }
    }
}

struct RowF64 {
    const double* A;
    long lda;
    __device__ double operator()(long r, long c) const { return A[r * lda + c]; }
};

struct RowI8 {
    const uint8_t* F;
    long ldf;
    const float* S;
    long nseg;
    __device__ double operator()(long r, long c) const
    {
        return (double)((float)((int)F[r * ldf + c] - 128) * S[r * nseg + (c >> 6)]);
    }
};

template <int R, int Q>
__global__ __launch_bounds__(256) void k_coarse_image(const double* __restrict__ A, long lda, long m,
                                                      const int64_t* __restrict__ perm,
                                                      const double* __restrict__ Wp,
                                                      const int64_t* __restrict__ gptr, long G,
                                                      double* __restrict__ Ct, long ldc)
{
    coarse_image_wg<R, Q>(RowF64{A, lda}, m, perm, Wp, gptr, G, Ct, ldc);
}

// The same image for the mixed operator's far field F: int8 offset-binary
// codes (value = (code - 128) * scale[r, col / 64]), near entries zero.
template <int R, int Q>
__global__ __launch_bounds__(256) void k_coarse_image_i8(const uint8_t* __restrict__ F, long ldf,
                                                         const float* __restrict__ S, long nseg,
                                                         long m, const int64_t* __restrict__ perm,
                                                         const double* __restrict__ Wp,
                                                         const int64_t* __restrict__ gptr, long G,
                                                         double* __restrict__ Ct, long ldc)
{
    coarse_image_wg<R, Q>(RowI8{F, ldf, S, nseg}, m, perm, Wp, gptr, G, Ct, ldc);
}

// Ct += the image of the near field's FP64 3 x 3 block-CSR (block row i,
// block columns col[...]): a thread per matrix row 3 i + a walks its block
// row; DOF d = 3 jb + b sits at listed position inv[d] of group grp[inv[d]].
template <int Q>
__global__ __launch_bounds__(256) void k_coarse_image_bsr(long nbr, const int64_t* __restrict__ rowptr,
                                                          const int64_t* __restrict__ col,
                                                          const double* __restrict__ vals,
                                                          const int64_t* __restrict__ inv,
                                                          const int64_t* __restrict__ grp,
                                                          const double* __restrict__ Wp,
                                                          double* __restrict__ Ct, long ldc)
{
    const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= 3 * nbr) return;
    const long i = row / 3;
    const int a = (int)(row - 3 * i);
    for (long k = rowptr[i]; k < rowptr[i + 1]; ++k) {
        const long jb = col[k];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const double v = vals[9 * k + 3 * a + b];
            const long t = inv[3 * jb + b];
            const long g = grp[t];
#pragma unroll
            for (int j = 0; j < Q; ++j) Ct[(g * Q + j) * ldc + row] += v * Wp[t * Q + j];
        }
    }
}

template <int Q>
__global__ __launch_bounds__(256) void k_coarse_expand(const double* __restrict__ u,
                                                       const int64_t* __restrict__ perm,
                                                       const double* __restrict__ Wp,
                                                       const int64_t* __restrict__ grp, long nt,
                                                       double alpha, double* __restrict__ x)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    const long g = grp[t];
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < Q; ++j) s += Wp[t * Q + j] * u[g * Q + j];
    x[perm[t]] += alpha * s;
}

constexpr int kRows = 4;

template <int Q>
int image_q(const double* A, int64_t lda, int64_t m, const int64_t* perm, const double* Wp,
            const int64_t* gptr, int64_t G, double* Ct, int64_t ldc, hipStream_t stream)
{
    hipLaunchKernelGGL((k_coarse_image<kRows, Q>), dim3((unsigned)((m + 4 * kRows - 1) / (4 * kRows))),
                       dim3(256), 0, stream, A, lda, m, perm, Wp, gptr, G, Ct, ldc);
    return launch_status();
}

template <int Q>
int image_i8_q(const uint8_t* F, int64_t ldf, const float* S, int64_t nseg, int64_t m,
               const int64_t* perm, const double* Wp, const int64_t* gptr, int64_t G, double* Ct,
               int64_t ldc, hipStream_t stream)
{
    hipLaunchKernelGGL((k_coarse_image_i8<kRows, Q>), dim3((unsigned)((m + 4 * kRows - 1) / (4 * kRows))),
                       dim3(256), 0, stream, F, ldf, S, nseg, m, perm, Wp, gptr, G, Ct, ldc);
    return launch_status();
}

template <int Q>
int image_bsr_q(int64_t nbr, const int64_t* rowptr, const int64_t* col, const double* vals,
                const int64_t* inv, const int64_t* grp, const double* Wp, double* Ct, int64_t ldc,
                hipStream_t stream)
{
    hipLaunchKernelGGL((k_coarse_image_bsr<Q>), dim3((unsigned)((3 * nbr + 255) / 256)), dim3(256), 0,
                       stream, nbr, rowptr, col, vals, inv, grp, Wp, Ct, ldc);
    return launch_status();
}

template <int Q>
int expand_q(const double* u, const int64_t* perm, const double* Wp, const int64_t* grp, int64_t nt,
             double alpha, double* x, hipStream_t stream)
{
    hipLaunchKernelGGL((k_coarse_expand<Q>), dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, stream,
                       u, perm, Wp, grp, nt, alpha, x);
    return launch_status();
}

}  // namespace

int mbem_coarse_image(const double* A, int64_t lda, int64_t m, const int64_t* perm,
                      const double* Wp, const int64_t* gptr, int64_t G, int q, double* Ct,
                      int64_t ldc, hipStream_t stream)
{
    if (m < 0 || G < 0 || ldc < m || (q != 3 && q != 6 && q != 12)) return MBEM_EINVAL;
    if (m == 0 || G == 0) return MBEM_OK;
    if (!A || !perm || !Wp || !gptr || !Ct) return MBEM_EINVAL;
    switch (q) {
    case 3: return image_q<3>(A, lda, m, perm, Wp, gptr, G, Ct, ldc, stream);
    case 6: return image_q<6>(A, lda, m, perm, Wp, gptr, G, Ct, ldc, stream);
    default: return image_q<12>(A, lda, m, perm, Wp, gptr, G, Ct, ldc, stream);
    }
}

int mbem_coarse_image_i8(const uint8_t* F, int64_t ldf, const float* scale, int64_t nseg, int64_t m,
                         const int64_t* perm, const double* Wp, const int64_t* gptr, int64_t G, int q,
                         double* Ct, int64_t ldc, hipStream_t stream)
{
    if (m < 0 || G < 0 || ldc < m || nseg < 0 || (q != 3 && q != 6 && q != 12)) return MBEM_EINVAL;
    if (m == 0 || G == 0) return MBEM_OK;
    if (!F || !scale || !perm || !Wp || !gptr || !Ct) return MBEM_EINVAL;
    switch (q) {
    case 3: return image_i8_q<3>(F, ldf, scale, nseg, m, perm, Wp, gptr, G, Ct, ldc, stream);
    case 6: return image_i8_q<6>(F, ldf, scale, nseg, m, perm, Wp, gptr, G, Ct, ldc, stream);
    default: return image_i8_q<12>(F, ldf, scale, nseg, m, perm, Wp, gptr, G, Ct, ldc, stream);
    }
}

int mbem_coarse_image_bsr(int64_t nbr, const int64_t* rowptr, const int64_t* col, const double* vals,
                          const int64_t* inv, const int64_t* grp, const double* Wp, int q, double* Ct,
                          int64_t ldc, hipStream_t stream)
{
    if (nbr < 0 || ldc < 3 * nbr || (q != 3 && q != 6 && q != 12)) return MBEM_EINVAL;
    if (nbr == 0) return MBEM_OK;
    if (!rowptr || !col || !vals || !inv || !grp || !Wp || !Ct) return MBEM_EINVAL;
    switch (q) {
    case 3: return image_bsr_q<3>(nbr, rowptr, col, vals, inv, grp, Wp, Ct, ldc, stream);
    case 6: return image_bsr_q<6>(nbr, rowptr, col, vals, inv, grp, Wp, Ct, ldc, stream);
    default: return image_bsr_q<12>(nbr, rowptr, col, vals, inv, grp, Wp, Ct, ldc, stream);
    }
}

int mbem_coarse_expand(const double* u, const int64_t* perm, const double* Wp, const int64_t* grp,
                       int64_t nt, int q, double alpha, double* x, hipStream_t stream)
{
    if (nt < 0 || (q != 3 && q != 6 && q != 12)) return MBEM_EINVAL;
    if (nt == 0) return MBEM_OK;
    if (!u || !perm || !Wp || !grp || !x) return MBEM_EINVAL;
    switch (q) {
    case 3: return expand_q<3>(u, perm, Wp, grp, nt, alpha, x, stream);
    case 6: return expand_q<6>(u, perm, Wp, grp, nt, alpha, x, stream);
    default: return expand_q<12>(u, perm, Wp, grp, nt, alpha, x, stream);
    }
}
