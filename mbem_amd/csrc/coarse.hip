// Coarse space of the two-level block-Jacobi preconditioner on gfx950.
//
// W is n x K with K = G q: column g q + j lives on the DOFs of group g (one
// block-Jacobi chunk's DOFs of one slot kind) and holds there the j-th of q
// shape functions (constants, rotations, linears of the group's geometry).
// The DOFs are listed group by group in ``perm`` (gptr[g] .. gptr[g+1]) and
// Wp[t q + j] is the j-th value at the t-th listed DOF.  The kernels:
//   * the coarse image Ct = (A W)^T (K x n, row-major): one pass over A, a
//     workgroup per (group, 128 rows), the group's listed DOFs staged
//     through LDS for all of its rows (below);
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
// A workgroup takes one group g and one block of 4 R P consecutive rows;
// each of its four waves takes P passes of R rows.  Workgroups go to the
// eight XCDs in turn (block b to XCD b mod 8), so the block index is mapped
// for each XCD to run one row block's groups one after another: the groups
// of a row block read the same rows of the matrix (each in short runs), and
// these reads then meet in that XCD's L2 instead of each fetching its rows'
// lines from HBM again.  In a pass
// the group's listed positions go through LDS 256 at a time (their column
// and their q shape values, loaded once for the workgroup's four waves);
// each wave reads its R rows at those columns (ascending within a group, so
// the gathered reads walk each row forward in runs), keeps q partial sums
// per row in registers over the group's tiles and reduces them across the
// wave once per pass.  The group loop lives in the grid: each workgroup's
// chain of dependent tile loads is the group's few tiles times P, not every
// group's.  Rows past m are clamped to m - 1 and not stored; every thread
// reaches every barrier (the only exit is the whole workgroup's).  ROW loads
// one matrix entry as a double (the FP64 rows of A).
constexpr int kTile = 256;

template <int R, int Q, int P, class ROW>
__device__ __forceinline__ void coarse_image_wg(const ROW& row, long m,
                                                const int64_t* __restrict__ perm,
                                                const double* __restrict__ Wp,
                                                const int64_t* __restrict__ gptr, long G,
                                                double* __restrict__ Ct, long ldc)
{
    __shared__ long s_col[kTile];
    __shared__ double s_w[kTile * Q];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long b = blockIdx.x, local = b >> 3;
    const long g = local % G;
    const long rb = (b & 7) + 8 * (local / G);
    const long wg0 = rb * 4 * R * P;
    if (wg0 >= m) return;                       // the whole workgroup
    const long t0 = gptr[g], t1 = gptr[g + 1];
    for (int pass = 0; pass < P; ++pass) {
        const long r0 = wg0 + ((long)wave * P + pass) * R;
        long rr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) rr[r] = (r0 + r < m) ? r0 + r : m - 1;
        double acc[R][Q];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int j = 0; j < Q; ++j) acc[r][j] = 0.0;
        for (long tb = t0; tb < t1; tb += kTile) {
            const long nt = (t1 - tb < kTile) ? t1 - tb : kTile;
            __syncthreads();                    // the last tile's readers are done
            if (threadIdx.x < nt) s_col[threadIdx.x] = perm[tb + threadIdx.x];
            // shape values stored value-major (s_w[j][u]): a wave's reads of
            // one value are then lane-contiguous (position-major, 8 Q bytes
            // apart, they met in a quarter of the LDS banks)
            for (long e = threadIdx.x; e < nt * Q; e += blockDim.x) {
                const long u = e / Q, j = e - u * Q;
                s_w[j * kTile + u] = Wp[tb * Q + e];
            }
            __syncthreads();
            for (long u = lane; u < nt; u += 64) {
                const long col = s_col[u];
                double a[R];
#pragma unroll
                for (int r = 0; r < R; ++r) a[r] = row(rr[r], col);
#pragma unroll
                for (int j = 0; j < Q; ++j) {
                    const double w = s_w[j * kTile + u];
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

template <int R, int Q, int P>
__global__ __launch_bounds__(256) void k_coarse_image(const double* __restrict__ A, long lda, long m,
                                                      const int64_t* __restrict__ perm,
                                                      const double* __restrict__ Wp,
                                                      const int64_t* __restrict__ gptr, long G,
                                                      double* __restrict__ Ct, long ldc)
{
    coarse_image_wg<R, Q, P>(RowF64{A, lda}, m, perm, Wp, gptr, G, Ct, ldc);
}

// The same image for the mixed operator's far field F: int8 offset-binary
// codes (value = (code - 128) * scale[r, col / 64]), near entries zero.  Its
// own kernel, in FP32: the operator it images is the int8 field (1.8e-4 from
// A), so FP32 products and sums (1e-7) change nothing that matters, and they
// take fewer instructions per entry than the FP64 form (27 VALU and 8 LDS
// instructions an entry measured there, 7.2 ms at n = 60,000): each entry
// is one FP32 value, the weights are FP32 in LDS, the cross-lane sums FP32,
// and a wave's row bases are wave-uniform (scalar registers), so a load is
// that base plus a 32-bit column.  A workgroup holds up to 1,024 of its
// group's positions in LDS at once (a larger group goes in tiles) and takes
// 4 waves x P passes x R rows; blocks map to XCDs as in coarse_image_wg.
constexpr int kTileI8 = 1024;

template <int R, int Q, int P>
__global__ __launch_bounds__(256) void k_coarse_image_i8(const uint8_t* __restrict__ F, long ldf,
                                                         const float* __restrict__ S, long nseg,
                                                         long m, const int64_t* __restrict__ perm,
                                                         const double* __restrict__ Wp,
                                                         const int64_t* __restrict__ gptr, long G,
                                                         double* __restrict__ Ct, long ldc)
{
    __shared__ int s_col[kTileI8];
    __shared__ float s_w[Q][kTileI8];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long b = blockIdx.x, local = b >> 3;
    const long g = local % G;
    const long rb = (b & 7) + 8 * (local / G);
    const long wg0 = rb * 4 * R * P;
    if (wg0 >= m) return;                       // the whole workgroup
    const long t0 = gptr[g], t1 = gptr[g + 1];
    for (int pass = 0; pass < P; ++pass) {
        const long r0 = wg0 + ((long)wave * P + pass) * R;
        const uint8_t* fr[R];
        const float* sr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const long rr = (r0 + r < m) ? r0 + r : m - 1;
            fr[r] = F + rr * ldf;
            sr[r] = S + rr * nseg;
        }
        float acc[R][Q];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int j = 0; j < Q; ++j) acc[r][j] = 0.0f;
        for (long tb = t0; tb < t1; tb += kTileI8) {
            const int nt = (int)((t1 - tb < kTileI8) ? t1 - tb : kTileI8);
            // a group of at most kTileI8 positions is staged once for all passes
            if (pass == 0 || t1 - t0 > kTileI8) {
                __syncthreads();                // the last tile's readers are done
                for (int u = threadIdx.x; u < nt; u += blockDim.x) s_col[u] = (int)perm[tb + u];
                for (long e = threadIdx.x; e < (long)nt * Q; e += blockDim.x) {
                    const long u = e / Q, j = e - u * Q;
                    s_w[j][u] = (float)Wp[tb * Q + e];
                }
                __syncthreads();
            }
            for (int u = lane; u < nt; u += 64) {
                const int col = s_col[u];
                float a[R];
#pragma unroll
                for (int r = 0; r < R; ++r) a[r] = ((float)fr[r][col] - 128.0f) * sr[r][col >> 6];
#pragma unroll
                for (int j = 0; j < Q; ++j) {
                    const float w = s_w[j][u];
#pragma unroll
                    for (int r = 0; r < R; ++r) acc[r][j] += a[r] * w;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int j = 0; j < Q; ++j) {
                float v = acc[r][j];
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
                if (lane == 0 && r0 + r < m) Ct[(g * Q + j) * ldc + r0 + r] = (double)v;
            }
        }
    }
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

constexpr int kRows = 4;       // rows per pass of a wave
constexpr int kPasses = 8;     // passes per wave: 128 rows per workgroup

// row blocks rounded up to a multiple of the eight XCDs, times the groups
// (blocks past the last row exit at once)
inline int64_t image_blocks(int64_t m, int64_t G)
{
    const int64_t rows = 4 * kRows * kPasses;
    const int64_t nrb = (m + rows - 1) / rows;
    const int64_t b = ((nrb + 7) / 8) * 8 * G;
    return b > 0x7fffffff ? -1 : b;
}

inline dim3 image_grid(int64_t m, int64_t G) { return dim3((unsigned)image_blocks(m, G)); }

// the int8 image: 64 rows per workgroup (a row block's int8 rows, 3.8 MB at
// n = 60,000, fit one XCD's L2 while its groups read them)
constexpr int kPassesI8 = 4;

inline int64_t image_blocks_i8(int64_t m, int64_t G)
{
    const int64_t rows = 4 * kRows * kPassesI8;
    const int64_t nrb = (m + rows - 1) / rows;
    const int64_t b = ((nrb + 7) / 8) * 8 * G;
    return b > 0x7fffffff ? -1 : b;
}

inline dim3 image_grid_i8(int64_t m, int64_t G) { return dim3((unsigned)image_blocks_i8(m, G)); }

template <int Q>
int image_q(const double* A, int64_t lda, int64_t m, const int64_t* perm, const double* Wp,
            const int64_t* gptr, int64_t G, double* Ct, int64_t ldc, hipStream_t stream)
{
    hipLaunchKernelGGL((k_coarse_image<kRows, Q, kPasses>), image_grid(m, G), dim3(256), 0, stream,
                       A, lda, m, perm, Wp, gptr, (long)G, Ct, ldc);
    return launch_status();
}

template <int Q>
int image_i8_q(const uint8_t* F, int64_t ldf, const float* S, int64_t nseg, int64_t m,
               const int64_t* perm, const double* Wp, const int64_t* gptr, int64_t G, double* Ct,
               int64_t ldc, hipStream_t stream)
{
    hipLaunchKernelGGL((k_coarse_image_i8<kRows, Q, kPassesI8>), image_grid_i8(m, G), dim3(256), 0,
                       stream, F, ldf, S, nseg, m, perm, Wp, gptr, (long)G, Ct, ldc);
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

// The coarse projection with Q^T stored in FP32 (FP64 vectors and sums):
// s = Q^T r and rest = r - Q s.  Q^T is K x n row-major (a row per coarse
// direction).  part[c K + row] = sum over the 512 columns of chunk c of
// Q^T[row, j] r[j]: a wave per (row, chunk), 16-byte loads of four floats
// and two 16-byte loads of r when rows and r are 16-byte aligned.
constexpr long kQChunk = 512;

__global__ __launch_bounds__(256) void k_qdot_f32(const float* __restrict__ Q, long ldq, long K,
                                                  long n, const double* __restrict__ r,
                                                  double* __restrict__ part, int vec)
{
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long c = blockIdx.y;
    if (row >= K) return;
    const long j0 = c * kQChunk;
    const float* q = Q + row * ldq;
    double a = 0.0;
    if (vec && j0 + kQChunk <= n) {
#pragma unroll
        for (int t = 0; t < (int)(kQChunk / 256); ++t) {
            const long j = j0 + t * 256 + 4 * lane;
            const float4 qv = *reinterpret_cast<const float4*>(q + j);
            const double2 r0 = *reinterpret_cast<const double2*>(r + j);
            const double2 r1 = *reinterpret_cast<const double2*>(r + j + 2);
            a += ((double)qv.x * r0.x + (double)qv.y * r0.y) + ((double)qv.z * r1.x + (double)qv.w * r1.y);
        }
    } else {
        const long j1 = (j0 + kQChunk < n) ? j0 + kQChunk : n;
        for (long j = j0 + lane; j < j1; j += 64) a += (double)q[j] * r[j];
    }
    a = wave_sum(a);
    if (lane == 0) part[c * K + row] = a;
}

// s[row] = sum_c part[c K + row]: a wave per row, lanes over the parts
__global__ __launch_bounds__(256) void k_qreduce(const double* __restrict__ part, long K, long nparts,
                                                 double* __restrict__ s)
{
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (row >= K) return;
    double a = 0.0;
    for (long c = lane; c < nparts; c += 64) a += part[c * K + row];
    a = wave_sum(a);
    if (lane == 0) s[row] = a;
}

// rest[j] = r[j] - sum_row Q^T[row, j] s[row]: a workgroup per 64 columns,
// its four waves taking every fourth row, combined in LDS in a fixed order
__global__ __launch_bounds__(256) void k_qaxpy_f32(const float* __restrict__ Q, long ldq, long K,
                                                   long n, const double* __restrict__ s,
                                                   const double* __restrict__ r,
                                                   double* __restrict__ rest)
{
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const long j = (long)blockIdx.x * 64 + lane;
    double a = 0.0;
    if (j < n) {
#pragma unroll 4
        for (long row = wv; row < K; row += 4) a += (double)Q[row * ldq + j] * s[row];
    }
    red[wv][lane] = a;
    __syncthreads();
    if (wv == 0 && j < n) rest[j] = r[j] - ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]));
}

}  // namespace

int mbem_coarse_image(const double* A, int64_t lda, int64_t m, const int64_t* perm,
                      const double* Wp, const int64_t* gptr, int64_t G, int q, double* Ct,
                      int64_t ldc, hipStream_t stream)
{
    if (m < 0 || G < 0 || ldc < m || (q != 3 && q != 6 && q != 12)) return MBEM_EINVAL;
    if (image_blocks(m, G) < 0) return MBEM_EINVAL;            // past a 32-bit grid
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
    if (image_blocks_i8(m, G) < 0 || m > 0x7fffffffL) return MBEM_EINVAL;            // past a 32-bit grid
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

int64_t mbem_coarse_project_work(int64_t K, int64_t n)
{
    int64_t parts = (n + kQChunk - 1) / kQChunk;
    if (parts < 1) parts = 1;
    return parts * (K > 0 ? K : 1);
}

int mbem_coarse_project_f32(const float* Qt, int64_t ldq, int64_t K, int64_t n, const double* r,
                            double* s, double* rest, double* work, hipStream_t stream)
{
    if (K < 0 || n < 0 || (K > 0 && n > 0 && (!Qt || !r || !s || !rest || !work || ldq < n)))
        return MBEM_EINVAL;
    if (K == 0 || n == 0) return MBEM_OK;
    const int64_t parts = (n + kQChunk - 1) / kQChunk;
    if (parts > 65535 || (K + 3) / 4 > 0x7fffffffL) return MBEM_EINVAL;
    const int vec = (ldq % 4 == 0) && (reinterpret_cast<uintptr_t>(Qt) % 16 == 0) &&
                    (reinterpret_cast<uintptr_t>(r) % 16 == 0);
    hipLaunchKernelGGL(k_qdot_f32, dim3((unsigned)((K + 3) / 4), (unsigned)parts), dim3(256), 0, stream,
                       Qt, (long)ldq, (long)K, (long)n, r, work, vec);
    hipLaunchKernelGGL(k_qreduce, dim3((unsigned)((K + 3) / 4)), dim3(256), 0, stream, work, (long)K,
                       (long)parts, s);
    hipLaunchKernelGGL(k_qaxpy_f32, dim3((unsigned)((n + 63) / 64)), dim3(256), 0, stream, Qt, (long)ldq,
                       (long)K, (long)n, s, r, rest);
    return launch_status();
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
