// Batched inversion of the FP32 block-Jacobi chunks on gfx950: blocked
// Gauss-Jordan with partial pivoting, every chunk of the preconditioner in
// one launch sequence (two launches per 32-column panel), in place of a
// strided-batched getrf + getri per size class.
//
// Each chunk's working matrix W (m x m FP32, row pitch ldw, a multiple of 32,
// zero padding) is transformed in place.  Partial pivoting never moves a row
// in memory: a row map gives the physical row of each logical row, and a
// pivot step swaps two entries of it.  Per panel of 32 pivot columns:
//
//   * k_gj_panel (one workgroup of 512 threads per chunk): the m x 32 panel
//     is held in registers, one logical row in each of up to four slots per
//     thread.  Step t (pivot column kk = k + t): the largest |W[i, kk]| over
//     logical rows i >= kk (smallest i on ties) is the pivot p; its row,
//     scaled by r = 1 / pivot (r in the pivot column), becomes logical row
//     kk; old row kk becomes row p; every other row i loses f_i times the
//     scaled row and takes -f_i r in the pivot column.  After 32 steps the
//     panel columns hold the 32 steps' transform G; they go back to W, and
//     the pivot rows (every column) are saved.
//   * k_gj_update (64 x 64 tiles of every chunk): the same 32 steps on every
//     other column at once, W[i, j] += G[i, :] S[:, j] for logical rows i
//     outside the panel, W[i, j] = G[i, :] S[:, j] for the pivot rows (S the
//     saved pivot rows): a rank-32 update, FP32 FMAs.
//
// Then k_gj_sigma composes sigma, the pivots' column swaps applied in
// reverse order (the in-place Gauss-Jordan's row interchanges undone on the
// columns), and k_gj_finish writes the inverse: logical row i, column j of
// the result is W[map[i], sigma(j)], into the packed row-major inverse (pitch
// ldo, pads zero), a workgroup per 16 rows of a chunk.
//
// A zero pivot records info = its column + 1 (the first) and continues with
// r = 0, so nothing becomes NaN; the caller reads info and raises.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mbem_hip.h"

namespace {

constexpr int kB = 32;           // pivot columns per panel
constexpr int kNT = 512;         // threads of the panel and finish kernels
constexpr int kMaxM = 2048;      // rows held in registers: 4 slots of 512 threads
constexpr int kTile = 64;        // rows and columns of an update tile

typedef float flt4 __attribute__((ext_vector_type(4)));

// per-chunk descriptor (int64): offset of the chunk's DOF list in idx, size m,
// working matrix offset and pitch, output offset and pitch, row map / pivot
// offset, saved pivot rows offset (kB x ldw floats)
enum { D_SRC, D_M, D_WOFF, D_LDW, D_OOFF, D_LDO, D_MAP, D_SOFF, D_N };

inline int launch_status() { return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP; }

// W_c[i, j] = (float) A[idx_c[i], idx_c[j]] for i, j < m, zero for
// m <= j < ldw; map_c[i] = i; info_c = 0.  One wave per row.
__global__ __launch_bounds__(256) void k_gj_gather(const double* __restrict__ A, long lda,
                                                   const int64_t* __restrict__ idx,
                                                   const int64_t* __restrict__ desc,
                                                   float* __restrict__ W, int* __restrict__ rowmap,
                                                   int* __restrict__ info)
{
    const int64_t* d = desc + (long)blockIdx.y * D_N;
    const long m = d[D_M];
    const long i = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (blockIdx.x == 0 && threadIdx.x == 0) info[blockIdx.y] = 0;
    if (i >= m) return;
    const int64_t* id = idx + d[D_SRC];
    const double* a = A + id[i] * lda;
    const long ldw = d[D_LDW];
    float* w = W + d[D_WOFF] + i * ldw;
    for (long j = lane; j < ldw; j += 64) w[j] = (j < m) ? (float)a[id[j]] : 0.0f;
    if (lane == 0) rowmap[d[D_MAP] + i] = (int)i;
}

// (value, index) pair kept by the larger value, the smaller index on ties
__device__ __forceinline__ void arg_better(float& bv, int& bi, const float ov, const int oi)
{
    if (ov > bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
    }
}

// one panel of every chunk (see the file comment); R row slots per thread
template <int R>
__global__ __launch_bounds__(kNT) void k_gj_panel(const int64_t* __restrict__ desc, long k,
                                                  float* __restrict__ W, int* __restrict__ rowmap,
                                                  int* __restrict__ piv, float* __restrict__ saved,
                                                  int* __restrict__ info)
{
    const int64_t* d = desc + (long)blockIdx.x * D_N;
    const long m = d[D_M];
    if (k >= m) return;                      // the whole workgroup: no barrier is skipped
    const int bw = (int)((m - k < kB) ? m - k : kB);
    const long ldw = d[D_LDW];
    float* Wc = W + d[D_WOFF];
    int* map = rowmap + d[D_MAP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NW = kNT / 64;
    __shared__ float red_v[2][NW];
    __shared__ int red_i[2][NW];
    __shared__ float zs[2][kB];              // the scaled pivot row
    __shared__ float ys[2][kB];              // the old row kk (moves to the pivot's place)
    __shared__ int zr[2], yr[2];             // their physical rows
    __shared__ int prow[kB];                 // physical rows of the panel's pivot rows

    float P[R][kB];
    int rm[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const long i = tid + (long)kNT * s;
        rm[s] = 0;
#pragma unroll
        for (int c = 0; c < kB; ++c) P[s][c] = 0.0f;
        if (i < m) {
            rm[s] = map[i];
            const flt4* src = reinterpret_cast<const flt4*>(Wc + (long)rm[s] * ldw + k);
#pragma unroll
            for (int q = 0; q < kB / 4; ++q) {
                const flt4 v = src[q];
                P[s][4 * q] = v.x;
                P[s][4 * q + 1] = v.y;
                P[s][4 * q + 2] = v.z;
                P[s][4 * q + 3] = v.w;
            }
        }
    }

    // Step t works on register column 0: after every step each row's
    // registers rotate left by one (the step's column to the end), so the
    // loop body is the same for every step -- one copy of the code, static
    // register indexes -- and after kB rotations every column is back in
    // place.  (Unrolled, the 32 steps' code missed the instruction cache:
    // 11 us a step against about 1.)
#pragma unroll 1
    for (int t = 0; t < kB; ++t) {
        // bw is the same for the whole workgroup (the barriers inside are reached by all)
        if (t < bw) {
            const int par = t & 1;           // LDS buffers alternate between steps
            const long kk = k + t;
            // pivot: this thread's rows, the wave, then the eight waves' results
            float bv = -1.0f;
            int bi = 0x7fffffff;
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const long i = tid + (long)kNT * s;
                if (i < m && i >= kk) arg_better(bv, bi, fabsf(P[s][0]), (int)i);
            }
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1)
                arg_better(bv, bi, __shfl_xor(bv, off, 64), __shfl_xor(bi, off, 64));
            if (lane == 0) {
                red_v[par][wave] = bv;
                red_i[par][wave] = bi;
            }
            __syncthreads();
            float pv = red_v[par][0];
            int p = red_i[par][0];
#pragma unroll
            for (int w = 1; w < NW; ++w) arg_better(pv, p, red_v[par][w], red_i[par][w]);
            // no candidate compared (all NaN): row kk itself, flagged like a zero pivot
            const bool bad = (p < kk || p >= m);
            if (bad) p = (int)kk;
            // the pivot row's owner publishes it scaled; the owner of row kk its old row
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const long i = tid + (long)kNT * s;
                if (i == p) {
                    const float a = P[s][0];
                    float r = 0.0f;
                    if (a != 0.0f && !bad) r = 1.0f / a;
                    else atomicCAS(info + blockIdx.x, 0, (int)(kk + 1));
                    zs[par][0] = r;
#pragma unroll
                    for (int c = 1; c < kB; ++c) zs[par][c] = P[s][c] * r;
                    zr[par] = rm[s];
                }
                if (i == kk && p != kk) {
#pragma unroll
                    for (int c = 0; c < kB; ++c) ys[par][c] = P[s][c];
                    yr[par] = rm[s];
                }
            }
            if (tid == 0) piv[d[D_MAP] + kk] = p;
            __syncthreads();
            const float r = zs[par][0];
#pragma unroll
            for (int s = 0; s < R; ++s) {
                const long i = tid + (long)kNT * s;
                if (i >= m) continue;
                if (i == kk) {
#pragma unroll
                    for (int c = 0; c < kB; ++c) P[s][c] = zs[par][c];
                    rm[s] = zr[par];
                } else {
                    if (i == p) {            // p != kk here: old row kk moves to p
#pragma unroll
                        for (int c = 0; c < kB; ++c) P[s][c] = ys[par][c];
                        rm[s] = yr[par];
                    }
                    const float f = P[s][0];
#pragma unroll
                    for (int c = 1; c < kB; ++c) P[s][c] -= f * zs[par][c];
                    P[s][0] = -f * r;
                }
            }
        }
#pragma unroll
        for (int s = 0; s < R; ++s) {
            const float first = P[s][0];
#pragma unroll
            for (int c = 0; c < kB - 1; ++c) P[s][c] = P[s][c + 1];
            P[s][kB - 1] = first;
        }
    }

    // the panel (now the steps' transform) and the row map back; the pivot
    // rows' physical rows for the copy below
#pragma unroll
    for (int s = 0; s < R; ++s) {
        const long i = tid + (long)kNT * s;
        if (i >= m) continue;
        flt4* dst = reinterpret_cast<flt4*>(Wc + (long)rm[s] * ldw + k);
#pragma unroll
        for (int q = 0; q < kB / 4; ++q)
            dst[q] = flt4{P[s][4 * q], P[s][4 * q + 1], P[s][4 * q + 2], P[s][4 * q + 3]};
        map[i] = rm[s];
        if (i >= k && i < k + bw) prow[i - k] = rm[s];
    }
    __syncthreads();
    // the pivot rows, every column (the update reads them while it rewrites W)
    float* S = saved + d[D_SOFF];
    const long n4 = ldw / 4;
    for (long e = tid; e < (long)bw * n4; e += kNT) {
        const long t = e / n4, q = e - t * n4;
        reinterpret_cast<flt4*>(S + t * ldw)[q] = reinterpret_cast<const flt4*>(Wc + (long)prow[t] * ldw)[q];
    }
}

// the panel's steps on every other column (see the file comment): one
// 64 x 64 tile of logical rows x columns per workgroup, 4 x 4 per thread
__global__ __launch_bounds__(256) void k_gj_update(const int64_t* __restrict__ desc, long k, long tiles,
                                                   float* __restrict__ W, const int* __restrict__ rowmap,
                                                   const float* __restrict__ saved)
{
    const int64_t* d = desc + (long)blockIdx.y * D_N;
    const long m = d[D_M];
    const long i0 = (long)(blockIdx.x / tiles) * kTile, j0 = (long)(blockIdx.x % tiles) * kTile;
    if (k >= m || i0 >= m || j0 >= m) return;       // the whole workgroup
    const long kend = (k + kB < m) ? k + kB : m;    // panel columns [k, kend)
    const int bw = (int)(kend - k);
    const long ldw = d[D_LDW];
    float* Wc = W + d[D_WOFF];
    const int* map = rowmap + d[D_MAP];
    const float* S = saved + d[D_SOFF];
    __shared__ float gs[kTile][kB + 1];     // the tile's rows of the transform (padded)
    __shared__ float ss[kB][kTile];         // the saved pivot rows, the tile's columns
    __shared__ int pr[kTile];
    const int tid = threadIdx.x;
    for (int e = tid; e < kTile * (kB / 4); e += 256) {
        const int r = e / (kB / 4), q = e % (kB / 4);
        const long i = i0 + r;
        flt4 v = {0.0f, 0.0f, 0.0f, 0.0f};
        if (i < m) v = reinterpret_cast<const flt4*>(Wc + (long)map[i] * ldw + k)[q];
        gs[r][4 * q] = v.x;
        gs[r][4 * q + 1] = v.y;
        gs[r][4 * q + 2] = v.z;
        gs[r][4 * q + 3] = v.w;
    }
    for (int r = tid; r < kTile; r += 256) pr[r] = (i0 + r < m) ? map[i0 + r] : 0;
    for (int e = tid; e < kB * (kTile / 4); e += 256) {
        const int t = e / (kTile / 4), q = e % (kTile / 4);
        const long j = j0 + 4 * q;
        flt4 v = {0.0f, 0.0f, 0.0f, 0.0f};
        if (t < bw && j < ldw) v = *reinterpret_cast<const flt4*>(S + t * ldw + j);
        *reinterpret_cast<flt4*>(&ss[t][4 * q]) = v;
    }
    __syncthreads();
    const int tx = tid & 15, ty = tid >> 4;
    float acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0f;
#pragma unroll 8
    for (int t = 0; t < kB; ++t) {
        const flt4 b = *reinterpret_cast<const flt4*>(&ss[t][4 * tx]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float a = gs[4 * ty + u][t];
            acc[u][0] += a * b.x;
            acc[u][1] += a * b.y;
            acc[u][2] += a * b.z;
            acc[u][3] += a * b.w;
        }
    }
    // four columns j .. j + 3 lie all inside or all outside the panel (k is a
    // multiple of 4, and kend is too unless it is m, past which all is padding)
    const long j = j0 + 4 * tx;
    if (j >= m || (j >= k && j < kend)) return;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long i = i0 + 4 * ty + u;
        if (i >= m) continue;
        flt4* w = reinterpret_cast<flt4*>(Wc + (long)pr[4 * ty + u] * ldw + j);
        if (i >= k && i < kend) {
            *w = flt4{acc[u][0], acc[u][1], acc[u][2], acc[u][3]};
        } else {
            const flt4 c = *w;
            *w = flt4{c.x + acc[u][0], c.y + acc[u][1], c.z + acc[u][2], c.w + acc[u][3]};
        }
    }
}

// sigma, the pivots' column swaps composed in reverse order, written over
// the chunk's pivots (working memory).  One wave per chunk; the serial part
// runs on LDS.
__global__ __launch_bounds__(64) void k_gj_sigma(const int64_t* __restrict__ desc,
                                                 int* __restrict__ piv)
{
    const int64_t* d = desc + (long)blockIdx.x * D_N;
    const long m = d[D_M];
    __shared__ int sig[kMaxM];
    __shared__ int pv[kMaxM];
    int* p = piv + d[D_MAP];
    const int lane = threadIdx.x;
    for (long j = lane; j < m; j += 64) {
        sig[j] = (int)j;
        pv[j] = p[j];
    }
    __syncthreads();
    if (lane == 0) {
        for (long t = m - 1; t >= 0; --t) {
            const int q = pv[t];
            const int a = sig[t];
            sig[t] = sig[q];
            sig[q] = a;
        }
    }
    __syncthreads();
    for (long j = lane; j < m; j += 64) p[j] = sig[j];
}

// out_c[i, j] = W_c[map[i], sigma(j)] for rows i0 .. i0 + kFinRows of chunk
// blockIdx.y, sigma read from piv (k_gj_sigma) into LDS once per block
constexpr int kFinRows = 16;
__global__ __launch_bounds__(256) void k_gj_finish(const int64_t* __restrict__ desc,
                                                   const float* __restrict__ W,
                                                   const int* __restrict__ rowmap,
                                                   const int* __restrict__ sigma,
                                                   float* __restrict__ out)
{
    const int64_t* d = desc + (long)blockIdx.y * D_N;
    const long m = d[D_M];
    const long i0 = (long)blockIdx.x * kFinRows;
    if (i0 >= m) return;                       // uniform across the block
    __shared__ int sig[kMaxM];
    const int tid = threadIdx.x;
    for (long j = tid; j < m; j += 256) sig[j] = sigma[d[D_MAP] + j];
    __syncthreads();
    const long ldw = d[D_LDW], ldo = d[D_LDO];
    const float* Wc = W + d[D_WOFF];
    const int* map = rowmap + d[D_MAP];
    float* o = out + d[D_OOFF];
    const int lane = tid & 63, wave = tid >> 6;
    const long i1 = (i0 + kFinRows < m) ? i0 + kFinRows : m;
    for (long i = i0 + wave; i < i1; i += 4) {
        const float* w = Wc + (long)map[i] * ldw;
        for (long j = lane; j < ldo; j += 64) o[i * ldo + j] = (j < m) ? w[sig[j]] : 0.0f;
    }
}

// the host copy of the descriptors against the sizes the caller allocated
int check_desc(const int64_t* dh, int64_t nchunks, int64_t mmax, int64_t idx_len, int64_t w_len,
               int64_t map_len, int64_t saved_len, int64_t out_len)
{
    if (nchunks < 0 || nchunks > 65535 || mmax < 0 || mmax > kMaxM || (nchunks > 0 && !dh))
        return MBEM_EINVAL;
    for (int64_t c = 0; c < nchunks; ++c) {
        const int64_t* d = dh + c * D_N;
        const int64_t m = d[D_M], ldw = d[D_LDW], ldo = d[D_LDO];
        if (m < 0 || m > mmax || ldw < m || ldw % kB || ldw > (1 << 20) || ldo < m ||
            ldo > (1 << 20))
            return MBEM_EINVAL;
        if (d[D_SRC] < 0 || d[D_SRC] + m > idx_len) return MBEM_EINVAL;
        if (d[D_WOFF] < 0 || d[D_WOFF] % 4 || d[D_WOFF] + m * ldw > w_len) return MBEM_EINVAL;
        if (d[D_OOFF] < 0 || d[D_OOFF] + m * ldo > out_len) return MBEM_EINVAL;
        if (d[D_MAP] < 0 || d[D_MAP] + m > map_len) return MBEM_EINVAL;
        if (d[D_SOFF] < 0 || d[D_SOFF] % 4 || d[D_SOFF] + (int64_t)kB * ldw > saved_len)
            return MBEM_EINVAL;
    }
    return MBEM_OK;
}

}  // namespace

extern "C" {

int mbem_bj_desc_fields(void) { return D_N; }
int mbem_bj_panel_width(void) { return kB; }

int mbem_bj_gather_f32(const double* A, int64_t lda, const int64_t* idx, int64_t idx_len,
                       const int64_t* desc_host, const int64_t* desc, int64_t nchunks, int64_t mmax,
                       float* W, int64_t w_len, int32_t* rowmap, int64_t map_len, int32_t* info,
                       hipStream_t stream)
{
    int st = check_desc(desc_host, nchunks, mmax, idx_len, w_len, map_len, INT64_MAX, INT64_MAX);
    if (st != MBEM_OK) return st;
    if (nchunks == 0 || mmax == 0) return MBEM_OK;
    if (!A || !idx || !desc || !W || !rowmap || !info || lda < 1) return MBEM_EINVAL;
    if (reinterpret_cast<uintptr_t>(W) & 15) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_gj_gather, dim3((unsigned)((mmax + 3) / 4), (unsigned)nchunks), dim3(256), 0, stream,
                       A, (long)lda, idx, desc, W, rowmap, info);
    return launch_status();
}

int mbem_bj_invert_f32(const int64_t* desc_host, const int64_t* desc, int64_t nchunks, int64_t mmax,
                       float* W, int64_t w_len, int32_t* rowmap, int32_t* piv, int64_t map_len,
                       float* saved, int64_t saved_len, int32_t* info, float* out, int64_t out_len,
                       hipStream_t stream)
{
    int st = check_desc(desc_host, nchunks, mmax, INT64_MAX, w_len, map_len, saved_len, out_len);
    if (st != MBEM_OK) return st;
    if (nchunks == 0 || mmax == 0) return MBEM_OK;
    if (!desc || !W || !rowmap || !piv || !saved || !info || !out) return MBEM_EINVAL;
    if ((reinterpret_cast<uintptr_t>(W) & 15) || (reinterpret_cast<uintptr_t>(saved) & 15))
        return MBEM_EINVAL;
    const int R = (int)((mmax + kNT - 1) / kNT);
    const long tiles = (mmax + kTile - 1) / kTile;
    const dim3 pgrid((unsigned)nchunks), ugrid((unsigned)(tiles * tiles), (unsigned)nchunks);
    for (long k = 0; k < mmax; k += kB) {
        switch (R) {
        case 1: hipLaunchKernelGGL((k_gj_panel<1>), pgrid, dim3(kNT), 0, stream, desc, k, W, rowmap, piv, saved, info); break;
        case 2: hipLaunchKernelGGL((k_gj_panel<2>), pgrid, dim3(kNT), 0, stream, desc, k, W, rowmap, piv, saved, info); break;
        case 3: hipLaunchKernelGGL((k_gj_panel<3>), pgrid, dim3(kNT), 0, stream, desc, k, W, rowmap, piv, saved, info); break;
        default: hipLaunchKernelGGL((k_gj_panel<4>), pgrid, dim3(kNT), 0, stream, desc, k, W, rowmap, piv, saved, info); break;
        }
        hipLaunchKernelGGL(k_gj_update, ugrid, dim3(256), 0, stream, desc, k, tiles, W,
                           (const int*)rowmap, (const float*)saved);
    }
    // sigma composed once per chunk (into piv), then the rows copied by
    // every CU: one workgroup per 16 rows of a chunk
    hipLaunchKernelGGL(k_gj_sigma, pgrid, dim3(64), 0, stream, desc, piv);
    hipLaunchKernelGGL(k_gj_finish, dim3((unsigned)((mmax + kFinRows - 1) / kFinRows), (unsigned)nchunks),
                       dim3(256), 0, stream, desc, (const float*)W, (const int*)rowmap,
                       (const int*)piv, out);
    return launch_status();
}

}  // extern "C"
