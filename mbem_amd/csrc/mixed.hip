// Mixed-precision operator of the dense solve on gfx950.
//
// The FP64 matrix A splits as A = F + N: N holds the near-field 3x3 blocks
// (source triangle within k (eps + h) of the collocation point, self blocks
// included) in FP64 block-CSR; F is every other entry, stored in BF16 or FP32
// (or int8 with per-segment scales) in a dense row-major copy whose near
// entries are zero.  y = F x + N x: the far part reads 2, 4 (or 1) bytes per
// entry and accumulates in FP32, the near part is exact FP64.  Also here: the FP32 block-Jacobi inverses
// (strided-batched sgetrf/sgetri) and their apply.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <stdint.h>

#include "../../include/mbem_hip.h"

namespace {

typedef float flt4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));

inline int launch_status() { return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP; }

__device__ __forceinline__ float wave_sum_f(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// float -> bf16 bits, round to nearest even (finite inputs; NaN kept NaN)
__device__ __forceinline__ unsigned short f32_to_bf16(float f)
{
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf16_lo(unsigned int w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(unsigned int w) { return __uint_as_float(w & 0xffff0000u); }

// out[r, c] = lo(A[r, c]); BF16 when BF = true, else FP32
template <bool BF>
__global__ __launch_bounds__(256) void k_cast(const double* __restrict__ A, long lda, long m, long n,
                                              void* __restrict__ out, long ldo)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long r = blockIdx.y;
    for (long c = t * 2; c < n; c += (long)gridDim.x * blockDim.x * 2) {
        double a0 = A[r * lda + c];
        double a1 = (c + 1 < n) ? A[r * lda + c + 1] : 0.0;
        if (BF) {
            unsigned short* o = reinterpret_cast<unsigned short*>(out) + r * ldo + c;
            o[0] = f32_to_bf16((float)a0);
            if (c + 1 < n) o[1] = f32_to_bf16((float)a1);
        } else {
            float* o = reinterpret_cast<float*>(out) + r * ldo + c;
            o[0] = (float)a0;
            if (c + 1 < n) o[1] = (float)a1;
        }
    }
    (void)m;
}

// near blocks: vals[b] (3x3 FP64) = A[3 i + a, 3 col[b] + c] for block b of
// block-row i; the same entries of the low-precision copy are set to zero
template <bool BF>
__global__ __launch_bounds__(256) void k_near_extract(const double* __restrict__ A, long lda,
                                                      long nbr, const int64_t* __restrict__ rowptr,
                                                      const int64_t* __restrict__ col,
                                                      double* __restrict__ vals,
                                                      void* __restrict__ lo, long ldo)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nnzb = rowptr[nbr];
    if (t >= 9 * nnzb) return;
    long b = t / 9;
    int q = (int)(t % 9), a = q / 3, c = q % 3;
    // block-row of b: binary search in rowptr
    long lo_i = 0, hi_i = nbr;
    while (hi_i - lo_i > 1) {
        long mid = (lo_i + hi_i) >> 1;
        if (rowptr[mid] <= b) lo_i = mid; else hi_i = mid;
    }
    long r = 3 * lo_i + a, cc = 3 * col[b] + c;
    vals[t] = A[r * lda + cc];
    if (BF) reinterpret_cast<unsigned short*>(lo)[r * ldo + cc] = 0;
    else reinterpret_cast<float*>(lo)[r * ldo + cc] = 0.0f;
}

// y = alpha (F x) + beta y with F stored low precision and x in FP32;
// FP32 products and accumulation, FP64 result.  One wave per R rows; each
// lane keeps two 16-byte row loads in flight per row (columns c and c + 512
// for BF16), which measured fastest on gfx950 for n = 60,000 (6.4 TB/s; 6.8
// TB/s, 1.065 ms for BF16, once x was staged through LDS).
template <int R, bool BF>
__device__ __forceinline__ void lo_fma8(float& acc, const u32x4 w, const flt4 xa, const flt4 xb)
{
    acc += bf16_lo(w.x) * xa.x + bf16_hi(w.x) * xa.y + bf16_lo(w.y) * xa.z + bf16_hi(w.y) * xa.w
           + bf16_lo(w.z) * xb.x + bf16_hi(w.z) * xb.y + bf16_lo(w.w) * xb.z + bf16_hi(w.w) * xb.w;
}

__device__ __forceinline__ void lo_fma4(float& acc, const flt4 w, const flt4 xa)
{
    acc += w.x * xa.x + w.y * xa.y + w.z * xa.z + w.w * xa.w;
}

// x goes through LDS 128 vectors (1,024 BF16 or 512 FP32 columns) at a time,
// double-buffered, one barrier per tile, as in k_gemv_i8; each lane keeps the
// column order of the plain loop (bit-identical sums).  No wave returns
// early (clamped rows, guarded stores).  Requires 256 threads per block.
template <int R, bool BF>
__global__ __launch_bounds__(256) void k_gemv_lo(const void* __restrict__ F, long ldf, long m, long n,
                                                 const float* __restrict__ x,
                                                 double* __restrict__ y, double alpha, double beta)
{
    constexpr int kXq = BF ? 256 : 128;  // flt4 of x per tile
    __shared__ flt4 xs[2][kXq];
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long r0 = wave * R;
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0f;
    // 16 bytes per load: 8 BF16 or 4 FP32 entries
    constexpr int E = BF ? 8 : 4;
    const long nvec = n / E;
    const unsigned short* Fb = reinterpret_cast<const unsigned short*>(F);
    const float* Ff = reinterpret_cast<const float*>(F);
    long rows[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rows[r] = ((r0 + r < m) ? r0 + r : m - 1) * ldf;
    const long ntile = nvec / 128;
    if (ntile > 0) {
        if (threadIdx.x < kXq) xs[0][threadIdx.x] = reinterpret_cast<const flt4*>(x)[threadIdx.x];
        __syncthreads();
    }
    for (long k = 0; k < ntile; ++k) {
        const bool more = k + 1 < ntile;
        flt4 xn;
        if (more && threadIdx.x < kXq)
            xn = reinterpret_cast<const flt4*>(x + (k + 1) * 128 * E)[threadIdx.x];
        const long c = (k * 128 + lane) * E, c2 = c + 64 * E;
        const flt4* xt = xs[k & 1];
        if (BF) {
            u32x4 w[R], w2[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                w[r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(Fb + rows[r] + c));
                w2[r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(Fb + rows[r] + c2));
            }
            const flt4 xa = xt[2 * lane], xb = xt[2 * lane + 1];
            const flt4 xc = xt[128 + 2 * lane], xd = xt[128 + 2 * lane + 1];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                lo_fma8<R, BF>(acc[r], w[r], xa, xb);
                lo_fma8<R, BF>(acc[r], w2[r], xc, xd);
            }
        } else {
            flt4 w[R], w2[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                w[r] = __builtin_nontemporal_load(reinterpret_cast<const flt4*>(Ff + rows[r] + c));
                w2[r] = __builtin_nontemporal_load(reinterpret_cast<const flt4*>(Ff + rows[r] + c2));
            }
            const flt4 xa = xt[lane], xc = xt[64 + lane];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                lo_fma4(acc[r], w[r], xa);
                lo_fma4(acc[r], w2[r], xc);
            }
        }
        if (more) {
            if (threadIdx.x < kXq) xs[(k + 1) & 1][threadIdx.x] = xn;
            __syncthreads();
        }
    }
    long v = ntile * 128 + lane;
    for (; v + 64 < nvec; v += 128) {
        const long c = v * E, c2 = c + 64 * E;
        if (BF) {
            u32x4 w[R], w2[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                w[r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(Fb + rows[r] + c));
                w2[r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(Fb + rows[r] + c2));
            }
            const flt4 xa = *reinterpret_cast<const flt4*>(x + c), xb = *reinterpret_cast<const flt4*>(x + c + 4);
            const flt4 xc = *reinterpret_cast<const flt4*>(x + c2), xd = *reinterpret_cast<const flt4*>(x + c2 + 4);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                lo_fma8<R, BF>(acc[r], w[r], xa, xb);
                lo_fma8<R, BF>(acc[r], w2[r], xc, xd);
            }
        } else {
            flt4 w[R], w2[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                w[r] = __builtin_nontemporal_load(reinterpret_cast<const flt4*>(Ff + rows[r] + c));
                w2[r] = __builtin_nontemporal_load(reinterpret_cast<const flt4*>(Ff + rows[r] + c2));
            }
            const flt4 xa = *reinterpret_cast<const flt4*>(x + c), xc = *reinterpret_cast<const flt4*>(x + c2);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                lo_fma4(acc[r], w[r], xa);
                lo_fma4(acc[r], w2[r], xc);
            }
        }
    }
    for (; v < nvec; v += 64) {
        const long c = v * E;
        if (BF) {
            const flt4 xa = *reinterpret_cast<const flt4*>(x + c), xb = *reinterpret_cast<const flt4*>(x + c + 4);
#pragma unroll
            for (int r = 0; r < R; ++r)
                lo_fma8<R, BF>(acc[r], __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(Fb + rows[r] + c)), xa, xb);
        } else {
            const flt4 xa = *reinterpret_cast<const flt4*>(x + c);
#pragma unroll
            for (int r = 0; r < R; ++r)
                lo_fma4(acc[r], __builtin_nontemporal_load(reinterpret_cast<const flt4*>(Ff + rows[r] + c)), xa);
        }
    }
    // tail columns
    for (long c = nvec * E + lane; c < n; c += 64) {
        const float xc = x[c];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float a = BF ? bf16_lo((unsigned int)Fb[rows[r] + c]) : Ff[rows[r] + c];
            acc[r] += a * xc;
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        float t = wave_sum_f(acc[r]);
        if (lane == 0 && r0 + r < m) {
            double* p = y + r0 + r;
            *p = (beta == 0.0) ? alpha * (double)t : alpha * (double)t + beta * (*p);
        }
    }
}

// ---- int8 far field ------------------------------------------------------
// F[r, c] = s[r, c / 64] * (u[r, c] - 128): offset-binary bytes u = q + 128,
// q = rint(A / s) in [-127, 127], and one FP32 scale s = max|A| / 127 per row
// and 64-column segment.  Every entry's error is at most s / 2, a fixed
// fraction of the segment's largest entry (a floating 8-bit format instead
// errs by a fixed fraction of each entry, largest on the largest entries).
constexpr long kSeg = 64;

// With R3, the code's rounding remainder is kept too: d = A - s q is exact
// in FP64 (s q has at most 31 significant bits; for q != 0, A and s q are
// within a factor 2 of each other), and e = rint(d / s * 2^24), |e| <= 2^23
// - 1, so s (q + e 2^-24) is A to within s 2^-25 (and FP64's rounding of
// d / s): about 2e-10 of the segment's largest entry, from 4 bytes instead
// of 8.  The remainders are packed four to three 32-bit words, little-endian
// 24-bit two's complement: a row's group of columns 4 k .. 4 k + 3 is the 12
// bytes at 12 k.  (32-bit remainders kept the solution error of the FP64
// operator, 4.5e-9 at n = 60,000; 24 bits give 5.4e-9, 20 bits 4.6e-8.)
//
// one wave per (row, 1024 columns), in 4 passes of 256: lane l encodes 4
// consecutive columns, so a 64-column segment is one 16-lane row of the wave
// and its largest magnitude is a reduction inside that row; loads (32 bytes
// a lane), 4-byte code stores and 12-byte remainder stores are contiguous
// across the wave
__global__ __launch_bounds__(256) void k_cast_i8(const double* __restrict__ A, long lda, long m,
                                                 long n, unsigned char* __restrict__ out, long ldo,
                                                 float* __restrict__ scale, long nseg,
                                                 unsigned char* __restrict__ R3, long ldr)
{
    const int lane = threadIdx.x & 63;
    const long nblk = (n + 1023) / 1024;
    const long w = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (w >= m * nblk) return;
    const long r = w / nblk;
    const long cb = (w % nblk) * 1024;
    const double* a = A + r * lda;
    unsigned char* o = out + r * ldo;
    // 16-byte loads when the row start is 16-byte aligned (A itself is)
    const bool vec = (lda & 1) == 0 && (reinterpret_cast<uintptr_t>(A) & 15) == 0;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const long c = cb + 256 * p + 4 * lane;
        double v[4];
        if (vec && c + 3 < n) {
            const dbl2 a01 = *reinterpret_cast<const dbl2*>(a + c);
            const dbl2 a23 = *reinterpret_cast<const dbl2*>(a + c + 2);
            v[0] = a01.x; v[1] = a01.y; v[2] = a23.x; v[3] = a23.y;
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] = (c + i < n) ? a[c + i] : 0.0;
        }
        double amax = fmax(fmax(fabs(v[0]), fabs(v[1])), fmax(fabs(v[2]), fabs(v[3])));
        amax = fmax(amax, __shfl_xor(amax, 1, 64));
        amax = fmax(amax, __shfl_xor(amax, 2, 64));
        amax = fmax(amax, __shfl_xor(amax, 4, 64));
        amax = fmax(amax, __shfl_xor(amax, 8, 64));
        const float s = (float)(amax / 127.0);
        unsigned int q[4];
        int e[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            double t = (s > 0.0f) ? rint(v[i] / (double)s) : 0.0;
            t = fmin(fmax(t, -127.0), 127.0);
            q[i] = (unsigned int)((int)t + 128);
            double f = (s > 0.0f) ? rint((v[i] - (double)s * t) / (double)s * 16777216.0) : 0.0;
            f = fmin(fmax(f, -8388607.0), 8388607.0);
            e[i] = (int)f;
        }
        if (c + 3 < n) {
            *reinterpret_cast<unsigned int*>(o + c) = q[0] | (q[1] << 8) | (q[2] << 16) | (q[3] << 24);
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (c + i < n) o[c + i] = (unsigned char)q[i];
        }
        if (R3 && c < n) {
            // the whole group (entries past n are zero: their v is): the
            // row pitch covers 3 bytes for every column up to n rounded to 4
            const unsigned u0 = (unsigned)e[0] & 0xffffffu, u1 = (unsigned)e[1] & 0xffffffu;
            const unsigned u2 = (unsigned)e[2] & 0xffffffu, u3 = (unsigned)e[3] & 0xffffffu;
            unsigned* w = reinterpret_cast<unsigned*>(R3 + r * ldr + 3 * c);
            w[0] = u0 | (u1 << 24);
            w[1] = (u1 >> 8) | (u2 << 16);
            w[2] = (u2 >> 16) | (u3 << 8);
        }
        if ((lane & 15) == 0 && c < n) scale[r * nseg + c / kSeg] = s;
    }
}

// near blocks of A held out while the int8 copy is built: RESTORE = false
// copies them to vals and zeroes them in A; RESTORE = true writes them back
template <bool RESTORE>
__global__ __launch_bounds__(256) void k_near_hold(double* __restrict__ A, long lda, long nbr,
                                                   const int64_t* __restrict__ rowptr,
                                                   const int64_t* __restrict__ col,
                                                   double* __restrict__ vals)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nnzb = rowptr[nbr];
    if (t >= 9 * nnzb) return;
    long b = t / 9;
    int q = (int)(t % 9), a = q / 3, c = q % 3;
    long lo_i = 0, hi_i = nbr;
    while (hi_i - lo_i > 1) {
        long mid = (lo_i + hi_i) >> 1;
        if (rowptr[mid] <= b) lo_i = mid; else hi_i = mid;
    }
    double* p = A + (3 * lo_i + a) * lda + 3 * col[b] + c;
    if (RESTORE) {
        *p = vals[t];
    } else {
        vals[t] = *p;
        *p = 0.0;
    }
}

__device__ __forceinline__ float ub0(unsigned int w) { return (float)(w & 0xffu); }
__device__ __forceinline__ float ub1(unsigned int w) { return (float)((w >> 8) & 0xffu); }
__device__ __forceinline__ float ub2(unsigned int w) { return (float)((w >> 16) & 0xffu); }
__device__ __forceinline__ float ub3(unsigned int w) { return (float)(w >> 24); }

// sum over 16 offset-binary bytes (one 16-byte load) of u_j x_j
__device__ __forceinline__ float u8_dot16(const u32x4 w, const flt4 a, const flt4 b, const flt4 c,
                                          const flt4 d)
{
    return ub0(w.x) * a.x + ub1(w.x) * a.y + ub2(w.x) * a.z + ub3(w.x) * a.w + ub0(w.y) * b.x +
           ub1(w.y) * b.y + ub2(w.y) * b.z + ub3(w.y) * b.w + ub0(w.z) * c.x + ub1(w.z) * c.y +
           ub2(w.z) * c.z + ub3(w.z) * c.w + ub0(w.w) * d.x + ub1(w.w) * d.y + ub2(w.w) * d.z +
           ub3(w.w) * d.w;
}

__device__ __forceinline__ float sum16(const flt4 a, const flt4 b, const flt4 c, const flt4 d)
{
    return ((a.x + a.y) + (a.z + a.w)) + ((b.x + b.y) + (b.z + b.w)) + ((c.x + c.y) + (c.z + c.w)) +
           ((d.x + d.y) + (d.z + d.w));
}

// y = alpha (F x) + beta y, F int8 with per-segment scales, x FP32: one wave
// per R rows with U 16-byte loads per row in flight (columns 1024 apart).  The
// row bases are wave-uniform (scalar registers) and each load is that base
// plus one 32-bit per-lane offset, which keeps the address arithmetic off the
// vector ALU.  A lane's 16 columns lie in one 64-column segment, so its
// partial sum is s (sum u x - 128 sum x) with the lane's own scale.
//
// x goes through LDS a tile of 1024 U columns at a time, double-buffered: the
// workgroup's 256 threads load the next tile (U 16-byte loads each) while its
// four waves multiply the current one, then one barrier.  Every wave would
// otherwise read all of x from the caches for its R rows; staging cuts those
// global loads fourfold (0.593 against 0.621 ms at n = 60,000, bit-identical
// sums).  Waves past the last row keep going with clamped rows, since every
// thread must reach the barriers; the tile loop's trip count is the same for
// the whole workgroup.  Requires 256 threads per block.
template <int R, int U>
__global__ __launch_bounds__(256) void k_gemv_i8(const unsigned char* __restrict__ F, long ldf,
                                                 const float* __restrict__ S, long nseg, long m,
                                                 long n, const float* __restrict__ x,
                                                 double* __restrict__ y, double alpha, double beta)
{
    constexpr unsigned kTile = 64 * U;  // 16-column vectors per tile
    __shared__ flt4 xs[2][kTile * 4];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long r0 = ((long)blockIdx.x * (blockDim.x >> 6) + wv) * R;
    const unsigned char* fb[R];
    const float* sb[R];
    float acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        acc[r] = 0.0f;
        const long rr = (r0 + r < m) ? r0 + r : m - 1;
        fb[r] = F + rr * ldf;
        sb[r] = S + rr * nseg;
    }
    // columns within a row fit 32 bits (the entry point checks n < 2^31)
    const unsigned nvec = (unsigned)(n / 16);
    const unsigned ntile = nvec / kTile;
    if (ntile > 0) {
        const flt4* xp = reinterpret_cast<const flt4*>(x);
#pragma unroll
        for (int i = 0; i < U; ++i) xs[0][threadIdx.x * U + i] = xp[threadIdx.x * U + i];
        __syncthreads();
    }
    for (unsigned k = 0; k < ntile; ++k) {
        const unsigned base = k * kTile;
        const bool more = k + 1 < ntile;
        flt4 xn[U];
        if (more) {
            const flt4* xp = reinterpret_cast<const flt4*>(x + (size_t)(base + kTile) * 16);
#pragma unroll
            for (int i = 0; i < U; ++i) xn[i] = xp[threadIdx.x * U + i];
        }
        u32x4 w[U][R];
        float s[U][R];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const unsigned c = (base + 64 * u + lane) * 16;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                w[u][r] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(fb[r] + c));
                s[u][r] = sb[r][c / kSeg];
            }
        }
        const flt4* xt = xs[k & 1];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const flt4* xl = xt + (64 * u + lane) * 4;
            const flt4 xa = xl[0], xb = xl[1], xc = xl[2], xd = xl[3];
            const float o = 128.0f * sum16(xa, xb, xc, xd);
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += s[u][r] * (u8_dot16(w[u][r], xa, xb, xc, xd) - o);
        }
        if (more) {
#pragma unroll
            for (int i = 0; i < U; ++i) xs[(k + 1) & 1][threadIdx.x * U + i] = xn[i];
            __syncthreads();
        }
    }
    for (unsigned v = ntile * kTile + lane; v < nvec; v += 64) {
        const unsigned c = v * 16;
        const flt4* xp = reinterpret_cast<const flt4*>(x + c);
        const flt4 xa = xp[0], xb = xp[1], xc = xp[2], xd = xp[3];
        const float o = 128.0f * sum16(xa, xb, xc, xd);
#pragma unroll
        for (int r = 0; r < R; ++r)
            acc[r] += sb[r][c / kSeg] *
                      (u8_dot16(__builtin_nontemporal_load(reinterpret_cast<const u32x4*>(fb[r] + c)),
                                xa, xb, xc, xd) - o);
    }
    for (unsigned c = nvec * 16 + lane; c < (unsigned)n; c += 64) {
        const float xc = x[c];
#pragma unroll
        for (int r = 0; r < R; ++r)
            acc[r] += sb[r][c / kSeg] * (((float)fb[r][c] - 128.0f) * xc);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        float t = wave_sum_f(acc[r]);
        if (lane == 0 && r0 + r < m) {
            double* p = y + r0 + r;
            *p = (beta == 0.0) ? alpha * (double)t : alpha * (double)t + beta * (*p);
        }
    }
}

// a group's three remainder words (four 24-bit two's-complement values,
// little-endian) from a 4-byte aligned address
struct Rem3 {
    unsigned w0, w1, w2;
};

__device__ __forceinline__ Rem3 load_rem3(const unsigned char* p)
{
    const unsigned* q = reinterpret_cast<const unsigned*>(p);
    return Rem3{__builtin_nontemporal_load(q), __builtin_nontemporal_load(q + 1),
                __builtin_nontemporal_load(q + 2)};
}

// the i-th 24-bit value of a group, sign-extended
__device__ __forceinline__ int rem24(const Rem3& e, int i)
{
    const unsigned v = (i == 0) ? e.w0
                       : (i == 1) ? __builtin_amdgcn_alignbit(e.w1, e.w0, 24)
                       : (i == 2) ? __builtin_amdgcn_alignbit(e.w2, e.w1, 16)
                                  : (e.w2 >> 8);
    return ((int)(v << 8)) >> 8;
}

// sum over a group of four columns of (u - 128 + e 2^-24) x in FP64, given
// sx = the group's sum of x: the codes' offset is taken out once per group
__device__ __forceinline__ double group_i8e(const unsigned w, const Rem3& e, const dbl2 xa,
                                            const dbl2 xb, const double sx)
{
    const double ux = (double)(w & 0xffu) * xa.x + (double)((w >> 8) & 0xffu) * xa.y +
                      (double)((w >> 16) & 0xffu) * xb.x + (double)(w >> 24) * xb.y;
    const double ex = (double)rem24(e, 0) * xa.x + (double)rem24(e, 1) * xa.y +
                      (double)rem24(e, 2) * xb.x + (double)rem24(e, 3) * xb.y;
    return (ux - 128.0 * sx) + ex * 0x1p-24;
}

// y = alpha (F x) + beta y in FP64 for the split far field
// F[r, c] = s (u - 128 + e 2^-24) (codes u, 24-bit remainders e, the
// per-segment scales s of k_cast_i8), FP64 x: the outer operator of the
// mixed method.  One wave per R rows.  In each 1024-column step lane l takes
// four groups of four columns, c = c0 + 256 g + 4 l (g < 4): per group and
// row one 4-byte load of codes and three 4-byte loads of remainders,
// contiguous across the wave (256 and 768 bytes).  A group lies in segment c0 / 64 + 4 g + l / 16;
// a row's 16 scales of the step are one load (lane l holds segment
// c0 / 64 + l % 16) and each group takes its own by a lane permute.  x goes
// through LDS 2048 columns at a time, double-buffered as in k_gemv_i8 (the
// same barrier discipline: every wave runs the same tile loop, rows past m
// are clamped and not stored).  Requires 256 threads per block.
template <int R>
__global__ __launch_bounds__(256) void k_gemv_i8e(const unsigned char* __restrict__ F, long ldf,
                                                  const unsigned char* __restrict__ R3, long ldr,
                                                  const float* __restrict__ S, long nseg, long m,
                                                  long n, const double* __restrict__ x,
                                                  double* __restrict__ y, double alpha, double beta)
{
    constexpr unsigned kTile = 2048;            // columns of x per LDS buffer
    __shared__ dbl2 xs[2][kTile / 2];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long r0 = ((long)blockIdx.x * (blockDim.x >> 6) + wv) * R;
    const unsigned char* fb[R];
    const unsigned char* eb[R];
    const float* sb[R];
    double acc[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        acc[r] = 0.0;
        const long rr = (r0 + r < m) ? r0 + r : m - 1;
        fb[r] = F + rr * ldf;
        eb[r] = R3 + rr * ldr;
        sb[r] = S + rr * nseg;
    }
    // columns within a row fit 32 bits (the entry point checks n < 2^31)
    const unsigned ntile = (unsigned)(n / kTile);
    if (ntile > 0) {
        const dbl2* xp = reinterpret_cast<const dbl2*>(x);
#pragma unroll
        for (int i = 0; i < 4; ++i) xs[0][threadIdx.x + 256 * i] = xp[threadIdx.x + 256 * i];
        __syncthreads();
    }
    for (unsigned k = 0; k < ntile; ++k) {
        const unsigned base = k * kTile;
        const bool more = k + 1 < ntile;
        dbl2 xn[4];
        if (more) {
            const dbl2* xp = reinterpret_cast<const dbl2*>(x + base + kTile);
#pragma unroll
            for (int i = 0; i < 4; ++i) xn[i] = xp[threadIdx.x + 256 * i];
        }
        const dbl2* xt = xs[k & 1];
        // the two steps of a tile one after the other: unrolled, both steps'
        // loads are hoisted together and R = 4 rows take all 256 VGPRs (one
        // wave per SIMD) instead of 150 (three)
#pragma unroll 1
        for (int st = 0; st < 2; ++st) {
            const unsigned c0 = base + 1024 * st;
            unsigned w[4][R];
            Rem3 ev[4][R];
            float sv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                sv[r] = sb[r][c0 / 64 + (lane & 15)];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const unsigned c = c0 + 256 * g + 4 * lane;
                    w[g][r] = __builtin_nontemporal_load(reinterpret_cast<const unsigned*>(fb[r] + c));
                    ev[g][r] = load_rem3(eb[r] + 3 * c);
                }
            }
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const dbl2* xl = xt + (1024 * st + 256 * g + 4 * lane) / 2;
                const dbl2 xa = xl[0], xb = xl[1];
                const double sx = (xa.x + xa.y) + (xb.x + xb.y);
                const int src = 4 * g + (lane >> 4);
#pragma unroll
                for (int r = 0; r < R; ++r)
                    acc[r] += (double)__shfl(sv[r], src, 64) * group_i8e(w[g][r], ev[g][r], xa, xb, sx);
            }
        }
        if (more) {
#pragma unroll
            for (int i = 0; i < 4; ++i) xs[(k + 1) & 1][threadIdx.x + 256 * i] = xn[i];
            __syncthreads();
        }
    }
    // columns past the last tile: groups of four, x from global memory
    const unsigned ng = (unsigned)(n / 4);
    for (unsigned q = ntile * (kTile / 4) + lane; q < ng; q += 64) {
        const unsigned c = 4 * q;
        const dbl2* xp = reinterpret_cast<const dbl2*>(x + c);
        const dbl2 xa = xp[0], xb = xp[1];
        const double sx = (xa.x + xa.y) + (xb.x + xb.y);
#pragma unroll
        for (int r = 0; r < R; ++r)
            acc[r] += (double)sb[r][c / 64] *
                      group_i8e(*reinterpret_cast<const unsigned*>(fb[r] + c),
                                load_rem3(eb[r] + 3 * c), xa, xb, sx);
    }
    for (unsigned c = 4 * ng + lane; c < (unsigned)n; c += 64) {
        const double xc = x[c];
        const unsigned g4 = c & ~3u, i = c & 3u;
#pragma unroll
        for (int r = 0; r < R; ++r)
            acc[r] += (double)sb[r][c / 64] *
                      ((((double)fb[r][c] - 128.0) +
                        (double)rem24(load_rem3(eb[r] + 3 * g4), (int)i) * 0x1p-24) * xc);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const double t = wave_sum_d(acc[r]);
        if (lane == 0 && r0 + r < m) {
            double* p = y + r0 + r;
            *p = (beta == 0.0) ? alpha * t : alpha * t + beta * (*p);
        }
    }
}

// xf[i] = (float) x[i]
__global__ __launch_bounds__(256) void k_to_f32(const double* __restrict__ x, long n, float* __restrict__ xf)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xf[i] = (float)x[i];
}

// y[3i + a] += alpha * sum_b vals[b][a][c] x[3 col[b] + c]: one wave per block-row
__global__ __launch_bounds__(256) void k_bsr_gemv(long nbr, const int64_t* __restrict__ rowptr,
                                                  const int64_t* __restrict__ col,
                                                  const double* __restrict__ vals,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ y, double alpha)
{
    const int lane = threadIdx.x & 63;
    const long i = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= nbr) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (long b = rowptr[i] + lane; b < rowptr[i + 1]; b += 64) {
        const double* v = vals + 9 * b;
        const double* xs = x + 3 * col[b];
        double x0 = xs[0], x1 = xs[1], x2 = xs[2];
        a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
        a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
        a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
    a0 = wave_sum_d(a0);
    a1 = wave_sum_d(a1);
    a2 = wave_sum_d(a2);
    if (lane == 0) {
        y[3 * i] += alpha * a0;
        y[3 * i + 1] += alpha * a1;
        y[3 * i + 2] += alpha * a2;
    }
}

// FP32 chunk gather: M[i, j] = (float) A[idx[i], idx[j]]
__global__ __launch_bounds__(256) void k_gather_block_f32(const double* __restrict__ A, long lda,
                                                          const int64_t* __restrict__ idx, long m,
                                                          float* __restrict__ M)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * m) return;
    long i = t / m, j = t % m;
    M[t] = (float)A[idx[i] * lda + idx[j]];
}

// block-Jacobi apply with FP32 inverses (FP64 accumulation of FP32 products)
template <int R>
__global__ __launch_bounds__(256) void k_bj_apply_f32(const int64_t* __restrict__ wtab, long nwaves,
                                                      const int64_t* __restrict__ cptr,
                                                      const int64_t* __restrict__ idx,
                                                      const int64_t* __restrict__ iptr,
                                                      const float* __restrict__ inv,
                                                      const double* __restrict__ r,
                                                      double* __restrict__ z)
{
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= nwaves) return;
    const long c = wtab[2 * wave], row0 = wtab[2 * wave + 1];
    const long i0 = cptr[c], m = cptr[c + 1] - i0;
    const int64_t* id = idx + i0;
    const float* B = inv + iptr[c];
    double acc[R];
#pragma unroll
    for (int q = 0; q < R; ++q) acc[q] = 0.0;
    for (long j = lane; j < m; j += 64) {
        double rv = r[id[j]];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            long i = row0 + q;
            if (i < m) acc[q] += (double)__builtin_nontemporal_load(B + i * m + j) * rv;
        }
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
        double t = wave_sum_d(acc[q]);
        long i = row0 + q;
        if (lane == 0 && i < m) z[id[i]] = t;
    }
}

// FP32 block-Jacobi on padded rows: chunk c's inverse is stored with its m
// rows padded to ldm = pptr[c+1] - pptr[c] (a multiple of 4, pads zero), and
// the right-hand side is gathered once per apply into the same padded chunk
// order in FP32 (k_gather_f32), so each lane reads 16-byte float4s of both;
// the products accumulate in FP64.  One wave per R rows of one chunk.
__global__ __launch_bounds__(256) void k_gather_f32(const int64_t* __restrict__ gidx, long np,
                                                    const double* __restrict__ r,
                                                    float* __restrict__ rr)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int64_t g = gidx[p];
    rr[p] = (g >= 0) ? (float)r[g] : 0.0f;
}

template <int R>
__global__ __launch_bounds__(256) void k_bj_apply_f32v(const int64_t* __restrict__ wtab, long nwaves,
                                                       const int64_t* __restrict__ cptr,
                                                       const int64_t* __restrict__ pptr,
                                                       const int64_t* __restrict__ idx,
                                                       const int64_t* __restrict__ iptr,
                                                       const float* __restrict__ inv,
                                                       const float* __restrict__ rr,
                                                       double* __restrict__ z)
{
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (wave >= nwaves) return;
    const long c = wtab[2 * wave], row0 = wtab[2 * wave + 1];
    const long m = cptr[c + 1] - cptr[c];
    const long ldm = pptr[c + 1] - pptr[c];
    const flt4* x = reinterpret_cast<const flt4*>(rr + pptr[c]);
    const flt4* B = reinterpret_cast<const flt4*>(inv + iptr[c]);
    const long n4 = ldm / 4;
    double acc[R];
#pragma unroll
    for (int q = 0; q < R; ++q) acc[q] = 0.0;
    for (long j = lane; j < n4; j += 64) {
        const flt4 xv = x[j];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const long i = row0 + q;
            if (i < m) {
                const flt4 b = __builtin_nontemporal_load(B + i * n4 + j);
                acc[q] += (double)b.x * (double)xv.x + (double)b.y * (double)xv.y +
                          (double)b.z * (double)xv.z + (double)b.w * (double)xv.w;
            }
        }
    }
    const int64_t* id = idx + cptr[c];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        double t = wave_sum_d(acc[q]);
        long i = row0 + q;
        if (lane == 0 && i < m) z[id[i]] = t;
    }
}

// near-field pattern: row block i is near source block j when
// |pts_i - cen_j|^2 < rad_j^2.  One wave per row block; lanes scan the
// sources 64 at a time and a ballot keeps the columns in ascending order.
template <bool FILL>
__global__ __launch_bounds__(256) void k_near_pattern(const double* __restrict__ pts,
                                                      const double* __restrict__ cen,
                                                      const double* __restrict__ rad, long nb,
                                                      int64_t* __restrict__ counts,
                                                      const int64_t* __restrict__ rowptr,
                                                      int64_t* __restrict__ col)
{
    const int lane = threadIdx.x & 63;
    const long i = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= nb) return;
    const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
    long pos = FILL ? rowptr[i] : 0;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (long s0 = 0; s0 < nb; s0 += 64) {
        const long s = s0 + lane;
        bool near = false;
        if (s < nb) {
            double dx = px - cen[3 * s], dy = py - cen[3 * s + 1], dz = pz - cen[3 * s + 2];
            double r = rad[s];
            near = dx * dx + dy * dy + dz * dz < r * r;
        }
        const unsigned long long mask = __ballot(near);
        if (FILL && near) col[pos + __popcll(mask & below)] = s;
        pos += __popcll(mask);
    }
    if (!FILL && lane == 0) counts[i] = pos;
}

rocblas_handle g_handle32 = nullptr;

}  // namespace

extern "C" {

int mbem_cast_lo(const double* A, int64_t lda, int64_t m, int64_t n, void* out, int64_t ldo,
                 int bf16, hipStream_t stream)
{
    if (m < 0 || n < 0 || (m > 0 && n > 0 && (!A || !out || lda < n || ldo < n))) return MBEM_EINVAL;
    if (m == 0 || n == 0) return MBEM_OK;
    if (m > 65535L * 1024L) return MBEM_EINVAL;
    long per = (n + 1) / 2;
    unsigned gx = (unsigned)((per + 255) / 256);
    if (gx > 64) gx = 64;
    // grid.y = rows (at most 2^31-1 on gfx950)
    dim3 grid(gx, (unsigned)m), block(256);
    if (bf16)
        hipLaunchKernelGGL((k_cast<true>), grid, block, 0, stream, A, lda, m, n, out, ldo);
    else
        hipLaunchKernelGGL((k_cast<false>), grid, block, 0, stream, A, lda, m, n, out, ldo);
    return launch_status();
}

int mbem_near_pattern(const double* pts, const double* cen, const double* rad, int64_t nb,
                      int64_t* counts, const int64_t* rowptr, int64_t* col, hipStream_t stream)
{
    if (nb < 0 || (nb > 0 && (!pts || !cen || !rad))) return MBEM_EINVAL;
    if (nb == 0) return MBEM_OK;
    if ((counts == nullptr) == (col == nullptr)) return MBEM_EINVAL;   // exactly one pass
    dim3 grid((unsigned)((nb + 3) / 4)), block(256);
    if (counts)
        hipLaunchKernelGGL((k_near_pattern<false>), grid, block, 0, stream, pts, cen, rad, nb, counts, rowptr, col);
    else {
        if (!rowptr) return MBEM_EINVAL;
        hipLaunchKernelGGL((k_near_pattern<true>), grid, block, 0, stream, pts, cen, rad, nb, counts, rowptr, col);
    }
    return launch_status();
}

int mbem_near_extract(const double* A, int64_t lda, int64_t nbr, const int64_t* rowptr,
                      const int64_t* col, int64_t nnzb, double* vals, void* lo, int64_t ldo,
                      int bf16, hipStream_t stream)
{
    if (nbr < 0 || nnzb < 0 || (nnzb > 0 && (!A || !rowptr || !col || !vals || !lo))) return MBEM_EINVAL;
    if (nnzb == 0) return MBEM_OK;
    dim3 grid((unsigned)((9 * nnzb + 255) / 256)), block(256);
    if (bf16)
        hipLaunchKernelGGL((k_near_extract<true>), grid, block, 0, stream, A, lda, nbr, rowptr, col, vals, lo, ldo);
    else
        hipLaunchKernelGGL((k_near_extract<false>), grid, block, 0, stream, A, lda, nbr, rowptr, col, vals, lo, ldo);
    return launch_status();
}

int mbem_vec_to_f32(const double* x, int64_t n, float* xf, hipStream_t stream)
{
    if (n < 0 || (n > 0 && (!x || !xf))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, x, n, xf);
    return launch_status();
}

int mbem_gemv_lo(const void* F, int64_t ldf, int64_t m, int64_t n, const float* x, double* y,
                 double alpha, double beta, int bf16, hipStream_t stream)
{
    if (m < 0 || n < 0 || (m > 0 && (!F || !y || ldf < n)) || (n > 0 && !x)) return MBEM_EINVAL;
    if (m == 0) return MBEM_OK;
    // 16-byte aligned rows and x for the vector loads
    if ((reinterpret_cast<uintptr_t>(F) & 15) || (reinterpret_cast<uintptr_t>(x) & 15)) return MBEM_EINVAL;
    if (bf16 ? (ldf % 8) : (ldf % 4)) return MBEM_EINVAL;
    constexpr int R = 4;
    long waves = (m + R - 1) / R;
    dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    if (bf16)
        hipLaunchKernelGGL((k_gemv_lo<R, true>), grid, block, 0, stream, F, ldf, m, n, x, y, alpha, beta);
    else
        hipLaunchKernelGGL((k_gemv_lo<R, false>), grid, block, 0, stream, F, ldf, m, n, x, y, alpha, beta);
    return launch_status();
}

int64_t mbem_i8_segments(int64_t n) { return n <= 0 ? 0 : (n + kSeg - 1) / kSeg; }

int mbem_near_hold(double* A, int64_t lda, int64_t nbr, const int64_t* rowptr, const int64_t* col,
                   int64_t nnzb, double* vals, int restore, hipStream_t stream)
{
    if (nbr < 0 || nnzb < 0 || (nnzb > 0 && (!A || !rowptr || !col || !vals))) return MBEM_EINVAL;
    if (nnzb == 0) return MBEM_OK;
    dim3 grid((unsigned)((9 * nnzb + 255) / 256)), block(256);
    if (restore)
        hipLaunchKernelGGL((k_near_hold<true>), grid, block, 0, stream, A, lda, nbr, rowptr, col, vals);
    else
        hipLaunchKernelGGL((k_near_hold<false>), grid, block, 0, stream, A, lda, nbr, rowptr, col, vals);
    return launch_status();
}

int mbem_cast_i8(const double* A, int64_t lda, int64_t m, int64_t n, uint8_t* out, int64_t ldo,
                 float* scale, int64_t nseg, hipStream_t stream)
{
    if (m < 0 || n < 0 || (m > 0 && n > 0 && (!A || !out || !scale || lda < n || ldo < n)))
        return MBEM_EINVAL;
    if (nseg != mbem_i8_segments(n)) return MBEM_EINVAL;
    if (m == 0 || n == 0) return MBEM_OK;
    // 16-byte stores of the packed codes
    if ((reinterpret_cast<uintptr_t>(out) & 15) || (ldo % 16)) return MBEM_EINVAL;
    const long waves = m * ((n + 1023) / 1024);
    const long blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffL) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_cast_i8, dim3((unsigned)blocks), dim3(256), 0, stream, A, lda, m, n, out, ldo, scale, nseg,
                       (unsigned char*)nullptr, 0L);
    return launch_status();
}

int64_t mbem_i8e_resid_bytes(int64_t n) { return n > 0 ? 3 * ((n + 3) / 4 * 4) : 0; }

int mbem_cast_i8e(const double* A, int64_t lda, int64_t m, int64_t n, uint8_t* out, int64_t ldo,
                  float* scale, int64_t nseg, uint8_t* resid, int64_t ldr, hipStream_t stream)
{
    if (m < 0 || n < 0 ||
        (m > 0 && n > 0 && (!A || !out || !scale || !resid || lda < n || ldo < n ||
                            ldr < mbem_i8e_resid_bytes(n))))
        return MBEM_EINVAL;
    if (nseg != mbem_i8_segments(n)) return MBEM_EINVAL;
    if (m == 0 || n == 0) return MBEM_OK;
    // 16-byte stores of the packed codes, 4-byte stores of the remainders
    if ((reinterpret_cast<uintptr_t>(out) & 15) || (ldo % 16) ||
        (reinterpret_cast<uintptr_t>(resid) & 3) || (ldr % 4))
        return MBEM_EINVAL;
    const long waves = m * ((n + 1023) / 1024);
    const long blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffL) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_cast_i8, dim3((unsigned)blocks), dim3(256), 0, stream, A, lda, m, n, out, ldo, scale, nseg,
                       resid, (long)ldr);
    return launch_status();
}

int mbem_gemv_i8(const uint8_t* F, int64_t ldf, const float* scale, int64_t nseg, int64_t m,
                 int64_t n, const float* x, double* y, double alpha, double beta, hipStream_t stream)
{
    if (m < 0 || n < 0 || (m > 0 && (!F || !y || !scale || ldf < n)) || (n > 0 && !x)) return MBEM_EINVAL;
    // the scale rows hold at least the segments the n columns touch
    if (nseg < mbem_i8_segments(n)) return MBEM_EINVAL;
    if (m == 0) return MBEM_OK;
    if ((reinterpret_cast<uintptr_t>(F) & 15) || (reinterpret_cast<uintptr_t>(x) & 15) || (ldf % 16))
        return MBEM_EINVAL;
    if (n > INT32_MAX) return MBEM_EINVAL;     // 32-bit column offsets
    constexpr int R = MBEM_I8_ROWS_PER_WAVE;
    long waves = (m + R - 1) / R;
    dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    hipLaunchKernelGGL((k_gemv_i8<R, MBEM_I8_LOADS_PER_ROW>), grid, block, 0, stream, F, ldf, scale, nseg, m, n, x, y, alpha, beta);
    return launch_status();
}

int mbem_gemv_i8e(const uint8_t* F, int64_t ldf, const uint8_t* resid, int64_t ldr,
                  const float* scale, int64_t nseg, int64_t m, int64_t n, const double* x,
                  double* y, double alpha, double beta, hipStream_t stream)
{
    if (m < 0 || n < 0 ||
        (m > 0 && (!F || !resid || !y || !scale || ldf < n || ldr < mbem_i8e_resid_bytes(n))) ||
        (n > 0 && !x))
        return MBEM_EINVAL;
    if (nseg < mbem_i8_segments(n)) return MBEM_EINVAL;
    if (m == 0) return MBEM_OK;
    // 4-byte code and remainder loads, 16-byte x loads, at multiples of 4 columns
    if ((reinterpret_cast<uintptr_t>(F) & 15) || (ldf % 16) ||
        (reinterpret_cast<uintptr_t>(resid) & 3) || (ldr % 4) ||
        (reinterpret_cast<uintptr_t>(x) & 15))
        return MBEM_EINVAL;
    if (n > INT32_MAX / 3) return MBEM_EINVAL;     // 32-bit remainder byte offsets (3 per column)
    constexpr int R = MBEM_I8E_ROWS_PER_WAVE;
    const long waves = (m + R - 1) / R;
    const long blocks = (waves + 3) / 4;
    if (blocks > 0x7fffffffL) return MBEM_EINVAL;
    hipLaunchKernelGGL((k_gemv_i8e<R>), dim3((unsigned)blocks), dim3(256), 0, stream, F, ldf, resid, ldr,
                       scale, nseg, m, n, x, y, alpha, beta);
    return launch_status();
}

int mbem_bsr_gemv(int64_t nbr, const int64_t* rowptr, const int64_t* col, const double* vals,
                  const double* x, double* y, double alpha, hipStream_t stream)
{
    if (nbr < 0 || (nbr > 0 && (!rowptr || !col || !vals || !x || !y))) return MBEM_EINVAL;
    if (nbr == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_bsr_gemv, dim3((unsigned)((nbr + 3) / 4)), dim3(256), 0, stream, nbr, rowptr, col, vals, x, y, alpha);
    return launch_status();
}

int mbem_gather_block_f32(const double* A, int64_t lda, const int64_t* idx, int64_t m, float* M,
                          hipStream_t stream)
{
    if (m < 0 || (m > 0 && (!A || !idx || !M))) return MBEM_EINVAL;
    if (m == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_gather_block_f32, dim3((unsigned)((m * m + 255) / 256)), dim3(256), 0, stream, A, lda, idx, m, M);
    return launch_status();
}

int mbem_invert_batched_f32(float* M, int64_t m, int64_t batch, int32_t* ipiv, int32_t* info,
                            hipStream_t stream)
{
    // rocblas_int sizes per matrix (m * m < 2^31 elements); the batch stride is 64-bit
    if (m < 0 || batch < 0 || (m > 0 && batch > 0 && (!M || !ipiv || !info)) || m * m > INT32_MAX ||
        batch > INT32_MAX)
        return MBEM_EINVAL;
    if (m == 0 || batch == 0) return MBEM_OK;
    if (!g_handle32 && rocblas_create_handle(&g_handle32) != rocblas_status_success) return MBEM_EHIP;
    if (rocblas_set_stream(g_handle32, stream) != rocblas_status_success) return MBEM_EHIP;
    rocblas_int mm = (rocblas_int)m;
    rocblas_stride sA = (rocblas_stride)(m * m), sP = (rocblas_stride)m;
    if (rocsolver_sgetrf_strided_batched(g_handle32, mm, mm, M, mm, sA, ipiv, sP, info, (rocblas_int)batch) != rocblas_status_success)
        return MBEM_EHIP;
    if (rocsolver_sgetri_strided_batched(g_handle32, mm, M, mm, sA, ipiv, sP, info, (rocblas_int)batch) != rocblas_status_success)
        return MBEM_EHIP;
    return launch_status();
}

int mbem_gather_f32(const int64_t* gidx, int64_t np, const double* r, float* rr, hipStream_t stream)
{
    if (np < 0 || (np > 0 && (!gidx || !r || !rr))) return MBEM_EINVAL;
    if (np == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_gather_f32, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, stream, gidx, np, r, rr);
    return launch_status();
}

int mbem_bj_apply_f32v(const int64_t* wave_table, int64_t nwaves, const int64_t* chunk_ptr,
                       const int64_t* pad_ptr, const int64_t* idx, const int64_t* inv_ptr,
                       const float* inv, const float* rr, double* z, hipStream_t stream)
{
    if (nwaves < 0 || (nwaves > 0 && (!wave_table || !chunk_ptr || !pad_ptr || !idx || !inv_ptr ||
                                      !inv || !rr || !z)))
        return MBEM_EINVAL;
    if (nwaves == 0) return MBEM_OK;
    if ((reinterpret_cast<uintptr_t>(inv) & 15) || (reinterpret_cast<uintptr_t>(rr) & 15)) return MBEM_EINVAL;
    constexpr int R = MBEM_BJ32_ROWS_PER_WAVE;
    hipLaunchKernelGGL((k_bj_apply_f32v<R>), dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, stream, wave_table, nwaves, chunk_ptr, pad_ptr, idx, inv_ptr, inv, rr, z);
    return launch_status();
}

int mbem_bj_apply_f32(const int64_t* wave_table, int64_t nwaves, const int64_t* chunk_ptr,
                      const int64_t* idx, const int64_t* inv_ptr, const float* inv, const double* r,
                      double* z, hipStream_t stream)
{
    if (nwaves < 0 || (nwaves > 0 && (!wave_table || !chunk_ptr || !idx || !inv_ptr || !inv || !r || !z))) return MBEM_EINVAL;
    if (nwaves == 0) return MBEM_OK;
    hipLaunchKernelGGL((k_bj_apply_f32<MBEM_BJ_ROWS_PER_WAVE>), dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, stream, wave_table, nwaves, chunk_ptr, idx, inv_ptr, inv, r, z);
    return launch_status();
}

}  // extern "C"
