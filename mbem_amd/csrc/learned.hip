// Kernels of the learned FGMRES preconditioner: a point-cloud network over the
// boundary elements (one point per triangle), trained and evaluated on gfx950.
//
// Network (P points, width W = 128, inputs F = 16 per point):
//   h1 = relu(X W1^T + b1)                      P x 128
//   a  = relu(h1 W2^T + b2);  m_i = mean_{j in kNN(i)} a_j   (k neighbours)
//   h3 = relu([h1 | m] W3^T + b3)               P x 128
//   g  = mean_i h3_i                            128
//   h4 = relu(h3 W4a^T + (W4b g + b4))          P x 128
//   y  = h4 W5^T + b5                           P x 6K
// The output holds two 3P x K bases Phi, Psi of a low-rank correction
// Phi (Psi^T r) added to the block-Jacobi preconditioner (k_lr_* below).
// Training: FP32 throughout (VALU SGEMM below, split-K for weight gradients,
// deterministic reductions, Adam).  Inference: the same graph with the dense
// layers on the BF16 matrix cores (k_hgemm), FP32 accumulation.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mbem_hip.h"

namespace {

inline int launch_status() { return hipGetLastError() == hipSuccess ? MBEM_OK : MBEM_EHIP; }

typedef float flt4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));

// ---------------------------------------------------------------------------
// k nearest neighbours (brute force, one thread per point, sorted insertion,
// ties broken by index so the graph is deterministic; the point itself is
// excluded).  K <= 32.
template <int K>
__global__ __launch_bounds__(256) void k_knn(const float* __restrict__ pts, long P,
                                             int32_t* __restrict__ nbr)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ float sp[3 * 256];
    float px = 0.f, py = 0.f, pz = 0.f;
    if (i < P) {
        px = pts[3 * i];
        py = pts[3 * i + 1];
        pz = pts[3 * i + 2];
    }
    float bd[K];
    int bi[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        bd[q] = 3.0e38f;
        bi[q] = -1;
    }
    for (long j0 = 0; j0 < P; j0 += 256) {
        const long jl = j0 + threadIdx.x;
        __syncthreads();
        if (jl < P) {
            sp[3 * threadIdx.x] = pts[3 * jl];
            sp[3 * threadIdx.x + 1] = pts[3 * jl + 1];
            sp[3 * threadIdx.x + 2] = pts[3 * jl + 2];
        }
        __syncthreads();
        const long jn = (P - j0 < 256) ? P - j0 : 256;
        for (long t = 0; t < jn; ++t) {
            const long j = j0 + t;
            if (j == i) continue;
            const float dx = sp[3 * t] - px, dy = sp[3 * t + 1] - py, dz = sp[3 * t + 2] - pz;
            const float d = dx * dx + dy * dy + dz * dz;
            if (d < bd[K - 1] || (d == bd[K - 1] && (int)j < bi[K - 1])) {
                // insert keeping (distance, index) ascending
                float cd = d;
                int ci = (int)j;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const bool before = (cd < bd[q]) || (cd == bd[q] && ci < bi[q]);
                    if (before) {
                        const float td = bd[q];
                        const int ti = bi[q];
                        bd[q] = cd;
                        bi[q] = ci;
                        cd = td;
                        ci = ti;
                    }
                }
            }
        }
    }
    if (i < P) {
#pragma unroll
        for (int q = 0; q < K; ++q) nbr[i * K + q] = bi[q];
    }
}

// ---------------------------------------------------------------------------
// FP32 GEMM with free strides: C[i, j] = alpha sum_k A(i, k) B(k, j) + beta C[i, j]
// (+ bias[j]) (relu), A(i, k) = A[i sa0 + k sa1], B(k, j) = B[k sb0 + j sb1],
// C row-major with ldc.  64 x 64 tile, 256 threads, 4 x 4 outputs per thread,
// K slices of 16 through LDS.  Split-K over grid.z writes partial tiles to
// C + z * cstride (beta, bias and relu only when gridDim.z == 1).
__global__ __launch_bounds__(256) void k_sgemm(const float* __restrict__ A, long sa0, long sa1,
                                               const float* __restrict__ B, long sb0, long sb1,
                                               float* __restrict__ C, long ldc, long cstride,
                                               long M, long N, long K, float alpha, float beta,
                                               const float* __restrict__ bias, int relu, long kchunk)
{
    __shared__ float sA[16][64 + 4];
    __shared__ float sB[16][64 + 4];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const long i0 = (long)blockIdx.x * 64, j0 = (long)blockIdx.y * 64;
    const long kb = (long)blockIdx.z * kchunk;
    const long ke = (kb + kchunk < K) ? kb + kchunk : K;
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
    for (long k0 = kb; k0 < ke; k0 += 16) {
        // A tile: 64 rows x 16 k; B tile: 16 k x 64 columns (4 elements per thread each)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = threadIdx.x + q * 256;
            const int r = e >> 4, kk = e & 15;
            const long gi = i0 + r, gk = k0 + kk;
            sA[kk][r] = (gi < M && gk < ke) ? A[gi * sa0 + gk * sa1] : 0.f;
            const int kk2 = e >> 6, c = e & 63;
            const long gk2 = k0 + kk2, gj = j0 + c;
            sB[kk2][c] = (gk2 < ke && gj < N) ? B[gk2 * sb0 + gj * sb1] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = sA[kk][ty + 16 * a];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = sB[kk][tx + 16 * b];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] += av[a] * bv[b];
        }
        __syncthreads();
    }
    const bool split = gridDim.z > 1;
    float* Cz = C + (split ? (long)blockIdx.z * cstride : 0);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const long gi = i0 + ty + 16 * a;
        if (gi >= M) continue;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const long gj = j0 + tx + 16 * b;
            if (gj >= N) continue;
            float v = alpha * acc[a][b];
            if (!split) {
                if (beta != 0.f) v += beta * Cz[gi * ldc + gj];
                if (bias) v += bias[gj];
                if (relu) v = v > 0.f ? v : 0.f;
            }
            Cz[gi * ldc + gj] = v;
        }
    }
}

// out[t] = beta out[t] + sum_z part[z * stride + t] (split-K reduction)
__global__ __launch_bounds__(256) void k_sum_parts(const float* __restrict__ part, long nz,
                                                   long stride, long n, float* __restrict__ out,
                                                   float beta)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float s = 0.f;
    for (long z = 0; z < nz; ++z) s += part[z * stride + t];
    out[t] = (beta != 0.f ? beta * out[t] : 0.f) + s;
}

// column sums of a P x C row-major matrix: partial sums over row slices,
// then a second pass (deterministic).  part: nslice x C.
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ X, long P, long C,
                                                     long rows_per, float* __restrict__ part)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long s = blockIdx.y;
    if (c >= C) return;
    const long r0 = s * rows_per;
    const long r1 = (r0 + rows_per < P) ? r0 + rows_per : P;
    float acc = 0.f;
    for (long r = r0; r < r1; ++r) acc += X[r * C + c];
    part[s * C + c] = acc;
}

// m[i, :] = (1/k) sum_q a[nbr[i, q], :]
__global__ __launch_bounds__(256) void k_knn_mean(const float* __restrict__ a, long P, long C,
                                                  const int32_t* __restrict__ nbr, int k,
                                                  float* __restrict__ m)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * C) return;
    const long i = t / C, c = t % C;
    float s = 0.f;
    for (int q = 0; q < k; ++q) s += a[(long)nbr[i * k + q] * C + c];
    m[t] = s / (float)k;
}

// backward of the kNN mean through the reverse graph (CSR of in-neighbours):
// da[j, :] = (1/k) sum_{i: j in kNN(i)} dm[i, :]
__global__ __launch_bounds__(256) void k_knn_mean_bwd(const float* __restrict__ dm, long P, long C,
                                                      const int32_t* __restrict__ rptr,
                                                      const int32_t* __restrict__ ridx, int k,
                                                      float* __restrict__ da)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P * C) return;
    const long j = t / C, c = t % C;
    float s = 0.f;
    for (int q = rptr[j]; q < rptr[j + 1]; ++q) s += dm[(long)ridx[q] * C + c];
    da[t] = s / (float)k;
}

// dY *= (Y > 0)   (relu backward, Y the activation)
__global__ __launch_bounds__(256) void k_relu_bwd(const float* __restrict__ Y, float* __restrict__ dY,
                                                  long n)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n && !(Y[t] > 0.f)) dY[t] = 0.f;
}

// out[:, 0:C1] = A, out[:, C1:C1+C2] = B  (row concatenation of two P-row blocks)
__global__ __launch_bounds__(256) void k_concat(const float* __restrict__ A, long C1,
                                                const float* __restrict__ B, long C2, long P,
                                                float* __restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long C = C1 + C2;
    if (t >= P * C) return;
    const long i = t / C, c = t % C;
    out[t] = (c < C1) ? A[i * C1 + c] : B[i * C2 + (c - C1)];
}

// the inverse of k_concat for gradients: dA += d[:, :C1] (or =), dB = d[:, C1:]
__global__ __launch_bounds__(256) void k_split(const float* __restrict__ d, long C1, long C2, long P,
                                               float* __restrict__ dA, int accumulate_a,
                                               float* __restrict__ dB)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long C = C1 + C2;
    if (t >= P * C) return;
    const long i = t / C, c = t % C;
    if (c < C1) {
        float* p = dA + i * C1 + c;
        *p = accumulate_a ? *p + d[t] : d[t];
    } else if (dB) {
        dB[i * C2 + (c - C1)] = d[t];
    }
}

// X[i, c] += v[c]
__global__ __launch_bounds__(256) void k_add_row(float* __restrict__ X, long P, long C,
                                                 const float* __restrict__ v, float scale)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < P * C) X[t] += scale * v[t % C];
}

// relative loss and its gradient: y (P x 3) network output, t target, s scale:
// diff = s y - t; sums[0] += |diff|^2 (per block partials), sums[1] += |t|^2.
__global__ __launch_bounds__(256) void k_loss_part(const float* __restrict__ y,
                                                   const float* __restrict__ tgt, long n, float s,
                                                   float* __restrict__ part)
{
    __shared__ float r0[256], r1[256];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float d2 = 0.f, t2 = 0.f;
    if (t < n) {
        const float d = s * y[t] - tgt[t];
        d2 = d * d;
        t2 = tgt[t] * tgt[t];
    }
    r0[threadIdx.x] = d2;
    r1[threadIdx.x] = t2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = r0[0];
        part[2 * blockIdx.x + 1] = r1[0];
    }
}

// dy = 2 s (s y - t) / tsum  (tsum = |t|^2 from the partials, read on device)
__global__ __launch_bounds__(256) void k_loss_grad(const float* __restrict__ y,
                                                   const float* __restrict__ tgt, long n, float s,
                                                   const float* __restrict__ part, long nparts,
                                                   float* __restrict__ dy, float* __restrict__ loss)
{
    __shared__ float tot[2];
    if (threadIdx.x == 0) {
        float a = 0.f, b = 0.f;
        for (long q = 0; q < nparts; ++q) {
            a += part[2 * q];
            b += part[2 * q + 1];
        }
        tot[0] = a;
        tot[1] = b > 0.f ? b : 1.f;
        if (blockIdx.x == 0 && loss) loss[0] = a / tot[1];
    }
    __syncthreads();
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dy[t] = 2.f * s * (s * y[t] - tgt[t]) / tot[1];
}

// Adam (Kingma & Ba 2015), bias-corrected step sizes passed in
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, long n,
                                              float lr_t, float b1, float b2, float eps, float wd)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t] + wd * w[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    w[t] -= lr_t * mt / (sqrtf(vt) + eps);
}

// ---------------------------------------------------------------------------
// BF16 matrix-core GEMM for inference: C (P x N) = act(A (P x K) W^T + bias),
// A FP32 activations and W FP32 weights rounded to BF16 on the way into LDS,
// v_mfma_f32_16x16x16bf16 with FP32 accumulation.  K a multiple of 16 (the
// layer widths); N a multiple of 16 or the 3-column head (padded).
__device__ __forceinline__ unsigned short to_bf16(float f)
{
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));

// block: 4 waves, tile 64 points x 64 outputs; wave w owns rows 16w..16w+15
// and all 4 column tiles.  Operand layout of v_mfma_f32_16x16x16bf16_1k
// (wave64): lane l holds A[i = l % 16][k = 4 (l / 16) .. +3] and
// B[k = 4 (l / 16) .. +3][j = l % 16]; result element v is
// D[i = 4 (l / 16) + v][j = l % 16].
__global__ __launch_bounds__(256) void k_hgemm(const float* __restrict__ A, long lda,
                                               const float* __restrict__ W, long K, long N,
                                               long P, const float* __restrict__ bias, int relu,
                                               float* __restrict__ C, long ldc)
{
    __shared__ unsigned short sA[64][16 + 2];
    __shared__ unsigned short sW[64][16 + 2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long i0 = (long)blockIdx.x * 64, j0 = (long)blockIdx.y * 64;
    flt4 acc[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[b] = flt4{0.f, 0.f, 0.f, 0.f};
    for (long k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = threadIdx.x + q * 256;
            const int r = e >> 4, kk = e & 15;
            const long gi = i0 + r, gj = j0 + r, gk = k0 + kk;
            sA[r][kk] = to_bf16((gi < P && gk < K) ? A[gi * lda + gk] : 0.f);
            sW[r][kk] = to_bf16((gj < N && gk < K) ? W[gj * K + gk] : 0.f);
        }
        __syncthreads();
        const int fi = lane & 15, fk = (lane >> 4) * 4;
        bf16x4 a;
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = (short)sA[wave * 16 + fi][fk + u];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            bf16x4 w;
#pragma unroll
            for (int u = 0; u < 4; ++u) w[u] = (short)sW[b * 16 + fi][fk + u];
            acc[b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, acc[b], 0, 0, 0);
        }
        __syncthreads();
    }
    const int fj = lane & 15, fr = (lane >> 4) * 4;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const long gj = j0 + b * 16 + fj;
        if (gj >= N) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long gi = i0 + wave * 16 + fr + v;
            if (gi >= P) continue;
            float val = acc[b][v] + (bias ? bias[gj] : 0.f);
            if (relu) val = val > 0.f ? val : 0.f;
            C[gi * ldc + gj] = val;
        }
    }
}

// ---------------------------------------------------------------------------
// Low-rank correction of the learned preconditioner.  The network's output
// Y (P x 6K, row-major) holds, for point p and component a, the rows
//   Phi[3p + a, :] = Y[p, a K .. a K + K - 1]
//   Psi[3p + a, :] = Y[p, 3K + a K .. 3K + a K + K - 1]
// of two 3P x K bases; the correction is Phi (Psi^T r).

// C[b, k] = sum_i Bas[i, k] V[b, i] over i < 3P (Bas = Phi or Psi through its
// column offset off = 0 or 3K); partial sums over point slices (FP64 accumulation).
template <typename T>
__global__ __launch_bounds__(256) void k_lr_coef_part(const float* __restrict__ Y, long P, long K,
                                                      long off, const T* __restrict__ V, long ldv,
                                                      long B, long pts_per, double* __restrict__ part)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;   // (b, k) pair
    if (q >= B * K) return;
    const long b = q / K, k = q % K;
    const long p0 = (long)blockIdx.y * pts_per;
    const long p1 = (p0 + pts_per < P) ? p0 + pts_per : P;
    const long W = 6 * K;
    double acc = 0.0;
    for (long p = p0; p < p1; ++p) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            acc += (double)Y[p * W + off + a * K + k] * (double)V[b * ldv + 3 * p + a];
    }
    part[(long)blockIdx.y * B * K + q] = acc;
}

__global__ __launch_bounds__(256) void k_lr_coef_sum(const double* __restrict__ part, long ns,
                                                     long BK, double* __restrict__ C)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= BK) return;
    double s = 0.0;
    for (long z = 0; z < ns; ++z) s += part[z * BK + q];
    C[q] = s;
}

// Out[b, 3p + a] (+)= alpha d[3p + a] sum_k Y[p, off + a K + k] C[b, k]
// (d = 1 when no row scale is given)
template <typename T>
__global__ __launch_bounds__(256) void k_lr_expand(const float* __restrict__ Y, long P, long K,
                                                   long off, const double* __restrict__ C, long B,
                                                   T* __restrict__ Out, long ldo, double alpha,
                                                   int accumulate,
                                                   const double* __restrict__ rowscale)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * 3 * P) return;
    const long b = t / (3 * P), i = t % (3 * P);
    const long p = i / 3, a = i % 3;
    const float* y = Y + p * 6 * K + off + a * K;
    const double* c = C + b * K;
    double s = 0.0;
    for (long k = 0; k < K; ++k) s += (double)y[k] * c[k];
    if (rowscale) s *= rowscale[i];
    T* o = Out + b * ldo + i;
    *o = accumulate ? (T)((double)*o + alpha * s) : (T)(alpha * s);
}

// dY[p, off + a K + k] = sum_b U[b, 3p + a] C[b, k]  (the outer-product gradient)
__global__ __launch_bounds__(256) void k_lr_outer(const float* __restrict__ U, long ldu,
                                                  const double* __restrict__ C, long B, long P,
                                                  long K, long off, float* __restrict__ dY)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * P * K) return;
    const long i = t / K, k = t % K;
    const long p = i / 3, a = i % 3;
    double s = 0.0;
    for (long b = 0; b < B; ++b) s += (double)U[b * ldu + i] * C[b * K + k];
    dY[p * 6 * K + off + a * K + k] = (float)s;
}

// batch loss: L = (1/B) sum_b |pred_b - t_b|^2 / |t_b|^2 and dpred (one block per b)
__global__ __launch_bounds__(256) void k_lr_loss(const float* __restrict__ pred,
                                                 const float* __restrict__ tgt, long n, long B,
                                                 float* __restrict__ dpred, float* __restrict__ lb)
{
    __shared__ double r0[256], r1[256];
    const long b = blockIdx.x;
    const float* pr = pred + b * n;
    const float* tg = tgt + b * n;
    double d2 = 0.0, t2 = 0.0;
    for (long i = threadIdx.x; i < n; i += 256) {
        const double d = (double)pr[i] - (double)tg[i];
        d2 += d * d;
        t2 += (double)tg[i] * tg[i];
    }
    r0[threadIdx.x] = d2;
    r1[threadIdx.x] = t2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    const double tt = r1[0] > 0.0 ? r1[0] : 1.0;
    for (long i = threadIdx.x; i < n; i += 256)
        dpred[b * n + i] = (float)(2.0 * ((double)pr[i] - (double)tg[i]) / (tt * (double)B));
    if (threadIdx.x == 0) lb[b] = (float)(r0[0] / tt / (double)B);
}

}  // namespace

extern "C" {

int mbem_knn(const float* pts, int64_t P, int k, int32_t* nbr, hipStream_t stream)
{
    if (P < 2 || !pts || !nbr || (k != 8 && k != 16) || k >= P) return MBEM_EINVAL;
    dim3 grid((unsigned)((P + 255) / 256));
    if (k == 8) hipLaunchKernelGGL(k_knn<8>, grid, dim3(256), 0, stream, pts, P, nbr);
    else hipLaunchKernelGGL(k_knn<16>, grid, dim3(256), 0, stream, pts, P, nbr);
    return launch_status();
}

int mbem_sgemm(const float* A, int64_t sa0, int64_t sa1, const float* B, int64_t sb0, int64_t sb1,
               float* C, int64_t ldc, int64_t M, int64_t N, int64_t K, float alpha, float beta,
               const float* bias, int relu, int64_t splits, float* work, hipStream_t stream)
{
    if (M < 0 || N < 0 || K < 0 || splits < 1 || !A || !B || !C || ldc < N) return MBEM_EINVAL;
    if (M == 0 || N == 0) return MBEM_OK;
    const long gx = (M + 63) / 64, gy = (N + 63) / 64;
    if (gy > 65535 || splits > 65535) return MBEM_EINVAL;
    if (splits == 1) {
        hipLaunchKernelGGL(k_sgemm, dim3((unsigned)gx, (unsigned)gy, 1), dim3(256), 0, stream, A, sa0, sa1,
                           B, sb0, sb1, C, ldc, 0L, M, N, K, alpha, beta, bias, relu, K > 0 ? K : 1);
        return launch_status();
    }
    // split-K: partial tiles in work (splits x M x N, ld = N), then summed into C
    if (!work || bias || relu || ldc != N) return MBEM_EINVAL;
    long kchunk = (K + splits - 1) / splits;
    kchunk = ((kchunk + 15) / 16) * 16;
    hipLaunchKernelGGL(k_sgemm, dim3((unsigned)gx, (unsigned)gy, (unsigned)splits), dim3(256), 0, stream,
                       A, sa0, sa1, B, sb0, sb1, work, N, M * N, M, N, K, alpha, 0.f,
                       (const float*)nullptr, 0, kchunk);
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((M * N + 255) / 256)), dim3(256), 0, stream, work,
                       splits, M * N, M * N, C, beta);
    return launch_status();
}

int mbem_colsum(const float* X, int64_t P, int64_t C, float* out, float* work, int64_t nslices,
                float beta, hipStream_t stream)
{
    if (P < 1 || C < 1 || nslices < 1 || nslices > 65535 || !X || !out || !work) return MBEM_EINVAL;
    const long rows_per = (P + nslices - 1) / nslices;
    hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)((C + 255) / 256), (unsigned)nslices), dim3(256), 0,
                       stream, X, P, C, rows_per, work);
    hipLaunchKernelGGL(k_sum_parts, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, stream, work,
                       nslices, C, C, out, beta);
    return launch_status();
}

int mbem_knn_mean(const float* a, int64_t P, int64_t C, const int32_t* nbr, int k, float* m,
                  hipStream_t stream)
{
    if (P < 1 || C < 1 || k < 1 || !a || !nbr || !m) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_knn_mean, dim3((unsigned)((P * C + 255) / 256)), dim3(256), 0, stream, a, P, C,
                       nbr, k, m);
    return launch_status();
}

int mbem_knn_mean_bwd(const float* dm, int64_t P, int64_t C, const int32_t* rptr,
                      const int32_t* ridx, int k, float* da, hipStream_t stream)
{
    if (P < 1 || C < 1 || k < 1 || !dm || !rptr || !ridx || !da) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_knn_mean_bwd, dim3((unsigned)((P * C + 255) / 256)), dim3(256), 0, stream, dm,
                       P, C, rptr, ridx, k, da);
    return launch_status();
}

int mbem_relu_bwd(const float* Y, float* dY, int64_t n, hipStream_t stream)
{
    if (n < 0 || (n > 0 && (!Y || !dY))) return MBEM_EINVAL;
    if (n == 0) return MBEM_OK;
    hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, Y, dY, n);
    return launch_status();
}

int mbem_concat(const float* A, int64_t C1, const float* B, int64_t C2, int64_t P, float* out,
                hipStream_t stream)
{
    if (P < 1 || C1 < 1 || C2 < 1 || !A || !B || !out) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_concat, dim3((unsigned)((P * (C1 + C2) + 255) / 256)), dim3(256), 0, stream, A,
                       C1, B, C2, P, out);
    return launch_status();
}

int mbem_split(const float* d, int64_t C1, int64_t C2, int64_t P, float* dA, int accumulate_a,
               float* dB, hipStream_t stream)
{
    if (P < 1 || C1 < 1 || C2 < 1 || !d || !dA) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_split, dim3((unsigned)((P * (C1 + C2) + 255) / 256)), dim3(256), 0, stream, d,
                       C1, C2, P, dA, accumulate_a, dB);
    return launch_status();
}

int mbem_add_row(float* X, int64_t P, int64_t C, const float* v, float scale, hipStream_t stream)
{
    if (P < 1 || C < 1 || !X || !v) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_add_row, dim3((unsigned)((P * C + 255) / 256)), dim3(256), 0, stream, X, P, C,
                       v, scale);
    return launch_status();
}

int mbem_loss(const float* y, const float* tgt, int64_t n, float s, float* part, float* dy,
              float* loss, hipStream_t stream)
{
    if (n < 1 || !y || !tgt || !part || !dy) return MBEM_EINVAL;
    const long nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_loss_part, dim3((unsigned)nb), dim3(256), 0, stream, y, tgt, n, s, part);
    hipLaunchKernelGGL(k_loss_grad, dim3((unsigned)nb), dim3(256), 0, stream, y, tgt, n, s, part, nb, dy,
                       loss);
    return launch_status();
}

int mbem_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr_t, float b1,
              float b2, float eps, float wd, hipStream_t stream)
{
    if (n < 1 || !w || !g || !m || !v) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, w, g, m, v, n,
                       lr_t, b1, b2, eps, wd);
    return launch_status();
}

int mbem_hgemm(const float* A, int64_t lda, const float* W, int64_t K, int64_t N, int64_t P,
               const float* bias, int relu, float* C, int64_t ldc, hipStream_t stream)
{
    if (P < 1 || N < 1 || K < 1 || (K % 16) || !A || !W || !C || ldc < N) return MBEM_EINVAL;
    const long gx = (P + 63) / 64, gy = (N + 63) / 64;
    hipLaunchKernelGGL(k_hgemm, dim3((unsigned)gx, (unsigned)gy), dim3(256), 0, stream, A, lda, W, K, N, P,
                       bias, relu, C, ldc);
    return launch_status();
}

int mbem_lr_coef(const float* Y, int64_t P, int64_t K, int64_t off, const void* V, int64_t ldv,
                 int64_t B, int v_fp64, double* part, int64_t nslices, double* C, hipStream_t stream)
{
    if (P < 1 || K < 1 || B < 1 || nslices < 1 || nslices > 65535 || !Y || !V || !part || !C)
        return MBEM_EINVAL;
    if (off != 0 && off != 3 * K) return MBEM_EINVAL;
    const long per = (P + nslices - 1) / nslices;
    dim3 grid((unsigned)((B * K + 255) / 256), (unsigned)nslices);
    if (v_fp64)
        hipLaunchKernelGGL(k_lr_coef_part<double>, grid, dim3(256), 0, stream, Y, P, K, off,
                           (const double*)V, ldv, B, per, part);
    else
        hipLaunchKernelGGL(k_lr_coef_part<float>, grid, dim3(256), 0, stream, Y, P, K, off,
                           (const float*)V, ldv, B, per, part);
    hipLaunchKernelGGL(k_lr_coef_sum, dim3((unsigned)((B * K + 255) / 256)), dim3(256), 0, stream, part,
                       nslices, B * K, C);
    return launch_status();
}

int mbem_lr_expand(const float* Y, int64_t P, int64_t K, int64_t off, const double* C, int64_t B,
                   void* Out, int64_t ldo, int out_fp64, double alpha, int accumulate,
                   const double* rowscale, hipStream_t stream)
{
    if (P < 1 || K < 1 || B < 1 || !Y || !C || !Out) return MBEM_EINVAL;
    if (off != 0 && off != 3 * K) return MBEM_EINVAL;
    const long n = B * 3 * P;
    if (out_fp64)
        hipLaunchKernelGGL(k_lr_expand<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, Y,
                           P, K, off, C, B, (double*)Out, ldo, alpha, accumulate, rowscale);
    else
        hipLaunchKernelGGL(k_lr_expand<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, Y,
                           P, K, off, C, B, (float*)Out, ldo, alpha, accumulate, rowscale);
    return launch_status();
}

int mbem_lr_outer(const float* U, int64_t ldu, const double* C, int64_t B, int64_t P, int64_t K,
                  int64_t off, float* dY, hipStream_t stream)
{
    if (P < 1 || K < 1 || B < 1 || !U || !C || !dY) return MBEM_EINVAL;
    if (off != 0 && off != 3 * K) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_lr_outer, dim3((unsigned)((3 * P * K + 255) / 256)), dim3(256), 0, stream, U, ldu,
                       C, B, P, K, off, dY);
    return launch_status();
}

int mbem_lr_loss(const float* pred, const float* tgt, int64_t n, int64_t B, float* dpred,
                 float* loss_b, hipStream_t stream)
{
    if (n < 1 || B < 1 || B > 65535 || !pred || !tgt || !dpred || !loss_b) return MBEM_EINVAL;
    hipLaunchKernelGGL(k_lr_loss, dim3((unsigned)B), dim3(256), 0, stream, pred, tgt, n, B, dpred, loss_b);
    return launch_status();
}

}  // extern "C"
