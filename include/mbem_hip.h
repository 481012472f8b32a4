/*
 * mbem_hip.h -- C ABI of libmbem_hip.so, the MI355X (gfx950) implementation of
 * mbem's dense mollified-BEM hot path.
 *
 * Conventions (every entry point):
 *   - pointers named like arrays are DEVICE pointers unless stated; the small
 *     coefficient vectors (g[3], c[6]) are HOST pointers, read at the call;
 *   - sizes and leading dimensions are element counts (int64_t);
 *   - work is enqueued on `stream` and returns at once (stream-ordered);
 *   - nothing is allocated except by mbem_invert / mbem_lu_* (rocSOLVER's
 *     own workspace) -- every buffer is caller-owned;
 *   - return MBEM_OK, MBEM_EINVAL (bad size, null pointer, non-positive mu),
 *     or MBEM_EHIP (launch / runtime error).  No exception crosses the ABI.
 *
 * Source record: mbem_prepare_sources packs each source triangle into a
 * structure-of-arrays record of mbem_src_fields() doubles (vertices, the
 * in-plane frame ex / ey / nhat of tri_kernels.py:136-153, the caller's
 * normal, eps, a non-degenerate flag), column s of a (fields, ld_src) array.
 * Every kernel below reads sources through that record.
 *
 * Layout: matrices are row-major; a kernel block is written at out[(3f+i) *
 * ld_out + 3s+j] (tri_kernels.py:40-41), so `out` may point inside a larger
 * system matrix.  Output updates are out = scale * K + beta * out.
 */
#ifndef MBEM_HIP_H
#define MBEM_HIP_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBEM_OK 0
#define MBEM_EINVAL 1
#define MBEM_EHIP 2
/* the call needs more device scratch than the library holds
 * (mbem_multidot_noalloc) */
#define MBEM_ESCRATCH 3

/* rows of a block-Jacobi chunk one wavefront applies (mbem_bj_apply tables) */
#define MBEM_BJ_ROWS_PER_WAVE 4

/* rows of a padded FP32 block-Jacobi chunk one wavefront applies
 * (mbem_bj_apply_f32v tables) */
#define MBEM_BJ32_ROWS_PER_WAVE 8

/* rows of the int8 far field one wavefront multiplies (mbem_gemv_i8), and
 * the 16-byte loads per row each lane keeps in flight (measured at n = 60,000
 * before x was staged through LDS: 8 x 2 623 us, 8 x 3 625 us, 4 x 4 644 us,
 * 4 x 3 652 us; with LDS staging 8 x 2 593 us) */
#define MBEM_I8_ROWS_PER_WAVE 8
#define MBEM_I8_LOADS_PER_ROW 2

/* rows of the split int8 far field one wavefront multiplies (mbem_gemv_i8e) */
#define MBEM_I8E_ROWS_PER_WAVE 4

/* doubles per source record */
int mbem_src_fields(void);

/* Host function (no device work): from a triangle mesh's vertices (nv,3)
 * and triangles (nt,3, int64 vertex indices) write tri_verts (nt,3,3), the
 * unit normals (nt,3) and the centroids (nt,3), each row-major: the source
 * arrays and centroid collocation points of a P0 patch, bit for bit what
 * numpy gives for vertices[triangles], cross / norm and mean (plain IEEE
 * FP64 operations, no contraction).  Replaces the host geometry of
 * basis.py:91-150 (_source_arrays and the centroids).  MBEM_EINVAL on a
 * triangle index outside [0, nv). */
int mbem_tri_geometry(const double *verts, int64_t nv, const int64_t *tris, int64_t nt,
                      double *tri_verts, double *normals, double *centroids);

/* Pack sources: tri_verts (ns,3,3), normals (ns,3), eps (ns) -> src
 * (mbem_src_fields(), ld_src).  Replaces tri_kernels._precompute_frames
 * (tri_kernels.py:1052-1065) and the per-source _tri_frame (:136-153). */
int mbem_prepare_sources(const double *tri_verts, const double *normals, const double *eps,
                         int64_t ns, double *src, int64_t ld_src, hipStream_t stream);

/* U (force -> displacement) block with coefficients g = u_coeffs(mu, lam)
 * applied in-loop.  Replaces tri_kernels.u_matrix_direct
 * (tri_kernels.py:546-565). */
int mbem_u_block(const double *x_field, int64_t nf, const double *src, int64_t ld_src, int64_t ns,
                 const double *g, double *out, int64_t ld_out, double scale, double beta,
                 hipStream_t stream);

/* T (slip -> displacement) block with c = t_coeffs(mu, lam).  Replaces
 * tri_kernels.t_matrix_direct (tri_kernels.py:568-591).  With rowsum != NULL
 * also rowsum[9f + 3i + j] += rs_scale * sum_s K[3f+i, 3s+j], the calibration
 * row sum of backends/dense.py:497-499; store = 0 computes the row sum only
 * (out unused). */
int mbem_t_block(const double *x_field, int64_t nf, const double *src, int64_t ld_src, int64_t ns,
                 const double *c, double *out, int64_t ld_out, double scale, double beta,
                 double *rowsum, double rs_scale, int store, hipStream_t stream);

/* Geometry-only basis stacks: U (3, 3nf, ld_out) and T (6, 3nf, ld_out), stack
 * k at out + k * stack_stride.  Replace tri_kernels.u_basis_matrices /
 * t_basis_matrices (tri_kernels.py:451-492) and their _serial twins
 * (:502-543). */
int mbem_u_basis(const double *x_field, int64_t nf, const double *src, int64_t ld_src, int64_t ns,
                 double *out, int64_t ld_out, int64_t stack_stride, hipStream_t stream);
int mbem_t_basis(const double *x_field, int64_t nf, const double *src, int64_t ld_src, int64_t ns,
                 double *out, int64_t ld_out, int64_t stack_stride, hipStream_t stream);

/* Matrix-free displacement, out (nf,3).  Replace tri_kernels.u_disp_contract /
 * t_disp_contract (tri_kernels.py:976-1049). */
int mbem_u_disp_contract(const double *x_field, int64_t nf, const double *src, int64_t ld_src,
                         int64_t ns, const double *density, const double *g, double *out,
                         double scale, double beta, hipStream_t stream);
int mbem_t_disp_contract(const double *x_field, int64_t nf, const double *src, int64_t ld_src,
                         int64_t ns, const double *density, const double *c, double *out,
                         double scale, double beta, hipStream_t stream);

/* Matrix-free stress, out (nf,3,3).  Replace tri_kernels.kelvin_stress_contract
 * (tri_kernels.py:1068-1108), dd_stress_contract (:1111-1154, total stress)
 * and eigenstress_contract (:1217-1260, +C:eps*, full sum). */
int mbem_kelvin_stress_contract(const double *x_field, int64_t nf, const double *src,
                                int64_t ld_src, int64_t ns, const double *density, double mu,
                                double lam, double *out, double scale, double beta,
                                hipStream_t stream);
int mbem_dd_stress_contract(const double *x_field, int64_t nf, const double *src, int64_t ld_src,
                            int64_t ns, const double *density, double mu, double lam, double *out,
                            double scale, double beta, hipStream_t stream);
int mbem_eigenstress_contract(const double *x_field, int64_t nf, const double *src,
                              int64_t ld_src, int64_t ns, const double *density, double mu,
                              double lam, double *out, double scale, double beta,
                              hipStream_t stream);

/* Eigenstress on a near list (evaluate.py:298-326 _eigen_near_list): point f
 * sums source chunks pt_chunk[pt_ptr[f] .. pt_ptr[f+1]), chunk c being the
 * sources ch_src[ch_ptr[c] .. ch_ptr[c+1]).  All index arrays on device. */
int mbem_eigenstress_near(const double *x_field, int64_t nf, const double *src, int64_t ld_src,
                          int64_t ns, const double *density, const int64_t *pt_ptr,
                          const int64_t *pt_chunk, const int64_t *ch_ptr, const int64_t *ch_src,
                          double mu, double lam, double *out, double scale, double beta,
                          hipStream_t stream);

/* Free term on a diagonal block (model/equations.py:566-582 add_block_diagonal
 * at P0): A[3c+a, 3c+b] += scale * D[c,a,b], A pointing at the block. */
int mbem_add_block_diag(double *A, int64_t lda, int64_t nc, const double *D, double scale,
                        hipStream_t stream);

/* y = alpha A x + beta y, A (m, n) row-major with leading dimension lda: the
 * FGMRES matvec (la/solver.py:178) and the block Gauss-Seidel lower
 * couplings (la/preconditioner.py:530-541). */
int mbem_gemv(const double *A, int64_t lda, int64_t m, int64_t n, const double *x, double *y,
              double alpha, double beta, hipStream_t stream);

/* Arnoldi orthogonalization passes (la/solver.py:179-183): h = V w for the k
 * rows of V (ldv), and w += alpha V^T h.  mbem_multidot needs a device work
 * buffer of mbem_multidot_work(k, n) doubles. */
int64_t mbem_multidot_work(int64_t k, int64_t n);
int mbem_multidot(const double *V, int64_t ldv, int64_t k, int64_t n, const double *w, double *h,
                  double *work, hipStream_t stream);
int mbem_multiaxpy(const double *V, int64_t ldv, int64_t k, int64_t n, const double *h, double *w,
                   double alpha, hipStream_t stream);

/* h = V w as mbem_multidot, allocating nothing: the partial sums go to a
 * device array the library holds, MBEM_NOALLOC_SCRATCH_DOUBLES doubles
 * (8 MiB) on each GPU it runs on, until the process ends.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops).  Limit: ceil(n / 512) * k <=
 * MBEM_NOALLOC_SCRATCH_DOUBLES, else MBEM_ESCRATCH; the first call on a
 * device looks up that device's copy.  Calls that overlap on one GPU race on
 * the array; results equal mbem_multidot's bit for bit. */
#define MBEM_NOALLOC_SCRATCH_DOUBLES (1 << 20)
int mbem_multidot_noalloc(const double *V, int64_t ldv, int64_t k, int64_t n, const double *w,
                          double *h, hipStream_t stream);

/* y = a x + b y */
int mbem_axpby(int64_t n, double a, const double *x, double b, double *y, hipStream_t stream);

/* w /= ||w|| with ||w||^2 = *sumsq read on the device, skipped when the norm
 * is not above 1e-300: the Arnoldi normalization V[j+1] = w / H[j+1, j]
 * (la/solver.py:183-184). */
int mbem_normalize(int64_t n, const double *sumsq, double *w, hipStream_t stream);

/* M (m,m) = A[idx, idx]: a block-Jacobi chunk of the assembled operator
 * (la/preconditioner.py:445 eval_block(ch, ch)). */
int mbem_gather_block(const double *A, int64_t lda, const int64_t *idx, int64_t m, double *M,
                      hipStream_t stream);

/* In-place inverse (LAPACK getrf + getri through rocSOLVER) of a chunk: the
 * factorization step of the block-Jacobi rung (la/preconditioner.py:448
 * lu_factor), kept as an explicit inverse so the apply is one GEMV. */
int mbem_invert(double *M, int64_t m, int32_t *ipiv, int32_t *info, hipStream_t stream);

/* In-place inverses of `batch` equal-size m x m chunks stored back to back
 * (stride m*m), one strided-batched getrf + getri; ipiv batch*m int32, info
 * batch int32 (device). */
int mbem_invert_batched(double *M, int64_t m, int64_t batch, int32_t *ipiv, int32_t *info,
                        hipStream_t stream);

/* The same inverses written to C (same layout as M, FP32 when f32 != 0,
 * else FP64): getrf on M in place, then C = A^-1 by solving A C = I with the
 * factors (rocSOLVER getri_outofplace), which takes far fewer launches than
 * getri at the chunk sizes of the block-Jacobi rung.  M != C. */
int mbem_invert_batched_oop(void *M, int64_t m, int64_t batch, int32_t *ipiv, int32_t *info,
                            void *C, int32_t f32, hipStream_t stream);

/* Padded chunk for one strided-batched inverse of unequal chunks: M (mp x mp,
 * FP64 or FP32) = [A[idx, idx] 0; 0 I]; its inverse's leading m x m block is
 * the chunk's inverse.  And a 2-D strided copy (4- or 8-byte elements) that
 * moves that block into the packed inverse store. */
int mbem_gather_block_pad(const double *A, int64_t lda, const int64_t *idx, int64_t m, int64_t mp,
                          void *M, int f32, hipStream_t stream);
int mbem_copy2d(const void *src, int64_t lds, void *dst, int64_t ldd, int64_t rows, int64_t cols,
                int esz, hipStream_t stream);

/* Dense LU of a row-major n x n operator in place and the solve A x = b:
 * the reference dense solve (backends/dense.py:532-571, la/lu.py:110-239 with
 * lu="double").  64-bit sizes and pivots (ipiv n int64, info 1 int64 on the
 * device): the dense path's matrices pass 2^31 elements. */
int mbem_lu_factor(double *A, int64_t n, int64_t *ipiv, int64_t *info, hipStream_t stream);
int mbem_lu_solve(const double *LU, int64_t n, const int64_t *ipiv, double *b, int64_t nrhs,
                  hipStream_t stream);

/* Blocked triangular solves on the factors of mbem_lu_factor (the solve of
 * the reference's LU path, la/lu.py:180-239, and the condition estimate's
 * solves): mbem_lu_prepare inverts the 64 x 64 diagonal blocks of L and U
 * once per factorization into inv_l, inv_u (mbem_lu_inv_doubles(n) doubles
 * each); mbem_lu_solve_blocked then solves A x = b (adjoint == 0) or
 * A^T x = b (adjoint != 0), one launch per diagonal block, with perm the
 * row order of the pivots (P^T v = v[perm], perm from ipiv by applying its
 * interchanges in order to 0 .. n-1) and work 2 n doubles.  b and x may be
 * the same vector. */
int64_t mbem_lu_inv_doubles(int64_t n);
int mbem_lu_prepare(const double *LU, int64_t n, double *inv_l, double *inv_u, hipStream_t stream);
int mbem_lu_solve_blocked(const double *LU, int64_t n, const int64_t *perm, const double *inv_l,
                          const double *inv_u, const double *b, double *x, double *work,
                          int adjoint, hipStream_t stream);

/* The same in single precision for lu="single"/"gpu" (la/lu.py:119-239:
 * single-precision factors, every solve refined in double by the caller):
 * sgetrf of the row-major FP32 copy (n x n, in place), the diagonal-block
 * inverses formed in FP64 from the FP32 factors, and the blocked solves
 * reading FP32 factors with FP64 vectors and accumulation. */
int mbem_lu_factor_f32(float *A, int64_t n, int64_t *ipiv, int64_t *info, hipStream_t stream);
int mbem_lu_prepare_f32(const float *LU, int64_t n, double *inv_l, double *inv_u, hipStream_t stream);
int mbem_lu_solve_blocked_f32(const float *LU, int64_t n, const int64_t *perm,
                              const double *inv_l, const double *inv_u, const double *b,
                              double *x, double *work, int adjoint, hipStream_t stream);

/* A^T x = b with the factors of mbem_lu_factor, in place, nrhs columns: the
 * transposed solves of the 1-norm condition estimate that LAPACK's gecon
 * makes on the reference's factors (la/lu.py:241-248). */
int mbem_lu_solve_adjoint(const double *LU, int64_t n, const int64_t *ipiv, double *b,
                          int64_t nrhs, hipStream_t stream);

/* colsum[c] = sum_r |A[r, c]| for an (m, n) row-major block, summed in a
 * fixed order through work (mbem_colsum_work(n) doubles): the 1-norm
 * max_c colsum[c] of the factored matrix (la/lu.py:142-145 anorm). */
int64_t mbem_colsum_work(int64_t n);
int mbem_colsum_abs(const double *A, int64_t lda, int64_t m, int64_t n, double *work,
                    double *colsum, hipStream_t stream);

/* Block-Jacobi apply (la/preconditioner.py:456-462 + _bj_solve :601-630):
 * z[idx[c][i]] = sum_j inv_c[i,j] r[idx[c][j]] for every chunk c.  chunk_ptr
 * (nchunks+1) delimits each chunk's DOF list in idx, inv_ptr its (m_c, m_c)
 * row-major inverse in inv; wave_table (nwaves, 2) = (chunk, first row) with
 * MBEM_BJ_ROWS_PER_WAVE rows per wave.  All on device. */
int mbem_bj_apply(const int64_t *wave_table, int64_t nwaves, const int64_t *chunk_ptr,
                  const int64_t *idx, const int64_t *inv_ptr, const double *inv, const double *r,
                  double *z, hipStream_t stream);

/* ---- mixed precision ------------------------------------------------- */

/* out = lo(A) for an (m, n) row-major FP64 block: BF16 bits (bf16 != 0, round
 * to nearest even) or FP32, leading dimension ldo.  Far-field storage of the
 * FGMRES matvec (la/solver.py:178 operator, stored low). */
int mbem_cast_lo(const double *A, int64_t lda, int64_t m, int64_t n, void *out, int64_t ldo,
                 int bf16, hipStream_t stream);

/* Near-field pattern over nb unknown blocks: block i is near source j when
 * |pts_i - cen_j| < rad_j.  Two passes: with counts != NULL write the
 * per-row counts; with col != NULL (and rowptr = exclusive prefix sum of the
 * counts) write the ascending column lists. */
int mbem_near_pattern(const double *pts, const double *cen, const double *rad, int64_t nb,
                      int64_t *counts, const int64_t *rowptr, int64_t *col, hipStream_t stream);

/* Near-field blocks: vals[b] (3x3 FP64) = A's block (row-block i, col[b]) for
 * b in rowptr[i] .. rowptr[i+1]; the same entries of the low copy are zeroed,
 * so A = lo + near exactly in the near blocks. */
int mbem_near_extract(const double *A, int64_t lda, int64_t nbr, const int64_t *rowptr,
                      const int64_t *col, int64_t nnzb, double *vals, void *lo, int64_t ldo,
                      int bf16, hipStream_t stream);

/* xf = (float) x: the FP32 copy of a Krylov vector that mbem_gemv_lo reads. */
int mbem_vec_to_f32(const double *x, int64_t n, float *xf, hipStream_t stream);

/* y = alpha F x + beta y, F stored BF16 / FP32 (16-byte aligned rows: ldf a
 * multiple of 8 / 4), FP32 x (16-byte aligned), FP32 products and
 * accumulation, FP64 y. */
int mbem_gemv_lo(const void *F, int64_t ldf, int64_t m, int64_t n, const float *x, double *y,
                 double alpha, double beta, int bf16, hipStream_t stream);

/* int8 far-field storage: F[r, c] = scale[r * nseg + c / 64] * (out[r, c] - 128)
 * with offset-binary bytes out = q + 128, q = rint(A / s) clamped to
 * [-127, 127], and one FP32 scale s = (float)(max |A| / 127) per row and
 * 64-column segment (q = 0 where s = 0), nseg = mbem_i8_segments(n); out
 * 16-byte aligned with ldo a multiple of 16.  Same role as mbem_cast_lo
 * (la/solver.py:178 operator, stored low). */
int64_t mbem_i8_segments(int64_t n);
int mbem_cast_i8(const double *A, int64_t lda, int64_t m, int64_t n, uint8_t *out, int64_t ldo,
                 float *scale, int64_t nseg, hipStream_t stream);

/* mbem_cast_i8 plus the code's rounding remainder e = rint((A - s q) / s *
 * 2^24) clamped to +-(2^23 - 1), so that s (q + e 2^-24) is A to within
 * s 2^-25 (2e-10 of the segment's largest entry).  The remainders are packed
 * four to three little-endian 32-bit words (24-bit two's complement): row r's
 * columns 4k .. 4k+3 are the 12 bytes at resid + r * ldr + 12 k.  resid
 * 4-byte aligned, ldr (bytes) a multiple of 4 and at least
 * mbem_i8e_resid_bytes(n).  Same role as mbem_cast_i8 (la/solver.py:178
 * operator, stored in 4 bytes an entry); its product is mbem_gemv_i8e. */
int64_t mbem_i8e_resid_bytes(int64_t n);
int mbem_cast_i8e(const double *A, int64_t lda, int64_t m, int64_t n, uint8_t *out, int64_t ldo,
                  float *scale, int64_t nseg, uint8_t *resid, int64_t ldr, hipStream_t stream);

/* Hold the near blocks out of A while its int8 copy is built: restore == 0
 * copies them to vals (9 doubles per block, as mbem_near_extract) and zeroes
 * them in A; restore != 0 writes vals back, leaving A as it was. */
int mbem_near_hold(double *A, int64_t lda, int64_t nbr, const int64_t *rowptr, const int64_t *col,
                   int64_t nnzb, double *vals, int restore, hipStream_t stream);

/* y = alpha F x + beta y for the int8 far field above (ldf a multiple of
 * 16, F and x 16-byte aligned), FP32 x, FP32 products and accumulation,
 * FP64 y. */
int mbem_gemv_i8(const uint8_t *F, int64_t ldf, const float *scale, int64_t nseg, int64_t m,
                 int64_t n, const float *x, double *y, double alpha, double beta,
                 hipStream_t stream);

/* y = alpha F x + beta y for the split far field of mbem_cast_i8e,
 * F[r, c] = scale * (out[r, c] - 128 + e[r, c] 2^-24), FP64 x (16-byte
 * aligned) and FP64 sums: the FGMRES operator of la/solver.py:178 (A v) with
 * the near field added by mbem_bsr_gemv, read from 4 bytes an entry instead
 * of the FP64 matrix's 8. */
int mbem_gemv_i8e(const uint8_t *F, int64_t ldf, const uint8_t *resid, int64_t ldr,
                  const float *scale, int64_t nseg, int64_t m, int64_t n, const double *x,
                  double *y, double alpha, double beta, hipStream_t stream);

/* y += alpha N x for a 3x3 block-CSR N in FP64 (the near field). */
int mbem_bsr_gemv(int64_t nbr, const int64_t *rowptr, const int64_t *col, const double *vals,
                  const double *x, double *y, double alpha, hipStream_t stream);

/* FP32 block-Jacobi chunk inverses, every chunk in one launch sequence
 * (blocked Gauss-Jordan with partial pivoting; replaces the per-chunk
 * lu_factor of la/preconditioner.py:448 for the mixed method's FP32
 * inverses).  desc holds mbem_bj_desc_fields() int64 per chunk: offset of
 * its DOF list in idx, size m (<= mmax <= 2048), working matrix offset (a
 * multiple of 4) and row pitch (a multiple of mbem_bj_panel_width(), >= m),
 * output offset and pitch (>= m), row-map offset, saved-rows offset (a
 * multiple of 4; panel width x pitch floats).  desc_host is the same table in
 * host memory: it is checked against the lengths given (elements) before
 * anything is launched.  Gather: W_c = (float) A[idx_c, idx_c], the row map
 * reset, info zeroed (a caller changing A may do so once this is queued).
 * Invert: out_c (row-major, pads zero) = W_c^-1; info[c] = 1 + the first
 * pivot column found zero (or NaN), 0 when none; W, rowmap, piv and saved
 * are working memory. */
int mbem_bj_desc_fields(void);
int mbem_bj_panel_width(void);
int mbem_bj_gather_f32(const double *A, int64_t lda, const int64_t *idx, int64_t idx_len,
                       const int64_t *desc_host, const int64_t *desc, int64_t nchunks,
                       int64_t mmax, float *W, int64_t w_len, int32_t *rowmap, int64_t map_len,
                       int32_t *info, hipStream_t stream);
int mbem_bj_invert_f32(const int64_t *desc_host, const int64_t *desc, int64_t nchunks,
                       int64_t mmax, float *W, int64_t w_len, int32_t *rowmap, int32_t *piv,
                       int64_t map_len, float *saved, int64_t saved_len, int32_t *info,
                       float *out, int64_t out_len, hipStream_t stream);

/* FP32 block-Jacobi: chunk gather, strided-batched sgetrf + sgetri, apply. */
int mbem_gather_block_f32(const double *A, int64_t lda, const int64_t *idx, int64_t m, float *M,
                          hipStream_t stream);
int mbem_invert_batched_f32(float *M, int64_t m, int64_t batch, int32_t *ipiv, int32_t *info,
                            hipStream_t stream);
int mbem_bj_apply_f32(const int64_t *wave_table, int64_t nwaves, const int64_t *chunk_ptr,
                      const int64_t *idx, const int64_t *inv_ptr, const float *inv, const double *r,
                      double *z, hipStream_t stream);

/* FP32 block-Jacobi on padded rows (the form la/preconditioner.py uses):
 * chunk c holds m = chunk_ptr[c+1] - chunk_ptr[c] DOFs idx[chunk_ptr[c]..],
 * its inverse at inv + inv_ptr[c] with rows of ldm = pad_ptr[c+1] - pad_ptr[c]
 * floats (a multiple of 4, pad columns zero).  mbem_gather_f32 fills rr[p] =
 * r[gidx[p]] (0 where gidx[p] < 0: the pads) in that padded chunk order;
 * mbem_bj_apply_f32v then sets z[idx[i]] = sum_j inv_c[i, j] rr_c[j] for the
 * (chunk, first row) pairs of wave_table, MBEM_BJ32_ROWS_PER_WAVE rows each.
 * inv and rr 16-byte aligned. */
int mbem_gather_f32(const int64_t *gidx, int64_t np, const double *r, float *rr, hipStream_t stream);
int mbem_bj_apply_f32v(const int64_t *wave_table, int64_t nwaves, const int64_t *chunk_ptr,
                       const int64_t *pad_ptr, const int64_t *idx, const int64_t *inv_ptr,
                       const float *inv, const float *rr, double *z, hipStream_t stream);

/* ---- Batched Laplace-state solves (rheology.py:479-524, dense.py:650-687) ----
 * The reference solves its Laplace states one after another: a re-assembly
 * or recombination of the bases at each state's moduli (rebuild_for_materials)
 * and a solve per state (runner.py:357-391).  These entries serve one solve
 * of every (state, right-hand side) pair at once: with A(s) = (A0 + g(s) A1)
 * D(s) and y = D x, column c of a row-major n x K block is system
 * B(g_c) y_c = b_c.  None of them replaces a reference function one-for-one;
 * together they stand in for that state loop. */

/* ---- Lagrange P0/P1/P2 nodal densities (csrc/nodal.hip) -----------------
 * Replace tri_nodal.py's numba drivers: the matrices and basis stacks of
 * u_basis_matrices / t_basis_matrices / u_matrix_direct / t_matrix_direct
 * (tri_nodal.py:1257-1306, via _u/_t_basis_nodal and _u/_t_direct_nodal
 * :958-1043), the contractions u_disp_contract / t_disp_contract /
 * kelvin_stress_contract / dd_stress_contract / eigenstress_contract
 * (:1309-1363, via :1055-1172) and traction_matrix (:1370-1404, via
 * _traction_matrix_nodal :1175-1227).  Device pointers; columns
 * 3 (K s + k) + j; densities (K ns, 3).
 *
 * The spec of a launch: ispec (host, 18 int64) = {order, K, q_split,
 * n_terms, nq, table degree for n = -1,1,3,5,7, edge degree for
 * m = -1,1,3,5, kernel degree for n = 1,3,5,7} (-1 = none), fspec (host, 4)
 * = {D*, D*_distant, small-|u| / rho, series |u| / rho}, tables (device,
 * mbem_nodal_table_doubles(order, n_terms, nq) doubles) = series
 * coefficients C(-m/2, j) (4 x n_terms) | node lattice (K x 3) | Gauss
 * barycentric points (nq x 3) | weights (nq) | shape values (nq x K). */
int64_t mbem_nodal_geo_doubles(int64_t ns);
int64_t mbem_nodal_table_doubles(int64_t order, int64_t n_terms, int64_t nq);
/* nodal frames of the sources (tri_nodal.py:418-470): geo[30 ns] */
int mbem_nodal_frames(const double *tri_verts, int64_t ns, double *geo, hipStream_t stream);
/* kind 0 U, 1 T; coeffs (host: 3 u_coeffs or 6 t_coeffs) gives the direct
 * matrix out (3 nf x ld), coeffs NULL the basis stack: block b at
 * out + b bstride (3 or 6 blocks). */
int mbem_nodal_matrix(int kind, const double *x_field, int64_t nf, const double *tri_verts,
                      const double *normals, const double *eps, int64_t ns, const double *geo,
                      const int64_t *ispec, const double *fspec, const double *tables,
                      const double *coeffs, double *out, int64_t ld, int64_t bstride,
                      hipStream_t stream);
/* kind 0 U / 1 T displacement (out nf x 3, coeffs on the host), 2 force ->
 * stress, 3 slip -> total stress, 4 eigenstress (out nf x 9); (mu, nu, lam)
 * for the stresses. */
int mbem_nodal_contract(int kind, const double *x_field, int64_t nf, const double *tri_verts,
                        const double *normals, const double *eps, int64_t ns, const double *geo,
                        const int64_t *ispec, const double *fspec, const double *tables,
                        const double *density, const double *coeffs, double mu, double nu,
                        double lam, double *out, hipStream_t stream);
/* traction rows accumulated into out (3 nf x ld, ld >= 3 n_cols): which 0
 * slip -> elastic, 1 force, 2 eigenstress, 3 slip -> total; ids (device,
 * ns x K) must lie in [0, n_cols) (checked by the caller). */
int mbem_nodal_traction(int which, const double *x_field, const double *n_field, int64_t nf,
                        const double *tri_verts, const double *normals, const double *eps,
                        int64_t ns, const double *geo, const int64_t *ispec, const double *fspec,
                        const double *tables, const int64_t *ids, int64_t n_cols, double mu,
                        double nu, double lam, double *out, int64_t ld, hipStream_t stream);

/* Tile configurations of the FP64 MFMA GEMM (the autotuner's candidates):
 * the count, and {BM, BN, BK, waves along M, waves along N} of one. */
int mbem_mgemm_configs(void);
int mbem_mgemm_config(int cfg, int32_t *out5);
/* 1 if configuration cfg has a form within the LDS: the two-operator form of
 * mbem_gemm_g (dual != 0) or the one-operator form of mbem_gemm_rm and
 * mbem_gemm_grouped (dual == 0); 0 otherwise. */
int mbem_mgemm_config_fits(int cfg, int dual);

/* Y = alpha (A0 X + A1 X diag(g)) + beta Y: A0, A1 row-major m x kd (even
 * lda, 16-byte aligned), X kd x ncols (even ldx, aligned), g (ncols).
 * FP64 matrix cores, tile configuration cfg. */
int mbem_gemm_g(const double *A0, const double *A1, int64_t lda, int64_t m, int64_t kd,
                const double *X, int64_t ldx, int64_t ncols, const double *g, double *Y,
                int64_t ldy, double alpha, double beta, int cfg, hipStream_t stream);

/* Row-major Y (m x ncols) = alpha A (m x kd) X (kd x ncols) + beta Y on the
 * project's FP64 matrix-core kernel, tile configuration cfg (even lda and
 * ldx, 16-byte aligned A and X): the batched operator as one GEMM on
 * [A0 | A1] [X ; X diag(g)], the same product as mbem_dgemm_rowmajor. */
int mbem_gemm_rm(const double *A, int64_t lda, int64_t m, int64_t kd, const double *X, int64_t ldx,
                 int64_t ncols, double *Y, int64_t ldy, double alpha, double beta, int cfg,
                 hipStream_t stream);

/* Row-major C (m x n) = alpha A (m x k) B (k x n) + beta C through the
 * library DGEMM (rocBLAS, matrix cores, 64-bit sizes): the autotuner's library
 * candidate for the batched operator, applied to [A0 | A1] [X ; X diag(g)]. */
int mbem_dgemm_rowmajor(int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                        const double *B, int64_t ldb, double beta, double *C, int64_t ldc,
                        hipStream_t stream);
/* C (m x n) = alpha A B^T + beta C, A (m x k) and B (n x k) row-major (the
 * coarse space's Gram matrices). */
int mbem_dgemm_abt(int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                   const double *B, int64_t ldb, double beta, double *C, int64_t ldc,
                   hipStream_t stream);

/* Grouped GEMM over a device work table of nwork items (mbem_work_item_bytes
 * each: chunk inverse offset and leading dimension, chunk size, block offset,
 * column count, tile row, tile column): Y[rows of chunk, cols] = Inv X[...].
 * The batched block-Jacobi apply (preconditioner.py:601-630 for many
 * columns). */
int mbem_work_item_bytes(void);
int mbem_gemm_grouped(const double *inv, const double *X, int64_t ldx, double *Y, int64_t ldy,
                      const void *work, int64_t nwork, int cfg, hipStream_t stream);

/* Column-wise multi-dot of cnt (<= 8) blocks V_i (stride bstride) with W:
 * partial sums over nslots row slices (a multiple of 4), then reduced into
 * out[c * ldo + o0 + i] (set, or added when accumulate); copy (cnt x K per
 * column) optional.  The Arnoldi projections of K runs (solver.py:115-245). */
int mbem_bdot(const double *V, int64_t bstride, int64_t cnt, const double *W, int64_t n, int64_t K,
              double *partial, int64_t nslots, hipStream_t stream);
int mbem_breduce(const double *partial, int64_t nslots, int64_t cnt, int64_t K, double *out,
                 int64_t ldo, int64_t o0, int accumulate, double *copy, hipStream_t stream);

/* Y += sign * sum_{i < cnt} coef[c * ldc + i] B_i (cnt read from cnt_dev when
 * non-null); Y = X diag(scale); Y = X; y[:] = v. */
int mbem_baxpy(const double *B, int64_t bstride, int64_t cnt, const int *cnt_dev,
               const double *coef, int64_t ldc, double sign, double *Y, int64_t n, int64_t K,
               hipStream_t stream);
int mbem_bscale(const double *X, const double *scale, double *Y, int64_t n, int64_t K,
                hipStream_t stream);
int mbem_bcopy(const double *X, double *Y, int64_t nk, hipStream_t stream);
int mbem_fill(double *y, int64_t n, double v, hipStream_t stream);

/* FGMRES bookkeeping on the device, one thread per run: cycle start
 * (beta, 1/beta, gvec), Arnoldi step j (H column j = hp1 + hp2, the two
 * CGS passes, K x (m+1) each; H[j+1, j]; Givens rotation j; gvec; residual
 * estimate; breakdown zeroes the next basis scale), and
 * y = R^{-1} gvec over the steps done (solver.py:115-245, _solve_upper
 * solver.py:503-510). */
int mbem_fg_start(int64_t K, int64_t m, const double *sumsq, double *beta, double *scale,
                  double *gvec, int *jdone, hipStream_t stream);
int mbem_fg_givens(int64_t K, int64_t m, int64_t j, const double *hp1, const double *hp2, double *H,
                   double *cs, double *sn, double *gvec, const double *sumsq, double *scale,
                   double *res, int *jdone, hipStream_t stream);
int mbem_fg_solve(int64_t K, int64_t m, const double *H, const double *gvec, const int *jdone,
                  double *y, hipStream_t stream);

/* out = A0[idx, idx] + g A1[idx, idx] (m x m): a chunk of B(g); out = A[idx, idx]
 * for a full permutation into cluster order. */
int mbem_gather_comb(const double *A0, const double *A1, int64_t lda, const int64_t *idx, int64_t m,
                     double g, double *out, hipStream_t stream);
int mbem_permute(const double *A, int64_t lda, const int64_t *idx, int64_t n, double *out,
                 int64_t ldo, hipStream_t stream);
/* The same chunk padded to mp x mp with an identity block, for one batched
 * inverse over chunks of unequal size. */
int mbem_gather_comb_pad(const double *A0, const double *A1, int64_t lda, const int64_t *idx,
                         int64_t m, int64_t mp, double g, double *out, hipStream_t stream);

/* HIP graphs: capture the launches made on a (non-null) stream, instantiate,
 * replay, destroy; and an async copy that a graph can hold. */
int mbem_graph_begin(hipStream_t stream);
int mbem_graph_end(hipStream_t stream, void **exec_out);
int mbem_graph_launch(void *exec, hipStream_t stream);
int mbem_graph_destroy(void *exec);
int mbem_memcpy_async(void *dst, const void *src, int64_t bytes, hipStream_t stream);

/* ---- Learned FGMRES preconditioner (no counterpart in the reference) ----
 * A point-cloud network over the boundary elements that adds a correction to
 * the block-Jacobi output inside FGMRES (the reference's preconditioner,
 * preconditioner.py:314-543, stays the fallback and the comparison).  FP32
 * training kernels (forward, backward, Adam) and BF16 matrix-core inference. */

/* k (8 or 16) nearest neighbours of each of P points (FP32 xyz), self
 * excluded, ties by index: nbr (P x k). */
int mbem_knn(const float *pts, int64_t P, int k, int32_t *nbr, hipStream_t stream);

/* FP32 GEMM with free strides: C = alpha A B + beta C (+ bias[j]) (relu),
 * A(i,k) = A[i sa0 + k sa1], B(k,j) = B[k sb0 + j sb1], C row-major (ldc).
 * splits > 1: split-K through work (splits x M x N floats; no bias/relu). */
int mbem_sgemm(const float *A, int64_t sa0, int64_t sa1, const float *B, int64_t sb0, int64_t sb1,
               float *C, int64_t ldc, int64_t M, int64_t N, int64_t K, float alpha, float beta,
               const float *bias, int relu, int64_t splits, float *work, hipStream_t stream);

/* out[c] = beta out[c] + sum_i X[i, c] (X: P x C; work: nslices x C floats). */
int mbem_colsum(const float *X, int64_t P, int64_t C, float *out, float *work, int64_t nslices,
                float beta, hipStream_t stream);

/* kNN mean m_i = (1/k) sum_q a[nbr[i,q]] and its backward through the
 * reverse graph (CSR rptr/ridx of in-neighbours). */
int mbem_knn_mean(const float *a, int64_t P, int64_t C, const int32_t *nbr, int k, float *m,
                  hipStream_t stream);
int mbem_knn_mean_bwd(const float *dm, int64_t P, int64_t C, const int32_t *rptr,
                      const int32_t *ridx, int k, float *da, hipStream_t stream);

/* relu backward (dY *= Y > 0); column concatenation and its gradient split;
 * X[i, :] += scale v. */
int mbem_relu_bwd(const float *Y, float *dY, int64_t n, hipStream_t stream);
int mbem_concat(const float *A, int64_t C1, const float *B, int64_t C2, int64_t P, float *out,
                hipStream_t stream);
int mbem_split(const float *d, int64_t C1, int64_t C2, int64_t P, float *dA, int accumulate_a,
               float *dB, hipStream_t stream);
int mbem_add_row(float *X, int64_t P, int64_t C, const float *v, float scale, hipStream_t stream);

/* Relative loss |s y - t|^2 / |t|^2 (loss[0]) and dy (part: 2 x ceil(n/256)). */
int mbem_loss(const float *y, const float *tgt, int64_t n, float s, float *part, float *dy,
              float *loss, hipStream_t stream);

/* Adam step with bias-corrected rate lr_t and decoupled weight decay wd. */
int mbem_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr_t, float b1,
              float b2, float eps, float wd, hipStream_t stream);

/* Inference GEMM on the BF16 matrix cores: C (P x N) = act(A (P x K) W^T +
 * bias), FP32 in and out, BF16 operands, FP32 accumulation; K % 16 == 0. */
int mbem_hgemm(const float *A, int64_t lda, const float *W, int64_t K, int64_t N, int64_t P,
               const float *bias, int relu, float *C, int64_t ldc, hipStream_t stream);

/* Low-rank correction from the network output Y (P x 6K): Phi = Y[:, 0:3K],
 * Psi = Y[:, 3K:6K] as 3P x K bases (row 3p + a = Y[p, a K + off ..]).
 * C (B x K) = Bas^T V_b (V FP64 or FP32, rows ldv apart; part: nslices x B x K
 * doubles); Out_b (+)= alpha diag(rowscale) Bas C_b (rowscale optional, FP64
 * per unknown); dY[:, off..] = sum_b U_b C_b^T; batch
 * loss (1/B) sum_b |pred_b - t_b|^2 / |t_b|^2 with its gradient. */
int mbem_lr_coef(const float *Y, int64_t P, int64_t K, int64_t off, const void *V, int64_t ldv,
                 int64_t B, int v_fp64, double *part, int64_t nslices, double *C, hipStream_t stream);
int mbem_lr_expand(const float *Y, int64_t P, int64_t K, int64_t off, const double *C, int64_t B,
                   void *Out, int64_t ldo, int out_fp64, double alpha, int accumulate,
                   const double *rowscale, hipStream_t stream);
int mbem_lr_outer(const float *U, int64_t ldu, const double *C, int64_t B, int64_t P, int64_t K,
                  int64_t off, float *dY, hipStream_t stream);
int mbem_lr_loss(const float *pred, const float *tgt, int64_t n, int64_t B, float *dpred,
                 float *loss_b, hipStream_t stream);

/* Coarse space of the two-level block-Jacobi preconditioner (an addition to
 * the reference's ladder, la/preconditioner.py:421-470, used by the
 * mixed-precision solve): W (n x G q) has q shape functions (q = 3, 6 or 12)
 * on each group g of DOFs, the groups' DOFs listed in perm (group g at
 * perm[gptr[g] .. gptr[g+1])) with values Wp[t q + j].
 * mbem_coarse_image: Ct = (A W)^T, (G q) x m row-major with leading
 * dimension ldc >= m, one pass over the m rows of A (row-major, lda).
 * mbem_coarse_expand: x[perm[t]] += alpha sum_j Wp[t q + j] u[grp[t] q + j]
 * for the nt listed DOFs. */
int mbem_coarse_image(const double *A, int64_t lda, int64_t m, const int64_t *perm,
                      const double *Wp, const int64_t *gptr, int64_t G, int q, double *Ct,
                      int64_t ldc, hipStream_t stream);
/* The coarse image of the mixed operator F + N (la/mixed.py): Ct = (F W)^T
 * from the int8 far field (codes, FP32 scales per row and 64 columns, near
 * entries zero; overwrites Ct), then Ct += (N W)^T from the near field's
 * FP64 3 x 3 block-CSR; inv[d] is the listed position of DOF d (the inverse
 * of perm), grp[t] the group of listed position t. */
int mbem_coarse_image_i8(const uint8_t *F, int64_t ldf, const float *scale, int64_t nseg, int64_t m,
                         const int64_t *perm, const double *Wp, const int64_t *gptr, int64_t G, int q,
                         double *Ct, int64_t ldc, hipStream_t stream);
int mbem_coarse_image_bsr(int64_t nbr, const int64_t *rowptr, const int64_t *col, const double *vals,
                          const int64_t *inv, const int64_t *grp, const double *Wp, int q, double *Ct,
                          int64_t ldc, hipStream_t stream);
/* The coarse space's factors from its Gram matrix G (K x K, overwritten),
 * on the device with no host wait: D = diag(sqrt(G_ii)) (d, K doubles),
 * Cholesky L L^T = D^-1 G D^-1 + delta I (rocSOLVER potrf), L^-1 (trtri),
 * then Tt = L^-1 D^-1 and T = Tt^T (both K x K row-major), so that Q = C T
 * has orthonormal columns for C^T C = G.  info: 2 ints (potrf, trtri). */
int mbem_coarse_factor(double *G, int64_t K, double delta, double *d, int32_t *info, double *Tt,
                       double *T, hipStream_t stream);
/* G = X X^T (k x k) for X (k x n) row-major with leading dimension ldx,
 * split-K on matrix cores (strided-batched DGEMM over n-slices, then a sum):
 * the coarse space's Gram matrix.  work: mbem_gram_work(k, n) doubles. */
int64_t mbem_gram_work(int64_t k, int64_t n);
int mbem_gram(const double *X, int64_t ldx, int64_t k, int64_t n, double *G, double *work,
              hipStream_t stream);
/* The coarse projection with Q^T (K x n, row-major, pitch ldq) stored in
 * FP32, FP64 vectors and sums: s = Q^T r, rest = r - Q s.  work:
 * mbem_coarse_project_work(K, n) doubles.  The deflation step of the
 * two-level block-Jacobi (no counterpart in the reference's preconditioner,
 * la/preconditioner.py:421-470, which it wraps). */
int64_t mbem_coarse_project_work(int64_t K, int64_t n);
int mbem_coarse_project_f32(const float *Qt, int64_t ldq, int64_t K, int64_t n, const double *r,
                            double *s, double *rest, double *work, hipStream_t stream);
int mbem_coarse_expand(const double *u, const int64_t *perm, const double *Wp, const int64_t *grp,
                       int64_t nt, int q, double alpha, double *x, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MBEM_HIP_H */
