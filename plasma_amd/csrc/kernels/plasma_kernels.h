/*
 * plasma_kernels.h -- the thin C ABI between the C host code and the HIP
 * kernels for gfx950.
 *
 * Every launcher enqueues on the given stream and returns 0, or a negative
 * PlasmaError code when the launch itself fails.  Launchers take
 * descriptors by value (their `matrix` is device memory) and the device
 * status word of the sequence: every kernel returns immediately when that
 * word is non-zero, which is how the tile tasks of the reference skip their
 * work after a failure (reference core_blas/core_zpotrf.c:74).
 *
 * "Lower-equivalent" convention: for uplo == PlasmaUpper the kernels work on
 * L = U^T, reading the stored upper tiles transposed, so one algorithm
 * serves both triangles.
 */
#ifndef PLASMA_KERNELS_H
#define PLASMA_KERNELS_H

#include <hip/hip_runtime_api.h>
#include "plasma_descriptor.h"
#include "plasma_batched.h"

#ifdef __cplusplus
extern "C" {
#endif

/* LAPACK band array (device) <-> a band descriptor's tiles, kd the band
 * width of the stored triangle (upper: superdiagonals).  to_tiles writes
 * every element of the band's tiles (zero outside the band). */
int plasma_kern_band_layout(plasma_desc_t A, void *pAB, int ldab, int kd,
                            int upper, int to_tiles, const int *status,
                            hipStream_t s);

/* Fill n ints with v (a kernel, never hipMemset: safe under graph capture). */
int plasma_kern_fill_int(int *p, int v, int n, hipStream_t s);
/* *dst = *src on the stream (one int; a kernel, capture-safe). */
int plasma_kern_copy_int(int *dst, const int *src, hipStream_t s);

/*
 * LAPACK column-major <-> tile layout (reference compute/pztr2desc.c,
 * pzdesc2tr.c, pzge2desc.c, pzdesc2ge.c).  region: PlasmaGeneral copies all
 * tiles of the view, PlasmaLower the tiles m >= n, PlasmaUpper m <= n.
 * pA is the device address of the view's first element, leading dim lda.
 */
int plasma_kern_lapack2tile(plasma_enum_t region, const void *pA, int lda,
                            plasma_desc_t A, const int *status,
                            hipStream_t s);
int plasma_kern_tile2lapack(plasma_enum_t region, plasma_desc_t A, void *pA,
                            int lda, const int *status, hipStream_t s);

/*
 * Cholesky factorization of diagonal tile k (reference core_zpotrf.c:54-82
 * called from pzpotrf.c:41-45), which also writes W_kk = L(k,k)^{-1} (lower,
 * zeros above, ld = A.mb) to inv + k * A.mb * A.mb.  On a non positive (or
 * NaN) pivot at local column j the status word is set to nb*k + j + 1 if it
 * was 0.
 */
int plasma_kern_potrf_diag(plasma_enum_t uplo, plasma_desc_t A, int k,
                           void *inv, int *status, hipStream_t s);

/* Diagonal tile k factored with only the inverses of its 32 x 32 diagonal
 * blocks written to W_kk (the rest of W_kk is left as it was), for the
 * blocked panel solve below. */
int plasma_kern_potrf_diag_blocks(plasma_enum_t uplo, plasma_desc_t A, int k,
                                  void *inv, int *status, hipStream_t s);
/* Panel solve P(m) = A(m,k) L_kk^{-T} by 32-column block substitution with
 * the diagonal-block inverses; same output as plasma_kern_trsm_panel.
 * plasma_panel_blocked_ok: whether its LDS fits A's tile size. */
int plasma_panel_blocked_ok(plasma_desc_t A);
int plasma_kern_trsm_panel_blocked(plasma_enum_t uplo, plasma_desc_t A, int k,
                                   const void *inv, void *panel, int pbase,
                                   int pcols, int pcol, const int *status,
                                   hipStream_t s);

/* Public BLAS-3 (kernels/blas3.hip).  C = alpha op(A) op(B) + beta C on
 * general tile descriptors (views allowed), transa / transb PlasmaNoTrans or
 * (Conj)Trans; tri = PlasmaLower / PlasmaUpper restricts the update to that
 * triangle of C (syrk), PlasmaGeneral updates all of C.  alpha == 0 reads no
 * operand, beta == 0 reads no C. */
int plasma_kern_gemm(int transa, int transb, double alpha, plasma_desc_t A,
                     plasma_desc_t B, double beta, plasma_desc_t C,
                     plasma_enum_t tri, const int *status, hipStream_t s);
/* Diagonal tiles of a triangular A (uplo, diag) as lower-equivalent lower
 * triangles (unit diagonal for PlasmaUnit, zeros above) into tiles (0, k) of
 * the one-tile-row descriptor D (D.m = A.mb, D.n = A.n). */
int plasma_kern_trsm_diag_prep(plasma_enum_t uplo, plasma_enum_t diag,
                               plasma_desc_t A, plasma_desc_t D,
                               const int *status, hipStream_t s);
/* W_k = D(0,k)^{-1} for the lower triangular tiles of D, at
 * inv + k * mb * mb (leading dimension mb). */
int plasma_kern_trtri_tiles(plasma_desc_t D, void *inv, const int *status,
                            hipStream_t s);

/* W_kk = L(k,k)^{-1} for every diagonal tile of an already factored A, one
 * workgroup per tile (used by potrs without a preceding factorization). */
int plasma_kern_trtri_diag(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                           const int *status, hipStream_t s);

/* Panel buffers.  The solved panels of a group of g consecutive tile
 * columns pbase .. pbase+g-1 share one buffer of (mt-1) tiles of
 * mb x pcols elements (pcols = g * nb): tile m (> pbase) at
 * panel + (m - pbase - 1) * mb * pcols, column major, ld = mb, whose
 * columns [j * nb, (j + 1) * nb) hold P(m) of tile column pbase + j. */

/* Panel solve P(m) = A(m,k) L(k,k)^{-T} = A(m,k) W_kk^T for m > k
 * (reference pzpotrf.c:47-57; upper: :90-99), as an MFMA product into
 * column block pcol of the group buffer.  A(m,k) itself is written later,
 * by the copy-back of plasma_kern_trailing_update. */
int plasma_kern_trsm_panel(plasma_enum_t uplo, plasma_desc_t A, int k,
                           const void *inv, void *panel, int pbase, int pcols,
                           int pcol, const int *status, hipStream_t s);

/* Trailing update A(m,n) -= P(m) P(n)^T for n in [n_first, n_end) and
 * m in [n, mt), lower triangle only on diagonal tiles, with P(m) the first
 * kdepth columns of the group buffer tile m: the syrk and gemm tasks of
 * kdepth / nb consecutive steps in one product (reference pzpotrf.c:58-78;
 * upper: :100-125).  The same launch copies the panels of the group's first
 * copy_cols tile columns into A(m, pbase + j), m > pbase + j; an empty
 * column range with copy_cols > 0 is a copy only.  max_wgs > 0 caps the
 * grid (its workgroups loop over the jobs), leaving workgroup slots free
 * for the lookahead stream. */
int plasma_kern_trailing_update(plasma_enum_t uplo, plasma_desc_t A,
                                int pbase, int kdepth, int n_first,
                                int n_end, const void *panel, int pcols,
                                int copy_cols, int max_wgs,
                                const int *status, hipStream_t s);

/*
 * Triangular sweeps of potrs (reference pztrsm.c:113-177 and :43-107).
 * backward == 0: solve L Y = B (forward); backward == 1: L^T X = Y.
 * trsm_diag_left computes S = W_kk B(k,:) (forward) or W_kk^T B(k,:)
 * (backward) into `scratch` (mb * B.nt * nb elements, ld = mb);
 * sweep_update subtracts L(m,k) S from the block rows still to be solved
 * (m > k forward, m < k backward) and copies S into B(k,:).
 */
int plasma_kern_trsm_diag_left(plasma_enum_t uplo, int backward,
                               plasma_desc_t A, plasma_desc_t B, int k,
                               const void *inv, void *scratch,
                               const int *status, hipStream_t s);
int plasma_kern_sweep_update(plasma_enum_t uplo, int backward,
                             plasma_desc_t A, plasma_desc_t B, int k,
                             const void *scratch, const int *status,
                             hipStream_t s);
/* the same, updating only block rows [k + 1, mhi) forward or [mlo, k)
 * backward (band factors: the rows within the band of tile k) */
int plasma_kern_sweep_update_range(plasma_enum_t uplo, int backward,
                                   plasma_desc_t A, plasma_desc_t B, int k,
                                   int mlo, int mhi, const void *scratch,
                                   const int *status, hipStream_t s);

/*
 * Sweeps for repeated solves with one factor.  plasma_kern_scale_panels
 * stores M(m,k) = L(m,k) W_kk for every m > k into M (packed strictly lower
 * tiles, column by column; (mt-1) mt / 2 tiles of mb * nb elements).  With
 * L = M diag(L_kk): forward Z = M^{-1} B is plasma_kern_msweep_update(0, k)
 * for k = 0 .. mt-2; plasma_kern_diag_normal replaces every B(k,:) by
 * W_kk^T W_kk B(k,:) (scratch: mt * mb * nt * nb elements); the backward
 * sweep M^T X = V is plasma_kern_msweep_update(1, k) for k = mt-1 .. 1.
 */
int plasma_kern_scale_panels(plasma_enum_t uplo, plasma_desc_t A,
                             const void *inv, void *M, const int *status,
                             hipStream_t s);
int plasma_kern_msweep_update(int backward, plasma_desc_t A, plasma_desc_t B,
                              int k, const void *M, const int *status,
                              hipStream_t s);
int plasma_kern_diag_normal(plasma_desc_t A, plasma_desc_t B, const void *inv,
                            void *scratch, const int *status, hipStream_t s);

/*
 * Batched SPD solves of uneven order (no reference counterpart).  System i
 * of order n[i] <= 512 is A + aoff[i] (column major, ld n[i], the uplo
 * triangle), its right-hand sides B + boff[i] (n[i] x nrhs, ld n[i]); W
 * holds one slab of plasma_kern_vbatched_wslab(precision) bytes per
 * workgroup (the diagonal-block inverses of the system it is solving).  A
 * persistent grid of `grid` workgroups takes systems from *ticket (zero on
 * entry) in the order `order` (NULL: index order); info[i] gets 0,
 * LAPACK's info, or -3 for an order outside [0, 512].  All pointers are
 * device memory.  plasma_kern_vbatched_lds: LDS bytes per workgroup
 * (dynamic and static), for sizing the grid.
 */
size_t plasma_kern_vbatched_lds(plasma_enum_t precision);
size_t plasma_kern_vbatched_wslab(plasma_enum_t precision);
int plasma_kern_vbatched_posv(plasma_enum_t precision, int upper, int batch,
                              const int *n, void *A, const size_t *aoff,
                              void *W, int nrhs, void *B, const size_t *boff,
                              int *info, const int *order, int *ticket,
                              int grid, hipStream_t s);

/*
 * Super-block sweeps (kernels/ssweep.hip), nb a multiple of 64, S tiles per
 * super-block.  plasma_kern_sbinv: V (ceil(mt / S) slabs of (S nb)^2
 * elements, ld S nb) = the inverses of the diagonal super-blocks of the
 * lower-equivalent factor, from the tile inverses inv; Tscr holds
 * plasma_sbinv_scratch_elems(A, S) elements.  plasma_kern_ssweep: step K of
 * the forward (backward = 0: B <- L^{-1} B) or backward (B <- L^{-T} B)
 * sweep, two launches; Y holds S nb x (B.nt nb) elements.
 */
size_t plasma_sbinv_scratch_elems(plasma_desc_t A, int S);
int plasma_kern_sbinv(plasma_enum_t uplo, plasma_desc_t A, const void *inv,
                      void *V, void *Tscr, int S, const int *status,
                      hipStream_t s);
int plasma_kern_ssweep(plasma_enum_t uplo, int backward, plasma_desc_t A,
                       plasma_desc_t B, const void *V, void *Y, int S, int K,
                       const int *status, hipStream_t s);

/*
 * Scratch of the allocation-free batched solve (kernels/noalloc.hip): one
 * file-scope device array per device, its address looked up at
 * plasma_noalloc_setup (called by plasma_init with the device current);
 * plasma_noalloc_regions hands out its parts.  plasma_kern_vbatched_order
 * writes the systems largest order first into `order`, zeroes the ticket
 * ctrl[0] and sets ctrl[1] when an order is outside [0, 512]; n, order and
 * ctrl are device memory, batch <= PLASMA_VBATCH_NOALLOC_MAX.
 */
typedef struct {
    void *W;                 /* grid_max slabs of wslab_bytes */
    size_t wslab_bytes;
    int grid_max;
    int *ctrl;               /* [0] ticket, [1] invalid-order flag */
    int *n, *order, *info;   /* PLASMA_VBATCH_NOALLOC_MAX entries each */
    size_t *aoff, *boff;
    size_t bytes;            /* the whole array */
} plasma_noalloc_t;
int plasma_noalloc_setup(int device);
int plasma_noalloc_regions(int device, plasma_noalloc_t *r);
int plasma_kern_vbatched_order(const int *n, int batch, int *order,
                               int *ctrl, hipStream_t s);

/*
 * Adjoint kernels (no reference counterpart).  ltl_lower: Q = the symmetric
 * matrix whose lower triangle is tril(L^T Lbar), L the factor (lower-
 * equivalent), Lbar lower triangular in a general n x n descriptor (for
 * uplo == PlasmaUpper the descriptor holds Ubar = Lbar^T).  tile_axpby_t:
 * Z(I,J) = a X(I,J) + b Y(J,I)^T over square general descriptors (X unused
 * when a == 0); with mirror, the symmetric matrix whose lower triangle is
 * a X (elements above the diagonal b Y^T, Y = X, a = b).  syr2k_lapack: C -= scale (U V^T + V U^T) on LAPACK-layout
 * device arrays (U, V: n x k).
 */
int plasma_kern_ltl_lower(plasma_enum_t uplo, plasma_desc_t L,
                          plasma_desc_t Lbar, plasma_desc_t Q,
                          const int *status, hipStream_t s);
int plasma_kern_tile_axpby_t(double a, plasma_desc_t X, double b,
                             plasma_desc_t Y, plasma_desc_t Z, int mirror,
                             const int *status, hipStream_t s);
/* One panel step of the blocked reverse sweep of the Cholesky adjoint
 * (lower storage; L the factor, Lb a general descriptor whose lower tiles
 * hold Lbar): P = Lbar(k+1.., k) W_kk (W = the inverse of tile (k,k),
 * column major, ld mb; P: (mt-k-1) tiles of mb x nb), Lbar(k,k) -= sum_t
 * P(t)^T L(k+1+t, k) (split over up to maxsplit tile ranges, part: maxsplit
 * x nb x nb elements), then Lbar(k+1.., k) = P. */
int plasma_kern_adjoint_panel_step(plasma_desc_t L, plasma_desc_t Lb, int k,
                                   const void *W, void *P, void *part,
                                   int maxsplit, const int *status,
                                   hipStream_t s);
/* Lb(k,k) = W^T Q W for the symmetric Q (column major, ld Lb.mb, order
 * of tile k) and W the inverse of the factor's tile (k,k); T1: mb x mb
 * scratch. */
int plasma_kern_adjoint_diag_sandwich(plasma_desc_t Lb, int k, const void *W,
                                      const void *Q, void *T1,
                                      const int *status, hipStream_t s);
/* plasma_dhposv (kernels/gmres.hip, FP64): symmetric diagonal
 * equilibration s_i = a_ii^{-1/2} (1 where a_ii <= 0), A <- S A S on the
 * stored tiles, B <- S B; and the column-wise vector operations of GMRES
 * with nrhs right-hand sides in lockstep.  A vector set holds the storage
 * pointers of up to PLASMA_VECSET_MAX descriptors with the geometry of the
 * descriptor passed beside it.
 *   coldots:    out[i * nrhs + c] = sum_r V_i(r, c) Z(r, c)
 *   colcombine: Y = beta Y + alpha sum_i coef[i * nrhs + c] V_i (beta = 0
 *               does not read Y; coef on the device) */
#define PLASMA_VECSET_MAX 32
typedef struct {
    const double *p[PLASMA_VECSET_MAX];
    int nv;
} plasma_vecset_t;
int plasma_kern_equil_diag(plasma_enum_t uplo, plasma_desc_t A, double *s,
                           const int *status, hipStream_t st);
int plasma_kern_scale_sym(plasma_enum_t uplo, plasma_desc_t A,
                          const double *s, const int *status, hipStream_t st);
int plasma_kern_scale_rows(plasma_desc_t B, const double *s,
                           const int *status, hipStream_t st);
int plasma_kern_coldots(plasma_desc_t Z, plasma_vecset_t V, double *out,
                        const int *status, hipStream_t st);
int plasma_kern_colcombine(plasma_desc_t Y, double beta, double alpha,
                           plasma_vecset_t V, const double *coef,
                           const int *status, hipStream_t st);
int plasma_kern_syr2k_lapack(plasma_enum_t precision, int n, int k,
                             double scale, const void *U, int ldu,
                             const void *V, int ldv, void *C, int ldc,
                             hipStream_t s);

/*
 * FP32 trailing update on bf16 MFMA (no reference counterpart): the panel
 * columns of step k are split into three bf16 planes per panel tile
 * (plasma_split_bytes(A, pcols) bytes per panel buffer, row stride
 * plasma_split_pstride(pcols)); the update then forms each product from six
 * bf16 products summed in FP32, or with `single` from the leading terms
 * alone (bf16 inputs, FP32 accumulation).  kdepth must be a multiple of 32.
 */
size_t plasma_split_pstride(int pcols);
size_t plasma_split_bytes(plasma_desc_t A, int pcols);
int plasma_kern_split_panel(plasma_desc_t A, int k, int pbase, int pcols,
                            const void *panel, void *planes,
                            const int *status, hipStream_t s);
int plasma_kern_trailing_update_split(plasma_enum_t uplo, plasma_desc_t A,
                                      int pbase, int kdepth, int n_first,
                                      int n_end, const void *panel,
                                      const void *planes, int pcols,
                                      int copy_cols, int max_wgs, int single,
                                      const int *status, hipStream_t s);

/* FP64 trailing update with operands loaded straight into LDS
 * (kernels/update64.hip): same arguments and results as
 * plasma_kern_trailing_update; plasma_update64_ok(A, kdepth) tells whether
 * it applies (FP64, square tiles, mb a multiple of 128, kdepth of 16). */
int plasma_update64_ok(plasma_desc_t A, int kdepth);
int plasma_kern_trailing_update64(plasma_enum_t uplo, plasma_desc_t A,
                                  int pbase, int kdepth, int n_first,
                                  int n_end, const void *panel, int pcols,
                                  int copy_cols, int max_wgs,
                                  const int *status, hipStream_t s);

/* FP32 trailing updates on fp16 MFMA (kernels/split16.hip).  Rows are
 * scaled by sigma_r = 2^-ceil(log2 sqrt(a_rr)) (plasma_kern_split16_scale,
 * plasma_split16_scale_elems(A) floats, from the diagonal of A before the
 * factorization), so every scaled factor entry has magnitude <= 1; each
 * panel element is then stored as np fp16 terms (np = 2: hi + lo, three
 * products per update product; np = 1: one product) in MFMA-fragment
 * order (plasma_split16_bytes(A, pcols, np) bytes per panel buffer).  The
 * update scales C by sigma_r sigma_c on entry and back on exit. */
size_t plasma_split16_bytes(plasma_desc_t A, int pcols, int np);
size_t plasma_split16_scale_elems(plasma_desc_t A);
int plasma_kern_split16_scale(plasma_enum_t uplo, plasma_desc_t A,
                              float *rscale, const int *status,
                              hipStream_t s);
int plasma_kern_split16_panel(plasma_desc_t A, int k, int pbase, int pcols,
                              const void *panel, void *planes, int np,
                              const float *rscale, const int *status,
                              hipStream_t s);
int plasma_kern_trailing_update_split16(
    plasma_enum_t uplo, plasma_desc_t A, int pbase, int kdepth, int n_first,
    int n_end, const void *panel, const void *planes, int pcols,
    int copy_cols, int max_wgs, int np, const float *rscale,
    const int *status, hipStream_t s);

/*
 * Mixed-precision helpers (reference compute/zcposv.c:363-539).
 */
/* FP64 -> FP32 (round to nearest), sets *status = 1 on overflow
 * (reference core_zlag2c.c:48-79); or FP32 -> FP64 (core_clag2z.c). */
int plasma_kern_lag2(plasma_desc_t src, plasma_desc_t dst, int *status,
                     hipStream_t s);
/* B <- A (general, reference pzlacpy.c:118-140) */
int plasma_kern_lacpy(plasma_desc_t A, plasma_desc_t B, const int *status,
                      hipStream_t s);
/* B <- alpha A + beta B (reference pzgeadd.c:41-49) */
int plasma_kern_geadd(double alpha, plasma_desc_t A, double beta,
                      plasma_desc_t B, const int *status, hipStream_t s);
/* ||A||_inf of a symmetric matrix stored in one triangle of a general
 * descriptor (reference pzlanhe.c:83-131: per-tile row and column sums, then
 * a reduction); work: plasma_lansy_work_doubles(A) doubles of device memory;
 * result -> *value (device). */
size_t plasma_lansy_work_doubles(plasma_desc_t A);
/* Norm of a tile matrix (kernels/norms.hip): uplo PlasmaGeneral for any
 * matrix, PlasmaLower / PlasmaUpper for a symmetric one held in that
 * triangle; norm Max / One / Inf / Frobenius; *value (device) written in
 * stream order; work: plasma_norm_work_doubles(A) doubles of device
 * memory.  amax: per column (n values) or per row (mt * mb slots, the
 * first m meaningful) max |a_ij| into device values. */
size_t plasma_norm_work_doubles(plasma_desc_t A);

/* Pieces of the inverse from the factor (kernels/potri.hip): B (general)
 * set to I; the first zero on the diagonal of the lower-equivalent factor
 * A into *status as its 1-based index (if *status is 0); the uplo triangle
 * of the symmetric Q (general) into A's stored tiles. */
int plasma_kern_identity(plasma_desc_t B, const int *status, hipStream_t s);
int plasma_kern_diag_zero(plasma_enum_t uplo, plasma_desc_t A, int *status,
                          hipStream_t s);
int plasma_kern_copy_tri(plasma_enum_t uplo, plasma_desc_t Q,
                         plasma_desc_t A, const int *status, hipStream_t s);
int plasma_kern_norm(plasma_enum_t norm, plasma_enum_t uplo, plasma_desc_t A,
                     double *work, double *value, const int *status,
                     hipStream_t s);
int plasma_kern_amax(plasma_enum_t colrow, plasma_desc_t A, double *values,
                     const int *status, hipStream_t s);
int plasma_kern_lansy_inf(plasma_enum_t uplo, plasma_desc_t A, double *work,
                          double *value, const int *status, hipStream_t s);
/* values[j] = max_i |A(i,j)| (reference pdzamax.c:36-52) */
int plasma_kern_amax_cols(plasma_desc_t A, double *values,
                          const int *status, hipStream_t s);
/* R <- R - A X, A symmetric stored in the uplo triangle of a general
 * descriptor (reference pzhemm.c:44-80 with alpha = -1, beta = 1).  work
 * (device, work_bytes; may be NULL): room for split-K partial sums of R,
 * used when R has few 64 x 64 blocks; the partials are added in a fixed
 * order, so results do not depend on scheduling. */
int plasma_kern_symm_residual(plasma_enum_t uplo, plasma_desc_t A,
                              plasma_desc_t X, plasma_desc_t R, void *work,
                              size_t work_bytes, const int *status,
                              hipStream_t s);
/* *flag = all_j (rnorm[j] <= xnorm[j] * cte), cte = anorm * eps * sqrt(n)
 * (reference zcposv.c:275-287 and :476) */
int plasma_kern_conv_test(const double *rnorm, const double *xnorm, int nrhs,
                          const double *anorm, double eps_sqrtn, int *flag,
                          const int *status, hipStream_t s);

/*
 * Complex precisions (kernels/zchol.hip), PlasmaComplexDouble and
 * PlasmaComplexFloat, lower-equivalent L = U^H for PlasmaUpper.
 *
 * zpotrf_diag: factor diagonal tile k (LAPACK zpotrf semantics; info
 * nb*k + j + 1 into the status word) and write W_kk = L(k,k)^{-1} to
 * inv + k * mb * mb (ld mb, zeros above).  ztrtri_diag: W_kk of every
 * diagonal tile of an already factored A.
 * zpanel: P(m) = L(m,k) W_kk^H, m > k, into slot m-k-1 of panel_buf
 * ((mt-1) slots of mb x nb, ld mb).  zupdate: L(m,n) -= P(m) P(n)^H for
 * the columns n in [n_first, n_end) (n_first > k), n <= m, and with copy
 * the panels of step k into L(m,k).
 * zsweep_diag: S = W_kk B(k,:) (forward) or W_kk^H B(k,:) (backward) into
 * S (mb x B.n, ld mb); zsweep_update: B(m,:) -= L(m,k) S (m > k) or
 * L(k,m)^H S (m < k), and B(k,:) = S.
 * zhemm_residual: R <- R - A X, A Hermitian in its uplo triangle.
 * zlanhe_tiles: the row-sum work array of plasma_kern_lansy_inf;
 * zamax_cols: values[j] = max_i |Re a_ij| + |Im a_ij|.
 */
int plasma_kern_zpotrf_diag(plasma_enum_t uplo, plasma_desc_t A, int k,
                            void *inv, int *status, hipStream_t s);
int plasma_kern_ztrtri_diag(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                            int *status, hipStream_t s);
int plasma_kern_zpanel(plasma_enum_t uplo, plasma_desc_t A, int k,
                       const void *inv, void *panel_buf, const int *status,
                       hipStream_t s);
int plasma_kern_zupdate(plasma_enum_t uplo, plasma_desc_t A, int k,
                        int n_first, int n_end, int copy,
                        const void *panel_buf, const int *status,
                        hipStream_t s);
int plasma_kern_zsweep_diag(int backward, plasma_desc_t B, int k,
                            const void *inv, void *S, const int *status,
                            hipStream_t s);
int plasma_kern_zsweep_update(plasma_enum_t uplo, int backward,
                              plasma_desc_t A, plasma_desc_t B, int k,
                              const void *S, const int *status,
                              hipStream_t s);
int plasma_kern_zhemm_residual(plasma_enum_t uplo, plasma_desc_t A,
                               plasma_desc_t X, plasma_desc_t R,
                               const int *status, hipStream_t s);
int plasma_kern_zlanhe_tiles(plasma_enum_t uplo, plasma_desc_t A,
                             double *work, const int *status, hipStream_t s);
int plasma_kern_zamax_cols(plasma_desc_t A, double *values,
                           const int *status, hipStream_t s);

/* A complex descriptor seen as a real one of twice the rows: element
 * (i, j) of the complex view is the pair (2i, j), (2i+1, j) of the real
 * view, tile by tile (interleaved storage, every offset doubled).  Lets the
 * element-wise kernels (layout, conversion, copy, add) serve complex data;
 * a real descriptor is returned unchanged. */
static inline plasma_desc_t plasma_desc_real_view(plasma_desc_t A)
{
    if (A.precision != PlasmaComplexDouble &&
        A.precision != PlasmaComplexFloat)
        return A;
    A.precision = A.precision == PlasmaComplexDouble ? PlasmaRealDouble
                                                     : PlasmaRealFloat;
    A.mb *= 2;
    A.gm *= 2;
    A.i *= 2;
    A.m *= 2;
    A.A21 *= 2;
    A.A12 *= 2;
    A.A22 *= 2;
    return A;
}

static inline int plasma_is_complex(plasma_enum_t precision)
{
    return precision == PlasmaComplexDouble ||
           precision == PlasmaComplexFloat;
}

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_KERNELS_H */
