/*
 * plasma_oracle.c -- CPU restatement of PLASMA's tile Cholesky solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle and CPU baseline for the
 * MI355X library.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it.
 *
 * Layout: the reference's tile descriptor (include/plasma_descriptor.h:
 * general tiles column by column then the edge regions A21 / A12 / A22,
 * triangular descriptors storing one triangle of tiles).  Algorithms: one
 * OpenMP task per tile operation, depend clauses on the tiles, as in
 * compute/pzpotrf.c, pztrsm.c and zcposv.c.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <complex.h>
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "plasma_oracle.h"

enum { ORA_NOTRANS = 111, ORA_TRANS = 112, ORA_CONJTRANS = 113,
       ORA_LEFT = 141, ORA_RIGHT = 142,
       ORA_NONUNIT = 131, ORA_UNIT = 132 };

static int ora_backend = ORA_BLAS_BUILTIN;
static int mkl_layer_refused = 0;
static int ora_threads = 0;
static int ora_planted = 0;

void ora_plant_mistake(int which) { ora_planted = which; }
void ora_set_threads(int n) { ora_threads = n; }
int ora_get_threads(void) { return ora_threads > 0 ? ora_threads : omp_get_max_threads(); }
static int ora_nthreads(void) { return ora_get_threads(); }

static void ora_fail(volatile int *status, int info)
{
    #pragma omp critical(ora_status)
    {
        if (*status == 0) *status = info;
    }
}

/* ---------------------------------------------------------------------
 * Vendor BLAS / LAPACK through dlopen (the reference links any vendor's
 * CBLAS + LAPACKE; MKL 2021.4 is what this image provides).
 * --------------------------------------------------------------------- */
typedef void (*gemm_d_t)(int, int, int, int, int, int, double, const double *,
                         int, const double *, int, double, double *, int);
typedef void (*trsm_d_t)(int, int, int, int, int, int, int, double,
                         const double *, int, double *, int);
typedef void (*syrk_d_t)(int, int, int, int, int, double, const double *, int,
                         double, double *, int);
typedef void (*symm_d_t)(int, int, int, int, int, double, const double *, int,
                         const double *, int, double, double *, int);
typedef int (*potrf_d_t)(int, char, int, double *, int);
typedef void (*gemm_s_t)(int, int, int, int, int, int, float, const float *,
                         int, const float *, int, float, float *, int);
typedef void (*trsm_s_t)(int, int, int, int, int, int, int, float,
                         const float *, int, float *, int);
typedef void (*syrk_s_t)(int, int, int, int, int, float, const float *, int,
                         float, float *, int);
typedef void (*symm_s_t)(int, int, int, int, int, float, const float *, int,
                         const float *, int, float, float *, int);
typedef int (*potrf_s_t)(int, char, int, float *, int);

static struct {
    gemm_d_t dgemm; trsm_d_t dtrsm; syrk_d_t dsyrk; symm_d_t dsymm;
    potrf_d_t dpotrf;
    gemm_s_t sgemm; trsm_s_t strsm; syrk_s_t ssyrk; symm_s_t ssymm;
    potrf_s_t spotrf;
} mkl;

static int mkl_load(void)
{
    static int tried = 0, ok = 0;
    if (tried) return ok;
    tried = 1;
    const char *names[] = {getenv("ORACLE_MKL"),
                           "/opt/conda/lib/libmkl_rt.so.1",
                           "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so.1",
                           "libmkl_rt.so"};
    void *h = NULL;
    for (int i = 0; i < 5 && h == NULL; i++)
        if (names[i] != NULL) h = dlopen(names[i], RTLD_NOW | RTLD_GLOBAL);
    if (h == NULL) return 0;
    /* Tile kernels must be sequential: parallelism comes from the tasks.
     * MKL keeps the layer it started with once any of its routines has
     * run, and its Intel-OpenMP layer gives wrong results in a process
     * where a GNU libgomp was loaded globally first (torch does this):
     * the layer forks teams through libiomp5's __kmpc_fork_call while its
     * omp_get_thread_num / omp_get_num_threads bind to libgomp's, so every
     * thread computes the same share.  The MKL backend is refused unless
     * the sequential layer is the one in force. */
    int (*set_layer)(int) = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer");
    int layer = set_layer ? set_layer(1 /* MKL_THREADING_SEQUENTIAL */) : -1;
    if (layer != 1) {
        fprintf(stderr, "oracle: MKL threading layer is %d, not sequential "
                "(an MKL routine ran before the oracle selected its layer; "
                "set MKL_THREADING_LAYER=SEQUENTIAL); MKL backend refused\n",
                layer);
        mkl_layer_refused = 1;
        return 0;
    }
    mkl.dgemm = (gemm_d_t)dlsym(h, "cblas_dgemm");
    mkl.dtrsm = (trsm_d_t)dlsym(h, "cblas_dtrsm");
    mkl.dsyrk = (syrk_d_t)dlsym(h, "cblas_dsyrk");
    mkl.dsymm = (symm_d_t)dlsym(h, "cblas_dsymm");
    mkl.dpotrf = (potrf_d_t)dlsym(h, "LAPACKE_dpotrf_work");
    mkl.sgemm = (gemm_s_t)dlsym(h, "cblas_sgemm");
    mkl.strsm = (trsm_s_t)dlsym(h, "cblas_strsm");
    mkl.ssyrk = (syrk_s_t)dlsym(h, "cblas_ssyrk");
    mkl.ssymm = (symm_s_t)dlsym(h, "cblas_ssymm");
    mkl.spotrf = (potrf_s_t)dlsym(h, "LAPACKE_spotrf_work");
    ok = mkl.dgemm && mkl.dtrsm && mkl.dsyrk && mkl.dsymm && mkl.dpotrf &&
         mkl.sgemm && mkl.strsm && mkl.ssyrk && mkl.ssymm && mkl.spotrf;
    return ok;
}

int ora_set_blas(int backend)
{
    if (backend == ORA_BLAS_MKL && mkl_load()) ora_backend = ORA_BLAS_MKL;
    else ora_backend = ORA_BLAS_BUILTIN;
    return ora_backend;
}
int ora_get_blas(void) { return ora_backend; }
/* 1 when MKL was found but refused because its threading layer is not the
 * sequential one */
int ora_mkl_refused(void) { return mkl_layer_refused; }

/* ---------------------------------------------------------------------
 * Tile descriptor (reference include/plasma_descriptor.h:40-246,
 * control/descriptor.c:134-264); views are not needed here (i = j = 0).
 * type: ORA_LOWER / ORA_UPPER triangular, anything else general.
 * --------------------------------------------------------------------- */
typedef struct {
    int type, mb, nb, gm, gn, m, n, mt, nt;
    size_t A21, A12, A22, es;
    void *mat;
} ora_desc_t;

static int ora_desc_create(ora_desc_t *A, int type, size_t es, int nb, int m,
                           int n)
{
    memset(A, 0, sizeof(*A));
    A->type = (type == ORA_LOWER || type == ORA_UPPER) ? type : 0;
    A->mb = A->nb = nb;
    A->gm = A->m = m;
    A->gn = A->n = n;
    A->mt = (m + nb - 1) / nb;
    A->nt = (n + nb - 1) / nb;
    A->es = es;
    size_t lm1 = (size_t)(m / nb), ln1 = (size_t)(n / nb), elems;
    if (A->type) {
        size_t full = (ln1 * (1 + lm1)) / 2;
        A->A21 = A->A12 = (size_t)nb * nb * full;
        A->A22 = (size_t)(m - m % nb) * (n % nb) + A->A12;
        elems = full * nb * nb + (size_t)m * (n % nb);
    }
    else {
        A->A21 = (size_t)(m - m % nb) * (n - n % nb);
        A->A12 = (size_t)(m % nb) * (n - n % nb) + A->A21;
        A->A22 = (size_t)(m - m % nb) * (n % nb) + A->A12;
        elems = (size_t)m * n;
    }
    A->mat = calloc(elems ? elems : 1, es);
    return A->mat == NULL;
}

static void ora_desc_destroy(ora_desc_t *A) { free(A->mat); A->mat = NULL; }

static void *ora_tile_addr(const ora_desc_t *A, int m, int n)
{
    size_t lm1 = (size_t)(A->gm / A->mb), ln1 = (size_t)(A->gn / A->nb), off;
    if ((size_t)m < lm1) {
        if ((size_t)n < ln1) {
            if (A->type == ORA_UPPER)
                off = (size_t)A->mb * A->nb * (m + ((size_t)n * (n + 1)) / 2);
            else if (A->type == ORA_LOWER)
                off = (size_t)A->mb * A->nb *
                      ((m - n) + ((size_t)n * (2 * lm1 - n + 1)) / 2);
            else
                off = (size_t)A->mb * A->nb * (m + lm1 * n);
        }
        else
            off = A->A12 + (size_t)A->mb * (A->gn % A->nb) * m;
    }
    else {
        if ((size_t)n < ln1) off = A->A21 + (size_t)A->nb * (A->gm % A->mb) * n;
        else                 off = A->A22;
    }
    return (char *)A->mat + off * A->es;
}

static int ora_mmain(const ora_desc_t *A, int k)
{
    return (k < A->gm / A->mb) ? A->mb : A->gm % A->mb;
}
static int ora_mview(const ora_desc_t *A, int k)
{
    if (k < A->mt - 1) return A->mb;
    return (A->m % A->mb) ? A->m % A->mb : A->mb;
}
static int ora_nview(const ora_desc_t *A, int k)
{
    if (k < A->nt - 1) return A->nb;
    return (A->n % A->nb) ? A->n % A->nb : A->nb;
}

/* ---------------------------------------------------------------------
 * Band tile storage: bt x nt full nb x nb tiles (ld nb), bt = 1 +
 * ceil(kd / nb) tile rows per band column (reference zpbtrf.c:104-110,
 * descriptor.c:199-211); tile (m, n) of the band at slot (kut - 1 + m - n,
 * n), kut = bt for upper storage, 1 for lower.
 * --------------------------------------------------------------------- */
typedef struct {
    int nb, n, nt, kd, bt, upper;
    size_t es;
    void *mat;
} ora_band_t;

static int ora_band_create(ora_band_t *A, int upper, size_t es, int nb,
                           int n, int kd)
{
    A->nb = nb;
    A->n = n;
    A->nt = (n + nb - 1) / nb;
    A->kd = kd;
    A->bt = 1 + (kd + nb - 1) / nb;
    A->upper = upper;
    A->es = es;
    A->mat = calloc((size_t)A->bt * A->nt * nb * nb, es);
    return A->mat == NULL;
}

static void ora_band_destroy(ora_band_t *A) { free(A->mat); A->mat = NULL; }

static void *ora_band_tile(const ora_band_t *A, int m, int n)
{
    const int slot = A->upper ? (A->bt - 1) + m - n : m - n;
    return (char *)A->mat +
           ((size_t)slot + (size_t)A->bt * n) * A->nb * A->nb * A->es;
}

static int ora_band_view(const ora_band_t *A, int k)
{
    if (k < A->nt - 1) return A->nb;
    return (A->n % A->nb) ? A->n % A->nb : A->nb;
}

static int imin_(int a, int b) { return a < b ? a : b; }
static int imax_(int a, int b) { return a > b ? a : b; }

/* ------------------------- double precision -------------------------- */
#define REAL double
#define FN(name) d_##name
#define CBLAS(name) mkl.d##name
#define MKL(name) mkl.d##name
#include "ora_tile.inc"
#undef REAL
#undef FN
#undef CBLAS
#undef MKL

/* ------------------------- single precision -------------------------- */
#define REAL float
#define FN(name) s_##name
#define CBLAS(name) mkl.s##name
#define MKL(name) mkl.s##name
#include "ora_tile.inc"
#undef REAL
#undef FN
#undef CBLAS
#undef MKL

int ora_dpbtrf(int uplo, int n, int kd, double *AB, int ldab, int nb)
{
    return d_band(uplo, n, kd, -1, AB, ldab, NULL, 1, nb, 1);
}
int ora_spbtrf(int uplo, int n, int kd, float *AB, int ldab, int nb)
{
    return s_band(uplo, n, kd, -1, AB, ldab, NULL, 1, nb, 1);
}
int ora_dpbsv(int uplo, int n, int kd, int nrhs, double *AB, int ldab,
              double *B, int ldb, int nb, int factor)
{
    return d_band(uplo, n, kd, nrhs, AB, ldab, B, ldb, nb, factor);
}
int ora_spbsv(int uplo, int n, int kd, int nrhs, float *AB, int ldab,
              float *B, int ldb, int nb, int factor)
{
    return s_band(uplo, n, kd, nrhs, AB, ldab, B, ldb, nb, factor);
}
int ora_dpotrf(int uplo, int n, double *A, int lda, int nb)
{
    return d_potrf(uplo, n, A, lda, nb);
}
int ora_spotrf(int uplo, int n, float *A, int lda, int nb)
{
    return s_potrf(uplo, n, A, lda, nb);
}
int ora_dposv(int uplo, int n, int nrhs, double *A, int lda, double *B,
              int ldb, int nb)
{
    return d_solve(uplo, n, nrhs, A, lda, B, ldb, nb, 1);
}
int ora_sposv(int uplo, int n, int nrhs, float *A, int lda, float *B,
              int ldb, int nb)
{
    return s_solve(uplo, n, nrhs, A, lda, B, ldb, nb, 1);
}
int ora_dpotri(int uplo, int n, double *A, int lda, int nb)
{
    return d_inverse(uplo, n, A, lda, nb, 0);
}
int ora_dpoinv(int uplo, int n, double *A, int lda, int nb)
{
    return d_inverse(uplo, n, A, lda, nb, 1);
}
int ora_spotri(int uplo, int n, float *A, int lda, int nb)
{
    return s_inverse(uplo, n, A, lda, nb, 0);
}
int ora_dpotrs(int uplo, int n, int nrhs, const double *A, int lda,
               double *B, int ldb, int nb)
{
    return d_solve(uplo, n, nrhs, (double *)A, lda, B, ldb, nb, 0);
}

/* ---------------------------------------------------------------------
 * Tile layout round trip (reference compute/pztr2desc.c, pzdesc2tr.c,
 * pzge2desc.c, pzdesc2ge.c).
 * --------------------------------------------------------------------- */
long ora_tile_storage_elems(int uplo, int m, int n, int nb)
{
    int type = (uplo == ORA_LOWER || uplo == ORA_UPPER) ? uplo : 0;
    size_t lm1 = (size_t)(m / nb), ln1 = (size_t)(n / nb);
    if (type) return (long)(((ln1 * (1 + lm1)) / 2) * nb * nb + (size_t)m * (n % nb));
    return (long)m * n;
}

void ora_lapack2tile(int uplo, int m, int n, const double *pA, int lda,
                     double *tiles, int nb)
{
    ora_desc_t A;
    if (ora_desc_create(&A, uplo, sizeof(double), nb, m, n)) return;
    d_lapack2tile(uplo, pA, lda, A, 1);
    memcpy(tiles, A.mat, (size_t)ora_tile_storage_elems(uplo, m, n, nb) * sizeof(double));
    ora_desc_destroy(&A);
}

void ora_tile2lapack(int uplo, int m, int n, const double *tiles, double *pA,
                     int lda, int nb)
{
    ora_desc_t A;
    if (ora_desc_create(&A, uplo, sizeof(double), nb, m, n)) return;
    memcpy(A.mat, tiles, (size_t)ora_tile_storage_elems(uplo, m, n, nb) * sizeof(double));
    d_lapack2tile(uplo, pA, lda, A, 0);
    ora_desc_destroy(&A);
}

/* ---------------------------------------------------------------------
 * Precision conversion (reference core_zlag2c.c:48-79 -> LAPACK dlag2s:
 * info = 1 when an entry exceeds the FP32 range; core_clag2z.c).
 * --------------------------------------------------------------------- */
/* Tile BLAS-3 on LAPACK arrays (reference compute/zgemm.c, zsyrk.c,
 * ztrsm.c with pzgemm / pzsyrk / pztrsm). */
int ora_dgemm(int ta, int tb, int m, int n, int k, double alpha,
              const double *A, int lda, const double *B, int ldb, double beta,
              double *C, int ldc, int nb)
{
    return d_blas3(0, ta, tb, 0, 0, m, n, k, alpha, A, lda, B, ldb, beta, C,
                   ldc, nb);
}
int ora_dsyrk(int uplo, int trans, int n, int k, double alpha,
              const double *A, int lda, double beta, double *C, int ldc,
              int nb)
{
    return d_blas3(1, uplo, trans, 0, 0, n, n, k, alpha, A, lda, NULL, 1,
                   beta, C, ldc, nb);
}
int ora_dtrsm(int side, int uplo, int trans, int diag, int m, int n,
              double alpha, const double *A, int lda, double *B, int ldb,
              int nb)
{
    return d_blas3(2, side, uplo, trans, diag, m, n, 0, alpha, A, lda, NULL,
                   1, 0, B, ldb, nb);
}
int ora_sgemm(int ta, int tb, int m, int n, int k, float alpha,
              const float *A, int lda, const float *B, int ldb, float beta,
              float *C, int ldc, int nb)
{
    return s_blas3(0, ta, tb, 0, 0, m, n, k, alpha, A, lda, B, ldb, beta, C,
                   ldc, nb);
}
int ora_strsm(int side, int uplo, int trans, int diag, int m, int n,
              float alpha, const float *A, int lda, float *B, int ldb, int nb)
{
    return s_blas3(2, side, uplo, trans, diag, m, n, 0, alpha, A, lda, NULL,
                   1, 0, B, ldb, nb);
}

int ora_dlag2s(int m, int n, const double *A, int lda, float *As, int ldas)
{
    double rmax = (double)FLT_MAX;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double v = A[(size_t)j * lda + i];
            if (v < -rmax || v > rmax) return 1;
            As[(size_t)j * ldas + i] = (float)v;
        }
    return 0;
}

void ora_slag2d(int m, int n, const float *As, int ldas, double *A, int lda)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++)
            A[(size_t)j * lda + i] = (double)As[(size_t)j * ldas + i];
}

/* ---------------------------------------------------------------------
 * Mixed-precision iterative refinement (reference compute/zcposv.c:363-539
 * generated as dsposv).  Tile storage: A, B, X, R general FP64; As, Xs
 * general FP32.  Steps and constants as in the reference: itermax = 30,
 * eps = dlamch('E') = 2^-53, cte = ||A||_inf eps sqrt(n), convergence when
 * every column has max|r| <= max|x| * cte.  A failure of the FP32 factor or
 * an FP64 -> FP32 overflow fails the whole sequence (no FP64 fallback), as
 * in the reference.
 * --------------------------------------------------------------------- */

/* ||A||_inf of the symmetric matrix held in the uplo triangle
 * (reference compute/pzlanhe.c:83-131). */
static double d_lansy_inf(int uplo, int n, const double *A, int lda)
{
    double best = 0;
    for (int i = 0; i < n; i++) {
        double s = 0;
        for (int j = 0; j < n; j++) {
            int r = i, c = j;
            if ((uplo == ORA_LOWER && r < c) || (uplo == ORA_UPPER && r > c)) {
                r = j;
                c = i;
            }
            s += fabs(A[(size_t)c * lda + r]);
        }
        if (s > best || isnan(s)) best = s;
    }
    return best;
}

/* column max |x| (reference compute/pdzamax.c:36-52, core_dzamax.c) */
static void d_amax_cols(int m, int n, const double *X, int ldx, double *v)
{
    for (int j = 0; j < n; j++) {
        double t = fabs(X[(size_t)j * ldx]);
        for (int i = 1; i < m; i++) {
            double a = fabs(X[(size_t)j * ldx + i]);
            if (a > t) t = a;
        }
        v[j] = t;
    }
}

/* ---------------------------------------------------------------------
 * Norms and column / row max-norms, the public plasma_dlansy /
 * plasma_dlange / plasma_damax (reference compute/zlansy.c:81,
 * compute/zlange.c:76, compute/dzamax.c; tile algorithms pzlansy.c,
 * pzlange.c, pdzamax.c).  The reference forms per-tile partial results
 * and combines them; the values here are the same quantities summed in
 * LAPACK order: Max = max |a_ij|, One = max column sum of |a_ij|, Inf = max
 * row sum, Frobenius = sqrt(sum a_ij^2) by the scaled sum of squares of
 * LAPACK's dlassq (no overflow for large entries).  NaN propagates as in
 * LAPACK's disnan tests.
 * --------------------------------------------------------------------- */
enum { ORA_ONE = 171, ORA_FRO = 174, ORA_INF = 175, ORA_MAX = 177,
       ORA_COLWISE = 401, ORA_ROWWISE = 402 };

static void ssq_add(double x, double *scale, double *ssq)
{
    double a = fabs(x);
    if (a == 0.0) return;
    if (isnan(a)) { *scale = a; *ssq = 1.0; return; }
    if (*scale < a) {
        *ssq = 1.0 + *ssq * (*scale / a) * (*scale / a);
        *scale = a;
    }
    else {
        *ssq += (a / *scale) * (a / *scale);
    }
}

static double ge_elem(int uplo, const double *A, int lda, int i, int j)
{
    /* element (i, j) of a symmetric matrix stored in the uplo triangle, or
     * of a general matrix (uplo = 0) */
    if ((uplo == ORA_LOWER && i < j) || (uplo == ORA_UPPER && i > j)) {
        int t = i; i = j; j = t;
    }
    return A[(size_t)j * lda + i];
}

static double norm_any(int norm, int uplo, int m, int n, const double *A,
                       int lda)
{
    if (m == 0 || n == 0) return 0.0;
    double value = 0.0;
    if (norm == ORA_MAX) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                double a = fabs(ge_elem(uplo, A, lda, i, j));
                if (a > value || isnan(a)) value = a;
            }
    }
    else if (norm == ORA_ONE) {
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int i = 0; i < m; i++) s += fabs(ge_elem(uplo, A, lda, i, j));
            if (s > value || isnan(s)) value = s;
        }
    }
    else if (norm == ORA_INF) {
        for (int i = 0; i < m; i++) {
            double s = 0;
            for (int j = 0; j < n; j++) s += fabs(ge_elem(uplo, A, lda, i, j));
            if (s > value || isnan(s)) value = s;
        }
    }
    else if (norm == ORA_FRO) {
        double scale = 0.0, ssq = 1.0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++)
                ssq_add(ge_elem(uplo, A, lda, i, j), &scale, &ssq);
        value = scale * sqrt(ssq);
    }
    else {
        value = -1.0;
    }
    return value;
}

double ora_dlansy(int norm, int uplo, int n, const double *A, int lda)
{
    return norm_any(norm, uplo, n, n, A, lda);
}

double ora_dlange(int norm, int m, int n, const double *A, int lda)
{
    return norm_any(norm, 0, m, n, A, lda);
}

/* values[j] = max_i |a_ij| (Columnwise) or values[i] = max_j |a_ij|
 * (Rowwise); reference compute/pdzamax.c with core_dzamax.c, where a NaN
 * entry does not replace a larger value (plain > comparison). */
void ora_damax(int colrow, int m, int n, const double *A, int lda,
               double *values)
{
    if (colrow == ORA_COLWISE) {
        for (int j = 0; j < n; j++) {
            double t = 0.0;
            for (int i = 0; i < m; i++) {
                double a = fabs(A[(size_t)j * lda + i]);
                if (a > t) t = a;
            }
            values[j] = t;
        }
    }
    else {
        for (int i = 0; i < m; i++) values[i] = 0.0;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                double a = fabs(A[(size_t)j * lda + i]);
                if (a > values[i]) values[i] = a;
            }
    }
}

/* reference zcposv.c:275-287 */
static int conv(const double *rn, const double *xn, int n, double cte)
{
    for (int i = 0; i < n; i++)
        if (rn[i] > xn[i] * cte) return 0;
    return 1;
}

/* Solve As Xs = Rs in FP32 on tiles: two sweeps with the FP32 factor. */
static int s_sweeps(int uplo, ora_desc_t As, ora_desc_t Xs)
{
    volatile int status = 0;
    #pragma omp parallel num_threads(ora_nthreads())
    #pragma omp master
    {
        s_pxtrsm_left(uplo, uplo == ORA_UPPER ? ORA_TRANS : ORA_NOTRANS, As, Xs,
                      &status);
        s_pxtrsm_left(uplo, uplo == ORA_UPPER ? ORA_NOTRANS : ORA_TRANS, As, Xs,
                      &status);
    }
    return status;
}

int ora_dsposv(int uplo, int n, int nrhs, const double *pA, int lda,
               const double *pB, int ldb, double *pX, int ldx, int *iter,
               int nb)
{
    const int itermax = 30;
    *iter = 0;
    if (n == 0 || nrhs == 0) return 0;

    /* FP64 pieces (norms, residual, update) run on LAPACK-layout copies;
     * the FP32 factor and sweeps run on tiles.  Results equal the
     * reference's tile computation up to the summation order of the
     * residual product. */
    double *A = malloc((size_t)n * n * sizeof(double));
    double *B = malloc((size_t)n * nrhs * sizeof(double));
    double *X = malloc((size_t)n * nrhs * sizeof(double));
    double *R = malloc((size_t)n * nrhs * sizeof(double));
    float *As = malloc((size_t)n * n * sizeof(float));
    float *Xs = malloc((size_t)n * nrhs * sizeof(float));
    double *xn = malloc(nrhs * sizeof(double)), *rn = malloc(nrhs * sizeof(double));
    int info = 0;
    for (int j = 0; j < n; j++)
        memcpy(A + (size_t)j * n, pA + (size_t)j * lda, n * sizeof(double));
    for (int j = 0; j < nrhs; j++)
        memcpy(B + (size_t)j * n, pB + (size_t)j * ldb, n * sizeof(double));

    const double eps = ldexp(1.0, -53);
    const double anorm = d_lansy_inf(uplo, n, A, n);
    ora_desc_t tAs, tXs;
    ora_desc_create(&tAs, 0, sizeof(float), nb, n, n);
    ora_desc_create(&tXs, 0, sizeof(float), nb, n, nrhs);

    /* Xs = lag2s(B), As = lag2s(A) (reference :445-449) */
    if (ora_dlag2s(n, nrhs, B, n, Xs, n) || ora_dlag2s(n, n, A, n, As, n)) {
        info = 1;
        goto done;
    }
    /* FP32 tile factorization and solve (reference :452-460) */
    {
        volatile int status = 0;
        s_lapack2tile(0, As, n, tAs, 1);
        #pragma omp parallel num_threads(ora_nthreads())
        #pragma omp master
        {
            s_pxpotrf(uplo, tAs, &status);
        }
        if (status) { info = status; goto done; }
    }
    s_lapack2tile(0, Xs, n, tXs, 1);
    if ((info = s_sweeps(uplo, tAs, tXs))) goto done;
    s_lapack2tile(0, Xs, n, tXs, 0);
    ora_slag2d(n, nrhs, Xs, n, X, n);

    /* R = B - A X; norms; convergence (reference :466-481) */
    memcpy(R, B, (size_t)n * nrhs * sizeof(double));
    d_k_symm(uplo, n, nrhs, -1.0, A, n, X, n, 1.0, R, n);
    d_amax_cols(n, nrhs, X, n, xn);
    d_amax_cols(n, nrhs, R, n, rn);
    const double cte = anorm * eps * sqrt((double)n);
    if (conv(rn, xn, nrhs, cte)) {
        *iter = 0;
        goto done;
    }
    for (int it = 0; it < itermax; it++) {
        /* reference :487-516 */
        if (ora_dlag2s(n, nrhs, R, n, Xs, n)) { info = 1; goto done; }
        s_lapack2tile(0, Xs, n, tXs, 1);
        if ((info = s_sweeps(uplo, tAs, tXs))) goto done;
        s_lapack2tile(0, Xs, n, tXs, 0);
        ora_slag2d(n, nrhs, Xs, n, R, n);
        for (size_t e = 0; e < (size_t)n * nrhs; e++) X[e] += R[e];
        memcpy(R, B, (size_t)n * nrhs * sizeof(double));
        d_k_symm(uplo, n, nrhs, -1.0, A, n, X, n, 1.0, R, n);
        d_amax_cols(n, nrhs, X, n, xn);
        d_amax_cols(n, nrhs, R, n, rn);
        if (conv(rn, xn, nrhs, cte)) {
            *iter = it + 1;
            goto done;
        }
    }
    /* FP64 fallback (reference :524-538) */
    *iter = -itermax - 1;
    memcpy(X, B, (size_t)n * nrhs * sizeof(double));
    info = d_solve(uplo, n, nrhs, A, n, X, n, nb, 1);

done:
    if (info == 0)
        for (int j = 0; j < nrhs; j++)
            memcpy(pX + (size_t)j * ldx, X + (size_t)j * n, n * sizeof(double));
    ora_desc_destroy(&tAs);
    ora_desc_destroy(&tXs);
    free(A); free(B); free(X); free(R); free(As); free(Xs); free(xn); free(rn);
    return info;
}

/* ---------------------------------------------------------------------
 * Cholesky adjoint (no reference counterpart; the MI355X library's
 * plasma_dpotrf_adjoint).  Reverse-mode sweep of the scalar column
 * Cholesky
 *     L(j,j) = sqrt(A(j,j) - sum_{k<j} L(j,k)^2)
 *     L(i,j) = (A(i,j) - sum_{k<j} L(i,k) L(j,k)) / L(j,j),   i > j
 * in exact reverse order, accumulating Lbar in place (I. Murray,
 * "Differentiation of the Cholesky decomposition", 2016, the unblocked
 * reverse sweep).  This gives the gradient with respect to the lower
 * triangle of A; the symmetric gradient written out (full n x n, both
 * triangles) halves its off-diagonal entries.  Input: the factor in the
 * uplo triangle of L (U = L^T for upper) and Lbar in the same triangle
 * (Ubar = Lbar^T for upper); the other triangles are not read.  Plain
 * loops in double, O(n^3): small orders only.
 * --------------------------------------------------------------------- */
int ora_dpotrf_adjoint(int uplo, int n, const double *pL, int ldl,
                       double *pLbar, int ldlb)
{
    if (n <= 0) return 0;
    double *L = malloc((size_t)n * n * sizeof(double));
    double *G = malloc((size_t)n * n * sizeof(double));   /* Lbar, lower */
    double *Ab = calloc((size_t)n * n, sizeof(double));
    if (!L || !G || !Ab) { free(L); free(G); free(Ab); return -1; }
    const int up = (uplo == ORA_UPPER);
    /* lower-equivalent copies, column-major, (i, j) at i + j n */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            int keep = i >= j;
            size_t src = up ? (size_t)j + (size_t)i * ldl
                            : (size_t)i + (size_t)j * ldl;
            size_t srcb = up ? (size_t)j + (size_t)i * ldlb
                             : (size_t)i + (size_t)j * ldlb;
            L[i + (size_t)j * n] = keep ? pL[src] : 0.0;
            G[i + (size_t)j * n] = keep ? pLbar[srcb] : 0.0;
        }
#define L_(i, j) L[(i) + (size_t)(j) * n]
#define G_(i, j) G[(i) + (size_t)(j) * n]
#define A_(i, j) Ab[(i) + (size_t)(j) * n]
    for (int j = n - 1; j >= 0; j--) {
        const double djj = L_(j, j);
        for (int i = n - 1; i > j; i--) {
            /* L(i,j) = t / L(j,j) */
            const double tb = G_(i, j) / djj;
            G_(j, j) -= G_(i, j) * L_(i, j) / djj;
            A_(i, j) += tb;
            for (int k = 0; k < j; k++) {
                G_(i, k) -= tb * L_(j, k);
                G_(j, k) -= tb * L_(i, k);
            }
        }
        /* L(j,j) = sqrt(s) */
        const double sb = G_(j, j) / (2.0 * djj);
        A_(j, j) += sb;
        for (int k = 0; k < j; k++) G_(j, k) -= 2.0 * sb * L_(j, k);
    }
    /* symmetric gradient: off-diagonal entries split over the pair */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            double v = (i == j) ? A_(i, i)
                     : 0.5 * (i > j ? A_(i, j) : A_(j, i));
            pLbar[i + (size_t)j * ldlb] = v;
        }
#undef L_
#undef G_
#undef A_
    free(L);
    free(G);
    free(Ab);
    return 0;
}

/* ---------------------------------------------------------------------
 * Complex precisions (reference compute/zpotrf.c, zpotrs.c, zposv.c,
 * zcposv.c, zlag2c.c, clag2z.c): the tile algorithms of ora_ztile.inc on
 * double complex (z) and float complex (c) data.
 * --------------------------------------------------------------------- */
#define CPLX double complex
#define RPART double
#define ZF(name) z_##name
#define CONJ conj
#define CREAL creal
#define CIMAG cimag
#define CABS cabs
#include "ora_ztile.inc"
#undef CPLX
#undef RPART
#undef ZF
#undef CONJ
#undef CREAL
#undef CIMAG
#undef CABS

#define CPLX float complex
#define RPART float
#define ZF(name) c_##name
#define CONJ conjf
#define CREAL crealf
#define CIMAG cimagf
#define CABS cabsf
#include "ora_ztile.inc"
#undef CPLX
#undef RPART
#undef ZF
#undef CONJ
#undef CREAL
#undef CIMAG
#undef CABS

int ora_zpotrf(int uplo, int n, void *A, int lda, int nb)
{
    return z_solve(uplo, n, 0, A, lda, NULL, 1, nb, 1);
}
int ora_cpotrf(int uplo, int n, void *A, int lda, int nb)
{
    return c_solve(uplo, n, 0, A, lda, NULL, 1, nb, 1);
}
int ora_zposv(int uplo, int n, int nrhs, void *A, int lda, void *B, int ldb,
              int nb)
{
    return z_solve(uplo, n, nrhs, A, lda, B, ldb, nb, 1);
}
int ora_cposv(int uplo, int n, int nrhs, void *A, int lda, void *B, int ldb,
              int nb)
{
    return c_solve(uplo, n, nrhs, A, lda, B, ldb, nb, 1);
}
/* potrs: A holds the factor (only read) */
int ora_zpotrs(int uplo, int n, int nrhs, const void *L, int lda, void *B,
               int ldb, int nb)
{
    return z_solve(uplo, n, nrhs, (double complex *)L, lda, B, ldb, nb, 0);
}
int ora_cpotrs(int uplo, int n, int nrhs, const void *L, int lda, void *B,
               int ldb, int nb)
{
    return c_solve(uplo, n, nrhs, (float complex *)L, lda, B, ldb, nb, 0);
}

/* LAPACK zlag2c: 1 when the real or imaginary part of an entry is outside
 * [-FLT_MAX, FLT_MAX] (conversion stops there) */
int ora_zlag2c(int m, int n, const void *vA, int lda, void *vAs, int ldas)
{
    const double complex *A = vA;
    float complex *As = vAs;
    const double rmax = (double)FLT_MAX;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            const double complex v = A[(size_t)j * lda + i];
            const double re = creal(v), im = cimag(v);
            if (re < -rmax || re > rmax || im < -rmax || im > rmax) return 1;
            As[(size_t)j * ldas + i] = (float)re + (float)im * I;
        }
    return 0;
}

void ora_clag2z(int m, int n, const void *vAs, int ldas, void *vA, int lda)
{
    const float complex *As = vAs;
    double complex *A = vA;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            const float complex v = As[(size_t)j * ldas + i];
            A[(size_t)j * lda + i] = (double)crealf(v) + (double)cimagf(v) * I;
        }
}

double ora_zlanhe_inf(int uplo, int n, const void *A, int lda)
{
    return z_lanhe_inf(uplo, n, A, lda);
}

void ora_dzamax(int m, int n, const void *A, int lda, double *values)
{
    z_amax_cols(m, n, A, lda, values);
}

/* reference compute/zcposv.c:363-539, as ora_dsposv with the complex
 * pieces: zlanhe, zlag2c / clag2z, the single complex tile factor and
 * sweeps, the double complex Hermitian residual and dzamax */
int ora_zcposv(int uplo, int n, int nrhs, const void *vA, int lda,
               const void *vB, int ldb, void *vX, int ldx, int *iter, int nb)
{
    typedef double complex zc_t;
    typedef float complex cc_t;
    const zc_t *pA = vA, *pB = vB;
    zc_t *pX = vX;
    const int itermax = 30;
    *iter = 0;
    if (n == 0 || nrhs == 0) return 0;
    zc_t *A = malloc((size_t)n * n * sizeof(zc_t));
    zc_t *B = malloc((size_t)n * nrhs * sizeof(zc_t));
    zc_t *X = malloc((size_t)n * nrhs * sizeof(zc_t));
    zc_t *R = malloc((size_t)n * nrhs * sizeof(zc_t));
    cc_t *As = malloc((size_t)n * n * sizeof(cc_t));
    cc_t *Xs = malloc((size_t)n * nrhs * sizeof(cc_t));
    double *xn = malloc(nrhs * sizeof(double)), *rn = malloc(nrhs * sizeof(double));
    int info = 0;
    for (int j = 0; j < n; j++)
        memcpy(A + (size_t)j * n, pA + (size_t)j * lda, n * sizeof(zc_t));
    for (int j = 0; j < nrhs; j++)
        memcpy(B + (size_t)j * n, pB + (size_t)j * ldb, n * sizeof(zc_t));

    const double eps = ldexp(1.0, -53);
    const double anorm = z_lanhe_inf(uplo, n, A, n);
    ora_desc_t tAs, tXs;
    ora_desc_create(&tAs, 0, sizeof(cc_t), nb, n, n);
    ora_desc_create(&tXs, 0, sizeof(cc_t), nb, n, nrhs);

    if (ora_zlag2c(n, nrhs, B, n, Xs, n) || ora_zlag2c(n, n, A, n, As, n)) {
        info = 1;
        goto done;
    }
    {
        volatile int status = 0;
        c_lapack2tile(0, As, n, tAs, 1);
        #pragma omp parallel num_threads(ora_nthreads())
        #pragma omp master
        {
            c_pxpotrf(uplo, tAs, &status);
        }
        if (status) { info = status; goto done; }
    }
    c_lapack2tile(0, Xs, n, tXs, 1);
    if ((info = c_sweeps(uplo, tAs, tXs))) goto done;
    c_lapack2tile(0, Xs, n, tXs, 0);
    ora_clag2z(n, nrhs, Xs, n, X, n);

    memcpy(R, B, (size_t)n * nrhs * sizeof(zc_t));
    z_hemm_res(uplo, n, nrhs, A, n, X, n, R, n);
    z_amax_cols(n, nrhs, X, n, xn);
    z_amax_cols(n, nrhs, R, n, rn);
    const double cte = anorm * eps * sqrt((double)n);
    if (conv(rn, xn, nrhs, cte)) {
        *iter = 0;
        goto done;
    }
    for (int it = 0; it < itermax; it++) {
        if (ora_zlag2c(n, nrhs, R, n, Xs, n)) { info = 1; goto done; }
        c_lapack2tile(0, Xs, n, tXs, 1);
        if ((info = c_sweeps(uplo, tAs, tXs))) goto done;
        c_lapack2tile(0, Xs, n, tXs, 0);
        ora_clag2z(n, nrhs, Xs, n, R, n);
        for (size_t e = 0; e < (size_t)n * nrhs; e++) X[e] += R[e];
        memcpy(R, B, (size_t)n * nrhs * sizeof(zc_t));
        z_hemm_res(uplo, n, nrhs, A, n, X, n, R, n);
        z_amax_cols(n, nrhs, X, n, xn);
        z_amax_cols(n, nrhs, R, n, rn);
        if (conv(rn, xn, nrhs, cte)) {
            *iter = it + 1;
            goto done;
        }
    }
    *iter = -itermax - 1;
    memcpy(X, B, (size_t)n * nrhs * sizeof(zc_t));
    info = z_solve(uplo, n, nrhs, A, n, X, n, nb, 1);

done:
    if (info == 0)
        for (int j = 0; j < nrhs; j++)
            memcpy(pX + (size_t)j * ldx, X + (size_t)j * n, n * sizeof(zc_t));
    ora_desc_destroy(&tAs);
    ora_desc_destroy(&tXs);
    free(A); free(B); free(X); free(R); free(As); free(Xs); free(xn); free(rn);
    return info;
}
