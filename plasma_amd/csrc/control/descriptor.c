/*
 * descriptor.c -- tile descriptor construction, validation and views.
 *
 * Field values, validation rules and error codes follow the reference
 * (control/descriptor.c:16-397).  Storage created by *_create lives in
 * device memory from the workspace pool; its size is the reference's:
 * gm * gn elements for a general descriptor, and for a triangular one the
 * (ln1 * (1 + lm1)) / 2 full tiles plus lm * (ln % nb) edge elements.
 */
#include "internal.h"

size_t plasma_desc_storage_bytes(plasma_desc_t A)
{
    size_t es = plasma_element_size(A.precision);
    if (A.type == PlasmaUpper || A.type == PlasmaLower) {
        size_t full_m = (size_t)(A.gm / A.mb), full_n = (size_t)(A.gn / A.nb);
        size_t ntiles = (full_n * (1 + full_m)) / 2;
        return (ntiles * A.mb * A.nb + (size_t)A.gm * (A.gn % A.nb)) * es;
    }
    return (size_t)A.gm * A.gn * es;
}

static int desc_allocate(plasma_desc_t *A)
{
    size_t bytes = plasma_desc_storage_bytes(*A);
    A->matrix = plasma_dev_alloc(bytes);
    if (A->matrix == NULL) {
        plasma_error("device allocation failed");
        return PlasmaErrorOutOfMemory;
    }
    return PlasmaSuccess;
}

int plasma_desc_general_create(plasma_enum_t precision, int mb, int nb,
                               int lm, int ln, int i, int j, int m, int n,
                               plasma_desc_t *A)
{
    if (plasma_context_self() == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    int retval = plasma_desc_general_init(precision, NULL, mb, nb, lm, ln,
                                          i, j, m, n, A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_init() failed");
        return retval;
    }
    if (plasma_desc_check(*A) != PlasmaSuccess) {
        plasma_error("plasma_desc_check() failed");
        return PlasmaErrorIllegalValue;
    }
    return desc_allocate(A);
}

int plasma_desc_triangular_create(plasma_enum_t precision, plasma_enum_t uplo,
                                  int mb, int nb, int lm, int ln,
                                  int i, int j, int m, int n,
                                  plasma_desc_t *A)
{
    if (plasma_context_self() == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    int retval = plasma_desc_triangular_init(precision, uplo, NULL, mb, nb,
                                             lm, ln, i, j, m, n, A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_triangular_init() failed");
        return retval;
    }
    if (plasma_desc_check(*A) != PlasmaSuccess) {
        plasma_error("plasma_desc_check() failed");
        return PlasmaErrorIllegalValue;
    }
    return desc_allocate(A);
}

int plasma_desc_destroy(plasma_desc_t *A)
{
    if (plasma_context_self() == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    plasma_dev_free(A->matrix);
    A->matrix = NULL;
    return PlasmaSuccess;
}

static void desc_fill_common(plasma_desc_t *A, int mb, int nb, int lm, int ln,
                             int i, int j, int m, int n)
{
    A->mb = mb;
    A->nb = nb;
    A->gm = lm;
    A->gn = ln;
    A->gmt = (lm % mb == 0) ? lm / mb : lm / mb + 1;
    A->gnt = (ln % nb == 0) ? ln / nb : ln / nb + 1;
    A->i = i;
    A->j = j;
    A->m = m;
    A->n = n;
    A->mt = (m == 0) ? 0 : (i + m - 1) / mb - i / mb + 1;
    A->nt = (n == 0) ? 0 : (j + n - 1) / nb - j / nb + 1;
    A->kl = m - 1;
    A->ku = n - 1;
    A->klt = A->mt;
    A->kut = A->nt;
}

int plasma_desc_general_init(plasma_enum_t precision, void *matrix,
                             int mb, int nb, int lm, int ln, int i, int j,
                             int m, int n, plasma_desc_t *A)
{
    if (mb <= 0 || nb <= 0)
        return PlasmaErrorIllegalValue;
    A->type = PlasmaGeneral;
    A->uplo = PlasmaGeneral;
    A->precision = precision;
    A->matrix = matrix;
    /* edge regions follow the full tiles (reference descriptor.c:141-143) */
    A->A21 = (size_t)(lm - lm % mb) * (ln - ln % nb);
    A->A12 = (size_t)(lm % mb) * (ln - ln % nb) + A->A21;
    A->A22 = (size_t)(lm - lm % mb) * (ln % nb) + A->A12;
    desc_fill_common(A, mb, nb, lm, ln, i, j, m, n);
    return PlasmaSuccess;
}

int plasma_desc_triangular_init(plasma_enum_t precision, plasma_enum_t uplo,
                                void *matrix, int mb, int nb, int lm, int ln,
                                int i, int j, int m, int n, plasma_desc_t *A)
{
    if (mb <= 0 || nb <= 0)
        return PlasmaErrorIllegalValue;
    if (lm != ln)
        plasma_error("invalid lm or ln");
    A->type = uplo;
    A->uplo = uplo;
    A->precision = precision;
    A->matrix = matrix;
    size_t full_m = (size_t)(lm / mb), full_n = (size_t)(ln / nb);
    size_t ntiles = (full_n * (1 + full_m)) / 2;
    /* reference descriptor.c:232-235 */
    A->A21 = (size_t)mb * nb * ntiles;
    A->A12 = (size_t)mb * nb * ntiles;
    A->A22 = (size_t)(lm - lm % mb) * (ln % nb) + A->A12;
    desc_fill_common(A, mb, nb, lm, ln, i, j, m, n);
    return PlasmaSuccess;
}

int plasma_desc_general_check(plasma_desc_t A)
{
    if (A.precision != PlasmaRealFloat && A.precision != PlasmaRealDouble &&
        A.precision != PlasmaComplexFloat &&
        A.precision != PlasmaComplexDouble) {
        plasma_error("invalid matrix type");
        return PlasmaErrorIllegalValue;
    }
    if (A.mb <= 0 || A.nb <= 0) {
        plasma_error("negative tile dimension");
        return PlasmaErrorIllegalValue;
    }
    if (A.m < 0 || A.n < 0) {
        plasma_error("negative matrix dimension");
        return PlasmaErrorIllegalValue;
    }
    if (A.gm < A.m || A.gn < A.n) {
        plasma_error("invalid leading dimensions");
        return PlasmaErrorIllegalValue;
    }
    if ((A.i > 0 && A.i >= A.gm) || (A.j > 0 && A.j >= A.gn)) {
        plasma_error("beginning of the matrix out of bounds");
        return PlasmaErrorIllegalValue;
    }
    if (A.i + A.m > A.gm || A.j + A.n > A.gn) {
        plasma_error("submatrix out of bounds");
        return PlasmaErrorIllegalValue;
    }
    if (A.i % A.mb != 0 || A.j % A.nb != 0) {
        plasma_error("submatrix not aligned to a tile");
        return PlasmaErrorIllegalValue;
    }
    return PlasmaSuccess;
}

/* The band parameters as the reference derives them: the band's tile rows
 * below (klt) and above (kut) the diagonal, 1 for the diagonal itself. */
int plasma_desc_general_band_init(plasma_enum_t precision,
                                  plasma_enum_t uplo, void *matrix, int mb,
                                  int nb, int lm, int ln, int i, int j,
                                  int m, int n, int kl, int ku,
                                  plasma_desc_t *A)
{
    int retval = plasma_desc_general_init(precision, matrix, mb, nb, lm, ln,
                                          i, j, m, n, A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_init() failed");
        return retval;
    }
    A->type = PlasmaGeneralBand;
    A->uplo = uplo;
    A->kl = kl;
    A->ku = ku;
    if (uplo == PlasmaGeneral) {
        A->klt = 1 + (i + kl + mb - 1) / mb - i / mb;
        A->kut = 1 + (i + ku + kl + nb - 1) / nb - i / nb;
    }
    else if (uplo == PlasmaUpper) {
        A->klt = 1;
        A->kut = 1 + (i + ku + nb - 1) / nb - i / nb;
    }
    else {
        A->klt = 1 + (i + kl + mb - 1) / mb - i / mb;
        A->kut = 1;
    }
    return PlasmaSuccess;
}

int plasma_desc_general_band_create(plasma_enum_t precision,
                                    plasma_enum_t uplo, int mb, int nb,
                                    int lm, int ln, int i, int j, int m,
                                    int n, int kl, int ku, plasma_desc_t *A)
{
    if (plasma_context_self() == NULL) {
        plasma_error("PLASMA not initialized");
        return PlasmaErrorNotInitialized;
    }
    int retval = plasma_desc_general_band_init(precision, uplo, NULL, mb, nb,
                                               lm, ln, i, j, m, n, kl, ku, A);
    if (retval != PlasmaSuccess) {
        plasma_error("plasma_desc_general_band_init() failed");
        return retval;
    }
    if (plasma_desc_check(*A) != PlasmaSuccess) {
        plasma_error("plasma_desc_check() failed");
        return PlasmaErrorIllegalValue;
    }
    return desc_allocate(A);
}

/* Checks of a band descriptor, with the reference's messages and order. */
int plasma_desc_general_band_check(plasma_desc_t A)
{
    if (A.precision != PlasmaRealFloat && A.precision != PlasmaRealDouble &&
        A.precision != PlasmaComplexFloat &&
        A.precision != PlasmaComplexDouble) {
        plasma_error("invalid matrix type");
        return PlasmaErrorIllegalValue;
    }
    if (A.mb <= 0 || A.nb <= 0) {
        plasma_error("negative tile dimension");
        return PlasmaErrorIllegalValue;
    }
    if (A.m < 0 || A.n < 0) {
        plasma_error("negative matrix dimension");
        return PlasmaErrorIllegalValue;
    }
    if (A.gn < A.n) {
        plasma_error("invalid leading column dimensions");
        return PlasmaErrorIllegalValue;
    }
    if ((A.uplo == PlasmaGeneral &&
         A.gm < A.mb * ((2 * A.kl + A.ku + A.mb) / A.mb)) ||
        (A.uplo == PlasmaUpper && A.gm < A.mb * ((A.ku + A.mb) / A.mb)) ||
        (A.uplo == PlasmaLower && A.gm < A.mb * ((A.kl + A.mb) / A.mb))) {
        plasma_error("invalid leading row dimensions");
        return PlasmaErrorIllegalValue;
    }
    if ((A.i > 0 && A.i >= A.gm) || (A.j > 0 && A.j >= A.gn)) {
        plasma_error("beginning of the matrix out of bounds");
        return PlasmaErrorIllegalValue;
    }
    if (A.j + A.n > A.gn) {
        plasma_error("submatrix out of bounds");
        return PlasmaErrorIllegalValue;
    }
    if (A.i % A.mb != 0 || A.j % A.nb != 0) {
        plasma_error("submatrix not aligned to a tile");
        return PlasmaErrorIllegalValue;
    }
    if (A.kl + 1 > A.m || A.ku + 1 > A.n) {
        plasma_error("band width larger than matrix dimension");
        return PlasmaErrorIllegalValue;
    }
    return PlasmaSuccess;
}

int plasma_desc_check(plasma_desc_t A)
{
    if (A.type == PlasmaGeneral || A.type == PlasmaUpper ||
        A.type == PlasmaLower)
        return plasma_desc_general_check(A);
    if (A.type == PlasmaGeneralBand)
        return plasma_desc_general_band_check(A);
    plasma_error("invalid matrix type");
    return PlasmaErrorIllegalValue;
}

plasma_desc_t plasma_desc_view(plasma_desc_t A, int i, int j, int m, int n)
{
    if (A.i + i + m > A.gm)
        plasma_fatal_error("rows out of bound");
    if (A.j + j + n > A.gn)
        plasma_fatal_error("columns out of bound");
    plasma_desc_t B = A;
    B.i = A.i + i;
    B.j = A.j + j;
    B.m = m;
    B.n = n;
    B.mt = (m == 0) ? 0 : (B.i + m - 1) / A.mb - B.i / A.mb + 1;
    B.nt = (n == 0) ? 0 : (B.j + n - 1) / A.nb - B.j / A.nb + 1;
    return B;
}
