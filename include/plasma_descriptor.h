/*
 * plasma_descriptor.h -- the PLASMA tile matrix descriptor.
 *
 * plasma_desc_t keeps the reference's field order and meaning
 * (include/plasma_descriptor.h:40-76) so that code written against PLASMA's
 * tile API compiles and links unchanged.  The one change of meaning is where
 * the storage lives: `matrix` points to DEVICE memory (HBM of the MI355X the
 * library was initialised on).  The storage format inside that buffer is
 * PLASMA's tile layout, unchanged:
 *
 *   - every tile is stored contiguously, column major, with leading
 *     dimension mb (tile rows of full height) or gm%mb (the last, short
 *     tile row);
 *   - a general descriptor stores the full tiles column by column (region
 *     A11), then the short last tile row (A21), the short last tile column
 *     (A12) and the corner tile (A22);
 *   - a triangular descriptor stores only the tiles of one triangle
 *     (type == PlasmaLower: m >= n, type == PlasmaUpper: m <= n), column by
 *     column, followed by the edge regions.
 *
 * The address helpers below compute the same offsets as the reference
 * (general: include/plasma_descriptor.h:93-118; triangular: :121-151) but in
 * size_t throughout, and they compile for both the host and the GPU so the
 * HIP kernels address tiles with exactly the host's formula.
 */
#ifndef PLASMA_DESCRIPTOR_H
#define PLASMA_DESCRIPTOR_H

#include <stddef.h>
#include "plasma_types.h"

#if defined(__HIPCC__)
#define PLASMA_HD __host__ __device__
#else
#define PLASMA_HD
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* matrix properties */
    plasma_enum_t type;       /* PlasmaGeneral, PlasmaUpper, PlasmaLower ... */
    plasma_enum_t uplo;       /* triangle of a band matrix */
    plasma_enum_t precision;  /* PlasmaRealDouble, PlasmaRealFloat ... */

    /* storage: device pointer and element offsets of the edge regions */
    void  *matrix;
    size_t A21;
    size_t A12;
    size_t A22;

    /* tile size */
    int mb;
    int nb;

    /* the whole stored matrix */
    int gm;
    int gn;
    int gmt;
    int gnt;

    /* the submatrix (view) this descriptor addresses */
    int i;
    int j;
    int m;
    int n;
    int mt;
    int nt;

    /* band parameters (kept for layout compatibility) */
    int kl;
    int ku;
    int klt;
    int kut;
} plasma_desc_t;

static inline PLASMA_HD size_t plasma_element_size(int type)
{
    switch (type) {
    case PlasmaByte:          return 1;
    case PlasmaInteger:       return sizeof(int);
    case PlasmaRealFloat:     return sizeof(float);
    case PlasmaRealDouble:    return sizeof(double);
    case PlasmaComplexFloat:  return 2 * sizeof(float);
    case PlasmaComplexDouble: return 2 * sizeof(double);
    default:                  return 0;
    }
}

/* Element offset of tile (m, n) of the view in a general descriptor. */
static inline PLASMA_HD size_t plasma_tile_offset_general(plasma_desc_t A,
                                                          int m, int n)
{
    int tm = m + A.i / A.mb;          /* tile coordinates in the stored matrix */
    int tn = n + A.j / A.nb;
    int full_mt = A.gm / A.mb;        /* tile rows / columns of full size */
    int full_nt = A.gn / A.nb;
    if (tm < full_mt) {
        if (tn < full_nt)
            return (size_t)A.mb * A.nb * ((size_t)tm + (size_t)full_mt * tn);
        return A.A12 + (size_t)A.mb * (A.gn % A.nb) * tm;
    }
    if (tn < full_nt)
        return A.A21 + (size_t)A.nb * (A.gm % A.mb) * tn;
    return A.A22;
}

/* Element offset of tile (m, n) in a triangular (PlasmaLower/Upper) one. */
static inline PLASMA_HD size_t plasma_tile_offset_triangle(plasma_desc_t A,
                                                           int m, int n)
{
    int tm = m + A.i / A.mb;
    int tn = n + A.j / A.nb;
    int full_mt = A.gm / A.mb;
    int full_nt = A.gn / A.nb;
    if (tm < full_mt) {
        if (tn < full_nt) {
            size_t tile;
            if (A.type == PlasmaUpper)   /* tiles tm <= tn, column by column */
                tile = (size_t)tm + ((size_t)tn * (tn + 1)) / 2;
            else                         /* tiles tm >= tn, column by column */
                tile = (size_t)(tm - tn)
                     + ((size_t)tn * (size_t)(2 * full_mt - tn + 1)) / 2;
            return (size_t)A.mb * A.nb * tile;
        }
        return A.A12 + (size_t)A.mb * (A.gn % A.nb) * tm;
    }
    if (tn < full_nt)
        return A.A21 + (size_t)A.nb * (A.gm % A.mb) * tn;
    return A.A22;
}

static inline PLASMA_HD size_t plasma_tile_offset(plasma_desc_t A, int m, int n)
{
    if (A.type == PlasmaUpper || A.type == PlasmaLower)
        return plasma_tile_offset_triangle(A, m, n);
    if (A.type == PlasmaGeneralBand)
        return plasma_tile_offset_general(A, (A.kut - 1) + m - n, n);
    return plasma_tile_offset_general(A, m, n);
}

/* Address of tile (m, n) (reference include/plasma_descriptor.h:159-175). */
static inline PLASMA_HD void *plasma_tile_addr(plasma_desc_t A, int m, int n)
{
    return (void *)((char *)A.matrix
                    + plasma_tile_offset(A, m, n)
                      * plasma_element_size(A.precision));
}

/* Leading dimension of tile row k (reference :182-194). */
static inline PLASMA_HD int plasma_tile_mmain(plasma_desc_t A, int k)
{
    if (A.type == PlasmaGeneralBand)
        return A.mb;
    return (A.i / A.mb + k < A.gm / A.mb) ? A.mb : A.gm % A.mb;
}

/* Stored width of tile column k (reference :201-207). */
static inline PLASMA_HD int plasma_tile_nmain(plasma_desc_t A, int k)
{
    return (A.j / A.nb + k < A.gn / A.nb) ? A.nb : A.gn % A.nb;
}

/* Rows of the view inside tile row k (reference :214-223). */
static inline PLASMA_HD int plasma_tile_mview(plasma_desc_t A, int k)
{
    if (k < A.mt - 1)
        return A.mb;
    int r = (A.i + A.m) % A.mb;
    return r == 0 ? A.mb : r;
}

/* Columns of the view inside tile column k (reference :230-239). */
static inline PLASMA_HD int plasma_tile_nview(plasma_desc_t A, int k)
{
    if (k < A.nt - 1)
        return A.nb;
    int r = (A.j + A.n) % A.nb;
    return r == 0 ? A.nb : r;
}

/*
 * Descriptor management (reference control/descriptor.c).  *_create
 * allocate the tile storage in device memory from the library's workspace
 * pool; *_destroy returns it.  *_init only fill the struct around a
 * caller-provided device buffer.
 */
int plasma_desc_general_create(plasma_enum_t precision, int mb, int nb,
                               int lm, int ln, int i, int j, int m, int n,
                               plasma_desc_t *A);

int plasma_desc_triangular_create(plasma_enum_t precision, plasma_enum_t uplo,
                                  int mb, int nb, int lm, int ln,
                                  int i, int j, int m, int n,
                                  plasma_desc_t *A);

int plasma_desc_destroy(plasma_desc_t *A);

int plasma_desc_general_init(plasma_enum_t precision, void *matrix,
                             int mb, int nb, int lm, int ln, int i, int j,
                             int m, int n, plasma_desc_t *A);

int plasma_desc_triangular_init(plasma_enum_t precision, plasma_enum_t uplo,
                                void *matrix, int mb, int nb, int lm, int ln,
                                int i, int j, int m, int n, plasma_desc_t *A);

/* Band descriptors (reference control/descriptor.c:50-66, :178-212,
 * :322-371): general storage of lm rows (a multiple of mb covering the
 * band) by ln columns; tile (m, n) of the band matrix is stored at tile
 * ((kut - 1) + m - n, n).  uplo PlasmaLower keeps kl subdiagonals, Upper
 * ku superdiagonals, General both. */
int plasma_desc_general_band_create(plasma_enum_t precision,
                                    plasma_enum_t uplo, int mb, int nb,
                                    int lm, int ln, int i, int j, int m,
                                    int n, int kl, int ku, plasma_desc_t *A);
int plasma_desc_general_band_init(plasma_enum_t precision,
                                  plasma_enum_t uplo, void *matrix, int mb,
                                  int nb, int lm, int ln, int i, int j,
                                  int m, int n, int kl, int ku,
                                  plasma_desc_t *A);
int plasma_desc_general_band_check(plasma_desc_t A);

int plasma_desc_check(plasma_desc_t A);
int plasma_desc_general_check(plasma_desc_t A);

plasma_desc_t plasma_desc_view(plasma_desc_t A, int i, int j, int m, int n);

/* Bytes of tile storage a descriptor created with these arguments needs. */
size_t plasma_desc_storage_bytes(plasma_desc_t A);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_DESCRIPTOR_H */
