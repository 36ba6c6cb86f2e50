/*
 * plasma_types.h -- enumerations and scalar types of the PLASMA C ABI.
 *
 * Every numeric value below is part of the binary interface that callers of
 * PLASMA compile against, so each one equals the value in the reference
 * (include/plasma_types.h:49-164): data types 0..5, transposition 111..113,
 * triangle 121..124, diagonal 131..132, side 141..142, norms 171..178,
 * direction 391..392, storage 401..402, error codes 0 / -101..-109 and the
 * context parameters PlasmaTuning..PlasmaHouseholderMode.
 */
#ifndef PLASMA_TYPES_H
#define PLASMA_TYPES_H

#include <limits.h>

/* Library version of this MI355X implementation of the PLASMA Cholesky path. */
#define PLASMA_VERSION_MAJOR 25
#define PLASMA_VERSION_MINOR 5
#define PLASMA_VERSION_PATCH 27

/* Element types (reference include/plasma_types.h:49-57). */
enum {
    PlasmaByte            = 0,
    PlasmaInteger         = 1,
    PlasmaRealFloat       = 2,
    PlasmaRealDouble      = 3,
    PlasmaComplexFloat    = 4,
    PlasmaComplexDouble   = 5,
    PlasmaDatatypeUnknown = INT_MIN
};

/* CBLAS / LAPACK constants (reference include/plasma_types.h:70-118). */
enum {
    PlasmaInvalid       = -1,

    PlasmaNoTrans       = 111,
    PlasmaTrans         = 112,
    PlasmaConjTrans     = 113,
    Plasma_ConjTrans    = PlasmaConjTrans,

    PlasmaUpper         = 121,
    PlasmaLower         = 122,
    PlasmaGeneral       = 123,
    PlasmaGeneralBand   = 124,

    PlasmaNonUnit       = 131,
    PlasmaUnit          = 132,

    PlasmaLeft          = 141,
    PlasmaRight         = 142,

    PlasmaOneNorm       = 171,
    PlasmaRealOneNorm   = 172,
    PlasmaTwoNorm       = 173,
    PlasmaFrobeniusNorm = 174,
    PlasmaInfNorm       = 175,
    PlasmaRealInfNorm   = 176,
    PlasmaMaxNorm       = 177,
    PlasmaRealMaxNorm   = 178,

    PlasmaForward       = 391,
    PlasmaBackward      = 392,

    PlasmaColumnwise    = 401,
    PlasmaRowwise       = 402,

    PlasmaConstUnknown  = INT_MIN
};

/* Complex element types of the z / c routines (reference
 * include/plasma_types.h:170-171): interleaved (re, im) pairs. */
typedef float  _Complex plasma_complex32_t;
typedef double _Complex plasma_complex64_t;

/* Return codes (reference include/plasma_types.h:123-137). */
enum {
    PlasmaSuccess             = 0,
    PlasmaErrorNotInitialized = -101,
    PlasmaErrorNotSupported   = -102,
    PlasmaErrorIllegalValue   = -103,
    PlasmaErrorOutOfMemory    = -104,
    PlasmaErrorNullParameter  = -105,
    PlasmaErrorInternal       = -106,
    PlasmaErrorSequence       = -107,
    PlasmaErrorComponent      = -108,
    PlasmaErrorEnvironment    = -109,
    PlasmaErrorUnknown        = INT_MIN
};

enum {
    PlasmaDisabled     = 0,
    PlasmaEnabled      = 1,
    PlasmaAbledUnknown = INT_MAX
};

enum {
    PlasmaInplace,
    PlasmaOutplace,
    PlasmaInoutUnknown = INT_MAX
};

enum {
    PlasmaFlatHouseholder,
    PlasmaTreeHouseholder,
    PlasmaHouseholderUnknown = INT_MAX
};

/*
 * Context parameters for plasma_set / plasma_get (reference
 * include/plasma_types.h:157-165).  The values after PlasmaHouseholderMode
 * are extensions of this implementation; their numbers start well above the
 * reference's range so they can never collide with it.
 */
enum {
    PlasmaTuning,
    PlasmaNb,
    PlasmaIb,
    PlasmaInplaceOutplace,
    PlasmaNumPanelThreads,
    PlasmaHouseholderMode,
    PlasmaLookahead        = 1001, /* depth of the panel lookahead, 0..4 */
    PlasmaGraphCapture     = 1002, /* PlasmaEnabled: replay hipGraphs */
    PlasmaDevice           = 1003, /* HIP device ordinal (read only) */
    PlasmaUpdateDepth      = 1004, /* tile columns per trailing update, 1..4 */
    PlasmaSplitUpdate      = 1005, /* FP32 updates on 16-bit MFMA (below) */
    PlasmaHalfInput        = 1006, /* input format of plasma_dhposv's factor */
    PlasmaParamUnknown     = INT_MAX
};

/*
 * Values of PlasmaSplitUpdate (extension).  FP32 trailing updates run on the
 * 16-bit matrix cores from exact splits of the panel entries:
 *   PlasmaSplitFp16: rows scaled by powers of two so every factor entry is
 *                    at most 1 in magnitude, two fp16 terms, three products
 *                    (22-bit operands, FP32 accumulation);
 *   PlasmaSplitBf16: three bf16 terms, six products (FP32-exact products);
 *   PlasmaEnabled:   the default split (PlasmaSplitBf16; plasma_get
 *                    returns the split in force);
 *   PlasmaDisabled:  FP32 MFMA.
 * Values of PlasmaHalfInput: PlasmaSplitBf16 (default) or PlasmaSplitFp16,
 * the one-term input format of plasma_dhposv's low-precision factor.
 */
enum {
    PlasmaSplitBf16 = 2,
    PlasmaSplitFp16 = 3
};

typedef int plasma_enum_t;

#ifdef __cplusplus
extern "C" {
#endif

/* 'U'/'u' -> PlasmaUpper, 'L'/'l' -> PlasmaLower, else PlasmaGeneral
 * (reference control/constants.c:124-130). */
plasma_enum_t plasma_uplo_const(char lapack_char);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_TYPES_H */
