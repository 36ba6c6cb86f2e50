/*
 * internal.h -- private state and helpers of the MI355X PLASMA library.
 *
 * The host side is C.  It owns the context (device, streams, status words,
 * device workspace pool, graph cache) and schedules the tile algorithms as
 * stream-ordered kernel launches through the thin C ABI declared in
 * kernels/plasma_kernels.h.
 */
#ifndef PLASMA_INTERNAL_H
#define PLASMA_INTERNAL_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#include "plasma.h"
#include "kernels/plasma_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* Device status words: one per live sequence (see control/async.c). */
#define PLASMA_STATUS_SLOTS 256

/* Cached device allocations kept by the workspace pool. */
#define PLASMA_POOL_ENTRIES 64

typedef struct {
    void  *ptr;
    size_t bytes;
    int    in_use;
} plasma_pool_entry_t;

/* Autotuned schedule table (control/tuning.c). */
#define PLASMA_TUNE_ROWS 64
typedef struct {
    int precision, n_max, nb, depth, lookahead;
    int mode;                 /* 0 any, 1 eager launches, 2 graph replay */
    int split;                /* FP32 update operands: 0 as set, -1 FP32,
                                 PlasmaSplitBf16, PlasmaSplitFp16 */
} plasma_tune_row_t;

/* Deepest panel lookahead (PlasmaLookahead): the panel stream runs up to
 * this many column groups ahead of the bulk trailing update. */
#define PLASMA_MAX_LOOKAHEAD 4

/* Cached factorization graphs (compute/pxpotrf.c). */
#define PLASMA_GRAPHS 4
typedef struct {
    int used;
    unsigned long stamp;
    plasma_desc_t A;          /* the key: descriptor, uplo, inverse   */
    int uplo;                 /* workspace argument and schedule      */
    const void *inv_arg;
    int lookahead, group, split;
    hipGraphExec_t exec;
    void *own_inv;            /* buffers the graph's launches address */
    void *panel[PLASMA_MAX_LOOKAHEAD + 1];
} plasma_graph_entry_t;

/* profiler records (control/profile.c) */
#define PLASMA_PROF_MAX 16384
typedef struct {
    hipEvent_t beg, end;
    double flops, bytes;
    int cls;
} plasma_prof_rec_t;

struct plasma_context_s {
    /* reference parameters (control/context.c:165-180) */
    int tuning;
    int nb;
    int ib;
    int inplace_outplace;
    int max_threads;
    int max_panel_threads;
    int householder_mode;

    /* extensions */
    int lookahead;        /* panel lookahead depth, 0 disables */
    int group;            /* tile columns whose panels update together */
    int split_update;     /* FP32 trailing updates on 16-bit MFMA:
                             PlasmaSplitFp16 / PlasmaSplitBf16 /
                             PlasmaEnabled (default) / PlasmaDisabled */
    int half_input;       /* plasma_dhposv's factor input format:
                             PlasmaSplitFp16 or PlasmaSplitBf16 */
    int low_factor;       /* set by plasma_dhposv around its factorization:
                             FP32 trailing updates from one-term 16-bit
                             inputs (PlasmaLowBf16 / PlasmaLowFp16), 0 off */
    int graph_capture;    /* replay cached hipGraphs for the sync drivers */

    /* device */
    int device;
    int num_cus;
    hipStream_t stream;        /* main stream: trailing updates, sweeps */
    hipStream_t panel_stream;  /* lookahead stream: diagonal potrf + panel */
    int owns_streams;

    /* status words, one per sequence with work in flight; a slot is free
     * when slot_seq is NULL (control/async.c) */
    int *d_status;                                   /* device array */
    plasma_sequence_t *slot_seq[PLASMA_STATUS_SLOTS];
    plasma_request_t *slot_req[PLASMA_STATUS_SLOTS];
    int slot_next;

    /* device scratch for reductions (norms, convergence flags) */
    double *d_scalars;          /* PLASMA_SCALARS doubles */

    /* workspace pool */
    plasma_pool_entry_t pool[PLASMA_POOL_ENTRIES];

    /* factorization graphs; their launches use the status word
     * d_status[PLASMA_STATUS_SLOTS], copied from and back to the
     * sequence's word around each replay */
    plasma_graph_entry_t graphs[PLASMA_GRAPHS];
    unsigned long graph_clock;

    /* autotuned schedules */
    plasma_tune_row_t tune[PLASMA_TUNE_ROWS];
    int ntune;

    /* kernel profiler */
    int profile;
    int nprof;
    plasma_prof_rec_t *prof;
    double prof_ms[PlasmaProfClasses];
    double prof_flops[PlasmaProfClasses];
    double prof_bytes[PlasmaProfClasses];
    long prof_count[PlasmaProfClasses];
};

#define PLASMA_SCALARS 64

/* the split PlasmaSplitUpdate = PlasmaEnabled selects (plasma_get reports
 * the mode in force) */
#define PLASMA_DEFAULT_SPLIT PlasmaSplitBf16

/* values of plasma_context_t.low_factor */
#define PlasmaLowBf16 1
#define PlasmaLowFp16 2


void plasma_prof_begin(hipStream_t s, int cls, double flops, double bytes,
                       int *slot);
void plasma_prof_end(hipStream_t s, int slot);

/* context.c */
plasma_context_t *plasma_context_self(void);
int plasma_hip_check(hipError_t err, const char *what);

/* async.c: the device status word of a sequence (allocated on first use) */
int *plasma_sequence_dstatus(plasma_sequence_t *sequence,
                             plasma_request_t *request);

/* control/tuning.c: load the table at init; apply it to a factorization
 * of order n (sets nb, update depth and lookahead in the context) */
void plasma_tuning_init(plasma_context_t *ctx);
void plasma_tune_potrf(plasma_context_t *ctx, plasma_enum_t precision, int n);

/* compute/pxpotrf.c: drop every cached graph (streams drained first) */
void plasma_graph_release_all(plasma_context_t *ctx);

/* workspace.c: device memory from the pool */
void *plasma_dev_alloc(size_t bytes);
void  plasma_dev_free(void *ptr);
void  plasma_pool_release_all(plasma_context_t *ctx);

/* Pointer kind of a caller's LAPACK-layout array. */
int plasma_is_device_ptr(const void *p);

/* A caller's LAPACK array: used in place when it is device memory, staged
 * through a device buffer (ld = m) otherwise.  stage_in copies it in on the
 * context stream when copy_in is set; stage_out copies it back (when staged
 * and copy_back is set, synchronously) and frees the staging buffer. */
typedef struct {
    void *user;
    int user_ld;
    void *dev;
    int ld;
    int m, n;
    size_t es;
    int staged;
} lapack_arg_t;
int plasma_arg_stage_in(lapack_arg_t *a, void *p, int ld, int m, int n,
                        size_t es, int copy_in);
void plasma_arg_stage_out(lapack_arg_t *a, int copy_back);

/*
 * Parallel tile algorithms (compute/p*.c).  Each enqueues its launches on
 * the context streams and returns; the main stream is the join point.
 */
/* inv: the diagonal-tile inverses (mt * mb * mb elements of A's
 * precision).  pxpotrf writes them when inv != NULL (and uses a temporary
 * buffer otherwise); pxpotrs reads them, computing them first when
 * inv == NULL. */
void plasma_pxpotrf(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                    plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_pxpotrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                    const void *inv, plasma_sequence_t *sequence,
                    plasma_request_t *request);
/* complex descriptors (compute/zchol.c), reached through the two above */
void plasma_pzpotrf(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                    plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_pzpotrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                    const void *inv, plasma_sequence_t *sequence,
                    plasma_request_t *request);
void plasma_pxtrsm_left(plasma_enum_t uplo, plasma_enum_t trans,
                        plasma_desc_t A, plasma_desc_t B, const void *inv,
                        plasma_sequence_t *sequence,
                        plasma_request_t *request);
/* bytes of the inverse workspace for A */
size_t plasma_inv_bytes(plasma_desc_t A);
/* Repeated solves with one factor: M(m,k) = L(m,k) W_kk for m > k
 * (plasma_mpanel_bytes(A) bytes), computed once by pxscale_panels; then
 * pxpotrs_m solves with one launch per sweep step. */
size_t plasma_mpanel_bytes(plasma_desc_t A);
void plasma_pxscale_panels(plasma_enum_t uplo, plasma_desc_t A,
                           const void *inv, void *M,
                           plasma_sequence_t *sequence,
                           plasma_request_t *request);
void plasma_pxpotrs_m(plasma_desc_t A, plasma_desc_t B, const void *inv,
                      const void *M, plasma_sequence_t *sequence,
                      plasma_request_t *request);
/* Repeated solves by super-blocks of plasma_sweep_tiles(A) tiles (0: the
 * path does not apply): V = the diagonal super-blocks' inverses
 * (plasma_sbinv_bytes(A) bytes), computed once by pxsbinv; then pxpotrs_sb
 * solves with two launches per super-block and sweep. */
int plasma_sweep_tiles(plasma_desc_t A);
size_t plasma_sbinv_bytes(plasma_desc_t A);
void plasma_pxsbinv(plasma_enum_t uplo, plasma_desc_t A, const void *inv,
                    void *V, plasma_sequence_t *sequence,
                    plasma_request_t *request);
void plasma_pxpotrs_sb(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                       const void *V, plasma_sequence_t *sequence,
                       plasma_request_t *request);
/* Band Cholesky and its solve on a band descriptor (PlasmaGeneralBand);
 * inv: the diagonal-tile inverses (A.mt tiles), NULL in pxpbtrf for a
 * private buffer.  plasma_band_tiles: tile rows of a band column. */
int plasma_band_tiles(plasma_enum_t uplo, plasma_desc_t A);
void plasma_pxpbtrf(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                    plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_pxpbtrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                    const void *inv, plasma_sequence_t *sequence,
                    plasma_request_t *request);
/* Either form for a factor solved many times: rsolve_prepare computes it
 * into one pool block (NULL: none, the plain sweeps run), rsolve solves. */
void *plasma_rsolve_prepare(plasma_enum_t uplo, plasma_desc_t A,
                            const void *inv, plasma_sequence_t *sequence,
                            plasma_request_t *request);
void plasma_rsolve(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                   const void *inv, const void *form,
                   plasma_sequence_t *sequence, plasma_request_t *request);

/* layout translation between a LAPACK array (host or device) and tiles */
void plasma_pxlapack2desc(plasma_enum_t region, const void *pA, int lda,
                          plasma_desc_t A, plasma_sequence_t *sequence,
                          plasma_request_t *request);
void plasma_pxdesc2lapack(plasma_enum_t region, plasma_desc_t A, void *pA,
                          int lda, plasma_sequence_t *sequence,
                          plasma_request_t *request);

/* precision-generic drivers behind plasma_{d,s}{potrf,potrs,posv} */
int plasma_xpotrf(plasma_enum_t precision, plasma_enum_t uplo, int n,
                  void *pA, int lda);
int plasma_xpotrs(plasma_enum_t precision, plasma_enum_t uplo, int n,
                  int nrhs, void *pA, int lda, void *pB, int ldb);
int plasma_xposv(plasma_enum_t precision, plasma_enum_t uplo, int n,
                 int nrhs, void *pA, int lda, void *pB, int ldb);
void plasma_omp_xpotrf(plasma_enum_t precision, plasma_enum_t uplo,
                       plasma_desc_t A, plasma_sequence_t *sequence,
                       plasma_request_t *request);
void plasma_omp_xpotrs(plasma_enum_t precision, plasma_enum_t uplo,
                       plasma_desc_t A, plasma_desc_t B,
                       plasma_sequence_t *sequence, plasma_request_t *request);
void plasma_omp_xposv(plasma_enum_t precision, plasma_enum_t uplo,
                      plasma_desc_t A, plasma_desc_t B,
                      plasma_sequence_t *sequence, plasma_request_t *request);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_INTERNAL_H */
