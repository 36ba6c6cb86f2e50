/*
 * pxpotrf.c -- tile Cholesky and its triangular sweeps as lookahead-
 * scheduled, stream-ordered launch sequences.
 *
 * The reference inserts one OpenMP task per tile operation and lets the
 * runtime order them by their depend clauses (compute/pzpotrf.c:37-80 for
 * lower, :84-126 for upper; 341 376 gemm tasks at n = 32768, nb = 256).
 * Here every step k issues one launch per task class covering all tiles of
 * that step:
 *     potrf(k)            diagonal tile, plus W_kk = L(k,k)^{-1}
 *     trsm(k)             panel tiles m > k: A(m,k) W_kk^T (MFMA)
 *     update(k, n0, n1)   trailing tiles of columns n0..n1 (MFMA)
 * Steps are taken in groups of g consecutive tile columns (g = 2 by
 * default): the panel stream factors a group column by column (each column
 * first updated by the group's earlier panels), and the trailing matrix is
 * then updated once by all g panels together, a product with K = g * nb
 * instead of g products with K = nb (the MFMA update is 15-20 % faster at
 * K = 512 than at K = 256 on gfx950).  With lookahead depth d the panel
 * stream runs up to d groups ahead: it applies the d previous groups'
 * panels to a group's columns itself and factors the group, while the main
 * stream's bulk update of each group covers only the columns beyond that
 * window (pxpotrf_groups).  The panel solves write into panel buffers that
 * the updates read; the copy into A rides along with the bulk update, off
 * the critical path.  Events carry the dependencies the depend clauses
 * expressed, so the critical path potrf -> trsm -> next column overlaps the
 * bulk updates.  The streams are joined again before returning.
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static int issue(int err, plasma_sequence_t *sequence,
                 plasma_request_t *request)
{
    if (err != 0)
        plasma_request_fail(sequence, request, err);
    return err;
}

size_t plasma_inv_bytes(plasma_desc_t A)
{
    return (size_t)A.mt * A.mb * A.mb * plasma_element_size(A.precision);
}

/* Algorithmic flops of the launches (reference test/flops.h conventions:
 * potrf n^3/3 + n^2/2 + n/6, trsm m n^2, syrk k n (n + 1), gemm 2 m n k). */
static double potrf_flops(double n)
{
    return n * n * n / 3.0 + n * n / 2.0 + n / 6.0;
}

static double panel_flops(plasma_desc_t A, int k)
{
    double f = 0, nb = A.nb;
    for (int m = k + 1; m < A.mt; m++)
        f += plasma_tile_mview(A, m) * nb * nb;
    return f;
}

static double update_flops(plasma_desc_t A, int kdepth, int n0, int n1)
{
    double f = 0, kd = kdepth;
    for (int n = n0; n < n1 && n < A.mt; n++) {
        double nv = plasma_tile_mview(A, n);
        f += kd * nv * (nv + 1);                           /* syrk (n, n) */
        for (int m = n + 1; m < A.mt; m++)
            f += 2.0 * plasma_tile_mview(A, m) * nv * kd;  /* gemm */
    }
    return f;
}

static int k_potrf(plasma_enum_t uplo, plasma_desc_t A, int k, void *inv,
                   int *st, hipStream_t s)
{
    int slot;
    plasma_prof_begin(s, PlasmaProfPotrf, potrf_flops(plasma_tile_mview(A, k)),
                      0, &slot);
    int err = plasma_kern_potrf_diag(uplo, A, k, inv, st, s);
    plasma_prof_end(s, slot);
    return err;
}

/* How the trailing updates of an FP32 factorization run (use_split). */
enum {
    SPLIT_NONE = 0,     /* FP32 MFMA */
    SPLIT_BF16X3,       /* three bf16 terms, six products (FP32-exact) */
    SPLIT_BF16X1,       /* bf16 inputs, one product */
    SPLIT_FP16X2,       /* row-scaled two-term fp16, three products */
    SPLIT_FP16X1        /* row-scaled fp16 inputs, one product */
};

static int split_is_fp16(int split)
{
    return split == SPLIT_FP16X2 || split == SPLIT_FP16X1;
}

/* the column-group context of the launches below */
typedef struct {
    plasma_enum_t uplo;
    plasma_desc_t A;
    void *inv;
    int *st;
    int pcols;          /* columns of a panel-buffer tile: g * nb */
    int split;          /* SPLIT_*: operand terms after the FP32 panel */
    int blocked;        /* diagonal-block inverses only; blocked panel solve */
    const float *rscale;    /* fp16 terms: row scales (in panel buffer 0) */
} group_t;

/* FP32 factorizations whose tiles are a multiple of 32 run their trailing
 * updates on the 16-bit matrix cores: the low-precision input factor of the
 * GMRES-IR solver takes one product of bf16 or fp16 inputs; otherwise
 * PlasmaSplitUpdate chooses the split (PlasmaEnabled selects
 * PLASMA_DEFAULT_SPLIT when set) or FP32 MFMA (PlasmaDisabled). */
static int use_split(plasma_desc_t A)
{
    plasma_context_t *ctx = plasma_context_self();
    if (A.precision != PlasmaRealFloat || A.nb % 32 != 0 || A.mb != A.nb)
        return SPLIT_NONE;
    if (ctx->low_factor == PlasmaLowBf16)
        return SPLIT_BF16X1;
    if (ctx->low_factor == PlasmaLowFp16)
        return SPLIT_FP16X1;
    switch (ctx->split_update) {
    case PlasmaSplitBf16:  return SPLIT_BF16X3;
    case PlasmaSplitFp16:  return SPLIT_FP16X2;
    default:               return SPLIT_NONE;
    }
}

/* Bytes of the FP32 part of a panel buffer, rounded up so that the split
 * planes that follow it are 256-byte aligned. */
static size_t panel_fp_bytes(plasma_desc_t A, int pcols)
{
    const size_t b = (size_t)imax(A.mt - 1, 1) * A.mb * pcols *
                     plasma_element_size(A.precision);
    return (b + 255) / 256 * 256;
}

/* Bytes of the operand terms of a panel buffer. */
static size_t planes_bytes(plasma_desc_t A, int pcols, int split)
{
    switch (split) {
    case SPLIT_BF16X3:
    case SPLIT_BF16X1: return plasma_split_bytes(A, pcols);
    case SPLIT_FP16X2: return plasma_split16_bytes(A, pcols, 2);
    case SPLIT_FP16X1: return plasma_split16_bytes(A, pcols, 1);
    default:           return 0;
    }
}

static void *planes_of(const group_t *g, const void *panel)
{
    return (char *)panel + panel_fp_bytes(g->A, g->pcols);
}

/* The diagonal tile of column k: with the blocked panel solve only the
 * 32 x 32 diagonal blocks of W_kk are formed on the critical path. */
static int k_potrf_g(const group_t *g, int k, hipStream_t s)
{
    if (!g->blocked)
        return k_potrf(g->uplo, g->A, k, g->inv, g->st, s);
    int slot;
    plasma_prof_begin(s, PlasmaProfPotrf,
                      potrf_flops(plasma_tile_mview(g->A, k)), 0, &slot);
    int err = plasma_kern_potrf_diag_blocks(g->uplo, g->A, k, g->inv, g->st,
                                            s);
    plasma_prof_end(s, slot);
    return err;
}

static int k_panel(const group_t *g, int k, void *panel, int pbase,
                   hipStream_t s)
{
    int slot;
    plasma_prof_begin(s, PlasmaProfPanel, panel_flops(g->A, k), 0, &slot);
    int err = g->blocked
        ? plasma_kern_trsm_panel_blocked(g->uplo, g->A, k, g->inv, panel,
                                         pbase, g->pcols,
                                         (k - pbase) * g->A.nb, g->st, s)
        : plasma_kern_trsm_panel(g->uplo, g->A, k, g->inv, panel, pbase,
                                 g->pcols, (k - pbase) * g->A.nb, g->st, s);
    if (!err && split_is_fp16(g->split))
        err = plasma_kern_split16_panel(g->A, k, pbase, g->pcols, panel,
                                        planes_of(g, panel),
                                        g->split == SPLIT_FP16X2 ? 2 : 1,
                                        g->rscale, g->st, s);
    else if (!err && g->split)
        err = plasma_kern_split_panel(g->A, k, pbase, g->pcols, panel,
                                      planes_of(g, panel), g->st, s);
    plasma_prof_end(s, slot);
    return err;
}

/* FP64 trailing updates with their operands loaded straight into LDS
 * (kernels/update64.hip) unless PLASMA_FP64_UPDATE=staged selects the
 * register-staged MFMA kernel (kept for comparison and for tile sizes that
 * are not a multiple of 128). */
static int update64_on(void)
{
    static int flag = -1;
    if (flag < 0) {
        const char *e = getenv("PLASMA_FP64_UPDATE");
        flag = !(e != NULL && strcmp(e, "staged") == 0);
    }
    return flag;
}

static int k_update(const group_t *g, int pbase, int ncols, int n0, int n1,
                    const void *panel, int copy_cols, int max_wgs,
                    hipStream_t s)
{
    int slot;
    plasma_context_t *ctx = plasma_context_self();
    const int kdepth = ncols * g->A.nb;
    plasma_prof_begin(s, PlasmaProfUpdate,
                      ctx->profile ? update_flops(g->A, kdepth, n0, n1) : 0,
                      0, &slot);
    int err;
    if (split_is_fp16(g->split))
        err = plasma_kern_trailing_update_split16(
            g->uplo, g->A, pbase, kdepth, n0, n1, panel, planes_of(g, panel),
            g->pcols, copy_cols, max_wgs, g->split == SPLIT_FP16X2 ? 2 : 1,
            g->rscale, g->st, s);
    else if (g->split)
        err = plasma_kern_trailing_update_split(
            g->uplo, g->A, pbase, kdepth, n0, n1, panel, planes_of(g, panel),
            g->pcols, copy_cols, max_wgs, g->split == SPLIT_BF16X1, g->st, s);
    else if (update64_on() && plasma_update64_ok(g->A, kdepth))
        err = plasma_kern_trailing_update64(g->uplo, g->A, pbase, kdepth, n0,
                                            n1, panel, g->pcols, copy_cols,
                                            max_wgs, g->st, s);
    else
        err = plasma_kern_trailing_update(g->uplo, g->A, pbase, kdepth, n0,
                                          n1, panel, g->pcols, copy_cols,
                                          max_wgs, g->st, s);
    plasma_prof_end(s, slot);
    return err;
}

/* Factor the columns [c0, c0 + w) of a group into its panel buffer: each
 * column is first updated by the group's earlier columns, then factored
 * and its panel solved. */
static int factor_group(const group_t *g, int c0, int w, void *panel,
                        hipStream_t s, plasma_sequence_t *sequence,
                        plasma_request_t *request)
{
    for (int j = 0; j < w; j++) {
        const int col = c0 + j;
        if (j > 0 && issue(k_update(g, c0, j, col, col + 1, panel, 0, 0, s),
                           sequence, request))
            return 1;
        if (issue(k_potrf_g(g, col, s), sequence, request) ||
            issue(k_panel(g, col, panel, c0, s), sequence, request))
            return 1;
    }
    return 0;
}

/* Events of one factorization: the fork, a ring of "group factored" events
 * (panel stream), a ring of "bulk update done" events (main stream), and the
 * join.  Rings of depth + 2 slots: a wait always refers to the latest record
 * of its slot (the slot is recorded again only depth + 2 groups later). */
#define EV_RING (PLASMA_MAX_LOOKAHEAD + 2)
#define PLASMA_FACTOR_EVENTS (2 * EV_RING + 2)

/* The launch sequence proper (main and panel streams joined on return).
 *
 * depth 0: one stream; group q is factored (factor_group) and then updates
 * the whole trailing matrix.
 *
 * depth d >= 1 (lookahead): the panel stream factors group t as soon as
 * every update of its columns is done, and the main stream runs the bulk
 * updates beside it:
 *   panel stream, group t:  wait bulk(t - d - 1);
 *                           narrow(i -> t) for i = t - d .. t - 1 (the
 *                           updates of group t's columns by the d previous
 *                           groups, left-looking, from their panel buffers);
 *                           factor_group(t) into buffer t % (d + 1)
 *   main stream, group t:   wait factor(t);
 *                           bulk(t): columns of groups >= t + d + 1, and
 *                           group t's panels copied back into A
 * so the panel stream runs up to d groups ahead of the bulk updates
 * (reference compute/pzpotrf.c:37-80: the DAG lets the runtime do the same
 * with the potrf / trsm / syrk tasks of the next columns).  Races are
 * excluded by construction: bulk(j) writes groups >= j + d + 1, so the
 * bulk updates that touch group t are exactly those with j <= t - d - 1,
 * all waited for; buffer t % (d + 1) was last read by group t - d - 1,
 * whose narrow updates precede it on the panel stream and whose bulk
 * update is waited for. */
static int pxpotrf_groups(const group_t *gp, int depth, void *panel[],
                          hipEvent_t ev[], plasma_sequence_t *sequence,
                          plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    hipStream_t s0 = ctx->stream, s1 = ctx->panel_stream;
    const plasma_desc_t A = gp->A;
    const int mt = A.mt;
    const int gw = imax(1, imin(ctx->group, mt));
    const int ngroups = (mt + gw - 1) / gw;
    const group_t g = *gp;

    if (depth == 0) {
        for (int c0 = 0; c0 < mt; c0 += gw) {
            const int w = imin(gw, mt - c0);
            if (factor_group(&g, c0, w, panel[0], s0, sequence, request) ||
                issue(k_update(&g, c0, w, c0 + w, mt, panel[0], w, 0, s0),
                      sequence, request))
                return 1;
        }
        return 0;
    }

    hipEvent_t ev_fork = ev[0], ev_join = ev[1];
    hipEvent_t *ev_factor = ev + 2, *ev_bulk = ev + 2 + EV_RING;
    /* The bulk update's grid: one workgroup per job (0: uncapped), so that
     * the hardware dispatcher hands the lookahead stream's workgroups the
     * slots of finishing jobs.  A persistent grid capped below the chip's
     * slots was measured against it: FP32 dsposv n = 32768, cond 1e6, 2
     * CUs - 64 workgroups 97.3 ms, 2 CUs - 128 96.8 ms, 2 CUs 101.8 ms,
     * uncapped 92.6 ms; FP64 capped 262.9 ms, uncapped 249.5 ms. */
    const int bulk_wgs = 0;

    /* fork: the panel stream starts after everything queued so far */
    hipEventRecord(ev_fork, s0);
    hipStreamWaitEvent(s1, ev_fork, 0);

    int err = 0;
    for (int t = 0; t < ngroups && !err; t++) {
        const int c0 = t * gw, w = imin(gw, mt - c0);
        void *pt = panel[t % (depth + 1)];
        /* panel stream: the updates of group t's columns, then group t */
        if (t - depth - 1 >= 0)
            hipStreamWaitEvent(s1, ev_bulk[(t - depth - 1) % EV_RING], 0);
        for (int i = imax(0, t - depth); i < t && !err; i++) {
            const int ci = i * gw;
            err = issue(k_update(&g, ci, imin(gw, mt - ci), c0, c0 + w,
                                 panel[i % (depth + 1)], 0, 0, s1),
                        sequence, request);
        }
        if (!err)
            err = factor_group(&g, c0, w, pt, s1, sequence, request);
        hipEventRecord(ev_factor[t % EV_RING], s1);
        /* main stream: group t's bulk update beyond its window, and its
         * panels back into A */
        hipStreamWaitEvent(s0, ev_factor[t % EV_RING], 0);
        if (!err)
            err = issue(k_update(&g, c0, w, imin(mt, (t + depth + 1) * gw),
                                 mt, pt, w, bulk_wgs, s0),
                        sequence, request);
        hipEventRecord(ev_bulk[t % EV_RING], s0);
    }
    /* join */
    hipEventRecord(ev_join, s1);
    hipStreamWaitEvent(s0, ev_join, 0);
    return err;
}

/* The launch sequence of the factorization: on return every launch is
 * enqueued (main stream joined with the panel stream).  Returns non-zero
 * when a launch failed (the sequence is marked).  ev: three events the
 * caller owns for the duration of the call (and of a capture). */
static int pxpotrf_launch(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                          int want_w, void *panel[], int lookahead, int *st,
                          hipEvent_t ev[], plasma_sequence_t *sequence,
                          plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    const int gw = imax(1, imin(ctx->group, A.mt));
    /* Blocked panel solve (diagonal-block inverses only on the critical
     * path) in FP32, where it shortens the serial chain.  Without
     * lookahead (n = 32768): diagonal tiles 36.4 -> 24.6 ms, panels 6.5 ->
     * 13.1 ms, factor 152.3 -> 147.4 ms.  With lookahead 1 and the fp16
     * split update, the chain of diagonal tile, panel and next column is
     * the critical path of the last half of the factorization: diagonal
     * tiles 27.6 -> 17.9 ms, panels 14.9 -> 26.6 ms, dsposv 98.6 -> 97.8
     * ms (with the uncapped bulk grid: 93.6 -> 92.5 ms).  In FP64 its
     * panel slows more than the diagonal tile gains (252.5 -> 281.8 ms),
     * so FP64 keeps the full-inverse panel GEMM. */
    const int blocked = A.precision == PlasmaRealFloat &&
                        plasma_panel_blocked_ok(A);
    const int split = use_split(A);
    const int pcols = gw * A.nb;
    /* fp16 terms: the row scales follow the planes of panel buffer 0 */
    float *rscale = split_is_fp16(split)
        ? (float *)((char *)panel[0] + panel_fp_bytes(A, pcols) +
                    planes_bytes(A, pcols, split))
        : NULL;
    const group_t g = {uplo, A, inv, st, pcols, split, blocked, rscale};
    int err = 0;
    if (rscale != NULL)
        err = issue(plasma_kern_split16_scale(uplo, A, rscale, st,
                                              ctx->stream),
                    sequence, request);
    if (!err)
        err = pxpotrf_groups(&g, lookahead, panel, ev, sequence, request);
    /* the whole W_kk for a caller that keeps it (the blocked schedule only
     * formed its diagonal blocks): all tiles at once, off the critical
     * path */
    if (!err && g.blocked && want_w) {
        int slot;
        plasma_prof_begin(ctx->stream, PlasmaProfPotrf, 0, 0, &slot);
        err = issue(plasma_kern_trtri_diag(uplo, A, inv, st, ctx->stream),
                    sequence, request);
        plasma_prof_end(ctx->stream, slot);
    }
    return err;
}

/* Workspace of one factorization: the inverses (when the caller passes
 * none) and the panel buffers (depth + 1 with lookahead depth `depth`).
 * Returns non-zero when out of memory, with nothing left allocated. */
static int pxpotrf_buffers(plasma_desc_t A, void *inv, int depth,
                           void **own_inv, void *panel[])
{
    plasma_context_t *ctx = plasma_context_self();
    const int gw = imax(1, imin(ctx->group, A.mt));
    const int split = use_split(A);
    const size_t pbytes =
        panel_fp_bytes(A, gw * A.nb) + planes_bytes(A, gw * A.nb, split) +
        (split_is_fp16(split) ? plasma_split16_scale_elems(A) * sizeof(float)
                              : 0);
    *own_inv = NULL;
    for (int i = 0; i <= PLASMA_MAX_LOOKAHEAD; i++)
        panel[i] = NULL;
    if (inv == NULL && (*own_inv = plasma_dev_alloc(plasma_inv_bytes(A))) ==
                           NULL)
        return 1;
    for (int i = 0; i <= depth; i++) {
        panel[i] = plasma_dev_alloc(pbytes);
        if (panel[i] == NULL) {
            while (i-- > 0) {
                plasma_dev_free(panel[i]);
                panel[i] = NULL;
            }
            plasma_dev_free(*own_inv);
            *own_inv = NULL;
            return 1;
        }
    }
    return 0;
}

static void pxpotrf_free(void *own_inv, void *panel[])
{
    for (int i = 0; i <= PLASMA_MAX_LOOKAHEAD; i++)
        plasma_dev_free(panel[i]);
    plasma_dev_free(own_inv);
}

static int events_create(hipEvent_t ev[PLASMA_FACTOR_EVENTS])
{
    for (int i = 0; i < PLASMA_FACTOR_EVENTS; i++)
        if (plasma_hip_check(hipEventCreateWithFlags(&ev[i],
                                                     hipEventDisableTiming),
                             "hipEventCreate")) {
            while (i-- > 0)
                hipEventDestroy(ev[i]);
            return 1;
        }
    return 0;
}

static void events_destroy(hipEvent_t ev[PLASMA_FACTOR_EVENTS])
{
    for (int i = 0; i < PLASMA_FACTOR_EVENTS; i++)
        hipEventDestroy(ev[i]);
}

static void graph_drop(plasma_graph_entry_t *e)
{
    if (!e->used)
        return;
    hipGraphExecDestroy(e->exec);
    pxpotrf_free(e->own_inv, e->panel);
    memset(e, 0, sizeof(*e));
}

void plasma_graph_release_all(plasma_context_t *ctx)
{
    hipStreamSynchronize(ctx->panel_stream);
    hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < PLASMA_GRAPHS; i++)
        graph_drop(&ctx->graphs[i]);
}

/*
 * Graph mode (PlasmaGraphCapture enabled): the launch sequence of a
 * factorization is captured once into a hipGraph (the two streams fork
 * and join inside the capture) and replayed on later calls with the same
 * descriptor, inverse workspace and schedule; the graph owns its panel
 * buffers.  Its launches read and write the context's graph status word,
 * which is loaded from the sequence's word before a replay and stored back
 * after it, so one graph serves every sequence.
 */
/* Same descriptor, field by field: plasma_desc_t has padding after its
 * enums that the init functions never write, so a byte compare of two
 * descriptors of the same matrix can differ. */
static int desc_same(const plasma_desc_t *a, const plasma_desc_t *b)
{
    return a->type == b->type && a->uplo == b->uplo &&
           a->precision == b->precision && a->matrix == b->matrix &&
           a->A21 == b->A21 && a->A12 == b->A12 && a->A22 == b->A22 &&
           a->mb == b->mb && a->nb == b->nb && a->gm == b->gm &&
           a->gn == b->gn && a->gmt == b->gmt && a->gnt == b->gnt &&
           a->i == b->i && a->j == b->j && a->m == b->m && a->n == b->n &&
           a->mt == b->mt && a->nt == b->nt && a->kl == b->kl &&
           a->ku == b->ku && a->klt == b->klt && a->kut == b->kut;
}

static plasma_graph_entry_t *graph_get(plasma_enum_t uplo, plasma_desc_t A,
                                       void *inv, int lookahead,
                                       plasma_sequence_t *sequence,
                                       plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    plasma_graph_entry_t *e = NULL, *victim = &ctx->graphs[0];
    for (int i = 0; i < PLASMA_GRAPHS; i++) {
        plasma_graph_entry_t *c = &ctx->graphs[i];
        if (c->used && c->uplo == (int)uplo && c->inv_arg == inv &&
            c->lookahead == lookahead && c->group == ctx->group &&
            c->split == use_split(A) && desc_same(&c->A, &A))
            e = c;
        if (!c->used || (victim->used && c->stamp < victim->stamp))
            victim = c;
    }
    if (e != NULL)
        return e;

    /* capture a new graph into the least recently used entry */
    if (victim->used) {
        hipStreamSynchronize(ctx->panel_stream);
        hipStreamSynchronize(ctx->stream);
        graph_drop(victim);
    }
    e = victim;
    void *own_inv, *panel[PLASMA_MAX_LOOKAHEAD + 1];
    if (pxpotrf_buffers(A, inv, lookahead, &own_inv, panel)) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return NULL;
    }
    hipEvent_t ev[PLASMA_FACTOR_EVENTS];
    if (events_create(ev)) {
        pxpotrf_free(own_inv, panel);
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
        return NULL;
    }
    hipStream_t s0 = ctx->stream;
    int *gword = ctx->d_status + PLASMA_STATUS_SLOTS;
    hipGraph_t graph = NULL;
    int err = plasma_hip_check(
        hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed),
        "hipStreamBeginCapture");
    if (!err) {
        int lerr = pxpotrf_launch(uplo, A, inv != NULL ? inv : own_inv,
                                  inv != NULL, panel, lookahead, gword, ev,
                                  sequence, request);
        err = plasma_hip_check(hipStreamEndCapture(s0, &graph),
                               "hipStreamEndCapture") || lerr;
    }
    if (!err)
        err = plasma_hip_check(hipGraphInstantiate(&e->exec, graph, NULL,
                                                   NULL, 0),
                               "hipGraphInstantiate");
    if (graph != NULL)
        hipGraphDestroy(graph);
    events_destroy(ev);
    if (err) {
        pxpotrf_free(own_inv, panel);
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
        return NULL;
    }
    e->used = 1;
    e->A = A;
    e->uplo = uplo;
    e->inv_arg = inv;
    e->lookahead = lookahead;
    e->group = ctx->group;
    e->split = use_split(A);
    e->own_inv = own_inv;
    for (int i = 0; i <= PLASMA_MAX_LOOKAHEAD; i++)
        e->panel[i] = panel[i];
    return e;
}

void plasma_pxpotrf(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                    plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    if (plasma_is_complex(A.precision)) {
        plasma_pzpotrf(uplo, A, inv, sequence, request);
        return;
    }
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream;
    const int gw = imax(1, imin(ctx->group, A.mt));
    /* lookahead depth: at most the number of groups after the first.  FP32
     * factorizations run at most 2: a tuning run crashed the host process
     * (SIGSEGV) in an FP32 factorization with lookahead 3 (n = 32768,
     * update depth 4, graph mode) and the cause is not found; depth 3 never
     * beat depth 1 in any measurement. */
    const int la_max = A.precision == PlasmaRealFloat ? 2 : ctx->lookahead;
    const int lookahead = imin(imin(ctx->lookahead, la_max),
                               (A.mt + gw - 1) / gw - 1);

    if (ctx->graph_capture == PlasmaEnabled && !ctx->profile) {
        plasma_graph_entry_t *e = graph_get(uplo, A, inv, lookahead,
                                            sequence, request);
        if (e == NULL)
            return;
        e->stamp = ++ctx->graph_clock;
        int *gword = ctx->d_status + PLASMA_STATUS_SLOTS;
        if (issue(plasma_kern_copy_int(gword, st, s0), sequence, request) ||
            plasma_hip_check(hipGraphLaunch(e->exec, s0), "hipGraphLaunch") ||
            issue(plasma_kern_copy_int(st, gword, s0), sequence, request))
            plasma_request_fail(sequence, request, PlasmaErrorInternal);
        return;
    }

    void *own_inv, *panel[PLASMA_MAX_LOOKAHEAD + 1];
    if (pxpotrf_buffers(A, inv, lookahead, &own_inv, panel)) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    hipEvent_t ev[PLASMA_FACTOR_EVENTS];
    if (events_create(ev)) {
        pxpotrf_free(own_inv, panel);
        plasma_request_fail(sequence, request, PlasmaErrorInternal);
        return;
    }
    pxpotrf_launch(uplo, A, inv != NULL ? inv : own_inv, inv != NULL, panel,
                   lookahead, st, ev, sequence, request);
    events_destroy(ev);
    pxpotrf_free(own_inv, panel);
}

/*
 * Triangular sweeps (reference compute/pztrsm.c, Left side).  In the
 * lower-equivalent convention the four (uplo, trans) cases used by potrs
 * reduce to a forward sweep with L or a backward sweep with L^T.
 */
void plasma_pxtrsm_left(plasma_enum_t uplo, plasma_enum_t trans,
                        plasma_desc_t A, plasma_desc_t B, const void *inv,
                        plasma_sequence_t *sequence,
                        plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream;
    /* lower+NoTrans and upper+Trans solve with L; the others with L^T */
    const int backward = (uplo == PlasmaLower) ? (trans != PlasmaNoTrans)
                                               : (trans == PlasmaNoTrans);
    const int mt = B.mt;
    size_t es = plasma_element_size(B.precision);
    void *scratch = plasma_dev_alloc((size_t)A.mb * B.nt * B.nb * es);
    if (scratch == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    for (int i = 0; i < mt; i++) {
        const int k = backward ? mt - 1 - i : i;
        int slot;
        double nbk = plasma_tile_mview(B, k);
        int nupd = backward ? k : mt - 1 - k;
        plasma_prof_begin(s0, PlasmaProfSweep,
                          nbk * nbk * B.n + 2.0 * nbk * nupd * B.mb * B.n, 0,
                          &slot);
        int err = issue(plasma_kern_trsm_diag_left(uplo, backward, A, B, k,
                                                   inv, scratch, st, s0),
                        sequence, request);
        if (!err)
            err = issue(plasma_kern_sweep_update(uplo, backward, A, B, k,
                                                 scratch, st, s0),
                        sequence, request);
        plasma_prof_end(s0, slot);
        if (err)
            break;
    }
    plasma_dev_free(scratch);
}

void plasma_pxpotrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                    const void *inv, plasma_sequence_t *sequence,
                    plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    if (plasma_is_complex(A.precision)) {
        plasma_pzpotrs(uplo, A, B, inv, sequence, request);
        return;
    }
    void *own_inv = NULL;
    if (inv == NULL) {
        plasma_context_t *ctx = plasma_context_self();
        int *st = plasma_sequence_dstatus(sequence, request);
        own_inv = plasma_dev_alloc(plasma_inv_bytes(A));
        if (own_inv == NULL) {
            plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
            return;
        }
        if (issue(plasma_kern_trtri_diag(uplo, A, own_inv, st, ctx->stream),
                  sequence, request)) {
            plasma_dev_free(own_inv);
            return;
        }
        inv = own_inv;
    }
    /* reference compute/zpotrs.c:270-283 */
    plasma_pxtrsm_left(uplo, uplo == PlasmaUpper ? PlasmaConjTrans
                                                 : PlasmaNoTrans,
                       A, B, inv, sequence, request);
    plasma_pxtrsm_left(uplo, uplo == PlasmaUpper ? PlasmaNoTrans
                                                 : PlasmaConjTrans,
                       A, B, inv, sequence, request);
    plasma_dev_free(own_inv);
}

/*
 * Repeated solves with one factor (the refinement loop of dsposv).  With
 * L = M D, M(m,k) = L(m,k) W_kk (m > k), M(k,k) = I and D = diag(L_kk):
 *     L Y = B      <=>  Z = M^{-1} B,  Y = D^{-1} Z
 *     L^T X = Y    <=>  M^T X = D^{-T} Y = W^T W Z
 * so each sweep step is one launch (B(k,:) is final when step k runs)
 * instead of a diagonal solve plus an update; the two diagonal products
 * run over all block rows at once.  The reference's pztrsm sweeps
 * (compute/pztrsm.c:113-177) compute the same solution.
 */
size_t plasma_mpanel_bytes(plasma_desc_t A)
{
    return (size_t)A.mt * (A.mt - 1) / 2 * A.mb * A.nb *
           plasma_element_size(A.precision);
}

void plasma_pxscale_panels(plasma_enum_t uplo, plasma_desc_t A,
                           const void *inv, void *M,
                           plasma_sequence_t *sequence,
                           plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    /* 2 mv nb^2 flops per strictly lower tile (M(m,k) = L(m,k) W_kk) */
    double flops = 0;
    for (int k = 0; k + 1 < A.mt; k++)
        flops += 2.0 * A.nb * A.nb * (A.m - (k + 1) * A.mb);
    int slot;
    plasma_prof_begin(ctx->stream, PlasmaProfSweep, flops, 0, &slot);
    issue(plasma_kern_scale_panels(uplo, A, inv, M, st, ctx->stream),
          sequence, request);
    plasma_prof_end(ctx->stream, slot);
}

void plasma_pxpotrs_m(plasma_desc_t A, plasma_desc_t B, const void *inv,
                      const void *M, plasma_sequence_t *sequence,
                      plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream;
    const int mt = B.mt;
    size_t es = plasma_element_size(B.precision);
    void *scratch = plasma_dev_alloc((size_t)mt * A.mb * B.nt * B.nb * es);
    if (scratch == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    int err = 0, slot;
    const double sweep_flops = 2.0 * B.m * B.m * B.n;  /* both sweeps */
    plasma_prof_begin(s0, PlasmaProfSweep, sweep_flops, 0, &slot);
    for (int k = 0; k + 1 < mt && !err; k++)
        err = issue(plasma_kern_msweep_update(0, A, B, k, M, st, s0),
                    sequence, request);
    if (!err)
        err = issue(plasma_kern_diag_normal(A, B, inv, scratch, st, s0),
                    sequence, request);
    for (int k = mt - 1; k > 0 && !err; k--)
        err = issue(plasma_kern_msweep_update(1, A, B, k, M, st, s0),
                    sequence, request);
    plasma_prof_end(s0, slot);
    plasma_dev_free(scratch);
}

/*
 * Super-block sweeps (kernels/ssweep.hip).  A super-block is S tiles, about
 * 2048 rows: large enough that a sweep takes 2 mt / S launches of S times
 * the work of a tile step, small enough that building V (the inverses of
 * the diagonal super-blocks, S^2 nb^2 elements each) stays cheap.  dsposv
 * at n = 32768, cond 1e6 (four solves): 512 rows 111.6 ms, 1024 rows
 * 106.8 ms, 2048 rows 105.0 ms; the per-tile sweeps 121.0 ms.
 * PLASMA_SWEEP_ROWS sets the rows per super-block (0 keeps the per-tile
 * sweeps).
 */
int plasma_sweep_tiles(plasma_desc_t A)
{
    static int rows = -1;
    if (rows < 0) {
        const char *e = getenv("PLASMA_SWEEP_ROWS");
        rows = e != NULL ? atoi(e) : 2048;
    }
    if (A.mb != A.nb || A.mb % 64 != 0 || rows < 2 * A.mb)
        return 0;
    const int S = imin(rows / A.mb, A.mt);
    return S >= 2 ? S : 0;
}

size_t plasma_sbinv_bytes(plasma_desc_t A)
{
    const int S = plasma_sweep_tiles(A);
    if (S == 0)
        return 0;
    const size_t q = (size_t)(A.mt + S - 1) / S, rows = (size_t)S * A.mb;
    return q * rows * rows * plasma_element_size(A.precision);
}

void plasma_pxsbinv(plasma_enum_t uplo, plasma_desc_t A, const void *inv,
                    void *V, plasma_sequence_t *sequence,
                    plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    const int S = plasma_sweep_tiles(A);
    if (S == 0) {
        plasma_request_fail(sequence, request, PlasmaErrorNotSupported);
        return;
    }
    int *st = plasma_sequence_dstatus(sequence, request);
    void *Tscr = plasma_dev_alloc(plasma_sbinv_scratch_elems(A, S) *
                                  plasma_element_size(A.precision));
    if (Tscr == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    /* recursive doubling: at block size h, 2 (h nb)^3 flops per product,
     * two products per pair of blocks */
    double flops = 0;
    for (int h = 1; h < S; h *= 2) {
        const double hb = (double)h * A.mb;
        flops += 4.0 * hb * hb * hb * ((A.mt + 2 * h - 1) / (2 * h));
    }
    int slot;
    plasma_prof_begin(ctx->stream, PlasmaProfSweep, flops, 0, &slot);
    issue(plasma_kern_sbinv(uplo, A, inv, V, Tscr, S, st, ctx->stream),
          sequence, request);
    plasma_prof_end(ctx->stream, slot);
    plasma_dev_free(Tscr);
}

void plasma_pxpotrs_sb(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                       const void *V, plasma_sequence_t *sequence,
                       plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    const int S = plasma_sweep_tiles(A);
    if (S == 0 || B.mb != A.mb) {
        plasma_request_fail(sequence, request, PlasmaErrorNotSupported);
        return;
    }
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream;
    const int q = (A.mt + S - 1) / S;
    void *Y = plasma_dev_alloc((size_t)S * A.mb * B.nt * B.nb *
                               plasma_element_size(B.precision));
    if (Y == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    int err = 0, slot;
    plasma_prof_begin(s0, PlasmaProfSweep, 2.0 * B.m * B.m * B.n, 0, &slot);
    for (int K = 0; K < q && !err; K++)
        err = issue(plasma_kern_ssweep(uplo, 0, A, B, V, Y, S, K, st, s0),
                    sequence, request);
    for (int K = q - 1; K >= 0 && !err; K--)
        err = issue(plasma_kern_ssweep(uplo, 1, A, B, V, Y, S, K, st, s0),
                    sequence, request);
    plasma_prof_end(s0, slot);
    plasma_dev_free(Y);
}

/* The repeated-solve form of a factor: the super-block inverses when those
 * sweeps apply, else the scaled panels; NULL when the block could not be
 * allocated (plasma_rsolve then runs the plain sweeps). */
void *plasma_rsolve_prepare(plasma_enum_t uplo, plasma_desc_t A,
                            const void *inv, plasma_sequence_t *sequence,
                            plasma_request_t *request)
{
    /* complex factors are solved by the plain sweeps */
    if (plasma_is_complex(A.precision))
        return NULL;
    const int sb = plasma_sweep_tiles(A) > 0;
    void *form = plasma_dev_alloc(sb ? plasma_sbinv_bytes(A)
                                     : plasma_mpanel_bytes(A));
    if (form == NULL)
        return NULL;
    if (sb)
        plasma_pxsbinv(uplo, A, inv, form, sequence, request);
    else
        plasma_pxscale_panels(uplo, A, inv, form, sequence, request);
    return form;
}

void plasma_rsolve(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                   const void *inv, const void *form,
                   plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (form == NULL)
        plasma_pxpotrs(uplo, A, B, inv, sequence, request);
    else if (plasma_sweep_tiles(A) > 0)
        plasma_pxpotrs_sb(uplo, A, B, form, sequence, request);
    else
        plasma_pxpotrs_m(A, B, inv, form, sequence, request);
}

/*
 * Band Cholesky (reference compute/pzpbtrf.c:37-70 lower, :72-117 upper).
 * Column by column on a band descriptor: the diagonal tile k (with W_kk),
 * its panel (the band's tiles below it, into a panel buffer) and the update
 * of the band's window (k, k + bt) with the panel copied back, where bt is
 * the number of tile rows of a band column (diagonal included).  Every
 * launch takes a copy of the descriptor cut to that window (mt = nt =
 * min(mt, k + bt)), so the tile loops of the dense kernels never leave the
 * band: the tiles outside it are not stored.  The factor keeps the band's
 * zeros exactly (a zero outside the band is only ever multiplied into
 * another zero's place).
 */
static plasma_desc_t band_window(plasma_desc_t A, int t_end)
{
    plasma_desc_t W = A;
    if (t_end < A.mt) {
        W.mt = W.nt = t_end;
        W.m = W.n = t_end * A.mb;
    }
    return W;
}

int plasma_band_tiles(plasma_enum_t uplo, plasma_desc_t A)
{
    return uplo == PlasmaUpper ? A.kut : A.klt;
}

void plasma_pxpbtrf(plasma_enum_t uplo, plasma_desc_t A, void *inv,
                    plasma_sequence_t *sequence, plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream;
    const int bt = plasma_band_tiles(uplo, A);
    const size_t es = plasma_element_size(A.precision);
    void *panel = plasma_dev_alloc((size_t)imax(bt - 1, 1) * A.mb * A.nb *
                                   es);
    void *own_inv = inv == NULL ? plasma_dev_alloc(plasma_inv_bytes(A))
                                : NULL;
    if (panel == NULL || (inv == NULL && own_inv == NULL)) {
        plasma_dev_free(panel);
        plasma_dev_free(own_inv);
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    for (int k = 0; k < A.mt; k++) {
        const int t_end = imin(A.mt, k + bt);
        const group_t g = {uplo, band_window(A, t_end),
                           inv != NULL ? inv : own_inv, st, A.nb,
                           SPLIT_NONE, 0, NULL};
        if (factor_group(&g, k, 1, panel, s0, sequence, request) ||
            issue(k_update(&g, k, 1, k + 1, t_end, panel, 1, 0, s0),
                  sequence, request))
            break;
    }
    plasma_dev_free(panel);
    plasma_dev_free(own_inv);
}

/* B <- A^{-1} B with the band factor (reference compute/pztbsm.c, the two
 * sweeps of zpbtrs.c:280-294): per tile the diagonal product with W_kk and
 * the update of the band's block rows only. */
void plasma_pxpbtrs(plasma_enum_t uplo, plasma_desc_t A, plasma_desc_t B,
                    const void *inv, plasma_sequence_t *sequence,
                    plasma_request_t *request)
{
    if (sequence->status != PlasmaSuccess)
        return;
    plasma_context_t *ctx = plasma_context_self();
    int *st = plasma_sequence_dstatus(sequence, request);
    hipStream_t s0 = ctx->stream;
    const int bt = plasma_band_tiles(uplo, A);
    const size_t es = plasma_element_size(B.precision);
    void *scratch = plasma_dev_alloc((size_t)A.mb * B.nt * B.nb * es);
    if (scratch == NULL) {
        plasma_request_fail(sequence, request, PlasmaErrorOutOfMemory);
        return;
    }
    int err = 0, slot;
    plasma_prof_begin(s0, PlasmaProfSweep,
                      4.0 * B.n * (double)A.m * imin(A.m, bt * A.mb), 0,
                      &slot);
    for (int k = 0; k < B.mt && !err; k++)
        err = issue(plasma_kern_trsm_diag_left(uplo, 0, A, B, k, inv,
                                               scratch, st, s0),
                    sequence, request) ||
              issue(plasma_kern_sweep_update_range(uplo, 0, A, B, k, 0,
                                                   k + bt, scratch, st, s0),
                    sequence, request);
    for (int k = B.mt - 1; k >= 0 && !err; k--)
        err = issue(plasma_kern_trsm_diag_left(uplo, 1, A, B, k, inv,
                                               scratch, st, s0),
                    sequence, request) ||
              issue(plasma_kern_sweep_update_range(uplo, 1, A, B, k,
                                                   k - bt + 1, k, scratch,
                                                   st, s0),
                    sequence, request);
    plasma_prof_end(s0, slot);
    plasma_dev_free(scratch);
}
