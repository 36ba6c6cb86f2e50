/*
 * tuning.c -- the autotuned schedule table.
 *
 * The reference picks the tile size of a factorization from Lua tables when
 * tuning is enabled (plasma_tune_potrf, control/tuning.c:367-374, called by
 * compute/zpotrf.c:105-106, zposv.c:130-131 and, in single precision,
 * zcposv.c:160-161), and leaves the context untouched when the tables are
 * absent.  Here the table is plain text, written by the tuner
 * (plasma_amd/tune.py) from measurements on the GPU (precision d, s, z or
 * c; the complex factorization reads only nb):
 *
 *     # routine precision n_max nb update_depth lookahead [mode [split]]
 *     potrf d 8192 256 2 1 graph
 *     potrf s 32768 256 4 2 graph fp16
 *
 * The optional mode, "eager", "graph" or "any", says which launch mode the
 * row was measured in (PlasmaGraphCapture disabled or enabled); "any" or no
 * mode serves both.  An FP32 row may also name the operands of its trailing
 * updates, the MFMA form tuned beside the schedule: "fp32" (FP32 MFMA),
 * "bf16" (three-term bf16 split, six products) or "fp16" (row-scaled
 * two-term fp16, three products).  A factorization of order n in precision
 * p uses, among the rows of p for the context's mode (those of the other
 * mode when p has none), the first with n <= n_max (rows sorted by n_max),
 * or the last.  Like the reference, a tuned call leaves its choice in the
 * context (nb, update depth, lookahead, and the split when the row names
 * one).  The table is read at plasma_init from
 * $PLASMA_TUNING_FILE, else from tuning/mi355x.txt next to the library's
 * directory; plasma_tuning_load() replaces it.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <libgen.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static int row_cmp(const void *a, const void *b)
{
    const plasma_tune_row_t *x = a, *y = b;
    if (x->precision != y->precision)
        return x->precision - y->precision;
    return x->n_max - y->n_max;
}

int plasma_tuning_load(const char *path)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL)
        return PlasmaErrorNotInitialized;
    ctx->ntune = 0;
    if (path == NULL)
        return PlasmaSuccess;
    FILE *f = fopen(path, "r");
    if (f == NULL)
        return PlasmaErrorIllegalValue;
    char line[256];
    while (fgets(line, sizeof(line), f) != NULL &&
           ctx->ntune < PLASMA_TUNE_ROWS) {
        char routine[32], prec, mode[16] = "", split[16] = "";
        int n_max, nb, depth, la;
        const int got = sscanf(line, "%31s %c %d %d %d %d %15s %15s",
                               routine, &prec, &n_max, &nb, &depth, &la, mode,
                               split);
        if (line[0] == '#' || got < 6 || strcmp(routine, "potrf") != 0)
            continue;
        if ((prec != 'd' && prec != 's' && prec != 'z' && prec != 'c') ||
            n_max < 1 || nb < 1 ||
            depth < 1 || depth > 4 || la < 0 || la > 4)
            continue;
        int m = 0;
        if (got >= 7 && strcmp(mode, "eager") == 0)
            m = 1;
        else if (got >= 7 && strcmp(mode, "graph") == 0)
            m = 2;
        else if (got >= 7 && strcmp(mode, "any") != 0)
            continue;
        /* FP32 rows may name the trailing update's operand format */
        int sp = 0;
        if (got == 8) {
            if (strcmp(split, "fp32") == 0)
                sp = -1;
            else if (strcmp(split, "bf16") == 0)
                sp = PlasmaSplitBf16;
            else if (strcmp(split, "fp16") == 0)
                sp = PlasmaSplitFp16;
            else
                continue;
        }
        plasma_tune_row_t *r = &ctx->tune[ctx->ntune++];
        r->precision = prec == 'd'   ? PlasmaRealDouble
                       : prec == 's' ? PlasmaRealFloat
                       : prec == 'z' ? PlasmaComplexDouble
                                     : PlasmaComplexFloat;
        r->n_max = n_max;
        r->nb = nb;
        r->depth = depth;
        r->lookahead = la;
        r->mode = m;
        r->split = sp;
    }
    fclose(f);
    qsort(ctx->tune, ctx->ntune, sizeof(ctx->tune[0]), row_cmp);
    return PlasmaSuccess;
}

void plasma_tuning_init(plasma_context_t *ctx)
{
    (void)ctx;
    const char *env = getenv("PLASMA_TUNING_FILE");
    if (env != NULL) {
        plasma_tuning_load(env);
        return;
    }
    Dl_info info;
    if (dladdr((void *)plasma_tuning_init, &info) == 0 ||
        info.dli_fname == NULL)
        return;
    char lib[1024], path[1200];
    snprintf(lib, sizeof(lib), "%s", info.dli_fname);
    snprintf(path, sizeof(path), "%s/../tuning/mi355x.txt", dirname(lib));
    FILE *f = fopen(path, "r");
    if (f == NULL)
        return;
    fclose(f);
    plasma_tuning_load(path);
}

void plasma_tune_potrf(plasma_context_t *ctx, plasma_enum_t precision, int n)
{
    const int mode = ctx->graph_capture == PlasmaEnabled ? 2 : 1;
    const plasma_tune_row_t *pick = NULL;
    for (int any = 0; any < 2 && pick == NULL; any++)
        for (int i = 0; i < ctx->ntune; i++) {
            const plasma_tune_row_t *r = &ctx->tune[i];
            if (r->precision != precision ||
                (!any && r->mode != 0 && r->mode != mode))
                continue;
            pick = r;
            if (n <= r->n_max)
                break;
        }
    if (pick == NULL)
        return;
    ctx->nb = pick->nb;
    ctx->group = pick->depth;
    ctx->lookahead = pick->lookahead;
    if (pick->split != 0)
        ctx->split_update = pick->split < 0 ? PlasmaDisabled : pick->split;
}
