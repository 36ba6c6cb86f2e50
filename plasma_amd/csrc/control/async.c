/*
 * async.c -- sequences, requests and their device status words.
 *
 * plasma_request_fail / plasma_request_init / plasma_sequence_init keep the
 * reference's behaviour (control/async.c:14-35).  Because the work of a
 * sequence runs on the GPU, a numerical failure is first recorded by a
 * kernel in the sequence's device status word; plasma_sequence_sync folds
 * it into sequence->status and frees the word.  Words are assigned to
 * sequences on first use from a table of PLASMA_STATUS_SLOTS entries.  When
 * every entry belongs to a sequence that was not synchronized yet, the
 * device is drained and each failure found is folded into its sequence
 * before the table is reused, so no sequence loses its info.  A sequence
 * with work in flight must therefore stay a live object until it is
 * synchronized (the reference's sequences live until their parallel region
 * ends).
 */
#include <string.h>

#include "internal.h"

int plasma_request_fail(plasma_sequence_t *sequence,
                        plasma_request_t *request, int status)
{
    if (sequence != NULL) {
        sequence->request = request;
        sequence->status = status;
    }
    if (request != NULL)
        request->status = status;
    return status;
}

int plasma_request_init(plasma_request_t *request)
{
    request->status = PlasmaSuccess;
    return PlasmaSuccess;
}

static int find_slot(plasma_context_t *ctx, const plasma_sequence_t *seq)
{
    for (int i = 0; i < PLASMA_STATUS_SLOTS; i++)
        if (ctx->slot_seq[i] == seq)
            return i;
    return -1;
}

int plasma_sequence_init(plasma_sequence_t *sequence)
{
    sequence->status = PlasmaSuccess;
    sequence->request = NULL;
    plasma_context_t *ctx = plasma_context_self();
    if (ctx != NULL) {
        int slot = find_slot(ctx, sequence);
        if (slot >= 0) {
            /* a previous sequence at this address: clear its word, in stream
             * order after any work still using it */
            plasma_kern_fill_int(ctx->d_status + slot, 0, 1, ctx->stream);
            ctx->slot_req[slot] = NULL;
        }
    }
    return PlasmaSuccess;
}

/* Every slot is taken by a sequence with work in flight (PLASMA_STATUS_SLOTS
 * sequences that were never synchronized): drain the device, fold each
 * non-zero word into its sequence (and request), then free every slot.
 * The sequences stay valid objects of their callers until they are
 * synchronized, so their status can be written here. */
static void recycle_slots(plasma_context_t *ctx)
{
    int words[PLASMA_STATUS_SLOTS];
    hipStreamSynchronize(ctx->panel_stream);
    hipStreamSynchronize(ctx->stream);
    if (plasma_hip_check(hipMemcpy(words, ctx->d_status, sizeof(words),
                                   hipMemcpyDeviceToHost),
                         "hipMemcpy(status words)"))
        for (int i = 0; i < PLASMA_STATUS_SLOTS; i++)
            words[i] = PlasmaErrorInternal;
    for (int i = 0; i < PLASMA_STATUS_SLOTS; i++) {
        plasma_sequence_t *seq = ctx->slot_seq[i];
        if (seq != NULL && words[i] != 0 && seq->status == PlasmaSuccess)
            plasma_request_fail(seq, ctx->slot_req[i], words[i]);
        ctx->slot_seq[i] = NULL;
        ctx->slot_req[i] = NULL;
    }
    ctx->slot_next = 0;
}

int *plasma_sequence_dstatus(plasma_sequence_t *sequence,
                             plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    int slot = find_slot(ctx, sequence);
    if (slot < 0) {
        /* the next free slot, searched round robin from slot_next */
        for (int i = 0; i < PLASMA_STATUS_SLOTS && slot < 0; i++) {
            const int c = (ctx->slot_next + i) % PLASMA_STATUS_SLOTS;
            if (ctx->slot_seq[c] == NULL)
                slot = c;
        }
        if (slot < 0) {
            recycle_slots(ctx);
            slot = 0;
        }
        ctx->slot_next = (slot + 1) % PLASMA_STATUS_SLOTS;
        ctx->slot_seq[slot] = sequence;
        ctx->slot_req[slot] = NULL;
        /* cleared in stream order after any earlier user of the word */
        plasma_kern_fill_int(ctx->d_status + slot, 0, 1, ctx->stream);
    }
    if (request != NULL)
        ctx->slot_req[slot] = request;
    return ctx->d_status + slot;
}

int plasma_sequence_sync(plasma_sequence_t *sequence)
{
    plasma_context_t *ctx = plasma_context_self();
    if (ctx == NULL || sequence == NULL)
        return PlasmaErrorNotInitialized;
    int e1 = plasma_hip_check(hipStreamSynchronize(ctx->panel_stream),
                              "hipStreamSynchronize");
    int e2 = plasma_hip_check(hipStreamSynchronize(ctx->stream),
                              "hipStreamSynchronize");
    if (e1 || e2) {
        if (sequence->status == PlasmaSuccess)
            sequence->status = PlasmaErrorInternal;
        return sequence->status;
    }
    int slot = find_slot(ctx, sequence);
    if (slot >= 0) {
        if (sequence->status == PlasmaSuccess) {
            int word = 0;
            if (plasma_hip_check(hipMemcpy(&word, ctx->d_status + slot,
                                           sizeof(int),
                                           hipMemcpyDeviceToHost),
                                 "hipMemcpy(status)"))
                word = PlasmaErrorInternal;
            if (word != 0)
                plasma_request_fail(sequence, ctx->slot_req[slot], word);
        }
        /* no work of the sequence is in flight any more: its word is free
         * (later work of the same sequence takes a fresh one) */
        ctx->slot_seq[slot] = NULL;
        ctx->slot_req[slot] = NULL;
    }
    return sequence->status;
}
