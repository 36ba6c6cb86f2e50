/*
 * async.c -- sequences, requests and their device status words.
 *
 * plasma_request_fail / plasma_request_init / plasma_sequence_init keep the
 * reference's behaviour (control/async.c:14-35).  Because the work of a
 * sequence runs on the GPU, a numerical failure is first recorded by a
 * kernel in the sequence's device status word; plasma_sequence_sync folds
 * it into sequence->status.  Words are assigned to sequence addresses on
 * first use from a table of PLASMA_STATUS_SLOTS entries.
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

int *plasma_sequence_dstatus(plasma_sequence_t *sequence,
                             plasma_request_t *request)
{
    plasma_context_t *ctx = plasma_context_self();
    int slot = find_slot(ctx, sequence);
    if (slot < 0) {
        if (ctx->slot_next >= PLASMA_STATUS_SLOTS) {
            /* table full: drain the device, then recycle every word */
            hipStreamSynchronize(ctx->panel_stream);
            hipStreamSynchronize(ctx->stream);
            memset(ctx->slot_seq, 0, sizeof(ctx->slot_seq));
            memset(ctx->slot_req, 0, sizeof(ctx->slot_req));
            ctx->slot_next = 0;
        }
        slot = ctx->slot_next++;
        ctx->slot_seq[slot] = sequence;
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
    if (slot >= 0 && sequence->status == PlasmaSuccess) {
        int word = 0;
        if (plasma_hip_check(hipMemcpy(&word, ctx->d_status + slot,
                                       sizeof(int), hipMemcpyDeviceToHost),
                             "hipMemcpy(status)"))
            word = PlasmaErrorInternal;
        if (word != 0)
            plasma_request_fail(sequence, ctx->slot_req[slot], word);
    }
    return sequence->status;
}
