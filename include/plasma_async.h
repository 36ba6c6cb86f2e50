/*
 * plasma_async.h -- sequence / request objects of the tile asynchronous API.
 *
 * The two structs keep the reference's layout (include/plasma_async.h:20-27)
 * because callers allocate them.  Semantics on MI355X: work submitted by a
 * plasma_omp_* call is enqueued on the library's HIP streams and the call
 * returns immediately.  Numerical failures (a non positive definite minor,
 * an FP64->FP32 overflow) are recorded by the kernels in a device status
 * word that belongs to the sequence; plasma_sequence_sync() waits for all
 * work of the sequence and copies that word into sequence->status.
 * Argument errors are detected on the host and set sequence->status at once.
 */
#ifndef PLASMA_ASYNC_H
#define PLASMA_ASYNC_H

#include "plasma_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    plasma_enum_t status;            /* error code */
} plasma_request_t;

typedef struct {
    plasma_enum_t status;            /* error code */
    plasma_request_t *request;       /* failed request */
} plasma_sequence_t;

/* reference control/async.c:14-20 */
int plasma_request_fail(plasma_sequence_t *sequence,
                        plasma_request_t *request, int status);
/* reference control/async.c:23-27 */
int plasma_request_init(plasma_request_t *request);
/* reference control/async.c:30-35; also clears the sequence's device word */
int plasma_sequence_init(plasma_sequence_t *sequence);

/*
 * Extension (no reference counterpart; the reference completes at the end
 * of its omp parallel region): block until every operation enqueued for
 * this sequence has finished, then fold the device status word into
 * sequence->status.  Returns the final sequence->status.
 */
int plasma_sequence_sync(plasma_sequence_t *sequence);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_ASYNC_H */
