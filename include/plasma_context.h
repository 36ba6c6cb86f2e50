/*
 * plasma_context.h -- library initialisation and run-time parameters.
 *
 * Same entry points and return conventions as the reference
 * (include/plasma_context.h:51-54, control/context.c:30-159):
 *   plasma_init      PlasmaSuccess, or PlasmaErrorNotInitialized when the
 *                    library is already initialised;
 *   plasma_finalize  PlasmaSuccess, or PlasmaErrorNotInitialized;
 *   plasma_set/get   PlasmaTuning, PlasmaNb (default 256), PlasmaIb (64),
 *                    PlasmaNumPanelThreads, PlasmaHouseholderMode, plus the
 *                    extensions PlasmaLookahead, PlasmaGraphCapture and
 *                    PlasmaDevice (plasma_types.h).
 * plasma_init binds the library to the current HIP device (or the ordinal
 * in the environment variable PLASMA_DEVICE), creates its streams and
 * device status words, and returns PlasmaErrorEnvironment when no GPU is
 * usable.  The context itself is private to the library.
 */
#ifndef PLASMA_CONTEXT_H
#define PLASMA_CONTEXT_H

#include "plasma_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct plasma_context_s plasma_context_t;

int plasma_init(void);
int plasma_finalize(void);
int plasma_set(plasma_enum_t param, int value);
int plasma_get(plasma_enum_t param, int *value);

/* The global context, or NULL before plasma_init (reference :196-202). */
plasma_context_t *plasma_context_self(void);

/* PLASMA version triple (reference control/version.c). */
void plasma_version(int *major, int *minor, int *micro);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_CONTEXT_H */
