/*
 * plasma_mi355x.h -- extensions of the MI355X implementation that have no
 * counterpart in the reference.
 */
#ifndef PLASMA_MI355X_H
#define PLASMA_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel classes timed by the profiler. */
enum {
    PlasmaProfUpdate = 0,   /* MFMA trailing update (syrk + gemm tiles)   */
    PlasmaProfPotrf  = 1,   /* diagonal tile factorization                */
    PlasmaProfPanel  = 2,   /* panel triangular solve                     */
    PlasmaProfSweep  = 3,   /* potrs sweeps (diagonal solve + update)     */
    PlasmaProfLayout = 4,   /* LAPACK <-> tile translation                */
    PlasmaProfMixed  = 5,   /* conversions, residual, norms of dsposv     */
    PlasmaProfAdjoint = 6,  /* Cholesky / potrs adjoint products          */
    PlasmaProfClasses = 7
};

/*
 * Kernel profiler.  When enabled, every launch of the classes above is
 * bracketed by HIP events recorded on the stream it runs on, together with
 * its algorithmic flop and byte counts.  plasma_profile_query synchronizes,
 * folds the recorded launches into per-class totals and returns them.
 * Enabling clears the totals.  Intended for benchmarks: the events add a
 * few microseconds per launch.
 */
int plasma_profile_enable(int on);
int plasma_profile_query(int kernel_class, double *milliseconds,
                         double *flops, double *bytes, long *launches);

/*
 * Autotuned schedules (no counterpart in the reference's C interface; it
 * replaces the Lua tables behind plasma_tune_potrf, control/tuning.c:
 * 367-374).  Replaces the table read at plasma_init with the file at path
 * (NULL: no table).  Lines "potrf <d|s> <n_max> <nb> <update_depth>
 * <lookahead>"; with PlasmaTuning enabled a factorization of order n takes
 * nb, update depth and lookahead from the first row of its precision with
 * n <= n_max (else the last one).  Returns PlasmaErrorIllegalValue when the
 * file cannot be read.
 */
int plasma_tuning_load(const char *path);

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_MI355X_H */
