/*
 * plasma_error.h -- diagnostics on stderr.
 *
 * Same observable behaviour as the reference (include/plasma_error.h:21-69):
 * warnings and errors print "PLASMA <KIND> at <line> of <func>() in <file>:
 * <msg>" to stderr, and a fatal error terminates the process with
 * EXIT_FAILURE (used when a routine is called before plasma_init).
 */
#ifndef PLASMA_ERROR_H
#define PLASMA_ERROR_H

#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

static inline void plasma_report_(const char *kind, const char *func,
                                  int line, const char *file,
                                  const char *msg)
{
    fprintf(stderr, "PLASMA %s at %d of %s() in %s: %s\n",
            kind, line, func, file, msg);
}

static inline void plasma_report_code_(const char *func, int line,
                                       const char *file, const char *msg,
                                       int code)
{
    fprintf(stderr, "PLASMA ERROR at %d of %s() in %s: %s %d\n",
            line, func, file, msg, code);
}

static inline void plasma_fatal_(const char *func, int line,
                                 const char *file, const char *msg)
{
    plasma_report_("FATAL ERROR", func, line, file, msg);
    exit(EXIT_FAILURE);
}

#define plasma_warning(msg) \
    plasma_report_("WARNING", __func__, __LINE__, __FILE__, msg)
#define plasma_error(msg) \
    plasma_report_("ERROR", __func__, __LINE__, __FILE__, msg)
#define plasma_error_with_code(msg, code) \
    plasma_report_code_(__func__, __LINE__, __FILE__, msg, code)
#define plasma_fatal_error(msg) \
    plasma_fatal_(__func__, __LINE__, __FILE__, msg)

#ifdef __cplusplus
}
#endif

#endif /* PLASMA_ERROR_H */
