/*
 * Error reporting shared by the translation units of libpumas_hip: the
 * message of the last error, per host thread, as the reference's error
 * handler message (src/pumas.c:871-880), returned with its code.
 */
#ifndef PUMAS_STATUS_H
#define PUMAS_STATUS_H

#include "../../include/pumas_gpu.h"

/* Records a formatted message for pumas_gpu_error_message and returns rc. */
enum pumas_return pg_fail(enum pumas_return rc, const char * fmt, ...);

/* Clears the message at the start of an API call. */
void pg_clear_error(void);

/* Device bytes per history that the transport scheduler allocates (history
 * record and queue slots). */
size_t pg_bytes_per_history(void);

#define HIP_TRY(expr)                                                          \
        do {                                                                   \
                hipError_t e_ = (expr);                                        \
                if (e_ != hipSuccess)                                          \
                        return pg_fail(PUMAS_RETURN_INTERNAL_ERROR,            \
                            "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                            __FILE__, __LINE__);                               \
        } while (0)

#endif
