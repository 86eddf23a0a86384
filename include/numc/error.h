/* Error reporting: every call returns 0 or a negative NumcErrorCode and
 * records a per-thread message (numc's include/numc/error.h:10-57). */
#ifndef NUMC_ERROR_H
#define NUMC_ERROR_H

#include "numc/export.h"
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  NUMC_ERR_NONE = 0,
  NUMC_ERR_NULL = -1,     /* NULL pointer where not allowed */
  NUMC_ERR_MALLOC = -2,   /* host or device allocation failed */
  NUMC_ERR_SHAPE = -3,    /* shape / dimension mismatch */
  NUMC_ERR_TYPE = -4,     /* dtype mismatch or unsupported dtype */
  NUMC_ERR_BOUNDS = -5,   /* index out of bounds */
  NUMC_ERR_VALUE = -6,    /* invalid argument value */
  NUMC_ERR_INTERNAL = -99 /* device / runtime failure */
} NumcErrorCode;

typedef struct {
  int code;
  const char *msg;
} NumcError;

NUMC_API int numc_set_error_v(int code, const char *func, const char *file,
                              int line, const char *fmt, ...);
NUMC_API int numc_set_error(int code, const char *msg);
NUMC_API NumcError numc_get_error(void);
NUMC_API void numc_log_error(const NumcError *err);

#define NUMC_SET_ERROR(code, ...) \
  numc_set_error_v(code, __func__, __FILE__, __LINE__, __VA_ARGS__)

#ifdef __cplusplus
}
#endif

#endif
