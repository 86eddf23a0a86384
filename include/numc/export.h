/* Symbol visibility for the numc C ABI (same macro as numc's
 * include/numc/export.h: default visibility on ELF, hidden otherwise). */
#ifndef NUMC_EXPORT_H
#define NUMC_EXPORT_H
#ifndef NUMC_API
#if defined(__GNUC__) && __GNUC__ >= 4
#define NUMC_API __attribute__((visibility("default")))
#else
#define NUMC_API
#endif
#endif
#endif
