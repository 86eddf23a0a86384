/* Umbrella header (numc's include/numc/numc.h) plus the MI355X extensions. */
#ifndef NUMC_NUMC_H
#define NUMC_NUMC_H
#include "numc/export.h"
#include "numc/dtype.h"
#include "numc/error.h"
#include "numc/array.h"
#include "numc/math.h"
#include "numc_amd.h"
#endif
