// generic_inverters.h -- the solvers of the GPU path in one include.
#ifndef GLX_COMPAT_INVERTERS
#define GLX_COMPAT_INVERTERS

#include "generic_bicgstab.h"
#include "generic_cg.h"
#include "generic_cg_m.h"
#include "generic_gcr.h"

#endif
