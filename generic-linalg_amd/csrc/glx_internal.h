// glx_internal.h -- library-internal hooks shared by the C ABI (glx_abi.hip)
// and the C++ compatibility layer (glx_compat.cpp).  Plain C++, no HIP.
#pragma once

#include "../../include/glx.h"

namespace glx {

// Registers a callback whose extra_info describes a general stencil in some
// other form (the compat layer's stencil_2d): `translate` fills a
// glx_stencil_2d_nc from it and returns 0, so solves given the callback run
// on the device.  kind is one of GLX_OP_CSTENCIL*.
void register_cst_callback(void* fn, int kind, int (*translate)(void* extra, glx_stencil_2d_nc* out));

}  // namespace glx
