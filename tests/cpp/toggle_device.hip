// The reference's toggle-switch propensity (tests/test_fsp_mat.cu:57-84) as a
// __device__ function, and the __device__ pointer the host fetches with
// hipMemcpyFromSymbol (tests/test_fsp_mat.cu:86) -- in a translation unit of
// its own, so the program is linked with -fgpu-rdc.
#include <hip/hip_runtime.h>

#include "cme_util.h"

__device__ double toggle_propensity_device(int *x, int reaction)
{
    const double ayx = 2.6e-3, axy = 6.1e-3, nyx = 3e0, nxy = 2.1e0;
    switch (reaction) {
    case 0: return 1.0;
    case 1: return 1.0 / (1.0 + ayx * pow(1.0 * x[1], nyx));
    case 2: return 1.0 * x[0];
    case 3: return 1.0;
    case 4: return 1.0 / (1.0 + axy * pow(1.0 * x[0], nxy));
    case 5: return 1.0 * x[1];
    default: return 0.0;
    }
}

__device__ cuFSP::PropFun toggle_prop_pointer = &toggle_propensity_device;
