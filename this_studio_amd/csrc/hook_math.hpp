// Shared arithmetic of the external force/torque hooks.  Both the built-in
// device hook program (dem.hip) and the exported launchers (hook.hip) use
// these, compiled without FMA contraction, so a force added through either
// route rounds identically.
#pragma once
#include <hip/hip_runtime.h>

__device__ __forceinline__ double hk_len(double x, double y, double z)
{
#pragma clang fp contract(off)
    return sqrt(x * x + y * y + z * z);
}

// myVec.h:112-128: zero for squared length below 1e-20
__device__ __forceinline__ void hk_unit(double x, double y, double z, double& nx, double& ny, double& nz)
{
#pragma clang fp contract(off)
    const double l2 = x * x + y * y + z * z;
    if (l2 < 1.0e-20) {
        nx = ny = nz = 0.0;
        return;
    }
    const double inv = 1.0 / sqrt(l2);
    nx = x * inv;
    ny = y * inv;
    nz = z * inv;
}
