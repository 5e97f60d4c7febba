// External force / torque launchers for the addExternalForceTorque hook,
// acting on FP64 interleaved insertion-ordered ball views.  Restated from
// DEMSPH/externalForceTorque.cu:4-161 (constant force, global damping,
// spherical-cap buoyancy with quadratic drag).
#include "common.hpp"
#include "hook_math.hpp"
#include "../../include/externalForceTorque.h"

namespace {

__global__ void k_add_constant(double3* force, const double3* fe, size_t num)
{
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    force[i].x += fe[i].x;
    force[i].y += fe[i].y;
    force[i].z += fe[i].z;
}

__global__ void k_add_damping(double3* force, double3* torque, const double3* v, const double3* w, double c, size_t num)
{
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    const double3 f = force[i], t = torque[i];
    double nx, ny, nz;
    const double lf = hk_len(f.x, f.y, f.z);
    hk_unit(v[i].x, v[i].y, v[i].z, nx, ny, nz);
    force[i].x = f.x - nx * (c * lf);
    force[i].y = f.y - ny * (c * lf);
    force[i].z = f.z - nz * (c * lf);
    const double lt = hk_len(t.x, t.y, t.z);
    hk_unit(w[i].x, w[i].y, w[i].z, nx, ny, nz);
    torque[i].x = t.x - nx * (c * lt);
    torque[i].y = t.y - ny * (c * lt);
    torque[i].z = t.z - nz * (c * lt);
}

__global__ void k_add_buoyancy(double3* force, const double3* pos, const double3* vel, const double* radius,
                               const double* invMass, double rho, double cd, double3 g, double3 vf, double level,
                               size_t num)
{
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    if (invMass && invMass[i] <= 0.0) return;
    const double r = radius[i];
    const double dd = level - pos[i].z;
    double Vs;
    if (dd <= -r) return;
    if (dd >= r) Vs = (4.0 / 3.0) * M_PI * r * r * r;
    else {
        const double h = dd + r;
        Vs = (M_PI * h * h * (3.0 * r - h)) / 3.0;
    }
    if (Vs <= 0.0) return;
    const double Vb = (4.0 / 3.0) * M_PI * r * r * r;
    const double alpha = (Vb > 1.e-30) ? (Vs / Vb) : 0.0;
    const double rx = vel[i].x - vf.x, ry = vel[i].y - vf.y, rz = vel[i].z - vf.z;
    const double v2 = rx * rx + ry * ry + rz * rz;
    double dx = 0.0, dy = 0.0, dz = 0.0;
    if (v2 > 1.e-30 && cd > 0.0) {
        const double k = 0.5 * rho * cd * (alpha * (M_PI * r * r)) * sqrt(v2);
        dx = -k * rx; dy = -k * ry; dz = -k * rz;
    }
    force[i].x += -rho * Vs * g.x + dx;
    force[i].y += -rho * Vs * g.y + dy;
    force[i].z += -rho * Vs * g.z + dz;
}

}  // namespace

extern "C" void launchAddConstantForce(double3* force, const double3* force_external, const size_t num,
                                       const size_t gridD, const size_t blockD, hipStream_t stream)
{
    if (num == 0) return;
    hipLaunchKernelGGL(k_add_constant, dim3((unsigned)gridD), dim3((unsigned)blockD), 0, stream, force, force_external,
                       num);
}

extern "C" void launchAddGlobalDampingForceTorque(double3* force, double3* torque, const double3* velocity,
                                                  const double3* angularVelocity, const double dampCoeff,
                                                  const size_t num, const size_t gridD, const size_t blockD,
                                                  hipStream_t stream)
{
    if (num == 0) return;
    hipLaunchKernelGGL(k_add_damping, dim3((unsigned)gridD), dim3((unsigned)blockD), 0, stream, force, torque, velocity,
                       angularVelocity, dampCoeff, num);
}

extern "C" void launchAddBuoyancyDrag(double3* force, double3* position, double3* velocity, double* radius,
                                      double* inverseMass, const double rhoFluid, const double Cd, const double3 gravity,
                                      const double3 fluidVelocity, const double waterLevel, const size_t numBall,
                                      const size_t gridD, const size_t blockD, hipStream_t stream)
{
    if (numBall == 0) return;
    hipLaunchKernelGGL(k_add_buoyancy, dim3((unsigned)gridD), dim3((unsigned)blockD), 0, stream, force, position,
                       velocity, radius, inverseMass, rhoFluid, Cd, gravity, fluidVelocity, waterLevel, numBall);
}
