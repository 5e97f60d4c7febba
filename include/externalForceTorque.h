/*
 * externalForceTorque.h -- HIP launchers of the reference's external force
 * kernels (DEMSPH/externalForceTorque.h:6-35, .cu:4-161), same names,
 * argument order and meaning, with hipStream_t for cudaStream_t.  They act
 * on the FP64 insertion-ordered device views of a DEM scene
 * (demsph_dem_views), e.g. from inside an addExternalForceTorque hook.
 */
#ifndef DEMSPH_EXTERNAL_FORCE_TORQUE_H
#define DEMSPH_EXTERNAL_FORCE_TORQUE_H

#include <stddef.h>
#include <hip/hip_runtime.h>

#ifdef __cplusplus
extern "C" {
#endif

/* force[i] += force_external[i]  (externalForceTorque.cu:4-24) */
void launchAddConstantForce(double3* force, const double3* force_external, const size_t num, const size_t gridD,
                            const size_t blockD, hipStream_t stream);

/* F -= c |F| v/|v|, T -= c |T| w/|w|  (externalForceTorque.cu:26-57) */
void launchAddGlobalDampingForceTorque(double3* force, double3* torque, const double3* velocity,
                                       const double3* angularVelocity, const double dampCoeff, const size_t num,
                                       const size_t gridD, const size_t blockD, hipStream_t stream);

/* Buoyancy of the submerged cap below z = waterLevel plus quadratic drag
 * against fluidVelocity, skipped for inverseMass <= 0
 * (externalForceTorque.cu:59-160). */
void launchAddBuoyancyDrag(double3* force, double3* position, double3* velocity, double* radius, double* inverseMass,
                           const double rhoFluid, const double Cd, const double3 gravity, const double3 fluidVelocity,
                           const double waterLevel, const size_t numBall, const size_t gridD, const size_t blockD,
                           hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif
