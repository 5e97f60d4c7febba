// Internal C++ interfaces between the scene implementations of the library
// (used by the DEM-SPH coupling); not part of the C ABI.
#pragma once

#include "common.hpp"

struct demsph_dem;
struct demsph_sph;
struct demsph_couple;

namespace demsph {

// The DEM scene's current ball buffers (cell-sorted slot order) and its
// force / torque accumulators.
struct DemBallRefs {
    int n;
    const double *px, *py, *pz, *vx, *vy, *vz, *wx, *wy, *wz, *r, *im;
    const int* orig;
    double *fx, *fy, *fz, *tx, *ty, *tz;
};

DemBallRefs dem_ball_refs(demsph_dem* d);
hipStream_t dem_stream(demsph_dem* d);
void dem_set_stream(demsph_dem* d, hipStream_t s);
// Throws unless the scene is initialised and solve_begin() has run.
void dem_require_ready(demsph_dem* d);

// One DEM step split at the points where external forces enter
// (loop_ball / loop_ball_wall, DEMSolver.h:1376-1405, 1452-1485):
//   begin_step: iStep++, time += dt, pair rebuild every 5th step; returns time
//   evaluate:   contact, bond and wall forces (overwrite the accumulators)
//   hooks_and_half: the external-force hook at hookTime, then a half kick
//                   (first: kick + drift)
//   end_block:  deferred capacity checks
double dem_begin_step(demsph_dem* d, double dt);
void dem_evaluate(demsph_dem* d, double h, double dt);
void dem_hooks_and_half(demsph_dem* d, bool first, const double* g, double h, double hookTime);
void dem_end_block(demsph_dem* d);
// Wall-candidate overflow since the last call (0: none; read back and
// cleared), and the wall history grown to hold `need` candidates per ball
// (an error above the maximum).
int dem_take_wall_overflow(demsph_dem* d);
void dem_grow_walls(demsph_dem* d, int need);
// A DEM scene keeps the couplings that hold it: when it is destroyed first,
// each of them is detached (it then refuses to step and touches nothing of
// that scene).
void dem_attach_couple(demsph_dem* d, demsph_couple* c);
void dem_detach_couple(demsph_dem* d, demsph_couple* c);
void couple_detach_dem(demsph_couple* c);
// demsph_dem_autotune once per size class in the process (ball count to a
// quarter octave, walls, bonds, periodic axes, precision): later scenes of
// the class take the cached "pairLanes" choice without trials.
void dem_autotune_cached(demsph_dem* d, const double* g, double dt, int trialSteps);
void dem_autotune_clear_cache();
// Device checkpoint of a DEM scene for replaying a block of steps (every
// array of the state export, the host values that size them and the step
// stamp counter); restore grows both pair lists to at least growPairs.
void dem_block_save(demsph_dem* d);
void dem_block_restore(demsph_dem* d, int growPairs);

// Batched serving: device bytes a staged scene will allocate (admission
// under a memory budget), sizes, step-end timestamps and streams.
size_t sph_bytes_estimate(demsph_sph* s, const double* minb, const double* maxb);
size_t dem_bytes_estimate(demsph_dem* d, const double* minb, const double* maxb);
size_t couple_bytes_estimate(demsph_sph* f, demsph_dem* d, const double* fluidMin, const double* fluidMax);
int sph_num_fluid(demsph_sph* s);
int dem_num_balls(demsph_dem* d);
void sph_set_stamps(demsph_sph* s, const StepStamps& st);
void dem_set_stamps(demsph_dem* d, const StepStamps& st);
hipStream_t sph_stream(demsph_sph* s);

}  // namespace demsph
