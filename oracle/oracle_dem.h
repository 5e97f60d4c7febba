/*
 * FP64 CPU oracle of the DEMSolver path (test infrastructure only; see
 * oracle.h).  Restates DEMSPH/kernel/DEMSolver.h step sequencing and the
 * kernels of ballNeighborSearchKernel.cu, contactKernel.cu(h),
 * ballIntegrationKernel.cu, wallIntegrationKernel.cu(h) and the hook
 * kernels of DEMSPH/externalForceTorque.cu, in insertion order, with the
 * reference's pass structure and summation order.
 */
#ifndef DEMSPH_ORACLE_DEM_H
#define DEMSPH_ORACLE_DEM_H

#include <stddef.h>
#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct o_dem o_dem;

o_dem* o_dem_create(void);
void o_dem_destroy(o_dem* d);

/* Material-pair rows (DEMSolver.h:1637-1726); p holds the row values in the
 * reference's argument order: linear 11 (kn ks kr kt dn ds dr dt mus mur mut),
 * Hertz 8 (E* G* e kr/ks kt/ks mus mur mut), bonded 6 (gamma E kn/ks sigma C mu). */
void o_dem_set_linear(o_dem* d, int a, int b, const double* p);
void o_dem_set_hertz(o_dem* d, int a, int b, const double* p);
void o_dem_set_bonded(o_dem* d, int a, int b, const double* p);

/* addBall / addFixedBall (DEMSolver.h:1728-1768). */
void o_dem_add_ball(o_dem* d, const double* pos, const double* vel, const double* w, double r, double density,
                    int mat, int clump);
void o_dem_add_fixed_ball(o_dem* d, const double* pos, double r, int mat, int clump);
/* addClump / addFixedClump (DEMSolver.h:1796-1875); inertia = xx yy zz xy xz yz. */
void o_dem_add_clump(o_dem* d, int np, const double* pts, const double* radii, const double* centroid,
                     const double* vel, const double* w, double mass, const double* inertia, int mat);
void o_dem_add_fixed_clump(o_dem* d, int np, const double* pts, const double* radii, const double* centroid, int mat);
/* addMeshWall (DEMSolver.h:1882-1935).  Returns -1 when ignored. */
int o_dem_add_mesh_wall(o_dem* d, int nv, const double* verts, int nt, const int* t0, const int* t1, const int* t2,
                        const double* pos, const double* vel, const double* w, int mat);
/* bondedInteraction::add (DEMSolver.h:93-134). */
void o_dem_add_bonds(o_dem* d, int n, const int* ob0, const int* ob1, int npos, const double* pos);

/* setPeriodicBoundary(X, Y, Z) (DEMSolver.h:1944-1959), intended semantics:
 * wrapped stencil, periodic images in the ball-pair list, nearest images for
 * bonds, balls (not clump pebbles) wrapped into the domain at each neighbour
 * search.  Call before o_dem_initialize. */
void o_dem_set_periodic(o_dem* d, int x, int y, int z);
/* initialize() (DEMSolver.h:245-276).  Returns -2 when a periodic axis has
 * fewer than three cells. */
int o_dem_initialize(o_dem* d, const double* minb, const double* maxb);
/* neighborSearch() (DEMSolver.h:698-777). */
void o_dem_neighbor_search(o_dem* d);
/* calculateBallContactForceTorque(timeStep) (DEMSolver.h:599-695). */
void o_dem_contact(o_dem* d, double dt);
/* addTriangleContactForceTorque(timeStep) (DEMSolver.h:428-597). */
void o_dem_triangle_contact(o_dem* d, double dt);
/* interaction1stHalf / 2ndHalf (DEMSolver.h:779-906). */
void o_dem_first_half(o_dem* d, const double* g, double hdt);
void o_dem_second_half(o_dem* d, const double* g, double hdt);
/* Start of solve() (DEMSolver.h:1994-1997): neighbour search and force
 * evaluation with zero time step. */
void o_dem_solve_begin(o_dem* d);
/* nsteps iterations of loop_ball / loop_ball_wall (DEMSolver.h:1376-1405,
 * 1452-1485), including the hook program at t - dt/2 and t. */
void o_dem_step(o_dem* d, const double* g, double dt, int nsteps);

/* Hook program (DEMSPH/externalForceTorque.cu), applied in order at each
 * addExternalForceTorque call.  Constant force per ball scaled by
 * min(time / ramp, 1) (ramp <= 0: factor 1). */
void o_dem_hook_clear(o_dem* d);
void o_dem_hook_constant_force(o_dem* d, const double* f, double ramp);
void o_dem_hook_damping(o_dem* d, double c);
void o_dem_hook_buoyancy(o_dem* d, double rho, double cd, const double* g, const double* vf, double level);

/* Clump force sum: 0 = intended (pebble forces), 1 = reproduce D1.  D1 reads
 * the clump array out of range in the reference and is not reproducible;
 * the switch exists so that tests can show the difference is rejected. */
void o_dem_set_policy(o_dem* d, int d1);

long o_dem_get(o_dem* d, const char* name, void* out);
int o_dem_set(o_dem* d, const char* name, const void* in);
void o_dem_grids(o_dem* d, o_grid* balls, o_grid* tris);
size_t o_dem_step_count(o_dem* d);

/* Closed-form helpers for tests. */
int o_sphere_triangle(const double* c, double r, const double* v0, const double* v1, const double* v2, double* cp);

/* ------------------------------------------------------------ coupling
 * Access for the DEM-SPH coupling (oracle_sph.c): the ball arrays
 * (interleaved x3 for vectors, insertion order) and one step of loop_ball /
 * loop_ball_wall (DEMSolver.h:1376-1405, 1452-1485) split where external
 * forces enter. */
typedef struct {
    int n;
    double *pos, *vel, *w, *radius, *invMass, *force, *torque;
} o_dem_balls;
void o_dem_couple_view(o_dem* d, o_dem_balls* v);
double o_dem_couple_begin(o_dem* d, double dt);
void o_dem_couple_evaluate(o_dem* d, double h, double dt);
void o_dem_couple_kick(o_dem* d, int first, const double* g, double h, double hookTime);
/* nsteps coupled steps of a fluid scene and a DEM scene (oracle_sph.c). */
void o_couple_step(o_sph* s, o_dem* d, const double* g, double dt, int nsteps);
/* Coupling results of the last evaluation; returns the element count. */
long o_couple_get(o_sph* s, const char* name, double* out);

#ifdef __cplusplus
}
#endif
#endif
