/*
 * demsph.h -- C ABI of the MI355X-native DEMSPH particle path (libdemsph.so).
 *
 * Plain C types only: pointers, sizes, doubles, ints and an opaque stream
 * handle (a hipStream_t passed as void*).  Every entry point returns an int
 * status (DEMSPH_OK on success) and never throws; demsph_last_error()
 * describes the last failure of the calling thread.
 *
 * Each declaration cites the reference interface it replaces
 * (kaiqideng/this_studio, paths relative to DEMSPH/).  The reference exposes
 * void CUDA launchers over raw device pointers and host-side C++ solver
 * classes; this ABI keeps the solver-level semantics behind opaque scene
 * handles that own their device memory in a cell-sorted SoA layout.
 *
 * Host arrays of vectors are interleaved (x,y,z per element), like the
 * reference's double3 arrays.  Results are returned in insertion order.
 */
#ifndef DEMSPH_H
#define DEMSPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEMSPH_OK 0
#define DEMSPH_ERR_INVALID 1   /* bad argument or call order */
#define DEMSPH_ERR_CAPACITY 2  /* a capacity limit was exceeded and could not be grown */
#define DEMSPH_ERR_HIP 3       /* HIP runtime error */
#define DEMSPH_ERR_INTERNAL 4

/* Thread-local description of the last failure (never NULL). */
const char* demsph_last_error(void);
/* Library version string. */
const char* demsph_version(void);

/* ------------------------------------------------------------------ cells
 * Drop-in for launchUpdateSpatialGridCellHashStartEnd
 * (kernel/neighborSearchKernel.cuh:16-31, .cu:137-174) and
 * buildHashStartEnd (kernel/buildHashStartEnd.cu:36-83).
 * Device pointers: position (n interleaved doubles x3), hashIndex[n],
 * hashValue[n], cellHashStart[numGrids], cellHashEnd[numGrids].
 * On return hashValue holds the sorted hashes and hashIndex the stable sort
 * permutation (lexicographic (hash, index)); cell arrays hold the first and
 * one-past-last sorted position of each cell, -1 for empty cells.
 * grid: min[3], max[3], cellSize[3], gridSize[3] exactly as spatialGrid
 * stores them.  Temporary storage is allocated by the library and cached. */
int demsph_build_cells(const double* position, int* hashIndex, int* hashValue,
                       int* cellHashStart, int* cellHashEnd,
                       const double* minBound, const double* maxBound,
                       const double* cellSize, const int* gridSize,
                       size_t numGrids, size_t numObject, void* stream);

/* demsph_build_cells without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as demsph_build_cells;
 * the sort scratch is a fixed array inside the library (about 28 MiB per
 * GPU, held until the process ends), so numObject is limited to
 * demsph_build_cells_noalloc_max_objects() (2,097,152) and larger inputs
 * return DEMSPH_ERR_CAPACITY.  The caller passes every input and output.
 * All calls on one GPU share that scratch: calls that overlap in time on
 * one GPU (other streams or threads) race and must be serialised. */
int demsph_build_cells_noalloc(const double* position, int* hashIndex, int* hashValue,
                               int* cellHashStart, int* cellHashEnd,
                               const double* minBound, const double* maxBound,
                               const double* cellSize, const int* gridSize,
                               size_t numGrids, size_t numObject, void* stream);
size_t demsph_build_cells_noalloc_max_objects(void);

/* spatialGrid::set (kernel/boundary.h:70-94): computes gridSize and
 * cellSize from the domain and the one-dimensional cell size.  Returns
 * DEMSPH_ERR_INVALID where the reference silently ignores the call. */
int demsph_grid_geometry(const double* minBound, const double* maxBound, double cellSizeOneDim,
                         double* cellSize, int* gridSize);

/* ------------------------------------------------------------------ WCSPH
 * Opaque WCSPHSolver scene (kernel/WCSPHSolver.h:18-636). */
typedef struct demsph_sph demsph_sph;

/* WCSPHSolver(cudaStream_t) (WCSPHSolver.h:21-24).  stream may be NULL
 * (the library creates its own). */
int demsph_sph_create(demsph_sph** out, void* stream);
int demsph_sph_destroy(demsph_sph* s);

/* addSPH / addDummy (WCSPHSolver.h:523-541), batched.  Entries with
 * density <= 0 or soundSpeed <= 0 are skipped as in the reference.
 * mass = spacing^3 * density, smoothing length = 1.3 * spacing. */
int demsph_sph_add_fluid(demsph_sph* s, size_t n, const double* position, const double* velocity,
                         const double* soundSpeed, const double* spacing, const double* density,
                         const double* viscosity);
int demsph_sph_add_dummy(demsph_sph* s, size_t n, const double* position, const double* velocity,
                         const double* soundSpeed, const double* spacing, const double* density,
                         const double* viscosity);

/* Equation of state.  eos 0: p = c^2 * drho (the reference,
 * WCSPHIntegrationKernel.cu:266-282); eos 1: Tait p = B((rho/rho0)^gamma-1),
 * B = rho0 c^2/gamma.  Default 0. */
int demsph_sph_set_eos(demsph_sph* s, int eos, double gamma);

/* Per-particle neighbour capacity of the fluid and fluid-dummy lists
 * (grows automatically on overflow).  Default 128 / 64. */
int demsph_sph_set_capacity(demsph_sph* s, int fluidCapacity, int dummyCapacity);

/* initialize() (WCSPHSolver.h:137-147): grid with cell 2*max(h), dummy
 * sort, dummy normals (WCSPHIntegrationKernel.cu:4-40).  gravity[3]. */
int demsph_sph_initialize(demsph_sph* s, const double* minBound, const double* maxBound,
                          const double* gravity);

/* neighborSearch() (WCSPHSolver.h:150-285): hash, sort, cell ranges,
 * fluid-fluid and fluid-dummy lists. */
int demsph_sph_neighbor_search(demsph_sph* s);
/* interaction1stHalf: acceleration, half kick, drift
 * (launchWCSPH1stHalfIntegration, WCSPHIntegrationKernel.cu:308-376). */
int demsph_sph_first_half(demsph_sph* s, double timeStep);
/* interaction2ndHalf: density change, EOS, acceleration, half kick
 * (launchWCSPH2ndHalfIntegration, WCSPHIntegrationKernel.cu:378-472). */
int demsph_sph_second_half(demsph_sph* s, double timeStep);
/* nsteps iterations of the solve() loop body (WCSPHSolver.h:607-621).
 * Neighbour-list overflow is detected on the device, the block of steps is
 * replayed from a device checkpoint with a larger capacity. */
int demsph_sph_step(demsph_sph* s, double timeStep, int nsteps);
/* Waits for all work of the scene. */
int demsph_sph_synchronize(demsph_sph* s);

/* Sizes: numFluid, numDummy, numGrids. */
int demsph_sph_sizes(demsph_sph* s, size_t* numFluid, size_t* numDummy, size_t* numGrids);

/* Copies a field to host memory in insertion order.  Fields (element type,
 * count):  "position" "velocity" (double, 3n)  "densityChange" "density"
 * "pressure" (double, n)  "dummyNormal" (double, 3*numDummy)
 * "hashValue" (int, n: cell hash of the last neighbour search, by index)
 * "hashIndex" (int, n: sorted permutation)  "cellHashStart" "cellHashEnd"
 * (int, numGrids)  "neighborCount" "dummyNeighborCount" (int, n)
 * "objectPointing" "dummyObjectPointing" (int, total pairs: the reference's
 * pair list, grouped by pointed particle in insertion order).
 * *count receives the element count; out may be NULL to query it. */
int demsph_sph_get(demsph_sph* s, const char* field, void* out, size_t capacity, size_t* count);

/* Frame files: the reference's outputWCSPHVTU (WCSPHSolver.h:360-443,
 * dir/SPH_FFFF.vtu) and outputDummyVTU (:445-520, dir/dummy_FFFF.vtu), the
 * same ASCII VTU layout, particles in insertion order, values %.10f from
 * FP64 (density, pressure and dummy normals are stored in FP32).  what:
 * DEMSPH_VTU_SPH | DEMSPH_VTU_DUMMY; dir is created when missing. */
#define DEMSPH_VTU_SPH 1
#define DEMSPH_VTU_DUMMY 2
int demsph_sph_write_vtu(demsph_sph* s, const char* dir, int what, long frame, long step, double time);
/* The same without waiting for the fluid frame: the scene's arrays are
 * copied in stream order to pinned host memory and a background thread
 * writes the file while the scene keeps stepping (no counterpart in the
 * reference, whose writers block).  Dummy frames are written before the
 * call returns.  At most two frames wait for the thread; a third call
 * waits for a slot.  demsph_frames_wait() waits for every pending frame
 * and reports the first failure since its last call. */
int demsph_sph_write_vtu_async(demsph_sph* s, const char* dir, int what, long frame, long step, double time);
int demsph_frames_wait(void);

/* Scene state to host memory and back (no counterpart in the reference; for
 * evicting a served job and resuming it): the fluid's dynamic state in its
 * current cell-sorted order.  A scene staged with the same particles and
 * initialised on the same domain continues from an imported state bit for
 * bit as the exporting scene does.  *bytes receives the size; out may be
 * NULL to query it. */
int demsph_sph_export_state(demsph_sph* s, void* out, size_t capacity, size_t* bytes);
int demsph_sph_import_state(demsph_sph* s, const void* in, size_t bytes);

/* Overwrites position / velocity (double, 3n) or densityChange (double, n)
 * by insertion index (density and pressure follow from the EOS). */
int demsph_sph_set(demsph_sph* s, const char* field, const void* in, size_t count);

/* Kernel timing: when enabled, every kernel launch of the scene is
 * bracketed by HIP events on the scene stream and accumulated per kernel
 * name.  demsph_sph_timing reads (name, launches, total ms) triples. */
int demsph_sph_enable_timing(demsph_sph* s, int enable);
int demsph_sph_timing(demsph_sph* s, int index, char* name, size_t nameCap, long* launches, double* totalMs);

/* Step graphs (no counterpart in the reference, whose solve() loop launches
 * every kernel from the host): demsph_sph_step runs pairs of steps as
 * replays of a captured two-step hipGraph, the first step after
 * initialisation and an odd remainder eagerly.  On by default; off while
 * kernel timing is enabled.  demsph_sph_graph_captures counts the captures
 * made so far (one per time step size and buffer parity until the lists are
 * reallocated). */
int demsph_sph_set_graphs(demsph_sph* s, int enable);
int demsph_sph_graph_captures(demsph_sph* s, long* count);

/* Kernel tuning knobs (no counterpart in the reference).  "pairLanes": lanes
 * per fluid particle in the momentum and continuity passes, 1, 2, 4 or 8,
 * or 0 to choose from the particle count (more lanes below a chip's worth
 * of resident waves, where each particle's chain of neighbour gathers bounds
 * the pass).  The neighbour sums are split over the lanes and combined in a
 * fixed order, so a setting gives deterministic results; different settings
 * differ by FP32 rounding.  "listLanes": lanes per fluid particle in the
 * neighbour-list build (same values); every setting builds the same lists.
 * "listPrefilter": 1 (default) tests list candidates first in FP32 on a
 * 16-byte copy of the positions and only those within 1e-3 of the cut (or
 * far from the grid origin) in exact FP64, 0 tests all in FP64; both give
 * the same lists.  "pairRecords": 1 (default) lets the pair passes of a
 * one-phase scene gather 32-byte records holding each neighbour's position
 * as an FP32 offset from its cell's origin (its cell coded in the list
 * entry; the exported lists are unchanged), 0 gathers the FP64 positions;
 * results differ by FP32 rounding.  demsph_sph_tune_get reads the value in
 * effect ("pairRecords": 1 only where the records are in use). */
int demsph_sph_tune(demsph_sph* s, const char* knob, int value);
int demsph_sph_tune_get(demsph_sph* s, const char* knob, int* value);

/* Autotuner (no counterpart in the reference: its block shapes are fixed).
 * Times the variants on the initialised scene's own steps, each trial one
 * untimed step and then trialSteps timed steps (HIP events) from a
 * checkpoint that is restored afterwards, so the scene's state and its
 * later results are those of a scene never tuned with the chosen knobs.
 * mode 0 tries only result-invariant knobs (listLanes 1, 2, 4, 8 with and
 * without listPrefilter: every variant builds the same lists); mode 1 also
 * tries pairLanes (FP32 summation order changes).  A variant within 2 % of
 * the built-in rule keeps the rule.  The cell size is not tuned: it fixes
 * the list order and the sort (2 max h, WCSPHSolver.h:127-135).  The choice
 * is set on the scene and cached per size class (uniform phase, fluid count
 * to a quarter octave, dummy count to a half octave, mode) for the process;
 * with useCache a cached class is applied without timing.  Not while
 * coupled stepping is in flight on the scene's stream. */
typedef struct demsph_sph_tune_report {
    int listLanes, listPrefilter, pairLanes; /* now in effect */
    int rulesListLanes, rulesPairLanes;      /* the built-in lane rules */
    int configs;                             /* trials timed (0 when cached) */
    int cached;                              /* 1: taken from the cache */
    double listUs, pairUs;                   /* per step: list build; the three pair passes */
    double rulesListUs, rulesPairUs;         /* the same under the rules (0 if not timed) */
} demsph_sph_tune_report;
int demsph_sph_autotune(demsph_sph* s, double dt, int trialSteps, int mode, int useCache,
                        demsph_sph_tune_report* report);
/* Empties the process-wide autotuner caches (WCSPH choices and the DEM
 * pair-lane choices the server's admission tuning keeps per size class). */
int demsph_sph_autotune_clear_cache(void);

/* ------------------------------------------------------------------ DEM
 * Opaque DEMSolver scene (kernel/DEMSolver.h:144-2079). */
typedef struct demsph_dem demsph_dem;

/* DEMSolver(cudaStream_t) (DEMSolver.h:147-151).  stream may be NULL (the
 * legacy default stream, as the reference's apps use). */
int demsph_dem_create(demsph_dem** out, void* stream);
int demsph_dem_destroy(demsph_dem* d);

/* Material-pair rows, appended (DEMSolver.h:1637-1726):
 * setLinearPair(a, b, kn, ks, kr, kt, dn, ds, dr, dt, mus, mur, mut),
 * setHertzianPair(a, b, E*, G*, e, kr/ks, kt/ks, mus, mur, mut),
 * setBondedPair(a, b, gamma, E, kn/ks, sigma_t, C, mu). */
int demsph_dem_set_linear_pair(demsph_dem* d, int a, int b, double kn, double ks, double kr, double kt, double dn,
                               double ds, double dr, double dt, double mus, double mur, double mut);
int demsph_dem_set_hertzian_pair(demsph_dem* d, int a, int b, double E, double G, double e, double krks,
                                 double ktks, double mus, double mur, double mut);
int demsph_dem_set_bonded_pair(demsph_dem* d, int a, int b, double gamma, double E, double knks, double sigmat,
                               double C, double mu);

/* addBall (DEMSolver.h:1728-1750), batched: mass = 4/3 pi r^3 density,
 * inverse mass 0 when mass <= 0.  clumpID may be NULL (all -1). */
int demsph_dem_add_balls(demsph_dem* d, size_t n, const double* position, const double* velocity,
                         const double* angularVelocity, const double* radius, const double* density,
                         const int* materialID, const int* clumpID);
/* addFixedBall (DEMSolver.h:1752-1768), batched. */
int demsph_dem_add_fixed_balls(demsph_dem* d, size_t n, const double* position, const double* radius,
                               const int* materialID, const int* clumpID);
/* addClump (DEMSolver.h:1796-1845): pebbles get the clump's mean density,
 * velocity v + w x (p - centroid) when mass > 1e-20; inertia = xx yy zz xy xz yz. */
int demsph_dem_add_clump(demsph_dem* d, size_t numPebbles, const double* points, const double* radii,
                         const double* centroid, const double* velocity, const double* angularVelocity,
                         double mass, const double* inertia, int materialID);
/* addFixedClump (DEMSolver.h:1847-1875). */
int demsph_dem_add_fixed_clump(demsph_dem* d, size_t numPebbles, const double* points, const double* radii,
                               const double* centroid, int materialID);
/* addMeshWall (DEMSolver.h:1882-1935).  As in the reference the wall starts
 * at rest (velocity arguments are not stored) and an index out of range
 * makes the call a no-op (return DEMSPH_OK, nothing added). */
int demsph_dem_add_mesh_wall(demsph_dem* d, size_t numVertices, const double* vertices, size_t numTriangles,
                             const int* tri0, const int* tri1, const int* tri2, const double* position,
                             const double* velocity, const double* angularVelocity, int materialID);
/* addBondedInteraction (DEMSolver.h:1937-1942, bondedInteraction::add
 * :93-134): (min,max) deduplicated, point = midpoint, normal = unit(pa-pb).
 * ballPosition: numPositions interleaved doubles x3. */
int demsph_dem_add_bonded_interactions(demsph_dem* d, size_t n, const int* objectPointed,
                                       const int* objectPointing, size_t numPositions, const double* ballPosition);

/* setPeriodicBoundary(X, Y, Z) (DEMSolver.h:1944-1959, ghost passes
 * :279-426, neighborSearchKernel.cu:29-77), before initialize.  Intended
 * semantics, not the reference's seven ghost passes (which list some
 * contacts twice and miss others): on a periodic axis the 27-cell stencil
 * wraps around the domain, each contact with a periodic image is listed
 * once in the ball-pair list with its image code (field "pairShift": (sx+1)
 * + 3 (sy+1) + 9 (sz+1), s = -1, 0, +1 periods; 13 = no image), bonds join
 * nearest images, and balls that are not clump pebbles are wrapped into
 * [min, max) at every pair rebuild.  A periodic axis needs at least three
 * ball cells (DEMSPH_ERR_INVALID at initialize otherwise). */
int demsph_dem_set_periodic_boundary(demsph_dem* d, int x, int y, int z);

/* Contact force arithmetic: 0 = FP32 forces with FP64 overlap, contact
 * geometry and spring history (default); 1 = FP64 throughout. */
int demsph_dem_set_precision(demsph_dem* d, int fp64Forces);

/* initialize() (DEMSolver.h:245-276): grids, device upload.  May be called
 * again after adding bonds (the reference's addInitialCondition path). */
int demsph_dem_initialize(demsph_dem* d, const double* minBound, const double* maxBound);
/* neighborSearch() (DEMSolver.h:698-777). */
int demsph_dem_neighbor_search(demsph_dem* d);
/* calculateBallContactForceTorque(timeStep) (DEMSolver.h:599-695): ball
 * pairs, bonds and the per-ball reduction (forces are overwritten). */
int demsph_dem_contact(demsph_dem* d, double timeStep);
/* addTriangleContactForceTorque(timeStep) (DEMSolver.h:428-597): ball-wall
 * contacts added to the forces, then rigid wall motion over timeStep. */
int demsph_dem_triangle_contact(demsph_dem* d, double timeStep);
/* interaction1stHalf / 2ndHalf (DEMSolver.h:779-906). */
int demsph_dem_first_half(demsph_dem* d, const double* gravity, double halfTimeStep);
int demsph_dem_second_half(demsph_dem* d, const double* gravity, double halfTimeStep);
/* The start of solve() (DEMSolver.h:1994-1997): neighbour search and one
 * force evaluation with zero time step; resets the step counter. */
int demsph_dem_solve_begin(demsph_dem* d);
/* nsteps iterations of loop_ball / loop_ball_wall (DEMSolver.h:1376-1405,
 * 1452-1485): rebuild every 5th step, two force evaluations per step,
 * external forces (hook) at t - dt/2 and t, kick-drift-kick.  The pair
 * lists of rebuilds inside the call are sized on the device (the reference
 * reads the total back at every rebuild, interaction.h:164-175): a list that
 * outgrows its capacity is truncated, and the call goes back to a device
 * checkpoint taken at its start and runs again with both lists grown, with
 * the same results.  With a host hook (demsph_dem_set_hook) the rebuilds
 * read the total back instead, so the hook sees each step once. */
int demsph_dem_step(demsph_dem* d, const double* gravity, double timeStep, int nsteps);
int demsph_dem_synchronize(demsph_dem* d);
/* Steps taken since solve_begin, and the simulated time. */
int demsph_dem_clock(demsph_dem* d, size_t* iStep, double* time);

/* External force hook (DEMSolver.h:1372-1374 addExternalForceTorque): the
 * callback runs on the host twice per step with device views valid (see
 * demsph_dem_views); forces/torques it adds are integrated.  NULL removes it. */
typedef void (*demsph_dem_hook_fn)(void* user, size_t iStep, double time);
int demsph_dem_set_hook(demsph_dem* d, demsph_dem_hook_fn fn, void* user);

/* Device views in insertion order, FP64 interleaved x3 (the reference's
 * getBall().force()/torque()/position()/velocity()/angularVelocity(),
 * radius(), inverseMass(), particle.h:234-241).  Valid during the hook; the
 * library scatters them before and gathers force/torque after the hook. */
typedef struct {
    double* force;
    double* torque;
    double* position;
    double* velocity;
    double* angularVelocity;
    double* radius;
    double* inverseMass;
    size_t numBalls;
} demsph_dem_ball_views;
int demsph_dem_views(demsph_dem* d, demsph_dem_ball_views* views);

/* Built-in external force program, applied in order at each hook point on
 * the device (DEMSPH/externalForceTorque.cu:4-161): a per-ball constant force
 * (n x3, insertion order) scaled by min(time/ramp, 1) (ramp <= 0: 1), global
 * damping F -= c|F| v/|v|, T -= c|T| w/|w|, and buoyancy + quadratic drag
 * against the plane z = waterLevel. */
int demsph_dem_hook_clear(demsph_dem* d);
int demsph_dem_hook_constant_force(demsph_dem* d, const double* force, double ramp);
int demsph_dem_hook_global_damping(demsph_dem* d, double coefficient);
int demsph_dem_hook_buoyancy_drag(demsph_dem* d, double rhoFluid, double Cd, const double* gravity,
                                  const double* fluidVelocity, double waterLevel);

/* Sizes: balls, clumps, walls, triangles, ball pairs, bonds. */
int demsph_dem_sizes(demsph_dem* d, size_t* numBalls, size_t* numClumps, size_t* numWalls, size_t* numTriangles,
                     size_t* numPairs, size_t* numBonds);
/* Copies a field in insertion order: "position" "velocity" "angularVelocity" "force" "torque"
 * (double 3n), "radius" "inverseMass" (double n), "objectPointed"
 * "objectPointing" (int, ball pairs as the reference lists them),
 * "slidingSpring" "rollingSpring" "torsionSpring" (double 3/pair),
 * "overlap" (double/pair), "triPointed" "triPointing" "triCancel" (int),
 * "triSlidingSpring" "triRollingSpring" "triTorsionSpring" (double 3/pair), "bondPointed" "bondPointing"
 * "bondIsBonded" (int), "bondNormalForce" "bondTorsionTorque" (double),
 * "bondShearForce" "bondBendingTorque" "bondPoint" "bondNormal" (double 3),
 * "clumpPosition" "clumpVelocity" "clumpAngularVelocity" (double 3/clump),
 * "clumpOrientation" (double 4/clump), "wallPosition" (double 3/wall),
 * "wallOrientation" (double 4/wall), "vertexGlobalPosition" (double 3/vertex),
 * "hashValue" "hashIndex" "cellHashStart" "cellHashEnd" (int). */
int demsph_dem_get(demsph_dem* d, const char* field, void* out, size_t capacity, size_t* count);

/* Frame files: the reference's outputBallVTU (DEMSolver.h:908-993,
 * dir/ball_FFFF.vtu) and outputMeshWallVTU (:995-1098,
 * dir/triangleWall_FFFF.vtu; nothing without triangles), the same ASCII VTU
 * layout, balls in insertion order, values %.10f.  what: DEMSPH_VTU_BALL |
 * DEMSPH_VTU_TRIANGLE_WALL | DEMSPH_VTU_BALL_TRIANGLE; dir is created when
 * missing. */
#define DEMSPH_VTU_BALL 1
#define DEMSPH_VTU_TRIANGLE_WALL 2
/* outputBallTriangleInteractionVTU (DEMSolver.h:1286-1370,
 * dir/ballTriangleInteraction_FFFF.vtu): every candidate ball-triangle
 * contact of the last force evaluation (point, overlap, normal, force,
 * torque, springs), by ball in insertion order, then in candidate order.
 * Needs demsph_dem_record_contacts on before that evaluation. */
#define DEMSPH_VTU_BALL_TRIANGLE 4
/* outputBallInteractionVTU (DEMSolver.h:1100-1185,
 * dir/ballInteraction_FFFF.vtu): every ball pair of the current list,
 * same fields, by pointed ball in insertion order, then in list order.
 * Needs demsph_dem_record_contacts on before the evaluation. */
#define DEMSPH_VTU_BALL_INTERACTION 8
/* outputBondedInteractionVTU (DEMSolver.h:1186-1285,
 * dir/bondedInteraction_FFFF.vtu): every bond in bond-list order with its
 * point, state, forces, torques and maximum stresses. */
#define DEMSPH_VTU_BONDED_INTERACTION 16
int demsph_dem_write_vtu(demsph_dem* d, const char* dir, int what, long frame, long step, double time);
/* The same with the ball frame written by the background frame thread (as
 * demsph_sph_write_vtu_async); the other kinds are written before the call
 * returns. */
int demsph_dem_write_vtu_async(demsph_dem* d, const char* dir, int what, long frame, long step, double time);
/* Keep each ball pair's and ball-triangle contact's point, normal, force
 * and torque from every force evaluation for the frame writer (off by
 * default: 96 bytes per pair or candidate slot and a store per entry). */
int demsph_dem_record_contacts(demsph_dem* d, int enable);

/* DEM scene state to host memory and back (no counterpart in the reference;
 * for evicting a served job and resuming it): ball state in
 * its current slot order with forces, the current pair list with its
 * contact history, clump and wall poses, wall-contact history, bond state,
 * step counter and time.  The importing scene must be staged with the same
 * balls, clumps, walls and bonds, initialised, and of the same precision. */
int demsph_dem_export_state(demsph_dem* d, void* out, size_t capacity, size_t* bytes);
int demsph_dem_import_state(demsph_dem* d, const void* in, size_t bytes);

/* Overwrites "position" "velocity" "angularVelocity" (double 3n, insertion
 * order) or "wallVelocity" "wallAngularVelocity" (double 3/wall; the
 * reference's getMeshWallBody() write path). */
int demsph_dem_set(demsph_dem* d, const char* field, const void* in, size_t count);

/* Step graphs (no counterpart in the reference): demsph_dem_step replays
 * each whole cycle of a pair rebuild and its five steps as one captured
 * hipGraph (and the four steps after a rebuild as another, at the start of
 * a call) when the scene runs on a stream of its own (not the null stream),
 * has no hooks, its walls are static and timing is off.  On by default;
 * bit-identical to eager steps.  demsph_dem_graph_captures counts the
 * captures. */
int demsph_dem_set_graphs(demsph_dem* d, int enable);
int demsph_dem_graph_captures(demsph_dem* d, long* count);
/* DEM kernel tuning knob (no counterpart in the reference): "pairLanes",
 * lanes per ball in the pair-list build: 0 (default: 8 up to 32,768 balls,
 * else 1), 1, 2, 4 or 8.  Every setting builds the same pair lists in the
 * same order.  "pairCapacity": entries of each of the two pair lists,
 * set between initialize and solve_begin (they grow on demand from there).
 * demsph_dem_tune_get also reads "wallCapacity", the wall-triangle
 * candidates kept per ball (16, or 64 once a ball has found more: the step
 * call replays from its device checkpoint with the wall history grown; more
 * than 64 is DEMSPH_ERR_CAPACITY with the scene back at the checkpoint), and
 * "wallReplays", the step calls run again for it.
 * "wallSlots": 1 (default) lets scenes of up to 32,768 balls evaluate
 * ball-wall contacts of cells with at most four candidate triangles in
 * registers, 0 keeps every ball on the general path (same results).
 * demsph_dem_tune_get reads the value in effect, and "pairReplays", the
 * step calls run again after a pair list outgrew its capacity. */
int demsph_dem_tune(demsph_dem* d, const char* knob, int value);
int demsph_dem_tune_get(demsph_dem* d, const char* knob, int* value);
/* Times each "pairLanes" setting (1, 2, 4, 8) over trialSteps >= 5 steps of
 * gravity g and step dt from the scene's current state, keeps the fastest
 * and restores the state (no counterpart in the reference).  The scene must
 * have run solve_begin; *lanes (may be NULL) receives the choice. */
int demsph_dem_autotune(demsph_dem* d, const double* g, double dt, int trialSteps, int* lanes);
int demsph_dem_enable_timing(demsph_dem* d, int enable);
int demsph_dem_timing(demsph_dem* d, int index, char* name, size_t nameCap, long* launches, double* totalMs);

/* ------------------------------------------------------------------ coupling
 * Two-way DEM-SPH coupling.  The reference has none (its DEM and WCSPH
 * solvers run separately; buoyancy and drag enter the DEM only through the
 * analytic hook of externalForceTorque.cu:66-161), so this model is the
 * project's own: every coupled ball is a solid particle of the fluid phase
 * (mass rho0 * 4/3 pi r^3, pressure interpolated from the fluid at its
 * centre, moving with its surface velocity) for the fluid particles within
 * their support 2h, with the Riemann interface pressure of
 * WCSPHIntegrationKernel.cu:42-148.  The ball receives exactly the opposite
 * of every fluid particle's pair force (and the torque of its viscous
 * part), so momentum is exchanged pairwise.  Both scenes must be
 * initialised (the DEM scene through solve_begin) and are advanced with one
 * time step on the fluid scene's stream.  Destroy the coupling before
 * either scene; the DEM scene then returns to its own stream. */
typedef struct demsph_couple demsph_couple;
int demsph_couple_create(demsph_couple** out, demsph_sph* fluid, demsph_dem* dem);
int demsph_couple_destroy(demsph_couple* c);
/* Per fluid particle capacity of the fluid -> ball list (default 32). */
int demsph_couple_set_capacity(demsph_couple* c, int ballsPerFluid);
/* nsteps coupled kick-drift-kick steps of both scenes (DEM pair rebuild
 * every 5th step as in DEMSolver.h:1384).  Both scenes are checkpointed on
 * the device first; a fluid list or fluid -> ball list that outgrows its
 * capacity inside the block grows, both scenes go back to the checkpoint
 * and the block runs again, with the results of ample lists. */
int demsph_couple_step(demsph_couple* c, const double* gravity, double timeStep, int nsteps);
/* "couplingForce" "couplingTorque" (double 3 per ball, DEM insertion order,
 * last evaluation), "ballPressure" (double per ball), and
 * "fluidCouplingAcceleration" (double 3 per fluid particle, insertion order,
 * the coupling part of the last fluid momentum pass), "blockReplays" (one
 * double: coupled blocks run again after a list outgrew its capacity). */
int demsph_couple_get(demsph_couple* c, const char* field, void* out, size_t capacity, size_t* count);

/* ------------------------------------------------------------ serving
 * Batched serving of independent scenes on one GPU.  No counterpart in the
 * reference, which runs one scene per process with host synchronisation
 * inside its step loop (interaction.h:164-175).
 *
 * A job is a staged, not yet initialised WCSPH scene, DEM scene, or both
 * (coupled), with its bounds, gravity, time step, step count and arrival
 * time (seconds after demsph_server_run starts).  Submitting hands the scene
 * handles to the server, which destroys them when the job ends.  Jobs are
 * admitted in arrival order (FIFO) once they have arrived, one of `slots`
 * slots is free and their estimated device memory fits `memoryBudgetBytes`
 * beside the admitted jobs; a job is initialised on admission, advanced on
 * its own stream in blocks of steps (WCSPH by two-step graph replays when
 * `graphs` is non-zero), its final state is copied to the host and its
 * memory released.  Each step ends with a device timestamp; a job's step
 * latencies are the intervals between consecutive step completions.
 * slots = 1 with graphs = 0 serves the same jobs one at a time with one
 * launch per stage from the host (the plain path). */
typedef struct demsph_server demsph_server;
typedef struct {
    int kind;           /* 0 WCSPH, 1 DEM, 2 coupled */
    int status;         /* DEMSPH_OK or the job's error code */
    long particles;     /* fluid particles + balls */
    long steps, stepsDone;
    size_t bytes;       /* admission estimate of the job's device memory */
    /* seconds after run() started (-1: never): arrival, admission, first
     * step enqueued (set-up done), last step completed, memory released */
    double arrival, admitted, started, stepped, finished;
} demsph_job_info;
typedef struct {
    double wall;                    /* run() duration, s */
    double particleSteps;           /* sum over completed jobs of particles * steps */
    double particleStepsPerSecond;  /* particleSteps / wall */
    double p50, p99;                /* step latency percentiles over all completed jobs, s */
    int jobsDone, jobsFailed;
    size_t peakBytes;               /* peak admitted memory estimate */
    int peakActive;                 /* peak number of jobs on the device together */
} demsph_server_summary;
int demsph_server_create(demsph_server** out, int slots, double memoryBudgetBytes, int graphs);
int demsph_server_destroy(demsph_server* v);
/* Steps per step() call inside a job (default 100). */
int demsph_server_set_block(demsph_server* v, long steps);
/* Autotune every job's fluid scene on admission (demsph_sph_autotune with
 * the cache on; mode -1 off (default), 0 result-invariant knobs, 1 also
 * pair lanes), and every job's DEM scene (demsph_dem_autotune's "pairLanes",
 * result-invariant, once per size class of ball count, walls, bonds,
 * periodic axes and precision, with trialSteps of at least 5).  Jobs of
 * fewer than minSteps steps (demsph_server_set_autotune_min_steps, default
 * 2,000) are not tuned: a size class's trials cost some 40-50 steps of the
 * scene.  Set before demsph_server_run. */
int demsph_server_set_autotune(demsph_server* v, int mode, int trialSteps);
int demsph_server_set_autotune_min_steps(demsph_server* v, long minSteps);
/* fluid and/or dem non-NULL; fluidMin/Max with a fluid, demMin/Max with a
 * DEM scene (a coupled DEM scene usually shares the fluid bounds). */
int demsph_server_submit(demsph_server* v, demsph_sph* fluid, demsph_dem* dem, const double* fluidMin,
                         const double* fluidMax, const double* demMin, const double* demMax, const double* gravity,
                         double timeStep, long steps, double arrival, int* jobId);
/* Serves every submitted job; returns when all have ended.  Once per server. */
int demsph_server_run(demsph_server* v);
/* Progress of demsph_server_run (no counterpart in the reference): fn is
 * called from a slot thread each time a job finishes (completed or failed),
 * one call at a time, with the jobs finished so far, the jobs submitted and
 * the seconds since the run started.  NULL removes it; set before the run. */
typedef void (*demsph_server_progress_fn)(void* user, int jobsFinished, int jobsTotal, double seconds);
int demsph_server_set_progress(demsph_server* v, demsph_server_progress_fn fn, void* user);
int demsph_server_job_info(demsph_server* v, int jobId, demsph_job_info* info, char* error, size_t errorCap);
/* stepsDone - 1 intervals, s. */
int demsph_server_job_latencies(demsph_server* v, int jobId, double* out, size_t capacity, size_t* count);
/* Final state in insertion order: "position" "velocity" (fluid), "ballPosition"
 * "ballVelocity" (DEM); double, 3 per particle. */
int demsph_server_job_result(demsph_server* v, int jobId, const char* field, double* out, size_t capacity,
                             size_t* count);
int demsph_server_stats(demsph_server* v, demsph_server_summary* s);

#ifdef __cplusplus
}
#endif
#endif
