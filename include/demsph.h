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

/* Overwrites position / velocity (double, 3n) or densityChange (double, n)
 * by insertion index (density and pressure follow from the EOS). */
int demsph_sph_set(demsph_sph* s, const char* field, const void* in, size_t count);

/* Kernel timing: when enabled, every kernel launch of the scene is
 * bracketed by HIP events on the scene stream and accumulated per kernel
 * name.  demsph_sph_timing reads (name, launches, total ms) triples. */
int demsph_sph_enable_timing(demsph_sph* s, int enable);
int demsph_sph_timing(demsph_sph* s, int index, char* name, size_t nameCap, long* launches, double* totalMs);

#ifdef __cplusplus
}
#endif
#endif
