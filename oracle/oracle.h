/*
 * FP64 CPU oracle for the DEMSPH per-step particle path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it; the product path never does.
 *
 * It restates, in plain C and double precision, the algorithms of the
 * reference solver (kaiqideng/this_studio, DEMSPH/kernel, the .cu sources).  Every
 * function names the reference file:line it follows.  Arrays are host
 * arrays owned by the caller unless a scene object owns them.  Vectors are
 * stored interleaved (x,y,z per element), the reference's double3 AoS.
 *
 * Policy on reference defects (see the function comments):
 *   D1  clump force sum uses the pebble force array (intended); switchable.
 *   D2  pointing ranges are reset every rebuild (intended).
 *   D3/D11/D16/D6/D7/D9/D10 are reproduced.
 *
 * Parity status of each function is recorded in the header comment of the
 * function ("pinned by" / "parity unpinned").
 */
#ifndef DEMSPH_ORACLE_H
#define DEMSPH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y, z; } o_vec3;

/* ------------------------------------------------------------------ grid */
typedef struct {
    double minx, miny, minz;
    double maxx, maxy, maxz;
    double cx, cy, cz;       /* cell size per axis */
    int nx, ny, nz;          /* grid size per axis */
} o_grid;

/* spatialGrid::set (DEMSPH/kernel/boundary.h:70-94).  Returns 0 on success,
 * -1 when the reference would silently ignore the call. */
int o_grid_set(o_grid* g, const double* minb, const double* maxb, double cellOneDim);

/* calculateHash kernel (neighborSearchKernel.cu:6-27): clamp then hash. */
void o_hash(const o_grid* g, const double* pos, int n, int* hash);

/* thrust::sort_by_key(hash, iota) (buildHashStartEnd.cu:36-83): stable,
 * i.e. lexicographic (hash, index).  Writes sorted hashes and indices. */
void o_sort_by_hash(const int* hash, int n, int numGrids, int* sortedHash, int* sortedIndex);

/* memset(0xFF) + findStartAndEnd (neighborSearchKernel.cu:105-106,
 * buildHashStartEnd.cu:14-34). */
void o_cell_ranges(const int* sortedHash, int n, int numGrids, int* start, int* end);

/* The 27-cell stencil of a query point (ballNeighborSearchKernel.cu:23-33):
 * unclamped truncation, then clamp to the grid and shrink at faces. */
void o_stencil(const o_grid* g, const double* p, int* cell, int* lo, int* hi);

/* ------------------------------------------------------------------ SPH */
/* Full SPH neighbour list, count pass (SPHNeighborSearchKernel.cu:6-55 for
 * A==B sets with excludeSelf=1; :117-167 for the fluid->dummy list with
 * excludeSelf=0).  Returns the total pair count. */
long o_sph_count(const o_grid* g, const double* posA, const double* hA, int nA,
                 const double* posB, const double* hB, const int* sortedIndexB,
                 const int* cellStart, const int* cellEnd, int excludeSelf, int* count);
/* Write pass (SPHNeighborSearchKernel.cu:57-115, :169-228). prefix is inclusive. */
void o_sph_write(const o_grid* g, const double* posA, const double* hA, int nA,
                 const double* posB, const double* hB, const int* sortedIndexB,
                 const int* cellStart, const int* cellEnd, int excludeSelf,
                 const int* prefix, int* pointed, int* pointing);

/* Wendland C2 gradient (WCSPHIntegrationKernel.cuh:15-32). */
void o_grad_wendland(const double* rij, double h, double* out);
/* Wendland C2 value (WCSPHIntegrationKernel.cuh:5-13). */
double o_wendland(double r, double h);

typedef struct o_sph o_sph;

/* WCSPHSolver scene (WCSPHSolver.h:18-636).  eos: 0 = linear p=c^2*drho
 * (reference), 1 = Tait gamma (extension).  Fluid fields are copied. */
o_sph* o_sph_create(int nf, const double* pos, const double* vel, const double* c,
                    const double* spacing, const double* rho0, const double* nu,
                    int nd, const double* dpos, const double* dvel, const double* dc,
                    const double* dspacing, const double* drho0, const double* dnu,
                    const double* minb, const double* maxb, const double* gravity,
                    int eos, double gamma, int nthreads);
void o_sph_destroy(o_sph* s);
/* initialize(): grid, dummy hash, dummy-dummy list, dummy normals
 * (WCSPHSolver.h:112-147, WCSPHIntegrationKernel.cu:4-40). */
int o_sph_initialize(o_sph* s);
/* neighborSearch() (WCSPHSolver.h:150-285). */
void o_sph_neighbor_search(o_sph* s);
/* interaction1stHalf / 2ndHalf (WCSPHIntegrationKernel.cu:308-472). */
void o_sph_first_half(o_sph* s, double dt);
void o_sph_second_half(o_sph* s, double dt);
/* One solve() loop iteration (WCSPHSolver.h:607-621), repeated nsteps. */
void o_sph_step(o_sph* s, double dt, int nsteps);
/* Field access.  name: "pos","vel","acc","drho","rho","p","dnormal",
 * "hash","index","cellStart","cellEnd","count_ff","count_fd","pointing_ff",
 * "pointing_fd","count_dd","pointing_dd".  Returns element count (doubles or
 * ints), copies into out when out != NULL. */
long o_sph_get(o_sph* s, const char* name, void* out);
int o_sph_set(o_sph* s, const char* name, const void* in);
void o_sph_grid(o_sph* s, o_grid* out);
/* OpenMP threads the oracle runs with. */
int o_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
