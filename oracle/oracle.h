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
 * What pins each function (tests/test_oracle_*.py unless noted; a pin is a
 * check the restatement could fail: a golden vector of the reference's own
 * run, a closed form, an invariant or brute force):
 *   o_grid_set            test_grid_set_closed_form (incl. a binary64 truncation case)
 *   o_hash                test_hash_known_cells_and_clamps (hand-computed cells, all clamps)
 *   o_sort_by_hash        test_sort_is_stable_lexicographic (numpy lexsort of (hash, index))
 *   o_cell_ranges         test_cell_ranges_brute_force
 *   o_stencil, o_sph_count/write
 *                         test_sph_pairs_match_brute_force (O(N^2) within the same stencil),
 *                         test_sph_pair_set_independent_of_cell_size (invariant)
 *   o_wendland            test_wendland_unit_integral (integral of W = 1)
 *   o_grad_wendland       test_grad_wendland_is_derivative_of_w, test_lattice_gradient_moment_is_identity
 *   o_sph dummy normals   test_dummy_normals_point_into_the_fluid_on_flat_walls (symmetry)
 *   o_sph step (acceleration, density change, EOS, kick/drift)
 *                         test_rest_lattice_has_no_density_change_in_interior,
 *                         test_hydrostatic_column_balances_gravity_in_interior,
 *                         test_tait_hydrostatic_column (Tait inversion + balance),
 *                         test_oracle_reproduces_reference_dam_break_frame1 (reference frame, step 670)
 *   o_dem neighbour search + history carry
 *                         test_ball_pairs_match_brute_force_stencil, test_spring_history_survives_rebuild
 *   linear contact, ball-wall contact, edge/vertex cancel
 *                         test_elastic_bounce_matches_reference_frames (reference frames, 10 decimals),
 *                         test_linear_collision_restitution_closed_form
 *   Hertz-Mindlin         test_hertz_collision_matches_closed_form
 *   Coulomb cap, rolling  test_sliding_ball_coulomb_cap_and_rolling_transition
 *   parallel bond         test_bonded_beam_tip_deflection_is_euler_bernoulli (elastic part)
 *   bond breakage (sigma > sigma_t, tau > C - mu sigma)
 *                         test_bond_breaks_at_closed_form_evaluation (tension, cohesion with
 *                         friction, bending, torsion, shear: the breaking evaluation in closed form)
 *   o_sphere_triangle     test_closest_point_on_triangle_brute_force
 *   mesh walls            test_cylinder_mesh_matches_reference_wall_frame (reference wall frame)
 *   wall motion (position, orientation, vertices, circumcentres, D3)
 *                         test_moving_wall_kinematics_closed_form
 *   clumps                test_rigid_clump_torque_free_rotation,
 *                         test_clump_force_sum_closed_form (also D1 against the intended sum),
 *                         test_clump_torque_sum_closed_form
 *   hooks                 test_buoyancy_hook_is_archimedes, test_global_damping_hook_closed_form,
 *                         constant force through the beam's Euler-Bernoulli deflection
 *   periodic boundaries (o_dem_set_periodic: wrapped stencil, images, wrap, nearest-image bonds)
 *                         test_periodic_pairs_match_minimum_image_brute_force,
 *                         test_periodic_sheet_is_translation_invariant,
 *                         test_rigid_translation_through_the_face (tests/test_oracle_periodic.py)
 *   scene builders        test_full_size_cantilever_keeps_every_lattice_bond,
 *                         test_full_size_ice_sheet_keeps_every_touching_bond (tests/test_scenes.py)
 * Parity unpinned: none of the functions above.  Clump-on-wall contact
 * reuses the pinned ball-wall and clump code paths and has no pin of its own.
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
void o_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
