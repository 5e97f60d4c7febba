/*
 * FP64 CPU oracle: spatial grid, hash/sort/cell ranges, SPH neighbour
 * lists and the WCSPH solver step.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   o_grid_set/o_hash/o_sort_by_hash/o_cell_ranges: closed-form hash formula,
 *     std sort by (hash,index), brute-force cell ranges on random inputs.
 *   o_sph_count/o_sph_write: brute-force O(N^2) pair enumeration restricted
 *     to the same 27-cell stencil and cut; cell-size independence.
 *   o_grad_wendland/o_wendland: unit integral, finite-difference gradient,
 *     lattice gradient moment.
 *   o_sph_* step: rest lattice (zero density change, a = g in the interior),
 *     hydrostatic column, damBreak fixture (SPH_0000 lattice count).
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "oracle_common.h"
#include "oracle_dem.h"

int o_grid_set(o_grid* g, const double* minb, const double* maxb, double c)
{
    if (maxb[0] <= minb[0] || maxb[1] <= minb[1] || maxb[2] <= minb[2]) return -1;
    if (c < 1.e-20) return -1;
    double dx = maxb[0] - minb[0], dy = maxb[1] - minb[1], dz = maxb[2] - minb[2];
    g->minx = minb[0]; g->miny = minb[1]; g->minz = minb[2];
    g->maxx = maxb[0]; g->maxy = maxb[1]; g->maxz = maxb[2];
    g->nx = dx > c ? (int)(dx / c) : 1;
    g->ny = dy > c ? (int)(dy / c) : 1;
    g->nz = dz > c ? (int)(dz / c) : 1;
    g->cx = dx / (double)g->nx;
    g->cy = dy / (double)g->ny;
    g->cz = dz / (double)g->nz;
    return 0;
}

/* neighborSearchKernel.cu:6-27 with neighborSearchKernel.cuh:4-14 */
void o_hash(const o_grid* g, const double* pos, int n, int* hash)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        if (px < g->minx) px = g->minx;
        else if (px >= g->maxx) px = g->maxx - 0.5 * g->cx;
        if (py < g->miny) py = g->miny;
        else if (py >= g->maxy) py = g->maxy - 0.5 * g->cy;
        if (pz < g->minz) pz = g->minz;
        else if (pz >= g->maxz) pz = g->maxz - 0.5 * g->cz;
        int ix = (int)((px - g->minx) / g->cx);
        int iy = (int)((py - g->miny) / g->cy);
        int iz = (int)((pz - g->minz) / g->cz);
        hash[i] = o_hash_cell(g, ix, iy, iz);
    }
}

/* Stable sort by hash (lexicographic (hash, index)) as the reference's
 * thrust::sort_by_key over an iota value array (buildHashStartEnd.cu:50-76).
 * Counting sort for hashes in [0,numGrids); anything outside is ordered by
 * value through a fallback comparison sort. */
static int cmp_pair(const void* a, const void* b)
{
    const long long* x = (const long long*)a;
    const long long* y = (const long long*)b;
    return (*x > *y) - (*x < *y);
}

void o_sort_by_hash(const int* hash, int n, int numGrids, int* sortedHash, int* sortedIndex)
{
    int inRange = 1;
    for (int i = 0; i < n; i++)
        if (hash[i] < 0 || hash[i] >= numGrids) { inRange = 0; break; }
    /* sparse keys: a comparison sort is cheaper than a pass over every cell */
    if (inRange && (long)n * 8 < (long)numGrids) inRange = 0;
    if (inRange) {
        int* cnt = (int*)calloc((size_t)numGrids + 1, sizeof(int));
        for (int i = 0; i < n; i++) cnt[hash[i] + 1]++;
        for (int h = 0; h < numGrids; h++) cnt[h + 1] += cnt[h];
        for (int i = 0; i < n; i++) {
            int k = cnt[hash[i]]++;
            sortedHash[k] = hash[i];
            sortedIndex[k] = i;
        }
        free(cnt);
        return;
    }
    long long* keys = (long long*)malloc(sizeof(long long) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) keys[i] = ((long long)hash[i] << 32) | (unsigned)i;
    qsort(keys, (size_t)n, sizeof(long long), cmp_pair);
    for (int i = 0; i < n; i++) {
        sortedHash[i] = (int)(keys[i] >> 32);
        sortedIndex[i] = (int)(keys[i] & 0xffffffffLL);
    }
    free(keys);
}

/* buildHashStartEnd.cu:14-34 after the 0xFF memsets (neighborSearchKernel.cu:105-106) */
void o_cell_ranges(const int* sh, int n, int numGrids, int* start, int* end)
{
    memset(start, 0xFF, sizeof(int) * (size_t)numGrids);
    memset(end, 0xFF, sizeof(int) * (size_t)numGrids);
    for (int i = 0; i < n; i++) {
        int h = sh[i];
        if (h < 0 || h >= numGrids) continue;
        if (i == 0 || sh[i - 1] != h) start[h] = i;
        if (i == n - 1 || sh[i + 1] != h) end[h] = i + 1;
    }
}

/* ballNeighborSearchKernel.cu:23-33 / SPHNeighborSearchKernel.cu:21-31 */
void o_stencil(const o_grid* g, const double* p, int* cell, int* lo, int* hi)
{
    int gx = (int)((p[0] - g->minx) / g->cx);
    int gy = (int)((p[1] - g->miny) / g->cy);
    int gz = (int)((p[2] - g->minz) / g->cz);
    lo[0] = lo[1] = lo[2] = -1;
    hi[0] = hi[1] = hi[2] = 1;
    if (gx <= 0) { gx = 0; lo[0] = 0; }
    if (gx >= g->nx - 1) { gx = g->nx - 1; hi[0] = 0; }
    if (gy <= 0) { gy = 0; lo[1] = 0; }
    if (gy >= g->ny - 1) { gy = g->ny - 1; hi[1] = 0; }
    if (gz <= 0) { gz = 0; lo[2] = 0; }
    if (gz >= g->nz - 1) { gz = g->nz - 1; hi[2] = 0; }
    cell[0] = gx; cell[1] = gy; cell[2] = gz;
}

/* One SPH stencil walk; write==NULL counts only. */
static int sph_walk(const o_grid* g, const double* posA, const double* hA, int A,
                    const double* posB, const double* hB, const int* sortedIndexB,
                    const int* cellStart, const int* cellEnd, int excludeSelf,
                    int* pointed, int* pointing, int base)
{
    int c[3], lo[3], hi[3];
    o_vec3 pa = vld(posA, A);
    double ha = hA[A];
    o_stencil(g, &posA[3 * (long)A], c, lo, hi);
    int count = 0;
    for (int zz = lo[2]; zz <= hi[2]; zz++)
        for (int yy = lo[1]; yy <= hi[1]; yy++)
            for (int xx = lo[0]; xx <= hi[0]; xx++) {
                int hb = o_hash_cell(g, c[0] + xx, c[1] + yy, c[2] + zz);
                int s = cellStart[hb];
                if (s == -1) continue;
                int e = cellEnd[hb];
                for (int i = s; i < e; i++) {
                    int B = sortedIndexB[i];
                    if (excludeSelf && A == B) continue;
                    double cut = 2.0 * fmax(ha, hB[B]);
                    o_vec3 r = vsub(pa, vld(posB, B));
                    if ((cut * cut - vdot(r, r)) >= 0.) {
                        if (pointing) {
                            pointed[base + count] = A;
                            pointing[base + count] = B;
                        }
                        count++;
                    }
                }
            }
    return count;
}

long o_sph_count(const o_grid* g, const double* posA, const double* hA, int nA,
                 const double* posB, const double* hB, const int* sortedIndexB,
                 const int* cellStart, const int* cellEnd, int excludeSelf, int* count)
{
    long total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
    for (int A = 0; A < nA; A++) {
        count[A] = sph_walk(g, posA, hA, A, posB, hB, sortedIndexB, cellStart, cellEnd, excludeSelf, NULL, NULL, 0);
        total += count[A];
    }
    return total;
}

void o_sph_write(const o_grid* g, const double* posA, const double* hA, int nA,
                 const double* posB, const double* hB, const int* sortedIndexB,
                 const int* cellStart, const int* cellEnd, int excludeSelf,
                 const int* prefix, int* pointed, int* pointing)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int A = 0; A < nA; A++) {
        int base = A > 0 ? prefix[A - 1] : 0;
        sph_walk(g, posA, hA, A, posB, hB, sortedIndexB, cellStart, cellEnd, excludeSelf, pointed, pointing, base);
    }
}

/* WCSPHIntegrationKernel.cuh:15-32 */
static o_vec3 grad_w(o_vec3 rij, double h)
{
    double r = vlen(rij);
    if (r < 1.e-20 || r >= 2.0 * h) return v3(0, 0, 0);
    double q = r / h;
    double sigma = 21.0 / (16.0 * O_PI * h * h * h);
    double t = 1.0 - 0.5 * q;
    double dWdq = (-2.0 * pow(t, 3) * (1.0 + 2.0 * q) + 2.0 * pow(t, 4));
    double dWdr = sigma * dWdq / h;
    double f = dWdr / r;
    return vmul(rij, f);
}

void o_grad_wendland(const double* rij, double h, double* out)
{
    o_vec3 g = grad_w(v3(rij[0], rij[1], rij[2]), h);
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}

/* WCSPHIntegrationKernel.cuh:5-13 */
double o_wendland(double r, double h)
{
    double q = r / h;
    if (q >= 2.0) return 0.0;
    double sigma = 21.0 / (16.0 * O_PI * h * h * h);
    double t = 1.0 - 0.5 * q;
    return sigma * pow(t, 4) * (1.0 + 2.0 * q);
}

/* ------------------------------------------------------------------ scene */
struct o_sph {
    int nf, nd, eos, nthreads;
    double gamma;
    o_vec3 g;
    double minb[3], maxb[3];
    o_grid grid;
    int numGrids;
    /* fluid */
    double *pos, *vel, *acc, *drho, *rho, *p, *c, *m, *rho0, *h, *nu;
    /* dummy */
    double *dpos, *dvel, *dnormal, *dc, *dm, *drho0, *dh, *dnu, *ddens;
    int *hash, *index, *shash, *cellStart, *cellEnd, *dhash, *dindex, *dshash;
    int *count_ff, *prefix_ff, *pointed_ff, *pointing_ff;
    long cap_ff, n_ff;
    int *count_fd, *prefix_fd, *pointed_fd, *pointing_fd;
    long cap_fd, n_fd;
    int *count_dd, *prefix_dd, *pointed_dd, *pointing_dd;
    long cap_dd, n_dd;
    /* coupled DEM balls (solids), insertion order, copied from the DEM scene */
    int ns;
    double *spos, *svel, *sw, *sr, *sP, *szero, *sF, *acs;
    int *bhash, *bindex, *bshash;
    int *count_fs, *prefix_fs, *pointed_fs, *pointing_fs;
    long cap_fs, n_fs;
};

static double* dup_d(const double* src, long n, int stride)
{
    double* d = (double*)malloc(sizeof(double) * (size_t)(n * stride > 0 ? n * stride : 1));
    if (src) memcpy(d, src, sizeof(double) * (size_t)(n * stride));
    else memset(d, 0, sizeof(double) * (size_t)(n * stride));
    return d;
}

/* WCSPHSolver::addSPH / addDummy (WCSPHSolver.h:523-541): m = dx^3 rho0,
 * h = 1.3 dx, rho = rho0, p = 0, drho = 0.  Entries with rho0 <= 0 or c <= 0
 * are ignored by the reference; the caller filters them. */
o_sph* o_sph_create(int nf, const double* pos, const double* vel, const double* c,
                    const double* spacing, const double* rho0, const double* nu,
                    int nd, const double* dpos, const double* dvel, const double* dc,
                    const double* dspacing, const double* drho0, const double* dnu,
                    const double* minb, const double* maxb, const double* gravity,
                    int eos, double gamma, int nthreads)
{
    o_sph* s = (o_sph*)calloc(1, sizeof(o_sph));
    s->nf = nf; s->nd = nd; s->eos = eos; s->gamma = gamma; s->nthreads = nthreads;
    s->g = v3(gravity[0], gravity[1], gravity[2]);
    memcpy(s->minb, minb, sizeof(double) * 3);
    memcpy(s->maxb, maxb, sizeof(double) * 3);
    s->pos = dup_d(pos, nf, 3); s->vel = dup_d(vel, nf, 3); s->acc = dup_d(NULL, nf, 3);
    s->drho = dup_d(NULL, nf, 1); s->rho = dup_d(rho0, nf, 1); s->p = dup_d(NULL, nf, 1);
    s->c = dup_d(c, nf, 1); s->rho0 = dup_d(rho0, nf, 1); s->nu = dup_d(nu, nf, 1);
    s->m = dup_d(NULL, nf, 1); s->h = dup_d(NULL, nf, 1);
    for (int i = 0; i < nf; i++) {
        s->m[i] = spacing[i] * spacing[i] * spacing[i] * rho0[i];
        s->h[i] = 1.3 * spacing[i];
    }
    s->dpos = dup_d(dpos, nd, 3); s->dvel = dup_d(dvel, nd, 3); s->dnormal = dup_d(NULL, nd, 3);
    s->dc = dup_d(dc, nd, 1); s->drho0 = dup_d(drho0, nd, 1); s->dnu = dup_d(dnu, nd, 1);
    s->ddens = dup_d(drho0, nd, 1);
    s->dm = dup_d(NULL, nd, 1); s->dh = dup_d(NULL, nd, 1);
    for (int i = 0; i < nd; i++) {
        s->dm[i] = dspacing[i] * dspacing[i] * dspacing[i] * drho0[i];
        s->dh[i] = 1.3 * dspacing[i];
    }
    int nfa = nf > 0 ? nf : 1, nda = nd > 0 ? nd : 1;
    s->hash = (int*)malloc(sizeof(int) * nfa); s->index = (int*)malloc(sizeof(int) * nfa);
    s->shash = (int*)malloc(sizeof(int) * nfa);
    s->dhash = (int*)malloc(sizeof(int) * nda); s->dindex = (int*)malloc(sizeof(int) * nda);
    s->dshash = (int*)malloc(sizeof(int) * nda);
    s->count_ff = (int*)calloc(nfa, sizeof(int)); s->prefix_ff = (int*)calloc(nfa, sizeof(int));
    s->count_fd = (int*)calloc(nfa, sizeof(int)); s->prefix_fd = (int*)calloc(nfa, sizeof(int));
    s->count_dd = (int*)calloc(nda, sizeof(int)); s->prefix_dd = (int*)calloc(nda, sizeof(int));
    return s;
}

void o_sph_destroy(o_sph* s)
{
    if (!s) return;
    void* ptrs[] = {s->pos, s->vel, s->acc, s->drho, s->rho, s->p, s->c, s->m, s->rho0, s->h, s->nu,
                    s->dpos, s->dvel, s->dnormal, s->dc, s->dm, s->drho0, s->dh, s->dnu, s->ddens,
                    s->hash, s->index, s->shash, s->cellStart, s->cellEnd, s->dhash, s->dindex, s->dshash,
                    s->count_ff, s->prefix_ff, s->pointed_ff, s->pointing_ff,
                    s->count_fd, s->prefix_fd, s->pointed_fd, s->pointing_fd,
                    s->count_dd, s->prefix_dd, s->pointed_dd, s->pointing_dd,
                    s->spos, s->svel, s->sw, s->sr, s->sP, s->szero, s->sF, s->acs, s->bhash, s->bindex, s->bshash,
                    s->count_fs, s->prefix_fs, s->pointed_fs, s->pointing_fs};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(s);
}

static void inclusive_scan(const int* c, int n, int* p)
{
    long acc = 0;
    for (int i = 0; i < n; i++) { acc += c[i]; p[i] = (int)acc; }
}

/* launchUpdateSpatialGridCellHashStartEnd (neighborSearchKernel.cu:79-116) */
static void build_cells(o_sph* s, const double* pos, int n, int* hash, int* shash, int* index)
{
    o_hash(&s->grid, pos, n, hash);
    o_sort_by_hash(hash, n, s->numGrids, shash, index);
    o_cell_ranges(shash, n, s->numGrids, s->cellStart, s->cellEnd);
}

static void build_list(o_sph* s, const double* posA, const double* hA, int nA,
                       const double* posB, const double* hB, const int* idxB, int excludeSelf,
                       int* count, int* prefix, int** pointed, int** pointing, long* cap, long* npairs)
{
    long tot = o_sph_count(&s->grid, posA, hA, nA, posB, hB, idxB, s->cellStart, s->cellEnd, excludeSelf, count);
    inclusive_scan(count, nA, prefix);
    if (tot > *cap || *pointing == NULL) {
        free(*pointed); free(*pointing);
        *cap = tot > 0 ? tot : 1;
        *pointed = (int*)malloc(sizeof(int) * (size_t)*cap);
        *pointing = (int*)malloc(sizeof(int) * (size_t)*cap);
    }
    o_sph_write(&s->grid, posA, hA, nA, posB, hB, idxB, s->cellStart, s->cellEnd, excludeSelf, prefix, *pointed, *pointing);
    *npairs = tot;
}

/* calDummyParticleNormalKernel (WCSPHIntegrationKernel.cu:4-40) */
static void dummy_normals(o_sph* s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->nd; i++) {
        o_vec3 ri = vld(s->dpos, i);
        double hi = s->dh[i];
        o_vec3 th = v3(0, 0, 0);
        int st = i > 0 ? s->prefix_dd[i - 1] : 0, en = s->prefix_dd[i];
        for (int k = st; k < en; k++) {
            int j = s->pointing_dd[k];
            o_vec3 dw = grad_w(vsub(ri, vld(s->dpos, j)), 0.5 * (hi + s->dh[j]));
            th = vsub(th, vmul(dw, s->dm[j] / s->ddens[j]));
        }
        o_vec3 n = v3(0, 0, 0);
        if (vlen2(th) > 1.e-10) n = vdiv(th, vlen(th));
        vst(s->dnormal, i, n);
    }
}

int o_sph_initialize(o_sph* s)
{
    double c = 0.0;
    for (int i = 0; i < s->nf; i++) c = s->h[i] * 2.0 > c ? s->h[i] * 2.0 : c;
    double cd = 0.0;
    for (int i = 0; i < s->nd; i++) cd = s->dh[i] * 2.0 > cd ? s->dh[i] * 2.0 : cd;
    if (s->nd > 0) c = c > cd ? c : cd;
    if (o_grid_set(&s->grid, s->minb, s->maxb, c) != 0) return -1;
    s->numGrids = s->grid.nx * s->grid.ny * s->grid.nz;
    free(s->cellStart); free(s->cellEnd);
    s->cellStart = (int*)malloc(sizeof(int) * (size_t)s->numGrids);
    s->cellEnd = (int*)malloc(sizeof(int) * (size_t)s->numGrids);
    if (s->nd > 0) {
        build_cells(s, s->dpos, s->nd, s->dhash, s->dshash, s->dindex);
        build_list(s, s->dpos, s->dh, s->nd, s->dpos, s->dh, s->dindex, 1, s->count_dd, s->prefix_dd,
                   &s->pointed_dd, &s->pointing_dd, &s->cap_dd, &s->n_dd);
        dummy_normals(s);
    }
    return 0;
}

/* WCSPHSolver::neighborSearch (WCSPHSolver.h:150-285).  Dummies are re-hashed
 * into the same grid arrays after the fluid list is written (D20). */
void o_sph_neighbor_search(o_sph* s)
{
    build_cells(s, s->pos, s->nf, s->hash, s->shash, s->index);
    build_list(s, s->pos, s->h, s->nf, s->pos, s->h, s->index, 1, s->count_ff, s->prefix_ff,
               &s->pointed_ff, &s->pointing_ff, &s->cap_ff, &s->n_ff);
    if (s->nd == 0) { memset(s->count_fd, 0, sizeof(int) * (size_t)s->nf); memset(s->prefix_fd, 0, sizeof(int) * (size_t)s->nf); s->n_fd = 0; return; }
    build_cells(s, s->dpos, s->nd, s->dhash, s->dshash, s->dindex);
    build_list(s, s->pos, s->h, s->nf, s->dpos, s->dh, s->dindex, 0, s->count_fd, s->prefix_fd,
               &s->pointed_fd, &s->pointing_fd, &s->cap_fd, &s->n_fd);
}

/* Inverse EOS used for the dummy density in the reference (linear,
 * WCSPHIntegrationKernel.cu:132,249) and its Tait counterpart. */
static double dummy_density(const o_sph* s, double PR, double cj, double rho0j)
{
    if (s->eos == 1) {
        double B = rho0j * cj * cj / s->gamma;
        double base = 1.0 + PR / B;
        if (base < 1.e-12) base = 1.e-12;
        return rho0j * pow(base, 1.0 / s->gamma);
    }
    return PR / (cj * cj) + rho0j;
}

/* DEM-SPH coupling (no reference counterpart; the project's model, the
 * FP64 statement of the GPU library's): coupled ball j is a solid particle of
 * fluid i's phase with mass rho0_i 4/3 pi r_j^3, pressure P_j interpolated
 * from the fluid, density from the inverse EOS of P_j, moving with its
 * surface velocity v_j + w_j x (r_j n), n = unit(x_i - x_j); kernel support
 * 2 h_i; the Riemann interface pressure of the fluid-fluid branch with
 * P_R = P_j.  Returns fluid i's acceleration contribution split into the
 * central pressure part *ap and the viscous part *av, and the lever r_j n. */
static void fs_pair(const o_sph* s, int i, int j, o_vec3* ap, o_vec3* av, o_vec3* lever, double* drho_rate)
{
    o_vec3 ri = vld(s->pos, i), vi = vld(s->vel, i);
    double ci = s->c[i], hi = s->h[i], nui = s->nu[i], rhoi = s->rho[i], PL = s->p[i], r0 = s->rho0[i];
    o_vec3 rj = vld(s->spos, j), vj = vld(s->svel, j), wj = vld(s->sw, j);
    double rad = s->sr[j], PR = s->sP[j];
    o_vec3 rij = vsub(ri, rj);
    o_vec3 n = vnormalize(rij);
    o_vec3 l = vmul(n, rad);
    o_vec3 vs = vadd(vj, vcross(wj, l));
    o_vec3 e = vneg(n);
    double UL = vdot(vi, e), UR = vdot(vs, e);
    double rhoR = dummy_density(s, PR, ci, r0);
    double mj = r0 * (4.0 / 3.0 * O_PI) * rad * rad * rad;
    o_vec3 dw = grad_w(rij, hi);
    double cbar = ci, rbar = 0.5 * (rhoi + rhoR);
    double beta = fmin(3.0 * fmax(UL - UR, 0.0), cbar);
    double Ps = 0.5 * (PL + PR) + 0.5 * beta * rbar * (UL - UR);
    *ap = vneg(vmul(dw, 2.0 * mj * (Ps / (rhoi * rhoR))));
    *av = vmul(vsub(vi, vs), 2.0 * nui * vdot(n, dw) * mj / (rhoi * rhoR));
    *lever = l;
    if (drho_rate) {
        o_vec3 vstar = vmul(vadd(vi, vs), 0.5);
        vstar = vadd(vstar, vmul(e, 0.5 * ((PL - PR) / (rbar * cbar))));
        *drho_rate = mj / rhoR * vdot(vsub(vi, vstar), dw);
    }
}

/* updateWCSPHAccelerationKernel (WCSPHIntegrationKernel.cu:42-148) */
static void acceleration(o_sph* s)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->nf; i++) {
        o_vec3 ri = vld(s->pos, i), vi = vld(s->vel, i);
        double ci = s->c[i], hi = s->h[i], nui = s->nu[i], rhoi = s->rho[i], PL = s->p[i];
        o_vec3 a = v3(0, 0, 0);
        int st = i > 0 ? s->prefix_ff[i - 1] : 0, en = s->prefix_ff[i];
        for (int k = st; k < en; k++) {
            int j = s->pointing_ff[k];
            o_vec3 rj = vld(s->pos, j), vj = vld(s->vel, j);
            double cj = s->c[j], hj = s->h[j], nuj = s->nu[j], rhoj = s->rho[j], PR = s->p[j], mj = s->m[j];
            o_vec3 rij = vsub(ri, rj);
            o_vec3 e = vneg(vnormalize(rij));
            double UL = vdot(vi, e), UR = vdot(vj, e);
            o_vec3 dw = grad_w(rij, 0.5 * (hi + hj));
            double cbar = 0.5 * (ci + cj), rbar = 0.5 * (rhoi + rhoj);
            double beta = fmin(3.0 * fmax(UL - UR, 0.0), cbar);
            double Ps = 0.5 * (PL + PR) + 0.5 * beta * rbar * (UL - UR);
            a = vsub(a, vmul(dw, 2.0 * mj * (Ps / (rhoi * rhoj))));
            double nuij = 0.5 * (nui + nuj);
            a = vadd(a, vmul(vsub(vi, vj), 2.0 * nuij * vdot(vnormalize(rij), dw) * mj / (rhoi * rhoj)));
        }
        st = i > 0 ? s->prefix_fd[i - 1] : 0; en = s->prefix_fd[i];
        for (int k = st; k < en; k++) {
            int j = s->pointing_fd[k];
            o_vec3 rj = vld(s->dpos, j), vj = vld(s->dvel, j), nw = vld(s->dnormal, j);
            double cj = s->dc[j], hj = s->dh[j], nuj = s->dnu[j], mj = s->dm[j];
            double UL = vdot(vneg(nw), vi);
            double UR = -UL + 2.0 * vdot(vneg(nw), vj);
            double grj = vdot(s->g, vsub(rj, ri));
            if (grj < 0.) grj = 0.;
            double PR = PL + rhoi * grj;
            double rhoj = dummy_density(s, PR, cj, s->drho0[j]);
            o_vec3 rij = vsub(ri, rj);
            o_vec3 dw = grad_w(rij, 0.5 * (hi + hj));
            double cbar = 0.5 * (ci + cj), rbar = 0.5 * (rhoi + rhoj);
            double beta = fmin(3.0 * fmax(UL - UR, 0.0), cbar);
            double Ps = 0.5 * (PL + PR) + 0.5 * beta * rbar * (UL - UR);
            a = vsub(a, vmul(dw, 2.0 * mj * (Ps / (rhoi * rhoj))));
            double nuij = 0.5 * (nui + nuj);
            a = vadd(a, vmul(vsub(vi, vj), 2.0 * nuij * vdot(vnormalize(rij), dw) * mj / (rhoi * rhoj)));
        }
        if (s->ns > 0) {
            o_vec3 ac = v3(0, 0, 0);
            st = i > 0 ? s->prefix_fs[i - 1] : 0; en = s->prefix_fs[i];
            for (int k = st; k < en; k++) {
                o_vec3 ap, av, l;
                fs_pair(s, i, s->pointing_fs[k], &ap, &av, &l, NULL);
                ac = vadd(ac, vadd(ap, av));
            }
            vst(s->acs, i, ac);
            a = vadd(a, ac);
        }
        vst(s->acc, i, vadd(a, s->g));
    }
}

/* updateWCSPHDensityChangeKernel (WCSPHIntegrationKernel.cu:172-264) */
static void density_change(o_sph* s, double dt)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s->nf; i++) {
        o_vec3 ri = vld(s->pos, i), vi = vld(s->vel, i);
        double ci = s->c[i], hi = s->h[i], rhoi = s->rho[i], PL = s->p[i];
        double d = 0.0;
        int st = i > 0 ? s->prefix_ff[i - 1] : 0, en = s->prefix_ff[i];
        for (int k = st; k < en; k++) {
            int j = s->pointing_ff[k];
            o_vec3 rj = vld(s->pos, j), vj = vld(s->vel, j);
            double cj = s->c[j], hj = s->h[j], rhoj = s->rho[j], PR = s->p[j], mj = s->m[j];
            o_vec3 rij = vsub(ri, rj);
            o_vec3 dw = grad_w(rij, 0.5 * (hi + hj));
            o_vec3 e = vneg(vnormalize(rij));
            double cbar = 0.5 * (ci + cj), rbar = 0.5 * (rhoi + rhoj);
            o_vec3 vs = vmul(vadd(vi, vj), 0.5);
            vs = vadd(vs, vmul(e, 0.5 * ((PL - PR) / (rbar * cbar))));
            d += mj / rhoj * vdot(vsub(vi, vs), dw);
        }
        st = i > 0 ? s->prefix_fd[i - 1] : 0; en = s->prefix_fd[i];
        for (int k = st; k < en; k++) {
            int j = s->pointing_fd[k];
            o_vec3 rj = vld(s->dpos, j), vj = vld(s->dvel, j);
            double cj = s->dc[j], hj = s->dh[j], mj = s->dm[j];
            double grj = vdot(s->g, vsub(rj, ri));
            if (grj < 0.) grj = 0.;
            double PR = PL + rhoi * grj;
            double rhoj = dummy_density(s, PR, cj, s->drho0[j]);
            o_vec3 rij = vsub(ri, rj);
            o_vec3 dw = grad_w(rij, 0.5 * (hi + hj));
            o_vec3 e = vneg(vnormalize(rij));
            double cbar = 0.5 * (ci + cj), rbar = 0.5 * (rhoi + rhoj);
            o_vec3 vs = vmul(vadd(vi, vj), 0.5);
            vs = vadd(vs, vmul(e, 0.5 * ((PL - PR) / (rbar * cbar))));
            d += mj / rhoj * vdot(vsub(vi, vs), dw);
        }
        if (s->ns > 0) {
            st = i > 0 ? s->prefix_fs[i - 1] : 0; en = s->prefix_fs[i];
            for (int k = st; k < en; k++) {
                o_vec3 ap, av, l;
                double rate;
                fs_pair(s, i, s->pointing_fs[k], &ap, &av, &l, &rate);
                d += rate;
            }
        }
        d *= 2.0 * rhoi;
        s->drho[i] += d * dt;
    }
}

/* SPHDensityPressureIntegrationKernel (WCSPHIntegrationKernel.cu:266-282),
 * plus the Tait option p = B((rho/rho0)^gamma - 1), B = rho0 c^2 / gamma. */
static void eos(o_sph* s)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s->nf; i++) {
        double ch = s->drho[i], r0 = s->rho0[i];
        if (ch < -0.1 * r0) ch = -0.1 * r0;
        s->rho[i] = ch + r0;
        double c = s->c[i];
        if (s->eos == 1) {
            double B = r0 * c * c / s->gamma;
            s->p[i] = B * (pow(s->rho[i] / r0, s->gamma) - 1.0);
        } else {
            s->p[i] = c * c * ch;
        }
    }
}

static void kick(o_sph* s, double h)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < 3 * s->nf; i++) s->vel[i] += h * s->acc[i];
}

static void drift(o_sph* s, double dt)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < 3 * s->nf; i++) s->pos[i] += dt * s->vel[i];
}

/* launchWCSPH1stHalfIntegration (WCSPHIntegrationKernel.cu:308-376) */
void o_sph_first_half(o_sph* s, double dt)
{
    acceleration(s);
    kick(s, 0.5 * dt);
    drift(s, dt);
}

/* launchWCSPH2ndHalfIntegration (WCSPHIntegrationKernel.cu:378-472) */
void o_sph_second_half(o_sph* s, double dt)
{
    density_change(s, dt);
    eos(s);
    acceleration(s);
    kick(s, 0.5 * dt);
}

void o_sph_step(o_sph* s, double dt, int nsteps)
{
    for (int k = 0; k < nsteps; k++) {
        o_sph_neighbor_search(s);
        o_sph_first_half(s, dt);
        o_sph_second_half(s, dt);
    }
}

void o_sph_grid(o_sph* s, o_grid* out) { *out = s->grid; }

long o_sph_get(o_sph* s, const char* name, void* out)
{
#define GETD(nm, arr, cnt) if (!strcmp(name, nm)) { if (out) memcpy(out, arr, sizeof(double) * (size_t)(cnt)); return (long)(cnt); }
#define GETI(nm, arr, cnt) if (!strcmp(name, nm)) { if (out && (cnt) > 0) memcpy(out, arr, sizeof(int) * (size_t)(cnt)); return (long)(cnt); }
    GETD("pos", s->pos, 3L * s->nf)
    GETD("vel", s->vel, 3L * s->nf)
    GETD("acc", s->acc, 3L * s->nf)
    GETD("drho", s->drho, s->nf)
    GETD("rho", s->rho, s->nf)
    GETD("p", s->p, s->nf)
    GETD("dnormal", s->dnormal, 3L * s->nd)
    GETD("h", s->h, s->nf)
    GETD("m", s->m, s->nf)
    GETI("hash", s->hash, s->nf)
    GETI("index", s->index, s->nf)
    GETI("dindex", s->dindex, s->nd)
    GETI("cellStart", s->cellStart, s->numGrids)
    GETI("cellEnd", s->cellEnd, s->numGrids)
    GETI("count_ff", s->count_ff, s->nf)
    GETI("count_fd", s->count_fd, s->nf)
    GETI("count_dd", s->count_dd, s->nd)
    GETI("pointing_ff", s->pointing_ff, s->n_ff)
    GETI("pointing_fd", s->pointing_fd, s->n_fd)
    GETI("pointing_dd", s->pointing_dd, s->n_dd)
#undef GETD
#undef GETI
    return -1;
}

int o_sph_set(o_sph* s, const char* name, const void* in)
{
    if (!strcmp(name, "pos")) { memcpy(s->pos, in, sizeof(double) * 3 * (size_t)s->nf); return 0; }
    if (!strcmp(name, "vel")) { memcpy(s->vel, in, sizeof(double) * 3 * (size_t)s->nf); return 0; }
    if (!strcmp(name, "drho")) {
        memcpy(s->drho, in, sizeof(double) * (size_t)s->nf);
        eos(s);
        return 0;
    }
    return -1;
}

/* Threads an OpenMP parallel region of the oracle uses (the CPU baseline's
 * "cores"). */
#ifdef _OPENMP
#include <omp.h>
#endif
int o_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Sets the threads of the oracle's parallel regions (the CPU baseline's
 * one-thread run). */
void o_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------- coupling */
static void couple_load(o_sph* s, o_dem* d)
{
    o_dem_balls b;
    o_dem_couple_view(d, &b);
    if (s->ns != b.n) {
        void* old[] = {s->spos, s->svel, s->sw, s->sr, s->sP, s->szero, s->sF, s->acs, s->bhash, s->bindex,
                       s->bshash};
        for (size_t q = 0; q < sizeof(old) / sizeof(old[0]); q++) free(old[q]);
        s->ns = b.n;
        int na = b.n > 0 ? b.n : 1;
        s->spos = dup_d(NULL, na, 3); s->svel = dup_d(NULL, na, 3); s->sw = dup_d(NULL, na, 3);
        s->sr = dup_d(NULL, na, 1); s->sP = dup_d(NULL, na, 1); s->szero = dup_d(NULL, na, 1);
        s->sF = dup_d(NULL, na, 6); s->acs = dup_d(NULL, s->nf > 0 ? s->nf : 1, 3);
        s->bhash = (int*)malloc(sizeof(int) * na); s->bindex = (int*)malloc(sizeof(int) * na);
        s->bshash = (int*)malloc(sizeof(int) * na);
        if (!s->count_fs) {
            int nfa = s->nf > 0 ? s->nf : 1;
            s->count_fs = (int*)calloc(nfa, sizeof(int)); s->prefix_fs = (int*)calloc(nfa, sizeof(int));
        }
    }
    memcpy(s->spos, b.pos, sizeof(double) * 3 * (size_t)b.n);
    memcpy(s->svel, b.vel, sizeof(double) * 3 * (size_t)b.n);
    memcpy(s->sw, b.w, sizeof(double) * 3 * (size_t)b.n);
    memcpy(s->sr, b.radius, sizeof(double) * (size_t)b.n);
}

/* Fluid -> ball lists on the fluid grid: every ball within 2 h_i of fluid
 * particle i (hB = 0 makes the cut 2 max(h_i, 0)). */
static void couple_lists(o_sph* s)
{
    build_cells(s, s->spos, s->ns, s->bhash, s->bshash, s->bindex);
    build_list(s, s->pos, s->h, s->nf, s->spos, s->szero, s->bindex, 0, s->count_fs, s->prefix_fs,
               &s->pointed_fs, &s->pointing_fs, &s->cap_fs, &s->n_fs);
}

/* Shepard interpolation of the fluid pressure at each ball centre over the
 * fluid particles that list the ball. */
static void couple_pressure(o_sph* s)
{
    double* num = (double*)calloc((size_t)s->ns, sizeof(double));
    double* den = (double*)calloc((size_t)s->ns, sizeof(double));
    for (int i = 0; i < s->nf; i++) {
        int st = i > 0 ? s->prefix_fs[i - 1] : 0, en = s->prefix_fs[i];
        for (int k = st; k < en; k++) {
            int j = s->pointing_fs[k];
            o_vec3 rij = vsub(vld(s->pos, i), vld(s->spos, j));
            double w = o_wendland(sqrt(vdot(rij, rij)), s->h[i]) * (s->m[i] / s->rho[i]);
            num[j] += w * s->p[i];
            den[j] += w;
        }
    }
    for (int j = 0; j < s->ns; j++) s->sP[j] = den[j] > 0.0 ? num[j] / den[j] : 0.0;
    free(num);
    free(den);
}

/* The reaction -m_i (ap + av) and torque (r_j n) x (-m_i av) of every
 * fluid-ball pair on the ball, added to the DEM accumulators. */
static void couple_forces(o_sph* s, o_dem* d)
{
    o_dem_balls b;
    o_dem_couple_view(d, &b);
    memset(s->sF, 0, sizeof(double) * 6 * (size_t)s->ns);
    for (int i = 0; i < s->nf; i++) {
        int st = i > 0 ? s->prefix_fs[i - 1] : 0, en = s->prefix_fs[i];
        for (int k = st; k < en; k++) {
            int j = s->pointing_fs[k];
            o_vec3 ap, av, l;
            fs_pair(s, i, j, &ap, &av, &l, NULL);
            o_vec3 F = vmul(vadd(ap, av), -s->m[i]);
            o_vec3 T = vcross(l, vmul(av, -s->m[i]));
            for (int q = 0; q < 3; q++) {
                s->sF[6 * j + q] += (&F.x)[q];
                s->sF[6 * j + 3 + q] += (&T.x)[q];
            }
        }
    }
    for (int j = 0; j < s->ns; j++)
        for (int q = 0; q < 3; q++) {
            b.force[3 * j + q] += s->sF[6 * j + q];
            b.torque[3 * j + q] += s->sF[6 * j + 3 + q];
        }
}

/* One coupled step per iteration, the sequence of the GPU library's
 * demsph_couple_step: DEM step begin | fluid search | balls, lists,
 * pressures | DEM forces + reaction | DEM hook, kick, drift | fluid
 * momentum, kick, drift | ball kinematics | continuity + EOS | pressures |
 * DEM forces + reaction | DEM hook, kick | fluid momentum, kick. */
void o_couple_step(o_sph* s, o_dem* d, const double* g, double dt, int nsteps)
{
    double h = 0.5 * dt;
    for (int k = 0; k < nsteps; k++) {
        double t = o_dem_couple_begin(d, dt);
        o_sph_neighbor_search(s);
        couple_load(s, d);
        couple_lists(s);
        couple_pressure(s);
        o_dem_couple_evaluate(d, h, dt);
        couple_forces(s, d);
        o_dem_couple_kick(d, 1, g, h, t - h);
        acceleration(s);
        kick(s, h);
        drift(s, dt);
        couple_load(s, d);
        density_change(s, dt);
        eos(s);
        couple_pressure(s);
        o_dem_couple_evaluate(d, h, dt);
        couple_forces(s, d);
        o_dem_couple_kick(d, 0, g, h, t);
        acceleration(s);
        kick(s, h);
    }
}

/* "couplingForce" "couplingTorque" (3 per ball), "ballPressure" (1 per
 * ball), "fluidCouplingAcceleration" (3 per fluid particle). */
long o_couple_get(o_sph* s, const char* name, double* out)
{
    if (!strcmp(name, "couplingForce") || !strcmp(name, "couplingTorque")) {
        int off = name[8] == 'F' ? 0 : 3;
        if (out)
            for (int j = 0; j < s->ns; j++)
                for (int q = 0; q < 3; q++) out[3 * j + q] = s->sF[6 * j + off + q];
        return 3L * s->ns;
    }
    if (!strcmp(name, "ballPressure")) {
        if (out) memcpy(out, s->sP, sizeof(double) * (size_t)s->ns);
        return s->ns;
    }
    if (!strcmp(name, "fluidCouplingAcceleration")) {
        if (out) memcpy(out, s->acs, sizeof(double) * 3 * (size_t)s->nf);
        return 3L * s->nf;
    }
    return -1;
}

