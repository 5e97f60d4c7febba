/*
 * FP64 CPU oracle of the DEMSolver particle path.  TEST INFRASTRUCTURE ONLY.
 *
 * Restated from the reference (kaiqideng/this_studio, DEMSPH/kernel):
 *   pair lists + history carry   ballNeighborSearchKernel.cu:6-339
 *   pointing ranges              DEMSolver.h:59-78 (D2: reset every build)
 *   contact models               contactKernel.cuh:64-290, contactKernel.cu:135-259
 *   sphere-triangle              contactKernel.cuh:292-501, contactKernel.cu:261-533
 *   parallel bonds               contactKernel.cuh:4-62, contactKernel.cu:535-655
 *   reduction                    contactKernel.cu:657-714
 *   integration                  ballIntegrationKernel.cu:3-183, wallIntegrationKernel.cu(h)
 *   step sequencing              DEMSolver.h:428-906, 1376-1485, 1961-2049
 *   hook kernels                 DEMSPH/externalForceTorque.cu:4-161
 *
 * Parity pins (tests/test_oracle_dem.py): elasticBounce frames (linear
 * contact, free fall, wall classification), Hertz closed-form force and
 * contact time, linear restitution closed form, Coulomb cap on a sliding
 * ball, cantilever beam deflection (Euler-Bernoulli), brute-force pair
 * sets, history continuity across rebuilds, brute-force closest point on a
 * triangle, rigid clump invariants, Archimedes buoyancy.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "oracle_common.h"
#include "oracle_dem.h"

/* ------------------------------------------------------------- growable */
typedef struct { double* p; long n, cap; } dvec;
typedef struct { int* p; long n, cap; } ivec;
static void dpush(dvec* v, double x)
{
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 64; v->p = (double*)realloc(v->p, sizeof(double) * v->cap); }
    v->p[v->n++] = x;
}
static void dpush3(dvec* v, o_vec3 a) { dpush(v, a.x); dpush(v, a.y); dpush(v, a.z); }
static void ipush(ivec* v, int x)
{
    if (v->n == v->cap) { v->cap = v->cap ? 2 * v->cap : 64; v->p = (int*)realloc(v->p, sizeof(int) * v->cap); }
    v->p[v->n++] = x;
}

/* ------------------------------------------------------------- params */
enum { H_E = 0, H_G, H_RES, H_KRKS, H_KTKS, H_MUS, H_MUR, H_MUT, H_N };
enum { L_KN = 0, L_KS, L_KR, L_KT, L_DN, L_DS, L_DR, L_DT, L_MUS, L_MUR, L_MUT, L_N };
enum { B_GAMMA = 0, B_E, B_KNKS, B_SIG, B_C, B_MU, B_N };

typedef struct { int a, b; double p[11]; } row_t;

/* solidInteraction (DEMSolver.h:16-79) */
typedef struct {
    long n, cap;          /* numActivated_, device capacity */
    long nold;
    int *pointed, *pointing, *hashIndex, *cancel;
    int* shift;           /* periodic image code of the pointing object per pair */
    double *sl, *ro, *to; /* springs */
    double *F, *T, *P, *N, *ov;
    int *opointed, *ohashIndex, *ocancel;
    double *osl, *oro, *oto;
    int *count, *prefix;  /* per pointed object */
    int *istart, *iend;   /* per pointing object */
    int npointed, npointing;
} inter_t;

struct o_dem {
    /* balls (particle.h ball) */
    dvec pos, vel, w, force, torque, radius, invMass;
    ivec mat, clump;
    /* clumps */
    dvec cpos, cvel, cw, cforce, ctorque, cq, cinvI, cinvMass;
    ivec cmat, cps, cpe;
    /* wall bodies */
    dvec wpos, wvel, ww, wq;
    ivec wmat;
    /* vertices / triangles */
    dvec vlocal, vglobal;
    ivec vwall;
    ivec t0, t1, t2, twall;
    dvec tcc;
    /* bonds */
    ivec bpointed, bpointing, bonded;
    dvec bpoint, bnormal, bshear, bbend, bFn, bTt, bsig, btau;
    /* materials */
    row_t *lin, *hz, *bd;
    int nlin, nhz, nbd;
    int nMat, cap;
    double *H, *L, *B;
    /* grids */
    o_grid bg, tg;
    int bgValid, tgValid;
    long bng, tng;
    int *bhash, *bindex, *bshash, *bcs, *bce;
    int *thash, *tindex, *tshash, *tcs, *tce;
    inter_t bb, bt;
    /* loop state */
    size_t iStep;
    double time;
    int policyD1;
    /* periodic axes (setPeriodicBoundary, DEMSolver.h:1944-1959) and periods */
    int per[3];
    double Lp[3];
    /* hook program */
    int nhook;
    struct { int kind; double c, rho, cd, level, ramp; o_vec3 g, vf; double* f; long nf; } hook[8];
};

o_dem* o_dem_create(void) { return (o_dem*)calloc(1, sizeof(o_dem)); }

static void inter_free(inter_t* I)
{
    void* p[] = {I->pointed, I->pointing, I->hashIndex, I->cancel, I->shift, I->sl, I->ro, I->to, I->F, I->T, I->P, I->N, I->ov,
                 I->opointed, I->ohashIndex, I->ocancel, I->osl, I->oro, I->oto, I->count, I->prefix, I->istart, I->iend};
    for (size_t k = 0; k < sizeof(p) / sizeof(p[0]); k++) free(p[k]);
    memset(I, 0, sizeof(*I));
}

void o_dem_destroy(o_dem* d)
{
    if (!d) return;
    dvec* dv[] = {&d->pos, &d->vel, &d->w, &d->force, &d->torque, &d->radius, &d->invMass, &d->cpos, &d->cvel, &d->cw,
                  &d->cforce, &d->ctorque, &d->cq, &d->cinvI, &d->cinvMass, &d->wpos, &d->wvel, &d->ww, &d->wq,
                  &d->vlocal, &d->vglobal, &d->tcc, &d->bpoint, &d->bnormal, &d->bshear, &d->bbend, &d->bFn, &d->bTt,
                  &d->bsig, &d->btau};
    for (size_t k = 0; k < sizeof(dv) / sizeof(dv[0]); k++) free(dv[k]->p);
    ivec* iv[] = {&d->mat, &d->clump, &d->cmat, &d->cps, &d->cpe, &d->wmat, &d->vwall, &d->t0, &d->t1, &d->t2,
                  &d->twall, &d->bpointed, &d->bpointing, &d->bonded};
    for (size_t k = 0; k < sizeof(iv) / sizeof(iv[0]); k++) free(iv[k]->p);
    free(d->lin); free(d->hz); free(d->bd); free(d->H); free(d->L); free(d->B);
    free(d->bhash); free(d->bindex); free(d->bshash); free(d->bcs); free(d->bce);
    free(d->thash); free(d->tindex); free(d->tshash); free(d->tcs); free(d->tce);
    inter_free(&d->bb); inter_free(&d->bt);
    for (int k = 0; k < d->nhook; k++) free(d->hook[k].f);
    free(d);
}

static void add_row(row_t** rows, int* n, int a, int b, const double* p, int np)
{
    *rows = (row_t*)realloc(*rows, sizeof(row_t) * (*n + 1));
    (*rows)[*n].a = a; (*rows)[*n].b = b;
    memset((*rows)[*n].p, 0, sizeof((*rows)[*n].p));
    memcpy((*rows)[*n].p, p, sizeof(double) * np);
    (*n)++;
}
void o_dem_set_linear(o_dem* d, int a, int b, const double* p) { add_row(&d->lin, &d->nlin, a, b, p, 11); }
void o_dem_set_hertz(o_dem* d, int a, int b, const double* p) { add_row(&d->hz, &d->nhz, a, b, p, 8); }
void o_dem_set_bonded(o_dem* d, int a, int b, const double* p) { add_row(&d->bd, &d->nbd, a, b, p, 6); }

/* contactPairParameterIndex (contactParameters.h:123-141) */
static int pidx(int a, int b, int nMat, int cap)
{
    if (nMat <= 0 || cap <= 0) return 0;
    if (a < 0 || b < 0 || a >= nMat || b >= nMat) return cap - 1;
    int i = a, j = b;
    if (i > j) { int t = i; i = j; j = t; }
    long long idx = ((long long)i * (2LL * nMat - i + 1LL)) / 2LL + (long long)(j - i);
    if (idx < 0) idx = 0;
    if (idx >= cap) idx = cap - 1;
    return (int)idx;
}

/* contactModelParameters::buildFromTables (contactKernel.cu:6-104) */
static void build_tables(o_dem* d)
{
    int mx = -1;
    for (int k = 0; k < d->nhz; k++) { if (d->hz[k].a > mx) mx = d->hz[k].a; if (d->hz[k].b > mx) mx = d->hz[k].b; }
    for (int k = 0; k < d->nlin; k++) { if (d->lin[k].a > mx) mx = d->lin[k].a; if (d->lin[k].b > mx) mx = d->lin[k].b; }
    for (int k = 0; k < d->nbd; k++) { if (d->bd[k].a > mx) mx = d->bd[k].a; if (d->bd[k].b > mx) mx = d->bd[k].b; }
    free(d->H); free(d->L); free(d->B);
    d->H = d->L = d->B = NULL;
    d->nMat = mx < 0 ? 0 : mx + 1;
    d->cap = d->nMat > 0 ? d->nMat * (d->nMat + 1) / 2 + 1 : 0;
    if (d->cap == 0) return;
    int c = d->cap;
    d->H = (double*)calloc((size_t)H_N * c, sizeof(double));
    d->L = (double*)calloc((size_t)L_N * c, sizeof(double));
    d->B = (double*)calloc((size_t)B_N * c, sizeof(double));
    for (int i = 0; i < c; i++) { d->H[H_RES * c + i] = 1.0; d->B[B_GAMMA * c + i] = 1.0; d->B[B_KNKS * c + i] = 1.0; }
    for (int k = 0; k < d->nhz; k++) {
        int i = pidx(d->hz[k].a, d->hz[k].b, d->nMat, c);
        for (int q = 0; q < H_N; q++) d->H[q * c + i] = d->hz[k].p[q];
    }
    for (int k = 0; k < d->nlin; k++) {
        int i = pidx(d->lin[k].a, d->lin[k].b, d->nMat, c);
        for (int q = 0; q < L_N; q++) d->L[q * c + i] = d->lin[k].p[q];
    }
    for (int k = 0; k < d->nbd; k++) {
        int i = pidx(d->bd[k].a, d->bd[k].b, d->nMat, c);
        for (int q = 0; q < B_N; q++) d->B[q * c + i] = d->bd[k].p[q];
    }
}
static double Hp(const o_dem* d, int ip, int q) { return d->cap ? d->H[q * d->cap + ip] : (q == H_RES ? 1.0 : 0.0); }
static double Lp(const o_dem* d, int ip, int q) { return d->cap ? d->L[q * d->cap + ip] : 0.0; }
static double Bp(const o_dem* d, int ip, int q)
{
    return d->cap ? d->B[q * d->cap + ip] : ((q == B_GAMMA || q == B_KNKS) ? 1.0 : 0.0);
}

/* ------------------------------------------------------------- scene API */
void o_dem_add_ball(o_dem* d, const double* pos, const double* vel, const double* w, double r, double density,
                    int mat, int clump)
{
    double mass = 4. / 3. * r * r * r * O_PI * density;
    double inv = mass > 0.0 ? 1.0 / mass : 0.0;
    dpush3(&d->pos, v3(pos[0], pos[1], pos[2]));
    dpush3(&d->vel, v3(vel[0], vel[1], vel[2]));
    dpush3(&d->w, v3(w[0], w[1], w[2]));
    dpush3(&d->force, v3(0, 0, 0));
    dpush3(&d->torque, v3(0, 0, 0));
    dpush(&d->radius, r);
    dpush(&d->invMass, inv);
    ipush(&d->mat, mat);
    ipush(&d->clump, clump);
}

void o_dem_add_fixed_ball(o_dem* d, const double* pos, double r, int mat, int clump)
{
    double z[3] = {0, 0, 0};
    dpush3(&d->pos, v3(pos[0], pos[1], pos[2]));
    dpush3(&d->vel, v3(z[0], z[1], z[2]));
    dpush3(&d->w, v3(0, 0, 0));
    dpush3(&d->force, v3(0, 0, 0));
    dpush3(&d->torque, v3(0, 0, 0));
    dpush(&d->radius, r);
    dpush(&d->invMass, 0.0);
    ipush(&d->mat, mat);
    ipush(&d->clump, clump);
}

static void push_body(dvec* pos, dvec* vel, dvec* w, dvec* q, o_vec3 p, o_vec3 v, o_vec3 om)
{
    dpush3(pos, p); dpush3(vel, v); dpush3(w, om);
    dpush(q, 1.0); dpush(q, 0.0); dpush(q, 0.0); dpush(q, 0.0);
}

void o_dem_add_clump(o_dem* d, int np, const double* pts, const double* radii, const double* centroid,
                     const double* vel, const double* w, double mass, const double* inertia, int mat)
{
    int cid = (int)d->cps.n;
    int ps = (int)d->radius.n;
    o_vec3 c = v3(centroid[0], centroid[1], centroid[2]);
    o_vec3 om = v3(w[0], w[1], w[2]);
    double volume = 0.0, invMass = 0.0;
    o_vec3* pv = (o_vec3*)malloc(sizeof(o_vec3) * (np > 0 ? np : 1));
    for (int i = 0; i < np; i++) pv[i] = v3(vel[0], vel[1], vel[2]);
    if (mass > 1.e-20) {
        invMass = 1.0 / mass;
        for (int i = 0; i < np; i++) {
            volume += 4.0 / 3.0 * O_PI * pow(radii[i], 3.0);
            pv[i] = vadd(pv[i], vcross(om, vsub(v3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), c)));
        }
    }
    double dens = volume > 0. ? mass / volume : 0.0;
    for (int i = 0; i < np; i++) {
        double vv[3] = {pv[i].x, pv[i].y, pv[i].z};
        o_dem_add_ball(d, &pts[3 * i], vv, w, radii[i], dens, mat, cid);
    }
    free(pv);
    o_sym I = {inertia[0], inertia[1], inertia[2], inertia[3], inertia[4], inertia[5]};
    o_sym iI = sym_inverse(I);
    dpush3(&d->cpos, c); dpush3(&d->cvel, v3(vel[0], vel[1], vel[2])); dpush3(&d->cw, om);
    dpush3(&d->cforce, v3(0, 0, 0)); dpush3(&d->ctorque, v3(0, 0, 0));
    dpush(&d->cq, 1.0); dpush(&d->cq, 0.0); dpush(&d->cq, 0.0); dpush(&d->cq, 0.0);
    dpush(&d->cinvI, iI.xx); dpush(&d->cinvI, iI.yy); dpush(&d->cinvI, iI.zz);
    dpush(&d->cinvI, iI.xy); dpush(&d->cinvI, iI.xz); dpush(&d->cinvI, iI.yz);
    dpush(&d->cinvMass, invMass);
    ipush(&d->cmat, mat); ipush(&d->cps, ps); ipush(&d->cpe, ps + np);
}

void o_dem_add_fixed_clump(o_dem* d, int np, const double* pts, const double* radii, const double* centroid, int mat)
{
    int cid = (int)d->cps.n;
    int ps = (int)d->radius.n;
    for (int i = 0; i < np; i++) o_dem_add_fixed_ball(d, &pts[3 * i], radii[i], mat, cid);
    dpush3(&d->cpos, v3(centroid[0], centroid[1], centroid[2]));
    dpush3(&d->cvel, v3(0, 0, 0)); dpush3(&d->cw, v3(0, 0, 0));
    dpush3(&d->cforce, v3(0, 0, 0)); dpush3(&d->ctorque, v3(0, 0, 0));
    dpush(&d->cq, 1.0); dpush(&d->cq, 0.0); dpush(&d->cq, 0.0); dpush(&d->cq, 0.0);
    for (int k = 0; k < 6; k++) dpush(&d->cinvI, 0.0);
    dpush(&d->cinvMass, 0.0);
    ipush(&d->cmat, mat); ipush(&d->cps, ps); ipush(&d->cpe, ps + np);
}

/* triangleCircumcenter (wallIntegrationKernel.cuh:4-29) */
static o_vec3 circumcenter(o_vec3 a, o_vec3 b, o_vec3 c)
{
    o_vec3 ab = vsub(b, a), ac = vsub(c, a);
    o_vec3 n = vcross(ab, ac);
    double n2 = vdot(n, n);
    if (n2 < 1e-30) return vdiv(vadd(vadd(a, b), c), 3.0);
    o_vec3 t1 = vmul(vcross(n, ab), vdot(ac, ac));
    o_vec3 t2 = vmul(vcross(ac, n), vdot(ab, ab));
    double inv = 1.0 / (2.0 * n2);
    return vadd(a, vmul(vadd(t1, t2), inv));
}

int o_dem_add_mesh_wall(o_dem* d, int nv, const double* verts, int nt, const int* t0, const int* t1, const int* t2,
                        const double* pos, const double* vel, const double* w, int mat)
{
    (void)vel; (void)w; /* D4: the reference stores zero velocities */
    if (nt <= 0) return -1;
    int mx = -1;
    for (int i = 0; i < nt; i++) {
        if (t0[i] > mx) mx = t0[i];
        if (t1[i] > mx) mx = t1[i];
        if (t2[i] > mx) mx = t2[i];
    }
    if (mx >= nv) return -1;
    int vs = (int)d->vwall.n;
    int wi = (int)d->wmat.n;
    o_vec3 p = v3(pos[0], pos[1], pos[2]);
    push_body(&d->wpos, &d->wvel, &d->ww, &d->wq, p, v3(0, 0, 0), v3(0, 0, 0));
    ipush(&d->wmat, mat);
    for (int i = 0; i < nv; i++) {
        o_vec3 v = v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
        dpush3(&d->vlocal, v);
        dpush3(&d->vglobal, vadd(v, p));
        ipush(&d->vwall, wi);
    }
    for (int i = 0; i < nt; i++) {
        o_vec3 a = vadd(v3(verts[3 * t0[i]], verts[3 * t0[i] + 1], verts[3 * t0[i] + 2]), p);
        o_vec3 b = vadd(v3(verts[3 * t1[i]], verts[3 * t1[i] + 1], verts[3 * t1[i] + 2]), p);
        o_vec3 c = vadd(v3(verts[3 * t2[i]], verts[3 * t2[i] + 1], verts[3 * t2[i] + 2]), p);
        ipush(&d->t0, t0[i] + vs); ipush(&d->t1, t1[i] + vs); ipush(&d->t2, t2[i] + vs);
        ipush(&d->twall, wi);
        dpush3(&d->tcc, circumcenter(a, b, c));
    }
    return 0;
}

/* bondedInteraction::add (DEMSolver.h:93-134): dedup of (min,max) pairs,
 * bond point = midpoint, normal = normalize(p[a]-p[b]), isBonded = 1. */
static int nearest_image(const o_dem* d, o_vec3 pi, o_vec3 pj);
static o_vec3 shift_point(const o_dem* d, o_vec3 p, int code);
static int any_periodic(const o_dem* d);

void o_dem_add_bonds(o_dem* d, int n, const int* ob0, const int* ob1, int npos, const double* pos)
{
    for (int k = 0; k < n; k++) {
        int a = ob0[k], b = ob1[k];
        if (a < 0 || b < 0) continue;
        if (a >= npos || b >= npos) continue;
        if (a == b) continue;
        if (a > b) { int t = a; a = b; b = t; }
        int found = 0;
        for (long j = 0; j < d->bpointed.n; j++)
            if (d->bpointed.p[j] == a && d->bpointing.p[j] == b) { found = 1; break; }
        if (found) continue;
        ipush(&d->bpointed, a); ipush(&d->bpointing, b); ipush(&d->bonded, 1);
        o_vec3 pa = vld(pos, a), pb = vld(pos, b);
        /* a bond across a periodic face joins a to the nearest image of b */
        if (any_periodic(d) && d->bgValid) pb = shift_point(d, pb, nearest_image(d, pa, pb));
        dpush3(&d->bpoint, vmul(vadd(pa, pb), 0.5));
        dpush3(&d->bnormal, vnormalize(vsub(pa, pb)));
        dpush3(&d->bshear, v3(0, 0, 0)); dpush3(&d->bbend, v3(0, 0, 0));
        dpush(&d->bFn, 0.0); dpush(&d->bTt, 0.0); dpush(&d->bsig, 0.0); dpush(&d->btau, 0.0);
    }
}

/* ------------------------------------------------------------- interactions */
static void inter_alloc(inter_t* I, long cap, int npointed, int npointing)
{
    I->cap = cap > 0 ? cap : 1;
    long c = I->cap;
    I->pointed = (int*)realloc(I->pointed, sizeof(int) * c);
    I->pointing = (int*)realloc(I->pointing, sizeof(int) * c);
    I->hashIndex = (int*)realloc(I->hashIndex, sizeof(int) * c);
    I->cancel = (int*)realloc(I->cancel, sizeof(int) * c);
    I->shift = (int*)realloc(I->shift, sizeof(int) * c);
    I->sl = (double*)realloc(I->sl, sizeof(double) * 3 * c);
    I->ro = (double*)realloc(I->ro, sizeof(double) * 3 * c);
    I->to = (double*)realloc(I->to, sizeof(double) * 3 * c);
    I->F = (double*)realloc(I->F, sizeof(double) * 3 * c);
    I->T = (double*)realloc(I->T, sizeof(double) * 3 * c);
    I->P = (double*)realloc(I->P, sizeof(double) * 3 * c);
    I->N = (double*)realloc(I->N, sizeof(double) * 3 * c);
    I->ov = (double*)realloc(I->ov, sizeof(double) * c);
    if (!I->count || I->npointed != npointed) {
        free(I->count); free(I->prefix);
        I->count = (int*)calloc(npointed > 0 ? npointed : 1, sizeof(int));
        I->prefix = (int*)calloc(npointed > 0 ? npointed : 1, sizeof(int));
        I->npointed = npointed;
    }
    if (!I->istart || I->npointing != npointing) {
        free(I->istart); free(I->iend);
        I->istart = (int*)malloc(sizeof(int) * (npointing > 0 ? npointing : 1));
        I->iend = (int*)malloc(sizeof(int) * (npointing > 0 ? npointing : 1));
        for (int k = 0; k < npointing; k++) I->istart[k] = I->iend[k] = -1;
        I->npointing = npointing;
    }
}

/* updateOldPairOldSpring (DEMSolver.h:30-46) */
static void inter_save_old(inter_t* I)
{
    long n = I->n;
    I->nold = n;
    if (n == 0) return;
    I->opointed = (int*)realloc(I->opointed, sizeof(int) * n);
    I->ohashIndex = (int*)realloc(I->ohashIndex, sizeof(int) * n);
    I->ocancel = (int*)realloc(I->ocancel, sizeof(int) * n);
    I->osl = (double*)realloc(I->osl, sizeof(double) * 3 * n);
    I->oro = (double*)realloc(I->oro, sizeof(double) * 3 * n);
    I->oto = (double*)realloc(I->oto, sizeof(double) * 3 * n);
    memcpy(I->opointed, I->pointed, sizeof(int) * n);
    memcpy(I->ohashIndex, I->hashIndex, sizeof(int) * n);
    memcpy(I->ocancel, I->cancel, sizeof(int) * n);
    memcpy(I->osl, I->sl, sizeof(double) * 3 * n);
    memcpy(I->oro, I->ro, sizeof(double) * 3 * n);
    memcpy(I->oto, I->to, sizeof(double) * 3 * n);
}

/* Spring history carry of the write kernels (ballNeighborSearchKernel.cu:127-144). */
static void carry(inter_t* I, int A, int B, long w)
{
    I->sl[3 * w] = I->sl[3 * w + 1] = I->sl[3 * w + 2] = 0.0;
    I->ro[3 * w] = I->ro[3 * w + 1] = I->ro[3 * w + 2] = 0.0;
    I->to[3 * w] = I->to[3 * w + 1] = I->to[3 * w + 2] = 0.0;
    if (I->istart[B] == -1) return;
    for (int j = I->istart[B]; j < I->iend[B]; j++) {
        int j1 = I->ohashIndex[j];
        if (I->opointed[j1] == A) {
            memcpy(&I->sl[3 * w], &I->osl[3 * (long)j1], sizeof(double) * 3);
            memcpy(&I->ro[3 * w], &I->oro[3 * (long)j1], sizeof(double) * 3);
            memcpy(&I->to[3 * w], &I->oto[3 * (long)j1], sizeof(double) * 3);
            break;
        }
    }
}

/* buildObjectPointingInteractionStartEnd (DEMSolver.h:59-78) with the
 * pointing ranges reset before the build (D2: intended semantics). */
static void inter_build_pointing(inter_t* I)
{
    for (int k = 0; k < I->npointing; k++) I->istart[k] = I->iend[k] = -1;
    long n = I->n;
    if (n == 0) return;
    int* key = (int*)malloc(sizeof(int) * n);
    int* sk = (int*)malloc(sizeof(int) * n);
    memcpy(key, I->pointing, sizeof(int) * n);
    o_sort_by_hash(key, (int)n, I->npointing, sk, I->hashIndex);
    o_cell_ranges(sk, (int)n, I->npointing, I->istart, I->iend);
    free(key); free(sk);
}

/* ------------------------------------------------------------- periodic boundaries
 * Intended semantics of setPeriodicBoundary (DEMSolver.h:1944-1959,
 * :279-426, neighborSearchKernel.cu:29-77): on a periodic axis the 27-cell
 * stencil wraps around the domain, a ball in the first cell layer meets the
 * balls of the last layer through their periodic images, and every contact
 * is listed once (half list by insertion index, as the ball-ball list).
 * The reference's seven ghost passes list some contacts twice and miss
 * others (see reference_periodic_pairs in oracle/__init__.py); this oracle
 * follows the intended rule.  A pair stores the image of its pointing ball
 * as a code (sx+1) + 3 (sy+1) + 9 (sz+1), s = -1, 0, +1 periods per axis;
 * 13 = the ball itself.  Balls that are not clump pebbles are wrapped back
 * into [min, max) on periodic axes at every neighbour search. */
#define O_NOSHIFT 13
static double shifted(double x, int s, double L) { return s == 0 ? x : (s > 0 ? x + L : x - L); }
static o_vec3 shift_point(const o_dem* d, o_vec3 p, int code)
{
    if (code == O_NOSHIFT) return p;
    return v3(shifted(p.x, code % 3 - 1, d->Lp[0]), shifted(p.y, code / 3 % 3 - 1, d->Lp[1]),
              shifted(p.z, code / 9 - 1, d->Lp[2]));
}
static o_vec3 image_pos(const o_dem* d, int j, int code) { return shift_point(d, vld(d->pos.p, j), code); }
/* Image of ball j nearest to ball i on the periodic axes (bonds). */
static int nearest_image(const o_dem* d, o_vec3 pi, o_vec3 pj)
{
    double dd[3] = {pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
    int s[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        if (!d->per[a]) continue;
        if (dd[a] > 0.5 * d->Lp[a]) s[a] = 1;
        else if (dd[a] < -0.5 * d->Lp[a]) s[a] = -1;
    }
    return (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
}
static int any_periodic(const o_dem* d) { return d->per[0] || d->per[1] || d->per[2]; }

void o_dem_set_periodic(o_dem* d, int x, int y, int z)
{
    d->per[0] = x != 0; d->per[1] = y != 0; d->per[2] = z != 0;
}

/* A position just below the maximum face can round into the cell past the
 * last one, (x - min) / cellSize >= n; such a ball takes its image just
 * below the minimum face instead (which hashes into cell 0), so every
 * wrapped ball hashes inside the grid. */
static void wrap_positions(o_dem* d)
{
    const double lo[3] = {d->bg.minx, d->bg.miny, d->bg.minz}, hi[3] = {d->bg.maxx, d->bg.maxy, d->bg.maxz};
    const double cs[3] = {d->bg.cx, d->bg.cy, d->bg.cz};
    const int nc[3] = {d->bg.nx, d->bg.ny, d->bg.nz};
    for (long i = 0; i < d->radius.n; i++) {
        if (d->clump.p[i] >= 0) continue;
        for (int a = 0; a < 3; a++) {
            if (!d->per[a]) continue;
            double* x = &d->pos.p[3 * i + a];
            if (*x >= hi[a]) *x -= d->Lp[a];
            else if (*x < lo[a]) *x += d->Lp[a];
            if ((int)((*x - lo[a]) / cs[a]) >= nc[a]) *x -= d->Lp[a];
        }
    }
}

/* ------------------------------------------------------------- init */
static void grid_arrays(int** h, int** idx, int** sh, int** cs, int** ce, long n, long ng)
{
    free(*h); free(*idx); free(*sh); free(*cs); free(*ce);
    *h = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    *idx = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    *sh = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    *cs = (int*)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    *ce = (int*)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    for (long k = 0; k < ng; k++) (*cs)[k] = (*ce)[k] = -1;
}

int o_dem_initialize(o_dem* d, const double* minb, const double* maxb)
{
    build_tables(d);
    long nb = d->radius.n, nt = d->t0.n;
    double c = 0.0;
    for (long i = 0; i < nb; i++) if (d->radius.p[i] * 2.0 > c) c = d->radius.p[i] * 2.0;
    /* ballSpatialGrid_.set ignores a zero cell size (boundary.h:75-78) */
    if (!d->bgValid || nb > 0) d->bgValid = (o_grid_set(&d->bg, minb, maxb, c) == 0) || d->bgValid;
    for (long i = 0; i < nt; i++) {
        o_vec3 a = vld(d->vlocal.p, d->t0.p[i]), b = vld(d->vlocal.p, d->t1.p[i]), cc = vld(d->vlocal.p, d->t2.p[i]);
        o_vec3 m = circumcenter(a, b, cc);
        double r0 = 2.0 * vlen(vsub(a, m)), r1 = 2.0 * vlen(vsub(b, m)), r2 = 2.0 * vlen(vsub(cc, m));
        if (r0 > c) c = r0;
        if (r1 > c) c = r1;
        if (r2 > c) c = r2;
    }
    d->tgValid = (o_grid_set(&d->tg, minb, maxb, c) == 0) || d->tgValid;
    d->bng = d->bgValid ? (long)d->bg.nx * d->bg.ny * d->bg.nz : 0;
    if (d->bgValid) {
        d->Lp[0] = d->bg.maxx - d->bg.minx; d->Lp[1] = d->bg.maxy - d->bg.miny; d->Lp[2] = d->bg.maxz - d->bg.minz;
        const int ncell[3] = {d->bg.nx, d->bg.ny, d->bg.nz};
        for (int a = 0; a < 3; a++)
            if (d->per[a] && ncell[a] < 3) return -2; /* a wrapped stencil needs three distinct cells */
    }
    d->tng = d->tgValid ? (long)d->tg.nx * d->tg.ny * d->tg.nz : 0;
    grid_arrays(&d->bhash, &d->bindex, &d->bshash, &d->bcs, &d->bce, nb, d->bng);
    grid_arrays(&d->thash, &d->tindex, &d->tshash, &d->tcs, &d->tce, nt, d->tng);
    if (nb > 0 && d->bb.cap == 0) inter_alloc(&d->bb, 6 * nb, (int)nb, (int)nb);
    if (nt > 0 && d->bt.npointing == 0) inter_alloc(&d->bt, 1, (int)nb, (int)nt);
    memset(d->force.p, 0, sizeof(double) * 3 * nb);
    memset(d->torque.p, 0, sizeof(double) * 3 * nb);
    return d->bgValid ? 0 : -1;
}

/* ------------------------------------------------------------- neighbours */
static int ball_walk(o_dem* d, int A, int write, long base)
{
    const o_grid* g = &d->bg;
    int c[3], lo[3], hi[3];
    o_stencil(g, &d->pos.p[3 * (long)A], c, lo, hi);
    const int ncell[3] = {g->nx, g->ny, g->nz};
    for (int a = 0; a < 3; a++)
        if (d->per[a]) { lo[a] = -1; hi[a] = 1; }
    o_vec3 pa = vld(d->pos.p, A);
    double ra = d->radius.p[A];
    int count = 0;
    for (int zz = lo[2]; zz <= hi[2]; zz++)
        for (int yy = lo[1]; yy <= hi[1]; yy++)
            for (int xx = lo[0]; xx <= hi[0]; xx++) {
                int cc[3] = {c[0] + xx, c[1] + yy, c[2] + zz}, sh[3] = {0, 0, 0};
                for (int a = 0; a < 3; a++) {
                    if (cc[a] < 0) { cc[a] += ncell[a]; sh[a] = -1; }
                    else if (cc[a] >= ncell[a]) { cc[a] -= ncell[a]; sh[a] = 1; }
                }
                const int code = (sh[0] + 1) + 3 * (sh[1] + 1) + 9 * (sh[2] + 1);
                int hb = o_hash_cell(g, cc[0], cc[1], cc[2]);
                int s = d->bcs[hb];
                if (s == -1) continue;
                int e = d->bce[hb];
                for (int i = s; i < e; i++) {
                    int B = d->bindex[i];
                    if (A >= B) continue;
                    if (d->invMass.p[A] == 0 && d->invMass.p[B] == 0) continue;
                    if (d->clump.p[B] >= 0 && d->clump.p[A] == d->clump.p[B]) continue;
                    double cut = 1.1 * (ra + d->radius.p[B]);
                    o_vec3 r = vsub(pa, image_pos(d, B, code));
                    if ((cut * cut - vdot(r, r)) >= 0.) {
                        if (write) {
                            long w = base + count;
                            d->bb.pointed[w] = A;
                            d->bb.pointing[w] = B;
                            d->bb.shift[w] = code;
                            carry(&d->bb, A, B, w);
                        }
                        count++;
                    }
                }
            }
    return count;
}

/* DEMSolver::neighborSearch (DEMSolver.h:698-777) */
void o_dem_neighbor_search(o_dem* d)
{
    long nb = d->radius.n;
    if (nb == 0) return;
    if (any_periodic(d)) wrap_positions(d);
    o_hash(&d->bg, d->pos.p, (int)nb, d->bhash);
    o_sort_by_hash(d->bhash, (int)nb, (int)d->bng, d->bshash, d->bindex);
    o_cell_ranges(d->bshash, (int)nb, (int)d->bng, d->bcs, d->bce);
    inter_save_old(&d->bb);
    long tot = 0;
    for (int A = 0; A < nb; A++) { d->bb.count[A] = ball_walk(d, A, 0, 0); tot += d->bb.count[A]; d->bb.prefix[A] = (int)tot; }
    d->bb.n = tot;
    if (tot > d->bb.cap) inter_alloc(&d->bb, tot, (int)nb, (int)nb);
    for (int A = 0; A < nb; A++) ball_walk(d, A, 1, A > 0 ? d->bb.prefix[A - 1] : 0);
    inter_build_pointing(&d->bb);
}

/* ------------------------------------------------------------- contact models */
/* integrateSlidingOrRollingSpring (contactKernel.cuh:64-97) */
static o_vec3 spring_slide(o_vec3 prev, o_vec3 sv, o_vec3 n, o_vec3 Fn, double mu, double k, double c, double dt)
{
    o_vec3 s = v3(0, 0, 0);
    if (k > 0.) {
        o_vec3 p1 = vsub(prev, vmul(n, vdot(prev, n)));
        double ap1 = vlen(p1);
        if (ap1 > 1.e-10) p1 = vmul(p1, vlen(prev) / ap1);
        s = vadd(p1, vmul(sv, dt));
        if (mu > 0.) {
            o_vec3 f = vsub(vmul(s, -k), vmul(sv, c));
            double af = vlen(f), an = vlen(Fn);
            if (af > mu * an) {
                double ratio = mu * an / af;
                f = vmul(f, ratio);
                s = vdiv(vneg(vadd(f, vmul(sv, c))), k);
            }
        }
    }
    return s;
}

/* integrateTorsionSpring (contactKernel.cuh:99-126) */
static o_vec3 spring_torsion(o_vec3 prev, o_vec3 tv, o_vec3 n, o_vec3 Fn, double mu, double k, double c, double dt)
{
    o_vec3 s = v3(0, 0, 0);
    if (k > 0.) {
        s = vmul(n, vdot(vadd(prev, vmul(tv, dt)), n));
        if (mu > 0.) {
            o_vec3 f = vsub(vmul(s, -k), vmul(tv, c));
            double af = vlen(f), an = vlen(Fn);
            if (af > mu * an) {
                double ratio = mu * an / af;
                f = vmul(f, ratio);
                s = vdiv(vneg(vadd(f, vmul(tv, c))), k);
            }
        }
    }
    return s;
}

/* LinearContact / HertzianMindlinContact (contactKernel.cuh:128-290) and
 * the model selection of contactKernel.cu:218-252 (linear if k_n > 1e-20). */
static void contact_model(const o_dem* d, int ip, o_vec3 vc, o_vec3 wij, o_vec3 n, double delta, double m,
                          double R, double dt, o_vec3* F, o_vec3* T, o_vec3* es, o_vec3* er, o_vec3* et)
{
    double kn, ks, kr, kt, cn, cs, cr, ct, mus, mur, mut;
    int linear = Lp(d, ip, L_KN) > 1.e-20;
    if (!(delta > 0)) {
        *es = *er = *et = v3(0, 0, 0);
        return;
    }
    if (linear) {
        kn = Lp(d, ip, L_KN); ks = Lp(d, ip, L_KS); kr = Lp(d, ip, L_KR); kt = Lp(d, ip, L_KT);
        cn = 2. * Lp(d, ip, L_DN) * sqrt(m * kn);
        cs = 2. * Lp(d, ip, L_DS) * sqrt(m * ks);
        cr = 2. * Lp(d, ip, L_DR) * sqrt(m * kr);
        ct = 2. * Lp(d, ip, L_DT) * sqrt(m * kt);
        mus = Lp(d, ip, L_MUS); mur = Lp(d, ip, L_MUR); mut = Lp(d, ip, L_MUT);
    } else {
        double logR = log(Hp(d, ip, H_RES));
        double D = -logR / sqrt(logR * logR + O_PI * O_PI);
        double E = Hp(d, ip, H_E), G = Hp(d, ip, H_G);
        kn = 4. / 3. * E * sqrt(R * delta);
        ks = 8. * G * sqrt(R * delta);
        cn = 2. * sqrt(5. / 6.) * D * sqrt(m * kn);
        cs = 2. * sqrt(5. / 6.) * D * sqrt(m * ks);
        kr = ks * Hp(d, ip, H_KRKS);
        kt = ks * Hp(d, ip, H_KTKS);
        cr = 2. * sqrt(5. / 6.) * D * sqrt(m * kr);
        ct = 2. * sqrt(5. / 6.) * D * sqrt(m * kt);
        mus = Hp(d, ip, H_MUS); mur = Hp(d, ip, H_MUR); mut = Hp(d, ip, H_MUT);
    }
    o_vec3 vn = vmul(n, vdot(vc, n));
    o_vec3 Fn = vsub(vmul(n, kn * delta), vmul(vn, cn));
    o_vec3 vs = vsub(vc, vn);
    *es = spring_slide(*es, vs, n, Fn, mus, ks, cs, dt);
    o_vec3 Fs = vsub(vmul(*es, -ks), vmul(vs, cs));
    o_vec3 vr = vmul(vcross(n, wij), -R);
    *er = spring_slide(*er, vr, n, Fn, mur, kr, cr, dt);
    o_vec3 Fr = vsub(vmul(*er, -kr), vmul(vr, cr));
    o_vec3 Tr = vmul(vcross(n, Fr), R);
    double D2 = 2 * R;
    o_vec3 vt = vmul(n, D2 * vdot(wij, n));
    *et = spring_torsion(*et, vt, n, Fn, mut, kt, ct, dt);
    o_vec3 Ft = vsub(vmul(*et, -kt), vmul(vt, ct));
    o_vec3 Tt = vmul(Ft, D2);
    *F = vadd(Fn, Fs);
    *T = vadd(Tr, Tt);
}

/* ParallelBondedContact (contactKernel.cuh:4-62) */
static int bond_model(double* Fn, double* Tt, o_vec3* Fs, o_vec3* Tb, double* sig, double* tau, o_vec3 n0, o_vec3 n,
                      o_vec3 vc, o_vec3 wa, o_vec3 wb, double ra, double rb, double dt, double gam, double E,
                      double knks, double sigt, double C, double mu)
{
    o_vec3 th1 = vcross(n0, n);
    *Fs = vrotate(*Fs, th1);
    *Tb = vrotate(*Tb, th1);
    o_vec3 th2 = vmul(n, vdot(vmul(vmul(vadd(wa, wb), 0.5), dt), n));
    *Fs = vrotate(*Fs, th2);
    *Tb = vrotate(*Tb, th2);
    double rmin = ra < rb ? ra : rb;
    double Rb = gam * rmin;
    double A = Rb * Rb * O_PI;
    double I = Rb * Rb * Rb * Rb / 4. * O_PI;
    double J = 2 * I;
    double kn = E / (ra + rb);
    double ks = kn / knks;
    o_vec3 dn = vmul(n, vdot(vc, n) * dt);
    o_vec3 dtg = vsub(vmul(vc, dt), dn);
    *Fn -= vdot(vmul(vmul(dn, kn), A), n);
    *Fs = vsub(*Fs, vmul(vmul(dtg, ks), A));
    o_vec3 wr = vsub(wa, wb);
    o_vec3 rn = vmul(n, vdot(wr, n) * dt);
    o_vec3 rt = vsub(vmul(wr, dt), rn);
    *Tt -= vdot(vmul(vmul(rn, ks), J), n);
    *Tb = vsub(*Tb, vmul(vmul(rt, kn), I));
    *sig = -*Fn / A + vlen(*Tb) / I * Rb;
    *tau = vlen(*Fs) / A + fabs(*Tt) / J * Rb;
    int bonded = 1;
    if (sigt > 0 && *sig > sigt) bonded = 0;
    else if (C > 0 && *tau > C - mu * *sig) bonded = 0;
    return bonded;
}

/* calculateBallContactForceTorque (DEMSolver.h:599-695) */
void o_dem_contact(o_dem* d, double dt)
{
    long nb = d->radius.n;
    memset(d->force.p, 0, sizeof(double) * 3 * nb);
    memset(d->torque.p, 0, sizeof(double) * 3 * nb);
    inter_t* I = &d->bb;
    /* updateBallContactKernel + calBallContactForceTorqueKernel (contactKernel.cu:135-259) */
#pragma omp parallel for schedule(static)
    for (long k = 0; k < I->n; k++) {
        int i = I->pointed[k], j = I->pointing[k];
        o_vec3 ri = vld(d->pos.p, i), rj = image_pos(d, j, I->shift[k]);
        double ai = d->radius.p[i], aj = d->radius.p[j];
        o_vec3 rij = vsub(ri, rj);
        o_vec3 n = vnormalize(rij);
        double delta = ai + aj - vlen(rij);
        o_vec3 rc = vadd(rj, vmul(n, aj - 0.5 * delta));
        vst(I->P, k, rc); vst(I->N, k, n); I->ov[k] = delta;
        double R = ai * aj / (ai + aj);
        double m = 1. / (d->invMass.p[i] + d->invMass.p[j]);
        o_vec3 vi = vld(d->vel.p, i), vj = vld(d->vel.p, j), wi = vld(d->w.p, i), wj = vld(d->w.p, j);
        o_vec3 vc = vsub(vadd(vi, vcross(wi, vsub(rc, ri))), vadd(vj, vcross(wj, vsub(rc, rj))));
        o_vec3 wij = vsub(wi, wj);
        o_vec3 F = v3(0, 0, 0), T = v3(0, 0, 0);
        o_vec3 es = vld(I->sl, k), er = vld(I->ro, k), et = vld(I->to, k);
        int ip = pidx(d->mat.p[i], d->mat.p[j], d->nMat, d->cap);
        contact_model(d, ip, vc, wij, n, delta, m, R, dt, &F, &T, &es, &er, &et);
        vst(I->F, k, F); vst(I->T, k, T);
        vst(I->sl, k, es); vst(I->ro, k, er); vst(I->to, k, et);
    }
    /* calBondedForceTorqueKernel (contactKernel.cu:535-655), bond order */
    for (long b = 0; b < d->bonded.n; b++) {
        if (d->bonded.p[b] == 0) {
            d->bFn.p[b] = 0; d->bTt.p[b] = 0;
            vst(d->bshear.p, b, v3(0, 0, 0)); vst(d->bbend.p, b, v3(0, 0, 0));
            continue;
        }
        int i = d->bpointed.p[b], j = d->bpointing.p[b];
        o_vec3 ri = vld(d->pos.p, i), rj = vld(d->pos.p, j);
        if (any_periodic(d)) rj = image_pos(d, j, nearest_image(d, ri, rj));
        double ai = d->radius.p[i], aj = d->radius.p[j];
        o_vec3 n0 = vld(d->bnormal.p, b);
        o_vec3 n = vnormalize(vsub(ri, rj));
        double delta = ai + aj - vlen(vsub(ri, rj));
        o_vec3 rc = vadd(rj, vmul(n, aj - 0.5 * delta));
        vst(d->bpoint.p, b, rc); vst(d->bnormal.p, b, n);
        int flag = 0;
        long kc = 0;
        int st = i > 0 ? I->prefix[i - 1] : 0, en = I->prefix[i];
        for (int k = st; k < en; k++)
            if (I->pointing[k] == j) {
                flag = 1; kc = k;
                rc = vld(I->P, k); n = vld(I->N, k);
                vst(d->bpoint.p, b, rc); vst(d->bnormal.p, b, n);
                break;
            }
        o_vec3 vi = vld(d->vel.p, i), vj = vld(d->vel.p, j), wi = vld(d->w.p, i), wj = vld(d->w.p, j);
        o_vec3 vc = vsub(vadd(vi, vcross(wi, vsub(rc, ri))), vadd(vj, vcross(wj, vsub(rc, rj))));
        int ip = pidx(d->mat.p[i], d->mat.p[j], d->nMat, d->cap);
        double Fn = d->bFn.p[b], Tt = d->bTt.p[b], sig = d->bsig.p[b], tau = d->btau.p[b];
        o_vec3 Fs = vld(d->bshear.p, b), Tb = vld(d->bbend.p, b);
        d->bonded.p[b] = bond_model(&Fn, &Tt, &Fs, &Tb, &sig, &tau, n0, n, vc, wi, wj, ai, aj, dt, Bp(d, ip, B_GAMMA),
                                    Bp(d, ip, B_E), Bp(d, ip, B_KNKS), Bp(d, ip, B_SIG), Bp(d, ip, B_C),
                                    Bp(d, ip, B_MU));
        d->bFn.p[b] = Fn; d->bTt.p[b] = Tt; d->bsig.p[b] = sig; d->btau.p[b] = tau;
        vst(d->bshear.p, b, Fs); vst(d->bbend.p, b, Tb);
        o_vec3 Fb = vadd(vmul(n, Fn), Fs);
        o_vec3 Tbn = vadd(vmul(n, Tt), Tb);
        if (!flag) {
            vst(d->force.p, i, vadd(vld(d->force.p, i), Fb));
            vst(d->torque.p, i, vadd(vld(d->torque.p, i), vadd(Tbn, vcross(vsub(rc, ri), Fs))));
            vst(d->force.p, j, vsub(vld(d->force.p, j), Fb));
            vst(d->torque.p, j, vadd(vld(d->torque.p, j), vadd(vneg(Tbn), vcross(vsub(rc, rj), vneg(Fs)))));
        } else {
            vst(I->F, kc, vadd(vld(I->F, kc), Fb));
            vst(I->T, kc, vadd(vld(I->T, kc), Tbn));
        }
    }
    /* sumObjectPointed / sumObjectPointing (contactKernel.cu:657-714) */
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nb; i++) {
        o_vec3 ri = vld(d->pos.p, i), F = v3(0, 0, 0), T = v3(0, 0, 0);
        int st = i > 0 ? I->prefix[i - 1] : 0, en = I->prefix[i];
        for (int k = st; k < en; k++) {
            o_vec3 Fc = vld(I->F, k);
            F = vadd(F, Fc);
            T = vadd(T, vadd(vld(I->T, k), vcross(vsub(vld(I->P, k), ri), Fc)));
        }
        vst(d->force.p, i, vadd(vld(d->force.p, i), F));
        vst(d->torque.p, i, vadd(vld(d->torque.p, i), T));
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < nb; i++) {
        o_vec3 ri = vld(d->pos.p, i), F = v3(0, 0, 0), T = v3(0, 0, 0);
        if (I->istart[i] != -1)
            for (int k = I->istart[i]; k < I->iend[i]; k++) {
                int k1 = I->hashIndex[k];
                o_vec3 Fc = vld(I->F, k1);
                /* lever arm about the image of i that the pair met */
                o_vec3 rim = I->shift[k1] == O_NOSHIFT ? ri : image_pos(d, (int)i, I->shift[k1]);
                F = vsub(F, Fc);
                T = vsub(T, vld(I->T, k1));
                T = vsub(T, vcross(vsub(vld(I->P, k1), rim), Fc));
            }
        vst(d->force.p, i, vadd(vld(d->force.p, i), F));
        vst(d->torque.p, i, vadd(vld(d->torque.p, i), T));
    }
}

/* ------------------------------------------------------------- triangles */
typedef enum { ST_NONE = 0, ST_FACE, ST_EDGE, ST_VERTEX } st_type;

static double clamp01(double x) { return x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x); }
static double dist2(o_vec3 a, o_vec3 b) { o_vec3 t = vsub(a, b); return vdot(t, t); }

/* closestPointOnSegment (contactKernel.cuh:315-330) */
static o_vec3 seg_closest(o_vec3 a, o_vec3 b, o_vec3 p, double* t)
{
    o_vec3 ab = vsub(b, a);
    double ab2 = vdot(ab, ab);
    if (ab2 <= 1e-30) { *t = 0.0; return a; }
    *t = clamp01(vdot(vsub(p, a), ab) / ab2);
    return vadd(a, vmul(ab, *t));
}

/* classifySphereTriangleContact (contactKernel.cuh:367-501): Ericson's
 * closest-point region tests; degenerate triangles as three segments. */
static st_type classify(o_vec3 c, double r, o_vec3 v0, o_vec3 v1, o_vec3 v2, o_vec3* cp)
{
    double r2 = r * r;
    double eps2 = 1e-12 * fmax(1.0, r2);
    o_vec3 ab = vsub(v1, v0), ac = vsub(v2, v0);
    o_vec3 n = vcross(ab, ac);
    double area2 = vdot(n, n);
    if (area2 <= 1e-20) {
        double t01, t02, t12;
        o_vec3 q01 = seg_closest(v0, v1, c, &t01), q02 = seg_closest(v0, v2, c, &t02), q12 = seg_closest(v1, v2, c, &t12);
        double dmin = dist2(c, q01);
        *cp = q01;
        st_type type = (t01 > 0.0 && t01 < 1.0) ? ST_EDGE : ST_VERTEX;
        double d02 = dist2(c, q02);
        if (d02 < dmin) { dmin = d02; *cp = q02; type = (t02 > 0.0 && t02 < 1.0) ? ST_EDGE : ST_VERTEX; }
        double d12 = dist2(c, q12);
        if (d12 < dmin) { dmin = d12; *cp = q12; type = (t12 > 0.0 && t12 < 1.0) ? ST_EDGE : ST_VERTEX; }
        return (dmin <= r2 + eps2) ? type : ST_NONE;
    }
    o_vec3 ap = vsub(c, v0);
    double d1 = vdot(ab, ap), d2 = vdot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) { *cp = v0; return dist2(c, v0) <= r2 + eps2 ? ST_VERTEX : ST_NONE; }
    o_vec3 bp = vsub(c, v1);
    double d3 = vdot(ab, bp), d4 = vdot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) { *cp = v1; return dist2(c, v1) <= r2 + eps2 ? ST_VERTEX : ST_NONE; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double t = d1 / (d1 - d3);
        *cp = vadd(v0, vmul(ab, t));
        return dist2(c, *cp) <= r2 + eps2 ? ST_EDGE : ST_NONE;
    }
    o_vec3 cp2 = vsub(c, v2);
    double d5 = vdot(ab, cp2), d6 = vdot(ac, cp2);
    if (d6 >= 0.0 && d5 <= d6) { *cp = v2; return dist2(c, v2) <= r2 + eps2 ? ST_VERTEX : ST_NONE; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double t = d2 / (d2 - d6);
        *cp = vadd(v0, vmul(ac, t));
        return dist2(c, *cp) <= r2 + eps2 ? ST_EDGE : ST_NONE;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        *cp = vadd(v1, vmul(vsub(v2, v1), t));
        return dist2(c, *cp) <= r2 + eps2 ? ST_EDGE : ST_NONE;
    }
    double sum = va + vb + vc;
    if (fabs(sum) <= 1e-30) return ST_NONE;
    double inv = 1.0 / sum;
    double v = vb * inv, w = vc * inv;
    *cp = vadd(vadd(v0, vmul(ab, v)), vmul(ac, w));
    return dist2(c, *cp) <= r2 + eps2 ? ST_FACE : ST_NONE;
}

int o_sphere_triangle(const double* c, double r, const double* v0, const double* v1, const double* v2, double* cp)
{
    o_vec3 q = v3(0, 0, 0);
    st_type t = classify(v3(c[0], c[1], c[2]), r, v3(v0[0], v0[1], v0[2]), v3(v1[0], v1[1], v1[2]),
                         v3(v2[0], v2[1], v2[2]), &q);
    cp[0] = q.x; cp[1] = q.y; cp[2] = q.z;
    return (int)t;
}

static int tri_walk(o_dem* d, int A, int write, long base)
{
    const o_grid* g = &d->tg;
    int c[3], lo[3], hi[3];
    o_stencil(g, &d->pos.p[3 * (long)A], c, lo, hi);
    o_vec3 pa = vld(d->pos.p, A);
    double ra = d->radius.p[A];
    int count = 0;
    for (int zz = lo[2]; zz <= hi[2]; zz++)
        for (int yy = lo[1]; yy <= hi[1]; yy++)
            for (int xx = lo[0]; xx <= hi[0]; xx++) {
                int hb = o_hash_cell(g, c[0] + xx, c[1] + yy, c[2] + zz);
                int s = d->tcs[hb];
                if (s == -1) continue;
                int e = d->tce[hb];
                for (int i = s; i < e; i++) {
                    int B = d->tindex[i];
                    o_vec3 v0 = vld(d->vglobal.p, d->t0.p[B]), v1 = vld(d->vglobal.p, d->t1.p[B]),
                           v2 = vld(d->vglobal.p, d->t2.p[B]);
                    double minx = fmin(v0.x, fmin(v1.x, v2.x)), miny = fmin(v0.y, fmin(v1.y, v2.y)),
                           minz = fmin(v0.z, fmin(v1.z, v2.z));
                    double maxx = fmax(v0.x, fmax(v1.x, v2.x)), maxy = fmax(v0.y, fmax(v1.y, v2.y)),
                           maxz = fmax(v0.z, fmax(v1.z, v2.z));
                    double cx = fmin(fmax(pa.x, minx), maxx), cy = fmin(fmax(pa.y, miny), maxy),
                           cz = fmin(fmax(pa.z, minz), maxz);
                    double dx = pa.x - cx, dy = pa.y - cy, dz = pa.z - cz;
                    if (dx * dx + dy * dy + dz * dz <= ra * ra) {
                        if (write) {
                            long w = base + count;
                            d->bt.pointed[w] = A;
                            d->bt.pointing[w] = B;
                            carry(&d->bt, A, B, w);
                        }
                        count++;
                    }
                }
            }
    return count;
}

static void copy_springs(inter_t* I, long to, long from)
{
    memcpy(&I->sl[3 * to], &I->sl[3 * from], sizeof(double) * 3);
    memcpy(&I->ro[3 * to], &I->ro[3 * from], sizeof(double) * 3);
    memcpy(&I->to[3 * to], &I->to[3 * from], sizeof(double) * 3);
}

/* updateBallTriangleContact (contactKernel.cu:261-418) */
static void update_ball_triangle(o_dem* d)
{
    inter_t* I = &d->bt;
    long nb = d->radius.n;
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < nb; i++) {
        int st = i > 0 ? I->prefix[i - 1] : 0, en = I->prefix[i];
        for (int c = st; c < en; c++) {
            I->cancel[c] = 0;
            int j = I->pointing[c];
            double ri_ = d->radius.p[i];
            o_vec3 ri = vld(d->pos.p, i);
            o_vec3 p0 = vld(d->vglobal.p, d->t0.p[j]), p1 = vld(d->vglobal.p, d->t1.p[j]), p2 = vld(d->vglobal.p, d->t2.p[j]);
            o_vec3 rc;
            st_type type = classify(ri, ri_, p0, p1, p2, &rc);
            vst(I->P, c, rc);
            vst(I->N, c, vnormalize(vsub(ri, rc)));
            I->ov[c] = ri_ - vlen(vsub(ri, rc));
            if (type == ST_NONE) {
                vst(I->sl, c, v3(0, 0, 0)); vst(I->ro, c, v3(0, 0, 0)); vst(I->to, c, v3(0, 0, 0));
                I->cancel[c] = 1;
                continue;
            }
            if (type == ST_FACE) continue;
            for (int c1 = st; c1 < en; c1++) {
                if (c1 == c) continue;
                int j1 = I->pointing[c1];
                o_vec3 q0 = vld(d->vglobal.p, d->t0.p[j1]), q1 = vld(d->vglobal.p, d->t1.p[j1]),
                       q2 = vld(d->vglobal.p, d->t2.p[j1]);
                o_vec3 rc1;
                st_type t1 = classify(ri, ri_, q0, q1, q2, &rc1);
                if (t1 == ST_NONE) continue;
                int hit = 0;
                if (t1 == ST_FACE) {
                    hit = vlen2(vcross(vsub(rc, q0), vsub(q1, q0))) < 1.e-20 ||
                          vlen2(vcross(vsub(rc, q1), vsub(q2, q1))) < 1.e-20 ||
                          vlen2(vcross(vsub(rc, q0), vsub(q2, q0))) < 1.e-20;
                } else if (t1 == ST_EDGE) {
                    if (type == t1) hit = (c1 < c) && vlen2(vsub(rc, rc1)) < 1.e-20;
                    else hit = vlen2(vsub(rc, rc1)) < 1.e-20;
                } else {
                    if (type == t1) hit = (c1 < c) && vlen2(vsub(rc, rc1)) < 1.e-20;
                }
                if (hit) {
                    copy_springs(I, c, c1);
                    I->cancel[c] = 1;
                    break;
                }
            }
        }
    }
}

/* calBallWallContactForceTorqueKernel (contactKernel.cu:420-533) */
static void ball_wall_force(o_dem* d, double dt)
{
    inter_t* I = &d->bt;
    long nb = d->radius.n;
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < nb; i++) {
        if (d->invMass.p[i] < 1.e-20) continue;
        o_vec3 ri = vld(d->pos.p, i), vi = vld(d->vel.p, i), wi = vld(d->w.p, i);
        double R = d->radius.p[i], m = 1. / d->invMass.p[i];
        int st = i > 0 ? I->prefix[i - 1] : 0, en = I->prefix[i];
        for (int c = st; c < en; c++) {
            vst(I->F, c, v3(0, 0, 0)); vst(I->T, c, v3(0, 0, 0));
            if (I->cancel[c] == 1) continue;
            int j = I->pointing[c];
            int wj = d->twall.p[j];
            o_vec3 rc = vld(I->P, c), n = vld(I->N, c);
            double delta = I->ov[c];
            o_vec3 rj = vld(d->wpos.p, wj), vj = vld(d->wvel.p, wj), wwj = vld(d->ww.p, wj);
            o_vec3 vc = vsub(vadd(vi, vcross(wi, vsub(rc, ri))), vadd(vj, vcross(wwj, vsub(rc, rj))));
            o_vec3 wij = vsub(wi, wwj);
            o_vec3 F = v3(0, 0, 0), T = v3(0, 0, 0);
            o_vec3 es = vld(I->sl, c), er = vld(I->ro, c), et = vld(I->to, c);
            int ip = pidx(d->mat.p[i], d->wmat.p[wj], d->nMat, d->cap);
            contact_model(d, ip, vc, wij, n, delta, m, R, dt, &F, &T, &es, &er, &et);
            vst(I->F, c, F); vst(I->T, c, T);
            vst(I->sl, c, es); vst(I->ro, c, er); vst(I->to, c, et);
            vst(d->force.p, i, vadd(vld(d->force.p, i), F));
            vst(d->torque.p, i, vadd(vld(d->torque.p, i), vadd(T, vcross(vsub(rc, ri), F))));
        }
    }
}

/* addTriangleContactForceTorque (DEMSolver.h:428-597) */
void o_dem_triangle_contact(o_dem* d, double dt)
{
    long nt = d->t0.n, nb = d->radius.n;
    if (nt == 0) return;
    o_hash(&d->tg, d->tcc.p, (int)nt, d->thash);
    o_sort_by_hash(d->thash, (int)nt, (int)d->tng, d->tshash, d->tindex);
    o_cell_ranges(d->tshash, (int)nt, (int)d->tng, d->tcs, d->tce);
    inter_t* I = &d->bt;
    inter_save_old(I);
    long tot = 0;
    for (int A = 0; A < nb; A++) { I->count[A] = tri_walk(d, A, 0, 0); tot += I->count[A]; I->prefix[A] = (int)tot; }
    I->n = tot;
    if (tot > I->cap) inter_alloc(I, tot, (int)nb, (int)nt);
    for (int A = 0; A < nb; A++) tri_walk(d, A, 1, A > 0 ? I->prefix[A - 1] : 0);
    inter_build_pointing(I);
    if (I->n > 0) {
        update_ball_triangle(d);
        ball_wall_force(d, dt);
    }
    /* launchWallIntegration (wallIntegrationKernel.cu:55-74) with full dt (D3) */
    for (long w = 0; w < d->wmat.n; w++) {
        vst(d->wpos.p, w, vadd(vld(d->wpos.p, w), vmul(vld(d->wvel.p, w), dt)));
        o_quat q = {d->wq.p[4 * w], d->wq.p[4 * w + 1], d->wq.p[4 * w + 2], d->wq.p[4 * w + 3]};
        q = qintegrate(q, vld(d->ww.p, w), dt);
        d->wq.p[4 * w] = q.q0; d->wq.p[4 * w + 1] = q.q1; d->wq.p[4 * w + 2] = q.q2; d->wq.p[4 * w + 3] = q.q3;
    }
    for (long v = 0; v < d->vwall.n; v++) {
        int w = d->vwall.p[v];
        o_quat q = {d->wq.p[4 * w], d->wq.p[4 * w + 1], d->wq.p[4 * w + 2], d->wq.p[4 * w + 3]};
        vst(d->vglobal.p, v, vadd(vld(d->wpos.p, w), qrotate(q, vld(d->vlocal.p, v))));
    }
    for (long t = 0; t < nt; t++)
        vst(d->tcc.p, t, circumcenter(vld(d->vglobal.p, d->t0.p[t]), vld(d->vglobal.p, d->t1.p[t]),
                                      vld(d->vglobal.p, d->t2.p[t])));
}

/* ------------------------------------------------------------- integration */
/* sumClumpForceTorqueKernel (ballIntegrationKernel.cu:3-28), intended form */
static void clump_sum(o_dem* d)
{
    for (long c = 0; c < d->cps.n; c++) {
        o_vec3 F = v3(0, 0, 0), T = v3(0, 0, 0), rc = vld(d->cpos.p, c);
        for (int i = d->cps.p[c]; i < d->cpe.p[c]; i++) {
            o_vec3 r = vld(d->pos.p, i);
            o_vec3 Fi = vld(d->force.p, i);
            if (d->policyD1) Fi = (i < d->cps.n) ? vld(d->cforce.p, i) : v3(0, 0, 0);
            F = vadd(F, Fi);
            T = vadd(T, vadd(vld(d->torque.p, i), vcross(vsub(r, rc), Fi)));
        }
        vst(d->cforce.p, c, F);
        vst(d->ctorque.p, c, T);
    }
}

/* clumpVelocityAngularVelocityIntegrationKernel (ballIntegrationKernel.cu:58-79) */
static void clump_kick(o_dem* d, o_vec3 g, double h)
{
    for (long c = 0; c < d->cps.n; c++) {
        double im = d->cinvMass.p[c];
        if (!(im > 1.e-20)) continue;
        vst(d->cvel.p, c, vadd(vld(d->cvel.p, c), vmul(vadd(vmul(vld(d->cforce.p, c), im), g), h)));
        o_quat q = {d->cq.p[4 * c], d->cq.p[4 * c + 1], d->cq.p[4 * c + 2], d->cq.p[4 * c + 3]};
        o_sym I = {d->cinvI.p[6 * c], d->cinvI.p[6 * c + 1], d->cinvI.p[6 * c + 2], d->cinvI.p[6 * c + 3],
                   d->cinvI.p[6 * c + 4], d->cinvI.p[6 * c + 5]};
        vst(d->cw.p, c, vadd(vld(d->cw.p, c), vmul(sym_mulv(sym_rotate(q, I), vld(d->ctorque.p, c)), h)));
    }
}

/* setPebbleVelocityAngularVelocityKernel (ballIntegrationKernel.cu:81-99) */
static void pebble_velocity(o_dem* d)
{
    for (long i = 0; i < d->radius.n; i++) {
        int c = d->clump.p[i];
        if (c < 0) continue;
        o_vec3 w = vld(d->cw.p, c);
        vst(d->w.p, i, w);
        vst(d->vel.p, i, vadd(vld(d->cvel.p, c), vcross(w, vsub(vld(d->pos.p, i), vld(d->cpos.p, c)))));
    }
}

/* ballVelocityAngularVelocityIntegrationKernel (ballIntegrationKernel.cu:30-56) */
static void ball_kick(o_dem* d, o_vec3 g, double h)
{
#pragma omp parallel for schedule(static)
    for (long i = 0; i < d->radius.n; i++) {
        if (d->clump.p[i] >= 0) continue;
        double im = d->invMass.p[i];
        if (im < 1.e-20) continue;
        vst(d->vel.p, i, vadd(vld(d->vel.p, i), vmul(vadd(vmul(vld(d->force.p, i), im), g), h)));
        double r = d->radius.p[i];
        if (r < 1.e-20) continue;
        double I = 0.4 * r * r / im;
        vst(d->w.p, i, vadd(vld(d->w.p, i), vmul(vdiv(vld(d->torque.p, i), I), h)));
    }
}

void o_dem_first_half(o_dem* d, const double* gg, double hdt)
{
    o_vec3 g = v3(gg[0], gg[1], gg[2]);
    if (d->cps.n > 0) {
        clump_sum(d);
        clump_kick(d, g, hdt);
        for (long c = 0; c < d->cps.n; c++) {
            vst(d->cpos.p, c, vadd(vld(d->cpos.p, c), vmul(vld(d->cvel.p, c), 2.0 * hdt)));
            o_quat q = {d->cq.p[4 * c], d->cq.p[4 * c + 1], d->cq.p[4 * c + 2], d->cq.p[4 * c + 3]};
            q = qintegrate(q, vld(d->cw.p, c), 2.0 * hdt);
            d->cq.p[4 * c] = q.q0; d->cq.p[4 * c + 1] = q.q1; d->cq.p[4 * c + 2] = q.q2; d->cq.p[4 * c + 3] = q.q3;
        }
        pebble_velocity(d);
    }
    ball_kick(d, g, hdt);
    for (long k = 0; k < 3 * d->radius.n; k++) d->pos.p[k] += 2.0 * hdt * d->vel.p[k];
}

void o_dem_second_half(o_dem* d, const double* gg, double hdt)
{
    o_vec3 g = v3(gg[0], gg[1], gg[2]);
    if (d->cps.n > 0) {
        clump_sum(d);
        clump_kick(d, g, hdt);
        pebble_velocity(d);
    }
    ball_kick(d, g, hdt);
}

/* ------------------------------------------------------------- hooks */
void o_dem_hook_clear(o_dem* d)
{
    for (int k = 0; k < d->nhook; k++) free(d->hook[k].f);
    d->nhook = 0;
}
void o_dem_hook_constant_force(o_dem* d, const double* f, double ramp)
{
    if (d->nhook >= 8) return;
    long n = d->radius.n;
    d->hook[d->nhook].kind = 0;
    d->hook[d->nhook].f = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    memcpy(d->hook[d->nhook].f, f, sizeof(double) * 3 * n);
    d->hook[d->nhook].nf = n;
    d->hook[d->nhook].ramp = ramp;
    d->nhook++;
}
void o_dem_hook_damping(o_dem* d, double c)
{
    if (d->nhook >= 8) return;
    d->hook[d->nhook].kind = 1;
    d->hook[d->nhook].c = c;
    d->hook[d->nhook].f = NULL;
    d->nhook++;
}
void o_dem_hook_buoyancy(o_dem* d, double rho, double cd, const double* g, const double* vf, double level)
{
    if (d->nhook >= 8) return;
    d->hook[d->nhook].kind = 2;
    d->hook[d->nhook].rho = rho;
    d->hook[d->nhook].cd = cd;
    d->hook[d->nhook].g = v3(g[0], g[1], g[2]);
    d->hook[d->nhook].vf = v3(vf[0], vf[1], vf[2]);
    d->hook[d->nhook].level = level;
    d->hook[d->nhook].f = NULL;
    d->nhook++;
}

/* sphereSubmergedVolumePlaneZ (externalForceTorque.cu:66-77) */
static double submerged(double zc, double r, double level)
{
    double dd = level - zc;
    if (dd <= -r) return 0.0;
    if (dd >= r) return (4.0 / 3.0) * O_PI * r * r * r;
    double h = dd + r;
    return (O_PI * h * h * (3.0 * r - h)) / 3.0;
}

static void run_hooks(o_dem* d, double time)
{
    long n = d->radius.n;
    for (int k = 0; k < d->nhook; k++) {
        if (d->hook[k].kind == 0) {
            /* addConstantForce (externalForceTorque.cu:4-12) with a ramp factor */
            double c = 1.0;
            if (d->hook[k].ramp > 0) { c = time / d->hook[k].ramp; if (c > 1.0) c = 1.0; }
            for (long i = 0; i < n && i < d->hook[k].nf; i++)
                vst(d->force.p, i, vadd(vld(d->force.p, i), vmul(vld(d->hook[k].f, i), c)));
        } else if (d->hook[k].kind == 1) {
            /* addGlobalDampingForceTorque (externalForceTorque.cu:26-37) */
            double c = d->hook[k].c;
            for (long i = 0; i < n; i++) {
                o_vec3 f = vld(d->force.p, i), t = vld(d->torque.p, i);
                vst(d->force.p, i, vsub(f, vmul(vnormalize(vld(d->vel.p, i)), c * vlen(f))));
                vst(d->torque.p, i, vsub(t, vmul(vnormalize(vld(d->w.p, i)), c * vlen(t))));
            }
        } else {
            /* addBuoyancyDragKernel (externalForceTorque.cu:79-133) */
            double rho = d->hook[k].rho, cd = d->hook[k].cd, level = d->hook[k].level;
            o_vec3 g = d->hook[k].g, vf = d->hook[k].vf;
            for (long i = 0; i < n; i++) {
                if (d->invMass.p[i] <= 0.0) continue;
                o_vec3 r = vld(d->pos.p, i), v = vld(d->vel.p, i);
                double rad = d->radius.p[i];
                double Vs = submerged(r.z, rad, level);
                if (Vs <= 0.0) continue;
                o_vec3 Fb = v3(-rho * Vs * g.x, -rho * Vs * g.y, -rho * Vs * g.z);
                double Vb = (4.0 / 3.0) * O_PI * rad * rad * rad;
                double alpha = (Vb > 1.e-30) ? (Vs / Vb) : 0.0;
                o_vec3 vr = vsub(v, vf);
                double v2 = vr.x * vr.x + vr.y * vr.y + vr.z * vr.z;
                o_vec3 Fd = v3(0, 0, 0);
                if (v2 > 1.e-30 && cd > 0.0) {
                    double vm = sqrt(v2);
                    double Aeff = alpha * (O_PI * rad * rad);
                    double kk = 0.5 * rho * cd * Aeff * vm;
                    Fd = v3(-kk * vr.x, -kk * vr.y, -kk * vr.z);
                }
                o_vec3 f = vld(d->force.p, i);
                f.x += Fb.x + Fd.x; f.y += Fb.y + Fd.y; f.z += Fb.z + Fd.z;
                vst(d->force.p, i, f);
            }
        }
    }
}

void o_dem_set_policy(o_dem* d, int d1) { d->policyD1 = d1; }

/* ------------------------------------------------------------- loop */
void o_dem_solve_begin(o_dem* d)
{
    o_dem_neighbor_search(d);
    o_dem_contact(d, 0.);
    if (d->t0.n > 0) o_dem_triangle_contact(d, 0.);
    d->iStep = 0;
    d->time = 0.0;
}

void o_dem_step(o_dem* d, const double* g, double dt, int nsteps)
{
    double h = 0.5 * dt;
    int walls = d->t0.n > 0;
    for (int k = 0; k < nsteps; k++) {
        d->iStep++;
        d->time += dt;
        if (d->iStep % 5 == 0) o_dem_neighbor_search(d);
        o_dem_contact(d, h);
        if (walls) o_dem_triangle_contact(d, dt);
        run_hooks(d, d->time - h);
        o_dem_first_half(d, g, h);
        o_dem_contact(d, h);
        if (walls) o_dem_triangle_contact(d, dt);
        run_hooks(d, d->time);
        o_dem_second_half(d, g, h);
    }
}

/* ------------------------------------------------------------- coupling */
void o_dem_couple_view(o_dem* d, o_dem_balls* v)
{
    v->n = (int)d->radius.n;
    v->pos = d->pos.p; v->vel = d->vel.p; v->w = d->w.p; v->radius = d->radius.p; v->invMass = d->invMass.p;
    v->force = d->force.p; v->torque = d->torque.p;
}

double o_dem_couple_begin(o_dem* d, double dt)
{
    d->iStep++;
    d->time += dt;
    if (d->iStep % 5 == 0) o_dem_neighbor_search(d);
    return d->time;
}

void o_dem_couple_evaluate(o_dem* d, double h, double dt)
{
    o_dem_contact(d, h);
    if (d->t0.n > 0) o_dem_triangle_contact(d, dt);
}

void o_dem_couple_kick(o_dem* d, int first, const double* g, double h, double hookTime)
{
    run_hooks(d, hookTime);
    if (first) o_dem_first_half(d, g, h);
    else o_dem_second_half(d, g, h);
}

size_t o_dem_step_count(o_dem* d) { return d->iStep; }

void o_dem_grids(o_dem* d, o_grid* balls, o_grid* tris)
{
    if (balls) *balls = d->bg;
    if (tris) *tris = d->tg;
}

long o_dem_get(o_dem* d, const char* name, void* out)
{
#define GD(nm, arr, cnt) if (!strcmp(name, nm)) { if (out && (cnt) > 0) memcpy(out, arr, sizeof(double) * (size_t)(cnt)); return (long)(cnt); }
#define GI(nm, arr, cnt) if (!strcmp(name, nm)) { if (out && (cnt) > 0) memcpy(out, arr, sizeof(int) * (size_t)(cnt)); return (long)(cnt); }
    long nb = d->radius.n;
    GD("pos", d->pos.p, 3 * nb) GD("vel", d->vel.p, 3 * nb) GD("angVel", d->w.p, 3 * nb)
    GD("force", d->force.p, 3 * nb) GD("torque", d->torque.p, 3 * nb) GD("radius", d->radius.p, nb)
    GD("invMass", d->invMass.p, nb)
    GI("pointed", d->bb.pointed, d->bb.n) GI("pointing", d->bb.pointing, d->bb.n)
    GI("pairShift", d->bb.shift, d->bb.n)
    GD("sliding", d->bb.sl, 3 * d->bb.n) GD("rolling", d->bb.ro, 3 * d->bb.n) GD("torsion", d->bb.to, 3 * d->bb.n)
    GD("overlap", d->bb.ov, d->bb.n) GD("contactForce", d->bb.F, 3 * d->bb.n)
    GD("contactTorque", d->bb.T, 3 * d->bb.n) GD("contactPoint", d->bb.P, 3 * d->bb.n)
    GI("neighborCount", d->bb.count, nb)
    GI("tri_pointed", d->bt.pointed, d->bt.n) GI("tri_pointing", d->bt.pointing, d->bt.n)
    GI("tri_cancel", d->bt.cancel, d->bt.n) GD("tri_overlap", d->bt.ov, d->bt.n)
    GD("tri_point", d->bt.P, 3 * d->bt.n) GD("tri_normal", d->bt.N, 3 * d->bt.n)
    GD("tri_force", d->bt.F, 3 * d->bt.n) GD("tri_sliding", d->bt.sl, 3 * d->bt.n)
    GI("bond_pointed", d->bpointed.p, d->bpointed.n) GI("bond_pointing", d->bpointing.p, d->bpointing.n)
    GI("bond_isBonded", d->bonded.p, d->bonded.n) GD("bond_normalForce", d->bFn.p, d->bFn.n)
    GD("bond_torsionTorque", d->bTt.p, d->bTt.n) GD("bond_shearForce", d->bshear.p, d->bshear.n)
    GD("bond_bendingTorque", d->bbend.p, d->bbend.n) GD("bond_point", d->bpoint.p, d->bpoint.n)
    GD("bond_normal", d->bnormal.p, d->bnormal.n) GD("bond_maxNormalStress", d->bsig.p, d->bsig.n)
    GD("bond_maxShearStress", d->btau.p, d->btau.n)
    GD("clump_pos", d->cpos.p, d->cpos.n) GD("clump_vel", d->cvel.p, d->cvel.n) GD("clump_angVel", d->cw.p, d->cw.n)
    GD("clump_q", d->cq.p, d->cq.n) GD("clump_force", d->cforce.p, d->cforce.n)
    GD("wall_pos", d->wpos.p, d->wpos.n) GD("wall_q", d->wq.p, d->wq.n) GD("vertex_global", d->vglobal.p, d->vglobal.n)
    GD("tri_circumcenter", d->tcc.p, d->tcc.n)
    GI("hash", d->bhash, nb) GI("index", d->bindex, nb) GI("cellStart", d->bcs, d->bng) GI("cellEnd", d->bce, d->bng)
    GI("tri_index", d->tindex, d->t0.n)
#undef GD
#undef GI
    return -1;
}

int o_dem_set(o_dem* d, const char* name, const void* in)
{
    long nb = d->radius.n;
    if (!strcmp(name, "pos")) { memcpy(d->pos.p, in, sizeof(double) * 3 * nb); return 0; }
    if (!strcmp(name, "vel")) { memcpy(d->vel.p, in, sizeof(double) * 3 * nb); return 0; }
    if (!strcmp(name, "angVel")) { memcpy(d->w.p, in, sizeof(double) * 3 * nb); return 0; }
    if (!strcmp(name, "wall_vel")) { memcpy(d->wvel.p, in, sizeof(double) * d->wvel.n); return 0; }
    if (!strcmp(name, "wall_angVel")) { memcpy(d->ww.p, in, sizeof(double) * d->ww.n); return 0; }
    return -1;
}
