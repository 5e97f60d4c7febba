// WCSPH scene for MI355X: cell-sorted state, per-step hash/sort/reorder,
// chunked ELL neighbour lists and the three Riemann pair passes of one step.
//
// Reference: DEMSPH/kernel/WCSPHSolver.h (step loop :607-621, neighbour
// search :150-285), SPHNeighborSearchKernel.cu (lists), and
// WCSPHIntegrationKernel.cu (acceleration :42-148, density change :172-264,
// EOS :266-282, dummy normals :4-40, KDK sequence :308-472).
//
// Layout in HBM (slot order = lexicographic (cell hash, insertion index),
// rebuilt every step; "orig" maps slot -> insertion index, "rank" the
// inverse).  Everything a neighbour contributes to a pair is packed into two
// records so a pair costs three wide gathers:
//   FluidPos  {x, y, z: f64; rho, p: f32}     32 B  (pair record 1)
//   float4    {vx, vy, vz: f32; phase bits}  16 B  (pair record 2)
//   velocity  3 x f64[n]                     own-slot only (integration)
//   drho      f32[n]                         own-slot only
// Dummies (sorted once): DummyPos {x, y, z: f64; nx, ny: f32} and float4
// {vx, vy, vz, nz}.
// Neighbour lists are ELL in 32-byte chunks: entries 8c..8c+7 of slot i sit
// at int offset (c*n + i)*8.  The list kernel fills a chunk in registers and
// stores it whole, so every store covers full 32-byte sectors; the lanes of
// a wavefront read one chunk of consecutive slots as consecutive 16-byte
// words.  Each slot's list holds neighbour slots in the reference's exact
// walk order (z, y, x cells, then ascending insertion index), so the pair
// set and its per-particle order are identical to the reference's.
//
// Pair arithmetic is FP32 (positions differenced in FP64 first); state
// integration is FP64.
#include "common.hpp"
#include "internal.hpp"
#include "timing.hpp"
#include "vtu.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

namespace demsph {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxPhases = 256;
constexpr int kChunk = 8;  // list entries per 32-byte chunk

struct Phase {
    double h;       // smoothing length, FP64 for the exact neighbour cut
    float hf;       // smoothing length
    float m;        // mass
    float c;        // sound speed
    float nu;       // kinematic viscosity
    float rho0;     // reference density
    float c2;       // c^2
    float taitB;    // rho0 c^2 / gamma
    float pad;
};

struct alignas(32) FluidPos {
    double x, y, z;
    float rho, p;
};
static_assert(sizeof(FluidPos) == 32, "FluidPos is one 32-byte record");

struct alignas(32) DummyPos {
    double x, y, z;
    float nx, ny;
};
static_assert(sizeof(DummyPos) == 32, "DummyPos is one 32-byte record");

// A coupled DEM ball seen from the fluid: centre (FP64), radius and the
// fluid pressure interpolated at its centre.  Velocity and spin travel in
// two float4 records.
struct alignas(32) SolidPos {
    double x, y, z;
    float r, P;
};
static_assert(sizeof(SolidPos) == 32, "SolidPos is one 32-byte record");

// The 32-byte record a pair pass gathers for a fluid neighbour (uniform
// scenes): the position as an FP32 offset from the origin of the particle's
// cell at the last sort, density, pressure and velocity.  Two 16-byte loads
// per neighbour where the FP64 position, density and pressure (32 bytes)
// plus the velocity (12) take three; the pair passes issue about one
// vector-L1 line access per CU per cycle, so the loads per neighbour set
// their time.  A list entry then carries the neighbour's cell relative to
// the particle's stencil centre, 2 bits per axis above the slot, and
// x_i - x_j = (o_i - o_j) - (c_j - c_i) cs: the offsets are below 2 cells,
// so the difference keeps ~1e-9 m absolute (3e-7 of a typical pair
// distance), against 3e-8 relative when rounding the FP64 difference.
//
// The two 16-byte halves of the records lie in two arrays of nf float4 (one
// buffer: {ox, oy, oz, rho} at [0, nf), {p, vx, vy, vz} at [nf, 2 nf)): a
// lane group's neighbours are mostly consecutive slots, so each gather
// instruction touches half the cache lines of 32-byte records.
struct alignas(32) GRec {
    float ox, oy, oz, rho;
    float p, vx, vy, vz;
};
static_assert(sizeof(GRec) == 32, "GRec is two 16-byte halves");
__device__ __forceinline__ void st_grec(float4* base, int half, int k, const GRec& r)
{
    base[k] = make_float4(r.ox, r.oy, r.oz, r.rho);
    base[half + k] = make_float4(r.p, r.vx, r.vy, r.vz);
}
constexpr int kCodeShift = 25;                      // entry = slot | code << 25
constexpr int kSlotMask = (1 << kCodeShift) - 1;    // slots below 2^25
__host__ __device__ constexpr int pack_cell(int ix, int iy, int iz) { return ix | iy << 11 | iz << 22; }

__device__ __forceinline__ unsigned phase_bits(float w) { return __float_as_uint(w); }

struct SphArgs {
    int nf, nd;
    const Phase* phases;
    Phase phf, phd;  // the single fluid / dummy phase of uniform scenes
    // fluid, current buffers
    const FluidPos* P;
    const float4* V;
    const double *vx, *vy, *vz;
    const float* drho;
    // fluid, destination buffers
    FluidPos* oP;
    float4* oV;
    double *ovx, *ovy, *ovz;
    float* odrho;
    // lists
    const int *nbr_ff, *cnt_ff, *nbr_fd, *cnt_fd;
    int K_ff, K_fd;
    // gather records (null: FP64 records): read and written by the passes,
    // each particle's packed cell at the last sort and stencil centre
    const float4* GRin;  // split halves, nf float4 apart
    float4* GRout;
    const int *pcell, *pcen;
    GridGeom grid;
    // dummies (sorted once)
    const DummyPos* DP;
    const float4* DV;
    const uint8_t* dphase;
    // coupled DEM balls (solids), sorted into the fluid grid each step
    int ns;
    const SolidPos* SP;
    const float4 *SV, *SW;  // velocity, angular velocity
    const int *nbr_fs, *cnt_fs;
    int K_fs;
    float* acs;             // per fluid slot: coupling acceleration of the last pass (3 floats), or null
    // physics
    float gx, gy, gz;
    double dt;
    int eos;
    float gamma, inv_gamma;
};

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Wendland C2 gradient factor f with gradW = f * r_ij
// (WCSPHIntegrationKernel.cuh:15-32).  The reference's
// sigma/h * [-2 t^3 (1+2q) + 2 t^4] / r equals -5 sigma t^3 / h^2 with
// t = 1 - q/2, q = r/h, sigma = 21 / (16 pi h^3).  With a uniform h every
// factor but t is loop-invariant and hoisted by the compiler.
// Evaluated unconditionally and selected, so that a pair loop has no
// divergent branch between its gathers and their use.
struct GradK {
    float half_invh, coef, two_h;
};
__device__ __forceinline__ GradK grad_k(float h)
{
    GradK k;
    const float invh = 1.0f / h;
    k.half_invh = 0.5f * invh;  // exact: t below equals 1 - 0.5 r invh
    const float sigma = (float)(21.0 / (16.0 * M_PI)) * invh * invh * invh;
    k.coef = -5.0f * sigma * invh * invh;
    k.two_h = 2.0f * h;
    return k;
}
__device__ __forceinline__ float grad_factor(float r, const GradK& k)
{
    const float t = 1.0f - r * k.half_invh;
    const float f = k.coef * (t * t * t);
    return (r < 1.e-20f || r >= k.two_h) ? 0.0f : f;
}
__device__ __forceinline__ float grad_factor(float r, float h) { return grad_factor(r, grad_k(h)); }


// Inverse EOS for the dummy density (WCSPHIntegrationKernel.cu:132,249):
// linear rho = P/c^2 + rho0; Tait rho = rho0 (1 + P/B)^(1/gamma).
__device__ __forceinline__ float dummy_rho(const SphArgs& a, const Phase& pj, float PR)
{
    if (a.eos == 1) {
        float base = 1.0f + PR / pj.taitB;
        base = fmaxf(base, 1.e-12f);
        return pj.rho0 * __powf(base, a.inv_gamma);
    }
    return PR / pj.c2 + pj.rho0;
}

// EOS of SPHDensityPressureIntegrationKernel (WCSPHIntegrationKernel.cu:266-282)
// with the Tait option.
__device__ __forceinline__ void eos_eval(int eos, float gamma, const Phase& ph, float drho, float& rho, float& p)
{
    float ch = drho;
    if (ch < -0.1f * ph.rho0) ch = -0.1f * ph.rho0;
    rho = ch + ph.rho0;
    if (eos == 1) {
        const double ratio = (double)rho / (double)ph.rho0;
        p = (float)((double)ph.taitB * (pow(ratio, (double)gamma) - 1.0));
    } else {
        p = ph.c2 * ch;
    }
}

// ----------------------------------------------------------- DEM-SPH coupling
// A coupled DEM ball j is a solid particle of the fluid phase for the fluid
// particles within its support: mass rho0 * 4/3 pi r_j^3, pressure P_j
// interpolated from the fluid at its centre, density from the inverse EOS
// of P_j, moving with its surface velocity v_j + w_j x (r_j n) where n points
// from the ball centre to the fluid particle.  The interface pressure is the
// Riemann one of the fluid-fluid branch (WCSPHIntegrationKernel.cu:42-112)
// with P_R = P_j.  The reference has no coupling; this is the project's
// model.  The fluid particle's acceleration contribution is returned split
// into its central pressure part `ap` and viscous part `av`; the ball
// receives -m_i (ap + av) and the torque (r_j n) x (-m_i av).  Contraction
// is off so that the fluid pass and the ball pass, which both call this,
// compute bit-identical values (exact pairwise momentum exchange).
struct FsGeom {
    float rx, ry, rz, r, inv_r, f;   // r_ij = x_i - x_j, |r_ij|, 1/|r_ij|, kernel factor
    float vsx, vsy, vsz;             // ball surface velocity at the pair
    float lx, ly, lz;                // lever r_j n
    float rhoR, mj;                  // ball density and fluid-equivalent mass
};

__device__ __forceinline__ FsGeom fs_geom(const SphArgs& a, const Phase& pi, double xi, double yi, double zi,
                                          const SolidPos& S, float4 SV, float4 SW)
{
#pragma clang fp contract(off)
    FsGeom g;
    g.rx = (float)(xi - S.x); g.ry = (float)(yi - S.y); g.rz = (float)(zi - S.z);
    const float r2 = g.rx * g.rx + g.ry * g.ry + g.rz * g.rz;
    g.inv_r = (r2 < 1.e-20f) ? 0.0f : rsq(r2);
    g.r = r2 * g.inv_r;
    const float h = pi.hf;
    if (g.r < 1.e-20f || g.r >= 2.0f * h) {
        g.f = 0.0f;
    } else {
        const float invh = 1.0f / h;
        const float t = 1.0f - 0.5f * g.r * invh;
        const float sigma = (float)(21.0 / (16.0 * M_PI)) * invh * invh * invh;
        g.f = -5.0f * sigma * invh * invh * (t * t * t);
    }
    const float s = S.r * g.inv_r;
    g.lx = s * g.rx; g.ly = s * g.ry; g.lz = s * g.rz;
    g.vsx = SV.x + (SW.y * g.lz - SW.z * g.ly);
    g.vsy = SV.y + (SW.z * g.lx - SW.x * g.lz);
    g.vsz = SV.z + (SW.x * g.ly - SW.y * g.lx);
    if (a.eos == 1) {
        const float base = fmaxf(1.0f + S.P / pi.taitB, 1.e-12f);
        g.rhoR = pi.rho0 * __powf(base, a.inv_gamma);
    } else {
        g.rhoR = S.P / pi.c2 + pi.rho0;
    }
    g.mj = pi.rho0 * ((float)(4.0 / 3.0 * M_PI) * (S.r * S.r * S.r));
    return g;
}

__device__ __forceinline__ void fs_momentum(const SphArgs& a, const Phase& pi, double xi, double yi, double zi,
                                            float vix, float viy, float viz, float rhoi, float PL, const SolidPos& S,
                                            float4 SV, float4 SW, float3& ap, float3& av, float3& lever)
{
#pragma clang fp contract(off)
    const FsGeom g = fs_geom(a, pi, xi, yi, zi, S, SV, SW);
    const float UL = -(vix * g.rx + viy * g.ry + viz * g.rz) * g.inv_r;
    const float UR = -(g.vsx * g.rx + g.vsy * g.ry + g.vsz * g.rz) * g.inv_r;
    const float cbar = pi.c, rbar = 0.5f * (rhoi + g.rhoR);
    const float beta = fminf(3.0f * fmaxf(UL - UR, 0.0f), cbar);
    const float Ps = 0.5f * (PL + S.P) + 0.5f * beta * rbar * (UL - UR);
    const float mr = g.mj / (rhoi * g.rhoR);
    const float kp = -2.0f * mr * Ps * g.f;
    const float kv = 2.0f * pi.nu * (g.f * g.r) * mr;
    ap = make_float3(kp * g.rx, kp * g.ry, kp * g.rz);
    av = make_float3(kv * (vix - g.vsx), kv * (viy - g.vsy), kv * (viz - g.vsz));
    lever = make_float3(g.lx, g.ly, g.lz);
}

__device__ __forceinline__ float fs_density(const SphArgs& a, const Phase& pi, double xi, double yi, double zi,
                                            float vix, float viy, float viz, float rhoi, float PL, const SolidPos& S,
                                            float4 SV, float4 SW)
{
    const FsGeom g = fs_geom(a, pi, xi, yi, zi, S, SV, SW);
    const float cbar = pi.c, rbar = 0.5f * (rhoi + g.rhoR);
    const float t = 0.5f * ((vix - g.vsx) * g.rx + (viy - g.vsy) * g.ry + (viz - g.vsz) * g.rz) +
                    0.5f * (PL - S.P) * g.r * rcp(rbar * cbar);
    return g.mj * rcp(g.rhoR) * g.f * t;
}

// ----------------------------------------------------------- hash & reorder
// Cell hash of every slot, scattered by insertion index for the stable sort.
// When prev_skey is given (the sorted keys that produced the current cell
// ranges), the first slot of each previously occupied cell also clears that
// cell's range, which replaces a fill of the whole grid.
__global__ __launch_bounds__(kThreads) void k_hash(int n, GridGeom g, long long numGrids,
                                                   const FluidPos* __restrict__ P, const int* __restrict__ orig,
                                                   uint32_t* __restrict__ key, const uint32_t* __restrict__ prev_skey,
                                                   int2* __restrict__ cells)
{
    const int s = blockIdx.x * kThreads + threadIdx.x;
    if (s >= n) return;
    if (prev_skey) {
        const uint32_t h = prev_skey[s];
        if ((long long)h < numGrids && (s == 0 || prev_skey[s - 1] != h)) cells[h] = make_int2(-1, -1);
    }
    const FluidPos p = P[s];
    key[orig[s]] = (uint32_t)cell_hash(g, p.x, p.y, p.z);
}

// FP32 copy of a position relative to the grid origin for the list
// prefilter; w = 1 marks a point farther than R from the origin on some
// axis, where the FP32 rounding could reach the prefilter's band.
__device__ __forceinline__ float4 pos_f32(const GridGeom& g, double R, double x, double y, double z)
{
    const double rx = x - g.minx, ry = y - g.miny, rz = z - g.minz;
    const bool far = fabs(rx) > R || fabs(ry) > R || fabs(rz) > R;
    return make_float4((float)rx, (float)ry, (float)rz, far ? 1.f : 0.f);
}

// Origin of a packed cell, and a particle's gather record in it.
__device__ __forceinline__ void cell_origin(const GridGeom& g, int pc, double& x, double& y, double& z)
{
    x = g.minx + (double)(pc & 2047) * g.cx;
    y = g.miny + (double)((pc >> 11) & 2047) * g.cy;
    z = g.minz + (double)((pc >> 22) & 1023) * g.cz;
}
__device__ __forceinline__ GRec grec(const GridGeom& g, int pc, const FluidPos& p, const float4& v)
{
    double x0, y0, z0;
    cell_origin(g, pc, x0, y0, z0);
    GRec r;
    r.ox = (float)(p.x - x0); r.oy = (float)(p.y - y0); r.oz = (float)(p.z - z0);
    r.rho = p.rho; r.p = p.p;
    r.vx = v.x; r.vy = v.y; r.vz = v.z;
    return r;
}

struct ReorderArgs {
    int n;
    long long numGrids;
    const uint32_t* skey;  // sorted keys
    const int* perm;       // sorted insertion indices
    int* rank;             // insertion index -> slot (updated)
    int* orig_out;
    const FluidPos* P;
    const float4* V;
    const double *vx, *vy, *vz;
    const float* drho;
    FluidPos* oP;
    float4* oV;
    double *ovx, *ovy, *ovz;
    float* odrho;
    int2* cells;  // (start, end) per cell
    float4* oPf;  // FP32 positions for the list prefilter (null: none)
    GridGeom grid;
    double R;
    float4* oGR;  // gather records (split halves, n apart) and packed cells (null: none)
    int* opcell;
};

__global__ __launch_bounds__(kThreads) void k_reorder(ReorderArgs a)
{
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= a.n) return;
    const int o = a.perm[k];
    const int s = a.rank[o];
    a.rank[o] = k;
    a.orig_out[k] = o;
    const FluidPos p = a.P[s];
    a.oP[k] = p;
    if (a.oPf) a.oPf[k] = pos_f32(a.grid, a.R, p.x, p.y, p.z);
    a.oV[k] = a.V[s];
    a.ovx[k] = a.vx[s]; a.ovy[k] = a.vy[s]; a.ovz[k] = a.vz[s];
    a.odrho[k] = a.drho[s];
    // findStartAndEnd (buildHashStartEnd.cu:14-34)
    const uint32_t h = a.skey[k];
    if (a.oGR) {
        const int nxy = a.grid.nx * a.grid.ny;
        const int hh = (int)h, iz = hh / nxy, r = hh - iz * nxy, iy = r / a.grid.nx, ix = r - iy * a.grid.nx;
        const int pc = pack_cell(ix, iy, iz);
        a.opcell[k] = pc;
        const float4 v = a.V[s];
        st_grec(a.oGR, a.n, k, grec(a.grid, pc, p, v));
    }
    if ((long long)h < a.numGrids) {
        if (k == 0 || a.skey[k - 1] != h) a.cells[h].x = k;
        if (k == a.n - 1 || a.skey[k + 1] != h) a.cells[h].y = k + 1;
    }
}

// ----------------------------------------------------------- neighbour lists
struct ListArgs {
    int nf, nd;
    GridGeom grid;
    const Phase* phases;
    double hf, hd;  // uniform fluid / dummy smoothing lengths
    const FluidPos* P;
    const float4* V;
    const int2* cells;  // fluid cell ranges
    const DummyPos* DP;
    const uint8_t* dphase;
    const int2* dcells;          // dummy cell ranges
    const uint8_t* dummy_near;   // per cell: a dummy cell lies in its 27-cell stencil
    int *nbr_ff, *cnt_ff, *nbr_fd, *cnt_fd;
    int K_ff, K_fd;  // multiples of kChunk
    int* maxcount;   // [0] fluid, [1] dummy, [2] coupled balls
    // coupled balls: cell ranges over the fluid grid, records, lists
    const int2* scells;
    const SolidPos* SP;
    int *nbr_fs, *cnt_fs;
    int K_fs;
    // FP32 prefilter of the lane walks: positions relative to the grid
    // origin (w = 1: far from it), and the relative half-width of the band
    // around cut^2 in which the exact FP64 test decides
    const float4 *Pf, *DPf;
    float band;
    int* pcen;  // coded lists: each particle's packed stencil centre (null: plain slots)
};

enum WalkKind { kFluid = 0, kDummy = 1, kSolid = 2 };

// One 32-byte list chunk held in registers while it fills.
struct Chunk {
    int e0, e1, e2, e3, e4, e5, e6, e7;
    __device__ __forceinline__ void put(int q, int j)
    {
        e0 = q == 0 ? j : e0; e1 = q == 1 ? j : e1; e2 = q == 2 ? j : e2; e3 = q == 3 ? j : e3;
        e4 = q == 4 ? j : e4; e5 = q == 5 ? j : e5; e6 = q == 6 ? j : e6; e7 = q == 7 ? j : e7;
    }
    __device__ __forceinline__ void store(int* dst) const
    {
        reinterpret_cast<int4*>(dst)[0] = make_int4(e0, e1, e2, e3);
        reinterpret_cast<int4*>(dst)[1] = make_int4(e4, e5, e6, e7);
    }
};

// Walks the 27-cell stencil of one point over one sorted set and appends
// every member within 2*max(hA,hB) to the point's chunked list
// (SPHNeighborSearchKernel.cu:57-115 with excludeSelf, :169-228 without).
// The three x-cells of a stencil row are consecutive hashes, so their
// members form one contiguous run of slots, walked in ascending order: the
// nine row runs are resolved first (all cell-range loads in flight at
// once), then each run is walked kWalkAhead candidates at a time with
// their position loads issued before any test.
#ifndef DEMSPH_WALK_AHEAD
#define DEMSPH_WALK_AHEAD 4
#endif
constexpr int kWalkAhead = DEMSPH_WALK_AHEAD;
// Candidates per lane per batch in the lane-group walk.  Three beat four
// and two (256k points, 4 lanes: 130, 155 and 135 us; 741k: 324, 393 and
// 332 us): fewer values in flight keep more waves resident and waste less
// of each row's last batch.
constexpr int kLaneAhead = 3;

template <bool UNIF, int KIND, class PosT, bool CODED = false>
__device__ __forceinline__ int walk_list(const ListArgs& a, const PosT* __restrict__ B,
                                         const int2* __restrict__ cells, double x, double y, double z, double hA,
                                         int self, int* __restrict__ out, int K)
{
    const GridGeom& g = a.grid;
    // balls couple to the fluid particles within the fluid's support 2 h_A
    const double cut_u = KIND == kSolid ? 2.0 * hA : 2.0 * fmax(hA, KIND == kDummy ? a.hd : a.hf);
    const size_t cstride = (size_t)a.nf * kChunk;
    const int nchunks = K / kChunk;
    const Stencil st = make_stencil(g, x, y, z);
    int rs[9], re[9], rb0[9], rb1[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int dz = k / 3 - 1, dy = k % 3 - 1;
        const bool row_ok = dz >= st.loz && dz <= st.hiz && dy >= st.loy && dy <= st.hiy;
        const int row = (st.cz + dz) * g.ny * g.nx + (st.cy + dy) * g.nx + st.cx;
        int s = -1, e = -1;
        int2 rr[3];
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const bool ok = row_ok && dx >= st.lox && dx <= st.hix;
            const int2 c = cells[ok ? row + dx : 0];
            const int2 r = ok ? c : make_int2(-1, -1);
            rr[dx + 1] = r;
            if (r.x != -1) {
                if (s < 0) s = r.x;
                e = r.y;
            }
        }
        rs[k] = s < 0 ? 0 : s;
        re[k] = s < 0 ? 0 : e;
        // a slot's x-cell in the row: -1 below b0, 0 below b1, +1 above
        rb0[k] = rr[0].x != -1 ? rr[0].y : rs[k];
        rb1[k] = rr[1].x != -1 ? rr[1].y : rb0[k];
    }
    Chunk buf{0, 0, 0, 0, 0, 0, 0, 0};
    int count = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int e = re[k];
        const int rowcode = ((k / 3) << 4) | ((k % 3) << 2);  // (dz + 1) << 4 | (dy + 1) << 2
        for (int j0 = rs[k]; j0 < e; j0 += kWalkAhead) {
            double qx[kWalkAhead], qy[kWalkAhead], qz[kWalkAhead];
#pragma unroll
            for (int u = 0; u < kWalkAhead; u++) {
                const PosT* b = B + min(j0 + u, e - 1);
                qx[u] = b->x; qy[u] = b->y; qz[u] = b->z;
            }
#pragma unroll
            for (int u = 0; u < kWalkAhead; u++) {
                const int j = j0 + u;
                if (j >= e || (KIND == kFluid && j == self)) continue;
                double cut = cut_u;
                if (!UNIF && KIND != kSolid) {
                    const double hB = KIND == kDummy ? a.phases[a.dphase[j]].h : a.phases[phase_bits(a.V[j].w)].h;
                    cut = 2.0 * fmax(hA, hB);
                }
                if (!within_cut(cut, x - qx[u], y - qy[u], z - qz[u])) continue;
                const int c = count & (kChunk - 1);
                buf.put(c, CODED ? j | (rowcode | (j >= rb0[k]) + (j >= rb1[k])) << kCodeShift : j);
                count++;
                if (c == kChunk - 1) {
                    const int cc = (count >> 3) - 1;
                    if (cc < nchunks) buf.store(out + (size_t)cc * cstride);
                }
            }
        }
    }
    if ((count & (kChunk - 1)) && (count >> 3) < nchunks) buf.store(out + (size_t)(count >> 3) * cstride);
    return count;
}

// walk_list over L lanes of one point: in each run the lanes test
// consecutive candidates (kLaneAhead per lane in flight), and a ballot over
// the lane group ranks the hits, so the list holds the same entries in the
// same order as walk_list's; entries are stored one int at a time.
//
// PRE: candidates are first tested in FP32 on the prefilter copy Bf; only
// those within the band around cut^2 (or marked far from the origin) take
// the exact FP64 test, so the decisions are those of within_cut (the host
// sizes the band above the FP32 rounding of points within R of the origin).
// The FP32 decisions of a batch are branch-free; the exact test sits behind
// one wave-uniform branch that a batch takes only when some lane has a
// candidate in the band.  List slots are 32-bit element offsets from the
// point's first chunk (the host keeps K * nf below 2^31).  The nine rows
// are walked in a rolled loop with the next row's three cell ranges loaded
// ahead: against nine unrolled rows with all ranges loaded up front, this
// keeps the kernel under 128 VGPRs and a tenth of the code size, and builds
// the lists 20 % faster (256k points: 153 vs 191 us).
template <bool UNIF, int KIND, class PosT, int L, bool PRE, bool CODED = false>
__device__ __forceinline__ int walk_list_lanes(const ListArgs& a, const PosT* __restrict__ B,
                                               const float4* __restrict__ Bf, const int2* __restrict__ cells,
                                               double x, double y, double z, float4 xf, double hA, int self,
                                               int* __restrict__ out, int K, int q)
{
    const GridGeom& g = a.grid;
    const double cut_u = KIND == kSolid ? 2.0 * hA : 2.0 * fmax(hA, KIND == kDummy ? a.hd : a.hf);
    const unsigned cstride = (unsigned)a.nf * (unsigned)kChunk;
    const Stencil st = make_stencil(g, x, y, z);
    // row k's three cell ranges, loaded one row ahead of the walk
    const auto row_load = [&](int k, int2* c) {
        const int dz = k / 3 - 1, dy = k % 3 - 1;
        const bool row_ok = dz >= st.loz && dz <= st.hiz && dy >= st.loy && dy <= st.hiy;
        const int row = (st.cz + dz) * g.ny * g.nx + (st.cy + dy) * g.nx + st.cx;
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const bool ok = row_ok && dx >= st.lox && dx <= st.hix;
            const int2 v = cells[ok ? row + dx : 0];
            c[dx + 1] = ok ? v : make_int2(-1, -1);
        }
    };
    int2 nxt[3];
    row_load(0, nxt);
    // this lane's L-bit window of a ballot: the 32-bit half holding the
    // group, then a bit-field extract at the group's offset in that half
    const int lane = threadIdx.x & 63;
    const unsigned goff = (unsigned)((lane - q) & 31);
    const bool hi_half = lane >= 32;
    // a point far from the origin takes the exact test for every candidate
    const float c2_u = (float)(cut_u * cut_u);
    const float band_u = PRE && xf.w != 0.f ? INFINITY : a.band * c2_u;
    int count = 0;
#pragma unroll 1
    for (int k = 0; k < 9; k++) {
        const int2 cur[3] = {nxt[0], nxt[1], nxt[2]};
        if (k < 8) row_load(k + 1, nxt);
        int s0 = -1, e = -1;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if (cur[c].x != -1) {
                if (s0 < 0) s0 = cur[c].x;
                e = cur[c].y;
            }
        }
        if (s0 < 0) { s0 = 0; e = 0; }
        // coded lists: a slot's x-cell in the row is -1 below b0, 0 below b1, +1 above
        const int b0 = cur[0].x != -1 ? cur[0].y : s0;
        const int b1 = cur[1].x != -1 ? cur[1].y : b0;
        const int rowcode = ((k / 3) << 4) | ((k % 3) << 2);  // (dz + 1) << 4 | (dy + 1) << 2
        for (int j0 = s0; j0 < e; j0 += L * kLaneAhead) {
            double qx[kLaneAhead], qy[kLaneAhead], qz[kLaneAhead];
            float4 qf[kLaneAhead];
#pragma unroll
            for (int u = 0; u < kLaneAhead; u++) {
                const unsigned jj = (unsigned)min(j0 + u * L + q, e - 1);
                if (PRE) {
                    qf[u] = Bf[jj];
                } else {
                    const PosT* b = B + jj;
                    qx[u] = b->x; qy[u] = b->y; qz[u] = b->z;
                }
            }
            bool hit[kLaneAhead];
            double cutv[kLaneAhead];
            bool need_any = false;
            bool need[kLaneAhead];
#pragma unroll
            for (int u = 0; u < kLaneAhead; u++) {
                const int j = j0 + u * L + q;
                // bitwise, not short-circuit, operators: no branches per candidate
                const bool valid = (j < e) & !(KIND == kFluid && j == self);
                double cut = cut_u;
                if (!UNIF && KIND != kSolid) {
                    const int jc = min(j, e - 1);
                    const double hB =
                        KIND == kDummy ? a.phases[a.dphase[jc]].h : a.phases[phase_bits(a.V[jc].w)].h;
                    cut = 2.0 * fmax(hA, hB);
                }
                cutv[u] = cut;
                if (PRE) {
                    const float c2 = UNIF || KIND == kSolid ? c2_u : (float)(cut * cut);
                    const float band = UNIF || KIND == kSolid ? band_u : (xf.w != 0.f ? INFINITY : a.band * c2);
                    const float fx = xf.x - qf[u].x, fy = xf.y - qf[u].y, fz = xf.z - qf[u].z;
                    const float d2 = fx * fx + fy * fy + fz * fz;
                    need[u] = valid & ((qf[u].w != 0.f) | (fabsf(d2 - c2) <= band));
                    hit[u] = valid & (d2 < c2);
                    need_any |= need[u];
                } else {
                    hit[u] = valid & within_cut(cut, x - qx[u], y - qy[u], z - qz[u]);
                }
            }
            if (PRE && __ballot(need_any) != 0ull) {
#pragma unroll
                for (int u = 0; u < kLaneAhead; u++) {
                    if (need[u]) {
                        const PosT* b = B + (unsigned)(j0 + u * L + q);
                        hit[u] = within_cut(cutv[u], x - b->x, y - b->y, z - b->z);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kLaneAhead; u++) {
                const unsigned long long bal = __ballot(hit[u]);
                const unsigned half = hi_half ? (unsigned)(bal >> 32) : (unsigned)bal;
                const unsigned gm = __builtin_amdgcn_ubfe(half, goff, L);
                const int kk = count + __popc(gm & ((1u << q) - 1u));
                const int j = j0 + u * L + q;
                if (hit[u] & (kk < K))
                    out[(unsigned)(kk >> 3) * cstride + (unsigned)(kk & (kChunk - 1))] =
                        CODED ? j | (rowcode | (j >= b0) + (j >= b1)) << kCodeShift : j;
                count += __popc(gm);
            }
        }
    }
    return count;
}

template <bool UNIF, int LANES, bool PRE, bool CODED>
__global__ __launch_bounds__(kThreads) void k_build_lists(ListArgs a)
{
    const int gt = xcd_block() * kThreads + threadIdx.x;
    const int i = gt / LANES, q = gt % LANES;
    if (i >= a.nf) return;  // uniform over a lane group
    const FluidPos p = a.P[i];
    const float4 pf = PRE ? a.Pf[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const double h = UNIF ? a.hf : a.phases[phase_bits(a.V[i].w)].h;
    int* of = a.nbr_ff + (size_t)i * kChunk;
    const int c1 = LANES == 1 ? walk_list<UNIF, kFluid, FluidPos, CODED>(a, a.P, a.cells, p.x, p.y, p.z, h, i, of,
                                                                          a.K_ff)
                              : walk_list_lanes<UNIF, kFluid, FluidPos, LANES, PRE, CODED>(
                                    a, a.P, a.Pf, a.cells, p.x, p.y, p.z, pf, h, i, of, a.K_ff, q);
    int c2 = 0;
    if (a.nd > 0 && a.dummy_near[cell_hash(a.grid, p.x, p.y, p.z)]) {
        int* od = a.nbr_fd + (size_t)i * kChunk;
        c2 = LANES == 1 ? walk_list<UNIF, kDummy>(a, a.DP, a.dcells, p.x, p.y, p.z, h, -1, od, a.K_fd)
                        : walk_list_lanes<UNIF, kDummy, DummyPos, LANES, PRE>(
                              a, a.DP, a.DPf, a.dcells, p.x, p.y, p.z, pf, h, -1, od, a.K_fd, q);
    }
    if (q != 0) return;
    a.cnt_ff[i] = c1;
    a.cnt_fd[i] = c2;
    if (CODED) {
        const Stencil st = make_stencil(a.grid, p.x, p.y, p.z);
        a.pcen[i] = pack_cell(st.cx, st.cy, st.cz);
    }
    if (c1 > a.K_ff) atomicMax(&a.maxcount[0], c1);
    if (c2 > a.K_fd) atomicMax(&a.maxcount[1], c2);
}

// Fluid -> coupled-ball lists over the balls sorted into the fluid grid.
__global__ __launch_bounds__(kThreads) void k_build_fs_lists(ListArgs a)
{
    const int i = xcd_block() * kThreads + threadIdx.x;
    if (i >= a.nf) return;
    const FluidPos p = a.P[i];
    const double h = a.phases[phase_bits(a.V[i].w)].h;
    const int c = walk_list<true, kSolid>(a, a.SP, a.scells, p.x, p.y, p.z, h, -1, a.nbr_fs + (size_t)i * kChunk,
                                          a.K_fs);
    a.cnt_fs[i] = c;
    if (c > a.K_fs) atomicMax(&a.maxcount[2], c);
}

// ----------------------------------------------------------- pair passes
// Pair loops consume one list chunk at a time: the chunk's eight neighbour
// records are gathered independently.  Entries past the list's end are
// replaced by a valid slot with a zero kernel factor, which adds exact zeros
// and keeps the loop free of branches.
#ifndef DEMSPH_PAIR_BATCH
#define DEMSPH_PAIR_BATCH 4
#endif
// Neighbour records gathered together per batch: 4 keeps the pair kernels
// under 128 VGPRs; 8 doubles the gathers in flight per wavefront at lower
// occupancy.
constexpr int kBatch = DEMSPH_PAIR_BATCH;
static_assert(kBatch >= 1 && kBatch <= 8, "pair batch within a chunk");

// LANES consecutive lanes share one particle: lane q takes the list chunks
// q, q + LANES, ... (LANES = 1: every chunk in order), so a small scene
// keeps more gathers in flight per particle.
template <int LANES, class F>
__device__ __forceinline__ void for_chunks(const int* __restrict__ list, int cnt, int K, int nf, int i, int q,
                                           int safe, F&& body)
{
    const int n = min(cnt, K);
    const int* L = list + (size_t)i * kChunk;
    const size_t cs = (size_t)nf * kChunk;
    for (int c = q; c * kChunk < n; c += LANES) {
        const int base = c * kChunk;
        const int4* src = reinterpret_cast<const int4*>(L + (size_t)c * cs);
        const int4 lo = src[0];
        const int4 hi = base + 4 < n ? src[1] : make_int4(safe, safe, safe, safe);
        const int jall[kChunk] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (int b0 = 0; b0 < kChunk; b0 += kBatch) {
            if (base + b0 >= n) break;
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                if (b0 + u >= kChunk) break;
                const bool valid = base + b0 + u < n;
                body(valid ? jall[b0 + u] : safe, valid);
            }
        }
    }
}

// for_chunks in two phases per batch: every entry's records are gathered
// (load) before any is used (use), so kBatch independent gathers are in
// flight per lane.  A chunk's eight indices are read at once; batches of
// kBatch entries follow in order.  Same entries, same order as for_chunks.
template <int LANES, class LoadF, class UseF>
__device__ __forceinline__ void for_chunks_gathered(const int* __restrict__ list, int cnt, int K, int nf, int i, int q,
                                                    int safe, LoadF&& load, UseF&& use)
{
    const int n = min(cnt, K);
    const int* L = list + (size_t)i * kChunk;
    const size_t cs = (size_t)nf * kChunk;
    for (int c = q; c * kChunk < n; c += LANES) {
        const int base = c * kChunk;
        const int4* src = reinterpret_cast<const int4*>(L + (size_t)c * cs);
        const int4 lo = src[0];
        const int4 hi = base + 4 < n ? src[1] : make_int4(safe, safe, safe, safe);
        const int jall[kChunk] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
        for (int b0 = 0; b0 < kChunk; b0 += kBatch) {
            if (base + b0 >= n) break;
            int js[kBatch];
            decltype(load(0)) rec[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                if (b0 + u >= kChunk) break;
                js[u] = base + b0 + u < n ? jall[b0 + u] : safe;
                rec[u] = load(js[u]);
            }
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                if (b0 + u >= kChunk) break;
                use(rec[u], js[u], base + b0 + u < n);
            }
        }
    }
}

struct FluidRec {
    FluidPos P;
    float4 V;
};
struct DummyRec {
    DummyPos D;
    float4 V;
};

// Sum over the LANES lanes of a particle (a fixed butterfly order).
template <int LANES, class T>
__device__ __forceinline__ T lane_group_sum(T v)
{
#pragma unroll
    for (int o = LANES / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// base[j] addressed as a uniform base plus a 32-bit byte offset, so that a
// gather costs one address instruction (records below 2^32 bytes).
template <class T>
__device__ __forceinline__ T ld_at(const T* __restrict__ base, int j)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (unsigned)j * (unsigned)sizeof(T));
}
// gather record j from its two halves (the second half's base is uniform too)
__device__ __forceinline__ GRec ld_grec(const float4* __restrict__ base, int half, int j)
{
    const float4 a = ld_at(base, j), b = ld_at(base + half, j);
    return GRec{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}

// Per-particle values of a pair pass.
struct PairSelf {
    double x, y, z;
    float vx, vy, vz, w;  // w: phase bits
    float rho, p, inv_rho;
    Phase ph;
    GradK gff, gfd;  // uniform scenes: kernel constants of fluid and dummy pairs
    float ox, oy, oz;     // gather records: offset from the stencil centre's origin
    float csx, csy, csz;  // cell size
};

template <bool UNIF>
__device__ __forceinline__ PairSelf pair_self(const SphArgs& a, int i)
{
    const FluidPos Pi = a.P[i];
    const float4 Vi = a.V[i];
    PairSelf s;
    s.x = Pi.x; s.y = Pi.y; s.z = Pi.z;
    s.vx = Vi.x; s.vy = Vi.y; s.vz = Vi.z; s.w = Vi.w;
    s.ph = UNIF ? a.phf : a.phases[phase_bits(Vi.w)];
    s.rho = Pi.rho; s.p = Pi.p;
    s.inv_rho = 1.0f / s.rho;
    if (UNIF) {
        s.gff = grad_k(0.5f * (s.ph.hf + a.phf.hf));
        s.gfd = grad_k(0.5f * (s.ph.hf + a.phd.hf));
    }
    if (a.GRin) {
        double x0, y0, z0;
        cell_origin(a.grid, a.pcen[i], x0, y0, z0);
        s.ox = (float)(s.x - x0); s.oy = (float)(s.y - y0); s.oz = (float)(s.z - z0);
        s.csx = (float)a.grid.cx; s.csy = (float)a.grid.cy; s.csz = (float)a.grid.cz;
    }
    return s;
}

// momentum_ff on a gather record; code holds the neighbour's cell relative
// to the stencil centre, 2 bits (offset + 1) per axis.
__device__ __forceinline__ void momentum_ff_g(const SphArgs& a, const PairSelf& s, const GRec& R, int code, bool valid,
                                              float& ax, float& ay, float& az)
{
    const Phase& pj = a.phf;
    const float dcx = (float)(code & 3) - 1.0f, dcy = (float)((code >> 2) & 3) - 1.0f, dcz = (float)(code >> 4) - 1.0f;
    const float rx = (s.ox - R.ox) - dcx * s.csx, ry = (s.oy - R.oy) - dcy * s.csy, rz = (s.oz - R.oz) - dcz * s.csz;
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float inv_r = (r2 < 1.e-20f) ? 0.0f : rsq(r2);
    const float r = r2 * inv_r;
    const float g = grad_factor(r, s.gff);
    const float f = valid ? g : 0.0f;
    const float dvx = s.vx - R.vx, dvy = s.vy - R.vy, dvz = s.vz - R.vz;
    const float dU = -(dvx * rx + dvy * ry + dvz * rz) * inv_r;
    const float cbar = 0.5f * (s.ph.c + pj.c);
    const float beta = fminf(3.0f * fmaxf(dU, 0.0f), cbar);
    const float Ps = 0.5f * (s.p + R.p) + 0.25f * beta * (s.rho + R.rho) * dU;
    const float fm = f * (pj.m * s.inv_rho) * rcp(R.rho);
    const float kp = -2.0f * Ps * fm;
    const float kv = (s.ph.nu + pj.nu) * r * fm;
    ax += kp * rx + kv * dvx;
    ay += kp * ry + kv * dvy;
    az += kp * rz + kv * dvz;
}

// continuity_ff on a gather record.
__device__ __forceinline__ float continuity_ff_g(const SphArgs& a, const PairSelf& s, const GRec& R, int code,
                                                 bool valid)
{
    const Phase& pj = a.phf;
    const float dcx = (float)(code & 3) - 1.0f, dcy = (float)((code >> 2) & 3) - 1.0f, dcz = (float)(code >> 4) - 1.0f;
    const float rx = (s.ox - R.ox) - dcx * s.csx, ry = (s.oy - R.oy) - dcy * s.csy, rz = (s.oz - R.oz) - dcz * s.csz;
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float inv_r = (r2 < 1.e-20f) ? 0.0f : rsq(r2);
    const float r = r2 * inv_r;
    const float g = grad_factor(r, s.gff);
    const float f = valid ? g : 0.0f;
    const float cbar = 0.5f * (s.ph.c + pj.c);
    const float dvr = (s.vx - R.vx) * rx + (s.vy - R.vy) * ry + (s.vz - R.vz) * rz;
    const float q = (s.p - R.p) * r * rcp((s.rho + R.rho) * (0.5f * cbar));
    return (0.5f * pj.m) * f * rcp(R.rho) * (dvr + q);
}

// One fluid neighbour's term of updateWCSPHAccelerationKernel
// (WCSPHIntegrationKernel.cu:80-110).
template <bool UNIF>
__device__ __forceinline__ void momentum_ff(const SphArgs& a, const PairSelf& s, const FluidRec& R, bool valid,
                                            float& ax, float& ay, float& az)
{
    const FluidPos& Pj = R.P;
    const float4& Vj = R.V;
    const Phase pj = UNIF ? a.phf : a.phases[phase_bits(Vj.w)];
    const float rx = (float)(s.x - Pj.x), ry = (float)(s.y - Pj.y), rz = (float)(s.z - Pj.z);
    const float r2 = rx * rx + ry * ry + rz * rz;
    // e = -r_ij/|r_ij| (zero below |r|^2 = 1e-20, myVec.h:112-128)
    const float inv_r = (r2 < 1.e-20f) ? 0.0f : rsq(r2);
    const float r = r2 * inv_r;
    const float g = UNIF ? grad_factor(r, s.gff) : grad_factor(r, 0.5f * (s.ph.hf + pj.hf));
    const float f = valid ? g : 0.0f;
    const float dvx = s.vx - Vj.x, dvy = s.vy - Vj.y, dvz = s.vz - Vj.z;
    // U_L - U_R with U_L = v_i . e and U_R = v_j . e
    const float dU = -(dvx * rx + dvy * ry + dvz * rz) * inv_r;
    const float cbar = 0.5f * (s.ph.c + pj.c);
    const float beta = fminf(3.0f * fmaxf(dU, 0.0f), cbar);
    // P* = (P_L + P_R)/2 + beta rho_bar (U_L - U_R)/2, rho_bar = (rho_i + rho_j)/2
    const float Ps = 0.5f * (s.p + Pj.p) + 0.25f * beta * (s.rho + Pj.rho) * dU;
    // f m_j / (rho_i rho_j)
    const float fm = f * (pj.m * s.inv_rho) * rcp(Pj.rho);
    const float kp = -2.0f * Ps * fm;
    // 2 nu_bar (r_hat . gradW) m_j / (rho_i rho_j), r_hat . gradW = f r
    const float kv = (s.ph.nu + pj.nu) * r * fm;
    ax += kp * rx + kv * dvx;
    ay += kp * ry + kv * dvy;
    az += kp * rz + kv * dvz;
}

// One dummy neighbour's term (WCSPHIntegrationKernel.cu:112-146).
template <bool UNIF>
__device__ __forceinline__ void momentum_fd(const SphArgs& a, const PairSelf& s, const DummyRec& R, int j, bool valid,
                                            float& ax, float& ay, float& az)
{
    const DummyPos& Dj = R.D;
    const float4& DVj = R.V;
    const Phase pj = UNIF ? a.phd : a.phases[a.dphase[j]];
    const float rx = (float)(s.x - Dj.x), ry = (float)(s.y - Dj.y), rz = (float)(s.z - Dj.z);
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float inv_r = (r2 < 1.e-20f) ? 0.0f : rsq(r2);
    const float r = r2 * inv_r;
    const float g = UNIF ? grad_factor(r, s.gfd) : grad_factor(r, 0.5f * (s.ph.hf + pj.hf));
    const float f = valid ? g : 0.0f;
    const float nx = Dj.nx, ny = Dj.ny, nz = DVj.w;
    const float UL = -(nx * s.vx + ny * s.vy + nz * s.vz);
    const float UR = -UL - 2.0f * (nx * DVj.x + ny * DVj.y + nz * DVj.z);
    const float grj = fmaxf(-(a.gx * rx + a.gy * ry + a.gz * rz), 0.0f);
    const float PR = s.p + s.rho * grj;
    const float rhoj = dummy_rho(a, pj, PR);
    const float cbar = 0.5f * (s.ph.c + pj.c), rbar = 0.5f * (s.rho + rhoj);
    const float beta = fminf(3.0f * fmaxf(UL - UR, 0.0f), cbar);
    const float Ps = 0.5f * (s.p + PR) + 0.5f * beta * rbar * (UL - UR);
    const float mr = pj.m * s.inv_rho * rcp(rhoj);
    const float kp = -2.0f * mr * Ps * f;
    const float kv = 2.0f * (0.5f * (s.ph.nu + pj.nu)) * (f * r) * mr;
    ax += kp * rx + kv * (s.vx - DVj.x);
    ay += kp * ry + kv * (s.vy - DVj.y);
    az += kp * rz + kv * (s.vz - DVj.z);
}

// Lane-group sum, coupled balls, gravity, half kick (:150-159) and, in the
// first half, the drift (:161-170).
template <bool FIRST, int LANES>
__device__ __forceinline__ void momentum_finish(const SphArgs& a, const PairSelf& s, int i, int q, float ax, float ay,
                                                float az)
{
    if (LANES > 1) {
        ax = lane_group_sum<LANES>(ax);
        ay = lane_group_sum<LANES>(ay);
        az = lane_group_sum<LANES>(az);
    }
    if (a.ns > 0) {
        float cx = 0.f, cy = 0.f, cz = 0.f;
        for_chunks<LANES>(a.nbr_fs, a.cnt_fs[i], a.K_fs, a.nf, i, q, 0, [&](int j, bool valid) {
            float3 ap, av, lv;
            fs_momentum(a, s.ph, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.rho, s.p, a.SP[j], a.SV[j], a.SW[j], ap, av, lv);
            cx += valid ? ap.x + av.x : 0.0f;
            cy += valid ? ap.y + av.y : 0.0f;
            cz += valid ? ap.z + av.z : 0.0f;
        });
        if (LANES > 1) {
            cx = lane_group_sum<LANES>(cx);
            cy = lane_group_sum<LANES>(cy);
            cz = lane_group_sum<LANES>(cz);
        }
        ax += cx; ay += cy; az += cz;
        if (a.acs && q == 0) {
            a.acs[3 * (size_t)i] = cx;
            a.acs[3 * (size_t)i + 1] = cy;
            a.acs[3 * (size_t)i + 2] = cz;
        }
    }
    if (q != 0) return;
    ax += a.gx; ay += a.gy; az += a.gz;
    const double hdt = 0.5 * a.dt;
    const double nvx = a.vx[i] + hdt * (double)ax;
    const double nvy = a.vy[i] + hdt * (double)ay;
    const double nvz = a.vz[i] + hdt * (double)az;
    a.ovx[i] = nvx; a.ovy[i] = nvy; a.ovz[i] = nvz;
    const float4 nV = make_float4((float)nvx, (float)nvy, (float)nvz, s.w);
    a.oV[i] = nV;
    if (FIRST) {
        FluidPos o;
        o.x = s.x + a.dt * nvx;
        o.y = s.y + a.dt * nvy;
        o.z = s.z + a.dt * nvz;
        o.rho = s.rho;
        o.p = s.p;
        a.oP[i] = o;
        if (a.GRout) st_grec(a.GRout, a.nf, i, grec(a.grid, a.pcell[i], o, nV));
    }
}

// Entry-split pair loops.  Lane q of a particle's LANES lanes takes entries
// q*E .. q*E+E-1 of every list chunk (E = 8 / LANES): the lanes of a group
// read each chunk together and finish within one entry per chunk of each
// other.  Against the chunk split (lane q takes whole chunks q, q + LANES,
// ...) this balances 8 lanes over lists of about ten chunks; the first
// kPreEntries entries of a lane are loaded with the particle's own record,
// ahead of the gathers they index.
template <int E>
__device__ __forceinline__ void load_entries(const int* __restrict__ p, int* v)
{
    if (E == 4) {
        const int4 x = *reinterpret_cast<const int4*>(p);
        v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
    } else if (E == 2) {
        const int2 x = *reinterpret_cast<const int2*>(p);
        v[0] = x.x; v[1] = x.y;
    } else {
        v[0] = p[0];
    }
}

template <int LANES, class LoadF, class UseF>
__device__ __forceinline__ void for_entries(const int* __restrict__ list, int n, int nf, int i, int q, int safe,
                                            LoadF&& load, UseF&& use)
{
    static_assert(LANES == 2 || LANES == 4 || LANES == 8, "entry-split pair loops take 2, 4 or 8 lanes");
    constexpr int E = kChunk / LANES;
    constexpr int CB = E >= kBatch ? 1 : kBatch / E;  // chunks per batch
    const int* L = list + (size_t)i * kChunk + q * E;
    const size_t cs = (size_t)nf * kChunk;
    const int nch = (n + kChunk - 1) / kChunk;
    for (int c0 = 0; c0 < nch; c0 += CB) {
        int js[CB * E];
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const int c = c0 + cb;
            int v[E];
            if (c < nch) {
                load_entries<E>(L + (size_t)c * cs, v);
            } else {
#pragma unroll
                for (int e = 0; e < E; e++) v[e] = safe;
            }
#pragma unroll
            for (int e = 0; e < E; e++) js[cb * E + e] = c * kChunk + q * E + e < n ? v[e] : safe;
        }
        decltype(load(0)) rec[CB * E];
#pragma unroll
        for (int k = 0; k < CB * E; k++) rec[k] = load(js[k]);
#pragma unroll
        for (int k = 0; k < CB * E; k++) use(rec[k], js[k], (c0 + k / E) * kChunk + q * E + k % E < n);
    }
}

// The first kPreEntries list entries of a lane (E per chunk) within the
// list capacity, loaded in one go whatever the particle's count.
constexpr int kPreEntries = 16;
template <int LANES>
__device__ __forceinline__ void preload_entries(const int* __restrict__ list, int K, int nf, int i, int q,
                                                int (&pre)[kPreEntries])
{
    constexpr int E = kChunk / LANES;
    constexpr int PC = kPreEntries / E;
    const int kch = K / kChunk;  // chunks of capacity (uniform)
    const int* L = list + (size_t)i * kChunk + q * E;
    const size_t cs = (size_t)nf * kChunk;
#pragma unroll
    for (int c = 0; c < PC; c++) {
        if (c < kch) {
            load_entries<E>(L + (size_t)c * cs, &pre[c * E]);
        } else {
#pragma unroll
            for (int e = 0; e < E; e++) pre[c * E + e] = 0;
        }
    }
}

// for_entries with the first kPreEntries entries of the lane taken from
// preload_entries.  Same entries, same order.
template <int LANES, class LoadF, class UseF>
__device__ __forceinline__ void for_entries_pre(const int (&pre)[kPreEntries], const int* __restrict__ list, int n,
                                                int nf, int i, int q, int safe, LoadF&& load, UseF&& use)
{
    constexpr int E = kChunk / LANES;
    constexpr int CB = E >= kBatch ? 1 : kBatch / E;  // chunks per batch
    constexpr int PC = kPreEntries / E;               // preloaded chunks
    const int nch = (n + kChunk - 1) / kChunk;
#pragma unroll
    for (int c0 = 0; c0 < PC; c0 += CB) {
        if (c0 >= nch) break;
        int js[CB * E];
#pragma unroll
        for (int k = 0; k < CB * E; k++) {
            const int c = c0 + k / E, e = k % E;
            js[k] = c * kChunk + q * E + e < n ? pre[c * E + e] : safe;
        }
        decltype(load(0)) rec[CB * E];
#pragma unroll
        for (int k = 0; k < CB * E; k++) rec[k] = load(js[k]);
#pragma unroll
        for (int k = 0; k < CB * E; k++) use(rec[k], js[k], (c0 + k / E) * kChunk + q * E + k % E < n);
    }
    if (nch <= PC) return;
    const int* L = list + (size_t)i * kChunk + q * E;
    const size_t cs = (size_t)nf * kChunk;
    for (int c0 = PC; c0 < nch; c0 += CB) {
        int js[CB * E];
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
            const int c = c0 + cb;
            int v[E];
            load_entries<E>(L + (size_t)min(c, nch - 1) * cs, v);
#pragma unroll
            for (int e = 0; e < E; e++) js[cb * E + e] = c * kChunk + q * E + e < n ? v[e] : safe;
        }
        decltype(load(0)) rec[CB * E];
#pragma unroll
        for (int k = 0; k < CB * E; k++) rec[k] = load(js[k]);
#pragma unroll
        for (int k = 0; k < CB * E; k++) use(rec[k], js[k], (c0 + k / E) * kChunk + q * E + k % E < n);
    }
}

// updateWCSPHAccelerationKernel (WCSPHIntegrationKernel.cu:42-148) fused
// with the half kick (:150-159) and, in the first half, the drift (:161-170).
template <bool FIRST, bool UNIF, int LANES, bool GR = false>
__global__ __launch_bounds__(kThreads) void k_momentum(SphArgs a)
{
    const int g = xcd_block() * kThreads + threadIdx.x;
    const int i = g / LANES, q = g % LANES;
    if (i >= a.nf) return;  // uniform over a lane group
    const PairSelf s = pair_self<UNIF>(a, i);
    float ax = 0.f, ay = 0.f, az = 0.f;
    const auto load_f = [&](int j) { return FluidRec{ld_at(a.P, j), ld_at(a.V, j)}; };
    const auto load_d = [&](int j) { return DummyRec{ld_at(a.DP, j), ld_at(a.DV, j)}; };
    const auto use_f = [&](const FluidRec& R, int, bool valid) { momentum_ff<UNIF>(a, s, R, valid, ax, ay, az); };
    const auto use_d = [&](const DummyRec& R, int j, bool valid) { momentum_fd<UNIF>(a, s, R, j, valid, ax, ay, az); };
    if constexpr (GR) {
        static_assert(UNIF && LANES >= 2, "gather records: uniform scenes, entry-split lanes");
        int pre[kPreEntries];
        preload_entries<LANES>(a.nbr_ff, a.K_ff, a.nf, i, q, pre);
        for_entries_pre<LANES>(pre, a.nbr_ff, min(a.cnt_ff[i], a.K_ff), a.nf, i, q, i,
                               [&](int e) { return ld_grec(a.GRin, a.nf, e & kSlotMask); },
                               [&](const GRec& R, int e, bool valid) {
                                   momentum_ff_g(a, s, R, e >> kCodeShift, valid, ax, ay, az);
                               });
        if (a.nd > 0) for_entries<LANES>(a.nbr_fd, min(a.cnt_fd[i], a.K_fd), a.nf, i, q, 0, load_d, use_d);
    } else if constexpr (LANES >= 2) {
        int pre[kPreEntries];
        preload_entries<LANES>(a.nbr_ff, a.K_ff, a.nf, i, q, pre);
        for_entries_pre<LANES>(pre, a.nbr_ff, min(a.cnt_ff[i], a.K_ff), a.nf, i, q, i, load_f, use_f);
        if (a.nd > 0) for_entries<LANES>(a.nbr_fd, min(a.cnt_fd[i], a.K_fd), a.nf, i, q, 0, load_d, use_d);
    } else {
        for_chunks_gathered<LANES>(a.nbr_ff, a.cnt_ff[i], a.K_ff, a.nf, i, q, i, load_f, use_f);
        if (a.nd > 0) for_chunks_gathered<LANES>(a.nbr_fd, a.cnt_fd[i], a.K_fd, a.nf, i, q, 0, load_d, use_d);
    }
    momentum_finish<FIRST, LANES>(a, s, i, q, ax, ay, az);
}

// One fluid neighbour's term of updateWCSPHDensityChangeKernel
// (WCSPHIntegrationKernel.cu:207-231).  With gradW = f r_ij and
// e = -r_ij/r: (v_i - v*) . gradW = f [ (v_i - v_j) . r_ij / 2 + (P_L - P_R) r / (2 rho_bar c_bar) ].
template <bool UNIF>
__device__ __forceinline__ float continuity_ff(const SphArgs& a, const PairSelf& s, const FluidRec& R, bool valid)
{
    const FluidPos& Pj = R.P;
    const float4& Vj = R.V;
    const Phase pj = UNIF ? a.phf : a.phases[phase_bits(Vj.w)];
    const float rx = (float)(s.x - Pj.x), ry = (float)(s.y - Pj.y), rz = (float)(s.z - Pj.z);
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float inv_r = (r2 < 1.e-20f) ? 0.0f : rsq(r2);
    const float r = r2 * inv_r;
    const float g = UNIF ? grad_factor(r, s.gff) : grad_factor(r, 0.5f * (s.ph.hf + pj.hf));
    const float f = valid ? g : 0.0f;
    const float cbar = 0.5f * (s.ph.c + pj.c);
    const float dvr = (s.vx - Vj.x) * rx + (s.vy - Vj.y) * ry + (s.vz - Vj.z) * rz;
    // (P_L - P_R) r / (rho_bar c_bar), rho_bar = (rho_i + rho_j)/2
    const float q = (s.p - Pj.p) * r * rcp((s.rho + Pj.rho) * (0.5f * cbar));
    return (0.5f * pj.m) * f * rcp(Pj.rho) * (dvr + q);
}

// One dummy neighbour's term (WCSPHIntegrationKernel.cu:233-262).
template <bool UNIF>
__device__ __forceinline__ float continuity_fd(const SphArgs& a, const PairSelf& s, const DummyRec& R, int j, bool valid)
{
    const DummyPos& Dj = R.D;
    const float4& DVj = R.V;
    const Phase pj = UNIF ? a.phd : a.phases[a.dphase[j]];
    const float rx = (float)(s.x - Dj.x), ry = (float)(s.y - Dj.y), rz = (float)(s.z - Dj.z);
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float inv_r = (r2 < 1.e-20f) ? 0.0f : rsq(r2);
    const float r = r2 * inv_r;
    const float g = UNIF ? grad_factor(r, s.gfd) : grad_factor(r, 0.5f * (s.ph.hf + pj.hf));
    const float f = valid ? g : 0.0f;
    const float grj = fmaxf(-(a.gx * rx + a.gy * ry + a.gz * rz), 0.0f);
    const float PR = s.p + s.rho * grj;
    const float rhoj = dummy_rho(a, pj, PR);
    const float cbar = 0.5f * (s.ph.c + pj.c), rbar = 0.5f * (s.rho + rhoj);
    const float t = 0.5f * ((s.vx - DVj.x) * rx + (s.vy - DVj.y) * ry + (s.vz - DVj.z) * rz) +
                    0.5f * (s.p - PR) * r * rcp(rbar * cbar);
    return pj.m * rcp(rhoj) * f * t;
}

// Coupled balls, lane-group sum and the EOS update (:266-282).
template <int LANES>
__device__ __forceinline__ void continuity_finish(const SphArgs& a, const PairSelf& s, int i, int q, float d)
{
    if (a.ns > 0) {
        for_chunks<LANES>(a.nbr_fs, a.cnt_fs[i], a.K_fs, a.nf, i, q, 0, [&](int j, bool valid) {
            const float t = fs_density(a, s.ph, s.x, s.y, s.z, s.vx, s.vy, s.vz, s.rho, s.p, a.SP[j], a.SV[j], a.SW[j]);
            d += valid ? t : 0.0f;
        });
    }
    if (LANES > 1) d = lane_group_sum<LANES>(d);
    if (q != 0) return;
    const float nd = a.drho[i] + (float)(a.dt * (double)(2.0f * s.rho * d));
    a.odrho[i] = nd;
    FluidPos o;
    o.x = s.x; o.y = s.y; o.z = s.z;
    eos_eval(a.eos, a.gamma, s.ph, nd, o.rho, o.p);
    a.oP[i] = o;
    if (a.GRout) st_grec(a.GRout, a.nf, i, grec(a.grid, a.pcell[i], o, make_float4(s.vx, s.vy, s.vz, s.w)));
}

// updateWCSPHDensityChangeKernel (WCSPHIntegrationKernel.cu:172-264) fused
// with the EOS update (:266-282).
template <bool UNIF, int LANES, bool GR = false>
__global__ __launch_bounds__(kThreads) void k_continuity(SphArgs a)
{
    const int g = xcd_block() * kThreads + threadIdx.x;
    const int i = g / LANES, q = g % LANES;
    if (i >= a.nf) return;  // uniform over a lane group
    const PairSelf s = pair_self<UNIF>(a, i);
    float d = 0.f;
    const auto load_f = [&](int j) { return FluidRec{ld_at(a.P, j), ld_at(a.V, j)}; };
    const auto load_d = [&](int j) { return DummyRec{ld_at(a.DP, j), ld_at(a.DV, j)}; };
    const auto use_f = [&](const FluidRec& R, int, bool valid) { d += continuity_ff<UNIF>(a, s, R, valid); };
    const auto use_d = [&](const DummyRec& R, int j, bool valid) { d += continuity_fd<UNIF>(a, s, R, j, valid); };
    if constexpr (GR) {
        static_assert(UNIF && LANES >= 2, "gather records: uniform scenes, entry-split lanes");
        int pre[kPreEntries];
        preload_entries<LANES>(a.nbr_ff, a.K_ff, a.nf, i, q, pre);
        for_entries_pre<LANES>(pre, a.nbr_ff, min(a.cnt_ff[i], a.K_ff), a.nf, i, q, i,
                               [&](int e) { return ld_grec(a.GRin, a.nf, e & kSlotMask); },
                               [&](const GRec& R, int e, bool valid) { d += continuity_ff_g(a, s, R, e >> kCodeShift, valid); });
        if (a.nd > 0) for_entries<LANES>(a.nbr_fd, min(a.cnt_fd[i], a.K_fd), a.nf, i, q, 0, load_d, use_d);
    } else if constexpr (LANES >= 2) {
        int pre[kPreEntries];
        preload_entries<LANES>(a.nbr_ff, a.K_ff, a.nf, i, q, pre);
        for_entries_pre<LANES>(pre, a.nbr_ff, min(a.cnt_ff[i], a.K_ff), a.nf, i, q, i, load_f, use_f);
        if (a.nd > 0) for_entries<LANES>(a.nbr_fd, min(a.cnt_fd[i], a.K_fd), a.nf, i, q, 0, load_d, use_d);
    } else {
        for_chunks_gathered<LANES>(a.nbr_ff, a.cnt_ff[i], a.K_ff, a.nf, i, q, i, load_f, use_f);
        if (a.nd > 0) for_chunks_gathered<LANES>(a.nbr_fd, a.cnt_fd[i], a.K_fd, a.nf, i, q, 0, load_d, use_d);
    }
    continuity_finish<LANES>(a, s, i, q, d);
}

// ----------------------------------------------------------- coupled balls
// Cell hash of each DEM ball (DEM slot order) on the fluid grid.
__global__ void k_solid_hash(int ns, GridGeom g, const double* __restrict__ px, const double* __restrict__ py,
                             const double* __restrict__ pz, uint32_t* __restrict__ key)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    key[s] = (uint32_t)cell_hash(g, px[s], py[s], pz[s]);
}

struct SolidGatherArgs {
    int ns;
    long long numGrids;
    const uint32_t* skey;  // sorted keys (null: refresh kinematics only)
    const int* s2d;        // solid slot -> DEM slot
    const double *px, *py, *pz, *vx, *vy, *vz, *wx, *wy, *wz, *r;
    SolidPos* SP;
    float4 *SV, *SW;
    int2* cells;  // null when refreshing
};

// Ball records in fluid-grid order (the pressure P is left as it is) and,
// after a sort, the balls' cell ranges on the fluid grid.
__global__ void k_solid_gather(SolidGatherArgs a)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.ns) return;
    const int s = a.s2d[k];
    SolidPos& P = a.SP[k];
    P.x = a.px[s]; P.y = a.py[s]; P.z = a.pz[s];
    P.r = (float)a.r[s];
    a.SV[k] = make_float4((float)a.vx[s], (float)a.vy[s], (float)a.vz[s], 0.f);
    a.SW[k] = make_float4((float)a.wx[s], (float)a.wy[s], (float)a.wz[s], 0.f);
    if (a.cells) {
        const uint32_t h = a.skey[k];
        if ((long long)h < a.numGrids) {
            if (k == 0 || a.skey[k - 1] != h) a.cells[h].x = k;
            if (k == a.ns - 1 || a.skey[k + 1] != h) a.cells[h].y = k + 1;
        }
    }
}

// One wavefront per coupled ball: lane l of the wave owns ball slot
// blockIdx.x * kBallsPerBlock + wave, and the 64 lanes split that ball's
// fluid candidates or list entries.
constexpr int kBallsPerBlock = 4;
__device__ __forceinline__ int wave_lane() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_ball() { return blockIdx.x * kBallsPerBlock + (threadIdx.x >> 6); }

template <class T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

struct RevArgs {
    int ns;
    GridGeom grid;
    const int2* cells;  // fluid cell ranges (fluid slots sorted by cell)
    const FluidPos* P;
    const float4* V;
    const Phase* phases;
    const SolidPos* SP;
    int* rcnt;
    const int* rstart;
    int* rfluid;
    int rcap;  // entries rfluid holds; a list overflow (reported by the fluid side) may exceed it
    double hf; // the fluid's smoothing length when it has one phase
    // FP32 prefilter of the fluid positions (null: every test in FP64), as
    // in the list walks: decisions outside the band are exact
    const float4* Pf;
    float band;
    double R;
};

// Reverse index of the fluid -> ball lists: for ball j, the fluid slots i
// whose support holds the ball (the predicate of k_build_fs_lists with the
// same operands: within_cut(2 h_i, x_i - x_j)).  Cell adjacency is
// symmetric, so the walk over the fluid cells of the ball's 27-cell stencil
// finds every such i; rows are walked in cell order, so the slots come out
// ascending.  COUNT: rcnt[j] = their number; else they fill
// rfluid[rstart[j] ...] through a wave ballot (lane order = slot order).
template <bool COUNT, bool UNIF>
__global__ __launch_bounds__(256) void k_fs_rev_walk(RevArgs a)
{
    const int lane = wave_lane();
    const int j = wave_ball();
    if (j >= a.ns) return;  // uniform over the wave
    const GridGeom& g = a.grid;
    const double bx = a.SP[j].x, by = a.SP[j].y, bz = a.SP[j].z;
    const Stencil st = make_stencil(g, bx, by, bz);
    int off = COUNT ? 0 : a.rstart[j];
    const float4 bf = pos_f32(g, a.R, bx, by, bz);
    const float c2 = (float)(4.0 * a.hf * a.hf);
    const float band2 = bf.w != 0.f ? INFINITY : a.band * c2;
    for (int k = 0; k < 9; k++) {
        const int dz = k / 3 - 1, dy = k % 3 - 1;
        if (dz < st.loz || dz > st.hiz || dy < st.loy || dy > st.hiy) continue;
        const int row = (st.cz + dz) * g.ny * g.nx + (st.cy + dy) * g.nx + st.cx;
        int s = -1, e = -1;
        for (int dx = st.lox; dx <= st.hix; dx++) {
            const int2 c = a.cells[row + dx];
            if (c.x != -1) {
                if (s < 0) s = c.x;
                e = c.y;
            }
        }
        if (s < 0) continue;
        for (int base = s; base < e; base += 64) {
            const int i = base + lane;
            bool hit = false;
            if (UNIF && a.Pf) {
                if (i < e) {
                    const float4 q = a.Pf[i];
                    const float fx = q.x - bf.x, fy = q.y - bf.y, fz = q.z - bf.z;
                    const float d2 = fx * fx + fy * fy + fz * fz;
                    hit = d2 < c2;
                    if ((q.w != 0.f) | (fabsf(d2 - c2) <= band2)) {
                        const FluidPos p = a.P[i];
                        hit = within_cut(2.0 * a.hf, p.x - bx, p.y - by, p.z - bz);
                    }
                }
            } else if (i < e) {
                const FluidPos p = a.P[i];
                const double h = UNIF ? a.hf : a.phases[phase_bits(a.V[i].w)].h;
                hit = within_cut(2.0 * h, p.x - bx, p.y - by, p.z - bz);
            }
            const unsigned long long m = __ballot(hit);
            if (!COUNT && hit) {
                const int pos = off + __popcll(m & ((1ull << lane) - 1ull));
                if (pos < a.rcap) a.rfluid[pos] = i;
            }
            off += __popcll(m);
        }
    }
    if (COUNT && lane == 0) a.rcnt[j] = off;
}

// Wendland C2 value (WCSPHIntegrationKernel.cuh:5-13), FP32.
__device__ __forceinline__ float wendland(float r, float h)
{
    const float q = r / h;
    if (q >= 2.0f) return 0.0f;
    const float t = 1.0f - 0.5f * q;
    return (float)(21.0 / (16.0 * M_PI)) / (h * h * h) * (t * t * t * t) * (2.0f * q + 1.0f);
}

// P_j = sum V_i p_i W_ij / sum V_i W_ij over the ball's fluid neighbours
// (Shepard interpolation of the fluid pressure at the ball centre), one
// wavefront per ball: lane l sums entries l, l + 64, ... of the ball's
// reverse list, then the wave reduces (a fixed order, so the result is
// deterministic).
__global__ __launch_bounds__(256) void k_solid_pressure(int ns, const int* __restrict__ rstart,
                                                        const int* __restrict__ rfluid, const FluidPos* __restrict__ P,
                                                        const float4* __restrict__ V, const Phase* __restrict__ phases,
                                                        SolidPos* __restrict__ SP, int rcap)
{
    const int lane = wave_lane();
    const int j = wave_ball();
    if (j >= ns) return;
    const double sx = SP[j].x, sy = SP[j].y, sz = SP[j].z;
    float num = 0.f, den = 0.f;
    const int qe = min(rstart[j + 1], rcap);
    for (int q = rstart[j] + lane; q < qe; q += 64) {
        const int i = rfluid[q];
        const FluidPos f = P[i];
        const Phase& ph = phases[phase_bits(V[i].w)];
        const float rx = (float)(f.x - sx), ry = (float)(f.y - sy), rz = (float)(f.z - sz);
        const float w = wendland(sqrtf(rx * rx + ry * ry + rz * rz), ph.hf) * (ph.m / f.rho);
        num += w * f.p;
        den += w;
    }
    num = wave_sum(num);
    den = wave_sum(den);
    if (lane == 0) SP[j].P = den > 0.f ? num / den : 0.f;
}

struct CoupleForceArgs {
    SphArgs a;  // fluid state and ball records, as the fluid pass sees them
    const int *rstart, *rfluid, *s2d;
    int rcap;                             // entries rfluid holds
    double *fx, *fy, *fz, *tx, *ty, *tz;  // DEM force / torque accumulators (DEM slot order)
    double* Fc;                           // per ball slot: coupling force and torque (6 doubles)
};

// The reaction of every fluid-ball pair of the fluid pass on the ball:
// -m_i (ap + av) and (r_j n) x (-m_i av), summed in FP64 by one wavefront
// per ball (lane stripes of the ball's reverse list, then a fixed-order
// wave reduction) and added to the DEM accumulators.
__global__ __launch_bounds__(256) void k_couple_forces(CoupleForceArgs c)
{
    const SphArgs& a = c.a;
    const int lane = wave_lane();
    const int j = wave_ball();
    if (j >= a.ns) return;
    const SolidPos S = a.SP[j];
    const float4 SV = a.SV[j], SW = a.SW[j];
    double Fx = 0, Fy = 0, Fz = 0, Tx = 0, Ty = 0, Tz = 0;
    const int qe = min(c.rstart[j + 1], c.rcap);
    for (int q = c.rstart[j] + lane; q < qe; q += 64) {
        const int i = c.rfluid[q];
        const FluidPos Pi = a.P[i];
        const float4 Vi = a.V[i];
        const Phase pi = a.phases[phase_bits(Vi.w)];
        float3 ap, av, lv;
        fs_momentum(a, pi, Pi.x, Pi.y, Pi.z, Vi.x, Vi.y, Vi.z, Pi.rho, Pi.p, S, SV, SW, ap, av, lv);
        const double m = (double)pi.m;
        Fx -= m * (double)(ap.x + av.x);
        Fy -= m * (double)(ap.y + av.y);
        Fz -= m * (double)(ap.z + av.z);
        const double vx = -m * (double)av.x, vy = -m * (double)av.y, vz = -m * (double)av.z;
        Tx += (double)lv.y * vz - (double)lv.z * vy;
        Ty += (double)lv.z * vx - (double)lv.x * vz;
        Tz += (double)lv.x * vy - (double)lv.y * vx;
    }
    Fx = wave_sum(Fx); Fy = wave_sum(Fy); Fz = wave_sum(Fz);
    Tx = wave_sum(Tx); Ty = wave_sum(Ty); Tz = wave_sum(Tz);
    if (lane != 0) return;
    const int s = c.s2d[j];
    c.fx[s] += Fx; c.fy[s] += Fy; c.fz[s] += Fz;
    c.tx[s] += Tx; c.ty[s] += Ty; c.tz[s] += Tz;
    double* o = c.Fc + 6 * (size_t)j;
    o[0] = Fx; o[1] = Fy; o[2] = Fz; o[3] = Tx; o[4] = Ty; o[5] = Tz;
}

// ----------------------------------------------------------- dummy normals
// calDummyParticleNormalKernel (WCSPHIntegrationKernel.cu:4-40) over the
// dummy-dummy pairs of the 27-cell walk, FP64 (computed once).
struct NormalArgs {
    int nd;
    GridGeom grid;
    const Phase* phases;
    DummyPos* DP;
    float4* DV;
    const uint8_t* dphase;
    const int2* dcells;
};

__device__ __forceinline__ void grad_w64(double rx, double ry, double rz, double h, double& gx, double& gy, double& gz)
{
    gx = gy = gz = 0.0;
    const double r = sqrt(rx * rx + ry * ry + rz * rz);
    if (r < 1.e-20 || r >= 2.0 * h) return;
    const double q = r / h;
    const double sigma = 21.0 / (16.0 * M_PI * h * h * h);
    const double t = 1.0 - 0.5 * q;
    const double dWdq = (-2.0 * t * t * t * (1.0 + 2.0 * q) + 2.0 * t * t * t * t);
    const double f = sigma * dWdq / h / r;
    gx = f * rx; gy = f * ry; gz = f * rz;
}

__global__ __launch_bounds__(kThreads) void k_dummy_normals(NormalArgs a)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= a.nd) return;
    const double x = a.DP[i].x, y = a.DP[i].y, z = a.DP[i].z;
    const Phase pi = a.phases[a.dphase[i]];
    const Stencil st = make_stencil(a.grid, x, y, z);
    double tx = 0, ty = 0, tz = 0;
    for (int zz = st.loz; zz <= st.hiz; zz++)
        for (int yy = st.loy; yy <= st.hiy; yy++)
            for (int xx = st.lox; xx <= st.hix; xx++) {
                const int hb = (st.cz + zz) * a.grid.ny * a.grid.nx + (st.cy + yy) * a.grid.nx + st.cx + xx;
                const int2 rg = a.dcells[hb];
                const int b0 = rg.x, b1 = rg.y;
                if (b0 == -1) continue;
                for (int j = b0; j < b1; j++) {
                    if (j == i) continue;
                    const Phase pj = a.phases[a.dphase[j]];
                    const double cut = 2.0 * fmax(pi.h, pj.h);
                    const double rx = x - a.DP[j].x, ry = y - a.DP[j].y, rz = z - a.DP[j].z;
                    if (!within_cut(cut, rx, ry, rz)) continue;
                    double gx, gy, gz;
                    grad_w64(rx, ry, rz, 0.5 * (pi.h + pj.h), gx, gy, gz);
                    // m_j / rho_j with the dummy density equal to rho0 (WCSPHSolver.h:533-541)
                    const double w = (double)pj.m / (double)pj.rho0;
                    tx -= w * gx; ty -= w * gy; tz -= w * gz;
                }
            }
    const double l2 = tx * tx + ty * ty + tz * tz;
    float nx = 0.f, ny = 0.f, nz = 0.f;
    if (l2 > 1.e-10) {
        // myVec.h:46-50: division through the reciprocal of the length
        const double inv = 1.0 / sqrt(l2);
        nx = (float)(tx * inv); ny = (float)(ty * inv); nz = (float)(tz * inv);
    }
    // Only this thread touches slot i's normal words; other threads read
    // the position words of the same records.
    a.DP[i].nx = nx;
    a.DP[i].ny = ny;
    a.DV[i].w = nz;
}

// Dummies: hash by insertion index, then gather into sorted order once.
struct DummyReorderArgs {
    int n;
    long long numGrids;
    const uint32_t* skey;
    const int* perm;
    const double *px, *py, *pz, *vx, *vy, *vz;
    const uint8_t* phase;
    DummyPos* DP;
    float4* DV;
    uint8_t* ophase;
    int* orig;
    int2* cells;
    float4* DPf;  // FP32 positions for the list prefilter (null: none)
    GridGeom grid;
    double R;
};

__global__ __launch_bounds__(kThreads) void k_dummy_reorder(DummyReorderArgs a)
{
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= a.n) return;
    const int o = a.perm[k];
    a.orig[k] = o;
    DummyPos d;
    d.x = a.px[o]; d.y = a.py[o]; d.z = a.pz[o];
    d.nx = 0.f; d.ny = 0.f;
    a.DP[k] = d;
    if (a.DPf) a.DPf[k] = pos_f32(a.grid, a.R, d.x, d.y, d.z);
    a.DV[k] = make_float4((float)a.vx[o], (float)a.vy[o], (float)a.vz[o], 0.f);
    a.ophase[k] = a.phase[o];
    const uint32_t h = a.skey[k];
    if ((long long)h < a.numGrids) {
        if (k == 0 || a.skey[k - 1] != h) a.cells[h].x = k;
        if (k == a.n - 1 || a.skey[k + 1] != h) a.cells[h].y = k + 1;
    }
}

// Per cell: 1 when any cell of its 27-cell stencil (clamped at the grid
// faces, as the list walk clamps it) holds a dummy.
__global__ void k_dummy_near(GridGeom g, long long numGrids, const int2* __restrict__ dcells, uint8_t* __restrict__ near)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numGrids) return;
    const int cx = (int)(c % g.nx), cy = (int)((c / g.nx) % g.ny), cz = (int)(c / ((long long)g.nx * g.ny));
    uint8_t any = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int x = cx + dx, y = cy + dy, z = cz + dz;
                if (x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz) continue;
                if (dcells[((long long)z * g.ny + y) * g.nx + x].x != -1) any = 1;
            }
    near[c] = any;
}

__global__ void k_dummy_hash(int n, GridGeom g, const double* __restrict__ px, const double* __restrict__ py,
                             const double* __restrict__ pz, uint32_t* __restrict__ key)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    key[s] = (uint32_t)cell_hash(g, px[s], py[s], pz[s]);
}

__global__ void k_iota(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// Fluid records from insertion-ordered host uploads (slot == index).
__global__ void k_fluid_init(int n, const double* __restrict__ pos, const double* __restrict__ vel,
                             const uint8_t* __restrict__ phase, FluidPos* P, float4* V, double* vx, double* vy,
                             double* vz, float* drho)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    FluidPos p;
    p.x = pos[3 * (size_t)i]; p.y = pos[3 * (size_t)i + 1]; p.z = pos[3 * (size_t)i + 2];
    p.rho = 0.f; p.p = 0.f;
    P[i] = p;
    const double ux = vel[3 * (size_t)i], uy = vel[3 * (size_t)i + 1], uz = vel[3 * (size_t)i + 2];
    vx[i] = ux; vy[i] = uy; vz[i] = uz;
    V[i] = make_float4((float)ux, (float)uy, (float)uz, __uint_as_float((unsigned)phase[i]));
    drho[i] = 0.f;
}

// rho, p of every slot from its drho (after initialisation or a host write).
__global__ void k_eos_all(int n, const Phase* phases, const float4* V, const float* drho, FluidPos* P, int eos,
                          float gamma)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float rho, p;
    eos_eval(eos, gamma, phases[phase_bits(V[i].w)], drho[i], rho, p);
    P[i].rho = rho;
    P[i].p = p;
}

}  // namespace

// ============================================================= host scene
struct SphHost {
    std::vector<double> pos, vel, c, spacing, rho0, nu;
    void add(size_t n, const double* p, const double* v, const double* cc, const double* dx, const double* r0,
             const double* vis)
    {
        for (size_t i = 0; i < n; i++) {
            if (r0[i] <= 0.) continue;  // WCSPHSolver.h:525-526
            if (cc[i] <= 0.) continue;
            for (int d = 0; d < 3; d++) {
                pos.push_back(p[3 * i + d]);
                vel.push_back(v ? v[3 * i + d] : 0.0);
            }
            c.push_back(cc[i]);
            spacing.push_back(dx[i]);
            rho0.push_back(r0[i]);
            nu.push_back(vis[i]);
        }
    }
    size_t size() const { return c.size(); }
};

}  // namespace demsph

using namespace demsph;

namespace demsph {
// What the fluid passes see of the coupled balls during coupled stepping.
struct SolidView {
    int ns = 0;
    const SolidPos* SP = nullptr;
    const float4 *SV = nullptr, *SW = nullptr;
    const int *nbr_fs = nullptr, *cnt_fs = nullptr;
    int K_fs = 0;
    float* acs = nullptr;
};
}  // namespace demsph

struct demsph_sph {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool initialized = false;
    SphHost hf, hd;
    int nf = 0, nd = 0;
    GridGeom grid{};
    long long numGrids = 0;
    int eos = 0;
    double gamma = 7.0;
    double g[3] = {0, 0, 0};
    int K_ff = 128, K_fd = 64;
    std::vector<Phase> phases;
    std::vector<uint8_t> fphase_h, dphase_h;
    Phase* d_phases = nullptr;
    bool uniform = false;  // one fluid phase and at most one dummy phase
    bool fluid_one_phase = false;
    bool fluid_uniform() const { return fluid_one_phase; }
    SolidView solids;      // coupled balls (ns == 0 outside coupled stepping)
    Phase phf{}, phd{};

    // fluid state, double-buffered per record group
    FluidPos* P[2] = {};
    float4* V[2] = {};
    double *vx[2] = {}, *vy[2] = {}, *vz[2] = {};
    float* drho[2] = {};
    int* orig[2] = {};
    int cP = 0, cV = 0, cD = 0, cO = 0;  // current buffer of P, V (with FP64 velocity), drho, orig
    int* rank = nullptr;
    // sort
    uint32_t *key = nullptr, *skey = nullptr, *ktmp = nullptr;
    int *perm = nullptr, *vtmp = nullptr;
    void* sort_temp = nullptr;
    int2* cells = nullptr;  // (start, end) per cell
    bool cells_track_skey = false;  // cell ranges were last written from the fluid keys in skey
    // lists
    int *nbr_ff = nullptr, *cnt_ff = nullptr, *nbr_fd = nullptr, *cnt_fd = nullptr;
    int* maxcount = nullptr;      // device [2]
    int* maxcount_h = nullptr;    // pinned [2]
    bool lists_valid = false;
    // 32-byte gather records for the pair passes (uniform scenes): knob,
    // whether the current lists carry cell codes, the two record buffers,
    // each particle's packed cell at the last sort and stencil centre
    bool pair_records = true;
    bool lists_coded = false;
    GRec* GR[2] = {};  // nf records each, stored as split halves
    float4* gr(int b) const { return reinterpret_cast<float4*>(GR[b]); }
    int *pcell = nullptr, *pcen = nullptr;
    // dummies
    DummyPos* DP = nullptr;
    float4* DV = nullptr;
    uint8_t* dphase = nullptr;
    int* dorig = nullptr;
    int2* dcells = nullptr;
    uint8_t* dummy_near = nullptr;
    // checkpoint for replay on list overflow
    FluidPos* ck_P = nullptr;
    float4* ck_V = nullptr;
    double *ck_vx = nullptr, *ck_vy = nullptr, *ck_vz = nullptr;
    float* ck_drho = nullptr;
    int *ck_orig = nullptr, *ck_rank = nullptr, *ck_stamp = nullptr;
    KernelTimer timer;
    // Captured two-step graphs.  A step flips the P, V and orig buffers once
    // and drho twice, so two steps return every ping-pong index to where it
    // started and a replay sees the same buffers as the capture.  Everything
    // else a capture bakes in (pointers, sizes, grid, phases, capacities) is
    // fixed between alloc_lists() / release(), which drop the cache.
    struct GraphKey {
        double dt;
        int parity;  // cP | cV << 1 | cD << 2 | cO << 3
        int ns, K_fs;
        const void *SP, *nbr_fs;
        bool operator==(const GraphKey& o) const
        {
            return dt == o.dt && parity == o.parity && ns == o.ns && K_fs == o.K_fs && SP == o.SP &&
                   nbr_fs == o.nbr_fs;
        }
    };
    struct GraphEntry {
        GraphKey key;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    bool use_graphs = true;
    long graph_captures = 0;
    StepStamps stamps;  // step-end timestamps (batched serving)
    int pair_lanes_set = 0;  // 0: chosen from the particle count
    int list_lanes_set = 0;
    // FP32 prefilter of the list walks (exact decisions; see pos_f32)
    bool prefilter = true;
    static constexpr float kPrefilterBand = 1.0e-3f;
    double pref_R = 0.0;
    float4 *Pf = nullptr, *DPf = nullptr;

    std::vector<demsph_couple*> couples;  // couplings holding this scene

    ~demsph_sph();

    void drop_graphs()
    {
        if (!graphs.empty() && stream) (void)hipStreamSynchronize(stream);
        for (auto& e : graphs) graph_exec_destroy(e.exec);
        graphs.clear();
    }

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (int b = 0; b < 2; b++) {
            dfree(P[b]); dfree(V[b]); dfree(vx[b]); dfree(vy[b]); dfree(vz[b]); dfree(drho[b]); dfree(orig[b]);
        }
        dfree(rank); dfree(key); dfree(skey); dfree(ktmp); dfree(perm); dfree(vtmp);
        dev_free(sort_temp);
        sort_temp = nullptr;
        dfree(cells); dfree(nbr_ff); dfree(cnt_ff); dfree(nbr_fd); dfree(cnt_fd); dfree(Pf); dfree(DPf);
        dfree(GR[0]); dfree(GR[1]); dfree(pcell); dfree(pcen);
        dfree(maxcount);
        host_free(maxcount_h);
        maxcount_h = nullptr;
        dfree(DP); dfree(DV); dfree(dphase); dfree(dorig); dfree(dcells); dfree(dummy_near); dfree(d_phases);
        dfree(ck_P); dfree(ck_V); dfree(ck_vx); dfree(ck_vy); dfree(ck_vz); dfree(ck_drho); dfree(ck_orig);
        dfree(ck_rank); dfree(ck_stamp);
        if (own_stream) stream_destroy(stream);
        stream = nullptr;
    }

    uint8_t phase_of(double h, double m, double c, double nu, double r0)
    {
        for (size_t k = 0; k < phases.size(); k++) {
            const Phase& q = phases[k];
            if (q.h == h && (double)q.m == (double)(float)m && q.c == (float)c && q.nu == (float)nu &&
                q.rho0 == (float)r0)
                return (uint8_t)k;
        }
        if (phases.size() >= (size_t)kMaxPhases)
            throw InvalidArg("more than 256 distinct (spacing, density, sound speed, viscosity) tuples");
        Phase q{};
        q.h = h;
        q.hf = (float)h;
        q.m = (float)m;
        q.c = (float)c;
        q.nu = (float)nu;
        q.rho0 = (float)r0;
        q.c2 = (float)(c * c);
        q.taitB = (float)(r0 * c * c / gamma);
        phases.push_back(q);
        return (uint8_t)(phases.size() - 1);
    }

    SphArgs args() const
    {
        SphArgs a{};
        a.nf = nf; a.nd = nd; a.phases = d_phases; a.phf = phf; a.phd = phd;
        a.P = P[cP]; a.V = V[cV]; a.vx = vx[cV]; a.vy = vy[cV]; a.vz = vz[cV]; a.drho = drho[cD];
        a.oP = P[1 - cP]; a.oV = V[1 - cV]; a.ovx = vx[1 - cV]; a.ovy = vy[1 - cV]; a.ovz = vz[1 - cV];
        a.odrho = drho[1 - cD];
        a.nbr_ff = nbr_ff; a.cnt_ff = cnt_ff; a.nbr_fd = nbr_fd; a.cnt_fd = cnt_fd;
        a.K_ff = K_ff; a.K_fd = K_fd;
        a.pcell = pcell; a.pcen = pcen; a.grid = grid;
        a.DP = DP; a.DV = DV; a.dphase = dphase;
        a.ns = solids.ns; a.SP = solids.SP; a.SV = solids.SV; a.SW = solids.SW;
        a.nbr_fs = solids.nbr_fs; a.cnt_fs = solids.cnt_fs; a.K_fs = solids.K_fs; a.acs = solids.acs;
        a.gx = (float)g[0]; a.gy = (float)g[1]; a.gz = (float)g[2];
        a.eos = eos;
        a.gamma = (float)gamma;
        a.inv_gamma = (float)(1.0 / gamma);
        return a;
    }

    static int round_chunk(int k) { return std::max(kChunk, (k + kChunk - 1) / kChunk * kChunk); }

    void alloc_lists()
    {
        drop_graphs();
        dfree(nbr_ff); dfree(nbr_fd);
        K_ff = round_chunk(K_ff);
        K_fd = round_chunk(K_fd);
        // the list walks address entries by 32-bit offsets
        if ((long long)std::max(K_ff, K_fd) * std::max(nf, 1) >= (1ll << 31))
            throw CapacityError("neighbour lists of " + std::to_string(std::max(K_ff, K_fd)) + " entries for " +
                                std::to_string(nf) + " fluid particles exceed 2^31 entries");
        nbr_ff = dmalloc<int>((size_t)K_ff * (nf > 0 ? nf : 1));
        nbr_fd = dmalloc<int>((size_t)K_fd * (nf > 0 ? nf : 1));
        lists_valid = false;
    }

    // grid cell = 2 * max smoothing length of fluid and dummies (WCSPHSolver.h:127-135)
    double staged_cell_size() const
    {
        double cmax = 0.0;
        for (double dx : hf.spacing) cmax = std::max(cmax, 1.3 * dx * 2.0);
        double cd2 = 0.0;
        for (double dx : hd.spacing) cd2 = std::max(cd2, 1.3 * dx * 2.0);
        if (!hd.spacing.empty()) cmax = std::max(cmax, cd2);
        return cmax;
    }

    // Device bytes initialize() and the first steps allocate for the staged
    // particles: double-buffered fluid records, sort buffers, the chunked
    // lists at the current capacities, the replay checkpoint, the dummy
    // records, the per-cell ranges and the upload staging.  Used to admit scenes under a memory
    // budget.
    size_t bytes_estimate(const double* minb, const double* maxb) const
    {
        const size_t n = hf.size(), m = hd.size();
        GridGeom gg{};
        const size_t cells = grid_set(gg, minb, maxb, staged_cell_size()) ? (size_t)gg.numGrids() : 0;
        const size_t rec = sizeof(FluidPos) + sizeof(float4) + 3 * sizeof(double) + sizeof(float) + sizeof(int);
        const size_t per_fluid = 2 * rec + rec + sizeof(int)             // two buffers, checkpoint, rank
                                 + sizeof(float4)                         // FP32 list-prefilter copy
                                 + 2 * sizeof(GRec) + 2 * sizeof(int)     // gather records, cells
                                 + 2 * sizeof(int)                        // list counts
                                 + sizeof(int) * (size_t)(round_chunk(K_ff) + round_chunk(K_fd));
        const size_t per_sort = 3 * sizeof(uint32_t) + 2 * sizeof(int);
        const size_t per_dummy = sizeof(DummyPos) + 2 * sizeof(float4) + 1 + sizeof(int);
        // initialisation uploads positions, velocities and phases through
        // temporary buffers, which the runtime's allocator keeps cached
        const size_t upload = 6 * sizeof(double) + 1;
        return n * (per_fluid + upload) + m * (per_dummy + upload) + std::max(n, m) * per_sort +
               cells * (2 * sizeof(int2) + 1) + radix_sort_temp_bytes((int)std::max<size_t>(std::max(n, m), 1)) +
               (size_t(4) << 20);
    }

    void initialize(const double* minb, const double* maxb, const double* gravity)
    {
        if (initialized) release_state_keep_host();
        if (!stream) {
            DEMSPH_HIP(stream_create(&stream));
            own_stream = true;
        }
        g[0] = gravity[0]; g[1] = gravity[1]; g[2] = gravity[2];
        nf = (int)hf.size();
        nd = (int)hd.size();
        if (nf == 0) throw InvalidArg("no fluid particles");
        if (!grid_set(grid, minb, maxb, staged_cell_size())) throw InvalidArg("degenerate domain or smoothing length");
        numGrids = grid.numGrids();
        if (numGrids > INT32_MAX / 2) throw InvalidArg("grid too large");

        phases.clear();
        fphase_h.resize(nf);
        dphase_h.resize(nd);
        for (int i = 0; i < nf; i++) {
            const double dx = hf.spacing[i];
            fphase_h[i] = phase_of(1.3 * dx, dx * dx * dx * hf.rho0[i], hf.c[i], hf.nu[i], hf.rho0[i]);
        }
        for (int i = 0; i < nd; i++) {
            const double dx = hd.spacing[i];
            dphase_h[i] = phase_of(1.3 * dx, dx * dx * dx * hd.rho0[i], hd.c[i], hd.nu[i], hd.rho0[i]);
        }
        bool fu = true, du = true;
        for (int i = 1; i < nf; i++) fu = fu && fphase_h[i] == fphase_h[0];
        for (int i = 1; i < nd; i++) du = du && dphase_h[i] == dphase_h[0];
        uniform = fu && du;
        fluid_one_phase = fu;
        phf = phases[fphase_h[0]];
        phd = nd > 0 ? phases[dphase_h[0]] : phf;
        d_phases = dmalloc<Phase>(phases.size());
        DEMSPH_HIP(hipMemcpyAsync(d_phases, phases.data(), sizeof(Phase) * phases.size(), hipMemcpyHostToDevice, stream));

        const size_t n = nf;
        for (int b = 0; b < 2; b++) {
            P[b] = dmalloc<FluidPos>(n); V[b] = dmalloc<float4>(n);
            vx[b] = dmalloc<double>(n); vy[b] = dmalloc<double>(n); vz[b] = dmalloc<double>(n);
            drho[b] = dmalloc<float>(n); orig[b] = dmalloc<int>(n);
        }
        rank = dmalloc<int>(n);
        const int nsort = std::max(nf, nd);
        key = dmalloc<uint32_t>(nsort); skey = dmalloc<uint32_t>(nsort); ktmp = dmalloc<uint32_t>(nsort);
        perm = dmalloc<int>(nsort); vtmp = dmalloc<int>(nsort);
        DEMSPH_HIP(dev_malloc(&sort_temp, radix_sort_temp_bytes(nsort)));
        cells = dmalloc<int2>(numGrids);
        Pf = dmalloc<float4>(nf);
        DPf = dmalloc<float4>(nd > 0 ? nd : 1);
        // Points within pref_R of the grid origin on every axis carry FP32
        // copies whose rounding stays inside the prefilter band for pairs at
        // the smallest cut: |d2_f32 - d2| <= 2 sqrt(3) cut 2^-22 R + O(2^-24)
        // cut^2 <= band cut^2 / 2 for R = 500 cut_min.
        {
            double hmin = 1e300;
            for (const Phase& q : phases) hmin = std::min(hmin, q.h);
            pref_R = 500.0 * 2.0 * hmin;
        }
        cnt_ff = dmalloc<int>(n); cnt_fd = dmalloc<int>(n);
        GR[0] = dmalloc<GRec>(n); GR[1] = dmalloc<GRec>(n);
        pcell = dmalloc<int>(n); pcen = dmalloc<int>(n);
        maxcount = dmalloc<int>(3);
        DEMSPH_HIP(host_malloc((void**)&maxcount_h, 3 * sizeof(int)));
        alloc_lists();
        cP = cV = cD = cO = 0;

        // upload fluid in insertion order (slot == index)
        DEMSPH_HIP(hipStreamSynchronize(stream));
        {
            double* upos = dmalloc<double>(3 * n);
            double* uvel = dmalloc<double>(3 * n);
            uint8_t* uph = dmalloc<uint8_t>(n);
            // uploads ordered on the scene's stream ahead of the kernels that read them
            DEMSPH_HIP(hipMemcpyAsync(upos, hf.pos.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, stream));
            DEMSPH_HIP(hipMemcpyAsync(uvel, hf.vel.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, stream));
            DEMSPH_HIP(hipMemcpyAsync(uph, fphase_h.data(), n, hipMemcpyHostToDevice, stream));
            const dim3 gb(blocks_for(n, 256)), tb(256);
            hipLaunchKernelGGL(k_fluid_init, gb, tb, 0, stream, (int)n, upos, uvel, uph, P[0], V[0], vx[0], vy[0],
                               vz[0], drho[0]);
            hipLaunchKernelGGL(k_iota, gb, tb, 0, stream, orig[0], (int)n);
            hipLaunchKernelGGL(k_iota, gb, tb, 0, stream, rank, (int)n);
            hipLaunchKernelGGL(k_eos_all, gb, tb, 0, stream, (int)n, d_phases, V[0], drho[0], P[0], eos,
                               (float)gamma);
            DEMSPH_HIP(hipGetLastError());
            DEMSPH_HIP(hipStreamSynchronize(stream));
            dfree(upos); dfree(uvel); dfree(uph);
        }

        // dummies: sort once (they are static), cell ranges, normals
        dcells = dmalloc<int2>(numGrids);
        fill_i32(reinterpret_cast<int*>(dcells), 2 * numGrids, -1, stream);
        dummy_near = dmalloc<uint8_t>(numGrids);
        const size_t m = nd;
        DP = dmalloc<DummyPos>(m); DV = dmalloc<float4>(m);
        dphase = dmalloc<uint8_t>(m); dorig = dmalloc<int>(m);
        if (nd > 0) {
            std::vector<double> tx(m), ty(m), tz(m);
            double *ux = dmalloc<double>(m), *uy = dmalloc<double>(m), *uz = dmalloc<double>(m);
            double *wx = dmalloc<double>(m), *wy = dmalloc<double>(m), *wz = dmalloc<double>(m);
            uint8_t* uph = dmalloc<uint8_t>(m);
            auto up3 = [&](const std::vector<double>& src, double* dx_, double* dy_, double* dz_) {
                for (size_t i = 0; i < m; i++) { tx[i] = src[3 * i]; ty[i] = src[3 * i + 1]; tz[i] = src[3 * i + 2]; }
                DEMSPH_HIP(hipMemcpyAsync(dx_, tx.data(), m * sizeof(double), hipMemcpyHostToDevice, stream));
                DEMSPH_HIP(hipMemcpyAsync(dy_, ty.data(), m * sizeof(double), hipMemcpyHostToDevice, stream));
                DEMSPH_HIP(hipMemcpyAsync(dz_, tz.data(), m * sizeof(double), hipMemcpyHostToDevice, stream));
                DEMSPH_HIP(hipStreamSynchronize(stream));  // tx, ty, tz are reused
            };
            up3(hd.pos, ux, uy, uz);
            up3(hd.vel, wx, wy, wz);
            DEMSPH_HIP(hipMemcpyAsync(uph, dphase_h.data(), m, hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(k_dummy_hash, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, stream, (int)m, grid,
                               ux, uy, uz, key);
            radix_sort_pairs(key, nullptr, skey, perm, ktmp, vtmp, sort_temp, (int)m, key_bits(numGrids), stream);
            DummyReorderArgs ra{};
            ra.n = (int)m; ra.numGrids = numGrids; ra.skey = skey; ra.perm = perm;
            ra.px = ux; ra.py = uy; ra.pz = uz; ra.vx = wx; ra.vy = wy; ra.vz = wz; ra.phase = uph;
            ra.DP = DP; ra.DV = DV; ra.ophase = dphase;
            ra.orig = dorig; ra.cells = dcells;
            ra.DPf = DPf; ra.grid = grid; ra.R = pref_R;
            hipLaunchKernelGGL(k_dummy_reorder, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, stream, ra);
            NormalArgs na{};
            na.nd = (int)m; na.grid = grid; na.phases = d_phases; na.DP = DP; na.DV = DV;
            na.dphase = dphase; na.dcells = dcells;
            hipLaunchKernelGGL(k_dummy_normals, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, stream, na);
            DEMSPH_HIP(hipGetLastError());
            DEMSPH_HIP(hipStreamSynchronize(stream));
            dfree(ux); dfree(uy); dfree(uz); dfree(wx); dfree(wy); dfree(wz); dfree(uph);
        }
        hipLaunchKernelGGL(k_dummy_near, dim3(blocks_for(numGrids, 256)), dim3(256), 0, stream, grid, numGrids, dcells,
                           dummy_near);
        DEMSPH_HIP(hipGetLastError());
        // the fluid cell ranges start empty; skey now holds dummy keys
        fill_i32(reinterpret_cast<int*>(cells), 2 * numGrids, -1, stream);
        cells_track_skey = false;
        DEMSPH_HIP(hipStreamSynchronize(stream));
        initialized = true;
        lists_valid = false;
    }

    void release_state_keep_host()
    {
        hipStream_t s = stream;
        bool own = own_stream;
        stream = nullptr;  // keep the stream alive across re-initialisation
        own_stream = false;
        (void)hipStreamSynchronize(s);
        release();
        stream = s;
        own_stream = own;
        initialized = false;
    }

    void require_init() const
    {
        if (!initialized) throw InvalidArg("scene not initialized");
    }

    // hash -> sort -> reorder -> cell ranges -> lists; counts list overflow
    // into maxcount (device).
    void search_enqueue()
    {
        const int n = nf;
        const dim3 gb(blocks_for(n, kThreads)), tb(kThreads);
        if (!cells_track_skey) {
            timer.begin("sph_fill_cells", stream);
            fill_i32(reinterpret_cast<int*>(cells), 2 * numGrids, -1, stream);
            timer.end(stream);
        }
        DEMSPH_LAUNCH(timer, "sph_hash", k_hash, gb, tb, 0, stream, n, grid, numGrids, P[cP], orig[cO], key,
                      cells_track_skey ? skey : nullptr, cells);
        timer.begin("sph_sort", stream);
        radix_sort_pairs(key, nullptr, skey, perm, ktmp, vtmp, sort_temp, n, key_bits(numGrids), stream);
        timer.end(stream);
        ReorderArgs ra{};
        ra.n = n; ra.numGrids = numGrids; ra.skey = skey; ra.perm = perm; ra.rank = rank; ra.orig_out = orig[1 - cO];
        ra.P = P[cP]; ra.V = V[cV]; ra.vx = vx[cV]; ra.vy = vy[cV]; ra.vz = vz[cV]; ra.drho = drho[cD];
        ra.oP = P[1 - cP]; ra.oV = V[1 - cV]; ra.ovx = vx[1 - cV]; ra.ovy = vy[1 - cV]; ra.ovz = vz[1 - cV];
        ra.odrho = drho[1 - cD];
        ra.cells = cells;
        ra.oPf = prefilter ? Pf : nullptr; ra.grid = grid; ra.R = pref_R;
        const bool coded = records_usable();
        ra.oGR = coded ? gr(0) : nullptr; ra.opcell = pcell;
        DEMSPH_LAUNCH(timer, "sph_reorder", k_reorder, gb, tb, 0, stream, ra);
        cP ^= 1; cV ^= 1; cD ^= 1; cO ^= 1;
        cells_track_skey = true;
        ListArgs la{};
        la.nf = nf; la.nd = nd; la.grid = grid; la.phases = d_phases; la.hf = phf.h; la.hd = phd.h;
        la.P = P[cP]; la.V = V[cV]; la.cells = cells;
        la.DP = DP; la.dphase = dphase; la.dcells = dcells; la.dummy_near = dummy_near;
        la.nbr_ff = nbr_ff; la.cnt_ff = cnt_ff; la.nbr_fd = nbr_fd; la.cnt_fd = cnt_fd;
        la.K_ff = K_ff; la.K_fd = K_fd; la.maxcount = maxcount;
        la.Pf = Pf; la.DPf = DPf; la.band = kPrefilterBand;
        la.pcen = coded ? pcen : nullptr;
        lists_coded = coded;
        if (uniform) lists_lanes<true>(la);
        else lists_lanes<false>(la);
        lists_valid = true;
    }

    // Lanes per fluid particle in the list build: the fewest of 4, 8 that
    // put 600k threads in flight (measured on MI355X: 8 lanes at 2.7k and
    // 35k particles, 4 at 256k and 741k, where 4 beat 2 by 10 and 4 %).
    int list_lanes() const
    {
        if (list_lanes_set > 0) return list_lanes_set;
        return list_lanes_rule();
    }
    int list_lanes_rule() const { return lanes_for(600000, 4); }
    int lanes_for(long long fill, int least = 2) const
    {
        int L = least;
        while (L < 8 && (long long)nf * L < fill) L *= 2;
        return L;
    }
    template <bool U, int L>
    void lists_launch(const ListArgs& la)
    {
        const dim3 gb(blocks_for((long long)nf * L, kThreads)), tb(kThreads);
        if constexpr (U) {
            if (la.pcen) {
                if (L > 1 && prefilter)
                    DEMSPH_LAUNCH(timer, "sph_lists", (k_build_lists<U, L, true, true>), gb, tb, 0, stream, la);
                else DEMSPH_LAUNCH(timer, "sph_lists", (k_build_lists<U, L, false, true>), gb, tb, 0, stream, la);
                return;
            }
        }
        if (L > 1 && prefilter)
            DEMSPH_LAUNCH(timer, "sph_lists", (k_build_lists<U, L, true, false>), gb, tb, 0, stream, la);
        else DEMSPH_LAUNCH(timer, "sph_lists", (k_build_lists<U, L, false, false>), gb, tb, 0, stream, la);
    }
    // Gather records and coded lists need one phase, slots below 2^25 and a
    // grid of at most 2048 x 2048 x 1024 cells, and the entry-split passes.
    bool records_usable() const
    {
        return pair_records && uniform && nf > 0 && nf <= kSlotMask && grid.nx <= 2048 && grid.ny <= 2048 &&
               grid.nz <= 1024 && pair_lanes() >= 2;
    }
    template <bool U>
    void lists_lanes(const ListArgs& la)
    {
        switch (list_lanes()) {
        case 2: lists_launch<U, 2>(la); break;
        case 4: lists_launch<U, 4>(la); break;
        case 8: lists_launch<U, 8>(la); break;
        default: lists_launch<U, 1>(la); break;
        }
    }

    // Lanes per fluid particle in the pair passes: 4 or 8, the fewer that
    // put 400k threads in flight (measured on MI355X with the entry split:
    // 8 lanes at 32k particles, 4 at 256k and 741k, where 4 beat 2 by 5 and
    // 9 % and 8 by 2 and 4 %).  One lane keeps every sum in list order.
    int pair_lanes() const
    {
        if (pair_lanes_set > 0) return pair_lanes_set;
        return lanes_for(400000, 4);
    }

    template <bool FIRST, bool U, int L>
    void momentum_launch(const char* name, const SphArgs& a)
    {
        const dim3 gb(blocks_for((long long)nf * L, kThreads)), tb(kThreads);
        if constexpr (U && L >= 2) {
            if (a.GRin) {
                DEMSPH_LAUNCH(timer, name, (k_momentum<FIRST, U, L, true>), gb, tb, 0, stream, a);
                return;
            }
        }
        DEMSPH_LAUNCH(timer, name, (k_momentum<FIRST, U, L>), gb, tb, 0, stream, a);
    }
    template <bool FIRST, bool U>
    void momentum_lanes(const char* name, const SphArgs& a)
    {
        switch (pair_lanes()) {
        case 2: momentum_launch<FIRST, U, 2>(name, a); break;
        case 4: momentum_launch<FIRST, U, 4>(name, a); break;
        case 8: momentum_launch<FIRST, U, 8>(name, a); break;
        default: momentum_launch<FIRST, U, 1>(name, a); break;
        }
    }
    template <bool U, int L>
    void continuity_launch(const SphArgs& a)
    {
        const dim3 gb(blocks_for((long long)nf * L, kThreads)), tb(kThreads);
        if constexpr (U && L >= 2) {
            if (a.GRin) {
                DEMSPH_LAUNCH(timer, "sph_continuity", (k_continuity<U, L, true>), gb, tb, 0, stream, a);
                return;
            }
        }
        DEMSPH_LAUNCH(timer, "sph_continuity", (k_continuity<U, L>), gb, tb, 0, stream, a);
    }
    template <bool U>
    void continuity_lanes(const SphArgs& a)
    {
        switch (pair_lanes()) {
        case 2: continuity_launch<U, 2>(a); break;
        case 4: continuity_launch<U, 4>(a); break;
        case 8: continuity_launch<U, 8>(a); break;
        default: continuity_launch<U, 1>(a); break;
        }
    }

    void first_half_enqueue(double dt)
    {
        SphArgs a = args();
        a.dt = dt;
        if (lists_coded) { a.GRin = gr(0); a.GRout = gr(1); }
        if (uniform) momentum_lanes<true, true>("sph_momentum_first", a);
        else momentum_lanes<true, false>("sph_momentum_first", a);
        cP ^= 1; cV ^= 1;
    }

    void continuity_enqueue(double dt)
    {
        SphArgs a = args();
        a.dt = dt;
        if (lists_coded) { a.GRin = gr(1); a.GRout = gr(0); }
        if (uniform) continuity_lanes<true>(a);
        else continuity_lanes<false>(a);
        cP ^= 1; cD ^= 1;
    }

    void momentum2_enqueue(double dt)
    {
        SphArgs a = args();
        a.dt = dt;
        if (lists_coded) a.GRin = gr(0);
        if (uniform) momentum_lanes<false, true>("sph_momentum_second", a);
        else momentum_lanes<false, false>("sph_momentum_second", a);
        cV ^= 1;
    }

    void second_half_enqueue(double dt)
    {
        continuity_enqueue(dt);
        momentum2_enqueue(dt);
    }

    // A fill kernel rather than a memset node, so that the reset may sit in a
    // captured graph.
    void reset_overflow() { fill_i32(maxcount, 3, 0, stream); }

    // Returns true when a list overflowed since reset_overflow(); grows the
    // capacities in that case.  Synchronises the stream.
    bool check_overflow()
    {
        DEMSPH_HIP(hipMemcpyAsync(maxcount_h, maxcount, 3 * sizeof(int), hipMemcpyDeviceToHost, stream));
        DEMSPH_HIP(hipStreamSynchronize(stream));
        timer.harvest();
        bool grew = false;
        if (maxcount_h[0] > K_ff) { K_ff = maxcount_h[0] + maxcount_h[0] / 4 + 8; grew = true; }
        if (maxcount_h[1] > K_fd) { K_fd = maxcount_h[1] + maxcount_h[1] / 4 + 8; grew = true; }
        if (grew) alloc_lists();
        return grew;
    }

    void ensure_checkpoint()
    {
        if (ck_P) return;
        const size_t n = nf;
        ck_P = dmalloc<FluidPos>(n); ck_V = dmalloc<float4>(n);
        ck_vx = dmalloc<double>(n); ck_vy = dmalloc<double>(n); ck_vz = dmalloc<double>(n);
        ck_drho = dmalloc<float>(n); ck_orig = dmalloc<int>(n); ck_rank = dmalloc<int>(n);
        ck_stamp = dmalloc<int>(1);
    }

    void checkpoint(bool save)
    {
        const size_t n = nf;
        auto cpb = [&](void* ck, void* cur, size_t bytes) {
            if (save) DEMSPH_HIP(hipMemcpyAsync(ck, cur, bytes, hipMemcpyDeviceToDevice, stream));
            else DEMSPH_HIP(hipMemcpyAsync(cur, ck, bytes, hipMemcpyDeviceToDevice, stream));
        };
        cpb(ck_P, P[cP], n * sizeof(FluidPos));
        cpb(ck_V, V[cV], n * sizeof(float4));
        cpb(ck_vx, vx[cV], n * 8); cpb(ck_vy, vy[cV], n * 8); cpb(ck_vz, vz[cV], n * 8);
        cpb(ck_drho, drho[cD], n * 4);
        cpb(ck_orig, orig[cO], n * 4);
        cpb(ck_rank, rank, n * 4);
        // a replayed block stamps its steps again from the same count
        if (stamps.counter) cpb(ck_stamp, stamps.counter, sizeof(int));
    }

    void neighbor_search()
    {
        require_init();
        // A truncated list is rebuilt with the grown capacity; the sort is
        // deterministic, so repeating it on sorted state is idempotent.
        for (;;) {
            reset_overflow();
            search_enqueue();
            if (!check_overflow()) return;
        }
    }

    void one_step(double dt)
    {
        search_enqueue();
        first_half_enqueue(dt);
        second_half_enqueue(dt);
        step_stamp(stamps, stream);
    }

    void set_stamps(const StepStamps& st)
    {
        drop_graphs();  // captured graphs hold the previous stamp buffer
        stamps = st;
    }

    GraphKey graph_key(double dt) const
    {
        GraphKey k{};
        k.dt = dt;
        k.parity = cP | cV << 1 | cD << 2 | cO << 3;
        k.ns = solids.ns;
        k.K_fs = solids.K_fs;
        k.SP = solids.SP;
        k.nbr_fs = solids.nbr_fs;
        return k;
    }

    // The executable graph of two steps from the current state, captured on
    // first use.  Capturing enqueues nothing; the caller launches the graph.
    hipGraphExec_t pair_graph(double dt)
    {
        const GraphKey key = graph_key(dt);
        for (auto& e : graphs)
            if (e.key == key) return e.exec;
        const int sP = cP, sV = cV, sD = cD, sO = cO;
        hipGraph_t g = nullptr;
        // No other thread may allocate, free or synchronise the device
        // until the graph is instantiated (see capture_lock).
        CaptureGuard capture_guard(capture_lock());
        DEMSPH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            one_step(dt);
            one_step(dt);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            cP = sP; cV = sV; cD = sD; cO = sO;
            throw;
        }
        DEMSPH_HIP(hipStreamEndCapture(stream, &g));
        if (cP != sP || cV != sV || cD != sD || cO != sO) {
            (void)hipGraphDestroy(g);
            throw std::logic_error("two steps did not restore the buffer parity");
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        DEMSPH_HIP(e);
        if (graphs.size() >= 4) {
            DEMSPH_HIP(hipStreamSynchronize(stream));
            graph_exec_destroy(graphs.front().exec);
            graphs.erase(graphs.begin());
        }
        graphs.push_back({key, ex});
        graph_captures++;
        return ex;
    }

    // nsteps steps: eagerly while the cell ranges are not yet tied to the
    // sorted keys (the first step after initialisation) and for an odd
    // remainder, as replays of the two-step graph otherwise.
    void enqueue_steps(double dt, int nsteps)
    {
        int k = 0;
        if (use_graphs && !timer.enabled) {
            if (!cells_track_skey && k < nsteps) { one_step(dt); k++; }
            for (; k + 2 <= nsteps; k += 2) DEMSPH_HIP(hipGraphLaunch(pair_graph(dt), stream));
        }
        for (; k < nsteps; k++) one_step(dt);
    }

    void step(double dt, int nsteps)
    {
        require_init();
        if (dt <= 0.0) throw InvalidArg("time step must be positive");
        if (nsteps <= 0) return;
        ensure_checkpoint();
        for (int attempt = 0; attempt < 8; attempt++) {
            checkpoint(true);
            const int sP = cP, sV = cV, sD = cD, sO = cO;
            reset_overflow();
            enqueue_steps(dt, nsteps);
            if (!check_overflow()) return;
            // Replay the block from the checkpoint with larger lists.
            cP = sP; cV = sV; cD = sD; cO = sO;
            checkpoint(false);
        }
        throw std::runtime_error("neighbour list capacity did not converge");
    }

    // ------------------------------------------------------------ autotuner
    // Times the list-build and pair-pass variants on this scene's own steps,
    // from a checkpoint that is restored after every trial, and keeps the
    // fastest.  Mode 0 tries only knobs that leave every result bit-identical
    // (list lanes, the FP32 prefilter: the lists are the same entries in the
    // same order); mode 1 also tries pair lanes, which change the FP32
    // summation order of the pair passes.  The cell size is not a knob: it
    // fixes the pair sets' order and the sort (2 max h, WCSPHSolver.h:127-135).
    // Choices are kept per size class in a process-wide cache, so that scenes
    // of one class get the same variants whichever is tuned first.
    struct TuneChoice {
        int list_lanes = 0, prefilter = 1, pair_lanes = 0;
    };
    struct TuneOutcome {
        TuneChoice chosen, rules;
        int configs = 0;
        bool cached = false;
        double list_us = 0.0, pair_us = 0.0;              // chosen variants, per step
        double rules_list_us = 0.0, rules_pair_us = 0.0;  // built-in lane rules, per step
    };
    using TuneKey = std::array<int, 4>;
    static std::mutex& tune_mutex()
    {
        static std::mutex m;
        return m;
    }
    static std::map<TuneKey, TuneChoice>& tune_cache()
    {
        static std::map<TuneKey, TuneChoice> c;
        return c;
    }
    // Size class: uniform phase, fluid count to a quarter octave, dummy
    // count to a half octave, mode.
    TuneKey tune_key(int mode) const
    {
        const auto oct = [](double v, double per) { return (int)std::lround(per * std::log2(v + 1.0)); };
        return {uniform ? 1 : 0, oct(nf, 4.0), oct(nd, 2.0), mode};
    }

    void apply_choice(const TuneChoice& c, int mode)
    {
        drop_graphs();
        list_lanes_set = c.list_lanes;
        prefilter = c.prefilter != 0;
        if (mode == 1) pair_lanes_set = c.pair_lanes;
    }

    // Mean ms per step of the list build and of the three pair passes over
    // `steps` eager steps from the checkpoint (one untimed step first).
    std::pair<double, double> time_trial(double dt, int steps, int sP, int sV, int sD, int sO)
    {
        for (int attempt = 0; attempt < 8; attempt++) {
            cP = sP; cV = sV; cD = sD; cO = sO;
            checkpoint(false);
            reset_overflow();
            timer.enabled = false;
            one_step(dt);
            timer.enabled = true;
            timer.reset();
            for (int k = 0; k < steps; k++) one_step(dt);
            if (check_overflow()) continue;  // grown lists: time the trial again
            double lists = 0.0, pairs = 0.0;
            for (const auto& kv : timer.totals()) {
                if (kv.first == "sph_lists") lists += kv.second.second;
                else if (kv.first == "sph_momentum_first" || kv.first == "sph_momentum_second" ||
                         kv.first == "sph_continuity")
                    pairs += kv.second.second;
            }
            timer.reset();
            return {lists / steps, pairs / steps};
        }
        throw std::runtime_error("neighbour list capacity did not converge");
    }

    TuneOutcome autotune(double dt, int steps, int mode, bool use_cache)
    {
        require_init();
        if (dt <= 0.0) throw InvalidArg("time step must be positive");
        if (steps < 1 || steps > 1000) throw InvalidArg("trial steps must be in 1..1000");
        if (mode != 0 && mode != 1) throw InvalidArg("mode must be 0 (result-invariant knobs) or 1 (also pair lanes)");
        TuneOutcome out;
        out.rules = TuneChoice{list_lanes_rule(), 1, lanes_for(400000, 4)};
        const TuneKey key = tune_key(mode);
        if (use_cache) {
            std::lock_guard<std::mutex> lk(tune_mutex());
            const auto it = tune_cache().find(key);
            if (it != tune_cache().end()) {
                out.chosen = it->second;
                out.cached = true;
                apply_choice(out.chosen, mode);
                if (mode == 0) out.chosen.pair_lanes = pair_lanes();
                return out;
            }
        }
        DEMSPH_HIP(hipStreamSynchronize(stream));
        timer.harvest();
        drop_graphs();
        std::map<std::string, std::pair<long, double>> saved_totals;
        timer.swap_totals(saved_totals);
        const bool saved_enabled = timer.enabled;
        const StepStamps saved_stamps = stamps;
        const int saved_list = list_lanes_set, saved_pair = pair_lanes_set;
        const bool saved_pre = prefilter;
        stamps = StepStamps{};  // trials leave no step timestamps
        ensure_checkpoint();
        checkpoint(true);
        const int sP = cP, sV = cV, sD = cD, sO = cO;
        const auto restore = [&] {
            cP = sP; cV = sV; cD = sD; cO = sO;
            checkpoint(false);
            DEMSPH_HIP(hipStreamSynchronize(stream));
            timer.harvest();
            timer.swap_totals(saved_totals);
            timer.enabled = saved_enabled;
            stamps = saved_stamps;
            lists_valid = false;  // the lists hold the last trial's step
        };
        TuneChoice best;
        try {
            // List build: every lane count, with and without the prefilter
            // (one lane has no prefilter variant).  Near-ties keep the rule.
            double best_l = 0.0, pair_sum = 0.0;
            for (int L : {1, 2, 4, 8}) {
                for (int pre = 0; pre < 2; pre++) {
                    if (L == 1 && pre) continue;
                    list_lanes_set = L;
                    prefilter = pre != 0;
                    const auto t = time_trial(dt, steps, sP, sV, sD, sO);
                    out.configs++;
                    const bool rule = L == out.rules.list_lanes && pre == 1;
                    if (rule) out.rules_list_us = t.first * 1e3;
                    if (out.configs == 1 || t.first < best_l) {
                        best_l = t.first;
                        best.list_lanes = L;
                        best.prefilter = pre;
                    }
                    pair_sum += t.second;
                }
            }
            // the pair passes ran with the scene's pair lanes in every trial
            out.pair_us = pair_sum / out.configs * 1e3;
            if (out.rules_list_us <= best_l * 1e3 * 1.02) {
                best.list_lanes = out.rules.list_lanes;
                best.prefilter = 1;
                best_l = out.rules_list_us * 1e-3;
            }
            out.list_us = best_l * 1e3;
            list_lanes_set = best.list_lanes;
            prefilter = best.prefilter != 0;
            if (mode == 1) {
                double best_p = 0.0;
                int k = 0;
                for (int L : {1, 2, 4, 8}) {
                    pair_lanes_set = L;
                    const auto t = time_trial(dt, steps, sP, sV, sD, sO);
                    out.configs++;
                    if (L == out.rules.pair_lanes) out.rules_pair_us = t.second * 1e3;
                    if (k++ == 0 || t.second < best_p) {
                        best_p = t.second;
                        best.pair_lanes = L;
                    }
                }
                if (out.rules_pair_us <= best_p * 1e3 * 1.02) {
                    best.pair_lanes = out.rules.pair_lanes;
                    best_p = out.rules_pair_us * 1e-3;
                }
                out.pair_us = best_p * 1e3;
            } else {
                best.pair_lanes = pair_lanes();
                if (best.pair_lanes == out.rules.pair_lanes) out.rules_pair_us = out.pair_us;
            }
        } catch (...) {
            list_lanes_set = saved_list;
            pair_lanes_set = saved_pair;
            prefilter = saved_pre;
            restore();
            throw;
        }
        restore();
        apply_choice(best, mode);
        out.chosen = best;
        {
            std::lock_guard<std::mutex> lk(tune_mutex());
            tune_cache()[key] = best;
        }
        return out;
    }

    void refresh_eos()
    {
        hipLaunchKernelGGL(k_eos_all, dim3(blocks_for(nf, 256)), dim3(256), 0, stream, nf, d_phases, V[cV], drho[cD],
                           P[cP], eos, (float)gamma);
        DEMSPH_HIP(hipGetLastError());
    }
};

namespace {

template <class T>
std::vector<T> fetch(const T* d, size_t n, hipStream_t s)
{
    std::vector<T> h(n);
    if (n) DEMSPH_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
    DEMSPH_HIP(hipStreamSynchronize(s));
    return h;
}

template <class T>
void emit(const std::vector<T>& v, void* out, size_t cap, size_t* count)
{
    if (count) *count = v.size();
    if (!out) return;
    if (cap < v.size()) throw InvalidArg("output capacity too small");
    std::memcpy(out, v.data(), v.size() * sizeof(T));
}

}  // namespace

extern "C" int demsph_sph_create(demsph_sph** out, void* stream)
{
    return guarded([&] {
        if (!out) throw InvalidArg("null output handle");
        auto* s = new demsph_sph();
        s->stream = static_cast<hipStream_t>(stream);
        s->own_stream = false;
        *out = s;
    });
}

extern "C" int demsph_sph_destroy(demsph_sph* s)
{
    return guarded([&] { delete s; });
}

extern "C" int demsph_sph_add_fluid(demsph_sph* s, size_t n, const double* position, const double* velocity,
                                    const double* soundSpeed, const double* spacing, const double* density,
                                    const double* viscosity)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        if (n && (!position || !soundSpeed || !spacing || !density || !viscosity)) throw InvalidArg("null array");
        s->hf.add(n, position, velocity, soundSpeed, spacing, density, viscosity);
    });
}

extern "C" int demsph_sph_add_dummy(demsph_sph* s, size_t n, const double* position, const double* velocity,
                                    const double* soundSpeed, const double* spacing, const double* density,
                                    const double* viscosity)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        if (n && (!position || !soundSpeed || !spacing || !density || !viscosity)) throw InvalidArg("null array");
        s->hd.add(n, position, velocity, soundSpeed, spacing, density, viscosity);
    });
}

extern "C" int demsph_sph_set_eos(demsph_sph* s, int eos, double gamma)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        if (eos != 0 && eos != 1) throw InvalidArg("eos must be 0 (linear) or 1 (Tait)");
        if (eos == 1 && !(gamma > 0.0)) throw InvalidArg("Tait gamma must be positive");
        if (s->initialized) throw InvalidArg("set the EOS before initialize()");
        s->eos = eos;
        s->gamma = gamma;
    });
}

extern "C" int demsph_sph_set_capacity(demsph_sph* s, int fluidCapacity, int dummyCapacity)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        if (fluidCapacity < 1 || dummyCapacity < 1) throw InvalidArg("capacities must be positive");
        s->K_ff = fluidCapacity;
        s->K_fd = dummyCapacity;
        if (s->initialized) s->alloc_lists();
    });
}

extern "C" int demsph_sph_initialize(demsph_sph* s, const double* minBound, const double* maxBound,
                                     const double* gravity)
{
    return guarded([&] {
        if (!s || !minBound || !maxBound || !gravity) throw InvalidArg("null argument");
        s->initialize(minBound, maxBound, gravity);
    });
}

extern "C" int demsph_sph_neighbor_search(demsph_sph* s)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        s->neighbor_search();
    });
}

extern "C" int demsph_sph_first_half(demsph_sph* s, double timeStep)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        s->require_init();
        if (!s->lists_valid) throw InvalidArg("neighbor_search() must precede the integration halves");
        s->first_half_enqueue(timeStep);
    });
}

extern "C" int demsph_sph_second_half(demsph_sph* s, double timeStep)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        s->require_init();
        if (!s->lists_valid) throw InvalidArg("neighbor_search() must precede the integration halves");
        s->second_half_enqueue(timeStep);
    });
}

extern "C" int demsph_sph_step(demsph_sph* s, double timeStep, int nsteps)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        s->step(timeStep, nsteps);
    });
}

extern "C" int demsph_sph_synchronize(demsph_sph* s)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        if (s->stream) DEMSPH_HIP(hipStreamSynchronize(s->stream));
        s->timer.harvest();
    });
}

extern "C" int demsph_sph_sizes(demsph_sph* s, size_t* numFluid, size_t* numDummy, size_t* numGrids)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        if (numFluid) *numFluid = s->initialized ? (size_t)s->nf : s->hf.size();
        if (numDummy) *numDummy = s->initialized ? (size_t)s->nd : s->hd.size();
        if (numGrids) *numGrids = (size_t)s->numGrids;
    });
}

extern "C" int demsph_sph_get(demsph_sph* s, const char* field, void* out, size_t capacity, size_t* count)
{
    return guarded([&] {
        if (!s || !field) throw InvalidArg("null argument");
        s->require_init();
        const std::string f = field;
        hipStream_t st = s->stream;
        const size_t n = s->nf;
        auto orig = fetch(s->orig[s->cO], n, st);
        if (f == "position" || f == "density" || f == "pressure") {
            auto h = fetch(s->P[s->cP], n, st);
            const bool pos = f == "position";
            std::vector<double> v(pos ? 3 * n : n);
            for (size_t i = 0; i < n; i++) {
                const size_t o = orig[i];
                if (pos) {
                    v[3 * o] = h[i].x; v[3 * o + 1] = h[i].y; v[3 * o + 2] = h[i].z;
                } else {
                    v[o] = f == "density" ? h[i].rho : h[i].p;
                }
            }
            return emit(v, out, capacity, count);
        }
        if (f == "velocity") {
            auto hx = fetch(s->vx[s->cV], n, st), hy = fetch(s->vy[s->cV], n, st), hz = fetch(s->vz[s->cV], n, st);
            std::vector<double> v(3 * n);
            for (size_t i = 0; i < n; i++) {
                const size_t o = orig[i];
                v[3 * o] = hx[i]; v[3 * o + 1] = hy[i]; v[3 * o + 2] = hz[i];
            }
            return emit(v, out, capacity, count);
        }
        if (f == "densityChange") {
            auto h = fetch(s->drho[s->cD], n, st);
            std::vector<double> v(n);
            for (size_t i = 0; i < n; i++) v[orig[i]] = h[i];
            return emit(v, out, capacity, count);
        }
        if (f == "dummyNormal") {
            const size_t m = s->nd;
            auto dor = fetch(s->dorig, m, st);
            auto hp = fetch(s->DP, m, st);
            auto hv = fetch(s->DV, m, st);
            std::vector<double> v(3 * m);
            for (size_t i = 0; i < m; i++) {
                v[3 * dor[i]] = hp[i].nx; v[3 * dor[i] + 1] = hp[i].ny; v[3 * dor[i] + 2] = hv[i].w;
            }
            return emit(v, out, capacity, count);
        }
        if (f == "hashValue") {
            auto k = fetch(s->skey, n, st);
            std::vector<int> v(k.begin(), k.end());
            return emit(v, out, capacity, count);
        }
        if (f == "hashIndex") return emit(orig, out, capacity, count);
        if (f == "cellHashStart" || f == "cellHashEnd") {
            auto c = fetch(s->cells, (size_t)s->numGrids, st);
            std::vector<int> v(c.size());
            for (size_t k = 0; k < c.size(); k++) v[k] = f == "cellHashStart" ? c[k].x : c[k].y;
            return emit(v, out, capacity, count);
        }
        const bool ff = (f == "neighborCount" || f == "objectPointing");
        const bool fd = (f == "dummyNeighborCount" || f == "dummyObjectPointing");
        if (ff || fd) {
            if (!s->lists_valid) throw InvalidArg("no neighbour lists yet");
            const int K = ff ? s->K_ff : s->K_fd;
            auto cnt = fetch(ff ? s->cnt_ff : s->cnt_fd, n, st);
            std::vector<int> byo(n);
            for (size_t i = 0; i < n; i++) byo[orig[i]] = cnt[i];
            if (f == "neighborCount" || f == "dummyNeighborCount") return emit(byo, out, capacity, count);
            auto nbr = fetch(ff ? s->nbr_ff : s->nbr_fd, (size_t)K * n, st);
            std::vector<int> rankh(n);
            for (size_t i = 0; i < n; i++) rankh[orig[i]] = (int)i;
            std::vector<int> dor;
            if (fd) dor = fetch(s->dorig, (size_t)s->nd, st);
            std::vector<int> v;
            for (size_t o = 0; o < n; o++) {
                const int i = rankh[o];
                const int c = std::min(cnt[i], K);
                for (int k = 0; k < c; k++) {
                    int j = nbr[(((size_t)(k / kChunk) * n + i) * kChunk) + (k % kChunk)];
                    if (ff && s->lists_coded) j &= kSlotMask;
                    v.push_back(ff ? orig[j] : dor[j]);
                }
            }
            return emit(v, out, capacity, count);
        }
        throw InvalidArg("unknown field '" + f + "'");
    });
}

// outputWCSPHVTU (WCSPHSolver.h:360-443) and outputDummyVTU (:445-520):
// particles in insertion order with smoothing length, density, pressure and
// velocity (fluid) or smoothing length, velocity and normal (dummies).
extern "C" int demsph_sph_write_vtu(demsph_sph* s, const char* dir, int what, long frame, long step, double time)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        s->require_init();
        if (!dir) throw InvalidArg("null directory");
        if (what & ~(DEMSPH_VTU_SPH | DEMSPH_VTU_DUMMY)) throw InvalidArg("unknown frame kinds");
        if (frame < 0 || step < 0) throw InvalidArg("negative frame or step");
        const auto field = [&](const char* name, size_t n) {
            std::vector<double> v(n);
            size_t got = 0;
            if (demsph_sph_get(s, name, v.data(), n, &got) != DEMSPH_OK || got != n)
                throw std::runtime_error(std::string("reading ") + name + " for a frame failed");
            return v;
        };
        if (what & DEMSPH_VTU_SPH) {
            const size_t n = s->nf;
            std::vector<double> h(n);
            for (size_t i = 0; i < n; i++) h[i] = s->phases[s->fphase_h[i]].h;
            vtu::Writer w(dir, "SPH", frame);
            w.begin(time, step, n, n);
            w.points(field("position", 3 * n));
            w.vertex_cells(n);
            w.open_data("PointData", "smoothLength");
            w.reals("smoothLength", h);
            w.reals("density", field("density", n));
            w.reals("pressure", field("pressure", n));
            w.vectors("velocity", field("velocity", 3 * n));
            w.close_data("PointData");
            w.end();
        }
        if (what & DEMSPH_VTU_DUMMY) {
            const size_t m = s->nd;
            std::vector<double> h(m);
            for (size_t i = 0; i < m; i++) h[i] = s->phases[s->dphase_h[i]].h;
            vtu::Writer w(dir, "dummy", frame);
            w.begin(time, step, m, m);
            w.points(s->hd.pos);
            w.vertex_cells(m);
            w.open_data("PointData", "smoothLength");
            w.reals("smoothLength", h);
            w.vectors("velocity", s->hd.vel);
            w.vectors("normal", m ? field("dummyNormal", 3 * m) : std::vector<double>());
            w.close_data("PointData");
            w.end();
        }
    });
}

// demsph_sph_write_vtu without waiting: the fluid arrays are copied to
// pinned host memory in stream order, and the frame thread restores
// insertion order and writes the file while the scene keeps stepping.  The
// dummy frame (static particles) is written before returning.
extern "C" int demsph_sph_write_vtu_async(demsph_sph* s, const char* dir, int what, long frame, long step,
                                          double time)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        s->require_init();
        if (!dir) throw InvalidArg("null directory");
        if (what & ~(DEMSPH_VTU_SPH | DEMSPH_VTU_DUMMY)) throw InvalidArg("unknown frame kinds");
        if (frame < 0 || step < 0) throw InvalidArg("negative frame or step");
        if (what & DEMSPH_VTU_DUMMY) {
            const int rc = demsph_sph_write_vtu(s, dir, DEMSPH_VTU_DUMMY, frame, step, time);
            if (rc != DEMSPH_OK) throw std::runtime_error(demsph_last_error());
        }
        if (!(what & DEMSPH_VTU_SPH)) return;
        const size_t n = s->nf;
        hipStream_t st = s->stream;
        auto snap = std::make_shared<PinnedSnap>();
        const FluidPos* P = snap->take(s->P[s->cP], n, st);
        const double* vx = snap->take(s->vx[s->cV], n, st);
        const double* vy = snap->take(s->vy[s->cV], n, st);
        const double* vz = snap->take(s->vz[s->cV], n, st);
        const int* orig = snap->take(s->orig[s->cO], n, st);
        snap->mark(st);
        std::vector<double> h(n);
        for (size_t i = 0; i < n; i++) h[i] = s->phases[s->fphase_h[i]].h;
        vtu::post([snap, P, vx, vy, vz, orig, h = std::move(h), d = std::string(dir), n, frame, step, time] {
            snap->wait();
            std::vector<double> pos(3 * n), vel(3 * n), rho(n), pr(n);
            for (size_t i = 0; i < n; i++) {
                const size_t o = (size_t)orig[i];
                pos[3 * o] = P[i].x; pos[3 * o + 1] = P[i].y; pos[3 * o + 2] = P[i].z;
                vel[3 * o] = vx[i]; vel[3 * o + 1] = vy[i]; vel[3 * o + 2] = vz[i];
                rho[o] = P[i].rho;
                pr[o] = P[i].p;
            }
            vtu::Writer w(d, "SPH", frame);
            w.begin(time, step, n, n);
            w.points(pos);
            w.vertex_cells(n);
            w.open_data("PointData", "smoothLength");
            w.reals("smoothLength", h);
            w.reals("density", rho);
            w.reals("pressure", pr);
            w.vectors("velocity", vel);
            w.close_data("PointData");
            w.end();
        });
    });
}

// Dynamic fluid state to host memory and back, in the current cell-sorted
// order (the stable sort of the next step depends on it): positions with
// density and pressure, velocities (FP32 record and FP64), density change,
// and the slot <-> insertion maps.  Header: magic, version, nf, nd,
// numGrids (the importing scene must be staged with the same particles).
namespace {
constexpr uint32_t kStateMagic = 0x48505344u;  // "DSPH"
struct StateHeader {
    uint32_t magic, version;
    uint64_t nf, nd, numGrids;
};
size_t state_bytes(size_t n)
{
    return sizeof(StateHeader) + n * (sizeof(FluidPos) + sizeof(float4) + 3 * sizeof(double) + sizeof(float) +
                                      2 * sizeof(int));
}
}  // namespace

extern "C" int demsph_sph_export_state(demsph_sph* s, void* out, size_t capacity, size_t* bytes)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        s->require_init();
        const size_t n = s->nf, total = state_bytes(n);
        if (bytes) *bytes = total;
        if (!out) return;
        if (capacity < total) throw InvalidArg("state buffer too small");
        hipStream_t st = s->stream;
        char* o = static_cast<char*>(out);
        const StateHeader h{kStateMagic, 1u, (uint64_t)n, (uint64_t)s->nd, (uint64_t)s->numGrids};
        std::memcpy(o, &h, sizeof h);
        o += sizeof h;
        const auto get = [&](const void* dev, size_t b) {
            if (b) DEMSPH_HIP(hipMemcpyAsync(o, dev, b, hipMemcpyDeviceToHost, st));
            o += b;
        };
        get(s->P[s->cP], n * sizeof(FluidPos));
        get(s->V[s->cV], n * sizeof(float4));
        get(s->vx[s->cV], n * 8); get(s->vy[s->cV], n * 8); get(s->vz[s->cV], n * 8);
        get(s->drho[s->cD], n * 4);
        get(s->orig[s->cO], n * 4);
        get(s->rank, n * 4);
        DEMSPH_HIP(hipStreamSynchronize(st));
    });
}

extern "C" int demsph_sph_import_state(demsph_sph* s, const void* in, size_t bytes)
{
    return guarded([&] {
        if (!s || !in) throw InvalidArg("null argument");
        s->require_init();
        StateHeader h;
        if (bytes < sizeof h) throw InvalidArg("state too short");
        std::memcpy(&h, in, sizeof h);
        if (h.magic != kStateMagic || h.version != 1u) throw InvalidArg("not a fluid scene state");
        if (h.nf != (uint64_t)s->nf || h.nd != (uint64_t)s->nd || h.numGrids != (uint64_t)s->numGrids)
            throw InvalidArg("the state belongs to a scene of other particle counts or grid");
        const size_t n = s->nf;
        if (bytes != state_bytes(n)) throw InvalidArg("state size does not match the scene");
        hipStream_t st = s->stream;
        DEMSPH_HIP(hipStreamSynchronize(st));
        const char* p = static_cast<const char*>(in) + sizeof h;
        const auto put = [&](void* dev, size_t b) {
            if (b) DEMSPH_HIP(hipMemcpyAsync(dev, p, b, hipMemcpyHostToDevice, st));
            p += b;
        };
        put(s->P[s->cP], n * sizeof(FluidPos));
        put(s->V[s->cV], n * sizeof(float4));
        put(s->vx[s->cV], n * 8); put(s->vy[s->cV], n * 8); put(s->vz[s->cV], n * 8);
        put(s->drho[s->cD], n * 4);
        put(s->orig[s->cO], n * 4);
        put(s->rank, n * 4);
        DEMSPH_HIP(hipStreamSynchronize(st));
        s->lists_valid = false;  // the lists belong to the replaced state
        s->cells_track_skey = false;
    });
}

extern "C" int demsph_frames_wait(void)
{
    return guarded([&] {
        const std::string e = vtu::drain();
        if (!e.empty()) throw std::runtime_error("frame output: " + e);
    });
}

extern "C" int demsph_sph_set(demsph_sph* s, const char* field, const void* in, size_t count)
{
    return guarded([&] {
        if (!s || !field || !in) throw InvalidArg("null argument");
        s->require_init();
        const std::string f = field;
        const size_t n = s->nf;
        hipStream_t st = s->stream;
        auto orig = fetch(s->orig[s->cO], n, st);  // slot -> insertion index
        const double* src = static_cast<const double*>(in);
        if (f == "position") {
            if (count != 3 * n) throw InvalidArg("expected 3*numFluid values");
            auto h = fetch(s->P[s->cP], n, st);
            for (size_t i = 0; i < n; i++) {
                const size_t o = orig[i];
                h[i].x = src[3 * o]; h[i].y = src[3 * o + 1]; h[i].z = src[3 * o + 2];
            }
            copy_sync(s->P[s->cP], h.data(), n * sizeof(FluidPos), hipMemcpyHostToDevice, s->stream);
            s->lists_valid = false;
            return;
        }
        if (f == "velocity") {
            if (count != 3 * n) throw InvalidArg("expected 3*numFluid values");
            auto hv = fetch(s->V[s->cV], n, st);
            std::vector<double> hx(n), hy(n), hz(n);
            for (size_t i = 0; i < n; i++) {
                const size_t o = orig[i];
                hx[i] = src[3 * o]; hy[i] = src[3 * o + 1]; hz[i] = src[3 * o + 2];
                hv[i].x = (float)hx[i]; hv[i].y = (float)hy[i]; hv[i].z = (float)hz[i];
            }
            const int b = s->cV;
            copy_sync(s->vx[b], hx.data(), n * 8, hipMemcpyHostToDevice, s->stream);
            copy_sync(s->vy[b], hy.data(), n * 8, hipMemcpyHostToDevice, s->stream);
            copy_sync(s->vz[b], hz.data(), n * 8, hipMemcpyHostToDevice, s->stream);
            copy_sync(s->V[b], hv.data(), n * sizeof(float4), hipMemcpyHostToDevice, s->stream);
            return;
        }
        if (f == "densityChange") {
            if (count != n) throw InvalidArg("expected numFluid values");
            std::vector<float> h(n);
            for (size_t i = 0; i < n; i++) h[i] = (float)src[orig[i]];
            copy_sync(s->drho[s->cD], h.data(), n * 4, hipMemcpyHostToDevice, s->stream);
            s->refresh_eos();
            DEMSPH_HIP(hipStreamSynchronize(st));
            return;
        }
        throw InvalidArg("unknown field '" + f + "'");
    });
}

extern "C" int demsph_sph_enable_timing(demsph_sph* s, int enable)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        if (s->stream) DEMSPH_HIP(hipStreamSynchronize(s->stream));
        s->timer.harvest();
        s->timer.enabled = enable != 0;
        s->timer.reset();
    });
}

extern "C" int demsph_sph_set_graphs(demsph_sph* s, int enable)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        s->use_graphs = enable != 0;
        if (!s->use_graphs) s->drop_graphs();
    });
}

extern "C" int demsph_sph_tune(demsph_sph* s, const char* knob, int value)
{
    return guarded([&] {
        if (!s || !knob) throw InvalidArg("null argument");
        const std::string k = knob;
        if (k == "listPrefilter") {
            if (value != 0 && value != 1) throw InvalidArg("listPrefilter must be 0 or 1");
            s->drop_graphs();
            s->prefilter = value != 0;
            return;
        }
        if (k == "pairRecords") {
            if (value != 0 && value != 1) throw InvalidArg("pairRecords must be 0 or 1");
            s->drop_graphs();
            s->pair_records = value != 0;
            return;
        }
        if (k == "pairLanes" || k == "listLanes") {
            if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
                throw InvalidArg(k + " must be 0 (automatic), 1, 2, 4 or 8");
            s->drop_graphs();
            (k == "pairLanes" ? s->pair_lanes_set : s->list_lanes_set) = value;
            return;
        }
        throw InvalidArg("unknown tuning knob '" + k + "'");
    });
}

extern "C" int demsph_sph_tune_get(demsph_sph* s, const char* knob, int* value)
{
    return guarded([&] {
        if (!s || !knob || !value) throw InvalidArg("null argument");
        const std::string k = knob;
        if (k == "pairLanes" || k == "listLanes") {
            *value = k == "pairLanes" ? s->pair_lanes() : s->list_lanes();
            return;
        }
        if (k == "listPrefilter") {
            *value = s->prefilter ? 1 : 0;
            return;
        }
        if (k == "pairRecords") {
            // 1 when the gather records are in use (the knob and the scene allow them)
            *value = s->records_usable() ? 1 : 0;
            return;
        }
        throw InvalidArg("unknown tuning knob '" + k + "'");
    });
}

extern "C" int demsph_sph_autotune(demsph_sph* s, double dt, int trialSteps, int mode, int useCache,
                                   demsph_sph_tune_report* report)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        const auto o = s->autotune(dt, trialSteps, mode, useCache != 0);
        if (!report) return;
        report->listLanes = o.chosen.list_lanes;
        report->listPrefilter = o.chosen.prefilter;
        report->pairLanes = o.chosen.pair_lanes;
        report->rulesListLanes = o.rules.list_lanes;
        report->rulesPairLanes = o.rules.pair_lanes;
        report->configs = o.configs;
        report->cached = o.cached ? 1 : 0;
        report->listUs = o.list_us;
        report->pairUs = o.pair_us;
        report->rulesListUs = o.rules_list_us;
        report->rulesPairUs = o.rules_pair_us;
    });
}

extern "C" int demsph_sph_autotune_clear_cache(void)
{
    return guarded([&] {
        {
            std::lock_guard<std::mutex> lk(demsph_sph::tune_mutex());
            demsph_sph::tune_cache().clear();
        }
        dem_autotune_clear_cache();
    });
}

extern "C" int demsph_sph_graph_captures(demsph_sph* s, long* count)
{
    return guarded([&] {
        if (!s || !count) throw InvalidArg("null argument");
        *count = s->graph_captures;
    });
}

extern "C" int demsph_sph_timing(demsph_sph* s, int index, char* name, size_t nameCap, long* launches, double* totalMs)
{
    return guarded([&] {
        if (!s) throw InvalidArg("null scene");
        if (s->stream) DEMSPH_HIP(hipStreamSynchronize(s->stream));
        s->timer.harvest();
        const auto& t = s->timer.totals();
        if (index < 0 || (size_t)index >= t.size()) throw InvalidArg("timing index out of range");
        auto it = t.begin();
        std::advance(it, index);
        if (name && nameCap) {
            std::strncpy(name, it->first.c_str(), nameCap - 1);
            name[nameCap - 1] = 0;
        }
        if (launches) *launches = it->second.first;
        if (totalMs) *totalMs = it->second.second;
    });
}

// ============================================================= coupling
// DEM-SPH coupling: one fluid scene and one DEM scene advanced together on
// the fluid scene's stream.  Each coupled step interleaves the two KDK
// sequences so that both force evaluations see the same state:
//   DEM step begin (pair rebuild every 5th) | fluid search | balls onto the
//   fluid grid, fluid->ball lists, reverse index, ball pressures
//   DEM forces + coupling reaction (start state) | DEM hook, kick, drift
//   fluid momentum + kick + drift (start state)
//   ball kinematics refreshed | fluid continuity | ball pressures
//   DEM forces + coupling reaction (end state) | DEM hook, kick
//   fluid momentum + kick (end state)
struct demsph_couple {
    demsph_sph* f = nullptr;
    demsph_dem* d = nullptr;
    bool ready = false;
    int ns = 0;
    static constexpr int kDefaultKfs = 32;
    int K_fs = kDefaultKfs;
    uint32_t *key = nullptr, *skey = nullptr, *ktmp = nullptr;
    int *perm = nullptr, *vtmp = nullptr;  // perm: ball slot -> DEM slot
    void* sort_temp = nullptr;
    int2* cells = nullptr;
    SolidPos* SP = nullptr;
    float4 *SV = nullptr, *SW = nullptr;
    int *nbr_fs = nullptr, *cnt_fs = nullptr;
    int *rcnt = nullptr, *rstart = nullptr, *rfluid = nullptr;
    void* scan_temp = nullptr;
    double* Fc = nullptr;  // 6 per ball slot (force, torque) of the last evaluation
    float* acs = nullptr;  // 3 per fluid slot, coupling acceleration of the last fluid pass
    long replays = 0;      // coupled blocks run again after a list outgrew its capacity

    // The DEM scene runs on the fluid's stream while coupled; it gets its
    // own stream back when the coupling ends, so that it never keeps a
    // handle to a stream the fluid scene destroys.
    hipStream_t dem_own_stream = nullptr;
    bool dem_switched = false;

    ~demsph_couple()
    {
        detach_fluid();
        if (d) {
            if (dem_switched) dem_set_stream(d, dem_own_stream);
            dem_detach_couple(d, this);
        }
    }

    // The fluid scene is going away (or this coupling is): free what lives
    // on its stream and forget it.
    void detach_fluid()
    {
        if (!f) return;
        release();
        f->couples.erase(std::remove(f->couples.begin(), f->couples.end(), this), f->couples.end());
        if (d && dem_switched) {
            dem_set_stream(d, dem_own_stream);
            dem_switched = false;
        }
        f = nullptr;
    }

    void require_scenes() const
    {
        if (!f || !d) throw InvalidArg("a scene of this coupling was destroyed");
    }

    void release()
    {
        if (f && f->stream) (void)hipStreamSynchronize(f->stream);
        dfree(key); dfree(skey); dfree(ktmp); dfree(perm); dfree(vtmp);
        dev_free(sort_temp);
        sort_temp = nullptr;
        dfree(cells); dfree(SP); dfree(SV); dfree(SW); dfree(nbr_fs); dfree(cnt_fs);
        dfree(rcnt); dfree(rstart); dfree(rfluid);
        dev_free(scan_temp);
        scan_temp = nullptr;
        dfree(Fc); dfree(acs);
        ready = false;
    }

    void setup()
    {
        require_scenes();
        f->require_init();
        dem_require_ready(d);
        if (!dem_switched) {
            dem_own_stream = dem_stream(d);
            if (dem_own_stream != f->stream) DEMSPH_HIP(hipStreamSynchronize(dem_own_stream));
            dem_switched = true;
        }
        dem_set_stream(d, f->stream);
        const DemBallRefs b = dem_ball_refs(d);
        if (b.n == 0) throw InvalidArg("the DEM scene has no balls");
        if (ready && b.n == ns) return;
        release();
        ns = b.n;
        K_fs = demsph_sph::round_chunk(K_fs);
        const size_t nf = f->nf;
        key = dmalloc<uint32_t>(ns); skey = dmalloc<uint32_t>(ns); ktmp = dmalloc<uint32_t>(ns);
        perm = dmalloc<int>(ns); vtmp = dmalloc<int>(ns);
        DEMSPH_HIP(dev_malloc(&sort_temp, radix_sort_temp_bytes(ns)));
        cells = dmalloc<int2>(f->numGrids);
        SP = dmalloc<SolidPos>(ns); SV = dmalloc<float4>(ns); SW = dmalloc<float4>(ns);
        nbr_fs = dmalloc<int>((size_t)K_fs * nf); cnt_fs = dmalloc<int>(nf);
        rcnt = dmalloc<int>(ns + 1); rstart = dmalloc<int>(ns + 1);
        rfluid = dmalloc<int>(rcap());
        DEMSPH_HIP(dev_malloc(&scan_temp, scan_temp_bytes(ns + 1)));
        Fc = dmalloc<double>(6 * (size_t)ns);
        acs = dmalloc<float>(3 * nf);
        DEMSPH_HIP(hipMemsetAsync(Fc, 0, 6 * (size_t)ns * sizeof(double), f->stream));
        DEMSPH_HIP(hipMemsetAsync(acs, 0, 3 * nf * sizeof(float), f->stream));
        DEMSPH_HIP(hipMemsetAsync(SP, 0, ns * sizeof(SolidPos), f->stream));
        // size the fluid lists before the first block (a coupled block cannot be replayed)
        f->neighbor_search();
        ready = true;
    }

    SolidView view() const
    {
        SolidView v;
        v.ns = ns; v.SP = SP; v.SV = SV; v.SW = SW; v.nbr_fs = nbr_fs; v.cnt_fs = cnt_fs; v.K_fs = K_fs; v.acs = acs;
        return v;
    }

    void gather(hipStream_t st, bool with_cells)
    {
        const DemBallRefs b = dem_ball_refs(d);
        SolidGatherArgs g{};
        g.ns = ns; g.numGrids = f->numGrids; g.skey = skey; g.s2d = perm;
        g.px = b.px; g.py = b.py; g.pz = b.pz; g.vx = b.vx; g.vy = b.vy; g.vz = b.vz;
        g.wx = b.wx; g.wy = b.wy; g.wz = b.wz; g.r = b.r;
        g.SP = SP; g.SV = SV; g.SW = SW; g.cells = with_cells ? cells : nullptr;
        DEMSPH_LAUNCH(f->timer, "couple_gather", k_solid_gather, dim3(blocks_for(ns, 256)), dim3(256), 0, st, g);
    }

    // The reverse lists hold the same pairs as the fluid lists, at most
    // K_fs per fluid particle.
    int rcap() const { return (int)std::min<size_t>((size_t)K_fs * (size_t)f->nf, (size_t)INT32_MAX); }

    void pressure(hipStream_t st)
    {
        DEMSPH_LAUNCH(f->timer, "couple_pressure", k_solid_pressure, dim3(blocks_for(ns, kBallsPerBlock)),
                      dim3(64 * kBallsPerBlock), 0, st, ns, rstart, rfluid, f->P[f->cP], f->V[f->cV], f->d_phases, SP,
                      rcap());
    }

    // Balls onto the fluid grid, fluid->ball lists and their reverse index,
    // ball pressures; at the start of a step, after the fluid search.
    void prepare(hipStream_t st)
    {
        const DemBallRefs b = dem_ball_refs(d);
        const int nf = f->nf;
        timer_begin(st, "couple_sort");
        hipLaunchKernelGGL(k_solid_hash, dim3(blocks_for(ns, 256)), dim3(256), 0, st, ns, f->grid, b.px, b.py, b.pz,
                           key);
        DEMSPH_HIP(hipGetLastError());
        radix_sort_pairs(key, nullptr, skey, perm, ktmp, vtmp, sort_temp, ns, key_bits(f->numGrids), st);
        fill_i32(reinterpret_cast<int*>(cells), 2 * f->numGrids, -1, st);
        f->timer.end(st);
        gather(st, true);
        ListArgs la{};
        la.nf = nf; la.nd = f->nd; la.grid = f->grid; la.phases = f->d_phases; la.hf = f->phf.h; la.hd = f->phd.h;
        la.P = f->P[f->cP]; la.V = f->V[f->cV];
        la.scells = cells; la.SP = SP; la.nbr_fs = nbr_fs; la.cnt_fs = cnt_fs; la.K_fs = K_fs; la.maxcount = f->maxcount;
        DEMSPH_LAUNCH(f->timer, "couple_lists", k_build_fs_lists, dim3(blocks_for(nf, kThreads)), dim3(kThreads), 0,
                      st, la);
        timer_begin(st, "couple_reverse");
        RevArgs ra{};
        ra.ns = ns; ra.grid = f->grid; ra.cells = f->cells; ra.P = f->P[f->cP]; ra.V = f->V[f->cV];
        ra.phases = f->d_phases; ra.SP = SP; ra.rcnt = rcnt; ra.rstart = rstart; ra.rfluid = rfluid; ra.rcap = rcap();
        ra.hf = f->phf.h;
        // the fluid's FP32 copy is current when the search wrote it
        ra.Pf = f->prefilter ? f->Pf : nullptr;
        ra.band = demsph_sph::kPrefilterBand;
        ra.R = f->pref_R;
        const dim3 wb(blocks_for(ns, kBallsPerBlock)), wt(64 * kBallsPerBlock);
        // one fluid phase: every fluid particle has the same support
        const bool unif = f->fluid_uniform();
        if (unif) hipLaunchKernelGGL((k_fs_rev_walk<true, true>), wb, wt, 0, st, ra);
        else hipLaunchKernelGGL((k_fs_rev_walk<true, false>), wb, wt, 0, st, ra);
        DEMSPH_HIP(hipGetLastError());
        exclusive_scan(rcnt, rstart, ns, scan_temp, rstart + ns, st);
        if (unif) hipLaunchKernelGGL((k_fs_rev_walk<false, true>), wb, wt, 0, st, ra);
        else hipLaunchKernelGGL((k_fs_rev_walk<false, false>), wb, wt, 0, st, ra);
        DEMSPH_HIP(hipGetLastError());
        f->timer.end(st);
        pressure(st);
    }

    void timer_begin(hipStream_t st, const char* name) { f->timer.begin(name, st); }

    void forces(hipStream_t st)
    {
        const DemBallRefs b = dem_ball_refs(d);
        CoupleForceArgs c{};
        c.a = f->args();
        c.rstart = rstart; c.rfluid = rfluid; c.s2d = perm; c.rcap = rcap();
        c.fx = b.fx; c.fy = b.fy; c.fz = b.fz; c.tx = b.tx; c.ty = b.ty; c.tz = b.tz;
        c.Fc = Fc;
        DEMSPH_LAUNCH(f->timer, "couple_forces", k_couple_forces, dim3(blocks_for(ns, kBallsPerBlock)),
                      dim3(64 * kBallsPerBlock), 0, st, c);
    }

    // A block of coupled steps.  Both scenes are checkpointed on the device
    // first; when a fluid list, the fluid-ball list or the DEM wall-contact
    // capacity overflows, it grows, both scenes go back to the checkpoint and
    // the block runs again, so the results are those of ample capacities.
    void step(const double* g, double dt, int nsteps)
    {
        if (!(dt > 0.0)) throw InvalidArg("time step must be positive");
        setup();
        if (nsteps <= 0) return;
        const double h = 0.5 * dt;
        hipStream_t st = f->stream;
        f->ensure_checkpoint();
        for (int attempt = 0; attempt < 8; attempt++) {
            f->checkpoint(true);
            const int sP = f->cP, sV = f->cV, sD = f->cD, sO = f->cO;
            dem_block_save(d);
            f->reset_overflow();
            f->solids = view();
            try {
                for (int k = 0; k < nsteps; k++) {
                    const double t = dem_begin_step(d, dt);
                    f->search_enqueue();
                    prepare(st);
                    dem_evaluate(d, h, dt);
                    forces(st);
                    dem_hooks_and_half(d, true, g, h, t - h);
                    f->first_half_enqueue(dt);
                    gather(st, false);
                    f->continuity_enqueue(dt);
                    pressure(st);
                    dem_evaluate(d, h, dt);
                    forces(st);
                    dem_hooks_and_half(d, false, g, h, t);
                    f->momentum2_enqueue(dt);
                    step_stamp(f->stamps, st);
                }
            } catch (...) {
                f->solids = SolidView();
                throw;
            }
            f->solids = SolidView();
            const bool grew = f->check_overflow();  // reads all three list maxima, grows the fluid lists
            const int mfs = f->maxcount_h[2];
            const int wneed = dem_take_wall_overflow(d);  // ball-wall candidates above the DEM capacity
            if (!grew && mfs <= K_fs && wneed == 0) return;
            if (mfs > K_fs) {
                K_fs = mfs + mfs / 4 + 8;
                ready = false;  // setup() reallocates the fluid-ball lists
            }
            f->cP = sP; f->cV = sV; f->cD = sD; f->cO = sO;
            f->checkpoint(false);
            dem_block_restore(d, 0);
            if (wneed > 0) dem_grow_walls(d, wneed);  // above the maximum: an error, both scenes at the checkpoint
            replays++;
            setup();
        }
        throw std::runtime_error("coupled list capacities did not converge");
    }
};

demsph_sph::~demsph_sph()
{
    // couplings destroyed after this scene must not touch it
    for (auto* c : std::vector<demsph_couple*>(couples)) c->detach_fluid();
    release();
}

namespace demsph {
void couple_detach_dem(demsph_couple* c)
{
    if (c->f && c->f->stream) (void)hipStreamSynchronize(c->f->stream);
    if (c->dem_switched) c->dem_switched = false;  // the scene's stream is its owner's business now
    c->d = nullptr;
    c->ready = false;
}
}  // namespace demsph

extern "C" int demsph_couple_create(demsph_couple** out, demsph_sph* fluid, demsph_dem* dem)
{
    return guarded([&] {
        if (!out || !fluid || !dem) throw InvalidArg("null argument");
        auto* c = new demsph_couple();
        c->f = fluid;
        c->d = dem;
        fluid->couples.push_back(c);
        dem_attach_couple(dem, c);
        *out = c;
    });
}

extern "C" int demsph_couple_destroy(demsph_couple* c)
{
    return guarded([&] { delete c; });
}

extern "C" int demsph_couple_set_capacity(demsph_couple* c, int ballsPerFluid)
{
    return guarded([&] {
        if (!c) throw InvalidArg("null coupling");
        if (ballsPerFluid < 1) throw InvalidArg("capacity must be positive");
        c->K_fs = ballsPerFluid;
        c->ready = false;
    });
}

extern "C" int demsph_couple_step(demsph_couple* c, const double* gravity, double timeStep, int nsteps)
{
    return guarded([&] {
        if (!c || !gravity) throw InvalidArg("null argument");
        c->step(gravity, timeStep, nsteps);
    });
}

extern "C" int demsph_couple_get(demsph_couple* c, const char* field, void* out, size_t capacity, size_t* count)
{
    return guarded([&] {
        if (!c || !field) throw InvalidArg("null argument");
        c->require_scenes();
        if (!c->ready) throw InvalidArg("no coupled step has run");
        const std::string fld = field;
        hipStream_t st = c->f->stream;
        if (fld == "couplingForce" || fld == "couplingTorque" || fld == "ballPressure") {
            const DemBallRefs b = dem_ball_refs(c->d);
            auto perm = fetch(c->perm, (size_t)c->ns, st);
            auto orig = fetch(b.orig, (size_t)b.n, st);
            const bool p = fld == "ballPressure";
            std::vector<double> v(p ? c->ns : 3 * (size_t)c->ns);
            if (p) {
                auto sp = fetch(c->SP, (size_t)c->ns, st);
                for (int k = 0; k < c->ns; k++) v[orig[perm[k]]] = sp[k].P;
            } else {
                auto fc = fetch(c->Fc, 6 * (size_t)c->ns, st);
                const int off = fld == "couplingForce" ? 0 : 3;
                for (int k = 0; k < c->ns; k++)
                    for (int q = 0; q < 3; q++) v[3 * (size_t)orig[perm[k]] + q] = fc[6 * (size_t)k + off + q];
            }
            return emit(v, out, capacity, count);
        }
        if (fld == "blockReplays") {  // coupled blocks run again after a list outgrew its capacity
            const std::vector<double> v{(double)c->replays};
            return emit(v, out, capacity, count);
        }
        if (fld == "fluidCouplingAcceleration") {
            const size_t n = c->f->nf;
            auto a = fetch(c->acs, 3 * n, st);
            auto orig = fetch(c->f->orig[c->f->cO], n, st);
            std::vector<double> v(3 * n);
            for (size_t i = 0; i < n; i++)
                for (int q = 0; q < 3; q++) v[3 * (size_t)orig[i] + q] = a[3 * i + q];
            return emit(v, out, capacity, count);
        }
        throw InvalidArg("unknown field '" + fld + "'");
    });
}

// ------------------------------------------------------------ internal API
namespace demsph {

size_t sph_bytes_estimate(demsph_sph* s, const double* minb, const double* maxb)
{
    return s->bytes_estimate(minb, maxb);
}
int sph_num_fluid(demsph_sph* s) { return s->initialized ? s->nf : (int)s->hf.size(); }
void sph_set_stamps(demsph_sph* s, const StepStamps& st) { s->set_stamps(st); }
hipStream_t sph_stream(demsph_sph* s) { return s->stream; }

// Fluid-ball lists and their reverse index at the default capacity, ball
// records and a fluid-grid cell array, as demsph_couple::setup allocates.
size_t couple_bytes_estimate(demsph_sph* f, demsph_dem* d, const double* fluidMin, const double* fluidMax)
{
    const size_t nf = (size_t)sph_num_fluid(f), ns = (size_t)dem_num_balls(d);
    GridGeom gg{};
    const size_t cells = grid_set(gg, fluidMin, fluidMax, f->staged_cell_size()) ? (size_t)gg.numGrids() : 0;
    const int K_fs = demsph_couple::kDefaultKfs;
    return nf * (2 * sizeof(int) * (size_t)K_fs + sizeof(int) + 3 * sizeof(float)) +
           ns * (5 * 4 + sizeof(SolidPos) + 2 * sizeof(float4) + 3 * sizeof(int) + 6 * sizeof(double)) +
           cells * sizeof(int2) + radix_sort_temp_bytes((int)std::max<size_t>(ns, 1)) + (size_t(1) << 20);
}

}  // namespace demsph
