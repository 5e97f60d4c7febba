// DEM scene for MI355X: balls in cell-sorted SoA order (reordered at each
// pair-list rebuild), CSR half pair lists per owner with spring history
// carried by insertion indices, a counting-sort reverse index for the
// partner-side reduction, per-bond state, per-ball wall contacts, rigid
// clumps and mesh walls, and the external-force hook.
//
// Reference: DEMSPH/kernel/DEMSolver.h (sequencing :428-906, :1376-1485,
// :1961-2049), ballNeighborSearchKernel.cu, contactKernel.cu(h),
// ballIntegrationKernel.cu, wallIntegrationKernel.cu(h), and the hook
// kernels of DEMSPH/externalForceTorque.cu.
//
// Precision: positions, velocities, angular velocities, overlaps, contact
// geometry, spring history, bond state, clump and wall state are FP64; the
// contact force model runs in FP32 (default) or FP64, selected per scene.
#include "common.hpp"
#include "dem_kernels.hpp"
#include "hook_math.hpp"
#include "internal.hpp"
#include "timing.hpp"
#include "vtu.hpp"

#include <memory>

#include <algorithm>
#include <array>
#include <cmath>
#include <chrono>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <set>
#include <utility>
#include <vector>

namespace demsph {
namespace {

using namespace dem;
// Threads per block of the per-ball and per-pair kernels.  Their threads
// walk short dependent chains (pairs, candidate triangles), so spreading a
// scene over more CUs beats fuller blocks: 64 against 256 threads, 46.8 vs
// 52.4 us per step for 4,096 balls and 167 vs 173 us for 112,768 (results
// identical).
constexpr int kT = 64;

struct Balls {
    int n;
    double *px, *py, *pz, *vx, *vy, *vz, *wx, *wy, *wz, *r, *im;
    int *mat, *clump, *orig;
};

struct Forces {
    double *fx, *fy, *fz, *tx, *ty, *tz;
};

struct Pairs {
    int* start;    // n + 1 (owner CSR)
    int cap;       // entries allocated; a list longer than cap is truncated (and replayed)
    int* partner;  // cap
    int* owner;    // cap, the owner slot of each entry (written with the list)
    uint8_t* shift;  // cap, periodic image code of the partner (periodic scenes only, else null)
    double *sx, *sy, *sz, *rx, *ry, *rz, *ox, *oy, *oz;  // sliding, rolling, torsion springs
    void *Fx, *Fy, *Fz, *Rx, *Ry, *Rz;                   // pair force, partner-side torque (precision T)
    void *Ix, *Iy, *Iz;                                  // owner-side torque (precision T)
    double* ov;                                           // overlap
};

__device__ __forceinline__ D3 ld3(const double* x, const double* y, const double* z, int i) { return {x[i], y[i], z[i]}; }
__device__ __forceinline__ void st3(double* x, double* y, double* z, int i, D3 v) { x[i] = v.x; y[i] = v.y; z[i] = v.z; }

// ----------------------------------------------------------------- periodic boundaries
// setPeriodicBoundary (DEMSolver.h:1944-1959) with the intended semantics:
// on a periodic axis the 27-cell stencil wraps around the domain and the
// balls of the first and last cell layers meet through periodic images, each
// contact listed once (the half list by insertion index of the ball pairs).
// A pair stores the image of its partner as a code (sx+1) + 3 (sy+1) +
// 9 (sz+1), s = -1, 0, +1 periods per axis; kNoShift = the ball itself.
constexpr int kNoShift = 13;
struct Periodic {
    int on[3];   // periodic axes
    double L[3]; // periods (max - min of the grid)
};
__device__ __forceinline__ double shifted(double x, int s, double L) { return s == 0 ? x : (s > 0 ? x + L : x - L); }
__device__ __forceinline__ D3 shift_point(const D3& p, int code, const Periodic& per)
{
    if (code == kNoShift) return p;
    return D3{shifted(p.x, code % 3 - 1, per.L[0]), shifted(p.y, code / 3 % 3 - 1, per.L[1]),
              shifted(p.z, code / 9 - 1, per.L[2])};
}
// Image of rj nearest to ri on the periodic axes (bonds across a face).
__device__ __forceinline__ int nearest_image(const D3& ri, const D3& rj, const Periodic& per)
{
    const double d[3] = {ri.x - rj.x, ri.y - rj.y, ri.z - rj.z};
    int s[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (!per.on[k]) continue;
        if (d[k] > 0.5 * per.L[k]) s[k] = 1;
        else if (d[k] < -0.5 * per.L[k]) s[k] = -1;
    }
    return (s[0] + 1) + 3 * (s[1] + 1) + 9 * (s[2] + 1);
}

// ----------------------------------------------------------------- rebuild
struct ReorderArgs {
    Balls src, dst;
    const uint32_t* skey;
    const int* perm;
    int* rank;
    int* rank_old;  // insertion index -> slot before this reorder
    int *rcnt, *rcursor;  // reverse-index counters, cleared here for the build after the pair write
    long long numGrids;
    int *cellStart, *cellEnd;
};

__global__ __launch_bounds__(kT) void k_reorder(ReorderArgs a)
{
    const int k = blockIdx.x * kT + threadIdx.x;
    if (k >= a.src.n) return;
    const int o = a.perm[k];
    const int s = a.rank[o];
    a.rank_old[o] = s;
    a.rank[o] = k;
    a.rcnt[k] = 0;
    a.rcursor[k] = 0;
    if (k == 0) a.rcnt[a.src.n] = 0;
    a.dst.orig[k] = o;
    a.dst.px[k] = a.src.px[s]; a.dst.py[k] = a.src.py[s]; a.dst.pz[k] = a.src.pz[s];
    a.dst.vx[k] = a.src.vx[s]; a.dst.vy[k] = a.src.vy[s]; a.dst.vz[k] = a.src.vz[s];
    a.dst.wx[k] = a.src.wx[s]; a.dst.wy[k] = a.src.wy[s]; a.dst.wz[k] = a.src.wz[s];
    a.dst.r[k] = a.src.r[s]; a.dst.im[k] = a.src.im[s];
    a.dst.mat[k] = a.src.mat[s]; a.dst.clump[k] = a.src.clump[s];
    const uint32_t h = a.skey[k];
    if ((long long)h < a.numGrids) {
        if (k == 0 || a.skey[k - 1] != h) a.cellStart[h] = k;
        if (k == a.src.n - 1 || a.skey[k + 1] != h) a.cellEnd[h] = k + 1;
    }
}

// Balls that are not clump pebbles are moved back into [min, max) on the
// periodic axes before each pair rebuild (one period at most: a ball moves
// far less than a period between rebuilds).  A position just below the
// maximum face can round into the cell past the last one,
// (x - min) / cellSize >= n; such a ball takes its image just below the
// minimum face instead (cell 0), so every wrapped ball hashes inside the grid.
__global__ __launch_bounds__(kT) void k_wrap_periodic(Balls b, GridGeom g, Periodic per)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= b.n || b.clump[i] >= 0) return;
    double* x[3] = {b.px, b.py, b.pz};
    const double lo[3] = {g.minx, g.miny, g.minz}, hi[3] = {g.maxx, g.maxy, g.maxz};
    const double cs[3] = {g.cx, g.cy, g.cz};
    const int nc[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (!per.on[k]) continue;
        double v = x[k][i];
        if (v >= hi[k]) v = v - per.L[k];
        else if (v < lo[k]) v = v + per.L[k];
        if ((int)((v - lo[k]) / cs[k]) >= nc[k]) v = v - per.L[k];
        x[k][i] = v;
    }
}

// Walks the 27-cell stencil of ball A and visits every ball B that the
// reference pairs with A (ballNeighborSearchKernel.cu:6-59): insertion
// index of A below that of B, not both fixed, not one clump, within
// 1.1 (rA + rB).  Visit order = the reference's write order.  PER: the
// stencil wraps on the periodic axes (three offsets each, no face shrink)
// and B is tested and visited as the image it is met through.
template <bool PER, class F>
__device__ __forceinline__ void walk_pair_rows(const Balls& b, const GridGeom& g, const int* cs, const int* ce, int A,
                                               int q0, int q1, const Periodic& per, F&& visit)
{
    const double xa = b.px[A], ya = b.py[A], za = b.pz[A], ra = b.r[A];
    const int oa = b.orig[A];
    const bool fa = b.im[A] == 0.0;
    const int ca = b.clump[A];
    Stencil st = make_stencil(g, xa, ya, za);
    if (PER) {
        if (per.on[0]) { st.lox = -1; st.hix = 1; }
        if (per.on[1]) { st.loy = -1; st.hiy = 1; }
        if (per.on[2]) { st.loz = -1; st.hiz = 1; }
    }
    const int nyr = st.hiy - st.loy + 1;
    const int nrows = (st.hiz - st.loz + 1) * nyr;
    const int qe = min(q1, nrows);
    for (int q = q0; q < qe; q++) {
        int cz = st.cz + st.loz + q / nyr, cy = st.cy + st.loy + q % nyr;
        int sz = 0, sy = 0;
        if (PER) {
            if (cz < 0) { cz += g.nz; sz = -1; } else if (cz >= g.nz) { cz -= g.nz; sz = 1; }
            if (cy < 0) { cy += g.ny; sy = -1; } else if (cy >= g.ny) { cy -= g.ny; sy = 1; }
        }
        const int row = cz * g.ny * g.nx + cy * g.nx;
        for (int xx = st.lox; xx <= st.hix; xx++) {
            int cx = st.cx + xx, sx = 0;
            if (PER) {
                if (cx < 0) { cx += g.nx; sx = -1; } else if (cx >= g.nx) { cx -= g.nx; sx = 1; }
            }
            const int code = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
            const int hb = row + cx;
            const int s0 = cs[hb];
            if (s0 == -1) continue;
            const int s1 = ce[hb];
            for (int j = s0; j < s1; j++) {
                if (oa >= b.orig[j]) continue;
                if (fa && b.im[j] == 0.0) continue;
                const int cb = b.clump[j];
                if (cb >= 0 && ca == cb) continue;
                const double cut = 1.1 * (ra + b.r[j]);
                if (PER && code != kNoShift) {
                    const D3 pj = shift_point(D3{b.px[j], b.py[j], b.pz[j]}, code, per);
                    if (within_cut(cut, xa - pj.x, ya - pj.y, za - pj.z)) visit(j, code);
                } else if (within_cut(cut, xa - b.px[j], ya - b.py[j], za - b.pz[j])) {
                    visit(j, kNoShift);
                }
            }
        }
    }
}

// The pair walk of ball A split over a group of L lanes (L = 1 or 8): each
// lane walks a run of stencil rows ((z, y) in the serial walk's order, x
// cells and slots within a row in order), so the rows of a group, taken in
// lane order, give the serial walk's visit order.
// Up to kPairLanesMaxBalls balls a scene uses 8 lanes per ball (measured on
// MI355X: a 4,096-ball collapse in a box, 48 -> 44 us per step; the rebuild
// kernels 47 -> 22 us); larger scenes have threads enough and use 1 (a
// 112,768-ball bonded beam with clumps, one process: 168 us per step with 1
// lane, 170 with 2, 171 with 4, 173 with 8).  Lane q walks rows [9q/L, 9(q+1)/L).  "pairLanes" overrides it.
constexpr int kPairLanesMaxBalls = 32768;
inline int pair_walk_lanes(int n, int set) { return set > 0 ? set : (n <= kPairLanesMaxBalls ? 8 : 1); }
template <int L>
__device__ __forceinline__ void pair_lane_rows(int q, int& q0, int& q1)
{
    q0 = 9 * q / L;
    q1 = 9 * (q + 1) / L;
}
template <int L>
__device__ __forceinline__ int group_sum(int v)
{
#pragma unroll
    for (int m = L / 2; m > 0; m >>= 1) v += __shfl_xor(v, m, L);
    return v;
}
template <int L>
__device__ __forceinline__ int group_exclusive(int v, int q)
{
    int inc = v;
#pragma unroll
    for (int d = 1; d < L; d <<= 1) {
        const int t = __shfl_up(inc, d, L);
        if (q >= d) inc += t;
    }
    return inc - v;
}

// L lanes per ball (whole groups exit together, so the group shuffles see
// only their own group's lanes).
template <int L, bool PER>
__global__ __launch_bounds__(kT) void k_count_pairs(Balls b, GridGeom g, const int* cs, const int* ce, int* cnt,
                                                    Periodic per)
{
    const int t = xcd_block() * kT + threadIdx.x;
    const int i = t / L, q = t % L;
    if (i >= b.n) return;
    int q0, q1;
    pair_lane_rows<L>(q, q0, q1);
    int c = 0;
    walk_pair_rows<PER>(b, g, cs, ce, i, q0, q1, per, [&](int, int) { c++; });
    c = group_sum<L>(c);
    if (q == 0) cnt[i] = c;
}

struct WritePairsArgs {
    Balls b;
    GridGeom g;
    const int *cs, *ce;
    Pairs cur, old;
    const int* old_orig;  // old slot -> insertion index
    const int* rank_old;  // insertion index -> old slot
    int has_old;
    Periodic per;
};

// writeBallInteractionsKernel (ballNeighborSearchKernel.cu:61-151) with the
// history carried by insertion indices: a new pair (A,B) takes the springs
// of the old pair (A,B) if it existed, else zero.
template <int L, bool PER>
__global__ __launch_bounds__(kT) void k_write_pairs(WritePairsArgs a)
{
    const int t = xcd_block() * kT + threadIdx.x;
    const int i = t / L, q = t % L;
    if (i >= a.b.n) return;
    int q0, q1;
    pair_lane_rows<L>(q, q0, q1);
    // this lane's rows start after the pairs of the group's earlier lanes
    int w = a.cur.start[i];
    if (L > 1) {
        int c = 0;
        walk_pair_rows<PER>(a.b, a.g, a.cs, a.ce, i, q0, q1, a.per, [&](int, int) { c++; });
        w += group_exclusive<L>(c, q);
    }
    const int oa = a.b.orig[i];
    int os0 = 0, os1 = 0;
    if (a.has_old) {
        const int s = a.rank_old[oa];
        os0 = min(a.old.start[s], a.old.cap);
        os1 = min(a.old.start[s + 1], a.old.cap);
    }
    walk_pair_rows<PER>(a.b, a.g, a.cs, a.ce, i, q0, q1, a.per, [&](int j, int code) {
        if (w >= a.cur.cap) {  // over capacity: counted, not stored; the block is replayed
            w++;
            return;
        }
        a.cur.partner[w] = j;
        a.cur.owner[w] = i;
        if (PER) a.cur.shift[w] = (uint8_t)code;
        double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const int ob = a.b.orig[j];
        for (int p = os0; p < os1; p++) {
            if (a.old_orig[a.old.partner[p]] == ob) {
                s[0] = a.old.sx[p]; s[1] = a.old.sy[p]; s[2] = a.old.sz[p];
                s[3] = a.old.rx[p]; s[4] = a.old.ry[p]; s[5] = a.old.rz[p];
                s[6] = a.old.ox[p]; s[7] = a.old.oy[p]; s[8] = a.old.oz[p];
                break;
            }
        }
        a.cur.sx[w] = s[0]; a.cur.sy[w] = s[1]; a.cur.sz[w] = s[2];
        a.cur.rx[w] = s[3]; a.cur.ry[w] = s[4]; a.cur.rz[w] = s[5];
        a.cur.ox[w] = s[6]; a.cur.oy[w] = s[7]; a.cur.oz[w] = s[8];
        w++;
    });
}

// Reverse index: pairs grouped by partner, ascending pair id within each
// partner (the reference sorts pairs by pointing object, DEMSolver.h:59-78).
__global__ __launch_bounds__(kT) void k_rev_count(int n, const int* start, int cap, const int* partner, int* rcnt)
{
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const int p1 = min(start[i + 1], cap);
    for (int p = start[i]; p < p1; p++) atomicAdd(&rcnt[partner[p]], 1);
}
__global__ __launch_bounds__(kT) void k_rev_fill(int n, const int* start, int cap, const int* partner,
                                                 const int* rstart, int* rcursor, int* rpair)
{
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const int p1 = min(start[i + 1], cap);
    for (int p = start[i]; p < p1; p++) {
        const int j = partner[p];
        rpair[rstart[j] + atomicAdd(&rcursor[j], 1)] = p;
    }
}
__global__ __launch_bounds__(kT) void k_rev_sort(int n, const int* rstart, int* rpair)
{
    const int j = blockIdx.x * kT + threadIdx.x;
    if (j >= n) return;
    const int s0 = rstart[j], s1 = rstart[j + 1];
    for (int a = s0 + 1; a < s1; a++) {
        const int v = rpair[a];
        int b = a - 1;
        while (b >= s0 && rpair[b] > v) {
            rpair[b + 1] = rpair[b];
            b--;
        }
        rpair[b + 1] = v;
    }
}

// A pair list longer than its capacity: the needed length, for the replay.
__global__ void k_pair_capacity(const int* total, int cap, int* over)
{
    if (threadIdx.x == 0 && blockIdx.x == 0 && *total > cap) atomicMax(over, *total);
}

// One launch copies every segment of a table (the device checkpoint of a
// block of steps): segment blockIdx.y, 4-byte words, grid-stride in x.
struct CopySeg {
    const void* src;
    void* dst;
    unsigned long long bytes;
};
// The table travels as a kernel argument: 128 segments of 24 bytes plus the
// count and the ~256 bytes of arguments HIP appends stay below the 4 KB
// kernel-argument segment (160 segments made 4,104 bytes).
constexpr int kMaxCopySegs = 128;
struct CopyTable {
    CopySeg s[kMaxCopySegs];
    int n;
};
__global__ __launch_bounds__(256) void k_multi_copy(CopyTable t)
{
    const int seg = blockIdx.y;
    if (seg >= t.n) return;
    const CopySeg c = t.s[seg];
    const size_t words = c.bytes / 4;
    const uint32_t* src = static_cast<const uint32_t*>(c.src);
    uint32_t* dst = static_cast<uint32_t*>(c.dst);
    for (size_t w = (size_t)blockIdx.x * 256 + threadIdx.x; w < words; w += (size_t)gridDim.x * 256) dst[w] = src[w];
    // byte arrays (periodic image codes) may end in a partial word
    if (blockIdx.x == 0 && threadIdx.x < (c.bytes & 3))
        static_cast<unsigned char*>(c.dst)[4 * words + threadIdx.x] =
            static_cast<const unsigned char*>(c.src)[4 * words + threadIdx.x];
}

// ----------------------------------------------------------------- forces
struct PairForceArgs {
    Balls b;
    Pairs p;
    Tables tb;
    double dt;
    // optional record of every pair of the evaluation, 12 doubles per pair
    // slot: contact point, normal, force, torque (without the lever term)
    double* rec;
    Periodic per;  // used when p.shift is set
};

// updateBallContactKernel + calBallContactForceTorqueKernel
// (contactKernel.cu:135-259), one thread per pair entry p of the owner CSR
// (entries past the list's end or capacity return).  Each entry leaves the
// pair force, the owner-side and the partner-side torque; the gather sums
// them per ball in list order (sumObjectPointedForceTorque,
// sumObjectPointingForceTorqueFromInteraction, :657-714).
template <class T>
__device__ __forceinline__ void pair_force_one(const PairForceArgs& a, int p)
{
    const int i = a.p.owner[p];
    const D3 ri = ld3(a.b.px, a.b.py, a.b.pz, i);
    const D3 vi = ld3(a.b.vx, a.b.vy, a.b.vz, i);
    const D3 wi = ld3(a.b.wx, a.b.wy, a.b.wz, i);
    const double ai = a.b.r[i], imi = a.b.im[i];
    const int mi = a.b.mat[i];
    const int j = a.p.partner[p];
    const D3 rj = a.p.shift ? shift_point(ld3(a.b.px, a.b.py, a.b.pz, j), a.p.shift[p], a.per)
                            : ld3(a.b.px, a.b.py, a.b.pz, j);
    const double aj = a.b.r[j];
    const D3 rij = ri - rj;
    const D3 n = normalize(rij);
    const double delta = ai + aj - len(rij);
    const D3 rc = rj + n * (aj - 0.5 * delta);
    a.p.ov[p] = delta;
    D3 es{a.p.sx[p], a.p.sy[p], a.p.sz[p]}, er{a.p.rx[p], a.p.ry[p], a.p.rz[p]}, et{a.p.ox[p], a.p.oy[p], a.p.oz[p]};
    V3<T> F{0, 0, 0}, Tq{0, 0, 0}, Fs{0, 0, 0};
    if (delta > 0) {
        const double R = ai * aj / (ai + aj);
        const double m = 1. / (imi + a.b.im[j]);
        const D3 vj = ld3(a.b.vx, a.b.vy, a.b.vz, j), wj = ld3(a.b.wx, a.b.wy, a.b.wz, j);
        const D3 vc = (vi + cross(wi, rc - ri)) - (vj + cross(wj, rc - rj));
        const int ip = pair_index(mi, a.b.mat[j], a.tb.nMat, a.tb.cap);
        contact_model<T>(a.tb, ip, m, R, delta, cvt<T>(vc), cvt<T>(wi - wj), cvt<T>(n), a.dt, es, er, et, F, Tq, Fs);
    } else {
        es = er = et = D3{0, 0, 0};
    }
    a.p.sx[p] = es.x; a.p.sy[p] = es.y; a.p.sz[p] = es.z;
    a.p.rx[p] = er.x; a.p.ry[p] = er.y; a.p.rz[p] = er.z;
    a.p.ox[p] = et.x; a.p.oy[p] = et.y; a.p.oz[p] = et.z;
    // T_c + (r_c - r) x F_c (contactKernel.cu:676, :707-708), with the
    // normal part of F_c (parallel to the lever arm) left out
    const V3<T> li = cvt<T>(rc - ri), lj = cvt<T>(rc - rj);
    const V3<T> Ti = Tq + cross(li, Fs);
    const V3<T> Tj = Tq + cross(lj, Fs);
    ((T*)a.p.Fx)[p] = F.x; ((T*)a.p.Fy)[p] = F.y; ((T*)a.p.Fz)[p] = F.z;
    ((T*)a.p.Ix)[p] = Ti.x; ((T*)a.p.Iy)[p] = Ti.y; ((T*)a.p.Iz)[p] = Ti.z;
    ((T*)a.p.Rx)[p] = Tj.x; ((T*)a.p.Ry)[p] = Tj.y; ((T*)a.p.Rz)[p] = Tj.z;
    if (a.rec) {
        double* r = a.rec + (size_t)p * 12;
        r[0] = rc.x; r[1] = rc.y; r[2] = rc.z;
        r[3] = n.x; r[4] = n.y; r[5] = n.z;
        r[6] = F.x; r[7] = F.y; r[8] = F.z;
        r[9] = Tq.x; r[10] = Tq.y; r[11] = Tq.z;
    }
}

struct BondArgs {
    int nb;
    Balls b;
    const int* rank;
    const int *bi, *bj;  // insertion indices
    int* bonded;
    double *bpx, *bpy, *bpz, *bnx, *bny, *bnz, *bsx, *bsy, *bsz, *btx, *bty, *btz, *bFn, *bTt, *bsig, *btau;
    void *Fx, *Fy, *Fz, *Tix, *Tiy, *Tiz, *Tjx, *Tjy, *Tjz;  // per-bond contributions to i and j (T)
    Tables tb;
    double dt;
    int periodic;  // bonds join the nearest images on periodic axes
    Periodic per;
};

// calBondedForceTorqueKernel (contactKernel.cu:535-655).  The bond force
// goes to both balls through the per-ball gather (no atomics); the lever
// arm of the shear force is taken about each ball centre.
template <class T>
__device__ __forceinline__ void bond_forces_one(const BondArgs& a, int k)
{
    if (a.bonded[k] == 0) {
        a.bFn[k] = 0; a.bTt[k] = 0;
        a.bsx[k] = a.bsy[k] = a.bsz[k] = 0;
        a.btx[k] = a.bty[k] = a.btz[k] = 0;
        ((T*)a.Fx)[k] = ((T*)a.Fy)[k] = ((T*)a.Fz)[k] = 0;
        ((T*)a.Tix)[k] = ((T*)a.Tiy)[k] = ((T*)a.Tiz)[k] = 0;
        ((T*)a.Tjx)[k] = ((T*)a.Tjy)[k] = ((T*)a.Tjz)[k] = 0;
        return;
    }
    const int i = a.rank[a.bi[k]], j = a.rank[a.bj[k]];
    const D3 ri = ld3(a.b.px, a.b.py, a.b.pz, i);
    D3 rj = ld3(a.b.px, a.b.py, a.b.pz, j);
    if (a.periodic) rj = shift_point(rj, nearest_image(ri, rj, a.per), a.per);
    const double ai = a.b.r[i], aj = a.b.r[j];
    const D3 n0 = ld3(a.bnx, a.bny, a.bnz, k);
    const D3 n = normalize(ri - rj);
    const double delta = ai + aj - len(ri - rj);
    const D3 rc = rj + n * (aj - 0.5 * delta);
    st3(a.bpx, a.bpy, a.bpz, k, rc);
    st3(a.bnx, a.bny, a.bnz, k, n);
    const D3 vi = ld3(a.b.vx, a.b.vy, a.b.vz, i), vj = ld3(a.b.vx, a.b.vy, a.b.vz, j);
    const D3 wi = ld3(a.b.wx, a.b.wy, a.b.wz, i), wj = ld3(a.b.wx, a.b.wy, a.b.wz, j);
    const D3 vc = (vi + cross(wi, rc - ri)) - (vj + cross(wj, rc - rj));
    const int ip = pair_index(a.b.mat[i], a.b.mat[j], a.tb.nMat, a.tb.cap);
    double Fn = a.bFn[k], Tt = a.bTt[k], sig = a.bsig[k], tau = a.btau[k];
    D3 Fs = ld3(a.bsx, a.bsy, a.bsz, k), Tb = ld3(a.btx, a.bty, a.btz, k);
    a.bonded[k] = bond_model(Fn, Tt, Fs, Tb, sig, tau, n0, n, vc, wi, wj, ai, aj, a.dt, Bp(a.tb, ip, B_GAMMA),
                             Bp(a.tb, ip, B_E), Bp(a.tb, ip, B_KNKS), Bp(a.tb, ip, B_SIG), Bp(a.tb, ip, B_C),
                             Bp(a.tb, ip, B_MU));
    a.bFn[k] = Fn; a.bTt[k] = Tt; a.bsig[k] = sig; a.btau[k] = tau;
    st3(a.bsx, a.bsy, a.bsz, k, Fs);
    st3(a.btx, a.bty, a.btz, k, Tb);
    const D3 Fb = n * Fn + Fs;
    const D3 Tn = n * Tt + Tb;
    const D3 Ti = Tn + cross(rc - ri, Fs);
    const D3 Tj = -Tn + cross(rc - rj, -Fs);
    ((T*)a.Fx)[k] = (T)Fb.x; ((T*)a.Fy)[k] = (T)Fb.y; ((T*)a.Fz)[k] = (T)Fb.z;
    ((T*)a.Tix)[k] = (T)Ti.x; ((T*)a.Tiy)[k] = (T)Ti.y; ((T*)a.Tiz)[k] = (T)Ti.z;
    ((T*)a.Tjx)[k] = (T)Tj.x; ((T*)a.Tjy)[k] = (T)Tj.y; ((T*)a.Tjz)[k] = (T)Tj.z;
}

// ----------------------------------------------------------------- walls
// Squared distance from a point to an axis-aligned box against r^2
// (ballNeighborSearchKernel.cu:202-222), unfused like the FP64 oracle.
__device__ __forceinline__ bool box_within(const D3& p, double r2, double2 bx, double2 by, double2 bz)
{
#pragma clang fp contract(off)
    const double cx = fmin(fmax(p.x, bx.x), bx.y), cy = fmin(fmax(p.y, by.x), by.y), cz = fmin(fmax(p.z, bz.x), bz.y);
    const double dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
    return dx * dx + dy * dy + dz * dz <= r2;
}

// Bounding box of each triangle in sorted-slot order (fmin/fmax are exact,
// so the boxes equal the ones built from the vertices during the walk).
__global__ void k_tri_boxes(int nt, const int* __restrict__ tsorted, const int* __restrict__ t0,
                            const int* __restrict__ t1, const int* __restrict__ t2, const double* __restrict__ gx,
                            const double* __restrict__ gy, const double* __restrict__ gz, double2* __restrict__ box)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nt) return;
    const int t = tsorted[s];
    const int i0 = t0[t], i1 = t1[t], i2 = t2[t];
    box[3 * (size_t)s] = make_double2(fmin(gx[i0], fmin(gx[i1], gx[i2])), fmax(gx[i0], fmax(gx[i1], gx[i2])));
    box[3 * (size_t)s + 1] = make_double2(fmin(gy[i0], fmin(gy[i1], gy[i2])), fmax(gy[i0], fmax(gy[i1], gy[i2])));
    box[3 * (size_t)s + 2] = make_double2(fmin(gz[i0], fmin(gz[i1], gz[i2])), fmax(gz[i0], fmax(gz[i1], gz[i2])));
}

// Per triangle-grid cell, the triangles a ball hashed into that cell can
// touch: the members of the cell's 27-cell stencil (in the walk order of
// ballNeighborSearchKernel.cu:153-339) whose bounding box comes within
// `reach` (the largest ball radius plus a rounding margin) of the cell's
// box.  Cells on the grid faces are unbounded outward, because positions
// beyond the domain are clamped into them.  Any triangle that passes a
// ball's exact box test is in its cell's list, in the same relative order,
// so the ball walk over this list yields the reference's candidate list.
struct CellCandArgs {
    GridGeom tg;
    int ntc;
    double reach;
    const int *tcs, *tce, *tsorted;
    const double2* tbox;
    const int* start;  // write pass
    int* count;        // count pass
    int* out_tri;      // write pass: triangle ids
    double2* out_box;  // write pass: their boxes
};

template <bool WRITE>
__global__ __launch_bounds__(kT) void k_cell_candidates(CellCandArgs a)
{
    const int c = blockIdx.x * kT + threadIdx.x;
    if (c >= a.ntc) return;
    const GridGeom& g = a.tg;
    const int cx = c % g.nx, cy = (c / g.nx) % g.ny, cz = c / (g.nx * g.ny);
    const int lox = cx > 0 ? -1 : 0, hix = cx < g.nx - 1 ? 1 : 0;
    const int loy = cy > 0 ? -1 : 0, hiy = cy < g.ny - 1 ? 1 : 0;
    const int loz = cz > 0 ? -1 : 0, hiz = cz < g.nz - 1 ? 1 : 0;
    const double inf = HUGE_VAL;
    const double blx = cx > 0 ? g.minx + cx * g.cx - a.reach : -inf;
    const double bhx = cx < g.nx - 1 ? g.minx + (cx + 1) * g.cx + a.reach : inf;
    const double bly = cy > 0 ? g.miny + cy * g.cy - a.reach : -inf;
    const double bhy = cy < g.ny - 1 ? g.miny + (cy + 1) * g.cy + a.reach : inf;
    const double blz = cz > 0 ? g.minz + cz * g.cz - a.reach : -inf;
    const double bhz = cz < g.nz - 1 ? g.minz + (cz + 1) * g.cz + a.reach : inf;
    const int base = WRITE ? a.start[c] : 0;
    int k = 0;
    for (int zz = loz; zz <= hiz; zz++)
        for (int yy = loy; yy <= hiy; yy++) {
            const int row = (cz + zz) * g.ny * g.nx + (cy + yy) * g.nx + cx;
            int s0 = -1, s1 = -1;
            for (int xx = lox; xx <= hix; xx++) {
                const int c0 = a.tcs[row + xx], c1 = a.tce[row + xx];
                if (c0 != -1) {
                    if (s0 < 0) s0 = c0;
                    s1 = c1;
                }
            }
            for (int sl = s0; sl < s1; sl++) {
                const double2 bx = a.tbox[3 * (size_t)sl], by = a.tbox[3 * (size_t)sl + 1],
                              bz = a.tbox[3 * (size_t)sl + 2];
                if (bx.x <= bhx && bx.y >= blx && by.x <= bhy && by.y >= bly && bz.x <= bhz && bz.y >= blz) {
                    if (WRITE) {
                        const size_t o = (size_t)base + k;
                        a.out_tri[o] = a.tsorted[sl];
                        a.out_box[3 * o] = bx;
                        a.out_box[3 * o + 1] = by;
                        a.out_box[3 * o + 2] = bz;
                    }
                    k++;
                }
            }
        }
    if (!WRITE) a.count[c] = k;
}

struct WallArgs {
    Balls b;
    Tables tb;
    double dt;
    GridGeom tg;
    const int *tcs, *tce, *tsorted;  // triangle cells over circumcentres
    const double2* tbox;             // per sorted slot: (min,max) x, y, z of the triangle
    const int* cstart;               // per triangle cell: candidate range [cstart[c], cstart[c+1])
    const int* ctri;                 // candidate triangle ids
    const double2* cbox;             // their boxes
    const int *t0, *t1, *t2, *twall;
    const double *vgx, *vgy, *vgz;     // vertex global positions
    const double *wpx, *wpy, *wpz, *wvx, *wvy, *wvz, *wwx, *wwy, *wwz;
    const int* wmat;
    int KW, nballs;
    // per-ball history keyed by insertion index, entry k at k*nballs + o
    int* hcnt;
    int *htri, *hcancel;
    double *hsx, *hsy, *hsz, *hrx, *hry, *hrz, *hox, *hoy, *hoz;
    double *hov;
    // optional record of every candidate contact of the evaluation, 12
    // doubles per history entry: point, normal, force, torque (the torque
    // at the contact, without the lever term)
    double* rec;
    void *ofx, *ofy, *ofz, *otx, *oty, *otz;  // wall force/torque per slot (precision T)
    int* overflow;
};

// Slots of the register path (the cells of a box of walls list 2 to 6
// triangles: two per face).
constexpr int kWallSlots = 4;
// Wall-contact candidates per ball: the default capacity, and the larger one
// a scene moves to (history arrays grown, the block replayed) when a ball
// finds more triangles within reach.  More than kWallCapMax is an error.
constexpr int kWallCap = 16;
constexpr int kWallCapMax = 64;

// wall_contacts_ball for a ball whose cell lists at most NS candidate
// triangles: every per-candidate value sits in a register slot indexed at
// compile time (slot u = the cell's candidate u), so nothing goes through
// scratch memory, and the loads of all slots are issued together (boxes,
// ids and the first history entries; then triangle corners, wall ids and
// matched springs; then corner positions and wall states), a few memory
// latencies per ball instead of one per dependent load.  A slot is live
// when its triangle box is within reach and capacity remains; live slots,
// in order, are the general path's compacted candidates, so every result
// and its order are the same.
template <class T, int NS>
__device__ __forceinline__ void wall_contacts_slots(const WallArgs& a, int i, int o, const D3& ri, double rad,
                                                    int c0, int c1)
{
    constexpr int QH = 4;  // history entries loaded with the candidates
    const int N = a.nballs;
    const int cap = min(kWallCap, a.KW);
    const double r2 = rad * rad;
    // candidates and the first history entries
    double2 bx[NS], by[NS], bz[NS];
    int ct[NS];
#pragma unroll
    for (int u = 0; u < NS; u++) {
        bx[u] = by[u] = bz[u] = make_double2(0.0, 0.0);
        ct[u] = -1;
        if (c0 + u < c1) {
            const double2* q = a.cbox + 3 * (size_t)(c0 + u);
            bx[u] = q[0]; by[u] = q[1]; bz[u] = q[2];
            ct[u] = a.ctri[c0 + u];
        }
    }
    const int oldc = a.hcnt[o];
    int ht[QH];
#pragma unroll
    for (int q = 0; q < QH; q++) ht[q] = q < a.KW ? a.htri[(size_t)q * N + o] : -1;
    bool live[NS];
    int tri[NS], cidx[NS];
    int cnt = 0, hits = 0;
#pragma unroll
    for (int u = 0; u < NS; u++) {
        const bool in = c0 + u < c1 && box_within(ri, r2, bx[u], by[u], bz[u]);
        live[u] = in && hits < cap;
        hits += in ? 1 : 0;
        tri[u] = live[u] ? ct[u] : -1;
        cidx[u] = cnt;
        cnt += live[u] ? 1 : 0;
    }
    if (hits > cap) atomicMax(a.overflow, hits);
    // corners, walls and matched springs
    int v0[NS], v1[NS], v2[NS], wl[NS];
    D3 es[NS], er[NS], et[NS];
#pragma unroll
    for (int u = 0; u < NS; u++) {
        es[u] = er[u] = et[u] = D3{0, 0, 0};
        v0[u] = v1[u] = v2[u] = wl[u] = 0;
        if (!live[u]) continue;
        const int t = tri[u];
        v0[u] = a.t0[t]; v1[u] = a.t1[t]; v2[u] = a.t2[t]; wl[u] = a.twall[t];
        int hq = -1;
#pragma unroll
        for (int q = QH - 1; q >= 0; q--)
            if (q < oldc && ht[q] == t) hq = q;
        for (int q = QH; hq < 0 && q < oldc; q++)
            if (a.htri[(size_t)q * N + o] == t) hq = q;
        if (hq >= 0) {
            const size_t h = (size_t)hq * N + o;
            es[u] = D3{a.hsx[h], a.hsy[h], a.hsz[h]};
            er[u] = D3{a.hrx[h], a.hry[h], a.hrz[h]};
            et[u] = D3{a.hox[h], a.hoy[h], a.hoz[h]};
        }
    }
    // corner positions and classification
    D3 p0[NS], p1[NS], p2[NS], cp[NS];
    int type[NS], cancel[NS];
#pragma unroll
    for (int u = 0; u < NS; u++) {
        cp[u] = D3{0, 0, 0};
        type[u] = ST_NONE;
        cancel[u] = 0;
        p0[u] = ld3(a.vgx, a.vgy, a.vgz, v0[u]);
        p1[u] = ld3(a.vgx, a.vgy, a.vgz, v1[u]);
        p2[u] = ld3(a.vgx, a.vgy, a.vgz, v2[u]);
        if (live[u]) type[u] = classify(ri, rad, p0[u], p1[u], p2[u], cp[u]);
    }
#pragma unroll
    for (int u = 0; u < NS; u++) {
        if (!live[u]) continue;
        if (type[u] == ST_NONE) {
            es[u] = er[u] = et[u] = D3{0, 0, 0};
            cancel[u] = 1;
            continue;
        }
        if (type[u] == ST_FACE) continue;
        bool done = false;
#pragma unroll
        for (int u1 = 0; u1 < NS; u1++) {
            if (u1 == u || done || !live[u1]) continue;
            const int t1 = type[u1];
            if (t1 == ST_NONE) continue;
            bool hit = false;
            if (t1 == ST_FACE) {
                const D3 q0 = p0[u1], q1 = p1[u1], q2 = p2[u1];
                const D3 e01 = cross(cp[u] - q0, q1 - q0), e12 = cross(cp[u] - q1, q2 - q1),
                         e02 = cross(cp[u] - q0, q2 - q0);
                hit = dot(e01, e01) < 1.e-20 || dot(e12, e12) < 1.e-20 || dot(e02, e02) < 1.e-20;
            } else {
                const D3 d = cp[u] - cp[u1];
                const bool same = dot(d, d) < 1.e-20;
                if (t1 == ST_EDGE) hit = (type[u] == t1) ? (u1 < u && same) : same;
                else hit = (type[u] == t1) && u1 < u && same;
            }
            if (hit) {
                es[u] = es[u1]; er[u] = er[u1]; et[u] = et[u1];
                cancel[u] = 1;
                done = true;
            }
        }
    }
    V3<T> Fw{0, 0, 0}, Tw{0, 0, 0};
    const double im = a.b.im[i];
    if (!(im < 1.e-20)) {
        const D3 vi = ld3(a.b.vx, a.b.vy, a.b.vz, i), wi = ld3(a.b.wx, a.b.wy, a.b.wz, i);
        const double m = 1. / im;
        const int mi = a.b.mat[i];
#pragma unroll
        for (int u = 0; u < NS; u++) {
            if (!live[u] || cancel[u]) continue;
            const int w = wl[u];
            const D3 rc = cp[u];
            const D3 n = normalize(ri - rc);
            const double delta = rad - len(ri - rc);
            const D3 rj = ld3(a.wpx, a.wpy, a.wpz, w), vj = ld3(a.wvx, a.wvy, a.wvz, w), wj = ld3(a.wwx, a.wwy, a.wwz, w);
            const D3 vc = (vi + cross(wi, rc - ri)) - (vj + cross(wj, rc - rj));
            const int ip = pair_index(mi, a.wmat[w], a.tb.nMat, a.tb.cap);
            V3<T> F{0, 0, 0}, Tq{0, 0, 0}, Fs{0, 0, 0};
            contact_model<T>(a.tb, ip, m, rad, delta, cvt<T>(vc), cvt<T>(wi - wj), cvt<T>(n), a.dt, es[u], er[u], et[u],
                             F, Tq, Fs);
            Fw = Fw + F;
            Tw = Tw + Tq + cross(cvt<T>(rc - ri), Fs);
            if (a.rec) {
                double* r = a.rec + ((size_t)cidx[u] * N + o) * 12;
                r[6] = F.x; r[7] = F.y; r[8] = F.z;
                r[9] = Tq.x; r[10] = Tq.y; r[11] = Tq.z;
            }
        }
    }
    a.hcnt[o] = cnt;
#pragma unroll
    for (int u = 0; u < NS; u++) {
        if (!live[u]) continue;
        const size_t h = (size_t)cidx[u] * N + o;
        a.htri[h] = tri[u];
        a.hcancel[h] = cancel[u];
        a.hsx[h] = es[u].x; a.hsy[h] = es[u].y; a.hsz[h] = es[u].z;
        a.hrx[h] = er[u].x; a.hry[h] = er[u].y; a.hrz[h] = er[u].z;
        a.hox[h] = et[u].x; a.hoy[h] = et[u].y; a.hoz[h] = et[u].z;
        a.hov[h] = rad - len(ri - cp[u]);
        if (a.rec) {
            double* r = a.rec + h * 12;
            const D3 n = normalize(ri - cp[u]);
            r[0] = cp[u].x; r[1] = cp[u].y; r[2] = cp[u].z;
            r[3] = n.x; r[4] = n.y; r[5] = n.z;
            if (cancel[u] || im < 1.e-20)
                for (int k = 6; k < 12; k++) r[k] = 0.0;
        }
    }
    ((T*)a.ofx)[i] = Fw.x; ((T*)a.ofy)[i] = Fw.y; ((T*)a.ofz)[i] = Fw.z;
    ((T*)a.otx)[i] = Tw.x; ((T*)a.oty)[i] = Tw.y; ((T*)a.otz)[i] = Tw.z;
}

// Scenes up to this many balls put at most one wave of pair work and one of
// wall work on each SIMD, so the register path's lower occupancy costs
// nothing there.
constexpr int kWallSlotBalls = 32768;

// count/writeBallTriangleInteractions (ballNeighborSearchKernel.cu:153-339),
// updateBallTriangleContact (contactKernel.cu:261-418) and
// calBallWallContactForceTorqueKernel (:420-533), fused per ball.  The
// candidate list, its spring history and cancel flags live in registers
// and are written back as the new history.
template <class T, bool SLOTS = false, int KMAX = kWallCap>
__device__ __forceinline__ void wall_contacts_ball(const WallArgs& a, int i)
{
    const int o = a.b.orig[i];
    const int N = a.nballs;
    const D3 ri = ld3(a.b.px, a.b.py, a.b.pz, i);
    const double rad = a.b.r[i];
    int tri[KMAX];
    int cnt = 0;
    // Candidate walk: the triangles of the ball's cell list (the stencil
    // members near the cell, in walk order) tested against the ball with
    // the distance to each triangle's bounding box (no skin), four ahead.
    const Stencil st = make_stencil(a.tg, ri.x, ri.y, ri.z);
    const int cell = (st.cz * a.tg.ny + st.cy) * a.tg.nx + st.cx;
    const double r2 = rad * rad;
    const int c0 = a.cstart[cell], c1 = a.cstart[cell + 1];
    if (SLOTS && c1 - c0 <= kWallSlots) {
        wall_contacts_slots<T, kWallSlots>(a, i, o, ri, rad, c0, c1);
        return;
    }
    for (int j0 = c0; j0 < c1; j0 += 4) {
        double2 bx[4], by[4], bz[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double2* q = a.cbox + 3 * (size_t)min(j0 + u, c1 - 1);
            bx[u] = q[0]; by[u] = q[1]; bz[u] = q[2];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            if (j0 + u >= c1) break;
            if (box_within(ri, r2, bx[u], by[u], bz[u])) {
                if (cnt < KMAX && cnt < a.KW) tri[cnt] = a.ctri[j0 + u];
                cnt++;
            }
        }
    }
    if (cnt > KMAX || cnt > a.KW) {
        atomicMax(a.overflow, cnt);
        cnt = min(KMAX, a.KW);
    }
    // spring history carry by triangle id
    D3 es[KMAX], er[KMAX], et[KMAX];
    int cancel[KMAX], type[KMAX];
    D3 cp[KMAX];
    const int oldc = a.hcnt[o];
    for (int c = 0; c < cnt; c++) {
        es[c] = er[c] = et[c] = D3{0, 0, 0};
        for (int q = 0; q < oldc; q++) {
            const size_t h = (size_t)q * N + o;
            if (a.htri[h] == tri[c]) {
                es[c] = D3{a.hsx[h], a.hsy[h], a.hsz[h]};
                er[c] = D3{a.hrx[h], a.hry[h], a.hrz[h]};
                et[c] = D3{a.hox[h], a.hoy[h], a.hoz[h]};
                break;
            }
        }
    }
    // classification and cancellation of duplicate edge/vertex contacts
    for (int c = 0; c < cnt; c++) {
        const int t = tri[c];
        const D3 p0 = ld3(a.vgx, a.vgy, a.vgz, a.t0[t]), p1 = ld3(a.vgx, a.vgy, a.vgz, a.t1[t]),
                 p2 = ld3(a.vgx, a.vgy, a.vgz, a.t2[t]);
        type[c] = classify(ri, rad, p0, p1, p2, cp[c]);
        cancel[c] = 0;
    }
    for (int c = 0; c < cnt; c++) {
        if (type[c] == ST_NONE) {
            es[c] = er[c] = et[c] = D3{0, 0, 0};
            cancel[c] = 1;
            continue;
        }
        if (type[c] == ST_FACE) continue;
        for (int c1 = 0; c1 < cnt; c1++) {
            if (c1 == c) continue;
            const int t1 = type[c1];
            if (t1 == ST_NONE) continue;
            bool hit = false;
            if (t1 == ST_FACE) {
                const int t = tri[c1];
                const D3 q0 = ld3(a.vgx, a.vgy, a.vgz, a.t0[t]), q1 = ld3(a.vgx, a.vgy, a.vgz, a.t1[t]),
                         q2 = ld3(a.vgx, a.vgy, a.vgz, a.t2[t]);
                const D3 e01 = cross(cp[c] - q0, q1 - q0), e12 = cross(cp[c] - q1, q2 - q1), e02 = cross(cp[c] - q0, q2 - q0);
                hit = dot(e01, e01) < 1.e-20 || dot(e12, e12) < 1.e-20 || dot(e02, e02) < 1.e-20;
            } else {
                const D3 d = cp[c] - cp[c1];
                const bool same = dot(d, d) < 1.e-20;
                if (t1 == ST_EDGE) hit = (type[c] == t1) ? (c1 < c && same) : same;
                else hit = (type[c] == t1) && c1 < c && same;
            }
            if (hit) {
                es[c] = es[c1]; er[c] = er[c1]; et[c] = et[c1];
                cancel[c] = 1;
                break;
            }
        }
    }
    // forces (balls with inverse mass below 1e-20 skip, their springs keep)
    V3<T> Fw{0, 0, 0}, Tw{0, 0, 0};
    const double im = a.b.im[i];
    if (!(im < 1.e-20)) {
        const D3 vi = ld3(a.b.vx, a.b.vy, a.b.vz, i), wi = ld3(a.b.wx, a.b.wy, a.b.wz, i);
        const double m = 1. / im;
        for (int c = 0; c < cnt; c++) {
            if (cancel[c]) continue;
            const int t = tri[c];
            const int w = a.twall[t];
            const D3 rc = cp[c];
            const D3 n = normalize(ri - rc);
            const double delta = rad - len(ri - rc);
            const D3 rj = ld3(a.wpx, a.wpy, a.wpz, w), vj = ld3(a.wvx, a.wvy, a.wvz, w), wj = ld3(a.wwx, a.wwy, a.wwz, w);
            const D3 vc = (vi + cross(wi, rc - ri)) - (vj + cross(wj, rc - rj));
            const int ip = pair_index(a.b.mat[i], a.wmat[w], a.tb.nMat, a.tb.cap);
            V3<T> F{0, 0, 0}, Tq{0, 0, 0}, Fs{0, 0, 0};
            contact_model<T>(a.tb, ip, m, rad, delta, cvt<T>(vc), cvt<T>(wi - wj), cvt<T>(n), a.dt, es[c], er[c], et[c],
                             F, Tq, Fs);
            Fw = Fw + F;
            Tw = Tw + Tq + cross(cvt<T>(rc - ri), Fs);
            if (a.rec) {
                double* r = a.rec + ((size_t)c * N + o) * 12;
                r[6] = F.x; r[7] = F.y; r[8] = F.z;
                r[9] = Tq.x; r[10] = Tq.y; r[11] = Tq.z;
            }
        }
    }
    // new history
    a.hcnt[o] = cnt;
    for (int c = 0; c < cnt; c++) {
        const size_t h = (size_t)c * N + o;
        a.htri[h] = tri[c];
        a.hcancel[h] = cancel[c];
        a.hsx[h] = es[c].x; a.hsy[h] = es[c].y; a.hsz[h] = es[c].z;
        a.hrx[h] = er[c].x; a.hry[h] = er[c].y; a.hrz[h] = er[c].z;
        a.hox[h] = et[c].x; a.hoy[h] = et[c].y; a.hoz[h] = et[c].z;
        a.hov[h] = rad - len(ri - cp[c]);
        if (a.rec) {
            double* r = a.rec + h * 12;
            const D3 n = normalize(ri - cp[c]);
            r[0] = cp[c].x; r[1] = cp[c].y; r[2] = cp[c].z;
            r[3] = n.x; r[4] = n.y; r[5] = n.z;
            if (cancel[c] || im < 1.e-20)
                for (int k = 6; k < 12; k++) r[k] = 0.0;
        }
    }
    ((T*)a.ofx)[i] = Fw.x; ((T*)a.ofy)[i] = Fw.y; ((T*)a.ofz)[i] = Fw.z;
    ((T*)a.otx)[i] = Tw.x; ((T*)a.oty)[i] = Tw.y; ((T*)a.otz)[i] = Tw.z;
}
template <class T, int KMAX>
__global__ __launch_bounds__(kT) void k_wall_contacts(WallArgs a)
{
    const int i = xcd_block() * kT + threadIdx.x;
    if (i >= a.b.n) return;
    wall_contacts_ball<T, false, KMAX>(a, i);
}

// ----------------------------------------------------------------- gather
struct GatherArgs {
    Balls b;
    const int *pstart;                            // owner CSR of the pairs
    int pcap;
    const void *pIx, *pIy, *pIz;                  // owner-side pair torques
    const void *pFx, *pFy, *pFz, *pRx, *pRy, *pRz;  // pair results
    const int *rstart, *rpair;
    const int *istart, *ibond;                    // bond incidence by insertion index (id*2 + side)
    const void *bFx, *bFy, *bFz, *bTix, *bTiy, *bTiz, *bTjx, *bTjy, *bTjz;
    const void *wfx, *wfy, *wfz, *wtx, *wty, *wtz;  // walls (nullable)
    Forces out;
};

template <class T>
__device__ __forceinline__ T at(const void* p, int i) { return ((const T*)p)[i]; }

// sumObjectPointingForceTorqueFromInteraction (contactKernel.cu:683-714) and
// the per-ball totals: owner sums - partner-side pair terms + bonds + walls.
template <class T>
__device__ __forceinline__ V3<T> at3(const void* x, const void* y, const void* z, int i)
{
    return V3<T>{at<T>(x, i), at<T>(y, i), at<T>(z, i)};
}

// Entries of each sum loaded per round trip: the loads of a batch are
// independent, so a ball's short sums cost a few memory latencies instead of
// one per entry; the adds still run one entry at a time in list order, and
// entries past the end are loaded from the last valid index and not added.
constexpr int kGatherBatch = 4;

template <class T>
__device__ __forceinline__ void gather_ball(const GatherArgs& a, int j)
{
    constexpr int B = kGatherBatch;
    // owner side: the sum over the ball's own entries, in list order
    V3<T> Fo{0, 0, 0}, To{0, 0, 0};
    const int p0 = a.pstart[j], p1 = min(a.pstart[j + 1], a.pcap);
    const int q0 = a.rstart[j], q1 = a.rstart[j + 1];
    for (int p = p0; p < p1; p += B) {
        V3<T> f[B], t[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            const int pk = min(p + k, p1 - 1);
            f[k] = at3<T>(a.pFx, a.pFy, a.pFz, pk);
            t[k] = at3<T>(a.pIx, a.pIy, a.pIz, pk);
        }
#pragma unroll
        for (int k = 0; k < B; k++)
            if (p + k < p1) {
                Fo = Fo + f[k];
                To = To + t[k];
            }
    }
    T fx = Fo.x, fy = Fo.y, fz = Fo.z;
    T tx = To.x, ty = To.y, tz = To.z;
    // partner side, in ascending pair order
    for (int q = q0; q < q1; q += B) {
        int pp[B];
#pragma unroll
        for (int k = 0; k < B; k++) pp[k] = a.rpair[min(q + k, q1 - 1)];
        V3<T> f[B], r[B];
#pragma unroll
        for (int k = 0; k < B; k++) {
            f[k] = at3<T>(a.pFx, a.pFy, a.pFz, pp[k]);
            r[k] = at3<T>(a.pRx, a.pRy, a.pRz, pp[k]);
        }
#pragma unroll
        for (int k = 0; k < B; k++)
            if (q + k < q1) {
                fx -= f[k].x; fy -= f[k].y; fz -= f[k].z;
                tx -= r[k].x; ty -= r[k].y; tz -= r[k].z;
            }
    }
    if (a.istart) {
        const int o = a.b.orig[j];
        const int e0 = a.istart[o], e1 = a.istart[o + 1];
        for (int q = e0; q < e1; q += B) {
            int ee[B];
#pragma unroll
            for (int k = 0; k < B; k++) ee[k] = a.ibond[min(q + k, e1 - 1)];
            V3<T> f[B], t[B];
#pragma unroll
            for (int k = 0; k < B; k++) {
                const int b = ee[k] >> 1;
                f[k] = at3<T>(a.bFx, a.bFy, a.bFz, b);
                t[k] = (ee[k] & 1) == 0 ? at3<T>(a.bTix, a.bTiy, a.bTiz, b) : at3<T>(a.bTjx, a.bTjy, a.bTjz, b);
            }
#pragma unroll
            for (int k = 0; k < B; k++) {
                if (q + k >= e1) continue;
                if ((ee[k] & 1) == 0) {
                    fx += f[k].x; fy += f[k].y; fz += f[k].z;
                } else {
                    fx -= f[k].x; fy -= f[k].y; fz -= f[k].z;
                }
                tx += t[k].x; ty += t[k].y; tz += t[k].z;
            }
        }
    }
    if (a.wfx) {
        fx += at<T>(a.wfx, j); fy += at<T>(a.wfy, j); fz += at<T>(a.wfz, j);
        tx += at<T>(a.wtx, j); ty += at<T>(a.wty, j); tz += at<T>(a.wtz, j);
    }
    a.out.fx[j] = fx; a.out.fy[j] = fy; a.out.fz[j] = fz;
    a.out.tx[j] = tx; a.out.ty[j] = ty; a.out.tz[j] = tz;
}
template <class T>
__global__ __launch_bounds__(kT) void k_gather(GatherArgs a)
{
    const int j = xcd_block() * kT + threadIdx.x;
    if (j >= a.b.n) return;
    gather_ball<T>(a, j);
}

// ----------------------------------------------------------------- integration
struct ClumpArgs {
    int nc;
    double *px, *py, *pz, *vx, *vy, *vz, *wx, *wy, *wz, *q0, *q1, *q2, *q3;
    const double *ixx, *iyy, *izz, *ixy, *ixz, *iyz, *im;
    const int *ps, *pe;  // pebble insertion-index ranges
    double *Fx, *Fy, *Fz, *Tx, *Ty, *Tz;
    Balls b;
    Forces f;
    const int* rank;
    double gx, gy, gz, h;
};

// sumClumpForceTorqueKernel (ballIntegrationKernel.cu:3-28, pebble forces),
// clumpVelocityAngularVelocityIntegrationKernel (:58-79) and, in the first
// half, the clump drift and orientation update (:101-121, :230-238).
template <bool FIRST>
__global__ __launch_bounds__(kT) void k_clump_half(ClumpArgs a)
{
    const int c = blockIdx.x * kT + threadIdx.x;
    if (c >= a.nc) return;
    const D3 rc = ld3(a.px, a.py, a.pz, c);
    D3 F{0, 0, 0}, T{0, 0, 0};
    for (int o = a.ps[c]; o < a.pe[c]; o++) {
        const int s = a.rank[o];
        const D3 r = ld3(a.b.px, a.b.py, a.b.pz, s);
        const D3 Fi = ld3(a.f.fx, a.f.fy, a.f.fz, s);
        F = F + Fi;
        T = T + (ld3(a.f.tx, a.f.ty, a.f.tz, s) + cross(r - rc, Fi));
    }
    st3(a.Fx, a.Fy, a.Fz, c, F);
    st3(a.Tx, a.Ty, a.Tz, c, T);
    const double im = a.im[c];
    D3 v = ld3(a.vx, a.vy, a.vz, c), w = ld3(a.wx, a.wy, a.wz, c);
    Quat q{a.q0[c], a.q1[c], a.q2[c], a.q3[c]};
    if (im > 1.e-20) {
        v = v + (F * im + D3{a.gx, a.gy, a.gz}) * a.h;
        const Sym I{a.ixx[c], a.iyy[c], a.izz[c], a.ixy[c], a.ixz[c], a.iyz[c]};
        w = w + sym_mul(sym_rotate(q, I), T) * a.h;
        st3(a.vx, a.vy, a.vz, c, v);
        st3(a.wx, a.wy, a.wz, c, w);
    }
    if (FIRST) {
        st3(a.px, a.py, a.pz, c, rc + v * (2.0 * a.h));
        q = qintegrate(q, w, 2.0 * a.h);
        a.q0[c] = q.q0; a.q1[c] = q.q1; a.q2[c] = q.q2; a.q3[c] = q.q3;
    }
}

struct BallHalfArgs {
    Balls b;
    Forces f;
    int nc;
    const double *cpx, *cpy, *cpz, *cvx, *cvy, *cvz, *cwx, *cwy, *cwz;  // clumps (nullable)
    double gx, gy, gz, h;
};

// setPebbleVelocityAngularVelocityKernel (ballIntegrationKernel.cu:81-99),
// ballVelocityAngularVelocityIntegrationKernel (:30-56) and, in the first
// half, positionIntegrationKernel (:112-121) for every ball.
template <bool FIRST>
__device__ __forceinline__ void ball_half_ball(const BallHalfArgs& a, int i)
{
    D3 v = ld3(a.b.vx, a.b.vy, a.b.vz, i);
    const int c = a.b.clump[i];
    // A clump id without a clump object only excludes contacts; such a ball
    // is neither kicked nor driven by a clump (ballIntegrationKernel.cu:44).
    if (c >= 0 && c < a.nc) {
        const D3 w = ld3(a.cwx, a.cwy, a.cwz, c);
        st3(a.b.wx, a.b.wy, a.b.wz, i, w);
        v = ld3(a.cvx, a.cvy, a.cvz, c) + cross(w, ld3(a.b.px, a.b.py, a.b.pz, i) - ld3(a.cpx, a.cpy, a.cpz, c));
        st3(a.b.vx, a.b.vy, a.b.vz, i, v);
    } else if (c < 0) {
        const double im = a.b.im[i];
        if (!(im < 1.e-20)) {
            v = v + (ld3(a.f.fx, a.f.fy, a.f.fz, i) * im + D3{a.gx, a.gy, a.gz}) * a.h;
            st3(a.b.vx, a.b.vy, a.b.vz, i, v);
            const double r = a.b.r[i];
            if (!(r < 1.e-20)) {
                const double I = 0.4 * r * r / im;
                D3 w = ld3(a.b.wx, a.b.wy, a.b.wz, i);
                w = w + divide(ld3(a.f.tx, a.f.ty, a.f.tz, i), I) * a.h;
                st3(a.b.wx, a.b.wy, a.b.wz, i, w);
            }
        }
    }
    if (FIRST) {
        const double dt = 2.0 * a.h;
        a.b.px[i] += dt * v.x;
        a.b.py[i] += dt * v.y;
        a.b.pz[i] += dt * v.z;
    }
}
template <bool FIRST>
__global__ __launch_bounds__(kT) void k_ball_half(BallHalfArgs a)
{
    const int i = blockIdx.x * kT + threadIdx.x;
    if (i >= a.b.n) return;
    ball_half_ball<FIRST>(a, i);
}

// The two per-ball loops of a force evaluation as one launch: contacts over
// the ball's owned pairs, then its wall contacts.  Neither reads what the
// other writes.
// One launch for the force terms of an evaluation, in three block ranges:
// ball pairs, ball-wall contacts and bonds.  The three write disjoint
// outputs, so their latency chains
// overlap on the device instead of running back to back.  WALLS: 0 none,
// 1 the general candidate path, 2 register slots for cells of up to
// kWallSlots candidates (more registers, fewer waves per SIMD: used for
// scenes too small to fill the device), 3 the general path at the larger
// capacity kWallCapMax.
//
// Each range is padded to a multiple of the XCD count and every XCD takes an
// eighth of each range (consecutive blocks of it, for L2 reuse): blocks are
// dispatched to XCDs round-robin, so hardware block b runs on XCD b % 8 as
// the (b / 8)-th of that XCD's blocks.  Ranges laid end to end instead would
// crowd the heavier pair blocks onto a few XCDs.
template <class T, int WALLS, bool BONDS>
__global__ __launch_bounds__(kT) void k_forces(PairForceArgs pa, WallArgs wa, BondArgs ba, int qp, int qw, int qb)
{
    const int xcd = blockIdx.x % kNumXcd, pos = blockIdx.x / kNumXcd;
    if (pos < qp) {
        const int p = (xcd * qp + pos) * kT + threadIdx.x;
        if (p < min(pa.p.start[pa.b.n], pa.p.cap)) pair_force_one<T>(pa, p);
    } else if (WALLS && pos < qp + qw) {
        const int i = (xcd * qw + pos - qp) * kT + threadIdx.x;
        if (i < wa.b.n) {
            if (WALLS == 3) wall_contacts_ball<T, false, kWallCapMax>(wa, i);
            else wall_contacts_ball<T, WALLS == 2, kWallCap>(wa, i);
        }
    } else if (BONDS && pos < qp + qw + qb) {
        const int k = (xcd * qb + pos - qp - qw) * kT + threadIdx.x;
        if (k < ba.nb) bond_forces_one<T>(ba, k);
    }
}

// A ball's force and torque gather and its half kick (with the drift in the
// first half) as one launch: the gather reads only pair, bond and wall
// terms, and the kick only the ball's own state.  Used without clumps (a
// clump's kick sums its pebbles' forces first) and without hooks (they act
// between the two).
template <class T, bool FIRST>
__global__ __launch_bounds__(kT) void k_gather_half(GatherArgs ga, BallHalfArgs ha)
{
    const int i = xcd_block() * kT + threadIdx.x;
    if (i >= ga.b.n) return;
    gather_ball<T>(ga, i);
    ball_half_ball<FIRST>(ha, i);
}

// launchWallIntegration / vertex update / circumcentres
// (wallIntegrationKernel.cu:3-53), full dt per call.
struct WallMoveArgs {
    int nw, nv, nt;
    double *px, *py, *pz, *q0, *q1, *q2, *q3;
    const double *vx, *vy, *vz, *wx, *wy, *wz;
    const double *lx, *ly, *lz;
    double *gx, *gy, *gz;
    const int* vwall;
    const int *t0, *t1, *t2;
    double* tcc;  // interleaved x3
    double dt;
};
__global__ void k_wall_move(WallMoveArgs a)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= a.nw) return;
    a.px[w] += a.dt * a.vx[w];
    a.py[w] += a.dt * a.vy[w];
    a.pz[w] += a.dt * a.vz[w];
    Quat q{a.q0[w], a.q1[w], a.q2[w], a.q3[w]};
    q = qintegrate(q, ld3(a.wx, a.wy, a.wz, w), a.dt);
    a.q0[w] = q.q0; a.q1[w] = q.q1; a.q2[w] = q.q2; a.q3[w] = q.q3;
}
__global__ void k_vertex_update(WallMoveArgs a)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.nv) return;
    const int w = a.vwall[v];
    const Quat q{a.q0[w], a.q1[w], a.q2[w], a.q3[w]};
    st3(a.gx, a.gy, a.gz, v, ld3(a.px, a.py, a.pz, w) + qrotate(q, ld3(a.lx, a.ly, a.lz, v)));
}
__global__ void k_circumcenters(WallMoveArgs a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nt) return;
    const int i0 = a.t0[t], i1 = a.t1[t], i2 = a.t2[t];
    const double p0[3] = {a.gx[i0], a.gy[i0], a.gz[i0]}, p1[3] = {a.gx[i1], a.gy[i1], a.gz[i1]},
                 p2[3] = {a.gx[i2], a.gy[i2], a.gz[i2]};
    circumcenter(p0, p1, p2, &a.tcc[3 * (size_t)t]);
}

// ----------------------------------------------------------------- hooks
// addConstantForce (externalForceTorque.cu:4-12), per insertion index.
__device__ __forceinline__ void hook_const_one(const Balls& b, const Forces& f, const double* fo, double s, int i)
{
#pragma clang fp contract(off)
    const int o = b.orig[i];
    const double ax = fo[3 * (size_t)o] * s, ay = fo[3 * (size_t)o + 1] * s, az = fo[3 * (size_t)o + 2] * s;
    f.fx[i] += ax;
    f.fy[i] += ay;
    f.fz[i] += az;
}
// addGlobalDampingForceTorque (externalForceTorque.cu:26-37)
__device__ __forceinline__ void hook_damping_one(const Balls& b, const Forces& f, double c, int i)
{
#pragma clang fp contract(off)
    const D3 F = ld3(f.fx, f.fy, f.fz, i), T = ld3(f.tx, f.ty, f.tz, i);
    double nx, ny, nz;
    const double cf = c * hk_len(F.x, F.y, F.z);
    hk_unit(b.vx[i], b.vy[i], b.vz[i], nx, ny, nz);
    st3(f.fx, f.fy, f.fz, i, D3{F.x - nx * cf, F.y - ny * cf, F.z - nz * cf});
    const double ct = c * hk_len(T.x, T.y, T.z);
    hk_unit(b.wx[i], b.wy[i], b.wz[i], nx, ny, nz);
    st3(f.tx, f.ty, f.tz, i, D3{T.x - nx * ct, T.y - ny * ct, T.z - nz * ct});
}
// addBuoyancyDragKernel (externalForceTorque.cu:66-133)
__device__ __forceinline__ void hook_buoyancy_one(const Balls& b, const Forces& f, double rho, double cd, D3 g, D3 vf,
                                                  double level, int i)
{
#pragma clang fp contract(off)
    if (b.im[i] <= 0.0) return;
    const double zc = b.pz[i], r = b.r[i];
    const double dd = level - zc;
    double Vs;
    if (dd <= -r) return;
    if (dd >= r) Vs = (4.0 / 3.0) * M_PI * r * r * r;
    else {
        const double hh = dd + r;
        Vs = (M_PI * hh * hh * (3.0 * r - hh)) / 3.0;
    }
    if (Vs <= 0.0) return;
    const D3 Fb{-rho * Vs * g.x, -rho * Vs * g.y, -rho * Vs * g.z};
    const double Vb = (4.0 / 3.0) * M_PI * r * r * r;
    const double alpha = (Vb > 1.e-30) ? (Vs / Vb) : 0.0;
    const D3 vr = ld3(b.vx, b.vy, b.vz, i) - vf;
    const double v2 = vr.x * vr.x + vr.y * vr.y + vr.z * vr.z;
    D3 Fd{0, 0, 0};
    if (v2 > 1.e-30 && cd > 0.0) {
        const double vm = sqrt(v2);
        const double A = alpha * (M_PI * r * r);
        const double k = 0.5 * rho * cd * A * vm;
        Fd = D3{-k * vr.x, -k * vr.y, -k * vr.z};
    }
    f.fx[i] += Fb.x + Fd.x;
    f.fy[i] += Fb.y + Fd.y;
    f.fz[i] += Fb.z + Fd.z;
}

__global__ void k_hook_const(Balls b, Forces f, const double* fo, double s)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < b.n) hook_const_one(b, f, fo, s, i);
}
__global__ void k_hook_damping(Balls b, Forces f, double c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < b.n) hook_damping_one(b, f, c, i);
}
__global__ void k_hook_buoyancy(Balls b, Forces f, double rho, double cd, D3 g, D3 vf, double level)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < b.n) hook_buoyancy_one(b, f, rho, cd, g, vf, level, i);
}

// The built-in hook program applied by the gather right after a ball's
// totals are stored (same per-ball arithmetic and order as the separate
// hook launches): up to kMaxHooksFused entries.
constexpr int kMaxHooksFused = 4;
struct HookDev {
    int kind;  // 0 constant force (scaled by s), 1 damping, 2 buoyancy and drag
    double c, rho, cd, level, s;
    D3 g, vf;
    const double* fo;
};
struct HookTable {
    int n;
    HookDev op[kMaxHooksFused];
};
template <class T>
__global__ __launch_bounds__(kT) void k_gather_hooks(GatherArgs a, HookTable ht)
{
    const int j = xcd_block() * kT + threadIdx.x;
    if (j >= a.b.n) return;
    gather_ball<T>(a, j);
    for (int k = 0; k < ht.n; k++) {
        const HookDev& h = ht.op[k];
        if (h.kind == 0) hook_const_one(a.b, a.out, h.fo, h.s, j);
        else if (h.kind == 1) hook_damping_one(a.b, a.out, h.c, j);
        else hook_buoyancy_one(a.b, a.out, h.rho, h.cd, h.g, h.vf, h.level, j);
    }
}

// Views in insertion order for a user hook (AoS x3 doubles).
struct ViewArgs {
    Balls b;
    Forces f;
    double *F, *T, *P, *V, *W, *R, *IM;
};
__global__ void k_views_scatter(ViewArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.b.n) return;
    const size_t o = a.b.orig[i];
    a.F[3 * o] = a.f.fx[i]; a.F[3 * o + 1] = a.f.fy[i]; a.F[3 * o + 2] = a.f.fz[i];
    a.T[3 * o] = a.f.tx[i]; a.T[3 * o + 1] = a.f.ty[i]; a.T[3 * o + 2] = a.f.tz[i];
    a.P[3 * o] = a.b.px[i]; a.P[3 * o + 1] = a.b.py[i]; a.P[3 * o + 2] = a.b.pz[i];
    a.V[3 * o] = a.b.vx[i]; a.V[3 * o + 1] = a.b.vy[i]; a.V[3 * o + 2] = a.b.vz[i];
    a.W[3 * o] = a.b.wx[i]; a.W[3 * o + 1] = a.b.wy[i]; a.W[3 * o + 2] = a.b.wz[i];
    a.R[o] = a.b.r[i];
    a.IM[o] = a.b.im[i];
}
__global__ void k_views_gather(ViewArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.b.n) return;
    const size_t o = a.b.orig[i];
    a.f.fx[i] = a.F[3 * o]; a.f.fy[i] = a.F[3 * o + 1]; a.f.fz[i] = a.F[3 * o + 2];
    a.f.tx[i] = a.T[3 * o]; a.f.ty[i] = a.T[3 * o + 1]; a.f.tz[i] = a.T[3 * o + 2];
}

}  // namespace
}  // namespace demsph

// ====================================================================== host
using namespace demsph;
using namespace demsph::dem;

namespace {

struct HostBalls {
    std::vector<double> pos, vel, w, r, im;
    std::vector<int> mat, clump;
    size_t size() const { return r.size(); }
    void add(const double* p, const double* v, const double* ww, double rr, double inv, int m, int c)
    {
        for (int k = 0; k < 3; k++) {
            pos.push_back(p[k]);
            vel.push_back(v ? v[k] : 0.0);
            w.push_back(ww ? ww[k] : 0.0);
        }
        r.push_back(rr);
        im.push_back(inv);
        mat.push_back(m);
        clump.push_back(c);
    }
};

struct HostClumps {
    std::vector<double> pos, vel, w, invI, im;  // invI: 6 per clump (xx yy zz xy xz yz)
    std::vector<int> mat, ps, pe;
    size_t size() const { return im.size(); }
};

struct HostWalls {
    std::vector<double> pos;  // 3 per wall
    std::vector<int> mat;
    std::vector<double> vlocal;  // 3 per vertex
    std::vector<int> vwall;
    std::vector<int> t0, t1, t2, twall;
    std::vector<double> tcc;  // 3 per triangle
};

struct Row {
    int kind, a, b;
    double p[11];
};

// Owned device array helper.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t m)
    {
        dfree(p);
        p = dmalloc<T>(m);
        n = m;
    }
    void free_() { dfree(p); n = 0; }
    ~DBuf() { dfree(p); }
};

}  // namespace

struct demsph_dem {
    hipStream_t stream = nullptr;
    bool initialized = false;
    int fp64 = 0;
    HostBalls hb;
    HostClumps hc;
    HostWalls hw;
    std::vector<Row> rows;
    std::vector<int> bond_a, bond_b;
    std::vector<double> bond_point, bond_normal;  // host staging of new bonds
    std::set<std::pair<int, int>> bond_set;
    int n = 0, nc = 0, nw = 0, nv = 0, nt = 0, nbond = 0, nbond_dev = 0;
    GridGeom bg{}, tg{};
    long long bng = 0, tng = 0;
    bool bgValid = false, tgValid = false;
    int nMat = 0, cap = 0;
    int wall_slots_set = 1;  // register-slot wall path for small scenes (tuning knob "wallSlots")
    KernelTimer timer;
    StepStamps stamps;  // step-end timestamps (batched serving)

    // balls: two buffers, cur is the live one
    DBuf<double> bpx[2], bpy[2], bpz[2], bvx[2], bvy[2], bvz[2], bwx[2], bwy[2], bwz[2], br[2], bim[2];
    DBuf<int> bmat[2], bclump[2], borig[2];
    int cb = 0;
    DBuf<int> rank, rank_old;
    DBuf<double> fx, fy, fz, tx, ty, tz;
    // sort / cells
    DBuf<uint32_t> key, skey, ktmp;
    DBuf<int> perm, vtmp;
    void* sort_temp = nullptr;
    DBuf<int> cellStart, cellEnd;
    // pairs: two lists, cp is the live one
    int pcap[2] = {0, 0};
    DBuf<int> pstart[2], ppartner[2], powner[2];
    DBuf<double> psx[2], psy[2], psz[2], prx[2], pry[2], prz[2], pox[2], poy[2], poz[2], pov[2];
    DBuf<double> pFx[2], pFy[2], pFz[2], pRx[2], pRy[2], pRz[2];  // hold float or double
    DBuf<double> pIx[2], pIy[2], pIz[2];                            // hold float or double
    int cp = 0;
    bool has_list = false;
    int npairs = 0;
    // periodic boundaries (setPeriodicBoundary): image code per pair
    bool periodic = false;
    Periodic per{};
    DBuf<uint8_t> pshift[2];
    DBuf<int> cnt, rcnt, rstart, rcursor, rpair;
    void* scan_temp = nullptr;
    DBuf<int> total;  // device int
    int* total_h = nullptr;
    // material tables
    DBuf<double> tH, tL, tB;
    // bonds
    DBuf<int> dbi, dbj, dbonded, dinc_start, dinc_bond;
    DBuf<double> dbpx, dbpy, dbpz, dbnx, dbny, dbnz, dbsx, dbsy, dbsz, dbtx, dbty, dbtz, dbFn, dbTt, dbsig, dbtau;
    DBuf<double> dbFx, dbFy, dbFz, dbTix, dbTiy, dbTiz, dbTjx, dbTjy, dbTjz;  // hold float or double
    // clumps
    DBuf<double> cpx, cpy, cpz, cvx, cvy, cvz, cwx, cwy, cwz, cq0, cq1, cq2, cq3, cixx, ciyy, cizz, cixy, cixz, ciyz,
        cim, cFx, cFy, cFz, cTx, cTy, cTz;
    DBuf<int> cps, cpe;
    // walls
    DBuf<double> wpx, wpy, wpz, wvx, wvy, wvz, wwx, wwy, wwz, wq0, wq1, wq2, wq3;
    DBuf<int> wmat;
    DBuf<double> vlx, vly, vlz, vgx, vgy, vgz;
    DBuf<int> vwall, dt0, dt1, dt2, dtwall;
    DBuf<double> tcc;
    DBuf<uint32_t> tkey, tskey, tktmp;
    DBuf<int> tperm, tvtmp, tcs, tce;
    DBuf<double> tbox;            // 6 per triangle, sorted-slot order
    DBuf<int> tcand_cnt, tcand_start, tcand_tri;  // per-cell candidate lists
    DBuf<double> tcand_box;       // 6 per candidate
    void* tscan_temp = nullptr;
    double rmax = 0.0;            // largest ball radius
    bool tri_cells_valid = false;  // triangle cells and boxes match the current vertices
    // Walls start at rest with identity orientation (addMeshWall stores no
    // velocities, DEMSolver.h:1901-1911); until a wall velocity is set,
    // their motion is an exact no-op and the triangle cells stay valid.
    bool walls_static = true;
    void* tsort_temp = nullptr;
    int KW = 16;
    int pair_lanes_set = 0;  // lanes per ball in the pair-list build; 0: by ball count
    bool record_contacts = false;  // keep every contact of the last evaluation (frame files)
    DBuf<double> crec, prec;       // wall contacts per history slot, ball pairs per pair slot
    DBuf<int> hcnt, htri, hcancel;
    DBuf<double> hsx, hsy, hsz, hrx, hry, hrz, hox, hoy, hoz, hov;
    DBuf<double> wfx, wfy, wfz, wtx, wty, wtz;  // hold float or double
    DBuf<int> overflow;
    int* overflow_h = nullptr;
    // Pair lists built inside step blocks: the longest list needed beyond
    // capacity (0: none), and the device checkpoint the block replays from.
    DBuf<int> pair_over;
    int* block_h = nullptr;  // pinned: [0] pair_over, [1] pairs of the current list
    DBuf<unsigned char> ckbuf;
    long pair_replays = 0;  // blocks replayed with grown pair lists
    long wall_replays = 0;  // step calls run again after the wall capacity grew
    struct HostCk {
        size_t iStep;
        double time;
        int cb, cp, npairs;
        bool has_list, walls_valid, tri_cells_valid, walls_static;
    } hck{};
    bool walls_valid = false;  // wall forces of the last triangle evaluation are held
    // hooks
    struct HookOp {
        int kind;
        double c, rho, cd, level, ramp;
        double g[3], vf[3];
        DBuf<double>* f;
    };
    std::vector<HookOp> hooks;
    std::vector<DBuf<double>*> hook_bufs;
    demsph_dem_hook_fn hook_fn = nullptr;
    void* hook_user = nullptr;
    DBuf<double> vF, vT, vP, vV, vW, vR, vIM;
    // clock
    size_t iStep = 0;
    double time = 0.0;
    // Captured graphs of the four steps between two pair rebuilds.  The
    // rebuild flips the ball and pair buffers (cb, cp) and may reallocate the
    // pair arrays, so a graph is keyed by what its launches bake in.
    // Every buffer a capture bakes in that can be reallocated between
    // captures is part of the key (the contact records are reallocated with
    // either pair list).
    // kind 4: the four steps after a rebuild; kind 5: a rebuild and the
    // five steps of its cycle (lists sized on the device).
    struct GraphKey {
        double dt, gx, gy, gz;
        int kind, cb, cp, nbond, fp64, cap0, cap1;
        const void *pairs, *pairs1, *rev, *cand, *stamp, *prec, *crec;
        bool operator==(const GraphKey& o) const
        {
            return dt == o.dt && gx == o.gx && gy == o.gy && gz == o.gz && kind == o.kind && cb == o.cb &&
                   cp == o.cp && nbond == o.nbond && fp64 == o.fp64 && cap0 == o.cap0 && cap1 == o.cap1 &&
                   pairs == o.pairs && pairs1 == o.pairs1 && rev == o.rev && cand == o.cand && stamp == o.stamp &&
                   prec == o.prec && crec == o.crec;
        }
    };
    struct GraphEntry {
        GraphKey key;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    bool use_graphs = true;
    long graph_captures = 0;

    void drop_graphs()
    {
        if (!graphs.empty()) (void)hipStreamSynchronize(stream);
        for (auto& e : graphs) graph_exec_destroy(e.exec);
        graphs.clear();
    }

    std::vector<demsph_couple*> couples;  // couplings holding this scene

    ~demsph_dem()
    {
        // couplings destroyed after this scene must not touch it
        for (auto* c : std::vector<demsph_couple*>(couples)) couple_detach_dem(c);
        (void)hipStreamSynchronize(stream);
        drop_graphs();
        dev_free(sort_temp);
        dev_free(tsort_temp);
        dev_free(tscan_temp);
        tscan_temp = nullptr;
        dev_free(scan_temp);
        host_free(total_h);
        host_free(overflow_h);
        host_free(block_h);
        for (auto* b : hook_bufs) delete b;
    }

    Balls balls(int which) const
    {
        Balls b;
        b.n = n;
        b.px = bpx[which].p; b.py = bpy[which].p; b.pz = bpz[which].p;
        b.vx = bvx[which].p; b.vy = bvy[which].p; b.vz = bvz[which].p;
        b.wx = bwx[which].p; b.wy = bwy[which].p; b.wz = bwz[which].p;
        b.r = br[which].p; b.im = bim[which].p;
        b.mat = bmat[which].p; b.clump = bclump[which].p; b.orig = borig[which].p;
        return b;
    }
    Balls cur() const { return balls(cb); }
    Forces forces() const { return Forces{fx.p, fy.p, fz.p, tx.p, ty.p, tz.p}; }
    Pairs pairs(int which) const
    {
        Pairs p;
        p.start = pstart[which].p; p.partner = ppartner[which].p; p.owner = powner[which].p; p.cap = pcap[which];
        p.sx = psx[which].p; p.sy = psy[which].p; p.sz = psz[which].p;
        p.rx = prx[which].p; p.ry = pry[which].p; p.rz = prz[which].p;
        p.ox = pox[which].p; p.oy = poy[which].p; p.oz = poz[which].p;
        p.Fx = pFx[which].p; p.Fy = pFy[which].p; p.Fz = pFz[which].p;
        p.Rx = pRx[which].p; p.Ry = pRy[which].p; p.Rz = pRz[which].p;
        p.Ix = pIx[which].p; p.Iy = pIy[which].p; p.Iz = pIz[which].p;
        p.ov = pov[which].p;
        p.shift = periodic ? pshift[which].p : nullptr;
        return p;
    }
    Tables tables() const { return Tables{nMat, cap, tH.p, tL.p, tB.p}; }

    void alloc_pairs(int which, int c)
    {
        c = std::max(c, 1);
        ppartner[which].alloc(c);
        powner[which].alloc(c);
        psx[which].alloc(c); psy[which].alloc(c); psz[which].alloc(c);
        prx[which].alloc(c); pry[which].alloc(c); prz[which].alloc(c);
        pox[which].alloc(c); poy[which].alloc(c); poz[which].alloc(c);
        pov[which].alloc(c);
        pFx[which].alloc(c); pFy[which].alloc(c); pFz[which].alloc(c);
        pRx[which].alloc(c); pRy[which].alloc(c); pRz[which].alloc(c);
        pIx[which].alloc(c); pIy[which].alloc(c); pIz[which].alloc(c);
        if (periodic) pshift[which].alloc(c);
        pcap[which] = c;
        if (record_contacts) prec.alloc(12 * (size_t)std::max(pcap[0], pcap[1]));
    }

    // contactModelParameters::buildFromTables (contactKernel.cu:6-104)
    void build_tables()
    {
        int mx = -1;
        for (auto& r : rows) mx = std::max(mx, std::max(r.a, r.b));
        nMat = mx < 0 ? 0 : mx + 1;
        cap = nMat > 0 ? nMat * (nMat + 1) / 2 + 1 : 0;
        std::vector<double> H((size_t)H_N * std::max(cap, 1), 0.0), L((size_t)L_N * std::max(cap, 1), 0.0),
            B((size_t)B_N * std::max(cap, 1), 0.0);
        for (int i = 0; i < cap; i++) {
            H[H_RES * cap + i] = 1.0;
            B[B_GAMMA * cap + i] = 1.0;
            B[B_KNKS * cap + i] = 1.0;
        }
        for (auto& r : rows) {
            const int i = pair_index(r.a, r.b, nMat, cap);
            if (r.kind == 0) for (int q = 0; q < L_N; q++) L[q * cap + i] = r.p[q];
            if (r.kind == 1) for (int q = 0; q < 8; q++) H[q * cap + i] = r.p[q];
            if (r.kind == 2) for (int q = 0; q < B_N; q++) B[q * cap + i] = r.p[q];
        }
        for (int i = 0; i < cap; i++) {
            // dissipation of the Hertz model (contactKernel.cu:239-240)
            const double logR = std::log(H[H_RES * cap + i]);
            H[H_D * cap + i] = -logR / std::sqrt(logR * logR + M_PI * M_PI);
        }
        tH.alloc(H.size()); tL.alloc(L.size()); tB.alloc(B.size());
        // Uploads are ordered on the scene's stream (whatever the runtime
        // does with pageable sources); waiting keeps the host arrays alive.
        DEMSPH_HIP(hipMemcpyAsync(tH.p, H.data(), H.size() * 8, hipMemcpyHostToDevice, stream));
        DEMSPH_HIP(hipMemcpyAsync(tL.p, L.data(), L.size() * 8, hipMemcpyHostToDevice, stream));
        DEMSPH_HIP(hipMemcpyAsync(tB.p, B.data(), B.size() * 8, hipMemcpyHostToDevice, stream));
        DEMSPH_HIP(hipStreamSynchronize(stream));
    }

    // Upload on the scene's stream; h must outlive the copy (the host
    // staging members do, until initialize() synchronises).
    template <class T>
    void up(DBuf<T>& d, const std::vector<T>& h, size_t count)
    {
        d.alloc(count);
        if (count) DEMSPH_HIP(hipMemcpyAsync(d.p, h.data(), count * sizeof(T), hipMemcpyHostToDevice, stream));
    }
    void up3(DBuf<double>& x, DBuf<double>& y, DBuf<double>& z, const std::vector<double>& h, size_t count)
    {
        std::vector<double> a(count), b(count), c(count);
        for (size_t i = 0; i < count; i++) { a[i] = h[3 * i]; b[i] = h[3 * i + 1]; c[i] = h[3 * i + 2]; }
        up(x, a, count); up(y, b, count); up(z, c, count);
        DEMSPH_HIP(hipStreamSynchronize(stream));  // a, b, c are local
    }

    void upload_bonds()
    {
        // new bonds are appended to the device arrays (existing state kept)
        const int nb_new = (int)bond_a.size();
        if (nb_new == nbond_dev) return;
        for (int k = nbond_dev; k < nb_new; k++)
            if (bond_a[k] >= n || bond_b[k] >= n) throw InvalidArg("bond references a ball index beyond the scene");
        std::vector<double> keep[16];
        std::vector<int> keep_bonded;
        DBuf<double>* st[16] = {&dbpx, &dbpy, &dbpz, &dbnx, &dbny, &dbnz, &dbsx, &dbsy, &dbsz, &dbtx, &dbty, &dbtz,
                                &dbFn, &dbTt, &dbsig, &dbtau};
        for (int k = 0; k < 16; k++) {
            keep[k].assign(nb_new, 0.0);
            if (nbond_dev) copy_sync(keep[k].data(), st[k]->p, nbond_dev * 8, hipMemcpyDeviceToHost, stream);
        }
        keep_bonded.assign(nb_new, 1);
        if (nbond_dev) copy_sync(keep_bonded.data(), dbonded.p, nbond_dev * 4, hipMemcpyDeviceToHost, stream);
        for (int k = nbond_dev; k < nb_new; k++) {
            keep[0][k] = bond_point[3 * k]; keep[1][k] = bond_point[3 * k + 1]; keep[2][k] = bond_point[3 * k + 2];
            keep[3][k] = bond_normal[3 * k]; keep[4][k] = bond_normal[3 * k + 1]; keep[5][k] = bond_normal[3 * k + 2];
        }
        for (int k = 0; k < 16; k++) up(*st[k], keep[k], nb_new);
        up(dbonded, keep_bonded, nb_new);
        up(dbi, bond_a, nb_new);
        up(dbj, bond_b, nb_new);
        for (auto* b : {&dbFx, &dbFy, &dbFz, &dbTix, &dbTiy, &dbTiz, &dbTjx, &dbTjy, &dbTjz}) b->alloc(nb_new);
        // incidence by insertion index: (bond id, side) in bond order
        std::vector<int> ist(n + 1, 0), ib(2 * (size_t)nb_new);
        for (int k = 0; k < nb_new; k++) { ist[bond_a[k] + 1]++; ist[bond_b[k] + 1]++; }
        for (int i = 0; i < n; i++) ist[i + 1] += ist[i];
        std::vector<int> cur(ist.begin(), ist.end() - 1);
        for (int k = 0; k < nb_new; k++) {
            ib[cur[bond_a[k]]++] = 2 * k;
            ib[cur[bond_b[k]]++] = 2 * k + 1;
        }
        up(dinc_start, ist, n + 1);
        up(dinc_bond, ib, ib.size());
        nbond_dev = nb_new;
        nbond = nb_new;
    }

    // Device bytes initialize() allocates for the staged scene: double-
    // buffered ball state, sort buffers, cell ranges and the two CSR pair
    // lists with their history at the initial capacity of 6 entries per ball
    // (DEMSolver.h:174), plus a margin for walls and clumps.  Used to admit
    // scenes under a memory budget.
    size_t bytes_estimate(const double* minb, const double* maxb) const
    {
        const size_t nb = hb.size();
        double c = 0.0;
        for (double r : hb.r) c = std::max(c, r * 2.0);
        GridGeom gg{};
        const size_t cells = grid_set(gg, minb, maxb, c) ? (size_t)gg.numGrids() : 0;
        const size_t per_ball = 2 * (11 * 8 + 3 * 4)          // state, two buffers
                                + 2 * 4 + 12 * 8              // ranks, force and torque accumulators
                                + 5 * 4 + 4 * 4 + 2 * 4       // sort keys, counts and scans, pair starts
                                + 6 * (2 * (4 + 16 * 8) + 4); // pairs with history (two lists), reverse index
        return nb * per_ball + 2 * cells * sizeof(int) + radix_sort_temp_bytes((int)std::max<size_t>(nb, 1)) +
               hw.t0.size() * 256 + hc.size() * 512 + (size_t(1) << 20);
    }

    void initialize(const double* minb, const double* maxb)
    {
        DEMSPH_HIP(hipStreamSynchronize(stream));
        build_tables();
        if (initialized) {
            upload_bonds();
            return;
        }
        n = (int)hb.size();
        nc = (int)hc.size();
        nw = (int)hw.mat.size();
        nv = (int)hw.vwall.size();
        nt = (int)hw.t0.size();
        if (n == 0) throw InvalidArg("no balls");
        // grids (DEMSolver.h:245-263)
        double c = 0.0;
        for (double r : hb.r) c = std::max(c, r * 2.0);
        bgValid = grid_set(bg, minb, maxb, c);
        if (!bgValid) throw InvalidArg("degenerate domain or ball radius");
        for (int t = 0; t < nt; t++) {
            const double* vl = hw.vlocal.data();
            const double* a = &vl[3 * (size_t)hw.t0[t]];
            const double* b = &vl[3 * (size_t)hw.t1[t]];
            const double* cc = &vl[3 * (size_t)hw.t2[t]];
            double m[3];
            circumcenter(a, b, cc, m);
            for (const double* p : {a, b, cc}) {
                const double d = 2.0 * std::sqrt((p[0] - m[0]) * (p[0] - m[0]) + (p[1] - m[1]) * (p[1] - m[1]) +
                                                 (p[2] - m[2]) * (p[2] - m[2]));
                c = std::max(c, d);
            }
        }
        if (nt > 0) {
            tgValid = grid_set(tg, minb, maxb, c);
            if (!tgValid) throw InvalidArg("degenerate triangle grid");
        }
        bng = bg.numGrids();
        tng = nt > 0 ? tg.numGrids() : 0;
        per.L[0] = bg.maxx - bg.minx; per.L[1] = bg.maxy - bg.miny; per.L[2] = bg.maxz - bg.minz;
        {
            const int ncell[3] = {bg.nx, bg.ny, bg.nz};
            for (int k = 0; k < 3; k++)
                if (per.on[k] && ncell[k] < 3)
                    throw InvalidArg("a periodic axis needs at least three ball cells (a wrapped stencil)");
        }
        if (bng > INT32_MAX / 2 || tng > INT32_MAX / 2) throw InvalidArg("grid too large");
        // balls in insertion order (slot == index)
        for (int b = 0; b < 2; b++) {
            for (auto* d : {&bpx[b], &bpy[b], &bpz[b], &bvx[b], &bvy[b], &bvz[b], &bwx[b], &bwy[b], &bwz[b], &br[b], &bim[b]})
                d->alloc(n);
            for (auto* d : {&bmat[b], &bclump[b], &borig[b]}) d->alloc(n);
        }
        cb = 0;
        up3(bpx[0], bpy[0], bpz[0], hb.pos, n);
        up3(bvx[0], bvy[0], bvz[0], hb.vel, n);
        up3(bwx[0], bwy[0], bwz[0], hb.w, n);
        up(br[0], hb.r, n);
        up(bim[0], hb.im, n);
        up(bmat[0], hb.mat, n);
        up(bclump[0], hb.clump, n);
        rank.alloc(n); rank_old.alloc(n);
        iota_i32(borig[0].p, n, stream);
        iota_i32(rank.p, n, stream);
        for (auto* d : {&fx, &fy, &fz, &tx, &ty, &tz}) d->alloc(n);
        key.alloc(n); skey.alloc(n); ktmp.alloc(n); perm.alloc(n); vtmp.alloc(n);
        DEMSPH_HIP(dev_malloc(&sort_temp, radix_sort_temp_bytes(n)));
        cellStart.alloc(bng); cellEnd.alloc(bng);
        cnt.alloc(n + 1); rcnt.alloc(n + 1); rstart.alloc(n + 1); rcursor.alloc(n);
        DEMSPH_HIP(dev_malloc(&scan_temp, scan_temp_bytes(n + 1)));
        total.alloc(2);
        DEMSPH_HIP(host_malloc((void**)&total_h, 2 * sizeof(int)));
        for (int w = 0; w < 2; w++) {
            pstart[w].alloc(n + 1);
            alloc_pairs(w, 6 * n);  // initial capacity as DEMSolver.h:174
        }
        rpair.alloc(6 * n);
        has_list = false;
        // clumps
        if (nc > 0) {
            up3(cpx, cpy, cpz, hc.pos, nc);
            up3(cvx, cvy, cvz, hc.vel, nc);
            up3(cwx, cwy, cwz, hc.w, nc);
            std::vector<double> q0(nc, 1.0), qz(nc, 0.0);
            up(cq0, q0, nc); up(cq1, qz, nc); up(cq2, qz, nc); up(cq3, qz, nc);
            std::vector<double> comp[6];
            for (int k = 0; k < 6; k++) {
                comp[k].resize(nc);
                for (int c2 = 0; c2 < nc; c2++) comp[k][c2] = hc.invI[6 * c2 + k];
            }
            up(cixx, comp[0], nc); up(ciyy, comp[1], nc); up(cizz, comp[2], nc);
            up(cixy, comp[3], nc); up(cixz, comp[4], nc); up(ciyz, comp[5], nc);
            up(cim, hc.im, nc);
            up(cps, hc.ps, nc); up(cpe, hc.pe, nc);
            for (auto* d : {&cFx, &cFy, &cFz, &cTx, &cTy, &cTz}) d->alloc(nc);
        }
        // walls
        if (nt > 0) {
            up3(wpx, wpy, wpz, hw.pos, nw);
            std::vector<double> z(nw, 0.0), one(nw, 1.0);
            up(wvx, z, nw); up(wvy, z, nw); up(wvz, z, nw);
            up(wwx, z, nw); up(wwy, z, nw); up(wwz, z, nw);
            up(wq0, one, nw); up(wq1, z, nw); up(wq2, z, nw); up(wq3, z, nw);
            up(wmat, hw.mat, nw);
            up3(vlx, vly, vlz, hw.vlocal, nv);
            std::vector<double> vg(3 * (size_t)nv);
            for (int v = 0; v < nv; v++)
                for (int k = 0; k < 3; k++) vg[3 * v + k] = hw.vlocal[3 * v + k] + hw.pos[3 * hw.vwall[v] + k];
            up3(vgx, vgy, vgz, vg, nv);
            up(vwall, hw.vwall, nv);
            up(dt0, hw.t0, nt); up(dt1, hw.t1, nt); up(dt2, hw.t2, nt); up(dtwall, hw.twall, nt);
            up(tcc, hw.tcc, 3 * (size_t)nt);
            tkey.alloc(nt); tskey.alloc(nt); tktmp.alloc(nt); tperm.alloc(nt); tvtmp.alloc(nt);
            tbox.alloc(6 * (size_t)nt);
            tri_cells_valid = false;
            walls_static = true;
            DEMSPH_HIP(dev_malloc(&tsort_temp, radix_sort_temp_bytes(nt)));
            tcs.alloc(tng); tce.alloc(tng);
            tcand_cnt.alloc(tng + 1); tcand_start.alloc(tng + 1);
            DEMSPH_HIP(dev_malloc(&tscan_temp, scan_temp_bytes((int)tng + 1)));
            rmax = 0.0;
            for (double rr : hb.r) rmax = std::max(rmax, rr);
            const size_t hn = (size_t)KW * n;
            hcnt.alloc(n);
            DEMSPH_HIP(hipMemsetAsync(hcnt.p, 0, n * sizeof(int), stream));
            htri.alloc(hn); hcancel.alloc(hn);
            for (auto* d : {&hsx, &hsy, &hsz, &hrx, &hry, &hrz, &hox, &hoy, &hoz, &hov}) d->alloc(hn);
            for (auto* d : {&wfx, &wfy, &wfz, &wtx, &wty, &wtz}) d->alloc(n);
            if (record_contacts) crec.alloc(12 * hn);
        }
        overflow.alloc(1);
        DEMSPH_HIP(hipMemsetAsync(overflow.p, 0, sizeof(int), stream));
        DEMSPH_HIP(host_malloc((void**)&overflow_h, sizeof(int)));
        pair_over.alloc(1);
        fill_i32(pair_over.p, 1, 0, stream);
        DEMSPH_HIP(host_malloc((void**)&block_h, 3 * sizeof(int)));
        nbond_dev = 0;
        upload_bonds();
        DEMSPH_HIP(hipStreamSynchronize(stream));
        initialized = true;
    }

    void require_init() const
    {
        if (!initialized) throw InvalidArg("scene not initialized");
    }

    // ---------------------------------------------------------- stages
    template <bool PER>
    void count_pairs_as(int lanes, dim3 gl, const Balls& nb)
    {
        const dim3 tb(kT);
        if (lanes == 8)
            DEMSPH_LAUNCH(timer, "dem_count_pairs", (k_count_pairs<8, PER>), gl, tb, 0, stream, nb, bg, cellStart.p,
                          cellEnd.p, cnt.p, per);
        else if (lanes == 4)
            DEMSPH_LAUNCH(timer, "dem_count_pairs", (k_count_pairs<4, PER>), gl, tb, 0, stream, nb, bg, cellStart.p,
                          cellEnd.p, cnt.p, per);
        else if (lanes == 2)
            DEMSPH_LAUNCH(timer, "dem_count_pairs", (k_count_pairs<2, PER>), gl, tb, 0, stream, nb, bg, cellStart.p,
                          cellEnd.p, cnt.p, per);
        else
            DEMSPH_LAUNCH(timer, "dem_count_pairs", (k_count_pairs<1, PER>), gl, tb, 0, stream, nb, bg, cellStart.p,
                          cellEnd.p, cnt.p, per);
    }
    void launch_count_pairs(int lanes, dim3 gl, const Balls& nb)
    {
        if (periodic) count_pairs_as<true>(lanes, gl, nb);
        else count_pairs_as<false>(lanes, gl, nb);
    }
    template <bool PER>
    void write_pairs_as(int lanes, dim3 gl, const WritePairsArgs& wa)
    {
        const dim3 tb(kT);
        if (lanes == 8) DEMSPH_LAUNCH(timer, "dem_write_pairs", (k_write_pairs<8, PER>), gl, tb, 0, stream, wa);
        else if (lanes == 4) DEMSPH_LAUNCH(timer, "dem_write_pairs", (k_write_pairs<4, PER>), gl, tb, 0, stream, wa);
        else if (lanes == 2) DEMSPH_LAUNCH(timer, "dem_write_pairs", (k_write_pairs<2, PER>), gl, tb, 0, stream, wa);
        else DEMSPH_LAUNCH(timer, "dem_write_pairs", (k_write_pairs<1, PER>), gl, tb, 0, stream, wa);
    }
    void launch_write_pairs(int lanes, dim3 gl, const WritePairsArgs& wa)
    {
        if (periodic) write_pairs_as<true>(lanes, gl, wa);
        else write_pairs_as<false>(lanes, gl, wa);
    }

    // neighborSearch (DEMSolver.h:698-777)
    // sync_resize: read the pair total back and grow the lists before writing
    // them (resizePair, interaction.h:164-175), as the API's neighbour search
    // does.  Otherwise (inside step blocks) nothing is read back: a list
    // longer than its capacity is truncated and noted in pair_over, and the
    // block is replayed from its checkpoint with larger lists.
    void neighbor_search(bool sync_resize = true)
    {
        const dim3 gb(blocks_for(n, kT)), tb(kT);
        Balls b = cur();
        if (periodic) DEMSPH_LAUNCH(timer, "dem_wrap_periodic", k_wrap_periodic, gb, tb, 0, stream, b, bg, per);
        timer.begin("dem_hash_sort", stream);
        hash_soa_by_orig(n, bg, b.px, b.py, b.pz, b.orig, key.p, stream);
        radix_sort_pairs(key.p, nullptr, skey.p, perm.p, ktmp.p, vtmp.p, sort_temp, n, key_bits(bng), stream);
        fill2_i32(cellStart.p, cellEnd.p, bng, -1, stream);
        timer.end(stream);
        ReorderArgs ra{};
        ra.src = b;
        ra.dst = balls(1 - cb);
        ra.skey = skey.p; ra.perm = perm.p; ra.rank = rank.p; ra.numGrids = bng;
        ra.rank_old = rank_old.p; ra.rcnt = rcnt.p; ra.rcursor = rcursor.p;
        ra.cellStart = cellStart.p; ra.cellEnd = cellEnd.p;
        DEMSPH_LAUNCH(timer, "dem_reorder", k_reorder, gb, tb, 0, stream, ra);
        const int old_cb = cb;
        cb = 1 - cb;
        Balls nb = cur();
        const int lanes = pair_walk_lanes(n, pair_lanes_set);
        const dim3 gl(blocks_for((long long)n * lanes, kT));
        launch_count_pairs(lanes, gl, nb);
        // the owner CSR goes into the list that becomes current
        const int np = 1 - cp;
        // start[n] = total, written by the scan
        exclusive_scan(cnt.p, pstart[np].p, n, scan_temp, pstart[np].p + n, stream);
        int tot = -1;  // on the device only
        if (sync_resize) {
            DEMSPH_HIP(hipMemcpyAsync(total_h, pstart[np].p + n, sizeof(int), hipMemcpyDeviceToHost, stream));
            DEMSPH_HIP(hipStreamSynchronize(stream));  // resizePair (interaction.h:164-175)
            timer.harvest();
            tot = total_h[0];
            if (tot > pcap[np]) alloc_pairs(np, tot + tot / 4);
            if (tot > (int)rpair.n) rpair.alloc(std::max(pcap[0], pcap[1]));
        } else {
            hipLaunchKernelGGL(k_pair_capacity, dim3(1), dim3(64), 0, stream, pstart[np].p + n, pcap[np], pair_over.p);
            DEMSPH_HIP(hipGetLastError());
        }
        WritePairsArgs wa{};
        wa.b = nb; wa.g = bg; wa.cs = cellStart.p; wa.ce = cellEnd.p;
        wa.cur = pairs(np);
        wa.old = pairs(cp);
        wa.old_orig = borig[old_cb].p;
        wa.rank_old = rank_old.p;
        wa.has_old = has_list ? 1 : 0;
        wa.per = per;
        launch_write_pairs(lanes, gl, wa);
        cp = np;
        has_list = true;
        npairs = tot;
        // reverse index by partner (DEMSolver.h:59-78, reset every build)
        timer.begin("dem_reverse_index", stream);
        // rcnt and rcursor were cleared by k_reorder
        hipLaunchKernelGGL(k_rev_count, gb, tb, 0, stream, n, pstart[cp].p, pcap[cp], ppartner[cp].p, rcnt.p);
        exclusive_scan(rcnt.p, rstart.p, n + 1, scan_temp, nullptr, stream);
        hipLaunchKernelGGL(k_rev_fill, gb, tb, 0, stream, n, pstart[cp].p, pcap[cp], ppartner[cp].p, rstart.p,
                           rcursor.p, rpair.p);
        hipLaunchKernelGGL(k_rev_sort, gb, tb, 0, stream, n, rstart.p, rpair.p);
        DEMSPH_HIP(hipGetLastError());
        timer.end(stream);
    }

    // Ball-pair and bond forces plus the reduction: calculateBallContactForceTorque
    // (DEMSolver.h:599-695); with walls, ball-wall contacts are added and the
    // walls then move (addTriangleContactForceTorque, :428-597).
    // hook_t: the time of the external-force hooks that follow this
    // evaluation (NaN: none); the built-in hook program then runs inside the
    // gather, and a user hook after it.
    void evaluate(double springDt, double wallDt, bool do_pairs = true, bool do_walls = true,
                  double hook_t = std::numeric_limits<double>::quiet_NaN())
    {
        const bool walls = do_walls && nt > 0;
        if (do_pairs) evaluate_pairs(springDt, walls, wallDt);
        else if (walls) evaluate_walls(wallDt);
        gather(nt > 0 && (walls || walls_valid), hook_t);
        if (walls) {
            walls_valid = true;
            if (!walls_static) move_walls(wallDt);
        }
    }

    PairForceArgs pair_args(double springDt) const
    {
        PairForceArgs pa{};
        pa.b = cur(); pa.p = pairs(cp); pa.tb = tables(); pa.dt = springDt;
        pa.rec = record_contacts ? prec.p : nullptr;
        pa.per = per;
        return pa;
    }

    BondArgs bond_args(double springDt) const
    {
        BondArgs ba{};
        ba.nb = nbond; ba.b = cur(); ba.rank = rank.p; ba.bi = dbi.p; ba.bj = dbj.p; ba.bonded = dbonded.p;
        ba.bpx = dbpx.p; ba.bpy = dbpy.p; ba.bpz = dbpz.p; ba.bnx = dbnx.p; ba.bny = dbny.p; ba.bnz = dbnz.p;
        ba.bsx = dbsx.p; ba.bsy = dbsy.p; ba.bsz = dbsz.p; ba.btx = dbtx.p; ba.bty = dbty.p; ba.btz = dbtz.p;
        ba.bFn = dbFn.p; ba.bTt = dbTt.p; ba.bsig = dbsig.p; ba.btau = dbtau.p;
        ba.Fx = dbFx.p; ba.Fy = dbFy.p; ba.Fz = dbFz.p;
        ba.Tix = dbTix.p; ba.Tiy = dbTiy.p; ba.Tiz = dbTiz.p; ba.Tjx = dbTjx.p; ba.Tjy = dbTjy.p; ba.Tjz = dbTjz.p;
        ba.tb = tables(); ba.dt = springDt;
        ba.periodic = periodic ? 1 : 0; ba.per = per;
        return ba;
    }

    template <class T, int WALLS>
    void forces_launch(bool bonds, const PairForceArgs& pa, const WallArgs& wa, const BondArgs& ba, int nbp, int nbw,
                       int nbb)
    {
        // blocks per XCD of each range (k_forces)
        const int qp = (nbp + kNumXcd - 1) / kNumXcd, qw = (nbw + kNumXcd - 1) / kNumXcd,
                  qb = (nbb + kNumXcd - 1) / kNumXcd;
        const dim3 gb(kNumXcd * (qp + qw + qb)), tb(kT);
        if (bonds) DEMSPH_LAUNCH(timer, "dem_forces", (k_forces<T, WALLS, true>), gb, tb, 0, stream, pa, wa, ba, qp, qw, qb);
        else DEMSPH_LAUNCH(timer, "dem_forces", (k_forces<T, WALLS, false>), gb, tb, 0, stream, pa, wa, ba, qp, qw, qb);
    }
    template <class T>
    void forces_as(int walls, bool bonds, const PairForceArgs& pa, const WallArgs& wa, const BondArgs& ba, int nbp,
                   int nbw, int nbb)
    {
        if (walls == 3) forces_launch<T, 3>(bonds, pa, wa, ba, nbp, nbw, nbb);
        else if (walls == 2) forces_launch<T, 2>(bonds, pa, wa, ba, nbp, nbw, nbb);
        else if (walls == 1) forces_launch<T, 1>(bonds, pa, wa, ba, nbp, nbw, nbb);
        else forces_launch<T, 0>(bonds, pa, wa, ba, nbp, nbw, nbb);
    }

    // Ball-pair forces, and with them the bond forces and (with_walls) the
    // ball-wall contacts, in one launch.
    void evaluate_pairs(double springDt, bool with_walls = false, double wallDt = 0.0)
    {
        const int nbp = blocks_for(pcap[cp], kT);  // one thread per list entry
        const PairForceArgs pa = pair_args(springDt);
        WallArgs wa{};
        int walls = 0, nbw = 0;
        if (with_walls) {
            prepare_walls();
            wa = wall_args(wallDt);
            walls = KW > kWallCap ? 3 : wall_slots_set && n <= kWallSlotBalls ? 2 : 1;
            nbw = blocks_for(n, kT);
        }
        BondArgs ba{};
        const int nbb = nbond > 0 ? blocks_for(nbond, kT) : 0;
        if (nbb > 0) ba = bond_args(springDt);
        if (fp64) forces_as<double>(walls, nbb > 0, pa, wa, ba, nbp, nbw, nbb);
        else forces_as<float>(walls, nbb > 0, pa, wa, ba, nbp, nbw, nbb);
    }

    // Triangle cells and per-cell candidate lists, when the walls moved.
    void prepare_walls()
    {
        {
            if (!tri_cells_valid) {
                timer.begin("dem_triangle_cells", stream);
                hash_aos(tcc.p, nt, tg, tkey.p, stream);
                radix_sort_pairs(tkey.p, nullptr, tskey.p, tperm.p, tktmp.p, tvtmp.p, tsort_temp, nt, key_bits(tng),
                                 stream);
                fill2_i32(tcs.p, tce.p, tng, -1, stream);
                cell_ranges(tskey.p, nt, tng, tcs.p, tce.p, stream);
                hipLaunchKernelGGL(k_tri_boxes, dim3(blocks_for(nt, 256)), dim3(256), 0, stream, nt, tperm.p, dt0.p,
                                   dt1.p, dt2.p, vgx.p, vgy.p, vgz.p, reinterpret_cast<double2*>(tbox.p));
                DEMSPH_HIP(hipGetLastError());
                build_tri_candidates();
                timer.end(stream);
                tri_cells_valid = walls_static;
            }
        }
    }

    WallArgs wall_args(double wallDt) const
    {
        const Balls b = cur();
        {
            WallArgs wa{};
            wa.b = b; wa.tb = tables(); wa.dt = wallDt; wa.tg = tg; wa.tcs = tcs.p; wa.tce = tce.p; wa.tsorted = tperm.p;
            wa.tbox = reinterpret_cast<const double2*>(tbox.p);
            wa.cstart = tcand_start.p; wa.ctri = tcand_tri.p;
            wa.cbox = reinterpret_cast<const double2*>(tcand_box.p);
            wa.t0 = dt0.p; wa.t1 = dt1.p; wa.t2 = dt2.p; wa.twall = dtwall.p;
            wa.vgx = vgx.p; wa.vgy = vgy.p; wa.vgz = vgz.p;
            wa.wpx = wpx.p; wa.wpy = wpy.p; wa.wpz = wpz.p; wa.wvx = wvx.p; wa.wvy = wvy.p; wa.wvz = wvz.p;
            wa.wwx = wwx.p; wa.wwy = wwy.p; wa.wwz = wwz.p; wa.wmat = wmat.p;
            wa.KW = KW; wa.nballs = n;
            wa.rec = record_contacts ? crec.p : nullptr;
            wa.hcnt = hcnt.p; wa.htri = htri.p; wa.hcancel = hcancel.p;
            wa.hsx = hsx.p; wa.hsy = hsy.p; wa.hsz = hsz.p; wa.hrx = hrx.p; wa.hry = hry.p; wa.hrz = hrz.p;
            wa.hox = hox.p; wa.hoy = hoy.p; wa.hoz = hoz.p; wa.hov = hov.p;
            wa.ofx = wfx.p; wa.ofy = wfy.p; wa.ofz = wfz.p; wa.otx = wtx.p; wa.oty = wty.p; wa.otz = wtz.p;
            wa.overflow = overflow.p;
            return wa;
        }
    }

    void evaluate_walls(double wallDt)
    {
        const dim3 gb(blocks_for(n, kT)), tb(kT);
        prepare_walls();
        const WallArgs wa = wall_args(wallDt);
        if (KW > kWallCap) {
            if (fp64) DEMSPH_LAUNCH(timer, "dem_wall_contacts", (k_wall_contacts<double, kWallCapMax>), gb, tb, 0, stream, wa);
            else DEMSPH_LAUNCH(timer, "dem_wall_contacts", (k_wall_contacts<float, kWallCapMax>), gb, tb, 0, stream, wa);
        } else if (fp64) {
            DEMSPH_LAUNCH(timer, "dem_wall_contacts", (k_wall_contacts<double, kWallCap>), gb, tb, 0, stream, wa);
        } else {
            DEMSPH_LAUNCH(timer, "dem_wall_contacts", (k_wall_contacts<float, kWallCap>), gb, tb, 0, stream, wa);
        }
    }

    // Per-cell triangle candidate lists (count, scan, one host read of the
    // total for the capacity, write).
    void build_tri_candidates()
    {
        CellCandArgs ca{};
        ca.tg = tg; ca.ntc = (int)tng;
        const double cmax = std::max(tg.cx, std::max(tg.cy, tg.cz));
        ca.reach = rmax * (1.0 + 1e-9) + 1e-9 * cmax;
        ca.tcs = tcs.p; ca.tce = tce.p; ca.tsorted = tperm.p; ca.tbox = reinterpret_cast<const double2*>(tbox.p);
        ca.count = tcand_cnt.p;
        const dim3 gb(blocks_for(tng, kT)), tb(kT);
        hipLaunchKernelGGL(k_cell_candidates<false>, gb, tb, 0, stream, ca);
        DEMSPH_HIP(hipGetLastError());
        exclusive_scan(tcand_cnt.p, tcand_start.p, (int)tng, tscan_temp, tcand_start.p + tng, stream);
        DEMSPH_HIP(hipMemcpyAsync(total_h + 1, tcand_start.p + tng, sizeof(int), hipMemcpyDeviceToHost, stream));
        DEMSPH_HIP(hipStreamSynchronize(stream));
        const size_t tot = (size_t)total_h[1];
        if (tot > tcand_tri.n || tcand_tri.p == nullptr) {
            tcand_tri.alloc(std::max<size_t>(tot, 1));
            tcand_box.alloc(6 * std::max<size_t>(tot, 1));
        }
        ca.start = tcand_start.p; ca.out_tri = tcand_tri.p; ca.out_box = reinterpret_cast<double2*>(tcand_box.p);
        hipLaunchKernelGGL(k_cell_candidates<true>, gb, tb, 0, stream, ca);
        DEMSPH_HIP(hipGetLastError());
    }

    GatherArgs gather_args(bool with_walls) const
    {
        const Balls b = cur();
        GatherArgs ga{};
        ga.b = b;
        ga.pstart = pstart[cp].p; ga.pcap = pcap[cp];
        ga.pIx = pIx[cp].p; ga.pIy = pIy[cp].p; ga.pIz = pIz[cp].p;
        ga.pFx = pFx[cp].p; ga.pFy = pFy[cp].p; ga.pFz = pFz[cp].p; ga.pRx = pRx[cp].p; ga.pRy = pRy[cp].p; ga.pRz = pRz[cp].p;
        ga.rstart = rstart.p; ga.rpair = rpair.p;
        if (nbond > 0) {
            ga.istart = dinc_start.p; ga.ibond = dinc_bond.p;
            ga.bFx = dbFx.p; ga.bFy = dbFy.p; ga.bFz = dbFz.p;
            ga.bTix = dbTix.p; ga.bTiy = dbTiy.p; ga.bTiz = dbTiz.p; ga.bTjx = dbTjx.p; ga.bTjy = dbTjy.p; ga.bTjz = dbTjz.p;
        }
        if (with_walls) {
            ga.wfx = wfx.p; ga.wfy = wfy.p; ga.wfz = wfz.p; ga.wtx = wtx.p; ga.wty = wty.p; ga.wtz = wtz.p;
        }
        ga.out = forces();
        return ga;
    }

    void gather(bool with_walls, double hook_t = std::numeric_limits<double>::quiet_NaN())
    {
        const dim3 gb(blocks_for(n, kT)), tb(kT);
        const GatherArgs ga = gather_args(with_walls);
        const bool hooked = !std::isnan(hook_t);
        if (hooked && !hook_fn && !hooks.empty() && hooks.size() <= (size_t)kMaxHooksFused) {
            const HookTable ht = hook_table(hook_t);
            if (fp64) DEMSPH_LAUNCH(timer, "dem_gather", k_gather_hooks<double>, gb, tb, 0, stream, ga, ht);
            else DEMSPH_LAUNCH(timer, "dem_gather", k_gather_hooks<float>, gb, tb, 0, stream, ga, ht);
            return;
        }
        if (fp64) DEMSPH_LAUNCH(timer, "dem_gather", k_gather<double>, gb, tb, 0, stream, ga);
        else DEMSPH_LAUNCH(timer, "dem_gather", k_gather<float>, gb, tb, 0, stream, ga);
        if (hooked) run_hooks(hook_t);
    }

    // The built-in hook program at time t, for the gather (the constant
    // force's ramp factor is taken on the host, as run_hooks does).
    HookTable hook_table(double t) const
    {
        HookTable ht{};
        ht.n = (int)hooks.size();
        for (int k = 0; k < ht.n; k++) {
            const HookOp& h = hooks[k];
            HookDev& o = ht.op[k];
            o.kind = h.kind;
            if (h.kind == 0) {
                double sc = 1.0;
                if (h.ramp > 0) sc = std::min(t / h.ramp, 1.0);
                o.s = sc;
                o.fo = (const double*)h.f->p;
            } else if (h.kind == 1) {
                o.c = h.c;
            } else {
                o.rho = h.rho; o.cd = h.cd; o.level = h.level;
                o.g = D3{h.g[0], h.g[1], h.g[2]};
                o.vf = D3{h.vf[0], h.vf[1], h.vf[2]};
            }
        }
        return ht;
    }

    void move_walls(double dt)
    {
        WallMoveArgs m{};
        m.nw = nw; m.nv = nv; m.nt = nt;
        m.px = wpx.p; m.py = wpy.p; m.pz = wpz.p; m.q0 = wq0.p; m.q1 = wq1.p; m.q2 = wq2.p; m.q3 = wq3.p;
        m.vx = wvx.p; m.vy = wvy.p; m.vz = wvz.p; m.wx = wwx.p; m.wy = wwy.p; m.wz = wwz.p;
        m.lx = vlx.p; m.ly = vly.p; m.lz = vlz.p; m.gx = vgx.p; m.gy = vgy.p; m.gz = vgz.p;
        m.vwall = vwall.p; m.t0 = dt0.p; m.t1 = dt1.p; m.t2 = dt2.p; m.tcc = tcc.p; m.dt = dt;
        timer.begin("dem_wall_motion", stream);
        hipLaunchKernelGGL(k_wall_move, dim3(blocks_for(nw, 64)), dim3(64), 0, stream, m);
        hipLaunchKernelGGL(k_vertex_update, dim3(blocks_for(nv, 256)), dim3(256), 0, stream, m);
        hipLaunchKernelGGL(k_circumcenters, dim3(blocks_for(nt, 256)), dim3(256), 0, stream, m);
        DEMSPH_HIP(hipGetLastError());
        timer.end(stream);
    }

    void run_hooks(double t)
    {
        const dim3 gb(blocks_for(n, kT)), tb(kT);
        Balls b = cur();
        for (auto& h : hooks) {
            if (h.kind == 0) {
                double s = 1.0;
                if (h.ramp > 0) s = std::min(t / h.ramp, 1.0);
                hipLaunchKernelGGL(k_hook_const, gb, tb, 0, stream, b, forces(), (const double*)h.f->p, s);
            } else if (h.kind == 1) {
                hipLaunchKernelGGL(k_hook_damping, gb, tb, 0, stream, b, forces(), h.c);
            } else {
                hipLaunchKernelGGL(k_hook_buoyancy, gb, tb, 0, stream, b, forces(), h.rho, h.cd, D3{h.g[0], h.g[1], h.g[2]},
                                   D3{h.vf[0], h.vf[1], h.vf[2]}, h.level);
            }
            DEMSPH_HIP(hipGetLastError());
        }
        if (hook_fn) {
            ensure_views();
            ViewArgs v = view_args();
            hipLaunchKernelGGL(k_views_scatter, gb, tb, 0, stream, v);
            DEMSPH_HIP(hipGetLastError());
            hook_fn(hook_user, iStep, t);
            hipLaunchKernelGGL(k_views_gather, gb, tb, 0, stream, v);
            DEMSPH_HIP(hipGetLastError());
        }
    }

    void ensure_views()
    {
        if (vF.p) return;
        for (auto* d : {&vF, &vT, &vP, &vV, &vW}) d->alloc(3 * (size_t)n);
        vR.alloc(n);
        vIM.alloc(n);
    }
    ViewArgs view_args() const
    {
        return ViewArgs{cur(), forces(), vF.p, vT.p, vP.p, vV.p, vW.p, vR.p, vIM.p};
    }

    void half(bool first, const double* g, double h)
    {
        const dim3 tb(kT);
        if (nc > 0) {
            ClumpArgs c{};
            c.nc = nc;
            c.px = cpx.p; c.py = cpy.p; c.pz = cpz.p; c.vx = cvx.p; c.vy = cvy.p; c.vz = cvz.p;
            c.wx = cwx.p; c.wy = cwy.p; c.wz = cwz.p; c.q0 = cq0.p; c.q1 = cq1.p; c.q2 = cq2.p; c.q3 = cq3.p;
            c.ixx = cixx.p; c.iyy = ciyy.p; c.izz = cizz.p; c.ixy = cixy.p; c.ixz = cixz.p; c.iyz = ciyz.p;
            c.im = cim.p; c.ps = cps.p; c.pe = cpe.p;
            c.Fx = cFx.p; c.Fy = cFy.p; c.Fz = cFz.p; c.Tx = cTx.p; c.Ty = cTy.p; c.Tz = cTz.p;
            c.b = cur(); c.f = forces(); c.rank = rank.p;
            c.gx = g[0]; c.gy = g[1]; c.gz = g[2]; c.h = h;
            if (first) DEMSPH_LAUNCH(timer, "dem_clump_half", k_clump_half<true>, dim3(blocks_for(nc, kT)), tb, 0, stream, c);
            else DEMSPH_LAUNCH(timer, "dem_clump_half", k_clump_half<false>, dim3(blocks_for(nc, kT)), tb, 0, stream, c);
        }
        const BallHalfArgs a = half_args(g, h);
        if (first) DEMSPH_LAUNCH(timer, "dem_ball_half", k_ball_half<true>, dim3(blocks_for(n, kT)), tb, 0, stream, a);
        else DEMSPH_LAUNCH(timer, "dem_ball_half", k_ball_half<false>, dim3(blocks_for(n, kT)), tb, 0, stream, a);
    }

    BallHalfArgs half_args(const double* g, double h) const
    {
        BallHalfArgs a{};
        a.b = cur(); a.f = forces();
        if (nc > 0) {
            a.cpx = cpx.p; a.cpy = cpy.p; a.cpz = cpz.p; a.cvx = cvx.p; a.cvy = cvy.p; a.cvz = cvz.p;
            a.cwx = cwx.p; a.cwy = cwy.p; a.cwz = cwz.p;
        }
        a.nc = nc;
        a.gx = g[0]; a.gy = g[1]; a.gz = g[2]; a.h = h;
        return a;
    }

    // Without clumps or hooks, one force evaluation and the half kick that
    // follows it run as two launches (pairs with walls, gather with kick)
    // instead of four; each ball's arithmetic is unchanged.
    bool fused_ok() const { return nc == 0 && hooks.empty() && !hook_fn; }

    void evaluate_half(bool first, const double* g, double h, double dt)
    {
        const dim3 gb(blocks_for(n, kT)), tb(kT);
        const bool walls = nt > 0;
        evaluate_pairs(h, walls, dt);
        const GatherArgs ga = gather_args(walls);
        const BallHalfArgs ha = half_args(g, h);
        if (fp64) {
            if (first) DEMSPH_LAUNCH(timer, "dem_gather_half", (k_gather_half<double, true>), gb, tb, 0, stream, ga, ha);
            else DEMSPH_LAUNCH(timer, "dem_gather_half", (k_gather_half<double, false>), gb, tb, 0, stream, ga, ha);
        } else {
            if (first) DEMSPH_LAUNCH(timer, "dem_gather_half", (k_gather_half<float, true>), gb, tb, 0, stream, ga, ha);
            else DEMSPH_LAUNCH(timer, "dem_gather_half", (k_gather_half<float, false>), gb, tb, 0, stream, ga, ha);
        }
        if (walls) {
            walls_valid = true;
            if (!walls_static) move_walls(dt);
        }
    }

    // The largest wall-candidate count of the balls that overflowed the
    // capacity since the last call (0: none), read back and cleared.
    int take_wall_overflow()
    {
        if (nt == 0) return 0;
        DEMSPH_HIP(hipMemcpyAsync(overflow_h, overflow.p, sizeof(int), hipMemcpyDeviceToHost, stream));
        DEMSPH_HIP(hipStreamSynchronize(stream));
        timer.harvest();
        const int c = overflow_h[0];
        if (c > 0) fill_i32(overflow.p, 1, 0, stream);
        return c;
    }

    // Wall-contact history arrays with room for `need` candidates per ball
    // (entry q of ball o stays at q * n + o, so the kept entries are copied
    // as they are).  More than kWallCapMax is an error.
    void grow_walls(int need)
    {
        if (need > kWallCapMax)
            throw CapacityError("a ball had " + std::to_string(need) +
                                " wall-triangle candidates, above the per-ball maximum of " +
                                std::to_string(kWallCapMax));
        if (need <= KW) return;
        const size_t keep = (size_t)KW * n, m = (size_t)kWallCapMax * n;
        std::vector<void*> old;
        const auto grow = [&](auto& b, size_t per) {
            using E = std::remove_reference_t<decltype(*b.p)>;
            E* q = dmalloc<E>(m * per);
            DEMSPH_HIP(hipMemcpyAsync(q, b.p, keep * per * sizeof(E), hipMemcpyDeviceToDevice, stream));
            old.push_back(b.p);
            b.p = q;
            b.n = m * per;
        };
        grow(htri, 1);
        grow(hcancel, 1);
        for (auto* b : {&hsx, &hsy, &hsz, &hrx, &hry, &hrz, &hox, &hoy, &hoz, &hov}) grow(*b, 1);
        if (record_contacts && crec.p) grow(crec, 12);
        // the copies read the old blocks: they are released only after them
        DEMSPH_HIP(hipStreamSynchronize(stream));
        for (void* q : old) dev_free(q);
        KW = kWallCapMax;
        drop_graphs();  // captured launches hold the old arrays and capacity
    }

    // Wall history arrays for kw candidates per ball, contents not kept (a
    // state import overwrites them).
    void set_wall_capacity(int kw)
    {
        if (nt == 0) {
            KW = kw;
            return;
        }
        const size_t hn = (size_t)kw * n;
        htri.alloc(hn); hcancel.alloc(hn);
        for (auto* b : {&hsx, &hsy, &hsz, &hrx, &hry, &hrz, &hox, &hoy, &hoz, &hov}) b->alloc(hn);
        if (record_contacts && crec.p) crec.alloc(12 * hn);
        KW = kw;
        drop_graphs();
    }

    // A wall-candidate overflow is an error here (paths that cannot replay).
    void check_overflow()
    {
        const int c = take_wall_overflow();
        if (c > KW)
            throw CapacityError("a ball had " + std::to_string(c) + " wall-triangle candidates, above the per-ball capacity of " +
                                std::to_string(KW) + " (this block's wall contacts are incomplete)");
    }

    void solve_begin()
    {
        require_init();
        neighbor_search();
        // the first evaluation has zero time step: running it again after the
        // wall capacity grew changes nothing but the completed wall contacts
        for (;;) {
            evaluate(0.0, 0.0);
            const int c = take_wall_overflow();
            if (c <= KW) break;
            grow_walls(c);
        }
        iStep = 0;
        time = 0.0;
    }

    // Steps between rebuilds can replay a captured graph when nothing in
    // them involves the host: no hooks (they take host times or call back),
    // static or no walls (moving walls re-sort their triangles with a host
    // read), a stream of the scene's own (the null stream cannot be
    // captured) and no kernel timing.
    bool graphs_ok() const
    {
        return use_graphs && !timer.enabled && stream != nullptr && hooks.empty() && !hook_fn &&
               (nt == 0 || (walls_static && tri_cells_valid && walls_valid));
    }

    void plain_steps(const double* g, double dt, int count)
    {
        const double h = 0.5 * dt;
        for (int k = 0; k < count; k++) {
            if (fused_ok()) {
                evaluate_half(true, g, h, dt);
                evaluate_half(false, g, h, dt);
            } else {
                evaluate(h, dt);
                half(true, g, h);
                evaluate(h, dt);
                half(false, g, h);
            }
            step_stamp(stamps, stream);
        }
    }

    // kind 4: the four steps after a rebuild.  kind 5: a rebuild with lists
    // sized on the device and the five steps of its cycle; capturing runs
    // the host side of the rebuild (buffer flips), which is put back: the
    // caller flips after launching.
    hipGraphExec_t cycle_graph(int kind, const double* g, double dt)
    {
        GraphKey key{};
        key.dt = dt; key.gx = g[0]; key.gy = g[1]; key.gz = g[2];
        key.kind = kind; key.cb = cb; key.cp = cp; key.nbond = nbond; key.fp64 = fp64 ? 1 : 0;
        key.cap0 = pcap[0]; key.cap1 = pcap[1];
        key.pairs = ppartner[cp].p; key.pairs1 = ppartner[1 - cp].p; key.rev = rpair.p; key.cand = tcand_tri.p;
        key.stamp = stamps.buf;
        key.prec = record_contacts ? prec.p : nullptr;
        key.crec = record_contacts ? crec.p : nullptr;
        for (auto& e : graphs)
            if (e.key == key) return e.exec;
        const int cb0 = cb, cp0 = cp, np0 = npairs;
        const bool has0 = has_list;
        hipGraph_t gr = nullptr;
        // No other thread may allocate, free or synchronise the device
        // until the graph is instantiated (see capture_lock).
        CaptureGuard capture_guard(capture_lock());
        DEMSPH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            if (kind == 5) neighbor_search(false);
            plain_steps(g, dt, kind);
        } catch (...) {
            (void)hipStreamEndCapture(stream, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            cb = cb0; cp = cp0; npairs = np0; has_list = has0;
            throw;
        }
        cb = cb0; cp = cp0; npairs = np0; has_list = has0;
        DEMSPH_HIP(hipStreamEndCapture(stream, &gr));
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
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

    // One block of loop_ball / loop_ball_wall (DEMSolver.h:1376-1405,
    // 1452-1485) with the pair lists sized on the device: whole cycles of a
    // rebuild and five steps replay one graph, the four steps after a
    // rebuild another, other steps run eagerly.
    void run_block(const double* g, double dt, int nsteps)
    {
        const double h = 0.5 * dt;
        for (int k = 0; k < nsteps;) {
            if (graphs_ok() && iStep % 5 == 4 && nsteps - k >= 5) {
                DEMSPH_HIP(hipGraphLaunch(cycle_graph(5, g, dt), stream));
                for (int q = 0; q < 5; q++) {
                    iStep++;
                    time += dt;
                }
                cb = 1 - cb;
                cp = 1 - cp;
                has_list = true;
                npairs = -1;
                k += 5;
                continue;
            }
            // right after a rebuild (or at the start) the next four steps rebuild nothing
            if (graphs_ok() && iStep % 5 == 0 && nsteps - k >= 4) {
                DEMSPH_HIP(hipGraphLaunch(cycle_graph(4, g, dt), stream));
                for (int q = 0; q < 4; q++) {
                    iStep++;
                    time += dt;
                }
                k += 4;
                continue;
            }
            iStep++;
            time += dt;
            if (iStep % 5 == 0) {
                // a host callback must not see its steps twice: with one,
                // rebuilds size their lists on the host and nothing replays
                neighbor_search(hook_fn != nullptr);
                if (!hook_fn) npairs = -1;
            }
            if (fused_ok()) {
                evaluate_half(true, g, h, dt);
                evaluate_half(false, g, h, dt);
            } else {
                evaluate(h, dt, true, true, time - h);  // with the hooks at t - dt/2
                half(true, g, h);
                evaluate(h, dt, true, true, time);
                half(false, g, h);
            }
            step_stamp(stamps, stream);
            k++;
        }
    }

    // nsteps steps with one host synchronisation at the end.  A block that
    // rebuilds the pair lists, or that has walls, saves a device checkpoint
    // first; a pair list that outgrew its capacity, or a ball with more
    // wall-triangle candidates than the wall capacity, sends the block back
    // to the checkpoint with the lists or the wall history grown to fit, and
    // it runs again.  (With a host hook nothing is replayed: the rebuilds
    // size their lists on the host, and a wall overflow is an error.)
    void step(const double* g, double dt, int nsteps)
    {
        require_init();
        if (!has_list) throw InvalidArg("solve_begin() (or neighbor_search) must precede step()");
        if (!(dt > 0)) throw InvalidArg("time step must be positive");
        if (nsteps <= 0) return;
        const bool rebuilds = !hook_fn && nsteps >= 5 - (int)(iStep % 5);
        const bool walls = nt > 0;
        const bool replayable = !hook_fn && (rebuilds || walls);
        for (int attempt = 0; attempt < 8; attempt++) {
            if (replayable) {
                dem_block_save(this);
                fill_i32(pair_over.p, 1, 0, stream);
            }
            run_block(g, dt, nsteps);
            DEMSPH_HIP(hipMemcpyAsync(block_h, pair_over.p, sizeof(int), hipMemcpyDeviceToHost, stream));
            DEMSPH_HIP(hipMemcpyAsync(block_h + 1, pstart[cp].p + n, sizeof(int), hipMemcpyDeviceToHost, stream));
            if (walls) DEMSPH_HIP(hipMemcpyAsync(block_h + 2, overflow.p, sizeof(int), hipMemcpyDeviceToHost, stream));
            DEMSPH_HIP(hipStreamSynchronize(stream));
            timer.harvest();
            const int need = rebuilds ? block_h[0] : 0;
            const int wneed = walls ? block_h[2] : 0;  // nonzero only above the capacity
            if (wneed > 0) fill_i32(overflow.p, 1, 0, stream);
            if (need == 0 && wneed == 0) {
                npairs = block_h[1];
                return;
            }
            if (!replayable)
                throw CapacityError("a ball had " + std::to_string(wneed) +
                                    " wall-triangle candidates, above the per-ball capacity of " + std::to_string(KW) +
                                    " (a scene with a host hook does not replay its steps)");
            // back to the checkpoint with the lists or the wall history grown
            if (need > 0) pair_replays++;
            if (wneed > 0) wall_replays++;
            drop_graphs();
            dem_block_restore(this, need > 0 ? need + need / 4 + 64 : 0);
            if (wneed > 0) grow_walls(wneed);  // above the maximum: an error, the state at the checkpoint
        }
        throw std::runtime_error("pair list and wall capacities did not converge");
    }
};

// ------------------------------------------------------------ internal API
namespace demsph {

DemBallRefs dem_ball_refs(demsph_dem* d)
{
    const Balls b = d->cur();
    const Forces f = d->forces();
    return DemBallRefs{b.n, b.px, b.py, b.pz, b.vx, b.vy, b.vz, b.wx, b.wy, b.wz, b.r, b.im, b.orig,
                       f.fx, f.fy, f.fz, f.tx, f.ty, f.tz};
}
hipStream_t dem_stream(demsph_dem* d) { return d->stream; }
void dem_attach_couple(demsph_dem* d, demsph_couple* c) { d->couples.push_back(c); }
void dem_detach_couple(demsph_dem* d, demsph_couple* c)
{
    d->couples.erase(std::remove(d->couples.begin(), d->couples.end(), c), d->couples.end());
}
void dem_set_stream(demsph_dem* d, hipStream_t s) { d->stream = s; }
void dem_require_ready(demsph_dem* d)
{
    d->require_init();
    if (!d->has_list) throw InvalidArg("solve_begin() (or neighbor_search) must precede coupled stepping");
}
double dem_begin_step(demsph_dem* d, double dt)
{
    d->iStep++;
    d->time += dt;
    if (d->iStep % 5 == 0) d->neighbor_search();
    return d->time;
}
void dem_evaluate(demsph_dem* d, double h, double dt) { d->evaluate(h, dt); }
void dem_hooks_and_half(demsph_dem* d, bool first, const double* g, double h, double hookTime)
{
    d->run_hooks(hookTime);
    d->half(first, g, h);
}
void dem_end_block(demsph_dem* d) { d->check_overflow(); }
int dem_take_wall_overflow(demsph_dem* d) { return d->take_wall_overflow(); }
void dem_grow_walls(demsph_dem* d, int need) { d->grow_walls(need); }
size_t dem_bytes_estimate(demsph_dem* d, const double* minb, const double* maxb)
{
    return d->bytes_estimate(minb, maxb);
}
int dem_num_balls(demsph_dem* d) { return d->initialized ? d->n : (int)d->hb.size(); }
void dem_set_stamps(demsph_dem* d, const StepStamps& st)
{
    d->drop_graphs();  // captured graphs hold the previous stamp buffer
    d->stamps = st;
}

}  // namespace demsph

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
    if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(T));
}

// symMatrix inverse (myMat.h:163-179), host side for addClump.
void sym_inverse(const double* A, double* r)
{
    const double xx = A[0], yy = A[1], zz = A[2], xy = A[3], xz = A[4], yz = A[5];
    const double det_inv = 1.0 / (xx * (yy * zz - yz * yz) + xy * (yz * xz - xy * zz) + xz * (xy * yz - yy * xz));
    r[0] = (yy * zz - yz * yz) * det_inv;
    r[1] = (xx * zz - xz * xz) * det_inv;
    r[2] = (xx * yy - xy * xy) * det_inv;
    r[3] = -(xy * zz - xz * yz) * det_inv;
    r[4] = (xy * yz - xz * yy) * det_inv;
    r[5] = -(xx * yz - xz * xy) * det_inv;
}

}  // namespace

#define DEM_GUARD(...) return guarded([&] { if (!d) throw InvalidArg("null scene"); __VA_ARGS__; })

extern "C" int demsph_dem_create(demsph_dem** out, void* stream)
{
    return guarded([&] {
        if (!out) throw InvalidArg("null output handle");
        auto* d = new demsph_dem();
        d->stream = static_cast<hipStream_t>(stream);
        *out = d;
    });
}

extern "C" int demsph_dem_destroy(demsph_dem* d) { return guarded([&] { delete d; }); }

static void add_row(demsph_dem* d, int kind, int a, int b, std::initializer_list<double> v)
{
    if (d->initialized) throw InvalidArg("material rows must be set before initialize()");
    Row r{};
    r.kind = kind; r.a = a; r.b = b;
    int k = 0;
    for (double x : v) r.p[k++] = x;
    d->rows.push_back(r);
}

extern "C" int demsph_dem_set_linear_pair(demsph_dem* d, int a, int b, double kn, double ks, double kr, double kt,
                                          double dn, double ds, double dr, double dt, double mus, double mur, double mut)
{
    DEM_GUARD(add_row(d, 0, a, b, {kn, ks, kr, kt, dn, ds, dr, dt, mus, mur, mut}));
}
extern "C" int demsph_dem_set_hertzian_pair(demsph_dem* d, int a, int b, double E, double G, double e, double krks,
                                            double ktks, double mus, double mur, double mut)
{
    DEM_GUARD(add_row(d, 1, a, b, {E, G, e, krks, ktks, mus, mur, mut}));
}
extern "C" int demsph_dem_set_bonded_pair(demsph_dem* d, int a, int b, double gamma, double E, double knks,
                                          double sigmat, double C, double mu)
{
    DEM_GUARD(add_row(d, 2, a, b, {gamma, E, knks, sigmat, C, mu}));
}

extern "C" int demsph_dem_add_balls(demsph_dem* d, size_t n, const double* position, const double* velocity,
                                    const double* angularVelocity, const double* radius, const double* density,
                                    const int* materialID, const int* clumpID)
{
    DEM_GUARD({
        if (d->initialized) throw InvalidArg("balls must be added before initialize()");
        if (n && (!position || !radius || !density || !materialID)) throw InvalidArg("null array");
        for (size_t i = 0; i < n; i++) {
            const double r = radius[i];
            const double mass = 4. / 3. * r * r * r * M_PI * density[i];
            const double inv = mass > 0.0 ? 1.0 / mass : 0.0;
            d->hb.add(&position[3 * i], velocity ? &velocity[3 * i] : nullptr,
                      angularVelocity ? &angularVelocity[3 * i] : nullptr, r, inv, materialID[i],
                      clumpID ? clumpID[i] : -1);
        }
    });
}

extern "C" int demsph_dem_add_fixed_balls(demsph_dem* d, size_t n, const double* position, const double* radius,
                                          const int* materialID, const int* clumpID)
{
    DEM_GUARD({
        if (d->initialized) throw InvalidArg("balls must be added before initialize()");
        if (n && (!position || !radius || !materialID)) throw InvalidArg("null array");
        for (size_t i = 0; i < n; i++)
            d->hb.add(&position[3 * i], nullptr, nullptr, radius[i], 0.0, materialID[i], clumpID ? clumpID[i] : -1);
    });
}

extern "C" int demsph_dem_add_clump(demsph_dem* d, size_t np, const double* points, const double* radii,
                                    const double* centroid, const double* velocity, const double* angularVelocity,
                                    double mass, const double* inertia, int materialID)
{
    DEM_GUARD({
        if (d->initialized) throw InvalidArg("clumps must be added before initialize()");
        if (!points || !radii || !centroid || !velocity || !angularVelocity || !inertia) throw InvalidArg("null array");
        const int cid = (int)d->hc.size();
        const int ps = (int)d->hb.size();
        double volume = 0.0, inv = 0.0;
        std::vector<double> pv(3 * np);
        for (size_t i = 0; i < np; i++)
            for (int k = 0; k < 3; k++) pv[3 * i + k] = velocity[k];
        const double* w = angularVelocity;
        if (mass > 1.e-20) {
            inv = 1.0 / mass;
            for (size_t i = 0; i < np; i++) {
                volume += 4.0 / 3.0 * M_PI * std::pow(radii[i], 3.0);
                const double rx = points[3 * i] - centroid[0], ry = points[3 * i + 1] - centroid[1],
                             rz = points[3 * i + 2] - centroid[2];
                pv[3 * i] += w[1] * rz - w[2] * ry;
                pv[3 * i + 1] += w[2] * rx - w[0] * rz;
                pv[3 * i + 2] += w[0] * ry - w[1] * rx;
            }
        }
        const double dens = volume > 0. ? mass / volume : 0.0;
        for (size_t i = 0; i < np; i++) {
            const double r = radii[i];
            const double m = 4. / 3. * r * r * r * M_PI * dens;
            d->hb.add(&points[3 * i], &pv[3 * i], w, r, m > 0.0 ? 1.0 / m : 0.0, materialID, cid);
        }
        for (int k = 0; k < 3; k++) {
            d->hc.pos.push_back(centroid[k]);
            d->hc.vel.push_back(velocity[k]);
            d->hc.w.push_back(w[k]);
        }
        double ii[6];
        sym_inverse(inertia, ii);
        for (int k = 0; k < 6; k++) d->hc.invI.push_back(ii[k]);
        d->hc.im.push_back(inv);
        d->hc.mat.push_back(materialID);
        d->hc.ps.push_back(ps);
        d->hc.pe.push_back(ps + (int)np);
    });
}

extern "C" int demsph_dem_add_fixed_clump(demsph_dem* d, size_t np, const double* points, const double* radii,
                                          const double* centroid, int materialID)
{
    DEM_GUARD({
        if (d->initialized) throw InvalidArg("clumps must be added before initialize()");
        if (!points || !radii || !centroid) throw InvalidArg("null array");
        const int cid = (int)d->hc.size();
        const int ps = (int)d->hb.size();
        for (size_t i = 0; i < np; i++) d->hb.add(&points[3 * i], nullptr, nullptr, radii[i], 0.0, materialID, cid);
        for (int k = 0; k < 3; k++) {
            d->hc.pos.push_back(centroid[k]);
            d->hc.vel.push_back(0.0);
            d->hc.w.push_back(0.0);
        }
        for (int k = 0; k < 6; k++) d->hc.invI.push_back(0.0);
        d->hc.im.push_back(0.0);
        d->hc.mat.push_back(materialID);
        d->hc.ps.push_back(ps);
        d->hc.pe.push_back(ps + (int)np);
    });
}

extern "C" int demsph_dem_add_mesh_wall(demsph_dem* d, size_t nv, const double* vertices, size_t nt, const int* tri0,
                                        const int* tri1, const int* tri2, const double* position,
                                        const double* velocity, const double* angularVelocity, int materialID)
{
    DEM_GUARD({
        (void)velocity;
        (void)angularVelocity;  // the reference stores zero wall velocities (DEMSolver.h:1901-1911)
        if (d->initialized) throw InvalidArg("walls must be added before initialize()");
        if (nt == 0) return;
        if (!vertices || !tri0 || !tri1 || !tri2 || !position) throw InvalidArg("null array");
        int mx = -1;
        for (size_t i = 0; i < nt; i++) mx = std::max(mx, std::max(tri0[i], std::max(tri1[i], tri2[i])));
        if (mx >= (int)nv) return;  // DEMSolver.h:1891-1897: ignored
        const int vs = (int)d->hw.vwall.size();
        const int wi = (int)d->hw.mat.size();
        for (int k = 0; k < 3; k++) d->hw.pos.push_back(position[k]);
        d->hw.mat.push_back(materialID);
        for (size_t i = 0; i < nv; i++) {
            for (int k = 0; k < 3; k++) d->hw.vlocal.push_back(vertices[3 * i + k]);
            d->hw.vwall.push_back(wi);
        }
        for (size_t i = 0; i < nt; i++) {
            double a[3], b[3], c[3], m[3];
            for (int k = 0; k < 3; k++) {
                a[k] = vertices[3 * (size_t)tri0[i] + k] + position[k];
                b[k] = vertices[3 * (size_t)tri1[i] + k] + position[k];
                c[k] = vertices[3 * (size_t)tri2[i] + k] + position[k];
            }
            circumcenter(a, b, c, m);
            d->hw.t0.push_back(tri0[i] + vs);
            d->hw.t1.push_back(tri1[i] + vs);
            d->hw.t2.push_back(tri2[i] + vs);
            d->hw.twall.push_back(wi);
            for (int k = 0; k < 3; k++) d->hw.tcc.push_back(m[k]);
        }
    });
}

extern "C" int demsph_dem_add_bonded_interactions(demsph_dem* d, size_t n, const int* ob0, const int* ob1,
                                                  size_t npos, const double* pos)
{
    DEM_GUARD({
        if (n && (!ob0 || !ob1 || !pos)) throw InvalidArg("null array");
        for (size_t k = 0; k < n; k++) {
            int a = ob0[k], b = ob1[k];
            if (a < 0 || b < 0) continue;
            if ((size_t)a >= npos || (size_t)b >= npos) continue;
            if (a == b) continue;
            if (a > b) std::swap(a, b);
            if (!d->bond_set.insert({a, b}).second) continue;
            d->bond_a.push_back(a);
            d->bond_b.push_back(b);
            const double* pa = &pos[3 * (size_t)a];
            double pb[3] = {pos[3 * (size_t)b], pos[3 * (size_t)b + 1], pos[3 * (size_t)b + 2]};
            if (d->periodic && d->initialized) {
                // a bond across a periodic face joins a to the nearest image of b
                for (int q = 0; q < 3; q++) {
                    if (!d->per.on[q]) continue;
                    const double dd = pa[q] - pb[q];
                    if (dd > 0.5 * d->per.L[q]) pb[q] += d->per.L[q];
                    else if (dd < -0.5 * d->per.L[q]) pb[q] -= d->per.L[q];
                }
            }
            for (int q = 0; q < 3; q++) d->bond_point.push_back(0.5 * (pa[q] + pb[q]));
            const double nx = pa[0] - pb[0], ny = pa[1] - pb[1], nz = pa[2] - pb[2];
            const double l2 = nx * nx + ny * ny + nz * nz;
            const double inv = l2 < 1.0e-20 ? 0.0 : 1.0 / std::sqrt(l2);
            d->bond_normal.push_back(nx * inv);
            d->bond_normal.push_back(ny * inv);
            d->bond_normal.push_back(nz * inv);
        }
    });
}

extern "C" int demsph_dem_set_precision(demsph_dem* d, int fp64Forces) { DEM_GUARD(d->fp64 = fp64Forces ? 1 : 0); }

extern "C" int demsph_dem_initialize(demsph_dem* d, const double* minBound, const double* maxBound)
{
    DEM_GUARD({
        if (!minBound || !maxBound) throw InvalidArg("null bounds");
        d->initialize(minBound, maxBound);
    });
}

extern "C" int demsph_dem_neighbor_search(demsph_dem* d) { DEM_GUARD(d->require_init(); d->neighbor_search()); }

extern "C" int demsph_dem_contact(demsph_dem* d, double timeStep)
{
    DEM_GUARD({
        d->require_init();
        if (!d->has_list) throw InvalidArg("neighbor_search() must precede contact()");
        // ball pairs and bonds; forces are overwritten (the reference clears them first)
        d->walls_valid = false;
        d->evaluate(timeStep, 0.0, true, false);
    });
}

extern "C" int demsph_dem_triangle_contact(demsph_dem* d, double timeStep)
{
    DEM_GUARD({
        d->require_init();
        if (!d->has_list) throw InvalidArg("neighbor_search() must precede triangle_contact()");
        d->evaluate(0.0, timeStep, false, true);
    });
}

extern "C" int demsph_dem_first_half(demsph_dem* d, const double* gravity, double halfTimeStep)
{
    DEM_GUARD(d->require_init(); d->half(true, gravity, halfTimeStep));
}
extern "C" int demsph_dem_second_half(demsph_dem* d, const double* gravity, double halfTimeStep)
{
    DEM_GUARD(d->require_init(); d->half(false, gravity, halfTimeStep));
}
extern "C" int demsph_dem_solve_begin(demsph_dem* d) { DEM_GUARD(d->solve_begin()); }
extern "C" int demsph_dem_step(demsph_dem* d, const double* gravity, double timeStep, int nsteps)
{
    DEM_GUARD({
        if (!gravity) throw InvalidArg("null gravity");
        d->step(gravity, timeStep, nsteps);
    });
}
extern "C" int demsph_dem_synchronize(demsph_dem* d)
{
    DEM_GUARD(DEMSPH_HIP(hipStreamSynchronize(d->stream)); d->timer.harvest());
}
extern "C" int demsph_dem_clock(demsph_dem* d, size_t* iStep, double* time)
{
    DEM_GUARD({
        if (iStep) *iStep = d->iStep;
        if (time) *time = d->time;
    });
}
extern "C" int demsph_dem_set_hook(demsph_dem* d, demsph_dem_hook_fn fn, void* user)
{
    DEM_GUARD(d->hook_fn = fn; d->hook_user = user);
}
extern "C" int demsph_dem_views(demsph_dem* d, demsph_dem_ball_views* v)
{
    DEM_GUARD({
        d->require_init();
        if (!v) throw InvalidArg("null views");
        d->ensure_views();
        v->force = d->vF.p; v->torque = d->vT.p; v->position = d->vP.p; v->velocity = d->vV.p;
        v->angularVelocity = d->vW.p; v->radius = d->vR.p; v->inverseMass = d->vIM.p; v->numBalls = d->n;
    });
}
extern "C" int demsph_dem_hook_clear(demsph_dem* d)
{
    DEM_GUARD({
        for (auto* b : d->hook_bufs) delete b;
        d->hook_bufs.clear();
        d->hooks.clear();
    });
}
extern "C" int demsph_dem_hook_constant_force(demsph_dem* d, const double* force, double ramp)
{
    DEM_GUARD({
        d->require_init();
        if (!force) throw InvalidArg("null force");
        auto* b = new DBuf<double>();
        d->hook_bufs.push_back(b);
        b->alloc(3 * (size_t)d->n);
        copy_sync(b->p, force, 3 * (size_t)d->n * 8, hipMemcpyHostToDevice, d->stream);
        demsph_dem::HookOp h{};
        h.kind = 0; h.ramp = ramp; h.f = b;
        d->hooks.push_back(h);
    });
}
extern "C" int demsph_dem_hook_global_damping(demsph_dem* d, double c)
{
    DEM_GUARD({
        demsph_dem::HookOp h{};
        h.kind = 1; h.c = c;
        d->hooks.push_back(h);
    });
}
extern "C" int demsph_dem_hook_buoyancy_drag(demsph_dem* d, double rho, double cd, const double* g, const double* vf,
                                             double level)
{
    DEM_GUARD({
        if (!g || !vf) throw InvalidArg("null vector");
        demsph_dem::HookOp h{};
        h.kind = 2; h.rho = rho; h.cd = cd; h.level = level;
        for (int k = 0; k < 3; k++) { h.g[k] = g[k]; h.vf[k] = vf[k]; }
        d->hooks.push_back(h);
    });
}

extern "C" int demsph_dem_sizes(demsph_dem* d, size_t* nb, size_t* nc, size_t* nw, size_t* nt, size_t* np, size_t* nbd)
{
    DEM_GUARD({
        if (nb) *nb = d->initialized ? d->n : d->hb.size();
        if (nc) *nc = d->initialized ? d->nc : d->hc.size();
        if (nw) *nw = d->initialized ? d->nw : d->hw.mat.size();
        if (nt) *nt = d->initialized ? d->nt : d->hw.t0.size();
        if (np) *np = d->npairs;
        if (nbd) *nbd = d->bond_a.size();
    });
}

extern "C" int demsph_dem_get(demsph_dem* d, const char* field, void* out, size_t capacity, size_t* count)
{
    DEM_GUARD({
        d->require_init();
        if (!field) throw InvalidArg("null field");
        const std::string f = field;
        hipStream_t st = d->stream;
        const int cb = d->cb;
        const size_t n = d->n;
        auto orig = fetch(d->borig[cb].p, n, st);
        auto vec = [&](const DBuf<double>& x, const DBuf<double>& y, const DBuf<double>& z) {
            auto hx = fetch(x.p, n, st), hy = fetch(y.p, n, st), hz = fetch(z.p, n, st);
            std::vector<double> v(3 * n);
            for (size_t i = 0; i < n; i++) {
                v[3 * (size_t)orig[i]] = hx[i]; v[3 * (size_t)orig[i] + 1] = hy[i]; v[3 * (size_t)orig[i] + 2] = hz[i];
            }
            return v;
        };
        auto scal = [&](const DBuf<double>& x) {
            auto hx = fetch(x.p, n, st);
            std::vector<double> v(n);
            for (size_t i = 0; i < n; i++) v[orig[i]] = hx[i];
            return v;
        };
        auto plain3 = [&](const DBuf<double>& x, const DBuf<double>& y, const DBuf<double>& z, size_t m) {
            auto hx = fetch(x.p, m, st), hy = fetch(y.p, m, st), hz = fetch(z.p, m, st);
            std::vector<double> v(3 * m);
            for (size_t i = 0; i < m; i++) { v[3 * i] = hx[i]; v[3 * i + 1] = hy[i]; v[3 * i + 2] = hz[i]; }
            return v;
        };
        auto plain4 = [&](const DBuf<double>& a, const DBuf<double>& b, const DBuf<double>& c, const DBuf<double>& e,
                          size_t m) {
            auto ha = fetch(a.p, m, st), hb = fetch(b.p, m, st), hc = fetch(c.p, m, st), he = fetch(e.p, m, st);
            std::vector<double> v(4 * m);
            for (size_t i = 0; i < m; i++) { v[4 * i] = ha[i]; v[4 * i + 1] = hb[i]; v[4 * i + 2] = hc[i]; v[4 * i + 3] = he[i]; }
            return v;
        };
        if (f == "position") return emit(vec(d->bpx[cb], d->bpy[cb], d->bpz[cb]), out, capacity, count);
        if (f == "velocity") return emit(vec(d->bvx[cb], d->bvy[cb], d->bvz[cb]), out, capacity, count);
        if (f == "angularVelocity") return emit(vec(d->bwx[cb], d->bwy[cb], d->bwz[cb]), out, capacity, count);
        if (f == "force") return emit(vec(d->fx, d->fy, d->fz), out, capacity, count);
        if (f == "torque") return emit(vec(d->tx, d->ty, d->tz), out, capacity, count);
        if (f == "radius") return emit(scal(d->br[cb]), out, capacity, count);
        if (f == "inverseMass") return emit(scal(d->bim[cb]), out, capacity, count);
        if (f == "hashValue") {
            auto k = fetch(d->skey.p, n, st);
            return emit(std::vector<int>(k.begin(), k.end()), out, capacity, count);
        }
        if (f == "hashIndex") return emit(orig, out, capacity, count);
        if (f == "cellHashStart") return emit(fetch(d->cellStart.p, (size_t)d->bng, st), out, capacity, count);
        if (f == "cellHashEnd") return emit(fetch(d->cellEnd.p, (size_t)d->bng, st), out, capacity, count);
        const bool pairf = f == "objectPointed" || f == "objectPointing" || f == "slidingSpring" ||
                           f == "rollingSpring" || f == "torsionSpring" || f == "overlap" || f == "pairShift";
        if (pairf) {
            if (!d->has_list) throw InvalidArg("no pair list yet");
            const int c = d->cp;
            auto start = fetch(d->pstart[c].p, n + 1, st);
            auto part = fetch(d->ppartner[c].p, (size_t)d->npairs, st);
            std::vector<int> rankh(n);
            for (size_t i = 0; i < n; i++) rankh[orig[i]] = (int)i;
            std::vector<double> s0, s1, s2;
            int comps = 0;
            if (f == "slidingSpring") { s0 = fetch(d->psx[c].p, d->npairs, st); s1 = fetch(d->psy[c].p, d->npairs, st); s2 = fetch(d->psz[c].p, d->npairs, st); comps = 3; }
            if (f == "rollingSpring") { s0 = fetch(d->prx[c].p, d->npairs, st); s1 = fetch(d->pry[c].p, d->npairs, st); s2 = fetch(d->prz[c].p, d->npairs, st); comps = 3; }
            if (f == "torsionSpring") { s0 = fetch(d->pox[c].p, d->npairs, st); s1 = fetch(d->poy[c].p, d->npairs, st); s2 = fetch(d->poz[c].p, d->npairs, st); comps = 3; }
            if (f == "overlap") { s0 = fetch(d->pov[c].p, d->npairs, st); comps = 1; }
            std::vector<uint8_t> shifts;
            if (f == "pairShift" && d->periodic) shifts = fetch(d->pshift[c].p, (size_t)d->npairs, st);
            std::vector<int> vi;
            std::vector<double> vd;
            for (size_t o = 0; o < n; o++) {
                const int i = rankh[o];
                for (int p = start[i]; p < start[i + 1]; p++) {
                    if (f == "objectPointed") vi.push_back((int)o);
                    else if (f == "objectPointing") vi.push_back(orig[part[p]]);
                    else if (f == "pairShift") vi.push_back(d->periodic ? (int)shifts[p] : kNoShift);
                    else {
                        vd.push_back(s0[p]);
                        if (comps == 3) { vd.push_back(s1[p]); vd.push_back(s2[p]); }
                    }
                }
            }
            if (f == "objectPointed" || f == "objectPointing" || f == "pairShift") return emit(vi, out, capacity, count);
            return emit(vd, out, capacity, count);
        }
        const bool tris = f == "triSlidingSpring" || f == "triRollingSpring" || f == "triTorsionSpring";
        const bool trif = f == "triPointed" || f == "triPointing" || f == "triCancel" || tris || f == "triOverlap";
        if (trif) {
            std::vector<int> vi;
            std::vector<double> vd;
            if (d->nt > 0) {
                const size_t hn = (size_t)d->KW * n;
                auto hc = fetch(d->hcnt.p, n, st);
                auto ht = fetch(d->htri.p, hn, st);
                auto hcan = fetch(d->hcancel.p, hn, st);
                std::vector<double> sx, sy, sz, ov;
                if (f == "triSlidingSpring") { sx = fetch(d->hsx.p, hn, st); sy = fetch(d->hsy.p, hn, st); sz = fetch(d->hsz.p, hn, st); }
                if (f == "triRollingSpring") { sx = fetch(d->hrx.p, hn, st); sy = fetch(d->hry.p, hn, st); sz = fetch(d->hrz.p, hn, st); }
                if (f == "triTorsionSpring") { sx = fetch(d->hox.p, hn, st); sy = fetch(d->hoy.p, hn, st); sz = fetch(d->hoz.p, hn, st); }
                if (f == "triOverlap") ov = fetch(d->hov.p, hn, st);
                for (size_t o = 0; o < n; o++)
                    for (int c = 0; c < hc[o]; c++) {
                        const size_t h = (size_t)c * n + o;
                        if (f == "triPointed") vi.push_back((int)o);
                        else if (f == "triPointing") vi.push_back(ht[h]);
                        else if (f == "triCancel") vi.push_back(hcan[h]);
                        else if (f == "triOverlap") vd.push_back(ov[h]);
                        else { vd.push_back(sx[h]); vd.push_back(sy[h]); vd.push_back(sz[h]); }
                    }
            }
            if (tris || f == "triOverlap") return emit(vd, out, capacity, count);
            return emit(vi, out, capacity, count);
        }
        const size_t nb = d->nbond;
        if (f == "bondPointed") return emit(d->bond_a, out, capacity, count);
        if (f == "bondPointing") return emit(d->bond_b, out, capacity, count);
        if (f == "bondIsBonded") return emit(fetch(d->dbonded.p, nb, st), out, capacity, count);
        if (f == "bondNormalForce") return emit(fetch(d->dbFn.p, nb, st), out, capacity, count);
        if (f == "bondTorsionTorque") return emit(fetch(d->dbTt.p, nb, st), out, capacity, count);
        if (f == "bondMaxNormalStress") return emit(fetch(d->dbsig.p, nb, st), out, capacity, count);
        if (f == "bondMaxShearStress") return emit(fetch(d->dbtau.p, nb, st), out, capacity, count);
        if (f == "bondShearForce") return emit(plain3(d->dbsx, d->dbsy, d->dbsz, nb), out, capacity, count);
        if (f == "bondBendingTorque") return emit(plain3(d->dbtx, d->dbty, d->dbtz, nb), out, capacity, count);
        if (f == "bondPoint") return emit(plain3(d->dbpx, d->dbpy, d->dbpz, nb), out, capacity, count);
        if (f == "bondNormal") return emit(plain3(d->dbnx, d->dbny, d->dbnz, nb), out, capacity, count);
        const size_t nc = d->nc;
        if (f == "clumpPosition") return emit(plain3(d->cpx, d->cpy, d->cpz, nc), out, capacity, count);
        if (f == "clumpVelocity") return emit(plain3(d->cvx, d->cvy, d->cvz, nc), out, capacity, count);
        if (f == "clumpAngularVelocity") return emit(plain3(d->cwx, d->cwy, d->cwz, nc), out, capacity, count);
        if (f == "clumpOrientation") return emit(plain4(d->cq0, d->cq1, d->cq2, d->cq3, nc), out, capacity, count);
        if (f == "wallPosition") return emit(plain3(d->wpx, d->wpy, d->wpz, d->nw), out, capacity, count);
        if (f == "wallOrientation") return emit(plain4(d->wq0, d->wq1, d->wq2, d->wq3, d->nw), out, capacity, count);
        if (f == "vertexGlobalPosition") return emit(plain3(d->vgx, d->vgy, d->vgz, d->nv), out, capacity, count);
        if (f == "triangleCircumcenter") return emit(fetch(d->tcc.p, 3 * (size_t)d->nt, st), out, capacity, count);
        throw InvalidArg("unknown field '" + f + "'");
    });
}

// outputBallVTU (DEMSolver.h:908-993) and outputMeshWallVTU (:995-1098):
// balls in insertion order with radius, material, clump id, velocity and
// angular velocity; the walls' vertices at their current global positions
// and their triangles with wall and material ids.
extern "C" int demsph_dem_record_contacts(demsph_dem* d, int enable)
{
    DEM_GUARD({
        if (d->stream) DEMSPH_HIP(hipStreamSynchronize(d->stream));
        d->drop_graphs();  // captured evaluations hold the record pointer
        d->record_contacts = enable != 0;
        if (d->record_contacts && d->initialized && d->nt > 0) d->crec.alloc(12 * (size_t)d->KW * d->n);
        if (d->record_contacts && d->initialized)
            d->prec.alloc(12 * (size_t)std::max(1, std::max(d->pcap[0], d->pcap[1])));
    });
}

// Dynamic DEM state to host memory and back: the balls in their current
// slot order with ranks and forces, the current pair list with its contact
// history and reverse index, clump and wall poses, wall-contact history and
// bond state, plus the step counter and time.  Static data (materials,
// triangles, bond pairs, clump inertia, hooks) comes from the importing
// scene's own staging; the triangle cells are rebuilt from the imported
// vertices at the next evaluation.
namespace {
constexpr uint32_t kDemStateMagic = 0x4d454444u;  // "DDEM"
struct DemStateHeader {
    uint32_t magic, version;
    int32_t n, nc, nw, nv, nt, nbond, KW, fp64;
    int32_t npairs, has_list, walls_static, walls_valid;
    uint64_t iStep;
    double time;
};

// Every array of the state in one fixed order, as (device pointer, bytes).
template <class F>
void dem_state_arrays(demsph_dem* d, F&& io)
{
    const size_t n = d->n, nc = d->nc, nw = d->nw, nv = d->nv, nt = d->nt, nb = d->nbond_dev;
    const size_t np = d->has_list ? (size_t)d->npairs : 0, hn = (size_t)d->KW * n;
    const int cb = d->cb, cp = d->cp;
    for (auto* b : {&d->bpx[cb], &d->bpy[cb], &d->bpz[cb], &d->bvx[cb], &d->bvy[cb], &d->bvz[cb], &d->bwx[cb],
                    &d->bwy[cb], &d->bwz[cb], &d->br[cb], &d->bim[cb], &d->fx, &d->fy, &d->fz, &d->tx, &d->ty, &d->tz})
        io(b->p, n * 8);
    for (auto* b : {&d->bmat[cb], &d->bclump[cb], &d->borig[cb], &d->rank}) io(b->p, n * 4);
    if (d->has_list) {
        io(d->pstart[cp].p, (n + 1) * 4);
        io(d->ppartner[cp].p, np * 4);
        io(d->powner[cp].p, np * 4);
        if (d->periodic) io(d->pshift[cp].p, np);
        for (auto* b : {&d->psx[cp], &d->psy[cp], &d->psz[cp], &d->prx[cp], &d->pry[cp], &d->prz[cp], &d->pox[cp],
                        &d->poy[cp], &d->poz[cp], &d->pov[cp], &d->pFx[cp], &d->pFy[cp], &d->pFz[cp], &d->pRx[cp],
                        &d->pRy[cp], &d->pRz[cp], &d->pIx[cp], &d->pIy[cp], &d->pIz[cp]})
            io(b->p, np * 8);
        io(d->rstart.p, (n + 1) * 4);
        io(d->rpair.p, np * 4);
    }
    if (nc > 0)
        for (auto* b : {&d->cpx, &d->cpy, &d->cpz, &d->cvx, &d->cvy, &d->cvz, &d->cwx, &d->cwy, &d->cwz, &d->cq0,
                        &d->cq1, &d->cq2, &d->cq3, &d->cFx, &d->cFy, &d->cFz, &d->cTx, &d->cTy, &d->cTz})
            io(b->p, nc * 8);
    if (nt > 0) {
        for (auto* b : {&d->wpx, &d->wpy, &d->wpz, &d->wvx, &d->wvy, &d->wvz, &d->wwx, &d->wwy, &d->wwz, &d->wq0,
                        &d->wq1, &d->wq2, &d->wq3})
            io(b->p, nw * 8);
        for (auto* b : {&d->vgx, &d->vgy, &d->vgz}) io(b->p, nv * 8);
        io(d->tcc.p, 3 * nt * 8);
        io(d->hcnt.p, n * 4);
        io(d->htri.p, hn * 4);
        io(d->hcancel.p, hn * 4);
        for (auto* b : {&d->hsx, &d->hsy, &d->hsz, &d->hrx, &d->hry, &d->hrz, &d->hox, &d->hoy, &d->hoz, &d->hov})
            io(b->p, hn * 8);
        for (auto* b : {&d->wfx, &d->wfy, &d->wfz, &d->wtx, &d->wty, &d->wtz}) io(b->p, n * 8);
    }
    if (nb > 0) {
        for (auto* b : {&d->dbpx, &d->dbpy, &d->dbpz, &d->dbnx, &d->dbny, &d->dbnz, &d->dbsx, &d->dbsy, &d->dbsz,
                        &d->dbtx, &d->dbty, &d->dbtz, &d->dbFn, &d->dbTt, &d->dbsig, &d->dbtau, &d->dbFx, &d->dbFy,
                        &d->dbFz, &d->dbTix, &d->dbTiy, &d->dbTiz, &d->dbTjx, &d->dbTjy, &d->dbTjz})
            io(b->p, nb * 8);
        io(d->dbonded.p, nb * 4);
    }
}

// The block checkpoint: every array of the state export copied to one
// device buffer (segments 256-byte aligned) by k_multi_copy launches, with
// the host-side values that select and size them.
void dem_block_copy(demsph_dem* d, bool save)
{
    std::vector<std::pair<void*, size_t>> segs;
    size_t total = 0;
    const auto seg = [&](void* p, size_t b) {
        if (!b) return;
        segs.push_back({p, b});
        total += (b + 255) & ~size_t(255);
    };
    dem_state_arrays(d, seg);
    // a replayed block stamps its steps again from the same count
    if (d->stamps.counter) seg(d->stamps.counter, sizeof(int));
    if (save && d->ckbuf.n < total) d->ckbuf.alloc(total + total / 4);
    unsigned char* ck = d->ckbuf.p;
    size_t off = 0;
    for (size_t s0 = 0; s0 < segs.size(); s0 += kMaxCopySegs) {
        CopyTable t{};
        size_t most = 0;
        t.n = (int)std::min<size_t>(kMaxCopySegs, segs.size() - s0);
        for (int k = 0; k < t.n; k++) {
            const auto& sg = segs[s0 + k];
            unsigned char* c = ck + off;
            t.s[k].src = save ? static_cast<const void*>(sg.first) : static_cast<const void*>(c);
            t.s[k].dst = save ? static_cast<void*>(c) : sg.first;
            t.s[k].bytes = sg.second;
            most = std::max(most, sg.second);
            off += (sg.second + 255) & ~size_t(255);
        }
        const int gx = (int)std::min<size_t>(64, std::max<size_t>(1, (most / 4 + 255) / 256));
        hipLaunchKernelGGL(k_multi_copy, dim3(gx, t.n), dim3(256), 0, d->stream, t);
        DEMSPH_HIP(hipGetLastError());
    }
}

}  // namespace

namespace demsph {
void dem_block_save(demsph_dem* d)
{
    if (d->npairs < 0) throw std::logic_error("block checkpoint with an unknown pair count");
    d->hck = {d->iStep, d->time, d->cb, d->cp, d->npairs, d->has_list, d->walls_valid, d->tri_cells_valid,
              d->walls_static};
    dem_block_copy(d, true);
}

void dem_block_restore(demsph_dem* d, int grow_pairs)
{
    const auto& h = d->hck;
    d->iStep = h.iStep; d->time = h.time; d->cb = h.cb; d->cp = h.cp; d->npairs = h.npairs;
    d->has_list = h.has_list; d->walls_valid = h.walls_valid; d->tri_cells_valid = h.tri_cells_valid;
    d->walls_static = h.walls_static;
    // all work using the old lists is done (the caller synchronised)
    for (int w = 0; w < 2; w++)
        if (d->pcap[w] < grow_pairs) d->alloc_pairs(w, grow_pairs);
    if ((int)d->rpair.n < std::max(d->pcap[0], d->pcap[1])) d->rpair.alloc(std::max(d->pcap[0], d->pcap[1]));
    dem_block_copy(d, false);
}
}  // namespace demsph

namespace {
DemStateHeader dem_state_header(const demsph_dem* d)
{
    DemStateHeader h{};
    h.magic = kDemStateMagic; h.version = 2u;
    h.n = d->n; h.nc = d->nc; h.nw = d->nw; h.nv = d->nv; h.nt = d->nt; h.nbond = d->nbond_dev;
    h.KW = d->KW; h.fp64 = d->fp64 ? 1 : 0;
    h.npairs = d->has_list ? d->npairs : 0;
    h.has_list = d->has_list ? 1 : 0;
    h.walls_static = d->walls_static ? 1 : 0;
    h.walls_valid = d->walls_valid ? 1 : 0;
    h.iStep = d->iStep; h.time = d->time;
    return h;
}
}  // namespace

extern "C" int demsph_dem_export_state(demsph_dem* d, void* out, size_t capacity, size_t* bytes)
{
    DEM_GUARD({
        d->require_init();
        const DemStateHeader h = dem_state_header(d);
        size_t total = sizeof h;
        dem_state_arrays(d, [&](const void*, size_t b) { total += b; });
        if (bytes) *bytes = total;
        if (!out) return;
        if (capacity < total) throw InvalidArg("state buffer too small");
        char* o = static_cast<char*>(out);
        std::memcpy(o, &h, sizeof h);
        o += sizeof h;
        dem_state_arrays(d, [&](const void* dev, size_t b) {
            if (b) DEMSPH_HIP(hipMemcpyAsync(o, dev, b, hipMemcpyDeviceToHost, d->stream));
            o += b;
        });
        DEMSPH_HIP(hipStreamSynchronize(d->stream));
    });
}

extern "C" int demsph_dem_import_state(demsph_dem* d, const void* in, size_t bytes)
{
    DEM_GUARD({
        if (!in) throw InvalidArg("null state");
        d->require_init();
        DemStateHeader h;
        if (bytes < sizeof h) throw InvalidArg("state too short");
        std::memcpy(&h, in, sizeof h);
        if (h.magic != kDemStateMagic || h.version != 2u) throw InvalidArg("not a DEM scene state (or of another layout version)");
        if (h.n != d->n || h.nc != d->nc || h.nw != d->nw || h.nv != d->nv || h.nt != d->nt ||
            h.nbond != d->nbond_dev || h.fp64 != (d->fp64 ? 1 : 0))
            throw InvalidArg("the state belongs to a scene of other counts or precision");
        if (h.KW != kWallCap && h.KW != kWallCapMax) throw InvalidArg("corrupt DEM state header (wall capacity)");
        if (h.npairs < 0 || (h.has_list == 0 && h.npairs != 0)) throw InvalidArg("corrupt DEM state header");
        // the size the header implies, computed before anything is resized
        const size_t np = (size_t)h.npairs;
        size_t total = sizeof h;
        {
            const bool had = d->has_list;
            const int had_np = d->npairs, had_kw = d->KW;
            d->has_list = h.has_list != 0;
            d->npairs = h.npairs;
            d->KW = h.KW;
            dem_state_arrays(d, [&](const void*, size_t b) { total += b; });
            d->has_list = had;
            d->npairs = had_np;
            d->KW = had_kw;
        }
        if (bytes != total) throw InvalidArg("state size does not match the scene");
        DEMSPH_HIP(hipStreamSynchronize(d->stream));
        d->drop_graphs();
        if (h.KW != d->KW) d->set_wall_capacity(h.KW);  // the state's wall history layout
        if (h.has_list) {
            if ((size_t)d->pcap[d->cp] < np) d->alloc_pairs(d->cp, h.npairs + h.npairs / 4);
            if (d->rpair.n < np) d->rpair.alloc(std::max<size_t>(np, (size_t)d->pcap[d->cp]));
        }
        d->has_list = h.has_list != 0;
        d->npairs = h.npairs;
        d->iStep = h.iStep;
        d->time = h.time;
        d->walls_static = h.walls_static != 0;
        d->walls_valid = h.walls_valid != 0;
        d->tri_cells_valid = false;
        const char* p = static_cast<const char*>(in) + sizeof h;
        dem_state_arrays(d, [&](void* dev, size_t b) {
            if (b) DEMSPH_HIP(hipMemcpyAsync(dev, p, b, hipMemcpyHostToDevice, d->stream));
            p += b;
        });
        DEMSPH_HIP(hipStreamSynchronize(d->stream));
    });
}

extern "C" int demsph_dem_write_vtu(demsph_dem* d, const char* dir, int what, long frame, long step, double time)
{
    DEM_GUARD({
        d->require_init();
        if (!dir) throw InvalidArg("null directory");
        if (frame < 0 || step < 0) throw InvalidArg("negative frame or step");
        hipStream_t st = d->stream;
        if (what & ~(DEMSPH_VTU_BALL | DEMSPH_VTU_TRIANGLE_WALL | DEMSPH_VTU_BALL_TRIANGLE | DEMSPH_VTU_BALL_INTERACTION |
                     DEMSPH_VTU_BONDED_INTERACTION))
            throw InvalidArg("unknown frame kinds");
        if (what & DEMSPH_VTU_BONDED_INTERACTION) {
            // outputBondedInteractionVTU (DEMSolver.h:1186-1285): every bond in
            // the bond list's order with its point, state and loads
            const size_t nb = d->nbond;
            const auto v3 = [&](const DBuf<double>& x, const DBuf<double>& y, const DBuf<double>& z) {
                const auto hx = fetch(x.p, nb, st), hy = fetch(y.p, nb, st), hz = fetch(z.p, nb, st);
                std::vector<double> v(3 * nb);
                for (size_t i = 0; i < nb; i++) { v[3 * i] = hx[i]; v[3 * i + 1] = hy[i]; v[3 * i + 2] = hz[i]; }
                return v;
            };
            vtu::Writer w(dir, "bondedInteraction", frame);
            w.begin(time, step, nb, nb);
            w.points(v3(d->dbpx, d->dbpy, d->dbpz));
            w.vertex_cells(nb);
            w.open_data("PointData", "overlap");  // the reference's name for the active scalars
            w.ints("is bonded", fetch(d->dbonded.p, nb, st));
            w.reals("normal force", fetch(d->dbFn.p, nb, st));
            w.reals("torsion torque", fetch(d->dbTt.p, nb, st));
            w.reals("maximum normal stress", fetch(d->dbsig.p, nb, st));
            w.reals("maximum shear stress", fetch(d->dbtau.p, nb, st));
            w.vectors("contact normal", v3(d->dbnx, d->dbny, d->dbnz));
            w.vectors("shearing force", v3(d->dbsx, d->dbsy, d->dbsz));
            w.vectors("bending torque", v3(d->dbtx, d->dbty, d->dbtz));
            w.close_data("PointData");
            w.end();
        }
        // one interaction frame: point, overlap, normal, force, torque and
        // the three springs per entry (DEMSolver.h:1100-1185, 1286-1370)
        const auto interactions = [&](const char* stem, const std::vector<double>& pt, const std::vector<double>& ov,
                                      const std::vector<double>& nr, const std::vector<double>& fc,
                                      const std::vector<double>& tc, const std::vector<double>& ss,
                                      const std::vector<double>& sr, const std::vector<double>& stq) {
            const size_t m = ov.size();
            vtu::Writer w(dir, stem, frame);
            w.begin(time, step, m, m);
            w.points(pt);
            w.vertex_cells(m);
            w.open_data("PointData", "overlap");
            w.reals("overlap", ov);
            w.vectors("contact normal", nr);
            w.vectors("contact force", fc);
            w.vectors("contact torque", tc);
            w.vectors("sliding spring", ss);
            w.vectors("rolling spring", sr);
            w.vectors("torsion spring", stq);
            w.close_data("PointData");
            w.end();
        };
        if (what & DEMSPH_VTU_BALL_INTERACTION) {
            // outputBallInteractionVTU: every pair of the current list, by
            // pointed ball in insertion order, then in the list's order
            if (!d->record_contacts)
                throw InvalidArg("ball interaction frames need demsph_dem_record_contacts before the evaluation");
            const size_t n = d->n;
            std::vector<double> pt, ov, nr, fc, tc, ss, sr, stq;
            if (d->has_list && d->npairs > 0) {
                const int c = d->cp;
                const size_t np = (size_t)d->npairs;
                const auto orig = fetch(d->borig[d->cb].p, n, st);
                const auto start = fetch(d->pstart[c].p, n + 1, st);
                const auto rec = fetch(d->prec.p, 12 * np, st);
                const auto o_v = fetch(d->pov[c].p, np, st);
                const auto sx = fetch(d->psx[c].p, np, st), sy = fetch(d->psy[c].p, np, st), sz = fetch(d->psz[c].p, np, st);
                const auto rx = fetch(d->prx[c].p, np, st), ry = fetch(d->pry[c].p, np, st), rz = fetch(d->prz[c].p, np, st);
                const auto ox = fetch(d->pox[c].p, np, st), oy = fetch(d->poy[c].p, np, st), oz = fetch(d->poz[c].p, np, st);
                std::vector<int> rankh(n);
                for (size_t i = 0; i < n; i++) rankh[orig[i]] = (int)i;
                for (size_t o = 0; o < n; o++) {
                    const int i = rankh[o];
                    for (int p = start[i]; p < start[i + 1]; p++) {
                        const double* r = &rec[12 * (size_t)p];
                        pt.insert(pt.end(), r, r + 3);
                        nr.insert(nr.end(), r + 3, r + 6);
                        fc.insert(fc.end(), r + 6, r + 9);
                        tc.insert(tc.end(), r + 9, r + 12);
                        ov.push_back(o_v[p]);
                        ss.insert(ss.end(), {sx[p], sy[p], sz[p]});
                        sr.insert(sr.end(), {rx[p], ry[p], rz[p]});
                        stq.insert(stq.end(), {ox[p], oy[p], oz[p]});
                    }
                }
            }
            interactions("ballInteraction", pt, ov, nr, fc, tc, ss, sr, stq);
        }
        if (what & DEMSPH_VTU_BALL) {
            const int cb = d->cb;
            const size_t n = d->n;
            const auto orig = fetch(d->borig[cb].p, n, st);
            const auto vec = [&](const DBuf<double>& x, const DBuf<double>& y, const DBuf<double>& z) {
                const auto hx = fetch(x.p, n, st), hy = fetch(y.p, n, st), hz = fetch(z.p, n, st);
                std::vector<double> v(3 * n);
                for (size_t i = 0; i < n; i++) {
                    const size_t o = 3 * (size_t)orig[i];
                    v[o] = hx[i]; v[o + 1] = hy[i]; v[o + 2] = hz[i];
                }
                return v;
            };
            vtu::Writer w(dir, "ball", frame);
            w.begin(time, step, n, n);
            w.points(vec(d->bpx[cb], d->bpy[cb], d->bpz[cb]));
            w.vertex_cells(n);
            w.open_data("PointData", "radius");
            w.reals("radius", d->hb.r);
            w.ints("materialID", d->hb.mat);
            w.ints("clumpID", d->hb.clump);
            w.vectors("velocity", vec(d->bvx[cb], d->bvy[cb], d->bvz[cb]));
            w.vectors("angularVelocity", vec(d->bwx[cb], d->bwy[cb], d->bwz[cb]));
            w.close_data("PointData");
            w.end();
        }
        if ((what & DEMSPH_VTU_TRIANGLE_WALL) && d->nt > 0) {
            const size_t nv = d->nv, nt = d->nt;
            const auto gx = fetch(d->vgx.p, nv, st), gy = fetch(d->vgy.p, nv, st), gz = fetch(d->vgz.p, nv, st);
            std::vector<double> xyz(3 * nv);
            for (size_t i = 0; i < nv; i++) { xyz[3 * i] = gx[i]; xyz[3 * i + 1] = gy[i]; xyz[3 * i + 2] = gz[i]; }
            std::vector<int> tmat(nt, 0);
            for (size_t t = 0; t < nt; t++) {
                const int wl = d->hw.twall[t];
                if (wl >= 0 && (size_t)wl < d->hw.mat.size()) tmat[t] = d->hw.mat[(size_t)wl];
            }
            vtu::Writer w(dir, "triangleWall", frame);
            w.begin(time, step, nv, nt);
            w.points(xyz);
            w.triangle_cells(d->hw.t0, d->hw.t1, d->hw.t2);
            w.open_data("CellData", "materialID");
            w.ints("wallID", d->hw.twall);
            w.ints("materialID", tmat);
            w.close_data("CellData");
            w.end();
        }
        if ((what & DEMSPH_VTU_BALL_TRIANGLE) && d->nt > 0) {
            // outputBallTriangleInteractionVTU (DEMSolver.h:1286-1370): every
            // candidate contact of the last evaluation, by ball in insertion
            // order and then in the candidate walk's order
            if (!d->record_contacts)
                throw InvalidArg("ball-triangle frames need demsph_dem_record_contacts before the evaluation");
            const size_t n = d->n, hn = (size_t)d->KW * n;
            std::vector<int> cnt;
            std::vector<double> rec, sx, sy, sz, rx, ry, rz, ox, oy, oz, ov;
            if (d->nt > 0) {
                cnt = fetch(d->hcnt.p, n, st);
                rec = fetch(d->crec.p, 12 * hn, st);
                sx = fetch(d->hsx.p, hn, st); sy = fetch(d->hsy.p, hn, st); sz = fetch(d->hsz.p, hn, st);
                rx = fetch(d->hrx.p, hn, st); ry = fetch(d->hry.p, hn, st); rz = fetch(d->hrz.p, hn, st);
                ox = fetch(d->hox.p, hn, st); oy = fetch(d->hoy.p, hn, st); oz = fetch(d->hoz.p, hn, st);
                ov = fetch(d->hov.p, hn, st);
            }
            std::vector<double> pt, nr, fc, tc, ss, sr, stq, overlap;
            for (size_t o = 0; o < cnt.size(); o++) {
                for (int c = 0; c < cnt[o] && c < d->KW; c++) {
                    const size_t h = (size_t)c * n + o;
                    const double* r = &rec[12 * h];
                    pt.insert(pt.end(), r, r + 3);
                    nr.insert(nr.end(), r + 3, r + 6);
                    fc.insert(fc.end(), r + 6, r + 9);
                    tc.insert(tc.end(), r + 9, r + 12);
                    overlap.push_back(ov[h]);
                    ss.insert(ss.end(), {sx[h], sy[h], sz[h]});
                    sr.insert(sr.end(), {rx[h], ry[h], rz[h]});
                    stq.insert(stq.end(), {ox[h], oy[h], oz[h]});
                }
            }
            interactions("ballTriangleInteraction", pt, overlap, nr, fc, tc, ss, sr, stq);
        }
    });
}

// demsph_dem_write_vtu without waiting for the ball frame: the ball arrays
// are copied to pinned host memory in stream order and the frame thread
// writes the file while the scene keeps stepping.  The other kinds are
// written before returning.
extern "C" int demsph_dem_write_vtu_async(demsph_dem* d, const char* dir, int what, long frame, long step,
                                          double time)
{
    DEM_GUARD({
        d->require_init();
        if (!dir) throw InvalidArg("null directory");
        if (frame < 0 || step < 0) throw InvalidArg("negative frame or step");
        const int rest = what & ~DEMSPH_VTU_BALL;
        if (rest) {
            const int rc = demsph_dem_write_vtu(d, dir, rest, frame, step, time);
            if (rc != DEMSPH_OK) throw std::runtime_error(demsph_last_error());
        }
        if (!(what & DEMSPH_VTU_BALL)) return;
        const int cb = d->cb;
        const size_t n = d->n;
        hipStream_t st = d->stream;
        auto snap = std::make_shared<PinnedSnap>();
        const double* a[9] = {snap->take(d->bpx[cb].p, n, st), snap->take(d->bpy[cb].p, n, st),
                              snap->take(d->bpz[cb].p, n, st), snap->take(d->bvx[cb].p, n, st),
                              snap->take(d->bvy[cb].p, n, st), snap->take(d->bvz[cb].p, n, st),
                              snap->take(d->bwx[cb].p, n, st), snap->take(d->bwy[cb].p, n, st),
                              snap->take(d->bwz[cb].p, n, st)};
        const int* orig = snap->take(d->borig[cb].p, n, st);
        snap->mark(st);
        vtu::post([snap, a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5], a6 = a[6], a7 = a[7],
                   a8 = a[8], orig, r = d->hb.r, mat = d->hb.mat, clump = d->hb.clump, dd = std::string(dir), n, frame,
                   step, time] {
            snap->wait();
            const double* c[9] = {a0, a1, a2, a3, a4, a5, a6, a7, a8};
            std::vector<double> v[3];
            for (auto& x : v) x.resize(3 * n);
            for (size_t i = 0; i < n; i++) {
                const size_t o = 3 * (size_t)orig[i];
                for (int k = 0; k < 3; k++) {
                    v[k][o] = c[3 * k][i]; v[k][o + 1] = c[3 * k + 1][i]; v[k][o + 2] = c[3 * k + 2][i];
                }
            }
            vtu::Writer w(dd, "ball", frame);
            w.begin(time, step, n, n);
            w.points(v[0]);
            w.vertex_cells(n);
            w.open_data("PointData", "radius");
            w.reals("radius", r);
            w.ints("materialID", mat);
            w.ints("clumpID", clump);
            w.vectors("velocity", v[1]);
            w.vectors("angularVelocity", v[2]);
            w.close_data("PointData");
            w.end();
        });
    });
}

extern "C" int demsph_dem_set(demsph_dem* d, const char* field, const void* in, size_t count)
{
    DEM_GUARD({
        d->require_init();
        if (!field || !in) throw InvalidArg("null argument");
        const std::string f = field;
        const double* src = static_cast<const double*>(in);
        hipStream_t st = d->stream;
        const int cb = d->cb;
        if (f == "position" || f == "velocity" || f == "angularVelocity") {
            const size_t n = d->n;
            if (count != 3 * n) throw InvalidArg("expected 3*numBalls values");
            auto orig = fetch(d->borig[cb].p, n, st);
            std::vector<double> hx(n), hy(n), hz(n);
            for (size_t i = 0; i < n; i++) {
                hx[i] = src[3 * (size_t)orig[i]]; hy[i] = src[3 * (size_t)orig[i] + 1]; hz[i] = src[3 * (size_t)orig[i] + 2];
            }
            DBuf<double>* t[3];
            if (f == "position") { t[0] = &d->bpx[cb]; t[1] = &d->bpy[cb]; t[2] = &d->bpz[cb]; }
            else if (f == "velocity") { t[0] = &d->bvx[cb]; t[1] = &d->bvy[cb]; t[2] = &d->bvz[cb]; }
            else { t[0] = &d->bwx[cb]; t[1] = &d->bwy[cb]; t[2] = &d->bwz[cb]; }
            copy_sync(t[0]->p, hx.data(), n * 8, hipMemcpyHostToDevice, d->stream);
            copy_sync(t[1]->p, hy.data(), n * 8, hipMemcpyHostToDevice, d->stream);
            copy_sync(t[2]->p, hz.data(), n * 8, hipMemcpyHostToDevice, d->stream);
            return;
        }
        if (f == "wallVelocity" || f == "wallAngularVelocity") {
            const size_t m = d->nw;
            if (count != 3 * m) throw InvalidArg("expected 3*numWalls values");
            if (m == 0) return;
            std::vector<double> hx(m), hy(m), hz(m);
            for (size_t i = 0; i < m; i++) { hx[i] = src[3 * i]; hy[i] = src[3 * i + 1]; hz[i] = src[3 * i + 2]; }
            for (size_t i = 0; i < 3 * m; i++)
                if (src[i] != 0.0) {
                    d->walls_static = false;
                    d->tri_cells_valid = false;
                }
            DEMSPH_HIP(hipStreamSynchronize(st));
            if (f == "wallVelocity") {
                copy_sync(d->wvx.p, hx.data(), m * 8, hipMemcpyHostToDevice, d->stream);
                copy_sync(d->wvy.p, hy.data(), m * 8, hipMemcpyHostToDevice, d->stream);
                copy_sync(d->wvz.p, hz.data(), m * 8, hipMemcpyHostToDevice, d->stream);
            } else {
                copy_sync(d->wwx.p, hx.data(), m * 8, hipMemcpyHostToDevice, d->stream);
                copy_sync(d->wwy.p, hy.data(), m * 8, hipMemcpyHostToDevice, d->stream);
                copy_sync(d->wwz.p, hz.data(), m * 8, hipMemcpyHostToDevice, d->stream);
            }
            return;
        }
        throw InvalidArg("unknown field '" + f + "'");
    });
}

extern "C" int demsph_dem_set_periodic_boundary(demsph_dem* d, int x, int y, int z)
{
    DEM_GUARD({
        if (d->initialized) throw InvalidArg("setPeriodicBoundary must precede initialize");
        d->per.on[0] = x != 0; d->per.on[1] = y != 0; d->per.on[2] = z != 0;
        d->periodic = d->per.on[0] || d->per.on[1] || d->per.on[2];
    });
}

extern "C" int demsph_dem_tune(demsph_dem* d, const char* knob, int value)
{
    DEM_GUARD({
        if (!knob) throw InvalidArg("null knob");
        const std::string k(knob);
        if (k == "pairCapacity") {
            // entries of each pair list; lists grow on demand from there
            if (!d->initialized || d->has_list) throw InvalidArg("pairCapacity is set between initialize and solve_begin");
            if (value < 1) throw InvalidArg("pairCapacity must be positive");
            for (int w = 0; w < 2; w++) d->alloc_pairs(w, value);
            d->rpair.alloc(value);
            return;
        }
        if (k == "wallSlots") {
            if (value != 0 && value != 1) throw InvalidArg("wallSlots must be 0 or 1");
            d->wall_slots_set = value;
            d->drop_graphs();
            return;
        }
        if (k != "pairLanes") throw InvalidArg("unknown knob '" + k + "'");
        if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8)
            throw InvalidArg("pairLanes must be 0 (by ball count), 1, 2, 4 or 8");
        d->pair_lanes_set = value;
    });
}

extern "C" int demsph_dem_tune_get(demsph_dem* d, const char* knob, int* value)
{
    DEM_GUARD({
        if (!knob || !value) throw InvalidArg("null argument");
        const std::string k(knob);
        if (k == "pairCapacity") {
            *value = d->pcap[d->cp];
            return;
        }
        if (k == "pairReplays") {
            *value = (int)d->pair_replays;
            return;
        }
        if (k == "wallSlots") {
            *value = d->wall_slots_set;
            return;
        }
        if (k == "wallCapacity") {  // wall-triangle candidates per ball
            *value = d->KW;
            return;
        }
        if (k == "wallReplays") {
            *value = (int)d->wall_replays;
            return;
        }
        if (k != "pairLanes") throw InvalidArg("unknown knob '" + k + "'");
        *value = pair_walk_lanes(d->initialized ? d->n : (int)d->hb.size(), d->pair_lanes_set);
    });
}

// Times every "pairLanes" setting on the same steps from the same state
// (the scene state export serves as the checkpoint), keeps the fastest and
// restores the state, so the scene continues as if it had not been tuned.
// pairLanes changes only the pair rebuild, so each trial is scored by the
// HIP-event time of its rebuild kernels (count and write), taken as the
// minimum over two repeats; a setting must beat the built-in rule by 3 % to
// replace it.  Step timestamps and timing totals recorded before the call are
// kept, and any failure restores the state and the knob before it returns.
// A hook set on the scene runs during the trial steps too.
namespace {
using DemTuneKey = std::array<int, 5>;
std::mutex& dem_tune_mutex()
{
    static std::mutex m;
    return m;
}
std::map<DemTuneKey, int>& dem_tune_cache()
{
    static std::map<DemTuneKey, int> c;
    return c;
}
}  // namespace

namespace demsph {
void dem_autotune_cached(demsph_dem* d, const double* g, double dt, int trialSteps)
{
    const DemTuneKey key{(int)std::lround(4.0 * std::log2((double)d->n + 1.0)), d->nt > 0 ? 1 : 0,
                         d->nbond > 0 ? 1 : 0, d->periodic ? 1 : 0, d->fp64 ? 1 : 0};
    {
        std::lock_guard<std::mutex> lk(dem_tune_mutex());
        const auto it = dem_tune_cache().find(key);
        if (it != dem_tune_cache().end()) {
            d->drop_graphs();
            d->pair_lanes_set = it->second;
            return;
        }
    }
    int lanes = 0;
    if (demsph_dem_autotune(d, g, dt, trialSteps, &lanes) != DEMSPH_OK)
        throw std::runtime_error(std::string("DEM autotune: ") + demsph_last_error());
    std::lock_guard<std::mutex> lk(dem_tune_mutex());
    dem_tune_cache()[key] = lanes;
}
void dem_autotune_clear_cache()
{
    std::lock_guard<std::mutex> lk(dem_tune_mutex());
    dem_tune_cache().clear();
}
}  // namespace demsph

extern "C" int demsph_dem_autotune(demsph_dem* d, const double* g, double dt, int trialSteps, int* lanes)
{
    DEM_GUARD({
        d->require_init();
        if (!d->has_list) throw InvalidArg("solve_begin() must precede autotune");
        if (!g || !(dt > 0.0)) throw InvalidArg("gravity and a positive time step are required");
        if (trialSteps < 5) throw InvalidArg("trials need at least 5 steps (one pair rebuild)");
        const auto ok = [](int rc) {
            if (rc != DEMSPH_OK) throw InvalidArg(std::string("autotune: ") + demsph_last_error());
        };
        size_t bytes = 0;
        ok(demsph_dem_export_state(d, nullptr, 0, &bytes));
        std::vector<char> ck(bytes);
        ok(demsph_dem_export_state(d, ck.data(), ck.size(), &bytes));
        DEMSPH_HIP(hipStreamSynchronize(d->stream));
        d->timer.harvest();
        std::map<std::string, std::pair<long, double>> saved_totals;
        d->timer.swap_totals(saved_totals);
        const bool saved_enabled = d->timer.enabled;
        const StepStamps saved_stamps = d->stamps;
        const int saved_lanes = d->pair_lanes_set;
        d->stamps = StepStamps{};  // trials leave no step timestamps
        const auto restore = [&](int knob) {
            ok(demsph_dem_import_state(d, ck.data(), ck.size()));
            DEMSPH_HIP(hipStreamSynchronize(d->stream));
            d->timer.harvest();
            d->timer.swap_totals(saved_totals);
            d->timer.enabled = saved_enabled;
            d->stamps = saved_stamps;
            d->pair_lanes_set = knob;
        };
        try {
            const int rule = pair_walk_lanes(d->n, 0);
            double rule_us = 0.0, best_us = 0.0;
            int best = 0;
            for (int L : {1, 2, 4, 8}) {
                double us = 0.0;
                for (int rep = 0; rep < 2; rep++) {
                    ok(demsph_dem_import_state(d, ck.data(), ck.size()));
                    d->pair_lanes_set = L;
                    d->timer.reset();
                    d->timer.enabled = true;
                    d->step(g, dt, trialSteps);
                    DEMSPH_HIP(hipStreamSynchronize(d->stream));
                    d->timer.harvest();
                    d->timer.enabled = false;
                    double ms = 0.0;
                    long launches = 0;
                    for (const char* k : {"dem_count_pairs", "dem_write_pairs"}) {
                        const auto it = d->timer.totals().find(k);
                        if (it != d->timer.totals().end()) {
                            ms += it->second.second;
                            launches = std::max(launches, it->second.first);
                        }
                    }
                    if (launches == 0) throw std::runtime_error("autotune: no pair rebuild was timed");
                    const double per = 1e3 * ms / (double)launches;
                    us = rep == 0 ? per : std::min(us, per);
                }
                if (L == rule) rule_us = us;
                if (best == 0 || us < best_us) {
                    best = L;
                    best_us = us;
                }
            }
            if (best != rule && best_us > 0.97 * rule_us) best = rule;
            d->timer.reset();
            restore(best);
            if (lanes) *lanes = best;
        } catch (...) {
            d->timer.reset();
            restore(saved_lanes);
            throw;
        }
    });
}

extern "C" int demsph_dem_set_graphs(demsph_dem* d, int enable)
{
    DEM_GUARD({
        d->use_graphs = enable != 0;
        if (!d->use_graphs) d->drop_graphs();
    });
}

extern "C" int demsph_dem_graph_captures(demsph_dem* d, long* count)
{
    DEM_GUARD({
        if (!count) throw InvalidArg("null argument");
        *count = d->graph_captures;
    });
}

extern "C" int demsph_dem_enable_timing(demsph_dem* d, int enable)
{
    DEM_GUARD({
        DEMSPH_HIP(hipStreamSynchronize(d->stream));
        d->timer.harvest();
        d->timer.enabled = enable != 0;
        d->timer.reset();
    });
}

extern "C" int demsph_dem_timing(demsph_dem* d, int index, char* name, size_t nameCap, long* launches, double* totalMs)
{
    DEM_GUARD({
        DEMSPH_HIP(hipStreamSynchronize(d->stream));
        d->timer.harvest();
        const auto& t = d->timer.totals();
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
