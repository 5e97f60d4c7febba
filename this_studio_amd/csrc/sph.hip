// WCSPH scene for MI355X: cell-sorted SoA state, per-step hash/sort/reorder,
// ELL neighbour lists and the three Riemann pair passes of one step.
//
// Reference: DEMSPH/kernel/WCSPHSolver.h (step loop :607-621, neighbour
// search :150-285), SPHNeighborSearchKernel.cu (lists), and
// WCSPHIntegrationKernel.cu (acceleration :42-148, density change :172-264,
// EOS :266-282, dummy normals :4-40, KDK sequence :308-472).
//
// Layout in HBM (slot order = lexicographic (cell hash, insertion index),
// rebuilt every step; "orig" maps slot -> insertion index, "rank" the
// inverse):
//   position   3 x double[n]   (FP64: integration)
//   velocity   3 x double[n]   (FP64: integration)
//   drho, rho, p   float[n]    (density perturbation and its EOS images)
//   phase      uint8[n]        (index into a table of h, m, c, nu, rho0)
// Neighbour lists are ELL, column-major: entry k of slot i at k*n + i, so
// the lanes of a wavefront read consecutive words.  Each slot's list holds
// neighbour slots in the reference's exact walk order (z, y, x cells, then
// ascending insertion index), so the pair set and its per-particle order are
// identical to the reference's.
//
// Pair arithmetic is FP32 (positions differenced in FP64 first); state
// integration is FP64.
#include "common.hpp"
#include "timing.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

namespace demsph {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxPhases = 256;

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

struct SphArgs {
    int nf, nd;
    GridGeom grid;
    const Phase* phases;
    // fluid, current buffers
    const double *px, *py, *pz, *vx, *vy, *vz;
    const float *drho, *rho, *p;
    const uint8_t* phase;
    // fluid, destination buffers
    double *opx, *opy, *opz, *ovx, *ovy, *ovz;
    float *odrho, *orho, *op;
    // lists
    const int *nbr_ff, *cnt_ff, *nbr_fd, *cnt_fd;
    int K_ff, K_fd;
    // dummies (sorted once)
    const double *dpx, *dpy, *dpz, *dvx, *dvy, *dvz;
    const float *dnx, *dny, *dnz;
    const uint8_t* dphase;
    // physics
    float gx, gy, gz;
    double dt;
    int eos;
    float gamma, inv_gamma;
};

// Wendland C2 gradient factor f with gradW = f * r_ij
// (WCSPHIntegrationKernel.cuh:15-32).  The reference's
// sigma/h * [-2 t^3 (1+2q) + 2 t^4] / r equals -5 sigma t^3 / h^2 with
// t = 1 - q/2, q = r/h, sigma = 21 / (16 pi h^3).
__device__ __forceinline__ float grad_factor(float r, float h)
{
    if (r < 1.e-20f || r >= 2.0f * h) return 0.0f;
    const float invh = 1.0f / h;
    const float t = 1.0f - 0.5f * r * invh;
    const float sigma = (float)(21.0 / (16.0 * M_PI)) * invh * invh * invh;
    return -5.0f * sigma * t * t * t * invh * invh;
}

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
__device__ __forceinline__ void eos_eval(const SphArgs& a, const Phase& ph, float drho, float& rho, float& p)
{
    float ch = drho;
    if (ch < -0.1f * ph.rho0) ch = -0.1f * ph.rho0;
    rho = ch + ph.rho0;
    if (a.eos == 1) {
        const double ratio = (double)rho / (double)ph.rho0;
        p = (float)((double)ph.taitB * (pow(ratio, (double)a.gamma) - 1.0));
    } else {
        p = ph.c2 * ch;
    }
}

// ----------------------------------------------------------- hash & reorder
__global__ __launch_bounds__(kThreads) void k_hash_by_orig(int n, GridGeom g, const double* __restrict__ px,
                                                           const double* __restrict__ py,
                                                           const double* __restrict__ pz,
                                                           const int* __restrict__ orig, uint32_t* __restrict__ key)
{
    const int s = blockIdx.x * kThreads + threadIdx.x;
    if (s >= n) return;
    key[orig[s]] = (uint32_t)cell_hash(g, px[s], py[s], pz[s]);
}

struct ReorderArgs {
    int n;
    long long numGrids;
    const uint32_t* skey;  // sorted keys
    const int* perm;       // sorted insertion indices
    int* rank;             // insertion index -> slot (updated)
    int* orig_out;
    const double *px, *py, *pz, *vx, *vy, *vz;
    const float* drho;
    const uint8_t* phase;
    double *opx, *opy, *opz, *ovx, *ovy, *ovz;
    float *odrho, *orho, *op;
    uint8_t* ophase;
    int *cellStart, *cellEnd;
    const Phase* phases;
    int eos;
    float gamma;
};

__global__ __launch_bounds__(kThreads) void k_reorder(ReorderArgs a)
{
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= a.n) return;
    const int o = a.perm[k];
    const int s = a.rank[o];
    a.rank[o] = k;
    a.orig_out[k] = o;
    a.opx[k] = a.px[s]; a.opy[k] = a.py[s]; a.opz[k] = a.pz[s];
    a.ovx[k] = a.vx[s]; a.ovy[k] = a.vy[s]; a.ovz[k] = a.vz[s];
    const uint8_t ph = a.phase[s];
    a.ophase[k] = ph;
    if (a.odrho) {
        const float d = a.drho[s];
        a.odrho[k] = d;
        SphArgs e;
        e.eos = a.eos;
        e.gamma = a.gamma;
        float rho, p;
        eos_eval(e, a.phases[ph], d, rho, p);
        a.orho[k] = rho;
        a.op[k] = p;
    }
    // findStartAndEnd (buildHashStartEnd.cu:14-34)
    const uint32_t h = a.skey[k];
    if ((long long)h < a.numGrids) {
        if (k == 0 || a.skey[k - 1] != h) a.cellStart[h] = k;
        if (k == a.n - 1 || a.skey[k + 1] != h) a.cellEnd[h] = k + 1;
    }
}

// ----------------------------------------------------------- neighbour lists
// Walks the 27-cell stencil of one point over one sorted set and appends
// every member within 2*max(hA,hB) to an ELL list (SPHNeighborSearchKernel.cu
// :57-115 with excludeSelf, :169-228 without).
__device__ __forceinline__ int walk_list(const GridGeom& g, double x, double y, double z, double hA, int self,
                                         const int* __restrict__ cs, const int* __restrict__ ce,
                                         const double* __restrict__ bx, const double* __restrict__ by,
                                         const double* __restrict__ bz, const uint8_t* __restrict__ bph,
                                         const Phase* __restrict__ phases, int* __restrict__ out, int stride, int K)
{
    const Stencil st = make_stencil(g, x, y, z);
    int count = 0;
    for (int zz = st.loz; zz <= st.hiz; zz++) {
        for (int yy = st.loy; yy <= st.hiy; yy++) {
            const int row = (st.cz + zz) * g.ny * g.nx + (st.cy + yy) * g.nx + st.cx;
            for (int xx = st.lox; xx <= st.hix; xx++) {
                const int hb = row + xx;
                const int b0 = cs[hb];
                if (b0 == -1) continue;
                const int b1 = ce[hb];
                for (int j = b0; j < b1; j++) {
                    if (j == self) continue;
                    const double cut = 2.0 * fmax(hA, phases[bph[j]].h);
                    if (within_cut(cut, x - bx[j], y - by[j], z - bz[j])) {
                        if (count < K) out[(size_t)count * stride] = j;
                        count++;
                    }
                }
            }
        }
    }
    return count;
}

struct ListArgs {
    int nf, nd;
    GridGeom grid;
    const Phase* phases;
    const double *px, *py, *pz;
    const uint8_t* phase;
    const int *cs, *ce;    // fluid cells
    const double *dpx, *dpy, *dpz;
    const uint8_t* dphase;
    const int *dcs, *dce;  // dummy cells
    int *nbr_ff, *cnt_ff, *nbr_fd, *cnt_fd;
    int K_ff, K_fd;
    int* maxcount;  // [0] fluid, [1] dummy
};

__global__ __launch_bounds__(kThreads) void k_build_lists(ListArgs a)
{
    const int i = xcd_block() * kThreads + threadIdx.x;
    if (i >= a.nf) return;
    const double x = a.px[i], y = a.py[i], z = a.pz[i];
    const double h = a.phases[a.phase[i]].h;
    const int c1 = walk_list(a.grid, x, y, z, h, i, a.cs, a.ce, a.px, a.py, a.pz, a.phase, a.phases, a.nbr_ff + i,
                             a.nf, a.K_ff);
    a.cnt_ff[i] = c1;
    int c2 = 0;
    if (a.nd > 0)
        c2 = walk_list(a.grid, x, y, z, h, -1, a.dcs, a.dce, a.dpx, a.dpy, a.dpz, a.dphase, a.phases, a.nbr_fd + i,
                       a.nf, a.K_fd);
    a.cnt_fd[i] = c2;
    if (c1 > a.K_ff) atomicMax(&a.maxcount[0], c1);
    if (c2 > a.K_fd) atomicMax(&a.maxcount[1], c2);
}

// ----------------------------------------------------------- pair passes
// updateWCSPHAccelerationKernel (WCSPHIntegrationKernel.cu:42-148) fused
// with the half kick (:150-159) and, in the first half, the drift (:161-170).
template <bool FIRST>
__global__ __launch_bounds__(kThreads) void k_momentum(SphArgs a)
{
    const int i = xcd_block() * kThreads + threadIdx.x;
    if (i >= a.nf) return;
    const double xi = a.px[i], yi = a.py[i], zi = a.pz[i];
    const float vix = (float)a.vx[i], viy = (float)a.vy[i], viz = (float)a.vz[i];
    const Phase pi = a.phases[a.phase[i]];
    const float rhoi = a.rho[i], PL = a.p[i];
    float ax = 0.f, ay = 0.f, az = 0.f;

    const int nff = min(a.cnt_ff[i], a.K_ff);
    for (int k = 0; k < nff; k++) {
        const int j = a.nbr_ff[(size_t)k * a.nf + i];
        const float rx = (float)(xi - a.px[j]), ry = (float)(yi - a.py[j]), rz = (float)(zi - a.pz[j]);
        const Phase pj = a.phases[a.phase[j]];
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrtf(r2);
        const float f = grad_factor(r, 0.5f * (pi.hf + pj.hf));
        if (f == 0.0f) continue;
        const float vjx = (float)a.vx[j], vjy = (float)a.vy[j], vjz = (float)a.vz[j];
        const float rhoj = a.rho[j], PR = a.p[j];
        // e = -r_ij/|r_ij| (zero below |r|^2 = 1e-20, myVec.h:112-128)
        const float inv_r = (r2 < 1.e-20f) ? 0.0f : 1.0f / r;
        const float UL = -(vix * rx + viy * ry + viz * rz) * inv_r;
        const float UR = -(vjx * rx + vjy * ry + vjz * rz) * inv_r;
        const float cbar = 0.5f * (pi.c + pj.c), rbar = 0.5f * (rhoi + rhoj);
        const float beta = fminf(3.0f * fmaxf(UL - UR, 0.0f), cbar);
        const float Ps = 0.5f * (PL + PR) + 0.5f * beta * rbar * (UL - UR);
        const float mr = pj.m / (rhoi * rhoj);
        const float kp = -2.0f * mr * Ps * f;
        // r_hat . gradW = f * r
        const float kv = 2.0f * (0.5f * (pi.nu + pj.nu)) * (r2 < 1.e-20f ? 0.0f : f * r) * mr;
        ax += kp * rx + kv * (vix - vjx);
        ay += kp * ry + kv * (viy - vjy);
        az += kp * rz + kv * (viz - vjz);
    }
    const int nfd = min(a.cnt_fd[i], a.K_fd);
    for (int k = 0; k < nfd; k++) {
        const int j = a.nbr_fd[(size_t)k * a.nf + i];
        const float rx = (float)(xi - a.dpx[j]), ry = (float)(yi - a.dpy[j]), rz = (float)(zi - a.dpz[j]);
        const Phase pj = a.phases[a.dphase[j]];
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrtf(r2);
        const float f = grad_factor(r, 0.5f * (pi.hf + pj.hf));
        if (f == 0.0f) continue;
        const float nx = a.dnx[j], ny = a.dny[j], nz = a.dnz[j];
        const float vjx = (float)a.dvx[j], vjy = (float)a.dvy[j], vjz = (float)a.dvz[j];
        const float UL = -(nx * vix + ny * viy + nz * viz);
        const float UR = -UL - 2.0f * (nx * vjx + ny * vjy + nz * vjz);
        float grj = -(a.gx * rx + a.gy * ry + a.gz * rz);
        if (grj < 0.f) grj = 0.f;
        const float PR = PL + rhoi * grj;
        const float rhoj = dummy_rho(a, pj, PR);
        const float cbar = 0.5f * (pi.c + pj.c), rbar = 0.5f * (rhoi + rhoj);
        const float beta = fminf(3.0f * fmaxf(UL - UR, 0.0f), cbar);
        const float Ps = 0.5f * (PL + PR) + 0.5f * beta * rbar * (UL - UR);
        const float mr = pj.m / (rhoi * rhoj);
        const float kp = -2.0f * mr * Ps * f;
        const float kv = 2.0f * (0.5f * (pi.nu + pj.nu)) * (r2 < 1.e-20f ? 0.0f : f * r) * mr;
        ax += kp * rx + kv * (vix - vjx);
        ay += kp * ry + kv * (viy - vjy);
        az += kp * rz + kv * (viz - vjz);
    }
    ax += a.gx; ay += a.gy; az += a.gz;
    const double hdt = 0.5 * a.dt;
    const double nvx = a.vx[i] + hdt * (double)ax;
    const double nvy = a.vy[i] + hdt * (double)ay;
    const double nvz = a.vz[i] + hdt * (double)az;
    a.ovx[i] = nvx; a.ovy[i] = nvy; a.ovz[i] = nvz;
    if (FIRST) {
        a.opx[i] = xi + a.dt * nvx;
        a.opy[i] = yi + a.dt * nvy;
        a.opz[i] = zi + a.dt * nvz;
    }
}

// updateWCSPHDensityChangeKernel (WCSPHIntegrationKernel.cu:172-264) fused
// with the EOS update (:266-282).  With gradW = f r_ij and e = -r_ij/r:
// (v_i - v*) . gradW = f [ (v_i - v_j) . r_ij / 2 + (P_L - P_R) r / (2 rho_bar c_bar) ].
__global__ __launch_bounds__(kThreads) void k_continuity(SphArgs a)
{
    const int i = xcd_block() * kThreads + threadIdx.x;
    if (i >= a.nf) return;
    const double xi = a.px[i], yi = a.py[i], zi = a.pz[i];
    const float vix = (float)a.vx[i], viy = (float)a.vy[i], viz = (float)a.vz[i];
    const Phase pi = a.phases[a.phase[i]];
    const float rhoi = a.rho[i], PL = a.p[i];
    float d = 0.f;
    const int nff = min(a.cnt_ff[i], a.K_ff);
    for (int k = 0; k < nff; k++) {
        const int j = a.nbr_ff[(size_t)k * a.nf + i];
        const float rx = (float)(xi - a.px[j]), ry = (float)(yi - a.py[j]), rz = (float)(zi - a.pz[j]);
        const Phase pj = a.phases[a.phase[j]];
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrtf(r2);
        const float f = grad_factor(r, 0.5f * (pi.hf + pj.hf));
        if (f == 0.0f) continue;
        const float vjx = (float)a.vx[j], vjy = (float)a.vy[j], vjz = (float)a.vz[j];
        const float rhoj = a.rho[j], PR = a.p[j];
        const float cbar = 0.5f * (pi.c + pj.c), rbar = 0.5f * (rhoi + rhoj);
        const float er = (r2 < 1.e-20f) ? 0.0f : r;
        const float t = 0.5f * ((vix - vjx) * rx + (viy - vjy) * ry + (viz - vjz) * rz) +
                        0.5f * (PL - PR) * er / (rbar * cbar);
        d += pj.m / rhoj * f * t;
    }
    const int nfd = min(a.cnt_fd[i], a.K_fd);
    for (int k = 0; k < nfd; k++) {
        const int j = a.nbr_fd[(size_t)k * a.nf + i];
        const float rx = (float)(xi - a.dpx[j]), ry = (float)(yi - a.dpy[j]), rz = (float)(zi - a.dpz[j]);
        const Phase pj = a.phases[a.dphase[j]];
        const float r2 = rx * rx + ry * ry + rz * rz;
        const float r = sqrtf(r2);
        const float f = grad_factor(r, 0.5f * (pi.hf + pj.hf));
        if (f == 0.0f) continue;
        const float vjx = (float)a.dvx[j], vjy = (float)a.dvy[j], vjz = (float)a.dvz[j];
        float grj = -(a.gx * rx + a.gy * ry + a.gz * rz);
        if (grj < 0.f) grj = 0.f;
        const float PR = PL + rhoi * grj;
        const float rhoj = dummy_rho(a, pj, PR);
        const float cbar = 0.5f * (pi.c + pj.c), rbar = 0.5f * (rhoi + rhoj);
        const float er = (r2 < 1.e-20f) ? 0.0f : r;
        const float t = 0.5f * ((vix - vjx) * rx + (viy - vjy) * ry + (viz - vjz) * rz) +
                        0.5f * (PL - PR) * er / (rbar * cbar);
        d += pj.m / rhoj * f * t;
    }
    const float nd = a.drho[i] + (float)(a.dt * (double)(2.0f * rhoi * d));
    a.odrho[i] = nd;
    float rho, p;
    eos_eval(a, pi, nd, rho, p);
    a.orho[i] = rho;
    a.op[i] = p;
}

// ----------------------------------------------------------- dummy normals
// calDummyParticleNormalKernel (WCSPHIntegrationKernel.cu:4-40) over the
// dummy-dummy pairs of the 27-cell walk, FP64 (computed once).
struct NormalArgs {
    int nd;
    GridGeom grid;
    const Phase* phases;
    const double *dpx, *dpy, *dpz;
    const uint8_t* dphase;
    const int *dcs, *dce;
    float *dnx, *dny, *dnz;
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
    const double x = a.dpx[i], y = a.dpy[i], z = a.dpz[i];
    const Phase pi = a.phases[a.dphase[i]];
    const Stencil st = make_stencil(a.grid, x, y, z);
    double tx = 0, ty = 0, tz = 0;
    for (int zz = st.loz; zz <= st.hiz; zz++)
        for (int yy = st.loy; yy <= st.hiy; yy++)
            for (int xx = st.lox; xx <= st.hix; xx++) {
                const int hb = (st.cz + zz) * a.grid.ny * a.grid.nx + (st.cy + yy) * a.grid.nx + st.cx + xx;
                const int b0 = a.dcs[hb];
                if (b0 == -1) continue;
                const int b1 = a.dce[hb];
                for (int j = b0; j < b1; j++) {
                    if (j == i) continue;
                    const Phase pj = a.phases[a.dphase[j]];
                    const double cut = 2.0 * fmax(pi.h, pj.h);
                    const double rx = x - a.dpx[j], ry = y - a.dpy[j], rz = z - a.dpz[j];
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
    a.dnx[i] = nx; a.dny[i] = ny; a.dnz[i] = nz;
}

// Dummies: hash by insertion index, then gather into sorted order once.
struct DummyReorderArgs {
    int n;
    long long numGrids;
    const uint32_t* skey;
    const int* perm;
    const double *px, *py, *pz, *vx, *vy, *vz;
    const uint8_t* phase;
    double *opx, *opy, *opz, *ovx, *ovy, *ovz;
    uint8_t* ophase;
    int* orig;
    int *cellStart, *cellEnd;
};

__global__ __launch_bounds__(kThreads) void k_dummy_reorder(DummyReorderArgs a)
{
    const int k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= a.n) return;
    const int o = a.perm[k];
    a.orig[k] = o;
    a.opx[k] = a.px[o]; a.opy[k] = a.py[o]; a.opz[k] = a.pz[o];
    a.ovx[k] = a.vx[o]; a.ovy[k] = a.vy[o]; a.ovz[k] = a.vz[o];
    a.ophase[k] = a.phase[o];
    const uint32_t h = a.skey[k];
    if ((long long)h < a.numGrids) {
        if (k == 0 || a.skey[k - 1] != h) a.cellStart[h] = k;
        if (k == a.n - 1 || a.skey[k + 1] != h) a.cellEnd[h] = k + 1;
    }
}

__global__ void k_iota(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_eos_all(int n, const Phase* phases, const uint8_t* phase, const float* drho, float* rho, float* p,
                          int eos, float gamma)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SphArgs e;
    e.eos = eos;
    e.gamma = gamma;
    eos_eval(e, phases[phase[i]], drho[i], rho[i], p[i]);
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

    // fluid state, double-buffered per field group
    double *px[2] = {}, *py[2] = {}, *pz[2] = {};
    double *vx[2] = {}, *vy[2] = {}, *vz[2] = {};
    float *drho[2] = {}, *rho[2] = {}, *p[2] = {};
    uint8_t* phase[2] = {};
    int* orig[2] = {};
    int cx = 0, cv = 0, cd = 0, cp = 0;  // current buffer of position, velocity, density, phase/orig
    int* rank = nullptr;
    // sort
    uint32_t *key = nullptr, *skey = nullptr, *ktmp = nullptr;
    int *perm = nullptr, *vtmp = nullptr;
    void* sort_temp = nullptr;
    int *cellStart = nullptr, *cellEnd = nullptr;
    // lists
    int *nbr_ff = nullptr, *cnt_ff = nullptr, *nbr_fd = nullptr, *cnt_fd = nullptr;
    int* maxcount = nullptr;      // device [2]
    int* maxcount_h = nullptr;    // pinned [2]
    bool lists_valid = false;
    // dummies
    double *dpx = nullptr, *dpy = nullptr, *dpz = nullptr, *dvx = nullptr, *dvy = nullptr, *dvz = nullptr;
    float *dnx = nullptr, *dny = nullptr, *dnz = nullptr;
    uint8_t* dphase = nullptr;
    int* dorig = nullptr;
    int *dcellStart = nullptr, *dcellEnd = nullptr;
    // checkpoint for replay on list overflow
    double *ck_px = nullptr, *ck_py = nullptr, *ck_pz = nullptr, *ck_vx = nullptr, *ck_vy = nullptr, *ck_vz = nullptr;
    float* ck_drho = nullptr;
    uint8_t* ck_phase = nullptr;
    int *ck_orig = nullptr, *ck_rank = nullptr;
    KernelTimer timer;

    ~demsph_sph() { release(); }

    void release()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (int b = 0; b < 2; b++) {
            dfree(px[b]); dfree(py[b]); dfree(pz[b]); dfree(vx[b]); dfree(vy[b]); dfree(vz[b]);
            dfree(drho[b]); dfree(rho[b]); dfree(p[b]); dfree(phase[b]); dfree(orig[b]);
        }
        dfree(rank); dfree(key); dfree(skey); dfree(ktmp); dfree(perm); dfree(vtmp);
        if (sort_temp) (void)hipFree(sort_temp);
        sort_temp = nullptr;
        dfree(cellStart); dfree(cellEnd); dfree(nbr_ff); dfree(cnt_ff); dfree(nbr_fd); dfree(cnt_fd);
        dfree(maxcount);
        if (maxcount_h) (void)hipHostFree(maxcount_h);
        maxcount_h = nullptr;
        dfree(dpx); dfree(dpy); dfree(dpz); dfree(dvx); dfree(dvy); dfree(dvz); dfree(dnx); dfree(dny); dfree(dnz);
        dfree(dphase); dfree(dorig); dfree(dcellStart); dfree(dcellEnd); dfree(d_phases);
        dfree(ck_px); dfree(ck_py); dfree(ck_pz); dfree(ck_vx); dfree(ck_vy); dfree(ck_vz); dfree(ck_drho);
        dfree(ck_phase); dfree(ck_orig); dfree(ck_rank);
        if (own_stream && stream) (void)hipStreamDestroy(stream);
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
        a.nf = nf; a.nd = nd; a.grid = grid; a.phases = d_phases;
        a.px = px[cx]; a.py = py[cx]; a.pz = pz[cx];
        a.vx = vx[cv]; a.vy = vy[cv]; a.vz = vz[cv];
        a.drho = drho[cd]; a.rho = rho[cd]; a.p = p[cd];
        a.phase = phase[cp];
        a.opx = px[1 - cx]; a.opy = py[1 - cx]; a.opz = pz[1 - cx];
        a.ovx = vx[1 - cv]; a.ovy = vy[1 - cv]; a.ovz = vz[1 - cv];
        a.odrho = drho[1 - cd]; a.orho = rho[1 - cd]; a.op = p[1 - cd];
        a.nbr_ff = nbr_ff; a.cnt_ff = cnt_ff; a.nbr_fd = nbr_fd; a.cnt_fd = cnt_fd;
        a.K_ff = K_ff; a.K_fd = K_fd;
        a.dpx = dpx; a.dpy = dpy; a.dpz = dpz; a.dvx = dvx; a.dvy = dvy; a.dvz = dvz;
        a.dnx = dnx; a.dny = dny; a.dnz = dnz; a.dphase = dphase;
        a.gx = (float)g[0]; a.gy = (float)g[1]; a.gz = (float)g[2];
        a.eos = eos;
        a.gamma = (float)gamma;
        a.inv_gamma = (float)(1.0 / gamma);
        return a;
    }

    void alloc_lists()
    {
        dfree(nbr_ff); dfree(nbr_fd);
        nbr_ff = dmalloc<int>((size_t)K_ff * (nf > 0 ? nf : 1));
        nbr_fd = dmalloc<int>((size_t)K_fd * (nf > 0 ? nf : 1));
        lists_valid = false;
    }

    void initialize(const double* minb, const double* maxb, const double* gravity)
    {
        if (initialized) release_state_keep_host();
        if (!stream) {
            DEMSPH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
            own_stream = true;
        }
        g[0] = gravity[0]; g[1] = gravity[1]; g[2] = gravity[2];
        nf = (int)hf.size();
        nd = (int)hd.size();
        if (nf == 0) throw InvalidArg("no fluid particles");
        // grid cell = 2 * max smoothing length of fluid and dummies (WCSPHSolver.h:127-135)
        double cmax = 0.0;
        for (double dx : hf.spacing) cmax = std::max(cmax, 1.3 * dx * 2.0);
        double cd2 = 0.0;
        for (double dx : hd.spacing) cd2 = std::max(cd2, 1.3 * dx * 2.0);
        if (nd > 0) cmax = std::max(cmax, cd2);
        if (!grid_set(grid, minb, maxb, cmax)) throw InvalidArg("degenerate domain or smoothing length");
        numGrids = grid.numGrids();
        if (numGrids > INT32_MAX / 2) throw InvalidArg("grid too large");

        phases.clear();
        fphase_h.resize(nf);
        dphase_h.resize(nd);
        std::vector<double> fm(nf), dm(nd);
        for (int i = 0; i < nf; i++) {
            const double dx = hf.spacing[i];
            fphase_h[i] = phase_of(1.3 * dx, dx * dx * dx * hf.rho0[i], hf.c[i], hf.nu[i], hf.rho0[i]);
        }
        for (int i = 0; i < nd; i++) {
            const double dx = hd.spacing[i];
            dphase_h[i] = phase_of(1.3 * dx, dx * dx * dx * hd.rho0[i], hd.c[i], hd.nu[i], hd.rho0[i]);
        }
        d_phases = dmalloc<Phase>(phases.size());
        DEMSPH_HIP(hipMemcpyAsync(d_phases, phases.data(), sizeof(Phase) * phases.size(), hipMemcpyHostToDevice, stream));

        const size_t n = nf;
        for (int b = 0; b < 2; b++) {
            px[b] = dmalloc<double>(n); py[b] = dmalloc<double>(n); pz[b] = dmalloc<double>(n);
            vx[b] = dmalloc<double>(n); vy[b] = dmalloc<double>(n); vz[b] = dmalloc<double>(n);
            drho[b] = dmalloc<float>(n); rho[b] = dmalloc<float>(n); p[b] = dmalloc<float>(n);
            phase[b] = dmalloc<uint8_t>(n); orig[b] = dmalloc<int>(n);
        }
        rank = dmalloc<int>(n);
        const int nsort = std::max(nf, nd);
        key = dmalloc<uint32_t>(nsort); skey = dmalloc<uint32_t>(nsort); ktmp = dmalloc<uint32_t>(nsort);
        perm = dmalloc<int>(nsort); vtmp = dmalloc<int>(nsort);
        DEMSPH_HIP(hipMalloc(&sort_temp, radix_sort_temp_bytes(nsort)));
        cellStart = dmalloc<int>(numGrids); cellEnd = dmalloc<int>(numGrids);
        cnt_ff = dmalloc<int>(n); cnt_fd = dmalloc<int>(n);
        maxcount = dmalloc<int>(2);
        DEMSPH_HIP(hipHostMalloc((void**)&maxcount_h, 2 * sizeof(int), hipHostMallocDefault));
        alloc_lists();
        cx = cv = cd = cp = 0;

        // upload fluid in insertion order (slot == index)
        std::vector<double> tx(n), ty(n), tz(n);
        auto up3 = [&](const std::vector<double>& src, double* dx_, double* dy_, double* dz_, size_t cnt) {
            tx.resize(cnt); ty.resize(cnt); tz.resize(cnt);
            for (size_t i = 0; i < cnt; i++) { tx[i] = src[3 * i]; ty[i] = src[3 * i + 1]; tz[i] = src[3 * i + 2]; }
            DEMSPH_HIP(hipMemcpy(dx_, tx.data(), cnt * sizeof(double), hipMemcpyHostToDevice));
            DEMSPH_HIP(hipMemcpy(dy_, ty.data(), cnt * sizeof(double), hipMemcpyHostToDevice));
            DEMSPH_HIP(hipMemcpy(dz_, tz.data(), cnt * sizeof(double), hipMemcpyHostToDevice));
        };
        DEMSPH_HIP(hipStreamSynchronize(stream));
        up3(hf.pos, px[0], py[0], pz[0], n);
        up3(hf.vel, vx[0], vy[0], vz[0], n);
        DEMSPH_HIP(hipMemcpy(phase[0], fphase_h.data(), n, hipMemcpyHostToDevice));
        DEMSPH_HIP(hipMemsetAsync(drho[0], 0, n * sizeof(float), stream));
        hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, stream, orig[0], (int)n);
        hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, stream, rank, (int)n);
        hipLaunchKernelGGL(k_eos_all, dim3(blocks_for(n, 256)), dim3(256), 0, stream, (int)n, d_phases, phase[0],
                           drho[0], rho[0], p[0], eos, (float)gamma);
        DEMSPH_HIP(hipGetLastError());

        // dummies: sort once (they are static), cell ranges, normals
        dcellStart = dmalloc<int>(numGrids); dcellEnd = dmalloc<int>(numGrids);
        fill_i32(dcellStart, numGrids, -1, stream);
        fill_i32(dcellEnd, numGrids, -1, stream);
        const size_t m = nd;
        dpx = dmalloc<double>(m); dpy = dmalloc<double>(m); dpz = dmalloc<double>(m);
        dvx = dmalloc<double>(m); dvy = dmalloc<double>(m); dvz = dmalloc<double>(m);
        dnx = dmalloc<float>(m); dny = dmalloc<float>(m); dnz = dmalloc<float>(m);
        dphase = dmalloc<uint8_t>(m); dorig = dmalloc<int>(m);
        if (nd > 0) {
            double *ux = dmalloc<double>(m), *uy = dmalloc<double>(m), *uz = dmalloc<double>(m);
            double *wx = dmalloc<double>(m), *wy = dmalloc<double>(m), *wz = dmalloc<double>(m);
            uint8_t* uph = dmalloc<uint8_t>(m);
            int* uorig = dmalloc<int>(m);
            up3(hd.pos, ux, uy, uz, m);
            up3(hd.vel, wx, wy, wz, m);
            DEMSPH_HIP(hipMemcpy(uph, dphase_h.data(), m, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_iota, dim3(blocks_for(m, 256)), dim3(256), 0, stream, uorig, (int)m);
            hipLaunchKernelGGL(k_hash_by_orig, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, stream, (int)m, grid,
                               ux, uy, uz, uorig, key);
            radix_sort_pairs(key, nullptr, skey, perm, ktmp, vtmp, sort_temp, (int)m, key_bits(numGrids), stream);
            DummyReorderArgs ra{};
            ra.n = (int)m; ra.numGrids = numGrids; ra.skey = skey; ra.perm = perm;
            ra.px = ux; ra.py = uy; ra.pz = uz; ra.vx = wx; ra.vy = wy; ra.vz = wz; ra.phase = uph;
            ra.opx = dpx; ra.opy = dpy; ra.opz = dpz; ra.ovx = dvx; ra.ovy = dvy; ra.ovz = dvz; ra.ophase = dphase;
            ra.orig = dorig; ra.cellStart = dcellStart; ra.cellEnd = dcellEnd;
            hipLaunchKernelGGL(k_dummy_reorder, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, stream, ra);
            NormalArgs na{};
            na.nd = (int)m; na.grid = grid; na.phases = d_phases; na.dpx = dpx; na.dpy = dpy; na.dpz = dpz;
            na.dphase = dphase; na.dcs = dcellStart; na.dce = dcellEnd; na.dnx = dnx; na.dny = dny; na.dnz = dnz;
            hipLaunchKernelGGL(k_dummy_normals, dim3(blocks_for(m, kThreads)), dim3(kThreads), 0, stream, na);
            DEMSPH_HIP(hipGetLastError());
            DEMSPH_HIP(hipStreamSynchronize(stream));
            dfree(ux); dfree(uy); dfree(uz); dfree(wx); dfree(wy); dfree(wz); dfree(uph); dfree(uorig);
        }
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
        DEMSPH_LAUNCH(timer, "sph_hash", k_hash_by_orig, gb, tb, 0, stream, n, grid, px[cx], py[cx], pz[cx],
                      orig[cp], key);
        timer.begin("sph_sort", stream);
        radix_sort_pairs(key, nullptr, skey, perm, ktmp, vtmp, sort_temp, n, key_bits(numGrids), stream);
        timer.end(stream);
        timer.begin("sph_fill_cells", stream);
        fill_i32(cellStart, numGrids, -1, stream);
        fill_i32(cellEnd, numGrids, -1, stream);
        timer.end(stream);
        ReorderArgs ra{};
        ra.n = n; ra.numGrids = numGrids; ra.skey = skey; ra.perm = perm; ra.rank = rank; ra.orig_out = orig[1 - cp];
        ra.px = px[cx]; ra.py = py[cx]; ra.pz = pz[cx]; ra.vx = vx[cv]; ra.vy = vy[cv]; ra.vz = vz[cv];
        ra.drho = drho[cd]; ra.phase = phase[cp];
        ra.opx = px[1 - cx]; ra.opy = py[1 - cx]; ra.opz = pz[1 - cx];
        ra.ovx = vx[1 - cv]; ra.ovy = vy[1 - cv]; ra.ovz = vz[1 - cv];
        ra.odrho = drho[1 - cd]; ra.orho = rho[1 - cd]; ra.op = p[1 - cd]; ra.ophase = phase[1 - cp];
        ra.cellStart = cellStart; ra.cellEnd = cellEnd; ra.phases = d_phases; ra.eos = eos; ra.gamma = (float)gamma;
        DEMSPH_LAUNCH(timer, "sph_reorder", k_reorder, gb, tb, 0, stream, ra);
        cx ^= 1; cv ^= 1; cd ^= 1; cp ^= 1;
        ListArgs la{};
        la.nf = nf; la.nd = nd; la.grid = grid; la.phases = d_phases;
        la.px = px[cx]; la.py = py[cx]; la.pz = pz[cx]; la.phase = phase[cp]; la.cs = cellStart; la.ce = cellEnd;
        la.dpx = dpx; la.dpy = dpy; la.dpz = dpz; la.dphase = dphase; la.dcs = dcellStart; la.dce = dcellEnd;
        la.nbr_ff = nbr_ff; la.cnt_ff = cnt_ff; la.nbr_fd = nbr_fd; la.cnt_fd = cnt_fd;
        la.K_ff = K_ff; la.K_fd = K_fd; la.maxcount = maxcount;
        DEMSPH_LAUNCH(timer, "sph_lists", k_build_lists, gb, tb, 0, stream, la);
        lists_valid = true;
    }

    void first_half_enqueue(double dt)
    {
        SphArgs a = args();
        a.dt = dt;
        const dim3 gb(blocks_for(nf, kThreads)), tb(kThreads);
        DEMSPH_LAUNCH(timer, "sph_momentum_first", k_momentum<true>, gb, tb, 0, stream, a);
        cx ^= 1; cv ^= 1;
    }

    void second_half_enqueue(double dt)
    {
        const dim3 gb(blocks_for(nf, kThreads)), tb(kThreads);
        SphArgs a = args();
        a.dt = dt;
        DEMSPH_LAUNCH(timer, "sph_continuity", k_continuity, gb, tb, 0, stream, a);
        cd ^= 1;
        a = args();
        a.dt = dt;
        DEMSPH_LAUNCH(timer, "sph_momentum_second", k_momentum<false>, gb, tb, 0, stream, a);
        cv ^= 1;
    }

    void reset_overflow() { DEMSPH_HIP(hipMemsetAsync(maxcount, 0, 2 * sizeof(int), stream)); }

    // Returns true when a list overflowed since reset_overflow(); grows the
    // capacities in that case.  Synchronises the stream.
    bool check_overflow()
    {
        DEMSPH_HIP(hipMemcpyAsync(maxcount_h, maxcount, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
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
        if (ck_px) return;
        const size_t n = nf;
        ck_px = dmalloc<double>(n); ck_py = dmalloc<double>(n); ck_pz = dmalloc<double>(n);
        ck_vx = dmalloc<double>(n); ck_vy = dmalloc<double>(n); ck_vz = dmalloc<double>(n);
        ck_drho = dmalloc<float>(n); ck_phase = dmalloc<uint8_t>(n); ck_orig = dmalloc<int>(n); ck_rank = dmalloc<int>(n);
    }

    void checkpoint(bool save)
    {
        const size_t n = nf;
        auto cp8 = [&](double* ck, double* cur) {
            if (save) DEMSPH_HIP(hipMemcpyAsync(ck, cur, n * 8, hipMemcpyDeviceToDevice, stream));
            else DEMSPH_HIP(hipMemcpyAsync(cur, ck, n * 8, hipMemcpyDeviceToDevice, stream));
        };
        cp8(ck_px, px[cx]); cp8(ck_py, py[cx]); cp8(ck_pz, pz[cx]);
        cp8(ck_vx, vx[cv]); cp8(ck_vy, vy[cv]); cp8(ck_vz, vz[cv]);
        auto cpb = [&](void* ck, void* cur, size_t bytes) {
            if (save) DEMSPH_HIP(hipMemcpyAsync(ck, cur, bytes, hipMemcpyDeviceToDevice, stream));
            else DEMSPH_HIP(hipMemcpyAsync(cur, ck, bytes, hipMemcpyDeviceToDevice, stream));
        };
        cpb(ck_drho, drho[cd], n * 4);
        cpb(ck_phase, phase[cp], n);
        cpb(ck_orig, orig[cp], n * 4);
        cpb(ck_rank, rank, n * 4);
        if (!save) {
            hipLaunchKernelGGL(k_eos_all, dim3(blocks_for(n, 256)), dim3(256), 0, stream, (int)n, d_phases,
                               phase[cp], drho[cd], rho[cd], p[cd], eos, (float)gamma);
            DEMSPH_HIP(hipGetLastError());
        }
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

    void step(double dt, int nsteps)
    {
        require_init();
        if (dt <= 0.0) throw InvalidArg("time step must be positive");
        if (nsteps <= 0) return;
        ensure_checkpoint();
        for (int attempt = 0; attempt < 8; attempt++) {
            checkpoint(true);
            const int scx = cx, scv = cv, scd = cd, scp = cp;
            reset_overflow();
            for (int k = 0; k < nsteps; k++) {
                search_enqueue();
                first_half_enqueue(dt);
                second_half_enqueue(dt);
            }
            if (!check_overflow()) return;
            // Replay the block from the checkpoint with larger lists.
            cx = scx; cv = scv; cd = scd; cp = scp;
            checkpoint(false);
        }
        throw std::runtime_error("neighbour list capacity did not converge");
    }
};

namespace {

template <class T>
void scatter_by_orig(const std::vector<T>& slot, const std::vector<int>& orig, std::vector<T>& out, int comps)
{
    out.resize(slot.size());
    const size_t n = orig.size();
    for (size_t s = 0; s < n; s++)
        for (int c = 0; c < comps; c++) out[(size_t)orig[s] * comps + c] = slot[s * comps + c];
}

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
        auto orig = fetch(s->orig[s->cp], n, st);
        auto vec3 = [&](double* const* xs, double* const* ys, double* const* zs, int b) {
            auto hx = fetch(xs[b], n, st), hy = fetch(ys[b], n, st), hz = fetch(zs[b], n, st);
            std::vector<double> v(3 * n);
            for (size_t i = 0; i < n; i++) {
                const size_t o = orig[i];
                v[3 * o] = hx[i]; v[3 * o + 1] = hy[i]; v[3 * o + 2] = hz[i];
            }
            return v;
        };
        auto scal = [&](const float* d) {
            auto h = fetch(d, n, st);
            std::vector<double> v(n);
            for (size_t i = 0; i < n; i++) v[orig[i]] = h[i];
            return v;
        };
        if (f == "position") return emit(vec3(s->px, s->py, s->pz, s->cx), out, capacity, count);
        if (f == "velocity") return emit(vec3(s->vx, s->vy, s->vz, s->cv), out, capacity, count);
        if (f == "densityChange") return emit(scal(s->drho[s->cd]), out, capacity, count);
        if (f == "density") return emit(scal(s->rho[s->cd]), out, capacity, count);
        if (f == "pressure") return emit(scal(s->p[s->cd]), out, capacity, count);
        if (f == "dummyNormal") {
            const size_t m = s->nd;
            auto dor = fetch(s->dorig, m, st);
            auto hx = fetch(s->dnx, m, st), hy = fetch(s->dny, m, st), hz = fetch(s->dnz, m, st);
            std::vector<double> v(3 * m);
            for (size_t i = 0; i < m; i++) {
                v[3 * dor[i]] = hx[i]; v[3 * dor[i] + 1] = hy[i]; v[3 * dor[i] + 2] = hz[i];
            }
            return emit(v, out, capacity, count);
        }
        if (f == "hashValue") {
            auto k = fetch(s->skey, n, st);
            std::vector<int> v(k.begin(), k.end());
            return emit(v, out, capacity, count);
        }
        if (f == "hashIndex") return emit(orig, out, capacity, count);
        if (f == "cellHashStart") return emit(fetch(s->cellStart, (size_t)s->numGrids, st), out, capacity, count);
        if (f == "cellHashEnd") return emit(fetch(s->cellEnd, (size_t)s->numGrids, st), out, capacity, count);
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
                    const int j = nbr[(size_t)k * n + i];
                    v.push_back(ff ? orig[j] : dor[j]);
                }
            }
            return emit(v, out, capacity, count);
        }
        throw InvalidArg("unknown field '" + f + "'");
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
        auto rankv = fetch(s->orig[s->cp], n, st);  // slot -> orig
        const double* src = static_cast<const double*>(in);
        if (f == "position" || f == "velocity") {
            if (count != 3 * n) throw InvalidArg("expected 3*numFluid values");
            std::vector<double> hx(n), hy(n), hz(n);
            for (size_t i = 0; i < n; i++) {
                const size_t o = rankv[i];
                hx[i] = src[3 * o]; hy[i] = src[3 * o + 1]; hz[i] = src[3 * o + 2];
            }
            const int b = (f == "position") ? s->cx : s->cv;
            double* dx = (f == "position") ? s->px[b] : s->vx[b];
            double* dy = (f == "position") ? s->py[b] : s->vy[b];
            double* dz = (f == "position") ? s->pz[b] : s->vz[b];
            DEMSPH_HIP(hipMemcpy(dx, hx.data(), n * 8, hipMemcpyHostToDevice));
            DEMSPH_HIP(hipMemcpy(dy, hy.data(), n * 8, hipMemcpyHostToDevice));
            DEMSPH_HIP(hipMemcpy(dz, hz.data(), n * 8, hipMemcpyHostToDevice));
            s->lists_valid = false;
            return;
        }
        if (f == "densityChange") {
            if (count != n) throw InvalidArg("expected numFluid values");
            std::vector<float> h(n);
            for (size_t i = 0; i < n; i++) h[i] = (float)src[rankv[i]];
            DEMSPH_HIP(hipMemcpy(s->drho[s->cd], h.data(), n * 4, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_eos_all, dim3(blocks_for(n, 256)), dim3(256), 0, st, (int)n, s->d_phases,
                               s->phase[s->cp], s->drho[s->cd], s->rho[s->cd], s->p[s->cd], s->eos, (float)s->gamma);
            DEMSPH_HIP(hipGetLastError());
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
