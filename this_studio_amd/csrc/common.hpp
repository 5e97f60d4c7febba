// Shared device/host helpers for the MI355X (gfx950) DEMSPH kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <stdexcept>
#include <vector>

#include "../../include/demsph.h"

namespace demsph {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);

#define DEMSPH_HIP(call)                                                                  \
    do {                                                                                  \
        hipError_t _e = (call);                                                           \
        if (_e != hipSuccess) {                                                           \
            throw ::demsph::HipFailure(std::string(#call) + ": " + hipGetErrorString(_e)); \
        }                                                                                 \
    } while (0)

struct HipFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct InvalidArg : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// A per-particle list or candidate capacity was exceeded where the work
// cannot be replayed with a larger one.
struct CapacityError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Runs f, translating exceptions to C-ABI status codes.
template <class F>
int guarded(F&& f)
{
    try {
        f();
        return DEMSPH_OK;
    } catch (const InvalidArg& e) {
        set_error(e.what());
        return DEMSPH_ERR_INVALID;
    } catch (const CapacityError& e) {
        set_error(e.what());
        return DEMSPH_ERR_CAPACITY;
    } catch (const HipFailure& e) {
        set_error(e.what());
        return DEMSPH_ERR_HIP;
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return DEMSPH_ERR_HIP;
    } catch (const std::exception& e) {
        set_error(e.what());
        return DEMSPH_ERR_INTERNAL;
    }
}

// ------------------------------------------------- capture-safe runtime calls
// Scenes served together capture step graphs on their own streams from
// several host threads.  HIP rejects some runtime calls made by one thread
// while another captures (allocation, release, device-wide and
// legacy-stream synchronisation), and the rejection can also invalidate the
// other thread's capture.  Every such call in this library and every
// capture hold this one process-wide lock; captures only enqueue, so they
// hold it briefly.  Recursive: a capture that allocates does not deadlock.
std::recursive_mutex& capture_lock();
using CaptureGuard = std::lock_guard<std::recursive_mutex>;

// Device and pinned host memory.  While a scene server runs (block_cache_
// begin .. block_cache_end), released blocks are cached by size class and
// reused instead of being returned with hipFree / hipHostFree, which wait
// for the work of every stream on the device and so would stall all the
// other jobs' streams at each job's set-up and teardown.  Outside serving
// (or for blocks not from the cache) the calls go straight to HIP.
hipError_t dev_malloc(void** p, size_t bytes);
void dev_free(void* p);
hipError_t host_malloc(void** p, size_t bytes);
void host_free(void* p);
// A server starts / ends using the cache; capBytes bounds what it holds
// (live and cached); the last server to end releases the cached blocks.
void block_cache_begin(size_t capBytes);
void block_cache_end();
inline hipError_t stream_create(hipStream_t* s)
{
    CaptureGuard g(capture_lock());
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}
inline void stream_destroy(hipStream_t s)
{
    if (!s) return;
    CaptureGuard g(capture_lock());
    (void)hipStreamDestroy(s);
}
inline void graph_exec_destroy(hipGraphExec_t e)
{
    if (!e) return;
    CaptureGuard g(capture_lock());
    (void)hipGraphExecDestroy(e);
}
// Host <-> device copy ordered on stream s, complete on return.
inline void copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s)
{
    if (!bytes) return;
    DEMSPH_HIP(hipMemcpyAsync(dst, src, bytes, kind, s));
    DEMSPH_HIP(hipStreamSynchronize(s));
}

// Pinned host copies of device arrays taken in stream order; complete once
// wait() returns.  Freed when the owner goes away (on any thread).
struct PinnedSnap {
    std::vector<void*> bufs;
    hipEvent_t ev = nullptr;
    int device = 0;
    PinnedSnap() = default;
    PinnedSnap(const PinnedSnap&) = delete;
    PinnedSnap& operator=(const PinnedSnap&) = delete;
    ~PinnedSnap()
    {
        for (void* b : bufs) host_free(b);
        if (ev) {
            CaptureGuard g(capture_lock());
            (void)hipEventDestroy(ev);
        }
    }
    template <class T>
    const T* take(const T* dev, size_t n, hipStream_t s)
    {
        void* p = nullptr;
        DEMSPH_HIP(host_malloc(&p, (n ? n : 1) * sizeof(T)));
        bufs.push_back(p);
        if (n) DEMSPH_HIP(hipMemcpyAsync(p, dev, n * sizeof(T), hipMemcpyDeviceToHost, s));
        return static_cast<const T*>(p);
    }
    // call after the last take()
    void mark(hipStream_t s)
    {
        DEMSPH_HIP(hipGetDevice(&device));
        {
            CaptureGuard g(capture_lock());
            DEMSPH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        }
        DEMSPH_HIP(hipEventRecord(ev, s));
    }
    // Under the capture lock: a synchronisation from this thread while
    // another thread captures a graph can invalidate that capture.
    void wait() const
    {
        CaptureGuard g(capture_lock());
        DEMSPH_HIP(hipSetDevice(device));
        DEMSPH_HIP(hipEventSynchronize(ev));
    }
};

// ---------------------------------------------------------------- memory
template <class T>
T* dmalloc(size_t n)
{
    void* p = nullptr;
    if (n == 0) n = 1;
    DEMSPH_HIP(dev_malloc(&p, n * sizeof(T)));
    return static_cast<T*>(p);
}
template <class T>
void dfree(T*& p)
{
    dev_free(p);
    p = nullptr;
}

// ---------------------------------------------------------------- geometry
// Uniform grid geometry of the reference's spatialGrid (boundary.h:5-106).
struct GridGeom {
    double minx, miny, minz;
    double maxx, maxy, maxz;
    double cx, cy, cz;
    int nx, ny, nz;
    __host__ __device__ long long numGrids() const { return (long long)nx * ny * nz; }
};

// spatialGrid::set (boundary.h:70-94).  Returns false when the reference
// would ignore the call.
bool grid_set(GridGeom& g, const double* minb, const double* maxb, double cellOneDim);

// Number of bits needed to hold hashes in [0, numGrids].
int key_bits(long long numGrids);

// calculateHash kernel body (neighborSearchKernel.cu:6-27 with
// neighborSearchKernel.cuh:4-14).  Contraction is disabled so that the
// clamp and the cell truncation round exactly as the FP64 oracle does.
__device__ __forceinline__ int cell_hash(const GridGeom& g, double px, double py, double pz)
{
#pragma clang fp contract(off)
    if (px < g.minx) px = g.minx;
    else if (px >= g.maxx) px = g.maxx - 0.5 * g.cx;
    if (py < g.miny) py = g.miny;
    else if (py >= g.maxy) py = g.maxy - 0.5 * g.cy;
    if (pz < g.minz) pz = g.minz;
    else if (pz >= g.maxz) pz = g.maxz - 0.5 * g.cz;
    int ix = (int)((px - g.minx) / g.cx);
    int iy = (int)((py - g.miny) / g.cy);
    int iz = (int)((pz - g.minz) / g.cz);
    return iz * g.ny * g.nx + iy * g.nx + ix;
}

// 27-cell stencil of a query point (ballNeighborSearchKernel.cu:23-33,
// SPHNeighborSearchKernel.cu:21-31): unclamped truncation, grid clamp and
// face shrink.
struct Stencil {
    int cx, cy, cz;
    int lox, loy, loz, hix, hiy, hiz;
};
__device__ __forceinline__ Stencil make_stencil(const GridGeom& g, double px, double py, double pz)
{
#pragma clang fp contract(off)
    Stencil s;
    int gx = (int)((px - g.minx) / g.cx);
    int gy = (int)((py - g.miny) / g.cy);
    int gz = (int)((pz - g.minz) / g.cz);
    s.lox = s.loy = s.loz = -1;
    s.hix = s.hiy = s.hiz = 1;
    if (gx <= 0) { gx = 0; s.lox = 0; }
    if (gx >= g.nx - 1) { gx = g.nx - 1; s.hix = 0; }
    if (gy <= 0) { gy = 0; s.loy = 0; }
    if (gy >= g.ny - 1) { gy = g.ny - 1; s.hiy = 0; }
    if (gz <= 0) { gz = 0; s.loz = 0; }
    if (gz >= g.nz - 1) { gz = g.nz - 1; s.hiz = 0; }
    s.cx = gx; s.cy = gy; s.cz = gz;
    return s;
}

// Exact FP64 neighbour cut test of the reference: (cut*cut - |r|^2) >= 0
// with |r|^2 = x*x + y*y + z*z evaluated left to right without fusion
// (SPHNeighborSearchKernel.cu:47-49, ballNeighborSearchKernel.cu:51-53).
__device__ __forceinline__ bool within_cut(double cut, double rx, double ry, double rz)
{
#pragma clang fp contract(off)
    double d2 = rx * rx + ry * ry + rz * rz;
    return (cut * cut - d2) >= 0.0;
}

// ---------------------------------------------------------------- launch
constexpr int kNumXcd = 8;

// Maps the hardware block id to a logical block id so that each XCD (which
// receives blocks round-robin) works on one contiguous range of logical
// blocks: neighbouring particles then share one L2.
__device__ __forceinline__ int xcd_block()
{
    const int b = blockIdx.x, G = gridDim.x;
    const int q = G / kNumXcd, r = G % kNumXcd;
    const int xcd = b % kNumXcd, pos = b / kNumXcd;
    return (xcd < r) ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}

inline int blocks_for(long long n, int threads) { return (int)((n + threads - 1) / threads); }

// ---------------------------------------------------------------- sort/scan
// Bytes of temporary storage radix_sort_pairs needs for n keys.
size_t radix_sort_temp_bytes(int n);
// Stable LSD radix sort of (keys, vals) on bits [0, nbits).  vals == nullptr
// sorts an implicit iota.  Results end in keys_out / vals_out.  All
// buffers hold n elements; temp holds radix_sort_temp_bytes(n) bytes.
void radix_sort_pairs(const uint32_t* keys_in, const int* vals_in, uint32_t* keys_out, int* vals_out,
                      uint32_t* keys_tmp, int* vals_tmp, void* temp, int n, int nbits, hipStream_t s);
// Exclusive scan of n int32 values; temp holds scan_temp_bytes(n) bytes.
// Writes the total to *total_dev when non-null.
size_t scan_temp_bytes(int n);
void exclusive_scan(const int* in, int* out, int n, void* temp, int* total_dev, hipStream_t s);

// Fill n ints with v.
void fill_i32(int* p, long long n, int v, hipStream_t s);
// Fills p[0, n) and q[0, n) with v in one launch.
void fill2_i32(int* p, int* q, long long n, int v, hipStream_t s);
// p[i] = i.
void iota_i32(int* p, int n, hipStream_t s);
// key[orig[s]] = cell hash of SoA point s (orig == nullptr: key[s]).
void hash_soa_by_orig(int n, const GridGeom& g, const double* px, const double* py, const double* pz, const int* orig,
                      uint32_t* key, hipStream_t s);
// key[i] = cell hash of interleaved point i.
void hash_aos(const double* pos, int n, const GridGeom& g, uint32_t* key, hipStream_t s);
// findStartAndEnd over sorted keys; start/end must hold -1 beforehand.
void cell_ranges(const uint32_t* sorted_keys, int n, long long numGrids, int* start, int* end, hipStream_t s);

// Step-end timestamps: a one-thread kernel appends the device's constant-rate
// wall clock to buf[counter % cap] and advances the device counter, so that
// replays of a captured graph keep appending.  buf == nullptr: no stamp.
struct StepStamps {
    unsigned long long* buf = nullptr;
    int* counter = nullptr;
    int cap = 0;
};
void step_stamp(const StepStamps& st, hipStream_t s);

}  // namespace demsph
