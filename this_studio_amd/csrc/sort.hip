// Stable LSD radix sort, device-wide scan and grid helpers for gfx950.
//
// The reference sorts (cell hash, particle index) pairs with Thrust's
// radix sort_by_key over an iota value array (buildHashStartEnd.cu:50-76),
// so its permutation is the lexicographic (hash, index) order.  This file
// implements that order natively: 8- or 9-bit digits, one histogram kernel
// and one scatter kernel per digit, a wave64 ballot "match-any" to rank equal digits
// stably inside a wavefront, and LDS prefix sums across the wavefronts of a
// workgroup.  Only the bits that hold hashes in [0, numGrids] are sorted.
#include "common.hpp"
#include "vtu.hpp"

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <tuple>
#include <thread>
#include <unordered_map>
#include <vector>

namespace demsph {

namespace {
// Digits of 8 bits, or of 9 when that saves a pass (17- and 18-bit keys,
// grids of up to 262,144 cells, sort in two passes instead of three).
constexpr int kMaxBins = 512;
constexpr int kSortThreads = 256;             // 4 wavefronts
constexpr int kWaves = kSortThreads / 64;
constexpr int kItems = 8;                     // keys per lane
constexpr int kTile = kSortThreads * kItems;  // keys per workgroup

int sort_blocks(int n) { return (n + kTile - 1) / kTile; }
// Up to this many tiles (64k keys) a scatter workgroup reads the whole
// histogram table itself instead of waiting for a device-wide scan.
constexpr int kFusedScatterTiles = 32;

// ----------------------------------------------------------------- scan
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int wave_inclusive_sum(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ int block_exclusive(int v, int& total)
{
    __shared__ int wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int inc = wave_inclusive_sum(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int wprefix = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; w++) {
        if (w < wave) wprefix += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    total = tot;
    return wprefix + inc - v;
}


// Histogram of one digit per tile, written digit-major so that one
// exclusive scan over [digit][block] yields every scatter base.
template <int BITS>
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint32_t* __restrict__ keys, int n, int shift,
                                                             uint32_t* __restrict__ hist, int nblocks)
{
    constexpr int BINS = 1 << BITS;
    __shared__ uint32_t h[BINS];
    for (int i = threadIdx.x; i < BINS; i += kSortThreads) h[i] = 0;
    __syncthreads();
    const int base = blockIdx.x * kTile;
#pragma unroll
    for (int it = 0; it < kItems; it++) {
        const int e = base + it * kSortThreads + threadIdx.x;
        if (e < n) atomicAdd(&h[(keys[e] >> shift) & (BINS - 1)], 1u);
    }
    __syncthreads();
    for (int d = threadIdx.x; d < BINS; d += kSortThreads) hist[(size_t)d * nblocks + blockIdx.x] = h[d];
}

// Per-digit exclusive scan over the tiles, one wave per digit:
// offs[d][t] = sum of hist[d][t'] over t' < t, dtot[d] = digit d's total.
// Replaces a device-wide scan of the whole digit-major table (three
// launches); the scatter adds the digit bases itself.
template <int BINS>
__global__ __launch_bounds__(kSortThreads) void k_radix_tile_scan(const uint32_t* __restrict__ hist, int nblocks,
                                                                  uint32_t* __restrict__ offs,
                                                                  uint32_t* __restrict__ dtot)
{
    const int lane = threadIdx.x & 63;
    const int d = blockIdx.x * kWaves + (threadIdx.x >> 6);
    if (d >= BINS) return;  // wave-uniform
    const uint32_t* h = hist + (size_t)d * nblocks;
    uint32_t* o = offs + (size_t)d * nblocks;
    int carry = 0;
    for (int t0 = 0; t0 < nblocks; t0 += 64) {
        const int t = t0 + lane;
        const int v = t < nblocks ? (int)h[t] : 0;
        const int inc = wave_inclusive_sum(v);
        if (t < nblocks) o[t] = (uint32_t)(carry + inc - v);
        carry += __shfl(inc, 63, 64);
    }
    if (lane == 0) dtot[d] = (uint32_t)carry;
}

// Stable scatter.  Tile element order is wave-major, then item, then lane,
// which is the global index order inside the tile.
//
// FUSED (few tiles): `table` is the digit-major histogram itself and each
// workgroup derives its scatter bases from it (thread d sums digit d over
// all tiles and over the tiles before its own, then a workgroup scan over
// the digits).  Otherwise `table` holds the per-digit tile offsets of
// k_radix_tile_scan and `dtot` the digit totals, whose workgroup scan gives
// the digit bases.
template <bool FUSED, int BITS>
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                                const int* __restrict__ vals_in,
                                                                uint32_t* __restrict__ keys_out,
                                                                int* __restrict__ vals_out, int n, int shift,
                                                                const uint32_t* __restrict__ table,
                                                                const uint32_t* __restrict__ dtot, int nblocks)
{
    constexpr int BINS = 1 << BITS;
    constexpr int DPT = BINS / kSortThreads;  // digits per thread in the fused bases: 1 or 2
    static_assert(DPT == 1 || DPT == 2, "8- or 9-bit digits");
    __shared__ uint32_t wave_count[kWaves][BINS];
    __shared__ uint32_t block_base[BINS];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * BINS; i += kSortThreads) (&wave_count[0][0])[i] = 0;
    if (FUSED) {
        // thread t owns digits DPT t .. DPT t + DPT - 1
        uint32_t tot[DPT], before[DPT];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < DPT; k++) {
            const int d = DPT * threadIdx.x + k;
            tot[k] = 0;
            before[k] = 0;
            for (int b = 0; b < nblocks; b++) {
                const uint32_t c = table[(size_t)d * nblocks + b];
                tot[k] += c;
                before[k] += b < (int)blockIdx.x ? c : 0u;
            }
            sum += tot[k];
        }
        int total;
        uint32_t run = (uint32_t)block_exclusive((int)sum, total);
#pragma unroll
        for (int k = 0; k < DPT; k++) {
            block_base[DPT * threadIdx.x + k] = run + before[k];
            run += tot[k];
        }
    } else {
        uint32_t tot[DPT];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < DPT; k++) {
            tot[k] = dtot[DPT * threadIdx.x + k];
            sum += tot[k];
        }
        int total;
        uint32_t run = (uint32_t)block_exclusive((int)sum, total);
#pragma unroll
        for (int k = 0; k < DPT; k++) {
            const int d = DPT * threadIdx.x + k;
            block_base[d] = run + table[(size_t)d * nblocks + blockIdx.x];
            run += tot[k];
        }
    }
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const int wbase = blockIdx.x * kTile + wave * (64 * kItems);
    uint32_t key[kItems];
    int val[kItems];
    uint32_t pos[kItems];
#pragma unroll
    for (int it = 0; it < kItems; it++) {
        const int e = wbase + it * 64 + lane;
        const bool valid = e < n;
        key[it] = valid ? keys_in[e] : 0u;
        val[it] = valid ? (vals_in ? vals_in[e] : e) : 0;
        const uint32_t digit = (key[it] >> shift) & (BINS - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < BITS; b++) {
            const bool bit = (digit >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const uint32_t before = wave_count[wave][digit];
        pos[it] = before + (uint32_t)__popcll(peers & lt_mask);
        // The lowest lane of each group of equal digits advances the counter.
        const bool leader = valid && ((peers & lt_mask) == 0ull);
        __builtin_amdgcn_wave_barrier();
        if (leader) wave_count[wave][digit] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        if (!valid) pos[it] = 0xffffffffu;
    }
    __syncthreads();
    // Exclusive prefix of the per-wave counts, per digit.
    for (int d = threadIdx.x; d < BINS; d += kSortThreads) {
        uint32_t run = block_base[d];
#pragma unroll
        for (int w = 0; w < kWaves; w++) {
            const uint32_t c = wave_count[w][d];
            wave_count[w][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kItems; it++) {
        if (pos[it] == 0xffffffffu) continue;
        const uint32_t digit = (key[it] >> shift) & (BINS - 1);
        const uint32_t dst = wave_count[wave][digit] + pos[it];
        keys_out[dst] = key[it];
        vals_out[dst] = val[it];
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ partial)
{
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++)
        if (base + i < n) s += in[base + i];
    int total;
    (void)block_exclusive(s, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Single workgroup: exclusive scan of the per-tile partial sums.
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(int* __restrict__ partial, int m, int* __restrict__ total_out)
{
    int carry = 0;
    for (int base = 0; base < m; base += kScanThreads) {
        const int i = base + threadIdx.x;
        const int v = i < m ? partial[i] : 0;
        int total;
        const int ex = block_exclusive(v, total);
        if (i < m) partial[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_final(const int* __restrict__ in, int* __restrict__ out, int n,
                                                             const int* __restrict__ partial)
{
    const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    int total;
    int run = block_exclusive(s, total) + partial[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
}

// The whole exclusive scan in one workgroup of 1024 threads, for inputs of
// up to kScanSingleMax values: one launch instead of three.  Beyond one pass
// of the workgroup the three-kernel scan is faster (a 64k-value digit table
// scanned this way made a 256k-key sort 147 us instead of 44 us).
constexpr int kScan1Threads = 1024;
constexpr int kScanSingleMax = kScan1Threads * kScanItems;
__global__ __launch_bounds__(kScan1Threads) void k_scan_single(const int* __restrict__ in, int* __restrict__ out,
                                                               int n, int* __restrict__ total_out)
{
    __shared__ int wsum[kScan1Threads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int carry = 0;
    for (int base = 0; base < n; base += kScan1Threads * kScanItems) {
        const int b = base + threadIdx.x * kScanItems;
        int v[kScanItems];
        int sum = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            v[i] = b + i < n ? in[b + i] : 0;
            sum += v[i];
        }
        const int inc = wave_inclusive_sum(sum);
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        int wprefix = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < kScan1Threads / 64; w++) {
            if (w < wave) wprefix += wsum[w];
            tot += wsum[w];
        }
        __syncthreads();
        int run = carry + wprefix + inc - sum;
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            if (b + i < n) out[b + i] = run;
            run += v[i];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

__global__ void k_fill_i32(int* p, long long n, int v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill2_i32(int* p, int* q, long long n, int v)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        p[i] = v;
        q[i] = v;
    }
}

__global__ void k_step_stamp(unsigned long long* buf, int* counter, int cap)
{
    if (threadIdx.x != 0) return;
    const int i = *counter;
    buf[i % cap] = wall_clock64();
    *counter = i + 1;
}

__global__ void k_iota_i32(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void k_hash_soa_by_orig(int n, GridGeom g, const double* __restrict__ px, const double* __restrict__ py,
                                   const double* __restrict__ pz, const int* __restrict__ orig,
                                   uint32_t* __restrict__ key)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    key[orig ? orig[s] : s] = (uint32_t)cell_hash(g, px[s], py[s], pz[s]);
}

__global__ void k_hash_aos_by_index(const double* __restrict__ pos, int n, GridGeom g, uint32_t* __restrict__ key)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = (uint32_t)cell_hash(g, pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]);
}

// Radix histogram offsets are uint32; scan them as int (n < 2^31).
}  // namespace

size_t scan_temp_bytes(int n)
{
    const int m = (n + kScanTile - 1) / kScanTile;
    return sizeof(int) * (size_t)(m + 1);
}

void exclusive_scan(const int* in, int* out, int n, void* temp, int* total_dev, hipStream_t s)
{
    if (n <= 0) {
        // A fill kernel rather than a memset node keeps the scan capturable.
        if (total_dev) fill_i32(total_dev, 1, 0, s);
        return;
    }
    if (n <= kScanSingleMax) {
        hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScan1Threads), 0, s, in, out, n, total_dev);
        DEMSPH_HIP(hipGetLastError());
        return;
    }
    const int m = (n + kScanTile - 1) / kScanTile;
    int* partial = static_cast<int*>(temp);
    hipLaunchKernelGGL(k_scan_reduce, dim3(m), dim3(kScanThreads), 0, s, in, n, partial);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, s, partial, m, total_dev);
    hipLaunchKernelGGL(k_scan_final, dim3(m), dim3(kScanThreads), 0, s, in, out, n, partial);
    DEMSPH_HIP(hipGetLastError());
}

// The fixed scratch of demsph_build_cells_noalloc repeats this layout.
size_t radix_sort_temp_bytes(int n)
{
    const int nb = sort_blocks(n > 0 ? n : 1);
    const size_t hist = sizeof(uint32_t) * (size_t)kMaxBins * nb;
    return hist * 2 + sizeof(uint32_t) * kMaxBins + 256;
}

namespace {
template <int BITS>
void radix_passes(const uint32_t* keys_in, const int* vals_in, uint32_t* keys_out, int* vals_out, uint32_t* keys_tmp,
                  int* vals_tmp, void* temp, int n, int passes, hipStream_t s)
{
    constexpr int BINS = 1 << BITS;
    const int nb = sort_blocks(n);
    uint32_t* hist = static_cast<uint32_t*>(temp);
    uint32_t* offs = hist + (size_t)kMaxBins * nb;
    uint32_t* dtot = offs + (size_t)kMaxBins * nb;  // kMaxBins digit totals
    // Ping-pong so that the last pass writes keys_out / vals_out.
    const uint32_t* ksrc = keys_in;
    const int* vsrc = vals_in;
    for (int p = 0; p < passes; p++) {
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* kdst = to_out ? keys_out : keys_tmp;
        int* vdst = to_out ? vals_out : vals_tmp;
        const int shift = p * BITS;
        hipLaunchKernelGGL(k_radix_hist<BITS>, dim3(nb), dim3(kSortThreads), 0, s, ksrc, n, shift, hist, nb);
        if (nb <= kFusedScatterTiles) {
            hipLaunchKernelGGL((k_radix_scatter<true, BITS>), dim3(nb), dim3(kSortThreads), 0, s, ksrc, vsrc, kdst,
                               vdst, n, shift, hist, nullptr, nb);
        } else {
            hipLaunchKernelGGL(k_radix_tile_scan<BINS>, dim3(BINS / kWaves), dim3(kSortThreads), 0, s, hist, nb, offs,
                               dtot);
            hipLaunchKernelGGL((k_radix_scatter<false, BITS>), dim3(nb), dim3(kSortThreads), 0, s, ksrc, vsrc, kdst,
                               vdst, n, shift, offs, dtot, nb);
        }
        DEMSPH_HIP(hipGetLastError());
        ksrc = kdst;
        vsrc = vdst;
    }
}
}  // namespace

void radix_sort_pairs(const uint32_t* keys_in, const int* vals_in, uint32_t* keys_out, int* vals_out,
                      uint32_t* keys_tmp, int* vals_tmp, void* temp, int n, int nbits, hipStream_t s)
{
    if (n <= 0) return;
    if (nbits < 1) nbits = 1;
    const int p8 = (nbits + 7) / 8, p9 = (nbits + 8) / 9;
    if (p9 < p8) radix_passes<9>(keys_in, vals_in, keys_out, vals_out, keys_tmp, vals_tmp, temp, n, p9, s);
    else radix_passes<8>(keys_in, vals_in, keys_out, vals_out, keys_tmp, vals_tmp, temp, n, p8, s);
}

void iota_i32(int* p, int n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota_i32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n);
    DEMSPH_HIP(hipGetLastError());
}

void hash_soa_by_orig(int n, const GridGeom& g, const double* px, const double* py, const double* pz, const int* orig,
                      uint32_t* key, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_hash_soa_by_orig, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, g, px, py, pz, orig, key);
    DEMSPH_HIP(hipGetLastError());
}

void hash_aos(const double* pos, int n, const GridGeom& g, uint32_t* key, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_hash_aos_by_index, dim3(blocks_for(n, 256)), dim3(256), 0, s, pos, n, g, key);
    DEMSPH_HIP(hipGetLastError());
}

void step_stamp(const StepStamps& st, hipStream_t s)
{
    if (!st.buf) return;
    hipLaunchKernelGGL(k_step_stamp, dim3(1), dim3(64), 0, s, st.buf, st.counter, st.cap);
    DEMSPH_HIP(hipGetLastError());
}

std::recursive_mutex& capture_lock()
{
    static std::recursive_mutex m;
    return m;
}

// ------------------------------------------------------------ frame thread
namespace vtu {
namespace {
struct FrameQueue {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::function<void()>> jobs;
    int running = 0;
    bool started = false;
    std::string error;
};
// Never destroyed: the detached writer thread may outlive static teardown.
FrameQueue& frame_queue()
{
    static FrameQueue* q = new FrameQueue();
    return *q;
}
void frame_thread()
{
    FrameQueue& q = frame_queue();
    for (;;) {
        std::function<void()> job;
        {
            std::unique_lock<std::mutex> lk(q.mu);
            q.cv.wait(lk, [&] { return !q.jobs.empty(); });
            job = std::move(q.jobs.front());
            q.jobs.pop_front();
            q.running = 1;
        }
        std::string err;
        try {
            job();
        } catch (const std::exception& e) {
            err = e.what();
        } catch (...) {
            err = "frame output failed";
        }
        job = nullptr;  // releases the job's snapshot before it is reported done
        {
            std::lock_guard<std::mutex> lk(q.mu);
            q.running = 0;
            if (!err.empty() && q.error.empty()) q.error = err;
        }
        q.cv.notify_all();
    }
}
}  // namespace

void post(std::function<void()> job)
{
    FrameQueue& q = frame_queue();
    std::unique_lock<std::mutex> lk(q.mu);
    if (!q.started) {
        std::thread(frame_thread).detach();
        q.started = true;
    }
    q.cv.wait(lk, [&] { return q.jobs.size() + (size_t)q.running < 2; });
    q.jobs.push_back(std::move(job));
    q.cv.notify_all();
}

std::string drain()
{
    FrameQueue& q = frame_queue();
    std::unique_lock<std::mutex> lk(q.mu);
    q.cv.wait(lk, [&] { return q.jobs.empty() && q.running == 0; });
    std::string e;
    e.swap(q.error);
    return e;
}
}  // namespace vtu

// ------------------------------------------------------------ block cache
namespace {

// Size classes: 256 B, then 2^k x {1, 1.25, 1.5, 1.75}, so a cached block
// is at most 25 % larger than the request it serves.
size_t size_class(size_t b)
{
    if (b <= 256) return 256;
    const int k = 63 - __builtin_clzll((unsigned long long)b);
    const size_t base = (size_t)1 << k, step = base >> 2;
    return base + (b - base + step - 1) / step * step;
}

struct BlockCache {
    struct Live {
        size_t bytes;
        int device;
        bool host;
    };
    using Key = std::tuple<bool, int, size_t>;  // host, device, size class
    std::mutex mu;
    int users = 0;
    size_t cap = 0;
    size_t held = 0;  // device bytes owned by the cache, live and free
    size_t free_bytes = 0;
    std::unordered_map<void*, Live> live;
    std::multimap<Key, void*> free_blocks;
};
BlockCache g_cache;

// Takes every free block out of the cache (caller holds g_cache.mu).
std::vector<std::pair<void*, bool>> take_free_blocks()
{
    std::vector<std::pair<void*, bool>> out;
    for (const auto& kv : g_cache.free_blocks) out.emplace_back(kv.second, std::get<0>(kv.first));
    g_cache.free_blocks.clear();
    g_cache.held -= g_cache.free_bytes;
    g_cache.free_bytes = 0;
    return out;
}

// Returns blocks to HIP (without g_cache.mu: the capture lock comes second).
void release_blocks(const std::vector<std::pair<void*, bool>>& blocks)
{
    if (blocks.empty()) return;
    CaptureGuard cg(capture_lock());
    for (const auto& b : blocks) {
        if (b.second) (void)hipHostFree(b.first);
        else (void)hipFree(b.first);
    }
}

hipError_t cached_alloc(void** p, size_t bytes, bool host)
{
    *p = nullptr;
    size_t cls = bytes;
    int dev = 0;
    bool pooled = false;
    std::vector<std::pair<void*, bool>> evict;
    {
        std::lock_guard<std::mutex> lk(g_cache.mu);
        if (g_cache.users > 0) {
            pooled = true;
            if (hipGetDevice(&dev) != hipSuccess) dev = 0;
            cls = size_class(bytes);
            const auto it = g_cache.free_blocks.find(BlockCache::Key{host, dev, cls});
            if (it != g_cache.free_blocks.end()) {
                *p = it->second;
                g_cache.free_blocks.erase(it);
                if (!host) g_cache.free_bytes -= cls;
                g_cache.live[*p] = BlockCache::Live{cls, dev, host};
                return hipSuccess;
            }
            if (!host && g_cache.held + cls > g_cache.cap) evict = take_free_blocks();
        }
    }
    release_blocks(evict);
    hipError_t e;
    {
        CaptureGuard cg(capture_lock());
        e = host ? hipHostMalloc(p, cls, hipHostMallocDefault) : hipMalloc(p, cls);
    }
    if (e == hipErrorOutOfMemory && pooled) {
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lk(g_cache.mu);
            evict = take_free_blocks();
        }
        release_blocks(evict);
        CaptureGuard cg(capture_lock());
        e = host ? hipHostMalloc(p, cls, hipHostMallocDefault) : hipMalloc(p, cls);
    }
    if (e == hipSuccess && pooled) {
        std::lock_guard<std::mutex> lk(g_cache.mu);
        g_cache.live[*p] = BlockCache::Live{cls, dev, host};
        if (!host) g_cache.held += cls;
    }
    return e;
}

// Returns a block: to the cache while a server runs (if the cache made it),
// else to HIP.  A block the cache never saw is of the caller's kind.
void cached_free(void* p, bool host_if_unknown)
{
    if (!p) return;
    bool host = host_if_unknown;
    {
        std::lock_guard<std::mutex> lk(g_cache.mu);
        const auto it = g_cache.live.find(p);
        if (it != g_cache.live.end()) {
            const BlockCache::Live L = it->second;
            g_cache.live.erase(it);
            host = L.host;
            if (g_cache.users > 0) {
                g_cache.free_blocks.emplace(BlockCache::Key{L.host, L.device, L.bytes}, p);
                if (!L.host) g_cache.free_bytes += L.bytes;
                return;
            }
            if (!L.host) g_cache.held -= L.bytes;
        }
    }
    release_blocks({{p, host}});
}

}  // namespace

hipError_t dev_malloc(void** p, size_t bytes) { return cached_alloc(p, bytes, false); }
void dev_free(void* p) { cached_free(p, false); }
hipError_t host_malloc(void** p, size_t bytes) { return cached_alloc(p, bytes, true); }
void host_free(void* p) { cached_free(p, true); }

void block_cache_begin(size_t capBytes)
{
    std::lock_guard<std::mutex> lk(g_cache.mu);
    g_cache.users++;
    g_cache.cap = std::max(g_cache.cap, capBytes);
}

void block_cache_end()
{
    std::vector<std::pair<void*, bool>> evict;
    {
        std::lock_guard<std::mutex> lk(g_cache.mu);
        if (g_cache.users > 0 && --g_cache.users == 0) {
            evict = take_free_blocks();
            g_cache.cap = 0;
        }
    }
    release_blocks(evict);
}

void fill_i32(int* p, long long n, int v, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
    DEMSPH_HIP(hipGetLastError());
}

void fill2_i32(int* p, int* q, long long n, int v, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill2_i32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, q, n, v);
    DEMSPH_HIP(hipGetLastError());
}

bool grid_set(GridGeom& g, const double* minb, const double* maxb, double c)
{
    if (maxb[0] <= minb[0] || maxb[1] <= minb[1] || maxb[2] <= minb[2]) return false;
    if (c < 1.e-20) return false;
    const double dx = maxb[0] - minb[0], dy = maxb[1] - minb[1], dz = maxb[2] - minb[2];
    g.minx = minb[0]; g.miny = minb[1]; g.minz = minb[2];
    g.maxx = maxb[0]; g.maxy = maxb[1]; g.maxz = maxb[2];
    g.nx = dx > c ? (int)(dx / c) : 1;
    g.ny = dy > c ? (int)(dy / c) : 1;
    g.nz = dz > c ? (int)(dz / c) : 1;
    g.cx = dx / (double)g.nx;
    g.cy = dy / (double)g.ny;
    g.cz = dz / (double)g.nz;
    return true;
}

int key_bits(long long numGrids)
{
    int b = 1;
    while (b < 32 && (1ll << b) <= numGrids) b++;
    return b;
}

// ----------------------------------------------------------------- errors
namespace {
thread_local std::string g_last_error = "no error";
}
void set_error(const std::string& msg) { g_last_error = msg; }

// ----------------------------------------------------------------- raw cells
namespace {
// findStartAndEnd (buildHashStartEnd.cu:14-34) over sorted hashes.
__global__ void k_cell_ranges(const uint32_t* __restrict__ sh, int n, long long numGrids, int* __restrict__ start,
                              int* __restrict__ end)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long h = (int)sh[i];
    if (h < 0 || h >= numGrids) return;
    if (i == 0 || sh[i - 1] != sh[i]) start[h] = i;
    if (i == n - 1 || sh[i + 1] != sh[i]) end[h] = i + 1;
}

}  // namespace

void cell_ranges(const uint32_t* sorted_keys, int n, long long numGrids, int* start, int* end, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_cell_ranges, dim3(blocks_for(n, 256)), dim3(256), 0, s, sorted_keys, n, numGrids, start, end);
    DEMSPH_HIP(hipGetLastError());
}

namespace {
struct RawCellCache {
    std::mutex mu;
    int device = -1;
    size_t cap = 0;
    uint32_t* keys = nullptr;
    uint32_t* keys_tmp = nullptr;
    int* vals_tmp = nullptr;
    void* temp = nullptr;
    size_t temp_bytes = 0;
};
RawCellCache g_raw;
}  // namespace

}  // namespace demsph

using namespace demsph;

extern "C" const char* demsph_last_error(void) { return g_last_error.c_str(); }
extern "C" const char* demsph_version(void) { return "demsph-mi355x 0.1 (gfx950)"; }

extern "C" int demsph_grid_geometry(const double* minBound, const double* maxBound, double cellSizeOneDim,
                                    double* cellSize, int* gridSize)
{
    return guarded([&] {
        if (!minBound || !maxBound || !cellSize || !gridSize) throw InvalidArg("null argument");
        GridGeom g;
        if (!grid_set(g, minBound, maxBound, cellSizeOneDim)) throw InvalidArg("degenerate domain or cell size");
        cellSize[0] = g.cx; cellSize[1] = g.cy; cellSize[2] = g.cz;
        gridSize[0] = g.nx; gridSize[1] = g.ny; gridSize[2] = g.nz;
    });
}

extern "C" int demsph_build_cells(const double* position, int* hashIndex, int* hashValue, int* cellHashStart,
                                  int* cellHashEnd, const double* minBound, const double* maxBound,
                                  const double* cellSize, const int* gridSize, size_t numGrids, size_t numObject,
                                  void* stream)
{
    return guarded([&] {
        if (!minBound || !maxBound || !cellSize || !gridSize) throw InvalidArg("null grid argument");
        if (numObject > 0 && (!position || !hashIndex || !hashValue)) throw InvalidArg("null particle array");
        if (numGrids > 0 && (!cellHashStart || !cellHashEnd)) throw InvalidArg("null cell array");
        if (numObject > (size_t)INT32_MAX || numGrids > (size_t)INT32_MAX) throw InvalidArg("size exceeds int32 range");
        hipStream_t s = static_cast<hipStream_t>(stream);
        GridGeom g;
        g.minx = minBound[0]; g.miny = minBound[1]; g.minz = minBound[2];
        g.maxx = maxBound[0]; g.maxy = maxBound[1]; g.maxz = maxBound[2];
        g.cx = cellSize[0]; g.cy = cellSize[1]; g.cz = cellSize[2];
        g.nx = gridSize[0]; g.ny = gridSize[1]; g.nz = gridSize[2];
        fill_i32(cellHashStart, (long long)numGrids, -1, s);
        fill_i32(cellHashEnd, (long long)numGrids, -1, s);
        const int n = (int)numObject;
        if (n == 0) return;
        std::lock_guard<std::mutex> lock(g_raw.mu);
        int dev = 0;
        DEMSPH_HIP(hipGetDevice(&dev));
        if (g_raw.device != dev || g_raw.cap < numObject) {
            // Older buffers may still be read by queued work on another stream.
            {
                CaptureGuard cg(capture_lock());
                DEMSPH_HIP(hipDeviceSynchronize());
            }
            dfree(g_raw.keys); dfree(g_raw.keys_tmp); dfree(g_raw.vals_tmp);
            dev_free(g_raw.temp);
            g_raw.cap = numObject + numObject / 4 + 1024;
            g_raw.keys = dmalloc<uint32_t>(g_raw.cap);
            g_raw.keys_tmp = dmalloc<uint32_t>(g_raw.cap);
            g_raw.vals_tmp = dmalloc<int>(g_raw.cap);
            g_raw.temp_bytes = radix_sort_temp_bytes((int)g_raw.cap);
            DEMSPH_HIP(dev_malloc(&g_raw.temp, g_raw.temp_bytes));
            g_raw.device = dev;
        }
        hash_aos(position, n, g, g_raw.keys, s);
        radix_sort_pairs(g_raw.keys, nullptr, reinterpret_cast<uint32_t*>(hashValue), hashIndex, g_raw.keys_tmp,
                         g_raw.vals_tmp, g_raw.temp, n, key_bits((long long)numGrids), s);
        cell_ranges(reinterpret_cast<const uint32_t*>(hashValue), n, (long long)numGrids, cellHashStart, cellHashEnd, s);
    });
}

// ----------------------------------------------------------------- cells without allocation
// demsph_build_cells_noalloc runs the same kernels as demsph_build_cells with
// its sort scratch (hash keys, the ping-pong keys and values, the digit
// table) in one file-scope device array instead of memory the library
// allocates.  Every kernel writes the scratch it later reads (the histogram
// kernel writes every digit of every tile), so nothing is cleared between
// calls.  Calls that overlap on one GPU share the array and race.
namespace {
constexpr size_t kNoAllocMaxObjects = size_t(1) << 21;  // 2,097,152 objects
constexpr size_t kNoAllocKeyBytes = kNoAllocMaxObjects * sizeof(uint32_t);
// radix_sort_temp_bytes(kNoAllocMaxObjects): two digit tables of 512 bins
// x 1,024 tiles, the digit totals and a pad.
constexpr size_t kNoAllocTempBytes =
    2 * sizeof(uint32_t) * (size_t)kMaxBins * ((kNoAllocMaxObjects + kTile - 1) / kTile) + sizeof(uint32_t) * kMaxBins +
    256;
constexpr size_t kNoAllocBytes = 3 * kNoAllocKeyBytes + kNoAllocTempBytes;  // about 28 MiB per GPU
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {
struct NoAllocScratch {
    std::mutex mu;
    std::map<int, unsigned char*> addr;  // device ordinal -> address of `scratch` there
};
NoAllocScratch g_noalloc;

// Address of `scratch` on the current device, looked up once per device.
unsigned char* noalloc_scratch()
{
    int dev = 0;
    DEMSPH_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_noalloc.mu);
    auto it = g_noalloc.addr.find(dev);
    if (it != g_noalloc.addr.end()) return it->second;
    void* p = nullptr;
    DEMSPH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned a null scratch address");
    g_noalloc.addr[dev] = static_cast<unsigned char*>(p);
    return static_cast<unsigned char*>(p);
}
}  // namespace

extern "C" size_t demsph_build_cells_noalloc_max_objects(void) { return kNoAllocMaxObjects; }

extern "C" int demsph_build_cells_noalloc(const double* position, int* hashIndex, int* hashValue, int* cellHashStart,
                                          int* cellHashEnd, const double* minBound, const double* maxBound,
                                          const double* cellSize, const int* gridSize, size_t numGrids,
                                          size_t numObject, void* stream)
{
    return guarded([&] {
        if (!minBound || !maxBound || !cellSize || !gridSize) throw InvalidArg("null grid argument");
        if (numObject > 0 && (!position || !hashIndex || !hashValue)) throw InvalidArg("null particle array");
        if (numGrids > 0 && (!cellHashStart || !cellHashEnd)) throw InvalidArg("null cell array");
        if (numGrids > (size_t)INT32_MAX) throw InvalidArg("numGrids exceeds int32 range");
        if (numObject > kNoAllocMaxObjects)
            throw CapacityError("demsph_build_cells_noalloc: numObject " + std::to_string(numObject) +
                                " exceeds the fixed scratch capacity of " + std::to_string(kNoAllocMaxObjects) +
                                " objects (use demsph_build_cells)");
        unsigned char* base = noalloc_scratch();
        hipStream_t s = static_cast<hipStream_t>(stream);
        GridGeom g;
        g.minx = minBound[0]; g.miny = minBound[1]; g.minz = minBound[2];
        g.maxx = maxBound[0]; g.maxy = maxBound[1]; g.maxz = maxBound[2];
        g.cx = cellSize[0]; g.cy = cellSize[1]; g.cz = cellSize[2];
        g.nx = gridSize[0]; g.ny = gridSize[1]; g.nz = gridSize[2];
        fill_i32(cellHashStart, (long long)numGrids, -1, s);
        fill_i32(cellHashEnd, (long long)numGrids, -1, s);
        const int n = (int)numObject;
        if (n == 0) return;
        uint32_t* keys = reinterpret_cast<uint32_t*>(base);
        uint32_t* keys_tmp = reinterpret_cast<uint32_t*>(base + kNoAllocKeyBytes);
        int* vals_tmp = reinterpret_cast<int*>(base + 2 * kNoAllocKeyBytes);
        void* temp = base + 3 * kNoAllocKeyBytes;
        hash_aos(position, n, g, keys, s);
        radix_sort_pairs(keys, nullptr, reinterpret_cast<uint32_t*>(hashValue), hashIndex, keys_tmp, vals_tmp, temp, n,
                         key_bits((long long)numGrids), s);
        cell_ranges(reinterpret_cast<const uint32_t*>(hashValue), n, (long long)numGrids, cellHashStart, cellHashEnd, s);
    });
}
