// primitives.hip - HBM arena, scans, stable LSD radix sort, fills, reductions.
//
// These replace the Thrust primitives and cached_allocator the reference builds
// on (LouvainCUDA/cached_allocator.h:12-108; exclusive_scan at
// NeighbourCompute.cu:113 and OneLevelKernel.cu:746; sort_by_key at
// NeighbourCompute.cu:82, RenumberKernel.cu:123,133, InducedGraphKernel.cu:91).
// All kernels assume 64-lane wavefronts (gfx950).
#include <cstdio>
#include <cstring>
#include <type_traits>

#include "engine.hpp"

namespace lv {

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    fprintf(stderr, "[louvain_amd] HIP error %s (%d) in %s at %s:%d\n", hipGetErrorString(e),
            (int)e, what, file, line);
    return e == hipErrorOutOfMemory ? LV_ERR_NOMEM : LV_ERR_DEVICE;
}

// ---------------------------------------------------------------------------
// Arena

Arena::~Arena()
{
    if (base_ && owned_) (void)hipFree(base_);
}

void Arena::adopt(void *base, size_t bytes)
{
    base_ = (char *)base;
    cap_ = bytes & ~(size_t)255;
    owned_ = false;
    used_.clear();
    free_.clear();
    if (cap_) free_[0] = cap_;
    used_bytes_ = high_ = 0;
}

int Arena::reserve(size_t bytes)
{
    bytes = (bytes + 255) & ~(size_t)255;
    if (base_ && cap_ >= bytes) return LV_OK;
    if (!owned_) return LV_ERR_NOMEM;  // a borrowed block cannot grow
    if (!used_.empty()) return LV_ERR_ARG;  // cannot grow under live allocations
    if (base_) {
        (void)hipFree(base_);
        base_ = nullptr;
        cap_ = 0;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return LV_ERR_NOMEM;
    }
    base_ = (char *)p;
    cap_ = bytes;
    free_.clear();
    free_[0] = cap_;
    used_bytes_ = 0;
    return LV_OK;
}

void *Arena::alloc(size_t bytes)
{
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= bytes) {
            size_t off = it->first, sz = it->second;
            free_.erase(it);
            if (sz > bytes) free_[off + bytes] = sz - bytes;
            used_[off] = bytes;
            used_bytes_ += bytes;
            if (used_bytes_ > high_) high_ = used_bytes_;
            return base_ + off;
        }
    }
    return nullptr;
}

void Arena::release(void *p)
{
    if (!p) return;
    size_t off = (size_t)((char *)p - base_);
    auto it = used_.find(off);
    if (it == used_.end()) return;
    size_t sz = it->second;
    used_.erase(it);
    used_bytes_ -= sz;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && nx->first == off + sz) {
        sz += nx->second;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += sz;
            return;
        }
    }
    free_[off] = sz;
}

void Arena::restore(const std::map<size_t, size_t> &m)
{
    std::vector<size_t> extra;
    for (const auto &u : used_)
        if (m.find(u.first) == m.end()) extra.push_back(u.first);
    for (size_t off : extra) release(base_ + off);
}

void Arena::release_all()
{
    used_.clear();
    free_.clear();
    if (cap_) free_[0] = cap_;
    used_bytes_ = 0;
}

// ---------------------------------------------------------------------------
// small helpers

int bits_for(uint64_t maxval)
{
    int b = 0;
    while (b < 64 && (maxval >> b) != 0) ++b;
    return b;
}

static inline int grid_for(int64_t n, int block, int64_t cap = 1 << 16)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

double q_from_aggregates(int64_t sum_in, uint64_t d2_lo, uint64_t d2_hi, int64_t m, double gamma)
{
    if (m <= 0) return 0.0;
    double dm = (double)m;
    double a = (double)sum_in / dm;
    double d2 = (double)d2_hi * 18446744073709551616.0 + (double)d2_lo;
    double b = d2 / ((4.0 * dm) * dm);
    return a - gamma * b;
}

template <typename T>
__global__ void k_fill(T *p, int64_t n, T v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_iota(int32_t *p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int32_t)i;
}

__global__ void k_copy_i32(int32_t *d, const int32_t *s, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        d[i] = s[i];
}

__device__ inline long long wave_sum_ll(long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void k_sum_i64(const int64_t *p, int64_t n, long long *out)
{
    long long acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        acc += p[i];
    acc = wave_sum_ll(acc);
    if ((threadIdx.x & 63) == 0 && acc != 0) atomicAdd((ull *)out, (ull)acc);
}

// *out += sum of p[idx[i]] over i < n (one atomic per wave).
__global__ void k_sum_gather_i32(const int32_t *idx, int64_t n, const int32_t *p, long long *out)
{
    long long acc = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        acc += p[idx[i]];
    acc = wave_sum_ll(acc);
    if ((threadIdx.x & 63) == 0 && acc != 0) atomicAdd((ull *)out, (ull)acc);
}

void sum_gather_i32(hipStream_t s, const int32_t *idx, int64_t n, const int32_t *p, int64_t *out_dev)
{
    if (n > 0)
        hipLaunchKernelGGL(k_sum_gather_i32, grid_for(n, 256, 1024), 256, 0, s, idx, n, p,
                           (long long *)out_dev);
}

void fill_i32(hipStream_t s, int32_t *p, int64_t n, int32_t v)
{
    if (n > 0) hipLaunchKernelGGL(k_fill<int32_t>, grid_for(n, 256), 256, 0, s, p, n, v);
}

void fill_i64(hipStream_t s, int64_t *p, int64_t n, int64_t v)
{
    if (n > 0) hipLaunchKernelGGL(k_fill<int64_t>, grid_for(n, 256), 256, 0, s, p, n, v);
}

// Mailbox: the segments' words into page-locked host memory, each thread's
// stores released at system scope, then the flag word set to seq, which the
// host polls.  One block; the segments are a few hundred words at most.
__global__ __launch_bounds__(256) void k_mailbox(MboxArgs a)
{
    for (int k = 0; k < a.nseg; ++k) {
        const MboxSeg g = a.seg[k];
        for (uint32_t i = threadIdx.x; i < g.words; i += 256) a.box[g.off + i] = g.src[i];
    }
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(a.flag, a.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void launch_mailbox(hipStream_t s, const MboxArgs &a)
{
    hipLaunchKernelGGL(k_mailbox, 1, 256, 0, s, a);
}

void iota_i32(hipStream_t s, int32_t *p, int64_t n)
{
    if (n > 0) hipLaunchKernelGGL(k_iota, grid_for(n, 256), 256, 0, s, p, n);
}

void copy_i32(hipStream_t s, int32_t *dst, const int32_t *src, int64_t n)
{
    if (n > 0) hipLaunchKernelGGL(k_copy_i32, grid_for(n, 256), 256, 0, s, dst, src, n);
}

void sum_i64(hipStream_t s, const int64_t *p, int64_t n, int64_t *out_dev)
{
    if (n > 0)
        hipLaunchKernelGGL(k_sum_i64, grid_for(n, 256, 4096), 256, 0, s, p, n,
                           (long long *)out_dev);
}

// ---------------------------------------------------------------------------
// Exclusive scan: per-tile reduce, scan of tile sums, per-tile scan + offset.

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

template <typename T>
__device__ inline T wave_incl_scan(T v, int lane)
{
    for (int off = 1; off < 64; off <<= 1) {
        T o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

// Exclusive scan of one value per thread across the block; returns the
// exclusive prefix and writes the block total to *total.
template <typename T>
__device__ inline T block_excl_scan(T v, T *total)
{
    __shared__ T s_w[kScanBlock / 64];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T inc = wave_incl_scan(v, lane);
    if (lane == 63) s_w[wave] = inc;
    __syncthreads();
    T wpre = 0, tot = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
        if (w < wave) wpre += s_w[w];
        tot += s_w[w];
    }
    __syncthreads();
    *total = tot;
    return wpre + inc - v;
}

template <typename T>
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const T *in, int64_t n, T *partials)
{
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    T acc = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j)
        if (base + j < n) acc += in[base + j];
    T tot;
    (void)block_excl_scan(acc, &tot);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

template <typename T>
__global__ __launch_bounds__(kScanBlock) void k_scan_partials(T *partials, int64_t nb, T *total)
{
    T carry = 0;
    for (int64_t base = 0; base < nb; base += kScanTile) {
        int64_t i0 = base + (int64_t)threadIdx.x * kScanItems;
        T loc[kScanItems];
        T acc = 0;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            loc[j] = (i0 + j < nb) ? partials[i0 + j] : 0;
            acc += loc[j];
        }
        T tot;
        T pre = block_excl_scan(acc, &tot) + carry;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            if (i0 + j < nb) partials[i0 + j] = pre;
            pre += loc[j];
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <typename T>
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const T *in, T *out, int64_t n,
                                                          const T *partials)
{
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    T loc[kScanItems];
    T acc = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        loc[j] = (base + j < n) ? in[base + j] : 0;
        acc += loc[j];
    }
    T tot;
    T pre = block_excl_scan(acc, &tot) + partials[blockIdx.x];
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) out[base + j] = pre;
        pre += loc[j];
    }
}

// One block for a short array (hub plans, segment plans): tiles in sequence
// with a carry, one launch instead of three.  in may equal out.
constexpr int64_t kScanSingleMax = 8 * kScanTile;

template <typename T>
__global__ __launch_bounds__(kScanBlock) void k_scan_single(const T *in, T *out, int64_t n,
                                                           T *total)
{
    T carry = 0;
    for (int64_t base = 0; base < n; base += kScanTile) {
        const int64_t i0 = base + (int64_t)threadIdx.x * kScanItems;
        T loc[kScanItems];
        T acc = 0;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            loc[j] = (i0 + j < n) ? in[i0 + j] : 0;
            acc += loc[j];
        }
        T tot;
        T pre = block_excl_scan(acc, &tot) + carry;
#pragma unroll
        for (int j = 0; j < kScanItems; ++j) {
            if (i0 + j < n) out[i0 + j] = pre;
            pre += loc[j];
        }
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <typename T>
static int scan_exclusive(Arena &a, hipStream_t s, const T *in, T *out, int64_t n, T *total_dev)
{
    if (n <= 0) {
        if (total_dev) LV_HIP(hipMemsetAsync(total_dev, 0, sizeof(T), s));
        return LV_OK;
    }
    if (n <= kScanSingleMax) {
        hipLaunchKernelGGL(k_scan_single<T>, 1, kScanBlock, 0, s, in, out, n, total_dev);
        LV_HIP(hipGetLastError());
        return LV_OK;
    }
    int64_t nb = (n + kScanTile - 1) / kScanTile;
    T *partials = nullptr;
    LV_TRY(arena_alloc(a, partials, nb));
    hipLaunchKernelGGL(k_scan_reduce<T>, (int)nb, kScanBlock, 0, s, in, n, partials);
    hipLaunchKernelGGL(k_scan_partials<T>, 1, kScanBlock, 0, s, partials, nb, total_dev);
    hipLaunchKernelGGL(k_scan_apply<T>, (int)nb, kScanBlock, 0, s, in, out, n, partials);
    LV_HIP(hipGetLastError());
    // The arena only recycles the block for later launches on the same stream.
    arena_free(a, partials);
    return LV_OK;
}

int scan_exclusive_i64(Arena &a, hipStream_t s, const int64_t *in, int64_t *out, int64_t n,
                       int64_t *total_dev)
{
    return scan_exclusive<int64_t>(a, s, in, out, n, total_dev);
}

int scan_exclusive_i32(Arena &a, hipStream_t s, const int32_t *in, int32_t *out, int64_t n,
                       int32_t *total_dev)
{
    return scan_exclusive<int32_t>(a, s, in, out, n, total_dev);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.
//
// Each block owns a tile of `rounds` x 256 consecutive items.  The histogram
// kernel writes digit-major counts hist[d * nblocks + b]; after an exclusive
// scan each (digit, block) pair owns a contiguous output range.  The scatter
// kernel walks its tile in rounds of 256 items (item order = round, wave,
// lane) and ranks each item among equal digits of its wave with eight
// ballots, so the order of equal digits is preserved.

constexpr int kRadixBlock = 256;

// Workgroups are dispatched to the 8 XCDs round-robin.  The radix kernels work
// on logical block xcd_block(blockIdx): consecutive logical blocks - the ones
// whose digit-major histogram entries share a cache line - then run on one
// XCD, so each line is assembled in one L2 instead of from eight.
__device__ __forceinline__ int64_t xcd_block(int64_t bid, int64_t nblocks)
{
    const int64_t q = nblocks >> 3, r = nblocks & 7, x = bid & 7, i = bid >> 3;
    return x * q + (x < r ? x : r) + i;
}

template <typename K>
__global__ __launch_bounds__(kRadixBlock) void k_radix_hist(const K *keys, int64_t n, int shift,
                                                           int rounds, int64_t *hist,
                                                           int64_t nblocks)
{
    __shared__ int s_h[256];
    s_h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t lb = xcd_block(blockIdx.x, nblocks);
    int64_t base = lb * rounds * kRadixBlock;
    // eight independent loads in flight before their eight LDS increments
    constexpr int U = 8;
    for (int r = 0; r < rounds; r += U) {
        int d[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = base + (int64_t)(r + u) * kRadixBlock + threadIdx.x;
            d[u] = (r + u < rounds && i < n) ? (int)((keys[i] >> shift) & 255) : -1;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (d[u] >= 0) atomicAdd(&s_h[d[u]], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * nblocks + lb] = s_h[threadIdx.x];
}

// Scatter of one digit pass.  A block round covers kRadixTile consecutive
// items: wave w owns the round's w-th quarter as kRadixSub sub-rounds of 64
// lanes.  Each sub-round ranks its items among equal digits with eight
// ballots and the wave's running per-digit counters in LDS, so item order
// (round, wave, sub-round, lane) is preserved: the sort is stable.  The round
// is then reordered by digit in LDS and written out by consecutive threads,
// so each digit's items of the round leave as one contiguous run.
constexpr int kRadixSub = 8;  // sub-rounds of 64 items per wave per round
constexpr int kRadixTile = kRadixBlock * kRadixSub;

// Exclusive scan of one int per thread over a 256-thread block.
__device__ __forceinline__ int block_exclusive_scan_256(int x, int *s_wave)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    int pre = 0;
    for (int w = 0; w < wave; ++w) pre += s_wave[w];
    return pre + inc - x;
}

// Keys without values (V = NoVal): only the keys move.
struct NoVal {};

template <typename K, typename V>
__global__ __launch_bounds__(kRadixBlock) void k_radix_scatter(const K *kin, const V *vin, K *kout,
                                                              V *vout, int64_t n, int shift,
                                                              int rounds, const int64_t *hscan,
                                                              int64_t nblocks)
{
    constexpr bool HV = !std::is_same<V, NoVal>::value;
    constexpr int NW = kRadixBlock / 64;
    __shared__ int64_t s_base[256];  // next output position of each digit for this block
    __shared__ int s_cnt[NW][256];   // per-wave digit counts, then tile offsets
    __shared__ int s_start[256];     // first tile position of each digit
    __shared__ int s_wave[NW];
    __shared__ K s_k[kRadixTile];
    __shared__ V s_v[HV ? kRadixTile : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t lb = xcd_block(blockIdx.x, nblocks);
    s_base[tid] = hscan[(int64_t)tid * nblocks + lb];
    const ull lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const int64_t base = lb * rounds * kRadixBlock;
    const int64_t end = min(n, base + (int64_t)rounds * kRadixBlock);
    for (int64_t r0 = base; r0 < end; r0 += kRadixTile) {
#pragma unroll
        for (int w = 0; w < NW; ++w) s_cnt[w][tid] = 0;
        __syncthreads();
        K k[kRadixSub];
        V v[kRadixSub];
        int d[kRadixSub], rank[kRadixSub];
        bool valid[kRadixSub];
#pragma unroll
        for (int u = 0; u < kRadixSub; ++u) {
            const int64_t i = r0 + wave * 64 * kRadixSub + u * 64 + lane;
            valid[u] = i < end;
            k[u] = valid[u] ? kin[i] : (K)0;
            if constexpr (HV) v[u] = valid[u] ? vin[i] : (V)0;
            d[u] = (int)((k[u] >> shift) & 255);
        }
#pragma unroll
        for (int u = 0; u < kRadixSub; ++u) {
            ull mask = __ballot(valid[u]);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                int bit = (d[u] >> b) & 1;
                ull bb = __ballot(bit);
                mask &= bit ? bb : ~bb;
            }
            const int before = ((volatile int *)s_cnt[wave])[d[u]];
            rank[u] = before + __popcll(mask & lt);
            if (valid[u] && __popcll(mask & lt) == 0) s_cnt[wave][d[u]] = before + __popcll(mask);
            __builtin_amdgcn_wave_barrier();
        }
        __syncthreads();
        int tot = 0, woff[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            woff[w] = tot;
            tot += s_cnt[w][tid];
        }
        const int start = block_exclusive_scan_256(tot, s_wave);
        s_start[tid] = start;
#pragma unroll
        for (int w = 0; w < NW; ++w) s_cnt[w][tid] = start + woff[w];
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kRadixSub; ++u) {
            if (valid[u]) {
                const int lp = s_cnt[wave][d[u]] + rank[u];
                s_k[lp] = k[u];
                if constexpr (HV) s_v[lp] = v[u];
            }
        }
        __syncthreads();
        const int items = (int)min((int64_t)kRadixTile, end - r0);
        for (int t = tid; t < items; t += kRadixBlock) {
            const K kk = s_k[t];
            const int dd = (int)((kk >> shift) & 255);
            const int64_t dst = s_base[dd] + (t - s_start[dd]);
            kout[dst] = kk;
            if constexpr (HV) vout[dst] = s_v[t];
        }
        __syncthreads();
        s_base[tid] += tot;
    }
}

template <typename K, typename V>
static int radix_sort(Arena &a, hipStream_t s, K *keys, V *vals, K *ktmp, V *vtmp, int64_t n,
                      int bits, K **kout, V **vout, int lo = 0)
{
    *kout = keys;
    *vout = vals;
    if (n <= 1 || bits <= lo) return LV_OK;
    // at least one full round per block, at most 8192 blocks: the digit-major
    // histogram (256 x blocks) and its scan stay small next to the data
    int rounds = kRadixSub;
    while ((n + (int64_t)rounds * kRadixBlock - 1) / ((int64_t)rounds * kRadixBlock) > 8192)
        rounds *= 2;
    int64_t nblocks = (n + (int64_t)rounds * kRadixBlock - 1) / ((int64_t)rounds * kRadixBlock);
    int64_t *hist = nullptr;
    LV_TRY(arena_alloc(a, hist, 256 * nblocks));
    K *ka = keys, *kb = ktmp;
    V *va = vals, *vb = vtmp;
    for (int shift = lo; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist<K>, (int)nblocks, kRadixBlock, 0, s, ka, n, shift, rounds,
                           hist, nblocks);
        int rc = scan_exclusive_i64(a, s, hist, hist, 256 * nblocks, nullptr);
        if (rc != LV_OK) {
            arena_free(a, hist);
            return rc;
        }
        hipLaunchKernelGGL((k_radix_scatter<K, V>), (int)nblocks, kRadixBlock, 0, s, ka, va, kb,
                           vb, n, shift, rounds, hist, nblocks);
        K *tk = ka;
        ka = kb;
        kb = tk;
        V *tv = va;
        va = vb;
        vb = tv;
    }
    LV_HIP(hipGetLastError());
    arena_free(a, hist);
    *kout = ka;
    *vout = va;
    return LV_OK;
}

int radix_sort_u64_i64(Arena &a, hipStream_t s, uint64_t *keys, int64_t *vals, uint64_t *ktmp,
                       int64_t *vtmp, int64_t n, int bits, uint64_t **keys_out,
                       int64_t **vals_out)
{
    return radix_sort<uint64_t, int64_t>(a, s, keys, vals, ktmp, vtmp, n, bits, keys_out,
                                         vals_out);
}

int radix_sort_u32_i32(Arena &a, hipStream_t s, uint32_t *keys, int32_t *vals, uint32_t *ktmp,
                       int32_t *vtmp, int64_t n, int bits, uint32_t **keys_out,
                       int32_t **vals_out)
{
    return radix_sort<uint32_t, int32_t>(a, s, keys, vals, ktmp, vtmp, n, bits, keys_out,
                                         vals_out);
}

int radix_sort_u32_u64(Arena &a, hipStream_t s, uint32_t *keys, uint64_t *vals, uint32_t *ktmp,
                       uint64_t *vtmp, int64_t n, int bits, uint32_t **keys_out,
                       uint64_t **vals_out)
{
    return radix_sort<uint32_t, uint64_t>(a, s, keys, vals, ktmp, vtmp, n, bits, keys_out,
                                          vals_out);
}

int radix_sort_u64_keys(Arena &a, hipStream_t s, uint64_t *keys, uint64_t *ktmp, int64_t n,
                        int lo, uint64_t **keys_out)
{
    NoVal *none = nullptr, *out = nullptr;
    return radix_sort<uint64_t, NoVal>(a, s, keys, none, ktmp, none, n, 64, keys_out, &out, lo);
}

int radix_sort_u32_u32(Arena &a, hipStream_t s, uint32_t *keys, uint32_t *vals, uint32_t *ktmp,
                       uint32_t *vtmp, int64_t n, int bits, uint32_t **keys_out,
                       uint32_t **vals_out)
{
    return radix_sort<uint32_t, uint32_t>(a, s, keys, vals, ktmp, vtmp, n, bits, keys_out,
                                          vals_out);
}

}  // namespace lv
