// louvain_kernels.hip - the Louvain kernels for gfx950 (64-lane wavefronts).
//
// Every weight sum is an exact int64 count of units 2^-s; gains and Q are
// formed in FP64 from exact integers (build flag -ffp-contract=off keeps the
// multiply and subtract of W - ((gamma*deg)*f) separate).
#include <climits>

#include "device.hpp"

namespace lv {

// ---------------------------------------------------------------------------
// Input validation and weight units (the reference reads float weights and
// truncates sums to int, LouvainCUDA/InitKernel.cu:119, StatusCUDA.h:31-41).

// Float weights also leave the bit pattern of the largest one (valid weights
// are non-negative, where bit patterns order as values), which bounds the
// units and so decides their width.
__global__ void k_weight_exponent(const void *w, int wtype, int64_t e, DevFlags *f)
{
    int smax = 0;
    int bad = 0;
    uint32_t mb = 0;
    GRID_STRIDE(i, e)
    {
        if (wtype == LV_WEIGHT_INT32) {
            if (((const int32_t *)w)[i] < 0) bad |= 1;
        } else {
            const uint32_t x = ((const uint32_t *)w)[i];
            int s = f32_unit_exponent(x);
            if (s < 0) {
                bad |= 1;
            } else {
                smax = max(smax, s);
                mb = max(mb, x & 0x7fffffffu);
            }
        }
    }
    smax = wave_max_i(smax);
    bad = wave_or_i(bad);
    mb = (uint32_t)wave_max_i((int)mb);
    if ((threadIdx.x & 63) == 0) {
        if (smax) atomicMax(&f->scale, smax);
        if (mb) atomicMax(&f->max_bits, mb);
        if (bad) atomicOr(&f->err, 1);
    }
}

// Units of every weight into wu (8 bytes) or wu32 (4 bytes, when the largest
// fits), or neither (int32 weights in place: only the sums and checks).
__global__ __launch_bounds__(256) void k_weight_units(const void *w, int wtype, int64_t e, int scale,
                                                     int64_t *wu, uint32_t *wu32, DevFlags *f)
{
    ull tot = 0, msum = 0, umax = 0;
    int bad = 0;
    GRID_STRIDE(i, e)
    {
        const int64_t u = weight_units(((const uint32_t *)w)[i], wtype, scale, bad);
        if (wu32)
            wu32[i] = (uint32_t)u;
        else if (wu)
            wu[i] = u;
        umax = max(umax, (ull)u);
        tot = sat_add(tot, (ull)(u == 0 ? 1 : u));
        msum = sat_add(msum, (ull)u);
    }
    block_units_flush(tot, msum, umax, bad, f);
}

__global__ void k_check_endpoints(const int32_t *src, const int32_t *dst, int64_t e, int32_t n,
                                  DevFlags *f)
{
    int bad = 0;
    GRID_STRIDE(i, e)
    {
        int32_t a = src[i], b = dst[i];
        if (a < 0 || a >= n || b < 0 || b >= n) bad |= 1;
    }
    bad = wave_or_i(bad);
    if ((threadIdx.x & 63) == 0 && bad) atomicOr(&f->err, 1);
}

__global__ void k_ids_to_keys(const int32_t *ids, uint32_t *keys, int32_t n)
{
    GRID_STRIDE(i, n) keys[i] = (uint32_t)ids[i] ^ 0x80000000u;  // signed order
}

__global__ void k_check_sorted_unique(const uint32_t *keys, int64_t n, DevFlags *f)
{
    int bad = 0;
    GRID_STRIDE(i, n)
    {
        if (i > 0 && keys[i] == keys[i - 1]) bad |= 1;
    }
    bad = wave_or_i(bad);
    if ((threadIdx.x & 63) == 0 && bad) atomicOr(&f->err, 1);
}

__global__ void k_map_endpoints(const uint32_t *skeys, const int32_t *spos, int32_t n, int32_t *ep,
                                int64_t e, DevFlags *f)
{
    int bad = 0;
    GRID_STRIDE(i, e)
    {
        uint32_t key = (uint32_t)ep[i] ^ 0x80000000u;
        int32_t lo = 0, hi = n - 1, found = -1;
        while (lo <= hi) {
            int32_t mid = lo + ((hi - lo) >> 1);
            uint32_t k = skeys[mid];
            if (k == key) {
                found = spos[mid];
                break;
            }
            if (k < key)
                lo = mid + 1;
            else
                hi = mid - 1;
        }
        if (found < 0) {
            bad |= 1;
            found = 0;
        }
        ep[i] = found;
    }
    bad = wave_or_i(bad);
    if ((threadIdx.x & 63) == 0 && bad) atomicOr(&f->err, 1);
}

void launch_weight_exponent(hipStream_t s, const void *w, int wtype, int64_t e, DevFlags *f)
{
    if (e > 0) hipLaunchKernelGGL(k_weight_exponent, grid_for(e, 256, 8192), 256, 0, s, w, wtype, e, f);
}

void launch_weight_units(hipStream_t s, const void *w, int wtype, int64_t e, int scale,
                         int64_t *wu, uint32_t *wu32, DevFlags *f)
{
    if (e > 0)
        hipLaunchKernelGGL(k_weight_units, grid_for(e, 256, 8192), 256, 0, s, w, wtype, e, scale,
                           wu, wu32, f);
}

void launch_check_endpoints(hipStream_t s, const int32_t *src, const int32_t *dst, int64_t e,
                            int32_t n, DevFlags *f)
{
    if (e > 0) hipLaunchKernelGGL(k_check_endpoints, grid_for(e, 256, 8192), 256, 0, s, src, dst, e, n, f);
}

void launch_check_sorted_unique(hipStream_t s, const uint32_t *keys, int64_t n, DevFlags *f)
{
    if (n > 1) hipLaunchKernelGGL(k_check_sorted_unique, grid_for(n, 256), 256, 0, s, keys, n, f);
}

void launch_map_endpoints(hipStream_t s, const uint32_t *sorted_ids, const int32_t *sorted_pos,
                          int32_t n, int32_t *ep, int64_t e, DevFlags *f)
{
    if (e > 0)
        hipLaunchKernelGGL(k_map_endpoints, grid_for(e, 256), 256, 0, s, sorted_ids, sorted_pos, n,
                           ep, e, f);
}

void launch_ids_to_keys(hipStream_t s, const int32_t *ids, uint32_t *keys, int32_t n)
{
    if (n > 0) hipLaunchKernelGGL(k_ids_to_keys, grid_for(n, 256), 256, 0, s, ids, keys, n);
}

// ---------------------------------------------------------------------------
// CSR build (CalculateNeighbour, LouvainCUDA/NeighbourCompute.cu:42-118, plus the
// degree and self-loop part of initValuesKernel, InitKernel.cu:37-77,101-126).
// Every non-self edge e gives two directed entries, coded 2e (row src, col dst)
// and 2e+1 (row dst, col src); a stable radix sort by row groups them (no
// atomics on the counters of high-degree rows).  Self-loops stay out of the
// rows: selfsum[i] = sum of (w==0?1:w) over the vertex's loops, k_i gains 4x
// that, loop_i = the weight of the vertex's last loop in input order.

__global__ void k_csr_entries(const int32_t *src, const int32_t *dst, int64_t e, int32_t n,
                              uint32_t *keys, uint32_t *codes)
{
    GRID_STRIDE(i, e)
    {
        int32_t a = src[i], b = dst[i];
        bool self = (a == b);
        keys[2 * i] = self ? (uint32_t)n : (uint32_t)a;  // self-loops sort past every row
        keys[2 * i + 1] = self ? (uint32_t)n : (uint32_t)b;
        codes[2 * i] = (uint32_t)(2 * i);
        codes[2 * i + 1] = (uint32_t)(2 * i + 1);
    }
}

// payload = col << wbits | w (the host checks that w < 2^wbits and col < 2^(64-wbits)).
// Both entries of every edge as one word, row << (64 - rb) | col << wb | w with
// rb = bits of n (self-loops keyed n, past every row), cb = bits of n - 1 and
// wb = 64 - rb - cb (when every weight fits wb bits).
__global__ void k_csr_entries1(const int32_t *src, const int32_t *dst, WUnits w, int64_t e,
                               int32_t n, int rb, int cb, uint64_t *words)
{
    const int wb = 64 - rb - cb;
    GRID_STRIDE(i, e)
    {
        const int32_t a = src[i], b = dst[i];
        const uint64_t u = (uint64_t)w[i];
        const bool self = (a == b);
        const uint64_t ka = self ? (uint64_t)n : (uint64_t)(uint32_t)a;
        const uint64_t kb = self ? (uint64_t)n : (uint64_t)(uint32_t)b;
        words[2 * i] = (ka << (64 - rb)) | ((uint64_t)(uint32_t)b << wb) | u;
        words[2 * i + 1] = (kb << (64 - rb)) | ((uint64_t)(uint32_t)a << wb) | u;
    }
}

__global__ __launch_bounds__(256) void k_csr_unpack1(const uint64_t *words, int64_t nnz, int rb,
                                                    int cb, int32_t *col, int64_t *cw, uint32_t *cw32,
                                                    int64_t *kdeg, DevFlags *f)
{
    const int wb = 64 - rb - cb;
    const uint64_t cmask = (1ull << cb) - 1ull, wmask = (1ull << wb) - 1ull;
    constexpr int64_t SPAN = 256 * kSpanRounds;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int bad = 0;
    SegCarry sc;
    for (int64_t base = (int64_t)blockIdx.x * SPAN; base < nnz; base += (int64_t)gridDim.x * SPAN) {
        const int64_t w0 = base + (int64_t)wave * 64 * kSpanRounds;
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t p = w0 + rr * 64 + lane;
            const bool valid = p < nnz;
            long long row = 0;
            ull wp = 0;
            if (valid) {
                const uint64_t x = words[p];
                const int64_t u = (int64_t)(x & wmask);
                const int32_t cc = (int32_t)((x >> wb) & cmask);
                col[p] = cc;
                put_w(cw, cw32, p, u);
                row = (long long)(x >> (64 - rb));
                wp = (ull)(u == 0 ? 1 : u);
                // rows must list their columns strictly increasing (err bit 4)
                if (p > 0) {
                    const uint64_t px = words[p - 1];
                    if ((px >> (64 - rb)) == (uint64_t)row && (int32_t)((px >> wb) & cmask) >= cc)
                        bad = 4;
                }
            }
            sc.add(valid, row, wp, (ull *)kdeg);
        }
        sc.flush((ull *)kdeg);
    }
    block_flag_or(bad, f);
}

void launch_csr_entries1(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                         int64_t e, int32_t n, int rb, int cb, uint64_t *words)
{
    if (e > 0)
        hipLaunchKernelGGL(k_csr_entries1, grid_for(e, 256, 1 << 20), 256, 0, s, src, dst, w, e, n,
                           rb, cb, words);
}

void launch_csr_unpack1(hipStream_t s, const uint64_t *words, int64_t nnz, int rb, int cb,
                        int32_t *col, int64_t *cw, uint32_t *cw32, int64_t *kdeg, DevFlags *f)
{
    if (nnz > 0)
        hipLaunchKernelGGL(k_csr_unpack1, grid_for(nnz, 256 * kSpanRounds, 8192), 256, 0, s, words,
                           nnz, rb, cb, col, cw, cw32, kdeg, f);
}

__global__ void k_csr_entries_packed(const int32_t *src, const int32_t *dst, WUnits w,
                                     int64_t e, int32_t n, int wbits, uint32_t *keys,
                                     uint64_t *payload)
{
    GRID_STRIDE(i, e)
    {
        int32_t a = src[i], b = dst[i];
        uint64_t u = (uint64_t)w[i];
        bool self = (a == b);
        keys[2 * i] = self ? (uint32_t)n : (uint32_t)a;
        keys[2 * i + 1] = self ? (uint32_t)n : (uint32_t)b;
        payload[2 * i] = ((uint64_t)(uint32_t)b << wbits) | u;
        payload[2 * i + 1] = ((uint64_t)(uint32_t)a << wbits) | u;
    }
}

__global__ __launch_bounds__(256) void k_csr_unpack(const uint32_t *skeys, const uint64_t *spay,
                                                   int64_t nnz, int wbits, int32_t *col,
                                                   int64_t *cw, uint32_t *cw32, int64_t *kdeg)
{
    constexpr int64_t SPAN = 256 * kSpanRounds;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    SegCarry sc;
    for (int64_t base = (int64_t)blockIdx.x * SPAN; base < nnz; base += (int64_t)gridDim.x * SPAN) {
        const int64_t w0 = base + (int64_t)wave * 64 * kSpanRounds;
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t p = w0 + rr * 64 + lane;
            const bool valid = p < nnz;
            long long row = 0;
            ull wp = 0;
            if (valid) {
                uint64_t x = spay[p];
                int64_t u = (int64_t)(x & ((1ull << wbits) - 1ull));
                col[p] = (int32_t)(x >> wbits);
                put_w(cw, cw32, p, u);
                row = (long long)skeys[p];
                wp = (ull)(u == 0 ? 1 : u);
            }
            sc.add(valid, row, wp, (ull *)kdeg);
        }
        sc.flush((ull *)kdeg);
    }
}

void launch_csr_entries_packed(hipStream_t s, const int32_t *src, const int32_t *dst,
                               WUnits w, int64_t e, int32_t n, int wbits, uint32_t *keys,
                               uint64_t *payload)
{
    if (e > 0)
        hipLaunchKernelGGL(k_csr_entries_packed, grid_for(e, 256, 1 << 20), 256, 0, s, src, dst, w,
                           e, n, wbits, keys, payload);
}

void launch_csr_unpack(hipStream_t s, const uint32_t *skeys, const uint64_t *spay, int64_t nnz,
                       int wbits, int32_t *col, int64_t *cw, uint32_t *cw32, int64_t *kdeg)
{
    if (nnz > 0)
        hipLaunchKernelGGL(k_csr_unpack, grid_for(nnz, 256 * kSpanRounds, 1 << 20), 256, 0, s, skeys, spay, nnz,
                           wbits, col, cw, cw32, kdeg);
}

__global__ void k_self_loops(const int32_t *src, const int32_t *dst, WUnits w, int64_t e,
                             int64_t *kdeg, int64_t *selfsum, ull *lastloop)
{
    GRID_STRIDE(i, e)
    {
        int32_t a = src[i];
        if (a != dst[i]) continue;
        int64_t u = w[i];
        ull wp = (ull)(u == 0 ? 1 : u);
        atomicAdd((ull *)&selfsum[a], wp);
        atomicAdd((ull *)&kdeg[a], 4ull * wp);
        atomicMax(&lastloop[a], (ull)(i + 1));
    }
}

// ---------------------------------------------------------------------------
// CSR build for an edge list sorted by source (every generator and most files):
// the entries (row = src) of non-self edges are already grouped by row, so only
// the reverse entries (row = dst) go through the row sort, and row v is its
// reverse entries (in edge order) followed by its forward ones.

// Every lane of the wavefront calls this: out[v] = val for v in the lane's
// range (lo, hi]; a range longer than a few entries is written by the whole
// wavefront.
__device__ __forceinline__ void wave_range_fill(long long lo, long long hi, long long val,
                                                int64_t *out)
{
    const int lane = threadIdx.x & 63;
    const bool wide = hi - lo > 8;
    if (!wide)
        for (long long v = lo + 1; v <= hi; ++v) out[v] = val;
    ull m = __ballot(wide);
    while (m) {
        const int l = __ffsll((unsigned long long)m) - 1;
        const long long a = __shfl(lo, l, 64), b = __shfl(hi, l, 64), x = __shfl(val, l, 64);
        for (long long v = a + 1 + lane; v <= b; v += 64) out[v] = x;
        m &= m - 1;
    }
}

// The two ends of "out[v] = first position r with min(keys[r], n) >= v" over
// ascending keys: 0 up to the first key, total past the last one.  (The
// ranges between keys are written where the keys change.)
template <typename K>
__global__ void k_range_ends(const K *keys, int64_t total, int32_t n, int kshift, int64_t *out)
{
    const int64_t k0 = min((int64_t)(keys[0] >> kshift), (int64_t)n);
    const int64_t k1 = min((int64_t)(keys[total - 1] >> kshift), (int64_t)n);
    GRID_STRIDE(v, (int64_t)n + 1)
    {
        if (v <= k0)
            out[v] = 0;
        else if (v > k1)
            out[v] = total;
    }
}

// Edges go through these kernels in tiles of kEdgeTile consecutive edges, one
// tile per 256-thread block iteration.  Instead of a per-edge scan of
// self-loop flags, the pass over the list counts self-loop edges per tile
// (tself) and per source vertex (vself); after exclusive scans of both,
// tscan[t] = self-loop edges before tile t and vscan[v] = self-loop edges
// with src < v (= before fs[v] in a list sorted by source).
static_assert(kEdgeTile == 256 * kSpanRounds, "an edge tile is one block iteration of the fills");

// Wave w of the block owns edges [tile + w*64*kSpanRounds, +64*kSpanRounds) of
// tile `t`, round rr holding lanes' edges tile + w*64*kSpanRounds + rr*64 +
// lane.  Given which of them are self-loops, returns the self-loop edges
// before the wave's span and sets m[rr] to the ballot of each round, so the
// self-loop edges before edge (rr, lane) are the returned count plus the
// bits of m[0..rr) and of m[rr] below the lane.  Every thread of the block
// calls it once per tile; a tile without self-loops costs two loads.
__device__ __forceinline__ int32_t tile_self_base(const int32_t *tscan, int64_t t,
                                                  const bool (&sl)[kSpanRounds],
                                                  ull (&m)[kSpanRounds], int32_t *s_wc)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t off = tscan[t];
    const bool any = tscan[t + 1] != off;  // the same for the whole block
    if (any) {
        int wc = 0;
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            m[rr] = __ballot(sl[rr]);
            wc += __popcll(m[rr]);
        }
        if (lane == 0) s_wc[wave] = wc;
        __syncthreads();
        for (int w = 0; w < wave; ++w) off += s_wc[w];
        __syncthreads();
    } else {
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) m[rr] = 0;
    }
    return off;
}

// One pass over the edge list: err bit 8 when src is not non-decreasing, bit
// 16 when the list is not canonical and strictly sorted (src <= dst and
// (src, dst) strictly increasing - then every row lists each neighbour once
// and in increasing order); a self-loop edge's weight goes to selfsum/kdeg of
// its vertex, its position to lastloop and its count to tself/vself; and, for
// a sorted list, fs[v] = the first edge with src >= v between the first and
// the last source (k_range_ends writes the two ends).
__global__ __launch_bounds__(256) void k_src_sorted(const int32_t *src, const int32_t *dst,
                                                   WUnits w, int64_t e, int32_t *tself,
                                                   int32_t *vself, int64_t *fs, int64_t *kdeg,
                                                   int64_t *selfsum, ull *lastloop, DevFlags *f)
{
    __shared__ int32_t s_wc[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int bad = 0;
    for (int64_t base = (int64_t)blockIdx.x * kEdgeTile; base < e;
         base += (int64_t)gridDim.x * kEdgeTile) {
        int wc = 0;
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t i = base + rr * 256 + threadIdx.x;
            long long lo = 0, hi = 0;
            bool sl = false;
            if (i < e) {
                const int32_t a = src[i], b = dst[i];
                const int32_t pa = i > 0 ? src[i - 1] : -1;
                if (a < pa) bad |= 8;
                if (a > b || (a == pa && b <= dst[i - 1])) bad |= 16;
                sl = a == b;
                if (sl) {
                    const int64_t u = w[i];
                    const ull wp = (ull)(u == 0 ? 1 : u);
                    atomicAdd((ull *)&selfsum[a], wp);
                    atomicAdd((ull *)&kdeg[a], 4ull * wp);
                    atomicMax(&lastloop[a], (ull)(i + 1));
                    atomicAdd(&vself[a], 1);
                }
                if (i > 0 && a > pa) {  // fs[v] = i for v in (pa, a]; in bounds for any order
                    lo = pa;
                    hi = a;
                }
            }
            wc += __popcll(__ballot(sl));
            wave_range_fill(lo, hi, (long long)i, fs);
        }
        if (lane == 0) s_wc[wave] = wc;
        __syncthreads();
        if (threadIdx.x == 0) tself[base / kEdgeTile] = s_wc[0] + s_wc[1] + s_wc[2] + s_wc[3];
        __syncthreads();
    }
    block_flag_or(bad, f);
}

// Reverse entries of non-self edges, compacted (edge i goes to i minus the
// self-loop edges before it): with ONE, one word per entry dst << (64 - kb) |
// src << wb | w, kb = bits of the largest id and wb = 64 - 2 kb (when every
// weight fits wb bits), so the row sort moves 8 bytes per entry instead of 12;
// otherwise key dst and payload src << wbits | w.
template <bool ONE>
__global__ __launch_bounds__(256) void k_rev_entries(const int32_t *src, const int32_t *dst,
                                                    WUnits w, const int32_t *tscan,
                                                    int64_t e, int bits, uint32_t *keys,
                                                    uint64_t *out)
{
    __shared__ int32_t s_wc[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const ull below = (1ull << lane) - 1ull;
    for (int64_t base = (int64_t)blockIdx.x * kEdgeTile; base < e;
         base += (int64_t)gridDim.x * kEdgeTile) {
        const int64_t w0 = base + (int64_t)wave * 64 * kSpanRounds;
        int32_t a[kSpanRounds], b[kSpanRounds];
        bool sl[kSpanRounds];
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t i = w0 + rr * 64 + lane;
            a[rr] = i < e ? src[i] : 0;
            b[rr] = i < e ? dst[i] : 0;
            sl[rr] = i < e && a[rr] == b[rr];
        }
        ull m[kSpanRounds];
        int32_t off = tile_self_base(tscan, base / kEdgeTile, sl, m, s_wc);
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t i = w0 + rr * 64 + lane;
            if (i < e && !sl[rr]) {
                const int64_t r = i - (off + __popcll(m[rr] & below));
                if constexpr (ONE) {
                    out[r] = ((uint64_t)(uint32_t)b[rr] << (64 - bits)) |
                             ((uint64_t)(uint32_t)a[rr] << (64 - 2 * bits)) | (uint64_t)w[i];
                } else {
                    keys[r] = (uint32_t)b[rr];
                    out[r] = ((uint64_t)(uint32_t)a[rr] << bits) | (uint64_t)w[i];
                }
            }
            off += __popcll(m[rr]);
        }
    }
}

// rowptr[v] = reverse entries of rows < v + forward (non-self) entries of rows
// < v; ffwd[v] = rowptr[v] + (reverse entries of v) - (non-self edges before
// fs[v]), so forward edge i of row v lands at ffwd[v] + i - (self-loop edges
// before i).
__global__ void k_rowptr_sum(const int64_t *rs, const int64_t *fs, const int32_t *vscan, int32_t n,
                             int64_t *rowptr, int64_t *ffwd)
{
    GRID_STRIDE(v, (int64_t)n + 1)
    {
        const int64_t nf = fs[v] - vscan[v];
        const int64_t rp = rs[v] + nf;
        rowptr[v] = rp;
        if (v < n) ffwd[v] = rp + (rs[v + 1] - rs[v]) - nf;
    }
}

// The fills sum k per row and, with CHECK, test that every row's columns
// strictly increase (err bit 4): within the reverse part, within the forward
// part (skipping self-loop edges), and from the last reverse column to the
// first forward one.  Without CHECK the list is canonical and strictly sorted.
// Reverse entries come as (skey[r], spay[r] = src << wbits | w), or with skey
// null as one word spay[r] = row << vshift | src << wbits | w; columns are
// (spay >> wbits) & cmask either way.
template <bool CHECK>
__global__ __launch_bounds__(256) void k_fill_rev(const uint32_t *skey, const uint64_t *spay,
                                                 int64_t nr, int wbits, int vshift, uint64_t cmask,
                                                 const int64_t *rs, const int64_t *rowptr,
                                                 int32_t *col, int64_t *cw, uint32_t *cw32, int pk,
                                                 int64_t *kdeg, DevFlags *f)
{
    int bad = 0;
    constexpr int64_t SPAN = 256 * kSpanRounds;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    SegCarry sc;
    for (int64_t base = (int64_t)blockIdx.x * SPAN; base < nr; base += (int64_t)gridDim.x * SPAN) {
        const int64_t w0 = base + (int64_t)wave * 64 * kSpanRounds;
        // all rounds' entries, then all their row offsets, then the stores: the
        // loads of a span are in flight together
        uint64_t xs[kSpanRounds];
        uint32_t vs[kSpanRounds];
        int64_t ro[kSpanRounds];
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t r = w0 + rr * 64 + lane;
            xs[rr] = r < nr ? spay[r] : 0;
            vs[rr] = r < nr ? (skey ? skey[r] : (uint32_t)(xs[rr] >> vshift)) : 0;
        }
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t r = w0 + rr * 64 + lane;
            ro[rr] = r < nr ? rowptr[vs[rr]] - rs[vs[rr]] : 0;
        }
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t r = w0 + rr * 64 + lane;
            const bool valid = r < nr;
            long long row = 0;
            ull wp = 0;
            if (valid) {
                const uint64_t x = xs[rr];
                const uint32_t v = vs[rr];
                const int64_t u = (int64_t)(x & ((1ull << wbits) - 1ull));
                const int64_t pos = ro[rr] + r;
                const int32_t cc = (int32_t)((x >> wbits) & cmask);
                put_entry(col, cw, cw32, pk, pos, cc, u);
                row = v;
                wp = (ull)(u == 0 ? 1 : u);
                if constexpr (CHECK) {
                    if (r > 0) {
                        const uint64_t px = spay[r - 1];
                        const uint32_t pv = skey ? skey[r - 1] : (uint32_t)(px >> vshift);
                        if (pv == v && (int32_t)((px >> wbits) & cmask) >= cc) bad = 4;
                    }
                }
            }
            sc.add(valid, row, wp, (ull *)kdeg);
        }
        sc.flush((ull *)kdeg);
    }
    if constexpr (CHECK) block_flag_or(bad, f);
}

template <bool CHECK>
__global__ __launch_bounds__(256) void k_fill_fwd(const int32_t *src, const int32_t *dst,
                                                 WUnits w, const int32_t *tscan,
                                                 const int32_t *vscan, int64_t e,
                                                 const int64_t *fs, const int64_t *rs,
                                                 const uint64_t *spay, int wbits, uint64_t cmask,
                                                 const int64_t *ffwd, int32_t *col, int64_t *cw, uint32_t *cw32,
                                                 int pk, int64_t *kdeg, DevFlags *f)
{
    __shared__ int32_t s_wc[4];
    int bad = 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const ull below = (1ull << lane) - 1ull;
    SegCarry sc;
    for (int64_t base = (int64_t)blockIdx.x * kEdgeTile; base < e;
         base += (int64_t)gridDim.x * kEdgeTile) {
        const int64_t w0 = base + (int64_t)wave * 64 * kSpanRounds;
        int32_t sv[kSpanRounds], sb[kSpanRounds];
        bool sl[kSpanRounds];
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t i = w0 + rr * 64 + lane;
            sv[rr] = i < e ? src[i] : 0;
            sb[rr] = i < e ? dst[i] : 0;
            sl[rr] = i < e && sv[rr] == sb[rr];
        }
        ull m[kSpanRounds];
        int32_t off = tile_self_base(tscan, base / kEdgeTile, sl, m, s_wc);
        // every round's row offset and weight in flight before the stores
        int64_t fo[kSpanRounds], uw[kSpanRounds];
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t i = w0 + rr * 64 + lane;
            const bool fwd = i < e && !sl[rr];
            fo[rr] = fwd ? ffwd[sv[rr]] : 0;
            uw[rr] = fwd ? w[i] : 0;
        }
#pragma unroll
        for (int rr = 0; rr < kSpanRounds; ++rr) {
            const int64_t i = w0 + rr * 64 + lane;
            const bool valid = i < e;
            long long row = 0;
            ull wp = 0;
            if (valid) {
                const int32_t v = sv[rr], b = sb[rr];
                row = v;  // a self-loop edge keeps its row's key with weight 0, so the
                          // runs of equal keys stay contiguous for the segmented sum
                if (v != b) {
                    const int64_t sli = off + __popcll(m[rr] & below);  // self-loops before i
                    const int64_t pos = fo[rr] + (i - sli);
                    const int64_t u = uw[rr];
                    put_entry(col, cw, cw32, pk, pos, b, u);
                    wp = (ull)(u == 0 ? 1 : u);
                    if constexpr (CHECK) {
                        const int64_t f0 = fs[v];
                        const int64_t rank = (i - f0) - (sli - vscan[v]);
                        if (rank > 0) {
                            int64_t k = i - 1;  // the previous non-self edge of the row
                            while (src[k] == dst[k]) --k;
                            if (dst[k] >= b) bad = 4;
                        } else if (rs[v + 1] > rs[v] &&
                                   (int32_t)((spay[rs[v + 1] - 1] >> wbits) & cmask) >= b) {
                            bad = 4;
                        }
                    }
                }
            }
            off += __popcll(m[rr]);
            sc.add(valid, row, wp, (ull *)kdeg);
        }
        sc.flush((ull *)kdeg);
    }
    if constexpr (CHECK) block_flag_or(bad, f);
}

void launch_src_sorted(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                       int64_t e, int32_t n, int32_t *tself, int32_t *vself, int64_t *fs,
                       int64_t *kdeg, int64_t *selfsum, ull *lastloop, DevFlags *f)
{
    fill_i32(s, vself, n, 0);
    if (e > 0) {
        hipLaunchKernelGGL(k_src_sorted, grid_for(e, kEdgeTile, 8192), 256, 0, s, src, dst, w, e,
                           tself, vself, fs, kdeg, selfsum, lastloop, f);
        hipLaunchKernelGGL(k_range_ends<int32_t>, grid_for((int64_t)n + 1, 256, 1 << 20), 256, 0,
                           s, src, e, n, 0, fs);
    } else {
        fill_i64(s, fs, (int64_t)n + 1, 0);
    }
}

void launch_rev_entries(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                        const int32_t *tscan, int64_t e, int wbits, uint32_t *keys,
                        uint64_t *payload)
{
    if (e > 0)
        hipLaunchKernelGGL(k_rev_entries<false>, grid_for(e, kEdgeTile, 1 << 20), 256, 0, s, src,
                           dst, w, tscan, e, wbits, keys, payload);
}

// Reverse entries as (skey, spay = src << wbits | w), or with skey null as one
// word spay = row << vshift | src << wbits | w; kb = bits of the largest id.
void launch_rowptr_sum(hipStream_t s, const int64_t *rs, const int64_t *fs, const int32_t *vscan,
                       int32_t n, int64_t *rowptr, int64_t *ffwd)
{
    hipLaunchKernelGGL(k_rowptr_sum, grid_for((int64_t)n + 1, 256, 1 << 20), 256, 0, s, rs, fs,
                       vscan, n, rowptr, ffwd);
}

// The reverse fill on srev and the forward fill on sfwd (after rowptr/ffwd of
// launch_rowptr_sum): they write disjoint entries and meet only in the atomic
// row sums of kdeg, so they may run concurrently.
void launch_csr_sorted_fill(hipStream_t srev, hipStream_t sfwd, bool check, const int32_t *src,
                            const int32_t *dst, WUnits w, const int32_t *tscan,
                            const int32_t *vscan, int64_t e, int32_t n, int wbits, int vshift,
                            int kb, const int64_t *fs, const uint32_t *skey, const uint64_t *spay,
                            int64_t nr, const int64_t *rs, const int64_t *rowptr,
                            const int64_t *ffwd, int32_t *col, int64_t *cw, uint32_t *cw32, int pk,
                            int64_t *kdeg, DevFlags *f)
{
    const uint64_t cmask = kb >= 64 ? ~0ull : ((1ull << kb) - 1ull);
    constexpr int SPAN = 256 * kSpanRounds;
    if (nr > 0) {
        if (check)
            hipLaunchKernelGGL(k_fill_rev<true>, grid_for(nr, SPAN, 8192), 256, 0, srev, skey, spay,
                               nr, wbits, vshift, cmask, rs, rowptr, col, cw, cw32, pk, kdeg, f);
        else
            hipLaunchKernelGGL(k_fill_rev<false>, grid_for(nr, SPAN, 1 << 20), 256, 0, srev, skey,
                               spay, nr, wbits, vshift, cmask, rs, rowptr, col, cw, cw32, pk, kdeg,
                               f);
    }
    if (e > 0) {
        if (check)
            hipLaunchKernelGGL(k_fill_fwd<true>, grid_for(e, kEdgeTile, 8192), 256, 0, sfwd, src,
                               dst, w, tscan, vscan, e, fs, rs, spay, wbits, cmask, ffwd, col, cw,
                               cw32, pk, kdeg, f);
        else
            hipLaunchKernelGGL(k_fill_fwd<false>, grid_for(e, kEdgeTile, 1 << 20), 256, 0, sfwd,
                               src, dst, w, tscan, vscan, e, fs, rs, spay, wbits, cmask, ffwd, col,
                               cw, cw32, pk, kdeg, f);
    }
}

void launch_rev_entries1(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                         const int32_t *tscan, int64_t e, int kb, uint64_t *words)
{
    if (e > 0)
        hipLaunchKernelGGL(k_rev_entries<true>, grid_for(e, kEdgeTile, 1 << 20), 256, 0, s, src,
                           dst, w, tscan, e, kb, nullptr, words);
}

template <typename K>
__global__ void k_row_starts(const K *keys, int64_t total, int32_t n, int kshift, int64_t *out);

void launch_row_starts64(hipStream_t s, const uint64_t *keys, int64_t total, int32_t n,
                         int kshift, int64_t *rowptr)
{
    if (total > 0) {
        hipLaunchKernelGGL(k_row_starts<uint64_t>, grid_for(total, 256, 1 << 20), 256, 0, s, keys,
                           total, n, kshift, rowptr);
        hipLaunchKernelGGL(k_range_ends<uint64_t>, grid_for((int64_t)n + 1, 256, 1 << 20), 256, 0,
                           s, keys, total, n, kshift, rowptr);
    } else {
        fill_i64(s, rowptr, (int64_t)n + 1, 0);
    }
}

// out[v] = the first position r with min(keys[r] >> kshift, n) >= v (v in
// [0, n]) of ascending keys, between the first and the last key: each change
// of key writes the rows it passes (k_range_ends writes the two ends).
template <typename K>
__global__ void k_row_starts(const K *keys, int64_t total, int32_t n, int kshift, int64_t *out)
{
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = base + threadIdx.x;
        long long lo = 0, hi = 0;
        if (r < total && r > 0) {
            lo = min((int64_t)(keys[r - 1] >> kshift), (int64_t)n);
            hi = min((int64_t)(keys[r] >> kshift), (int64_t)n);
        }
        wave_range_fill(lo, hi, (long long)r, out);
    }
}

__global__ __launch_bounds__(256) void k_csr_gather(const uint32_t *skeys, const uint32_t *scodes,
                                                   int64_t nnz, const int32_t *src,
                                                   const int32_t *dst, WUnits w,
                                                   int32_t *col, int64_t *cw, uint32_t *cw32, int64_t *kdeg)
{
    // four entries per thread, 256 apart, so each wave covers 64 consecutive
    // entries (rows sorted within the wave for the segmented atomics)
    for (int64_t base = (int64_t)blockIdx.x * 1024; base < nnz; base += (int64_t)gridDim.x * 1024) {
        int64_t ei[4];
        uint32_t code[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t p = base + r * 256 + threadIdx.x;
            code[r] = (p < nnz) ? scodes[p] : 0u;
            ei[r] = (int64_t)(code[r] >> 1);
        }
        int32_t cv[4];
        int64_t u[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t p = base + r * 256 + threadIdx.x;
            if (p < nnz) {
                cv[r] = (code[r] & 1u) ? src[ei[r]] : dst[ei[r]];
                u[r] = w[ei[r]];
            } else {
                cv[r] = 0;
                u[r] = 0;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t p = base + r * 256 + threadIdx.x;
            const bool valid = p < nnz;
            long long row = 0;
            if (valid) {
                col[p] = cv[r];
                put_w(cw, cw32, p, u[r]);
                row = (long long)skeys[p];
            }
            seg_atomic_add(valid, row, (ull)(u[r] == 0 ? 1 : u[r]), (ull *)kdeg);
        }
    }
}

__global__ void k_loop_gather(const ull *lastloop, WUnits w, int32_t n, int64_t *loopw)
{
    GRID_STRIDE(i, n)
    {
        ull j = lastloop[i];
        loopw[i] = j ? w[j - 1] : 0;
    }
}

void launch_csr_entries(hipStream_t s, const int32_t *src, const int32_t *dst, int64_t e,
                        int32_t n, uint32_t *keys, uint32_t *codes)
{
    if (e > 0)
        hipLaunchKernelGGL(k_csr_entries, grid_for(e, 256, 1 << 20), 256, 0, s, src, dst, e, n,
                           keys, codes);
}

void launch_self_loops(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                       int64_t e, int64_t *kdeg, int64_t *selfsum, ull *lastloop)
{
    if (e > 0)
        hipLaunchKernelGGL(k_self_loops, grid_for(e, 256, 1 << 20), 256, 0, s, src, dst, w, e, kdeg,
                           selfsum, lastloop);
}

void launch_rowptr_search(hipStream_t s, const uint32_t *skeys, int64_t total, int32_t n,
                          int64_t *rowptr)
{
    if (total > 0) {
        hipLaunchKernelGGL(k_row_starts<uint32_t>, grid_for(total, 256, 1 << 20), 256, 0, s, skeys,
                           total, n, 0, rowptr);
        hipLaunchKernelGGL(k_range_ends<uint32_t>, grid_for((int64_t)n + 1, 256, 1 << 20), 256, 0,
                           s, skeys, total, n, 0, rowptr);
    } else
        fill_i64(s, rowptr, (int64_t)n + 1, 0);
}

void launch_csr_gather(hipStream_t s, const uint32_t *skeys, const uint32_t *scodes, int64_t nnz,
                       const int32_t *src, const int32_t *dst, WUnits w, int32_t *col,
                       int64_t *cw, uint32_t *cw32, int64_t *kdeg)
{
    if (nnz > 0)
        hipLaunchKernelGGL(k_csr_gather, grid_for(nnz, 1024, 1 << 20), 256, 0, s, skeys, scodes,
                           nnz, src, dst, w, col, cw, cw32, kdeg);
}

void launch_loop_gather(hipStream_t s, const ull *lastloop, WUnits w, int32_t n,
                        int64_t *loopw)
{
    if (n > 0) hipLaunchKernelGGL(k_loop_gather, grid_for(n, 256), 256, 0, s, lastloop, w, n, loopw);
}

// ---------------------------------------------------------------------------
// Per-level init (initCUDA, LouvainCUDA/InitKernel.cu:79-126; communityInitKernel,
// OneLevelKernel.cu:80-91) and the degree buckets of the sweep.

__global__ __launch_bounds__(256) void k_init_vertex(
    int32_t n, const int64_t *rowptr, const int64_t *rowend, const int64_t *kdeg,
    const int64_t *loopw, const int64_t *selfsum, int64_t *deg, int64_t *in, int32_t *size,
    ull *ddeg, int32_t *dsize,
    int32_t *C, int32_t *next, uint32_t *tier_key, int32_t *verts, long long *tier_cnt,
    long long *tier_nnz, ull *ksum, uint32_t *deg32, int32_t hub_q)
{
    __shared__ long long s_cnt[kTiers + 1], s_nnz[kTiers + 1];
    __shared__ ull s_k[4];
    if (threadIdx.x <= kTiers) {
        s_cnt[threadIdx.x] = 0;
        s_nnz[threadIdx.x] = 0;
    }
    __syncthreads();
    ull kt = 0;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        const bool valid = i < n;
        int t = 0;
        int64_t d = 0;
        if (valid) {
            const int64_t k = kdeg[i];
            kt = sat_add(kt, (ull)k);
            deg[i] = k;
            deg32[i] = (uint32_t)k;  // used only when the level's total degree fits 32 bits
            in[i] = loopw[i];
            size[i] = 1;
            ddeg[i] = 0;
            dsize[i] = 0;
            C[i] = (int32_t)i;
            next[i] = (int32_t)i;
            d = rowend[i] - rowptr[i];
            t = tier_of(d, d + (selfsum[i] > 0 ? 1 : 0), hub_q);
            tier_key[i] = (uint32_t)t;
            verts[i] = (int32_t)i;
        }
        wave_bucket_add(valid, t, d, s_cnt, s_nnz);  // one LDS atomic pair per tier and wave
    }
    kt = wave_sum_sat(kt);
    if ((threadIdx.x & 63) == 0) s_k[threadIdx.x >> 6] = kt;
    __syncthreads();
    if (threadIdx.x <= kTiers) {
        if (s_cnt[threadIdx.x]) atomicAdd((ull *)&tier_cnt[threadIdx.x], (ull)s_cnt[threadIdx.x]);
        if (s_nnz[threadIdx.x]) atomicAdd((ull *)&tier_nnz[threadIdx.x], (ull)s_nnz[threadIdx.x]);
    }
    if (threadIdx.x == 0) {  // total degree (saturating at kUnitLimit), one atomic per block
        const ull t = sat_add(sat_add(s_k[0], s_k[1]), sat_add(s_k[2], s_k[3]));
        if (t) atomicAdd(ksum, t);
    }
}

void launch_init_vertex(hipStream_t s, int32_t n, const int64_t *rowptr, const int64_t *rowend,
                        const int64_t *kdeg, const int64_t *loopw, const int64_t *selfsum,
                        int64_t *deg, int64_t *in, int32_t *size, ull *ddeg, int32_t *dsize,
                        int32_t *C, int32_t *next, uint32_t *tier_key, int32_t *verts,
                        long long *tier_cnt, long long *tier_nnz, ull *ksum, uint32_t *deg32,
                        int32_t hub_q)
{
    if (n > 0)
        hipLaunchKernelGGL(k_init_vertex, grid_for(n, 256, 1024), 256, 0, s, n, rowptr, rowend, kdeg,
                           loopw, selfsum, deg, in, size, ddeg, dsize, C, next, tier_key, verts,
                           tier_cnt, tier_nnz, ksum, deg32, hub_q);
}

// ---------------------------------------------------------------------------
// Exact Q aggregates (modularity kernel + thrust::reduce,
// LouvainCUDA/OneLevelKernel.cu:427-451,875-880) and the Jacobi apply step
// (communitySetKernel/mergeSetKernel, OneLevelKernel.cu:93-136).

// Per-block partial of (sum_in, sum deg^2) written to part[blockIdx.x]; the
// partials are summed by k_qfinal (no atomics on shared accumulators).
__device__ __forceinline__ void block_qpartial(long long sin, ull lo, ull hi, QAcc *part)
{
    __shared__ long long s_in[4];
    __shared__ ull s_lo[4], s_hi[4];
    sin = wave_sum(sin);
    wave_sum_u128(lo, hi);
    if ((threadIdx.x & 63) == 0) {
        s_in[threadIdx.x >> 6] = sin;
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long ti = 0;
        ull tl = 0, th = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            ti += s_in[w];
            tl += s_lo[w];
            th += s_hi[w] + (tl < s_lo[w] ? 1ull : 0ull);
        }
        part[blockIdx.x].sum_in = ti;
        part[blockIdx.x].d2_lo = tl;
        part[blockIdx.x].d2_hi = th;
        part[blockIdx.x].pad = 0;
    }
}

__global__ __launch_bounds__(256) void k_qfinal(const QAcc *part, int nb, QAcc *acc)
{
    long long sin = 0;
    ull lo = 0, hi = 0;
    for (int b = threadIdx.x; b < nb; b += 256) {
        sin += part[b].sum_in;
        lo += part[b].d2_lo;
        hi += part[b].d2_hi + (lo < part[b].d2_lo ? 1ull : 0ull);
    }
    __shared__ QAcc one[1];
    block_qpartial(sin, lo, hi, one);
    __syncthreads();
    if (threadIdx.x == 0) *acc = one[0];
}

__global__ __launch_bounds__(256) void k_qagg(int32_t n, const int64_t *deg, const int64_t *in, QAcc *part)
{
    long long sin = 0;
    ull lo = 0, hi = 0;
    GRID_STRIDE(c, n)
    {
        sin += in[c];
        acc_sq(lo, hi, (ull)deg[c]);
    }
    block_qpartial(sin, lo, hi, part);
}

__global__ __launch_bounds__(256) void k_qagg_ull(int32_t n, const ull *dc, const ull *lc, QAcc *part)
{
    long long sin = 0;
    ull lo = 0, hi = 0;
    GRID_STRIDE(c, n)
    {
        sin += (long long)lc[c];
        acc_sq(lo, hi, dc[c]);
    }
    block_qpartial(sin, lo, hi, part);
}

__global__ __launch_bounds__(256) void k_apply(int32_t n, int64_t *deg, const int64_t *in,
                                              int32_t *size, ull *ddeg, int32_t *dsize,
                                              int size_shift, uint32_t *deg32, QAcc *part)
{
    long long sin = 0;
    ull lo = 0, hi = 0;
    GRID_STRIDE(c, n)
    {
        const int64_t dv = apply_deltas(c, deg, size, ddeg, dsize, size_shift, deg32);
        sin += in[c];
        acc_sq(lo, hi, (ull)dv);
    }
    block_qpartial(sin, lo, hi, part);
}

__global__ void k_lab(int32_t n, const int32_t *C, const int32_t *size, const int64_t *deg, Lab *lab,
                      Lab8 *lab8)
{
    GRID_STRIDE(v, n)
    {
        const int32_t c = C[v];
        const int64_t d = deg[c];
        Lab L;
        L.c = c;
        L.sz = size[c];
        L.deg = d;
        lab[v] = L;
        if (lab8) {
            Lab8 N;
            N.c = c;
            N.deg = (uint32_t)d;
            lab8[v] = N;
        }
    }
}

void launch_lab(hipStream_t s, int32_t n, const int32_t *C, const int32_t *size,
                const int64_t *deg, Lab *lab, Lab8 *lab8)
{
    if (n > 0)
        hipLaunchKernelGGL(k_lab, grid_for(n, 256, 1 << 20), 256, 0, s, n, C, size, deg, lab, lab8);
}

static int qblocks(int32_t n) { return grid_for(n, 256, kQParts); }

void launch_qagg(hipStream_t s, int32_t n, const int64_t *deg, const int64_t *in, QAcc *part,
                 QAcc *acc)
{
    int nb = qblocks(n);
    hipLaunchKernelGGL(k_qagg, nb, 256, 0, s, n, deg, in, part);
    hipLaunchKernelGGL(k_qfinal, 1, 256, 0, s, part, nb, acc);
}

void launch_qagg_ull(hipStream_t s, int32_t n, const ull *dc, const ull *lc, QAcc *part, QAcc *acc)
{
    int nb = qblocks(n);
    hipLaunchKernelGGL(k_qagg_ull, nb, 256, 0, s, n, dc, lc, part);
    hipLaunchKernelGGL(k_qfinal, 1, 256, 0, s, part, nb, acc);
}

void launch_apply(hipStream_t s, int32_t n, int64_t *deg, const int64_t *in, int32_t *size,
                  ull *ddeg, int32_t *dsize, int size_shift, uint32_t *deg32, QAcc *part,
                  QAcc *acc)
{
    int nb = qblocks(n);
    hipLaunchKernelGGL(k_apply, nb, 256, 0, s, n, deg, in, size, ddeg, dsize, size_shift, deg32,
                       part);
    hipLaunchKernelGGL(k_qfinal, 1, 256, 0, s, part, nb, acc);
}

// ---------------------------------------------------------------------------
// The local-moving sweep (neighbourCommunitySetKernel + sort/reduce_by_key +
// SetNewCommunityWarpShuffle + oneLevelKernel, LouvainCUDA/OneLevelKernel.cu:28-78,
// 248-422, and on the last pass mergeIsolatedNodeKernel :138-171).
//
// For each vertex v the distinct communities of its neighbours are gathered
// into a hash table keyed by community id (LDS for q <= 2048, a global table
// for hubs), summing the weights W_vc; a self-loop adds the own community
// with W = 0.  Every occupied slot counts towards n_v; slots with W != 0 give
// g = W - ((gamma*deg[c])*f), f = k_v/(2m).  The choice is the lowest c among
// the maximal g when that maximum is > 0, otherwise the own community - the
// outcome of the reference's ascending scan with its replace rule.

// Per-vertex parameters: 2m, resolution and pass mode (0 frozen, 1 sweep,
// 2 last pass) - of the single graph, or of the vertex's component in a set.
struct VParam {
    double m2, gamma;
    int mode;
};

template <bool B>
__device__ __forceinline__ VParam vparam(const SweepArgs &a, int32_t v)
{
    if constexpr (B) {
        const int32_t c = a.vcomp[v];
        return {a.cm2[c], a.cgam[c], a.cmode[c]};
    } else {
        return {a.m2, a.gamma, a.last ? 2 : 1};
    }
}

// Gathers of a community total and of a neighbour's label record, from the
// narrow arrays when the level has them (the choice is uniform per launch).
__device__ __forceinline__ int64_t deg_of(const SweepArgs &a, int32_t c)
{
    return a.deg32 ? (int64_t)a.deg32[c] : a.deg[c];
}

__device__ __forceinline__ Lab lab_of(const SweepArgs &a, int32_t j)
{
    if (a.lab8) {
        const Lab8 L = a.lab8[j];
        return Lab{L.c, 0, (int64_t)L.deg};
    }
    return a.lab[j];
}

// Weight of row entry p, streamed past the caches (read once per pass); W4:
// the level keeps 4-byte weights (w32).  The width is a template parameter of
// the row-walking kernels so their unrolled loads carry no branch.
template <bool W4>
__device__ __forceinline__ int64_t w_of(const SweepArgs &a, int64_t p)
{
    if constexpr (W4)
        return (int64_t)__builtin_nontemporal_load(&a.w32[p]);
    else
        return __builtin_nontemporal_load(&a.w[p]);
}

// Row entry p as (column j, weight w) for a level's entry layout WM: 0 col +
// 8-byte weights, 1 col + 4-byte weights, 2 one word col << cpk | weight (a
// level 0 whose ids and largest weight fit 32 bits); streamed past the caches.
template <int WM>
__device__ __forceinline__ void entry_at(const SweepArgs &a, int64_t p, int32_t &j, int64_t &w)
{
    if constexpr (WM == 2) {
        const uint32_t x = __builtin_nontemporal_load((const uint32_t *)&a.col[p]);
        j = (int32_t)(x >> a.cpk);
        w = (int64_t)(x & ((1u << a.cpk) - 1u));
    } else {
        j = __builtin_nontemporal_load(&a.col[p]);
        w = w_of<WM == 1>(a, p);
    }
}

template <int WM>
__device__ __forceinline__ int32_t col_at(const SweepArgs &a, int64_t p)
{
    if constexpr (WM == 2)
        return (int32_t)(__builtin_nontemporal_load((const uint32_t *)&a.col[p]) >> a.cpk);
    else
        return __builtin_nontemporal_load(&a.col[p]);
}

// The entry layout of a level (see entry_at).
__host__ __device__ __forceinline__ int entry_mode(const SweepArgs &a)
{
    return a.cpk ? 2 : (a.w32 ? 1 : 0);
}

// The vertex's own record, with the size of its community (the wide records
// are always written; narrow levels gather neighbours from lab8).
__device__ __forceinline__ Lab own_lab(const SweepArgs &a, int32_t v) { return a.lab[v]; }

__device__ __forceinline__ void consider(int32_t c, int64_t W, int64_t degc, double gamma,
                                         double f, int &cnt, double &bg, int32_t &bc, int64_t &bw)
{
    ++cnt;
    if (W != 0) {
        double g = (double)W - ((gamma * (double)degc) * f);
        if (g > 0.0 && better(g, c, bg, bc)) {
            bg = g;
            bc = c;
            bw = W;
        }
    }
}

// The candidates of an LDS table, slots s0, s0 + S, ... below cap, with the
// community totals gathered four at a time (the order of consider calls does
// not change the result).
template <int S, typename V>
__device__ __forceinline__ void consider_slots(const SweepArgs &a, const int32_t *key, const V *val,
                                               int s0, int cap, double gamma, double f, int &cnt,
                                               double &bg, int32_t &bc, int64_t &bw)
{
    for (int s = s0; s < cap; s += 4 * S) {
        int32_t c[4];
        int64_t d[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = (s + u * S < cap) ? key[s + u * S] : -1;
#pragma unroll
        for (int u = 0; u < 4; ++u) d[u] = (c[u] >= 0) ? deg_of(a, c[u]) : 0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (c[u] >= 0) consider(c[u], (int64_t)val[s + u * S], d[u], gamma, f, cnt, bg, bc, bw);
    }
}

// One butterfly step of the argmax (largest gain, lowest community id) and
// the candidate count.
template <int OFF>
__device__ __forceinline__ void reduce_step(int &cnt, double &bg, int32_t &bc, int64_t &bw)
{
    cnt += xor_lanes<OFF>(cnt);
    const double og = xor_lanes<OFF>(bg);
    const int32_t oc = xor_lanes<OFF>(bc);
    const int64_t ow = xor_lanes<OFF>(bw);
    if (better(og, oc, bg, bc)) {
        bg = og;
        bc = oc;
        bw = ow;
    }
}

// Every lane of each aligned group of W lanes ends with the group's count and
// best candidate (DPP / ds_swizzle / ds_bpermute butterflies, see xor_lanes).
template <int W>
__device__ __forceinline__ void group_reduce(int &cnt, double &bg, int32_t &bc, int64_t &bw)
{
    if constexpr (W > 32) reduce_step<32>(cnt, bg, bc, bw);
    if constexpr (W > 16) reduce_step<16>(cnt, bg, bc, bw);
    if constexpr (W > 8) reduce_step<8>(cnt, bg, bc, bw);
    if constexpr (W > 4) reduce_step<4>(cnt, bg, bc, bw);
    if constexpr (W > 2) reduce_step<2>(cnt, bg, bc, bw);
    if constexpr (W > 1) reduce_step<1>(cnt, bg, bc, bw);
}

// The block's candidate count and best candidate from its waves' results
// (each after group_reduce<64>), passed through `scratch`: LDS of a table that
// no wave reads any more, so the kernel needs no LDS of its own for them (a
// 32 KiB table then leaves room for five blocks per CU).  True on thread 0,
// which holds the result; the caller syncs before the table is reused.
struct WaveBest {
    double g;
    long long w;
    int cnt;
    int32_t c;
};

template <int NW>
__device__ __forceinline__ bool block_best(void *scratch, int &cnt, double &bg, int32_t &bc,
                                           int64_t &bw)
{
    WaveBest *r = reinterpret_cast<WaveBest *>(scratch);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();  // every wave is done with the table under the scratch
    if (lane == 0) r[wave] = WaveBest{bg, (long long)bw, cnt, bc};
    __syncthreads();
    if (threadIdx.x != 0) return false;
    cnt = r[0].cnt;
    bg = r[0].g;
    bc = r[0].c;
    bw = r[0].w;
    for (int w = 1; w < NW; ++w) {
        cnt += r[w].cnt;
        if (better(r[w].g, r[w].c, bg, bc)) {
            bg = r[w].g;
            bc = r[w].c;
            bw = r[w].w;
        }
    }
    return true;
}

// Guard, move and deltas of one vertex (oneLevelKernel + remove/insert,
// OneLevelKernel.cu:52-78,400-421); on the last pass the deltas are the
// merge-isolated ones only (:109-171,921-927).  The deltas of an ordinary
// pass are formed afterwards from C and next by k_move_deltas; the last pass
// adds its (rare) merge-isolated deltas here.
__device__ __forceinline__ void finalize_vertex(const SweepArgs &a, int32_t v, int32_t own,
                                                int32_t so, int n, int32_t bc, int64_t bw, int mode)
{
    if (a.pcnt) a.pcnt[v] = n;
    if (mode == 0) {  // frozen component: keeps its communities
        a.next[v] = own;
        return;
    }
    const int32_t c0 = bc != INT_MAX ? bc : own;
    int32_t choice = c0;
    if (so == 1 && a.size[c0] == 1 && c0 > own) choice = own;
    if (mode == 1) {
        a.next[v] = choice;
        return;
    }
    const int32_t to = (n == 1 && so == 1 && c0 != own && own < c0) ? c0 : own;
    a.next[v] = to != own ? to : choice;
    if (to == own) return;
    const int64_t k = a.kdeg[v];
    if (a.size_shift) {  // one atomic per side: delta_deg * 2^shift + delta_size
        const ull pk = ((ull)k << a.size_shift) + 1ull;
        atomicAdd(&a.ddeg[own], 0ull - pk);
        atomicAdd(&a.ddeg[to], pk);
    } else {
        atomicAdd(&a.ddeg[own], (ull)(-k));
        atomicAdd(&a.ddeg[to], (ull)k);
        atomicAdd(&a.dsize[own], -1);
        atomicAdd(&a.dsize[to], 1);
    }
}

// Move deltas of an ordinary pass: every vertex whose next community differs
// from its current one moves k and one member from C[v] to next[v].  A block
// combines the deltas of its kMoveBlock vertices by community in LDS before
// they reach the global words, so a community joined by 10^5 movers in one
// pass takes one atomic per block instead of one per mover (same-address
// device atomics serialise; a scattered stream of them runs about 40x faster
// than the same number on one address).  A community that finds no slot
// within kDeltaProbes probes adds directly.  The community left is a singleton
// on the first pass of a level (own = v) and goes straight to global memory.
constexpr int kMoveBlock = 2048;
constexpr int kMoveLog2 = 12;
constexpr int kDeltaProbes = 8;

__device__ __forceinline__ void move_add(const SweepArgs &a, int32_t c, int64_t dk, int32_t ds)
{
    if (a.size_shift) {
        atomicAdd(&a.ddeg[c], ((ull)dk << a.size_shift) + (ull)(int64_t)ds);
    } else {
        atomicAdd(&a.ddeg[c], (ull)dk);
        atomicAdd(&a.dsize[c], ds);
    }
}

__device__ __forceinline__ void move_combine(const SweepArgs &a, int32_t *key, ull *dk, int32_t *ds,
                                             int32_t c, int64_t k, int32_t s)
{
    constexpr uint32_t mask = (1u << kMoveLog2) - 1u;
    uint32_t sl = hslot(c, kMoveLog2);
    for (int p = 0; p < kDeltaProbes; ++p) {
        int32_t cur = ((volatile int32_t *)key)[sl];
        if (cur == -1) {
            cur = atomicCAS(&key[sl], -1, c);
            if (cur == -1) cur = c;
        }
        if (cur == c) {
            atomicAdd(&dk[sl], (ull)k);
            atomicAdd(&ds[sl], s);
            return;
        }
        sl = (sl + 1) & mask;
    }
    move_add(a, c, k, s);
}

template <bool B>
__global__ __launch_bounds__(256) void k_move_deltas(SweepArgs a, int32_t n, int own_direct)
{
    constexpr int CAP = 1 << kMoveLog2;
    __shared__ int32_t s_key[CAP];
    __shared__ ull s_dk[CAP];
    __shared__ int32_t s_ds[CAP];
    for (int64_t base = (int64_t)blockIdx.x * kMoveBlock; base < n;
         base += (int64_t)gridDim.x * kMoveBlock) {
        for (int s = threadIdx.x; s < CAP; s += 256) {
            s_key[s] = -1;
            s_dk[s] = 0;
            s_ds[s] = 0;
        }
        __syncthreads();
        for (int u = 0; u < kMoveBlock / 256; ++u) {
            const int64_t v = base + u * 256 + threadIdx.x;
            if (v >= n) break;
            const int32_t own = a.C[v], to = a.next[v];
            if (to == own) continue;
            if constexpr (B) {
                if (a.cmode[a.vcomp[v]] != 1) continue;  // frozen, or last pass (done inline)
            }
            const int64_t k = a.kdeg[v];
            if (own_direct)
                move_add(a, own, -k, -1);
            else
                move_combine(a, s_key, s_dk, s_ds, own, -k, -1);
            move_combine(a, s_key, s_dk, s_ds, to, k, 1);
        }
        __syncthreads();
        for (int s = threadIdx.x; s < CAP; s += 256) {
            const int32_t c = s_key[s];
            if (c >= 0 && (s_dk[s] != 0 || s_ds[s] != 0)) move_add(a, c, (int64_t)s_dk[s], s_ds[s]);
        }
        __syncthreads();
    }
}

void launch_move_deltas(hipStream_t s, const SweepArgs &a, int32_t n, bool own_direct, bool batch)
{
    if (n <= 0) return;
    const int g = grid_for(n, kMoveBlock, 1 << 20);
    if (batch)
        hipLaunchKernelGGL(k_move_deltas<true>, g, 256, 0, s, a, n, own_direct ? 1 : 0);
    else
        hipLaunchKernelGGL(k_move_deltas<false>, g, 256, 0, s, a, n, own_direct ? 1 : 0);
}

// Inserts row entries [beg, end) with stride S starting at beg+lane: four
// independent row loads, then four independent gathers, then the inserts.
// With DL the gather is the neighbour's 16-byte label record and the
// community total goes into the table (dg); otherwise only C[target] is
// gathered and the totals are read per candidate later.
template <int S, bool DL, int WM, typename V, typename D>
__device__ __forceinline__ void insert_row(const SweepArgs &a, int32_t *key, V *val, D *dg,
                                           int l2, int64_t beg, int64_t end, int lane)
{
    for (int64_t p0 = beg + lane; p0 < end; p0 += 4 * S) {
        int32_t j[4];
        int64_t w[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t p = p0 + (int64_t)u * S;
            j[u] = -1;
            w[u] = 0;
            if (p < end) entry_at<WM>(a, p, j[u], w[u]);
        }
        if constexpr (DL) {
            Lab L[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j[u] >= 0) L[u] = lab_of(a, j[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j[u] >= 0) lds_insert_d(key, val, dg, l2, L[u].c, w[u], L[u].deg);
        } else {
            int32_t c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j[u] >= 0) c[u] = a.C[j[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j[u] >= 0) lds_insert(key, val, l2, c[u], w[u]);
        }
    }
}

// G lanes per vertex, 256/G vertices per block, an LDS table of up to
// 2^LOG2CAP slots per vertex (2^MINLOG2 .. 2^LOG2CAP sized to the row).
template <int G, int LOG2CAP, int MINLOG2, bool DL, bool B, int WM, bool N32>
__global__ __launch_bounds__(256) void k_sweep_group(SweepArgs a, const int32_t *__restrict__ verts,
                                                    int32_t nv)
{
    using V = TabV<N32>;
    using D = TabD<N32>;
    constexpr int CAP = 1 << LOG2CAP;
    constexpr int GPB = 256 / G;
    __shared__ int32_t s_key[GPB * CAP];
    __shared__ V s_val[GPB * CAP];
    __shared__ D s_dg[DL ? GPB * CAP : 1];
    const int lg = threadIdx.x % G, grp = threadIdx.x / G;
    int32_t *key = s_key + grp * CAP;
    V *val = s_val + grp * CAP;
    D *dg = s_dg + (DL ? grp * CAP : 0);
    for (int64_t base = (int64_t)blockIdx.x * GPB; base < nv; base += (int64_t)gridDim.x * GPB) {
        const int64_t vi = base + grp;
        const bool active = vi < nv;
        int32_t v = 0;
        Lab me{0, 0, 0};
        VParam vp{1.0, 1.0, 0};
        int64_t beg = 0, end = 0;
        bool loop = false;
        int l2 = MINLOG2;
        if (active) {
            v = verts[vi];
            vp = vparam<B>(a, v);
            me = own_lab(a, v);
            beg = a.rowptr[v];
            end = a.rowend[v];
            loop = a.selfsum[v] > 0;
            if (vp.mode == 0) end = beg, loop = false;  // frozen: nothing to gather
            if (MINLOG2 < LOG2CAP) l2 = table_log2(end - beg + (loop ? 1 : 0), MINLOG2);
        }
        const int cap = 1 << l2;
        for (int s = lg; s < cap; s += G) {
            key[s] = -1;
            val[s] = 0;
        }
        __syncthreads();
        if (active) {
            insert_row<G, DL, WM>(a, key, val, dg, l2, beg, end, lg);
            if (lg == 0 && loop) {
                if constexpr (DL)
                    lds_insert_d(key, val, dg, l2, me.c, 0, me.deg);
                else
                    lds_insert(key, val, l2, me.c, 0);
            }
        }
        __syncthreads();
        int cnt = 0;
        double bg = 0.0;
        int32_t bc = INT_MAX;
        int64_t bw = 0;
        if (active) {
            const double f = (double)a.kdeg[v] / vp.m2;
            if constexpr (DL) {
                for (int s = lg; s < cap; s += G) {
                    const int32_t c = key[s];
                    if (c >= 0)
                        consider(c, (int64_t)val[s], (int64_t)dg[s], vp.gamma, f, cnt, bg, bc, bw);
                }
            } else {
                consider_slots<G>(a, key, val, lg, cap, vp.gamma, f, cnt, bg, bc, bw);
            }
        }
        group_reduce<G>(cnt, bg, bc, bw);
        if (active && lg == 0) finalize_vertex(a, v, me.c, me.sz, cnt, bc, bw, vp.mode);
        __syncthreads();
    }
}

// One vertex per block of BLOCK threads with an LDS table of up to
// 2^LOG2CAP slots sized to the row.
template <int BLOCK, int LOG2CAP, int MINLOG2, bool B, int WM, bool N32>
__global__ __launch_bounds__(BLOCK) void k_sweep_block(SweepArgs a, const int32_t *__restrict__ verts,
                                                      int32_t nv)
{
    using V = TabV<N32>;
    constexpr int CAP = 1 << LOG2CAP;
    constexpr int NW = BLOCK / 64;
    __shared__ __align__(16) int32_t s_key[CAP];
    __shared__ V s_val[CAP];
    const int tid = threadIdx.x;
    for (int64_t vi = blockIdx.x; vi < nv; vi += gridDim.x) {
        const int32_t v = verts[vi];
        const VParam vp = vparam<B>(a, v);
        const Lab me = own_lab(a, v);
        const int64_t beg = a.rowptr[v];
        const int64_t end = vp.mode == 0 ? beg : a.rowend[v];  // frozen: nothing to gather
        const bool loop = vp.mode != 0 && a.selfsum[v] > 0;
        const int l2 = table_log2(end - beg + (loop ? 1 : 0), MINLOG2);
        const int cap = 1 << l2;
        for (int s = tid; s < cap; s += BLOCK) {
            s_key[s] = -1;
            s_val[s] = 0;
        }
        __syncthreads();
        insert_row<BLOCK, false, WM>(a, s_key, s_val, (TabD<N32> *)nullptr, l2, beg, end, tid);
        if (tid == 0 && loop) lds_insert(s_key, s_val, l2, me.c, 0);
        __syncthreads();
        int cnt = 0;
        double bg = 0.0;
        int32_t bc = INT_MAX;
        int64_t bw = 0;
        const double f = (double)a.kdeg[v] / vp.m2;
        consider_slots<BLOCK>(a, s_key, s_val, tid, cap, vp.gamma, f, cnt, bg, bc, bw);
        group_reduce<64>(cnt, bg, bc, bw);
        if (block_best<NW>(s_key, cnt, bg, bc, bw))
            finalize_vertex(a, v, me.c, me.sz, cnt, bc, bw, vp.mode);
        __syncthreads();
    }
}

template <typename KernelT>
static void launch_grid(KernelT k, int64_t blocks, int block, hipStream_t s, const SweepArgs &a,
                        const int32_t *verts, int32_t nv)
{
    int g = (int)(blocks > (1 << 20) ? (1 << 20) : (blocks < 1 ? 1 : blocks));
    hipLaunchKernelGGL(k, g, block, 0, s, a, verts, nv);
}

template <bool B, int WM, bool N32>
static void launch_sweep_tier_t(hipStream_t s, int tier, const SweepArgs &a, const int32_t *verts,
                                int32_t nv)
{
    switch (tier) {
    case 1: launch_grid(k_sweep_group<4, 3, 3, true, B, WM, N32>, (nv + 63) / 64, 256, s, a, verts, nv); break;
    case 2: launch_grid(k_sweep_group<8, 4, 4, true, B, WM, N32>, (nv + 31) / 32, 256, s, a, verts, nv); break;
    case 3: launch_grid(k_sweep_group<16, 5, 5, true, B, WM, N32>, (nv + 15) / 16, 256, s, a, verts, nv); break;
    case 4: launch_grid(k_sweep_group<32, 6, 6, true, B, WM, N32>, (nv + 7) / 8, 256, s, a, verts, nv); break;
    case 5: launch_grid(k_sweep_group<64, 7, 7, true, B, WM, N32>, (nv + 3) / 4, 256, s, a, verts, nv); break;
    case 6: launch_grid(k_sweep_group<64, 8, 8, true, B, WM, N32>, (nv + 3) / 4, 256, s, a, verts, nv); break;
    case 7: launch_grid(k_sweep_group<64, 9, 9, true, B, WM, N32>, (nv + 3) / 4, 256, s, a, verts, nv); break;
    case 8: launch_grid(k_sweep_group<64, 10, 10, false, B, WM, N32>, (nv + 3) / 4, 256, s, a, verts, nv); break;
    case 9: launch_grid(k_sweep_block<256, 11, 11, B, WM, N32>, nv, 256, s, a, verts, nv); break;
    case 10: launch_grid(k_sweep_block<256, 12, 11, B, WM, N32>, nv, 256, s, a, verts, nv); break;
    case 11: launch_grid(k_sweep_block<512, 13, 12, B, WM, N32>, nv, 512, s, a, verts, nv); break;
    default: break;
    }
}

// Tiers 1..11 (tier 12, the hubs, runs through run_hub_sweep).
void launch_sweep_tier(hipStream_t s, int tier, const SweepArgs &a, const int32_t *verts, int32_t nv,
                       bool batch)
{
    if (nv <= 0) return;
    // 4-byte tables where the level is narrow and keeps 4-byte (or packed) weights
    const int wm = entry_mode(a);
    const bool n32 = wm != 0 && a.lab8 != nullptr;
    if (batch) {  // (graph sets keep separate columns and weights)
        if (n32)
            launch_sweep_tier_t<true, 1, true>(s, tier, a, verts, nv);
        else if (wm == 1)
            launch_sweep_tier_t<true, 1, false>(s, tier, a, verts, nv);
        else
            launch_sweep_tier_t<true, 0, false>(s, tier, a, verts, nv);
    } else {
        if (wm == 2 && n32)
            launch_sweep_tier_t<false, 2, true>(s, tier, a, verts, nv);
        else if (wm == 2)
            launch_sweep_tier_t<false, 2, false>(s, tier, a, verts, nv);
        else if (n32)
            launch_sweep_tier_t<false, 1, true>(s, tier, a, verts, nv);
        else if (wm == 1)
            launch_sweep_tier_t<false, 1, false>(s, tier, a, verts, nv);
        else
            launch_sweep_tier_t<false, 0, false>(s, tier, a, verts, nv);
    }
}

// ---------------------------------------------------------------------------
// First pass of a level.  After init every community is one vertex (C = id,
// size 1, deg = k), so on rows that list each neighbour once a neighbour j is
// its own candidate with W = w(v, j): the argmax needs no aggregation table,
// and the only gather per entry is deg[j].  n_v = row length + has_loop (the
// own community enters through a self-loop with W = 0).

template <int G, bool B, int WM>
__global__ __launch_bounds__(256) void k_sweep1_group(SweepArgs a, const int32_t *__restrict__ verts,
                                                     int32_t nv)
{
    constexpr int GPB = 256 / G;
    const int lg = threadIdx.x % G, grp = threadIdx.x / G;
    for (int64_t base = (int64_t)blockIdx.x * GPB; base < nv; base += (int64_t)gridDim.x * GPB) {
        const int64_t vi = base + grp;
        const bool active = vi < nv;
        int cnt = 0;
        double bg = 0.0;
        int32_t bc = INT_MAX;
        int64_t bw = 0;
        int32_t v = 0;
        VParam vp{1.0, 1.0, 0};
        if (active) {
            v = verts[vi];
            vp = vparam<B>(a, v);
            const int64_t beg = a.rowptr[v];
            const int64_t end = vp.mode == 0 ? beg : a.rowend[v];
            const double f = (double)a.kdeg[v] / vp.m2;
            for (int64_t p0 = beg + lg; p0 < end; p0 += 4 * G) {
                int32_t j[4];
                int64_t w[4], d[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t p = p0 + (int64_t)u * G;
                    j[u] = -1;
                    w[u] = 0;
                    if (p < end) entry_at<WM>(a, p, j[u], w[u]);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) d[u] = (j[u] >= 0) ? deg_of(a, j[u]) : 0;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (j[u] >= 0) consider(j[u], w[u], d[u], vp.gamma, f, cnt, bg, bc, bw);
            }
        }
        group_reduce<G>(cnt, bg, bc, bw);
        if (active && lg == 0) {
            if (vp.mode != 0 && a.selfsum[v] > 0) ++cnt;
            finalize_vertex(a, v, v, 1, cnt, bc, bw, vp.mode);
        }
    }
}

template <int BLOCK, bool B, int WM>
__global__ __launch_bounds__(BLOCK) void k_sweep1_block(SweepArgs a, const int32_t *__restrict__ verts,
                                                       int32_t nv)
{
    constexpr int NW = BLOCK / 64;
    __shared__ int r_cnt[NW];
    __shared__ double r_g[NW];
    __shared__ int32_t r_c[NW];
    __shared__ long long r_w[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int64_t vi = blockIdx.x; vi < nv; vi += gridDim.x) {
        const int32_t v = verts[vi];
        const VParam vp = vparam<B>(a, v);
        const int64_t beg = a.rowptr[v];
        const int64_t end = vp.mode == 0 ? beg : a.rowend[v];
        const double f = (double)a.kdeg[v] / vp.m2;
        int cnt = 0;
        double bg = 0.0;
        int32_t bc = INT_MAX;
        int64_t bw = 0;
        for (int64_t p0 = beg + tid; p0 < end; p0 += 4 * BLOCK) {
            int32_t j[4];
            int64_t w[4], d[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t p = p0 + (int64_t)u * BLOCK;
                j[u] = -1;
                w[u] = 0;
                if (p < end) entry_at<WM>(a, p, j[u], w[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) d[u] = (j[u] >= 0) ? deg_of(a, j[u]) : 0;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (j[u] >= 0) consider(j[u], w[u], d[u], vp.gamma, f, cnt, bg, bc, bw);
        }
        group_reduce<64>(cnt, bg, bc, bw);
        if (lane == 0) {
            r_cnt[wave] = cnt;
            r_g[wave] = bg;
            r_c[wave] = bc;
            r_w[wave] = bw;
        }
        __syncthreads();
        if (tid == 0) {
            int tc = r_cnt[0];
            double g0 = r_g[0];
            int32_t c0 = r_c[0];
            int64_t w0 = r_w[0];
            for (int w = 1; w < NW; ++w) {
                tc += r_cnt[w];
                if (better(r_g[w], r_c[w], g0, c0)) {
                    g0 = r_g[w];
                    c0 = r_c[w];
                    w0 = r_w[w];
                }
            }
            if (vp.mode != 0 && a.selfsum[v] > 0) ++tc;
            finalize_vertex(a, v, v, 1, tc, c0, w0, vp.mode);
        }
        __syncthreads();
    }
}

template <bool B, int WM>
static void launch_sweep1_tier_t(hipStream_t s, int tier, const SweepArgs &a, const int32_t *verts,
                                 int32_t nv)
{
    switch (tier) {
    case 1: launch_grid(k_sweep1_group<4, B, WM>, (nv + 63) / 64, 256, s, a, verts, nv); break;
    case 2: launch_grid(k_sweep1_group<8, B, WM>, (nv + 31) / 32, 256, s, a, verts, nv); break;
    case 3: launch_grid(k_sweep1_group<16, B, WM>, (nv + 15) / 16, 256, s, a, verts, nv); break;
    case 4: launch_grid(k_sweep1_group<32, B, WM>, (nv + 7) / 8, 256, s, a, verts, nv); break;
    case 5:
    case 6:
    case 7:
    case 8: launch_grid(k_sweep1_group<64, B, WM>, (nv + 3) / 4, 256, s, a, verts, nv); break;
    case 9:
    case 10: launch_grid(k_sweep1_block<256, B, WM>, nv, 256, s, a, verts, nv); break;
    case 11: launch_grid(k_sweep1_block<512, B, WM>, nv, 512, s, a, verts, nv); break;
    default: break;
    }
}

void launch_sweep1_tier(hipStream_t s, int tier, const SweepArgs &a, const int32_t *verts,
                        int32_t nv, bool batch)
{
    if (nv <= 0) return;
    const int wm = entry_mode(a);
    if (batch)
        wm == 1 ? launch_sweep1_tier_t<true, 1>(s, tier, a, verts, nv)
                : launch_sweep1_tier_t<true, 0>(s, tier, a, verts, nv);
    else if (wm == 2)
        launch_sweep1_tier_t<false, 2>(s, tier, a, verts, nv);
    else
        wm == 1 ? launch_sweep1_tier_t<false, 1>(s, tier, a, verts, nv)
                : launch_sweep1_tier_t<false, 0>(s, tier, a, verts, nv);
}

// Hubs: one block per chunk of a hub's entries (entry t == row length is the
// self-loop's own community, W = 0), then one wavefront per hub over its chunks.
template <bool B, int WM>
__global__ __launch_bounds__(256) void k_hub1_chunk(SweepArgs a, HubArgs h)
{
    __shared__ int r_cnt[4];
    __shared__ double r_g[4];
    __shared__ int32_t r_c[4];
    __shared__ long long r_w[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t ci = blockIdx.x;
    const int32_t j = h.chunk_hub[ci];
    const int32_t v = h.hv[j];
    const int64_t q = h.hq[j];
    const int64_t rb = a.rowptr[v], d = a.rowend[v] - rb;
    const VParam vp = vparam<B>(a, v);
    const double f = (double)a.kdeg[v] / vp.m2;
    const int64_t cs = chunk_span(h.hP[j], h.chunk_mul);
    const int64_t t0 = (int64_t)h.chunk_k[ci] * cs;
    const int64_t t1 = vp.mode == 0 ? t0 : min(q, t0 + cs);
    int cnt = 0;
    double bg = 0.0;
    int32_t bc = INT_MAX;
    int64_t bw = 0;
    for (int64_t p0 = t0 + tid; p0 < t1; p0 += 4 * 256) {
        int32_t jj[4];
        int64_t w[4], dg[4];
        bool own[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t t = p0 + (int64_t)u * 256;
            own[u] = (t == d);
            jj[u] = -1;
            w[u] = 0;
            if (t < t1 && t < d) entry_at<WM>(a, rb + t, jj[u], w[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) dg[u] = (jj[u] >= 0) ? deg_of(a, jj[u]) : 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (jj[u] >= 0)
                consider(jj[u], w[u], dg[u], vp.gamma, f, cnt, bg, bc, bw);
            else if (own[u] && p0 + (int64_t)u * 256 < t1)
                ++cnt;
        }
    }
    group_reduce<64>(cnt, bg, bc, bw);
    if (lane == 0) {
        r_cnt[wave] = cnt;
        r_g[wave] = bg;
        r_c[wave] = bc;
        r_w[wave] = bw;
    }
    __syncthreads();
    if (tid == 0) {
        int tc = r_cnt[0];
        double g0 = r_g[0];
        int32_t c0 = r_c[0];
        int64_t w0 = r_w[0];
        for (int w = 1; w < 4; ++w) {
            tc += r_cnt[w];
            if (better(r_g[w], r_c[w], g0, c0)) {
                g0 = r_g[w];
                c0 = r_c[w];
                w0 = r_w[w];
            }
        }
        h.cn[ci] = tc;
        h.cg[ci] = g0;
        h.cc[ci] = c0;
        h.cw[ci] = w0;
    }
}

template <bool B>
__global__ __launch_bounds__(256) void k_hub1_final(SweepArgs a, HubArgs h, int32_t nh)
{
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    int cnt = 0;
    double bg = 0.0;
    int32_t bc = INT_MAX;
    int64_t bw = 0;
    if (j < nh) {
        for (int64_t k = h.cbase[j] + lane; k < h.cbase[j + 1]; k += 64) {
            cnt += h.cn[k];
            if (better(h.cg[k], h.cc[k], bg, bc)) {
                bg = h.cg[k];
                bc = h.cc[k];
                bw = h.cw[k];
            }
        }
    }
    group_reduce<64>(cnt, bg, bc, bw);
    if (j < nh && lane == 0) {
        const int32_t v = h.hv[j];
        finalize_vertex(a, v, v, 1, cnt, bc, bw, vparam<B>(a, v).mode);
    }
}

int run_hub_sweep1(hipStream_t s, const SweepArgs &a, const HubArgs &h, int32_t nh, int64_t nchunks,
                   bool batch)
{
    if (nh <= 0) return LV_OK;
    const int wm = entry_mode(a);
    if (batch) {
        if (wm == 1)
            hipLaunchKernelGGL((k_hub1_chunk<true, 1>), (int)nchunks, 256, 0, s, a, h);
        else
            hipLaunchKernelGGL((k_hub1_chunk<true, 0>), (int)nchunks, 256, 0, s, a, h);
        hipLaunchKernelGGL(k_hub1_final<true>, (int)((nh + 3) / 4), 256, 0, s, a, h, nh);
    } else {
        if (wm == 2)
            hipLaunchKernelGGL((k_hub1_chunk<false, 2>), (int)nchunks, 256, 0, s, a, h);
        else if (wm == 1)
            hipLaunchKernelGGL((k_hub1_chunk<false, 1>), (int)nchunks, 256, 0, s, a, h);
        else
            hipLaunchKernelGGL((k_hub1_chunk<false, 0>), (int)nchunks, 256, 0, s, a, h);
        hipLaunchKernelGGL(k_hub1_final<false>, (int)((nh + 3) / 4), 256, 0, s, a, h, nh);
    }
    LV_HIP(hipGetLastError());
    return LV_OK;
}

__global__ void k_rows_check(const uint32_t *skeys, const int32_t *col, int64_t nnz, DevFlags *f)
{
    int bad = 0;
    GRID_STRIDE(p, nnz)
    {
        if (p > 0 && skeys[p] == skeys[p - 1] && col[p] <= col[p - 1]) bad = 4;
    }
    block_flag_or(bad, f);
}

void launch_rows_check(hipStream_t s, const uint32_t *skeys, const int32_t *col, int64_t nnz,
                       DevFlags *f)
{
    if (nnz > 1)
        hipLaunchKernelGGL(k_rows_check, grid_for(nnz, 256, 8192), 256, 0, s, skeys, col, nnz, f);
}

// ---------------------------------------------------------------------------
// Hub tier (q > 2048): partition + LDS aggregation instead of one global hash
// table per hub, so the work is streaming reads/writes plus LDS atomics.

__global__ void k_hub_setup(const int32_t *hv, int32_t nh, const int64_t *rowptr,
                            const int64_t *rowend, const int64_t *selfsum, int32_t target,
                            int chunk_mul, int64_t *hq,
                            int32_t *hP, int64_t *hP64, int64_t *hch, int64_t *hpc)
{
    GRID_STRIDE(j, nh)
    {
        int32_t v = hv[j];
        int64_t q = rowend[v] - rowptr[v] + (selfsum[v] > 0 ? 1 : 0);
        int64_t P = (q + target - 1) / target;
        if (P < 1) P = 1;
        if (P > kHubMaxParts) P = kHubMaxParts;
        hq[j] = q;
        hP[j] = (int32_t)P;
        hP64[j] = P;
        const int64_t cs = chunk_span((int)P, chunk_mul);
        hch[j] = (q + cs - 1) / cs;
        hpc[j] = hch[j] * P;
    }
}

// owner[i] = the j with base[j] <= i < base[j+1]; rank[i] = i - base[j].
__global__ void k_expand_owner(const int64_t *base, int32_t nseg, int64_t total, int32_t *owner,
                               int32_t *rank)
{
    GRID_STRIDE(i, total)
    {
        int32_t lo = 0, hi = nseg;  // base[lo] <= i < base[hi]
        while (hi - lo > 1) {
            int32_t mid = (lo + hi) >> 1;
            if (base[mid] <= i) lo = mid; else hi = mid;
        }
        owner[i] = lo;
        if (rank) rank[i] = (int32_t)(i - base[lo]);
    }
}

__global__ void k_widen_i32(const int32_t *in, int64_t n, int64_t *out)
{
    GRID_STRIDE(i, n) out[i] = in[i];
}

void launch_expand_owner(hipStream_t s, const int64_t *base, int32_t nseg, int64_t total,
                         int32_t *owner, int32_t *rank)
{
    if (total > 0)
        hipLaunchKernelGGL(k_expand_owner, grid_for(total, 256, 1 << 20), 256, 0, s, base, nseg,
                           total, owner, rank);
}

void launch_widen_i32(hipStream_t s, const int32_t *in, int64_t n, int64_t *out)
{
    if (n > 0) hipLaunchKernelGGL(k_widen_i32, grid_for(n, 256), 256, 0, s, in, n, out);
}

void launch_hub_setup(hipStream_t s, const int32_t *hv, int32_t nh, const int64_t *rowptr,
                      const int64_t *rowend, const int64_t *selfsum, int32_t target,
                      int chunk_mul, int64_t *hq, int32_t *hP, int64_t *hP64, int64_t *hch,
                      int64_t *hpc)
{
    if (nh > 0)
        hipLaunchKernelGGL(k_hub_setup, grid_for(nh, 256), 256, 0, s, hv, nh, rowptr, rowend, selfsum,
                           target, chunk_mul, hq, hP, hP64, hch, hpc);
}

// Entry t of hub j: t < d is row entry rowptr[v]+t, t == d the virtual
// own-community entry (weight 0) of a hub with a self-loop.
template <int WM>
__global__ __launch_bounds__(256) void k_hub_count(SweepArgs a, HubArgs h)
{
    __shared__ int hist[kHubMaxParts];
    const int32_t ci = blockIdx.x;
    const int32_t j = h.chunk_hub[ci];
    const int32_t v = h.hv[j];
    const int64_t q = h.hq[j];
    const int P = h.hP[j];
    const int64_t rb = a.rowptr[v], d = a.rowend[v] - rb;
    const int32_t own = a.C[v];
    for (int t = threadIdx.x; t < P; t += 256) hist[t] = 0;
    __syncthreads();
    const int64_t cs = chunk_span(P, h.chunk_mul);
    const int64_t t0 = (int64_t)h.chunk_k[ci] * cs;
    const int64_t t1 = min(q, t0 + cs);
    int32_t *tc = h.tmp_c + h.ebase[j];
    // four community gathers in flight per thread before their stores and counts
    for (int64_t p0 = t0 + threadIdx.x; p0 < t1; p0 += 4 * 256) {
        int32_t y[4], c[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t t = p0 + (int64_t)u * 256;
            y[u] = (t < t1 && t < d) ? col_at<WM>(a, rb + t) : -1;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) c[u] = y[u] >= 0 ? a.C[y[u]] : own;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t t = p0 + (int64_t)u * 256;
            if (t < t1) {
                tc[t] = c[u];
                atomicAdd(&hist[hub_part(c[u], P)], 1);
            }
        }
    }
    __syncthreads();
    int32_t *ch = h.chist + h.hbase[j] + (int64_t)h.chunk_k[ci] * P;
    for (int t = threadIdx.x; t < P; t += 256) {
        ch[t] = hist[t];  // the scatter reserves the chunk's ranges from these
        if (hist[t]) atomicAdd((ull *)&h.pcount[h.pbase[j] + t], (ull)hist[t]);
    }
}

// Partitioned hub entry e as (community, weight).
__device__ __forceinline__ PEntry hub_entry(const HubArgs &h, int64_t e)
{
    if (h.buf4) {
        const uint32_t x = h.buf4[e];
        return PEntry{(int32_t)(x >> h.pwb), 0, (int64_t)(x & ((1u << h.pwb) - 1u))};
    }
    if (h.buf8) {
        const PEntry8 x = h.buf8[e];
        return PEntry{x.c, 0, (int64_t)x.w};
    }
    return h.buf[e];
}

template <int WM>
__global__ __launch_bounds__(256) void k_hub_scatter(SweepArgs a, HubArgs h)
{
    __shared__ int hist[kHubMaxParts];
    __shared__ long long lbase[kHubMaxParts];
    const int32_t ci = blockIdx.x;
    const int32_t j = h.chunk_hub[ci];
    const int32_t v = h.hv[j];
    const int64_t q = h.hq[j];
    const int P = h.hP[j];
    const int64_t rb = a.rowptr[v], d = a.rowend[v] - rb;
    for (int t = threadIdx.x; t < P; t += 256) hist[t] = 0;
    __syncthreads();
    const int64_t cs = chunk_span(P, h.chunk_mul);
    const int64_t t0 = (int64_t)h.chunk_k[ci] * cs;
    const int64_t t1 = min(q, t0 + cs);
    const int32_t *tc = h.tmp_c + h.ebase[j];
    const int32_t *ch = h.chist + h.hbase[j] + (int64_t)h.chunk_k[ci] * P;
    for (int t = threadIdx.x; t < P; t += 256) {
        const int32_t k = ch[t];
        if (k) {  // the run [old - k, old) of the count left in pcount
            const int64_t p = h.pbase[j] + t;
            lbase[t] = h.poff[p] + (long long)atomicAdd((ull *)&h.pcount[p], (ull)(-(int64_t)k)) - k;
        }
    }
    __syncthreads();
    for (int64_t t = t0 + threadIdx.x; t < t1; t += 256) {
        int32_t c = tc[t];
        int part = hub_part(c, P);
        int64_t pos = lbase[part] + atomicAdd(&hist[part], 1);
        int64_t w = 0;
        if (t < d) {
            int32_t jx;
            entry_at<WM>(a, rb + t, jx, w);
        }
        if (h.buf4) {
            h.buf4[pos] = ((uint32_t)c << h.pwb) | (uint32_t)w;
        } else if (h.buf8) {
            h.buf8[pos] = PEntry8{c, (uint32_t)w};
        } else {
            PEntry en;
            en.c = c;
            en.pad = 0;
            en.w = w;
            h.buf[pos] = en;
        }
    }
}


template <bool B, bool N32, int L2>
__global__ __launch_bounds__(256) void k_hub_agg(SweepArgs a, HubArgs h)
{
    constexpr int CAP = 1 << L2;
    __shared__ __align__(16) int32_t s_key[CAP];
    __shared__ TabV<N32> s_val[CAP];
    const int tid = threadIdx.x;
    const int64_t pi = blockIdx.x;
    const int64_t beg = h.poff[pi], end = h.poff[pi + 1];
    const int32_t j = h.part_hub[pi];
    const int32_t v = h.hv[j];
    const VParam vp = vparam<B>(a, v);
    const double f = (double)a.kdeg[v] / vp.m2;
    // 2 slots per entry up to the LDS capacity; a full table (possible only
    // with more than CAP/2 entries) sends the partition to its global table
    const int l2 = (end - beg) * 2 >= CAP ? L2 : table_log2(end - beg, 6);
    const int cap = 1 << l2;
    for (int s = tid; s < cap; s += 256) {
        s_key[s] = -1;
        s_val[s] = 0;
    }
    __syncthreads();
    // each thread folds its runs of one community (a partition can hold 10^5
    // entries of one popular community) into one insert
    KeyRun run;
    for (int64_t e0 = beg + tid; e0 < end; e0 += 4 * 256) {
        PEntry en[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * 256 < end) en[u] = hub_entry(h, e0 + u * 256);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * 256 < end && run.add(en[u].c, en[u].w))
                (void)lds_insert_bounded(s_key, s_val, l2, run.pk, run.pw);
    }
    if (run.k >= 0) (void)lds_insert_bounded(s_key, s_val, l2, run.k, run.w);
    __syncthreads();
    int cnt = 0;
    double bg = 0.0;
    int32_t bc = INT_MAX;
    int64_t bw = 0;
    consider_slots<256>(a, s_key, s_val, tid, cap, vp.gamma, f, cnt, bg, bc, bw);
    group_reduce<64>(cnt, bg, bc, bw);
    const bool lead = block_best<4>(s_key, cnt, bg, bc, bw);
    // An insert that found no slot left the table full, which more than CAP
    // entries can do: then (or when exactly CAP keys filled it) the partition
    // is redone in its global region.  The flag sits past block_best's scratch.
    int *redo = s_key + 32;
    if (lead) *redo = (cnt == cap && end - beg > cap) ? 1 : 0;
    __syncthreads();
    if (*redo) {
        cnt = 0;
        bg = 0.0;
        bc = INT_MAX;
        bw = 0;
        const int64_t gcap = 2 * (end - beg);
        int32_t *gk = h.fbk + 2 * beg;
        ull *gv = h.fbv + 2 * beg;
        for (int64_t s = tid; s < gcap; s += 256) {
            gk[s] = -1;
            gv[s] = 0;
        }
        __threadfence();
        __syncthreads();
        for (int64_t e = beg + tid; e < end; e += 256) {
            const PEntry en = hub_entry(h, e);
            glb_insert_mod(gk, gv, gcap, en.c, en.w);
        }
        __threadfence();
        __syncthreads();
        for (int64_t s = tid; s < gcap; s += 256) {
            int32_t c = __hip_atomic_load(&gk[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (c >= 0)
                consider(c, (int64_t)__hip_atomic_load(&gv[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                         deg_of(a, c), vp.gamma, f, cnt, bg, bc, bw);
        }
        group_reduce<64>(cnt, bg, bc, bw);
        if (!block_best<4>(s_key, cnt, bg, bc, bw)) return;
    } else if (!lead) {
        return;
    }
    h.pn[pi] = cnt;
    h.pg[pi] = bg;
    h.pc[pi] = bc;
    h.pw[pi] = bw;
}

// One wavefront per hub: combine its partitions, then guard/move as usual.
template <bool B>
__global__ __launch_bounds__(256) void k_hub_final(SweepArgs a, HubArgs h, int32_t nh)
{
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    int cnt = 0;
    double bg = 0.0;
    int32_t bc = INT_MAX;
    int64_t bw = 0;
    if (j < nh) {
        const int64_t base = h.pbase[j];
        for (int p = lane; p < h.hP[j]; p += 64) {
            cnt += h.pn[base + p];
            if (better(h.pg[base + p], h.pc[base + p], bg, bc)) {
                bg = h.pg[base + p];
                bc = h.pc[base + p];
                bw = h.pw[base + p];
            }
        }
    }
    group_reduce<64>(cnt, bg, bc, bw);
    if (j < nh && lane == 0) {
        const int32_t v = h.hv[j];
        const Lab me = own_lab(a, v);
        finalize_vertex(a, v, me.c, me.sz, cnt, bc, bw, vparam<B>(a, v).mode);
    }
}

int run_hub_sweep(Arena &A, hipStream_t s, const SweepArgs &a, const HubArgs &h, int32_t nh,
                  int64_t nparts, int64_t nchunks, bool batch)
{
    if (nh <= 0) return LV_OK;
    // pcount is zero here (the plan clears it, every scatter counts it back down)
    const int wm = entry_mode(a);
    if (wm == 2)
        hipLaunchKernelGGL(k_hub_count<2>, (int)nchunks, 256, 0, s, a, h);
    else
        hipLaunchKernelGGL(k_hub_count<0>, (int)nchunks, 256, 0, s, a, h);
    LV_TRY(scan_exclusive_i64(A, s, h.pcount, (int64_t *)h.poff, nparts, (int64_t *)h.poff + nparts));
    if (wm == 2)
        hipLaunchKernelGGL(k_hub_scatter<2>, (int)nchunks, 256, 0, s, a, h);
    else if (wm == 1)
        hipLaunchKernelGGL(k_hub_scatter<1>, (int)nchunks, 256, 0, s, a, h);
    else
        hipLaunchKernelGGL(k_hub_scatter<0>, (int)nchunks, 256, 0, s, a, h);
    // a partition's sums stay below a narrow level's total degree: 4-byte table
    const bool n32 = a.lab8 != nullptr;
    constexpr int LB = kHubAggLog2, LS = kHubAggLog2Small;
    if (batch) {
        if (n32 && h.small_tables)
            hipLaunchKernelGGL((k_hub_agg<true, true, LS>), (int)nparts, 256, 0, s, a, h);
        else if (n32)
            hipLaunchKernelGGL((k_hub_agg<true, true, LB>), (int)nparts, 256, 0, s, a, h);
        else
            hipLaunchKernelGGL((k_hub_agg<true, false, LB>), (int)nparts, 256, 0, s, a, h);
        hipLaunchKernelGGL(k_hub_final<true>, (int)((nh + 3) / 4), 256, 0, s, a, h, nh);
    } else {
        if (n32 && h.small_tables)
            hipLaunchKernelGGL((k_hub_agg<false, true, LS>), (int)nparts, 256, 0, s, a, h);
        else if (n32)
            hipLaunchKernelGGL((k_hub_agg<false, true, LB>), (int)nparts, 256, 0, s, a, h);
        else
            hipLaunchKernelGGL((k_hub_agg<false, false, LB>), (int)nparts, 256, 0, s, a, h);
        hipLaunchKernelGGL(k_hub_final<false>, (int)((nh + 3) / 4), 256, 0, s, a, h, nh);
    }
    LV_HIP(hipGetLastError());
    return LV_OK;
}

// ---------------------------------------------------------------------------
// Dendrogram flattening and the final modularity.

__global__ void k_compose(int32_t *f, int32_t n, const int32_t *map)
{
    GRID_STRIDE(i, n) f[i] = map[f[i]];
}

__global__ void k_gather_partition(const int32_t *perm, const int32_t *ids, const int32_t *cbp,
                                   int32_t n, int32_t *node, int32_t *comm)
{
    GRID_STRIDE(j, n)
    {
        int32_t p = perm[j];
        node[j] = ids ? ids[p] : p;
        comm[j] = cbp[p];
    }
}

// Final modularity aggregates: d_c += w at both ends, l_c += w when both ends
// are in c.  Each block combines a chunk of edges in an LDS table first (big
// communities would otherwise serialise on their counters).
constexpr int kFqLog2 = 12;
constexpr int kFqChunk = 2048;

__global__ __launch_bounds__(256) void k_final_q(const int32_t *src, const int32_t *dst,
                                                WUnits w, int64_t e, const int32_t *comm,
                                                ull *lc, ull *dc)
{
    constexpr int CAP = 1 << kFqLog2;
    __shared__ int32_t s_key[CAP];
    __shared__ ull s_d[CAP], s_l[CAP];
    for (int64_t base = (int64_t)blockIdx.x * kFqChunk; base < e;
         base += (int64_t)gridDim.x * kFqChunk) {
        for (int s = threadIdx.x; s < CAP; s += 256) {
            s_key[s] = -1;
            s_d[s] = 0;
            s_l[s] = 0;
        }
        __syncthreads();
        for (int64_t i = base + threadIdx.x; i < e && i < base + kFqChunk; i += 256) {
            int32_t a = comm[src[i]], b = comm[dst[i]];
            ull u = (ull)w[i];
            if (u == 0) continue;
            if (!lds_insert_bounded(s_key, s_d, kFqLog2, a, (int64_t)u)) {
                atomicAdd(&dc[a], u);
                if (a == b) atomicAdd(&lc[a], u);
            } else if (a == b) {
                // same slot: add the internal weight too
                uint32_t mask = CAP - 1, sl = hslot(a, kFqLog2);
                while (((volatile int32_t *)s_key)[sl] != a) sl = (sl + 1) & mask;
                atomicAdd(&s_l[sl], u);
            }
            if (!lds_insert_bounded(s_key, s_d, kFqLog2, b, (int64_t)u)) atomicAdd(&dc[b], u);
        }
        __syncthreads();
        for (int s = threadIdx.x; s < CAP; s += 256) {
            int32_t k = s_key[s];
            if (k >= 0) {
                if (s_d[s]) atomicAdd(&dc[k], s_d[s]);
                if (s_l[s]) atomicAdd(&lc[k], s_l[s]);
            }
        }
        __syncthreads();
    }
}

void launch_compose(hipStream_t s, int32_t *f, int32_t n, const int32_t *map)
{
    if (n > 0) hipLaunchKernelGGL(k_compose, grid_for(n, 256), 256, 0, s, f, n, map);
}

void launch_gather_partition(hipStream_t s, const int32_t *perm, const int32_t *ids,
                             const int32_t *comm_by_pos, int32_t n, int32_t *node, int32_t *comm)
{
    if (n > 0)
        hipLaunchKernelGGL(k_gather_partition, grid_for(n, 256), 256, 0, s, perm, ids, comm_by_pos,
                           n, node, comm);
}

// ---------------------------------------------------------------------------
// Cluster sizes and circle layout of a partition (calculateClusterSize and
// layoutProcess, LouvainCUDA/LouvainCPP.cpp:555-651).  With positions stably
// sorted by cluster, position j of the sorted order is the reference's
// clusterSize[c-1] + indexInCluster: the node's slot on the circle.

__global__ void k_layout_first(const uint32_t *skey, int32_t n, int32_t ncomm, int32_t *first,
                               DevFlags *f)
{
    int bad = 0;
    GRID_STRIDE(j, n)
    {
        const uint32_t c = skey[j];
        if (c >= (uint32_t)ncomm) {
            bad |= 1;
            continue;
        }
        if (j == 0 || skey[j - 1] != c) first[c] = (int32_t)j;
    }
    bad = wave_or_i(bad);
    if ((threadIdx.x & 63) == 0 && bad) atomicOr(&f->err, 1);
}

__global__ void k_layout_place(const uint32_t *skey, const int32_t *perm, int32_t n, int32_t ncomm,
                               const int32_t *first, int32_t *count, int32_t *index_in_cluster,
                               float *x, float *y)
{
    GRID_STRIDE(j, n)
    {
        const uint32_t c = skey[j];
        if (c >= (uint32_t)ncomm) continue;
        const int32_t v = perm[j];
        index_in_cluster[v] = (int32_t)j - first[c];
        if (j == n - 1 || skey[j + 1] != c) count[c] = (int32_t)j + 1 - first[c];
        const int32_t i = (int32_t)j;
        x[v] = (float)(50 * cos(2 * M_PI * i / n));
        y[v] = (float)(50 + 50 * sin(2 * M_PI * i / n));
    }
}

void launch_layout(hipStream_t s, const uint32_t *skey, const int32_t *perm, int32_t n,
                   int32_t ncomm, int32_t *first, int32_t *count, int32_t *index_in_cluster,
                   float *x, float *y, DevFlags *f)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_layout_first, grid_for(n, 256), 256, 0, s, skey, n, ncomm, first, f);
    hipLaunchKernelGGL(k_layout_place, grid_for(n, 256), 256, 0, s, skey, perm, n, ncomm, first,
                       count, index_in_cluster, x, y);
}

void launch_final_q(hipStream_t s, const int32_t *src, const int32_t *dst, WUnits w,
                    int64_t e, const int32_t *comm, ull *lc, ull *dc)
{
    if (e > 0)
        hipLaunchKernelGGL(k_final_q, grid_for(e, kFqChunk, 4096), 256, 0, s, src, dst, w, e, comm,
                           lc, dc);
}

}  // namespace lv
