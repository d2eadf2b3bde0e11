// device.hpp - device helpers shared by the engine's kernel files (64-lane
// wavefronts, gfx950).
#pragma once

#include <climits>
#include <type_traits>

#include "engine.hpp"

namespace lv {

static inline int grid_for(int64_t n, int block, int64_t cap = 1 << 16)
{
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

#define GRID_STRIDE(i, n)                                                          \
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); \
         i += (int64_t)gridDim.x * blockDim.x)

// Smallest s >= 0 with w * 2^s integral, from the float32 bit pattern; -1 for
// a negative, infinite or NaN weight.
__device__ __forceinline__ int f32_unit_exponent(uint32_t x)
{
    if ((x & 0x7fffffffu) == 0) return 0;  // +0 or -0
    if (x >> 31) return -1;
    uint32_t ex = (x >> 23) & 0xffu, mant = x & 0x7fffffu;
    if (ex == 0xffu) return -1;
    uint32_t M = ex ? (mant | 0x800000u) : mant;
    int e2 = ex ? (int)ex - 150 : -149;  // w = M * 2^e2
    int s = -e2 - __builtin_ctz(M);
    return s > 0 ? s : 0;
}

// Weight in units of 2^-scale from its raw 32-bit pattern (int32 or float32);
// sets bit 2 of bad when it reaches the exact range limit.
__device__ __forceinline__ int64_t weight_units(uint32_t x, int wtype, int scale, int &bad)
{
    int64_t u;
    if (wtype == LV_WEIGHT_INT32) {
        u = (int32_t)x;
    } else if ((x & 0x7fffffffu) == 0) {
        u = 0;
    } else {
        uint32_t ex = (x >> 23) & 0xffu, mant = x & 0x7fffffu;
        uint64_t M = ex ? (mant | 0x800000u) : mant;
        int sh = (ex ? (int)ex - 150 : -149) + scale;
        if (sh > 36) {
            bad |= 2;
            u = 0;
        } else {
            u = (int64_t)(sh >= 0 ? (M << sh) : (M >> (-sh)));
        }
    }
    if (u >= kUnitLimit) bad |= 2;
    return u;
}

// Butterfly exchange: lane l receives v from lane l ^ OFF (OFF < 64; within an
// aligned group of G > OFF lanes this stays inside the group).  Offsets 1 and 2
// are DPP quad permutations (no LDS hardware), 4-16 ds_swizzle in bitmask mode
// (and 0x1f, xor OFF within each 32-lane half), 32 a ds_bpermute across the
// halves.  All lanes of the wave must be active.
template <int OFF>
__device__ __forceinline__ int xor_lanes32(int v)
{
    if constexpr (OFF == 1)
        return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    else if constexpr (OFF == 2)
        return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    else if constexpr (OFF < 32)
        return __builtin_amdgcn_ds_swizzle(v, 0x1F | (OFF << 10));
    else
        return __shfl_xor(v, OFF, 64);
}

template <int OFF, typename T>
__device__ __forceinline__ T xor_lanes(T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, xor_lanes32<OFF>(__builtin_bit_cast(int, v)));
    } else {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const int lo = xor_lanes32<OFF>((int)(unsigned)u);
        const int hi = xor_lanes32<OFF>((int)(unsigned)(u >> 32));
        return __builtin_bit_cast(T, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    }
}

__device__ __forceinline__ long long wave_sum(long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int wave_max_i(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// OR of an error-flag word over the wavefront (flags of different lanes may
// hold different bits).
__device__ __forceinline__ int wave_or_i(int v)
{
    for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, 64);
    return v;
}

// Sums of weight units saturate at kUnitLimit (both operands are at most
// kUnitLimit, so a + b never wraps): reaching the limit signals the range error.
__device__ __forceinline__ ull sat_add(ull a, ull b)
{
    const ull s = a + b;
    return s >= (ull)kUnitLimit ? (ull)kUnitLimit : s;
}

__device__ __forceinline__ ull wave_sum_sat(ull v)
{
    for (int off = 32; off > 0; off >>= 1) v = sat_add(v, __shfl_xor(v, off, 64));
    return v;
}

// End of a 256-thread block of a unit-summing kernel: the block's totals
// reach the device flags with one atomic each (one address shared by the
// whole grid, so per-wave atomics would serialise).
__device__ __forceinline__ void block_units_flush(ull tot, ull msum, ull umax, int bad, DevFlags *f)
{
    __shared__ ull s_t[4], s_m[4], s_x[4];
    __shared__ int s_b[4];
    tot = wave_sum_sat(tot);
    msum = wave_sum_sat(msum);
    for (int off = 32; off > 0; off >>= 1) umax = max(umax, __shfl_xor(umax, off, 64));
    bad = wave_or_i(bad);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_t[wave] = tot;
        s_m[wave] = msum;
        s_x[wave] = umax;
        s_b[wave] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ull T = 0, M = 0, X = 0;
        int B = 0;
        for (int w = 0; w < 4; ++w) {
            T = sat_add(T, s_t[w]);
            M = sat_add(M, s_m[w]);
            X = max(X, s_x[w]);
            B |= s_b[w];
        }
        if (T) {
            const ull old = atomicAdd((ull *)&f->total_units, T);
            if (T >= (ull)kUnitLimit || old + T >= (ull)kUnitLimit) B |= 2;
        }
        if (M) atomicAdd((ull *)&f->m_units, M);
        if (X) atomicMax(&f->max_units, X);
        if (B) atomicOr(&f->err, B);
    }
}

// End of a kernel that raises error flags: the block's flags reach the device
// word with one atomic (all threads of the block call it).  A flag such as
// "rows not increasing" can be set in every wavefront of a big launch, where
// per-wave atomics on the one word would serialise.
__device__ __forceinline__ void block_flag_or(int bad, DevFlags *f)
{
    __shared__ int s_or;
    if (threadIdx.x == 0) s_or = 0;
    __syncthreads();
    bad = wave_or_i(bad);
    if ((threadIdx.x & 63) == 0 && bad) atomicOr(&s_or, bad);
    __syncthreads();
    if (threadIdx.x == 0 && s_or) atomicOr(&f->err, s_or);
}

// 128-bit accumulate of deg^2 as (lo, hi).
__device__ __forceinline__ void acc_sq(ull &lo, ull &hi, ull d)
{
    ull plo = d * d, phi = __umul64hi(d, d);
    lo += plo;
    hi += phi + (lo < plo ? 1ull : 0ull);
}

__device__ __forceinline__ void wave_sum_u128(ull &lo, ull &hi)
{
    for (int off = 32; off > 0; off >>= 1) {
        ull olo = __shfl_xor(lo, off, 64), ohi = __shfl_xor(hi, off, 64);
        lo += olo;
        hi += ohi + (lo < olo ? 1ull : 0ull);
    }
}

__device__ __forceinline__ void atomic_add_u128(ull *lo_p, ull *hi_p, ull lo, ull hi)
{
    ull old = atomicAdd(lo_p, lo);
    ull carry = (old + lo < old) ? 1ull : 0ull;
    if (hi + carry) atomicAdd(hi_p, hi + carry);
}

// Adds val into base[key] for keys sorted within the wave: lanes holding the
// same key (contiguous runs) are combined first, one atomic per run.  Invalid
// lanes may only follow the valid ones (a lane that must not add in the middle
// of a run keeps the run's key with val 0).
__device__ __forceinline__ void seg_atomic_add(bool valid, long long key, ull val, ull *base)
{
    const int lane = threadIdx.x & 63;
    long long k = valid ? key : LLONG_MIN;
    ull v = valid ? val : 0ull;
    for (int off = 1; off < 64; off <<= 1) {
        ull ov = __shfl_up(v, off, 64);
        long long ok = __shfl_up(k, off, 64);
        if (lane >= off && ok == k) v += ov;
    }
    long long nk = __shfl_down(k, 1, 64);
    bool tail = (lane == 63) || (nk != k);
    if (valid && tail && v != 0) atomicAdd(&base[k], v);
}

// Adds the pass deltas of community c to its degree and size and clears them
// (size_shift as in SweepArgs); deg32, when given, follows deg.
__device__ __forceinline__ int64_t apply_deltas(int64_t c, int64_t *deg, int32_t *size, ull *ddeg,
                                                int32_t *dsize, int size_shift, uint32_t *deg32)
{
    int64_t dv = deg[c];
    const ull dd = ddeg[c];
    if (size_shift) {
        if (dd) {
            const int64_t ds = (int64_t)(dd << (64 - size_shift)) >> (64 - size_shift);
            dv += (int64_t)(dd - (ull)ds) >> size_shift;
            deg[c] = dv;
            if (deg32) deg32[c] = (uint32_t)dv;
            size[c] += (int32_t)ds;
            ddeg[c] = 0;
        }
    } else {
        const int32_t ds = dsize[c];
        if (dd | (ull)(uint32_t)ds) {
            dv += (int64_t)dd;
            deg[c] = dv;
            if (deg32) deg32[c] = (uint32_t)dv;
            size[c] += ds;
            ddeg[c] = 0;
            dsize[c] = 0;
        }
    }
    return dv;
}

// Adds (1, val) of every valid lane to (cnt[key], sum[key]) in LDS with one
// atomic pair per distinct key of the wave (all lanes of the wave call it).
__device__ __forceinline__ void wave_bucket_add(bool valid, int key, long long val, long long *cnt,
                                                long long *sum)
{
    const int lane = threadIdx.x & 63;
    ull rem = __ballot(valid);
    while (rem) {
        const int leader = __ffsll((long long)rem) - 1;
        const int k = __shfl(key, leader, 64);
        const bool mine = valid && key == k;
        const ull m = __ballot(mine);
        const long long t = wave_sum(mine ? val : 0);
        if (lane == 0) {
            atomicAdd((ull *)&cnt[k], (ull)__popcll(m));
            if (sum && t) atomicAdd((ull *)&sum[k], (ull)t);
        }
        rem &= ~m;
    }
}

// Running segmented sum of a wavefront over consecutive rounds of 64 keyed
// values whose keys never decrease along the wavefront's span: each run of
// the span reaches base[key] with one atomic (seg_atomic_add takes one per
// run per round, so a row of 10^6 entries would take 1.6e4 atomics on one
// address, where device atomics serialise).  All lanes call add() every round
// and flush() after the span; invalid lanes only at the end of the data.
constexpr int kSpanRounds = 8;

// Stores weight u of CSR entry p: in the 4-byte array when the level keeps
// one (every weight of the level below 2^32), else in the 8-byte one.
__device__ __forceinline__ void put_w(int64_t *w, uint32_t *w32, int64_t p, int64_t u)
{
    if (w32)
        w32[p] = (uint32_t)u;
    else
        w[p] = u;
}

// Stores CSR entry p (column cc, weight u): with pk > 0 as one word
// cc << pk | u in col, else the column and the weight (put_w).
__device__ __forceinline__ void put_entry(int32_t *col, int64_t *w, uint32_t *w32, int pk,
                                          int64_t p, int32_t cc, int64_t u)
{
    if (pk) {
        col[p] = (int32_t)(((uint32_t)cc << pk) | (uint32_t)u);
    } else {
        col[p] = cc;
        put_w(w, w32, p, u);
    }
}

struct SegCarry {
    long long key = LLONG_MIN;
    ull sum = 0;

    __device__ __forceinline__ void add(bool valid, long long k_in, ull val, ull *base)
    {
        const int lane = threadIdx.x & 63;
        long long k = valid ? k_in : LLONG_MIN;
        ull v = valid ? val : 0ull;
        if (lane == 0) {
            if (k == key)
                v += sum;
            else if (key != LLONG_MIN && sum != 0)
                atomicAdd(&base[key], sum);
        }
        for (int off = 1; off < 64; off <<= 1) {
            const ull ov = __shfl_up(v, off, 64);
            const long long ok = __shfl_up(k, off, 64);
            if (lane >= off && ok == k) v += ov;
        }
        const long long nk = __shfl_down(k, 1, 64);
        const bool tail = (lane == 63) || (nk != k);
        if (valid && tail && lane != 63 && v != 0) atomicAdd(&base[k], v);
        key = __shfl(k, 63, 64);
        sum = __shfl(v, 63, 64);
    }

    __device__ __forceinline__ void flush(ull *base)
    {
        if ((threadIdx.x & 63) == 0 && key != LLONG_MIN && sum != 0) atomicAdd(&base[key], sum);
        key = LLONG_MIN;
        sum = 0;
    }
};

// Tier of a row of d entries and q candidates; rows with q > hub_q (at most
// kTierMaxQ[kTiers - 2]) take the hub partitions; below that the first tier whose
// bound holds q.
__device__ __forceinline__ int tier_of(int64_t d, int64_t q, int64_t hub_q)
{
    if (d == 0) return 0;  // no neighbours: the vertex can never move
    if (q > hub_q) return kTiers;
#pragma unroll
    for (int t = 0; t < kTiers - 1; ++t)
        if (q <= kTierMaxQ[t]) return t + 1;
    return kTiers;
}

__device__ __forceinline__ uint32_t hslot(int32_t c, int log2cap)
{
    return ((uint32_t)c * 0x9E3779B1u) >> (32 - log2cap);
}

// LDS tables keep their sums (V) and community totals (D) in 8 bytes, or in 4
// on narrow levels (total degree below 2^32 bounds every sum and total).
template <bool N32>
using TabV = typename std::conditional<N32, uint32_t, ull>::type;
template <bool N32>
using TabD = typename std::conditional<N32, uint32_t, int64_t>::type;

template <typename V>
__device__ __forceinline__ void lds_insert(int32_t *key, V *val, int log2cap, int32_t c, int64_t w)
{
    const uint32_t mask = (1u << log2cap) - 1u;
    uint32_t s = hslot(c, log2cap);
    for (;;) {
        int32_t k = ((volatile int32_t *)key)[s];
        if (k == c) break;
        if (k == -1) {
            int32_t prev = atomicCAS(&key[s], -1, c);
            if (prev == -1 || prev == c) break;
        }
        s = (s + 1) & mask;
    }
    if (w != 0) atomicAdd(&val[s], (V)w);
}

// As lds_insert; the lane that creates a slot also stores the key's
// community total d in dg[slot].
template <typename V, typename D>
__device__ __forceinline__ void lds_insert_d(int32_t *key, V *val, D *dg, int log2cap, int32_t c,
                                             int64_t w, int64_t d)
{
    const uint32_t mask = (1u << log2cap) - 1u;
    uint32_t s = hslot(c, log2cap);
    for (;;) {
        int32_t k = ((volatile int32_t *)key)[s];
        if (k == c) break;
        if (k == -1) {
            int32_t prev = atomicCAS(&key[s], -1, c);
            if (prev == -1) {
                dg[s] = (D)d;
                break;
            }
            if (prev == c) break;
        }
        s = (s + 1) & mask;
    }
    if (w != 0) atomicAdd(&val[s], (V)w);
}

__device__ __forceinline__ bool better(double g1, int32_t c1, double g2, int32_t c2)
{
    return g1 > g2 || (g1 == g2 && c1 < c2);
}

// Table size for a row of q inserts: the smallest power of two >= 2q, at least
// 2^lo (never above the kernel's LDS capacity, which the tier bound ensures).
__device__ __forceinline__ int table_log2(int64_t q, int lo)
{
    int l = lo;
    while ((int64_t(1) << l) < 2 * q) ++l;
    return l;
}

__device__ __forceinline__ uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

// Partition of community c among P; independent of hslot's multiplicative hash
// so the keys of one partition spread over its table.
__device__ __forceinline__ int hub_part(int32_t c, int P)
{
    return (int)(((uint64_t)fmix32((uint32_t)c) * (uint64_t)P) >> 32);
}

// Entries per chunk of a row or segment cut into P hash partitions: at least
// kHubChunk, and mul entries per partition when mul > 0 (longer write runs per
// partition for large P, fewer and longer blocks).
__host__ __device__ inline int64_t chunk_span(int P, int mul)
{
    const int64_t c = (int64_t)P * mul;
    return c > kHubChunk ? c : kHubChunk;
}

// Insert with at most 2^log2cap probes; false when the table is full.
template <typename V>
__device__ __forceinline__ bool lds_insert_bounded(int32_t *key, V *val, int log2cap, int32_t c,
                                                   int64_t w)
{
    const uint32_t mask = (1u << log2cap) - 1u;
    uint32_t s = hslot(c, log2cap);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        int32_t k = ((volatile int32_t *)key)[s];
        if (k == -1) {
            int32_t prev = atomicCAS(&key[s], -1, c);
            if (prev == -1 || prev == c) {
                if (w != 0) atomicAdd(&val[s], (V)w);
                return true;
            }
        } else if (k == c) {
            if (w != 0) atomicAdd(&val[s], (V)w);
            return true;
        }
        s = (s + 1) & mask;
    }
    return false;
}

// A thread's run of equal keys and its summed weight, so that a run costs
// one LDS table insert instead of one per entry (the entries of one key meet
// on one LDS word, where atomics serialise).  add()
// returns true when the entry ends the previous run, which is then (pk, pw)
// and must be added by the caller; after the last entry, (k, w) is the run
// still open (k < 0: none).
struct KeyRun {
    int32_t k = -1, pk = -1;
    int64_t w = 0, pw = 0;
    __device__ __forceinline__ bool add(int32_t key, int64_t wt)
    {
        if (key == k) {
            w += wt;
            return false;
        }
        pk = k;
        pw = w;
        k = key;
        w = wt;
        return pk >= 0;
    }
};

// Global open-addressing table of `cap` slots (cap > distinct keys).
__device__ __forceinline__ void glb_insert_mod(int32_t *key, ull *val, int64_t cap, int32_t c,
                                               int64_t w)
{
    int64_t s = (int64_t)(fmix32((uint32_t)c ^ 0x5bd1e995u) % (uint64_t)cap);
    for (;;) {
        int32_t k = __hip_atomic_load(&key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == c) break;
        if (k == -1) {
            int32_t prev = atomicCAS(&key[s], -1, c);
            if (prev == -1 || prev == c) break;
        }
        s = (s + 1 == cap) ? 0 : s + 1;
    }
    if (w != 0) atomicAdd(&val[s], (ull)w);
}


}  // namespace lv
