// coarsen.hip - renumbering and coarsening by communities (gfx950).
//
// Renumbering (renumberCUDA, LouvainCUDA/RenumberKernel.cu:63-220): positions are
// stably radix-sorted by their community; the head of each group is the
// community's smallest member, and the new id of a community is the rank of that
// head among all heads in position order - the reference's first-occurrence
// rule.  The sorted positions are also the order of the reference's Partition
// arrays (:121-133,191,202-203) and the member lists coarsening walks.
//
// Coarsening (inducedGraphCUDA, LouvainCUDA/InducedGraphKernel.cu:28-113): every
// community's row entries are gathered contiguously in member order as
// (newc[target], w' = w==0?1:w); each community segment is then aggregated by
// neighbour community - in one LDS table for segments of up to 2048 entries, or
// by hash partitions of about 2048 entries for larger ones - which gives row x
// of the next level: one entry (y, E'(x,y) + E'(y,x)) per neighbour community
// y != x.  Entries with y == x and the members' self-loops form the next
// level's self-loop E'(x,x) = internal/2 + sum(selfsum).
#include "device.hpp"

namespace lv {

// Weight of member-order entry e (4-byte array on narrow levels).
__device__ __forceinline__ int64_t entry_w(const CoarsenArgs &a, int64_t e)
{
    return a.ew32 ? (int64_t)a.ew32[e] : a.ew[e];
}

// Member-order entry e as (next-level id y, weight w').
__device__ __forceinline__ void entry_of(const CoarsenArgs &a, int64_t e, int32_t &y, int64_t &w)
{
    if (a.ep) {
        const uint32_t x = a.ep[e];
        y = (int32_t)(x >> a.ewb);
        w = (int64_t)(x & ((1u << a.ewb) - 1u));
    } else {
        y = a.ey[e];
        w = entry_w(a, e);
    }
}

// Partition entry e as (community, weight).
__device__ __forceinline__ PEntry part_entry(const BigSegArgs &B, int64_t e)
{
    if (B.pe4) {
        const uint32_t x = B.pe4[e];
        return PEntry{(int32_t)(x >> B.pwb), 0, (int64_t)(x & ((1u << B.pwb) - 1u))};
    }
    if (B.pe8) {
        const PEntry8 x = B.pe8[e];
        return PEntry{x.c, 0, (int64_t)x.w};
    }
    return B.pe[e];
}

// ---------------------------------------------------------------------------
// renumbering

__global__ void k_cr_heads(const uint32_t *Cs, int32_t n, int32_t *head)
{
    GRID_STRIDE(j, n) head[j] = (j == 0 || Cs[j] != Cs[j - 1]) ? 1 : 0;
}

// hscan = exclusive scan of head; group of sorted element j = hscan[j]+head[j]-1.
__global__ void k_cr_groups(const int32_t *head, const int32_t *hscan, const int32_t *perm,
                            int32_t n, int32_t *gstart, int32_t *lflag)
{
    GRID_STRIDE(j, n)
    {
        if (head[j]) {
            gstart[hscan[j]] = (int32_t)j;
            lflag[perm[j]] = 1;
        }
        if (j == n - 1) gstart[hscan[j] + head[j]] = n;
    }
}

__global__ void k_cr_newid(const int32_t *gstart, const int32_t *perm, const int32_t *rank,
                           int32_t K, int32_t *newid_g)
{
    GRID_STRIDE(g, K) newid_g[g] = rank[perm[gstart[g]]];
}

__global__ void k_cr_newc(const int32_t *head, const int32_t *hscan, const int32_t *perm,
                          const int32_t *newid_g, int32_t n, int32_t *newc)
{
    GRID_STRIDE(j, n) newc[perm[j]] = newid_g[hscan[j] + head[j] - 1];
}

void launch_cr_heads(hipStream_t s, const uint32_t *Cs, int32_t n, int32_t *head)
{
    if (n > 0) hipLaunchKernelGGL(k_cr_heads, grid_for(n, 256), 256, 0, s, Cs, n, head);
}

void launch_cr_groups(hipStream_t s, const int32_t *head, const int32_t *hscan, const int32_t *perm,
                      int32_t n, int32_t *gstart, int32_t *lflag)
{
    if (n > 0)
        hipLaunchKernelGGL(k_cr_groups, grid_for(n, 256), 256, 0, s, head, hscan, perm, n, gstart,
                           lflag);
}

void launch_cr_newid(hipStream_t s, const int32_t *gstart, const int32_t *perm, const int32_t *rank,
                     int32_t K, int32_t *newid_g)
{
    if (K > 0)
        hipLaunchKernelGGL(k_cr_newid, grid_for(K, 256), 256, 0, s, gstart, perm, rank, K, newid_g);
}

void launch_cr_newc(hipStream_t s, const int32_t *head, const int32_t *hscan, const int32_t *perm,
                    const int32_t *newid_g, int32_t n, int32_t *newc)
{
    if (n > 0)
        hipLaunchKernelGGL(k_cr_newc, grid_for(n, 256), 256, 0, s, head, hscan, perm, newid_g, n,
                           newc);
}

// ---------------------------------------------------------------------------
// member-order gather

// Per member j (member order): its row length and row start, so the gather
// below reaches a member's entries without dependent lookups.
__global__ void k_cr_memdeg(const int32_t *perm, const int64_t *rowptr, const int64_t *rowend,
                            int32_t n, int64_t *mdeg, int64_t *mstart)
{
    GRID_STRIDE(j, n)
    {
        int32_t a = perm[j];
        const int64_t b = rowptr[a];
        mdeg[j] = rowend[a] - b;
        mstart[j] = b;
    }
}

// selfl[g] = sum of selfsum over the members of group g (members are sorted by group).
__global__ void k_cr_selfl(const int32_t *perm, const int64_t *selfsum, const int32_t *head,
                           const int32_t *hscan, int32_t n, ull *selfl)
{
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
         base += (int64_t)gridDim.x * blockDim.x) {
        int64_t j = base + threadIdx.x;
        bool valid = j < n;
        long long g = 0;
        ull v = 0;
        if (valid) {
            g = hscan[j] + head[j] - 1;
            v = (ull)selfsum[perm[j]];
        }
        seg_atomic_add(valid, g, v, selfl);
    }
}

// bound[b] = the last member j with estart[j] <= min(b*T, total-1).
__global__ void k_lbs_bounds(const int64_t *estart, int32_t n, int64_t total, int T, int64_t nb,
                             int32_t *bound)
{
    GRID_STRIDE(b, nb + 1)
    {
        int64_t o = b * (int64_t)T;
        if (o > total - 1) o = total - 1;
        int32_t lo = 0, hi = n;  // estart[lo] <= o < estart[hi]
        while (hi - lo > 1) {
            int32_t mid = (lo + hi) >> 1;
            if (estart[mid] <= o) lo = mid; else hi = mid;
        }
        bound[b] = lo;
    }
}

constexpr int kGatherT = 1024;  // output entries per block (4 per thread)
constexpr int kGatherLds = 2048; // member offsets a block can stage in LDS

// ey/ew[o] for every output slot o of member order: the o-th entry of the
// member rows, as (newc[target], w').  The
// block's member offsets estart[lo0..hi0+1] are staged in LDS when they fit;
// a member's row start comes precomputed (mstart), so the chain
// of dependent global loads per entry is: member record, row entry, newc.
__global__ __launch_bounds__(256) void k_cr_gather(const int64_t *estart, const int32_t *bound,
                                                  const int64_t *mstart, const int32_t *col,
                                                  const int64_t *cw, const uint32_t *cw32,
                                                  const int32_t *newc,
                                                  int64_t total, int32_t *ey, int64_t *ew,
                                                  uint32_t *ew32, uint32_t *ep, int ewb, int cpk)
{
    __shared__ int64_t s_est[kGatherLds];
    const int64_t b = blockIdx.x;
    const int32_t lo0 = bound[b], hi0 = bound[b + 1];
    const int32_t span = hi0 - lo0 + 2;  // estart[lo0 .. hi0+1]
    const bool staged = span <= kGatherLds;
    if (staged) {
        for (int i = threadIdx.x; i < span; i += 256) s_est[i] = estart[lo0 + i];
    }
    __syncthreads();
    constexpr int R = kGatherT / 256;
    int32_t jj[R];
    int64_t off[R];
    bool ok[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t o = b * kGatherT + r * 256 + threadIdx.x;
        ok[r] = o < total;
        jj[r] = 0;
        off[r] = 0;
        if (ok[r]) {
            if (staged) {
                int32_t l = 0, h = span - 1;  // s_est[l] <= o < s_est[h]
                while (h - l > 1) {
                    int32_t mid = (l + h) >> 1;
                    if (s_est[mid] <= o) l = mid; else h = mid;
                }
                jj[r] = lo0 + l;
                off[r] = o - s_est[l];
            } else {
                int32_t l = lo0, h = hi0 + 1;
                while (h - l > 1) {
                    int32_t mid = (l + h) >> 1;
                    if (estart[mid] <= o) l = mid; else h = mid;
                }
                jj[r] = l;
                off[r] = o - estart[l];
            }
        }
    }
    int64_t pp[R];
#pragma unroll
    for (int r = 0; r < R; ++r) pp[r] = ok[r] ? mstart[jj[r]] + off[r] : 0;
    int32_t cc[R];
    int64_t uu[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        cc[r] = ok[r] ? __builtin_nontemporal_load(&col[pp[r]]) : 0;
        if (cpk) {  // the level's entries are col << cpk | weight
            uu[r] = (int64_t)((uint32_t)cc[r] & ((1u << cpk) - 1u));
            cc[r] = (int32_t)((uint32_t)cc[r] >> cpk);
        } else {
            uu[r] = !ok[r] ? 0
                    : cw32 ? (int64_t)__builtin_nontemporal_load(&cw32[pp[r]])
                           : __builtin_nontemporal_load(&cw[pp[r]]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (ok[r]) cc[r] = newc[cc[r]];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int64_t o = b * kGatherT + r * 256 + threadIdx.x;
        if (ok[r]) {
            const int64_t wp = uu[r] == 0 ? 1 : uu[r];
            if (ep) {
                __builtin_nontemporal_store(((uint32_t)cc[r] << ewb) | (uint32_t)wp, &ep[o]);
                continue;
            }
            __builtin_nontemporal_store(cc[r], &ey[o]);
            if (ew32)
                __builtin_nontemporal_store((uint32_t)wp, &ew32[o]);
            else
                __builtin_nontemporal_store(wp, &ew[o]);
        }
    }
}

void launch_cr_memdeg(hipStream_t s, const int32_t *perm, const int64_t *rowptr,
                      const int64_t *rowend, int32_t n, int64_t *mdeg, int64_t *mstart)
{
    if (n > 0)
        hipLaunchKernelGGL(k_cr_memdeg, grid_for(n, 256), 256, 0, s, perm, rowptr, rowend, n, mdeg,
                           mstart);
}

void launch_cr_selfl(hipStream_t s, const int32_t *perm, const int64_t *selfsum,
                     const int32_t *head, const int32_t *hscan, int32_t n, ull *selfl)
{
    if (n > 0)
        hipLaunchKernelGGL(k_cr_selfl, grid_for(n, 256, 1 << 20), 256, 0, s, perm, selfsum, head,
                           hscan, n, selfl);
}

void launch_cr_gather(hipStream_t s, const int64_t *estart, int32_t *bound, const int64_t *mstart,
                      const int32_t *col, const int64_t *cw, const uint32_t *cw32,
                      const int32_t *newc, int32_t n,
                      int64_t total, int32_t *ey, int64_t *ew, uint32_t *ew32, uint32_t *ep,
                      int ewb, int cpk)
{
    if (total <= 0) return;
    const int64_t nb = (total + kGatherT - 1) / kGatherT;
    hipLaunchKernelGGL(k_lbs_bounds, grid_for(nb + 1, 256), 256, 0, s, estart, n, total, kGatherT,
                       nb, bound);
    hipLaunchKernelGGL(k_cr_gather, (int)nb, 256, 0, s, estart, bound, mstart, col, cw, cw32, newc,
                       total, ey, ew, ew32, ep, ewb, cpk);
}

// ---------------------------------------------------------------------------
// segments: group g's entries are [sb[g], sb[g] + slen[g]) of ey/ew

// Groups whose new id is -1 (components of a graph set that stop at this
// level) are put in tier 0: they produce no next-level row.
__global__ __launch_bounds__(256) void k_cr_seg_tier(const int32_t *gstart, const int64_t *estart,
                                                    const int32_t *newid_g, int32_t K, int64_t *sb,
                                                    int64_t *slen, uint32_t *tkey, int32_t *sid,
                                                    long long *tier_cnt, int32_t hub_q)
{
    __shared__ long long s_cnt[kTiers + 1];
    if (threadIdx.x <= kTiers) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < K;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = base + threadIdx.x;
        const bool valid = g < K;
        int t = 0;
        if (valid) {
            int64_t b = estart[gstart[g]], e = estart[gstart[g + 1]];
            sb[g] = b;
            slen[g] = e - b;
            t = newid_g[g] < 0 ? 0 : tier_of(e - b, e - b, hub_q);
            tkey[g] = (uint32_t)t;
            sid[g] = (int32_t)g;
        }
        wave_bucket_add(valid, t, 0, s_cnt, nullptr);
    }
    __syncthreads();
    if (threadIdx.x <= kTiers && s_cnt[threadIdx.x])
        atomicAdd((ull *)&tier_cnt[threadIdx.x], (ull)s_cnt[threadIdx.x]);
}

void launch_cr_seg_tier(hipStream_t s, const int32_t *gstart, const int64_t *estart,
                        const int32_t *newid_g, int32_t K, int64_t *sb, int64_t *slen,
                        uint32_t *tkey, int32_t *sid, long long *tier_cnt, int32_t hub_q)
{
    if (K > 0)
        hipLaunchKernelGGL(k_cr_seg_tier, grid_for(K, 256, 1024), 256, 0, s, gstart, estart,
                           newid_g, K, sb, slen, tkey, sid, tier_cnt, hub_q);
}

// Writes the occupied slots (key >= 0) of a 256-thread block's LDS table of
// `cap` <= 256 * SPT slots compactly to oy/ow, in slot order, and returns how
// many on every thread.  Each thread stages its slots in registers, so the
// cross-wave prefix goes through the table's own first words and the kernel
// keeps no counter in LDS (a 32 KiB table leaves room for five blocks per CU).
// Ends with a barrier: the caller may reuse the table at once.
template <int SPT, typename V>
__device__ __forceinline__ int table_compact(int32_t *key, const V *val, int cap, int32_t *oy,
                                             int64_t *ow, uint32_t *ow32)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int32_t ky[SPT];
    V vv[SPT];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int s = tid + k * 256;
        ky[k] = s < cap ? key[s] : -1;
        vv[k] = ky[k] >= 0 ? val[s] : (V)0;
        mine += ky[k] >= 0 ? 1 : 0;
    }
    int inc = mine;
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(inc, off, 64);
        if (lane >= off) inc += o;
    }
    __syncthreads();  // the table is read: its first words take the wave totals
    if (lane == 63) key[wave] = inc;
    __syncthreads();
    int pos = inc - mine, total = 0;
    for (int w = 0; w < 4; ++w) {
        const int t = key[w];
        if (w < wave) pos += t;
        total += t;
    }
#pragma unroll
    for (int k = 0; k < SPT; ++k)
        if (ky[k] >= 0) {
            oy[pos] = ky[k];
            put_w(ow, ow32, pos, (int64_t)vv[k]);
            ++pos;
        }
    __syncthreads();
    return total;
}

// Segments of up to 2048 entries (or hub_q if lower): G lanes (or a block) per
// segment, LDS table sized to the segment; uniques are written at sb[g] in ey's index space.
// N32: a narrow level, whose sums all stay below 2^32, keeps 4-byte table sums.
template <int G, int LOG2CAP, int MINLOG2, bool N32>
__global__ __launch_bounds__(256) void k_cr_agg_group(CoarsenArgs a, const int32_t *__restrict__ segs,
                                                     int32_t ns)
{
    using V = TabV<N32>;
    constexpr int CAP = 1 << LOG2CAP;
    constexpr int GPB = 256 / G;
    __shared__ int32_t s_key[GPB * CAP];
    __shared__ V s_val[GPB * CAP];
    int *s_cnt = nullptr;  // groups narrower than a wave count their uniques in LDS
    if constexpr (G < 64) {
        __shared__ int cnt_lds[GPB];
        s_cnt = cnt_lds;
    }
    const int lg = threadIdx.x % G, grp = threadIdx.x / G;
    int32_t *key = s_key + grp * CAP;
    V *val = s_val + grp * CAP;
    for (int64_t base = (int64_t)blockIdx.x * GPB; base < ns; base += (int64_t)gridDim.x * GPB) {
        const int64_t vi = base + grp;
        const bool active = vi < ns;
        int32_t g = 0, x = 0;
        int64_t b = 0, L = 0;
        int l2 = MINLOG2;
        if (active) {
            g = segs[vi];
            x = a.newid_g[g];
            b = a.sb[g];
            L = a.slen[g];
            if (MINLOG2 < LOG2CAP) l2 = table_log2(L, MINLOG2);
        }
        const int cap = 1 << l2;
        for (int s = lg; s < cap; s += G) {
            key[s] = -1;
            val[s] = 0;
        }
        if constexpr (G < 64)
            if (lg == 0) s_cnt[grp] = 0;
        __syncthreads();
        ull inter = 0, tot = 0;
        if (active) {
            for (int64_t e = b + lg; e < b + L; e += G) {
                int32_t y;
                int64_t w;
                entry_of(a, e, y, w);
                tot += (ull)w;
                if (y == x)
                    inter += (ull)w;
                else
                    lds_insert(key, val, l2, y, w);
            }
        }
        __syncthreads();
        int nun = 0;  // (G == 64) uniques of the wave's group
        if (active) {
            if constexpr (G == 64) {
                const ull below = (1ull << lg) - 1ull;
                for (int s = lg; s < cap; s += 64) {
                    const int32_t y = key[s];
                    const ull m = __ballot(y >= 0);
                    if (y >= 0) {
                        const int pos = nun + __popcll(m & below);
                        a.uy[b + pos] = y;
                        put_w(a.uw, a.uw32, b + pos, (int64_t)val[s]);
                    }
                    nun += __popcll(m);
                }
            } else {
                for (int s = lg; s < cap; s += G) {
                    int32_t y = key[s];
                    if (y >= 0) {
                        int pos = atomicAdd(&s_cnt[grp], 1);
                        a.uy[b + pos] = y;
                        put_w(a.uw, a.uw32, b + pos, (int64_t)val[s]);
                    }
                }
            }
        }
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) {
            inter += __shfl_xor(inter, off, G);
            tot += __shfl_xor(tot, off, G);
        }
        __syncthreads();
        if (active && lg == 0) {
            if constexpr (G == 64)
                a.nuniq[g] = nun;
            else
                a.nuniq[g] = s_cnt[grp];
            a.internal[g] = inter;
            a.ssum[g] = tot;
        }
        __syncthreads();
    }
}

template <int LOG2CAP, int MINLOG2, bool N32>
__global__ __launch_bounds__(256) void k_cr_agg_block(CoarsenArgs a, const int32_t *__restrict__ segs,
                                                     int32_t ns)
{
    constexpr int CAP = 1 << LOG2CAP;
    __shared__ __align__(16) int32_t s_key[CAP];
    __shared__ TabV<N32> s_val[CAP];
    const int tid = threadIdx.x;
    for (int64_t vi = blockIdx.x; vi < ns; vi += gridDim.x) {
        const int32_t g = segs[vi];
        const int32_t x = a.newid_g[g];
        const int64_t b = a.sb[g], L = a.slen[g];
        const int l2 = table_log2(L, MINLOG2);
        const int cap = 1 << l2;
        for (int s = tid; s < cap; s += 256) {
            s_key[s] = -1;
            s_val[s] = 0;
        }
        __syncthreads();
        ull inter = 0, tot = 0;
        for (int64_t e = b + tid; e < b + L; e += 256) {
            int32_t y;
            int64_t w;
            entry_of(a, e, y, w);
            tot += (ull)w;
            if (y == x)
                inter += (ull)w;
            else
                lds_insert(s_key, s_val, l2, y, w);
        }
        __syncthreads();
        const int nun = table_compact<CAP / 256>(s_key, s_val, cap, a.uy + b,
                                                 a.uw32 ? nullptr : a.uw + b,
                                                 a.uw32 ? a.uw32 + b : nullptr);
        inter = (ull)wave_sum((long long)inter);
        tot = (ull)wave_sum((long long)tot);
        ull *red = reinterpret_cast<ull *>(s_key);  // the table is free again
        if ((tid & 63) == 0) {
            red[tid >> 6] = inter;
            red[4 + (tid >> 6)] = tot;
        }
        __syncthreads();
        if (tid == 0) {
            a.nuniq[g] = nun;
            a.internal[g] = red[0] + red[1] + red[2] + red[3];
            a.ssum[g] = red[4] + red[5] + red[6] + red[7];
        }
        __syncthreads();
    }
}

template <typename KernelT>
static void launch_seg(KernelT k, int64_t blocks, hipStream_t s, const CoarsenArgs &a,
                       const int32_t *segs, int32_t ns)
{
    int g = (int)(blocks > (1 << 20) ? (1 << 20) : (blocks < 1 ? 1 : blocks));
    hipLaunchKernelGGL(k, g, 256, 0, s, a, segs, ns);
}

template <bool N32>
static void launch_cr_agg_tier_t(hipStream_t s, int tier, const CoarsenArgs &a,
                                 const int32_t *segs, int32_t ns)
{
    switch (tier) {
    case 1: launch_seg(k_cr_agg_group<4, 3, 3, N32>, (ns + 63) / 64, s, a, segs, ns); break;
    case 2: launch_seg(k_cr_agg_group<8, 4, 4, N32>, (ns + 31) / 32, s, a, segs, ns); break;
    case 3: launch_seg(k_cr_agg_group<16, 5, 5, N32>, (ns + 15) / 16, s, a, segs, ns); break;
    case 4: launch_seg(k_cr_agg_group<32, 6, 6, N32>, (ns + 7) / 8, s, a, segs, ns); break;
    case 5: launch_seg(k_cr_agg_group<64, 7, 7, N32>, (ns + 3) / 4, s, a, segs, ns); break;
    case 6: launch_seg(k_cr_agg_group<64, 8, 8, N32>, (ns + 3) / 4, s, a, segs, ns); break;
    case 7: launch_seg(k_cr_agg_group<64, 9, 9, N32>, (ns + 3) / 4, s, a, segs, ns); break;
    case 8: launch_seg(k_cr_agg_group<64, 10, 10, N32>, (ns + 3) / 4, s, a, segs, ns); break;
    case 9: launch_seg(k_cr_agg_block<11, 11, N32>, ns, s, a, segs, ns); break;
    case 10: launch_seg(k_cr_agg_block<12, 11, N32>, ns, s, a, segs, ns); break;
    default: break;
    }
}

// Narrow levels (every sum below the total degree, < 2^32) aggregate in 4-byte sums.
void launch_cr_agg_tier(hipStream_t s, int tier, const CoarsenArgs &a, const int32_t *segs, int32_t ns)
{
    if (ns <= 0) return;
    if (a.n32)
        launch_cr_agg_tier_t<true>(s, tier, a, segs, ns);
    else
        launch_cr_agg_tier_t<false>(s, tier, a, segs, ns);
}

// ---------------------------------------------------------------------------
// segments of more than 2048 entries (or hub_q): hash partitions of about kHubPartTarget
// entries, each aggregated in LDS (global table when it has more than 4096
// distinct keys); uniques land at poff[p] of buy/buw.

__global__ void k_cr_big_setup(const int32_t *bsegs, int32_t nbs, const int64_t *slen,
                               int32_t target, int chunk_mul, int32_t *bP, int64_t *bP64,
                               int64_t *bch, int64_t *blen, int64_t *bpc, int32_t *nuniq)
{
    GRID_STRIDE(j, nbs)
    {
        nuniq[bsegs[j]] = 0;  // the partitions' aggregation adds into it
        int64_t L = slen[bsegs[j]];
        int64_t P = (L + target - 1) / target;
        if (P < 1) P = 1;
        if (P > kHubMaxParts) P = kHubMaxParts;
        const int64_t cs = chunk_span((int)P, chunk_mul);
        bP[j] = (int32_t)P;
        bP64[j] = P;
        bch[j] = (L + cs - 1) / cs;
        bpc[j] = bch[j] * P;
        blen[j] = L;
    }
}

__global__ __launch_bounds__(256) void k_cr_big_count(CoarsenArgs a, BigSegArgs B)
{
    __shared__ int hist[kHubMaxParts];
    __shared__ ull s_inter[4], s_tot[4];
    const int32_t ci = blockIdx.x;
    const int32_t j = B.chunk_big[ci];
    const int32_t g = B.bsegs[j];
    const int32_t x = a.newid_g[g];
    const int P = B.bP[j];
    const int64_t b = a.sb[g], L = a.slen[g];
    for (int t = threadIdx.x; t < P; t += 256) hist[t] = 0;
    __syncthreads();
    const int64_t cs = chunk_span(P, B.chunk_mul);
    const int64_t t0 = (int64_t)B.chunk_k[ci] * cs;
    const int64_t t1 = min(L, t0 + cs);
    ull inter = 0, tot = 0;
    for (int64_t t = t0 + threadIdx.x; t < t1; t += 256) {
        int32_t y;
        int64_t w0;
        entry_of(a, b + t, y, w0);
        const ull w = (ull)w0;
        tot += w;
        if (y == x)
            inter += w;
        else
            atomicAdd(&hist[hub_part(y, P)], 1);
    }
    inter = (ull)wave_sum((long long)inter);
    tot = (ull)wave_sum((long long)tot);
    if ((threadIdx.x & 63) == 0) {
        s_inter[threadIdx.x >> 6] = inter;
        s_tot[threadIdx.x >> 6] = tot;
    }
    __syncthreads();
    int32_t *ch = B.chist + B.hbase[j] + (int64_t)B.chunk_k[ci] * P;
    for (int t = threadIdx.x; t < P; t += 256) {
        ch[t] = hist[t];  // the scatter reserves the chunk's ranges from these
        if (hist[t]) atomicAdd((ull *)&B.pcount[B.pbase[j] + t], (ull)hist[t]);
    }
    if (threadIdx.x == 0) {
        ull it = s_inter[0] + s_inter[1] + s_inter[2] + s_inter[3];
        if (it) atomicAdd(&a.internal[g], it);
        ull tt = s_tot[0] + s_tot[1] + s_tot[2] + s_tot[3];
        if (tt) atomicAdd(&a.ssum[g], tt);
    }
}

__global__ __launch_bounds__(256) void k_cr_big_scatter(CoarsenArgs a, BigSegArgs B)
{
    __shared__ int hist[kHubMaxParts];
    __shared__ long long lbase[kHubMaxParts];
    const int32_t ci = blockIdx.x;
    const int32_t j = B.chunk_big[ci];
    const int32_t g = B.bsegs[j];
    const int32_t x = a.newid_g[g];
    const int P = B.bP[j];
    const int64_t b = a.sb[g], L = a.slen[g];
    for (int t = threadIdx.x; t < P; t += 256) hist[t] = 0;
    __syncthreads();
    const int64_t cs = chunk_span(P, B.chunk_mul);
    const int64_t t0 = (int64_t)B.chunk_k[ci] * cs;
    const int64_t t1 = min(L, t0 + cs);
    const int32_t *ch = B.chist + B.hbase[j] + (int64_t)B.chunk_k[ci] * P;
    for (int t = threadIdx.x; t < P; t += 256) {
        const int32_t k = ch[t];
        if (k) {  // the run [old - k, old) of the count left in pcount
            const int64_t p = B.pbase[j] + t;
            lbase[t] = B.poff[p] + (long long)atomicAdd((ull *)&B.pcount[p], (ull)(-(int64_t)k)) - k;
        }
    }
    __syncthreads();
    for (int64_t t = t0 + threadIdx.x; t < t1; t += 256) {
        int32_t y;
        int64_t w;
        entry_of(a, b + t, y, w);
        if (y == x) continue;
        int part = hub_part(y, P);
        int64_t pos = lbase[part] + atomicAdd(&hist[part], 1);
        if (B.pe4) {
            B.pe4[pos] = ((uint32_t)y << B.pwb) | (uint32_t)w;
        } else if (B.pe8) {
            PEntry8 en;
            en.c = y;
            en.w = (uint32_t)w;
            B.pe8[pos] = en;
        } else {
            PEntry en;
            en.c = y;
            en.pad = 0;
            en.w = w;
            B.pe[pos] = en;
        }
    }
}

// Writes the occupied slots of a partition's LDS table (as table_compact) into
// the row of its segment g: the block reserves its place with one atomic on
// nuniq[g], so the partitions of a segment fill [sb[g], sb[g] + nuniq[g]) in
// the order they finish.  Returns -1 without reserving or writing when the
// table is full and the partition has more entries than slots (some inserts
// may have found no slot).  Ends with a barrier.
template <int SPT, typename V>
__device__ __forceinline__ int table_to_row(int32_t *key, const V *val, int cap, int64_t entries,
                                            const CoarsenArgs &a, int32_t g)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int32_t ky[SPT];
    V vv[SPT];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        const int s = tid + k * 256;
        ky[k] = s < cap ? key[s] : -1;
        vv[k] = ky[k] >= 0 ? val[s] : (V)0;
        mine += ky[k] >= 0 ? 1 : 0;
    }
    int inc = mine;
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(inc, off, 64);
        if (lane >= off) inc += o;
    }
    __syncthreads();  // the table is read: its first words take the wave totals
    if (lane == 63) key[wave] = inc;
    __syncthreads();
    int pos = inc - mine, total = 0;
    for (int w = 0; w < 4; ++w) {
        const int t = key[w];
        if (w < wave) pos += t;
        total += t;
    }
    if (total == cap && entries > cap) {  // the same on every thread
        __syncthreads();
        return -1;
    }
    if (tid == 0) key[4] = atomicAdd(&a.nuniq[g], total);
    __syncthreads();
    const int64_t base = a.sb[g] + key[4] + pos;
    int k2 = 0;
#pragma unroll
    for (int k = 0; k < SPT; ++k)
        if (ky[k] >= 0) {
            a.uy[base + k2] = ky[k];
            put_w(a.uw, a.uw32, base + k2, (int64_t)vv[k]);
            ++k2;
        }
    __syncthreads();
    return total;
}

template <bool N32, int L2>
__global__ __launch_bounds__(256) void k_cr_big_agg(CoarsenArgs a, BigSegArgs B)
{
    constexpr int CAP = 1 << L2;
    __shared__ __align__(16) int32_t s_key[CAP];
    __shared__ TabV<N32> s_val[CAP];
    const int tid = threadIdx.x;
    const int64_t pi = blockIdx.x;
    const int64_t beg = B.poff[pi], end = B.poff[pi + 1];
    const int32_t g = B.bsegs[B.part_big[pi]];  // (big segments are never stopped groups)
    const int l2 = (end - beg) * 2 >= CAP ? L2 : table_log2(end - beg, 6);
    const int cap = 1 << l2;
    for (int s = tid; s < cap; s += 256) {
        s_key[s] = -1;
        s_val[s] = 0;
    }
    __syncthreads();
    // each thread folds its runs of one neighbour community (a partition can
    // hold 10^5 entries of one) into one insert
    KeyRun run;
    for (int64_t e0 = beg + tid; e0 < end; e0 += 4 * 256) {
        PEntry en[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * 256 < end) en[u] = part_entry(B, e0 + u * 256);
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (e0 + u * 256 < end && run.add(en[u].c, en[u].w))
                (void)lds_insert_bounded(s_key, s_val, l2, run.pk, run.pw);
    }
    if (run.k >= 0) (void)lds_insert_bounded(s_key, s_val, l2, run.k, run.w);
    __syncthreads();
    if (table_to_row<CAP / 256>(s_key, s_val, cap, end - beg, a, g) >= 0) return;
    // An insert that found no slot left the table full, which more than CAP
    // entries can do: then (or when exactly CAP keys filled it) the partition
    // is aggregated in its global table instead; its uniques are counted, their
    // place reserved, then written.
    const int64_t gcap = 2 * (end - beg);
    int32_t *gk = B.fbk + 2 * beg;
    ull *gv = B.fbv + 2 * beg;
    for (int64_t s = tid; s < gcap; s += 256) {
        gk[s] = -1;
        gv[s] = 0;
    }
    __threadfence();
    __syncthreads();
    for (int64_t e = beg + tid; e < end; e += 256) {
        const PEntry en = part_entry(B, e);
        glb_insert_mod(gk, gv, gcap, en.c, en.w);
    }
    __threadfence();
    __syncthreads();
    int mine = 0;
    for (int64_t s = tid; s < gcap; s += 256)
        mine += __hip_atomic_load(&gk[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= 0 ? 1 : 0;
    mine = (int)wave_sum((long long)mine);
    if ((tid & 63) == 0) s_key[tid >> 6] = mine;
    __syncthreads();
    if (tid == 0) {
        s_key[4] = atomicAdd(&a.nuniq[g], s_key[0] + s_key[1] + s_key[2] + s_key[3]);
        s_key[5] = 0;
    }
    __syncthreads();
    const int64_t base = a.sb[g] + s_key[4];
    for (int64_t s = tid; s < gcap; s += 256) {
        int32_t y = __hip_atomic_load(&gk[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (y >= 0) {
            const int pos = atomicAdd(&s_key[5], 1);
            a.uy[base + pos] = y;
            put_w(a.uw, a.uw32, base + pos,
                  (int64_t)__hip_atomic_load(&gv[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        }
    }
}

void launch_cr_big_setup(hipStream_t s, const int32_t *bsegs, int32_t nbs, const int64_t *slen,
                         int32_t target, int chunk_mul, int32_t *bP, int64_t *bP64, int64_t *bch,
                         int64_t *blen, int64_t *bpc, int32_t *nuniq)
{
    if (nbs > 0)
        hipLaunchKernelGGL(k_cr_big_setup, grid_for(nbs, 256), 256, 0, s, bsegs, nbs, slen, target,
                           chunk_mul, bP, bP64, bch, blen, bpc, nuniq);
}

int run_cr_big(Arena &A, hipStream_t s, const CoarsenArgs &a, const BigSegArgs &B, int32_t nbs,
               int64_t nparts, int64_t nchunks)
{
    if (nbs <= 0) return LV_OK;
    fill_i64(s, B.pcount, nparts + 1, 0);
    hipLaunchKernelGGL(k_cr_big_count, (int)nchunks, 256, 0, s, a, B);
    LV_TRY(scan_exclusive_i64(A, s, B.pcount, B.poff, nparts, B.poff + nparts));
    hipLaunchKernelGGL(k_cr_big_scatter, (int)nchunks, 256, 0, s, a, B);
    // the partitions add their uniques into nuniq of their segment (zeroed by
    // the setup) and write them into the segment's row
    if (B.n32 && B.small_tables)  // narrow level, partitions of about 1024 entries
        hipLaunchKernelGGL((k_cr_big_agg<true, kHubAggLog2Small>), (int)nparts, 256, 0, s, a, B);
    else if (B.n32)  // narrow level
        hipLaunchKernelGGL((k_cr_big_agg<true, kHubAggLog2>), (int)nparts, 256, 0, s, a, B);
    else
        hipLaunchKernelGGL((k_cr_big_agg<false, kHubAggLog2>), (int)nparts, 256, 0, s, a, B);
    LV_HIP(hipGetLastError());
    return LV_OK;
}

// ---------------------------------------------------------------------------
// next level: per-vertex k / loop / selfsum and row bounds

// Next-level vertex x = newid_g[g]: its loop, k and row [sb[g], sb[g] + nuniq[g])
// of uy/uw; sums 2m and the level's entries.
__global__ __launch_bounds__(256) void k_cr_vertex(const int32_t *newid_g, const ull *ssum,
                                                  const ull *internal, const ull *selfl,
                                                  const int64_t *sb, const int32_t *nuniq, int32_t K,
                                                  int64_t *loop2, int64_t *self2, int64_t *kdeg2,
                                                  int64_t *rowptr2, int64_t *rowend2,
                                                  long long *msum, long long *nsum)
{
    __shared__ long long s_acc[4], s_nn[4];
    long long acc = 0, nn = 0;
    GRID_STRIDE(g, K)
    {
        const int32_t x = newid_g[g];
        if (x < 0) continue;  // stopped component
        const int64_t lp = (int64_t)((internal[g] + 2 * selfl[g]) / 2);
        loop2[x] = lp;
        self2[x] = lp;
        kdeg2[x] = (int64_t)(ssum[g] - internal[g]) + 4 * lp;
        rowptr2[x] = sb[g];
        rowend2[x] = sb[g] + nuniq[g];
        acc += (long long)(ssum[g] + 2 * selfl[g]);
        nn += nuniq[g];
    }
    acc = wave_sum(acc);  // the level's total is below 2^61: no wrap
    nn = wave_sum(nn);
    if ((threadIdx.x & 63) == 0) {
        s_acc[threadIdx.x >> 6] = acc;
        s_nn[threadIdx.x >> 6] = nn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // one atomic per block on each shared total
        const long long t = s_acc[0] + s_acc[1] + s_acc[2] + s_acc[3];
        if (t) atomicAdd((ull *)msum, (ull)t);
        const long long u = s_nn[0] + s_nn[1] + s_nn[2] + s_nn[3];
        if (u) atomicAdd((ull *)nsum, (ull)u);
    }
}

void launch_cr_vertex(hipStream_t s, const int32_t *newid_g, const ull *ssum, const ull *internal,
                      const ull *selfl, const int64_t *sb, const int32_t *nuniq, int32_t K,
                      int64_t *loop2, int64_t *self2, int64_t *kdeg2, int64_t *rowptr2,
                      int64_t *rowend2, long long *msum, long long *nsum)
{
    if (K > 0)
        hipLaunchKernelGGL(k_cr_vertex, grid_for(K, 256, 4096), 256, 0, s, newid_g, ssum, internal,
                           selfl, sb, nuniq, K, loop2, self2, kdeg2, rowptr2, rowend2, msum, nsum);
}

}  // namespace lv
