// batch.hip - graph sets (gfx950): many independent graphs, each possibly at
// several resolutions, run as the components of one union graph.
//
// Component c is graph c / nres at resolution c % nres.  Its vertices form one
// contiguous id range of every level's union graph and no edge crosses two
// components, so the sweep, renumbering and coarsening kernels of the single
// graph serve the set unchanged; what is per component is 2m, the resolution,
// the pass mode (sweep / last pass / frozen), the exact Q aggregates and the
// level-loop decision.  Renumbering the union by first occurrence numbers each
// component's communities consecutively, in component order, which is every
// component's own first-occurrence numbering shifted by an offset (koff).
#include "device.hpp"

namespace lv {

// Per-component sums over data grouped by component (vertices, edges,
// communities) reach their component's word once per block, not once per 64
// items: device atomics on one word serialise (~11 ns each), and a component
// of 10^6 vertices would otherwise take 1.6e4 of them per pass.  Each wave
// takes a contiguous range of 64-item rounds (the waves of a block take
// consecutive ranges), carries its current run across rounds, and the block
// merges its waves' last runs in LDS before adding them.
static_assert(sizeof(QAcc) == 4 * sizeof(ull), "block_flush indexes QAcc as four words");

__device__ __forceinline__ void wave_rounds(int64_t items, int64_t &r0, int64_t &r1)
{
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t wv = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t rounds = (items + 63) >> 6;
    const int64_t per = (rounds + waves - 1) / waves;
    r0 = min(rounds, wv * per);
    r1 = min(rounds, r0 + per);
}

// The last runs of a block's waves (keys ascending by wave, LLONG_MIN for a
// wave with none) merged and added: one atomic per distinct key.  Needs
// blockDim.x / 64 <= 16 and every thread of the block.
template <int NV, bool U128>
__device__ __forceinline__ void block_flush(long long key, const ull (&v)[NV], ull *const (&dst)[NV])
{
    __shared__ long long s_key[16];
    __shared__ ull s_v[16][NV];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0) {
        s_key[wave] = key;
        for (int i = 0; i < NV; ++i) s_v[wave][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long k = LLONG_MIN;
        ull acc[NV] = {};
        for (int w = 0; w <= nw; ++w) {
            const long long kw = w < nw ? s_key[w] : LLONG_MIN;
            if (w < nw && kw == LLONG_MIN) continue;
            if (w < nw && kw == k) {
                if constexpr (U128) {  // (sum, 128-bit lo/hi)
                    acc[0] += s_v[w][0];
                    const ull lo = acc[1] + s_v[w][1];
                    acc[2] += s_v[w][2] + (lo < acc[1] ? 1ull : 0ull);
                    acc[1] = lo;
                } else {
                    for (int i = 0; i < NV; ++i) acc[i] += s_v[w][i];
                }
                continue;
            }
            if (k != LLONG_MIN) {
                if constexpr (U128) {
                    if (acc[0]) atomicAdd(dst[0] + 4 * k, acc[0]);
                    if (acc[1] | acc[2]) atomic_add_u128(dst[1] + 4 * k, dst[2] + 4 * k, acc[1], acc[2]);
                } else {
                    for (int i = 0; i < NV; ++i)
                        if (acc[i]) atomicAdd(dst[i] + k, acc[i]);
                }
            }
            if (w < nw) {
                k = kw;
                for (int i = 0; i < NV; ++i) acc[i] = s_v[w][i];
            }
        }
    }
}

// SegCarry's run of a wave added through block_flush.
__device__ __forceinline__ void seg_block_flush(SegCarry &sc, ull *base)
{
    const ull v[1] = {sc.key != LLONG_MIN ? sc.sum : 0ull};
    ull *const d[1] = {base};
    block_flush<1, false>(sc.key, v, d);
}

// Unit exponent of every graph of the set: gscale[g] = the smallest s making
// all of graph g's weights integral in units of 2^-s (so every component
// computes in the units a single-graph run of its graph would use); negative,
// infinite or NaN weights flag invalid input.
__global__ void k_b_graph_exponent(const uint32_t *bw, int wtype, const int64_t *beoff,
                                   int32_t ngraphs, int64_t eb, int *gscale, DevFlags *f)
{
    int bad = 0;
    const int lane = threadIdx.x & 63;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < eb;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = base + threadIdx.x;
        const bool valid = i < eb;
        int32_t g = INT_MIN;
        int sc = 0;
        if (valid) {
            int32_t lo = 0, hi = ngraphs;  // beoff[lo] <= i < beoff[hi]
            while (hi - lo > 1) {
                const int32_t mid = (lo + hi) >> 1;
                if (beoff[mid] <= i) lo = mid; else hi = mid;
            }
            g = lo;
            if (wtype == LV_WEIGHT_INT32) {
                if ((int32_t)bw[i] < 0) bad |= 1;
            } else {
                sc = f32_unit_exponent(bw[i]);
                if (sc < 0) {
                    bad |= 1;
                    sc = 0;
                }
            }
        }
        // runs of equal graphs (edges are grouped by graph): max, one atomic per run
        for (int off = 1; off < 64; off <<= 1) {
            const int os = __shfl_up(sc, off, 64);
            const int32_t og = __shfl_up(g, off, 64);
            if (lane >= off && og == g) sc = max(sc, os);
        }
        const int32_t ng = __shfl_down(g, 1, 64);
        // only a run that raises its graph's maximum takes the atomic: every
        // wave of a large graph would otherwise serialise on one word
        if (valid && (lane == 63 || ng != g) && sc > 0 &&
            sc > __hip_atomic_load(&gscale[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&gscale[g], sc);
    }
    bad = wave_or_i(bad);
    if (lane == 0 && bad) atomicOr(&f->err, 1);
}

// Union edge list: union edge j of component c is edge j - ueoff[c] of graph
// g = c / nres, with endpoints shifted by uvoff[c] and its weight in graph g's
// units; endpoints outside [0, n_g) flag invalid input, weights beyond the
// exact range flag a range error.
__global__ __launch_bounds__(256) void k_b_union_edges(const int32_t *bsrc, const int32_t *bdst, const uint32_t *bw,
                                int wtype, const int *gscale, const int64_t *beoff,
                                const int32_t *bn, const int64_t *ueoff, const int64_t *uvoff,
                                int32_t ncomp, int32_t nres, int64_t etot, int32_t *usrc,
                                int32_t *udst, int64_t *uw, DevFlags *f)
{
    int bad = 0;
    ull tot = 0, umax = 0;
    GRID_STRIDE(j, etot)
    {
        int32_t lo = 0, hi = ncomp;  // ueoff[lo] <= j < ueoff[hi]
        while (hi - lo > 1) {
            const int32_t mid = (lo + hi) >> 1;
            if (ueoff[mid] <= j) lo = mid; else hi = mid;
        }
        const int32_t g = lo / nres;
        const int64_t i = beoff[g] + (j - ueoff[lo]);
        const int32_t a = bsrc[i], b = bdst[i], ng = bn[g];
        if (a < 0 || a >= ng || b < 0 || b >= ng) bad |= 1;
        usrc[j] = a + (int32_t)uvoff[lo];
        udst[j] = b + (int32_t)uvoff[lo];
        const int64_t u = weight_units(bw[i], wtype, gscale[g], bad);
        uw[j] = u;
        umax = max(umax, (ull)u);
        tot = sat_add(tot, (ull)(u == 0 ? 1 : u));
    }
    block_units_flush(tot, 0, umax, bad, f);
}

// The level-0 CSR of a graph set is built once over the graphs (vertex ranges
// bvoff, each graph's entries contiguous from geoff[g]) and then written out for
// every resolution: component c = g * nres + r holds graph g's rows, its
// vertices at nres * bvoff[g] + r * n_g and its entries at nres * geoff[g] +
// r * nnz_g, columns moved by the same vertex offset.
__device__ __forceinline__ int32_t graph_of(const int64_t *bvoff, int32_t ngraphs, int64_t v)
{
    int32_t lo = 0, hi = ngraphs;  // bvoff[lo] <= v < bvoff[hi]
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (bvoff[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void k_b_graph_entry_off(const int64_t *bvoff, const int64_t *rowptr, int32_t ngraphs,
                                    int64_t *geoff)
{
    GRID_STRIDE(g, (int64_t)ngraphs + 1) geoff[g] = rowptr[bvoff[g]];
}

__global__ __launch_bounds__(256) void k_b_replicate_entries(
    const int64_t *bvoff, const int64_t *geoff, int32_t ngraphs, int32_t nres, int64_t nnz,
    const int32_t *__restrict__ col, const int64_t *__restrict__ w, const uint32_t *__restrict__ w32,
    int32_t *__restrict__ ucol, int64_t *__restrict__ uw, uint32_t *__restrict__ uw32)
{
    GRID_STRIDE(q, nnz)
    {
        const int32_t y = col[q];
        const int32_t g = graph_of(bvoff, ngraphs, y);  // rows and columns share their graph
        const int64_t vb = bvoff[g], ng = bvoff[g + 1] - vb;
        const int64_t eg0 = geoff[g], eg = geoff[g + 1] - eg0;
        const int64_t p0 = (int64_t)nres * eg0 + (q - eg0);
        const int32_t y0 = (int32_t)((int64_t)nres * vb + (y - vb));
        if (w32) {
            const uint32_t x = w32[q];
            for (int32_t r = 0; r < nres; ++r) {
                ucol[p0 + r * eg] = y0 + (int32_t)(r * ng);
                uw32[p0 + r * eg] = x;
            }
        } else {
            const int64_t x = w[q];
            for (int32_t r = 0; r < nres; ++r) {
                ucol[p0 + r * eg] = y0 + (int32_t)(r * ng);
                uw[p0 + r * eg] = x;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_b_replicate_vertices(
    const int64_t *bvoff, const int64_t *geoff, int32_t ngraphs, int32_t nres, int32_t nb,
    const int64_t *rowptr, const int64_t *kdeg, const int64_t *loopw, const int64_t *selfsum,
    int64_t *urowptr, int64_t *ukdeg, int64_t *uloopw, int64_t *uselfsum, int64_t unnz)
{
    GRID_STRIDE(i, nb)
    {
        const int32_t g = graph_of(bvoff, ngraphs, i);
        const int64_t vb = bvoff[g], ng = bvoff[g + 1] - vb;
        const int64_t eg0 = geoff[g], eg = geoff[g + 1] - eg0;
        const int64_t v0 = (int64_t)nres * vb + (i - vb);
        const int64_t p0 = (int64_t)nres * eg0 + (rowptr[i] - eg0);
        const int64_t k = kdeg[i], lp = loopw[i], ss = selfsum[i];
        for (int32_t r = 0; r < nres; ++r) {
            const int64_t v = v0 + r * ng;
            urowptr[v] = p0 + r * eg;
            ukdeg[v] = k;
            uloopw[v] = lp;
            uselfsum[v] = ss;
        }
        if (i == 0) urowptr[(int64_t)nres * nb] = unnz;
    }
}

// cm[c] += sum of w over the union edges of component c (edge ranges ueoff).
// cm[c] = sum of w over component c's edges.
__global__ void k_b_edge_sum(const int64_t *w, const int64_t *ueoff, int32_t ncomp, int64_t etot,
                             ull *cm)
{
    int64_t r0, r1;
    wave_rounds(etot, r0, r1);
    SegCarry sc;
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t j = (r << 6) + (threadIdx.x & 63);
        const bool valid = j < etot;
        long long key = 0;
        ull v = 0;
        if (valid) {
            int32_t lo = 0, hi = ncomp;
            while (hi - lo > 1) {
                const int32_t mid = (lo + hi) >> 1;
                if (ueoff[mid] <= j) lo = mid; else hi = mid;
            }
            key = lo;
            v = (ull)w[j];
        }
        sc.add(valid, key, v, cm);
    }
    seg_block_flush(sc, cm);
}

// Per component Q aggregates of the community state (sum of internals, sum of
// deg^2); with `apply` the pass deltas are added first (k_apply per component).
__global__ __launch_bounds__(256) void k_b_qagg(int32_t n, int64_t *deg, const int64_t *in,
                                               int32_t *size, ull *ddeg, int32_t *dsize,
                                               int size_shift, uint32_t *deg32,
                                               const int32_t *vcomp, int apply, QAcc *cq)
{
    const int lane = threadIdx.x & 63;
    int64_t r0, r1;
    wave_rounds(n, r0, r1);
    long long ckey = LLONG_MIN;  // the wave's carried run (as lane 63 holds it)
    long long csin = 0;
    ull clo = 0, chi = 0;
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t c = (r << 6) + lane;
        const bool valid = c < n;
        long long k = LLONG_MIN, sin = 0;
        ull lo = 0, hi = 0;
        if (valid) {
            const int64_t dv = apply ? apply_deltas(c, deg, size, ddeg, dsize, size_shift, deg32) : deg[c];
            k = vcomp[c];
            sin = in[c];
            acc_sq(lo, hi, (ull)dv);
        }
        if (lane == 0 && ckey != LLONG_MIN) {  // the carried run joins lane 0's or is added
            if (k == ckey) {
                sin += csin;
                const ull l2 = lo + clo;
                hi += chi + (l2 < lo ? 1ull : 0ull);
                lo = l2;
            } else {
                if (csin) atomicAdd((ull *)&cq[ckey].sum_in, (ull)csin);
                if (clo | chi) atomic_add_u128(&cq[ckey].d2_lo, &cq[ckey].d2_hi, clo, chi);
            }
        }
        for (int off = 1; off < 64; off <<= 1) {
            const long long os = __shfl_up(sin, off, 64);
            const ull ol = __shfl_up(lo, off, 64), oh = __shfl_up(hi, off, 64);
            const long long ok = __shfl_up(k, off, 64);
            if (lane >= off && ok == k) {
                sin += os;
                lo += ol;
                hi += oh + (lo < ol ? 1ull : 0ull);
            }
        }
        const long long nk = __shfl_down(k, 1, 64);
        if (valid && lane != 63 && nk != k) {  // a run that ends inside the round
            if (sin) atomicAdd((ull *)&cq[k].sum_in, (ull)sin);
            if (lo | hi) atomic_add_u128(&cq[k].d2_lo, &cq[k].d2_hi, lo, hi);
        }
        ckey = __shfl(k, 63, 64);
        csin = __shfl(sin, 63, 64);
        clo = __shfl(lo, 63, 64);
        chi = __shfl(hi, 63, 64);
    }
    const ull v[3] = {(ull)csin, clo, chi};
    ull *const d[3] = {(ull *)&cq[0].sum_in, &cq[0].d2_lo, &cq[0].d2_hi};
    block_flush<3, true>(ckey, v, d);
}

// Communities (renumbering groups) per component: group g's component is its
// head's; groups are in community order, so components ascend.
__global__ void k_b_groups(const int32_t *gstart, const int32_t *perm, const int32_t *vcomp,
                           int32_t K, ull *kcnt)
{
    int64_t r0, r1;
    wave_rounds(K, r0, r1);
    SegCarry sc;
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t g = (r << 6) + (threadIdx.x & 63);
        const bool valid = g < K;
        const long long key = valid ? vcomp[perm[gstart[g]]] : 0;
        sc.add(valid, key, 1ull, kcnt);
    }
    seg_block_flush(sc, kcnt);
}

// Next-level vertex id of every group: its community's number within its
// component plus the component's first vertex at the next level, or -1 when
// the component stops at this level.
__global__ void k_b_next_ids(const int32_t *gstart, const int32_t *perm, const int32_t *vcomp,
                             const int32_t *newid_g, int32_t K, const int64_t *koff,
                             const int64_t *vnext, int32_t *newid_next)
{
    GRID_STRIDE(g, K)
    {
        const int32_t comp = vcomp[perm[gstart[g]]];
        const int64_t vn = vnext[comp];
        newid_next[g] = vn < 0 ? -1 : (int32_t)(newid_g[g] - koff[comp] + vn);
    }
}

// cm[comp] += kdeg - 2*loop over the vertices of each component (= sum over the
// component's communities of their entry weights plus twice their members'
// self-loops, i.e. 2m of the coarsened component).
__global__ void k_b_vertex_sum(const int64_t *kdeg, const int64_t *loopw, const int32_t *vcomp,
                               int32_t n, ull *cm)
{
    int64_t r0, r1;
    wave_rounds(n, r0, r1);
    SegCarry sc;
    for (int64_t r = r0; r < r1; ++r) {
        const int64_t x = (r << 6) + (threadIdx.x & 63);
        const bool valid = x < n;
        long long key = 0;
        ull v = 0;
        if (valid) {
            key = vcomp[x];
            v = (ull)(kdeg[x] - 2 * loopw[x]);
        }
        sc.add(valid, key, v, cm);
    }
    seg_block_flush(sc, cm);
}

// One level of the flattening: fin holds each level-0 vertex's id at level l
// (a union vertex id); components with more than l+1 levels move on to their
// next-level vertex, the others keep their community number at level l.
__global__ void k_b_flatten(int32_t *fin, int32_t n0, const int32_t *vcomp0, const int32_t *clev,
                            int32_t l, const int32_t *newc_l, const int64_t *koff_l,
                            const int64_t *vnext_l)
{
    GRID_STRIDE(v, n0)
    {
        const int32_t comp = vcomp0[v];
        const int32_t L = clev[comp];
        if (L <= l) continue;
        const int64_t local = (int64_t)newc_l[fin[v]] - koff_l[comp];
        fin[v] = (int32_t)(L > l + 1 ? local + vnext_l[comp] : local);
    }
}

void launch_b_graph_exponent(hipStream_t s, const uint32_t *bw, int wtype, const int64_t *beoff,
                             int32_t ngraphs, int64_t eb, int *gscale, DevFlags *f)
{
    if (eb > 0)
        hipLaunchKernelGGL(k_b_graph_exponent, grid_for(eb, 256, 1 << 20), 256, 0, s, bw, wtype,
                           beoff, ngraphs, eb, gscale, f);
}

void launch_b_union_edges(hipStream_t s, const int32_t *bsrc, const int32_t *bdst,
                          const uint32_t *bw, int wtype, const int *gscale, const int64_t *beoff,
                          const int32_t *bn, const int64_t *ueoff, const int64_t *uvoff,
                          int32_t ncomp, int32_t nres, int64_t etot, int32_t *usrc, int32_t *udst,
                          int64_t *uw, DevFlags *f)
{
    if (etot > 0)
        hipLaunchKernelGGL(k_b_union_edges, grid_for(etot, 256, 8192), 256, 0, s, bsrc, bdst, bw,
                           wtype, gscale, beoff, bn, ueoff, uvoff, ncomp, nres, etot, usrc, udst,
                           uw, f);
}

void launch_b_replicate_csr(hipStream_t s, const int64_t *bvoff, int64_t *geoff, int32_t ngraphs,
                            int32_t nres, int32_t nb, int64_t nnz, const int64_t *rowptr,
                            const int32_t *col, const int64_t *w, const uint32_t *w32,
                            const int64_t *kdeg, const int64_t *loopw, const int64_t *selfsum,
                            int64_t *urowptr, int32_t *ucol, int64_t *uw, uint32_t *uw32,
                            int64_t *ukdeg, int64_t *uloopw, int64_t *uselfsum)
{
    hipLaunchKernelGGL(k_b_graph_entry_off, grid_for((int64_t)ngraphs + 1, 256), 256, 0, s, bvoff,
                       rowptr, ngraphs, geoff);
    if (nnz > 0)
        hipLaunchKernelGGL(k_b_replicate_entries, grid_for(nnz, 256, 1 << 16), 256, 0, s, bvoff,
                           geoff, ngraphs, nres, nnz, col, w, w32, ucol, uw, uw32);
    if (nb > 0)
        hipLaunchKernelGGL(k_b_replicate_vertices, grid_for(nb, 256, 1 << 16), 256, 0, s, bvoff,
                           geoff, ngraphs, nres, nb, rowptr, kdeg, loopw, selfsum, urowptr, ukdeg,
                           uloopw, uselfsum, (int64_t)nres * nnz);
}

void launch_b_edge_sum(hipStream_t s, const int64_t *w, const int64_t *ueoff, int32_t ncomp,
                       int64_t etot, ull *cm)
{
    if (etot > 0)
        hipLaunchKernelGGL(k_b_edge_sum, grid_for(etot, 256, 1024), 256, 0, s, w, ueoff, ncomp,
                           etot, cm);
}

void launch_b_qagg(hipStream_t s, int32_t n, int64_t *deg, const int64_t *in, int32_t *size,
                   ull *ddeg, int32_t *dsize, int size_shift, uint32_t *deg32,
                   const int32_t *vcomp, bool apply, QAcc *cq)
{
    if (n > 0)
        hipLaunchKernelGGL(k_b_qagg, grid_for(n, 256, 1024), 256, 0, s, n, deg, in, size, ddeg,
                           dsize, size_shift, deg32, vcomp, apply ? 1 : 0, cq);
}

void launch_b_groups(hipStream_t s, const int32_t *gstart, const int32_t *perm,
                     const int32_t *vcomp, int32_t K, ull *kcnt)
{
    if (K > 0)
        hipLaunchKernelGGL(k_b_groups, grid_for(K, 256, 1024), 256, 0, s, gstart, perm, vcomp,
                           K, kcnt);
}

void launch_b_next_ids(hipStream_t s, const int32_t *gstart, const int32_t *perm,
                       const int32_t *vcomp, const int32_t *newid_g, int32_t K,
                       const int64_t *koff, const int64_t *vnext, int32_t *newid_next)
{
    if (K > 0)
        hipLaunchKernelGGL(k_b_next_ids, grid_for(K, 256), 256, 0, s, gstart, perm, vcomp, newid_g,
                           K, koff, vnext, newid_next);
}

void launch_b_vertex_sum(hipStream_t s, const int64_t *kdeg, const int64_t *loopw,
                         const int32_t *vcomp, int32_t n, ull *cm)
{
    if (n > 0)
        hipLaunchKernelGGL(k_b_vertex_sum, grid_for(n, 256, 1024), 256, 0, s, kdeg, loopw, vcomp,
                           n, cm);
}

void launch_b_flatten(hipStream_t s, int32_t *fin, int32_t n0, const int32_t *vcomp0,
                      const int32_t *clev, int32_t l, const int32_t *newc_l, const int64_t *koff_l,
                      const int64_t *vnext_l)
{
    if (n0 > 0)
        hipLaunchKernelGGL(k_b_flatten, grid_for(n0, 256), 256, 0, s, fin, n0, vcomp0, clev, l,
                           newc_l, koff_l, vnext_l);
}

}  // namespace lv
