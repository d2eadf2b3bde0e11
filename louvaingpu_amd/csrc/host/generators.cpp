// generators.cpp - deterministic synthetic inputs for the benchmark
// configurations (the reference ships no generator).
//
// PRNG: SplitMix64 (Steele, Lea, Flood 2014) derives one seed per stream;
// each stream is xoshiro256** (Blackman, Vigna 2018).  Streams are per SBM row
// and per block of 2^20 R-MAT draws, so results do not depend on threading.
//
// SBM: planted partition, block(v) = v / block_size; every pair u < v is an
//   edge with p_in (same block) or p_out, sampled per row with geometric
//   skipping (Batagelj, Brandes 2005).  Edges come out canonical (u < v) in
//   lexicographic order, unit weights.
// R-MAT: edge_factor * 2^scale draws; each of `scale` bit levels picks a
//   quadrant with probabilities a, b, c, 1-a-b-c (Chakrabarti, Zhan, Faloutsos
//   2004), no noise.  Draws are canonicalised to (min, max); duplicates and
//   self-loops are merged by summing weights: int32 weights count the draws,
//   float32 weights sum per-draw U[1,10] values in FP64 and store the sum as
//   float32.  Output is lexicographically sorted.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro256 {
    uint64_t s[4];
    Xoshiro256(uint64_t seed, uint64_t stream)
    {
        uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (stream + 1));
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    uint64_t next()
    {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    // uniform in (0, 1]
    double open0() { return (double)((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }
    // uniform in [0, 1)
    double closed0() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

int num_threads()
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Appends v in [lo, hi) where each is taken with probability p.
template <typename F>
void skip_sample(Xoshiro256 &rng, int64_t lo, int64_t hi, double p, F emit)
{
    if (lo >= hi || p <= 0.0) return;
    if (p >= 1.0) {
        for (int64_t v = lo; v < hi; ++v) emit(v);
        return;
    }
    const double lq = std::log1p(-p);
    int64_t v = lo - 1;
    for (;;) {
        double r = rng.open0();
        double sk = std::floor(std::log(r) / lq);
        if (sk >= (double)(hi - v)) break;
        v += 1 + (int64_t)sk;
        if (v >= hi) break;
        emit(v);
    }
}

// Stable LSD radix sort of 64-bit keys (11-bit digits) carrying a payload.
template <typename V>
void radix_sort_pairs(std::vector<uint64_t> &key, std::vector<V> &val, int bits)
{
    const size_t n = key.size();
    std::vector<uint64_t> k2(n);
    std::vector<V> v2(n);
    const int D = 11, R = 1 << D;
    const int T = std::max(1, std::min(num_threads(), 64));
    std::vector<size_t> cnt((size_t)T * R);
    for (int shift = 0; shift < bits; shift += D) {
        std::fill(cnt.begin(), cnt.end(), 0);
#pragma omp parallel num_threads(T)
        {
            int t = 0;
#ifdef _OPENMP
            t = omp_get_thread_num();
#endif
            size_t lo = n * (size_t)t / (size_t)T, hi = n * (size_t)(t + 1) / (size_t)T;
            size_t *c = &cnt[(size_t)t * R];
            for (size_t i = lo; i < hi; ++i) c[(key[i] >> shift) & (R - 1)]++;
#pragma omp barrier
#pragma omp single
            {
                size_t run = 0;
                for (int d = 0; d < R; ++d)
                    for (int u = 0; u < T; ++u) {
                        size_t x = cnt[(size_t)u * R + d];
                        cnt[(size_t)u * R + d] = run;
                        run += x;
                    }
            }
            for (size_t i = lo; i < hi; ++i) {
                size_t p = c[(key[i] >> shift) & (R - 1)]++;
                k2[p] = key[i];
                v2[p] = val[i];
            }
        }
        key.swap(k2);
        val.swap(v2);
    }
}

}  // namespace

struct lvg_graph {
    int32_t n = 0;
    int wtype = 0;  // 0 int32, 1 float32
    std::vector<int32_t> src, dst;
    std::vector<int32_t> wi;
    std::vector<float> wf;
};

extern "C" {

lvg_graph *lvg_sbm(int32_t n, int32_t block_size, double p_in, double p_out, uint64_t seed)
{
    if (n < 0 || block_size <= 0) return nullptr;
    lvg_graph *g = new (std::nothrow) lvg_graph();
    if (!g) return nullptr;
    g->n = n;
    g->wtype = 0;
    const int T = std::max(1, std::min(num_threads(), 64));
    std::vector<std::vector<int32_t>> ps((size_t)T), pd((size_t)T);
#pragma omp parallel num_threads(T)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        int64_t lo = (int64_t)n * t / T, hi = (int64_t)n * (t + 1) / T;
        for (int64_t u = lo; u < hi; ++u) {
            Xoshiro256 rng(seed, (uint64_t)u);
            int64_t bend = std::min<int64_t>(((u / block_size) + 1) * (int64_t)block_size, n);
            auto emit = [&](int64_t v) {
                ps[(size_t)t].push_back((int32_t)u);
                pd[(size_t)t].push_back((int32_t)v);
            };
            skip_sample(rng, u + 1, bend, p_in, emit);
            skip_sample(rng, bend, n, p_out, emit);
        }
    }
    size_t total = 0;
    for (int t = 0; t < T; ++t) total += ps[(size_t)t].size();
    g->src.reserve(total);
    g->dst.reserve(total);
    for (int t = 0; t < T; ++t) {
        g->src.insert(g->src.end(), ps[(size_t)t].begin(), ps[(size_t)t].end());
        g->dst.insert(g->dst.end(), pd[(size_t)t].begin(), pd[(size_t)t].end());
    }
    g->wi.assign(total, 1);
    return g;
}

lvg_graph *lvg_rmat(int32_t scale, int64_t edge_factor, double a, double b, double c,
                    uint64_t seed, int wtype)
{
    if (scale < 1 || scale > 30 || edge_factor < 0 || (wtype != 0 && wtype != 1)) return nullptr;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return nullptr;
    lvg_graph *g = new (std::nothrow) lvg_graph();
    if (!g) return nullptr;
    const int64_t n = int64_t(1) << scale;
    const int64_t draws = edge_factor * n;
    g->n = (int32_t)n;
    g->wtype = wtype;
    const double two53 = 9007199254740992.0;
    const uint64_t ta = (uint64_t)(a * two53), tb = (uint64_t)((a + b) * two53),
                   tc = (uint64_t)((a + b + c) * two53);
    std::vector<uint64_t> key((size_t)draws);
    std::vector<float> w(wtype == 1 ? (size_t)draws : 0);
    const int64_t chunk = int64_t(1) << 20;
    const int64_t nchunks = (draws + chunk - 1) / chunk;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t ch = 0; ch < nchunks; ++ch) {
        Xoshiro256 rng(seed, (uint64_t)ch);
        int64_t lo = ch * chunk, hi = std::min(draws, lo + chunk);
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t u = 0, v = 0;
            for (int lev = 0; lev < scale; ++lev) {
                uint64_t r = rng.next() >> 11;
                uint64_t bu = 0, bv = 0;
                if (r < ta) {
                } else if (r < tb) {
                    bv = 1;
                } else if (r < tc) {
                    bu = 1;
                } else {
                    bu = 1;
                    bv = 1;
                }
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            if (u > v) std::swap(u, v);
            key[(size_t)i] = (u << scale) | v;
            if (wtype == 1) w[(size_t)i] = (float)(1.0 + 9.0 * rng.closed0());
        }
    }
    // sort by (u, v): v occupies bits [0, scale), u bits [scale, 2*scale)
    if (wtype == 1) {
        radix_sort_pairs(key, w, 2 * scale);
    } else {
        std::vector<uint8_t> dummy(key.size());
        radix_sort_pairs(key, dummy, 2 * scale);
    }
    size_t i = 0, nk = key.size();
    while (i < nk) {
        size_t j = i;
        double sum = 0.0;
        while (j < nk && key[j] == key[i]) {
            if (wtype == 1) sum += (double)w[j];
            ++j;
        }
        g->src.push_back((int32_t)(key[i] >> scale));
        g->dst.push_back((int32_t)(key[i] & (uint64_t)(n - 1)));
        if (wtype == 1)
            g->wf.push_back((float)sum);
        else
            g->wi.push_back((int32_t)(j - i));
        i = j;
    }
    return g;
}

int32_t lvg_num_nodes(const lvg_graph *g) { return g ? g->n : -1; }
int64_t lvg_num_edges(const lvg_graph *g) { return g ? (int64_t)g->src.size() : -1; }
int lvg_wtype(const lvg_graph *g) { return g ? g->wtype : -1; }

int lvg_copy(const lvg_graph *g, int32_t *src, int32_t *dst, void *w)
{
    if (!g || !src || !dst || !w) return -4;
    size_t e = g->src.size();
    std::memcpy(src, g->src.data(), e * sizeof(int32_t));
    std::memcpy(dst, g->dst.data(), e * sizeof(int32_t));
    if (g->wtype == 1)
        std::memcpy(w, g->wf.data(), e * sizeof(float));
    else
        std::memcpy(w, g->wi.data(), e * sizeof(int32_t));
    return 0;
}

void lvg_free(lvg_graph *g) { delete g; }

}  // extern "C"
