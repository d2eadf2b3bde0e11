// CPU check of the upload's host encoders (louvaingpu_amd/csrc/engine/encode.cpp),
// built and run by test_encode.py: the AVX2 and scalar forms against each other
// and against plain restatements (offsets by counting, packing by shifts), on
// sorted sources with skewed runs and empty rows, one huge row, sparse sources,
// and corrupted lists, cut into 1 to 17 parts; weights at every width with
// escapes and escape caps; the width choice on drawn distributions.
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../louvaingpu_amd/csrc/engine/encode.hpp"

using namespace lv::enc;

static int fails = 0;
#define CHECK(cond, ...)                      \
    do {                                      \
        if (!(cond)) {                        \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");                \
            ++fails;                          \
        }                                     \
    } while (0)

// off[v] = number of sources below v (the definition, by counting)
static std::vector<int64_t> offsets_ref(const std::vector<int32_t> &src, int32_t n)
{
    std::vector<int64_t> cnt((size_t)n + 1, 0), off((size_t)n + 1, 0);
    for (int32_t v : src) ++cnt[(size_t)v];
    for (int32_t v = 0; v < n; ++v) off[(size_t)v + 1] = off[(size_t)v] + cnt[(size_t)v];
    return off;
}

template <typename O>
static bool run_parts(const std::vector<int32_t> &src, int32_t n, int T, Impl impl,
                      std::vector<O> &off)
{
    const int64_t e = (int64_t)src.size();
    off.assign((size_t)n + 1, (O)-7);
    bool ok = true;
    for (int t = 0; t < T; ++t) {
        const int64_t a = e * t / T, b = e * (t + 1) / T;
        if (a < b) ok = offsets_part(src.data(), a, b, e, n, off.data(), impl) && ok;
    }
    return ok;
}

static void check_offsets(const char *name, const std::vector<int32_t> &src, int32_t n, bool valid)
{
    const std::vector<int64_t> ref = valid ? offsets_ref(src, n) : std::vector<int64_t>();
    for (int T : {1, 2, 3, 7, 16, 17}) {
        std::vector<int32_t> o32s, o32v;
        std::vector<int64_t> o64v;
        const bool s = run_parts(src, n, T, Impl::Scalar, o32s);
        const bool v = run_parts(src, n, T, Impl::Avx2, o32v);
        const bool w = run_parts(src, n, T, Impl::Avx2, o64v);
        CHECK(s == valid && v == valid && w == valid, "%s T=%d: accepted %d %d %d, expected %d",
              name, T, s, v, w, valid);
        if (!valid || !s || !v || !w) continue;
        bool same = true;
        for (size_t i = 0; i < ref.size(); ++i)
            same = same && o32s[i] == ref[i] && o32v[i] == ref[i] && o64v[i] == ref[i];
        CHECK(same, "%s T=%d: offsets differ from counting", name, T);
    }
}

static std::vector<int32_t> skewed_sorted(std::mt19937_64 &rng, int32_t n, int64_t e)
{
    // sources with geometric run lengths, gaps of empty rows and a few huge runs
    std::vector<int32_t> src;
    src.reserve((size_t)e);
    int32_t v = (int32_t)(rng() % 5);
    while ((int64_t)src.size() < e && v < n) {
        int64_t run = 1 + (int64_t)(std::geometric_distribution<int>(0.05)(rng));
        if (rng() % 97 == 0) run *= 200;
        for (int64_t k = 0; k < run && (int64_t)src.size() < e; ++k) src.push_back(v);
        v += 1 + (rng() % 4 == 0 ? (int32_t)(rng() % 40) : 0);
    }
    return src;
}

int main()
{
    std::mt19937_64 rng(7);
    // offsets
    {
        const int32_t n = 200000;
        std::vector<int32_t> src = skewed_sorted(rng, n, 1 << 20);
        check_offsets("skewed", src, n, true);
        check_offsets("skewed, n larger", src, n + 1000, true);
        std::vector<int32_t> one(100000, 3);
        check_offsets("one row", one, 10, true);
        check_offsets("one edge", std::vector<int32_t>{0}, 1, true);
        std::vector<int32_t> sparse;
        for (int i = 0; i < 5000; ++i) sparse.push_back(i * 397);
        check_offsets("sparse", sparse, 5000 * 397, true);
        std::vector<int32_t> bad = src;
        bad[bad.size() / 3] = bad[bad.size() / 3 - 1] - 1;
        check_offsets("decrease", bad, n, false);
        bad = src;
        bad[bad.size() / 2] = -1000;  // a decrease to a negative value
        check_offsets("decrease below zero", bad, n, false);
        bad = src;
        bad.back() = n;
        check_offsets("past n", bad, n, false);
        bad = src;
        bad[0] = -1;
        check_offsets("negative first", bad, n, false);
        bad = src;
        for (size_t i = bad.size() / 7; i < bad.size(); ++i) bad[i] = 0;  // a drop at a cut
        check_offsets("drop at a part boundary", bad, n, false);
    }
    // packing
    {
        std::vector<uint32_t> x(1 << 18);
        for (auto &y : x) {
            const uint64_t r = rng();
            y = r % 11 == 0 ? (uint32_t)(r >> 33) : (uint32_t)(r % 300);
            if (r % 5 == 0) y = 1;
        }
        x[77] = 0xffffffffu;  // a negative int32 weight: escaped too
        for (int wb : {2, 4, 8}) {
            const uint32_t m = (1u << wb) - 1u;
            for (int64_t a : {(int64_t)0, (int64_t)640}) {
                const int64_t b = (int64_t)x.size() - 3;
                std::vector<uint8_t> ps(x.size()), pv(x.size());
                std::vector<int64_t> es, ev;
                const bool s = pack_part(x.data(), a, b, ps.data(), es, 1 << 30, wb, Impl::Scalar);
                const bool v = pack_part(x.data(), a, b, pv.data(), ev, 1 << 30, wb, Impl::Avx2);
                CHECK(s && v, "pack wb=%d a=%lld refused", wb, (long long)a);
                const size_t lo = (size_t)(a * wb / 8), hi = (size_t)((b * wb + 7) / 8);
                CHECK(std::memcmp(ps.data() + lo, pv.data() + lo, hi - lo) == 0 && es == ev,
                      "pack wb=%d a=%lld: AVX2 and scalar differ", wb, (long long)a);
                std::vector<int64_t> eref;
                bool dec = true;
                for (int64_t i = a; i < b; ++i) {
                    const uint32_t got = (pv[(size_t)((i * wb) >> 3)] >> ((i * wb) & 7)) & m;
                    dec = dec && got == (x[(size_t)i] < m ? x[(size_t)i] : m);
                    if (x[(size_t)i] >= m) eref.push_back(i);
                }
                CHECK(dec && ev == eref, "pack wb=%d a=%lld: decode or escapes wrong", wb,
                      (long long)a);
                std::vector<int64_t> ec;
                const bool capped =
                    pack_part(x.data(), a, b, pv.data(), ec, (int64_t)eref.size() - 1, wb, Impl::Avx2);
                CHECK(!capped, "pack wb=%d: escape cap not enforced", wb);
            }
        }
        // the width choice on drawn weights
        for (int wb : {2, 4, 8}) {
            std::vector<uint32_t> y(1 << 20);
            const uint32_t top = wb == 2 ? 3 : wb == 4 ? 15 : 255;
            for (auto &v : y) v = (uint32_t)(rng() % top);
            CHECK(pack_width_of(y.data(), (int64_t)y.size()) == wb, "width choice for values below %u",
                  top);
        }
    }
    std::printf("%s (%d failures)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
