// encode.cpp - host side of the encoded edge-list upload (plain C++17; AVX2
// where the CPU has it): row offsets of a source-sorted src and int weights
// packed 2, 4 or 8 bits each.  upload.hip runs these on its host threads; the
// CPU tests compile this file on its own (tests/test_encode.py).
#include "encode.hpp"

#include <immintrin.h>

#include <algorithm>
#include <cstring>

namespace lv {
namespace enc {

namespace {

// Row offsets from a part [a, b) of a source-sorted list: the rows
// (src[a-1], src[b-1]] start in this part, and the last part also ends the
// empty rows after the last source at e; the parts write disjoint entries when
// src is sorted.  Every write is checked first (each row change must follow
// the last row written, and stay below n), so an unsorted or out-of-range list
// only returns false.
template <typename O>
bool offsets_scalar(const int32_t *src, int64_t a, int64_t b, int64_t e, int32_t n, O *off)
{
    int32_t prev = a == 0 ? -1 : src[a - 1];
    if (a == 0 && src[0] < 0) return false;
    for (int64_t i = a; i < b; ++i) {
        const int32_t v = src[i];
        if (v == prev) continue;
        if (v < prev || v >= n) return false;
        for (int32_t u = prev + 1; u <= v; ++u) off[u] = (O)i;
        prev = v;
    }
    if (b == e)
        for (int64_t u = (int64_t)prev + 1; u <= n; ++u) off[u] = (O)e;
    return true;
}

// The same with eight sources per compare: only the row changes (set bits of
// src[i] > src[i-1]) take the scalar path, so the irregular run lengths of a
// skewed graph cost no branch per edge.
template <typename O>
__attribute__((target("avx2"))) bool offsets_avx2(const int32_t *src, int64_t a, int64_t b,
                                                  int64_t e, int32_t n, O *off)
{
    if (a == 0 && src[0] < 0) return false;
    int32_t last = a == 0 ? -1 : src[a - 1];  // rows up to last are written
    int64_t i = a;
    auto change = [&](int64_t j) {  // src[j] != src[j-1] (or j == 0)
        const int32_t v = src[j];
        const int32_t p = j == 0 ? -1 : src[j - 1];
        if (p != last || v <= p || v >= n) return false;
        for (int32_t u = p + 1; u <= v; ++u) off[u] = (O)j;
        last = v;
        return true;
    };
    if (i < b && (i == 0 || src[i] != src[i - 1])) {
        if (!change(i)) return false;
    }
    ++i;
    int down = 0;
    for (; i + 8 <= b; i += 8) {
        const __m256i cur = _mm256_loadu_si256((const __m256i *)(src + i));
        const __m256i prv = _mm256_loadu_si256((const __m256i *)(src + i - 1));
        down |= _mm256_movemask_epi8(_mm256_cmpgt_epi32(prv, cur));
        unsigned m = (unsigned)_mm256_movemask_ps(_mm256_castsi256_ps(_mm256_cmpgt_epi32(cur, prv)));
        while (m) {
            const int k = __builtin_ctz(m);
            m &= m - 1;
            if (!change(i + k)) return false;
        }
    }
    for (; i < b; ++i) {
        if (src[i] < src[i - 1]) return false;
        if (src[i] != src[i - 1] && !change(i)) return false;
    }
    if (down) return false;
    if (b == e)
        for (int64_t u = (int64_t)last + 1; u <= n; ++u) off[u] = (O)e;
    return true;
}

// Int weights below m = 2^wb - 1 packed wb bits each (edge i at bit wb * i,
// little-endian within a byte); the others get m and their positions go to
// esc, at most cap of them (false beyond).  a is a multiple of 8 / wb.
bool pack_scalar(const uint32_t *x, int64_t a, int64_t b, uint8_t *o, std::vector<int64_t> &esc,
                 int64_t cap, int wb)
{
    const uint32_t m = (1u << wb) - 1u;
    const int per = 8 / wb;
    for (int64_t i = a; i < b; i += per) {
        uint32_t byte = 0;
        for (int q = 0; q < per && i + q < b; ++q) {
            const uint32_t y = x[i + q];
            if (y >= m) {
                if ((int64_t)esc.size() >= cap) return false;
                esc.push_back(i + q);
            }
            byte |= (y < m ? y : m) << (q * wb);
        }
        o[(i * wb) >> 3] = (uint8_t)byte;
    }
    return true;
}

// The same, 32 weights at a time: min(x, m) as 32 ordered bytes, then two
// (wb 4) or four (wb 2) of them folded into each output byte.
__attribute__((target("avx2"))) bool pack_avx2(const uint32_t *x, int64_t a, int64_t b,
                                               uint8_t *o, std::vector<int64_t> &esc, int64_t cap,
                                               int wb)
{
    const uint32_t m = (1u << wb) - 1u;
    const __m256i vm = _mm256_set1_epi32((int)m), vm1 = _mm256_set1_epi32((int)(m - 1));
    const __m256i order = _mm256_setr_epi32(0, 4, 1, 5, 2, 6, 3, 7);
    int64_t i = a;
    for (; i + 32 <= b; i += 32) {
        __m256i c[4];
        uint32_t em = 0;
#pragma GCC unroll 4
        for (int k = 0; k < 4; ++k) {
            const __m256i v = _mm256_loadu_si256((const __m256i *)(x + i + 8 * k));
            // unsigned v < m: max(v, m - 1) == m - 1
            const __m256i fits = _mm256_cmpeq_epi32(_mm256_max_epu32(v, vm1), vm1);
            em |= (~(uint32_t)_mm256_movemask_ps(_mm256_castsi256_ps(fits)) & 0xffu) << (8 * k);
            c[k] = _mm256_min_epu32(v, vm);
        }
        const __m256i p = _mm256_permutevar8x32_epi32(
            _mm256_packus_epi16(_mm256_packus_epi32(c[0], c[1]), _mm256_packus_epi32(c[2], c[3])),
            order);  // the 32 values as bytes, in order
        uint8_t *out = o + ((i * wb) >> 3);
        if (wb == 8) {
            _mm256_storeu_si256((__m256i *)out, p);
        } else if (wb == 4) {
            // 16-bit lane (v0, v1) -> v0 | v1 << 4
            const __m256i t =
                _mm256_or_si256(_mm256_and_si256(p, _mm256_set1_epi16(0x000f)),
                                _mm256_and_si256(_mm256_srli_epi16(p, 4), _mm256_set1_epi16(0x00f0)));
            const __m256i q = _mm256_packus_epi16(t, t);
            const uint64_t u0 = (uint64_t)_mm256_extract_epi64(q, 0);
            const uint64_t u1 = (uint64_t)_mm256_extract_epi64(q, 2);
            memcpy(out, &u0, 8);
            memcpy(out + 8, &u1, 8);
        } else {
            // 32-bit lane (v0, v1, v2, v3) -> v0 | v1 << 2 | v2 << 4 | v3 << 6
            const __m256i t = _mm256_or_si256(
                _mm256_or_si256(_mm256_and_si256(p, _mm256_set1_epi32(0x03)),
                                _mm256_and_si256(_mm256_srli_epi32(p, 6), _mm256_set1_epi32(0x0c))),
                _mm256_or_si256(_mm256_and_si256(_mm256_srli_epi32(p, 12), _mm256_set1_epi32(0x30)),
                                _mm256_and_si256(_mm256_srli_epi32(p, 18), _mm256_set1_epi32(0xc0))));
            const __m256i q16 = _mm256_packus_epi32(t, t);
            const __m256i q = _mm256_packus_epi16(q16, q16);
            const uint32_t u0 = (uint32_t)_mm256_extract_epi32(q, 0);
            const uint32_t u1 = (uint32_t)_mm256_extract_epi32(q, 4);
            memcpy(out, &u0, 4);
            memcpy(out + 4, &u1, 4);
        }
        while (em) {
            const int k = __builtin_ctz(em);
            em &= em - 1;
            if ((int64_t)esc.size() >= cap) return false;
            esc.push_back(i + k);
        }
    }
    return pack_scalar(x, i, b, o, esc, cap, wb);
}



// The packing width (2, 4 or 8 bits) whose cost a sample estimates lowest:
// the packed bytes plus 64 per escape (12 bytes to copy, and a host-side
// record and a re-read of its weight, about a cache line).  The sample is 256
// runs of 64 consecutive weights spread over the list (whole cache lines, so
// it costs a few hundred line reads).
int pack_width(const uint32_t *x, int64_t e)
{
    constexpr int64_t kRuns = 256, kRun = 64;
    int64_t ns = 0, c3 = 0, c15 = 0, c255 = 0;
    for (int64_t r = 0; r < kRuns; ++r) {
        const int64_t a = (e - std::min<int64_t>(e, kRun)) * r / (kRuns - 1) & ~(int64_t)15;
        for (int64_t i = a; i < std::min(e, a + kRun); ++i, ++ns) {
            const uint32_t y = x[i];
            c3 += y >= 3u;
            c15 += y >= 15u;
            c255 += y >= 255u;
        }
    }
    const double f = (double)e / (double)std::max<int64_t>(ns, 1);
    const double b2 = e / 4.0 + 64.0 * f * c3, b4 = e / 2.0 + 64.0 * f * c15,
                 b8 = (double)e + 64.0 * f * c255;
    return (b2 <= b4 && b2 <= b8) ? 2 : (b4 <= b8 ? 4 : 8);
}

bool use_avx2(Impl impl)
{
    static const bool have = __builtin_cpu_supports("avx2");
    return have && impl != Impl::Scalar;
}

}  // namespace

bool offsets_part(const int32_t *src, int64_t a, int64_t b, int64_t e, int32_t n, int32_t *off,
                  Impl impl)
{
    return use_avx2(impl) ? offsets_avx2(src, a, b, e, n, off)
                          : offsets_scalar(src, a, b, e, n, off);
}

bool offsets_part(const int32_t *src, int64_t a, int64_t b, int64_t e, int32_t n, int64_t *off,
                  Impl impl)
{
    return use_avx2(impl) ? offsets_avx2(src, a, b, e, n, off)
                          : offsets_scalar(src, a, b, e, n, off);
}

bool pack_part(const uint32_t *x, int64_t a, int64_t b, uint8_t *o, std::vector<int64_t> &esc,
               int64_t cap, int wb, Impl impl)
{
    return use_avx2(impl) ? pack_avx2(x, a, b, o, esc, cap, wb)
                          : pack_scalar(x, a, b, o, esc, cap, wb);
}

int pack_width_of(const uint32_t *x, int64_t e) { return pack_width(x, e); }

}  // namespace enc
}  // namespace lv
