// encode.hpp - host side of the encoded edge-list upload (see encode.cpp).
#pragma once

#include <cstdint>
#include <vector>

namespace lv {
namespace enc {

enum class Impl { Auto, Scalar, Avx2 };  // Auto and Avx2: AVX2 where the CPU has it

// Row offsets off[0..n] of the part [a, b) of a source-sorted src (rows
// (src[a-1], src[b-1]] start in the part; the part ending at e also writes the
// empty rows after the last source).  Parts write disjoint entries; an
// unsorted or out-of-range list returns false without writing out of bounds.
bool offsets_part(const int32_t *src, int64_t a, int64_t b, int64_t e, int32_t n, int32_t *off,
                  Impl impl = Impl::Auto);
bool offsets_part(const int32_t *src, int64_t a, int64_t b, int64_t e, int32_t n, int64_t *off,
                  Impl impl = Impl::Auto);

// Int weights x[a, b) packed wb (2, 4 or 8) bits each into o (edge i at bit
// wb * i); weights of 2^wb - 1 and more get that value and their positions go
// to esc, at most cap of them (false beyond).  a is a multiple of 64.
bool pack_part(const uint32_t *x, int64_t a, int64_t b, uint8_t *o, std::vector<int64_t> &esc,
               int64_t cap, int wb, Impl impl = Impl::Auto);

// The packing width a sample of x[0, e) estimates cheapest: packed bytes plus
// 64 per escape.
int pack_width_of(const uint32_t *x, int64_t e);

}  // namespace enc
}  // namespace lv
