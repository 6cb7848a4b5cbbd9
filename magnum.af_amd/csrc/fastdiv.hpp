// Division by a run-time invariant divisor without the GPU's software integer division.
//
// Granlund & Montgomery, "Division by invariant integers using multiplication" (PLDI 1994),
// round-up variant: for a divisor 1 <= d <= 2^31 let s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1.  Then for every dividend n < 2^31
//     n / d = (umulhi(n, m) + n) >> s,
// where umulhi is the high word of the 64-bit product; the sum stays below 2^32.  On gfx950 the
// quotient is one v_mul_hi_u32, one add and one shift instead of a ~40-instruction sequence
// (~100 for 64-bit operands).  Plain C++ with no HIP dependency, so host tests use the same code.
#pragma once
#include <cstdint>

namespace mafx {

struct FastDiv {
    uint32_t d; // divisor
    uint32_t m; // magic multiplier
    uint32_t s; // post-shift
};

inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t s = 0;
    while ((uint64_t{1} << s) < d) ++s;
    const uint64_t m = ((uint64_t{1} << 32) * ((uint64_t{1} << s) - d)) / d + 1;
    return FastDiv{d, static_cast<uint32_t>(m), s};
}

// n / f.d for n < 2^31
#if defined(__HIPCC__)
__host__ __device__
#endif
    inline uint32_t
    fdiv(uint32_t n, const FastDiv& f) {
    const uint32_t hi = static_cast<uint32_t>((static_cast<uint64_t>(n) * f.m) >> 32);
    return (hi + n) >> f.s;
}

// The dispatcher hands the work-groups of a launch to the 8 XCDs of gfx950 round-robin by linear
// id, and each XCD has its own L2.  Renumbering id -> xcd_swizzle(id, total) gives XCD x one
// contiguous range of logical ids (sizes differ by at most one), so tiles that share cache lines
// or stencil neighbours run behind one L2.  A bijection of [0, total).
#if defined(__HIPCC__)
__host__ __device__
#endif
    inline uint32_t
    xcd_swizzle(uint32_t id, uint32_t total) {
    constexpr uint32_t kXcd = 8;
    const uint32_t x = id % kXcd, q = total / kXcd, rem = total % kXcd;
    return x * q + (x < rem ? x : rem) + id / kXcd;
}

} // namespace mafx
