// Checks mafx::fdiv against integer division: every divisor up to 5000 against edge and strided
// dividends, plus large and power-of-two divisors, for dividends below 2^31.
#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>

#include "fastdiv.hpp"

static int check(uint32_t d, uint32_t n) {
    const mafx::FastDiv f = mafx::make_fastdiv(d);
    if (mafx::fdiv(n, f) != n / d) {
        std::printf("FAIL d=%u n=%u got %u want %u\n", d, n, mafx::fdiv(n, f), n / d);
        return 1;
    }
    return 0;
}

int main() {
    const uint32_t top = 0x7fffffffu;
    int bad = 0;
    std::mt19937 rng(7);
    for (uint32_t d = 1; d <= 5000 && !bad; ++d) {
        for (uint32_t q = 0; q < 70; ++q) {
            const uint64_t base = static_cast<uint64_t>(q) * d;
            for (int o = -1; o <= 1; ++o) {
                const int64_t n = static_cast<int64_t>(base) + o;
                if (n >= 0 && n <= top) bad |= check(d, static_cast<uint32_t>(n));
            }
        }
        for (int k = 0; k < 64; ++k) bad |= check(d, rng() & top);
        bad |= check(d, top);
        bad |= check(d, top - d + 1);
    }
    for (int s = 0; s <= 31 && !bad; ++s) {
        const uint64_t p = uint64_t{1} << s;
        for (int o = -1; o <= 1; ++o) {
            const uint64_t d = p + o;
            if (d < 1 || d > (uint64_t{1} << 31)) continue;
            for (int k = 0; k < 2000; ++k) bad |= check(static_cast<uint32_t>(d), rng() & top);
            bad |= check(static_cast<uint32_t>(d), top);
        }
    }
    for (int k = 0; k < 200000 && !bad; ++k) {
        const uint32_t d = 1 + (rng() & top);
        bad |= check(d, rng() & top);
    }
    // xcd_swizzle is a bijection of [0, total) mapping XCD x = id mod 8 to one contiguous range
    for (uint32_t total = 1; total <= 3000 && !bad; total += (total < 100 ? 1 : 37)) {
        std::vector<int> seen(total, 0);
        for (uint32_t id = 0; id < total; ++id) {
            const uint32_t l = mafx::xcd_swizzle(id, total);
            if (l >= total || seen[l]++) {
                std::printf("FAIL xcd_swizzle total=%u id=%u -> %u\n", total, id, l);
                bad = 1;
                break;
            }
        }
        for (uint32_t id = 8; id < total && !bad; ++id)
            if (mafx::xcd_swizzle(id, total) != mafx::xcd_swizzle(id - 8, total) + 1) {
                std::printf("FAIL xcd_swizzle range total=%u id=%u\n", total, id);
                bad = 1;
            }
    }
    if (!bad) std::printf("ok\n");
    return bad;
}
