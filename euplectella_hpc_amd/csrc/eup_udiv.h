// Division by a fixed divisor as a multiply and a shift, callable from host and device code.
// Free of HIP headers so that a plain C++ compiler can build it (the CPU test does).
#pragma once

#include <stdint.h>

#ifndef EUP_HD
#define EUP_HD inline
#endif

namespace eup {

// Division of n < 2^31 by a fixed divisor 1 <= d < 2^31 as a multiply and a shift:
// with s = ceil(log2 d) and M = floor(2^(31+s) / d) + 1 (< 2^32), n / d = (n M) >> (31 + s).
// Exact because M - 2^(31+s)/d lies in (0, 1] and n < 2^31 <= 2^(31+s)/d, so the error term
// n (M - 2^(31+s)/d) / 2^(31+s) stays below 1/d.  Uniform operands compile to three scalar
// instructions instead of a reciprocal-based integer division.
struct UDiv {
    uint32_t mul, shift;
    static UDiv make(uint32_t d)
    {
        uint32_t s = 0;
        while ((1ull << s) < d) ++s;
        UDiv u;
        u.mul = (uint32_t)(((1ull << (31 + s)) / d) + 1ull);
        u.shift = 31 + s;
        return u;
    }
    EUP_HD uint32_t div(uint32_t n) const { return (uint32_t)(((uint64_t)n * mul) >> shift); }
};

} // namespace eup
