// PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically
// Good Algorithms for Random Number Generation", XSH-RR 64/32 variant) with
// Brown's arbitrary-stride jump-ahead, and the [0,1) float mapping used by
// tiny-cuda-nn's parameter initialisation (23 mantissa bits, 1 + u - 1).
#pragma once

#include <cstdint>
#include <cstring>

namespace nlod {

struct Pcg32 {
    static constexpr uint64_t kMult = 0x5851f42d4c957f2dULL;
    uint64_t state = 0x853c49e6748fea9bULL;
    uint64_t inc = 0xda3e39cb94b95bdbULL;

    Pcg32() = default;
    explicit Pcg32(uint64_t initstate, uint64_t initseq = 1u) { seed(initstate, initseq); }

    void seed(uint64_t initstate, uint64_t initseq = 1u) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next_uint();
        state += initstate;
        next_uint();
    }
    uint32_t next_uint() {
        const uint64_t old = state;
        state = old * kMult + inc;
        const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31u));
    }
    float next_float() {
        const uint32_t u = (next_uint() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &u, 4);
        return f - 1.0f;
    }
    // state after `delta` steps (Brown 1994, square-and-multiply on the LCG)
    void advance(uint64_t delta) {
        uint64_t cur_mult = kMult, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        while (delta > 0) {
            if (delta & 1u) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1u) * cur_plus;
            cur_mult *= cur_mult;
            delta >>= 1u;
        }
        state = acc_mult * state + acc_plus;
    }
};

}  // namespace nlod
