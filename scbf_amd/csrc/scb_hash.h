// SplitMix64 output function of the counter addresses, and its low 15 bits in
// fewer instructions. Device code; tests/cpp/hash_check.cpp compiles it on the
// host (with portable stand-ins for the 24-bit multiply and the funnel shift)
// and checks the short form against the full one.
#pragma once

#include <cstdint>

// filter.rs:629-635 (SplitMix64 output function)
__device__ __forceinline__ uint64_t xs64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Low 15 bits of xs64(x), enough for a slot address inside a block of at most
// 2^15 slots. The second multiply only has to be exact in bits 0..45 of its
// product (bit b < 15 of x ^ (x >> 31) reads bits b and b + 31): the low word
// comes from one 32x32->64 multiply and bits 32..45 from the two cross
// products, whose low 14 bits depend only on the operands' low 14 bits; the
// full-rate 24-bit multiplier (which reads the low 24 bits of each operand)
// gives them exactly, with no masking of the operands.
__device__ __forceinline__ uint32_t xs64_low15(uint64_t x) {
    constexpr uint64_t kC2 = 0x94D049BB133111EBull;
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    const uint32_t yl = uint32_t(x), yh = uint32_t(x >> 32);
    const uint64_t p = uint64_t(yl) * uint32_t(kC2);
    const uint32_t hi = uint32_t(p >> 32) + uint32_t(__umul24(yh, uint32_t(kC2) & 0xFFFFFFu)) +
                        uint32_t(__umul24(yl, uint32_t(kC2 >> 32) & 0xFFFFFFu));
    // low word of z ^ (z >> 31) for z = hi:lo; (z >> 31) is one funnel shift
    return uint32_t(p) ^ __builtin_amdgcn_alignbit(hi, uint32_t(p), 31);
}
