// Host check of the counter-address hash (scbf_amd/csrc/scb_hash.h): xs64
// gives the published SplitMix64 outputs for seed 0, and the short form
// xs64_low15 equals the low 15 bits of xs64 on random, structured and
// carry-heavy inputs, with the 24-bit multiply modelled as the hardware does
// it (the low 32 bits of the product of the operands' low 24 bits).
#include <cstdint>
#include <cstdio>
#include <random>

static int host_umul24(uint32_t a, uint32_t b) {
    return int(uint32_t((uint64_t(a & 0xFFFFFFu) * uint64_t(b & 0xFFFFFFu)) & 0xFFFFFFFFull));
}
// v_alignbit_b32: bits [sh, sh + 32) of hi:lo
static uint32_t host_alignbit(uint32_t hi, uint32_t lo, uint32_t sh) {
    return uint32_t(((uint64_t(hi) << 32) | lo) >> (sh & 31));
}
#define __umul24 host_umul24
#define __builtin_amdgcn_alignbit host_alignbit
#define __device__
#define __forceinline__ inline
#include "scb_hash.h"

int main() {
    // SplitMix64 from state 0: output i is xs64(i * golden gamma)
    const uint64_t published[4] = {0xE220A8397B1DCDAFull, 0x6E789E6AA1B965F4ull, 0x06C45D188009454Full,
                                   0xF88BB8A8724C81ECull};
    for (uint64_t i = 0; i < 4; ++i)
        if (xs64(i * 0x9E3779B97F4A7C15ull) != published[i]) {
            std::printf("xs64 output %llu differs\n", (unsigned long long)i);
            return 1;
        }
    uint64_t n = 0;
    auto check = [&](uint64_t x) {
        ++n;
        if ((xs64_low15(x) & 0x7FFFu) != (xs64(x) & 0x7FFFu)) {
            std::printf("xs64_low15 differs at x = 0x%016llx\n", (unsigned long long)x);
            return false;
        }
        return true;
    };
    std::mt19937_64 rng(12345);
    for (int i = 0; i < 4000000; ++i)
        if (!check(rng())) return 1;
    // chains as the bin kernels walk them, and single bits, runs and their complements
    uint64_t h = 1;
    for (int i = 0; i < 1000000; ++i, h = xs64(h))
        if (!check(h)) return 1;
    for (int a = 0; a < 64; ++a)
        for (int b = a; b < 64; ++b) {
            const uint64_t run = (b == 63 ? ~0ull : (1ull << (b + 1)) - 1) & ~((1ull << a) - 1);
            const uint64_t v[4] = {1ull << a, run, ~run, (1ull << a) | (1ull << b)};
            for (uint64_t x : v)
                if (!check(x) || !check(x - 0x9E3779B97F4A7C15ull)) return 1;
        }
    std::printf("xs64_low15 exact on %llu inputs\n", (unsigned long long)n);
    return 0;
}
