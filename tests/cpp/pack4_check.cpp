// Host check of pack4_bases (scbf_amd/csrc/scb_bases.h) against the per-base
// rule: every byte value at every position of the word, with random neighbours.
#include <cstdint>
#include <cstdio>
#include <random>

// v_perm_b32: byte i of the result is byte sel_i of the 64-bit value
// {src0 (high dword), src1 (low dword)} for selectors 0..7
static uint32_t host_perm(uint32_t src0, uint32_t src1, uint32_t sel) {
    const uint64_t v = (uint64_t(src0) << 32) | src1;
    uint32_t out = 0;
    for (int i = 0; i < 4; ++i) {
        const uint32_t s = (sel >> (8 * i)) & 0xFF;
        out |= uint32_t((v >> (8 * s)) & 0xFF) << (8 * i);
    }
    return out;
}
static uint32_t host_bitreverse32(uint32_t x) {  // used by the header's revcomp, not checked here
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}
static uint32_t host_alignbit(uint32_t hi, uint32_t lo, uint32_t sh) {  // used by the header's streams
    return uint32_t(((uint64_t(hi) << 32) | lo) >> (sh & 31));
}
#define __builtin_amdgcn_perm host_perm
#define __builtin_bitreverse32 host_bitreverse32
#define __builtin_amdgcn_alignbit host_alignbit
#define __device__
#define __forceinline__ inline
#include "scb_bases.h"

static bool valid(uint32_t ch) {
    const uint32_t u = ch & 0xDF;
    return u == 'A' || u == 'C' || u == 'G' || u == 'T';
}

int main() {
    std::mt19937 rng(7);
    long bad_cases = 0;
    for (int pos = 0; pos < 4; ++pos) {
        for (uint32_t ch = 0; ch < 256; ++ch) {
            for (int rep = 0; rep < 64; ++rep) {
                uint32_t x = rng();
                x = (x & ~(0xFFu << (8 * pos))) | (ch << (8 * pos));
                uint32_t codes, bad, tg;
                pack4_bases(x, codes, bad, tg);
                uint32_t wc = 0, wb = 0, wt = 0;
                for (int t = 0; t < 4; ++t) {
                    const uint32_t c = (x >> (8 * t)) & 0xFF, code = (c >> 1) & 3;
                    wc |= code << (2 * t);
                    wb |= uint32_t(!valid(c)) << t;
                    wt |= uint32_t(valid(c) && (code >> 1)) << t;
                }
                if (codes != wc || bad != wb || tg != wt) {
                    if (bad_cases++ < 5)
                        printf("x %08x: codes %02x/%02x bad %x/%x tg %x/%x\n", x, codes, wc, bad, wb, tg, wt);
                }
            }
        }
    }
    if (!bad_cases) printf("pack4_bases exact for every byte at every position\n");
    return bad_cases ? 1 : 0;
}
