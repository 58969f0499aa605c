// pack4_bases: four ASCII bases of one 32-bit word encoded at once. Device code
// of the scan; tests/cpp/pack4_check.cpp compiles it on the host (with a
// portable v_perm_b32) and checks it against the per-base rule for every byte.
#pragma once

#include <cstdint>

// Byte t of x is base t. Sets bits 2t..2t+1 of codes to the base's 2-bit code
// (A 0, C 1, T 2, G 3, either case: (ch >> 1) & 3), bit t of bad when the byte
// is not A, C, G or T in either case, and bit t of tg for a valid T or G.
__device__ __forceinline__ void pack4_bases(uint32_t x, uint32_t &codes, uint32_t &bad, uint32_t &tg) {
    const uint32_t idx = (x >> 1) & 0x07070707u;  // A, C, T, G -> 0, 1, 2, 3 in either case
    // the letter each index stands for; indices 4..7 select 0xFF, which no
    // byte & 0xDF equals
    const uint32_t want = __builtin_amdgcn_perm(0xFFFFFFFFu, 0x47544341u, idx);
    const uint32_t diff = (x & 0xDFDFDFDFu) ^ want;  // a zero byte is a valid base
    // bit 7 of each byte: set when the byte of diff is nonzero (exact per byte)
    const uint32_t nz = (((diff & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | diff) & 0x80808080u;
    // gathers: each multiplier moves byte t's bit to its own output bit, and
    // no two partial products overlap, so nothing carries
    bad = (nz * 0x00204081u) >> 28;
    codes = ((idx & 0x03030303u) * 0x01041040u) >> 24;
    tg = ((((idx & 0x02020202u) << 6) * 0x00204081u) >> 28) & ~bad;
}
