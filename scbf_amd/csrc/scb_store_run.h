// store_run: a run of up to 16 bytes written with the fewest aligned stores.
// Device code of the query kernel; tests/cpp/store_run_check.cpp compiles it on
// the host (with a portable alignbit) and checks every alignment and length.
#pragma once

#include <cstdint>

// Writes the first len <= 16 bytes of v to global memory at dst, which may
// have any alignment: a byte and a short up to the next dword boundary, one
// store of whole dwords, then a short and a byte. A run of estimates then
// costs two or three memory requests instead of one per estimate.
__device__ __forceinline__ void store_run(uint8_t *dst, uint4 v, uint32_t len) {
    struct __attribute__((aligned(4))) D2 {
        uint32_t x, y;
    };
    struct __attribute__((aligned(4))) D3 {
        uint32_t x, y, z;
    };
    struct __attribute__((aligned(4))) D4 {
        uint32_t x, y, z, w;
    };
    const uint32_t head = (4u - uint32_t(reinterpret_cast<uintptr_t>(dst) & 3)) & 3;
    if (len < head) {  // at most two bytes, all before the next dword boundary
        if (len > 0) dst[0] = uint8_t(v.x);
        if (len > 1) dst[1] = uint8_t(v.x >> 8);
        return;
    }
    if (head & 1) dst[0] = uint8_t(v.x);
    if (head & 2) *reinterpret_cast<uint16_t *>(dst + (head & 1)) = uint16_t(v.x >> (8 * (head & 1)));
    // the bytes after the head, from a dword boundary
    const uint32_t sh = 8 * head;
    const uint32_t w0 = __builtin_amdgcn_alignbit(v.y, v.x, sh), w1 = __builtin_amdgcn_alignbit(v.z, v.y, sh);
    const uint32_t w2 = __builtin_amdgcn_alignbit(v.w, v.z, sh), w3 = __builtin_amdgcn_alignbit(0u, v.w, sh);
    uint8_t *p = dst + head;
    const uint32_t body = len - head, nd = body >> 2, tail = body & 3;
    if (nd == 4) {
        const D4 d{w0, w1, w2, w3};
        __builtin_memcpy(p, &d, 16);
    } else if (nd == 3) {
        const D3 d{w0, w1, w2};
        __builtin_memcpy(p, &d, 12);
    } else if (nd == 2) {
        const D2 d{w0, w1};
        __builtin_memcpy(p, &d, 8);
    } else if (nd == 1) {
        *reinterpret_cast<uint32_t *>(p) = w0;
    }
    if (tail) {
        const uint32_t t = nd == 0 ? w0 : nd == 1 ? w1 : nd == 2 ? w2 : w3;
        uint8_t *q = p + 4 * nd;
        if (tail & 2) *reinterpret_cast<uint16_t *>(q) = uint16_t(t);
        if (tail & 1) q[tail & 2] = uint8_t(t >> (8 * (tail & 2)));
    }
}
