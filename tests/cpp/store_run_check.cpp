// Host check of store_run (scbf_amd/csrc/scb_store_run.h): for every start
// alignment and every length 0..16, exactly the bytes [dst, dst + len) are
// written, with the right values, and nothing around them.
#include <cstdint>
#include <cstdio>
#include <cstring>

struct uint4 {
    uint32_t x, y, z, w;
};
// v_alignbit_b32: the low 32 bits of (hi:lo) >> (sh & 31)
static uint32_t host_alignbit(uint32_t hi, uint32_t lo, uint32_t sh) {
    return uint32_t(((uint64_t(hi) << 32) | lo) >> (sh & 31));
}
#define __builtin_amdgcn_alignbit host_alignbit
#define __device__
#define __forceinline__ inline
#include "scb_store_run.h"

int main() {
    int bad = 0;
    for (uint32_t off = 0; off < 8; ++off) {
        for (uint32_t len = 0; len <= 16; ++len) {
            alignas(16) uint8_t buf[64];
            memset(buf, 0xEE, sizeof buf);
            const uint4 v{0x03020100u, 0x07060504u, 0x0B0A0908u, 0x0F0E0D0Cu};
            store_run(buf + 16 + off, v, len);
            for (uint32_t b = 0; b < sizeof buf; ++b) {
                const bool in = b >= 16 + off && b < 16 + off + len;
                const uint8_t want = in ? uint8_t(b - 16 - off) : 0xEE;
                if (buf[b] != want) {
                    printf("offset %u length %u: byte %u is %02x, expected %02x\n", off, len, b, buf[b], want);
                    ++bad;
                    break;
                }
            }
        }
    }
    if (!bad) printf("store_run exact for all alignments and lengths\n");
    return bad ? 1 : 0;
}
