// Base encoding and 2-bit base streams shared by the scan and the bin kernels.
// Device code; tests/cpp/pack4_check.cpp and tests/cpp/smer_stream_check.cpp
// compile it on the host (with portable stand-ins for the two AMDGPU builtins)
// and check it against direct per-base definitions.
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

// Reverse complement of a len-base 2-bit word (packed-seq read_revcomp_kmer).
__device__ __forceinline__ uint64_t revcomp(uint64_t x, uint32_t len) {
    uint32_t lo = uint32_t(x), hi = uint32_t(x >> 32);
    uint64_t r = (uint64_t(__builtin_bitreverse32(lo)) << 32) | __builtin_bitreverse32(hi);
    r = ((r >> 1) & 0x5555555555555555ull) | ((r & 0x5555555555555555ull) << 1);
    r ^= 0xAAAAAAAAAAAAAAAAull;
    return r >> (64 - 2 * len);
}

__device__ __forceinline__ uint64_t kmask(uint32_t len) {
    return len >= 32 ? ~0ull : ((1ull << (2 * len)) - 1);
}

// Streams the bases of one super-k-mer (from 4 pre-aligned words holding its
// first 64 bases) and calls f(j, canonical s-mer) for each of its s-mers in
// position order (filter.rs:224-226, 423-426 with canonical_kmer 622-627).
// The first s-mer is read directly from the words; the rest roll one base at a
// time. one_word (uniform) asserts 17 <= s <= 32 and nbases - s <= 16: every
// base after the first s-mer then sits in one word, and the s-mer halves roll
// as 32-bit words (the low half needs no mask).
template <class F>
__device__ __forceinline__ void for_each_smer(uint32_t A0, uint32_t A1, uint32_t A2, uint32_t A3,
                                              uint32_t nbases, uint32_t s, F &&f, bool one_word = false) {
    const uint64_t mask = kmask(s);
    const uint32_t top = 2 * (s - 1);
    uint64_t fw = ((uint64_t(A1) << 32) | A0) & mask;
    uint64_t rc = revcomp(fw, s);
    f(0u, fw < rc ? fw : rc);
    if (one_word) {
        // bases s .. s + 15, base s in bits 0-1
        uint32_t cur = s < 32 ? __builtin_amdgcn_alignbit(A2, A1, 2 * (s & 15)) : A2;
        uint32_t fl = uint32_t(fw), fh = uint32_t(fw >> 32), rl = uint32_t(rc), rh = uint32_t(rc >> 32);
        const uint32_t mh = uint32_t(mask >> 32), th = top - 32;
        for (uint32_t i = s; i < nbases; ++i) {
            const uint32_t c = cur & 3;
            cur >>= 2;
            fl = __builtin_amdgcn_alignbit(fh, fl, 2);
            fh = (fh >> 2) | (c << th);
            rh = __builtin_amdgcn_alignbit(rh, rl, 30) & mh;
            rl = (rl << 2) | (c ^ 2);
            const uint64_t f64 = (uint64_t(fh) << 32) | fl, r64 = (uint64_t(rh) << 32) | rl;
            f(i + 1 - s, f64 < r64 ? f64 : r64);
        }
        return;
    }
    // word stream positioned at base s (s is uniform across the wave)
    if (s >= 16) A0 = A1, A1 = A2, A2 = A3;
    if (s >= 32) A0 = A1, A1 = A2;
    uint32_t cur = A0 >> (2 * (s & 15));
    for (uint32_t i = s; i < nbases; ++i) {
        uint32_t c = cur & 3;
        cur >>= 2;
        if ((i & 15) == 15) {
            cur = A1;
            A1 = A2;
            A2 = A3;
        }
        fw = (fw >> 2) | (uint64_t(c) << top);
        rc = ((rc << 2) | (c ^ 2)) & mask;
        f(i + 1 - s, fw < rc ? fw : rc);
    }
}

// for_each_smer that also rolls the k-mer ending at the same base (bases
// packed first base lowest, as canonical_at reads them): f(j, canonical s-mer
// j, canonical k-mer ending at s-mer j's last base). The k-mer is complete
// from j = k - s on (j + 1 >= k - s + 1); before that it is not meaningful.
// It starts as the first s-mer placed at the top of a k-base word, so the
// bases shifted in after it complete it.
// one_word as for for_each_smer (then k >= s >= 17 and both rolls use 32-bit halves).
template <class F>
__device__ __forceinline__ void for_each_smer_kmer(uint32_t A0, uint32_t A1, uint32_t A2, uint32_t A3,
                                                   uint32_t nbases, uint32_t s, uint32_t k, F &&f,
                                                   bool one_word = false) {
    const uint64_t mask = kmask(s), kmk = kmask(k);
    const uint32_t top = 2 * (s - 1), ktop = 2 * (k - 1);
    uint64_t fw = ((uint64_t(A1) << 32) | A0) & mask;
    uint64_t rc = revcomp(fw, s);
    uint64_t fk = fw << (2 * (k - s)), rk = rc;
    f(0u, fw < rc ? fw : rc, fk < rk ? fk : rk);
    if (one_word) {
        uint32_t cur = s < 32 ? __builtin_amdgcn_alignbit(A2, A1, 2 * (s & 15)) : A2;
        uint32_t fl = uint32_t(fw), fh = uint32_t(fw >> 32), rl = uint32_t(rc), rh = uint32_t(rc >> 32);
        uint32_t kl = uint32_t(fk), kh = uint32_t(fk >> 32), ql = uint32_t(rk), qh = uint32_t(rk >> 32);
        const uint32_t mh = uint32_t(mask >> 32), th = top - 32;
        const uint32_t kmh = uint32_t(kmk >> 32), kth = ktop - 32;
        for (uint32_t i = s; i < nbases; ++i) {
            const uint32_t c = cur & 3;
            cur >>= 2;
            fl = __builtin_amdgcn_alignbit(fh, fl, 2);
            fh = (fh >> 2) | (c << th);
            rh = __builtin_amdgcn_alignbit(rh, rl, 30) & mh;
            rl = (rl << 2) | (c ^ 2);
            kl = __builtin_amdgcn_alignbit(kh, kl, 2);
            kh = (kh >> 2) | (c << kth);
            qh = __builtin_amdgcn_alignbit(qh, ql, 30) & kmh;
            ql = (ql << 2) | (c ^ 2);
            const uint64_t f64 = (uint64_t(fh) << 32) | fl, r64 = (uint64_t(rh) << 32) | rl;
            const uint64_t k64 = (uint64_t(kh) << 32) | kl, q64 = (uint64_t(qh) << 32) | ql;
            f(i + 1 - s, f64 < r64 ? f64 : r64, k64 < q64 ? k64 : q64);
        }
        return;
    }
    if (s >= 16) A0 = A1, A1 = A2, A2 = A3;
    if (s >= 32) A0 = A1, A1 = A2;
    uint32_t cur = A0 >> (2 * (s & 15));
    for (uint32_t i = s; i < nbases; ++i) {
        uint32_t c = cur & 3;
        cur >>= 2;
        if ((i & 15) == 15) {
            cur = A1;
            A1 = A2;
            A2 = A3;
        }
        fw = (fw >> 2) | (uint64_t(c) << top);
        rc = ((rc << 2) | (c ^ 2)) & mask;
        fk = (fk >> 2) | (uint64_t(c) << ktop);
        rk = ((rk << 2) | (c ^ 2)) & kmk;
        f(i + 1 - s, fw < rc ? fw : rc, fk < rk ? fk : rk);
    }
}
