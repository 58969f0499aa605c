// Host check of the 2-bit base streams (scbf_amd/csrc/scb_bases.h): for every
// k <= 32 and s <= k and random records of up to 62 bases, for_each_smer and
// for_each_smer_kmer yield the canonical s-mers and k-mers that a direct
// evaluation from the bases gives.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

static uint32_t host_bitreverse32(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}
static uint32_t host_perm(uint32_t, uint32_t, uint32_t) { return 0; }  // not used here
// v_alignbit_b32: bits [sh, sh + 32) of hi:lo
static uint32_t host_alignbit(uint32_t hi, uint32_t lo, uint32_t sh) {
    return uint32_t(((uint64_t(hi) << 32) | lo) >> (sh & 31));
}
#define __builtin_bitreverse32 host_bitreverse32
#define __builtin_amdgcn_perm host_perm
#define __builtin_amdgcn_alignbit host_alignbit
#define __device__
#define __forceinline__ inline
#include "scb_bases.h"

// canonical form of bases [a, a + len) (first base in the low bits; complement c ^ 2)
static uint64_t canon(const std::vector<uint32_t> &b, uint32_t a, uint32_t len) {
    uint64_t fw = 0, rc = 0;
    for (uint32_t i = 0; i < len; ++i) {
        fw |= uint64_t(b[a + i]) << (2 * i);
        rc |= uint64_t(b[a + len - 1 - i] ^ 2u) << (2 * i);
    }
    return fw < rc ? fw : rc;
}

int main() {
    std::mt19937 rng(11);
    long bad = 0, checked = 0, checked_one_word = 0;
    for (uint32_t k = 1; k <= 32; ++k) {
        for (uint32_t s = 1; s <= k; ++s) {
            for (int rep = 0; rep < 40; ++rep) {
                const uint32_t nk = 1 + rng() % 32;
                const uint32_t nbases = nk + k - 1;
                if (nbases > 62) continue;
                std::vector<uint32_t> b(64);
                for (auto &x : b) x = rng() & 3u;
                uint32_t A[4] = {0, 0, 0, 0};
                for (uint32_t i = 0; i < 64; ++i) A[i >> 4] |= b[i] << (2 * (i & 15));
                std::vector<uint64_t> sm;
                for_each_smer(A[0], A[1], A[2], A[3], nbases, s, [&](uint32_t j, uint64_t x) {
                    if (j != sm.size()) ++bad;
                    sm.push_back(x);
                });
                if (s >= 17 && nbases - s <= 16) {  // the one-word stream gives the same s-mers
                    uint32_t n1 = 0;
                    for_each_smer(
                        A[0], A[1], A[2], A[3], nbases, s,
                        [&](uint32_t j, uint64_t x) {
                            if (j != n1++ || x != sm[j]) {
                                if (bad++ < 5) printf("k %u s %u j %u: one-word s-mer mismatch\n", k, s, j);
                            }
                        },
                        true);
                    if (n1 != sm.size()) ++bad;
                    ++checked_one_word;
                }
                for_each_smer_kmer(A[0], A[1], A[2], A[3], nbases, s, k,
                                   [&](uint32_t j, uint64_t smer, uint64_t kmer) {
                                       ++checked;
                                       if (smer != canon(b, j, s) || smer != sm[j]) {
                                           if (bad++ < 5) printf("k %u s %u j %u: s-mer mismatch\n", k, s, j);
                                       }
                                       if (j + 1 >= k - s + 1 && kmer != canon(b, j + s - k, k)) {
                                           if (bad++ < 5) printf("k %u s %u j %u: k-mer mismatch\n", k, s, j);
                                       }
                                   });
                if (sm.size() != nbases - s + 1) ++bad;
                if (s >= 17 && nbases - s <= 16) {  // one-word s-mer and k-mer streams
                    for_each_smer_kmer(
                        A[0], A[1], A[2], A[3], nbases, s, k,
                        [&](uint32_t j, uint64_t smer, uint64_t kmer) {
                            if (smer != canon(b, j, s) ||
                                (j + 1 >= k - s + 1 && kmer != canon(b, j + s - k, k))) {
                                if (bad++ < 5) printf("k %u s %u j %u: one-word k-mer stream mismatch\n", k, s, j);
                            }
                        },
                        true);
                }
            }
        }
    }
    if (!checked_one_word) ++bad;
    if (!bad) printf("s-mer and k-mer streams exact (%ld positions, %ld one-word records)\n", checked, checked_one_word);
    return bad ? 1 : 0;
}
