// gfx950 kernels of the SuperCountingBloom hot path and the Engine that drives
// them. All arithmetic is integer and bit-exact with the reference CPU path
// (super-counting-bloom 0.1.0, src/filter.rs); each kernel cites the loop it
// replaces.
//
// Device data layout (HBM):
//   table   : 2^cse counters of 1/2/4 bytes, physical block order
//             P(b) = ((b & shard_mask) << sub_log) | (b >> shard_shift), which is
//             the .scb payload order (filter.rs:575-580), so save/load are copies.
//   pk      : 2-bit codes of the batch, 16 bases per u32, base i at bits 2*(i%16)
//   rec     : one u64 per super-k-mer: window index (32 bits) | (n_kmers-1) (5) |
//             physical block (27)
//   sorted  : rec grouped by bin (a run of consecutive physical blocks whose
//             counters fit in LDS)
//
// Batched pipeline (one stream, no host sync inside):
//   k_scan      per 1952-window tile: ASCII -> codes, ACGT-run and read-boundary
//               masks, canonical ntHash of every m-mer, minimizer of every k-mer
//               window, super-k-mer cut, block hash -> rec (appended to one of
//               64 segments)
//   k_seg_hist / k_small_scan / k_partition  rec -> records grouped by bin
//               (tile-aggregated counting sort: a coarse pass, then a fine pass)
//   k_plan_sums / k_plan_write  bin offsets + work list (hot bins split)
//   k_bin_insert / k_bin_query / k_bin_spectrum  persistent work queue, one bin
//               per item, counters staged in LDS, written back once (insert) or
//               read once (query, spectrum)
// Plain pipeline: k_scan alone, one thread per super-k-mer, global atomics/loads.

#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <thread>
#include <type_traits>

#include "scb_bases.h"
#include "scb_engine.h"
#include "scb_hash.h"
#include "scb_store_run.h"

namespace scb {
namespace {

// ----------------------------------------------------------------------------
// Constants and small device helpers
// ----------------------------------------------------------------------------
// A tile's 256 threads load 16 bases each for 2048 bases of span, which hold
// its windows, the 32 bases before them and 64 after (a window of k <= 32 and
// the hashes of the w - 1 windows after the tile reach 62 bases past its end).
constexpr int kTile = 1952;            // windows owned by one scan tile
constexpr int kTileThreads = 256;
constexpr int kLeft = 32;              // LDS holds bases [t0 - 32, t0 + kTile + 64)
constexpr int kSpan = kTile + 96;
static_assert(kSpan == 16 * 128, "two waves load the span, 16 bases per thread");
static_assert(kTile % 32 == 0 && kTile <= 8 * kTileThreads, "at most eight windows per thread, whole bitmap words");
constexpr int kSpanWords = kSpan / 32 + 2;
constexpr int kPkWords = kSpan / 16 + 8;
constexpr int kHashN = kTile + 64;
constexpr int kWinN = kTile + 32;
static_assert(kWinN <= kHashN, "window positions reuse a hash-position array");
constexpr int kGPad = kSpan + kSpan / 16;  // padded XOR-prefix array length
constexpr int kGArr = 2 * kGPad > 2 * kHashN ? 2 * kGPad : 2 * kHashN;  // prefixes, then four u16 arrays
constexpr uint16_t kNoWin = 0xFFFF;
constexpr int kPkPadWords = 16;        // zero words after the packed batch

enum StatSlot {
    kStRecords = 0,
    kStValid = 1,
    kStSmers = 2,
    kStQueried = 3,
    kStPositive = 4,
    kStChecksum = 5,
    kStNWork = 6,
    kStWorkHead = 7,
    kStOverflow = 8,
    kStDistinct = 9,
    kStGlobalKmers = 10,
    kStBadOffsets = 11,
    kStEmitted = 12,
    kStSetFull = 13,  // a key found no free slot in the global spectrum set
    kStCount = 16
};

struct DevGeo {
    uint32_t k, m, s, nh, w, win, bse, shard_mask, shard_shift, sub_log;
    uint64_t block_mask;
};

// Every base of a super-k-mer after its first s-mer fits in one 32-bit word
// (a record has at most w k-mers, so at most w + k - 1 - s such bases), with
// 17 <= s: the one-word s-mer stream applies (for_each_smer).
__device__ __forceinline__ bool smer_tail_one_word(const DevGeo &g) {
    return g.s >= 17 && g.w + g.k - 1 - g.s <= 16;
}

DevGeo dev_geo(const Geometry &g) {
    DevGeo d;
    d.k = g.k, d.m = g.m, d.s = g.s, d.nh = g.nh, d.w = g.w, d.win = g.win, d.bse = g.bse;
    d.shard_mask = g.shard_mask, d.shard_shift = g.shard_shift, d.sub_log = g.sub_log;
    d.block_mask = g.block_mask;
    return d;
}

// Calls f(address) for the n_hashes chain values xs64(x), xs64(xs64(x)), ...
// of one canonical s-mer (filter.rs:224-233, 478-481); amask < 2^15.
template <class F>
__device__ __forceinline__ void for_each_address(uint64_t smer, uint32_t nh, uint32_t amask, F &&f) {
    uint64_t h = smer;
    for (uint32_t t = 1; t < nh; ++t) {
        h = xs64(h);
        f(uint32_t(h) & amask);
    }
    f(xs64_low15(h) & amask);
}

// for_each_address with the chain unrolled for NH = n_hashes (NH = 0: any
// n_hashes, read at run time).
template <int NH, class F>
__device__ __forceinline__ void for_each_address_n(uint64_t h, uint32_t nh, uint32_t amask, F &&f) {
    if constexpr (NH == 0) {
        for_each_address(h, nh, amask, f);
    } else {
#pragma unroll
        for (int t = 0; t + 1 < NH; ++t) {
            h = xs64(h);
            f(uint32_t(h) & amask);
        }
        f(xs64_low15(h) & amask);
    }
}

// n_hashes as a compile-time constant for the s-mer loops: with_nh(nh, f)
// calls f(NhTag<NH>{}) with NH = nh for 1..4 and NH = 0 (run-time count)
// otherwise. The uniform switch runs once per record, so the per-s-mer code
// holds no scalar compares and branches on n_hashes.
template <int N>
struct NhTag {
    static constexpr int value = N;
};
template <class F>
__device__ __forceinline__ void with_nh(uint32_t nh, F &&f) {
    switch (nh) {
        case 4: f(NhTag<4>{}); break;
        case 3: f(NhTag<3>{}); break;
        case 2: f(NhTag<2>{}); break;
        case 1: f(NhTag<1>{}); break;
        default: f(NhTag<0>{}); break;
    }
}

// Minimum of the counters base[a] over the addresses of one canonical s-mer
// (filter.rs:478-489), with all NH loads issued before the first is used
// (NH = 0: any n_hashes, one load at a time); amask < 2^15.
template <int NH, class T>
__device__ __forceinline__ uint32_t min_at_nh(uint64_t h, uint32_t nh, uint32_t amask, const T *base, uint32_t kmax) {
    if constexpr (NH == 0) {
        uint32_t est = kmax;
        for_each_address(h, nh, amask, [&](uint32_t ad) {
            const uint32_t v = base[ad];
            est = v < est ? v : est;
        });
        return est;
    } else {
        uint32_t v[NH];
#pragma unroll
        for (int t = 0; t + 1 < NH; ++t) {
            h = xs64(h);
            v[t] = base[uint32_t(h) & amask];
        }
        v[NH - 1] = base[xs64_low15(h) & amask];
        uint32_t m = v[0];
#pragma unroll
        for (int t = 1; t < NH; ++t) m = v[t] < m ? v[t] : m;
        return m;
    }
}

// Canonical k-mer of len bases starting at base index b of a packed array.
__device__ __forceinline__ uint64_t canonical_at(const uint32_t *words, uint64_t b, uint32_t len) {
    uint64_t wi = b >> 4;
    uint32_t sh = 2 * uint32_t(b & 15);
    uint32_t w0 = words[wi], w1 = words[wi + 1], w2 = words[wi + 2];
    uint32_t lo = __builtin_amdgcn_alignbit(w1, w0, sh);
    uint32_t hi = __builtin_amdgcn_alignbit(w2, w1, sh);
    uint64_t fw = ((uint64_t(hi) << 32) | lo) & kmask(len);
    uint64_t rc = revcomp(fw, len);
    return fw < rc ? fw : rc;
}

__device__ __forceinline__ uint32_t rotl32(uint32_t x, uint32_t r) {
    return __builtin_amdgcn_alignbit(x, x, (32 - (r & 31)) & 31);
}
__device__ __forceinline__ uint32_t rotr32(uint32_t x, uint32_t r) {
    return __builtin_amdgcn_alignbit(x, x, r & 31);
}

// Inclusive scan over the 64 lanes of a wave (XOR or +) with DPP moves: rows
// of 16 lanes by shifts of 1, 2, 4 and 8, then the last lane of rows 0 and 2
// into rows 1 and 3, then lane 31 into rows 2 and 3. A lane without a source
// takes 0, the identity of both operations. Every lane of the wave must be
// active.
template <bool XOR>
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t x) {
    auto step = [&](uint32_t y) { return XOR ? (x ^ y) : (x + y); };
    x = step(__builtin_amdgcn_update_dpp(0u, x, 0x111, 0xF, 0xF, false));  // row_shr:1
    x = step(__builtin_amdgcn_update_dpp(0u, x, 0x112, 0xF, 0xF, false));  // row_shr:2
    x = step(__builtin_amdgcn_update_dpp(0u, x, 0x114, 0xF, 0xF, false));  // row_shr:4
    x = step(__builtin_amdgcn_update_dpp(0u, x, 0x118, 0xF, 0xF, false));  // row_shr:8
    x = step(__builtin_amdgcn_update_dpp(0u, x, 0x142, 0xA, 0xF, false));  // row_bcast:15
    x = step(__builtin_amdgcn_update_dpp(0u, x, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return x;
}

// ntHash seeds (low 32 bits of the original ntHash seeds), by 2-bit code
// A=0 C=1 T=2 G=3. Third-party arithmetic (seq-hash 0.1.2), restated.
constexpr uint32_t seed_of(uint32_t c) {
    return c == 0 ? 0x95c60474u : c == 1 ? 0x62a02b4cu : c == 2 ? 0x4be24456u : 0x82572324u;
}

// The ntHash XOR-prefix terms of the scan by (strand, code, rotation), entry
// strand * 128 + code * 32 + r: forward rotr(seed(code), r), reverse
// rotl(seed(code ^ 2), r). Built at compile time.
struct SeedTerms {
    uint32_t v[256];
};
constexpr uint32_t rotl_c(uint32_t x, uint32_t r) { return r == 0 ? x : (x << r) | (x >> (32 - r)); }
constexpr SeedTerms make_seed_terms() {
    SeedTerms s{};
    for (uint32_t e = 0; e < 256; ++e) {
        const uint32_t strand = e >> 7, c = (e >> 5) & 3, r = e & 31;
        s.v[e] = strand ? rotl_c(seed_of(c ^ 2), r) : rotl_c(seed_of(c), (32 - r) & 31);
    }
    return s;
}
__device__ constexpr SeedTerms kSeedTerms = make_seed_terms();

__device__ __forceinline__ uint32_t get_bits(const uint32_t *bm, int li, int n) {
    if (n <= 0) return 0;
    uint32_t w0 = bm[li >> 5], w1 = bm[(li >> 5) + 1];
    uint32_t v = __builtin_amdgcn_alignbit(w1, w0, uint32_t(li & 31));
    return n >= 32 ? v : (v & ((1u << n) - 1));
}

template <int BITS>
struct Ctr;
template <>
struct Ctr<8> {
    using T = uint8_t;
    static constexpr uint32_t kMax = 0xFFu;
};
template <>
struct Ctr<16> {
    using T = uint16_t;
    static constexpr uint32_t kMax = 0xFFFFu;
};
template <>
struct Ctr<32> {
    using T = uint32_t;
    static constexpr uint32_t kMax = 0xFFFFFFFFu;
};

// counter.rs:20-23 applied through a 32-bit CAS (plain path and hot-bin merge):
// adds `add` saturating to the sub-word counter at element index idx.
template <int BITS>
__device__ __forceinline__ void sat_add_global(void *table, uint64_t idx, uint32_t add) {
    if (BITS == 32) {
        uint32_t *p = reinterpret_cast<uint32_t *>(table) + idx;
        uint32_t old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (old != 0xFFFFFFFFu) {
            uint64_t sum = uint64_t(old) + add;
            uint32_t nv = sum > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(sum);
            uint32_t prev = atomicCAS(p, old, nv);
            if (prev == old) break;
            old = prev;
        }
    } else {
        constexpr uint32_t per = 32 / BITS;
        constexpr uint32_t mx = Ctr<BITS>::kMax;
        uint32_t *p = reinterpret_cast<uint32_t *>(table) + idx / per;
        uint32_t sh = uint32_t(idx % per) * BITS;
        uint32_t old = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (;;) {
            uint32_t cur = (old >> sh) & mx;
            if (cur == mx) break;
            uint32_t nc = cur + add > mx ? mx : cur + add;
            uint32_t nv = (old & ~(mx << sh)) | (nc << sh);
            uint32_t prev = atomicCAS(p, old, nv);
            if (prev == old) break;
            old = prev;
        }
    }
}

template <int BITS>
__device__ __forceinline__ uint32_t load_ctr(const void *base, uint64_t idx) {
    return reinterpret_cast<const typename Ctr<BITS>::T *>(base)[idx];
}

// Loads 5 packed words starting at base b and aligns them so that A0 holds
// bases b..b+15.
__device__ __forceinline__ void load_aligned(const uint32_t *words, uint64_t b, uint32_t &A0,
                                             uint32_t &A1, uint32_t &A2, uint32_t &A3) {
    uint64_t wi = b >> 4;
    uint32_t sh = 2 * uint32_t(b & 15);
    uint32_t W0 = words[wi], W1 = words[wi + 1], W2 = words[wi + 2], W3 = words[wi + 3],
             W4 = words[wi + 4];
    A0 = __builtin_amdgcn_alignbit(W1, W0, sh);
    A1 = __builtin_amdgcn_alignbit(W2, W1, sh);
    A2 = __builtin_amdgcn_alignbit(W3, W2, sh);
    A3 = __builtin_amdgcn_alignbit(W4, W3, sh);
}

// Sliding minimum over the last `win` s-mer estimates (filter.rs:485-507; min
// is associative, so a shift register gives the deque's result).
template <int WINC>
struct WinMin {
    // m[i] = minimum of the last i + 1 estimates; each push is WINC - 1 mins
    uint32_t m[WINC];
    __device__ __forceinline__ WinMin() {
#pragma unroll
        for (int i = 0; i < WINC; ++i) m[i] = 0xFFFFFFFFu;
    }
    __device__ __forceinline__ void add(uint32_t est) {
#pragma unroll
        for (int i = WINC - 1; i > 0; --i) m[i] = m[i - 1] < est ? m[i - 1] : est;
        m[0] = est;
    }
    // window of `win` <= WINC estimates
    __device__ __forceinline__ uint32_t push(uint32_t est, uint32_t win) {
        add(est);
        uint32_t r = m[0];
#pragma unroll
        for (int i = 1; i < WINC; ++i) r = uint32_t(i) + 1 == win ? m[i] : r;
        return r;
    }
    // window of exactly WINC estimates
    __device__ __forceinline__ uint32_t push_exact(uint32_t est) {
        add(est);
        return m[WINC - 1];
    }
};

// ----------------------------------------------------------------------------
// Fill kernel (used instead of hipMemsetAsync so that every step can be
// captured in a graph)
// ----------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

constexpr uint32_t kHistSmall = 4096;
constexpr uint32_t kHistBigCap = 1u << 20;

// ---- distinct-k-mer spectrum (filter.rs:407-460; fastx.rs:255-308, 392-410) ----
// A canonical k-mer always lands in the same counter block (its minimizer is
// strand-symmetric), so every occurrence of a key is processed by the work
// items of one bin. A single-item bin dedups in an LDS hash set of kSetCap keys
// (32 KiB: two spectrum workgroups share a CU); a key whose first kSetProbes
// slots there hold other keys goes to a global open-addressing set, as do all
// keys of split bins and of bins with more than kSetDirect k-mers. The
// estimate of a key is the same at every occurrence, so "first wins" needs no
// order.
constexpr uint32_t kSetLog = 12, kSetCap = 1u << kSetLog;
constexpr uint32_t kSetProbes = 32;
constexpr uint32_t kSetDirect = 4 * kSetCap;
constexpr uint32_t kGlobalProbes = 1024;      // a global insert that finds no slot raises kStSetFull
constexpr uint32_t kSpecHist = 1024;          // LDS histogram entries of the spectrum bin kernel
constexpr uint64_t kEmptyKey = ~0ull;          // keys use at most 62 bits (k <= 31)

__device__ __forceinline__ uint32_t key_slot(uint64_t key, uint32_t log2cap) {
    return uint32_t((key * 0x9E3779B97F4A7C15ull) >> (64 - log2cap));
}

// Inserts key into an open-addressing set of 2^log2cap slots, visiting at most
// max_probes slots of its probe sequence: 1 if this call inserted it, 0 if it
// was there, 2 if every visited slot holds another key. Slots only change from
// empty to a key, so a key never occupies two slots, and a key that found its
// window full never enters this set later (the window stays full): with a
// second set taking the keys that return 2, each key is in exactly one set.
__device__ __forceinline__ uint32_t set_insert_bounded(unsigned long long *set, uint32_t log2cap, uint64_t key,
                                                       uint32_t max_probes) {
    const uint32_t mask = (1u << log2cap) - 1;
    uint32_t h = key_slot(key, log2cap);
    for (uint32_t p = 0; p < max_probes; ++p) {
        const unsigned long long cur = set[h];
        if (cur == key) return 0;
        if (cur == kEmptyKey) {
            const unsigned long long prev = atomicCAS(&set[h], kEmptyKey, (unsigned long long)key);
            if (prev == kEmptyKey) return 1;
            if (prev == key) return 0;
        }
        h = (h + 1) & mask;
    }
    return 2;
}

// Returns true if key was not in the set (this call inserted it).
__device__ __forceinline__ bool set_insert(unsigned long long *set, uint32_t log2cap, uint64_t key) {
    const uint32_t mask = (1u << log2cap) - 1;
    uint32_t h = key_slot(key, log2cap);
    for (;;) {
        unsigned long long cur = set[h];
        if (cur == key) return false;
        if (cur == kEmptyKey) {
            unsigned long long prev = atomicCAS(&set[h], kEmptyKey, (unsigned long long)key);
            if (prev == kEmptyKey) return true;
            if (prev == key) return false;
        }
        h = (h + 1) & mask;
    }
}

struct HistOut {
    unsigned long long *dense;
    uint32_t dense_n;
    uint32_t *big_keys;
    unsigned long long *big_counts;
    uint64_t *stats;  // stats[kStOverflow] is raised when big_keys is full
};

template <uint32_t SMALL = kHistSmall>
__device__ __forceinline__ void spec_hist_add(const HistOut &a, uint32_t *s_hist, uint32_t v) {
    if (v < SMALL) {
        atomicAdd(&s_hist[v], 1u);
    } else if (v < a.dense_n) {
        atomicAdd(&a.dense[v], 1ull);
    } else {
        uint32_t hsh = uint32_t(xs64(v)) & (kHistBigCap - 1);
        for (uint32_t probe = 0; probe < kHistBigCap; ++probe) {
            uint32_t key = a.big_keys[hsh];
            if (key == 0) {
                uint32_t prev = atomicCAS(&a.big_keys[hsh], 0u, v);
                if (prev == 0 || prev == v) {
                    atomicAdd(&a.big_counts[hsh], 1ull);
                    return;
                }
                key = prev;
            }
            if (key == v) {
                atomicAdd(&a.big_counts[hsh], 1ull);
                return;
            }
            hsh = (hsh + 1) & (kHistBigCap - 1);
        }
        atomicAdd((unsigned long long *)&a.stats[kStOverflow], 1ull);
    }
}

// ----------------------------------------------------------------------------
// Scan tile (front end shared by both pipelines)
// ----------------------------------------------------------------------------
constexpr int kSeg = 64;         // record segments (one reservation counter each)
constexpr int kStSpread = 16;    // stats[kStSpread + slot * kSeg + (tile & 63)]
enum SpreadSlot { kSpValid = 0, kSpSmers = 1, kSpQueried = 2, kSpPositive = 3, kSpChecksum = 4, kSpCount = 5 };
constexpr int kStatsLen = kStSpread + kSpCount * kSeg;

struct ScanArgs {
    const uint8_t *bases;
    const uint64_t *offsets;
    uint64_t n_reads, n_bases;
    uint32_t *pk;          // packed codes out (batched) or null
    uint32_t *valid_bits;  // per-window validity out or null
    uint64_t *rec;         // records out (batched): kSeg segments of seg_cap records
    uint32_t *seg_count;   // records written per segment
    uint64_t seg_cap;
    uint64_t *stats;
    void *table;           // plain paths
    void *out;             // plain query abundances (position-indexed) or null
    uint32_t aligned16;    // bases pointer is 16-byte aligned
    const uint32_t *tile_r0;  // first record starting after each tile's left edge
    uint32_t ntiles;
    // plain spectrum: the key set (a power of two of slots, at most half full) and the histogram
    unsigned long long *gset;
    uint32_t glog2;
    HistOut hist;
};

// For every scan tile, the first record whose start lies after the tile's left
// edge t*kTile - kLeft (a binary search per tile, all tiles in parallel); and
// the offsets contract of a batch (offsets[0] = 0, non-decreasing, offsets[n] =
// n_bases), checked by one thread per record: a violation sets
// stats[kStBadOffsets] and the call returns an error.
__global__ void k_tile_r0(const uint64_t *offsets, uint64_t n_reads, uint64_t n_bases, uint32_t ntiles,
                          uint32_t *r0, uint64_t *stats) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bool bad = t < n_reads && offsets[t] > offsets[t + 1];
    if (t == 0) bad |= offsets[0] != 0 || offsets[n_reads] != n_bases;
    if (bad) atomicOr(reinterpret_cast<unsigned long long *>(stats) + kStBadOffsets, 1ull);
    if (t >= ntiles) return;
    int64_t lb = int64_t(t) * kTile - kLeft;
    uint64_t lo = 0, hi = n_reads;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (int64_t(offsets[mid]) > lb) hi = mid; else lo = mid + 1;
    }
    r0[t] = uint32_t(lo);
}

enum ScanMode { kModeRecords = 0, kModePlainInsert = 1, kModePlainQuery = 2, kModePlainSpectrum = 3 };

// One tile owns windows [t0, t0 + kTile). Window j is valid iff its k bases are
// ACGT and belong to one record (the maximal-ACGT-run split of api.rs:676-726,
// 799-853 and fastx.rs:14-19). Its minimizer is the canonical-ntHash argmin
// over its w m-mers (simd-minimizers restatement, third-party), and a
// super-k-mer starts where the minimizer position changes or a run starts
// (filter.rs:608-620, 210-216).
//
// Phases: (1) 16 bases per thread from one 16-byte load -> 2-bit word and
// non-ACGT / T-G masks; (2) ntHash as XOR prefixes G(i) over the span, so the
// hash of the m-mer at i is rotl(Gf(i+m)^Gf(i), i+m-1) + rotr(Gr(i+m)^Gr(i), i);
// (3) van Herk/Gil-Werman block prefix/suffix minima of (key16 << 16 | pos)
// for the leftmost and (key16 << 16 | ~pos) for the rightmost argmin; (4) each
// window takes the min of one suffix and one prefix value.
template <int MODE, int BITS, int WINC, int WREG = 0>
__global__ __launch_bounds__(kTileThreads) void k_scan(ScanArgs a, DevGeo g) {
    __shared__ uint32_t s_bad[kSpanWords], s_tg[kSpanWords], s_brk[kSpanWords];
    __shared__ uint32_t s_pk[kPkWords];
    // Gf | Gr (padded: index li + li/16 keeps the 16-per-thread writes
    // bank-conflict free); before them the seed table below, after them the
    // block prefix/suffix minimum arrays (and, with register windows, the keys)
    __shared__ __align__(16) uint32_t s_G[kGArr];
    // high 16 bits of each m-mer hash (the argmin key): its own array for the
    // block minima, which use all four u16 arrays of s_G; with register windows
    // it lives in s_G (LDS of 20 KiB per tile: eight tiles share a CU)
    __shared__ __align__(16) uint16_t s_Kb[WREG > 0 ? 8 : kHashN];
    __shared__ uint32_t s_wx[2][kTileThreads / 64];
    __shared__ uint32_t s_wc[kTileThreads / 64];
    // bit b set: window b + 1 ends the super-k-mer before it (invalid or a new start)
    __shared__ uint32_t s_bnd[kTile / 32 + 2];
    __shared__ uint64_t s_red[3][kTileThreads / 64];
    __shared__ uint64_t s_base;
    // plain spectrum: counts of small estimates of the tile's distinct keys
    __shared__ uint32_t s_hsm[MODE == kModePlainSpectrum ? kHistSmall : 1];
    // ntHash XOR-prefix terms by (strand, code, rotation): forward rotr(seed(c), r),
    // reverse rotl(seed(c ^ 2), r); rows of 33 words put the four codes of one
    // rotation in different banks. Read before the prefixes are written to s_G.
    constexpr int kSeedRow = 33;
    static_assert(2 * 4 * kSeedRow <= kGArr, "the seed table fits in s_G");
    uint32_t *s_seed = s_G;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    constexpr int nwave = kTileThreads / 64;
    constexpr int kChunks = kSpan / 16;
    const int64_t N = int64_t(a.n_bases);
    const int k = int(g.k), m = int(g.m), w = int(g.w);
    uint16_t *s_bad16 = reinterpret_cast<uint16_t *>(s_bad);
    uint16_t *s_tg16 = reinterpret_cast<uint16_t *>(s_tg);
    uint32_t *s_Gf = s_G, *s_Gr = s_G + kGPad;
    // positions of the leftmost / rightmost minimum of every block prefix and suffix
    uint16_t *s_PL = reinterpret_cast<uint16_t *>(s_G);
    uint16_t *s_SL = s_PL + kHashN, *s_PR = s_PL + 2 * kHashN, *s_SR = s_PL + 3 * kHashN;
    static_assert(kGArr * 2 >= 4 * kHashN, "four u16 arrays fit in s_G");
    // register windows use only s_PL (start list) and s_SL (minimizer positions)
    uint16_t *s_K = WREG > 0 ? s_PR : s_Kb;
    // minimizer position per window overwrites s_SL in place: window hj reads
    // s_SL only at index hj before writing it
    uint16_t *s_mpos = s_SL;
    const int NH = kTile + 2 * w - 1;
    uint64_t acc_smers = 0, acc_pos = 0, acc_sum = 0;
    uint32_t wave_valid = 0;    // valid windows of the wave (uniform)
    uint32_t tile_records = 0;  // records the tile emits

    // 16 bases per thread are prefetched one tile ahead.
    auto fetch = [&](int64_t tile_lb, uint32_t (&wv)[4]) {
        const int64_t g0 = tile_lb + 16 * tid;
        if (a.aligned16 && g0 >= 0 && g0 + 16 <= N) {
            uint4 v = *reinterpret_cast<const uint4 *>(a.bases + g0);
            wv[0] = v.x, wv[1] = v.y, wv[2] = v.z, wv[3] = v.w;
        } else {  // unaligned or at the batch edges: byte loads
#pragma unroll 1
            for (int q = 0; q < 4; ++q) {
                uint32_t x = 0;
#pragma unroll 1
                for (int t = 0; t < 4; ++t) {
                    int64_t gi = g0 + 4 * q + t;
                    uint32_t ch = (gi >= 0 && gi < N) ? a.bases[gi] : uint32_t('N');
                    x |= ch << (8 * t);
                }
                if (q == 0) wv[0] = x;
                if (q == 1) wv[1] = x;
                if (q == 2) wv[2] = x;
                if (q == 3) wv[3] = x;
            }
        }
    };
    // One tile per workgroup (a persistent loop lets the compiler hoist
    // per-thread invariants into ~60 more VGPRs and halves occupancy).
    {
        const uint32_t tile = blockIdx.x;
        uint32_t wv[4] = {0, 0, 0, 0};
        if (tid < kChunks) fetch(int64_t(tile) * kTile - kLeft, wv);
        const int64_t t0 = int64_t(tile) * kTile;
        const int64_t LB = t0 - kLeft;
        for (int i = tid; i < kSpanWords; i += kTileThreads) s_brk[i] = 0;
        if (MODE == kModePlainSpectrum)
            for (int i = tid; i < int(kHistSmall); i += kTileThreads) s_hsm[i] = 0;
        // the first read start after the span's left edge (uniform), and this
        // thread's first candidate start, loaded now so that the load overlaps
        // the work before the read boundaries are marked (the index is clamped,
        // the test is made where the value is used)
        const uint32_t r0 = a.tile_r0[tile];
        const uint64_t rfirst = uint64_t(r0) + tid;
        const int64_t ofirst = int64_t(a.offsets[min<uint64_t>(rfirst, a.n_reads - 1)]);
        static_assert(kTileThreads == 256, "one seed term per thread");
        s_seed[(tid >> 5) * kSeedRow + (tid & 31)] = kSeedTerms.v[tid];

        // (1) codes and masks of this thread's 16 bases
        uint32_t cw = 0;
        if (tid < kChunks) {
            uint32_t badm = 0, tgm = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {  // four bases per word at once
                uint32_t c4, b4, t4;
                pack4_bases(wv[q], c4, b4, t4);
                cw |= c4 << (8 * q);
                badm |= b4 << (4 * q);
                tgm |= t4 << (4 * q);
            }
            s_pk[tid] = cw;
            s_bad16[tid] = uint16_t(badm);
            s_tg16[tid] = uint16_t(tgm);
            if (MODE == kModeRecords && tid >= kLeft / 16 && tid < kLeft / 16 + kTile / 16)
                a.pk[(LB >> 4) + tid] = cw;  // LB is a multiple of 16
        } else if (tid < kPkWords) {
            s_pk[tid] = 0;
            if (tid < 2 * kSpanWords) s_bad16[tid] = 0xFFFF, s_tg16[tid] = 0;
        }

        // (2) ntHash XOR prefixes: the chunk-local exclusive prefixes stay in
        // registers until the prefix of the preceding chunks is known, then go to
        // LDS complete, so each m-mer hash reads two words per strand. Base
        // 16 * tid + t is rotated by r0 + t (r0 = 16 for odd threads): its terms
        // come from the seed table.
        __syncthreads();  // the seed table is complete
        uint32_t tf[16], tr[16];
        uint32_t xf = 0, xr = 0;
        if (tid < kChunks) {
            const uint32_t r0 = 16 * (uint32_t(tid) & 1);
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const uint32_t ix = ((cw >> (2 * t)) & 3) * kSeedRow + r0 + uint32_t(t);
                tf[t] = xf, tr[t] = xr;
                xf ^= s_seed[ix];
                xr ^= s_seed[4 * kSeedRow + ix];
            }
        }
        // the chunks' XOR totals, scanned over the wave (the waves that load
        // bases are whole: kChunks is a multiple of 64)
        static_assert(kChunks % 64 == 0, "whole waves load the span");
        uint32_t sf = 0, sr = 0;
        if (tid < kChunks) {
            sf = wave_scan_incl<true>(xf);
            sr = wave_scan_incl<true>(xr);
            if (lane == 63) s_wx[0][wid] = sf, s_wx[1][wid] = sr;
        }
        __syncthreads();
        {   // record starts inside the span mark read boundaries
            uint64_t r = rfirst;
            int64_t o = ofirst;
            while (r < a.n_reads && o < LB + kSpan) {
                const int li = int(o - LB);
                if (li >= 0) atomicOr(&s_brk[li >> 5], 1u << (li & 31));  // (offsets that decrease: an error)
                r += kTileThreads;
                if (r < a.n_reads) o = int64_t(a.offsets[r]);
            }
        }
        if (tid < kChunks) {
            uint32_t pf = 0, pr = 0;
            for (int i = 0; i < wid; ++i) pf ^= s_wx[0][i], pr ^= s_wx[1][i];
            const uint32_t of = pf ^ sf ^ xf, orr = pr ^ sr ^ xr;  // XOR of the chunks before this one
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                s_Gf[17 * tid + t] = tf[t] ^ of;
                s_Gr[17 * tid + t] = tr[t] ^ orr;
            }
        }
        __syncthreads();

        // canonical hash of the m-mer at p = t0-1+hp, hp in [0, NH)
        auto mmer_key = [&](int hp) {
            const int li = hp + kLeft - 1, lm = li + m;
            const int pi = li + (li >> 4), pm = lm + (lm >> 4);
            const uint32_t gf = s_Gf[pm] ^ s_Gf[pi];
            const uint32_t gr = s_Gr[pm] ^ s_Gr[pi];
            const uint32_t fw = rotl32(gf, uint32_t(lm - 1) & 31);
            const uint32_t rc = rotr32(gr, uint32_t(li) & 31);
            return uint16_t((fw + rc) >> 16);
        };
        if constexpr (WREG > 0) {
            // the keys overwrite part of the prefixes: held in registers until
            // every thread has read its prefixes
            constexpr int NHW = kTile + 2 * WREG - 1, KR = (NHW + kTileThreads - 1) / kTileThreads;
            // thread tid takes hp = tid + q kTileThreads: kTileThreads is a multiple
            // of 32 and of 16, so every q shares the rotation amounts and the
            // padded prefix index moves by a constant (folded into the LDS offsets)
            static_assert(kTileThreads % 32 == 0, "keys of one thread share their rotations");
            constexpr int kStride = kTileThreads + kTileThreads / 16;
            const int li0 = tid + kLeft - 1, lm0 = li0 + m;
            const uint32_t *gfi = s_Gf + li0 + (li0 >> 4), *gfm = s_Gf + lm0 + (lm0 >> 4);
            const uint32_t *gri = s_Gr + li0 + (li0 >> 4), *grm = s_Gr + lm0 + (lm0 >> 4);
            const uint32_t rot_f = uint32_t(1 - lm0), rot_r = uint32_t(li0);  // rotl by lm - 1, rotr by li
            auto key_at = [&](int q) {
                const int o = q * kStride;
                const uint32_t gf = gfm[o] ^ gfi[o], gr = grm[o] ^ gri[o];
                return __builtin_amdgcn_alignbit(gf, gf, rot_f) + __builtin_amdgcn_alignbit(gr, gr, rot_r);
            };
            uint32_t key[KR];  // the key is the high half
#pragma unroll
            for (int q = 0; q < KR; ++q) {
                if ((q + 1) * kTileThreads <= NHW) key[q] = key_at(q);                       // every thread
                else key[q] = tid + q * kTileThreads < NHW ? key_at(q) : 0u;                  // the last keys
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < KR; ++q)
                if (tid + q * kTileThreads < NHW) s_K[tid + q * kTileThreads] = uint16_t(key[q] >> 16);
        } else {
            for (int hp = tid; hp < NH; hp += kTileThreads) s_K[hp] = mmer_key(hp);
        }
        __syncthreads();

        if constexpr (WREG > 0) {
            // (3+4) with w = WREG known at compile time: thread tid takes the 8
            // windows hj in [8 tid, 8 tid + 8); their 8 + w - 1 keys come from LDS
            // in 16-byte loads, and the minimum of window i of the group is
            // min(suffix minimum of keys [i, 8), prefix minimum of keys [8, i + w)):
            // w >= 9 makes every window reach key 8. Leftmost and rightmost
            // minima pack (key, i) and (key, 31 - i) with the key's index i in
            // the group (a constant), so a min keeps the smaller key and, among
            // equal keys, the leftmost or the rightmost index.
            static_assert(WREG >= 9 && WREG <= 25, "register windows need 9 <= w <= 25");
            constexpr int NK = 8 + WREG - 1, NV = (NK + 7) / 8;
            // the groups cover the tile's windows and the w after it, hj < kTile + w
            // (threads kTile / 8 .. of the last wave, otherwise idle here): a group
            // past kTile + w - 1 reads keys beyond the NHW computed ones, which only
            // its windows past kTile + w - 1 use, and those are never read
            constexpr int kGroups = (kTile + WREG + 7) / 8;
            static_assert(kGroups <= kTileThreads, "one group per thread");
            static_assert(8 * (kGroups - 1) + 8 * NV <= kHashN, "a group's key loads stay inside the key array");
            if (tid < kGroups) {
                const int h0 = tid * 8;
                uint32_t kw[NV * 4];  // two 16-bit keys per word
                const uint4 *src = reinterpret_cast<const uint4 *>(s_K + h0);
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const uint4 x = src[v];
                    kw[4 * v] = x.x, kw[4 * v + 1] = x.y, kw[4 * v + 2] = x.z, kw[4 * v + 3] = x.w;
                }
                // one byte permute: the key's two bytes above the index's two
                auto packed = [&](int i, bool right) {
                    return __builtin_amdgcn_perm(kw[i >> 1], uint32_t(right ? 31 - i : i),
                                                 (i & 1) ? 0x07060100u : 0x05040100u);
                };
                // the group's windows read bits [base, base + 7 + k) of the masks:
                // three words per mask, then one funnel shift per window
                const int base = h0 + kLeft - 1;
                const uint32_t bsh = uint32_t(base & 31), bwi = uint32_t(base >> 5);
                auto span64 = [&](const uint32_t *bm, uint32_t &lo, uint32_t &hi) {
                    const uint32_t w0 = bm[bwi], w1 = bm[bwi + 1], w2 = bm[bwi + 2];
                    lo = __builtin_amdgcn_alignbit(w1, w0, bsh);
                    hi = __builtin_amdgcn_alignbit(w2, w1, bsh);
                };
                uint32_t bad_lo, bad_hi, brk_lo, brk_hi, tg_lo, tg_hi;
                span64(s_bad, bad_lo, bad_hi);
                span64(s_brk, brk_lo, brk_hi);
                span64(s_tg, tg_lo, tg_hi);
                const uint32_t mk = k >= 32 ? ~0u : (1u << k) - 1, mk1 = (1u << (k - 1)) - 1;
                uint32_t sl[8], sr[8], ql[WREG - 1], qr[WREG - 1];
#pragma unroll
                for (int i = 7; i >= 0; --i) {
                    sl[i] = i == 7 ? packed(i, false) : min(packed(i, false), sl[i + 1]);
                    sr[i] = i == 7 ? packed(i, true) : min(packed(i, true), sr[i + 1]);
                }
#pragma unroll
                for (int j = 0; j < WREG - 1; ++j) {
                    ql[j] = j == 0 ? packed(8, false) : min(packed(8 + j, false), ql[j - 1]);
                    qr[j] = j == 0 ? packed(8, true) : min(packed(8 + j, true), qr[j - 1]);
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    // bits [li, li + k) with li = base + i (and the breaks after li)
                    const uint32_t bad = (__builtin_amdgcn_alignbit(bad_hi, bad_lo, i) & mk) |
                                         (__builtin_amdgcn_alignbit(brk_hi, brk_lo, i + 1) & mk1);
                    const bool left =  // leftmost on ties: more than k / 2 of T or G
                        __popc(__builtin_amdgcn_alignbit(tg_hi, tg_lo, i) & mk) > uint32_t(k >> 1);
                    const uint32_t pl = min(sl[i], ql[i + WREG - 9]), pr = min(sr[i], qr[i + WREG - 9]);
                    // 31 - x is ~x in the low five bits
                    const uint32_t pos = uint32_t(h0) + ((left ? pl : ~pr) & 31u);
                    s_mpos[h0 + i] = bad ? kNoWin : uint16_t(pos);
                }
            }
            __syncthreads();
        } else {
            // (3) van Herk/Gil-Werman: per block of w hash positions, the positions
            // of the leftmost and rightmost minimum of every prefix and suffix
            {
                const int nblk = (NH + w - 1) / w;
                for (int b = tid; b < nblk; b += kTileThreads) {
                    const int h0 = b * w, h1 = min(h0 + w, NH);
                    uint32_t cl = ~0u, cr = ~0u;
                    for (int hp = h0; hp < h1; ++hp) {
                        uint32_t key = uint32_t(s_K[hp]) << 16;
                        cl = min(cl, key | uint32_t(hp));
                        cr = min(cr, key | uint32_t(0xFFFF - hp));
                        s_PL[hp] = uint16_t(cl), s_PR[hp] = uint16_t(0xFFFF - (cr & 0xFFFF));
                    }
                    cl = cr = ~0u;
                    for (int hp = h1 - 1; hp >= h0; --hp) {
                        uint32_t key = uint32_t(s_K[hp]) << 16;
                        cl = min(cl, key | uint32_t(hp));
                        cr = min(cr, key | uint32_t(0xFFFF - hp));
                        s_SL[hp] = uint16_t(cl), s_SR[hp] = uint16_t(0xFFFF - (cr & 0xFFFF));
                    }
                }
            }
            __syncthreads();

            // (4) minimizer position of every window j = t0-1+hj, hj in [0, kTile + w)
            // Branch-free: both tie rules read the same shape of data, so the rule
            // only selects the arrays and the direction of the final comparison.
            for (int hj = tid; hj < kTile + w; hj += kTileThreads) {
                const int li = hj + kLeft - 1;
                const uint32_t bad = get_bits(s_bad, li, k) | get_bits(s_brk, li + 1, k - 1);
                const bool left = 2 * int(__popc(get_bits(s_tg, li, k))) > k;  // leftmost minimum on ties
                const int e = hj + w - 1;
                const uint32_t c1 = (left ? s_SL : s_SR)[hj], c2 = (left ? s_PL : s_PR)[e];
                const uint32_t k1 = s_K[c1], k2 = s_K[c2];
                // leftmost: the suffix candidate c1 (left of c2) wins ties; rightmost: c2 wins ties
                const uint32_t pos = left ? (k1 <= k2 ? c1 : c2) : (k2 <= k1 ? c2 : c1);
                s_mpos[hj] = bad ? kNoWin : uint16_t(pos);
            }
            __syncthreads();
        }

        // owned windows hj in [1, kTile]: thread tid takes hj = 1 + 8 tid + u,
        // u in [0, 8), whose validity, super-k-mer starts and boundaries are one
        // byte each (bit u), written as bytes of the bitmaps
        uint32_t starts = 0;  // bit u set if window 1 + 8 tid + u starts a super-k-mer
        if (MODE == kModeRecords && wid == 0) {
            // boundaries of the windows after the tile (hj in [kTile + 1, kTile + w))
            const int hj = kTile + 1 + lane;
            bool b = false;
            if (hj < kTile + w) {
                uint16_t mp = s_mpos[hj], pm = s_mpos[hj - 1];
                b = mp == kNoWin || pm != mp;
            }
            uint64_t bm = __ballot(b);
            if (lane == 0) s_bnd[kTile / 32] = uint32_t(bm), s_bnd[kTile / 32 + 1] = uint32_t(bm >> 32);
        }
        if (tid < kTile / 8) {
            // minimizer positions of windows 8 tid .. 8 tid + 8 (16-byte aligned)
            const uint4 q = *reinterpret_cast<const uint4 *>(s_mpos + 8 * tid);
            const uint32_t mpw[5] = {q.x, q.y, q.z, q.w, uint32_t(s_mpos[8 * tid + 8])};
            uint32_t vbyte = 0, bbyte = 0;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t pm = (mpw[u >> 1] >> (16 * (u & 1))) & 0xFFFFu;
                const uint32_t mp = (mpw[(u + 1) >> 1] >> (16 * ((u + 1) & 1))) & 0xFFFFu;
                const bool valid = mp != kNoWin;
                // a boundary: invalid, or a new minimizer position (a start)
                const bool bnd = !valid || pm != mp;
                wave_valid += uint32_t(__popcll(__ballot(valid)));
                vbyte |= uint32_t(valid) << u;
                bbyte |= uint32_t(bnd) << u;
            }
            starts = vbyte & bbyte;
            if (MODE == kModeRecords) reinterpret_cast<uint8_t *>(s_bnd)[tid] = uint8_t(bbyte);
            if (a.valid_bits) {
                // windows j0 .. j0 + 7 (t0 is a multiple of 32): every byte of each
                // bitmap word that starts before the batch end is written
                const int64_t j0 = t0 + 8 * tid;
                if ((j0 & ~int64_t(31)) < N) reinterpret_cast<uint8_t *>(a.valid_bits)[j0 >> 3] = uint8_t(vbyte);
            }
        }

        const uint32_t my = __popc(starts);
        if (MODE == kModeRecords) {
            // exclusive scan of the per-thread start counts (wave scan + wave totals)
            const uint32_t incl = wave_scan_incl<false>(my);
            if (lane == 63) s_wc[wid] = incl;
            __syncthreads();
            uint32_t before = 0, total = 0;
            for (int i = 0; i < nwave; ++i) {
                if (i < wid) before += s_wc[i];
                total += s_wc[i];
            }
            tile_records = total;
            if (tid == 0) {
                uint32_t seg = tile & (kSeg - 1);
                s_base = seg * a.seg_cap + atomicAdd(&a.seg_count[seg], total);
            }
            // compact the start windows into a list (s_PL is dead after phase 4) so
            // that every thread emits the same number of records, in slot order
            uint16_t *s_list = s_PL;
            {
                uint32_t at = before + incl - my;
                for (uint32_t bits = starts; bits; bits &= bits - 1)
                    s_list[at++] = uint16_t(1 + 8 * tid + (__ffs(bits) - 1));
            }
            __syncthreads();
            const uint64_t base = s_base;
            for (uint32_t i = tid; i < total; i += kTileThreads) {
                const int hj = s_list[i];
                const uint16_t mp = s_mpos[hj];
                // k-mers of the super-k-mer: up to the next boundary window (at most w)
                const uint32_t wi = uint32_t(hj) >> 5;
                const uint64_t bw = (uint64_t(s_bnd[wi + 1]) << 32 | s_bnd[wi]) >> (hj & 31);
                const uint32_t nk = bw ? min(uint32_t(__builtin_ctzll(bw)) + 1, uint32_t(w)) : uint32_t(w);
                uint64_t val = canonical_at(s_pk, uint64_t(mp) + kLeft - 1, uint32_t(m));
                uint64_t block = xs64(val) & g.block_mask;
                uint64_t pblk = ((block & g.shard_mask) << g.sub_log) | (block >> g.shard_shift);
                uint64_t j = uint64_t(t0 - 1 + hj);
                a.rec[base + i] = j | (uint64_t(nk - 1) << 32) | (pblk << 37);
            }
        } else {
            // Plain path: this thread applies its own super-k-mers straight to HBM.
            for (uint32_t bits = starts; bits; bits &= bits - 1) {
                int u = __ffs(bits) - 1;
                int hj = 1 + 8 * tid + u;
                uint16_t mp = s_mpos[hj];
                int e = hj + 1;
                while (e < hj + w && s_mpos[e] == mp) ++e;
                uint32_t nk = uint32_t(e - hj);
                uint64_t val = canonical_at(s_pk, uint64_t(mp) + kLeft - 1, uint32_t(m));
                uint64_t block = xs64(val) & g.block_mask;
                uint64_t pblk = ((block & g.shard_mask) << g.sub_log) | (block >> g.shard_shift);
                uint64_t cbase = pblk << g.bse;
                const uint64_t amask = (1ull << g.bse) - 1;
                uint32_t A0, A1, A2, A3;
                load_aligned(s_pk, uint64_t(hj + kLeft - 1), A0, A1, A2, A3);
                uint32_t nbases = nk + g.k - 1;
                if (MODE == kModePlainSpectrum) {
                    // filter.rs:407-460 (as filter.rs:302-335): the estimate of every
                    // k-mer, counted when its canonical key enters the key set
                    WinMin<WINC> wm;
                    for_each_smer_kmer(A0, A1, A2, A3, nbases, g.s, g.k, [&](uint32_t j, uint64_t smer, uint64_t key) {
                        uint64_t h = xs64(smer);
                        uint32_t est = Ctr<BITS>::kMax;
                        for (uint32_t t = 0; t < g.nh; ++t) {
                            uint32_t v = load_ctr<BITS>(a.table, cbase + (h & amask));
                            est = v < est ? v : est;
                            if (t + 1 < g.nh) h = xs64(h);
                        }
                        const uint32_t ab = wm.push(est, g.win);
                        if (j + 1 >= g.win && set_insert(a.gset, a.glog2, key)) {
                            ++acc_pos;  // distinct keys
                            spec_hist_add(a.hist, s_hsm, ab);
                        }
                    });
                } else if (MODE == kModePlainInsert) {
                    // filter.rs:224-233
                    for_each_smer(A0, A1, A2, A3, nbases, g.s, [&](uint32_t, uint64_t smer) {
                        uint64_t h = xs64(smer);
                        for (uint32_t t = 0; t < g.nh; ++t) {
                            sat_add_global<BITS>(a.table, cbase + (h & amask), 1);
                            if (t + 1 < g.nh) h = xs64(h);
                        }
                    });
                    acc_smers += nk + g.k - g.s;
                } else {
                    // filter.rs:478-507 + 562-571
                    WinMin<WINC> wm;
                    uint64_t j0 = uint64_t(t0 - 1 + hj);
                    for_each_smer(A0, A1, A2, A3, nbases, g.s, [&](uint32_t j, uint64_t smer) {
                        uint64_t h = xs64(smer);
                        uint32_t est = Ctr<BITS>::kMax;
                        for (uint32_t t = 0; t < g.nh; ++t) {
                            uint32_t v = load_ctr<BITS>(a.table, cbase + (h & amask));
                            est = v < est ? v : est;
                            if (t + 1 < g.nh) h = xs64(h);
                        }
                        uint32_t ab = wm.push(est, g.win);
                        if (j + 1 >= g.win) {
                            uint64_t q = j + 1 - g.win;
                            if (a.out) reinterpret_cast<typename Ctr<BITS>::T *>(a.out)[j0 + q] =
                                typename Ctr<BITS>::T(ab);
                            acc_pos += ab > 0;
                            acc_sum += ab;
                        }
                    });
                }
            }
        }
    }

    // Block reductions of the statistics, added into one of kSeg spread slots
    // (per-tile counts fit 32 bits; only the plain query's estimate sum needs 64).
    // The valid windows come per wave from the ballots. With records, the
    // s-mers need no sum: a super-k-mer's nk windows are the valid windows from
    // its start to the next boundary, so over the batch the records' k-mers are
    // its valid windows, and each record adds k - s s-mers to its k-mers.
    {
        uint32_t v1 = uint32_t(MODE == kModePlainQuery || MODE == kModePlainSpectrum ? acc_pos : acc_smers);
        uint64_t v2 = acc_sum;
        if (MODE != kModeRecords) {
            for (int off = 32; off > 0; off >>= 1) {
                v1 += __shfl_down(v1, off, 64);
                if (MODE == kModePlainQuery) v2 += __shfl_down(v2, off, 64);
            }
        }
        if (lane == 0) s_red[0][wid] = wave_valid, s_red[1][wid] = v1, s_red[2][wid] = MODE == kModePlainQuery ? v2 : 0;
    }
    __syncthreads();
    if (tid == 0) {
        uint64_t t[3] = {0, 0, 0};
        for (int q = 0; q < 3; ++q)
            for (int i = 0; i < nwave; ++i) t[q] += s_red[q][i];
        if (MODE == kModeRecords) t[1] = t[0] + uint64_t(g.k - g.s) * tile_records;
        unsigned long long *st = (unsigned long long *)a.stats + kStSpread + (blockIdx.x & (kSeg - 1));
        constexpr bool kQueried = MODE == kModePlainQuery || MODE == kModePlainSpectrum;
        if (t[0]) atomicAdd(&st[kSeg * (kQueried ? kSpQueried : kSpValid)], t[0]);
        if (MODE == kModePlainQuery) {
            if (t[1]) atomicAdd(&st[kSeg * kSpPositive], t[1]);
            if (t[2]) atomicAdd(&st[kSeg * kSpChecksum], t[2]);
        } else if (MODE == kModePlainSpectrum) {
            if (t[1]) atomicAdd((unsigned long long *)a.stats + kStDistinct, t[1]);
        } else if (t[1]) {
            atomicAdd(&st[kSeg * kSpSmers], t[1]);
        }
    }
    if (MODE == kModePlainSpectrum)  // the tile's small estimates into the histogram
        for (int i = tid; i < int(kHistSmall); i += kTileThreads)
            if (s_hsm[i]) atomicAdd(&a.hist.dense[i], (unsigned long long)s_hsm[i]);
}

// ----------------------------------------------------------------------------
// Bin planning: exclusive scan of per-bin record counts and the work list.
// ----------------------------------------------------------------------------
// Bin plan in two launches of nbins/1024 workgroups, one bin per thread: block
// sums of records and work items, then each block adds the sums of the blocks
// before it to its own exclusive scan and writes offsets, cursors and its part
// of the work list (a bin above per_item records becomes several items).
constexpr int kPlanThreads = 1024;

__device__ __forceinline__ uint32_t items_of(uint32_t c, uint32_t per_item, uint32_t per_shift) {
    return per_shift < 32 ? (c + per_item - 1) >> per_shift : (c + per_item - 1) / per_item;
}

// Inclusive scan over the workgroup of two values; returns the totals.
__device__ __forceinline__ void block_scan2(uint64_t &a, uint64_t &b, uint64_t *s_a, uint64_t *s_b, uint32_t tid,
                                           uint64_t &tot_a, uint64_t &tot_b) {
    const uint32_t lane = tid & 63, wid = tid >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t ya = __shfl_up(a, off, 64), yb = __shfl_up(b, off, 64);
        if (lane >= uint32_t(off)) a += ya, b += yb;
    }
    if (lane == 63) s_a[wid] = a, s_b[wid] = b;
    __syncthreads();
    uint64_t pa = 0, pb = 0;
    tot_a = tot_b = 0;
    for (uint32_t w = 0; w < kPlanThreads / 64; ++w) {
        if (w < wid) pa += s_a[w], pb += s_b[w];
        tot_a += s_a[w], tot_b += s_b[w];
    }
    a += pa, b += pb;
}

__global__ __launch_bounds__(kPlanThreads) void k_plan_sums(const uint32_t *__restrict__ bin_count, uint64_t nbins,
                                                            uint32_t per_item, uint32_t per_shift,
                                                            uint64_t *__restrict__ blk) {
    __shared__ uint64_t s_a[kPlanThreads / 64], s_b[kPlanThreads / 64];
    const uint32_t tid = threadIdx.x;
    const uint64_t bin = uint64_t(blockIdx.x) * kPlanThreads + tid;
    const uint32_t c = bin < nbins ? bin_count[bin] : 0;
    uint64_t a = c, b = items_of(c, per_item, per_shift), ta, tb;
    block_scan2(a, b, s_a, s_b, tid, ta, tb);
    if (tid == 0) blk[2 * blockIdx.x] = ta, blk[2 * blockIdx.x + 1] = tb;
}

// A work item: x = bin, y/z = its first and one-past-last record in the
// sorted array, w = the bin's k-mers (bit 31 set when the bin is split over
// several items), so that a bin kernel starts an item with one 16-byte load.
__global__ __launch_bounds__(kPlanThreads) void k_plan_write(const uint32_t *__restrict__ bin_count, uint64_t nbins,
                                                             uint32_t per_item, uint32_t per_shift,
                                                             const uint64_t *__restrict__ blk,
                                                             uint32_t *__restrict__ bin_off,
                                                             uint32_t *__restrict__ cursor,
                                                             uint4 *__restrict__ work,
                                                             const uint32_t *__restrict__ bin_kmers,
                                                             uint64_t *__restrict__ stats) {
    __shared__ uint64_t s_a[kPlanThreads / 64], s_b[kPlanThreads / 64];
    __shared__ uint64_t s_base[2];
    const uint32_t tid = threadIdx.x;
    if (tid < 64) {  // records and items of the blocks before this one
        uint64_t ra = 0, rb = 0;
        for (uint32_t q = tid; q < blockIdx.x; q += 64) ra += blk[2 * q], rb += blk[2 * q + 1];
        for (int off = 32; off > 0; off >>= 1) ra += __shfl_down(ra, off, 64), rb += __shfl_down(rb, off, 64);
        if (tid == 0) s_base[0] = ra, s_base[1] = rb;
    }
    const uint64_t bin = uint64_t(blockIdx.x) * kPlanThreads + tid;
    const uint32_t c = bin < nbins ? bin_count[bin] : 0;
    const uint32_t parts = items_of(c, per_item, per_shift);
    uint64_t a = c, b = parts, ta, tb;
    block_scan2(a, b, s_a, s_b, tid, ta, tb);  // (synchronizes: s_base is visible)
    const uint64_t ro = s_base[0] + a - c, io = s_base[1] + b - parts;
    if (bin < nbins) {
        bin_off[bin] = uint32_t(ro);
        cursor[bin] = uint32_t(ro);
        const uint32_t info = (bin_kmers ? bin_kmers[bin] & 0x7FFFFFFFu : 0u) | (parts > 1 ? 0x80000000u : 0u);
        for (uint32_t p = 0; p < parts; ++p) {
            const uint32_t r0 = uint32_t(ro) + p * per_item;
            work[io + p] = make_uint4(uint32_t(bin), r0, min(uint32_t(ro) + c, r0 + per_item), info);
        }
    }
    if (blockIdx.x == gridDim.x - 1 && tid == kPlanThreads - 1) {
        stats[kStNWork] = s_base[1] + tb;
        stats[kStWorkHead] = 0;
        stats[kStRecords] = s_base[0] + ta;
    }
}

// Counting sort of the records by bin without one global atomic per record:
// tiles of kPartTile records rank themselves in an LDS histogram and reserve
// one contiguous range per key with a single global atomicAdd. When there are
// more bins than LDS entries, a coarse pass (bin >> fine_bits) runs first; the
// fine pass then sees tiles whose records share a few coarse buckets.
constexpr int kPartTile = 8192;
constexpr int kPartThreads = 1024;
constexpr int kPartItems = kPartTile / kPartThreads;
constexpr uint32_t kPartRange = 2048;

// Records laid out as nseg segments of seg_cap slots; counts[s] are valid.
struct SegIn {
    const uint64_t *src;
    const uint32_t *counts;
    uint32_t nseg;
    uint64_t seg_cap;
};

// Key of a record: (bin >> key_shift) - lo, where lo is the first bin of the
// coarse bucket of the tile's first record (tile_based) or 0.
struct KeySpec {
    uint32_t bin_shift, key_shift, fine_bits, tile_based;
    uint32_t copies;  // counters per key (spreads contended coarse counters); 1 for bins
};

__device__ __forceinline__ uint32_t rec_bin(uint64_t r, uint32_t bin_shift) {
    return uint32_t((r >> 37) >> bin_shift);
}

// Numbers the non-empty tiles of all segments densely: s_pre[s] is the first
// tile of segment s (s_pre[nseg] the total), so that a grid-stride loop visits
// only tiles holding records. Run by the whole workgroup (nseg <= 64); s_n
// (if given) receives the record count of every segment.
__device__ __forceinline__ uint32_t dense_tiles(const SegIn &in, uint32_t *s_pre, uint32_t tid,
                                                uint32_t *s_n = nullptr) {
    if (tid < 64) {
        const uint32_t n = tid < in.nseg ? uint32_t(min<uint64_t>(in.counts[tid], in.seg_cap)) : 0;
        if (s_n) s_n[tid] = n;
        uint32_t incl = (n + kPartTile - 1) / kPartTile;
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t y = __shfl_up(incl, off, 64);
            if (tid >= uint32_t(off)) incl += y;
        }
        s_pre[tid + 1] = incl;
        if (tid == 0) s_pre[0] = 0;
    }
    __syncthreads();
    return s_pre[in.nseg];
}

// Segment of dense tile u (s_pre ascending): the last s with s_pre[s] <= u.
__device__ __forceinline__ uint32_t tile_segment(const uint32_t *s_pre, uint32_t nseg, uint32_t u) {
    uint32_t lo = 0, hi = nseg;  // s_pre[lo] <= u < s_pre[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (s_pre[mid] <= u) lo = mid; else hi = mid;
    }
    return lo;
}

// Records of dense tile u (s_pre and s_n from dense_tiles), item j of this
// thread at j * kPartThreads + tid; cnt = records in the tile. Indices are
// clamped to the tile (which holds at least one record), so the loads carry
// no condition and stay in flight until the records are used; thread 0's first
// item is the tile's first record.
__device__ __forceinline__ void load_part_tile(const SegIn &in, const uint32_t *s_pre, const uint32_t *s_n,
                                               uint32_t u, uint32_t tid, uint64_t (&r)[kPartItems],
                                               uint32_t &cnt) {
    const uint32_t seg = tile_segment(s_pre, in.nseg, u);
    const uint32_t t0 = (u - s_pre[seg]) * uint32_t(kPartTile);
    cnt = min<uint32_t>(kPartTile, s_n[seg] - t0);
    const uint64_t *tile = in.src + seg * in.seg_cap + t0;
#pragma unroll
    for (int j = 0; j < kPartItems; ++j) r[j] = tile[min(uint32_t(j) * kPartThreads + tid, cnt - 1)];
}

// The key window of a tile whose first record is r0 (see KeySpec).
__device__ __forceinline__ uint32_t window_lo(uint64_t r0, const KeySpec &ks) {
    return ks.tile_based ? (rec_bin(r0, ks.bin_shift) >> ks.fine_bits) << ks.fine_bits : 0u;
}

// Histogram of keys (records per key, and k-mers per key when kmers != null).
// The workgroup's tiles u = blockIdx.x + i * gridDim.x share u % copies (the
// grid is a multiple of the copy count), so counts accumulate in LDS across
// tiles and are flushed once per run of tiles with the same key window.
__global__ __launch_bounds__(kPartThreads) void k_seg_hist(SegIn in, KeySpec ks, uint32_t *hist,
                                                           uint32_t *kmers) {
    __shared__ uint32_t s_h[kPartRange], s_k[kPartRange];
    __shared__ uint32_t s_lo;
    __shared__ uint32_t s_pre[65], s_n[64];
    const uint32_t tid = threadIdx.x;
    const uint32_t total = dense_tiles(in, s_pre, tid, s_n);
    const uint32_t copy = blockIdx.x % ks.copies;
    uint32_t lo = 0xFFFFFFFFu;  // key window of the counts held in LDS (none yet)
    auto flush = [&]() {
        for (uint32_t i = tid; i < kPartRange; i += kPartThreads) {
            if (s_h[i]) atomicAdd(&hist[(lo + i) * ks.copies + copy], s_h[i]);
            if (kmers && s_k[i]) atomicAdd(&kmers[lo + i], s_k[i]);
            s_h[i] = 0, s_k[i] = 0;
        }
    };
    for (uint32_t i = tid; i < kPartRange; i += kPartThreads) s_h[i] = 0, s_k[i] = 0;
    // the next tile's records load while the current one is counted
    uint64_t rn[kPartItems];
    uint32_t cnt_n = 0;
    if (blockIdx.x < total) load_part_tile(in, s_pre, s_n, blockIdx.x, tid, rn, cnt_n);
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        uint64_t r[kPartItems];
#pragma unroll
        for (int j = 0; j < kPartItems; ++j) r[j] = rn[j];
        const uint32_t cnt = cnt_n;
        if (u + gridDim.x < total) load_part_tile(in, s_pre, s_n, u + gridDim.x, tid, rn, cnt_n);
        if (tid == 0) s_lo = window_lo(r[0], ks);
        __syncthreads();
        const uint32_t tlo = s_lo;
        if (tlo != lo) {  // uniform: a new key window
            if (lo != 0xFFFFFFFFu) flush();
            lo = tlo;
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < kPartItems; ++j) {
            if (uint32_t(j) * kPartThreads + tid >= cnt) continue;
            const uint32_t key = rec_bin(r[j], ks.bin_shift) >> ks.key_shift;
            const uint32_t nk = uint32_t((r[j] >> 32) & 31) + 1;
            if (key - lo < kPartRange) {
                atomicAdd(&s_h[key - lo], 1u);
                if (kmers) atomicAdd(&s_k[key - lo], nk);
            } else {
                atomicAdd(&hist[key * ks.copies + copy], 1u);
                if (kmers) atomicAdd(&kmers[key], nk);
            }
        }
        __syncthreads();
    }
    if (lo != 0xFFFFFFFFu) flush();
}

// Exclusive scan of a histogram into cursors (one workgroup; each thread scans
// a contiguous run of n/1024 entries); the total becomes the record count.
__global__ __launch_bounds__(1024) void k_small_scan(const uint32_t *hist, uint32_t n, uint32_t *cursor,
                                                     uint64_t *stats, uint32_t *total_out) {
    __shared__ uint32_t s_wave[16];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t b0 = tid * per, b1 = min(n, b0 + per);
    uint32_t sum = 0;
    for (uint32_t i = b0; i < b1; ++i) sum += hist[i];
    // exclusive prefix of the per-thread sums: wave scans, then the wave totals
    uint32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= uint32_t(off)) incl += y;
    }
    if (lane == 63) s_wave[wid] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (uint32_t i = 0; i < 16; ++i) {
        const uint32_t w = s_wave[i];
        before += i < wid ? w : 0u;
        total += w;
    }
    uint32_t run = before + incl - sum;
    for (uint32_t i = b0; i < b1; ++i) {
        cursor[i] = run;
        run += hist[i];
    }
    if (tid == 0) {
        stats[kStRecords] = total;
        *total_out = total;
    }
}

// Persistent: a workgroup takes tiles u = blockIdx.x + i * gridDim.x and
// loads the records of its next tile into registers while the current one is
// ranked, reserved and scattered, so neither the record loads nor the cursor
// reservations of consecutive tiles wait on each other.
__global__ __launch_bounds__(kPartThreads) void k_partition(SegIn in, KeySpec ks, uint32_t *cursor,
                                                            uint64_t *dst) {
    __shared__ uint32_t s_hist[kPartRange];
    __shared__ uint32_t s_lo;
    __shared__ uint32_t s_pre[65], s_n[64];
    const uint32_t tid = threadIdx.x;
    const uint32_t total = dense_tiles(in, s_pre, tid, s_n);
    uint64_t rn[kPartItems];
    uint32_t cnt_n = 0;
    if (blockIdx.x < total) load_part_tile(in, s_pre, s_n, blockIdx.x, tid, rn, cnt_n);
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        uint64_t r[kPartItems];
#pragma unroll
        for (int j = 0; j < kPartItems; ++j) r[j] = rn[j];
        const uint32_t cnt = cnt_n;
        if (u + gridDim.x < total) load_part_tile(in, s_pre, s_n, u + gridDim.x, tid, rn, cnt_n);
        for (uint32_t i = tid; i < kPartRange; i += kPartThreads) s_hist[i] = 0;
        if (tid == 0) s_lo = window_lo(r[0], ks);
        __syncthreads();
        const uint32_t lo = s_lo;
        // key << 13 | rank in the tile's run of that key (key < 2^11, rank < 2^13)
        static_assert(kPartRange <= (1u << 11) && kPartTile <= (1 << 13), "key and rank share a word");
        uint32_t kr[kPartItems];
#pragma unroll
        for (int j = 0; j < kPartItems; ++j) {
            uint32_t i = uint32_t(j) * kPartThreads + tid;
            kr[j] = 0xFFFFFFFFu;
            if (i < cnt) {
                uint32_t kk = (rec_bin(r[j], ks.bin_shift) >> ks.key_shift) - lo;
                if (kk < kPartRange) {
                    kr[j] = kk << 13 | atomicAdd(&s_hist[kk], 1u);
                } else {  // outside the tile's range: reserve one slot directly
                    dst[atomicAdd(&cursor[(kk + lo) * ks.copies + u % ks.copies], 1u)] = r[j];
                }
            }
        }
        __syncthreads();
        const uint32_t copy = uint32_t(u % ks.copies);
        for (uint32_t kk = tid; kk < kPartRange; kk += kPartThreads) {
            uint32_t c = s_hist[kk];
            if (c) s_hist[kk] = atomicAdd(&cursor[(kk + lo) * ks.copies + copy], c);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kPartItems; ++j)
            if (kr[j] != 0xFFFFFFFFu) dst[s_hist[kr[j] >> 13] + (kr[j] & 0x1FFFu)] = r[j];
        __syncthreads();
    }
}


// Copies nwords 32-bit words global -> LDS with U independent 16-byte loads in
// flight per thread (one memory round trip per U*16 bytes instead of one per
// word).
template <int U>
__device__ __forceinline__ void stage_to_lds(uint32_t *lds, const uint32_t *src, uint32_t nwords,
                                             uint32_t tid, uint32_t nthr) {
    if ((nwords & 3) == 0) {
        const uint4 *sv = reinterpret_cast<const uint4 *>(src);
        uint4 *dv = reinterpret_cast<uint4 *>(lds);
        const uint32_t nvec = nwords >> 2;
        uint32_t v0 = tid;
        // whole rounds carry no conditions, so the compiler keeps all U loads
        // in flight instead of sinking each into its guarded store
        for (; v0 + (U - 1) * nthr < nvec; v0 += nthr * U) {
            uint4 x[U];
#pragma unroll
            for (int u = 0; u < U; ++u) x[u] = sv[v0 + u * nthr];
#pragma unroll
            for (int u = 0; u < U; ++u) dv[v0 + u * nthr] = x[u];
        }
        for (; v0 < nvec; v0 += nthr) dv[v0] = sv[v0];
    } else {
        for (uint32_t wi = tid; wi < nwords; wi += nthr) lds[wi] = src[wi];
    }
}

__device__ __forceinline__ void zero_lds(uint32_t *lds, uint32_t nwords, uint32_t tid, uint32_t nthr) {
    if ((nwords & 3) == 0) {
        uint4 *dv = reinterpret_cast<uint4 *>(lds);
        for (uint32_t i = tid; i < (nwords >> 2); i += nthr) dv[i] = make_uint4(0, 0, 0, 0);
    } else {
        for (uint32_t i = tid; i < nwords; i += nthr) lds[i] = 0;
    }
}

// Two 16-bit counters per word plus two 16-bit counts, each saturated at 65535
// (one packed add with clamp).
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t sat_add_u16x2(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t,
                              __builtin_elementwise_add_sat(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b)));
}

// One 32-bit table word plus the increment counts of its counters, saturated
// per counter (counter.rs:20-23 applied count times).
template <int BITS>
__device__ __forceinline__ uint32_t sat_merge_word(uint32_t old, const uint32_t *c) {
    if (BITS == 32) {
        uint64_t sum = uint64_t(old) + c[0];
        return sum > 0xFFFFFFFFull ? 0xFFFFFFFFu : uint32_t(sum);
    } else {
        constexpr uint32_t per = 32 / BITS;
        constexpr uint32_t mx = Ctr<BITS>::kMax;
        uint32_t nv = 0;
#pragma unroll
        for (uint32_t q = 0; q < per; ++q) {
            uint32_t cur = (old >> (q * BITS)) & mx;
            uint32_t sum = cur + (c[q] > mx ? mx : c[q]);
            nv |= (sum > mx ? mx : sum) << (q * BITS);
        }
        return nv;
    }
}

// Work items hold at most this many records (their order lives in LDS as u16).
constexpr uint32_t kMaxItemRecords = 4096;
// Workgroup sizes of the bin kernels (their launch bounds).
constexpr int kInsertThreads = 1024, kQueryThreads = 512, kSpectrumThreads = 512;

// Orders the records [rb, rb + n) of a work item by k-mer count, longest
// first, into s_perm (indices relative to rb). A wave's s-mer loop runs as
// long as its longest record; with the lanes of a wave on records of equal
// length no lane idles. Lanes of equal length are matched with five ballots
// and counted by one leader each, so the LDS atomics never collide. R rounds
// of nthr threads cover the largest item; all R keys are loaded up front.
// With s_lead given, only the records i with s_lead[i] == i are ordered, and
// their number is written to *s_total. order_by_hi takes the records' high
// words (hi[r] of record r * nthr + tid; any value past n) from registers.
template <int R>
__device__ __forceinline__ void order_by_hi(const uint32_t (&hi)[R], uint32_t n, uint16_t *s_perm, uint32_t *s_cls,
                                            uint32_t tid, uint32_t nthr, uint32_t *s_smers = nullptr,
                                            uint32_t k_minus_s = 0, const uint16_t *s_lead = nullptr,
                                            uint32_t *s_total = nullptr) {
    const uint32_t lane = tid & 63;
    const uint64_t lt = (uint64_t(1) << lane) - 1;
    if (tid < 32) s_cls[tid] = 0;
    uint32_t key[R];
#pragma unroll
    for (int r = 0; r < R; ++r) key[r] = n ? 31 - (hi[r] & 31) : 0;  // longest first
    __syncthreads();
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (uint32_t(r) * nthr >= n) break;  // uniform
            const uint32_t i = r * nthr + tid;
            const bool act = i < n && (!s_lead || s_lead[i] == i);
            uint64_t m = __ballot(act);
#pragma unroll
            for (int bit = 0; bit < 5; ++bit) {
                const uint64_t b = __ballot((key[r] >> bit) & 1);
                m &= ((key[r] >> bit) & 1) ? b : ~b;
            }
            const uint32_t rank = uint32_t(__popcll(m & lt));
            const uint32_t leader = act ? uint32_t(__builtin_ctzll(m)) : 0;
            uint32_t base = 0;
            if (act && rank == 0) base = atomicAdd(&s_cls[key[r]], uint32_t(__popcll(m)));
            base = __shfl(base, int(leader), 64);
            if (pass == 1 && act) s_perm[base + rank] = uint16_t(i);
        }
        __syncthreads();
        if (pass == 0) {
            if (tid < 64) {  // the first wave: class c in lane c, prefix by shuffles
                const uint32_t v = lane < 32 ? s_cls[lane] : 0u;
                uint32_t run = v;
#pragma unroll
                for (int o = 1; o < 32; o <<= 1) {
                    const uint32_t y = __shfl_up(run, o, 64);
                    run += lane >= uint32_t(o) ? y : 0u;
                }
                if (lane < 32) s_cls[lane] = run - v;
                uint32_t smers = lane < 32 ? v * (32 - lane + k_minus_s) : 0u;  // class c: records of 32 - c k-mers
#pragma unroll
                for (int o = 16; o > 0; o >>= 1) smers += __shfl_xor(smers, o, 64);
                const uint32_t total = __shfl(run, 31, 64);
                if (lane == 0) {
                    if (s_smers) *s_smers = smers;
                    if (s_total) *s_total = total;
                }
            }
            __syncthreads();
        }
    }
}

// The high words of records [rb, rb + n) for order_by_hi (record r * nthr + tid
// in hi[r]; the clamped index keeps the load in bounds).
template <int R>
__device__ __forceinline__ void load_hi(const uint64_t *sorted, uint32_t rb, uint32_t n, uint32_t (&hi)[R],
                                        uint32_t tid, uint32_t nthr) {
#pragma unroll
    for (int r = 0; r < R; ++r) hi[r] = n ? uint32_t(sorted[rb + min(r * nthr + tid, n - 1)] >> 32) : 0u;
}
template <int R>
__device__ __forceinline__ void order_by_length(const uint64_t *sorted, uint32_t rb, uint32_t n, uint16_t *s_perm,
                                                uint32_t *s_cls, uint32_t tid, uint32_t nthr,
                                                uint32_t *s_smers = nullptr, uint32_t k_minus_s = 0) {
    uint32_t hi[R];
    load_hi<R>(sorted, rb, n, hi, tid, nthr);
    order_by_hi<R>(hi, n, s_perm, s_cls, tid, nthr, s_smers, k_minus_s);
}

// The five packed words holding bases [b & ~15, (b & ~15) + 80) and the four
// words of 16 bases each starting exactly at base b.
// One 16-byte load (dword-aligned: global memory allows it), and the fifth
// word only for a record whose bases reach past the first four: one or two
// memory requests per record instead of five.
__device__ __forceinline__ void load_words5(const uint32_t *pk, uint64_t rec, uint32_t k, uint32_t (&W)[5]) {
    struct __attribute__((aligned(4))) Q {
        uint32_t x, y, z, w;
    } q;
    const uint64_t b = rec & 0xFFFFFFFFull, wi = b >> 4;
    const uint32_t end = uint32_t(b & 15) + uint32_t((rec >> 32) & 31) + k;  // past the record's last base
    __builtin_memcpy(&q, pk + wi, sizeof(q));
    W[0] = q.x, W[1] = q.y, W[2] = q.z, W[3] = q.w;
    W[4] = end > 64 ? pk[wi + 4] : 0u;
}
__device__ __forceinline__ void align_words(const uint32_t (&W)[5], uint64_t b, uint32_t &A0, uint32_t &A1,
                                            uint32_t &A2, uint32_t &A3) {
    const uint32_t sh = 2 * uint32_t(b & 15);
    A0 = __builtin_amdgcn_alignbit(W[1], W[0], sh);
    A1 = __builtin_amdgcn_alignbit(W[2], W[1], sh);
    A2 = __builtin_amdgcn_alignbit(W[3], W[2], sh);
    A3 = __builtin_amdgcn_alignbit(W[4], W[3], sh);
}

// Walks the records [rb, rb + n) of a work item in the order of s_perm, one
// per thread per round, calling f(record, packed words). The next record and
// its packed words are in flight while the current one is processed: a record
// and its bases are two dependent global loads, which would otherwise stall
// every round. Indices past the end are clamped, so every load is in bounds.
template <class F>
__device__ __forceinline__ void walk_records(const uint64_t *sorted, const uint32_t *pk, uint32_t k, uint32_t rb, uint32_t n,
                                             const uint16_t *s_perm, uint32_t tid, uint32_t nthr, F &&f) {
    if (tid >= n) return;
    const uint32_t last = n - 1;
    uint64_t rc = sorted[rb + s_perm[tid]];
    uint64_t rn = sorted[rb + s_perm[min(tid + nthr, last)]];
    uint32_t W[5];
    load_words5(pk, rc, k, W);
    for (uint32_t i = tid; i < n; i += nthr) {
        uint32_t V[5];
        load_words5(pk, rn, k, V);
        const uint32_t pnn = s_perm[min(i + 2 * nthr, last)];
        const uint64_t rnn = sorted[rb + pnn];
        f(rc, W);
        rc = rn;
        rn = rnn;
#pragma unroll
        for (int q = 0; q < 5; ++q) W[q] = V[q];
    }
}

// walk_records for the insert: f(record high word, bases from the first in
// A0..A3, index in the item). With use_first, this thread's first record
// (s_perm[tid]) is passed in registers (its bases first, its high word
// first_hi) and neither it nor its bases are read from HBM.
template <class F>
__device__ __forceinline__ void walk_records_aligned(const uint64_t *sorted, const uint32_t *pk, uint32_t k, uint32_t rb,
                                                     uint32_t n, const uint16_t *s_perm, uint32_t tid, uint32_t nthr,
                                                     F &&f, bool use_first, uint4 first, uint32_t first_hi) {
    if (tid >= n) return;
    const uint32_t last = n - 1;
    uint32_t pc = s_perm[tid], pn = s_perm[min(tid + nthr, last)];
    uint64_t rn = sorted[rb + pn];
    uint32_t hc, A0, A1, A2, A3;
    if (use_first) {
        hc = first_hi, A0 = first.x, A1 = first.y, A2 = first.z, A3 = first.w;
    } else {
        const uint64_t rc = sorted[rb + pc];
        uint32_t W[5];
        load_words5(pk, rc, k, W);
        align_words(W, rc & 0xFFFFFFFFull, A0, A1, A2, A3);
        hc = uint32_t(rc >> 32);
    }
    for (uint32_t i = tid; i < n; i += nthr) {
        uint32_t V[5];
        load_words5(pk, rn, k, V);
        const uint32_t pnn = s_perm[min(i + 2 * nthr, last)];
        const uint64_t rnn = sorted[rb + pnn];
        f(hc, A0, A1, A2, A3, pc);
        align_words(V, rn & 0xFFFFFFFFull, A0, A1, A2, A3);
        hc = uint32_t(rn >> 32);
        rn = rnn;
        pc = pn;
        pn = pnn;
    }
}

struct BinArgs {
    const uint64_t *sorted;
    const uint32_t *bin_off, *bin_count;
    const uint4 *work;          // work items (k_plan_write)
    const uint32_t *bin_kmers;  // k-mers per bin (insert: selects the items counted by distinct record)
    uint64_t *stats;
    const uint32_t *pk;
    void *table;
    void *out;
    uint32_t bin_log;      // slots per bin = 2^bin_log
    uint32_t per_item;
    uint32_t fresh;        // insert into a table whose contents are undefined (see k_zero_cold)
};

// Before the first insert into an uncleared table: zero the bins that the
// insert kernel will not write whole, i.e. bins without records and bins split
// over several work items (those merge with atomics into the table).
__global__ __launch_bounds__(256) void k_zero_cold(const uint32_t *bin_count, uint64_t nbins, uint32_t per_item,
                                                   uint32_t *table, uint32_t words_per_bin) {
    for (uint64_t b = blockIdx.x; b < nbins; b += gridDim.x) {
        uint32_t c = bin_count[b];
        if (c != 0 && c <= per_item) continue;
        uint32_t *p = table + b * words_per_bin;
        if ((words_per_bin & 3) == 0) {
            uint4 *v = reinterpret_cast<uint4 *>(p);
            for (uint32_t i = threadIdx.x; i < words_per_bin / 4; i += blockDim.x) v[i] = make_uint4(0, 0, 0, 0);
        } else {
            for (uint32_t i = threadIdx.x; i < words_per_bin; i += blockDim.x) p[i] = 0;
        }
    }
}

// Insert work items hold at most this many records (two per thread), so that
// the distinct-record search below fits in the LDS of the bin's counts.
constexpr uint32_t kInsertItemRecords = 2 * kInsertThreads;
constexpr uint32_t kRecDedupSlots = 2 * kInsertItemRecords;  // open addressing, load <= 1/2
// 16-byte keys and the high words of the records, then the table (words)
constexpr uint32_t kRecDedupWords = 5 * kInsertItemRecords + kRecDedupSlots;
static_assert(kInsertItemRecords <= 4096, "table entries hold 12-bit record indices");

// Identical records of a work item: the same bases, k-mer count and block
// (the record's high word), hence the same s-mers at the same addresses.
// s_lead[i] becomes the first record of i's key to claim a table slot (i for
// that record, the leader), and s_mult the number of records of each leader
// (16-bit halves of words, record l in half l & 1 of word l >> 1). Identical
// keys walk the same probe sequence and a slot is never released, so each key
// has exactly one leader; at most n of the 2n slots are taken, so every probe
// ends. Record i's key (its bases, aligned and masked to its length) stays at
// dedup_key(s_ws, i) and its high word at dedup_hi(s_ws, i) until s_ws is
// written again.
// The insert's search pays while distinct records are a minority (50-60x
// coverage of short reads: ~40% of a bin's records); long noisy reads make
// most records distinct, and then it costs more than it saves. A workgroup
// whose item had more than kDedupMaxLeaders / 8 of its records distinct skips
// the search for its next items, 8 after the first such item and twice as many
// after each further one (at most 256), until an item with fewer distinct
// records resets it (decided per workgroup, from its own items).
constexpr uint32_t kDedupMaxLeaders = 5;  // eighths

__device__ __forceinline__ const uint4 &dedup_key(const uint32_t *s_ws, uint32_t i) {
    return reinterpret_cast<const uint4 *>(s_ws)[i];
}
__device__ __forceinline__ uint32_t dedup_hi(const uint32_t *s_ws, uint32_t i) {
    return s_ws[4 * kInsertItemRecords + i];
}
template <int R>
__device__ __forceinline__ void find_duplicates(const uint64_t *sorted, const uint32_t *pk, uint32_t rb, uint32_t n,
                                                uint32_t k, uint32_t *s_ws, uint16_t *s_lead, uint32_t *s_mult,
                                                uint32_t tid, uint32_t nthr, uint32_t (&hi_out)[R]) {
    uint4 *s_key = reinterpret_cast<uint4 *>(s_ws);
    uint32_t *s_hi = s_ws + 4 * kInsertItemRecords;
    uint32_t *s_ht = s_hi + kInsertItemRecords;
    for (uint32_t i = tid; i < kRecDedupSlots; i += nthr) s_ht[i] = 0;
    for (uint32_t i = tid; i < kInsertItemRecords / 2; i += nthr) s_mult[i] = 0;
    uint32_t h[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = r * nthr + tid;
        const uint64_t rec = sorted[rb + min(i, n - 1)];  // clamped: in bounds
        uint32_t W[5];
        load_words5(pk, rec, k, W);
        uint32_t A[4];
        align_words(W, rec & 0xFFFFFFFFull, A[0], A[1], A[2], A[3]);
        const uint32_t hi = uint32_t(rec >> 32);
        hi_out[r] = hi;
        const int bits = 2 * int((hi & 31) + k);  // 2 bits for each of the nk + k - 1 bases
        uint32_t x = hi * 0x9E3779B1u;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int b = bits - 32 * t;
            A[t] &= b >= 32 ? ~0u : b <= 0 ? 0u : (1u << b) - 1;
            x = (x ^ A[t]) * 0x85EBCA77u;
            x ^= x >> 15;
        }
        h[r] = x;
        if (i < n) {
            s_key[i] = make_uint4(A[0], A[1], A[2], A[3]);
            s_hi[i] = hi;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = r * nthr + tid;
        if (i >= n) continue;
        const uint32_t entry = 0x80000000u | (h[r] & 0x7FFFF000u) | i;  // valid | tag | index
        uint32_t slot = h[r] & (kRecDedupSlots - 1), lead = i;
        for (;;) {
            const uint32_t old = atomicCAS(&s_ht[slot], 0u, entry);
            if (old == 0) break;
            if (((old ^ entry) & 0x7FFFF000u) == 0) {
                const uint32_t li = old & 0xFFFu;
                const uint4 p = s_key[li], q = s_key[i];
                if (s_hi[li] == s_hi[i] && p.x == q.x && p.y == q.y && p.z == q.z && p.w == q.w) {
                    lead = li;
                    break;
                }
            }
            slot = (slot + 1) & (kRecDedupSlots - 1);
        }
        s_lead[i] = uint16_t(lead);
        atomicAdd(&s_mult[lead >> 1], 1u << (16 * (lead & 1)));
    }
    __syncthreads();
}

// filter.rs:199-237 for every super-k-mer of one bin: n_hashes saturating +1
// per s-mer, counted in LDS (u32 per slot) and merged once as
// min(MAX, old + count), which equals applying the +1s one by one (saturating
// increments commute).
template <int BITS>
// Counts are 16-bit halves of LDS words (a 2^15-slot bin is 64 KiB, so two
// workgroups share a CU): slot s lives in word s mod (slots/2), low half for
// the first half of the bin and high half for the second. With several blocks
// per bin a record's half is fixed by its block, so an increment costs what a
// 32-bit counter would. A work item whose records make at most 65535
// increments cannot carry out of a half; larger items are counted and merged
// in sub-batches of records that each stay within that bound.
__global__ __launch_bounds__(kInsertThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_bin_insert(BinArgs a,
                                                                                                       DevGeo g) {
    extern __shared__ __align__(16) uint32_t s_cnt[];
    __shared__ uint64_t s_item;
    __shared__ uint16_t s_perm[kInsertItemRecords], s_lead[kInsertItemRecords];
    __shared__ uint32_t s_mult[kInsertItemRecords / 2];
    __shared__ uint32_t s_cls[32];
    __shared__ uint32_t s_smers, s_nwalk;
    const uint32_t tid = threadIdx.x, nthr = blockDim.x;
    const uint32_t bin_slots = 1u << a.bin_log;
    const uint32_t lshift = a.bin_log - g.bse;
    const uint32_t amask = (1u << g.bse) - 1;
    const uint64_t nwork = a.stats[kStNWork];
    unsigned long long *head = (unsigned long long *)&a.stats[kStWorkHead];
    uint32_t skip = 0, backoff = 4;  // items left that skip the distinct-record search (uniform)
    // records per sub-batch: a record has at most w k-mers, so w + k - s s-mers
    const uint32_t rec_incr = (g.w + g.k - g.s) * g.nh;
    const bool one_word = smer_tail_one_word(g);
    const uint32_t sub_recs = max(1u, 65535u / rec_incr);
    constexpr uint32_t per = 32 / BITS;
    const uint32_t nwords = bin_slots / per;
    const uint32_t half_log = a.bin_log - 1, half_words = bin_slots / 2;
    // 16-bit count of slot sl
    auto count_of = [&](uint32_t sl) { return (s_cnt[sl & (half_words - 1)] >> (16 * (sl >> half_log))) & 0xFFFFu; };
    for (;;) {
        if (tid == 0) s_item = atomicAdd(head, 1ull);
        __syncthreads();
        uint64_t item = s_item;
        if (item >= nwork) break;
        const uint4 wk = a.work[item];
        const uint64_t bin = wk.x;
        const uint32_t rb = wk.y, re = wk.z;
        const uint32_t n = re - rb;
        const bool multi = wk.w >> 31;
        // A bin that is one work item and makes at most 65535 increments (its
        // k-mers plus k - s per record, times n_hashes) is counted by distinct
        // record: each leader adds its multiplicity c at every increment, which
        // equals c single increments (and stays within one 16-bit pass).
        const bool dedup = skip == 0 && !multi && n > 1 && a.bin_kmers &&
                           (uint64_t(wk.w & 0x7FFFFFFFu) + uint64_t(g.k - g.s) * n) * g.nh <= 65535u;
        skip -= skip > 0;
        uint32_t nwalk = n;
        bool one = true;
        // with duplicates found, the leader this thread walks first (s_perm
        // entry tid): its bases and high word
        uint4 lkey = make_uint4(0, 0, 0, 0);
        uint32_t lhi = 0;
        if (dedup) {
            constexpr int R = kInsertItemRecords / kInsertThreads;
            uint32_t hi[R];
            find_duplicates<R>(a.sorted, a.pk, rb, n, g.k, s_cnt, s_lead, s_mult, tid, nthr, hi);  // (synchronizes)
            order_by_hi<R>(hi, n, s_perm, s_cls, tid, nthr, nullptr, 0, s_lead, &s_nwalk);  // (synchronizes)
            nwalk = s_nwalk;
            if (8 * nwalk > kDedupMaxLeaders * n) backoff = min(2 * backoff, 256u), skip = backoff;
            else backoff = 4;
            // the keys share the LDS of the counts: the first round of the walk
            // takes its leaders' bases from them into registers (no second read
            // from HBM) before the counts are cleared
            if (tid < nwalk) {
                const uint32_t li = s_perm[tid];
                lkey = dedup_key(s_cnt, li), lhi = dedup_hi(s_cnt, li);
            }
            __syncthreads();
            zero_lds(s_cnt, bin_slots / 2, tid, nthr);
            __syncthreads();
        } else {
            zero_lds(s_cnt, bin_slots / 2, tid, nthr);
            order_by_length<kInsertItemRecords / kInsertThreads>(a.sorted, rb, n, s_perm, s_cls, tid, nthr,
                                                                 &s_smers, g.k - g.s);  // (synchronizes)
            one = uint64_t(s_smers) * g.nh <= 65535u;
        }
        const uint32_t step = one ? nwalk : sub_recs;
        uint32_t *tw = reinterpret_cast<uint32_t *>(a.table) + bin * uint64_t(nwords);
        // the n_hashes increments of every s-mer of one record, mult each (the
        // number of identical records); A0..A3 hold its bases from the first,
        // hi is the record's high word
        auto count_record = [&](uint32_t A0, uint32_t A1, uint32_t A2, uint32_t A3, uint32_t hi, uint32_t mult) {
            const uint32_t nk = (hi & 31) + 1;
            const uint32_t lbase = ((hi >> 5) & ((1u << lshift) - 1)) << g.bse;
            with_nh(g.nh, [&](auto nht) {
                constexpr int NH = decltype(nht)::value;
                if (lshift > 0) {  // the record's block lies in one half of the bin
                    const uint32_t wbase = lbase & (half_words - 1), inc = mult << (16 * (lbase >> half_log));
                    for_each_smer(
                        A0, A1, A2, A3, nk + g.k - 1, g.s,
                        [&](uint32_t, uint64_t smer) {
                            for_each_address_n<NH>(smer, g.nh, amask,
                                                   [&](uint32_t ad) { atomicAdd(&s_cnt[wbase + ad], inc); });
                        },
                        one_word);
                } else {  // one block per bin: its slots span both halves
                    for_each_smer(
                        A0, A1, A2, A3, nk + g.k - 1, g.s,
                        [&](uint32_t, uint64_t smer) {
                            for_each_address_n<NH>(smer, g.nh, amask, [&](uint32_t ad) {
                                atomicAdd(&s_cnt[ad & (half_words - 1)], mult << (16 * (ad >> half_log)));
                            });
                        },
                        one_word);
                }
            });
        };
        for (uint32_t r0 = 0; r0 < nwalk; r0 += step) {
            // with duplicates (one pass, step = nwalk), the first round's leaders
            // come from registers and later rounds (rare: an item has at most
            // 2 nthr records) from HBM
            walk_records_aligned(
                a.sorted, a.pk, g.k, rb, min(step, nwalk - r0), s_perm + r0, tid, nthr,
                [&](uint32_t hi, uint32_t A0, uint32_t A1, uint32_t A2, uint32_t A3, uint32_t ri) {
                    count_record(A0, A1, A2, A3, hi, dedup ? (s_mult[ri >> 1] >> (16 * (ri & 1))) & 0xFFFFu : 1u);
                },
                dedup, lkey, lhi);
            __syncthreads();
            // merge this batch's counts: 32-bit table words of the bin
            const bool write_all = a.fresh && r0 == 0;  // the table holds no counts yet
            if (!multi && (nwords & 7) == 0) {
                // 16-byte read-modify-write with kU loads in flight per thread
                // (at least two vectors per bin, so none straddles the halves)
                constexpr int kU = BITS == 16 ? 4 : 2;  // fewer registers for the u8 and u32 merges
                uint4 *tv = reinterpret_cast<uint4 *>(tw);
                const uint32_t nvec = nwords >> 2;
                for (uint32_t v0 = tid; v0 < nvec; v0 += nthr * kU) {
                    uint4 old[kU];
                    if (write_all) {
#pragma unroll
                        for (int u = 0; u < kU; ++u) old[u] = make_uint4(0, 0, 0, 0);
                    } else {
                        // all in flight; a vector none of whose slots received a
                        // count is neither read nor written (a sparse batch into a
                        // large table touches a fraction of it)
#pragma unroll
                        for (int u = 0; u < kU; ++u) {
                            const uint32_t vi = min(v0 + u * nthr, nvec - 1);
                            const uint32_t s0 = vi * 4 * per, sh = 16 * (s0 >> half_log);
                            const uint4 *cw = reinterpret_cast<const uint4 *>(s_cnt + (s0 & (half_words - 1)));
                            uint32_t orw = 0;
#pragma unroll
                            for (uint32_t q = 0; q < per; ++q) {
                                const uint4 x = cw[q];
                                orw |= x.x | x.y | x.z | x.w;
                            }
                            old[u] = ((orw >> sh) & 0xFFFFu) ? tv[vi] : make_uint4(0, 0, 0, 0);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kU; ++u) {
                        const uint32_t vi = v0 + u * nthr;
                        if (vi >= nvec) continue;
                        // the vector's 4 * per slots are consecutive words of one half
                        uint32_t c[4 * per];
                        uint32_t any = write_all;
                        const uint32_t s0 = vi * 4 * per, sh = 16 * (s0 >> half_log);
                        const uint4 *cw = reinterpret_cast<const uint4 *>(s_cnt + (s0 & (half_words - 1)));
                        if (BITS == 16) {
                            // counts of slots s0..s0+7: the low (first half of the
                            // bin) or high halves of eight LDS words, paired into
                            // table order by a byte permute, then one clamped
                            // packed add per table word
                            const uint4 x0 = cw[0], x1 = cw[1];
                            const uint32_t sel = sh ? 0x07060302u : 0x05040100u;
                            const uint32_t p0 = __builtin_amdgcn_perm(x0.y, x0.x, sel);
                            const uint32_t p1 = __builtin_amdgcn_perm(x0.w, x0.z, sel);
                            const uint32_t p2 = __builtin_amdgcn_perm(x1.y, x1.x, sel);
                            const uint32_t p3 = __builtin_amdgcn_perm(x1.w, x1.z, sel);
                            if (!(any | p0 | p1 | p2 | p3)) continue;
                            tv[vi] = make_uint4(sat_add_u16x2(old[u].x, p0), sat_add_u16x2(old[u].y, p1),
                                                sat_add_u16x2(old[u].z, p2), sat_add_u16x2(old[u].w, p3));
                            continue;
                        }
#pragma unroll
                        for (uint32_t q = 0; q < per; ++q) {
                            const uint4 x = cw[q];
                            c[4 * q] = (x.x >> sh) & 0xFFFFu, c[4 * q + 1] = (x.y >> sh) & 0xFFFFu;
                            c[4 * q + 2] = (x.z >> sh) & 0xFFFFu, c[4 * q + 3] = (x.w >> sh) & 0xFFFFu;
                            any |= c[4 * q] | c[4 * q + 1] | c[4 * q + 2] | c[4 * q + 3];
                        }
                        if (!any) continue;
                        uint4 nv;
                        nv.x = sat_merge_word<BITS>(old[u].x, c);
                        nv.y = sat_merge_word<BITS>(old[u].y, c + per);
                        nv.z = sat_merge_word<BITS>(old[u].z, c + 2 * per);
                        nv.w = sat_merge_word<BITS>(old[u].w, c + 3 * per);
                        tv[vi] = nv;
                    }
                }
            } else {
                for (uint32_t wi = tid; wi < nwords; wi += nthr) {
                    uint32_t c[per];
                    uint32_t any = !multi && write_all;
#pragma unroll
                    for (uint32_t q = 0; q < per; ++q) c[q] = count_of(wi * per + q), any |= c[q];
                    if (!any) continue;
                    if (!multi) {
                        tw[wi] = sat_merge_word<BITS>(write_all ? 0u : tw[wi], c);
                    } else {
#pragma unroll
                        for (uint32_t q = 0; q < per; ++q)
                            if (c[q])
                                sat_add_global<BITS>(a.table, (bin * uint64_t(nwords) + wi) * per + q,
                                                     c[q] > Ctr<BITS>::kMax && BITS < 32 ? Ctr<BITS>::kMax : c[q]);
                    }
                }
            }
            __syncthreads();
            if (r0 + step < nwalk) {
                zero_lds(s_cnt, bin_slots / 2, tid, nthr);
                __syncthreads();
            }
        }
    }
}

// Records per work item of the query kernel (its LDS holds the 64 KiB bin,
// the record order and a 16-byte estimate slot per thread: two workgroups
// share a CU).
constexpr uint32_t kQueryItemRecords = 3584;

// filter.rs:462-508 (+ 562-571) for every super-k-mer of one bin: counters of
// the bin staged in LDS at native width, min over hashes per s-mer, sliding
// min over k-s+1 s-mers per k-mer, abundances written at their window index.
template <int BITS, int WINC, bool EXACT>
__global__ __launch_bounds__(kQueryThreads) void k_bin_query(BinArgs a, DevGeo g) {
    extern __shared__ __align__(16) uint32_t s_tab[];
    __shared__ uint64_t s_item;
    __shared__ uint64_t s_red[3][kQueryThreads / 64];
    __shared__ uint16_t s_perm[kQueryItemRecords];
    __shared__ uint32_t s_cls[32];
    __shared__ uint4 s_slot[kQueryThreads];  // a thread's next 16 bytes of estimates of its record
    using T = typename Ctr<BITS>::T;
    constexpr uint32_t kV = 16 / sizeof(T);
    const uint32_t tid = threadIdx.x, nthr = blockDim.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t bin_slots = 1u << a.bin_log;
    const uint32_t lshift = a.bin_log - g.bse;
    const uint32_t amask = (1u << g.bse) - 1;
    const uint32_t nwords = bin_slots * (BITS / 8) / 4;
    const uint64_t nwork = a.stats[kStNWork];
    unsigned long long *head = (unsigned long long *)&a.stats[kStWorkHead];
    const T *lt = reinterpret_cast<const T *>(s_tab);
    const bool one_word = smer_tail_one_word(g);
    uint64_t queried = 0, positive = 0, checksum = 0;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(head, 1ull);
        __syncthreads();
        uint64_t item = s_item;
        if (item >= nwork) break;
        const uint4 wk = a.work[item];
        const uint64_t bin = wk.x;
        const uint32_t rb = wk.y, re = wk.z;
        const uint32_t *tw = reinterpret_cast<const uint32_t *>(a.table) + bin * uint64_t(nwords);
        // the records' high words load with the bin's table words (one round trip)
        constexpr int R = kQueryItemRecords / kQueryThreads;
        uint32_t hi[R];
        load_hi<R>(a.sorted, rb, re - rb, hi, tid, nthr);
        stage_to_lds<8>(s_tab, tw, nwords, tid, nthr);
        order_by_hi<R>(hi, re - rb, s_perm, s_cls, tid, nthr);  // (synchronizes)
        walk_records(a.sorted, a.pk, g.k, rb, re - rb, s_perm, tid, nthr, [&](uint64_t rec, const uint32_t (&W)[5]) {
            const uint64_t j0 = rec & 0xFFFFFFFFull;
            const uint32_t nk = uint32_t((rec >> 32) & 31) + 1;
            const uint32_t lbase = uint32_t((rec >> 37) & ((1u << lshift) - 1)) << g.bse;
            uint32_t A0, A1, A2, A3;
            align_words(W, j0, A0, A1, A2, A3);
            WinMin<WINC> wm;
            T *out = reinterpret_cast<T *>(a.out);
            T *slot = reinterpret_cast<T *>(&s_slot[tid]);
            // per-record sums: at most 32 k-mers, so 8/16-bit estimates sum in 32 bits
            uint32_t rec_pos = 0, rec_sum = 0;
            uint64_t rec_sum64 = 0;
            with_nh(g.nh, [&](auto nht) {
                constexpr int NH = decltype(nht)::value;
                for_each_smer(
                    A0, A1, A2, A3, nk + g.k - 1, g.s,
                    [&](uint32_t j, uint64_t smer) {
                        const uint32_t est = min_at_nh<NH>(smer, g.nh, amask, lt + lbase, Ctr<BITS>::kMax);
                        const uint32_t ab = EXACT ? wm.push_exact(est) : wm.push(est, g.win);
                        if (j + 1 >= g.win) {
                            const uint32_t q = j + 1 - g.win;
                            slot[q & (kV - 1)] = T(ab);
                            if ((q & (kV - 1)) == kV - 1 && out)  // a full slot: kV estimates in one store
                                store_run(reinterpret_cast<uint8_t *>(out + j0 + q - (kV - 1)), s_slot[tid], 16);
                            rec_pos += ab > 0;
                            if (BITS < 32) rec_sum += ab;
                            else rec_sum64 += ab;
                        }
                    },
                    one_word);
            });
            // the estimates left in the slot, in two or three stores
            const uint32_t rest = nk & (kV - 1);
            if (rest && out)
                store_run(reinterpret_cast<uint8_t *>(out + j0 + (nk - rest)), s_slot[tid], rest * uint32_t(sizeof(T)));
            queried += nk;
            positive += rec_pos;
            checksum += BITS < 32 ? uint64_t(rec_sum) : rec_sum64;
        });
        __syncthreads();
    }
    uint64_t vals[3] = {queried, positive, checksum};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        uint64_t v = vals[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) s_red[q][wid] = v;
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long *st = (unsigned long long *)a.stats;
        uint64_t t[3] = {0, 0, 0};
        for (int q = 0; q < 3; ++q)
            for (uint32_t i = 0; i < nthr / 64; ++i) t[q] += s_red[q][i];
        if (t[0]) atomicAdd(&st[kStQueried], t[0]);
        if (t[1]) atomicAdd(&st[kStPositive], t[1]);
        if (t[2]) atomicAdd(&st[kStChecksum], t[2]);
    }
}

// One workgroup: total k-mers of the bins that must use the global set.
__global__ __launch_bounds__(1024) void k_spectrum_plan(const uint32_t *bin_count, const uint32_t *bin_kmers,
                                                        uint64_t nbins, uint32_t per_item, uint64_t *stats) {
    __shared__ unsigned long long s_sum;
    if (threadIdx.x == 0) s_sum = 0;
    __syncthreads();
    unsigned long long loc = 0;
    for (uint64_t b = threadIdx.x; b < nbins; b += blockDim.x)
        if (bin_count[b] > per_item || bin_kmers[b] > kSetDirect) loc += bin_kmers[b];
    atomicAdd(&s_sum, loc);
    __syncthreads();
    if (threadIdx.x == 0) stats[kStGlobalKmers] = s_sum;
}

// Histogram of estimates: values below kHistSmall are staged in LDS by the
// caller, values below dense_n go to a dense global array, larger (32-bit)
// values to an open-addressing table of kHistBigCap entries.
struct SpecArgs {
    BinArgs b;
    const uint32_t *bin_kmers;
    unsigned long long *gset;
    uint32_t glog2;
    HistOut h;
    uint32_t stage;              // counters staged in LDS
    // non-null: the first occurrence of each key in this call is appended as
    // (key, estimate) at the slot counted in stats[kStEmitted], instead of
    // being counted into the histogram
    uint64_t *emit_keys;
    uint32_t *emit_vals;
    uint64_t emit_cap;
};

// Appends (key, v) to the emit arrays: the emitting lanes of a wave take
// consecutive slots with one atomic add by the lowest of them.
__device__ __forceinline__ void emit_pair(const SpecArgs &a, uint64_t key, uint32_t v) {
    const uint64_t m = __ballot(1);
    const uint32_t lane = __lane_id();
    const uint32_t leader = uint32_t(__ffsll((unsigned long long)m) - 1);
    const uint32_t rank = uint32_t(__popcll(m & ((uint64_t(1) << lane) - 1)));
    unsigned long long base = 0;
    if (lane == leader)
        base = atomicAdd((unsigned long long *)&a.b.stats[kStEmitted], (unsigned long long)__popcll(m));
    base = __shfl(base, int(leader), 64);
    const uint64_t i = base + rank;
    if (i < a.emit_cap) {
        a.emit_keys[i] = key;
        a.emit_vals[i] = v;
    }
}

template <int BITS, int WINC>
__global__ __launch_bounds__(kSpectrumThreads) void k_bin_spectrum(SpecArgs sa, DevGeo g) {
    extern __shared__ __align__(16) uint8_t s_raw[];
    __shared__ uint64_t s_item;
    __shared__ uint64_t s_red[3][kSpectrumThreads / 64];
    __shared__ uint16_t s_perm[kMaxItemRecords];
    __shared__ uint32_t s_cls[32];
    __shared__ uint32_t s_nwalk;
    using T = typename Ctr<BITS>::T;
    const BinArgs &a = sa.b;
    const uint32_t tid = threadIdx.x, nthr = blockDim.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t bin_slots = 1u << a.bin_log;
    const uint32_t lshift = a.bin_log - g.bse;
    const uint32_t amask = (1u << g.bse) - 1;
    const uint32_t tab_bytes = sa.stage ? bin_slots * (BITS / 8) : 0;
    const uint32_t nwords = tab_bytes / 4;
    uint32_t *s_tab = reinterpret_cast<uint32_t *>(s_raw);
    unsigned long long *s_set = reinterpret_cast<unsigned long long *>(s_raw + tab_bytes);
    uint32_t *s_hist = reinterpret_cast<uint32_t *>(s_raw + tab_bytes + kSetCap * 8);  // kSpecHist counts
    const T *lt = reinterpret_cast<const T *>(s_raw);
    const uint64_t nwork = a.stats[kStNWork];
    unsigned long long *head = (unsigned long long *)&a.stats[kStWorkHead];
    for (uint32_t i = tid; i < kSpecHist; i += nthr) s_hist[i] = 0;
    uint64_t queried = 0, distinct = 0;
    // the distinct-record search of the insert, in the LDS of the staged table
    // and the key set (which are filled after it): its table and keys, then the
    // multiplicities and the leaders
    uint32_t *s_ws = reinterpret_cast<uint32_t *>(s_raw);
    uint32_t *s_mult = s_ws + kRecDedupWords;
    uint16_t *s_lead = reinterpret_cast<uint16_t *>(s_mult + kInsertItemRecords / 2);
    const bool dedup_fits = tab_bytes + kSetCap * 8 >= 4 * (kRecDedupWords + kInsertItemRecords);
    for (;;) {
        if (tid == 0) s_item = atomicAdd(head, 1ull);
        __syncthreads();
        uint64_t item = s_item;
        if (item >= nwork) break;
        const uint4 wk = a.work[item];
        const uint64_t bin = wk.x;
        const uint32_t rb = wk.y, re = wk.z;
        const bool global = (wk.w >> 31) || (wk.w & 0x7FFFFFFFu) > kSetDirect;
        const uint32_t *tw = reinterpret_cast<const uint32_t *>(a.table) + bin * uint64_t(bin_slots * (BITS / 8) / 4);
        const uint32_t n = re - rb;
        // Identical records (same bases, k-mer count and block) have the same
        // keys with the same estimates: only the first of them (the leader) is
        // walked; the others add their k-mers to the queried count.
        const bool dedup = dedup_fits && n > 1 && n <= kInsertItemRecords;
        uint32_t nwalk = n;
        uint4 lkey = make_uint4(0, 0, 0, 0);  // this thread's first leader: bases and high word
        uint32_t lhi = 0;
        if (dedup) {
            constexpr int R = kInsertItemRecords / kSpectrumThreads;
            uint32_t hi[R];
            find_duplicates<R>(a.sorted, a.pk, rb, n, g.k, s_ws, s_lead, s_mult, tid, nthr, hi);  // (synchronizes)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const uint32_t i = r * nthr + tid;
                if (i < n && s_lead[i] != i) queried += (hi[r] & 31) + 1;
            }
            order_by_hi<R>(hi, n, s_perm, s_cls, tid, nthr, nullptr, 0, s_lead, &s_nwalk);  // (synchronizes)
            nwalk = s_nwalk;
            if (tid < nwalk) {
                const uint32_t li = s_perm[tid];
                lkey = dedup_key(s_ws, li), lhi = dedup_hi(s_ws, li);
            }
            __syncthreads();
            stage_to_lds<8>(s_tab, tw, nwords, tid, nthr);
            if (!global)
                for (uint32_t i = tid; i < kSetCap; i += nthr) s_set[i] = kEmptyKey;
            __syncthreads();
        } else {
            constexpr int R = kMaxItemRecords / kSpectrumThreads;
            uint32_t hi[R];
            load_hi<R>(a.sorted, rb, n, hi, tid, nthr);
            stage_to_lds<8>(s_tab, tw, nwords, tid, nthr);
            if (!global)
                for (uint32_t i = tid; i < kSetCap; i += nthr) s_set[i] = kEmptyKey;
            __syncthreads();
            order_by_hi<R>(hi, n, s_perm, s_cls, tid, nthr);  // (synchronizes)
        }
        const T *tbase = sa.stage ? lt : reinterpret_cast<const T *>(a.table) + (bin << a.bin_log);
        walk_records_aligned(a.sorted, a.pk, g.k, rb, nwalk, s_perm, tid, nthr, [&](uint32_t hi, uint32_t A0, uint32_t A1,
                                                                                   uint32_t A2, uint32_t A3, uint32_t) {
            const uint32_t nk = (hi & 31) + 1;
            const uint32_t lbase = ((hi >> 5) & ((1u << lshift) - 1)) << g.bse;
            WinMin<WINC> wm;
            // blocks of at most 2^15 slots use the unrolled chains; larger blocks
            // (addresses above 15 bits) the run-time loop below
            with_nh(g.bse <= 15 ? g.nh : 0u, [&](auto nht) {
                constexpr int NH = decltype(nht)::value;
                // the canonical k-mer key rolls along with the s-mers
                for_each_smer_kmer(
                    A0, A1, A2, A3, nk + g.k - 1, g.s, g.k,
                    [&](uint32_t j, uint64_t smer, uint64_t key) {
                        uint32_t est = Ctr<BITS>::kMax;
                        if (NH > 0 || g.bse <= 15) {
                            est = min_at_nh<NH>(smer, g.nh, amask, tbase + lbase, Ctr<BITS>::kMax);
                        } else {
                            uint64_t h = xs64(smer);
                            for (uint32_t t = 0; t < g.nh; ++t) {
                                uint32_t v = tbase[lbase + (uint32_t(h) & amask)];
                                est = v < est ? v : est;
                                if (t + 1 < g.nh) h = xs64(h);
                            }
                        }
                        uint32_t ab = wm.push(est, g.win);
                        if (j + 1 >= g.win) {
                            uint32_t r = global ? 2u : set_insert_bounded(s_set, kSetLog, key, kSetProbes);
                            if (r == 2) r = set_insert_bounded(sa.gset, sa.glog2, key, kGlobalProbes);
                            if (r == 2)  // the global set is too small: the call runs again with a larger one
                                atomicOr((unsigned long long *)&a.stats[kStSetFull], 1ull);
                            ++queried;
                            if (r == 1) {
                                ++distinct;
                                if (sa.emit_keys) emit_pair(sa, key, ab);
                                else spec_hist_add<kSpecHist>(sa.h, s_hist, ab);
                            }
                        }
                    },
                    smer_tail_one_word(g));
            });
        }, dedup, lkey, lhi);
        __syncthreads();
    }
    for (uint32_t i = tid; i < kSpecHist; i += nthr)
        if (s_hist[i]) atomicAdd(&sa.h.dense[i], (unsigned long long)s_hist[i]);
    uint64_t vals[2] = {queried, distinct};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        uint64_t v = vals[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) s_red[q][wid] = v;
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long *st = (unsigned long long *)a.stats;
        uint64_t t[2] = {0, 0};
        for (int q = 0; q < 2; ++q)
            for (uint32_t i = 0; i < nthr / 64; ++i) t[q] += s_red[q][i];
        if (t[0]) atomicAdd(&st[kStQueried], t[0]);
        if (t[1]) atomicAdd(&st[kStDistinct], t[1]);
    }
}

// Key sets kept across spectrum calls (the per-worker maps of fastx.rs:255-308
// merged first-wins, fastx.rs:392-400): the pairs one call emitted go into a
// global open-addressing set; a key the set did not hold yet is counted into
// the set's histogram. A key's estimate is the same in every call (the index
// is frozen), so first-wins needs no order.
__global__ __launch_bounds__(256) void k_keyset_merge(const uint64_t *keys, const uint32_t *vals, uint64_t n,
                                                      unsigned long long *set, uint32_t log2cap, HistOut h,
                                                      unsigned long long *distinct) {
    __shared__ uint32_t s_hist[kHistSmall];
    __shared__ unsigned long long s_new;
    for (uint32_t i = threadIdx.x; i < kHistSmall; i += blockDim.x) s_hist[i] = 0;
    if (threadIdx.x == 0) s_new = 0;
    __syncthreads();
    uint32_t fresh = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        if (set_insert(set, log2cap, keys[i])) {
            ++fresh;
            spec_hist_add(h, s_hist, vals[i]);
        }
    if (fresh) atomicAdd(&s_new, (unsigned long long)fresh);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHistSmall; i += blockDim.x)
        if (s_hist[i]) atomicAdd(&h.dense[i], (unsigned long long)s_hist[i]);
    if (threadIdx.x == 0 && s_new) atomicAdd(distinct, s_new);
}

// Moves every key of a full set into a larger one.
__global__ __launch_bounds__(256) void k_keyset_rehash(const unsigned long long *old, uint64_t old_cap,
                                                       unsigned long long *set, uint32_t log2cap) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < old_cap;
         i += uint64_t(gridDim.x) * blockDim.x) {
        const unsigned long long k = old[i];
        if (k != kEmptyKey) (void)set_insert(set, log2cap, k);
    }
}

// ---- reservoir-sampled spectrum (bin/sampled_spectrum.rs:150-308) ------------
// The reference keeps the first `cap` valid windows of the input in a
// reservoir; the w-th valid window (w > cap, windows counted from 1 in input
// order) then draws r = next_u64() % w from one SplitMix64 stream and replaces
// slot r when r < cap. The stream's t-th draw is the SplitMix64 output of
// seed + t * 0x9E3779B97F4A7C15, i.e. xs64(seed + (t - 1) * 0x9E3779B97F4A7C15),
// and window w makes draw t = w - cap, so every window's target slot is a
// function of w alone: slot s ends up holding the largest w that targeted it.
// Windows are numbered from a per-word rank of the query's validity bitmap;
// one pass takes, per slot, the largest w (atomic max), a second pass stores
// that window's estimate.
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr int kRankThreads = 1024;

// Reservoir slot written by the w-th valid window, or cap for none.
__device__ __forceinline__ uint64_t reservoir_slot(uint64_t w, uint64_t cap, uint64_t seed) {
    if (w <= cap) return w - 1;
    const uint64_t r = xs64(seed + (w - cap - 1) * kGolden) % w;
    return r < cap ? r : cap;
}

// Valid windows per block of kRankThreads bitmap words.
__global__ __launch_bounds__(kRankThreads) void k_rank_sums(const uint32_t *bits, uint64_t nwords,
                                                            uint32_t *block_sums) {
    __shared__ uint32_t s_w[kRankThreads / 64];
    const uint64_t i = uint64_t(blockIdx.x) * kRankThreads + threadIdx.x;
    uint32_t c = i < nwords ? uint32_t(__popc(bits[i])) : 0u;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int q = 0; q < kRankThreads / 64; ++q) t += s_w[q];
        block_sums[blockIdx.x] = t;
    }
}

// Exclusive scan of the block sums in place (one workgroup, chunks of 1024).
__global__ __launch_bounds__(kRankThreads) void k_rank_scan_blocks(uint32_t *sums, uint64_t n) {
    __shared__ uint32_t s_w[kRankThreads / 64];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (uint64_t base = 0; base < n; base += kRankThreads) {
        const uint64_t i = base + threadIdx.x;
        const uint32_t v = i < n ? sums[i] : 0u;
        uint32_t incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(incl, off, 64);
            if (lane >= uint32_t(off)) incl += y;
        }
        if (lane == 63) s_w[wid] = incl;
        __syncthreads();
        uint32_t before = s_carry, total = 0;
        for (uint32_t q = 0; q < kRankThreads / 64; ++q) {
            if (q < wid) before += s_w[q];
            total += s_w[q];
        }
        if (i < n) sums[i] = before + incl - v;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += total;
        __syncthreads();
    }
}

// rank[i] = valid windows in words before word i.
__global__ __launch_bounds__(kRankThreads) void k_rank_words(const uint32_t *bits, uint64_t nwords,
                                                             const uint32_t *block_base, uint32_t *rank) {
    __shared__ uint32_t s_w[kRankThreads / 64];
    const uint64_t i = uint64_t(blockIdx.x) * kRankThreads + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t v = i < nwords ? uint32_t(__popc(bits[i])) : 0u;
    uint32_t incl = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= uint32_t(off)) incl += y;
    }
    if (lane == 63) s_w[wid] = incl;
    __syncthreads();
    uint32_t before = block_base[blockIdx.x];
    for (uint32_t q = 0; q < wid; ++q) before += s_w[q];
    if (i < nwords) rank[i] = before + incl - v;
}

// Per slot, the largest window number that targets it (one thread per word).
__global__ __launch_bounds__(256) void k_reservoir_hits(const uint32_t *bits, uint64_t nwords, const uint32_t *rank,
                                                        uint64_t w0, uint64_t cap, uint64_t seed,
                                                        unsigned long long *slot_win) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nwords;
         i += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t w = w0 + rank[i];
        for (uint32_t b = bits[i]; b; b &= b - 1) {
            ++w;
            const uint64_t slot = reservoir_slot(w, cap, seed);
            if (slot < cap) atomicMax(&slot_win[slot], (unsigned long long)w);
        }
    }
}

// The estimate of each slot's final window (when that window is in this batch).
template <int BITS>
__global__ __launch_bounds__(256) void k_reservoir_take(const uint32_t *bits, uint64_t nwords, const uint32_t *rank,
                                                        uint64_t w0, uint64_t cap, uint64_t seed,
                                                        const unsigned long long *slot_win, const void *values,
                                                        uint32_t *slot_val) {
    const typename Ctr<BITS>::T *v = reinterpret_cast<const typename Ctr<BITS>::T *>(values);
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nwords;
         i += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t w = w0 + rank[i];
        for (uint32_t b = bits[i]; b; b &= b - 1) {
            ++w;
            const uint64_t slot = reservoir_slot(w, cap, seed);
            if (slot < cap && slot_win[slot] == w) slot_val[slot] = uint32_t(v[i * 32 + uint32_t(__ffs(b) - 1)]);
        }
    }
}

// filter.rs:50-55, 267-275: histogram of every counter value. Values below
// kSmall are counted in LDS; larger 8/16-bit values go straight to a dense
// global array; larger 32-bit values go to an open-addressing table.

template <int BITS>
__global__ __launch_bounds__(256) void k_counter_hist(const void *table, uint64_t slots,
                                                       unsigned long long *dense,
                                                       uint32_t *big_keys,
                                                       unsigned long long *big_counts,
                                                       uint64_t *stats) {
    __shared__ uint32_t s_h[kHistSmall];
    for (uint32_t i = threadIdx.x; i < kHistSmall; i += blockDim.x) s_h[i] = 0;
    __syncthreads();
    const typename Ctr<BITS>::T *t = reinterpret_cast<const typename Ctr<BITS>::T *>(table);
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < slots;
         i += uint64_t(gridDim.x) * blockDim.x) {
        uint32_t v = t[i];
        if (v < kHistSmall) {
            atomicAdd(&s_h[v], 1u);
        } else if (BITS < 32) {
            atomicAdd(&dense[v], 1ull);
        } else {
            uint32_t hsh = uint32_t(xs64(v)) & (kHistBigCap - 1);
            for (uint32_t probe = 0; probe < kHistBigCap; ++probe) {
                uint32_t key = big_keys[hsh];
                if (key == 0) {
                    uint32_t prev = atomicCAS(&big_keys[hsh], 0u, v);
                    if (prev == 0 || prev == v) {
                        atomicAdd(&big_counts[hsh], 1ull);
                        break;
                    }
                    key = prev;
                }
                if (key == v) {
                    atomicAdd(&big_counts[hsh], 1ull);
                    break;
                }
                hsh = (hsh + 1) & (kHistBigCap - 1);
                if (probe + 1 == kHistBigCap) atomicAdd((unsigned long long *)&stats[kStOverflow], 1ull);
            }
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHistSmall; i += blockDim.x)
        if (s_h[i]) atomicAdd(&dense[i], (unsigned long long)s_h[i]);
}

// ----------------------------------------------------------------------------
// Host side
// ----------------------------------------------------------------------------
#define SCB_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return Status{8, std::string("device error: ") + #call + ": " + hipGetErrorString(e_)}; \
    } while (0)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    bool owned = true;  // false: a view of memory this buffer does not own; it never grows
    hipError_t reserve(size_t want) {
        if (want <= n) return hipSuccess;
        if (!owned) return hipErrorOutOfMemory;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), std::max<size_t>(want, 1) * sizeof(T));
        if (e == hipSuccess) n = want;
        return e;
    }
    ~DevBuf() {
        if (p && owned) (void)hipFree(p);
    }
    void view(T *at, size_t count) { p = at, n = count, owned = false; }
};

// Page-locked host buffer that only grows (host side of batch copies).
template <class T>
struct HostBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t reserve(size_t want) {
        if (want <= n) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&p), std::max<size_t>(want, 1) * sizeof(T),
                                     hipHostMallocDefault);
        if (e == hipSuccess) n = want;
        return e;
    }
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
};

}  // namespace

namespace {
void fill(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st);

uint32_t spectrum_dense_n(const Geometry &g) { return g.bits == 8 ? 256 : (g.bits == 16 ? 65536 : kHistSmall); }

// Device histogram tables of one spectrum (per call, or kept by a KeySet).
struct HistTables {
    DevBuf<unsigned long long> dense, big_counts;
    DevBuf<uint32_t> big_keys;
    uint32_t dense_n = 0;
    hipError_t reserve(const Geometry &g) {
        dense_n = spectrum_dense_n(g);
        hipError_t e;
        if ((e = dense.reserve(dense_n)) != hipSuccess) return e;
        if ((e = big_keys.reserve(g.bits == 32 ? kHistBigCap : 1)) != hipSuccess) return e;
        return big_counts.reserve(g.bits == 32 ? kHistBigCap : 1);
    }
    void clear(const Geometry &g, hipStream_t st) {
        fill(reinterpret_cast<uint32_t *>(dense.p), uint64_t(dense_n) * 2, 0, st);
        if (g.bits == 32) {
            fill(big_keys.p, kHistBigCap, 0, st);
            fill(reinterpret_cast<uint32_t *>(big_counts.p), 2 * kHistBigCap, 0, st);
        }
    }
    HistOut out(uint64_t *stats) const { return HistOut{dense.p, dense_n, big_keys.p, big_counts.p, stats}; }
    // rows (value, count) sorted by value; synchronizes the stream
    Status rows(const Geometry &g, hipStream_t st, std::vector<std::pair<uint64_t, uint64_t>> &out) const {
        std::vector<unsigned long long> d(dense_n);
        std::vector<uint32_t> keys;
        std::vector<unsigned long long> cnts;
        SCB_HIP(hipMemcpyAsync(d.data(), dense.p, uint64_t(dense_n) * 8, hipMemcpyDeviceToHost, st));
        if (g.bits == 32) {
            keys.resize(kHistBigCap);
            cnts.resize(kHistBigCap);
            SCB_HIP(hipMemcpyAsync(keys.data(), big_keys.p, kHistBigCap * 4, hipMemcpyDeviceToHost, st));
            SCB_HIP(hipMemcpyAsync(cnts.data(), big_counts.p, kHistBigCap * 8, hipMemcpyDeviceToHost, st));
        }
        SCB_HIP(hipStreamSynchronize(st));
        out.clear();
        for (uint64_t v = 0; v < dense_n; ++v)
            if (d[v]) out.emplace_back(v, d[v]);
        for (uint32_t i = 0; i < keys.size(); ++i)
            if (keys[i]) out.emplace_back(keys[i], cnts[i]);
        std::sort(out.begin(), out.end());
        return Status::ok();
    }
};
}  // namespace

class Workspace {
   public:
    hipStream_t stream = nullptr;
    DevBuf<uint32_t> pk, valid;
    DevBuf<uint64_t> rec, sorted, stats, plan_blk;
    DevBuf<uint4> work;
    DevBuf<uint32_t> bin_count, bin_off, cursor, ccursor, coarse_hist, seg_count, tile_r0;
    uint64_t *final_sorted = nullptr;  // records grouped by bin after the partition passes
    DevBuf<uint8_t> h_bases_dev;   // staging for host batches
    DevBuf<uint64_t> h_offsets_dev;
    DevBuf<uint8_t> out_dev;
    DevBuf<uint32_t> out_bits;
    HostBuf<uint8_t> h_vals;       // query results on their way back
    HostBuf<uint32_t> h_bits;
    HostBuf<uint64_t> h_offs;
    DevBuf<unsigned long long> hist_dense, hist_big_counts;
    DevBuf<uint32_t> hist_big_keys;
    DevBuf<uint32_t> spec_bin_kmers;
    DevBuf<unsigned long long> spec_gset;
    DevBuf<uint64_t> emit_keys;    // (key, estimate) pairs a spectrum pass emits
    DevBuf<uint32_t> emit_vals;
    HistTables hist;               // spectrum histogram
    uint64_t *h_stats = nullptr;   // pinned
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // A captured pipeline (insert or query) and the key of the call it replays.
    struct GraphSlot {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        uint64_t key = 0, eager_key = 0;
        void drop() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            exec = nullptr, graph = nullptr, key = 0;
        }
    } graph_insert, graph_query;
    bool capturing = false;    // launches are being captured (no event records)
    bool stages_timed = true;  // the last call recorded all four stage events
    // A view workspace runs on the null stream, has no events, and its device
    // buffers point into memory it does not own (none of them can grow); its
    // h_stats belongs to the caller.
    bool is_view = false;
    void make_view() {
        is_view = true;
        for (auto *b : {&pk, &valid, &bin_count, &bin_off, &cursor, &ccursor, &coarse_hist, &seg_count, &tile_r0,
                        &out_bits, &hist_big_keys, &spec_bin_kmers})
            b->owned = false;
        for (auto *b : {&rec, &sorted, &stats, &plan_blk, &h_offsets_dev}) b->owned = false;
        h_bases_dev.owned = out_dev.owned = work.owned = false;
        hist_dense.owned = hist_big_counts.owned = spec_gset.owned = false;
        emit_keys.owned = emit_vals.owned = false;
        hist.dense.owned = hist.big_counts.owned = hist.big_keys.owned = false;
    }
    // Hash of the device buffers a pipeline uses: their addresses and sizes.
    uint64_t buffer_signature() const {
        uint64_t h = 0x243F6A8885A308D3ull;
        auto mix = [&](const void *p, size_t n) {
            h = (h ^ reinterpret_cast<uintptr_t>(p)) * 0x100000001B3ull;
            h = (h ^ n) * 0x100000001B3ull;
        };
        mix(pk.p, pk.n), mix(rec.p, rec.n), mix(sorted.p, sorted.n), mix(work.p, work.n);
        mix(stats.p, stats.n), mix(plan_blk.p, plan_blk.n), mix(bin_count.p, bin_count.n);
        mix(bin_off.p, bin_off.n), mix(cursor.p, cursor.n), mix(ccursor.p, ccursor.n);
        mix(coarse_hist.p, coarse_hist.n), mix(seg_count.p, seg_count.n), mix(tile_r0.p, tile_r0.n);
        mix(spec_bin_kmers.p, spec_bin_kmers.n);
        return h;
    }
    ~Workspace() {
        graph_insert.drop();
        graph_query.drop();
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (h_stats && !is_view) (void)hipHostFree(h_stats);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

Status make_workspace(std::unique_ptr<Workspace> &ws) {
    ws.reset(new Workspace);
    SCB_HIP(hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking));
    SCB_HIP(hipHostMalloc(reinterpret_cast<void **>(&ws->h_stats), kStatsLen * sizeof(uint64_t),
                          hipHostMallocDefault));
    SCB_HIP(ws->stats.reserve(kStatsLen));
    SCB_HIP(ws->seg_count.reserve(kSeg + 1));
    for (auto &e : ws->ev) SCB_HIP(hipEventCreate(&e));
    return Status::ok();
}

thread_local StageTimes g_stage_times;

void mark(Workspace &ws, int i) {
    if (ws.ev[i] && !ws.capturing) (void)hipEventRecord(ws.ev[i], ws.stream);
}

// Elapsed ms between recorded events (after the stream has been synchronized).
void collect_times(Workspace &ws, bool batched) {
    StageTimes t{};
    float ms = 0;
    if (!ws.stages_timed) {  // a replayed graph: only the total is known
        if (hipEventElapsedTime(&ms, ws.ev[0], ws.ev[3]) == hipSuccess) t.total_ms = ms;
        t.batched = batched;
        g_stage_times = t;
        return;
    }
    if (hipEventElapsedTime(&ms, ws.ev[0], ws.ev[1]) == hipSuccess) t.scan_ms = ms;
    if (batched) {
        if (hipEventElapsedTime(&ms, ws.ev[1], ws.ev[2]) == hipSuccess) t.plan_scatter_ms = ms;
        if (hipEventElapsedTime(&ms, ws.ev[2], ws.ev[3]) == hipSuccess) t.bin_ms = ms;
        if (hipEventElapsedTime(&ms, ws.ev[0], ws.ev[3]) == hipSuccess) t.total_ms = ms;
    } else {
        t.total_ms = t.scan_ms;
    }
    t.batched = batched;
    g_stage_times = t;
}

void fill(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st) {
    if (!n) return;
    uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(unsigned(blocks)), dim3(256), 0, st, p, n, v);
}

// Fills of several buffers in one launch: a pipeline starts by clearing five
// or six small buffers, and one launch each costs more than the stores.
struct FillJobs {
    static constexpr int kMax = 8;
    uint32_t *p[kMax];
    uint64_t n[kMax];
    uint32_t v[kMax];
};

__global__ void k_fill_jobs(FillJobs j) {
    const int q = blockIdx.y;
    uint32_t *p = j.p[q];
    const uint64_t n = j.n[q];
    const uint32_t v = j.v[q];
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

class FillBatch {
   public:
    explicit FillBatch(hipStream_t st) : st_(st) {}
    void add(uint32_t *p, uint64_t n, uint32_t v) {
        if (!n) return;
        if (count_ == FillJobs::kMax) flush();
        jobs_.p[count_] = p, jobs_.n[count_] = n, jobs_.v[count_] = v;
        ++count_;
        maxn_ = std::max(maxn_, n);
    }
    void flush() {
        if (!count_) return;
        const uint64_t blocks = std::min<uint64_t>((maxn_ + 255) / 256, 4096);
        hipLaunchKernelGGL(k_fill_jobs, dim3(unsigned(blocks), unsigned(count_)), dim3(256), 0, st_, jobs_);
        count_ = 0, maxn_ = 0;
    }

   private:
    hipStream_t st_;
    FillJobs jobs_{};
    int count_ = 0;
    uint64_t maxn_ = 0;
};

int cu_count(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
    return prop.multiProcessorCount;
}

template <class K>
int persistent_grid(K kernel, int threads, size_t lds, int dev) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return per_cu * cu_count(dev);
}

}  // namespace

// ----------------------------------------------------------------------------
// Engine
// ----------------------------------------------------------------------------
int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// Counter tables of freed engines are kept for reuse by the next engine of
// the same size on the same device (a builder per batch of work would
// otherwise pay hipMalloc/hipFree of the whole table every time). At most
// kTableCacheMax tables are held; an allocation that fails drops the cache and
// retries.
namespace {
constexpr size_t kTableCacheMax = 2;
struct TableCache {
    std::mutex mu;
    struct Entry {
        int device;
        uint64_t bytes;
        void *p;
    };
    std::vector<Entry> free;
};
TableCache &table_cache() {
    static TableCache *c = new TableCache;  // never destroyed, like the workspace pool
    return *c;
}
hipError_t table_alloc(int device, uint64_t bytes, void **out) {
    TableCache &c = table_cache();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        for (size_t i = c.free.size(); i-- > 0;) {
            if (c.free[i].device == device && c.free[i].bytes == bytes) {
                *out = c.free[i].p;
                c.free.erase(c.free.begin() + i);
                return hipSuccess;
            }
        }
    }
    hipError_t e = hipMalloc(out, bytes);
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        (void)hipGetLastError();
        std::vector<TableCache::Entry> drop;
        {
            std::lock_guard<std::mutex> lk(c.mu);
            drop.swap(c.free);
        }
        for (auto &d : drop) (void)hipFree(d.p);
        e = hipMalloc(out, bytes);
    }
    return e;
}
void table_release(int device, uint64_t bytes, void *p) {
    TableCache &c = table_cache();
    void *evict = nullptr;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        c.free.push_back({device, bytes, p});
        if (c.free.size() > kTableCacheMax) {
            evict = c.free.front().p;
            c.free.erase(c.free.begin());
        }
    }
    if (evict) (void)hipFree(evict);
}
}  // namespace

static_assert(kMaxItemRecords <= 65536, "record order indices are u16");
static_assert(kQueryItemRecords % kQueryThreads == 0, "query items are walked in whole rounds");

// ---- tuning -------------------------------------------------------------------
Tuning Engine::default_tuning(const Geometry &g) {
    Tuning t;
    // keep query bins within LDS at every width (u32 counters: 2^15 slots = 128 KiB)
    t.insert_bin_log = std::min<uint32_t>(15, g.cse);
    t.query_bin_log = std::min<uint32_t>(15, g.cse);
    t.spectrum_bin_log = std::min<uint32_t>(14, g.cse);
    t.insert_item_records = kInsertItemRecords;
    t.query_item_records = kQueryItemRecords;
    t.spectrum_item_records = kMaxItemRecords;
    return t;
}

Status Engine::set_tuning(const Tuning &t) {
    const Geometry &g = geo_;
    auto bad = [](const std::string &m) { return Status{1, "invalid config: tuning: " + m}; };
    const uint32_t lo = std::max<uint32_t>(g.bse, 1);
    if (t.insert_bin_log < lo || t.insert_bin_log > std::min<uint32_t>(15, g.cse))
        return bad("insert_bin_log must lie in [max(block_slots_exponent, 1), min(15, counter_slots_exponent)]");
    if (t.query_bin_log < g.bse || t.query_bin_log > g.cse || (uint64_t(1) << t.query_bin_log) * g.bytes > (128u << 10))
        return bad("query_bin_log must lie in [block_slots_exponent, counter_slots_exponent] with a bin of at most "
                   "128 KiB");
    if (t.spectrum_bin_log < std::min<uint32_t>(g.bse, 14) || t.spectrum_bin_log > std::min<uint32_t>(14, g.cse))
        return bad("spectrum_bin_log must lie in [min(block_slots_exponent, 14), min(14, counter_slots_exponent)]");
    if (t.insert_item_records < 1 || t.insert_item_records > kInsertItemRecords)
        return bad("insert_item_records must lie in [1, " + std::to_string(kInsertItemRecords) + "]");
    if (t.query_item_records < 1 || t.query_item_records > kQueryItemRecords)
        return bad("query_item_records must lie in [1, " + std::to_string(kQueryItemRecords) + "]");
    if (t.spectrum_item_records < 1 || t.spectrum_item_records > kMaxItemRecords)
        return bad("spectrum_item_records must lie in [1, " + std::to_string(kMaxItemRecords) + "]");
    if (t.use_graph > 1) return bad("use_graph must be 0 or 1");
    tune_ = t;
    return Status::ok();
}

namespace {
struct TuningCache {
    std::mutex mu;
    bool loaded = false;
    std::vector<std::pair<TuningKey, Tuning>> entries;
};
TuningCache &tuning_cache() {
    static TuningCache *c = new TuningCache;
    return *c;
}

// Integer value of "name": in one JSON object's text (no nesting of the same
// name); false when absent.
bool json_int(const std::string &obj, const char *name, uint64_t &out) {
    const std::string key = std::string("\"") + name + "\"";
    size_t p = obj.find(key);
    if (p == std::string::npos) return false;
    p = obj.find(':', p + key.size());
    if (p == std::string::npos) return false;
    ++p;
    while (p < obj.size() && (obj[p] == ' ' || obj[p] == '\t' || obj[p] == '\n' || obj[p] == '\r')) ++p;
    if (p >= obj.size() || obj[p] < '0' || obj[p] > '9') return false;
    out = std::strtoull(obj.c_str() + p, nullptr, 10);
    return true;
}

// Entries are the objects at depth 2 of {"entries": [ {...}, ... ]}; an object's
// nested objects (timings) are part of its text and hold no key of its own.
Status parse_tuning_json(const std::string &text, std::vector<std::pair<TuningKey, Tuning>> &out) {
    const size_t arr = text.find("\"entries\"");
    if (arr == std::string::npos) return Status{1, "invalid config: tuning cache has no \"entries\""};
    size_t p = text.find('[', arr);
    if (p == std::string::npos) return Status{1, "invalid config: tuning cache \"entries\" is not a list"};
    int depth = 0;
    size_t start = 0;
    bool in_str = false;
    for (++p; p < text.size(); ++p) {
        const char c = text[p];
        if (in_str) {
            if (c == '\\') ++p;
            else if (c == '"') in_str = false;
            continue;
        }
        if (c == '"') in_str = true;
        else if (c == '{') {
            if (depth++ == 0) start = p;
        } else if (c == '}') {
            if (--depth == 0) {
                const std::string obj = text.substr(start, p - start + 1);
                // the entry's own fields precede its nested timing object
                const size_t nested = obj.find('{', 1);
                const std::string head = nested == std::string::npos ? obj : obj.substr(0, nested);
                uint64_t v[14];
                const char *names[14] = {"k", "m", "s", "n_hashes", "counter_bits", "counter_slots_exponent",
                                         "block_slots_exponent", "insert_bin_log", "query_bin_log",
                                         "spectrum_bin_log", "insert_item_records", "query_item_records",
                                         "spectrum_item_records", "use_graph"};
                for (int i = 0; i < 14; ++i)
                    if (!json_int(head, names[i], v[i]))
                        return Status{1, std::string("invalid config: tuning cache entry without \"") + names[i] +
                                             "\""};
                TuningKey key{uint32_t(v[0]), uint32_t(v[1]), uint32_t(v[2]), uint32_t(v[3]), uint32_t(v[4]),
                              uint32_t(v[5]), uint32_t(v[6])};
                Tuning t;
                t.insert_bin_log = uint32_t(v[7]), t.query_bin_log = uint32_t(v[8]);
                t.spectrum_bin_log = uint32_t(v[9]), t.insert_item_records = uint32_t(v[10]);
                t.query_item_records = uint32_t(v[11]), t.spectrum_item_records = uint32_t(v[12]);
                t.use_graph = uint32_t(v[13]);
                out.emplace_back(key, t);
            }
        } else if (c == ']' && depth == 0) {
            break;
        }
    }
    return Status::ok();
}

Status load_tuning_file(const char *path, std::vector<std::pair<TuningKey, Tuning>> &out) {
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return Status{7, std::string("I/O error for '") + path + "': cannot open the tuning cache"};
    std::string text;
    char buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, got);
    std::fclose(fp);
    return parse_tuning_json(text, out);
}

// The cache a process starts with: SCB_TUNING_CACHE, else tuning.json beside
// libscb.so (absent files leave the defaults).
void load_default_tuning_cache(TuningCache &c) {
    c.loaded = true;
    std::string path;
    if (const char *e = std::getenv("SCB_TUNING_CACHE")) {
        path = e;
    } else {
        Dl_info info;
        if (dladdr(reinterpret_cast<void *>(&load_tuning_cache), &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.rfind('/');
            path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/tuning.json";
        }
    }
    if (path.empty()) return;
    std::vector<std::pair<TuningKey, Tuning>> entries;
    if (load_tuning_file(path.c_str(), entries).good()) c.entries.swap(entries);
}

// The cache entry of the geometry, if any.
bool cached_tuning(const Geometry &g, Tuning &t);
}  // namespace

Status load_tuning_cache(const char *path) {
    std::vector<std::pair<TuningKey, Tuning>> entries;
    Status s = load_tuning_file(path, entries);
    if (!s.good()) return s;
    TuningCache &c = tuning_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    c.entries.swap(entries);
    c.loaded = true;
    return Status::ok();
}

namespace {
bool cached_tuning(const Geometry &g, Tuning &t) {
    TuningCache &c = tuning_cache();
    std::lock_guard<std::mutex> lk(c.mu);
    if (!c.loaded) load_default_tuning_cache(c);
    const TuningKey key{g.k, g.m, g.s, g.nh, g.bits, g.cse, g.bse};
    bool have = false;
    for (auto &e : c.entries)
        if (e.first == key) t = e.second, have = true;
    return have;
}
}  // namespace

Status Engine::create(const Geometry &g, std::unique_ptr<Engine> &out) {
    if (device_count() <= 0) return Status{8, "device error: no HIP device is visible"};
    std::unique_ptr<Engine> e(new Engine);
    e->geo_ = g;
    SCB_HIP(hipGetDevice(&e->device_));
    SCB_HIP(table_alloc(e->device_, std::max<uint64_t>(g.table_bytes(), 4), &e->d_table_));
    e->zero_pending_ = true;
    // keep query bins within LDS at every width (u32 counters: 2^15 slots = 128 KiB)
    e->tune_ = default_tuning(g);
    Tuning cached;
    if (cached_tuning(g, cached)) (void)e->set_tuning(cached);  // an entry the kernels cannot run is ignored
    out = std::move(e);
    return Status::ok();
}

Engine::~Engine() {
    // every call on the engine has synchronized its stream before returning
    if (d_table_) table_release(device_, std::max<uint64_t>(geo_.table_bytes(), 4), d_table_);
}

Status Engine::ensure_zeroed() const {
    std::lock_guard<std::mutex> lk(zero_mu_);
    if (!zero_pending_) return Status::ok();
    SCB_HIP(hipSetDevice(device_));
    SCB_HIP(hipMemset(d_table_, 0, std::max<uint64_t>(geo_.table_bytes(), 4)));
    SCB_HIP(hipDeviceSynchronize());
    zero_pending_ = false;
    return Status::ok();
}

// Scratch workspaces are pooled per device for the whole process, so a new
// builder or index reuses the batch buffers of the previous one instead of
// allocating them again. The pool is never destroyed (the HIP runtime may
// already be gone when static destructors run).
namespace {
struct WorkspacePool {
    std::mutex mu;
    std::vector<std::pair<int, std::unique_ptr<Workspace>>> free;
};
WorkspacePool &workspace_pool() {
    static WorkspacePool *pool = new WorkspacePool;
    return *pool;
}
}  // namespace

std::unique_ptr<Workspace> Engine::acquire() const {
    WorkspacePool &pool = workspace_pool();
    {
        std::lock_guard<std::mutex> lk(pool.mu);
        for (size_t i = pool.free.size(); i-- > 0;) {
            if (pool.free[i].first == device_) {
                auto ws = std::move(pool.free[i].second);
                pool.free.erase(pool.free.begin() + i);
                return ws;
            }
        }
    }
    std::unique_ptr<Workspace> ws;
    if (!make_workspace(ws).good()) return nullptr;
    return ws;
}

void Engine::release(std::unique_ptr<Workspace> ws) const {
    if (!ws) return;
    WorkspacePool &pool = workspace_pool();
    std::lock_guard<std::mutex> lk(pool.mu);
    pool.free.emplace_back(device_, std::move(ws));
}

Status Engine::download(void *host, uint64_t bytes) const {
    if (bytes != geo_.table_bytes()) return Status{6, "internal state error: table size mismatch"};
    Status z = ensure_zeroed();
    if (!z.good()) return z;
    SCB_HIP(hipMemcpy(host, d_table_, bytes, hipMemcpyDeviceToHost));
    return Status::ok();
}

Status Engine::upload(const void *host, uint64_t bytes) {
    if (bytes != geo_.table_bytes()) return Status{6, "internal state error: table size mismatch"};
    SCB_HIP(hipMemcpy(d_table_, host, bytes, hipMemcpyHostToDevice));
    std::lock_guard<std::mutex> lk(zero_mu_);
    zero_pending_ = false;
    return Status::ok();
}

namespace {

struct Launch {
    const Geometry &geo;
    DevGeo dg;
    Workspace &ws;
    void *table;
    int device;
};

template <int MODE>
Status launch_scan(Launch &L, const DeviceBatch &b, uint64_t seg_cap, uint32_t *valid_bits, void *out,
                   unsigned long long *gset = nullptr, uint32_t glog2 = 0, const HistOut *hist = nullptr) {
    ScanArgs a{};
    a.gset = gset;
    a.glog2 = glog2;
    if (hist) a.hist = *hist;
    a.bases = b.d_bases;
    a.offsets = b.d_offsets;
    a.n_reads = b.n_reads;
    a.n_bases = b.n_bases;
    a.pk = L.ws.pk.p;
    a.valid_bits = valid_bits;
    a.rec = L.ws.rec.p;
    a.seg_count = L.ws.seg_count.p;
    a.seg_cap = seg_cap;
    a.stats = L.ws.stats.p;
    a.table = L.table;
    a.out = out;
    a.aligned16 = (reinterpret_cast<uintptr_t>(b.d_bases) & 15) == 0;
    unsigned tiles = unsigned((b.n_bases + kTile - 1) / kTile);
    if (!tiles) return Status::ok();
    bool wide = L.geo.win > 8;
    hipStream_t st = L.ws.stream;
    SCB_HIP(L.ws.tile_r0.reserve(tiles));
    hipLaunchKernelGGL(k_tile_r0, dim3(unsigned((std::max<uint64_t>(tiles, b.n_reads) + 255) / 256)), dim3(256), 0,
                       st, b.d_offsets, b.n_reads, b.n_bases, tiles, L.ws.tile_r0.p, L.ws.stats.p);
    a.tile_r0 = L.ws.tile_r0.p;
    a.ntiles = tiles;
    // persistent tiles: a few resident workgroups per CU, each prefetching its next tile
    dim3 grid(tiles), blk(kTileThreads);
#define SCB_SCAN(BITS)                                                                         \
    do {                                                                                       \
        if (wide) hipLaunchKernelGGL((k_scan<MODE, BITS, 32>), grid, blk, 0, st, a, L.dg);     \
        else hipLaunchKernelGGL((k_scan<MODE, BITS, 8>), grid, blk, 0, st, a, L.dg);           \
    } while (0)
    if constexpr (MODE == kModeRecords) {
        // records do not depend on the counter width or the fimpera window; the
        // minimizer window w has register-resident variants for common values
        switch (L.geo.w) {
            case 11: hipLaunchKernelGGL((k_scan<kModeRecords, 16, 8, 11>), grid, blk, 0, st, a, L.dg); break;
            case 21: hipLaunchKernelGGL((k_scan<kModeRecords, 16, 8, 21>), grid, blk, 0, st, a, L.dg); break;
            default: hipLaunchKernelGGL((k_scan<kModeRecords, 16, 8>), grid, blk, 0, st, a, L.dg); break;
        }
    } else {
        switch (L.geo.bits) {
            case 8: SCB_SCAN(8); break;
            case 16: SCB_SCAN(16); break;
            default: SCB_SCAN(32); break;
        }
    }
#undef SCB_SCAN
    SCB_HIP(hipGetLastError());
    return Status::ok();
}

// Bits of the bin index sorted by the fine pass when 2^lb bins need two passes
// (lb > 11): the bits are split evenly (at least 7 fine), so the coarse pass
// ranks each 8192-record tile over at most 2^((lb + 1) / 2) buckets and its
// scattered runs stay long (with 2^11 coarse buckets a tile held ~4 records
// per bucket and wrote each 128-byte line several times).
uint32_t sort_fine_bits(uint32_t lb) { return std::max<uint32_t>(7, (lb + 1) / 2); }

void launch_plan(Workspace &ws, uint64_t nbins, uint32_t per_item, hipStream_t st, const uint32_t *bin_kmers) {
    const unsigned nblk = unsigned((nbins + kPlanThreads - 1) / kPlanThreads);
    const uint32_t per_shift = (per_item & (per_item - 1)) == 0 ? uint32_t(__builtin_ctz(per_item)) : 32u;
    hipLaunchKernelGGL(k_plan_sums, dim3(nblk), dim3(kPlanThreads), 0, st, ws.bin_count.p, nbins, per_item, per_shift,
                       ws.plan_blk.p);
    hipLaunchKernelGGL(k_plan_write, dim3(nblk), dim3(kPlanThreads), 0, st, ws.bin_count.p, nbins, per_item,
                       per_shift, ws.plan_blk.p, ws.bin_off.p, ws.cursor.p, ws.work.p, bin_kmers, ws.stats.p);
}

// Batched front half shared by insert, query and spectrum: scan to records,
// then group the records by bin (one or two tile-aggregated counting-sort
// passes) and build the work list.
Status batched_front(Launch &L, const DeviceBatch &b, uint32_t bin_log, uint32_t per_item,
                     uint32_t *valid_bits, uint64_t &nbins_out, uint32_t *bin_kmers = nullptr) {
    const Geometry &g = L.geo;
    Workspace &ws = L.ws;
    const uint64_t nbins = uint64_t(1) << (g.cse - bin_log);
    const uint32_t bin_shift = bin_log - g.bse;
    const uint64_t tiles = (b.n_bases + kTile - 1) / kTile;
    const uint64_t words = tiles * (kTile / 16) + kPkPadWords;  // every tile writes kTile/16 words
    const uint64_t seg_cap = (tiles + kSeg - 1) / kSeg * kTile;  // a tile emits at most kTile records
    const uint64_t cap = seg_cap * kSeg;
    SCB_HIP(ws.pk.reserve(words));
    SCB_HIP(ws.rec.reserve(cap));
    SCB_HIP(ws.sorted.reserve(cap));
    SCB_HIP(ws.bin_count.reserve(nbins));
    SCB_HIP(ws.bin_off.reserve(nbins));
    SCB_HIP(ws.cursor.reserve(nbins));
    SCB_HIP(ws.work.reserve(nbins + cap / per_item + 2));
    SCB_HIP(ws.plan_blk.reserve(2 * ((nbins + kPlanThreads - 1) / kPlanThreads)));
    hipStream_t st = ws.stream;
    // coarse pass by bin >> fine_bits (then a fine pass by bin) when the bins
    // exceed one tile's key range
    const bool two_pass = nbins > kPartRange;
    const uint32_t lb = uint32_t(__builtin_ctzll(nbins));
    const uint32_t fine_bits = sort_fine_bits(lb);
    const uint32_t ncoarse = two_pass ? uint32_t(nbins >> fine_bits) : 0;
    constexpr uint32_t kCopies = 32;  // coarse counters are hot: spread each over 32 copies
    if (two_pass) {
        SCB_HIP(ws.coarse_hist.reserve(ncoarse * kCopies));
        SCB_HIP(ws.ccursor.reserve(ncoarse * kCopies));
    }
    FillBatch fb(st);
    fb.add(reinterpret_cast<uint32_t *>(ws.stats.p), 2 * kStatsLen, 0);
    fb.add(ws.seg_count.p, kSeg + 1, 0);
    fb.add(ws.bin_count.p, nbins, 0);
    if (bin_kmers) fb.add(bin_kmers, nbins, 0);
    fb.add(ws.pk.p + (words - kPkPadWords), kPkPadWords, 0);
    if (two_pass) fb.add(ws.coarse_hist.p, ncoarse * kCopies, 0);
    fb.flush();
    mark(ws, 0);
    Status s = launch_scan<kModeRecords>(L, b, seg_cap, valid_bits, nullptr);
    if (!s.good()) return s;
    mark(ws, 1);
    const int pgrid = cu_count(L.device) * 16;
    const dim3 pg(pgrid), pb(kPartThreads);
    // histograms accumulate many tiles per workgroup; their grid is a multiple of
    // the coarse copy count (each workgroup's tiles then share one copy)
    const dim3 hg(unsigned((cu_count(L.device) * 6 + 31) / 32 * 32));
    SegIn segs{ws.rec.p, ws.seg_count.p, kSeg, seg_cap};
    if (two_pass) {
        KeySpec coarse{bin_shift, fine_bits, 0, 0, kCopies};
        hipLaunchKernelGGL(k_seg_hist, hg, pb, 0, st, segs, coarse, ws.coarse_hist.p, (uint32_t *)nullptr);
        hipLaunchKernelGGL(k_small_scan, dim3(1), dim3(1024), 0, st, ws.coarse_hist.p, ncoarse * kCopies,
                           ws.ccursor.p, ws.stats.p, ws.seg_count.p + kSeg);
        hipLaunchKernelGGL(k_partition, pg, pb, 0, st, segs, coarse, ws.ccursor.p, ws.sorted.p);
        SegIn comp{ws.sorted.p, ws.seg_count.p + kSeg, 1, cap};
        KeySpec fine{bin_shift, 0, fine_bits, 1, 1};
        hipLaunchKernelGGL(k_seg_hist, hg, pb, 0, st, comp, fine, ws.bin_count.p, bin_kmers);
        launch_plan(ws, nbins, per_item, st, bin_kmers);
        hipLaunchKernelGGL(k_partition, pg, pb, 0, st, comp, fine, ws.cursor.p, ws.rec.p);
        ws.final_sorted = ws.rec.p;
    } else {
        KeySpec one{bin_shift, 0, 0, 0, 1};
        hipLaunchKernelGGL(k_seg_hist, hg, pb, 0, st, segs, one, ws.bin_count.p, bin_kmers);
        launch_plan(ws, nbins, per_item, st, bin_kmers);
        hipLaunchKernelGGL(k_partition, pg, pb, 0, st, segs, one, ws.cursor.p, ws.sorted.p);
        ws.final_sorted = ws.sorted.p;
    }
    SCB_HIP(hipGetLastError());
    mark(ws, 2);
    nbins_out = nbins;
    return Status::ok();
}

Status finish(Workspace &ws) {
    SCB_HIP(hipMemcpyAsync(ws.h_stats, ws.stats.p, kStatsLen * sizeof(uint64_t),
                           hipMemcpyDeviceToHost, ws.stream));
    SCB_HIP(hipStreamSynchronize(ws.stream));
    // fold the per-tile spread counters into their totals
    const int direct[kSpCount] = {kStValid, kStSmers, kStQueried, kStPositive, kStChecksum};
    for (int q = 0; q < kSpCount; ++q)
        for (int i = 0; i < kSeg; ++i) ws.h_stats[direct[q]] += ws.h_stats[kStSpread + q * kSeg + i];
    if (ws.h_stats[kStBadOffsets])
        return Status{6, "internal state error: batch offsets must start at 0, never decrease and end at the "
                         "batch's base count"};
    return Status::ok();
}

template <class K>
Status set_lds(K kernel, size_t bytes) {
    SCB_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
    return Status::ok();
}

Status check_batch(const Geometry &, const DeviceBatch &b) {
    if (b.n_bases >= kMaxBatchBases)
        return Status{6, "internal state error: device batch exceeds 2^32 bases"};
    return Status::ok();
}

// The batched path packs the physical block into 27 bits of a record and
// merges whole 32-bit table words per bin.
bool batched_ok(const Geometry &g, uint32_t bin_log) {
    return g.bse <= bin_log && g.nb <= (uint64_t(1) << 27) && (uint64_t(1) << bin_log) * g.bytes >= 4;
}

// Insert bin size for one batch: the tuned size, halved while the batch's
// expected records per bin exceed 90% of a work item. A bin with more records
// than one item is split over several items, which merge into the table with
// global atomics instead of one read-modify-write of the bin; for a large batch
// smaller bins that stay whole cost less than that merge. Records are expected
// at the random-minimizer density 2 / (w + 1) per valid window, plus one per
// read; the choice only affects speed, never the counters.
uint32_t insert_bin_log_for(const Geometry &g, uint32_t tuned, uint32_t per_item, uint64_t n_bases,
                            uint64_t n_reads) {
    const uint64_t lost = n_reads * uint64_t(g.k - 1);
    const double windows = n_bases > lost ? double(n_bases - lost) : 0.0;
    const double records = windows * 2.0 / double(g.w + 1) + double(n_reads);
    uint32_t bl = tuned;
    const uint32_t floor_log = std::max<uint32_t>(g.bse, 12);
    while (bl > floor_log && batched_ok(g, bl - 1) &&
           records / double(uint64_t(1) << (g.cse - bl)) > 0.9 * double(per_item))
        --bl;
    return bl;
}

}  // namespace

namespace {
uint64_t key_of(std::initializer_list<uint64_t> v) {
    uint64_t h = 0xCBF29CE484222325ull;
    for (uint64_t x : v) {
        h = (h ^ x) * 0x100000001B3ull;
        h ^= h >> 29;
    }
    return h;
}

// Runs body (the launches of one batched pipeline, ending before the stats
// copy) eagerly, or from a HIP graph: the first call with a key runs eagerly
// and leaves its buffers sized; the next call with the same key (and so the
// same buffers) is captured and launched; later ones replay the graph. Stage
// events are recorded only around a replay (total time).
template <class Body>
Status run_graphable(Workspace &ws, bool use_graph, Workspace::GraphSlot &slot, uint64_t args_key, Body &&body) {
    ws.stages_timed = true;
    if (!use_graph || ws.is_view) return body();
    const uint64_t key = key_of({args_key, ws.buffer_signature()});
    if (slot.exec && slot.key == key) {
        ws.stages_timed = false;
        if (ws.ev[0]) (void)hipEventRecord(ws.ev[0], ws.stream);
        SCB_HIP(hipGraphLaunch(slot.exec, ws.stream));
        if (ws.ev[3]) (void)hipEventRecord(ws.ev[3], ws.stream);
        return Status::ok();
    }
    if (slot.eager_key != key) {
        Status s = body();
        slot.eager_key = key_of({args_key, ws.buffer_signature()});
        return s;
    }
    slot.drop();
    ws.capturing = true;
    hipError_t e = hipStreamBeginCapture(ws.stream, hipStreamCaptureModeRelaxed);
    Status s = e == hipSuccess ? body() : Status{8, std::string("device error: hipStreamBeginCapture: ") +
                                                       hipGetErrorString(e)};
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(ws.stream, &g);
    ws.capturing = false;
    if (s.good() && e2 != hipSuccess)
        s = Status{8, std::string("device error: hipStreamEndCapture: ") + hipGetErrorString(e2)};
    if (s.good()) {
        hipError_t e3 = hipGraphInstantiate(&slot.exec, g, nullptr, nullptr, 0);
        if (e3 != hipSuccess) s = Status{8, std::string("device error: hipGraphInstantiate: ") + hipGetErrorString(e3)};
    }
    if (!s.good()) {
        if (g) (void)hipGraphDestroy(g);
        slot.exec = nullptr;
        slot.eager_key = 0;
        return s;
    }
    slot.graph = g;
    slot.key = key;
    ws.stages_timed = false;
    if (ws.ev[0]) (void)hipEventRecord(ws.ev[0], ws.stream);
    SCB_HIP(hipGraphLaunch(slot.exec, ws.stream));
    if (ws.ev[3]) (void)hipEventRecord(ws.ev[3], ws.stream);
    return Status::ok();
}
}  // namespace

Status Engine::insert(const DeviceBatch &b, Method method, InsertResult &r) {
    r = InsertResult{};
    method = resolve(method, b.n_bases);
    Status s = check_batch(geo_, b);
    if (!s.good()) return s;
    if (b.n_bases < geo_.k || b.n_reads == 0) return Status::ok();
    SCB_HIP(hipSetDevice(device_));
    // the 16-bit LDS counts of a bin need at least two slots (one word)
    bool batched = method == kBatched && batched_ok(geo_, tune_.insert_bin_log) && tune_.insert_bin_log >= 1;
    bool fresh = false;
    {
        std::lock_guard<std::mutex> lk(zero_mu_);
        fresh = zero_pending_;
    }
    if (fresh && !batched) {
        s = ensure_zeroed();
        if (!s.good()) return s;
        fresh = false;
    }
    auto ws = acquire();
    if (!ws) return Status{8, "device error: cannot create a workspace"};
    Launch L{geo_, dev_geo(geo_), *ws, d_table_, device_};
    hipStream_t st = ws->stream;
    if (!batched) {
        fill(reinterpret_cast<uint32_t *>(ws->stats.p), 2 * kStatsLen, 0, st);
        mark(*ws, 0);
        s = launch_scan<kModePlainInsert>(L, b, 0, nullptr, nullptr);
        mark(*ws, 1);
    } else {
        const uint32_t per_item =
            std::max<uint32_t>(1, std::min<uint32_t>(tune_.insert_item_records, kInsertItemRecords));
        const uint32_t bin_log = insert_bin_log_for(geo_, tune_.insert_bin_log, per_item, b.n_bases, b.n_reads);
        const uint64_t args =
            key_of({1, uint64_t(uintptr_t(b.d_bases)), uint64_t(uintptr_t(b.d_offsets)), b.n_reads, b.n_bases,
                    uint64_t(uintptr_t(d_table_)), uint64_t(fresh), bin_log, tune_.insert_item_records});
        s = run_graphable(*ws, tune_.use_graph, ws->graph_insert, args, [&]() -> Status {
        Status s;
        uint64_t nbins = 0;
        // k-mers per bin select the bins counted by distinct record
        SCB_HIP(ws->spec_bin_kmers.reserve(uint64_t(1) << (geo_.cse - bin_log)));
        s = batched_front(L, b, bin_log, per_item, nullptr, nbins, ws->spec_bin_kmers.p);
        if (s.good()) {
            BinArgs a{};
            a.sorted = ws->final_sorted;
            a.bin_off = ws->bin_off.p;
            a.bin_count = ws->bin_count.p;
            a.work = ws->work.p;
            a.stats = ws->stats.p;
            a.pk = ws->pk.p;
            a.table = d_table_;
            a.bin_log = bin_log;
            a.per_item = per_item;
            a.bin_kmers = ws->spec_bin_kmers.p;
            a.fresh = fresh;
            if (fresh) {
                const uint32_t words_per_bin = uint32_t(((uint64_t(1) << bin_log) * geo_.bytes) / 4);
                hipLaunchKernelGGL(k_zero_cold, dim3(unsigned(std::min<uint64_t>(nbins, 8192))), dim3(256), 0, st,
                                   ws->bin_count.p, nbins, a.per_item, reinterpret_cast<uint32_t *>(d_table_),
                                   words_per_bin);
            }
            // 16-bit counts; the distinct-record search uses the same LDS first
            size_t lds = std::max<size_t>((size_t(1) << bin_log) * 2, size_t(kRecDedupWords) * 4);
            const int thr = kInsertThreads;
#define SCB_INS(BITS)                                                                           \
    do {                                                                                        \
        s = set_lds(k_bin_insert<BITS>, lds);                                                   \
        if (!s.good()) break;                                                                   \
        int grid = persistent_grid(k_bin_insert<BITS>, thr, lds, device_);                      \
        hipLaunchKernelGGL(k_bin_insert<BITS>, dim3(grid), dim3(thr), lds, st, a, L.dg);        \
    } while (0)
            switch (geo_.bits) {
                case 8: SCB_INS(8); break;
                case 16: SCB_INS(16); break;
                default: SCB_INS(32); break;
            }
#undef SCB_INS
            if (s.good()) {
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) s = Status{8, std::string("device error: k_bin_insert: ") + hipGetErrorString(e)};
                mark(*ws, 3);
            }
        }
        return s;
        });
    }
    if (s.good()) s = finish(*ws);
    if (s.good() && fresh) {
        std::lock_guard<std::mutex> lk(zero_mu_);
        zero_pending_ = false;
    }
    if (s.good()) {
        collect_times(*ws, batched);
        r.inserted_kmers = ws->h_stats[kStValid];
        r.inserted_smers = ws->h_stats[kStSmers];
        r.super_kmers = ws->h_stats[kStRecords];
    }
    release(std::move(ws));
    return s;
}

Status Engine::query(const DeviceBatch &b, Method method, void *d_values, uint32_t *d_valid_bits,
                     QueryResult &r) const {
    r = QueryResult{};
    method = resolve(method, b.n_bases);
    Status s = check_batch(geo_, b);
    if (!s.good()) return s;
    s = ensure_zeroed();
    if (!s.good()) return s;
    if (b.n_bases == 0 || b.n_reads == 0) return Status::ok();
    SCB_HIP(hipSetDevice(device_));
    auto ws = acquire();
    if (!ws) return Status{8, "device error: cannot create a workspace"};
    s = run_query(*ws, b, method, d_values, d_valid_bits, r);
    release(std::move(ws));
    return s;
}

// The query pipeline on one workspace (pooled, or a view of the fixed scratch
// array); the stream is synchronized before it returns.
Status Engine::run_query(Workspace &wsr, const DeviceBatch &b, Method method, void *d_values,
                         uint32_t *d_valid_bits, QueryResult &r) const {
    Status s;
    Workspace *ws = &wsr;
    Launch L{geo_, dev_geo(geo_), *ws, d_table_, device_};
    hipStream_t st = ws->stream;
    if (d_valid_bits) fill(d_valid_bits, (b.n_bases + 31) / 32, 0, st);
    bool batched = method == kBatched && batched_ok(geo_, tune_.query_bin_log);
    if (b.n_bases < geo_.k) {
        fill(reinterpret_cast<uint32_t *>(ws->stats.p), 2 * kStatsLen, 0, st);
    } else if (!batched) {
        fill(reinterpret_cast<uint32_t *>(ws->stats.p), 2 * kStatsLen, 0, st);
        // the plain scan writes validity itself
        mark(*ws, 0);
        s = launch_scan<kModePlainQuery>(L, b, 0, d_valid_bits, d_values);
        mark(*ws, 1);
    } else {
        const uint64_t args =
            key_of({2, uint64_t(uintptr_t(b.d_bases)), uint64_t(uintptr_t(b.d_offsets)), b.n_reads, b.n_bases,
                    uint64_t(uintptr_t(d_table_)), uint64_t(uintptr_t(d_values)), uint64_t(uintptr_t(d_valid_bits)),
                    tune_.query_bin_log, tune_.query_item_records});
        s = run_graphable(*ws, tune_.use_graph, ws->graph_query, args, [&]() -> Status {
        Status s;
        uint64_t nbins = 0;
        const uint32_t bin_log = tune_.query_bin_log;
        const uint32_t per_item =
            std::max<uint32_t>(1, std::min<uint32_t>(tune_.query_item_records, kQueryItemRecords));
        s = batched_front(L, b, bin_log, per_item, d_valid_bits, nbins);
        if (s.good()) {
            BinArgs a{};
            a.sorted = ws->final_sorted;
            a.bin_off = ws->bin_off.p;
            a.bin_count = ws->bin_count.p;
            a.work = ws->work.p;
            a.stats = ws->stats.p;
            a.pk = ws->pk.p;
            a.table = d_table_;
            a.out = d_values;
            a.bin_log = bin_log;
            a.per_item = per_item;
            size_t lds = (size_t(1) << bin_log) * geo_.bytes;
            const int thr = kQueryThreads;
#define SCB_QRY(BITS, WC, EX)                                                                          \
    do {                                                                                               \
        s = set_lds(k_bin_query<BITS, WC, EX>, lds);                                                   \
        if (!s.good()) break;                                                                          \
        int grid = persistent_grid(k_bin_query<BITS, WC, EX>, thr, lds, device_);                      \
        hipLaunchKernelGGL((k_bin_query<BITS, WC, EX>), dim3(grid), dim3(thr), lds, st, a, L.dg);      \
    } while (0)
#define SCB_QRY_W(BITS)                                                                                \
    switch (geo_.win) {  /* fimpera window k - s + 1: exact up to 8, else up to 32 */                  \
        case 1: SCB_QRY(BITS, 1, true); break;                                                         \
        case 2: SCB_QRY(BITS, 2, true); break;                                                         \
        case 3: SCB_QRY(BITS, 3, true); break;                                                         \
        case 4: SCB_QRY(BITS, 4, true); break;                                                         \
        case 5: SCB_QRY(BITS, 5, true); break;                                                         \
        case 6: SCB_QRY(BITS, 6, true); break;                                                         \
        case 7: SCB_QRY(BITS, 7, true); break;                                                         \
        case 8: SCB_QRY(BITS, 8, true); break;                                                         \
        default: SCB_QRY(BITS, 32, false); break;                                                      \
    }
            switch (geo_.bits) {
                case 8: SCB_QRY_W(8); break;
                case 16: SCB_QRY_W(16); break;
                default: SCB_QRY_W(32); break;
            }
#undef SCB_QRY_W
#undef SCB_QRY
            if (s.good()) {
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) s = Status{8, std::string("device error: k_bin_query: ") + hipGetErrorString(e)};
                mark(*ws, 3);
            }
        }
        return s;
        });
    }
    if (s.good()) s = finish(*ws);
    if (s.good()) {
        if (b.n_bases >= geo_.k && ws->ev[0]) collect_times(*ws, batched);
        r.queried = ws->h_stats[kStQueried];
        r.positive = ws->h_stats[kStPositive];
        r.checksum = ws->h_stats[kStChecksum];
        r.super_kmers = ws->h_stats[kStRecords];
    }
    return s;
}

// ---- allocation-free query ---------------------------------------------------
// query_noalloc runs the same kernels as query with every per-call device
// buffer carved from `scratch`, a fixed array that each GPU holds from the
// first call until the process ends. Host code never uses the array's name as
// an address: hipGetSymbolAddress gives each device's own copy once, with that
// device current, and that pointer is what the kernels receive.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mu;
uint8_t *g_scratch_ptr[kNoAllocDevices] = {};
std::mutex g_noalloc_call_mu[kNoAllocDevices];  // one call at a time uses a device's array

// Address of the current device's copy of `scratch`, looked up on the
// device's first call.
Status scratch_address(int dev, uint8_t *&out) {
    if (dev < 0 || dev >= kNoAllocDevices)
        return Status{8, "device error: device index beyond the allocation-free entry point's table"};
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return Status{8, std::string("device error: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        g_scratch_ptr[dev] = static_cast<uint8_t *>(p);
    }
    out = g_scratch_ptr[dev];
    return Status::ok();
}

// Points the buffers run_query uses at 256-byte aligned pieces of the scratch
// array, sized as batched_front and launch_scan size them for this batch.
// Returns the bytes needed (which may exceed cap: then nothing may run).
uint64_t carve_query(uint8_t *base, const Geometry &g, uint64_t n_bases, uint32_t bin_log, uint32_t per_item,
                     bool batched, Workspace &ws) {
    uint64_t off = 0;
    auto take = [&](auto &buf, uint64_t count) {
        using T = std::remove_reference_t<decltype(*buf.p)>;
        off = (off + 255) & ~uint64_t(255);
        buf.view(reinterpret_cast<T *>(base + off), count);
        off += std::max<uint64_t>(count, 1) * sizeof(T);
    };
    const uint64_t tiles = (n_bases + kTile - 1) / kTile;
    take(ws.stats, kStatsLen);
    take(ws.seg_count, kSeg + 1);
    take(ws.tile_r0, std::max<uint64_t>(tiles, 1));
    if (batched) {
        const uint64_t nbins = uint64_t(1) << (g.cse - bin_log);
        const uint64_t words = tiles * (kTile / 16) + kPkPadWords;
        const uint64_t seg_cap = (tiles + kSeg - 1) / kSeg * kTile;
        const uint64_t cap = seg_cap * kSeg;
        take(ws.pk, words);
        take(ws.rec, cap);
        take(ws.sorted, cap);
        take(ws.bin_count, nbins);
        take(ws.bin_off, nbins);
        take(ws.cursor, nbins);
        take(ws.work, nbins + cap / per_item + 2);
        take(ws.plan_blk, 2 * ((nbins + kPlanThreads - 1) / kPlanThreads));
        if (nbins > kPartRange) {
            const uint32_t lb = uint32_t(__builtin_ctzll(nbins));
            const uint32_t fine_bits = sort_fine_bits(lb);
            const uint64_t ncoarse = nbins >> fine_bits;
            take(ws.coarse_hist, ncoarse * 32);
            take(ws.ccursor, ncoarse * 32);
        }
    }
    return (off + 255) & ~uint64_t(255);
}
}  // namespace

Status Engine::query_noalloc(const DeviceBatch &b, Method method, void *d_values, uint32_t *d_valid_bits,
                             QueryResult &r) const {
    r = QueryResult{};
    method = resolve(method, b.n_bases);
    if (b.n_bases > kNoAllocMaxBases)
        return Status{6, "internal state error: the allocation-free query takes at most " +
                             std::to_string(kNoAllocMaxBases) + " bases per call (" + std::to_string(b.n_bases) +
                             " given)"};
    Status s = ensure_zeroed();
    if (!s.good()) return s;
    if (b.n_bases == 0 || b.n_reads == 0) return Status::ok();
    SCB_HIP(hipSetDevice(device_));
    uint8_t *base = nullptr;
    s = scratch_address(device_, base);
    if (!s.good()) return s;
    std::lock_guard<std::mutex> lk(g_noalloc_call_mu[device_]);
    Workspace ws;  // null stream, no events, no owned buffers
    ws.make_view();
    uint64_t host_stats[kStatsLen];
    ws.h_stats = host_stats;
    const bool batched = method == kBatched && batched_ok(geo_, tune_.query_bin_log) && b.n_bases >= geo_.k;
    const uint32_t per_item =
        std::max<uint32_t>(1, std::min<uint32_t>(tune_.query_item_records, kQueryItemRecords));
    const uint64_t need = carve_query(base, geo_, b.n_bases, tune_.query_bin_log, per_item, batched, ws);
    if (need > kNoAllocScratchBytes)
        return Status{6, "internal state error: the batch needs " + std::to_string(need) +
                             " bytes of scratch, more than the " + std::to_string(kNoAllocScratchBytes) +
                             " bytes of the allocation-free query"};
    s = run_query(ws, b, method, d_values, d_valid_bits, r);
    if (!s.good()) (void)hipStreamSynchronize(nullptr);  // nothing may still use the array
    return s;
}


// Distinct canonical k-mers accumulated over several spectrum calls: a device
// open-addressing set of the keys (grown and rehashed before it would pass
// half full) and the histogram of their estimates.
class KeySet {
   public:
    int device = 0;
    DevBuf<unsigned long long> set;
    uint32_t log2 = 0;
    uint64_t size = 0;     // keys held
    uint64_t queried = 0;  // k-mers queried over all calls
    HistTables hist;
    DevBuf<uint64_t> stats;  // [0] keys held, [kStOverflow] big-value table overflow
};

void KeySetDeleter::operator()(KeySet *k) const { delete k; }

Status Engine::keyset_create(KeySetPtr &out) const {
    SCB_HIP(hipSetDevice(device_));
    KeySetPtr ks(new KeySet);
    ks->device = device_;
    ks->log2 = 16;
    SCB_HIP(ks->set.reserve(uint64_t(1) << ks->log2));
    SCB_HIP(ks->hist.reserve(geo_));
    SCB_HIP(ks->stats.reserve(kStatsLen));
    fill(reinterpret_cast<uint32_t *>(ks->set.p), (uint64_t(1) << ks->log2) * 2, 0xFFFFFFFFu, nullptr);
    ks->hist.clear(geo_, nullptr);
    fill(reinterpret_cast<uint32_t *>(ks->stats.p), 2 * kStatsLen, 0, nullptr);
    SCB_HIP(hipGetLastError());
    SCB_HIP(hipStreamSynchronize(nullptr));
    out = std::move(ks);
    return Status::ok();
}

Status Engine::spectrum_checks(const DeviceBatch &b) const {
    Status s = check_batch(geo_, b);
    if (!s.good()) return s;
    s = ensure_zeroed();
    if (!s.good()) return s;
    if (geo_.nb > (uint64_t(1) << 27))
        return Status{6, "internal state error: the GPU spectrum supports at most 2^27 counter blocks"};
    if (geo_.bse >= 32)  // slot addresses inside a block are 32-bit in the spectrum kernel
        return Status{6, "internal state error: the GPU spectrum supports blocks of at most 2^31 slots"};
    return Status::ok();
}

// One spectrum pass over a batch (filter.rs:407-460 per record): records
// binned by counter bin, each bin's keys deduplicated in LDS (or, for split or
// oversized bins, in a per-call global set); first occurrences are counted
// into `hist` or, with emit, appended as (key, estimate) to ws.emit_*. The
// stream is synchronized on return; h_stats hold the counts.
Status Engine::spectrum_pass(Workspace &ws, const DeviceBatch &b, const void *hist, bool emit) const {
    Launch L{geo_, dev_geo(geo_), ws, d_table_, device_};
    hipStream_t st = ws.stream;
    const uint32_t bin_log = geo_.bse <= 14 ? std::min<uint32_t>(tune_.spectrum_bin_log, geo_.cse) : geo_.bse;
    const uint32_t stage = (geo_.bse <= 14 && (uint64_t(1) << bin_log) * geo_.bytes >= 4) ? 1 : 0;
    const uint32_t per_item = std::max<uint32_t>(1, std::min<uint32_t>(tune_.spectrum_item_records, kMaxItemRecords));
    const uint64_t nbins = uint64_t(1) << (geo_.cse - bin_log);
    SCB_HIP(ws.spec_bin_kmers.reserve(nbins));
    if (emit) {
        SCB_HIP(ws.emit_keys.reserve(b.n_bases + 1));
        SCB_HIP(ws.emit_vals.reserve(b.n_bases + 1));
    }
    uint64_t nb_unused = 0;
    Status s = batched_front(L, b, bin_log, per_item, nullptr, nb_unused, ws.spec_bin_kmers.p);
    if (!s.good()) return s;
    hipLaunchKernelGGL(k_spectrum_plan, dim3(1), dim3(1024), 0, st, ws.bin_count.p, ws.spec_bin_kmers.p, nbins,
                       per_item, ws.stats.p);
    if ((s = finish(ws)).code) return s;
    // Global set: a slot per k-mer of the bins that use it directly, twice
    // over, plus room for the keys of LDS-set bins whose probe windows fill
    // (a 64th of the batch's k-mers). A pass whose global set overflows runs
    // again with a slot per k-mer of the batch, four times over.
    const uint64_t gk = ws.h_stats[kStGlobalKmers], all = ws.h_stats[kStValid];
    SpecArgs sa{};
    sa.b.sorted = ws.final_sorted;
    sa.b.bin_off = ws.bin_off.p;
    sa.b.bin_count = ws.bin_count.p;
    sa.b.work = ws.work.p;
    sa.b.stats = ws.stats.p;
    sa.b.pk = ws.pk.p;
    sa.b.table = d_table_;
    sa.b.bin_log = bin_log;
    sa.b.per_item = per_item;
    sa.bin_kmers = ws.spec_bin_kmers.p;
    if (hist) sa.h = *static_cast<const HistOut *>(hist);
    sa.h.stats = ws.stats.p;
    sa.stage = stage;
    if (emit) {
        sa.emit_keys = ws.emit_keys.p;
        sa.emit_vals = ws.emit_vals.p;
        sa.emit_cap = b.n_bases + 1;
    }
    size_t lds = (stage ? (size_t(1) << bin_log) * geo_.bytes : 0) + kSetCap * 8 + kSpecHist * 4;
    const int thr = kSpectrumThreads;
    const bool wide = geo_.win > 8;
    for (int attempt = 0; attempt < 2; ++attempt) {
        const uint64_t want = attempt == 0 ? 2 * (gk + all / 64) + 2 : 4 * all + 2;
        uint32_t glog2 = 10;
        while ((uint64_t(1) << glog2) < want) ++glog2;
        SCB_HIP(ws.spec_gset.reserve(uint64_t(1) << glog2));
        fill(reinterpret_cast<uint32_t *>(ws.spec_gset.p), (uint64_t(1) << glog2) * 2, 0xFFFFFFFFu, st);
        if (attempt > 0) {
            // start over: the kernel's counts, the work queue and the histogram
            for (int slot : {int(kStWorkHead), int(kStQueried), int(kStDistinct), int(kStEmitted), int(kStOverflow),
                             int(kStSetFull)})
                fill(reinterpret_cast<uint32_t *>(ws.stats.p + slot), 2, 0, st);
            if (hist) ws.hist.clear(geo_, st);
        }
        sa.gset = ws.spec_gset.p;
        sa.glog2 = glog2;
#define SCB_SPEC(BITS, WC)                                                                      \
    do {                                                                                        \
        s = set_lds(k_bin_spectrum<BITS, WC>, lds);                                             \
        if (!s.good()) break;                                                                   \
        int grid = persistent_grid(k_bin_spectrum<BITS, WC>, thr, lds, device_);                \
        hipLaunchKernelGGL((k_bin_spectrum<BITS, WC>), dim3(grid), dim3(thr), lds, st, sa, L.dg); \
    } while (0)
        switch (geo_.bits) {
            case 8: if (wide) SCB_SPEC(8, 32); else SCB_SPEC(8, 8); break;
            case 16: if (wide) SCB_SPEC(16, 32); else SCB_SPEC(16, 8); break;
            default: if (wide) SCB_SPEC(32, 32); else SCB_SPEC(32, 8); break;
        }
#undef SCB_SPEC
        if (!s.good()) return s;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return Status{8, std::string("device error: k_bin_spectrum: ") + hipGetErrorString(e)};
        mark(ws, 3);
        if ((s = finish(ws)).code) return s;
        if (!ws.h_stats[kStSetFull]) break;
        if (attempt == 1) return Status{8, "device error: the spectrum's global key set overflowed"};
    }
    collect_times(ws, true);  // the bin stage includes the plan's host round trip
    if (ws.h_stats[kStOverflow]) return Status{8, "device error: spectrum has more than 2^20 distinct large estimates"};
    if (emit && ws.h_stats[kStEmitted] > b.n_bases + 1)
        return Status{6, "internal state error: spectrum emitted more keys than the batch has k-mers"};
    return Status::ok();
}

// The plain spectrum: one thread per super-k-mer estimates its k-mers from the
// table in HBM and inserts their canonical keys into one key set for the batch
// (a slot per base, so it stays at most half full); a key counts when it
// enters the set.
Status Engine::spectrum_plain(Workspace &ws, const DeviceBatch &b, const void *hist) const {
    Launch L{geo_, dev_geo(geo_), ws, d_table_, device_};
    hipStream_t st = ws.stream;
    fill(reinterpret_cast<uint32_t *>(ws.stats.p), 2 * kStatsLen, 0, st);
    uint32_t glog2 = 10;
    while ((uint64_t(1) << glog2) < 2 * b.n_bases + 2) ++glog2;
    SCB_HIP(ws.spec_gset.reserve(uint64_t(1) << glog2));
    fill(reinterpret_cast<uint32_t *>(ws.spec_gset.p), (uint64_t(1) << glog2) * 2, 0xFFFFFFFFu, st);
    HistOut h = *static_cast<const HistOut *>(hist);
    h.stats = ws.stats.p;
    mark(ws, 0);
    Status s = launch_scan<kModePlainSpectrum>(L, b, 0, nullptr, nullptr, ws.spec_gset.p, glog2, &h);
    if (!s.good()) return s;
    mark(ws, 1);
    if ((s = finish(ws)).code) return s;
    collect_times(ws, false);
    if (ws.h_stats[kStOverflow]) return Status{8, "device error: spectrum has more than 2^20 distinct large estimates"};
    return Status::ok();
}

Status Engine::spectrum(const DeviceBatch &b, SpectrumResult &r, Method method) const {
    r = SpectrumResult{};
    Status s = spectrum_checks(b);
    if (!s.good()) return s;
    if (b.n_bases < geo_.k || b.n_reads == 0) return Status::ok();
    const bool plain = resolve(method, b.n_bases) == kPlain;
    SCB_HIP(hipSetDevice(device_));
    auto wsp = acquire();
    if (!wsp) return Status{8, "device error: cannot create a workspace"};
    Workspace &ws = *wsp;
    do {
        hipError_t e = ws.hist.reserve(geo_);
        if (e != hipSuccess) {
            s = Status{8, std::string("device error: ") + hipGetErrorString(e)};
            break;
        }
        ws.hist.clear(geo_, ws.stream);
        HistOut h = ws.hist.out(ws.stats.p);
        if ((s = plain ? spectrum_plain(ws, b, &h) : spectrum_pass(ws, b, &h, false)).code) break;
        if ((s = ws.hist.rows(geo_, ws.stream, r.rows)).code) break;
        r.queried = ws.h_stats[kStQueried];
        r.distinct = ws.h_stats[kStDistinct];
    } while (false);
    release(std::move(wsp));
    return s;
}

Status Engine::spectrum_accumulate(const DeviceBatch &b, KeySet &ks) const {
    Status s = spectrum_checks(b);
    if (!s.good()) return s;
    if (ks.device != device_) return Status{6, "internal state error: the key set belongs to another device"};
    if (b.n_bases < geo_.k || b.n_reads == 0) return Status::ok();
    SCB_HIP(hipSetDevice(device_));
    auto wsp = acquire();
    if (!wsp) return Status{8, "device error: cannot create a workspace"};
    Workspace &ws = *wsp;
    hipStream_t st = ws.stream;
    do {
        if ((s = spectrum_pass(ws, b, nullptr, true)).code) break;
        const uint64_t ne = ws.h_stats[kStEmitted];
        ks.queried += ws.h_stats[kStQueried];
        if (!ne) break;
        // keep the set at most half full: grow (powers of two) and rehash
        uint32_t want = ks.log2;
        while ((uint64_t(1) << want) < 2 * (ks.size + ne)) ++want;
        if (want > 31) {
            s = Status{8, "device error: the spectrum key set would exceed 2^30 distinct keys"};
            break;
        }
        if (want != ks.log2) {
            DevBuf<unsigned long long> bigger;
            hipError_t e = bigger.reserve(uint64_t(1) << want);
            if (e != hipSuccess) {
                s = Status{8, std::string("device error: growing the spectrum key set: ") + hipGetErrorString(e)};
                break;
            }
            fill(reinterpret_cast<uint32_t *>(bigger.p), (uint64_t(1) << want) * 2, 0xFFFFFFFFu, st);
            const uint64_t old_cap = uint64_t(1) << ks.log2;
            hipLaunchKernelGGL(k_keyset_rehash, dim3(unsigned(std::min<uint64_t>((old_cap + 255) / 256, 16384))),
                               dim3(256), 0, st, ks.set.p, old_cap, bigger.p, want);
            SCB_HIP(hipStreamSynchronize(st));  // the old set is freed below
            std::swap(ks.set.p, bigger.p);
            std::swap(ks.set.n, bigger.n);
            ks.log2 = want;
        }
        hipLaunchKernelGGL(k_keyset_merge, dim3(unsigned(std::min<uint64_t>((ne + 255) / 256, 16384))), dim3(256), 0,
                           st, ws.emit_keys.p, ws.emit_vals.p, ne, ks.set.p, ks.log2, ks.hist.out(ks.stats.p),
                           reinterpret_cast<unsigned long long *>(ks.stats.p));
        SCB_HIP(hipGetLastError());
        uint64_t head[kStOverflow + 1];
        SCB_HIP(hipMemcpyAsync(head, ks.stats.p, sizeof(head), hipMemcpyDeviceToHost, st));
        SCB_HIP(hipStreamSynchronize(st));
        ks.size = head[0];
        if (head[kStOverflow]) s = Status{8, "device error: spectrum has more than 2^20 distinct large estimates"};
    } while (false);
    release(std::move(wsp));
    return s;
}

Status Engine::keyset_rows(const KeySet &ks, SpectrumResult &r) const {
    r = SpectrumResult{};
    SCB_HIP(hipSetDevice(device_));
    Status s = ks.hist.rows(geo_, nullptr, r.rows);
    r.queried = ks.queried;
    r.distinct = ks.size;
    return s;
}

Status Engine::kmer_map(const DeviceBatch &b, std::vector<std::pair<uint64_t, uint32_t>> &pairs,
                        uint64_t &queried) const {
    pairs.clear();
    queried = 0;
    Status s = spectrum_checks(b);
    if (!s.good()) return s;
    if (b.n_bases < geo_.k || b.n_reads == 0) return Status::ok();
    SCB_HIP(hipSetDevice(device_));
    auto wsp = acquire();
    if (!wsp) return Status{8, "device error: cannot create a workspace"};
    Workspace &ws = *wsp;
    do {
        if ((s = spectrum_pass(ws, b, nullptr, true)).code) break;
        const uint64_t ne = ws.h_stats[kStEmitted];
        queried = ws.h_stats[kStQueried];
        std::vector<uint64_t> keys(ne);
        std::vector<uint32_t> vals(ne);
        hipError_t e = hipMemcpy(keys.data(), ws.emit_keys.p, ne * 8, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(vals.data(), ws.emit_vals.p, ne * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            s = Status{8, std::string("device error: copying the k-mer map: ") + hipGetErrorString(e)};
            break;
        }
        pairs.resize(ne);
        for (uint64_t i = 0; i < ne; ++i) pairs[i] = {keys[i], vals[i]};
    } while (false);
    release(std::move(wsp));
    return s;
}

Status Engine::counter_histogram(std::vector<std::pair<uint64_t, uint64_t>> &rows) const {
    rows.clear();
    Status z = ensure_zeroed();
    if (!z.good()) return z;
    SCB_HIP(hipSetDevice(device_));
    auto ws = acquire();
    if (!ws) return Status{8, "device error: cannot create a workspace"};
    hipStream_t st = ws->stream;
    uint64_t dense_n = geo_.bits == 8 ? 256 : (geo_.bits == 16 ? 65536 : kHistSmall);
    Status s;
    do {
        hipError_t e;
        if ((e = ws->hist_dense.reserve(dense_n)) != hipSuccess ||
            (e = ws->hist_big_keys.reserve(geo_.bits == 32 ? kHistBigCap : 1)) != hipSuccess ||
            (e = ws->hist_big_counts.reserve(geo_.bits == 32 ? kHistBigCap : 1)) != hipSuccess) {
            s = Status{8, std::string("device error: ") + hipGetErrorString(e)};
            break;
        }
        fill(reinterpret_cast<uint32_t *>(ws->hist_dense.p), dense_n * 2, 0, st);
        fill(reinterpret_cast<uint32_t *>(ws->stats.p), 2 * kStatsLen, 0, st);
        if (geo_.bits == 32) {
            fill(ws->hist_big_keys.p, kHistBigCap, 0, st);
            fill(reinterpret_cast<uint32_t *>(ws->hist_big_counts.p), 2 * kHistBigCap, 0, st);
        }
        int grid = cu_count(device_) * 8;
#define SCB_HIST(BITS)                                                                         \
    hipLaunchKernelGGL(k_counter_hist<BITS>, dim3(grid), dim3(256), 0, st, d_table_, geo_.slots, \
                       ws->hist_dense.p, ws->hist_big_keys.p, ws->hist_big_counts.p, ws->stats.p)
        switch (geo_.bits) {
            case 8: SCB_HIST(8); break;
            case 16: SCB_HIST(16); break;
            default: SCB_HIST(32); break;
        }
#undef SCB_HIST
        if ((e = hipGetLastError()) != hipSuccess) {
            s = Status{8, std::string("device error: k_counter_hist: ") + hipGetErrorString(e)};
            break;
        }
        std::vector<unsigned long long> dense(dense_n);
        std::vector<uint32_t> keys;
        std::vector<unsigned long long> cnts;
        if ((e = hipMemcpyAsync(dense.data(), ws->hist_dense.p, dense_n * 8, hipMemcpyDeviceToHost, st)) != hipSuccess) {
            s = Status{8, std::string("device error: ") + hipGetErrorString(e)};
            break;
        }
        if (geo_.bits == 32) {
            keys.resize(kHistBigCap);
            cnts.resize(kHistBigCap);
            (void)hipMemcpyAsync(keys.data(), ws->hist_big_keys.p, kHistBigCap * 4, hipMemcpyDeviceToHost, st);
            (void)hipMemcpyAsync(cnts.data(), ws->hist_big_counts.p, kHistBigCap * 8, hipMemcpyDeviceToHost, st);
        }
        s = finish(*ws);
        if (!s.good()) break;
        if (ws->h_stats[kStOverflow]) {
            s = Status{8, "device error: counter histogram has more than 2^20 distinct large values"};
            break;
        }
        for (uint64_t v = 0; v < dense_n; ++v)
            if (dense[v]) rows.emplace_back(v, dense[v]);
        for (uint32_t i = 0; i < keys.size(); ++i)
            if (keys[i]) rows.emplace_back(keys[i], cnts[i]);
        std::sort(rows.begin(), rows.end());
    } while (false);
    release(std::move(ws));
    return s;
}

// ---- host-memory batches ---------------------------------------------------
namespace {

// Splits [0, n) into chunks whose byte size stays below limit (a single record
// larger than limit is an error).
template <class Fn>
Status for_each_chunk(const uint64_t *offsets, uint64_t n, uint64_t limit, Fn &&fn) {
    uint64_t r = 0;
    while (r < n) {
        uint64_t e = r;
        while (e < n && offsets[e + 1] - offsets[r] <= limit) ++e;
        if (e == r) return Status{6, "internal state error: a record exceeds the device batch limit"};
        Status s = fn(r, e);
        if (!s.good()) return s;
        r = e;
    }
    return Status::ok();
}

}  // namespace

Status Engine::insert_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n, Method method,
                           InsertResult &r) {
    r = InsertResult{};
    SCB_HIP(hipSetDevice(device_));
    auto st = acquire();  // staging buffers, kept with the pooled workspace
    if (!st) return Status{8, "device error: cannot create a workspace"};
    const uint64_t limit = std::min<uint64_t>(kMaxBatchBases, uint64_t(1) << 30);
    Status s = for_each_chunk(offsets, n, limit, [&](uint64_t r0, uint64_t r1) -> Status {
        const uint64_t nb = offsets[r1] - offsets[r0];
        SCB_HIP(st->h_bases_dev.reserve(nb + 4));
        SCB_HIP(st->h_offsets_dev.reserve(r1 - r0 + 1));
        SCB_HIP(st->h_offs.reserve(r1 - r0 + 1));
        for (uint64_t i = r0; i <= r1; ++i) st->h_offs.p[i - r0] = offsets[i] - offsets[r0];
        SCB_HIP(hipMemcpy(st->h_bases_dev.p, bases + offsets[r0], nb, hipMemcpyHostToDevice));
        SCB_HIP(hipMemcpy(st->h_offsets_dev.p, st->h_offs.p, (r1 - r0 + 1) * 8, hipMemcpyHostToDevice));
        DeviceBatch b{st->h_bases_dev.p, st->h_offsets_dev.p, r1 - r0, nb};
        InsertResult part;
        Status ps = insert(b, method, part);
        if (!ps.good()) return ps;
        r.inserted_kmers += part.inserted_kmers;
        r.inserted_smers += part.inserted_smers;
        r.super_kmers += part.super_kmers;
        return Status::ok();
    });
    release(std::move(st));
    return s;
}

namespace {
// Runs fn(i0, i1) over [0, n) in up to 32 host threads (chunks of a multiple of
// 32 positions, so no two threads share a validity word).
template <class Fn>
void parallel_ranges(uint64_t n, Fn &&fn) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = n < (uint64_t(1) << 20) ? 1u : std::min(32u, hw);
    if (nt == 1) {
        fn(uint64_t(0), n);
        return;
    }
    const uint64_t step = ((n + nt - 1) / nt + 31) & ~uint64_t(31);
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < nt; ++t) {
        const uint64_t i0 = std::min(n, t * step), i1 = std::min(n, i0 + step);
        if (i0 < i1) ts.emplace_back([&fn, i0, i1] { fn(i0, i1); });
    }
    for (auto &t : ts) t.join();
}
}  // namespace

Status Engine::query_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n, Method method,
                          uint64_t *values, uint8_t *valid, QueryResult &r) const {
    r = QueryResult{};
    SCB_HIP(hipSetDevice(device_));
    auto st = acquire();  // staging buffers, kept with the pooled workspace
    if (!st) return Status{8, "device error: cannot create a workspace"};
    const uint64_t limit = std::min<uint64_t>(kMaxBatchBases, uint64_t(1) << 29);
    const uint32_t W = geo_.bytes;
    Status s = for_each_chunk(offsets, n, limit, [&](uint64_t r0, uint64_t r1) -> Status {
        const uint64_t nb = offsets[r1] - offsets[r0];
        SCB_HIP(st->h_bases_dev.reserve(nb + 4));
        SCB_HIP(st->h_offsets_dev.reserve(r1 - r0 + 1));
        SCB_HIP(st->h_offs.reserve(r1 - r0 + 1));
        const bool want = values || valid;
        if (want) {
            SCB_HIP(st->out_dev.reserve(nb * W + 4));
            SCB_HIP(st->out_bits.reserve((nb + 31) / 32 + 1));
            SCB_HIP(st->h_vals.reserve(nb * W + 4));
            SCB_HIP(st->h_bits.reserve((nb + 31) / 32 + 1));
        }
        for (uint64_t i = r0; i <= r1; ++i) st->h_offs.p[i - r0] = offsets[i] - offsets[r0];
        SCB_HIP(hipMemcpy(st->h_bases_dev.p, bases + offsets[r0], nb, hipMemcpyHostToDevice));
        SCB_HIP(hipMemcpy(st->h_offsets_dev.p, st->h_offs.p, (r1 - r0 + 1) * 8, hipMemcpyHostToDevice));
        DeviceBatch b{st->h_bases_dev.p, st->h_offsets_dev.p, r1 - r0, nb};
        QueryResult part;
        Status qs = query(b, method, want ? st->out_dev.p : nullptr, want ? st->out_bits.p : nullptr, part);
        if (!qs.good()) return qs;
        r.queried += part.queried;
        r.positive += part.positive;
        r.checksum += part.checksum;
        r.super_kmers += part.super_kmers;
        if (want) {
            SCB_HIP(hipMemcpy(st->h_vals.p, st->out_dev.p, nb * W, hipMemcpyDeviceToHost));
            SCB_HIP(hipMemcpy(st->h_bits.p, st->out_bits.p, (nb + 31) / 32 * 4, hipMemcpyDeviceToHost));
            const uint64_t o = offsets[r0];
            const uint8_t *hv = st->h_vals.p;
            const uint32_t *hb = st->h_bits.p;
            // position-indexed u64 values (0 where no valid window) and validity bytes
            parallel_ranges(nb, [&](uint64_t i0, uint64_t i1) {
                for (uint64_t i = i0; i < i1; ++i) {
                    const bool ok = (hb[i >> 5] >> (i & 31)) & 1;
                    if (valid) valid[o + i] = ok;
                    if (values) {
                        uint64_t v = 0;
                        if (ok) {
                            if (W == 1) v = hv[i];
                            else if (W == 2) v = reinterpret_cast<const uint16_t *>(hv)[i];
                            else v = reinterpret_cast<const uint32_t *>(hv)[i];
                        }
                        values[o + i] = v;
                    }
                }
            });
        }
        return Status::ok();
    });
    release(std::move(st));
    return s;
}

namespace {
// Host records staged chunk by chunk (chunks of at most `limit` bases, split at
// record boundaries) into the staging workspace; fn(DeviceBatch) runs per chunk.
template <class Fn>
Status for_each_staged_chunk(Workspace &st, const uint8_t *bases, const uint64_t *offsets, uint64_t n,
                             uint64_t limit, Fn &&fn) {
    return for_each_chunk(offsets, n, limit, [&](uint64_t r0, uint64_t r1) -> Status {
        const uint64_t nb = offsets[r1] - offsets[r0];
        SCB_HIP(st.h_bases_dev.reserve(nb + 4));
        SCB_HIP(st.h_offsets_dev.reserve(r1 - r0 + 1));
        SCB_HIP(st.h_offs.reserve(r1 - r0 + 1));
        for (uint64_t i = r0; i <= r1; ++i) st.h_offs.p[i - r0] = offsets[i] - offsets[r0];
        SCB_HIP(hipMemcpy(st.h_bases_dev.p, bases + offsets[r0], nb, hipMemcpyHostToDevice));
        SCB_HIP(hipMemcpy(st.h_offsets_dev.p, st.h_offs.p, (r1 - r0 + 1) * 8, hipMemcpyHostToDevice));
        return fn(DeviceBatch{st.h_bases_dev.p, st.h_offsets_dev.p, r1 - r0, nb});
    });
}
// Spectrum chunks of host input: the per-call pass holds about 8 bytes of
// records and 12 bytes of emitted keys per base.
constexpr uint64_t kSpectrumChunkBases = uint64_t(1) << 28;
}  // namespace

// One call when the input fits one device batch; larger inputs go through a
// key set chunk by chunk, so a key seen in an earlier chunk is not counted again.
Status Engine::spectrum_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n,
                             SpectrumResult &r, Method method) const {
    r = SpectrumResult{};
    SCB_HIP(hipSetDevice(device_));
    const uint64_t nb = offsets[n] - offsets[0];
    auto st = acquire();
    if (!st) return Status{8, "device error: cannot create a workspace"};
    Status s;
    if (nb < kMaxBatchBases) {
        s = for_each_staged_chunk(*st, bases, offsets, n, kMaxBatchBases,
                                  [&](const DeviceBatch &b) { return spectrum(b, r, method); });
    } else {
        KeySetPtr ks;
        s = keyset_create(ks);
        if (s.good())
            s = for_each_staged_chunk(*st, bases, offsets, n, kSpectrumChunkBases,
                                      [&](const DeviceBatch &b) { return spectrum_accumulate(b, *ks); });
        if (s.good()) s = keyset_rows(*ks, r);
    }
    release(std::move(st));
    return s;
}

Status Engine::spectrum_accumulate_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n,
                                        KeySet &ks) const {
    SCB_HIP(hipSetDevice(device_));
    auto st = acquire();
    if (!st) return Status{8, "device error: cannot create a workspace"};
    Status s = for_each_staged_chunk(*st, bases, offsets, n, kSpectrumChunkBases,
                                     [&](const DeviceBatch &b) { return spectrum_accumulate(b, ks); });
    release(std::move(st));
    return s;
}

// Pairs of every chunk, then one pass that keeps the first pair of each key
// (all pairs of a key carry the same estimate), sorted by key.
Status Engine::kmer_map_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n,
                             std::vector<std::pair<uint64_t, uint32_t>> &pairs, uint64_t &queried) const {
    pairs.clear();
    queried = 0;
    SCB_HIP(hipSetDevice(device_));
    auto st = acquire();
    if (!st) return Status{8, "device error: cannot create a workspace"};
    Status s = for_each_staged_chunk(*st, bases, offsets, n, kSpectrumChunkBases, [&](const DeviceBatch &b) {
        std::vector<std::pair<uint64_t, uint32_t>> part;
        uint64_t q = 0;
        Status ps = kmer_map(b, part, q);
        if (!ps.good()) return ps;
        queried += q;
        pairs.insert(pairs.end(), part.begin(), part.end());
        return Status::ok();
    });
    release(std::move(st));
    if (!s.good()) return s;
    std::stable_sort(pairs.begin(), pairs.end(),
                     [](const std::pair<uint64_t, uint32_t> &x, const std::pair<uint64_t, uint32_t> &y) {
                         return x.first < y.first;
                     });
    pairs.erase(std::unique(pairs.begin(), pairs.end(),
                            [](const std::pair<uint64_t, uint32_t> &x, const std::pair<uint64_t, uint32_t> &y) {
                                return x.first == y.first;
                            }),
                pairs.end());
    return Status::ok();
}

// ---- reservoir-sampled spectrum -------------------------------------------------
class Reservoir {
   public:
    uint64_t cap = 0, seed = 0, seen = 0;
    DevBuf<unsigned long long> slot_win;  // largest window number per slot
    DevBuf<uint32_t> slot_val;            // estimate of that window
    DevBuf<uint8_t> values;               // per-batch query outputs
    DevBuf<uint32_t> bits, rank, block_sums;
};

void ReservoirDeleter::operator()(Reservoir *r) const { delete r; }

Status Engine::reservoir_create(uint64_t cap, uint64_t seed, ReservoirPtr &out) const {
    if (cap == 0) return Status{1, "invalid config: sample_size must be positive"};
    SCB_HIP(hipSetDevice(device_));
    ReservoirPtr r(new Reservoir);
    r->cap = cap, r->seed = seed;
    SCB_HIP(r->slot_win.reserve(cap));
    SCB_HIP(r->slot_val.reserve(cap));
    fill(reinterpret_cast<uint32_t *>(r->slot_win.p), 2 * cap, 0, nullptr);
    fill(r->slot_val.p, cap, 0, nullptr);
    SCB_HIP(hipGetLastError());
    SCB_HIP(hipStreamSynchronize(nullptr));
    out = std::move(r);
    return Status::ok();
}

// Queries the batch (every window's estimate and validity), numbers its valid
// windows after the ones seen before, and updates the reservoir.
Status Engine::sample_accumulate(const DeviceBatch &b, Reservoir &rs) const {
    const uint64_t npos = b.n_bases;
    if (npos == 0 || b.n_reads == 0) return Status::ok();
    SCB_HIP(hipSetDevice(device_));
    const uint64_t nwords = (npos + 31) / 32;
    const uint64_t nblocks = (nwords + kRankThreads - 1) / kRankThreads;
    SCB_HIP(rs.values.reserve(npos * geo_.bytes + 4));
    SCB_HIP(rs.bits.reserve(nwords + 1));
    SCB_HIP(rs.rank.reserve(nwords + 1));
    SCB_HIP(rs.block_sums.reserve(nblocks + 1));
    QueryResult qr;
    Status s = query(b, kBatched, rs.values.p, rs.bits.p, qr);
    if (!s.good()) return s;
    hipStream_t st = nullptr;
    hipLaunchKernelGGL(k_rank_sums, dim3(unsigned(nblocks)), dim3(kRankThreads), 0, st, rs.bits.p, nwords,
                       rs.block_sums.p);
    hipLaunchKernelGGL(k_rank_scan_blocks, dim3(1), dim3(kRankThreads), 0, st, rs.block_sums.p, nblocks);
    hipLaunchKernelGGL(k_rank_words, dim3(unsigned(nblocks)), dim3(kRankThreads), 0, st, rs.bits.p, nwords,
                       rs.block_sums.p, rs.rank.p);
    const unsigned grid = unsigned(std::min<uint64_t>((nwords + 255) / 256, 65536));
    hipLaunchKernelGGL(k_reservoir_hits, dim3(grid), dim3(256), 0, st, rs.bits.p, nwords, rs.rank.p, rs.seen, rs.cap,
                       rs.seed, rs.slot_win.p);
    switch (geo_.bits) {
        case 8:
            hipLaunchKernelGGL(k_reservoir_take<8>, dim3(grid), dim3(256), 0, st, rs.bits.p, nwords, rs.rank.p,
                               rs.seen, rs.cap, rs.seed, rs.slot_win.p, rs.values.p, rs.slot_val.p);
            break;
        case 16:
            hipLaunchKernelGGL(k_reservoir_take<16>, dim3(grid), dim3(256), 0, st, rs.bits.p, nwords, rs.rank.p,
                               rs.seen, rs.cap, rs.seed, rs.slot_win.p, rs.values.p, rs.slot_val.p);
            break;
        default:
            hipLaunchKernelGGL(k_reservoir_take<32>, dim3(grid), dim3(256), 0, st, rs.bits.p, nwords, rs.rank.p,
                               rs.seen, rs.cap, rs.seed, rs.slot_win.p, rs.values.p, rs.slot_val.p);
            break;
    }
    SCB_HIP(hipGetLastError());
    SCB_HIP(hipStreamSynchronize(st));
    rs.seen += qr.queried;
    return Status::ok();
}

Status Engine::sample_accumulate_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n,
                                      Reservoir &rs) const {
    SCB_HIP(hipSetDevice(device_));
    auto st = acquire();
    if (!st) return Status{8, "device error: cannot create a workspace"};
    Status s = for_each_staged_chunk(*st, bases, offsets, n, kSpectrumChunkBases,
                                     [&](const DeviceBatch &b) { return sample_accumulate(b, rs); });
    release(std::move(st));
    return s;
}

Status Engine::reservoir_values(const Reservoir &rs, std::vector<uint32_t> &vals, uint64_t &seen) const {
    SCB_HIP(hipSetDevice(device_));
    const uint64_t n = std::min(rs.seen, rs.cap);
    vals.resize(n);
    if (n) SCB_HIP(hipMemcpy(vals.data(), rs.slot_val.p, n * 4, hipMemcpyDeviceToHost));
    seen = rs.seen;
    return Status::ok();
}

StageTimes last_stage_times() { return g_stage_times; }

namespace {
constexpr uint32_t kStatChunk = 4096;  // positions per workgroup

// filter.rs:552-556 per segment: count, positive count and wrapping sum of
// the valid windows' estimates.
template <int BITS>
__global__ __launch_bounds__(256) void k_segment_stats(const void *values, const uint32_t *valid_bits,
                                                       const uint64_t *segs, uint32_t nseg, uint64_t npos,
                                                       unsigned long long *out) {
    using T = typename Ctr<BITS>::T;
    __shared__ uint64_t s_red[3][4];
    const uint64_t c0 = uint64_t(blockIdx.x) * kStatChunk, c1 = min(npos, c0 + kStatChunk);
    // first segment ending after c0
    uint32_t lo = 0, hi = nseg;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (segs[2 * mid + 1] <= c0) lo = mid + 1; else hi = mid;
    }
    for (uint32_t sg = lo; sg < nseg && segs[2 * sg] < c1; ++sg) {
        const uint64_t b = max(c0, segs[2 * sg]), e = min(c1, segs[2 * sg + 1]);
        uint64_t n = 0, pos = 0, sum = 0;
        for (uint64_t p = b + threadIdx.x; p < e; p += blockDim.x) {
            if ((valid_bits[p >> 5] >> (p & 31)) & 1) {
                uint64_t v = reinterpret_cast<const T *>(values)[p];
                ++n;
                pos += v > 0;
                sum += v;
            }
        }
        uint64_t vals[3] = {n, pos, sum};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            uint64_t v = vals[q];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
            if ((threadIdx.x & 63) == 0) s_red[q][threadIdx.x >> 6] = v;
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            uint64_t t = 0;
            for (uint32_t i = 0; i < blockDim.x / 64; ++i) t += s_red[threadIdx.x][i];
            if (t) atomicAdd(&out[3 * sg + threadIdx.x], (unsigned long long)t);
        }
        __syncthreads();
    }
}
}  // namespace

Status fill_u32_async(uint32_t *d_p, uint64_t n, uint32_t value, void *stream) {
    fill(d_p, n, value, reinterpret_cast<hipStream_t>(stream));
    SCB_HIP(hipGetLastError());
    return Status::ok();
}

Status segment_stats_async(uint32_t counter_bits, const void *d_values, const uint32_t *d_valid_bits,
                           const uint64_t *d_segs, uint32_t nseg, uint64_t n_positions, uint64_t *d_out,
                           void *stream) {
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    fill(reinterpret_cast<uint32_t *>(d_out), 6 * uint64_t(nseg), 0, st);
    if (!nseg || !n_positions) return Status::ok();
    dim3 grid(unsigned((n_positions + kStatChunk - 1) / kStatChunk)), blk(256);
    auto *o = reinterpret_cast<unsigned long long *>(d_out);
    switch (counter_bits) {
        case 8: hipLaunchKernelGGL(k_segment_stats<8>, grid, blk, 0, st, d_values, d_valid_bits, d_segs, nseg, n_positions, o); break;
        case 16: hipLaunchKernelGGL(k_segment_stats<16>, grid, blk, 0, st, d_values, d_valid_bits, d_segs, nseg, n_positions, o); break;
        default: hipLaunchKernelGGL(k_segment_stats<32>, grid, blk, 0, st, d_values, d_valid_bits, d_segs, nseg, n_positions, o); break;
    }
    SCB_HIP(hipGetLastError());
    return Status::ok();
}

Status make_geometry(uint64_t k, uint64_t m, uint64_t s, uint64_t nh, uint32_t bits, uint32_t cse,
                     uint32_t bse, Geometry &g) {
    g = Geometry{};
    g.k = uint32_t(k), g.m = uint32_t(m), g.s = uint32_t(s), g.nh = uint32_t(nh);
    g.w = g.k - g.m + 1;
    g.win = g.k - g.s + 1;
    g.bits = bits;
    g.bytes = bits / 8;
    g.cse = cse, g.bse = bse;
    g.slots = uint64_t(1) << cse;
    g.nb = g.slots >> bse;
    g.block_mask = g.nb - 1;
    g.shard_count = std::min<uint64_t>(g.nb, 1024);
    g.subblocks = g.nb / g.shard_count;
    g.shard_mask = uint32_t(g.shard_count - 1);
    g.shard_shift = uint32_t(__builtin_ctzll(g.shard_count));
    g.sub_log = uint32_t(__builtin_ctzll(g.subblocks));
    return Status::ok();
}

}  // namespace scb
