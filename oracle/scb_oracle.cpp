// CPU oracle for the SuperCountingBloom hot path: TEST INFRASTRUCTURE ONLY.
//
// A restatement, written from the reference's source text, of the CPU algorithm
// of super-counting-bloom 0.1.0 (Malfoy/SCBF). Every function cites the
// reference file:line it follows. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use this library, as the checker or as the timed
// CPU baseline; the product library never links it.
//
// Parity status (what pins each part; the tests named here live in tests/):
//   xorshift_u64 ............. pinned: SplitMix64 published output sequence for
//                              seed 0 (test_oracle_known_answers.py)
//   config validate .......... pinned: reference messages + rejection matrix
//                              (tests/pipeline_matrix.rs:187-262, config.rs:50-107)
//   layout / P(b) / .scb ..... pinned: closed form of filter.rs:90-116, api.rs:388-608;
//                              round trip + rejection tests (api.rs:1129-1156, 1215-1225)
//   run splitting, counts .... pinned: api.rs:1058-1072, 1158-1202; fastx.rs:448-466;
//                              pipeline_matrix.rs:61-82 (exact counts re-run here)
//   insert / query / stream /  pinned given the super-k-mer list: brute-force
//   map / spectrum ........... definition check (oracle_filter_estimate_ascii_brute),
//                              no-underestimation, stream == vector, repeats monotone,
//                              saturation (filter.rs:680-756, backend_matrix.rs)
//   canonical_kmer (F1/F2) ... PARITY UNPINNED at the encoding: packed-seq 4.4.2 is
//                              absent; rc-symmetry and filter.rs:748-756 checked only
//   nthash / minimizer scan /  PARITY UNPINNED: simd-minimizers 2.3.1 + seq-hash 0.1.2
//   super_kmers (F3-F8) ...... are absent; checked against a brute-force scan and the
//                              invariants that hold for any F3-F6 (partition,
//                              minimizer-in-window, rc symmetry, density)
//   f64 estimators ........... pinned by api.rs:1113-1127 and hand-computed closed
//                              forms (test_oracle_estimators.py)
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off; Rust never fuses mul-add).

#include "scb_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

// ============================================================================
// Third-party restatement (UNPINNED). packed-seq 4.4.2, seq-hash 0.1.2 and
// simd-minimizers 2.3.1 are not in /root/reference; their published behaviour
// is restated here and nowhere else. Assumptions F1..F8:
//   F1 base code = (ascii >> 1) & 3: A=0 C=1 T=2 G=3; complement = code ^ 2
//   F2 read_kmer(len,pos): base pos+j at bits [2j,2j+1]; read_revcomp_kmer:
//      bits 2j hold the complement of base pos+len-1-j
//   F3 canonical ntHash, 32-bit: seeds = low 32 bits of the original ntHash
//      seeds, rotate-by-1 rolling, canonical = fw + rc (wrapping)
//   F4 canonical_minimizers(m, w = k-m+1): window = w m-mers = one k-mer
//   F5 argmin key = high 16 bits of the hash; ties broken by position
//   F6 strand rule: #{T,G} in the k-bp window > k/2 -> leftmost tie, else rightmost
//   F7 super-k-mer starts = first window of each run of equal minimizer
//      positions (pinned by its use in filter.rs:210-217)
//   F8 values_u64 = min(read_kmer(m,pos), read_revcomp_kmer(m,pos))
// ============================================================================

inline uint8_t base_code(uint8_t c) { return (c >> 1) & 3; }  // F1

// ntHash seeds (Mohamadi et al. 2016), low 32 bits, indexed by the F1 code.
constexpr uint32_t kNtSeed[4] = {
    0x95c60474u,  // A: 0x3c8bfbb395c60474
    0x62a02b4cu,  // C: 0x3193c18562a02b4c
    0x4be24456u,  // T: 0x295549f54be24456
    0x82572324u,  // G: 0x20323ed082572324
};

inline uint32_t rotl32(uint32_t x, unsigned r) {
    r &= 31;
    return r ? (x << r) | (x >> (32 - r)) : x;
}
inline uint32_t rotr32(uint32_t x, unsigned r) {
    r &= 31;
    return r ? (x >> r) | (x << (32 - r)) : x;
}

// 2-bit packed sequence, 32 bases per u64 word, first base least significant (F2).
struct Packed {
    std::vector<uint64_t> words;
    size_t len = 0;
    Packed() = default;
    Packed(const uint8_t *ascii, size_t n) { assign(ascii, n); }
    void assign(const uint8_t *ascii, size_t n) {
        len = n;
        words.assign(n / 32 + 2, 0);
        for (size_t i = 0; i < n; ++i)
            words[i >> 5] |= uint64_t(base_code(ascii[i])) << (2 * (i & 31));
    }
    uint8_t code(size_t i) const { return (words[i >> 5] >> (2 * (i & 31))) & 3; }
    // packed_seq::PackedSeq::read_kmer (F2)
    uint64_t read_kmer(size_t klen, size_t pos) const {
        size_t w = pos >> 5, off = 2 * (pos & 31);
        uint64_t x = words[w] >> off;
        if (off) x |= words[w + 1] << (64 - off);
        return klen == 32 ? x : (x & ((uint64_t(1) << (2 * klen)) - 1));
    }
    // packed_seq::PackedSeq::read_revcomp_kmer (F2)
    uint64_t read_revcomp_kmer(size_t klen, size_t pos) const {
        uint64_t x = read_kmer(klen, pos);
        x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
        x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
        x = __builtin_bswap64(x);
        x ^= 0xAAAAAAAAAAAAAAAAull;
        return x >> (64 - 2 * klen);
    }
};

// filter.rs:622-627
inline uint64_t canonical_kmer(const Packed &seq, size_t len, size_t pos) {
    return std::min(seq.read_kmer(len, pos), seq.read_revcomp_kmer(len, pos));
}

// filter.rs:629-635 (the SplitMix64 output function)
inline uint64_t xorshift_u64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// F3, rolling form.
void nthash_canonical(const Packed &seq, size_t m, std::vector<uint32_t> &out) {
    size_t L = seq.len;
    out.clear();
    if (L < m) return;
    out.resize(L - m + 1);
    uint32_t fw = 0, rc = 0;
    for (size_t i = 0; i < m; ++i) {
        uint8_t c = seq.code(i);
        fw ^= rotl32(kNtSeed[c], unsigned(m - 1 - i));
        rc ^= rotl32(kNtSeed[c ^ 2], unsigned(i));
    }
    out[0] = fw + rc;
    for (size_t p = 0; p + m < L; ++p) {
        uint8_t co = seq.code(p), ci = seq.code(p + m);
        fw = rotl32(fw, 1) ^ rotl32(kNtSeed[co], unsigned(m)) ^ kNtSeed[ci];
        rc = rotr32(rc, 1) ^ rotr32(kNtSeed[co ^ 2], 1) ^ rotl32(kNtSeed[ci ^ 2], unsigned(m - 1));
        out[p + 1] = fw + rc;
    }
}

// F3, direct definition (used to check the rolling form).
void nthash_canonical_direct(const Packed &seq, size_t m, std::vector<uint32_t> &out) {
    size_t L = seq.len;
    out.clear();
    if (L < m) return;
    out.resize(L - m + 1);
    for (size_t p = 0; p + m <= L; ++p) {
        uint32_t fw = 0, rc = 0;
        for (size_t i = 0; i < m; ++i) {
            uint8_t c = seq.code(p + i);
            fw ^= rotl32(kNtSeed[c], unsigned(m - 1 - i));
            rc ^= rotl32(kNtSeed[c ^ 2], unsigned(i));
        }
        out[p] = fw + rc;
    }
}

// F4-F6: minimizer position of every k-mer window, O(L) with two monotone
// deques (one keeps the leftmost minimum, the other the rightmost).
void minimizer_positions(const Packed &seq, size_t k, size_t m, std::vector<uint32_t> &pos_out,
                         std::vector<uint32_t> &hash_buf) {
    size_t L = seq.len;
    pos_out.clear();
    if (L < k) return;
    size_t w = k - m + 1, nwin = L - k + 1;
    nthash_canonical(seq, m, hash_buf);
    pos_out.resize(nwin);
    std::vector<uint32_t> dl(hash_buf.size()), dr(hash_buf.size());
    size_t lh = 0, lt = 0, rh = 0, rt = 0;
    auto key = [&](size_t p) { return hash_buf[p] >> 16; };  // F5
    size_t tg = 0;
    for (size_t i = 0; i < k - 1; ++i) tg += seq.code(i) >> 1;
    for (size_t p = 0; p < hash_buf.size(); ++p) {
        uint32_t kp = key(p);
        while (lt > lh && key(dl[lt - 1]) > kp) --lt;   // leftmost among equals survives
        dl[lt++] = uint32_t(p);
        while (rt > rh && key(dr[rt - 1]) >= kp) --rt;  // rightmost among equals survives
        dr[rt++] = uint32_t(p);
        if (p + 1 >= w) {
            size_t j = p + 1 - w;  // window index
            while (dl[lh] < j) ++lh;
            while (dr[rh] < j) ++rh;
            tg += seq.code(j + k - 1) >> 1;
            bool left = 2 * tg > k;  // F6: #TG > k/2 (k odd)
            pos_out[j] = left ? dl[lh] : dr[rh];
            tg -= seq.code(j) >> 1;
        }
    }
}

// F4-F6, brute force over each window.
void minimizer_positions_brute(const Packed &seq, size_t k, size_t m, std::vector<uint32_t> &pos_out) {
    size_t L = seq.len;
    pos_out.clear();
    if (L < k) return;
    size_t w = k - m + 1;
    std::vector<uint32_t> h;
    nthash_canonical_direct(seq, m, h);
    for (size_t j = 0; j + k <= L; ++j) {
        size_t tg = 0;
        for (size_t i = 0; i < k; ++i) tg += seq.code(j + i) >> 1;
        uint32_t best = 0xFFFFFFFFu;
        size_t lpos = 0, rpos = 0;
        for (size_t p = j; p < j + w; ++p) {
            uint32_t kp = h[p] >> 16;
            if (kp < best) {
                best = kp;
                lpos = rpos = p;
            } else if (kp == best) {
                rpos = p;
            }
        }
        pos_out.push_back(uint32_t(2 * tg > k ? lpos : rpos));
    }
}

// filter.rs:608-620 -> simd_minimizers::canonical_minimizers(m, k-m+1)
//   .super_kmers(&mut starts).run(seq, &mut pos).values_u64()  (F7, F8)
struct ScanBuf {
    std::vector<uint32_t> pos, hash;
};
void super_kmers(const Packed &seq, size_t k, size_t m, std::vector<uint32_t> &starts,
                 std::vector<uint64_t> &mins, ScanBuf &buf) {
    starts.clear();
    mins.clear();
    minimizer_positions(seq, k, m, buf.pos, buf.hash);
    for (size_t j = 0; j < buf.pos.size(); ++j) {
        if (j == 0 || buf.pos[j] != buf.pos[j - 1]) {
            starts.push_back(uint32_t(j));
            mins.push_back(canonical_kmer(seq, m, buf.pos[j]));
        }
    }
}

// ============================================================================
// Config (config.rs:3-107)
// ============================================================================
struct Config {
    size_t k, m, s, n_hashes;
    unsigned bits;
    unsigned cse, bse;
    size_t threads, queue;
};

Config from_c(const oracle_config *c) {
    return Config{size_t(c->k), size_t(c->m), size_t(c->s), size_t(c->n_hashes), c->counter_bits,
                  c->counter_slots_exponent, c->block_slots_exponent, size_t(c->threads),
                  size_t(c->queue)};
}

// config.rs:50-92 then checked_pow2 (103-107) as used by FrozenFilterLayout::from_config
bool validate(const Config &c, std::string &msg) {
    if (c.k == 0 || c.k > 32) return msg = "k must be in 1..=32", false;
    if (c.m == 0 || c.m > c.k) return msg = "m must be in 1..=k", false;
    if (c.s == 0 || c.s > c.k || c.s > 32) return msg = "s must be in 1..=k and <= 32", false;
    if (c.k % 2 == 0) return msg = "canonical SIMD minimizers require odd k", false;
    if (c.n_hashes == 0) return msg = "n_hashes must be positive", false;
    if (c.bse >= c.cse)
        return msg = "block_slots_exponent must be smaller than counter_slots_exponent", false;
    if (c.threads == 0) return msg = "threads must be positive", false;
    if (c.queue == 0) return msg = "queue must be positive", false;
    return true;
}
bool checked_layout(const Config &c, std::string &msg) {
    if (c.cse >= 64) return msg = "counter_slots_exponent is too large", false;
    if (c.bse >= 64) return msg = "block_slots_exponent is too large", false;
    return true;
}

// FrozenFilterLayout (filter.rs:77-116)
struct Layout {
    uint64_t shard_count, subblocks, block_slots, shard_mask;
    uint32_t shard_shift;
    uint64_t block_slots_mask, nb_blocks, block_mask, n_hashes;
    bool operator==(const Layout &o) const {
        return shard_count == o.shard_count && subblocks == o.subblocks &&
               block_slots == o.block_slots && shard_mask == o.shard_mask &&
               shard_shift == o.shard_shift && block_slots_mask == o.block_slots_mask &&
               nb_blocks == o.nb_blocks && block_mask == o.block_mask && n_hashes == o.n_hashes;
    }
};
Layout layout_from_config(const Config &c) {
    Layout l;
    uint64_t counter_slots = uint64_t(1) << c.cse, block_slots = uint64_t(1) << c.bse;
    l.nb_blocks = counter_slots / block_slots;
    l.shard_count = std::min<uint64_t>(l.nb_blocks, 1024);
    l.subblocks = l.nb_blocks / l.shard_count;
    l.block_slots = block_slots;
    l.shard_mask = l.shard_count - 1;
    l.shard_shift = uint32_t(__builtin_ctzll(l.shard_count));
    l.block_slots_mask = block_slots - 1;
    l.block_mask = l.nb_blocks - 1;
    l.n_hashes = c.n_hashes;
    return l;
}

bool is_acgt(uint8_t b) {  // api.rs:929-931
    uint8_t u = (b >= 'a' && b <= 'z') ? uint8_t(b - 32) : b;
    return u == 'A' || u == 'C' || u == 'G' || u == 'T';
}

// ============================================================================
// Filter (filter.rs:17-606), counters held in physical (shard-major) order:
// index = (shard * subblocks + subblock) * block_slots + address.
// ============================================================================
struct FilterBase {
    Config cfg;
    Layout lay;
    virtual ~FilterBase() = default;
    virtual uint64_t insert_packed(const Packed &seq) = 0;
    virtual void estimate_abundances(const Packed &seq, std::vector<uint64_t> &out) const = 0;
    virtual void stream(const Packed &seq, uint64_t &queried, uint64_t &checksum) const = 0;
    // (canonical k-mer, estimate) of every k-mer of seq in position order
    virtual uint64_t estimate_pairs(const Packed &seq,
                                    std::vector<std::pair<uint64_t, uint64_t>> &out) const = 0;
    virtual uint64_t get(uint64_t block, uint64_t addr) const = 0;
    virtual uint64_t max_value() const = 0;
    virtual void to_u32(uint32_t *out) const = 0;
    virtual void from_u32(const uint32_t *in) = 0;
    virtual size_t slots() const = 0;
    virtual void histogram(std::unordered_map<uint64_t, uint64_t> &h) const = 0;
    virtual const void *data(size_t &bytes) const = 0;
    virtual void write_le(FILE *fp) const = 0;
    virtual bool read_le(FILE *fp) = 0;
    uint64_t phys_block(uint64_t block) const {  // filter.rs:217-219 -> .scb order
        uint64_t shard = block & lay.shard_mask, sub = block >> lay.shard_shift;
        return shard * lay.subblocks + sub;
    }
};

template <class C>
struct Filter final : FilterBase {
    std::vector<C> counters;
    std::unique_ptr<std::mutex[]> locks;  // one per shard (filter.rs:58, 130-134)

    Filter(const Config &c) {
        cfg = c;
        lay = layout_from_config(c);
        counters.assign(size_t(lay.nb_blocks * lay.block_slots), C(0));
        locks.reset(new std::mutex[lay.shard_count]);
    }
    static constexpr C kMax = C(~C(0));
    uint64_t max_value() const override { return kMax; }
    size_t slots() const override { return counters.size(); }
    uint64_t get(uint64_t block, uint64_t addr) const override {
        return counters[size_t(phys_block(block) * lay.block_slots + addr)];
    }

    // filter.rs:199-237
    uint64_t insert_packed(const Packed &seq) override {
        if (seq.len < cfg.k) return 0;
        thread_local std::vector<uint32_t> starts;
        thread_local std::vector<uint64_t> mins;
        thread_local ScanBuf buf;
        uint64_t inserted_smers = 0;
        super_kmers(seq, cfg.k, cfg.m, starts, mins, buf);
        for (size_t i = 0; i < starts.size(); ++i) {
            size_t start_kmer = starts[i];
            size_t end_kmer = i + 1 < starts.size() ? starts[i + 1] : seq.len + 1 - cfg.k;
            uint64_t block = xorshift_u64(mins[i]) & lay.block_mask;
            uint64_t shard = block & lay.shard_mask;
            std::lock_guard<std::mutex> guard(locks[shard]);
            C *blk = &counters[size_t(phys_block(block) * lay.block_slots)];
            size_t end_smer = end_kmer + (cfg.k - cfg.s);
            for (size_t pos = start_kmer; pos < end_smer; ++pos) {
                uint64_t smer = canonical_kmer(seq, cfg.s, pos);
                uint64_t hash = xorshift_u64(smer);
                for (size_t t = 0; t < cfg.n_hashes; ++t) {
                    C &ctr = blk[hash & lay.block_slots_mask];
                    if (ctr != kMax) ++ctr;  // counter.rs:20-23 saturating_increment
                    hash = xorshift_u64(hash);
                }
                ++inserted_smers;
            }
        }
        return inserted_smers;
    }

    // filter.rs:562-571
    C estimate_smer(const C *blk, uint64_t smer) const {
        uint64_t hash = xorshift_u64(smer);
        C est = kMax;
        for (size_t t = 0; t < cfg.n_hashes; ++t) {
            est = std::min(est, blk[hash & lay.block_slots_mask]);
            hash = xorshift_u64(hash);
        }
        return est;
    }

    // Shared skeleton of filter.rs:407-460, 462-508, 510-560: s-mer estimates of
    // one super-k-mer then the monotone-deque sliding minimum of width k-s+1;
    // emit(kmer_pos, estimate) is called in position order.
    template <class Emit>
    void super_kmer_window_min(const C *blk, const Packed &seq, size_t start_kmer, size_t end_kmer,
                               Emit &&emit) const {
        size_t window = cfg.k - cfg.s + 1;
        size_t kmer_count = end_kmer - start_kmer;
        size_t smer_count = kmer_count + window - 1;
        thread_local std::vector<C> est;
        est.clear();
        for (size_t off = 0; off < smer_count; ++off)
            est.push_back(estimate_smer(blk, canonical_kmer(seq, cfg.s, start_kmer + off)));
        thread_local std::deque<size_t> dq;
        dq.clear();
        for (size_t off = 0; off < est.size(); ++off) {
            while (!dq.empty() && est[dq.back()] >= est[off]) dq.pop_back();
            dq.push_back(off);
            while (!dq.empty() && dq.front() + window <= off) dq.pop_front();
            if (off + 1 >= window) emit(start_kmer + off + 1 - window, est[dq.front()]);
        }
    }

    template <class PerSuper>
    void for_each_super_kmer(const Packed &seq, PerSuper &&per) const {
        thread_local std::vector<uint32_t> starts;
        thread_local std::vector<uint64_t> mins;
        thread_local ScanBuf buf;
        super_kmers(seq, cfg.k, cfg.m, starts, mins, buf);
        for (size_t i = 0; i < starts.size(); ++i) {
            size_t start_kmer = starts[i];
            size_t end_kmer = i + 1 < starts.size() ? starts[i + 1] : seq.len + 1 - cfg.k;
            uint64_t block = xorshift_u64(mins[i]) & lay.block_mask;
            const C *blk = &counters[size_t(phys_block(block) * lay.block_slots)];
            per(blk, start_kmer, end_kmer);
        }
    }

    // filter.rs:337-369
    void estimate_abundances(const Packed &seq, std::vector<uint64_t> &out) const override {
        out.clear();
        if (seq.len < cfg.k) return;
        out.reserve(seq.len + 1 - cfg.k);
        for_each_super_kmer(seq, [&](const C *blk, size_t a, size_t b) {
            super_kmer_window_min(blk, seq, a, b, [&](size_t, C e) { out.push_back(e); });
        });
    }

    // filter.rs:371-405
    void stream(const Packed &seq, uint64_t &queried, uint64_t &checksum) const override {
        queried = 0;
        checksum = 0;
        if (seq.len < cfg.k) return;
        for_each_super_kmer(seq, [&](const C *blk, size_t a, size_t b) {
            uint64_t lq = 0, lc = 0;
            super_kmer_window_min(blk, seq, a, b, [&](size_t, C e) {
                lc += uint64_t(e);
                ++lq;
            });
            queried += lq;
            checksum += lc;
        });
    }

    // filter.rs:302-335 + 407-460: the (canonical k-mer, estimate) pairs that
    // estimate_super_kmer_in_shard offers to the map in position order; the
    // caller inserts them first-wins (filter.rs:454)
    uint64_t estimate_pairs(const Packed &seq,
                            std::vector<std::pair<uint64_t, uint64_t>> &out) const override {
        if (seq.len < cfg.k) return 0;
        uint64_t queried = 0;
        for_each_super_kmer(seq, [&](const C *blk, size_t a, size_t b) {
            super_kmer_window_min(blk, seq, a, b, [&](size_t kpos, C e) {
                out.emplace_back(canonical_kmer(seq, cfg.k, kpos), uint64_t(e));
                ++queried;
            });
        });
        return queried;
    }

    void to_u32(uint32_t *out) const override {
        for (size_t i = 0; i < counters.size(); ++i) out[i] = counters[i];
    }
    void from_u32(const uint32_t *in) override {
        for (size_t i = 0; i < counters.size(); ++i) counters[i] = C(in[i]);
    }
    // filter.rs:50-55: every counter value is counted. Values below 2^16 are
    // counted densely, by worker threads over slices of the table; the rest
    // (32-bit counters only) go to a map.
    void histogram(std::unordered_map<uint64_t, uint64_t> &h) const override {
        const size_t n = counters.size();
        const size_t nt = std::max<size_t>(1, std::min<size_t>(16, n >> 24));
        std::vector<std::vector<uint64_t>> dense(nt, std::vector<uint64_t>(65536, 0));
        std::vector<std::unordered_map<uint64_t, uint64_t>> wide(nt);
        std::vector<std::thread> pool;
        for (size_t t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                const size_t a = n * t / nt, b = n * (t + 1) / nt;
                for (size_t i = a; i < b; ++i) {
                    const uint64_t c = counters[i];
                    if (c < 65536) ++dense[t][c];
                    else ++wide[t][c];
                }
            });
        for (auto &th : pool) th.join();
        for (size_t t = 0; t < nt; ++t) {
            for (size_t v = 0; v < 65536; ++v)
                if (dense[t][v]) h[v] += dense[t][v];
            for (auto &kv : wide[t]) h[kv.first] += kv.second;
        }
    }
    const void *data(size_t &bytes) const override {
        bytes = counters.size() * sizeof(C);
        return counters.data();
    }
    // counter.rs:31-135: little-endian, shard-major (filter.rs:575-580)
    void write_le(FILE *fp) const override {
        fwrite(counters.data(), sizeof(C), counters.size(), fp);  // host is little-endian
    }
    bool read_le(FILE *fp) override {
        return fread(counters.data(), sizeof(C), counters.size(), fp) == counters.size();
    }
};

FilterBase *make_filter(const Config &c) {
    switch (c.bits) {
        case 8: return new Filter<uint8_t>(c);
        case 16: return new Filter<uint16_t>(c);
        default: return new Filter<uint32_t>(c);
    }
}

// api.rs:676-726, with a callback per maximal ACGT run of length >= k
template <class Fn>
void for_each_valid_run(const uint8_t *seq, size_t len, size_t k, Fn &&fn) {
    if (len < k) return;
    bool in_run = false;
    size_t start = 0;
    for (size_t i = 0; i < len; ++i) {
        if (is_acgt(seq[i])) {
            if (!in_run) in_run = true, start = i;
        } else if (in_run) {
            in_run = false;
            if (i - start >= k) fn(start, i);
        }
    }
    if (in_run && len - start >= k) fn(start, len);
}

Packed pack_run(const uint8_t *seq, size_t a, size_t b) {
    // api.rs:718-722: uppercase then PackedSeqVec::from_ascii. F1's code is case
    // insensitive, so packing the raw bytes gives the same words.
    return Packed(seq + a, b - a);
}

// The distinct-k-mer map of fastx.rs:255-308, split by a hash of the key into
// kParts maps so that merges run in parallel; each key lives in exactly one of
// them, so first-wins per map is first-wins over the whole key set.
struct Spectrum {
    static constexpr int kParts = 64;
    std::unordered_map<uint64_t, uint64_t> map[kParts];
    static int part(uint64_t key) { return int((key * 0x9E3779B97F4A7C15ull) >> 58); }
    size_t size() const {
        size_t n = 0;
        for (auto &m : map) n += m.size();
        return n;
    }
};

// Bounded MPMC channel of owned records (crossbeam_channel::bounded semantics).
class Channel {
   public:
    explicit Channel(size_t cap) : cap_(cap) {}
    void send(std::vector<uint8_t> &&v) {
        std::unique_lock<std::mutex> lk(mu_);
        not_full_.wait(lk, [&] { return q_.size() < cap_; });
        q_.push_back(std::move(v));
        not_empty_.notify_one();
    }
    bool recv(std::vector<uint8_t> &out) {
        std::unique_lock<std::mutex> lk(mu_);
        not_empty_.wait(lk, [&] { return !q_.empty() || closed_; });
        if (q_.empty()) return false;
        out = std::move(q_.front());
        q_.pop_front();
        not_full_.notify_one();
        return true;
    }
    void close() {
        std::lock_guard<std::mutex> lk(mu_);
        closed_ = true;
        not_empty_.notify_all();
    }

   private:
    size_t cap_;
    std::deque<std::vector<uint8_t>> q_;
    std::mutex mu_;
    std::condition_variable not_full_, not_empty_;
    bool closed_ = false;
};

void copy_msg(char *dst, size_t cap, const std::string &s) {
    if (!dst || !cap) return;
    size_t n = std::min(cap - 1, s.size());
    memcpy(dst, s.data(), n);
    dst[n] = 0;
}

// Rust's `f64 as u64`: saturating, NaN -> 0.
uint64_t f64_as_u64(double x) {
    if (std::isnan(x) || x <= 0.0) return 0;
    if (x >= 18446744073709551616.0) return UINT64_MAX;
    return uint64_t(x);
}

// api.rs:958-1011
std::vector<double> deconvolved_counts(const uint64_t *values, const uint64_t *counts, size_t rows,
                                       uint64_t n_hashes, uint64_t max_abundance_u) {
    size_t max_abundance = size_t(max_abundance_u);
    if (max_abundance == 0) return {};
    uint64_t total_counters = 0;
    for (size_t i = 0; i < rows; ++i) total_counters += counts[i];
    if (total_counters == 0) return std::vector<double>(max_abundance, 0.0);
    std::vector<double> probabilities(max_abundance + 1, 0.0);
    for (size_t i = 0; i < rows; ++i) {
        if (values[i] <= max_abundance)
            probabilities[size_t(values[i])] += double(counts[i]) / double(total_counters);
    }
    double p0 = probabilities[0];
    if (p0 <= 0.0) return {};
    std::vector<double> normalized(max_abundance + 1, 0.0);
    normalized[0] = 1.0;
    for (size_t a = 1; a <= max_abundance; ++a) normalized[a] = probabilities[a] / p0;
    std::vector<double> log_coefficients(max_abundance + 1, 0.0);
    for (size_t a = 1; a <= max_abundance; ++a) {
        double coefficient = double(a) * normalized[a];
        for (size_t prev = 1; prev < a; ++prev)
            coefficient -= double(prev) * log_coefficients[prev] * normalized[a - prev];
        log_coefficients[a] = coefficient / double(a);
    }
    double hash_count = double(std::max<uint64_t>(n_hashes, 1));
    std::vector<double> out;
    out.reserve(max_abundance);
    for (size_t a = 1; a <= max_abundance; ++a)
        out.push_back(std::fmax(double(total_counters) * log_coefficients[a] / hash_count, 0.0));
    return out;
}

// Record-parallel forms of the per-record calls above, for inputs of millions
// of records: records [0, n) are dealt to `threads` workers in contiguous
// ranges, as add_fasta_typed's and query_fasta's worker pools share records
// (api.rs:278-313, 621-666). Inserts lock the shard of every super-k-mer
// (filter.rs:220-222) and saturating increments commute, so the table equals
// the one-record-at-a-time result.
template <class Fn>
void for_record_ranges(size_t n, size_t threads, Fn &&fn) {
    threads = std::max<size_t>(1, std::min(threads, n));
    if (threads == 1) {
        fn(size_t(0), size_t(0), n);
        return;
    }
    std::vector<std::thread> pool;
    for (size_t t = 0; t < threads; ++t)
        pool.emplace_back([&, t] { fn(t, n * t / threads, n * (t + 1) / threads); });
    for (auto &th : pool) th.join();
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
struct oracle_filter {
    std::unique_ptr<FilterBase> f;
};
struct oracle_spectrum {
    Spectrum s;
};

extern "C" {

uint64_t oracle_xorshift_u64(uint64_t x) { return xorshift_u64(x); }

uint64_t oracle_read_kmer(const uint8_t *ascii, size_t len, size_t klen, size_t pos) {
    return Packed(ascii, len).read_kmer(klen, pos);
}
uint64_t oracle_read_revcomp_kmer(const uint8_t *ascii, size_t len, size_t klen, size_t pos) {
    return Packed(ascii, len).read_revcomp_kmer(klen, pos);
}
uint64_t oracle_canonical_kmer(const uint8_t *ascii, size_t len, size_t klen, size_t pos) {
    Packed p(ascii, len);
    return canonical_kmer(p, klen, pos);
}

void oracle_nthash_canonical(const uint8_t *ascii, size_t len, size_t m, uint32_t *out) {
    std::vector<uint32_t> h;
    nthash_canonical(Packed(ascii, len), m, h);
    std::copy(h.begin(), h.end(), out);
}
void oracle_nthash_canonical_direct(const uint8_t *ascii, size_t len, size_t m, uint32_t *out) {
    std::vector<uint32_t> h;
    nthash_canonical_direct(Packed(ascii, len), m, h);
    std::copy(h.begin(), h.end(), out);
}
void oracle_minimizer_positions(const uint8_t *ascii, size_t len, size_t k, size_t m, uint32_t *out) {
    std::vector<uint32_t> pos, h;
    minimizer_positions(Packed(ascii, len), k, m, pos, h);
    std::copy(pos.begin(), pos.end(), out);
}
void oracle_minimizer_positions_brute(const uint8_t *ascii, size_t len, size_t k, size_t m,
                                      uint32_t *out) {
    std::vector<uint32_t> pos;
    minimizer_positions_brute(Packed(ascii, len), k, m, pos);
    std::copy(pos.begin(), pos.end(), out);
}
size_t oracle_super_kmers(const uint8_t *ascii, size_t len, size_t k, size_t m, uint32_t *starts,
                          uint64_t *minimizers, size_t cap) {
    std::vector<uint32_t> st;
    std::vector<uint64_t> mn;
    ScanBuf buf;
    super_kmers(Packed(ascii, len), k, m, st, mn, buf);
    for (size_t i = 0; i < st.size() && i < cap; ++i) {
        starts[i] = st[i];
        minimizers[i] = mn[i];
    }
    return st.size();
}

int oracle_config_validate(const oracle_config *cfg, char *msg, size_t cap) {
    std::string m;
    Config c = from_c(cfg);
    if (!validate(c, m) || !checked_layout(c, m)) {
        copy_msg(msg, cap, m);
        return 1;
    }
    return 0;
}

oracle_filter *oracle_filter_new(const oracle_config *cfg) {
    std::string m;
    Config c = from_c(cfg);
    if (!validate(c, m) || !checked_layout(c, m)) return nullptr;
    auto *o = new oracle_filter;
    o->f.reset(make_filter(c));
    return o;
}
void oracle_filter_free(oracle_filter *f) { delete f; }
size_t oracle_filter_counter_slots(const oracle_filter *f) { return f->f->slots(); }
void oracle_filter_counters_u32(const oracle_filter *f, uint32_t *out) { f->f->to_u32(out); }
void oracle_filter_set_counters_u32(oracle_filter *f, const uint32_t *in) { f->f->from_u32(in); }
// The counter table itself (physical order, native width); valid until the
// filter is freed.
const void *oracle_filter_data(const oracle_filter *f, size_t *bytes) { return f->f->data(*bytes); }

uint64_t oracle_filter_insert_packed(oracle_filter *f, const uint8_t *ascii, size_t len) {
    return f->f->insert_packed(Packed(ascii, len));
}

void oracle_filter_insert_ascii(oracle_filter *f, const uint8_t *seq, size_t len, uint64_t *kmers,
                                uint64_t *smers) {
    uint64_t ik = 0, is = 0;
    size_t k = f->f->cfg.k;
    for_each_valid_run(seq, len, k, [&](size_t a, size_t b) {
        ik += b - a + 1 - k;  // api.rs:723
        is += f->f->insert_packed(pack_run(seq, a, b));
    });
    if (kmers) *kmers = ik;
    if (smers) *smers = is;
}

size_t oracle_filter_estimate_packed(const oracle_filter *f, const uint8_t *ascii, size_t len,
                                     uint64_t *out) {
    std::vector<uint64_t> v;
    f->f->estimate_abundances(Packed(ascii, len), v);
    std::copy(v.begin(), v.end(), out);
    return v.size();
}

void oracle_filter_stream_packed(const oracle_filter *f, const uint8_t *ascii, size_t len,
                                 uint64_t *queried, uint64_t *checksum) {
    f->f->stream(Packed(ascii, len), *queried, *checksum);
}

// api.rs:799-853
size_t oracle_filter_estimate_ascii(const oracle_filter *f, const uint8_t *seq, size_t len,
                                    uint64_t *values, uint8_t *valid) {
    size_t k = f->f->cfg.k;
    if (len < k) return 0;
    size_t n = len + 1 - k;
    for (size_t i = 0; i < n; ++i) values[i] = 0, valid[i] = 0;
    std::vector<uint64_t> est;
    for_each_valid_run(seq, len, k, [&](size_t a, size_t b) {
        f->f->estimate_abundances(pack_run(seq, a, b), est);
        for (size_t off = 0; off < est.size(); ++off) {
            values[a + off] = est[off];
            valid[a + off] = 1;
        }
    });
    return n;
}

// Definition-level brute force: window q valid iff its k bytes are ACGT; its
// block is that of the super-k-mer (of its run) containing it; its abundance is
// the min over s-mers p in [q, q+k-s] and hashes t of C[block][h_t(p)].
size_t oracle_filter_estimate_ascii_brute(const oracle_filter *f, const uint8_t *seq, size_t len,
                                          uint64_t *values, uint8_t *valid) {
    const FilterBase &F = *f->f;
    size_t k = F.cfg.k, s = F.cfg.s;
    if (len < k) return 0;
    size_t n = len + 1 - k;
    for (size_t i = 0; i < n; ++i) values[i] = 0, valid[i] = 0;
    for_each_valid_run(seq, len, k, [&](size_t a, size_t b) {
        Packed run = pack_run(seq, a, b);
        std::vector<uint32_t> st;
        std::vector<uint64_t> mn;
        ScanBuf buf;
        super_kmers(run, k, F.cfg.m, st, mn, buf);
        size_t si = 0;
        for (size_t q = 0; q + k <= run.len; ++q) {
            while (si + 1 < st.size() && st[si + 1] <= q) ++si;
            uint64_t block = xorshift_u64(mn[si]) & F.lay.block_mask;
            uint64_t best = F.max_value();
            for (size_t p = q; p <= q + k - s; ++p) {
                uint64_t h = xorshift_u64(canonical_kmer(run, s, p));
                for (size_t t = 0; t < F.cfg.n_hashes; ++t) {
                    best = std::min(best, F.get(block, h & F.lay.block_slots_mask));
                    h = xorshift_u64(h);
                }
            }
            values[a + q] = best;
            valid[a + q] = 1;
        }
    });
    return n;
}

void oracle_filter_insert_records(oracle_filter *f, const uint8_t *bases, const uint64_t *offsets,
                                  size_t n, size_t threads, uint64_t *kmers, uint64_t *smers) {
    std::vector<uint64_t> ik(std::max<size_t>(threads, 1), 0), is(std::max<size_t>(threads, 1), 0);
    for_record_ranges(n, threads, [&](size_t t, size_t a, size_t b) {
        for (size_t r = a; r < b; ++r)
            for_each_valid_run(bases + offsets[r], size_t(offsets[r + 1] - offsets[r]), f->f->cfg.k,
                               [&](size_t x, size_t y) {
                                   const uint8_t *rec = bases + offsets[r];
                                   ik[t] += y - x + 1 - f->f->cfg.k;  // api.rs:723
                                   is[t] += f->f->insert_packed(pack_run(rec, x, y));
                               });
    });
    uint64_t tk = 0, ts = 0;
    for (size_t t = 0; t < ik.size(); ++t) tk += ik[t], ts += is[t];
    if (kmers) *kmers = tk;
    if (smers) *smers = ts;
}

// api.rs:799-853 per record; values/valid are indexed by position
// (offsets[r] + window), entries of positions where no window starts are 0.
void oracle_filter_estimate_records(const oracle_filter *f, const uint8_t *bases,
                                    const uint64_t *offsets, size_t n, size_t threads,
                                    uint64_t *values, uint8_t *valid) {
    memset(values, 0, size_t(offsets[n]) * sizeof(uint64_t));
    memset(valid, 0, size_t(offsets[n]));
    for_record_ranges(n, threads, [&](size_t, size_t a, size_t b) {
        for (size_t r = a; r < b; ++r) {
            const size_t len = size_t(offsets[r + 1] - offsets[r]);
            if (len >= f->f->cfg.k)
                oracle_filter_estimate_ascii(f, bases + offsets[r], len, values + offsets[r],
                                             valid + offsets[r]);
        }
    });
}

size_t oracle_filter_counter_histogram(const oracle_filter *f, uint64_t *values, uint64_t *counts,
                                       size_t cap) {
    std::unordered_map<uint64_t, uint64_t> h;
    f->f->histogram(h);
    std::vector<std::pair<uint64_t, uint64_t>> rows(h.begin(), h.end());
    std::sort(rows.begin(), rows.end());
    for (size_t i = 0; i < rows.size() && i < cap; ++i) values[i] = rows[i].first, counts[i] = rows[i].second;
    return rows.size();
}

oracle_spectrum *oracle_spectrum_new(void) { return new oracle_spectrum; }
void oracle_spectrum_free(oracle_spectrum *sp) { delete sp; }
uint64_t oracle_spectrum_add_chunk(oracle_spectrum *sp, const oracle_filter *f, const uint8_t *ascii,
                                   size_t len) {
    if (len < f->f->cfg.k) return 0;  // fastx.rs:284-286
    std::vector<std::pair<uint64_t, uint64_t>> pairs;
    uint64_t q = f->f->estimate_pairs(Packed(ascii, len), pairs);
    for (auto &kv : pairs) sp->s.map[Spectrum::part(kv.first)].try_emplace(kv.first, kv.second);
    return q;
}
size_t oracle_spectrum_distinct(const oracle_spectrum *sp) { return sp->s.size(); }
// estimate_distinct_kmers (fastx.rs:255-308) over records split into their
// DnaChunks (maximal ACGT runs of at least k bases, fastx.rs:14-19, 234-236).
// Worker t walks its records in order and keeps its pairs per key part; then
// each part's map takes the pairs of worker 0, 1, ... in order, first wins.
// That is the result of per-worker first-wins maps (filter.rs:454) merged
// first-wins in worker order (fastx.rs:392-400). Returns the queried k-mers.
uint64_t oracle_spectrum_add_records(oracle_spectrum *sp, const oracle_filter *f, const uint8_t *bases,
                                     const uint64_t *offsets, size_t n, size_t threads) {
    constexpr int P = Spectrum::kParts;
    const size_t T = std::max<size_t>(1, std::min(threads, std::max<size_t>(n, 1)));
    std::vector<std::vector<std::vector<std::pair<uint64_t, uint64_t>>>> parts(
        T, std::vector<std::vector<std::pair<uint64_t, uint64_t>>>(P));
    std::vector<uint64_t> q(T, 0);
    for_record_ranges(n, T, [&](size_t t, size_t a, size_t b) {
        std::vector<std::pair<uint64_t, uint64_t>> pairs;
        for (size_t r = a; r < b; ++r) {
            const uint8_t *rec = bases + offsets[r];
            for_each_valid_run(rec, size_t(offsets[r + 1] - offsets[r]), f->f->cfg.k,
                               [&](size_t x, size_t y) {
                                   pairs.clear();
                                   q[t] += f->f->estimate_pairs(pack_run(rec, x, y), pairs);
                                   for (auto &kv : pairs) parts[t][Spectrum::part(kv.first)].push_back(kv);
                               });
        }
    });
    for_record_ranges(P, T, [&](size_t, size_t a, size_t b) {
        for (size_t p = a; p < b; ++p)
            for (size_t t = 0; t < T; ++t) {
                for (auto &kv : parts[t][p]) sp->s.map[p].try_emplace(kv.first, kv.second);
                std::vector<std::pair<uint64_t, uint64_t>>().swap(parts[t][p]);
            }
    });
    uint64_t total = 0;
    for (uint64_t x : q) total += x;
    return total;
}
// The merged map itself: (key, estimate) pairs sorted by key.
size_t oracle_spectrum_pairs(const oracle_spectrum *sp, uint64_t *keys, uint64_t *values, size_t cap) {
    std::vector<std::pair<uint64_t, uint64_t>> all;
    for (auto &m : sp->s.map) all.insert(all.end(), m.begin(), m.end());
    std::sort(all.begin(), all.end());
    for (size_t i = 0; i < all.size() && i < cap; ++i) keys[i] = all[i].first, values[i] = all[i].second;
    return all.size();
}
// fastx.rs:402-410
size_t oracle_spectrum_rows(const oracle_spectrum *sp, uint64_t *values, uint64_t *counts, size_t cap) {
    std::unordered_map<uint64_t, uint64_t> h;
    for (auto &m : sp->s.map)
        for (auto &kv : m) ++h[kv.second];
    std::vector<std::pair<uint64_t, uint64_t>> rows(h.begin(), h.end());
    std::sort(rows.begin(), rows.end());
    for (size_t i = 0; i < rows.size() && i < cap; ++i) values[i] = rows[i].first, counts[i] = rows[i].second;
    return rows.size();
}

// bin/sampled_spectrum.rs:150-308: reservoir sampling of the valid k-mer
// windows of the records in input order (SplitMix64 draws, gen_below = next %
// upper), then each sampled k-mer queried on its own through
// estimate_sequence_abundances (api.rs:250-262), the first window's estimate
// counted; rows scaled as write_scaled_spectrum (sampled_spectrum.rs:225-253).
// rows holds 4 u64 per row: estimate, sampled k-mers, linear scaled windows,
// estimated distinct k-mers. stats: total windows seen, sampled k-mers,
// queried k-mers, invalid queries.
namespace {
// ReservoirSampler (sampled_spectrum.rs:255-308) over the valid windows of n
// records in order (sample_sequence_kmers, 168-199): the retained k-mers
// (uppercased) and, per slot, the 1-based number of the window it holds.
struct SampledWindows {
    std::vector<std::vector<uint8_t>> kmers;
    std::vector<uint64_t> window;
    uint64_t seen = 0;
};
SampledWindows reservoir_sample(const uint8_t *bases, const uint64_t *offsets, size_t n, size_t k,
                                uint64_t sample_size, uint64_t seed) {
    SampledWindows out;
    uint64_t state = seed;
    auto next_u64 = [&]() {  // sampled_spectrum.rs:289-295
        state += 0x9E3779B97F4A7C15ull;
        uint64_t v = state;
        v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
        v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
        return v ^ (v >> 31);
    };
    auto consider = [&](const uint8_t *kmer) {  // sampled_spectrum.rs:262-279
        ++out.seen;
        size_t slot;
        if (out.kmers.size() < sample_size) {
            slot = out.kmers.size();
        } else {
            uint64_t r = next_u64() % out.seen;  // gen_below (297-299)
            if (r >= sample_size) return;
            slot = size_t(r);
        }
        std::vector<uint8_t> up(kmer, kmer + k);
        for (auto &c : up)
            if (c >= 'a' && c <= 'z') c = uint8_t(c - 32);
        if (slot == out.kmers.size()) {
            out.kmers.push_back(std::move(up));
            out.window.push_back(out.seen);
        } else {
            out.kmers[slot] = std::move(up);
            out.window[slot] = out.seen;
        }
    };
    for (size_t r = 0; r < n; ++r) {
        const uint8_t *rec = bases + offsets[r];
        for_each_valid_run(rec, size_t(offsets[r + 1] - offsets[r]), k, [&](size_t a, size_t b) {
            for (size_t o = a; o + k <= b; ++o) consider(rec + o);
        });
    }
    return out;
}
}  // namespace

// The sampler alone: window numbers held by slots 0..min(seen, sample_size);
// returns the number of slots filled, *seen the windows considered.
size_t oracle_reservoir_windows(const uint8_t *bases, const uint64_t *offsets, size_t n, size_t k,
                                uint64_t sample_size, uint64_t seed, uint64_t *windows, uint64_t *seen) {
    SampledWindows sw = reservoir_sample(bases, offsets, n, k, sample_size, seed);
    std::copy(sw.window.begin(), sw.window.end(), windows);
    *seen = sw.seen;
    return sw.window.size();
}

size_t oracle_sampled_spectrum(const oracle_filter *f, const uint8_t *bases, const uint64_t *offsets, size_t n,
                               uint64_t sample_size, uint64_t seed, uint64_t *rows, size_t cap, uint64_t *stats) {
    SampledWindows sw = reservoir_sample(bases, offsets, n, f->f->cfg.k, sample_size, seed);
    const std::vector<std::vector<uint8_t>> &kmers = sw.kmers;
    const uint64_t seen = sw.seen;
    std::unordered_map<uint64_t, uint64_t> hist;  // sampled_spectrum.rs:201-221
    uint64_t queried = 0, invalid = 0;
    std::vector<uint64_t> v(1);
    std::vector<uint8_t> ok(1);
    for (auto &km : kmers) {
        ++queried;
        size_t got = oracle_filter_estimate_ascii(f, km.data(), km.size(), v.data(), ok.data());
        if (got && ok[0]) ++hist[v[0]];
        else ++invalid;
    }
    std::vector<std::pair<uint64_t, uint64_t>> h(hist.begin(), hist.end());
    std::sort(h.begin(), h.end());
    const double sample_count = double(std::max<size_t>(kmers.size(), 1));
    const double scale = double(seen) / sample_count;
    for (size_t i = 0; i < h.size() && i < cap; ++i) {
        const uint64_t lin = f64_as_u64(std::round(double(h[i].second) * scale));
        rows[4 * i] = h[i].first;
        rows[4 * i + 1] = h[i].second;
        rows[4 * i + 2] = lin;
        rows[4 * i + 3] = h[i].first == 0 ? 0 : f64_as_u64(std::round(double(lin) / double(h[i].first)));
    }
    stats[0] = seen, stats[1] = kmers.size(), stats[2] = queried, stats[3] = invalid;
    return h.size();
}

// api.rs:388-414, 466-480, 515-529, 545-559
int oracle_save(const oracle_filter *f, uint64_t inserted_kmers, const char *path, char *msg,
                size_t cap) {
    FILE *fp = fopen(path, "wb");
    if (!fp) {
        copy_msg(msg, cap, std::string("I/O error for '") + path + "': cannot create");
        return 7;
    }
    const Config &c = f->f->cfg;
    const Layout &l = f->f->lay;
    auto u8 = [&](uint8_t v) { fwrite(&v, 1, 1, fp); };
    auto u32 = [&](uint32_t v) { fwrite(&v, 4, 1, fp); };
    auto u64 = [&](uint64_t v) { fwrite(&v, 8, 1, fp); };
    fwrite("SCBIDX\0\0", 1, 8, fp);
    u32(2);
    u64(c.k), u64(c.m), u64(c.s), u64(c.n_hashes);
    u8(uint8_t(c.bits)), u8(uint8_t(c.cse)), u8(uint8_t(c.bse));
    u64(c.threads), u64(c.queue);
    u64(inserted_kmers);
    u64(l.shard_count), u64(l.subblocks), u64(l.block_slots), u64(l.shard_mask);
    u32(l.shard_shift);
    u64(l.block_slots_mask), u64(l.nb_blocks), u64(l.block_mask), u64(l.n_hashes);
    f->f->write_le(fp);
    fclose(fp);
    return 0;
}

// api.rs:416-464, 482-513, 531-543, 561-608
int oracle_load(const char *path, oracle_filter **out, oracle_config *cfg_out, uint64_t *ik,
                char *msg, size_t cap) {
    *out = nullptr;
    FILE *fp = fopen(path, "rb");
    std::string p(path);
    if (!fp) {
        copy_msg(msg, cap, "I/O error for '" + p + "': No such file or directory (os error 2)");
        return 7;
    }
    auto eof = [&]() {
        copy_msg(msg, cap, "I/O error for '" + p + "': failed to fill whole buffer");
        fclose(fp);
        return 7;
    };
    auto bad = [&](const std::string &m) {
        copy_msg(msg, cap, "invalid index format: " + m);
        fclose(fp);
        return 3;
    };
    char magic[8];
    if (fread(magic, 1, 8, fp) != 8) return eof();
    if (memcmp(magic, "SCBIDX\0\0", 8) != 0) return bad("bad SuperCountingBloom index magic");
    uint32_t version;
    if (fread(&version, 4, 1, fp) != 1) return eof();
    if (version != 2) return bad("unsupported index format version " + std::to_string(version));
    uint64_t k, m, s, nh, threads, queue;
    uint8_t bits, cse, bse;
    if (fread(&k, 8, 1, fp) != 1 || fread(&m, 8, 1, fp) != 1 || fread(&s, 8, 1, fp) != 1 ||
        fread(&nh, 8, 1, fp) != 1 || fread(&bits, 1, 1, fp) != 1)
        return eof();
    if (bits != 8 && bits != 16 && bits != 32)
        return bad("unsupported counter width " + std::to_string(bits));
    if (fread(&cse, 1, 1, fp) != 1 || fread(&bse, 1, 1, fp) != 1 || fread(&threads, 8, 1, fp) != 1 ||
        fread(&queue, 8, 1, fp) != 1)
        return eof();
    Config c{size_t(k), size_t(m), size_t(s), size_t(nh), bits, cse, bse, size_t(threads), size_t(queue)};
    std::string vm;
    if (!validate(c, vm)) {
        copy_msg(msg, cap, "invalid config: " + vm);
        fclose(fp);
        return 1;
    }
    uint64_t inserted;
    if (fread(&inserted, 8, 1, fp) != 1) return eof();
    Layout st{};
    if (fread(&st.shard_count, 8, 1, fp) != 1 || fread(&st.subblocks, 8, 1, fp) != 1 ||
        fread(&st.block_slots, 8, 1, fp) != 1 || fread(&st.shard_mask, 8, 1, fp) != 1 ||
        fread(&st.shard_shift, 4, 1, fp) != 1 || fread(&st.block_slots_mask, 8, 1, fp) != 1 ||
        fread(&st.nb_blocks, 8, 1, fp) != 1 || fread(&st.block_mask, 8, 1, fp) != 1 ||
        fread(&st.n_hashes, 8, 1, fp) != 1)
        return eof();
    if (!checked_layout(c, vm)) {
        copy_msg(msg, cap, "invalid config: " + vm);
        fclose(fp);
        return 1;
    }
    if (!(st == layout_from_config(c))) return bad("stored filter layout does not match config");
    std::unique_ptr<FilterBase> F(make_filter(c));
    if (!F->read_le(fp)) return eof();
    int extra = fgetc(fp);
    if (extra != EOF) return bad("trailing bytes after serialized index");
    fclose(fp);
    auto *o = new oracle_filter;
    o->f = std::move(F);
    *out = o;
    cfg_out->k = k, cfg_out->m = m, cfg_out->s = s, cfg_out->n_hashes = nh;
    cfg_out->counter_bits = bits, cfg_out->counter_slots_exponent = cse;
    cfg_out->block_slots_exponent = bse, cfg_out->threads = threads, cfg_out->queue = queue;
    *ik = inserted;
    return 0;
}

size_t oracle_deconvolved_counts(const uint64_t *values, const uint64_t *counts, size_t rows,
                                 uint64_t n_hashes, uint64_t max_abundance, double *out) {
    auto d = deconvolved_counts(values, counts, rows, n_hashes, max_abundance);
    std::copy(d.begin(), d.end(), out);
    return d.size();
}

// api.rs:933-956, 1013-1033
size_t oracle_filter_only_spectrum(const uint64_t *values, const uint64_t *counts, size_t rows,
                                   uint64_t n_hashes, uint64_t max_abundance,
                                   uint64_t kmer_count_hint, uint64_t *out_counts) {
    auto d = deconvolved_counts(values, counts, rows, n_hashes, max_abundance);
    double total = 0.0;
    for (double x : d) total += x;
    if (total <= 0.0) {
        uint64_t pos = 0;
        for (size_t i = 0; i < rows; ++i)
            if (values[i] > 0) pos += counts[i];
        double positive = double(std::max<uint64_t>(pos, 1));
        for (uint64_t a = 1; a <= max_abundance; ++a) {
            uint64_t c = 0;
            for (size_t i = 0; i < rows; ++i)
                if (values[i] == a) c = counts[i];
            out_counts[a - 1] = f64_as_u64(std::round(double(kmer_count_hint) * double(c) / positive));
        }
        return size_t(max_abundance);
    }
    double scale = double(kmer_count_hint) / total;
    for (uint64_t a = 1; a <= max_abundance; ++a)
        out_counts[a - 1] = f64_as_u64(std::round(d[size_t(a - 1)] * scale));
    return size_t(max_abundance);
}

// api.rs:375-385
uint64_t oracle_filter_only_count_hint(const uint64_t *values, const uint64_t *counts, size_t rows,
                                       uint64_t n_hashes, uint64_t max_abundance) {
    auto d = deconvolved_counts(values, counts, rows, n_hashes, max_abundance);
    double total = 0.0;
    for (double x : d) total += x;
    return f64_as_u64(std::fmax(std::round(total), 1.0));
}

// api.rs:855-891
size_t oracle_summarize(const uint64_t *values, const uint8_t *valid, size_t n, double *mean,
                        double *median) {
    std::vector<uint64_t> v;
    for (size_t i = 0; i < n; ++i)
        if (valid[i]) v.push_back(values[i]);
    if (v.empty()) {
        *mean = NAN;
        *median = NAN;
        return 0;
    }
    uint64_t sum = 0;
    for (uint64_t x : v) sum += x;
    *mean = double(sum) / double(v.size());
    std::sort(v.begin(), v.end());
    size_t mid = v.size() / 2;
    *median = v.size() % 2 == 0 ? (double(v[mid - 1]) + double(v[mid])) / 2.0 : double(v[mid]);
    return v.size();
}

// fastx.rs:172-204 (stream_estimates_from_fastx_typed), 206-253 (build_filter),
// 310-369 (stream_query_kmers), with split_non_actg chunking (fastx.rs:14-19).
int oracle_cpu_baseline(const oracle_config *cfgc, const uint8_t *ib, const uint64_t *io, size_t ni,
                        const uint8_t *qb, const uint64_t *qo, size_t nq, oracle_baseline_result *out) {
    std::string m;
    Config cfg = from_c(cfgc);
    if (!validate(cfg, m) || !checked_layout(cfg, m)) return 1;
    memset(out, 0, sizeof(*out));
    std::unique_ptr<FilterBase> F(make_filter(cfg));
    auto produce = [&](const uint8_t *bases, const uint64_t *offs, size_t n, Channel &ch,
                       uint64_t *chunks, uint64_t *kmers) {
        for (size_t r = 0; r < n; ++r) {
            const uint8_t *rec = bases + offs[r];
            size_t len = size_t(offs[r + 1] - offs[r]);
            // split_non_actg: every maximal ACGT run; runs shorter than k are skipped
            size_t i = 0;
            while (i < len) {
                while (i < len && !is_acgt(rec[i])) ++i;
                size_t a = i;
                while (i < len && is_acgt(rec[i])) ++i;
                if (i - a >= cfg.k) {
                    if (chunks) ++*chunks;
                    if (kmers) *kmers += i - a + 1 - cfg.k;
                    ch.send(std::vector<uint8_t>(rec + a, rec + i));
                }
            }
        }
        ch.close();
    };
    auto t0 = std::chrono::steady_clock::now();
    {
        Channel ch(cfg.queue);
        std::vector<std::thread> workers;
        std::vector<uint64_t> smers(cfg.threads, 0);
        for (size_t t = 0; t < cfg.threads; ++t)
            workers.emplace_back([&, t] {
                std::vector<uint8_t> rec;
                Packed p;
                while (ch.recv(rec)) {
                    p.assign(rec.data(), rec.size());
                    smers[t] += F->insert_packed(p);
                }
            });
        produce(ib, io, ni, ch, &out->chunks, &out->indexed_kmers);
        for (auto &w : workers) w.join();
        for (uint64_t s : smers) out->inserted_smers += s;
    }
    auto t1 = std::chrono::steady_clock::now();
    {
        Channel ch(cfg.queue);
        std::vector<std::thread> workers;
        struct Acc {
            uint64_t q = 0, pos = 0, sum = 0;
        };
        std::vector<Acc> acc(cfg.threads);
        for (size_t t = 0; t < cfg.threads; ++t)
            workers.emplace_back([&, t] {
                std::vector<uint8_t> rec;
                std::vector<uint64_t> est;
                Packed p;
                while (ch.recv(rec)) {
                    p.assign(rec.data(), rec.size());
                    F->estimate_abundances(p, est);
                    acc[t].q += est.size();
                    for (uint64_t e : est) {
                        if (e > 0) ++acc[t].pos;
                        acc[t].sum += e;
                    }
                }
            });
        produce(qb, qo, nq, ch, nullptr, nullptr);
        for (auto &w : workers) w.join();
        for (auto &a : acc) {
            out->queried_kmers += a.q;
            out->positive_estimates += a.pos;
            out->estimate_checksum += a.sum;
        }
    }
    auto t2 = std::chrono::steady_clock::now();
    out->build_s = std::chrono::duration<double>(t1 - t0).count();
    out->query_s = std::chrono::duration<double>(t2 - t1).count();
    return 0;
}

}  // extern "C"
