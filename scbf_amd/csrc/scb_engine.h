// Internal interface between the C ABI (scb_api.cpp) and the HIP engine
// (scb_kernels.hip). Not installed; no torch types anywhere.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace scb {

// Filter geometry derived from SuperCountingBloomConfig and FrozenFilterLayout
// (config.rs:20-107, filter.rs:90-116).
struct Geometry {
    uint32_t k = 0, m = 0, s = 0, nh = 0;  // k-mer, minimizer, s-mer lengths, n_hashes
    uint32_t w = 0;                         // minimizer window in m-mers: k - m + 1
    uint32_t win = 0;                       // fimpera window in s-mers: k - s + 1
    uint32_t bits = 16, bytes = 2;          // counter width
    uint32_t cse = 0, bse = 0;              // log2 counter slots, log2 block slots
    uint64_t slots = 0, nb = 0, block_mask = 0;
    uint64_t shard_count = 0, subblocks = 0;
    uint32_t shard_mask = 0, shard_shift = 0, sub_log = 0;
    uint64_t table_bytes() const { return slots * bytes; }
};

// Status of an engine call: code 0 = ok, otherwise an SCB_ERR_* value.
struct Status {
    int code = 0;
    std::string msg;
    static Status ok() { return {}; }
    bool good() const { return code == 0; }
};

// A batch of records already in device memory: d_bases holds n_bases bytes,
// d_offsets n_reads + 1 u64 (d_offsets[0] = 0, d_offsets[n_reads] = n_bases).
struct DeviceBatch {
    const uint8_t *d_bases = nullptr;
    const uint64_t *d_offsets = nullptr;
    uint64_t n_reads = 0, n_bases = 0;
};

struct InsertResult {
    uint64_t inserted_kmers = 0;   // valid windows (api.rs:723)
    uint64_t inserted_smers = 0;   // filter.rs:232
    uint64_t super_kmers = 0;
};

struct QueryResult {
    uint64_t queried = 0, positive = 0, checksum = 0;  // fastx.rs:325-334 / api.rs:284-296
    uint64_t super_kmers = 0;
};

struct SpectrumResult {
    uint64_t queried = 0, distinct = 0;
    std::vector<std::pair<uint64_t, uint64_t>> rows;  // (estimate, #distinct keys), sorted
};

// kAuto: the plain method for batches below plain_bases() bases, where its
// one thread per super-k-mer finishes before the batched pipeline's fixed cost
// (a plan over every bin and, for the query, every bin with records staged),
// and the batched method above.
enum Method { kBatched = 0, kPlain = 1, kAuto = 2 };

// Tunables of the batched pipeline, per engine: defaults from the geometry,
// replaced by a tuning cache entry for the geometry (written by the autotuner,
// scbf_amd/tune.py) or by set_tuning. Every valid setting gives the same
// counters and estimates; only the speed differs.
struct Tuning {
    uint32_t insert_bin_log = 15;         // slots per insert bin (16-bit LDS counts)
    uint32_t query_bin_log = 15;          // slots per query bin (native-width LDS copy)
    uint32_t spectrum_bin_log = 14;       // slots per spectrum bin (bse <= 14)
    uint32_t insert_item_records = 2048;  // hot-bin split sizes (records per work item)
    uint32_t query_item_records = 3584;
    uint32_t spectrum_item_records = 4096;
    uint32_t use_graph = 0;               // replay repeated device-batch pipelines from a HIP graph
};

// Geometry key of a tuning cache entry.
struct TuningKey {
    uint32_t k, m, s, nh, bits, cse, bse;
    bool operator==(const TuningKey &o) const {
        return k == o.k && m == o.m && s == o.s && nh == o.nh && bits == o.bits && cse == o.cse && bse == o.bse;
    }
};
// Replaces the process-wide tuning cache with the entries of a JSON file (the
// format tune.py writes); engines created afterwards use them.
Status load_tuning_cache(const char *path);

class Workspace;  // device scratch buffers + stream, one per concurrent call

// Distinct canonical k-mers accumulated over several spectrum calls (the
// per-worker maps of fastx.rs:255-308 merged as fastx.rs:392-400), held in
// device memory; created by Engine::keyset_create.
class KeySet;
struct KeySetDeleter {
    void operator()(KeySet *k) const;
};
using KeySetPtr = std::unique_ptr<KeySet, KeySetDeleter>;

// Reservoir of sampled valid k-mer windows and their estimates
// (bin/sampled_spectrum.rs:150-308), held in device memory.
class Reservoir;
struct ReservoirDeleter {
    void operator()(Reservoir *r) const;
};
using ReservoirPtr = std::unique_ptr<Reservoir, ReservoirDeleter>;

// Counter table resident in HBM in physical (.scb payload) block order
// P(b) = (b & (shard_count-1)) * subblocks + (b >> log2(shard_count)).
class Engine {
   public:
    static Status create(const Geometry &g, std::unique_ptr<Engine> &out);
    ~Engine();

    const Geometry &geometry() const { return geo_; }
    int device() const { return device_; }
    const Tuning &tuning() const { return tune_; }
    // checks the setting against the geometry and the kernels' LDS limits
    Status set_tuning(const Tuning &t);
    static Tuning default_tuning(const Geometry &g);

    // the method kAuto resolves to for a batch of n_bases bases
    Method resolve(Method m, uint64_t n_bases) const { return m != kAuto ? m : n_bases < plain_bases() ? kPlain : kBatched; }
    // Below this many bases the plain method is faster: 2^17 bases plus one per
    // 256 table bytes (the batched query stages the bins it touches; measured on
    // the MI355X at 2^30 16-bit slots: plain 0.11 ms against batched 0.42 ms at
    // 150 kbp, 0.60 against 0.91 ms at 4.5 Mbp, 1.56 against 0.98 ms at 15 Mbp).
    uint64_t plain_bases() const { return (uint64_t(1) << 17) + geo_.table_bytes() / 256; }

    Status insert(const DeviceBatch &b, Method method, InsertResult &r);
    // d_values: n_bases counters of geo.bytes each (or null); d_valid_bits:
    // n_bases bits rounded up to u32 words (or null).
    Status query(const DeviceBatch &b, Method method, void *d_values, uint32_t *d_valid_bits,
                 QueryResult &r) const;
    // query with every per-call device buffer taken from a fixed scratch array
    // (kNoAllocScratchBytes per device, held until the process ends): no
    // hipMalloc or hipFree; b.n_bases <= kNoAllocMaxBases; runs on the null
    // stream; calls on one device are serialized.
    Status query_noalloc(const DeviceBatch &b, Method method, void *d_values, uint32_t *d_valid_bits,
                         QueryResult &r) const;
    // kAuto as for query; spectra are the same with every method
    Status spectrum(const DeviceBatch &b, SpectrumResult &r, Method method = kBatched) const;
    // Spectrum over several batches: each call adds the batch's distinct keys
    // to ks (keys already in it are not counted again); keyset_rows gives the
    // rows, the distinct keys and the k-mers queried so far.
    Status keyset_create(KeySetPtr &out) const;
    Status spectrum_accumulate(const DeviceBatch &b, KeySet &ks) const;
    Status keyset_rows(const KeySet &ks, SpectrumResult &r) const;
    // filter.rs:302-335: (canonical k-mer, estimate) of every distinct k-mer of
    // the batch (first occurrence; every occurrence has the same estimate)
    Status kmer_map(const DeviceBatch &b, std::vector<std::pair<uint64_t, uint32_t>> &pairs,
                    uint64_t &queried) const;
    // Reservoir sampling of the valid windows of batches in input order (a
    // capacity of cap windows, SplitMix64 draws from seed) with each sampled
    // window's estimate; reservoir_values gives the min(seen, cap) estimates by
    // slot and the number of windows seen.
    Status reservoir_create(uint64_t cap, uint64_t seed, ReservoirPtr &out) const;
    Status sample_accumulate(const DeviceBatch &b, Reservoir &rs) const;
    Status sample_accumulate_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n, Reservoir &rs) const;
    Status reservoir_values(const Reservoir &rs, std::vector<uint32_t> &vals, uint64_t &seen) const;
    Status counter_histogram(std::vector<std::pair<uint64_t, uint64_t>> &rows) const;

    // Host <-> HBM copies of the whole table (physical order), and of batches.
    Status download(void *host, uint64_t bytes) const;
    Status upload(const void *host, uint64_t bytes);

    // Host-memory convenience: stages records into device memory and runs the
    // device path in chunks of at most max_chunk_bases.
    Status insert_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n, Method method,
                       InsertResult &r);
    // values/valid indexed by position; either may be null.
    Status query_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n, Method method,
                      uint64_t *values, uint8_t *valid, QueryResult &r) const;
    // any size: inputs beyond one device batch are counted chunk by chunk in a key set
    Status spectrum_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n,
                         SpectrumResult &r, Method method = kBatched) const;
    Status spectrum_accumulate_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n, KeySet &ks) const;
    // pairs sorted by key
    Status kmer_map_host(const uint8_t *bases, const uint64_t *offsets, uint64_t n,
                         std::vector<std::pair<uint64_t, uint32_t>> &pairs, uint64_t &queried) const;

   private:
    Engine() = default;
    // A new table is not cleared when it is allocated: the first batched insert
    // writes every counter word of the table (merging its counts into zero
    // instead of reading the table), and any other first use clears it.
    Status ensure_zeroed() const;
    Geometry geo_;
    Tuning tune_;
    void *d_table_ = nullptr;
    int device_ = 0;
    mutable std::mutex zero_mu_;
    mutable bool zero_pending_ = false;
    std::unique_ptr<Workspace> acquire() const;
    void release(std::unique_ptr<Workspace> ws) const;
    Status run_query(Workspace &ws, const DeviceBatch &b, Method method, void *d_values, uint32_t *d_valid_bits,
                     QueryResult &r) const;
    Status spectrum_checks(const DeviceBatch &b) const;
    // hist: the kernel-side histogram targets (HistOut) or null with emit
    Status spectrum_pass(Workspace &ws, const DeviceBatch &b, const void *hist, bool emit) const;
    Status spectrum_plain(Workspace &ws, const DeviceBatch &b, const void *hist) const;
};

// Device time of the stages of the calling thread's last insert/query call,
// measured with HIP events on the engine's stream.
struct StageTimes {
    float scan_ms = 0, plan_scatter_ms = 0, bin_ms = 0, total_ms = 0;
    bool batched = false;
};
StageTimes last_stage_times();

// Stream-ordered helpers for host code that drives its own HIP streams
// (stream is a hipStream_t).
Status fill_u32_async(uint32_t *d_p, uint64_t n, uint32_t value, void *stream);
// Statistics of query outputs per segment of positions: segs holds nseg sorted,
// disjoint [begin, end) position pairs; out receives, per segment, the number
// of valid windows, of valid windows with a positive estimate and the wrapping
// sum of the estimates (3 u64 each).
Status segment_stats_async(uint32_t counter_bits, const void *d_values, const uint32_t *d_valid_bits,
                           const uint64_t *d_segs, uint32_t nseg, uint64_t n_positions, uint64_t *d_out,
                           void *stream);

Status make_geometry(uint64_t k, uint64_t m, uint64_t s, uint64_t nh, uint32_t bits, uint32_t cse,
                     uint32_t bse, Geometry &g);
int device_count();

// Maximum bases per device batch (positions are 32-bit inside the kernels).
constexpr uint64_t kMaxBatchBases = (uint64_t(1) << 32) - 4096;

// Allocation-free query: the fixed per-device scratch array and the largest
// batch it serves (about 16.3 bytes per base plus the bin plan).
constexpr uint64_t kNoAllocScratchBytes = uint64_t(96) << 20;
constexpr uint64_t kNoAllocMaxBases = uint64_t(1) << 22;

}  // namespace scb
