// scb.hpp: header-only C++17 mirror of the super-counting-bloom crate API
// (src/lib.rs:15-31) on top of the C ABI in scb.h. Type and method names,
// argument meaning and error kinds follow the Rust crate; failures throw
// scb::SuperCountingBloomError carrying the SCB_ERR_* kind and the reference's
// Display text.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <exception>
#include <functional>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "scb.h"

namespace scb {

// error.rs:3-30
class SuperCountingBloomError : public std::runtime_error {
   public:
    SuperCountingBloomError(int kind, const std::string &msg) : std::runtime_error(msg), kind_(kind) {}
    int kind() const { return kind_; }

   private:
    int kind_;
};

inline void check(int rc) {
    if (rc == SCB_OK) return;
    size_t n = scb_last_error(nullptr, 0);
    std::string msg(n, '\0');
    scb_last_error(msg.data(), n + 1);
    throw SuperCountingBloomError(rc, msg);
}

using ApproxSpectrum = std::vector<std::pair<uint64_t, uint64_t>>;  // spectrum.rs:9

enum class CounterBits : uint8_t { Bits8 = 8, Bits16 = 16, Bits32 = 32 };  // config.rs:3-18

// config.rs:20-47
struct SuperCountingBloomConfig {
    size_t k = 31, m = 21, s = 25, n_hashes = 4;
    CounterBits counter_bits = CounterBits::Bits16;
    uint8_t counter_slots_exponent = 30, block_slots_exponent = 9;
    size_t threads = 1, queue = 4096;

    static SuperCountingBloomConfig defaults() {
        scb_config c;
        scb_config_default(&c);
        return from_c(c);
    }
    void validate() const {
        scb_config c = to_c();
        check(scb_config_validate(&c));
    }
    scb_config to_c() const {
        scb_config c{};
        c.k = k, c.m = m, c.s = s, c.n_hashes = n_hashes;
        c.counter_bits = uint8_t(counter_bits);
        c.counter_slots_exponent = counter_slots_exponent;
        c.block_slots_exponent = block_slots_exponent;
        c.threads = threads, c.queue = queue;
        return c;
    }
    static SuperCountingBloomConfig from_c(const scb_config &c) {
        SuperCountingBloomConfig r;
        r.k = c.k, r.m = c.m, r.s = c.s, r.n_hashes = c.n_hashes;
        r.counter_bits = CounterBits(c.counter_bits);
        r.counter_slots_exponent = c.counter_slots_exponent;
        r.block_slots_exponent = c.block_slots_exponent;
        r.threads = c.threads, r.queue = c.queue;
        return r;
    }
};

using AddReport = scb_add_report;          // api.rs:66-73
using QueryFastaReport = scb_query_report;  // api.rs:75-82
using BuildReport = scb_build_report;       // fastx.rs:21-34 (build half)
using Tuning = scb_tuning;                  // extension: tunables of the batched method
using SampledRow = scb_sampled_row;         // sampled_spectrum.rs:225-253
using SampledStats = scb_sampled_stats;     // sampled_spectrum.rs:113-146

// extension: which device method runs a batch (results are the same with every method)
enum class Method : int { Batched = SCB_METHOD_BATCHED, Plain = SCB_METHOD_PLAIN, Auto = SCB_METHOD_AUTO };

// A batch of records in the layout every batched entry point takes: record i
// is bases[offsets[i], offsets[i+1]).
struct Records {
    std::vector<uint8_t> bases;
    std::vector<uint64_t> offsets{0};
    void add(const std::string &seq) {
        bases.insert(bases.end(), seq.begin(), seq.end());
        offsets.push_back(bases.size());
    }
    size_t size() const { return offsets.size() - 1; }
    uint64_t total_bases() const { return offsets.back(); }
    std::string record(size_t i) const {
        return std::string(bases.begin() + offsets[i], bases.begin() + offsets[i + 1]);
    }
};

// estimate_sequence_abundances over a batch (api.rs:250-262 per record): the
// estimate of the window starting at each position of the concatenated records.
struct SequenceEstimates {
    std::vector<uint64_t> values;
    std::vector<uint8_t> valid;
    QueryFastaReport report;
    std::optional<uint64_t> at(uint64_t pos) const {
        return valid[pos] ? std::optional<uint64_t>(values[pos]) : std::nullopt;
    }
};

// Distinct-k-mer spectrum (fastx.rs:255-308, 392-410): rows (estimate, #keys).
struct DistinctSpectrum {
    ApproxSpectrum rows;
    uint64_t queried = 0, distinct = 0;
};

// The HashMap<KmerKey, C> of estimate_packed_sequence (filter.rs:302-335).
struct KmerEstimates {
    std::unordered_map<uint64_t, uint64_t> map;
    uint64_t queried = 0;
};

struct SampledSpectrum {  // sampled_spectrum.rs:150-253
    std::vector<SampledRow> rows;
    SampledStats stats;
};

struct SequenceAbundanceSummary {  // api.rs:84-92
    std::string name;
    size_t sequence_len, total_windows, valid_kmers;
    std::optional<double> mean_abundance, median_abundance;
};

struct FilterOnlySpectrumEstimate {  // api.rs:94-99
    const char *method;
    uint64_t kmer_count_hint;
    ApproxSpectrum rows;
};

class SuperCountingBloomIndex;

// api.rs:101-208
class SuperCountingBloomBuilder {
   public:
    explicit SuperCountingBloomBuilder(const SuperCountingBloomConfig &cfg) {
        scb_config c = cfg.to_c();
        check(scb_builder_new(&c, &h_));
    }
    ~SuperCountingBloomBuilder() {
        if (h_) scb_builder_free(h_);
    }
    SuperCountingBloomBuilder(SuperCountingBloomBuilder &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    SuperCountingBloomBuilder(const SuperCountingBloomBuilder &) = delete;
    SuperCountingBloomBuilder &operator=(const SuperCountingBloomBuilder &) = delete;

    uint64_t add_sequence(const std::string &sequence) {
        uint64_t n = 0;
        check(scb_builder_add_sequence(h_, reinterpret_cast<const uint8_t *>(sequence.data()),
                                       sequence.size(), &n));
        return n;
    }
    AddReport add_sequences(const uint8_t *bases, const uint64_t *offsets, size_t n) {
        AddReport r{};
        check(scb_builder_add_sequences(h_, bases, offsets, n, &r));
        return r;
    }
    AddReport add_sequences(const Records &r) { return add_sequences(r.bases.data(), r.offsets.data(), r.size()); }
    // HBM-resident bases and offsets (scb_builder_add_sequences_device)
    AddReport add_sequences_device(const void *d_bases, const void *d_offsets, size_t n, uint64_t total_bases) {
        AddReport r{};
        check(scb_builder_add_sequences_device(h_, static_cast<const uint8_t *>(d_bases),
                                               static_cast<const uint64_t *>(d_offsets), n, total_bases, &r));
        return r;
    }
    AddReport add_fasta(const std::string &path) {
        AddReport r{};
        check(scb_builder_add_fasta(h_, path.c_str(), &r));
        return r;
    }
    inline SuperCountingBloomIndex into_index() &&;
    void set_method(Method m) { check(scb_builder_set_method(h_, int(m))); }
    void set_tuning(const Tuning &t) { check(scb_builder_set_tuning(h_, &t)); }
    Tuning tuning() const {
        Tuning t{};
        scb_builder_tuning(h_, &t);
        return t;
    }
    // The counter table in .scb payload order, counter_bits/8 bytes per counter.
    std::vector<uint8_t> counter_bytes() const {
        std::vector<uint8_t> v(table_bytes(config()));
        check(scb_builder_counters(h_, v.data(), v.size()));
        return v;
    }
    static size_t table_bytes(const SuperCountingBloomConfig &c) {
        return (size_t(1) << c.counter_slots_exponent) * (size_t(c.counter_bits) / 8);
    }
    uint64_t inserted_kmers() const { return scb_builder_inserted_kmers(h_); }
    uint64_t inserted_smers() const { return scb_builder_inserted_smers(h_); }
    void set_threads(size_t threads) { check(scb_builder_set_threads(h_, threads)); }
    void clear_threads() { scb_builder_clear_threads(h_); }
    size_t threads() const { return scb_builder_threads(h_); }
    SuperCountingBloomConfig config() const {
        scb_config c;
        scb_builder_config(h_, &c);
        return SuperCountingBloomConfig::from_c(c);
    }

   private:
    scb_builder *h_ = nullptr;
};

// api.rs:210-386
class SuperCountingBloomIndex {
   public:
    explicit SuperCountingBloomIndex(scb_index *h) : h_(h) {}
    ~SuperCountingBloomIndex() {
        if (h_) scb_index_free(h_);
    }
    SuperCountingBloomIndex(SuperCountingBloomIndex &&o) noexcept : h_(o.h_) { o.h_ = nullptr; }
    SuperCountingBloomIndex(const SuperCountingBloomIndex &) = delete;
    SuperCountingBloomIndex &operator=(const SuperCountingBloomIndex &) = delete;

    static SuperCountingBloomIndex build_from_fastx(const std::string &path, const SuperCountingBloomConfig &cfg,
                                                    scb_build_report *report = nullptr) {
        scb_config c = cfg.to_c();
        scb_index *h = nullptr;
        check(scb_index_build_from_fastx(path.c_str(), &c, &h, report));
        return SuperCountingBloomIndex(h);
    }
    static SuperCountingBloomIndex load(const std::string &path) {
        scb_index *h = nullptr;
        check(scb_index_load(path.c_str(), &h));
        return SuperCountingBloomIndex(h);
    }
    void save(const std::string &path) const { check(scb_index_save(h_, path.c_str())); }
    SuperCountingBloomConfig config() const {
        scb_config c;
        scb_index_config(h_, &c);
        return SuperCountingBloomConfig::from_c(c);
    }
    uint64_t inserted_kmers() const { return scb_index_inserted_kmers(h_); }

    std::vector<std::optional<uint64_t>> estimate_sequence_abundances(const std::string &seq) const {
        size_t k = config().k;
        size_t n = seq.size() < k ? 0 : seq.size() + 1 - k;
        std::vector<uint64_t> v(n ? n : 1);
        std::vector<uint8_t> ok(n ? n : 1);
        size_t got = 0;
        check(scb_index_estimate_sequence_abundances(h_, reinterpret_cast<const uint8_t *>(seq.data()),
                                                     seq.size(), v.data(), ok.data(), &got));
        std::vector<std::optional<uint64_t>> out(n);
        for (size_t i = 0; i < n; ++i)
            if (ok[i]) out[i] = v[i];
        return out;
    }
    SequenceAbundanceSummary summarize_sequence(const std::string &name, const std::string &seq) const {
        scb_summary s{};
        check(scb_index_summarize_sequence(h_, reinterpret_cast<const uint8_t *>(seq.data()), seq.size(), &s));
        SequenceAbundanceSummary r{name, size_t(s.sequence_len), size_t(s.total_windows), size_t(s.valid_kmers),
                                   std::nullopt, std::nullopt};
        if (!std::isnan(s.mean_abundance)) r.mean_abundance = s.mean_abundance;
        if (!std::isnan(s.median_abundance)) r.median_abundance = s.median_abundance;
        return r;
    }
    // estimate_sequence_abundances of every record of a batch in one device pass
    SequenceEstimates estimate_sequences(const Records &r) const {
        SequenceEstimates e;
        e.values.resize(r.total_bases() ? r.total_bases() : 1);
        e.valid.resize(e.values.size());
        e.report = QueryFastaReport{};
        check(scb_index_estimate_sequences(h_, r.bases.data(), r.offsets.data(), r.size(), e.values.data(),
                                           e.valid.data(), &e.report));
        e.values.resize(r.total_bases()), e.valid.resize(r.total_bases());
        return e;
    }
    // HBM-resident form: d_values holds counter_bits/8 bytes per position, d_valid_bits one bit
    QueryFastaReport estimate_sequences_device(const void *d_bases, const void *d_offsets, size_t n,
                                               uint64_t total_bases, void *d_values, uint32_t *d_valid_bits) const {
        QueryFastaReport q{};
        check(scb_index_estimate_sequences_device(h_, static_cast<const uint8_t *>(d_bases),
                                                  static_cast<const uint64_t *>(d_offsets), n, total_bases,
                                                  d_values, d_valid_bits, &q));
        return q;
    }
    // estimate_packed_sequence (filter.rs:302-335) over the records' ACGT runs
    KmerEstimates estimate_kmer_map(const Records &r) const {
        size_t cap = r.total_bases() ? r.total_bases() : 1, n = 0;
        std::vector<uint64_t> keys(cap), vals(cap);
        KmerEstimates out;
        check(scb_index_estimate_kmer_map(h_, r.bases.data(), r.offsets.data(), r.size(), keys.data(), vals.data(),
                                          cap, &n, &out.queried));
        out.map.reserve(n);
        for (size_t i = 0; i < n; ++i) out.map.emplace(keys[i], vals[i]);
        return out;
    }
    DistinctSpectrum spectrum(const Records &r) const {
        return rows_of([&](uint64_t *v, uint64_t *c, size_t cap, size_t *n, uint64_t *q, uint64_t *d) {
            return scb_index_spectrum(h_, r.bases.data(), r.offsets.data(), r.size(), v, c, cap, n, q, d);
        });
    }
    // sample_inserted_kmers + query_sampled_kmers (sampled_spectrum.rs:150-221)
    SampledSpectrum sampled_spectrum(const Records &r, uint64_t sample_size, uint64_t seed) const {
        SampledSpectrum s{};
        std::vector<SampledRow> rows(1 << 16);
        size_t n = 0;
        check(scb_index_sampled_spectrum(h_, r.bases.data(), r.offsets.data(), r.size(), sample_size, seed,
                                         rows.data(), rows.size(), &n, &s.stats));
        if (n > rows.size()) {
            rows.resize(n);
            check(scb_index_sampled_spectrum(h_, r.bases.data(), r.offsets.data(), r.size(), sample_size, seed,
                                             rows.data(), rows.size(), &n, &s.stats));
        }
        rows.resize(n);
        s.rows = std::move(rows);
        return s;
    }
    void set_method(Method m) { check(scb_index_set_method(h_, int(m))); }
    void set_tuning(const Tuning &t) { check(scb_index_set_tuning(h_, &t)); }
    Tuning tuning() const {
        Tuning t{};
        scb_index_tuning(h_, &t);
        return t;
    }
    std::vector<uint8_t> counter_bytes() const {
        std::vector<uint8_t> v(SuperCountingBloomBuilder::table_bytes(config()));
        check(scb_index_counters(h_, v.data(), v.size()));
        return v;
    }
    QueryFastaReport query_fasta(const std::string &path) const {
        QueryFastaReport r{};
        check(scb_index_query_fasta(h_, path.c_str(), &r));
        return r;
    }
    ApproxSpectrum counter_histogram() const {
        size_t n = 0;
        std::vector<uint64_t> v(1 << 16), c(1 << 16);
        check(scb_index_counter_histogram(h_, v.data(), c.data(), v.size(), &n));
        if (n > v.size()) {
            v.resize(n), c.resize(n);
            check(scb_index_counter_histogram(h_, v.data(), c.data(), v.size(), &n));
        }
        ApproxSpectrum rows;
        for (size_t i = 0; i < n; ++i) rows.emplace_back(v[i], c[i]);
        return rows;
    }
    FilterOnlySpectrumEstimate estimate_filter_only_spectrum(uint64_t max_abundance, uint64_t hint) const {
        std::vector<uint64_t> counts(max_abundance ? max_abundance : 1);
        check(scb_index_filter_only_spectrum(h_, nullptr, nullptr, 0, max_abundance, hint, counts.data()));
        FilterOnlySpectrumEstimate e{"compound_poisson_counter_deconvolution", hint, {}};
        for (uint64_t a = 1; a <= max_abundance; ++a) e.rows.emplace_back(a, counts[a - 1]);
        return e;
    }
    uint64_t filter_only_count_hint() const {
        uint64_t h = 0;
        check(scb_index_filter_only_count_hint(h_, nullptr, nullptr, 0, 255, &h));
        return h;
    }
    // api.rs:345-368, 375-385 on a given counter histogram
    FilterOnlySpectrumEstimate estimate_filter_only_spectrum_from_counter_histogram(const ApproxSpectrum &hist,
                                                                                    uint64_t max_abundance,
                                                                                    uint64_t hint) const {
        std::vector<uint64_t> hv, hc, counts(max_abundance ? max_abundance : 1);
        for (auto &r : hist) hv.push_back(r.first), hc.push_back(r.second);
        check(scb_index_filter_only_spectrum(h_, hv.data(), hc.data(), hv.size(), max_abundance, hint,
                                             counts.data()));
        FilterOnlySpectrumEstimate e{"compound_poisson_counter_deconvolution", hint, {}};
        for (uint64_t a = 1; a <= max_abundance; ++a) e.rows.emplace_back(a, counts[a - 1]);
        return e;
    }
    uint64_t filter_only_count_hint_from_counter_histogram(const ApproxSpectrum &hist, uint64_t max_abundance) const {
        std::vector<uint64_t> hv, hc;
        for (auto &r : hist) hv.push_back(r.first), hc.push_back(r.second);
        uint64_t h = 0;
        check(scb_index_filter_only_count_hint(h_, hv.data(), hc.data(), hv.size(), max_abundance, &h));
        return h;
    }
    scb_index *handle() const { return h_; }

    // Calls f(values, counts, cap, &n_rows, &queried, &distinct), growing the
    // row buffers once when the first call reports more rows than fit.
    template <class F>
    static DistinctSpectrum rows_of(F &&f) {
        std::vector<uint64_t> v(1 << 16), c(1 << 16);
        size_t n = 0;
        DistinctSpectrum s;
        check(f(v.data(), c.data(), v.size(), &n, &s.queried, &s.distinct));
        if (n > v.size()) {
            v.resize(n), c.resize(n);
            check(f(v.data(), c.data(), v.size(), &n, &s.queried, &s.distinct));
        }
        for (size_t i = 0; i < n; ++i) s.rows.emplace_back(v[i], c[i]);
        return s;
    }

   private:
    scb_index *h_ = nullptr;
};

// estimate_distinct_kmers over many batches (fastx.rs:255-308, 392-400): one
// device key set across add() calls; the index must outlive the accumulator.
class SpectrumAccumulator {
   public:
    explicit SpectrumAccumulator(const SuperCountingBloomIndex &idx) { check(scb_spectrum_acc_new(idx.handle(), &h_)); }
    ~SpectrumAccumulator() {
        if (h_) scb_spectrum_acc_free(h_);
    }
    SpectrumAccumulator(const SpectrumAccumulator &) = delete;
    SpectrumAccumulator &operator=(const SpectrumAccumulator &) = delete;
    void add(const Records &r) { check(scb_spectrum_acc_add(h_, r.bases.data(), r.offsets.data(), r.size())); }
    void add_device(const void *d_bases, const void *d_offsets, size_t n, uint64_t total_bases) {
        check(scb_spectrum_acc_add_device(h_, static_cast<const uint8_t *>(d_bases),
                                          static_cast<const uint64_t *>(d_offsets), n, total_bases));
    }
    DistinctSpectrum rows() const {
        return SuperCountingBloomIndex::rows_of(
            [&](uint64_t *v, uint64_t *c, size_t cap, size_t *n, uint64_t *q, uint64_t *d) {
                return scb_spectrum_acc_rows(h_, v, c, cap, n, q, d);
            });
    }

   private:
    scb_spectrum_acc *h_ = nullptr;
};

// extension: the serving loop over a frozen index (requests coalesced into
// device launches by worker threads). The index must outlive the server; the
// buffers of a request must stay valid until its ticket is waited for.
using ServeConfig = scb_serve_config;
using Served = scb_served;
class Server {
   public:
    explicit Server(const SuperCountingBloomIndex &idx, const ServeConfig *cfg = nullptr) {
        check(scb_server_new(idx.handle(), cfg, &h_));
    }
    ~Server() {
        if (h_) scb_server_free(h_);
    }
    Server(const Server &) = delete;
    Server &operator=(const Server &) = delete;
    static ServeConfig default_config() {
        ServeConfig c;
        scb_serve_config_default(&c);
        return c;
    }
    // values: counter_bits/8 bytes per position; valid_bits: one bit per position
    uint64_t submit_abundances(const Records &r, void *values, uint32_t *valid_bits) {
        uint64_t t = 0;
        check(scb_server_submit_abundances(h_, r.bases.data(), r.offsets.data(), r.size(), values, valid_bits, &t));
        return t;
    }
    uint64_t submit_spectrum(const Records &r, uint64_t *values, uint64_t *counts, size_t cap) {
        uint64_t t = 0;
        check(scb_server_submit_spectrum(h_, r.bases.data(), r.offsets.data(), r.size(), values, counts, cap, &t));
        return t;
    }
    // The request's outcome; a failed request throws its own error.
    Served wait(uint64_t ticket) {
        Served s{};
        check(scb_server_wait(h_, ticket, &s));
        if (s.status != SCB_OK) throw SuperCountingBloomError(s.status, s.message);
        return s;
    }
    // Any finished request not yet waited for: its ticket (0 when none finished
    // within timeout_us) and outcome; a failed request throws its own error.
    uint64_t wait_any(uint64_t timeout_us, Served &s) {
        uint64_t t = 0;
        s = Served{};
        check(scb_server_wait_any(h_, timeout_us, &t, &s));
        if (t && s.status != SCB_OK) throw SuperCountingBloomError(s.status, s.message);
        return t;
    }

   private:
    scb_server *h_ = nullptr;
};

inline SuperCountingBloomIndex SuperCountingBloomBuilder::into_index() && {
    scb_index *idx = nullptr;
    scb_builder *b = h_;
    h_ = nullptr;
    check(scb_builder_into_index(b, &idx));
    return SuperCountingBloomIndex(idx);
}

// fastx.rs:81-131
struct SuperCountingBloomReport {
    ApproxSpectrum spectrum;
    scb_spectrum_stats stats;
};

inline SuperCountingBloomReport estimate_spectrum_from_fastx_pair(const std::string &index_path,
                                                                  const std::string &query_path,
                                                                  const SuperCountingBloomConfig &cfg) {
    scb_config c = cfg.to_c();
    std::vector<uint64_t> v(1 << 17), n(1 << 17);
    size_t rows = 0;
    SuperCountingBloomReport r{};
    check(scb_estimate_spectrum_from_fastx_pair(index_path.c_str(), query_path.c_str(), &c, v.data(), n.data(),
                                                v.size(), &rows, &r.stats));
    for (size_t i = 0; i < rows && i < v.size(); ++i) r.spectrum.emplace_back(v[i], n[i]);
    return r;
}
inline SuperCountingBloomReport estimate_spectrum_from_fastx(const std::string &path,
                                                             const SuperCountingBloomConfig &cfg) {
    return estimate_spectrum_from_fastx_pair(path, path, cfg);
}
inline scb_stream_stats stream_estimates_from_fastx_pair(const std::string &index_path,
                                                         const std::string &query_path,
                                                         const SuperCountingBloomConfig &cfg) {
    scb_config c = cfg.to_c();
    scb_stream_stats s{};
    check(scb_stream_estimates_from_fastx_pair(index_path.c_str(), query_path.c_str(), &c, &s));
    return s;
}
inline scb_stream_stats stream_estimates_from_fastx(const std::string &path, const SuperCountingBloomConfig &cfg) {
    return stream_estimates_from_fastx_pair(path, path, cfg);
}
inline void write_spectrum(const ApproxSpectrum &rows, const char *path) {  // spectrum.rs:11-31
    std::vector<uint64_t> v, n;
    for (auto &r : rows) v.push_back(r.first), n.push_back(r.second);
    check(scb_write_spectrum(v.data(), n.data(), rows.size(), path));
}
inline uint64_t xorshift_u64(uint64_t x) { return scb_xorshift_u64(x); }
// canonical_kmer of the ACGT window seq[pos, pos+k)
inline uint64_t canonical_kmer(const std::string &seq, size_t k, size_t pos) {
    return scb_canonical_kmer(reinterpret_cast<const uint8_t *>(seq.data()), seq.size(), k, pos);
}
inline int device_count() { return scb_device_count(); }

// summarize_abundances (api.rs:855-891) of per-window estimates
inline SequenceAbundanceSummary summarize_abundances(const std::string &name, size_t sequence_len,
                                                     const std::vector<std::optional<uint64_t>> &windows) {
    std::vector<uint64_t> v(windows.size() ? windows.size() : 1);
    std::vector<uint8_t> ok(v.size());
    for (size_t i = 0; i < windows.size(); ++i) v[i] = windows[i].value_or(0), ok[i] = windows[i].has_value();
    scb_summary s{};
    check(scb_summarize_abundances(v.data(), ok.data(), windows.size(), &s));
    SequenceAbundanceSummary r{name, sequence_len, size_t(s.total_windows), size_t(s.valid_kmers), std::nullopt,
                               std::nullopt};
    if (!std::isnan(s.mean_abundance)) r.mean_abundance = s.mean_abundance;
    if (!std::isnan(s.median_abundance)) r.median_abundance = s.median_abundance;
    return r;
}

// for_each_fastx_record (api.rs:893-927); an exception thrown by visit stops
// the walk and is rethrown here.
inline void for_each_fastx_record(const std::string &path,
                                  const std::function<void(const std::string &, const std::string &)> &visit) {
    struct Ctx {
        const std::function<void(const std::string &, const std::string &)> *fn;
        std::exception_ptr err;
    } ctx{&visit, nullptr};
    auto tramp = [](void *user, const char *name, const uint8_t *seq, size_t len) -> int {
        auto *c = static_cast<Ctx *>(user);
        try {
            (*c->fn)(name, std::string(reinterpret_cast<const char *>(seq), len));
            return 0;
        } catch (...) {
            c->err = std::current_exception();
            return 1;
        }
    };
    int rc = scb_for_each_fastx_record(path.c_str(), tramp, &ctx);
    if (ctx.err) std::rethrow_exception(ctx.err);
    check(rc);
}

// run() of bin/sampled_spectrum.rs:99-148: build from a FASTA/FASTQ file, then
// sample and query its k-mers
inline SampledSpectrum sampled_spectrum_from_fastx(const std::string &path, const SuperCountingBloomConfig &cfg,
                                                   uint64_t sample_size, uint64_t seed) {
    scb_config c = cfg.to_c();
    SampledSpectrum s{};
    // one row per distinct estimate among the sampled k-mers: at most sample_size
    std::vector<SampledRow> rows(std::max<uint64_t>(1, std::min<uint64_t>(sample_size, uint64_t(1) << 22)));
    size_t n = 0;
    check(scb_sampled_spectrum_from_fastx(path.c_str(), &c, sample_size, seed, rows.data(), rows.size(), &n,
                                          &s.stats));
    if (n > rows.size())
        throw SuperCountingBloomError(SCB_ERR_INTERNAL_STATE, "more than 2^22 distinct sampled estimates");
    rows.resize(n);
    s.rows = std::move(rows);
    return s;
}
inline void write_sampled_spectrum(const std::vector<SampledRow> &rows, const char *path) {
    check(scb_write_sampled_spectrum(rows.data(), rows.size(), path));
}

// The filter-only estimators on a histogram with no index (api.rs:933-1033, 375-385)
inline std::vector<uint64_t> filter_only_spectrum_from_histogram(uint64_t n_hashes, const ApproxSpectrum &hist,
                                                                 uint64_t max_abundance, uint64_t hint) {
    std::vector<uint64_t> hv, hc, counts(max_abundance ? max_abundance : 1);
    for (auto &r : hist) hv.push_back(r.first), hc.push_back(r.second);
    check(scb_filter_only_spectrum_from_histogram(n_hashes, hv.data(), hc.data(), hv.size(), max_abundance, hint,
                                                  counts.data()));
    counts.resize(max_abundance);
    return counts;
}
inline uint64_t filter_only_count_hint_from_histogram(uint64_t n_hashes, const ApproxSpectrum &hist,
                                                      uint64_t max_abundance) {
    std::vector<uint64_t> hv, hc;
    for (auto &r : hist) hv.push_back(r.first), hc.push_back(r.second);
    uint64_t h = 0;
    check(scb_filter_only_count_hint_from_histogram(n_hashes, hv.data(), hc.data(), hv.size(), max_abundance, &h));
    return h;
}

}  // namespace scb
