// scb.hpp: header-only C++17 mirror of the super-counting-bloom crate API
// (src/lib.rs:15-31) on top of the C ABI in scb.h. Type and method names,
// argument meaning and error kinds follow the Rust crate; failures throw
// scb::SuperCountingBloomError carrying the SCB_ERR_* kind and the reference's
// Display text.
#pragma once

#include <cmath>
#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
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
    AddReport add_fasta(const std::string &path) {
        AddReport r{};
        check(scb_builder_add_fasta(h_, path.c_str(), &r));
        return r;
    }
    inline SuperCountingBloomIndex into_index() &&;
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
    scb_index *handle() const { return h_; }

   private:
    scb_index *h_ = nullptr;
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

}  // namespace scb
