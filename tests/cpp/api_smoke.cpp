// C++ mirror of the reference's builder test (src/api.rs:1158-1202) through
// scb.hpp, then the batched hot path through the same wrapper; exits 0 when
// every assertion holds. Needs a GPU.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>

#include "scb.hpp"

#define REQUIRE(c)                                                  \
    do {                                                            \
        if (!(c)) {                                                 \
            std::fprintf(stderr, "assertion failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                               \
        }                                                           \
    } while (0)

// Reads of a seeded random genome, some with N runs and lower-case stretches.
static scb::Records make_reads(uint64_t seed, size_t n, size_t len, const std::string &genome) {
    scb::Records r;
    uint64_t x = seed;
    auto next = [&] { return x = x * 6364136223846793005ULL + 1442695040888963407ULL, x >> 33; };
    for (size_t i = 0; i < n; ++i) {
        std::string s = genome.substr(next() % (genome.size() - len), len);
        if (next() % 8 == 0) s.replace(next() % (len - 3), 3, "NNN");
        if (next() % 16 == 0)
            for (size_t j = 0; j < 20; ++j) s[j] = char(s[j] | 0x20);
        r.add(s);
    }
    r.add("");
    r.add("ACGT");
    return r;
}

// The batched hot path through scb.hpp: every method gives the same counters
// and estimates, the batch query equals the per-sequence one, the k-mer map,
// the spectrum, the multi-batch accumulator and the serving loop agree with
// each other, and the sampled spectrum draws the requested sample.
static int hot_path(const std::string &dir) {
    std::string genome;
    uint64_t x = 7;
    for (size_t i = 0; i < 200000; ++i) {
        x = x * 6364136223846793005ULL + 1442695040888963407ULL;
        genome.push_back("ACGT"[(x >> 40) & 3]);
    }
    const scb::Records ins = make_reads(11, 4000, 150, genome), qry = make_reads(12, 3000, 150, genome);
    scb::SuperCountingBloomConfig cfg;
    cfg.counter_slots_exponent = 22, cfg.threads = 1;
    const size_t k = cfg.k;
    std::vector<uint8_t> tables[3];
    std::vector<scb::SuperCountingBloomIndex> idx;
    for (int m = 0; m < 3; ++m) {
        scb::SuperCountingBloomBuilder b(cfg);
        b.set_method(scb::Method(m));
        auto rep = b.add_sequences(ins);
        REQUIRE(rep.records_processed == ins.size() && rep.inserted_kmers > 0);
        tables[m] = b.counter_bytes();
        idx.push_back(std::move(b).into_index());
    }
    REQUIRE(tables[0] == tables[1] && tables[0] == tables[2]);
    REQUIRE(idx[0].counter_bytes() == tables[0]);
    auto est = idx[0].estimate_sequences(qry);
    for (int m = 1; m < 3; ++m) {
        idx[m].set_method(scb::Method(m));
        auto e = idx[m].estimate_sequences(qry);
        REQUIRE(e.values == est.values && e.valid == est.valid);
        REQUIRE(e.report.estimate_checksum == est.report.estimate_checksum);
    }
    REQUIRE(est.report.records_processed == qry.size() && est.report.valid_kmers > 0);
    uint64_t valid = 0;
    for (size_t r = 0; r < qry.size(); ++r) {  // batch == per sequence, and the window count
        std::string s = qry.record(r);
        auto one = idx[0].estimate_sequence_abundances(s);
        for (size_t j = 0; j < one.size(); ++j) {
            REQUIRE(one[j] == est.at(qry.offsets[r] + j));
            valid += one[j].has_value();
        }
        if (r == 5) {
            auto a = scb::summarize_abundances("q", s.size(), one), b = idx[0].summarize_sequence("q", s);
            REQUIRE(a.valid_kmers == b.valid_kmers && a.mean_abundance == b.mean_abundance &&
                    a.median_abundance == b.median_abundance);
        }
    }
    REQUIRE(valid == est.report.valid_kmers);
    // the k-mer map holds each valid window's estimate under its canonical key
    auto km = idx[0].estimate_kmer_map(qry);
    REQUIRE(km.queried == valid);
    for (size_t r = 0; r < qry.size(); ++r) {
        std::string s = qry.record(r);
        for (char &c : s) c = char(c & ~0x20);
        for (size_t j = 0; j + k <= s.size(); ++j) {
            auto v = est.at(qry.offsets[r] + j);
            if (!v) continue;
            auto it = km.map.find(scb::canonical_kmer(s, k, j));
            REQUIRE(it != km.map.end() && it->second == *v);
        }
    }
    // the spectrum is the histogram of the map's estimates
    auto sp = idx[0].spectrum(qry);
    REQUIRE(sp.distinct == km.map.size() && sp.queried == valid);
    std::map<uint64_t, uint64_t> hist;
    for (auto &kv : km.map) ++hist[kv.second];
    REQUIRE(scb::ApproxSpectrum(hist.begin(), hist.end()) == sp.rows);
    // two batches into one accumulator: a key counted once
    scb::Records a, b;
    for (size_t r = 0; r < qry.size(); ++r) (r % 2 ? a : b).add(qry.record(r));
    scb::SpectrumAccumulator acc(idx[0]);
    acc.add(a);
    acc.add(b);
    acc.add(a);
    auto ar = acc.rows();
    const uint64_t qa = idx[0].spectrum(a).queried, qb = idx[0].spectrum(b).queried;
    REQUIRE(qa + qb == sp.queried);
    REQUIRE(ar.rows == sp.rows && ar.distinct == sp.distinct && ar.queried == 2 * qa + qb);
    // the serving loop answers as the direct calls
    {
        scb::Server srv(idx[0]);
        std::vector<uint16_t> vals(qry.total_bases());
        std::vector<uint32_t> bits((qry.total_bases() + 31) / 32);
        std::vector<uint64_t> sv(1 << 16), sc(1 << 16);
        uint64_t t1 = srv.submit_abundances(qry, vals.data(), bits.data());
        uint64_t t2 = srv.submit_spectrum(qry, sv.data(), sc.data(), sv.size());
        auto s1 = srv.wait(t1);
        auto s2 = srv.wait(t2);
        REQUIRE(s1.report.estimate_checksum == est.report.estimate_checksum);
        for (uint64_t p = 0; p < qry.total_bases(); ++p) {
            bool ok = (bits[p / 32] >> (p % 32)) & 1;
            REQUIRE(ok == bool(est.valid[p]) && (!ok || vals[p] == est.values[p]));
        }
        REQUIRE(s2.n_rows == sp.rows.size() && s2.distinct == sp.distinct);
        for (size_t i = 0; i < sp.rows.size(); ++i) REQUIRE(sv[i] == sp.rows[i].first && sc[i] == sp.rows[i].second);
        // wait_any hands out each finished ticket once
        const uint64_t t3 = srv.submit_abundances(qry, vals.data(), bits.data());
        scb::Served s3{};
        uint64_t got = 0;
        for (int i = 0; i < 200 && !got; ++i) got = srv.wait_any(50000, s3);
        REQUIRE(got == t3 && s3.report.estimate_checksum == est.report.estimate_checksum);
        REQUIRE(srv.wait_any(1000, s3) == 0);
    }
    // sampled spectrum: a reservoir of 5000 valid windows, estimates drawn from the map's
    auto ss = idx[0].sampled_spectrum(qry, 5000, 42);
    uint64_t sampled = 0;
    for (auto &row : ss.rows) {
        sampled += row.sampled_kmers;
        REQUIRE(hist.count(row.estimated_count));
    }
    REQUIRE(sampled == 5000 && ss.stats.sampled_kmers == 5000);
    // the filter-only estimators agree with and without an index
    auto h = idx[0].counter_histogram();
    uint64_t hint = scb::filter_only_count_hint_from_histogram(cfg.n_hashes, h, 64);
    REQUIRE(hint == idx[0].filter_only_count_hint_from_counter_histogram(h, 64));
    auto fo = idx[0].estimate_filter_only_spectrum_from_counter_histogram(h, 64, hint);
    auto fo2 = scb::filter_only_spectrum_from_histogram(cfg.n_hashes, h, 64, hint);
    for (uint64_t i = 0; i < 64; ++i) REQUIRE(fo.rows[i].second == fo2[i]);
    // record names through the FASTX walk; a throwing visitor stops it
    std::string fa = dir + "/cpp_names.fa";
    std::ofstream(fa) << ">one x\nAC\nGT\n>two\nTT\n>three\nGG\n";
    std::vector<std::string> names;
    scb::for_each_fastx_record(fa, [&](const std::string &n, const std::string &s) { names.push_back(n + ":" + s); });
    REQUIRE((names == std::vector<std::string>{"one:ACGT", "two:TT", "three:GG"}));
    names.clear();
    try {
        scb::for_each_fastx_record(fa, [&](const std::string &n, const std::string &) {
            names.push_back(n);
            if (n == "two") throw std::runtime_error("stop");
        });
        REQUIRE(false);
    } catch (const std::runtime_error &e) {
        REQUIRE(std::string(e.what()) == "stop" && names.size() == 2);
    }
    return 0;
}

int main(int argc, char **argv) {
    std::string dir = argc > 1 ? argv[1] : "/tmp";
    std::string reads = dir + "/cpp_builder_reads.fa", index = dir + "/cpp_builder_index.scb";
    std::ofstream(reads) << ">r1\nACGTACGTACGT\n>r2\nACGTNNACGTAC\n";
    scb::SuperCountingBloomConfig cfg;
    cfg.k = 5, cfg.m = 3, cfg.s = 4, cfg.n_hashes = 2, cfg.counter_bits = scb::CounterBits::Bits16;
    cfg.counter_slots_exponent = 12, cfg.block_slots_exponent = 5, cfg.threads = 2, cfg.queue = 1;
    scb::SuperCountingBloomBuilder b(cfg);
    REQUIRE(b.threads() == 2);
    b.set_threads(1);
    REQUIRE(b.threads() == 1);
    REQUIRE(b.add_sequence("ACGTACNNACGTAC") == 4);
    auto rep = b.add_fasta(reads);
    REQUIRE(rep.records_processed == 2 && rep.records_indexed == 2 && rep.inserted_kmers == 10);
    REQUIRE(b.inserted_kmers() == 14 && b.inserted_smers() > 0);
    try {
        b.set_threads(0);
        REQUIRE(false);
    } catch (const scb::SuperCountingBloomError &e) {
        REQUIRE(e.kind() == SCB_ERR_INVALID_CONFIG);
    }
    auto idx = std::move(b).into_index();
    REQUIRE(idx.inserted_kmers() == 14);
    auto q = idx.query_fasta(reads);
    REQUIRE(q.records_processed == 2 && q.total_windows == 16 && q.valid_kmers == 10);
    REQUIRE(q.positive_estimates > 0 && q.estimate_checksum > 0);
    idx.save(index);
    auto loaded = scb::SuperCountingBloomIndex::load(index);
    auto q2 = loaded.query_fasta(reads);
    REQUIRE(q2.estimate_checksum == q.estimate_checksum && q2.valid_kmers == q.valid_kmers);
    auto ab = loaded.estimate_sequence_abundances("ACGTACNNACGTAC");
    REQUIRE(ab.size() == 10 && ab[0] && !ab[2] && ab[9]);
    if (hot_path(dir)) return 1;
    std::printf("cpp api ok\n");
    return 0;
}
