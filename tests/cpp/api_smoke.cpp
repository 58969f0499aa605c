// C++ mirror of the reference's builder test (src/api.rs:1158-1202) through
// scb.hpp; exits 0 when every assertion holds. Needs a GPU.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>

#include "scb.hpp"

#define REQUIRE(c)                                                  \
    do {                                                            \
        if (!(c)) {                                                 \
            std::fprintf(stderr, "assertion failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                               \
        }                                                           \
    } while (0)

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
    std::printf("cpp api ok\n");
    return 0;
}
