// Internal: the opaque handle types behind include/scb.h and the error
// reporting shared by the C ABI translation units (scb_api.cpp, scb_serve.cpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../../include/scb.h"
#include "scb_engine.h"

struct scb_builder {
    scb_config cfg;
    std::unique_ptr<scb::Engine> engine;
    uint64_t inserted_kmers = 0, inserted_smers = 0;
    scb::Method method = scb::kAuto;
    // Sequences of add_sequence not sent to the device yet, as one host batch
    // (inserted_kmers already counts their k-mers); flushed before anything
    // reads the builder's table or s-mer count.
    std::vector<uint8_t> pend_bases;
    std::vector<uint64_t> pend_offsets{0};
    uint64_t pend_kmers = 0;
};

struct scb_index {
    scb_config cfg;
    std::unique_ptr<scb::Engine> engine;
    uint64_t inserted_kmers = 0;
    scb::Method method = scb::kAuto;
};

struct scb_spectrum_acc {
    const scb_index *idx = nullptr;
    scb::KeySetPtr ks;
};

namespace scb_host {
// Records msg as the calling thread's last error (scb_last_error) and returns code.
int fail(int code, const std::string &msg);
int fail(const scb::Status &s);
}  // namespace scb_host
