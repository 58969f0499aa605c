// C ABI of libscb.so (declared in include/scb.h): the host side of the
// SuperCountingBloom engine. Mirrors the reference crate's public API
// (super-counting-bloom 0.1.0, src/lib.rs:15-31): config validation and error
// messages, builder/index lifecycle, reports, FASTA/FASTQ pipelines, .scb
// persistence and the f64 host estimators. All counting work runs in the HIP
// engine (scb_kernels.hip); nothing here computes counters on the CPU.

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scb.h"
#include "scb_engine.h"
#include "scb_fastx.h"
#include "scb_handles.h"

using scb::Engine;
using scb::Status;

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace scb_host {
int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}
int fail(const Status &s) { return fail(s.code, s.msg); }
}  // namespace scb_host

namespace {

using scb_host::fail;

int invalid_config(const std::string &m) { return fail(SCB_ERR_INVALID_CONFIG, "invalid config: " + m); }

// config.rs:50-92
int validate(const scb_config &c) {
    if (c.k == 0 || c.k > 32) return invalid_config("k must be in 1..=32");
    if (c.m == 0 || c.m > c.k) return invalid_config("m must be in 1..=k");
    if (c.s == 0 || c.s > c.k || c.s > 32) return invalid_config("s must be in 1..=k and <= 32");
    if (c.k % 2 == 0) return invalid_config("canonical SIMD minimizers require odd k");
    if (c.n_hashes == 0) return invalid_config("n_hashes must be positive");
    if (c.block_slots_exponent >= c.counter_slots_exponent)
        return invalid_config("block_slots_exponent must be smaller than counter_slots_exponent");
    if (c.threads == 0) return invalid_config("threads must be positive");
    if (c.queue == 0) return invalid_config("queue must be positive");
    return SCB_OK;
}

// config.rs:94-107 (checked_pow2) as reached through FrozenFilterLayout::from_config
int check_layout(const scb_config &c) {
    if (c.counter_slots_exponent >= 64) return invalid_config("counter_slots_exponent is too large");
    if (c.block_slots_exponent >= 64) return invalid_config("block_slots_exponent is too large");
    return SCB_OK;
}

int check_bits(const scb_config &c) {
    if (c.counter_bits != 8 && c.counter_bits != 16 && c.counter_bits != 32)
        return invalid_config("counter_bits must be 8, 16 or 32");
    return SCB_OK;
}

int full_check(const scb_config &c) {
    int rc = validate(c);
    if (rc) return rc;
    if ((rc = check_bits(c))) return rc;
    return check_layout(c);
}

size_t online_cpus() {
    unsigned n = std::thread::hardware_concurrency();
    return n ? n : 1;
}

int make_engine(const scb_config &c, std::unique_ptr<Engine> &out) {
    scb::Geometry g;
    Status s = scb::make_geometry(c.k, c.m, c.s, c.n_hashes, c.counter_bits,
                                  c.counter_slots_exponent, c.block_slots_exponent, g);
    if (!s.good()) return fail(s);
    s = Engine::create(g, out);
    if (!s.good()) return fail(s);
    return SCB_OK;
}

bool is_acgt(uint8_t b) {  // api.rs:929-931
    uint8_t u = b & 0xDF;
    return u == 'A' || u == 'C' || u == 'G' || u == 'T';
}

// Maximal ACGT runs of length >= k (api.rs:688-717; fastx.rs:14-19, 234-236):
// returns (#runs, #k-mers).
std::pair<uint64_t, uint64_t> valid_runs(const uint8_t *seq, size_t len, size_t k) {
    uint64_t runs = 0, kmers = 0;
    size_t i = 0;
    while (i < len) {
        while (i < len && !is_acgt(seq[i])) ++i;
        size_t a = i;
        while (i < len && is_acgt(seq[i])) ++i;
        if (i - a >= k && i > a) ++runs, kmers += i - a + 1 - k;
    }
    return {runs, kmers};
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// api.rs:917-927
std::string record_name(const std::string &header, uint64_t idx) {
    size_t e = 0;
    while (e < header.size() && !isspace(static_cast<unsigned char>(header[e]))) ++e;
    if (e == 0) return "record_" + std::to_string(idx);
    return header.substr(0, e);
}

// Host batch of records (concatenated bytes + offsets).
struct HostBatch {
    std::vector<uint8_t> bases;
    std::vector<uint64_t> offsets{0};
    void clear() {
        bases.clear();
        offsets.assign(1, 0);
    }
    void add(const uint8_t *p, size_t n) {
        bases.insert(bases.end(), p, p + n);
        offsets.push_back(bases.size());
    }
    size_t n() const { return offsets.size() - 1; }
};

// Bases per batch of a FASTX file: 256 MiB, or SCB_FILE_BATCH_BYTES (at least
// 4096) from the environment for hosts that should hold less.
size_t file_batch_bytes() {
    const char *e = std::getenv("SCB_FILE_BATCH_BYTES");
    if (e && *e) {
        char *end = nullptr;
        unsigned long long v = std::strtoull(e, &end, 10);
        if (end && *end == 0 && v >= 4096) return size_t(v);
    }
    return size_t(256) << 20;
}

// Walks a FASTX file in batches of records. split=true yields the DnaChunks of
// split_non_actg (fastx.rs:14-19: maximal ACGT runs, those shorter than k
// dropped); split=false yields whole records (keep_non_actg, api.rs:24-27).
// A parser thread fills the next batch while fn processes the current one
// (two batches in flight), so parsing overlaps the device work.
template <class Fn>
int for_each_batch(const char *path, Fn &&fn) {
    scb_fastx::Reader rd;
    int rc = rd.open(path);
    if (rc) return rc;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::unique_ptr<HostBatch>> free_list, full;
    bool parsed_all = false, stop = false;
    int parse_rc = SCB_OK;
    std::string parse_msg;
    for (int i = 0; i < 2; ++i) free_list.emplace_back(new HostBatch);
    const size_t batch_bytes = file_batch_bytes();
    std::thread parser([&] {
        std::unique_ptr<HostBatch> hb;
        auto take_free = [&]() -> bool {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return !free_list.empty() || stop; });
            if (stop) return false;
            hb = std::move(free_list.back());
            free_list.pop_back();
            hb->clear();
            hb->bases.reserve(batch_bytes + (batch_bytes >> 4));
            return true;
        };
        auto publish = [&]() {
            std::lock_guard<std::mutex> lk(mu);
            full.push_back(std::move(hb));
        };
        if (!take_free()) return;
        while (rd.next_seq(hb->bases)) {
            // whole records: the device scan drops every window with a non-ACGT
            // base, so a record gives the k-mers of its ACGT runs (fastx.rs:14-19)
            hb->offsets.push_back(hb->bases.size());
            if (hb->bases.size() >= batch_bytes) {
                publish();
                cv.notify_all();
                if (!take_free()) return;
            }
        }
        int e = rd.error();  // records the message in this thread; hand it over
        std::lock_guard<std::mutex> lk(mu);
        if (e) {
            parse_rc = e;
            size_t n = scb_last_error(nullptr, 0);
            parse_msg.resize(n);
            scb_last_error(&parse_msg[0], n + 1);
        }
        if (hb && hb->n()) full.push_back(std::move(hb));
        parsed_all = true;
        cv.notify_all();
    });
    for (;;) {
        std::unique_ptr<HostBatch> hb;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return !full.empty() || parsed_all; });
            if (full.empty()) break;
            hb = std::move(full.front());
            full.erase(full.begin());
        }
        rc = fn(*hb);
        {
            std::lock_guard<std::mutex> lk(mu);
            free_list.push_back(std::move(hb));
            if (rc) stop = true;
        }
        cv.notify_all();
        if (rc) break;
    }
    parser.join();
    if (rc) return rc;
    if (parse_rc) return fail(parse_rc, parse_msg);
    return SCB_OK;
}

}  // namespace

namespace {

uint64_t sat_add(uint64_t a, uint64_t b) { return a + b < a ? UINT64_MAX : a + b; }

// Records with at least one valid k-mer (api.rs:638-640), counted by host
// threads over record ranges.
uint64_t count_indexed(const uint8_t *bases, const uint64_t *offsets, size_t n, size_t k) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = offsets[n] - offsets[0] < (uint64_t(1) << 22) ? 1u : std::min(16u, hw);
    std::vector<uint64_t> part(nt, 0);
    auto run = [&](unsigned t) {
        const size_t i0 = n * t / nt, i1 = n * (t + 1) / nt;
        uint64_t c = 0;
        for (size_t i = i0; i < i1; ++i)
            c += valid_runs(bases + offsets[i], size_t(offsets[i + 1] - offsets[i]), k).second > 0;
        part[t] = c;
    };
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < nt; ++t) ts.emplace_back(run, t);
    run(0);
    for (auto &t : ts) t.join();
    uint64_t c = 0;
    for (uint64_t v : part) c += v;
    return c;
}

// ACGT runs of at least k bases over all records (fastx.rs:14-19 chunks),
// counted by host threads over record ranges.
uint64_t count_chunks(const uint8_t *bases, const uint64_t *offsets, size_t n, size_t k) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = offsets[n] - offsets[0] < (uint64_t(1) << 22) ? 1u : std::min(16u, hw);
    std::vector<uint64_t> part(nt, 0);
    auto run = [&](unsigned t) {
        const size_t i0 = n * t / nt, i1 = n * (t + 1) / nt;
        uint64_t c = 0;
        for (size_t i = i0; i < i1; ++i)
            c += valid_runs(bases + offsets[i], size_t(offsets[i + 1] - offsets[i]), k).first;
        part[t] = c;
    };
    std::vector<std::thread> ts;
    for (unsigned t = 1; t < nt; ++t) ts.emplace_back(run, t);
    run(0);
    for (auto &t : ts) t.join();
    uint64_t c = 0;
    for (uint64_t v : part) c += v;
    return c;
}

// add_sequence calls are gathered into one host batch of up to this many bases
// (or records) before it goes to the device: a call then costs a host copy and
// a host count of its k-mers, not a device round trip.
constexpr size_t kDeferBases = size_t(64) << 20, kDeferRecords = size_t(1) << 20;

// Sends the gathered add_sequence records to the device. The pending batch is
// dropped either way: after a device error the builder's table is undefined.
int builder_flush(scb_builder *b) {
    if (b->pend_offsets.size() <= 1) return SCB_OK;
    scb::InsertResult r;
    Status s = b->engine->insert_host(b->pend_bases.data(), b->pend_offsets.data(), b->pend_offsets.size() - 1,
                                      b->method, r);
    const uint64_t want = b->pend_kmers;
    b->pend_bases.clear();
    b->pend_offsets.assign(1, 0);
    b->pend_kmers = 0;
    if (!s.good()) return fail(s);
    if (r.inserted_kmers != want)
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: the device inserted " +
                                                std::to_string(r.inserted_kmers) + " k-mers of " +
                                                std::to_string(want) + " counted on the host");
    b->inserted_smers = sat_add(b->inserted_smers, r.inserted_smers);
    return SCB_OK;
}


int builder_insert_host(scb_builder *b, const uint8_t *bases, const uint64_t *offsets, size_t n,
                        scb_add_report *rep) {
    if (int rc = builder_flush(b)) return rc;
    scb::InsertResult r;
    // the host count overlaps the device insert
    uint64_t indexed = 0;
    std::thread counter;
    if (rep) counter = std::thread([&] { indexed = count_indexed(bases, offsets, n, size_t(b->cfg.k)); });
    Status s = b->engine->insert_host(bases, offsets, n, b->method, r);
    if (counter.joinable()) counter.join();
    if (!s.good()) return fail(s);
    if (rep) {
        rep->records_processed += n;
        rep->records_indexed += indexed;
        rep->inserted_kmers += r.inserted_kmers;
        rep->inserted_smers += r.inserted_smers;
    }
    b->inserted_kmers = sat_add(b->inserted_kmers, r.inserted_kmers);  // api.rs:135-136
    b->inserted_smers = sat_add(b->inserted_smers, r.inserted_smers);
    return SCB_OK;
}

int check_offsets(const uint64_t *offsets, size_t n) {
    if (!offsets) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: offsets is NULL");
    if (offsets[0] != 0) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: offsets[0] must be 0");
    for (size_t i = 0; i < n; ++i)
        if (offsets[i + 1] < offsets[i])
            return fail(SCB_ERR_INTERNAL_STATE, "internal state error: offsets must be non-decreasing");
    return SCB_OK;
}

// ---- .scb header (api.rs:388-414, 466-480, 515-529) --------------------------
struct LayoutRec {
    uint64_t shard_count, subblocks, block_slots, shard_mask;
    uint32_t shard_shift;
    uint64_t block_slots_mask, nb_blocks, block_mask, n_hashes;
    bool operator==(const LayoutRec &o) const {
        return shard_count == o.shard_count && subblocks == o.subblocks &&
               block_slots == o.block_slots && shard_mask == o.shard_mask &&
               shard_shift == o.shard_shift && block_slots_mask == o.block_slots_mask &&
               nb_blocks == o.nb_blocks && block_mask == o.block_mask && n_hashes == o.n_hashes;
    }
};

// FrozenFilterLayout::from_config (filter.rs:90-116)
LayoutRec layout_of(const scb_config &c) {
    LayoutRec l;
    uint64_t slots = uint64_t(1) << c.counter_slots_exponent;
    l.block_slots = uint64_t(1) << c.block_slots_exponent;
    l.nb_blocks = slots / l.block_slots;
    l.shard_count = std::min<uint64_t>(l.nb_blocks, 1024);
    l.subblocks = l.nb_blocks / l.shard_count;
    l.shard_mask = l.shard_count - 1;
    l.shard_shift = uint32_t(__builtin_ctzll(l.shard_count));
    l.block_slots_mask = l.block_slots - 1;
    l.block_mask = l.nb_blocks - 1;
    l.n_hashes = c.n_hashes;
    return l;
}

std::string layout_debug(const LayoutRec &l) {
    char buf[512];
    snprintf(buf, sizeof buf,
             "FrozenFilterLayout { shard_count: %llu, subblocks: %llu, block_slots: %llu, "
             "shard_mask: %llu, shard_shift: %u, block_slots_mask: %llu, nb_blocks: %llu, "
             "block_mask: %llu, n_hashes: %llu }",
             (unsigned long long)l.shard_count, (unsigned long long)l.subblocks,
             (unsigned long long)l.block_slots, (unsigned long long)l.shard_mask, l.shard_shift,
             (unsigned long long)l.block_slots_mask, (unsigned long long)l.nb_blocks,
             (unsigned long long)l.block_mask, (unsigned long long)l.n_hashes);
    return buf;
}

int io_error(const std::string &path, const std::string &msg) {
    return fail(SCB_ERR_IO, "I/O error for '" + path + "': " + msg);
}

std::string os_error(int e) { return std::string(strerror(e)) + " (os error " + std::to_string(e) + ")"; }

// ---- f64 estimators (api.rs:958-1033) -----------------------------------------
uint64_t f64_as_u64(double x) {  // Rust `as u64`: saturating, NaN -> 0
    if (std::isnan(x) || x <= 0.0) return 0;
    if (x >= 18446744073709551616.0) return UINT64_MAX;
    return uint64_t(x);
}

std::vector<double> deconvolved_counts(const uint64_t *values, const uint64_t *counts, size_t rows,
                                       uint64_t n_hashes, uint64_t max_abundance_u) {
    size_t max_abundance = size_t(max_abundance_u);
    if (max_abundance == 0) return {};
    uint64_t total = 0;
    for (size_t i = 0; i < rows; ++i) total += counts[i];
    if (total == 0) return std::vector<double>(max_abundance, 0.0);
    std::vector<double> prob(max_abundance + 1, 0.0);
    for (size_t i = 0; i < rows; ++i)
        if (values[i] <= max_abundance) prob[size_t(values[i])] += double(counts[i]) / double(total);
    double p0 = prob[0];
    if (p0 <= 0.0) return {};
    std::vector<double> norm(max_abundance + 1, 0.0);
    norm[0] = 1.0;
    for (size_t a = 1; a <= max_abundance; ++a) norm[a] = prob[a] / p0;
    std::vector<double> logc(max_abundance + 1, 0.0);
    for (size_t a = 1; a <= max_abundance; ++a) {
        double c = double(a) * norm[a];
        for (size_t p = 1; p < a; ++p) c -= double(p) * logc[p] * norm[a - p];
        logc[a] = c / double(a);
    }
    double hc = double(std::max<uint64_t>(n_hashes, 1));
    std::vector<double> out;
    out.reserve(max_abundance);
    for (size_t a = 1; a <= max_abundance; ++a) out.push_back(std::fmax(double(total) * logc[a] / hc, 0.0));
    return out;
}

int index_histogram(const scb_index *idx, std::vector<std::pair<uint64_t, uint64_t>> &rows) {
    Status s = idx->engine->counter_histogram(rows);
    if (!s.good()) return fail(s);
    return SCB_OK;
}

}  // namespace

extern "C" {

size_t scb_last_error(char *buf, size_t cap) {
    if (buf && cap) {
        size_t n = std::min(cap - 1, g_last_error.size());
        memcpy(buf, g_last_error.data(), n);
        buf[n] = 0;
    }
    return g_last_error.size();
}

void scb_config_default(scb_config *cfg) {
    memset(cfg, 0, sizeof *cfg);
    cfg->k = 31, cfg->m = 21, cfg->s = 25, cfg->n_hashes = 4;
    cfg->counter_bits = 16, cfg->counter_slots_exponent = 30, cfg->block_slots_exponent = 9;
    cfg->threads = online_cpus();
    cfg->queue = 4096;
}

int scb_config_validate(const scb_config *cfg) {
    int rc = validate(*cfg);
    if (rc) return rc;
    if ((rc = check_bits(*cfg))) return rc;
    return SCB_OK;
}

int scb_device_count(void) { return scb::device_count(); }

void scb_last_stage_times(double *ms) {
    scb::StageTimes t = scb::last_stage_times();
    ms[0] = t.scan_ms, ms[1] = t.plan_scatter_ms, ms[2] = t.bin_ms, ms[3] = t.total_ms;
}

uint64_t scb_xorshift_u64(uint64_t x) {  // filter.rs:629-635
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint64_t scb_canonical_kmer(const uint8_t *seq, size_t len, size_t klen, size_t pos) {
    // filter.rs:622-627 with packed-seq's 2-bit code (A=0 C=1 T=2 G=3)
    if (klen == 0 || klen > 32 || pos + klen > len) return 0;
    uint64_t fw = 0, rc = 0;
    for (size_t i = 0; i < klen; ++i) {
        uint64_t c = (seq[pos + i] >> 1) & 3;
        fw |= c << (2 * i);
        rc |= (c ^ 2) << (2 * (klen - 1 - i));
    }
    return std::min(fw, rc);
}

// ---- builder ----------------------------------------------------------------
int scb_builder_new(const scb_config *cfg, scb_builder **out) {
    *out = nullptr;
    int rc = full_check(*cfg);
    if (rc) return rc;
    std::unique_ptr<scb_builder> b(new scb_builder);
    b->cfg = *cfg;
    if ((rc = make_engine(*cfg, b->engine))) return rc;
    *out = b.release();
    return SCB_OK;
}

void scb_builder_free(scb_builder *b) { delete b; }

// api.rs:123-138: returns the k-mers of the sequence (valid windows), counted on
// the host; the insert itself joins the builder's pending batch.
int scb_builder_add_sequence(scb_builder *b, const uint8_t *seq, size_t len, uint64_t *inserted_kmers) {
    const uint64_t kmers = valid_runs(seq, len, size_t(b->cfg.k)).second;
    if (len) {
        b->pend_bases.insert(b->pend_bases.end(), seq, seq + len);
        b->pend_offsets.push_back(b->pend_bases.size());
        b->pend_kmers += kmers;
    }
    b->inserted_kmers = sat_add(b->inserted_kmers, kmers);  // api.rs:135-136
    if (inserted_kmers) *inserted_kmers = kmers;
    if (b->pend_bases.size() >= kDeferBases || b->pend_offsets.size() > kDeferRecords) return builder_flush(b);
    return SCB_OK;
}

int scb_builder_add_sequences(scb_builder *b, const uint8_t *bases, const uint64_t *offsets, size_t n,
                              scb_add_report *report) {
    int rc = check_offsets(offsets, n);
    if (rc) return rc;
    scb_add_report rep{};
    double t0 = now_s();
    if ((rc = builder_insert_host(b, bases, offsets, n, &rep))) return rc;
    rep.add_seconds = now_s() - t0;
    if (report) *report = rep;
    return SCB_OK;
}

int scb_builder_add_sequences_device(scb_builder *b, const uint8_t *d_bases, const uint64_t *d_offsets,
                                     size_t n, uint64_t total_bases, scb_add_report *report) {
    if (total_bases >= scb::kMaxBatchBases)
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: device batch exceeds 2^32 bases");
    if (int rc = builder_flush(b)) return rc;
    double t0 = now_s();
    scb::DeviceBatch db{d_bases, d_offsets, n, total_bases};
    scb::InsertResult r;
    Status s = b->engine->insert(db, b->method, r);
    if (!s.good()) return fail(s);
    b->inserted_kmers = sat_add(b->inserted_kmers, r.inserted_kmers);
    b->inserted_smers = sat_add(b->inserted_smers, r.inserted_smers);
    if (report) {
        *report = scb_add_report{};
        report->inserted_kmers = r.inserted_kmers;
        report->inserted_smers = r.inserted_smers;
        report->add_seconds = now_s() - t0;
    }
    return SCB_OK;
}

// api.rs:140-158, 621-674
int scb_builder_add_fasta(scb_builder *b, const char *path, scb_add_report *report) {
    int rc = validate(b->cfg);
    if (rc) return rc;
    scb_add_report rep{};
    double t0 = now_s();
    rc = for_each_batch(path, [&](HostBatch &hb) -> int {
        return builder_insert_host(b, hb.bases.data(), hb.offsets.data(), hb.n(), &rep);
    });
    if (rc) return rc;
    rep.add_seconds = now_s() - t0;
    if (report) *report = rep;
    return SCB_OK;
}

int scb_builder_into_index(scb_builder *b, scb_index **out) {
    *out = nullptr;
    std::unique_ptr<scb_builder> owned(b);
    if (int rc = builder_flush(b)) return rc;
    std::unique_ptr<scb_index> idx(new scb_index);
    idx->cfg = b->cfg;
    idx->engine = std::move(b->engine);
    idx->inserted_kmers = b->inserted_kmers;
    idx->method = b->method;
    if (!idx->engine)
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: mutable filter is still shared while freezing");
    *out = idx.release();
    return SCB_OK;
}

uint64_t scb_builder_inserted_kmers(const scb_builder *b) { return b->inserted_kmers; }
uint64_t scb_builder_inserted_smers(const scb_builder *b) {
    (void)builder_flush(const_cast<scb_builder *>(b));  // an error stays in scb_last_error
    return b->inserted_smers;
}

int scb_builder_set_threads(scb_builder *b, size_t threads) {  // api.rs:187-195
    if (threads == 0) return invalid_config("threads must be positive");
    b->cfg.threads = threads;
    return SCB_OK;
}
void scb_builder_clear_threads(scb_builder *b) { b->cfg.threads = online_cpus(); }
size_t scb_builder_threads(const scb_builder *b) { return size_t(b->cfg.threads); }
void scb_builder_config(const scb_builder *b, scb_config *out) { *out = b->cfg; }

int scb_builder_set_method(scb_builder *b, int method) {
    if (method != SCB_METHOD_BATCHED && method != SCB_METHOD_PLAIN && method != SCB_METHOD_AUTO)
        return invalid_config("method must be SCB_METHOD_BATCHED, SCB_METHOD_PLAIN or SCB_METHOD_AUTO");
    b->method = scb::Method(method);
    return SCB_OK;
}

namespace {
scb::Tuning from_c(const scb_tuning &c) {
    scb::Tuning t;
    t.insert_bin_log = c.insert_bin_log, t.query_bin_log = c.query_bin_log, t.spectrum_bin_log = c.spectrum_bin_log;
    t.insert_item_records = c.insert_item_records, t.query_item_records = c.query_item_records;
    t.spectrum_item_records = c.spectrum_item_records, t.use_graph = c.use_graph;
    return t;
}
void to_c(const scb::Tuning &t, scb_tuning *c) {
    *c = scb_tuning{};
    c->insert_bin_log = t.insert_bin_log, c->query_bin_log = t.query_bin_log, c->spectrum_bin_log = t.spectrum_bin_log;
    c->insert_item_records = t.insert_item_records, c->query_item_records = t.query_item_records;
    c->spectrum_item_records = t.spectrum_item_records, c->use_graph = t.use_graph;
}
}  // namespace

int scb_tuning_cache_load(const char *path) {
    if (!path) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null path");
    Status s = scb::load_tuning_cache(path);
    return s.good() ? SCB_OK : fail(s);
}

int scb_builder_set_tuning(scb_builder *b, const scb_tuning *t) {
    if (!t) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null tuning");
    Status s = b->engine->set_tuning(from_c(*t));
    return s.good() ? SCB_OK : fail(s);
}

void scb_builder_tuning(const scb_builder *b, scb_tuning *out) { to_c(b->engine->tuning(), out); }

int scb_index_set_tuning(scb_index *idx, const scb_tuning *t) {
    if (!t) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null tuning");
    Status s = idx->engine->set_tuning(from_c(*t));
    return s.good() ? SCB_OK : fail(s);
}

void scb_index_tuning(const scb_index *idx, scb_tuning *out) { to_c(idx->engine->tuning(), out); }

int scb_builder_counters(const scb_builder *b, void *out, size_t bytes) {
    if (int rc = builder_flush(const_cast<scb_builder *>(b))) return rc;
    Status s = b->engine->download(out, bytes);
    return s.good() ? SCB_OK : fail(s);
}

// ---- index -----------------------------------------------------------------
// api.rs:211-223, 738-779 with fastx.rs:206-253, 378-390
int scb_index_build_from_fastx(const char *path, const scb_config *cfg, scb_index **out,
                               scb_build_report *report) {
    *out = nullptr;
    int rc = full_check(*cfg);
    if (rc) return rc;
    double t0 = now_s();
    std::unique_ptr<scb_index> idx(new scb_index);
    idx->cfg = *cfg;
    if ((rc = make_engine(*cfg, idx->engine))) return rc;
    scb_build_report rep{};
    rc = for_each_batch(path, [&](HostBatch &hb) -> int {
        scb::InsertResult r;
        // the chunk count overlaps the device insert
        uint64_t chunks = 0;
        std::thread counter([&] { chunks = count_chunks(hb.bases.data(), hb.offsets.data(), hb.n(), size_t(cfg->k)); });
        Status s = idx->engine->insert_host(hb.bases.data(), hb.offsets.data(), hb.n(), idx->method, r);
        counter.join();
        if (!s.good()) return fail(s);
        rep.chunks += chunks;
        rep.indexed_kmers += r.inserted_kmers;
        rep.inserted_smers += r.inserted_smers;
        return SCB_OK;
    });
    if (rc) return rc;
    idx->inserted_kmers = rep.indexed_kmers;
    rep.counter_slots = uint64_t(1) << cfg->counter_slots_exponent;
    rep.blocks = rep.counter_slots >> cfg->block_slots_exponent;
    rep.build_seconds = now_s() - t0;
    if (report) *report = rep;
    *out = idx.release();
    return SCB_OK;
}

void scb_index_free(scb_index *idx) { delete idx; }
void scb_index_config(const scb_index *idx, scb_config *out) { *out = idx->cfg; }
uint64_t scb_index_inserted_kmers(const scb_index *idx) { return idx->inserted_kmers; }

int scb_index_set_method(scb_index *idx, int method) {
    if (method != SCB_METHOD_BATCHED && method != SCB_METHOD_PLAIN && method != SCB_METHOD_AUTO)
        return invalid_config("method must be SCB_METHOD_BATCHED, SCB_METHOD_PLAIN or SCB_METHOD_AUTO");
    idx->method = scb::Method(method);
    return SCB_OK;
}

int scb_index_counters(const scb_index *idx, void *out, size_t bytes) {
    Status s = idx->engine->download(out, bytes);
    return s.good() ? SCB_OK : fail(s);
}

// api.rs:233-240, 388-414
int scb_index_save(const scb_index *idx, const char *path) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return io_error(path, os_error(errno));
    const scb_config &c = idx->cfg;
    LayoutRec l = layout_of(c);
    std::vector<uint8_t> hdr;
    auto put = [&](const void *p, size_t n) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        hdr.insert(hdr.end(), b, b + n);
    };
    auto u64 = [&](uint64_t v) { put(&v, 8); };
    auto u32 = [&](uint32_t v) { put(&v, 4); };
    auto u8 = [&](uint8_t v) { put(&v, 1); };
    put("SCBIDX\0\0", 8);
    u32(2);
    u64(c.k), u64(c.m), u64(c.s), u64(c.n_hashes);
    u8(c.counter_bits), u8(c.counter_slots_exponent), u8(c.block_slots_exponent);
    u64(c.threads), u64(c.queue);
    u64(idx->inserted_kmers);
    u64(l.shard_count), u64(l.subblocks), u64(l.block_slots), u64(l.shard_mask);
    u32(l.shard_shift);
    u64(l.block_slots_mask), u64(l.nb_blocks), u64(l.block_mask), u64(l.n_hashes);
    bool ok = fwrite(hdr.data(), 1, hdr.size(), fp) == hdr.size();
    uint64_t bytes = idx->engine->geometry().table_bytes();
    std::vector<uint8_t> buf;
    try {
        buf.resize(size_t(bytes));
    } catch (...) {
        fclose(fp);
        return fail(SCB_ERR_IO, "I/O error for '" + std::string(path) + "': out of host memory");
    }
    Status s = idx->engine->download(buf.data(), bytes);  // table is already in .scb order
    if (!s.good()) {
        fclose(fp);
        return fail(s);
    }
    ok = ok && fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
    ok = (fclose(fp) == 0) && ok;
    if (!ok) return io_error(path, os_error(errno ? errno : EIO));
    return SCB_OK;
}

// api.rs:242-248, 416-464, 482-543, 561-608
int scb_index_load(const char *path, scb_index **out) {
    *out = nullptr;
    FILE *fp = fopen(path, "rb");
    if (!fp) return io_error(path, os_error(errno));
    std::unique_ptr<FILE, int (*)(FILE *)> guard(fp, fclose);
    auto rd = [&](void *p, size_t n) { return fread(p, 1, n, fp) == n; };
    auto eof = [&]() { return io_error(path, "failed to fill whole buffer"); };
    auto bad = [&](const std::string &m) { return fail(SCB_ERR_INVALID_INDEX_FORMAT, "invalid index format: " + m); };
    char magic[8];
    if (!rd(magic, 8)) return eof();
    if (memcmp(magic, "SCBIDX\0\0", 8) != 0) return bad("bad SuperCountingBloom index magic");
    uint32_t version;
    if (!rd(&version, 4)) return eof();
    if (version != 2) return bad("unsupported index format version " + std::to_string(version));
    scb_config c{};
    uint8_t bits;
    if (!rd(&c.k, 8) || !rd(&c.m, 8) || !rd(&c.s, 8) || !rd(&c.n_hashes, 8) || !rd(&bits, 1)) return eof();
    if (bits != 8 && bits != 16 && bits != 32) return bad("unsupported counter width " + std::to_string(bits));
    c.counter_bits = bits;
    if (!rd(&c.counter_slots_exponent, 1) || !rd(&c.block_slots_exponent, 1) || !rd(&c.threads, 8) ||
        !rd(&c.queue, 8))
        return eof();
    int rc = validate(c);
    if (rc) return rc;
    uint64_t inserted;
    if (!rd(&inserted, 8)) return eof();
    LayoutRec st{};
    if (!rd(&st.shard_count, 8) || !rd(&st.subblocks, 8) || !rd(&st.block_slots, 8) ||
        !rd(&st.shard_mask, 8) || !rd(&st.shard_shift, 4) || !rd(&st.block_slots_mask, 8) ||
        !rd(&st.nb_blocks, 8) || !rd(&st.block_mask, 8) || !rd(&st.n_hashes, 8))
        return eof();
    if ((rc = check_layout(c))) return rc;
    LayoutRec ex = layout_of(c);
    if (!(st == ex))
        return bad("stored filter layout does not match config: stored " + layout_debug(st) +
                   ", expected " + layout_debug(ex));
    std::unique_ptr<scb_index> idx(new scb_index);
    idx->cfg = c;
    idx->inserted_kmers = inserted;
    uint64_t bytes = (uint64_t(1) << c.counter_slots_exponent) * (bits / 8);
    std::vector<uint8_t> buf;
    try {
        buf.resize(size_t(bytes));
    } catch (...) {
        return io_error(path, "out of host memory");
    }
    if (!rd(buf.data(), buf.size())) return eof();
    int extra = fgetc(fp);
    if (extra != EOF) return bad("trailing bytes after serialized index");
    if ((rc = make_engine(c, idx->engine))) return rc;
    Status s = idx->engine->upload(buf.data(), bytes);
    if (!s.good()) return fail(s);
    *out = idx.release();
    return SCB_OK;
}

int scb_index_estimate_sequence_abundances(const scb_index *idx, const uint8_t *seq, size_t len,
                                           uint64_t *values, uint8_t *valid, size_t *n_windows) {
    size_t k = size_t(idx->cfg.k);
    if (n_windows) *n_windows = len < k ? 0 : len + 1 - k;
    if (len < k) return SCB_OK;  // api.rs:804-806
    uint64_t offs[2] = {0, len};
    std::vector<uint64_t> v(len);
    std::vector<uint8_t> ok(len);
    scb::QueryResult r;
    Status s = idx->engine->query_host(seq, offs, 1, idx->method, v.data(), ok.data(), r);
    if (!s.good()) return fail(s);
    size_t nw = len + 1 - k;
    if (values) std::copy(v.begin(), v.begin() + nw, values);
    if (valid) std::copy(ok.begin(), ok.begin() + nw, valid);
    return SCB_OK;
}

int scb_index_estimate_sequences(const scb_index *idx, const uint8_t *bases, const uint64_t *offsets,
                                 size_t n, uint64_t *values, uint8_t *valid, scb_query_report *report) {
    int rc = check_offsets(offsets, n);
    if (rc) return rc;
    scb::QueryResult r;
    Status s = idx->engine->query_host(bases, offsets, n, idx->method, values, valid, r);
    if (!s.good()) return fail(s);
    if (report) {
        *report = scb_query_report{};
        report->records_processed = n;
        size_t k = size_t(idx->cfg.k);
        for (size_t i = 0; i < n; ++i) {
            uint64_t len = offsets[i + 1] - offsets[i];
            if (len >= k) report->total_windows += len + 1 - k;
        }
        report->valid_kmers = r.queried;
        report->positive_estimates = r.positive;
        report->estimate_checksum = r.checksum;
    }
    return SCB_OK;
}

int scb_index_estimate_sequences_device(const scb_index *idx, const uint8_t *d_bases,
                                        const uint64_t *d_offsets, size_t n, uint64_t total_bases,
                                        void *d_values, uint32_t *d_valid_bits,
                                        scb_query_report *report) {
    if (total_bases >= scb::kMaxBatchBases)
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: device batch exceeds 2^32 bases");
    scb::DeviceBatch db{d_bases, d_offsets, n, total_bases};
    scb::QueryResult r;
    Status s = idx->engine->query(db, idx->method, d_values, d_valid_bits, r);
    if (!s.good()) return fail(s);
    if (report) {
        *report = scb_query_report{};
        report->valid_kmers = r.queried;
        report->positive_estimates = r.positive;
        report->estimate_checksum = r.checksum;
    }
    return SCB_OK;
}

// No counterpart in the reference: scb_index_estimate_sequences_device for
// callers that cannot allocate device memory.
int scb_index_estimate_sequences_device_noalloc(const scb_index *idx, const uint8_t *d_bases,
                                                const uint64_t *d_offsets, size_t n, uint64_t total_bases,
                                                void *d_values, uint32_t *d_valid_bits,
                                                scb_query_report *report) {
    if (!idx) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null index");
    scb::DeviceBatch db{d_bases, d_offsets, n, total_bases};
    scb::QueryResult r;
    Status s = idx->engine->query_noalloc(db, idx->method, d_values, d_valid_bits, r);
    if (!s.good()) return fail(s);
    if (report) {
        *report = scb_query_report{};
        report->valid_kmers = r.queried;
        report->positive_estimates = r.positive;
        report->estimate_checksum = r.checksum;
    }
    return SCB_OK;
}

// api.rs:855-891 (summarize_abundances): mean of the valid values as u64 sum /
// count, median of the sorted values (mean of the middle two for an even count).
int scb_summarize_abundances(const uint64_t *values, const uint8_t *valid, size_t n, scb_summary *out) {
    if (!out || (n && (!values || !valid))) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null argument");
    std::vector<uint64_t> vals;
    for (size_t i = 0; i < n; ++i)
        if (valid[i]) vals.push_back(values[i]);
    out->total_windows = n;
    out->valid_kmers = vals.size();
    if (vals.empty()) {
        out->mean_abundance = NAN;
        out->median_abundance = NAN;
        return SCB_OK;
    }
    uint64_t sum = 0;
    for (uint64_t x : vals) sum += x;  // wrapping, as u64 sum in a release build
    out->mean_abundance = double(sum) / double(vals.size());
    std::sort(vals.begin(), vals.end());
    size_t mid = vals.size() / 2;
    out->median_abundance =
        vals.size() % 2 == 0 ? (double(vals[mid - 1]) + double(vals[mid])) / 2.0 : double(vals[mid]);
    return SCB_OK;
}

// api.rs:264-271
int scb_index_summarize_sequence(const scb_index *idx, const uint8_t *seq, size_t len, scb_summary *out) {
    size_t k = size_t(idx->cfg.k);
    size_t nw = len < k ? 0 : len + 1 - k;
    std::vector<uint64_t> v(nw);
    std::vector<uint8_t> ok(nw);
    size_t got = 0;
    int rc = scb_index_estimate_sequence_abundances(idx, seq, len, v.data(), ok.data(), &got);
    if (rc) return rc;
    rc = scb_summarize_abundances(v.data(), ok.data(), nw, out);
    out->sequence_len = len;
    return rc;
}

// api.rs:273-327
int scb_index_query_fasta(const scb_index *idx, const char *path, scb_query_report *report) {
    scb_query_report rep{};
    size_t k = size_t(idx->cfg.k);
    int rc = for_each_batch(path, [&](HostBatch &hb) -> int {
        scb::QueryResult r;
        Status s = idx->engine->query_host(hb.bases.data(), hb.offsets.data(), hb.n(), idx->method,
                                           nullptr, nullptr, r);
        if (!s.good()) return fail(s);
        rep.records_processed += hb.n();
        for (size_t i = 0; i < hb.n(); ++i) {
            uint64_t len = hb.offsets[i + 1] - hb.offsets[i];
            if (len >= k) rep.total_windows += len + 1 - k;
        }
        rep.valid_kmers += r.queried;
        rep.positive_estimates += r.positive;
        rep.estimate_checksum += r.checksum;
        return SCB_OK;
    });
    if (rc) return rc;
    if (report) *report = rep;
    return SCB_OK;
}

int scb_index_counter_histogram(const scb_index *idx, uint64_t *values, uint64_t *counts, size_t cap,
                                size_t *n_rows) {
    std::vector<std::pair<uint64_t, uint64_t>> rows;
    int rc = index_histogram(idx, rows);
    if (rc) return rc;
    for (size_t i = 0; i < rows.size() && i < cap; ++i) values[i] = rows[i].first, counts[i] = rows[i].second;
    if (n_rows) *n_rows = rows.size();
    return SCB_OK;
}

int scb_index_spectrum(const scb_index *idx, const uint8_t *bases, const uint64_t *offsets, size_t n,
                       uint64_t *values, uint64_t *counts, size_t cap, size_t *n_rows,
                       uint64_t *queried, uint64_t *distinct) {
    int rc = check_offsets(offsets, n);
    if (rc) return rc;
    scb::SpectrumResult r;
    Status s = idx->engine->spectrum_host(bases, offsets, n, r, idx->method);
    if (!s.good()) return fail(s);
    for (size_t i = 0; i < r.rows.size() && i < cap; ++i) values[i] = r.rows[i].first, counts[i] = r.rows[i].second;
    if (n_rows) *n_rows = r.rows.size();
    if (queried) *queried = r.queried;
    if (distinct) *distinct = r.distinct;
    return SCB_OK;
}

int scb_index_spectrum_device(const scb_index *idx, const uint8_t *d_bases, const uint64_t *d_offsets,
                              size_t n, uint64_t total_bases, uint64_t *values, uint64_t *counts,
                              size_t cap, size_t *n_rows, uint64_t *queried, uint64_t *distinct) {
    scb::DeviceBatch db{d_bases, d_offsets, n, total_bases};
    scb::SpectrumResult r;
    Status s = idx->engine->spectrum(db, r, idx->method);
    if (!s.good()) return fail(s);
    for (size_t i = 0; i < r.rows.size() && i < cap; ++i) values[i] = r.rows[i].first, counts[i] = r.rows[i].second;
    if (n_rows) *n_rows = r.rows.size();
    if (queried) *queried = r.queried;
    if (distinct) *distinct = r.distinct;
    return SCB_OK;
}

// filter.rs:302-335 (estimate_packed_sequence) over the DnaChunks of n records:
// every distinct canonical k-mer with its estimate, sorted by key.
int scb_index_estimate_kmer_map(const scb_index *idx, const uint8_t *bases, const uint64_t *offsets, size_t n,
                                uint64_t *keys, uint64_t *values, size_t cap, size_t *n_keys, uint64_t *queried) {
    int rc = check_offsets(offsets, n);
    if (rc) return rc;
    std::vector<std::pair<uint64_t, uint32_t>> pairs;
    uint64_t q = 0;
    Status s = idx->engine->kmer_map_host(bases, offsets, n, pairs, q);
    if (!s.good()) return fail(s);
    if (cap && (!keys || !values)) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null output buffers");
    for (size_t i = 0; i < pairs.size() && i < cap; ++i) keys[i] = pairs[i].first, values[i] = pairs[i].second;
    if (n_keys) *n_keys = pairs.size();
    if (queried) *queried = q;
    return SCB_OK;
}

int scb_spectrum_acc_new(const scb_index *idx, scb_spectrum_acc **out) {
    if (!idx || !out) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null argument");
    std::unique_ptr<scb_spectrum_acc> a(new scb_spectrum_acc);
    a->idx = idx;
    Status s = idx->engine->keyset_create(a->ks);
    if (!s.good()) return fail(s);
    *out = a.release();
    return SCB_OK;
}

void scb_spectrum_acc_free(scb_spectrum_acc *acc) { delete acc; }

int scb_spectrum_acc_add(scb_spectrum_acc *acc, const uint8_t *bases, const uint64_t *offsets, size_t n) {
    int rc = check_offsets(offsets, n);
    if (rc) return rc;
    Status s = acc->idx->engine->spectrum_accumulate_host(bases, offsets, n, *acc->ks);
    return s.good() ? SCB_OK : fail(s);
}

int scb_spectrum_acc_add_device(scb_spectrum_acc *acc, const uint8_t *d_bases, const uint64_t *d_offsets,
                                size_t n, uint64_t total_bases) {
    if (total_bases >= scb::kMaxBatchBases)
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: device batch exceeds 2^32 bases");
    Status s = acc->idx->engine->spectrum_accumulate(scb::DeviceBatch{d_bases, d_offsets, n, total_bases}, *acc->ks);
    return s.good() ? SCB_OK : fail(s);
}

int scb_spectrum_acc_rows(const scb_spectrum_acc *acc, uint64_t *values, uint64_t *counts, size_t cap,
                          size_t *n_rows, uint64_t *queried, uint64_t *distinct) {
    scb::SpectrumResult r;
    Status s = acc->idx->engine->keyset_rows(*acc->ks, r);
    if (!s.good()) return fail(s);
    for (size_t i = 0; i < r.rows.size() && i < cap; ++i) values[i] = r.rows[i].first, counts[i] = r.rows[i].second;
    if (n_rows) *n_rows = r.rows.size();
    if (queried) *queried = r.queried;
    if (distinct) *distinct = r.distinct;
    return SCB_OK;
}

// api.rs:933-956, 1013-1033 on a given histogram
int scb_filter_only_spectrum_from_histogram(uint64_t n_hashes, const uint64_t *hv, const uint64_t *hc, size_t hr,
                                            uint64_t max_abundance, uint64_t hint, uint64_t *counts) {
    if ((hr && (!hv || !hc)) || (max_abundance && !counts))
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null argument");
    auto d = deconvolved_counts(hv, hc, hr, n_hashes, max_abundance);
    double total = 0.0;
    for (double x : d) total += x;
    if (total <= 0.0) {
        uint64_t pos = 0;
        for (size_t i = 0; i < hr; ++i)
            if (hv[i] > 0) pos += hc[i];
        double positive = double(std::max<uint64_t>(pos, 1));
        for (uint64_t a = 1; a <= max_abundance; ++a) {
            uint64_t c = 0;
            for (size_t i = 0; i < hr; ++i)
                if (hv[i] == a) c = hc[i];
            counts[a - 1] = f64_as_u64(std::round(double(hint) * double(c) / positive));
        }
        return SCB_OK;
    }
    double scale = double(hint) / total;
    for (uint64_t a = 1; a <= max_abundance; ++a) counts[a - 1] = f64_as_u64(std::round(d[size_t(a - 1)] * scale));
    return SCB_OK;
}

// api.rs:375-385 on a given histogram
int scb_filter_only_count_hint_from_histogram(uint64_t n_hashes, const uint64_t *hv, const uint64_t *hc, size_t hr,
                                              uint64_t max_abundance, uint64_t *hint) {
    if ((hr && (!hv || !hc)) || !hint) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null argument");
    auto d = deconvolved_counts(hv, hc, hr, n_hashes, max_abundance);
    double total = 0.0;
    for (double x : d) total += x;
    *hint = f64_as_u64(std::fmax(std::round(total), 1.0));
    return SCB_OK;
}

// api.rs:337-368: the index's own counter histogram unless one is given
int scb_index_filter_only_spectrum(const scb_index *idx, const uint64_t *hv, const uint64_t *hc,
                                   size_t hr, uint64_t max_abundance, uint64_t hint, uint64_t *counts) {
    std::vector<std::pair<uint64_t, uint64_t>> rows;
    std::vector<uint64_t> vv, cc;
    if (!hv) {
        int rc = index_histogram(idx, rows);
        if (rc) return rc;
        for (auto &r : rows) vv.push_back(r.first), cc.push_back(r.second);
        hv = vv.data(), hc = cc.data(), hr = vv.size();
    }
    return scb_filter_only_spectrum_from_histogram(idx->cfg.n_hashes, hv, hc, hr, max_abundance, hint, counts);
}

// api.rs:370-385
int scb_index_filter_only_count_hint(const scb_index *idx, const uint64_t *hv, const uint64_t *hc,
                                     size_t hr, uint64_t max_abundance, uint64_t *hint) {
    std::vector<std::pair<uint64_t, uint64_t>> rows;
    std::vector<uint64_t> vv, cc;
    if (!hv) {
        int rc = index_histogram(idx, rows);
        if (rc) return rc;
        for (auto &r : rows) vv.push_back(r.first), cc.push_back(r.second);
        hv = vv.data(), hc = cc.data(), hr = vv.size();
    }
    return scb_filter_only_count_hint_from_histogram(idx->cfg.n_hashes, hv, hc, hr, max_abundance, hint);
}

// ---- FASTX pipelines ----------------------------------------------------------
// fastx.rs:133-170
int scb_estimate_spectrum_from_fastx_pair(const char *index_path, const char *query_path,
                                          const scb_config *cfg, uint64_t *values, uint64_t *counts,
                                          size_t cap, size_t *n_rows, scb_spectrum_stats *stats) {
    int rc = full_check(*cfg);
    if (rc) return rc;
    double t_total = now_s();
    scb_index *idx = nullptr;
    scb_build_report br{};
    if ((rc = scb_index_build_from_fastx(index_path, cfg, &idx, &br))) return rc;
    std::unique_ptr<scb_index, void (*)(scb_index *)> guard(idx, scb_index_free);
    double tq = now_s();
    uint64_t queried = 0, distinct = 0;
    // fastx.rs:255-308: the query file is read batch by batch; one device key
    // set spans all batches, so a key is counted once however many batches
    // hold it (the merged per-worker maps of fastx.rs:392-400)
    scb::KeySetPtr ks;
    Status s = idx->engine->keyset_create(ks);
    if (!s.good()) return fail(s);
    rc = for_each_batch(query_path, [&](HostBatch &hb) -> int {
        Status bs = idx->engine->spectrum_accumulate_host(hb.bases.data(), hb.offsets.data(), hb.n(), *ks);
        return bs.good() ? SCB_OK : fail(bs);
    });
    if (rc) return rc;
    double tqs = now_s();
    scb::SpectrumResult r;
    s = idx->engine->keyset_rows(*ks, r);
    if (!s.good()) return fail(s);
    queried = r.queried;
    distinct = r.distinct;
    double tqe = now_s();
    for (size_t i = 0; i < r.rows.size() && i < cap; ++i) values[i] = r.rows[i].first, counts[i] = r.rows[i].second;
    if (n_rows) *n_rows = r.rows.size();
    if (stats) {
        *stats = scb_spectrum_stats{};
        stats->chunks = br.chunks;
        stats->indexed_kmers = br.indexed_kmers;
        stats->inserted_smers = br.inserted_smers;
        stats->queried_kmers = queried;
        stats->distinct_keys = distinct;
        stats->counter_slots = br.counter_slots;
        stats->blocks = br.blocks;
        stats->build_seconds = br.build_seconds;
        stats->query_seconds = tqs - tq;
        stats->spectrum_seconds = tqe - tqs;
        stats->total_seconds = now_s() - t_total;
    }
    return SCB_OK;
}

// fastx.rs:172-204
int scb_stream_estimates_from_fastx_pair(const char *index_path, const char *query_path,
                                         const scb_config *cfg, scb_stream_stats *stats) {
    int rc = full_check(*cfg);
    if (rc) return rc;
    double t_total = now_s();
    scb_index *idx = nullptr;
    scb_build_report br{};
    if ((rc = scb_index_build_from_fastx(index_path, cfg, &idx, &br))) return rc;
    std::unique_ptr<scb_index, void (*)(scb_index *)> guard(idx, scb_index_free);
    double tq = now_s();
    scb_stream_stats st{};
    rc = for_each_batch(query_path, [&](HostBatch &hb) -> int {
        scb::QueryResult r;
        Status s = idx->engine->query_host(hb.bases.data(), hb.offsets.data(), hb.n(), idx->method,
                                           nullptr, nullptr, r);
        if (!s.good()) return fail(s);
        st.queried_kmers += r.queried;
        st.positive_estimates += r.positive;
        st.estimate_checksum += r.checksum;
        return SCB_OK;
    });
    if (rc) return rc;
    st.chunks = br.chunks;
    st.indexed_kmers = br.indexed_kmers;
    st.inserted_smers = br.inserted_smers;
    st.counter_slots = br.counter_slots;
    st.blocks = br.blocks;
    st.build_seconds = br.build_seconds;
    st.query_seconds = now_s() - tq;
    st.total_seconds = now_s() - t_total;
    if (stats) *stats = st;
    return SCB_OK;
}

// ---- reservoir-sampled spectrum (bin/sampled_spectrum.rs) ----------------------
namespace {
// query_sampled_kmers' histogram (sampled_spectrum.rs:201-221) of the sampled
// estimates and write_scaled_spectrum's scaling (225-253), in f64 as there.
std::vector<scb_sampled_row> sampled_rows(const std::vector<uint32_t> &vals, uint64_t seen) {
    std::vector<std::pair<uint64_t, uint64_t>> h;
    {
        std::vector<uint32_t> v(vals);
        std::sort(v.begin(), v.end());
        for (size_t i = 0; i < v.size();) {
            size_t j = i;
            while (j < v.size() && v[j] == v[i]) ++j;
            h.emplace_back(v[i], j - i);
            i = j;
        }
    }
    const double sample_count = double(std::max<size_t>(vals.size(), 1));
    const double scale = double(seen) / sample_count;
    std::vector<scb_sampled_row> rows;
    for (auto &r : h) {
        scb_sampled_row o{};
        o.estimated_count = r.first;
        o.sampled_kmers = r.second;
        o.linear_scaled_windows = f64_as_u64(std::round(double(r.second) * scale));
        o.estimated_distinct_kmers =
            r.first == 0 ? 0 : f64_as_u64(std::round(double(o.linear_scaled_windows) / double(r.first)));
        rows.push_back(o);
    }
    return rows;
}

int sampled_out(const scb_index *idx, const scb::Reservoir &rs, scb_sampled_row *rows, size_t cap, size_t *n_rows,
                scb_sampled_stats *stats) {
    std::vector<uint32_t> vals;
    uint64_t seen = 0;
    Status s = idx->engine->reservoir_values(rs, vals, seen);
    if (!s.good()) return fail(s);
    auto out = sampled_rows(vals, seen);
    if (cap && !rows) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null row buffer");
    for (size_t i = 0; i < out.size() && i < cap; ++i) rows[i] = out[i];
    if (n_rows) *n_rows = out.size();
    if (stats) {
        stats->total_inserted_kmers = seen;
        stats->sampled_kmers = vals.size();
        stats->queried_kmers = vals.size();
        stats->invalid_queries = 0;  // every sampled window lies in an ACGT run
    }
    return SCB_OK;
}
}  // namespace

int scb_index_sampled_spectrum(const scb_index *idx, const uint8_t *bases, const uint64_t *offsets, size_t n,
                               uint64_t sample_size, uint64_t seed, scb_sampled_row *rows, size_t cap,
                               size_t *n_rows, scb_sampled_stats *stats) {
    if (sample_size == 0) return fail(SCB_ERR_INVALID_CONFIG, "invalid config: sample_size must be positive");
    int rc = check_offsets(offsets, n);
    if (rc) return rc;
    double t0 = now_s();
    scb::ReservoirPtr rs;
    Status s = idx->engine->reservoir_create(sample_size, seed, rs);
    if (s.good()) s = idx->engine->sample_accumulate_host(bases, offsets, n, *rs);
    if (!s.good()) return fail(s);
    if (stats) *stats = scb_sampled_stats{};
    rc = sampled_out(idx, *rs, rows, cap, n_rows, stats);
    if (stats) stats->sample_seconds = now_s() - t0;
    return rc;
}

// bin/sampled_spectrum.rs:99-148 (run): build the index from the file, then
// sample its valid windows and query them, in one pass of the file's records.
int scb_sampled_spectrum_from_fastx(const char *path, const scb_config *cfg, uint64_t sample_size, uint64_t seed,
                                    scb_sampled_row *rows, size_t cap, size_t *n_rows, scb_sampled_stats *stats) {
    if (sample_size == 0) return fail(SCB_ERR_INVALID_CONFIG, "invalid config: sample_size must be positive");
    int rc = full_check(*cfg);
    if (rc) return rc;
    scb_index *idx = nullptr;
    scb_build_report br{};
    if ((rc = scb_index_build_from_fastx(path, cfg, &idx, &br))) return rc;
    std::unique_ptr<scb_index, void (*)(scb_index *)> guard(idx, scb_index_free);
    double t0 = now_s();
    scb::ReservoirPtr rs;
    Status s = idx->engine->reservoir_create(sample_size, seed, rs);
    if (!s.good()) return fail(s);
    rc = for_each_batch(path, [&](HostBatch &hb) -> int {
        Status bs = idx->engine->sample_accumulate_host(hb.bases.data(), hb.offsets.data(), hb.n(), *rs);
        return bs.good() ? SCB_OK : fail(bs);
    });
    if (rc) return rc;
    scb_sampled_stats st{};
    rc = sampled_out(idx, *rs, rows, cap, n_rows, &st);
    if (rc) return rc;
    st.indexed_kmers = br.indexed_kmers;
    st.inserted_smers = br.inserted_smers;
    st.build_seconds = br.build_seconds;
    st.sample_seconds = now_s() - t0;
    if (stats) *stats = st;
    return SCB_OK;
}

// write_scaled_spectrum (sampled_spectrum.rs:225-253): CSV with its header
int scb_write_sampled_spectrum(const scb_sampled_row *rows, size_t n, const char *path) {
    FILE *fp = path ? std::fopen(path, "wb") : stdout;
    if (!fp) return fail(SCB_ERR_IO, std::string("I/O error for '") + path + "': " + std::strerror(errno));
    std::fprintf(fp, "estimated_count,sampled_kmers,linear_scaled_windows,estimated_distinct_kmers\n");
    for (size_t i = 0; i < n; ++i)
        std::fprintf(fp, "%llu,%llu,%llu,%llu\n", (unsigned long long)rows[i].estimated_count,
                     (unsigned long long)rows[i].sampled_kmers, (unsigned long long)rows[i].linear_scaled_windows,
                     (unsigned long long)rows[i].estimated_distinct_kmers);
    int bad = std::ferror(fp);
    if (path) bad |= std::fclose(fp);
    else std::fflush(fp);
    if (bad) return fail(SCB_ERR_IO, std::string("I/O error for '") + (path ? path : "<stream>") + "': write failed");
    return SCB_OK;
}

// spectrum.rs:9-31
int scb_write_spectrum(const uint64_t *values, const uint64_t *counts, size_t n, const char *path) {
    FILE *fp = path ? fopen(path, "wb") : stdout;
    if (!fp) return fail(SCB_ERR_IO, "I/O error for '" + std::string(path) + "': " + os_error(errno));
    bool ok = fprintf(fp, "estimated_count\tkmers\n") > 0;
    for (size_t i = 0; i < n && ok; ++i)
        ok = fprintf(fp, "%llu\t%llu\n", (unsigned long long)values[i], (unsigned long long)counts[i]) > 0;
    if (path) ok = (fclose(fp) == 0) && ok;
    else fflush(fp);
    if (!ok) return fail(SCB_ERR_IO, "I/O error for '<stream>': write failed");
    return SCB_OK;
}

// api.rs:893-915
int scb_for_each_fastx_record(const char *path, scb_record_visitor visit, void *user) {
    scb_fastx::Reader rd;
    int rc = rd.open(path);
    if (rc) return rc;
    std::string hdr;
    std::vector<uint8_t> seq;
    uint64_t idx = 0;
    while (rd.next(hdr, seq)) {
        ++idx;
        std::string name = record_name(hdr, idx);
        if ((rc = visit(user, name.c_str(), seq.data(), seq.size()))) return rc;
    }
    return rd.error();
}

}  // extern "C"
