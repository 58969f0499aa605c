// Serving loop over a frozen index resident in HBM (declared in
// include/scb.h; no counterpart in the reference, whose index is queried by one
// caller at a time through api.rs:250-327).
//
// Requests queue in FIFO order. Each worker thread owns a HIP stream and a
// staging area (device batch buffers plus page-locked host buffers) sized at
// creation, so serving allocates nothing per request. A worker takes the
// oldest request; abundance requests behind it are coalesced into the same
// device batch while they fit. Every request (or chunk of an oversized one)
// starts at a device position congruent to its own first position modulo 32,
// so its validity bits are whole-word aligned with the caller's bitmap; the
// bytes in between form filler records whose results are never copied out.
// Per-request reports come from one segmented reduction on the device.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/scb.h"
#include "scb_engine.h"
#include "scb_handles.h"

using scb::Status;
using scb_host::fail;

namespace {

uint64_t now_ns() {
    return uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
                        std::chrono::steady_clock::now().time_since_epoch())
                        .count());
}

struct Request {
    uint64_t ticket = 0;
    uint32_t kind = SCB_REQ_ABUNDANCES;
    const uint8_t *bases = nullptr;
    const uint64_t *offsets = nullptr;
    size_t n = 0;
    uint64_t nb = 0;  // bases over all records
    void *values = nullptr;
    uint32_t *valid_bits = nullptr;
    uint64_t *rows_v = nullptr, *rows_c = nullptr;
    size_t cap = 0;
    bool done = false;
    bool claimed = false;  // a scb_server_wait call is blocked on this ticket
    scb_served res{};
};
using ReqPtr = std::shared_ptr<Request>;

// Device bytes per staged base: bases, counters, validity, and the pipeline
// scratch of the batched query (records + sorted records at 8 bytes each per
// possible super-k-mer start, packed bases).
constexpr double kScratchPerBase = 1.0 + 4.0 + 0.125 + 16.0 + 0.25 + 1.0;

struct Staging {
    hipStream_t st = nullptr;
    uint64_t bases_cap = 0, rec_cap = 0, seg_cap = 0;
    uint8_t *d_bases = nullptr;
    uint64_t *d_offsets = nullptr, *d_segs = nullptr, *d_stats = nullptr;
    void *d_values = nullptr;
    uint32_t *d_bits = nullptr;
    uint64_t *h_offsets = nullptr, *h_segs = nullptr, *h_stats = nullptr;
    uint32_t *h_bits = nullptr;
};

Status alloc_staging(Staging &s, uint64_t bases_cap, uint64_t rec_cap, uint64_t seg_cap, uint32_t W) {
#define SCB_S(call)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) return Status{8, std::string("device error: ") + #call + ": " + hipGetErrorString(e_)}; \
    } while (0)
    s.bases_cap = bases_cap, s.rec_cap = rec_cap, s.seg_cap = seg_cap;
    const uint64_t words = (bases_cap + 31) / 32 + 2;
    SCB_S(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
    SCB_S(hipMalloc(reinterpret_cast<void **>(&s.d_bases), bases_cap + 16));
    SCB_S(hipMalloc(reinterpret_cast<void **>(&s.d_offsets), (rec_cap + 1) * 8));
    SCB_S(hipMalloc(reinterpret_cast<void **>(&s.d_segs), seg_cap * 16));
    SCB_S(hipMalloc(reinterpret_cast<void **>(&s.d_stats), seg_cap * 24));
    SCB_S(hipMalloc(&s.d_values, ((bases_cap * W + 3) / 4 + 1) * 4));
    SCB_S(hipMalloc(reinterpret_cast<void **>(&s.d_bits), words * 4));
    SCB_S(hipHostMalloc(reinterpret_cast<void **>(&s.h_offsets), (rec_cap + 1) * 8, hipHostMallocDefault));
    SCB_S(hipHostMalloc(reinterpret_cast<void **>(&s.h_segs), seg_cap * 16, hipHostMallocDefault));
    SCB_S(hipHostMalloc(reinterpret_cast<void **>(&s.h_stats), seg_cap * 24, hipHostMallocDefault));
    SCB_S(hipHostMalloc(reinterpret_cast<void **>(&s.h_bits), words * 4, hipHostMallocDefault));
#undef SCB_S
    return Status::ok();
}

void free_staging(Staging &s) {
    if (s.st) (void)hipStreamSynchronize(s.st);
    for (void *p : {static_cast<void *>(s.d_bases), static_cast<void *>(s.d_offsets),
                    static_cast<void *>(s.d_segs), static_cast<void *>(s.d_stats), s.d_values,
                    static_cast<void *>(s.d_bits)})
        if (p) (void)hipFree(p);
    for (void *p : {static_cast<void *>(s.h_offsets), static_cast<void *>(s.h_segs),
                    static_cast<void *>(s.h_stats), static_cast<void *>(s.h_bits)})
        if (p) (void)hipHostFree(p);
    if (s.st) (void)hipStreamDestroy(s.st);
    s = Staging{};
}

// One request's records [r0, r1) placed in the device batch.
struct Piece {
    Request *req;
    size_t r0, r1;
    uint64_t dev_pos;  // device position of the piece's first base
    uint64_t req_pos;  // position of that base within the request
    uint64_t len;
};

}  // namespace

struct scb_server {
    const scb_index *idx = nullptr;
    scb_serve_config cfg{};
    uint32_t W = 2;
    uint64_t max_record = 0;  // longest record an abundance request may hold
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ReqPtr> queue;
    std::unordered_map<uint64_t, ReqPtr> live;  // submitted and not yet waited for
    std::deque<uint64_t> finished;             // finished, not yet waited for
    uint64_t next_ticket = 1;
    bool stopping = false;
    std::vector<Staging> staging;
    std::vector<std::thread> workers;

    void worker(size_t wi);
    void run_abundances(Staging &s, std::vector<Request *> &batch);
    void run_spectrum(Staging &s, Request *r);
};

namespace {

void set_error(Request *r, const Status &s) {
    r->res.status = s.code;
    std::snprintf(r->res.message, sizeof(r->res.message), "%s", s.msg.c_str());
}

}  // namespace

// Runs the pieces placed so far as one query launch and scatters the results.
void scb_server::run_abundances(Staging &s, std::vector<Request *> &batch) {
    const scb::Engine &eng = *idx->engine;
    std::vector<Piece> pieces;
    uint64_t p = 0, nrec = 0;
    s.h_offsets[0] = 0;
    Status st = Status::ok();
    auto flush = [&]() -> Status {
        if (pieces.empty()) return Status::ok();
        const uint64_t nb = p;
        const uint64_t words = (nb + 31) / 32;
        Status r = scb::fill_u32_async(reinterpret_cast<uint32_t *>(s.d_values), (nb * W + 3) / 4, 0, s.st);
        if (!r.good()) return r;
        if (hipMemcpyAsync(s.d_offsets, s.h_offsets, (nrec + 1) * 8, hipMemcpyHostToDevice, s.st) != hipSuccess ||
            hipMemcpyAsync(s.d_segs, s.h_segs, pieces.size() * 16, hipMemcpyHostToDevice, s.st) != hipSuccess ||
            hipStreamSynchronize(s.st) != hipSuccess)
            return Status{8, "device error: staging a served batch failed"};
        scb::DeviceBatch db{s.d_bases, s.d_offsets, nrec, nb};
        scb::QueryResult qr;
        r = eng.query(db, idx->method, s.d_values, s.d_bits, qr);
        if (!r.good()) return r;
        r = scb::segment_stats_async(W * 8, s.d_values, s.d_bits, s.d_segs, uint32_t(pieces.size()), nb, s.d_stats,
                                     s.st);
        if (!r.good()) return r;
        for (const Piece &pc : pieces)
            if (pc.req->values &&
                hipMemcpyAsync(static_cast<uint8_t *>(pc.req->values) + pc.req_pos * W,
                               static_cast<const uint8_t *>(s.d_values) + pc.dev_pos * W, pc.len * W,
                               hipMemcpyDeviceToHost, s.st) != hipSuccess)
                return Status{8, "device error: copying served abundances failed"};
        if (hipMemcpyAsync(s.h_bits, s.d_bits, words * 4, hipMemcpyDeviceToHost, s.st) != hipSuccess ||
            hipMemcpyAsync(s.h_stats, s.d_stats, pieces.size() * 24, hipMemcpyDeviceToHost, s.st) != hipSuccess ||
            hipStreamSynchronize(s.st) != hipSuccess)
            return Status{8, "device error: copying served results failed"};
        for (size_t i = 0; i < pieces.size(); ++i) {
            const Piece &pc = pieces[i];
            Request *q = pc.req;
            q->res.report.valid_kmers += s.h_stats[3 * i];
            q->res.report.positive_estimates += s.h_stats[3 * i + 1];
            q->res.report.estimate_checksum += s.h_stats[3 * i + 2];
            if (!q->valid_bits || !pc.len) continue;
            // device word w holds request word w - dev_pos/32 + req_pos/32 (dev_pos = req_pos mod 32)
            const uint64_t dw0 = pc.dev_pos >> 5, dw1 = (pc.dev_pos + pc.len - 1) >> 5;
            const uint64_t shift = (pc.req_pos >> 5) - dw0;
            for (uint64_t w = dw0; w <= dw1; ++w) {
                uint32_t mask = ~0u;
                if (w == dw0) mask &= ~0u << (pc.dev_pos & 31);
                if (w == dw1) {
                    uint32_t e = uint32_t((pc.dev_pos + pc.len) & 31);
                    if (e) mask &= ~0u >> (32 - e);
                }
                uint32_t &dst = q->valid_bits[w + shift];
                dst = mask == ~0u ? s.h_bits[w] : ((dst & ~mask) | (s.h_bits[w] & mask));
            }
        }
        pieces.clear();
        p = 0, nrec = 0;
        return Status::ok();
    };
    // place a record range of one request, starting a new launch when it does not fit
    auto place = [&](Request *q, size_t r0, size_t r1) -> Status {
        const uint64_t req_pos = q->offsets[r0] - q->offsets[0];
        const uint64_t len = q->offsets[r1] - q->offsets[r0];
        uint64_t pad = (req_pos - p) & 31;
        if (p + pad + len > s.bases_cap || nrec + (r1 - r0) + 1 > s.rec_cap || pieces.size() + 1 > s.seg_cap) {
            Status f = flush();
            if (!f.good()) return f;
            pad = req_pos & 31;
        }
        if (pad) {  // filler record: its results are masked out
            p += pad;
            s.h_offsets[++nrec] = p;
        }
        if (len &&
            hipMemcpyAsync(s.d_bases + p, q->bases + q->offsets[r0], len, hipMemcpyHostToDevice, s.st) != hipSuccess)
            return Status{8, "device error: copying request bases failed"};
        for (size_t r = r0; r < r1; ++r) s.h_offsets[++nrec] = p + (q->offsets[r + 1] - q->offsets[r0]);
        s.h_segs[2 * pieces.size()] = p;
        s.h_segs[2 * pieces.size() + 1] = p + len;
        pieces.push_back(Piece{q, r0, r1, p, req_pos, len});
        p += len;
        return Status::ok();
    };
    for (Request *q : batch) {
        q->res.report = scb_query_report{};
        q->res.report.records_processed = q->n;
        const uint64_t k = idx->cfg.k;
        for (size_t i = 0; i < q->n; ++i) {
            uint64_t len = q->offsets[i + 1] - q->offsets[i];
            if (len >= k) q->res.report.total_windows += len + 1 - k;
        }
        // split into record ranges that fit one launch
        size_t r0 = 0;
        while (st.good() && r0 < q->n) {
            size_t r1 = r0;
            uint64_t room = s.bases_cap - 32, recs = s.rec_cap - 2;
            while (r1 < q->n && q->offsets[r1 + 1] - q->offsets[r0] <= room && r1 - r0 < recs) ++r1;
            if (r1 == r0) r1 = r0 + 1;  // a single record always fits (checked at submit)
            st = place(q, r0, r1);
            r0 = r1;
        }
        if (!st.good()) break;
    }
    if (st.good()) st = flush();
    if (!st.good())
        for (Request *q : batch) set_error(q, st);
}

void scb_server::run_spectrum(Staging &s, Request *r) {
    const scb::Engine &eng = *idx->engine;
    scb::SpectrumResult sr;
    Status st = Status::ok();
    if (r->nb + 16 <= s.bases_cap && r->n + 1 <= s.rec_cap) {
        for (size_t i = 0; i <= r->n; ++i) s.h_offsets[i] = r->offsets[i] - r->offsets[0];
        if ((r->nb && hipMemcpyAsync(s.d_bases, r->bases + r->offsets[0], r->nb, hipMemcpyHostToDevice, s.st) !=
                          hipSuccess) ||
            hipMemcpyAsync(s.d_offsets, s.h_offsets, (r->n + 1) * 8, hipMemcpyHostToDevice, s.st) != hipSuccess ||
            hipStreamSynchronize(s.st) != hipSuccess)
            st = Status{8, "device error: staging a spectrum request failed"};
        if (st.good()) st = eng.spectrum(scb::DeviceBatch{s.d_bases, s.d_offsets, r->n, r->nb}, sr, idx->method);
    } else {
        st = eng.spectrum_host(r->bases, r->offsets, r->n, sr, idx->method);  // larger than the staging area
    }
    if (!st.good()) {
        set_error(r, st);
        return;
    }
    for (size_t i = 0; i < sr.rows.size() && i < r->cap; ++i) r->rows_v[i] = sr.rows[i].first, r->rows_c[i] = sr.rows[i].second;
    r->res.n_rows = sr.rows.size();
    r->res.queried = sr.queried;
    r->res.distinct = sr.distinct;
}

void scb_server::worker(size_t wi) {
    (void)hipSetDevice(idx->engine->device());
    Staging &s = staging[wi];
    for (;;) {
        std::vector<ReqPtr> batch;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stopping || !queue.empty(); });
            if (queue.empty()) return;  // stopping and drained
            batch.push_back(queue.front());
            queue.pop_front();
            if (batch[0]->kind == SCB_REQ_ABUNDANCES && batch[0]->nb + 32 <= s.bases_cap) {
                uint64_t tb = batch[0]->nb + 32, tr = batch[0]->n + 2;
                while (!queue.empty() && batch.size() < s.seg_cap) {
                    const ReqPtr &q = queue.front();
                    if (q->kind != SCB_REQ_ABUNDANCES || tb + q->nb + 32 > s.bases_cap || tr + q->n + 2 > s.rec_cap)
                        break;
                    tb += q->nb + 32, tr += q->n + 2;
                    batch.push_back(q);
                    queue.pop_front();
                }
            }
        }
        const uint64_t t0 = now_ns();
        std::vector<Request *> raw;
        for (auto &r : batch) {
            r->res.start_ns = t0;
            r->res.batch_requests = batch.size();
            raw.push_back(r.get());
        }
        if (raw[0]->kind == SCB_REQ_SPECTRUM) run_spectrum(s, raw[0]);
        else run_abundances(s, raw);
        const uint64_t t1 = now_ns();
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto &r : batch) {
                r->res.done_ns = t1;
                r->done = true;
                finished.push_back(r->ticket);
            }
        }
        cv_done.notify_all();
    }
}

extern "C" {

void scb_serve_config_default(scb_serve_config *cfg) {
    if (!cfg) return;
    cfg->max_batch_bases = uint64_t(1) << 27;
    cfg->max_batch_requests = 1024;
    cfg->hbm_budget_bytes = uint64_t(16) << 30;
    cfg->workers = 4;
    cfg->max_pending = 4096;
}

int scb_server_new(const scb_index *idx, const scb_serve_config *cfg_in, scb_server **out) {
    if (!idx || !out) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null argument");
    scb_serve_config cfg;
    scb_serve_config_default(&cfg);
    if (cfg_in) cfg = *cfg_in;
    if (cfg.workers == 0) return fail(SCB_ERR_INVALID_CONFIG, "invalid config: workers must be positive");
    if (cfg.max_pending == 0) return fail(SCB_ERR_INVALID_CONFIG, "invalid config: max_pending must be positive");
    if (cfg.max_batch_requests == 0)
        return fail(SCB_ERR_INVALID_CONFIG, "invalid config: max_batch_requests must be positive");
    // admission by HBM budget: staging and pipeline scratch of every worker
    uint64_t fit = uint64_t(double(cfg.hbm_budget_bytes) / (kScratchPerBase * cfg.workers));
    uint64_t bases_cap = std::min<uint64_t>(cfg.max_batch_bases, fit);
    bases_cap = std::min<uint64_t>(bases_cap, scb::kMaxBatchBases - 4096);
    if (bases_cap < 4096)
        return fail(SCB_ERR_INVALID_CONFIG, "invalid config: hbm_budget_bytes leaves less than 4096 bases per worker");
    std::unique_ptr<scb_server> srv(new scb_server);
    srv->idx = idx;
    srv->cfg = cfg;
    srv->W = idx->cfg.counter_bits / 8;
    srv->max_record = bases_cap - 64;
    const uint64_t seg_cap = std::min<uint64_t>(cfg.max_batch_requests, bases_cap / 64);
    const uint64_t rec_cap = bases_cap / 8 + 2 * seg_cap + 8;
    if (hipSetDevice(idx->engine->device()) != hipSuccess)
        return fail(SCB_ERR_DEVICE, "device error: cannot select the index's device");
    srv->staging.resize(cfg.workers);
    for (auto &s : srv->staging) {
        Status st = alloc_staging(s, bases_cap, rec_cap, seg_cap, srv->W);
        if (!st.good()) {
            for (auto &t : srv->staging) free_staging(t);
            return fail(st);
        }
    }
    for (size_t i = 0; i < cfg.workers; ++i) srv->workers.emplace_back([p = srv.get(), i] { p->worker(i); });
    *out = srv.release();
    return SCB_OK;
}

void scb_server_free(scb_server *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true;
    }
    srv->cv_work.notify_all();
    for (auto &t : srv->workers) t.join();
    for (auto &s : srv->staging) free_staging(s);
    delete srv;
}

namespace {

int submit(scb_server *srv, ReqPtr r, uint64_t *ticket) {
    if (!srv || !ticket) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null argument");
    if (!r->offsets) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: offsets is NULL");
    if (r->offsets[0] != 0) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: offsets[0] must be 0");
    uint64_t longest = 0;
    for (size_t i = 0; i < r->n; ++i) {
        if (r->offsets[i + 1] < r->offsets[i])
            return fail(SCB_ERR_INTERNAL_STATE, "internal state error: offsets must be non-decreasing");
        longest = std::max<uint64_t>(longest, r->offsets[i + 1] - r->offsets[i]);
    }
    if (r->kind == SCB_REQ_ABUNDANCES && longest > srv->max_record)
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: a record is longer than the serving batch (" +
                                                std::to_string(srv->max_record) + " bases)");
    r->nb = r->offsets[r->n];
    r->res.kind = r->kind;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stopping) return fail(SCB_ERR_CHANNEL_CLOSED, "channel closed: the server is stopping");
        if (srv->queue.size() >= srv->cfg.max_pending)
            return fail(SCB_ERR_BUSY, "server busy: " + std::to_string(srv->queue.size()) + " requests queued");
        r->ticket = srv->next_ticket++;
        r->res.submit_ns = now_ns();
        srv->live.emplace(r->ticket, r);
        srv->queue.push_back(r);
        *ticket = r->ticket;
    }
    srv->cv_work.notify_one();
    return SCB_OK;
}

}  // namespace

int scb_server_submit_abundances(scb_server *srv, const uint8_t *bases, const uint64_t *offsets, size_t n,
                                 void *values, uint32_t *valid_bits, uint64_t *ticket) {
    auto r = std::make_shared<Request>();
    r->kind = SCB_REQ_ABUNDANCES;
    r->bases = bases, r->offsets = offsets, r->n = n;
    r->values = values, r->valid_bits = valid_bits;
    return submit(srv, std::move(r), ticket);
}

int scb_server_submit_spectrum(scb_server *srv, const uint8_t *bases, const uint64_t *offsets, size_t n,
                               uint64_t *values, uint64_t *counts, size_t cap, uint64_t *ticket) {
    if (cap && (!values || !counts)) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null row buffers");
    auto r = std::make_shared<Request>();
    r->kind = SCB_REQ_SPECTRUM;
    r->bases = bases, r->offsets = offsets, r->n = n;
    r->rows_v = values, r->rows_c = counts, r->cap = cap;
    return submit(srv, std::move(r), ticket);
}

namespace {

int take(scb_server *srv, uint64_t ticket, scb_served *out) {
    auto it = srv->live.find(ticket);
    if (it == srv->live.end())
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: unknown or already waited ticket");
    ReqPtr r = it->second;
    srv->live.erase(it);
    auto f = std::find(srv->finished.begin(), srv->finished.end(), ticket);
    if (f != srv->finished.end()) srv->finished.erase(f);
    if (out) *out = r->res;
    if (r->res.status != SCB_OK) return fail(r->res.status, r->res.message);
    return SCB_OK;
}

}  // namespace

int scb_server_wait(scb_server *srv, uint64_t ticket, scb_served *out) {
    if (!srv) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->live.find(ticket);
    if (it == srv->live.end())
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: unknown or already waited ticket");
    ReqPtr r = it->second;
    if (r->claimed)
        return fail(SCB_ERR_INTERNAL_STATE, "internal state error: the ticket is already being waited for");
    // scb_server_wait_any skips claimed tickets, so this call alone takes it
    r->claimed = true;
    srv->cv_done.wait(lk, [&] { return r->done; });
    return take(srv, ticket, out);
}

int scb_server_wait_any(scb_server *srv, uint64_t timeout_us, uint64_t *ticket, scb_served *out) {
    if (!srv || !ticket) return fail(SCB_ERR_INTERNAL_STATE, "internal state error: null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    *ticket = 0;
    // the oldest finished request that no scb_server_wait call has claimed
    uint64_t t = 0;
    auto pick = [&] {
        for (uint64_t f : srv->finished) {
            auto it = srv->live.find(f);
            if (it != srv->live.end() && !it->second->claimed) {
                t = f;
                return true;
            }
        }
        return false;
    };
    if (!srv->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), pick)) return SCB_OK;
    *ticket = t;
    return take(srv, t, out);
}

void *scb_host_alloc(size_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

void scb_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

}  // extern "C"
