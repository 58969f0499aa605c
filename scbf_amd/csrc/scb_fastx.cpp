// FASTA/FASTQ input (declared in scb_fastx.h). A read-ahead thread fills 4 MiB
// blocks from the decompressor while the caller splits them into records, so
// decompression and parsing overlap. zstd is decoded through the system's
// libzstd.so.1 (loaded at first use; its stable streaming API is declared
// below), gzip (also multi-member) through zlib.
#include "scb_fastx.h"

#include <dlfcn.h>
#include <zlib.h>

#include <cerrno>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

#include "../../include/scb.h"
#include "scb_handles.h"

namespace scb_fastx {

using scb_host::fail;

// ---- byte sources ----------------------------------------------------------

class Source {
   public:
    virtual ~Source() = default;
    // Bytes read (> 0), 0 at the end, < 0 on error (message in err).
    virtual long read(char *buf, size_t cap, std::string &err) = 0;
};

namespace {

class PlainSource : public Source {
   public:
    explicit PlainSource(FILE *f) : f_(f) {}
    ~PlainSource() override { fclose(f_); }
    long read(char *buf, size_t cap, std::string &err) override {
        size_t n = fread(buf, 1, cap, f_);
        if (n == 0 && ferror(f_)) {
            err = std::string("read error: ") + strerror(errno);
            return -1;
        }
        return long(n);
    }

   private:
    FILE *f_;
};

class GzipSource : public Source {
   public:
    explicit GzipSource(gzFile g) : g_(g) { gzbuffer(g_, 1 << 20); }
    ~GzipSource() override { gzclose(g_); }
    long read(char *buf, size_t cap, std::string &err) override {
        int n = gzread(g_, buf, unsigned(std::min<size_t>(cap, 1u << 30)));
        if (n < 0) {
            int code = 0;
            const char *m = gzerror(g_, &code);
            err = std::string("gzip error: ") + (m ? m : "unknown");
            return -1;
        }
        return n;
    }

   private:
    gzFile g_;
};

// zstd streaming decompression API (zstd.h, stable since v1.0).
struct ZstdIn {
    const void *src;
    size_t size, pos;
};
struct ZstdOut {
    void *dst;
    size_t size, pos;
};
struct ZstdApi {
    void *(*create)() = nullptr;
    size_t (*free_ds)(void *) = nullptr;
    size_t (*init)(void *) = nullptr;
    size_t (*decompress)(void *, ZstdOut *, ZstdIn *) = nullptr;
    unsigned (*is_error)(size_t) = nullptr;
    const char *(*error_name)(size_t) = nullptr;
    bool ok = false;
};

const ZstdApi &zstd_api() {
    static ZstdApi api = [] {
        ZstdApi a;
        void *h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) return a;
        a.create = reinterpret_cast<void *(*)()>(dlsym(h, "ZSTD_createDStream"));
        a.free_ds = reinterpret_cast<size_t (*)(void *)>(dlsym(h, "ZSTD_freeDStream"));
        a.init = reinterpret_cast<size_t (*)(void *)>(dlsym(h, "ZSTD_initDStream"));
        a.decompress = reinterpret_cast<size_t (*)(void *, ZstdOut *, ZstdIn *)>(dlsym(h, "ZSTD_decompressStream"));
        a.is_error = reinterpret_cast<unsigned (*)(size_t)>(dlsym(h, "ZSTD_isError"));
        a.error_name = reinterpret_cast<const char *(*)(size_t)>(dlsym(h, "ZSTD_getErrorName"));
        a.ok = a.create && a.free_ds && a.init && a.decompress && a.is_error && a.error_name;
        return a;
    }();
    return api;
}

class ZstdSource : public Source {
   public:
    ZstdSource(FILE *f, const ZstdApi &api) : f_(f), api_(api), in_buf_(1 << 20) {
        ds_ = api_.create();
        if (ds_) api_.init(ds_);
    }
    ~ZstdSource() override {
        if (ds_) api_.free_ds(ds_);
        fclose(f_);
    }
    long read(char *buf, size_t cap, std::string &err) override {
        if (!ds_) {
            err = "zstd error: cannot create a decompression stream";
            return -1;
        }
        ZstdOut out{buf, cap, 0};
        while (out.pos == 0) {
            if (in_.pos == in_.size) {
                size_t n = fread(in_buf_.data(), 1, in_buf_.size(), f_);
                if (n == 0) {
                    if (ferror(f_)) {
                        err = std::string("read error: ") + strerror(errno);
                        return -1;
                    }
                    if (!frame_done_) {
                        err = "zstd error: truncated frame";
                        return -1;
                    }
                    return 0;
                }
                in_ = ZstdIn{in_buf_.data(), n, 0};
            }
            size_t r = api_.decompress(ds_, &out, &in_);
            if (api_.is_error(r)) {
                err = std::string("zstd error: ") + api_.error_name(r);
                return -1;
            }
            frame_done_ = r == 0;  // 0: a frame was completely decoded and flushed
        }
        return long(out.pos);
    }

   private:
    FILE *f_;
    const ZstdApi &api_;
    void *ds_ = nullptr;
    std::vector<char> in_buf_;
    ZstdIn in_{nullptr, 0, 0};
    bool frame_done_ = true;
};

int open_failed(const std::string &path, int e) {
    return fail(SCB_ERR_FASTX_OPEN, "failed to open FASTA/FASTQ '" + path + "': " + std::string(strerror(e)) +
                                        " (os error " + std::to_string(e) + ")");
}

// Opens path and picks the decoder from its first bytes.
int open_source(const std::string &path, std::unique_ptr<Source> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return open_failed(path, errno ? errno : ENOENT);
    unsigned char magic[4] = {0, 0, 0, 0};
    size_t got = fread(magic, 1, 4, f);
    rewind(f);
    if (got >= 2 && magic[0] == 0x1f && magic[1] == 0x8b) {
        fclose(f);
        gzFile g = gzopen(path.c_str(), "rb");
        if (!g) return open_failed(path, errno ? errno : ENOENT);
        out.reset(new GzipSource(g));
    } else if (got == 4 && magic[0] == 0x28 && magic[1] == 0xb5 && magic[2] == 0x2f && magic[3] == 0xfd) {
        const ZstdApi &api = zstd_api();
        if (!api.ok) {
            fclose(f);
            return fail(SCB_ERR_FASTX_OPEN, "failed to open FASTA/FASTQ '" + path +
                                                "': zstd input needs libzstd.so.1, which cannot be loaded");
        }
        out.reset(new ZstdSource(f, api));
    } else {
        out.reset(new PlainSource(f));
    }
    return SCB_OK;
}

}  // namespace

// ---- read-ahead ------------------------------------------------------------

class ReadAhead {
   public:
    static constexpr size_t kBlock = size_t(4) << 20;
    static constexpr size_t kDepth = 4;

    explicit ReadAhead(std::unique_ptr<Source> src) : src_(std::move(src)) {
        for (size_t i = 0; i < kDepth; ++i) free_.emplace_back(kBlock);
        th_ = std::thread([this] { run(); });
    }
    ~ReadAhead() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_free_.notify_all();
        th_.join();
    }
    // Next filled block (swapped into out); false at the end or on error.
    bool take(std::vector<char> &out) {
        std::unique_lock<std::mutex> lk(mu_);
        cv_ready_.wait(lk, [&] { return !ready_.empty() || done_; });
        if (ready_.empty()) return false;
        if (!out.empty() || out.capacity()) free_.push_back(std::move(out));
        out = std::move(ready_.front());
        ready_.pop_front();
        lk.unlock();
        cv_free_.notify_one();
        return true;
    }
    // The read error, recorded as the calling thread's last error.
    int error() {
        std::lock_guard<std::mutex> lk(mu_);
        return err_code_ ? fail(err_code_, err_msg_) : SCB_OK;
    }

   private:
    void run() {
        for (;;) {
            std::vector<char> b;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_free_.wait(lk, [&] { return !free_.empty() || stop_; });
                if (stop_) return;
                b = std::move(free_.back());
                free_.pop_back();
            }
            b.resize(kBlock);
            size_t n = 0;
            std::string err;
            bool end = false;
            while (n < kBlock) {
                long r = src_->read(b.data() + n, kBlock - n, err);
                if (r < 0) {
                    std::lock_guard<std::mutex> lk(mu_);
                    err_msg_ = err;
                    err_code_ = SCB_ERR_IO;
                    end = true;
                    break;
                }
                if (r == 0) {
                    end = true;
                    break;
                }
                n += size_t(r);
            }
            b.resize(n);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (n) ready_.push_back(std::move(b));
                if (end) done_ = true;
            }
            cv_ready_.notify_one();
            if (end) return;
        }
    }
    std::unique_ptr<Source> src_;
    std::mutex mu_;
    std::condition_variable cv_ready_, cv_free_;
    std::deque<std::vector<char>> ready_;
    std::vector<std::vector<char>> free_;
    bool done_ = false, stop_ = false;
    int err_code_ = SCB_OK;
    std::string err_msg_;
    std::thread th_;
};

// ---- records -----------------------------------------------------------------

Reader::Reader() = default;
Reader::~Reader() = default;

int Reader::open(const char *path) {
    path_ = path ? path : "";
    std::unique_ptr<Source> src;
    int rc = open_source(path_, src);
    if (rc) return rc;
    ra_.reset(new ReadAhead(std::move(src)));
    return SCB_OK;
}

int Reader::error() { return ra_ ? ra_->error() : SCB_OK; }

bool Reader::next_line(const char *&p, size_t &n) {
    if (pos_ >= block_.size()) {
        if (!ra_ || !ra_->take(block_)) return false;
        pos_ = 0;
    }
    const char *start = block_.data() + pos_;
    const char *end = block_.data() + block_.size();
    const char *nl = static_cast<const char *>(memchr(start, '\n', size_t(end - start)));
    if (nl) {  // the whole line is in this block
        p = start, n = size_t(nl - start);
        pos_ = size_t(nl - block_.data()) + 1;
    } else {  // it continues in the next blocks (or ends the input without a '\n')
        span_.assign(start, end);
        pos_ = block_.size();
        while (ra_->take(block_)) {
            pos_ = 0;
            start = block_.data(), end = start + block_.size();
            nl = static_cast<const char *>(memchr(start, '\n', size_t(end - start)));
            if (nl) {
                span_.append(start, nl);
                pos_ = size_t(nl - start) + 1;
                break;
            }
            span_.append(start, end);
            pos_ = block_.size();
        }
        if (span_.empty() && pos_ >= block_.size() && !nl) return false;
        p = span_.data(), n = span_.size();
    }
    if (n && p[n - 1] == '\r') --n;
    return true;
}

bool Reader::find_header() {
    for (;;) {
        if (!have_line_ && !next_line(lp_, ln_)) return false;
        have_line_ = false;
        if (ln_ && (lp_[0] == '>' || lp_[0] == '@')) return true;
    }
}

bool Reader::next(std::string &header, std::vector<uint8_t> &seq) {
    header.clear();
    seq.clear();
    if (!find_header()) return false;
    header.assign(lp_ + 1, lp_ + ln_);
    return next_seq_body(seq, lp_[0] == '@');
}

bool Reader::next_seq(std::vector<uint8_t> &dst) {
    if (!find_header()) return false;
    return next_seq_body(dst, lp_[0] == '@');
}

// The sequence of the record whose header was just read, appended to dst.
bool Reader::next_seq_body(std::vector<uint8_t> &dst, bool fastq) {
    const char *p;
    size_t n;
    if (fastq) {
        if (!next_line(p, n)) return true;
        dst.insert(dst.end(), reinterpret_cast<const uint8_t *>(p), reinterpret_cast<const uint8_t *>(p) + n);
        if (!next_line(p, n)) return true;  // '+'
        if (!next_line(p, n)) return true;  // quality
        return true;
    }
    while (next_line(p, n)) {
        if (n && p[0] == '>') {  // the next record's header: pushed back
            lp_ = p, ln_ = n;
            have_line_ = true;
            break;
        }
        dst.insert(dst.end(), reinterpret_cast<const uint8_t *>(p), reinterpret_cast<const uint8_t *>(p) + n);
    }
    return true;
}

}  // namespace scb_fastx
