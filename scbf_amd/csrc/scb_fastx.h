// FASTA/FASTQ input for the host pipelines (replaces helicase 0.1.1's
// FastxParser, used at api.rs:893-927 and fastx.rs:14-19, 228-243): plain,
// gzip and zstd input detected by their magic bytes, decompressed by a
// read-ahead thread while the caller parses records.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace scb_fastx {

class Source;
class ReadAhead;

class Reader {
   public:
    Reader();
    ~Reader();
    // SCB_OK, or SCB_ERR_FASTX_OPEN with the reference's message.
    int open(const char *path);
    // Next record (header without '>'/'@', sequence lines joined); false at
    // the end of input or on a read error (then error() is nonzero).
    bool next(std::string &header, std::vector<uint8_t> &seq);
    // The same record walk, appending only the sequence to dst (no header, no
    // intermediate copies of lines that lie in one read-ahead block).
    bool next_seq(std::vector<uint8_t> &dst);
    // SCB_OK, or the error code of a failed read (its message becomes the
    // calling thread's last error).
    int error();
    const std::string &path() const { return path_; }

   private:
    // The next line without its '\n' (and a trailing '\r'): a view into the
    // current block, or into span_ for a line that crosses blocks; valid until
    // the next call. False at the end of input.
    bool next_line(const char *&p, size_t &n);
    // Reads up to a header line (the pushed-back one first); false at the end.
    bool find_header();
    bool next_seq_body(std::vector<uint8_t> &dst, bool fastq);
    std::string path_;
    std::unique_ptr<ReadAhead> ra_;
    std::string span_;
    const char *lp_ = nullptr;  // the current (or pushed-back) line
    size_t ln_ = 0;
    bool have_line_ = false;
    std::vector<char> block_;
    size_t pos_ = 0;
};

}  // namespace scb_fastx
