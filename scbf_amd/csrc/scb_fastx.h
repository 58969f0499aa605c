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
    // SCB_OK, or the error code of a failed read (its message becomes the
    // calling thread's last error).
    int error();
    const std::string &path() const { return path_; }

   private:
    bool getline(std::string &out);
    std::string path_;
    std::unique_ptr<ReadAhead> ra_;
    std::string line_;
    bool have_line_ = false;
    std::vector<char> block_;
    size_t pos_ = 0;
};

}  // namespace scb_fastx
