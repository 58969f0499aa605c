/*
 * scb.h: C ABI of the MI355X SuperCountingBloom engine (libscb.so).
 *
 * Drop-in boundary for the hot path of the Rust crate super-counting-bloom
 * 0.1.0 (insert + query + spectrum). Each entry point names the reference
 * interface it replaces (file:line in the reference crate). A Rust crate would
 * bind these symbols through `extern "C"`; the C++ wrapper
 * in scb.hpp mirrors the crate's types and method names on top of them.
 *
 * Conventions
 *  - Every fallible call returns int: SCB_OK (0), or 1..7 mapping one-to-one to
 *    the variants of SuperCountingBloomError (error.rs:3-11), or SCB_ERR_DEVICE
 *    (8, extension) for a HIP runtime failure. scb_last_error() returns the
 *    message of the last failure on the calling thread, formatted like the
 *    reference's Display impl (error.rs:13-27).
 *  - Handles are owned by the caller and freed with *_free. Output buffers are
 *    caller-allocated.
 *  - Builder calls must be serialized by the caller (the Rust API takes
 *    &mut self, api.rs:123,140). Index calls are safe to call concurrently
 *    (the Rust API takes &self, api.rs:250,273): each call uses its own
 *    device workspace.
 *  - Calls block until their results are in the caller's memory. The *_device
 *    variants take device pointers (HBM-resident inputs and outputs).
 *  - Batched inputs are concatenated bytes plus n+1 offsets (offsets[0] = 0,
 *    offsets[n] = total bytes). Per-window outputs of the batched calls are
 *    indexed by position: window j of record i is at offsets[i] + j.
 */
#ifndef SCB_H
#define SCB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error.rs:3-11 (+ SCB_ERR_DEVICE, an extension for HIP runtime failures) */
enum {
    SCB_OK = 0,
    SCB_ERR_INVALID_CONFIG = 1,
    SCB_ERR_FASTX_OPEN = 2,
    SCB_ERR_INVALID_INDEX_FORMAT = 3,
    SCB_ERR_CHANNEL_CLOSED = 4,
    SCB_ERR_WORKER_PANIC = 5,
    SCB_ERR_INTERNAL_STATE = 6,
    SCB_ERR_IO = 7,
    SCB_ERR_DEVICE = 8,
    SCB_ERR_BUSY = 9 /* serving loop: the request queue is full (admission control) */
};

/* Insert/query/spectrum execution method (extension). BATCHED: super-k-mers
 * binned by counter block and applied block-by-block in LDS. PLAIN: one fused
 * launch per batch, one thread per super-k-mer, global atomics / loads. AUTO
 * (the default): PLAIN below a batch size that grows with the table
 * (scb_builder_set_method), BATCHED above. All are bit-identical. */
enum { SCB_METHOD_BATCHED = 0, SCB_METHOD_PLAIN = 1, SCB_METHOD_AUTO = 2 };

/* SuperCountingBloomConfig (config.rs:20-31); counter_bits is 8, 16 or 32. */
typedef struct scb_config {
    uint64_t k, m, s, n_hashes;
    uint8_t counter_bits;
    uint8_t counter_slots_exponent;
    uint8_t block_slots_exponent;
    uint8_t _pad[5];
    uint64_t threads, queue;
} scb_config;

/* AddReport (api.rs:66-73) */
typedef struct scb_add_report {
    uint64_t records_processed, records_indexed, inserted_kmers, inserted_smers;
    double add_seconds;
} scb_add_report;

/* QueryFastaReport (api.rs:75-82) */
typedef struct scb_query_report {
    uint64_t records_processed, total_windows, valid_kmers, positive_estimates, estimate_checksum;
} scb_query_report;

/* BuildIndexReport without the index (api.rs:56-64) */
typedef struct scb_build_report {
    uint64_t chunks, indexed_kmers, inserted_smers, counter_slots, blocks;
    double build_seconds;
} scb_build_report;

/* SuperCountingBloomStats (fastx.rs:21-34) */
typedef struct scb_spectrum_stats {
    uint64_t chunks, indexed_kmers, inserted_smers, queried_kmers, distinct_keys, counter_slots,
        blocks;
    double build_seconds, query_seconds, spectrum_seconds, total_seconds;
} scb_spectrum_stats;

/* StreamEstimateStats (fastx.rs:42-55) */
typedef struct scb_stream_stats {
    uint64_t chunks, indexed_kmers, inserted_smers, queried_kmers, positive_estimates,
        estimate_checksum, counter_slots, blocks;
    double build_seconds, query_seconds, total_seconds;
} scb_stream_stats;

/* SequenceAbundanceSummary (api.rs:84-92); mean/median are NaN for None */
typedef struct scb_summary {
    uint64_t sequence_len, total_windows, valid_kmers;
    double mean_abundance, median_abundance;
} scb_summary;

typedef struct scb_builder scb_builder; /* SuperCountingBloomBuilder (api.rs:43-48) */
typedef struct scb_index scb_index;     /* SuperCountingBloomIndex (api.rs:50-54) */

/* ---- errors and config --------------------------------------------------- */
/* Copies the calling thread's last error message (NUL-terminated, truncated to
 * cap); returns its full length. */
size_t scb_last_error(char *buf, size_t cap);
/* SuperCountingBloomConfig::default (config.rs:33-47); threads = online CPUs */
void scb_config_default(scb_config *cfg);
/* SuperCountingBloomConfig::validate (config.rs:50-92) */
int scb_config_validate(const scb_config *cfg);
/* number of HIP devices visible (0 without a GPU; never fails) */
int scb_device_count(void);

/* extension: device milliseconds of the calling thread's last insert or query
 * call, measured with HIP events on the engine's stream: ms[0] scan kernel,
 * ms[1] bin planning + scatter, ms[2] per-bin kernel, ms[3] total (the plain
 * method reports its single fused kernel in ms[0] and ms[3]). */
void scb_last_stage_times(double *ms);

/* extension: tunables of the batched pipeline (no counterpart in the
 * reference; every valid setting gives identical results). An engine starts
 * from defaults for its geometry, replaced by the matching entry of the tuning
 * cache: the JSON file named by SCB_TUNING_CACHE, else tuning.json beside
 * libscb.so, or one loaded with scb_tuning_cache_load (written by the
 * autotuner scbf_amd/tune.py). */
typedef struct scb_tuning {
    uint32_t insert_bin_log;        /* slots per insert bin, log2: [max(bse,1), min(15,cse)] */
    uint32_t query_bin_log;         /* slots per query bin, log2: [bse, cse], bin <= 128 KiB */
    uint32_t spectrum_bin_log;      /* slots per spectrum bin, log2 (used when bse <= 14) */
    uint32_t insert_item_records;   /* records per insert work item: [1, 2048] */
    uint32_t query_item_records;    /* records per query work item: [1, 3584] */
    uint32_t spectrum_item_records; /* records per spectrum work item: [1, 4096] */
    uint32_t use_graph;             /* 1: replay repeated device-batch insert/query pipelines
                                       from a captured HIP graph */
    uint32_t _pad;
} scb_tuning;
/* Replaces the process-wide tuning cache (engines created afterwards use it). */
int scb_tuning_cache_load(const char *path);

/* ---- builder ------------------------------------------------------------- */
/* SuperCountingBloomBuilder::new (api.rs:102-121): allocates and zeroes the
 * 2^counter_slots_exponent-counter table in HBM. */
int scb_builder_new(const scb_config *cfg, scb_builder **out);
void scb_builder_free(scb_builder *b);
/* SuperCountingBloomBuilder::add_sequence (api.rs:123-138). *inserted_kmers
 * (the sequence's valid k-mer windows) is counted on the host at once; the
 * sequence joins a pending host batch that goes to the device when it holds
 * 64 MiB or 2^20 sequences, or before any call that reads the builder
 * (inserted_smers, counters, into_index, the batch and FASTA inserts). A
 * device error of a pending batch is returned by that later call. */
int scb_builder_add_sequence(scb_builder *b, const uint8_t *seq, size_t len,
                             uint64_t *inserted_kmers);
/* Batched add_sequence over n host records (api.rs:123-138 per record;
 * AddReport fields as add_fasta_typed sums them, api.rs:633-662) */
int scb_builder_add_sequences(scb_builder *b, const uint8_t *bases, const uint64_t *offsets,
                              size_t n, scb_add_report *report);
/* Same with HBM-resident inputs: d_offsets holds n+1 device u64 with
 * d_offsets[0] = 0, never decreasing, d_offsets[n] = total_bases < 2^32 (checked
 * on the device: a violation returns SCB_ERR_INTERNAL_STATE and the batch's
 * counts are not defined). records_* fields of the report are not computed
 * (left 0). */
int scb_builder_add_sequences_device(scb_builder *b, const uint8_t *d_bases,
                                     const uint64_t *d_offsets, size_t n, uint64_t total_bases,
                                     scb_add_report *report);
/* SuperCountingBloomBuilder::add_fasta (api.rs:140-158) */
int scb_builder_add_fasta(scb_builder *b, const char *path, scb_add_report *report);
/* SuperCountingBloomBuilder::into_index (api.rs:160-177): consumes the builder
 * (it is freed whatever the result). */
int scb_builder_into_index(scb_builder *b, scb_index **out);
/* api.rs:179-207 */
uint64_t scb_builder_inserted_kmers(const scb_builder *b);
uint64_t scb_builder_inserted_smers(const scb_builder *b);
int scb_builder_set_threads(scb_builder *b, size_t threads);
void scb_builder_clear_threads(scb_builder *b);
size_t scb_builder_threads(const scb_builder *b);
void scb_builder_config(const scb_builder *b, scb_config *out);
/* extension: the device method. SCB_METHOD_AUTO (default) runs the plain
 * method (one thread per super-k-mer, global counter accesses) for batches
 * below 2^17 bases + one base per 256 table bytes, where its lower fixed cost
 * wins, and the batched method (sort by counter bin, bins staged in LDS)
 * above; SCB_METHOD_BATCHED and SCB_METHOD_PLAIN force one. Results are the
 * same with every method. The index takes the builder's method; see
 * scb_index_set_method. */
int scb_builder_set_method(scb_builder *b, int method);
/* extension: copy the counter table (physical = .scb payload order) to host */
int scb_builder_counters(const scb_builder *b, void *out, size_t bytes);
/* extension: the engine's tuning (an invalid setting is SCB_ERR_INVALID_CONFIG) */
int scb_builder_set_tuning(scb_builder *b, const scb_tuning *t);
void scb_builder_tuning(const scb_builder *b, scb_tuning *out);

/* ---- index --------------------------------------------------------------- */
/* SuperCountingBloomIndex::build_from_fastx (api.rs:211-223; fastx.rs:206-253,
 * 378-390); report may be NULL */
int scb_index_build_from_fastx(const char *path, const scb_config *cfg, scb_index **out,
                               scb_build_report *report);
void scb_index_free(scb_index *idx);
void scb_index_config(const scb_index *idx, scb_config *out);   /* api.rs:225-227 */
uint64_t scb_index_inserted_kmers(const scb_index *idx);         /* api.rs:229-231 */
int scb_index_set_method(scb_index *idx, int method);            /* extension */
int scb_index_counters(const scb_index *idx, void *out, size_t bytes); /* extension */
int scb_index_set_tuning(scb_index *idx, const scb_tuning *t);       /* extension */
void scb_index_tuning(const scb_index *idx, scb_tuning *out);         /* extension */
/* SuperCountingBloomIndex::save / load (api.rs:233-248, 388-608): byte-exact
 * .scb files (139-byte header, then shard-major little-endian counters) */
int scb_index_save(const scb_index *idx, const char *path);
int scb_index_load(const char *path, scb_index **out);
/* SuperCountingBloomIndex::estimate_sequence_abundances (api.rs:250-262,
 * 799-853): writes len-k+1 entries (none if len<k) to values/valid (valid=0
 * is None); *n_windows receives the count. */
int scb_index_estimate_sequence_abundances(const scb_index *idx, const uint8_t *seq, size_t len,
                                           uint64_t *values, uint8_t *valid, size_t *n_windows);
/* Batched estimate_sequence_abundances over n host records: values/valid are
 * indexed by position (offsets[i] + j, j < len_i-k+1) and hold
 * offsets[n] entries; report (may be NULL) is the QueryFastaReport of the batch
 * (api.rs:283-296). */
int scb_index_estimate_sequences(const scb_index *idx, const uint8_t *bases,
                                 const uint64_t *offsets, size_t n, uint64_t *values,
                                 uint8_t *valid, scb_query_report *report);
/* Same with HBM-resident inputs and outputs (d_offsets as for
 * scb_builder_add_sequences_device, checked the same way): d_values holds
 * total_bases counters of counter_bits/8 bytes each, d_valid_bits total_bases bits
 * (rounded up to 32-bit words); either may be NULL to skip it (then only the
 * report is produced: the stream path of fastx.rs:310-369). report's
 * records_processed/total_windows are not computed (left 0). */
int scb_index_estimate_sequences_device(const scb_index *idx, const uint8_t *d_bases,
                                        const uint64_t *d_offsets, size_t n,
                                        uint64_t total_bases, void *d_values,
                                        uint32_t *d_valid_bits, scb_query_report *report);
/* No counterpart in the reference. scb_index_estimate_sequences_device for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc. Same arguments and
 * results; its kernels take every per-call buffer from a fixed device array of
 * SCB_NOALLOC_SCRATCH_BYTES that each GPU holds from its first call until the
 * process ends, and the call makes no hipMalloc or hipFree. total_bases may be
 * at most SCB_NOALLOC_MAX_BASES (larger batches are refused with
 * SCB_ERR_INTERNAL_STATE and a message). Calls on one GPU share that GPU's
 * array: the library serializes them with a lock (one call uses the array at a
 * time), each runs on the null stream and returns after it has synchronized,
 * and the array keeps the last call's contents. */
#define SCB_NOALLOC_SCRATCH_BYTES ((uint64_t)96 << 20)
#define SCB_NOALLOC_MAX_BASES ((uint64_t)1 << 22)
int scb_index_estimate_sequences_device_noalloc(const scb_index *idx, const uint8_t *d_bases,
                                                const uint64_t *d_offsets, size_t n,
                                                uint64_t total_bases, void *d_values,
                                                uint32_t *d_valid_bits, scb_query_report *report);
/* SuperCountingBloomIndex::summarize_sequence (api.rs:264-271, 855-891) */
int scb_index_summarize_sequence(const scb_index *idx, const uint8_t *seq, size_t len,
                                 scb_summary *out);
/* SuperCountingBloomIndex::query_fasta (api.rs:273-327) */
int scb_index_query_fasta(const scb_index *idx, const char *path, scb_query_report *report);
/* SuperCountingBloomIndex::counter_histogram (api.rs:329-335; filter.rs:267-275):
 * rows sorted by value; *n_rows receives the row count (rows beyond cap are not
 * written). */
int scb_index_counter_histogram(const scb_index *idx, uint64_t *values, uint64_t *counts,
                                size_t cap, size_t *n_rows);
/* Distinct-k-mer spectrum of a batch of query records against the index
 * (fastx.rs:255-308, 392-410 with the DnaChunk split of fastx.rs:14-19):
 * rows (estimate, #distinct canonical k-mers) sorted by estimate. *queried
 * receives the number of queried k-mers, *distinct the number of keys. Host
 * input of any size (beyond one device batch it is counted chunk by chunk in
 * a key set); the _device form takes one batch of fewer than 2^32 bases. */
int scb_index_spectrum(const scb_index *idx, const uint8_t *bases, const uint64_t *offsets,
                       size_t n, uint64_t *values, uint64_t *counts, size_t cap, size_t *n_rows,
                       uint64_t *queried, uint64_t *distinct);
int scb_index_spectrum_device(const scb_index *idx, const uint8_t *d_bases,
                              const uint64_t *d_offsets, size_t n, uint64_t total_bases,
                              uint64_t *values, uint64_t *counts, size_t cap, size_t *n_rows,
                              uint64_t *queried, uint64_t *distinct);
/* FrozenSuperCountingBloom::estimate_packed_sequence (filter.rs:302-335, 407-460)
 * over the DnaChunks of n host records (maximal ACGT runs of at least k bases):
 * every distinct canonical k-mer with its estimate (the HashMap<KmerKey, C> of
 * the reference, first occurrence wins; all occurrences of a key have the same
 * estimate), as pairs sorted by key. *n_keys receives the number of keys (pairs
 * beyond cap are not written), *queried the number of k-mers queried. */
int scb_index_estimate_kmer_map(const scb_index *idx, const uint8_t *bases, const uint64_t *offsets,
                                size_t n, uint64_t *keys, uint64_t *values, size_t cap,
                                size_t *n_keys, uint64_t *queried);

/* Distinct-k-mer spectrum over many batches (estimate_distinct_kmers,
 * fastx.rs:255-308, with its per-worker maps merged first-wins, 392-400): a
 * device key set that grows as needed; a key is counted once over all batches.
 * Rows are (estimate, #distinct keys) sorted by estimate, as scb_index_spectrum.
 * The index must outlive the accumulator; calls on one accumulator are
 * serialized by the caller. */
typedef struct scb_spectrum_acc scb_spectrum_acc;
int scb_spectrum_acc_new(const scb_index *idx, scb_spectrum_acc **out);
void scb_spectrum_acc_free(scb_spectrum_acc *acc);
int scb_spectrum_acc_add(scb_spectrum_acc *acc, const uint8_t *bases, const uint64_t *offsets, size_t n);
int scb_spectrum_acc_add_device(scb_spectrum_acc *acc, const uint8_t *d_bases, const uint64_t *d_offsets,
                                size_t n, uint64_t total_bases);
int scb_spectrum_acc_rows(const scb_spectrum_acc *acc, uint64_t *values, uint64_t *counts, size_t cap,
                          size_t *n_rows, uint64_t *queried, uint64_t *distinct);
/* estimate_filter_only_spectrum(_from_counter_histogram) (api.rs:337-368,
 * 933-1033); writes rows 1..=max_abundance to counts (counts[a-1]). When
 * hist_values is NULL the index's counter_histogram is used. */
int scb_index_filter_only_spectrum(const scb_index *idx, const uint64_t *hist_values,
                                   const uint64_t *hist_counts, size_t hist_rows,
                                   uint64_t max_abundance, uint64_t kmer_count_hint,
                                   uint64_t *counts);
/* filter_only_count_hint(_from_counter_histogram) (api.rs:370-385); max_abundance
 * is 255 for filter_only_count_hint(). */
int scb_index_filter_only_count_hint(const scb_index *idx, const uint64_t *hist_values,
                                     const uint64_t *hist_counts, size_t hist_rows,
                                     uint64_t max_abundance, uint64_t *hint);

/* The same estimators on a given histogram and n_hashes, with no index and no
 * device (extension; host arithmetic only): api.rs:933-1033 and 375-385. */
int scb_filter_only_spectrum_from_histogram(uint64_t n_hashes, const uint64_t *hist_values,
                                            const uint64_t *hist_counts, size_t hist_rows,
                                            uint64_t max_abundance, uint64_t kmer_count_hint,
                                            uint64_t *counts);
int scb_filter_only_count_hint_from_histogram(uint64_t n_hashes, const uint64_t *hist_values,
                                              const uint64_t *hist_counts, size_t hist_rows,
                                              uint64_t max_abundance, uint64_t *hint);
/* summarize_abundances (api.rs:855-891) over n windows given as values + validity
 * (valid[i] == 0 is None); sequence_len is left for the caller. Host only. */
int scb_summarize_abundances(const uint64_t *values, const uint8_t *valid, size_t n, scb_summary *out);

/* ---- FASTX pipelines (fastx.rs:81-204) ------------------------------------ */
/* estimate_spectrum_from_fastx_pair (fastx.rs:88-105, 133-170); query_path may
 * equal index_path (estimate_spectrum_from_fastx, fastx.rs:81-86) */
int scb_estimate_spectrum_from_fastx_pair(const char *index_path, const char *query_path,
                                          const scb_config *cfg, uint64_t *values,
                                          uint64_t *counts, size_t cap, size_t *n_rows,
                                          scb_spectrum_stats *stats);
/* stream_estimates_from_fastx_pair (fastx.rs:114-131, 172-204) */
int scb_stream_estimates_from_fastx_pair(const char *index_path, const char *query_path,
                                         const scb_config *cfg, scb_stream_stats *stats);
/* write_spectrum (spectrum.rs:11-31): TSV with header "estimated_count\tkmers";
 * path NULL writes to stdout */
int scb_write_spectrum(const uint64_t *values, const uint64_t *counts, size_t n_rows,
                       const char *path);
/* for_each_fastx_record (api.rs:893-927): calls visit(user, name, seq, len) per
 * record (keep_non_actg); a nonzero return stops the walk and is returned. */
typedef int (*scb_record_visitor)(void *user, const char *name, const uint8_t *seq, size_t len);
int scb_for_each_fastx_record(const char *path, scb_record_visitor visit, void *user);

/* ---- reservoir-sampled spectrum (bin/sampled_spectrum.rs:150-308) ---------- */
/* A row of write_scaled_spectrum (sampled_spectrum.rs:225-253). */
typedef struct scb_sampled_row {
    uint64_t estimated_count, sampled_kmers, linear_scaled_windows, estimated_distinct_kmers;
} scb_sampled_row;
/* The CLI's stderr metrics (sampled_spectrum.rs:113-146). Sampling and
 * querying are one fused device pass here (sample_seconds covers both). */
typedef struct scb_sampled_stats {
    uint64_t indexed_kmers, inserted_smers, total_inserted_kmers, sampled_kmers, queried_kmers,
        invalid_queries;
    double build_seconds, sample_seconds;
} scb_sampled_stats;
/* sample_inserted_kmers + query_sampled_kmers (sampled_spectrum.rs:150-221)
 * over n host records against the index: a reservoir of sample_size valid
 * k-mer windows in input order (SplitMix64 from seed), each sampled k-mer's
 * estimate, rows sorted by estimate. sample_size 0 is SCB_ERR_INVALID_CONFIG. */
int scb_index_sampled_spectrum(const scb_index *idx, const uint8_t *bases, const uint64_t *offsets,
                               size_t n, uint64_t sample_size, uint64_t seed, scb_sampled_row *rows,
                               size_t cap, size_t *n_rows, scb_sampled_stats *stats);
/* run() of bin/sampled_spectrum.rs:99-148: build the index from the FASTA/FASTQ
 * file, then sample and query its valid windows. */
int scb_sampled_spectrum_from_fastx(const char *path, const scb_config *cfg, uint64_t sample_size,
                                    uint64_t seed, scb_sampled_row *rows, size_t cap, size_t *n_rows,
                                    scb_sampled_stats *stats);
/* CSV "estimated_count,sampled_kmers,linear_scaled_windows,estimated_distinct_kmers"
 * (sampled_spectrum.rs:228-252); path NULL writes to stdout */
int scb_write_sampled_spectrum(const scb_sampled_row *rows, size_t n_rows, const char *path);

/* ---- serving loop --------------------------------------------------------
 * No counterpart in the reference. A frozen index stays resident in HBM and
 * answers concurrent requests: worker threads coalesce queued abundance
 * requests into one device launch (up to max_batch_bases / max_batch_requests),
 * run spectrum requests one per launch, and refuse new requests with
 * SCB_ERR_BUSY once max_pending are queued. Each request's results equal
 * scb_index_estimate_sequences / scb_index_spectrum on the same records
 * (api.rs:250-262, 273-327; fastx.rs:133-170 per request). */
typedef struct scb_server scb_server;

enum { SCB_REQ_ABUNDANCES = 0, SCB_REQ_SPECTRUM = 1 };

typedef struct scb_serve_config {
    uint64_t max_batch_bases;    /* bases coalesced into one launch (default 2^27) */
    uint64_t max_batch_requests; /* requests coalesced into one launch (default 1024) */
    uint64_t hbm_budget_bytes;   /* device memory for staging and pipeline scratch of all
                                    workers (default 16 GiB); max_batch_bases is lowered to fit */
    uint32_t workers;            /* launches in flight at once (default 4) */
    uint32_t max_pending;        /* queued requests before submit answers SCB_ERR_BUSY (4096) */
} scb_serve_config;

typedef struct scb_served {
    int status;                   /* SCB_OK or the error code of this request */
    uint32_t kind;                /* SCB_REQ_ABUNDANCES or SCB_REQ_SPECTRUM */
    scb_query_report report;      /* abundances: as scb_index_estimate_sequences */
    uint64_t n_rows, queried, distinct;  /* spectrum: as scb_index_spectrum */
    uint64_t submit_ns, start_ns, done_ns;  /* CLOCK_MONOTONIC nanoseconds */
    uint64_t batch_requests;      /* requests served by the same launch */
    char message[192];            /* error text when status != SCB_OK */
} scb_served;

void scb_serve_config_default(scb_serve_config *cfg);
/* The index must outlive the server. */
int scb_server_new(const scb_index *idx, const scb_serve_config *cfg, scb_server **out);
/* Finishes the queued requests, then stops the workers. */
void scb_server_free(scb_server *srv);
/* values: total_bases counters of counter_bits/8 bytes, indexed by position over
 * the concatenated records, 0 where no valid k-mer window starts; valid_bits:
 * (total_bases + 31) / 32 words, bit p % 32 of word p / 32 set for a valid
 * window at p. Inputs and outputs must stay valid until the ticket is waited
 * for. Records longer than max_batch_bases - 64 are refused. */
int scb_server_submit_abundances(scb_server *srv, const uint8_t *bases, const uint64_t *offsets,
                                 size_t n, void *values, uint32_t *valid_bits, uint64_t *ticket);
/* Spectrum rows (value, #distinct keys) as scb_index_spectrum. */
int scb_server_submit_spectrum(scb_server *srv, const uint8_t *bases, const uint64_t *offsets,
                               size_t n, uint64_t *values, uint64_t *counts, size_t cap,
                               uint64_t *ticket);
/* Blocks until the request has finished; each ticket is waited for once. A
 * ticket that another scb_server_wait call is blocked on is refused
 * (SCB_ERR_INTERNAL_STATE), and scb_server_wait_any never returns it. */
int scb_server_wait(scb_server *srv, uint64_t ticket, scb_served *out);
/* Waits up to timeout_us for any finished request not yet waited for and not
 * claimed by a blocked scb_server_wait call; *ticket is 0 when none finished in
 * time. */
int scb_server_wait_any(scb_server *srv, uint64_t timeout_us, uint64_t *ticket, scb_served *out);
/* Page-locked host memory for request buffers (faster host<->HBM copies). */
void *scb_host_alloc(size_t bytes);
void scb_host_free(void *p);

/* ---- low-level primitives (filter.rs:622-635) ----------------------------- */
uint64_t scb_xorshift_u64(uint64_t x);
/* canonical_kmer over an ASCII ACGT sequence (2-bit packed as packed-seq) */
uint64_t scb_canonical_kmer(const uint8_t *seq, size_t len, size_t klen, size_t pos);

#ifdef __cplusplus
}
#endif
#endif
