/*
 * CPU oracle for the SuperCountingBloom hot path: TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain restatement of the reference crate's CPU algorithm
 * (super-counting-bloom 0.1.0, src/filter.rs, src/api.rs, src/fastx.rs,
 * src/counter.rs, src/config.rs) plus an isolated restatement of the
 * third-party arithmetic it calls (packed-seq 4.4.2, simd-minimizers 2.3.1,
 * seq-hash 0.1.2; see "third-party restatement" in scb_oracle.cpp).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker (or as the timed CPU baseline). The
 * product library (scbf_amd/libscb.so) never links or calls it.
 *
 * Parity status per function is listed in scb_oracle.cpp's header comment.
 */
#ifndef SCB_ORACLE_H
#define SCB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_config {
    uint64_t k, m, s, n_hashes;
    uint8_t counter_bits;          /* 8, 16 or 32 */
    uint8_t counter_slots_exponent;
    uint8_t block_slots_exponent;
    uint8_t _pad[5];
    uint64_t threads, queue;
} oracle_config;

/* ---- primitives ---------------------------------------------------------- */
uint64_t oracle_xorshift_u64(uint64_t x);                       /* filter.rs:629-635 */
/* canonical_kmer over an ASCII ACGT sequence (filter.rs:622-627 + packed-seq F1/F2) */
uint64_t oracle_canonical_kmer(const uint8_t *ascii, size_t len, size_t klen, size_t pos);
uint64_t oracle_read_kmer(const uint8_t *ascii, size_t len, size_t klen, size_t pos);
uint64_t oracle_read_revcomp_kmer(const uint8_t *ascii, size_t len, size_t klen, size_t pos);
/* canonical ntHash of the m-mer at every position (F3); out has len-m+1 entries */
void oracle_nthash_canonical(const uint8_t *ascii, size_t len, size_t m, uint32_t *out);
void oracle_nthash_canonical_direct(const uint8_t *ascii, size_t len, size_t m, uint32_t *out);
/* per-window minimizer position (F4-F6); out has len-k+1 entries */
void oracle_minimizer_positions(const uint8_t *ascii, size_t len, size_t k, size_t m, uint32_t *out);
void oracle_minimizer_positions_brute(const uint8_t *ascii, size_t len, size_t k, size_t m, uint32_t *out);
/* super_kmers (filter.rs:608-620); returns the number of super-k-mers, writes at
 * most cap entries */
size_t oracle_super_kmers(const uint8_t *ascii, size_t len, size_t k, size_t m,
                          uint32_t *starts, uint64_t *minimizers, size_t cap);

/* ---- config -------------------------------------------------------------- */
/* returns 0 if valid, else 1 and writes the reference's message (config.rs:50-107) */
int oracle_config_validate(const oracle_config *cfg, char *msg, size_t msg_cap);

/* ---- filter -------------------------------------------------------------- */
typedef struct oracle_filter oracle_filter;
oracle_filter *oracle_filter_new(const oracle_config *cfg);
void oracle_filter_free(oracle_filter *f);
/* counters in physical (.scb, shard-major) order, widened to u32 */
size_t oracle_filter_counter_slots(const oracle_filter *f);
void oracle_filter_counters_u32(const oracle_filter *f, uint32_t *out);
void oracle_filter_set_counters_u32(oracle_filter *f, const uint32_t *in);
/* SuperCountingBloom::insert_packed_sequence on an all-ACGT sequence
 * (filter.rs:199-237); returns inserted_smers */
uint64_t oracle_filter_insert_packed(oracle_filter *f, const uint8_t *ascii, size_t len);
/* insert_ascii_sequence_typed (api.rs:676-726): run split + insert */
void oracle_filter_insert_ascii(oracle_filter *f, const uint8_t *seq, size_t len,
                                uint64_t *inserted_kmers, uint64_t *inserted_smers);
/* estimate_abundances_packed_sequence on an all-ACGT sequence (filter.rs:337-369,
 * 462-508); out has len-k+1 entries; returns the count written */
size_t oracle_filter_estimate_packed(const oracle_filter *f, const uint8_t *ascii, size_t len,
                                     uint64_t *out);
/* stream_estimate_packed_sequence (filter.rs:371-405, 510-560) */
void oracle_filter_stream_packed(const oracle_filter *f, const uint8_t *ascii, size_t len,
                                 uint64_t *queried, uint64_t *checksum);
/* estimate_sequence_abundances_typed (api.rs:799-853): Option<u64> as value+valid;
 * returns len-k+1 (or 0 if len<k) */
size_t oracle_filter_estimate_ascii(const oracle_filter *f, const uint8_t *seq, size_t len,
                                    uint64_t *values, uint8_t *valid);
/* Brute force: abundance of every window straight from the definition (min over
 * k-s+1 s-mers x n_hashes counters of the window's block); same output shape */
size_t oracle_filter_estimate_ascii_brute(const oracle_filter *f, const uint8_t *seq, size_t len,
                                          uint64_t *values, uint8_t *valid);
/* Record-parallel insert / abundances over n records (bases + n+1 offsets) on
 * `threads` workers, as the worker pools of add_fasta_typed / query_fasta
 * (api.rs:278-313, 621-666): per record, the run split of api.rs:676-726 /
 * 799-853. Abundances are indexed by position (offsets[r] + window); values
 * and valid hold offsets[n] entries. */
void oracle_filter_insert_records(oracle_filter *f, const uint8_t *bases, const uint64_t *offsets,
                                  size_t n, size_t threads, uint64_t *kmers, uint64_t *smers);
void oracle_filter_estimate_records(const oracle_filter *f, const uint8_t *bases,
                                    const uint64_t *offsets, size_t n, size_t threads,
                                    uint64_t *values, uint8_t *valid);
/* the counter table in physical order at its native width (bytes receives its size) */
const void *oracle_filter_data(const oracle_filter *f, size_t *bytes);
/* counter_histogram (filter.rs:267-275): rows sorted by value; returns row count */
size_t oracle_filter_counter_histogram(const oracle_filter *f, uint64_t *values, uint64_t *counts,
                                       size_t cap);

/* ---- distinct-k-mer spectrum (filter.rs:302-335, 407-460; fastx.rs:255-308, 392-410) */
typedef struct oracle_spectrum oracle_spectrum;
oracle_spectrum *oracle_spectrum_new(void);
void oracle_spectrum_free(oracle_spectrum *sp);
/* queries one DnaChunk (all-ACGT run); returns queried k-mers */
uint64_t oracle_spectrum_add_chunk(oracle_spectrum *sp, const oracle_filter *f,
                                   const uint8_t *ascii, size_t len);
/* estimate_distinct_kmers (fastx.rs:255-308) over the DnaChunks of n records on
 * `threads` workers, merged first-wins (fastx.rs:392-400); returns queried k-mers */
uint64_t oracle_spectrum_add_records(oracle_spectrum *sp, const oracle_filter *f, const uint8_t *bases,
                                     const uint64_t *offsets, size_t n, size_t threads);
size_t oracle_spectrum_distinct(const oracle_spectrum *sp);
size_t oracle_spectrum_rows(const oracle_spectrum *sp, uint64_t *values, uint64_t *counts, size_t cap);
/* the merged map as (key, estimate) pairs sorted by key; returns the pair count */
size_t oracle_spectrum_pairs(const oracle_spectrum *sp, uint64_t *keys, uint64_t *values, size_t cap);

/* ---- reservoir-sampled spectrum (bin/sampled_spectrum.rs:150-308) -------- */
/* Sequential reservoir over the valid windows of n records, sampled k-mers
 * queried one by one; rows: 4 u64 each (estimate, sampled, linear scaled
 * windows, estimated distinct); stats: seen, sampled, queried, invalid.
 * Returns the row count. */
size_t oracle_sampled_spectrum(const oracle_filter *f, const uint8_t *bases, const uint64_t *offsets, size_t n,
                               uint64_t sample_size, uint64_t seed, uint64_t *rows, size_t cap, uint64_t *stats);
/* The sampler alone: the 1-based window number each slot holds (windows has
 * min(seen, sample_size) entries); returns the slot count, *seen the windows. */
size_t oracle_reservoir_windows(const uint8_t *bases, const uint64_t *offsets, size_t n, size_t k,
                                uint64_t sample_size, uint64_t seed, uint64_t *windows, uint64_t *seen);

/* ---- .scb serialization (api.rs:388-608) --------------------------------- */
/* returns 0 on success; message written on error */
int oracle_save(const oracle_filter *f, uint64_t inserted_kmers, const char *path, char *msg,
                size_t msg_cap);
/* error kinds: 0 ok, 3 InvalidIndexFormat, 1 InvalidConfig, 7 Io */
int oracle_load(const char *path, oracle_filter **out, oracle_config *cfg, uint64_t *inserted_kmers,
                char *msg, size_t msg_cap);

/* ---- f64 host estimators (api.rs:337-385, 855-891, 933-1033) -------------- */
/* deconvolved_counts; returns count written (0 when the reference returns empty) */
size_t oracle_deconvolved_counts(const uint64_t *values, const uint64_t *counts, size_t rows,
                                 uint64_t n_hashes, uint64_t max_abundance, double *out);
/* estimate_filter_only_spectrum_from_counter_histogram: rows 1..=max */
size_t oracle_filter_only_spectrum(const uint64_t *values, const uint64_t *counts, size_t rows,
                                   uint64_t n_hashes, uint64_t max_abundance,
                                   uint64_t kmer_count_hint, uint64_t *out_counts);
uint64_t oracle_filter_only_count_hint(const uint64_t *values, const uint64_t *counts, size_t rows,
                                       uint64_t n_hashes, uint64_t max_abundance);
/* summarize_abundances: returns valid count; mean/median NaN when none */
size_t oracle_summarize(const uint64_t *values, const uint8_t *valid, size_t n, double *mean,
                        double *median);

/* ---- restated multithreaded CPU path (fastx.rs:206-369, api.rs:621-674) --- */
typedef struct oracle_baseline_result {
    uint64_t chunks, indexed_kmers, inserted_smers, queried_kmers, positive_estimates,
        estimate_checksum;
    double build_s, query_s;
} oracle_baseline_result;
/* Build the filter from the index records, then stream-query the query records,
 * with 1 producer thread, a bounded(queue) channel of owned records and
 * cfg->threads workers, exactly as stream_estimates_from_fastx_pair. Records are
 * given as concatenated bytes + offsets (n+1 entries) and are split into maximal
 * ACGT runs (the split_non_actg chunks) by the producer. */
int oracle_cpu_baseline(const oracle_config *cfg, const uint8_t *index_bases,
                        const uint64_t *index_offsets, size_t n_index, const uint8_t *query_bases,
                        const uint64_t *query_offsets, size_t n_query, oracle_baseline_result *out);

#ifdef __cplusplus
}
#endif
#endif
