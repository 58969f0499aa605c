"""ctypes binding of libscb.so (include/scb.h). Loads the in-tree library and
fails loudly when it is missing: there is no CPU fallback."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libscb.so")

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
szp = C.POINTER(C.c_size_t)


class ScbConfig(C.Structure):
    _fields_ = [("k", C.c_uint64), ("m", C.c_uint64), ("s", C.c_uint64), ("n_hashes", C.c_uint64),
                ("counter_bits", C.c_uint8), ("counter_slots_exponent", C.c_uint8),
                ("block_slots_exponent", C.c_uint8), ("_pad", C.c_uint8 * 5),
                ("threads", C.c_uint64), ("queue", C.c_uint64)]


class ScbAddReport(C.Structure):
    _fields_ = [("records_processed", C.c_uint64), ("records_indexed", C.c_uint64),
                ("inserted_kmers", C.c_uint64), ("inserted_smers", C.c_uint64),
                ("add_seconds", C.c_double)]


class ScbQueryReport(C.Structure):
    _fields_ = [("records_processed", C.c_uint64), ("total_windows", C.c_uint64),
                ("valid_kmers", C.c_uint64), ("positive_estimates", C.c_uint64),
                ("estimate_checksum", C.c_uint64)]


class ScbBuildReport(C.Structure):
    _fields_ = [("chunks", C.c_uint64), ("indexed_kmers", C.c_uint64), ("inserted_smers", C.c_uint64),
                ("counter_slots", C.c_uint64), ("blocks", C.c_uint64), ("build_seconds", C.c_double)]


class ScbSpectrumStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("chunks", "indexed_kmers", "inserted_smers", "queried_kmers",
                                          "distinct_keys", "counter_slots", "blocks")] + \
               [(n, C.c_double) for n in ("build_seconds", "query_seconds", "spectrum_seconds",
                                          "total_seconds")]


class ScbStreamStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("chunks", "indexed_kmers", "inserted_smers", "queried_kmers",
                                          "positive_estimates", "estimate_checksum", "counter_slots",
                                          "blocks")] + \
               [(n, C.c_double) for n in ("build_seconds", "query_seconds", "total_seconds")]


class ScbSummary(C.Structure):
    _fields_ = [("sequence_len", C.c_uint64), ("total_windows", C.c_uint64), ("valid_kmers", C.c_uint64),
                ("mean_abundance", C.c_double), ("median_abundance", C.c_double)]


class ScbSampledRow(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("estimated_count", "sampled_kmers", "linear_scaled_windows",
                                          "estimated_distinct_kmers")]


class ScbSampledStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("indexed_kmers", "inserted_smers", "total_inserted_kmers",
                                          "sampled_kmers", "queried_kmers", "invalid_queries")] + \
               [("build_seconds", C.c_double), ("sample_seconds", C.c_double)]


class ScbTuning(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("insert_bin_log", "query_bin_log", "spectrum_bin_log",
                                          "insert_item_records", "query_item_records", "spectrum_item_records",
                                          "use_graph", "_pad")]


class ScbServeConfig(C.Structure):
    _fields_ = [("max_batch_bases", C.c_uint64), ("max_batch_requests", C.c_uint64),
                ("hbm_budget_bytes", C.c_uint64), ("workers", C.c_uint32), ("max_pending", C.c_uint32)]


class ScbServed(C.Structure):
    _fields_ = [("status", C.c_int), ("kind", C.c_uint32), ("report", ScbQueryReport),
                ("n_rows", C.c_uint64), ("queried", C.c_uint64), ("distinct", C.c_uint64),
                ("submit_ns", C.c_uint64), ("start_ns", C.c_uint64), ("done_ns", C.c_uint64),
                ("batch_requests", C.c_uint64), ("message", C.c_char * 192)]


VISITOR = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p, u8p, C.c_size_t)

# name -> (restype, argtypes)
SIGNATURES = {
    "scb_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
    "scb_config_default": (None, [C.POINTER(ScbConfig)]),
    "scb_config_validate": (C.c_int, [C.POINTER(ScbConfig)]),
    "scb_device_count": (C.c_int, []),
    "scb_last_stage_times": (None, [C.POINTER(C.c_double)]),
    "scb_builder_new": (C.c_int, [C.POINTER(ScbConfig), C.POINTER(C.c_void_p)]),
    "scb_builder_free": (None, [C.c_void_p]),
    "scb_builder_add_sequence": (C.c_int, [C.c_void_p, u8p, C.c_size_t, u64p]),
    "scb_builder_add_sequences": (C.c_int, [C.c_void_p, u8p, u64p, C.c_size_t, C.POINTER(ScbAddReport)]),
    "scb_builder_add_sequences_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                                   C.c_uint64, C.POINTER(ScbAddReport)]),
    "scb_builder_add_fasta": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(ScbAddReport)]),
    "scb_builder_into_index": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "scb_builder_inserted_kmers": (C.c_uint64, [C.c_void_p]),
    "scb_builder_inserted_smers": (C.c_uint64, [C.c_void_p]),
    "scb_builder_set_threads": (C.c_int, [C.c_void_p, C.c_size_t]),
    "scb_builder_clear_threads": (None, [C.c_void_p]),
    "scb_builder_threads": (C.c_size_t, [C.c_void_p]),
    "scb_builder_config": (None, [C.c_void_p, C.POINTER(ScbConfig)]),
    "scb_builder_set_method": (C.c_int, [C.c_void_p, C.c_int]),
    "scb_builder_counters": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "scb_index_build_from_fastx": (C.c_int, [C.c_char_p, C.POINTER(ScbConfig), C.POINTER(C.c_void_p),
                                             C.POINTER(ScbBuildReport)]),
    "scb_index_free": (None, [C.c_void_p]),
    "scb_index_config": (None, [C.c_void_p, C.POINTER(ScbConfig)]),
    "scb_index_inserted_kmers": (C.c_uint64, [C.c_void_p]),
    "scb_index_set_method": (C.c_int, [C.c_void_p, C.c_int]),
    "scb_index_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(ScbTuning)]),
    "scb_index_tuning": (None, [C.c_void_p, C.POINTER(ScbTuning)]),
    "scb_builder_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(ScbTuning)]),
    "scb_builder_tuning": (None, [C.c_void_p, C.POINTER(ScbTuning)]),
    "scb_tuning_cache_load": (C.c_int, [C.c_char_p]),
    "scb_index_counters": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "scb_index_save": (C.c_int, [C.c_void_p, C.c_char_p]),
    "scb_index_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "scb_index_estimate_sequence_abundances": (C.c_int, [C.c_void_p, u8p, C.c_size_t, u64p, u8p, szp]),
    "scb_index_estimate_sequences": (C.c_int, [C.c_void_p, u8p, u64p, C.c_size_t, u64p, u8p,
                                               C.POINTER(ScbQueryReport)]),
    "scb_index_estimate_sequences_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                                      C.c_uint64, C.c_void_p, C.c_void_p,
                                                      C.POINTER(ScbQueryReport)]),
    "scb_index_estimate_sequences_device_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                                              C.c_uint64, C.c_void_p, C.c_void_p,
                                                              C.POINTER(ScbQueryReport)]),
    "scb_index_summarize_sequence": (C.c_int, [C.c_void_p, u8p, C.c_size_t, C.POINTER(ScbSummary)]),
    "scb_summarize_abundances": (C.c_int, [u64p, u8p, C.c_size_t, C.POINTER(ScbSummary)]),
    "scb_index_estimate_kmer_map": (C.c_int, [C.c_void_p, u8p, u64p, C.c_size_t, u64p, u64p, C.c_size_t, szp,
                                              u64p]),
    "scb_spectrum_acc_new": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "scb_index_sampled_spectrum": (C.c_int, [C.c_void_p, u8p, u64p, C.c_size_t, C.c_uint64, C.c_uint64,
                                             C.POINTER(ScbSampledRow), C.c_size_t, szp, C.POINTER(ScbSampledStats)]),
    "scb_sampled_spectrum_from_fastx": (C.c_int, [C.c_char_p, C.POINTER(ScbConfig), C.c_uint64, C.c_uint64,
                                                  C.POINTER(ScbSampledRow), C.c_size_t, szp,
                                                  C.POINTER(ScbSampledStats)]),
    "scb_write_sampled_spectrum": (C.c_int, [C.POINTER(ScbSampledRow), C.c_size_t, C.c_char_p]),
    "scb_spectrum_acc_free": (None, [C.c_void_p]),
    "scb_spectrum_acc_add": (C.c_int, [C.c_void_p, u8p, u64p, C.c_size_t]),
    "scb_spectrum_acc_add_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint64]),
    "scb_spectrum_acc_rows": (C.c_int, [C.c_void_p, u64p, u64p, C.c_size_t, szp, u64p, u64p]),
    "scb_filter_only_spectrum_from_histogram": (C.c_int, [C.c_uint64, u64p, u64p, C.c_size_t, C.c_uint64,
                                                          C.c_uint64, u64p]),
    "scb_filter_only_count_hint_from_histogram": (C.c_int, [C.c_uint64, u64p, u64p, C.c_size_t, C.c_uint64,
                                                            u64p]),
    "scb_index_query_fasta": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(ScbQueryReport)]),
    "scb_index_counter_histogram": (C.c_int, [C.c_void_p, u64p, u64p, C.c_size_t, szp]),
    "scb_index_spectrum": (C.c_int, [C.c_void_p, u8p, u64p, C.c_size_t, u64p, u64p, C.c_size_t, szp,
                                     u64p, u64p]),
    "scb_index_spectrum_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_uint64,
                                            u64p, u64p, C.c_size_t, szp, u64p, u64p]),
    "scb_index_filter_only_spectrum": (C.c_int, [C.c_void_p, u64p, u64p, C.c_size_t, C.c_uint64,
                                                 C.c_uint64, u64p]),
    "scb_index_filter_only_count_hint": (C.c_int, [C.c_void_p, u64p, u64p, C.c_size_t, C.c_uint64, u64p]),
    "scb_estimate_spectrum_from_fastx_pair": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(ScbConfig), u64p,
                                                        u64p, C.c_size_t, szp, C.POINTER(ScbSpectrumStats)]),
    "scb_stream_estimates_from_fastx_pair": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(ScbConfig),
                                                       C.POINTER(ScbStreamStats)]),
    "scb_write_spectrum": (C.c_int, [u64p, u64p, C.c_size_t, C.c_char_p]),
    "scb_for_each_fastx_record": (C.c_int, [C.c_char_p, VISITOR, C.c_void_p]),
    "scb_serve_config_default": (None, [C.POINTER(ScbServeConfig)]),
    "scb_server_new": (C.c_int, [C.c_void_p, C.POINTER(ScbServeConfig), C.POINTER(C.c_void_p)]),
    "scb_server_free": (None, [C.c_void_p]),
    "scb_server_submit_abundances": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                               C.c_void_p, u64p]),
    "scb_server_submit_spectrum": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                             C.c_void_p, C.c_size_t, u64p]),
    "scb_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(ScbServed)]),
    "scb_server_wait_any": (C.c_int, [C.c_void_p, C.c_uint64, u64p, C.POINTER(ScbServed)]),
    "scb_host_alloc": (C.c_void_p, [C.c_size_t]),
    "scb_host_free": (None, [C.c_void_p]),
    "scb_xorshift_u64": (C.c_uint64, [C.c_uint64]),
    "scb_canonical_kmer": (C.c_uint64, [u8p, C.c_size_t, C.c_size_t, C.c_size_t]),
}

_lib = None


def lib():
    """Returns the loaded libscb.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python scbf_amd/build.py` (the HIP "
                               "library is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    L = lib()
    n = L.scb_last_error(None, 0)
    buf = C.create_string_buffer(n + 1)
    L.scb_last_error(buf, n + 1)
    return buf.value.decode("utf-8", "replace")
