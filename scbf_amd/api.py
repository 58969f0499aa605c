"""Python host API over libscb.so, mirroring the public API of the reference
crate super-counting-bloom 0.1.0 (src/lib.rs:15-31): same type and method
names, argument meaning and error kinds. Every call goes through the C ABI of
include/scb.h; counting runs on the GPU."""
import ctypes as C
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple
import threading

import numpy as np

from ._lib import (VISITOR, ScbAddReport, ScbBuildReport, ScbConfig, ScbQueryReport, ScbSampledRow,
                   ScbSampledStats, ScbServeConfig, ScbServed, ScbSpectrumStats, ScbStreamStats, ScbSummary,
                   ScbTuning, last_error, lib, u8p, u32p, u64p)

ApproxSpectrum = List[Tuple[int, int]]

# error.rs:3-11 (+ Device, the HIP runtime failure kind of this engine, and Busy,
# the admission refusal of the serving loop)
ERROR_KINDS = {1: "InvalidConfig", 2: "FastxOpen", 3: "InvalidIndexFormat", 4: "ChannelClosed",
               5: "WorkerPanic", 6: "InternalState", 7: "Io", 8: "Device", 9: "Busy"}

METHOD_BATCHED = 0
METHOD_PLAIN = 1
METHOD_AUTO = 2  # the default: plain below a size that depends on the table, batched above


class SuperCountingBloomError(Exception):
    """error.rs:3-30; `kind` names the variant, str(e) is its Display text."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code
        self.kind = ERROR_KINDS.get(code, "Unknown")


def _check(rc):
    if rc != 0:
        raise SuperCountingBloomError(rc, last_error())


class CounterBits:  # config.rs:3-18
    Bits8 = 8
    Bits16 = 16
    Bits32 = 32


@dataclass
class SuperCountingBloomConfig:  # config.rs:20-47
    k: int = 31
    m: int = 21
    s: int = 25
    n_hashes: int = 4
    counter_bits: int = CounterBits.Bits16
    counter_slots_exponent: int = 30
    block_slots_exponent: int = 9
    threads: int = field(default_factory=lambda: max(1, os.cpu_count() or 1))
    queue: int = 4096

    def _c(self):
        c = ScbConfig()
        c.k, c.m, c.s, c.n_hashes = self.k, self.m, self.s, self.n_hashes
        c.counter_bits = self.counter_bits
        c.counter_slots_exponent = self.counter_slots_exponent
        c.block_slots_exponent = self.block_slots_exponent
        c.threads, c.queue = self.threads, self.queue
        return c

    @staticmethod
    def _from_c(c):
        return SuperCountingBloomConfig(c.k, c.m, c.s, c.n_hashes, c.counter_bits,
                                        c.counter_slots_exponent, c.block_slots_exponent, c.threads,
                                        c.queue)

    def validate(self):  # config.rs:50-92
        if not (0 <= self.counter_slots_exponent < 256 and 0 <= self.block_slots_exponent < 256):
            raise SuperCountingBloomError(1, "invalid config: exponent out of u8 range")
        for v in (self.k, self.m, self.s, self.n_hashes, self.threads, self.queue):
            if v < 0:
                raise SuperCountingBloomError(1, "invalid config: negative value")
        c = self._c()
        _check(lib().scb_config_validate(C.byref(c)))


@dataclass
class AddReport:  # api.rs:66-73
    records_processed: int = 0
    records_indexed: int = 0
    inserted_kmers: int = 0
    inserted_smers: int = 0
    add_seconds: float = 0.0


@dataclass
class QueryFastaReport:  # api.rs:75-82
    records_processed: int = 0
    total_windows: int = 0
    valid_kmers: int = 0
    positive_estimates: int = 0
    estimate_checksum: int = 0

    def __eq__(self, other):
        return (self.records_processed, self.total_windows, self.valid_kmers, self.positive_estimates,
                self.estimate_checksum) == (other.records_processed, other.total_windows,
                                            other.valid_kmers, other.positive_estimates,
                                            other.estimate_checksum)


@dataclass
class SequenceAbundanceSummary:  # api.rs:84-92
    name: str
    sequence_len: int
    total_windows: int
    valid_kmers: int
    mean_abundance: Optional[float]
    median_abundance: Optional[float]


@dataclass
class FilterOnlySpectrumEstimate:  # api.rs:94-99
    method: str
    kmer_count_hint: int
    rows: ApproxSpectrum


@dataclass
class SuperCountingBloomStats:  # fastx.rs:21-34
    chunks: int = 0
    indexed_kmers: int = 0
    inserted_smers: int = 0
    queried_kmers: int = 0
    distinct_keys: int = 0
    counter_slots: int = 0
    blocks: int = 0
    build_seconds: float = 0.0
    query_seconds: float = 0.0
    spectrum_seconds: float = 0.0
    total_seconds: float = 0.0


@dataclass
class SuperCountingBloomReport:  # fastx.rs:36-40
    spectrum: ApproxSpectrum
    stats: SuperCountingBloomStats


@dataclass
class StreamEstimateStats:  # fastx.rs:42-55
    chunks: int = 0
    indexed_kmers: int = 0
    inserted_smers: int = 0
    queried_kmers: int = 0
    positive_estimates: int = 0
    estimate_checksum: int = 0
    counter_slots: int = 0
    blocks: int = 0
    build_seconds: float = 0.0
    query_seconds: float = 0.0
    total_seconds: float = 0.0


@dataclass
class StreamEstimateReport:  # fastx.rs:57-60
    stats: StreamEstimateStats


@dataclass
class BuildIndexReport:  # api.rs:56-64
    index: "SuperCountingBloomIndex"
    chunks: int
    indexed_kmers: int
    inserted_smers: int
    counter_slots: int
    blocks: int
    build_seconds: float


def _bytes(seq) -> np.ndarray:
    if isinstance(seq, np.ndarray):
        return np.ascontiguousarray(seq, dtype=np.uint8)
    if isinstance(seq, str):
        seq = seq.encode()
    return np.frombuffer(bytes(seq), dtype=np.uint8)


def _ptr(a, t=u8p):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def concat_records(records: Sequence) -> Tuple[np.ndarray, np.ndarray]:
    """Concatenated bytes + n+1 offsets for a list of byte strings."""
    arrs = [_bytes(r) for r in records]
    offsets = np.zeros(len(arrs) + 1, dtype=np.uint64)
    if arrs:
        offsets[1:] = np.cumsum([a.size for a in arrs], dtype=np.uint64)
    bases = np.concatenate(arrs) if arrs else np.zeros(0, dtype=np.uint8)
    return np.ascontiguousarray(bases, dtype=np.uint8), offsets


def _path(p):
    return os.fspath(p).encode()


@dataclass
class Tuning:
    """Tunables of the batched pipeline (include/scb.h scb_tuning); results are
    identical for every valid setting."""
    insert_bin_log: int = 15
    query_bin_log: int = 15
    spectrum_bin_log: int = 14
    insert_item_records: int = 2048
    query_item_records: int = 3584
    spectrum_item_records: int = 4096
    use_graph: int = 0

    def _c(self):
        return ScbTuning(self.insert_bin_log, self.query_bin_log, self.spectrum_bin_log, self.insert_item_records,
                         self.query_item_records, self.spectrum_item_records, self.use_graph, 0)

    @staticmethod
    def _from_c(c):
        return Tuning(c.insert_bin_log, c.query_bin_log, c.spectrum_bin_log, c.insert_item_records,
                      c.query_item_records, c.spectrum_item_records, c.use_graph)


def tuning_cache_load(path):
    """Replaces the process-wide tuning cache; engines created afterwards use it."""
    _check(lib().scb_tuning_cache_load(_path(path)))


class SuperCountingBloomBuilder:  # api.rs:43-48, 101-208
    def __init__(self, cfg: SuperCountingBloomConfig):
        self._h = C.c_void_p()
        c = cfg._c()
        _check(lib().scb_builder_new(C.byref(c), C.byref(self._h)))

    @classmethod
    def new(cls, cfg):
        return cls(cfg)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().scb_builder_free(h)
            self._h = C.c_void_p()

    def _handle(self):
        if not self._h:
            raise SuperCountingBloomError(6, "internal state error: builder was consumed by into_index")
        return self._h

    def add_sequence(self, sequence) -> int:
        a = _bytes(sequence)
        out = C.c_uint64()
        _check(lib().scb_builder_add_sequence(self._handle(), _ptr(a), a.size, C.byref(out)))
        return out.value

    def add_sequences(self, records) -> AddReport:
        bases, offsets = concat_records(records) if not isinstance(records, tuple) else records
        rep = ScbAddReport()
        _check(lib().scb_builder_add_sequences(self._handle(), _ptr(bases), _ptr(offsets, u64p),
                                               offsets.size - 1, C.byref(rep)))
        return AddReport(rep.records_processed, rep.records_indexed, rep.inserted_kmers,
                         rep.inserted_smers, rep.add_seconds)

    def add_sequences_device(self, d_bases: int, d_offsets: int, n: int, total_bases: int) -> AddReport:
        rep = ScbAddReport()
        _check(lib().scb_builder_add_sequences_device(self._handle(), C.c_void_p(d_bases),
                                                      C.c_void_p(d_offsets), n, total_bases,
                                                      C.byref(rep)))
        return AddReport(0, 0, rep.inserted_kmers, rep.inserted_smers, rep.add_seconds)

    def add_fasta(self, path) -> AddReport:
        rep = ScbAddReport()
        _check(lib().scb_builder_add_fasta(self._handle(), _path(path), C.byref(rep)))
        return AddReport(rep.records_processed, rep.records_indexed, rep.inserted_kmers,
                         rep.inserted_smers, rep.add_seconds)

    def into_index(self) -> "SuperCountingBloomIndex":
        out = C.c_void_p()
        h = self._handle()
        self._h = C.c_void_p()
        _check(lib().scb_builder_into_index(h, C.byref(out)))
        return SuperCountingBloomIndex._wrap(out)

    def inserted_kmers(self) -> int:
        return lib().scb_builder_inserted_kmers(self._handle())

    def inserted_smers(self) -> int:
        return lib().scb_builder_inserted_smers(self._handle())

    def set_threads(self, threads: int):
        if threads < 0:
            raise SuperCountingBloomError(1, "invalid config: threads must be positive")
        _check(lib().scb_builder_set_threads(self._handle(), threads))

    def clear_threads(self):
        lib().scb_builder_clear_threads(self._handle())

    def threads(self) -> int:
        return lib().scb_builder_threads(self._handle())

    def config(self) -> SuperCountingBloomConfig:
        c = ScbConfig()
        lib().scb_builder_config(self._handle(), C.byref(c))
        return SuperCountingBloomConfig._from_c(c)

    # extensions
    def set_method(self, method: int):
        _check(lib().scb_builder_set_method(self._handle(), method))

    def set_tuning(self, t: "Tuning"):
        c = t._c()
        _check(lib().scb_builder_set_tuning(self._handle(), C.byref(c)))

    def tuning(self) -> "Tuning":
        c = ScbTuning()
        lib().scb_builder_tuning(self._handle(), C.byref(c))
        return Tuning._from_c(c)

    def counters(self) -> np.ndarray:
        cfg = self.config()
        dt = {8: np.uint8, 16: np.uint16, 32: np.uint32}[cfg.counter_bits]
        out = np.empty(1 << cfg.counter_slots_exponent, dtype=dt)
        _check(lib().scb_builder_counters(self._handle(), out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out


class SuperCountingBloomIndex:  # api.rs:50-54, 210-386
    def __init__(self):
        raise TypeError("use SuperCountingBloomIndex.build_from_fastx / load / Builder.into_index")

    @classmethod
    def _wrap(cls, h):
        self = object.__new__(cls)
        self._h = h
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().scb_index_free(h)
            self._h = None

    @classmethod
    def build_from_fastx(cls, path, cfg: SuperCountingBloomConfig) -> BuildIndexReport:
        out = C.c_void_p()
        rep = ScbBuildReport()
        c = cfg._c()
        _check(lib().scb_index_build_from_fastx(_path(path), C.byref(c), C.byref(out), C.byref(rep)))
        return BuildIndexReport(cls._wrap(out), rep.chunks, rep.indexed_kmers, rep.inserted_smers,
                                rep.counter_slots, rep.blocks, rep.build_seconds)

    @classmethod
    def load(cls, path) -> "SuperCountingBloomIndex":
        out = C.c_void_p()
        _check(lib().scb_index_load(_path(path), C.byref(out)))
        return cls._wrap(out)

    def save(self, path):
        _check(lib().scb_index_save(self._h, _path(path)))

    def config(self) -> SuperCountingBloomConfig:
        c = ScbConfig()
        lib().scb_index_config(self._h, C.byref(c))
        return SuperCountingBloomConfig._from_c(c)

    def inserted_kmers(self) -> int:
        return lib().scb_index_inserted_kmers(self._h)

    def estimate_sequence_abundances(self, sequence) -> List[Optional[int]]:
        vals, valid = self.estimate_sequence_abundances_np(sequence)
        return [int(v) if ok else None for v, ok in zip(vals, valid)]

    def estimate_sequence_abundances_np(self, sequence):
        a = _bytes(sequence)
        k = self.config().k
        n = max(0, a.size - k + 1)
        vals = np.zeros(max(n, 1), dtype=np.uint64)
        valid = np.zeros(max(n, 1), dtype=np.uint8)
        got = C.c_size_t()
        _check(lib().scb_index_estimate_sequence_abundances(self._h, _ptr(a), a.size, _ptr(vals, u64p),
                                                            _ptr(valid), C.byref(got)))
        return vals[:n], valid[:n].astype(bool)

    def estimate_sequences(self, bases: np.ndarray, offsets: np.ndarray):
        """Batched abundances: (values, valid, QueryFastaReport), position-indexed."""
        vals = np.zeros(max(bases.size, 1), dtype=np.uint64)
        valid = np.zeros(max(bases.size, 1), dtype=np.uint8)
        rep = ScbQueryReport()
        _check(lib().scb_index_estimate_sequences(self._h, _ptr(bases), _ptr(offsets, u64p),
                                                  offsets.size - 1, _ptr(vals, u64p), _ptr(valid),
                                                  C.byref(rep)))
        return vals[:bases.size], valid[:bases.size].astype(bool), _qrep(rep)

    def estimate_sequences_device(self, d_bases: int, d_offsets: int, n: int, total_bases: int,
                                  d_values: Optional[int], d_valid_bits: Optional[int]):
        rep = ScbQueryReport()
        _check(lib().scb_index_estimate_sequences_device(
            self._h, C.c_void_p(d_bases), C.c_void_p(d_offsets), n, total_bases,
            C.c_void_p(d_values) if d_values else None,
            C.c_void_p(d_valid_bits) if d_valid_bits else None, C.byref(rep)))
        return _qrep(rep)

    def summarize_sequence(self, name, sequence) -> SequenceAbundanceSummary:
        a = _bytes(sequence)
        s = ScbSummary()
        _check(lib().scb_index_summarize_sequence(self._h, _ptr(a), a.size, C.byref(s)))
        mean = None if math.isnan(s.mean_abundance) else s.mean_abundance
        med = None if math.isnan(s.median_abundance) else s.median_abundance
        return SequenceAbundanceSummary(name, s.sequence_len, s.total_windows, s.valid_kmers, mean, med)

    def query_fasta(self, path) -> QueryFastaReport:
        rep = ScbQueryReport()
        _check(lib().scb_index_query_fasta(self._h, _path(path), C.byref(rep)))
        return _qrep(rep)

    def counter_histogram(self) -> ApproxSpectrum:
        cap = 1 << 16
        while True:
            v = np.zeros(cap, dtype=np.uint64)
            c = np.zeros(cap, dtype=np.uint64)
            n = C.c_size_t()
            _check(lib().scb_index_counter_histogram(self._h, _ptr(v, u64p), _ptr(c, u64p), cap,
                                                     C.byref(n)))
            if n.value <= cap:
                return [(int(a), int(b)) for a, b in zip(v[:n.value], c[:n.value])]
            cap = n.value

    def spectrum(self, records) -> Tuple[ApproxSpectrum, int, int]:
        bases, offsets = concat_records(records) if not isinstance(records, tuple) else records
        cap = 1 << 17
        v = np.zeros(cap, dtype=np.uint64)
        c = np.zeros(cap, dtype=np.uint64)
        n, q, d = C.c_size_t(), C.c_uint64(), C.c_uint64()
        _check(lib().scb_index_spectrum(self._h, _ptr(bases), _ptr(offsets, u64p), offsets.size - 1,
                                        _ptr(v, u64p), _ptr(c, u64p), cap, C.byref(n), C.byref(q),
                                        C.byref(d)))
        rows = [(int(a), int(b)) for a, b in zip(v[:min(cap, n.value)], c[:min(cap, n.value)])]
        return rows, q.value, d.value

    def estimate_kmer_map(self, records) -> Tuple[dict, int]:
        """filter.rs:302-335: {canonical k-mer: estimate} of every distinct k-mer
        of the records' DnaChunks, and the number of k-mers queried."""
        keys, vals, queried = self.estimate_kmer_map_np(records)
        return dict(zip(keys.tolist(), vals.tolist())), queried

    def estimate_kmer_map_np(self, records):
        """The same as sorted key and estimate arrays."""
        bases, offsets = concat_records(records) if not isinstance(records, tuple) else records
        cap = max(1, int(offsets[-1]))
        keys = np.zeros(cap, dtype=np.uint64)
        vals = np.zeros(cap, dtype=np.uint64)
        n, q = C.c_size_t(), C.c_uint64()
        _check(lib().scb_index_estimate_kmer_map(self._h, _ptr(bases), _ptr(offsets, u64p), offsets.size - 1,
                                                 _ptr(keys, u64p), _ptr(vals, u64p), cap, C.byref(n), C.byref(q)))
        return keys[:n.value], vals[:n.value], q.value

    def sampled_spectrum(self, records, sample_size: int = 1_000_000, seed: int = 1) -> "SampledSpectrum":
        """bin/sampled_spectrum.rs:150-253 over the records against this index."""
        bases, offsets = concat_records(records) if not isinstance(records, tuple) else records
        cap = 1 << 17
        rows = (ScbSampledRow * cap)()
        n = C.c_size_t()
        st = ScbSampledStats()
        _check(lib().scb_index_sampled_spectrum(self._h, _ptr(bases), _ptr(offsets, u64p), offsets.size - 1,
                                                sample_size, seed, rows, cap, C.byref(n), C.byref(st)))
        return _sampled(rows, n.value, st)

    def spectrum_accumulator(self) -> "SpectrumAccumulator":
        return SpectrumAccumulator(self)

    def estimate_filter_only_spectrum(self, max_abundance, kmer_count_hint) -> FilterOnlySpectrumEstimate:
        return self.estimate_filter_only_spectrum_from_counter_histogram(None, max_abundance,
                                                                         kmer_count_hint)

    def estimate_filter_only_spectrum_from_counter_histogram(self, counter_histogram, max_abundance,
                                                             kmer_count_hint):
        hv, hc, hr = _hist_arrays(counter_histogram)
        out = np.zeros(max(1, max_abundance), dtype=np.uint64)
        _check(lib().scb_index_filter_only_spectrum(self._h, _ptr(hv, u64p), _ptr(hc, u64p), hr,
                                                    max_abundance, kmer_count_hint, _ptr(out, u64p)))
        rows = [(a, int(out[a - 1])) for a in range(1, max_abundance + 1)]
        return FilterOnlySpectrumEstimate("compound_poisson_counter_deconvolution", kmer_count_hint, rows)

    def filter_only_count_hint(self) -> int:
        return self.filter_only_count_hint_from_counter_histogram(None, 255)

    def filter_only_count_hint_from_counter_histogram(self, counter_histogram, max_abundance) -> int:
        hv, hc, hr = _hist_arrays(counter_histogram)
        out = C.c_uint64()
        _check(lib().scb_index_filter_only_count_hint(self._h, _ptr(hv, u64p), _ptr(hc, u64p), hr,
                                                      max_abundance, C.byref(out)))
        return out.value

    # extensions
    def set_method(self, method: int):
        _check(lib().scb_index_set_method(self._h, method))

    def set_tuning(self, t: "Tuning"):
        c = t._c()
        _check(lib().scb_index_set_tuning(self._h, C.byref(c)))

    def tuning(self) -> "Tuning":
        c = ScbTuning()
        lib().scb_index_tuning(self._h, C.byref(c))
        return Tuning._from_c(c)

    def counters(self) -> np.ndarray:
        cfg = self.config()
        dt = {8: np.uint8, 16: np.uint16, 32: np.uint32}[cfg.counter_bits]
        out = np.empty(1 << cfg.counter_slots_exponent, dtype=dt)
        _check(lib().scb_index_counters(self._h, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out


class SpectrumAccumulator:
    """Distinct-k-mer spectrum over many batches against one index
    (fastx.rs:255-308 with merged per-worker maps): add() batches, then rows()."""

    def __init__(self, index: "SuperCountingBloomIndex"):
        self._index = index
        self._h = C.c_void_p()
        _check(lib().scb_spectrum_acc_new(index._h, C.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().scb_spectrum_acc_free(h)
            self._h = C.c_void_p()

    def add(self, records):
        bases, offsets = concat_records(records) if not isinstance(records, tuple) else records
        _check(lib().scb_spectrum_acc_add(self._h, _ptr(bases), _ptr(offsets, u64p), offsets.size - 1))

    def add_device(self, d_bases: int, d_offsets: int, n: int, total_bases: int):
        _check(lib().scb_spectrum_acc_add_device(self._h, C.c_void_p(d_bases), C.c_void_p(d_offsets), n,
                                                 total_bases))

    def rows(self) -> Tuple[ApproxSpectrum, int, int]:
        cap = 1 << 17
        while True:
            v = np.zeros(cap, dtype=np.uint64)
            c = np.zeros(cap, dtype=np.uint64)
            n, q, d = C.c_size_t(), C.c_uint64(), C.c_uint64()
            _check(lib().scb_spectrum_acc_rows(self._h, _ptr(v, u64p), _ptr(c, u64p), cap, C.byref(n), C.byref(q),
                                               C.byref(d)))
            if n.value <= cap:
                return [(int(a), int(b)) for a, b in zip(v[:n.value], c[:n.value])], q.value, d.value
            cap = n.value


@dataclass
class SampledSpectrum:
    """write_scaled_spectrum's rows (estimated_count, sampled_kmers,
    linear_scaled_windows, estimated_distinct_kmers) and the CLI's metrics."""
    rows: List[Tuple[int, int, int, int]]
    total_inserted_kmers: int
    sampled_kmers: int
    queried_kmers: int
    invalid_queries: int
    indexed_kmers: int = 0
    inserted_smers: int = 0
    build_seconds: float = 0.0
    sample_seconds: float = 0.0

    def write(self, output=None):
        arr = (ScbSampledRow * max(1, len(self.rows)))()
        for i, r in enumerate(self.rows):
            arr[i] = ScbSampledRow(*r)
        _check(lib().scb_write_sampled_spectrum(arr, len(self.rows), None if output is None else _path(output)))


def _sampled(rows, n, st):
    out = [(rows[i].estimated_count, rows[i].sampled_kmers, rows[i].linear_scaled_windows,
            rows[i].estimated_distinct_kmers) for i in range(min(n, len(rows)))]
    return SampledSpectrum(out, st.total_inserted_kmers, st.sampled_kmers, st.queried_kmers, st.invalid_queries,
                           st.indexed_kmers, st.inserted_smers, st.build_seconds, st.sample_seconds)


def sampled_spectrum_from_fastx(path, cfg: SuperCountingBloomConfig, sample_size: int = 1_000_000,
                                seed: int = 1) -> SampledSpectrum:
    """run() of bin/sampled_spectrum.rs: index the file, then sample and query its windows."""
    cap = 1 << 17
    rows = (ScbSampledRow * cap)()
    n = C.c_size_t()
    st = ScbSampledStats()
    c = cfg._c()
    _check(lib().scb_sampled_spectrum_from_fastx(_path(path), C.byref(c), sample_size, seed, rows, cap,
                                                 C.byref(n), C.byref(st)))
    return _sampled(rows, n.value, st)


def _hist_arrays(h):
    if h is None:
        return None, None, 0
    hv = np.array([a for a, _ in h], dtype=np.uint64)
    hc = np.array([b for _, b in h], dtype=np.uint64)
    return hv, hc, len(h)


def filter_only_spectrum_from_histogram(counter_histogram, n_hashes: int, max_abundance: int,
                                        kmer_count_hint: int) -> ApproxSpectrum:
    """Rows (a, estimated k-mers) for a = 1..=max_abundance of a given counter
    histogram (api.rs:933-1033), host arithmetic only."""
    hv, hc, hr = _hist_arrays(counter_histogram)
    out = np.zeros(max(1, max_abundance), dtype=np.uint64)
    _check(lib().scb_filter_only_spectrum_from_histogram(n_hashes, _ptr(hv, u64p), _ptr(hc, u64p), hr,
                                                         max_abundance, kmer_count_hint, _ptr(out, u64p)))
    return [(a, int(out[a - 1])) for a in range(1, max_abundance + 1)]


def filter_only_count_hint_from_histogram(counter_histogram, n_hashes: int, max_abundance: int = 255) -> int:
    """api.rs:375-385 on a given counter histogram, host arithmetic only."""
    hv, hc, hr = _hist_arrays(counter_histogram)
    out = C.c_uint64()
    _check(lib().scb_filter_only_count_hint_from_histogram(n_hashes, _ptr(hv, u64p), _ptr(hc, u64p), hr,
                                                           max_abundance, C.byref(out)))
    return out.value


def summarize_abundances(values, valid):
    """(total_windows, valid_kmers, mean or None, median or None) of per-window
    values and validity (api.rs:855-891)."""
    v = np.ascontiguousarray(values, dtype=np.uint64)
    ok = np.ascontiguousarray(valid, dtype=np.uint8)
    s = ScbSummary()
    _check(lib().scb_summarize_abundances(_ptr(v, u64p), _ptr(ok), v.size, C.byref(s)))
    mean = None if math.isnan(s.mean_abundance) else s.mean_abundance
    med = None if math.isnan(s.median_abundance) else s.median_abundance
    return s.total_windows, s.valid_kmers, mean, med


def _qrep(r):
    return QueryFastaReport(r.records_processed, r.total_windows, r.valid_kmers, r.positive_estimates,
                            r.estimate_checksum)


def estimate_spectrum_from_fastx(path, cfg) -> SuperCountingBloomReport:  # fastx.rs:81-86
    return estimate_spectrum_from_fastx_pair(path, path, cfg)


def estimate_spectrum_from_fastx_pair(index_path, query_path, cfg) -> SuperCountingBloomReport:
    cap = 1 << 17
    v = np.zeros(cap, dtype=np.uint64)
    c = np.zeros(cap, dtype=np.uint64)
    n = C.c_size_t()
    st = ScbSpectrumStats()
    cc = cfg._c()
    _check(lib().scb_estimate_spectrum_from_fastx_pair(_path(index_path), _path(query_path), C.byref(cc),
                                                       _ptr(v, u64p), _ptr(c, u64p), cap, C.byref(n),
                                                       C.byref(st)))
    rows = [(int(a), int(b)) for a, b in zip(v[:min(cap, n.value)], c[:min(cap, n.value)])]
    stats = SuperCountingBloomStats(*[getattr(st, f[0]) for f in ScbSpectrumStats._fields_])
    return SuperCountingBloomReport(rows, stats)


def stream_estimates_from_fastx(path, cfg) -> StreamEstimateReport:  # fastx.rs:107-112
    return stream_estimates_from_fastx_pair(path, path, cfg)


def stream_estimates_from_fastx_pair(index_path, query_path, cfg) -> StreamEstimateReport:
    st = ScbStreamStats()
    cc = cfg._c()
    _check(lib().scb_stream_estimates_from_fastx_pair(_path(index_path), _path(query_path), C.byref(cc),
                                                      C.byref(st)))
    return StreamEstimateReport(StreamEstimateStats(*[getattr(st, f[0]) for f in ScbStreamStats._fields_]))


def write_spectrum(spectrum: ApproxSpectrum, output=None):  # spectrum.rs:11-31
    v = np.array([a for a, _ in spectrum], dtype=np.uint64)
    c = np.array([b for _, b in spectrum], dtype=np.uint64)
    _check(lib().scb_write_spectrum(_ptr(v, u64p), _ptr(c, u64p), len(spectrum),
                                    _path(output) if output is not None else None))


def for_each_fastx_record(path, visit):  # api.rs:893-915
    err = []

    def cb(_user, name, seq, n):
        try:
            visit(name.decode("utf-8", "replace"), C.string_at(seq, n) if n else b"")
            return 0
        except Exception as e:  # propagate after the walk stops
            err.append(e)
            return 6

    rc = lib().scb_for_each_fastx_record(_path(path), VISITOR(cb), None)
    if err:
        raise err[0]
    _check(rc)


# ---- serving loop (no counterpart in the reference) -------------------------
REQ_ABUNDANCES = 0
REQ_SPECTRUM = 1


@dataclass
class ServeConfig:
    """Coalescing limits, HBM budget and admission limit of a SuperCountingBloomServer."""
    max_batch_bases: int = 1 << 27
    max_batch_requests: int = 1024
    hbm_budget_bytes: int = 16 << 30
    workers: int = 4
    max_pending: int = 4096

    def _c(self):
        return ScbServeConfig(self.max_batch_bases, self.max_batch_requests, self.hbm_budget_bytes,
                              self.workers, self.max_pending)


@dataclass
class Served:
    """One finished request: an abundance request carries values (counter width,
    position-indexed, 0 where no valid window starts), the validity bitmap and the
    report of scb_index_estimate_sequences; a spectrum request carries rows,
    queried and distinct. Times are CLOCK_MONOTONIC nanoseconds."""
    ticket: int
    kind: int
    report: Optional[QueryFastaReport] = None
    values: Optional[np.ndarray] = None
    valid_bits: Optional[np.ndarray] = None
    total_bases: int = 0
    rows: Optional[ApproxSpectrum] = None
    queried: int = 0
    distinct: int = 0
    submit_ns: int = 0
    start_ns: int = 0
    done_ns: int = 0
    batch_requests: int = 0

    @property
    def valid(self) -> np.ndarray:
        bits = np.unpackbits(self.valid_bits.view(np.uint8), bitorder="little")
        return bits[:self.total_bases].astype(bool)


def host_array(n: int, dtype) -> np.ndarray:
    """numpy array over page-locked host memory (freed with the array)."""
    dt = np.dtype(dtype)
    nbytes = max(1, n * dt.itemsize)
    p = lib().scb_host_alloc(nbytes)
    if not p:
        raise MemoryError(f"scb_host_alloc({nbytes}) failed")
    buf = (C.c_uint8 * nbytes).from_address(p)
    arr = np.frombuffer(buf, dtype=dt, count=n)
    import weakref
    weakref.finalize(buf, lib().scb_host_free, C.c_void_p(p))
    return arr


class SuperCountingBloomServer:
    """Serves concurrent abundance and spectrum requests from an index resident in
    HBM (include/scb.h serving loop). Inputs and outputs of a request are kept
    alive here until it is waited for."""

    def __init__(self, index: "SuperCountingBloomIndex", cfg: Optional[ServeConfig] = None):
        self._index = index
        self._h = C.c_void_p()
        c = (cfg or ServeConfig())._c()
        _check(lib().scb_server_new(index._h, C.byref(c), C.byref(self._h)))
        self._dtype = {8: np.uint8, 16: np.uint16, 32: np.uint32}[index.config().counter_bits]
        self._live = {}
        # Held from a submit call until its ticket is registered in _live, and taken
        # before a finished ticket is looked up: a request can finish (and be returned
        # by wait_any in another thread) before its submit call has returned.
        self._reg = threading.Lock()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().scb_server_free(self._h)
            self._h = C.c_void_p()
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit_abundances(self, bases: np.ndarray, offsets: np.ndarray, values: Optional[np.ndarray] = None,
                          valid_bits: Optional[np.ndarray] = None) -> int:
        bases = np.ascontiguousarray(bases, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        nb = int(offsets[-1])
        if values is None:
            values = np.empty(max(nb, 1), dtype=self._dtype)
        if valid_bits is None:
            valid_bits = np.empty(max((nb + 31) // 32, 1), dtype=np.uint32)
        if values.dtype != self._dtype or values.size < nb or valid_bits.dtype != np.uint32 \
                or valid_bits.size < (nb + 31) // 32:
            raise ValueError("output buffers are too small or of the wrong dtype")
        t = C.c_uint64()
        with self._reg:
            _check(lib().scb_server_submit_abundances(self._h, bases.ctypes.data, offsets.ctypes.data,
                                                      offsets.size - 1, values.ctypes.data, valid_bits.ctypes.data,
                                                      C.byref(t)))
            self._live[t.value] = (REQ_ABUNDANCES, bases, offsets, values, valid_bits, nb)
        return t.value

    def submit_spectrum(self, bases: np.ndarray, offsets: np.ndarray, cap: int = 1 << 17) -> int:
        bases = np.ascontiguousarray(bases, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        rv = np.zeros(cap, dtype=np.uint64)
        rc = np.zeros(cap, dtype=np.uint64)
        t = C.c_uint64()
        with self._reg:
            _check(lib().scb_server_submit_spectrum(self._h, bases.ctypes.data, offsets.ctypes.data,
                                                    offsets.size - 1, rv.ctypes.data, rc.ctypes.data, cap,
                                                    C.byref(t)))
            self._live[t.value] = (REQ_SPECTRUM, bases, offsets, rv, rc, cap)
        return t.value

    def _take(self, ticket):
        with self._reg:
            return self._live.pop(ticket, None)

    def _result(self, ticket, r: ScbServed) -> Served:
        kind, _b, _o, x, y, z = self._take(ticket)
        out = Served(ticket, kind, submit_ns=r.submit_ns, start_ns=r.start_ns, done_ns=r.done_ns,
                     batch_requests=r.batch_requests)
        if kind == REQ_ABUNDANCES:
            out.report, out.values, out.valid_bits, out.total_bases = _qrep(r.report), x[:z], y, z
        else:
            n = min(z, r.n_rows)
            out.rows = [(int(a), int(b)) for a, b in zip(x[:n], y[:n])]
            out.queried, out.distinct = r.queried, r.distinct
        return out

    def wait(self, ticket: int) -> Served:
        r = ScbServed()
        rc = lib().scb_server_wait(self._h, ticket, C.byref(r))
        if rc != 0 and r.done_ns:  # a failed request was taken
            self._take(ticket)
        _check(rc)
        return self._result(ticket, r)

    def wait_any(self, timeout_s: float) -> Optional[Served]:
        r = ScbServed()
        t = C.c_uint64()
        rc = lib().scb_server_wait_any(self._h, int(timeout_s * 1e6), C.byref(t), C.byref(r))
        if rc != 0 and t.value:
            self._take(t.value)
        _check(rc)
        return self._result(t.value, r) if t.value else None


def xorshift_u64(x: int) -> int:  # filter.rs:629-635
    return lib().scb_xorshift_u64(x)


def canonical_kmer(sequence, length: int, pos: int) -> int:  # filter.rs:622-627
    a = _bytes(sequence)
    return lib().scb_canonical_kmer(_ptr(a), a.size, length, pos)


def device_count() -> int:
    return lib().scb_device_count()
