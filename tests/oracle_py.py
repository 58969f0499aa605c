"""ctypes binding of the CPU oracle (oracle/liboracle.so). Test infrastructure:
used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as
the checker; the product library never calls it."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class OracleConfig(C.Structure):
    _fields_ = [("k", C.c_uint64), ("m", C.c_uint64), ("s", C.c_uint64), ("n_hashes", C.c_uint64),
                ("counter_bits", C.c_uint8), ("counter_slots_exponent", C.c_uint8),
                ("block_slots_exponent", C.c_uint8), ("_pad", C.c_uint8 * 5),
                ("threads", C.c_uint64), ("queue", C.c_uint64)]


class BaselineResult(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("chunks", "indexed_kmers", "inserted_smers", "queried_kmers",
                                          "positive_estimates", "estimate_checksum")] + \
               [("build_s", C.c_double), ("query_s", C.c_double)]


_L = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        L = C.CDLL(ORACLE_LIB)
        sig = {
            "oracle_xorshift_u64": (C.c_uint64, [C.c_uint64]),
            "oracle_canonical_kmer": (C.c_uint64, [u8p, C.c_size_t, C.c_size_t, C.c_size_t]),
            "oracle_read_kmer": (C.c_uint64, [u8p, C.c_size_t, C.c_size_t, C.c_size_t]),
            "oracle_read_revcomp_kmer": (C.c_uint64, [u8p, C.c_size_t, C.c_size_t, C.c_size_t]),
            "oracle_nthash_canonical": (None, [u8p, C.c_size_t, C.c_size_t, u32p]),
            "oracle_nthash_canonical_direct": (None, [u8p, C.c_size_t, C.c_size_t, u32p]),
            "oracle_minimizer_positions": (None, [u8p, C.c_size_t, C.c_size_t, C.c_size_t, u32p]),
            "oracle_minimizer_positions_brute": (None, [u8p, C.c_size_t, C.c_size_t, C.c_size_t, u32p]),
            "oracle_super_kmers": (C.c_size_t, [u8p, C.c_size_t, C.c_size_t, C.c_size_t, u32p, u64p,
                                                C.c_size_t]),
            "oracle_config_validate": (C.c_int, [C.POINTER(OracleConfig), C.c_char_p, C.c_size_t]),
            "oracle_filter_new": (C.c_void_p, [C.POINTER(OracleConfig)]),
            "oracle_filter_free": (None, [C.c_void_p]),
            "oracle_filter_counter_slots": (C.c_size_t, [C.c_void_p]),
            "oracle_filter_counters_u32": (None, [C.c_void_p, u32p]),
            "oracle_filter_set_counters_u32": (None, [C.c_void_p, u32p]),
            "oracle_filter_insert_packed": (C.c_uint64, [C.c_void_p, u8p, C.c_size_t]),
            "oracle_filter_insert_ascii": (None, [C.c_void_p, u8p, C.c_size_t, u64p, u64p]),
            "oracle_filter_estimate_packed": (C.c_size_t, [C.c_void_p, u8p, C.c_size_t, u64p]),
            "oracle_filter_stream_packed": (None, [C.c_void_p, u8p, C.c_size_t, u64p, u64p]),
            "oracle_filter_estimate_ascii": (C.c_size_t, [C.c_void_p, u8p, C.c_size_t, u64p, u8p]),
            "oracle_filter_estimate_ascii_brute": (C.c_size_t, [C.c_void_p, u8p, C.c_size_t, u64p, u8p]),
            "oracle_filter_counter_histogram": (C.c_size_t, [C.c_void_p, u64p, u64p, C.c_size_t]),
            "oracle_filter_insert_records": (None, [C.c_void_p, u8p, u64p, C.c_size_t, C.c_size_t, u64p, u64p]),
            "oracle_filter_estimate_records": (None, [C.c_void_p, u8p, u64p, C.c_size_t, C.c_size_t, u64p, u8p]),
            "oracle_filter_data": (C.c_void_p, [C.c_void_p, C.POINTER(C.c_size_t)]),
            "oracle_spectrum_add_records": (C.c_uint64, [C.c_void_p, C.c_void_p, u8p, u64p, C.c_size_t,
                                                         C.c_size_t]),
            "oracle_spectrum_new": (C.c_void_p, []),
            "oracle_spectrum_free": (None, [C.c_void_p]),
            "oracle_spectrum_add_chunk": (C.c_uint64, [C.c_void_p, C.c_void_p, u8p, C.c_size_t]),
            "oracle_spectrum_distinct": (C.c_size_t, [C.c_void_p]),
            "oracle_spectrum_rows": (C.c_size_t, [C.c_void_p, u64p, u64p, C.c_size_t]),
            "oracle_spectrum_pairs": (C.c_size_t, [C.c_void_p, u64p, u64p, C.c_size_t]),
            "oracle_reservoir_windows": (C.c_size_t, [u8p, u64p, C.c_size_t, C.c_size_t, C.c_uint64, C.c_uint64,
                                                      u64p, u64p]),
            "oracle_sampled_spectrum": (C.c_size_t, [C.c_void_p, u8p, u64p, C.c_size_t, C.c_uint64, C.c_uint64,
                                                     u64p, C.c_size_t, u64p]),
            "oracle_save": (C.c_int, [C.c_void_p, C.c_uint64, C.c_char_p, C.c_char_p, C.c_size_t]),
            "oracle_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(OracleConfig), u64p,
                                      C.c_char_p, C.c_size_t]),
            "oracle_deconvolved_counts": (C.c_size_t, [u64p, u64p, C.c_size_t, C.c_uint64, C.c_uint64,
                                                       C.POINTER(C.c_double)]),
            "oracle_filter_only_spectrum": (C.c_size_t, [u64p, u64p, C.c_size_t, C.c_uint64, C.c_uint64,
                                                         C.c_uint64, u64p]),
            "oracle_filter_only_count_hint": (C.c_uint64, [u64p, u64p, C.c_size_t, C.c_uint64, C.c_uint64]),
            "oracle_summarize": (C.c_size_t, [u64p, u8p, C.c_size_t, C.POINTER(C.c_double),
                                              C.POINTER(C.c_double)]),
            "oracle_cpu_baseline": (C.c_int, [C.POINTER(OracleConfig), u8p, u64p, C.c_size_t, u8p, u64p,
                                              C.c_size_t, C.POINTER(BaselineResult)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        _L = L
    return _L


def worker_threads():
    """Oracle worker threads: the CPUs this process may use, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _b(seq):
    if isinstance(seq, np.ndarray):
        return np.ascontiguousarray(seq, dtype=np.uint8)
    if isinstance(seq, str):
        seq = seq.encode()
    return np.frombuffer(bytes(seq), dtype=np.uint8).copy()


def _p(a, t=u8p):
    return a.ctypes.data_as(t)


def make_cfg(k, m, s, n_hashes, bits, cse, bse, threads=1, queue=4):
    c = OracleConfig()
    c.k, c.m, c.s, c.n_hashes = k, m, s, n_hashes
    c.counter_bits, c.counter_slots_exponent, c.block_slots_exponent = bits, cse, bse
    c.threads, c.queue = threads, queue
    return c


def validate(cfg):
    buf = C.create_string_buffer(256)
    rc = lib().oracle_config_validate(C.byref(cfg), buf, 256)
    return rc, buf.value.decode()


def xorshift_u64(x):
    return lib().oracle_xorshift_u64(x)


def canonical_kmer(seq, klen, pos):
    a = _b(seq)
    return lib().oracle_canonical_kmer(_p(a), a.size, klen, pos)


def read_kmer(seq, klen, pos):
    a = _b(seq)
    return lib().oracle_read_kmer(_p(a), a.size, klen, pos)


def read_revcomp_kmer(seq, klen, pos):
    a = _b(seq)
    return lib().oracle_read_revcomp_kmer(_p(a), a.size, klen, pos)


def nthash(seq, m, direct=False):
    a = _b(seq)
    out = np.zeros(max(1, a.size - m + 1), dtype=np.uint32)
    fn = lib().oracle_nthash_canonical_direct if direct else lib().oracle_nthash_canonical
    fn(_p(a), a.size, m, _p(out, u32p))
    return out[:max(0, a.size - m + 1)]


def minimizer_positions(seq, k, m, brute=False):
    a = _b(seq)
    out = np.zeros(max(1, a.size - k + 1), dtype=np.uint32)
    fn = lib().oracle_minimizer_positions_brute if brute else lib().oracle_minimizer_positions
    fn(_p(a), a.size, k, m, _p(out, u32p))
    return out[:max(0, a.size - k + 1)]


def super_kmers(seq, k, m):
    a = _b(seq)
    cap = max(1, a.size)
    st = np.zeros(cap, dtype=np.uint32)
    mn = np.zeros(cap, dtype=np.uint64)
    n = lib().oracle_super_kmers(_p(a), a.size, k, m, _p(st, u32p), _p(mn, u64p), cap)
    return st[:n], mn[:n]


class Filter:
    """SuperCountingBloom restated on the CPU (filter.rs + api.rs)."""

    def __init__(self, cfg=None, handle=None):
        self.cfg = cfg
        self.h = handle if handle is not None else lib().oracle_filter_new(C.byref(cfg))
        if not self.h:
            raise ValueError("invalid oracle config")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_filter_free(self.h)
            self.h = None

    def insert_packed(self, seq):
        a = _b(seq)
        return lib().oracle_filter_insert_packed(self.h, _p(a), a.size)

    def insert_ascii(self, seq):
        a = _b(seq)
        k, s = C.c_uint64(), C.c_uint64()
        lib().oracle_filter_insert_ascii(self.h, _p(a), a.size, C.byref(k), C.byref(s))
        return k.value, s.value

    def estimate_packed(self, seq):
        a = _b(seq)
        out = np.zeros(max(1, a.size), dtype=np.uint64)
        n = lib().oracle_filter_estimate_packed(self.h, _p(a), a.size, _p(out, u64p))
        return out[:n]

    def stream_packed(self, seq):
        a = _b(seq)
        q, c = C.c_uint64(), C.c_uint64()
        lib().oracle_filter_stream_packed(self.h, _p(a), a.size, C.byref(q), C.byref(c))
        return q.value, c.value

    def estimate_ascii(self, seq, brute=False):
        a = _b(seq)
        n = max(1, a.size)
        v = np.zeros(n, dtype=np.uint64)
        ok = np.zeros(n, dtype=np.uint8)
        fn = lib().oracle_filter_estimate_ascii_brute if brute else lib().oracle_filter_estimate_ascii
        got = fn(self.h, _p(a), a.size, _p(v, u64p), _p(ok))
        return v[:got], ok[:got].astype(bool)

    def insert_records(self, bases, offsets, threads=None):
        """Record-parallel insert of n records; returns (inserted_kmers, inserted_smers)."""
        b = _b(bases)
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        k, s = C.c_uint64(), C.c_uint64()
        lib().oracle_filter_insert_records(self.h, _p(b), _p(o, u64p), o.size - 1, threads or worker_threads(),
                                           C.byref(k), C.byref(s))
        return k.value, s.value

    def estimate_records(self, bases, offsets, threads=None):
        """Position-indexed abundances (u64) and validity (bool) of n records."""
        b = _b(bases)
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = max(1, int(o[-1]))
        v = np.empty(n, dtype=np.uint64)
        ok = np.empty(n, dtype=np.uint8)
        lib().oracle_filter_estimate_records(self.h, _p(b), _p(o, u64p), o.size - 1, threads or worker_threads(),
                                             _p(v, u64p), _p(ok))
        return v[:int(o[-1])], ok[:int(o[-1])].view(bool)

    def table(self):
        """Zero-copy view of the counter table (physical order, native width);
        valid while this filter lives."""
        nbytes = C.c_size_t()
        ptr = lib().oracle_filter_data(self.h, C.byref(nbytes))
        dt = {8: np.uint8, 16: np.uint16, 32: np.uint32}[self.cfg.counter_bits]
        buf = (C.c_uint8 * nbytes.value).from_address(ptr)
        return np.frombuffer(buf, dtype=dt)

    def sampled_spectrum(self, bases, offsets, sample_size, seed):
        """bin/sampled_spectrum.rs: rows [(estimate, sampled, linear_scaled, distinct)]
        and (seen, sampled, queried, invalid)."""
        b = _b(bases)
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        cap = 1 << 16
        rows = np.zeros(4 * cap, dtype=np.uint64)
        st = np.zeros(4, dtype=np.uint64)
        n = lib().oracle_sampled_spectrum(self.h, _p(b), _p(o, u64p), o.size - 1, sample_size, seed,
                                          _p(rows, u64p), cap, _p(st, u64p))
        r = rows[:4 * min(n, cap)].reshape(-1, 4)
        return [tuple(int(x) for x in row) for row in r], tuple(int(x) for x in st)

    def counters(self):
        n = lib().oracle_filter_counter_slots(self.h)
        out = np.zeros(n, dtype=np.uint32)
        lib().oracle_filter_counters_u32(self.h, _p(out, u32p))
        return out

    def set_counters(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint32)
        lib().oracle_filter_set_counters_u32(self.h, _p(a, u32p))

    def counter_histogram(self):
        cap = 1 << 20
        v = np.zeros(cap, dtype=np.uint64)
        c = np.zeros(cap, dtype=np.uint64)
        n = lib().oracle_filter_counter_histogram(self.h, _p(v, u64p), _p(c, u64p), cap)
        return [(int(a), int(b)) for a, b in zip(v[:n], c[:n])]

    def save(self, path, inserted_kmers):
        buf = C.create_string_buffer(512)
        return lib().oracle_save(self.h, inserted_kmers, os.fspath(path).encode(), buf, 512), buf.value.decode()


def load(path):
    h = C.c_void_p()
    cfg = OracleConfig()
    ik = C.c_uint64()
    buf = C.create_string_buffer(1024)
    rc = lib().oracle_load(os.fspath(path).encode(), C.byref(h), C.byref(cfg), C.byref(ik), buf, 1024)
    if rc:
        return rc, buf.value.decode(), None, None
    return 0, "", Filter(cfg, h.value), ik.value


class Spectrum:
    def __init__(self):
        self.h = lib().oracle_spectrum_new()

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_spectrum_free(self.h)
            self.h = None

    def add_chunk(self, filt, seq):
        a = _b(seq)
        return lib().oracle_spectrum_add_chunk(self.h, filt.h, _p(a), a.size)

    def add_records(self, filt, bases, offsets, threads=None):
        """estimate_distinct_kmers over the DnaChunks of n records; returns queried k-mers."""
        b = _b(bases)
        o = np.ascontiguousarray(offsets, dtype=np.uint64)
        return lib().oracle_spectrum_add_records(self.h, filt.h, _p(b), _p(o, u64p), o.size - 1,
                                                 threads or worker_threads())

    def distinct(self):
        return lib().oracle_spectrum_distinct(self.h)

    def pairs(self):
        """The merged map as sorted key and estimate arrays."""
        n = self.distinct()
        k = np.zeros(max(1, n), dtype=np.uint64)
        v = np.zeros(max(1, n), dtype=np.uint64)
        got = lib().oracle_spectrum_pairs(self.h, _p(k, u64p), _p(v, u64p), n)
        return k[:got], v[:got]

    def rows(self):
        cap = 1 << 20
        v = np.zeros(cap, dtype=np.uint64)
        c = np.zeros(cap, dtype=np.uint64)
        n = lib().oracle_spectrum_rows(self.h, _p(v, u64p), _p(c, u64p), cap)
        return [(int(a), int(b)) for a, b in zip(v[:n], c[:n])]


def _hist(h):
    v = np.array([a for a, _ in h] or [0], dtype=np.uint64)
    c = np.array([b for _, b in h] or [0], dtype=np.uint64)
    return v, c, len(h)


def deconvolved_counts(hist, n_hashes, max_abundance):
    v, c, n = _hist(hist)
    out = np.zeros(max(1, max_abundance), dtype=np.float64)
    got = lib().oracle_deconvolved_counts(_p(v, u64p), _p(c, u64p), n, n_hashes, max_abundance,
                                          out.ctypes.data_as(C.POINTER(C.c_double)))
    return out[:got]


def filter_only_spectrum(hist, n_hashes, max_abundance, hint):
    v, c, n = _hist(hist)
    out = np.zeros(max(1, max_abundance), dtype=np.uint64)
    lib().oracle_filter_only_spectrum(_p(v, u64p), _p(c, u64p), n, n_hashes, max_abundance, hint,
                                      _p(out, u64p))
    return [(a, int(out[a - 1])) for a in range(1, max_abundance + 1)]


def filter_only_count_hint(hist, n_hashes, max_abundance=255):
    v, c, n = _hist(hist)
    return lib().oracle_filter_only_count_hint(_p(v, u64p), _p(c, u64p), n, n_hashes, max_abundance)


def summarize(values, valid):
    v = np.ascontiguousarray(values, dtype=np.uint64)
    ok = np.ascontiguousarray(valid, dtype=np.uint8)
    mean, med = C.c_double(), C.c_double()
    n = lib().oracle_summarize(_p(v, u64p), _p(ok), v.size, C.byref(mean), C.byref(med))
    return n, mean.value, med.value


def cpu_baseline(cfg, ib, io, qb, qo):
    r = BaselineResult()
    rc = lib().oracle_cpu_baseline(C.byref(cfg), _p(ib), _p(io, u64p), io.size - 1, _p(qb), _p(qo, u64p),
                                   qo.size - 1, C.byref(r))
    if rc:
        raise ValueError("invalid config")
    return r


def reservoir_windows(bases, offsets, k, sample_size, seed):
    """bin/sampled_spectrum.rs's reservoir: window number per slot, and windows seen."""
    b = _b(bases)
    o = np.ascontiguousarray(offsets, dtype=np.uint64)
    out = np.zeros(max(1, min(int(o[-1]) + 1, sample_size)), dtype=np.uint64)
    seen = C.c_uint64()
    n = lib().oracle_reservoir_windows(_p(b), _p(o, u64p), o.size - 1, k, sample_size, seed, _p(out, u64p),
                                       C.byref(seen))
    return out[:n], seen.value


def insert_records(filt, bases, offsets):
    """Inserts every record (api.rs:676-726 per record); returns (kmers, smers)."""
    tk = ts = 0
    for i in range(offsets.size - 1):
        a, b = int(offsets[i]), int(offsets[i + 1])
        k, s = filt.insert_ascii(bases[a:b])
        tk += k
        ts += s
    return tk, ts


def estimate_records(filt, bases, offsets, k):
    """Position-indexed abundances of every record (api.rs:799-853 per record)."""
    vals = np.zeros(bases.size, dtype=np.uint64)
    valid = np.zeros(bases.size, dtype=bool)
    for i in range(offsets.size - 1):
        a, b = int(offsets[i]), int(offsets[i + 1])
        if b - a < k:
            continue
        v, ok = filt.estimate_ascii(bases[a:b])
        vals[a:a + v.size] = v
        valid[a:a + v.size] = ok
    return vals, valid
