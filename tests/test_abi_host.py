"""CPU tests of libscb.so: it loads, exports every symbol include/scb.h declares,
and its host-side logic (config validation, error kinds and messages, .scb
header checks, FASTX record walk, spectrum writer, primitives) behaves like the
reference. No counting work runs here (that needs the GPU)."""
import ctypes as C
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import oracle_py as O
import scbf_amd as scb
from scbf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "scb.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(scb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    scb.api.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (scb_[a-z0-9_]+)", out))
    declared = declared_functions()
    assert declared, "no declarations parsed"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    # and the ctypes table binds exactly the declared functions
    assert sorted(_lib.SIGNATURES) == declared


def test_library_is_built_for_gfx950():
    # the offload bundle of our own build names its target triple
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_default_config_matches_reference():  # config.rs:33-47
    c = _lib.ScbConfig()
    scb.api.lib().scb_config_default(C.byref(c))
    assert (c.k, c.m, c.s, c.n_hashes, c.counter_bits, c.counter_slots_exponent,
            c.block_slots_exponent, c.queue) == (31, 21, 25, 4, 16, 30, 9, 4096)
    assert c.threads >= 1


@pytest.mark.parametrize("kw,msg", [
    (dict(k=0), "k must be in 1..=32"), (dict(k=10), "canonical SIMD minimizers require odd k"),
    (dict(k=33), "k must be in 1..=32"), (dict(m=0), "m must be in 1..=k"),
    (dict(m=10), "m must be in 1..=k"), (dict(s=0), "s must be in 1..=k and <= 32"),
    (dict(s=10), "s must be in 1..=k and <= 32"), (dict(n_hashes=0), "n_hashes must be positive"),
    (dict(block_slots_exponent=16), "block_slots_exponent must be smaller than counter_slots_exponent"),
    (dict(threads=0), "threads must be positive"), (dict(queue=0), "queue must be positive"),
])
def test_config_validation_matches_reference(kw, msg):  # pipeline_matrix.rs:187-262
    base = dict(k=9, m=5, s=7, n_hashes=3, counter_bits=8, counter_slots_exponent=16,
                block_slots_exponent=6, threads=1, queue=1)
    base.update(kw)
    cfg = scb.SuperCountingBloomConfig(**base)
    with pytest.raises(scb.SuperCountingBloomError) as e:
        cfg.validate()
    assert e.value.kind == "InvalidConfig" and str(e.value) == "invalid config: " + msg
    # the oracle agrees on the message
    oc = O.make_cfg(base["k"], base["m"], base["s"], base["n_hashes"], 8, base["counter_slots_exponent"],
                    base["block_slots_exponent"], base["threads"], base["queue"])
    assert O.validate(oc) == (1, msg)


def test_builder_without_device_fails_loudly():
    if scb.device_count() > 0:
        pytest.skip("a GPU is visible")
    cfg = scb.SuperCountingBloomConfig(k=9, m=5, s=7, n_hashes=3, counter_bits=8,
                                       counter_slots_exponent=16, block_slots_exponent=6, threads=1, queue=1)
    with pytest.raises(scb.SuperCountingBloomError) as e:
        scb.SuperCountingBloomBuilder(cfg)
    assert e.value.kind == "Device"


def _write_scb(path, mutate=None):
    f = O.Filter(O.make_cfg(9, 5, 7, 3, 16, 12, 6, 2, 4))
    f.insert_ascii(b"ACGTACGTACGTTTGCA")
    f.save(path, 9)
    if mutate:
        path.write_bytes(mutate(path.read_bytes()))


@pytest.mark.parametrize("mutate,kind,text", [
    (lambda d: b"not an index", "InvalidIndexFormat", "invalid index format: bad SuperCountingBloom index magic"),
    (lambda d: d[:8] + struct.pack("<I", 1) + d[12:], "InvalidIndexFormat",
     "invalid index format: unsupported index format version 1"),
    (lambda d: d[:44] + b"\x07" + d[45:], "InvalidIndexFormat", "invalid index format: unsupported counter width 7"),
    (lambda d: d[:12] + struct.pack("<Q", 10) + d[20:], "InvalidConfig",
     "invalid config: canonical SIMD minimizers require odd k"),
    (lambda d: d[:71] + struct.pack("<Q", 3) + d[79:], "InvalidIndexFormat",
     "invalid index format: stored filter layout does not match config"),
    (lambda d: d + b"x", "InvalidIndexFormat", "invalid index format: trailing bytes after serialized index"),
    (lambda d: d[:-3], "Io", "failed to fill whole buffer"),
    (lambda d: d[:50], "Io", "failed to fill whole buffer"),
])
def test_load_rejections(tmp_path, mutate, kind, text):  # api.rs:416-464, 1215-1225
    p = tmp_path / "bad.scb"
    _write_scb(p, mutate)
    with pytest.raises(scb.SuperCountingBloomError) as e:
        scb.SuperCountingBloomIndex.load(p)
    assert e.value.kind == kind and text in str(e.value)


def test_load_missing_file(tmp_path):
    with pytest.raises(scb.SuperCountingBloomError) as e:
        scb.SuperCountingBloomIndex.load(tmp_path / "nope.scb")
    assert e.value.kind == "Io" and "No such file or directory" in str(e.value)


def test_fastx_record_iteration_preserves_record_names(tmp_path):  # api.rs:1074-1093
    p = tmp_path / "records.fa"
    p.write_bytes(b">contig_1 description\nACGTNN\n>contig_2\nTGCATG\n")
    names, seqs = [], []
    scb.for_each_fastx_record(p, lambda n, s: (names.append(n), seqs.append(s)))
    assert names == ["contig_1", "contig_2"] and seqs == [b"ACGTNN", b"TGCATG"]


def test_fastx_multiline_fastq_gzip_and_empty_header(tmp_path):
    import gzip
    p = tmp_path / "mixed.fq.gz"
    with gzip.open(p, "wb") as f:
        f.write(b"@r1 x\nACGT\n+\nIIII\n@\nGGCC\n+\nIIII\n")
    names, seqs = [], []
    scb.for_each_fastx_record(p, lambda n, s: (names.append(n), seqs.append(s)))
    assert names == ["r1", "record_2"] and seqs == [b"ACGT", b"GGCC"]
    q = tmp_path / "ml.fa"
    q.write_bytes(b">a\nAC\nGT\r\n\n>b\nTT\n")
    names, seqs = [], []
    scb.for_each_fastx_record(q, lambda n, s: (names.append(n), seqs.append(s)))
    assert seqs == [b"ACGT", b"TT"]


def _py_fasta(data: bytes):
    """A plain FASTA parse: '>' lines start records, '\r\n' and '\n' end lines,
    the remaining lines of a record are joined."""
    names, seqs = [], []
    for line in data.split(b"\n"):
        if line.endswith(b"\r"):
            line = line[:-1]
        if line.startswith(b">"):
            names.append(line[1:].split(b" ")[0].decode())
            seqs.append(b"")
        elif names:
            seqs[-1] += line
    return names, seqs


def test_fastx_lines_across_read_blocks(tmp_path):
    """Lines longer than the 4 MiB read-ahead block, a '\r\n' split across two
    blocks and a last line with no '\n' read like a plain parse."""
    rng = np.random.default_rng(5)
    alpha = np.frombuffer(b"ACGTN", np.uint8)
    parts = []
    block = 4 << 20
    for i, n in enumerate((10, 9_000_000, 70, 3_000_000, 1, 5_000_000)):
        hdr = b">r%d desc\n" % i
        body = alpha[rng.integers(0, 5, n)].tobytes()
        parts += [hdr, body, b"\r\n"]
    data = b"".join(parts)
    # move a '\r\n' onto the boundary: the '\r' ends one block, the '\n' starts the next
    k = data.rfind(b"\r\n", 0, block - 1)
    assert k > 0
    data = data[:k] + b"A" * (block - 1 - k) + data[k:]
    assert data[block - 1:block + 1] == b"\r\n"
    data = data + b">last\nACGTACGT"  # no final newline
    p = tmp_path / "long.fa"
    p.write_bytes(data)
    names, seqs = _records(p)
    want_n, want_s = _py_fasta(data)
    assert names == want_n and [len(s) for s in seqs] == [len(s) for s in want_s]
    assert seqs == want_s


def _zstd_compress(data: bytes) -> bytes:
    import ctypes as C
    z = C.CDLL("libzstd.so.1")
    z.ZSTD_compressBound.restype = C.c_size_t
    z.ZSTD_compressBound.argtypes = [C.c_size_t]
    z.ZSTD_compress.restype = C.c_size_t
    z.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
    cap = z.ZSTD_compressBound(len(data))
    out = C.create_string_buffer(cap)
    n = z.ZSTD_compress(out, cap, data, len(data), 3)
    assert n < cap
    return out.raw[:n]


def _records(p):
    names, seqs = [], []
    scb.for_each_fastx_record(p, lambda n, s: (names.append(n), seqs.append(s)))
    return names, seqs


def test_fastx_zstd_and_multi_member_inputs(tmp_path):
    """zstd (one and several frames) and multi-member gzip read like the plain file."""
    import gzip
    import random
    rng = random.Random(5)
    recs = [(f"r{i} d", "".join(rng.choice("ACGTN") for _ in range(rng.randint(0, 300)))) for i in range(3000)]
    text = "".join(f">{n}\n{s[:70]}\n{s[70:]}\n" for n, s in recs).encode()
    plain = tmp_path / "r.fa"
    plain.write_bytes(text)
    want = _records(plain)
    assert want[0] == [f"r{i}" for i in range(3000)]
    half = len(text) // 2
    cuts = text.index(b">", half)
    (tmp_path / "r.fa.zst").write_bytes(_zstd_compress(text))
    (tmp_path / "r2.fa.zst").write_bytes(_zstd_compress(text[:cuts]) + _zstd_compress(text[cuts:]))
    (tmp_path / "r.fa.gz").write_bytes(gzip.compress(text[:cuts]) + gzip.compress(text[cuts:]))
    for name in ("r.fa.zst", "r2.fa.zst", "r.fa.gz"):
        assert _records(tmp_path / name) == want, name


def test_fastx_corrupt_compressed_input_is_an_io_error(tmp_path):
    import gzip
    data = gzip.compress(b">a\nACGT\n" * 10000)
    p = tmp_path / "bad.fa.gz"
    p.write_bytes(data[:len(data) // 2] + bytes(64) + data[len(data) // 2:])
    with pytest.raises(scb.SuperCountingBloomError) as e:
        _records(p)
    assert e.value.kind == "Io" and "gzip error" in str(e.value)
    z = _zstd_compress(b">a\nACGT\n" * 10000)
    q = tmp_path / "bad.fa.zst"
    q.write_bytes(z[:len(z) - 7])
    with pytest.raises(scb.SuperCountingBloomError) as e:
        _records(q)
    assert e.value.kind == "Io" and "zstd error" in str(e.value)


def test_fastx_open_error(tmp_path):
    with pytest.raises(scb.SuperCountingBloomError) as e:
        scb.for_each_fastx_record(tmp_path / "missing.fa", lambda n, s: None)
    assert e.value.kind == "FastxOpen"
    assert str(e.value).startswith(f"failed to open FASTA/FASTQ '{tmp_path / 'missing.fa'}': ")


def test_write_spectrum(tmp_path):  # spectrum.rs:9-31, pipeline_matrix.rs:84-95
    p = tmp_path / "s.tsv"
    scb.write_spectrum([(1, 10), (3, 2)], p)
    assert p.read_text() == "estimated_count\tkmers\n1\t10\n3\t2\n"


def test_primitives_match_oracle():
    rng = np.random.default_rng(1)
    for x in rng.integers(0, 2**63, size=50, dtype=np.uint64):
        assert scb.xorshift_u64(int(x)) == O.xorshift_u64(int(x))
    seq = b"ACGTTGCAGGCATTACAGCATTTACGGACTAGGACCAT"
    for L in (1, 7, 21, 25, 31, 32):
        for pos in (0, 3):
            assert scb.canonical_kmer(seq, L, pos) == O.canonical_kmer(seq, L, pos)


def test_sampled_spectrum_host_checks_and_csv(tmp_path):
    """bin/sampled_spectrum.rs:100-105: a zero sample size is an InvalidConfig
    before any work; write_scaled_spectrum's CSV (225-253)."""
    L = scb.api.lib()
    n = C.c_size_t()
    rc = L.scb_index_sampled_spectrum(None, None, None, 0, 0, 1, None, 0, C.byref(n), None)
    assert rc == 1 and _lib.last_error() == "invalid config: sample_size must be positive"
    s = scb.SampledSpectrum([(1, 5, 50, 50), (3, 2, 20, 7)], 70, 7, 7, 0)
    p = tmp_path / "s.csv"
    s.write(p)
    assert p.read_text() == ("estimated_count,sampled_kmers,linear_scaled_windows,estimated_distinct_kmers\n"
                             "1,5,50,50\n3,2,20,7\n")
