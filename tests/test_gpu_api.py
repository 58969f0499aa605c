"""GPU tests of the crate-level API through the C ABI: the reference's own
pipeline/api tests (tests/pipeline_matrix.rs, src/api.rs tests, src/fastx.rs
tests) re-run against libscb.so, plus spectrum, save/load and histogram parity
with the oracle."""
import numpy as np
import pytest

from scbf_amd import synth

pytestmark = pytest.mark.gpu

READS = b">r1\nACGTACGTACGTACGTACGTACGTACGT\n>r2\nTGCATGCAACGTACGTTGCATGCAACGT\n"   # pipeline_matrix.rs:27-35
NEG = b">n1\nTTTTTTTTTTTTTTTTTTTTTTTTTTTT\n>n2\nCCCCCCCCCCCCCCCCCCCCCCCCCCCC\n"     # pipeline_matrix.rs:37-45


def pm_cfg(scb, bits, threads, queue):
    return scb.SuperCountingBloomConfig(k=9, m=5, s=7, n_hashes=3, counter_bits=bits, counter_slots_exponent=16,
                                        block_slots_exponent=6, threads=threads, queue=queue)


CASES = [(b, t, q) for (t, q) in ((1, 1), (2, 2), (4, 8), (3, 3)) for b in (8, 16, 32)]


@pytest.mark.parametrize("bits,threads,queue", CASES)
def test_pipeline_matrix(gpu, tmp_path, bits, threads, queue):
    scb = gpu
    reads = tmp_path / "reads.fa"
    reads.write_bytes(READS)
    neg = tmp_path / "neg.fa"
    neg.write_bytes(NEG)
    cfg = pm_cfg(scb, bits, threads, queue)
    rep = scb.estimate_spectrum_from_fastx(reads, cfg)                       # 61-70
    assert (rep.stats.chunks, rep.stats.indexed_kmers, rep.stats.queried_kmers) == (2, 40, 40)
    assert rep.spectrum
    st = scb.stream_estimates_from_fastx(reads, cfg).stats                   # 72-82
    assert (st.chunks, st.indexed_kmers, st.queried_kmers, st.positive_estimates) == (2, 40, 40, 40)
    assert st.estimate_checksum > 0
    out = tmp_path / "spectrum.tsv"                                           # 84-95
    scb.write_spectrum(rep.spectrum, out)
    text = out.read_text()
    assert text.startswith("estimated_count\tkmers\n") and len(text.splitlines()) > 1
    pos = scb.stream_estimates_from_fastx_pair(reads, reads, cfg).stats     # 97-117
    negr = scb.stream_estimates_from_fastx_pair(reads, neg, cfg).stats
    assert pos.indexed_kmers == negr.indexed_kmers
    assert pos.queried_kmers == 40 and negr.queried_kmers == 40 and pos.positive_estimates == 40
    assert pos.estimate_checksum >= negr.estimate_checksum
    assert scb.estimate_spectrum_from_fastx_pair(reads, reads, cfg).spectrum


def test_fastx_smoke_all_widths(gpu, tmp_path):  # fastx.rs:448-466
    scb = gpu
    p = tmp_path / "reads.fa"
    p.write_bytes(b">r1\nACGTACGTACGTACGTACGT\n>r2\nACGTACGTACGTACGTACGT\n")
    for bits in (8, 16, 32):
        rep = scb.estimate_spectrum_from_fastx(p, pm_cfg(scb, bits, 2, 2))
        assert rep.spectrum and (rep.stats.chunks, rep.stats.indexed_kmers, rep.stats.queried_kmers) == (2, 24, 24)


def small_cfg(scb, **kw):  # api.rs:1044-1056
    c = dict(k=5, m=3, s=4, n_hashes=2, counter_bits=8, counter_slots_exponent=12, block_slots_exponent=5,
             threads=1, queue=2)
    c.update(kw)
    return scb.SuperCountingBloomConfig(**c)


def test_api_unit_tests(gpu, tmp_path):
    scb = gpu
    p = tmp_path / "r.fa"
    p.write_bytes(b">r\nACGTACGTACGT\n")
    idx = scb.SuperCountingBloomIndex.build_from_fastx(p, small_cfg(scb)).index
    ab = idx.estimate_sequence_abundances(b"ACGTACNNACGTAC")                    # 1058-1072
    assert len(ab) == 10 and all(a is not None for a in ab[:2])
    assert all(a is None for a in ab[2:8]) and all(a is not None for a in ab[8:])
    summ = idx.summarize_sequence("query", b"ACGTACGT")                        # 1095-1111
    assert (summ.name, summ.sequence_len, summ.total_windows, summ.valid_kmers) == ("query", 8, 4, 4)
    assert summ.mean_abundance > 0 and summ.median_abundance > 0
    assert idx.filter_only_count_hint() > 0                                    # 1113-1127
    sp = idx.estimate_filter_only_spectrum(8, 10)
    assert sp.method == "compound_poisson_counter_deconvolution"
    assert sp.rows[0][0] == 1 and sp.rows[-1][0] == 8


@pytest.mark.parametrize("bits", [8, 16, 32])
def test_save_load_roundtrip(gpu, tmp_path, bits):  # api.rs:1129-1156
    scb = gpu
    reads = tmp_path / "rt.fa"
    reads.write_bytes(b">r1\nACGTACGTACGTACGT\n>r2\nTTTTACGTACGTAAAA\n")
    idx = scb.SuperCountingBloomIndex.build_from_fastx(reads, small_cfg(scb, counter_bits=bits)).index
    path = tmp_path / "rt.scb"
    idx.save(path)
    loaded = scb.SuperCountingBloomIndex.load(path)
    assert loaded.config() == idx.config()
    assert loaded.counter_histogram() == idx.counter_histogram()
    q = b"ACGTACGTNNACGTACGT"
    assert loaded.estimate_sequence_abundances(q) == idx.estimate_sequence_abundances(q)
    assert path.stat().st_size == 139 + (1 << 12) * bits // 8


def test_builder_add_sequence_add_fasta_query_and_save_load(gpu, tmp_path):  # api.rs:1158-1202
    scb = gpu
    reads = tmp_path / "b.fa"
    reads.write_bytes(b">r1\nACGTACGTACGT\n>r2\nACGTNNACGTAC\n")
    b = scb.SuperCountingBloomBuilder(small_cfg(scb, counter_bits=16, threads=2, queue=1))
    assert b.threads() == 2
    b.set_threads(1)
    assert b.threads() == 1
    b.clear_threads()
    assert b.threads() >= 1
    b.set_threads(2)
    assert b.add_sequence(b"ACGTACNNACGTAC") == 4
    rep = b.add_fasta(reads)
    assert (rep.records_processed, rep.records_indexed, rep.inserted_kmers) == (2, 2, 10)
    assert b.inserted_kmers() == 14 and b.inserted_smers() > 0
    idx = b.into_index()
    assert idx.inserted_kmers() == 14
    q = idx.query_fasta(reads)
    assert (q.records_processed, q.total_windows, q.valid_kmers) == (2, 16, 10)
    assert q.positive_estimates > 0 and q.estimate_checksum > 0
    path = tmp_path / "b.scb"
    idx.save(path)
    loaded = scb.SuperCountingBloomIndex.load(path)
    assert loaded.inserted_kmers() == 14
    assert loaded.query_fasta(reads) == q


def test_builder_rejects_zero_threads(gpu):  # api.rs:1204-1213
    scb = gpu
    b = scb.SuperCountingBloomBuilder(small_cfg(scb))
    with pytest.raises(scb.SuperCountingBloomError) as e:
        b.set_threads(0)
    assert e.value.kind == "InvalidConfig"


SPEC_GEOS = [(9, 5, 7, 3, 8, 16, 6), (31, 21, 25, 4, 16, 20, 9), (21, 15, 17, 3, 32, 18, 10),
             (31, 11, 15, 2, 16, 18, 8), (5, 3, 4, 2, 8, 12, 5), (31, 21, 25, 4, 16, 22, 16)]


@pytest.mark.parametrize("method", [0, 1], ids=["batched", "plain"])
@pytest.mark.parametrize("geo", SPEC_GEOS, ids=[f"k{g[0]}s{g[2]}u{g[4]}c{g[5]}b{g[6]}" for g in SPEC_GEOS])
def test_spectrum_parity(gpu, oracle, geo, method):
    scb = gpu
    k, m, s, h, bits, cse, bse = geo
    print(f"spectrum case {geo} method {method}", flush=True)
    g = synth.genome(30_000, 21)
    ib, io = synth.fixed_reads(g, 1500, 150, 22, sub_rate=0.01, n_rate=0.002)
    qb, qo = synth.fixed_reads(g, 700, 150, 23, sub_rate=0.02, n_rate=0.002)
    cfg = scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                       block_slots_exponent=bse, threads=1, queue=4)
    b = scb.SuperCountingBloomBuilder(cfg)
    b.add_sequences((ib, io))
    idx = b.into_index()
    idx.set_method(method)
    # query as DnaChunks (split_non_actg): maximal ACGT runs >= k of every read
    chunks = []
    for i in range(qo.size - 1):
        r = bytes(qb[int(qo[i]):int(qo[i + 1])])
        for part in r.replace(b"N", b" ").split():
            if len(part) >= k:
                chunks.append(part)
    rows, queried, distinct = idx.spectrum(chunks)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    oracle.insert_records(of, ib, io)
    sp = oracle.Spectrum()
    oq = sum(sp.add_chunk(of, c) for c in chunks)
    assert queried == oq and distinct == sp.distinct()
    assert rows == sp.rows()
    assert sum(c for _, c in rows) == distinct   # fastx.rs:161


@pytest.mark.parametrize("bits", [16, 32])
def test_spectrum_key_set_overflow_parity(gpu, oracle, bits):
    """Four 2^14-slot spectrum bins, each with ~8,000 distinct keys: far more than
    the 4,096-key LDS set of a work item, so most keys find their probe window
    full and go to the global set, which then overflows its first size and the
    pass runs again with a larger one. Rows, queried and distinct keys must still
    equal the oracle's."""
    scb = gpu
    k, m, s, h, cse, bse = 31, 21, 25, 4, 16, 9
    print(f"spectrum overflow case u{bits}", flush=True)
    g = synth.genome(40_000, 31)
    ib, io = synth.fixed_reads(g, 600, 150, 32, sub_rate=0.01)
    qb, qo = synth.random_records(250, 150, 151, 33)  # random sequence: every k-mer distinct
    cfg = scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                       block_slots_exponent=bse, threads=1, queue=4)
    b = scb.SuperCountingBloomBuilder(cfg)
    b.add_sequences((ib, io))
    idx = b.into_index()
    idx.set_method(0)  # batched: the binned key sets
    chunks = [bytes(qb[int(qo[i]):int(qo[i + 1])]) for i in range(qo.size - 1)]
    chunks = chunks + chunks[:100]  # repeated keys as well
    rows, queried, distinct = idx.spectrum(chunks)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    oracle.insert_records(of, ib, io)
    sp = oracle.Spectrum()
    oq = sum(sp.add_chunk(of, c) for c in chunks)
    assert queried == oq and distinct == sp.distinct()
    assert distinct > 4 * 6000  # more keys than four LDS sets hold
    assert rows == sp.rows()
    # the accumulator's emitting pass takes the same overflow path
    acc = idx.spectrum_accumulator()
    acc.add(chunks)
    acc.add(chunks[:50])
    assert acc.rows() == (rows, queried + sum(max(0, len(c) - k + 1) for c in chunks[:50]), distinct)


SPEC_DUP_GEOS = SPEC_GEOS + [(9, 5, 7, 2, 32, 1, 0), (9, 5, 7, 3, 8, 4, 2)]  # bins of 8 and 4 bytes


@pytest.mark.parametrize("method", [0, 1], ids=["batched", "plain"])
@pytest.mark.parametrize("geo", SPEC_DUP_GEOS, ids=[f"k{g[0]}s{g[2]}u{g[4]}c{g[5]}b{g[6]}" for g in SPEC_DUP_GEOS])
def test_spectrum_repeated_chunks_parity(gpu, oracle, geo, method):
    """Every chunk five times over plus homopolymers and tandem repeats: most
    records of a bin repeat an earlier one and most keys are found present;
    queried, distinct and the rows equal the oracle's, also for bins of 8 and 4
    bytes."""
    scb = gpu
    k, m, s, h, bits, cse, bse = geo
    print(f"spectrum repeated case {geo} method {method}", flush=True)
    g = synth.genome(20_000, 31)
    ib, io = synth.fixed_reads(g, 1000, 150, 32, sub_rate=0.01)
    qb, qo = synth.fixed_reads(g, 300, 150, 33, sub_rate=0.01)
    reads = [bytes(qb[int(qo[i]):int(qo[i + 1])]) for i in range(qo.size - 1)]
    for unit in (b"A", b"AC", b"ACGT", b"AAC", b"CAGCAG"):
        reads.append((unit * 200)[:150])
    chunks = [r for r in reads for _ in range(5) if len(r) >= k]
    cfg = scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                       block_slots_exponent=bse, threads=1, queue=4)
    b = scb.SuperCountingBloomBuilder(cfg)
    b.add_sequences((ib, io))
    b.add_sequence(b"".join(reads[-5:]))
    idx = b.into_index()
    idx.set_method(method)
    rows, queried, distinct = idx.spectrum(chunks)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    oracle.insert_records(of, ib, io)
    of.insert_ascii(b"".join(reads[-5:]))
    sp = oracle.Spectrum()
    oq = sum(sp.add_chunk(of, c) for c in chunks)
    assert queried == oq and distinct == sp.distinct()
    assert rows == sp.rows()


def test_counter_histogram_after_load_matches_oracle(gpu, oracle, tmp_path):
    scb = gpu
    cfg_o = oracle.make_cfg(21, 15, 17, 3, 32, 14, 8, 1, 1)
    of = oracle.Filter(cfg_o)
    b, o = synth.random_records(200, 50, 300, 5)
    oracle.insert_records(of, b, o)
    for _ in range(5000):  # drive some counters above the 4096 LDS histogram range
        of.insert_packed(b"ACGTTGCATGTCAGTACGATC")
    p = tmp_path / "h.scb"
    of.save(p, 0)
    idx = scb.SuperCountingBloomIndex.load(p)
    assert idx.counter_histogram() == of.counter_histogram()
    assert np.array_equal(idx.counters().astype(np.uint32), of.counters())


def test_cpp_wrapper_builder_roundtrip(gpu, tmp_path):
    """include/scb.hpp (the C++ mirror of the crate API) over libscb.so."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "api_smoke"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{root}/include", f"{root}/tests/cpp/api_smoke.cpp",
                    f"-L{root}/scbf_amd", "-lscb", f"-Wl,-rpath,{root}/scbf_amd", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "cpp api ok" in out.stdout


def _write_fasta(path, records):
    path.write_bytes(b"".join(b">r%d\n%s\n" % (i, r) for i, r in enumerate(records)))


def _ragged_records(seed, n):
    """Records of 0..400 bases with N runs, runs shorter than k, lowercase stretches
    and other IUPAC letters: every record splits into several chunks or none."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        r = bytearray(rng.choice(np.frombuffer(b"ACGT", np.uint8), size=int(rng.integers(0, 400))).tobytes())
        for _ in range(int(rng.integers(0, 5))):
            if not r:
                break
            a = int(rng.integers(0, len(r)))
            run = rng.choice(np.frombuffer(b"NNNNRY", np.uint8), size=int(rng.integers(1, 8))).tobytes()
            r[a:a + len(run)] = run
        if r and rng.random() < 0.3:
            a = int(rng.integers(0, len(r)))
            r[a:a + 40] = bytes(r[a:a + 40]).lower()
        out.append(bytes(r))
    return out


@pytest.mark.parametrize("bits", [8, 16, 32])
def test_fastx_pipeline_with_n_runs_matches_oracle(gpu, oracle, tmp_path, bits):
    """Whole records go to the device and chunks (ACGT runs of at least k bases,
    fastx.rs:14-19) are counted on the host: chunk counts, indexed and queried
    k-mers, s-mers and the estimate checksum of the stream pipeline equal the
    oracle's restatement of the threaded pipeline (fastx.rs:172-204, 206-369)."""
    scb = gpu
    k, m, s, h, cse, bse = 21, 11, 15, 3, 18, 8
    ins, qry = _ragged_records(21, 400), _ragged_records(22, 300) + _ragged_records(21, 100)
    fi, fq = tmp_path / "index.fa", tmp_path / "query.fa"
    _write_fasta(fi, ins)
    _write_fasta(fq, qry)
    cfg = scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                       block_slots_exponent=bse, threads=2, queue=4)
    st = scb.stream_estimates_from_fastx_pair(fi, fq, cfg).stats
    build = scb.SuperCountingBloomIndex.build_from_fastx(fi, cfg)

    def concat(recs):
        offs = np.zeros(len(recs) + 1, dtype=np.uint64)
        offs[1:] = np.cumsum([len(r) for r in recs])
        return np.frombuffer(b"".join(recs), dtype=np.uint8).copy(), offs

    ib, io = concat(ins)
    qb, qo = concat(qry)
    ob = oracle.cpu_baseline(oracle.make_cfg(k, m, s, h, bits, cse, bse, 2, 4), ib, io, qb, qo)
    got = (st.chunks, st.indexed_kmers, st.inserted_smers, st.queried_kmers, st.positive_estimates,
           st.estimate_checksum)
    want = (ob.chunks, ob.indexed_kmers, ob.inserted_smers, ob.queried_kmers, ob.positive_estimates,
            ob.estimate_checksum)
    assert got == want
    assert (build.chunks, build.indexed_kmers, build.inserted_smers) == want[:3]
    assert st.chunks > len(ins)  # the N runs split records into several chunks


@pytest.mark.parametrize("bits", [8, 16, 32])
def test_host_estimators_and_scb_bytes_equal_oracle(gpu, oracle, tmp_path, bits):
    """On a real index at each counter width: the filter-only spectrum and count
    hint of the index's own histogram (api.rs:337-385, 933-1033), the same on
    given histograms, and summarize_sequence (api.rs:264-271, 855-891) equal the
    oracle's; scb_index_save writes exactly the oracle's .scb bytes, header
    (with non-default threads and queue) and payload (api.rs:388-480)."""
    scb = gpu
    k, m, s, h, cse, bse = (21, 15, 17, 3, 22, 10) if bits == 32 else (31, 21, 25, 4, 22, 9)
    g = synth.genome(200_000, 3)
    ib, io = synth.fixed_reads(g, 20_000, 150, 4, sub_rate=0.01, n_rate=0.001)
    cfg = scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                       block_slots_exponent=bse, threads=3, queue=77)
    b = scb.SuperCountingBloomBuilder(cfg)
    b.add_sequences((ib, io))
    idx = b.into_index()
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse, threads=3, queue=77))
    kmers, _ = of.insert_records(ib, io)
    assert idx.inserted_kmers() == kmers
    ohist = of.counter_histogram()
    assert idx.counter_histogram() == ohist
    hint = idx.filter_only_count_hint()
    assert hint == oracle.filter_only_count_hint(ohist, h, 255)
    assert idx.estimate_filter_only_spectrum(255, hint).rows == oracle.filter_only_spectrum(ohist, h, 255, hint)
    for max_ab, given_hint in ((8, 1000), (64, 12345)):
        got = idx.estimate_filter_only_spectrum_from_counter_histogram(ohist, max_ab, given_hint).rows
        assert got == oracle.filter_only_spectrum(ohist, h, max_ab, given_hint)
        assert idx.filter_only_count_hint_from_counter_histogram(ohist, max_ab) == \
            oracle.filter_only_count_hint(ohist, h, max_ab)
    for seq in (ib[:1500], ib[:k - 1], b"ACGTNNNN" * 20 + bytes(ib[3000:3400])):
        summ = idx.summarize_sequence("q", seq)
        v, ok = of.estimate_ascii(seq)
        n, mean, med = oracle.summarize(v, ok)
        assert (summ.total_windows, summ.valid_kmers) == (v.size, n)
        if n:
            assert (summ.mean_abundance, summ.median_abundance) == (mean, med)
        else:
            assert summ.mean_abundance is None and summ.median_abundance is None
    ours, theirs = tmp_path / "gpu.scb", tmp_path / "oracle.scb"
    idx.save(ours)
    rc, msg = of.save(theirs, kmers)
    assert rc == 0, msg
    a, bb = ours.read_bytes(), theirs.read_bytes()
    assert len(a) == len(bb) == 139 + (1 << cse) * bits // 8
    assert a[:139] == bb[:139], "header differs"
    assert a == bb, "payload differs"


def test_device_batches_with_bad_offsets_are_refused(gpu):
    """The offsets contract of the HBM-resident entry points (offsets[0] = 0,
    never decreasing, offsets[n] = total_bases) is checked on the device."""
    from hipmem import DeviceBuffer
    scb = gpu
    b = scb.SuperCountingBloomBuilder(small_cfg(scb, counter_bits=16))
    bases = np.frombuffer(b"ACGTACGTTGCAACGT" * 64, dtype=np.uint8).copy()
    db = DeviceBuffer.from_array(bases)
    for offs in ([0, 600, 500, 1024], [0, 512, 1000], [8, 512, 1024]):
        o = np.array(offs, dtype=np.uint64)
        do = DeviceBuffer.from_array(o)
        with pytest.raises(scb.SuperCountingBloomError) as e:
            b.add_sequences_device(db.ptr, do.ptr, o.size - 1, bases.size)
        assert e.value.kind == "InternalState" and "offsets" in str(e.value)
    good = DeviceBuffer.from_array(np.array([0, 512, 1024], dtype=np.uint64))
    assert b.add_sequences_device(db.ptr, good.ptr, 2, bases.size).inserted_kmers == 2 * (512 - 5 + 1)
    idx = b.into_index()
    bad = DeviceBuffer.from_array(np.array([0, 700, 300, 1024], dtype=np.uint64))
    with pytest.raises(scb.SuperCountingBloomError):
        idx.estimate_sequences_device(db.ptr, bad.ptr, 3, bases.size, None, None)
    rep = idx.estimate_sequences_device(db.ptr, good.ptr, 2, bases.size, None, None)
    assert rep.valid_kmers == 2 * (512 - 5 + 1)


def test_spectrum_refuses_blocks_beyond_2e31_slots(gpu):
    """counter_slots_exponent 33 with 2^32-slot blocks is a valid config
    (config.rs:76) that the GPU spectrum cannot address: a clear error."""
    scb = gpu
    b = scb.SuperCountingBloomBuilder(small_cfg(scb, k=9, m=5, s=7, counter_slots_exponent=33,
                                                block_slots_exponent=32))
    b.add_sequence(b"ACGTACGTTGCAACGTAGGT")
    idx = b.into_index()
    with pytest.raises(scb.SuperCountingBloomError) as e:
        idx.spectrum([b"ACGTACGTTGCAACGTAGGT"])
    assert e.value.kind == "InternalState" and "2^31" in str(e.value)
