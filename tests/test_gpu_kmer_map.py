"""GPU tests of the per-key estimate map (filter.rs:302-335) and of spectra
counted over many batches with one key set (fastx.rs:255-308, 392-400): every
(key, estimate) pair and spectrum row equals the oracle's."""
import numpy as np
import pytest

from scbf_amd import synth
from test_oracle import BACKEND_CASES

pytestmark = pytest.mark.gpu


def _cfg(scb, k, m, s, h, bits, cse, bse):
    return scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                        block_slots_exponent=bse, threads=2, queue=4)


@pytest.mark.parametrize("bits,k,m,s,seq", BACKEND_CASES)
def test_backend_matrix_map_query(gpu, oracle, bits, k, m, s, seq):
    """tests/backend_matrix.rs:71-79 through the C ABI: after two inserts, the
    map query counts len - k + 1 k-mers, the map is not empty and every
    estimate is positive; the pairs equal the oracle's map."""
    scb = gpu
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, 3, bits, 16, 6))
    b.add_sequence(seq)
    b.add_sequence(seq)
    idx = b.into_index()
    got, queried = idx.estimate_kmer_map([seq])
    assert queried == len(seq) + 1 - k
    assert got and all(v > 0 for v in got.values())
    of = oracle.Filter(oracle.make_cfg(k, m, s, 3, bits, 16, 6))
    of.insert_ascii(seq)
    of.insert_ascii(seq)
    sp = oracle.Spectrum()
    assert sp.add_chunk(of, seq) == queried
    ok, ov = sp.pairs()
    assert got == dict(zip(ok.tolist(), ov.tolist()))


GEOS = [(31, 21, 25, 4, 16, 22, 9), (21, 15, 17, 3, 32, 20, 10), (9, 5, 7, 3, 8, 16, 6), (31, 21, 25, 4, 16, 20, 15)]


@pytest.mark.parametrize("geo", GEOS, ids=[f"k{g[0]}u{g[4]}c{g[5]}b{g[6]}" for g in GEOS])
def test_kmer_map_parity(gpu, oracle, geo):
    """Reads with N runs, lowercase and repeats: the sorted (key, estimate)
    pairs equal the oracle's merged map; keys are canonical k-mers."""
    scb = gpu
    k, m, s, h, bits, cse, bse = geo
    g = synth.genome(40_000, 61)
    ib, io = synth.fixed_reads(g, 2000, 150, 62, sub_rate=0.01, n_rate=0.002)
    qb1, qo1 = synth.fixed_reads(g, 800, 150, 63, sub_rate=0.02, n_rate=0.002)
    qb2, qo2 = synth.random_records(200, 0, 400, 64, n_rate=0.01, lower_rate=0.02)
    qb, qo = synth.concat_sets([(qb1, qo1), (qb2, qo2), (qb1[:15000], qo1[:101])])
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.add_sequences((ib, io))
    idx = b.into_index()
    keys, vals, queried = idx.estimate_kmer_map_np((qb, qo))
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    of.insert_records(ib, io)
    sp = oracle.Spectrum()
    assert sp.add_records(of, qb, qo) == queried
    ok, ov = sp.pairs()
    assert np.array_equal(keys, ok) and np.array_equal(vals, ov)
    assert keys.size > 100 and np.all(keys[1:] > keys[:-1])
    # the spectrum is the histogram of the map (fastx.rs:402-410)
    rows, q, d = idx.spectrum((qb, qo))
    u, c = np.unique(vals, return_counts=True)
    assert rows == [(int(a), int(b)) for a, b in zip(u, c)] and (q, d) == (queried, keys.size)


@pytest.mark.parametrize("geo", GEOS[:3], ids=[f"k{g[0]}u{g[4]}" for g in GEOS[:3]])
def test_spectrum_accumulator_counts_each_key_once(gpu, oracle, geo):
    """Query reads split into batches that share most of their k-mers (a read
    set repeated and overlapping): the accumulated rows, distinct keys and
    queried k-mers equal the oracle's over all batches and one spectrum call
    over their concatenation; the key set grows past its first capacity."""
    scb = gpu
    k, m, s, h, bits, cse, bse = geo
    g = synth.genome(300_000, 71)
    ib, io = synth.fixed_reads(g, 10_000, 150, 72, sub_rate=0.01)
    qb, qo = synth.fixed_reads(g, 12_000, 150, 73, sub_rate=0.02, n_rate=0.001)
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.add_sequences((ib, io))
    idx = b.into_index()
    acc = idx.spectrum_accumulator()
    batches = [(0, 4000), (3000, 8000), (0, 12_000), (7000, 12_000), (11_990, 12_000)]
    for a, e in batches:
        acc.add((qb[int(qo[a]):int(qo[e])], (qo[a:e + 1] - qo[a]).astype(np.uint64)))
    rows, queried, distinct = acc.rows()
    all_b, all_o = synth.concat_sets([(qb[int(qo[a]):int(qo[e])], (qo[a:e + 1] - qo[a]).astype(np.uint64))
                                      for a, e in batches])
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    of.insert_records(ib, io)
    sp = oracle.Spectrum()
    oq = sp.add_records(of, all_b, all_o)
    assert (queried, distinct) == (oq, sp.distinct())
    assert rows == sp.rows()
    one_rows, one_q, one_d = idx.spectrum((all_b, all_o))
    assert (one_rows, one_q, one_d) == (rows, queried, distinct)
    assert distinct > (1 << 15)  # beyond the first key-set capacity (2^16 slots at half load)


def test_fastx_spectrum_streams_file_batches(gpu, oracle, tmp_path, monkeypatch):
    """estimate_spectrum_from_fastx_pair reads the query file in batches
    (1 MiB here through SCB_FILE_BATCH_BYTES, about 40 batches) with one key
    set: queried k-mers, distinct keys and rows equal the oracle's over the whole
    file, and equal the result with the default 256 MiB batches."""
    scb = gpu
    k, m, s, h, bits, cse, bse = 31, 21, 25, 4, 16, 22, 9
    g = synth.genome(200_000, 81)
    ib, io = synth.fixed_reads(g, 5000, 150, 82, sub_rate=0.01, n_rate=0.001)
    qb, qo = synth.fixed_reads(g, 140_000, 300, 83, sub_rate=0.01, n_rate=0.001)

    def fasta(path, b, o):
        with open(path, "wb") as f:
            for i in range(o.size - 1):
                f.write(b">r\n" + bytes(b[int(o[i]):int(o[i + 1])]) + b"\n")

    fi, fq = tmp_path / "i.fa", tmp_path / "q.fa"
    fasta(fi, ib, io)
    fasta(fq, qb, qo)
    cfg = _cfg(scb, k, m, s, h, bits, cse, bse)
    monkeypatch.setenv("SCB_FILE_BATCH_BYTES", str(1 << 20))
    small = scb.estimate_spectrum_from_fastx_pair(fi, fq, cfg)
    monkeypatch.delenv("SCB_FILE_BATCH_BYTES")
    whole = scb.estimate_spectrum_from_fastx_pair(fi, fq, cfg)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    of.insert_records(ib, io)
    sp = oracle.Spectrum()
    oq = sp.add_records(of, qb, qo)
    for rep in (small, whole):
        assert (rep.stats.queried_kmers, rep.stats.distinct_keys) == (oq, sp.distinct())
        assert rep.spectrum == sp.rows()
