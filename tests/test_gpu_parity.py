"""GPU parity: counters, per-k-mer abundances and stream reports of the HIP
engine (batched and plain methods) must equal the CPU oracle bit for bit on the
same seeded inputs."""
import zlib

import numpy as np
import pytest

from scbf_amd import synth

pytestmark = pytest.mark.gpu

METHODS = (0, 1)  # batched, plain

# (name, k, m, s, h, bits, cse, bse)
GEOMETRIES = [
    ("k9m5s7_u8", 9, 5, 7, 3, 8, 16, 6),        # tests/backend_matrix.rs:8-20
    ("k9m5s9_u16", 9, 5, 9, 3, 16, 16, 6),
    ("k11m5s8_u32", 11, 5, 8, 3, 32, 16, 6),
    ("k13m7s9_u16", 13, 7, 9, 3, 16, 16, 6),
    ("k5m3s4_u8", 5, 3, 4, 2, 8, 12, 5),          # api.rs:1044-1056
    ("default_small_u16", 31, 21, 25, 4, 16, 20, 9),
    ("k21m15s17_u32", 21, 15, 17, 3, 32, 20, 10),
    ("k31m11s15_u16_widewin", 31, 11, 15, 2, 16, 18, 8),  # fimpera window 17 (> 8)
    ("k31m21s25_u8_sat", 31, 21, 25, 4, 8, 14, 9),  # dense: saturates u8
    ("k31m21s25_u16_bse15", 31, 21, 25, 4, 16, 20, 15),  # 15-bit slot addresses (largest staged block)
    ("k15m9s11_u32_h1", 15, 9, 11, 1, 32, 18, 12),  # one hash per s-mer
    ("k31m21s30_u16_win2", 31, 21, 30, 3, 16, 20, 9),  # s-mers of 60 bits, fimpera window 2
    # a 2^15-slot table: even the smallest insert bins (2^12 slots) hold more
    # records than one work item, so items split at 2048 records merge into the
    # table with atomics, in sub-batches of ~960 records each
    ("k31m21s25_u16_onebin", 31, 21, 25, 4, 16, 15, 9),
    # tiny tables (config.rs:76 admits any bse < cse): one bin smaller than the
    # merge vectors, 16-bit counts saturating across insert sub-batches
    ("k9m5s7_u16_16slots", 9, 5, 7, 3, 16, 4, 1),
    ("k9m5s7_u8_16slots", 9, 5, 7, 3, 8, 4, 2),
    ("k9m5s7_u32_2slots", 9, 5, 7, 2, 32, 1, 0),
]


def _cfg(scb, k, m, s, h, bits, cse, bse):
    return scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits,
                                        counter_slots_exponent=cse, block_slots_exponent=bse,
                                        threads=2, queue=4)


def _records(seed, k):
    # mixed lengths (incl. < k), N runs, lowercase, empty records
    b1, o1 = synth.random_records(300, 0, 400, seed, n_rate=0.01, lower_rate=0.02)
    g = synth.genome(20000, seed)
    b2, o2 = synth.fixed_reads(g, 800, 150, seed, sub_rate=0.01, n_rate=0.001)
    bases = np.concatenate([b1, b2])
    offsets = np.concatenate([o1, o2[1:] + o1[-1]]).astype(np.uint64)
    return bases, offsets


@pytest.mark.parametrize("geo", GEOMETRIES, ids=[g[0] for g in GEOMETRIES])
@pytest.mark.parametrize("method", METHODS, ids=["batched", "plain"])
def test_insert_query_parity(gpu, oracle, geo, method):
    scb = gpu
    name, k, m, s, h, bits, cse, bse = geo
    print(f"case {name} method={method}", flush=True)
    bases, offsets = _records(zlib.crc32(name.encode()) & 0xFFFF, k)
    ocfg = oracle.make_cfg(k, m, s, h, bits, cse, bse)
    of = oracle.Filter(ocfg)
    ok_kmers, ok_smers = oracle.insert_records(of, bases, offsets)
    if name.endswith("_sat"):  # repeat the insert to saturate 8-bit counters
        for _ in range(3):
            oracle.insert_records(of, bases, offsets)

    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.set_method(method)
    rep = b.add_sequences((bases, offsets))
    if name.endswith("_sat"):
        for _ in range(3):
            b.add_sequences((bases, offsets))
    assert rep.inserted_kmers == ok_kmers
    assert rep.inserted_smers == ok_smers
    assert rep.records_processed == offsets.size - 1
    got = b.counters().astype(np.uint32)
    want = of.counters()
    assert np.array_equal(got, want), f"counters differ at {np.flatnonzero(got != want)[:10]}"

    idx = b.into_index()
    idx.set_method(method)
    qb, qo = _records(1000 + (zlib.crc32(name.encode()) & 0xFFFF), k)
    qb = np.concatenate([qb, bases])
    qo = np.concatenate([qo, offsets[1:] + qo[-1]]).astype(np.uint64)
    vals, valid, qrep = idx.estimate_sequences(qb, qo)
    wv, wok = oracle.estimate_records(of, qb, qo, k)
    assert np.array_equal(valid, wok)
    assert np.array_equal(vals[wok], wv[wok])
    assert qrep.valid_kmers == int(wok.sum())
    assert qrep.positive_estimates == int((wv[wok] > 0).sum())
    assert qrep.estimate_checksum == int(wv[wok].sum()) & 0xFFFFFFFFFFFFFFFF


@pytest.mark.parametrize("method", METHODS, ids=["batched", "plain"])
def test_single_sequence_api(gpu, oracle, method):
    scb = gpu
    k, m, s, h, bits, cse, bse = 5, 3, 4, 2, 8, 12, 5
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.set_method(method)
    assert b.add_sequence(b"ACGTACNNACGTAC") == 4   # api.rs:1178
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    of.insert_ascii(b"ACGTACNNACGTAC")
    idx = b.into_index()
    idx.set_method(method)
    got = idx.estimate_sequence_abundances(b"ACGTACNNACGTAC")
    v, ok = of.estimate_ascii(b"ACGTACNNACGTAC")
    assert got == [int(x) if o else None for x, o in zip(v, ok)]
    assert len(got) == 10
    assert all(x is not None for x in got[:2]) and all(x is None for x in got[2:8])


@pytest.mark.parametrize("bits", [8, 16, 32])
def test_counter_histogram(gpu, oracle, bits):
    scb = gpu
    k, m, s, h, cse, bse = 21, 15, 17, 3, 16, 8
    bases, offsets = _records(77, k)
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    for _ in range(40):
        b.add_sequences((bases[:5000], offsets[offsets <= 5000]))
    for _ in range(40):
        oracle.insert_records(of, bases[:5000], offsets[offsets <= 5000])
    idx = b.into_index()
    assert idx.counter_histogram() == of.counter_histogram()


@pytest.mark.parametrize("bits", [8, 16])
def test_repeated_records_parity(gpu, oracle, bits):
    """One read 300 times and others 3 times in one call: identical records are
    counted once with their multiplicity (bins that are one work item), so large
    multiplicities, both halves of the 16-bit LDS counts and 8-bit saturation are
    checked against the oracle's one-by-one increments."""
    scb = gpu
    k, m, s, h, cse, bse = 31, 21, 25, 4, 20, 9  # 32 bins of ~300 records: each one work item
    g = synth.genome(3000, 9)
    rb, ro = synth.fixed_reads(g, 30, 150, 9)
    parts = [np.tile(rb[ro[0]:ro[1]], 300)] + [np.tile(rb[ro[i]:ro[i + 1]], 3) for i in range(1, 30)]
    lens = [150] * 300 + [150] * (3 * 29)
    bases = np.concatenate(parts)
    offsets = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    assert offsets[-1] == bases.size
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    ok_kmers, _ = oracle.insert_records(of, bases, offsets)
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    rep = b.add_sequences((bases, offsets))
    assert rep.inserted_kmers == ok_kmers
    got = b.counters().astype(np.uint32)
    want = of.counters()
    assert np.array_equal(got, want), f"counters differ at {np.flatnonzero(got != want)[:10]}"
    if bits == 8:
        assert want.max() == 255  # the 300 copies saturate
    idx = b.into_index()
    vals, valid, _ = idx.estimate_sequences(bases, offsets)
    wv, wok = oracle.estimate_records(of, bases, offsets, k)
    assert np.array_equal(valid, wok)
    assert np.array_equal(vals[wok], wv[wok])


def test_c1_small_parity(gpu, oracle):
    """1 Mbp genome, error-free 150 bp reads with random strand (the small CPU-runnable case),
    reduced to 20k reads so the oracle stays fast."""
    scb = gpu
    k, m, s, h, bits, cse, bse = 31, 21, 25, 4, 16, 24, 9
    g = synth.genome(1_000_000, 11)
    bases, offsets = synth.fixed_reads(g, 20_000, 150, 12)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    ok_kmers, ok_smers = oracle.insert_records(of, bases, offsets)
    for method in METHODS:
        b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
        b.set_method(method)
        rep = b.add_sequences((bases, offsets))
        assert (rep.inserted_kmers, rep.inserted_smers) == (ok_kmers, ok_smers)
        assert np.array_equal(b.counters().astype(np.uint32), of.counters())
        idx = b.into_index()
        idx.set_method(method)
        vals, valid, _ = idx.estimate_sequences(bases, offsets)
        wv, wok = oracle.estimate_records(of, bases, offsets, k)
        assert np.array_equal(valid, wok) and np.array_equal(vals[wok], wv[wok])
        # no underestimation: every inserted k-mer has estimate >= 1
        assert (vals[wok] >= 1).all()


@pytest.mark.parametrize("method", METHODS, ids=["batched", "plain"])
def test_reused_table_memory_starts_from_zero(gpu, oracle, method):
    """Counter tables are recycled between builders of the same geometry and are
    not cleared at allocation: the first insert must define every counter, and a
    table that is read before any insert must read as zero."""
    scb = gpu
    k, m, s, h, bits, cse, bse = 31, 21, 25, 4, 16, 20, 9
    g = synth.genome(50_000, 31)
    big_b, big_o = synth.fixed_reads(g, 4000, 150, 32)
    small_b, small_o = synth.fixed_reads(g, 50, 150, 33)
    for _ in range(2):  # leave dirty tables of this size behind
        b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
        b.add_sequences((big_b, big_o))
        assert b.counters().any()
        del b
    # read before any insert
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    assert not b.counters().any()
    idx = b.into_index()
    vals, valid, rep = idx.estimate_sequences(small_b, small_o)
    assert rep.positive_estimates == 0 and not vals[valid].any()
    del idx
    # first insert is a small batch that touches few bins
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.set_method(method)
    b.add_sequences((small_b, small_o))
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    oracle.insert_records(of, small_b, small_o)
    assert np.array_equal(b.counters().astype(np.uint32), of.counters())
    # an insert that produces no k-mers leaves the table pending; the next read clears it
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.add_sequences((np.frombuffer(b"ACGT", dtype=np.uint8), np.array([0, 4], dtype=np.uint64)))
    assert not b.counters().any()


def _parity_roundtrip(scb, oracle, geo, ib, io, qb, qo, method):
    k, m, s, h, bits, cse, bse = geo
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    ok_kmers, ok_smers = oracle.insert_records(of, ib, io)
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.set_method(method)
    rep = b.add_sequences((ib, io))
    assert (rep.inserted_kmers, rep.inserted_smers) == (ok_kmers, ok_smers)
    got = b.counters().astype(np.uint32)
    want = of.counters()
    assert np.array_equal(got, want), f"counters differ at {np.flatnonzero(got != want)[:10]}"
    idx = b.into_index()
    idx.set_method(method)
    vals, valid, qrep = idx.estimate_sequences(qb, qo)
    wv, wok = oracle.estimate_records(of, qb, qo, k)
    assert np.array_equal(valid, wok)
    assert np.array_equal(vals[wok], wv[wok])
    assert qrep.estimate_checksum == int(wv[wok].sum()) & 0xFFFFFFFFFFFFFFFF


@pytest.mark.parametrize("cse", [22, 18], ids=["sparse", "hot_bins"])
@pytest.mark.parametrize("method", METHODS, ids=["batched", "plain"])
def test_long_reads_with_repeats_parity(gpu, oracle, cse, method):
    """Long noisy reads over a repeat-rich genome, 8-bit counters (the long-read
    workload at reduced size): records span many scan tiles, indels shift
    minimizers, the repeat saturates counters, and at cse=18 every bin holds more
    records than one work item."""
    scb = gpu
    print(f"long reads cse={cse} method={method}", flush=True)
    g = synth.insert_repeats(synth.genome(200_000, 51), 300, 200, 52)
    ib, io = synth.variable_reads(g, 1_500_000, 8000, 0.7, 500, 60_000, 53, sub=0.03, ins=0.01, dele=0.01)
    qb, qo = synth.variable_reads(g, 300_000, 8000, 0.7, 500, 60_000, 54, sub=0.03, ins=0.01, dele=0.01)
    _parity_roundtrip(scb, oracle, (31, 21, 25, 4, 8, cse, 9), ib, io, qb, qo, method)


@pytest.mark.parametrize("method", METHODS, ids=["batched", "plain"])
def test_weighted_mixture_u32_parity(gpu, oracle, method):
    """Abundance-weighted mixture of genomes, k21 m15 s17 h3, 32-bit counters,
    1024-slot blocks (the metagenome workload at reduced size)."""
    scb = gpu
    rng = np.random.default_rng(61)
    weights = np.exp(rng.normal(0.0, 2.0, size=6))
    reads_b, reads_o = [], [np.zeros(1, dtype=np.uint64)]
    for i, wgt in enumerate(weights / weights.sum()):
        g = synth.genome(40_000, 62 + i)
        n = max(1, int(wgt * 20_000))
        b, o = synth.fixed_reads(g, n, 150, 70 + i, sub_rate=0.005)
        reads_b.append(b)
        reads_o.append(o[1:] + reads_o[-1][-1])
    ib = np.concatenate(reads_b)
    io = np.concatenate(reads_o).astype(np.uint64)
    _parity_roundtrip(scb, oracle, (21, 15, 17, 3, 32, 22, 10), ib, io, ib, io, method)


@pytest.mark.parametrize("reads,cse", [(4000, 16), (2500, 20)], ids=["split_items", "one_item"])
@pytest.mark.parametrize("bits", [8, 16, 32])
def test_work_items_counted_in_sub_batches(gpu, oracle, bits, reads, cse):
    """Work items that make more than 65,535 counter increments are counted and
    merged in sub-batches (the 16-bit LDS counts cannot overflow), including the
    first insert into a new table and a second insert on top of it. 2,500 reads
    into 2^20 slots put ~1,600 records in each 2^15-slot bin: every bin is one
    work item (at most 2048 records) too large for distinct-record counting in
    one pass. 4,000 reads into 2^16 slots put ~5,200 records in each bin even at
    the smallest insert bin size (2^12 slots): bins split into items that merge
    with atomics."""
    scb = gpu
    k, m, s, h, bse = 31, 21, 25, 4, 9
    g = synth.genome(400_000, 81)
    b1, o1 = synth.fixed_reads(g, reads, 150, 82, sub_rate=0.01)
    b2, o2 = synth.fixed_reads(g, reads, 150, 83, sub_rate=0.01)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    oracle.insert_records(of, b1, o1)
    oracle.insert_records(of, b2, o2)
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.add_sequences((b1, o1))
    b.add_sequences((b2, o2))
    assert np.array_equal(b.counters().astype(np.uint32), of.counters())


@pytest.mark.parametrize("bits", [16, 32])
def test_insert_bins_sized_per_batch(gpu, oracle, bits):
    """A batch whose 2^15-slot insert bins would hold ~2,100 records each (more
    than one 2048-record work item) is inserted with 2^14-slot bins instead; a
    small batch into the same table keeps 2^15-slot bins. Counters after both
    equal the oracle's."""
    scb = gpu
    k, m, s, h, cse, bse = 21, 15, 17, 3, 20, 10
    g = synth.genome(200_000, 91)
    b1, o1 = synth.fixed_reads(g, 2000, 150, 92, sub_rate=0.005)
    b2, o2 = synth.fixed_reads(g, 300, 150, 93, sub_rate=0.005)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    oracle.insert_records(of, b1, o1)
    oracle.insert_records(of, b2, o2)
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.set_method(scb.METHOD_BATCHED)
    r1 = b.add_sequences((b1, o1))
    r2 = b.add_sequences((b2, o2))
    # records (super-k-mers) = (s-mers - k-mers) / (k - s); 32 bins of 2^15 slots
    rec1, rec2 = ((r.inserted_smers - r.inserted_kmers) // (k - s) for r in (r1, r2))
    assert rec1 / 32 > 2048  # the large batch's 2^15-slot bins would split
    assert rec2 / 32 < 0.9 * 2048
    assert np.array_equal(b.counters().astype(np.uint32), of.counters())


@pytest.mark.parametrize("bits", [8, 16])
def test_distinct_records_past_first_walk_round(gpu, oracle, bits):
    """Reads at 0.1x coverage of a 2 Mbp genome, one hash and k - s = 1: each
    2^15-slot insert bin holds ~1,390 records, nearly all distinct, in one work
    item counted by distinct record. The 1,024 threads walk the first 1,024
    leaders with bases kept from the duplicate search and the rest with bases
    read again from HBM; counters equal the oracle's. Abundances and reports of
    the query equal the oracle's when its bins hold more distinct records than
    two walk rounds and when they hold two rounds of leaders with copies."""
    scb = gpu
    k, m, s, h, cse, bse = 21, 15, 20, 1, 20, 10
    g = synth.genome(2_000_000, 94)
    bases, offsets = synth.fixed_reads(g, 1333, 150, 95, sub_rate=0.01)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    ok_kmers, ok_smers = oracle.insert_records(of, bases, offsets)
    b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    b.set_method(scb.METHOD_BATCHED)
    r = b.add_sequences((bases, offsets))
    assert (r.inserted_kmers, r.inserted_smers) == (ok_kmers, ok_smers)
    records = (r.inserted_smers - r.inserted_kmers) // (k - s)
    assert 1100 * 32 < records < 1800 * 32  # per 2^15-slot bin: past one round, within one item
    assert np.array_equal(b.counters().astype(np.uint32), of.counters())
    # queries: the same reads (more distinct records per bin than two rounds of
    # the query's 512 threads: every record is walked) and 40% of them three
    # times (~560 distinct records per bin in two rounds, their copies written
    # from the leaders' estimates)
    idx = b.into_index()
    n40 = (offsets.size - 1) * 2 // 5
    cut = int(offsets[n40])
    tb = np.concatenate([bases[:cut]] * 3)
    to = np.concatenate([offsets[:1]] + [offsets[1:n40 + 1] + c * cut for c in range(3)]).astype(np.uint64)
    for qb, qo in ((bases, offsets), (tb, to)):
        vals, valid, rep = idx.estimate_sequences(qb, qo)
        wv, wok = oracle.estimate_records(of, qb, qo, k)
        assert np.array_equal(valid, wok)
        assert np.array_equal(vals[wok], wv[wok])
        assert rep.valid_kmers == int(wok.sum())
        assert rep.positive_estimates == int((wv[wok] > 0).sum())
        assert rep.estimate_checksum == int(wv[wok].sum()) & 0xFFFFFFFFFFFFFFFF


def _low_complexity_reads(seed):
    """Homopolymers, short tandem repeats (ACGT is its own reverse complement),
    repeats broken by N, by lowercase and by single substitutions, repeats
    glued to random sequence, and reads of length around k and m: every window
    of a repeat has tied m-mer hashes, so the minimizer tie-break and the
    canonical strand choice decide the records."""
    rng = np.random.default_rng(seed)
    units = ["A", "C", "G", "T", "AC", "AT", "CG", "AAC", "ACG", "ACGT", "AAAAC", "CAGCAG", "ACGTTGCA"]
    reads = []
    for u in units:
        for n in (5, 9, 20, 21, 24, 25, 30, 31, 32, 33, 45, 62, 63, 150, 400):
            reads.append((u * (n // len(u) + 1))[:n])
    rand = lambda n: "".join(rng.choice(list("ACGT"), n))
    for u in units:
        r = (u * 200)[:200]
        reads.append(r[:90] + "N" + r[91:])
        reads.append(r[:70] + r[70:80].lower() + r[80:])
        reads.append(r[:100] + ("C" if r[100] != "C" else "G") + r[101:])
        reads.append(rand(60) + r[:120] + rand(40))
        reads.append(r[:50] + rand(33) + r[:50] + rand(7) + r[:90])
    return [np.frombuffer(x.encode(), dtype=np.uint8) for x in reads]


@pytest.mark.parametrize("geo", [(31, 21, 25, 4, 16, 22, 9), (21, 15, 17, 3, 32, 22, 10), (9, 5, 7, 3, 8, 16, 6)],
                         ids=["k31m21s25", "k21m15s17", "k9m5s7"])
@pytest.mark.parametrize("method", METHODS, ids=["batched", "plain"])
def test_low_complexity_reads_parity(gpu, oracle, geo, method):
    """Low-complexity reads, each inserted 3 times, queried forward and reverse
    complemented: counters and abundances equal the oracle's."""
    scb = gpu
    print(f"low complexity geo={geo} method={method}", flush=True)
    reads = _low_complexity_reads(91) * 3
    ib = np.concatenate(reads)
    io = np.concatenate([[0], np.cumsum([r.size for r in reads])]).astype(np.uint64)
    comp = np.frombuffer(bytes.maketrans(b"ACGTacgtN", b"TGCAtgcaN"), dtype=np.uint8)
    rc = [comp[r[::-1]] for r in reads[: len(reads) // 3]]
    qreads = reads[: len(reads) // 3] + rc
    qb = np.concatenate(qreads)
    qo = np.concatenate([[0], np.cumsum([r.size for r in qreads])]).astype(np.uint64)
    _parity_roundtrip(scb, oracle, geo, ib, io, qb, qo, method)
