"""GPU tests of add_sequence one sequence at a time (api.rs:123-138): each call
returns the sequence's k-mers at once, the inserts reach the device in
batches, and whatever reads the builder (its s-mer count, its counters,
into_index, a batch insert) sees every earlier call; counters, counts and
estimates equal the oracle's."""
import numpy as np
import pytest

from scbf_amd import synth

pytestmark = pytest.mark.gpu

GEO = (31, 21, 25, 4, 16, 22, 9)


def _cfg(scb):
    k, m, s, h, bits, cse, bse = GEO
    return scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                        block_slots_exponent=bse, threads=1, queue=4)


def test_single_sequences_equal_oracle(gpu, oracle):
    scb = gpu
    k, m, s, h, bits, cse, bse = GEO
    g = synth.genome(200_000, 41)
    rb, ro = synth.fixed_reads(g, 3000, 150, 42, sub_rate=0.01, n_rate=0.003)
    reads = [bytes(rb[int(ro[i]):int(ro[i + 1])]) for i in range(ro.size - 1)]
    reads += [b"", b"ACGT", b"acgtacgtacgtacgtacgtacgtacgtacgtacgt", b"N" * 40]  # empty, short, lower case, no k-mer
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    b = scb.SuperCountingBloomBuilder(_cfg(scb))
    tk = ts = 0
    for i, r in enumerate(reads):
        got = b.add_sequence(r)
        ok, os_ = of.insert_ascii(r)
        assert got == ok, i
        tk, ts = tk + ok, ts + os_
        if i == 1000:  # the s-mer count and the counters see every call so far
            assert b.inserted_smers() == ts
            assert np.array_equal(b.counters(), of.table())
    assert b.inserted_kmers() == tk
    # a batch insert after pending single sequences
    bb, bo = synth.fixed_reads(g, 500, 150, 43, sub_rate=0.01)
    rep = b.add_sequences((bb, bo))
    ok, os_ = of.insert_records(bb, bo)
    assert (rep.inserted_kmers, rep.inserted_smers) == (ok, os_)
    tk, ts = tk + ok, ts + os_
    for r in reads[:200]:  # pending again when the index is made
        tk += b.add_sequence(r)
        ts += of.insert_ascii(r)[1]
    assert b.inserted_kmers() == tk and b.inserted_smers() == ts
    idx = b.into_index()
    assert np.array_equal(idx.counters(), of.table())
    qb, qo = synth.fixed_reads(g, 400, 150, 44, sub_rate=0.02)
    vals, valid, _ = idx.estimate_sequences(qb, qo)
    wv, wok = of.estimate_records(qb, qo)
    assert np.array_equal(valid, wok) and np.array_equal(vals[wok], wv[wok])


def test_long_sequences_flush_on_their_own(gpu, oracle):
    """More than 64 MiB of single sequences: the pending batch goes to the
    device when it fills, and the counters still equal the oracle's."""
    scb = gpu
    k, m, s, h, bits, cse, bse = GEO
    g = synth.genome(3_000_000, 45)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    b = scb.SuperCountingBloomBuilder(_cfg(scb))
    tk = 0
    rng = np.random.default_rng(46)
    for _ in range(70):  # 70 x 1 MB
        start = int(rng.integers(0, g.size - 1_000_000))
        r = bytes(g[start:start + 1_000_000])
        got = b.add_sequence(r)
        assert got == of.insert_ascii(r)[0]
        tk += got
    assert b.inserted_kmers() == tk
    assert np.array_equal(b.counters(), of.table())
