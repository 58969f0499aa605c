"""GPU parity of the reservoir-sampled spectrum (bin/sampled_spectrum.rs:150-308):
the device's one-pass formulation (independent draws per window, the largest
window per slot) against the oracle's sequential reservoir and one-by-one
queries of the sampled k-mers: rows and counts equal."""
import numpy as np
import pytest

from scbf_amd import synth

pytestmark = pytest.mark.gpu

GEOS = [(31, 21, 25, 4, 16, 22, 9), (21, 15, 17, 3, 32, 20, 10), (9, 5, 7, 3, 8, 16, 6)]


def _cfg(scb, k, m, s, h, bits, cse, bse):
    return scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                        block_slots_exponent=bse, threads=2, queue=4)


@pytest.mark.parametrize("geo", GEOS, ids=[f"k{g[0]}u{g[4]}" for g in GEOS])
@pytest.mark.parametrize("size,seed", [(1, 1), (1000, 2), (50_000, 1), (10**7, 3)],
                         ids=["one", "1k", "50k", "all"])
def test_sampled_spectrum_equals_oracle(gpu, oracle, geo, size, seed):
    scb = gpu
    k, m, s, h, bits, cse, bse = geo
    g = synth.genome(100_000, 91)
    rb, ro = synth.fixed_reads(g, 3000, 150, 92, sub_rate=0.01, n_rate=0.002)
    xb, xo = synth.random_records(300, 0, 500, 93, n_rate=0.01, lower_rate=0.03)
    b, o = synth.concat_sets([(rb, ro), (xb, xo)])
    idx_b = scb.SuperCountingBloomBuilder(_cfg(scb, k, m, s, h, bits, cse, bse))
    idx_b.add_sequences((b, o))
    idx = idx_b.into_index()
    got = idx.sampled_spectrum((b, o), size, seed)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    of.insert_records(b, o)
    rows, (seen, sampled, queried, invalid) = of.sampled_spectrum(b, o, size, seed)
    assert (got.total_inserted_kmers, got.sampled_kmers, got.queried_kmers, got.invalid_queries) == \
        (seen, sampled, queried, invalid)
    assert got.rows == rows
    assert sampled == min(size, seen)


def test_sampled_spectrum_from_fastx_over_many_batches(gpu, oracle, tmp_path, monkeypatch):
    """The file is read in 64 KiB batches (SCB_FILE_BATCH_BYTES): windows are
    numbered across batches as one stream, so the result equals the oracle's
    over the whole file; the CSV has the reference's header."""
    scb = gpu
    k, m, s, h, bits, cse, bse = 31, 21, 25, 4, 16, 22, 9
    g = synth.genome(200_000, 95)
    b, o = synth.fixed_reads(g, 20_000, 150, 96, sub_rate=0.01, n_rate=0.001)
    p = tmp_path / "r.fa"
    with open(p, "wb") as f:
        for i in range(o.size - 1):
            f.write(b">r%d\n%s\n" % (i, bytes(b[int(o[i]):int(o[i + 1])])))
    monkeypatch.setenv("SCB_FILE_BATCH_BYTES", str(1 << 16))
    got = scb.sampled_spectrum_from_fastx(p, _cfg(scb, k, m, s, h, bits, cse, bse), 20_000, 17)
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    kmers, smers = of.insert_records(b, o)
    rows, (seen, sampled, queried, invalid) = of.sampled_spectrum(b, o, 20_000, 17)
    assert (got.indexed_kmers, got.inserted_smers) == (kmers, smers)
    assert (got.total_inserted_kmers, got.sampled_kmers) == (seen, sampled)
    assert got.rows == rows
    out = tmp_path / "s.csv"
    got.write(out)
    lines = out.read_text().splitlines()
    assert lines[0] == "estimated_count,sampled_kmers,linear_scaled_windows,estimated_distinct_kmers"
    assert [tuple(int(x) for x in l.split(",")) for l in lines[1:]] == rows
