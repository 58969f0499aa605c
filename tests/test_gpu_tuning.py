"""GPU tests of the tunables (include/scb.h scb_tuning): invalid settings are
refused, non-default settings give the oracle's counters and estimates, HIP
graph replays of the insert and query pipelines equal the eager calls on every
replay, and a tuning cache file sets the tuning of new engines."""
import json

import numpy as np
import pytest

from hipmem import DeviceBuffer
from scbf_amd import synth

pytestmark = pytest.mark.gpu

GEO = (31, 21, 25, 4, 16, 22, 9)


def _cfg(scb, geo=GEO):
    k, m, s, h, bits, cse, bse = geo
    return scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                        block_slots_exponent=bse, threads=1, queue=4)


@pytest.mark.parametrize("field,value", [("insert_bin_log", 8), ("insert_bin_log", 16), ("query_bin_log", 23),
                                         ("spectrum_bin_log", 15), ("insert_item_records", 0),
                                         ("query_item_records", 4000), ("spectrum_item_records", 5000),
                                         ("use_graph", 2)])
def test_invalid_tuning_is_refused(gpu, field, value):
    scb = gpu
    b = scb.SuperCountingBloomBuilder(_cfg(scb))
    t = b.tuning()
    setattr(t, field, value)
    with pytest.raises(scb.SuperCountingBloomError) as e:
        b.set_tuning(t)
    assert e.value.kind == "InvalidConfig" and field in str(e.value)


@pytest.mark.parametrize("t", [dict(insert_bin_log=12, insert_item_records=300, query_bin_log=13,
                                    query_item_records=700, spectrum_bin_log=12, spectrum_item_records=500),
                               dict(insert_bin_log=14, insert_item_records=1024, query_bin_log=16,
                                    query_item_records=1792, spectrum_bin_log=13, spectrum_item_records=2048)],
                         ids=["small_bins", "mixed"])
def test_tuned_pipelines_equal_oracle(gpu, oracle, t):
    scb = gpu
    k, m, s, h, bits, cse, bse = GEO
    g = synth.genome(200_000, 5)
    ib, io = synth.fixed_reads(g, 40_000, 150, 6, sub_rate=0.01, n_rate=0.001)
    qb, qo = synth.fixed_reads(g, 10_000, 150, 7, sub_rate=0.01, n_rate=0.001)
    b = scb.SuperCountingBloomBuilder(_cfg(scb))
    b.set_tuning(scb.Tuning(**t))
    assert b.tuning() == scb.Tuning(**t)
    b.add_sequences((ib, io))
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    of.insert_records(ib, io)
    assert np.array_equal(b.counters(), of.table())
    idx = b.into_index()
    assert idx.tuning() == scb.Tuning(**t)
    vals, valid, _ = idx.estimate_sequences(qb, qo)
    wv, wok = of.estimate_records(qb, qo)
    assert np.array_equal(valid, wok) and np.array_equal(vals[wok], wv[wok])
    rows, q, d = idx.spectrum((qb, qo))
    sp = oracle.Spectrum()
    assert (q, d) == (sp.add_records(of, qb, qo), sp.distinct()) and rows == sp.rows()


def test_graph_replays_equal_eager_calls(gpu):
    """use_graph: the first device-batch call runs eagerly, the second is
    captured, later ones replay. Builders in a row (a new table each time, the
    benchmark's step) and queries interleaved with eager calls and host
    synchronizations: every replay equals the eager result."""
    scb = gpu
    g = synth.genome(300_000, 9)
    ib, io = synth.fixed_reads(g, 30_000, 150, 10, sub_rate=0.01, n_rate=0.001)
    qb, qo = synth.fixed_reads(g, 20_000, 150, 11, sub_rate=0.01, n_rate=0.001)
    dib, dio = DeviceBuffer.from_array(ib), DeviceBuffer.from_array(io)
    dqb, dqo = DeviceBuffer.from_array(qb), DeviceBuffer.from_array(qo)
    N = int(qo[-1])
    vals, bits = DeviceBuffer(N * 2 + 4), DeviceBuffer((N + 31) // 32 * 4 + 4)

    def build(graph):
        b = scb.SuperCountingBloomBuilder(_cfg(scb))
        t = b.tuning()
        t.use_graph = graph
        b.set_tuning(t)
        rep = b.add_sequences_device(dib.ptr, dio.ptr, io.size - 1, int(io[-1]))
        return b, rep

    eager_b, eager_rep = build(0)
    want_counters = eager_b.counters()
    eager_idx = eager_b.into_index()

    def query(idx):
        vals.zero()
        bits.zero()
        rep = idx.estimate_sequences_device(dqb.ptr, dqo.ptr, qo.size - 1, N, vals.ptr, bits.ptr)
        return vals.to_array(np.uint16, N), bits.to_array(np.uint32, (N + 31) // 32), rep

    want = query(eager_idx)
    for i in range(4):
        b, rep = build(1)
        assert (rep.inserted_kmers, rep.inserted_smers) == (eager_rep.inserted_kmers, eager_rep.inserted_smers)
        assert np.array_equal(b.counters(), want_counters), f"insert replay {i} differs"
        del b
    idx_b, _ = build(1)
    idx = idx_b.into_index()
    t = idx.tuning()
    t.use_graph = 1
    idx.set_tuning(t)
    for i in range(5):
        got = query(idx)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), f"query replay {i} differs"
        assert got[2] == want[2]
        eager = query(eager_idx)  # eager work between the replays
        assert np.array_equal(eager[0], want[0])


def test_tuning_cache_sets_new_engines(gpu, tmp_path):
    scb = gpu
    k, m, s, h, bits, cse, bse = GEO
    good = dict(insert_bin_log=13, query_bin_log=14, spectrum_bin_log=12, insert_item_records=999,
                query_item_records=1234, spectrum_item_records=777, use_graph=0)
    entry = dict(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse, block_slots_exponent=bse,
                 **good, ms={"note": {"nested": 1}})
    bad = dict(entry, counter_slots_exponent=20, insert_bin_log=3)  # below block_slots_exponent: ignored
    p = tmp_path / "tuning.json"
    p.write_text(json.dumps({"format": "scb-tuning-1", "entries": [entry, bad]}))
    try:
        scb.tuning_cache_load(p)
        assert scb.SuperCountingBloomBuilder(_cfg(scb)).tuning() == scb.Tuning(**good)
        other = scb.SuperCountingBloomBuilder(_cfg(scb, (k, m, s, h, bits, 20, bse))).tuning()
        assert other.insert_bin_log == 15  # the invalid entry left the defaults
        p.write_text('{"format": "scb-tuning-1"}')
        with pytest.raises(scb.SuperCountingBloomError) as e:
            scb.tuning_cache_load(p)
        assert e.value.kind == "InvalidConfig"
    finally:
        p.write_text('{"format": "scb-tuning-1", "entries": []}')
        scb.tuning_cache_load(p)
