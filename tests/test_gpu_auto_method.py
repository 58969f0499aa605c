"""GPU tests of the default method choice (SCB_METHOD_AUTO): a batch below
2^17 bases + one per 256 table bytes runs the plain method, a larger one the
batched method, and both give the oracle's counters and estimates."""
import ctypes as C

import numpy as np
import pytest

from hipmem import DeviceBuffer
from scbf_amd import synth

pytestmark = pytest.mark.gpu

GEO = (31, 21, 25, 4, 16, 22, 9)  # 8 MiB table: plain below 2^17 + 2^15 bases


def _stages(scb):
    ms = (C.c_double * 4)()
    scb.api.lib().scb_last_stage_times(ms)
    return list(ms)  # scan, plan + scatter, bins, total


def test_auto_picks_plain_for_small_batches_and_batched_for_large(gpu, oracle):
    scb = gpu
    k, m, s, h, bits, cse, bse = GEO
    cfg = scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                       block_slots_exponent=bse, threads=1, queue=4)
    g = synth.genome(300_000, 21)
    small_b, small_o = synth.fixed_reads(g, 200, 150, 22, sub_rate=0.01, n_rate=0.001)      # 30 kbp
    large_b, large_o = synth.fixed_reads(g, 20_000, 150, 23, sub_rate=0.01, n_rate=0.001)   # 3 Mbp
    of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
    b = scb.SuperCountingBloomBuilder(cfg)  # the default method is auto
    for bb, oo, batched in ((small_b, small_o, False), (large_b, large_o, True)):
        db, do = DeviceBuffer.from_array(bb), DeviceBuffer.from_array(oo)
        b.add_sequences_device(db.ptr, do.ptr, oo.size - 1, int(oo[-1]))
        st = _stages(scb)
        assert (st[2] > 0) == batched, (int(oo[-1]), st)
        of.insert_records(bb, oo)
    assert np.array_equal(b.counters(), of.table())
    idx = b.into_index()
    for bb, oo, batched in ((small_b, small_o, False), (large_b, large_o, True)):
        N = int(oo[-1])
        db, do = DeviceBuffer.from_array(bb), DeviceBuffer.from_array(oo)
        vals, vbits = DeviceBuffer(N * 2 + 4), DeviceBuffer((N + 31) // 32 * 4 + 4)
        idx.estimate_sequences_device(db.ptr, do.ptr, oo.size - 1, N, vals.ptr, vbits.ptr)
        st = _stages(scb)
        assert (st[2] > 0) == batched, (N, st)
        wv, wok = of.estimate_records(bb, oo)
        got = vals.to_array(np.uint16, N)
        okb = np.unpackbits(vbits.to_array(np.uint32, (N + 31) // 32).view(np.uint8), bitorder="little")[:N]
        assert np.array_equal(okb.astype(bool), wok) and np.array_equal(got[wok], wv[wok])


def test_method_setting_is_checked(gpu):
    scb = gpu
    cfg = scb.SuperCountingBloomConfig(k=31, m=21, s=25, n_hashes=4, counter_bits=16, counter_slots_exponent=22,
                                       block_slots_exponent=9, threads=1, queue=4)
    b = scb.SuperCountingBloomBuilder(cfg)
    for m in (scb.METHOD_BATCHED, scb.METHOD_PLAIN, scb.METHOD_AUTO):
        b.set_method(m)
    with pytest.raises(scb.SuperCountingBloomError) as e:
        b.set_method(3)
    assert e.value.kind == "InvalidConfig" and "SCB_METHOD_AUTO" in str(e.value)
