"""The allocation-free query (scb_index_estimate_sequences_device_noalloc),
called through the C ABI, against scb_index_estimate_sequences_device on the
same HBM-resident inputs: equal abundances, validity bits and reports, on two
calls in a row (the second smaller, so stale scratch contents would show),
with no device memory taken by the second call; batches above the stated
maximum are refused with a message."""
import ctypes as C

import numpy as np
import pytest

from hipmem import DeviceBuffer, device_free_bytes
from scbf_amd import synth
from scbf_amd._lib import ScbQueryReport

MAX_BASES = 1 << 22  # SCB_NOALLOC_MAX_BASES

pytestmark = pytest.mark.gpu


def _cfg(scb, bits, cse):
    return scb.SuperCountingBloomConfig(k=31, m=21, s=25, n_hashes=4, counter_bits=bits, counter_slots_exponent=cse,
                                        block_slots_exponent=9, threads=1, queue=4)


def _call(lib, fn, idx, bases, offsets, bits):
    total = int(offsets[-1])
    db, do = DeviceBuffer.from_array(bases), DeviceBuffer.from_array(offsets)
    vals = DeviceBuffer(total * bits // 8 + 4).zero()
    words = (total + 31) // 32
    vbits = DeviceBuffer(words * 4 + 4).zero()
    rep = ScbQueryReport()
    rc = getattr(lib, fn)(idx._h, C.c_void_p(db.ptr), C.c_void_p(do.ptr), offsets.size - 1, total,
                          C.c_void_p(vals.ptr), C.c_void_p(vbits.ptr), C.byref(rep))
    assert rc == 0, f"{fn} failed: {rc}"
    dt = {8: np.uint8, 16: np.uint16, 32: np.uint32}[bits]
    return (vals.to_array(dt, total), vbits.to_array(np.uint32, words),
            (rep.valid_kmers, rep.positive_estimates, rep.estimate_checksum))


@pytest.mark.parametrize("bits,cse,method", [(16, 30, 0), (16, 30, 1), (8, 24, 0), (32, 32, 0)],
                         ids=["u16_2e30_batched", "u16_2e30_plain", "u8_2e24_batched", "u32_2e32_batched"])
def test_noalloc_query_equals_device_query(gpu, bits, cse, method):
    scb = gpu
    lib = scb.api.lib()
    print(f"no-alloc query u{bits} cse={cse} method={method}", flush=True)
    g = synth.genome(1_000_000, 71)
    ib, io = synth.fixed_reads(g, 60_000, 150, 72, sub_rate=0.01, n_rate=0.001)
    b = scb.SuperCountingBloomBuilder(_cfg(scb, bits, cse))
    b.add_sequences((ib, io))
    idx = b.into_index()
    idx.set_method(method)
    big_b, big_o = synth.fixed_reads(g, 27_000, 150, 73, sub_rate=0.01, n_rate=0.001)   # 4.05 Mbp
    small_b, small_o = synth.random_records(500, 0, 3000, 74, n_rate=0.01, lower_rate=0.01)
    assert int(big_o[-1]) <= MAX_BASES and int(small_o[-1]) < int(big_o[-1])
    for bases, offs in ((big_b, big_o), (small_b, small_o)):
        free0 = device_free_bytes()
        got = _call(lib, "scb_index_estimate_sequences_device_noalloc", idx, bases, offs, bits)
        free1 = device_free_bytes()
        want = _call(lib, "scb_index_estimate_sequences_device", idx, bases, offs, bits)
        assert np.array_equal(got[1], want[1]), "validity bits differ"
        assert np.array_equal(got[0], want[0]), "abundances differ"
        assert got[2] == want[2]
        assert got[2][0] > 0
        if offs is small_o:
            assert free1 == free0, "the allocation-free query took device memory"


def test_noalloc_query_refuses_oversized_batch(gpu):
    scb = gpu
    lib = scb.api.lib()
    b = scb.SuperCountingBloomBuilder(_cfg(scb, 16, 24))
    b.add_sequence(b"ACGT" * 40)
    idx = b.into_index()
    n = MAX_BASES + 1
    bases = np.full(n, ord("A"), dtype=np.uint8)
    offs = np.array([0, n], dtype=np.uint64)
    db, do = DeviceBuffer.from_array(bases), DeviceBuffer.from_array(offs)
    rc = lib.scb_index_estimate_sequences_device_noalloc(idx._h, C.c_void_p(db.ptr), C.c_void_p(do.ptr), 1, n,
                                                         None, None, None)
    assert rc == 6  # SCB_ERR_INTERNAL_STATE
    msg = C.create_string_buffer(256)
    lib.scb_last_error(msg, 256)
    assert b"allocation-free" in msg.value
