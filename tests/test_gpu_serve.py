"""GPU tests of the serving loop (include/scb.h, scb_server_*): every served
request must equal the direct index call on the same records, which the parity
tests pin to the oracle; small requests are also checked against the oracle."""
import zlib

import numpy as np
import pytest

from scbf_amd import synth

pytestmark = pytest.mark.gpu

GEO = (31, 21, 25, 4, 16, 20, 9)


def _index(scb, oracle=None):
    k, m, s, h, bits, cse, bse = GEO
    g = synth.genome(60_000, 41)
    ib, io = synth.fixed_reads(g, 3000, 150, 42, sub_rate=0.01, n_rate=0.002)
    cfg = scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                       block_slots_exponent=bse, threads=1, queue=4)
    b = scb.SuperCountingBloomBuilder(cfg)
    b.add_sequences((ib, io))
    of = None
    if oracle is not None:
        of = oracle.Filter(oracle.make_cfg(k, m, s, h, bits, cse, bse))
        oracle.insert_records(of, ib, io)
    return g, b.into_index(), of


def _requests(g, count, seed, max_records=60, max_len=2500):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(0, max_records + 1))
        recs = []
        for _ in range(n):
            L = int(rng.integers(0, max_len))
            if rng.random() < 0.7 and L < g.size:
                p = int(rng.integers(0, g.size - L + 1))
                r = g[p:p + L].copy()
            else:
                r = rng.choice(np.frombuffer(b"ACGTNacgt", dtype=np.uint8), size=L)
            recs.append(r)
        off = np.zeros(n + 1, dtype=np.uint64)
        if n:
            off[1:] = np.cumsum([r.size for r in recs])
        bases = np.concatenate(recs).astype(np.uint8) if n else np.zeros(0, dtype=np.uint8)
        out.append((bases, off))
    return out


def _check_abundances(idx, req, served):
    bases, off = req
    vals, valid, rep = idx.estimate_sequences(bases, off)
    nb = int(off[-1])
    assert served.total_bases == nb
    assert np.array_equal(served.valid, valid[:nb])
    assert np.array_equal(served.values[:nb].astype(np.uint64), vals[:nb])
    assert served.report == rep


@pytest.mark.parametrize("workers,max_batch", [(2, 1 << 27), (1, 1 << 27), (2, 8192)],
                         ids=["two_workers", "one_worker", "small_batches"])
def test_served_abundances_equal_direct_calls(gpu, workers, max_batch):
    scb = gpu
    g, idx, _ = _index(scb)
    reqs = _requests(g, 40, zlib.crc32(f"{workers}{max_batch}".encode()),
                     max_len=2500 if max_batch > 8192 else 3000)
    with scb.SuperCountingBloomServer(idx, scb.ServeConfig(max_batch_bases=max_batch, workers=workers)) as srv:
        tickets = [srv.submit_abundances(b, o) for b, o in reqs]
        got = {}
        while len(got) < len(tickets):
            s = srv.wait_any(5.0)
            assert s is not None, "no request finished within 5 s"
            got[s.ticket] = s
        for t, r in zip(tickets, reqs):
            _check_abundances(idx, r, got[t])
            assert got[t].submit_ns <= got[t].start_ns <= got[t].done_ns


def test_served_matches_oracle_and_coalesces(gpu, oracle):
    scb = gpu
    g, idx, of = _index(scb, oracle)
    reqs = _requests(g, 30, 7, max_records=8, max_len=400)
    with scb.SuperCountingBloomServer(idx, scb.ServeConfig(workers=1)) as srv:
        tickets = [srv.submit_abundances(b, o) for b, o in reqs]
        res = [srv.wait(t) for t in tickets]
    for (b, o), s in zip(reqs, res):
        wv, wok = oracle.estimate_records(of, b, o, GEO[0])
        assert np.array_equal(s.valid, wok)
        assert np.array_equal(s.values[:b.size][wok].astype(np.uint64), wv[wok])
        assert s.report.valid_kmers == int(wok.sum())
        assert s.report.estimate_checksum == int(wv[wok].sum()) & 0xFFFFFFFFFFFFFFFF
    assert sum(s.batch_requests for s in res) >= len(res)


def test_served_spectrum_equals_index_spectrum(gpu):
    scb = gpu
    g, idx, _ = _index(scb)
    reqs = _requests(g, 12, 99, max_records=40, max_len=1500)
    with scb.SuperCountingBloomServer(idx) as srv:
        ts = []
        for i, (b, o) in enumerate(reqs):  # interleave both kinds
            ts.append(("s", srv.submit_spectrum(b, o)))
            ts.append(("a", srv.submit_abundances(b, o)))
        for i, (kind, t) in enumerate(ts):
            s = srv.wait(t)
            b, o = reqs[i // 2]
            if kind == "s":
                rows, q, d = idx.spectrum((b, o))
                assert (s.rows, s.queried, s.distinct) == (rows, q, d)
            else:
                _check_abundances(idx, (b, o), s)


def test_admission_refuses_when_queue_is_full(gpu):
    scb = gpu
    g, idx, _ = _index(scb)
    big_b, big_o = synth.fixed_reads(g, 200_000, 150, 5)   # 30 Mbp keeps the worker busy
    small = _requests(g, 1, 3, max_records=4, max_len=300)[0]
    with scb.SuperCountingBloomServer(idx, scb.ServeConfig(workers=1, max_pending=1)) as srv:
        first = srv.submit_abundances(big_b, big_o)
        accepted, refused = [first], 0
        for _ in range(50):
            try:
                accepted.append(srv.submit_abundances(*small))
            except scb.SuperCountingBloomError as e:
                assert e.kind == "Busy"
                refused += 1
        assert refused > 0
        for t in accepted:
            srv.wait(t)


def test_overlong_record_is_refused(gpu):
    scb = gpu
    g, idx, _ = _index(scb)
    with scb.SuperCountingBloomServer(idx, scb.ServeConfig(max_batch_bases=8192)) as srv:
        rec = g[:9000].copy()
        with pytest.raises(scb.SuperCountingBloomError) as e:
            srv.submit_abundances(rec, np.array([0, rec.size], dtype=np.uint64))
        assert e.value.kind == "InternalState"
        # spectrum requests of any size are served (larger than staging: direct path)
        t = srv.submit_spectrum(rec, np.array([0, rec.size], dtype=np.uint64))
        rows, q, d = idx.spectrum((rec, np.array([0, rec.size], dtype=np.uint64)))
        s = srv.wait(t)
        assert (s.rows, s.queried, s.distinct) == (rows, q, d)


def test_wait_and_wait_any_from_several_threads(gpu):
    """Threads that wait for their own tickets and threads that take any finished
    ticket share one server: every ticket is delivered exactly once with the
    direct call's result. A ticket taken by wait_any before its owner waits is
    refused to the owner (InternalState); a ticket whose owner is blocked in
    wait is never handed to wait_any, and a second wait on it is refused."""
    import threading
    import time
    scb = gpu
    g, idx, _ = _index(scb)
    reqs = _requests(g, 48, 4242, max_records=30, max_len=2000)
    big_b, big_o = synth.fixed_reads(g, 100_000, 150, 6)
    with scb.SuperCountingBloomServer(idx, scb.ServeConfig(workers=2)) as srv:
        big = srv.submit_abundances(big_b, big_o)
        owner = {}

        def own_wait():
            owner["r"] = srv.wait(big)

        th = threading.Thread(target=own_wait)
        th.start()
        time.sleep(0.2)
        with pytest.raises(scb.SuperCountingBloomError) as e:
            srv.wait(big)  # claimed by the owner thread (or already taken by it)
        assert e.value.kind == "InternalState"
        th.join()
        assert owner["r"].ticket == big

        tickets = [srv.submit_abundances(b, o) for b, o in reqs]
        delivered, refused, errors = {}, [], []
        lock = threading.Lock()

        def waiter(ts):
            for t in ts:
                try:
                    r = srv.wait(t)
                    with lock:
                        assert t not in delivered
                        delivered[t] = r
                except scb.SuperCountingBloomError as ex:
                    with lock:
                        if ex.kind == "InternalState":
                            refused.append(t)
                        else:
                            errors.append(ex)

        def any_waiter():
            while True:
                with lock:
                    if len(delivered) == len(tickets):
                        return
                r = srv.wait_any(0.05)
                if r is not None:
                    with lock:
                        if r.ticket in delivered:
                            errors.append(AssertionError(f"ticket {r.ticket} delivered twice"))
                        delivered[r.ticket] = r

        threads = [threading.Thread(target=waiter, args=(tickets[i::4],)) for i in range(2)]
        threads += [threading.Thread(target=any_waiter) for _ in range(2)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=60)
            assert not t.is_alive(), "a waiting thread did not finish"
        assert not errors, errors
        assert sorted(delivered) == sorted(tickets)
        assert all(t in delivered for t in refused)
        for t, r in zip(tickets, reqs):
            _check_abundances(idx, r, delivered[t])


def test_wait_any_while_another_thread_submits(gpu):
    """One thread submits small requests while another takes finished tickets with
    wait_any: a request can finish before its submit call has returned, and its
    result must still be delivered once, equal to the direct call."""
    import threading
    scb = gpu
    g, idx, _ = _index(scb)
    reqs = _requests(g, 300, 777, max_records=3, max_len=300)
    submitted, delivered, errors = {}, {}, []
    done_submitting = threading.Event()

    def collector():
        try:
            while True:
                r = srv.wait_any(0.05)
                if r is not None:
                    assert r.ticket not in delivered, f"ticket {r.ticket} delivered twice"
                    delivered[r.ticket] = r
                elif done_submitting.is_set() and len(delivered) == len(submitted):
                    return
        except Exception as ex:  # reported by the main thread
            errors.append(ex)

    with scb.SuperCountingBloomServer(idx, scb.ServeConfig(workers=4)) as srv:
        th = threading.Thread(target=collector)
        th.start()
        for i, (b, o) in enumerate(reqs):
            submitted[srv.submit_abundances(b, o)] = i
        done_submitting.set()
        th.join(timeout=120)
        assert not th.is_alive(), "the collecting thread did not finish"
    assert not errors, errors
    assert sorted(delivered) == sorted(submitted)
    for t, i in submitted.items():
        _check_abundances(idx, reqs[i], delivered[t])
