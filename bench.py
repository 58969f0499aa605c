"""Benchmark of the SuperCountingBloom hot path on one MI355X.

Workload (default geometry k=31 m=21 s=25 h=4, 16-bit counters, 2^30 slots,
2^9 slots/block): seeded 5 Mbp uniform genome; insert 150 bp reads at 50x
(1,666,667 reads, 1% substitutions, 0.1% N), then query the same reads plus a
fresh 10x set (333,333 reads). One step = new builder (table zeroed in HBM) +
insert of the whole read set + into_index + query of the whole query set with
per-k-mer abundances and validity bits written to HBM. Inputs are resident in
HBM before the timed region; outputs stay in HBM.

Prints one JSON line. `value` is inserted+queried k-mers per second of the
batched pipeline; the plain one-launch-per-batch method is timed on the same
inputs and checked bit-exact against it; `cpu_baseline` is the oracle's
restatement of the reference's multithreaded CPU path (1 producer, bounded
queue, T workers, 1024 mutex shards) on a bounded sample of the same reads
(median of three runs, at T = host threads and at T = 1). `configs` holds the
same measurement for the long-read (8-bit) and skewed-mixture (32-bit, 2^32
slots) workloads, driven in 256 MB device batches; `served` sweeps the serving
loop's request rate up to the point where p99 passes 10 ms.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G k-mers/s insert+query, batched vs plain GPU (bit-exact); served p50/p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gen_workload(seed=1):
    from scbf_amd import synth
    _, ins_b, ins_o, q_b, q_o = synth.default_geometry_workload(seed)
    return ins_b, ins_o, q_b, q_o


def write_fastq(path, bases, offsets):
    """Fixed-length reads as FASTQ records ('@r', the bases, '+', quality 'I' per base)."""
    n = offsets.size - 1
    L = int(offsets[1] - offsets[0])
    if not np.all(np.diff(offsets.astype(np.int64)) == L):
        raise ValueError("write_fastq expects reads of one length")
    rec = np.empty((n, 2 * L + 7), dtype=np.uint8)
    rec[:, 0:3] = np.frombuffer(b"@r\n", np.uint8)
    rec[:, 3:3 + L] = bases.reshape(n, L)
    rec[:, 3 + L:6 + L] = np.frombuffer(b"\n+\n", np.uint8)
    rec[:, 6 + L:6 + 2 * L] = ord("I")
    rec[:, 6 + 2 * L] = ord("\n")
    rec.tofile(path)


def io_bench(scb, cfg, ins_b, ins_o, q_b, q_o, inserted, queried):
    """The file boundaries of the path on the same reads: build from a FASTQ file (parse, copy,
    insert), query a FASTQ file (parse, copy, query, reports), and the index written to and read
    back from a .scb file. Files live in a temporary directory that is removed afterwards."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="scb_io_")
    try:
        fi, fq, fs = (os.path.join(d, x) for x in ("insert.fq", "query.fq", "index.scb"))
        write_fastq(fi, ins_b, ins_o)
        write_fastq(fq, q_b, q_o)
        t0 = time.perf_counter()
        rep = scb.SuperCountingBloomIndex.build_from_fastx(fi, cfg)
        t1 = time.perf_counter()
        qr = rep.index.query_fasta(fq)
        t2 = time.perf_counter()
        rep.index.save(fs)
        t3 = time.perf_counter()
        loaded = scb.SuperCountingBloomIndex.load(fs)
        t4 = time.perf_counter()
        scb_bytes = os.path.getsize(fs)
        del loaded
        return {
            "fastq_build": {"G_kmers_per_s": round(rep.indexed_kmers / (t1 - t0) / 1e9, 3),
                            "file_MB_per_s": round(os.path.getsize(fi) / (t1 - t0) / 1e6, 1),
                            "seconds": round(t1 - t0, 3),
                            "kmers_equal_device_path": rep.indexed_kmers == inserted},
            "fastq_query": {"G_kmers_per_s": round(qr.valid_kmers / (t2 - t1) / 1e9, 3),
                            "file_MB_per_s": round(os.path.getsize(fq) / (t2 - t1) / 1e6, 1),
                            "seconds": round(t2 - t1, 3),
                            "kmers_equal_device_path": qr.valid_kmers == queried},
            "scb_save_GB_per_s": round(scb_bytes / (t3 - t2) / 1e9, 2),
            "scb_load_GB_per_s": round(scb_bytes / (t4 - t3) / 1e9, 2),
            "scb_bytes": scb_bytes,
            "note": "plain FASTQ of the benchmark reads and a .scb file in the temporary directory; "
                    "file reads may be served from the host page cache",
        }
    except Exception as e:  # a full or missing temporary directory is reported, not fatal
        return {"error": f"{type(e).__name__}: {e}"}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def pinned_bench(scb, cfg, arrays, dev, out_vals, out_bits, steps=3):
    """End to end from page-locked host memory: each step copies the insert and
    query reads (ASCII + offsets) host -> HBM, inserts, queries, and copies the
    abundances (counter width) and validity bits HBM -> host. One warm-up step,
    then the mean of `steps`; the copy time is measured on its own as well."""
    from scbf_amd.devmem import _hip, synchronize
    h = _hip()
    pins = {}
    for key, a in arrays.items():
        p = scb.host_array(a.size, a.dtype)
        p[:] = a
        pins[key] = p
    qN = dev["qN"]
    hv, hb = scb.host_array(qN, np.uint16), scb.host_array((qN + 31) // 32, np.uint32)

    def h2d():
        for key in ("ib", "io", "qb", "qo"):
            a = pins[key]
            if h.hipMemcpy(C.c_void_p(dev[key]), C.c_void_p(a.ctypes.data), a.nbytes, 1):
                raise RuntimeError("hipMemcpy H2D failed")

    def d2h():
        for dst, src in ((hv, out_vals), (hb, out_bits)):
            if h.hipMemcpy(C.c_void_p(dst.ctypes.data), C.c_void_p(src), dst.nbytes, 2):
                raise RuntimeError("hipMemcpy D2H failed")

    def step():
        h2d()
        ir, qr, _, _, idx = run_step(scb, cfg, scb.METHOD_BATCHED, dev, out_vals, out_bits)
        d2h()
        del idx
        return ir.inserted_kmers + qr.valid_kmers

    step()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        kmers = step()
    synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    t1 = time.perf_counter()
    for _ in range(steps):
        h2d()
        d2h()
    copy_ms = (time.perf_counter() - t1) * 1e3 / steps
    moved = sum(pins[k].nbytes for k in pins) + hv.nbytes + hb.nbytes
    return {"value": round(kmers / (ms / 1e3) / 1e9, 3), "unit": "G k-mers/s", "ms_per_step": round(ms, 3),
            "copy_ms": round(copy_ms, 3), "copy_GB_per_s": round(moved / (copy_ms / 1e3) / 1e9, 2),
            "bytes_copied": int(moved),
            "note": "inputs in page-locked host memory copied to HBM and abundances (u16) plus validity "
                    "bits copied back every step; this value includes the copies, the headline value "
                    "(inputs resident in HBM) does not"}


def spectrum_bench(scb, idx, dev, reps=3):
    """Distinct-k-mer spectrum of the query reads (resident in HBM) against the built index:
    scb_index_spectrum_device, one warm-up call, then the mean of `reps` calls."""
    lib = scb.api.lib()
    cap = 1 << 17
    vals = np.zeros(cap, dtype=np.uint64)
    cnts = np.zeros(cap, dtype=np.uint64)
    n, q, d = C.c_size_t(), C.c_uint64(), C.c_uint64()
    u64p = C.POINTER(C.c_uint64)

    def once():
        rc = lib.scb_index_spectrum_device(idx._h, C.c_void_p(dev["qb"]), C.c_void_p(dev["qo"]), dev["qn"],
                                           dev["qN"], vals.ctypes.data_as(u64p), cnts.ctypes.data_as(u64p),
                                           cap, C.byref(n), C.byref(q), C.byref(d))
        if rc:
            raise RuntimeError(f"scb_index_spectrum_device failed ({rc})")

    once()
    t0 = time.perf_counter()
    for _ in range(reps):
        once()
    t = (time.perf_counter() - t0) / reps
    st = stage_times(scb)
    return {"G_kmers_per_s": round(q.value / t / 1e9, 3), "ms": round(t * 1e3, 3),
            "queried_kmers": q.value, "distinct_keys": d.value, "rows": n.value,
            "stage_ms": {"scan": st[0], "plan_scatter": st[1], "bin": st[2], "total": st[3]}}


def cfg_c2(scb):
    return scb.SuperCountingBloomConfig(k=31, m=21, s=25, n_hashes=4, counter_bits=16,
                                        counter_slots_exponent=30, block_slots_exponent=9,
                                        threads=max(1, os.cpu_count() or 1), queue=4096)


def stage_times(scb):
    ms = (C.c_double * 4)()
    scb.api.lib().scb_last_stage_times(ms)
    return list(ms)


def run_step(scb, cfg, method, dev, out_vals, out_bits):
    """One insert + query pass; returns (ins_report, q_report, ins_stage, q_stage, index)."""
    b = scb.SuperCountingBloomBuilder(cfg)
    b.set_method(method)
    ir = b.add_sequences_device(dev["ib"], dev["io"], dev["in"], dev["iN"])
    ist = stage_times(scb)
    idx = b.into_index()
    idx.set_method(method)
    qr = idx.estimate_sequences_device(dev["qb"], dev["qo"], dev["qn"], dev["qN"], out_vals, out_bits)
    qst = stage_times(scb)
    return ir, qr, ist, qst, idx


class BatchedSet:
    """A read set resident in HBM as device batches of at most `limit` bases
    (split at record boundaries), each with its own offsets array, and output
    slices (abundances, validity bits) per batch."""

    def __init__(self, bases, offsets, limit, W):
        from scbf_amd.devmem import DeviceBuffer
        self.bases = DeviceBuffer.from_array(bases)
        self.batches = []
        r0, n = 0, offsets.size - 1
        out_bytes = bit_words = 0
        while r0 < n:
            r1 = int(np.searchsorted(offsets, offsets[r0] + limit, side="right")) - 1
            r1 = max(r0 + 1, min(n, r1))
            nb = int(offsets[r1] - offsets[r0])
            offs = DeviceBuffer.from_array((offsets[r0:r1 + 1] - offsets[r0]).astype(np.uint64))
            self.batches.append((int(offsets[r0]), r1 - r0, nb, offs, out_bytes, bit_words))
            out_bytes += (nb * W + 255) // 256 * 256
            bit_words += (nb + 31) // 32 + 64
            r0 = r1
        self.values = DeviceBuffer(out_bytes + 4)
        self.bits = DeviceBuffer(bit_words * 4 + 4)

    def insert(self, builder):
        ik = ismers = 0
        st = np.zeros(4)
        for o, n, nb, offs, _, _ in self.batches:
            r = builder.add_sequences_device(self.bases.ptr + o, offs.ptr, n, nb)
            ik += r.inserted_kmers
            ismers += r.inserted_smers
            st += np.array(stage_times_module())
        return ik, ismers, st

    def query(self, idx):
        q = pos = cs = 0
        st = np.zeros(4)
        for o, n, nb, offs, vo, bw in self.batches:
            r = idx.estimate_sequences_device(self.bases.ptr + o, offs.ptr, n, nb, self.values.ptr + vo,
                                              self.bits.ptr + 4 * bw)
            q += r.valid_kmers
            pos += r.positive_estimates
            cs = (cs + r.estimate_checksum) & 0xFFFFFFFFFFFFFFFF
            st += np.array(stage_times_module())
        return (q, pos, cs), st


def stage_times_module():
    import scbf_amd as scb
    return stage_times(scb)


def config_bench(scb, name, geo, bases, offsets, steps, limit=256_000_000, workload=""):
    """Throughput of one workload driven in device batches of `limit` bases:
    builder + insert of every batch + into_index + query of every batch (outputs
    in HBM), batched and plain methods; batched == plain is checked on the
    query reports and the counter histograms."""
    k, m, s, h, bits, cse, bse = geo
    cfg = scb.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                       block_slots_exponent=bse, threads=max(1, os.cpu_count() or 1), queue=4096)
    ds = BatchedSet(bases, offsets, limit, bits // 8)

    def step(method):
        from scbf_amd.devmem import synchronize
        synchronize()
        t0 = time.perf_counter()
        b = scb.SuperCountingBloomBuilder(cfg)
        b.set_method(method)
        ik, ismers, ist = ds.insert(b)
        idx = b.into_index()
        idx.set_method(method)
        rep, qst = ds.query(idx)
        synchronize()
        return (time.perf_counter() - t0) * 1e3, ik, ismers, rep, ist, qst, idx

    step(scb.METHOD_BATCHED)  # warm-up
    runs = [step(scb.METHOD_BATCHED) for _ in range(steps)]
    ms = float(np.median([r[0] for r in runs]))
    _, ik, ismers, rep, ist, qst, idx = runs[-1]
    kmers = ik + rep[0]
    hist_b = idx.counter_histogram()
    del runs, idx
    pms, pik, pismers, prep, _, _, pidx = step(scb.METHOD_PLAIN)
    exact = bool((pik, pismers, prep) == (ik, ismers, rep) and pidx.counter_histogram() == hist_b)
    del pidx
    out = {"workload": workload, "value": round(kmers / ms / 1e6, 3), "unit": "G k-mers/s",
           "ms_per_step": round(ms, 3), "steps": steps, "inserted_kmers": ik, "queried_kmers": rep[0],
           "records": int(offsets.size - 1), "bases": int(offsets[-1]), "device_batches": len(ds.batches),
           "batch_bases": limit,
           "stage_ms": {"insert": {"scan": round(float(ist[0]), 3), "plan_scatter": round(float(ist[1]), 3),
                                   "bin": round(float(ist[2]), 3), "total": round(float(ist[3]), 3)},
                        "query": {"scan": round(float(qst[0]), 3), "plan_scatter": round(float(qst[1]), 3),
                                  "bin": round(float(qst[2]), 3), "total": round(float(qst[3]), 3)}},
           "plain": {"value": round(kmers / pms / 1e6, 3), "ms_per_step": round(pms, 3)},
           "batched_over_plain": round(pms / ms, 2), "bit_exact_batched_vs_plain": exact}
    log(f"{name}: {out}")
    return out


def lines_touched(counters, line_bytes=64):
    per = line_bytes // counters.itemsize
    return int(np.count_nonzero(counters.reshape(-1, per).any(axis=1)))


def cpu_baseline(ins_b, ins_o, q_b, q_o, target_s=6.0, runs=3):
    """The oracle's restatement of the reference's threaded path, at T = the host's
    threads (at most 16) and at T = 1: a sample of the workload sized to about
    target_s of CPU time, then the median of `runs` runs of that sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py

    def sample(n_ins, n_q):
        ib, io = ins_b[:int(ins_o[n_ins])], ins_o[:n_ins + 1]
        # query sample: the same inserted reads + an equal share of fresh reads
        fresh0 = ins_o.size - 1
        qi = q_o[fresh0:fresh0 + n_q + 1]
        qb = np.concatenate([ib, q_b[int(qi[0]):int(qi[-1])]])
        qo = np.concatenate([io, qi[1:] - qi[0] + io[-1]]).astype(np.uint64)
        return ib, io, qb, qo

    def measure(threads):
        cfg = oracle_py.make_cfg(31, 21, 25, 4, 16, 30, 9, threads=threads, queue=4096)
        n = 5_000
        while True:
            ib, io, qb, qo = sample(n, n // 5)
            r = oracle_py.cpu_baseline(cfg, ib, io, qb, qo)
            t = r.build_s + r.query_s
            if t >= target_s / 2 or n >= ins_o.size - 1:
                break
            n = min(ins_o.size - 1, int(n * max(2.0, target_s / max(t, 1e-3))))
        rates, last = [], r
        for _ in range(runs):
            last = oracle_py.cpu_baseline(cfg, ib, io, qb, qo)
            rates.append((last.indexed_kmers + last.queried_kmers) / (last.build_s + last.query_s))
        return {"value": float(np.median(rates)) / 1e9, "runs": [round(x / 1e9, 5) for x in rates],
                "cores": threads,
                "sample": f"{n} inserted + {n + n // 5} queried 150 bp reads of the same workload "
                          f"({last.indexed_kmers} + {last.queried_kmers} k-mers; build {last.build_s:.2f} s, "
                          f"query {last.query_s:.2f} s), default geometry, 2^30 16-bit counters"}

    threads = max(1, min(16, os.cpu_count() or 1))
    multi = measure(threads)
    single = measure(1)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": multi["value"], "unit": "G k-mers/s", "cores": threads, "kind": "port",
            "sample": multi["sample"], "runs": multi["runs"], "statistic": f"median of {runs} runs",
            "single_thread": single, "host_cpus": os.cpu_count(), "cpu_model": model}


PMC_PROFILE = os.path.join(ROOT, "profiles", "pmc_default_geometry.md")


def pmc_figures():
    """Per-launch HBM traffic and VALU figures of the kernels, read from the JSON block of the
    committed rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE scaled by a copy kernel of known
    size). Counters cannot be read live without the profiler, so these come from that file."""
    try:
        with open(PMC_PROFILE) as f:
            text = f.read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block)
    except (OSError, IndexError, ValueError):
        return None


def copy_probe(mib=1024, reps=3):
    """Device-to-device copy of a known size: the calibration kernel for FETCH_SIZE/WRITE_SIZE."""
    import torch
    n = mib << 20
    src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(7)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n, "reps": reps}), flush=True)


def serve_bench(scb, idx, genome, qps_steps, min_requests=2000, pool_size=160, p99_limit_ms=10.0, seed=9, cfg=None):
    """Open-loop serving of the default-geometry index: Poisson arrivals at each
    request rate of the sweep, at least min_requests per rate; request sizes
    log-uniform (1-10,000 records of 31-20,000 bp; 70% genome substrings on a
    random strand, 30% random); 80% abundance and 20% spectrum requests. Once per
    rate, half way through, the worst-case request of 10,000 records of 20 kbp is
    submitted in place of a pool request (its own latency is reported; it counts in the percentiles).
    The sweep stops after the first rate whose p99 passes p99_limit_ms. Latency
    is scheduled arrival -> completion (CLOCK_MONOTONIC), so submit lateness
    counts."""
    import threading
    from scbf_amd import synth
    # the server runs the index's method: the default one, which takes the plain
    # path for coalesced batches below the table-dependent size and the batched
    # path above
    idx.set_method(scb.METHOD_AUTO)
    pool = synth.serving_requests(genome, pool_size, seed)
    worst = synth.serving_requests(genome, 1, seed + 1, n_range=(10_000, 10_000), len_range=(20_000, 20_000),
                                   genome_frac=1.0)[0]
    pool.append(worst)
    wi = len(pool) - 1
    pinned = []
    for b, o in pool:  # request bytes live in page-locked memory like a serving front end's
        hb = scb.host_array(b.size, np.uint8)
        hb[:] = b
        pinned.append((hb, o))
    W = np.dtype({8: np.uint8, 16: np.uint16, 32: np.uint32}[idx.config().counter_bits])
    outs = [(scb.host_array(max(1, int(o[-1])), W), scb.host_array(max(1, (int(o[-1]) + 31) // 32), np.uint32))
            for _, o in pool]
    busy = [False] * len(pool)
    rng = np.random.default_rng([seed, 1])
    steps = []
    with scb.SuperCountingBloomServer(idx, cfg) as srv:
        # warm-up: one request of each kind
        srv.wait(srv.submit_abundances(*pinned[0], *outs[0]))
        srv.wait(srv.submit_spectrum(*pinned[0]))
        for qps in qps_steps:
            arrivals = {}
            lat, kmers_done, refused = [], [0], [0]
            worst_ms = [None]
            lock = threading.Lock()
            n_sub = [0]
            stop = [False]

            failed = []

            def collect():
                try:
                    while True:
                        r = srv.wait_any(0.05)
                        if r is None:
                            with lock:
                                if stop[0] and len(lat) + refused[0] >= n_sub[0]:
                                    return
                            continue
                        with lock:
                            t_arr, pi, kind, owns = arrivals.pop(r.ticket)
                            lat.append((r.done_ns - t_arr) / 1e6)
                            if pi == wi and kind == 0:
                                worst_ms[0] = lat[-1]
                            kmers_done[0] += r.report.valid_kmers if kind == 0 else r.queried
                            if owns:
                                busy[pi] = False
                except BaseException as ex:  # re-raised by the submitting thread
                    failed.append(ex)

            th = threading.Thread(target=collect)
            th.start()
            # at least one second of arrivals and at least min_requests requests
            t0 = time.monotonic_ns()
            t_next = t0
            end = t0 + int(1e9)
            worst_sent = False
            for i in range(1 << 40):
                t_next += int(rng.exponential(1e9 / qps))
                if i >= min_requests and t_next >= end:
                    break
                now = time.monotonic_ns()
                if t_next > now:
                    time.sleep((t_next - now) / 1e9)
                if not worst_sent and i >= min_requests // 2:
                    pi, kind, worst_sent = wi, 0, True
                else:
                    pi = int(rng.integers(0, len(pool) - 1))
                    kind = 0 if rng.random() < 0.8 else 1
                owns = False
                try:
                    with lock:
                        if kind == 0:
                            if busy[pi]:
                                nb = int(pool[pi][1][-1])
                                vo, bo = np.empty(max(1, nb), W), np.empty(max(1, (nb + 31) // 32), np.uint32)
                            else:
                                vo, bo = outs[pi]
                                busy[pi] = owns = True
                            t = srv.submit_abundances(*pinned[pi], vo, bo)
                        else:
                            t = srv.submit_spectrum(*pinned[pi])
                        arrivals[t] = (t_next, pi, kind, owns)
                        n_sub[0] += 1
                except scb.SuperCountingBloomError:
                    with lock:
                        refused[0] += 1
                        n_sub[0] += 1
                        if owns:
                            busy[pi] = False
            with lock:
                stop[0] = True
            th.join()
            if failed:
                raise RuntimeError(f"serving collector failed at {qps} requests/s") from failed[0]
            wall = (time.monotonic_ns() - t0) / 1e9
            a = np.array(lat)
            steps.append({"qps": qps, "requests": int(a.size), "refused": refused[0],
                          "p50_ms": round(float(np.percentile(a, 50)), 3) if a.size else None,
                          "p99_ms": round(float(np.percentile(a, 99)), 3) if a.size else None,
                          "max_ms": round(float(a.max()), 3) if a.size else None,
                          "worst_case_request_ms": round(worst_ms[0], 3) if worst_ms[0] is not None else None,
                          "served_G_kmers_per_s": round(kmers_done[0] / wall / 1e9, 4)})
            log(f"served qps={qps}: {steps[-1]}")
            if steps[-1]["p99_ms"] is None or steps[-1]["p99_ms"] > p99_limit_ms:
                break
        # spot checks: served abundances equal the direct call, worst case included
        ok = True
        for pi in (1, wi):
            b, o = pool[pi]
            srv_r = srv.wait(srv.submit_abundances(b, o))
            vals, valid, rep = idx.estimate_sequences(b, o)
            ok &= bool(np.array_equal(srv_r.valid, valid)
                       and np.array_equal(srv_r.values[:b.size].astype(np.uint64), vals) and srv_r.report == rep)
    knee = None
    for st in steps:
        if st["p99_ms"] is not None and st["p99_ms"] <= p99_limit_ms:
            knee = st["qps"]
    sizes = [int(o[-1]) for _, o in pool[:-1]]
    return {"steps": steps, "min_requests_per_step": min_requests, "request_pool": len(pool),
            "mean_request_bases": round(float(np.mean(sizes))), "max_pool_request_bases": int(max(sizes)),
            "worst_case_request": {"records": int(worst[1].size - 1), "bases": int(worst[1][-1])},
            "mix": "80% abundances, 20% spectrum", "method": "auto", "p99_limit_ms": p99_limit_ms,
            "highest_qps_within_limit": knee, "served_equals_direct": ok}


def valu_roofline(pmc, stage_ms, kmers):
    """VALU lane-operations per k-mer (SQ_INSTS_VALU x 64 lanes, from the committed
    rocprofv3 summary) against the MI355X VALU peak: 256 CUs x 4 SIMDs x 16 lanes
    x 2.4 GHz = 39.3 T lane-ops/s, for the dominant kernel and the whole step."""
    peak = 256 * 4 * 16 * 2.4e9
    ks = pmc.get("kernels", {})
    by = pmc.get("kernels_by_stage", {})
    per_launch = {st: ks[kn]["valu_insts"] * 64 for st, kn in by.items() if kn in ks}
    # steps in the profiled run: one insert-bin launch per step; every kernel's
    # per-dispatch mean times its dispatches per step
    steps = ks.get(by.get("insert_bin", ""), {}).get("dispatches", 0)
    total = sum(v["valu_insts"] * 64 * v["dispatches"] / steps for v in ks.values()) if steps else None
    dom = max(stage_ms, key=lambda k: stage_ms[k])
    out = {"peak_T_lane_ops_s": round(peak / 1e12, 1), "source": os.path.relpath(PMC_PROFILE, ROOT)}
    if dom in per_launch:
        out["kernel"] = dom
        out["kernel_lane_ops_per_launch"] = per_launch[dom]
        out["kernel_achieved_T_lane_ops_s"] = round(per_launch[dom] / (stage_ms[dom] / 1e3) / 1e12, 2)
        out["kernel_frac"] = round(per_launch[dom] / (stage_ms[dom] / 1e3) / peak, 3)
    if total:
        step_ms = sum(stage_ms.values())
        out["pipeline_lane_ops_per_kmer"] = round(total / kmers, 1)
        out["pipeline_frac"] = round(total / (step_ms / 1e3) / peak, 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-io", action="store_true", help="skip the FASTQ and .scb file measurements")
    ap.add_argument("--no-spectrum", action="store_true", help="skip the distinct-k-mer spectrum timing")
    ap.add_argument("--no-configs", action="store_true", help="skip the long-read and mixture workloads")
    ap.add_argument("--serve-qps", default="100,200,400,800,1600,3200")
    ap.add_argument("--serve-min-requests", type=int, default=2000)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return

    import scbf_amd as scb
    from scbf_amd import synth
    from scbf_amd.devmem import DeviceBuffer, synchronize

    t0 = time.time()
    ins_b, ins_o, q_b, q_o = gen_workload()
    log(f"workload: {ins_o.size - 1} insert reads, {q_o.size - 1} query reads "
        f"({time.time() - t0:.1f} s to generate)")
    buf = {"ib": DeviceBuffer.from_array(ins_b), "io": DeviceBuffer.from_array(ins_o),
           "qb": DeviceBuffer.from_array(q_b), "qo": DeviceBuffer.from_array(q_o)}
    dev = {"ib": buf["ib"].ptr, "io": buf["io"].ptr, "in": ins_o.size - 1, "iN": int(ins_o[-1]),
           "qb": buf["qb"].ptr, "qo": buf["qo"].ptr, "qn": q_o.size - 1, "qN": int(q_o[-1])}
    qN = dev["qN"]
    out_vals = DeviceBuffer(qN * 2 + 4).zero()
    out_bits = DeviceBuffer((qN + 31) // 32 * 4 + 4).zero()
    synchronize()
    cfg = cfg_c2(scb)

    for _ in range(args.warmup):
        run_step(scb, cfg, scb.METHOD_BATCHED, dev, out_vals.ptr, out_bits.ptr)
    synchronize()
    ist_all, qst_all = [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ir, qr, ist, qst, idx = run_step(scb, cfg, scb.METHOD_BATCHED, dev, out_vals.ptr, out_bits.ptr)
        ist_all.append(ist)
        qst_all.append(qst)
        del idx
    synchronize()
    t_total = time.perf_counter() - t_start
    ms_per_step = t_total * 1e3 / args.steps
    kmers = ir.inserted_kmers + qr.valid_kmers
    value = kmers / (ms_per_step / 1e3) / 1e9
    ist_m = np.median(np.array(ist_all), axis=0).tolist()
    qst_m = np.median(np.array(qst_all), axis=0).tolist()
    log(f"batched: {ms_per_step:.2f} ms/step, {value:.2f} G k-mers/s; insert stages {ist_m}, "
        f"query stages {qst_m}")

    # keep one batched result for the checks
    ir_b, qr_b, _, _, idx_b = run_step(scb, cfg, scb.METHOD_BATCHED, dev, out_vals.ptr, out_bits.ptr)
    counters_b = idx_b.counters()
    vals_b = out_vals.to_array(np.uint16, qN)
    bits_b = out_bits.to_array(np.uint32, (qN + 31) // 32)
    served = None
    if not args.no_serve:
        try:
            served = serve_bench(scb, idx_b, synth.genome(5_000_000, 1),
                                 [float(q) for q in args.serve_qps.split(",")], args.serve_min_requests)
        except RuntimeError as e:  # reported in the line instead of numbers from a broken sweep
            served = {"error": f"{e}: {e.__cause__!r}", "steps": []}
    spectrum = None
    if not args.no_spectrum:
        spectrum = spectrum_bench(scb, idx_b, dev)
        log(f"spectrum: {spectrum}")
    del idx_b
    pinned = None
    if not args.no_host:
        pinned = pinned_bench(scb, cfg, {"ib": ins_b, "io": ins_o, "qb": q_b, "qo": q_o}, dev, out_vals.ptr,
                              out_bits.ptr)
        log(f"pinned end to end: {pinned}")

    plain = None
    bit_exact = None
    if not args.no_plain:
        out_vals.zero()
        out_bits.zero()
        synchronize()
        tp = time.perf_counter()
        ir_p, qr_p, ist_p, qst_p, idx_p = run_step(scb, cfg, scb.METHOD_PLAIN, dev, out_vals.ptr, out_bits.ptr)
        synchronize()
        plain_ms = (time.perf_counter() - tp) * 1e3
        plain = {"value": kmers / (plain_ms / 1e3) / 1e9, "ms_per_step": plain_ms,
                 "insert_ms": ist_p[3], "query_ms": qst_p[3]}
        log(f"plain: {plain_ms:.2f} ms/step, {plain['value']:.2f} G k-mers/s")
        if not args.no_check:
            counters_p = idx_p.counters()
            bit_exact = bool(np.array_equal(counters_b, counters_p)
                             and np.array_equal(vals_b, out_vals.to_array(np.uint16, qN))
                             and np.array_equal(bits_b, out_bits.to_array(np.uint32, (qN + 31) // 32))
                             and (ir_b.inserted_kmers, ir_b.inserted_smers, qr_b.valid_kmers,
                                  qr_b.positive_estimates, qr_b.estimate_checksum)
                             == (ir_p.inserted_kmers, ir_p.inserted_smers, qr_p.valid_kmers,
                                 qr_p.positive_estimates, qr_p.estimate_checksum))
            log(f"batched vs plain bit-exact: {bit_exact}")
        del idx_p

    # roofline: algorithmic bytes, read-modify-write model of the compulsory traffic:
    # insert = bases + 2*64 B per distinct counter line touched; query = bases + 64 B per
    # distinct line read + W bytes per abundance + 1 bit of validity per k-mer.
    W = 2
    L_touch = lines_touched(counters_b)
    n_sk_i = (ir.inserted_smers - ir.inserted_kmers) // (31 - 25)
    iN, qNb = dev["iN"], dev["qN"]
    b_insert_pipe = iN + 128 * L_touch
    b_query_pipe = qNb + 64 * L_touch + qr.valid_kmers * W + qr.valid_kmers / 8
    pipe_ms = ist_m[3] + qst_m[3]
    stages = {"insert_scan": (ist_m[0], iN * 1.25 + 8 * n_sk_i),
              "insert_bin": (ist_m[2], 128 * L_touch + 8 * n_sk_i + iN / 4),
              "query_scan": (qst_m[0], qNb * 1.25 + qNb / 8 + 8 * (qr.valid_kmers / 7.2)),
              "query_bin": (qst_m[2], 64 * L_touch + qr.valid_kmers * W + 8 * (qr.valid_kmers / 7.2) + qNb / 4)}
    dom = max(stages, key=lambda s: stages[s][0])
    dom_ms, dom_bytes = stages[dom]
    ach = dom_bytes / (dom_ms / 1e3) / 1e9
    roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
            "kernel_ms": round(dom_ms, 3), "kernel_bytes": int(dom_bytes),
            "pipeline_bytes": int(b_insert_pipe + b_query_pipe),
            "pipeline_device_ms": round(pipe_ms, 3),
            "pipeline_achieved_GBs": round((b_insert_pipe + b_query_pipe) / (pipe_ms / 1e3) / 1e9, 1),
            "pipeline_frac": round((b_insert_pipe + b_query_pipe) / (pipe_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_per_kmer": round((b_insert_pipe + b_query_pipe) / kmers, 2),
            "lines_touched": L_touch}
    pmc = pmc_figures()
    if pmc and dom in pmc.get("kernels_by_stage", {}):
        kp = pmc["kernels"][pmc["kernels_by_stage"][dom]]
        roof["traffic"] = int(kp["hbm_bytes"])
        roof["traffic_source"] = os.path.relpath(PMC_PROFILE, ROOT)
        roof["valu_busy"] = kp.get("valu_busy")
        if "write_kib" in kp:
            roof["kernel_write_bytes"] = int(kp["write_kib"] * 1024)
            roof["kernel_output_bytes"] = int(qr.valid_kmers * W) if dom == "query_bin" else None
    if pmc:
        roof["valu"] = valu_roofline(pmc, {k: v[0] for k, v in stages.items()}, kmers)

    # host-buffer path (PCIe-inclusive): the crate-style calls on numpy arrays in host memory,
    # abundances returned as u64 + validity bytes per position
    host_api = None
    if not args.no_host:
        synchronize()
        th0 = time.perf_counter()
        hb = scb.SuperCountingBloomBuilder(cfg)
        hrep = hb.add_sequences((ins_b, ins_o))
        hidx = hb.into_index()
        th1 = time.perf_counter()
        _, _, hq = hidx.estimate_sequences(q_b, q_o)
        th2 = time.perf_counter()
        host_api = {"value": round((hrep.inserted_kmers + hq.valid_kmers) / (th2 - th0) / 1e9, 3),
                    "unit": "G k-mers/s", "insert_s": round(th1 - th0, 4), "query_s": round(th2 - th1, 4),
                    "note": "scb_builder_add_sequences + scb_index_estimate_sequences on host arrays "
                            "(PCIe transfers and u64/u8 result expansion included)"}
        log(f"host-buffer path: {host_api}")
        del hidx

    io = None
    if not args.no_io:
        io = io_bench(scb, cfg, ins_b, ins_o, q_b, q_o, ir.inserted_kmers, qr.valid_kmers)
        log(f"file boundaries: {io}")
    del buf, out_vals, out_bits

    configs = None
    if not args.no_configs:
        configs = {}
        t1 = time.time()
        g3 = synth.repeat_genome()
        b3, o3 = synth.long_reads(g3, 20.0, 31)
        log(f"long-read workload: {o3.size - 1} reads, {int(o3[-1])} bases ({time.time() - t1:.1f} s)")
        configs["long_reads_u8"] = config_bench(
            scb, "long reads", (31, 21, 25, 4, 8, 30, 9), b3, o3, 3,
            workload="k31 m21 s25 h4, 8-bit counters, 2^30 slots, 2^9 slots/block; 20 Mbp genome + 300 bp "
                     "element x10,000; log-normal reads (median 8 kbp, sigma 0.7, 500 bp-60 kbp) at 20x, 3% subst, "
                     "1% ins, 1% del; insert and query the same reads in 256 MB device batches")
        del b3, o3, g3
        t1 = time.time()
        b4, o4 = synth.mixture_reads(50, 2_000_000, 2_000_000_000, 41)
        log(f"mixture workload: {o4.size - 1} reads, {int(o4[-1])} bases ({time.time() - t1:.1f} s)")
        configs["skewed_mixture_u32"] = config_bench(
            scb, "skewed mixture", (21, 15, 17, 3, 32, 32, 10), b4, o4, 2,
            workload="k21 m15 s17 h3, 32-bit counters, 2^32 slots (16 GiB), 2^10 slots/block; 50 genomes of 2 Mbp, "
                     "log-normal (sigma 2) abundances; 150 bp reads, 2e9 bases, 0.5% subst; insert and query the "
                     "same reads in 256 MB device batches")
        configs["skewed_mixture_u32_one_batch"] = config_bench(
            scb, "skewed mixture, one batch", (21, 15, 17, 3, 32, 32, 10), b4, o4, 2, limit=1 << 31,
            workload="the skewed mixture above (2e9 bases) inserted and queried as one device batch each")
        del b4, o4

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ins_b, ins_o, q_b, q_o)
        log(f"cpu baseline: {cpu}")

    tun = scb.SuperCountingBloomBuilder(cfg).tuning()
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "G k-mers/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u16",
        "data": "synthetic (seeded uniform genome, simulated reads)",
        "config": {"workload": "default geometry k31 m21 s25 h4, 16-bit counters, 2^30 slots, "
                               "2^9 slots/block; 5 Mbp genome; insert 50x 150 bp reads (1% subst, "
                               "0.1% N), query those + fresh 10x", "parallelism": "single GPU",
                   "insert_reads": dev["in"], "query_reads": dev["qn"], "inserted_kmers": ir.inserted_kmers,
                   "queried_kmers": qr.valid_kmers, "tuning": tun.__dict__},
        "stage_ms": {"insert": {"scan": ist_m[0], "plan_scatter": ist_m[1], "bin": ist_m[2], "total": ist_m[3]},
                     "query": {"scan": qst_m[0], "plan_scatter": qst_m[1], "bin": qst_m[2], "total": qst_m[3]}},
        "plain": plain,
        "batched_over_plain": round(value / plain["value"], 2) if plain else None,
        "bit_exact_batched_vs_plain": bit_exact,
        "roofline": roof,
        "configs": configs,
        "served": served,
        "host_buffers": host_api,
        "pinned_end_to_end": pinned,
        "files": io,
        "spectrum": spectrum,
        "served_p50_ms": served["steps"][0]["p50_ms"] if served and served["steps"] else None,
        "served_p99_ms": served["steps"][0]["p99_ms"] if served and served["steps"] else None,
        "cpu_baseline": cpu,
        "host": platform.node(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
