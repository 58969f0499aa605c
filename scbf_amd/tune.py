"""Autotuner of the batched pipeline.

For one geometry and a workload of that geometry, searches the engine's
runtime tunables (include/scb.h scb_tuning): the bin size and work-item size of
the insert, of the query and of the spectrum, then whether replaying the
insert and query launch sequences from a HIP graph pays. Each candidate is
timed on HBM-resident inputs with the engine's own HIP-event stage times
(insert and query) or the call's wall time (spectrum), median of `reps` runs,
and every candidate's time is kept in the output. The best setting per
geometry goes into a JSON tuning cache; libscb.so reads `tuning.json` beside
itself (or the file named by SCB_TUNING_CACHE) when an engine is created.

    python -m scbf_amd.tune --workload default --out scbf_amd/tuning.json

Workloads (seeded, synthetic; the benchmark's read models):
  default  k31 m21 s25 h4, 16-bit, 2^30 slots, 2^9-slot blocks: 5 Mbp genome,
           50x insert + (inserted + fresh 10x) query of 150 bp reads
  long     the same k/m/s/h, 8-bit, 2^30 slots: 20 Mbp genome with a 300 bp
           element x10,000, log-normal reads (median 8 kbp) at 5x
  mixture  k21 m15 s17 h3, 32-bit, 2^32 slots, 2^10-slot blocks: 50 genomes of
           2 Mbp, log-normal abundances, one 256 MB batch of 150 bp reads
Every setting gives identical counters and estimates; only the speed differs.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

from scbf_amd import api, synth  # noqa: E402
from scbf_amd.devmem import DeviceBuffer  # noqa: E402

# a candidate replaces the default only when it is at least 2% faster (the
# run-to-run spread of a stage is about 1-3%)
MARGIN = 0.98

GEOMETRIES = {
    "default": (31, 21, 25, 4, 16, 30, 9),
    "long": (31, 21, 25, 4, 8, 30, 9),
    "mixture": (21, 15, 17, 3, 32, 32, 10),
}


def workload(name):
    """(insert bases, offsets, query bases, offsets) of a named workload."""
    if name == "default":
        _, ib, io, qb, qo = synth.default_geometry_workload(1)
        return ib, io, qb, qo
    if name == "long":
        g = synth.repeat_genome()
        ib, io = synth.long_reads(g, 5.0, 31)
        return ib, io, ib, io
    if name == "mixture":
        ib, io = synth.mixture_reads(50, 2_000_000, 256_000_000, 41)
        return ib, io, ib, io
    raise ValueError(f"unknown workload {name}")


def _cfg(geo):
    k, m, s, h, bits, cse, bse = geo
    return api.SuperCountingBloomConfig(k=k, m=m, s=s, n_hashes=h, counter_bits=bits, counter_slots_exponent=cse,
                                        block_slots_exponent=bse, threads=1, queue=4096)


def _stage_total():
    ms = (C.c_double * 4)()
    api.lib().scb_last_stage_times(ms)
    return ms[3]


def candidates(geo):
    k, m, s, h, bits, cse, bse = geo
    W = bits // 8
    ins_logs = [x for x in (12, 13, 14, 15) if max(bse, 1) <= x <= min(15, cse)]
    q_logs = [x for x in (13, 14, 15, 16) if bse <= x <= cse and (1 << x) * W <= (128 << 10)]
    sp_logs = [x for x in (12, 13, 14) if min(bse, 14) <= x <= min(14, cse)] if bse <= 14 else [14]
    return {
        "insert": [(a, b) for a in ins_logs for b in (512, 1024, 1536, 2048)],
        "query": [(a, b) for a in q_logs for b in (1024, 1792, 2560, 3584)],
        "spectrum": [(a, b) for a in sp_logs for b in (1024, 2048, 4096)],
    }


class Bench:
    def __init__(self, geo, ib, io, qb, qo, reps):
        self.geo, self.reps = geo, reps
        self.cfg = _cfg(geo)
        self.dev = {"ib": DeviceBuffer.from_array(ib), "io": DeviceBuffer.from_array(io),
                    "qb": DeviceBuffer.from_array(qb), "qo": DeviceBuffer.from_array(qo)}
        self.n = {"i": (io.size - 1, int(io[-1])), "q": (qo.size - 1, int(qo[-1]))}
        W = geo[4] // 8
        self.vals = DeviceBuffer(int(qo[-1]) * W + 4)
        self.bits = DeviceBuffer((int(qo[-1]) + 31) // 32 * 4 + 4)

    def insert(self, t):
        """Device ms of an insert into a new table (the benchmark's builder per step)."""
        times = []
        for _ in range(self.reps + 1):
            b = api.SuperCountingBloomBuilder(self.cfg)
            b.set_tuning(t)
            b.add_sequences_device(self.dev["ib"].ptr, self.dev["io"].ptr, *self.n["i"])
            times.append(_stage_total())
            del b
        return float(np.median(times[1:]))

    def index(self, t):
        b = api.SuperCountingBloomBuilder(self.cfg)
        b.set_tuning(t)
        b.add_sequences_device(self.dev["ib"].ptr, self.dev["io"].ptr, *self.n["i"])
        return b.into_index()

    def query(self, idx, t):
        idx.set_tuning(t)
        times = []
        for _ in range(self.reps + 1):
            idx.estimate_sequences_device(self.dev["qb"].ptr, self.dev["qo"].ptr, *self.n["q"], self.vals.ptr,
                                          self.bits.ptr)
            times.append(_stage_total())
        return float(np.median(times[1:]))

    def spectrum(self, idx, t):
        idx.set_tuning(t)
        lib = api.lib()
        cap = 1 << 17
        v = np.zeros(cap, dtype=np.uint64)
        c = np.zeros(cap, dtype=np.uint64)
        n, q, d = C.c_size_t(), C.c_uint64(), C.c_uint64()
        u64p = C.POINTER(C.c_uint64)
        times = []
        for _ in range(self.reps + 1):
            t0 = time.perf_counter()
            rc = lib.scb_index_spectrum_device(idx._h, C.c_void_p(self.dev["qb"].ptr), C.c_void_p(self.dev["qo"].ptr),
                                               *self.n["q"], v.ctypes.data_as(u64p), c.ctypes.data_as(u64p), cap,
                                               C.byref(n), C.byref(q), C.byref(d))
            if rc:
                raise RuntimeError(f"spectrum failed: {api.last_error()}")
            times.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(times[1:]))

    def step(self, t):
        """Wall ms of builder + insert + into_index + query (the benchmark's step)."""
        times = []
        for _ in range(self.reps + 2):
            t0 = time.perf_counter()
            idx = self.index(t)
            idx.estimate_sequences_device(self.dev["qb"].ptr, self.dev["qo"].ptr, *self.n["q"], self.vals.ptr,
                                          self.bits.ptr)
            times.append((time.perf_counter() - t0) * 1e3)
            del idx
        return float(np.median(times[2:]))


def _reference(stage_ms, default_ms, bin_log, item):
    """Time a candidate has to beat: the current setting as measured within the
    candidate sweep (same conditions as its rivals) when the sweep holds it,
    else its separate measurement, which runs first and reads slower."""
    return stage_ms.get(f"{bin_log}/{item}", default_ms)


def tune(name, reps=3, log=print):
    geo = GEOMETRIES[name]
    t0 = time.time()
    ib, io, qb, qo = workload(name)
    log(f"{name}: {io.size - 1} insert / {qo.size - 1} query records ({time.time() - t0:.1f} s to generate)")
    bench = Bench(geo, ib, io, qb, qo, reps)
    base = api.SuperCountingBloomBuilder(_cfg(geo))
    best = base.tuning()  # the geometry's defaults
    best.use_graph = 0
    del base
    cand = candidates(geo)
    ms = {"insert": {}, "query": {}, "spectrum": {}}
    default_ms = {"insert": bench.insert(best)}
    for bl, it in cand["insert"]:
        t = api.Tuning(**{**best.__dict__, "insert_bin_log": bl, "insert_item_records": it})
        ms["insert"][f"{bl}/{it}"] = bench.insert(t)
    bl, it = min(cand["insert"], key=lambda c: ms["insert"][f"{c[0]}/{c[1]}"])
    if ms["insert"][f"{bl}/{it}"] < _reference(ms["insert"], default_ms["insert"], best.insert_bin_log,
                                              best.insert_item_records) * MARGIN:
        best.insert_bin_log, best.insert_item_records = bl, it
    bl, it = best.insert_bin_log, best.insert_item_records
    log(f"  insert: bin_log {bl}, item {it}: {ms['insert'].get(f'{bl}/{it}', default_ms['insert']):.3f} ms "
        f"(defaults {default_ms['insert']:.3f} ms)")
    idx = bench.index(best)
    default_ms["query"] = bench.query(idx, best)
    for bl, it in cand["query"]:
        t = api.Tuning(**{**best.__dict__, "query_bin_log": bl, "query_item_records": it})
        ms["query"][f"{bl}/{it}"] = bench.query(idx, t)
    bl, it = min(cand["query"], key=lambda c: ms["query"][f"{c[0]}/{c[1]}"])
    if ms["query"][f"{bl}/{it}"] < _reference(ms["query"], default_ms["query"], best.query_bin_log,
                                             best.query_item_records) * MARGIN:
        best.query_bin_log, best.query_item_records = bl, it
    bl, it = best.query_bin_log, best.query_item_records
    log(f"  query: bin_log {bl}, item {it}: {ms['query'].get(f'{bl}/{it}', default_ms['query']):.3f} ms "
        f"(defaults {default_ms['query']:.3f} ms)")
    default_ms["spectrum"] = bench.spectrum(idx, best)
    for bl, it in cand["spectrum"]:
        t = api.Tuning(**{**best.__dict__, "spectrum_bin_log": bl, "spectrum_item_records": it})
        ms["spectrum"][f"{bl}/{it}"] = bench.spectrum(idx, t)
    bl, it = min(cand["spectrum"], key=lambda c: ms["spectrum"][f"{c[0]}/{c[1]}"])
    if ms["spectrum"][f"{bl}/{it}"] < _reference(ms["spectrum"], default_ms["spectrum"], best.spectrum_bin_log,
                                                best.spectrum_item_records) * MARGIN:
        best.spectrum_bin_log, best.spectrum_item_records = bl, it
    bl, it = best.spectrum_bin_log, best.spectrum_item_records
    log(f"  spectrum: bin_log {bl}, item {it}: {ms['spectrum'].get(f'{bl}/{it}', default_ms['spectrum']):.3f} ms "
        f"(defaults {default_ms['spectrum']:.3f} ms)")
    del idx
    eager = bench.step(api.Tuning(**{**best.__dict__, "use_graph": 0}))
    graph = bench.step(api.Tuning(**{**best.__dict__, "use_graph": 1}))
    best.use_graph = 1 if graph < eager * MARGIN else 0
    log(f"  step: eager {eager:.3f} ms, HIP graph {graph:.3f} ms -> use_graph {best.use_graph}")
    k, m, s, h, bits, cse, bse = geo
    entry = {"workload": name, "k": k, "m": m, "s": s, "n_hashes": h, "counter_bits": bits,
             "counter_slots_exponent": cse, "block_slots_exponent": bse, **best.__dict__,
             "ms": {"defaults": default_ms, "step_eager": eager, "step_graph": graph,
                    "candidates": ms}}
    return entry


def merge(path, entries):
    """Writes entries into the cache at path, replacing those of the same geometry."""
    keys = ("k", "m", "s", "n_hashes", "counter_bits", "counter_slots_exponent", "block_slots_exponent")
    old = []
    if os.path.exists(path):
        with open(path) as f:
            old = json.load(f).get("entries", [])
    new_keys = {tuple(e[x] for x in keys) for e in entries}
    kept = [e for e in old if tuple(e[x] for x in keys) not in new_keys]
    with open(path, "w") as f:
        json.dump({"format": "scb-tuning-1", "entries": kept + entries}, f, indent=1)
        f.write("\n")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", action="append", choices=sorted(GEOMETRIES), required=True)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    # tune from the engine's defaults, not from an existing cache
    os.environ["SCB_TUNING_CACHE"] = os.devnull
    entries = []
    for name in args.workload:
        entries.append(tune(name, args.reps, log=lambda *a: print(*a, flush=True)))
        merge(args.out, entries)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
