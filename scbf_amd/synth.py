"""Seeded synthetic read sets for the benchmark workloads (numpy only).

Genomes are i.i.d. uniform ACGT; reads start at uniform positions and are
reverse-complemented with probability 1/2; errors are i.i.d. per base. Genome,
read and error streams use separate seeds, so the CPU baseline and the GPU read
identical bytes.
"""
import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)
_COMP = np.zeros(256, dtype=np.uint8)
for a, b in zip(b"ACGTN", b"TGCAN"):
    _COMP[a] = b


def genome(length, seed):
    rng = np.random.default_rng([seed, 1])
    return ACGT[rng.integers(0, 4, size=length, dtype=np.uint8)]


def insert_repeats(g, repeat_len, copies, seed):
    """Inserts one random element of repeat_len bases `copies` times."""
    rng = np.random.default_rng([seed, 2])
    elem = ACGT[rng.integers(0, 4, size=repeat_len, dtype=np.uint8)]
    pos = np.sort(rng.integers(0, g.size, size=copies))
    parts, prev = [], 0
    for p in pos:
        parts.append(g[prev:p])
        parts.append(elem)
        prev = p
    parts.append(g[prev:])
    return np.concatenate(parts)


def fixed_reads(g, n_reads, read_len, seed, sub_rate=0.0, n_rate=0.0):
    """n_reads reads of read_len bases: returns (bases u8, offsets u64[n+1])."""
    rng = np.random.default_rng([seed, 3])
    starts = rng.integers(0, g.size - read_len + 1, size=n_reads)
    idx = starts[:, None] + np.arange(read_len)[None, :]
    reads = g[idx]
    rc = rng.random(n_reads) < 0.5
    reads[rc] = _COMP[reads[rc][:, ::-1]]
    erng = np.random.default_rng([seed, 4])
    if sub_rate > 0:
        mask = erng.random(reads.shape) < sub_rate
        codes = np.searchsorted(ACGT, reads[mask])
        shift = erng.integers(1, 4, size=codes.size)
        reads[mask] = ACGT[(codes + shift) % 4]
    if n_rate > 0:
        reads[erng.random(reads.shape) < n_rate] = ord("N")
    bases = np.ascontiguousarray(reads.reshape(-1))
    offsets = np.arange(n_reads + 1, dtype=np.uint64) * np.uint64(read_len)
    return bases, offsets


def coverage_reads(g, coverage, read_len, seed, sub_rate=0.0, n_rate=0.0):
    n = int(round(g.size * coverage / read_len))
    return fixed_reads(g, n, read_len, seed, sub_rate, n_rate)


def variable_reads(g, total_bases, median, sigma, lo, hi, seed, sub=0.0, ins=0.0, dele=0.0):
    """Log-normal read lengths with substitutions, insertions and deletions."""
    rng = np.random.default_rng([seed, 5])
    lens = []
    acc = 0
    while acc < total_bases:
        L = int(np.clip(np.exp(rng.normal(np.log(median), sigma)), lo, hi))
        lens.append(L)
        acc += L
    out, offs = [], [0]
    for L in lens:
        s = int(rng.integers(0, g.size - L + 1))
        r = g[s:s + L].copy()
        if rng.random() < 0.5:
            r = _COMP[r[::-1]]
        u = rng.random(L)
        keep = u >= dele
        codes = np.searchsorted(ACGT, r)
        sub_m = (u >= dele) & (u < dele + sub)
        codes[sub_m] = (codes[sub_m] + rng.integers(1, 4, size=int(sub_m.sum()))) % 4
        r = ACGT[codes]
        ins_m = (u >= dele + sub) & (u < dele + sub + ins)
        pieces = []
        for i in np.flatnonzero(ins_m):
            pieces.append(i)
        if pieces:
            extra = ACGT[rng.integers(0, 4, size=len(pieces))]
            r = np.insert(r, pieces, extra)
            keep = np.insert(keep, pieces, True)
        r = r[keep]
        out.append(r)
        offs.append(offs[-1] + r.size)
    return np.concatenate(out), np.array(offs, dtype=np.uint64)


def random_records(n, lo, hi, seed, n_rate=0.0, lower_rate=0.0):
    """Random records of uniform length in [lo, hi], optionally with N and lowercase."""
    rng = np.random.default_rng([seed, 6])
    lens = rng.integers(lo, hi + 1, size=n)
    total = int(lens.sum())
    b = ACGT[rng.integers(0, 4, size=total, dtype=np.uint8)]
    if n_rate > 0:
        b[rng.random(total) < n_rate] = ord("N")
    if lower_rate > 0:
        m = (rng.random(total) < lower_rate) & (b != ord("N"))
        b[m] = b[m] + 32
    offs = np.zeros(n + 1, dtype=np.uint64)
    offs[1:] = np.cumsum(lens)
    return b, offs


def serving_requests(g, count, seed, n_range=(1, 10_000), len_range=(31, 20_000), genome_frac=0.7,
                     max_bases=None):
    """Requests of the served workload: each has n ~ log-uniform(n_range) records
    of length ~ log-uniform(len_range); a record is a genome substring on a random
    strand with probability genome_frac, else uniform random ACGT. Returns a list
    of (bases u8, offsets u64[n+1]); max_bases caps one request's total."""
    rng = np.random.default_rng([seed, 6])
    grc = _COMP[g[::-1]]
    out = []
    for _ in range(count):
        n = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1] + 1))))
        n = max(n_range[0], min(n, n_range[1]))
        lens = np.exp(rng.uniform(np.log(len_range[0]), np.log(len_range[1] + 1), size=n)).astype(np.int64)
        lens = np.clip(lens, len_range[0], min(len_range[1], g.size))
        if max_bases is not None:
            keep = np.cumsum(lens) <= max_bases
            keep[0] = True
            lens = lens[keep]
            n = lens.size
        from_g = rng.random(n) < genome_frac
        strand = rng.random(n) < 0.5
        starts = rng.integers(0, g.size - lens + 1)
        parts = []
        for i in range(n):
            L = int(lens[i])
            if from_g[i]:
                src = grc if strand[i] else g
                parts.append(src[starts[i]:starts[i] + L])
            else:
                parts.append(ACGT[rng.integers(0, 4, size=L, dtype=np.uint8)])
        bases = np.ascontiguousarray(np.concatenate(parts))
        offsets = np.zeros(n + 1, dtype=np.uint64)
        offsets[1:] = np.cumsum(lens)
        out.append((bases, offsets))
    return out


def concat_sets(parts):
    """Concatenates (bases, offsets) read sets into one."""
    bases = np.concatenate([b for b, _ in parts])
    offs = [np.zeros(1, dtype=np.uint64)]
    for _, o in parts:
        offs.append(o[1:].astype(np.uint64) + offs[-1][-1])
    return np.ascontiguousarray(bases), np.concatenate(offs).astype(np.uint64)


def default_geometry_workload(seed=1, insert_coverage=50, fresh_coverage=10):
    """Short-read workload of the default geometry: 5 Mbp genome; insert 150 bp
    reads at 50x (1% substitutions, 0.1% N); query = the inserted reads followed
    by a fresh 10x set. Returns (genome, insert bases, insert offsets, query
    bases, query offsets)."""
    g = genome(5_000_000, seed)
    ib, io = coverage_reads(g, insert_coverage, 150, seed + 1, sub_rate=0.01, n_rate=0.001)
    nb, no = coverage_reads(g, fresh_coverage, 150, seed + 2, sub_rate=0.01, n_rate=0.001)
    qb, qo = concat_sets([(ib, io), (nb, no)])
    return g, ib, io, qb, qo


def repeat_genome(seed=3, length=20_000_000, repeat_len=300, copies=10_000):
    """20 Mbp uniform genome with one 300 bp random element inserted 10,000 times."""
    return insert_repeats(genome(length, seed), repeat_len, copies, seed)


def long_reads(g, coverage, seed):
    """Log-normal read lengths (median 8 kbp, sigma 0.7, clipped to 500 bp-60 kbp)
    at the given coverage with 3% substitutions, 1% insertions, 1% deletions."""
    return variable_reads(g, int(g.size * coverage), 8000, 0.7, 500, 60_000, seed,
                          sub=0.03, ins=0.01, dele=0.01)


def mixture_reads(n_genomes, genome_len, total_bases, seed, read_len=150, sigma=2.0, sub_rate=0.005,
                  chunk_reads=1_000_000):
    """Reads of a mixture of n_genomes random genomes whose relative abundances are
    exp(N(0, sigma^2)); 150 bp reads, total_bases in all, 0.5% substitutions.
    Each genome's reads are drawn in chunks of at most chunk_reads (each chunk
    with its own seed), which bounds the generator's memory."""
    rng = np.random.default_rng([seed, 7])
    w = np.exp(rng.normal(0.0, sigma, size=n_genomes))
    w /= w.sum()
    n_reads = int(total_bases // read_len)
    counts = rng.multinomial(n_reads, w)
    parts = []
    for i, n in enumerate(counts):
        if n == 0:
            continue
        g = genome(genome_len, seed * 1000 + i)
        for c, a in enumerate(range(0, int(n), chunk_reads)):
            m = min(chunk_reads, int(n) - a)
            parts.append(fixed_reads(g, m, read_len, (seed * 1000 + 500 + i) * 4096 + c, sub_rate=sub_rate))
    return concat_sets(parts)
