"""Pins of the oracle's restatement of the reservoir-sampled spectrum
(bin/sampled_spectrum.rs:150-308), which the GPU's parallel formulation is
tested against:
* its SplitMix64 stream is the published one (splitmix64.c, seed 1234567);
* the sequential reservoir equals the closed form the GPU uses (window w > cap
  draws output number w - cap of the stream; slot s holds the largest window
  that targets it), computed here with Python integers;
* every window is retained with probability cap / seen (Algorithm R), checked
  over many seeds;
* with a reservoir at least as large as the input, the sampled spectrum is the
  histogram of every window's estimate, unscaled."""
import numpy as np
import pytest

import oracle_py as O
from scbf_amd import synth

M64 = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def splitmix_outputs(seed, count):
    out, s = [], seed
    for _ in range(count):
        s = (s + GAMMA) & M64
        z = ((s ^ (s >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out.append(z ^ (z >> 31))
    return out


def test_splitmix_stream_is_the_published_one():
    # Vigna's splitmix64.c, seeded with 1234567: the first five outputs
    assert splitmix_outputs(1234567, 5) == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                            4593380528125082431, 16408922859458223821]
    # the oracle's xorshift_u64 is that output function: draw t = xorshift(seed + (t - 1) * gamma)
    assert [O.xorshift_u64((1234567 + t * GAMMA) & M64) for t in range(5)] == splitmix_outputs(1234567, 5)


def _records(seed):
    b, o = synth.random_records(60, 0, 120, seed, n_rate=0.03, lower_rate=0.05)
    return b, o


@pytest.mark.parametrize("cap,seed", [(1, 1), (7, 2), (50, 99), (400, 5), (5000, 7), (64, 2**64 - 1)])
def test_reservoir_equals_independent_draw_closed_form(cap, seed):
    k = 11
    b, o = _records(seed & 0xFFFF)
    win, seen = O.reservoir_windows(b, o, k, cap, seed)
    assert seen > 0
    draws = splitmix_outputs(seed, max(0, seen - cap))
    want = list(range(1, min(seen, cap) + 1))
    for w in range(cap + 1, seen + 1):
        r = draws[w - cap - 1] % w
        if r < cap:
            want[r] = w  # later windows override: the largest w wins
    assert win.tolist() == want


def test_reservoir_retains_each_window_with_probability_cap_over_seen():
    k, cap = 11, 12
    b, o = synth.random_records(3, 40, 40, 3)   # 3 records of 40 bases: 90 windows
    hits = None
    trials = 3000
    for seed in range(trials):
        win, seen = O.reservoir_windows(b, o, k, cap, seed * 7919 + 1)
        if hits is None:
            hits = np.zeros(seen + 1)
        hits[win] += 1
    p = cap / seen
    expect = trials * p
    sd = np.sqrt(trials * p * (1 - p))
    dev = np.abs(hits[1:] - expect) / sd
    assert dev.max() < 5.0, dev.max()
    # chi-square over all windows: mean squared z close to 1
    assert 0.7 < float(np.mean(dev ** 2)) < 1.3


@pytest.mark.parametrize("bits", [8, 16, 32])
def test_full_reservoir_is_the_window_histogram(bits):
    k, m, s, h, cse, bse = 21, 15, 17, 3, 16, 8
    g = synth.genome(5000, 4)
    b, o = synth.fixed_reads(g, 200, 100, 5, sub_rate=0.02, n_rate=0.01)
    f = O.Filter(O.make_cfg(k, m, s, h, bits, cse, bse))
    f.insert_records(b, o)
    rows, (seen, sampled, queried, invalid) = f.sampled_spectrum(b, o, 10**6, 3)
    v, ok = f.estimate_records(b, o)
    assert seen == sampled == queried == int(ok.sum()) and invalid == 0
    u, c = np.unique(v[ok], return_counts=True)
    assert rows == [(int(a), int(n), int(n), 0 if a == 0 else int(np.floor(n / a + 0.5))) for a, n in zip(u, c)]
