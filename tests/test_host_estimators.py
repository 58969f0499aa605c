"""CPU tests of the product's host estimators (libscb.so, no device): the
filter-only spectrum and count hint (api.rs:337-385, 933-1033) and the
abundance summary (api.rs:855-891) equal the oracle's restatement bit for bit on
the same inputs, and agree with the exact rational evaluation of the
deconvolution (a closed form that shares no floating-point code with either)."""
from fractions import Fraction

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

import oracle_py as O
import scbf_amd as scb


def rational_deconvolution(hist, n_hashes, max_abundance):
    """api.rs:958-1011 in exact arithmetic: p_a = count_a / total, n_a = p_a / p_0,
    log-coefficients c_a = (a n_a - sum_{p<a} p c_p n_{a-p}) / a, then
    max(total c_a / n_hashes, 0); None where the reference returns empty."""
    total = sum(c for _, c in hist)
    if max_abundance == 0:
        return []
    if total == 0:
        return [Fraction(0)] * max_abundance
    prob = [Fraction(0)] * (max_abundance + 1)
    for v, c in hist:
        if v <= max_abundance:
            prob[v] += Fraction(c, total)
    if prob[0] == 0:
        return None
    norm = [p / prob[0] for p in prob]
    lc = [Fraction(0)] * (max_abundance + 1)
    for a in range(1, max_abundance + 1):
        c = a * norm[a]
        for q in range(1, a):
            c -= q * lc[q] * norm[a - q]
        lc[a] = c / a
    return [max(total * lc[a] / max(n_hashes, 1), Fraction(0)) for a in range(1, max_abundance + 1)]


histograms = st.lists(st.tuples(st.integers(0, 40), st.integers(1, 10**9)), min_size=1, max_size=12,
                      unique_by=lambda t: t[0]).map(sorted)


@settings(max_examples=300, deadline=None)
@given(hist=histograms, nh=st.integers(1, 6), mx=st.integers(0, 30), hint=st.integers(0, 10**12))
def test_filter_only_estimators_equal_oracle(hist, nh, mx, hint):
    got = scb.filter_only_spectrum_from_histogram(hist, nh, mx, hint)
    want = O.filter_only_spectrum(hist, nh, mx, hint) if mx else []
    assert got == want
    assert scb.filter_only_count_hint_from_histogram(hist, nh, mx) == O.filter_only_count_hint(hist, nh, mx)


def poisson_histogram(lam, n, vmax=14):
    from math import exp, factorial
    hist = [(v, int(round(n * exp(-lam) * lam ** v / factorial(v)))) for v in range(vmax + 1)]
    return [(v, c) for v, c in hist if c > 0]


@settings(max_examples=150, deadline=None)
@given(lam=st.floats(0.02, 3.0), n=st.integers(10**4, 10**12), nh=st.integers(1, 6), mx=st.integers(1, 12))
def test_count_hint_matches_rational_closed_form(lam, n, nh, mx):
    """Counter histograms of the shape a Bloom table has (Poisson slot loads):
    the f64 hint is within one count (and 1e-9 relative) of the exact rational
    evaluation. (The recursion amplifies rounding on histograms far from this
    shape; there only the oracle equality above applies.)"""
    hist = poisson_histogram(lam, n)
    exact = rational_deconvolution(hist, nh, mx)
    got = scb.filter_only_count_hint_from_histogram(hist, nh, mx)
    assert got == O.filter_only_count_hint(hist, nh, mx)
    want = max(1, round(sum(exact)))
    assert abs(got - want) <= max(1, want * 1e-9)


def test_filter_only_rows_on_a_counter_histogram_shape():
    """A histogram shaped like a real 2^30-slot table (mostly zeros, a Poisson
    tail): rows equal the oracle's and lie within one count of the rational
    closed form."""
    rng = np.random.default_rng(5)
    lam = 0.37
    counts = rng.poisson(lam, size=1 << 20)
    vals, cnts = np.unique(counts, return_counts=True)
    hist = [(int(v) , int(c) << 10) for v, c in zip(vals, cnts)]
    for nh in (1, 3, 4):
        hint = scb.filter_only_count_hint_from_histogram(hist, nh, 255)
        assert hint == O.filter_only_count_hint(hist, nh, 255)
        rows = scb.filter_only_spectrum_from_histogram(hist, nh, 255, hint)
        assert rows == O.filter_only_spectrum(hist, nh, 255, hint)
        exact = rational_deconvolution(hist, nh, 12)
        scale = Fraction(hint) / sum(rational_deconvolution(hist, nh, 255))
        for a in range(1, 13):
            assert abs(rows[a - 1][1] - float(exact[a - 1] * scale)) <= 1.0


def test_filter_only_fallback_without_zero_counters():  # api.rs:942-948
    hist = [(1, 2), (2, 6)]
    assert scb.filter_only_spectrum_from_histogram(hist, 3, 3, 80) == [(1, 20), (2, 60), (3, 0)]
    assert scb.filter_only_count_hint_from_histogram(hist, 3, 3) == 1
    assert scb.filter_only_spectrum_from_histogram([], 3, 2, 5) == O.filter_only_spectrum([], 3, 2, 5)


@settings(max_examples=200, deadline=None)
@given(vals=st.lists(st.integers(0, 2**64 - 1), max_size=40), data=st.data())
def test_summary_equals_oracle(vals, data):
    valid = data.draw(st.lists(st.booleans(), min_size=len(vals), max_size=len(vals)))
    total, n, mean, med = scb.summarize_abundances(vals, valid)
    on, omean, omed = O.summarize(vals, valid)
    assert (total, n) == (len(vals), on)
    if n == 0:
        assert mean is None and med is None and np.isnan(omean)
    else:
        assert mean == omean and med == omed


def test_summary_known_answers():  # api.rs:855-891
    assert scb.summarize_abundances([4, 1, 9, 2], [1, 1, 1, 1]) == (4, 4, 4.0, 3.0)
    assert scb.summarize_abundances([4, 1, 9], [1, 0, 1]) == (3, 2, 6.5, 6.5)
    assert scb.summarize_abundances([5], [0]) == (1, 0, None, None)
    assert scb.summarize_abundances([], []) == (0, 0, None, None)
