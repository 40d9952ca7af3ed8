"""Pins of the oracle's samplers (crates/qsv-core/src/sample.rs:33-173) on the CPU.

* prefix sum: the serial scan is numpy's sequential cumsum, bit for bit;
* cdf_sample: the same outcomes as numpy's searchsorted(side='left') on that CDF with the
  draws of an independent SplitMix64 (tests/golden/make_golden.py), and the generator
  advances by one draw per shot;
* alias table: the table reproduces the distribution exactly (prob[i] plus the spill of
  every bucket aliased to i equals n p_i / sum p), and the reference's own sampling tests
  (sample.rs:208-250, tests/sampling.rs) hold for it.
"""
import numpy as np

import oracle
from tests.golden.make_golden import SplitMix64 as PySM


def _dist(n, seed, zeros=0):
    rng = np.random.default_rng(seed)
    p = rng.random(n) ** 3
    if zeros:
        p[rng.choice(n, zeros, replace=False)] = 0.0
    return p


def test_prefix_sum_is_serial_cumsum():
    p = _dist(1 << 12, 1)
    assert np.array_equal(oracle.prefix_sum(p), np.cumsum(p))


def test_cdf_sample_matches_searchsorted():
    for n, seed in ((1, 3), (7, 4), (1 << 10, 5), (1 << 14, 6)):
        p = _dist(n, seed, zeros=n // 4)
        shots = 5000
        out, end = oracle.cdf_sample(p, shots, seed)
        cdf = np.cumsum(p)
        sm = PySM(seed)
        u = np.array([sm.unit() for _ in range(shots)]) * cdf[-1]
        want = np.minimum(np.searchsorted(cdf, u, side="left"), n - 1)
        assert np.array_equal(out.astype(np.int64), want), (n, seed)
        assert end == sm.s
        # zero-probability outcomes are never drawn (u = 0 exactly has probability 2^-53)
        assert np.all(p[out.astype(np.int64)] > 0)


def test_cdf_sample_planted_error_fails():
    # a sampler that returned the last index below u (an off-by-one) must fail the pin
    p = _dist(64, 9)
    out, _ = oracle.cdf_sample(p, 2000, 9)
    cdf = np.cumsum(p)
    sm = PySM(9)
    u = np.array([sm.unit() for _ in range(2000)]) * cdf[-1]
    wrong = np.maximum(np.searchsorted(cdf, u, side="left") - 1, 0)
    assert not np.array_equal(out.astype(np.int64), wrong)


def test_alias_table_reproduces_distribution():
    for n, seed in ((4, 1), (5, 2), (1 << 10, 3), (3000, 4)):
        p = _dist(n, seed, zeros=n // 5)
        prob, alias = oracle.alias_build(p)
        mass = prob.copy()
        np.add.at(mass, alias.astype(np.int64), 1.0 - prob)
        assert np.allclose(mass, n * p / p.sum(), rtol=0, atol=1e-9), (n, seed)
        assert np.all(mass[p == 0] < 1e-12)


def test_alias_histogram_reference_case():
    # sample.rs:208-225: [0.1, 0.6, 0.0, 0.3], seed 0xDEADBEEF, 400k shots, 0.01 tolerance
    w = np.array([0.1, 0.6, 0.0, 0.3])
    out, end = oracle.alias_sample(w, 400_000, 0xDEAD_BEEF)
    counts = np.bincount(out.astype(np.int64), minlength=4)
    assert counts[2] == 0
    assert np.all(np.abs(counts / 400_000 - w) < 0.01)
    sm = PySM(0xDEAD_BEEF)
    for _ in range(2 * 400_000):
        sm.next()
    assert end == sm.s  # two draws per shot


def test_cdf_and_alias_agree_statistically():
    # sample.rs:227-250
    w = np.array([0.25, 0.25, 0.4, 0.1])
    a, _ = oracle.alias_sample(w, 200_000, 1)
    b, _ = oracle.cdf_sample(w, 200_000, 2)
    fa = np.bincount(a.astype(np.int64), minlength=4) / 200_000
    fb = np.bincount(b.astype(np.int64), minlength=4) / 200_000
    assert np.all(np.abs(fa - fb) < 0.01)
