"""On-device sampling (qsvh_sample; Backend::sample, crates/qsv-core/src/backend/mod.rs:52-54)
against the oracle's cdf_sample (sample.rs:140-153) on the same probabilities and the same
SplitMix64 draws, and the reference's own statistical sampling tests
(crates/qsv-core/tests/sampling.rs).

The device adds the probabilities in a different order than the reference's serial scan,
so an outcome may differ only when the draw u falls within rounding distance of a CDF
boundary; every such difference is checked to be one (relative distance <= 1e-12)."""
import numpy as np
import pytest

import oracle
import rustatevec_amd as rs

pytestmark = pytest.mark.gpu


def _check_against_oracle(backend, st, shots, seed):
    p = backend.probabilities(st).astype(np.float64)
    got, end = backend.sample(st, shots, seed)
    want, want_end = oracle.cdf_sample(p, shots, seed)
    assert end == want_end
    assert got.shape == (shots,) and got.max() < len(p)
    diff = np.nonzero(got != want)[0]
    assert len(diff) <= max(2, shots // 100000), (len(diff), shots)
    if len(diff):
        cdf = np.cumsum(p)
        from tests.golden.make_golden import SplitMix64
        sm = SplitMix64(seed)
        u = np.array([sm.unit() for _ in range(shots)]) * cdf[-1]
        for k in diff:
            lo, hi = sorted((int(got[k]), int(want[k])))
            gap = np.min(np.abs(cdf[lo:hi + 1] - u[k]))
            assert gap <= 1e-12 * cdf[-1], (k, got[k], want[k], gap)
    return got


@pytest.mark.parametrize("dtype", [rs.C128, rs.C64])
def test_sample_matches_oracle_cdf(backend, dtype):
    for n, shots, seed in ((1, 1000, 3), (5, 20000, 4), (11, 50000, 5), (16, 200000, 6),
                           (20, 100000, 7)):
        st = backend.alloc(n, dtype)
        backend.init_random_normal(st, seed)
        _check_against_oracle(backend, st, shots, seed)
        st.free()


def test_sample_sparse_state_never_draws_zero_probability(backend):
    # |psi> supported on 3 of 2^12 basis states, uploaded
    n = 12
    re = np.zeros(1 << n)
    im = np.zeros(1 << n)
    re[5], im[700], re[4095] = 0.6, 0.48, 0.64
    st = backend.alloc(n, rs.C128)
    backend.upload(st, re=re, im=im)
    got = _check_against_oracle(backend, st, 100000, 11)
    assert set(np.unique(got).tolist()) <= {5, 700, 4095}
    st.free()


def test_sample_ghz_only_all_zero_or_all_one(backend):
    # tests/sampling.rs:7-27
    st = backend.execute(rs.ghz(6))
    out, _ = backend.sample(st, 50000, 42)
    assert set(np.unique(out).tolist()) <= {0, 63}
    assert abs(np.mean(out == 0) - 0.5) < 0.02
    st.free()


def test_sample_histogram_matches_probabilities(backend):
    # tests/sampling.rs:29-44
    st = backend.execute(rs.qft(6))
    p = backend.probabilities(st)
    out, _ = backend.sample(st, 400000, 7)
    f = np.bincount(out.astype(np.int64), minlength=64) / 400000
    assert np.all(np.abs(f - p) < 0.01)
    st.free()


def test_sample_large_state_marginals(backend):
    # 24-qubit complex64 random state: the sampled top-4-bit marginal matches the
    # probabilities' (a size-independent property of the full-size sampler)
    n = 24
    st = backend.alloc(n, rs.C64)
    backend.init_random_normal(st, 21)
    p = backend.probabilities(st).astype(np.float64)
    out, _ = backend.sample(st, 1_000_000, 22)
    marg = p.reshape(16, -1).sum(axis=1)
    f = np.bincount((out >> np.uint64(n - 4)).astype(np.int64), minlength=16) / 1e6
    assert np.all(np.abs(f - marg) < 0.003)
    st.free()


def test_sample_zero_shots_and_one_amplitude(backend):
    st = backend.alloc(0, rs.C128)
    backend.init_basis(st, 0)
    out, end = backend.sample(st, 10, 5)
    assert np.all(out == 0)
    out, end = backend.sample(st, 0, 5)
    assert out.shape == (0,) and end == 5
    st.free()
