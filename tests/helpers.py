"""Helpers shared by the tests."""
import numpy as np


def max_abs_diff(a, b):
    """crates/qsv-core/tests/equivalence.rs:11-16: max_i |a_i - b_i| (complex modulus)."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    assert a.shape == b.shape
    return float(np.abs(a - b).max()) if a.size else 0.0


def records_of(case):
    """Gate records (name, qubits, params) of a golden case."""
    return [(k, tuple(q), tuple(p)) for k, q, p in case["ops"]]
