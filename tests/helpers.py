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


def qft_closed_form(psi):
    """QFT of the reference's generator (circuits.rs:61-74) in closed form: H on qubit 0
    first and the final swaps make QFT|x> = 2^(-n/2) sum_y exp(2 pi i rev(x) rev(y) / 2^n) |y>
    with rev the n-bit reversal, i.e. numpy's inverse FFT between two bit reversals."""
    psi = np.asarray(psi, dtype=np.complex128)
    n = int(psi.size).bit_length() - 1
    idx = np.arange(psi.size)
    rev = np.zeros_like(idx)
    for b in range(n):
        rev |= ((idx >> b) & 1) << (n - 1 - b)
    return np.fft.ifft(psi[rev])[rev] * np.sqrt(psi.size)
