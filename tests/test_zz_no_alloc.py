"""The allocation-free sampler (qsvh_sample_noalloc) against qsvh_sample: the same outcomes
for the same state and seed (the same kernels on the same data, so bit for bit), on two
calls in a row -- the second on a smaller state, so the scratch holds the first call's
larger tables when it starts -- and the size limits rejected with a clear error.

It runs in a fresh process as well as after the other tests (the scratch is the library's
static device array, looked up when the context is created)."""
import ctypes as C

import numpy as np
import pytest

import rustatevec_amd as rs

pytestmark = pytest.mark.gpu


def _noalloc():
    return rs.lib().qsvh_sample_noalloc


def _hip():
    # the HIP runtime the engine links (torch bundles another one: two runtimes in one
    # process do not share devices), for the caller-owned output buffer
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class DeviceBuffer:
    """Device memory the caller owns (the entry point under test allocates none)."""

    def __init__(self, count):
        self.count = max(count, 1)
        self.ptr = C.c_void_p()
        assert _hip().hipMalloc(C.byref(self.ptr), 8 * self.count) == 0

    def read(self, count):
        out = np.zeros(max(count, 1), dtype=np.uint64)
        assert _hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, 8 * count, 2) == 0
        return out[:count]

    def free(self):
        _hip().hipFree(self.ptr)


def _sample_noalloc(backend, st, shots, seed):
    buf = DeviceBuffer(shots)
    rc = _noalloc()(st.handle, shots, seed, buf.ptr)
    assert rc == 0, rs.lib().qsvh_last_error(backend.ctx)
    assert rs.lib().qsvh_synchronize_ctx(backend.ctx) == 0
    out = buf.read(shots)
    buf.free()
    return out


def test_noalloc_sample_equals_the_allocating_sampler(backend):
    # first call: 20 qubits complex128; second: 9 qubits complex64 (smaller input); third:
    # the 20-qubit state again, after the small call left its tables in the scratch
    cases = ((20, rs.C128, 300000, 17), (9, rs.C64, 50000, 18), (20, rs.C128, 1000, 19))
    states = {}
    for n, dtype, shots, seed in cases:
        key = (n, dtype)
        if key not in states:
            st = backend.alloc(n, dtype)
            backend.init_random_normal(st, 100 + n)
            states[key] = st
        st = states[key]
        want, _ = backend.sample(st, shots, seed)
        got = _sample_noalloc(backend, st, shots, seed)
        print(f"noalloc sample n={n} dtype={dtype} shots={shots}: "
              f"{int(np.count_nonzero(got != want))} differences", flush=True)
        assert got.shape == want.shape and np.array_equal(got, want.astype(np.uint64))
    for st in states.values():
        st.free()


def test_noalloc_sample_limits(backend):
    st = backend.alloc(23, rs.C64)  # one qubit beyond the scratch
    backend.init_basis(st, 0)
    out = DeviceBuffer(4)
    rc = _noalloc()(st.handle, 4, 1, out.ptr)
    assert rc == 1  # QSVH_ERR_INVALID_ARG
    assert b"at most 22 qubits" in rs.lib().qsvh_last_error(backend.ctx)
    st.free()
    st = backend.alloc(10, rs.C64)
    backend.init_basis(st, 5)
    assert _noalloc()(st.handle, (1 << 31) + 1, 1, out.ptr) == 1
    assert _noalloc()(st.handle, 4, 1, None) == 1
    # zero shots: nothing written, no error
    assert _noalloc()(st.handle, 0, 1, None) == 0
    # a basis state samples only its index
    got = _sample_noalloc(backend, st, 4, 3)
    assert got.tolist() == [5, 5, 5, 5]
    out.free()
    st.free()
