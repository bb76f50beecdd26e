"""ph2hip_gather_noalloc (the gather that allocates no device memory)
against ph2hip_gather, both called through the C ABI on one device
register: two calls in a row, the second with fewer indices than the
first, so that the second result cannot come from what the first left in
the shared device array.  Bit-exact (both copy amplitudes)."""

import ctypes as C

import numpy as np
import pytest

from phase2_amd import _native as nat

pytestmark = pytest.mark.gpu

_u64p, _f64p = C.POINTER(C.c_uint64), C.POINTER(C.c_double)
NQB = 20


@pytest.fixture(scope="module")
def dev_state():
    print("no_alloc: init", flush=True)
    nat.init()
    L = nat.lib()
    for f in ("ph2hip_malloc", "ph2hip_free", "ph2hip_fill_hash",
              "ph2hip_gather", "ph2hip_gather_noalloc"):
        getattr(L, f).restype = C.c_int
    L.ph2hip_strerror.restype = C.c_char_p
    amp = C.c_void_p()
    assert L.ph2hip_malloc(C.byref(amp), C.c_size_t(16 << NQB)) == 0
    assert L.ph2hip_fill_hash(amp, C.c_uint64(1 << NQB),
                              C.c_uint64(77)) == 0
    yield L, amp
    L.ph2hip_free(amp)


def _call(L, fn, amp, idx):
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.full(2 * idx.shape[0], np.nan)
    rc = getattr(L, fn)(amp, C.c_uint32(NQB), idx.ctypes.data_as(_u64p),
                        C.c_size_t(idx.shape[0]), out.ctypes.data_as(_f64p))
    return rc, out


def test_noalloc_gather_equals_gather(dev_state):
    L, amp = dev_state
    rng = np.random.default_rng(5)
    want_all = nat.hash_state(NQB, 77)
    for n in (50000, 777):       # second call smaller than the first
        idx = rng.integers(0, 1 << NQB, size=n, dtype=np.uint64)
        idx[0] = (1 << NQB) - 1
        print(f"no_alloc: gather of {n} indices", flush=True)
        rc_a, ref = _call(L, "ph2hip_gather", amp, idx)
        rc_b, got = _call(L, "ph2hip_gather_noalloc", amp, idx)
        assert rc_a == 0 and rc_b == 0, L.ph2hip_strerror()
        assert np.array_equal(got, ref)
        assert np.array_equal(got.view(np.complex128), want_all[idx])


def test_noalloc_gather_limits(dev_state):
    L, amp = dev_state
    rc, out = _call(L, "ph2hip_gather_noalloc", amp,
                    np.zeros(65537, np.uint64))
    assert rc == -1
    assert b"at most 65536" in L.ph2hip_strerror()
    assert np.all(np.isnan(out))
    rc, _ = _call(L, "ph2hip_gather_noalloc", amp,
                  np.array([1 << NQB], np.uint64))
    assert rc == -1 and b"beyond" in L.ph2hip_strerror()
    # the maximum itself is accepted
    idx = np.arange(65536, dtype=np.uint64) * 16
    rc, got = _call(L, "ph2hip_gather_noalloc", amp, idx)
    assert rc == 0
    assert np.array_equal(got.view(np.complex128),
                          nat.hash_state(NQB, 77)[idx])
