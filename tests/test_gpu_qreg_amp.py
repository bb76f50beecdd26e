"""Single-amplitude access of a register (qreg_setamp / qreg_getamp,
phase2/qreg_cuda.c:61-88, whose callers are phase2/circ.c:222,346 and
phase2/phase2_run.c:137,159), called through the C-ABI of
libphase2_hip.so.

Writes are queued on the host and land in one scatter before the device
buffer is next used; reads come from a host copy of a 2^14-amplitude
window.  These tests pin what a caller sees: the same values, in the same
order, as synchronous device copies would give -- the last write to an
index wins, a read after a write returns the write whether or not the
window holds the index, a rotation sees every earlier write and drops the
window, qreg_zero discards queued writes, and more writes than the queue
holds (2^20) land in order.  Rotated states are compared bit for bit
with the oracle."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from phase2_amd import _native as nat

pytestmark = pytest.mark.gpu

_f64p = C.POINTER(C.c_double)


class Paulis(C.Structure):
    _fields_ = [("pak", C.c_uint64 * 2)]


@pytest.fixture(scope="module")
def L():
    nat.init()
    lib = nat.lib()
    lib.qreg_init.argtypes = [C.c_void_p, C.c_uint32]
    lib.qreg_init.restype = C.c_int
    lib.qreg_free.argtypes = [C.c_void_p]
    lib.qreg_zero.argtypes = [C.c_void_p]
    lib.qreg_error.argtypes = [C.c_void_p]
    lib.qreg_error.restype = C.c_int
    lib.qreg_clear_error.argtypes = [C.c_void_p]
    # _Complex double by value travels as two doubles (x86-64 SysV)
    lib.qreg_setamp.argtypes = [C.c_void_p, C.c_uint64, C.c_double,
                                C.c_double]
    lib.qreg_getamp.argtypes = [C.c_void_p, C.c_uint64, _f64p]
    lib.qreg_paulirot.argtypes = [C.c_void_p, Paulis, C.POINTER(Paulis),
                                  _f64p, C.c_size_t]
    return lib


class Reg:
    def __init__(self, L, nqb):
        self.L, self.nqb = L, nqb
        self.buf = C.create_string_buffer(512)
        assert L.qreg_init(self.buf, nqb) == 0
        self.z = np.zeros(2)

    def set(self, i, z):
        self.L.qreg_setamp(self.buf, int(i), float(z.real), float(z.imag))

    def get(self, i):
        self.L.qreg_getamp(self.buf, int(i), self.z.ctypes.data_as(_f64p))
        return complex(self.z[0], self.z[1])

    def state(self):
        return np.array([self.get(i) for i in range(1 << self.nqb)])

    def rot(self, xm, zm, phi):
        codes = (Paulis * len(phi))()
        for k in range(len(phi)):
            codes[k].pak[0], codes[k].pak[1] = int(xm[k]), int(zm[k])
        phi = np.ascontiguousarray(phi, np.float64)
        self.L.qreg_paulirot(self.buf, Paulis(), codes,
                             phi.ctypes.data_as(_f64p), len(phi))

    def err(self):
        return self.L.qreg_error(self.buf)

    def free(self):
        self.L.qreg_free(self.buf)


def _state(rng, nqb):
    n = 1 << nqb
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _terms(rng, nqb, n):
    m = (1 << nqb) - 1
    xm = rng.integers(0, 1 << 63, n, dtype=np.uint64) & np.uint64(m)
    zm = rng.integers(0, 1 << 63, n, dtype=np.uint64) & np.uint64(m)
    return xm, zm, rng.uniform(-1.0, 1.0, n)


def test_shuffled_writes_last_wins_then_rotate(L):
    print("qreg amp: shuffled writes, 16 qubits", flush=True)
    rng = np.random.default_rng(11)
    nqb = 16
    psi = _state(rng, nqb)
    r = Reg(L, nqb)
    try:
        r.L.qreg_zero(r.buf)
        order = rng.permutation(1 << nqb)
        # a stale value first for a tenth of the indices: the later write wins
        for i in order[: len(order) // 10]:
            r.set(i, -7.0 + 3.0j)
        for n, i in enumerate(order):
            r.set(i, psi[i])
            if n % 5000 == 0:  # a read in the middle lands the queue
                assert r.get(i) == psi[i]
                if n + 1 < len(order) // 10:
                    assert r.get(order[n + 1]) == -7.0 + 3.0j
                if n + 1 < len(order):
                    assert r.get(order[-1]) == 0.0
        assert r.err() == 0
        assert np.array_equal(r.state(), psi)
        xm, zm, phi = _terms(rng, nqb, 40)
        r.rot(xm, zm, phi)
        assert r.err() == 0
        want = po.oracle().rotate(psi, xm, zm, phi)
        assert np.array_equal(r.state(), want)
    finally:
        r.free()


def test_window_follows_writes_and_rotations(L):
    print("qreg amp: window coherence, 15 qubits", flush=True)
    rng = np.random.default_rng(12)
    nqb = 15
    psi = _state(rng, nqb)
    r = Reg(L, nqb)
    try:
        for i in range(1 << nqb):
            r.set(i, psi[i])
        cur = psi.copy()
        # inside the window just read, outside it, and across a rotation
        for i, j in ((5, 7), (5, 20000), (30000, 3)):
            assert r.get(i) == cur[i]
            v = complex(rng.standard_normal(), rng.standard_normal())
            r.set(j, v)
            cur[j] = v
            assert r.get(j) == v
            assert r.get(i) == cur[i]
        xm, zm, phi = _terms(rng, nqb, 9)
        r.rot(xm, zm, phi)
        cur = po.oracle().rotate(cur, xm, zm, phi)
        assert r.get(5) == cur[5] and r.get(20000) == cur[20000]
        r.set(6, 1.5 - 0.5j)
        r.rot(xm[:1], zm[:1], phi[:1])
        cur[6] = 1.5 - 0.5j
        cur = po.oracle().rotate(cur, xm[:1], zm[:1], phi[:1])
        assert np.array_equal(r.state(), cur)
        assert r.err() == 0
    finally:
        r.free()


def test_zero_discards_queued_writes(L):
    print("qreg amp: zero after writes", flush=True)
    r = Reg(L, 12)
    try:
        for i in range(0, 4096, 3):
            r.set(i, 2.0 + 1.0j)
        assert r.get(3) == 2.0 + 1.0j
        r.set(4, 1.0)
        r.L.qreg_zero(r.buf)
        assert not np.any(r.state())
        r.set(4095, -1.0j)
        assert r.get(4095) == -1.0j and r.get(4094) == 0.0
        assert r.err() == 0
    finally:
        r.free()


def test_more_writes_than_the_queue_holds(L):
    print("qreg amp: 2^20 + 2^18 writes, 21 qubits", flush=True)
    rng = np.random.default_rng(13)
    nqb = 21
    n = (1 << 20) + (1 << 18)
    idx = rng.permutation(1 << nqb)[:n]
    val = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = Reg(L, nqb)
    try:
        r.L.qreg_zero(r.buf)
        r.get(int(idx[0]))  # a window open while the queue overflows
        for i, v in zip(idx, val):
            r.set(i, v)
        assert r.get(int(idx[0])) == val[0]
        want = np.zeros(1 << nqb, np.complex128)
        want[idx] = val
        assert np.array_equal(r.state(), want)
        assert r.err() == 0
    finally:
        r.free()


def test_index_beyond_register_sets_error(L):
    print("qreg amp: index beyond the register", flush=True)
    r = Reg(L, 10)
    try:
        r.L.qreg_zero(r.buf)
        r.set(1 << 10, 1.0)
        assert r.err() == -1
        r.L.qreg_clear_error(r.buf)
        assert r.get(1 << 10) == 0.0
        assert r.err() == -1
        r.L.qreg_clear_error(r.buf)
        r.set(1023, 1.0)
        assert r.get(1023) == 1.0 and r.err() == 0
    finally:
        r.free()
