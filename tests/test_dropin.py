"""Drop-in check: the reference's own algorithm drivers (circ/trott.c,
trott2.c, qdrift.c, cmpsit.c), compiled unchanged against this library's
include/ and linked to libphase2_hip.so (oracle/Makefile.ref builds
oracle/_ref/libcirc_dropin.so with oracle/dropin.c as the caller), run on
the GPU and reproduce the reference's golden vectors: bit for bit for the
multidet overlaps and states, to 1e-12 for the coefficient-matrix
overlaps (a device tree reduction).  They also equal the library's own
drivers, and each step writer receives every overlap."""

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle as po
from phase2_amd import _native as nat
from phase2_amd import synth

pytestmark = pytest.mark.gpu

DROPIN = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / \
    "libcirc_dropin.so"
_u64p, _f64p = C.POINTER(C.c_uint64), C.POINTER(C.c_double)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


@pytest.fixture(scope="module")
def dl():
    if not DROPIN.is_file():
        pytest.skip("oracle/_ref/libcirc_dropin.so not built (needs the "
                    "reference sources)")
    print("dropin: init", flush=True)
    nat.init()
    lib = C.CDLL(str(DROPIN))
    for f in ("dropin_trott", "dropin_qdrift", "dropin_cmpsit"):
        getattr(lib, f).restype = C.c_int
    yield lib
    nat.set_default(nat.FUSED, 12, 6)


def _terms(hm):
    return (np.ascontiguousarray(hm.xm, dtype=np.uint64),
            np.ascontiguousarray(hm.zm, dtype=np.uint64),
            np.ascontiguousarray(hm.cf, dtype=np.float64))


def _trott(dl, order, hm, prep_kw, delta, steps):
    xm, zm, cf = _terms(hm)
    prep, keep = nat.make_prep(**prep_kw)
    out = np.zeros(steps, np.complex128)
    wrote = np.zeros(steps, np.complex128)
    st = np.zeros(1 << hm.nqb, np.complex128)
    rc = dl.dropin_trott(C.c_int(order), C.c_uint32(hm.nqb),
                         C.c_size_t(len(cf)), _p(xm, _u64p), _p(zm, _u64p),
                         _p(cf, _f64p), C.byref(prep), C.c_double(delta),
                         C.c_size_t(steps), _p(out, _f64p),
                         _p(wrote, _f64p), _p(st, _f64p))
    assert rc == 0
    assert np.array_equal(wrote, out)
    return out, st


def test_reference_drivers_are_the_ones_called(dl):
    """The drivers in the drop-in library are the reference's, not the
    library's functions of the same names."""
    for f in ("trott_simul", "trott2_simul", "qdrift_simul", "cmpsit_simul"):
        a = C.cast(getattr(dl, f), C.c_void_p).value
        b = C.cast(getattr(nat.lib(), f), C.c_void_p).value
        assert a != b, f


def test_reference_trott_c1_golden(dl, golden):
    """circ/trott.c on the full 16-qubit, 537-term, 64-step run."""
    g = golden("trott_n16.npz")
    w = synth.trott_n16()
    nat.set_default(nat.FUSED, 0, 3)
    print("dropin: trott n=16 64 steps", flush=True)
    ov, st = _trott(dl, 1, w["hm"], dict(det_idx=w["det_idx"],
                                         det_cf=w["det_cf"]),
                    w["delta"], w["steps"])
    assert np.array_equal(ov, g["overlaps"])
    assert po.state_hash(st) == str(g["hash"])
    own, _ = nat.trott(1, w["hm"], dict(det_idx=w["det_idx"],
                                        det_cf=w["det_cf"]),
                       w["delta"], w["steps"], want_state=False)
    assert np.array_equal(own, ov)


def test_reference_trott2_coeff_golden(dl, golden):
    g = golden("trott2_coeff_n12.npz")
    hm = synth.hamiltonian(12, 150, seed=3, decay=True)
    prep = dict(C_alpha=g["Ca"], n_sites=12, n_alpha=3, n_beta=0,
                closed_shell=1, tapered=0)
    nat.set_default(nat.FUSED, 10, 4)
    print("dropin: trott2 n=12 coefficient matrix", flush=True)
    ov, st = _trott(dl, 2, hm, prep, 0.05, 6)
    rel = np.max(np.abs(ov - g["overlaps"])) / np.max(np.abs(g["overlaps"]))
    assert rel <= 1e-12
    assert po.state_hash(st) == str(g["hash"])


def test_reference_qdrift_golden(dl, golden):
    g = golden("qdrift_n10.npz")
    hm = synth.hamiltonian(10, 120, seed=5)
    xm, zm, cf = _terms(hm)
    prep, keep = nat.make_prep(det_idx=g["didx"], det_cf=g["dcf"])
    nat.set_default(nat.FUSED, 10, 3)
    print("dropin: qdrift n=10", flush=True)
    out = np.zeros(4, np.complex128)
    wrote = np.zeros(4, np.complex128)
    st = np.zeros(1 << 10, np.complex128)
    rc = dl.dropin_qdrift(C.c_uint32(10), C.c_size_t(len(cf)), _p(xm, _u64p),
                          _p(zm, _u64p), _p(cf, _f64p), C.byref(prep),
                          C.c_size_t(4), C.c_size_t(150), C.c_double(0.05),
                          C.c_uint64(0x1234), _p(out, _f64p),
                          _p(wrote, _f64p), _p(st, _f64p))
    assert rc == 0
    assert np.array_equal(out, g["overlaps"])
    assert np.array_equal(wrote, out)
    assert po.state_hash(st) == str(g["hash"])


def test_reference_cmpsit_golden(dl, golden):
    g = golden("cmpsit_n10.npz")
    hm = synth.hamiltonian(10, 200, seed=6)
    xm, zm, cf = _terms(hm)
    prep, keep = nat.make_prep(det_idx=np.array([0b0000111100], np.uint64),
                               det_cf=np.array([1.0 + 0j]))
    nat.set_default(nat.FUSED, 9, 2)
    print("dropin: cmpsit n=10", flush=True)
    out = np.zeros(3, np.complex128)
    wrote = np.zeros(3, np.complex128)
    st = np.zeros(1 << 10, np.complex128)
    rc = dl.dropin_cmpsit(C.c_uint32(10), C.c_size_t(len(cf)), _p(xm, _u64p),
                          _p(zm, _u64p), _p(cf, _f64p), C.byref(prep),
                          C.c_size_t(3), C.c_size_t(2), C.c_size_t(40),
                          C.c_size_t(25), C.c_double(0.05),
                          C.c_double(0.01), C.c_uint64(99), _p(out, _f64p),
                          _p(wrote, _f64p), _p(st, _f64p))
    assert rc == 0
    assert np.array_equal(out, g["overlaps"])
    assert np.array_equal(wrote, out)
    assert po.state_hash(st) == str(g["hash"])
