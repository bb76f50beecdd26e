"""The persistent half-tile form with prefetch (ph2hip_set_fused_prefetch):
at tiles 12 and 13, prefetching 8 or 16 amplitudes per coset, with one to
three workgroups per CU (so every workgroup walks several tiles), the state
after a sweep block equals the plain kernel's, every double."""

import ctypes as C

import numpy as np
import pytest

from phase2_amd import _native as nat
from phase2_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    print("prefetch: init", flush=True)
    nat.init()
    nat.lib().ph2hip_set_fused_prefetch.restype = C.c_int
    yield
    for tb in (12, 13):
        nat.lib().ph2hip_set_fused_prefetch(C.c_uint32(tb), C.c_uint32(0),
                                            C.c_uint32(0))


def _set(tb, amps, wpc):
    assert nat.lib().ph2hip_set_fused_prefetch(
        C.c_uint32(tb), C.c_uint32(amps), C.c_uint32(wpc)) == 0


@pytest.mark.parametrize("tb,amps,wpc", [(12, 8, 3), (12, 16, 2),
                                         (13, 8, 1), (13, 16, 1)])
def test_prefetch_equals_plain(tb, amps, wpc):
    nqb = 24
    hm = synth.hamiltonian(nqb, 900, seed=11, decay=True)
    perm = nat.sort_lex(hm.xm, hm.zm)
    xm, zm, phi = hm.xm[perm][:400], hm.zm[perm][:400], 0.05 * hm.cf[perm][:400]
    ref = nat.Register(nqb, nat.PLAIN, 12, 3)
    fus = nat.Register(nqb, nat.FUSED, tb, 3)
    try:
        print(f"prefetch: plain n={nqb}", flush=True)
        ref.fill_hash(24)
        ref.rotate(xm, zm, phi)
        print(f"prefetch: tile {tb} amps {amps} wgs/cu {wpc}", flush=True)
        _set(tb, amps, wpc)
        fus.fill_hash(24)
        nat.prof_enable(True)
        nat.prof_records(reset=True)
        fus.rotate(xm, zm, phi)
        rec = nat.prof_records(reset=True)
        nat.prof_enable(False)
        fused = rec[rec["kind"] == 2]
        assert np.any(fused["direct"] & 4)     # half-tile passes ran
        assert fus.count_unequal(ref) == 0
    finally:
        nat.prof_enable(False)
        _set(tb, 0, 0)
        ref.free()
        fus.free()


def test_prefetch_rejects_bad_settings():
    L = nat.lib()
    assert L.ph2hip_set_fused_prefetch(C.c_uint32(11), C.c_uint32(8),
                                       C.c_uint32(2)) == -1
    assert L.ph2hip_set_fused_prefetch(C.c_uint32(12), C.c_uint32(4),
                                       C.c_uint32(2)) == -1
    assert L.ph2hip_set_fused_prefetch(C.c_uint32(12), C.c_uint32(8),
                                       C.c_uint32(0)) == -1
