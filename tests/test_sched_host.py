"""The greedy fusion cut on the host (ph2_sched_greedy; no GPU call): every
pass holds at least one term and fits its tile, also when the requested
contiguous low bits reach or exceed the tile width."""

import ctypes as C
import math

import numpy as np
import pytest

from phase2_amd import _native as nat, synth


class Sched(C.Structure):
    """struct ph2_sched (include/phase2/qreg_hip.h)."""
    _fields_ = [("npass", C.c_size_t), ("len", C.POINTER(C.c_uint32)),
                ("tile_bits", C.POINTER(C.c_uint32)),
                ("low_bits", C.POINTER(C.c_uint32)), ("cap", C.c_size_t)]


def _rot(xm, zm):
    rot = (nat.Rot * len(xm))()
    for k in range(len(xm)):
        rot[k] = nat.Rot(int(xm[k]), int(zm[k]), math.cos(0.1),
                         math.sin(0.1))
    return rot


def _greedy(rot, n, nqb, tile, low):
    L = nat.lib()
    s = Sched()
    rc = L.ph2_sched_greedy(rot, C.c_size_t(n), C.c_uint32(nqb),
                            C.c_uint32(tile), C.c_uint32(low), C.byref(s))
    out = [(s.len[p], s.tile_bits[p], s.low_bits[p]) for p in range(s.npass)]
    L.ph2_sched_free(C.byref(s))
    return rc, out


@pytest.mark.parametrize("nqb,tile,low", [(20, 12, 12), (20, 12, 13),
                                          (20, 12, 40), (16, 13, 13),
                                          (24, 12, 3), (9, 9, 9)])
def test_greedy_passes_fit(nqb, tile, low):
    hm = synth.hamiltonian(nqb, 300, seed=nqb, decay=True)
    perm = nat.sort_lex(hm.xm, hm.zm)
    xm, zm = hm.xm[perm], hm.zm[perm]
    rot = _rot(xm, zm)
    rc, passes = _greedy(rot, len(xm), nqb, tile, low)
    assert rc == 0
    assert sum(p[0] for p in passes) == len(xm)
    k = 0
    shape = (C.c_uint32 * 4)()
    for ln, tb, lb in passes:
        assert ln >= 1 and tb <= min(tile, nqb) and lb < tb
        sub = (nat.Rot * ln).from_buffer(rot, k * C.sizeof(nat.Rot))
        assert nat.lib().ph2hip_pass_shape(sub, C.c_size_t(ln),
                                           C.c_uint32(tb), C.c_uint32(lb),
                                           shape) == 0
        k += ln


def test_greedy_rejects_zero_tile():
    rot = _rot([1], [0])
    rc, passes = _greedy(rot, 1, 0, 12, 3)
    assert rc == -1 and passes == []
