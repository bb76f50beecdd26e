"""Graph replay of repeated fused sweeps on narrow registers
(ph2hip_set_replay): the same block applied again and again leaves the
same state as direct launches (bit for bit), replays only from the third
identical call, and any change of register, angle or schedule is a new
call.  The reference-written C1 golden (64 Trotter steps, every step the
same block) is reproduced with replay on."""

import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po
from phase2_amd import _native as nat
from phase2_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    print("replay: init", flush=True)
    nat.init()
    L = nat.lib()
    L.ph2hip_set_replay.restype = C.c_int
    L.ph2hip_replay_hits.restype = C.c_uint64
    yield
    L.ph2hip_set_replay(C.c_uint32(20))


def _hits():
    return int(nat.lib().ph2hip_replay_hits())


def _block(nqb, nterms, seed):
    hm = synth.hamiltonian(nqb, nterms, seed=seed, decay=True)
    perm = nat.sort_lex(hm.xm, hm.zm)
    return hm.xm[perm], hm.zm[perm], 0.07 * hm.cf[perm]


def _run(max_qb, blocks, nqb=14):
    L = nat.lib()
    assert L.ph2hip_set_replay(C.c_uint32(max_qb)) == 0
    reg = nat.Register(nqb, nat.FUSED, 11, 3)
    try:
        reg.fill_hash(5)
        h0 = _hits()
        for b in blocks:
            reg.rotate(*b)
        return reg.read(), _hits() - h0
    finally:
        reg.free()


def test_replay_same_state_and_counts():
    b1 = _block(14, 300, 1)
    b2 = _block(14, 280, 2)
    seq = [b1, b1, b1, b2, b1, b2, b2, b1]
    print("replay: direct", flush=True)
    direct, h_off = _run(0, seq)
    print("replay: replayed", flush=True)
    replay, h_on = _run(20, seq)
    assert h_off == 0
    # b1: calls 1 (remember), 2 (capture), 3, 5, 8 replayed; b2: 4, 6
    # (capture), 7 replayed
    assert h_on == 4
    assert np.array_equal(direct, replay)


def test_replay_keys_on_angles_and_register():
    L = nat.lib()
    assert L.ph2hip_set_replay(C.c_uint32(20)) == 0
    xm, zm, phi = _block(13, 200, 3)
    a = nat.Register(13, nat.FUSED, 11, 3)
    b = nat.Register(13, nat.FUSED, 11, 3)
    try:
        a.fill_hash(1)
        b.fill_hash(1)
        h0 = _hits()
        for _ in range(3):
            a.rotate(xm, zm, phi)
        assert _hits() - h0 == 1
        b.rotate(xm, zm, phi)                # other register: new call
        phi2 = phi.copy()
        phi2[17] = np.nextafter(phi2[17], 1.0)
        a.rotate(xm, zm, phi2)               # one angle one ulp off
        assert _hits() - h0 == 1
        ref = nat.Register(13, nat.PLAIN, 11, 3)
        try:
            ref.fill_hash(1)
            for _ in range(3):
                ref.rotate(xm, zm, phi)
            ref.rotate(xm, zm, phi2)
            assert a.count_unequal(ref) == 0
        finally:
            ref.free()
    finally:
        a.free()
        b.free()


def test_trott_c1_golden_with_replay(golden):
    g = golden("trott_n16.npz")
    w = synth.trott_n16()
    L = nat.lib()
    assert L.ph2hip_set_replay(C.c_uint32(20)) == 0
    nat.set_default(nat.FUSED, 0, 3)
    h0 = _hits()
    ov, st = nat.trott(1, w["hm"], dict(det_idx=w["det_idx"],
                                        det_cf=w["det_cf"]),
                       w["delta"], w["steps"])
    assert _hits() - h0 >= w["steps"] - 2
    assert np.array_equal(ov, g["overlaps"])
    assert po.state_hash(st) == str(g["hash"])
