"""Host construction of fused-pass descriptors (ph2hip_pass_shape: no GPU
call): the half-tile form keeps the pass shape and is available for
multi-group tile-12 passes; other tile widths never take it."""

import ctypes as C
import math

import numpy as np
import pytest

from phase2_amd import _native as nat, synth


def _shape(xm, zm, tile, low):
    rot = (nat.Rot * len(xm))()
    for k in range(len(xm)):
        rot[k] = nat.Rot(int(xm[k]), int(zm[k]), math.cos(0.1),
                         math.sin(0.1))
    sh = (C.c_uint32 * 4)()
    rc = nat.lib().ph2hip_pass_shape(rot, C.c_size_t(len(xm)),
                                     C.c_uint32(tile), C.c_uint32(low), sh)
    return rc, list(sh)


def _passes(n, seed, tile, low, maxlen=24):
    """Greedy cuts of lex-ordered terms: the longest prefixes that fit."""
    hm = synth.hamiltonian(n, 600, seed=seed, decay=True)
    perm = nat.sort_lex(hm.xm, hm.zm)
    xm, zm = hm.xm[perm], hm.zm[perm]
    out, k = [], 0
    while k < len(xm) and len(out) < 40:
        e = k + 1
        while e < len(xm) and e - k < maxlen and \
                _shape(xm[k:e + 1], zm[k:e + 1], tile, low)[0] == 0:
            e += 1
        out.append((xm[k:e], zm[k:e]))
        k = e
    return out


@pytest.mark.parametrize("n,seed,low,tile", [(20, 1, 4, 12), (24, 2, 3, 12),
                                             (28, 3, 5, 12), (22, 4, 4, 13),
                                             (30, 5, 3, 13)])
def test_half_form_keeps_shape(n, seed, low, tile):
    passes = _passes(n, seed, tile, low)
    nhalf = nmulti = 0
    for xm, zm in passes:
        try:
            nat.set_fused_half(0)
            rc0, s0 = _shape(xm, zm, tile, low)
            nat.set_fused_half(1 << tile)
            rc1, s1 = _shape(xm, zm, tile, low)
        finally:
            nat.set_fused_half(nat.HALF_TILES)
        assert rc0 == 0 and rc1 == 0
        assert s1[:3] == s0[:3]            # L, g, groups
        if s0[2] > 1:
            nmulti += 1
        if s1[3] & 4:
            nhalf += 1
            assert s1[3] & 3 == 3          # direct first and last group
        assert not s0[3] & 4
    assert nmulti > 0 and nhalf >= nmulti // 2


def test_half_form_only_at_tiles_12_and_13():
    xm, zm = _passes(20, 4, 11, 4)[0]
    try:
        nat.set_fused_half(nat.HALF_TILES)
        rc, s = _shape(xm, zm, 11, 4)
    finally:
        nat.set_fused_half(nat.HALF_TILES)
    assert rc == 0 and not s[3] & 4
    with pytest.raises(RuntimeError):
        nat.set_fused_half(1 << 11)
