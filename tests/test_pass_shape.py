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


# -- the half-tile exchange, emulated on the host ----------------------------

def _groups(xm, zm, tile, low):
    rot = (nat.Rot * len(xm))()
    for k in range(len(xm)):
        rot[k] = nat.Rot(int(xm[k]), int(zm[k]), 1.0, 0.0)
    out = (C.c_uint32 * (20 * 48))()
    L = nat.lib()
    L.ph2hip_pass_groups.restype = C.c_long
    ng = L.ph2hip_pass_groups(rot, C.c_size_t(len(xm)), C.c_uint32(tile),
                              C.c_uint32(low), out, C.c_size_t(48))
    assert ng > 0
    a = np.frombuffer(out, dtype=np.uint32)[:20 * ng].reshape(ng, 20)
    return [dict(u=[int(x) for x in r[0:4]], jpos=[int(x) for x in r[4:16]],
                 xh=int(r[16]), xk=int(r[17]), first=int(r[18]),
                 count=int(r[19])) for r in a]


def _half_slot(v, h, tb):
    # the kernel's map: bits h and tb-1 exchanged, bits 4 and up folded
    # into bits 0..3 by y ^ (y >> 1) ^ (y >> 3) (y = v >> 4), bit tb-1
    # dropped
    t = ((v >> h) ^ (v >> (tb - 1))) & 1
    v ^= t * ((1 << h) | (1 << (tb - 1)))
    v ^= ((v >> 4) ^ (v >> 5) ^ (v >> 7)) & 15
    return v & ((1 << (tb - 1)) - 1)


def _coset(g, j, w, tb, d):
    bu = 0
    for k in range(tb - d):
        bu |= ((j >> k) & 1) << g["jpos"][k]
    for i in range(d):
        if w >> i & 1:
            bu ^= g["u"][i]
    return bu


@pytest.mark.parametrize("n,seed,low,tile", [(20, 11, 3, 12), (24, 12, 4, 12),
                                             (22, 13, 3, 13), (30, 14, 5, 13)])
def test_half_tile_exchange_moves_every_amplitude(n, seed, low, tile):
    """For every exchange of a half-tile pass: each thread's coset lies in
    the half of the tile its wave handles, the write slots of group q fill
    each half exactly once, and group q+1 reads back from each slot the
    very tile coordinate it expects (the kernel's slot map, restated)."""
    d = 4
    nthreads = 1 << (tile - d)
    checked = 0
    for xm, zm in _passes(n, seed, tile, low, maxlen=40):
        rc, sh = _shape(xm, zm, tile, low)
        assert rc == 0
        if not sh[3] & 4 or sh[2] < 2:
            continue
        gs = _groups(xm, zm, tile, low)
        assert len(gs) == sh[2]
        assert sum(g["count"] for g in gs) == len(xm)
        for q in range(len(gs) - 1):
            g0, g1 = gs[q], gs[q + 1]
            h, xk = g0["xh"], g0["xk"]
            assert xk in (tile - d - 1, tile - d - 2)   # wave-select bits
            assert g0["jpos"][xk] == h and g1["jpos"][xk] == h
            table = {}
            for j in range(nthreads):
                half = (j >> xk) & 1
                for w in range(1 << d):
                    v = _coset(g0, j, w, tile, d)
                    assert (v >> h) & 1 == half
                    key = (half, _half_slot(v, h, tile))
                    assert key not in table
                    table[key] = v
            assert len(table) == 1 << tile
            for j in range(nthreads):
                half = (j >> xk) & 1
                for w in range(1 << d):
                    v = _coset(g1, j, w, tile, d)
                    assert table[(half, _half_slot(v, h, tile))] == v
            checked += 1
    assert checked > 0


def _exchange_conflicts(n, seed, low, tile, slot):
    """Lane groups served together in the exchanges of half-tile passes,
    and how many of them meet a bank conflict, by the bank rules measured
    on gfx950 with the LDS conflict counters: a 16-byte write serves lanes
    0..7 of a group of eight (lane bits 0..2) per clock over 32 banks
    (slots distinct mod 8), a 16-byte read lanes 0..15 (lane bits 0..3)
    over 64 banks (distinct mod 16).  Group q writes with its layout and
    group q+1 reads with its own.  The register offset of an access is
    uniform, so register 0 stands for all 16."""
    d = 4
    nthreads = 1 << (tile - d)
    total = bad = 0
    for xm, zm in _passes(n, seed, tile, low, maxlen=40):
        rc, sh = _shape(xm, zm, tile, low)
        if rc != 0 or not sh[3] & 4 or sh[2] < 2:
            continue
        gs = _groups(xm, zm, tile, low)
        for q in range(len(gs) - 1):
            h = gs[q]["xh"]
            for g, lanes, banks in ((gs[q], 8, 8), (gs[q + 1], 16, 16)):
                for j0 in range(0, nthreads, lanes):
                    sl = {slot(_coset(g, j, 0, tile, d), h, tile)
                          for j in range(j0, j0 + lanes)}
                    total += 1
                    bad += len({x % banks for x in sl}) < len(sl)
    return total, bad


def _whole_fold_half_slot(v, h, tb):
    # the fold of the whole-tile form (bank bit t mod 4 for tile bit t < 12)
    t = ((v >> h) ^ (v >> (tb - 1))) & 1
    v ^= t * ((1 << h) | (1 << (tb - 1)))
    v ^= ((v >> 4) ^ (v >> 8)) & 15
    return v & ((1 << (tb - 1)) - 1)


@pytest.mark.parametrize("n,seed,low,tile", [(24, 21, 3, 12),
                                             (30, 22, 3, 13)])
def test_half_tile_exchange_spreads_lanes_over_banks(n, seed, low, tile):
    """The lanes served together take coset-index bits on free tile bits
    whose bank vectors under the half-tile fold are independent (three
    mod 8 for writes, four mod 16 for reads), so nearly every lane group of
    an exchange is conflict-free; the same layouts under the whole-tile
    fold are not."""
    total, bad = _exchange_conflicts(n, seed, low, tile, _half_slot)
    assert total > 1000
    assert bad <= 0.03 * total, (bad, total)
    total_w, bad_w = _exchange_conflicts(n, seed, low, tile,
                                         _whole_fold_half_slot)
    assert bad_w > 3 * bad + 0.02 * total_w, (bad_w, bad, total_w)
