"""Correctness at the full sizes of the workloads, where the oracle cannot
follow the whole run:

* 33 qubits (128 GiB, indices beyond 2^32): a window of the lexicographic
  sweep that flips the top qubits, fused (learned schedule, greedy tile 12
  and 13) against the plain kernel on two resident registers, compared on
  the device double by double; and a sparse closed form whose flips span
  qubits 30..32, against a host dictionary simulation;
* 26 qubits: the 230,230-determinant coefficient-matrix state against the
  oracle's, and its norm by the Cauchy-Binet identity;
* 30 qubits: one composite-integrator step, fused against plain;
* the 512-register batched qDRIFT workload: batched launch against
  register-by-register fused passes, plus two registers against the oracle.

Bit-exact where the reference is (every rotation and multidet overlap);
1e-12 for the coefficient-matrix norm (a device tree reduction)."""

import math

import numpy as np
import pytest

from oracle import pyoracle as po
from phase2_amd import _native as nat
from phase2_amd import costmodel as cm
from phase2_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    print("fullsize: init", flush=True)
    nat.init()
    yield
    cm.uninstall()
    nat.set_default(nat.FUSED, 0, 3)


def test_fill_hash_matches_host_formula():
    reg = nat.Register(12, nat.PLAIN, 12, 3)
    try:
        reg.fill_hash(0xC0FFEE)
        st = reg.read()
    finally:
        reg.free()
    assert np.array_equal(st, nat.hash_state(12, 0xC0FFEE))


def test_count_unequal_counts_components():
    a = nat.Register(10, nat.PLAIN, 10, 3)
    b = nat.Register(10, nat.PLAIN, 10, 3)
    try:
        a.fill_hash(1)
        b.fill_hash(1)
        assert a.count_unequal(b) == 0
        b.prep_multidet(np.array([5], np.uint64), np.array([-0.0 + 0.0j]))
        a.prep_multidet(np.array([5], np.uint64), np.array([0.0 - 0.0j]))
        assert a.count_unequal(b) == 0           # -0 == +0
        b.prep_multidet(np.array([5, 9], np.uint64),
                        np.array([1.0 + 0j, 2.0j]))
        assert a.count_unequal(b) == 2           # re of 5, im of 9
    finally:
        a.free()
        b.free()


def _high_window(nterms=96):
    """A window of the 33-qubit lexicographic sweep whose flip masks touch
    qubits 30, 31 and 32 (and the low qubits)."""
    w = synth.trott_n33()
    hm = w["hm"]
    perm = nat.sort_lex(hm.xm, hm.zm)
    xm, zm, cf = hm.xm[perm], hm.zm[perm], hm.cf[perm]
    top = np.uint64(0b111 << 30)
    hits = np.nonzero((xm & top) == top)[0]
    assert hits.size > 0
    k0 = int(hits[0])
    s = slice(k0, k0 + nterms)
    for q in (30, 31, 32):
        assert np.any(xm[s] >> np.uint64(q) & np.uint64(1))
    return xm[s], zm[s], 0.02 * cf[s]


def test_fused_equals_plain_n33():
    """96 consecutive terms of the headline sweep, including terms that flip
    qubits 30..32, on a dense 128 GiB state: the fused passes (learned
    schedule; greedy cuts at tile 12 and tile 13) leave every one of the
    2^34 doubles equal to the plain kernel's."""
    xm, zm, phi = _high_window()
    params, _ = cm.load()
    ref = nat.Register(33, nat.PLAIN, 12, 3)
    fus = nat.Register(33, nat.FUSED, 0, 3)
    try:
        print("n33: plain sweep of 96 terms", flush=True)
        ref.fill_hash(33)
        ref.rotate(xm, zm, phi)
        fus.fill_hash(33)
        assert fus.count_unequal(ref) > 0        # the window changed it
        for label, tile, learned in (("learned", 0, True),
                                     ("greedy 13/3", 13, False),
                                     ("greedy 12/3", 12, False)):
            print(f"n33: fused sweep, {label}", flush=True)
            if learned:
                cm.install(params, min_qb=30)
            else:
                cm.uninstall()
            fus.configure(nat.FUSED, tile, 3)
            fus.fill_hash(33)
            nat.prof_enable(True)
            nat.prof_records(reset=True)
            fus.rotate(xm, zm, phi)
            rec = nat.prof_records(reset=True)
            nat.prof_enable(False)
            fused = rec[rec["kind"] == 2]
            assert fused.size > 0 and fused["nterms"].sum() == len(xm)
            if tile:
                assert set(fused["tile_bits"].tolist()) == {tile}
            else:
                assert 13 in set(fused["tile_bits"].tolist())
            assert np.any(fused["direct"] & 4)   # half-tile passes ran
            n = fus.count_unequal(ref)
            print(f"n33: {label}: {fused.size} passes, {n} unequal",
                  flush=True)
            assert n == 0
    finally:
        nat.prof_enable(False)
        cm.uninstall()
        ref.free()
        fus.free()


def _dict_rotate(state, x, z, phi):
    """exp(i phi P) on a sparse state {index: amplitude}:
    new[j] = cos(phi) a[j] + i sin(phi) i^r4(j^x) a[j^x], with
    r4(i) = popc(x & z) + 2 popc(i & z) mod 4 (P|i> = i^r4 |i^x>)."""
    c, s = math.cos(phi), math.sin(phi)
    y = bin(x & z).count("1")
    out = {}
    for j in set(state) | {i ^ x for i in state}:
        i = j ^ x
        r4 = (y + 2 * bin(i & z).count("1")) & 3
        out[j] = c * state.get(j, 0) + 1j * s * (1j ** r4) * state.get(i, 0)
    return out


def test_sparse_closed_form_n33():
    """Two determinants above 2^32 and 12 terms whose flip masks span a
    10-dimensional space that includes qubits 30, 31 and 32: every
    amplitude of the 2,048-element support matches a host dictionary
    simulation to 1e-14, and the norm outside the support is zero, for the
    plain kernel, the learned schedule and greedy tile 13."""
    rng = np.random.default_rng(33)
    qubits = [0, 2, 5, 9, 14, 20, 26, 30, 31, 32]
    dets = {(1 << 32) | (1 << 31) | 0b10110: 0.6 + 0.0j,
            (1 << 32) | (1 << 30) | (1 << 27) | 0b1101: 0.0 + 0.8j}
    terms = []
    for k in range(12):
        x = 0
        while x == 0:
            for q in qubits:
                if rng.random() < 0.35:
                    x |= 1 << q
        if k in (3, 7):
            x = (1 << 30) | (1 << 31) | (1 << 32)
        z = int(rng.integers(0, 1 << 33))
        terms.append((x, z, float(rng.normal()) * 0.7))
    want = dict(dets)
    for x, z, phi in terms:
        want = _dict_rotate(want, x, z, phi)
    keys = np.array(sorted(want), dtype=np.uint64)
    # two disjoint cosets of the 10-dimensional flip span, half of the
    # support above 2^32
    assert len(keys) == 2048
    assert np.count_nonzero(keys >> np.uint64(32)) == 1024
    expect = np.array([want[int(k)] for k in keys])
    xm = np.array([t[0] for t in terms], np.uint64)
    zm = np.array([t[1] for t in terms], np.uint64)
    phi = np.array([t[2] for t in terms])
    params, _ = cm.load()
    reg = nat.Register(33, nat.PLAIN, 12, 3)
    try:
        for label, mode, tile, learned in (("plain", nat.PLAIN, 12, False),
                                           ("learned", nat.FUSED, 0, True),
                                           ("greedy 13/3", nat.FUSED, 13,
                                            False)):
            print(f"n33 sparse: {label}", flush=True)
            if learned:
                cm.install(params, min_qb=30)
            else:
                cm.uninstall()
            reg.configure(mode, tile, 3)
            reg.prep_multidet(np.array(list(dets), np.uint64),
                              np.array(list(dets.values())))
            reg.rotate(xm, zm, phi)
            got = reg.gather(keys)
            err = np.max(np.abs(got - expect))
            resid = reg.norm2() - float(np.sum(np.abs(got) ** 2))
            print(f"n33 sparse: {label}: max error {err:.2e}, norm outside "
                  f"the support {resid:.2e}", flush=True)
            assert err <= 1e-14
            assert abs(resid) <= 1e-14
    finally:
        cm.uninstall()
        reg.free()


def test_expand_230230_n26():
    """The 26-qubit coefficient-matrix state (n_sites 26, 6 alpha electrons,
    230,230 determinants): the same state as the oracle's (every index and
    value), and norm 1 to 1e-12 by Cauchy-Binet (C has orthonormal
    columns), through the device inner product."""
    w = synth.trott2_n26()
    C = w["C_alpha"]
    print("n26: expand 230230 determinants", flush=True)
    st = nat.expand(26, 26, 6, 0, 1, 0, C)
    assert np.count_nonzero(st) == math.comb(26, 6)
    want = po.oracle().expand(26, 26, 6, 0, 1, 0, C)
    assert po.state_hash(st) == po.state_hash(want)
    z = nat.inner(st, 26, 6, 0, 1, 0, C)
    print("n26: <psi0|psi0> =", z, flush=True)
    assert abs(z - 1.0) <= 1e-12


def test_cmpsit_step_fused_equals_plain_n30():
    """One step of the 30-qubit composite integrator (2,000 deterministic
    and 512 sampled rotations per half-step): the state after the learned
    fused schedule equals the plain kernel's, every double."""
    w = synth.cmpsit_n30()
    prep = dict(det_idx=w["det_idx"], det_cf=w["det_cf"])
    args = (w["hm"], prep, 1, 1, w["length"], w["depth"], w["angle_det"],
            w["angle_rand"], w["seed"])
    params, _ = cm.load()
    try:
        print("n30: cmpsit step, plain", flush=True)
        nat.set_default(nat.PLAIN, 12, 3)
        ov_p, st_p = nat.cmpsit(*args)
        print("n30: cmpsit step, learned fused", flush=True)
        cm.install(params, min_qb=26)
        nat.set_default(nat.FUSED, 0, 3)
        ov_f, st_f = nat.cmpsit(*args)
    finally:
        cm.uninstall()
        nat.set_default(nat.FUSED, 0, 3)
    assert np.array_equal(ov_f, ov_p)
    assert np.array_equal(st_f, st_p)
    assert abs(np.vdot(st_f, st_f).real - 1.0) < 1e-12


def test_qdrift_batch_512_full():
    """The 512-register qDRIFT workload (16..26 qubits, 107 GiB in one
    buffer, depth 1000), one sample: the batched launch gives the same
    overlaps as register-by-register fused passes, bit for bit."""
    regs = synth.qdrift_batch(512, samples=1)
    print("c4: batched, 512 registers", flush=True)
    ob, sb = nat.batch_qdrift(regs, nat.BATCH_FUSED_BATCHED, 12, 3)
    print("c4: register by register", flush=True)
    oe, se = nat.batch_qdrift(regs, nat.BATCH_FUSED_EACH, 12, 3)
    print(f"c4: {sb['launches']} batched launches, {se['launches']} "
          "register-by-register", flush=True)
    assert all(np.array_equal(a, b) for a, b in zip(ob, oe))
    assert sb["launches"] < se["launches"]


def test_qdrift_batch_vs_oracle_large():
    """Two registers of 22 and 23 qubits (depth 100, two samples) in one
    batched launch against the oracle's qdrift on each alone."""
    regs = [r for r in synth.qdrift_batch(64, seed=3, samples=2, depth=100,
                                          nmin=22, nmax=23)[:6]
            if r["hm"].nqb in (22, 23)][:2]
    assert {r["hm"].nqb for r in regs} == {22, 23}
    print("c4: two large registers batched", flush=True)
    got, _ = nat.batch_qdrift(regs, nat.BATCH_FUSED_BATCHED, 12, 3)
    o = po.oracle()
    for r, ov in zip(regs, got):
        want, _ = o.qdrift(r["hm"], dict(det_idx=r["det_idx"],
                                        det_cf=r["det_cf"]),
                           r["samples"], r["depth"], r["step_size"],
                           r["seed"], want_state=False)
        assert np.array_equal(ov, want)
