"""Cost model: feature extraction (CPU), the HIP network against the numpy
oracle (GPU), and the learned schedule's bit identity (GPU)."""

import ctypes as C
import time

import numpy as np
import pytest

from oracle import mlp_oracle
from phase2_amd import _native as nat
from phase2_amd import costmodel as cm
from phase2_amd import synth


@pytest.mark.parametrize("direct", [2, 5, 7])
def test_features_match_c(direct):
    """The Python feature mirror equals ph2_learned_features, including the
    direct first/last-group and half-tile flags (bits 0, 1, 2)."""
    rng = np.random.default_rng(0)
    xm = rng.integers(0, 1 << 33, size=7, dtype=np.uint64)
    zm = rng.integers(0, 1 << 33, size=7, dtype=np.uint64)
    f = cm.c_features(xm, zm, 12, (6, 5, 3, direct))
    assert f[-3:].tolist() == [direct & 1, direct >> 1 & 1, direct >> 2 & 1]
    rec = np.zeros(1, dtype=nat.REC_DTYPE)
    rec["kind"], rec["nqb"], rec["tile_bits"], rec["low_bits"] = 2, 33, 12, 6
    rec["gdim"], rec["nterms"], rec["ngroups"] = 5, 7, 3
    rec["direct"] = direct
    rec["xunion"] = np.bitwise_or.reduce(xm)
    rec["zunion"] = np.bitwise_or.reduce(zm)
    rec["ms"] = 1.0
    X, y, _ = cm.features_from_records(rec)
    assert np.array_equal(X[0], f)
    assert abs(y[0] - (-3 * np.log(2))) < 1e-6


def test_random_schedule_feasible():
    hm = synth.hamiltonian(20, 400, seed=2)
    rng = np.random.default_rng(1)
    lens, tbs, lbs = cm.random_schedule(hm.xm[:300], rng)
    assert lens.sum() == 300 and lens.min() >= 1 and lens.max() <= 48
    assert set(lbs.tolist()) <= {3, 4, 5} and len(set(lbs.tolist())) > 1
    k = 0
    for ln, tb, lb in zip(lens, tbs, lbs):
        assert cm._span_growth(hm.xm[k:k + ln], int(lb))[-1] <= tb - lb
        k += ln


def test_oracle_gradient_finite_differences():
    rng = np.random.default_rng(3)
    nin, nhid = 6, 5
    p = rng.normal(size=nhid * nin + nhid + nhid * nhid + nhid + nhid + 1)
    X = rng.normal(size=(9, nin))
    t = rng.normal(size=9)
    _, g = mlp_oracle.loss_grad(p, X, t, nin, nhid)
    eps = 1e-6
    for k in rng.choice(len(p), 25, replace=False):
        e = np.zeros_like(p)
        e[k] = eps
        fd = (mlp_oracle.loss_grad(p + e, X, t, nin, nhid)[0] -
              mlp_oracle.loss_grad(p - e, X, t, nin, nhid)[0]) / (2 * eps)
        assert abs(fd - g[k]) < 1e-7 * max(1, abs(g[k]))


@pytest.mark.gpu
def test_mlp_forward_grad_vs_oracle():
    rng = np.random.default_rng(4)
    m = cm.MLP(seed=7)
    p = m.get_params().astype(np.float64)
    X = rng.random((300, cm.NIN)).astype(np.float32)
    t = rng.normal(size=300).astype(np.float32)
    y = m.forward(X)
    y_ref, _ = mlp_oracle.forward(p, X.astype(np.float64), cm.NIN, cm.NHID)
    assert np.max(np.abs(y - y_ref)) <= 1e-4 * max(1, np.max(np.abs(y_ref)))
    loss, g = m.grad(X, t)
    loss_ref, g_ref = mlp_oracle.loss_grad(p, X.astype(np.float64),
                                           t.astype(np.float64), cm.NIN,
                                           cm.NHID)
    assert abs(loss - loss_ref) <= 1e-4 * max(1, loss_ref)
    assert np.max(np.abs(g - g_ref)) <= 1e-4 * max(1, np.max(np.abs(g_ref)))
    # First Adam step (bias-corrected moments m = g, v = g^2): every
    # parameter moves by -lr * g / (|g| + eps), eps = 1e-8
    lr = 1e-3
    l0 = m.train_step(X, t, lr)
    assert abs(l0 - loss_ref) <= 1e-4 * max(1, loss_ref)
    p1 = m.get_params().astype(np.float64)
    g64 = g.astype(np.float64)
    want = p - lr * g64 / (np.abs(g64) + 1e-8)
    assert np.max(np.abs(p1 - want)) <= 1e-6
    # and further steps on the fixed batch reduce the loss
    for _ in range(100):
        l1 = m.train_step(X, t, lr)
    assert l1 < 0.8 * l0
    m.close()


@pytest.mark.gpu
def test_learned_schedule_bit_identical(oracle):
    """With the learned schedule installed (any parameters), fused sweeps
    stay bit-identical to the oracle."""
    nat.init()
    m = cm.MLP(seed=11)
    p = m.get_params()
    m.close()
    n = 20
    hm = synth.hamiltonian(n, 1500, seed=12, decay=True)
    perm = oracle.sort_lex(hm.xm, hm.zm)
    xm, zm = hm.xm[perm], hm.zm[perm]
    phi = 0.2 * hm.cf[perm] / np.max(np.abs(hm.cf))
    rng = np.random.default_rng(5)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    want = oracle.rotate(st, xm, zm, phi)
    try:
        cm.install(p, low=3, min_qb=16)
        got = nat.rotate(st, xm, zm, phi, nat.FUSED, 12, 4)
        again = nat.rotate(st, xm, zm, phi, nat.FUSED, 12, 4)
        hits, misses = cm.stats()
    finally:
        cm.uninstall()
    assert np.array_equal(got, want)
    assert np.array_equal(again, want)
    assert misses >= 1 and hits >= misses   # second sweep reuses schedules


def _time_sched(reg, xm, zm, phi, sched):
    """Wall seconds of one block run under an explicit schedule."""
    L = nat.lib()
    L.ph2_reg_rotate_sched.restype = C.c_int
    u32p = C.POINTER(C.c_uint32)
    lens, tbs, lbs = (np.ascontiguousarray(a, np.uint32) for a in sched)
    nat.synchronize()
    t0 = time.perf_counter()
    rc = L.ph2_reg_rotate_sched(
        C.c_void_p(reg.h), C.c_size_t(len(xm)),
        xm.ctypes.data_as(C.POINTER(C.c_uint64)),
        zm.ctypes.data_as(C.POINTER(C.c_uint64)),
        phi.ctypes.data_as(C.POINTER(C.c_double)),
        lens.ctypes.data_as(u32p), tbs.ctypes.data_as(u32p),
        lbs.ctypes.data_as(u32p), C.c_size_t(len(lens)))
    nat.synchronize()
    assert rc == 0
    return time.perf_counter() - t0


@pytest.mark.gpu
def test_learned_schedule_gain_over_greedy():
    """The committed network's schedule against the library's default
    greedy cut (tile 12, low 3) on 30-qubit sweep blocks of a Hamiltonian
    it was not trained on (generator seed 7).  Each block is timed under
    both schedules, interleaved, REPS times after a warm-up; a block's
    noise is the larger relative spread (max/min - 1) of its two series,
    floored at 0.5%.  Where the network predicts a margin over greedy of
    more than twice that noise, the learned schedule's best time must not
    exceed greedy's by more than the noise; at least one block must be such
    a block, and over all blocks the learned schedules must be faster.  The
    predicted total is never above greedy's (the greedy passes are among
    the candidates the dynamic program chooses from)."""
    nat.init()
    params, _ = cm.load()
    n, reps = 30, 5
    hm = synth.hamiltonian(n, 0, seed=7, decay=True, n_total=1024 * 14)
    perm = nat.sort_lex(hm.xm, hm.zm)
    xm, zm, phi = hm.xm[perm], hm.zm[perm], 0.02 * hm.cf[perm]
    blocks = [(np.ascontiguousarray(xm[s]), np.ascontiguousarray(zm[s]),
               np.ascontiguousarray(phi[s]))
              for s in (slice(b * 1024, (b + 1) * 1024)
                        for b in (2, 5, 8, 11))]
    rows = []
    try:
        cm.install(params, min_qb=n)
        scheds = []
        for bx, bz, _ in blocks:
            g = cm.schedule(bx, bz, n, greedy=(12, 3))
            lr = cm.schedule(bx, bz, n)
            assert lr[0].sum() == 1024 and g[0].sum() == 1024
            pg = cm.predict(bx, bz, n, g).sum()
            pl = cm.predict(bx, bz, n, lr).sum()
            assert pl <= pg * (1 + 1e-6)
            scheds.append((g, lr, pg, pl))
        reg = nat.Register(n, nat.FUSED, 12, 3)
        try:
            reg.prep_multidet(np.array([(1 << 15) - 1], np.uint64),
                              np.array([1.0 + 0j]))
            times = [([], []) for _ in blocks]
            for rep in range(reps + 1):
                for k, (bx, bz, bp) in enumerate(blocks):
                    for m in (0, 1):
                        dt = _time_sched(reg, bx, bz, bp, scheds[k][m])
                        if rep:
                            times[k][m].append(dt)
        finally:
            reg.free()
    finally:
        cm.uninstall()
    checked = 0
    tot_g = tot_l = 0.0
    for k, (tg, tl) in enumerate(times):
        tg, tl = np.array(tg), np.array(tl)
        noise = max(tg.max() / tg.min() - 1, tl.max() / tl.min() - 1, 0.005)
        g, lr, pg, pl = scheds[k]
        margin = 1 - pl / pg
        gain = tg.min() / tl.min()
        tot_g += tg.min()
        tot_l += tl.min()
        rows.append((k, len(g[0]), len(lr[0]), margin, noise, gain))
        print(f"block {k}: passes greedy {len(g[0])} learned {len(lr[0])}; "
              f"predicted margin {margin:.4f}, noise {noise:.4f}, "
              f"measured gain {gain:.4f}x", flush=True)
        if margin > 2 * noise:
            checked += 1
            assert tl.min() <= tg.min() * (1 + noise), rows[-1]
    print(f"total: greedy {tot_g:.4f} s, learned {tot_l:.4f} s, "
          f"gain {tot_g / tot_l:.4f}x, blocks checked {checked}")
    assert checked >= 1, rows
    assert tot_l < tot_g, rows


def test_load_rejects_other_feature_vectors(tmp_path):
    """A parameter file trained for another feature vector (or a truncated
    array) is refused before it can reach ph2_learned_install."""
    p, meta = cm.load()
    assert p.size == cm.n_params() and int(meta["nin"]) == cm.NIN
    bad = tmp_path / "old.npz"
    np.savez(bad, params=np.zeros(cm.n_params(cm.NIN - 1), np.float32),
             nin=cm.NIN - 1, nhid=cm.NHID)
    with pytest.raises(ValueError):
        cm.load(bad)
    with pytest.raises(ValueError):
        cm.install(p[:-1])
