"""Cost model: feature extraction (CPU), the HIP network against the numpy
oracle (GPU), and the learned schedule's bit identity (GPU)."""

import numpy as np
import pytest

from oracle import mlp_oracle
from phase2_amd import _native as nat
from phase2_amd import costmodel as cm
from phase2_amd import synth


def test_features_match_c():
    rng = np.random.default_rng(0)
    xm = rng.integers(0, 1 << 33, size=7, dtype=np.uint64)
    zm = rng.integers(0, 1 << 33, size=7, dtype=np.uint64)
    f = cm.c_features(xm, zm, 12, (6, 5, 3))
    rec = np.zeros(1, dtype=nat.REC_DTYPE)
    rec["kind"], rec["nqb"], rec["tile_bits"], rec["low_bits"] = 2, 33, 12, 6
    rec["gdim"], rec["nterms"], rec["ngroups"] = 5, 7, 3
    rec["xunion"] = np.bitwise_or.reduce(xm)
    rec["zunion"] = np.bitwise_or.reduce(zm)
    rec["ms"] = 1.0
    X, y, _ = cm.features_from_records(rec)
    assert np.array_equal(X[0], f)
    assert abs(y[0] - (-3 * np.log(2))) < 1e-6


def test_random_schedule_feasible():
    hm = synth.hamiltonian(20, 400, seed=2)
    rng = np.random.default_rng(1)
    lens, tbs = cm.random_schedule(hm.xm[:300], rng)
    assert lens.sum() == 300 and lens.min() >= 1 and lens.max() <= 48
    k = 0
    for ln, tb in zip(lens, tbs):
        assert cm._span_growth(hm.xm[k:k + ln], 4)[-1] <= tb - 4
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
    # Adam steps on a fixed batch reduce the loss
    l0 = m.train_step(X, t, 1e-3)
    for _ in range(50):
        l1 = m.train_step(X, t, 1e-3)
    assert l1 < 0.5 * l0
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
        cm.install(p, low=4, min_qb=16)
        got = nat.rotate(st, xm, zm, phi, nat.FUSED, 12, 4)
        again = nat.rotate(st, xm, zm, phi, nat.FUSED, 12, 4)
        hits, misses = cm.stats()
    finally:
        cm.uninstall()
    assert np.array_equal(got, want)
    assert np.array_equal(again, want)
    assert misses >= 1 and hits >= misses   # second sweep reuses schedules
