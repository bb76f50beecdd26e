"""The synthetic generators produce the term families they document."""

import math

import numpy as np

from phase2_amd import synth


def _check_terms(hm, n):
    x = [int(v) for v in hm.xm]
    z = [int(v) for v in hm.zm]
    assert abs(np.sum(np.abs(hm.cf)) - 1.0) < 1e-12
    assert np.all(np.diff(np.abs(hm.cf)) <= 0)      # |c| descending
    assert len(set(zip(x, z))) == len(x)             # distinct strings
    four = 0
    for xm, zm in zip(x, z):
        if xm == 0:
            continue
        four += 1
        q = [k for k in range(n) if xm >> k & 1]
        assert len(q) == 4
        p0, p1, p2, p3 = q
        y = zm & xm
        assert bin(y).count("1") % 2 == 0            # even number of Y
        between = sum(1 << k for k in range(p0 + 1, p1)) | \
            sum(1 << k for k in range(p2 + 1, p3))
        assert zm & ~xm == between
    return four


def test_hamiltonian_counts():
    hm = synth.hamiltonian(10, 60, seed=3)
    assert len(hm) == 10 + 45 + 60 + 1
    assert _check_terms(hm, 10) == 60


def test_hamiltonian_vec_structure():
    hm = synth.hamiltonian_vec(12, 300, seed=5)
    assert len(hm) == 12 + 66 + 300 + 1
    assert _check_terms(hm, 12) == 300


def test_workload_sizes():
    w = synth.trott_n16()
    assert len(w["hm"]) == 537
    regs = synth.qdrift_batch(8, seed=1)
    for r in regs:
        n = r["hm"].nqb
        assert 16 <= n <= 26 and 2000 <= len(r["hm"]) <= 10000
        assert r["det_idx"].max() < 2 ** n
        assert abs(np.sum(np.abs(r["det_cf"]) ** 2) - 1) < 1e-12


def test_coeff_matrix_orthonormal():
    C = synth.coeff_matrix(9, 4, seed=2)
    assert np.allclose(C.T @ C, np.eye(4), atol=1e-14)
    assert C.shape == (9, 4) and math.isfinite(C.sum())
