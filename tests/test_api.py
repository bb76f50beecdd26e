"""The widened Python entry beside phase2.run: Pauli-string parsing and
input checks (CPU), and run_multidet / qdrift_batch against run and the
oracle (GPU)."""

import numpy as np
import pytest

import phase2_amd
from phase2_amd import synth


def test_pauli_masks_grammar():
    xm, zm = phase2_amd.pauli_masks(
        ["X0 Y1", "Z2\tI3", "", "X0 Z0", "Y-0", "X+1", "  Z3\n"], 4)
    assert xm.tolist() == [3, 0, 0, 0, 1, 2, 0]
    assert zm.tolist() == [2, 4, 0, 1, 1, 0, 8]


@pytest.mark.parametrize("bad", ["W0", "X4", "X-1", "X", "X1a", "X0\rY1",
                                 "x0", "X" + "0" * 4096])
def test_pauli_masks_rejects(bad):
    with pytest.raises(ValueError):
        phase2_amd.pauli_masks([bad], 4)


@pytest.mark.parametrize("args", [
    (["Z0"], [1.0, 2.0], 0.1, {"0": 1.0}),      # lengths differ
    (["Z0"], [1.0], 0.1, {}),                  # empty state
    (["Z0"], [1.0], 0.1, {"02": 1.0}),         # bad bit string
    (["Z0"], [1.0], 0.1, {"01": 1.0, "1": 1}),  # uneven widths
    (["Z0"], [1.0], 0.1, {5: 1.0}),            # no width for an index
    (["Z3"], [1.0], 0.1, {"01": 1.0}),         # qubit beyond the state
])
def test_run_multidet_rejects(args):
    with pytest.raises(ValueError):
        phase2_amd.run_multidet(*args)


def _strings(xm, zm, n):
    out = []
    for x, z in zip(xm, zm):
        toks = []
        for q in range(n):
            b = (int(x) >> q & 1, int(z) >> q & 1)
            if b != (0, 0):
                toks.append({(1, 0): "X", (1, 1): "Y", (0, 1): "Z"}[b] +
                            str(q))
        out.append(" ".join(toks))
    return out


@pytest.mark.gpu
def test_run_multidet_single_det_equals_run():
    """One determinant with amplitude 1: exactly phase2.run's value."""
    rng = np.random.default_rng(21)
    for n, nfour in ((5, 10), (9, 150), (12, 600)):
        hm = synth.hamiltonian(n, nfour, seed=n)
        paulis = _strings(hm.xm, hm.zm, n)
        psi = "".join(rng.choice(["0", "1"], size=n))
        want = phase2_amd.run(paulis, list(hm.cf), 0.3, psi)
        got = phase2_amd.run_multidet(paulis, list(hm.cf), 0.3, {psi: 1.0})
        assert got == want


@pytest.mark.gpu
@pytest.mark.parametrize("n,nfour,ndet", [(10, 200, 5), (20, 1500, 16)])
def test_run_multidet_vs_oracle(oracle, n, nfour, ndet):
    """A superposition: the overlap sum_d conj(c_d) <i_d|U|psi> against the
    oracle's state (same order of rotations), summed in determinant order
    in numpy; 1e-12 relative (the complex products may round differently
    from the library's host sum)."""
    hm = synth.hamiltonian(n, nfour, seed=30 + n, decay=True)
    idx, cf = synth.multidet_random(n, ndet, n // 2, seed=n)
    dets = {int(i): complex(c) for i, c in zip(idx, cf)}
    got = phase2_amd.run_multidet(_strings(hm.xm, hm.zm, n), list(hm.cf),
                                  0.05, dets, nqb=n)
    st = np.zeros(1 << n, dtype=np.complex128)
    st[idx.astype(np.int64)] = cf
    st = oracle.rotate(st, hm.xm, hm.zm, 0.05 * hm.cf)
    want = complex(0)
    for i, c in zip(idx, cf):
        want += st[int(i)] * np.conj(c)
    assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


@pytest.mark.gpu
def test_qdrift_batch_api_vs_oracle(oracle):
    regs, want = [], []
    for i, n in enumerate((12, 13, 14)):
        hm = synth.hamiltonian(n, 60 + 5 * n, seed=300 + i)
        idx, cf = synth.multidet_random(n, 2, n // 2, seed=400 + i)
        regs.append(dict(paulis=_strings(hm.xm, hm.zm, n), coeffs=hm.cf,
                         nqb=n, dets={int(a): complex(b)
                                      for a, b in zip(idx, cf)},
                         samples=2, depth=90, step_size=0.02,
                         seed=0x4400 + i))
        w, _ = oracle.qdrift(hm, dict(det_idx=idx, det_cf=cf), 2, 90, 0.02,
                             0x4400 + i, want_state=False)
        want.append(w)
    for mode in ("batched", "fused", "plain"):
        got = phase2_amd.qdrift_batch(regs, mode)
        for g, w in zip(got, want):
            assert np.array_equal(g, w), mode
