"""Batched samplers fed by the device-side draws.

ph2_batch_cmpsit runs composite samples of many registers of uneven width
together (each half-step of every register in one round); every register
must give exactly the overlaps of cmpsit run on it alone with the same
seed -- against the oracle on fresh seeds (uneven steps and samples, so
registers start and end samples in different rounds), and against the
reference-written golden vectors through the batch API.  The qDRIFT batch
is pinned to its golden vector the same way."""

import numpy as np
import pytest

from phase2_amd import _native as nat
from phase2_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    print("batch_cmpsit: init", flush=True)
    nat.init()


def _cmpsit_regs(widths, steps, samples, length=30, depth=40):
    regs = []
    for i, (n, st, ns) in enumerate(zip(widths, steps, samples)):
        hm = synth.hamiltonian(n, 60 + 9 * n, seed=300 + i)
        idx, cf = synth.multidet_random(n, 3, n // 2, seed=400 + i)
        regs.append(dict(hm=hm, det_idx=idx, det_cf=cf, samples=ns,
                         steps=st, length=length + i, depth=depth + 3 * i,
                         angle_det=0.04, angle_rand=0.01 + 0.002 * i,
                         seed=0x7700 + i))
    return regs


def _oracle_cmpsit(oracle, r):
    want, _ = oracle.cmpsit(r["hm"], dict(det_idx=r["det_idx"],
                                          det_cf=r["det_cf"]),
                            r["samples"], r["steps"], r["length"],
                            r["depth"], r["angle_det"], r["angle_rand"],
                            r["seed"], want_state=False)
    return want


@pytest.mark.parametrize("mode,tile", [(nat.BATCH_PLAIN_EACH, 12),
                                       (nat.BATCH_FUSED_EACH, 12),
                                       (nat.BATCH_FUSED_BATCHED, 11),
                                       (nat.BATCH_FUSED_BATCHED, 0)])
def test_batch_cmpsit_vs_oracle(oracle, mode, tile):
    regs = _cmpsit_regs([12, 9, 14, 11, 13], steps=[2, 3, 1, 2, 0],
                        samples=[2, 1, 3, 2, 2])
    print(f"batch_cmpsit: mode {mode} tile {tile}", flush=True)
    got, st = nat.batch_cmpsit(regs, mode, tile, 4)
    print("batch_cmpsit", st, flush=True)
    for r, g in zip(regs, got):
        assert np.array_equal(g, _oracle_cmpsit(oracle, r)), r["hm"].nqb


def test_batch_cmpsit_pipelined_rounds(oracle):
    """Registers at least as wide as the tile: rounds are planned on the
    host from the device draws while the previous round runs."""
    regs = _cmpsit_regs([12, 14, 13, 12], steps=[2, 1, 3, 2],
                        samples=[2, 3, 1, 1], depth=60)
    got, st = nat.batch_cmpsit(regs, nat.BATCH_FUSED_BATCHED, 12, 3)
    print("batch_cmpsit pipelined", st, flush=True)
    for r, g in zip(regs, got):
        assert np.array_equal(g, _oracle_cmpsit(oracle, r))
    assert st["launches"] < st["passes"]
    assert st["draw_seconds"] > 0


def test_batch_cmpsit_golden(golden):
    """The reference's own cmpsit output (tests/golden/cmpsit_n10.npz),
    reproduced by a batch that holds that register between two others."""
    g = golden("cmpsit_n10.npz")
    r = dict(hm=synth.hamiltonian(10, 200, seed=6),
             det_idx=np.array([0b0000111100], np.uint64),
             det_cf=np.array([1.0 + 0j]), samples=3, steps=2, length=40,
             depth=25, angle_det=0.05, angle_rand=0.01, seed=99)
    others = _cmpsit_regs([11, 12], steps=[1, 2], samples=[2, 1])
    for mode in (nat.BATCH_FUSED_EACH, nat.BATCH_FUSED_BATCHED):
        got, _ = nat.batch_cmpsit([others[0], r, others[1]], mode, 9, 2)
        assert np.array_equal(got[1], g["overlaps"])


def test_batch_qdrift_golden(golden):
    """The reference's own qdrift output (tests/golden/qdrift_n10.npz)
    through the batch API (device draws)."""
    g = golden("qdrift_n10.npz")
    r = dict(hm=synth.hamiltonian(10, 120, seed=5), det_idx=g["didx"],
             det_cf=g["dcf"], samples=4, depth=150, step_size=0.05,
             seed=0x1234)
    got, _ = nat.batch_qdrift([r], nat.BATCH_FUSED_EACH, 10, 3)
    assert np.array_equal(got[0], g["overlaps"])


def test_batch_cmpsit_rejects_empty_sample_set():
    r = _cmpsit_regs([10], steps=[1], samples=[1])[0]
    r["length"] = len(r["hm"].cf)
    with pytest.raises(RuntimeError, match="ph2_batch_cmpsit"):
        nat.batch_cmpsit([r], nat.BATCH_FUSED_EACH, 9, 2)


@pytest.mark.parametrize("kind", ["qdrift", "cmpsit"])
def test_batch_rejects_determinant_beyond_register(kind):
    """A determinant index of 2^n or more is refused (it would alias a
    lower index, or land in the next register of the shared buffer)."""
    r = _cmpsit_regs([10, 11], steps=[1, 1], samples=[1, 1])
    idx = np.array(r[1]["det_idx"], dtype=np.uint64)
    idx[1] = np.uint64(1 << 11)
    r[1]["det_idx"] = idx
    if kind == "qdrift":
        for x in r:
            x["step_size"] = 0.05
        with pytest.raises(RuntimeError, match="ph2_batch_qdrift"):
            nat.batch_qdrift(r, nat.BATCH_FUSED_EACH, 9, 2)
    else:
        with pytest.raises(RuntimeError, match="ph2_batch_cmpsit"):
            nat.batch_cmpsit(r, nat.BATCH_FUSED_EACH, 9, 2)


def _graph_stats():
    import ctypes as C
    b, r = C.c_uint64(0), C.c_uint64(0)
    nat.lib().ph2hip_round_graph_stats(C.byref(b), C.byref(r))
    return b.value, r.value


def test_round_graph_built_once_and_replayed(oracle):
    """Pipelined batched rounds go through one graph built once (or again
    only when a round needs more pass indices) and replayed every round with
    new node arguments; every replay gives the oracle's overlaps."""
    regs = _cmpsit_regs([12, 13, 14, 12], steps=[3, 3, 3, 3],
                        samples=[2, 2, 2, 2], depth=50)
    b0, r0 = _graph_stats()
    got, st = nat.batch_cmpsit(regs, nat.BATCH_FUSED_BATCHED, 12, 3)
    b1, r1 = _graph_stats()
    rounds = 2 * 3 * 2
    print(f"round graph: {b1 - b0} builds, {r1 - r0} replays", flush=True)
    assert r1 - r0 == rounds
    assert 1 <= b1 - b0 <= rounds // 3
    for r, g in zip(regs, got):
        assert np.array_equal(g, _oracle_cmpsit(oracle, r))
