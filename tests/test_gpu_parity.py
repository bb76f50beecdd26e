"""HIP backend against the reference's golden vectors and the CPU oracle.

Bit-exact (``array_equal`` on the float64 view; -0.0 == +0.0) for every
rotation, sweep, multidet overlap and scatter index/value; the
coefficient-matrix inner product reduces in a tree on the device, so its
overlaps are compared at 1e-12 relative."""

import math

import numpy as np
import pytest

from oracle import pyoracle as po
from phase2_amd import _native as nat
from phase2_amd import synth

pytestmark = pytest.mark.gpu

REL = 1e-12
MODES = [(nat.PLAIN, 12, 6), (nat.FUSED, 12, 6), (nat.FUSED, 9, 0),
         (nat.FUSED, 13, 8), (nat.FUSED, 10, 4)]


def close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b)) <= rel * max(1.0, np.max(np.abs(b)))


@pytest.fixture(scope="module", autouse=True)
def gpu():
    print("gpu: init", flush=True)
    nat.init()
    yield
    nat.set_default(nat.FUSED, 12, 6)


@pytest.mark.parametrize("mode", MODES)
def test_rotate_n10_golden(golden, mode):
    g = golden("rotate_n10.npz")
    out = nat.rotate(g["state_in"], g["xm"], g["zm"], g["phi"], *mode)
    assert np.array_equal(out, g["state_out"])


@pytest.mark.parametrize("mode", MODES)
def test_rotate_n13_golden(golden, mode):
    g = golden("rotate_n13.npz")
    out = nat.rotate(g["state_in"], g["xm"], g["zm"], g["phi"], *mode)
    assert po.state_hash(out) == str(g["hash"])


@pytest.mark.parametrize("mode", MODES[:2] + [(nat.FUSED, 0, 4)])
def test_trott_n16_golden(golden, mode):
    """The full 537-term, 64-step 16-qubit run, including the automatic
    tile width (0: tile 10 at 16 qubits)."""
    g = golden("trott_n16.npz")
    w = synth.trott_n16()
    nat.set_default(*mode)
    ov, st = nat.trott(1, w["hm"], dict(det_idx=w["det_idx"],
                                        det_cf=w["det_cf"]),
                       w["delta"], w["steps"])
    assert np.array_equal(ov, g["overlaps"])
    assert po.state_hash(st) == str(g["hash"])


def test_trott2_coeff_golden(golden):
    g = golden("trott2_coeff_n12.npz")
    hm = synth.hamiltonian(12, 150, seed=3, decay=True)
    prep = dict(C_alpha=g["Ca"], n_sites=12, n_alpha=3, n_beta=0,
                closed_shell=1, tapered=0)
    nat.set_default(nat.FUSED, 10, 4)
    ov, st = nat.trott(2, hm, prep, 0.05, 6)
    assert close(ov, g["overlaps"])
    assert po.state_hash(st) == str(g["hash"])


def test_trott_open_tapered_golden(golden):
    g = golden("trott_open_n12.npz")
    hm = synth.hamiltonian(12, 120, seed=4)
    nat.set_default(nat.FUSED, 11, 5)
    ov, st = nat.trott(1, hm, dict(C_alpha=g["Ca"], C_beta=g["Cb"],
                                   n_sites=6, n_alpha=2, n_beta=3,
                                   closed_shell=0, tapered=0), 0.1, 5)
    assert close(ov, g["overlaps"])
    assert po.state_hash(st) == str(g["hash"])
    ov, st = nat.trott(1, hm, dict(C_alpha=g["Ca7"], C_beta=g["Cb7"],
                                   n_sites=7, n_alpha=2, n_beta=2,
                                   closed_shell=0, tapered=1), 0.1, 5)
    assert close(ov, g["overlaps_tap"])
    assert po.state_hash(st) == str(g["hash_tap"])


def test_qdrift_golden(golden):
    g = golden("qdrift_n10.npz")
    hm = synth.hamiltonian(10, 120, seed=5)
    nat.set_default(nat.FUSED, 10, 3)
    ov, st = nat.qdrift(hm, dict(det_idx=g["didx"], det_cf=g["dcf"]), 4,
                        150, 0.05, 0x1234)
    assert np.array_equal(ov, g["overlaps"])
    assert po.state_hash(st) == str(g["hash"])


def test_cmpsit_golden(golden):
    g = golden("cmpsit_n10.npz")
    hm = synth.hamiltonian(10, 200, seed=6)
    nat.set_default(nat.FUSED, 9, 2)
    ov, st = nat.cmpsit(hm, dict(det_idx=np.array([0b0000111100],
                                                  dtype=np.uint64),
                                 det_cf=np.array([1.0 + 0j])),
                        3, 2, 40, 25, 0.05, 0.01, 99)
    assert np.array_equal(ov, g["overlaps"])
    assert po.state_hash(st) == str(g["hash"])


def test_expand_inner_golden(golden):
    g = golden("expand.npz")
    for c in range(int(g["ncases"])):
        nqb, ns, na, nb, cl, tp = (int(v) for v in g[f"case{c}"])
        Ca, Cb = g[f"Ca{c}"], g[f"Cb{c}"]
        print(f"expand case {c}: nqb={nqb} n_sites={ns} na={na} nb={nb} "
              f"closed={cl} tapered={tp}", flush=True)
        st = nat.expand(nqb, ns, na, nb, cl, tp, Ca, None if cl else Cb)
        assert po.state_hash(st) == str(g[f"hash{c}"])
        r = np.random.default_rng(1000 + c)
        psi = r.normal(size=1 << nqb) + 1j * r.normal(size=1 << nqb)
        z = nat.inner(psi, ns, na, nb, cl, tp, Ca, None if cl else Cb)
        assert close(z, g[f"inner{c}"][0])


def test_expand_rejects_out_of_register():
    C = synth.coeff_matrix(8, 2, seed=1)
    with pytest.raises(RuntimeError):
        nat.expand(6, 8, 2, 0, 1, 0, C)  # orbital 7 is beyond 6 qubits


# -- phase2.run: the reference's own Python tests, against this backend -----

def test_phase2_run_golden(golden):
    import phase2_amd
    g = golden("phase2_run.npz")
    runs = [(["Z0"], [1.0], 0.3, "0"),
            (["X0 Y1", "Z1 Z2"], [0.3, -0.5], 0.2, "010"),
            ([f"X{k}" for k in range(10)], [0.1 * (k + 1) for k in range(10)],
             0.05, "0" * 10),
            (list(g["terms8"]), list(g["cf8"]), 0.3, "01101001")]
    for (p, c, d, s), want in zip(runs, g["out"]):
        assert phase2_amd.run(p, c, d, s) == want


def test_phase2_run_closed_forms():
    import phase2_amd as phase2
    TOL = 1e-12
    r = phase2.run(["Z0"], [1.0], 0.3, "0")
    assert isinstance(r, complex)
    assert abs(r - complex(math.cos(0.3), math.sin(0.3))) < TOL
    r = phase2.run(["Z0"], [1.0], 0.7, "1")
    assert abs(r - complex(math.cos(0.7), -math.sin(0.7))) < TOL
    assert abs(phase2.run(["X0"], [1.0], 0.5, "0") - math.cos(0.5)) < TOL
    assert abs(phase2.run([], [], 1.0, "00") - 1.0) < TOL
    assert abs(phase2.run(["X0", "X0"], [1.0, 1.0], 0.4, "0").real -
               math.cos(0.8)) < TOL
    r = phase2.run(["Z0 Z1"], [1.0], 0.8, "00")
    assert abs(r - complex(math.cos(0.8), math.sin(0.8))) < TOL
    h = [0.1 * (k + 1) for k in range(10)]
    r = phase2.run([f"X{k}" for k in range(10)], h, 0.05, "0" * 10)
    assert abs(r - math.prod(math.cos(0.05 * v) for v in h)) < TOL
    psi = "0101010101"
    r = phase2.run([f"Z{k}" for k in range(10)], h, 0.1, psi)
    ph = sum((1 if psi[k] == "0" else -1) * 0.1 * h[k] for k in range(10))
    assert abs(r - complex(math.cos(ph), math.sin(ph))) < TOL
    assert abs(phase2.run(["Z0"], [1.0], math.pi, "0") + 1.0) < TOL
    args = (["X0", "Z1"], [0.3, -0.7], 0.25, "01")
    assert phase2.run(*args) == phase2.run(*args)


@pytest.mark.parametrize("bad", [(["Z0"], [1.0, 2.0], 0.1, "0"),
                                 (["W0"], [1.0], 0.1, "0"),
                                 (["X5"], [1.0], 0.1, "0"),
                                 (["Z0"], [1.0], 0.1, "2"),
                                 (["X"], [1.0], 0.1, "0"),
                                 (["Z0"], [1.0], 0.1, "0" * 31),
                                 (["Z0"], [1.0], 0.1, "")])
def test_phase2_run_errors(bad):
    import phase2_amd
    with pytest.raises(ValueError):
        phase2_amd.run(*bad)


# -- against the oracle on fresh seeds ----------------------------------------

@pytest.mark.parametrize("n,K,seed", [(9, 200, 1), (14, 400, 2), (17, 700, 3)])
def test_rotate_vs_oracle(oracle, n, K, seed):
    rng = np.random.default_rng(seed)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    xm = rng.integers(0, 1 << n, size=K, dtype=np.uint64)
    zm = rng.integers(0, 1 << n, size=K, dtype=np.uint64)
    xm[::5] = 0
    xm[3] = 1 << (n - 1)  # top qubit only
    phi = rng.normal(size=K)
    want = oracle.rotate(st, xm, zm, phi)
    for mode in MODES:
        if mode[1] > n:
            continue
        print(f"rotate n={n} K={K} mode={mode}", flush=True)
        got = nat.rotate(st, xm, zm, phi, *mode)
        assert np.array_equal(got, want), mode


def test_small_registers_plain_fallback(oracle):
    for n in (1, 2, 5, 8):
        rng = np.random.default_rng(n)
        st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        xm = rng.integers(0, 1 << n, size=30, dtype=np.uint64)
        zm = rng.integers(0, 1 << n, size=30, dtype=np.uint64)
        phi = rng.normal(size=30)
        got = nat.rotate(st, xm, zm, phi, nat.FUSED, 12, 6)
        assert np.array_equal(got, oracle.rotate(st, xm, zm, phi))


def test_empty_and_zero_angle():
    st = np.zeros(1 << 10, dtype=np.complex128)
    st[5] = 0.6 - 0.8j
    assert np.array_equal(nat.rotate(st, [], [], []), st)
    out = nat.rotate(st, [3, 0], [1, 7], [0.0, 0.0])
    assert np.array_equal(out, st)


def test_trott_sampled_vs_oracle(oracle):
    hm = synth.hamiltonian(15, 500, seed=21, decay=True)
    idx, cf = synth.multidet_random(15, 5, 6, seed=3)
    prep = dict(det_idx=idx, det_cf=cf)
    nat.set_default(nat.FUSED, 12, 6)
    ov, st = nat.trott(2, hm, prep, 0.07, 3)
    ov_o, st_o = oracle.trott(2, hm, prep, 0.07, 3)
    assert np.array_equal(ov, ov_o)
    assert po.state_hash(st) == po.state_hash(st_o)


def test_qdrift_cmpsit_vs_oracle(oracle):
    hm = synth.hamiltonian(13, 300, seed=8)
    prep = dict(det_idx=np.array([0b1111000], dtype=np.uint64),
                det_cf=np.array([1.0 + 0j]))
    nat.set_default(nat.FUSED, 11, 5)
    ov, st = nat.qdrift(hm, prep, 3, 400, 0.02, 77)
    ov_o, st_o = oracle.qdrift(hm, prep, 3, 400, 0.02, 77)
    assert np.array_equal(ov, ov_o)
    assert po.state_hash(st) == po.state_hash(st_o)
    ov, st = nat.cmpsit(hm, prep, 2, 2, 60, 100, 0.04, 0.02, 78)
    ov_o, st_o = oracle.cmpsit(hm, prep, 2, 2, 60, 100, 0.04, 0.02, 78)
    assert np.array_equal(ov, ov_o)
    assert po.state_hash(st) == po.state_hash(st_o)


# -- size-independent properties at large registers ----------------------------

def test_fused_equals_plain_n24():
    """A block of the n=26 workload on 24 qubits (masks with bits >= 24
    dropped): fused passes leave the state bit-identical to the plain sweep,
    and undoing the block restores it to 1e-12."""
    hm = synth.hamiltonian(24, 3000, seed=9, decay=True)
    rng = np.random.default_rng(5)
    n = 24
    st = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n))
    st /= np.linalg.norm(st)
    phi = 0.05 * hm.cf[:1024] / np.max(np.abs(hm.cf))
    xm, zm = hm.xm[:1024], hm.zm[:1024]
    print("n=24 plain", flush=True)
    a = nat.rotate(st, xm, zm, phi, nat.PLAIN, 12, 6)
    print("n=24 fused", flush=True)
    b = nat.rotate(st, xm, zm, phi, nat.FUSED, 12, 6)
    assert np.array_equal(a, b)
    c = nat.rotate(st, xm, zm, phi, nat.FUSED, 13, 7)
    assert np.array_equal(a, c)
    assert abs(np.linalg.norm(a) - 1.0) < 1e-12
    back = nat.rotate(a, xm[::-1], zm[::-1], -phi[::-1], nat.FUSED, 12, 6)
    assert np.max(np.abs(back - st)) < 1e-12


def test_fused_undo_n33():
    """At the headline width (33 qubits, amplitude indices beyond 2^32, the
    128 GiB register) one 1024-term block of the lexicographic sweep under
    the learned schedule (tile 12 and 13, half-tile and whole-tile passes),
    then the same block reversed with negated angles, returns the
    64-determinant state: its overlap with itself stays 1 to 1e-12 and the
    norm to 1e-12.  An indexing error anywhere in the passes would move the
    state far from that."""
    from phase2_amd import costmodel as cm
    w = synth.trott_n33()
    hm = w["hm"]
    perm = nat.sort_lex(hm.xm, hm.zm)
    k0 = 42 * 1024
    xm = hm.xm[perm][k0:k0 + 1024]
    zm = hm.zm[perm][k0:k0 + 1024]
    cfb = hm.cf[perm][k0:k0 + 1024]
    phi = 0.3 * cfb / np.max(np.abs(cfb))
    cf = w["det_cf"] / np.linalg.norm(w["det_cf"])
    params, _ = cm.load()
    nat.init()
    reg = nat.Register(33, nat.FUSED, 0, 3)
    try:
        cm.install(params, min_qb=30)
        nat.prof_enable(True)
        nat.prof_records(reset=True)
        reg.prep_multidet(w["det_idx"], cf)
        print("n=33 block forward", flush=True)
        reg.rotate(xm, zm, phi)
        mid = reg.overlap(w["det_idx"], cf)
        print("n=33 block backward", flush=True)
        reg.rotate(xm[::-1], zm[::-1], -phi[::-1])
        ov = reg.overlap(w["det_idx"], cf)
        n2 = reg.norm2()
        rec = nat.prof_records(reset=True)
    finally:
        nat.prof_enable(False)
        cm.uninstall()
        reg.free()
    fused = rec[rec["kind"] == 2]
    forms = {(int(t), bool(d & 4)) for t, d in zip(fused["tile_bits"],
                                                  fused["direct"])}
    print("n=33 overlap after the block", mid, "after undoing it", ov,
          "norm2", n2, "pass forms", sorted(forms), flush=True)
    assert abs(mid - 1.0) > 1e-6                 # the block did something
    assert abs(ov - 1.0) < 1e-12
    assert abs(n2 - 1.0) < 1e-12
    assert (13, True) in forms


@pytest.mark.parametrize("pipe_tile", [10, 11, 12])
def test_fused_pipelined_form_bit_identical(oracle, pipe_tile):
    """The persistent pipelined fused kernel and every plain-kernel variant
    leave the state bit-identical to the oracle."""
    n = 18
    rng = np.random.default_rng(40 + pipe_tile)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    hm = synth.hamiltonian(n, 900, seed=pipe_tile, decay=True)
    perm = oracle.sort_lex(hm.xm, hm.zm)
    xm, zm = hm.xm[perm], hm.zm[perm]
    phi = 0.3 * hm.cf[perm] / np.max(np.abs(hm.cf))
    want = oracle.rotate(st, xm, zm, phi)
    try:
        nat.set_fused_form(1 << pipe_tile, 64)
        got = nat.rotate(st, xm, zm, phi, nat.FUSED, pipe_tile, 4)
        assert np.array_equal(got, want)
        for form in (4, 8, 2 + 16, 4 + 16):
            nat.set_plain_form(form)
            got = nat.rotate(st, xm[:40], zm[:40], phi[:40], nat.PLAIN)
            assert np.array_equal(got, oracle.rotate(st, xm[:40], zm[:40],
                                                     phi[:40])), form
    finally:
        nat.set_fused_form(0, 0)
        nat.set_plain_form(2)


@pytest.mark.parametrize("tile,low", [(10, 4), (11, 5), (12, 4), (12, 3)])
def test_fused_direct_load_bit_identical(oracle, tile, low):
    """Tiles with at least 4 contiguous low bits are loaded straight into LDS
    (lanes permuted within 256-byte runs to match the LDS swizzle), the
    others through registers; with the direct form on, eligible first and
    last micro-groups move their cosets between HBM and registers instead;
    with the half-tile form on (tile 12), passes that have one run through
    half a tile of LDS.  Every combination leaves the state bit-identical to
    the oracle."""
    n = 18
    rng = np.random.default_rng(60 + tile + low)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    hm = synth.hamiltonian(n, 900, seed=tile * 10 + low, decay=True)
    perm = oracle.sort_lex(hm.xm, hm.zm)
    xm, zm = hm.xm[perm], hm.zm[perm]
    phi = 0.3 * hm.cf[perm] / np.max(np.abs(hm.cf))
    want = oracle.rotate(st, xm, zm, phi)
    for direct in (True, False):
        for half in (0, nat.HALF_TILES):
            try:
                nat.set_fused_direct(direct)
                nat.set_fused_half(half)
                got = nat.rotate(st, xm, zm, phi, nat.FUSED, tile, low)
            finally:
                nat.set_fused_direct(True)
                nat.set_fused_half(nat.HALF_TILES)
            assert np.array_equal(got, want), (direct, half)


@pytest.mark.parametrize("n,tile,low,seed", [
    (16, 12, 4, 1), (18, 12, 4, 2), (18, 12, 3, 3), (20, 12, 5, 4),
    (20, 12, 2, 5), (17, 13, 4, 6), (19, 13, 3, 7), (21, 13, 5, 8)])
def test_fused_half_tile_bit_identical(oracle, n, tile, low, seed):
    """Tile-12 and tile-13 passes in the half-tile form (exchanges between
    micro-groups in two rounds through half a tile of LDS, split by a tile
    bit neither group touches; first and last groups straight between HBM
    and registers, whatever their lane layout) leave the state
    bit-identical to the oracle, and the form is in fact taken (launch
    records, flag 4)."""
    rng = np.random.default_rng(700 + seed)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    hm = synth.hamiltonian(n, 1500, seed=seed, decay=True)
    perm = oracle.sort_lex(hm.xm, hm.zm)
    xm, zm = hm.xm[perm], hm.zm[perm]
    phi = 0.3 * hm.cf[perm] / np.max(np.abs(hm.cf))
    want = oracle.rotate(st, xm, zm, phi)
    try:
        nat.set_fused_half(1 << tile)
        nat.prof_enable(True)
        nat.prof_records(reset=True)
        got = nat.rotate(st, xm, zm, phi, nat.FUSED, tile, low)
        rec = nat.prof_records(reset=True)
    finally:
        nat.prof_enable(False)
        nat.set_fused_half(nat.HALF_TILES)
    fused = rec[rec["kind"] == 2]
    half = fused[(fused["direct"] & 4) != 0]
    multi = half[half["ngroups"] > 1]
    print("half-tile passes", len(half), "of", len(fused), "multi-group",
          len(multi), flush=True)
    assert len(multi) > 0
    assert np.array_equal(got, want)


def test_fused_half_mask_rejects_other_tiles():
    with pytest.raises(RuntimeError):
        nat.set_fused_half(1 << 11)
    nat.set_fused_half(nat.HALF_TILES)
    assert nat.HALF_TILES == (1 << 12) | (1 << 13)


def _batch_regs(widths, depth=120, samples=2):
    regs = []
    for i, n in enumerate(widths):
        hm = synth.hamiltonian(n, 40 + 7 * n, seed=100 + i)
        idx, cf = synth.multidet_random(n, 3, n // 2, seed=200 + i)
        regs.append(dict(hm=hm, det_idx=idx, det_cf=cf, samples=samples,
                         depth=depth, step_size=0.03, seed=0x9000 + i))
    return regs


@pytest.mark.parametrize("tile,half", [(12, 0), (12, nat.HALF_TILES),
                                       (13, nat.HALF_TILES)])
def test_batch_qdrift_half_tile_vs_oracle(oracle, tile, half):
    """Batched tile-12 and tile-13 launches with and without the half-tile
    form (the passes that have it in one launch, the others in a second per
    pass index) give exactly each register's own qdrift overlaps."""
    widths = [12, 14, 13, 15, 12, 16] if tile == 12 else [13, 14, 15, 16,
                                                          13, 14]
    regs = _batch_regs(widths, depth=160)
    try:
        nat.set_fused_half(half)
        nat.prof_enable(True)
        nat.prof_records(reset=True)
        got, st = nat.batch_qdrift(regs, nat.BATCH_FUSED_BATCHED, tile, 4)
        rec = nat.prof_records(reset=True)
    finally:
        nat.prof_enable(False)
        nat.set_fused_half(nat.HALF_TILES)
    rec = rec[rec["kind"] == 3]
    nhalf = int(np.sum((rec["direct"] & 4) != 0))
    print("batched launches", len(rec), "half-tile", nhalf, flush=True)
    assert (nhalf > 0) == bool(half)
    for r, g in zip(regs, got):
        want, _ = oracle.qdrift(r["hm"], dict(det_idx=r["det_idx"],
                                              det_cf=r["det_cf"]),
                                r["samples"], r["depth"], r["step_size"],
                                r["seed"], want_state=False)
        assert np.array_equal(g, want)


@pytest.mark.parametrize("mode,tile", [(nat.BATCH_PLAIN_EACH, 12),
                                       (nat.BATCH_FUSED_EACH, 12),
                                       (nat.BATCH_FUSED_BATCHED, 11),
                                       (nat.BATCH_FUSED_BATCHED, 0)])
def test_batch_qdrift_vs_oracle(oracle, mode, tile):
    """Every register of a batch of uneven widths gives exactly the
    overlaps of qdrift run on it alone (same seed)."""
    regs = _batch_regs([12, 9, 14, 11, 13, 12])
    got, st = nat.batch_qdrift(regs, mode, tile, 4)
    print("batch", mode, tile, st, flush=True)
    for r, g in zip(regs, got):
        want, _ = oracle.qdrift(r["hm"], dict(det_idx=r["det_idx"],
                                              det_cf=r["det_cf"]),
                                r["samples"], r["depth"], r["step_size"],
                                r["seed"], want_state=False)
        assert np.array_equal(g, want)
    if mode == nat.BATCH_FUSED_BATCHED:
        # one launch per pass index instead of one per pass per register
        assert st["launches"] < st["passes"]


@pytest.mark.parametrize("tile", [11, 0])
def test_batch_qdrift_pipelined_rounds(oracle, tile):
    """All registers at least as wide as the tile: the batched rounds are
    pipelined (round s+1 drawn and planned on the host while round s runs),
    with uneven sample counts so registers drop out of later rounds; every
    register still matches qdrift run on it alone."""
    regs = _batch_regs([12, 14, 13, 12, 15], depth=150, samples=1)
    for r, ns in zip(regs, (3, 1, 2, 3, 2)):
        r["samples"] = ns
    got, st = nat.batch_qdrift(regs, nat.BATCH_FUSED_BATCHED, tile, 4)
    print("pipelined", tile, st, flush=True)
    for r, g in zip(regs, got):
        want, _ = oracle.qdrift(r["hm"], dict(det_idx=r["det_idx"],
                                              det_cf=r["det_cf"]),
                                r["samples"], r["depth"], r["step_size"],
                                r["seed"], want_state=False)
        assert np.array_equal(g, want)
    assert st["launches"] < st["passes"]
