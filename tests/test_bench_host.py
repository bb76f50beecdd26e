"""Host-side pieces of bench.py: kernel naming, the choice of the dominant
fused template from per-launch records, the roofline arithmetic, and the
one-GPU guard.  No GPU call."""

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402
from phase2_amd import _native  # noqa: E402


def test_fused_name():
    assert bench.fused_name(12) == "k_rot_fused<12, 4>"
    assert bench.fused_name(11) == "k_rot_fused<11, 3>"
    assert bench.fused_name(10, True) == "k_rot_fused_batch<10, 2, false>"
    assert bench.fused_name(12, True, True) == \
        "k_rot_fused_batch<12, 4, true>"
    assert bench.fused_name(12, half=True) == "k_rot_fused_half<12, 4>"


def test_auto_tile_mirrors_the_library_rule():
    # replayed registers (<= 20 qubits by default): tile 10, 11 from 19,
    # never above nqb - 2; wider: 12 from 17, 13 from 26
    assert [bench.auto_tile(0, n) for n in (10, 11, 12, 16, 18, 19, 20)] \
        == [9, 9, 10, 10, 10, 11, 11]
    assert [bench.auto_tile(0, n) for n in (21, 25, 26, 33)] == \
        [12, 12, 13, 13]
    assert bench.auto_tile(0, 16, replay_max_qb=0) == 11
    assert bench.auto_tile(12, 16) == 12


def _recs(rows):
    r = np.zeros(len(rows), dtype=_native.REC_DTYPE)
    for i, (kind, nqb, tb, nt, ms) in enumerate(rows):
        r[i]["kind"], r[i]["nqb"], r[i]["tile_bits"] = kind, nqb, tb
        r[i]["nterms"], r[i]["ms"], r[i]["amps"] = nt, ms, 1 << nqb
    return r


def test_dominant_fused_by_time_not_count():
    # three short tile-11 launches, two long tile-12 launches, one record of
    # another kind that must be ignored
    rec = _recs([(2, 33, 11, 10, 10.0)] * 3 + [(2, 33, 12, 20, 40.0)] * 2
                + [(0, 33, 0, 1, 500.0)])
    name, st, tiles = bench.dominant_fused(rec, {"fused": None})
    assert name == "k_rot_fused<12, 4>"
    assert tiles == {"k_rot_fused<11, 3>": 3, "k_rot_fused<12, 4>": 2}
    assert st["launches"] == 2 and st["ms"] == 80.0
    assert st["amps"] == 2 * 2.0 ** 33 and st["terms"] == 40
    rf = bench.roofline(name, st)
    # 32 B per amplitude, 40 ms per launch
    assert rf["algorithmic_bytes_per_launch"] == 32 * 2 ** 33
    assert rf["achieved"] == pytest.approx(32 * 2 ** 33 / 0.040 / 1e9)
    assert rf["frac"] == pytest.approx(rf["achieved"] / 8000.0)
    # 20 terms per launch, each 12 multiplies/adds per amplitude pair
    assert rf["terms_per_launch"] == 20
    ops = 12 * 20 * 2 ** 32
    assert rf["fp64"]["ops_per_launch"] == ops
    assert rf["fp64"]["frac"] == pytest.approx(ops / 0.040 / 39.3e12)
    assert rf["hbm_plus_fp64_frac"] == pytest.approx(
        rf["frac"] + rf["fp64"]["frac"])
    # batched records count work items, not rotations: no FP64 side
    rb = bench.roofline(name, st, rotations=False)
    assert "fp64" not in rb and rb["terms_per_launch"] is None


def test_half_tile_form_is_its_own_kernel():
    rec = _recs([(2, 30, 12, 10, 30.0)] * 2 + [(2, 30, 12, 10, 25.0)] * 3)
    rec["direct"][2:] = 7          # first, last and half-tile flags
    name, st, tiles = bench.dominant_fused(rec, {"fused": None})
    assert name == "k_rot_fused_half<12, 4>" and st["launches"] == 3
    assert tiles == {"k_rot_fused<12, 4>": 2, "k_rot_fused_half<12, 4>": 3}


def test_dominant_falls_back_without_records():
    agg = {"launches": 4, "ms": 8.0, "amps": 4 * 2.0 ** 20, "terms": 12}
    name, st, tiles = bench.dominant_fused(_recs([]), {"fused": agg})
    assert name is None and st is agg and tiles == {}


def test_dominant_prefers_plain_when_it_dominates():
    prof = {"records": _recs([(2, 20, 12, 5, 1.0)]),
            "plain_pair": {"launches": 9, "ms": 5.0, "amps": 9 * 2.0 ** 20,
                           "terms": 9},
            "plain_diag": {"launches": 0, "ms": 0.0, "amps": 0.0,
                           "terms": 0}}
    name, st, _ = bench._dominant(prof)
    assert name == "k_rot_pair" and st["ms"] == 5.0


def test_more_than_one_gpu_exits_with_message():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in (p.stdout + p.stderr)


def test_batched_kernel_forms():
    rec = _recs([(3, 20, 12, 64, 3.0)] * 2 + [(3, 20, 12, 64, 1.0)])
    rec["direct"][:2] = 4
    rec["amps"][:] = 3 << 30
    name, st, kernels = bench.dominant_fused(rec, {"batched": None},
                                             bench.REC_BATCH)
    assert name == "k_rot_fused_batch<12, 4, true>"
    assert st["amps"] == 2 * (3 << 30) and st["launches"] == 2
    assert kernels == {"k_rot_fused_batch<12, 4, false>": 1,
                       "k_rot_fused_batch<12, 4, true>": 2}
