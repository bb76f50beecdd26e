#!/usr/bin/env python3
"""Benchmark: Pauli amp-updates/s and % of HBM peak of phase2's state-vector
sweep on one MI355X, fused/batched against plain, with the reference CPU
backend timed beside it.

Workloads (synthetic, phase2_amd.synth; --workload):

trott_n33 (default, the headline)
    33 qubits (128 GiB complex128 register resident in HBM), 100,000 terms
    (33 Z, 528 ZZ, 99,438 four-index, identity) in the lexicographic order
    trott_init sorts them into, delta = 0.02, 64-determinant initial state.
    The sweep is applied in circ_cache blocks of 1024 consecutive terms
    through qreg_paulirot.  A step is one whole block; the timed steps cycle
    through 8 blocks spread evenly over the sweep (a full sweep is 98 blocks),
    because the blocks differ in how well they fuse (the first block, with
    all flip masks on qubits 0..11, fuses into one pass).
trott_n16
    1st-order Trotter at 16 qubits, 537 terms, delta 0.1; a step is one
    Trotter step (sweep + overlap), the whole trott_simul is timed.
trott2_n26
    2nd-order Trotter at 26 qubits, 20,352 terms, coefficient-matrix state
    (230,230 determinants); a step is one trott2 step (forward and reverse
    half-sweeps + coefficient-matrix overlap).
qdrift_batch
    512 qDRIFT registers of 16..26 qubits (107 GiB in one device buffer),
    depth 1000; a step is one sample of every register: draws made on the
    device before the first round, batched launches (one per pass index
    and form) through one graph built once and replayed every round.
cmpsit_n30
    composite integrator at 30 qubits, 60,000 terms, 2,000 deterministic +
    512 sampled rotations per half-step; a step is one cmpsit step.

value        whole-job amp-updates/s over the timed steps (inputs resident)
roofline     the dominant kernel: algorithmic bytes per launch = 32 B per
             amplitude swept (one read and one write of the register) divided
             by its mean launch duration (HIP events on the library stream)
plain        the plain one-term-per-pass kernel on terms of the same workload
cpu_baseline the reference CPU backend (oracle/_ref, built from the reference
             sources) under mpirun on this host, on a bounded sample
"""

from __future__ import annotations

import argparse
import json
import math
import os
import shutil
import struct
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("Pauli amp-updates/s + % HBM peak, 33-qb FP64 Trotter: "
          "fused/batched vs plain")
HBM_PEAK_GBS = 8000.0
BYTES_PER_AMP_PASS = 32
# FP64 multiplies and adds at full rate (MI355X: 78.6 TFLOP/s counts an FMA
# as two; a rotation uses no FMA, so one operation per lane per cycle)
FP64_PEAK_TOPS = 78.6 / 2
OPS_PER_PAIR = 12   # 8 multiplies and 4 adds per amplitude pair and term
WORKLOADS = ("trott_n33", "trott_n16", "trott2_n26", "qdrift_batch",
             "cmpsit_n30")
DEFAULT_STEPS = dict(trott_n33=8, trott_n16=64, trott2_n26=4,
                     qdrift_batch=16, cmpsit_n30=2)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# -- reference CPU baselines ---------------------------------------------------

def _mpi_runs(exe, args, ranks, timeout=240):
    """Run a reference program under mpirun with `ranks` ranks (then with
    one process if that fails); returns the last JSON line it printed."""
    mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
    cmds = []
    if ranks > 1 and Path(mpirun).exists():
        cmds.append([mpirun, "-launcher", "fork", "-n", str(ranks), str(exe)]
                    + args)
    cmds.append([str(exe)] + args)
    for cmd in cmds:
        try:
            r = subprocess.run(cmd, capture_output=True, text=True,
                               timeout=timeout)
        except subprocess.TimeoutExpired:
            log("cpu baseline: timeout", cmd[:6])
            continue
        lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
        if r.returncode == 0 and lines:
            return json.loads(lines[-1])
        log("cpu baseline: failed", r.returncode, r.stderr[-400:])
    return None


def cpu_rotations(nqb, xm, zm, phi, ranks, what):
    """Reference qreg_paulirot (oracle/_ref/ref_bench) on the rotations
    (xm, zm, phi) applied in order to a `nqb`-qubit register; falls back to
    the oracle port on one core when the reference build is absent."""
    n = len(phi)
    rec = np.zeros(n, dtype=[("x", "<u8"), ("z", "<u8"), ("phi", "<f8")])
    rec["x"], rec["z"], rec["phi"] = xm, zm, phi
    sample = f"{n} rotations ({what}) on a {nqb}-qubit register"
    exe = ROOT / "oracle" / "_ref" / "ref_bench"
    with tempfile.TemporaryDirectory() as td:
        f = Path(td) / "terms.bin"
        rec.tofile(f)
        d = _mpi_runs(exe, [str(f), str(nqb)], ranks) if exe.is_file() \
            else None
    if d is not None:
        return dict(value=d["amp_updates_per_s"], unit="amp-updates/s",
                    cores=d["ranks"], kind="reference",
                    sample=sample + f", {d['ranks']} MPI ranks x 1 thread, "
                    "reference qreg_paulirot incl. its per-flush "
                    "exchange/mix/add", seconds=d["seconds"])
    from oracle import pyoracle
    o = pyoracle.oracle()
    st = np.zeros(1 << nqb, dtype=np.complex128)
    st[0] = 1
    t0 = time.perf_counter()
    o.rotate(st, rec["x"], rec["z"], rec["phi"])
    dt = time.perf_counter() - t0
    return dict(value=n * 2.0 ** nqb / dt, unit="amp-updates/s", cores=1,
                kind="port", sample=sample + ", oracle port, 1 thread",
                seconds=dt)


def cpu_circ(algo, hm, det_idx, det_cf, ranks, what, steps=0, delta=0.0,
             samples=0, depth=0, seed=0):
    """Whole reference circuit run (oracle/_ref/ref_circ: trott_simul or
    qdrift_simul) under mpirun; None when the reference build is absent."""
    exe = ROOT / "oracle" / "_ref" / "ref_circ"
    if not exe.is_file():
        return None
    hdr = struct.pack("<IIQQQQQQd", algo, hm.nqb, len(hm), len(det_idx),
                      steps, samples, depth, seed, delta)
    t = np.zeros(len(hm), dtype=[("x", "<u8"), ("z", "<u8"), ("c", "<f8")])
    t["x"], t["z"], t["c"] = hm.xm, hm.zm, hm.cf
    d = np.zeros(len(det_idx), dtype=[("i", "<u8"), ("re", "<f8"),
                                      ("im", "<f8")])
    d["i"] = det_idx
    d["re"], d["im"] = np.real(det_cf), np.imag(det_cf)
    with tempfile.TemporaryDirectory() as td:
        f = Path(td) / "job.bin"
        with open(f, "wb") as fh:
            fh.write(hdr)
            fh.write(t.tobytes())
            fh.write(d.tobytes())
        r = _mpi_runs(exe, [str(f)], ranks)
    if r is None:
        return None
    return dict(value=r["amp_updates_per_s"], unit="amp-updates/s",
                cores=r["ranks"], kind="reference",
                sample=f"{what}, {r['ranks']} MPI ranks x 1 thread, whole "
                "reference circuit run", seconds=r["seconds"])


def cpu_ranks(args):
    if args.cpu_ranks > 0:
        return args.cpu_ranks
    ncpu = min(os.cpu_count() or 1, 16)
    return 1 << int(math.log2(ncpu))


# -- helpers -------------------------------------------------------------------

def lex(hm, nat):
    perm = nat.sort_lex(hm.xm, hm.zm)
    return hm.xm[perm], hm.zm[perm], hm.cf[perm]


def roofline(kernel, prof, rotations=True):
    """Roofline object of one profiled kernel kind.  With rotations (the
    records count rotations, not batched work items), also the FP64 side:
    algorithmic multiplies and adds per launch against the full-rate peak,
    and hbm_plus_fp64_frac = (HBM time at peak + FP64 time at peak) / the
    mean launch time, which is 1 for a pass that runs both resources at
    peak one after the other."""
    launches = max(1, prof["launches"])
    ms = prof["ms"] / launches
    nbytes = BYTES_PER_AMP_PASS * prof["amps"] / launches
    gbs = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {"bound": "hbm", "kernel": kernel, "achieved": gbs,
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
           "traffic": None, "algorithmic_bytes_per_launch": nbytes,
           "mean_launch_ms": ms, "launches": prof["launches"],
           "terms_per_launch": prof["terms"] / launches
           if "terms" in prof and rotations else None}
    if rotations and "terms" in prof and ms > 0:
        # every term of a pass touches every amplitude pair once
        ops = (OPS_PER_PAIR * (prof["amps"] / launches / 2)
               * (prof["terms"] / launches))
        tops = ops / (ms * 1e-3) / 1e12
        out["fp64"] = {"ops_per_launch": ops, "achieved": tops,
                       "peak": FP64_PEAK_TOPS,
                       "unit": "TFLOP/s (multiply/add, no FMA)",
                       "frac": tops / FP64_PEAK_TOPS}
        out["hbm_plus_fp64_frac"] = out["frac"] + out["fp64"]["frac"]
    return out


def traffic_for(kernel, nqb):
    """HBM bytes per launch measured by rocprofv3 (FETCH_SIZE, WRITE_SIZE,
    calibrated) for this kernel and width, from profiles/traffic.json."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.is_file():
        return None
    try:
        d = json.loads(f.read_text())
    except ValueError:
        return None
    e = d.get(f"{kernel}@n{nqb}")
    return None if e is None else e.get("bytes_per_launch")


def plain_rate(nat, nqb, xm, zm, phi):
    """The plain kernel on the given rotations on a fresh register."""
    reg = nat.Register(nqb, nat.PLAIN)
    reg.prep_multidet(np.array([1], np.uint64), np.array([1.0 + 0j]))
    reg.rotate(xm[:2], zm[:2], phi[:2])  # warm-up
    nat.synchronize()
    nat.prof_enable(True)
    nat.prof_read(reset=True)
    t0 = time.perf_counter()
    reg.rotate(xm, zm, phi)
    nat.synchronize()
    t1 = time.perf_counter()
    pp = nat.prof_read(reset=True)
    nat.prof_enable(False)
    reg.free()
    launches = pp["plain_pair"]["launches"] + pp["plain_diag"]["launches"]
    pms = (pp["plain_pair"]["ms"] + pp["plain_diag"]["ms"]) / max(1, launches)
    gbs = BYTES_PER_AMP_PASS * 2.0 ** nqb / (pms * 1e-3) / 1e9
    return dict(amp_updates_per_s=len(phi) * 2.0 ** nqb / (t1 - t0),
                kernel_ms_per_pass=pms, kernel_GBps=gbs,
                hbm_frac=gbs / HBM_PEAK_GBS, terms=len(phi),
                pair_launches=pp["plain_pair"]["launches"],
                diag_launches=pp["plain_diag"]["launches"])


def auto_tile(tile, nqb, replay_max_qb=20):
    """The library's tile width for a requested one (0: automatic), with
    graph replay of repeated sweeps on for registers of at most
    replay_max_qb qubits (the library's default)."""
    if tile:
        return tile
    if nqb <= replay_max_qb:
        return max(9, min(11 if nqb >= 19 else 10, nqb - 2))
    return max(9, 13 if nqb >= 26 else 12 if nqb > 16 else nqb - 5)


def fused_name(tile, batched=False, half=False):
    """Kernel name as rocprofv3 prints it (template arguments TB, D, and
    for the batched kernel the half-tile form)."""
    d = 4 if tile >= 12 else tile - 8
    if batched:
        return f"k_rot_fused_batch<{tile}, {d}, {'true' if half else 'false'}>"
    return f"k_rot_fused{'_half' if half else ''}<{tile}, {d}>"


REC_FUSED = 2   # launch-record kind of a single-register fused pass
REC_BATCH = 3   # launch-record kind of a batched fused pass
REC_HALF = 4    # launch-record flag: half-tile form


def fused_by_kernel(rec, kind=REC_FUSED):
    """Per kernel template (tile width, half-tile form or not), the
    fused-pass launches of a record array: launches, summed ms, amplitudes
    swept and terms (work items for batched launches)."""
    rec = rec[rec["kind"] == kind]
    half = (rec["direct"] & REC_HALF) != 0
    out = {}
    for t in np.unique(rec["tile_bits"]):
        for h in (False, True):
            r = rec[(rec["tile_bits"] == t) & (half == h)]
            if len(r):
                out[fused_name(int(t), kind == REC_BATCH, h)] = {
                    "launches": len(r), "ms": float(r["ms"].sum()),
                    "amps": float(np.sum(r["amps"].astype(float))),
                    "terms": int(r["nterms"].sum())}
    return out


def dominant_fused(rec, prof, kind=REC_FUSED):
    """(kernel name, stats) of the fused kernel template with the most
    launch time, and the launch count per template.  Falls back to the
    aggregate when the records are empty."""
    per = fused_by_kernel(rec, kind)
    if not per:
        return None, prof["fused" if kind == REC_FUSED else "batched"], {}
    k = max(per, key=lambda n: per[n]["ms"])
    return k, per[k], {n: v["launches"] for n, v in per.items()}


def spread(n, k):
    return np.unique(np.linspace(0, n - 1, min(n, k)).astype(int))


# -- workloads -----------------------------------------------------------------

def bench_trott_n33(args, nat, costmodel, schedule):
    from phase2_amd import synth
    nqb = args.nqb
    log(f"bench: generating workload nqb={nqb}")
    if nqb == 33:
        w = synth.trott_n33()
    else:
        # Same generator at a smaller width (for quick runs).
        hm = synth.hamiltonian(nqb, 0, seed=0, decay=True,
                               n_total=min(100_000, nqb + nqb * (nqb - 1) // 2
                                           + 1 + 20 * nqb * nqb))
        idx, cf = synth.multidet_random(nqb, 64, nqb // 2, seed=0)
        w = dict(hm=hm, det_idx=idx, det_cf=cf, delta=0.02, steps=8)
    xm, zm, cf = lex(w["hm"], nat)
    phi = w["delta"] * cf
    nterms = len(phi)
    nblocks = (nterms + 1023) // 1024
    S = 8
    blocks = sorted({min(nblocks - 1, int((i + 0.5) * nblocks / S))
                     for i in range(S)})

    def block(b):
        s = slice(b * 1024, min(nterms, (b + 1) * 1024))
        return xm[s], zm[s], phi[s]

    namp = 2.0 ** nqb
    log(f"bench: allocating {nqb}-qubit register ({16 * namp / 2**30:.0f} GiB)")
    reg = nat.Register(nqb, nat.FUSED, args.tile, args.low)
    reg.prep_multidet(w["det_idx"], w["det_cf"])

    # Warm-up.  With the learned schedule every timed block is visited once
    # first, so its schedule is computed and cached as in steps 2.. of a
    # Trotter run.  These first visits (scheduling plus one sweep of each
    # block) are timed apart from the steps, as first_visit_seconds.
    sched_s = 0.0
    nwarm = max(args.warmup, len(blocks) if schedule == "learned" else 0)
    for k in range(nwarm):
        ts = time.perf_counter()
        reg.rotate(*block(blocks[k % len(blocks)]))
        nat.synchronize()
        if schedule == "learned" and k < len(blocks):
            sched_s += time.perf_counter() - ts
        log(f"bench: warmup {k + 1}/{nwarm}")
    nat.synchronize()

    nat.prof_enable(True)
    nat.prof_read(reset=True)
    nat.prof_records(reset=True)
    nupd = 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        b = blocks[k % len(blocks)]
        bx = block(b)
        reg.rotate(*bx)
        nupd += len(bx[0]) * namp
        log(f"bench: step {k + 1}/{args.steps} (block {b})")
    nat.synchronize()
    elapsed = time.perf_counter() - t0
    prof = nat.prof_read(reset=True)
    rec = nat.prof_records(reset=True)
    nat.prof_enable(False)
    norm2 = reg.norm2()
    reg.free()
    value = nupd / elapsed
    fz = prof["fused"]
    kern, kp, tiles = dominant_fused(rec, prof)
    plain = None
    if not args.no_plain:
        # terms spread evenly through each timed block (the per-pass cost of
        # the plain kernel does not depend on the order)
        per = max(1, args.plain_terms // len(blocks))
        sel = np.concatenate([b * 1024 + spread(len(block(b)[0]), per)
                              for b in blocks])
        plain = plain_rate(nat, nqb, xm[sel], zm[sel], phi[sel])
    return dict(
        value=value, elapsed=elapsed, norm2=norm2, plain=plain,
        roofline=roofline(kern, kp),
        config={"workload": f"trott1_n{nqb}_lex_sweep_blocks", "nqb": nqb,
                "terms": nterms, "terms_per_step": 1024,
                "blocks_timed": blocks, "blocks_in_sweep": nblocks,
                "delta": w["delta"]},
        extra={"terms_per_pass": fz["terms"] / max(1, fz["launches"]),
               "launches_per_kernel": tiles,
               "first_visit_seconds": sched_s if schedule == "learned"
               else None},
        cpu=lambda ranks: cpu_n33_proxy(args, nat, ranks),
        nqb=nqb)


def cpu_n33_proxy(args, nat, ranks):
    """A 33-qubit register does not fit twice in host memory: the reference
    runs `cpu_terms` lex-ordered terms of the same generator (from the middle
    of the sweep) at `cpu_nqb` qubits."""
    from phase2_amd import synth
    n = args.cpu_nqb
    hm = synth.hamiltonian(n, 0, seed=0, decay=True,
                           n_total=n + n * (n - 1) // 2 + 1 + 4000)
    xm, zm, cf = lex(hm, nat)
    s = slice(len(cf) // 2, len(cf) // 2 + args.cpu_terms)
    return cpu_rotations(n, xm[s], zm[s], 0.02 * cf[s], ranks,
                         "lex-ordered terms of the same generator")


def _timed_driver(nat, run, steps, warmup):
    """run(steps) once with `warmup` steps untimed, then timed with
    profiling on; returns (elapsed, prof)."""
    if warmup:
        run(warmup)
    nat.synchronize()
    nat.prof_enable(True)
    nat.prof_read(reset=True)
    nat.prof_records(reset=True)
    t0 = time.perf_counter()
    run(steps)
    nat.synchronize()
    elapsed = time.perf_counter() - t0
    prof = nat.prof_read(reset=True)
    prof["records"] = nat.prof_records(reset=True)
    nat.prof_enable(False)
    return elapsed, prof


def _dominant(prof):
    """The kernel with most launch time: a fused template (per tile width,
    from the launch records) or the plain pair/diagonal kernel."""
    kern, kp, tiles = dominant_fused(prof["records"], prof)
    best = (kern, kp) if kern else (None, {"ms": -1.0})
    for k, name in (("plain_pair", "k_rot_pair"), ("plain_diag", "k_rot_diag")):
        if prof[k]["ms"] > best[1]["ms"]:
            best = (name, prof[k])
    return best[0], best[1], tiles


def bench_trott_n16(args, nat, costmodel, schedule):
    from phase2_amd import synth
    w = synth.trott_n16()
    hm = w["hm"]
    prep = dict(det_idx=w["det_idx"], det_cf=w["det_cf"])
    nat.set_default(nat.FUSED, args.tile, args.low)

    def run(k):
        return nat.trott(1, hm, prep, w["delta"], k, want_state=False)

    # profiling adds an event pair per launch, which matters for microsecond
    # launches: time without it, then profile a separate run
    if args.warmup:
        run(args.warmup)
    nat.synchronize()
    t0 = time.perf_counter()
    out, _ = run(args.steps)
    nat.synchronize()
    elapsed = time.perf_counter() - t0
    # the same run with the graph replay of repeated sweeps turned off
    import ctypes
    L = nat.lib()
    L.ph2hip_set_replay.restype = ctypes.c_int
    L.ph2hip_set_replay(ctypes.c_uint32(0))
    t0 = time.perf_counter()
    run(args.steps)
    nat.synchronize()
    no_replay = time.perf_counter() - t0
    L.ph2hip_set_replay(ctypes.c_uint32(20))
    _, prof = _timed_driver(nat, run, args.steps, 0)
    kern, kp, tiles = _dominant(prof)
    nat.set_default(nat.PLAIN)
    t1 = time.perf_counter()
    run(args.steps)
    nat.synchronize()
    tp = time.perf_counter() - t1
    nat.set_default(nat.FUSED, args.tile, args.low)
    nupd = args.steps * len(hm) * 2.0 ** hm.nqb
    xm, zm, cf = lex(hm, nat)
    return dict(
        value=nupd / elapsed, elapsed=elapsed, norm2=None,
        plain={"amp_updates_per_s": nupd / tp, "whole_run_seconds": tp},
        roofline=roofline(kern, kp),
        config={"workload": "trott1_n16_full", "nqb": 16, "terms": len(hm),
                "steps": args.steps, "delta": w["delta"]},
        extra={"overlap_last": [float(out[-1].real), float(out[-1].imag)],
               "launches_per_kernel": tiles,
               "ms_per_step_replay": 1e3 * elapsed / args.steps,
               "ms_per_step_no_replay": 1e3 * no_replay / args.steps},
        cpu=lambda ranks: cpu_circ(
            1, synth.trott_n16()["hm"], w["det_idx"], w["det_cf"], ranks,
            f"trott n=16, {args.steps} steps of {len(hm)} terms",
            steps=args.steps, delta=w["delta"]),
        nqb=16)


def bench_trott2_n26(args, nat, costmodel, schedule):
    from phase2_amd import synth
    w = synth.trott2_n26()
    hm = w["hm"]
    prep = dict(C_alpha=w["C_alpha"], n_sites=w["n_sites"],
                n_alpha=w["n_alpha"], n_beta=w["n_beta"],
                closed_shell=w["closed_shell"], tapered=w["tapered"])
    nat.set_default(nat.FUSED, args.tile, args.low)

    def run(k):
        return nat.trott(2, hm, prep, w["delta"], k, want_state=False)

    elapsed, prof = _timed_driver(nat, run, args.steps, args.warmup)
    kern, kp, tiles = _dominant(prof)
    nupd = args.steps * 2 * len(hm) * 2.0 ** hm.nqb
    xm, zm, cf = lex(hm, nat)
    sel = spread(len(cf), args.plain_terms)
    plain = None if args.no_plain else plain_rate(
        nat, hm.nqb, xm[sel], zm[sel], 0.5 * w["delta"] * cf[sel])
    return dict(
        value=nupd / elapsed, elapsed=elapsed, norm2=None, plain=plain,
        roofline=roofline(kern, kp),
        config={"workload": "trott2_n26_coeff_matrix", "nqb": 26,
                "terms": len(hm), "steps": args.steps, "delta": w["delta"],
                "determinants": 230230},
        extra={"terms_per_pass": prof["fused"]["terms"]
               / max(1, prof["fused"]["launches"]),
               "launches_per_kernel": tiles},
        cpu=lambda ranks: cpu_rotations(
            26, xm[:256], zm[:256], 0.5 * w["delta"] * cf[:256], ranks,
            "first 256 rotations of a trott2 forward half-sweep"),
        nqb=26)


def bench_cmpsit_n30(args, nat, costmodel, schedule):
    from phase2_amd import synth
    w = synth.cmpsit_n30()
    hm = w["hm"]
    prep = dict(det_idx=w["det_idx"], det_cf=w["det_cf"])
    nat.set_default(nat.FUSED, args.tile, args.low)

    def run(k):
        return nat.cmpsit(hm, prep, 1, k, w["length"], w["depth"],
                          w["angle_det"], w["angle_rand"], w["seed"],
                          want_state=False)

    elapsed, prof = _timed_driver(nat, run, args.steps, args.warmup)
    kern, kp, tiles = _dominant(prof)
    rot_per_step = 2 * (w["length"] + w["depth"])
    nupd = args.steps * rot_per_step * 2.0 ** hm.nqb
    det = np.argsort(-np.abs(hm.cf), kind="stable")[:w["length"]]
    dx, dz, dc = hm.xm[det], hm.zm[det], hm.cf[det]
    perm = nat.sort_lex(dx, dz)
    dx, dz, dc = dx[perm], dz[perm], dc[perm]
    sel = spread(len(dc), args.plain_terms)
    phi = 0.5 * w["angle_det"] * dc
    plain = None if args.no_plain else plain_rate(
        nat, hm.nqb, dx[sel], dz[sel], phi[sel])
    return dict(
        value=nupd / elapsed, elapsed=elapsed, norm2=None, plain=plain,
        roofline=roofline(kern, kp),
        config={"workload": "cmpsit_n30", "nqb": 30, "terms": len(hm),
                "length": w["length"], "depth": w["depth"],
                "steps": args.steps, "rotations_per_step": rot_per_step,
                "angle_det": w["angle_det"], "angle_rand": w["angle_rand"]},
        extra={"terms_per_pass": prof["fused"]["terms"]
               / max(1, prof["fused"]["launches"]),
               "launches_per_kernel": tiles},
        cpu=lambda ranks: cpu_rotations(
            30, dx[:16], dz[:16], phi[:16], ranks,
            "first 16 deterministic rotations of a cmpsit half-step"),
        nqb=30)


def bench_qdrift_batch(args, nat, costmodel, schedule):
    from phase2_amd import synth
    regs = synth.qdrift_batch(args.regs, samples=args.steps)
    # warm-up: a few registers, one sample
    warm = synth.qdrift_batch(16, seed=1, samples=1)
    if args.warmup:
        nat.batch_qdrift(warm, nat.BATCH_FUSED_BATCHED, args.tile, args.low)
    nat.synchronize()
    import ctypes
    gb, gr = ctypes.c_uint64(0), ctypes.c_uint64(0)
    nat.lib().ph2hip_round_graph_stats(ctypes.byref(gb), ctypes.byref(gr))
    g0 = (gb.value, gr.value)
    t0 = time.perf_counter()
    _, st = nat.batch_qdrift(regs, nat.BATCH_FUSED_BATCHED, 0, args.low)
    nat.synchronize()
    elapsed = time.perf_counter() - t0
    nat.lib().ph2hip_round_graph_stats(ctypes.byref(gb), ctypes.byref(gr))
    # kernel timing on a subset, without graph capture (profiling brackets
    # every launch with events)
    sub = [dict(r, samples=1) for r in regs[:64]]
    nat.prof_enable(True)
    nat.prof_read(reset=True)
    nat.prof_records(reset=True)
    nat.batch_qdrift(sub, nat.BATCH_FUSED_BATCHED, st["tile_bits"],
                     st["low_bits"])
    prof = nat.prof_read(reset=True)
    rec = nat.prof_records(reset=True)
    nat.prof_enable(False)
    kern, kp, kernels = dominant_fused(rec, prof, REC_BATCH)
    _, sp = nat.batch_qdrift(sub, nat.BATCH_PLAIN_EACH, 12, args.low)
    _, sf = nat.batch_qdrift(sub, nat.BATCH_FUSED_EACH, st["tile_bits"],
                             st["low_bits"])
    nsamp = sum(r["samples"] for r in regs)
    widths = [r["hm"].nqb for r in regs]
    return dict(
        value=st["amp_updates"] / elapsed, elapsed=elapsed, norm2=None,
        plain={"amp_updates_per_s": sp["amp_updates"] / sp["seconds"],
               "subset": "first 64 registers, 1 sample, one launch per "
               "rotation", "launches": sp["launches"]},
        roofline=roofline(kern or fused_name(st["tile_bits"], True), kp,
                          rotations=False),
        config={"workload": "qdrift_batch_512", "registers": len(regs),
                "widths": [int(min(widths)), int(max(widths))],
                "samples": args.steps, "depth": 1000,
                "buffer_GiB": sum(2.0 ** n for n in widths) * 16 / 2 ** 30},
        extra={"samples_per_s": nsamp / elapsed,
               "launches": st["launches"], "passes": st["passes"],
               "tile_bits": st["tile_bits"], "low_bits": st["low_bits"],
               "tune_seconds": st["tune_seconds"],
               "device_draw_seconds": st["draw_seconds"],
               "round_graph_builds": gb.value - g0[0],
               "round_graph_replays": gr.value - g0[1],
               "launches_per_kernel": kernels,
               "fused_each_amp_updates_per_s":
                   sf["amp_updates"] / sf["seconds"]},
        cpu=lambda ranks: cpu_qdrift_sample(regs, ranks),
        nqb=None)


def cpu_qdrift_sample(regs, ranks):
    """Reference qdrift_simul on the first registers of 20..22 qubits of
    the batch, one sample each."""
    pick = [r for r in regs if 20 <= r["hm"].nqb <= 22][:6]
    tot_upd = tot_s = 0.0
    res = None
    for r in pick:
        res = cpu_circ(3, r["hm"], r["det_idx"], r["det_cf"], ranks, "",
                       samples=1, depth=r["depth"], delta=r["step_size"],
                       seed=r["seed"])
        if res is None:
            return None
        tot_upd += res["value"] * res["seconds"]
        tot_s += res["seconds"]
    if not pick or tot_s == 0:
        return None
    widths = ",".join(str(r["hm"].nqb) for r in pick)
    return dict(value=tot_upd / tot_s, unit="amp-updates/s",
                cores=res["cores"], kind="reference",
                sample=f"qdrift_simul, 1 sample x 1000 draws on {len(pick)} "
                f"registers of the batch (widths {widths}), "
                f"{res['cores']} MPI ranks x 1 thread each",
                seconds=tot_s)


BENCH = dict(trott_n33=bench_trott_n33, trott_n16=bench_trott_n16,
             trott2_n26=bench_trott2_n26, qdrift_batch=bench_qdrift_batch,
             cmpsit_n30=bench_cmpsit_n30)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=WORKLOADS, default="trott_n33")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nqb", type=int, default=33,
                    help="register width of trott_n33 (smaller for quick runs)")
    ap.add_argument("--regs", type=int, default=512,
                    help="registers of qdrift_batch")
    ap.add_argument("--tile", type=int, default=0,
                    help="fused tile bits (0: the library's automatic width)")
    ap.add_argument("--low", type=int, default=3)
    ap.add_argument("--plain-terms", type=int, default=24)
    ap.add_argument("--schedule", choices=("auto", "greedy", "learned"),
                    default="auto",
                    help="fusion schedule; auto = learned when the trained "
                    "parameters (phase2_amd/costmodel_params.npz) exist")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-baseline", dest="cpu", action="store_true",
                    default=True)
    ap.add_argument("--no-cpu-baseline", dest="cpu", action="store_false")
    ap.add_argument("--cpu-nqb", type=int, default=28)
    ap.add_argument("--cpu-terms", type=int, default=64)
    ap.add_argument("--cpu-ranks", type=int, default=0)
    args = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, "
              f"WORLD_SIZE={world}): the register lives on a single MI355X "
              "and the library has no multi-GPU path.",
              file=sys.stderr, flush=True)
        sys.exit(2)
    if args.steps is None:
        args.steps = DEFAULT_STEPS[args.workload]

    # the library logs INFO lines to stdout (as the reference does); keep
    # stdout to the one JSON line
    os.environ.setdefault("PHASE2_LOG", "WARN")
    from phase2_amd import _native as nat
    from phase2_amd import costmodel
    nat.init()
    schedule = args.schedule
    if schedule == "auto":
        schedule = "learned" if costmodel.PARAMS_PATH.is_file() else "greedy"
        if schedule == "learned":
            try:
                costmodel.load()
            except ValueError as e:   # trained for other features
                log(f"bench: {e}; greedy schedule")
                schedule = "greedy"
    if schedule == "learned":
        params, _ = costmodel.load()
        costmodel.install(params, min_qb=26)

    r = BENCH[args.workload](args, nat, costmodel, schedule)
    if schedule == "learned":
        costmodel.uninstall()
    roof = r["roofline"]
    if r["nqb"] is not None:
        roof["traffic"] = traffic_for(roof["kernel"], r["nqb"])

    cpu = None
    if args.cpu:
        ranks = cpu_ranks(args)
        log(f"bench: CPU baseline, {ranks} ranks")
        cpu = r["cpu"](ranks)

    value = r["value"]
    plain = r["plain"]
    cfg = dict(r["config"])
    cfg.update(schedule=schedule, low_bits=args.low,
               tile_bits=auto_tile(args.tile, r["nqb"]) if r["nqb"]
               else args.tile)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": cfg,
        "roofline": roof,
        "fused": dict(amp_updates_per_s=value,
                      effective_GBps=BYTES_PER_AMP_PASS * value / 1e9,
                      **r["extra"]),
        "plain": plain,
        "fused_vs_plain": (value / plain["amp_updates_per_s"])
        if plain else None,
        "norm_drift": abs(1.0 - r["norm2"]) if r["norm2"] is not None
        else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
