#!/usr/bin/env python3
"""Benchmark: Pauli amp-updates/s and % of HBM peak of the 33-qubit FP64
1st-order Trotter sweep on one MI355X, fused (LDS coset tiles) against plain
(one HBM pass per term), with the reference CPU backend timed beside it.

Workload (synthetic, phase2_amd.synth.trott_n33): 33 qubits (128 GiB
complex128 register resident in HBM), 100,000 terms (33 Z, 528 ZZ, 99,438
four-index, identity), lexicographic order as trott_init sorts it,
delta = 0.02, 64-determinant initial state.  The sweep is applied in
circ_cache blocks of 1024 consecutive terms through qreg_paulirot.  A step is
one whole block; the timed steps cycle through 8 blocks spread evenly over
the sweep (a full sweep is 98 blocks), because the blocks differ in how
well they fuse (the first block, with all flip masks on qubits 0..11, fuses
into one pass).

value        whole-job amp-updates/s of the fused sweep over the timed steps
roofline     the fused kernel (k_rot_fused): algorithmic bytes per launch =
             32 B x 2^n (one read and one write of the register) divided by
             its mean launch duration (HIP events on the library stream)
plain        the plain kernel on terms spread evenly through the same blocks
cpu_baseline the reference CPU backend (oracle/_ref/ref_bench, built from the
             reference sources) under mpirun on this host
"""

from __future__ import annotations

import argparse
import json
import math
import os
import shutil
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(nqb: int):
    from phase2_amd import synth
    if nqb == 33:
        w = synth.trott_n33()
    else:
        # Same generator at a smaller width (for quick runs).
        hm = synth.hamiltonian(nqb, 0, seed=0, decay=True,
                               n_total=min(100_000, nqb + nqb * (nqb - 1) // 2
                                           + 1 + 20 * nqb * nqb))
        idx, cf = synth.multidet_random(nqb, 64, nqb // 2, seed=0)
        w = dict(hm=hm, det_idx=idx, det_cf=cf, delta=0.02, steps=8)
    return w


def cpu_baseline(nqb: int, nterms: int, ranks: int):
    """Reference CPU backend under mpirun on a bounded sample: the first
    `nterms` terms (lex order) of the same generator at `nqb` qubits."""
    from phase2_amd import synth, _native
    exe = ROOT / "oracle" / "_ref" / "ref_bench"
    hm = synth.hamiltonian(nqb, 0, seed=0, decay=True,
                           n_total=nqb + nqb * (nqb - 1) // 2 + 1 + 4000)
    perm = _native.sort_lex(hm.xm, hm.zm)
    # skip the first block's low-qubit terms: take terms from the middle
    sel = perm[len(perm) // 2: len(perm) // 2 + nterms]
    rec = np.zeros(nterms, dtype=[("x", "<u8"), ("z", "<u8"), ("phi", "<f8")])
    rec["x"], rec["z"] = hm.xm[sel], hm.zm[sel]
    rec["phi"] = 0.02 * hm.cf[sel]
    sample = (f"{nterms} rotations (lex-ordered terms of the same generator) "
              f"on a {nqb}-qubit register")
    with tempfile.TemporaryDirectory() as td:
        f = Path(td) / "terms.bin"
        rec.tofile(f)
        if exe.is_file():
            mpirun = shutil.which("mpirun") or "/opt/conda/bin/mpirun"
            cmds = []
            if ranks > 1 and Path(mpirun).exists():
                cmds.append(([mpirun, "-launcher", "fork", "-n", str(ranks),
                              str(exe), str(f), str(nqb)], ranks))
            cmds.append(([str(exe), str(f), str(nqb)], 1))
            for cmd, nr in cmds:
                try:
                    r = subprocess.run(cmd, capture_output=True, text=True,
                                       timeout=240)
                except subprocess.TimeoutExpired:
                    log("cpu baseline: timeout", cmd[:4])
                    continue
                lines = [ln for ln in r.stdout.splitlines()
                         if ln.startswith("{")]
                if r.returncode == 0 and lines:
                    d = json.loads(lines[-1])
                    return dict(value=d["amp_updates_per_s"],
                                unit="amp-updates/s", cores=d["ranks"],
                                kind="reference",
                                sample=sample + f", {d['ranks']} MPI ranks x "
                                "1 thread, reference qreg_paulirot incl. its "
                                "per-flush exchange/mix/add",
                                seconds=d["seconds"])
                log("cpu baseline: failed", r.returncode, r.stderr[-400:])
        # the oracle port, one core
        from oracle import pyoracle
        o = pyoracle.oracle()
        st = np.zeros(1 << nqb, dtype=np.complex128)
        st[0] = 1
        t0 = time.perf_counter()
        o.rotate(st, rec["x"], rec["z"], rec["phi"])
        dt = time.perf_counter() - t0
        return dict(value=nterms * 2.0 ** nqb / dt, unit="amp-updates/s",
                    cores=1, kind="port", sample=sample + ", 1 thread",
                    seconds=dt)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nqb", type=int, default=33)
    ap.add_argument("--tile", type=int, default=12)
    ap.add_argument("--low", type=int, default=4)
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
    ap.add_argument("--cpu-terms", type=int, default=16)
    ap.add_argument("--cpu-ranks", type=int, default=0)
    args = ap.parse_args(argv)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, "
              f"WORLD_SIZE={world}): the register lives on a single MI355X "
              "and there is no multi-GPU path in this build.",
              file=sys.stderr, flush=True)
        sys.exit(2)

    from phase2_amd import _native as nat

    nqb = args.nqb
    log(f"bench: generating workload nqb={nqb}")
    w = workload(nqb)
    hm = w["hm"]
    perm = nat.sort_lex(hm.xm, hm.zm)
    xm, zm = hm.xm[perm], hm.zm[perm]
    phi = w["delta"] * hm.cf[perm]
    nterms = len(phi)
    nblocks = (nterms + 1023) // 1024
    S = 8
    blocks = sorted({min(nblocks - 1, int((i + 0.5) * nblocks / S))
                     for i in range(S)})

    def block(b):
        s = slice(b * 1024, min(nterms, (b + 1) * 1024))
        return xm[s], zm[s], phi[s]

    namp = 2.0 ** nqb
    nat.init()
    from phase2_amd import costmodel
    schedule = args.schedule
    if schedule == "auto":
        schedule = "learned" if costmodel.PARAMS_PATH.is_file() else "greedy"
    if schedule == "learned":
        params, _ = costmodel.load()
        costmodel.install(params, low=args.low, min_qb=min(nqb, 28))
    log(f"bench: allocating {nqb}-qubit register ({16 * namp / 2**30:.0f} GiB)")
    reg = nat.Register(nqb, nat.FUSED, args.tile, args.low)
    reg.prep_multidet(w["det_idx"], w["det_cf"])

    # Warm-up.  With the learned schedule every timed block is visited once
    # first, so its schedule is computed and cached as in steps 2.. of a
    # Trotter run; the one-off scheduling time is reported separately.
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
    nupd = 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        b = blocks[k % len(blocks)]
        bx = block(b)
        reg.rotate(*bx)
        nupd += len(bx[0]) * namp
        log(f"bench: step {k + 1}/{args.steps} (block {b})")
    nat.synchronize()
    t1 = time.perf_counter()
    prof = nat.prof_read(reset=True)
    elapsed = t1 - t0
    value = nupd / elapsed
    fz = prof["fused"]
    fused_ms = fz["ms"] / max(1, fz["launches"])
    fused_gbs = BYTES_PER_AMP_PASS * namp / (fused_ms * 1e-3) / 1e9
    terms_per_pass = fz["terms"] / max(1, fz["launches"])

    plain = None
    if not args.no_plain:
        reg.configure(nat.PLAIN)
        # terms spread evenly through each block (the per-pass cost of the
        # plain kernel does not depend on the order)
        per = max(1, args.plain_terms // len(blocks))
        nt = 0
        nat.synchronize()
        p0 = time.perf_counter()
        for b in blocks:
            bx, bz, bp = block(b)
            sel = np.linspace(0, len(bx) - 1, per).astype(int)
            reg.rotate(bx[sel], bz[sel], bp[sel])
            nt += len(sel)
        nat.synchronize()
        p1 = time.perf_counter()
        pp = nat.prof_read(reset=True)
        launches = pp["plain_pair"]["launches"] + pp["plain_diag"]["launches"]
        pms = (pp["plain_pair"]["ms"] + pp["plain_diag"]["ms"]) / max(1, launches)
        pgbs = BYTES_PER_AMP_PASS * namp / (pms * 1e-3) / 1e9
        plain = dict(amp_updates_per_s=nt * namp / (p1 - p0),
                     kernel_ms_per_pass=pms, kernel_GBps=pgbs,
                     hbm_frac=pgbs / HBM_PEAK_GBS, terms=nt,
                     pair_launches=pp["plain_pair"]["launches"],
                     diag_launches=pp["plain_diag"]["launches"])
        reg.configure(nat.FUSED, args.tile, args.low)
    nat.prof_enable(False)

    norm2 = reg.norm2()
    reg.free()
    if schedule == "learned":
        costmodel.uninstall()

    cpu = None
    if args.cpu:
        ranks = args.cpu_ranks
        if ranks <= 0:
            ncpu = min(os.cpu_count() or 1, 16)
            ranks = 1 << int(math.log2(ncpu))
        log(f"bench: CPU baseline, {ranks} ranks, nqb={args.cpu_nqb}")
        cpu = cpu_baseline(args.cpu_nqb, args.cpu_terms, ranks)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"trott1_n{nqb}_lex_sweep_blocks",
            "nqb": nqb,
            "terms": nterms,
            "terms_per_step": 1024,
            "blocks_timed": blocks,
            "blocks_in_sweep": nblocks,
            "delta": w["delta"],
            "schedule": schedule,
            "tile_bits": args.tile,
            "low_bits": args.low,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_rot_fused<{args.tile}>",
            "achieved": fused_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": fused_gbs / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": BYTES_PER_AMP_PASS * namp,
            "mean_launch_ms": fused_ms,
            "launches": fz["launches"],
        },
        "fused": {
            "amp_updates_per_s": value,
            "terms_per_pass": terms_per_pass,
            "effective_GBps": BYTES_PER_AMP_PASS * value / 1e9,
        },
        "plain": plain,
        "fused_vs_plain": (value / plain["amp_updates_per_s"]) if plain else None,
        "norm_drift": abs(1.0 - norm2),
        "first_visit_seconds": sched_s if schedule == "learned" else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
