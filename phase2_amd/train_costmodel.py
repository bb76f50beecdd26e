"""Record fused-pass timings, train the cost model on the GPU, evaluate it.

    python -m phase2_amd.train_costmodel --out DIR [--extra-records GLOB]

1. Records: random cuts (phase2_amd.costmodel.random_schedule) of 1024-term
   blocks of lexicographically sorted synthetic Hamiltonians at 28, 30 and
   33 qubits, timed per pass with HIP events (ph2hip_prof_records), plus any
   record files given with --extra-records.
2. Training: the library's FP32 network (85 -> 64 -> 64 -> 1), Adam, on the
   GPU, 10% held out.
3. Evaluation on 33-qubit blocks never used for training: the same blocks
   are swept with (a) the greedy cut at tile 12 / low 4, (b) the greedy cut
   at every tile/low setting, keeping the fastest (exhaustive tuning of the
   fixed setting), and (c) the learned schedule; wall time per block after
   a synchronisation, and predicted-vs-measured error per pass.
Writes costmodel_params.npz, records.npy and report.json into DIR.
"""

from __future__ import annotations

import argparse
import glob
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import _native as nat
from . import costmodel as cm
from . import synth


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sorted_terms(hm):
    perm = nat.sort_lex(hm.xm, hm.zm)
    return hm.xm[perm], hm.zm[perm], 0.02 * hm.cf[perm]


def collect_width(nqb, hm, blocks, rng):
    xm, zm, phi = sorted_terms(hm)
    reg = nat.Register(nqb, nat.FUSED, 12, 4)
    reg.prep_multidet(np.array([(1 << (nqb // 2)) - 1], np.uint64),
                      np.array([1.0 + 0j]))
    rec = cm.collect(reg, xm, zm, phi, blocks, rng,
                     log=lambda m: log(f"n={nqb} {m}"))
    reg.free()
    return rec


def time_blocks(reg, xm, zm, phi, blocks):
    out = []
    for b in blocks:
        s = slice(b * 1024, (b + 1) * 1024)
        nat.synchronize()
        t0 = time.perf_counter()
        reg.rotate(xm[s], zm[s], phi[s])
        nat.synchronize()
        out.append(time.perf_counter() - t0)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--extra-records", default="")
    ap.add_argument("--blocks28", type=int, default=60)
    ap.add_argument("--blocks30", type=int, default=36)
    ap.add_argument("--blocks33", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=250)
    args = ap.parse_args(argv)
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    nat.init()
    rng = np.random.default_rng(2024)

    recs = []
    # 28 and 30 qubits: own Hamiltonians of the same generator
    for nqb, nb in ((28, args.blocks28), (30, args.blocks30)):
        hm = synth.hamiltonian(nqb, 0, seed=nqb, decay=True,
                               n_total=1024 * (nb + 4))
        blocks = list(range(2, nb + 2))
        recs.append(collect_width(nqb, hm, blocks, rng))
        log(f"records n={nqb}: {len(recs[-1])}")
    # 33 qubits: blocks of the headline sweep, held-out blocks excluded
    w = synth.trott_n33()
    xm33, zm33, phi33 = sorted_terms(w["hm"])
    nblk = (len(phi33) + 1023) // 1024
    eval_blocks = [9, 27, 44, 61, 78, 95]
    train33 = [b for b in np.linspace(3, nblk - 2, args.blocks33).astype(int)
               if b not in eval_blocks]
    reg = nat.Register(33, nat.FUSED, 12, 4)
    reg.prep_multidet(w["det_idx"], w["det_cf"])
    recs.append(cm.collect(reg, xm33, zm33, phi33, train33, rng,
                           log=lambda m: log(f"n=33 {m}")))
    for f in sorted(glob.glob(args.extra_records)) if args.extra_records else []:
        r = np.load(f, allow_pickle=False)
        recs.append(r[r["kind"] == 2])
        log(f"extra records {f}: {len(recs[-1])}")
    rec = np.concatenate(recs)
    np.save(out / "records.npy", rec)
    X, y, rec = cm.features_from_records(rec)
    log(f"training on {len(y)} pass timings")

    t0 = time.perf_counter()
    params, info = cm.train(X, y, epochs=args.epochs, log=log)
    train_s = time.perf_counter() - t0
    m = cm.MLP()
    m.set_params(params)
    va = info["val_idx"]
    pv = m.forward(X[va])
    rel_val = float(np.median(np.abs(np.exp(pv - y[va]) - 1.0)))
    m.close()
    cm.save(params, out / "costmodel_params.npz")

    # evaluation on held-out 33-qubit blocks
    res = {}
    t_greedy = time_blocks(reg, xm33, zm33, phi33, eval_blocks)
    res["greedy_12_4_s"] = t_greedy
    best = (None, 1e300)
    for tb, low in ((12, 3), (12, 4), (12, 5), (12, 6), (11, 3), (11, 4),
                    (11, 5), (10, 4)):
        reg.configure(nat.FUSED, tb, low)
        t = sum(time_blocks(reg, xm33, zm33, phi33, eval_blocks))
        res[f"greedy_{tb}_{low}_total_s"] = t
        if t < best[1]:
            best = ((tb, low), t)
    res["exhaustive_best"] = dict(tile=best[0][0], low=best[0][1],
                                  total_s=best[1])
    reg.configure(nat.FUSED, 12, 4)
    cm.install(params, low=4, min_qb=28)
    # first sweep computes and caches the schedules, the second is timed
    time_blocks(reg, xm33, zm33, phi33, eval_blocks)
    nat.prof_enable(True)
    nat.prof_records(reset=True)
    t_learned = time_blocks(reg, xm33, zm33, phi33, eval_blocks)
    lrec = nat.prof_records(reset=True)
    nat.prof_enable(False)
    Xl, yl, lrec = cm.features_from_records(lrec)
    m = cm.MLP()
    m.set_params(params)
    pl = m.forward(Xl)
    m.close()
    err = np.abs(np.exp(pl - yl) - 1.0)
    hits, misses = cm.stats()
    cm.uninstall()
    reg.free()
    res.update(
        learned_s=t_learned,
        greedy_total_s=float(sum(t_greedy)),
        learned_total_s=float(sum(t_learned)),
        learned_vs_greedy=float(sum(t_greedy) / sum(t_learned)),
        learned_vs_exhaustive=float(best[1] / sum(t_learned)),
        eval_blocks=eval_blocks,
        records=int(len(y)),
        train_seconds=train_s,
        val_median_rel_err=rel_val,
        eval_pass_median_rel_err=float(np.median(err)),
        eval_pass_p90_rel_err=float(np.quantile(err, 0.9)),
        eval_passes=int(len(err)),
        cache_hits=hits, cache_misses=misses,
        history=info["history"])
    (out / "report.json").write_text(json.dumps(res, indent=1))
    log(json.dumps({k: v for k, v in res.items() if k != "history"}))


if __name__ == "__main__":
    main()
