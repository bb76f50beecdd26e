"""Record fused-pass timings, train the cost model on the GPU, evaluate it.

    python -m phase2_amd.train_costmodel --out DIR [--no-eval]
    python -m phase2_amd.train_costmodel --out DIR --eval-only PARAMS

1. Records: random cuts (phase2_amd.costmodel.random_schedule: random pass
   length, tile width 12/13 and 3..5 contiguous low bits per pass) of
   1024-term blocks of lexicographically sorted synthetic Hamiltonians at 28
   (two seeds), 30 and 33 qubits, timed per pass with HIP events
   (ph2hip_prof_records), plus any record files given with --extra-records.
2. Training: the library's FP32 network (88 -> 64 -> 64 -> 1), Adam, on the
   GPU, 10% held out.
3. Evaluation on 33-qubit blocks never used for training: the same blocks
   are swept with (a) the greedy cut at tile 12 / low 3, (b) the greedy cut
   at four tile/low settings, keeping the fastest (exhaustive tuning of the
   fixed setting), and (c) the learned schedule; wall time per block after
   a synchronisation, and predicted-vs-measured error per pass.
4. Evaluation on the composite integrator at 30 qubits (synth.cmpsit_n30):
   greedy, exhaustive and learned schedules over whole cmpsit steps, with
   the overlaps compared bit for bit.
Writes costmodel_params.npz, records.npy and train.json (steps 1-2), and
report.json and report_cmpsit.json (steps 3-4) into DIR.
"""

from __future__ import annotations

import argparse
import ctypes as C
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


def time_blocks(reg, xm, zm, phi, blocks, what=""):
    out = []
    for b in blocks:
        s = slice(b * 1024, (b + 1) * 1024)
        nat.synchronize()
        t0 = time.perf_counter()
        reg.rotate(xm[s], zm[s], phi[s])
        nat.synchronize()
        out.append(time.perf_counter() - t0)
    log(f"eval {what}: {sum(out):.3f}s")
    return out


def eval_cmpsit(params, steps=1, log=log):
    """Composite integrator at 30 qubits (synth.cmpsit_n30, `steps` steps):
    greedy 12/3, greedy at four tile/low settings (exhaustive tuning of the
    fixed setting), learned; overlaps must agree bit for bit."""
    w = synth.cmpsit_n30()
    hm = w["hm"]
    prep = dict(det_idx=w["det_idx"], det_cf=w["det_cf"])

    def run():
        nat.synchronize()
        t0 = time.perf_counter()
        out, _ = nat.cmpsit(hm, prep, 1, steps, w["length"], w["depth"],
                            w["angle_det"], w["angle_rand"], w["seed"],
                            want_state=False)
        nat.synchronize()
        return out, time.perf_counter() - t0

    res = {}
    outs = {}
    best = (None, 1e300)
    for tb, low in ((12, 3), (12, 4), (13, 3), (13, 4)):
        nat.set_default(nat.FUSED, tb, low)
        outs[(tb, low)], t = run()
        res[f"greedy_{tb}_{low}_s"] = t
        log(f"cmpsit greedy {tb}/{low}: {t:.3f}s")
        if t < best[1]:
            best = ((tb, low), t)
    nat.set_default(nat.FUSED, 12, 3)
    cm.install(params, min_qb=28)
    nat.prof_enable(True)
    nat.prof_records(reset=True)
    out_l, t_l = run()
    lrec = nat.prof_records(reset=True)
    nat.prof_enable(False)
    hits, misses = cm.stats()
    cm.uninstall()
    log(f"cmpsit learned: {t_l:.3f}s")
    Xl, yl, _ = cm.features_from_records(lrec)
    m = cm.MLP()
    m.set_params(params)
    pl = m.forward(Xl)
    m.close()
    err = np.abs(np.exp(pl - yl) - 1.0)
    same = all(np.array_equal(o, out_l) for o in outs.values())
    res.update(
        steps=steps, learned_s=t_l,
        learned_vs_greedy=res["greedy_12_3_s"] / t_l,
        exhaustive_best=dict(tile=best[0][0], low=best[0][1], s=best[1]),
        learned_vs_exhaustive=best[1] / t_l,
        overlaps_bit_identical=bool(same),
        pass_median_rel_err=float(np.median(err)) if len(err) else None,
        pass_p90_rel_err=float(np.quantile(err, 0.9)) if len(err) else None,
        passes=int(len(err)), cache_hits=hits, cache_misses=misses)
    return res


EVAL_BLOCKS = [9, 27, 44, 61, 78, 95]


def eval_n33(params, log=log):
    """Held-out 33-qubit sweep blocks: greedy 12/3, greedy at four tile/low
    settings (exhaustive tuning of the fixed setting), learned."""
    w = synth.trott_n33()
    xm33, zm33, phi33 = sorted_terms(w["hm"])
    reg = nat.Register(33, nat.FUSED, 12, 4)
    reg.prep_multidet(w["det_idx"], w["det_cf"])
    res = {}
    best = (None, 1e300)
    for tb, low in ((12, 3), (12, 4), (13, 3), (13, 4)):
        reg.configure(nat.FUSED, tb, low)
        t = time_blocks(reg, xm33, zm33, phi33, EVAL_BLOCKS,
                        f"greedy {tb}/{low}")
        res[f"greedy_{tb}_{low}_s"] = t
        if sum(t) < best[1]:
            best = ((tb, low), sum(t))
    reg.configure(nat.FUSED, 12, 3)
    cm.install(params, min_qb=28)
    # first sweep computes and caches the schedules, the second is timed
    t_first = time_blocks(reg, xm33, zm33, phi33, EVAL_BLOCKS,
                          "learned, first visit")
    nat.prof_enable(True)
    nat.prof_records(reset=True)
    t_learned = time_blocks(reg, xm33, zm33, phi33, EVAL_BLOCKS, "learned")
    lrec = nat.prof_records(reset=True)
    nat.prof_enable(False)
    ov_l = reg.overlap(w["det_idx"], w["det_cf"])
    hits, misses = cm.stats()
    cm.uninstall()
    reg.free()
    Xl, yl, lrec = cm.features_from_records(lrec)
    m = cm.MLP()
    m.set_params(params)
    pl = m.forward(Xl)
    m.close()
    err = np.abs(np.exp(pl - yl) - 1.0)
    g = sum(res["greedy_12_3_s"])
    res.update(
        learned_s=t_learned, learned_first_visit_s=t_first,
        greedy_total_s=g, learned_total_s=float(sum(t_learned)),
        learned_vs_greedy=float(g / sum(t_learned)),
        exhaustive_best=dict(tile=best[0][0], low=best[0][1],
                             total_s=best[1]),
        learned_vs_exhaustive=float(best[1] / sum(t_learned)),
        eval_blocks=EVAL_BLOCKS,
        learned_passes=int(len(lrec)),
        learned_low_bits=np.bincount(lrec["low_bits"], minlength=13
                                     ).tolist(),
        learned_tile_bits=np.bincount(lrec["tile_bits"], minlength=13
                                      ).tolist(),
        eval_pass_median_rel_err=float(np.median(err)),
        eval_pass_p90_rel_err=float(np.quantile(err, 0.9)),
        cache_hits=hits, cache_misses=misses,
        overlap_after_learned=[ov_l.real, ov_l.imag])
    return res


def time_sched(reg, xm, zm, phi, sched, reps=3):
    """Wall time (least of reps) of one block under an explicit schedule."""
    L = nat.lib()
    L.ph2_reg_rotate_sched.restype = C.c_int
    u32p = C.POINTER(C.c_uint32)
    lens, tbs, lbs = (np.ascontiguousarray(a, np.uint32) for a in sched)
    best = 1e300
    for _ in range(reps):
        nat.synchronize()
        t0 = time.perf_counter()
        if L.ph2_reg_rotate_sched(
                C.c_void_p(reg.h), C.c_size_t(len(xm)),
                xm.ctypes.data_as(C.POINTER(C.c_uint64)),
                zm.ctypes.data_as(C.POINTER(C.c_uint64)),
                phi.ctypes.data_as(C.POINTER(C.c_double)),
                lens.ctypes.data_as(u32p), tbs.ctypes.data_as(u32p),
                lbs.ctypes.data_as(u32p), C.c_size_t(len(lens))) != 0:
            raise RuntimeError("ph2_reg_rotate_sched")
        nat.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def eval_dp(params, nqb=26, blocks=(3, 9, 15), log=log):
    """The learned schedule against the best schedule over the same
    candidate passes by their measured times: every candidate pass of a
    block is timed on its own, a shortest path over cut points picks the
    measured optimum, and the greedy cuts, the learned schedule and that
    optimum are timed as whole blocks (least of three)."""
    hm = synth.hamiltonian(nqb, 0, seed=4242, decay=True,
                           n_total=1024 * (max(blocks) + 2))
    xm, zm, phi = sorted_terms(hm)
    reg = nat.Register(nqb, nat.FUSED, 12, 3)
    reg.prep_multidet(np.array([(1 << (nqb // 2)) - 1], np.uint64),
                      np.array([1.0 + 0j]))
    cm.install(params, min_qb=nqb)
    rows = []
    try:
        for b in blocks:
            sl = slice(b * 1024, (b + 1) * 1024)
            bx, bz, bp = (np.ascontiguousarray(a[sl]) for a in (xm, zm, phi))
            cand = cm.candidates(bx, bz, nqb)
            t0 = time.perf_counter()
            ms = cm.measure_candidates(reg, bx, bz, bp, cand, reps=1)
            log(f"dp: block {b}: {len(cand)} candidates timed in "
                f"{time.perf_counter() - t0:.1f}s")
            opt, opt_ms = cm.dp_schedule(cand, ms, len(bx))
            lrn = cm.schedule(bx, bz, nqb)
            # the learned schedule's passes at their measured times
            key = {tuple(c): k for k, c in enumerate(cand.tolist())}
            i, lrn_ms = 0, 0.0
            for ln, tb, low in zip(*lrn):
                lrn_ms += ms[key[(i, i + int(ln), int(tb), int(low))]]
                i += int(ln)
            row = dict(block=b, candidates=int(len(cand)),
                       optimum_sum_ms=opt_ms, learned_sum_ms=float(lrn_ms),
                       optimum_passes=int(len(opt[0])),
                       learned_passes=int(len(lrn[0])))
            for name, sch in (("greedy_12_3", cm.schedule(bx, bz, nqb,
                                                           greedy=(12, 3))),
                              ("greedy_13_3", cm.schedule(bx, bz, nqb,
                                                           greedy=(13, 3))),
                              ("learned", lrn), ("optimum", opt)):
                row[name + "_s"] = time_sched(reg, bx, bz, bp, sch)
            log(f"dp: {row}")
            rows.append(row)
    finally:
        cm.uninstall()
        reg.free()
    tot = {k: float(sum(r[k] for r in rows)) for k in
           ("greedy_12_3_s", "greedy_13_3_s", "learned_s", "optimum_s")}
    return dict(nqb=nqb, blocks=list(blocks), rows=rows, totals=tot,
                learned_vs_optimum=tot["optimum_s"] / tot["learned_s"],
                learned_vs_greedy=tot["greedy_12_3_s"] / tot["learned_s"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--extra-records", default="")
    ap.add_argument("--blocks28", type=int, default=75)
    ap.add_argument("--blocks30", type=int, default=60)
    ap.add_argument("--blocks33", type=int, default=12)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--eval-only", default="",
                    help="parameter file: skip collection and training and "
                    "evaluate (held-out 33-qubit blocks, composite "
                    "integrator at 30 qubits)")
    ap.add_argument("--no-eval", action="store_true")
    ap.add_argument("--records", default="",
                    help="train on this saved records.npy instead of "
                    "collecting new timings")
    ap.add_argument("--cmpsit-steps", type=int, default=1)
    ap.add_argument("--eval-dp", action="store_true",
                    help="also compare the learned schedule with the "
                    "measured-optimal schedule over the same candidates")
    args = ap.parse_args(argv)
    out = Path(args.out)
    out.mkdir(parents=True, exist_ok=True)
    nat.init()
    if args.eval_only:
        params, _ = cm.load(args.eval_only)
        if args.eval_dp:
            res_d = eval_dp(params)
            (out / "report_dp.json").write_text(json.dumps(res_d, indent=1))
            log(json.dumps(res_d))
        res = eval_n33(params)
        (out / "report.json").write_text(json.dumps(res, indent=1))
        log(json.dumps(res))
        res_c = eval_cmpsit(params, args.cmpsit_steps)
        (out / "report_cmpsit.json").write_text(json.dumps(res_c, indent=1))
        log(json.dumps(res_c))
        return
    rng = np.random.default_rng(2024)

    recs = []
    train33 = []
    if args.records:
        recs.append(np.load(args.records, allow_pickle=False))
        log(f"records from {args.records}: {len(recs[-1])}")
    # 28 qubits (two Hamiltonians: one holds at most ~80 blocks of distinct
    # four-index strings) and 30 qubits, same generator
    for nqb, nb, seed in () if args.records else (
            (28, args.blocks28, 28), (28, args.blocks28, 128),
            (30, args.blocks30, 30)):
        hm = synth.hamiltonian(nqb, 0, seed=seed, decay=True,
                               n_total=1024 * (nb + 4))
        blocks = list(range(2, nb + 2))
        recs.append(collect_width(nqb, hm, blocks, rng))
        log(f"records n={nqb} seed {seed}: {len(recs[-1])}")
    # 33 qubits: blocks of the headline sweep, held-out blocks excluded
    if not args.records:
        w = synth.trott_n33()
        xm33, zm33, phi33 = sorted_terms(w["hm"])
        nblk = (len(phi33) + 1023) // 1024
        train33 = [int(b) for b in np.linspace(3, nblk - 2, args.blocks33
                                               ).astype(int)
                   if b not in EVAL_BLOCKS]
        reg = nat.Register(33, nat.FUSED, 12, 4)
        reg.prep_multidet(w["det_idx"], w["det_cf"])
        recs.append(cm.collect(reg, xm33, zm33, phi33, train33, rng,
                               log=lambda m: log(f"n=33 {m}")))
        log(f"records n=33: {len(recs[-1])}")
        reg.free()
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
    cm.save(params, out / "costmodel_params.npz", records=len(y))
    res = dict(records=int(len(y)), train_seconds=train_s,
               val_median_rel_err=rel_val, history=info["history"],
               train33_blocks=train33)
    (out / "train.json").write_text(json.dumps(res, indent=1))
    log(json.dumps({k: v for k, v in res.items() if k != "history"}))
    if args.no_eval:
        return
    res_n = eval_n33(params)
    (out / "report.json").write_text(json.dumps(res_n, indent=1))
    log(json.dumps(res_n))
    res_c = eval_cmpsit(params, args.cmpsit_steps)
    (out / "report_cmpsit.json").write_text(json.dumps(res_c, indent=1))
    log(json.dumps(res_c))
    if args.eval_dp:
        res_d = eval_dp(params)
        (out / "report_dp.json").write_text(json.dumps(res_d, indent=1))
        log(json.dumps(res_d))


if __name__ == "__main__":
    main()
