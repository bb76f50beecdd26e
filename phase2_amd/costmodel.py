"""Learned cost model of fused passes (include/phase2_learned.h).

An FP32 perceptron (88 -> 64 -> 64 -> 1, ReLU) predicts the time of a fused
pass from the mask-bit features of its terms; the scheduler in
csrc/sched_learned.c uses it to choose pass boundaries and tile widths.  The
network is trained on the GPU by the library's own kernels (mlp.hip:
forward, backward, Adam) from per-pass timings the library records while
profiling (ph2hip_prof_records).

Feature vector (must match ph2_learned_features): bits 0..39 of the OR of
the pass's flip masks, bits 0..39 of the OR of its phase masks, then
terms/48, micro-groups/48, gathered dims/13, contiguous low bits/13,
tile bits/13, whether the first / last micro-group moves its cosets
straight between HBM and registers, and whether the pass takes the
half-tile form.  Target: ln(ms) - (nqb - 30) ln 2.
"""

from __future__ import annotations

import ctypes as C
import math
import time
from pathlib import Path

import numpy as np

from . import _native as nat

NQ = 40
NIN = 2 * NQ + 8
NHID = 64
PARAMS_PATH = Path(__file__).resolve().parent / "costmodel_params.npz"
LN2 = math.log(2.0)

_f32p = C.POINTER(C.c_float)


def _p(a, t):
    return a.ctypes.data_as(t)


class MLP:
    """The library's FP32 network on the GPU (ph2hip_mlp_*)."""

    def __init__(self, nin=NIN, nhid=NHID, max_batch=8192, seed=1):
        L = nat.lib()
        L.ph2hip_mlp_create.restype = C.c_int
        L.ph2hip_mlp_nparams.restype = C.c_size_t
        L.ph2hip_mlp_strerror.restype = C.c_char_p
        for f in ("ph2hip_mlp_set_params", "ph2hip_mlp_get_params",
                  "ph2hip_mlp_forward", "ph2hip_mlp_grad",
                  "ph2hip_mlp_train_step"):
            getattr(L, f).restype = C.c_int
        self.L = L
        self.nin, self.nhid, self.max_batch = nin, nhid, max_batch
        h = C.c_void_p()
        if L.ph2hip_mlp_create(C.byref(h), C.c_uint32(nin), C.c_uint32(nhid),
                               C.c_uint32(max_batch), C.c_uint64(seed)) != 0:
            self._err("ph2hip_mlp_create")
        self.h = h
        self.np = int(L.ph2hip_mlp_nparams(h))

    def _err(self, what):
        msg = self.L.ph2hip_mlp_strerror().decode(errors="replace")
        raise RuntimeError(f"{what} failed: {msg} / {nat.lib().ph2_strerror()}")

    def get_params(self):
        p = np.zeros(self.np, np.float32)
        if self.L.ph2hip_mlp_get_params(self.h, _p(p, _f32p)) != 0:
            self._err("get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.np,)
        if self.L.ph2hip_mlp_set_params(self.h, _p(p, _f32p)) != 0:
            self._err("set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        y = np.zeros(X.shape[0], np.float32)
        if self.L.ph2hip_mlp_forward(self.h, _p(X, _f32p),
                                     C.c_size_t(X.shape[0]), _p(y, _f32p)) != 0:
            self._err("forward")
        return y

    def grad(self, X, t):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.np, np.float32)
        loss = C.c_float()
        if self.L.ph2hip_mlp_grad(self.h, _p(X, _f32p), _p(t, _f32p),
                                  C.c_size_t(X.shape[0]), C.byref(loss),
                                  _p(g, _f32p)) != 0:
            self._err("grad")
        return float(loss.value), g

    def train_step(self, X, t, lr):
        X = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        if self.L.ph2hip_mlp_train_step(self.h, _p(X, _f32p), _p(t, _f32p),
                                        C.c_size_t(X.shape[0]),
                                        C.c_float(lr), C.byref(loss)) != 0:
            self._err("train_step")
        return float(loss.value)

    def close(self):
        if self.h:
            self.L.ph2hip_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def unpack(p, nin=NIN, nhid=NHID):
    """Split packed parameters into (W1, b1, W2, b2, W3, b3)."""
    o = 0
    out = []
    for shape in ((nhid, nin), (nhid,), (nhid, nhid), (nhid,), (nhid,), (1,)):
        n = int(np.prod(shape))
        out.append(np.asarray(p[o:o + n]).reshape(shape))
        o += n
    return out


# -- features ------------------------------------------------------------------

def _bits(v, n=NQ):
    v = np.asarray(v, dtype=np.uint64)
    return ((v[:, None] >> np.arange(n, dtype=np.uint64)) & np.uint64(1)
            ).astype(np.float32)


def features_from_records(rec):
    """Feature matrix and targets of fused-pass records (REC_DTYPE)."""
    rec = rec[rec["kind"] == 2]
    X = np.zeros((len(rec), NIN), np.float32)
    X[:, :NQ] = _bits(rec["xunion"])
    X[:, NQ:2 * NQ] = _bits(rec["zunion"])
    X[:, 2 * NQ + 0] = rec["nterms"] / 48.0
    X[:, 2 * NQ + 1] = rec["ngroups"] / 48.0
    X[:, 2 * NQ + 2] = rec["gdim"] / 13.0
    X[:, 2 * NQ + 3] = rec["low_bits"] / 13.0
    X[:, 2 * NQ + 4] = rec["tile_bits"] / 13.0
    X[:, 2 * NQ + 5] = rec["direct"] & 1
    X[:, 2 * NQ + 6] = (rec["direct"] >> 1) & 1
    X[:, 2 * NQ + 7] = (rec["direct"] >> 2) & 1
    y = (np.log(rec["ms"]) - (rec["nqb"].astype(np.float64) - 30) * LN2
         ).astype(np.float32)
    return X, y, rec


def c_features(xm, zm, tile_bits, shape):
    """ph2_learned_features on one pass (for checking the Python mirror)."""
    n = len(xm)
    rot = (nat.Rot * max(n, 1))()
    for k in range(n):
        rot[k] = nat.Rot(int(xm[k]), int(zm[k]), 1.0, 0.0)
    sh = (C.c_uint32 * 4)(*shape)
    f = np.zeros(NIN, np.float32)
    nat.lib().ph2_learned_features(rot, C.c_size_t(n), C.c_uint32(tile_bits),
                                   sh, _p(f, _f32p))
    return f


# -- data collection -------------------------------------------------------------

def _span_growth(xs, low):
    """rank[k] = dim span{x_0..x_k} mod the low bits."""
    basis = {}
    mask = ~((1 << low) - 1)
    out = []
    for x in xs:
        v = int(x) & mask
        while v:
            h = v.bit_length() - 1
            if h in basis:
                v ^= basis[h]
            else:
                basis[h] = v
                break
        out.append(len(basis))
    return out


def random_schedule(xm, rng, tiles=(12, 13), lows=(3, 4, 5), maxlen=48):
    """Cut a block into passes of random feasible length, tile width and
    contiguous low bits."""
    lens, tbs, lbs = [], [], []
    i, n = 0, len(xm)
    while i < n:
        tb = int(rng.choice(tiles))
        low = int(rng.choice(lows))
        ranks = _span_growth(xm[i:i + maxlen], low)
        jmax = 1
        while jmax < len(ranks) and ranks[jmax] <= tb - low:
            jmax += 1
        ln = int(rng.integers(1, jmax + 1)) if rng.random() < 0.7 else jmax
        lens.append(ln)
        tbs.append(tb)
        lbs.append(low)
        i += ln
    return (np.array(lens, np.uint32), np.array(tbs, np.uint32),
            np.array(lbs, np.uint32))


def collect(reg, xm, zm, phi, blocks, rng, log=print):
    """Run random schedules over the given 1024-term blocks on a device
    register and return the per-pass timing records."""
    L = nat.lib()
    L.ph2_reg_rotate_sched.restype = C.c_int
    nat.prof_enable(True)
    nat.prof_records(reset=True)
    t0 = time.perf_counter()
    for bi, b in enumerate(blocks):
        s = slice(b * 1024, (b + 1) * 1024)
        bx, bz, bp = (np.ascontiguousarray(xm[s]), np.ascontiguousarray(zm[s]),
                      np.ascontiguousarray(phi[s]))
        lens, tbs, lbs = random_schedule(bx, rng)
        rc = L.ph2_reg_rotate_sched(
            C.c_void_p(reg.h), C.c_size_t(len(bx)),
            bx.ctypes.data_as(C.POINTER(C.c_uint64)),
            bz.ctypes.data_as(C.POINTER(C.c_uint64)),
            bp.ctypes.data_as(C.POINTER(C.c_double)),
            lens.ctypes.data_as(C.POINTER(C.c_uint32)),
            tbs.ctypes.data_as(C.POINTER(C.c_uint32)),
            lbs.ctypes.data_as(C.POINTER(C.c_uint32)), C.c_size_t(len(lens)))
        if rc != 0:
            raise RuntimeError(f"ph2_reg_rotate_sched: {L.ph2_strerror()}")
        if (bi + 1) % 8 == 0:
            nat.synchronize()
            log(f"collect: block {bi + 1}/{len(blocks)} "
                f"{time.perf_counter() - t0:.1f}s")
    rec = nat.prof_records(reset=True)
    nat.prof_enable(False)
    return rec


# -- training ------------------------------------------------------------------

def train(X, y, epochs=200, batch=512, lr=2e-3, seed=0, val_frac=0.1,
          log=print):
    """Train the network on the GPU; returns (params, history)."""
    rng = np.random.default_rng(seed)
    n = len(y)
    perm = rng.permutation(n)
    nval = max(1, int(n * val_frac))
    va, tr = perm[:nval], perm[nval:]
    m = MLP(seed=seed + 1)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(tr)
        lr_ep = lr * (0.5 * (1 + math.cos(math.pi * ep / epochs)))
        tot = 0.0
        for k in range(0, len(order), batch):
            idx = order[k:k + batch]
            tot += m.train_step(X[idx], y[idx], lr_ep) * len(idx)
        if ep % 20 == 0 or ep == epochs - 1:
            pv = m.forward(X[va])
            vl = float(np.mean(0.5 * (pv - y[va]) ** 2))
            hist.append((ep, tot / len(tr), vl))
            log(f"train: epoch {ep} loss {tot / len(tr):.5f} val {vl:.5f}")
    p = m.get_params()
    m.close()
    return p, dict(history=hist, val_idx=va)


def save(params, path=PARAMS_PATH, **meta):
    np.savez(path, params=np.asarray(params, np.float32), nin=NIN, nhid=NHID,
             **meta)


def n_params(nin=NIN, nhid=NHID):
    return nhid * nin + nhid + nhid * nhid + nhid + nhid + 1


def load(path=PARAMS_PATH):
    """Parameters and metadata of a saved network; ValueError if it was
    trained for another feature vector."""
    d = np.load(path, allow_pickle=False)
    params = d["params"]
    if int(d["nin"]) != NIN or int(d["nhid"]) != NHID or \
            params.size != n_params():
        raise ValueError(f"{path}: network {int(d['nin'])} -> "
                         f"{int(d['nhid'])}, this library needs {NIN} -> "
                         f"{NHID}")
    return params, {k: d[k] for k in d.files if k != "params"}


def install(params, low=3, min_qb=26):
    """Use the learned schedule for fused registers of >= min_qb qubits
    (passes with low .. low + 2 contiguous low bits)."""
    p = np.ascontiguousarray(params, np.float32)
    if p.size != n_params():
        raise ValueError(f"{p.size} parameters, the network has "
                         f"{n_params()}")
    L = nat.lib()
    L.ph2_learned_install.restype = C.c_int
    if L.ph2_learned_install(_p(p, _f32p), C.c_uint32(NHID), C.c_uint32(low),
                             C.c_uint32(min_qb)) != 0:
        raise RuntimeError(f"ph2_learned_install: {L.ph2_strerror()}")


def uninstall():
    nat.lib().ph2_learned_uninstall()


class Sched(C.Structure):
    """struct ph2_sched (include/phase2/qreg_hip.h)."""
    _fields_ = [("npass", C.c_size_t), ("len", C.POINTER(C.c_uint32)),
                ("tile_bits", C.POINTER(C.c_uint32)),
                ("low_bits", C.POINTER(C.c_uint32)), ("cap", C.c_size_t)]


def _rots(xm, zm, phi=None):
    n = len(xm)
    rot = (nat.Rot * max(n, 1))()
    for k in range(n):
        p = 0.0 if phi is None else float(phi[k])
        rot[k] = nat.Rot(int(xm[k]), int(zm[k]), math.cos(p), math.sin(p))
    return rot


def _sched_arrays(s):
    n = s.npass
    return (np.array(s.len[:n], np.uint32), np.array(s.tile_bits[:n], np.uint32),
            np.array(s.low_bits[:n], np.uint32))


def schedule(xm, zm, nqb, greedy=None):
    """Schedule of one block: the installed learned one, or greedy=(tile,
    low).  Returns (pass lengths, tile bits, low bits)."""
    L = nat.lib()
    rot = _rots(xm, zm)
    s = Sched()
    if greedy is None:
        L.ph2_learned_schedule.restype = C.c_int
        rc = L.ph2_learned_schedule(rot, C.c_size_t(len(xm)),
                                    C.c_uint32(nqb), C.byref(s))
    else:
        L.ph2_sched_greedy.restype = C.c_int
        rc = L.ph2_sched_greedy(rot, C.c_size_t(len(xm)), C.c_uint32(nqb),
                                C.c_uint32(greedy[0]), C.c_uint32(greedy[1]),
                                C.byref(s))
    if rc != 0:
        raise RuntimeError(f"schedule: {L.ph2_strerror()}")
    out = _sched_arrays(s)
    L.ph2_sched_free(C.byref(s))
    return out


def predict(xm, zm, nqb, sched):
    """Predicted milliseconds per pass of a schedule (installed network)."""
    L = nat.lib()
    lens, tbs, lbs = (np.ascontiguousarray(a, np.uint32) for a in sched)
    s = Sched(len(lens), lens.ctypes.data_as(C.POINTER(C.c_uint32)),
              tbs.ctypes.data_as(C.POINTER(C.c_uint32)),
              lbs.ctypes.data_as(C.POINTER(C.c_uint32)), len(lens))
    ms = np.zeros(len(lens))
    L.ph2_learned_predict.restype = C.c_int
    if L.ph2_learned_predict(_rots(xm, zm), C.c_size_t(len(xm)),
                             C.c_uint32(nqb), C.byref(s),
                             ms.ctypes.data_as(C.POINTER(C.c_double))) != 0:
        raise RuntimeError(f"ph2_learned_predict: {L.ph2_strerror()}")
    return ms


def stats():
    h, m = C.c_uint64(), C.c_uint64()
    nat.lib().ph2_learned_stats(C.byref(h), C.byref(m))
    return int(h.value), int(m.value)


# -- evaluation against the measured optimum -----------------------------------

def candidates(xm, zm, nqb, low=3):
    """The candidate passes the learned scheduler chooses from for one block
    (ph2_learned_candidates): an (m, 4) array of first term, end
    (exclusive), tile bits, low bits."""
    L = nat.lib()
    rot = _rots(xm, zm)
    cap = len(xm) * 2 * 3 * 16 + 16
    out = np.zeros(4 * cap, np.uint32)
    L.ph2_learned_candidates.restype = C.c_long
    m = L.ph2_learned_candidates(rot, C.c_size_t(len(xm)), C.c_uint32(nqb),
                                 C.c_uint32(low),
                                 out.ctypes.data_as(C.POINTER(C.c_uint32)),
                                 C.c_size_t(cap))
    if m < 0:
        raise RuntimeError("ph2_learned_candidates")
    return out[:4 * m].reshape(m, 4).copy()


def measure_candidates(reg, xm, zm, phi, cand, reps=2):
    """Kernel time (ms, HIP events, the least of reps) of every candidate
    pass run on its own on a device register."""
    L = nat.lib()
    L.ph2_reg_rotate_sched.restype = C.c_int
    u32p = C.POINTER(C.c_uint32)
    best = np.full(len(cand), np.inf)
    chunk = 20000
    for _ in range(reps):
        for c0 in range(0, len(cand), chunk):
            part = cand[c0:c0 + chunk]
            nat.prof_enable(True)
            nat.prof_records(reset=True)
            for i, j, tb, low in part:
                bx = np.ascontiguousarray(xm[i:j])
                bz = np.ascontiguousarray(zm[i:j])
                bp = np.ascontiguousarray(phi[i:j])
                sch = np.array([[j - i], [tb], [low]], np.uint32)
                rc = L.ph2_reg_rotate_sched(
                    C.c_void_p(reg.h), C.c_size_t(j - i),
                    bx.ctypes.data_as(C.POINTER(C.c_uint64)),
                    bz.ctypes.data_as(C.POINTER(C.c_uint64)),
                    bp.ctypes.data_as(C.POINTER(C.c_double)),
                    sch[0].ctypes.data_as(u32p), sch[1].ctypes.data_as(u32p),
                    sch[2].ctypes.data_as(u32p), C.c_size_t(1))
                if rc != 0:
                    raise RuntimeError("ph2_reg_rotate_sched: "
                                       f"{L.ph2_strerror()}")
            rec = nat.prof_records(reset=True)
            nat.prof_enable(False)
            rec = rec[rec["kind"] == 2]
            if len(rec) != len(part):
                raise RuntimeError(f"{len(rec)} records for {len(part)} "
                                   "candidates")
            best[c0:c0 + len(part)] = np.minimum(best[c0:c0 + len(part)],
                                                 rec["ms"])
    return best


def dp_schedule(cand, cost, n):
    """Cheapest cut of terms [0, n) into candidate passes (shortest path
    over cut points, candidates grouped by ascending first term); returns
    (schedule arrays, total cost)."""
    best = np.full(n + 1, np.inf)
    best[0] = 0.0
    frm = np.full(n + 1, -1, np.int64)
    for c, (i, j, tb, low) in enumerate(cand):
        v = best[i] + cost[c]
        if v < best[j]:
            best[j] = v
            frm[j] = c
    lens, tbs, lbs = [], [], []
    j = n
    while j > 0:
        i, _, tb, low = cand[frm[j]]
        lens.append(j - i)
        tbs.append(tb)
        lbs.append(low)
        j = i
    return (np.array(lens[::-1], np.uint32), np.array(tbs[::-1], np.uint32),
            np.array(lbs[::-1], np.uint32)), float(best[n])
