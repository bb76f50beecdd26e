"""ctypes bindings of libphase2_hip.so (the C library over the HIP kernels).

The library is built in-tree by ``make -C phase2_amd`` (``__graft_entry__.build``)
and loaded from this directory, or from ``PHASE2_LIB``.  There is no fallback:
if the library is missing, importing the entry points raises.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libphase2_hip.so"

_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)

PLAIN, FUSED = 0, 1


class Prep(C.Structure):
    """struct ph2_prep (include/phase2_flat.h)."""
    _fields_ = [
        ("kind", C.c_int),
        ("ndet", C.c_size_t),
        ("didx", _u64p),
        ("dcf", _f64p),
        ("n_sites", C.c_uint32),
        ("n_alpha", C.c_uint32),
        ("n_beta", C.c_uint32),
        ("closed_shell", C.c_int),
        ("tapered", C.c_int),
        ("Ca", _f64p),
        ("Cb", _f64p),
    ]


class Rot(C.Structure):
    """struct ph2hip_rot (include/phase2_hip.h)."""
    _fields_ = [("x", C.c_uint64), ("z", C.c_uint64), ("c", C.c_double),
                ("s", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("PHASE2_LIB", LIB_PATH))
        if not path.is_file():
            raise RuntimeError(
                f"{path} not found: build it with `make -C phase2_amd` "
                "(the HIP backend has no CPU fallback)")
        _lib = C.CDLL(str(path))
        _declare(_lib)
    return _lib


def _declare(L):
    L.ph2_init.restype = C.c_int
    L.ph2_strerror.restype = C.c_char_p
    L.ph2_set_default.restype = None
    L.ph2_set_default.argtypes = [C.c_int, C.c_uint32, C.c_uint32]
    L.phase2_run.restype = C.c_int
    L.phase2_run.argtypes = [C.POINTER(C.c_char_p), _f64p, C.c_size_t,
                             C.c_double, C.c_char_p, _f64p, _f64p]
    for name in ("ph2_trott", "ph2_qdrift", "ph2_cmpsit", "ph2_rotate",
                 "ph2_expand", "ph2_inner"):
        getattr(L, name).restype = C.c_int


def _p(a, t):
    return a.ctypes.data_as(t)


def _err(what):
    msg = lib().ph2_strerror().decode(errors="replace")
    raise RuntimeError(f"{what} failed: {msg}")


def make_prep(det_idx=None, det_cf=None, C_alpha=None, C_beta=None,
              n_sites=0, n_alpha=0, n_beta=0, closed_shell=1, tapered=0):
    keep = []
    p = Prep()
    if det_idx is not None:
        idx = np.ascontiguousarray(det_idx, dtype=np.uint64)
        cf = np.ascontiguousarray(
            np.asarray(det_cf, dtype=np.complex128).view(np.float64))
        keep += [idx, cf]
        p.kind, p.ndet = 1, idx.shape[0]
        p.didx, p.dcf = _p(idx, _u64p), _p(cf, _f64p)
    else:
        ca = np.ascontiguousarray(C_alpha, dtype=np.float64).reshape(-1)
        keep.append(ca)
        p.kind = 2
        p.n_sites, p.n_alpha, p.n_beta = n_sites, n_alpha, n_beta
        p.closed_shell, p.tapered = closed_shell, tapered
        p.Ca = _p(ca, _f64p)
        if C_beta is not None:
            cb = np.ascontiguousarray(C_beta, dtype=np.float64).reshape(-1)
            keep.append(cb)
            p.Cb = _p(cb, _f64p)
    return p, keep


def _terms(hm):
    return (np.ascontiguousarray(hm.xm, dtype=np.uint64),
            np.ascontiguousarray(hm.zm, dtype=np.uint64),
            np.ascontiguousarray(hm.cf, dtype=np.float64))


def init():
    if lib().ph2_init() != 0:
        _err("ph2_init")


def set_default(mode=FUSED, tile_bits=12, low_bits=3):
    lib().ph2_set_default(int(mode), int(tile_bits), int(low_bits))


def trott(algo, hm, prep_kw, delta, steps, want_state=True):
    xm, zm, cf = _terms(hm)
    prep, keep = make_prep(**prep_kw)
    out = np.zeros(steps, dtype=np.complex128)
    st = np.zeros(1 << hm.nqb, dtype=np.complex128) if want_state else None
    rc = lib().ph2_trott(C.c_int(algo), C.c_uint32(hm.nqb),
                         C.c_size_t(len(cf)), _p(xm, _u64p), _p(zm, _u64p),
                         _p(cf, _f64p), C.byref(prep), C.c_double(delta),
                         C.c_size_t(steps), _p(out, _f64p),
                         _p(st, _f64p) if st is not None else None)
    if rc != 0:
        _err("ph2_trott")
    return out, st


def qdrift(hm, prep_kw, samples, depth, step_size, seed, want_state=True):
    xm, zm, cf = _terms(hm)
    prep, keep = make_prep(**prep_kw)
    out = np.zeros(samples, dtype=np.complex128)
    st = np.zeros(1 << hm.nqb, dtype=np.complex128) if want_state else None
    rc = lib().ph2_qdrift(C.c_uint32(hm.nqb), C.c_size_t(len(cf)),
                          _p(xm, _u64p), _p(zm, _u64p), _p(cf, _f64p),
                          C.byref(prep), C.c_size_t(samples),
                          C.c_size_t(depth), C.c_double(step_size),
                          C.c_uint64(seed), _p(out, _f64p),
                          _p(st, _f64p) if st is not None else None)
    if rc != 0:
        _err("ph2_qdrift")
    return out, st


def cmpsit(hm, prep_kw, samples, steps, length, depth, angle_det,
           angle_rand, seed, want_state=True):
    xm, zm, cf = _terms(hm)
    prep, keep = make_prep(**prep_kw)
    out = np.zeros(samples, dtype=np.complex128)
    st = np.zeros(1 << hm.nqb, dtype=np.complex128) if want_state else None
    rc = lib().ph2_cmpsit(C.c_uint32(hm.nqb), C.c_size_t(len(cf)),
                          _p(xm, _u64p), _p(zm, _u64p), _p(cf, _f64p),
                          C.byref(prep), C.c_size_t(samples),
                          C.c_size_t(steps), C.c_size_t(length),
                          C.c_size_t(depth), C.c_double(angle_det),
                          C.c_double(angle_rand), C.c_uint64(seed),
                          _p(out, _f64p),
                          _p(st, _f64p) if st is not None else None)
    if rc != 0:
        _err("ph2_cmpsit")
    return out, st


def rotate(state, xm, zm, phi, mode=FUSED, tile_bits=12, low_bits=3):
    st = np.array(state, dtype=np.complex128, copy=True)
    nqb = int(st.shape[0]).bit_length() - 1
    xm = np.ascontiguousarray(xm, dtype=np.uint64)
    zm = np.ascontiguousarray(zm, dtype=np.uint64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    rc = lib().ph2_rotate(_p(st, _f64p), C.c_uint32(nqb),
                          C.c_size_t(xm.shape[0]), _p(xm, _u64p),
                          _p(zm, _u64p), _p(phi, _f64p), C.c_int(mode),
                          C.c_uint32(tile_bits), C.c_uint32(low_bits))
    if rc != 0:
        _err("ph2_rotate")
    return st


def expand(nqb, n_sites, n_alpha, n_beta, closed_shell, tapered, Ca,
           Cb=None):
    st = np.zeros(1 << nqb, dtype=np.complex128)
    ca = np.ascontiguousarray(Ca, dtype=np.float64).reshape(-1)
    cb = (np.ascontiguousarray(Cb, dtype=np.float64).reshape(-1)
          if Cb is not None else ca)
    rc = lib().ph2_expand(C.c_uint32(nqb), C.c_uint32(n_sites),
                          C.c_uint32(n_alpha), C.c_uint32(n_beta),
                          C.c_int(closed_shell), C.c_int(tapered),
                          _p(ca, _f64p), _p(cb, _f64p), _p(st, _f64p))
    if rc != 0:
        _err("ph2_expand")
    return st


def inner(state, n_sites, n_alpha, n_beta, closed_shell, tapered, Ca,
          Cb=None):
    st = np.ascontiguousarray(state, dtype=np.complex128)
    nqb = int(st.shape[0]).bit_length() - 1
    ca = np.ascontiguousarray(Ca, dtype=np.float64).reshape(-1)
    cb = (np.ascontiguousarray(Cb, dtype=np.float64).reshape(-1)
          if Cb is not None else ca)
    out = np.zeros(2)
    rc = lib().ph2_inner(C.c_uint32(nqb), C.c_uint32(n_sites),
                         C.c_uint32(n_alpha), C.c_uint32(n_beta),
                         C.c_int(closed_shell), C.c_int(tapered),
                         _p(ca, _f64p), _p(cb, _f64p), _p(st, _f64p),
                         _p(out, _f64p))
    if rc != 0:
        _err("ph2_inner")
    return complex(out[0], out[1])


# -- sorting, device-resident registers, profiling ---------------------------

def sort_lex(xm, zm):
    xm = np.ascontiguousarray(xm, dtype=np.uint64)
    zm = np.ascontiguousarray(zm, dtype=np.uint64)
    perm = np.zeros(xm.shape[0], dtype=np.uint64)
    L = lib()
    L.ph2_sort_lex.restype = C.c_int
    if L.ph2_sort_lex(C.c_size_t(xm.shape[0]), _p(xm, _u64p), _p(zm, _u64p),
                      _p(perm, _u64p)) != 0:
        raise MemoryError("ph2_sort_lex")
    return perm


def xoshiro_state(seed) -> np.ndarray:
    """The four state words after xoshiro256ss_init(seed) (host library)."""
    st = np.zeros(4, dtype=np.uint64)
    L = lib()
    L.xoshiro256ss_init.restype = None
    L.xoshiro256ss_init(_p(st, _u64p), C.c_uint64(seed))
    return st


class _Cdf(C.Structure):
    _fields_ = [("y", _f64p), ("len", C.c_size_t)]


def cdf_table(w) -> np.ndarray:
    """The cumulative table prob_cdf_from_array_strided builds over |w|."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    L = lib()
    L.prob_cdf_init.restype = C.c_int
    L.prob_cdf_from_array_strided.restype = C.c_int
    cdf = _Cdf()
    if L.prob_cdf_init(C.byref(cdf), C.c_size_t(w.shape[0])) != 0:
        raise ValueError("prob_cdf_init")
    try:
        if L.prob_cdf_from_array_strided(C.byref(cdf), _p(w, _f64p),
                                         C.c_size_t(8), None) != 0:
            raise ValueError("weights sum below DBL_EPSILON")
        return np.ctypeslib.as_array(cdf.y, shape=(w.shape[0],)).copy()
    finally:
        L.prob_cdf_free(C.byref(cdf))


def sample_cdf(tables, states, ndraws):
    """Device draws (ph2hip_sample_cdf): for each stream k, ndraws[k]
    indices into tables[k] (cumulative, as prob_cdf builds them) from the
    xoshiro256** state states[k] (4 words).  Returns (list of index arrays,
    the advanced states as an (n, 4) array)."""
    n = len(tables)
    tabs = [np.ascontiguousarray(t, dtype=np.float64) for t in tables]
    ptrs = (C.POINTER(C.c_double) * max(n, 1))(*[_p(t, _f64p) for t in tabs])
    lens = np.array([t.shape[0] for t in tabs], dtype=np.uint64)
    nd = np.ascontiguousarray(ndraws, dtype=np.uint64)
    st = np.ascontiguousarray(np.asarray(states, dtype=np.uint64).reshape(n, 4))
    out = np.zeros(max(1, int(nd.sum())), dtype=np.uint32)
    L = lib()
    L.ph2hip_sample_cdf.restype = C.c_int
    if L.ph2hip_sample_cdf(C.c_size_t(n), ptrs, _p(lens, _u64p),
                           _p(st, _u64p), _p(nd, _u64p),
                           _p(out, _u32p)) != 0:
        L.ph2hip_strerror.restype = C.c_char_p
        raise RuntimeError("ph2hip_sample_cdf: " +
                           L.ph2hip_strerror().decode())
    offs = np.concatenate([[0], np.cumsum(nd)]).astype(np.int64)
    return [out[offs[k]:offs[k + 1]].copy() for k in range(n)], st


class Prof(C.Structure):
    """struct ph2hip_prof: kinds 0 plain pair, 1 plain diagonal, 2 fused,
    3 batched fused."""
    _fields_ = [("launches", C.c_uint64 * 4), ("ms", C.c_double * 4),
                ("terms", C.c_uint64 * 4), ("amps", C.c_uint64 * 4)]


def prof_enable(on: bool):
    if lib().ph2hip_prof_enable(C.c_int(int(on))) != 0:
        _err("ph2hip_prof_enable")


def prof_read(reset=True) -> dict:
    p = Prof()
    if lib().ph2hip_prof_read(C.byref(p), C.c_int(int(reset))) != 0:
        _err("ph2hip_prof_read")
    kinds = ("plain_pair", "plain_diag", "fused", "batched")
    return {k: dict(launches=int(p.launches[i]), ms=float(p.ms[i]),
                    terms=int(p.terms[i]), amps=int(p.amps[i]))
            for i, k in enumerate(kinds)}


def synchronize():
    if lib().ph2hip_synchronize() != 0:
        _err("ph2hip_synchronize")


def hash_state(nqb, seed):
    """The state ph2hip_fill_hash writes, restated in numpy (uint64
    arithmetic wraps as in the kernel)."""
    g = np.uint64(0x9e3779b97f4a7c15)
    i = np.arange(1 << nqb, dtype=np.uint64)

    def mix(x):
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
        return x ^ (x >> np.uint64(31))

    def unit(h):
        return (h >> np.uint64(11)).astype(np.float64) * 2.0 ** -52 - 1.0

    with np.errstate(over="ignore"):
        re = unit(mix(np.uint64(seed) + (np.uint64(2) * i + np.uint64(1)) * g))
        im = unit(mix(np.uint64(seed) + (np.uint64(2) * i + np.uint64(2)) * g))
    return re + 1j * im


class Register:
    """A device-resident register (struct qreg behind ph2_reg_*)."""

    def __init__(self, nqb, mode=FUSED, tile_bits=12, low_bits=3):
        L = lib()
        L.ph2_reg_create.restype = C.c_void_p
        for f in ("ph2_reg_prep_multidet", "ph2_reg_rotate",
                  "ph2_reg_overlap_multidet", "ph2_reg_norm2", "ph2_reg_read",
                  "ph2_reg_configure"):
            getattr(L, f).restype = C.c_int
        self.nqb = nqb
        self.h = L.ph2_reg_create(C.c_uint32(nqb), C.c_int(mode),
                                  C.c_uint32(tile_bits), C.c_uint32(low_bits))
        if not self.h:
            _err(f"ph2_reg_create({nqb})")

    def configure(self, mode, tile_bits=12, low_bits=3):
        lib().ph2_reg_configure(C.c_void_p(self.h), C.c_int(mode),
                                C.c_uint32(tile_bits), C.c_uint32(low_bits))

    def prep_multidet(self, idx, cf):
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        cf = np.ascontiguousarray(np.asarray(cf, np.complex128).view(np.float64))
        if lib().ph2_reg_prep_multidet(C.c_void_p(self.h),
                                       C.c_size_t(idx.shape[0]),
                                       _p(idx, _u64p), _p(cf, _f64p)) != 0:
            _err("ph2_reg_prep_multidet")

    def rotate(self, xm, zm, phi):
        xm = np.ascontiguousarray(xm, dtype=np.uint64)
        zm = np.ascontiguousarray(zm, dtype=np.uint64)
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        if lib().ph2_reg_rotate(C.c_void_p(self.h), C.c_size_t(xm.shape[0]),
                                _p(xm, _u64p), _p(zm, _u64p),
                                _p(phi, _f64p)) != 0:
            _err("ph2_reg_rotate")

    def overlap(self, idx, cf):
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        cf = np.ascontiguousarray(np.asarray(cf, np.complex128).view(np.float64))
        out = np.zeros(2)
        if lib().ph2_reg_overlap_multidet(C.c_void_p(self.h),
                                          C.c_size_t(idx.shape[0]),
                                          _p(idx, _u64p), _p(cf, _f64p),
                                          _p(out, _f64p)) != 0:
            _err("ph2_reg_overlap_multidet")
        return complex(out[0], out[1])

    def norm2(self):
        out = np.zeros(1)
        if lib().ph2_reg_norm2(C.c_void_p(self.h), _p(out, _f64p)) != 0:
            _err("ph2_reg_norm2")
        return float(out[0])

    def read(self):
        st = np.zeros(1 << self.nqb, dtype=np.complex128)
        if lib().ph2_reg_read(C.c_void_p(self.h), _p(st, _f64p)) != 0:
            _err("ph2_reg_read")
        return st

    def gather(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        out = np.zeros(idx.shape[0], dtype=np.complex128)
        L = lib()
        L.ph2_reg_gather.restype = C.c_int
        if L.ph2_reg_gather(C.c_void_p(self.h), C.c_size_t(idx.shape[0]),
                            _p(idx, _u64p), _p(out, _f64p)) != 0:
            _err("ph2_reg_gather")
        return out

    def fill_hash(self, seed):
        L = lib()
        L.ph2_reg_fill_hash.restype = C.c_int
        if L.ph2_reg_fill_hash(C.c_void_p(self.h), C.c_uint64(seed)) != 0:
            _err("ph2_reg_fill_hash")

    def count_unequal(self, other):
        n = C.c_uint64(0)
        L = lib()
        L.ph2_reg_count_unequal.restype = C.c_int
        if L.ph2_reg_count_unequal(C.c_void_p(self.h), C.c_void_p(other.h),
                                   C.byref(n)) != 0:
            _err("ph2_reg_count_unequal")
        return int(n.value)

    def free(self):
        if self.h:
            lib().ph2_reg_free(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


REC_DTYPE = np.dtype([
    ("kind", "<u4"), ("nqb", "<u4"), ("tile_bits", "<u4"), ("low_bits", "<u4"),
    ("gdim", "<u4"), ("nterms", "<u4"), ("ngroups", "<u4"), ("direct", "<u4"),
    ("pivots", "<u8"), ("xunion", "<u8"), ("zunion", "<u8"), ("amps", "<u8"),
    ("ms", "<f8")])


def prof_records(reset=True, cap=65536) -> np.ndarray:
    """Per-launch records (struct ph2hip_launch_rec) kept while profiling."""
    out = np.zeros(cap, dtype=REC_DTYPE)
    L = lib()
    L.ph2hip_prof_records.restype = C.c_long
    n = L.ph2hip_prof_records(out.ctypes.data_as(C.c_void_p), C.c_size_t(cap),
                              C.c_int(int(reset)))
    if n < 0:
        _err("ph2hip_prof_records")
    return out[:n].copy()


def set_fused_form(pipe_mask: int, persistent_wgs: int = 0):
    """Bit TB of pipe_mask: persistent pipelined fused kernel for tile TB."""
    if lib().ph2hip_set_fused_form(C.c_uint32(pipe_mask),
                                   C.c_uint64(persistent_wgs)) != 0:
        _err("ph2hip_set_fused_form")


def set_fused_direct(on: bool):
    """First/last micro-groups straight between HBM and registers."""
    if lib().ph2hip_set_fused_direct(C.c_int(int(on))) != 0:
        _err("ph2hip_set_fused_direct")


HALF_TILES = (1 << 12) | (1 << 13)   # PH2HIP_HALF_TILES: default mask


def set_fused_half(tile_mask: int):
    """Bit tb: fused passes of tile width tb in the half-tile LDS form."""
    if lib().ph2hip_set_fused_half(C.c_uint32(tile_mask)) != 0:
        _err("ph2hip_set_fused_half")


def set_plain_form(form: int):
    """2, 4 or 8 pairs per thread; +16 for non-temporal stores."""
    lib().ph2hip_set_plain_form(C.c_int(form))


# -- batched qDRIFT ------------------------------------------------------------

class BatchReg(C.Structure):
    """struct ph2_batch_reg (include/phase2_batch.h)."""
    _fields_ = [("nqb", C.c_uint32), ("nterms", C.c_size_t),
                ("xm", _u64p), ("zm", _u64p), ("cf", _f64p),
                ("ndet", C.c_size_t), ("didx", _u64p), ("dcf", _f64p),
                ("samples", C.c_size_t), ("depth", C.c_size_t),
                ("step_size", C.c_double), ("seed", C.c_uint64)]


class BatchCmpsitReg(C.Structure):
    """struct ph2_batch_cmpsit_reg (include/phase2_batch.h)."""
    _fields_ = [("nqb", C.c_uint32), ("nterms", C.c_size_t),
                ("xm", _u64p), ("zm", _u64p), ("cf", _f64p),
                ("ndet", C.c_size_t), ("didx", _u64p), ("dcf", _f64p),
                ("samples", C.c_size_t), ("steps", C.c_size_t),
                ("length", C.c_size_t), ("depth", C.c_size_t),
                ("angle_det", C.c_double), ("angle_rand", C.c_double),
                ("seed", C.c_uint64)]


class BatchStats(C.Structure):
    """struct ph2_batch_stats (include/phase2_batch.h)."""
    _fields_ = [("seconds", C.c_double), ("tune_seconds", C.c_double),
                ("prep_seconds", C.c_double), ("draw_seconds", C.c_double),
                ("amp_updates", C.c_double), ("launches", C.c_uint64),
                ("passes", C.c_uint64), ("tile_bits", C.c_uint32),
                ("low_bits", C.c_uint32)]


BATCH_PLAIN_EACH, BATCH_FUSED_EACH, BATCH_FUSED_BATCHED = 0, 1, 2


def batch_qdrift(regs, mode=BATCH_FUSED_BATCHED, tile_bits=0, low_bits=3):
    """regs: list of dicts with hm, det_idx, det_cf, samples, depth,
    step_size, seed.  Returns (list of per-register overlap arrays, stats)."""
    keep = []
    arr = (BatchReg * len(regs))()
    total = 0
    for i, r in enumerate(regs):
        hm = r["hm"]
        xm, zm, cf = _terms(hm)
        idx = np.ascontiguousarray(r["det_idx"], dtype=np.uint64)
        dcf = np.ascontiguousarray(
            np.asarray(r["det_cf"], dtype=np.complex128).view(np.float64))
        keep += [xm, zm, cf, idx, dcf]
        arr[i] = BatchReg(hm.nqb, len(cf), _p(xm, _u64p), _p(zm, _u64p),
                          _p(cf, _f64p), idx.shape[0], _p(idx, _u64p),
                          _p(dcf, _f64p), r["samples"], r["depth"],
                          r["step_size"], r["seed"])
        total += r["samples"]
    out = np.zeros(max(total, 1), dtype=np.complex128)
    st = BatchStats()
    L = lib()
    L.ph2_batch_qdrift.restype = C.c_int
    rc = L.ph2_batch_qdrift(arr, C.c_size_t(len(regs)), C.c_int(mode),
                            C.c_uint32(tile_bits), C.c_uint32(low_bits),
                            _p(out, _f64p), C.byref(st))
    if rc != 0:
        _err("ph2_batch_qdrift")
    res, o = [], 0
    for r in regs:
        res.append(out[o:o + r["samples"]].copy())
        o += r["samples"]
    stats = {f: getattr(st, f) for f, _ in BatchStats._fields_}
    return res, stats


def _batch_prep(r, keep):
    idx = np.ascontiguousarray(r["det_idx"], dtype=np.uint64)
    dcf = np.ascontiguousarray(
        np.asarray(r["det_cf"], dtype=np.complex128).view(np.float64))
    keep += [idx, dcf]
    return idx, dcf


def batch_cmpsit(regs, mode=BATCH_FUSED_BATCHED, tile_bits=0, low_bits=3):
    """regs: list of dicts with hm, det_idx, det_cf, samples, steps, length,
    depth, angle_det, angle_rand, seed (ph2_batch_cmpsit).  Returns (list of
    per-register overlap arrays, stats)."""
    keep = []
    arr = (BatchCmpsitReg * len(regs))()
    total = 0
    for i, r in enumerate(regs):
        hm = r["hm"]
        xm, zm, cf = _terms(hm)
        keep += [xm, zm, cf]
        idx, dcf = _batch_prep(r, keep)
        arr[i] = BatchCmpsitReg(
            hm.nqb, len(cf), _p(xm, _u64p), _p(zm, _u64p), _p(cf, _f64p),
            idx.shape[0], _p(idx, _u64p), _p(dcf, _f64p), r["samples"],
            r["steps"], r["length"], r["depth"], r["angle_det"],
            r["angle_rand"], r["seed"])
        total += r["samples"]
    out = np.zeros(max(total, 1), dtype=np.complex128)
    st = BatchStats()
    L = lib()
    L.ph2_batch_cmpsit.restype = C.c_int
    rc = L.ph2_batch_cmpsit(arr, C.c_size_t(len(regs)), C.c_int(mode),
                            C.c_uint32(tile_bits), C.c_uint32(low_bits),
                            _p(out, _f64p), C.byref(st))
    if rc != 0:
        _err("ph2_batch_cmpsit")
    res, o = [], 0
    for r in regs:
        res.append(out[o:o + r["samples"]].copy())
        o += r["samples"]
    stats = {f: getattr(st, f) for f, _ in BatchStats._fields_}
    return res, stats
