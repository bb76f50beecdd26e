"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the reference CPU backend (oracle/_ref/libphase2_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the phase2_amd package.

Both libraries expose the same flat entry points (prefix ``orc_`` for the
restatement, ``refd_`` for the reference driver), so one wrapper serves both.
States are numpy complex128 arrays of length 2**nqb.
"""

from __future__ import annotations

import ctypes as C
import hashlib
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libphase2_ref.so"

_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)


class Prep(C.Structure):
    """Initial state: kind 1 = multidet, 2 = coefficient matrix."""
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


def _p(a, t):
    return a.ctypes.data_as(t)


def make_prep(det_idx=None, det_cf=None, C_alpha=None, C_beta=None,
              n_sites=0, n_alpha=0, n_beta=0, closed_shell=1, tapered=0):
    """Build a Prep plus the arrays it points into (keep both alive)."""
    keep = []
    p = Prep()
    if det_idx is not None:
        idx = np.ascontiguousarray(det_idx, dtype=np.uint64)
        cf = np.ascontiguousarray(
            np.asarray(det_cf, dtype=np.complex128).view(np.float64))
        keep += [idx, cf]
        p.kind = 1
        p.ndet = idx.shape[0]
        p.didx = _p(idx, _u64p)
        p.dcf = _p(cf, _f64p)
    else:
        ca = np.ascontiguousarray(C_alpha, dtype=np.float64).reshape(-1)
        cb = (np.ascontiguousarray(C_beta, dtype=np.float64).reshape(-1)
              if C_beta is not None else np.zeros(1))
        keep += [ca, cb]
        p.kind = 2
        p.n_sites, p.n_alpha, p.n_beta = n_sites, n_alpha, n_beta
        p.closed_shell, p.tapered = closed_shell, tapered
        p.Ca = _p(ca, _f64p)
        p.Cb = _p(cb, _f64p) if C_beta is not None else None
    return p, keep


def state_hash(state: np.ndarray) -> str:
    """sha256 of the state's bytes with -0.0 folded into +0.0."""
    s = np.ascontiguousarray(state, dtype=np.complex128).view(np.float64)
    s = s + 0.0
    return hashlib.sha256(s.tobytes()).hexdigest()


class Lib:
    """One of the two CPU libraries behind a common interface."""

    def __init__(self, path: Path, prefix: str):
        if not path.is_file():
            raise FileNotFoundError(f"{path} is not built")
        self.lib = C.CDLL(str(path))
        self.pre = prefix
        self.path = path

    def _f(self, name):
        return getattr(self.lib, self.pre + name)

    @staticmethod
    def _terms(hm):
        xm = np.ascontiguousarray(hm.xm, dtype=np.uint64)
        zm = np.ascontiguousarray(hm.zm, dtype=np.uint64)
        cf = np.ascontiguousarray(hm.cf, dtype=np.float64)
        return xm, zm, cf

    def rotate(self, state, xm, zm, phi):
        """Sequential rotations exp(i phi_k P_k), k = 0.., on a copy."""
        st = np.array(state, dtype=np.complex128, copy=True)
        nqb = int(st.shape[0]).bit_length() - 1
        xm = np.ascontiguousarray(xm, dtype=np.uint64)
        zm = np.ascontiguousarray(zm, dtype=np.uint64)
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        if self.pre == "orc_":
            f = self._f("rotate")
            f.restype = None
            f(_p(st, _f64p), C.c_uint32(nqb), C.c_size_t(xm.shape[0]),
              _p(xm, _u64p), _p(zm, _u64p), _p(phi, _f64p))
        else:
            f = self._f("paulirot")
            f.restype = C.c_int
            rc = f(C.c_uint32(nqb), _p(st, _f64p), C.c_size_t(xm.shape[0]),
                   _p(xm, _u64p), _p(zm, _u64p), _p(phi, _f64p))
            assert rc == 0
        return st

    def trott(self, algo, hm, prep_kw, delta, steps, want_state=True):
        xm, zm, cf = self._terms(hm)
        prep, keep = make_prep(**prep_kw)
        out = np.zeros(steps, dtype=np.complex128)
        st = (np.zeros(1 << hm.nqb, dtype=np.complex128) if want_state
              else None)
        f = self._f("trott")
        f.restype = C.c_int
        rc = f(C.c_int(algo), C.c_uint32(hm.nqb), C.c_size_t(len(cf)),
               _p(xm, _u64p), _p(zm, _u64p), _p(cf, _f64p), C.byref(prep),
               C.c_double(delta), C.c_size_t(steps), _p(out, _f64p),
               _p(st, _f64p) if st is not None else None)
        if rc != 0:
            raise RuntimeError(f"{self.pre}trott rc={rc}")
        return out, st

    def qdrift(self, hm, prep_kw, samples, depth, step_size, seed,
               want_state=True, want_drawn=False):
        xm, zm, cf = self._terms(hm)
        prep, keep = make_prep(**prep_kw)
        out = np.zeros(samples, dtype=np.complex128)
        st = (np.zeros(1 << hm.nqb, dtype=np.complex128) if want_state
              else None)
        f = self._f("qdrift")
        f.restype = C.c_int
        args = [C.c_uint32(hm.nqb), C.c_size_t(len(cf)), _p(xm, _u64p),
                _p(zm, _u64p), _p(cf, _f64p), C.byref(prep),
                C.c_size_t(samples), C.c_size_t(depth),
                C.c_double(step_size), C.c_uint64(seed), _p(out, _f64p),
                _p(st, _f64p) if st is not None else None]
        drawn = None
        if self.pre == "orc_":
            drawn = (np.zeros(samples * depth, dtype=np.uint64)
                     if want_drawn else None)
            args.append(_p(drawn, _u64p) if drawn is not None else None)
        rc = f(*args)
        if rc != 0:
            raise RuntimeError(f"{self.pre}qdrift rc={rc}")
        return (out, st, drawn) if want_drawn else (out, st)

    def cmpsit(self, hm, prep_kw, samples, steps, length, depth, angle_det,
               angle_rand, seed, want_state=True):
        xm, zm, cf = self._terms(hm)
        prep, keep = make_prep(**prep_kw)
        out = np.zeros(samples, dtype=np.complex128)
        st = (np.zeros(1 << hm.nqb, dtype=np.complex128) if want_state
              else None)
        f = self._f("cmpsit")
        f.restype = C.c_int
        rc = f(C.c_uint32(hm.nqb), C.c_size_t(len(cf)), _p(xm, _u64p),
               _p(zm, _u64p), _p(cf, _f64p), C.byref(prep),
               C.c_size_t(samples), C.c_size_t(steps), C.c_size_t(length),
               C.c_size_t(depth), C.c_double(angle_det),
               C.c_double(angle_rand), C.c_uint64(seed), _p(out, _f64p),
               _p(st, _f64p) if st is not None else None)
        if rc != 0:
            raise RuntimeError(f"{self.pre}cmpsit rc={rc}")
        return out, st

    def expand(self, nqb, n_sites, n_alpha, n_beta, closed_shell, tapered,
               Ca, Cb=None):
        st = np.zeros(1 << nqb, dtype=np.complex128)
        ca = np.ascontiguousarray(Ca, dtype=np.float64).reshape(-1)
        cb = (np.ascontiguousarray(Cb, dtype=np.float64).reshape(-1)
              if Cb is not None else ca)
        f = self._f("expand")
        f.restype = C.c_int
        rc = f(C.c_uint32(nqb), C.c_uint32(n_sites), C.c_uint32(n_alpha),
               C.c_uint32(n_beta), C.c_int(closed_shell), C.c_int(tapered),
               _p(ca, _f64p), _p(cb, _f64p), _p(st, _f64p))
        if rc != 0:
            raise RuntimeError(f"{self.pre}expand rc={rc}")
        return st

    def inner(self, state, n_sites, n_alpha, n_beta, closed_shell, tapered,
              Ca, Cb=None):
        st = np.ascontiguousarray(state, dtype=np.complex128)
        nqb = int(st.shape[0]).bit_length() - 1
        ca = np.ascontiguousarray(Ca, dtype=np.float64).reshape(-1)
        cb = (np.ascontiguousarray(Cb, dtype=np.float64).reshape(-1)
              if Cb is not None else ca)
        out = np.zeros(2)
        f = self._f("inner")
        f.restype = C.c_int
        rc = f(C.c_uint32(nqb), C.c_uint32(n_sites), C.c_uint32(n_alpha),
               C.c_uint32(n_beta), C.c_int(closed_shell), C.c_int(tapered),
               _p(ca, _f64p), _p(cb, _f64p), _p(st, _f64p), _p(out, _f64p))
        if rc != 0:
            raise RuntimeError(f"{self.pre}inner rc={rc}")
        return complex(out[0], out[1])

    def xoshiro(self, seed, n):
        out = np.zeros(n, dtype=np.uint64)
        f = self._f("xoshiro")
        f.restype = None
        f(C.c_uint64(seed), C.c_size_t(n), _p(out, _u64p))
        return out

    def cdf_inverse(self, w, ys):
        w = np.ascontiguousarray(w, dtype=np.float64)
        ys = np.ascontiguousarray(ys, dtype=np.float64)
        idx = np.zeros(ys.shape[0], dtype=np.uint64)
        lam = C.c_double(0.0)
        f = self._f("cdf_inverse")
        f.restype = C.c_int
        rc = f(C.c_size_t(w.shape[0]), _p(w, _f64p), C.c_size_t(ys.shape[0]),
               _p(ys, _f64p), _p(idx, _u64p), C.byref(lam))
        if rc != 0:
            raise RuntimeError("cdf_inverse failed")
        return idx, lam.value

    def sort_lex(self, xm, zm):
        xm = np.ascontiguousarray(xm, dtype=np.uint64)
        zm = np.ascontiguousarray(zm, dtype=np.uint64)
        perm = np.zeros(xm.shape[0], dtype=np.uint64)
        f = self._f("sort_lex")
        f.restype = C.c_int
        rc = f(C.c_size_t(xm.shape[0]), _p(xm, _u64p), _p(zm, _u64p),
               _p(perm, _u64p))
        assert rc == 0
        return perm

    def det(self, A):
        A = np.ascontiguousarray(A, dtype=np.float64)
        f = self._f("det")
        f.restype = C.c_double
        return f(_p(A, _f64p), C.c_uint32(A.shape[0]))


def oracle() -> Lib:
    return Lib(ORACLE_SO, "orc_")


def reference() -> Lib:
    return Lib(REF_SO, "refd_")


def phase2_run(nqb, idx, xm, zm, coeffs, delta):
    """Oracle phase2_run on masks (the string layer is the product's)."""
    lib = oracle().lib
    xm = np.ascontiguousarray(xm, dtype=np.uint64)
    zm = np.ascontiguousarray(zm, dtype=np.uint64)
    cf = np.ascontiguousarray(coeffs, dtype=np.float64)
    out = np.zeros(2)
    lib.orc_phase2_run.restype = C.c_int
    rc = lib.orc_phase2_run(C.c_uint32(nqb), C.c_uint64(idx),
                            C.c_size_t(xm.shape[0]), _p(xm, _u64p),
                            _p(zm, _u64p), _p(cf, _f64p), C.c_double(delta),
                            _p(out, _f64p))
    if rc != 0:
        raise ValueError("orc_phase2_run failed")
    return complex(out[0], out[1])


def build():
    """Compile liboracle.so (and the reference library when its sources are
    present in this container)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if Path(os.environ.get("PH2_REFERENCE", "/root/reference")).is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "-f", "Makefile.ref"],
                       check=True)
