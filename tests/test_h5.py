"""The simul.h5 layer (libphase2_h5.so, include/ph2run/data.h) against the
reference's own fixture files (copies in tests/golden/h5/) and their
contents as h5dump prints them (tests/golden/h5_contents.npz, written by
tests/golden/make_h5_golden.py): Hamiltonian, multidet and
coefficient-matrix loads, the state-prep dispatch and its rejections, the
per-step writer, attribute writes and the dangling-link case.  No GPU."""

import ctypes as C
import math
import shutil
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
H5LIB = ROOT / "phase2_amd" / "libphase2_h5.so"
FIX = Path(__file__).resolve().parent / "golden" / "h5"

pytestmark = pytest.mark.skipif(not H5LIB.is_file(),
                                reason="libphase2_h5.so not built")


class Paulis(C.Structure):
    _fields_ = [("pak", C.c_uint64 * 2)]


class Term(C.Structure):
    _fields_ = [("cf", C.c_double), ("op", Paulis)]


class Hamil(C.Structure):
    _fields_ = [("qb", C.c_uint32), ("terms", C.POINTER(Term)),
                ("len", C.c_size_t)]


class Det(C.Structure):
    _fields_ = [("idx", C.c_uint64), ("cf", C.c_double * 2)]


class Muldet(C.Structure):
    _fields_ = [("dets", C.POINTER(Det)), ("len", C.c_size_t)]


class Block(C.Structure):
    _fields_ = [("cf", C.c_double), ("C_alpha", C.POINTER(C.c_double)),
                ("C_beta", C.POINTER(C.c_double))]


class CoeffMatrix(C.Structure):
    _fields_ = [("nqb", C.c_uint32), ("n_sites", C.c_uint32),
                ("n_alpha", C.c_uint32), ("n_beta", C.c_uint32),
                ("closed_shell", C.c_int), ("tapered", C.c_int),
                ("n_components", C.c_size_t),
                ("C_alpha", C.POINTER(C.c_double)),
                ("C_beta", C.POINTER(C.c_double)),
                ("blocks", C.POINTER(Block))]


class Writer(C.Structure):
    _fields_ = [("fid", C.c_int64), ("dset", C.c_int64),
                ("mspace", C.c_int64), ("n_steps", C.c_size_t)]


@pytest.fixture(scope="module")
def h5():
    L = C.CDLL(str(H5LIB))
    L.data_open.restype = C.c_int64
    L.data_open.argtypes = [C.c_char_p]
    L.data_close.argtypes = [C.c_int64]
    for f in ("data_hamil_load", "data_muldet_load", "data_state_prep_kind",
              "data_coeff_matrix_load", "data_grp_create",
              "data_attr_write_dbl", "data_attr_read_dbl",
              "data_circ_writer_init", "data_circ_write_step",
              "ph2_h5_read_values"):
        getattr(L, f).restype = C.c_int
    # _Complex double by value: two doubles in SSE registers (x86-64 SysV)
    L.data_circ_write_step.argtypes = [C.POINTER(Writer), C.c_size_t,
                                       C.c_double, C.c_double]
    return L


@pytest.fixture(scope="module")
def contents(golden):
    return golden("h5_contents.npz")


def _open(h5, tmp_path, name):
    p = tmp_path / name
    shutil.copyfile(FIX / name, p)
    fid = h5.data_open(str(p).encode())
    assert fid >= 0
    return fid, p


@pytest.mark.parametrize("name", ["N4_multidet.h5", "N4_closed.h5",
                                  "N8_tapered.h5", "H2O_CAS56.h5"])
def test_hamil_load(h5, contents, tmp_path, name):
    fid, _ = _open(h5, tmp_path, name)
    hm = Hamil()
    try:
        assert h5.data_hamil_load(C.c_int64(fid), C.byref(hm)) == 0
    finally:
        h5.data_close(C.c_int64(fid))
    cf = contents[f"{name}:pauli_hamil/coeffs"]
    norm = contents[f"{name}:pauli_hamil/normalization"][0]
    pb = contents[f"{name}:pauli_hamil/paulis"].astype(int).reshape(
        len(cf), -1)
    assert hm.len == len(cf) and hm.qb == pb.shape[1]
    for t in range(len(cf)):
        x = sum(1 << q for q in range(pb.shape[1]) if pb[t, q] in (1, 2))
        z = sum(1 << q for q in range(pb.shape[1]) if pb[t, q] in (2, 3))
        assert hm.terms[t].op.pak[0] == x and hm.terms[t].op.pak[1] == z
        assert hm.terms[t].cf == cf[t] * norm
    C.CDLL(str(ROOT / "phase2_amd" / "libphase2_hip.so")).circ_hamil_free(
        C.byref(hm))


@pytest.mark.parametrize("name", ["N4_multidet.h5", "H2O_CAS56.h5"])
def test_muldet_load(h5, contents, tmp_path, name):
    fid, _ = _open(h5, tmp_path, name)
    md = Muldet()
    kind = C.c_int(0)
    try:
        assert h5.data_state_prep_kind(C.c_int64(fid), C.byref(kind)) == 0
        assert kind.value == 1                       # STPREP_MULTIDET
        assert h5.data_muldet_load(C.c_int64(fid), C.byref(md)) == 0
    finally:
        h5.data_close(C.c_int64(fid))
    cf = contents[f"{name}:state_prep/multidet/coeffs"].reshape(-1, 2)
    dets = contents[f"{name}:state_prep/multidet/dets"].astype(int).reshape(
        len(cf), -1)
    assert md.len == len(cf)
    for d in range(len(cf)):
        assert md.dets[d].idx == sum(int(b) << q for q, b in
                                     enumerate(dets[d]))
        assert list(md.dets[d].cf) == list(cf[d])


@pytest.mark.parametrize("name", ["N4_closed.h5", "N4_open.h5", "N4_csf.h5",
                                  "N8_tapered.h5", "N8_untapered.h5"])
def test_coeff_matrix_load(h5, contents, tmp_path, name):
    fid, _ = _open(h5, tmp_path, name)
    cm = CoeffMatrix()
    kind = C.c_int(0)
    try:
        assert h5.data_state_prep_kind(C.c_int64(fid), C.byref(kind)) == 0
        assert kind.value == 2                       # STPREP_COEFF_MATRIX
        assert h5.data_coeff_matrix_load(C.c_int64(fid), C.byref(cm)) == 0
    finally:
        h5.data_close(C.c_int64(fid))
    g = f"{name}:state_prep/coeff_matrix"
    for f in ("n_qubits", "n_sites", "n_alpha", "n_beta", "closed_shell",
              "tapered"):
        want = int(contents[f"{g}/{f}"][0])
        got = getattr(cm, "nqb" if f == "n_qubits" else f)
        assert got == want, f
    ns = cm.n_sites

    def arr(p, nocc):
        return np.ctypeslib.as_array(p, shape=(ns * nocc,))
    csf = f"{g}/csf/n_components" in contents.files
    if not csf:
        assert cm.n_components == 0
        assert np.array_equal(arr(cm.C_alpha, cm.n_alpha),
                              contents[f"{g}/C_alpha"])
        if not cm.closed_shell:
            assert np.array_equal(arr(cm.C_beta, cm.n_beta),
                                  contents[f"{g}/C_beta"])
        else:
            assert not cm.C_beta
    else:
        k = int(contents[f"{g}/csf/n_components"][0])
        assert cm.n_components == k
        for b in range(k):
            assert cm.blocks[b].cf == contents[f"{g}/csf/{b}/coefficient"][0]
            assert np.array_equal(arr(cm.blocks[b].C_alpha, cm.n_alpha),
                                  contents[f"{g}/csf/{b}/C_alpha"])
    C.CDLL(str(ROOT / "phase2_amd" / "libphase2_hip.so")
           ).data_coeff_matrix_free(C.byref(cm))


@pytest.mark.parametrize("name,fn", [("N4_both.h5", "kind"),
                                     ("N4_csf_empty.h5", "coeff")])
def test_rejected_state_preps(h5, tmp_path, name, fn):
    fid, _ = _open(h5, tmp_path, name)
    try:
        if fn == "kind":
            kind = C.c_int(7)
            assert h5.data_state_prep_kind(C.c_int64(fid),
                                           C.byref(kind)) == -1
            assert kind.value == 7                   # unchanged
        else:
            cm = CoeffMatrix()
            assert h5.data_coeff_matrix_load(C.c_int64(fid),
                                             C.byref(cm)) == -1
            assert not cm.blocks and not cm.C_alpha
    finally:
        h5.data_close(C.c_int64(fid))


def test_writer_nan_rows_and_reuse(h5, tmp_path):
    fid, p = _open(h5, tmp_path, "N4_multidet.h5")
    w = Writer()
    try:
        assert h5.data_circ_writer_init(C.c_int64(fid), b"circ_trott",
                                        C.c_size_t(4), C.byref(w)) == 0
        assert h5.data_circ_write_step(C.byref(w), 0,
                                       0.5, -0.25) == 0
        assert h5.data_circ_write_step(C.byref(w), 2,
                                       1.0, 2.0) == 0
        assert h5.data_circ_write_step(C.byref(w), 4,
                                       1.0, 2.0) == -1
        h5.data_circ_writer_close(C.byref(w))
        # a second run on the same file reuses the dataset
        assert h5.data_circ_writer_init(C.c_int64(fid), b"circ_trott",
                                        C.c_size_t(4), C.byref(w)) == 0
        assert h5.data_circ_write_step(C.byref(w), 1,
                                       3.0, 4.0) == 0
        h5.data_circ_writer_close(C.byref(w))
        # fid 0: a disabled writer
        assert h5.data_circ_writer_init(C.c_int64(0), b"x", C.c_size_t(2),
                                        C.byref(w)) == 0
        assert h5.data_circ_write_step(C.byref(w), 0,
                                       1.0, 1.0) == 0
    finally:
        h5.data_close(C.c_int64(fid))
    out = np.zeros(8)
    n = C.c_size_t(0)
    assert h5.ph2_h5_read_values(str(p).encode(), b"circ_trott",
                                 out.ctypes.data_as(C.POINTER(C.c_double)),
                                 C.c_size_t(4), C.byref(n)) == 0
    assert n.value == 4
    v = out.reshape(4, 2)
    assert list(v[0]) == [0.5, -0.25] and list(v[1]) == [3.0, 4.0]
    assert list(v[2]) == [1.0, 2.0] and all(math.isnan(x) for x in v[3])


def test_dangling_link_and_attributes(h5, tmp_path):
    """H2O_CAS56.h5 carries /circ_trott as a soft link to a missing
    group: creating the group replaces it.  An attribute is written once
    (a second write of the same name fails, as in the reference)."""
    fid, _ = _open(h5, tmp_path, "H2O_CAS56.h5")
    try:
        assert h5.data_grp_create(C.c_int64(fid), b"circ_trott") == 0
        assert h5.data_grp_create(C.c_int64(fid), b"circ_trott") == 0
        assert h5.data_attr_write_dbl(C.c_int64(fid), b"circ_trott",
                                      b"delta", C.c_double(0.125)) == 0
        assert h5.data_attr_write_dbl(C.c_int64(fid), b"circ_trott",
                                      b"delta", C.c_double(0.5)) == -1
        d = C.c_double(0.0)
        assert h5.data_attr_read_dbl(C.c_int64(fid), b"circ_trott",
                                     b"delta", C.byref(d)) == 0
        assert d.value == 0.125
    finally:
        h5.data_close(C.c_int64(fid))
    assert h5.data_open(str(tmp_path / "missing.h5").encode()) == -1
