"""The reference's examples pipeline on its own output (no GPU): the
reference CPU build's 256-step Trotter series on the bundled Hamiltonian
(tests/golden/ph2run_h5.npz, case example_trott256) gives a ground-state
energy inside the band its CI asserts.  This pins the energy restatement in
tests/pipeline_energy.py; tests/test_gpu_ph2run.py applies it to the
series this library writes."""

import ctypes as C
import shutil
from pathlib import Path

import pytest

from pipeline_energy import BAND, fft_energy

ROOT = Path(__file__).resolve().parents[1]
H5LIB = ROOT / "phase2_amd" / "libphase2_h5.so"
FIX = Path(__file__).resolve().parent / "golden" / "h5" / "examples_hamil.h5"


def hamil_attrs(path):
    """normalization and offset of /pauli_hamil, read by the simul.h5
    layer."""
    L = C.CDLL(str(H5LIB))
    L.data_open.restype = C.c_int64
    L.data_open.argtypes = [C.c_char_p]
    L.data_close.argtypes = [C.c_int64]
    L.data_attr_read_dbl.restype = C.c_int
    fid = L.data_open(str(path).encode())
    assert fid >= 0
    out = []
    try:
        for a in (b"normalization", b"offset"):
            d = C.c_double(0.0)
            assert L.data_attr_read_dbl(C.c_int64(fid), b"pauli_hamil", a,
                                        C.byref(d)) == 0
            out.append(d.value)
    finally:
        L.data_close(C.c_int64(fid))
    return out


@pytest.mark.skipif(not H5LIB.is_file(), reason="libphase2_h5.so not built")
def test_reference_series_energy_in_ci_band(golden, tmp_path):
    g = golden("ph2run_h5.npz")
    v = g["example_trott256__values"]
    assert len(v) == 256
    p = tmp_path / "h.h5"
    shutil.copyfile(FIX, p)
    norm, offset = hamil_attrs(p)
    e = fft_energy(v, 0.1, norm, offset)
    assert BAND[0] < e < BAND[1], e
    # one FFT bin is tau / (256 * 0.1 * norm) = 3.43 Ha wide; the peak
    # sits in the bin nearest the H2O ground state
    assert abs(e - (-75.642)) < 1e-3
