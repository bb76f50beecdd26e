"""ph2run, the reference's command-line driver, dropped in: its ph2run.c and
circ/*.c compiled unchanged against this library's include/ and linked to
libphase2_h5.so (the simul.h5 layer) and libphase2_hip.so
(oracle/_ref/ph2run_hip, built by oracle/Makefile.ref from the reference
sources).  On copies of the reference's fixture files it writes the same
per-step values as the reference's own CPU build did
(tests/golden/ph2run_h5.npz): bit for bit for multideterminant states, to
1e-12 relative for coefficient-matrix states (their overlap is a device
tree reduction); files the reference rejects are rejected too.  The
256-step series on the bundled Hamiltonian of the reference's examples
pipeline also gives a ground-state energy inside the band the reference's
CI asserts."""

import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from pipeline_energy import BAND, fft_energy
from test_example_pipeline import hamil_attrs

ROOT = Path(__file__).resolve().parents[1]
PH2RUN = ROOT / "oracle" / "_ref" / "ph2run_hip"
H5LIB = ROOT / "phase2_amd" / "libphase2_h5.so"
FIX = Path(__file__).resolve().parent / "golden" / "h5"

pytestmark = pytest.mark.gpu

CASES = [c.split("__")[0] for c in np.load(
    Path(__file__).resolve().parent / "golden" / "ph2run_h5.npz",
    allow_pickle=False).files if c.endswith("__status")]


def _values(path, grp):
    L = C.CDLL(str(H5LIB))
    L.ph2_h5_read_values.restype = C.c_int
    out = np.zeros(2 * 256)
    n = C.c_size_t(0)
    rc = L.ph2_h5_read_values(str(path).encode(), grp.encode(),
                              out.ctypes.data_as(C.POINTER(C.c_double)),
                              C.c_size_t(256), C.byref(n))
    assert rc == 0
    v = out[:2 * n.value].reshape(-1, 2)
    return v[:, 0] + 1j * v[:, 1]


@pytest.mark.parametrize("case", CASES)
def test_ph2run_dropin(golden, tmp_path, case):
    if not PH2RUN.is_file() or not H5LIB.is_file():
        pytest.skip("ph2run_hip or libphase2_h5.so not built")
    g = golden("ph2run_h5.npz")
    fix, grp = str(g[f"{case}__fixture"]), str(g[f"{case}__group"])
    args = [str(a) for a in g[f"{case}__args"]]
    p = tmp_path / fix
    shutil.copyfile(FIX / fix, p)
    print(f"ph2run: {case}: {fix} {' '.join(args)}", flush=True)
    r = subprocess.run([str(PH2RUN), "-S", str(p)] + args,
                       capture_output=True, text=True, timeout=120)
    if int(g[f"{case}__status"]) != 0:
        assert r.returncode != 0, r.stdout[-500:]
        return
    assert r.returncode == 0, (r.stdout + r.stderr)[-2000:]
    got, want = _values(p, grp), g[f"{case}__values"]
    assert got.shape == want.shape
    if fix in ("N4_multidet.h5", "H2O_CAS56.h5", "examples_hamil.h5"):
        assert np.array_equal(got, want), (got, want)   # multidet states
    else:
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    if case == "example_trott256":
        # the reference CI's acceptance of its examples pipeline
        norm, offset = hamil_attrs(p)
        e = fft_energy(got, 0.1, norm, offset)
        assert BAND[0] < e < BAND[1], e
        assert e == fft_energy(want, 0.1, norm, offset)
