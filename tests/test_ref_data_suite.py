"""The reference's own tests of its simul.h5 layer, run on this library:
test/t-data_{attr,coeff_matrix,dets_validate,hamil,multidet,open,
idempotence,trott_steps}.c compiled unchanged against include/ and linked to
libphase2_h5.so and libphase2_hip.so (oracle/_ref/t-data_*, built by
oracle/Makefile.ref from the reference sources).  Each runs in a scratch
directory whose test/data holds copies of the reference's fixture files
(tests/golden/h5/), as the reference's own runner lays them out, and must
exit 0.  These tests need no GPU: the data layer is host code.  (The two
t-data tests that call MPI collectives on the reference's MPI world are
not built.)"""

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = ROOT / "oracle" / "_ref"
FIX = Path(__file__).resolve().parent / "golden" / "h5"
TESTS = ["attr", "coeff_matrix", "dets_validate", "hamil", "multidet",
         "open", "idempotence", "trott_steps"]


@pytest.mark.parametrize("name", TESTS)
def test_reference_data_test(tmp_path, name):
    exe = REF / f"t-data_{name}"
    if not exe.is_file():
        pytest.skip(f"{exe.name} not built (needs the reference sources)")
    data = tmp_path / "test" / "data"
    data.mkdir(parents=True)
    for f in FIX.iterdir():
        shutil.copyfile(f, data / f.name)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, (r.stdout + r.stderr)[-2000:]
