"""The reference's own unit tests, run on this library: test/t-*.c compiled
unchanged against include/ and linked to libphase2_h5.so and
libphase2_hip.so (oracle/_ref/t-*, built by oracle/Makefile.ref from the
reference sources, with the 1e-14 tolerance the reference's tests give its
CPU and CUDA backends).  Each runs in a scratch directory
whose test/data holds copies of the reference's fixture files
(tests/golden/h5/), laid out as the reference's own runner expects, and
must exit 0.

The host-only tests (the simul.h5 layer, Pauli algebra, CDFs,
combinations, determinants, the circ data carriers) run without a GPU; the
rest build registers and run on the device.  Not built: the tests of the
reference's MPI collectives, logging subsystem, bench framework and test
runner."""

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = ROOT / "oracle" / "_ref"
FIX = Path(__file__).resolve().parent / "golden" / "h5"

HOST = ["data_attr", "data_coeff_matrix", "data_dets_validate", "data_hamil",
        "data_multidet", "data_open", "data_idempotence", "data_trott_steps",
        "bitstring_index", "circ", "circ_cache", "combinations", "det_small",
        "paulis", "prob"]
DEVICE = ["world", "qreg", "circ_prepst_coeff", "circ_trott", "circ_trott2",
          "circ_trott_coeff", "circ_trott2_coeff", "cmpsit", "qdrift",
          "ref-bendazzoli", "state_prep_coeff_csf", "state_prep_coeff_expand",
          "state_prep_coeff_large", "state_prep_coeff_sync"]


def _run(tmp_path, name):
    exe = REF / f"t-{name}"
    if not exe.is_file():
        pytest.skip(f"{exe.name} not built (needs the reference sources)")
    data = tmp_path / "test" / "data"
    shutil.copytree(FIX, data)
    # the tests open some fixtures read-write, as the reference's own runner
    # leaves them; the sources may be read-only
    for p in [data, *data.rglob("*")]:
        p.chmod(0o755 if p.is_dir() else 0o644)
    print(f"reference test: t-{name}", flush=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]


@pytest.mark.parametrize("name", HOST)
def test_reference_host_test(tmp_path, name):
    _run(tmp_path, name)


@pytest.mark.gpu
@pytest.mark.parametrize("name", DEVICE)
def test_reference_device_test(tmp_path, name):
    _run(tmp_path, name)
