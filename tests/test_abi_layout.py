"""Struct layout of the public C API against the reference's headers: the
same program (tests/abi/layout.c) prints every struct's size and every
member's offset and size, compiled once against this library's include/
and once against the reference's include/ (with the MPI headers it needs).
Identical output means code compiled against the reference headers -- the
algorithm drivers, ph2run -- lays out struct qreg, struct circ and the
driver structs exactly as this library does.  Needs the reference sources
and the MPI headers of /opt/conda (this container only)."""

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
MPI_INC = Path("/opt/conda/include")


def _layout(tmp_path, name, incs):
    exe = tmp_path / name
    subprocess.run(["gcc", "-std=c11", *[f"-I{i}" for i in incs],
                    str(ROOT / "tests" / "abi" / "layout.c"), "-o", str(exe)],
                   check=True)
    return subprocess.run([str(exe)], check=True, capture_output=True,
                          text=True).stdout.splitlines()


@pytest.mark.skipif(not (REF / "include").is_dir() or
                    not (MPI_INC / "mpi.h").is_file(),
                    reason="reference headers or MPI headers absent")
def test_struct_layout_matches_reference(tmp_path):
    ours = _layout(tmp_path, "ours", [ROOT / "include"])
    ref = _layout(tmp_path, "ref", [REF / "include", MPI_INC])
    assert len(ours) > 100
    assert ours == ref
