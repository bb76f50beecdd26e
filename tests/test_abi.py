"""The C-ABI library loads and exports every function the public headers
declare (no GPU call is made)."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "phase2_amd" / "libphase2_hip.so"

DECL = re.compile(r"^(?!static|typedef|#|struct \w+ \{|enum)"
                  r"[A-Za-z_][\w \t\*]*?\b([a-z_][a-z0-9_]*)\s*\(", re.M)


def declared():
    names = set()
    for h in sorted((ROOT / "include").rglob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in DECL.finditer(text):
            name = m.group(1)
            if name not in {"if", "while", "for", "sizeof", "return"}:
                names.add((h.relative_to(ROOT).as_posix(), name))
    return sorted(names)


def test_headers_declare_functions():
    names = {n for _, n in declared()}
    for must in ("qreg_init", "qreg_paulirot", "circ_step", "circ_measure",
                 "state_prep_coeff_expand_all", "trott_simul",
                 "qdrift_simul", "cmpsit_simul", "phase2_run",
                 "ph2hip_rotate_fused", "ph2_trott"):
        assert must in names


@pytest.fixture(scope="module")
def lib():
    if not LIB.is_file():
        pytest.fail(f"{LIB} not built (run __graft_entry__.build())")
    return ctypes.CDLL(str(LIB))


# headers whose functions live in the simul.h5 library (libphase2_h5.so)
H5_HEADERS = ("include/ph2run/data.h", "include/phase2_h5.h")
H5LIB = ROOT / "phase2_amd" / "libphase2_h5.so"


@pytest.mark.parametrize("header,name", declared())
def test_symbol_exported(lib, header, name):
    if header in H5_HEADERS:
        if not H5LIB.is_file():
            pytest.skip("libphase2_h5.so not built (no HDF5 headers)")
        lib = ctypes.CDLL(str(H5LIB))
    assert hasattr(lib, name), f"{name} from {header} not exported"


def test_no_undefined_library_symbols():
    out = subprocess.run(["nm", "-D", "--undefined-only", str(LIB)],
                         capture_output=True, text=True, check=True).stdout
    bad = [ln for ln in out.splitlines()
           if not re.search(r"@(GLIBC|GCC|hip|CXXABI|GLIBCXX)", ln)
           and not ln.strip().startswith("w ")]
    assert not bad, bad


def test_dropin_library_loads():
    """The drop-in check library (the reference's drivers, unchanged, on
    this library) loads without a GPU and exports its entry points and the
    reference driver functions."""
    import ctypes as C
    from pathlib import Path
    p = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / \
        "libcirc_dropin.so"
    if not p.is_file():
        pytest.skip("oracle/_ref/libcirc_dropin.so not built")
    lib = C.CDLL(str(p))
    for f in ("dropin_trott", "dropin_qdrift", "dropin_cmpsit", "trott_init",
              "trott_simul", "trott2_simul", "qdrift_simul", "cmpsit_simul"):
        assert hasattr(lib, f), f
