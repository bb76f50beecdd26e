"""Golden outputs of the simul.h5 path (ph2run), from the reference's own
build.

Copies of the reference's test fixtures (test/data/*.h5; data files), and
of the Hamiltonian of its examples pipeline (examples/data/hamil.h5, kept as
examples_hamil.h5), are kept in tests/golden/h5/.  For each case below the reference's ph2run
(oracle/_ref/ph2run_ref, built from its sources by oracle/Makefile.ref) runs
on a fresh copy, and the per-step values it writes are read back with
h5dump (17 significant digits, so every double is exact) into
tests/golden/ph2run_h5.npz, together with each case's command line.  A case
the reference rejects is recorded with status 1 and no values.  The
datasets and attributes of every fixture, read the same way, go to
tests/golden/h5_contents.npz (key "<fixture>:<path>") for the loader tests.

    python tests/golden/make_h5_golden.py     (needs /root/reference)
"""

import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_DATA = Path("/root/reference/test/data")
EXAMPLE_DATA = Path("/root/reference/examples/data")
PH2RUN = HERE.parents[1] / "oracle" / "_ref" / "ph2run_ref"
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"

FIXTURES = ["N4_multidet.h5", "N4_closed.h5", "N4_open.h5", "N4_csf.h5",
            "N8_tapered.h5", "N8_untapered.h5", "H2O_CAS56.h5",
            "N4_both.h5", "N4_csf_empty.h5"]
# fixtures that only the reference's own t-data_* tests read (copied, not
# dumped)
COPY_ONLY = ["case-d9f603dc.h5_solved"]
COPY_TREES = ["bendazzoli"]

# (name, fixture, group, ph2run arguments after -S FILE)
CASES = [
    ("md_trott", "N4_multidet.h5", "circ_trott",
     ["trott", "--delta", "0.3", "--steps", "5"]),
    ("md_trott2", "N4_multidet.h5", "circ_trott2",
     ["trott2", "--delta", "0.2", "--steps", "4"]),
    ("md_qdrift", "N4_multidet.h5", "circ_qdrift",
     ["qdrift", "--step-size", "0.05", "--depth", "40", "--samples", "3",
      "--seed", "7"]),
    ("md_cmpsit", "N4_multidet.h5", "circ_cmpsit",
     ["cmpsit", "--length", "1", "--depth", "12", "--steps", "3",
      "--angle-det", "0.1", "--angle-rand", "0.05", "--samples", "2",
      "--seed", "9"]),
    ("closed_trott", "N4_closed.h5", "circ_trott",
     ["trott", "--delta", "0.3", "--steps", "4"]),
    # state-prep fixtures without /pauli_hamil: ph2run rejects them
    ("open_no_hamil_rejected", "N4_open.h5", "circ_trott2",
     ["trott2", "--delta", "0.25", "--steps", "3"]),
    ("csf_no_hamil_rejected", "N4_csf.h5", "circ_trott",
     ["trott", "--delta", "0.3", "--steps", "3"]),
    ("tapered_trott", "N8_tapered.h5", "circ_trott",
     ["trott", "--delta", "0.2", "--steps", "3"]),
    ("untapered_trott", "N8_untapered.h5", "circ_trott",
     ["trott", "--delta", "0.2", "--steps", "3"]),
    ("h2o_trott", "H2O_CAS56.h5", "circ_trott",
     ["trott", "--delta", "0.1", "--steps", "4"]),
    ("both_rejected", "N4_both.h5", "circ_trott",
     ["trott", "--delta", "0.3", "--steps", "2"]),
    ("csf_empty_rejected", "N4_csf_empty.h5", "circ_trott",
     ["trott", "--delta", "0.3", "--steps", "2"]),
    # the examples pipeline as the reference's CI runs it
    # (examples/simul/trott/Makefile, CI.yml: TROTT_STEPS=256)
    ("example_trott256", "examples_hamil.h5", "circ_trott",
     ["trott", "--delta", "0.1", "--steps", "256"]),
]


def source(fixture):
    if fixture == "examples_hamil.h5":
        return EXAMPLE_DATA / "hamil.h5"
    return REF_DATA / fixture


def dump_values(path, grp):
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", f"/{grp}/values",
                          str(path)], capture_output=True, text=True,
                         check=True).stdout
    body = out.split("DATA {", 1)[1].rsplit("}", 2)[0]
    vals = []
    for line in body.splitlines():
        if ":" in line:
            vals += [float(v) for v in line.split(":", 1)[1].split(",")
                     if v.strip()]
    a = np.array(vals).reshape(-1, 2)
    return a[:, 0] + 1j * a[:, 1]


def dump_all(path):
    """Every dataset and attribute of a file, as flat float arrays."""
    out = subprocess.run([H5DUMP, "-m", "%.17g", str(path)],
                         capture_output=True, text=True, check=True).stdout
    res, stack, cur, buf = {}, [], None, None
    for line in out.splitlines():
        t = line.strip()
        if t.startswith(("GROUP", "DATASET", "ATTRIBUTE")):
            kind, name = t.split(" ", 1)
            name = name.split('"')[1]
            stack.append((kind, name))
        elif t.startswith("DATA {"):
            cur = "/".join(n for _, n in stack if n != "/")
            buf = []
        elif cur is not None and t == "}":
            if buf is not None:           # strings are left out
                res[cur] = np.array(buf, dtype=np.float64)
            cur = None
        elif cur is not None and ":" in t and buf is not None:
            try:
                buf += [float(v) for v in t.split(":", 1)[1].split(",")
                        if v.strip()]
            except ValueError:
                buf = None
        elif t == "}" and stack:
            stack.pop()
    return res


def main():
    for f in FIXTURES + COPY_ONLY + ["examples_hamil.h5"]:
        shutil.copyfile(source(f), HERE / "h5" / f)
    for d in COPY_TREES:
        shutil.copytree(REF_DATA / d, HERE / "h5" / d, dirs_exist_ok=True)
    res = {}
    with tempfile.TemporaryDirectory() as td:
        for name, fix, grp, args in CASES:
            p = Path(td) / f"{name}.h5"
            shutil.copyfile(source(fix), p)
            r = subprocess.run([str(PH2RUN), "-S", str(p)] + args,
                               capture_output=True, text=True)
            res[f"{name}__status"] = np.array(0 if r.returncode == 0 else 1)
            res[f"{name}__fixture"] = np.array(fix)
            res[f"{name}__group"] = np.array(grp)
            res[f"{name}__args"] = np.array(args)
            if r.returncode == 0:
                res[f"{name}__values"] = dump_values(p, grp)
            print(name, "rc", r.returncode,
                  res.get(f"{name}__values", np.zeros(0))[:2], flush=True)
    np.savez(HERE / "ph2run_h5.npz", **res)
    contents = {}
    for f in FIXTURES:
        for k, v in dump_all(REF_DATA / f).items():
            contents[f"{f}:{k}"] = v
    np.savez(HERE / "h5_contents.npz", **contents)


if __name__ == "__main__":
    sys.exit(main())
