"""phase2_amd: the phase2 state-vector hot path on one AMD MI355X (gfx950).

``run(paulis, coeffs, delta, psi)`` is the drop-in for ``phase2.run`` of the
reference Python package (python/phase2/__init__.py:71-131): it computes
<psi| prod_k exp(i*delta*coeffs[k]*P_k) |psi> through the C entry
``phase2_run`` of libphase2_hip.so, with Pauli strings in "X0 Z1" notation and
``psi`` a '0'/'1' string (character k is qubit k).  Invalid input raises
ValueError, as there.
"""

from __future__ import annotations

import ctypes

from . import _native

__all__ = ["run"]


def run(paulis: list[str], coeffs: list[float], delta: float,
        psi: str) -> complex:
    """Compute <psi| prod_k exp(i*delta*coeffs[k]*P_k) |psi> on the GPU.

    Raises ValueError if len(paulis) != len(coeffs) or the C entry point
    rejects the input (bad operator, qubit out of range, bad bit string,
    more than 30 qubits).
    """
    nterms = len(paulis)
    if nterms != len(coeffs):
        raise ValueError(
            f"len(paulis)={nterms} != len(coeffs)={len(coeffs)}")
    lib = _native.lib()
    c_paulis = (ctypes.c_char_p * nterms)(
        *(s.encode("ascii") for s in paulis))
    c_coeffs = (ctypes.c_double * nterms)(*coeffs)
    out_re = ctypes.c_double(0.0)
    out_im = ctypes.c_double(0.0)
    rc = lib.phase2_run(c_paulis, c_coeffs, ctypes.c_size_t(nterms),
                        ctypes.c_double(delta), psi.encode("ascii"),
                        ctypes.byref(out_re), ctypes.byref(out_im))
    if rc == -1:
        raise ValueError("phase2_run returned -1: invalid input")
    return complex(out_re.value, out_im.value)
