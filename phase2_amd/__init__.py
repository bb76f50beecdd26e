"""phase2_amd: the phase2 state-vector hot path on one AMD MI355X (gfx950).

``run(paulis, coeffs, delta, psi)`` is the drop-in for ``phase2.run`` of the
reference Python package (python/phase2/__init__.py:71-131): it computes
<psi| prod_k exp(i*delta*coeffs[k]*P_k) |psi> through the C entry
``phase2_run`` of libphase2_hip.so, with Pauli strings in "X0 Z1" notation and
``psi`` a '0'/'1' string (character k is qubit k).  Invalid input raises
ValueError, as there.

Beside it, without changing it:

``run_multidet(paulis, coeffs, delta, dets, nqb=None)``
    the same product on a multideterminant state sum_d c_d |i_d> of any
    width the device holds (``run`` stops at 30 qubits and basis states);
    returns sum_d conj(c_d) <i_d| U |psi>, the reference's multidet overlap
    (phase2/circ.c:338-350).
``qdrift_batch(registers, mode="batched")``
    many independent qDRIFT registers of uneven width at once (the batched
    launch); each register's overlaps equal qdrift_simul on it alone.
"""

from __future__ import annotations

import ctypes

from . import _native

__all__ = ["run", "run_multidet", "qdrift_batch", "pauli_masks"]


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


def pauli_masks(paulis: list[str], nqb: int):
    """Flip and phase masks of Pauli strings in "X0 Y1 Z5" notation, read
    as ``run`` reads them (phase2/phase2_run.c:60-103): tokens separated by
    blanks, tabs or newlines, each an operator in IXYZ followed by the qubit
    field as C's strtoul reads it (leading \\v, \\f or \\r skipped, an
    optional sign, decimal digits, the value modulo 2^64 and negated as an
    unsigned number after '-', a magnitude of 2^64 or more saturating), the
    qubit below nqb; a later operator on the same qubit replaces the earlier
    one.  Raises ValueError otherwise."""
    import re
    import numpy as np
    xm = np.zeros(len(paulis), dtype=np.uint64)
    zm = np.zeros(len(paulis), dtype=np.uint64)
    tok_re = re.compile(r"([IXYZ])[\v\f\r]*([+-]?)([0-9]+)")
    top = 1 << 64
    for k, s in enumerate(paulis):
        if len(s) >= 4096:
            raise ValueError("Pauli string of 4096 characters or more")
        x = z = 0
        for tok in re.split(r"[ \t\n]+", s):
            if not tok:
                continue
            m = tok_re.fullmatch(tok)
            q = -1
            if m:
                q = int(m.group(3))
                if q >= top:
                    q = top - 1          # ERANGE: ULONG_MAX
                elif m.group(2) == "-":
                    q = (top - q) % top  # unsigned negation
            if not m or q >= nqb:
                raise ValueError(f"bad Pauli token {tok!r} in {s!r}")
            b = 1 << q
            x &= ~b
            z &= ~b
            if m.group(1) in "XY":
                x |= b
            if m.group(1) in "YZ":
                z |= b
        xm[k], zm[k] = x, z
    return xm, zm


def run_multidet(paulis: list[str], coeffs: list[float], delta: float,
                 dets: dict, nqb: int | None = None) -> complex:
    """sum_d conj(c_d) <i_d| prod_k exp(i*delta*coeffs[k]*P_k) |psi> with
    |psi> = sum_d c_d |i_d>, the rotations applied in the given order in
    HBM-fused passes.  ``dets`` maps '0'/'1' strings (character k is qubit
    k) or integer indices to complex amplitudes; ``nqb`` defaults to the
    length of the bit strings.  Raises ValueError on invalid input."""
    import numpy as np
    if len(paulis) != len(coeffs):
        raise ValueError(
            f"len(paulis)={len(paulis)} != len(coeffs)={len(coeffs)}")
    if not dets:
        raise ValueError("empty state")
    keys = list(dets)
    if nqb is None:
        strs = [k for k in keys if isinstance(k, str)]
        if not strs:
            raise ValueError("nqb is needed with integer indices")
        nqb = len(strs[0])
    if nqb < 1 or nqb > 40:
        raise ValueError(f"nqb={nqb} out of range")
    idx = np.zeros(len(keys), dtype=np.uint64)
    for d, k in enumerate(keys):
        if isinstance(k, str):
            if len(k) != nqb or set(k) - {"0", "1"}:
                raise ValueError(f"bad basis string {k!r}")
            idx[d] = sum(1 << q for q, ch in enumerate(k) if ch == "1")
        else:
            if not 0 <= int(k) < (1 << nqb):
                raise ValueError(f"index {k} beyond {nqb} qubits")
            idx[d] = int(k)
    cf = np.array([complex(dets[k]) for k in keys], dtype=np.complex128)
    xm, zm = pauli_masks(paulis, nqb)
    phi = delta * np.asarray(coeffs, dtype=np.float64)
    _native.init()
    reg = _native.Register(nqb, _native.FUSED, 0, 3)
    try:
        reg.prep_multidet(idx, cf)
        for k0 in range(0, len(phi), 1024):   # circ_cache-sized blocks
            reg.rotate(xm[k0:k0 + 1024], zm[k0:k0 + 1024],
                       phi[k0:k0 + 1024])
        return reg.overlap(idx, cf)
    finally:
        reg.free()


def qdrift_batch(registers: list[dict], mode: str = "batched"):
    """Run many qDRIFT registers together.  Each register is a dict with
    paulis (list of strings) or xm/zm (masks), coeffs, nqb, dets (as in
    ``run_multidet``), samples, depth, step_size and seed; mode is
    "batched" (one launch per pass index, hipGraph per round, autotuned
    tiles), "fused" (fused passes register by register) or "plain" (one
    launch per rotation).  Returns one array of complex overlaps per
    register (one per sample)."""
    import numpy as np
    from . import synth
    modes = {"plain": _native.BATCH_PLAIN_EACH,
             "fused": _native.BATCH_FUSED_EACH,
             "batched": _native.BATCH_FUSED_BATCHED}
    if mode not in modes:
        raise ValueError(f"mode {mode!r} not in {sorted(modes)}")
    regs = []
    for r in registers:
        nqb = int(r["nqb"])
        if "paulis" in r:
            xm, zm = pauli_masks(r["paulis"], nqb)
        else:
            xm = np.asarray(r["xm"], dtype=np.uint64)
            zm = np.asarray(r["zm"], dtype=np.uint64)
        cf = np.asarray(r["coeffs"], dtype=np.float64)
        if not (len(xm) == len(zm) == len(cf)) or len(cf) == 0:
            raise ValueError("terms and coefficients differ in length")
        keys = list(r["dets"])
        idx = np.array([sum(1 << q for q, ch in enumerate(k) if ch == "1")
                        if isinstance(k, str) else int(k) for k in keys],
                       dtype=np.uint64)
        if np.any(idx >> np.uint64(nqb)):
            raise ValueError("determinant beyond the register")
        dcf = np.array([complex(r["dets"][k]) for k in keys])
        regs.append(dict(hm=synth.Hamiltonian(nqb, xm, zm, cf), det_idx=idx,
                         det_cf=dcf, samples=int(r["samples"]),
                         depth=int(r["depth"]),
                         step_size=float(r["step_size"]),
                         seed=int(r.get("seed", 0))))
    _native.init()
    out, _ = _native.batch_qdrift(regs, modes[mode], 0, 3)
    return out
