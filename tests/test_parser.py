"""The Pauli-string reader of phase2.run (phase2_amd.pauli_masks, and the C
reader in phase2_run.c that it mirrors) against the reference's grammar
(phase2/phase2_run.c:60-103): strtok on " \\t\\n", an operator letter, then
the qubit field through the C library's own strtoul(field, &end, 10),
called here through ctypes, so that the model below is glibc's behaviour
and not a restatement of it."""

import ctypes as C

import pytest

from phase2_amd import pauli_masks

_libc = C.CDLL(None)
_libc.strtoul.restype = C.c_ulong
_libc.strtoul.argtypes = [C.c_char_p, C.POINTER(C.c_char_p), C.c_int]
_libc.__errno_location.restype = C.POINTER(C.c_int)


def _reference_masks(s, nqb):
    """parse_pauli of the reference with strtok/strtoul; None on -1."""
    if len(s) >= 4096:
        return None
    x = z = 0
    toks = [t for t in s.replace("\t", " ").replace("\n", " ").split(" ")
            if t]
    for tok in toks:
        if tok[0] not in "IXYZ":
            return None
        field = tok[1:].encode("latin-1")
        buf = C.create_string_buffer(field)
        end = C.c_char_p()
        q = _libc.strtoul(buf, C.byref(end), 10)
        consumed = C.cast(end, C.c_void_p).value - C.addressof(buf)
        if consumed == 0 or consumed != len(field):
            return None
        if q >= nqb:
            return None
        b = 1 << q
        x &= ~b
        z &= ~b
        if tok[0] in "XY":
            x |= b
        if tok[0] in "YZ":
            z |= b
    return x, z


FIELDS = ["X1", "X\v1", "X\f\r2", "X+1", "X-0", "X+0", "Y-18446744073709551615",
          "Z-18446744073709551614", "X18446744073709551615",
          "X18446744073709551616", "X-18446744073709551616",
          "X99999999999999999999999", "X-99999999999999999999999", "X1\v",
          "X\v", "X-", "X+", "X", "X- 1", "X-\v1", "X1x", "X 1", "W0",
          "X0 Y1\tZ2\nI3", "X3 Z3", "Y2 Y2", "X00002", "X\v-2",
          "Z18446744073709551618", "X-18446744073709551612", ""]


@pytest.mark.parametrize("s", FIELDS)
def test_reader_matches_strtoul(s):
    want = _reference_masks(s, 4)
    if want is None:
        with pytest.raises(ValueError):
            pauli_masks([s], 4)
    else:
        xm, zm = pauli_masks([s], 4)
        assert (int(xm[0]), int(zm[0])) == want


def test_reader_length_limit():
    s = "X0 " * 1400
    assert len(s) >= 4096 and _reference_masks(s, 4) is None
    with pytest.raises(ValueError):
        pauli_masks([s], 4)
    s = ("X0 " * 1365)[:4095]
    assert pauli_masks([s], 4)[0][0] == 1


@pytest.mark.gpu
@pytest.mark.parametrize("s", FIELDS)
def test_c_reader_matches_strtoul(s):
    """The C reader inside phase2_run (through phase2_amd.run on the GPU)
    accepts and rejects what strtoul does, and a string it accepts rotates
    as its masks do in the oracle."""
    import phase2_amd
    from oracle import pyoracle as po
    want = _reference_masks(s, 4)
    if want is None:
        with pytest.raises(ValueError):
            phase2_amd.run([s], [0.7], 0.3, "0110")
        return
    got = phase2_amd.run([s, "X0 Y2"], [0.7, -0.4], 0.3, "0110")
    x, z = want
    assert got == po.phase2_run(4, 0b0110, [x, 0b101], [z, 0b100],
                                [0.7, -0.4], 0.3)
