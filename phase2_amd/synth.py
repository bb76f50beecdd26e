"""Synthetic Jordan-Wigner-style Hamiltonians and initial states.

The generator is frozen: every workload, golden fixture and benchmark in this
package is defined by it and a seed.  Changing any line below changes every
recorded vector.

Pauli strings use the symplectic encoding of the phase2 library
(include/phase2/paulis.h:37-58): ``xm`` is the flip mask (X or Y on the qubit),
``zm`` the phase mask (Z or Y on the qubit).  Qubit k is bit k.

Generator (numpy PCG64 stream seeded with ``seed``; ``N(0, s)`` below is a
normal draw with standard deviation ``s``):

* ``n`` single-qubit terms Z_p, p = 0..n-1, c ~ N(0, sigma_z);
* C(n,2) terms Z_p Z_q, p < q in lexicographic (p, q) order, c ~ N(0, sigma_zz);
* ``n_four`` four-index terms: p<q<r<s drawn uniformly (``rng.choice(n, 4,
  replace=False)`` sorted), an X/Y pattern on (p, q, r, s) drawn uniformly from
  the 8 patterns with an even number of Y, Z on every qubit strictly between p
  and q and strictly between r and s; c ~ N(0, sigma_4), times
  ``exp(-(r - p) / 8)`` when ``decay`` is set.  A string already drawn is
  rejected and redrawn, so all strings are distinct;
* one identity term, c ~ N(0, sigma_z), drawn last.

Terms are then ordered by |c| descending (stable) and scaled so that
sum |c| = 1.  Distinct strings make the library's lexicographic sort
(circ/trott.c:33) independent of sort stability.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# X/Y patterns on (p, q, r, s) with an even number of Y.
PATTERNS = ("XXXX", "XXYY", "XYXY", "XYYX", "YXXY", "YXYX", "YYXX", "YYYY")


@dataclass
class Hamiltonian:
    nqb: int
    xm: np.ndarray  # uint64 flip masks
    zm: np.ndarray  # uint64 phase masks
    cf: np.ndarray  # float64 coefficients

    def __len__(self) -> int:
        return int(self.cf.shape[0])


def _four_index(p, q, r, s, pat):
    xm = (1 << p) | (1 << q) | (1 << r) | (1 << s)
    zm = 0
    for bit, op in zip((p, q, r, s), PATTERNS[pat]):
        if op == "Y":
            zm |= 1 << bit
    for k in range(p + 1, q):
        zm |= 1 << k
    for k in range(r + 1, s):
        zm |= 1 << k
    return xm, zm


def hamiltonian(n: int, n_four: int, seed: int = 0, *, sigma_z: float = 0.5,
                sigma_zz: float = 0.05, sigma_4: float = 0.01,
                decay: bool = False, n_total: int | None = None) -> Hamiltonian:
    """Generate the synthetic Hamiltonian.

    ``n_total``, when given, fixes the total term count (identity included)
    and overrides ``n_four`` with ``n_total - n - C(n,2) - 1``.
    """
    if n < 4 or n > 64:
        raise ValueError("n must be in [4, 64]")
    nzz = n * (n - 1) // 2
    if n_total is not None:
        n_four = n_total - n - nzz - 1
    max_four = math.comb(n, 4) * len(PATTERNS)
    if n_four < 0 or n_four > max_four // 2:
        raise ValueError(f"n_four={n_four} out of range for n={n}")
    rng = np.random.Generator(np.random.PCG64(seed))
    xs, zs, cs = [], [], []
    for p in range(n):
        xs.append(0)
        zs.append(1 << p)
        cs.append(rng.normal(0.0, sigma_z))
    for p in range(n):
        for q in range(p + 1, n):
            xs.append(0)
            zs.append((1 << p) | (1 << q))
            cs.append(rng.normal(0.0, sigma_zz))
    seen = set()
    count = 0
    while count < n_four:
        p, q, r, s = sorted(int(v) for v in rng.choice(n, 4, replace=False))
        pat = int(rng.integers(len(PATTERNS)))
        c = rng.normal(0.0, sigma_4)
        if decay:
            c *= math.exp(-(r - p) / 8.0)
        xm, zm = _four_index(p, q, r, s, pat)
        if (xm, zm) in seen:
            continue
        seen.add((xm, zm))
        xs.append(xm)
        zs.append(zm)
        cs.append(c)
        count += 1
    xs.append(0)
    zs.append(0)
    cs.append(rng.normal(0.0, sigma_z))

    xm = np.array(xs, dtype=np.uint64)
    zm = np.array(zs, dtype=np.uint64)
    cf = np.array(cs, dtype=np.float64)
    order = np.argsort(-np.abs(cf), kind="stable")
    xm, zm, cf = xm[order], zm[order], cf[order]
    cf = cf / np.sum(np.abs(cf))
    return Hamiltonian(n, xm, zm, cf)


def multidet_random(n: int, ndet: int, ones: int, seed: int):
    """``ndet`` distinct random basis states of ``n`` qubits with ``ones`` bits
    set, with complex Gaussian amplitudes normalised to unit norm.
    Returns (idx uint64[ndet], cf complex128[ndet])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    idx = []
    seen = set()
    while len(idx) < ndet:
        bits = rng.choice(n, ones, replace=False)
        v = 0
        for b in bits:
            v |= 1 << int(b)
        if v in seen:
            continue
        seen.add(v)
        idx.append(v)
    a = rng.normal(size=ndet) + 1j * rng.normal(size=ndet)
    a = a / np.sqrt(np.sum(np.abs(a) ** 2))
    return np.array(idx, dtype=np.uint64), a.astype(np.complex128)


def coeff_matrix(n_sites: int, n_occ: int, seed: int) -> np.ndarray:
    """Real orthonormal columns: Q of the QR of a seeded Gaussian
    (n_sites, n_occ) matrix, row-major float64."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.normal(size=(n_sites, n_occ))
    q, _ = np.linalg.qr(a)
    return np.ascontiguousarray(q, dtype=np.float64)


# -- named workloads ---------------------------------------------------------

def trott_n16():
    """1st-order Trotter, n=16: 16 Z + 120 ZZ + 400 four-index + identity
    (537 terms), single determinant 0xFF, delta=0.1, 64 steps."""
    hm = hamiltonian(16, 400, seed=0)
    idx = np.array([0xFF], dtype=np.uint64)
    cf = np.array([1.0 + 0.0j])
    return dict(hm=hm, det_idx=idx, det_cf=cf, delta=0.1, steps=64)


def trott2_n26():
    """2nd-order Trotter, n=26: 26 Z + 325 ZZ + 20,000 four-index (decayed)
    + identity, coefficient-matrix state n_sites=26, n_alpha=6, n_beta=0,
    delta=0.05, 256 steps."""
    hm = hamiltonian(26, 20000, seed=0, decay=True)
    C = coeff_matrix(26, 6, seed=0)
    return dict(hm=hm, C_alpha=C, n_sites=26, n_alpha=6, n_beta=0,
                closed_shell=1, tapered=0, delta=0.05, steps=256)


def trott_n33():
    """1st-order Trotter, n=33: 100,000 terms (33 Z + 528 ZZ + 99,438
    four-index, decayed, + identity), 64-determinant random state with 16
    ones each, delta=0.02, 8 steps."""
    hm = hamiltonian(33, 0, seed=0, decay=True, n_total=100_000)
    idx, cf = multidet_random(33, 64, 16, seed=0)
    return dict(hm=hm, det_idx=idx, det_cf=cf, delta=0.02, steps=8)


def cmpsit_n30():
    """Composite integrator, n=30: 60,000 terms (30 Z + 435 ZZ + 59,534
    four-index, decayed, + identity), 64-determinant random state with 15
    ones each.  The 2,000 largest |c| form the deterministic part (angle
    delta = 0.02 on cf), the rest is sampled, 512 draws per half-step with
    angle lambda_r * delta / 512 (lambda_r = sum |c| of the sampled part),
    32 steps, one sample."""
    hm = hamiltonian(30, 0, seed=0, decay=True, n_total=60_000)
    idx, cf = multidet_random(30, 64, 15, seed=0)
    length, depth, delta = 2000, 512, 0.02
    lam_r = float(np.sum(np.abs(hm.cf[length:])))
    return dict(hm=hm, det_idx=idx, det_cf=cf, length=length, depth=depth,
                steps=32, samples=1, angle_det=delta,
                angle_rand=lam_r * delta / depth, seed=0x5EED0030)


def hamiltonian_vec(n: int, n_four: int, seed: int, *, sigma_z: float = 0.5,
                    sigma_zz: float = 0.05, sigma_4: float = 0.01,
                    decay: bool = True) -> Hamiltonian:
    """Vectorised variant of ``hamiltonian`` for many registers: same term
    families and distributions, but the four-index supports come from the
    first four columns of a random permutation of each row (numpy argsort
    of uniform keys), patterns and coefficients are drawn in bulk, and
    duplicate strings are merged away before topping up.  Used for the
    batched qDRIFT workload; its stream differs from ``hamiltonian``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p_, q_ = np.triu_indices(n, 1)
    zs = [np.uint64(1) << np.arange(n, dtype=np.uint64),
          (np.uint64(1) << p_.astype(np.uint64)) |
          (np.uint64(1) << q_.astype(np.uint64))]
    cs = [rng.normal(0.0, sigma_z, n), rng.normal(0.0, sigma_zz, len(p_))]
    pat_y = np.array([[c == "Y" for c in p] for p in PATTERNS])
    bits = np.uint64(1) << np.arange(n, dtype=np.uint64)
    below = np.concatenate([[0], np.cumsum(bits)]).astype(np.uint64)
    seen_x = np.zeros(0, np.uint64)
    seen_z = np.zeros(0, np.uint64)
    seen_c = np.zeros(0)
    while len(seen_x) < n_four:
        k = 2 * (n_four - len(seen_x)) + 16
        sup = np.sort(np.argsort(rng.random((k, n)), axis=1)[:, :4], axis=1)
        pat = rng.integers(len(PATTERNS), size=k)
        c = rng.normal(0.0, sigma_4, k)
        p, q, r, s = sup.T
        if decay:
            c = c * np.exp(-(r - p) / 8.0)
        xm = bits[p] | bits[q] | bits[r] | bits[s]
        y = pat_y[pat]
        zm = (np.where(y[:, 0], bits[p], 0) | np.where(y[:, 1], bits[q], 0) |
              np.where(y[:, 2], bits[r], 0) | np.where(y[:, 3], bits[s], 0))
        # Z strictly between p..q and r..s: bits (p, q) = below[q] ^ below[p+1]
        zm = zm | (below[q] ^ below[p + 1]) | (below[s] ^ below[r + 1])
        zm = zm.astype(np.uint64)
        allx = np.concatenate([seen_x, xm])
        allz = np.concatenate([seen_z, zm])
        allc = np.concatenate([seen_c, c])
        _, first = np.unique(np.stack([allx, allz], 1), axis=0,
                             return_index=True)
        keep = np.sort(first)[:n_four]
        seen_x, seen_z, seen_c = allx[keep], allz[keep], allc[keep]
    xm = np.concatenate([np.zeros(n + len(p_), np.uint64), seen_x,
                         np.zeros(1, np.uint64)])
    zm = np.concatenate(zs + [seen_z, np.zeros(1, np.uint64)])
    cf = np.concatenate(cs + [seen_c, rng.normal(0.0, sigma_z, 1)])
    order = np.argsort(-np.abs(cf), kind="stable")
    xm, zm, cf = xm[order], zm[order], cf[order]
    return Hamiltonian(n, xm, zm, cf / np.sum(np.abs(cf)))


def qdrift_batch(nreg: int = 512, seed: int = 0, samples: int = 16,
                 depth: int = 1000, step_size: float = 0.01,
                 nmin: int = 16, nmax: int = 26):
    """Batched qDRIFT workload: ``nreg`` registers of widths drawn uniformly
    from [nmin, nmax], each with its own Hamiltonian of 2,000-10,000 terms
    (``hamiltonian_vec``), a 4-determinant initial state, and its own
    xoshiro256** seed; ``samples`` samples of ``depth`` draws each."""
    rng = np.random.Generator(np.random.PCG64(seed))
    widths = rng.integers(nmin, nmax + 1, size=nreg)
    totals = rng.integers(2000, 10001, size=nreg)
    regs = []
    for i in range(nreg):
        n = int(widths[i])
        nfixed = n + n * (n - 1) // 2 + 1
        hm = hamiltonian_vec(n, int(totals[i]) - nfixed, seed=seed * 7919 + i)
        idx, cf = multidet_random(n, 4, n // 2, seed=seed * 104729 + i)
        regs.append(dict(hm=hm, det_idx=idx, det_cf=cf, samples=samples,
                         depth=depth, step_size=step_size,
                         seed=0x9E3779B97F4A7C15 ^ (seed << 20) ^ (i + 1)))
    return regs
