"""The device-side sampler (ph2hip_sample_cdf): xoshiro256** draws and the
inverse of the |c| CDF made on the GPU give the index streams of the host
sampler bit for bit.

* the raw generator: on a table of 2^20 equal steps the drawn index is the
  top 20 bits of each output, which must be the reference-written golden
  streams (tests/golden/xoshiro.npz, six seeds including 0 and 2^64-1);
* the inverse: the reference-written tables and queries of
  tests/golden/cdf.npz (lengths 1, 2, 3, 32, 37, 1000), drawn through the
  device, against the oracle's inverse of the same uniforms, plus tables
  with zero weights (ties) and 10,000 terms;
* continuation: a stream drawn in two calls equals one call, and the
  returned state equals the host generator's after the same draws."""

import numpy as np
import pytest

from oracle import pyoracle as po
from phase2_amd import _native as nat

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    print("sampler: init", flush=True)
    nat.init()


def _uniforms(seed, n):
    raw = po.oracle().xoshiro(int(seed), n)
    return (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def test_raw_stream_matches_golden(golden):
    g = golden("xoshiro.npz")
    k = 20
    steps = np.arange(1, (1 << k) + 1, dtype=np.float64) / (1 << k)
    tables = [steps] * len(g["seeds"])
    states = [nat.xoshiro_state(int(s)) for s in g["seeds"]]
    n = g["out"].shape[1]
    print(f"sampler: {len(tables)} golden streams of {n}", flush=True)
    got, _ = nat.sample_cdf(tables, states, [n] * len(tables))
    for s, out, idx in zip(g["seeds"], g["out"], got):
        assert np.array_equal(idx.astype(np.uint64),
                              out >> np.uint64(64 - k)), int(s)


def test_inverse_matches_oracle(golden):
    g = golden("cdf.npz")
    rng = np.random.default_rng(4)
    weights = [g[f"w{nw}"] for nw in (1, 2, 3, 32, 37, 1000)]
    ties = rng.normal(size=300)
    ties[rng.random(300) < 0.4] = 0.0     # equal table entries
    ties[:5] = 0.0
    ties[-7:] = 0.0
    weights += [ties, rng.normal(size=10_000) * np.exp(-rng.random(10_000))]
    seeds = [0x1234 + 7 * k for k in range(len(weights))]
    ndraw = [403, 1, 2048, 1000, 777, 3000, 4096, 5000]
    tables = [nat.cdf_table(w) for w in weights]
    print(f"sampler: {len(tables)} tables", flush=True)
    got, _ = nat.sample_cdf(tables, [nat.xoshiro_state(s) for s in seeds],
                            ndraw)
    o = po.oracle()
    for w, s, n, idx in zip(weights, seeds, ndraw, got):
        want, _ = o.cdf_inverse(w, _uniforms(s, n))
        assert np.array_equal(idx.astype(np.uint64), want), (len(w), s)


def test_stream_continues_across_calls():
    w = np.random.default_rng(9).normal(size=2000)
    t = nat.cdf_table(w)
    st = nat.xoshiro_state(77)
    one, st_one = nat.sample_cdf([t], [st], [2500])
    a, st_a = nat.sample_cdf([t], [st], [1200])
    b, st_b = nat.sample_cdf([t], st_a, [1300])
    assert np.array_equal(np.concatenate([a[0], b[0]]), one[0])
    assert np.array_equal(st_b, st_one)
    want, _ = po.oracle().cdf_inverse(w, _uniforms(77, 2500))
    assert np.array_equal(one[0].astype(np.uint64), want)


def test_rejects_bad_streams():
    t = nat.cdf_table(np.ones(4))
    with pytest.raises(RuntimeError, match="zero state"):
        nat.sample_cdf([t], [np.zeros(4, np.uint64)], [3])
    with pytest.raises(RuntimeError, match="table of 0"):
        nat.sample_cdf([np.zeros(0)], [nat.xoshiro_state(1)], [3])
