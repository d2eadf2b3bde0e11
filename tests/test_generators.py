"""Generators are deterministic (pinned digests) and produce canonical,
sorted, merged edge lists with the configured statistics."""
import hashlib
import json
import os

import numpy as np

from louvaingpu_amd import generators

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "generators.json")


def digest(n, src, dst, w):
    h = hashlib.sha256()
    h.update(np.int64(n).tobytes())
    for a in (src, dst, w):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


CASES = {
    "sbm_10k_seed0": lambda: generators.sbm_10k(0),
    "rmat16_int_seed1": lambda: generators.rmat(16, 16, 0.57, 0.19, 0.19, 1, "int"),
    "rmat14_float_seed2": lambda: generators.rmat(14, 16, 0.57, 0.19, 0.19, 2, "float"),
    "sbm_batch_first": lambda: generators.sbm(generators.batch_params(4, 3)[0][0], 100,
                                              generators.batch_params(4, 3)[0][1],
                                              generators.batch_params(4, 3)[0][1] / 50,
                                              generators.batch_params(4, 3)[0][2]),
}


def test_digests_are_pinned():
    with open(GOLDEN) as f:
        pinned = json.load(f)
    for name, fn in CASES.items():
        assert digest(*fn()) == pinned[name], name


def test_deterministic_across_calls():
    a = generators.rmat(12, 8, seed=5, weights="float")
    b = generators.rmat(12, 8, seed=5, weights="float")
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


def test_sbm_statistics():
    n, src, dst, w = generators.sbm_10k()
    # expected 24,750 intra + 24,750 inter edges, sd about 220
    assert abs(len(src) - 49500) < 6 * 220
    assert np.all(src < dst)
    key = src.astype(np.int64) * n + dst
    assert np.all(np.diff(key) > 0)
    intra = np.sum(src // 100 == dst // 100)
    assert abs(intra - 24750) < 6 * 160
    assert np.all(w == 1)


def test_rmat_canonical_sorted_merged():
    n, src, dst, w = generators.rmat(14, 16, 0.57, 0.19, 0.19, 1, "int")
    assert n == 1 << 14
    assert np.all(src <= dst)
    key = src.astype(np.int64) * n + dst
    assert np.all(np.diff(key) > 0)          # duplicates merged
    assert int(w.sum()) == 16 * n            # every draw counted once
    nf, sf, df, wf = generators.rmat(14, 16, 0.57, 0.19, 0.19, 1, "float")
    assert np.all(sf <= df)
    assert np.all(np.diff(sf.astype(np.int64) * nf + df) > 0)
    # each merged weight is a sum of draws from U[1, 10]
    assert np.all(wf >= 1.0)
    assert 16 * nf * 5.0 * 0.98 < float(wf.astype(np.float64).sum()) < 16 * nf * 5.5 * 1.02


def test_resolutions():
    g = generators.resolutions()
    assert len(g) == 16 and abs(g[0] - 0.25) < 1e-15 and abs(g[-1] - 4.0) < 1e-12
