"""lv_load_edges' encoded upload against plain copies of the same edge list.

The encoded upload sends a source-sorted list's row offsets instead of src and
int weights packed 2, 4 or 8 bits each (the width a sample of about 64 K
weights says sends the fewest bytes; weights that do not fit go to an escape
list); the device decodes them in its load checks.  Every case runs the same list with
upload=1 (plain copies of src, dst, w as the reference makes them,
LouvainCUDA/LouvainCPP.cpp:283-287) and upload=2 (encoded at any size), and
asserts identical levels, Partition arrays, final partition and modularity,
or the same error status.  The cases: each packing width, escaped weights (at
the escape value, above it, large and zero), rows missing before the first and after the last source,
decode tiles whose sources span more than the offsets window in LDS, an
unsorted list (its src goes raw, its weights still as bytes), float weights
(offsets only), invalid input found in escapes, dst and src; and the automatic
choice at R-MAT scale 18 against the oracle.
"""
import numpy as np
import pytest

import oracle
from conftest import gpu_case
from louvaingpu_amd import Engine, generators
from louvaingpu_amd._lib import LouvainError, LV_ERR_INVALID_INPUT

pytestmark = pytest.mark.gpu


def outcome(n, src, dst, w, upload, **cfg):
    with Engine(upload=upload, **cfg) as eng:
        eng.load_edges(n, src, dst, w)
        sent = eng.stats().upload_bytes
        eng.run()
        levels = []
        for l in range(eng.num_levels()):
            i = eng.level_info(l)
            node, comm = eng.level_partition(l)
            levels.append(((i.n, i.k, i.nnz, i.m_units, i.passes, i.scale, i.q, i.lower),
                           node.tobytes(), comm.tobytes()))
        return levels, eng.final_assignment(), eng.final_modularity(), sent


def assert_modes_agree(n, src, dst, w, **cfg):
    plain = outcome(n, src, dst, w, 1, **cfg)
    enc = outcome(n, src, dst, w, 2, **cfg)
    assert plain[0] == enc[0]
    assert np.array_equal(plain[1], enc[1])
    assert plain[2] == enc[2]
    assert plain[3] == 12 * len(src)
    return plain, enc


def pack_width(w):
    """The packing width the uploader picks: the lowest packed bytes plus 64
    per escape over 2, 4 and 8 bits, estimated from 256 runs of 64 consecutive
    weights spread over the list (run r starts at (e - 64) * r // 255, rounded
    down to a multiple of 16)."""
    e = len(w)
    u = np.asarray(w).view(np.uint32)
    run = min(e, 64)
    starts = [((e - run) * r // 255) & ~15 for r in range(256)]
    x = np.concatenate([u[a:min(e, a + 64)] for a in starts]).astype(np.int64)
    f = e / max(len(x), 1)
    b2 = e / 4 + 64 * f * (x >= 3).sum()
    b4 = e / 2 + 64 * f * (x >= 15).sum()
    b8 = e + 64 * f * (x >= 255).sum()
    return 2 if (b2 <= b4 and b2 <= b8) else (4 if b4 <= b8 else 8)


def sent_bytes(n, e, w=None, src_raw=False):
    """Bytes of an encoded upload: dst, offsets (or src), packed weights with
    12 bytes per escape (or raw float weights when w is None)."""
    s = 4 * e + (4 * e if src_raw else 4 * (n + 1))
    if w is None:
        return s + 4 * e
    wb = pack_width(w)
    nesc = int((np.asarray(w).view(np.uint32) >= (1 << wb) - 1).sum())
    return s + (e * wb + 7) // 8 + 12 * nesc


def rmat16():
    n, src, dst, w = generators.rmat(16, 16, 0.57, 0.19, 0.19, 3, "int")
    return n, src, dst, w.copy()


def test_escaped_weights_match_plain_and_oracle():
    gpu_case("upload: escaped weights")
    n, src, dst, w = rmat16()
    w[::997] = 255
    w[1::1009] = 256
    w[2::5003] = 70000
    w[3::7919] = 0
    w[4] = 1 << 20
    _, enc = assert_modes_agree(n, src, dst, w)
    assert enc[3] == sent_bytes(n, len(src), w)
    r = oracle.run(n, src, dst, w)
    assert np.array_equal(enc[1], r.final_comm())
    assert enc[2] == pytest.approx(r.final_modularity(), rel=1e-9, abs=1e-12)


def test_rows_missing_at_both_ends():
    gpu_case("upload: empty leading and trailing rows")
    n, src, dst, w = rmat16()
    # vertices 0..499 and the last 700 have no out-edges (nor, here, any edge)
    _, enc = assert_modes_agree(n + 1200, src + 500, dst + 500, w)
    assert enc[3] == sent_bytes(n + 1200, len(src), w)


def test_sparse_sources_take_the_global_search():
    gpu_case("upload: tiles spanning many vertices")
    rng = np.random.default_rng(11)
    n, e = 1 << 21, 20000  # a 4096-edge tile spans ~400k vertices
    src = np.sort(rng.integers(0, n, e)).astype(np.int32)
    dst = rng.integers(0, n, e).astype(np.int32)
    w = rng.integers(1, 400, e).astype(np.int32)
    assert_modes_agree(n, src, dst, w)


def test_unsorted_list_sends_src_raw():
    gpu_case("upload: unsorted list")
    n, src, dst, w = rmat16()
    p = np.random.default_rng(5).permutation(len(src))
    _, enc = assert_modes_agree(n, src[p], dst[p], w[p])
    assert enc[3] == sent_bytes(n, len(src), w[p], src_raw=True)


def test_float_weights_send_offsets_only():
    gpu_case("upload: float weights")
    n, src, dst, w = generators.rmat(15, 16, 0.57, 0.19, 0.19, 4, "float")
    _, enc = assert_modes_agree(n, src, dst, w)
    assert enc[3] == sent_bytes(n, len(src))


@pytest.mark.parametrize("wb", [2, 4, 8])
def test_each_packing_width(wb):
    """Weights drawn so that each width wins, with a few escapes of each kind
    (the escape value itself, larger, and 2^30)."""
    gpu_case("upload: %d-bit weights" % wb)
    n, src, dst, _ = rmat16()
    rng = np.random.default_rng(wb)
    top = {2: 3, 4: 15, 8: 255}[wb]
    w = rng.integers(0, top, len(src)).astype(np.int32)
    w[::4001] = top
    w[1::5003] = top + 1
    w[2::7001] = 1 << 30
    assert pack_width(w) == wb
    _, enc = assert_modes_agree(n, src, dst, w)
    assert enc[3] == sent_bytes(n, len(src), w)


@pytest.mark.parametrize("case", ["negative escape", "dst past n", "src past n", "src negative"])
def test_invalid_input_same_status(case):
    gpu_case("upload: invalid " + case)
    n, src, dst, w = rmat16()
    src = src.copy()
    dst = dst.copy()
    if case == "negative escape":
        w[10] = -5
    elif case == "dst past n":
        dst[7] = n
    elif case == "src past n":
        src[-1] = n  # still sorted: the host scan refuses it, src goes raw
    else:
        src[0] = -1
    for upload in (1, 2):
        with Engine(upload=upload) as eng:
            with pytest.raises(LouvainError) as ex:
                eng.load_edges(n, src, dst, w)
            assert ex.value.status == LV_ERR_INVALID_INPUT, upload
            # the context stays usable
            eng.load_edges(4, np.array([0, 1], np.int32), np.array([1, 2], np.int32),
                           np.array([300, 1], np.int32))
            eng.run()
            assert eng.final_assignment().tolist() == [0, 0, 0, 1]  # the oracle's


def test_automatic_choice_at_rmat18():
    gpu_case("upload: automatic at rmat18")
    n, src, dst, w = generators.rmat(18, 16, 0.57, 0.19, 0.19, 1, "int")
    plain = outcome(n, src, dst, w, 1)
    auto = outcome(n, src, dst, w, 0)
    assert plain[0] == auto[0] and np.array_equal(plain[1], auto[1]) and plain[2] == auto[2]
    assert auto[3] < plain[3]  # encoded: the process may use 4 or more CPUs
    r = oracle.run(n, src, dst, w, threads=8)
    assert np.array_equal(auto[1], r.final_comm())
