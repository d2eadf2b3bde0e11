"""Graph files: the reference CLI's text format (LouvainCUDA/LouvainCPP.cpp:
660-737) with its token rules, and the binary edge list; round trips of the
10k-vertex SBM through both.  CPU only."""
import os

import numpy as np
import pytest

from louvaingpu_amd import _lib, generators, graphio
from louvaingpu_amd._lib import LouvainError


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def test_text_count_header_and_whitespace_tokens(tmp_path):
    nodes = write(tmp_path, "nodes.txt", "4\n10\t11  12\r\n\n13\n")
    edges = write(tmp_path, "edges.txt", "3\n10 11 1.5\n11\t12 2\n  12 13   0.25 \n")
    n, src, dst, w, ids = graphio.load_text(nodes, edges)
    assert n == 4 and ids.tolist() == [10, 11, 12, 13]
    assert src.tolist() == [10, 11, 12] and dst.tolist() == [11, 12, 13]
    assert w.tolist() == [1.5, 2.0, 0.25]


def test_last_node_id_needs_trailing_whitespace(tmp_path):
    """The reference stores a token only if reading it did not hit the end of
    the file: `3\\n1 2 3` gives two nodes, `3\\n1 2 3\\n` three."""
    edges = write(tmp_path, "edges.txt", "0\n")
    a = write(tmp_path, "a.txt", "3\n1 2 3")
    b = write(tmp_path, "b.txt", "3\n1 2 3\n")
    assert graphio.load_text(a, edges)[4].tolist() == [1, 2]
    assert graphio.load_text(b, edges)[4].tolist() == [1, 2, 3]


def test_edges_keep_the_last_triple_either_way(tmp_path):
    nodes = write(tmp_path, "nodes.txt", "3\n0 1 2\n")
    for tail in ("", "\n", "  \n\n"):
        edges = write(tmp_path, "e.txt", "2\n0 1 3\n1 2 4" + tail)
        _, src, dst, w, _ = graphio.load_text(nodes, edges)
        assert src.tolist() == [0, 1] and dst.tolist() == [1, 2] and w.tolist() == [3.0, 4.0]


def test_atoi_atof_parsing(tmp_path):
    nodes = write(tmp_path, "nodes.txt", "3\n7x -2 +5\n")
    edges = write(tmp_path, "edges.txt", "1\n7 -2 1e-1\n")
    _, src, dst, w, ids = graphio.load_text(nodes, edges)
    assert ids.tolist() == [7, -2, 5]
    assert w[0] == np.float32(0.1)


@pytest.mark.parametrize("nodes_text,edges_text", [
    ("2\n1 2 3\n", "0\n"),                 # more ids than the count (an overrun there)
    ("2\n1 2\n", "1\n1 2 1\n2 1 1\n"),     # more edges than the count
    ("2\n1 2\n", "1\n1 2\n"),              # incomplete triple (stale token there)
    ("", "0\n"),                            # no count
])
def test_malformed_text_is_rejected(tmp_path, nodes_text, edges_text):
    nodes = write(tmp_path, "nodes.txt", nodes_text)
    edges = write(tmp_path, "edges.txt", edges_text)
    with pytest.raises(LouvainError) as ex:
        graphio.load_text(nodes, edges)
    assert ex.value.status == _lib.LV_ERR_INVALID_INPUT


def test_missing_file(tmp_path):
    with pytest.raises(LouvainError) as ex:
        graphio.load_text(str(tmp_path / "none.txt"), str(tmp_path / "none2.txt"))
    assert ex.value.status == _lib.LV_ERR_ARG


def test_sbm_10k_round_trips_both_formats(tmp_path):
    n, src, dst, w = generators.sbm_10k()
    ids = np.arange(n, dtype=np.int32) * 3 + 5
    wf = w.astype(np.float32)
    graphio.save_text(tmp_path / "nodes.txt", tmp_path / "edges.txt", ids, ids[src], ids[dst], wf)
    n2, s2, d2, w2, ids2 = graphio.load_text(tmp_path / "nodes.txt", tmp_path / "edges.txt")
    assert n2 == n and np.array_equal(ids2, ids)
    assert np.array_equal(s2, ids[src]) and np.array_equal(d2, ids[dst])
    assert np.array_equal(w2, wf)
    for idv in (None, ids):
        p = tmp_path / "g.lveb"
        graphio.save_binary(p, n, src, dst, w, ids=idv)
        assert os.path.getsize(p) == 32 + (0 if idv is None else 4 * n) + 12 * len(src)
        n3, s3, d3, w3, ids3 = graphio.load_binary(p)
        assert n3 == n and w3.dtype == np.int32
        assert np.array_equal(s3, src) and np.array_equal(d3, dst) and np.array_equal(w3, w)
        assert (ids3 is None) if idv is None else np.array_equal(ids3, ids)


def test_binary_float_weights_and_layout(tmp_path):
    n, src, dst, w = generators.rmat(10, 16, 0.57, 0.19, 0.19, 2, "float")
    p = tmp_path / "r.lveb"
    graphio.save_binary(p, n, src, dst, w)
    raw = p.read_bytes()
    assert raw[:4] == b"LVEB"
    assert int.from_bytes(raw[4:8], "little") == 1
    assert int.from_bytes(raw[8:12], "little") == n
    assert int.from_bytes(raw[12:16], "little") == 2  # float weights, no ids
    assert int.from_bytes(raw[16:24], "little") == len(src)
    assert raw[24:32] == bytes(8)
    assert np.array_equal(np.frombuffer(raw[32:32 + 4 * len(src)], "<i4"), src)
    _, s, d, w2, ids = graphio.load_binary(p)
    assert ids is None and w2.dtype == np.float32 and np.array_equal(w2, w)


def test_binary_rejects_bad_headers(tmp_path):
    p = tmp_path / "g.lveb"
    graphio.save_binary(p, 3, [0, 1], [1, 2], np.ones(2, np.int32))
    raw = bytearray(p.read_bytes())
    for mutate in (lambda b: b.__setitem__(0, ord("X")),        # magic
                   lambda b: b.__setitem__(4, 2),               # version
                   lambda b: b.__setitem__(12, 4),              # unknown flag
                   lambda b: b.__setitem__(16, 3),              # e does not match the size
                   lambda b: b.__setitem__(30, 1)):             # reserved bytes
        bad = bytearray(raw)
        mutate(bad)
        q = tmp_path / "bad.lveb"
        q.write_bytes(bytes(bad))
        with pytest.raises(LouvainError) as ex:
            graphio.load_binary(q)
        assert ex.value.status == _lib.LV_ERR_INVALID_INPUT
    q = tmp_path / "short.lveb"
    q.write_bytes(bytes(raw[:-1]))
    with pytest.raises(LouvainError):
        graphio.load_binary(q)
