"""CPU-side checks of the C-ABI library and the host mirror (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from zog_amd import _lib
from zog_amd import resource_graph as rg
from zog_amd.soa import SoA

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zog_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?(zg_\w+)\s*\(", text, re.M)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_header_constants_match_the_python_mirror():
    # every #define of the header that the Python binding names holds the
    # same value there (error codes, precision flags, tuning knobs, tables)
    text = open(HEADER).read()
    defines = {}
    for name, value in re.findall(r"^#define\s+(ZG_\w+)\s+\(?(-?(?:0x[0-9a-fA-F]+|\d+))\)?\s*(?:/\*.*)?$", text, re.M):
        defines[name] = int(value, 0)
    mirrored = {n: v for n, v in defines.items() if hasattr(_lib, n)}
    assert len(mirrored) >= 30
    for name, value in mirrored.items():
        assert getattr(_lib, name) == value, name
    for name in ("ZG_TUNE_LANE_ORDER", "ZG_TUNE_QUEUE_ORDER", "ZG_TUNE_PR_BLOCK", "ZG_ERR_SCRATCH"):
        assert name in mirrored, name
    # profile slots: the binding's names by slot index
    slots = {n[len("ZG_PROF_"):].lower(): v for n, v in defines.items() if n.startswith("ZG_PROF_") and n != "ZG_PROF_SLOTS"}
    assert len(_lib.PROF_SLOTS) == defines["ZG_PROF_SLOTS"]
    for name, value in slots.items():
        assert _lib.PROF_SLOTS[value] == name, name


def test_library_loads_and_exports_every_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def test_strerror_names_codes():
    for code in range(-8, 1):
        s = _lib.strerror(code)
        assert s and "unknown" not in s
    assert "unknown" in _lib.strerror(-99)


def test_null_handle_is_rejected():
    L = _lib.lib()
    assert L.zg_graph_info(None, None, None, None) == _lib.ZG_ERR_INVALID_HANDLE
    assert L.zg_bc_unweighted(None, None, 0, None, None) == _lib.ZG_ERR_INVALID_HANDLE
    assert L.zg_pagerank(None, 0.85, 10, 1e-4, 32, None, None) == _lib.ZG_ERR_INVALID_HANDLE
    assert L.zg_set_batch_groups(None, 4) == _lib.ZG_ERR_INVALID_HANDLE
    assert L.zg_graph_table(None, _lib.ZG_TABLE_ICH, None, 0, None) == _lib.ZG_ERR_INVALID_HANDLE
    L.zg_graph_destroy(None)
    assert L.zg_graph_create(3, None, None, None, 0, None) == _lib.ZG_ERR_INVALID_ARG
    assert L.zg_bc_f64(None, None, 0, None, None) == _lib.ZG_ERR_INVALID_HANDLE
    assert L.zg_graph_create_weighted(3, None, None, None, None, 0, None) == _lib.ZG_ERR_INVALID_ARG


def test_errors_raise():
    with pytest.raises(_lib.NodeNotFoundError):
        _lib.check(_lib.ZG_ERR_NODE_NOT_FOUND, "x")
    with pytest.raises(_lib.ZogError):
        _lib.check(_lib.ZG_ERR_GPU_OOM, "x")
    _lib.check(_lib.ZG_OK)


# --- host mirror of Zog.SoA ---------------------------------------------------

def test_soa_edge_arrays_order():
    # soa_test.exs:45-55
    b = SoA.directed().add_edge("A", "B", 1.0).add_edge("B", "C", 2.0)
    f, t, w = b.to_edge_arrays()
    assert f.tolist() == [0, 1] and t.tolist() == [1, 2] and w.tolist() == [1.0, 2.0]


def test_soa_undirected_both_arcs():
    b = SoA.undirected().add_edge("A", "B", 5.0)
    f, t, w = b.to_edge_arrays()
    assert list(zip(f.tolist(), t.tolist())) == [(0, 1), (1, 0)]
    assert b.edge_count() == 2 and b.node_count() == 2
    assert b.all_labels() == ["A", "B"]


def test_map_scores_and_scaling():
    b = SoA.undirected().add_edge("x", "y").add_edge("y", "z")
    assert rg._map_scores(b, [1.0, 2.0, 3.0]) == {"x": 1.0, "y": 2.0, "z": 3.0}
    assert rg._maybe_scale_undirected(b, np.array([2.0, 4.0])).tolist() == [1.0, 2.0]
    d = SoA.directed().add_edge("x", "y")
    assert rg._maybe_scale_undirected(d, np.array([2.0])).tolist() == [2.0]
    i = SoA.from_arrays("directed", 3, [0, 1], [1, 2])
    assert rg._map_scores(i, [0.5, 0.25, 0.25]) == {0: 0.5, 1: 0.25, 2: 0.25}


def _read(tmp_path, text, directed=True, integer_labels=False, name="e.txt"):
    p = tmp_path / name
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return rg._read_edgelist_native(str(p), directed, integer_labels)


def test_edgelist_parse_integer_labels(tmp_path):
    # resource_graph_test.exs:545-571 input; placeholder ids up to max id;
    # a comment, a header, a blank line, a non-numeric id (skipped), commas
    src, dst, w, n, labels = _read(tmp_path, "# comment\nsource target\n0 1 1.5\n1 2 2.0\n\n2 0 1.0\nx 3\n5,4\n",
                                   integer_labels=True)
    assert n == 6 and labels is None
    assert list(zip(src.tolist(), dst.tolist())) == [(0, 1), (1, 2), (2, 0), (5, 4)]
    assert w.tolist() == [1.5, 2.0, 1.0, 1.0]


def test_edgelist_parse_labels_first_seen(tmp_path):
    # a tab before a non-numeric weight (1.0), commas, a last line without
    # a newline, CRLF line ends, a header in any case
    src, dst, w, n, labels = _read(tmp_path, "Source TARGET\r\nb a\r\na c\tw\nc,b,2.5")
    assert labels == ["b", "a", "c"] and n == 3
    assert list(zip(src.tolist(), dst.tolist(), w.tolist())) == [(0, 1, 1.0), (1, 2, 1.0), (2, 0, 2.5)]


def test_edgelist_number_forms(tmp_path):
    # std.fmt.parseInt(u32, _, 10) and std.fmt.parseFloat(f64, _): '+',
    # underscores between digits, exponents, hex floats, inf; anything else
    # skips the line (ids) or gives weight 1.0
    text = "\n".join(["+1 2 1e-3", "1_0 3 0x1p3", "4 5 -inf", "_4 5 2", "4 5_ 2", "-1 5 2", "4294967296 1 2",
                      "6 7 1_5.0_0", "6 7 1__5", "6 7 abc", "6 7 nan(1)", "6 7 +2.5", "8 9 .5", "8 9 1."])
    src, dst, w, n, _ = _read(tmp_path, text, integer_labels=True)
    assert list(zip(src.tolist(), dst.tolist())) == [(1, 2), (10, 3), (4, 5), (6, 7), (6, 7), (6, 7), (6, 7),
                                                      (6, 7), (8, 9), (8, 9)]
    assert w[:3].tolist() == [1e-3, 8.0, float("-inf")]
    assert w[3:].tolist() == [15.0, 1.0, 1.0, 1.0, 2.5, 0.5, 1.0]
    assert n == 11


def test_undirected_dedupe_rule(tmp_path):
    # addParsedEdges (resource_graph.ex:953-988): a stable sort by (min, max),
    # the first arc of an unordered pair in file order keeps its weight,
    # both arcs unless a self-loop
    text = "2 1 1.0\n0 1 4.0\n1 2 2.0\n3 3 3.0\n1 0 5.0\n"
    src, dst, w, _, _ = _read(tmp_path, text, directed=False, integer_labels=True)
    assert list(zip(src.tolist(), dst.tolist())) == [(0, 1), (1, 0), (2, 1), (1, 2), (3, 3)]
    assert w.tolist() == [4.0, 4.0, 1.0, 1.0, 3.0]
    src, dst, w, _, _ = _read(tmp_path, text, directed=True, integer_labels=True)
    assert list(zip(src.tolist(), dst.tolist())) == [(2, 1), (0, 1), (1, 2), (3, 3), (1, 0)]
    assert w.tolist() == [1.0, 4.0, 2.0, 3.0, 5.0]


def test_edgelist_reference_fixture_and_errors(tmp_path):
    # the reference's fixture test/fixtures/wiki_vote_sample.txt read as
    # labels: ids in first-seen order, as tests/graphs.wiki_vote builds them
    from tests import graphs
    src, dst, w, n, labels = rg._read_edgelist_native(graphs.WIKI_VOTE, True, False)
    n2, s2, d2 = graphs.wiki_vote()
    assert n == n2 and np.array_equal(src, s2) and np.array_equal(dst, d2) and np.all(w == 1.0)
    with pytest.raises(FileNotFoundError):
        rg._read_edgelist_native(str(tmp_path / "missing.txt"), True, False)
    # an id whose node count would not fit in u32
    with pytest.raises(_lib.ZogError):
        _read(tmp_path, "4294967295 0\n", integer_labels=True)
    # empty file
    src, dst, w, n, labels = _read(tmp_path, "", integer_labels=False)
    assert len(src) == 0 and n == 0 and labels == []


def test_generators_deterministic():
    from zog_amd import generators as G
    a = G.erdos_renyi_directed(500, 0.02, 11)
    b = G.erdos_renyi_directed(500, 0.02, 11)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert np.all(a[1] != a[2])
    keys = a[1].astype(np.int64) * 500 + a[2]
    assert len(np.unique(keys)) == len(keys)
    n, s, d = G.rmat(10, 5000, 3)
    assert n == 1024 and len(s) == 5000
    assert len(np.unique(s.astype(np.int64) * n + d)) == 5000
    n, s, d = G.grid_2d(4, 5, seed=1, delete=3)
    assert n == 20 and len(s) == 2 * (4 * 4 + 3 * 5 - 3)
    n, s, d = G.barabasi_albert(50, 4, 5)
    assert len(s) == 2 * (6 + 46 * 4)
    assert np.array_equal(s, G.barabasi_albert(50, 4, 5)[1])
    # clique first, then each new node's 4 distinct targets in ascending order
    assert list(zip(s[:4], d[:4])) == [(0, 1), (1, 0), (0, 2), (2, 0)]
    t = d[12:20:2]
    assert np.all(s[12:20:2] == 4) and len(set(t.tolist())) == 4 and np.all(np.diff(t) > 0)
    deg = np.bincount(s, minlength=50)
    assert deg.min() >= 3 and deg.sum() == len(s)
