"""Kind / depth queries (ol_query_document, ol_query_subtree, ol_query_raw)
against the oracle's nodes: the predicate of filter_nodes
(crates/oneline-gpu/src/shader.wgsl:152-179), lines in ascending order, the
true total kept when the indices are truncated."""
import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def want(nodes, kind, depth, r0=0, r1=None):
    r1 = nodes.size if r1 is None else r1
    sel = np.ones(nodes.size, bool)
    if kind:
        sel &= nodes["kind"] == kind
    if depth is not None:
        sel &= nodes["depth"] == depth
    idx = np.flatnonzero(sel)
    return idx[(idx >= r0) & (idx < r1)].astype(np.uint32)


def mixed_text():
    body = datagen.document(1 << 20, seed=61, crlf=20, empty=30).tobytes()
    extra = (b"| | \n|\t|  \n| | | \r\n" + b"|" * 300 + b" deep\n" + b"|" * 260 + b" \n"
             + b"= h\n:c\n@e\n- l\n-. o\n// c\n-x\n/x\n")
    return extra + body + extra


@pytest.mark.parametrize("nodes_kept", [True, False])
def test_query_document_against_oracle(parser, nodes_kept):
    text = mixed_text()
    _, nodes = orc.parse(np.frombuffer(text, np.uint8))
    doc = parser.upload(text, nodes=nodes_kept)
    for kind in (None, 1, 2, 4, 6, 7):
        for depth in (None, 0, 1, 2, 3, 7, 300, 260):
            got = parser.query_document(doc, kind, depth, max_matches=1 << 22)
            w = want(nodes, kind, depth)
            assert got.total_matches == w.size, (kind, depth)
            assert np.array_equal(got.indices, w), (kind, depth)


def test_query_truncation_keeps_total(parser):
    text = mixed_text()
    _, nodes = orc.parse(np.frombuffer(text, np.uint8))
    doc = parser.upload(text, nodes=False)
    w = want(nodes, 2, None)
    for m in (0, 1, 7, w.size - 1, w.size, w.size + 5):
        got = parser.query_document(doc, 2, None, max_matches=m)
        assert got.total_matches == w.size
        assert np.array_equal(got.indices, w[: min(m, w.size)])


def test_query_subtree(parser):
    text = datagen.document(256 << 10, seed=62, crlf=10, empty=40).tobytes()
    _, nodes = orc.parse(np.frombuffer(text, np.uint8))
    parent, se, _, _ = orc.links(nodes)
    doc = parser.upload(text, nodes=False)
    tree = parser.build_tree(doc)
    rng = np.random.default_rng(3)
    roots = list(rng.integers(0, nodes.size, 40)) + [0, nodes.size - 1]
    roots += [int(i) for i in np.flatnonzero(nodes["kind"] == 0)[:3]]   # blank roots: the line alone
    for root in roots:
        for kind, depth in ((None, None), (4, None), (None, int(nodes["depth"][root]) + 1)):
            got = parser.query_subtree(doc, tree, int(root), kind, depth)
            w = want(nodes, kind, depth, int(root), int(se[root]))
            assert got.total_matches == w.size and np.array_equal(got.indices, w), (root, kind, depth)
    with pytest.raises(ol.OnelineError):
        parser.query_subtree(doc, tree, nodes.size, None, None)


def test_query_raw_and_empty(parser):
    text = mixed_text()
    doc = parser.upload(text, nodes=True)
    a = parser.query_document(doc, 4, 2)
    b = parser.query_raw(text, 4, 2)
    assert a.total_matches == b.total_matches and np.array_equal(a.indices, b.indices)
    e = parser.query_raw(b"", None, None)
    assert e.total_matches == 0 and e.indices.size == 0
    one = parser.query_raw(b"| | x", None, 2)
    assert one.total_matches == 1 and list(one.indices) == [0]


def test_query_subtree_across_blocks(parser):
    # a root off the 16-line boundary whose subtree spans several blocks of
    # 4,096 lines and ends inside a 16-line word
    rng = np.random.default_rng(64)
    bodies = [b"x", b"= h", b":c", b"@e", b"- l", b"-. o", b"// c", b"-x", b"/x", b""]
    lines = [b"top %d" % i for i in range(7)] + [b"root"]
    depth = 1
    for _ in range(10_001):
        depth = int(np.clip(depth + rng.integers(-1, 2), 1, 5))
        lines.append(b"| " * depth + bodies[rng.integers(len(bodies))])
    lines += [b"after", b"| tail"]
    text = b"\n".join(lines) + b"\n"
    _, nodes = orc.parse(np.frombuffer(text, np.uint8))
    _, se, _, _ = orc.links(nodes)
    assert se[7] == nodes.size - 2 and se[7] % 16 != 0
    doc = parser.upload(text, nodes=False)
    tree = parser.build_tree(doc)
    for root in (7, 8, 9, 4100, 8195):
        for kind, depth in ((None, None), (4, None), (None, 3), (2, 2), (6, 5)):
            got = parser.query_subtree(doc, tree, root, kind, depth)
            w = want(nodes, kind, depth, root, int(se[root]))
            assert got.total_matches == w.size and np.array_equal(got.indices, w), (root, kind, depth)


def test_query_with_given_offsets(parser):
    """GpuParser::query (crates/oneline-gpu/src/query.rs:288-314): the lines are
    the caller's offsets, each parsed from its bytes; against the oracle's
    per-line parse of the same spans."""
    text = mixed_text()
    t = np.frombuffer(text, np.uint8)
    offs, nodes = orc.parse(t)
    for kind in (None, 2, 4):
        for depth in (None, 0, 3, 300):
            got = parser.query(text, offs, kind, depth, max_matches=1 << 22)
            w = want(nodes, kind, depth)
            assert got.total_matches == w.size and np.array_equal(got.indices, w), (kind, depth)
    # offsets that are not the scanner's: every third line start, so spans
    # hold several lines; each span parsed as parse() parses a given line
    sub = offs[::3].copy()
    ends = np.append(sub[1:], t.size)
    sub_nodes = np.zeros(sub.size, orc.NODE)
    for i, (s, e) in enumerate(zip(sub, ends)):
        sub_nodes[i] = orc.parse_line(text[s:e], int(s))
    for kind, depth in ((None, None), (1, None), (None, 2), (4, 1)):
        got = parser.query(text, sub, kind, depth, max_matches=1 << 22)
        w = want(sub_nodes, kind, depth)
        assert got.total_matches == w.size and np.array_equal(got.indices, w), (kind, depth)
    # truncation keeps the true total; no offsets: nothing
    got = parser.query(text, offs, None, None, max_matches=5)
    assert got.total_matches == offs.size and np.array_equal(got.indices, np.arange(5, dtype=np.uint32))
    empty = parser.query(text, np.zeros(0, np.uint32), 2, None)
    assert empty.total_matches == 0 and empty.indices.size == 0


def test_per_call_profiles(parser):
    """The profile_* record: host steps in execution order whose sum is the
    wall time, and the call's GPU passes by name."""
    text = datagen.document(4 << 20, seed=3)
    doc, prof = parser.profile_upload(text, nodes=False)
    names = [k for k, _ in prof.gpu_passes]
    assert "scan_classify" in names
    steps = [k for k, _ in prof.host_steps]
    assert "upload: stream through the staging ring" in steps and "scan: launch" in steps
    assert prof.wall > 0 and abs(sum(d for _, d in prof.host_steps) - prof.wall) < 1e-6
    assert prof.gpu_total() > 0
    pm = parser.profile_search_many(doc, parser.compile_patterns(doc, datagen.patterns(256, seed=1)), 1 << 20)
    assert any(k == "many" for k, _ in pm.gpu_passes) and any(k.startswith("many:") for k, _ in pm.host_steps)
    tree, pt = parser.profile_build_tree_reusing(doc, None)
    assert [k for k, _ in pt.gpu_passes] == ["tree"] and tree.line_count() == doc.line_count()
    pq = parser.profile_query_document(doc, 2, None)
    assert any(k == "query" for k, _ in pq.gpu_passes)
    # a call outside a record leaves nothing behind
    parser.search_document(doc, b"the", 10)
    ps = parser.profile_search_document(doc, b"the", 10)
    assert [k for k, _ in ps.gpu_passes] == ["search"]
    tree.close()
    doc.close()
