"""Tree links on the HIP path (parent, subtree_end, first_child, next_sibling),
bit-exact against the oracle (crates/oneline-gpu/src/tests.rs:450-647,1061-1125)."""
import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu
NAMES = ("parent", "subtree_end", "first_child", "next_sibling")


def text_from_depths(lines):
    """One line per (depth, blank); blanks alternate empty and prefix-only
    (crates/oneline-gpu/src/tests.rs:456-473)."""
    out = []
    for i, (d, blank) in enumerate(lines):
        s = "|" * d
        if blank:
            if i % 2 == 0:
                s = ""
        else:
            s += " x"
        out.append(s + "\n")
    return "".join(out)


def gpu_links(parser, text):
    doc = parser.upload(text, nodes=False)
    tree = parser.build_tree(doc)
    assert tree.line_count() == doc.line_count()
    return parser.download_links(tree)


def assert_tree_matches(parser, text, what=""):
    _, nodes = orc.parse(ol.as_bytes_array(text))
    want = orc.links(nodes)
    got = gpu_links(parser, text)
    for name, g, w in zip(NAMES, (got.parent, got.subtree_end, got.first_child, got.next_sibling), want):
        assert g.size == w.size, f"{what}: {name} length"
        if not np.array_equal(g, w):
            i = int(np.nonzero(g != w)[0][0])
            raise AssertionError(f"{what}: {name} differs at line {i}: gpu {g[i]} cpu {w[i]}")


def test_known_answers(parser, known):
    for case in known["tree_depths"]:
        text = text_from_depths([(d, b) for d, b in case["lines"]])
        links = gpu_links(parser, text)
        assert links.parent.tolist() == case["parent"]
        assert links.subtree_end.tolist() == case["subtree_end"]
    for case in known["tree_text"]:
        links = gpu_links(parser, case["text"])
        for line, lo, hi in case["subtree"]:
            assert (line, int(links.subtree_end[line])) == (lo, hi)
        for line, par in case["parent_of"]:
            assert links.parent_of(line) == par


def test_generated_documents(parser):
    assert_tree_matches(parser, datagen.document(1 << 20, seed=0), "1 MiB seed 0")
    assert_tree_matches(parser, datagen.document(29 << 20, seed=21, crlf=10, empty=5), "29 MiB")


def test_run_and_tile_boundaries(parser):
    rng = np.random.default_rng(5)
    for n in (1, 2, 15, 16, 17, 31, 32, 33, 255, 256, 257, 1000, 2047, 2048, 2049, 4095, 4096, 4097,
              6145, 8191, 8192, 8193, 12289, 20481):
        lines = [(int(rng.integers(5)), bool(rng.random() < 0.15)) for _ in range(n)]
        assert_tree_matches(parser, text_from_depths(lines), f"{n} random lines")


def test_adversarial_shapes(parser):
    n = 4096 * 5 + 123
    rng = np.random.default_rng(11)
    walk = np.clip(np.cumsum(rng.integers(-1, 2, n)), 0, 31)
    shapes = {
        "random levels over all 32": [(int(rng.integers(32)), bool(rng.random() < 0.1)) for _ in range(n)],
        "random walk over all 32": [(int(d), bool(rng.random() < 0.05)) for d in walk],
        "flat depth 1": [(1, False)] * n,
        "flat depth 0": [(0, False)] * n,
        "header then flat": [(0, False)] + [(1, False)] * (n - 1),
        "flat then header at the end": [(1, False)] * (n - 1) + [(0, False)],
        "strictly increasing to the limit": [(i % 32, False) for i in range(n)],
        "strictly decreasing sawtooth": [(31 - i % 32, False) for i in range(n)],
        "alternating": [(i % 2, False) for i in range(n)],
        "only blanks": [(3, True)] * n,
        "blanks between levels": [(i % 4, i % 3 == 0) for i in range(n)],
        "deepest allowed": [(31, False)] * n,
        "long blank runs across tiles": [(i // 5000 % 3, (i % 5000) > 3) for i in range(n)],
        "deep then shallow across tiles": [(min(31, i // 700), False) for i in range(n)] + [(0, False)],
    }
    for name, lines in shapes.items():
        assert_tree_matches(parser, text_from_depths(lines), name)


def test_every_tile_geometry(parser):
    """The tuner's tree geometries (64 x rounds lines per wave): tile edges,
    random depths and blank runs under each."""
    rng = np.random.default_rng(9)
    n = 8192 * 3 + 77
    lines = [(int(rng.integers(8)), bool(rng.random() < 0.2)) for _ in range(n)]
    text = text_from_depths(lines)
    deep = text_from_depths([(min(31, i // 300), i % 97 == 0) for i in range(n)] + [(0, False)])
    every = text_from_depths([(int(rng.integers(32)), bool(rng.random() < 0.1)) for _ in range(n)])
    try:
        for rounds in (4, 1, 16):
            parser.tune_set("tree_rounds", rounds)
            assert_tree_matches(parser, text, f"random, {rounds} rounds")
            assert_tree_matches(parser, deep, f"deep runs, {rounds} rounds")
            assert_tree_matches(parser, every, f"random over all 32 levels, {rounds} rounds")
    finally:
        parser.tune_set("tree_rounds", 16)


def test_depth_limit_and_recovery(parser, known):
    c = known["tree_depth_limit"]
    parser.build_tree(parser.upload(c["ok"]))
    with pytest.raises(ol.OnelineError) as e:
        parser.build_tree(parser.upload(c["too_deep"]))
    assert c["error_mentions"] in str(e.value)
    parser.build_tree(parser.upload(c["blank_deep_ok"]))
    assert_tree_matches(parser, "| a\n|| b\n| c\n", "after an error")


def test_empty_document(parser):
    tree = parser.build_tree(parser.upload(b""))
    assert tree.line_count() == 0
    assert parser.download_links(tree).parent.size == 0


def test_reuse_matches_fresh_builds(parser):
    docs = [datagen.document(3 << 20, seed=1), datagen.document(600 << 10, seed=2),
            datagen.document(6 << 20, seed=3, crlf=10, empty=5), datagen.document(3 << 20, seed=1),
            b"| a\n|| b\n", b"", datagen.document(3 << 20, seed=4)]
    tree = None
    for i, text in enumerate(docs):
        doc = parser.upload(text, nodes=False)
        if i == 4:
            deep = parser.upload("|" * 40 + " x\n")
            with pytest.raises(ol.OnelineError):
                parser.build_tree_reusing(deep, tree)
            tree = None
        tree = parser.build_tree_reusing(doc, tree)
        _, nodes = orc.parse(ol.as_bytes_array(text))
        want = orc.links(nodes)
        got = parser.download_links(tree)
        for name, g, w in zip(NAMES, (got.parent, got.subtree_end, got.first_child, got.next_sibling), want):
            assert np.array_equal(g, w), f"doc {i}: {name}"
