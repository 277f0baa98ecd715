"""The benchmark's 2 GiB synthetic document (seed 0, 1% CRLF, 0.5% empty
lines).  Parse and tree: every line offset, class and link against the
oracle's parallel restatement (itself pinned against the sequential oracle in
test_oracle.py).  Searches: the dictionary's hits are grouped by pattern with
ascending offsets, a sample of them spell their pattern, per-pattern counts
and offsets equal those of the literal kernel (a different algorithm) for a
sample of patterns, and at 65,536 patterns every (offset, id) of sampled
windows equals the oracle's; the query kernels against the downloaded nodes; the
AST bridge's structural pass against the oracle.
Smaller documents are compared with the oracle hit by hit in
test_gpu_many.py and test_gpu_search.py."""
import ctypes
import os

import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

N = 1 << 31


@pytest.fixture(scope="module")
def big(parser):
    print(f"[fullsize] generating and uploading {N} bytes", flush=True)
    text = datagen.document(N, seed=0, crlf=10, empty=5)
    doc = parser.upload(text, nodes=False)
    yield text, doc
    doc.close()


def test_dictionary_1024_against_literal_searches(parser, big):
    text, doc = big
    pats = datagen.patterns(1024, seed=0xC3B)
    pset = parser.compile_patterns(doc, pats)
    r = parser.search_many(doc, pset, 1 << 24)
    pset.close()
    assert r.total_hits == int(r.counts.sum()) == r.hits.size > 1_000_000
    pid = r.hits["pattern"].astype(np.int64)
    off = r.hits["offset"].astype(np.int64)
    # pattern-major, offsets ascending within a pattern
    key = pid << 32 | off
    assert np.all(np.diff(key) > 0)
    assert np.array_equal(np.bincount(pid, minlength=len(pats)), r.counts)
    # a sample of hits spell their pattern
    rng = np.random.default_rng(5)
    for k in rng.choice(r.hits.size, 20_000, replace=False):
        p = pats[pid[k]]
        assert bytes(text[off[k]:off[k] + len(p)]) == p, (int(pid[k]), int(off[k]))
    # counts and offsets of a sample of patterns from the literal kernel
    for i in list(rng.choice(len(pats), 48, replace=False)) + [int(np.argmax(r.counts))]:
        lit = parser.search_document(doc, pats[i], 1 << 24)
        assert lit.total_matches == r.counts[i], (i, pats[i])
        assert np.array_equal(lit.offsets.astype(np.int64), off[pid == i]), (i, pats[i])


def test_dictionary_65536_counts_against_literal_searches(parser, big):
    _, doc = big
    pats = datagen.patterns(65536, seed=0xC3B)
    pset = parser.compile_patterns(doc, pats)
    r = parser.search_many(doc, pset, 0)
    pset.close()
    assert r.total_hits == int(r.counts.sum()) > 100_000_000 and r.hits.size == 0
    rng = np.random.default_rng(6)
    sample = list(rng.choice(len(pats), 96, replace=False)) + list(np.argsort(r.counts)[-4:])
    for i in sample:
        lit = parser.search_document(doc, pats[i], 0)
        assert lit.total_matches == r.counts[i], (int(i), pats[i])


def test_queries_against_downloaded_nodes(parser, big):
    # the query kernels at 2 GiB (lean document: keys, blank lines parsed again;
    # kept nodes: depths from the FlatNodes) against the predicate on the
    # downloaded nodes (the parse itself is checked against the oracle by the
    # benchmark at this size)
    text, doc = big
    kept = parser.upload(text, nodes=True)
    _, cls, nodes = kept.download(nodes=True)
    assert np.array_equal(cls, nodes["kind"])
    for kind, depth in ((4, None), (None, 3), (2, 5), (7, None), (None, 0)):
        sel = np.ones(nodes.size, bool)
        if kind is not None:
            sel &= nodes["kind"] == kind
        if depth is not None:
            sel &= nodes["depth"] == depth
        w = np.flatnonzero(sel).astype(np.uint32)
        for d in (doc, kept):
            got = parser.query_document(d, kind, depth, max_matches=1 << 26)
            assert got.total_matches == w.size and np.array_equal(got.indices, w), (kind, depth)
    kept.close()


def test_parse_and_tree_against_oracle(parser, big):
    text, doc = big
    threads = min(16, os.cpu_count() or 1)
    print(f"[fullsize] oracle parse + tree on {threads} threads", flush=True)
    w_off, w_nodes = orc.parse_par(text, threads)
    off, cls, _ = doc.download()
    assert off.size == w_off.size > 30_000_000
    assert np.array_equal(off, w_off), "offsets"
    assert np.array_equal(cls, w_nodes["kind"].astype(np.uint8)), "classes"
    tree = parser.build_tree(doc)
    got = parser.download_links(tree)
    want = orc.links_par(w_nodes, threads)
    for name, g, w in zip(("parent", "subtree_end", "first_child", "next_sibling"),
                          (got.parent, got.subtree_end, got.first_child, got.next_sibling), want):
        assert np.array_equal(g, w), name
    # the same tree rebuilt asynchronously with the three links of the benchmark's step
    del got
    links = ol.LINK_PARENT | ol.LINK_FIRST_CHILD | ol.LINK_NEXT_SIBLING
    parser.reparse_async(doc)
    parser.rebuild_tree_async(doc, tree, links)
    parser.check_tree(tree)
    L = tree.line_count()
    arrs = [np.zeros(L, np.uint32) for _ in range(3)]
    lib = ol.load_library()
    ol._check(lib.ol_tree_download(parser._ctx, tree._h, ol._ptr(arrs[0]), None, ol._ptr(arrs[1]),
                                   ol._ptr(arrs[2])))
    for name, g, w in zip(("parent", "first_child", "next_sibling"), arrs, (want[0], want[2], want[3])):
        assert np.array_equal(g, w), "async " + name
    # subtree_end was not written by that build: downloading it is an error
    with pytest.raises(ol.OnelineError):
        parser.download_links(tree)
    # the AST bridge's structural pass over the whole document (lean: depths of
    # blank lines and content spans parsed again from the text)
    print("[fullsize] bridge against the oracle", flush=True)
    flags, cont = parser.bridge(doc, tree)
    w_flags, w_cont = orc.bridge(text, w_nodes, want[0])
    assert np.array_equal(flags, w_flags), "bridge flags"
    assert np.array_equal(cont, w_cont), "bridge containers"
    tree.close()


def test_dictionary_65536_windows_against_oracle(parser, big):
    """Every (offset, id) of the offset-major device list inside sampled
    windows (and the text's last window) equals the oracle's."""
    text, doc = big
    pats = datagen.patterns(65536, seed=0xC3B)
    pset = parser.compile_patterns(doc, pats)
    lib = ol.load_library()
    hip = ctypes.CDLL("libamdhip64.so")
    tot = ctypes.c_uint64()
    ol._check(lib.ol_search_many_device(parser._ctx, doc._h, pset._h, None, None, 0, ctypes.byref(tot)))
    total = tot.value
    assert total > 100_000_000
    d_off, d_id = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_off), ctypes.c_size_t(total * 8)) == 0
    assert hip.hipMalloc(ctypes.byref(d_id), ctypes.c_size_t(total * 4)) == 0
    try:
        ol._check(lib.ol_search_many_device(parser._ctx, doc._h, pset._h, d_off, d_id, total, ctypes.byref(tot)))
        assert tot.value == total
        off = np.empty(total, np.uint64)
        ids = np.empty(total, np.uint32)
        assert hip.hipMemcpy(off.ctypes.data_as(ctypes.c_void_p), d_off, ctypes.c_size_t(total * 8), 2) == 0
        assert hip.hipMemcpy(ids.ctypes.data_as(ctypes.c_void_p), d_id, ctypes.c_size_t(total * 4), 2) == 0
    finally:
        hip.hipFree(d_off)
        hip.hipFree(d_id)
        pset.close()
    key = off.astype(np.uint64) << np.uint64(22) | ids.astype(np.uint64)
    assert np.all(np.diff(key) > 0), "sorted by (offset, id) without repeats"
    del key
    n = text.size
    W = 128 << 10
    rng = np.random.default_rng(7)
    starts = sorted(int(x) for x in rng.integers(0, n - W, 40)) + [n - W, 0]
    for w0 in starts:
        w1 = w0 + W
        lo, hi = np.searchsorted(off, w0), np.searchsorted(off, w1)
        wo, wi = orc.many_find_all(text[w0:min(n, w1 + 63)], pats)
        keep = wo < W
        assert np.array_equal(off[lo:hi] - w0, wo[keep]), f"window at {w0}: offsets"
        assert np.array_equal(ids[lo:hi], wi[keep]), f"window at {w0}: ids"
