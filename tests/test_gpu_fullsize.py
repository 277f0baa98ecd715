"""Searches over the benchmark's 2 GiB synthetic document, checked through
properties that need no CPU pass over the whole text: the dictionary's hits
are grouped by pattern with ascending offsets, a sample of them spell their
pattern in the text, and per-pattern counts and offsets equal those of the
literal search kernel (a different algorithm) for a sample of patterns; the
query kernels are checked at the same size against the downloaded nodes.
Smaller documents are compared with the oracle hit by hit in
test_gpu_many.py and test_gpu_search.py."""
import numpy as np
import pytest

from oneline_amd import datagen

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
