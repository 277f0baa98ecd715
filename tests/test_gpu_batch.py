"""Batches (ol_batch_*): many documents in one launch sequence, each result
equal to the oracle run on that document alone -- line offsets relative to
the document, classes, tree links local to it, dictionary matches that stay
inside it."""
import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

EDGE_DOCS = [
    b"",
    b"x",
    b"\n",
    b"| | a",                      # unterminated, nested
    b"\r",
    b"line\r\n",
    b"",
    b"| x\n|| y\n||| z\n| w",   # no final newline
    b"= head\n- item\n-. ord\n// c\n@el\n:conn\n",
    b"\n\n\n",
    b"|" * 31 + b" deep\n" + b"a" * 5000,   # deepest level the tree tracks, then a long line
]


def edge_docs():
    return list(EDGE_DOCS)


def generated_docs(count, seed, median, hi):
    sizes = datagen.lognormal_sizes(count, seed, median, 2.0, 64, hi)
    return [datagen.document(int(sz), seed=datagen.mix(seed, i)).tobytes() for i, sz in enumerate(sizes)]


def check_doc(batch, i, doc):
    off, cls, links = batch.document(i)
    want_off, nodes = orc.parse(ol.as_bytes_array(doc)) if doc else (np.zeros(0, np.uint32), None)
    assert np.array_equal(off, want_off), f"doc {i}: offsets"
    if not doc:
        return
    assert np.array_equal(cls, nodes["kind"].astype(np.uint8)), f"doc {i}: classes"
    want = orc.links(nodes)
    for name, g, w in zip(("parent", "subtree_end", "first_child", "next_sibling"),
                          (links.parent, links.subtree_end, links.first_child, links.next_sibling), want):
        assert np.array_equal(g, w), f"doc {i}: {name}"


def check_all(batch, docs):
    total, st, cnt = batch.lines()
    assert total == sum(int(c) for c in cnt)
    for i, d in enumerate(docs):
        check_doc(batch, i, d)


def test_edge_documents(parser):
    docs = edge_docs()
    b = parser.batch(docs)
    assert len(b) == len(docs) and b.text_bytes() == sum(len(d) for d in docs)
    b.run()
    check_all(b, docs)
    _, st, cnt = b.lines()
    assert cnt[0] == 0 and cnt[6] == 0 and st[0] == 0


def test_generated_documents_all_tree_tiles(parser):
    docs = generated_docs(150, seed=1, median=8 << 10, hi=2 << 20) + edge_docs()
    b = parser.batch(docs)
    for rounds in (16, 4, 1):
        parser.tune_set("tree_rounds", rounds)
        b.run()
        check_all(b, docs)
    parser.tune_set("tree_rounds", 16)


def test_staged_upload_of_a_large_batch(parser):
    """A batch above the staging threshold (32 MiB) goes through several copy
    workers and many 8 MiB slots.  Documents end one byte before, at and after
    slot boundaries, every other one without its final newline (the batch
    appends one), with empty documents between."""
    slot = 8 << 20
    docs, pos = [], 0
    for i in range(6):
        end = (i + 1) * slot + (i % 3) - 1   # where the document's bytes (and newline) end
        span = end - pos
        d = datagen.document(span, seed=300 + i).tobytes()
        if i % 2:
            d = d[:-2] + b"x"   # no final newline: the batch adds it back, same span
        docs += [d, b""]
        pos = end
    docs += edge_docs()
    b = parser.batch(docs)
    assert b.text_bytes() == sum(len(d) for d in docs) and b.text_bytes() > (40 << 20)
    b.run()
    check_all(b, docs)


def test_many_tiny_documents(parser):
    """Thousands of documents of 64..300 bytes: many documents per scan tile."""
    rng = np.random.default_rng(5)
    docs = [datagen.document(int(n), seed=100 + i).tobytes() for i, n in enumerate(rng.integers(64, 300, 4000))]
    b = parser.batch(docs)
    b.run()
    check_all(b, docs)


def test_scan_tile_sizes(parser):
    docs = generated_docs(60, seed=3, median=16 << 10, hi=1 << 20)
    b = parser.batch(docs)
    for bpt in (64, 128, 256):
        parser.tune_set("scan_bpt", bpt)
        b.run()
        check_all(b, docs)
    parser.tune_set("scan_bpt", 128)


def test_dictionary_matches_stay_inside_documents(parser):
    docs = generated_docs(80, seed=7, median=4 << 10, hi=256 << 10) + edge_docs()
    pats = datagen.patterns(48, seed=0xC3B)
    # patterns that would match across a document boundary (end of one, start of the next)
    for i in range(0, 20, 2):
        a, c = docs[i], docs[i + 1]
        if len(a) >= 3 and len(c) >= 2:
            pats.append(a[-3:] + (b"\n" if not a.endswith(b"\n") else b"") + c[:2])
            pats.append(a[-2:] + c[:3])
    pats.append(b"\n")
    b = parser.batch(docs)
    pset = parser.compile_patterns(None, pats)
    b.run(pset)
    for i, d in enumerate(docs):
        off, ids = b.matches(i)
        want_off, want_id = orc.many_find_all(ol.as_bytes_array(d), pats) if d else (np.zeros(0), np.zeros(0))
        assert np.array_equal(off, want_off.astype(np.uint32)), f"doc {i}: match offsets"
        assert np.array_equal(ids, want_id.astype(np.uint32)), f"doc {i}: match ids"


def test_graph_replays_equal_eager_runs(parser):
    docs = generated_docs(40, seed=9, median=16 << 10, hi=1 << 20)
    pats = datagen.patterns(64, seed=0xC3B)
    b = parser.batch(docs)
    pset = parser.compile_patterns(None, pats)
    b.run(pset)                 # sizes
    want = [b.document(i) for i in range(len(docs))]
    want_m = [b.matches(i) for i in range(len(docs))]
    check_all(b, docs)
    other = parser.upload(datagen.document(1 << 20, seed=44), nodes=False)
    for rep in range(3):
        b.run(pset, graph=True)
        parser.synchronize()
        for i in range(len(docs)):
            off, cls, links = b.document(i)
            w_off, w_cls, w_links = want[i]
            assert np.array_equal(off, w_off) and np.array_equal(cls, w_cls), f"replay {rep} doc {i}"
            for g, w in zip((links.parent, links.subtree_end, links.first_child, links.next_sibling),
                            (w_links.parent, w_links.subtree_end, w_links.first_child, w_links.next_sibling)):
                assert np.array_equal(g, w), f"replay {rep} doc {i}: links"
            mo, mi = b.matches(i)
            assert np.array_equal(mo, want_m[i][0]) and np.array_equal(mi, want_m[i][1]), f"replay {rep} doc {i}"
        # eager work on the same context between replays
        parser.build_tree(other)
        b.run(pset, graph=False)
        parser.synchronize()


def test_empty_and_all_empty_batches(parser):
    b = parser.batch([])
    b.run()
    assert b.lines()[0] == 0
    b = parser.batch([b"", b""])
    b.run()
    total, st, cnt = b.lines()
    assert total == 0 and list(cnt) == [0, 0]


def test_graph_replay_after_the_context_scratch_grew(parser):
    """A captured batch graph keeps its own scratch: growing every scratch
    buffer of the context (a larger upload, a dictionary search with written
    matches, a literal search) between replays changes none of its results."""
    docs = generated_docs(30, seed=12, median=8 << 10, hi=256 << 10)
    pats = datagen.patterns(128, seed=0xC3B)
    b = parser.batch(docs)
    pset = parser.compile_patterns(None, pats)
    b.run(pset)
    want = [b.document(i) for i in range(len(docs))]
    want_m = [b.matches(i) for i in range(len(docs))]
    b.run(pset, graph=True)
    parser.synchronize()
    big = parser.upload(datagen.document(96 << 20, seed=45), nodes=False)
    parser.search_many(big, pset, 1 << 20)            # candidate store and status words grow
    parser.search_document(big, b"the", 1 << 10)
    parser.build_tree(big)
    for rep in range(2):
        b.run(pset, graph=True)
        parser.synchronize()
        for i in range(len(docs)):
            off, cls, links = b.document(i)
            assert np.array_equal(off, want[i][0]) and np.array_equal(cls, want[i][1]), f"replay {rep} doc {i}"
            for g, w in zip((links.parent, links.subtree_end, links.first_child, links.next_sibling),
                            (want[i][2].parent, want[i][2].subtree_end, want[i][2].first_child,
                             want[i][2].next_sibling)):
                assert np.array_equal(g, w), f"replay {rep} doc {i}: links"
            mo, mi = b.matches(i)
            assert np.array_equal(mo, want_m[i][0]) and np.array_equal(mi, want_m[i][1]), f"replay {rep} doc {i}"
    big.close()
    b.close()


def test_c4_workload_after_graph_replays(parser):
    """The batched workload at its size: 10,000 documents (log-normal sizes,
    median 16 KiB, sigma 2, 64 B .. 64 MiB, seed 1) with the 4,096-pattern
    dictionary, run eagerly, replayed as a graph (with eager work between the
    replays), then a seeded sample of documents checked against the oracle
    run on each document alone, and the totals checked against the per-
    document results."""
    count = 10_000
    sizes = datagen.lognormal_sizes(count, 1, 16384.0, 2.0, 64, 64 << 20)
    print(f"[c4] {count} documents, {int(sum(int(x) for x in sizes))} bytes", flush=True)
    docs = [datagen.document(int(sz), seed=datagen.mix(1, i)) for i, sz in enumerate(sizes)]
    pats = datagen.patterns(4096, seed=0xC3B)
    pset = parser.compile_patterns(None, pats)
    b = parser.batch(docs)
    b.run(pset)                                    # sizing run
    total, st, cnt = b.lines()
    assert total == int(cnt.astype(np.int64).sum())
    other = parser.upload(datagen.document(8 << 20, seed=77), nodes=False)
    for _ in range(2):
        b.run(pset, graph=True)
        parser.synchronize()
        parser.search_many(other, pset, 1 << 16)  # eager work on the context between replays
        b.run(pset, graph=False)
    b.run(pset, graph=True)
    parser.synchronize()
    rng = np.random.default_rng(21)
    order = rng.permutation(count)
    checked, nbytes = 0, 0
    for i in order:
        d = docs[int(i)]
        if nbytes + d.size > (96 << 20):
            continue
        nbytes += d.size
        check_doc(b, int(i), d.tobytes())
        mo, mi = b.matches(int(i))
        wo, wi = orc.many_find_all(d, pats)
        assert np.array_equal(mo, wo.astype(np.uint32)) and np.array_equal(mi, wi.astype(np.uint32)), f"doc {i}"
        checked += 1
        if checked == 60:
            break
    assert checked == 60
    # the largest document too
    j = int(np.argmax(sizes))
    check_doc(b, j, docs[j].tobytes())
    other.close()
    b.close()
