"""Serving (ol_server_*): continuous batching of parse and dictionary-search
requests on slots with persistent arenas, resident documents and
dictionaries.  Each result equals the oracle's and is the same whether the
request ran alone or inside a mix of concurrent requests."""
import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def oracle_parse(doc):
    off, nodes = orc.parse(ol.as_bytes_array(doc))
    return off, nodes["kind"].astype(np.uint8), orc.links(nodes)


def assert_parse_equal(got, want, what):
    off, cls, links = got
    w_off, w_cls, w_links = want
    assert np.array_equal(off, w_off), f"{what}: offsets"
    assert np.array_equal(cls, w_cls), f"{what}: classes"
    for g, w in zip((links.parent, links.subtree_end, links.first_child, links.next_sibling), w_links):
        assert np.array_equal(g, w), f"{what}: links"


def mixed_docs(count, seed):
    sizes = datagen.lognormal_sizes(count, seed, 16384.0, 2.0, 64, 4 << 20)
    return [datagen.document(int(sz), seed=datagen.mix(seed, i)).tobytes() for i, sz in enumerate(sizes)]


def test_result_alone_equals_result_in_mix():
    srv = ol.GpuServer(slots=4, max_batch_docs=64)
    try:
        pats = datagen.patterns(1024, seed=0xC3B)
        d0 = srv.add_dictionary(pats)
        d1 = srv.add_dictionary(datagen.patterns(256, seed=77))
        x = datagen.document(1 << 20, seed=31, crlf=10, empty=5).tobytes()
        want = oracle_parse(x)
        w_off, w_id = orc.many_find_all(ol.as_bytes_array(x), pats)

        # alone
        rp = srv.submit(x, ol.PARSE)
        rs = srv.submit(x, ol.SEARCH, d0)
        srv.wait(rp)
        srv.wait(rs)
        alone_p = srv.parse_result(rp)
        alone_s = srv.search_result(rs)
        assert_parse_equal(alone_p, want, "alone")
        assert np.array_equal(alone_s[0], w_off.astype(np.uint32)) and np.array_equal(alone_s[1], w_id)

        # in a mix of concurrent requests
        docs = mixed_docs(300, seed=4)
        rng = np.random.default_rng(2)
        ids = []
        for i, d in enumerate(docs):
            kind = ol.SEARCH if rng.random() < 0.3 else ol.PARSE
            ids.append((srv.submit(d, kind, int(rng.integers(0, 2)) if kind == ol.SEARCH else 0), kind, i))
            if i == 150:
                mp = srv.submit(x, ol.PARSE)
                ms = srv.submit(x, ol.SEARCH, d0)
        for rid, _, _ in ids:
            srv.wait(rid)
        srv.wait(mp)
        srv.wait(ms)
        assert_parse_equal(srv.parse_result(mp), want, "in the mix")
        mix_s = srv.search_result(ms)
        assert np.array_equal(mix_s[0], alone_s[0]) and np.array_equal(mix_s[1], alone_s[1])

        # a sample of the mix against the oracle
        for rid, kind, i in ids[::25]:
            if kind == ol.PARSE:
                assert_parse_equal(srv.parse_result(rid), oracle_parse(docs[i]), f"mix doc {i}")
        # the requests ran batched, and batches overlapped in time
        spans = sorted(srv.timing(rid)[1:] for rid, _, _ in ids)
        starts = sorted({s for s, _ in spans})
        assert len(starts) < len(ids), "requests were batched"
        overlapped = any(spans[k + 1][0] < spans[k][1] and spans[k + 1][0] != spans[k][0]
                         for k in range(len(spans) - 1))
        assert overlapped, "batches overlapped on the device"
    finally:
        srv.close()


def test_hbm_budget_admission():
    # arenas sized for 4 MiB batches: two slots fit a 512 MiB budget, sixteen do not
    with pytest.raises(ol.OnelineError):
        ol.GpuServer(hbm_budget=512 << 20, slots=16, max_batch_bytes=4 << 20)
    srv = ol.GpuServer(hbm_budget=512 << 20, slots=2, max_batch_bytes=4 << 20, max_batch_docs=16)
    try:
        arenas = srv.hbm_used()
        assert arenas == 2 * srv.arena_bytes() <= 256 << 20
        with pytest.raises(ol.OnelineError):          # larger than a batch
            srv.submit(np.zeros(5 << 20, np.uint8) + ord("a"), ol.PARSE)
        docs = [datagen.document(3 << 20, seed=50 + i).tobytes() for i in range(6)]
        ids = [srv.submit(d, ol.PARSE) for d in docs]
        for rid, d in zip(ids, docs):
            srv.wait(rid)
            assert_parse_equal(srv.parse_result(rid), oracle_parse(d), "budgeted")
        # one 3 MiB document per 4 MiB batch: each batch ran alone on a slot
        assert len({srv.timing(rid)[1] for rid in ids}) == len(ids)
        assert {srv.slot_of(rid) for rid in ids} <= {0, 1}
        # a resident document is charged at its device size, and released
        big = datagen.document(8 << 20, seed=60).tobytes()
        did = srv.put_document(big)
        assert srv.hbm_used() > arenas + (8 << 20)
        srv.drop_document(did)
        assert srv.hbm_used() == arenas
        with pytest.raises(ol.OnelineError):          # beyond the budget
            srv.put_document(np.zeros(400 << 20, np.uint8) + ord("a"))
        assert srv.hbm_used() == arenas
    finally:
        srv.close()


def test_resident_documents():
    """Requests on resident documents skip the upload and the parse; their
    results equal the oracle's and those of the same bytes submitted."""
    srv = ol.GpuServer(slots=2, max_batch_bytes=64 << 20)
    try:
        pats = datagen.patterns(4096, seed=0xC3B)
        d0 = srv.add_dictionary(pats)
        x = datagen.document(5 << 20, seed=33, crlf=10, empty=5).tobytes()
        did = srv.put_document(x)
        rp = srv.submit_resident(did, ol.PARSE)
        rs = srv.submit_resident(did, ol.SEARCH, d0)
        bp = srv.submit(x, ol.PARSE)
        bs = srv.submit(x, ol.SEARCH, d0)
        for r in (rp, rs, bp, bs):
            srv.wait(r)
        want = oracle_parse(x)
        assert_parse_equal(srv.parse_result(rp), want, "resident")
        assert_parse_equal(srv.parse_result(bp), want, "submitted")
        w_off, w_id = orc.many_find_all(ol.as_bytes_array(x), pats)
        for r in (rs, bs):
            off, ids = srv.search_result(r)
            assert np.array_equal(off, w_off.astype(np.uint32)) and np.array_equal(ids, w_id)
        # dropped while a request on it is queued: the request still completes
        r2 = srv.submit_resident(did, ol.SEARCH, d0)
        srv.drop_document(did)
        srv.wait(r2)
        assert np.array_equal(srv.search_result(r2)[0], w_off.astype(np.uint32))
        with pytest.raises(ol.OnelineError):
            srv.submit_resident(did, ol.PARSE)
    finally:
        srv.close()


def test_c5_workload_against_oracle():
    """The serving workload's shape: 8 resident dictionaries of 1,024 to 65,536
    patterns, documents log-normal (median 64 KiB, sigma 2, up to 256 MiB,
    seed 2) plus a 64 MiB one, 70% parse / 30% search requests submitted
    together; a sample of each kind checked against the oracle."""
    srv = ol.GpuServer(hbm_budget=96 << 30, slots=4, max_batch_docs=256, max_batch_bytes=256 << 20)
    try:
        dict_sizes = (1024, 2048, 4096, 8192, 16384, 32768, 65536, 1024)
        dicts = [datagen.patterns(cnt, seed=0xC3B + 17 * k) for k, cnt in enumerate(dict_sizes)]
        for p in dicts:
            srv.add_dictionary(p)
        sizes = list(datagen.lognormal_sizes(120, 2, 65536.0, 2.0, 64, 256 << 20)) + [64 << 20]
        docs = [datagen.document(int(sz), seed=datagen.mix(2, i)) for i, sz in enumerate(sizes)]
        print(f"[c5] {len(docs)} documents, {sum(int(x) for x in sizes)} bytes", flush=True)
        rng = np.random.default_rng(12)
        reqs = []
        for i, d in enumerate(docs):
            if rng.random() < 0.3:
                k = int(rng.integers(0, len(dicts)))
                reqs.append((srv.submit(d, ol.SEARCH, k), i, k))
            else:
                reqs.append((srv.submit(d, ol.PARSE), i, None))
        for rid, _, _ in reqs:
            srv.wait(rid, timeout=300)
        checked_p = checked_s = 0
        for rid, i, k in reqs:
            d = docs[i]
            if k is None and (checked_p < 25 or d.size > (32 << 20)):
                assert_parse_equal(srv.parse_result(rid), oracle_parse(d), f"c5 doc {i}")
                checked_p += 1
            elif k is not None and checked_s < 12 and d.size <= (4 << 20):
                off, ids = srv.search_result(rid)
                w_off, w_id = orc.many_find_all(d, dicts[k])
                assert np.array_equal(off, w_off.astype(np.uint32)) and np.array_equal(ids, w_id), f"c5 doc {i}"
                checked_s += 1
        assert checked_p >= 20 and checked_s >= 5
    finally:
        srv.close()


def test_bad_requests():
    srv = ol.GpuServer(slots=1)
    try:
        with pytest.raises(ol.OnelineError):
            srv.submit(b"abc\n", ol.SEARCH, 0)   # no dictionary yet
        with pytest.raises(ol.OnelineError):
            srv.submit(b"abc\n", 7)
        rid = srv.submit(b"", ol.PARSE)
        srv.wait(rid)
        off, cls, links = srv.parse_result(rid)
        assert off.size == 0
    finally:
        srv.close()


def test_open_loop_load():
    srv = ol.GpuServer(slots=4)
    try:
        srv.add_dictionary(datagen.patterns(1024, seed=0xC3B))
        srv.add_dictionary(datagen.patterns(4096, seed=0xC3C))
        docs = mixed_docs(64, seed=8)
        st = srv.load(docs, qps=500.0, requests=600, search_fraction=0.3, seed=3)
        assert st["requests"] == 600 and st["rejected"] == 0
        assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
        assert st["gbps"] > 0 and st["batches"] <= st["batched_requests"]
    finally:
        srv.close()
