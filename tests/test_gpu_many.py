"""Dictionary search on the HIP path against the oracle, in both output forms:
the offset-major (offset, id) list and the drop-in ManyResult (per-pattern
counts, hits sorted by (pattern, offset)) (crates/oneline-gpu/src/tests.rs:732-974)."""
import ctypes

import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def device_list(parser, doc, pset, cap):
    """Offset-major list via ol_search_many_device (device buffers via hipMalloc)."""
    lib = ol.load_library()
    hip = ctypes.CDLL("libamdhip64.so")
    d_off = ctypes.c_void_p()
    d_id = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_off), ctypes.c_size_t(max(1, cap) * 8)) == 0
    assert hip.hipMalloc(ctypes.byref(d_id), ctypes.c_size_t(max(1, cap) * 4)) == 0
    try:
        tot = ctypes.c_uint64()
        ol._check(lib.ol_search_many_device(parser._ctx, doc._h, pset._h, d_off, d_id, cap, ctypes.byref(tot)))
        keep = min(cap, tot.value)
        off = np.zeros(keep, np.uint64)
        ids = np.zeros(keep, np.uint32)
        if keep:
            assert hip.hipMemcpy(ctypes.c_void_p(off.ctypes.data), d_off, ctypes.c_size_t(keep * 8), 2) == 0
            assert hip.hipMemcpy(ctypes.c_void_p(ids.ctypes.data), d_id, ctypes.c_size_t(keep * 4), 2) == 0
        return tot.value, off, ids
    finally:
        hip.hipFree(d_off)
        hip.hipFree(d_id)


def assert_many_matches(parser, text, patterns, what=""):
    text = ol.as_bytes_array(text)
    want_off, want_id = orc.many_find_all(text, patterns)
    doc = parser.upload(text, nodes=False)
    pset = parser.compile_patterns(doc, patterns)
    assert len(pset) == len(patterns)
    tot, off, ids = device_list(parser, doc, pset, want_off.size)
    assert tot == want_off.size, f"{what}: total"
    assert np.array_equal(off, want_off), f"{what}: offsets"
    assert np.array_equal(ids, want_id), f"{what}: ids"
    counts = np.bincount(want_id, minlength=len(patterns)).astype(np.uint32)
    counted = parser.search_many(doc, pset, 0)
    assert np.array_equal(counted.counts, counts), f"{what}: counts (count-only)"
    assert counted.total_hits == want_off.size and counted.hits.size == 0
    full = parser.search_many(doc, pset, max(1, want_off.size))
    assert np.array_equal(full.counts, counts), f"{what}: counts"
    order = np.lexsort((want_off, want_id))
    assert np.array_equal(full.hits["pattern"], want_id[order]), f"{what}: hit patterns"
    assert np.array_equal(full.hits["offset"], want_off[order].astype(np.uint32)), f"{what}: hit offsets"


def test_corpus(parser, known):
    c = known["many_corpus"]
    assert_many_matches(parser, c["text"].encode(), [p.encode() for p in c["patterns"]], "nested/overlapping")
    for text, pats in c["also"]:
        assert_many_matches(parser, text.encode(), [p.encode() for p in pats], text)


def test_lengths_1_to_64_at_all_alignments(parser):
    text = b""
    pats = []
    for ln in range(1, 65):
        p = bytes(ord("a") + (i * 7 + ln) % 26 for i in range(ln))
        for shift in range(4):
            text += b"#" * shift + p + b"\n"
        pats.append(p)
    assert_many_matches(parser, text, pats, "lengths 1..=64")


def test_generated_text_and_dictionaries(parser):
    text = datagen.document(3 << 20, seed=12)
    rng = np.random.default_rng(77)
    for n in (1, 2, 7, 64, 500, 2048):
        pats = []
        for i in range(n):
            if i % 3 == 2:
                ln = int(rng.integers(1, 65))
                at = int(rng.integers(0, text.size - ln))
                pats.append(bytes(text[at:at + ln]))
            else:
                pats.append(datagen.vocab_word(int(rng.integers(0, 10_000))))
        assert_many_matches(parser, text, pats, f"{n} patterns")


def test_short_classes_share_one_filter(parser):
    """1- and 2-byte patterns pass through one 2-byte filter (a 1-byte pattern
    sets every second byte): each class alone, both, both with longer
    patterns, and matches on the text's last byte and first bytes."""
    text = bytearray(datagen.document(1 << 20, seed=21))
    text[0:2] = b"Qz"
    text[-1:] = b"Q"    # a 1-byte match on the last byte (no byte after it)
    text = bytes(text)
    ones = [b"Q", b"e", b"\x01"]
    twos = [b"Qz", b"th", b"\x01\x02", b"zQ"]
    longer = [datagen.vocab_word(i) for i in range(40)]
    for what, pats in (("1-byte", ones), ("2-byte", twos), ("1- and 2-byte", ones + twos),
                       ("1-byte and longer", ones + longer), ("2-byte and longer", longer + twos),
                       ("all", twos + longer + ones)):
        assert_many_matches(parser, text, pats, what)


def test_benchmark_dictionaries(parser):
    text = datagen.document(4 << 20, seed=0, crlf=10, empty=5)
    for count in (1024, 65536):
        assert_many_matches(parser, text, datagen.patterns(count, seed=0xC3B), f"{count}-pattern set")


def test_validation_and_truncation(parser):
    doc = parser.upload(datagen.document(2 << 20, seed=3))
    with pytest.raises(ol.OnelineError):
        parser.compile_patterns(doc, [])
    with pytest.raises(ol.OnelineError):
        parser.compile_patterns(doc, [b""])
    with pytest.raises(ol.OnelineError):
        parser.compile_patterns(doc, [b"a" * 65])
    pset = parser.compile_patterns(doc, [datagen.vocab_word(0), datagen.vocab_word(1)])
    all_ = parser.search_many(doc, pset, 0)
    assert all_.total_hits > 1000
    few = parser.search_many(doc, pset, 50)
    assert np.array_equal(few.counts, all_.counts) and few.total_hits == all_.total_hits
    assert few.hits.size == 50
    for _ in range(3):
        assert np.array_equal(parser.search_many(doc, pset, 0).counts, all_.counts)
    empty = parser.upload(b"")
    assert parser.search_many(empty, pset, 8).counts.tolist() == [0, 0]


def test_start_before_text_and_saturation(parser):
    text = b"bcdq" + b"\0" * 5000 + b"xyz"
    assert_many_matches(parser, text, [b"\0bcdq", b"zzzz"], "start before the text")
    assert_many_matches(parser, b"a" * (1 << 18), [b"aaaa", b"a", b"aa", b"zq001x"], "saturated")


def test_dense_candidates_few_matches(parser):
    """Every position passes the prefix filter (the filter keeps only the low 5
    bits of a key's fourth byte: 'a' and 'A' agree there) but few match, so a
    wave tile's candidates take several list chunks while its matches are
    stored."""
    text = bytearray(b"a" * 20000)
    for at in (0, 700, 2047, 2048, 5000, 19995):
        text[at:at + 5] = b"aaaAx"
    assert_many_matches(parser, bytes(text), [b"aaaAx", b"aaA", b"zzzz"], "dense candidates")


def test_positions_with_one_to_six_matches(parser):
    """Nested prefixes match 1..6 patterns at one position: up to three ids
    travel with a stored candidate, more are found again by the write pass."""
    pats = [b"ab", b"abc", b"abcd", b"abcde", b"abcdefgh", b"abcdefghij", b"bcd", b"zz"]
    text = (b"abcdefghijk " * 500 + b"abcde abcd abc ab zz\n") * 40
    assert_many_matches(parser, text, pats, "nested prefixes")


def test_duplicate_patterns_count_again(parser):
    """17,000 copies of one pattern: a position with more matches than a
    stored candidate's count field holds is counted again by the write pass."""
    text = b"q" * 3000 + b"abc" + b"q" * 100 + b"abcabc" + b"q" * 10
    assert_many_matches(parser, text, [b"abc"] * 17000 + [b"bca", b"qqqq"], "duplicates")


def test_many_tiles_per_workgroup_equal_single_pattern_searches(parser):
    """64 MiB (4,096 tiles, more than one per resident workgroup): each
    pattern's hits equal its own literal search (the reference's
    many_patterns_match_single_pattern_search, crates/oneline-simd/src/search.rs:290-312)."""
    text = datagen.document(64 << 20, seed=5, crlf=10, empty=5)
    pats = datagen.patterns(48, seed=0xC3B) + [datagen.vocab_word(i) for i in (0, 1, 2, 7, 50, 999)]
    doc = parser.upload(text, nodes=False)
    pset = parser.compile_patterns(doc, pats)
    per = [parser.search_document(doc, p, 1 << 26) for p in pats]
    want_off = np.concatenate([np.asarray(r.offsets, np.uint64) for r in per])
    want_id = np.concatenate([np.full(len(r.offsets), i, np.uint32) for i, r in enumerate(per)])
    order = np.lexsort((want_id, want_off))
    want_off, want_id = want_off[order], want_id[order]
    assert want_off.size > 100_000
    tot, off, ids = device_list(parser, doc, pset, want_off.size)
    assert tot == want_off.size
    assert np.array_equal(off, want_off)
    assert np.array_equal(ids, want_id)


def merged_entries(patterns):
    """Entries merged runs would hold for `patterns`, by the compiler's rule
    (runs merge when this is at most 2 x count + 65,536): each pattern once in
    the run of its own class and key, plus once more in the run of every
    longer class's key that starts with it."""
    from collections import Counter
    groups = Counter((min(len(p), 4), bytes(p[:4])) for p in patterns)
    total = len(patterns)
    for (klen, key) in groups:
        for shorter in range(klen - 1, 0, -1):
            total += groups.get((shorter, key[:shorter]), 0)
    return total


def test_runs_around_the_stored_bits(parser):
    """Merged runs of 63, 64 and 65 entries on one key (62..64 patterns on
    "abcd" plus the 3-byte "abc" merged in), with up to 21 matches at one
    position: a stored candidate records which of up to 64 entries match, a
    longer run is probed again by the write pass."""
    tail = b"efghijklmnopqrstuvwxyz0123456789ABCDEFGHIJKLMNOPQRSTUVWXYZ!?"
    assert len(tail) == 60
    text = b"".join(b"abcd" + tail[:k] + b"#" + b"." * (k % 5) for k in range(0, 61, 3)) * 30
    for n in (63, 64, 65):
        pats = [b"abc"] + [b"abcd" + tail[:i % 61] for i in range(n - 1)]
        assert merged_entries(pats) <= 2 * len(pats) + 65536
        assert_many_matches(parser, text, pats, f"run of {n}")


def test_per_class_runs_when_merging_would_copy_too_much(parser):
    """300 copies of b"e" under 400 longer keys that start with it: merged runs
    would copy it 120,000 times, so each run keeps its own class and the
    classes' runs are merged by id at every position."""
    rng = np.random.default_rng(5)
    words = [w for w in (datagen.vocab_word(i) for i in range(10_000)) if w[:1] == b"e" and len(w) >= 4]
    longs = list(dict.fromkeys(words))[:300]
    while len(longs) < 400:
        w = b"e" + bytes(rng.integers(97, 123, size=int(rng.integers(3, 12))).astype(np.uint8))
        if w not in longs:
            longs.append(w)
    pats = [b"e"] * 300 + longs + [b"ex", b"exa", b"q", b"ee"]
    assert merged_entries(pats) > 2 * len(pats) + 65536
    text = datagen.document(1 << 20, seed=21)
    assert_many_matches(parser, text, pats, "per-class runs")
