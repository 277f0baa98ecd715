"""Single-literal search on the HIP path against the oracle
(crates/oneline-gpu/src/tests.rs:280-448)."""
import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def assert_search_matches(parser, text, pattern, doc=None):
    text = ol.as_bytes_array(text)
    pattern = bytes(pattern)
    want = orc.find_all(text, pattern).astype(np.uint32)
    cap = max(1, want.size)
    doc = doc or parser.upload(text, nodes=False)
    r = parser.search_document(doc, pattern, cap)
    assert r.total_matches == want.size, f"count for {pattern!r}"
    assert np.array_equal(r.offsets, want), f"offsets for {pattern!r}"
    one = parser.search_raw(text, pattern, cap)
    assert np.array_equal(one.offsets, want), f"one-shot offsets for {pattern!r}"


def test_known_answers(parser, known):
    for case in known["find_all"]:
        if "offsets" in case and case["text"]:
            r = parser.search_raw(case["text"], case["pattern"].encode(), 16)
            assert r.offsets.tolist() == case["offsets"], case["src"]
            assert r.total_matches == len(case["offsets"])


def test_corpus(parser, known):
    for text, pat in known["search_corpus"]["cases"]:
        assert_search_matches(parser, text.encode(), pat.encode())


def test_generated_documents(parser):
    text = datagen.document(5 << 20, seed=17, crlf=10, empty=5)
    doc = parser.upload(text, nodes=False)
    pats = [b"|", b"| |", b"\n|", b"- ", b"-. ", b"//", b"zzzz-absent", b"\r\n"]
    pats += [datagen.vocab_word(k) for k in (0, 1, 2, 50, 999, 9999)]
    for i, ln in enumerate((1, 2, 3, 4, 5, 8, 17, 33, 63, 64)):
        at = 1000 + i * 99_991
        pats.append(bytes(text[at:at + ln]))
    for p in pats:
        assert_search_matches(parser, text, p, doc)


def test_truncation_reports_true_total(parser):
    text = datagen.document(2 << 20, seed=2)
    want = orc.find_all(text, b"|")
    assert want.size > 10_000
    doc = parser.upload(text, nodes=False)
    r = parser.search_document(doc, b"|", 100)
    assert r.total_matches == want.size
    assert r.offsets.size == 100 and np.isin(r.offsets, want).all()
    none = parser.search_document(doc, b"|", 0)
    assert none.total_matches == want.size and none.offsets.size == 0


def test_bad_patterns_and_recovery(parser, known):
    c = known["search_rejects"]
    doc = parser.upload(c["text"])
    for ln in c["bad_lengths"]:
        with pytest.raises(ol.OnelineError):
            parser.search_document(doc, b"a" * ln, 8)
    parser.search_document(doc, b"a" * c["ok_length"], 8)
    with pytest.raises(ol.OnelineError):
        parser.search_raw(c["text"], b"", 8)
    assert parser.search_document(doc, c["after_error"]["pattern"].encode(), 8).offsets.tolist() == \
        c["after_error"]["offsets"]


def test_empty_document(parser):
    assert parser.search_document(parser.upload(b""), b"a", 8).total_matches == 0
    assert parser.search_raw(b"", b"a", 8).total_matches == 0


def test_anchor_positions_and_nul(parser):
    text = b"abcq" + b"abc" * 40_000 + b"abcq"
    for p in (b"abcq", b"qabc", b"cabcqabc", b"abcz"):
        assert_search_matches(parser, text, p)
    assert_search_matches(parser, b"qxyz", b"abcdq")
    text = b"bcdq" + b"\0" * 5000 + b"xyz"
    assert_search_matches(parser, text, b"\0bcdq")
    assert_search_matches(parser, b"xa", b"a\0")


def test_irregular_matches_in_consecutive_tiles(parser):
    """Matches in every tile of an 8 MiB generated document, unevenly spread
    over the lanes and vectors: each tile's matches are written only after
    the next tile is counted, so the offsets of one tile must not leak into
    the next (a periodic text would hide that)."""
    text = datagen.document(8 << 20, seed=21, crlf=10, empty=5)
    doc = parser.upload(text, nodes=False)
    pats = [datagen.vocab_word(i) for i in (0, 1, 3, 20, 200)] + datagen.patterns(12, seed=0xC3B)
    for pat in pats:
        assert_search_matches(parser, text, pat, doc=doc)


def test_rare_window_inside_the_pattern(parser):
    """The kernel tests the pattern's rarest window (here at offsets 2 and
    60); matches start before it, at the very start and end of the text."""
    body = b"ab" * 60_000
    text = b"abxyzQ" + body + b"abxyzQ" + body + b"abxyzQ"
    for p in (b"abxyzQ", b"bxyzQ", b"babxyzQ", b"xyzQab"):
        assert_search_matches(parser, text, p)
    long = (b"ab" * 30) + b"xyzQ"   # 64 bytes, the rare bytes last
    text = long + body + long + b"b" + long
    assert_search_matches(parser, text, long)
    assert_search_matches(parser, text[:-1], long)   # the last one cut short


def test_matches_across_tile_edges(parser):
    rec = b"%-46sMARKER%-47s\n" % (b"header", b"x")
    assert len(rec) == 100
    text = rec * 90_000 + b"end"
    want = np.arange(90_000, dtype=np.uint64) * 100 + 46
    r = parser.search_raw(text, b"MARKER", 90_000)
    assert np.array_equal(r.offsets, want.astype(np.uint32))
    assert parser.search_raw(text, b"end", 4).offsets.tolist() == [9_000_000]
    for pat in (b"ab", b"aba", b"bab", b"ababab", b"x"):
        assert_search_matches(parser, (b"ab" * 70_000) + b"a", pat)
