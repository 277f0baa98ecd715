"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip in the restatement (a dropped
trailing-CR trim, blank runs counted into the prefix when no '|' follows, an
off-by-one in the line split, '<' for '<=' in the tree, a missed overlapping
match, a duplicate pattern reported once) fails at least one of them:

* the reference's inline known answers (tests/golden/known_answers.json);
* an independently written byte-loop parser (after the reference's test oracle,
  crates/oneline-simd/src/line.rs:89-142), exhaustively over the 8-symbol
  alphabet up to length 6 and on random lines (line.rs:144-180);
* brute-force definitions of the tree relations on random small inputs
  (crates/oneline-core/src/links.rs:8-21 for parent / subtree_end; the
  first_child / next_sibling closed forms against "first line whose parent is
  i" / "next line with the same parent");
* naive search and brute-force dictionary search.
"""
import itertools
import random

import numpy as np
import pytest

from oracle import pyoracle as orc

NONE = orc.NONE
BLANK, TEXT, HEADING, ELEMENT, LIST, CONNECT, ORDERED, COMMENT = range(8)


# ---- independent byte-loop parser ------------------------------------------------
def ref_parse_line(b: bytes, start: int):
    """Written the other way round from the oracle: one forward scan remembering
    where the last '|' was (after crates/oneline-simd/src/line.rs:89-142)."""
    n = len(b)
    while n and b[n - 1] in (0x0D, 0x0A):
        n -= 1
    if n == 0:
        return (0, BLANK, start, 0)
    depth = 0
    prefix_end = 0
    i = 0
    while i < n:
        c = b[i]
        if c == 0x7C:
            depth += 1
            i += 1
            prefix_end = i
        elif c in (0x20, 0x09) and depth > 0:
            i += 1
        else:
            break
    cs = prefix_end
    while cs < n and b[cs] in (0x20, 0x09):
        cs += 1
    if cs == n:
        return (depth, BLANK, start + cs, 0)
    rest = b[cs:n]
    c0 = rest[0]
    if c0 == ord("="):
        kind = HEADING
    elif c0 == ord(":"):
        kind = CONNECT
    elif c0 == ord("@"):
        kind = ELEMENT
    elif c0 == ord("-") and len(rest) > 1 and rest[1] in (0x20, 0x09):
        kind = LIST
    elif c0 == ord("-") and len(rest) > 2 and rest[1] == ord(".") and rest[2] in (0x20, 0x09):
        kind = ORDERED
    elif c0 == ord("/") and len(rest) > 1 and rest[1] == ord("/"):
        kind = COMMENT
    else:
        kind = TEXT
    return (depth, kind, start + cs, n - cs)


def ref_split(text: bytes):
    """Split-on-'\\n' reference (crates/oneline-simd/src/lib.rs:493-513)."""
    starts = []
    if text:
        starts.append(0)
        for i, c in enumerate(text):
            if c == 0x0A and i + 1 < len(text):
                starts.append(i + 1)
    nodes = []
    for k, s in enumerate(starts):
        e = starts[k + 1] if k + 1 < len(starts) else len(text)
        nodes.append(ref_parse_line(text[s:e], s))
    return starts, nodes


def test_parse_line_known_answers(known):
    for case in known["parse_line"]:
        d, k, _, _ = orc.parse_line(case["line"].encode())
        assert k == case["kind"], case
        if "depth" in case:
            assert d == case["depth"], case


def test_scan_known_answers(known):
    for case in known["scan_line_offsets"]:
        assert orc.scan_line_offsets(case["text"].encode()).tolist() == case["offsets"], case["src"]


def test_parse_document_known_answers(known):
    for case in known["parse_document"]:
        _, nodes = orc.parse(case["text"].encode())
        got = [[int(n["depth"]), int(n["kind"])] for n in nodes]
        assert got == case["depth_kind"], case["src"]


def test_parse_line_exhaustive_short_lines():
    # Exhaustive over the reference's 8-symbol alphabet up to length 6 (line.rs:164-180).
    alphabet = b"| -.:@/\n"
    for length in range(7):
        for combo in itertools.product(alphabet, repeat=length):
            line = bytes(combo)
            assert orc.parse_line(line, 7) == ref_parse_line(line, 7), line


def test_parse_line_random_lines():
    # The reference's random-line alphabet (line.rs:148): prefixes, blanks, markers, CR/LF runs.
    alphabet = b"|| \t=:@->*`0123456789.\r\nxy"
    rng = random.Random(4242)
    for _ in range(60_000):
        n = rng.randrange(14)
        line = bytes(rng.choice(alphabet) for _ in range(n))
        start = rng.randrange(1000)
        assert orc.parse_line(line, start) == ref_parse_line(line, start), line


def test_parse_matches_split_reference(known):
    cases = [c.encode() for c in known["parse_edge_corpus"]["cases"]]
    cases += ["a" * 100_000 + "\n|" + "b" * 300_000, "|" * 100_000, "x\r\n" * 3000, "\n" * 4000,
              "line\n" * 2000 + "tail without newline"]
    rng = random.Random(5)
    for _ in range(200):
        n = rng.randrange(200)
        cases.append(bytes(rng.choice(b"|| \t=:@-./\r\nab") for _ in range(n)))
    for text in cases:
        text = text.encode() if isinstance(text, str) else text
        starts, nodes = ref_split(text)
        off, got = orc.parse(text)
        assert off.tolist() == starts
        assert [tuple(int(x) for x in g) for g in got] == nodes
        assert orc.scan_line_offsets(text).tolist() == starts


# ---- tree -------------------------------------------------------------------------
def brute_links(depths, blank):
    n = len(depths)
    parent = [NONE] * n
    se = [i + 1 for i in range(n)]
    for i in range(n):
        if blank[i]:
            continue
        for j in range(i - 1, -1, -1):
            if not blank[j] and depths[j] < depths[i]:
                parent[i] = j
                break
        se[i] = n
        for j in range(i + 1, n):
            if not blank[j] and depths[j] <= depths[i]:
                se[i] = j
                break
    fc = [NONE] * n
    ns = [NONE] * n
    for i in range(n):
        if blank[i]:
            continue
        kids = [j for j in range(n) if not blank[j] and parent[j] == i]
        fc[i] = kids[0] if kids else NONE
        sib = [k for k in range(i + 1, n) if not blank[k] and parent[k] == parent[i]]
        ns[i] = sib[0] if sib else NONE
    return parent, se, fc, ns


def test_tree_known_answer(known):
    for case in known["tree_depths"]:
        depths = np.array([d for d, _ in case["lines"]], np.uint32)
        kinds = np.array([BLANK if b else TEXT for _, b in case["lines"]], np.uint32)
        p, se, fc, ns = orc.links_from(depths, kinds)
        assert p.tolist() == case["parent"], case["src"]
        assert se.tolist() == case["subtree_end"], case["src"]
    for case in known["tree_text"]:
        _, nodes = orc.parse(case["text"].encode())
        p, se, _, _ = orc.links(nodes)
        for line, lo, hi in case["subtree"]:
            assert (line, int(se[line])) == (lo, hi)
        for line, par in case["parent_of"]:
            assert int(p[line]) == par


@pytest.mark.parametrize("seed", range(12))
def test_tree_against_brute_force(seed):
    rng = random.Random(seed)
    for _ in range(60):
        n = rng.randrange(1, 60)
        maxd = rng.choice([1, 2, 3, 5, 8])
        depths = [rng.randrange(maxd) for _ in range(n)]
        blank = [rng.random() < 0.15 for _ in range(n)]
        kinds = np.array([BLANK if b else TEXT for b in blank], np.uint32)
        got = orc.links_from(np.array(depths, np.uint32), kinds)
        want = brute_links(depths, blank)
        for g, w, name in zip(got, want, ("parent", "subtree_end", "first_child", "next_sibling")):
            assert g.tolist() == w, (name, depths, blank)


# ---- search -------------------------------------------------------------------------
def naive(text: bytes, pat: bytes):
    if not pat or len(text) < len(pat):
        return []
    return [i for i in range(len(text) - len(pat) + 1) if text[i:i + len(pat)] == pat]


def test_find_all_known_answers(known):
    for case in known["find_all"]:
        got = orc.find_all(case["text"].encode(), case["pattern"].encode())
        if "offsets" in case:
            assert got.tolist() == case["offsets"], case["src"]
        if "count" in case:
            assert len(got) == case["count"], case["src"]


def test_find_all_against_naive():
    rng = random.Random(9)
    for _ in range(300):
        text = bytes(rng.choice(b"ab\0\n") for _ in range(rng.randrange(80)))
        pat = bytes(rng.choice(b"ab\0\n") for _ in range(rng.randrange(1, 6)))
        assert orc.find_all(text, pat).tolist() == naive(text, pat)


def brute_many(text: bytes, pats):
    out = []
    for p in range(len(text)):
        for i, q in enumerate(pats):
            if text[p:p + len(q)] == q and len(q) > 0 and p + len(q) <= len(text):
                out.append((p, i))
    return out


def test_many_against_brute_force(known):
    c = known["many_corpus"]
    text = c["text"].encode()
    pats = [p.encode() for p in c["patterns"]]
    oo, ii = orc.many_find_all(text, pats)
    assert list(zip(oo.tolist(), ii.tolist())) == brute_many(text, pats)
    # duplicates each report their own id
    dup = [i for i, p in enumerate(pats) if p == b"abcdef"]
    assert len(dup) == 2 and all((ii == d).any() for d in dup)
    rng = random.Random(3)
    for _ in range(150):
        text = bytes(rng.choice(b"abc\n") for _ in range(rng.randrange(120)))
        pats = [bytes(rng.choice(b"abc\n") for _ in range(rng.randrange(1, 7))) for _ in range(rng.randrange(1, 12))]
        oo, ii = orc.many_find_all(text, pats)
        assert list(zip(oo.tolist(), ii.tolist())) == brute_many(text, pats)


def test_many_each_pattern_equals_single_search():
    # The reference pins its dictionary search this way (crates/oneline-simd/src/search.rs:290-312).
    from oneline_amd import datagen
    text = bytes(datagen.document(1 << 18, seed=8))
    pats = [datagen.vocab_word(k) for k in (0, 1, 5, 200, 3500, 4000)] + [b"|", b"- "]
    oo, ii = orc.many_find_all(text, pats)
    for i, p in enumerate(pats):
        assert oo[ii == i].tolist() == orc.find_all(text, p).tolist()


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_parallel_restatement_equals_sequential(threads, known):
    """The multi-threaded CPU baseline (chunked parse, three-phase tree) gives
    the sequential oracle's outputs, on edge texts, generated documents split
    into several blocks, and random texts over the markup alphabet."""
    import numpy as np
    from oneline_amd import datagen
    cases = [b"", b"x", b"\n", b"| | a", b"a\nb\n", b"| x\n|| y\n||| z\n| w", b"\n\n\n", b"\r\n\r", b"|" * 40 + b"\n"]
    cases.append(known["tree_doc"].encode() if "tree_doc" in known else b"|=[ A ]\n|| x\n||- y\n|||- z\n|=[ B ]\n|| w\n")
    for i, size in enumerate((5000, 70000, 3 << 20)):
        cases.append(datagen.document(size, seed=40 + i, crlf=10, empty=60).tobytes())
    rng = random.Random(threads)
    for _ in range(6):
        cases.append(bytes(rng.choice(b"| \t\n\r-=.") for _ in range(rng.randrange(1, 200000))))
    # deep and flat runs across block boundaries (blocks are 32768 lines)
    cases.append(b"".join(b"|" * (i % 40) + b" x\n" for i in range(70000)))
    cases.append(b"x\n" * 70000 + b"| y\n" * 3)
    for c in cases:
        o1, n1 = orc.parse(c)
        o2, n2 = orc.parse_par(c, threads)
        assert np.array_equal(o1, o2) and np.array_equal(n1, n2), len(c)
        if n1.size:
            for a, b in zip(orc.links(n1), orc.links_par(n1, threads)):
                assert np.array_equal(a, b), len(c)


def test_find_all_par_equals_find_all_across_chunks():
    """The chunked literal search (find_all_par's 1 MiB chunks with a halo)
    finds what the sequential one finds, also for matches that straddle a
    chunk edge."""
    from oneline_amd import datagen
    t = datagen.document(5 << 20, seed=7).copy()
    pat = b"zqzqzq"
    for edge in (1 << 20, 2 << 20, 3 << 20):
        for k in (-5, -3, -1, 0):
            t[edge + k:edge + k + len(pat)] = np.frombuffer(pat, np.uint8)
    for p in (pat, b"zq", b"the", datagen.vocab_word(1), b"\n"):
        want = orc.find_all(t, p)
        for threads in (1, 3, 8):
            assert np.array_equal(orc.find_all_par(t, p, threads), want), (p, threads)


def test_aho_corasick_equals_closed_form():
    """The overlapping Aho-Corasick baseline reports every (start, id) of the
    closed form, duplicates each with its id, across chunk edges."""
    from oneline_amd import datagen
    t = datagen.document(3 << 20, seed=8).copy()
    pats = datagen.patterns(512, seed=9) + [b"a", b"ab", b"abc", b"abcd", b"abcd", b"bcd", b"\n|", b"zz" * 20]
    t[(1 << 20) - 3:(1 << 20) + 37] = np.frombuffer(b"zz" * 20, np.uint8)
    ac = orc.AhoCorasick(pats)
    wo, wi = orc.many_find_all(t, pats)
    for threads in (1, 4):
        o, i = ac.find(t, threads)
        assert np.array_equal(o, wo) and np.array_equal(i, wi), threads
        assert ac.count(t, threads) == wo.size
