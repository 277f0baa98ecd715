"""The AST bridge's structural pass (ol_tree_bridge) against the oracle's
restatement of crates/oneline-tomet/src/lib.rs:20-90 (orc_bridge).

The oracle follows the reference's sequential loops: literal_pipe_lines
(lib.rs:96-125), has_pipe_child (lib.rs:33-41), blocks (lib.rs:44-60) and the
stitching walk (lib.rs:66-86) with line numbers standing in for blocks.  It is
pinned here by
  - answers worked out by hand from those lines for the reference's own test
    documents (lib.rs:318-339 and the constructs of its differential
    generator, lib.rs:353-386) and for whitespace-only content (str::trim,
    Unicode White_Space), and
  - a second, independent formulation -- per line, walk up the parent chain
    with the last bare blank line before the line -- on random documents.
The HIP kernels use the second formulation; the GPU tests compare them with
the oracle.
"""
import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

NONE = 0xFFFFFFFF
N = -1   # NONE in the tables below

# (document, flags, container): flags 1 literal | 2 block | 4 cut | 8 pipe child | 16 bare blank
KNOWN = [
    # lib.rs:318-331: the element gets `|` content
    (b'@card(title:"tomet-shelf")\n| Lightweight and fast\n', [10, 2], [N, 0]),
    # lib.rs:333-339: a list item's deeper prose is literal text (its parent is not an element)
    (b'@note\n| lead prose\n| @card(title: "kept")\n| | - nested item\n| - item\n| | trailing detail\n',
     [10, 2, 10, 2, 2, 3], [N, 0, 0, 2, 0, 4]),
    # lib.rs:385: a bare blank line ends the element's content: the next `|` line is cut off
    (b'@card(title: "c")\n| inside\n\n| after a blank line\n', [10, 2, 16, 7], [N, 0, N, N]),
    # lib.rs:377-383: `|` lines under a heading or a list item are text
    (b'= Under 3\n| pipe under a heading\n', [2, 3], [N, 0]),
    (b'- an item\n| pipe under a list item\n', [2, 3], [N, 0]),
    # lib.rs:375: nested elements; a `|` line with nothing after it is literal and has no block
    (b'@outer\n| @inner\n| | deep text\n|\n| after inner\n', [10, 10, 2, 1, 2], [N, 0, 1, N, 0]),
    # a comment has no block: its child's block goes to the comment's parent
    (b'@a\n| // note\n| | child of comment\n', [10, 0, 3], [N, N, 0]),
    # content that trims to nothing (vertical tab, no-break space) has no block
    (b'@a\n|\x0b\n', [10, 0], [N, N]),
    (b'x\n|\xc2\xa0\n', [2, 1], [N, N]),
    # U+001C is not Unicode White_Space (Rust keeps it, Python's str.strip would not)
    (b'@a\n| \x1c\n', [10, 2], [N, 0]),
]


def oracle(text):
    off, nodes = orc.parse(text)
    parent = orc.links(nodes)[0]
    return orc.bridge(text, nodes, parent)


@pytest.mark.parametrize("doc,flags,cont", KNOWN)
def test_oracle_known_answers(doc, flags, cont):
    f, c = oracle(doc)
    assert f.tolist() == flags
    assert [N if x == NONE else int(x) for x in c] == cont


# Unicode White_Space (char::is_whitespace) as code points
WHITE = [0x09, 0x0A, 0x0B, 0x0C, 0x0D, 0x20, 0x85, 0xA0, 0x1680, *range(0x2000, 0x200B), 0x2028, 0x2029,
         0x202F, 0x205F, 0x3000]


def chain_form(text, nodes, parent):
    """Per line, from the parent chain and the last bare blank line before the
    line: the formulation the kernels use, written independently of the
    oracle's sequential sweep and stitching walk."""
    text = bytes(text)
    n = len(nodes)
    depth, kind = nodes["depth"].tolist(), nodes["kind"].tolist()
    par = [int(p) for p in parent]
    bare = [kind[i] == 0 and depth[i] == 0 for i in range(n)]
    white = {chr(c) for c in WHITE}

    def has_text(i):
        s, ln = int(nodes[i]["span_start"]), int(nodes[i]["span_len"])
        return any(ch not in white for ch in text[s:s + ln].decode("utf-8", "surrogateescape"))

    block = [kind[i] not in (0, 7) and has_text(i) for i in range(n)]
    flags, cont = [], []
    last = -1
    for i in range(n):
        b = last
        lit = cut = False
        if depth[i] > 0:
            x = i
            while True:
                p = par[x]
                if p == NONE or b > p or kind[p] != 3:
                    lit = True
                    break
                if depth[p] == 0:
                    break
                x = p
            cut = par[i] != NONE and b > par[i]
        c = NONE
        if block[i]:
            x = i
            while True:
                p = par[x]
                if p == NONE or b > p:
                    break
                if block[p]:
                    c = p
                    break
                x = p
        pipe = False
        if kind[i] == 3 and not lit:
            j = i + 1
            while j < n and kind[j] == 0 and not bare[j]:
                j += 1
            pipe = j < n and kind[j] != 0 and depth[j] > depth[i]
        flags.append(int(lit) | int(block[i]) << 1 | int(cut) << 2 | int(pipe) << 3 | int(bare[i]) << 4)
        cont.append(c)
        if bare[i]:
            last = i
    return np.array(flags, np.uint8), np.array(cont, np.uint32)


LINE_BODIES = [b"@card(title: \"t\")", b"@solo(a: 1)", b"- item", b"-. step", b"= Heading", b"==[ H ]",
               b": connect", b"// comment", b"plain text", b"x", b"\x0b", b"\xc2\xa0", b"\xe3\x80\x80 ",
               b"\x0c\t", b"\xe2\x80\x89", b"\x1c", b"@", b"/", b"-", b""]


def grammar_doc(rng, lines):
    """Lines of any depth (`|` runs with blanks between), kinds from the
    grammar, whitespace-only content, bare blank lines and `|` blank lines,
    occasional CRLF."""
    out = []
    d = 0
    for _ in range(lines):
        d = max(0, min(12, d + int(rng.integers(-2, 3))))
        r = rng.random()
        if r < 0.08:
            out.append(b"")                    # bare blank line
            continue
        body = LINE_BODIES[int(rng.integers(0, len(LINE_BODIES)))]
        pre = b"".join(b"|" + (b" " if rng.random() < 0.7 else b"") for _ in range(d))
        sep = b" " if (d and body and rng.random() < 0.8) else b""
        out.append(pre + sep + body + (b"\r" if rng.random() < 0.05 else b""))
    return b"\n".join(out) + b"\n"


def test_oracle_equals_chain_form_on_random_documents():
    rng = np.random.default_rng(7)
    for round_ in range(300):
        text = grammar_doc(rng, int(rng.integers(1, 80)))
        off, nodes = orc.parse(text)
        parent = orc.links(nodes)[0]
        f, c = orc.bridge(text, nodes, parent)
        f2, c2 = chain_form(text, nodes, parent)
        assert np.array_equal(f, f2), (round_, text)
        assert np.array_equal(c, c2), (round_, text)


def test_oracle_equals_chain_form_on_generated_profiles():
    for name in datagen.PROFILES:
        text = datagen.profile(name, 96 << 10, 11)
        off, nodes = orc.parse(text)
        if nodes["depth"].max(initial=0) >= 32:
            continue
        parent = orc.links(nodes)[0]
        f, c = orc.bridge(text, nodes, parent)
        f2, c2 = chain_form(text, nodes, parent)
        assert np.array_equal(f, f2), name
        assert np.array_equal(c, c2), name


# ---- the HIP path -------------------------------------------------------------------


def gpu_bridge(parser, text, lean):
    doc = parser.upload(text, nodes=not lean)
    tree = parser.build_tree(doc)
    return parser.bridge(doc, tree)


def assert_bridge(parser, text, what):
    want_f, want_c = oracle(text)
    for lean in (False, True):
        f, c = gpu_bridge(parser, text, lean)
        assert np.array_equal(f, want_f), f"{what} (lean={lean}): flags"
        assert np.array_equal(c, want_c), f"{what} (lean={lean}): container"


@pytest.mark.gpu
def test_gpu_known_answers(parser):
    for doc, flags, cont in KNOWN:
        assert_bridge(parser, doc, repr(doc))


@pytest.mark.gpu
def test_gpu_random_grammar_documents(parser):
    rng = np.random.default_rng(8)
    for k in range(40):
        assert_bridge(parser, grammar_doc(rng, int(rng.integers(1, 3000))), f"grammar {k}")


@pytest.mark.gpu
def test_gpu_blocks_and_chains_across_block_edges(parser):
    """Long chains and runs of blank lines that cross the 4,096-line blocks
    and 16-line thread boundaries: a bare blank far before a deep line, an
    element whose first child follows many `|` blank lines."""
    lines = [b"@root"]
    for d in range(1, 20):
        lines.append(b"| " * d + b"@e")
    lines += [b"|"] * 5000 + [b"| | child after pipe blanks"]
    lines += [b"@other"] + [b"| x"] * 4100 + [b""] + [b"| after bare"] * 3
    lines += [b"@el"] + [b"|"] * 4095 + [b"", b"| cut"]
    assert_bridge(parser, b"\n".join(lines) + b"\n", "block edges")


def long_prefix_doc():
    """Text lines whose content starts past the first 32-byte window (and the
    second, third and fourth): deep `| ` prefixes, blank runs inside the
    prefix, blanks-only prefixes at depth 0, and content that starts with or
    is only Unicode whitespace."""
    contents = [b"text", " x".encode(), "　".encode(), "  ".encode(), b"\x0bq", b"x"]
    lines = [b"@root"]
    for d in range(1, 32):
        for k, c in enumerate(contents):
            pre = b"| " * d
            if k % 3 == 1:
                pre = b"|\t" + b"| " * (d - 1)
            lines.append(pre + c)
    for blanks in (30, 60, 93, 126, 127, 200):
        lines.append(b"|" + b" " * blanks + b"deep blank run")
        lines.append(b"|" + b" " * blanks + "　".encode())
        lines.append(b" " * blanks + b"depth-0 blanks")
    lines.append(b"@after")
    return b"\n".join(lines) + b"\n"


def test_oracle_equals_chain_form_on_long_prefixes():
    text = long_prefix_doc()
    off, nodes = orc.parse(text)
    parent = orc.links(nodes)[0]
    f, c = orc.bridge(text, nodes, parent)
    f2, c2 = chain_form(text, nodes, parent)
    assert np.array_equal(f, f2) and np.array_equal(c, c2)
    # both block and no-block Text lines past the first window
    assert (f & ol.BRIDGE_BLOCK).any() and not (f & ol.BRIDGE_BLOCK).all()


@pytest.mark.gpu
def test_gpu_long_prefixes(parser):
    assert_bridge(parser, long_prefix_doc(), "long prefixes")


@pytest.mark.gpu
def test_gpu_generated_profiles(parser):
    for name in datagen.PROFILES:
        text = datagen.profile(name, 3 << 20, 3)
        off, nodes = orc.parse(text)
        if nodes["depth"].max(initial=0) >= 32:
            continue   # no tree (MAX_TREE_DEPTH)
        assert_bridge(parser, text, name)


@pytest.mark.gpu
def test_gpu_benchmark_document_64mib(parser):
    text = datagen.document(64 << 20, seed=4, crlf=10, empty=5)
    assert_bridge(parser, text, "64 MiB")


@pytest.mark.gpu
def test_gpu_empty_and_errors(parser):
    doc = parser.upload(b"")
    tree = parser.build_tree(doc)
    f, c = parser.bridge(doc, tree)
    assert f.size == 0 and c.size == 0
    a = parser.upload(b"@a\n| b\n")
    b = parser.upload(b"x\n")
    with pytest.raises(ol.OnelineError):
        parser.bridge(b, parser.build_tree(a))   # the tree belongs to another document
