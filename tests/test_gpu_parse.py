"""Scan + classify on the HIP path, bit-exact against the oracle
(the reference's own GPU tests: crates/oneline-gpu/src/tests.rs:27-279)."""
import ctypes

import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def assert_parse_matches(parser, text, what=""):
    text = ol.as_bytes_array(text)
    off, nodes = orc.parse(text)
    tree = parser.parse_raw(text)
    assert tree.line_offsets.size == off.size, f"{what}: line count"
    assert np.array_equal(tree.line_offsets, off), f"{what}: offsets"
    if not np.array_equal(tree.nodes, nodes):
        i = int(np.nonzero(tree.nodes != nodes)[0][0])
        raise AssertionError(f"{what}: node {i} gpu {tree.nodes[i]} cpu {nodes[i]}")


def test_known_answers(parser, known):
    for case in known["scan_line_offsets"]:
        assert parser.scan_line_offsets(case["text"]).tolist() == case["offsets"], case["src"]
    for case in known["parse_document"]:
        tree = parser.parse_raw(case["text"])
        got = [[int(n["depth"]), int(n["kind"])] for n in tree.nodes]
        assert got == case["depth_kind"], case["src"]
        offs = orc.scan_line_offsets(case["text"].encode())
        given = parser.parse(case["text"], offs)
        assert [[int(n["depth"]), int(n["kind"])] for n in given.nodes] == case["depth_kind"]


def test_edge_corpus(parser, known):
    for case in known["parse_edge_corpus"]["cases"] + known["scan_corpus"]["cases"]:
        assert_parse_matches(parser, case, repr(case))
        assert parser.scan_line_offsets(case).tolist() == orc.scan_line_offsets(case.encode()).tolist()


def test_parse_with_given_offsets(parser, known):
    for case in known["parse_edge_corpus"]["cases"]:
        off, nodes = orc.parse(case.encode())
        got = parser.parse(case, off)
        assert np.array_equal(got.nodes, nodes), repr(case)


def test_awkward_shapes(parser):
    cases = {
        "long lines": b"a" * 100_000 + b"\n|" + b"b" * 300_000,
        "one giant prefix": b"|" * 1_000_000,
        "crlf only": b"x\r\n" * 300_000,
        "all newlines": b"\n" * 400_000,
        "tail without newline": b"line\n" * 200_000 + b"tail without newline",
        "cr runs": b"a\r\r\r\n" * 5000 + b"\r\r",
        "long blank prefix": (b"| " * 2000 + b"x\n") * 50,
    }
    for name, text in cases.items():
        assert_parse_matches(parser, text, name)


def test_tile_boundaries(parser):
    base = datagen.document(1 << 20, seed=11)
    for n in (1, 15, 16, 17, 255, 256, 257, 16383, 16384, 16385, 32768, 65535, 65536, 65537, 200_000):
        t = base[:n].copy()
        assert_parse_matches(parser, t, f"n={n}")
        t[-1] = ord("x")   # unterminated last line
        assert_parse_matches(parser, t, f"n={n} unterminated")


def test_newline_at_every_position_near_a_tile_edge(parser):
    for p in range(16384 - 70, 16384 + 70, 7):
        t = np.full(40_000, ord("a"), np.uint8)
        t[p] = 0x0A
        t[p + 1:p + 5] = np.frombuffer(b"|| -", np.uint8)
        assert_parse_matches(parser, t, f"newline at {p}")


def test_generated_documents(parser):
    text = datagen.document(1 << 20, seed=0)
    assert_parse_matches(parser, text, "1 MiB seed 0")
    text = datagen.document(37 << 20, seed=3, crlf=10, empty=5)
    assert_parse_matches(parser, text, "37 MiB crlf/empty")


def test_resident_document_arrays(parser):
    text = datagen.document(9 << 20, seed=5, crlf=10, empty=5)
    off, nodes = orc.parse(text)
    doc = parser.upload(text)
    assert doc.line_count() == off.size and doc.text_len() == text.size
    o, c, nd = doc.download(nodes=True)
    assert np.array_equal(o, off)
    assert np.array_equal(c, nodes["kind"].astype(np.uint8))
    assert np.array_equal(nd, nodes)
    lean = parser.upload(text, nodes=False)
    o2, c2, _ = lean.download()
    assert np.array_equal(o2, off) and np.array_equal(c2, c)


def test_staged_upload_sizes(parser):
    """Uploads go through a pinned staging ring: one slot in order below
    32 MiB, parallel 8 MiB slots above.  Sizes around that threshold and
    around slot edges, with a distinct byte at each slot edge, arrive intact
    (the scan's offsets and classes equal the oracle's)."""
    slot, small = 8 << 20, 32 << 20
    base = datagen.document(small + slot + 64, seed=9, crlf=10, empty=5)
    for n in (small - 1, small, small + 1, small + slot + 3):
        text = base[:n].copy()
        for edge in range(slot, n, slot):
            text[edge - 1], text[edge] = ord("\n"), ord("=")
        off, nodes = orc.parse(text)
        doc = parser.upload(text, nodes=False)
        o, c, _ = doc.download()
        assert np.array_equal(o, off), f"{n} bytes: offsets"
        assert np.array_equal(c, nodes["kind"].astype(np.uint8)), f"{n} bytes: classes"


def test_empty_input(parser):
    assert parser.scan_line_offsets(b"").size == 0
    assert parser.parse_raw(b"").nodes.size == 0
    doc = parser.upload(b"")
    assert doc.line_count() == 0 and doc.text_len() == 0


def test_oversized_input_is_an_error(parser):
    lib = ol.load_library()
    buf = np.zeros(16, np.uint8)
    h = ctypes.c_void_p()
    rc = lib.ol_doc_upload(parser._ctx, ctypes.c_void_p(buf.ctypes.data), ol.MAX_GPU_INPUT_LEN + 1, ctypes.byref(h))
    assert rc == -2
    assert "bytes" in lib.ol_last_error().decode()
    # still usable
    assert parser.scan_line_offsets(b"a\nb").tolist() == [0, 2]


def test_random_lines_over_the_grammar_alphabet(parser):
    """Documents of random short lines over the reference's random-line alphabet
    (crates/oneline-simd/src/line.rs:148), so prefixes, blank runs, markers and
    CR runs land at every offset relative to lanes, masks and tiles."""
    rng = np.random.default_rng(4242)
    alphabet = np.frombuffer(b"|| \t=:@->*`0123456789.\rxy", np.uint8)
    for n_lines in (5_000, 60_000):
        lens = rng.integers(0, 14, n_lines)
        parts = []
        for ln in lens:
            parts.append(alphabet[rng.integers(0, alphabet.size, ln)].tobytes() + b"\n")
        assert_parse_matches(parser, b"".join(parts), f"{n_lines} random lines")


def test_prefixes_around_the_mask_window(parser):
    """Prefixes whose '|' / blank run ends just before, at and after 64 bytes."""
    lines = []
    for plen in range(50, 80):
        for unit in (b"|", b"| ", b"|\t", b"|  "):
            pre = (unit * plen)[:plen]
            for tail in (b"x", b"- y", b"", b"  ", b"-. z", b"//c"):
                lines.append(pre + tail + b"\n")
                lines.append(b" " + pre + tail + b"\r\n")
                lines.append(b"\t" * (plen % 7) + tail + b"\n")
    text = b"".join(lines)
    assert_parse_matches(parser, text, "prefix lengths 50..79")
    for shift in (1, 7, 33, 63):
        assert_parse_matches(parser, b"q" * shift + b"\n" + text, f"shifted {shift}")


def test_newline_bytes_at_every_vector_lane_and_neighbours(parser):
    """'\\n' at each byte of a 16-byte vector, next to bytes that differ from
    it by one bit (0x0B, 0x0E, 0x8A, 0x2A) -- the exact zero-byte test."""
    rng = np.random.default_rng(8)
    near = np.array([0x0B, 0x0E, 0x08, 0x8A, 0x2A, 0x09, 0x0A], np.uint8)
    text = rng.choice(near, size=3 << 20).astype(np.uint8)
    assert_parse_matches(parser, text, "near-newline bytes")


def test_ticket_ordered_scan(parser):
    """Scan tiles claimed from a ticket (ordinary launch) give the same results."""
    parser.tune_set("scan_ticket", 1)
    try:
        for bpt in (64, 128, 256):
            parser.tune_set("scan_bpt", bpt)
            assert_parse_matches(parser, datagen.document(24 << 20, seed=21, crlf=10, empty=5), f"ticket bpt={bpt}")
    finally:
        parser.tune_set("scan_ticket", 0)
        parser.tune_set("scan_bpt", 128)


def test_scan_floor_calibration(parser):
    """ol_calib_scan_floor times a kernel over the document without changing it."""
    text = datagen.document(8 << 20, seed=4)
    doc = parser.upload(text, nodes=False)
    before = doc.download()
    r = parser.calibrate_scan_floor(doc, 500, write=False, reps=3)
    rw = parser.calibrate_scan_floor(doc, 500, write=True, reps=3)
    assert r > 0 and rw > 0
    after = doc.download()
    assert all(np.array_equal(x, y) for x, y in zip(before[:2], after[:2]))
    with pytest.raises(ol.OnelineError):
        parser.calibrate_scan_floor(doc, 40000)


def test_adapter_name_names_the_device(parser):
    """GpuParser::adapter_name (crates/oneline-gpu/src/parser.rs:470): a
    device name (the runtime's, or a generic one when it reports none), the
    architecture, compute units and HBM size."""
    name = parser.adapter_name()
    assert name and not name.startswith(" ") and "gfx950" in name and " CUs, " in name, name
