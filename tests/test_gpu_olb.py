"""Parsed documents saved to and loaded from .olb files: the loaded document
gives the same lines, classes, FlatNodes, tree and queries as the original,
and damaged files are rejected."""
import os
import struct
import zlib

import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nodes_kept", [True, False])
def test_round_trip(parser, tmp_path, nodes_kept):
    text = datagen.document(3 << 20, seed=71, crlf=10, empty=20).tobytes() + b"| | tail without newline"
    doc = parser.upload(text, nodes=nodes_kept)
    path = str(tmp_path / "doc.olb")
    parser.save(doc, path)
    with open(path, "rb") as f:
        head = f.read(64)
    magic, ver, flags, n, lines, toff, noff, ooff, crc = struct.unpack("<8sIIQQQQQQ", head)
    assert magic == b"ONELINE\0" and ver == 1 and flags & 1
    assert n == len(text) and lines == doc.line_count() and toff == 64 and noff % 64 == 0 and ooff % 64 == 0
    assert os.path.getsize(path) == ooff + 4 * lines
    with open(path, "rb") as f:
        assert zlib.crc32(f.read()[64:]) == crc   # CRC-32 (IEEE) of everything after the header
    back = parser.load(path)
    assert back.line_count() == doc.line_count() and back.text_len() == len(text)
    off, nodes = orc.parse(np.frombuffer(text, np.uint8))
    o, c, nd = back.download(nodes=True)
    assert np.array_equal(o, off) and np.array_equal(c, nodes["kind"].astype(np.uint8))
    assert np.array_equal(nd["depth"], nodes["depth"]) and np.array_equal(nd["kind"], nodes["kind"])
    assert np.array_equal(nd["span_start"], nodes["span_start"]) and np.array_equal(nd["span_len"], nodes["span_len"])
    t = parser.download_links(parser.build_tree(back))
    for g, w in zip((t.parent, t.subtree_end, t.first_child, t.next_sibling), orc.links(nodes)):
        assert np.array_equal(g, w)
    q1, q2 = parser.query_document(doc, 4, 3), parser.query_document(back, 4, 3)
    assert q1.total_matches == q2.total_matches and np.array_equal(q1.indices, q2.indices)
    s1, s2 = parser.search_document(doc, b"ab", 1 << 20), parser.search_document(back, b"ab", 1 << 20)
    assert s1.total_matches == s2.total_matches and np.array_equal(s1.offsets, s2.offsets)


def test_empty_document(parser, tmp_path):
    path = str(tmp_path / "empty.olb")
    parser.save(parser.upload(b""), path)
    back = parser.load(path)
    assert back.line_count() == 0 and back.text_len() == 0


def test_damaged_files_are_rejected(parser, tmp_path):
    text = datagen.document(1 << 16, seed=72).tobytes()
    path = str(tmp_path / "d.olb")
    parser.save(parser.upload(text), path)
    raw = bytearray(open(path, "rb").read())
    flipped = bytearray(raw)
    flipped[100] ^= 1
    p2 = str(tmp_path / "flip.olb")
    open(p2, "wb").write(flipped)
    with pytest.raises(ol.OnelineError):
        parser.load(p2)
    parser.load(p2, verify=False)   # the checksum is the only thing wrong
    for bad in (b"XNELINE\0" + raw[8:], raw[:8] + struct.pack("<I", 2) + raw[12:], raw[:40]):
        p3 = str(tmp_path / "bad.olb")
        open(p3, "wb").write(bytes(bad))
        with pytest.raises(ol.OnelineError):
            parser.load(p3)
    with pytest.raises(ol.OnelineError):
        parser.load(str(tmp_path / "missing.olb"))
