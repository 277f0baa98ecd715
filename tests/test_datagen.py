"""The synthetic workload generator: exact sizes, determinism, the shape the
benchmark configurations state, and a golden hash of the 1 MiB seed-0 document."""
import hashlib
import json
import os

import numpy as np

from oneline_amd import datagen
from oracle import pyoracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_exact_size_and_terminated():
    for n in (1, 2, 3, 100, 4095, 1 << 20, (1 << 20) + 7, 3 << 20):
        d = datagen.document(n, seed=3)
        assert d.size == n and d[-1] == 0x0A


def test_deterministic_and_thread_independent():
    a = datagen.document(5 << 20, seed=7, crlf=10, empty=5, threads=1)
    b = datagen.document(5 << 20, seed=7, crlf=10, empty=5, threads=8)
    assert np.array_equal(a, b)
    c = datagen.document(5 << 20, seed=8, crlf=10, empty=5, threads=8)
    assert not np.array_equal(a, c)


def test_line_shape_of_the_1mib_document():
    text = datagen.document(1 << 20, seed=0)
    off, nodes = orc.parse(text)
    assert 15_000 < len(off) < 17_500
    lengths = np.diff(np.append(off, text.size)) - 1   # without '\n'
    assert lengths[:-1].min() >= 1 and lengths[:-1].max() <= 120 + 2 * 16
    assert nodes["depth"].max() <= 16
    kinds = np.bincount(nodes["kind"], minlength=8)
    assert kinds[0] < 5                    # no blank lines but a rare block-end filler
    assert all(kinds[k] > 1500 for k in range(1, 8))
    assert not (text == 0x0D).any()


def test_crlf_and_empty_rates():
    text = datagen.document(16 << 20, seed=0, crlf=10, empty=5)
    off, nodes = orc.parse(text)
    L = len(off)
    crlf = int(((text[1:] == 0x0A) & (text[:-1] == 0x0D)).sum())
    assert 0.008 < crlf / L < 0.012
    blank = int((nodes["kind"] == 0).sum())
    assert 0.004 < blank / L < 0.006


def test_patterns_shape():
    pats = datagen.patterns(1024, seed=0xC3B)
    assert len(pats) == 1024
    assert all(3 <= len(p) <= 32 for p in pats)
    assert all(all(97 <= c <= 122 or c == 32 for c in p) for p in pats)


def test_golden_hash_of_the_1mib_document():
    with open(os.path.join(ROOT, "tests", "golden", "c1_seed0.json")) as f:
        g = json.load(f)
    text = datagen.document(1 << 20, seed=0)
    assert hashlib.sha256(text.tobytes()).hexdigest() == g["text_sha256"]
    off, nodes = orc.parse(text)
    assert len(off) == g["lines"]
    assert hashlib.sha256(off.tobytes()).hexdigest() == g["offsets_sha256"]
    assert hashlib.sha256(nodes.tobytes()).hexdigest() == g["nodes_sha256"]
    links = orc.links(nodes)
    for name, arr in zip(("parent", "subtree_end", "first_child", "next_sibling"), links):
        assert hashlib.sha256(arr.tobytes()).hexdigest() == g[name + "_sha256"], name
