"""The allocation-free literal search (ol_search_device_noalloc) against the
allocating one (ol_search_device), through the C ABI: same totals and offsets
on two calls in a row, the second over a smaller document, with capacities
above and below the total.  Also checks the results against the oracle."""
import ctypes

import numpy as np
import pytest

import oneline_amd as ol
from oneline_amd import datagen

pytestmark = pytest.mark.gpu


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def _dev_alloc(hip, nbytes):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(16, nbytes))) == 0
    return p


def _download(hip, p, count):
    out = np.zeros(count, np.uint64)
    if count:
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(count * 8), 2) == 0
    return out


def _search(lib, parser, doc, pat, fn, d_off, cap):
    tot = ctypes.c_uint64()
    ol._check(getattr(lib, fn)(parser._ctx, doc._h, pat, len(pat), d_off, cap, ctypes.byref(tot)))
    return tot.value


@pytest.mark.parametrize("pat", [b"the", b"aiwh", datagen.vocab_word(3), b"| | -. "])
def test_noalloc_equals_allocating_search(parser, pat):
    from oracle import pyoracle as orc
    lib = ol.load_library()
    hip = _hip()
    big = datagen.document(24 << 20, seed=5)
    small = datagen.document(3 << 20, seed=6)
    for k, text in enumerate((big, small)):        # the second call: a smaller input
        print(f"[noalloc] call {k}: {text.size} bytes, pattern {pat!r}", flush=True)
        doc = parser.upload(text, nodes=False)
        want = orc.find_all(text, pat).astype(np.uint64)
        cap = want.size + 8
        d_a, d_b = _dev_alloc(hip, cap * 8), _dev_alloc(hip, cap * 8)
        try:
            ta = _search(lib, parser, doc, pat, "ol_search_device", d_a, cap)
            tb = _search(lib, parser, doc, pat, "ol_search_device_noalloc", d_b, cap)
            assert ta == tb == want.size
            a, b = _download(hip, d_a, ta), _download(hip, d_b, tb)
            assert np.array_equal(a, b)
            assert np.array_equal(b, want)
            # a capacity below the total keeps the first matches and the true total
            short = max(0, want.size // 3)
            tb2 = _search(lib, parser, doc, pat, "ol_search_device_noalloc", d_b, short)
            assert tb2 == want.size
            assert np.array_equal(_download(hip, d_b, short), want[:short])
            # the count alone
            assert _search(lib, parser, doc, pat, "ol_search_device_noalloc", None, 0) == want.size
        finally:
            hip.hipFree(d_a)
            hip.hipFree(d_b)
            doc.close()


def test_noalloc_rejects_bad_patterns(parser):
    lib = ol.load_library()
    doc = parser.upload(datagen.document(1 << 16, seed=1), nodes=False)
    tot = ctypes.c_uint64()
    assert lib.ol_search_device_noalloc(parser._ctx, doc._h, b"", 0, None, 0, ctypes.byref(tot)) == -3
    assert lib.ol_search_device_noalloc(parser._ctx, doc._h, b"x" * 65, 65, None, 0, ctypes.byref(tot)) == -3
    # the context stays usable
    assert _search(lib, parser, doc, b"the", "ol_search_device_noalloc", None, 0) >= 0
    doc.close()
