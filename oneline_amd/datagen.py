"""Synthetic workloads (ctypes over libol_datagen.so, built from csrc/datagen.c).

Document shape: see csrc/datagen.c.  Named workloads:

* ``document(1 << 20, seed=0)``                    -- 1 MiB, '\\n' only (the CPU-only case)
* ``document(1 << 31, seed=0, crlf=10, empty=5)``  -- 2 GiB, 1% CRLF, 0.5% empty lines
* ``patterns(count, seed)``                        -- dictionaries of 3..32-byte patterns
* ``lognormal_sizes(count, seed, median, sigma, lo, hi)`` -- batch / serving document sizes
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GEN_PATH = os.path.join(HERE, "libol_datagen.so")

_gen = None


def _lib():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: build it with `python -m oneline_amd.build datagen`")
        g = ctypes.CDLL(GEN_PATH)
        g.ol_gen_document.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint32,
                                      ctypes.c_uint32, ctypes.c_int]
        g.ol_gen_document.restype = None
        g.ol_gen_vocab_word.argtypes = [ctypes.c_uint32, ctypes.c_char_p]
        g.ol_gen_vocab_word.restype = ctypes.c_int
        g.ol_gen_patterns.argtypes = [ctypes.c_size_t, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]
        g.ol_gen_patterns.restype = None
        g.ol_gen_literal.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
        g.ol_gen_literal.restype = None
        g.ol_gen_lognormal_sizes.argtypes = [ctypes.c_size_t, ctypes.c_uint64, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p]
        g.ol_gen_lognormal_sizes.restype = None
        g.ol_gen_mix.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        g.ol_gen_mix.restype = ctypes.c_uint64
        _gen = g
    return _gen


def default_threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def document(n: int, seed: int = 0, crlf: int = 0, empty: int = 0, threads: int | None = None,
             out: np.ndarray | None = None) -> np.ndarray:
    """Exactly n bytes of synthetic markup; crlf / empty in permille of lines."""
    buf = out if out is not None else np.empty(n, np.uint8)
    assert buf.dtype == np.uint8 and buf.size >= n
    if n:
        _lib().ol_gen_document(buf.ctypes.data, n, seed, crlf, empty, threads or default_threads())
    return buf[:n]


def vocab_word(k: int) -> bytes:
    b = ctypes.create_string_buffer(16)
    ln = _lib().ol_gen_vocab_word(k, b)
    return b.raw[:ln]


def patterns(count: int, seed: int = 0xC3B) -> list:
    pool = np.zeros(max(1, count) * 32, np.uint8)
    offs = np.zeros(max(1, count), np.uint64)
    lens = np.zeros(max(1, count), np.uint32)
    _lib().ol_gen_patterns(count, seed, pool.ctypes.data, offs.ctypes.data, lens.ctypes.data)
    return [bytes(pool[int(o):int(o) + int(ln)]) for o, ln in zip(offs[:count], lens[:count])]


def literal(length: int, kind: int = 0, seed: int = 0xC3A) -> bytes:
    out = np.zeros(length, np.uint8)
    _lib().ol_gen_literal(length, kind, seed, out.ctypes.data)
    return bytes(out)


def lognormal_sizes(count: int, seed: int, median: float, sigma: float, lo: int, hi: int) -> np.ndarray:
    out = np.zeros(count, np.uint64)
    _lib().ol_gen_lognormal_sizes(count, seed, median, sigma, lo, hi, out.ctypes.data)
    return out


def mix(seed: int, idx: int) -> int:
    return int(_lib().ol_gen_mix(seed, idx))
