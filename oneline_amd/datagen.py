"""Synthetic workloads (ctypes over libol_datagen.so, built from csrc/datagen.c).

Document shape: see csrc/datagen.c.  Named workloads:

* ``document(1 << 20, seed=0)``                    -- 1 MiB, '\\n' only (the CPU-only case)
* ``document(1 << 31, seed=0, crlf=10, empty=5)``  -- 2 GiB, 1% CRLF, 0.5% empty lines
* ``patterns(count, seed)``                        -- dictionaries of 3..32-byte patterns
* ``lognormal_sizes(count, seed, median, sigma, lo, hi)`` -- batch / serving document sizes
* ``profile(name, target_bytes, seed)``            -- the reference's six document profiles
  (``PROFILES``: Repeat, ShortLines, Mixed, LongLines, Prose, Nested; csrc/profiles.c)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# OL_DATAGEN_LIB: another build of the same sources (the sanitizer build)
GEN_PATH = os.environ.get("OL_DATAGEN_LIB") or os.path.join(HERE, "libol_datagen.so")

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
        g.ol_gen_profile.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_void_p,
                                     ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        g.ol_gen_profile.restype = ctypes.c_int
        g.ol_gen_prose_vocab.argtypes = [ctypes.c_void_p]
        g.ol_gen_prose_vocab.restype = None
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


# Profile::ALL order and Profile::name, crates/oneline-core/src/datagen.rs:29-49
PROFILES = ("repeat", "short-lines", "mixed", "long-lines", "prose", "nested")

# a generated line adds at most ~4,100 bytes past the target (profiles.c)
_PROFILE_SLACK = 8192


def profile(name, target_bytes: int, seed: int) -> np.ndarray:
    """``datagen::generate(profile, target_bytes, seed)`` of the reference: at
    least target_bytes bytes, ending in '\n'.  ``name`` is a Profile::name or
    an index into PROFILES."""
    idx = PROFILES.index(name) if isinstance(name, str) else int(name)
    buf = np.empty(target_bytes + _PROFILE_SLACK, np.uint8)
    n = ctypes.c_size_t()
    rc = _lib().ol_gen_profile(idx, target_bytes, seed, buf.ctypes.data, buf.size, ctypes.byref(n))
    if rc != 0:
        raise ValueError(f"profile generation failed ({rc}) for {name!r}")
    return buf[:n.value]


def prose_vocabulary() -> list:
    """``datagen::prose_vocabulary()``: 4096 words, index 0 the most frequent."""
    raw = ctypes.create_string_buffer(4096 * 11)
    _lib().ol_gen_prose_vocab(raw)
    b = raw.raw
    return [b[11 * k:11 * k + 11].split(b"\0", 1)[0] for k in range(4096)]
