"""ctypes front of the CPU oracle (oracle/oracle.c, oracle/oracle_par.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# OL_ORACLE_LIB: another build of the same sources (the sanitizer build)
LIB = os.environ.get("OL_ORACLE_LIB") or os.path.join(HERE, "liboneline_oracle.so")
NONE = 0xFFFFFFFF
NODE = np.dtype([("depth", "<u4"), ("kind", "<u4"), ("span_start", "<u4"), ("span_len", "<u4")])

_o = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(LIB):
            build()
        o = ctypes.CDLL(LIB)
        P, S = ctypes.c_void_p, ctypes.c_size_t
        o.orc_parse_line.argtypes = [P, S, ctypes.c_uint32, P]
        o.orc_parse_line.restype = None
        o.orc_line_count.argtypes = [P, S]
        o.orc_line_count.restype = S
        o.orc_scan_line_offsets.argtypes = [P, S, P, S]
        o.orc_scan_line_offsets.restype = S
        o.orc_parse.argtypes = [P, S, P, P]
        o.orc_parse.restype = S
        o.orc_links_sequential.argtypes = [P, P, S, P, P]
        o.orc_links_sequential.restype = None
        o.orc_links_derived.argtypes = [P, P, P, S, P, P]
        o.orc_links_derived.restype = None
        o.orc_find_all.argtypes = [P, S, P, S, P, S]
        o.orc_find_all.restype = S
        o.orc_many_find_all.argtypes = [P, S, P, P, P, S, P, P, S]
        o.orc_many_find_all.restype = S
        o.orc_parse_par.argtypes = [P, S, P, P, ctypes.c_int]
        o.orc_parse_par.restype = S
        o.orc_links_par.argtypes = [P, S, P, P, P, P, ctypes.c_int]
        o.orc_links_par.restype = ctypes.c_int
        o.orc_find_all_par.argtypes = [P, S, P, S, ctypes.c_int, P, S]
        o.orc_find_all_par.restype = S
        o.orc_bridge.argtypes = [P, P, P, S, P, P]
        o.orc_bridge.restype = None
        o.orc_ac_build.argtypes = [P, P, P, S]
        o.orc_ac_build.restype = P
        o.orc_ac_free.argtypes = [P]
        o.orc_ac_free.restype = None
        o.orc_ac_find_par.argtypes = [P, P, S, ctypes.c_int, P, P, S]
        o.orc_ac_find_par.restype = S
        _o = o
    return _o


def _arr(text) -> np.ndarray:
    if isinstance(text, np.ndarray):
        return np.ascontiguousarray(text, np.uint8).reshape(-1)
    if isinstance(text, str):
        text = text.encode()
    return np.frombuffer(bytes(text), np.uint8) if len(text) else np.zeros(0, np.uint8)


def _p(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def parse_line(line: bytes, start: int = 0):
    a = _arr(line)
    out = np.zeros(1, NODE)
    lib().orc_parse_line(_p(a), a.size, start, _p(out))
    return tuple(int(x) for x in out[0])


def scan_line_offsets(text) -> np.ndarray:
    a = _arr(text)
    n = lib().orc_line_count(_p(a), a.size)
    out = np.zeros(n, np.uint32)
    lib().orc_scan_line_offsets(_p(a), a.size, _p(out), n)
    return out


def parse(text):
    """(offsets u32, nodes NODE) by splitting on '\\n'."""
    a = _arr(text)
    n = lib().orc_line_count(_p(a), a.size)
    off = np.zeros(n, np.uint32)
    nodes = np.zeros(n, NODE)
    lib().orc_parse(_p(a), a.size, _p(off), _p(nodes))
    return off, nodes


def links(nodes: np.ndarray):
    """(parent, subtree_end, first_child, next_sibling) for NODE records."""
    n = nodes.size
    depth = np.ascontiguousarray(nodes["depth"], np.uint32)
    kind = np.ascontiguousarray(nodes["kind"], np.uint32)
    return links_from(depth, kind)


def links_from(depth: np.ndarray, kind: np.ndarray):
    n = depth.size
    parent = np.zeros(n, np.uint32)
    se = np.zeros(n, np.uint32)
    fc = np.zeros(n, np.uint32)
    ns = np.zeros(n, np.uint32)
    depth = np.ascontiguousarray(depth, np.uint32)
    kind = np.ascontiguousarray(kind, np.uint32)
    lib().orc_links_sequential(_p(depth), _p(kind), n, _p(parent), _p(se))
    lib().orc_links_derived(_p(kind), _p(parent), _p(se), n, _p(fc), _p(ns))
    return parent, se, fc, ns


def line_count(text) -> int:
    a = _arr(text)
    return int(lib().orc_line_count(_p(a), a.size))


def parse_par(text, threads: int, out=None):
    """parse() on `threads` threads (the two-pass chunked parse); `out` may
    hold (offsets, nodes) arrays of line_count(text) entries to fill."""
    a = _arr(text)
    if out is None:
        n = lib().orc_line_count(_p(a), a.size)
        out = (np.empty(n, np.uint32), np.empty(n, NODE))
    off, nodes = out
    got = lib().orc_parse_par(_p(a), a.size, _p(off), _p(nodes), threads)
    assert got == off.size == nodes.size
    return off, nodes


def links_par(nodes: np.ndarray, threads: int, out=None):
    """links() on `threads` threads (three-phase nearest-smaller-values);
    `out` may hold the four u32 arrays to fill."""
    n = nodes.size
    nodes = np.ascontiguousarray(nodes)
    if out is None:
        out = tuple(np.empty(n, np.uint32) for _ in range(4))
    parent, se, fc, ns = out
    if lib().orc_links_par(_p(nodes), n, _p(parent), _p(se), _p(fc), _p(ns), threads):
        raise MemoryError("orc_links_par")
    return parent, se, fc, ns


def find_all(text, pattern) -> np.ndarray:
    a, p = _arr(text), _arr(pattern)
    cnt = lib().orc_find_all(_p(a), a.size, _p(p), p.size, None, 0)
    out = np.zeros(cnt, np.uint64)
    lib().orc_find_all(_p(a), a.size, _p(p), p.size, _p(out), cnt)
    return out


def many_find_all(text, patterns):
    """(offsets u64, ids u32) of every (start, id), sorted by start then id."""
    a = _arr(text)
    lens = np.array([len(p) for p in patterns], np.uint32)
    offs = np.zeros(len(patterns), np.uint64)
    if len(patterns):
        offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
    pool = _arr(b"".join(bytes(p) for p in patterns) + b"\0")
    cnt = lib().orc_many_find_all(_p(a), a.size, _p(pool), _p(offs), _p(lens), len(patterns), None, None, 0)
    oo = np.zeros(cnt, np.uint64)
    ii = np.zeros(cnt, np.uint32)
    lib().orc_many_find_all(_p(a), a.size, _p(pool), _p(offs), _p(lens), len(patterns), _p(oo), _p(ii), cnt)
    return oo, ii


def find_all_par(text, pattern, threads: int) -> np.ndarray:
    """find_all on `threads` threads over 1 MiB+ chunks with a halo (the
    reference's find_all_par); ascending starts."""
    a, p = _arr(text), _arr(pattern)
    cnt = lib().orc_find_all_par(_p(a), a.size, _p(p), p.size, threads, None, 0)
    out = np.zeros(cnt, np.uint64)
    lib().orc_find_all_par(_p(a), a.size, _p(p), p.size, threads, _p(out), cnt)
    return out


class AhoCorasick:
    """The overlapping Aho-Corasick DFA search (the reference's ManyPatterns)."""

    def __init__(self, patterns):
        lens = np.array([len(p) for p in patterns], np.uint32)
        offs = np.zeros(len(patterns), np.uint64)
        if len(patterns):
            offs[1:] = np.cumsum(lens[:-1], dtype=np.uint64)
        self._pool = _arr(b"".join(bytes(p) for p in patterns) + b"\0")
        self._lens, self._offs = lens, offs
        self._h = lib().orc_ac_build(_p(self._pool), _p(offs), _p(lens), len(patterns))

    def count(self, text, threads: int) -> int:
        a = _arr(text)
        return int(lib().orc_ac_find_par(self._h, _p(a), a.size, threads, None, None, 0))

    def find(self, text, threads: int, sort: bool = True):
        """(starts u64, ids u32); sorted by (start, id) when `sort`."""
        a = _arr(text)
        cnt = self.count(a, threads)
        off = np.zeros(cnt, np.uint64)
        ids = np.zeros(cnt, np.uint32)
        lib().orc_ac_find_par(self._h, _p(a), a.size, threads, _p(off), _p(ids), cnt)
        if sort:
            order = np.lexsort((ids, off))
            off, ids = off[order], ids[order]
        return off, ids

    def close(self):
        if self._h:
            lib().orc_ac_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bridge(text, nodes: np.ndarray, parent: np.ndarray):
    """The AST bridge's structural pass (crates/oneline-tomet/src/lib.rs:20-90):
    per-line flags (1 literal, 2 block, 4 cut, 8 pipe child, 16 bare blank) and
    the container line of each block (NONE at the top level or without a block)."""
    t = _arr(text)
    nodes = np.ascontiguousarray(nodes, NODE)
    parent = np.ascontiguousarray(parent, np.uint32)
    n = nodes.size
    flags = np.zeros(n, np.uint8)
    cont = np.zeros(n, np.uint32)
    lib().orc_bridge(_p(t), _p(nodes), _p(parent), n, _p(flags), _p(cont))
    return flags, cont
