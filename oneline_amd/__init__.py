"""oneline_amd -- MI355X-native engine for oneline's GPU parse-and-search path.

Python mirror of the reference's ``oneline-gpu`` public API
(crates/oneline-gpu/src/lib.rs:27-36) over the C ABI in ``include/oneline_hip.h``:
``GpuParser`` with ``scan_line_offsets``, ``parse_raw``, ``parse``, ``upload``,
``build_tree`` / ``build_tree_reusing``, ``download_links``, ``search_document``,
``search_raw``, ``compile_patterns`` and ``search_many``.  Same names, argument
meaning and error behaviour: bad input raises ``OnelineError`` (the reference's
``anyhow::Error``) and leaves the parser usable.

The native library must be built (``python -m oneline_amd.build``); nothing here
falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes
import os
import sys
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboneline_hip.so")

NONE = 0xFFFFFFFF
# ol_tree_bridge flags (include/oneline_hip.h)
BRIDGE_LITERAL, BRIDGE_BLOCK, BRIDGE_CUT, BRIDGE_PIPE_CHILD, BRIDGE_BARE_BLANK = 1, 2, 4, 8, 16
MAX_GPU_INPUT_LEN = 4294967292
MAX_PATTERN_LEN = 64
MIN_SET_PATTERN_LEN = 1
MAX_SET_PATTERN_LEN = 64
MAX_PATTERNS = 1 << 22
MAX_TREE_DEPTH = 32

# link arrays of a tree rebuild (OL_LINK_*)
LINK_PARENT, LINK_SUBTREE_END, LINK_FIRST_CHILD, LINK_NEXT_SIBLING = 1, 2, 4, 8
LINK_ALL = 15

# NodeKind, crates/oneline-core/src/lib.rs:41-51
BLANK, TEXT, HEADING, ELEMENT, LIST, CONNECT, ORDERED_LIST, COMMENT = range(8)

# FlatNode, crates/oneline-core/src/lib.rs:87-96
FLAT_NODE = np.dtype([("depth", "<u4"), ("kind", "<u4"), ("span_start", "<u4"), ("span_len", "<u4")])

OL_OK = 0
ERR_NAMES = {-1: "invalid", -2: "too_large", -3: "pattern", -4: "too_many", -5: "device",
             -6: "too_deep", -7: "nomem"}


class OnelineError(RuntimeError):
    """A failed engine call (bad input, device failure, ...)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{message} [{ERR_NAMES.get(code, code)}]")
        self.code = code


_lib = None


def load_library() -> ctypes.CDLL:
    """Loads liboneline_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m oneline_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    S = ctypes.c_size_t
    U64P = ctypes.POINTER(ctypes.c_uint64)
    U32P = ctypes.POINTER(ctypes.c_uint32)
    sig = {
        "ol_ctx_create": [ctypes.POINTER(P)],
        "ol_ctx_destroy": [P],
        "ol_ctx_adapter_name": [P, ctypes.c_char_p, S],
        "ol_ctx_is_software": [P],
        "ol_ctx_timestamps_supported": [P],
        "ol_ctx_synchronize": [P],
        "ol_free": [P],
        "ol_scan_line_offsets": [P, P, S, ctypes.POINTER(P), ctypes.POINTER(S)],
        "ol_parse_raw": [P, P, S, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.POINTER(S)],
        "ol_parse": [P, P, S, P, S, ctypes.POINTER(P)],
        "ol_doc_upload": [P, P, S, ctypes.POINTER(P)],
        "ol_doc_upload_lean": [P, P, S, ctypes.POINTER(P)],
        "ol_doc_free": [P],
        "ol_doc_line_count": [P],
        "ol_doc_text_len": [P],
        "ol_doc_download": [P, P, P, P, P],
        "ol_doc_reparse": [P, P],
        "ol_tree_build": [P, P, P, ctypes.POINTER(P)],
        "ol_tree_free": [P],
        "ol_tree_line_count": [P],
        "ol_tree_download": [P, P, P, P, P, P],
        "ol_tree_rebuild_async": [P, P, P],
        "ol_tree_rebuild_links_async": [P, P, P, ctypes.c_uint32],
        "ol_tree_check": [P, P],
        "ol_tree_bridge": [P, P, P, P, P],
        "ol_search_document": [P, P, P, S, S, U32P, ctypes.POINTER(P), ctypes.POINTER(S)],
        "ol_search_raw": [P, P, S, P, S, S, U32P, ctypes.POINTER(P), ctypes.POINTER(S)],
        "ol_search_device": [P, P, P, S, P, S, U64P],
        "ol_search_device_noalloc": [P, P, P, S, P, S, U64P],
        "ol_doc_save": [P, P, ctypes.c_char_p],
        "ol_doc_load": [P, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(P)],
        "ol_query_document": [P, P, ctypes.c_uint32, ctypes.c_uint32, S, U32P, ctypes.POINTER(P), ctypes.POINTER(S)],
        "ol_query_subtree": [P, P, P, S, ctypes.c_uint32, ctypes.c_uint32, S, U32P, ctypes.POINTER(P),
                             ctypes.POINTER(S)],
        "ol_query_raw": [P, P, S, ctypes.c_uint32, ctypes.c_uint32, S, U32P, ctypes.POINTER(P), ctypes.POINTER(S)],
        "ol_query": [P, P, S, P, S, ctypes.c_uint32, ctypes.c_uint32, S, U32P, ctypes.POINTER(P), ctypes.POINTER(S)],
        "ol_profile_call_begin": [P],
        "ol_profile_call_end": [P, ctypes.POINTER(ctypes.c_double), P, S, ctypes.POINTER(S), P, S, ctypes.POINTER(S)],
        "ol_dict_compile": [P, P, P, P, S, ctypes.POINTER(P)],
        "ol_dict_free": [P],
        "ol_dict_len": [P],
        "ol_search_many": [P, P, P, S, P, U64P, ctypes.POINTER(P), ctypes.POINTER(S)],
        "ol_search_many_device": [P, P, P, P, P, S, U64P],
        "ol_profile_enable": [P, ctypes.c_int],
        "ol_tune_set": [P, ctypes.c_char_p, ctypes.c_int],
        "ol_tune_get": [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)],
        "ol_profile_take": [P, P, S, P, P, S, ctypes.POINTER(S)],
        "ol_calib_scan_floor": [P, P, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
        "ol_batch_create": [P, P, P, S, ctypes.POINTER(P)],
        "ol_batch_run": [P, P, P, ctypes.c_int],
        "ol_batch_lines": [P, P, U64P, P, P],
        "ol_batch_download_doc": [P, P, S, P, P, P, P, P, P],
        "ol_batch_matches": [P, P, S, U64P, P, P, S],
        "ol_server_create": [P, ctypes.POINTER(P)],
        "ol_server_add_dict": [P, P, P, S, ctypes.POINTER(ctypes.c_uint32)],
        "ol_server_submit": [P, ctypes.c_int, ctypes.c_uint32, P, S, ctypes.c_int, U64P],
        "ol_server_wait": [P, ctypes.c_uint64, ctypes.c_double],
        "ol_server_result_lines": [P, ctypes.c_uint64, ctypes.POINTER(S), P, P, P, P, P, P],
        "ol_server_result_matches": [P, ctypes.c_uint64, ctypes.POINTER(S), P, P],
        "ol_server_timing": [P, ctypes.c_uint64, P, P, P],
        "ol_server_release": [P, ctypes.c_uint64],
        "ol_server_load": [P, P, P, S, ctypes.c_double, S, ctypes.c_double, ctypes.c_uint64, P],
        "ol_server_load_serial": [P, P, P, S, S, ctypes.c_double, ctypes.c_uint64, P],
        "ol_server_put_doc": [P, P, S, U64P],
        "ol_server_drop_doc": [P, ctypes.c_uint64],
        "ol_server_submit_doc": [P, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, U64P],
        "ol_server_slot_of": [P, ctypes.c_uint64, U32P],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    for name in ("ol_doc_line_count", "ol_doc_text_len", "ol_tree_line_count", "ol_dict_len",
                 "ol_batch_doc_count", "ol_batch_text_bytes"):
        getattr(lib, name).restype = ctypes.c_size_t
        getattr(lib, name).argtypes = [P]
    for name in ("ol_server_arena_bytes", "ol_server_hbm_used"):
        getattr(lib, name).restype = ctypes.c_size_t
        getattr(lib, name).argtypes = [P]
    lib.ol_server_last_error.argtypes = []
    lib.ol_server_last_error.restype = ctypes.c_char_p
    for name in ("ol_ctx_destroy", "ol_free", "ol_doc_free", "ol_tree_free", "ol_dict_free", "ol_batch_free",
                 "ol_server_destroy"):
        getattr(lib, name).restype = None
        getattr(lib, name).argtypes = [P]
    lib.ol_last_error.argtypes = []
    lib.ol_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != OL_OK:
        msg = _lib.ol_last_error().decode("utf-8", "replace")
        raise OnelineError(rc, msg)


def as_bytes_array(text) -> np.ndarray:
    """Zero-copy uint8 view of str/bytes/bytearray/ndarray input."""
    if isinstance(text, np.ndarray):
        return np.ascontiguousarray(text, dtype=np.uint8).reshape(-1)
    if isinstance(text, str):
        text = text.encode("utf-8")
    return np.frombuffer(bytes(text), dtype=np.uint8) if len(text) else np.zeros(0, np.uint8)


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def _take(ptr: ctypes.c_void_p, count: int, dtype) -> np.ndarray:
    """Copies a library-allocated array into numpy and frees it."""
    dtype = np.dtype(dtype)
    if not ptr.value:
        return np.zeros(0, dtype)
    try:
        if count == 0:
            return np.zeros(0, dtype)
        buf = (ctypes.c_char * (count * dtype.itemsize)).from_address(ptr.value)
        return np.frombuffer(buf, dtype=dtype, count=count).copy()
    finally:
        _lib.ol_free(ptr)


# ---- result types (crates/oneline-core/src/lib.rs:113-158, links.rs:22-47,
#      crates/oneline-gpu/src/search.rs:12-21, many.rs:75-94) -----------------------------
@dataclass
class FlatTree:
    source: np.ndarray
    nodes: np.ndarray          # FLAT_NODE records
    line_offsets: np.ndarray   # u32

    def content_of(self, i: int) -> bytes:
        nd = self.nodes[i]
        s = int(nd["span_start"])
        return bytes(self.source[s:s + int(nd["span_len"])])


@dataclass
class TreeLinks:
    parent: np.ndarray
    subtree_end: np.ndarray
    first_child: np.ndarray
    next_sibling: np.ndarray

    def __len__(self):
        return len(self.parent)

    def parent_of(self, line: int) -> Optional[int]:
        p = int(self.parent[line])
        return None if p == NONE else p

    def subtree(self, line: int) -> range:
        return range(line, int(self.subtree_end[line]))


@dataclass
class SearchResult:
    total_matches: int
    offsets: np.ndarray


@dataclass
class QueryResult:
    """QueryResult (crates/oneline-gpu/src/query.rs:11): every matching line
    counted; the first min(total, max_matches) line indices, ascending."""
    total_matches: int
    indices: np.ndarray


QUERY_ANY_KIND = 0             # kind 0 matches every kind (NodeKind.Blank is 0)
QUERY_ANY_DEPTH = 0xFFFFFFFF


@dataclass
class ManyResult:
    counts: np.ndarray
    total_hits: int
    hits: np.ndarray = field(default_factory=lambda: np.zeros(0, np.dtype([("pattern", "<u4"), ("offset", "<u4")])))


MANY_HIT = np.dtype([("pattern", "<u4"), ("offset", "<u4")])

PROFILE_STEP = np.dtype([("label", "S48"), ("ms", "<f8")])


@dataclass
class Profile:
    """Where the time of one call went (crates/oneline-gpu/src/timing.rs:109-123):
    host wall time from the start to the last host step, the host steps (time
    since the previous one, in execution order) and the GPU passes (HIP-event
    durations, in launch order).  Durations in seconds."""
    wall: float
    host_steps: List[tuple]
    gpu_passes: List[tuple]

    def gpu_total(self) -> float:
        return sum(d for _, d in self.gpu_passes)


class GpuDocument:
    """A document uploaded, scanned and classified once, resident in HBM
    (crates/oneline-gpu/src/document.rs:16-90)."""

    def __init__(self, parser: "GpuParser", handle: ctypes.c_void_p):
        self._parser = parser
        self._h = handle

    def line_count(self) -> int:
        return int(_lib.ol_doc_line_count(self._h))

    def text_len(self) -> int:
        return int(_lib.ol_doc_text_len(self._h))

    def download(self, nodes: bool = False):
        """(offsets u32, classes u8, nodes FLAT_NODE or None) of the resident document."""
        L = self.line_count()
        off = np.zeros(L, np.uint32)
        cls = np.zeros(L, np.uint8)
        nd = np.zeros(L, FLAT_NODE) if nodes else None
        _check(_lib.ol_doc_download(self._parser._ctx, self._h, _ptr(off), _ptr(cls),
                                    _ptr(nd) if nd is not None else None))
        return off, cls, nd

    def close(self) -> None:
        if self._h:
            _lib.ol_doc_free(self._h)
            self._h = None

    def __del__(self):
        if sys.is_finalizing():   # the runtime may be gone: the process reclaims it
            return
        try:
            self.close()
        except Exception:
            pass


class GpuTree:
    """parent / subtree_end / first_child / next_sibling of a document, resident
    in HBM (crates/oneline-gpu/src/tree.rs:29-46)."""

    def __init__(self, handle: ctypes.c_void_p):
        self._h = handle

    def line_count(self) -> int:
        return int(_lib.ol_tree_line_count(self._h))

    def _release(self) -> ctypes.c_void_p:
        h, self._h = self._h, None
        return h

    def close(self) -> None:
        if self._h:
            _lib.ol_tree_free(self._h)
            self._h = None

    def __del__(self):
        if sys.is_finalizing():   # the runtime may be gone: the process reclaims it
            return
        try:
            self.close()
        except Exception:
            pass


class GpuPatternSet:
    """A compiled dictionary (crates/oneline-gpu/src/many.rs:116-150)."""

    def __init__(self, handle: ctypes.c_void_p):
        self._h = handle

    def __len__(self) -> int:
        return int(_lib.ol_dict_len(self._h))

    def close(self) -> None:
        if self._h:
            _lib.ol_dict_free(self._h)
            self._h = None

    def __del__(self):
        if sys.is_finalizing():   # the runtime may be gone: the process reclaims it
            return
        try:
            self.close()
        except Exception:
            pass


class GpuBatch:
    """Many documents parsed, tree-linked and dictionary-searched by one launch
    sequence (ol_batch_*); per-document results equal the single-document calls."""

    def __init__(self, parser: "GpuParser", handle: ctypes.c_void_p, keep):
        self._p = parser
        self._h = handle
        self._keep = keep

    def __len__(self) -> int:
        return int(_lib.ol_batch_doc_count(self._h))

    def text_bytes(self) -> int:
        return int(_lib.ol_batch_text_bytes(self._h))

    def run(self, dictionary: Optional["GpuPatternSet"] = None, graph: bool = False,
            search_only: bool = False) -> None:
        """Runs (the first time: sizes, synchronously) parse + tree (+ search)."""
        flags = (1 if graph else 0) | (2 if search_only else 0)
        _check(_lib.ol_batch_run(self._p._ctx, self._h, dictionary._h if dictionary is not None else None, flags))

    def lines(self):
        """(total lines, per-document first line, per-document line count)."""
        n = len(self)
        st = np.zeros(max(1, n), np.uint32)
        cnt = np.zeros(max(1, n), np.uint32)
        tot = ctypes.c_uint64()
        _check(_lib.ol_batch_lines(self._p._ctx, self._h, ctypes.byref(tot), _ptr(st), _ptr(cnt)))
        return tot.value, st[:n], cnt[:n]

    def document(self, i: int):
        """(offsets, classes, TreeLinks) of document i, local to it."""
        _, st, cnt = self.lines()
        n = int(cnt[i])
        off = np.zeros(n, np.uint32)
        cls = np.zeros(n, np.uint8)
        par, se, fc, ns = (np.zeros(n, np.uint32) for _ in range(4))
        _check(_lib.ol_batch_download_doc(self._p._ctx, self._h, i, _ptr(off), _ptr(cls), _ptr(par), _ptr(se),
                                          _ptr(fc), _ptr(ns)))
        return off, cls, TreeLinks(par, se, fc, ns)

    def matches(self, i: int):
        """(offsets, ids) of document i's dictionary matches, (offset, id) order."""
        cnt = ctypes.c_uint64()
        _check(_lib.ol_batch_matches(self._p._ctx, self._h, i, ctypes.byref(cnt), None, None, 0))
        off = np.zeros(cnt.value, np.uint32)
        ids = np.zeros(cnt.value, np.uint32)
        _check(_lib.ol_batch_matches(self._p._ctx, self._h, i, ctypes.byref(cnt), _ptr(off), _ptr(ids),
                                     cnt.value))
        return off, ids

    def close(self) -> None:
        if self._h:
            _lib.ol_batch_free(self._h)
            self._h = None

    def __del__(self):
        if sys.is_finalizing():   # the runtime may be gone: the process reclaims it
            return
        try:
            self.close()
        except Exception:
            pass


class GpuParser:
    """GpuParser (crates/oneline-gpu/src/parser.rs:8-57) over HIP device 0."""

    def __init__(self):
        load_library()
        h = ctypes.c_void_p()
        _check(_lib.ol_ctx_create(ctypes.byref(h)))
        self._ctx = h

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            _lib.ol_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        if sys.is_finalizing():   # the runtime may be gone: the process reclaims it
            return
        try:
            self.close()
        except Exception:
            pass

    # -- info -------------------------------------------------------------------
    def adapter_name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _check(_lib.ol_ctx_adapter_name(self._ctx, buf, 256))
        return buf.value.decode()

    def is_software(self) -> bool:
        return bool(_lib.ol_ctx_is_software(self._ctx))

    def timestamps_supported(self) -> bool:
        return bool(_lib.ol_ctx_timestamps_supported(self._ctx))

    def synchronize(self) -> None:
        _check(_lib.ol_ctx_synchronize(self._ctx))

    # -- parse --------------------------------------------------------------------
    def scan_line_offsets(self, text) -> np.ndarray:
        t = as_bytes_array(text)
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        _check(_lib.ol_scan_line_offsets(self._ctx, _ptr(t), t.size, ctypes.byref(p), ctypes.byref(n)))
        return _take(p, n.value, np.uint32)

    def parse_raw(self, text) -> FlatTree:
        t = as_bytes_array(text)
        pn = ctypes.c_void_p()
        po = ctypes.c_void_p()
        n = ctypes.c_size_t()
        _check(_lib.ol_parse_raw(self._ctx, _ptr(t), t.size, ctypes.byref(pn), ctypes.byref(po), ctypes.byref(n)))
        return FlatTree(t, _take(pn, n.value, FLAT_NODE), _take(po, n.value, np.uint32))

    def parse(self, text, line_offsets) -> FlatTree:
        t = as_bytes_array(text)
        o = np.ascontiguousarray(line_offsets, dtype=np.uint32)
        pn = ctypes.c_void_p()
        _check(_lib.ol_parse(self._ctx, _ptr(t), t.size, _ptr(o), o.size, ctypes.byref(pn)))
        return FlatTree(t, _take(pn, o.size, FLAT_NODE), o.copy())

    def upload(self, text, nodes: bool = True) -> GpuDocument:
        t = as_bytes_array(text)
        h = ctypes.c_void_p()
        fn = _lib.ol_doc_upload if nodes else _lib.ol_doc_upload_lean
        _check(fn(self._ctx, _ptr(t), t.size, ctypes.byref(h)))
        return GpuDocument(self, h)

    def save(self, doc: GpuDocument, path: str) -> None:
        """Writes the parsed document as a .olb file (header, text, FlatNodes,
        line offsets, CRC-32)."""
        _check(_lib.ol_doc_save(self._ctx, doc._h, os.fsencode(path)))

    def load(self, path: str, verify: bool = True) -> GpuDocument:
        """A .olb file as a resident document, without scanning the text again."""
        h = ctypes.c_void_p()
        _check(_lib.ol_doc_load(self._ctx, os.fsencode(path), 1 if verify else 0, ctypes.byref(h)))
        return GpuDocument(self, h)

    # -- tree -----------------------------------------------------------------------
    def build_tree(self, doc: GpuDocument) -> GpuTree:
        return self.build_tree_reusing(doc, None)

    def build_tree_reusing(self, doc: GpuDocument, previous: Optional[GpuTree]) -> GpuTree:
        reuse = previous._release() if previous is not None else None
        h = ctypes.c_void_p()
        _check(_lib.ol_tree_build(self._ctx, doc._h, reuse, ctypes.byref(h)))
        return GpuTree(h)

    def reparse_async(self, doc: GpuDocument) -> None:
        """Scan + classify the resident text again into the resident arrays
        (no host round trip)."""
        _check(_lib.ol_doc_reparse(self._ctx, doc._h))

    def rebuild_tree_async(self, doc: GpuDocument, tree: GpuTree, links: int = LINK_ALL) -> None:
        """Links of `doc` into `tree` again, only the arrays in `links`
        (LINK_* bits); errors surface at check_tree / synchronize."""
        _check(_lib.ol_tree_rebuild_links_async(self._ctx, doc._h, tree._h, links))

    def check_tree(self, tree: GpuTree) -> None:
        _check(_lib.ol_tree_check(self._ctx, tree._h))

    def download_links(self, tree: GpuTree) -> TreeLinks:
        L = tree.line_count()
        arrs = [np.zeros(L, np.uint32) for _ in range(4)]
        _check(_lib.ol_tree_download(self._ctx, tree._h, *[_ptr(a) for a in arrs]))
        return TreeLinks(*arrs)

    # -- AST bridge ---------------------------------------------------------------------
    def bridge(self, doc: GpuDocument, tree: GpuTree) -> Tuple[np.ndarray, np.ndarray]:
        """The structural pass of to_tomet_ast_with_links
        (crates/oneline-tomet/src/lib.rs:20-90): per line the BRIDGE_* flags and
        the line whose block receives its block (NONE at the top level or
        without a block)."""
        L = doc.line_count()
        flags = np.zeros(L, np.uint8)
        container = np.zeros(L, np.uint32)
        _check(_lib.ol_tree_bridge(self._ctx, doc._h, tree._h, _ptr(flags), _ptr(container)))
        return flags, container

    # -- literal search ---------------------------------------------------------------
    def search_document(self, doc: GpuDocument, pattern: bytes, max_matches: int) -> SearchResult:
        pat = as_bytes_array(pattern)
        tot = ctypes.c_uint32()
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        if max_matches < 0:
            raise OnelineError(-1, "max_matches must be >= 0")
        _check(_lib.ol_search_document(self._ctx, doc._h, _ptr(pat), pat.size, max_matches,
                                       ctypes.byref(tot), ctypes.byref(p), ctypes.byref(n)))
        return SearchResult(tot.value, _take(p, n.value, np.uint32))

    def search_raw(self, text, pattern: bytes, max_matches: int) -> SearchResult:
        t = as_bytes_array(text)
        pat = as_bytes_array(pattern)
        tot = ctypes.c_uint32()
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        _check(_lib.ol_search_raw(self._ctx, _ptr(t), t.size, _ptr(pat), pat.size, max_matches,
                                  ctypes.byref(tot), ctypes.byref(p), ctypes.byref(n)))
        return SearchResult(tot.value, _take(p, n.value, np.uint32))

    # -- kind / depth queries (GpuParser::query_document / query_subtree / query_raw) ---
    @staticmethod
    def _query_args(kind, depth):
        k = QUERY_ANY_KIND if kind is None else int(kind)
        dp = QUERY_ANY_DEPTH if depth is None else int(depth)
        return k, dp

    def query_document(self, doc: GpuDocument, kind=None, depth=None, max_matches: int = 1 << 20) -> QueryResult:
        """Lines of `doc` with NodeKind `kind` and FlatNode depth `depth` (None:
        any); total plus the first min(total, max_matches) indices, ascending."""
        k, dp = self._query_args(kind, depth)
        tot, p, n = ctypes.c_uint32(), ctypes.c_void_p(), ctypes.c_size_t()
        _check(_lib.ol_query_document(self._ctx, doc._h, k, dp, max_matches, ctypes.byref(tot), ctypes.byref(p),
                                      ctypes.byref(n)))
        return QueryResult(tot.value, _take(p, n.value, np.uint32))

    def query_subtree(self, doc: GpuDocument, tree: "GpuTree", root: int, kind=None, depth=None,
                      max_matches: int = 1 << 20) -> QueryResult:
        """query_document restricted to the lines root .. subtree_end[root]."""
        k, dp = self._query_args(kind, depth)
        tot, p, n = ctypes.c_uint32(), ctypes.c_void_p(), ctypes.c_size_t()
        _check(_lib.ol_query_subtree(self._ctx, doc._h, tree._h, root, k, dp, max_matches, ctypes.byref(tot),
                                     ctypes.byref(p), ctypes.byref(n)))
        return QueryResult(tot.value, _take(p, n.value, np.uint32))

    def query(self, text, line_offsets, kind=None, depth=None, max_matches: int = 1 << 20) -> QueryResult:
        """GpuParser::query: the query over caller-given line offsets (no scan)."""
        t = as_bytes_array(text)
        o = np.ascontiguousarray(line_offsets, dtype=np.uint32)
        k, dp = self._query_args(kind, depth)
        tot, p, n = ctypes.c_uint32(), ctypes.c_void_p(), ctypes.c_size_t()
        _check(_lib.ol_query(self._ctx, _ptr(t), t.size, _ptr(o), o.size, k, dp, max_matches, ctypes.byref(tot),
                             ctypes.byref(p), ctypes.byref(n)))
        return QueryResult(tot.value, _take(p, n.value, np.uint32))

    def query_raw(self, text, kind=None, depth=None, max_matches: int = 1 << 20) -> QueryResult:
        """Upload, scan and query in one call."""
        t = as_bytes_array(text)
        k, dp = self._query_args(kind, depth)
        tot, p, n = ctypes.c_uint32(), ctypes.c_void_p(), ctypes.c_size_t()
        _check(_lib.ol_query_raw(self._ctx, _ptr(t), t.size, k, dp, max_matches, ctypes.byref(tot),
                                 ctypes.byref(p), ctypes.byref(n)))
        return QueryResult(tot.value, _take(p, n.value, np.uint32))

    # -- dictionary search --------------------------------------------------------------
    def compile_patterns(self, doc: Optional[GpuDocument], patterns: Sequence[bytes]) -> GpuPatternSet:
        arrs = [as_bytes_array(p) for p in patterns]
        keep = [np.concatenate([a, np.zeros(1, np.uint8)]) for a in arrs]   # non-null even when empty
        ptrs = (ctypes.c_void_p * max(1, len(keep)))(*[k.ctypes.data for k in keep])
        lens = (ctypes.c_size_t * max(1, len(keep)))(*[a.size for a in arrs])
        h = ctypes.c_void_p()
        _check(_lib.ol_dict_compile(self._ctx, doc._h if doc is not None else None,
                                    ctypes.cast(ptrs, ctypes.c_void_p), ctypes.cast(lens, ctypes.c_void_p),
                                    len(arrs), ctypes.byref(h)))
        return GpuPatternSet(h)

    def search_many(self, doc: GpuDocument, pset: GpuPatternSet, max_hits: int) -> ManyResult:
        counts = np.zeros(len(pset), np.uint32)
        tot = ctypes.c_uint64()
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        _check(_lib.ol_search_many(self._ctx, doc._h, pset._h, max_hits, _ptr(counts), ctypes.byref(tot),
                                   ctypes.byref(p), ctypes.byref(n)))
        return ManyResult(counts, tot.value, _take(p, n.value, MANY_HIT))

    # -- per-call profiles (the reference's profile_* methods) ---------------------------
    def calibrate_scan_floor(self, doc, lines_per_tile: int, write: bool = True, reps: int = 10) -> float:
        """Mean microseconds of a kernel moving the scan's bytes over doc's text
        with none of its work (ol_calib_scan_floor): the roofline's floor."""
        us = ctypes.c_double()
        _check(_lib.ol_calib_scan_floor(self._ctx, doc._h, lines_per_tile, 1 if write else 0, reps, ctypes.byref(us)))
        return us.value

    def profile_call(self, fn, *args, **kwargs):
        """(fn(*args), Profile) of one call through this parser."""
        _check(_lib.ol_profile_call_begin(self._ctx))
        try:
            res = fn(*args, **kwargs)
        finally:
            wall = ctypes.c_double()
            nh, ng = ctypes.c_size_t(), ctypes.c_size_t()
            host = np.zeros(64, PROFILE_STEP)
            gpu = np.zeros(64, PROFILE_STEP)
            _check(_lib.ol_profile_call_end(self._ctx, ctypes.byref(wall), _ptr(host), host.size, ctypes.byref(nh),
                                            _ptr(gpu), gpu.size, ctypes.byref(ng)))
        prof = Profile(wall.value / 1e3,
                       [(x["label"].decode(), float(x["ms"]) / 1e3) for x in host[:min(nh.value, host.size)]],
                       [(x["label"].decode(), float(x["ms"]) / 1e3) for x in gpu[:min(ng.value, gpu.size)]])
        return res, prof

    def profile_upload(self, text, nodes: bool = True):
        return self.profile_call(self.upload, text, nodes)

    def profile_query_document(self, doc, kind=None, depth=None, max_matches: int = 1 << 20) -> Profile:
        return self.profile_call(self.query_document, doc, kind, depth, max_matches)[1]

    def profile_query_raw(self, text, kind=None, depth=None, max_matches: int = 1 << 20) -> Profile:
        return self.profile_call(self.query_raw, text, kind, depth, max_matches)[1]

    def profile_search_document(self, doc, pattern: bytes, max_matches: int) -> Profile:
        return self.profile_call(self.search_document, doc, pattern, max_matches)[1]

    def profile_search_raw(self, text, pattern: bytes, max_matches: int) -> Profile:
        return self.profile_call(self.search_raw, text, pattern, max_matches)[1]

    def profile_search_many(self, doc, pset, max_hits: int) -> Profile:
        return self.profile_call(self.search_many, doc, pset, max_hits)[1]

    def profile_build_tree_reusing(self, doc, previous):
        return self.profile_call(self.build_tree_reusing, doc, previous)

    def profile_build_tree(self, doc) -> Profile:
        return self.profile_build_tree_reusing(doc, None)[1]

    # -- launch tuning ------------------------------------------------------------------
    def tune_set(self, name: str, value: int) -> None:
        _check(_lib.ol_tune_set(self._ctx, name.encode(), int(value)))

    def tune_get(self, name: str) -> int:
        v = ctypes.c_int()
        _check(_lib.ol_tune_get(self._ctx, name.encode(), ctypes.byref(v)))
        return v.value

    # -- profiling ----------------------------------------------------------------------
    def batch(self, documents: Sequence) -> GpuBatch:
        """Uploads the documents back to back (ol_batch_create)."""
        arrs = [as_bytes_array(d) for d in documents]
        ptrs = (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data if a.size else None for a in arrs])
        lens = (ctypes.c_size_t * max(1, len(arrs)))(*[a.size for a in arrs])
        h = ctypes.c_void_p()
        _check(_lib.ol_batch_create(self._ctx, ctypes.cast(ptrs, ctypes.c_void_p), ctypes.cast(lens, ctypes.c_void_p),
                                    len(arrs), ctypes.byref(h)))
        return GpuBatch(self, h, arrs)

    def profile_enable(self, on: bool = True) -> None:
        _check(_lib.ol_profile_enable(self._ctx, 1 if on else 0))

    def profile_take(self) -> dict:
        """{kernel name: (launches, total ms)} since the last call."""
        cap, stride = 64, 64
        names = ctypes.create_string_buffer(cap * stride)
        counts = np.zeros(cap, np.uint64)
        ms = np.zeros(cap, np.float64)
        n = ctypes.c_size_t()
        _check(_lib.ol_profile_take(self._ctx, names, stride, _ptr(counts), _ptr(ms), cap, ctypes.byref(n)))
        out = {}
        raw = names.raw
        for i in range(min(n.value, cap)):
            nm = raw[i * stride:(i + 1) * stride].split(b"\0", 1)[0].decode()
            out[nm] = (int(counts[i]), float(ms[i]))
        return out


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("slots", ctypes.c_uint32),
                ("max_batch_docs", ctypes.c_uint32), ("max_batch_bytes", ctypes.c_size_t)]


class ServeStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("requests", "parse_requests", "search_requests", "rejected", "bytes",
                                               "batches", "batched_requests")] + \
               [(n, ctypes.c_double) for n in ("seconds", "qps_offered", "qps_achieved", "gbps", "p50_ms", "p90_ms",
                                               "p99_ms", "max_ms", "mean_ms")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


PARSE, SEARCH = 0, 1


def _scheck(rc: int) -> None:
    if rc != 0:
        raise OnelineError(rc, _lib.ol_server_last_error().decode(errors="replace"))


class GpuServer:
    """Continuous-batching request loop (ol_server_*): parse and dictionary
    search requests batched across documents, several batches in flight."""

    def __init__(self, hbm_budget: int = 0, slots: int = 0, max_batch_docs: int = 0, max_batch_bytes: int = 0):
        load_library()
        cfg = ServerConfig(hbm_budget, slots, max_batch_docs, max_batch_bytes)
        h = ctypes.c_void_p()
        _scheck(_lib.ol_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._docs = {}      # request id -> document bytes (kept alive until completion)
        self._dict_keep = []

    def add_dictionary(self, patterns: Sequence[bytes]) -> int:
        arrs = [np.concatenate([as_bytes_array(p), np.zeros(1, np.uint8)]) for p in patterns]
        ptrs = (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
        lens = (ctypes.c_size_t * max(1, len(arrs)))(*[a.size - 1 for a in arrs])
        out = ctypes.c_uint32()
        _scheck(_lib.ol_server_add_dict(self._h, ctypes.cast(ptrs, ctypes.c_void_p),
                                        ctypes.cast(lens, ctypes.c_void_p), len(arrs), ctypes.byref(out)))
        return out.value

    def submit(self, doc, kind: int = PARSE, dictionary: int = 0, keep: bool = True) -> int:
        a = as_bytes_array(doc)
        rid = ctypes.c_uint64()
        _scheck(_lib.ol_server_submit(self._h, kind, dictionary, _ptr(a) if a.size else None, a.size,
                                      1 if keep else 0, ctypes.byref(rid)))
        self._docs[rid.value] = a
        return rid.value

    def put_document(self, doc) -> int:
        """Uploads, scans and builds the tree of `doc` once; it stays resident."""
        a = as_bytes_array(doc)
        out = ctypes.c_uint64()
        _scheck(_lib.ol_server_put_doc(self._h, _ptr(a) if a.size else None, a.size, ctypes.byref(out)))
        return out.value

    def drop_document(self, doc_id: int) -> None:
        _scheck(_lib.ol_server_drop_doc(self._h, doc_id))

    def submit_resident(self, doc_id: int, kind: int = PARSE, dictionary: int = 0, keep: bool = True) -> int:
        rid = ctypes.c_uint64()
        _scheck(_lib.ol_server_submit_doc(self._h, kind, dictionary, doc_id, 1 if keep else 0, ctypes.byref(rid)))
        return rid.value

    def arena_bytes(self) -> int:
        return int(_lib.ol_server_arena_bytes(self._h))

    def hbm_used(self) -> int:
        return int(_lib.ol_server_hbm_used(self._h))

    def slot_of(self, rid: int) -> int:
        v = ctypes.c_uint32()
        _scheck(_lib.ol_server_slot_of(self._h, rid, ctypes.byref(v)))
        return v.value

    def wait(self, rid: int, timeout: float = 120.0) -> None:
        _scheck(_lib.ol_server_wait(self._h, rid, timeout))

    def parse_result(self, rid: int):
        """(offsets, classes, TreeLinks) of a completed parse request."""
        n = ctypes.c_size_t()
        _scheck(_lib.ol_server_result_lines(self._h, rid, ctypes.byref(n), None, None, None, None, None, None))
        off = np.zeros(n.value, np.uint32)
        cls = np.zeros(n.value, np.uint8)
        par, se, fc, ns = (np.zeros(n.value, np.uint32) for _ in range(4))
        _scheck(_lib.ol_server_result_lines(self._h, rid, ctypes.byref(n), _ptr(off), _ptr(cls), _ptr(par),
                                            _ptr(se), _ptr(fc), _ptr(ns)))
        return off, cls, TreeLinks(par, se, fc, ns)

    def search_result(self, rid: int):
        """(offsets, ids) of a completed search request, (offset, id) order."""
        n = ctypes.c_size_t()
        _scheck(_lib.ol_server_result_matches(self._h, rid, ctypes.byref(n), None, None))
        off = np.zeros(n.value, np.uint32)
        ids = np.zeros(n.value, np.uint32)
        _scheck(_lib.ol_server_result_matches(self._h, rid, ctypes.byref(n), _ptr(off), _ptr(ids)))
        return off, ids

    def timing(self, rid: int):
        v = [ctypes.c_double() for _ in range(3)]
        _scheck(_lib.ol_server_timing(self._h, rid, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def release(self, rid: int) -> None:
        _scheck(_lib.ol_server_release(self._h, rid))
        self._docs.pop(rid, None)

    def load(self, docs: Sequence, qps: float, requests: int, search_fraction: float = 0.3,
             seed: int = 5) -> dict:
        """Open-loop Poisson load over `docs` (ol_server_load); latency and GB/s."""
        arrs = [as_bytes_array(d) for d in docs]
        ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        lens = (ctypes.c_size_t * len(arrs))(*[a.size for a in arrs])
        st = ServeStats()
        _scheck(_lib.ol_server_load(self._h, ctypes.cast(ptrs, ctypes.c_void_p), ctypes.cast(lens, ctypes.c_void_p),
                                    len(arrs), qps, requests, search_fraction, seed, ctypes.byref(st)))
        return st.as_dict()

    def load_serial(self, docs: Sequence, requests: int, search_fraction: float = 0.3, seed: int = 5) -> dict:
        """The same request sequence one at a time through the single-document
        entry points (ol_server_load_serial): the plain per-document path."""
        arrs = [as_bytes_array(d) for d in docs]
        ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        lens = (ctypes.c_size_t * len(arrs))(*[a.size for a in arrs])
        st = ServeStats()
        _scheck(_lib.ol_server_load_serial(self._h, ctypes.cast(ptrs, ctypes.c_void_p),
                                           ctypes.cast(lens, ctypes.c_void_p), len(arrs), requests, search_fraction,
                                           seed, ctypes.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if self._h:
            _lib.ol_server_destroy(self._h)
            self._h = None
            self._docs.clear()

    def __del__(self):
        if sys.is_finalizing():   # the runtime may be gone: the process reclaims it
            return
        try:
            self.close()
        except Exception:
            pass
