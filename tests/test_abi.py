"""The C-ABI library loads and exports every entry point include/oneline_hip.h
declares.  No compute call is made here (this suite also runs without a GPU)."""
import ctypes
import os
import re

import pytest

import oneline_amd as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if not name.endswith(".h"):
            continue
        with open(os.path.join(ROOT, "include", name)) as f:
            src = f.read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(ol_[a-z0-9_]+)\s*\(", src):
            syms.add(m.group(1))
    return syms


def test_library_exports_every_declared_symbol():
    lib = ol.load_library()
    syms = declared_symbols()
    assert len(syms) > 30
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing


def test_wrapper_binds_only_declared_symbols():
    import inspect
    src = inspect.getsource(ol)
    used = set(re.findall(r'"(ol_[a-z0-9_]+)"', src))
    assert used <= declared_symbols(), used - declared_symbols()


def test_datagen_library_exports():
    from oneline_amd import datagen
    g = datagen._lib()
    for s in ("ol_gen_document", "ol_gen_patterns", "ol_gen_vocab_word", "ol_gen_lognormal_sizes",
              "ol_gen_profile", "ol_gen_prose_vocab"):
        assert hasattr(g, s)


def test_context_creation_fails_cleanly_without_a_device():
    """Without a HIP device the engine reports an error instead of falling back."""
    lib = ol.load_library()
    n = ctypes.c_int(0)
    hip = ctypes.CDLL("libamdhip64.so")
    if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(ol.OnelineError) as e:
        ol.GpuParser()
    assert e.value.code == -5
