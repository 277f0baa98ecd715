"""The oracle and the workload generator (host C) under AddressSanitizer and
UndefinedBehaviorSanitizer: `make -C oracle asan` builds both with
-fsanitize=address,undefined, and the oracle and generator test modules run
against those builds in a child interpreter with libasan preloaded.  Any
sanitizer report fails the run (-fno-sanitize-recover, abort_on_error)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_and_generator_are_sanitizer_clean():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    asan_dir = os.path.join(ROOT, "oracle", "_asan")
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], check=True, capture_output=True,
                             text=True).stdout.strip()
    env = dict(os.environ)
    env.update({
        "OL_ORACLE_LIB": os.path.join(asan_dir, "liboneline_oracle.so"),
        "OL_DATAGEN_LIB": os.path.join(asan_dir, "libol_datagen.so"),
        "LD_PRELOAD": libasan,
        # the interpreter itself is not instrumented: its allocations are not leaks to report
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1:halt_on_error=1",
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
    })
    # the child really runs the instrumented builds
    probe = ("import ctypes; from oracle import pyoracle as o; from oneline_amd import datagen as d; "
             "o.lib(); d._lib(); ctypes.CDLL(None).__asan_init; "
             "print(o.LIB, d.GEN_PATH)")
    chk = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True,
                         timeout=120)
    assert chk.returncode == 0, chk.stderr[-2000:]
    assert chk.stdout.count("_asan") == 2, chk.stdout
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle.py"), os.path.join(ROOT, "tests", "test_datagen.py"),
                        # the 1 GiB line-count check is the same code as the 3 MiB ones, 4x slower here
                        "-k", "not reported_by_the_reference"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert " passed" in out
