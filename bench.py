"""Benchmark of the parse-and-search path on one MI355X.

Workload (the headline): one 2 GiB synthetic markup document (seed 0, 1% CRLF
lines, 0.5% empty lines, ~33M lines), resident in HBM.  One step = the full
parse + tree over it: newline scan, classification, and tree links
(parent / first_child / next_sibling), i.e. the kernels of ol_doc_reparse +
ol_tree_rebuild_async, with no host round trip inside the step.

value    = algorithmic bytes / second = (N + 17 L) / step time, N = text bytes,
           L = lines (4 B offset + 1 B class + 3 x 4 B links per line);
roofline = the dominant kernel (scan_classify): its algorithmic bytes per launch
           (N + 5 L) / its mean launch time from HIP events on the engine stream,
           against the 8 TB/s HBM peak; beside it the floor: the time of a
           kernel that only moves the scan's bytes (calib.hip);
cpu_baseline = the C oracle's multi-threaded restatement of oneline-simd's
           parallel parse (two-pass chunked) and tree (three-phase nearest
           smaller values) over the same document, on up to 16 host threads,
           median of 3 runs after a warm-up, checked against the GPU outputs.

Outputs are checked bit-exact against the oracle before timing (as the
reference's CLI does).  Extra fields report literal and dictionary search.
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(p, doc, tree, text, args):
    from oracle import pyoracle as orc
    threads = cpu_threads(args)
    n = text.size
    sample = text[: min(args.cpu_sample, n)]
    if sample.size < n:   # whole lines only
        last_nl = int(np.flatnonzero(sample == 0x0A)[-1]) + 1 if sample.size else 0
        sample = sample[:last_nl]
    L = orc.line_count(sample)
    out = (np.zeros(L, np.uint32), np.zeros(L, orc.NODE))
    lout = tuple(np.zeros(L, np.uint32) for _ in range(4))

    def run():
        t0 = time.perf_counter()
        orc.parse_par(sample, threads, out)
        orc.links_par(out[1], threads, lout)
        return time.perf_counter() - t0

    run()   # warm-up; then agreement with the GPU outputs, as the reference's CLI checks
    p.rebuild_tree_async(doc, tree)   # every link array (the timed step skips subtree_end)
    p.check_tree(tree)
    o, c, _ = doc.download()
    got = p.download_links(tree)
    assert np.array_equal(o[:L], out[0]) and np.array_equal(c[:L], out[1]["kind"].astype(np.uint8)), \
        "CPU baseline disagrees with the GPU (offsets / classes)"
    if sample.size == n:
        for g, w in zip((got.parent, got.subtree_end, got.first_child, got.next_sibling), lout):
            assert np.array_equal(g, w), "CPU baseline disagrees with the GPU (links)"
    del o, c, got
    times = sorted(run() for _ in range(3))
    cpu_s = times[1]
    cpu_gbs = (sample.size + 17 * L) / cpu_s / 1e9
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    log(f"[bench] cpu oracle ({threads} threads of {os.cpu_count()}, {cpu_model}): "
        f"{sample.size} B in {cpu_s:.3f}s = {cpu_gbs:.3f} GB/s")
    return {
        "value": round(cpu_gbs, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{'the whole' if sample.size == n else 'first ' + str(sample.size) + ' bytes of the'} "
                  f"document: parallel oracle parse (two-pass chunks) + tree (three-phase nearest smaller "
                  f"values) + first_child/next_sibling, {threads} threads, median of 3 after a warm-up",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=1 << 31, help="document bytes")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-search", action="store_true")
    ap.add_argument("--search", type=str, default="literal,dict1024,dict65536,query,bridge,olb",
                    help="search-side workloads to run (comma list of literal, dict<count>, query, bridge, olb; "
                         "dict<count>s: the same dictionary plus one 2-byte pattern that never occurs, "
                         "the cost of the 1- and 2-byte classes)")
    ap.add_argument("--cpu-sample", type=int, default=1 << 31, help="CPU baseline bytes (a prefix of the document)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0 = min(16, cores))")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-dropin", action="store_true", help="skip the FlatNode + subtree_end variant")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workload (10,000 documents)")
    ap.add_argument("--no-tune", action="store_true", help="use the default launch geometry")
    ap.add_argument("--batch-docs", type=int, default=10000)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving sweep")
    ap.add_argument("--serve-qps", type=str, default="250,1000,4000,16000,32000,64000")
    ap.add_argument("--serve-serial", type=int, default=1600,
                    help="requests of the per-document loop beside the serving sweep (0: skip it)")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--only", choices=("serve",), default=None,
                    help="run one secondary workload alone and print its record (for profiling)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="after timing, copy 1 GiB device-to-device twice (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the path's multi-GPU sharding is a design, not a build", file=sys.stderr)
        sys.exit(2)

    if args.only == "serve":
        print(json.dumps({"serve": serve_benchmark(args)}), flush=True)
        return

    import oneline_amd as ol
    from oneline_amd import datagen

    n = args.size
    log(f"[bench] generating {n} bytes (seed 0, crlf 1%, empty 0.5%)")
    t0 = time.time()
    text = datagen.document(n, seed=0, crlf=10, empty=5)
    log(f"[bench] generated in {time.time() - t0:.1f}s")

    p = ol.GpuParser()
    log(f"[bench] adapter {p.adapter_name()}")
    t0 = time.time()
    doc = p.upload(text, nodes=False)
    L = doc.line_count()
    upload_s = time.time() - t0
    tree = p.build_tree(doc)
    log(f"[bench] upload+scan {upload_s:.2f}s, lines {L}")
    # the same again with the staging ring already pinned: host -> HBM upload
    # plus the scan, as a caller with a resident context sees it
    t0 = time.time()
    doc2 = p.upload(text, nodes=False)
    upload2_s = time.time() - t0
    del doc2
    log(f"[bench] second upload+scan {upload2_s:.3f}s = {n / upload2_s / 1e9:.1f} GB/s")

    verified = False
    want = None
    if not args.no_verify:
        from oracle import pyoracle as orc
        log("[bench] verifying against the oracle (full document)")
        off, nodes = orc.parse(text)
        o, c, _ = doc.download()
        assert np.array_equal(o, off), "offsets differ"
        kinds = nodes["kind"].astype(np.uint8)
        assert np.array_equal(c, kinds), "classes differ"
        wl = orc.links(nodes)
        got = p.download_links(tree)
        for name, g, w in zip(("parent", "subtree_end", "first_child", "next_sibling"),
                              (got.parent, got.subtree_end, got.first_child, got.next_sibling), wl):
            assert np.array_equal(g, w), f"{name} differs"
        want = (off, kinds, wl[0], wl[2], wl[3])   # re-checked after the timed steps
        del nodes, o, c, got, wl
        log("[bench] verified bit-exact")

    lib = ol.load_library()

    # the step's outputs: offsets + classes (scan), parent / first_child /
    # next_sibling (tree); subtree_end is not part of the metric's bytes
    links = ol.LINK_PARENT | ol.LINK_FIRST_CHILD | ol.LINK_NEXT_SIBLING

    def step():
        ol._check(lib.ol_doc_reparse(p._ctx, doc._h))
        ol._check(lib.ol_tree_rebuild_links_async(p._ctx, doc._h, tree._h, links))

    tuning = None
    if not args.no_tune:
        from oneline_amd import autotune
        tuning = autotune.tune(p, "single", step, reps=2)
        log(f"[bench] autotune: scan_bpt {tuning['scan_bpt']}, tree_rounds {tuning['tree_rounds']}")

    for _ in range(args.warmup):
        step()
    p.synchronize()
    p.profile_take()
    p.profile_enable(True)
    p.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    p.synchronize()
    dt = time.perf_counter() - t0
    p.profile_enable(False)
    prof = p.profile_take()
    p.check_tree(tree)
    ms = dt / args.steps * 1e3
    alg_bytes = n + 17 * L
    value = alg_bytes / (dt / args.steps) / 1e9
    kname = "scan_classify"
    k_launch, k_ms = prof.get(kname, (0, 0.0))
    k_avg_ms = k_ms / max(1, k_launch)
    k_bytes = n + 5 * L
    k_gbs = k_bytes / (k_avg_ms / 1e3) / 1e9 if k_avg_ms > 0 else 0.0
    traffic = load_traffic(kname, n)
    prof_us = load_profile_mean("void scan_classify_kernel<256, 2>", n)
    kernels = {k: {"launches": c, "avg_us": round(t / max(1, c) * 1e3, 2)} for k, (c, t) in prof.items()}
    log(f"[bench] {ms:.3f} ms/step, {value:.1f} GB/s; kernels {kernels}")

    # profiled run uses events; re-time without events for the headline
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    p.synchronize()
    dt2 = time.perf_counter() - t0
    p.check_tree(tree)
    ms2 = dt2 / args.steps * 1e3
    value2 = alg_bytes / (dt2 / args.steps) / 1e9
    log(f"[bench] without events: {ms2:.3f} ms/step, {value2:.1f} GB/s")
    if want is not None:
        # the outputs of the last timed step, against the oracle
        o, c, _ = doc.download()
        L3 = tree.line_count()
        arrs = [np.zeros(L3, np.uint32) for _ in range(3)]
        ol._check(lib.ol_tree_download(p._ctx, tree._h, ol._ptr(arrs[0]), None, ol._ptr(arrs[1]), ol._ptr(arrs[2])))
        for name, g, w in zip(("offsets", "classes", "parent", "first_child", "next_sibling"),
                              (o, c, *arrs), want):
            assert np.array_equal(g, w), f"{name} of the timed step differs from the oracle"
        del o, c, arrs, want
        verified = True
        log("[bench] timed step's outputs bit-exact against the oracle")

    # the floor under the scan: a kernel that moves the scan's bytes (the text
    # in 32 KiB tiles; the document's mean lines per tile of u32 + 2 x u8
    # writes) with none of its work
    tiles32 = -(-n // 32768)
    lpt = max(1, round(L / tiles32))
    floor_r = p.calibrate_scan_floor(doc, lpt, write=False, reps=10)
    floor_rw = p.calibrate_scan_floor(doc, lpt, write=True, reps=10)
    floor = {"kernel": "stream_floor_kernel (calib.hip): the scan's reads and writes without its work",
             "lines_per_tile": lpt, "read_us": round(floor_r, 2), "read_write_us": round(floor_rw, 2),
             "read_write_gbs": round((n + 6 * lpt * tiles32) / (floor_rw / 1e6) / 1e9, 1)}
    log(f"[bench] floor: read {floor_r:.1f} us, read + write {floor_rw:.1f} us (scan {k_avg_ms * 1e3:.1f} us)")

    # the drop-in variant: the reference's outputs as it returns them -- the
    # 16-byte FlatNode per line (parse_raw / upload) and the tree with
    # subtree_end -- beside offsets, classes and the other three links
    dropin = None
    if not args.no_dropin:
        docn = p.upload(text, nodes=True)
        treen = p.build_tree(docn)

        def step_dropin():
            ol._check(lib.ol_doc_reparse(p._ctx, docn._h))
            ol._check(lib.ol_tree_rebuild_links_async(p._ctx, docn._h, treen._h, ol.LINK_ALL))

        for _ in range(args.warmup):
            step_dropin()
        p.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step_dropin()
        p.synchronize()
        p.check_tree(treen)
        dms = (time.perf_counter() - t0) / args.steps * 1e3
        db = n + 37 * L   # offsets 4 + class 1 + FlatNode 16 + four links 16, per line
        dropin = {"outputs": "offsets, classes, 16-byte FlatNodes, parent, subtree_end, first_child, next_sibling",
                  "ms_per_step": round(dms, 4), "algorithmic_bytes": db, "gbs": round(db / (dms / 1e3) / 1e9, 1),
                  "frac_of_hbm_peak": round(db / (dms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"[bench] drop-in variant (FlatNodes + subtree_end): {dms:.3f} ms/step, {dropin['gbs']} GB/s")
        treen.close()
        docn.close()

    if args.copy_calibration:
        copy_calibration(1 << 30)

    # CPU baseline: the oracle's parallel restatement over (a prefix of) the
    # same document, warm buffers, median of 3 after a warm-up run.
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p, doc, tree, text, args)

    extra = {}
    if not args.no_search:
        which = set(x for x in args.search.split(",") if x)
        extra = search_benchmarks(p, doc, text, which, args)
        if "query" in which:
            extra["query"] = query_benchmark(p, doc, tree, L)
        if "bridge" in which:
            extra["bridge"] = bridge_benchmark(p, doc, tree, text, L, args)
        if "olb" in which:
            extra["olb"] = olb_benchmark(p, text)
    tree.close()
    doc.close()
    del doc, tree, text
    if not args.no_batch:
        extra["batch"] = batch_benchmark(p, args)
    if not args.no_serve:
        extra["serve"] = serve_benchmark(args)

    out = {
        "metric": "parse & dict-search GB/s (frac of HBM peak); served p50/p99 at fixed QPS",
        "value": round(value2, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms2, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": "parse+tree of one 2 GiB synthetic document (seed 0, 1% CRLF, 0.5% empty lines)",
            "text_bytes": n,
            "lines": L,
            "algorithmic_bytes_per_step": alg_bytes,
            "frac_of_hbm_peak": round(value2 / HBM_PEAK_GBS, 4),
            "input_gbs": round(n / (dt2 / args.steps) / 1e9, 2),
            "verified_bit_exact": verified,
            "upload_s_pcie_inclusive": round(upload_s, 3),
            "upload_scan_gbs_warm": round(n / upload2_s / 1e9, 2),
            "kernels": kernels,
            "tuning": tuning and {"scan_bpt": tuning["scan_bpt"], "tree_rounds": tuning["tree_rounds"],
                                  "sweep_kernel_ms": tuning["sweep"]},
            "dropin_variant": dropin,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(k_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(k_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic.get("bytes_per_launch"),
            "traffic_source": traffic.get("source"),
            "kernel": kname,
            "kernel_avg_us": round(k_avg_ms * 1e3, 2),
            "kernel_algorithmic_bytes": k_bytes,
            "profile_kernel_avg_us": prof_us,
            "profile_frac": round(k_bytes / (prof_us / 1e6) / 1e9 / HBM_PEAK_GBS, 4) if prof_us else None,
            "profile_source": "profiles/scan_tree_2GiB_kernel_trace.txt (rocprofv3 --kernel-trace, last 20 launches)"
            if prof_us else None,
            "floor_us": floor["read_write_us"],
            "frac_of_floor": round(floor["read_write_us"] / (k_avg_ms * 1e3), 4) if k_avg_ms > 0 else None,
            "floor": floor,
        },
        "cpu_baseline": cpu,
    }
    out.update(extra)
    p.close()
    print(json.dumps(out), flush=True)


def load_traffic(kname, n):
    """HBM bytes per launch of `kname` as measured by rocprofv3 counters
    (FETCH_SIZE, WRITE_SIZE, scaled by a device copy of known size), kept in
    profiles/traffic.json; empty if absent or measured on another size."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(kname)
    except (OSError, ValueError):
        return {}
    if not rec or rec.get("text_bytes") != n:
        return {}
    return {"bytes_per_launch": rec["bytes_per_launch"], "source": f"profiles/traffic.json ({rec['profile']})"}


def load_profile_mean(kernel_prefix, n):
    """The mean duration (us) of a kernel in the committed rocprofv3 kernel
    trace of the headline (profiles/scan_tree_2GiB_kernel_trace.txt: its last
    20 launches, the timed steps), for the 2 GiB workload only."""
    if n != 1 << 31:
        return None
    path = os.path.join(ROOT, "profiles", "scan_tree_2GiB_kernel_trace.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(kernel_prefix):
                    return float(line.split()[-1])
    except (OSError, ValueError):
        return None
    return None


def copy_calibration(nbytes):
    """Two device-to-device copies of `nbytes`: a plain copy kernel of known
    size whose FETCH_SIZE / WRITE_SIZE readings calibrate the counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    for _ in range(2):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] copy calibration: 2 x {nbytes} bytes device-to-device")


def batch_benchmark(p, args):
    """Batched workload: documents with log-normal sizes (median 16 KiB, sigma
    2.0, 64 B .. 64 MiB, seed 1), parse + tree + a 4,096-pattern dictionary
    search as one launch sequence replayed as a hipGraph, against the plain
    per-document path (reparse, tree, dictionary search per document)."""
    import ctypes

    import oneline_amd as ol
    from oneline_amd import autotune, datagen
    from oracle import pyoracle as orc

    lib = ol.load_library()
    hip = ctypes.CDLL("libamdhip64.so")
    count = args.batch_docs
    sizes = datagen.lognormal_sizes(count, 1, 16384.0, 2.0, 64, 64 << 20)
    t0 = time.time()
    docs = [datagen.document(int(sz), seed=datagen.mix(1, i)) for i, sz in enumerate(sizes)]
    nbytes = int(sum(int(x) for x in sizes))
    log(f"[bench] batch: {count} documents, {nbytes} bytes, generated in {time.time() - t0:.1f}s")
    pats = datagen.patterns(4096, seed=0xC3B)
    pset = p.compile_patterns(None, pats)
    b = p.batch(docs)

    tuning = None
    if not args.no_tune:
        tuning = autotune.tune(p, "batch", lambda: b.run(pset), reps=2)
        log(f"[bench] batch autotune: scan_bpt {tuning['scan_bpt']}, tree_rounds {tuning['tree_rounds']}")
    b.run(pset)                          # sized for the chosen geometry
    total_lines, _, _ = b.lines()

    # parity on a sample of documents (the oracle, document alone)
    rng = np.random.default_rng(11)
    sample = rng.choice(count, size=min(count, 60), replace=False)
    for i in sample:
        d = docs[int(i)]
        off, cls, links = b.document(int(i))
        w_off, nodes = orc.parse(d)
        assert np.array_equal(off, w_off), f"batch doc {i}: offsets"
        assert np.array_equal(cls, nodes["kind"].astype(np.uint8)), f"batch doc {i}: classes"
        w = orc.links(nodes)
        for g, ww in zip((links.parent, links.subtree_end, links.first_child, links.next_sibling), w):
            assert np.array_equal(g, ww), f"batch doc {i}: links"
        mo, mi = b.matches(int(i))
        wo, wi = orc.many_find_all(d, pats)
        assert np.array_equal(mo, wo.astype(np.uint32)) and np.array_equal(mi, wi.astype(np.uint32)), \
            f"batch doc {i}: matches"
    log(f"[bench] batch: {len(sample)} sampled documents bit-exact against the oracle")

    # total matches (all documents)
    cnt = ctypes.c_uint64()
    matches = 0
    for i in range(count):
        ol._check(lib.ol_batch_matches(p._ctx, b._h, i, ctypes.byref(cnt), None, None, 0))
        matches += cnt.value
    alg = nbytes + 17 * total_lines + 12 * matches

    def timed(fn, reps):
        fn()
        p.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            fn()
        p.synchronize()
        return (time.perf_counter() - t) / reps

    reps = 10
    t_graph = timed(lambda: b.run(pset, graph=True), reps)
    t_eager = timed(lambda: b.run(pset, graph=False), reps)
    p.profile_take()
    p.profile_enable(True)
    b.run(pset, graph=False)
    p.synchronize()
    p.profile_enable(False)
    kern = {k: round(ms, 4) for k, (c, ms) in p.profile_take().items()}
    log(f"[bench] batch: graph {t_graph * 1e3:.3f} ms, eager {t_eager * 1e3:.3f} ms, kernels {kern}")
    b.close()

    # plain per-document path on the same documents and dictionary
    pdocs = [p.upload(d, nodes=False) for d in docs]
    ptrees = [p.build_tree(d) for d in pdocs]
    caps = []
    for d in pdocs:
        ol._check(lib.ol_search_many_device(p._ctx, d._h, pset._h, None, None, 0, ctypes.byref(cnt)))
        caps.append(cnt.value)
    bufs = []
    for c in caps:
        o, q = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(o), ctypes.c_size_t(max(1, c) * 8)) == 0
        assert hip.hipMalloc(ctypes.byref(q), ctypes.c_size_t(max(1, c) * 4)) == 0
        bufs.append((o, q))

    def per_doc():
        for d, tr, (o, q), c in zip(pdocs, ptrees, bufs, caps):
            ol._check(lib.ol_doc_reparse(p._ctx, d._h))
            ol._check(lib.ol_tree_rebuild_async(p._ctx, d._h, tr._h))
            ol._check(lib.ol_search_many_device(p._ctx, d._h, pset._h, o, q, max(1, c), ctypes.byref(cnt)))

    t_loop = timed(per_doc, 1)
    log(f"[bench] batch: per-document loop {t_loop * 1e3:.1f} ms")
    for o, q in bufs:
        hip.hipFree(o)
        hip.hipFree(q)
    for d in pdocs:
        d.close()
    for tr in ptrees:
        tr.close()
    return {
        "workload": f"{count} documents, log-normal sizes (median 16 KiB, sigma 2.0, 64 B..64 MiB, seed 1), "
                    "parse + tree + 4,096-pattern dictionary search",
        "documents": count, "text_bytes": nbytes, "lines": total_lines, "matches": matches,
        "algorithmic_bytes": alg,
        "graph_ms": round(t_graph * 1e3, 4), "graph_gbs": round(alg / t_graph / 1e9, 2),
        "eager_ms": round(t_eager * 1e3, 4), "eager_gbs": round(alg / t_eager / 1e9, 2),
        "per_document_loop_ms": round(t_loop * 1e3, 3), "per_document_loop_gbs": round(alg / t_loop / 1e9, 2),
        "speedup_vs_per_document": round(t_loop / t_graph, 2),
        "kernel_ms": kern,
        "tuning": tuning and {"scan_bpt": tuning["scan_bpt"], "tree_rounds": tuning["tree_rounds"],
                              "sweep_kernel_ms": tuning["sweep"]},
        "verified_sample": int(len(sample)),
    }


def serve_benchmark(args):
    """Serving: open-loop Poisson arrivals at each QPS of a sweep, 70% parse and
    30% dictionary-search requests; document sizes log-normal (median 64 KiB,
    sigma 2.0, 64 B .. 256 MiB, seed 2); each search uses one of 8 resident
    dictionaries of 1,024 .. 65,536 patterns.  p50/p99 latency (completion -
    scheduled arrival) and served GB/s ((document + result bytes) / wall).
    Beside it, the same request sequence run one request at a time through the
    single-document entry points (the plain per-document path), and the served
    capacity's ratio to it."""
    import oneline_amd as ol
    from oneline_amd import datagen

    pool = 400
    sizes = datagen.lognormal_sizes(pool, 2, 65536.0, 2.0, 64, 256 << 20)
    docs = [datagen.document(int(sz), seed=datagen.mix(2, i)) for i, sz in enumerate(sizes)]
    log(f"[bench] serve: {pool} documents, {int(sum(int(x) for x in sizes))} bytes in the pool")
    slots, budget = 4, 96 << 30
    srv = ol.GpuServer(hbm_budget=budget, slots=slots, max_batch_docs=256, max_batch_bytes=256 << 20)
    dict_sizes = (1024, 2048, 4096, 8192, 16384, 32768, 65536, 1024)
    for k, cnt in enumerate(dict_sizes):
        srv.add_dictionary(datagen.patterns(cnt, seed=0xC3B + 17 * k))
    srv.load(docs[:32], qps=500.0, requests=200, search_fraction=0.3, seed=1)   # warm up
    sweep = []
    for q in [float(x) for x in args.serve_qps.split(",")]:
        n = max(200, int(q * args.serve_seconds))
        st = srv.load(docs, qps=q, requests=n, search_fraction=0.3, seed=int(q))
        rec = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in st.items()}
        sweep.append(rec)
        log(f"[bench] serve qps {q:.0f}: achieved {st['qps_achieved']:.0f}, p50 {st['p50_ms']:.2f} ms, "
            f"p99 {st['p99_ms']:.2f} ms, {st['gbps']:.2f} GB/s, batches {st['batches']}")
    # the plain per-document path on the same kind of request sequence
    serial = None
    if args.serve_serial:
        serial = srv.load_serial(docs, requests=args.serve_serial, search_fraction=0.3, seed=7)
        serial = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in serial.items()}
        log(f"[bench] serve: per-document loop {serial['qps_achieved']:.0f} requests/s, {serial['gbps']:.2f} GB/s")
    best = max(sweep, key=lambda r: r["gbps"])
    arena = srv.arena_bytes()
    srv.close()
    return {
        "workload": "open-loop Poisson arrivals, 70% parse / 30% dictionary search, sizes log-normal (median "
                    "64 KiB, sigma 2.0, max 256 MiB, seed 2), 8 resident dictionaries of 1k..64k patterns",
        "slots": slots, "hbm_budget_bytes": budget, "arena_bytes_per_slot": arena, "sweep": sweep,
        "per_document_loop": serial,
        "capacity_gbps": best["gbps"], "capacity_qps": best["qps_achieved"],
        "speedup_vs_per_document": round(best["gbps"] / serial["gbps"], 2) if serial and serial["gbps"] else None,
    }


def query_benchmark(p, doc, tree, L):
    """Kind / depth queries over the resident document and one of its
    subtrees (the next step after the path: the tree's callers)."""
    res = []
    for kind, depth in ((4, None), (None, 3), (2, 5)):
        p.query_document(doc, kind, depth, max_matches=1 << 26)   # warm
        p.profile_enable(True)
        t0 = time.perf_counter()
        q = p.query_document(doc, kind, depth, max_matches=1 << 26)
        dt = time.perf_counter() - t0
        p.profile_enable(False)
        c, ms = p.profile_take().get("query", (1, 0.0))
        kms = ms / max(1, c)
        # bytes: the class bytes (a kind asked) and the key bytes (a depth asked),
        # each read by the count and the write pass, + 4 B per match written
        b = 2 * L * ((kind is not None) + (depth is not None)) + 4 * q.total_matches
        res.append({"kind": kind, "depth": depth, "matches": q.total_matches, "call_ms": round(dt * 1e3, 3),
                    "kernel_ms": round(kms, 4), "kernel_gbs": round(b / (kms / 1e3) / 1e9, 1)})
        log(f"[bench] query {res[-1]}")
    return res


def bridge_benchmark(p, doc, tree, text, L, args):
    """The AST bridge's structural pass over the resident document and its
    tree (the tree's consumer, crates/oneline-tomet/src/lib.rs:20-90), checked
    against the oracle, then timed with HIP events; the oracle's sequential
    restatement (one thread) is its CPU baseline."""
    import oneline_amd as ol
    from oracle import pyoracle as orc
    flags, cont = p.bridge(doc, tree)
    res = {"lines": L, "blocks": int(np.count_nonzero(flags & ol.BRIDGE_BLOCK)),
           "literal": int(np.count_nonzero(flags & ol.BRIDGE_LITERAL)),
           "cut": int(np.count_nonzero(flags & ol.BRIDGE_CUT))}
    if not args.no_verify or not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        _, nodes = orc.parse_par(text, threads)
        parent = orc.links_par(nodes, threads)[0]
        t0 = time.perf_counter()
        w_flags, w_cont = orc.bridge(text, nodes, parent)
        cpu_s = time.perf_counter() - t0
        res["verified_bit_exact"] = bool(np.array_equal(flags, w_flags) and np.array_equal(cont, w_cont))
        if not res["verified_bit_exact"]:
            raise SystemExit("bridge: the GPU differs from the oracle")
        res["cpu_ms"] = round(cpu_s * 1e3, 1)
        res["cpu_threads"] = 1
        del nodes, parent, w_flags, w_cont
    del flags, cont
    p.profile_enable(True)
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        p.bridge(doc, tree)
    dt = (time.perf_counter() - t0) / reps
    p.profile_enable(False)
    c, ms = p.profile_take().get("bridge", (1, 0.0))
    kms = ms / max(1, c)
    # bytes: classes, keys and parent read, flags and container written (11 per line)
    b = 11 * L
    res.update({"call_ms": round(dt * 1e3, 3), "kernel_ms": round(kms, 4), "kernel_gbs": round(b / (kms / 1e3) / 1e9, 1),
                "algorithmic_bytes": b})
    log(f"[bench] bridge {res}")
    return res


def olb_benchmark(p, text, size=256 << 20):
    """A parsed document saved as .olb and loaded back (map, upload text,
    offsets and nodes, no scan) against uploading and scanning its text."""
    import tempfile
    sample = text[:size]
    path = os.path.join(tempfile.gettempdir(), f"oneline_bench_{os.getpid()}.olb")
    try:
        t0 = time.perf_counter()
        d0 = p.upload(sample, nodes=True)
        p.synchronize()
        up_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        p.save(d0, path)
        save_s = time.perf_counter() - t0
        p.load(path, verify=False).close()   # warm: page cache, first-use allocations
        t0 = time.perf_counter()
        d1 = p.load(path, verify=False)
        p.synchronize()
        load_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        p.load(path, verify=True).close()
        p.synchronize()
        verify_s = time.perf_counter() - t0
        ok = d1.line_count() == d0.line_count()
        res = {"bytes": int(sample.size), "file_bytes": os.path.getsize(path), "upload_scan_ms": round(up_s * 1e3, 2),
               "save_ms": round(save_s * 1e3, 2), "load_ms": round(load_s * 1e3, 2),
               "load_crc_ms": round(verify_s * 1e3, 2), "lines_equal": ok}
    finally:
        if os.path.exists(path):
            os.remove(path)
    log(f"[bench] olb {res}")
    return res


def cpu_threads(args):
    return args.cpu_threads or min(16, os.cpu_count() or 1)


def search_benchmarks(p, doc, text, which, args):
    """Literal and dictionary search over the resident document (device
    output).  Beside each, the oracle's restatement of the reference's CPU
    search on the host cores (chunked memmem; the overlapping Aho-Corasick DFA),
    checked against the GPU's matches before it is timed."""
    import ctypes

    import oneline_amd as ol
    from oneline_amd import datagen
    from oracle import pyoracle as orc

    threads = cpu_threads(args)
    lib = ol.load_library()
    hip = ctypes.CDLL("libamdhip64.so")
    n = text.size
    res = {}

    def dalloc(nbytes):
        ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(max(16, nbytes))) == 0
        return ptr

    # literal (C3a): lengths 4..64, vocabulary phrases (kind 0) and random lowercase (kind 1)
    lits = []
    for kind in (0, 1) if "literal" in which else ():
        for ln in (4, 8, 16, 32, 64):
            pat = datagen.literal(ln, kind=kind, seed=0xC3A)
            tot = ctypes.c_uint64()
            ol._check(lib.ol_search_device(p._ctx, doc._h, pat, len(pat), None, 0, ctypes.byref(tot)))
            cap = tot.value
            d_off = dalloc(cap * 8)
            reps = 3
            p.profile_enable(True)
            t0 = time.perf_counter()
            for _ in range(reps):
                ol._check(lib.ol_search_device(p._ctx, doc._h, pat, len(pat), d_off, cap, ctypes.byref(tot)))
            dt = (time.perf_counter() - t0) / reps
            p.profile_enable(False)
            prof = p.profile_take()
            k = prof.get("search", (1, 0.0))
            kms = k[1] / max(1, k[0])
            b = n + 12 * tot.value   # text + (u64 offset, u32 id) per match
            rec = {"pattern": pat.decode(), "kind": ("vocabulary", "random")[kind], "matches": tot.value,
                   "call_ms": round(dt * 1e3, 3), "kernel_ms": round(kms, 4),
                   "kernel_gbs": round(b / (kms / 1e3) / 1e9, 1),
                   "frac_of_hbm_peak": round(b / (kms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
            if not args.no_cpu:
                # the reference's find_all_par, restated: agreement, then the median of 3
                got = np.zeros(cap, np.uint64)
                if cap:
                    assert hip.hipMemcpy(got.ctypes.data_as(ctypes.c_void_p), d_off, ctypes.c_size_t(cap * 8), 2) == 0
                assert np.array_equal(orc.find_all_par(text, pat, threads), got), f"CPU literal search disagrees ({pat})"
                ts = []
                for _ in range(3):
                    t1 = time.perf_counter()
                    orc.find_all_par(text, pat, threads)
                    ts.append(time.perf_counter() - t1)
                cms = sorted(ts)[1] * 1e3
                rec.update({"cpu_ms": round(cms, 2), "cpu_gbs": round(n / (cms / 1e3) / 1e9, 2), "cpu_threads": threads})
            lits.append(rec)
            hip.hipFree(d_off)
            log(f"[bench] literal {lits[-1]}")
    if lits:
        res["literal_search"] = lits

    tot = ctypes.c_uint64()
    dicts = sorted((int(x[4:].rstrip("s")), x.endswith("s")) for x in which if x.startswith("dict"))
    for count, short in dicts:
        pats = datagen.patterns(count, seed=0xC3B) + ([b"\x01\x02"] if short else [])
        pset = p.compile_patterns(doc, pats)
        ol._check(lib.ol_search_many_device(p._ctx, doc._h, pset._h, None, None, 0, ctypes.byref(tot)))
        cap = tot.value
        d_off = dalloc(cap * 8)
        d_id = dalloc(cap * 4)
        p.profile_enable(True)
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):
            ol._check(lib.ol_search_many_device(p._ctx, doc._h, pset._h, d_off, d_id, cap, ctypes.byref(tot)))
        dt = (time.perf_counter() - t0) / reps
        p.profile_enable(False)
        prof = p.profile_take()
        k = prof.get("many", (1, 0.0))
        kms = k[1] / max(1, k[0])
        b = n + 12 * tot.value + sum(len(x) for x in pats)
        rec = {"matches": tot.value, "call_ms": round(dt * 1e3, 3),
               "kernel_ms": round(kms, 4), "kernel_gbs": round(b / (kms / 1e3) / 1e9, 1),
               "frac_of_hbm_peak": round(b / (kms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
        if not args.no_cpu:
            # the reference's ManyPatterns (an Aho-Corasick DFA, overlapping),
            # restated: the total over the whole text and every (start, id) of
            # a 4 MiB window against the GPU's list, then the median of 3
            # counting runs over the whole text (the reference's count_par)
            ac = orc.AhoCorasick(pats)
            assert ac.count(text, threads) == tot.value, f"CPU dictionary search disagrees ({count} patterns)"
            w0, w1 = n // 3, n // 3 + (4 << 20)
            got_off = np.zeros(cap, np.uint64)
            assert hip.hipMemcpy(got_off.ctypes.data_as(ctypes.c_void_p), d_off, ctypes.c_size_t(cap * 8), 2) == 0
            lo, hi = int(np.searchsorted(got_off, w0)), int(np.searchsorted(got_off, w1))
            got_id = np.zeros(hi - lo, np.uint32)
            if hi > lo:
                assert hip.hipMemcpy(got_id.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(d_id.value + 4 * lo),
                                     ctypes.c_size_t((hi - lo) * 4), 2) == 0
            wo, wi = ac.find(text[w0:w1 + 63], threads)
            keep = wo < (w1 - w0)
            assert np.array_equal(got_off[lo:hi] - w0, wo[keep]) and np.array_equal(got_id, wi[keep]), \
                f"CPU dictionary search disagrees in a window ({count} patterns)"
            del got_off
            ts = []
            for _ in range(3):
                t1 = time.perf_counter()
                ac.count(text, threads)
                ts.append(time.perf_counter() - t1)
            ac.close()
            cms = sorted(ts)[1] * 1e3
            rec.update({"cpu_ms": round(cms, 2), "cpu_gbs": round(n / (cms / 1e3) / 1e9, 2), "cpu_threads": threads,
                        "cpu_sample": "the whole document, every overlapping (start, pattern) counted"})
        name = f"dict_search_{count}" + ("_short" if short else "")
        if short:
            rec["short_pattern"] = "one 2-byte pattern (0x01 0x02) that never occurs, added to the set"
        res[name] = rec
        hip.hipFree(d_off)
        hip.hipFree(d_id)
        pset.close()
        log(f"[bench] {name}: {res[name]}")
    return res


if __name__ == "__main__":
    main()
