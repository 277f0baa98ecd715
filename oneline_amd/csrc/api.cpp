// api.cpp -- the C ABI (include/oneline_hip.h): host orchestration of the HIP
// kernels.  Mirrors GpuParser and its resident types (crates/oneline-gpu/src/*.rs):
// one HIP stream per context, device buffers owned by the handles, scratch
// grown on demand and reused across calls, errors returned as status codes.
#include "internal.h"
#include "../../include/oneline_hip.h"

#include <algorithm>
#include <cstdio>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(OL_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
uint64_t round_up(uint64_t a, uint64_t b) { return ceil_div(a, b) * b; }

// A device buffer, owned (grown with hipMalloc on demand) or borrowed (a
// region of a serving slot's arena: never reallocated or freed).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    hipError_t ensure(size_t b)
    {
        if (b <= bytes && p)
            return hipSuccess;
        if (borrowed)
            return hipErrorOutOfMemory;   // the region was carved too small
        release();
        hipError_t e = hipMalloc(&p, b ? b : 16);
        if (e == hipSuccess)
            bytes = b;
        else
            p = nullptr;
        return e;
    }
    void lend(void *q, size_t b)
    {
        release();
        p = q;
        bytes = b;
        borrowed = true;
    }
    void release()
    {
        if (p && !borrowed)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        borrowed = false;
    }
};

}  // namespace

// Device scratch of one owner: a context, or a batch (whose launch sequence
// may be captured as a graph, which keeps the addresses).
//   head     never reallocated: u64 meta[8] (results the host reads back),
//            u32 ctr[16] (ctr[CTR_TREE_ERR]: the tree's depth flag,
//            ctr[CTR_DEV_ERR]: the sticky device error bits, OLE_*);
//   status   u64 words, grown on demand: word 0 the ticket counter of a
//            ticket-ordered scan, the per-tile words from STATUS_TILES on;
//   tree_sum the tree's per-tile level states; cands the dictionary search's
//            candidate store.
// `dirty` counts the leading status words that may be non-zero: cooperative
// scans and literal searches need their words zero and leave them zero; every
// other user (ticket-ordered scans, dictionary searches, queries, a graph
// replay) marks its range dirty, and the next launch that needs zeros clears
// it first.
static constexpr size_t HEAD_META = 0, HEAD_CTR = 8, HEAD_WORDS = 16;
static constexpr uint32_t CTR_TREE_ERR = 1, CTR_DEV_ERR = 4;
static constexpr size_t STATUS_TILES = 8;

struct Scratch {
    DevBuf head, status, tree_sum, cands;
    size_t dirty = 0;
    bool frozen = false;   // a captured graph holds the addresses: growing is an error
    uint64_t *meta() { return (uint64_t *)head.p + HEAD_META; }
    uint32_t *ctr() { return (uint32_t *)((uint64_t *)head.p + HEAD_CTR); }
    uint64_t *words() { return (uint64_t *)status.p; }
    size_t nwords() const { return status.bytes / 8; }
    void release()
    {
        head.release();
        status.release();
        tree_sum.release();
        cands.release();
        dirty = 0;
        frozen = false;
    }
};

struct ol_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string name;
    std::mutex mu;
    Scratch scr;
    uint64_t *noalloc = nullptr;           // this device's static scratch (ol_search_device_noalloc)
    // launch geometry picked by the autotuner (ol_tune_set)
    uint32_t scan_bpt = SCAN_BPT;          // bytes per lane of scan_classify (tile = 256 lanes x bpt)
    uint32_t scan_grid = 0;                // persistent workgroups of scan_classify (0 = all co-resident)
    uint32_t scan_max_grid[3] = {0, 0, 0}; // co-resident limit per bpt (64, 128, 256), from the occupancy API
    uint32_t tree_rounds = 16;             // rounds of 64 lines per tree wave (tile = 256 x rounds lines)
    bool scan_ticket = false;              // scan tiles from an atomic ticket (ordinary launch)
    // pinned staging ring for host -> HBM uploads: two slots per copy worker
    // (each allocated on first use), each with the event of its last DMA
    uint8_t *ring[2 * 8] = {};
    hipEvent_t ring_ev[2 * 8] = {};
    bool ring_used[2 * 8] = {};
    // profiling: per-kernel totals (ol_profile_enable / _take), and the
    // per-call record (ol_profile_call_begin / _end): host step marks and
    // the call's GPU passes in launch order
    bool profiling = false;
    bool call_prof = false;
    struct Rec {
        const char *name;
        hipEvent_t a, b;
    };
    std::vector<Rec> recs, call_recs;
    std::vector<std::pair<const char *, std::chrono::steady_clock::time_point>> marks;
    std::vector<hipEvent_t> pool;
};

// A host step of the per-call profile (the reference's StageTimer::mark,
// crates/oneline-gpu/src/timing.rs:24-28): the time since the previous mark
// is attributed to `label`.
static void mark(ol_ctx *c, const char *label)
{
    if (c->call_prof)
        c->marks.emplace_back(label, std::chrono::steady_clock::now());
}

struct ol_doc {
    uint64_t n = 0, lines = 0, cap = 0;
    bool with_nodes = true;
    uint8_t *text = nullptr;
    uint32_t *offsets = nullptr;
    uint8_t *classes = nullptr, *keys = nullptr;
    ol_node_dev *nodes = nullptr;
    uint32_t byte_freq[256];
    // literal search: a text sample (4 windows of 16 KiB spread over the
    // text) and, built at the first search, its 4-byte windows' counts
    // (sorted keys, counts)
    std::vector<uint8_t> sample;
    mutable std::mutex win_mu;
    mutable std::vector<uint32_t> win_key, win_count;
};

struct ol_tree {
    uint64_t lines = 0, cap = 0;
    uint32_t links = 0;   // OL_LINK_* arrays holding this document's links
    uint32_t *parent = nullptr, *fc = nullptr, *ns = nullptr, *se = nullptr;
};

struct ol_dict {
    uint32_t count = 0;
    DictDev d{};
    void *dev = nullptr;   // one allocation holding filter, tables, entries, pool
    size_t bytes = 0;      // its size
};

namespace {

hipEvent_t take_event(ol_ctx *c)
{
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

// Brackets one launch with events when profiling is on.
struct Prof {
    ol_ctx *c;
    const char *name;
    hipEvent_t a = nullptr;
    Prof(ol_ctx *c_, const char *n) : c(c_), name(n)
    {
        if (c->profiling || c->call_prof) {
            a = take_event(c);
            if (a)
                (void)hipEventRecord(a, c->stream);
        }
    }
    ~Prof()
    {
        if (a) {
            hipEvent_t b = take_event(c);
            if (b) {
                (void)hipEventRecord(b, c->stream);
                (c->call_prof ? c->call_recs : c->recs).push_back({name, a, b});
            } else {
                c->pool.push_back(a);
            }
        }
    }
};

// The head of a scratch (zeroed once, synchronously).
hipError_t scratch_init(Scratch &s)
{
    hipError_t e = s.head.ensure(HEAD_WORDS * 8);
    if (e == hipSuccess)
        e = hipMemset(s.head.p, 0, HEAD_WORDS * 8);
    return e;
}

// At least `words` status words; new memory counts as dirty.
hipError_t ensure_status(Scratch &s, size_t words)
{
    if (s.nwords() >= words && s.status.p)
        return hipSuccess;
    if (s.frozen)
        return hipErrorInvalidValue;   // a captured graph would point at freed memory
    const hipError_t e = s.status.ensure(words * 8);
    if (e == hipSuccess)
        s.dirty = s.nwords();
    return e;
}

// Clears status words [0, words) now (a launch on the stream).
hipError_t clear_status(ol_ctx *c, Scratch &s, size_t words)
{
    Prof p(c, "clear");
    return ol_launch_clear_u64(s.words(), words, c->stream);
}

// Makes every status word zero if some may not be.
hipError_t zero_status(ol_ctx *c, Scratch &s)
{
    if (!s.dirty)
        return hipSuccess;
    const hipError_t e = clear_status(c, s, s.dirty);
    if (e == hipSuccess)
        s.dirty = 0;
    return e;
}

void soil(Scratch &s, size_t words) { s.dirty = std::max(s.dirty, std::min(words, s.nwords())); }

std::string dev_error_text(uint32_t bits)
{
    std::string m;
    auto add = [&](const char *t) { m += (m.empty() ? "" : "; ") + std::string(t); };
    if (bits & OLE_PREFIX_WAIT)
        add("a tile of a cooperative launch gave up waiting for its prefix (the prefix scanner did not reach it)");
    if (bits & OLE_SCANNER)
        add("the prefix scanner of a cooperative launch gave up waiting for a tile count");
    if (bits & OLE_LOOKBACK)
        add("a decoupled look-back of a ticket-ordered launch gave up waiting for a predecessor tile's status");
    if (bits & OLE_DICT_SCAN)
        add("the dictionary search's tile-count scan gave up waiting for an earlier chunk");
    return m.empty() ? "device error" : m;
}

// Waits for the stream; reports (and then clears) the sticky device error
// bits of `s`, and copies meta[0] to *out when asked.
int sync_check(ol_ctx *c, Scratch &s, uint64_t *out = nullptr)
{
    uint64_t head[HEAD_WORDS];
    HIPCHK(hipMemcpyAsync(head, s.head.p, sizeof head, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t bits = ((const uint32_t *)(head + HEAD_CTR))[CTR_DEV_ERR];
    if (bits) {
        // the failed launch may have left status words set
        s.dirty = s.nwords();
        HIPCHK(hipMemsetAsync(s.ctr() + CTR_DEV_ERR, 0, 4, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return fail(OL_ERR_DEVICE, dev_error_text(bits));
    }
    if (out)
        *out = head[HEAD_META];
    return OL_OK;
}

int read_meta(ol_ctx *c, Scratch &s, uint64_t *out) { return sync_check(c, s, out); }

void free_doc_lines(ol_doc *d)
{
    if (d->offsets) (void)hipFree(d->offsets);
    if (d->classes) (void)hipFree(d->classes);
    if (d->keys) (void)hipFree(d->keys);
    if (d->nodes) (void)hipFree(d->nodes);
    d->offsets = nullptr;
    d->classes = d->keys = nullptr;
    d->nodes = nullptr;
    d->cap = 0;
}

int alloc_doc_lines(ol_doc *d, uint64_t cap)
{
    free_doc_lines(d);
    uint64_t capa = round_up(cap ? cap : 1, LINE_ALIGN);
    HIPCHK(hipMalloc(&d->offsets, capa * 4));
    HIPCHK(hipMalloc(&d->classes, capa));
    HIPCHK(hipMalloc(&d->keys, capa));
    if (d->with_nodes)
        HIPCHK(hipMalloc(&d->nodes, capa * sizeof(ol_node_dev)));
    d->cap = capa;
    return OL_OK;
}

// Scan + classify of a text into the per-line arrays set in `a` (text, n,
// cap, outputs and, for batches, the document fields); fills in the rest.
// Status words of a scan over n bytes in either tile order.
size_t scan_status_words(ol_ctx *c, uint64_t n, bool coop)
{
    const uint64_t tiles = ceil_div(n, (uint64_t)SCAN_THREADS * c->scan_bpt);
    return STATUS_TILES + tiles * (coop ? SCAN_STATUS_WORDS : LB_STRIDE);
}

int launch_scan_args(ol_ctx *c, Scratch &s, ScanArgs &a, bool cooperative = true)
{
    uint32_t tiles = (uint32_t)ceil_div(a.n, (uint64_t)SCAN_THREADS * c->scan_bpt);
    // tiles from a ticket unless the cooperative launch guarantees residency
    const bool coop = cooperative && !c->scan_ticket;
    const size_t words = scan_status_words(c, a.n, coop);
    HIPCHK(ensure_status(s, words));
    if (coop) {
        HIPCHK(zero_status(c, s));   // the launch leaves its words zero again
    } else {
        HIPCHK(clear_status(c, s, words));   // ticket counter and look-back words
        soil(s, words);
    }
    a.num_tiles = tiles;
    a.status = s.words() + STATUS_TILES;
    a.meta = s.meta();
    a.lb_err = s.ctr() + CTR_DEV_ERR;
    const int vi = c->scan_bpt == 64 ? 0 : (c->scan_bpt == 128 ? 1 : 2);
    if (!c->scan_max_grid[vi])
        c->scan_max_grid[vi] = ol_scan_max_grid(c->scan_bpt);
    if (!c->scan_max_grid[vi])
        return fail(OL_ERR_DEVICE, "occupancy query for the scan kernel failed");
    // cooperative: workgroups taking tiles plus the one that turns tile counts
    // into prefixes; ticket order: workgroups taking tiles
    uint32_t grid = c->scan_max_grid[vi];
    if (c->scan_grid && c->scan_grid < grid)
        grid = c->scan_grid;
    if (grid > tiles + (coop ? 1 : 0))
        grid = tiles + (coop ? 1 : 0);
    if (coop && grid < 2)
        return fail(OL_ERR_DEVICE, "the scan needs two resident workgroups");
    a.tile_ctr = coop ? nullptr : (uint32_t *)s.words();
    Prof p(c, "scan_classify");
    HIPCHK(ol_launch_scan(a, c->scan_bpt, grid, coop, c->stream));
    return OL_OK;
}

int launch_scan(ol_ctx *c, ol_doc *d)
{
    ScanArgs a{};
    a.text = d->text;
    a.n = (uint32_t)d->n;
    a.cap = d->cap;
    a.offsets = d->offsets;
    a.classes = d->classes;
    a.keys = d->keys;
    a.nodes = d->nodes;
    return launch_scan_args(c, c->scr, a);
}

// Host -> HBM upload through the context's pinned staging ring: STAGE_WORKERS
// threads each take every STAGE_WORKERS-th chunk of STAGE_SLOT bytes, fill one
// of their two pinned slots (`fill` writes bytes [off, off + len) of the
// source), and queue its DMA on the context stream; a slot is filled again
// once the event after its last DMA has completed, so host copies overlap the
// DMA of earlier chunks.  Kernels queued on the stream afterwards see the
// whole text.  Sources under STAGE_MIN bytes go through one slot, in order.
constexpr size_t STAGE_SLOT = 8u << 20;
constexpr int STAGE_WORKERS = 8;   // the context holds 2 * 8 slots
constexpr size_t STAGE_MIN = 32u << 20;

int stage_upload(ol_ctx *c, uint8_t *dst, uint64_t n, const std::function<void(uint8_t *, uint64_t, uint64_t)> &fill)
{
    if (!n)
        return OL_OK;
    const uint64_t chunks = (n + STAGE_SLOT - 1) / STAGE_SLOT;
    const int workers = n < STAGE_MIN ? 1 : (int)std::min<uint64_t>(STAGE_WORKERS, chunks);
    for (int i = 0; i < 2 * workers; i++) {
        if (!c->ring[i] && hipHostMalloc((void **)&c->ring[i], STAGE_SLOT, hipHostMallocDefault) != hipSuccess) {
            c->ring[i] = nullptr;
            return fail(OL_ERR_NOMEM, "pinned staging allocation failed");
        }
        if (!c->ring_ev[i] && hipEventCreateWithFlags(&c->ring_ev[i], hipEventDisableTiming) != hipSuccess) {
            c->ring_ev[i] = nullptr;
            return fail(OL_ERR_DEVICE, "staging event creation failed");
        }
    }
    std::atomic<bool> failed{false};
    auto work = [&](int w) {
        int k = 0;
        for (uint64_t ch = (uint64_t)w; ch < chunks && !failed.load(std::memory_order_relaxed);
             ch += (uint64_t)workers, k ^= 1) {
            const int slot = 2 * w + k;
            uint8_t *buf = c->ring[slot];
            if (c->ring_used[slot] && hipEventSynchronize(c->ring_ev[slot]) != hipSuccess) {
                failed = true;
                return;
            }
            const uint64_t off = ch * STAGE_SLOT, len = std::min<uint64_t>(STAGE_SLOT, n - off);
            fill(buf, off, len);
            if (hipMemcpyAsync(dst + off, buf, len, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                hipEventRecord(c->ring_ev[slot], c->stream) != hipSuccess) {
                failed = true;
                return;
            }
            c->ring_used[slot] = true;
        }
    };
    if (workers == 1) {
        work(0);
    } else {
        std::vector<std::thread> pool;
        pool.reserve(workers);
        try {
            for (int w = 0; w < workers; w++)
                pool.emplace_back(work, w);
        } catch (...) {
            failed = true;
        }
        for (auto &t : pool)
            t.join();
    }
    return failed ? fail(OL_ERR_DEVICE, "text upload failed") : OL_OK;
}

// A text sample for the literal search's window counts: the whole text up
// to 64 KiB, else 4 windows of 16 KiB spread over it (window edges cut no
// real window in two, since each is counted on its own).
void sample_text(const uint8_t *t, size_t n, std::vector<uint8_t> &out)
{
    const size_t SAMPLE = 64u << 10, WINDOWS = 4;
    out.clear();
    if (n <= SAMPLE) {
        out.assign(t, t + n);
        return;
    }
    const size_t w = SAMPLE / WINDOWS, span = n - w;
    for (size_t k = 0; k < WINDOWS; k++) {
        const size_t s = span / (WINDOWS - 1) * k;
        out.insert(out.end(), t + s, t + s + w);
    }
}

void sample_histogram(const uint8_t *t, size_t n, uint32_t *freq)
{
    // sample_byte_histogram, crates/oneline-gpu/src/search.rs:176-201
    std::memset(freq, 0, 256 * sizeof(uint32_t));
    const size_t SAMPLE = 1u << 20, WINDOWS = 4;
    if (n <= SAMPLE) {
        for (size_t i = 0; i < n; i++)
            freq[t[i]]++;
    } else {
        size_t w = SAMPLE / WINDOWS, span = n - w;
        for (size_t k = 0; k < WINDOWS; k++) {
            size_t s = span / (WINDOWS - 1) * k;
            for (size_t i = 0; i < w; i++)
                freq[t[s + i]]++;
        }
    }
}

int upload_doc(ol_ctx *c, const uint8_t *text, size_t len, bool with_nodes, ol_doc **out)
{
    if (!out || (!text && len))
        return fail(OL_ERR_INVALID, "null argument");
    if (len > OL_MAX_GPU_INPUT_LEN)
        return fail(OL_ERR_TOO_LARGE, "input is " + std::to_string(len) +
                                          " bytes; the GPU path addresses at most " +
                                          std::to_string(OL_MAX_GPU_INPUT_LEN) +
                                          " bytes (32-bit offsets)");
    ol_doc *d = new (std::nothrow) ol_doc();
    if (!d)
        return fail(OL_ERR_NOMEM, "out of host memory");
    d->n = len;
    d->with_nodes = with_nodes;
    std::memset(d->byte_freq, 0, sizeof d->byte_freq);
    if (len == 0) {
        *out = d;
        return OL_OK;
    }
    auto bail = [&](int rc) {
        free_doc_lines(d);
        if (d->text) (void)hipFree(d->text);
        delete d;
        return rc;
    };
    uint64_t padded = round_up(len, TEXT_ALIGN) + TEXT_SLACK;
    if (hipMalloc(&d->text, padded) != hipSuccess) {
        d->text = nullptr;
        return bail(fail(OL_ERR_NOMEM, "device allocation of the text failed"));
    }
    mark(c, "upload: create device buffer");
    int up = stage_upload(c, d->text, len, [text](uint8_t *buf, uint64_t off, uint64_t k) {
        std::memcpy(buf, text + off, k);
    });
    if (up)
        return bail(up);
    mark(c, "upload: stream through the staging ring");
    if (hipMemsetAsync(d->text + len, 0, padded - len, c->stream) != hipSuccess)
        return bail(fail(OL_ERR_DEVICE, "text upload failed"));
    uint64_t cap = std::min<uint64_t>(len, len / 16 + 4096);
    int rc = alloc_doc_lines(d, cap);
    if (rc)
        return bail(rc);
    for (int attempt = 0; attempt < 2; attempt++) {
        rc = launch_scan(c, d);
        if (rc)
            return bail(rc);
        mark(c, "scan: launch");
        uint64_t lines = 0;
        rc = read_meta(c, c->scr, &lines);
        if (rc)
            return bail(rc);
        mark(c, "scan: wait for the GPU + read the line count");
        d->lines = lines;
        if (lines <= d->cap)
            break;
        rc = alloc_doc_lines(d, lines);
        if (rc)
            return bail(rc);
    }
    sample_histogram(text, len, d->byte_freq);
    sample_text(text, len, d->sample);
    mark(c, "upload: sample byte histogram");
    *out = d;
    return OL_OK;
}

int ensure_tree(ol_tree *t, uint64_t lines)
{
    if (t->cap >= lines && t->parent)
        return OL_OK;
    if (t->parent) (void)hipFree(t->parent);
    if (t->fc) (void)hipFree(t->fc);
    if (t->ns) (void)hipFree(t->ns);
    if (t->se) (void)hipFree(t->se);
    t->parent = t->fc = t->ns = t->se = nullptr;
    t->cap = 0;
    uint64_t capa = round_up(lines ? lines : 1, LINE_ALIGN);
    HIPCHK(hipMalloc(&t->parent, capa * 4));
    HIPCHK(hipMalloc(&t->fc, capa * 4));
    HIPCHK(hipMalloc(&t->ns, capa * 4));
    HIPCHK(hipMalloc(&t->se, capa * 4));
    t->cap = capa;
    return OL_OK;
}

// The links a tree launch writes (OL_LINK_* bits); the others are left as
// they were.
int launch_tree(ol_ctx *c, const ol_doc *d, ol_tree *t, uint32_t links)
{
    uint32_t tiles = (uint32_t)ceil_div(d->lines ? d->lines : 1, (uint64_t)ol_tree_tile_lines(c->tree_rounds));
    HIPCHK(c->scr.tree_sum.ensure(ol_tree_sum_words(tiles) * 4));
    uint32_t *base = (uint32_t *)c->scr.tree_sum.p;
    TreeArgs a{};
    a.keys = d->keys;
    a.lines = d->lines;
    a.num_tiles = tiles;
    a.parent = (links & OL_LINK_PARENT) ? t->parent : nullptr;
    a.first_child = (links & OL_LINK_FIRST_CHILD) ? t->fc : nullptr;
    a.next_sibling = (links & OL_LINK_NEXT_SIBLING) ? t->ns : nullptr;
    a.subtree_end = (links & OL_LINK_SUBTREE_END) ? t->se : nullptr;
    ol_tree_set_sums(a, base, tiles);
    a.err = c->scr.ctr() + CTR_TREE_ERR;
    Prof p(c, "tree");
    HIPCHK(ol_launch_tree(a, c->tree_rounds, c->stream));
    return OL_OK;
}

int check_tree_error(ol_ctx *c, Scratch &s)
{
    uint32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, s.ctr() + CTR_TREE_ERR, 4, hipMemcpyDeviceToHost, c->stream));
    int rc = sync_check(c, s);
    if (rc)
        return rc;
    if (err)
        return fail(OL_ERR_TOO_DEEP,
                    "a content line is nested deeper than the " + std::to_string(OL_MAX_TREE_DEPTH) +
                        " levels the GPU tree tracks (MAX_TREE_DEPTH = " +
                        std::to_string(OL_MAX_TREE_DEPTH) + "); use the CPU builder");
    return OL_OK;
}

// Occurrences of every 4-byte window of the document's text sample (each of
// its 16 KiB windows counted on its own), sorted by window; built at the
// document's first literal search.
void build_window_counts(const ol_doc *d)
{
    std::lock_guard<std::mutex> g(d->win_mu);
    if (!d->win_key.empty() || d->sample.size() < 4)
        return;
    const size_t part = d->sample.size() <= (64u << 10) && d->n <= (64u << 10) ? d->sample.size() : (16u << 10);
    std::vector<uint32_t> all;
    all.reserve(d->sample.size());
    for (size_t p0 = 0; p0 < d->sample.size(); p0 += part) {
        const size_t end = std::min(d->sample.size(), p0 + part);
        for (size_t i = p0; i + 4 <= end; i++) {
            uint32_t w;
            std::memcpy(&w, d->sample.data() + i, 4);
            all.push_back(w);
        }
    }
    std::sort(all.begin(), all.end());
    for (size_t i = 0; i < all.size();) {
        size_t j = i + 1;
        while (j < all.size() && all[j] == all[i])
            j++;
        d->win_key.push_back(all[i]);
        d->win_count.push_back((uint32_t)(j - i));
        i = j;
    }
}

uint32_t window_count(const ol_doc *d, uint32_t w)   // after build_window_counts
{
    auto it = std::lower_bound(d->win_key.begin(), d->win_key.end(), w);
    return (it != d->win_key.end() && *it == w) ? d->win_count[it - d->win_key.begin()] : 0u;
}

// The pattern window the literal kernel tests.  Each 4-byte window j of the
// pattern costs (its occurrences in the document's text sample) x (2 + the
// pattern words that verification then reads: those holding a byte outside
// the window); the cheapest wins, then the one whose bytes are rarest in the
// sampled byte histogram (the sum of ln(count + 1) over its bytes, the
// reference's choose_window, crates/oneline-gpu/src/search.rs:205-221), then
// the earliest.  Patterns shorter than 4 bytes have one window.  The choice
// never changes the matches, only how many positions go on to the full
// comparison and what that costs; counting real windows, not bytes, matters
// on vocabulary text, where a word's letters are each common but the word is
// rarer or more frequent than their product says.  `hits` gets the chosen
// window's count in the sample.
uint32_t choose_window(const ol_doc *d, const uint8_t *pat, size_t m, uint32_t *hits)
{
    std::lock_guard<std::mutex> g(d->win_mu);   // the counts may be built by another thread
    const size_t k = std::min<size_t>(m, 4);
    uint32_t best = 0, best_count = 0;
    uint64_t best_cost = 0;
    double best_score = 0;
    for (size_t j = 0; j + k <= m; j++) {
        uint32_t cnt = 0;
        if (k == 4) {
            uint32_t w;
            std::memcpy(&w, pat + j, 4);
            cnt = window_count(d, w);
        }
        uint32_t words = 0;   // pattern words with a byte outside the window
        for (size_t w0 = 0; w0 < m; w0 += 4)
            for (size_t b = w0; b < std::min(m, w0 + 4); b++)
                if (b < j || b >= j + k) {
                    words++;
                    break;
                }
        const uint64_t cost = (uint64_t)cnt * (2 + words);
        double sc = 0;
        for (size_t b = 0; b < k; b++)
            sc += std::log((double)d->byte_freq[pat[j + b]] + 1.0);
        if (j == 0 || cost < best_cost || (cost == best_cost && sc < best_score)) {
            best = (uint32_t)j;
            best_count = cnt;
            best_cost = cost;
            best_score = sc;
        }
    }
    *hits = best_count;
    return best;
}

// The literal kernel's arguments for pattern `pat` over document `d`, except
// the scratch words.  Allocates nothing (the window counts of the document's
// sample are used when an earlier search built them).
void literal_args(const ol_doc *d, const uint8_t *pat, size_t m, uint64_t *out64, uint32_t *out32, uint64_t cap,
                  SearchArgs &a)
{
    a = SearchArgs{};
    uint32_t hits = 0;
    a.wo = choose_window(d, pat, m, &hits);
    // a window that occurs often (more than once in 8 KiB of the sample) in
    // a pattern longer than it makes the kernel wait on verifications: it then
    // takes 4 vectors per lane, which leaves registers for more waves
    a.dense = m > 4 && d->sample.size() && (uint64_t)hits * 8192 > d->sample.size();
    const uint32_t tiles = (uint32_t)ceil_div(d->n, ol_literal_tile(a.dense));
    a.text = d->text;
    a.n = (uint32_t)d->n;
    a.num_tiles = tiles;
    a.m = (uint32_t)m;
    uint32_t k = (uint32_t)std::min<size_t>(m, 4);
    a.win = 0;
    for (uint32_t i = 0; i < k; i++)
        a.win |= (uint32_t)pat[a.wo + i] << (8 * i);
    a.win_mask = (k == 4) ? 0xFFFFFFFFu : ((1u << (8 * k)) - 1u);
    for (uint32_t w = 0; w < 16; w++) {
        a.pw[w] = a.vmask[w] = 0;
        for (uint32_t b = 0; b < 4; b++) {
            const uint32_t j = 4 * w + b;
            if (j >= m)
                continue;
            a.pw[w] |= (uint32_t)pat[j] << (8 * b);
            if (j < a.wo || j >= a.wo + k)
                a.vmask[w] |= 0xFFu << (8 * b);
        }
    }
    a.cap = cap;
    a.out64 = out64;
    a.out32 = out32;
}

int run_search(ol_ctx *c, const ol_doc *d, const uint8_t *pat, size_t m, uint64_t *out64,
               uint32_t *out32, uint64_t cap, uint64_t *total)
{
    build_window_counts(d);
    SearchArgs a;
    literal_args(d, pat, m, out64, out32, cap, a);
    mark(c, "search: choose window");
    Scratch &s = c->scr;
    HIPCHK(ensure_status(s, STATUS_TILES + (size_t)a.num_tiles * SCAN_STATUS_WORDS));
    HIPCHK(zero_status(c, s));   // each tile zeroes its words again
    a.status = s.words() + STATUS_TILES;
    a.meta = s.meta();
    a.lb_err = s.ctr() + CTR_DEV_ERR;
    {
        Prof p(c, "search");
        HIPCHK(ol_launch_search(a, c->stream));
    }
    mark(c, "search: launch");
    const int rc = read_meta(c, s, total);
    mark(c, "search: wait for the GPU + read the total");
    return rc;
}

static_assert(OL_NOALLOC_MAX_TILES * 16384 >= OL_MAX_GPU_INPUT_LEN, "the no-allocation scratch covers every document");

// ol_search_device on the static scratch of this device (no allocation): the
// kernel is the same; the total and the error word are read back from the
// array through its device address.
int run_search_noalloc(ol_ctx *c, const ol_doc *d, const uint8_t *pat, size_t m, uint64_t *out64, uint64_t cap,
                       uint64_t *total)
{
    if (!c->noalloc)
        return fail(OL_ERR_DEVICE, "the no-allocation scratch of this device is not set up");
    SearchArgs a;
    literal_args(d, pat, m, out64, nullptr, cap, a);
    if (a.num_tiles > OL_NOALLOC_MAX_TILES)
        return fail(OL_ERR_TOO_LARGE, "the document needs " + std::to_string(a.num_tiles) +
                                          " literal tiles; the no-allocation scratch holds " +
                                          std::to_string(OL_NOALLOC_MAX_TILES));
    uint64_t *const sc = c->noalloc;
    a.status = sc + OL_NOALLOC_STATUS;
    a.meta = sc;
    a.lb_err = (uint32_t *)(sc + 8);
    HIPCHK(ol_launch_search(a, c->stream));
    uint64_t head[OL_NOALLOC_STATUS];
    HIPCHK(hipMemcpyAsync(head, sc, sizeof head, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t bits = (uint32_t)head[8];
    if (bits) {
        // a tile gave up: its status words may be left set; clear the array
        HIPCHK(ol_launch_clear_u64(sc, OL_NOALLOC_WORDS, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return fail(OL_ERR_DEVICE, dev_error_text(bits));
    }
    *total = head[0];
    return OL_OK;
}

int check_pattern(size_t len)
{
    if (len == 0 || len > OL_MAX_PATTERN_LEN)
        return fail(OL_ERR_PATTERN, "pattern must be 1..=" + std::to_string(OL_MAX_PATTERN_LEN) +
                                        " bytes, got " + std::to_string(len));
    return OL_OK;
}

int search_doc_host(ol_ctx *c, const ol_doc *d, const uint8_t *pat, size_t m, size_t max_matches,
                    uint32_t *total, uint32_t **offsets, size_t *n_offsets)
{
    if (!total || !offsets || !n_offsets || (!pat && m))
        return fail(OL_ERR_INVALID, "null argument");
    int rc = check_pattern(m);
    if (rc)
        return rc;
    if (max_matches > 0xFFFFFFFFull)
        return fail(OL_ERR_TOO_MANY, "max_matches " + std::to_string(max_matches) + " does not fit in u32");
    *total = 0;
    *offsets = nullptr;
    *n_offsets = 0;
    if (d->n == 0)
        return OL_OK;
    uint64_t cap = std::min<uint64_t>(max_matches, d->n);
    DevBuf out;
    HIPCHK(out.ensure((cap ? cap : 1) * 4));
    uint64_t tot = 0;
    rc = run_search(c, d, pat, m, nullptr, cap ? (uint32_t *)out.p : nullptr, cap, &tot);
    if (rc) {
        out.release();
        return rc;
    }
    uint64_t keep = std::min<uint64_t>(tot, cap);
    uint32_t *h = (uint32_t *)std::malloc((keep ? keep : 1) * 4);
    if (!h) {
        out.release();
        return fail(OL_ERR_NOMEM, "out of host memory");
    }
    if (keep) {
        hipError_t e = hipMemcpy(h, out.p, keep * 4, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            std::free(h);
            out.release();
            return fail(OL_ERR_DEVICE, std::string("download: ") + hipGetErrorString(e));
        }
    }
    out.release();
    mark(c, "search: download offsets");
    *total = (uint32_t)tot;
    *offsets = h;
    *n_offsets = keep;
    return OL_OK;
}

// The dictionary search's per-wave-tile state in the scratch status words:
// the scan's chunk totals first (the only words cleared), then tile_count
// (u32), tile_off and tile_list.  The candidate store (when matches are
// written) holds one candidate in four text positions per wave tile (16 bytes
// each, with the ids of its first three matches: 4 bytes per text byte); a
// wave tile with more candidates that have matches is filtered and counted
// again by the write pass.
size_t many_status_words(uint64_t n)
{
    const uint64_t wtiles = ceil_div(n, SEARCH_TILE) * (SEARCH_TILE / MANY_WT_BYTES);
    return wtiles / MANY_SCAN_CHUNK + 1 + (wtiles + 1) / 2 + 2 * wtiles;
}

hipError_t setup_many(ol_ctx *c, Scratch &s, ManyArgs &a, uint64_t n, bool writes)
{
    static_assert(OL_MAX_PATTERNS <= (1u << 22), "stored candidates keep 22-bit pattern ids");
    const uint32_t tiles = (uint32_t)ceil_div(n, SEARCH_TILE);
    const uint32_t wtiles = tiles * (SEARCH_TILE / MANY_WT_BYTES);
    const size_t nsums = wtiles / MANY_SCAN_CHUNK + 1;
    const size_t words = many_status_words(n);
    hipError_t e = ensure_status(s, words);
    if (e != hipSuccess || (e = clear_status(c, s, nsums)) != hipSuccess)   // the scan's chunk totals
        return e;
    soil(s, words);
    uint64_t *st = s.words();
    a.num_tiles = tiles;
    a.num_wtiles = wtiles;
    a.scan_sums = st;
    a.tile_count = (uint32_t *)(st + nsums);
    a.tile_off = st + nsums + (wtiles + 1) / 2;
    a.tile_list = a.tile_off + wtiles;
    a.meta = s.meta();
    a.lb_err = s.ctr() + CTR_DEV_ERR;
    a.cands = nullptr;
    if (!writes)
        return hipSuccess;
    const size_t cb = (size_t)wtiles * MANY_WT_BYTES * 4;
    if (s.cands.bytes < cb && s.frozen)
        return hipErrorInvalidValue;
    if ((e = s.cands.ensure(cb)) != hipSuccess)
        return e;
    a.cands = (uint32_t *)s.cands.p;
    return hipSuccess;
}

// Kind / depth query of lines [r0, r1) of a resident document: the first
// min(total, max_matches) matching line indices, ascending, library-allocated.
int run_query(ol_ctx *c, const ol_doc *d, uint32_t r0, uint32_t r1, uint32_t kind, uint32_t depth,
              size_t max_matches, uint32_t *total, uint32_t **indices, size_t *n_indices)
{
    if (!total || !indices || !n_indices)
        return fail(OL_ERR_INVALID, "null argument");
    if (max_matches > 0xFFFFFFFFull)
        return fail(OL_ERR_TOO_MANY, "max_matches " + std::to_string(max_matches) + " does not fit in u32");
    *total = 0;
    *indices = nullptr;
    *n_indices = 0;
    if (r1 <= r0)
        return OL_OK;
    const uint32_t blocks = ol_query_blocks(r0, r1);
    Scratch &s = c->scr;
    HIPCHK(ensure_status(s, 2 * (size_t)blocks + 8));
    soil(s, 2 * (size_t)blocks + 8);
    const uint64_t cap = std::min<uint64_t>(max_matches, r1 - r0);
    DevBuf out;
    HIPCHK(out.ensure((cap ? cap : 1) * 4));
    QueryArgs a{};
    a.text = d->text;
    a.n = (uint32_t)d->n;
    a.lines = d->lines;
    a.offsets = d->offsets;
    a.classes = d->classes;
    a.keys = d->keys;
    a.nodes = d->nodes;
    a.kind = kind;
    a.depth = depth;
    a.r0 = r0;
    a.r1 = r1;
    a.block_count = (uint32_t *)s.words();
    a.block_off = s.words() + blocks;
    a.out = cap ? (uint32_t *)out.p : nullptr;
    a.cap = cap;
    {
        Prof p(c, "query");
        const hipError_t e = ol_launch_query(a, s.meta(), c->stream);
        if (e != hipSuccess) {
            out.release();
            return fail(OL_ERR_DEVICE, std::string("query: ") + hipGetErrorString(e));
        }
    }
    mark(c, "query: launch");
    uint64_t tot = 0;
    int rc = read_meta(c, s, &tot);
    mark(c, "query: wait for the GPU + read the total");
    if (rc) {
        out.release();
        return rc;
    }
    const uint64_t keep = std::min<uint64_t>(tot, cap);
    uint32_t *h = (uint32_t *)std::malloc((keep ? keep : 1) * 4);
    if (!h) {
        out.release();
        return fail(OL_ERR_NOMEM, "out of host memory");
    }
    if (keep && hipMemcpy(h, out.p, keep * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        std::free(h);
        out.release();
        return fail(OL_ERR_DEVICE, "query download failed");
    }
    out.release();
    mark(c, "query: copy out");
    *total = (uint32_t)tot;
    *indices = h;
    *n_indices = (size_t)keep;
    return OL_OK;
}

// ---- .olb files: a parsed document on disk ------------------------------------
// Layout (.agents/tasks/phase4-zero-copy-binary-serialization.md:13-36 of the
// reference, a specification it never built): a 64-byte header
//   0x00 magic "ONELINE\0" | 0x08 version u32 = 1 | 0x0C flags u32
//   0x10 text_len u64 | 0x18 line_count u64
//   0x20 text_offset u64 | 0x28 nodes_offset u64 | 0x30 line_offsets_offset u64
//   0x38 checksum u64
// then the text, the FlatNode array (16 bytes per line) and the u32 line
// offsets, each section starting on a 64-byte boundary.  The checksum (low 32
// bits) is CRC-32 (IEEE 802.3, reflected, polynomial 0xEDB88320) of every byte
// from text_offset to the end of the file; flag bit 0 says it is present.
constexpr uint32_t OLB_VERSION = 1, OLB_FLAG_CRC = 1;
constexpr size_t OLB_HEADER = 64;

uint32_t crc32_update(uint32_t crc, const uint8_t *p, size_t n)
{
    static uint32_t table[8][256];
    static std::once_flag once;
    std::call_once(once, [] {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t c = i;
            for (int k = 0; k < 8; k++)
                c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[0][i] = c;
        }
        for (uint32_t i = 0; i < 256; i++)
            for (int t = 1; t < 8; t++)
                table[t][i] = (table[t - 1][i] >> 8) ^ table[0][table[t - 1][i] & 0xFF];
    });
    crc = ~crc;
    while (n >= 8) {   // slicing by 8
        uint32_t lo, hi;
        std::memcpy(&lo, p, 4);
        std::memcpy(&hi, p + 4, 4);
        lo ^= crc;
        crc = table[7][lo & 0xFF] ^ table[6][(lo >> 8) & 0xFF] ^ table[5][(lo >> 16) & 0xFF] ^ table[4][lo >> 24] ^
              table[3][hi & 0xFF] ^ table[2][(hi >> 8) & 0xFF] ^ table[1][(hi >> 16) & 0xFF] ^ table[0][hi >> 24];
        p += 8;
        n -= 8;
    }
    while (n--)
        crc = table[0][(crc ^ *p++) & 0xFF] ^ (crc >> 8);
    return ~crc;
}

struct OlbLayout {
    uint64_t text_off, nodes_off, offs_off, size;
};

OlbLayout olb_layout(uint64_t n, uint64_t lines)
{
    OlbLayout l;
    l.text_off = OLB_HEADER;
    l.nodes_off = round_up(l.text_off + n, 64);
    l.offs_off = round_up(l.nodes_off + lines * 16, 64);
    l.size = l.offs_off + lines * 4;
    return l;
}

int save_olb(ol_ctx *c, const ol_doc *d, const char *path)
{
    const OlbLayout l = olb_layout(d->n, d->lines);
    std::vector<uint8_t> buf;
    try {
        buf.assign(l.size, 0);
    } catch (...) {
        return fail(OL_ERR_NOMEM, "out of host memory");
    }
    if (d->n)
        HIPCHK(hipMemcpyAsync(buf.data() + l.text_off, d->text, d->n, hipMemcpyDeviceToHost, c->stream));
    if (d->lines) {
        HIPCHK(hipMemcpyAsync(buf.data() + l.offs_off, d->offsets, d->lines * 4, hipMemcpyDeviceToHost, c->stream));
        if (d->nodes) {
            HIPCHK(hipMemcpyAsync(buf.data() + l.nodes_off, d->nodes, d->lines * 16, hipMemcpyDeviceToHost, c->stream));
        } else {   // a lean document: parse its lines again into a temporary node array
            DevBuf nodes;
            HIPCHK(nodes.ensure(d->lines * 16));
            HIPCHK(ol_launch_parse_given(d->text, (uint32_t)d->n, d->offsets, d->lines, (ol_node_dev *)nodes.p,
                                         c->stream));
            HIPCHK(hipMemcpyAsync(buf.data() + l.nodes_off, nodes.p, d->lines * 16, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            nodes.release();
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    uint8_t *h = buf.data();
    std::memcpy(h, "ONELINE\0", 8);
    const uint32_t ver = OLB_VERSION, flags = OLB_FLAG_CRC;
    std::memcpy(h + 0x08, &ver, 4);
    std::memcpy(h + 0x0C, &flags, 4);
    std::memcpy(h + 0x10, &d->n, 8);
    std::memcpy(h + 0x18, &d->lines, 8);
    std::memcpy(h + 0x20, &l.text_off, 8);
    std::memcpy(h + 0x28, &l.nodes_off, 8);
    std::memcpy(h + 0x30, &l.offs_off, 8);
    const uint64_t crc = crc32_update(0, h + l.text_off, l.size - l.text_off);
    std::memcpy(h + 0x38, &crc, 8);
    FILE *f = std::fopen(path, "wb");
    if (!f)
        return fail(OL_ERR_INVALID, std::string("cannot open ") + path + " for writing");
    const bool ok = std::fwrite(h, 1, l.size, f) == l.size;
    if (std::fclose(f) != 0 || !ok)
        return fail(OL_ERR_DEVICE, std::string("write to ") + path + " failed");
    return OL_OK;
}

int load_olb(ol_ctx *c, const char *path, bool verify, ol_doc **out)
{
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0)
        return fail(OL_ERR_INVALID, std::string("cannot open ") + path);
    struct stat st {};
    if (::fstat(fd, &st) != 0 || (size_t)st.st_size < OLB_HEADER) {
        ::close(fd);
        return fail(OL_ERR_INVALID, std::string(path) + " is not a .olb file (too short)");
    }
    const size_t size = (size_t)st.st_size;
    void *map = ::mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    ::close(fd);
    if (map == MAP_FAILED)
        return fail(OL_ERR_INVALID, std::string("cannot map ") + path);
    const uint8_t *h = (const uint8_t *)map;
    auto bad = [&](const std::string &why) {
        ::munmap(map, size);
        return fail(OL_ERR_INVALID, std::string(path) + ": " + why);
    };
    uint32_t ver, flags;
    uint64_t n, lines, toff, noff, ooff, crc;
    std::memcpy(&ver, h + 0x08, 4);
    std::memcpy(&flags, h + 0x0C, 4);
    std::memcpy(&n, h + 0x10, 8);
    std::memcpy(&lines, h + 0x18, 8);
    std::memcpy(&toff, h + 0x20, 8);
    std::memcpy(&noff, h + 0x28, 8);
    std::memcpy(&ooff, h + 0x30, 8);
    std::memcpy(&crc, h + 0x38, 8);
    if (std::memcmp(h, "ONELINE\0", 8) != 0)
        return bad("bad magic");
    if (ver != OLB_VERSION)
        return bad("unsupported version " + std::to_string(ver));
    if (n > OL_MAX_GPU_INPUT_LEN)
        return bad("text of " + std::to_string(n) + " bytes is too large for the GPU path");
    if (lines > n || toff < OLB_HEADER || toff > size || n > size - toff || noff > size || lines * 16 > size - noff ||
        ooff > size || lines * 4 > size - ooff)
        return bad("sections do not fit the file");
    if (verify && (flags & OLB_FLAG_CRC) && crc32_update(0, h + toff, size - toff) != (uint32_t)crc)
        return bad("checksum mismatch");
    ol_doc *d = new (std::nothrow) ol_doc();
    if (!d)
        return bad("out of host memory");
    d->n = n;
    d->with_nodes = true;
    std::memset(d->byte_freq, 0, sizeof d->byte_freq);
    auto fail_doc = [&](int rc) {
        free_doc_lines(d);
        if (d->text)
            (void)hipFree(d->text);
        delete d;
        ::munmap(map, size);
        return rc;
    };
    if (n) {
        const uint64_t padded = round_up(n, TEXT_ALIGN) + TEXT_SLACK;
        if (hipMalloc(&d->text, padded) != hipSuccess) {
            d->text = nullptr;
            return fail_doc(fail(OL_ERR_NOMEM, "device allocation of the text failed"));
        }
        if (hipMemcpyAsync(d->text, h + toff, n, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemsetAsync(d->text + n, 0, padded - n, c->stream) != hipSuccess)
            return fail_doc(fail(OL_ERR_DEVICE, "text upload failed"));
        int rc = alloc_doc_lines(d, lines);
        if (rc)
            return fail_doc(rc);
        if (lines) {
            if (hipMemcpyAsync(d->offsets, h + ooff, lines * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                hipMemcpyAsync(d->nodes, h + noff, lines * 16, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                ol_launch_nodes_to_lines(d->nodes, lines, d->classes, d->keys, c->stream) != hipSuccess)
                return fail_doc(fail(OL_ERR_DEVICE, "upload of the parsed lines failed"));
        }
        d->lines = lines;
        sample_histogram(h + toff, n, d->byte_freq);
        sample_text(h + toff, n, d->sample);
        if (hipStreamSynchronize(c->stream) != hipSuccess)
            return fail_doc(fail(OL_ERR_DEVICE, "upload failed"));
    }
    ::munmap(map, size);
    *out = d;
    return OL_OK;
}

int run_many(ol_ctx *c, const ol_doc *d, const ol_dict *s, uint64_t *out_off, uint32_t *out_id,
             uint64_t cap, uint64_t *total)
{
    ManyArgs a{};
    a.text = d->text;
    a.n = (uint32_t)d->n;
    a.d = s->d;
    a.cap = cap;
    a.out_off = out_off;
    a.out_id = out_id;
    HIPCHK(setup_many(c, c->scr, a, d->n, out_off || out_id));
    {
        Prof p(c, "many");
        HIPCHK(ol_launch_many(a, c->stream));
    }
    mark(c, "many: launch");
    const int rc = read_meta(c, c->scr, total);
    mark(c, "many: wait for the GPU + read the total");
    return rc;
}

uint32_t ceil_log2(uint64_t v)
{
    uint32_t b = 0;
    while ((1ull << b) < v)
        b++;
    return b;
}

}  // namespace

// ============================================================================
extern "C" {

const char *ol_last_error(void) { return g_err.c_str(); }

void ol_free(void *p) { std::free(p); }

int ol_ctx_create(ol_ctx **out)
{
    if (!out)
        return fail(OL_ERR_INVALID, "null argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(OL_ERR_DEVICE, "no HIP device available");
    ol_ctx *c = new (std::nothrow) ol_ctx();
    if (!c)
        return fail(OL_ERR_NOMEM, "out of host memory");
    c->device = 0;
    if (hipSetDevice(0) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(OL_ERR_DEVICE, "stream creation failed");
    }
    if (ol_noalloc_scratch(0, &c->noalloc) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(OL_ERR_DEVICE, "the no-allocation scratch's device address is not available");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess) {
        // the marketing name can come back empty: then the runtime's device
        // name, else a generic one; the architecture, CUs and HBM size follow
        std::string nm = prop.name;
        if (nm.empty()) {
            char buf[256] = {0};
            if (hipDeviceGetName(buf, sizeof(buf) - 1, 0) == hipSuccess)
                nm = buf;
        }
        if (nm.empty())
            nm = "AMD GPU";
        c->name = nm + " (" + prop.gcnArchName + ", " + std::to_string(prop.multiProcessorCount) + " CUs, " +
                  std::to_string((prop.totalGlobalMem + (1ull << 29)) >> 30) + " GiB)";
    }
    if (scratch_init(c->scr) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(OL_ERR_NOMEM, "scratch allocation failed");
    }
    *out = c;
    return OL_OK;
}

void ol_ctx_destroy(ol_ctx *c)
{
    if (!c)
        return;
    (void)hipStreamSynchronize(c->stream);
    for (auto &r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : c->pool)
        (void)hipEventDestroy(e);
    c->scr.release();
    for (int i = 0; i < 16; i++) {
        if (c->ring[i]) (void)hipHostFree(c->ring[i]);
        if (c->ring_ev[i]) (void)hipEventDestroy(c->ring_ev[i]);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int ol_ctx_adapter_name(ol_ctx *c, char *buf, size_t cap)
{
    if (!c || !buf || cap == 0)
        return fail(OL_ERR_INVALID, "null argument");
    std::snprintf(buf, cap, "%s", c->name.c_str());
    return OL_OK;
}

int ol_ctx_is_software(ol_ctx *c) { return c ? 0 : fail(OL_ERR_INVALID, "null argument"); }

int ol_ctx_timestamps_supported(ol_ctx *c) { return c ? 1 : fail(OL_ERR_INVALID, "null argument"); }

int ol_ctx_synchronize(ol_ctx *c)
{
    if (!c)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    return sync_check(c, c->scr);
}

int ol_doc_upload(ol_ctx *c, const uint8_t *text, size_t len, ol_doc **out)
{
    if (!c)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    return upload_doc(c, text, len, true, out);
}

int ol_doc_upload_lean(ol_ctx *c, const uint8_t *text, size_t len, ol_doc **out)
{
    if (!c)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    return upload_doc(c, text, len, false, out);
}

void ol_doc_free(ol_doc *d)
{
    if (!d)
        return;
    free_doc_lines(d);
    if (d->text)
        (void)hipFree(d->text);
    delete d;
}

size_t ol_doc_line_count(const ol_doc *d) { return d ? d->lines : 0; }
size_t ol_doc_text_len(const ol_doc *d) { return d ? d->n : 0; }

int ol_doc_download(ol_ctx *c, const ol_doc *d, uint32_t *offsets, uint8_t *classes, ol_flat_node *nodes)
{
    if (!c || !d)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (d->lines == 0)
        return OL_OK;
    if (nodes && !d->nodes)
        return fail(OL_ERR_INVALID, "document was uploaded without nodes");
    if (offsets)
        HIPCHK(hipMemcpyAsync(offsets, d->offsets, d->lines * 4, hipMemcpyDeviceToHost, c->stream));
    if (classes)
        HIPCHK(hipMemcpyAsync(classes, d->classes, d->lines, hipMemcpyDeviceToHost, c->stream));
    if (nodes)
        HIPCHK(hipMemcpyAsync(nodes, d->nodes, d->lines * 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return OL_OK;
}

int ol_doc_reparse(ol_ctx *c, ol_doc *d)
{
    if (!c || !d)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (d->n == 0)
        return OL_OK;
    return launch_scan(c, d);
}

int ol_calib_scan_floor(ol_ctx *c, const ol_doc *d, uint32_t lines_per_tile, int write, int reps, double *us)
{
    if (!c || !d || !us || reps <= 0)
        return fail(OL_ERR_INVALID, "null argument or no repetitions");
    if (lines_per_tile > 32768)
        return fail(OL_ERR_INVALID, "a 32 KiB tile holds at most 32768 lines");
    std::lock_guard<std::mutex> g(c->mu);
    *us = 0.0;
    const uint32_t tiles = (uint32_t)((d->n + 32767) / 32768);   // the text is padded to 64 KiB
    if (tiles == 0)
        return OL_OK;
    const size_t lines = (size_t)tiles * lines_per_tile;
    struct Held {   // released on every return
        DevBuf off, cls, key, sink;
        hipEvent_t a = nullptr, b = nullptr;
        ~Held()
        {
            off.release();
            cls.release();
            key.release();
            sink.release();
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } h;
    HIPCHK(h.sink.ensure(64));
    if (write) {
        HIPCHK(h.off.ensure(lines * 4 + 4));
        HIPCHK(h.cls.ensure(lines + 4));
        HIPCHK(h.key.ensure(lines + 4));
    }
    HIPCHK(hipEventCreate(&h.a));
    HIPCHK(hipEventCreate(&h.b));
    auto launch = [&]() {
        return ol_launch_stream_floor(d->text, tiles, lines_per_tile, write != 0, (uint32_t *)h.off.p,
                                      (uint8_t *)h.cls.p, (uint8_t *)h.key.p, (uint32_t *)h.sink.p, c->stream);
    };
    HIPCHK(launch());   // warm-up
    HIPCHK(hipEventRecord(h.a, c->stream));
    for (int i = 0; i < reps; i++)
        HIPCHK(launch());
    HIPCHK(hipEventRecord(h.b, c->stream));
    HIPCHK(hipEventSynchronize(h.b));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, h.a, h.b));
    *us = (double)ms * 1e3 / reps;
    return OL_OK;
}

int ol_scan_line_offsets(ol_ctx *c, const uint8_t *text, size_t len, uint32_t **offsets, size_t *line_count)
{
    if (!c || !offsets || !line_count)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    *offsets = nullptr;
    *line_count = 0;
    ol_doc *d = nullptr;
    int rc = upload_doc(c, text, len, false, &d);
    if (rc)
        return rc;
    uint32_t *h = (uint32_t *)std::malloc((d->lines ? d->lines : 1) * 4);
    if (!h) {
        ol_doc_free(d);
        return fail(OL_ERR_NOMEM, "out of host memory");
    }
    if (d->lines && hipMemcpy(h, d->offsets, d->lines * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        std::free(h);
        ol_doc_free(d);
        return fail(OL_ERR_DEVICE, "download failed");
    }
    *offsets = h;
    *line_count = d->lines;
    ol_doc_free(d);
    return OL_OK;
}

int ol_parse_raw(ol_ctx *c, const uint8_t *text, size_t len, ol_flat_node **nodes, uint32_t **offsets,
                 size_t *line_count)
{
    if (!c || !nodes || !offsets || !line_count)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    *nodes = nullptr;
    *offsets = nullptr;
    *line_count = 0;
    ol_doc *d = nullptr;
    int rc = upload_doc(c, text, len, true, &d);
    if (rc)
        return rc;
    size_t L = d->lines;
    uint32_t *ho = (uint32_t *)std::malloc((L ? L : 1) * 4);
    ol_flat_node *hn = (ol_flat_node *)std::malloc((L ? L : 1) * 16);
    if (!ho || !hn) {
        std::free(ho);
        std::free(hn);
        ol_doc_free(d);
        return fail(OL_ERR_NOMEM, "out of host memory");
    }
    if (L && (hipMemcpy(ho, d->offsets, L * 4, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(hn, d->nodes, L * 16, hipMemcpyDeviceToHost) != hipSuccess)) {
        std::free(ho);
        std::free(hn);
        ol_doc_free(d);
        return fail(OL_ERR_DEVICE, "download failed");
    }
    ol_doc_free(d);
    *nodes = hn;
    *offsets = ho;
    *line_count = L;
    return OL_OK;
}

int ol_parse(ol_ctx *c, const uint8_t *text, size_t len, const uint32_t *offsets, size_t line_count,
             ol_flat_node **nodes)
{
    if (!c || !nodes || (!offsets && line_count) || (!text && len))
        return fail(OL_ERR_INVALID, "null argument");
    if (len > OL_MAX_GPU_INPUT_LEN)
        return fail(OL_ERR_TOO_LARGE, "input is " + std::to_string(len) + " bytes; too large");
    std::lock_guard<std::mutex> g(c->mu);
    *nodes = nullptr;
    ol_flat_node *hn = (ol_flat_node *)std::malloc((line_count ? line_count : 1) * 16);
    if (!hn)
        return fail(OL_ERR_NOMEM, "out of host memory");
    if (line_count == 0) {
        *nodes = hn;
        return OL_OK;
    }
    DevBuf t, o, nd;
    auto done = [&](int rc) {
        t.release();
        o.release();
        nd.release();
        if (rc)
            std::free(hn);
        return rc;
    };
    if (t.ensure(len + 16) != hipSuccess || o.ensure(line_count * 4) != hipSuccess ||
        nd.ensure(line_count * 16) != hipSuccess)
        return done(fail(OL_ERR_NOMEM, "device allocation failed"));
    if ((len && hipMemcpyAsync(t.p, text, len, hipMemcpyHostToDevice, c->stream) != hipSuccess) ||
        hipMemcpyAsync(o.p, offsets, line_count * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return done(fail(OL_ERR_DEVICE, "upload failed"));
    {
        Prof p(c, "parse_given");
        if (ol_launch_parse_given((const uint8_t *)t.p, (uint32_t)len, (const uint32_t *)o.p, line_count,
                                  (ol_node_dev *)nd.p, c->stream) != hipSuccess)
            return done(fail(OL_ERR_DEVICE, "launch failed"));
    }
    if (hipMemcpyAsync(hn, nd.p, line_count * 16, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return done(fail(OL_ERR_DEVICE, "download failed"));
    *nodes = hn;
    return done(OL_OK);
}

// ---- tree -------------------------------------------------------------------------
int ol_tree_build(ol_ctx *c, const ol_doc *d, ol_tree *reuse, ol_tree **out)
{
    if (!c || !d || !out)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    ol_tree *t = reuse ? reuse : new (std::nothrow) ol_tree();
    if (!t)
        return fail(OL_ERR_NOMEM, "out of host memory");
    *out = nullptr;
    t->lines = d->lines;
    if (d->lines == 0) {
        *out = t;
        return OL_OK;
    }
    int rc = ensure_tree(t, d->lines);
    mark(c, "tree: prepare buffers");
    if (!rc)
        rc = launch_tree(c, d, t, OL_LINK_ALL);
    mark(c, "tree: launch");
    if (!rc)
        rc = check_tree_error(c, c->scr);
    mark(c, "tree: wait for the GPU + read status");
    if (rc) {
        ol_tree_free(t);
        return rc;
    }
    t->links = OL_LINK_ALL;
    *out = t;
    return OL_OK;
}

int ol_tree_rebuild_links_async(ol_ctx *c, const ol_doc *d, ol_tree *t, uint32_t links)
{
    if (!c || !d || !t)
        return fail(OL_ERR_INVALID, "null argument");
    if (links == 0 || (links & ~OL_LINK_ALL))
        return fail(OL_ERR_INVALID, "links must be a non-empty set of OL_LINK_* bits");
    std::lock_guard<std::mutex> g(c->mu);
    if (t->cap < d->lines && d->lines)
        return fail(OL_ERR_INVALID, "tree capacity is smaller than the document");
    t->lines = d->lines;
    t->links = links;
    if (d->lines == 0)
        return OL_OK;
    return launch_tree(c, d, t, links);
}

int ol_tree_rebuild_async(ol_ctx *c, const ol_doc *d, ol_tree *t)
{
    return ol_tree_rebuild_links_async(c, d, t, OL_LINK_ALL);
}

int ol_tree_check(ol_ctx *c, const ol_tree *t)
{
    if (!c || !t)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (t->lines == 0)
        return sync_check(c, c->scr);
    return check_tree_error(c, c->scr);
}

void ol_tree_free(ol_tree *t)
{
    if (!t)
        return;
    if (t->parent) (void)hipFree(t->parent);
    if (t->fc) (void)hipFree(t->fc);
    if (t->ns) (void)hipFree(t->ns);
    if (t->se) (void)hipFree(t->se);
    delete t;
}

size_t ol_tree_line_count(const ol_tree *t) { return t ? t->lines : 0; }

int ol_tree_download(ol_ctx *c, const ol_tree *t, uint32_t *parent, uint32_t *subtree_end,
                     uint32_t *first_child, uint32_t *next_sibling)
{
    if (!c || !t)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    size_t b = t->lines * 4;
    if (!b)
        return OL_OK;
    const uint32_t want = (parent ? OL_LINK_PARENT : 0u) | (subtree_end ? OL_LINK_SUBTREE_END : 0u) |
                          (first_child ? OL_LINK_FIRST_CHILD : 0u) | (next_sibling ? OL_LINK_NEXT_SIBLING : 0u);
    if (want & ~t->links)
        return fail(OL_ERR_INVALID, "the tree's last build did not write every requested link array");
    if (parent) HIPCHK(hipMemcpyAsync(parent, t->parent, b, hipMemcpyDeviceToHost, c->stream));
    if (subtree_end) HIPCHK(hipMemcpyAsync(subtree_end, t->se, b, hipMemcpyDeviceToHost, c->stream));
    if (first_child) HIPCHK(hipMemcpyAsync(first_child, t->fc, b, hipMemcpyDeviceToHost, c->stream));
    if (next_sibling) HIPCHK(hipMemcpyAsync(next_sibling, t->ns, b, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return OL_OK;
}

// ---- AST bridge: structural pass ---------------------------------------------------
int ol_tree_bridge(ol_ctx *c, const ol_doc *d, const ol_tree *t, uint8_t *flags, uint32_t *container)
{
    if (!c || !d || !t)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (t->lines != d->lines)
        return fail(OL_ERR_INVALID, "the tree has " + std::to_string(t->lines) + " lines, the document " +
                                        std::to_string(d->lines));
    if (t->lines && !(t->links & OL_LINK_PARENT))
        return fail(OL_ERR_INVALID, "the tree's last build did not write parent");
    const uint64_t lines = d->lines;
    if (lines == 0)
        return OL_OK;
    const uint32_t blocks = ol_bridge_blocks(lines);
    Scratch &s = c->scr;
    HIPCHK(ensure_status(s, (size_t)blocks + 8));
    soil(s, (size_t)blocks + 8);
    DevBuf fl, ct;
    hipError_t e = fl.ensure((size_t)blocks * 4096);
    if (e == hipSuccess)
        e = ct.ensure((size_t)blocks * 4096 * 4);
    if (e != hipSuccess) {
        fl.release();
        return fail(OL_ERR_NOMEM, "device allocation of the bridge outputs failed");
    }
    BridgeArgs a{};
    a.text = d->text;
    a.n = (uint32_t)d->n;
    a.lines = lines;
    a.offsets = d->offsets;
    a.classes = d->classes;
    a.keys = d->keys;
    a.nodes = d->nodes;
    a.parent = t->parent;
    a.flags = (uint8_t *)fl.p;
    a.container = (uint32_t *)ct.p;
    a.blk_last = (uint32_t *)s.words();
    a.blk_in = a.blk_last + blocks;
    {
        Prof p(c, "bridge");
        e = ol_launch_bridge(a, c->stream);
    }
    mark(c, "bridge: launch");
    if (e == hipSuccess && flags)
        e = hipMemcpyAsync(flags, fl.p, lines, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && container)
        e = hipMemcpyAsync(container, ct.p, lines * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(c->stream);
    mark(c, "bridge: wait for the GPU + download");
    fl.release();
    ct.release();
    if (e != hipSuccess)
        return fail(OL_ERR_DEVICE, std::string("bridge: ") + hipGetErrorString(e));
    return OL_OK;
}

// ---- literal search -------------------------------------------------------------
int ol_search_document(ol_ctx *c, const ol_doc *d, const uint8_t *pattern, size_t len, size_t max_matches,
                       uint32_t *total, uint32_t **offsets, size_t *n_offsets)
{
    if (!c || !d)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    return search_doc_host(c, d, pattern, len, max_matches, total, offsets, n_offsets);
}

int ol_search_raw(ol_ctx *c, const uint8_t *text, size_t text_len, const uint8_t *pattern, size_t len,
                  size_t max_matches, uint32_t *total, uint32_t **offsets, size_t *n_offsets)
{
    if (!c)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    int rc = check_pattern(len);
    if (rc)
        return rc;
    if (text_len > OL_MAX_GPU_INPUT_LEN)
        return fail(OL_ERR_TOO_LARGE, "input is " + std::to_string(text_len) + " bytes; too large");
    if (!total || !offsets || !n_offsets)
        return fail(OL_ERR_INVALID, "null argument");
    if (text_len == 0) {
        *total = 0;
        *offsets = nullptr;
        *n_offsets = 0;
        return OL_OK;
    }
    // The one-shot path needs only the text on the device, no line scan.
    ol_doc tmp;
    tmp.n = text_len;
    sample_histogram(text, text_len, tmp.byte_freq);
    sample_text(text, text_len, tmp.sample);
    uint64_t padded = round_up(text_len, TEXT_ALIGN) + TEXT_SLACK;
    HIPCHK(hipMalloc(&tmp.text, padded));
    if (hipMemcpyAsync(tmp.text, text, text_len, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemsetAsync(tmp.text + text_len, 0, padded - text_len, c->stream) != hipSuccess) {
        (void)hipFree(tmp.text);
        return fail(OL_ERR_DEVICE, "upload failed");
    }
    rc = search_doc_host(c, &tmp, pattern, len, max_matches, total, offsets, n_offsets);
    (void)hipFree(tmp.text);
    tmp.text = nullptr;
    return rc;
}

int ol_doc_save(ol_ctx *c, const ol_doc *d, const char *path)
{
    if (!c || !d || !path)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    return save_olb(c, d, path);
}

int ol_doc_load(ol_ctx *c, const char *path, int verify, ol_doc **out)
{
    if (!c || !path || !out)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    return load_olb(c, path, verify != 0, out);
}

int ol_query_document(ol_ctx *c, const ol_doc *d, uint32_t kind, uint32_t depth, size_t max_matches,
                      uint32_t *total, uint32_t **indices, size_t *n_indices)
{
    if (!c || !d)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    return run_query(c, d, 0, (uint32_t)d->lines, kind, depth, max_matches, total, indices, n_indices);
}

int ol_query_subtree(ol_ctx *c, const ol_doc *d, const ol_tree *t, size_t root, uint32_t kind, uint32_t depth,
                     size_t max_matches, uint32_t *total, uint32_t **indices, size_t *n_indices)
{
    if (!c || !d || !t)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (!(t->links & OL_LINK_SUBTREE_END))
        return fail(OL_ERR_INVALID, "the tree's last build did not write subtree_end");
    if (root >= t->lines || t->lines != d->lines)
        return fail(OL_ERR_INVALID, "line " + std::to_string(root) + " is out of range (document has " +
                                        std::to_string(d->lines) + " lines, tree has " + std::to_string(t->lines) +
                                        ")");
    uint32_t end = 0;
    HIPCHK(hipMemcpyAsync(&end, t->se + root, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return run_query(c, d, (uint32_t)root, end, kind, depth, max_matches, total, indices, n_indices);
}

int ol_query_raw(ol_ctx *c, const uint8_t *text, size_t len, uint32_t kind, uint32_t depth, size_t max_matches,
                 uint32_t *total, uint32_t **indices, size_t *n_indices)
{
    if (!c || (!text && len))
        return fail(OL_ERR_INVALID, "null argument");
    if (!total || !indices || !n_indices)
        return fail(OL_ERR_INVALID, "null argument");
    ol_doc *d = nullptr;
    {
        std::lock_guard<std::mutex> g(c->mu);
        int rc = upload_doc(c, text, len, false, &d);
        if (rc)
            return rc;
        rc = run_query(c, d, 0, (uint32_t)d->lines, kind, depth, max_matches, total, indices, n_indices);
        ol_doc_free(d);
        return rc;
    }
}

int ol_query(ol_ctx *c, const uint8_t *text, size_t len, const uint32_t *offsets, size_t line_count, uint32_t kind,
             uint32_t depth, size_t max_matches, uint32_t *total, uint32_t **indices, size_t *n_indices)
{
    if (!c || (!text && len) || (!offsets && line_count) || !total || !indices || !n_indices)
        return fail(OL_ERR_INVALID, "null argument");
    if (len > OL_MAX_GPU_INPUT_LEN)
        return fail(OL_ERR_TOO_LARGE, "input is " + std::to_string(len) + " bytes; the GPU path addresses at most " +
                                          std::to_string(OL_MAX_GPU_INPUT_LEN) + " bytes (32-bit offsets)");
    if (line_count > 0xFFFFFFFFull)
        return fail(OL_ERR_TOO_MANY, "line count does not fit u32");
    std::lock_guard<std::mutex> g(c->mu);
    *total = 0;
    *indices = nullptr;
    *n_indices = 0;
    if (line_count == 0)
        return OL_OK;
    // the caller's lines as a temporary document: text, offsets, FlatNodes
    // parsed from them, and the class / key bytes derived from those
    ol_doc d;
    d.n = len;
    d.with_nodes = true;
    auto done = [&](int rc) {
        free_doc_lines(&d);
        if (d.text)
            (void)hipFree(d.text);
        d.text = nullptr;
        return rc;
    };
    const uint64_t padded = round_up(len ? len : 1, TEXT_ALIGN) + TEXT_SLACK;
    if (hipMalloc(&d.text, padded) != hipSuccess) {
        d.text = nullptr;
        return fail(OL_ERR_NOMEM, "device allocation of the text failed");
    }
    int rc = stage_upload(c, d.text, len, [text](uint8_t *buf, uint64_t off, uint64_t k) {
        std::memcpy(buf, text + off, k);
    });
    if (rc)
        return done(rc);
    if (hipMemsetAsync(d.text + len, 0, padded - len, c->stream) != hipSuccess)
        return done(fail(OL_ERR_DEVICE, "text upload failed"));
    if ((rc = alloc_doc_lines(&d, line_count)))
        return done(rc);
    d.lines = line_count;
    if (hipMemcpyAsync(d.offsets, offsets, line_count * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return done(fail(OL_ERR_DEVICE, "offset upload failed"));
    mark(c, "query: upload text + offsets");
    {
        Prof p(c, "parse_given");
        if (ol_launch_parse_given(d.text, (uint32_t)len, d.offsets, line_count, d.nodes, c->stream) != hipSuccess ||
            ol_launch_nodes_to_lines(d.nodes, line_count, d.classes, d.keys, c->stream) != hipSuccess)
            return done(fail(OL_ERR_DEVICE, "parse launch failed"));
    }
    return done(run_query(c, &d, 0, (uint32_t)line_count, kind, depth, max_matches, total, indices, n_indices));
}

int ol_search_device(ol_ctx *c, const ol_doc *d, const uint8_t *pattern, size_t len, uint64_t *d_offsets,
                     size_t cap, uint64_t *total)
{
    if (!c || !d || !total || (!pattern && len))
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    int rc = check_pattern(len);
    if (rc)
        return rc;
    *total = 0;
    if (d->n == 0)
        return OL_OK;
    return run_search(c, d, pattern, len, d_offsets, nullptr, d_offsets ? cap : 0, total);
}

int ol_search_device_noalloc(ol_ctx *c, const ol_doc *d, const uint8_t *pattern, size_t len, uint64_t *d_offsets,
                             size_t cap, uint64_t *total)
{
    if (!c || !d || !total || (!pattern && len))
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    int rc = check_pattern(len);
    if (rc)
        return rc;
    *total = 0;
    if (d->n == 0)
        return OL_OK;
    return run_search_noalloc(c, d, pattern, len, d_offsets, d_offsets ? cap : 0, total);
}

// ---- dictionary ------------------------------------------------------------------
int ol_dict_compile(ol_ctx *c, const ol_doc *doc, const uint8_t *const *patterns, const size_t *lens,
                    size_t count, ol_dict **out)
{
    (void)doc;
    if (!c || !out || (count && (!patterns || !lens)))
        return fail(OL_ERR_INVALID, "null argument");
    if (count == 0 || count > OL_MAX_PATTERNS)
        return fail(OL_ERR_TOO_MANY, "need 1..=" + std::to_string(OL_MAX_PATTERNS) + " patterns, got " +
                                         std::to_string(count));
    for (size_t i = 0; i < count; i++) {
        if (lens[i] < OL_MIN_SET_PATTERN_LEN || lens[i] > OL_MAX_SET_PATTERN_LEN || !patterns[i])
            return fail(OL_ERR_PATTERN, "pattern " + std::to_string(i) + " is " + std::to_string(lens[i]) +
                                            " bytes; each must be " + std::to_string(OL_MIN_SET_PATTERN_LEN) +
                                            "..=" + std::to_string(OL_MAX_SET_PATTERN_LEN));
    }
    std::lock_guard<std::mutex> g(c->mu);
    // Key class and key of each pattern: class DK_LONG for 4 bytes and more
    // (key: the first 4), DK_3 / DK_2 / DK_1 for 3, 2 and 1 bytes (key: all
    // of them).  Patterns with equal class and key form a run, sorted by id.
    auto word = [&](const uint8_t *p, size_t from, size_t n) {
        uint32_t v = 0;
        for (size_t b = 0; b < n; b++)
            v |= (uint32_t)p[from + b] << (8 * b);
        return v;
    };
    std::vector<uint32_t> cls(count), key(count), order(count);
    uint32_t max_len = 0;
    for (size_t i = 0; i < count; i++) {
        const size_t L = lens[i];
        cls[i] = L >= 4 ? DK_LONG : (L == 3 ? DK_3 : (L == 2 ? DK_2 : DK_1));
        key[i] = word(patterns[i], 0, std::min<size_t>(L, 4));
        order[i] = (uint32_t)i;
        max_len = std::max<uint32_t>(max_len, (uint32_t)L);
    }
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        if (cls[a] != cls[b]) return cls[a] < cls[b];
        if (key[a] != key[b]) return key[a] < key[b];
        return a < b;
    });
    // A pattern's bytes 4..7 (w1) and what follows: bytes 8..11 (w2) when it
    // ends by byte 12, else the pool offset of its bytes 8.. (word aligned).
    std::vector<uint32_t> pool;
    auto tail = [&](uint32_t i, uint32_t &w1, uint32_t &aux) {
        const size_t L = lens[i];
        w1 = L > 4 ? word(patterns[i], 4, std::min<size_t>(L - 4, 4)) : 0u;
        if (L <= 12) {
            aux = L > 8 ? word(patterns[i], 8, L - 8) : 0u;
        } else {
            aux = (uint32_t)pool.size();
            for (size_t b = 8; b < L; b += 4)
                pool.push_back(word(patterns[i], b, std::min<size_t>(4, L - b)));
        }
    };
    auto info_of = [&](uint32_t i) { return dk_info(i, (uint32_t)lens[i], cls[i]); };
    // slots (one per distinct class and key) and the entries of runs
    struct Slot {
        uint32_t key, info, a, b;
    };
    DictDev dd{};
    // groups: the patterns of one (class, key), a range of `order`
    struct Group {
        uint32_t cls, key;
        size_t b, e;
    };
    std::vector<Group> groups;
    for (size_t e = 0; e < count;) {
        const uint32_t i = order[e];
        size_t f = e + 1;
        while (f < count && cls[order[f]] == cls[i] && key[order[f]] == key[i])
            f++;
        groups.push_back({cls[i], key[i], e, f});
        e = f;
    }
    // Merged runs: the run of a key also holds the shorter patterns that are
    // prefixes of it (the groups of classes 3, 2, 1 on its first 3, 2, 1
    // bytes), so one run per position holds all its candidates.  Short
    // patterns are copied into every longer key that extends them; when that
    // would more than double the entries, runs keep their own class only.
    std::unordered_map<uint64_t, uint32_t> gidx;
    gidx.reserve(groups.size() * 2);
    for (size_t g = 0; g < groups.size(); g++)
        gidx[(uint64_t)groups[g].cls << 32 | groups[g].key] = (uint32_t)g;
    auto prefix_groups = [&](const Group &g, auto &&visit) {
        for (uint32_t c = g.cls + 1; c <= DK_1; c++) {   // DK_3, DK_2, DK_1: 3, 2, 1 bytes
            const uint32_t nb = 4 - c;
            const auto it = gidx.find((uint64_t)c << 32 | (g.key & ((1u << (8 * nb)) - 1u)));
            if (it != gidx.end())
                visit(groups[it->second]);
        }
    };
    {
        uint64_t total = 0;
        for (const Group &g : groups) {
            total += g.e - g.b;
            prefix_groups(g, [&](const Group &s) { total += s.e - s.b; });
        }
        dd.merged = total <= 2 * (uint64_t)count + 65536 ? 1u : 0u;
    }
    std::vector<Slot> slots;
    std::vector<uint32_t> ents;   // 4 words per entry: info, w1, aux, 0
    std::vector<uint32_t> mem;
    for (const Group &g : groups) {
        mem.assign(order.begin() + g.b, order.begin() + g.e);
        if (dd.merged) {
            prefix_groups(g, [&](const Group &s) { mem.insert(mem.end(), order.begin() + s.b, order.begin() + s.e); });
            std::sort(mem.begin(), mem.end());
        }
        Slot sl{g.key, 0, 0, 0};
        if (mem.size() == 1) {   // one pattern (of the key's own class): decided from the slot
            sl.info = info_of(mem[0]);
            tail(mem[0], sl.a, sl.b);
        } else {                 // a run of entries, in id order
            sl.info = dk_info(0, 0, g.cls);
            sl.a = (uint32_t)(ents.size() / 4);
            sl.b = (uint32_t)mem.size();
            for (uint32_t p : mem) {
                uint32_t w1, aux;
                tail(p, w1, aux);
                ents.insert(ents.end(), {info_of(p), w1, aux, 0u});
            }
        }
        slots.push_back(sl);
    }
    pool.resize(pool.size() + 16, 0);
    if (ents.empty())
        ents.assign(4, 0);
    // Cuckoo placement: each slot goes to one of its key's two buckets (an
    // empty one, else it displaces a resident to that one's other bucket);
    // the table (at most about 45% full) doubles until every slot has a place.
    uint32_t bbits = 2;
    while (((size_t)DK_BUCKET << bbits) * 9 < slots.size() * 20)
        bbits++;
    std::vector<int32_t> place;   // slot index per table position, -1 empty
    for (;; bbits++) {
        const size_t nb = (size_t)1 << bbits;
        place.assign(nb * DK_BUCKET, -1);
        std::vector<uint8_t> fillc(nb, 0);
        bool ok = true;
        std::mt19937 rng(12345);
        for (size_t k = 0; k < slots.size() && ok; k++) {
            int32_t cur = (int32_t)k;
            uint32_t prev_bucket = ~0u;
            for (int step = 0; step < 500; step++) {
                const Slot &sl = slots[cur];
                const uint32_t cl = sl.info >> DK_CLASS_SHIFT;
                const uint32_t b1 = dk_bucket1(sl.key, cl, bbits), b2 = dk_bucket2(sl.key, cl, bbits);
                uint32_t tgt = fillc[b1] <= fillc[b2] ? b1 : b2;
                if (fillc[tgt] < DK_BUCKET) {
                    place[(size_t)tgt * DK_BUCKET + fillc[tgt]++] = cur;
                    cur = -1;
                    break;
                }
                // both full: evict a random resident of the bucket not just left
                tgt = (b1 == prev_bucket) ? b2 : (b2 == prev_bucket ? b1 : (rng() & 1 ? b1 : b2));
                const uint32_t victim_at = tgt * DK_BUCKET + (rng() % DK_BUCKET);
                const int32_t victim = place[victim_at];
                place[victim_at] = cur;
                cur = victim;
                prev_bucket = tgt;
            }
            if (cur >= 0)
                ok = false;
        }
        if (ok)
            break;
    }
    std::vector<uint32_t> tab(((size_t)DK_BUCKET << bbits) * 4, 0);   // empty: info 0 (not DK_VALID)
    for (size_t q = 0; q < place.size(); q++) {
        if (place[q] < 0)
            continue;
        const Slot &sl = slots[place[q]];
        tab[q * 4 + 0] = sl.key;
        tab[q * 4 + 1] = sl.info;
        tab[q * 4 + 2] = sl.a;
        tab[q * 4 + 3] = sl.b;
    }
    dd.max_len = max_len;
    dd.bbits = bbits;
    for (size_t i = 0; i < count; i++)
        dd.class_mask |= 1u << cls[i];
    // filters: 2 = a direct bitmap of the 2-byte window (a 2-byte pattern's
    // key; a 1-byte pattern's byte with each of the 256 second bytes), so a
    // position takes one lookup for both short classes; 4 = the prefix filter
    // of patterns of 3 bytes and more (4-byte keys, a 3-byte key with each of
    // the 32 values of the fourth byte's low 5 bits).  Layout: filter 4 first,
    // then 2 (the kernels find it right after filter 4), then a zero word that
    // absent filters point at.
    std::vector<std::vector<uint32_t>> filt(6);
    uint32_t fwords = 0;
    const bool any3 = (dd.class_mask >> DK_3) & 1u, anylong = (dd.class_mask >> DK_LONG) & 1u;
    // prefix filter keying: sparse (free bit offset) unless the filter would
    // hold more than DICT_PF_DENSE_BITS bits
    {
        uint64_t bits = 0;
        for (size_t i = 0; i < count; i++)
            bits += cls[i] == DK_3 ? 32 : cls[i] == DK_LONG ? 1 : 0;
        dd.pf_hashed = bits > DICT_PF_DENSE_BITS ? 1u : 0u;
    }
    const bool anyshort = ((dd.class_mask >> DK_1) & 1u) || ((dd.class_mask >> DK_2) & 1u);
    for (uint32_t f : {4u, 2u}) {
        const bool used = f == 4 ? (any3 || anylong) : anyshort;
        if (!used)
            continue;
        const uint32_t fb = f == 2 ? 16 : DICT_PBITS;
        dd.fbits[f] = fb;
        dd.foff[f] = fwords;
        const uint32_t words = (1u << fb) / 32;
        filt[f].assign(words, 0);
        fwords += words;
        for (size_t i = 0; i < count; i++) {
            if (f == 2) {
                if (cls[i] == DK_2)
                    filt[f][key[i] >> 5] |= 1u << (key[i] & 31);
                else if (cls[i] == DK_1)   // the byte, then any second byte
                    for (uint32_t b1 = 0; b1 < 256; b1++) {
                        const uint32_t k2 = key[i] | b1 << 8;
                        filt[f][k2 >> 5] |= 1u << (k2 & 31);
                    }
                continue;
            }
            auto setkey = [&](uint32_t w) {
                const uint32_t h = dict_ph(w);
                filt[f][dict_ph_word(h)] |= 1u << (h & 31);
            };
            if (dd.pf_hashed) {
                if (cls[i] == DK_3)   // every value of the window's byte 3
                    for (uint32_t c4 = 0; c4 < 32; c4++) setkey(key[i] | c4 << 24);
                else if (cls[i] == DK_LONG)
                    setkey(key[i]);
            } else if (cls[i] == DK_3) {   // every value of the window's byte 3
                filt[f][dict_pword(key[i])] = ~0u;
            } else if (cls[i] == DK_LONG) {
                filt[f][dict_pword(key[i])] |= 1u << dict_pbit(key[i]);
            }
        }
    }
    for (uint32_t f = 1; f <= 4; f++)
        if (!dd.fbits[f])
            dd.foff[f] = fwords;   // the zero word
    fwords += 1;
    dd.fwords = fwords;
    // One device allocation: [filter][table][entries][pool]
    size_t bytes = 0;
    auto placeb = [&](size_t n) {
        size_t o = round_up(bytes, 256);
        bytes = o + n;
        return o;
    };
    const size_t o_f = placeb((size_t)fwords * 4);
    const size_t o_t = placeb(tab.size() * 4);
    const size_t o_ent = placeb(ents.size() * 4);
    const size_t o_pool = placeb(pool.size() * 4);
    const size_t o_ids = placeb(ents.size());   // one u32 per 16-byte entry
    ol_dict *s = new (std::nothrow) ol_dict();
    if (!s)
        return fail(OL_ERR_NOMEM, "out of host memory");
    if (hipMalloc(&s->dev, bytes) != hipSuccess) {
        delete s;
        return fail(OL_ERR_NOMEM, "device allocation of the dictionary failed");
    }
    s->bytes = bytes;
    uint8_t *base = (uint8_t *)s->dev;
    std::vector<uint8_t> host(bytes, 0);
    for (uint32_t f = 1; f <= 4; f++)
        if (!filt[f].empty())
            std::memcpy(host.data() + o_f + (size_t)dd.foff[f] * 4, filt[f].data(), filt[f].size() * 4);
    std::memcpy(host.data() + o_t, tab.data(), tab.size() * 4);
    std::memcpy(host.data() + o_ent, ents.data(), ents.size() * 4);
    std::memcpy(host.data() + o_pool, pool.data(), pool.size() * 4);
    for (size_t e = 0; e < ents.size() / 4; e++) {
        const uint32_t id = ents[4 * e] & 0x3FFFFFu;
        std::memcpy(host.data() + o_ids + 4 * e, &id, 4);
    }
    if (hipMemcpy(base, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(s->dev);
        delete s;
        return fail(OL_ERR_DEVICE, "dictionary upload failed");
    }
    dd.filter = (const uint32_t *)(base + o_f);
    dd.table = (const uint4 *)(base + o_t);
    dd.ents = (const uint4 *)(base + o_ent);
    dd.pool = (const uint32_t *)(base + o_pool);
    dd.ent_ids = (const uint32_t *)(base + o_ids);
    s->d = dd;
    s->count = (uint32_t)count;
    *out = s;
    return OL_OK;
}

void ol_dict_free(ol_dict *s)
{
    if (!s)
        return;
    if (s->dev)
        (void)hipFree(s->dev);
    delete s;
}

size_t ol_dict_len(const ol_dict *s) { return s ? s->count : 0; }

int ol_search_many_device(ol_ctx *c, const ol_doc *d, const ol_dict *s, uint64_t *d_offsets, uint32_t *d_ids,
                          size_t cap, uint64_t *total)
{
    if (!c || !d || !s || !total)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    *total = 0;
    if (d->n == 0)
        return OL_OK;
    return run_many(c, d, s, d_offsets, d_ids, (d_offsets || d_ids) ? cap : 0, total);
}

int ol_search_many(ol_ctx *c, const ol_doc *d, const ol_dict *s, size_t max_hits, uint32_t *counts,
                   uint64_t *total_hits, ol_many_hit **hits, size_t *n_hits)
{
    if (!c || !d || !s || !counts || !total_hits || !hits || !n_hits)
        return fail(OL_ERR_INVALID, "null argument");
    if (max_hits > 0xFFFFFFFFull)
        return fail(OL_ERR_TOO_MANY, "max_hits " + std::to_string(max_hits) + " does not fit in u32");
    std::lock_guard<std::mutex> g(c->mu);
    std::memset(counts, 0, (size_t)s->count * 4);
    *total_hits = 0;
    *hits = nullptr;
    *n_hits = 0;
    if (d->n == 0)
        return OL_OK;
    uint64_t total = 0;
    int rc = run_many(c, d, s, nullptr, nullptr, 0, &total);
    if (rc)
        return rc;
    *total_hits = total;
    if (total == 0)
        return OL_OK;
    // Exact per-pattern counts come from the full offset-major list.
    std::vector<uint32_t> ids(total);
    std::vector<uint64_t> offs64(total);
    {
        DevBuf d64, did;
        if (d64.ensure(total * 8) != hipSuccess || did.ensure(total * 4) != hipSuccess)
            return fail(OL_ERR_NOMEM, "device allocation of the hit list failed");
        uint64_t again = 0;
        rc = run_many(c, d, s, (uint64_t *)d64.p, (uint32_t *)did.p, total, &again);
        if (!rc && (hipMemcpy(offs64.data(), d64.p, total * 8, hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(ids.data(), did.p, total * 4, hipMemcpyDeviceToHost) != hipSuccess))
            rc = fail(OL_ERR_DEVICE, "hit list download failed");
        d64.release();
        did.release();
        if (rc)
            return rc;
    }
    for (uint64_t i = 0; i < total; i++)
        counts[ids[i]]++;
    mark(c, "many: read back the hits + count per pattern");
    if (max_hits == 0)
        return OL_OK;
    // Stable counting sort by pattern: the list is offset-ordered, so the result
    // is sorted by (pattern, offset) (ManyResult::hits, crates/oneline-gpu/src/many.rs:89-93).
    std::vector<uint64_t> start(s->count + 1, 0);
    for (uint32_t p = 0; p < s->count; p++)
        start[p + 1] = start[p] + counts[p];
    uint64_t keep = std::min<uint64_t>(total, max_hits);
    ol_many_hit *h = (ol_many_hit *)std::malloc(keep * sizeof(ol_many_hit));
    if (!h)
        return fail(OL_ERR_NOMEM, "out of host memory");
    std::vector<uint64_t> pos(start.begin(), start.end() - 1);
    for (uint64_t i = 0; i < total; i++) {
        uint64_t at = pos[ids[i]]++;
        if (at < keep)
            h[at] = ol_many_hit{ids[i], (uint32_t)offs64[i]};
    }
    *hits = h;
    *n_hits = keep;
    mark(c, "many: sort the hits by pattern");
    return OL_OK;
}

// ---- launch tuning ----------------------------------------------------------------
int ol_tune_set(ol_ctx *c, const char *name, int value)
{
    if (!c || !name)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    std::string n(name);
    if (n == "scan_bpt") {
        if (value != 64 && value != 128 && value != 256)
            return fail(OL_ERR_INVALID, "scan_bpt must be 64, 128 or 256");
        c->scan_bpt = (uint32_t)value;
    } else if (n == "scan_grid") {
        if (value < 0)
            return fail(OL_ERR_INVALID, "scan_grid must be >= 0 (0 = all co-resident workgroups)");
        c->scan_grid = (uint32_t)value;
    } else if (n == "scan_ticket") {
        c->scan_ticket = value != 0;
    } else if (n == "tree_rounds") {
        if (value != 1 && value != 4 && value != 16)
            return fail(OL_ERR_INVALID, "tree_rounds must be 1, 4 or 16");
        c->tree_rounds = (uint32_t)value;
    } else {
        return fail(OL_ERR_INVALID, "unknown tuning parameter " + n);
    }
    return OL_OK;
}

int ol_tune_get(ol_ctx *c, const char *name, int *value)
{
    if (!c || !name || !value)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    std::string n(name);
    if (n == "scan_bpt")
        *value = (int)c->scan_bpt;
    else if (n == "scan_grid")
        *value = (int)c->scan_grid;
    else if (n == "tree_rounds")
        *value = (int)c->tree_rounds;
    else if (n == "scan_ticket")
        *value = c->scan_ticket ? 1 : 0;
    else
        return fail(OL_ERR_INVALID, "unknown tuning parameter " + n);
    return OL_OK;
}

// ---- profiling --------------------------------------------------------------------
int ol_profile_enable(ol_ctx *c, int on)
{
    if (!c)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    c->profiling = on != 0;
    return OL_OK;
}

int ol_profile_take(ol_ctx *c, char *names, size_t name_stride, uint64_t *counts, double *ms, size_t cap,
                    size_t *n)
{
    if (!c || !n)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<std::string> keys;
    std::vector<uint64_t> cnt;
    std::vector<double> sum;
    for (auto &r : c->recs) {
        float t = 0.f;
        (void)hipEventElapsedTime(&t, r.a, r.b);
        size_t i = 0;
        while (i < keys.size() && keys[i] != r.name)
            i++;
        if (i == keys.size()) {
            keys.push_back(r.name);
            cnt.push_back(0);
            sum.push_back(0.0);
        }
        cnt[i]++;
        sum[i] += t;
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->recs.clear();
    size_t m = std::min(cap, keys.size());
    for (size_t i = 0; i < m; i++) {
        if (names && name_stride) {
            std::snprintf(names + i * name_stride, name_stride, "%s", keys[i].c_str());
        }
        if (counts) counts[i] = cnt[i];
        if (ms) ms[i] = sum[i];
    }
    *n = keys.size();
    return OL_OK;
}

int ol_profile_call_begin(ol_ctx *c)
{
    if (!c)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    for (auto &r : c->call_recs) {
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->call_recs.clear();
    c->marks.clear();
    c->call_prof = true;
    mark(c, "start");
    return OL_OK;
}

int ol_profile_call_end(ol_ctx *c, double *wall_ms, ol_profile_step *host_steps, size_t host_cap, size_t *n_host,
                        ol_profile_step *gpu_passes, size_t gpu_cap, size_t *n_gpu)
{
    if (!c || !wall_ms || !n_host || !n_gpu)
        return fail(OL_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->mu);
    if (!c->call_prof)
        return fail(OL_ERR_INVALID, "no profiled call: ol_profile_call_begin first");
    c->call_prof = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    const auto &m = c->marks;
    *wall_ms = m.size() >= 2 ? std::chrono::duration<double, std::milli>(m.back().second - m.front().second).count()
                             : 0.0;
    *n_host = m.size() ? m.size() - 1 : 0;
    for (size_t i = 1; i < m.size() && i - 1 < host_cap && host_steps; i++) {
        std::snprintf(host_steps[i - 1].label, sizeof host_steps[i - 1].label, "%s", m[i].first);
        host_steps[i - 1].ms = std::chrono::duration<double, std::milli>(m[i].second - m[i - 1].second).count();
    }
    *n_gpu = c->call_recs.size();
    for (size_t i = 0; i < c->call_recs.size(); i++) {
        const auto &r = c->call_recs[i];
        if (gpu_passes && i < gpu_cap) {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, r.a, r.b);
            std::snprintf(gpu_passes[i].label, sizeof gpu_passes[i].label, "%s", r.name);
            gpu_passes[i].ms = t;
        }
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->call_recs.clear();
    c->marks.clear();
    return OL_OK;
}

}  // extern "C"

// ---- batches ------------------------------------------------------------------------
// Many documents in one launch sequence (no counterpart in the reference,
// whose every call handles one document).  The documents lie back to back in
// one HBM text, each followed by a '\n' unless it already ends in one (an
// unterminated last line keeps its start, class and key: the trailing '\n' is
// trimmed like any line end); empty documents occupy nothing and have no
// lines.  The scan runs over the whole text and writes document-relative
// offsets plus each document's first line; the tree runs one segment per
// document (tiles never cross documents); the dictionary search rejects
// matches that run past their document.  The launch sequence is sized once
// (host round trips) and then replayed, optionally as a captured hipGraph.

struct ol_batch {
    size_t count = 0;                     // documents, empty ones included
    std::vector<uint64_t> h_start, h_len; // per document: byte start in the batch text, length
    std::vector<uint32_t> h_dev;          // per document: index among the non-empty ones (or ~0u)
    uint32_t nd = 0;                      // non-empty documents
    uint64_t n = 0;                       // batch text bytes (documents and separators)
    uint64_t payload = 0;                 // document bytes
    uint8_t *text = nullptr;
    DevBuf d_starts, d_ends, d_line_start;   // [nd]
    DevBuf scan_tile_doc, search_tile_doc;
    uint32_t scan_tile_doc_bpt = 0;
    // the batch's own scratch: a captured graph keeps its addresses, so it is
    // sized for every launch form at the sizing run and then frozen
    mutable Scratch scr;
    hipStream_t stream = nullptr;   // the stream of the last run (a context's)
    // per-line outputs (global line order)
    uint64_t cap = 0, lines = 0;
    DevBuf offsets, classes, keys, parent, fc, ns, se;
    std::vector<uint32_t> h_line_start, h_line_count;   // [nd]
    // tree segments
    DevBuf seg_line0, seg_lines, seg_tile0, tile_seg;
    uint32_t tree_tiles = 0;
    // dictionary matches, sorted by (global offset, id)
    DevBuf m_off, m_id;
    uint64_t m_cap = 0, m_total = 0;
    std::vector<uint64_t> h_moff;
    std::vector<uint32_t> h_mid;
    bool h_matches = false;
    // the sized launch sequence
    bool prepared = false, prep_search_only = false;
    uint32_t prep_bpt = 0, prep_rounds = 0;
    const ol_dict *prep_dict = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t scan_node = nullptr;   // the captured scan (marked cooperative before instantiation)
    size_t graph_words = 0;   // status words a replay leaves set

    void drop_graph()
    {
        if ((exec || graph) && stream)
            (void)hipStreamSynchronize(stream);   // a replay may still run
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

// The documents of a batch laid back to back, each followed by '\n' unless it
// ends in one (empty documents take no bytes).
struct BatchLayout {
    const uint8_t *const *docs = nullptr;
    const size_t *lens = nullptr;
    size_t count = 0;
    std::vector<uint64_t> start;   // per document: byte start in the batch text
    std::vector<uint32_t> dev;     // per document: index among the non-empty ones (or ~0u)
    std::vector<size_t> order;     // the non-empty documents, in order
    uint32_t nd = 0;
    uint64_t n = 0, payload = 0;

    int build(const uint8_t *const *d, const size_t *l, size_t cnt)
    {
        docs = d;
        lens = l;
        count = cnt;
        start.assign(cnt, 0);
        dev.assign(cnt, ~0u);
        order.clear();
        nd = 0;
        n = payload = 0;
        for (size_t i = 0; i < cnt; i++) {
            if (l[i] && !d[i])
                return fail(OL_ERR_INVALID, "document " + std::to_string(i) + " is null");
            start[i] = n;
            payload += l[i];
            if (l[i]) {
                dev[i] = nd++;
                order.push_back(i);
                n += l[i] + (d[i][l[i] - 1] != '\n' ? 1 : 0);
            }
        }
        if (n > OL_MAX_GPU_INPUT_LEN)
            return fail(OL_ERR_TOO_LARGE, "batch text is " + std::to_string(n) + " bytes; at most " +
                                              std::to_string(OL_MAX_GPU_INPUT_LEN) + " (32-bit offsets)");
        return OL_OK;
    }
    // bytes [off, off + k) of the batch text into buf
    void fill(uint8_t *buf, uint64_t off, uint64_t k) const
    {
        // the last non-empty document starting at or before off
        size_t lo = 0, hi = order.size();
        while (hi - lo > 1) {
            const size_t mid = (lo + hi) / 2;
            if (start[order[mid]] <= off) lo = mid; else hi = mid;
        }
        uint64_t done = 0;
        for (size_t j = lo; done < k && j < order.size(); j++) {
            const size_t i = order[j];
            const uint64_t s0 = start[i];
            const uint64_t span = lens[i] + (docs[i][lens[i] - 1] != '\n' ? 1 : 0);
            const uint64_t at = off + done;
            if (at >= s0 + span)
                continue;
            const uint64_t in = at - s0;
            if (in < lens[i]) {
                const uint64_t m = std::min<uint64_t>(lens[i] - in, k - done);
                std::memcpy(buf + done, docs[i] + in, m);
                done += m;
            }
            if (done < k && off + done == s0 + lens[i] && span > lens[i])
                buf[done++] = '\n';
        }
    }
    // starts and ends of the non-empty documents
    void bounds(uint32_t *st, uint32_t *en) const
    {
        for (size_t j = 0; j < order.size(); j++) {
            if (st) st[j] = (uint32_t)start[order[j]];
            if (en) en[j] = (uint32_t)(start[order[j]] + lens[order[j]]);
        }
    }
    // scan tile t -> first non-empty document starting at or after it (T + 1 entries)
    void scan_tile_doc(uint64_t tile_bytes, uint32_t *td) const
    {
        const uint32_t T = (uint32_t)ceil_div(n, tile_bytes);
        uint32_t j = 0;
        for (uint32_t t = 0; t <= T; t++) {
            const uint64_t pos = (uint64_t)t * tile_bytes;
            while (j < nd && start[order[j]] < pos)
                j++;
            td[t] = j;
        }
    }
    // search tile t -> last non-empty document starting at or before it
    void search_tile_doc(uint32_t *sd) const
    {
        const uint32_t TS = (uint32_t)ceil_div(n ? n : 1, SEARCH_TILE);
        uint32_t j = 0;
        for (uint32_t t = 0; t < TS; t++) {
            const uint64_t p = (uint64_t)t * SEARCH_TILE;
            while (j + 1 < nd && start[order[j + 1]] <= p)
                j++;
            sd[t] = j;
        }
    }
};

// scan of the batch text (outputs sized by b->cap)
int batch_scan(ol_ctx *c, ol_batch *b, bool cooperative)
{
    ScanArgs a{};
    a.text = b->text;
    a.n = (uint32_t)b->n;
    a.cap = b->cap;
    a.offsets = (uint32_t *)b->offsets.p;
    a.classes = (uint8_t *)b->classes.p;
    a.keys = (uint8_t *)b->keys.p;
    a.ndocs = b->nd;
    a.doc_starts = (const uint32_t *)b->d_starts.p;
    a.tile_doc = (const uint32_t *)b->scan_tile_doc.p;
    a.doc_line_start = (uint32_t *)b->d_line_start.p;
    return launch_scan_args(c, b->scr, a, cooperative);
}

int batch_tree(ol_ctx *c, ol_batch *b)
{
    const size_t sum_bytes = ol_tree_sum_words(b->tree_tiles) * 4;
    if (b->scr.tree_sum.bytes < sum_bytes && b->scr.frozen)
        return fail(OL_ERR_DEVICE, "batch tree scratch would move under a captured graph");
    HIPCHK(b->scr.tree_sum.ensure(sum_bytes));
    uint32_t *base = (uint32_t *)b->scr.tree_sum.p;
    TreeArgs a{};
    a.keys = (const uint8_t *)b->keys.p;
    a.lines = b->lines;
    a.num_tiles = b->tree_tiles;
    a.parent = (uint32_t *)b->parent.p;
    a.first_child = (uint32_t *)b->fc.p;
    a.next_sibling = (uint32_t *)b->ns.p;
    a.subtree_end = (uint32_t *)b->se.p;
    ol_tree_set_sums(a, base, b->tree_tiles);
    a.err = b->scr.ctr() + CTR_TREE_ERR;
    a.nsegs = b->nd;
    a.seg_line0 = (const uint32_t *)b->seg_line0.p;
    a.seg_lines = (const uint32_t *)b->seg_lines.p;
    a.seg_tile0 = (const uint32_t *)b->seg_tile0.p;
    a.tile_seg = (const uint32_t *)b->tile_seg.p;
    Prof p(c, "tree");
    HIPCHK(ol_launch_tree(a, c->tree_rounds, c->stream));
    return OL_OK;
}

int batch_many(ol_ctx *c, ol_batch *b, const ol_dict *s, uint64_t cap)
{
    ManyArgs a{};
    a.text = b->text;
    a.n = (uint32_t)b->n;
    a.d = s->d;
    a.cap = cap;
    a.out_off = cap ? (uint64_t *)b->m_off.p : nullptr;
    a.out_id = cap ? (uint32_t *)b->m_id.p : nullptr;
    HIPCHK(setup_many(c, b->scr, a, b->n, cap != 0));
    a.ndocs = b->nd;
    a.doc_starts = (const uint32_t *)b->d_starts.p;
    a.doc_ends = (const uint32_t *)b->d_ends.p;
    a.tile_doc = (const uint32_t *)b->search_tile_doc.p;
    Prof p(c, "many");
    HIPCHK(ol_launch_many(a, c->stream));
    return OL_OK;
}

int alloc_batch_lines(ol_batch *b, uint64_t cap)
{
    const uint64_t capa = round_up(cap ? cap : 1, LINE_ALIGN);
    HIPCHK(b->offsets.ensure(capa * 4));
    HIPCHK(b->classes.ensure(capa));
    HIPCHK(b->keys.ensure(capa));
    HIPCHK(b->parent.ensure(capa * 4));
    HIPCHK(b->fc.ensure(capa * 4));
    HIPCHK(b->ns.ensure(capa * 4));
    HIPCHK(b->se.ensure(capa * 4));
    b->cap = capa;
    return OL_OK;
}

// Sizes and runs the whole sequence once, with host round trips.
// Parse + tree part of the sizing run.
int batch_prepare_parse(ol_ctx *c, ol_batch *b)
{
    // scan tiles -> first document starting at or after each tile
    const uint64_t stile = (uint64_t)SCAN_THREADS * c->scan_bpt;
    if (b->scan_tile_doc_bpt != c->scan_bpt) {
        const uint32_t T = (uint32_t)ceil_div(b->n, stile);
        std::vector<uint32_t> td(T + 1);
        uint32_t j = 0;
        // starts of the non-empty documents, ascending
        std::vector<uint32_t> st;
        st.reserve(b->nd);
        for (size_t i = 0; i < b->count; i++)
            if (b->h_dev[i] != ~0u)
                st.push_back((uint32_t)b->h_start[i]);
        j = 0;
        for (uint32_t t = 0; t <= T; t++) {
            const uint64_t pos = (uint64_t)t * stile;
            while (j < b->nd && st[j] < pos)
                j++;
            td[t] = j;
        }
        HIPCHK(b->scan_tile_doc.ensure(td.size() * 4));
        HIPCHK(hipMemcpy(b->scan_tile_doc.p, td.data(), td.size() * 4, hipMemcpyHostToDevice));
        b->scan_tile_doc_bpt = c->scan_bpt;
    }
    // scan (re-run with a larger capacity if the estimate was short)
    for (int attempt = 0; attempt < 2; attempt++) {
        int rc = batch_scan(c, b, true);
        if (rc)
            return rc;
        uint64_t lines = 0;
        rc = read_meta(c, b->scr, &lines);
        if (rc)
            return rc;
        b->lines = lines;
        if (lines <= b->cap)
            break;
        rc = alloc_batch_lines(b, lines);
        if (rc)
            return rc;
    }
    if (b->lines > b->cap)
        return fail(OL_ERR_DEVICE, "batch line count exceeds the sized capacity");
    // per-document line ranges -> tree segments
    b->h_line_start.assign(b->nd, 0);
    HIPCHK(hipMemcpy(b->h_line_start.data(), b->d_line_start.p, (size_t)b->nd * 4, hipMemcpyDeviceToHost));
    b->h_line_count.assign(b->nd, 0);
    const uint64_t TL = ol_tree_tile_lines(c->tree_rounds);
    std::vector<uint32_t> seg_tile0(b->nd), tile_seg;
    uint64_t tiles = 0;
    for (uint32_t j = 0; j < b->nd; j++) {
        const uint64_t end = (j + 1 < b->nd) ? b->h_line_start[j + 1] : b->lines;
        b->h_line_count[j] = (uint32_t)(end - b->h_line_start[j]);
        seg_tile0[j] = (uint32_t)tiles;
        const uint64_t k = ceil_div(b->h_line_count[j] ? b->h_line_count[j] : 1, TL);
        for (uint64_t q = 0; q < k; q++)
            tile_seg.push_back(j);
        tiles += k;
    }
    b->tree_tiles = (uint32_t)tiles;
    HIPCHK(b->seg_line0.ensure((size_t)b->nd * 4));
    HIPCHK(b->seg_lines.ensure((size_t)b->nd * 4));
    HIPCHK(b->seg_tile0.ensure((size_t)b->nd * 4));
    HIPCHK(b->tile_seg.ensure(tile_seg.size() * 4));
    HIPCHK(hipMemcpy(b->seg_line0.p, b->h_line_start.data(), (size_t)b->nd * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->seg_lines.p, b->h_line_count.data(), (size_t)b->nd * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->seg_tile0.p, seg_tile0.data(), (size_t)b->nd * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->tile_seg.p, tile_seg.data(), tile_seg.size() * 4, hipMemcpyHostToDevice));
    int rc = batch_tree(c, b);
    if (rc)
        return rc;
    rc = check_tree_error(c, b->scr);
    if (rc)
        return rc;
    return OL_OK;
}

int batch_prepare(ol_ctx *c, ol_batch *b, const ol_dict *dict, bool search_only)
{
    b->drop_graph();
    b->stream = c->stream;
    b->scr.frozen = false;
    b->prepared = false;
    b->h_matches = false;
    b->prep_search_only = search_only;
    if (b->nd == 0) {
        b->lines = 0;
        b->m_total = 0;
        b->prepared = true;
        b->prep_bpt = c->scan_bpt;
        b->prep_rounds = c->tree_rounds;
        b->prep_dict = dict;
        return OL_OK;
    }
    if (!search_only) {
        const int prc = batch_prepare_parse(c, b);
        if (prc)
            return prc;
    }
    // dictionary: count, size, write
    b->m_total = 0;
    if (dict) {
        int rc;
        rc = batch_many(c, b, dict, 0);
        if (rc)
            return rc;
        uint64_t total = 0;
        rc = read_meta(c, b->scr, &total);
        if (rc)
            return rc;
        HIPCHK(b->m_off.ensure((total ? total : 1) * 8));
        HIPCHK(b->m_id.ensure((total ? total : 1) * 4));
        b->m_cap = total;
        rc = batch_many(c, b, dict, total ? total : 1);
        if (rc)
            return rc;
        rc = read_meta(c, b->scr, &total);
        if (rc)
            return rc;
        b->m_total = total;
    }
    // every launch form of this geometry (cooperative scan; ticket-ordered
    // scan, as a context set to it uses) fits the scratch from now on: freeze it
    size_t words = STATUS_TILES;
    if (!search_only)
        words = std::max(scan_status_words(c, b->n, true), scan_status_words(c, b->n, false));
    if (dict)
        words = std::max(words, many_status_words(b->n));
    HIPCHK(ensure_status(b->scr, words));
    b->graph_words = words;
    b->scr.frozen = true;
    b->prepared = true;
    b->prep_bpt = c->scan_bpt;
    b->prep_rounds = c->tree_rounds;
    b->prep_dict = dict;
    return OL_OK;
}

// The sized sequence, launched asynchronously (no host round trip).
int batch_launch(ol_ctx *c, ol_batch *b, bool cooperative)
{
    if (b->nd == 0)
        return OL_OK;
    int rc = OL_OK;
    if (!b->prep_search_only) {
        rc = batch_scan(c, b, cooperative);
        // under capture: the scan's node is the stream's one dependency now
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        const hipGraphNode_t *deps = nullptr;
        size_t nd = 0;
        if (!rc && hipStreamGetCaptureInfo_v2(c->stream, &cs, nullptr, nullptr, &deps, &nd) == hipSuccess &&
            cs == hipStreamCaptureStatusActive && nd == 1)
            b->scan_node = deps[0];
        if (!rc)
            rc = batch_tree(c, b);
    }
    if (!rc && b->prep_dict)
        rc = batch_many(c, b, b->prep_dict, b->m_cap ? b->m_cap : 1);
    return rc;
}

}  // namespace

extern "C" {

int ol_batch_create(ol_ctx *c, const uint8_t *const *docs, const size_t *lens, size_t count, ol_batch **out)
{
    if (!c || !out || (count && (!docs || !lens)))
        return fail(OL_ERR_INVALID, "null argument");
    *out = nullptr;
    std::lock_guard<std::mutex> g(c->mu);
    BatchLayout L;
    int rc = L.build(docs, lens, count);
    if (rc)
        return rc;
    ol_batch *b = new (std::nothrow) ol_batch();
    if (!b)
        return fail(OL_ERR_NOMEM, "out of host memory");
    b->count = count;
    b->h_start = L.start;
    b->h_len.assign(lens, lens + count);
    b->h_dev = L.dev;
    b->nd = L.nd;
    b->n = L.n;
    b->payload = L.payload;
    auto bail = [&](int rc) {
        if (b->text) (void)hipFree(b->text);
        b->scr.release();
        delete b;
        return rc;
    };
    if (scratch_init(b->scr) != hipSuccess)
        return bail(fail(OL_ERR_NOMEM, "batch scratch allocation failed"));
    const uint64_t padded = round_up(b->n ? b->n : 1, TEXT_ALIGN) + TEXT_SLACK;
    if (hipMalloc(&b->text, padded) != hipSuccess) {
        b->text = nullptr;
        return bail(fail(OL_ERR_NOMEM, "device allocation of the batch text failed"));
    }
    // the documents back to back through the staging ring
    rc = stage_upload(c, b->text, b->n, [&L](uint8_t *buf, uint64_t off, uint64_t k) { L.fill(buf, off, k); });
    if (rc == OL_OK && hipStreamSynchronize(c->stream) != hipSuccess)
        rc = fail(OL_ERR_DEVICE, "batch upload failed");
    if (rc)
        return bail(rc);
    if (hipMemset(b->text + b->n, 0, padded - b->n) != hipSuccess)
        return bail(fail(OL_ERR_DEVICE, "batch upload failed"));
    // document bounds of the non-empty documents; search tiles -> document
    std::vector<uint32_t> st(b->nd), en(b->nd);
    L.bounds(st.data(), en.data());
    std::vector<uint32_t> std_((size_t)ceil_div(b->n ? b->n : 1, SEARCH_TILE));
    L.search_tile_doc(std_.data());
    if (b->d_starts.ensure((size_t)(b->nd ? b->nd : 1) * 4) != hipSuccess ||
        b->d_ends.ensure((size_t)(b->nd ? b->nd : 1) * 4) != hipSuccess ||
        b->d_line_start.ensure((size_t)(b->nd ? b->nd : 1) * 4) != hipSuccess ||
        b->search_tile_doc.ensure(std_.size() * 4) != hipSuccess)
        return bail(fail(OL_ERR_NOMEM, "batch index allocation failed"));
    if ((b->nd && (hipMemcpy(b->d_starts.p, st.data(), st.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(b->d_ends.p, en.data(), en.size() * 4, hipMemcpyHostToDevice) != hipSuccess)) ||
        hipMemcpy(b->search_tile_doc.p, std_.data(), std_.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(OL_ERR_DEVICE, "batch index upload failed"));
    rc = alloc_batch_lines(b, std::min<uint64_t>(b->n, b->n / 16 + b->nd + 4096));
    if (rc)
        return bail(rc);
    *out = b;
    return OL_OK;
}

void ol_batch_free(ol_batch *b)
{
    if (!b)
        return;
    b->drop_graph();
    if (b->text)
        (void)hipFree(b->text);
    for (DevBuf *d : {&b->d_starts, &b->d_ends, &b->d_line_start, &b->scan_tile_doc, &b->search_tile_doc,
                      &b->offsets, &b->classes, &b->keys, &b->parent, &b->fc, &b->ns, &b->se, &b->seg_line0,
                      &b->seg_lines, &b->seg_tile0, &b->tile_seg, &b->m_off, &b->m_id})
        d->release();
    b->scr.release();
    delete b;
}

size_t ol_batch_doc_count(const ol_batch *b) { return b ? b->count : 0; }
size_t ol_batch_text_bytes(const ol_batch *b) { return b ? b->payload : 0; }

int ol_batch_run(ol_ctx *c, ol_batch *b, const ol_dict *dict, int flags)
{
    if (!c || !b)
        return fail(OL_ERR_INVALID, "null argument");
    const bool graph = (flags & OL_BATCH_GRAPH) != 0;
    const bool search_only = (flags & OL_BATCH_SEARCH_ONLY) != 0;
    if (search_only && !dict)
        return fail(OL_ERR_INVALID, "a search-only run needs a dictionary");
    std::lock_guard<std::mutex> g(c->mu);
    if (!b->prepared || b->prep_bpt != c->scan_bpt || b->prep_rounds != c->tree_rounds || b->prep_dict != dict ||
        b->prep_search_only != search_only)
        return batch_prepare(c, b, dict, search_only);
    b->h_matches = false;
    b->stream = c->stream;
    if (!graph)
        return batch_launch(c, b, true);
    soil(b->scr, b->graph_words);   // a replay leaves its look-back words and chunk totals set
    if (!b->exec) {
        const bool prof = c->profiling;
        c->profiling = false;
        // the same cooperative scan as the eager sequence: the capture
        // records it as a plain kernel node, so the node is marked
        // cooperative before instantiation.  The replay's first node clears
        // the status words that the previous run's dictionary pass left set.
        b->scan_node = nullptr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = batch_launch(c, b, true);
        hipGraph_t gr = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &gr);
        c->profiling = prof;
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess)
            return fail(OL_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        if (!b->prep_search_only) {
            hipKernelNodeAttrValue v{};
            v.cooperative = 1;
            if (!b->scan_node ||
                hipGraphKernelNodeSetAttribute(b->scan_node, hipKernelNodeAttributeCooperative, &v) != hipSuccess) {
                (void)hipGraphDestroy(gr);
                return fail(OL_ERR_DEVICE, "graph capture: the scan's node could not be marked cooperative");
            }
        }
        b->graph = gr;
        HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(b->exec, c->stream));
    return OL_OK;
}

int ol_batch_lines(ol_ctx *c, const ol_batch *b, uint64_t *total, uint32_t *doc_line_start, uint32_t *doc_line_count)
{
    if (!c || !b)
        return fail(OL_ERR_INVALID, "null argument");
    if (!b->prepared)
        return fail(OL_ERR_INVALID, "the batch has not been run");
    if (b->prep_search_only)
        return fail(OL_ERR_INVALID, "the batch was run for its dictionary search only");
    if (total)
        *total = b->lines;
    // empty documents: no lines, starting where the next document's lines start
    uint64_t next = b->lines;
    for (size_t i = b->count; i-- > 0;) {
        const uint32_t j = b->h_dev[i];
        const uint64_t st = (j == ~0u) ? next : b->h_line_start[j];
        if (doc_line_start) doc_line_start[i] = (uint32_t)st;
        if (doc_line_count) doc_line_count[i] = (j == ~0u) ? 0u : b->h_line_count[j];
        next = st;
    }
    return OL_OK;
}

int ol_batch_download_doc(ol_ctx *c, const ol_batch *b, size_t doc, uint32_t *offsets, uint8_t *classes,
                          uint32_t *parent, uint32_t *subtree_end, uint32_t *first_child, uint32_t *next_sibling)
{
    if (!c || !b)
        return fail(OL_ERR_INVALID, "null argument");
    if (!b->prepared)
        return fail(OL_ERR_INVALID, "the batch has not been run");
    if (doc >= b->count)
        return fail(OL_ERR_INVALID, "document index out of range");
    if (b->prep_search_only)
        return fail(OL_ERR_INVALID, "the batch was run for its dictionary search only");
    std::lock_guard<std::mutex> g(c->mu);
    int rc = check_tree_error(c, b->scr);   // also the sticky device error bits
    if (rc)
        return rc;
    const uint32_t j = b->h_dev[doc];
    if (j == ~0u)
        return OL_OK;
    const size_t l0 = b->h_line_start[j], nl = b->h_line_count[j];
    auto get = [&](void *dst, const DevBuf &src, size_t width) -> hipError_t {
        if (!dst || !nl)
            return hipSuccess;
        return hipMemcpy(dst, (const uint8_t *)src.p + l0 * width, nl * width, hipMemcpyDeviceToHost);
    };
    HIPCHK(get(offsets, b->offsets, 4));
    HIPCHK(get(classes, b->classes, 1));
    HIPCHK(get(parent, b->parent, 4));
    HIPCHK(get(subtree_end, b->se, 4));
    HIPCHK(get(first_child, b->fc, 4));
    HIPCHK(get(next_sibling, b->ns, 4));
    return OL_OK;
}

int ol_batch_matches(ol_ctx *c, ol_batch *b, size_t doc, uint64_t *count, uint32_t *offsets, uint32_t *ids,
                     size_t cap)
{
    if (!c || !b || !count)
        return fail(OL_ERR_INVALID, "null argument");
    if (!b->prepared)
        return fail(OL_ERR_INVALID, "the batch has not been run");
    if (doc >= b->count)
        return fail(OL_ERR_INVALID, "document index out of range");
    std::lock_guard<std::mutex> g(c->mu);
    if (!b->h_matches) {
        const int rc = sync_check(c, b->scr);
        if (rc)
            return rc;
        b->h_moff.resize(b->m_total);
        b->h_mid.resize(b->m_total);
        if (b->m_total) {
            HIPCHK(hipMemcpy(b->h_moff.data(), b->m_off.p, b->m_total * 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(b->h_mid.data(), b->m_id.p, b->m_total * 4, hipMemcpyDeviceToHost));
        }
        b->h_matches = true;
    }
    const uint64_t s = b->h_start[doc], e = s + b->h_len[doc];
    const size_t lo = std::lower_bound(b->h_moff.begin(), b->h_moff.end(), s) - b->h_moff.begin();
    const size_t hi = std::lower_bound(b->h_moff.begin(), b->h_moff.end(), e) - b->h_moff.begin();
    *count = hi - lo;
    const size_t k = std::min(cap, hi - lo);
    for (size_t i = 0; i < k; i++) {
        if (offsets) offsets[i] = (uint32_t)(b->h_moff[lo + i] - s);
        if (ids) ids[i] = b->h_mid[lo + i];
    }
    return OL_OK;
}

}  // extern "C"

// ---- serving slots ---------------------------------------------------------------
// A slot is a context (stream, staging ring, launch tuning) plus one device
// arena, allocated when the slot is created.  Every batch the slot runs is
// carved from the arena by bounds (one line per text byte for the per-line
// arrays; a dictionary's match list after its count), so running a batch
// allocates no device memory.  Index arrays go up and results come down
// through a pinned host buffer with asynchronous copies on the slot's
// stream; each batch has one read-back in the middle (the line counts, or
// the match total), and the slot's thread waits only for its own stream.
#include "slot.h"

constexpr uint32_t SLOT_TREE_ROUNDS = 4;   // tree tiles of 1,024 lines: small documents waste less

struct ol_slot {
    ol_ctx *ctx = nullptr;
    uint8_t *arena = nullptr;
    size_t arena_bytes = 0;
    uint8_t *pin = nullptr;   // pinned host staging (grown on demand)
    size_t pin_bytes = 0;
};

namespace {

// Bump allocation from an arena, 256-byte aligned.
struct Carver {
    uint8_t *base;
    size_t cap, off = 0;
    Carver(uint8_t *b, size_t c) : base(b), cap(c) {}
    uint8_t *take(size_t b)
    {
        const size_t o = round_up(off, 256);
        if (o + b > cap)
            return nullptr;
        off = o + b;
        return base + o;
    }
    static size_t need(size_t &acc, size_t b)
    {
        acc = round_up(acc, 256) + b;
        return acc;
    }
};

// Region sizes of a parse batch over n text bytes and nd documents.
struct ParsePlan {
    uint64_t padded, cap, tiles_bound;
    uint32_t scan_tiles;
    size_t status_words, total;
};
ParsePlan parse_plan(uint32_t bpt, uint32_t rounds, uint64_t n, uint64_t nd)
{
    ParsePlan P;
    P.padded = round_up(n ? n : 1, TEXT_ALIGN) + TEXT_SLACK;
    P.cap = round_up(n ? n : 1, LINE_ALIGN);   // a line has at least one byte
    P.scan_tiles = (uint32_t)ceil_div(n ? n : 1, (uint64_t)SCAN_THREADS * bpt);
    P.tiles_bound = ceil_div(P.cap, (uint64_t)ol_tree_tile_lines(rounds)) + nd + 1;
    P.status_words = STATUS_TILES + (size_t)P.scan_tiles * LB_STRIDE;   // ticket-ordered scan
    size_t t = 0;
    Carver::need(t, P.padded);
    Carver::need(t, nd * 4);                    // starts
    Carver::need(t, ((size_t)P.scan_tiles + 1) * 4);
    Carver::need(t, nd * 4);                    // line starts
    Carver::need(t, P.cap * 4);                 // offsets
    Carver::need(t, P.cap);                     // classes
    Carver::need(t, P.cap);                     // keys
    for (int k = 0; k < 4; k++)
        Carver::need(t, P.cap * 4);             // links
    for (int k = 0; k < 3; k++)
        Carver::need(t, nd * 4);                // segments
    Carver::need(t, P.tiles_bound * 4);         // tile -> segment
    Carver::need(t, ol_tree_sum_words(P.tiles_bound) * 4);   // tree level states
    Carver::need(t, HEAD_WORDS * 8);
    Carver::need(t, P.status_words * 8);
    P.total = t;
    return P;
}

struct SearchPlan {
    uint64_t padded;
    uint32_t search_tiles;
    size_t status_words, cand_bytes, total;   // without the match list
};
SearchPlan search_plan(uint64_t n, uint64_t nd)
{
    SearchPlan P;
    P.padded = round_up(n ? n : 1, TEXT_ALIGN) + TEXT_SLACK;
    P.search_tiles = (uint32_t)ceil_div(n ? n : 1, SEARCH_TILE);
    P.status_words = many_status_words(n);
    P.cand_bytes = (size_t)P.search_tiles * SEARCH_TILE * 4;
    size_t t = 0;
    Carver::need(t, P.padded);
    Carver::need(t, nd * 4);   // starts
    Carver::need(t, nd * 4);   // ends
    Carver::need(t, (size_t)P.search_tiles * 4);
    Carver::need(t, HEAD_WORDS * 8);
    Carver::need(t, P.status_words * 8);
    Carver::need(t, P.cand_bytes);
    P.total = t;
    return P;
}

uint64_t bound_text(const size_t *lens, size_t count, uint64_t *nd)
{
    uint64_t n = 0, k = 0;
    for (size_t i = 0; i < count; i++)
        if (lens[i]) {
            n += lens[i] + 1;
            k++;
        }
    *nd = k;
    return n;
}

int ensure_pin(ol_slot *s, size_t bytes)
{
    if (bytes <= s->pin_bytes)
        return OL_OK;
    if (s->pin)
        (void)hipHostFree(s->pin);
    s->pin = nullptr;
    s->pin_bytes = 0;
    const size_t want = round_up(std::max<size_t>(bytes, (size_t)16 << 20), (size_t)1 << 20);
    if (hipHostMalloc((void **)&s->pin, want, hipHostMallocDefault) != hipSuccess) {
        s->pin = nullptr;
        return fail(OL_ERR_NOMEM, "pinned host allocation of " + std::to_string(want) + " bytes failed");
    }
    s->pin_bytes = want;
    return OL_OK;
}

// Lends the batch's scratch its arena regions (zeroing the head).
int lend_scratch(ol_ctx *c, Carver &cv, Scratch &scr, size_t status_words)
{
    uint8_t *h = cv.take(HEAD_WORDS * 8), *st = cv.take(status_words * 8);
    if (!h || !st)
        return fail(OL_ERR_NOMEM, "the batch does not fit the slot's arena");
    scr.head.lend(h, HEAD_WORDS * 8);
    scr.status.lend(st, status_words * 8);
    scr.dirty = status_words;
    HIPCHK(hipMemsetAsync(h, 0, HEAD_WORDS * 8, c->stream));
    return OL_OK;
}

// Waits for the slot's stream; the head's sticky error bits and meta[0].
int slot_sync(ol_ctx *c, Scratch &scr, uint64_t *head, uint64_t *meta)
{
    HIPCHK(hipMemcpyAsync(head, scr.head.p, HEAD_WORDS * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint32_t bits = ((const uint32_t *)(head + HEAD_CTR))[CTR_DEV_ERR];
    if (bits)
        return fail(OL_ERR_DEVICE, dev_error_text(bits));
    if (meta)
        *meta = head[HEAD_META];
    return OL_OK;
}

// The dictionary search of text [0, n) (documents `nd` of batch `b` when nd > 0,
// else one document) into results: count with stored candidates, one read-back
// of the total, the match list carved from what is left, the writes.
int slot_many(ol_slot *s, Carver &cv, ol_batch &b, const ol_dict *dict, std::vector<uint64_t> &moff,
              std::vector<uint32_t> &mid)
{
    ol_ctx *c = s->ctx;
    ManyArgs a{};
    a.text = b.text;
    a.n = (uint32_t)b.n;
    a.d = dict->d;
    HIPCHK(setup_many(c, b.scr, a, b.n, true));
    if (b.nd) {
        a.ndocs = b.nd;
        a.doc_starts = (const uint32_t *)b.d_starts.p;
        a.doc_ends = (const uint32_t *)b.d_ends.p;
        a.tile_doc = (const uint32_t *)b.search_tile_doc.p;
    }
    {
        Prof p(c, "many");
        HIPCHK(ol_launch_many(a, c->stream));   // count (storing candidates) + scan
    }
    uint64_t *head = (uint64_t *)s->pin;
    uint64_t total = 0;
    int rc = slot_sync(c, b.scr, head, &total);
    if (rc)
        return rc;
    moff.resize(total);
    mid.resize(total);
    if (!total)
        return OL_OK;
    uint8_t *mo = cv.take(total * 8), *mi = cv.take(total * 4);
    if (!mo || !mi)
        return fail(OL_ERR_NOMEM, std::to_string(total) + " matches do not fit what is left of the slot's arena");
    a.out_off = (uint64_t *)mo;
    a.out_id = (uint32_t *)mi;
    a.cap = total;
    {
        Prof p(c, "many_write");
        HIPCHK(ol_launch_many_write(a, c->stream));
    }
    rc = ensure_pin(s, total * 12 + HEAD_WORDS * 8);
    if (rc)
        return rc;
    uint8_t *pin = s->pin + HEAD_WORDS * 8;
    HIPCHK(hipMemcpyAsync(pin, mo, total * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(pin + total * 8, mi, total * 4, hipMemcpyDeviceToHost, c->stream));
    rc = slot_sync(c, b.scr, (uint64_t *)s->pin, nullptr);
    if (rc)
        return rc;
    std::memcpy(moff.data(), pin, total * 8);
    std::memcpy(mid.data(), pin + total * 8, total * 4);
    return OL_OK;
}

}  // namespace

int ol_slot_create(size_t arena_bytes, ol_slot **out)
{
    if (!out)
        return fail(OL_ERR_INVALID, "null argument");
    *out = nullptr;
    ol_slot *s = new (std::nothrow) ol_slot();
    if (!s)
        return fail(OL_ERR_NOMEM, "out of host memory");
    int rc = ol_ctx_create(&s->ctx);
    if (rc) {
        delete s;
        return rc;
    }
    // ticket-ordered scan tiles: several slots share the device
    (void)ol_tune_set(s->ctx, "scan_ticket", 1);
    (void)ol_tune_set(s->ctx, "tree_rounds", (int)SLOT_TREE_ROUNDS);
    if (hipMalloc((void **)&s->arena, arena_bytes) != hipSuccess) {
        ol_ctx_destroy(s->ctx);
        delete s;
        return fail(OL_ERR_NOMEM, "device allocation of a " + std::to_string(arena_bytes) + "-byte slot arena failed");
    }
    s->arena_bytes = arena_bytes;
    rc = ensure_pin(s, 16u << 20);
    if (rc) {
        ol_slot_destroy(s);
        return rc;
    }
    *out = s;
    return OL_OK;
}

void ol_slot_destroy(ol_slot *s)
{
    if (!s)
        return;
    if (s->ctx)
        (void)hipStreamSynchronize(s->ctx->stream);
    if (s->arena)
        (void)hipFree(s->arena);
    if (s->pin)
        (void)hipHostFree(s->pin);
    ol_ctx_destroy(s->ctx);
    delete s;
}

ol_ctx *ol_slot_ctx(ol_slot *s) { return s ? s->ctx : nullptr; }
size_t ol_slot_arena_bytes(const ol_slot *s) { return s ? s->arena_bytes : 0; }

size_t ol_slot_parse_bytes(const size_t *lens, size_t count)
{
    // the plan depends on the slot's tuning, which every slot shares
    uint64_t nd = 0;
    const uint64_t n = bound_text(lens, count, &nd);
    return parse_plan(SCAN_BPT, SLOT_TREE_ROUNDS, n, nd).total;
}

size_t ol_slot_search_bytes(const size_t *lens, size_t count)
{
    uint64_t nd = 0;
    const uint64_t n = bound_text(lens, count, &nd);
    return search_plan(n, nd).total;
}

size_t ol_doc_device_bytes(const ol_doc *d)
{
    if (!d || !d->text)
        return 0;
    return (size_t)(round_up(d->n, TEXT_ALIGN) + TEXT_SLACK + d->cap * (4 + 1 + 1 + (d->nodes ? 16 : 0)));
}

size_t ol_tree_device_bytes(const ol_tree *t) { return t ? (size_t)t->cap * 16 : 0; }
size_t ol_dict_device_bytes(const ol_dict *d) { return d ? d->bytes : 0; }

int ol_slot_parse(ol_slot *s, const uint8_t *const *docs, const size_t *lens, size_t count, SlotResult *res)
{
    if (!s || (count && (!docs || !lens || !res)))
        return fail(OL_ERR_INVALID, "null argument");
    ol_ctx *c = s->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    BatchLayout L;
    int rc = L.build(docs, lens, count);
    if (rc)
        return rc;
    for (size_t i = 0; i < count; i++)
        res[i] = SlotResult();
    if (L.nd == 0)
        return OL_OK;
    const ParsePlan P = parse_plan(c->scan_bpt, c->tree_rounds, L.n, L.nd);
    if (P.total > s->arena_bytes)
        return fail(OL_ERR_NOMEM, "a parse batch of " + std::to_string(L.n) + " bytes needs " +
                                      std::to_string(P.total) + " arena bytes; the slot has " +
                                      std::to_string(s->arena_bytes));
    Carver cv(s->arena, s->arena_bytes);
    ol_batch b;
    b.count = count;
    b.nd = L.nd;
    b.n = L.n;
    b.h_start = L.start;
    b.h_dev = L.dev;
    b.text = cv.take(P.padded);
    auto lend = [&](DevBuf &d, size_t bytes) {
        uint8_t *q = cv.take(bytes);
        if (q) d.lend(q, bytes);
        return q != nullptr;
    };
    bool ok = b.text && lend(b.d_starts, (size_t)L.nd * 4) && lend(b.scan_tile_doc, ((size_t)P.scan_tiles + 1) * 4) &&
              lend(b.d_line_start, (size_t)L.nd * 4) && lend(b.offsets, P.cap * 4) && lend(b.classes, P.cap) &&
              lend(b.keys, P.cap) && lend(b.parent, P.cap * 4) && lend(b.fc, P.cap * 4) && lend(b.ns, P.cap * 4) &&
              lend(b.se, P.cap * 4) && lend(b.seg_line0, (size_t)L.nd * 4) && lend(b.seg_lines, (size_t)L.nd * 4) &&
              lend(b.seg_tile0, (size_t)L.nd * 4) && lend(b.tile_seg, P.tiles_bound * 4) &&
              lend(b.scr.tree_sum, ol_tree_sum_words(P.tiles_bound) * 4);
    if (!ok)
        return fail(OL_ERR_NOMEM, "the batch does not fit the slot's arena");
    if ((rc = lend_scratch(c, cv, b.scr, P.status_words)))
        return rc;
    b.cap = P.cap;
    b.scan_tile_doc_bpt = c->scan_bpt;
    // index arrays up (pinned), text up (staging ring)
    const size_t idx_bytes = ((size_t)L.nd + P.scan_tiles + 1) * 4;
    if ((rc = ensure_pin(s, HEAD_WORDS * 8 + idx_bytes + (size_t)L.nd * 4)))
        return rc;
    uint32_t *up = (uint32_t *)(s->pin + HEAD_WORDS * 8);
    L.bounds(up, nullptr);
    L.scan_tile_doc((uint64_t)SCAN_THREADS * c->scan_bpt, up + L.nd);
    HIPCHK(hipMemcpyAsync(b.d_starts.p, up, (size_t)L.nd * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.scan_tile_doc.p, up + L.nd, ((size_t)P.scan_tiles + 1) * 4, hipMemcpyHostToDevice,
                          c->stream));
    rc = stage_upload(c, b.text, L.n, [&L](uint8_t *buf, uint64_t off, uint64_t k) { L.fill(buf, off, k); });
    if (rc)
        return rc;
    HIPCHK(hipMemsetAsync(b.text + L.n, 0, P.padded - L.n, c->stream));
    mark(c, "serve: upload");
    if ((rc = batch_scan(c, &b, false)))
        return rc;
    // read-back: line count and each document's first line
    uint32_t *ls = (uint32_t *)(s->pin + HEAD_WORDS * 8 + idx_bytes);
    HIPCHK(hipMemcpyAsync(ls, b.d_line_start.p, (size_t)L.nd * 4, hipMemcpyDeviceToHost, c->stream));
    uint64_t lines = 0;
    if ((rc = slot_sync(c, b.scr, (uint64_t *)s->pin, &lines)))
        return rc;
    if (lines > P.cap)
        return fail(OL_ERR_DEVICE, "batch line count exceeds its bound");
    b.lines = lines;
    std::vector<uint32_t> line0(ls, ls + L.nd), cnt(L.nd), tile0(L.nd), tseg;
    const uint64_t TL = ol_tree_tile_lines(c->tree_rounds);
    uint64_t tiles = 0;
    for (uint32_t j = 0; j < L.nd; j++) {
        const uint64_t end = (j + 1 < L.nd) ? line0[j + 1] : lines;
        cnt[j] = (uint32_t)(end - line0[j]);
        tile0[j] = (uint32_t)tiles;
        const uint64_t k = ceil_div(cnt[j] ? cnt[j] : 1, TL);
        tseg.insert(tseg.end(), k, j);
        tiles += k;
    }
    if (tiles > P.tiles_bound)
        return fail(OL_ERR_DEVICE, "batch tree tiles exceed their bound");
    b.tree_tiles = (uint32_t)tiles;
    // the pinned buffer for the rest of the batch (every copy so far has
    // completed, so it may move): the segments up, then 21 bytes per line down
    const size_t base = HEAD_WORDS * 8 + round_up(((size_t)3 * L.nd + tiles) * 4, 256);
    if ((rc = ensure_pin(s, base + lines * 21)))
        return rc;
    uint32_t *sg = (uint32_t *)(s->pin + HEAD_WORDS * 8);
    std::memcpy(sg, line0.data(), (size_t)L.nd * 4);
    std::memcpy(sg + L.nd, cnt.data(), (size_t)L.nd * 4);
    std::memcpy(sg + 2 * (size_t)L.nd, tile0.data(), (size_t)L.nd * 4);
    std::memcpy(sg + 3 * (size_t)L.nd, tseg.data(), tiles * 4);
    HIPCHK(hipMemcpyAsync(b.seg_line0.p, sg, (size_t)L.nd * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.seg_lines.p, sg + L.nd, (size_t)L.nd * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.seg_tile0.p, sg + 2 * (size_t)L.nd, (size_t)L.nd * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.tile_seg.p, sg + 3 * (size_t)L.nd, tiles * 4, hipMemcpyHostToDevice, c->stream));
    if ((rc = batch_tree(c, &b)))
        return rc;
    mark(c, "serve: scan + tree");
    // results down, after the segment area
    uint8_t *dn = s->pin + base;
    const struct {
        const DevBuf *src;
        size_t w;
    } arr[6] = {{&b.offsets, 4}, {&b.parent, 4}, {&b.se, 4}, {&b.fc, 4}, {&b.ns, 4}, {&b.classes, 1}};
    size_t at = 0;
    for (const auto &x : arr) {
        if (lines)
            HIPCHK(hipMemcpyAsync(dn + at, x.src->p, lines * x.w, hipMemcpyDeviceToHost, c->stream));
        at += lines * x.w;
    }
    uint64_t *head = (uint64_t *)s->pin;
    if ((rc = slot_sync(c, b.scr, head, nullptr)))
        return rc;
    if (((const uint32_t *)(head + HEAD_CTR))[CTR_TREE_ERR])
        return fail(OL_ERR_TOO_DEEP, "a content line is nested deeper than the " + std::to_string(OL_MAX_TREE_DEPTH) +
                                         " levels the GPU tree tracks");
    mark(c, "serve: download");
    const uint32_t *h_off = (const uint32_t *)dn, *h_par = h_off + lines, *h_se = h_par + lines,
                   *h_fc = h_se + lines, *h_ns = h_fc + lines;
    const uint8_t *h_cls = (const uint8_t *)(h_ns + lines);
    for (size_t i = 0; i < count; i++) {
        const uint32_t j = L.dev[i];
        if (j == ~0u)
            continue;
        const size_t l0 = line0[j], k = cnt[j];
        SlotResult &r = res[i];
        r.offsets.assign(h_off + l0, h_off + l0 + k);
        r.parent.assign(h_par + l0, h_par + l0 + k);
        r.se.assign(h_se + l0, h_se + l0 + k);
        r.fc.assign(h_fc + l0, h_fc + l0 + k);
        r.ns.assign(h_ns + l0, h_ns + l0 + k);
        r.classes.assign(h_cls + l0, h_cls + l0 + k);
    }
    return OL_OK;
}

int ol_slot_search(ol_slot *s, const uint8_t *const *docs, const size_t *lens, size_t count, const ol_dict *dict,
                   SlotResult *res)
{
    if (!s || !dict || (count && (!docs || !lens || !res)))
        return fail(OL_ERR_INVALID, "null argument");
    ol_ctx *c = s->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    BatchLayout L;
    int rc = L.build(docs, lens, count);
    if (rc)
        return rc;
    for (size_t i = 0; i < count; i++)
        res[i] = SlotResult();
    if (L.nd == 0)
        return OL_OK;
    const SearchPlan P = search_plan(L.n, L.nd);
    if (P.total > s->arena_bytes)
        return fail(OL_ERR_NOMEM, "a search batch of " + std::to_string(L.n) + " bytes needs " +
                                      std::to_string(P.total) + " arena bytes; the slot has " +
                                      std::to_string(s->arena_bytes));
    Carver cv(s->arena, s->arena_bytes);
    ol_batch b;
    b.count = count;
    b.nd = L.nd;
    b.n = L.n;
    b.text = cv.take(P.padded);
    auto lend = [&](DevBuf &d, size_t bytes) {
        uint8_t *q = cv.take(bytes);
        if (q) d.lend(q, bytes);
        return q != nullptr;
    };
    if (!b.text || !lend(b.d_starts, (size_t)L.nd * 4) || !lend(b.d_ends, (size_t)L.nd * 4) ||
        !lend(b.search_tile_doc, (size_t)P.search_tiles * 4))
        return fail(OL_ERR_NOMEM, "the batch does not fit the slot's arena");
    if ((rc = lend_scratch(c, cv, b.scr, P.status_words)))
        return rc;
    if (!lend(b.scr.cands, P.cand_bytes))
        return fail(OL_ERR_NOMEM, "the batch does not fit the slot's arena");
    const size_t idx_bytes = ((size_t)2 * L.nd + P.search_tiles) * 4;
    if ((rc = ensure_pin(s, HEAD_WORDS * 8 + idx_bytes)))
        return rc;
    uint32_t *up = (uint32_t *)(s->pin + HEAD_WORDS * 8);
    L.bounds(up, up + L.nd);
    L.search_tile_doc(up + 2 * (size_t)L.nd);
    HIPCHK(hipMemcpyAsync(b.d_starts.p, up, (size_t)L.nd * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.d_ends.p, up + L.nd, (size_t)L.nd * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(b.search_tile_doc.p, up + 2 * (size_t)L.nd, (size_t)P.search_tiles * 4,
                          hipMemcpyHostToDevice, c->stream));
    rc = stage_upload(c, b.text, L.n, [&L](uint8_t *buf, uint64_t off, uint64_t k) { L.fill(buf, off, k); });
    if (rc)
        return rc;
    HIPCHK(hipMemsetAsync(b.text + L.n, 0, P.padded - L.n, c->stream));
    mark(c, "serve: upload");
    std::vector<uint64_t> moff;
    std::vector<uint32_t> mid;
    if ((rc = slot_many(s, cv, b, dict, moff, mid)))
        return rc;
    mark(c, "serve: search + download");
    for (size_t i = 0; i < count; i++) {
        if (L.dev[i] == ~0u)
            continue;
        const uint64_t st = L.start[i], en = st + lens[i];
        const size_t lo = std::lower_bound(moff.begin(), moff.end(), st) - moff.begin();
        const size_t hi = std::lower_bound(moff.begin(), moff.end(), en) - moff.begin();
        SlotResult &r = res[i];
        r.m_off.resize(hi - lo);
        for (size_t k = lo; k < hi; k++)
            r.m_off[k - lo] = (uint32_t)(moff[k] - st);
        r.m_id.assign(mid.begin() + lo, mid.begin() + hi);
    }
    return OL_OK;
}

int ol_slot_search_doc(ol_slot *s, const ol_doc *doc, const ol_dict *dict, SlotResult *res)
{
    if (!s || !doc || !dict || !res)
        return fail(OL_ERR_INVALID, "null argument");
    ol_ctx *c = s->ctx;
    std::lock_guard<std::mutex> g(c->mu);
    *res = SlotResult();
    if (doc->n == 0)
        return OL_OK;
    const SearchPlan P = search_plan(doc->n, 0);
    Carver cv(s->arena, s->arena_bytes);
    ol_batch b;
    b.n = doc->n;
    b.text = doc->text;
    int rc = lend_scratch(c, cv, b.scr, P.status_words);
    if (rc)
        return rc;
    uint8_t *cq = cv.take(P.cand_bytes);
    if (!cq)
        return fail(OL_ERR_NOMEM, "the search does not fit the slot's arena");
    b.scr.cands.lend(cq, P.cand_bytes);
    std::vector<uint64_t> moff;
    if ((rc = slot_many(s, cv, b, dict, moff, res->m_id)))
        return rc;
    res->m_off.assign(moff.begin(), moff.end());
    return OL_OK;
}
