/*
 * oneline_hip.h -- C ABI of the MI355X (gfx950) parse-and-search engine.
 *
 * Drop-in boundary for the `oneline-gpu` crate's public API
 * (crates/oneline-gpu/src/lib.rs:27-36).  Every entry point below names the
 * reference item it replaces.  Plain pointers and sizes only; no C++ or torch
 * types cross this boundary.
 *
 * Conventions (mirroring the reference's behaviour):
 *   - every call returns an int status: OL_OK (0) or a negative OL_ERR_*;
 *     ol_last_error() returns the calling thread's last message;
 *   - bad input is an error, never a crash, and an error does not poison the
 *     context (crates/oneline-gpu/src/query.rs:112-118);
 *   - empty input returns empty results without touching the GPU
 *     (crates/oneline-gpu/src/parse.rs:88-90, document.rs:60-67);
 *   - truncated outputs keep the true totals (crates/oneline-gpu/src/tests.rs:375-397);
 *   - arrays returned through `T **out` are allocated by the library and
 *     released with ol_free(); arrays passed as `T *out` are caller-owned;
 *   - a context serialises its calls with an internal lock (the reference's
 *     Mutex-guarded scratch, crates/oneline-gpu/src/parser.rs:15-19), so it may
 *     be shared between threads.
 */
#ifndef ONELINE_HIP_H
#define ONELINE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OL_OK 0
#define OL_ERR_INVALID (-1)      /* bad argument */
#define OL_ERR_TOO_LARGE (-2)    /* input exceeds OL_MAX_GPU_INPUT_LEN */
#define OL_ERR_PATTERN (-3)      /* empty or too long pattern */
#define OL_ERR_TOO_MANY (-4)     /* too many patterns, or a count that does not fit */
#define OL_ERR_DEVICE (-5)       /* HIP runtime / device failure */
#define OL_ERR_TOO_DEEP (-6)     /* content line nested deeper than OL_MAX_TREE_DEPTH - 1 */
#define OL_ERR_NOMEM (-7)        /* host or device allocation failed, or HBM budget exceeded */

/* Limits.  MAX_GPU_INPUT_LEN: crates/oneline-gpu/src/util.rs:8.
 * MAX_PATTERN_LEN: crates/oneline-gpu/src/lib.rs:67.  MAX_TREE_DEPTH:
 * crates/oneline-gpu/src/tree.rs:12.  The dictionary limits are wider than the
 * reference's (crates/oneline-gpu/src/many.rs:14-18: 4..=64 bytes, 2,048 patterns):
 * 1..=64 bytes and up to OL_MAX_PATTERNS patterns. */
#define OL_MAX_GPU_INPUT_LEN 4294967292ull
#define OL_MAX_PATTERN_LEN 64
#define OL_MIN_SET_PATTERN_LEN 1
#define OL_MAX_SET_PATTERN_LEN 64
#define OL_MAX_PATTERNS (1u << 22)
#define OL_MAX_TREE_DEPTH 32
#define OL_NONE 0xFFFFFFFFu

/* FlatNode, crates/oneline-core/src/lib.rs:87-96 (16 bytes, same layout). */
typedef struct ol_flat_node {
    uint32_t depth;
    uint32_t kind;       /* NodeKind: 0 blank .. 7 comment (lib.rs:41-51) */
    uint32_t span_start;
    uint32_t span_len;
} ol_flat_node;

/* ManyHit, crates/oneline-gpu/src/many.rs:75-82. */
typedef struct ol_many_hit {
    uint32_t pattern;
    uint32_t offset;
} ol_many_hit;

typedef struct ol_ctx ol_ctx;    /* GpuParser      crates/oneline-gpu/src/parser.rs:8-57 */
typedef struct ol_doc ol_doc;    /* GpuDocument    crates/oneline-gpu/src/document.rs:16-31 */
typedef struct ol_tree ol_tree;  /* GpuTree        crates/oneline-gpu/src/tree.rs:29-46 */
typedef struct ol_dict ol_dict;  /* GpuPatternSet  crates/oneline-gpu/src/many.rs:116-150 */

/* ---- context ------------------------------------------------------------------ */
/* GpuParser::new (crates/oneline-gpu/src/parser.rs:94).  Uses HIP device 0. */
int ol_ctx_create(ol_ctx **out);
void ol_ctx_destroy(ol_ctx *ctx);
/* GpuParser::adapter_name (parser.rs:470). */
int ol_ctx_adapter_name(ol_ctx *ctx, char *buf, size_t cap);
/* GpuParser::is_software (parser.rs:476): always 0 on a HIP device. */
int ol_ctx_is_software(ol_ctx *ctx);
/* GpuParser::timestamps_supported (crates/oneline-gpu/src/document.rs:130). */
int ol_ctx_timestamps_supported(ol_ctx *ctx);
/* Waits for all work of the context; OL_ERR_DEVICE if a kernel of it set a
 * device error flag since the last check (the flags are sticky until read). */
int ol_ctx_synchronize(ol_ctx *ctx);
/* Thread-local message of the last failed call ("" if none). */
const char *ol_last_error(void);
/* Releases an array returned through a `T **out` parameter. */
void ol_free(void *p);

/* ---- one-shot parse ------------------------------------------------------------ */
/* GpuParser::scan_line_offsets (crates/oneline-gpu/src/scan.rs:127). */
int ol_scan_line_offsets(ol_ctx *ctx, const uint8_t *text, size_t len, uint32_t **offsets,
                         size_t *line_count);
/* GpuParser::parse_raw (crates/oneline-gpu/src/parse.rs:86): offsets + nodes. */
int ol_parse_raw(ol_ctx *ctx, const uint8_t *text, size_t len, ol_flat_node **nodes,
                 uint32_t **offsets, size_t *line_count);
/* GpuParser::parse (crates/oneline-gpu/src/parse.rs:104): nodes for caller-given
 * line offsets (line i spans offsets[i] .. offsets[i+1], the last to the end). */
int ol_parse(ol_ctx *ctx, const uint8_t *text, size_t len, const uint32_t *offsets,
             size_t line_count, ol_flat_node **nodes);

/* ---- resident documents ---------------------------------------------------------- */
/* GpuParser::upload (crates/oneline-gpu/src/document.rs:54): upload, scan and
 * classify once; text, offsets, kinds, keys and nodes stay in HBM. */
int ol_doc_upload(ol_ctx *ctx, const uint8_t *text, size_t len, ol_doc **out);
/* Same as ol_doc_upload without the resident 16-byte FlatNode array: offsets,
 * kinds and tree keys only (the per-line outputs the tree and searches need). */
int ol_doc_upload_lean(ol_ctx *ctx, const uint8_t *text, size_t len, ol_doc **out);
void ol_doc_free(ol_doc *doc);
/* GpuDocument::line_count / text_len (document.rs:35,47). */
size_t ol_doc_line_count(const ol_doc *doc);
size_t ol_doc_text_len(const ol_doc *doc);
/* Copies the resident per-line arrays to the host (any pointer may be NULL;
 * each array holds ol_doc_line_count entries).  Classes are NodeKind as u8. */
int ol_doc_download(ol_ctx *ctx, const ol_doc *doc, uint32_t *offsets, uint8_t *classes,
                    ol_flat_node *nodes);
/* Re-runs scan + classify over the resident text into the resident arrays
 * (same results as ol_doc_upload; asynchronous: no host round trip). */
int ol_doc_reparse(ol_ctx *ctx, ol_doc *doc);

/* ---- tree links ------------------------------------------------------------------- */
/* GpuParser::build_tree / build_tree_reusing (crates/oneline-gpu/src/tree.rs:53,60).
 * `reuse` (nullable) is consumed: its buffers are reused when large enough.
 * Fails with OL_ERR_TOO_DEEP if a content line is nested >= OL_MAX_TREE_DEPTH. */
int ol_tree_build(ol_ctx *ctx, const ol_doc *doc, ol_tree *reuse, ol_tree **out);
void ol_tree_free(ol_tree *tree);
/* GpuTree::line_count (tree.rs:43). */
size_t ol_tree_line_count(const ol_tree *tree);
/* GpuParser::download_links (tree.rs:278), plus the derived first_child and
 * next_sibling.  Any pointer may be NULL; each holds line_count entries. */
int ol_tree_download(ol_ctx *ctx, const ol_tree *tree, uint32_t *parent, uint32_t *subtree_end,
                     uint32_t *first_child, uint32_t *next_sibling);
/* Rebuilds the links of `doc` into `tree` (capacity at least the document's
 * lines) without a host round trip; the depth and device error flags are
 * checked by the next ol_tree_check.  Same arrays as ol_tree_build. */
int ol_tree_rebuild_async(ol_ctx *ctx, const ol_doc *doc, ol_tree *tree);
/* The same, writing only the arrays in `links` (OL_LINK_* bits): a caller that
 * needs parent / first_child / next_sibling skips the subtree_end writes.
 * ol_tree_download of an array the last build did not write fails. */
#define OL_LINK_PARENT 1u
#define OL_LINK_SUBTREE_END 2u
#define OL_LINK_FIRST_CHILD 4u
#define OL_LINK_NEXT_SIBLING 8u
#define OL_LINK_ALL 15u
int ol_tree_rebuild_links_async(ol_ctx *ctx, const ol_doc *doc, ol_tree *tree, uint32_t links);
int ol_tree_check(ol_ctx *ctx, const ol_tree *tree);

/* ---- AST bridge: structural pass ---------------------------------------------------- */
/* The line-structure half of to_tomet_ast_with_links
 * (crates/oneline-tomet/src/lib.rs:20-90): literal_pipe_lines (lib.rs:96-125),
 * has_pipe_child (lib.rs:33-41), which lines get a block (lib.rs:44-60) and
 * the stitching walk's parent for each block (lib.rs:66-86).  The blocks
 * themselves need tomet's own parser (lib.rs:128-310) and stay with the
 * caller.  Per line (host arrays of ol_doc_line_count entries; either may be
 * NULL):
 *   flags[i]     = OL_BRIDGE_* bits;
 *   container[i] = the line whose block receives line i's block in the
 *                  stitching walk (before finish() keeps or hoists it), or
 *                  OL_NONE when the block goes to the top level or line i has
 *                  no block.
 * `tree` must hold the links of `doc` with parent written. */
#define OL_BRIDGE_LITERAL 1u      /* a `|` line tomet reads as plain text */
#define OL_BRIDGE_BLOCK 2u        /* the line becomes a block (not blank, not a comment, text after trim) */
#define OL_BRIDGE_CUT 4u          /* a bare blank line lies between the line and its parent: parent NONE */
#define OL_BRIDGE_PIPE_CHILD 8u   /* an element whose `|` lines are its content */
#define OL_BRIDGE_BARE_BLANK 16u  /* an empty line with no `|` */
int ol_tree_bridge(ol_ctx *ctx, const ol_doc *doc, const ol_tree *tree, uint8_t *flags,
                   uint32_t *container);

/* ---- literal search --------------------------------------------------------------- */
/* GpuParser::search_document (crates/oneline-gpu/src/search.rs:84): 1..=64-byte
 * pattern; total = every overlapping occurrence; offsets = the first
 * min(total, max_matches) occurrences in ascending order (library-allocated). */
int ol_search_document(ol_ctx *ctx, const ol_doc *doc, const uint8_t *pattern, size_t len,
                       size_t max_matches, uint32_t *total, uint32_t **offsets, size_t *n_offsets);
/* GpuParser::search_raw (search.rs:107): uploads, then searches. */
int ol_search_raw(ol_ctx *ctx, const uint8_t *text, size_t text_len, const uint8_t *pattern,
                  size_t len, size_t max_matches, uint32_t *total, uint32_t **offsets,
                  size_t *n_offsets);
/* Device-resident form: u64 offsets written to device memory `d_offsets`
 * (capacity `cap`), *total = all matches.  Synchronous. */
int ol_search_device(ol_ctx *ctx, const ol_doc *doc, const uint8_t *pattern, size_t len,
                     uint64_t *d_offsets, size_t cap, uint64_t *total);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code that must not call hipMalloc: tight serving loops, harnesses
 * that forbid it): ol_search_device with the same kernel and results, whose
 * scratch (per-tile status words, the match total, the error word) is a
 * static device array of 4 MiB + 128 B per GPU instead of context memory, so
 * the call allocates nothing.  The array covers every document up to
 * OL_MAX_GPU_INPUT_LEN bytes (a larger tile count fails with
 * OL_ERR_TOO_LARGE).  All calls on one GPU share the array: calls that overlap
 * on one GPU (several contexts or threads at once) race; calls through one
 * context are serialised by its lock.  Synchronous, like ol_search_device. */
int ol_search_device_noalloc(ol_ctx *ctx, const ol_doc *doc, const uint8_t *pattern, size_t len,
                             uint64_t *d_offsets, size_t cap, uint64_t *total);

/* ---- .olb files ------------------------------------------------------------------------ */
/* A parsed document on disk, the ".olb" container the reference specifies
 * (.agents/tasks/phase4-zero-copy-binary-serialization.md:13-36; not built there):
 * a 64-byte header (magic "ONELINE\0", version 1, flags, text_len, line_count,
 * section offsets, CRC-32 of the sections), the text, the 16-byte FlatNodes and
 * the u32 line offsets, each on a 64-byte boundary.  Loading maps the file and
 * uploads text, offsets and nodes without scanning again.  No counterpart
 * entry point in the reference. */
int ol_doc_save(ol_ctx *ctx, const ol_doc *doc, const char *path);
/* verify != 0: check the CRC (OL_ERR_INVALID on mismatch, bad magic, version or sizes). */
int ol_doc_load(ol_ctx *ctx, const char *path, int verify, ol_doc **out);

/* ---- kind / depth queries ------------------------------------------------------------ */
/* The line filter of oneline-gpu (filter_nodes, crates/oneline-gpu/src/shader.wgsl:152-179):
 * a line matches when (kind == OL_QUERY_ANY_KIND or its NodeKind == kind) and
 * (depth == OL_QUERY_ANY_DEPTH or its FlatNode depth == depth).  As in the
 * reference, kind 0 means any kind (NodeKind::Blank is 0, so blank lines cannot
 * be selected by kind alone).  total = every matching line; indices = the first
 * min(total, max_matches) of them in ascending order (library-allocated, release
 * with ol_free).  The reference returns an arbitrary subset when truncating; this
 * one is the first lines. */
#define OL_QUERY_ANY_KIND 0u
#define OL_QUERY_ANY_DEPTH 0xFFFFFFFFu
/* GpuParser::query_document (crates/oneline-gpu/src/document.rs:94). */
int ol_query_document(ol_ctx *ctx, const ol_doc *doc, uint32_t kind, uint32_t depth, size_t max_matches,
                      uint32_t *total, uint32_t **indices, size_t *n_indices);
/* GpuParser::query_subtree (crates/oneline-gpu/src/tree.rs:295): lines
 * root .. subtree_end[root]; OL_ERR_INVALID when root is out of range or the
 * tree does not belong to the document. */
int ol_query_subtree(ol_ctx *ctx, const ol_doc *doc, const ol_tree *tree, size_t root, uint32_t kind,
                     uint32_t depth, size_t max_matches, uint32_t *total, uint32_t **indices,
                     size_t *n_indices);
/* GpuParser::query (crates/oneline-gpu/src/query.rs:288-314): the same query
 * over caller-given line offsets (computed elsewhere, e.g. by a CPU scanner):
 * line i spans offsets[i] .. offsets[i+1], the last one to the end of the
 * text, and is parsed from those bytes (no scan).  No offsets: no matches and
 * no GPU work.  Indices are line numbers of the given offsets, ascending. */
int ol_query(ol_ctx *ctx, const uint8_t *text, size_t len, const uint32_t *offsets, size_t line_count,
             uint32_t kind, uint32_t depth, size_t max_matches, uint32_t *total, uint32_t **indices,
             size_t *n_indices);
/* GpuParser::query_raw (crates/oneline-gpu/src/query.rs:232): uploads, scans, queries. */
int ol_query_raw(ol_ctx *ctx, const uint8_t *text, size_t len, uint32_t kind, uint32_t depth,
                 size_t max_matches, uint32_t *total, uint32_t **indices, size_t *n_indices);

/* ---- dictionary search ------------------------------------------------------------ */
/* GpuParser::compile_patterns (crates/oneline-gpu/src/many.rs:158).  `doc` is
 * accepted for signature parity and may be NULL (the set is not bound to it). */
int ol_dict_compile(ol_ctx *ctx, const ol_doc *doc, const uint8_t *const *patterns,
                    const size_t *lens, size_t count, ol_dict **out);
void ol_dict_free(ol_dict *dict);
/* GpuPatternSet::len (many.rs:141). */
size_t ol_dict_len(const ol_dict *dict);
/* GpuParser::search_many (many.rs:391): counts[dict_len] per pattern,
 * *total_hits = their sum, hits = min(total, max_hits) hits sorted by
 * (pattern, offset) (library-allocated; none when max_hits == 0). */
int ol_search_many(ol_ctx *ctx, const ol_doc *doc, const ol_dict *dict, size_t max_hits,
                   uint32_t *counts, uint64_t *total_hits, ol_many_hit **hits, size_t *n_hits);
/* Offset-major form: every (offset, pattern id) sorted by offset then id,
 * written to device memory (d_offsets u64, d_ids u32; capacity `cap`, the
 * first `cap` kept); *total = all matches.  Synchronous. */
int ol_search_many_device(ol_ctx *ctx, const ol_doc *doc, const ol_dict *dict, uint64_t *d_offsets,
                          uint32_t *d_ids, size_t cap, uint64_t *total);

/* ---- batches (no counterpart in the reference, which handles one document
 * per call): parse, tree and dictionary search of many documents as one
 * launch sequence over one HBM text, replayable as a captured hipGraph.
 * Results per document are those of the single-document calls. */
typedef struct ol_batch ol_batch;
/* Uploads `count` documents back to back (empty ones allowed). */
int ol_batch_create(ol_ctx *ctx, const uint8_t *const *docs, const size_t *lens, size_t count, ol_batch **out);
void ol_batch_free(ol_batch *batch);
size_t ol_batch_doc_count(const ol_batch *batch);
/* Sum of the document lengths. */
size_t ol_batch_text_bytes(const ol_batch *batch);
/* Parse + tree of every document, then (dict non-NULL) its dictionary search.
 * The first call for a given dictionary, flags and launch tuning sizes the
 * outputs (synchronous); later calls launch the same sequence asynchronously,
 * as a captured hipGraph with OL_BATCH_GRAPH.  OL_BATCH_SEARCH_ONLY skips the
 * parse and tree (a dictionary is then required, and only ol_batch_matches
 * has results).  ol_ctx_synchronize waits. */
#define OL_BATCH_GRAPH 1
#define OL_BATCH_SEARCH_ONLY 2
int ol_batch_run(ol_ctx *ctx, ol_batch *batch, const ol_dict *dict, int flags);
/* Total lines; per document its first line (in the batch's line order) and
 * line count (either pointer may be NULL; arrays of ol_batch_doc_count). */
int ol_batch_lines(ol_ctx *ctx, const ol_batch *batch, uint64_t *total, uint32_t *doc_line_start,
                   uint32_t *doc_line_count);
/* Per-line arrays of document `doc` (its line count entries each; any pointer
 * may be NULL): offsets relative to the document, classes (NodeKind as u8),
 * and tree links local to the document. */
int ol_batch_download_doc(ol_ctx *ctx, const ol_batch *batch, size_t doc, uint32_t *offsets, uint8_t *classes,
                          uint32_t *parent, uint32_t *subtree_end, uint32_t *first_child, uint32_t *next_sibling);
/* Dictionary matches of document `doc` sorted by (offset, id), offsets
 * relative to the document: *count = all of them, the first `cap` copied. */
int ol_batch_matches(ol_ctx *ctx, ol_batch *batch, size_t doc, uint64_t *count, uint32_t *offsets, uint32_t *ids,
                     size_t cap);

/* ---- serving (no counterpart in the reference): a continuous-batching
 * request loop.  Each worker slot is a context (its own HIP stream) plus one
 * device arena allocated at ol_server_create; a slot takes the oldest pending
 * requests that belong together (parse requests, or searches with the same
 * dictionary), as many as fit its arena and the batch limits, runs them as one
 * batch carved from the arena (no device allocation per batch, one read-back
 * per batch) and completes them; requests arriving meanwhile join the next
 * batch any slot forms, so up to `slots` batches run on the device at once.
 * Dictionaries and documents put with ol_server_put_doc stay resident in HBM.
 * The HBM budget is charged with the arenas, the dictionaries and the
 * resident documents at their actual device sizes.  Results equal the
 * single-document calls.  Documents passed to ol_server_submit must stay
 * valid until the request completes. */
typedef struct ol_server ol_server;
typedef struct ol_server_config {
    size_t hbm_budget;        /* device bytes: slot arenas + dictionaries + resident documents (0: 64 GiB) */
    uint32_t slots;           /* concurrent batches (0: 4, at most 16) */
    uint32_t max_batch_docs;  /* requests per batch (0: 256) */
    size_t max_batch_bytes;   /* document bytes per batch, and the largest request (0: 256 MiB) */
} ol_server_config;
#define OL_REQ_PARSE 0        /* offsets, classes and tree links */
#define OL_REQ_SEARCH 1       /* dictionary matches */
typedef struct ol_serve_stats {
    uint64_t requests, parse_requests, search_requests, rejected, bytes, batches, batched_requests;
    double seconds, qps_offered, qps_achieved, gbps, p50_ms, p90_ms, p99_ms, max_ms, mean_ms;
} ol_serve_stats;
/* Fails with OL_ERR_NOMEM when the slots' arenas (each sized for a batch of
 * max_batch_bytes) exceed the budget. */
int ol_server_create(const ol_server_config *cfg, ol_server **out);
/* Finishes the pending requests, then releases everything. */
void ol_server_destroy(ol_server *server);
/* Message of the calling thread's last failed ol_server_* call. */
const char *ol_server_last_error(void);
/* Device bytes of one slot's arena; device bytes charged to the budget now. */
size_t ol_server_arena_bytes(const ol_server *server);
size_t ol_server_hbm_used(ol_server *server);
/* Compiles a dictionary into HBM (OL_ERR_NOMEM beyond the budget). */
int ol_server_add_dict(ol_server *server, const uint8_t *const *patterns, const size_t *lens, size_t count,
                       uint32_t *dict_id);
/* Uploads, scans and builds the tree of a document once; text, per-line
 * arrays and links stay resident in HBM until ol_server_drop_doc (charged to
 * the budget at their device size; OL_ERR_NOMEM beyond it). */
int ol_server_put_doc(ol_server *server, const uint8_t *doc, size_t len, uint64_t *doc_id);
int ol_server_drop_doc(ol_server *server, uint64_t doc_id);
/* kind OL_REQ_PARSE or OL_REQ_SEARCH (dict_id from ol_server_add_dict).  A
 * request larger than max_batch_bytes fails with OL_ERR_NOMEM. */
int ol_server_submit(ol_server *server, int kind, uint32_t dict_id, const uint8_t *doc, size_t len,
                     int keep_results, uint64_t *req_id);
/* A request on a resident document: a parse returns its stored lines and
 * links (no upload, no scan), a search runs over its resident text. */
int ol_server_submit_doc(ol_server *server, int kind, uint32_t dict_id, uint64_t doc_id, int keep_results,
                         uint64_t *req_id);
/* Waits for a request (timeout_s <= 0: no limit); its status. */
int ol_server_wait(ol_server *server, uint64_t req_id, double timeout_s);
/* Results of a completed request (two calls: sizes with NULL arrays, then data). */
int ol_server_result_lines(ol_server *server, uint64_t req_id, size_t *lines, uint32_t *offsets, uint8_t *classes,
                           uint32_t *parent, uint32_t *subtree_end, uint32_t *first_child, uint32_t *next_sibling);
int ol_server_result_matches(ol_server *server, uint64_t req_id, size_t *count, uint32_t *offsets, uint32_t *ids);
/* Submission, batch start and completion times (seconds, steady clock), and
 * the slot that ran the request. */
int ol_server_timing(ol_server *server, uint64_t req_id, double *submit, double *start, double *done);
int ol_server_slot_of(ol_server *server, uint64_t req_id, uint32_t *slot);
int ol_server_release(ol_server *server, uint64_t req_id);
/* Open-loop load: `nreq` requests with exponential inter-arrival times at
 * `qps` (seeded), documents taken round-robin from `docs`, a search with
 * probability search_fraction against a uniformly chosen dictionary; waits
 * for all and reports latency (completion - scheduled arrival) percentiles
 * and served GB/s ((document + result bytes) / wall time). */
int ol_server_load(ol_server *server, const uint8_t *const *docs, const size_t *lens, size_t ndocs, double qps,
                   size_t nreq, double search_fraction, uint64_t seed, ol_serve_stats *out);
/* The same request sequence (same documents, kinds, dictionaries and seed)
 * run one request at a time through the single-document entry points on one
 * context -- upload, scan, tree and download for a parse; upload, dictionary
 * search and download for a search -- the plain per-document path the
 * batched server is measured against.  Fills requests, bytes, seconds,
 * qps_achieved, gbps and the latency fields (each request's own time). */
int ol_server_load_serial(ol_server *server, const uint8_t *const *docs, const size_t *lens, size_t ndocs,
                          size_t nreq, double search_fraction, uint64_t seed, ol_serve_stats *out);

/* ---- launch tuning (no counterpart in the reference: its tile sizes are
 * compile-time constants, e.g. crates/oneline-gpu/src/scan.rs:11).  The
 * autotuner (oneline_amd/autotune.py) sets these from measurements:
 *   "scan_bpt"      bytes per lane of the scan kernel: 64, 128 or 256
 *                   (tile = 256 lanes x bpt bytes);
 *   "scan_grid"     persistent workgroups of the scan kernel (0 = as many as
 *                   are co-resident on the device, the default);
 *   "tree_rounds"   rounds of 64 lines per tree wave: 1, 4 or 16 (tile =
 *                   256 x rounds lines; 16 by default);
 *   "scan_ticket"   1: scan tiles are claimed from an atomic ticket by an
 *                   ordinary launch (safe beside other work on the device),
 *                   0: static tiles under a cooperative launch (default). */
int ol_tune_set(ol_ctx *ctx, const char *name, int value);
int ol_tune_get(ol_ctx *ctx, const char *name, int *value);

/* ---- profiling (the reference's Profile / StageTimer, crates/oneline-gpu/src/timing.rs:13-123) --- */
/* When enabled, every kernel launch is bracketed by HIP events on the
 * context's stream.  ol_profile_take waits for the stream and returns, per
 * kernel name, the launch count and the summed milliseconds, then resets. */
int ol_profile_enable(ol_ctx *ctx, int on);
int ol_profile_take(ol_ctx *ctx, char *names, size_t name_stride, uint64_t *counts, double *ms,
                    size_t cap, size_t *n);

/* Roofline calibration (no counterpart in the reference): the mean time in
 * microseconds of `reps` launches of a kernel that moves the scan's bytes over
 * the document's text with none of its work -- the text read in the scan's
 * 32 KiB tiles, and when `write` is set, lines_per_tile u32 + 2 x u8 written
 * per tile at consecutive positions (the scan's offsets, classes, keys).  The
 * write targets are allocated for the call; the document is not changed. */
int ol_calib_scan_floor(ol_ctx *ctx, const ol_doc *doc, uint32_t lines_per_tile, int write, int reps, double *us);

/* Per-call profile: the reference's Profile {wall, host_steps, gpu_passes}
 * returned by its profile_* methods (crates/oneline-gpu/src/timing.rs:109-123).
 * ol_profile_call_begin starts a record on the context; the entry point calls
 * made until ol_profile_call_end add host steps (time since the previous step,
 * in execution order, labelled like the reference's marks) and GPU passes
 * (each kernel launch sequence with its HIP-event duration, in launch order).
 * ol_profile_call_end waits for the stream and returns the wall time from the
 * start to the last host step, and the steps and passes (the first `cap` of
 * each; *n_* = all of them). */
typedef struct ol_profile_step {
    char label[48];
    double ms;
} ol_profile_step;
int ol_profile_call_begin(ol_ctx *ctx);
int ol_profile_call_end(ol_ctx *ctx, double *wall_ms, ol_profile_step *host_steps, size_t host_cap, size_t *n_host,
                        ol_profile_step *gpu_passes, size_t gpu_cap, size_t *n_gpu);

#ifdef __cplusplus
}
#endif

#endif /* ONELINE_HIP_H */
