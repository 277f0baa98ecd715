// internal.h -- shared declarations of the HIP engine (device helpers, kernel
// argument blocks and host-side launchers).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define OL_NONE 0xFFFFFFFFu

// NodeKind values, crates/oneline-core/src/lib.rs:41-51.
enum : uint32_t {
    OLK_BLANK = 0, OLK_TEXT = 1, OLK_HEADING = 2, OLK_ELEMENT = 3,
    OLK_LIST = 4, OLK_CONNECT = 5, OLK_ORDERED = 6, OLK_COMMENT = 7
};

// Tree key byte per line: depth for content lines, KEY_DEEP when the depth
// does not fit a byte, KEY_BLANK for blank lines (outside the tree).
#define KEY_BLANK 255u
#define KEY_DEEP 254u

// Device text buffers are padded with zeros to a multiple of TEXT_ALIGN plus
// TEXT_SLACK bytes, so tile loads never need a bounds check.
#define TEXT_ALIGN (64u * 1024u)
#define TEXT_SLACK (64u * 1024u)
// Per-line device arrays are padded to a multiple of LINE_ALIGN entries.
#define LINE_ALIGN 4096u

struct ol_node_dev {
    uint32_t depth, kind, span_start, span_len;
};

// ---- decoupled look-back status words ------------------------------------
// One 64-bit word per tile: flag in the top two bits, value in the rest.
#define LB_FLAG_AGG 1ull
#define LB_FLAG_INCL 2ull
#define LB_SHIFT 62
#define LB_MASK ((1ull << LB_SHIFT) - 1)

// Look-back status words are self-describing (flag + value in one word), so
// they are read and written with RELAXED agent-scope atomics.  Acquire/release
// at agent scope would compile to whole-L2 write-back / invalidate on this
// multi-XCD part (each XCD has its own L2), which serialises the whole chip.
__device__ __forceinline__ uint64_t ld_relaxed_u64(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed_u64(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Read-modify-write forms of the same: performed where agent-scope atomics are
// resolved for the whole device, so a poll sees another agent's update even
// when its own cache holds the line (a plain relaxed load may be served from
// the local L2 of this XCD, which does not see other XCDs' stores until the
// line leaves it).  Used for every word one agent waits on another to write.
__device__ __forceinline__ uint64_t ld_coherent_u64(uint64_t *p)
{
    return __hip_atomic_fetch_or(p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent_u64(uint64_t *p, uint64_t v)
{
    (void)__hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A non-temporal 16-byte load, for text a pass reads once and does not read
// again: the caches keep less of it.  Reading 2 GiB in 32 KiB tiles, with 6
// bytes per line written, takes 385 us this way and 442 us with plain loads.
__device__ __forceinline__ uint4 ld_stream_u4(const uint4 *p)
{
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 w = __builtin_nontemporal_load((const u32x4 *)p);
    return make_uint4(w.x, w.y, w.z, w.w);
}

// ---- wave-wide scans on DPP lane moves (no LDS round trip) -----------------
// DPP controls of the GFX9 encoding: row_shr:n = 0x110 + n (within rows of 16
// lanes), row_bcast:15 = 0x142 (lane 15 of each row to the next row, written
// in the rows of row_mask), row_bcast:31 = 0x143 (lane 31 to rows 2 and 3),
// wave_shr:1 = 0x138.  A lane whose source is outside its row, or whose row is
// masked off, gets `old`.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_move(uint32_t old, uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, ROW_MASK, 0xF, false);
}

// Inclusive prefix sum over the 64 lanes (Hillis-Steele in rows, then the row
// totals carried with two broadcasts).
__device__ __forceinline__ uint32_t wave_incl_add(uint32_t x)
{
    x += dpp_move<0x111>(0u, x);
    x += dpp_move<0x112>(0u, x);
    x += dpp_move<0x114>(0u, x);
    x += dpp_move<0x118>(0u, x);
    x += dpp_move<0x142, 0xA>(0u, x);
    x += dpp_move<0x143, 0xC>(0u, x);
    return x;
}

// Inclusive prefix maximum over the 64 lanes (signed; -1 is the identity for
// values >= -1).
__device__ __forceinline__ int32_t wave_incl_max(int32_t x)
{
    auto mx = [](int32_t a, uint32_t b) { return a > (int32_t)b ? a : (int32_t)b; };
    x = mx(x, dpp_move<0x111>(~0u, (uint32_t)x));
    x = mx(x, dpp_move<0x112>(~0u, (uint32_t)x));
    x = mx(x, dpp_move<0x114>(~0u, (uint32_t)x));
    x = mx(x, dpp_move<0x118>(~0u, (uint32_t)x));
    x = mx(x, dpp_move<0x142, 0xA>(~0u, (uint32_t)x));
    x = mx(x, dpp_move<0x143, 0xC>(~0u, (uint32_t)x));
    return x;
}

// The value of the next lower lane (lane 0 gets `first`).
__device__ __forceinline__ uint32_t wave_shr1(uint32_t x, uint32_t first)
{
    return dpp_move<0x138>(first, x);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o, 64);
    return v;
}

// Orders a wave's LDS accesses around a point (a wave's LDS operations
// complete in order; the fence keeps the compiler from moving them).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Exclusive look-back over per-tile status words for the sum of the values of
// all tiles before `t` (one full wave calls this; every lane gets the result).
// Tile t must already have published its own aggregate.
//
// The status words are read with uncached (agent-scope) loads, and on this
// part their throughput, not their latency, is what limits a look-back: a wave
// reads one 64-tile window, keeps every ready entry up to the first
// not-ready one, and then polls only that tile, with one lane and a sleep,
// instead of re-reading the window.
// Status word of tile t is status[t * LB_STRIDE]: spreading the words of
// neighbouring tiles over memory channels.
#ifndef LB_STRIDE
#define LB_STRIDE 8
#endif

// Polls of one not-ready status word before the look-back gives up (about a
// second): it then sets *err and returns, so a launch whose workgroups are not
// all resident ends with an error instead of waiting forever.
constexpr uint32_t LB_SPIN_LIMIT = 1u << 22;

// Bits of the sticky device error word (one per scratch owner): kernels only
// OR bits into it; the host reads it after a launch sequence, reports the
// error and only then clears it, so a give-up is never lost to a later stage.
enum : uint32_t {
    OLE_PREFIX_WAIT = 1u,   // cooperative launch: a tile waited too long for its prefix word
    OLE_SCANNER = 2u,       // cooperative launch: the prefix scanner waited too long for a tile count
    OLE_LOOKBACK = 4u,      // ticket-ordered launch: a decoupled look-back waited too long for a predecessor
    OLE_DICT_SCAN = 8u,     // dictionary tile-count scan: a chunk waited too long for an earlier chunk
};

template <int NWIN = 1>
__device__ __forceinline__ uint64_t lookback_sum(const uint64_t *status, uint32_t t, uint32_t *err)
{
    const int lane = threadIdx.x & 63;
    uint64_t prefix = 0;
    int64_t top = (int64_t)t - 1;
    while (true) {
        // NWIN windows of 64 status words in flight at once: window k covers
        // top - 64k - 63 ..= top - 64k, lane l reading top - 64k - l.
        uint64_t st[NWIN];
#pragma unroll
        for (int k = 0; k < NWIN; k++) {
            const int64_t idx = top - 64 * k - lane;
            st[k] = (idx >= 0) ? ld_relaxed_u64(status + idx * LB_STRIDE) : (LB_FLAG_INCL << LB_SHIFT);
        }
        bool done = false, waited = false;
#pragma unroll
        for (int k = 0; k < NWIN; k++) {
            const uint64_t flag = st[k] >> LB_SHIFT;
            const uint64_t incl = __ballot(flag == LB_FLAG_INCL);
            const uint64_t notready = __ballot(flag == 0);
            const int fi = incl ? (int)__builtin_ctzll(incl) : 64;
            const int fn = notready ? (int)__builtin_ctzll(notready) : 64;
            if (fn < fi) {
                prefix += wave_sum_u64((lane < fn) ? (st[k] & LB_MASK) : 0);
                top -= 64 * k + fn;
                uint32_t spins = 0;
                if (lane == 0) {
                    while ((ld_coherent_u64(const_cast<uint64_t *>(status) + top * LB_STRIDE) >> LB_SHIFT) == 0 &&
                           ++spins < LB_SPIN_LIMIT)
                        __builtin_amdgcn_s_sleep(8);
                    if (spins >= LB_SPIN_LIMIT)
                        atomicOr(err, OLE_LOOKBACK);
                }
                if (__shfl(spins, 0, 64) >= LB_SPIN_LIMIT)
                    return prefix;   // gave up: the result is flagged invalid
                waited = true;
                break;
            }
            prefix += wave_sum_u64((lane <= fi) ? (st[k] & LB_MASK) : 0);
            if (fi < 64) {
                done = true;
                break;
            }
        }
        if (done)
            break;
        if (!waited)
            top -= 64 * NWIN;
    }
    return prefix;
}

// Status per tile of the prefix scanner's kernels: the tile's count (written
// by the tile's workgroup, read by the scanner) and its inclusive prefix
// (written by the scanner, read by the tile's workgroup), all accessed with
// the read-modify-write atomics.
constexpr uint32_t SCAN_STATUS_WORDS = 2;                  // u64 words per tile
constexpr uint32_t SCAN_PREFIX_WORD = 1;                   // offset of the prefix word
#ifndef SCAN_PREFIX_WINDOWS
#define SCAN_PREFIX_WINDOWS 8                              // 64-tile windows the prefix scanner reads per poll
#endif

// The prefix scanner of the persistent tile kernels (one wave of the one
// workgroup that takes no tiles; the scan and the literal search use it): the
// count word of tile t holds its line count (flag AGG) once the tile is
// counted; the scanner reads the count words in tile order, NWS windows of 64
// per poll, and writes every complete window's inclusive prefixes (flag INCL)
// to the tiles' prefix words, carrying the running total from window to
// window.  A window with a tile not counted yet is polled again.  Tiles wait
// only for their own prefix word, never for a chain of their neighbours'
// look-backs.
template <int NWS>
__device__ __forceinline__ void scan_prefixes(uint64_t *status, uint32_t T, uint32_t *err)
{
    const uint32_t lane = threadIdx.x & 63;
    uint64_t carry = 0;
    uint32_t t0 = 0, spins = 0;
    while (t0 < T) {
        uint64_t st[NWS];
#pragma unroll
        for (int k = 0; k < NWS; k++) {
            const uint32_t idx = t0 + 64 * k + lane;
            st[k] = idx < T ? ld_coherent_u64(status + (size_t)idx * SCAN_STATUS_WORDS) : (LB_FLAG_AGG << LB_SHIFT);
        }
        uint32_t done = 0;
#pragma unroll
        for (int k = 0; k < NWS; k++) {
            if (done != (uint32_t)k || __ballot((st[k] >> LB_SHIFT) == 0))
                break;
            const uint32_t inc = wave_incl_add((uint32_t)(st[k] & LB_MASK));   // a tile has < 2^17 lines
            const uint32_t idx = t0 + 64 * k + lane;
            if (idx < T)
                st_coherent_u64(status + (size_t)idx * SCAN_STATUS_WORDS + SCAN_PREFIX_WORD,
                                (LB_FLAG_INCL << LB_SHIFT) | (carry + inc));
            carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
            done++;
        }
        t0 += 64 * done;
        if (done == 0) {
            if (++spins >= LB_SPIN_LIMIT) {
                if (lane == 0)
                    atomicOr(err, OLE_SCANNER);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
        } else {
            spins = 0;
        }
    }
}

// Workgroups of kernel f that can be resident at once on the device, for the
// persistent launches whose look-backs need every workgroup resident: the
// occupancy API's count per CU, capped by the CU's LDS divided by the
// kernel's LDS rounded up to 4 KiB (margin for the allocation granule, which
// the API does not model).  0 on failure.
uint32_t ol_coresident(const void *f, int threads, size_t dyn_lds);

// ---- kernel argument blocks ----------------------------------------------
struct ScanArgs {
    const uint8_t *text;     // padded device text
    uint32_t n;              // text length (bytes)
    uint32_t num_tiles;
    uint64_t cap;            // capacity of the per-line outputs (lines)
    uint32_t *offsets;       // line start offsets (u32)          [cap]
    uint8_t *classes;        // NodeKind as u8                     [cap]
    uint8_t *keys;           // tree key byte                      [cap]
    ol_node_dev *nodes;      // FlatNode (nullable)                [cap]
    // Status words per tile.  Cooperative launch: two per tile (count,
    // prefix), zero at launch, and each tile's workgroup zeroes its own two
    // once it has its prefix, so the words are zero again after the launch
    // (no clear launch between calls, and a captured graph replays cleanly).
    // Ticket order: LB_STRIDE per tile, cleared before each launch.
    uint64_t *status;
    uint64_t *meta;          // meta[0] <- line count
    uint32_t *lb_err;        // sticky device error word (OLE_* bits)
    uint32_t *tile_ctr;      // non-null: tiles claimed from this ticket (zero at launch)
    // Batches (ndocs > 0): the text is the documents back to back, each ending
    // in '\n'.  Offsets are written relative to the line's document, and the
    // index of each document's first line to doc_line_start.
    uint32_t ndocs;
    const uint32_t *doc_starts;   // [ndocs] byte start of each document (ascending)
    const uint32_t *tile_doc;     // [num_tiles + 1] first document starting at or after tile t
    uint32_t *doc_line_start;     // [ndocs] out
};

struct TreeArgs {
    const uint8_t *keys;     // tree key byte per line (padded to LINE_ALIGN)
    uint64_t lines;          // line count
    uint32_t num_tiles;      // ceil(lines / ol_tree_tile_lines(rounds)), or the batch's tile count
    uint32_t *parent, *first_child, *next_sibling, *subtree_end;  // nullable each
    uint32_t *sum_last;      // per tile x level: last content line + 1 in the tile (0 = none)
    uint32_t *sum_first;     // per tile x level: first content line in the tile (OL_NONE = none)
    uint32_t *in_last;       // per tile x level: sum_last over all earlier tiles (max)
    uint32_t *out_first;     // per tile x level: sum_first over all later tiles (min)
    uint32_t *wsum_last;     // per wave (tile * TREE_WAVES + wave) x level, wave-major: as sum_last
    uint32_t *wsum_first;    //   and sum_first over the wave's own lines (written by the summary)
    uint32_t *err;           // set to 1 (by the walk) if a content line is too deep; zeroed by the summary
    // Batches (nsegs > 0): one tree per document (segment); a tile holds lines
    // of one document only, and every link is local to its document.
    uint32_t nsegs;
    const uint32_t *seg_line0;   // [nsegs] global index of the document's first line
    const uint32_t *seg_lines;   // [nsegs] its line count
    const uint32_t *seg_tile0;   // [nsegs] its first tile
    const uint32_t *tile_seg;    // [num_tiles] document of each tile
};

struct SearchArgs {
    const uint8_t *text;
    uint32_t n;
    uint32_t num_tiles;
    uint32_t m;              // pattern length 1..64
    uint32_t wo;             // offset in the pattern of the window the kernel tests
    uint32_t dense;          // the window is frequent: 4 vectors per lane (16 KiB tiles), else 8 (32 KiB)
    uint32_t win, win_mask;  // pattern bytes wo .. wo + min(m,4) and their mask
    uint32_t pw[16];         // the pattern as words (zero past its end)
    uint32_t vmask[16];      // per word, the bytes left to verify: inside the pattern, outside the window
    uint64_t cap;            // output capacity (matches)
    uint32_t *lb_err;        // sticky device error word (OLE_* bits)
    uint64_t *out64;         // offsets as u64 (nullable)
    uint32_t *out32;         // offsets as u32 (nullable)
    uint64_t *status;        // two words per tile (count, prefix): zero at launch, zeroed again by each tile
    uint64_t *meta;          // meta[0] <- total matches
};

// Compiled dictionary.  Every pattern has a key class and a key: DK_LONG for
// patterns of 4 bytes and more (key: their first 4 bytes), DK_3 / DK_2 / DK_1
// for patterns of 3, 2 and 1 bytes (key: all their bytes).  Patterns with the
// same class and key form a run (sorted by id).  One table maps (class, key)
// to a 16-byte slot at one of two places (cuckoo hashing: dk_bucket1,
// dk_bucket2), so a lookup is two 16-byte loads, both at once (a lane's load
// instructions, not its bytes, are what the L1 serves one by one):
//   slot = {key, info, a, b}; info = id | len << 22 | DK_VALID | class << 30;
//   one pattern (len > 0): a = its bytes 4..7, b = bytes 8..11 when it ends by
//     byte 12, else the pool offset of its bytes 8..; decided from the slot
//     alone up to 12 bytes;
//   a run (len == 0): a = first entry, b = its length; entries are
//     {info, a, b, 0} of each pattern, in id order.
// LDS filters: DK_1 and DK_2 direct bitmaps (indices 1, 2), and one prefix
// bitmap for DK_3 and DK_LONG (index 4): a word per hash of the window's
// bytes 0..2, a bit per low 5 bits of its byte 3 (dict_pword, dict_pbit); a
// 3-byte key sets its whole word.
enum : uint32_t { DK_LONG = 0, DK_3 = 1, DK_2 = 2, DK_1 = 3 };
constexpr uint32_t DK_BUCKET = 1;   // slots per bucket
constexpr uint32_t DK_LEN_SHIFT = 22, DK_VALID = 1u << 29, DK_CLASS_SHIFT = 30;
__host__ __device__ __forceinline__ uint32_t dk_info(uint32_t id, uint32_t len, uint32_t cls)
{
    return id | len << DK_LEN_SHIFT | DK_VALID | cls << DK_CLASS_SHIFT;
}
__host__ __device__ __forceinline__ uint32_t dk_bucket1(uint32_t key, uint32_t cls, uint32_t bits)
{
    return ((key ^ cls * 0x27D4EB2Fu) * 0x9E3779B1u) >> (32 - bits);
}
__host__ __device__ __forceinline__ uint32_t dk_bucket2(uint32_t key, uint32_t cls, uint32_t bits)
{
    uint32_t x = key ^ (cls + 1u) * 0x165667B1u;
    x ^= x >> 15;
    const uint32_t b = (x * 0x85EBCA6Bu) >> (32 - bits);
    return b ^ (b == dk_bucket1(key, cls, bits) ? 1u : 0u);   // never the first bucket
}

struct DictDev {
    uint32_t class_mask;     // bit k set if class k (DK_*) has patterns
    uint32_t fbits[6];       // log2 bits of the filters 1, 2 and 4 (0: absent)
    uint32_t foff[6];        // word offset of each filter in `filter`
    uint32_t fwords;         // total filter words
    uint32_t bbits;          // log2 buckets of the table
    const uint32_t *filter;
    const uint4 *table;      // (DK_BUCKET << bbits) slots
    const uint4 *ents;       // run entries
    const uint32_t *pool;    // pattern bytes 8.. of patterns longer than 12 bytes, word aligned, zero padded
    const uint32_t *ent_ids; // the pattern id of each entry (the entries' info words without length and class)
    uint32_t max_len;
    uint32_t pf_hashed;      // prefix filter keying: 0 dict_pword/dict_pbit, 1 dict_ph (dense filters)
    // Merged runs (1): the run of a key also holds every shorter pattern that
    // is a prefix of the key (a class-L run the 3-, 2- and 1-byte patterns on
    // its first 3, 2 and 1 bytes; a class-3 run the 2- and 1-byte ones; a
    // class-2 run the 1-byte ones), sorted by id.  A position's matches are
    // then all in the run of the first class whose slot holds its key, in
    // the order L, 3, 2, 1.  0: each run holds its own class only (the
    // compiler's choice when merging would copy too many short patterns).
    uint32_t merged;
};

// Hashes shared by the compiler (host) and the search kernel (device).
__host__ __device__ __forceinline__ uint32_t dict_mul24(uint32_t a, uint32_t b)
{
    return (a & 0xFFFFFFu) * (b & 0xFFFFFFu);
}
// Prefix filter: 2^DICT_PBITS bits, 64 KiB of LDS, keyed on a 4-byte window w
// (first byte lowest) cut to its bytes 0..2 and the low 5 bits of byte 3.
// Two keyings, chosen per dictionary:
// - sparse filters: the word is a multiplicative hash of bytes 0..2 (one
//   full-rate 24-bit multiply, its top DICT_PBITS - 5 bits), the bit is byte
//   3's low 5 bits, and a 3-byte key sets its whole word.  On the device the
//   bit costs nothing: a bit-field extract takes the low 5 bits of its offset
//   operand, and the window three positions on has this window's byte 3 as
//   its lowest byte;
// - dense filters: word and bit both from a hash of the cut window (a 24-bit
//   multiply-add of the window and the window shifted by 3, which drops byte
//   3's top bits), and a 3-byte key sets its 32 windows' bits.  One operation
//   more per position, but hot 3-byte prefixes of the text no longer share a
//   word with every pattern that starts with them.
constexpr uint32_t DICT_PBITS = 19;
__host__ __device__ __forceinline__ uint32_t dict_pword(uint32_t w) { return dict_mul24(w, 0x9E3779u) >> (32 - (DICT_PBITS - 5)); }
__host__ __device__ __forceinline__ uint32_t dict_pbit(uint32_t w) { return (w >> 24) & 31u; }
__host__ __device__ __forceinline__ uint32_t dict_ph(uint32_t w) { return dict_mul24(w, 0x9E3779u) + (w << 3); }
__host__ __device__ __forceinline__ uint32_t dict_ph_word(uint32_t h) { return h >> (32 - (DICT_PBITS - 5)); }
// The dense keying above this many prefix filter bits set (1/8 of them).  On
// the 2 GiB benchmark text, sparse / dense: 1,024 patterns 0.78 / 0.82 ms,
// 4,096 1.58 / 1.60 ms, 16,384 (6% of the bits) 4.9 / 5.2 ms, 65,536 (25%)
// 25.2 / 20.8 ms.
constexpr uint32_t DICT_PF_DENSE_BITS = (1u << DICT_PBITS) / 8;

struct ManyArgs {
    const uint8_t *text;
    uint32_t n;
    uint32_t num_tiles;      // SEARCH_TILE tiles (tile_doc)
    uint32_t num_wtiles;     // wave tiles of MANY_WT_BYTES: num_tiles * SEARCH_TILE / MANY_WT_BYTES
    DictDev d;
    uint64_t cap;
    uint64_t *out_off;       // u64 offsets (nullable)
    uint32_t *out_id;        // pattern ids (nullable)
    uint64_t *meta;          // meta[0] <- total matches (written by the tile-count scan)
    uint32_t *tile_count;    // [num_wtiles] matches per wave tile
    uint64_t *tile_off;      // [num_wtiles] first output index of each wave tile
    uint64_t *scan_sums;     // [num_wtiles / MANY_SCAN_CHUNK + 1] chunk totals of the scan (zero at launch)
    uint32_t *lb_err;        // sticky device error word (OLE_* bits)
    uint32_t job_tiles;      // write pass: wave tiles per wave job (set by the launcher)
    // Candidate store (nullable): the candidates of wave tile wt that have
    // matches go, 16 bytes each, to entries wt * MANY_WT_BYTES / 4 .. when
    // they fit; tile_list[wt] records their number, or ~0 when they were not
    // stored.
    uint32_t *cands;
    uint64_t *tile_list;     // [num_wtiles]
    // Batches (ndocs > 0): a match must end inside the document it starts in
    // (the bytes between documents belong to none).
    uint32_t ndocs;
    const uint32_t *doc_starts;   // [ndocs] ascending
    const uint32_t *doc_ends;     // [ndocs] exclusive end of each document's bytes
    const uint32_t *tile_doc;     // [num_tiles] last document starting at or before tile t
};

// ---- launchers (each returns the launch error, if any) ---------------------
hipError_t ol_launch_clear_u64(uint64_t *p, size_t count, hipStream_t s);
// Persistent scan: `grid` co-resident workgroups (cooperative launch), at most
// ol_scan_max_grid(bpt); bpt (bytes per lane per tile) is 64, 128 or 256, i.e.
// tiles of 1, 2 or 4 LDS-staged quarters of 256 x 64 bytes.
hipError_t ol_launch_scan(const ScanArgs &a, uint32_t bpt, uint32_t grid, bool cooperative, hipStream_t s);
uint32_t ol_scan_max_grid(uint32_t bpt);
// summary + scan + walk; `rounds` of 64 lines per wave: 1, 4 or 16 (tile = 256 x rounds lines)
hipError_t ol_launch_tree(const TreeArgs &a, uint32_t rounds, hipStream_t s);
uint32_t ol_tree_tile_lines(uint32_t rounds);
hipError_t ol_launch_search(const SearchArgs &a, hipStream_t s);
// The scan's traffic without its work (calib.hip): the text read in 32 KiB
// tiles, and with `write` lpt u32 + 2 x u8 written per tile.
hipError_t ol_launch_stream_floor(const uint8_t *text, uint32_t tiles, uint32_t lpt, bool write, uint32_t *off,
                                  uint8_t *cls, uint8_t *key, uint32_t *sink, hipStream_t s);
uint32_t ol_literal_tile(bool dense);   // text bytes per literal search tile
// Scratch of the allocation-free literal search: OL_NOALLOC_WORDS u64 words
// in a static device array (search.hip); [0, 8) meta, word 8 the sticky error
// word, status words from OL_NOALLOC_STATUS on, two per tile.
constexpr size_t OL_NOALLOC_STATUS = 16;
constexpr size_t OL_NOALLOC_MAX_TILES = (4294967292ull + 16383) / 16384;
constexpr size_t OL_NOALLOC_WORDS = OL_NOALLOC_STATUS + 2 * OL_NOALLOC_MAX_TILES;
// The array's device address on `device` (which must be current), resolved
// with hipGetSymbolAddress on the first call for that device.
hipError_t ol_noalloc_scratch(int device, uint64_t **out);
hipError_t ol_launch_many(const ManyArgs &a, hipStream_t s);
// The write pass alone, after an ol_launch_many of the same arguments that
// counted and stored candidates without outputs.
hipError_t ol_launch_many_write(const ManyArgs &a, hipStream_t s);

// Tile geometry of each kernel.
constexpr uint32_t SCAN_THREADS = 256;
constexpr uint32_t SCAN_BPT = 128;                          // default bytes per lane (32 KiB tiles)
constexpr uint32_t SCAN_TILE_MAX = SCAN_THREADS * 256;
constexpr uint32_t SCAN_HALO = 256;
constexpr uint32_t TREE_D = 32;
constexpr uint32_t TREE_WAVES = 4;   // waves per tree tile
// The tree's level states for `tiles` tiles in one buffer of
// ol_tree_sum_words(tiles) u32: per tile and level sum_last, sum_first,
// in_last, out_first (level-major), then per wave wsum_last, wsum_first.
inline size_t ol_tree_sum_words(size_t tiles) { return tiles * TREE_D * (4 + 2 * TREE_WAVES); }
inline void ol_tree_set_sums(TreeArgs &a, uint32_t *base, size_t tiles)
{
    const size_t part = tiles * TREE_D;
    a.sum_last = base;
    a.sum_first = base + part;
    a.in_last = base + 2 * part;
    a.out_first = base + 3 * part;
    a.wsum_last = base + 4 * part;
    a.wsum_first = base + (4 + TREE_WAVES) * part;
}
constexpr uint32_t SEARCH_THREADS = 256;
constexpr uint32_t SEARCH_BPT = 64;   // lanes keep 64-bit position masks
constexpr uint32_t SEARCH_TILE = SEARCH_THREADS * SEARCH_BPT;
constexpr uint32_t MANY_THREADS = 512;     // dictionary search: 8 independent waves per workgroup
#ifndef OL_MANY_WT_BYTES
#define OL_MANY_WT_BYTES 8192
#endif
constexpr uint32_t MANY_WT_BYTES = OL_MANY_WT_BYTES;   // dictionary wave tile: 4 parts of 64 lanes x 32 positions
constexpr uint32_t MANY_SCAN_CHUNK = 8192; // wave tiles per workgroup of the dictionary's tile scan
struct QueryArgs {
    const uint8_t *text;
    uint32_t n;
    uint64_t lines;
    const uint32_t *offsets;
    const uint8_t *classes, *keys;
    const ol_node_dev *nodes;   // nullable
    uint32_t kind, depth;
    uint32_t r0, r1;            // line range
    uint32_t *block_count;      // [blocks]
    uint64_t *block_off;        // [blocks]
    uint32_t *out;              // [cap] (nullable)
    uint64_t cap;
};
hipError_t ol_launch_query(const QueryArgs &a, uint64_t *total, hipStream_t s);
// The AST bridge's structural pass (scan.hip): per line the bridge flags
// (1 literal, 2 block, 4 cut, 8 pipe child, 16 bare blank) and the container
// line of each block.  flags and container cover whole blocks of 4,096 lines.
struct BridgeArgs {
    const uint8_t *text;
    uint32_t n;
    uint64_t lines;
    const uint32_t *offsets;
    const uint8_t *classes, *keys;
    const ol_node_dev *nodes;   // nullable
    const uint32_t *parent;
    uint8_t *flags;             // [ol_bridge_blocks * 4096]
    uint32_t *container;        // [ol_bridge_blocks * 4096]
    uint32_t *blk_last;         // [blocks] each block's last bare blank line + 1 (0: none)
    uint32_t *blk_in;           // [blocks] the last one before each block
};
hipError_t ol_launch_bridge(const BridgeArgs &a, hipStream_t s);
uint32_t ol_bridge_blocks(uint64_t lines);
uint32_t ol_query_blocks(uint32_t r0, uint32_t r1);   // count blocks of a query over lines [r0, r1)
hipError_t ol_launch_parse_given(const uint8_t *text, uint32_t n, const uint32_t *offsets,
                                 uint64_t lines, ol_node_dev *nodes, hipStream_t s);
hipError_t ol_launch_nodes_to_lines(const ol_node_dev *nodes, uint64_t lines, uint8_t *classes, uint8_t *keys,
                                    hipStream_t s);
size_t ol_many_lds_bytes(const DictDev &d);
