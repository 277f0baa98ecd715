// scan.hip -- newline scan + line classification in one pass over the text.
//
// Replaces the reference's three scan passes and its per-line parse pass
// (crates/oneline-gpu/src/scan.wgsl:20-237, crates/oneline-gpu/src/shader.wgsl:34-134)
// with a single kernel that reads every text byte from HBM once:
//
//   1. a workgroup claims the next tile (dynamic index, so predecessors are
//      always resident) and stages it, plus a 256-byte halo before it, in LDS
//      with coalesced 16-byte loads;
//   2. each lane turns its 64 contiguous bytes into a 64-bit terminator mask
//      (SWAR '\n' test), counts it, and the workgroup scans the counts;
//   3. wave 0 publishes the tile's line count in the tile's status word; one
//      workgroup of the grid does no tiles but turns the counts, in tile
//      order, into inclusive prefixes in the same words (scan_prefixes), and
//      a tile's wave 0 reads its prefix back from there;
//   4. each lane classifies the lines that END in its bytes (line = previous
//      terminator + 1 .. terminator) from the LDS copy and writes the line
//      offset, the kind byte, the tree key byte and optionally the 16-byte
//      FlatNode.
//
// Terminators are every '\n' plus the last byte of the text: a text that does
// not end in '\n' has one more, unterminated line (lines_in_chunk,
// crates/oneline-simd/src/lib.rs:38-44), and a text that does is unaffected.
#include "internal.h"

__global__ void clear_u64_kernel(uint64_t *p, size_t count)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride)
        p[i] = 0;
}

hipError_t ol_launch_clear_u64(uint64_t *p, size_t count, hipStream_t s)
{
    if (count == 0)
        return hipSuccess;
    size_t blocks = (count + 255) / 256;
    if (blocks > 2048)
        blocks = 2048;
    hipLaunchKernelGGL(clear_u64_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, count);
    return hipGetLastError();
}

// 4-bit mask of the bytes of a little-endian word equal to the byte replicated
// in `pat` (exact SWAR test: no false positives from borrows).
__device__ __forceinline__ uint32_t eq_nibble(uint32_t w, uint32_t pat)
{
    uint32_t x = w ^ pat;
    uint32_t nz = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x;   // bit 7 of each byte: byte != pat
    uint32_t z = ~nz & 0x80808080u;
    return (((z >> 7) * 0x00204081u) >> 21) & 0xFu;
}
__device__ __forceinline__ uint32_t nl_nibble(uint32_t w)
{
    const uint32_t x = w ^ 0x0A0A0A0Au;
    const uint32_t z = ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
    return ((z >> 7) | (z >> 14) | (z >> 21) | (z >> 28)) & 0xFu;
}

// 16-bit mask of the '\n' bytes of a 16-byte vector (bit i = byte i): an exact
// SWAR zero-byte test per word leaves 0x80 in each matching byte, and u8 dot
// products with the bit weights gather those flags, two words per result.
__device__ __forceinline__ uint32_t nl_mask16(const uint4 &v)
{
    auto z = [](uint32_t w) {
        const uint32_t x = w ^ 0x0A0A0A0Au;
        return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
    };
    const uint32_t lo = __builtin_amdgcn_udot4(z(v.y), 0x80402010u, __builtin_amdgcn_udot4(z(v.x), 0x08040201u, 0u, false), false);
    const uint32_t hi = __builtin_amdgcn_udot4(z(v.w), 0x80402010u, __builtin_amdgcn_udot4(z(v.z), 0x08040201u, 0u, false), false);
    return (lo >> 7) | ((hi >> 7) << 8);
}

__device__ __forceinline__ bool is_blank(uint32_t b) { return b == ' ' || b == '\t'; }

// classify_line, crates/oneline-simd/src/line.rs:59-81.  c1/c2 are the second and
// third content bytes, or 256 when the content is shorter.
__device__ __forceinline__ uint32_t classify(uint32_t c0, uint32_t c1, uint32_t c2)
{
    switch (c0) {
    case '=': return OLK_HEADING;
    case ':': return OLK_CONNECT;
    case '@': return OLK_ELEMENT;
    case '/': return c1 == '/' ? OLK_COMMENT : OLK_TEXT;
    case '-':
        if (is_blank(c1)) return OLK_LIST;
        if (c1 == '.' && is_blank(c2)) return OLK_ORDERED;
        return OLK_TEXT;
    default: return OLK_TEXT;
    }
}

// parse_line, crates/oneline-simd/src/line.rs:8-44, over text[s ..= e].
template <class Src>
__device__ __forceinline__ ol_node_dev parse_span(const Src &b, uint32_t s, uint32_t e_incl)
{
    ol_node_dev nd;
    uint32_t eff = e_incl + 1;
    while (eff > s) {
        uint32_t c = b(eff - 1);
        if (c != '\n' && c != '\r')
            break;
        eff--;
    }
    if (eff == s) {
        nd.depth = 0; nd.kind = OLK_BLANK; nd.span_start = s; nd.span_len = 0;
        return nd;
    }
    uint32_t pos = s;
    while (pos < eff && b(pos) == '|')
        pos++;
    uint32_t depth = pos - s;
    uint32_t cs = pos;
    while (cs < eff && is_blank(b(cs)))
        cs++;
    while (depth > 0 && cs < eff && b(cs) == '|') {
        pos = cs;
        uint32_t r0 = pos;
        while (pos < eff && b(pos) == '|')
            pos++;
        depth += pos - r0;
        cs = pos;
        while (cs < eff && is_blank(b(cs)))
            cs++;
    }
    nd.depth = depth;
    nd.span_start = cs;
    if (cs == eff) {
        nd.kind = OLK_BLANK;
        nd.span_len = 0;
        return nd;
    }
    uint32_t c0 = b(cs);
    uint32_t c1 = (cs + 1 < eff) ? b(cs + 1) : 256u;
    uint32_t c2 = (cs + 2 < eff) ? b(cs + 2) : 256u;
    nd.kind = classify(c0, c1, c2);
    nd.span_len = eff - cs;
    return nd;
}

__device__ __forceinline__ uint8_t key_of(const ol_node_dev &nd)
{
    if (nd.kind == OLK_BLANK)
        return (uint8_t)KEY_BLANK;
    return (uint8_t)(nd.depth < KEY_DEEP ? nd.depth : KEY_DEEP);
}

// Highest '\n' strictly before `end` (a multiple of 16), searched backwards one
// wave-wide KiB at a time; returns the position + 1, or 0 if there is none.
__device__ uint32_t find_line_start_global(const uint8_t *text, uint32_t end)
{
    const int lane = threadIdx.x & 63;
    int64_t hi = end;
    while (hi > 0) {
        int64_t at = hi - 16 * (int64_t)(lane + 1);
        int32_t best = -1;
        if (at >= 0) {
            uint4 v = *(const uint4 *)(text + at);
            uint32_t m = nl_nibble(v.x) | nl_nibble(v.y) << 4 | nl_nibble(v.z) << 8 | nl_nibble(v.w) << 12;
            if (m)
                best = (int32_t)(at + 31 - __clz(m));
        }
        // the lowest lane holds the highest address
        uint64_t found = __ballot(best >= 0);
        if (found) {
            int src = (int)__builtin_ctzll(found);
            int32_t p = __shfl(best, src, 64);
            return (uint32_t)p + 1;
        }
        hi -= 16 * 64;
    }
    return 0;
}

struct GlobalSrc {
    const uint8_t *g;
    __device__ __forceinline__ uint32_t operator()(uint32_t pos) const { return g[pos]; }
};

// High bit of each byte of x that is zero (exact SWAR test).
__device__ __forceinline__ uint32_t zbytes(uint32_t x)
{
    return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
}
// Bit 7 of each byte -> bits 0..3 (shifts only: a 32-bit multiply is quarter rate).
__device__ __forceinline__ uint32_t squeeze(uint32_t z)
{
    return ((z >> 7) | (z >> 14) | (z >> 21) | (z >> 28)) & 0xFu;
}

// Prefix run of `nb` bytes starting at line position `off` (words W[0..], the
// line's bytes from LDS, aligned so byte 0 is the line start): returns the
// length of the leading '|'/blank run when `from_pipe` (the line starts with
// '|'), else of the leading blank run; `pipes` gets the '|' count inside it.
// Returns nb when the run covers the whole window.
template <int NWORDS>
__device__ __forceinline__ uint32_t prefix_run(const uint32_t (&A)[NWORDS], uint32_t avail, bool from_pipe,
                                               uint32_t &pipes)
{
    uint64_t pipe = 0, stop = 0;
#pragma unroll
    for (int j = 0; j < NWORDS; j++) {
        const uint32_t w = A[j];
        const uint32_t zp = zbytes(w ^ 0x7C7C7C7Cu);
        const uint32_t zb = zbytes(w ^ 0x20202020u) | zbytes(w ^ 0x09090909u);
        const uint32_t keep = from_pipe ? (zp | zb) : zb;
        pipe |= (uint64_t)squeeze(zp) << (4 * j);
        stop |= (uint64_t)squeeze(~keep & 0x80808080u) << (4 * j);
    }
    constexpr uint32_t NB = 4 * NWORDS;
    if (avail < NB)
        stop |= ~((1ull << avail) - 1ull);
    const uint32_t r = stop ? (uint32_t)__builtin_ctzll(stop) : NB;
    pipes = (uint32_t)__popcll(pipe & ((r >= 64) ? ~0ull : ((1ull << r) - 1ull)));
    return r > NB ? NB : r;
}

// Loads NWORDS line words starting `off` bytes after the line start (LDS byte
// index rs of the line start).
template <int NWORDS>
__device__ __forceinline__ void load_words(const uint8_t *s_text, uint32_t pos, uint32_t (&A)[NWORDS])
{
    const uint32_t *wp = (const uint32_t *)(s_text + (pos & ~3u));
    const uint32_t sh = pos & 3u;
    uint32_t W[NWORDS + 1];
#pragma unroll
    for (int j = 0; j <= NWORDS; j++)
        W[j] = wp[j];
#pragma unroll
    for (int j = 0; j < NWORDS; j++)
        A[j] = __builtin_amdgcn_alignbyte(W[j + 1], W[j], sh);
}

// parse_line (crates/oneline-simd/src/line.rs:8-44) for the non-empty line
// text[s .. eff) held in LDS at byte index rs (48 bytes readable from there).
// A line starting with '|' has the leading run of '|'/blank bytes as its prefix
// (depth = the '|' count in it, content after it); a line starting with a blank
// skips the leading blank run at depth 0.  Fast tier: count leading "| | "
// words with one compare each, then decode an 8-byte tail with SWAR; general
// tier: SWAR over 48 bytes.  Returns false (the caller falls back to the byte
// loop) when the prefix is longer than that.
__device__ __forceinline__ bool decode_head(const uint8_t *s_text, uint32_t rs, uint32_t s, uint32_t eff,
                                            ol_node_dev &nd)
{
    const uint32_t len = eff - s;
    uint32_t depth = 0, r = 0;
    const uint32_t c0 = s_text[rs];
    if (c0 == '|' || c0 == ' ' || c0 == '\t') {
        const bool from_pipe = (c0 == '|');
        bool done = false;
        if (from_pipe) {
            uint32_t A[10];
            load_words<10>(s_text, rs, A);
            uint32_t eq = 0;
#pragma unroll
            for (int j = 0; j < 10; j++)
                eq |= (uint32_t)(A[j] == 0x207C207Cu) << j;   // "| | "
            uint32_t k = (uint32_t)__builtin_ctz(~eq);
            if (4 * k > len)
                k = len / 4;
            if (k <= 8) {
                uint32_t T[2];
                load_words<2>(s_text, rs + 4 * k, T);
                uint32_t pipes;
                const uint32_t rt = prefix_run<2>(T, len - 4 * k, true, pipes);
                if (rt < 8 || 4 * k + 8 >= len) {
                    r = 4 * k + rt;
                    depth = 2 * k + pipes;
                    done = true;
                }
            }
        }
        if (!done) {
            uint32_t A[12];
            load_words<12>(s_text, rs, A);
            uint32_t pipes;
            r = prefix_run<12>(A, len, from_pipe, pipes);
            if (r == 48 && len > 48)
                return false;
            depth = from_pipe ? pipes : 0;
        }
    }
    const uint32_t cs = s + r;
    nd.depth = depth;
    if (cs >= eff) {
        nd.span_start = eff;
        nd.kind = OLK_BLANK;
        nd.span_len = 0;
        return true;
    }
    nd.span_start = cs;
    const uint32_t d0 = s_text[rs + r];
    const uint32_t d1 = (cs + 1 < eff) ? (uint32_t)s_text[rs + r + 1] : 256u;
    const uint32_t d2 = (cs + 2 < eff) ? (uint32_t)s_text[rs + r + 2] : 256u;
    nd.kind = classify(d0, d1, d2);
    nd.span_len = eff - cs;
    return true;
}

// classify_line (crates/oneline-simd/src/line.rs:59-81) on the raw bytes at the
// content start, x = c0 | c1 << 8 | c2 << 16, as selects (no branches).  The
// bytes may run into the line's trailing '\r'/'\n' (classify_line would see
// "no byte" there): neither is '/', '.', ' ' or '\t', so the class is the
// same.  c1 is looked at only after c0 in {'/', '-'} and c2 only after
// "-.", so neither lies past the line's terminator when it matters.
__device__ __forceinline__ uint32_t classify_bytes(uint32_t x)
{
    const uint32_t c0 = x & 0xFFu, c01 = x & 0xFFFFu, c012 = x & 0xFFFFFFu;
    uint32_t k = OLK_TEXT;
    k = (c0 == '=') ? (uint32_t)OLK_HEADING : k;
    k = (c0 == ':') ? (uint32_t)OLK_CONNECT : k;
    k = (c0 == '@') ? (uint32_t)OLK_ELEMENT : k;
    k = (c01 == 0x2F2Fu) ? (uint32_t)OLK_COMMENT : k;                          // "//"
    k = ((c01 == 0x202Du) | (c01 == 0x092Du)) ? (uint32_t)OLK_LIST : k;         // "- ", "-\t"
    k = ((c012 == 0x202E2Du) | (c012 == 0x092E2Du)) ? (uint32_t)OLK_ORDERED : k; // "-. ", "-.\t"
    return k;
}

constexpr uint32_t RC_SLOW = 0x10000u;   // decode_fast: not a class | key << 8

// Class | key << 8 of the line staged in LDS at byte rs (its terminator in
// LDS), for the common shape: a prefix of "| " pairs (at most 19 of them), then
// content whose first byte is not '|', ' ' or '\t' -- or no content: the line
// is "\n" or "\r\n" after the prefix.  Those lines need no trailing trim:
// the first content byte is '\n' exactly when the content is empty, and
// classify_bytes reads the raw bytes.  Returns RC_SLOW for every other line
// (the general decoder handles it).  Straight-line code: the prefix words are
// compared all at once, the content bytes come from one more read.
__device__ __forceinline__ uint32_t decode_fast(const uint8_t *s_text, uint32_t rs)
{
    uint32_t A[10];
    load_words<10>(s_text, rs, A);
    uint32_t eq = 0;
#pragma unroll
    for (int j = 0; j < 10; j++)
        eq |= (uint32_t)(A[j] == 0x207C207Cu) << j;   // "| | "
    const uint32_t k = (uint32_t)__builtin_ctz(~eq);   // whole "| | " words (a '\n' ends the run)
    uint32_t T[2];
    load_words<2>(s_text, rs + 4 * (k < 10 ? k : 9u), T);
    const bool half = (T[0] & 0xFFFFu) == 0x207Cu;     // one more "| "
    const uint32_t x = half ? __builtin_amdgcn_alignbyte(T[1], T[0], 2) : T[0];
    const uint32_t x0 = x & 0xFFu, x01 = x & 0xFFFFu;
    const bool slow = (k >= 10) | (x0 == '|') | (x0 == ' ') | (x0 == '\t') | (x01 == 0x0D0Du);
    const bool blank = (x0 == '\n') | (x01 == 0x0A0Du);
    const uint32_t rc = blank ? (OLK_BLANK | (uint32_t)KEY_BLANK << 8)
                              : (classify_bytes(x) | (2 * k + (half ? 1u : 0u)) << 8);   // depth <= 19: the key is the depth
    return slow ? RC_SLOW : rc;
}

// One tile = THREADS lanes x 64*Q contiguous bytes.  The tile arrives in
// registers (coalesced 16-byte vectors), is staged whole in LDS behind a
// 256-byte halo, and the registers are then refilled with the workgroup's next
// tile (software pipelining).  Each lane owns 64*Q contiguous bytes: their
// '\n' masks are counted and scanned once per tile, the line ends are
// compacted into LDS in byte order, and lane i decodes line i into an LDS
// result slot (class | key << 8).  The tile's output position (its look-back
// over the per-tile status words) is resolved one iteration later, while the
// workgroup already works on its next tile, and only then are the staged
// results written (coalesced).  Tiles whose lines do not fit the LDS list, and
// calls that want the 16-byte FlatNodes, are resolved and written at once
// instead.  A line that starts before the staged window (longer than the
// halo) is written later by scan_fixup_kernel from global memory, so the
// decode touches LDS only.
template <int THREADS, int Q>
__global__ __launch_bounds__(THREADS) void scan_classify_kernel(ScanArgs a)
{
    constexpr int LB = 64 * Q;          // bytes per lane
    constexpr int TILE = THREADS * LB;
    constexpr int HALO = SCAN_HALO;
    constexpr int NW = THREADS / 64;
    constexpr int NV = LB / 16;         // 16-byte vectors per lane
    constexpr int LIST = TILE / 24 / 64 * 64;   // lines per tile decoded from the compacted list
    static_assert(NW >= 2, "needs two waves");
    static_assert(HALO % 16 == 0 && HALO / 16 <= THREADS, "halo");
    static_assert(TILE <= 65536, "tile positions are u16");

    // the staged window (halo + tile + 64 bytes of slack), then two 80-byte
    // line heads (s_head, below): one array, so every decode read is an LDS read
    constexpr int HEAD0 = HALO + TILE + 64;
    __shared__ __align__(16) uint8_t s_text[HEAD0 + 2 * 80];
    __shared__ __align__(16) uint16_t s_nm[TILE / 16];   // '\n' bits per 16-byte vector, byte order
    __shared__ uint16_t s_le[2][LIST];                   // tile position of each line end
    __shared__ uint16_t s_rc[2][LIST];                   // decoded class | key << 8
    // s_head[b]: the head of the tile's first line when it starts before the
    // staged window (a line longer than the halo): 80 bytes from its start
    // rounded down to 16, enough for decode_fast / decode_head
    uint8_t *const s_head0 = s_text + HEAD0;
    __shared__ uint32_t s_wcnt[NW];
    __shared__ int32_t s_wlast[NW];
    __shared__ uint64_t s_prefix, s_pprefix;
    __shared__ uint32_t s_first[2];
    constexpr int MAXD = 64;                              // document starts staged per tile
    __shared__ uint32_t s_ds[2][MAXD];                   // starts of the documents starting in the tile
    __shared__ uint32_t s_dj[2][2];                      // their index range [j0, j1)

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t G = gridDim.x - 1;   // workgroups that take tiles (static order)
    const uint32_t T = a.num_tiles;
    const uint32_t last = a.n - 1;   // always a terminator, '\n' or not

    uint4 v[NV];
    uint4 hv;
    auto load_tile = [&](uint32_t tt) {
        const uint32_t b = tt * TILE;
        const uint4 *src = (const uint4 *)(a.text + (size_t)b);
#pragma unroll
        for (int k = 0; k < NV; k++)
            v[k] = src[k * THREADS + tid];   // plain loads: non-temporal ones measured slower here (the halo and long-line reads miss)
        hv = make_uint4(0, 0, 0, 0);
        if (tt > 0 && tid < HALO / 16)
            hv = ((const uint4 *)(a.text + b - HALO))[tid];
    };
    // wave 0: the tile's exclusive prefix.  Static order: lane 0 waits for the
    // scanner's inclusive prefix of the tile (bounded wait: see LB_SPIN_LIMIT).
    // Ticket order: a decoupled look-back over the tiles' status words, and
    // the tile's own inclusive status after it.
    const bool ticket = a.tile_ctr != nullptr;
    // static order: the prefix once the scanner has written the tile's prefix
    // word (`w`: a first read of it, 0 if none).  s_ahead records whether the
    // first read found it missing: the workgroup is then ahead of the
    // scanner's front, i.e. of the slowest workgroups.
    __shared__ uint32_t s_ahead;
    auto finish_prefix = [&](uint32_t tt, uint32_t tot, uint64_t w) -> uint64_t {
        if (lane == 0) {
            uint32_t spins = 0;
            while ((w >> LB_SHIFT) != LB_FLAG_INCL && spins < LB_SPIN_LIMIT) {
                if (spins++)
                    __builtin_amdgcn_s_sleep(8);
                w = ld_coherent_u64(a.status + (size_t)tt * SCAN_STATUS_WORDS + SCAN_PREFIX_WORD);
            }
            s_ahead = spins > 1 ? 1u : 0u;
            if ((w >> LB_SHIFT) != LB_FLAG_INCL) {
                atomicOr(a.lb_err, OLE_PREFIX_WAIT);
                w = (LB_FLAG_INCL << LB_SHIFT) | tot;
            }
            // the scanner has consumed the count word and this tile the
            // prefix word: zero both for the next launch
            uint64_t *const st = a.status + (size_t)tt * SCAN_STATUS_WORDS;
            st_coherent_u64(st, 0ull);
            st_coherent_u64(st + SCAN_PREFIX_WORD, 0ull);
        }
        return (w & LB_MASK) - tot;
    };
    auto resolve = [&](uint32_t tt, uint32_t tot) -> uint64_t {
        if (ticket) {
            uint64_t prefix = 0;
            if (tt > 0) {
                prefix = lookback_sum<1>(a.status, tt, a.lb_err);
                if (lane == 0)
                    st_relaxed_u64(a.status + (size_t)tt * LB_STRIDE, (LB_FLAG_INCL << LB_SHIFT) | (prefix + tot));
            }
            return prefix;
        }
        return finish_prefix(tt, tot, 0);
    };
    // offsets[r] for the line starting at s (relative to its document in a
    // batch; the first line of a document also records r in doc_line_start)
    auto put_offset = [&](int pb, uint64_t r, uint32_t s) {
        if (a.ndocs == 0) {
            a.offsets[r] = s;
            return;
        }
        // the document holding s: the last one starting at or before s
        const uint32_t j0 = s_dj[pb][0], j1 = s_dj[pb][1];
        uint32_t lo = j0, hi = j1;   // first j in [j0, j1) with start > s
        if (j1 - j0 <= (uint32_t)MAXD) {
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_ds[pb][mid - j0] <= s)
                    lo = mid + 1;
                else
                    hi = mid;
            }
        } else {
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a.doc_starts[mid] <= s)
                    lo = mid + 1;
                else
                    hi = mid;
            }
        }
        const uint32_t d = lo - 1;   // a line starts in or after its document
        const uint32_t ds = a.doc_starts[d];
        a.offsets[r] = s - ds;
        if (s == ds)
            a.doc_line_start[d] = (uint32_t)r;
    };
    // writes a staged tile: lines i < tot of buffer pb, first line at pfirst.
    // Line indices fit 32 bits (a text has fewer than 2^32 lines).
    uint32_t *const offsets = a.offsets;
    uint8_t *const classes = a.classes;
    uint8_t *const keys = a.keys;
    auto write_staged = [&](int pb, uint32_t pbase, uint32_t pfirst, uint32_t tot, uint64_t running, bool is_last) {
        if (tid == 0 && is_last)
            a.meta[0] = running + tot;
        const uint32_t nw = (running + tot <= a.cap) ? tot : (running < a.cap ? (uint32_t)(a.cap - running) : 0u);
        const uint32_t r0 = (uint32_t)running;
        if (a.ndocs == 0) {
            for (uint32_t i = tid; i < nw; i += THREADS) {
                const uint32_t r = r0 + i;
                const uint32_t ls = i ? pbase + s_le[pb][i - 1] + 1u : pfirst;
                const uint32_t rc = s_rc[pb][i];
                offsets[r] = ls;
                classes[r] = (uint8_t)(rc & 0xFFu);
                keys[r] = (uint8_t)(rc >> 8);
            }
            return;
        }
        for (uint32_t i = tid; i < nw; i += THREADS) {
            const uint32_t r = r0 + i;
            const uint32_t ls = i ? pbase + s_le[pb][i - 1] + 1u : pfirst;
            put_offset(pb, r, ls);
            const uint32_t rc = s_rc[pb][i];
            classes[r] = (uint8_t)(rc & 0xFFu);
            keys[r] = (uint8_t)(rc >> 8);
        }
    };

    // the workgroup's previous tile, decoded and waiting for its prefix
    bool pending = false;
    uint32_t p_t = 0, p_total = 0;

    // Tile order: static (t = blockIdx + k * workers: needs every workgroup
    // resident, the cooperative launch; the last workgroup is the prefix
    // scanner), or claimed from a ticket (tile_ctr: a tile's look-back then
    // only waits on tiles held by running workgroups, whatever shares the
    // device; every workgroup takes tiles).  A workgroup knows its next tile
    // one iteration ahead, for the prefetch.
    __shared__ uint32_t s_claim[2];
    uint32_t t, tn;
    if (ticket) {
        if (tid == 0) {
            s_claim[0] = atomicAdd(a.tile_ctr, 1u);
            s_claim[1] = atomicAdd(a.tile_ctr, 1u);
        }
        __syncthreads();
        t = s_claim[0];
        tn = s_claim[1];
    } else {
        if (blockIdx.x == G) {
            if (wid == 0) {
                // every tile's prefix passes through this wave: it never
                // yields its issue slots to the tile waves (see below)
                __builtin_amdgcn_s_setprio(3);
                scan_prefixes<SCAN_PREFIX_WINDOWS>(a.status, T, a.lb_err);
            }
            return;
        }
        t = blockIdx.x;
        tn = t + G;
    }
    int b = 0;
    if (t < T)
        load_tile(t);
    while (t < T) {
        const uint32_t base = t * TILE;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            s_nm[k * THREADS + tid] = (uint16_t)nl_mask16(v[k]);
            ((uint4 *)(s_text + HALO))[k * THREADS + tid] = v[k];
        }
        if (tid < HALO / 16)
            ((uint4 *)s_text)[tid] = hv;
        if (wid != 0 && tn < T)
            load_tile(tn);
        if (last >= base && last - base < (uint32_t)TILE) {   // the tile holding the text's last byte
            __syncthreads();
            if (tid == 0)
                s_nm[(last - base) >> 4] |= (uint16_t)(1u << ((last - base) & 15));
        }
        __syncthreads();

        // this lane's bytes [tid*LB, tid*LB + LB): masks, count, last terminator
        uint64_t mk[Q];
        uint32_t cnt = 0;
        int32_t il = -1;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            mk[q] = ((const uint64_t *)s_nm)[tid * Q + q];
            cnt += (uint32_t)__popcll(mk[q]);
            if (mk[q])
                il = (int32_t)(tid * LB + 64 * q + 63 - __clzll(mk[q]));
        }
        const uint32_t ic = wave_incl_add(cnt);
        const int32_t ilm = wave_incl_max(il);
        const int32_t prev_l = (int32_t)wave_shr1((uint32_t)ilm, ~0u);
        if (lane == 63) {
            s_wcnt[wid] = ic;
            s_wlast[wid] = ilm;
        }
        __syncthreads();
        uint32_t total = 0, wb = 0;
        int32_t wl = -1;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const uint32_t c = s_wcnt[w];
            total += c;
            if (w < wid) {
                wb += c;
                wl = wl > s_wlast[w] ? wl : s_wlast[w];
            }
        }
        const uint32_t excl_c = wb + ic - cnt;
        const int32_t excl_l = wl > prev_l ? wl : prev_l;

        const bool listed = total <= (uint32_t)LIST;
        const bool staged = listed && !a.nodes;   // written one iteration later

        // Publish the tile's count at once, then compact the line ends.
        if (wid == 0 && lane == 0) {
            if (ticket)
                st_relaxed_u64(a.status + (size_t)t * LB_STRIDE, ((t == 0 ? LB_FLAG_INCL : LB_FLAG_AGG) << LB_SHIFT) | total);
            else
                st_coherent_u64(a.status + (size_t)t * SCAN_STATUS_WORDS, (LB_FLAG_AGG << LB_SHIFT) | total);
        }
        if (listed) {
            uint32_t at = excl_c;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                uint64_t m = mk[q];
                while (m) {
                    const int bit = __builtin_ctzll(m);
                    m &= m - 1;
                    s_le[b][at++] = (uint16_t)(tid * LB + 64 * q + bit);
                }
            }
        }
        __syncthreads();

        const uint32_t lds_lo = (base >= (uint32_t)HALO) ? base - HALO : 0u;
        uint32_t first_w1 = 0;   // wave 1: start of the tile's first line
        if (wid == 0) {
            // the previous tile's prefix (its predecessors have long published
            // their counts), and this tile's when it is written at once
            if (pending) {
                const uint64_t pp = resolve(p_t, p_total);
                if (lane == 0)
                    s_pprefix = pp;
            }
            if (!staged) {
                const uint64_t pr = resolve(t, total);
                if (lane == 0)
                    s_prefix = pr;
            }
            if (tn < T)
                load_tile(tn);
        } else if (wid == 1) {
            // Start of the first line that ends in this tile: after the last '\n'
            // before the tile (in the halo, else further back).
            if (t > 0 && total > 0) {
                const uint32_t w4 = ((const uint32_t *)s_text)[lane];
                const uint32_t m = nl_nibble(w4);
                const int32_t best = m ? (int32_t)(4 * lane + 31 - __clz(m)) : -1;
                const uint64_t found = __ballot(best >= 0);
                if (found) {
                    const int srcl = 63 - (int)__clzll(found);
                    const int32_t p = __shfl(best, srcl, 64);
                    first_w1 = base - HALO + (uint32_t)p + 1;
                } else {
                    first_w1 = find_line_start_global(a.text, base - HALO);
                }
                if (first_w1 < lds_lo) {   // the line's head from global memory (the text is padded)
                    if (lane < 5)
                        ((uint4 *)(s_head0 + 80 * b))[lane] = *(const uint4 *)(a.text + (first_w1 & ~15u) + 16u * lane);
                    wave_lds_sync();   // read back by this wave's lane 0 (line 0's decode)
                }
            }
            if (lane == 0)
                s_first[b] = first_w1;
            if (a.ndocs) {
                const uint32_t j0 = a.tile_doc[t], j1 = a.tile_doc[t + 1];
                if (lane < 2)
                    s_dj[b][lane] = lane ? j1 : j0;
                for (uint32_t j = j0 + lane; j < j1 && j < j0 + MAXD; j += 64)
                    s_ds[b][j - j0] = a.doc_starts[j];
            }
        }

        auto lds_byte = [&](uint32_t pos) -> uint32_t { return s_text[pos + (uint32_t)HALO - base]; };
        // LDS index in s_text of the line's head bytes: the staged window, or
        // s_head[b] for a first line starting below it
        auto head_of = [&](uint32_t s) -> uint32_t {
            return s >= lds_lo ? s + (uint32_t)HALO - base : (uint32_t)HEAD0 + 80u * b + (s & 15u);
        };
        // parse_line of text[s ..= e] (e in the tile).  Inlined: as a call it
        // spilled the registers live across it to scratch on every tile.
        // Only the tile's first line can start below the window; its bytes
        // there come from global memory (separate loops: a byte source that
        // picked LDS or global per access would make every read a flat one).
        auto decode = [&](uint32_t s, uint32_t e) __attribute__((always_inline)) -> ol_node_dev {
            const uint32_t lo = s > lds_lo ? s : lds_lo;
            uint32_t eff = e + 1;
            while (eff > lo) {
                const uint32_t c = lds_byte(eff - 1);
                if (c != '\n' && c != '\r')
                    break;
                eff--;
            }
            if (eff == lo) {
                while (eff > s) {
                    const uint32_t c = a.text[eff - 1];
                    if (c != '\n' && c != '\r')
                        break;
                    eff--;
                }
            }
            ol_node_dev nd;
            if (eff == s) {
                nd.depth = 0; nd.kind = OLK_BLANK; nd.span_start = s; nd.span_len = 0;
            } else if (!decode_head(s_text, head_of(s), s, eff, nd)) {
                nd = s >= lds_lo ? parse_span(lds_byte, s, e) : parse_span(GlobalSrc{a.text}, s, e);
            }
            return nd;
        };
        // Decode line (s, e) and write it at output index r.
        auto emit = [&](uint64_t r, uint32_t s, uint32_t e) {
            if (r >= a.cap)
                return;
            put_offset(b, r, s);
            const ol_node_dev nd = decode(s, e);
            a.classes[r] = (uint8_t)nd.kind;
            a.keys[r] = key_of(nd);
            if (a.nodes)
                a.nodes[r] = nd;
        };
        if (staged) {
            // Decode into LDS, by waves 1.. (wave 0 resolves the prefix): line i
            // goes to thread 64 + i mod (THREADS - 64), so line 0 is wave 1's,
            // which has just found its start.
            for (uint32_t i = (uint32_t)tid - 64u; tid >= 64 && i < total; i += THREADS - 64) {
                const uint32_t e = base + s_le[b][i];
                const uint32_t s = i ? base + s_le[b][i - 1] + 1u : first_w1;
                // (the text's last byte ends a line without being a '\n')
                uint32_t rc = (e == last) ? RC_SLOW : decode_fast(s_text, head_of(s));
                if (rc == RC_SLOW) {
                    const ol_node_dev nd = decode(s, e);
                    rc = nd.kind | (uint32_t)key_of(nd) << 8;
                }
                s_rc[b][i] = (uint16_t)rc;
            }
        }
        __syncthreads();

        // Static order couples the workgroups through the prefixes: a tile's
        // prefix waits for the counts of every lower tile.  The oldest
        // workgroups on a CU win instruction issue by age and run up to a
        // third faster than the youngest, then wait for them.  A workgroup
        // whose last prefix was not ready at the first read (ahead of the
        // front) lowers its waves' priority, one whose prefix was (at or
        // behind the front) raises it, which evens out the workgroups' pace:
        // 500 -> 484 us per 2 GiB.
        if (pending && !ticket) {
            if (s_ahead)
                __builtin_amdgcn_s_setprio(0);
            else
                __builtin_amdgcn_s_setprio(2);
        }
        if (pending)
            write_staged(b ^ 1, p_t * TILE, s_first[b ^ 1], p_total, s_pprefix, p_t == T - 1);
        if (!staged) {
            const uint64_t running = s_prefix;   // index of the tile's first line
            const uint32_t first = s_first[b];   // start of the tile's first line
            if (tid == 0 && t == T - 1)
                a.meta[0] = running + total;
            if (listed) {
                for (uint32_t i = tid; i < total; i += THREADS) {
                    const uint32_t e = base + s_le[b][i];
                    const uint32_t s = i ? base + s_le[b][i - 1] + 1u : first;
                    emit(running + i, s, e);
                }
            } else if (cnt) {
                // dense tile: each lane decodes the lines ending in its bytes
                uint64_t r = running + excl_c;
                uint32_t sp = (excl_l >= 0) ? base + (uint32_t)excl_l + 1 : first;
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    uint64_t m = mk[q];
                    while (m) {
                        const int bit = __builtin_ctzll(m);
                        m &= m - 1;
                        const uint32_t e = base + tid * LB + 64 * q + (uint32_t)bit;
                        emit(r, sp, e);
                        sp = e + 1;
                        r++;
                    }
                }
            }
        }
        pending = staged;
        p_t = t;
        p_total = total;
        if (ticket && tid == 0)
            s_claim[b] = atomicAdd(a.tile_ctr, 1u);
        __syncthreads();   // the next tile overwrites LDS
        t = tn;
        tn = ticket ? s_claim[b] : tn + G;
        b ^= 1;
    }
    if (pending) {
        if (wid == 0) {
            const uint64_t pp = resolve(p_t, p_total);
            if (lane == 0)
                s_pprefix = pp;
        }
        __syncthreads();
        write_staged(b ^ 1, p_t * TILE, s_first[b ^ 1], p_total, s_pprefix, p_t == T - 1);
    }
}

template <int Q>
static const void *scan_fn()
{
    return (const void *)scan_classify_kernel<SCAN_THREADS, Q>;
}

static const void *scan_fn_bpt(uint32_t bpt)
{
    switch (bpt) {
    case 64: return scan_fn<1>();
    case 128: return scan_fn<2>();
    case 256: return scan_fn<4>();
    default: return nullptr;
    }
}

// Co-resident workgroups of the scan kernel for a given bytes-per-lane.
uint32_t ol_scan_max_grid(uint32_t bpt)
{
    const void *f = scan_fn_bpt(bpt);
    return f ? ol_coresident(f, SCAN_THREADS, 0) : 0u;
}

hipError_t ol_launch_scan(const ScanArgs &a, uint32_t bpt, uint32_t grid, bool cooperative, hipStream_t s)
{
    const void *f = scan_fn_bpt(bpt);
    if (!f || grid == 0)
        return hipErrorInvalidValue;
    ScanArgs args = a;
    void *params[] = {&args};
    if (cooperative)
        return hipLaunchCooperativeKernel(f, dim3(grid), dim3(SCAN_THREADS), params, 0, s);
    // inside a captured graph: an ordinary launch of the same co-resident grid
    return hipLaunchKernel(f, dim3(grid), dim3(SCAN_THREADS), params, 0, s);
}

// ---- parse with caller-given line offsets ------------------------------------
// GpuParser::parse (crates/oneline-gpu/src/parse.rs:104-124 + shader.wgsl:34-134):
// line i spans offsets[i] .. offsets[i+1] (or the text end for the last line).

__global__ void parse_given_kernel(const uint8_t *text, uint32_t n, const uint32_t *offsets,
                                   uint64_t lines, ol_node_dev *nodes)
{
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= lines)
        return;
    uint32_t s = offsets[i];
    uint32_t e = (i + 1 < lines) ? offsets[i + 1] : n;   // exclusive
    ol_node_dev nd;
    if (e <= s || s >= n) {
        nd.depth = 0; nd.kind = OLK_BLANK; nd.span_start = s; nd.span_len = 0;
    } else {
        if (e > n)
            e = n;
        nd = parse_span(GlobalSrc{text}, s, e - 1);
    }
    nodes[i] = nd;
}

hipError_t ol_launch_parse_given(const uint8_t *text, uint32_t n, const uint32_t *offsets,
                                 uint64_t lines, ol_node_dev *nodes, hipStream_t s)
{
    if (lines == 0)
        return hipSuccess;
    uint64_t blocks = (lines + 255) / 256;
    hipLaunchKernelGGL(parse_given_kernel, dim3((unsigned)blocks), dim3(256), 0, s, text, n,
                       offsets, lines, nodes);
    return hipGetLastError();
}

// ---- kind / depth queries over a resident document ---------------------------
// GpuParser::query_document / query_subtree (crates/oneline-gpu/src/document.rs:94,
// crates/oneline-gpu/src/tree.rs:295) with the predicate of filter_nodes
// (crates/oneline-gpu/src/shader.wgsl:152-179): kind 0 matches every kind,
// depth 0xFFFFFFFF every depth; otherwise node.kind == kind and node.depth ==
// depth.  The matching line indices of [r0, r1) are written in ascending
// order: a count pass (per block of QUERY_BLOCK lines), an exclusive scan of
// the block counts, and a write pass that recomputes the predicate.  A line's
// depth is its tree key for content lines; blank lines (key 255) and lines at
// depth >= 254 are parsed again from the text when a depth is asked for (or
// read from the FlatNodes when the document keeps them).
constexpr uint32_t QUERY_THREADS = 256, QUERY_PER_THREAD = 16;
constexpr uint32_t QUERY_BLOCK = QUERY_THREADS * QUERY_PER_THREAD;

// Bit j of the result: byte j of `w` equals the byte `v` (SWAR zero-byte test
// of w ^ v, exact per byte: the low seven bits are added without carries into
// the next byte).
__device__ __forceinline__ uint32_t query_byte_eq4(uint32_t w, uint32_t v)
{
    const uint32_t x = w ^ (v * 0x01010101u);
    const uint32_t nz = (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
    const uint32_t e = (~nz & 0x80808080u) >> 7;   // bits 0, 8, 16, 24
    return ((e * 0x204081u) >> 21) & 0xFu;         // gathered into bits 0..3
}

// Bit j: byte j of `w` is >= KEY_DEEP (254 or 255: depth not in the key).
__device__ __forceinline__ uint32_t query_byte_deep4(uint32_t w)
{
    return query_byte_eq4(w | 0x01010101u, 0xFFu);   // 254 and 255 are the bytes with bits 1..7 set
}

__device__ __forceinline__ uint32_t query_eq16(uint4 v, uint32_t b)
{
    return query_byte_eq4(v.x, b) | query_byte_eq4(v.y, b) << 4 | query_byte_eq4(v.z, b) << 8 |
           query_byte_eq4(v.w, b) << 12;
}

__device__ __forceinline__ uint32_t query_deep16(uint4 v)
{
    return query_byte_deep4(v.x) | query_byte_deep4(v.y) << 4 | query_byte_deep4(v.z) << 8 |
           query_byte_deep4(v.w) << 12;
}

// First line of the thread's QUERY_PER_THREAD lines: blocks start on the
// 16-line boundary at or below r0, so a thread's lines are one aligned
// 16-byte word of the class and key arrays.
__device__ __forceinline__ uint32_t query_line0(const QueryArgs &a, uint32_t blk)
{
    return (a.r0 & ~(QUERY_PER_THREAD - 1)) + blk * QUERY_BLOCK + threadIdx.x * QUERY_PER_THREAD;
}

// Match bits of the lines i0 .. i0 + 15 (bit k = line i0 + k).  The classes
// and keys arrive in one 16-byte load each (both arrays are padded to
// LINE_ALIGN lines, and i0 < r1 <= lines); a byte load per line behind the
// range test compiled to sixteen round trips in a row.  Only lines whose key
// does not hold the depth (blank, depth >= 254) read more: the FlatNode when
// the document keeps them, else the line parsed again.
__device__ __forceinline__ uint32_t query_bits(const QueryArgs &a, uint32_t i0)
{
    if (i0 >= a.r1)
        return 0;
    uint32_t bits = 0xFFFFu;
    if (i0 < a.r0)
        bits &= 0xFFFFu << (a.r0 - i0);
    if (a.r1 - i0 < QUERY_PER_THREAD)
        bits &= (1u << (a.r1 - i0)) - 1u;
    if (a.kind != 0)
        bits &= a.kind < 256 ? query_eq16(*(const uint4 *)(a.classes + i0), a.kind) : 0u;
    if (a.depth == 0xFFFFFFFFu || bits == 0)
        return bits;
    const uint4 key = *(const uint4 *)(a.keys + i0);
    uint32_t eq = a.depth < KEY_DEEP ? query_eq16(key, a.depth) : 0u;
    uint32_t deep = bits & query_deep16(key);
    while (deep) {
        const uint32_t k = (uint32_t)__builtin_ctz(deep);
        deep &= deep - 1;
        const uint32_t i = i0 + k;
        uint32_t depth;
        if (a.nodes) {
            depth = a.nodes[i].depth;
        } else {
            const uint32_t s = a.offsets[i];
            const uint32_t e = (i + 1 < a.lines) ? a.offsets[i + 1] - 1 : a.n - 1;
            depth = parse_span(GlobalSrc{a.text}, s, e).depth;
        }
        eq |= (depth == a.depth) ? 1u << k : 0u;
    }
    return bits & eq;
}

__device__ __forceinline__ uint32_t query_block_excl(uint32_t v, uint32_t &total)
{
    __shared__ uint32_t s_w[QUERY_THREADS / 64];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_add(v);
    if (lane == 63)
        s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (uint32_t k = 0; k < QUERY_THREADS / 64; k++) {
        before += k < w ? s_w[k] : 0u;
        total += s_w[k];
    }
    return before + inc - v;
}

__global__ __launch_bounds__(QUERY_THREADS) void query_count_kernel(QueryArgs a)
{
    const uint32_t bits = query_bits(a, query_line0(a, blockIdx.x));
    uint32_t total;
    (void)query_block_excl((uint32_t)__builtin_popcount(bits), total);
    if (threadIdx.x == 0)
        a.block_count[blockIdx.x] = total;
}

__global__ __launch_bounds__(QUERY_THREADS) void query_write_kernel(QueryArgs a)
{
    const uint32_t i0 = query_line0(a, blockIdx.x);
    const uint32_t bits = query_bits(a, i0);
    uint32_t total;
    const uint32_t ex = query_block_excl((uint32_t)__builtin_popcount(bits), total);
    uint64_t r = a.block_off[blockIdx.x] + ex;
    uint32_t m = bits;
    while (m) {
        const uint32_t k = (uint32_t)__builtin_ctz(m);
        m &= m - 1;
        if (r < a.cap)
            a.out[r] = i0 + k;
        r++;
    }
}

// Exclusive scan of the block counts (one workgroup, chunks of 1024).
__global__ __launch_bounds__(1024) void query_scan_kernel(QueryArgs a, uint32_t blocks, uint64_t *total)
{
    __shared__ uint32_t s_w[16];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t carry = 0;
    for (uint32_t c0 = 0; c0 < blocks; c0 += 1024) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t v = i < blocks ? a.block_count[i] : 0u;
        const uint32_t inc = wave_incl_add(v);
        if (lane == 63)
            s_w[w] = inc;
        __syncthreads();
        uint32_t before = 0, tot = 0;
        for (uint32_t k = 0; k < 16; k++) {
            before += k < w ? s_w[k] : 0u;
            tot += s_w[k];
        }
        if (i < blocks)
            a.block_off[i] = carry + before + inc - v;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *total = carry;
}

// Count, scan and (when `a.out`) write; *total (device) gets the match count.
hipError_t ol_launch_query(const QueryArgs &a, uint64_t *total, hipStream_t s)
{
    if (a.r1 <= a.r0)
        return hipSuccess;
    const uint32_t blocks = ol_query_blocks(a.r0, a.r1);
    hipLaunchKernelGGL(query_count_kernel, dim3(blocks), dim3(QUERY_THREADS), 0, s, a);
    hipLaunchKernelGGL(query_scan_kernel, dim3(1), dim3(1024), 0, s, a, blocks, total);
    if (a.out && a.cap)
        hipLaunchKernelGGL(query_write_kernel, dim3(blocks), dim3(QUERY_THREADS), 0, s, a);
    return hipGetLastError();
}

uint32_t ol_query_blocks(uint32_t r0, uint32_t r1)
{
    const uint32_t lines = r1 - (r0 & ~(QUERY_PER_THREAD - 1));
    return (uint32_t)(((uint64_t)lines + QUERY_BLOCK - 1) / QUERY_BLOCK);
}

// ---- AST bridge: the structural pass of to_tomet_ast_with_links -------------
// crates/oneline-tomet/src/lib.rs:20-90 walks the lines in order
// (literal_pipe_lines keeps the last bare blank line seen and each line's
// literal flag depends on its parent's) and stitches blocks from the last
// line back.  Both are sequential there; here every line is independent
// given one number, b(i) = the last bare blank line before i (a prefix
// maximum over the lines):
//   - a line of depth > 0 is literal when, walking up its parent chain, a
//     parent is NONE, or lies before b(i) (a bare blank line between them:
//     the cut), or is not an element, before a depth-0 parent is reached.
//     The chain's lines are content lines, so no bare blank lies at them,
//     and b(i) is also the last bare blank before every chain line above a
//     cut, which is why one b serves the whole chain;
//   - the line's parent after the cut is NONE when b(i) > parent(i);
//   - a line's block goes to the nearest ancestor along the cut parents that
//     has a block (lines without one hand their children on, lib.rs:66-86);
//   - has_pipe_child(p) (lib.rs:33-41): the non-cut children of p are
//     literal exactly when p is not an element or is literal, and if any
//     child is not cut then the first is not (a cut of the first child cuts
//     all later ones): so p is a non-literal element whose next content line
//     is deeper, with no bare blank line before it.
// Three launches over blocks of BRIDGE_BLOCK lines: marks (block and
// bare-blank bits, each block's last bare blank; a line per lane), a prefix
// maximum of those over the blocks, and the per-line chains (16 consecutive
// lines per thread).
constexpr uint32_t BRIDGE_THREADS = 256, BRIDGE_PER_THREAD = 16;
constexpr uint32_t BRIDGE_BLOCK = BRIDGE_THREADS * BRIDGE_PER_THREAD;
enum : uint32_t { BR_LITERAL = 1u, BR_BLOCK = 2u, BR_CUT = 4u, BR_PIPE = 8u, BR_BARE = 16u };

// char::is_whitespace (Unicode White_Space) at text[p .. e): its length in
// bytes, 0 when it is not whitespace (or not valid UTF-8 there).
__device__ __forceinline__ uint32_t bridge_space(const uint8_t *t, uint32_t p, uint32_t e)
{
    const uint32_t c = t[p];
    if ((c >= 0x09u && c <= 0x0Du) || c == 0x20u)
        return 1;
    if (c < 0xC2u || c > 0xE3u)
        return 0;
    const uint32_t c1 = p + 1 < e ? t[p + 1] : 0u;
    if (c == 0xC2u)
        return (c1 == 0x85u || c1 == 0xA0u) ? 2u : 0u;   // U+0085, U+00A0
    const uint32_t c2 = p + 2 < e ? t[p + 2] : 0u;
    if (c == 0xE1u)
        return (c1 == 0x9Au && c2 == 0x80u) ? 3u : 0u;   // U+1680
    if (c == 0xE2u) {
        if (c1 == 0x80u && ((c2 >= 0x80u && c2 <= 0x8Au) || c2 == 0xA8u || c2 == 0xA9u || c2 == 0xAFu))
            return 3;                                    // U+2000..200A, U+2028, U+2029, U+202F
        return (c1 == 0x81u && c2 == 0x9Fu) ? 3u : 0u;   // U+205F
    }
    if (c == 0xE3u)
        return (c1 == 0x80u && c2 == 0x80u) ? 3u : 0u;   // U+3000
    return 0;
}

// Whether Text line i's content has a character that is not whitespace.
// Other kinds need no look: a Heading, Connect, Element or list line's
// content starts with `=`, `:`, `@` or `-`.  The content's first byte is the
// line's first byte that is not part of the prefix (`|` and blanks when the
// line starts with `|`, blanks otherwise; parse_line,
// crates/oneline-simd/src/line.rs:8-44), and a Text line's content is not
// empty, so that byte is inside the line.  It is found in 32-byte windows of
// independent loads (one for a prefix under 29 bytes, about 92% of the lines
// of the benchmark document); when it is plainly not whitespace (the usual
// case) that decides, otherwise the span is taken from the FlatNodes or the
// line parsed again and walked character by character.  Taking that path for
// every prefix longer than one window (a lane of most 64-line rounds) made
// the mark pass 1.36 ms at 2 GiB.
__device__ __forceinline__ bool bridge_text_block_slow(const BridgeArgs &a, uint32_t i)
{
    ol_node_dev nd;
    if (a.nodes) {
        nd = a.nodes[i];
    } else {
        const uint32_t s = a.offsets[i];
        const uint32_t e = (i + 1 < a.lines) ? a.offsets[i + 1] - 1 : a.n - 1;
        nd = parse_span(GlobalSrc{a.text}, s, e);
    }
    const uint32_t e = nd.span_start + nd.span_len;
    uint32_t p = nd.span_start, w = 1;
    while (p < e && (w = bridge_space(a.text, p, e)) != 0)
        p += w;
    return p < e;
}

__device__ __forceinline__ bool bridge_text_block(const BridgeArgs &a, uint32_t i)
{
    const uint32_t s = a.offsets[i];
    uint32_t s0 = s & ~3u, sh = s - s0;
    uint32_t piped = 2;   // whether the line starts with `|` (2: not read yet)
    // A prefix longer than the window continues in the next one (a prefix of
    // depth 16 or more fills 32 bytes): up to four windows, then the slow
    // path.  Every window read ends within 32 bytes of a byte of the line (the
    // text is padded past its end).
#pragma unroll 1
    for (uint32_t round = 0; round < 4; round++) {
        const uint32_t *wp = (const uint32_t *)(a.text + s0);
        uint32_t w[8];
#pragma unroll
        for (int j = 0; j < 8; j++)
            w[j] = wp[j];
        uint32_t pipes = 0, blanks = 0;   // bit b: byte s0 + b
#pragma unroll
        for (int j = 0; j < 8; j++) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const uint32_t c = (w[j] >> (8 * q)) & 0xFFu;
                pipes |= (c == '|' ? 1u : 0u) << (4 * j + q);
                blanks |= (c == ' ' || c == '\t' ? 1u : 0u) << (4 * j + q);
            }
        }
        if (piped == 2)
            piped = (pipes >> sh) & 1u;
        const uint32_t pre = piped ? (pipes | blanks) : blanks;
        const uint32_t rest = ~pre >> sh;   // bit k: byte s0 + sh + k is not prefix
        if (rest) {
            const uint32_t p = sh + (uint32_t)__builtin_ctz(rest);
            uint32_t word = w[0];
#pragma unroll
            for (int j = 1; j < 8; j++)
                word = (p >> 2) == (uint32_t)j ? w[j] : word;
            const uint32_t c0 = (word >> (8 * (p & 3))) & 0xFFu;
            // not the first byte of any Unicode White_Space character
            const bool plain = (c0 > 0x20u && c0 != 0xC2u && c0 != 0xE1u && c0 != 0xE2u && c0 != 0xE3u) ||
                               c0 < 0x09u || (c0 > 0x0Du && c0 < 0x20u);
            return plain ? true : bridge_text_block_slow(a, i);
        }
        s0 += 32;
        sh = 0;
        if (s0 >= a.n)
            break;
    }
    return bridge_text_block_slow(a, i);
}

// Lane per line (line blk * BRIDGE_BLOCK + r * 256 + tid in round r), so the
// class, offset and text loads of a wave are coalesced: a line's flags are
// its block bit and its bare-blank bit.
__global__ __launch_bounds__(BRIDGE_THREADS) void bridge_mark_kernel(BridgeArgs a)
{
    __shared__ uint32_t s_w[BRIDGE_THREADS / 64];
    uint32_t last = 0;   // the thread's last bare blank line + 1
    const uint32_t base = blockIdx.x * BRIDGE_BLOCK + threadIdx.x;
#pragma unroll 4
    for (uint32_t r = 0; r < BRIDGE_PER_THREAD; r++) {
        const uint32_t i = base + r * BRIDGE_THREADS;
        if (i >= a.lines)
            break;
        const uint32_t kind = a.classes[i];
        uint32_t fl = 0;
        if (kind == OLK_BLANK) {
            // depth 0 exactly when the line does not start with `|`
            if (a.text[a.offsets[i]] != '|') {
                fl = BR_BARE;
                last = i + 1;
            }
        } else if (kind == OLK_TEXT) {
            fl = bridge_text_block(a, i) ? BR_BLOCK : 0u;
        } else if (kind != OLK_COMMENT) {
            fl = BR_BLOCK;
        }
        a.flags[i] = (uint8_t)fl;
    }
    // the block's last bare blank line (+ 1, 0: none)
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t m = (uint32_t)wave_incl_max((int32_t)last);
    if (lane == 63)
        s_w[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t k = 0; k < BRIDGE_THREADS / 64; k++)
            t = max(t, s_w[k]);
        a.blk_last[blockIdx.x] = t;
    }
}

// Exclusive prefix maximum of blk_last into blk_in (one workgroup, chunks of 1024).
__global__ __launch_bounds__(1024) void bridge_scan_kernel(BridgeArgs a, uint32_t blocks)
{
    __shared__ uint32_t s_w[16];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < blocks; c0 += 1024) {
        const uint32_t i = c0 + threadIdx.x;
        const uint32_t v = i < blocks ? a.blk_last[i] : 0u;
        const uint32_t inc = (uint32_t)wave_incl_max((int32_t)v);
        if (lane == 63)
            s_w[w] = inc;
        __syncthreads();
        uint32_t before = carry, tot = carry;
        for (uint32_t k = 0; k < 16; k++) {
            before = k < w ? max(before, s_w[k]) : before;
            tot = max(tot, s_w[k]);
        }
        const uint32_t prev = __shfl_up(inc, 1, 64);
        if (i < blocks)
            a.blk_in[i] = max(before, lane ? prev : 0u);
        carry = tot;
        __syncthreads();
    }
}

// The rest of a parent chain, from parent p of a line whose last bare blank
// line before it is b1 - 1 (b1 = 0: none): literal (an element chain that
// ends at a depth-0 element is not), and the nearest ancestor with a block.
__device__ __forceinline__ uint32_t bridge_literal_from(const BridgeArgs &a, uint32_t p, uint32_t b1)
{
    for (uint32_t step = 0; step < TREE_D; step++) {
        // p is a depth > 0 element, not cut: its own parent decides
        const uint32_t q = a.parent[p];
        if (q == OL_NONE || b1 > q + 1 || a.classes[q] != OLK_ELEMENT)
            return BR_LITERAL;
        if (a.keys[q] == 0)
            return 0;
        p = q;
    }
    return 0;
}
__device__ __forceinline__ uint32_t bridge_container_from(const BridgeArgs &a, uint32_t p, uint32_t b1)
{
    for (uint32_t step = 0; step < TREE_D; step++) {
        // p has no block and is not cut: its own parent is next
        const uint32_t q = a.parent[p];
        if (q == OL_NONE || b1 > q + 1)
            return OL_NONE;
        if (a.flags[q] & BR_BLOCK)
            return q;
        p = q;
    }
    return OL_NONE;
}

__global__ __launch_bounds__(BRIDGE_THREADS) void bridge_link_kernel(BridgeArgs a)
{
    __shared__ uint32_t s_w[BRIDGE_THREADS / 64];
    const uint32_t i0 = blockIdx.x * BRIDGE_BLOCK + threadIdx.x * BRIDGE_PER_THREAD;
    const bool any = i0 < a.lines;
    uint4 fv = make_uint4(0, 0, 0, 0);
    if (any)
        fv = *(const uint4 *)(a.flags + i0);
    const uint32_t fw[4] = {fv.x, fv.y, fv.z, fv.w};
    // the thread's last bare blank line (+ 1), then the last one before its lines
    uint32_t last = 0;
#pragma unroll
    for (uint32_t k = 0; k < BRIDGE_PER_THREAD; k++)
        if ((fw[k >> 2] >> (8 * (k & 3))) & BR_BARE)
            last = i0 + k + 1;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = (uint32_t)wave_incl_max((int32_t)last);
    if (lane == 63)
        s_w[w] = inc;
    __syncthreads();
    uint32_t b1 = a.blk_in[blockIdx.x];
    for (uint32_t k = 0; k < BRIDGE_THREADS / 64; k++)
        b1 = k < w ? max(b1, s_w[k]) : b1;
    const uint32_t prev = __shfl_up(inc, 1, 64);
    b1 = max(b1, lane ? prev : 0u);
    if (!any)
        return;
    const uint4 cl = *(const uint4 *)(a.classes + i0), ky = *(const uint4 *)(a.keys + i0);
    const uint32_t cw[4] = {cl.x, cl.y, cl.z, cl.w}, kw[4] = {ky.x, ky.y, ky.z, ky.w};
    uint32_t par[BRIDGE_PER_THREAD];
#pragma unroll
    for (uint32_t q = 0; q < BRIDGE_PER_THREAD / 4; q++) {
        const uint4 v = *(const uint4 *)(a.parent + i0 + 4 * q);   // padded to 4,096 lines
        par[4 * q] = v.x;
        par[4 * q + 1] = v.y;
        par[4 * q + 2] = v.z;
        par[4 * q + 3] = v.w;
    }
    // each line's parent: class | key << 8 | flags << 16, the loads issued together
    uint32_t pa[BRIDGE_PER_THREAD];
#pragma unroll
    for (uint32_t k = 0; k < BRIDGE_PER_THREAD; k++) {
        const uint32_t p = (i0 + k < a.lines && par[k] != OL_NONE) ? par[k] : i0;
        pa[k] = (uint32_t)a.classes[p] | (uint32_t)a.keys[p] << 8 | (uint32_t)a.flags[p] << 16;
    }
    uint32_t out[4] = {0, 0, 0, 0}, cont[BRIDGE_PER_THREAD];
    for (uint32_t k = 0; k < BRIDGE_PER_THREAD; k++) {
        const uint32_t i = i0 + k;
        const uint32_t kind = (cw[k >> 2] >> (8 * (k & 3))) & 0xFFu, key = (kw[k >> 2] >> (8 * (k & 3))) & 0xFFu;
        uint32_t fl = (fw[k >> 2] >> (8 * (k & 3))) & 0xFFu;
        uint32_t ct = OL_NONE;
        if (i < a.lines) {
            // b1: the last bare blank line before line i, + 1 (0: none)
            const uint32_t p = par[k];
            const bool cut = p != OL_NONE && b1 > p + 1;
            if (key == KEY_BLANK) {
                if (!(fl & BR_BARE))
                    fl |= BR_LITERAL;   // a `|` line with no content: parent NONE
            } else if (key != 0) {
                // literal: up the parent chain to a depth-0 parent
                uint32_t lit;
                if (p == OL_NONE || cut || (pa[k] & 0xFFu) != OLK_ELEMENT)
                    lit = BR_LITERAL;
                else if (((pa[k] >> 8) & 0xFFu) == 0)
                    lit = 0;
                else
                    lit = bridge_literal_from(a, p, b1);
                fl |= lit | (cut ? BR_CUT : 0u);
                // the nearest ancestor with a block along the cut parents
                if ((fl & BR_BLOCK) && p != OL_NONE && !cut)
                    ct = ((pa[k] >> 16) & BR_BLOCK) ? p : bridge_container_from(a, p, b1);
            }
            if (kind == OLK_ELEMENT && !(fl & BR_LITERAL)) {
                // a `|` child: the next content line is deeper, no bare blank line before it
                uint32_t j = i + 1;
                bool blocked = false;
                while (j < a.lines && a.keys[j] == KEY_BLANK) {
                    if (a.flags[j] & BR_BARE) {
                        blocked = true;
                        break;
                    }
                    j++;
                }
                if (!blocked && j < a.lines && a.keys[j] < KEY_DEEP && a.keys[j] > key)
                    fl |= BR_PIPE;
            }
            if (fl & BR_BARE)
                b1 = i + 1;
        }
        out[k >> 2] |= fl << (8 * (k & 3));
        cont[k] = ct;
    }
    *(uint4 *)(a.flags + i0) = make_uint4(out[0], out[1], out[2], out[3]);
#pragma unroll
    for (uint32_t q = 0; q < BRIDGE_PER_THREAD / 4; q++)
        *(uint4 *)(a.container + i0 + 4 * q) = make_uint4(cont[4 * q], cont[4 * q + 1], cont[4 * q + 2], cont[4 * q + 3]);
}

hipError_t ol_launch_bridge(const BridgeArgs &a, hipStream_t s)
{
    if (a.lines == 0)
        return hipSuccess;
    const uint32_t blocks = ol_bridge_blocks(a.lines);
    hipLaunchKernelGGL(bridge_mark_kernel, dim3(blocks), dim3(BRIDGE_THREADS), 0, s, a);
    hipLaunchKernelGGL(bridge_scan_kernel, dim3(1), dim3(1024), 0, s, a, blocks);
    hipLaunchKernelGGL(bridge_link_kernel, dim3(blocks), dim3(BRIDGE_THREADS), 0, s, a);
    return hipGetLastError();
}

uint32_t ol_bridge_blocks(uint64_t lines) { return (uint32_t)((lines + BRIDGE_BLOCK - 1) / BRIDGE_BLOCK); }

// Classes and tree keys from FlatNodes (documents loaded from a .olb file,
// whose nodes were parsed when it was written).
__global__ void nodes_to_lines_kernel(const ol_node_dev *nodes, uint64_t lines, uint8_t *classes, uint8_t *keys)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < lines; i += stride) {
        const ol_node_dev nd = nodes[i];
        classes[i] = (uint8_t)nd.kind;
        keys[i] = key_of(nd);
    }
}

hipError_t ol_launch_nodes_to_lines(const ol_node_dev *nodes, uint64_t lines, uint8_t *classes, uint8_t *keys,
                                    hipStream_t s)
{
    if (lines == 0)
        return hipSuccess;
    uint64_t blocks = (lines + 255) / 256;
    if (blocks > 8192)
        blocks = 8192;
    hipLaunchKernelGGL(nodes_to_lines_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nodes, lines, classes, keys);
    return hipGetLastError();
}
