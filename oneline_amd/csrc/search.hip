// search.hip -- single-literal and dictionary search with ordered output.
//
// Both find the matches that START in each tile, count them, get the tile's
// output position from the tile counts before it, and write the matches
// there.  Output is therefore ascending by offset (then by pattern id) without
// any sort, and a capacity smaller than the total keeps the first matches.
//
// Literal: every overlapping occurrence (find_all,
// crates/oneline-simd/src/search.rs:14-99; GPU: crates/oneline-gpu/src/search.wgsl:100-173).
// Dictionary: every (start, id) where pattern id occurs, duplicates reporting
// each id (ManyPatterns, crates/oneline-simd/src/search.rs:206-283; GPU:
// crates/oneline-gpu/src/many_fused.wgsl:71-141).  Patterns are keyed by their
// first min(len, 4) bytes; a filter bitmap per key length sits in LDS, the
// hash tables, entries and pattern bytes are read through L2.
#include "internal.h"

// Block-wide exclusive sum of `v` (THREADS/64 waves); also returns the total.
template <int THREADS>
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t *s_w, uint32_t &total)
{
    constexpr int NW = THREADS / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o)
            inc += y;
    }
    if (lane == 63)
        s_w[wid] = inc;
    __syncthreads();
    uint32_t wb = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        if (w < wid)
            wb += s_w[w];
        total += s_w[w];
    }
    return wb + inc - v;
}

uint32_t ol_coresident(const void *f, int threads, size_t dyn_lds)
{
    int dev = 0, ncu = 0, per_cu = 0, lds_cu = 0;
    hipFuncAttributes fa{};
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
        hipFuncGetAttributes(&fa, f) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, threads, dyn_lds) != hipSuccess)
        return 0;
    const size_t lds = ((size_t)fa.sharedSizeBytes + dyn_lds + 4095) / 4096 * 4096;
    if (lds > 0) {
        const int by_lds = (int)((size_t)lds_cu / lds);
        if (by_lds < per_cu)
            per_cu = by_lds;
    }
    return per_cu > 0 ? (uint32_t)(per_cu * ncu) : 0u;
}

static uint32_t max_grid(const void *f, int threads, size_t lds) { return ol_coresident(f, threads, lds); }

// ---- single literal -----------------------------------------------------------
// Persistent, cooperative launch: the last workgroup is the prefix scanner
// (scan_prefixes); the others take tiles t = blockIdx + k * workers in
// increasing order.  No LDS staging: lane l holds the tile's 16-byte vectors
// k * 256 + l (k < 4, coalesced loads) plus the first word of the vector after
// each, so every 4-byte window of its 64 positions is an alignbyte of its own
// registers.  Windows that equal the pattern's first (up to) four bytes are
// verified against the rest from global memory (an L2 hit).  Matches are
// ordered by vector (k major), counted per vector in 16-bit fields, and kept
// as one 64-bit mask per lane until the next iteration, when the scanner has
// (almost always) produced the tile's prefix; then they are written.  The next
// tile is loaded while the current one is processed.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void search_kernel(SearchArgs a)
{
    constexpr int NV = 4;                     // vectors per lane
    constexpr int TILE = THREADS * NV * 16;
    constexpr int NW = THREADS / 64;
    __shared__ uint32_t s_w[2][NW];
    __shared__ uint64_t s_prefix;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t G = gridDim.x - 1;   // workgroups that take tiles
    const uint32_t T = a.num_tiles;
    if (blockIdx.x == G) {
        if (wid == 0)
            scan_prefixes<SCAN_PREFIX_WINDOWS>(a.status, T, a.lb_err);
        return;
    }
    uint4 v[NV];
    uint32_t nx[NV];
    auto load = [&](uint32_t tt) {
        const uint8_t *src = a.text + (size_t)tt * TILE;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            v[k] = ((const uint4 *)src)[k * THREADS + tid];
            nx[k] = *(const uint32_t *)(src + (size_t)(k * THREADS + tid + 1) * 16);   // text is padded
        }
    };
    const uint32_t mlen = a.m;
    const uint32_t pat4 = a.win, mask4 = a.win_mask;
    // the rest of a candidate at text position gp: pattern bytes 4 .. m
    auto verify = [&](uint32_t gp) -> bool {
        for (uint32_t i = 4; i < mlen; i += 4) {
            const uint32_t x = gp + i;
            const uint32_t *q = (const uint32_t *)(a.text + (x & ~3u));
            const uint32_t sx = x & 3u;
            const uint32_t w = sx ? __builtin_amdgcn_alignbyte(q[1], q[0], sx) : q[0];
            uint32_t pw = 0;
#pragma unroll
            for (int b = 0; b < 4; b++)
                pw |= (uint32_t)a.pat[(i + b) & 63] << (8 * b);
            const uint32_t nb = mlen - i;
            const uint32_t msk = nb >= 4 ? ~0u : ((1u << (8 * nb)) - 1u);
            if ((w ^ pw) & msk)
                return false;
        }
        return true;
    };
    // the previous tile: its matches, waiting for its prefix
    bool pend = false;
    uint32_t p_t = 0, p_total = 0;
    uint32_t p_off[NV];
    uint64_t p_match = 0;
    auto flush = [&]() {   // writes the previous tile's matches (all threads)
        if (wid == 0 && lane == 0) {
            uint64_t w = 0;
            uint32_t spins = 0;
            while (((w = ld_coherent_u64(a.status + (size_t)p_t * SCAN_STATUS_WORDS + SCAN_PREFIX_WORD)) >> LB_SHIFT) !=
                       LB_FLAG_INCL &&
                   ++spins < LB_SPIN_LIMIT)
                __builtin_amdgcn_s_sleep(8);
            if (spins >= LB_SPIN_LIMIT) {
                atomicOr(a.lb_err, 1u);
                w = (LB_FLAG_INCL << LB_SHIFT) | p_total;
            }
            s_prefix = (w & LB_MASK) - p_total;
        }
        __syncthreads();
        const uint64_t prefix = s_prefix;
        if (tid == 0 && p_t == T - 1)
            a.meta[0] = prefix + p_total;
        const uint32_t pbase = p_t * TILE;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            uint32_t m = (uint32_t)(p_match >> (16 * k)) & 0xFFFFu;
            uint64_t r = prefix + p_off[k];
            while (m) {
                const uint32_t bit = (uint32_t)__builtin_ctz(m);
                m &= m - 1;
                const uint32_t gp = pbase + (uint32_t)(k * THREADS + tid) * 16 + bit;
                if (r < a.cap) {
                    if (a.out64) a.out64[r] = gp;
                    if (a.out32) a.out32[r] = gp;
                }
                r++;
            }
        }
    };

    uint32_t t = blockIdx.x;
    if (t < T)
        load(t);
    for (; t < T; t += G) {
        const uint32_t base = t * TILE;
        uint4 cv[NV];
        uint32_t cn[NV];
#pragma unroll
        for (int k = 0; k < NV; k++) {
            cv[k] = v[k];
            cn[k] = nx[k];
        }
        if (t + G < T)
            load(t + G);
        // positions gp with gp + m <= n may match
        const uint32_t last_start = a.n >= mlen ? a.n - mlen : 0u;   // inclusive, when n >= m
        uint64_t match = 0;
        uint32_t cnt01 = 0, cnt23 = 0;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const uint32_t W[5] = {cv[k].x, cv[k].y, cv[k].z, cv[k].w, cn[k]};
            const uint32_t g = base + (uint32_t)(k * THREADS + tid) * 16;
            uint32_t hits = 0;
#pragma unroll
            for (int p = 0; p < 16; p++) {
                const uint32_t win = (p & 3) ? __builtin_amdgcn_alignbyte(W[(p >> 2) + 1], W[p >> 2], p & 3)
                                             : W[p >> 2];
                hits |= (uint32_t)((win & mask4) == pat4) << p;
            }
            // positions past the last possible start
            if (a.n < mlen || g > last_start)
                hits = 0;
            else if (last_start - g < 15)
                hits &= (2u << (last_start - g)) - 1u;
            if (hits && mlen > 4) {
                uint32_t q = hits;
                while (q) {
                    const uint32_t bit = (uint32_t)__builtin_ctz(q);
                    q &= q - 1;
                    if (!verify(g + bit))
                        hits &= ~(1u << bit);
                }
            }
            match |= (uint64_t)hits << (16 * k);
            const uint32_t c = (uint32_t)__builtin_popcount(hits);
            if (k == 0) cnt01 |= c;
            if (k == 1) cnt01 |= c << 16;
            if (k == 2) cnt23 |= c;
            if (k == 3) cnt23 |= c << 16;
        }
        // offsets of each vector's matches in the tile (vector order k major):
        // two 16-bit counts per word, block-scanned together
        const uint32_t i01 = wave_incl_add(cnt01), i23 = wave_incl_add(cnt23);
        if (lane == 63) {
            s_w[0][wid] = i01;
            s_w[1][wid] = i23;
        }
        __syncthreads();
        uint32_t b01 = 0, b23 = 0, t01 = 0, t23 = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            if (w < wid) {
                b01 += s_w[0][w];
                b23 += s_w[1][w];
            }
            t01 += s_w[0][w];
            t23 += s_w[1][w];
        }
        const uint32_t e01 = b01 + i01 - cnt01, e23 = b23 + i23 - cnt23;
        const uint32_t T0 = t01 & 0xFFFFu, T1 = t01 >> 16, T2 = t23 & 0xFFFFu, T3 = t23 >> 16;
        const uint32_t total = T0 + T1 + T2 + T3;
        if (tid == 0)
            st_coherent_u64(a.status + (size_t)t * SCAN_STATUS_WORDS, (LB_FLAG_AGG << LB_SHIFT) | total);
        if (pend)
            flush();
        pend = true;
        p_t = t;
        p_total = total;
        p_off[0] = e01 & 0xFFFFu;
        p_off[1] = T0 + (e01 >> 16);
        p_off[2] = T0 + T1 + (e23 & 0xFFFFu);
        p_off[3] = T0 + T1 + T2 + (e23 >> 16);
        p_match = match;
        __syncthreads();   // s_w and s_prefix are reused
    }
    if (pend)
        flush();
}

hipError_t ol_launch_search(const SearchArgs &a, hipStream_t s)
{
    static uint32_t cap = 0;   // co-resident limit (same device for the process)
    const void *f = (const void *)search_kernel<SEARCH_THREADS>;
    if (!cap)
        cap = max_grid(f, SEARCH_THREADS, 0);
    if (!cap)
        return hipErrorInvalidValue;
    // workgroups taking tiles plus the prefix scanner
    uint32_t grid = a.num_tiles + 1 < cap ? a.num_tiles + 1 : cap;
    if (grid < 2)
        return hipErrorInvalidValue;
    SearchArgs args = a;
    void *params[] = {&args};
    return hipLaunchCooperativeKernel(f, dim3(grid), dim3(SEARCH_THREADS), params, 0, s);
}

// ---- dictionary -----------------------------------------------------------------
// Three launches with no waits between workgroups:
//   count pass   per tile (256 lanes x 64 positions): filter, candidate list,
//                per-candidate match counts -> tile_count[t]; the candidates
//                that have matches are stored with their counts (4 bytes
//                each) in the tile's own region of the store when the tile's
//                candidates fit it;
//   scan         exclusive scan of tile_count -> tile_off, total;
//   write pass   per tile: its stored candidates (or filter and count again),
//                each candidate's matches written at tile_off[t] + (matches
//                of the earlier candidates).
// Filter: each lane tests its 64 positions against the LDS bitmaps of the
// dictionary's key classes (straight-line code, 4-byte windows from
// registers).  Probe: the class tables give, per class, the run of patterns
// sharing the key (sorted by id); the runs are merged by id and each pattern's
// remaining bytes compared word by word.  The candidate list goes through LDS
// in chunks of MANY_LCAP, one candidate per thread.  The text is read from
// global memory (the probes hit L2), so the LDS holds only the filters and
// the list.

__device__ __forceinline__ uint32_t filter_bit(const uint32_t *f, uint32_t i)
{
    return __builtin_amdgcn_ubfe(f[i >> 5], i & 31, 1);
}

constexpr int MANY_LCAP = 2048;     // candidates of a tile per list chunk

// Matches starting at tile position lp (global gp) with at most `rem` bytes
// (1..64) before the text or document end; written from out[r] in pattern-id
// order (up to cap) when `emit`.  Returns their number.
__device__ __forceinline__ uint32_t probe_position(const DictDev &d, const uint32_t *s_filter, const uint8_t *s_text,
                                                   uint32_t lp, uint64_t gp, uint32_t rem, bool emit, uint64_t *out_off,
                                                   uint32_t *out_id, uint64_t r, uint64_t cap)
{
    // s_text: the tile's first byte in global memory (word aligned)
    const uint32_t *wp = (const uint32_t *)(s_text + (lp & ~3u));
    const uint32_t sh = lp & 3u;
    const uint32_t a0 = wp[0], a1 = wp[1], a2 = wp[2];
    const uint32_t w0 = sh ? __builtin_amdgcn_alignbyte(a1, a0, sh) : a0;
    const uint32_t w1 = sh ? __builtin_amdgcn_alignbyte(a2, a1, sh) : a1;
    // One run per class (1..5): the patterns whose key matches the text here.
    // Named variables per class (no arrays: nothing may be indexed at run time).
    // First the table slots of every class are loaded together, then resolved.
    // A class's table is read only where its filter passes: the 3-byte keys'
    // filter for class 3, the 4-byte prefix filter for classes 4 and 5 (a
    // candidate passed at least one of them).
    const bool pass3 = !d.fbits[3] || filter_bit(s_filter + d.foff[3], dict_fh3(w0) >> 16);
    const bool pass45 = !d.fbits[4] || filter_bit(s_filter, dict_fh4(w0) >> (32 - d.fbits[4]));
#define OL_KEY(K, LO)                                                                       \
    const uint32_t klo##K = (LO), khi##K = (K) == 5 ? w1 : 0u;                              \
    const bool use##K = ((d.class_mask >> (K)) & 1u) && ((K) <= 4 ? (K) : 8u) <= rem &&    \
                        ((K) == 3 ? pass3 : (K) >= 4 ? pass45 : true);                       \
    uint32_t h##K = use##K ? dict_thash(klo##K, khi##K, d.tbits[K]) : 0u;                  \
    uint4 v##K = use##K ? d.table[K][h##K] : make_uint4(0, 0, 0, 0);
    OL_KEY(1, w0 & 0xFFu)
    OL_KEY(2, w0 & 0xFFFFu)
    OL_KEY(3, w0 & 0xFFFFFFu)
    OL_KEY(4, w0)
    OL_KEY(5, w0)
#undef OL_KEY
#define OL_RUN(K)                                                                           \
    while (v##K.z != 0 && !(v##K.x == klo##K && v##K.w == khi##K)) {                        \
        h##K = (h##K + 1) & ((1u << d.tbits[K]) - 1u);                                       \
        v##K = d.table[K][h##K];                                                             \
    }                                                                                        \
    uint32_t b##K = v##K.y, e##K = v##K.z != 0 ? v##K.z : v##K.y;
    OL_RUN(1)
    OL_RUN(2)
    OL_RUN(3)
    OL_RUN(4)
    OL_RUN(5)
#undef OL_RUN
    if (b1 >= e1 && b2 >= e2 && b3 >= e3 && b4 >= e4 && b5 >= e5)
        return 0;
    // text word at tile position x (any alignment)
    auto tword = [&](uint32_t x) -> uint32_t {
        const uint32_t *q = (const uint32_t *)(s_text + (x & ~3u));
        const uint32_t sx = x & 3u;
        return sx ? __builtin_amdgcn_alignbyte(q[1], q[0], sx) : q[0];
    };
    auto tail_mask = [](uint32_t nb) -> uint32_t { return nb >= 4 ? ~0u : ((1u << (8 * nb)) - 1u); };
    const uint4 none = make_uint4(0xFFFFFFFFu, 0, 0, 0);
    uint4 c1 = b1 < e1 ? d.ents[b1] : none, c2 = b2 < e2 ? d.ents[b2] : none, c3 = b3 < e3 ? d.ents[b3] : none,
          c4 = b4 < e4 ? d.ents[b4] : none, c5 = b5 < e5 ? d.ents[b5] : none;
    uint32_t count = 0;
    while (true) {
        // the run whose next entry has the smallest pattern id (runs are sorted by id)
        uint32_t bid = 0xFFFFFFFFu, bk = 0;
#define OL_PICK(K)                                                                          \
        if (b##K < e##K && c##K.x < bid) {                                                   \
            bid = c##K.x;                                                                    \
            bk = (K);                                                                        \
        }
        OL_PICK(1) OL_PICK(2) OL_PICK(3) OL_PICK(4) OL_PICK(5)
#undef OL_PICK
        if (!bk)
            break;
        uint4 en = none;
        uint32_t kb = 0;
#define OL_ADV(K)                                                                           \
        if (bk == (K)) {                                                                     \
            en = c##K;                                                                       \
            kb = (K) <= 4 ? (K) : 8u;                                                        \
            if (++b##K < e##K)                                                               \
                c##K = d.ents[b##K];                                                         \
        }
        OL_ADV(1) OL_ADV(2) OL_ADV(3) OL_ADV(4) OL_ADV(5)
#undef OL_ADV
        const uint32_t len = en.y;
        if (len > rem)
            continue;
        bool ok = true;
        if (len > kb) {
            ok = ((tword(lp + kb) ^ en.w) & tail_mask(len - kb)) == 0;
            for (uint32_t j = kb + 4; ok && j < len; j += 4)
                ok = ((tword(lp + j) ^ d.pool[en.z + (j >> 2)]) & tail_mask(len - j)) == 0;
        }
        if (!ok)
            continue;
        if (emit && r + count < cap) {
            if (out_off) out_off[r + count] = gp;
            if (out_id) out_id[r + count] = bid;
        }
        count++;
    }
    return count;
}


// Stored candidate: tile position (14 bits) | bytes before the text or
// document end (7 bits) << 14 | match count (11 bits, CAND_RECOUNT: count
// again) << 21.
constexpr uint32_t CAND_RECOUNT = 2047u;
constexpr uint64_t TILE_LIST_NONE = ~0ull;   // tile_list: candidates not stored

// Per-tile state of the dictionary search: the tile's text, and the lanes'
// document bounds (batches) for the end bound of a match.
struct ManyTile {
    const uint8_t *g_text;   // the tile's first byte (global memory)
    uint32_t base, lo, g0, lim;
    uint32_t di0, dlim0, dnext0;
};

template <int THREADS>
__device__ __forceinline__ ManyTile many_tile_setup(const ManyArgs &a, uint32_t t)
{
    constexpr int TILE = THREADS * 64;
    ManyTile m;
    m.base = t * TILE;
    m.g_text = a.text + m.base;
    m.lo = threadIdx.x * 64;
    m.g0 = m.base + m.lo;
    m.lim = (m.g0 < a.n) ? ((a.n - m.g0) < 64u ? a.n - m.g0 : 64u) : 0u;
    // match-end bound per position: the text end, or (batches) the end of
    // the position's document; positions between documents match nothing
    uint32_t dlim = a.n, dnext = 0xFFFFFFFFu, di = 0;
    if (a.ndocs) {
        di = a.tile_doc[t];   // then the last document starting at or before g0
        while (di + 1 < a.ndocs && a.doc_starts[di + 1] <= m.g0)
            di++;
        dlim = a.doc_ends[di];
        dnext = (di + 1 < a.ndocs) ? a.doc_starts[di + 1] : 0xFFFFFFFFu;
    }
    m.di0 = di;
    m.dlim0 = dlim;
    m.dnext0 = dnext;
    return m;
}

// The lane's 64 positions through the filters: bit p set when a pattern may
// start at g0 + p.  FS: the filters the dictionary has (bit 1: 1- and 2-byte
// keys, bit 3: 3-byte keys, bit 4: 4-byte prefixes); compile-time, so the 64
// lookups are straight-line code whose LDS reads the compiler can overlap.
template <int FS>
__device__ __forceinline__ uint64_t many_filter(const DictDev &d, const uint32_t *s_filter, const ManyTile &m)
{
    if (!m.lim)
        return 0;
    // Filter 4, when the dictionary has it, is first in LDS.  The 1- and
    // 2-byte filters are looked up together: an absent one reads the zero
    // word (index mask 0).
    const uint32_t *f1 = s_filter + d.foff[1], *f2 = s_filter + d.foff[2], *f3 = s_filter + d.foff[3];
    const uint32_t *f4 = s_filter;
    const uint32_t m1 = d.fbits[1] ? 0xFFu : 0u, m2 = d.fbits[2] ? 0xFFFFu : 0u;
    const uint32_t sh4 = 32 - d.fbits[4];   // filter 4 index = hash >> sh4
    uint32_t W[20];
    const uint4 *vsrc = (const uint4 *)(m.g_text + m.lo);
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const uint4 v = vsrc[j];
        W[4 * j] = v.x;
        W[4 * j + 1] = v.y;
        W[4 * j + 2] = v.z;
        W[4 * j + 3] = v.w;
    }
    uint32_t clo = 0, chi = 0;
#pragma unroll
    for (int p = 0; p < 64; p++) {
        const uint32_t w = (p & 3) ? __builtin_amdgcn_alignbyte(W[(p >> 2) + 1], W[p >> 2], p & 3) : W[p >> 2];
        uint32_t hit = 0;
        if (FS & 0x10) hit |= filter_bit(f4, dict_fh4(w) >> sh4);
        if (FS & 0x08) hit |= filter_bit(f3, dict_fh3(w) >> 16);
        if (FS & 0x02) hit |= filter_bit(f1, w & m1) | filter_bit(f2, w & m2);
        if (p < 32)
            clo |= hit << p;
        else
            chi |= hit << (p - 32);
    }
    uint64_t cand = (uint64_t)chi << 32 | clo;
    if (m.lim < 64)
        cand &= (1ull << m.lim) - 1ull;
    return cand;
}

// Candidates [c0, c0 + LCAP) of the tile (in position order; the lane's own
// start at index cex) into the LDS list: tile position and end bound.
__device__ __forceinline__ void many_list_chunk(const ManyArgs &a, const ManyTile &m, uint64_t cand, uint32_t cex,
                                                uint32_t c0, uint16_t *s_cand, uint8_t *s_rem)
{
    uint32_t di = m.di0, dlim = m.dlim0, dnext = m.dnext0;
    uint32_t at = cex;
    uint64_t q = cand;
    while (q && at < c0 + (uint32_t)MANY_LCAP) {
        const uint32_t p = (uint32_t)__builtin_ctzll(q);
        q &= q - 1;
        if (at >= c0) {
            const uint32_t gp = m.g0 + p;
            while (gp >= dnext) {
                di++;
                dlim = a.doc_ends[di];
                dnext = (di + 1 < a.ndocs) ? a.doc_starts[di + 1] : 0xFFFFFFFFu;
            }
            const uint32_t rm = gp < dlim ? dlim - gp : 0u;   // 0: no match may start here
            s_cand[at - c0] = (uint16_t)(m.lo + p);
            s_rem[at - c0] = (uint8_t)(rm < 64u ? rm : 64u);
        }
        at++;
    }
}

// Per-candidate exclusive offsets of s_cnt[0 .. n) (contiguous runs per
// thread), in place; returns the chunk's total (every thread).
template <int THREADS>
__device__ __forceinline__ uint32_t many_chunk_offsets(uint32_t *s_cnt, uint32_t n, uint32_t *s_w)
{
    const uint32_t tid = threadIdx.x;
    const uint32_t K = (n + THREADS - 1) / THREADS;
    const uint32_t i0 = min(n, tid * K), i1 = min(n, i0 + K);
    uint32_t loc = 0;
    for (uint32_t i = i0; i < i1; i++)
        loc += s_cnt[i];
    uint32_t total;
    uint32_t run = block_excl_sum<THREADS>(loc, s_w, total);
    for (uint32_t i = i0; i < i1; i++) {
        const uint32_t c = s_cnt[i];
        s_cnt[i] = run;
        run += c;
    }
    __syncthreads();
    return total;
}

// Count pass over one tile: filter, candidate list in chunks of LCAP,
// per-candidate match counts.  The candidates (with their counts) are stored
// for the write pass when the store has room.  Returns the tile's match count.
template <int THREADS, int FS>
__device__ __forceinline__ uint32_t many_count_tile(const ManyArgs &a, uint32_t t, const uint32_t *s_filter,
                                                    uint16_t *s_cand, uint8_t *s_rem, uint32_t *s_cnt,
                                                    uint32_t *s_w)
{
    const DictDev &d = a.d;
    const ManyTile m = many_tile_setup<THREADS>(a, t);
    const uint64_t cand = many_filter<FS>(d, s_filter, m);
    uint32_t total_c;
    const uint32_t cex = block_excl_sum<THREADS>((uint32_t)__popcll(cand), s_w, total_c);
    uint32_t own = 0, kept = 0;   // kept: candidates with matches stored so far (every thread)
    const uint64_t off = (total_c && a.cands && total_c <= a.cand_per_tile) ? (uint64_t)t * a.cand_per_tile
                                                                            : TILE_LIST_NONE;
    for (uint32_t c0 = 0; c0 < total_c; c0 += MANY_LCAP) {
        const uint32_t n = min((uint32_t)MANY_LCAP, total_c - c0);
        many_list_chunk(a, m, cand, cex, c0, s_cand, s_rem);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += THREADS) {
            const uint32_t lp = s_cand[i], rm = s_rem[i];
            const uint32_t c =
                rm ? probe_position(d, s_filter, m.g_text, lp, (uint64_t)m.base + lp, rm, false, nullptr, nullptr, 0, 0)
                   : 0u;
            own += c;
            s_cnt[i] = c;
        }
        __syncthreads();
        if (off != TILE_LIST_NONE) {
            // store the candidates that have matches, in position order: each
            // thread takes a contiguous run of the chunk
            const uint32_t K = (n + THREADS - 1) / THREADS;
            const uint32_t i0 = min(n, threadIdx.x * K), i1 = min(n, i0 + K);
            uint32_t nz = 0;
            for (uint32_t i = i0; i < i1; i++)
                nz += s_cnt[i] != 0;
            uint32_t chunk_kept;
            uint32_t at = kept + block_excl_sum<THREADS>(nz, s_w, chunk_kept);
            for (uint32_t i = i0; i < i1; i++) {
                const uint32_t c = s_cnt[i];
                if (c)
                    a.cands[off + at++] = s_cand[i] | (uint32_t)s_rem[i] << 14 | (c < CAND_RECOUNT ? c : CAND_RECOUNT)
                                                                                    << 21;
            }
            kept += chunk_kept;
        }
        __syncthreads();   // the list is rewritten by the next chunk
    }
    __syncthreads();   // s_w is reused
    uint32_t total;
    (void)block_excl_sum<THREADS>(own, s_w, total);
    if (threadIdx.x == 0)
        a.tile_list[t] = off == TILE_LIST_NONE ? TILE_LIST_NONE : (off << 24 | kept);
    __syncthreads();
    return total;
}

// Write pass over one tile: its matches from out0 on, in position then
// pattern-id order.  With stored candidates (those with matches): read them,
// offsets from their counts, one probe each.  Without: filter and count again
// first.
template <int THREADS, int FS>
__device__ __forceinline__ void many_write_tile(const ManyArgs &a, uint32_t t, const uint32_t *s_filter,
                                                uint16_t *s_cand, uint8_t *s_rem, uint32_t *s_cnt, uint32_t *s_w,
                                                uint64_t out0)
{
    const DictDev &d = a.d;
    const ManyTile m = many_tile_setup<THREADS>(a, t);
    const uint64_t stored = a.tile_list[t];
    uint64_t cand = 0;
    uint32_t total_c, cex = 0;
    if (stored != TILE_LIST_NONE) {
        total_c = (uint32_t)(stored & 0xFFFFFFu);
    } else {
        cand = many_filter<FS>(d, s_filter, m);
        cex = block_excl_sum<THREADS>((uint32_t)__popcll(cand), s_w, total_c);
    }
    uint64_t run = out0;
    for (uint32_t c0 = 0; c0 < total_c; c0 += MANY_LCAP) {
        const uint32_t n = min((uint32_t)MANY_LCAP, total_c - c0);
        if (stored != TILE_LIST_NONE) {
            const uint32_t *src = a.cands + (stored >> 24) + c0;
            for (uint32_t i = threadIdx.x; i < n; i += THREADS) {
                const uint32_t e = src[i];
                const uint32_t lp = e & 0x3FFFu, rm = (e >> 14) & 0x7Fu;
                uint32_t c = e >> 21;
                if (c == CAND_RECOUNT)
                    c = probe_position(d, s_filter, m.g_text, lp, (uint64_t)m.base + lp, rm, false, nullptr, nullptr, 0, 0);
                s_cand[i] = (uint16_t)lp;
                s_rem[i] = (uint8_t)rm;
                s_cnt[i] = c;
            }
        } else {
            many_list_chunk(a, m, cand, cex, c0, s_cand, s_rem);
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n; i += THREADS) {
                const uint32_t lp = s_cand[i], rm = s_rem[i];
                s_cnt[i] = rm ? probe_position(d, s_filter, m.g_text, lp, (uint64_t)m.base + lp, rm, false, nullptr,
                                               nullptr, 0, 0)
                              : 0u;
            }
        }
        __syncthreads();
        const uint32_t ctot = many_chunk_offsets<THREADS>(s_cnt, n, s_w);
        for (uint32_t i = threadIdx.x; i < n; i += THREADS) {
            const uint32_t lp = s_cand[i], rm = s_rem[i];
            const uint64_t r = run + s_cnt[i];
            if (rm && r < a.cap)
                probe_position(d, s_filter, m.g_text, lp, (uint64_t)m.base + lp, rm, true, a.out_off, a.out_id, r,
                               a.cap);
        }
        run += ctot;
        __syncthreads();   // the list is rewritten by the next chunk
    }
    __syncthreads();   // s_w is reused by the next tile
}

template <int THREADS, bool WRITE, int FS>
__global__ __launch_bounds__(THREADS) void many_tiles_kernel(ManyArgs a)
{
    extern __shared__ __align__(16) uint32_t s_filter[];
    __shared__ uint16_t s_cand[MANY_LCAP];
    __shared__ uint8_t s_rem[MANY_LCAP];
    __shared__ uint32_t s_cnt[MANY_LCAP];
    __shared__ uint32_t s_w[THREADS / 64];
    for (uint32_t i = threadIdx.x; i < a.d.fwords; i += THREADS)
        s_filter[i] = a.d.filter[i];
    __syncthreads();
    for (uint32_t t = blockIdx.x; t < a.num_tiles; t += gridDim.x) {
        if (WRITE) {
            many_write_tile<THREADS, FS>(a, t, s_filter, s_cand, s_rem, s_cnt, s_w, a.tile_off[t]);
        } else {
            const uint32_t total = many_count_tile<THREADS, FS>(a, t, s_filter, s_cand, s_rem, s_cnt, s_w);
            if (threadIdx.x == 0) {
                a.tile_count[t] = total;
                if (total)
                    atomicAdd((unsigned long long *)a.meta, (unsigned long long)total);
            }
        }
    }
}

// Exclusive scan of tile_count into tile_off: one workgroup per chunk of
// ST * K tiles.  Each publishes its chunk's total, then waits for the totals
// of the chunks before it (lower workgroups, dispatched first, which wait on
// nothing later) and adds them up; the last chunk's workgroup writes the
// grand total to meta[0].  Chunk totals live in `sums` (zero at launch).
template <int ST>
__global__ __launch_bounds__(ST) void many_scan_kernel(ManyArgs a, uint64_t *sums)
{
    constexpr int K = 8;
    constexpr int CH = ST * K;
    constexpr int NWS = ST / 64;
    __shared__ uint64_t s_wt[NWS];
    __shared__ uint64_t s_carry;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t T = a.num_tiles;
    const uint32_t c0 = blockIdx.x * CH;
    uint32_t loc[K];
    uint64_t m = 0;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint32_t i = c0 + tid * K + k;
        loc[k] = i < T ? a.tile_count[i] : 0u;
        m += loc[k];
    }
    uint64_t inc = m;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(inc, o, 64);
        if (lane >= o)
            inc += y;
    }
    if (lane == 63)
        s_wt[w] = inc;
    __syncthreads();
    uint64_t before = 0, tot = 0;
    for (int k = 0; k < NWS; k++) {
        if (k < w)
            before += s_wt[k];
        tot += s_wt[k];
    }
    // publish the chunk total, then the earlier chunks' (wave 0, one per lane)
    if (tid == 0)
        st_coherent_u64(sums + blockIdx.x, (LB_FLAG_AGG << LB_SHIFT) | tot);
    if (w == 0) {
        uint64_t carry = 0;
        for (uint32_t b0 = 0; b0 < blockIdx.x; b0 += 64) {
            uint64_t v = 0;
            if (b0 + lane < blockIdx.x) {
                uint32_t spins = 0;
                while (((v = ld_coherent_u64(sums + b0 + lane)) >> LB_SHIFT) == 0 && ++spins < LB_SPIN_LIMIT)
                    __builtin_amdgcn_s_sleep(4);
            }
            carry += wave_sum_u64(v & LB_MASK);
        }
        if (lane == 0)
            s_carry = carry;
    }
    __syncthreads();
    uint64_t run = s_carry + before + inc - m;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const uint32_t i = c0 + tid * K + k;
        if (i < T)
            a.tile_off[i] = run;
        run += loc[k];
    }
    if (tid == 0 && blockIdx.x == gridDim.x - 1)
        a.meta[0] = s_carry + tot;
}

size_t ol_many_lds_bytes(uint32_t fwords)
{
    return (size_t)fwords * 4;
}

template <bool WRITE, int FS>
static hipError_t launch_many_tiles_fs(const ManyArgs &a, hipStream_t s)
{
    const size_t lds = ol_many_lds_bytes(a.d.fwords);
    const void *f = (const void *)many_tiles_kernel<SEARCH_THREADS, WRITE, FS>;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    uint32_t cap = max_grid(f, SEARCH_THREADS, lds);
    if (!cap)
        return hipErrorInvalidValue;
    const uint32_t grid = a.num_tiles < 4 * cap ? a.num_tiles : 4 * cap;
    hipLaunchKernelGGL((many_tiles_kernel<SEARCH_THREADS, WRITE, FS>), dim3(grid), dim3(SEARCH_THREADS), lds, s, a);
    return hipGetLastError();
}

// The kernel for the filters the dictionary has: 4-byte prefixes (with the
// 3-byte keys folded in, the usual case), plus 3-byte keys, plus 1/2-byte keys.
template <bool WRITE>
static hipError_t launch_many_tiles(const ManyArgs &a, hipStream_t s)
{
    const int fs = (a.d.fbits[4] ? 0x10 : 0) | (a.d.fbits[3] ? 0x08 : 0) | (a.d.fbits[1] || a.d.fbits[2] ? 0x02 : 0);
    switch (fs) {
    case 0x10: return launch_many_tiles_fs<WRITE, 0x10>(a, s);
    case 0x18: return launch_many_tiles_fs<WRITE, 0x18>(a, s);
    case 0x12: return launch_many_tiles_fs<WRITE, 0x12>(a, s);
    case 0x1A: return launch_many_tiles_fs<WRITE, 0x1A>(a, s);
    case 0x08: return launch_many_tiles_fs<WRITE, 0x08>(a, s);
    case 0x0A: return launch_many_tiles_fs<WRITE, 0x0A>(a, s);
    case 0x02: return launch_many_tiles_fs<WRITE, 0x02>(a, s);
    default: return hipErrorInvalidValue;   // a dictionary has at least one filter
    }
}

// Counts (total in meta[0]); with outputs also the scan and the writes.
hipError_t ol_launch_many(const ManyArgs &a, hipStream_t s)
{
    if (a.num_tiles == 0)
        return hipSuccess;
    hipError_t e = launch_many_tiles<false>(a, s);
    if (e != hipSuccess || (!a.out_off && !a.out_id))
        return e;
    const uint32_t chunks = (a.num_tiles + 8 * 1024 - 1) / (8 * 1024);
    hipLaunchKernelGGL((many_scan_kernel<1024>), dim3(chunks), dim3(1024), 0, s, a, a.tile_list + a.num_tiles);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    return launch_many_tiles<true>(a, s);
}
