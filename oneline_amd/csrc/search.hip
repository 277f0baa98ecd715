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
// crates/oneline-gpu/src/many_fused.wgsl:71-141).  Patterns are keyed by class
// (their first 1, 2, 3, 4 or 8 bytes); one prefix filter (plus the 1- and
// 2-byte bitmaps, when used) sits in LDS, the hash tables, entries and pattern
// bytes are read through L2.
#include "internal.h"

#include <mutex>

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
// k * 256 + l (k < NV, coalesced loads) plus the first word of the vector
// after each, so every 4-byte window of its 16 NV positions is an alignbyte of
// its own registers.  Windows that equal the pattern's window (its rarest
// min(m, 4) bytes, at offset wo, chosen on the host) are verified against the
// whole pattern from global memory (an L2 hit).
// Matches are ordered by vector (k major), counted per vector in 16-bit
// fields, and kept as 16-bit masks per vector until the next iteration, when
// the scanner has (almost always) produced the tile's prefix; then they are
// written.  The next tile is loaded while the current one is processed.
template <int THREADS, int NV, bool FULL>
__global__ __launch_bounds__(THREADS) void search_kernel(SearchArgs a)
{
    static_assert(NV % 4 == 0, "vectors per lane: 16-bit counts, two per word");
    constexpr int TILE = THREADS * NV * 16;
    constexpr int NW = THREADS / 64;
    constexpr int NC = NV / 2;   // count words (two 16-bit fields each)
    constexpr int NM = NV / 4;   // match words (four 16-bit vector masks each)
    __shared__ uint32_t s_w[NC][NW];
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_ahead;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint32_t G = gridDim.x - 1;   // workgroups that take tiles
    const uint32_t T = a.num_tiles;
    if (blockIdx.x == G) {
        if (wid == 0) {
            __builtin_amdgcn_s_setprio(3);   // never yields to the tile waves
            scan_prefixes<SCAN_PREFIX_WINDOWS>(a.status, T, a.lb_err);
        }
        return;
    }
    uint4 v[NV];
    uint32_t nx[NV];
    auto load = [&](uint32_t tt) {
        const uint8_t *src = a.text + (size_t)tt * TILE;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            v[k] = ld_stream_u4((const uint4 *)src + k * THREADS + tid);
            // a wave's last lane loads the word after its vector (the text
            // is padded); the other lanes take it from the next lane's vector
            // when the tile is searched
            nx[k] = lane == 63 ? *(const uint32_t *)(src + (size_t)(k * THREADS + tid + 1) * 16) : 0u;
        }
    };
    const uint32_t mlen = a.m, wo = a.wo;
    const uint32_t pat4 = a.win, mask4 = a.win_mask;
    // a candidate starting at text position gp: the pattern bytes outside its
    // window (vmask; the window matched already).  The first 16 bytes are
    // loaded together, so most candidates cost one memory round trip.
    const uint32_t nw = (mlen + 3) / 4;
    auto tword = [&](uint32_t x) -> uint32_t {
        const uint32_t *q = (const uint32_t *)(a.text + (x & ~3u));
        const uint32_t sx = x & 3u;
        return sx ? __builtin_amdgcn_alignbyte(q[1], q[0], sx) : q[0];
    };
    auto verify = [&](uint32_t gp) -> bool {
        uint32_t bad = 0;
#pragma unroll
        for (uint32_t w = 0; w < 4; w++)
            if (w < nw && a.vmask[w])
                bad |= (tword(gp + 4 * w) ^ a.pw[w]) & a.vmask[w];
        for (uint32_t w = 4; !bad && w < nw; w++)
            if (a.vmask[w])
                bad |= (tword(gp + 4 * w) ^ a.pw[w]) & a.vmask[w];
        return bad == 0;
    };
    // the previous tile: its matches, waiting for its prefix
    bool pend = false;
    uint32_t p_t = 0, p_total = 0;
    uint32_t p_off[NV];
    uint64_t p_match[NM];
    auto flush = [&]() {   // writes the previous tile's matches (all threads)
        if (wid == 0 && lane == 0) {
            uint64_t w = 0;
            uint32_t spins = 0;
            while (((w = ld_coherent_u64(a.status + (size_t)p_t * SCAN_STATUS_WORDS + SCAN_PREFIX_WORD)) >> LB_SHIFT) !=
                       LB_FLAG_INCL &&
                   ++spins < LB_SPIN_LIMIT)
                __builtin_amdgcn_s_sleep(8);
            if (spins >= LB_SPIN_LIMIT) {
                atomicOr(a.lb_err, OLE_PREFIX_WAIT);
                w = (LB_FLAG_INCL << LB_SHIFT) | p_total;
            }
            s_prefix = (w & LB_MASK) - p_total;
            s_ahead = spins ? 1u : 0u;
            // count and prefix word consumed: zero them for the next launch
            uint64_t *const st = a.status + (size_t)p_t * SCAN_STATUS_WORDS;
            st_coherent_u64(st, 0ull);
            st_coherent_u64(st + SCAN_PREFIX_WORD, 0ull);
        }
        __syncthreads();
        // a workgroup ahead of the prefix front yields issue slots to the
        // ones behind it (the oldest workgroups on a CU otherwise win by age)
        if (s_ahead)
            __builtin_amdgcn_s_setprio(0);
        else
            __builtin_amdgcn_s_setprio(2);
        const uint64_t prefix = s_prefix;
        if (tid == 0 && p_t == T - 1)
            a.meta[0] = prefix + p_total;
        const uint32_t pbase = p_t * TILE;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            uint32_t m = (uint32_t)(p_match[k / 4] >> (16 * (k % 4))) & 0xFFFFu;
            uint64_t r = prefix + p_off[k];
            while (m) {
                const uint32_t bit = (uint32_t)__builtin_ctz(m);
                m &= m - 1;
                const uint32_t gp = pbase + (uint32_t)(k * THREADS + tid) * 16 + bit - wo;   // the match's start
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
#pragma unroll
        for (int k = 0; k < NV; k++)
            cn[k] = dpp_move<0x130>(cn[k], cv[k].x);   // wave_shl:1: the next lane's first word; lane 63 keeps its own
        // window positions q hold a match start q - wo when wo <= q and
        // q - wo + m <= n
        const uint32_t last_q = a.n >= mlen ? a.n - mlen + wo : 0u;   // inclusive, when n >= m
        uint64_t match[NM];
        uint32_t cnt[NC];
#pragma unroll
        for (int k = 0; k < NM; k++)
            match[k] = 0;
#pragma unroll
        for (int k = 0; k < NC; k++)
            cnt[k] = 0;
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const uint32_t W[5] = {cv[k].x, cv[k].y, cv[k].z, cv[k].w, cn[k]};
            const uint32_t g = base + (uint32_t)(k * THREADS + tid) * 16;
            // FULL: a 4-byte window (no mask).  The 32 KiB-tile kernel gets
            // windows rarer than one per 8 KiB: it first takes each position's
            // compare as a wave mask (one compare, the masks ORed by the scalar
            // unit) and builds the lanes' hit bits only when a lane of the wave
            // has a hit in this vector.
            auto window = [&](int p) -> uint32_t {
                const uint32_t win = (p & 3) ? __builtin_amdgcn_alignbyte(W[(p >> 2) + 1], W[p >> 2], p & 3)
                                             : W[p >> 2];
                return FULL ? win : win & mask4;
            };
            uint64_t any = ~0ull;
            if (NV == 8) {
                any = 0;
#pragma unroll
                for (int p = 0; p < 16; p++)
                    any |= __ballot(window(p) == pat4);
            }
            uint32_t hits = 0;
            if (any) {
#pragma unroll
                for (int p = 0; p < 16; p++)
                    hits |= (uint32_t)(window(p) == pat4) << p;
            }
            // window positions past the last possible one, or before wo
            if (a.n < mlen || g > last_q)
                hits = 0;
            else if (last_q - g < 15)
                hits &= (2u << (last_q - g)) - 1u;
            if (g < wo)
                hits = (wo - g) >= 16 ? 0u : hits & ~((1u << (wo - g)) - 1u);
            if (hits && mlen > 4) {
                uint32_t q = hits;
                while (q) {
                    const uint32_t bit = (uint32_t)__builtin_ctz(q);
                    q &= q - 1;
                    if (!verify(g + bit - wo))
                        hits &= ~(1u << bit);
                }
            }
            match[k / 4] |= (uint64_t)hits << (16 * (k % 4));
            cnt[k / 2] |= (uint32_t)__builtin_popcount(hits) << (16 * (k % 2));
        }
        // offsets of each vector's matches in the tile (vector order k major):
        // two 16-bit counts per word, block-scanned together
        uint32_t inc[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) {
            inc[k] = wave_incl_add(cnt[k]);
            if (lane == 63)
                s_w[k][wid] = inc[k];
        }
        __syncthreads();
        uint32_t run = 0;   // matches of the tile's earlier vectors (k major)
        uint32_t off[NV];   // this tile's: flush() still needs the previous tile's p_off
#pragma unroll
        for (int k = 0; k < NC; k++) {
            uint32_t before = 0, all = 0;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                if (w < wid)
                    before += s_w[k][w];
                all += s_w[k][w];
            }
            const uint32_t ex = before + inc[k] - cnt[k];   // this lane's place among the lanes, per field
            const uint32_t t_lo = all & 0xFFFFu, t_hi = all >> 16;
            off[2 * k] = run + (ex & 0xFFFFu);
            off[2 * k + 1] = run + t_lo + (ex >> 16);
            run += t_lo + t_hi;
        }
        const uint32_t total = run;
        if (tid == 0)
            st_coherent_u64(a.status + (size_t)t * SCAN_STATUS_WORDS, (LB_FLAG_AGG << LB_SHIFT) | total);
        if (pend)
            flush();
        pend = true;
        p_t = t;
        p_total = total;
#pragma unroll
        for (int k = 0; k < NV; k++)
            p_off[k] = off[k];
#pragma unroll
        for (int k = 0; k < NM; k++)
            p_match[k] = match[k];
        __syncthreads();   // s_w and s_prefix are reused
    }
    if (pend)
        flush();
}

template <int NV, bool FULL>
static hipError_t launch_search_nv(const SearchArgs &a, hipStream_t s)
{
    static uint32_t cap = 0;   // co-resident limit (same device for the process)
    const void *f = (const void *)search_kernel<SEARCH_THREADS, NV, FULL>;
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

uint32_t ol_literal_tile(bool dense) { return SEARCH_THREADS * (dense ? 4u : 8u) * 16u; }

// ---- the allocation-free literal search's scratch ---------------------------
// One static device array per GPU (each device holds its own copy until the
// process ends): meta (8 words: [0] the match total), the sticky error word
// (word 8, low half), then two status words per tile for every tile of the
// largest document (16 KiB tiles of OL_MAX_GPU_INPUT_LEN bytes): 4 MiB + 128 B.
// The literal kernel leaves every status word zero again, so the array needs
// clearing only after a call that gave up (ol_noalloc_clear).
__device__ __align__(256) static uint64_t scratch[OL_NOALLOC_WORDS];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
uint64_t *g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

hipError_t ol_noalloc_scratch(int device, uint64_t **out)
{
    *out = nullptr;
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        // the array's device address on this device (the current one): its
        // host-side name is not a device address and is never passed on
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_noalloc_ptr[device] = (uint64_t *)p;
    }
    *out = g_noalloc_ptr[device];
    return hipSuccess;
}

// 8 vectors per lane halve the per-tile synchronisation; 4 leave registers
// for more waves, which hide verifications when the window is frequent.
hipError_t ol_launch_search(const SearchArgs &a, hipStream_t s)
{
    if (a.win_mask == 0xFFFFFFFFu)
        return a.dense ? launch_search_nv<4, true>(a, s) : launch_search_nv<8, true>(a, s);
    return a.dense ? launch_search_nv<4, false>(a, s) : launch_search_nv<8, false>(a, s);
}

// ---- dictionary -----------------------------------------------------------------
// Three launches with no waits between workgroups, all working on wave tiles
// (64 lanes x 32 positions = 2 KiB of text; every wave independent, no
// workgroup barriers after the filter load):
//   count pass   per wave tile: filter, candidate list (per-wave LDS chunk),
//                per-candidate match counts, one candidate per lane ->
//                tile_count[wt]; the candidates that have matches are stored
//                with their counts (4 bytes each) in the wave tile's region of
//                the store when they fit it;
//   scan         exclusive scan of tile_count -> tile_off, total;
//   write pass   per wave tile: its stored candidates (or filter and count
//                again), each candidate's matches written at tile_off[wt] +
//                (matches of the earlier candidates).
// Filter: each lane tests its 32 positions against the LDS bitmaps
// (straight-line code, 4-byte windows from registers): one lookup per
// position in the 64 KiB prefix filter of classes 3..5, plus one in the
// direct 2-byte bitmap (which also holds the 1-byte patterns) when the
// dictionary has such patterns.  Two workgroups of 8
// waves share a CU.  Probe: the class tables give, per class, the run of
// patterns sharing the key (sorted by id); the runs are merged by id and each
// pattern's remaining bytes compared word by word.  The text is read from
// global memory (the probes hit the cache lines the filter just read).

// Matches starting at tile position lp (global gp) with at most `rem` bytes
// (1..64) before the text or document end; written from out[r] in pattern-id
// order (up to cap) when `emit`.  Returns their number.  GATE: look up the
// classes of 3 bytes and more only where the prefix filter (s_filter) passes;
// a candidate that passed it needs no gate.  KEEP: the ids of the first
// MANY_KEEP matches go to keep[].  SMALL: the dictionary may have 1- and
// 2-byte patterns (without them their lookups are compiled out).
//
// One lookup per key class present: the key's two slots are read for every
// class at once, with the text's bytes 0..15 at the position.  A slot holding one pattern decides it from its own words (up to
// 12 bytes; longer ones read their bytes 8.. from the pool); a run's entries
// are read one after another, the runs of the classes merged by pattern id.
constexpr uint32_t MANY_KEEP = 4;

// The slot of (cls, key) at its two places (info 0: none).  The loads are
// independent, so they are in flight together (an unused class reads slot 0
// and is masked afterwards).
__device__ __forceinline__ uint4 dk_lookup(const DictDev &d, uint32_t cls, uint32_t key, bool use)
{
    static_assert(DK_BUCKET == 1, "one slot per bucket");
    const uint32_t b1 = use ? dk_bucket1(key, cls, d.bbits) : 0u, b2 = use ? dk_bucket2(key, cls, d.bbits) : 0u;
    // both loads issued before any compare; bitwise tests and per-word
    // selects, so nothing branches on a loaded value
    const uint4 v0 = d.table[b1], v1 = d.table[b2];
    uint32_t fx = 0, fy = 0, fz = 0, fw = 0;
#define DK_SEL(V)                                                                           \
    {                                                                                        \
        const bool h = (V.x == key) & ((V.y & DK_VALID) != 0u) & ((V.y >> DK_CLASS_SHIFT) == cls); \
        fx = h ? V.x : fx;                                                                   \
        fy = h ? V.y : fy;                                                                   \
        fz = h ? V.z : fz;                                                                   \
        fw = h ? V.w : fw;                                                                   \
    }
    DK_SEL(v0) DK_SEL(v1)
#undef DK_SEL
    return make_uint4(fx, use ? fy : 0u, fz, fw);
}

template <bool GATE, bool KEEP = false, bool SMALL = true>
__device__ __forceinline__ uint32_t probe_position(const DictDev &d, const uint32_t *s_filter, const uint8_t *s_text,
                                                   uint32_t lp, uint64_t gp, uint32_t rem, bool emit, uint64_t *out_off,
                                                   uint32_t *out_id, uint64_t r, uint64_t cap,
                                                   uint32_t *keep = nullptr)
{
    // s_text: the tile's first byte in global memory (word aligned)
    const uint32_t *wp = (const uint32_t *)(s_text + (lp & ~3u));
    const uint32_t sh = lp & 3u;
    const uint32_t a0 = wp[0], a1 = wp[1], a2 = wp[2], a3 = wp[3];
    const uint32_t T0 = __builtin_amdgcn_alignbyte(a1, a0, sh);
    const uint32_t T1 = __builtin_amdgcn_alignbyte(a2, a1, sh);
    const uint32_t T2 = __builtin_amdgcn_alignbyte(a3, a2, sh);
    bool pass = true;
    if (GATE && d.fbits[4]) {
        const uint32_t ph = dict_ph(T0);
        pass = d.pf_hashed ? __builtin_amdgcn_ubfe(s_filter[dict_ph_word(ph)], ph, 1)
                           : __builtin_amdgcn_ubfe(s_filter[dict_pword(T0)], dict_pbit(T0), 1);
    }
    const uint32_t cm = d.class_mask;
    const bool useL = ((cm >> DK_LONG) & 1u) && rem >= 4 && pass;
    const bool use3 = ((cm >> DK_3) & 1u) && rem >= 3 && pass;
    const bool use2 = SMALL && ((cm >> DK_2) & 1u) && rem >= 2;
    const bool use1 = SMALL && ((cm >> DK_1) & 1u) && rem >= 1;
    const uint4 fL = dk_lookup(d, DK_LONG, T0, useL), f3 = dk_lookup(d, DK_3, T0 & 0xFFFFFFu, use3);
    const uint4 f2 = SMALL ? dk_lookup(d, DK_2, T0 & 0xFFFFu, use2) : make_uint4(0, 0, 0, 0);
    const uint4 f1 = SMALL ? dk_lookup(d, DK_1, T0 & 0xFFu, use1) : make_uint4(0, 0, 0, 0);
    // Cursor per class: the head pattern {info, a, b}, and for a run the next
    // entry and how many remain.  A run's first entry is read for every class
    // at once.
    auto is_run = [](const uint4 &f) { return (f.y & DK_VALID) && ((f.y >> DK_LEN_SHIFT) & 0x7Fu) == 0; };
#define OL_CUR(K)                                                                           \
    const bool run##K = is_run(f##K);                                                        \
    const uint4 e##K = d.ents[run##K ? f##K.z : 0u];                                         \
    uint32_t hi##K = run##K ? e##K.x : f##K.y, ha##K = run##K ? e##K.y : f##K.z,             \
             hb##K = run##K ? e##K.z : f##K.w;                                               \
    uint32_t nx##K = run##K ? f##K.z + 1u : 0u, left##K = run##K ? f##K.w - 1u : 0u;
    OL_CUR(L) OL_CUR(3) OL_CUR(2) OL_CUR(1)
#undef OL_CUR
    // text word at tile position x (any alignment)
    auto tword = [&](uint32_t x) -> uint32_t {
        const uint32_t *q = (const uint32_t *)(s_text + (x & ~3u));
        const uint32_t sx = x & 3u;
        return sx ? __builtin_amdgcn_alignbyte(q[1], q[0], sx) : q[0];
    };
    auto tail_mask = [](uint32_t nb) -> uint32_t { return nb >= 4 ? ~0u : ((1u << (8 * nb)) - 1u); };
    uint32_t count = 0;
    while (true) {
        // the class whose head has the smallest pattern id
        uint32_t bid = 0xFFFFFFFFu, bk = 0;
#define OL_PICK(K, N)                                                                       \
        if ((hi##K & DK_VALID) && (hi##K & 0x3FFFFFu) < bid) {                               \
            bid = hi##K & 0x3FFFFFu;                                                         \
            bk = (N);                                                                        \
        }
        OL_PICK(L, 1) OL_PICK(3, 2) OL_PICK(2, 3) OL_PICK(1, 4)
#undef OL_PICK
        if (!bk)
            break;
        uint32_t info = 0, pa = 0, pb = 0;
#define OL_ADV(K, N)                                                                        \
        if (bk == (N)) {                                                                     \
            info = hi##K;                                                                    \
            pa = ha##K;                                                                      \
            pb = hb##K;                                                                      \
            if (left##K) {                                                                   \
                const uint4 en = d.ents[nx##K++];                                            \
                left##K--;                                                                   \
                hi##K = en.x;                                                                \
                ha##K = en.y;                                                                \
                hb##K = en.z;                                                                \
            } else {                                                                         \
                hi##K = 0;                                                                   \
            }                                                                                \
        }
        OL_ADV(L, 1) OL_ADV(3, 2) OL_ADV(2, 3) OL_ADV(1, 4)
#undef OL_ADV
        const uint32_t len = (info >> DK_LEN_SHIFT) & 0x7Fu;
        if (len > rem)
            continue;
        bool ok = true;
        if (bk == 1 && len > 4) {   // a pattern of 5 bytes or more: bytes 4.. against the text
            ok = ((T1 ^ pa) & tail_mask(len - 4)) == 0;
            if (ok && len > 8) {
                if (len <= 12) {
                    ok = ((T2 ^ pb) & tail_mask(len - 8)) == 0;
                } else {
                    ok = ((T2 ^ d.pool[pb]) & tail_mask(len - 8)) == 0;
                    for (uint32_t j = 12; ok && j < len; j += 4)
                        ok = ((tword(lp + j) ^ d.pool[pb + ((j - 8) >> 2)]) & tail_mask(len - j)) == 0;
                }
            }
        }
        if (!ok)
            continue;
        if (emit && r + count < cap) {
            if (out_off) out_off[r + count] = gp;
            if (out_id) out_id[r + count] = bid;
        }
        if (KEEP) {
            if (count == 0) keep[0] = bid;
            if (count == 1) keep[1] = bid;
            if (count == 2) keep[2] = bid;
            if (count == 3) keep[3] = bid;
        }
        count++;
    }
    return count;
}


// The count pass's probe: the number of matches at tile position lp and the
// ids of the MANY_KEEP smallest (keep[], ascending; ~0u past the count).
// s_filter is the prefix filter in LDS.
// Order matters only for those ids, which are kept in four sorted registers,
// so each class's run is walked on its own, the next entry's load issued
// before the current one is verified (the write pass emits in id order with
// probe_position).  GATE and SMALL as for probe_position.
template <bool GATE, bool SMALL>
__device__ __forceinline__ uint32_t probe_count_keep(const DictDev &d, const uint32_t *s_filter, const uint8_t *s_text,
                                                     uint32_t lp, uint32_t rem, uint32_t *keep)
{
    const uint32_t *wp = (const uint32_t *)(s_text + (lp & ~3u));
    const uint32_t sh = lp & 3u;
    const uint32_t a0 = wp[0], a1 = wp[1], a2 = wp[2], a3 = wp[3];
    const uint32_t T0 = __builtin_amdgcn_alignbyte(a1, a0, sh);
    const uint32_t T1 = __builtin_amdgcn_alignbyte(a2, a1, sh);
    const uint32_t T2 = __builtin_amdgcn_alignbyte(a3, a2, sh);
    // The 3-byte lookup is made only where the prefix filter can hold a 3-byte
    // key of the window: sparse keying, the window's word is full (a 3-byte
    // key sets all of it); dense keying, the bits of two other values of
    // byte 3's low bits are set too (a 3-byte key sets all 32).
    const bool has_pf = d.fbits[4] != 0;
    auto hbit = [&](uint32_t w) -> bool {
        const uint32_t ph = dict_ph(w);
        return __builtin_amdgcn_ubfe(s_filter[dict_ph_word(ph)], ph, 1);
    };
    bool pass, gate3;
    if (!has_pf) {
        pass = gate3 = true;
    } else if (d.pf_hashed) {
        pass = !GATE || hbit(T0);
        gate3 = hbit(T0 ^ 1u << 24) && hbit(T0 ^ 2u << 24);
    } else {
        const uint32_t pfw = s_filter[dict_pword(T0)];
        pass = !GATE || __builtin_amdgcn_ubfe(pfw, dict_pbit(T0), 1);
        gate3 = pfw == ~0u;
    }
    const uint32_t cm = d.class_mask;
    const bool useL = ((cm >> DK_LONG) & 1u) && rem >= 4 && pass;
    const bool use3 = ((cm >> DK_3) & 1u) && rem >= 3 && pass && gate3;
    const bool use2 = SMALL && ((cm >> DK_2) & 1u) && rem >= 2;
    const bool use1 = SMALL && ((cm >> DK_1) & 1u) && rem >= 1;
    const uint4 fL = dk_lookup(d, DK_LONG, T0, useL), f3 = dk_lookup(d, DK_3, T0 & 0xFFFFFFu, use3);
    const uint4 f2 = SMALL ? dk_lookup(d, DK_2, T0 & 0xFFFFu, use2) : make_uint4(0, 0, 0, 0);
    const uint4 f1 = SMALL ? dk_lookup(d, DK_1, T0 & 0xFFu, use1) : make_uint4(0, 0, 0, 0);
    auto tword = [&](uint32_t x) -> uint32_t {
        const uint32_t *q = (const uint32_t *)(s_text + (x & ~3u));
        const uint32_t sx = x & 3u;
        return sx ? __builtin_amdgcn_alignbyte(q[1], q[0], sx) : q[0];
    };
    auto tail_mask = [](uint32_t nb) -> uint32_t { return nb >= 4 ? ~0u : ((1u << (8 * nb)) - 1u); };
    uint32_t count = 0, k0 = ~0u, k1 = ~0u, k2 = ~0u, k3 = ~0u;
    // a pattern {info, a, b} of class `cls` at this position: verify, count,
    // and insert its id into the sorted four
    // (selects, no branches, except for the pool words of patterns longer
    // than 12 bytes whose bytes 4..11 already match)
    auto consider = [&](uint32_t info, uint32_t pa, uint32_t pb, bool is_long) {
        const uint32_t len = (info >> DK_LEN_SHIFT) & 0x7Fu;
        const uint32_t m1 = is_long && len > 4 ? tail_mask(len - 4) : 0u;
        const uint32_t m2 = is_long && len > 8 && len <= 12 ? tail_mask(len - 8) : 0u;
        bool ok = (len <= rem) & (((T1 ^ pa) & m1) == 0u) & (((T2 ^ pb) & m2) == 0u);
        if (is_long && len > 12 && ok) {
            ok = ((T2 ^ d.pool[pb]) & tail_mask(len - 8)) == 0;
            for (uint32_t j = 12; ok && j < len; j += 4)
                ok = ((tword(lp + j) ^ d.pool[pb + ((j - 8) >> 2)]) & tail_mask(len - j)) == 0;
        }
        count += ok ? 1u : 0u;
        uint32_t x = ok ? (info & 0x3FFFFFu) : ~0u;   // ~0u leaves the four as they are
        const uint32_t n0 = min(k0, x); x = max(k0, x); k0 = n0;
        const uint32_t n1 = min(k1, x); x = max(k1, x); k1 = n1;
        const uint32_t n2 = min(k2, x); x = max(k2, x); k2 = n2;
        k3 = min(k3, x);
    };
    auto walk = [&](const uint4 &f, bool is_long) {
        if (!(f.y & DK_VALID))
            return;
        if ((f.y >> DK_LEN_SHIFT) & 0x7Fu) {   // one pattern, in the slot
            consider(f.y, f.z, f.w, is_long);
            return;
        }
        // four entries per step, their loads in flight together
        const uint32_t b = f.z, n = f.w, last = f.z + f.w - 1u;
        for (uint32_t i = 0; i < n; i += 4) {
            const uint4 e0 = d.ents[min(b + i, last)], e1 = d.ents[min(b + i + 1, last)],
                        e2 = d.ents[min(b + i + 2, last)], e3 = d.ents[min(b + i + 3, last)];
            consider(e0.x, e0.y, e0.z, is_long);
            if (i + 1 < n) consider(e1.x, e1.y, e1.z, is_long);
            if (i + 2 < n) consider(e2.x, e2.y, e2.z, is_long);
            if (i + 3 < n) consider(e3.x, e3.y, e3.z, is_long);
        }
    };
    walk(fL, true);
    walk(f3, false);
    if (SMALL) {
        walk(f2, false);
        walk(f1, false);
    }
    keep[0] = k0;
    keep[1] = k1;
    keep[2] = k2;
    keep[3] = k3;
    return count;
}


// ---- merged runs (DictDev::merged) -------------------------------------------
// The slot of the first class, in the order L, 3, 2, 1, that holds the key of
// the window T0 (info 0: none).  Its run holds every pattern that can start
// there.  FILT: the prefix filter s_filter and the 2-byte filter s_small are
// at hand (LDS): the 3-byte lookup then needs the filter's 3-byte gate (as in
// probe_count_keep), with GATE the classes of 3 bytes and more the window's
// filter bit, and the 2- and 1-byte classes the window's 2-byte filter bit.
// The short classes are looked up only where neither longer class holds the
// key (a longer key's run holds its 1- and 2-byte prefixes).
template <bool FILT, bool GATE, bool SMALL>
__device__ __forceinline__ uint4 merged_slot(const DictDev &d, const uint32_t *s_filter, const uint32_t *s_small,
                                             uint32_t T0, uint32_t rem)
{
    bool pass = true, gate3 = true;
    if (FILT && d.fbits[4]) {
        if (d.pf_hashed) {
            auto hbit = [&](uint32_t w) -> bool {
                const uint32_t ph = dict_ph(w);
                return __builtin_amdgcn_ubfe(s_filter[dict_ph_word(ph)], ph, 1);
            };
            pass = !GATE || hbit(T0);
            gate3 = hbit(T0 ^ 1u << 24) && hbit(T0 ^ 2u << 24);
        } else {
            const uint32_t pfw = s_filter[dict_pword(T0)];
            pass = !GATE || __builtin_amdgcn_ubfe(pfw, dict_pbit(T0), 1);
            gate3 = pfw == ~0u;
        }
    }
    const uint32_t cm = d.class_mask;
    const bool useL = ((cm >> DK_LONG) & 1u) && rem >= 4 && pass;
    const bool use3 = ((cm >> DK_3) & 1u) && rem >= 3 && pass && gate3;
    const uint4 fL = dk_lookup(d, DK_LONG, T0, useL), f3 = dk_lookup(d, DK_3, T0 & 0xFFFFFFu, use3);
    uint4 f = (fL.y & DK_VALID) ? fL : f3;
    if (SMALL && !(f.y & DK_VALID) && (!FILT || __builtin_amdgcn_ubfe(s_small[__builtin_amdgcn_ubfe(T0, 5, 11)], T0, 1))) {
        const bool use2 = ((cm >> DK_2) & 1u) && rem >= 2, use1 = ((cm >> DK_1) & 1u) && rem >= 1;
        const uint4 f2 = dk_lookup(d, DK_2, T0 & 0xFFFFu, use2), f1 = dk_lookup(d, DK_1, T0 & 0xFFu, use1);
        f = (f2.y & DK_VALID) ? f2 : f1;
    }
    return f;
}

// Whether pattern {info, pa, pb} of a run or slot that holds the position's
// key occurs at tile position lp: its length fits `rem`, its bytes 4..11
// equal the text words T1, T2 there, and its bytes 12.. the text (pool
// words, read only when bytes 4..11 match).  The key bytes need no check.
__device__ __forceinline__ bool pattern_here(const DictDev &d, const uint8_t *s_text, uint32_t lp, uint32_t T1,
                                             uint32_t T2, uint32_t info, uint32_t pa, uint32_t pb, uint32_t rem)
{
    auto tail_mask = [](uint32_t nb) -> uint32_t { return nb >= 4 ? ~0u : ((1u << (8 * nb)) - 1u); };
    const uint32_t len = (info >> DK_LEN_SHIFT) & 0x7Fu;
    const uint32_t m1 = len > 4 ? tail_mask(len - 4) : 0u;
    const uint32_t m2 = len > 8 && len <= 12 ? tail_mask(len - 8) : 0u;
    bool ok = (len <= rem) & (((T1 ^ pa) & m1) == 0u) & (((T2 ^ pb) & m2) == 0u);
    if (len > 12 && ok) {
        ok = (T2 ^ d.pool[pb]) == 0u;
        for (uint32_t j = 12; ok && j < len; j += 4) {
            const uint32_t x = lp + j;
            const uint32_t *q = (const uint32_t *)(s_text + (x & ~3u));
            const uint32_t sx = x & 3u;
            const uint32_t tw = sx ? __builtin_amdgcn_alignbyte(q[1], q[0], sx) : q[0];
            ok = ((tw ^ d.pool[pb + ((j - 8) >> 2)]) & tail_mask(len - j)) == 0u;
        }
    }
    return ok;
}

// The text words at tile position lp: bytes 0..3, 4..7, 8..11.
__device__ __forceinline__ void text_words(const uint8_t *s_text, uint32_t lp, uint32_t &T0, uint32_t &T1,
                                           uint32_t &T2)
{
    const uint32_t *wp = (const uint32_t *)(s_text + (lp & ~3u));
    const uint32_t sh = lp & 3u;
    const uint32_t a0 = wp[0], a1 = wp[1], a2 = wp[2], a3 = wp[3];
    T0 = __builtin_amdgcn_alignbyte(a1, a0, sh);
    T1 = __builtin_amdgcn_alignbyte(a2, a1, sh);
    T2 = __builtin_amdgcn_alignbyte(a3, a2, sh);
}

// Merged layout, one lane on its own: the matches at tile position lp
// (global gp), counted, and written in id order from out[r] (up to cap) when
// `emit`.  For the write pass's rare cases (runs longer than 64 entries,
// wave tiles whose candidates were not stored).
template <bool SMALL>
__device__ uint32_t probe_merged(const DictDev &d, const uint8_t *s_text, uint32_t lp, uint64_t gp, uint32_t rem,
                                 bool emit, uint64_t *out_off, uint32_t *out_id, uint64_t r, uint64_t cap)
{
    uint32_t T0, T1, T2;
    text_words(s_text, lp, T0, T1, T2);
    const uint4 f = merged_slot<false, false, SMALL>(d, nullptr, nullptr, T0, rem);
    if (!(f.y & DK_VALID))
        return 0;
    const uint32_t slen = (f.y >> DK_LEN_SHIFT) & 0x7Fu;
    const uint32_t n = slen ? 1u : f.w;
    uint32_t count = 0;
    for (uint32_t j = 0; j < n; j++) {
        const uint4 e = slen ? make_uint4(f.y, f.z, f.w, 0u) : d.ents[f.z + j];
        if (!pattern_here(d, s_text, lp, T1, T2, e.x, e.y, e.z, rem))
            continue;
        if (emit && r + count < cap) {
            if (out_off) out_off[r + count] = gp;
            if (out_id) out_id[r + count] = e.x & 0x3FFFFFu;
        }
        count++;
    }
    return count;
}

// The write pass's probe of one position for either layout.
template <bool MERGED, bool SMALL>
__device__ __forceinline__ uint32_t probe_any(const DictDev &d, const uint8_t *s_text, uint32_t lp, uint64_t gp,
                                              uint32_t rem, bool emit, uint64_t *out_off, uint32_t *out_id, uint64_t r,
                                              uint64_t cap)
{
    if (MERGED)
        return probe_merged<SMALL>(d, s_text, lp, gp, rem, emit, out_off, out_id, r, cap);
    return probe_position<false, false, SMALL>(d, nullptr, s_text, lp, gp, rem, emit, out_off, out_id, r, cap);
}

// Wave tiles: MANY_H parts of 64 lanes x MANY_LB positions (a lane's
// positions are contiguous within a part; parts follow each other).
constexpr uint32_t MANY_LB = 32;
constexpr uint32_t MANY_PART = 64 * MANY_LB;                // 2 KiB
constexpr uint32_t MANY_WT = MANY_WT_BYTES;
constexpr uint32_t MANY_H = MANY_WT / MANY_PART;            // parts per wave tile
constexpr uint32_t MANY_WPT = SEARCH_TILE / MANY_WT;        // wave tiles per search tile
constexpr uint32_t MANY_WCAP = MANY_WT / 4;                 // stored candidates per wave tile
constexpr uint32_t MANY_WLIST = 512;                        // per-wave LDS list chunk
constexpr uint32_t MANY_PWORDS = 1u << (DICT_PBITS - 5);    // prefix filter words
constexpr uint32_t MANY_JOB_MAX = MANY_WT <= 4096 ? 64 : 64 * 4096 / MANY_WT;   // wave tiles per write job, at most
static_assert(MANY_WPT * MANY_WT == SEARCH_TILE && MANY_H * MANY_PART == MANY_WT, "wave tiles");

// Stored candidate (16 bytes): x = wave-tile position (CAND_POS_BITS) | bytes
// before the text or document end (7 bits) | match count (the rest;
// CAND_RECOUNT: count again); y, z, w = the ids of its first four matches,
// 22 bits each (pattern ids are below OL_MAX_PATTERNS = 2^22); a candidate
// with more is probed again by the write pass.
constexpr uint32_t CAND_POS_BITS = __builtin_ctz(MANY_WT);
constexpr uint32_t CAND_CNT_SHIFT = CAND_POS_BITS + 7;
constexpr uint32_t CAND_RECOUNT = (1u << (32 - CAND_CNT_SHIFT)) - 1u;
constexpr uint64_t TILE_LIST_NONE = ~0ull;   // tile_list: candidates not stored
static_assert(32 - CAND_CNT_SHIFT >= 10, "count bits");
__device__ __forceinline__ uint4 cand_pack(uint32_t x, const uint32_t (&id)[4])
{
    return make_uint4(x, id[0] | id[1] << 22, id[1] >> 10 | id[2] << 12, id[2] >> 20 | id[3] << 2);
}
__device__ __forceinline__ void cand_ids(const uint4 &e, uint32_t (&id)[4])
{
    id[0] = e.y & 0x3FFFFFu;
    id[1] = e.y >> 22 | (e.z & 0xFFFu) << 10;
    id[2] = e.z >> 12 | (e.w & 0x3u) << 20;
    id[3] = e.w >> 2;
}

// A lane's share of wave tile wt (its positions lo + h * MANY_PART + p of the
// tile, p < MANY_LB, in part h): the text and the document bounds (batches)
// for the end bound of a match at its first position.
struct ManyLane {
    const uint8_t *g_text;   // the wave tile's first byte (global memory)
    uint32_t base, lo;
    uint32_t di0, dlim0, dnext0;
};

__device__ __forceinline__ ManyLane many_lane_setup(const ManyArgs &a, uint32_t wt)
{
    ManyLane m;
    m.base = wt * MANY_WT;
    m.g_text = a.text + m.base;
    m.lo = (threadIdx.x & 63u) * MANY_LB;
    const uint32_t g0 = m.base + m.lo;
    // match-end bound per position: the text end, or (batches) the end of
    // the position's document; positions between documents match nothing
    uint32_t dlim = a.n, dnext = 0xFFFFFFFFu, di = 0;
    if (a.ndocs) {
        di = a.tile_doc[wt / MANY_WPT];   // then the last document starting at or before g0
        while (di + 1 < a.ndocs && a.doc_starts[di + 1] <= g0)
            di++;
        dlim = a.doc_ends[di];
        dnext = (di + 1 < a.ndocs) ? a.doc_starts[di + 1] : 0xFFFFFFFFu;
    }
    m.di0 = di;
    m.dlim0 = dlim;
    m.dnext0 = dnext;
    return m;
}

// The lane's text words in part h: MANY_LB bytes and the next word (the
// padded text covers it).
struct ManyWords {
    uint32_t w[MANY_LB / 4 + 1];
};
__device__ __forceinline__ ManyWords many_load(const ManyArgs &a, uint32_t wt, uint32_t h)
{
    const uint8_t *p = a.text + (size_t)wt * MANY_WT + h * MANY_PART + (threadIdx.x & 63u) * MANY_LB;
    ManyWords W;
#pragma unroll
    for (int j = 0; j < (int)MANY_LB / 16; j++) {
        const uint4 v = ((const uint4 *)p)[j];
        W.w[4 * j] = v.x;
        W.w[4 * j + 1] = v.y;
        W.w[4 * j + 2] = v.z;
        W.w[4 * j + 3] = v.w;
    }
    W.w[MANY_LB / 4] = *(const uint32_t *)(p + MANY_LB);
    return W;
}

// One part's positions through the filters: bit p set when a pattern may
// start at the lane's position p.  FS: the filters the dictionary has (bit 4:
// the prefix filter at pf, bit 1: the 2-byte filter, which the host places
// right after the prefix filter in the layout that small points at);
// compile-time, so the lookups are straight-line code
// whose reads the compiler can overlap.  A position of the prefix filter
// costs a byte align, a 24-bit multiply, the word address, one read and the
// bit extract, whose offset is the window three positions on.
template <int FS>
__device__ __forceinline__ uint32_t many_filter(const DictDev &d, const uint32_t *pf, const uint32_t *small,
                                                const ManyWords &W, uint32_t lim)
{
    // the 2-byte filter (both short classes) follows the prefix filter
    const uint32_t *const f2 = small + ((FS & 0x10) ? MANY_PWORDS : 0u);
    uint32_t cand = 0;
    // the window at lane position q (q < MANY_LB), and a word whose lowest
    // byte is the byte at q (q < MANY_LB + 3)
    auto win = [&](int q) -> uint32_t {
        return (q & 3) ? __builtin_amdgcn_alignbyte(W.w[(q >> 2) + 1], W.w[q >> 2], q & 3) : W.w[q >> 2];
    };
    auto low_byte_at = [&](int q) -> uint32_t { return q < (int)MANY_LB ? win(q) : W.w[q >> 2] >> (8 * (q & 3)); };
    static_assert(MANY_LB % 4 == 0 && MANY_LB / 4 + 1 <= sizeof(W.w) / 4, "the words past the lane's bytes");
#pragma unroll
    for (int p = 0; p < (int)MANY_LB; p++) {
        const uint32_t w = win(p);
        uint32_t hit = 0;
        if ((FS & 0x30) == 0x10) {   // offset: the low 5 bits of byte 3
            hit |= __builtin_amdgcn_ubfe(pf[dict_pword(w)], low_byte_at(p + 3), 1);
        } else if (FS & 0x10) {      // offset: the hash's low 5 bits
            const uint32_t h = dict_ph(w);
            hit |= __builtin_amdgcn_ubfe(pf[dict_ph_word(h)], h, 1);
        }
        if (FS & 0x02)   // word: bits 5..15 of the window; bit: its low 5 bits
            hit |= __builtin_amdgcn_ubfe(f2[__builtin_amdgcn_ubfe(w, 5, 11)], w, 1);
        cand |= hit << p;
    }
    return lim >= MANY_LB ? cand : cand & ((1u << lim) - 1u);
}

// The wave tile's candidates: per part the lane's bits and its first index
// in the tile's list (position order: part, then lane, then position).
struct ManyCand {
    uint32_t q[MANY_H];
    uint32_t ex[MANY_H];
    uint32_t total;   // the tile's candidates (every lane)
};

// Filters wave tile wt part by part from the text words in W, loading each
// next part (and at the last part the first part of wave tile `next`, when
// there is one) into W before filtering, so the loads overlap the lookups and
// the probes that follow.
template <int FS>
__device__ __forceinline__ ManyCand many_candidates(const ManyArgs &a, uint32_t wt, uint32_t next, const ManyLane &m,
                                                    const uint32_t *pf, const uint32_t *small, ManyWords &W)
{
    ManyCand c;
    uint32_t run = 0;
#pragma unroll
    for (uint32_t h = 0; h < MANY_H; h++) {
        const ManyWords cur = W;
        if (h + 1 < MANY_H)
            W = many_load(a, wt, h + 1);
        else if (next < a.num_wtiles)
            W = many_load(a, next, 0);
        const uint32_t g0 = m.base + h * MANY_PART + m.lo;
        const uint32_t lim = g0 < a.n ? min(a.n - g0, MANY_LB) : 0u;
        c.q[h] = lim ? many_filter<FS>(a.d, pf, small, cur, lim) : 0u;
        const uint32_t nc = (uint32_t)__popc(c.q[h]);
        const uint32_t inc = wave_incl_add(nc);
        c.ex[h] = run + inc - nc;
        run += __shfl(inc, 63, 64);
    }
    c.total = run;
    return c;
}

// Candidates [c0, c0 + MANY_WLIST) of the wave tile into the wave's LDS list:
// wave-tile position and end bound (0: no match may start there).  Chunks are
// taken in order; each removes the bits it lists from `cand`.
__device__ __forceinline__ void many_list_chunk(const ManyArgs &a, const ManyLane &m, ManyCand &cand, uint32_t c0,
                                                uint16_t *l_pos, uint8_t *l_rem)
{
    if (a.ndocs == 0) {   // one document: every bound is the text end, nothing to load
        const uint32_t n = a.n;
#pragma unroll
        for (uint32_t h = 0; h < MANY_H; h++) {
            uint32_t at = cand.ex[h];
            uint32_t q = cand.q[h];
            while (q && at < c0 + MANY_WLIST) {
                const uint32_t lp = h * MANY_PART + m.lo + (uint32_t)__builtin_ctz(q);
                q &= q - 1;
                const uint32_t gp = m.base + lp;
                const uint32_t rm = gp < n ? n - gp : 0u;
                l_pos[at - c0] = (uint16_t)lp;
                l_rem[at - c0] = (uint8_t)(rm < 64u ? rm : 64u);
                at++;
            }
            cand.q[h] = q;
            cand.ex[h] = at;
        }
        return;
    }
    uint32_t di = m.di0, dlim = m.dlim0, dnext = m.dnext0;
#pragma unroll
    for (uint32_t h = 0; h < MANY_H; h++) {
        uint32_t at = cand.ex[h];
        uint32_t q = cand.q[h];
        while (q && at < c0 + MANY_WLIST) {
            const uint32_t lp = h * MANY_PART + m.lo + (uint32_t)__builtin_ctz(q);
            q &= q - 1;
            const uint32_t gp = m.base + lp;
            while (gp >= dnext) {
                di++;
                dlim = a.doc_ends[di];
                dnext = (di + 1 < a.ndocs) ? a.doc_starts[di + 1] : 0xFFFFFFFFu;
            }
            const uint32_t rm = gp < dlim ? dlim - gp : 0u;
            l_pos[at - c0] = (uint16_t)lp;
            l_rem[at - c0] = (uint8_t)(rm < 64u ? rm : 64u);
            at++;
        }
        // the bits listed are consumed: the next chunk starts at the first not listed
        cand.q[h] = q;
        cand.ex[h] = at;
    }
}


// Count pass: every wave takes wave tiles wt = wave + k * waves.
template <int THREADS, int FS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) void many_count_kernel(ManyArgs a)
{
    constexpr int NWV = THREADS / 64;
    __shared__ __align__(16) uint32_t s_pf[(FS & 0x10) ? MANY_PWORDS : 1];
    extern __shared__ __align__(16) uint32_t s_dyn[];
    __shared__ uint16_t s_pos[NWV][MANY_WLIST];
    __shared__ uint8_t s_rem[NWV][MANY_WLIST];
    constexpr uint32_t PW = (FS & 0x10) ? MANY_PWORDS : 0u;   // host layout: the prefix filter first
    if (FS & 0x10) {
        const uint4 *src = (const uint4 *)a.d.filter;
        for (uint32_t i = threadIdx.x; i < MANY_PWORDS / 4; i += THREADS)
            ((uint4 *)s_pf)[i] = src[i];
    }
    for (uint32_t i = PW + threadIdx.x; i < a.d.fwords; i += THREADS)
        s_dyn[i - PW] = a.d.filter[i];
    __syncthreads();
    const DictDev &d = a.d;
    const uint32_t *const pf = (FS & 0x10) ? s_pf : s_dyn;
    const uint32_t *const small = s_dyn - PW;   // filters 1, 2 at their word offsets
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint16_t *const l_pos = s_pos[wid];
    uint8_t *const l_rem = s_rem[wid];
    const uint32_t stride = gridDim.x * NWV;
    uint32_t wt = blockIdx.x * NWV + wid;
    ManyWords W;
    if (wt < a.num_wtiles)
        W = many_load(a, wt, 0);
    for (; wt < a.num_wtiles; wt += stride) {
        const ManyLane m = many_lane_setup(a, wt);
        ManyCand cand = many_candidates<FS>(a, wt, wt + stride, m, pf, small, W);
        const uint32_t total_c = cand.total;
        uint4 *const store = a.cands ? (uint4 *)a.cands + (size_t)wt * MANY_WCAP : nullptr;
        uint32_t kept = 0, tile_total = 0;
        for (uint32_t c0 = 0; c0 < total_c; c0 += MANY_WLIST) {
            const uint32_t n = min(MANY_WLIST, total_c - c0);
            many_list_chunk(a, m, cand, c0, l_pos, l_rem);
            wave_lds_sync();
            for (uint32_t i0 = 0; i0 < n; i0 += 64) {
                const uint32_t i = i0 + lane;
                uint32_t lp = 0, rm = 0, c = 0;
                uint32_t ids[MANY_KEEP] = {0, 0, 0, 0};
                if (i < n) {
                    lp = l_pos[i];
                    rm = l_rem[i];
                    if (rm)
                        c = probe_count_keep<(FS & 0x02) != 0, (FS & 0x02) != 0>(d, pf, m.g_text, lp, rm, ids);
                }
                tile_total += c;
                if (store) {
                    // keep the candidates with matches, in position order
                    const uint64_t has = __ballot(c != 0);
                    const uint32_t at = kept + (uint32_t)__popcll(has & ((1ull << lane) - 1ull));
                    if (c && at < MANY_WCAP)
                        store[at] = cand_pack(
                            lp | rm << CAND_POS_BITS | (c < CAND_RECOUNT ? c : CAND_RECOUNT) << CAND_CNT_SHIFT, ids);
                    kept += (uint32_t)__popcll(has);
                }
            }
            wave_lds_sync();   // the list is rewritten by the next chunk
        }
        const uint32_t tt = wave_incl_add(tile_total);
        const uint32_t tsum = __shfl(tt, 63, 64);
        if (lane == 0) {
            a.tile_count[wt] = tsum;
            a.tile_list[wt] = (store && kept <= MANY_WCAP) ? (uint64_t)kept : TILE_LIST_NONE;
        }
    }
}

// Stored candidate of the merged layout (16 bytes): x as above (position |
// bytes left | match count; CAND_RECOUNT: a run of more than 64 entries,
// probed again by the write pass); y = the first entry of the key's run, or
// CAND_SINGLE | the pattern id when the key has one pattern; z, w = which of
// the run's entries match (bit j: entry y + j), low word first.
constexpr uint32_t CAND_SINGLE = 1u << 31;
constexpr uint32_t MERGED_BITS = 64;   // run entries a stored candidate records
constexpr uint32_t MANY_STAGE = 512;   // write pass: matches a wave stages per round

// Count pass of the merged layout.  Per 64 candidates (one per lane): the
// key's slot; a single pattern is checked in the lane, and the runs' entries
// are checked 64 at a time over the whole wave (entry k of the concatenated
// runs on lane k mod 64, its candidate found from markers in LDS and a
// prefix maximum over the lanes), so a wave's work follows the sum of its
// candidates' runs, not their maximum.  A candidate collects its entries'
// results from the round's ballot.
template <int THREADS, int FS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu((FS & 0x12) == 0x12 ? 3 : 4))) void many_count_m_kernel(ManyArgs a)
{
    constexpr int NWV = THREADS / 64;
    constexpr bool SMALL = (FS & 0x02) != 0;
    __shared__ __align__(16) uint32_t s_pf[(FS & 0x10) ? MANY_PWORDS : 1];
    extern __shared__ __align__(16) uint32_t s_dyn[];
    __shared__ uint16_t s_pos[NWV][MANY_WLIST];
    __shared__ uint8_t s_rem[NWV][MANY_WLIST];
    __shared__ int32_t s_mark[NWV][64];
    constexpr uint32_t PW = (FS & 0x10) ? MANY_PWORDS : 0u;   // host layout: the prefix filter first
    if (FS & 0x10) {
        const uint4 *src = (const uint4 *)a.d.filter;
        for (uint32_t i = threadIdx.x; i < MANY_PWORDS / 4; i += THREADS)
            ((uint4 *)s_pf)[i] = src[i];
    }
    for (uint32_t i = PW + threadIdx.x; i < a.d.fwords; i += THREADS)
        s_dyn[i - PW] = a.d.filter[i];
    __syncthreads();
    const DictDev &d = a.d;
    const uint32_t *const pf = (FS & 0x10) ? s_pf : s_dyn;
    const uint32_t *const small = s_dyn - PW;   // filters 1, 2 at their word offsets
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    uint16_t *const l_pos = s_pos[wid];
    uint8_t *const l_rem = s_rem[wid];
    int32_t *const mark = s_mark[wid];
    const uint32_t stride = gridDim.x * NWV;
    uint32_t wt = blockIdx.x * NWV + wid;
    ManyWords W;
    if (wt < a.num_wtiles)
        W = many_load(a, wt, 0);
    for (; wt < a.num_wtiles; wt += stride) {
        const ManyLane m = many_lane_setup(a, wt);
        ManyCand cand = many_candidates<FS>(a, wt, wt + stride, m, pf, small, W);
        const uint32_t total_c = cand.total;
        uint4 *const store = a.cands ? (uint4 *)a.cands + (size_t)wt * MANY_WCAP : nullptr;
        uint32_t kept = 0, tile_total = 0;
        for (uint32_t c0 = 0; c0 < total_c; c0 += MANY_WLIST) {
            const uint32_t n = min(MANY_WLIST, total_c - c0);
            many_list_chunk(a, m, cand, c0, l_pos, l_rem);
            wave_lds_sync();
            for (uint32_t i0 = 0; i0 < n; i0 += 64) {
                const uint32_t i = i0 + lane;
                uint32_t lp = 0, rm = 0, T0 = 0, T1 = 0, T2 = 0;
                uint4 f = make_uint4(0, 0, 0, 0);
                if (i < n) {
                    lp = l_pos[i];
                    rm = l_rem[i];
                }
                if (rm) {
                    text_words(m.g_text, lp, T0, T1, T2);
                    f = merged_slot<true, SMALL, SMALL>(d, pf, small + PW, T0, rm);
                }
                const bool valid = (f.y & DK_VALID) != 0;
                const uint32_t slen = (f.y >> DK_LEN_SHIFT) & 0x7Fu;
                const uint32_t nrun = valid && !slen ? f.w : 0u;
                uint32_t c = valid && slen && pattern_here(d, m.g_text, lp, T1, T2, f.y, f.z, f.w, rm) ? 1u : 0u;
                uint64_t bits = c;
                // the runs' entries: entry k of the wave's concatenated runs
                // (this lane's at [P, P + nrun)) is checked on lane k mod 64
                const uint32_t incl = wave_incl_add(nrun);
                const uint32_t P = incl - nrun, M = __shfl(incl, 63, 64);
                int32_t carry = -1;
                for (uint32_t r0 = 0; r0 < M; r0 += 64) {
                    mark[lane] = -1;
                    wave_lds_sync();
                    if (nrun && P >= r0 && P < r0 + 64)
                        mark[P - r0] = (int32_t)lane;
                    wave_lds_sync();
                    const int32_t o = max(wave_incl_max(mark[lane]), carry);
                    carry = __shfl(o, 63, 64);
                    const int ol = o < 0 ? 0 : o;
                    const uint32_t oP = __shfl(P, ol, 64), ob = __shfl(f.z, ol, 64);
                    const uint32_t oT1 = __shfl(T1, ol, 64), oT2 = __shfl(T2, ol, 64);
                    const uint32_t olr = __shfl(lp | rm << 16, ol, 64);
                    const uint32_t k = r0 + lane;
                    bool ok = false;
                    if (k < M) {
                        const uint4 e = d.ents[ob + (k - oP)];
                        ok = pattern_here(d, m.g_text, olr & 0xFFFFu, oT1, oT2, e.x, e.y, e.z, olr >> 16);
                    }
                    const uint64_t okm = __ballot(ok);
                    // this lane's entries in the round: lanes [at, at + nrun)
                    const int32_t at = (int32_t)P - (int32_t)r0;
                    if (nrun && at < 64 && at + (int32_t)nrun > 0) {
                        const int32_t lo = at > 0 ? at : 0, hi = min(at + (int32_t)nrun, 64);
                        const uint64_t rmask = (hi >= 64 ? ~0ull : (1ull << hi) - 1ull) & ~((1ull << lo) - 1ull);
                        const uint64_t mine = okm & rmask;
                        c += (uint32_t)__popcll(mine);
                        if (at >= 0)
                            bits |= mine >> at;
                        else if (at > -64)
                            bits |= mine << (-at);
                    }
                    wave_lds_sync();   // mark[] is rewritten by the next round
                }
                tile_total += c;
                if (store) {
                    // keep the candidates with matches, in position order
                    const uint64_t has = __ballot(c != 0);
                    const uint32_t at = kept + (uint32_t)__popcll(has & below);
                    if (c && at < MANY_WCAP) {
                        const uint32_t cc = nrun > MERGED_BITS ? CAND_RECOUNT : c;
                        store[at] = make_uint4(lp | rm << CAND_POS_BITS | cc << CAND_CNT_SHIFT,
                                               slen ? (CAND_SINGLE | (f.y & 0x3FFFFFu)) : f.z, (uint32_t)bits,
                                               (uint32_t)(bits >> 32));
                    }
                    kept += (uint32_t)__popcll(has);
                }
            }
            wave_lds_sync();   // the list is rewritten by the next chunk
        }
        const uint32_t tt = wave_incl_add(tile_total);
        const uint32_t tsum = __shfl(tt, 63, 64);
        if (lane == 0) {
            a.tile_count[wt] = tsum;
            a.tile_list[wt] = (store && kept <= MANY_WCAP) ? (uint64_t)kept : TILE_LIST_NONE;
        }
    }
}

// A wave tile whose candidates were not stored: per part, each lane filters
// its positions again (filter read from global memory) and walks its own
// candidates, counting, then writing from out0 on.
template <int FS, bool MERGED>
__device__ __forceinline__ void many_write_unstored(const ManyArgs &a, uint32_t wt, uint64_t out0)
{
    const DictDev &d = a.d;
    const ManyLane m = many_lane_setup(a, wt);
    uint32_t di = m.di0, dlim = m.dlim0, dnext = m.dnext0;
    for (uint32_t h = 0; h < MANY_H; h++) {
        const uint32_t g0 = m.base + h * MANY_PART + m.lo;
        const uint32_t lim = g0 < a.n ? min(a.n - g0, MANY_LB) : 0u;
        const uint32_t cand = lim ? many_filter<FS>(d, d.filter, d.filter, many_load(a, wt, h), lim) : 0u;
        uint32_t own = 0;
        for (int pass = 0; pass < 2; pass++) {
            uint64_t r = 0;
            if (pass) {
                const uint32_t inc = wave_incl_add(own);
                r = out0 + (inc - own);
                out0 += __shfl(inc, 63, 64);
            }
            uint32_t pdi = di, pdlim = dlim, pdnext = dnext;
            uint32_t q = cand;
            while (q) {
                const uint32_t lp = h * MANY_PART + m.lo + (uint32_t)__builtin_ctz(q);
                q &= q - 1;
                const uint32_t gp = m.base + lp;
                while (gp >= pdnext) {
                    pdi++;
                    pdlim = a.doc_ends[pdi];
                    pdnext = (pdi + 1 < a.ndocs) ? a.doc_starts[pdi + 1] : 0xFFFFFFFFu;
                }
                const uint32_t rm = gp < pdlim ? min(pdlim - gp, 64u) : 0u;
                if (!rm)
                    continue;
                if (!pass)
                    own += probe_any<MERGED, (FS & 0x02) != 0>(d, m.g_text, lp, gp, rm, false, nullptr, nullptr, 0, 0);
                else if (r < a.cap)
                    r += probe_any<MERGED, (FS & 0x02) != 0>(d, m.g_text, lp, gp, rm, true, a.out_off, a.out_id, r,
                                                             a.cap);
            }
            if (pass) {
                di = pdi;
                dlim = pdlim;
                dnext = pdnext;
            }
        }
    }
}

// Write pass.  A wave takes job_tiles (at most MANY_JOB_MAX) consecutive wave
// tiles (a job; lane j holds tile j's stored count and output offset) and
// walks the stored candidates of all of them as one list, 64 per round, one
// candidate per lane.  A candidate with at most four matches writes the ids
// stored with it; one with more is probed again.  Lane k
// of a round finds its tile as the last tile starting at or before it
// (markers in LDS, a prefix maximum over the lanes); its output offset is
// its tile's offset plus the counts of the tile's earlier candidates (a
// running sum over the list minus that sum at the tile's first candidate).
// Tiles whose candidates were not stored are redone afterwards, one at a
// time.
template <int THREADS, int FS>
__global__ __launch_bounds__(THREADS) void many_write_kernel(ManyArgs a)
{
    constexpr int NWV = THREADS / 64;
    __shared__ uint32_t s_mark[NWV][64];
    __shared__ uint64_t s_first[NWV][64];
    __shared__ uint4 s_queue[NWV][128];   // candidates to probe again: {position, end bound, output index}
    const DictDev &d = a.d;
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t *const mark = s_mark[wid];
    uint64_t *const first = s_first[wid];
    uint4 *const queue = s_queue[wid];
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t qn = 0;   // queued candidates (every lane)
    // probes the first `count` queued candidates, one per lane
    auto probe_queued = [&](uint32_t count) {
        if (lane < count) {
            const uint4 q = queue[lane];
            probe_position<false, false, (FS & 0x02) != 0>(d, nullptr, a.text, q.x, q.x, q.y, true, a.out_off, a.out_id,
                                  (uint64_t)q.w << 32 | q.z, a.cap);
        }
    };
    const uint32_t JT = a.job_tiles;
    const uint32_t jobs = (a.num_wtiles + JT - 1) / JT;
    for (uint32_t job = blockIdx.x * NWV + wid; job < jobs; job += gridDim.x * NWV) {
        const uint32_t wt = job * JT + lane;
        const bool valid = lane < JT && wt < a.num_wtiles;
        const uint64_t st = valid ? a.tile_list[wt] : 0ull;
        const uint64_t off = valid ? a.tile_off[wt] : 0ull;
        const uint32_t n = st == TILE_LIST_NONE ? 0u : (uint32_t)st;
        const uint32_t inc = wave_incl_add(n);
        const uint32_t ex = inc - n;
        const uint32_t total = __shfl(inc, 63, 64);
        int32_t carry = -1;
        uint64_t run = 0;   // counts of the list's candidates before this round
        for (uint32_t r0 = 0; r0 < total; r0 += 64) {
            mark[lane] = 0xFFFFFFFFu;
            wave_lds_sync();
            if (n && ex >= r0 && ex < r0 + 64)
                mark[ex - r0] = lane;
            wave_lds_sync();
            const int32_t j = max(wave_incl_max((int32_t)mark[lane]), carry);
            carry = __shfl(j, 63, 64);
            const uint32_t k = r0 + lane;
            const bool act = k < total;
            const uint32_t tj = (uint32_t)(j < 0 ? 0 : j);
            const uint32_t i = k - (uint32_t)__shfl((int)ex, (int)tj, 64);
            const uint64_t off_j = __shfl(off, (int)tj, 64);
            const uint32_t wt_j = job * JT + tj;
            const uint32_t base = wt_j * MANY_WT;
            const uint8_t *const g_text = a.text + base;
            const uint4 ev = act ? ((const uint4 *)a.cands)[(size_t)wt_j * MANY_WCAP + i] : make_uint4(0, 0, 0, 0);
            const uint32_t e = ev.x;
            const uint32_t lp = e & (MANY_WT - 1u), rm = (e >> CAND_POS_BITS) & 0x7Fu;
            uint32_t c = e >> CAND_CNT_SHIFT;
            if (c == CAND_RECOUNT)
                c = probe_position<false, false, (FS & 0x02) != 0>(d, nullptr, g_text, lp, (uint64_t)base + lp, rm,
                                                                    false, nullptr, nullptr, 0, 0);
            const uint32_t cinc = wave_incl_add(c);
            const uint64_t sk = run + (cinc - c);   // counts of the list before candidate k
            if (act && i == 0)
                first[tj] = sk;
            wave_lds_sync();
            const uint64_t r = off_j + (sk - (act ? first[tj] : sk));
            // more matches than stored ids: queued, and probed 64 at a time so
            // that rounds without them never wait for a probe
            const bool again = act && c > MANY_KEEP && r < a.cap;
            const uint64_t qb = __ballot(again);
            if (again)
                queue[qn + (uint32_t)__popcll(qb & below)] = make_uint4(base + lp, rm, (uint32_t)r, (uint32_t)(r >> 32));
            qn += (uint32_t)__popcll(qb);
            if (act && c && c <= MANY_KEEP) {   // the stored ids
                const uint64_t gp = (uint64_t)base + lp;
                uint32_t id4[MANY_KEEP];
                cand_ids(ev, id4);
#pragma unroll
                for (uint32_t q = 0; q < MANY_KEEP; q++)
                    if (q < c && r + q < a.cap) {
                        if (a.out_off) a.out_off[r + q] = gp;
                        if (a.out_id) a.out_id[r + q] = id4[q];
                    }
            }
            run += __shfl(cinc, 63, 64);
            wave_lds_sync();   // mark[] is rewritten by the next round
            if (qn >= 64) {
                probe_queued(64);
                const uint4 rest = lane < qn - 64 ? queue[64 + lane] : make_uint4(0, 0, 0, 0);
                wave_lds_sync();
                if (lane < qn - 64)
                    queue[lane] = rest;
                qn -= 64;
                wave_lds_sync();
            }
        }
        if (qn) {
            probe_queued(qn);
            qn = 0;
            wave_lds_sync();
        }
        uint64_t redo = __ballot(valid && st == TILE_LIST_NONE);
        while (redo) {
            const uint32_t j = (uint32_t)__builtin_ctzll(redo);
            redo &= redo - 1;
            many_write_unstored<FS, false>(a, job * JT + j, __shfl(off, (int)j, 64));
        }
    }
}

// Write pass of the merged layout: as many_write_kernel, but a stored
// candidate names its matches (the single pattern's id, or its run's first
// entry and which entries match), so its ids are read from the entries' id
// array (4 bytes per entry, which L2 keeps better than the 16-byte entries),
// in id order, and nothing is probed again except runs longer than 64
// entries.
template <int THREADS, int FS>
__global__ __launch_bounds__(THREADS) void many_write_m_kernel(ManyArgs a)
{
    constexpr int NWV = THREADS / 64;
    constexpr bool SMALL = (FS & 0x02) != 0;
    __shared__ uint32_t s_mark[NWV][64];
    __shared__ uint64_t s_first[NWV][64];
    __shared__ uint4 s_queue[NWV][128];   // candidates to probe again: {position, end bound, output index}
    __shared__ uint32_t s_sgp[NWV][MANY_STAGE], s_sid[NWV][MANY_STAGE];   // a round's matches, staged
    const DictDev &d = a.d;
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    uint32_t *const mark = s_mark[wid];
    uint64_t *const first = s_first[wid];
    uint4 *const queue = s_queue[wid];
    uint32_t *const sgp = s_sgp[wid], *const sid = s_sid[wid];
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t qn = 0;   // queued candidates (every lane)
    auto probe_queued = [&](uint32_t count) {
        if (lane < count) {
            const uint4 q = queue[lane];
            probe_merged<SMALL>(d, a.text, q.x, q.x, q.y, true, a.out_off, a.out_id, (uint64_t)q.w << 32 | q.z, a.cap);
        }
    };
    const uint32_t JT = a.job_tiles;
    const uint32_t jobs = (a.num_wtiles + JT - 1) / JT;
    for (uint32_t job = blockIdx.x * NWV + wid; job < jobs; job += gridDim.x * NWV) {
        const uint32_t wt = job * JT + lane;
        const bool valid = lane < JT && wt < a.num_wtiles;
        const uint64_t st = valid ? a.tile_list[wt] : 0ull;
        const uint64_t off = valid ? a.tile_off[wt] : 0ull;
        const uint32_t n = st == TILE_LIST_NONE ? 0u : (uint32_t)st;
        const uint32_t inc = wave_incl_add(n);
        const uint32_t ex = inc - n;
        const uint32_t total = __shfl(inc, 63, 64);
        int32_t carry = -1;
        // round rr's list entry on this lane: its tile (within the job) and
        // index in the tile's stored candidates, and the stored record
        auto locate = [&](uint32_t rr, uint32_t &tj_out, uint32_t &i_out) {
            mark[lane] = 0xFFFFFFFFu;
            wave_lds_sync();
            if (n && ex >= rr && ex < rr + 64)
                mark[ex - rr] = lane;
            wave_lds_sync();
            const int32_t j = max(wave_incl_max((int32_t)mark[lane]), carry);
            carry = __shfl(j, 63, 64);
            tj_out = (uint32_t)(j < 0 ? 0 : j);
            i_out = rr + lane - (uint32_t)__shfl((int)ex, (int)tj_out, 64);
            wave_lds_sync();   // mark[] is rewritten by the next call
        };
        auto record = [&](uint32_t rr, uint32_t tj_, uint32_t i_) -> uint4 {
            return rr + lane < total ? ((const uint4 *)a.cands)[(size_t)(job * JT + tj_) * MANY_WCAP + i_]
                                     : make_uint4(0, 0, 0, 0);
        };
        uint32_t tj = 0, i = 0;
        if (total)
            locate(0, tj, i);
        uint4 ev = record(0, tj, i);
        uint64_t run = 0;   // counts of the list's candidates before this round
        for (uint32_t r0 = 0; r0 < total; r0 += 64) {
            // the next round's records load while this round is written
            uint32_t tjn = 0, in = 0;
            if (r0 + 64 < total)
                locate(r0 + 64, tjn, in);
            const uint4 evn = record(r0 + 64, tjn, in);
            const uint32_t k = r0 + lane;
            const bool act = k < total;
            const uint64_t off_j = __shfl(off, (int)tj, 64);
            const uint32_t wt_j = job * JT + tj;
            const uint32_t base = wt_j * MANY_WT;
            const uint8_t *const g_text = a.text + base;
            const uint32_t e = ev.x;
            const uint32_t lp = e & (MANY_WT - 1u), rm = (e >> CAND_POS_BITS) & 0x7Fu;
            uint32_t c = e >> CAND_CNT_SHIFT;
            const bool recount = c == CAND_RECOUNT;
            if (recount)
                c = probe_merged<SMALL>(d, g_text, lp, (uint64_t)base + lp, rm, false, nullptr, nullptr, 0, 0);
            const uint32_t cinc = wave_incl_add(c);
            const uint64_t sk = run + (cinc - c);   // counts of the list before candidate k
            if (act && i == 0)
                first[tj] = sk;
            wave_lds_sync();
            const uint64_t r = off_j + (sk - (act ? first[tj] : sk));
            // runs longer than the stored bits: queued, and probed 64 at a time
            const bool again = act && recount && c && r < a.cap;
            const uint64_t qb = __ballot(again);
            if (again)
                queue[qn + (uint32_t)__popcll(qb & below)] = make_uint4(base + lp, rm, (uint32_t)r, (uint32_t)(r >> 32));
            qn += (uint32_t)__popcll(qb);
            // The round's outputs [R0, R0 + span) go through LDS and out in
            // coalesced stores when its candidates fill them without a gap (no
            // queued run, no unstored wave tile between) and they fit the
            // stage; otherwise each lane writes its own.
            const uint64_t R0 = __shfl(r, 0, 64);   // lane 0 is active in every round
            const uint32_t lastl = min(63u, total - 1u - r0);
            const uint64_t span = __shfl(r + c, (int)lastl, 64) - R0;
            // A round whose candidates each have one match already writes
            // consecutive outputs lane by lane, so it skips the stage.
            const bool staged = __ballot(act && recount) == 0 && __ballot(act && c != 1) != 0 &&
                                span == __shfl(cinc, 63, 64) && span <= MANY_STAGE;
            if (act && c && !recount) {
                const uint32_t gp = base + lp;
                const uint32_t s0 = (uint32_t)(r - R0);
                if (ev.y & CAND_SINGLE) {
                    if (staged) {
                        sgp[s0] = gp;
                        sid[s0] = ev.y & 0x3FFFFFu;
                    } else if (r < a.cap) {
                        if (a.out_off) a.out_off[r] = gp;
                        if (a.out_id) a.out_id[r] = ev.y & 0x3FFFFFu;
                    }
                } else {
                    // the matching entries' ids, four loads in flight at a time
                    uint64_t bm = (uint64_t)ev.w << 32 | ev.z;
                    for (uint32_t q = 0; bm; q += 4) {
                        uint32_t b[4], nb = 0;
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            b[t] = bm ? (uint32_t)__builtin_ctzll(bm) : b[0];
                            nb += bm ? 1u : 0u;
                            bm &= bm ? bm - 1 : bm;
                        }
                        uint32_t id[4];
#pragma unroll
                        for (int t = 0; t < 4; t++)
                            id[t] = d.ent_ids[ev.y + b[t]];
#pragma unroll
                        for (uint32_t t = 0; t < 4; t++) {
                            if (t >= nb)
                                break;
                            if (staged) {
                                sgp[s0 + q + t] = gp;
                                sid[s0 + q + t] = id[t];
                            } else if (r + q + t < a.cap) {
                                if (a.out_off) a.out_off[r + q + t] = gp;
                                if (a.out_id) a.out_id[r + q + t] = id[t];
                            }
                        }
                    }
                }
            }
            if (staged) {
                wave_lds_sync();
                for (uint32_t j = lane; j < (uint32_t)span && R0 + j < a.cap; j += 64) {
                    if (a.out_off) a.out_off[R0 + j] = sgp[j];
                    if (a.out_id) a.out_id[R0 + j] = sid[j];
                }
            }
            run += __shfl(cinc, 63, 64);
            tj = tjn;
            i = in;
            ev = evn;
            wave_lds_sync();   // first[] and the stage are rewritten by the next round
            if (qn >= 64) {
                probe_queued(64);
                const uint4 rest = lane < qn - 64 ? queue[64 + lane] : make_uint4(0, 0, 0, 0);
                wave_lds_sync();
                if (lane < qn - 64)
                    queue[lane] = rest;
                qn -= 64;
                wave_lds_sync();
            }
        }
        if (qn) {
            probe_queued(qn);
            qn = 0;
            wave_lds_sync();
        }
        uint64_t redo = __ballot(valid && st == TILE_LIST_NONE);
        while (redo) {
            const uint32_t j = (uint32_t)__builtin_ctzll(redo);
            redo &= redo - 1;
            many_write_unstored<FS, true>(a, job * JT + j, __shfl(off, (int)j, 64));
        }
    }
}

// Exclusive scan of tile_count into tile_off: one workgroup per chunk of
// ST * K tiles.  Each publishes its chunk's total, then waits for the totals
// of the chunks before it (lower workgroups, dispatched first, which wait on
// nothing later) and adds them up; the last chunk's workgroup writes the
// grand total to meta[0].  Chunk totals live in `sums` (zero at launch).
template <int ST>
__global__ __launch_bounds__(ST) void many_scan_kernel(ManyArgs a)
{
    uint64_t *const sums = a.scan_sums;
    constexpr int K = 8;
    constexpr int CH = ST * K;
    static_assert(CH == MANY_SCAN_CHUNK, "scan chunk");
    constexpr int NWS = ST / 64;
    __shared__ uint64_t s_wt[NWS];
    __shared__ uint64_t s_carry;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t T = a.num_wtiles;
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
                if (spins >= LB_SPIN_LIMIT)
                    atomicOr(a.lb_err, OLE_DICT_SCAN);
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

// Dynamic LDS of the count pass: the filters after the prefix filter.
size_t ol_many_lds_bytes(const DictDev &d)
{
    return (size_t)(d.fwords - (d.fbits[4] ? MANY_PWORDS : 0u)) * 4;
}

template <bool WRITE, int FS, bool MERGED>
static hipError_t launch_many_fs(const ManyArgs &a, hipStream_t s)
{
    const size_t lds = WRITE ? 0 : ol_many_lds_bytes(a.d);
    const void *f = MERGED ? (WRITE ? (const void *)many_write_m_kernel<MANY_THREADS, FS>
                                    : (const void *)many_count_m_kernel<MANY_THREADS, FS>)
                           : (WRITE ? (const void *)many_write_kernel<MANY_THREADS, FS>
                                    : (const void *)many_count_kernel<MANY_THREADS, FS>);
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    const uint32_t cap = max_grid(f, MANY_THREADS, lds);
    if (!cap)
        return hipErrorInvalidValue;
    // persistent waves: workgroups beyond the co-resident ones would only
    // copy the filter again
    constexpr uint32_t NWV = MANY_THREADS / 64;
    if (WRITE) {
        // jobs of up to MANY_JOB_MAX wave tiles, fewer when the text is too
        // small to give every resident wave one
        ManyArgs w = a;
        const uint32_t per_wave = a.num_wtiles / (cap * NWV);
        w.job_tiles = per_wave < 1 ? 1 : per_wave > MANY_JOB_MAX ? MANY_JOB_MAX : per_wave;
        const uint32_t jobs = (a.num_wtiles + w.job_tiles - 1) / w.job_tiles;
        const uint32_t need = (jobs + NWV - 1) / NWV;
        if (MERGED)
            hipLaunchKernelGGL((many_write_m_kernel<MANY_THREADS, FS>), dim3(need < cap ? need : cap),
                               dim3(MANY_THREADS), 0, s, w);
        else
            hipLaunchKernelGGL((many_write_kernel<MANY_THREADS, FS>), dim3(need < cap ? need : cap),
                               dim3(MANY_THREADS), 0, s, w);
        return hipGetLastError();
    }
    const uint32_t need = (a.num_wtiles + NWV - 1) / NWV;
    if (MERGED)
        hipLaunchKernelGGL((many_count_m_kernel<MANY_THREADS, FS>), dim3(need < cap ? need : cap), dim3(MANY_THREADS),
                           lds, s, a);
    else
        hipLaunchKernelGGL((many_count_kernel<MANY_THREADS, FS>), dim3(need < cap ? need : cap), dim3(MANY_THREADS),
                           lds, s, a);
    return hipGetLastError();
}

// The kernels for the filters the dictionary has: the prefix filter (classes
// 3..5, the usual case), the 1/2-byte filters, or both.
template <bool WRITE>
static hipError_t launch_many(const ManyArgs &a, hipStream_t s)
{
    const int fs = (a.d.fbits[4] ? (a.d.pf_hashed ? 0x30 : 0x10) : 0) | (a.d.fbits[2] ? 0x02 : 0);
    const bool m = a.d.merged != 0;
    switch (fs) {
    case 0x10: return m ? launch_many_fs<WRITE, 0x10, true>(a, s) : launch_many_fs<WRITE, 0x10, false>(a, s);
    case 0x12: return m ? launch_many_fs<WRITE, 0x12, true>(a, s) : launch_many_fs<WRITE, 0x12, false>(a, s);
    case 0x30: return m ? launch_many_fs<WRITE, 0x30, true>(a, s) : launch_many_fs<WRITE, 0x30, false>(a, s);
    case 0x32: return m ? launch_many_fs<WRITE, 0x32, true>(a, s) : launch_many_fs<WRITE, 0x32, false>(a, s);
    case 0x02: return m ? launch_many_fs<WRITE, 0x02, true>(a, s) : launch_many_fs<WRITE, 0x02, false>(a, s);
    default: return hipErrorInvalidValue;   // a dictionary has at least one filter
    }
}

// Counts and their scan (total in meta[0]); with outputs also the writes.
// Without outputs but with a candidate store, the candidates are stored for a
// later ol_launch_many_write with the same arguments plus the outputs.
hipError_t ol_launch_many(const ManyArgs &a, hipStream_t s)
{
    if (a.num_wtiles == 0)
        return hipSuccess;
    if (a.num_wtiles != a.num_tiles * MANY_WPT)
        return hipErrorInvalidValue;
    hipError_t e = launch_many<false>(a, s);
    if (e != hipSuccess)
        return e;
    const uint32_t chunks = (a.num_wtiles + MANY_SCAN_CHUNK - 1) / MANY_SCAN_CHUNK;
    hipLaunchKernelGGL((many_scan_kernel<1024>), dim3(chunks), dim3(1024), 0, s, a);
    if ((e = hipGetLastError()) != hipSuccess || (!a.out_off && !a.out_id))
        return e;
    return launch_many<true>(a, s);
}

hipError_t ol_launch_many_write(const ManyArgs &a, hipStream_t s)
{
    if (a.num_wtiles == 0 || (!a.out_off && !a.out_id))
        return hipSuccess;
    return launch_many<true>(a, s);
}
