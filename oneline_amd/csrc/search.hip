// search.hip -- single-literal and dictionary search with ordered output.
//
// Both kernels share one shape: a workgroup claims the next text tile, stages
// it (plus a 80-byte halo for patterns running past the tile end) in LDS,
// finds the matches that START in each lane's 64 bytes, counts them, resolves
// the tile's output position with a decoupled look-back, and writes its matches
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

constexpr int HALO_AFTER = 80;

template <int THREADS, int BPT>
__device__ __forceinline__ void load_tile(uint8_t *s_text, const uint8_t *text, uint32_t base)
{
    const uint4 *src = (const uint4 *)(text + base);
    uint4 *dst = (uint4 *)s_text;
#pragma unroll
    for (int k = 0; k < BPT / 16; k++)
        dst[k * THREADS + threadIdx.x] = src[k * THREADS + threadIdx.x];
    if (threadIdx.x < HALO_AFTER / 16)
        dst[(THREADS * BPT) / 16 + threadIdx.x] = src[(THREADS * BPT) / 16 + threadIdx.x];
}

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

// Tile prefix via look-back; wave 0 computes it, everyone reads it.
__device__ __forceinline__ uint64_t tile_prefix(uint64_t *status, uint32_t t, uint64_t total,
                                                uint64_t *s_prefix, uint32_t *err)
{
    if ((threadIdx.x >> 6) == 0) {
        const int lane = threadIdx.x & 63;
        uint64_t prefix = 0;
        if (t == 0) {
            if (lane == 0)
                st_relaxed_u64(status, (LB_FLAG_INCL << LB_SHIFT) | total);
        } else {
            if (lane == 0)
                st_relaxed_u64(status + (size_t)t * LB_STRIDE, (LB_FLAG_AGG << LB_SHIFT) | total);
            prefix = lookback_sum(status, t, err);
            if (lane == 0)
                st_relaxed_u64(status + (size_t)t * LB_STRIDE, (LB_FLAG_INCL << LB_SHIFT) | (prefix + total));
        }
        if (lane == 0)
            *s_prefix = prefix;
    }
    __syncthreads();
    return *s_prefix;
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
// Persistent: tile t = blockIdx + k * gridDim (all workgroups co-resident, each
// walking its tiles in increasing order, so look-backs always make progress).
template <int THREADS, int BPT>
__global__ __launch_bounds__(THREADS) void search_kernel(SearchArgs a)
{
    constexpr int TILE = THREADS * BPT;
    static_assert(BPT == 64, "64-bit lane masks");
    __shared__ __align__(16) uint8_t s_text[TILE + HALO_AFTER];
    __shared__ uint32_t s_w[THREADS / 64];
    __shared__ uint64_t s_prefix;

    const int tid = threadIdx.x;
    for (uint32_t t = blockIdx.x; t < a.num_tiles; t += gridDim.x) {
        const uint32_t base = t * TILE;
        load_tile<THREADS, BPT>(s_text, a.text, base);
        __syncthreads();

        const uint32_t lo = tid * BPT;
        const uint32_t g0 = base + lo;
        // positions p with g0 + p + m <= n
        uint32_t lim = 0;
        if ((uint64_t)g0 + a.m <= a.n)
            lim = a.n - a.m - g0 + 1;
        if (lim > BPT)
            lim = BPT;
        uint64_t match = 0;
        if (lim) {
            const uint32_t *wsrc = (const uint32_t *)(s_text + lo);
            uint32_t w[BPT / 4 + 1];
#pragma unroll
            for (int i = 0; i < BPT / 4 + 1; i++)
                w[i] = wsrc[i];
            uint64_t cand = 0;
#pragma unroll
            for (int p = 0; p < BPT; p++) {
                uint32_t win = (p & 3) ? __builtin_amdgcn_alignbyte(w[(p >> 2) + 1], w[p >> 2], p & 3)
                                       : w[p >> 2];
                cand |= (uint64_t)((win & a.win_mask) == a.win) << p;
            }
            if (lim < BPT)
                cand &= (1ull << lim) - 1;
            while (cand) {
                int p = __builtin_ctzll(cand);
                cand &= cand - 1;
                bool ok = true;
                for (uint32_t i = 4; i < a.m; i++) {
                    if (s_text[lo + p + i] != a.pat[i]) {
                        ok = false;
                        break;
                    }
                }
                if (ok)
                    match |= 1ull << p;
            }
        }
        uint32_t total;
        const uint32_t excl = block_excl_sum<THREADS>((uint32_t)__popcll(match), s_w, total);
        const uint64_t prefix = tile_prefix(a.status, t, total, &s_prefix, a.lb_err);
        if (tid == 0 && t == a.num_tiles - 1)
            a.meta[0] = prefix + total;
        uint64_t r = prefix + excl;
        while (match) {
            int p = __builtin_ctzll(match);
            match &= match - 1;
            if (r < a.cap) {
                if (a.out64) a.out64[r] = g0 + p;
                if (a.out32) a.out32[r] = g0 + p;
            }
            r++;
        }
        __syncthreads();   // LDS is reused by the next tile
    }
}

hipError_t ol_launch_search(const SearchArgs &a, hipStream_t s)
{
    static uint32_t cap = 0;   // co-resident limit (same device for the process)
    const void *f = (const void *)search_kernel<SEARCH_THREADS, SEARCH_BPT>;
    if (!cap)
        cap = max_grid(f, SEARCH_THREADS, 0);
    if (!cap)
        return hipErrorInvalidValue;
    uint32_t grid = a.num_tiles < cap ? a.num_tiles : cap;
    SearchArgs args = a;
    void *params[] = {&args};
    return hipLaunchCooperativeKernel(f, dim3(grid), dim3(SEARCH_THREADS), params, 0, s);
}

// ---- dictionary -----------------------------------------------------------------
// Per tile (256 lanes x 64 positions, staged in LDS with 128 bytes after it):
//   1. filter: each lane tests its 64 positions against the LDS bitmaps of the
//      dictionary's key classes (straight-line code, 4-byte windows from
//      registers) and gets a 64-bit candidate mask;
//   2. the candidates are compacted into an LDS list in position order;
//   3. count: thread i probes candidates i, i + 256, ...: the class tables
//      give, per class, the run of patterns sharing the key (sorted by id);
//      the runs are merged by id and each pattern's remaining bytes compared;
//   4. an exclusive scan of the per-candidate counts places every
//      candidate's matches, the look-back places the tile, and (when output
//      is wanted) the candidates are probed again and their matches written.
// A tile with more candidates than the list holds is handled per lane.

constexpr int MANY_HALO = 128;      // bytes staged after the tile (8-byte keys, 64-byte patterns)
constexpr int MANY_LCAP = 2048;     // candidates per tile taken through the list

// Matches starting at tile position lp (global gp) with at most `rem` bytes
// (1..64) before the text or document end; written from out[r] in pattern-id
// order (up to cap) when `emit`.  Returns their number.
__device__ __forceinline__ uint32_t probe_position(const DictDev &d, const uint8_t *s_text, uint32_t lp, uint64_t gp,
                                                   uint32_t rem, bool emit, uint64_t *out_off, uint32_t *out_id,
                                                   uint64_t r, uint64_t cap)
{
    const uint32_t *wp = (const uint32_t *)(s_text + (lp & ~3u));
    const uint32_t sh = lp & 3u;
    const uint32_t a0 = wp[0], a1 = wp[1], a2 = wp[2];
    const uint32_t w0 = sh ? __builtin_amdgcn_alignbyte(a1, a0, sh) : a0;
    const uint32_t w1 = sh ? __builtin_amdgcn_alignbyte(a2, a1, sh) : a1;
    uint32_t beg[5], end[5], kl[5];
    int nr = 0;
#pragma unroll
    for (uint32_t k = 1; k <= 5; k++) {
        const uint32_t klen = k <= 4 ? k : 8u;   // bytes the key covers
        if (!((d.class_mask >> k) & 1u) || klen > rem)
            continue;
        const uint32_t lo = k == 1 ? (w0 & 0xFFu) : k == 2 ? (w0 & 0xFFFFu) : k == 3 ? (w0 & 0xFFFFFFu) : w0;
        const uint32_t hi = k == 5 ? w1 : 0u;
        const uint32_t tb = d.tbits[k];
        const uint32_t tmask = (1u << tb) - 1u;
        const uint4 *tab = d.table[k];
        uint32_t h = dict_thash(lo, hi, tb);
        while (true) {
            const uint4 sl = tab[h];
            if (sl.z == 0)
                break;
            if (sl.x == lo && sl.w == hi) {
                beg[nr] = sl.y;
                end[nr] = sl.z;
                kl[nr] = k == 4 ? 4u : klen;
                nr++;
                break;
            }
            h = (h + 1) & tmask;
        }
    }
    if (nr == 0)
        return 0;
    // text word at tile position x (any alignment)
    auto tword = [&](uint32_t x) -> uint32_t {
        const uint32_t *q = (const uint32_t *)(s_text + (x & ~3u));
        const uint32_t sx = x & 3u;
        return sx ? __builtin_amdgcn_alignbyte(q[1], q[0], sx) : q[0];
    };
    auto tail_mask = [](uint32_t nb) -> uint32_t { return nb >= 4 ? ~0u : ((1u << (8 * nb)) - 1u); };
    uint4 cur[5];
#pragma unroll
    for (int i = 0; i < 5; i++)
        if (i < nr && beg[i] < end[i])
            cur[i] = d.ents[beg[i]];
    uint32_t count = 0;
    while (true) {
        int best = -1;
        uint32_t bid = 0xFFFFFFFFu;
#pragma unroll
        for (int i = 0; i < 5; i++) {
            if (i < nr && beg[i] < end[i] && cur[i].x < bid) {
                bid = cur[i].x;
                best = i;
            }
        }
        if (best < 0)
            break;
        uint4 en = cur[0];
        uint32_t kb = kl[0];
#pragma unroll
        for (int i = 1; i < 5; i++)
            if (i == best) {
                en = cur[i];
                kb = kl[i];
            }
        // advance the chosen run
#pragma unroll
        for (int i = 0; i < 5; i++)
            if (i == best && ++beg[i] < end[i])
                cur[i] = d.ents[beg[i]];
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

__device__ __forceinline__ uint32_t filter_bit(const uint32_t *f, uint32_t i) { return (f[i >> 5] >> (i & 31)) & 1u; }

template <int THREADS>
__global__ __launch_bounds__(THREADS) void many_kernel(ManyArgs a)
{
    constexpr int TILE = THREADS * 64;
    extern __shared__ __align__(16) uint32_t s_dyn[];
    __shared__ uint16_t s_cand[MANY_LCAP];
    __shared__ uint8_t s_rem[MANY_LCAP];
    __shared__ uint32_t s_cnt[MANY_LCAP];
    __shared__ uint32_t s_w[THREADS / 64];
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_tile;

    uint8_t *s_text = (uint8_t *)s_dyn;                      // TILE + MANY_HALO bytes
    uint32_t *s_filter = s_dyn + (TILE + MANY_HALO) / 4;    // fwords
    const DictDev &d = a.d;
    const int tid = threadIdx.x;
    for (uint32_t i = tid; i < d.fwords; i += THREADS)
        s_filter[i] = d.filter[i];
    const uint32_t cm = d.class_mask;
    const uint32_t *f1 = s_filter + d.foff[1], *f2 = s_filter + d.foff[2], *f3 = s_filter + d.foff[3],
                   *f4 = s_filter + d.foff[4];
    const uint32_t fb4 = d.fbits[4];

    // Tiles are claimed in order from a ticket: a tile's look-back then only
    // waits on tiles already claimed by running workgroups, whatever the
    // residency of the launch (its LDS depends on the dictionary).
    while (true) {
        if (tid == 0)
            s_tile = atomicAdd(a.tile_ctr, 1u);
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= a.num_tiles)
            break;
        const uint32_t base = t * TILE;
        {
            const uint4 *src = (const uint4 *)(a.text + base);
            uint4 *dst = (uint4 *)s_text;
#pragma unroll
            for (int k = 0; k < 4; k++)
                dst[k * THREADS + tid] = src[k * THREADS + tid];
            if (tid < MANY_HALO / 16)
                dst[TILE / 16 + tid] = src[TILE / 16 + tid];
        }
        __syncthreads();

        const uint32_t lo = tid * 64;
        const uint32_t g0 = base + lo;
        const uint32_t lim = (g0 < a.n) ? ((a.n - g0) < 64u ? a.n - g0 : 64u) : 0u;

        // match-end bound per position: the text end, or (batches) the end of
        // the position's document; positions between documents match nothing
        uint32_t dlim = a.n, dnext = 0xFFFFFFFFu, di = 0;
        if (a.ndocs) {
            di = a.tile_doc[t];   // then the last document starting at or before g0
            while (di + 1 < a.ndocs && a.doc_starts[di + 1] <= g0)
                di++;
            dlim = a.doc_ends[di];
            dnext = (di + 1 < a.ndocs) ? a.doc_starts[di + 1] : 0xFFFFFFFFu;
        }
        const uint32_t di0 = di, dlim0 = dlim, dnext0 = dnext;
        auto rem_at = [&](uint32_t gp) -> uint32_t {   // 0: no match may start at gp
            while (gp >= dnext) {
                di++;
                dlim = a.doc_ends[di];
                dnext = (di + 1 < a.ndocs) ? a.doc_starts[di + 1] : 0xFFFFFFFFu;
            }
            const uint32_t rm = gp < dlim ? dlim - gp : 0u;
            return rm < 64u ? rm : 64u;
        };

        // 1. filter
        uint64_t cand = 0;
        {
            uint32_t W[17];
            const uint32_t *wsrc = (const uint32_t *)(s_text + lo);
#pragma unroll
            for (int j = 0; j < 17; j++)
                W[j] = wsrc[j];
#pragma unroll
            for (int p = 0; p < 64; p++) {
                const uint32_t w = (p & 3) ? __builtin_amdgcn_alignbyte(W[(p >> 2) + 1], W[p >> 2], p & 3) : W[p >> 2];
                uint32_t hit = 0;
                if (cm & 0x02u) hit |= filter_bit(f1, w & 0xFFu);
                if (cm & 0x04u) hit |= filter_bit(f2, w & 0xFFFFu);
                if (cm & 0x08u) hit |= filter_bit(f3, dict_fhash3(w, 16));
                if (cm & 0x30u) hit |= filter_bit(f4, dict_fhash4(w, fb4));
                cand |= (uint64_t)hit << p;
            }
            if (lim < 64)
                cand &= lim ? ((1ull << lim) - 1ull) : 0ull;
        }
        // positions outside documents (batches) are no candidates; remaining bytes per candidate
        uint32_t nc = (uint32_t)__popcll(cand);
        uint32_t total_c;
        const uint32_t cex = block_excl_sum<THREADS>(nc, s_w, total_c);
        const bool listed = total_c <= (uint32_t)MANY_LCAP;
        if (listed) {
            uint32_t at = cex;
            uint64_t m = cand;
            while (m) {
                const uint32_t p = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                s_cand[at] = (uint16_t)(lo + p);
                s_rem[at] = (uint8_t)rem_at(g0 + p);
                at++;
            }
        }
        __syncthreads();

        // 2-3. count
        uint32_t own = 0;   // per-lane count (unlisted tiles)
        if (listed) {
            for (uint32_t i = tid; i < total_c; i += THREADS) {
                const uint32_t lp = s_cand[i], rm = s_rem[i];
                s_cnt[i] = rm ? probe_position(d, s_text, lp, (uint64_t)base + lp, rm, false, nullptr, nullptr, 0, 0)
                              : 0u;
            }
        } else {
            di = di0;
            dlim = dlim0;
            dnext = dnext0;
            uint64_t m = cand;
            while (m) {
                const uint32_t p = (uint32_t)__builtin_ctzll(m);
                m &= m - 1;
                const uint32_t rm = rem_at(g0 + p);
                if (rm)
                    own += probe_position(d, s_text, lo + p, (uint64_t)g0 + p, rm, false, nullptr, nullptr, 0, 0);
            }
        }
        __syncthreads();
        // 4. place: per-candidate exclusive offsets (listed) or per-lane ones
        uint32_t total, lane_ex = 0;
        if (listed) {
            const uint32_t K = (total_c + THREADS - 1) / THREADS;
            const uint32_t i0 = min(total_c, tid * K), i1 = min(total_c, i0 + K);
            uint32_t loc = 0;
            for (uint32_t i = i0; i < i1; i++)
                loc += s_cnt[i];
            uint32_t run = block_excl_sum<THREADS>(loc, s_w, total);
            for (uint32_t i = i0; i < i1; i++) {
                const uint32_t c = s_cnt[i];
                s_cnt[i] = run;
                run += c;
            }
        } else {
            lane_ex = block_excl_sum<THREADS>(own, s_w, total);
        }
        const uint64_t prefix = tile_prefix(a.status, t, total, &s_prefix, a.lb_err);
        if (tid == 0 && t == a.num_tiles - 1)
            a.meta[0] = prefix + total;
        if (a.out_off || a.out_id) {
            if (listed) {
                for (uint32_t i = tid; i < total_c; i += THREADS) {
                    const uint32_t lp = s_cand[i], rm = s_rem[i];
                    const uint64_t r = prefix + s_cnt[i];
                    if (rm && r < a.cap)
                        probe_position(d, s_text, lp, (uint64_t)base + lp, rm, true, a.out_off, a.out_id, r, a.cap);
                }
            } else {
                di = di0;
                dlim = dlim0;
                dnext = dnext0;
                uint64_t r = prefix + lane_ex;
                uint64_t m = cand;
                while (m && r < a.cap) {
                    const uint32_t p = (uint32_t)__builtin_ctzll(m);
                    m &= m - 1;
                    const uint32_t rm = rem_at(g0 + p);
                    if (rm)
                        r += probe_position(d, s_text, lo + p, (uint64_t)g0 + p, rm, true, a.out_off, a.out_id, r,
                                            a.cap);
                }
            }
        }
        __syncthreads();   // LDS is reused by the next tile
    }
}

size_t ol_many_lds_bytes(uint32_t fwords)
{
    return (size_t)SEARCH_TILE + MANY_HALO + (size_t)fwords * 4;
}

hipError_t ol_launch_many(const ManyArgs &a, hipStream_t s)
{
    const size_t lds = ol_many_lds_bytes(a.d.fwords);
    const void *f = (const void *)many_kernel<SEARCH_THREADS>;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    uint32_t cap = max_grid(f, SEARCH_THREADS, lds);
    if (!cap)
        return hipErrorInvalidValue;
    uint32_t grid = a.num_tiles < cap ? a.num_tiles : cap;
    hipLaunchKernelGGL((many_kernel<SEARCH_THREADS>), dim3(grid), dim3(SEARCH_THREADS), lds, s, a);
    return hipGetLastError();
}
