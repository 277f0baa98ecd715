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
__device__ __forceinline__ uint32_t filter_index(uint32_t k, uint32_t key, uint32_t bits)
{
    if (k <= 2)
        return key;   // direct: 2^8 / 2^16 bits
    return (key * 0x9E3779B1u) >> (32 - bits);
}

__device__ __forceinline__ uint32_t table_hash(uint32_t key, uint32_t bits)
{
    return ((key ^ (key >> 15)) * 0x85EBCA6Bu) >> (32 - bits);
}

// Matches starting at local position lp (global gp).  When `emit`, writes them
// at out[r..] (up to cap) in pattern-id order.  Returns the number of matches.
// `lim` (> gp) bounds the match end: the text end, or the end of gp's document.
__device__ uint32_t probe_position(const DictDev &d, const uint32_t *s_filter, const uint8_t *s_text,
                                   uint32_t lp, uint32_t gp, uint32_t lim, uint32_t w, bool emit,
                                   uint64_t *out_off, uint32_t *out_id, uint64_t &r, uint64_t cap)
{
    const uint32_t rem = lim - gp;
    uint32_t beg[4], end[4], kk[4];
    int nr = 0;
#pragma unroll
    for (uint32_t k = 1; k <= 4; k++) {
        if (!((d.class_mask >> k) & 1u) || k > rem)
            continue;
        uint32_t key = (k == 4) ? w : (w & ((1u << (8 * k)) - 1u));
        uint32_t fi = filter_index(k, key, d.fbits[k]);
        if (!((s_filter[d.foff[k] + (fi >> 5)] >> (fi & 31)) & 1u))
            continue;
        const uint32_t tb = d.tbits[k];
        const uint32_t tmask = (1u << tb) - 1u;
        uint32_t h = table_hash(key, tb);
        const uint4 *tab = d.table[k];
        while (true) {
            uint4 sl = tab[h];
            if (sl.z == 0)
                break;
            if (sl.x == key) {
                beg[nr] = sl.y;
                end[nr] = sl.z;
                kk[nr] = k;
                nr++;
                break;
            }
            h = (h + 1) & tmask;
        }
    }
    if (nr == 0)
        return 0;
    uint32_t count = 0;
    while (true) {
        int best = -1;
        uint32_t bid = 0xFFFFFFFFu;
        for (int i = 0; i < nr; i++) {
            if (beg[i] < end[i]) {
                uint32_t id = d.ent_id[beg[i]];
                if (id < bid) {
                    bid = id;
                    best = i;
                }
            }
        }
        if (best < 0)
            break;
        const uint32_t e = beg[best]++;
        const uint32_t len = d.ent_len[e];
        if (len > rem)
            continue;
        const uint8_t *pb = d.pool + d.ent_off[e];
        bool ok = true;
        for (uint32_t i = kk[best]; i < len; i++) {
            if (pb[i] != s_text[lp + i]) {
                ok = false;
                break;
            }
        }
        if (!ok)
            continue;
        count++;
        if (emit) {
            if (r < cap) {
                if (out_off) out_off[r] = gp;
                if (out_id) out_id[r] = bid;
            }
            r++;
        }
    }
    return count;
}

template <int THREADS, int BPT>
__global__ __launch_bounds__(THREADS) void many_kernel(ManyArgs a)
{
    constexpr int TILE = THREADS * BPT;
    static_assert(BPT == 64, "64-bit lane masks");
    extern __shared__ __align__(16) uint32_t s_dyn[];
    __shared__ uint32_t s_w[THREADS / 64];
    __shared__ uint64_t s_prefix;

    uint8_t *s_text = (uint8_t *)s_dyn;                       // TILE + HALO_AFTER bytes
    uint32_t *s_filter = s_dyn + (TILE + HALO_AFTER) / 4;   // fwords

    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x;
    for (uint32_t i = tid; i < a.d.fwords; i += THREADS)
        s_filter[i] = a.d.filter[i];
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
        load_tile<THREADS, BPT>(s_text, a.text, base);
        __syncthreads();

        const uint32_t lo = tid * BPT;
        const uint32_t g0 = base + lo;
        uint32_t lim = (g0 < a.n) ? a.n - g0 : 0;
        if (lim > BPT)
            lim = BPT;

        const uint32_t *wsrc = (const uint32_t *)(s_text + lo);
        // match-end bound per position: the text end, or (batches) the end of
        // the position's document; positions between documents match nothing
        uint32_t dlim = a.n, dnext = 0xFFFFFFFFu;
        uint32_t di = 0;
        if (a.ndocs) {
            di = a.tile_doc[t];   // then the last document starting at or before g0
            while (di + 1 < a.ndocs && a.doc_starts[di + 1] <= g0)
                di++;
            dlim = a.doc_ends[di];
            dnext = (di + 1 < a.ndocs) ? a.doc_starts[di + 1] : 0xFFFFFFFFu;
        }
        auto bound = [&](uint32_t gp) -> uint32_t {   // 0: no match may start at gp
            while (gp >= dnext) {
                di++;
                dlim = a.doc_ends[di];
                dnext = (di + 1 < a.ndocs) ? a.doc_starts[di + 1] : 0xFFFFFFFFu;
            }
            return gp < dlim ? dlim : 0u;
        };
        const uint32_t di0 = di, dlim0 = dlim, dnext0 = dnext;
        uint64_t has = 0;
        uint32_t count = 0;
        uint64_t dummy = 0;
        for (uint32_t p = 0; p < lim; p++) {
            const uint32_t bl = bound(g0 + p);
            if (!bl)
                continue;
            uint32_t w0 = wsrc[p >> 2], w1 = wsrc[(p >> 2) + 1];
            uint32_t win = (p & 3) ? __builtin_amdgcn_alignbyte(w1, w0, p & 3) : w0;
            uint32_t c = probe_position(a.d, s_filter, s_text, lo + p, g0 + p, bl, win, false,
                                        nullptr, nullptr, dummy, 0);
            if (c) {
                has |= 1ull << p;
                count += c;
            }
        }
        uint32_t total;
        const uint32_t excl = block_excl_sum<THREADS>(count, s_w, total);
        const uint64_t prefix = tile_prefix(a.status, t, total, &s_prefix, a.lb_err);
        if (tid == 0 && t == a.num_tiles - 1)
            a.meta[0] = prefix + total;
        if (a.out_off || a.out_id) {
            uint64_t r = prefix + excl;
            di = di0;
            dlim = dlim0;
            dnext = dnext0;
            while (has && r < a.cap) {
                uint32_t p = (uint32_t)__builtin_ctzll(has);
                has &= has - 1;
                uint32_t w0 = wsrc[p >> 2], w1 = wsrc[(p >> 2) + 1];
                uint32_t win = (p & 3) ? __builtin_amdgcn_alignbyte(w1, w0, p & 3) : w0;
                probe_position(a.d, s_filter, s_text, lo + p, g0 + p, bound(g0 + p), win, true, a.out_off,
                               a.out_id, r, a.cap);
            }
        }
        __syncthreads();   // LDS is reused by the next tile
    }
}

size_t ol_many_lds_bytes(uint32_t fwords)
{
    return (size_t)SEARCH_TILE + HALO_AFTER + (size_t)fwords * 4;
}

hipError_t ol_launch_many(const ManyArgs &a, hipStream_t s)
{
    const size_t lds = ol_many_lds_bytes(a.d.fwords);
    const void *f = (const void *)many_kernel<SEARCH_THREADS, SEARCH_BPT>;
    uint32_t cap = max_grid(f, SEARCH_THREADS, lds);
    if (!cap)
        return hipErrorInvalidValue;
    uint32_t grid = a.num_tiles < cap ? a.num_tiles : cap;
    hipLaunchKernelGGL((many_kernel<SEARCH_THREADS, SEARCH_BPT>), dim3(grid), dim3(SEARCH_THREADS), lds, s, a);
    return hipGetLastError();
}
