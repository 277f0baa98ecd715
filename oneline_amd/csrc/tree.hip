// tree.hip -- parent / first_child / next_sibling / subtree_end from the
// per-line key bytes, one line per lane.
//
// The reference builds parent + subtree_end with seven passes of per-thread
// 32-line runs (crates/oneline-gpu/src/tree.wgsl:127-375) and checks them
// against the monotonic stack of TreeLinks::from_nodes_sequential
// (crates/oneline-core/src/links.rs:52-81).  Here every line is resolved
// independently from two per-level states:
//
//   F[d] = the last content line at depth d before a point (forward),
//   B[d] = the first content line at depth d after a point (backward).
//
//   parent[i]       = max over d < x_i of F[d]            (nearest previous smaller)
//   subtree_end[i]  = min over d <= x_i of B[d], else n   (nearest next smaller-or-equal)
//   first_child[i]  = the next content line j if x_j > x_i, else NONE
//   next_sibling[i] = k = subtree_end[i] if k < n and depth(parent[i]) < x_k
//                     (or parent[i] is NONE), else NONE
//
// The last rule restates "parent[k] == parent[i]": the content lines between i
// and k are deeper than x_i >= x_k, so parent[k] is the nearest line before i
// shallower than x_k, which is parent[i] exactly when parent[i] is shallower
// than x_k.  first_child restates "first content line in (i, subtree_end[i])".
//
// Work split: a wave resolves 64 consecutive lines per round, lane = line.  For
// each level d the wave takes ballot(key < d); a lane's nearest smaller line
// inside the round is then one masked clz / ctz, and only lines whose answer
// lies outside the round use F / B, kept one level per lane (prefix max / min
// over the lanes = over the levels, read with one shuffle).  Three launches:
//
//   tree_summary: per tile (4 waves x 16 rounds x 64 lines), F and B of the tile;
//   tree_scan:    per level, max over earlier tiles / min over later tiles;
//   tree_walk:    the four outputs, written once each and coalesced.
//
// Blank lines are outside the tree: parent NONE, subtree_end i + 1,
// first_child = next_sibling = NONE.  A content line nested TREE_D or more
// levels deep sets the error flag (the reference's MAX_TREE_DEPTH contract,
// crates/oneline-gpu/src/tree.rs:10-12,218-222).
#include "internal.h"

constexpr int TW = 256;              // threads per tile (4 waves); a tile = TW x TR lines
constexpr int D = TREE_D;            // levels tracked
static_assert(D == 32, "five key bits, lanes 0..32");
static_assert(TW / 64 == (int)TREE_WAVES, "per-wave summaries");

__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// The tile's document: its first global line, line count and the tile's
// index inside it (a single document when nsegs == 0).
struct TileSeg {
    uint64_t g0;
    uint64_t L;
    uint32_t lt;
};
__device__ __forceinline__ TileSeg tile_seg(const TreeArgs &a, uint32_t t)
{
    TileSeg ts{0, a.lines, t};
    if (a.nsegs) {
        const uint32_t sg = a.tile_seg[t];
        ts.g0 = a.seg_line0[sg];
        ts.L = a.seg_lines[sg];
        ts.lt = t - a.seg_tile0[sg];
    }
    return ts;
}

// Keys of this lane's TR rounds (lines past the end read as blank).  The
// loads are unconditional (past the end they read the document's last key),
// so all TR are in flight together; a guarded load per round compiled to TR
// round trips one after another.
template <int TR>
__device__ __forceinline__ void load_keys(const uint8_t *keys, const TileSeg &ts, uint32_t wbase, uint32_t (&key)[TR])
{
    const int lane = threadIdx.x & 63;
    if (ts.L == 0) {
#pragma unroll
        for (int r = 0; r < TR; r++)
            key[r] = KEY_BLANK;
        return;
    }
    const uint8_t *const kd = keys + ts.g0;
    if ((uint64_t)wbase + TR * 64 <= ts.L) {   // the whole wave inside the document
#pragma unroll
        for (int r = 0; r < TR; r++)
            key[r] = kd[wbase + r * 64 + lane];
        return;
    }
    const uint64_t last = ts.L - 1;
    uint32_t v[TR];
#pragma unroll
    for (int r = 0; r < TR; r++) {
        const uint64_t g = wbase + r * 64 + lane;
        v[r] = kd[g < last ? g : last];
    }
#pragma unroll
    for (int r = 0; r < TR; r++)
        key[r] = (wbase + r * 64 + lane < ts.L) ? v[r] : KEY_BLANK;
}

// Exclusive prefix OR over the lanes of a 64-bit value (lane 0 gets 0): DPP
// row shifts within rows of 16 lanes, then the row broadcasts carry rows 0-1
// into rows 1-3, as wave_incl_add does for sums.
__device__ __forceinline__ uint64_t wave_excl_or64(uint64_t v)
{
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
#define OL_OR_STEP(CTRL, RM)              \
    lo |= dpp_move<CTRL, RM>(0u, lo);     \
    hi |= dpp_move<CTRL, RM>(0u, hi);
    OL_OR_STEP(0x111, 0xF) OL_OR_STEP(0x112, 0xF) OL_OR_STEP(0x114, 0xF) OL_OR_STEP(0x118, 0xF)
    OL_OR_STEP(0x142, 0xA) OL_OR_STEP(0x143, 0xC)
#undef OL_OR_STEP
    return (uint64_t)wave_shr1(hi, 0u) << 32 | wave_shr1(lo, 0u);
}

// Level masks of a round (64 lines, lane = line, x = its key): lane d
// (0 <= d < 32) gets eq = {lines with key == d}, from five ballots of the key
// bits compared with d's bits (`dm[b]` is all ones in lanes whose bit b is
// set), and lt = {lines with key < d}, the OR of the lower lanes' eq; lane 32
// gets lt = {all content lines}, the OR of all 32.
__device__ __forceinline__ void level_masks(uint32_t x, const uint64_t (&dm)[5], uint64_t &lt, uint64_t &eq)
{
    uint64_t e = __ballot(x < (uint32_t)D);
#pragma unroll
    for (int b = 4; b >= 0; b--)
        e &= ~(__ballot((x >> b) & 1u) ^ dm[b]);   // key bit b == d's bit b
    eq = e;
    lt = wave_excl_or64(e);
}

__device__ __forceinline__ void lane_bit_masks(uint64_t (&dm)[5])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int b = 0; b < 5; b++)
        dm[b] = ((lane >> b) & 1) ? ~0ull : 0ull;
}

// Per wave and per tile: for each level the last content line (+1, 0 = none)
// and the first (NONE = none).  Lane l takes the TR consecutive lines
// wbase + l TR + k of the wave's range and notes, per level it meets, the
// run-local index of its last and first line there (LDS, [level][lane]) and a
// presence bit; 32 ballots of the presence bits then give lane d the runs that
// hold level d, and the last and first of them hold the wave's answers.  No
// lane waits on another lane's line.
template <int TR>
__global__ __launch_bounds__(TW) void tree_summary_kernel(TreeArgs a)
{
    static_assert(TR <= 64 && 64 * TR <= (D + 1) * 64 * 2, "a run fits the level rows");
    constexpr int NWV = TW / 64;
    __shared__ uint8_t s_run[NWV][2][D + 1][64];   // [last, first][level][lane]; row D takes non-content lines
    __shared__ uint32_t s_l[NWV][D], s_f[NWV][D];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t t = blockIdx.x;
    const TileSeg ts = tile_seg(a, t);
    const uint32_t wbase = ts.lt * (TW * TR) + w * (TR * 64);
    uint8_t (*const lastk)[64] = s_run[w][0];
    uint8_t (*const firstk)[64] = s_run[w][1];

    // keys: loaded a round of 64 lines at a time (coalesced), turned into runs
    uint32_t key[TR];
    {
        uint32_t kr[TR];
        load_keys<TR>(a.keys, ts, wbase, kr);
        uint8_t *const kb = &lastk[0][0];
#pragma unroll
        for (int r = 0; r < TR; r++)
            kb[r * 64 + lane] = (uint8_t)kr[r];
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < TR; k++)
            key[k] = kb[lane * TR + k];
        wave_lds_sync();
    }
    // the depth error flag is set by the walk (the next launch but one)
    if (t == 0 && threadIdx.x == 0)
        *a.err = 0u;

    // the last line of each level: written forward, the last write stays; the
    // first: the same addresses written backward
    uint32_t pres = 0;   // levels met in the run
    uint32_t at[TR];
#pragma unroll
    for (int k = 0; k < TR; k++) {
        const uint32_t x = key[k];
        const bool c = x < (uint32_t)D;
        at[k] = (c ? x : (uint32_t)D) * 64u + (uint32_t)lane;
        lastk[0][at[k]] = (uint8_t)k;
        pres |= c ? 1u << x : 0u;
    }
#pragma unroll
    for (int k = TR - 1; k >= 0; k--)
        firstk[0][at[k]] = (uint8_t)k;
    wave_lds_sync();
    // lane d: the runs that hold level d.  The presence words of each half
    // wave are a 32 x 32 bit matrix (lane = row, level = column), transposed
    // by swapping ever smaller off-diagonal blocks between lanes l and l ^ s;
    // lane d then holds column d of its half, and the other half's comes
    // from lane d + 32.
    uint32_t col = pres;
#pragma unroll
    for (int sh = 16; sh >= 1; sh >>= 1) {
        const uint32_t m = sh == 16 ? 0x0000FFFFu : sh == 8 ? 0x00FF00FFu : sh == 4 ? 0x0F0F0F0Fu
                         : sh == 2 ? 0x33333333u : 0x55555555u;   // the columns whose bit sh is 0
        const uint32_t o = (uint32_t)__shfl_xor((int)col, sh, 64);
        col = (lane & sh) ? ((col & ~m) | ((o & ~m) >> sh)) : ((col & m) | ((o & m) << sh));
    }
    const uint64_t md = (uint64_t)(uint32_t)__shfl_xor((int)col, 32, 64) << 32 | col;
    uint32_t last = 0, first = OL_NONE;
    if (lane < D && md) {
        const uint32_t rl = 63u - (uint32_t)__clzll(md), rf = (uint32_t)__builtin_ctzll(md);
        last = wbase + rl * TR + lastk[lane][rl] + 1u;   // index + 1
        first = wbase + rf * TR + firstk[lane][rf];
    }
    if (lane < D) {
        s_l[w][lane] = last;
        s_f[w][lane] = first;
        const size_t wi = ((size_t)t * TREE_WAVES + w) * D + lane;   // wave-major
        a.wsum_last[wi] = last;
        a.wsum_first[wi] = first;
    }
    __syncthreads();
    if (w == 0 && lane < D) {
        uint32_t l = 0, f = OL_NONE;
#pragma unroll
        for (int k = 0; k < TW / 64; k++) {
            l = umax(l, s_l[k][lane]);
            f = umin(f, s_f[k][lane]);
        }
        a.sum_last[(size_t)lane * a.num_tiles + t] = l;   // level-major
        a.sum_first[(size_t)lane * a.num_tiles + t] = f;
    }
}

__device__ __forceinline__ uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }

// Two workgroups per level, one per direction (arrays level-major, one row of
// num_tiles per level): in_last[t] = max of sum_last over the earlier tiles of
// t's document, out_first[t] = min of sum_first over its later tiles.  The scan runs over
// (document << 32 | value): the maximum over earlier tiles carries t's own
// document exactly when one of them belongs to it (documents are contiguous
// runs of tiles), and likewise the minimum over later tiles.  The row goes
// through LDS in chunks: coalesced loads, each thread scans a contiguous run
// of the chunk, a block scan joins the runs, coalesced stores.
template <int ST>
__global__ __launch_bounds__(ST) void tree_scan_kernel(TreeArgs a)
{
    constexpr int NWS = ST / 64;
    constexpr int K = 8;                 // tiles per thread per chunk (8,192 per chunk: one chunk at 2 GiB)
    constexpr int CH = ST * K;
    __shared__ uint64_t s_v[CH];
    __shared__ uint64_t s_w[NWS];
    const uint32_t T = a.num_tiles;
    // workgroup 2 d takes level d forward, 2 d + 1 backward
    const size_t row = (size_t)(blockIdx.x >> 1) * T;
    const bool fwd = (blockIdx.x & 1u) == 0;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    auto seg = [&](uint32_t t) -> uint64_t { return a.nsegs ? (uint64_t)a.tile_seg[t] : 0ull; };

    // forward: exclusive prefix max
    uint64_t carry = 0;
    for (uint32_t c0 = 0; fwd && c0 < T; c0 += CH) {
        const uint32_t n = (T - c0) < (uint32_t)CH ? T - c0 : (uint32_t)CH;
        for (uint32_t i = tid; i < (uint32_t)CH; i += ST)
            s_v[i] = i < n ? (seg(c0 + i) << 32 | a.sum_last[row + c0 + i]) : 0ull;
        __syncthreads();
        uint64_t loc[K], m = 0;
#pragma unroll
        for (int k = 0; k < K; k++) {
            loc[k] = s_v[tid * K + k];
            m = umax64(m, loc[k]);
        }
        uint64_t inc = m;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t y = __shfl_up(inc, o, 64);
            if (lane >= o)
                inc = umax64(inc, y);
        }
        if (lane == 63)
            s_w[w] = inc;
        __syncthreads();
        uint64_t run = __shfl_up(inc, 1, 64);
        if (lane == 0)
            run = 0;
        uint64_t tot = 0;
        for (int k = 0; k < NWS; k++) {
            if (k < w)
                run = umax64(run, s_w[k]);
            tot = umax64(tot, s_w[k]);
        }
        run = umax64(run, carry);
#pragma unroll
        for (int k = 0; k < K; k++) {
            s_v[tid * K + k] = run;
            run = umax64(run, loc[k]);
        }
        __syncthreads();
        for (uint32_t i = tid; i < n; i += ST) {
            const uint64_t r = s_v[i];
            a.in_last[row + c0 + i] = (r >> 32) == seg(c0 + i) ? (uint32_t)r : 0u;
        }
        carry = umax64(carry, tot);
        __syncthreads();
    }

    // backward: exclusive suffix min
    carry = ~0ull;
    for (uint32_t c1 = T; !fwd && c1 > 0;) {
        const uint32_t n = c1 < (uint32_t)CH ? c1 : (uint32_t)CH;
        const uint32_t c0 = c1 - n;
        for (uint32_t i = tid; i < (uint32_t)CH; i += ST)
            s_v[i] = i < n ? (seg(c0 + i) << 32 | a.sum_first[row + c0 + i]) : ~0ull;
        __syncthreads();
        uint64_t loc[K], m = ~0ull;
#pragma unroll
        for (int k = 0; k < K; k++) {
            loc[k] = s_v[tid * K + k];
            m = umin64(m, loc[k]);
        }
        uint64_t inc = m;   // inclusive suffix min over lanes >= lane
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t y = __shfl_down(inc, o, 64);
            if (lane + o < 64)
                inc = umin64(inc, y);
        }
        if (lane == 0)
            s_w[w] = inc;
        __syncthreads();
        uint64_t run = __shfl_down(inc, 1, 64);
        if (lane == 63)
            run = ~0ull;
        uint64_t tot = ~0ull;
        for (int k = 0; k < NWS; k++) {
            if (k > w)
                run = umin64(run, s_w[k]);
            tot = umin64(tot, s_w[k]);
        }
        run = umin64(run, carry);
#pragma unroll
        for (int k = K - 1; k >= 0; k--) {
            s_v[tid * K + k] = run;
            run = umin64(run, loc[k]);
        }
        __syncthreads();
        for (uint32_t i = tid; i < n; i += ST) {
            const uint64_t r = s_v[i];
            a.out_first[row + c0 + i] = (r >> 32) == seg(c0 + i) ? (uint32_t)r : OL_NONE;
        }
        carry = umin64(carry, tot);
        c1 = c0;
        __syncthreads();
    }
}

// Per round (lane k = line rb + k, key x): the nearest smaller-key line inside
// the round is a masked clz / ctz on the round's level masks (LDS).  Lines
// whose answer lies outside the round read two per-lane states, lane d holding
// level d: G[d] = last line before the round with key < d (+1, 0 = none) and
// its key, and H[d] = first line after the round with key <= d and its key.
// G moves forward round by round from the round's own masks (the last line of
// {key < d}); H is filled backwards over the rounds in LDS before the walk.
template <int TR>
__global__ __launch_bounds__(TW) void tree_walk_kernel(TreeArgs a)
{
    constexpr int NWV = TW / 64;
    __shared__ uint64_t s_m[NWV][TR][D + 1];   // lane d's {key < d} of each round
    __shared__ uint32_t s_h[NWV][TR][D];       // first line with key <= d: in the round, then after it
    __shared__ uint8_t s_hl[NWV][TR][D];       // that line's key

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t t = blockIdx.x;
    const uint32_t T = a.num_tiles;
    const TileSeg ts = tile_seg(a, t);
    // a document's lines are numbered in 32 bits (its offsets are u32), so the
    // per-line bound checks compare 32-bit values
    const uint32_t L = (uint32_t)ts.L;
    const uint32_t wbase = ts.lt * (TW * TR) + w * (TR * 64);
    uint32_t key[TR];
    load_keys<TR>(a.keys, ts, wbase, key);

    // pass 1: masks; per round the first line with key <= d; content lines
    // too deep for the tree.  The waves share no LDS: each works on its own
    // slices, and the other waves' per-level lines come from the summary.
    uint64_t dm[5];
    lane_bit_masks(dm);
    bool deep = false;
#pragma unroll
    for (int r = 0; r < TR; r++) {
        deep |= (key[r] != KEY_BLANK && key[r] >= (uint32_t)D);
        uint64_t lt, eq;
        level_masks(key[r], dm, lt, eq);
        if (lane <= D)
            s_m[w][r][lane] = lt;
        const uint32_t rb = wbase + r * 64;
        const uint64_t le = lt | eq;
        const int j = le ? (int)__builtin_ctzll(le) : 0;
        const uint32_t kj = __shfl(key[r], j, 64);
        if (lane < D) {
            s_h[w][r][lane] = le ? rb + (uint32_t)j : OL_NONE;
            s_hl[w][r][lane] = (uint8_t)kj;
        }
    }
    if (__ballot(deep) && lane == 0)
        atomicOr(a.err, 1u);
    wave_lds_sync();

    // per-level states at the wave's first line (F, last line + 1) and after
    // its last line (B, first line): the tile's, and the tile's other waves'
    uint32_t F = 0, Bw = OL_NONE;
    if (lane < D) {
        F = a.in_last[(size_t)lane * T + t];
        Bw = a.out_first[(size_t)lane * T + t];
        const uint32_t *const wl = a.wsum_last + (size_t)t * TREE_WAVES * D + lane;
        const uint32_t *const wf = a.wsum_first + (size_t)t * TREE_WAVES * D + lane;
        for (int k = 0; k < w; k++)
            F = umax(F, wl[k * D]);
        for (int k = w + 1; k < NWV; k++)
            Bw = umin(Bw, wf[k * D]);
    }
    // G = exclusive prefix max of F over levels, H = inclusive prefix min of B
    uint32_t pv = (lane < D) ? F : 0u, pl = (uint32_t)lane;
    uint32_t sv = (lane < D) ? Bw : OL_NONE, sl = (uint32_t)lane;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const uint32_t pv2 = __shfl_up(pv, o, 64), pl2 = __shfl_up(pl, o, 64);
        const uint32_t sv2 = __shfl_up(sv, o, 64), sl2 = __shfl_up(sl, o, 64);
        if (lane >= o && pv2 > pv) {
            pv = pv2;
            pl = pl2;
        }
        if (lane >= o && sv2 < sv) {
            sv = sv2;
            sl = sl2;
        }
    }
    uint32_t G = __shfl_up(pv, 1, 64), GL = __shfl_up(pl, 1, 64);
    if (lane == 0)
        G = 0;
    // H after each round, backwards over the rounds
    if (lane < D) {
        uint32_t cv = sv, cl = sl;
        for (int r = TR - 1; r >= 0; r--) {
            const uint32_t fv = s_h[w][r][lane];
            const uint32_t fl = s_hl[w][r][lane];
            s_h[w][r][lane] = cv;
            s_hl[w][r][lane] = (uint8_t)cl;
            if (fv < cv) {
                cv = fv;
                cl = fl;
            }
        }
    }
    wave_lds_sync();

    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t above = (lane == 63) ? 0ull : ~((2ull << lane) - 1ull);
    // the lane's output line in round 0: round r stores 64 r entries further,
    // an immediate offset on one address per output kept across the rounds
    const uint64_t wg = ts.g0 + wbase + lane;
    uint32_t *const P = a.parent ? a.parent + wg : nullptr;
    uint32_t *const FC = a.first_child ? a.first_child + wg : nullptr;
    uint32_t *const NS = a.next_sibling ? a.next_sibling + wg : nullptr;
    uint32_t *const SE = a.subtree_end ? a.subtree_end + wg : nullptr;

#pragma unroll
    for (int r = 0; r < TR; r++) {
        const uint32_t rb = wbase + r * 64;
        const uint32_t x = key[r];
        const bool valid = x < (uint32_t)D;
        const uint32_t xi = valid ? x : 0u;
        const uint64_t mlt = s_m[w][r][xi], mle = s_m[w][r][xi + 1];
        const uint64_t pm = mlt & below, sm = mle & above;
        const uint64_t fm = __ballot(x != KEY_BLANK) & above;
        const uint32_t jp = pm ? 63u - (uint32_t)__clzll(pm) : 0u;
        const uint32_t js = sm ? (uint32_t)__builtin_ctzll(sm) : 0u;
        const uint32_t jf = fm ? (uint32_t)__builtin_ctzll(fm) : 0u;
        const uint32_t kp = __shfl(x, jp, 64), ks = __shfl(x, js, 64);
        const uint32_t gv = __shfl(G, xi, 64), gl = __shfl(GL, xi, 64);
        const uint32_t hv = s_h[w][r][xi], hl = s_hl[w][r][xi];
        const uint32_t nv = s_h[w][r][D - 1];   // first content line after the round

        const uint32_t g = rb + lane;
        // content line (blank and too deep lines: below; too deep is reported
        // by the error flag)
        const uint32_t cpar = pm ? rb + jp : (gv ? gv - 1 : OL_NONE);
        const uint32_t plv = pm ? kp : gl;
        const uint32_t cse = sm ? rb + js : (hv != OL_NONE ? hv : L);
        const uint32_t slv = sm ? ks : hl;
        // the next content line is a child exactly when it comes before
        // subtree_end (the first later line with key <= x); NONE > L >= cse
        const uint32_t fi = fm ? rb + jf : nv;
        const uint32_t cfc = fi < cse ? fi : OL_NONE;
        const uint32_t cns = (cse < L && (cpar == OL_NONE || plv < slv)) ? cse : OL_NONE;
        const uint32_t par = valid ? cpar : OL_NONE;
        const uint32_t se = valid ? cse : g + 1;
        const uint32_t fc = valid ? cfc : OL_NONE;
        const uint32_t ns = valid ? cns : OL_NONE;
        if (g < L) {
            const uint32_t o = (uint32_t)(r * 64);   // the lane is in the pointers
            if (P) P[o] = par;
            if (SE) SE[o] = se;
            if (FC) FC[o] = fc;
            if (NS) NS[o] = ns;
        }
        // G after this round: the round's last line with key < d, if any
        const uint64_t ltd = s_m[w][r][lane < D ? lane : 0];
        const uint32_t jd = ltd ? 63u - (uint32_t)__clzll(ltd) : 0u;
        const uint32_t kd = __shfl(x, jd, 64);
        if (lane < D && ltd) {
            G = rb + jd + 1;
            GL = kd;
        }
    }
}

template <int TR>
static void launch_tree_tr(const TreeArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((tree_summary_kernel<TR>), dim3(a.num_tiles), dim3(TW), 0, s, a);
    hipLaunchKernelGGL((tree_scan_kernel<1024>), dim3(2 * D), dim3(1024), 0, s, a);
    hipLaunchKernelGGL((tree_walk_kernel<TR>), dim3(a.num_tiles), dim3(TW), 0, s, a);
}

hipError_t ol_launch_tree(const TreeArgs &a, uint32_t rounds, hipStream_t s)
{
    switch (rounds) {
    case 1: launch_tree_tr<1>(a, s); break;
    case 4: launch_tree_tr<4>(a, s); break;
    case 16: launch_tree_tr<16>(a, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

uint32_t ol_tree_tile_lines(uint32_t rounds) { return TW * rounds; }
