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

constexpr int TW = 256;              // threads per tile (4 waves)
constexpr int TR = 16;               // rounds of 64 lines per wave
constexpr int TLINES = TW * TR;      // lines per tile
constexpr int D = TREE_D;            // levels tracked
static_assert(D <= 63, "levels + 1 must fit a wave");

__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// Keys of this lane's 16 rounds (lines past the end read as blank).
__device__ __forceinline__ void load_keys(const uint8_t *keys, uint32_t wbase, uint64_t L, uint32_t (&key)[TR])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int r = 0; r < TR; r++) {
        const uint32_t g = wbase + r * 64 + lane;
        key[r] = (g < L) ? (uint32_t)keys[g] : KEY_BLANK;
    }
}

// Lane d (d <= D) gets ballot(key < d) of the round.
__device__ __forceinline__ uint64_t level_masks(uint32_t key)
{
    const int lane = threadIdx.x & 63;
    uint64_t my = 0;
#pragma unroll
    for (int d = 0; d <= D; d++) {
        const uint64_t m = __ballot(key < (uint32_t)d);
        my = (lane == d) ? m : my;
    }
    return my;
}

// Per tile: for each level the last content line (+1, 0 = none) and the first
// (NONE = none); error flag for content lines at depth >= D.
__global__ __launch_bounds__(TW) void tree_summary_kernel(TreeArgs a)
{
    __shared__ uint32_t s_l[TW / 64][D], s_f[TW / 64][D];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t t = blockIdx.x;
    const uint32_t wbase = t * TLINES + w * (TR * 64);
    uint32_t key[TR];
    load_keys(a.keys, wbase, a.lines, key);
    uint32_t last = 0, first = OL_NONE;
    bool deep = false;
#pragma unroll
    for (int r = 0; r < TR; r++) {
        const uint64_t my = level_masks(key[r]);
        const uint64_t up = __shfl_down(my, 1, 64);
        const uint64_t eq = up & ~my;   // lanes whose key == lane id
        const uint32_t rb = wbase + r * 64;
        if (lane < D && eq) {
            last = rb + 64 - (uint32_t)__clzll(eq);     // index + 1
            if (first == OL_NONE)
                first = rb + (uint32_t)__builtin_ctzll(eq);
        }
        deep |= (key[r] != KEY_BLANK && key[r] >= (uint32_t)D);
    }
    if (__ballot(deep) && lane == 0)
        atomicOr(a.err, 1u);
    if (lane < D) {
        s_l[w][lane] = last;
        s_f[w][lane] = first;
    }
    __syncthreads();
    if (w == 0 && lane < D) {
        uint32_t l = 0, f = OL_NONE;
#pragma unroll
        for (int k = 0; k < TW / 64; k++) {
            l = umax(l, s_l[k][lane]);
            f = umin(f, s_f[k][lane]);
        }
        a.sum_last[(size_t)t * D + lane] = l;
        a.sum_first[(size_t)t * D + lane] = f;
    }
}

// One workgroup per level: in_last[t] = max of sum_last over tiles < t,
// out_first[t] = min of sum_first over tiles > t.  Each thread owns a contiguous
// chunk of tiles.
template <int ST>
__global__ __launch_bounds__(ST) void tree_scan_kernel(TreeArgs a)
{
    constexpr int NWS = ST / 64;
    __shared__ uint32_t s_w[NWS];
    const uint32_t d = blockIdx.x;
    const uint32_t T = a.num_tiles;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t K = (T + ST - 1) / ST;
    const uint32_t t0 = tid * K, t1 = umin(T, t0 + K);

    // forward: exclusive prefix max
    uint32_t v = 0;
    for (uint32_t t = t0; t < t1; t++)
        v = umax(v, a.sum_last[(size_t)t * D + d]);
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, 64);
        if (lane >= o)
            inc = umax(inc, y);
    }
    if (lane == 63)
        s_w[w] = inc;
    __syncthreads();
    uint32_t run = __shfl_up(inc, 1, 64);
    if (lane == 0)
        run = 0;
    for (int k = 0; k < w; k++)
        run = umax(run, s_w[k]);
    for (uint32_t t = t0; t < t1; t++) {
        a.in_last[(size_t)t * D + d] = run;
        run = umax(run, a.sum_last[(size_t)t * D + d]);
    }
    __syncthreads();

    // backward: exclusive suffix min
    v = OL_NONE;
    for (uint32_t t = t0; t < t1; t++)
        v = umin(v, a.sum_first[(size_t)t * D + d]);
    inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_down(inc, o, 64);
        if (lane + o < 64)
            inc = umin(inc, y);
    }
    if (lane == 0)
        s_w[w] = inc;
    __syncthreads();
    run = __shfl_down(inc, 1, 64);
    if (lane == 63)
        run = OL_NONE;
    for (int k = w + 1; k < NWS; k++)
        run = umin(run, s_w[k]);
    for (uint32_t t = t1; t-- > t0;) {
        a.out_first[(size_t)t * D + d] = run;
        run = umin(run, a.sum_first[(size_t)t * D + d]);
    }
}

__global__ __launch_bounds__(TW) void tree_walk_kernel(TreeArgs a)
{
    constexpr int NWV = TW / 64;
    __shared__ uint64_t s_m[NWV][TR][D + 1];   // per round: lane d's ballot(key < d)
    __shared__ uint32_t s_b[NWV][TR][D];       // per round: first line per level (then: B after the round)
    __shared__ uint32_t s_l[NWV][D], s_f[NWV][D];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t t = blockIdx.x;
    const uint64_t L = a.lines;
    const uint32_t wbase = t * TLINES + w * (TR * 64);
    uint32_t key[TR];
    load_keys(a.keys, wbase, L, key);

    // pass 1: masks and per-round first line per level
    uint32_t last = 0, first = OL_NONE;
#pragma unroll
    for (int r = 0; r < TR; r++) {
        const uint64_t my = level_masks(key[r]);
        if (lane <= D)
            s_m[w][r][lane] = my;
        const uint64_t up = __shfl_down(my, 1, 64);
        const uint64_t eq = up & ~my;
        const uint32_t rb = wbase + r * 64;
        uint32_t f = OL_NONE;
        if (lane < D && eq) {
            last = rb + 64 - (uint32_t)__clzll(eq);
            f = rb + (uint32_t)__builtin_ctzll(eq);
            if (first == OL_NONE)
                first = f;
        }
        if (lane < D)
            s_b[w][r][lane] = f;
    }
    if (lane < D) {
        s_l[w][lane] = last;
        s_f[w][lane] = first;
    }
    __syncthreads();

    // states at the wave's first and last line
    uint32_t F = 0, Bw = OL_NONE;
    if (lane < D) {
        F = a.in_last[(size_t)t * D + lane];
        Bw = a.out_first[(size_t)t * D + lane];
        for (int k = 0; k < w; k++)
            F = umax(F, s_l[k][lane]);
        for (int k = w + 1; k < NWV; k++)
            Bw = umin(Bw, s_f[k][lane]);
        // B after each round (suffix over the later rounds of this wave)
        uint32_t bc = Bw;
        for (int r = TR - 1; r >= 0; r--) {
            const uint32_t f = s_b[w][r][lane];
            s_b[w][r][lane] = bc;
            bc = umin(bc, f);
        }
    }

    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t above = (lane == 63) ? 0ull : ~((2ull << lane) - 1ull);
    uint32_t *const P = a.parent, *const FC = a.first_child, *const NS = a.next_sibling, *const SE = a.subtree_end;

#pragma unroll
    for (int r = 0; r < TR; r++) {
        const uint32_t rb = wbase + r * 64;
        const uint32_t x = key[r];
        // exclusive prefix max over levels of F (value = line + 1, with its level)
        uint32_t pv = (lane < D) ? F : 0u, pl = (uint32_t)lane;
        // inclusive prefix min over levels of B after this round
        uint32_t sv = (lane < D) ? s_b[w][r][lane] : OL_NONE, sl = (uint32_t)lane;
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
        uint32_t pfx = __shfl_up(pv, 1, 64), pfl = __shfl_up(pl, 1, 64);         // max over d < lane
        if (lane == 0)
            pfx = 0;
        const uint32_t nxt = __shfl(sv, D - 1, 64), nxtl = __shfl(sl, D - 1, 64); // first content after the round

        const bool valid = x < (uint32_t)D;
        const uint32_t xi = valid ? x : 0u;
        const uint64_t mlt = s_m[w][r][xi], mle = s_m[w][r][xi + 1];
        const uint64_t pm = mlt & below, sm = mle & above;
        const uint64_t fm = __ballot(x != KEY_BLANK) & above;
        const uint32_t jp = pm ? 63u - (uint32_t)__clzll(pm) : 0u;
        const uint32_t js = sm ? (uint32_t)__builtin_ctzll(sm) : 0u;
        const uint32_t jf = fm ? (uint32_t)__builtin_ctzll(fm) : 0u;
        const uint32_t kp = __shfl(x, jp, 64), ks = __shfl(x, js, 64), kf = __shfl(x, jf, 64);
        const uint32_t opv = __shfl(pfx, xi, 64), opl = __shfl(pfl, xi, 64);
        const uint32_t osv = __shfl(sv, xi, 64), osl = __shfl(sl, xi, 64);

        const uint32_t g = rb + lane;
        uint32_t par, se, fc, ns;
        if (x == KEY_BLANK || !valid) {
            par = OL_NONE;
            se = g + 1;
            fc = OL_NONE;
            ns = OL_NONE;
        } else {
            par = pm ? rb + jp : (opv ? opv - 1 : OL_NONE);
            const uint32_t plv = pm ? kp : opl;
            se = sm ? rb + js : (osv != OL_NONE ? osv : (uint32_t)L);
            const uint32_t slv = sm ? ks : osl;
            const uint32_t fi = fm ? rb + jf : nxt;
            const uint32_t fl = fm ? kf : nxtl;
            fc = (fi != OL_NONE && fl > x) ? fi : OL_NONE;
            ns = (se < L && (par == OL_NONE || plv < slv)) ? se : OL_NONE;
        }
        if (g < L) {
            if (P) P[g] = par;
            if (SE) SE[g] = se;
            if (FC) FC[g] = fc;
            if (NS) NS[g] = ns;
        }
        // F after this round: the last line per level
        if (lane < D) {
            const uint64_t eq = s_m[w][r][lane + 1] & ~s_m[w][r][lane];
            if (eq)
                F = rb + 64 - (uint32_t)__clzll(eq);
        }
    }
}

hipError_t ol_launch_tree(const TreeArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(tree_summary_kernel, dim3(a.num_tiles), dim3(TW), 0, s, a);
    hipLaunchKernelGGL((tree_scan_kernel<1024>), dim3(D), dim3(1024), 0, s, a);
    hipLaunchKernelGGL(tree_walk_kernel, dim3(a.num_tiles), dim3(TW), 0, s, a);
    return hipGetLastError();
}

uint32_t ol_tree_tile_lines() { return TLINES; }
