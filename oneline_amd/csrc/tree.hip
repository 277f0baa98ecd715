// tree.hip -- parent / first_child / next_sibling / subtree_end from the
// per-line key bytes.
//
// The reference builds parent + subtree_end with seven passes (extract keys,
// per-run forward/backward summaries, per-level scans, two assign passes;
// crates/oneline-gpu/src/tree.wgsl:127-375).  Here:
//
//   * State.  F[d] = index of the last content line at depth d seen so far.
//     The monotonic stack of TreeLinks::from_nodes_sequential
//     (crates/oneline-core/src/links.rs:52-81) is exactly the set of levels d
//     with F[d] > max(F[0..d)): the lines not yet closed by a later line of
//     depth <= theirs.  F combines across line ranges by element-wise max.
//   * Three launches (reduce-then-scan; no inter-workgroup waiting, which on
//     this multi-XCD part would need cross-L2 traffic per tile):
//       tree_summary: per tile (threads x TREE_RUN lines), F of the tile alone;
//       tree_scan:    per level, the exclusive max-scan of the tile summaries,
//                     i.e. the F in force at each tile's first line;
//       tree_walk:    the links (below).
//   * Walk.  Each lane owns a run of TREE_RUN consecutive lines; an exclusive
//     max-scan over the runs of a tile (one level per pass) gives each run its
//     incoming F.  The lane then replays the monotonic stack over its run: a
//     local stack in LDS for its own lines, and the incoming stack read from F
//     ("virtual" entries, popped by lowering a level bound).  When line i
//     arrives: the previous content line p gets first_child = i unless i
//     closes it; every popped line j gets subtree_end = i; the last popped one
//     is the line directly above i's parent, so it gets next_sibling = i and
//     the others NONE; parent = the top after popping.
//   * Exactly-once writes.  A line's subtree_end / next_sibling are written
//     by whichever lane pops it (possibly in a later tile), its first_child
//     by the lane of the next content line; the last tile closes the lines
//     still open at the end of the document.
//
// Blank lines are outside the tree: parent NONE, subtree_end i + 1 and
// (new outputs) first_child = next_sibling = NONE.  A content line nested
// TREE_D or more levels deep sets the error flag (the reference's
// MAX_TREE_DEPTH contract, crates/oneline-gpu/src/tree.rs:10-12,218-222).
#include "internal.h"

// Loads the R keys of this lane's run (keys past the last line read as blank).
template <int R>
__device__ __forceinline__ void load_run_keys(const uint8_t *keys, uint32_t rbase, uint64_t L, uint8_t (&key)[R])
{
    static_assert(R == 16, "one 16-byte key load per run");
    uint4 v = *(const uint4 *)(keys + rbase);
    uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < R; k++) {
        uint32_t x = (w[k >> 2] >> (8 * (k & 3))) & 0xFF;
        key[k] = (rbase + k < L) ? (uint8_t)x : (uint8_t)KEY_BLANK;
    }
}

// Fills s_run[d][tid] with this run's last local index + 1 at depth d (0 = none);
// returns true if the run holds a content line that is too deep.
template <int THREADS, int R, int D>
__device__ __forceinline__ bool run_summary(const uint8_t (&key)[R], uint16_t (*s_run)[THREADS])
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int d = 0; d < D; d++)
        s_run[d][tid] = 0;
    bool deep = false;
#pragma unroll
    for (int k = 0; k < R; k++) {
        uint32_t x = key[k];
        if (x < (uint32_t)D)
            s_run[x][tid] = (uint16_t)(tid * R + k + 1);
        else if (x != KEY_BLANK)
            deep = true;
    }
    return deep;
}

// Exclusive max-scan of s_run over the runs, one level per pass; s_agg[d] gets
// the tile total.
template <int THREADS, int D>
__device__ __forceinline__ void run_scan(uint16_t (*s_run)[THREADS], uint16_t *s_agg)
{
    constexpr int NW = THREADS / 64;
    constexpr int RPL = THREADS / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int d = wid; d < D; d += NW) {
        uint16_t v[RPL];
        uint32_t lmax = 0;
#pragma unroll
        for (int j = 0; j < RPL; j++) {
            v[j] = s_run[d][lane * RPL + j];
            lmax = lmax > v[j] ? lmax : v[j];
        }
        uint32_t inc = lmax;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(inc, o, 64);
            if (lane >= o)
                inc = inc > y ? inc : y;
        }
        uint32_t run = __shfl_up(inc, 1, 64);
        if (lane == 0)
            run = 0;
#pragma unroll
        for (int j = 0; j < RPL; j++) {
            s_run[d][lane * RPL + j] = (uint16_t)run;
            run = run > v[j] ? run : v[j];
        }
        if (lane == 63)
            s_agg[d] = (uint16_t)inc;
    }
}

template <int THREADS, int R, int D>
__global__ __launch_bounds__(THREADS) void tree_summary_kernel(TreeArgs a)
{
    constexpr int TL = THREADS * R;
    __shared__ uint16_t s_run[D][THREADS];
    __shared__ uint16_t s_agg[D];
    const uint32_t t = blockIdx.x;
    const uint32_t tbase = t * TL;
    uint8_t key[R];
    load_run_keys<R>(a.keys, tbase + threadIdx.x * R, a.lines, key);
    if (run_summary<THREADS, R, D>(key, s_run))
        atomicOr(a.err, 1u);
    __syncthreads();
    // tile total per level: max over runs
    constexpr int NW = THREADS / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int d = wid; d < D; d += NW) {
        uint32_t m = 0;
        for (int j = lane; j < THREADS; j += 64)
            m = m > s_run[d][j] ? m : s_run[d][j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            uint32_t y = __shfl_xor(m, o, 64);
            m = m > y ? m : y;
        }
        if (lane == 0)
            s_agg[d] = (uint16_t)m;
    }
    __syncthreads();
    if (threadIdx.x < D) {
        uint32_t v = s_agg[threadIdx.x];
        a.summary[(size_t)t * D + threadIdx.x] = v ? tbase + v : 0u;
    }
}

// One workgroup per level: exclusive max-scan of the tile summaries.
template <int SCAN_T>
__global__ __launch_bounds__(SCAN_T) void tree_scan_kernel(TreeArgs a, uint32_t D)
{
    const uint32_t d = blockIdx.x;
    const uint32_t T = a.num_tiles;
    __shared__ uint32_t s_w[SCAN_T / 64];
    __shared__ uint32_t s_carry;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0)
        s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < T; base += SCAN_T) {
        uint32_t i = base + tid;
        uint32_t v = (i < T) ? a.summary[(size_t)i * D + d] : 0u;
        uint32_t inc = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(inc, o, 64);
            if (lane >= o)
                inc = inc > y ? inc : y;
        }
        if (lane == 63)
            s_w[wid] = inc;
        __syncthreads();
        uint32_t carry = s_carry, wb = 0, tot = 0;
        for (int w = 0; w < SCAN_T / 64; w++) {
            if (w < wid)
                wb = wb > s_w[w] ? wb : s_w[w];
            tot = tot > s_w[w] ? tot : s_w[w];
        }
        uint32_t excl = __shfl_up(inc, 1, 64);
        if (lane == 0)
            excl = 0;
        excl = excl > wb ? excl : wb;
        excl = excl > carry ? excl : carry;
        if (i < T)
            a.incoming[(size_t)i * D + d] = excl;
        __syncthreads();
        if (tid == 0)
            s_carry = carry > tot ? carry : tot;
        __syncthreads();
    }
}

#define TREE_UNSET 0xFFFFFFFEu   // staging slot not written by this tile (never a valid value)

template <int THREADS, int R, int D>
__global__ __launch_bounds__(THREADS) void tree_walk_kernel(TreeArgs a)
{
    constexpr int TL = THREADS * R;
    static_assert(D <= 64 && TL <= 65535, "level / index width");

    __shared__ uint16_t s_run[D][THREADS];   // exclusive scan over runs (local index + 1)
    __shared__ uint16_t s_agg[D];            // tile aggregate (local index + 1)
    __shared__ uint32_t s_fin[D];            // incoming from earlier tiles (global index + 1)
    __shared__ uint16_t s_stack[R][THREADS]; // local stack: (k << 8) | depth
    // Staged parent / first_child / next_sibling / subtree_end.  Line
    // rel = run * R + k sits at slot k * (THREADS + 1) + run: lanes writing
    // their k-th lines hit consecutive banks, and the padding keeps the
    // line-order copy-out nearly conflict-free.
    constexpr int SLOTS = R * (THREADS + 1);
    __shared__ uint32_t s_out[4][SLOTS];

    const int tid = threadIdx.x;
    const uint32_t t = blockIdx.x;
    const uint64_t L = a.lines;
    const uint32_t tbase = t * TL;
    const uint32_t rbase = tbase + tid * R;

    uint8_t key[R];
    load_run_keys<R>(a.keys, rbase, L, key);
    run_summary<THREADS, R, D>(key, s_run);
    if (tid < D)
        s_fin[tid] = a.incoming[(size_t)t * D + tid];
    for (int i = tid; i < 4 * SLOTS; i += THREADS)
        (&s_out[0][0])[i] = TREE_UNSET;
    __syncthreads();
    run_scan<THREADS, D>(s_run, s_agg);
    __syncthreads();

    uint32_t *const outs[4] = {a.parent, a.first_child, a.next_sibling, a.subtree_end};
    // Lines of this tile are staged in LDS and written out coalesced at the end;
    // lines of earlier tiles (closed from here) are written straight to memory.
    auto slot = [](uint32_t rel) -> uint32_t { return (rel % R) * (THREADS + 1) + rel / R; };
    auto put = [&](int which, uint32_t line, uint32_t v) {
        uint32_t rel = line - tbase;
        if (rel < (uint32_t)TL)
            s_out[which][slot(rel)] = v;
        else if (outs[which])
            outs[which][line] = v;
    };
    enum { P = 0, FC = 1, NS = 2, SE = 3 };

    // Incoming F of this run at level d, as global index + 1 (0 = none).
    auto fin = [&](int d) -> uint32_t {
        uint32_t l = s_run[d][tid];
        return l ? tbase + l : s_fin[d];
    };
    // Top of the incoming stack among levels < vb: (depth, index + 1).
    auto vtop = [&](int vb, uint32_t &vd, uint32_t &vj) {
        vd = 0;
        vj = 0;
        for (int d = 0; d < vb; d++) {
            uint32_t f = fin(d);
            if (f > vj) {
                vj = f;
                vd = (uint32_t)d;
            }
        }
    };

    int vb = D;
    uint32_t cvd, cvj;
    vtop(vb, cvd, cvj);
    int sp = 0;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const uint32_t g = rbase + k;
        if (g >= L)
            break;
        const uint32_t x = key[k];
        if (x == KEY_BLANK) {
            put(P, g, OL_NONE);
            put(FC, g, OL_NONE);
            put(NS, g, OL_NONE);
            put(SE, g, g + 1);
            continue;
        }
        if (x >= (uint32_t)D)
            continue;   // too deep: flagged by the summary pass, the build fails
        // previous content line: the stack top
        uint32_t p, pd;
        if (sp > 0) {
            uint32_t e = s_stack[sp - 1][tid];
            p = rbase + (e >> 8);
            pd = e & 0xFF;
        } else {
            p = cvj ? cvj - 1 : OL_NONE;
            pd = cvd;
        }
        if (p != OL_NONE)
            put(FC, p, (pd >= x) ? OL_NONE : g);
        uint32_t pend = OL_NONE;
        while (sp > 0) {
            uint32_t e = s_stack[sp - 1][tid];
            if ((e & 0xFF) < x)
                break;
            uint32_t jj = rbase + (e >> 8);
            put(SE, jj, g);
            if (pend != OL_NONE)
                put(NS, pend, OL_NONE);
            pend = jj;
            sp--;
        }
        uint32_t par;
        if (sp > 0) {
            par = rbase + (s_stack[sp - 1][tid] >> 8);
        } else {
            while (cvj && cvd >= x) {
                uint32_t jj = cvj - 1;
                put(SE, jj, g);
                if (pend != OL_NONE)
                    put(NS, pend, OL_NONE);
                pend = jj;
                vb = (int)cvd;
                vtop(vb, cvd, cvj);
            }
            if (vb > (int)x)
                vb = (int)x;   // the cached top (depth < x) stays valid
            par = cvj ? cvj - 1 : OL_NONE;
        }
        if (pend != OL_NONE)
            put(NS, pend, g);
        put(P, g, par);
        s_stack[sp][tid] = (uint16_t)((k << 8) | x);
        sp++;
    }
    __syncthreads();

    // The last tile closes what is still open at the end of the document.
    if (t == a.num_tiles - 1 && tid == 0) {
        auto ffin = [&](int d) -> uint32_t { return s_agg[d] ? tbase + s_agg[d] : s_fin[d]; };
        int b = D;
        bool first = true;
        while (true) {
            uint32_t vd = 0, vj = 0;
            for (int d = 0; d < b; d++) {
                uint32_t f = ffin(d);
                if (f > vj) {
                    vj = f;
                    vd = (uint32_t)d;
                }
            }
            if (!vj)
                break;
            uint32_t jj = vj - 1;
            put(SE, jj, (uint32_t)L);
            put(NS, jj, OL_NONE);
            if (first)
                put(FC, jj, OL_NONE);
            first = false;
            b = (int)vd;
        }
    }
    __syncthreads();

    // Coalesced copy-out of the staged slots this tile wrote.
#pragma unroll
    for (int w = 0; w < 4; w++) {
        uint32_t *dst = outs[w];
        if (!dst)
            continue;
        for (int i = tid; i < TL; i += THREADS) {
            uint32_t v = s_out[w][slot((uint32_t)i)];
            if (v != TREE_UNSET && tbase + i < L)
                dst[tbase + i] = v;
        }
    }
}

template <int THREADS>
static void launch_tree_cfg(const TreeArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((tree_summary_kernel<THREADS, TREE_RUN, TREE_D>), dim3(a.num_tiles), dim3(THREADS), 0, s, a);
    hipLaunchKernelGGL((tree_scan_kernel<1024>), dim3(TREE_D), dim3(1024), 0, s, a, (uint32_t)TREE_D);
    hipLaunchKernelGGL((tree_walk_kernel<THREADS, TREE_RUN, TREE_D>), dim3(a.num_tiles), dim3(THREADS), 0, s, a);
}

hipError_t ol_launch_tree(const TreeArgs &a, uint32_t threads, hipStream_t s)
{
    switch (threads) {
    case 64:
        launch_tree_cfg<64>(a, s);
        break;
    case 128:
        launch_tree_cfg<128>(a, s);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
