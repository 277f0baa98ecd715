// tree.hip -- parent / first_child / next_sibling / subtree_end from the
// per-line key bytes, in one pass over the lines.
//
// The reference builds parent + subtree_end with seven passes (extract keys,
// per-run forward/backward summaries, per-level scans, two assign passes;
// crates/oneline-gpu/src/tree.wgsl:127-375).  Here one kernel does it:
//
//   * State.  F[d] = index of the last content line at depth d seen so far.
//     The monotonic stack of TreeLinks::from_nodes_sequential
//     (crates/oneline-core/src/links.rs:52-81) is exactly the set of levels d
//     with F[d] > max(F[0..d)): the lines not yet closed by a later line of
//     depth <= theirs.  F is combined across tiles by element-wise max, so a
//     decoupled look-back over per-tile F vectors gives every tile the stack
//     in force at its first line.
//   * Tile.  TREE_TILE lines; each lane owns a run of TREE_RUN consecutive
//     lines.  Runs publish their own F, an exclusive max-scan over the runs
//     (one level per pass) gives each run its incoming F.
//   * Walk.  Each lane replays the monotonic stack over its run: a local
//     stack in LDS for its own lines, and the incoming stack read from F
//     ("virtual" entries, popped by lowering a level bound).  When line i
//     arrives: the previous content line p gets first_child = i unless i
//     closes it; every popped line j gets subtree_end = i; the last popped
//     one is the line directly above i's parent, so it gets next_sibling = i
//     and the others NONE; parent = the top after popping.
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

__device__ __forceinline__ void fence_release_agent()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
}

template <int THREADS, int R, int D>
__global__ __launch_bounds__(THREADS) void tree_kernel(TreeArgs a)
{
    constexpr int TL = THREADS * R;
    constexpr int NW = THREADS / 64;
    constexpr int RPL = THREADS / 64;   // runs per lane in the level scan
    static_assert(R == 16, "one 16-byte key load per run");
    static_assert(D <= 64 && TL <= 65535, "level / index width");

    __shared__ uint16_t s_run[D][THREADS];   // last local index + 1 in run, then the exclusive scan
    __shared__ uint16_t s_agg[D];            // tile aggregate (local index + 1)
    __shared__ uint32_t s_fin[D];            // incoming from earlier tiles (global index + 1)
    __shared__ uint16_t s_stack[R][THREADS]; // local stack: (k << 8) | depth
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0)
        s_tile = atomicAdd(a.tile_ctr, 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    if (t >= a.num_tiles)
        return;
    const uint64_t L = a.lines;
    const uint32_t tbase = t * TL;
    const uint32_t rbase = tbase + tid * R;

    uint8_t key[R];
    {
        uint4 v = *(const uint4 *)(a.keys + rbase);
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < R; k++) {
            uint32_t x = (w[k >> 2] >> (8 * (k & 3))) & 0xFF;
            key[k] = (rbase + k < L) ? (uint8_t)x : (uint8_t)KEY_BLANK;
        }
    }

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
    if (deep)
        atomicOr(a.err, 1u);
    __syncthreads();

    // Exclusive max-scan over the runs, one level at a time.
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
    __syncthreads();

    // Publish the tile's F and look back for the F in force before it.
    if (wid == 0) {
        uint32_t aggv = 0;
        if (lane < D)
            aggv = s_agg[lane] ? tbase + s_agg[lane] : 0u;
        if (t == 0) {
            if (lane < D) {
                a.incl[lane] = aggv;
                s_fin[lane] = 0;
            }
            fence_release_agent();
            if (lane == 0)
                st_release_u32(a.flags, 2u);
        } else {
            if (lane < D)
                a.agg[(size_t)t * D + lane] = aggv;
            fence_release_agent();
            if (lane == 0)
                st_release_u32(a.flags + t, 1u);
            uint32_t acc = 0;
            int64_t j = (int64_t)t - 1;
            while (j >= 0) {
                uint32_t f = ld_acquire_u32(a.flags + j);
                f = __shfl(f, 0, 64);
                if (f == 0) {
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                if (lane < D) {
                    uint32_t y = (f == 2u) ? a.incl[(size_t)j * D + lane] : a.agg[(size_t)j * D + lane];
                    acc = acc > y ? acc : y;
                }
                if (f == 2u)
                    break;
                j--;
            }
            if (lane < D) {
                s_fin[lane] = acc;
                a.incl[(size_t)t * D + lane] = acc > aggv ? acc : aggv;
            }
            fence_release_agent();
            if (lane == 0)
                st_release_u32(a.flags + t, 2u);
        }
    }
    __syncthreads();

    uint32_t *const parent = a.parent, *const fc = a.first_child;
    uint32_t *const ns = a.next_sibling, *const se = a.subtree_end;

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
            if (parent) parent[g] = OL_NONE;
            if (fc) fc[g] = OL_NONE;
            if (ns) ns[g] = OL_NONE;
            if (se) se[g] = g + 1;
            continue;
        }
        if (x >= (uint32_t)D)
            continue;   // too deep: flagged above, the build fails
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
        if (p != OL_NONE && fc)
            fc[p] = (pd >= x) ? OL_NONE : g;
        uint32_t pend = OL_NONE;
        while (sp > 0) {
            uint32_t e = s_stack[sp - 1][tid];
            if ((e & 0xFF) < x)
                break;
            uint32_t jj = rbase + (e >> 8);
            if (se) se[jj] = g;
            if (pend != OL_NONE && ns) ns[pend] = OL_NONE;
            pend = jj;
            sp--;
        }
        uint32_t par;
        if (sp > 0) {
            par = rbase + (s_stack[sp - 1][tid] >> 8);
        } else {
            while (cvj && cvd >= x) {
                uint32_t jj = cvj - 1;
                if (se) se[jj] = g;
                if (pend != OL_NONE && ns) ns[pend] = OL_NONE;
                pend = jj;
                vb = (int)cvd;
                vtop(vb, cvd, cvj);
            }
            if (vb > (int)x)
                vb = (int)x;   // the cached top (depth < x) stays valid
            par = cvj ? cvj - 1 : OL_NONE;
        }
        if (pend != OL_NONE && ns)
            ns[pend] = g;
        if (parent)
            parent[g] = par;
        s_stack[sp][tid] = (uint16_t)((k << 8) | x);
        sp++;
    }

    // The last tile closes what is still open at the end of the document.
    if (t == a.num_tiles - 1) {
        __syncthreads();
        if (tid == 0) {
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
                if (se) se[jj] = (uint32_t)L;
                if (ns) ns[jj] = OL_NONE;
                if (first && fc)
                    fc[jj] = OL_NONE;
                first = false;
                b = (int)vd;
            }
        }
    }
}

hipError_t ol_launch_tree(const TreeArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((tree_kernel<TREE_THREADS, TREE_RUN, TREE_D>), dim3(a.num_tiles),
                       dim3(TREE_THREADS), 0, s, a);
    return hipGetLastError();
}
