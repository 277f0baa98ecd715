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
//   3. wave 0 publishes the tile's line count and resolves its prefix by a
//      decoupled look-back over the per-tile status words;
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
__device__ __forceinline__ uint32_t nl_nibble(uint32_t w) { return eq_nibble(w, 0x0A0A0A0Au); }

// '\n', '|' and blank (space / tab) masks of 64 LDS bytes.
__device__ __forceinline__ void masks64(const uint4 *src, uint64_t &nl, uint64_t &pipe, uint64_t &blank)
{
    nl = pipe = blank = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint4 v = src[k];
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
        uint64_t qn = 0, qp = 0, qb = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            qn |= (uint64_t)eq_nibble(w[i], 0x0A0A0A0Au) << (4 * i);
            qp |= (uint64_t)eq_nibble(w[i], 0x7C7C7C7Cu) << (4 * i);
            qb |= (uint64_t)(eq_nibble(w[i], 0x20202020u) | eq_nibble(w[i], 0x09090909u)) << (4 * i);
        }
        nl |= qn << (16 * k);
        pipe |= qp << (16 * k);
        blank |= qb << (16 * k);
    }
}

// 64 mask bits starting at bit `b` of a mask array.
__device__ __forceinline__ uint64_t window64(const uint64_t *m, uint32_t b)
{
    uint32_t w = b >> 6, sh = b & 63;
    uint64_t lo = m[w] >> sh;
    return sh ? (lo | (m[w + 1] << (64 - sh))) : lo;
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

// parse_line (crates/oneline-simd/src/line.rs:8-44) for text[s ..= e] from the
// '|' / blank bit masks instead of a byte loop.  With P = '|' bits and
// B = blank bits from the line start: a line starting with '|' has its prefix
// = the leading run of P|B (depth = '|' count in it, content after it); one
// starting with a blank skips the leading blank run at depth 0.  Returns false
// (caller falls back to the byte loop) when the prefix runs past the 64-bit
// window.  `rel` is the bit index of s in the mask arrays.
template <class Src>
__device__ __forceinline__ bool line_head(const uint64_t *pm, const uint64_t *bm, const Src &b, uint32_t rel,
                                          uint32_t s, uint32_t e_incl, ol_node_dev &nd)
{
    uint32_t eff = e_incl + 1;
    while (eff > s) {
        uint32_t c = b(eff - 1);
        if (c != '\n' && c != '\r')
            break;
        eff--;
    }
    if (eff == s) {
        nd.depth = 0; nd.kind = OLK_BLANK; nd.span_start = s; nd.span_len = 0;
        return true;
    }
    const uint32_t len = eff - s;
    uint64_t P = window64(pm, rel), B = window64(bm, rel);
    if (len < 64) {
        const uint64_t lm = (1ull << len) - 1;
        P &= lm;
        B &= lm;
    }
    uint32_t depth = 0, r = 0;
    if (P & 1) {
        const uint64_t nq = ~(P | B);
        if (!nq)
            return false;                       // prefix longer than the window
        r = (uint32_t)__builtin_ctzll(nq);
        depth = (uint32_t)__popcll(P & ((1ull << r) - 1));
    } else if (B & 1) {
        const uint64_t nb = ~B;
        if (!nb)
            return false;
        r = (uint32_t)__builtin_ctzll(nb);
    }
    const uint32_t cs = s + r;
    nd.depth = depth;
    nd.span_start = cs;
    if (cs >= eff) {
        nd.span_start = eff;
        nd.kind = OLK_BLANK;
        nd.span_len = 0;
        return true;
    }
    const uint32_t c0 = b(cs);
    const uint32_t c1 = (cs + 1 < eff) ? b(cs + 1) : 256u;
    const uint32_t c2 = (cs + 2 < eff) ? b(cs + 2) : 256u;
    nd.kind = classify(c0, c1, c2);
    nd.span_len = eff - cs;
    return true;
}

template <int THREADS, int BPT>
__global__ __launch_bounds__(THREADS) void scan_classify_kernel(ScanArgs a)
{
    constexpr int TILE = THREADS * BPT;
    constexpr int HALO = SCAN_HALO;
    constexpr int NW = THREADS / 64;
    constexpr int NM = BPT / 64;   // 64-bit terminator masks per lane
    static_assert(BPT % 64 == 0 && NM >= 1, "BPT");
    static_assert(NW >= 2, "needs two waves");

    constexpr int MW = (HALO + TILE) / 64 + 1;   // mask words (+1 so a window may read past the end)
    __shared__ __align__(16) uint8_t s_text[HALO + TILE];
    __shared__ uint64_t s_pm[MW];   // '|' bits of the staged window
    __shared__ uint64_t s_bm[MW];   // blank bits of the staged window
    __shared__ uint32_t s_wcnt[NW];
    __shared__ int32_t s_wlast[NW];
    __shared__ uint32_t s_tile;
    __shared__ uint64_t s_prefix;
    __shared__ uint32_t s_first;

    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (tid == 0)
        s_tile = atomicAdd(a.tile_ctr, 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    const uint32_t base = t * TILE;

    {
        const uint4 *src = (const uint4 *)(a.text + base);
        uint4 *dst = (uint4 *)(s_text + HALO);
        uint4 v[BPT / 16];
#pragma unroll
        for (int k = 0; k < BPT / 16; k++)
            v[k] = src[k * THREADS + tid];
#pragma unroll
        for (int k = 0; k < BPT / 16; k++)
            dst[k * THREADS + tid] = v[k];
        if (tid < HALO / 16) {
            uint4 h = make_uint4(0, 0, 0, 0);
            if (t > 0)
                h = ((const uint4 *)(a.text + base - HALO))[tid];
            ((uint4 *)s_text)[tid] = h;
        }
    }
    __syncthreads();

    // Lane-contiguous BPT bytes -> NM masks of terminators, plus the '|' and
    // blank masks the line heads are decoded from.
    const uint32_t lo = tid * BPT;
    uint64_t mask[NM];
    uint32_t cnt = 0;
    int32_t mylast = -1;
    {
        const uint4 *mine = (const uint4 *)(s_text + HALO + lo);
        const uint32_t last = a.n - 1;   // always a terminator
        if (tid < HALO / 64) {
            uint64_t nl, pm, bm;
            masks64((const uint4 *)(s_text + 64 * tid), nl, pm, bm);
            s_pm[tid] = pm;
            s_bm[tid] = bm;
        }
        if (tid == 0) {
            s_pm[MW - 1] = 0;
            s_bm[MW - 1] = 0;
        }
#pragma unroll
        for (int j = 0; j < NM; j++) {
            uint64_t mk, pm, bm;
            masks64(mine + 4 * j, mk, pm, bm);
            s_pm[(HALO + lo) / 64 + j] = pm;
            s_bm[(HALO + lo) / 64 + j] = bm;
            const uint32_t b0 = base + lo + 64 * j;
            if (last >= b0 && last < b0 + 64)
                mk |= 1ull << (last - b0);
            mask[j] = mk;
            cnt += (uint32_t)__popcll(mk);
            if (mk)
                mylast = (int32_t)(lo + 64 * j + 63 - __clzll(mk));
        }
    }

    uint32_t ic = cnt;
    int32_t il = mylast;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t c = __shfl_up(ic, o, 64);
        int32_t l = __shfl_up(il, o, 64);
        if (lane >= o) {
            ic += c;
            il = il > l ? il : l;
        }
    }
    int32_t prev_l = __shfl_up(il, 1, 64);
    if (lane == 0)
        prev_l = -1;
    if (lane == 63) {
        s_wcnt[wid] = ic;
        s_wlast[wid] = il;
    }
    __syncthreads();
    uint32_t wbc = 0, total = 0;
    int32_t wbl = -1;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        if (w < wid) {
            wbc += s_wcnt[w];
            wbl = wbl > s_wlast[w] ? wbl : s_wlast[w];
        }
        total += s_wcnt[w];
    }
    const uint32_t excl_c = wbc + ic - cnt;
    const int32_t excl_l = wbl > prev_l ? wbl : prev_l;

    if (wid == 0) {
        uint64_t prefix = 0;
        if (t == 0) {
            if (lane == 0)
                st_relaxed_u64(a.status, (LB_FLAG_INCL << LB_SHIFT) | total);
        } else {
            if (lane == 0)
                st_relaxed_u64(a.status + t, (LB_FLAG_AGG << LB_SHIFT) | total);
            prefix = lookback_sum(a.status, t);
            if (lane == 0)
                st_relaxed_u64(a.status + t, (LB_FLAG_INCL << LB_SHIFT) | (prefix + total));
        }
        if (lane == 0)
            s_prefix = prefix;
    } else if (wid == 1) {
        // Start of the first line that ends in this tile: after the last '\n'
        // before the tile (in the halo, else further back in global memory).
        uint32_t start = 0;
        if (t > 0 && total > 0) {
            uint32_t w4 = ((const uint32_t *)s_text)[lane];   // halo bytes 4*lane .. +3
            uint32_t m = nl_nibble(w4);
            int32_t best = m ? (int32_t)(4 * lane + 31 - __clz(m)) : -1;
            uint64_t found = __ballot(best >= 0);
            if (found) {
                int src = 63 - (int)__clzll(found);
                int32_t p = __shfl(best, src, 64);
                start = base - HALO + (uint32_t)p + 1;
            } else {
                start = find_line_start_global(a.text, base - HALO);
            }
        }
        if (lane == 0)
            s_first = start;
    }
    __syncthreads();

    const uint64_t prefix = s_prefix;
    if (tid == 0 && t == a.num_tiles - 1)
        a.meta[0] = prefix + total;
    if (!cnt)
        return;

    // Lines that start inside the staged window (all but possibly the tile's
    // first) read only LDS; the accessor has no global branch, so the reads
    // compile to ds_read (a mixed accessor makes the compiler select between
    // the two base pointers and emit slower generic-address loads).
    const uint32_t lds_lo = (t > 0) ? base - HALO : 0u;
    auto lds_byte = [&](uint32_t pos) -> uint32_t { return s_text[pos + (uint32_t)HALO - base]; };
    auto any_byte = [&](uint32_t pos) -> uint32_t {
        uint32_t rel = pos + (uint32_t)HALO - base;
        return (rel < (uint32_t)(HALO + TILE)) ? (uint32_t)s_text[rel] : (uint32_t)a.text[pos];
    };
    uint64_t r = prefix + excl_c;
    uint32_t s = (excl_l >= 0) ? base + (uint32_t)excl_l + 1 : s_first;
#pragma unroll
    for (int j = 0; j < NM; j++) {
        uint64_t m = mask[j];
        while (m) {
            int bit = __builtin_ctzll(m);
            m &= m - 1;
            uint32_t e = base + lo + 64 * j + (uint32_t)bit;
            if (r < a.cap) {
                ol_node_dev nd;
                if (s < lds_lo || !line_head(s_pm, s_bm, lds_byte, s + (uint32_t)HALO - base, s, e, nd))
                    nd = (s >= lds_lo) ? parse_span(lds_byte, s, e) : parse_span(any_byte, s, e);
                a.offsets[r] = s;
                a.classes[r] = (uint8_t)nd.kind;
                a.keys[r] = key_of(nd);
                if (a.nodes)
                    a.nodes[r] = nd;
            }
            s = e + 1;
            r++;
        }
    }
}

hipError_t ol_launch_scan(const ScanArgs &a, uint32_t bpt, hipStream_t s)
{
    switch (bpt) {
    case 64:
        hipLaunchKernelGGL((scan_classify_kernel<SCAN_THREADS, 64>), dim3(a.num_tiles), dim3(SCAN_THREADS), 0, s, a);
        break;
    case 128:
        hipLaunchKernelGGL((scan_classify_kernel<SCAN_THREADS, 128>), dim3(a.num_tiles), dim3(SCAN_THREADS), 0, s, a);
        break;
    case 256:
        hipLaunchKernelGGL((scan_classify_kernel<SCAN_THREADS, 256>), dim3(a.num_tiles), dim3(SCAN_THREADS), 0, s, a);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---- parse with caller-given line offsets ------------------------------------
// GpuParser::parse (crates/oneline-gpu/src/parse.rs:104-124 + shader.wgsl:34-134):
// line i spans offsets[i] .. offsets[i+1] (or the text end for the last line).
struct GlobalSrc {
    const uint8_t *g;
    __device__ __forceinline__ uint32_t operator()(uint32_t pos) const { return g[pos]; }
};

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
