/*
 * oracle.c -- CPU restatement of the reference's parse-and-search semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * load it; the product library (oneline_amd/liboneline_hip.so) never links or
 * calls it, and the product path fails loudly when its HIP build is missing.
 *
 * Every function restates one piece of the reference (Rust, tomet-lang/oneline)
 * from its source; the file:line it follows is given above each function.  The
 * reference cannot be compiled here (no cargo/rustc; its crates are not vendored),
 * so the restatement is pinned instead by:
 *   - the reference's own inline known answers (tests/golden/known_answers.json),
 *   - a second, independently written byte-loop parser (tests/test_oracle.py),
 *     checked exhaustively over the 8-symbol alphabet up to length 6,
 *   - brute-force definitions of the tree relations and of search on small inputs.
 *
 * Build: oracle/Makefile -> oracle/liboneline_oracle.so (plain C, gcc).
 */
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define ORC_NONE 0xFFFFFFFFu

/* NodeKind values, crates/oneline-core/src/lib.rs:41-51 */
enum {
    K_BLANK = 0, K_TEXT = 1, K_HEADING = 2, K_ELEMENT = 3,
    K_LIST = 4, K_CONNECT = 5, K_ORDERED = 6, K_COMMENT = 7
};

/* FlatNode, crates/oneline-core/src/lib.rs:87-96: four u32, 16 bytes. */
typedef struct {
    uint32_t depth, kind, span_start, span_len;
} orc_node;

static int is_blank(uint8_t b) { return b == ' ' || b == '\t'; }

/* classify_line + FIRST_BYTE, crates/oneline-simd/src/line.rs:46-81.
 * `c` is the content (after prefix and blanks, CR/LF trimmed), n >= 1. */
static uint32_t classify(const uint8_t *c, size_t n)
{
    switch (c[0]) {
    case '=': return K_HEADING;
    case ':': return K_CONNECT;
    case '@': return K_ELEMENT;
    case '/': return (n > 1 && c[1] == '/') ? K_COMMENT : K_TEXT;
    case '-':
        if (n > 1 && is_blank(c[1])) return K_LIST;
        if (n > 2 && c[1] == '.' && is_blank(c[2])) return K_ORDERED;
        return K_TEXT;
    default: return K_TEXT;
    }
}

/* parse_line, crates/oneline-simd/src/line.rs:8-44.
 * Trailing run of CR/LF is dropped; depth = number of '|', where blank runs
 * between '|' runs belong to the prefix only when another '|' follows. */
void orc_parse_line(const uint8_t *s, size_t len, uint32_t start, orc_node *o)
{
    size_t eff = len;
    while (eff > 0 && (s[eff - 1] == '\r' || s[eff - 1] == '\n'))
        eff--;
    if (eff == 0) {
        o->depth = 0; o->kind = K_BLANK; o->span_start = start; o->span_len = 0;
        return;
    }
    size_t pos = 0;
    while (pos < eff && s[pos] == '|')
        pos++;
    size_t depth = pos;
    size_t bl = 0;
    while (pos + bl < eff && is_blank(s[pos + bl]))
        bl++;
    while (depth > 0 && pos + bl < eff && s[pos + bl] == '|') {
        pos += bl;
        size_t run = 0;
        while (pos + run < eff && s[pos + run] == '|')
            run++;
        depth += run;
        pos += run;
        bl = 0;
        while (pos + bl < eff && is_blank(s[pos + bl]))
            bl++;
    }
    size_t cs = pos + bl;
    o->depth = (uint32_t)depth;
    o->span_start = start + (uint32_t)cs;
    if (cs == eff) {
        o->kind = K_BLANK;
        o->span_len = 0;
        return;
    }
    o->kind = classify(s + cs, eff - cs);
    o->span_len = (uint32_t)(eff - cs);
}

/* Number of lines: one per '\n' plus one for an unterminated tail.
 * lines_in_chunk, crates/oneline-simd/src/lib.rs:38-44. */
size_t orc_line_count(const uint8_t *t, size_t n)
{
    if (n == 0)
        return 0;
    size_t c = 0;
    for (size_t i = 0; i < n; i++)
        c += (t[i] == '\n');
    return c + (t[n - 1] != '\n');
}

/* scan_line_offsets, crates/oneline-simd/src/lib.rs:49-66 (its test oracle,
 * lib.rs:493-513, is the same split-on-'\n' rule): 0, then p+1 for every
 * '\n' at p < n-1.  Writes min(count, cap) offsets; returns count. */
size_t orc_scan_line_offsets(const uint8_t *t, size_t n, uint32_t *out, size_t cap)
{
    if (n == 0)
        return 0;
    size_t k = 0;
    if (k < cap) out[k] = 0;
    k++;
    for (size_t i = 0; i + 1 < n; i++) {
        if (t[i] == '\n') {
            if (k < cap) out[k] = (uint32_t)(i + 1);
            k++;
        }
    }
    return k;
}

/* Whole-document parse: split on '\n', parse each piece
 * (crates/oneline-simd/src/lib.rs:493-513).  offsets/nodes hold line_count
 * entries (orc_line_count).  Returns the line count. */
size_t orc_parse(const uint8_t *t, size_t n, uint32_t *offsets, orc_node *nodes)
{
    if (n == 0)
        return 0;
    size_t k = 0, s = 0;
    for (size_t i = 0; i < n; i++) {
        if (t[i] == '\n') {
            if (offsets) offsets[k] = (uint32_t)s;
            if (nodes) orc_parse_line(t + s, i + 1 - s, (uint32_t)s, &nodes[k]);
            k++;
            s = i + 1;
        }
    }
    if (s < n) {
        if (offsets) offsets[k] = (uint32_t)s;
        if (nodes) orc_parse_line(t + s, n - s, (uint32_t)s, &nodes[k]);
        k++;
    }
    return k;
}

/* TreeLinks::from_nodes_sequential, crates/oneline-core/src/links.rs:52-81.
 * Monotonic stack over content lines; blank lines get parent NONE and
 * subtree_end i+1.  `kinds`/`depths` are per-line arrays. */
void orc_links_sequential(const uint32_t *depths, const uint32_t *kinds, size_t n,
                          uint32_t *parent, uint32_t *subtree_end)
{
    uint32_t *stack = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    size_t sp = 0;
    for (size_t i = 0; i < n; i++) {
        parent[i] = ORC_NONE;
        subtree_end[i] = (uint32_t)(i + 1);
    }
    for (size_t i = 0; i < n; i++) {
        if (kinds[i] == K_BLANK)
            continue;
        while (sp > 0 && depths[stack[sp - 1]] >= depths[i]) {
            subtree_end[stack[sp - 1]] = (uint32_t)i;
            sp--;
        }
        parent[i] = sp ? stack[sp - 1] : ORC_NONE;
        stack[sp++] = (uint32_t)i;
    }
    for (size_t k = 0; k < sp; k++)
        subtree_end[stack[k]] = (uint32_t)n;
    free(stack);
}

/* first_child / next_sibling.  The reference stores only parent and
 * subtree_end (crates/oneline-core/src/links.rs:23-26); these two follow by
 * closed form from its definitions:
 *   first_child[i]  = first content line j in (i, subtree_end[i]), else NONE
 *   next_sibling[i] = k = subtree_end[i] if k < n and parent[k] == parent[i],
 *                     else NONE
 * Both are NONE for blank lines. */
void orc_links_derived(const uint32_t *kinds, const uint32_t *parent,
                       const uint32_t *subtree_end, size_t n,
                       uint32_t *first_child, uint32_t *next_sibling)
{
    /* next content line after i, computed right to left */
    uint32_t next_content = ORC_NONE;
    for (size_t ii = n; ii-- > 0;) {
        if (kinds[ii] == K_BLANK) {
            first_child[ii] = ORC_NONE;
            next_sibling[ii] = ORC_NONE;
            continue;
        }
        uint32_t se = subtree_end[ii];
        first_child[ii] = (next_content != ORC_NONE && next_content < se) ? next_content : ORC_NONE;
        next_sibling[ii] = (se < n && parent[se] == parent[ii]) ? se : ORC_NONE;
        next_content = (uint32_t)ii;
    }
}

/* All overlapping occurrences of pat in t, ascending; the naive
 * all-positions compare of crates/oneline-simd/src/search.rs:106-114, which
 * the reference uses to pin find_all (search.rs:21-26, 88-99).
 * Empty pattern or text shorter than it -> none.  Writes min(count, cap)
 * offsets; returns count. */
size_t orc_find_all(const uint8_t *t, size_t n, const uint8_t *pat, size_t m,
                    uint64_t *out, size_t cap)
{
    if (m == 0 || n < m)
        return 0;
    size_t k = 0;
    for (size_t i = 0; i + m <= n; i++) {
        if (t[i] == pat[0] && memcmp(t + i, pat, m) == 0) {
            if (k < cap) out[k] = i;
            k++;
        }
    }
    return k;
}

/* Dictionary search: every (start, id) with t[start..start+len_id) == pat_id,
 * sorted by start then id.  This is the closed form of what the reference's
 * ManyPatterns (aho-corasick 1.1.5, MatchKind::Standard, overlapping
 * iteration; crates/oneline-simd/src/search.rs:206-283) reports: every
 * occurrence of every pattern, a duplicate pattern reporting each of its ids
 * (crates/oneline-gpu/src/tests.rs:805-826).  Patterns are given as one byte
 * pool plus per-pattern offset and length.
 *
 * Implementation: patterns sorted lexicographically (shorter prefix first,
 * then id).  From each start the range of patterns that have t[start..start+k)
 * as prefix is narrowed one byte at a time (binary search on byte k-1); the
 * patterns of length exactly k at the front of the range are the matches of
 * length k.  Writes min(count, cap) hits; returns count. */
typedef struct {
    const uint8_t *pool;
    const uint64_t *poff;
    const uint32_t *plen;
} orc_pats;

static orc_pats g_pats;     /* qsort context (single-threaded oracle) */

static int cmp_pat(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    uint32_t lx = g_pats.plen[x], ly = g_pats.plen[y];
    uint32_t l = lx < ly ? lx : ly;
    int c = memcmp(g_pats.pool + g_pats.poff[x], g_pats.pool + g_pats.poff[y], l);
    if (c) return c;
    if (lx != ly) return lx < ly ? -1 : 1;
    return x < y ? -1 : (x > y);
}

size_t orc_many_find_all(const uint8_t *t, size_t n, const uint8_t *pool,
                         const uint64_t *poff, const uint32_t *plen, size_t np,
                         uint64_t *out_off, uint32_t *out_id, size_t cap)
{
    if (np == 0 || n == 0)
        return 0;
    uint32_t *order = (uint32_t *)malloc(np * sizeof(uint32_t));
    for (size_t i = 0; i < np; i++)
        order[i] = (uint32_t)i;
    g_pats.pool = pool; g_pats.poff = poff; g_pats.plen = plen;
    qsort(order, np, sizeof(uint32_t), cmp_pat);
    size_t k = 0, idcap = 64;
    uint32_t *ids = (uint32_t *)malloc(idcap * sizeof(uint32_t));
    for (size_t p = 0; p < n; p++) {
        size_t nid = 0;
        size_t lo = 0, hi = np;
        for (size_t d = 0; lo < hi && p + d < n; d++) {
            /* [lo, hi): patterns whose first d bytes equal t[p..p+d).  Those of
             * length d (a match, recorded at the previous step) sort first. */
            while (lo < hi && plen[order[lo]] == d)
                lo++;
            uint8_t c = t[p + d];
            size_t a = lo, b = hi;
            while (a < b) {   /* first with byte d >= c */
                size_t mid = (a + b) / 2;
                if (pool[poff[order[mid]] + d] < c) a = mid + 1; else b = mid;
            }
            size_t lo2 = a;
            b = hi;
            while (a < b) {   /* first with byte d > c */
                size_t mid = (a + b) / 2;
                if (pool[poff[order[mid]] + d] <= c) a = mid + 1; else b = mid;
            }
            lo = lo2;
            hi = a;
            for (size_t q = lo; q < hi && plen[order[q]] == d + 1; q++) {
                if (nid == idcap) {
                    idcap *= 2;
                    ids = (uint32_t *)realloc(ids, idcap * sizeof(uint32_t));
                }
                ids[nid++] = order[q];
            }
        }
        /* ids of different lengths interleave: sort this position's ids */
        for (size_t a = 1; a < nid; a++) {
            uint32_t v = ids[a];
            size_t b = a;
            while (b > 0 && ids[b - 1] > v) { ids[b] = ids[b - 1]; b--; }
            ids[b] = v;
        }
        for (size_t a = 0; a < nid; a++) {
            if (k < cap) { out_off[k] = p; out_id[k] = ids[a]; }
            k++;
        }
    }
    free(ids); free(order);
    return k;
}

/* ---- the AST bridge's structural pass ----------------------------------------
 * to_tomet_ast_with_links, crates/oneline-tomet/src/lib.rs:20-90, up to the
 * point where it needs tomet's own parser (the per-line blocks and
 * finish/nest_sections, lib.rs:143-310): which lines are literal `|` text,
 * where a bare blank line cuts a line from its element, which lines have a
 * block, which elements get `|` content, and the block each block is attached
 * to by the stitching walk.  Per line:
 *   flags[i]     = 1 literal | 2 has a block | 4 cut by a bare blank line
 *                  | 8 has a `|` child (has_pipe_child) | 16 bare blank line
 *   container[i] = the line whose block receives line i's block, NONE at the
 *                  top level or when line i has no block.
 *
 * char::is_whitespace (Unicode White_Space) at s: the character's length in
 * bytes, 0 when it is not whitespace (or not valid UTF-8 there). */
static size_t orc_utf8_space(const uint8_t *s, size_t left)
{
    const uint8_t c = s[0];
    if ((c >= 0x09 && c <= 0x0D) || c == 0x20)
        return 1;
    if (c == 0xC2 && left >= 2 && (s[1] == 0x85 || s[1] == 0xA0))
        return 2;                                   /* U+0085, U+00A0 */
    if (left >= 3) {
        if (c == 0xE1 && s[1] == 0x9A && s[2] == 0x80)
            return 3;                               /* U+1680 */
        if (c == 0xE2 && s[1] == 0x80 &&
            ((s[2] >= 0x80 && s[2] <= 0x8A) || s[2] == 0xA8 || s[2] == 0xA9 || s[2] == 0xAF))
            return 3;                               /* U+2000..200A, U+2028, U+2029, U+202F */
        if (c == 0xE2 && s[1] == 0x81 && s[2] == 0x9F)
            return 3;                               /* U+205F */
        if (c == 0xE3 && s[1] == 0x80 && s[2] == 0x80)
            return 3;                               /* U+3000 */
    }
    return 0;
}

/* content_of(node).trim() is not empty, crates/oneline-core/src/lib.rs:136-140 */
static int orc_has_text(const uint8_t *text, const orc_node *nd)
{
    size_t p = nd->span_start, e = (size_t)nd->span_start + nd->span_len;
    while (p < e) {
        size_t w = orc_utf8_space(text + p, e - p);
        if (!w)
            return 1;
        p += w;
    }
    return 0;
}

void orc_bridge(const uint8_t *text, const orc_node *nodes, const uint32_t *parent, size_t n,
                uint8_t *flags, uint32_t *container)
{
    uint8_t *literal = (uint8_t *)calloc(n ? n : 1, 1);
    uint8_t *pipe = (uint8_t *)calloc(n ? n : 1, 1);
    uint32_t *par = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    /* literal_pipe_lines, lib.rs:96-125 */
    size_t last_blank = (size_t)-1;
    for (size_t i = 0; i < n; i++) {
        par[i] = parent[i];
        if (nodes[i].kind == K_BLANK && nodes[i].depth == 0) {
            last_blank = i;
            continue;
        }
        if (nodes[i].depth == 0)
            continue;
        const uint32_t p = parent[i];
        if (p == ORC_NONE) {
            literal[i] = 1;
        } else if (last_blank != (size_t)-1 && last_blank > p) {
            par[i] = ORC_NONE;
            literal[i] = 1;
        } else {
            literal[i] = nodes[p].kind != K_ELEMENT || literal[p];
        }
    }
    /* has_pipe_child, lib.rs:33-41 */
    for (size_t i = 0; i < n; i++)
        if (!literal[i] && nodes[i].depth > 0 && par[i] != ORC_NONE && par[i] < n)
            pipe[par[i]] = 1;
    /* blocks, lib.rs:44-60: none for Blank and Comment lines and for content
     * that trims to nothing */
    for (size_t i = 0; i < n; i++) {
        const int block = nodes[i].kind != K_BLANK && nodes[i].kind != K_COMMENT && orc_has_text(text, &nodes[i]);
        flags[i] = (uint8_t)(literal[i] | block << 1 | (par[i] != parent[i]) << 2 | pipe[i] << 3 |
                             (nodes[i].kind == K_BLANK && nodes[i].depth == 0) << 4);
        container[i] = ORC_NONE;
    }
    /* the stitching walk, lib.rs:66-86, with line numbers for blocks: from the
     * last line back, a line with a block takes the pending blocks of its
     * children and is itself pending at its parent; a line without one hands
     * its children's on.  Pending lists as linked lists (head, tail, next). */
    uint32_t *head = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *tail = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    uint32_t *next = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    for (size_t i = 0; i < n; i++)
        head[i] = tail[i] = next[i] = ORC_NONE;
    for (size_t ii = n; ii-- > 0;) {
        uint32_t fh, ft;   /* the finished list */
        if (flags[ii] & 2) {
            for (uint32_t c = head[ii]; c != ORC_NONE; c = next[c])
                container[c] = (uint32_t)ii;
            fh = ft = (uint32_t)ii;
            next[ii] = ORC_NONE;
        } else {
            fh = head[ii];
            ft = tail[ii];
        }
        const uint32_t p = par[ii];
        if (fh == ORC_NONE || p == ORC_NONE)
            continue;   /* nothing finished, or it goes to the top level (container NONE) */
        if (head[p] == ORC_NONE) {
            head[p] = fh;
        } else {
            next[tail[p]] = fh;
        }
        tail[p] = ft;
    }
    free(head); free(tail); free(next);
    free(literal); free(pipe); free(par);
}
