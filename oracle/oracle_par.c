/*
 * oracle_par.c -- multi-threaded CPU restatement of oneline-simd's parallel
 * parse and tree build, the CPU baseline timed beside the GPU path.
 *
 * TEST INFRASTRUCTURE ONLY (like oracle.c): loaded by tests/ and by bench.py's
 * cpu_baseline leg, never by the product library.  Its outputs are checked
 * against the sequential restatement in oracle.c (tests/test_oracle.py).
 *
 *   orc_parse_par  -- the two-pass parallel parse of crates/oneline-simd/src/lib.rs
 *                     (partition_chunks :114-137 cuts the text at line starts,
 *                     lines_in_chunk :38-44 counts each chunk, fill_chunk writes
 *                     each chunk at its prefix), with orc_parse_line per line;
 *   orc_links_par  -- parent and subtree_end as all-nearest-smaller-values in
 *                     three phases, crates/oneline-simd/src/links.rs:1-180
 *                     (block-local stacks, a serial combine of the block
 *                     summaries into a persistent stack, a per-block resolve of
 *                     the lines whose answer lies before their block), then
 *                     first_child / next_sibling by their closed forms;
 *   orc_find_all_par  -- the chunked literal search of crates/oneline-simd/src/search.rs:29-52;
 *   orc_ac_*          -- the overlapping Aho-Corasick DFA search of
 *                     crates/oneline-simd/src/search.rs:204-283 (the
 *                     aho-corasick 1.1.5 crate there, restated here).
 *
 * Threads are plain pthreads, one per chunk / block range.
 */
#define _GNU_SOURCE   /* memmem */
#include <pthread.h>
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define ORC_NONE 0xFFFFFFFFu
#define PENDING (0xFFFFFFFFu - 1u)
#define BLOCK ((size_t)1 << 15)
#define MAX_THREADS 256

typedef struct {
    uint32_t depth, kind, span_start, span_len;
} orc_node;

void orc_parse_line(const uint8_t *s, size_t len, uint32_t start, orc_node *o);

/* Runs fn(arg_i) for i < nt on nt threads (the caller runs the last one). */
static void run_threads(int nt, void *(*fn)(void *), void *args, size_t stride)
{
    pthread_t th[MAX_THREADS];
    int started[MAX_THREADS];
    for (int i = 0; i < nt - 1; i++)
        started[i] = pthread_create(&th[i], NULL, fn, (char *)args + (size_t)i * stride) == 0;
    fn((char *)args + (size_t)(nt - 1) * stride);
    for (int i = 0; i < nt - 1; i++) {
        if (started[i])
            pthread_join(th[i], NULL);
        else
            fn((char *)args + (size_t)i * stride);
    }
}

static int clamp_threads(int nt)
{
    return nt < 1 ? 1 : (nt > MAX_THREADS ? MAX_THREADS : nt);
}

/* ---- parse ---------------------------------------------------------------- */

typedef struct {
    const uint8_t *t;
    size_t n, lo, hi, count, base;
    uint32_t *offsets;
    orc_node *nodes;
} parse_job;

/* lines starting in [lo, hi): lo itself (a line start) and after every '\n'
 * before hi - 1 (the '\n' at hi - 1, if any, ends the chunk's last line) */
static void *count_job(void *p)
{
    parse_job *j = (parse_job *)p;
    size_t c = 0;
    if (j->lo < j->hi) {
        c = 1;
        const uint8_t *q = j->t + j->lo, *e = j->t + j->hi - 1;
        while (q < e) {
            const uint8_t *r = memchr(q, '\n', (size_t)(e - q));
            if (!r)
                break;
            c++;
            q = r + 1;
        }
    }
    j->count = c;
    return NULL;
}

static void *fill_job(void *p)
{
    parse_job *j = (parse_job *)p;
    size_t k = j->base, s = j->lo;
    while (s < j->hi) {
        const uint8_t *r = memchr(j->t + s, '\n', j->n - s);
        const size_t end = r ? (size_t)(r - j->t) + 1 : j->n;   /* exclusive, terminator included */
        if (j->offsets) j->offsets[k] = (uint32_t)s;
        if (j->nodes) orc_parse_line(j->t + s, end - s, (uint32_t)s, &j->nodes[k]);
        k++;
        s = end;
    }
    return NULL;
}

/* Parallel parse into offsets/nodes (orc_line_count entries); returns the
 * line count.  Output identical to orc_parse. */
size_t orc_parse_par(const uint8_t *t, size_t n, uint32_t *offsets, orc_node *nodes, int threads)
{
    if (n == 0)
        return 0;
    int nt = clamp_threads(threads);
    if ((size_t)nt > n / 4096 + 1)
        nt = (int)(n / 4096 + 1);
    parse_job jobs[MAX_THREADS];
    size_t prev = 0;
    for (int i = 0; i < nt; i++) {
        size_t hi = (i == nt - 1) ? n : n / (size_t)nt * (size_t)(i + 1);
        if (hi < prev)
            hi = prev;
        if (hi < n) {   /* move the cut to the next line start */
            const uint8_t *r = memchr(t + hi, '\n', n - hi);
            hi = r ? (size_t)(r - t) + 1 : n;
        }
        jobs[i] = (parse_job){t, n, prev, hi, 0, 0, offsets, nodes};
        prev = hi;
    }
    run_threads(nt, count_job, jobs, sizeof jobs[0]);
    size_t base = 0;
    for (int i = 0; i < nt; i++) {
        jobs[i].base = base;
        base += jobs[i].count;
    }
    run_threads(nt, fill_job, jobs, sizeof jobs[0]);
    return base;
}

/* ---- links ---------------------------------------------------------------- */

typedef struct {
    uint32_t key, idx, link;
} stack_entry;

typedef struct {
    const uint32_t *keys;   /* per line; blank lines have key UINT32_MAX */
    size_t n;
    int or_equal;           /* answer: key <= x (else key < x) */
    uint32_t *res;          /* nearest previous answer, or PENDING / NONE */
    size_t b0, b1;          /* block range of this job */
    uint32_t **summ;        /* per block: its stack after the block, equal keys merged */
    size_t *summ_len;
    const stack_entry *pool;
    const uint32_t *heads;  /* per block: global stack top at the block start */
} ansv_job;

static int ok_key(uint32_t v, uint32_t x, int or_equal)
{
    return or_equal ? v <= x : v < x;
}

static void *ansv_local(void *p)
{
    ansv_job *j = (ansv_job *)p;
    uint32_t *stack = (uint32_t *)malloc(BLOCK * sizeof(uint32_t));
    for (size_t b = j->b0; b < j->b1; b++) {
        const size_t lo = b * BLOCK, hi = lo + BLOCK < j->n ? lo + BLOCK : j->n;
        size_t sp = 0;
        for (size_t i = lo; i < hi; i++) {
            const uint32_t x = j->keys[i];
            while (sp && !ok_key(j->keys[stack[sp - 1]], x, j->or_equal))
                sp--;
            j->res[i] = sp ? stack[sp - 1] : PENDING;
            stack[sp++] = (uint32_t)i;
        }
        /* summary: among equal keys only the latest can answer a later block */
        uint32_t *sm = (uint32_t *)malloc((sp ? sp : 1) * sizeof(uint32_t));
        size_t m = 0;
        for (size_t k = 0; k < sp; k++)
            if (k + 1 == sp || j->keys[stack[k + 1]] != j->keys[stack[k]])
                sm[m++] = stack[k];
        j->summ[b] = sm;
        j->summ_len[b] = m;
    }
    free(stack);
    return NULL;
}

static void *ansv_resolve(void *p)
{
    ansv_job *j = (ansv_job *)p;
    for (size_t b = j->b0; b < j->b1; b++) {
        const size_t lo = b * BLOCK, hi = lo + BLOCK < j->n ? lo + BLOCK : j->n;
        uint32_t q = j->heads[b];
        for (size_t i = lo; i < hi; i++) {
            if (j->res[i] != PENDING)
                continue;
            const uint32_t x = j->keys[i];
            while (q != ORC_NONE && !ok_key(j->pool[q].key, x, j->or_equal))
                q = j->pool[q].link;
            j->res[i] = q == ORC_NONE ? ORC_NONE : j->pool[q].idx;
        }
    }
    return NULL;
}

/* For every i, the nearest j < i with ok(keys[j], keys[i]), else NONE. */
static int nearest_previous(const uint32_t *keys, size_t n, int or_equal, uint32_t *res, int nt)
{
    const size_t nb = (n + BLOCK - 1) / BLOCK;
    if ((size_t)nt > nb)
        nt = (int)(nb ? nb : 1);
    uint32_t **summ = (uint32_t **)calloc(nb ? nb : 1, sizeof(uint32_t *));
    size_t *summ_len = (size_t *)calloc(nb ? nb : 1, sizeof(size_t));
    uint32_t *heads = (uint32_t *)malloc((nb ? nb : 1) * sizeof(uint32_t));
    if (!summ || !summ_len || !heads)
        return -1;
    ansv_job jobs[MAX_THREADS];
    for (int i = 0; i < nt; i++)
        jobs[i] = (ansv_job){keys, n, or_equal, res, nb * (size_t)i / (size_t)nt, nb * (size_t)(i + 1) / (size_t)nt,
                             summ, summ_len, NULL, heads};
    run_threads(nt, ansv_local, jobs, sizeof jobs[0]);
    /* combine: fold each block's summary into the persistent global stack */
    size_t total = 0;
    for (size_t b = 0; b < nb; b++)
        total += summ_len[b];
    stack_entry *pool = (stack_entry *)malloc((total ? total : 1) * sizeof(stack_entry));
    if (!pool)
        return -1;
    size_t used = 0;
    uint32_t head = ORC_NONE;
    for (size_t b = 0; b < nb; b++) {
        heads[b] = head;
        if (summ_len[b]) {
            const uint32_t min_key = keys[summ[b][0]];   /* the block minimum is its stack bottom */
            while (head != ORC_NONE && !ok_key(pool[head].key, min_key, or_equal))
                head = pool[head].link;
            for (size_t k = 0; k < summ_len[b]; k++) {
                pool[used] = (stack_entry){keys[summ[b][k]], summ[b][k], head};
                head = (uint32_t)used++;
            }
        }
    }
    for (int i = 0; i < nt; i++)
        jobs[i].pool = pool;
    run_threads(nt, ansv_resolve, jobs, sizeof jobs[0]);
    for (size_t b = 0; b < nb; b++)
        free(summ[b]);
    free(summ);
    free(summ_len);
    free(heads);
    free(pool);
    return 0;
}

typedef struct {
    const orc_node *nodes;
    size_t n, lo, hi;
    uint32_t *keys, *rkeys;
    const uint32_t *parent_raw, *rev;
    uint32_t *parent, *subtree_end, *first_child, *next_sibling;
    uint32_t first_content;   /* first content line in [lo, hi), or NONE */
    uint32_t next_after;      /* first content line at or after hi, or NONE */
} links_job;

static void *keys_job(void *p)
{
    links_job *j = (links_job *)p;
    for (size_t i = j->lo; i < j->hi; i++) {
        const uint32_t k = j->nodes[i].kind == 0 ? 0xFFFFFFFFu
                                                 : (j->nodes[i].depth < 0xFFFFFFFEu ? j->nodes[i].depth : 0xFFFFFFFEu);
        j->keys[i] = k;
        j->rkeys[j->n - 1 - i] = k;
    }
    j->first_content = ORC_NONE;
    for (size_t i = j->lo; i < j->hi; i++)
        if (j->nodes[i].kind != 0) {
            j->first_content = (uint32_t)i;
            break;
        }
    return NULL;
}

static void *translate_job(void *p)
{
    links_job *j = (links_job *)p;
    for (size_t i = j->lo; i < j->hi; i++) {
        if (j->keys[i] == 0xFFFFFFFFu) {
            j->parent[i] = ORC_NONE;
            j->subtree_end[i] = (uint32_t)(i + 1);
        } else {
            j->parent[i] = j->parent_raw[i];
            const uint32_t r = j->rev[j->n - 1 - i];
            j->subtree_end[i] = r == ORC_NONE ? (uint32_t)j->n : (uint32_t)(j->n - 1 - r);
        }
    }
    return NULL;
}

/* first_child / next_sibling by their closed forms (oracle.c, orc_links_derived),
 * right to left within the range, starting from the first content line after it */
static void *derived_job(void *p)
{
    links_job *j = (links_job *)p;
    uint32_t next_content = j->next_after;
    for (size_t ii = j->hi; ii-- > j->lo;) {
        if (j->keys[ii] == 0xFFFFFFFFu) {
            j->first_child[ii] = ORC_NONE;
            j->next_sibling[ii] = ORC_NONE;
            continue;
        }
        const uint32_t se = j->subtree_end[ii];
        j->first_child[ii] = (next_content != ORC_NONE && next_content < se) ? next_content : ORC_NONE;
        j->next_sibling[ii] = (se < j->n && j->parent[se] == j->parent[ii]) ? se : ORC_NONE;
        next_content = (uint32_t)ii;
    }
    return NULL;
}

/* Parallel parent / subtree_end / first_child / next_sibling of the parsed
 * nodes; any output may be NULL except parent and subtree_end.  Output
 * identical to orc_links_sequential + orc_links_derived.  Returns 0, or -1 if
 * memory ran out. */
int orc_links_par(const orc_node *nodes, size_t n, uint32_t *parent, uint32_t *subtree_end,
                  uint32_t *first_child, uint32_t *next_sibling, int threads)
{
    if (n == 0)
        return 0;
    int nt = clamp_threads(threads);
    if ((size_t)nt > n / 4096 + 1)
        nt = (int)(n / 4096 + 1);
    uint32_t *keys = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *rkeys = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *praw = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *rev = (uint32_t *)malloc(n * sizeof(uint32_t));
    if (!keys || !rkeys || !praw || !rev)
        return -1;
    links_job jobs[MAX_THREADS];
    for (int i = 0; i < nt; i++)
        jobs[i] = (links_job){nodes, n, n * (size_t)i / (size_t)nt, n * (size_t)(i + 1) / (size_t)nt, keys, rkeys,
                              praw, rev, parent, subtree_end, first_child, next_sibling, ORC_NONE, ORC_NONE};
    run_threads(nt, keys_job, jobs, sizeof jobs[0]);
    /* parent: nearest previous strictly smaller key; subtree_end: nearest next
     * key <= own = nearest previous in the reversed keys with <= */
    if (nearest_previous(keys, n, 0, praw, nt) || nearest_previous(rkeys, n, 1, rev, nt))
        return -1;
    run_threads(nt, translate_job, jobs, sizeof jobs[0]);
    if (first_child && next_sibling) {
        uint32_t after = ORC_NONE;
        for (int i = nt; i-- > 0;) {
            jobs[i].next_after = after;
            if (jobs[i].first_content != ORC_NONE)
                after = jobs[i].first_content;
        }
        run_threads(nt, derived_job, jobs, sizeof jobs[0]);
    }
    free(keys);
    free(rkeys);
    free(praw);
    free(rev);
    return 0;
}

/* ---- literal search: find_all_par ----------------------------------------
 * crates/oneline-simd/src/search.rs:29-52: the text in chunks of
 * max(n / (threads * 4), 1 MiB) bytes, each searched with a (len - 1)-byte
 * halo past its end, keeping the starts inside the chunk; every overlapping
 * occurrence.  The per-chunk search is glibc's memmem (the reference uses the
 * memchr crate's memmem: a SIMD prefilter plus Two-Way), restarted one byte
 * after each match. */
#define MIN_CHUNK ((size_t)1 << 20)

typedef struct {
    const uint8_t *t, *pat;
    size_t n, m, start, end;
    uint64_t *hits;
    size_t count, cap;
} lit_job;

static void *lit_job_run(void *p)
{
    lit_job *j = (lit_job *)p;
    const size_t wend = j->end + j->m - 1 < j->n ? j->end + j->m - 1 : j->n;
    const uint8_t *at = j->t + j->start, *lim = j->t + wend;
    while (at + j->m <= lim) {
        const uint8_t *q = (const uint8_t *)memmem(at, (size_t)(lim - at), j->pat, j->m);
        if (!q || (size_t)(q - j->t) >= j->end)
            break;
        if (j->count == j->cap) {
            size_t nc = j->cap ? 2 * j->cap : 1024;
            uint64_t *h = (uint64_t *)realloc(j->hits, nc * sizeof(uint64_t));
            if (!h)
                break;
            j->hits = h;
            j->cap = nc;
        }
        j->hits[j->count++] = (uint64_t)(q - j->t);
        at = q + 1;
    }
    return NULL;
}

/* Every start of `pat` in t (ascending); the first `cap` go to out; returns
 * how many there are. */
size_t orc_find_all_par(const uint8_t *t, size_t n, const uint8_t *pat, size_t m, int threads, uint64_t *out,
                        size_t cap)
{
    if (m == 0 || n < m)
        return 0;
    const int nt = clamp_threads(threads);
    size_t chunk = n / ((size_t)nt * 4);
    if (chunk < MIN_CHUNK)
        chunk = MIN_CHUNK;
    const size_t nj = (n + chunk - 1) / chunk;
    lit_job *jobs = (lit_job *)calloc(nj, sizeof(lit_job));
    for (size_t k = 0; k < nj; k++) {
        jobs[k].t = t;
        jobs[k].pat = pat;
        jobs[k].n = n;
        jobs[k].m = m;
        jobs[k].start = k * chunk;
        jobs[k].end = (k + 1) * chunk < n ? (k + 1) * chunk : n;
    }
    for (size_t k0 = 0; k0 < nj; k0 += (size_t)nt) {
        const int g = (int)((nj - k0) < (size_t)nt ? nj - k0 : (size_t)nt);
        run_threads(g, lit_job_run, jobs + k0, sizeof(lit_job));
    }
    size_t total = 0;
    for (size_t k = 0; k < nj; k++) {
        for (size_t i = 0; i < jobs[k].count; i++, total++)
            if (out && total < cap)
                out[total] = jobs[k].hits[i];
        free(jobs[k].hits);
    }
    free(jobs);
    return total;
}

/* ---- dictionary search: an Aho-Corasick DFA, overlapping ------------------
 * The CPU baseline of crates/oneline-simd/src/search.rs:204-283: the
 * aho-corasick crate (1.1.5) built as a DFA with MatchKind::Standard and
 * searched with find_overlapping_iter over chunks of max(n / (threads * 4),
 * 1 MiB) bytes with a (longest - 1)-byte halo, keeping matches that start in
 * the chunk; every (start, pattern) is reported, duplicates each with its own
 * id.  Restated from the published algorithm (Aho & Corasick 1975): a trie of
 * the patterns, failure links by breadth-first search, the transition table
 * completed into a DFA over byte classes (the bytes the patterns use, plus one
 * class for every other byte), and per state the patterns ending there plus a
 * link to the nearest state on its failure chain that has outputs. */
typedef struct orc_ac {
    uint32_t nstates, nclass, longest;
    uint8_t cls[256];
    uint32_t *delta;       /* nstates x nclass */
    uint32_t *out_first;   /* per state: first index into out_ids (ORC_NONE: none of its own) */
    uint32_t *out_count;
    uint32_t *out_link;    /* nearest state on the failure chain with outputs (0: none) */
    uint32_t *out_ids;
    uint32_t *plen;
} orc_ac;

orc_ac *orc_ac_build(const uint8_t *pool, const uint64_t *poff, const uint32_t *plen, size_t np)
{
    orc_ac *a = (orc_ac *)calloc(1, sizeof(orc_ac));
    if (!a || np == 0)
        return a;
    int used[256] = {0};
    size_t total = 0;
    for (size_t i = 0; i < np; i++) {
        total += plen[i];
        if (plen[i] > a->longest)
            a->longest = plen[i];
        for (uint32_t b = 0; b < plen[i]; b++)
            used[pool[poff[i] + b]] = 1;
    }
    a->nclass = 1;
    for (int b = 0; b < 256; b++)
        a->cls[b] = used[b] ? (uint8_t)a->nclass++ : 0;
    const uint32_t nc = a->nclass;
    const size_t maxs = total + 1;
    uint32_t *trie = (uint32_t *)calloc(maxs * nc, sizeof(uint32_t));   /* 0: no edge (state 0 is the root) */
    uint32_t *fail = (uint32_t *)calloc(maxs, sizeof(uint32_t));
    uint32_t *nown = (uint32_t *)calloc(maxs, sizeof(uint32_t));
    uint32_t *endstate = (uint32_t *)malloc(np * sizeof(uint32_t));
    uint32_t ns = 1;
    for (size_t i = 0; i < np; i++) {
        uint32_t s = 0;
        for (uint32_t b = 0; b < plen[i]; b++) {
            const uint32_t c = a->cls[pool[poff[i] + b]];
            if (!trie[(size_t)s * nc + c])
                trie[(size_t)s * nc + c] = ns++;
            s = trie[(size_t)s * nc + c];
        }
        endstate[i] = s;
        nown[s]++;
    }
    /* patterns per state, in id order */
    uint32_t *first = (uint32_t *)malloc(ns * sizeof(uint32_t));
    uint32_t acc = 0;
    for (uint32_t s = 0; s < ns; s++) {
        first[s] = acc;
        acc += nown[s];
    }
    uint32_t *ids = (uint32_t *)malloc((np ? np : 1) * sizeof(uint32_t));
    uint32_t *fillp = (uint32_t *)calloc(ns, sizeof(uint32_t));
    for (size_t i = 0; i < np; i++) {
        const uint32_t s = endstate[i];
        ids[first[s] + fillp[s]++] = (uint32_t)i;
    }
    /* breadth-first: failure links, the DFA, output links */
    uint32_t *queue = (uint32_t *)malloc(ns * sizeof(uint32_t));
    uint32_t *link = (uint32_t *)calloc(ns, sizeof(uint32_t));
    size_t qh = 0, qt = 0;
    for (uint32_t c = 0; c < nc; c++) {
        const uint32_t t = trie[c];
        if (t) {
            fail[t] = 0;
            queue[qt++] = t;
        }
    }
    while (qh < qt) {
        const uint32_t s = queue[qh++];
        link[s] = nown[fail[s]] ? fail[s] : link[fail[s]];
        for (uint32_t c = 0; c < nc; c++) {
            const uint32_t t = trie[(size_t)s * nc + c];
            if (t) {
                fail[t] = trie[(size_t)fail[s] * nc + c];
                queue[qt++] = t;
            } else {
                trie[(size_t)s * nc + c] = trie[(size_t)fail[s] * nc + c];
            }
        }
    }
    a->nstates = ns;
    a->delta = (uint32_t *)realloc(trie, (size_t)ns * nc * sizeof(uint32_t));
    a->out_first = first;
    a->out_count = (uint32_t *)realloc(nown, ns * sizeof(uint32_t));
    a->out_link = link;
    a->out_ids = ids;
    a->plen = (uint32_t *)malloc(np * sizeof(uint32_t));
    memcpy(a->plen, plen, np * sizeof(uint32_t));
    free(fail);
    free(endstate);
    free(fillp);
    free(queue);
    return a;
}

void orc_ac_free(orc_ac *a)
{
    if (!a)
        return;
    free(a->delta);
    free(a->out_first);
    free(a->out_count);
    free(a->out_link);
    free(a->out_ids);
    free(a->plen);
    free(a);
}

typedef struct {
    const orc_ac *a;
    const uint8_t *t;
    size_t n, start, end;
    int collect;
    uint64_t *off;
    uint32_t *id;
    size_t count, cap;
} ac_job;

static void *ac_job_run(void *p)
{
    ac_job *j = (ac_job *)p;
    const orc_ac *a = j->a;
    const size_t wend = j->end + a->longest - 1 < j->n ? j->end + a->longest - 1 : j->n;
    const uint32_t nc = a->nclass;
    uint32_t s = 0;
    for (size_t i = j->start; i < wend; i++) {
        s = a->delta[(size_t)s * nc + a->cls[j->t[i]]];
        for (uint32_t q = a->out_count[s] ? s : a->out_link[s]; q; q = a->out_link[q]) {
            for (uint32_t k = 0; k < a->out_count[q]; k++) {
                const uint32_t pid = a->out_ids[a->out_first[q] + k];
                const size_t st = i + 1 - a->plen[pid];
                if (st >= j->end)
                    continue;   /* starts in the next chunk */
                if (j->collect) {
                    if (j->count == j->cap) {
                        size_t ncap = j->cap ? 2 * j->cap : 4096;
                        uint64_t *o = (uint64_t *)realloc(j->off, ncap * sizeof(uint64_t));
                        uint32_t *d = (uint32_t *)realloc(j->id, ncap * sizeof(uint32_t));
                        if (o) j->off = o;
                        if (d) j->id = d;
                        if (!o || !d)
                            return NULL;
                        j->cap = ncap;
                    }
                    j->off[j->count] = st;
                    j->id[j->count] = pid;
                }
                j->count++;
            }
        }
    }
    return NULL;
}

/* Every (start, pattern) of the dictionary in t, overlapping, over chunks on
 * `threads` threads; with out_off / out_id the first `cap` are copied in chunk
 * order (within a chunk in the automaton's order: by end position, then along
 * the output chain).  Returns how many there are. */
size_t orc_ac_find_par(const orc_ac *a, const uint8_t *t, size_t n, int threads, uint64_t *out_off,
                       uint32_t *out_id, size_t cap)
{
    if (!a || !a->nstates || n == 0)
        return 0;
    const int nt = clamp_threads(threads);
    size_t chunk = n / ((size_t)nt * 4);
    if (chunk < MIN_CHUNK)
        chunk = MIN_CHUNK;
    const size_t nj = (n + chunk - 1) / chunk;
    ac_job *jobs = (ac_job *)calloc(nj, sizeof(ac_job));
    for (size_t k = 0; k < nj; k++) {
        jobs[k].a = a;
        jobs[k].t = t;
        jobs[k].n = n;
        jobs[k].start = k * chunk;
        jobs[k].end = (k + 1) * chunk < n ? (k + 1) * chunk : n;
        jobs[k].collect = out_off || out_id;
    }
    for (size_t k0 = 0; k0 < nj; k0 += (size_t)nt) {
        const int g = (int)((nj - k0) < (size_t)nt ? nj - k0 : (size_t)nt);
        run_threads(g, ac_job_run, jobs + k0, sizeof(ac_job));
    }
    size_t total = 0;
    for (size_t k = 0; k < nj; k++) {
        for (size_t i = 0; jobs[k].collect && i < jobs[k].count; i++)
            if (total + i < cap) {
                if (out_off) out_off[total + i] = jobs[k].off[i];
                if (out_id) out_id[total + i] = jobs[k].id[i];
            }
        total += jobs[k].count;
        free(jobs[k].off);
        free(jobs[k].id);
    }
    free(jobs);
    return total;
}
