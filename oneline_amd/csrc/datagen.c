/*
 * datagen.c -- deterministic synthetic markup for the parse-and-search path.
 *
 * The random source is the reference's xorshift64* generator
 * (crates/oneline-core/src/datagen.rs:53-82: state = seed * 0x9E3779B97F4A7C15 | 1,
 * below(n) = (next >> 11) % n).  The document shape is new (the reference's
 * profiles are Repeat/ShortLines/Mixed/...); it is the workload the benchmark is
 * quoted on:
 *
 *   - a document is cut into 1 MiB blocks; block b is generated from its own
 *     stream Rng(mix(seed, b)), so blocks can be produced in parallel and the
 *     bytes do not depend on the thread count;
 *   - per line: depth does a random walk (-1/0/+1 with p = 0.2/0.6/0.2, clamped
 *     to 0..16, starting at 0 in each block) and is written as "| " per level;
 *     kind is uniform over the 7 non-blank markers ("" text, "=", ":", "@",
 *     "- ", "-. ", "//"); the line length without terminator is uniform in
 *     8..120 bytes; the payload is Zipf(s = 1.1) words of a fixed 10,000-word
 *     vocabulary (lowercase, 3..12 letters) joined by single spaces and cut to
 *     fit; a line is empty with probability empty_permille/1000 and ends in
 *     "\r\n" with probability crlf_permille/1000, else "\n";
 *   - a line that would cross the block end is replaced by a plain text line
 *     that fills the block exactly, so every block (and the document) ends in
 *     '\n' and the document is exactly n bytes.
 *
 * Built with gcc into oneline_amd/libol_datagen.so (host code only).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define VOCAB 10000
#define BLOCK (1u << 20)
#define MAX_DEPTH 16

typedef struct { uint64_t s; } rng_t;

static void rng_init(rng_t *r, uint64_t seed) { r->s = (seed * 0x9E3779B97F4A7C15ull) | 1ull; }

static uint64_t rng_next(rng_t *r)
{
    uint64_t x = r->s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    r->s = x;
    return x * 0x2545F4914F6CDD1Dull;
}

static uint64_t rng_below(rng_t *r, uint64_t n) { return (rng_next(r) >> 11) % n; }

/* splitmix64 finaliser: independent stream seeds per (seed, index). */
static uint64_t mix(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (idx + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static char g_words[VOCAB][13];
static uint8_t g_wlen[VOCAB];
static double g_cdf[VOCAB];
#define GUIDE_BITS 16
static uint32_t g_guide[(1u << GUIDE_BITS) + 1];
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void init_tables(void)
{
    rng_t r;
    rng_init(&r, 0x10C0ABull);
    for (int k = 0; k < VOCAB; k++) {
        int len = 3 + (int)rng_below(&r, 10);
        for (int c = 0; c < len; c++)
            g_words[k][c] = (char)('a' + rng_below(&r, 26));
        g_words[k][len] = 0;
        g_wlen[k] = (uint8_t)len;
    }
    double total = 0.0;
    for (int k = 0; k < VOCAB; k++)
        total += 1.0 / pow((double)(k + 1), 1.1);
    double acc = 0.0;
    for (int k = 0; k < VOCAB; k++) {
        acc += 1.0 / pow((double)(k + 1), 1.1);
        g_cdf[k] = acc / total;
    }
    g_cdf[VOCAB - 1] = 1.0;
    /* guide[g] = first rank whose cdf exceeds g / 2^GUIDE_BITS */
    uint32_t k = 0;
    for (uint32_t g = 0; g <= (1u << GUIDE_BITS); g++) {
        double u = (double)g / (double)(1u << GUIDE_BITS);
        while (k < VOCAB - 1 && g_cdf[k] <= u)
            k++;
        g_guide[g] = k;
    }
}

static void ensure_tables(void) { pthread_once(&g_once, init_tables); }

/* Zipf(1.1) rank in 0..VOCAB-1 (0 = most frequent). */
static uint32_t zipf_rank(rng_t *r)
{
    double u = (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
    uint32_t g = (uint32_t)(u * (double)(1u << GUIDE_BITS));
    uint32_t k = g_guide[g];
    while (k < VOCAB - 1 && g_cdf[k] <= u)
        k++;
    return k;
}

/* Writes exactly `len` payload bytes: Zipf words joined by ' ', cut to fit. */
static void payload(rng_t *r, char *out, size_t len)
{
    size_t p = 0;
    while (p < len) {
        if (p > 0) {
            out[p++] = ' ';
            if (p == len)
                break;
        }
        uint32_t w = zipf_rank(r);
        size_t wl = g_wlen[w];
        if (wl > len - p)
            wl = len - p;
        memcpy(out + p, g_words[w], wl);
        p += wl;
    }
}

static const char *const MARKERS[7] = {"", "=", ":", "@", "- ", "-. ", "//"};

static void gen_block(uint8_t *out, size_t n, uint64_t seed, uint32_t crlf_pm, uint32_t empty_pm)
{
    rng_t r;
    rng_init(&r, seed);
    int depth = 0;
    size_t pos = 0;
    char line[2 * MAX_DEPTH + 3 + 128 + 2];
    while (pos < n) {
        uint64_t step = rng_below(&r, 10);
        if (step < 2) {
            if (depth > 0) depth--;
        } else if (step < 4) {
            if (depth < MAX_DEPTH) depth++;
        }
        int empty = rng_below(&r, 1000) < empty_pm;
        int crlf = rng_below(&r, 1000) < crlf_pm;
        size_t ll = 0;
        if (!empty) {
            int kind = (int)rng_below(&r, 7);
            size_t total = 8 + (size_t)rng_below(&r, 113);
            for (int d = 0; d < depth; d++) {
                line[ll++] = '|';
                line[ll++] = ' ';
            }
            size_t ml = strlen(MARKERS[kind]);
            memcpy(line + ll, MARKERS[kind], ml);
            ll += ml;
            size_t body = total > ll ? total - ll : 1;
            payload(&r, line + ll, body);
            ll += body;
        }
        if (crlf)
            line[ll++] = '\r';
        line[ll++] = '\n';
        if (pos + ll <= n) {
            memcpy(out + pos, line, ll);
            pos += ll;
        } else {
            size_t rem = n - pos;
            if (rem >= 2)
                payload(&r, (char *)out + pos, rem - 1);
            out[n - 1] = '\n';
            pos = n;
        }
    }
}

typedef struct {
    uint8_t *out;
    size_t n;
    uint64_t seed;
    uint32_t crlf_pm, empty_pm;
    size_t nblocks;
    size_t next;
    pthread_mutex_t mu;
} job_t;

static void *worker(void *arg)
{
    job_t *j = (job_t *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        size_t b = j->next++;
        pthread_mutex_unlock(&j->mu);
        if (b >= j->nblocks)
            break;
        size_t off = b * (size_t)BLOCK;
        size_t len = j->n - off < BLOCK ? j->n - off : BLOCK;
        gen_block(j->out + off, len, mix(j->seed, b), j->crlf_pm, j->empty_pm);
    }
    return NULL;
}

/* One document of exactly n bytes (see the header for the shape). */
void ol_gen_document(uint8_t *out, size_t n, uint64_t seed, uint32_t crlf_permille,
                     uint32_t empty_permille, int threads)
{
    ensure_tables();
    if (n == 0)
        return;
    job_t j;
    j.out = out; j.n = n; j.seed = seed;
    j.crlf_pm = crlf_permille; j.empty_pm = empty_permille;
    j.nblocks = (n + BLOCK - 1) / BLOCK;
    j.next = 0;
    pthread_mutex_init(&j.mu, NULL);
    if (threads < 1) threads = 1;
    if ((size_t)threads > j.nblocks) threads = (int)j.nblocks;
    if (threads == 1) {
        worker(&j);
    } else {
        pthread_t *t = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
        for (int i = 0; i < threads; i++)
            pthread_create(&t[i], NULL, worker, &j);
        for (int i = 0; i < threads; i++)
            pthread_join(t[i], NULL);
        free(t);
    }
    pthread_mutex_destroy(&j.mu);
}

/* Vocabulary word k (0 = most frequent under the Zipf draw); returns its length. */
int ol_gen_vocab_word(uint32_t k, char *buf)
{
    ensure_tables();
    if (k >= VOCAB)
        return -1;
    memcpy(buf, g_words[k], g_wlen[k]);
    return g_wlen[k];
}

/* A dictionary of `count` patterns, 3..32 bytes.  Even indices come from the
 * payload vocabulary (a uniformly drawn word, or with probability 1/2 a
 * bigram "w1 w2", cut to 32 bytes); odd indices are uniform random lowercase
 * strings of uniform length 3..32.  pool must hold 32 * count bytes; offs and
 * lens get each pattern's place in it. */
void ol_gen_patterns(size_t count, uint64_t seed, uint8_t *pool, uint64_t *offs, uint32_t *lens)
{
    ensure_tables();
    rng_t r;
    rng_init(&r, seed);
    size_t at = 0;
    for (size_t i = 0; i < count; i++) {
        char buf[64];
        size_t len = 0;
        if ((i & 1) == 0) {
            uint32_t a = (uint32_t)rng_below(&r, VOCAB);
            memcpy(buf, g_words[a], g_wlen[a]);
            len = g_wlen[a];
            if (rng_below(&r, 2)) {
                uint32_t b = (uint32_t)rng_below(&r, VOCAB);
                buf[len++] = ' ';
                memcpy(buf + len, g_words[b], g_wlen[b]);
                len += g_wlen[b];
            }
            if (len > 32)
                len = 32;
        } else {
            len = 3 + (size_t)rng_below(&r, 30);
            for (size_t c = 0; c < len; c++)
                buf[c] = (char)('a' + rng_below(&r, 26));
        }
        memcpy(pool + at, buf, len);
        offs[i] = at;
        lens[i] = (uint32_t)len;
        at += len;
    }
}

/* A literal of exactly `len` bytes made of Zipf-drawn vocabulary words joined
 * by spaces (kind 0), or uniform random lowercase (kind 1). */
void ol_gen_literal(size_t len, int kind, uint64_t seed, uint8_t *out)
{
    ensure_tables();
    rng_t r;
    rng_init(&r, seed);
    if (kind == 0)
        payload(&r, (char *)out, len);
    else
        for (size_t c = 0; c < len; c++)
            out[c] = (uint8_t)('a' + rng_below(&r, 26));
}

/* Log-normal document sizes: exp(ln(median) + sigma * z), z standard normal
 * (Box-Muller on two draws), clipped to [lo, hi]. */
void ol_gen_lognormal_sizes(size_t count, uint64_t seed, double median, double sigma,
                            uint64_t lo, uint64_t hi, uint64_t *out)
{
    rng_t r;
    rng_init(&r, seed);
    for (size_t i = 0; i < count; i++) {
        double u1 = ((double)(rng_next(&r) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
        double u2 = (double)(rng_next(&r) >> 11) * (1.0 / 9007199254740992.0);
        double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        double v = exp(log(median) + sigma * z);
        uint64_t s = v < (double)lo ? lo : (v > (double)hi ? hi : (uint64_t)v);
        out[i] = s;
    }
}

/* Stream seed of item `idx` derived from `seed` (documents of a batch). */
uint64_t ol_gen_mix(uint64_t seed, uint64_t idx) { return mix(seed, idx); }
