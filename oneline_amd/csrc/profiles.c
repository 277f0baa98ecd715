/*
 * profiles.c -- the reference's six synthetic document profiles, restated
 * byte for byte: Profile and generate() of crates/oneline-core/src/datagen.rs
 * (lines 8-49 and 152-253), the workloads the reference's own parity tests
 * loop over (crates/oneline-gpu/src/tests.rs:161,305,501).
 *
 * The random source is the xorshift64* generator of datagen.rs:51-82, shared
 * with datagen.c.  Every draw is taken in the order the reference takes it,
 * including the ones a Rust match guard makes (`4 if rng.chance(2)`,
 * `_ if rng.chance(1)`) and the CR draw of every line end, also when its
 * percentage is 0.  Prose's word pick is (u^3 * 4096) with u = k / 2^20 in
 * IEEE double, evaluated left to right with no fused multiply-add
 * (-ffp-contract=off); Nested's section length is pow(10, 1 + 5u) from the C
 * library, truncated.
 *
 * Built with gcc into oneline_amd/libol_datagen.so (host code only).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

enum { PF_REPEAT, PF_SHORT_LINES, PF_MIXED, PF_LONG_LINES, PF_PROSE, PF_NESTED, PF_COUNT };

typedef struct { uint64_t s; } prng_t;

static void prng_init(prng_t *r, uint64_t seed) { r->s = (seed * 0x9E3779B97F4A7C15ull) | 1ull; }

static uint64_t prng_next(prng_t *r)
{
    uint64_t x = r->s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    r->s = x;
    return x * 0x2545F4914F6CDD1Dull;
}

static uint64_t prng_below(prng_t *r, uint64_t n) { return (prng_next(r) >> 11) % n; }
static uint64_t prng_range(prng_t *r, uint64_t lo, uint64_t hi) { return lo + prng_below(r, hi - lo); }
static int prng_chance(prng_t *r, uint64_t percent) { return prng_below(r, 100) < percent; }

/* The fixed word list of the non-Prose profiles (UTF-8 bytes). */
static const char *const k_words[] = {
    "lorem", "ipsum", "dolor", "sit", "amet", "parse", "line", "node", "depth", "tomet", "SIMD", "GPU",
    "\xe7\xb8\xa6\xe6\xa3\x92", "\xe9\x9a\x8e\xe5\xb1\xa4", "\xe6\xa7\x8b\xe6\x96\x87",
    "\xe8\xa7\xa3\xe6\x9e\x90", "na\xc3\xafve", "caf\xc3\xa9", "\xe3\x83\x87\xe3\x83\xbc\xe3\x82\xbf",
    "|pipe|", "a", "of", "the",
};
#define N_WORDS (sizeof(k_words) / sizeof(k_words[0]))

/* Prose vocabulary: 4096 words of 2..10 lowercase letters from Rng(0xC0FFEE). */
#define PROSE_WORDS 4096
typedef struct { char w[PROSE_WORDS][11]; uint8_t len[PROSE_WORDS]; } prose_vocab_t;

static void prose_vocab(prose_vocab_t *v)
{
    prng_t r;
    prng_init(&r, 0xC0FFEEull);
    for (int k = 0; k < PROSE_WORDS; k++) {
        const int len = (int)prng_range(&r, 2, 11);
        for (int c = 0; c < len; c++)
            v->w[k][c] = (char)('a' + prng_below(&r, 26));
        v->w[k][len] = 0;
        v->len[k] = (uint8_t)len;
    }
}

/* Output buffer: appends stop (and flag) at the capacity. */
typedef struct { uint8_t *p; size_t len, cap; int full; } obuf_t;

static void put(obuf_t *b, const char *s, size_t n)
{
    if (b->full || n > b->cap - b->len) {
        b->full = 1;
        return;
    }
    memcpy(b->p + b->len, s, n);
    b->len += n;
}
static void puts0(obuf_t *b, const char *s) { put(b, s, strlen(s)); }

/* Words until target bytes were added (push_text, datagen.rs:126-142). */
static void put_text(obuf_t *b, prng_t *r, size_t target, const prose_vocab_t *v)
{
    const size_t start = b->len;
    while (!b->full && b->len - start < target) {
        if (b->len > start)
            put(b, " ", 1);
        if (v) {
            const double u = (double)prng_below(r, 1u << 20) / (double)(1u << 20);
            const size_t k = (size_t)((u * u * u) * (double)PROSE_WORDS) % PROSE_WORDS;
            put(b, v->w[k], v->len[k]);
        } else {
            puts0(b, k_words[prng_below(r, N_WORDS)]);
        }
    }
}

static void put_prefix(obuf_t *b, uint32_t depth)
{
    for (uint32_t i = 0; i < depth; i++)
        put(b, "|", 1);
    if (depth > 0)
        put(b, " ", 1);
}

static void put_eol(obuf_t *b, prng_t *r, uint64_t crlf_percent)
{
    if (prng_chance(r, crlf_percent))
        put(b, "\r", 1);
    put(b, "\n", 1);
}

static const char k_repeat[] =
    "|=[ Header ]\n| Is this syntax necessary?\n||@card(title: \"x\")\n||:rule(allow:list(card))\n"
    "|| Content.\n|||- Item A\n|||- Item B\n";

/*
 * generate(profile, target_bytes, seed) into out[0, cap): *len gets the
 * document length (at least target_bytes, ending in '\n').  Returns 0, -1 for
 * a bad profile, -2 if cap was too small (a line adds at most 4,100 bytes past
 * target_bytes, so cap >= target_bytes + 8192 always suffices).
 */
int ol_gen_profile(int profile, size_t target, uint64_t seed, uint8_t *out, size_t cap, size_t *len)
{
    obuf_t b = {out, 0, cap, 0};
    *len = 0;
    if (profile < 0 || profile >= PF_COUNT)
        return -1;
    if (profile == PF_REPEAT) {
        while (!b.full && b.len < target)
            put(&b, k_repeat, sizeof(k_repeat) - 1);
        *len = b.len;
        return b.full ? -2 : 0;
    }
    prose_vocab_t local;
    const prose_vocab_t *v = NULL;
    if (profile == PF_PROSE) {
        prose_vocab(&local);
        v = &local;
    }
    prng_t r;
    prng_init(&r, seed);
    uint64_t until_heading = 0, section = 0;
    uint32_t depth = 1;
    while (!b.full && b.len < target) {
        /* depth random walk (datagen.rs:170-176); the chance(2) draw happens
           only on a roll of 4 */
        const uint64_t roll = prng_below(&r, 10);
        if (roll <= 1 && depth > 0)
            depth--;
        else if ((roll == 2 || roll == 3) && depth < 7)
            depth++;
        else if (roll == 4 && prng_chance(&r, 2))
            depth = (uint32_t)prng_below(&r, 3);

        if (profile == PF_NESTED) {
            if (until_heading == 0) {
                char h[48];
                section++;
                snprintf(h, sizeof h, "=[ Section %llu ]", (unsigned long long)section);
                puts0(&b, h);
                put_eol(&b, &r, 0);
                const double u = (double)prng_below(&r, 1u << 20) / (double)(1u << 20);
                until_heading = (uint64_t)pow(10.0, 1.0 + 5.0 * u);
                continue;
            }
            until_heading--;
            if (depth < 1)
                depth = 1;
        }

        size_t body;
        if (profile == PF_SHORT_LINES)
            body = (size_t)prng_range(&r, 4, 24);
        else if (profile == PF_LONG_LINES)
            body = (size_t)prng_range(&r, 200, 1500);
        else if (prng_chance(&r, 1))
            body = (size_t)prng_range(&r, 800, 4000);   /* heavy tail: 1% of lines */
        else
            body = (size_t)prng_range(&r, 5, 120);

        const uint64_t crlf = profile == PF_MIXED ? 2 : 0;
        const uint64_t kind = prng_below(&r, 100);
        if (kind <= 4) {   /* blank, sometimes with a bare prefix */
            if (prng_chance(&r, 30))
                put_prefix(&b, depth);
        } else if (kind <= 9) {
            put_prefix(&b, depth);
            puts0(&b, "=[ ");
            put_text(&b, &r, body < 60 ? body : 60, v);
            puts0(&b, " ]");
        } else if (kind <= 14) {
            put_prefix(&b, depth);
            puts0(&b, ":rule(");
            put_text(&b, &r, body < 60 ? body : 60, v);
            puts0(&b, ")");
        } else if (kind <= 22) {
            put_prefix(&b, depth);
            puts0(&b, "@card(title: \"");
            put_text(&b, &r, body < 30 ? body : 30, v);
            puts0(&b, "\")");
        } else if (kind <= 37) {
            put_prefix(&b, depth);
            puts0(&b, prng_chance(&r, 70) ? "- " : "-. ");
            put_text(&b, &r, body, v);
        } else {
            put_prefix(&b, depth);
            put_text(&b, &r, body, v);
        }
        put_eol(&b, &r, crlf);
    }
    *len = b.len;
    return b.full ? -2 : 0;
}

/* The Prose vocabulary (prose_vocabulary, datagen.rs:113-124): word k at
   out[11 k], NUL-terminated, for k < 4096. */
void ol_gen_prose_vocab(char *out)
{
    prose_vocab_t v;
    prose_vocab(&v);
    memcpy(out, v.w, sizeof v.w);
}
