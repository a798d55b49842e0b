/*
 * gen.c -- deterministic workload generators for the suffix-array benchmarks and
 * tests.  xoshiro256** seeded through splitmix64; every overlay draws from its own
 * stream so that changing one overlay does not shift the others.
 *
 * Workloads (sizes are arguments, the defaults live in libsais_amd/workloads.py):
 *   sagen_iid          i.i.d. symbols from an alphabet (1 MiB and 4 GiB ACGT strings)
 *   sagen_dna_like     order-3 Markov ACGT + tandem repeats + copied segments (256 MiB)
 *   sagen_markov       order-k Markov ACGT with seeded transition tables (serving jobs)
 *   sagen_periodic     period-p string with one mutated byte at n/2
 *   sagen_fibonacci    prefix of the Fibonacci word over {a,b}
 *   sagen_gsa_batch    separator-terminated string set with log-normal lengths
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed, uint64_t stream)
{
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r)
{
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

static inline double rng_unif(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

/* uniform integer in [lo, hi] */
static inline uint64_t rng_range(rng_t *r, uint64_t lo, uint64_t hi)
{
    return lo + (uint64_t)(rng_unif(r) * (double)(hi - lo + 1));
}

static double rng_normal(rng_t *r)
{
    double u1 = rng_unif(r), u2 = rng_unif(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static const uint8_t ACGT[4] = {'A', 'C', 'G', 'T'};

void sagen_iid(uint8_t *out, uint64_t n, uint64_t seed, const uint8_t *alphabet, int k)
{
    rng_t r;
    rng_seed(&r, seed, 0);
    uint64_t i = 0;
    if (k == 4) {
        /* 32 symbols per 64-bit draw */
        for (; i + 32 <= n; i += 32) {
            uint64_t x = rng_next(&r);
            for (int j = 0; j < 32; j++) out[i + j] = alphabet[(x >> (2 * j)) & 3];
        }
    }
    for (; i < n; i++) out[i] = alphabet[rng_next(&r) % (uint64_t)k];
}

/* Order-k Markov chain over ACGT: per context, weights U(0.5,1.5) scaled by the GC
 * target (gc/2 for C and G, (1-gc)/2 for A and T), normalised.  gc < 0 draws the
 * GC target from U(0.35, 0.60). */
static void markov_fill(uint8_t *out, uint64_t n, rng_t *r, int order, double gc)
{
    int nctx = 1 << (2 * order);
    double *cum = (double *)malloc((size_t)nctx * 4 * sizeof(double));
    if (!cum) return;
    if (gc < 0) gc = 0.35 + 0.25 * rng_unif(r);
    double base[4] = {(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2};
    for (int x = 0; x < nctx; x++) {
        double w[4], s = 0;
        for (int b = 0; b < 4; b++) { w[b] = (0.5 + rng_unif(r)) * base[b]; s += w[b]; }
        double c = 0;
        for (int b = 0; b < 4; b++) { c += w[b] / s; cum[x * 4 + b] = c; }
        cum[x * 4 + 3] = 1.0;
    }
    uint32_t ctx = 0, mask = (uint32_t)nctx - 1;
    for (uint64_t i = 0; i < n; i++) {
        double u = rng_unif(r);
        const double *c = &cum[(i < (uint64_t)order ? 0 : ctx) * 4];
        int b = (u >= c[0]) + (u >= c[1]) + (u >= c[2]);
        out[i] = ACGT[b];
        ctx = ((ctx << 2) | (uint32_t)b) & mask;
    }
    free(cum);
}

void sagen_markov(uint8_t *out, uint64_t n, uint64_t seed, int order)
{
    rng_t r;
    rng_seed(&r, seed, 0);
    markov_fill(out, n, &r, order, -1.0);
}

static int base_index(uint8_t c) { return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : 3; }

/* 256 MiB DNA-like workload: order-3 Markov, then 2% tandem repeats (period
 * U{2..200}, copies U{5..50}) and 5% copied segments (length U{1000..50000}, 1%
 * substitutions to a different base), each overlay on its own stream. */
int sagen_dna_like(uint8_t *out, uint64_t n, uint64_t seed)
{
    rng_t r;
    rng_seed(&r, seed, 0);
    markov_fill(out, n, &r, 3, -1.0);

    rng_t tr;
    rng_seed(&tr, seed, 1);
    uint64_t budget = (uint64_t)(0.02 * (double)n), done = 0;
    uint8_t unit[200];
    while (done < budget) {
        uint64_t p = rng_range(&tr, 2, 200), c = rng_range(&tr, 5, 50), len = p * c;
        if (len >= n) break;
        for (uint64_t j = 0; j < p; j++) unit[j] = ACGT[rng_next(&tr) & 3];
        uint64_t pos = rng_range(&tr, 0, n - len);
        for (uint64_t j = 0; j < len; j++) out[pos + j] = unit[j % p];
        done += len;
    }

    rng_t cr;
    rng_seed(&cr, seed, 2);
    budget = (uint64_t)(0.05 * (double)n);
    done = 0;
    uint8_t *tmp = (uint8_t *)malloc(50000);
    if (!tmp) return -2;
    while (done < budget) {
        uint64_t len = rng_range(&cr, 1000, 50000);
        if (len >= n) break;
        uint64_t src = rng_range(&cr, 0, n - len), dst = rng_range(&cr, 0, n - len);
        memcpy(tmp, out + src, len);
        for (uint64_t j = 0; j < len; j++) {
            if (rng_unif(&cr) < 0.01) {
                int b = base_index(tmp[j]);
                tmp[j] = ACGT[(b + (int)rng_range(&cr, 1, 3)) & 3];
            }
        }
        memcpy(out + dst, tmp, len);
        done += len;
    }
    free(tmp);
    return 0;
}

/* T[i] = base[i mod p] with base drawn i.i.d. ACGT; T[n/2] advanced A->C->G->T->A. */
void sagen_periodic(uint8_t *out, uint64_t n, uint64_t period, uint64_t seed)
{
    rng_t r;
    rng_seed(&r, seed, 0);
    uint8_t *base = (uint8_t *)malloc(period);
    if (!base) return;
    for (uint64_t j = 0; j < period; j++) base[j] = ACGT[rng_next(&r) & 3];
    for (uint64_t i = 0; i < n; i++) out[i] = base[i % period];
    if (n > 0) out[n / 2] = ACGT[(base_index(out[n / 2]) + 1) & 3];
    free(base);
}

/* Prefix of the infinite Fibonacci word (fixed point of a->ab, b->a):
 * f_k = f_{k-1} f_{k-2}, built in place by appending the shorter prefix. */
void sagen_fibonacci(uint8_t *out, uint64_t n)
{
    if (n == 0) return;
    out[0] = 'a';
    if (n == 1) return;
    out[1] = 'b';
    uint64_t a = 1, b = 2; /* |f_{k-2}|, |f_{k-1}| ; f_{k-1} = out[0..b) */
    while (b < n) {
        uint64_t len = a;
        if (b + len > n) len = n - b;
        memcpy(out + b, out, len);
        uint64_t nb = b + a;
        a = b;
        b = nb;
    }
}

/* Separator-terminated string set.  Lengths are log-normal with the given median,
 * clipped to [min_len, max_len]; sigma is solved by bisection so that the clipped
 * expected total matches target_total.  A fraction `shared_frac` of strings start
 * with a prefix of length U(1024, 20480) of one global random 20 KiB string.
 * Writes the concatenation (each string followed by a 0 byte) into out (capacity
 * cap) and the start offset of every string into offsets[0..count]; returns the
 * total number of bytes, or -1 if cap is too small. */
int64_t sagen_gsa_batch(uint8_t *out, uint64_t cap, int64_t *offsets, int count, uint64_t seed,
                        double median, double min_len, double max_len, double target_total,
                        double shared_frac)
{
    rng_t lr;
    rng_seed(&lr, seed, 0);
    double *z = (double *)malloc((size_t)count * sizeof(double));
    if (!z) return -2;
    for (int i = 0; i < count; i++) z[i] = rng_normal(&lr);
    double mu = log(median), lo = 0.0, hi = 6.0;
    for (int it = 0; it < 100; it++) {
        double s = 0.5 * (lo + hi), tot = 0;
        for (int i = 0; i < count; i++) {
            double L = exp(mu + s * z[i]);
            tot += L < min_len ? min_len : L > max_len ? max_len : L;
        }
        if (tot < target_total) lo = s; else hi = s;
    }
    double sigma = 0.5 * (lo + hi);

    rng_t pr, sr, cr;
    rng_seed(&pr, seed, 1);
    rng_seed(&sr, seed, 2);
    rng_seed(&cr, seed, 3);
    uint8_t shared[20480];
    for (int j = 0; j < 20480; j++) shared[j] = ACGT[rng_next(&sr) & 3];

    uint64_t pos = 0;
    for (int i = 0; i < count; i++) {
        double L = exp(mu + sigma * z[i]);
        L = L < min_len ? min_len : L > max_len ? max_len : L;
        uint64_t len = (uint64_t)L;
        if (pos + len + 1 > cap) { free(z); return -1; }
        offsets[i] = (int64_t)pos;
        uint64_t pre = 0;
        if (rng_unif(&pr) < shared_frac) {
            pre = rng_range(&pr, 1024, 20480);
            if (pre > len) pre = len;
            memcpy(out + pos, shared, pre);
        }
        for (uint64_t j = pre; j < len; j++) out[pos + j] = ACGT[rng_next(&cr) & 3];
        out[pos + len] = 0;
        pos += len + 1;
    }
    offsets[count] = (int64_t)pos;
    free(z);
    return (int64_t)pos;
}
