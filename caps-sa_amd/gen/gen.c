/*
 * Seeded synthetic texts for the suffix-array workloads.  xoshiro256**
 * seeded by SplitMix64; every generator is a pure function of its arguments.
 *
 *   uniform:  i.i.d. symbols over {0..sigma-1}
 *   markov:   order-3 Markov chain over {0,1,2,3} (each of the 64 context rows
 *             a Dirichlet(1,1,1,1) draw, the first 3 symbols uniform), then
 *             planted repeats until >= repeat_frac of the positions have been
 *             overwritten: destination d uniform in [1,n), source uniform in
 *             [0,d), length Geometric(mean_len) clipped to the text, copied
 *             left to right, each copied symbol replaced with probability
 *             sub_rate by a uniformly chosen different symbol
 *   zipf16:   i.i.d. u16 symbols, value r-1 for rank r ~ Zipf(s) over
 *             `alphabet` ranks
 *   lognormal lengths: exp(ln(median) + sigma*z), rounded, clipped
 *   genome:   segments with lognormal lengths rescaled to the total, each
 *             ending with spacer 4; i.i.d. bases {0..3}; runs of N (5) with
 *             Geometric(n_mean) lengths placed until n_frac is covered
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed) {
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&seed);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r) {
    uint64_t *s = r->s;
    uint64_t res = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}

static inline double rng_u01(rng_t *r) { return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

static inline uint64_t rng_below(rng_t *r, uint64_t m) {
    return (uint64_t)(((unsigned __int128)rng_next(r) * m) >> 64);
}

static double rng_normal(rng_t *r) {
    double u1 = 1.0 - rng_u01(r), u2 = rng_u01(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Geometric on {1,2,...} with the given mean. */
static uint64_t rng_geometric(rng_t *r, double mean) {
    if (mean <= 1.0) return 1;
    double q = 1.0 / mean;
    double u = 1.0 - rng_u01(r); /* (0,1] */
    return 1 + (uint64_t)floor(log(u) / log(1.0 - q));
}

uint64_t caps_gen_next(uint64_t seed, uint64_t k) {
    rng_t r;
    rng_seed(&r, seed);
    uint64_t v = 0;
    for (uint64_t i = 0; i <= k; i++) v = rng_next(&r);
    return v;
}

void caps_gen_uniform_u8(uint8_t *out, uint64_t n, uint32_t sigma, uint64_t seed) {
    rng_t r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = (uint8_t)rng_below(&r, sigma);
}

void caps_gen_uniform_u16(uint16_t *out, uint64_t n, uint32_t sigma, uint64_t seed) {
    rng_t r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < n; i++) out[i] = (uint16_t)rng_below(&r, sigma);
}

/* Returns the number of positions overwritten by planted repeats. */
uint64_t caps_gen_markov_repeats(uint8_t *out, uint64_t n, uint64_t seed, double repeat_frac, double mean_len,
                                 double sub_rate) {
    rng_t r;
    rng_seed(&r, seed);
    double cdf[64][4];
    for (int c = 0; c < 64; c++) {
        double e[4], s = 0;
        for (int j = 0; j < 4; j++) {
            e[j] = -log(1.0 - rng_u01(&r));
            s += e[j];
        }
        double acc = 0;
        for (int j = 0; j < 4; j++) {
            acc += e[j] / s;
            cdf[c][j] = acc;
        }
        cdf[c][3] = 1.0;
    }
    for (uint64_t i = 0; i < n && i < 3; i++) out[i] = (uint8_t)rng_below(&r, 4);
    for (uint64_t i = 3; i < n; i++) {
        uint32_t ctx = ((uint32_t)out[i - 3] << 4) | ((uint32_t)out[i - 2] << 2) | out[i - 1];
        double u = rng_u01(&r);
        int j = 0;
        while (j < 3 && u >= cdf[ctx][j]) j++;
        out[i] = (uint8_t)j;
    }
    if (n < 2 || repeat_frac <= 0) return 0;
    uint64_t words = n / 64 + 1;
    uint64_t *cov = (uint64_t *)calloc(words, 8);
    if (!cov) return 0;
    uint64_t target = (uint64_t)ceil(repeat_frac * (double)n), covered = 0;
    while (covered < target) {
        uint64_t d = 1 + rng_below(&r, n - 1);
        uint64_t len = rng_geometric(&r, mean_len);
        if (len > n - d) len = n - d;
        uint64_t s = rng_below(&r, d);
        for (uint64_t i = 0; i < len; i++) {
            uint8_t c = out[s + i];
            if (rng_u01(&r) < sub_rate) c = (uint8_t)((c + 1 + rng_below(&r, 3)) & 3);
            out[d + i] = c;
            uint64_t p = d + i;
            uint64_t m = 1ull << (p & 63);
            if (!(cov[p >> 6] & m)) {
                cov[p >> 6] |= m;
                covered++;
            }
        }
    }
    free(cov);
    return covered;
}

void caps_gen_zipf_u16(uint16_t *out, uint64_t n, uint32_t alphabet, double s, uint64_t seed) {
    rng_t r;
    rng_seed(&r, seed);
    double *cdf = (double *)malloc(alphabet * sizeof(double));
    double acc = 0;
    for (uint32_t k = 0; k < alphabet; k++) {
        acc += 1.0 / pow((double)(k + 1), s);
        cdf[k] = acc;
    }
    for (uint32_t k = 0; k < alphabet; k++) cdf[k] /= acc;
    cdf[alphabet - 1] = 1.0;
    for (uint64_t i = 0; i < n; i++) {
        double u = rng_u01(&r);
        uint32_t lo = 0, hi = alphabet - 1;
        while (lo < hi) {
            uint32_t mid = (lo + hi) / 2;
            if (u < cdf[mid]) hi = mid;
            else lo = mid + 1;
        }
        out[i] = (uint16_t)lo; /* rank lo+1 -> value lo */
    }
    free(cdf);
}

void caps_gen_lognormal_lengths(uint64_t *out, uint64_t count, double median, double sigma, uint64_t lo,
                                uint64_t hi, uint64_t seed) {
    rng_t r;
    rng_seed(&r, seed);
    for (uint64_t i = 0; i < count; i++) {
        double v = exp(log(median) + sigma * rng_normal(&r));
        uint64_t x = (uint64_t)llround(v);
        if (x < lo) x = lo;
        if (x > hi) x = hi;
        out[i] = x;
    }
}

/* Genome-like segmented text of exactly n symbols in n_seg segments.  ends
 * receives n_seg cumulative ends (the last is n).  Returns the count of
 * positions p with text[p] < 4. */
uint64_t caps_gen_genome(uint8_t *out, uint64_t *ends, uint64_t n, uint64_t n_seg, double median, double sigma,
                         double n_frac, double n_mean, uint64_t seed) {
    rng_t r;
    rng_seed(&r, seed);
    if (n_seg == 0 || n < 2 * n_seg) return 0;
    double *w = (double *)malloc(n_seg * sizeof(double));
    double tot = 0;
    for (uint64_t i = 0; i < n_seg; i++) {
        w[i] = exp(log(median) + sigma * rng_normal(&r));
        tot += w[i];
    }
    uint64_t *len = (uint64_t *)malloc(n_seg * 8), acc = 0, big = 0;
    for (uint64_t i = 0; i < n_seg; i++) {
        len[i] = (uint64_t)floor(w[i] / tot * (double)n);
        if (len[i] < 2) len[i] = 2;
        acc += len[i];
        if (len[i] > len[big]) big = i;
    }
    /* the largest segment absorbs the rounding difference */
    int64_t diff = (int64_t)n - (int64_t)acc;
    if ((int64_t)len[big] + diff < 2) diff = 2 - (int64_t)len[big];
    len[big] = (uint64_t)((int64_t)len[big] + diff);
    acc = 0;
    for (uint64_t i = 0; i < n_seg; i++) {
        acc += len[i];
        ends[i] = acc;
    }
    free(len);
    if (acc != n) {
        free(w);
        return 0;
    }
    free(w);
    uint64_t start = 0;
    for (uint64_t i = 0; i < n_seg; i++) {
        for (uint64_t p = start; p + 1 < ends[i]; p++) out[p] = (uint8_t)rng_below(&r, 4);
        out[ends[i] - 1] = 4;
        start = ends[i];
    }
    uint64_t target = (uint64_t)(n_frac * (double)n), covered = 0;
    while (covered < target) {
        uint64_t len = rng_geometric(&r, n_mean);
        uint64_t d = rng_below(&r, n);
        for (uint64_t p = d; p < n && p < d + len; p++)
            if (out[p] < 4) {
                out[p] = 5;
                covered++;
            }
    }
    uint64_t acgt = 0;
    for (uint64_t p = 0; p < n; p++) acgt += out[p] < 4;
    return acgt;
}
