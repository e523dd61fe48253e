/*
 * synth.c -- seeded synthetic polygons for the clipping workloads.
 *
 * Random numbers: xoshiro256** seeded through splitmix64; a uniform double is
 * (u >> 11) * 2^-53.  Every ring is star-shaped about its centre with
 * strictly increasing angles, hence simple and counter-clockwise.
 *
 * Constructions (sizes and parameters are chosen by the caller):
 *  - random star: angles = n sorted U[0,2pi) draws, radii U[rmin,rmax];
 *  - harmonic star: phases phi_k ~ U[0,2pi) for k = 1..32 are drawn first,
 *    then n sorted angles, then per-vertex jitter U[-0.02,0.02];
 *    r(t) = scale * (1 + 0.3 * sum_k k^-1.2 sin(k t + phi_k) + jitter);
 *  - convex: angles sorted, radius 1 (times scale).
 * A rotation is applied about the centre before translation.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} sc_rng;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void rng_seed(sc_rng *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t rng_next(sc_rng *r) { /* xoshiro256** */
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

static double rng_u01(sc_rng *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

static const double TWO_PI = 6.283185307179586476925286766559;

static int sorted_angles(sc_rng *r, double *th, int n) {
    for (int i = 0; i < n; i++) th[i] = TWO_PI * rng_u01(r);
    qsort(th, (size_t)n, sizeof(double), cmp_double);
    for (int i = 1; i < n; i++)
        if (!(th[i] > th[i - 1])) return -1; /* repeated angle: not a simple ring */
    return 0;
}

static void place(double *xy, int n, const double *th, const double *rad, double cx, double cy,
                  double rot) {
    double cr = cos(rot), sr = sin(rot);
    for (int i = 0; i < n; i++) {
        double x = rad[i] * cos(th[i]), y = rad[i] * sin(th[i]);
        xy[2 * i] = cx + (cr * x - sr * y);
        xy[2 * i + 1] = cy + (sr * x + cr * y);
    }
}

/* 0 on success; -1 if the draw repeated an angle (reseed) */
int segclip_synth_star(double *xy, int n, uint64_t seed, double cx, double cy, double rmin,
                       double rmax, double rot) {
    if (n < 3) return -2;
    sc_rng r;
    rng_seed(&r, seed);
    double *th = malloc(sizeof(double) * (size_t)n), *rad = malloc(sizeof(double) * (size_t)n);
    if (!th || !rad) { free(th); free(rad); return -3; }
    int rc = sorted_angles(&r, th, n);
    for (int i = 0; i < n; i++) rad[i] = rmin + (rmax - rmin) * rng_u01(&r);
    if (rc == 0) place(xy, n, th, rad, cx, cy, rot);
    free(th);
    free(rad);
    return rc;
}

int segclip_synth_harmonic(double *xy, int n, uint64_t seed, double cx, double cy, double scale,
                           double rot) {
    if (n < 3) return -2;
    sc_rng r;
    rng_seed(&r, seed);
    double phi[33];
    for (int k = 1; k <= 32; k++) phi[k] = TWO_PI * rng_u01(&r);
    double amp[33];
    for (int k = 1; k <= 32; k++) amp[k] = pow((double)k, -1.2);
    double *th = malloc(sizeof(double) * (size_t)n), *rad = malloc(sizeof(double) * (size_t)n);
    if (!th || !rad) { free(th); free(rad); return -3; }
    int rc = sorted_angles(&r, th, n);
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int k = 1; k <= 32; k++) s += amp[k] * sin((double)k * th[i] + phi[k]);
        double jit = -0.02 + 0.04 * rng_u01(&r);
        rad[i] = scale * (1.0 + 0.3 * s + jit);
    }
    if (rc == 0) place(xy, n, th, rad, cx, cy, rot);
    free(th);
    free(rad);
    return rc;
}

int segclip_synth_convex(double *xy, int n, uint64_t seed, double cx, double cy, double scale,
                         double rot) {
    if (n < 3) return -2;
    sc_rng r;
    rng_seed(&r, seed);
    double *th = malloc(sizeof(double) * (size_t)n), *rad = malloc(sizeof(double) * (size_t)n);
    if (!th || !rad) { free(th); free(rad); return -3; }
    int rc = sorted_angles(&r, th, n);
    for (int i = 0; i < n; i++) rad[i] = scale;
    if (rc == 0) place(xy, n, th, rad, cx, cy, rot);
    free(th);
    free(rad);
    return rc;
}

/* uniform doubles / normals from the same generator (for per-pair parameters) */
void segclip_synth_uniform(double *out, int n, uint64_t seed) {
    sc_rng r;
    rng_seed(&r, seed);
    for (int i = 0; i < n; i++) out[i] = rng_u01(&r);
}

/* number of edges with x1 == x2 in a closed ring (vertical edges) */
int segclip_count_vertical(const double *xy, int n) {
    int c = 0;
    for (int i = 0; i < n; i++) {
        int j = (i + 1 < n) ? i + 1 : 0;
        if (xy[2 * i] == xy[2 * j]) c++;
    }
    return c;
}
