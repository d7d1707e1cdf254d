/*
 * Seeded synthetic-system generator for the transcorrelated integral path.
 *
 * Produces the inputs that cutc_int() consumes (reference interface:
 * src/cutc_int.c:15-20, array shapes from src/cutc_module.f90:23-35):
 *   - nuclei, either given or drawn uniformly in a cube with a minimum spacing;
 *   - normalised Cartesian Gaussian AOs (s, p, d), one exponent per shell,
 *     log-uniform in [0.1, 50];
 *   - quadrature points: a uniformly chosen nucleus + isotropic direction x
 *     radius ~ Exp(1 bohr); weights lognormal(mu = -2, sigma = 1);
 *   - AO values (and gradients) on the points, Fortran layout aos(n, n_ao, 4);
 *   - Boys-Handy Jastrow tables c/m/n/o(size_bh = 6, n_nuc).
 *
 * Every stage draws from its own xoshiro256** stream derived from
 * (seed, stream id), so one stage's size never shifts another stage's draws.
 * Host-only C (gcc), no GPU code.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s[4];
} rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void rng_init(rng_t *g, uint64_t seed, uint64_t stream)
{
    uint64_t x = seed ^ (0xD1B54A32D192ED03ull * (stream + 1));
    for (int i = 0; i < 4; i++)
        g->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t rng_next(rng_t *g)
{
    uint64_t *s = g->s;
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
}

/* uniform in the open interval (0, 1) */
static double rng_unif(rng_t *g)
{
    return ((double)(rng_next(g) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static double rng_normal(rng_t *g)
{
    double u1 = rng_unif(g), u2 = rng_unif(g);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

enum { ST_NUC = 1, ST_BASIS = 2, ST_GRID = 16, ST_JAST = 3 };

/* Nuclei uniform in [0, box]^3 with pairwise distance >= min_dist (rejection).
 * Returns 0, or -1 when 1e6 draws did not place them all. */
int synth_nuclei_cube(uint64_t seed, int n_nuc, double box, double min_dist, double *rn)
{
    rng_t g;
    rng_init(&g, seed, ST_NUC);
    int placed = 0;
    long tries = 0;
    while (placed < n_nuc) {
        if (++tries > 1000000)
            return -1;
        double p[3] = {box * rng_unif(&g), box * rng_unif(&g), box * rng_unif(&g)};
        int ok = 1;
        for (int a = 0; a < placed && ok; a++) {
            double dx = p[0] - rn[3 * a], dy = p[1] - rn[3 * a + 1], dz = p[2] - rn[3 * a + 2];
            if (dx * dx + dy * dy + dz * dz < min_dist * min_dist)
                ok = 0;
        }
        if (ok) {
            memcpy(rn + 3 * placed, p, sizeof p);
            placed++;
        }
    }
    return 0;
}

static double dfact_odd(int n) /* (2n-1)!! with (-1)!! = 1 */
{
    double r = 1.0;
    for (int k = 2 * n - 1; k > 1; k -= 2)
        r *= k;
    return r;
}

/* Expand shells (centre, L) into Cartesian AOs, truncating at n_ao.
 * Component order: p = x, y, z; d = xx, xy, xz, yy, yz, zz.
 * One exponent per shell, log-uniform in [0.1, 50].
 * Returns the number of AOs written (== n_ao unless the shells are too few). */
int synth_basis(uint64_t seed, int n_shell, const int *shell_nuc, const int *shell_L, int n_ao,
                int *ao_nuc, int *ao_pow, double *ao_alpha, double *ao_norm)
{
    static const int cart[3][6][3] = {
        {{0, 0, 0}},
        {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}},
        {{2, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 2, 0}, {0, 1, 1}, {0, 0, 2}},
    };
    static const int ncart[3] = {1, 3, 6};
    const double lo = log(0.1), hi = log(50.0);
    rng_t g;
    rng_init(&g, seed, ST_BASIS);
    int k = 0;
    for (int s = 0; s < n_shell && k < n_ao; s++) {
        int L = shell_L[s];
        if (L < 0 || L > 2)
            return -1;
        double alpha = exp(lo + (hi - lo) * rng_unif(&g));
        for (int c = 0; c < ncart[L] && k < n_ao; c++, k++) {
            const int *e = cart[L][c];
            ao_nuc[k] = shell_nuc[s];
            ao_pow[3 * k] = e[0];
            ao_pow[3 * k + 1] = e[1];
            ao_pow[3 * k + 2] = e[2];
            ao_alpha[k] = alpha;
            ao_norm[k] = pow(2.0 * alpha / 3.141592653589793, 0.75) * pow(4.0 * alpha, 0.5 * L) /
                         sqrt(dfact_odd(e[0]) * dfact_odd(e[1]) * dfact_odd(e[2]));
        }
    }
    return k;
}

/* n points: a uniformly chosen nucleus + isotropic unit vector x Exp(1) radius;
 * weights lognormal(-2, 1).  r is (3, n) xyz-interleaved. */
void synth_grid(uint64_t seed, int stream, int n_nuc, const double *rn, int n, double *r, double *w)
{
    rng_t g;
    rng_init(&g, seed, ST_GRID + (uint64_t)stream);
    for (int i = 0; i < n; i++) {
        int a = (int)(rng_unif(&g) * n_nuc);
        if (a >= n_nuc)
            a = n_nuc - 1;
        double cz = 2.0 * rng_unif(&g) - 1.0;
        double ph = 6.283185307179586 * rng_unif(&g);
        double sz = sqrt(fmax(0.0, 1.0 - cz * cz));
        double rad = -log(rng_unif(&g));
        r[3 * i] = rn[3 * a] + rad * sz * cos(ph);
        r[3 * i + 1] = rn[3 * a + 1] + rad * sz * sin(ph);
        r[3 * i + 2] = rn[3 * a + 2] + rad * cz;
        w[i] = exp(-2.0 + rng_normal(&g));
    }
}

static inline double ipow(double x, int e)
{
    double r = 1.0;
    for (int i = 0; i < e; i++)
        r *= x;
    return r;
}

/* AO values and gradients, aos(n, n_ao, 4): aos[p + n*(i + n_ao*c)],
 * c = 0 value, 1..3 d/dx, d/dy, d/dz. */
void synth_aos(int n, const double *r, int n_ao, const double *rn, const int *ao_nuc, const int *ao_pow,
               const double *ao_alpha, const double *ao_norm, double *aos)
{
    const size_t ld = (size_t)n * n_ao;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; p++) {
        for (int i = 0; i < n_ao; i++) {
            const double *A = rn + 3 * ao_nuc[i];
            double x = r[3 * p] - A[0], y = r[3 * p + 1] - A[1], z = r[3 * p + 2] - A[2];
            int ex = ao_pow[3 * i], ey = ao_pow[3 * i + 1], ez = ao_pow[3 * i + 2];
            double al = ao_alpha[i];
            double e = ao_norm[i] * exp(-al * (x * x + y * y + z * z));
            double px = ipow(x, ex), py = ipow(y, ey), pz = ipow(z, ez);
            /* d/dx x^e exp(-a x^2) = (e x^(e-1) - 2 a x^(e+1)) exp(...) */
            double dx = (ex ? ex * ipow(x, ex - 1) : 0.0) - 2.0 * al * px * x;
            double dy = (ey ? ey * ipow(y, ey - 1) : 0.0) - 2.0 * al * py * y;
            double dz = (ez ? ez * ipow(z, ez - 1) : 0.0) - 2.0 * al * pz * z;
            size_t o = (size_t)p + (size_t)n * i;
            aos[o] = e * px * py * pz;
            aos[o + ld] = e * dx * py * pz;
            aos[o + 2 * ld] = e * px * dy * pz;
            aos[o + 3 * ld] = e * px * py * dz;
        }
    }
}

/* Boys-Handy tables, size_bh = 6 terms per nucleus, Fortran (size_bh, n_nuc):
 * (m, n, o) = (0,0,1) (0,0,2) (2,0,0) (3,0,0) (2,2,0) (0,0,3);
 * c(0,0,1) = 0.5, the others U[-0.5, 0.5]. */
void synth_jastrow(uint64_t seed, int n_nuc, double *c, int *m, int *n, int *o)
{
    static const int t[6][3] = {{0, 0, 1}, {0, 0, 2}, {2, 0, 0}, {3, 0, 0}, {2, 2, 0}, {0, 0, 3}};
    rng_t g;
    rng_init(&g, seed, ST_JAST);
    for (int a = 0; a < n_nuc; a++) {
        for (int p = 0; p < 6; p++) {
            int k = p + 6 * a;
            m[k] = t[p][0];
            n[k] = t[p][1];
            o[k] = t[p][2];
            c[k] = (p == 0) ? 0.5 : rng_unif(&g) - 0.5;
        }
    }
}
