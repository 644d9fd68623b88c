/*
 * bd_oracle.c -- FP64 CPU restatement of the reference BD timestep.
 * TEST INFRASTRUCTURE ONLY (see bd_oracle.h).  Every function cites the
 * reference lines it restates.  Build: oracle/Makefile (gcc -O2
 * -ffp-contract=off, i.e. no FMA, like the reference's x86-64 build).
 */
#define _GNU_SOURCE
#include "bd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* growable int list = std::vector<int> of the extended neighbour lists */
typedef struct { int n, cap; int32_t *a; } ilist;

static void il_push(ilist *l, int32_t v) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 8;
        l->a = (int32_t *)realloc(l->a, sizeof(int32_t) * (size_t)l->cap);
    }
    l->a[l->n++] = v;
}

/* erwListe_rem (dm:1398-1417): erase the first occurrence, keep order */
static void il_remove(ilist *l, int32_t v) {
    int i;
    for (i = 0; i < l->n; ++i)
        if (l->a[i] == v) break;
    if (i == l->n) return; /* the reference erases end() here (UB); skip */
    memmove(l->a + i, l->a + i + 1, sizeof(int32_t) * (size_t)(l->n - i - 1));
    l->n--;
}

/* ------------------------------------------------------------------ */
/* mt19937_64 (Matsumoto & Nishimura 2004; C++ [rand.predef]) */
#define MT_NN 312
#define MT_MM 156
typedef struct { uint64_t mt[MT_NN]; int mti; } mt64;

static void mt_seed(mt64 *g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < MT_NN; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->mti = MT_NN;
}

static uint64_t mt_next(mt64 *g) {
    static const uint64_t A = 0xB5026F5AA96619E9ULL;
    static const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    if (g->mti >= MT_NN) {
        int i;
        uint64_t x;
        for (i = 0; i < MT_NN - MT_MM; ++i) {
            x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
            g->mt[i] = g->mt[i + MT_MM] ^ (x >> 1) ^ ((x & 1ULL) ? A : 0ULL);
        }
        for (; i < MT_NN - 1; ++i) {
            x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
            g->mt[i] = g->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ ((x & 1ULL) ? A : 0ULL);
        }
        x = (g->mt[MT_NN - 1] & UM) | (g->mt[0] & LM);
        g->mt[MT_NN - 1] = g->mt[MT_MM - 1] ^ (x >> 1) ^ ((x & 1ULL) ? A : 0ULL);
        g->mti = 0;
    }
    uint64_t x = g->mt[g->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

/* libstdc++ generate_canonical<double,53> over a 64-bit engine: one draw,
   (double)r / 2^64, clamped below 1.  Used by uniform_real_distribution
   (zufall.cpp:21-24). */
static double mt_canonical(mt64 *g) {
    double r = (double)mt_next(g) / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

/* zufall_gleichverteilt_vonbis (zufall.cpp:35-39) */
static double mt_vonbis(mt64 *g, double lo, double hi) {
    return lo + (hi - lo) * mt_canonical(g);
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy
   as 1, 2, 3"); the counter-based noise of the GPU path. */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void bdo_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
        philox_round(c, k);
    }
    out[0] = c[0]; out[1] = c[1]; out[2] = c[2]; out[3] = c[3];
}

/* noise sample: counter (pid, step lo, step hi, 0x4244u), key = seed.
   disc: radius sqrt(5 u1), angle 2 pi u2 (uniform on the disc of radius
   sqrt 5, the distribution of dm:1321-1325 without rejection).
   gauss: Box-Muller with variance var. */
void bdo_philox_noise(uint64_t seed, uint32_t pid, uint64_t step, int gauss,
                      double var, double out[2]) {
    uint32_t ctr[4] = {pid, (uint32_t)step, (uint32_t)(step >> 32), 0x4244u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t w[4];
    bdo_philox4x32(ctr, key, w);
    const double inv53 = 1.0 / 9007199254740992.0;
    double u1 = (double)((((uint64_t)w[1] << 32) | w[0]) >> 11) * inv53;
    double u2 = (double)((((uint64_t)w[3] << 32) | w[2]) >> 11) * inv53;
    double r;
    if (gauss)
        r = sqrt(-2.0 * var * log(1.0 - u1));
    else
        r = sqrt(5.0 * u1);
    double a = 2.0 * M_PI * u2;
    out[0] = r * cos(a);
    out[1] = r * sin(a);
}

/* ------------------------------------------------------------------ */
/* in-place complex DFT of any length n, unnormalised, sign -1 / +1 (FFTW
   sign convention used at dm:600-602): the definition summed directly,
   y[k] = sum_j x[j] w^(jk mod n), w = cs[1] + sign i sn[1].  For the meshes
   with a factor 3 (the reference's default FFT_samples = 384,
   parameter.h:24); O(n^2) per transform. */
static void dft1d(double *re, double *im, int n, int stride, int sign,
                  const double *cs, const double *sn) {
    double *yr = (double *)malloc(sizeof(double) * (size_t)n);
    double *yi = (double *)malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k < n; ++k) {
        double ar = 0.0, ai = 0.0;
        for (int j = 0; j < n; ++j) {
            const int e = (int)(((long long)j * k) % n);
            const double wr = cs[e], wi = sign * sn[e];
            const double xr = re[j * stride], xi = im[j * stride];
            ar += xr * wr - xi * wi;
            ai += xr * wi + xi * wr;
        }
        yr[k] = ar;
        yi[k] = ai;
    }
    for (int k = 0; k < n; ++k) { re[k * stride] = yr[k]; im[k * stride] = yi[k]; }
    free(yr);
    free(yi);
}

/* radix-2 in-place complex FFT, unnormalised, sign -1 / +1 (FFTW sign
   convention used at dm:600-602).  The reference calls FFTW3 (external,
   unpinned); this is the DFT definition evaluated by Cooley-Tukey.  A
   length that is not a power of two goes to dft1d. */
static void fft1d(double *re, double *im, int n, int stride, int sign,
                  const double *cs, const double *sn) {
    if (n & (n - 1)) {
        dft1d(re, im, n, stride, sign, cs, sn);
        return;
    }
    int j = 0;
    for (int i = 0; i < n - 1; ++i) {
        if (i < j) {
            double t = re[i * stride]; re[i * stride] = re[j * stride]; re[j * stride] = t;
            t = im[i * stride]; im[i * stride] = im[j * stride]; im[j * stride] = t;
        }
        int m = n >> 1;
        while (m >= 1 && j >= m) { j -= m; m >>= 1; }
        j += m;
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len) {
            for (int k = 0; k < half; ++k) {
                double wr = cs[k * step], wi = sign * sn[k * step];
                int a = (s + k) * stride, b = (s + k + half) * stride;
                double tr = re[b] * wr - im[b] * wi;
                double ti = re[b] * wi + im[b] * wr;
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] += tr; im[a] += ti;
            }
        }
    }
}

void bdo_fft2d(double *data, int n, int sign) {
    double *cs = (double *)malloc(sizeof(double) * (size_t)n);
    double *sn = (double *)malloc(sizeof(double) * (size_t)n);
    double *re = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *im = (double *)malloc(sizeof(double) * (size_t)n * n);
    for (int k = 0; k < n; ++k) {
        cs[k] = cos(2.0 * M_PI * k / n);
        sn[k] = sin(2.0 * M_PI * k / n); /* w = cs + sign*i*sn */
    }
    for (size_t i = 0; i < (size_t)n * n; ++i) { re[i] = data[2 * i]; im[i] = data[2 * i + 1]; }
    for (int r = 0; r < n; ++r) fft1d(re + (size_t)r * n, im + (size_t)r * n, n, 1, sign, cs, sn);
    for (int c = 0; c < n; ++c) fft1d(re + c, im + c, n, n, sign, cs, sn);
    for (size_t i = 0; i < (size_t)n * n; ++i) { data[2 * i] = re[i]; data[2 * i + 1] = im[i]; }
    free(cs); free(sn); free(re); free(im);
}

/* ------------------------------------------------------------------ */
struct bdo_state {
    bdo_params p;
    /* derived constants, parameter.h:170-191 */
    double zh16, B, h, dq, drho;
    int nc;
    int32_t *git[2];   /* [n][2] */
    double *rel[2];    /* [n][2] */
    double *Fw[2], *Fk[2], *Fn[2]; /* WCA, capillary, noise, [n][2] */
    double *ext_noise[2];
    ilist *nb[2];      /* erwNachbarn1/2, [nc*nc] */
    double *rhox, *Fx, *Fy;        /* Z*Z complex, interleaved */
    double *sinxG, *sinyG;          /* Z*Z */
    mt64 rng;
    int64_t step;
};

static int nsp(const bdo_state *s, int sp) { return sp == 0 ? s->p.n1 : s->p.n2; }

/* iw (am:242-250); out-of-range indices (x == L, reference UB) wrap */
static inline size_t iwz(const bdo_state *s, int i, int j) {
    int Z = s->p.Z;
    if (i >= Z) i -= Z;
    if (j >= Z) j -= Z;
    if (i < 0) i += Z;
    if (j < 0) j += Z;
    return (size_t)i * (size_t)Z + (size_t)j;
}

/* G (dm:1384-1394): discrete screened Green's function, index-wrapped */
static double greens(const bdo_state *s, int j, int k) {
    int Z = s->p.Z;
    double dx = s->h;
    double qx = s->dq * (((int)(j + Z * 0.5)) % Z - 0.5 * Z);
    double qy = s->dq * (((int)(k + Z * 0.5)) % Z - 0.5 * Z);
    if (j == 0 && k == 0 && s->p.lambda > 0.2 * s->p.L) return 0.0;
    return 1.0 / (4 * sin(0.5 * qx * dx) * sin(0.5 * qx * dx) / dx / dx +
                  4 * sin(0.5 * qy * dx) * sin(0.5 * qy * dx) / dx / dx +
                  1.0 / (s->p.lambda * s->p.lambda));
}

bdo_state *bdo_create(const bdo_params *p) {
    bdo_state *s = (bdo_state *)calloc(1, sizeof(bdo_state));
    s->p = *p;
    /* parameter.h:170, 179-191 */
    s->zh16 = pow(2.0, 1.0 / 6.0);
    s->nc = (int)floor(p->L / s->zh16);
    s->B = p->L / ((double)s->nc);
    s->h = p->L / p->Z;
    s->dq = 2.0 * M_PI / p->L;
    s->drho = 1.0 / (s->h * s->h); /* gr:28 */
    for (int sp = 0; sp < 2; ++sp) {
        int n = nsp(s, sp);
        size_t m = (size_t)(n > 0 ? n : 1) * 2;
        s->git[sp] = (int32_t *)calloc(m, sizeof(int32_t));
        s->rel[sp] = (double *)calloc(m, sizeof(double));
        s->Fw[sp] = (double *)calloc(m, sizeof(double));
        s->Fk[sp] = (double *)calloc(m, sizeof(double));
        s->Fn[sp] = (double *)calloc(m, sizeof(double));
        s->ext_noise[sp] = (double *)calloc(m, sizeof(double));
        s->nb[sp] = (ilist *)calloc((size_t)s->nc * s->nc, sizeof(ilist));
    }
    size_t Z2 = (size_t)p->Z * p->Z;
    s->rhox = (double *)calloc(2 * Z2, sizeof(double));
    s->Fx = (double *)calloc(2 * Z2, sizeof(double));
    s->Fy = (double *)calloc(2 * Z2, sizeof(double));
    s->sinxG = (double *)calloc(Z2, sizeof(double));
    s->sinyG = (double *)calloc(Z2, sizeof(double));
    /* kapkraefte_init table (dm:578-594) */
    int Z = p->Z;
    double dx = s->h;
    for (int j = 0; j < Z; ++j)
        for (int k = 0; k < Z; ++k) {
            double q = s->dq * ((j + (int)(Z * 0.5)) % Z - 0.5 * Z);
            s->sinxG[iwz(s, j, k)] = p->kap * sin(q * dx) / dx * greens(s, j, k) / Z / Z;
            q = s->dq * ((k + (int)(Z * 0.5)) % Z - 0.5 * Z);
            s->sinyG[iwz(s, j, k)] = p->kap * sin(q * dx) / dx * greens(s, j, k) / Z / Z;
        }
    mt_seed(&s->rng, p->seed);
    return s;
}

void bdo_destroy(bdo_state *s) {
    if (!s) return;
    for (int sp = 0; sp < 2; ++sp) {
        free(s->git[sp]); free(s->rel[sp]); free(s->Fw[sp]); free(s->Fk[sp]);
        free(s->Fn[sp]); free(s->ext_noise[sp]);
        for (int c = 0; c < s->nc * s->nc; ++c) free(s->nb[sp][c].a);
        free(s->nb[sp]);
    }
    free(s->rhox); free(s->Fx); free(s->Fy); free(s->sinxG); free(s->sinyG);
    free(s);
}

void bdo_derived(const bdo_state *s, double *o) {
    o[0] = s->zh16; o[1] = s->nc; o[2] = s->B; o[3] = s->h; o[4] = s->dq; o[5] = s->drho;
}

int bdo_set_positions(bdo_state *s, int species, int n, const int32_t *git, const double *rel) {
    int sp = species - 1;
    if (sp < 0 || sp > 1 || n != nsp(s, sp)) return -1;
    memcpy(s->git[sp], git, sizeof(int32_t) * 2 * (size_t)n);
    memcpy(s->rel[sp], rel, sizeof(double) * 2 * (size_t)n);
    return 0;
}

int bdo_get_positions(const bdo_state *s, int species, int n, int32_t *git, double *rel) {
    int sp = species - 1;
    if (sp < 0 || sp > 1 || n != nsp(s, sp)) return -1;
    if (git) memcpy(git, s->git[sp], sizeof(int32_t) * 2 * (size_t)n);
    if (rel) memcpy(rel, s->rel[sp], sizeof(double) * 2 * (size_t)n);
    return 0;
}

/* push particle i into the 9 lists around cell (k,l) (dm:487-498) */
static void lists_add(bdo_state *s, int sp, int i, int k, int l) {
    int nc = s->nc;
    int x[3] = {(k - 1 + nc) % nc, k, (k + 1) % nc};
    int y[3] = {(l - 1 + nc) % nc, l, (l + 1) % nc};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) il_push(&s->nb[sp][x[a] * nc + y[b]], i);
}

static void lists_rem(bdo_state *s, int sp, int i, int k, int l) {
    int nc = s->nc;
    int x[3] = {(k - 1 + nc) % nc, k, (k + 1) % nc};
    int y[3] = {(l - 1 + nc) % nc, l, (l + 1) % nc};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) il_remove(&s->nb[sp][x[a] * nc + y[b]], i);
}

/* WCA_init (dm:459-526) */
void bdo_wca_init(bdo_state *s) {
    for (int sp = 0; sp < 2; ++sp) {
        for (int c = 0; c < s->nc * s->nc; ++c) s->nb[sp][c].n = 0;
        for (int i = 0; i < nsp(s, sp); ++i)
            lists_add(s, sp, i, s->git[sp][2 * i], s->git[sp][2 * i + 1]);
    }
}

/* init_zufallspos (dm:1647-1678), incl. the (int)x/B precedence of dm:1655 */
void bdo_init_zufallspos(bdo_state *s) {
    for (int sp = 0; sp < 2; ++sp)
        for (int i = 0; i < nsp(s, sp); ++i) {
            double x = mt_vonbis(&s->rng, 0.0, s->p.L);
            double y = mt_vonbis(&s->rng, 0.0, s->p.L);
            int jc = (int)((int)x / s->B);
            int kc = (int)((int)y / s->B);
            s->git[sp][2 * i] = jc;
            s->git[sp][2 * i + 1] = kc;
            s->rel[sp][2 * i] = x - jc * s->B;
            s->rel[sp][2 * i + 1] = y - kc * s->B;
        }
    bdo_wca_init(s);
}

/* abstand2_{11,12,22} (am:100-166) */
static double abstand2(const bdo_state *s, int spi, int i, int spj, int j) {
    double B = s->B, L = s->p.L;
    double xi = s->git[spi][2 * i] * B + s->rel[spi][2 * i];
    double xj = s->git[spj][2 * j] * B + s->rel[spj][2 * j];
    double yi = s->git[spi][2 * i + 1] * B + s->rel[spi][2 * i + 1];
    double yj = s->git[spj][2 * j + 1] * B + s->rel[spj][2 * j + 1];
    double t1 = xi - xj, t2 = xi - xj + L, t3 = xi - xj - L;
    double m23 = (t3 * t3 < t2 * t2) ? t3 * t3 : t2 * t2;
    double dx2 = (m23 < t1 * t1) ? m23 : t1 * t1;
    t1 = yi - yj; t2 = yi - yj + L; t3 = yi - yj - L;
    m23 = (t3 * t3 < t2 * t2) ? t3 * t3 : t2 * t2;
    double dy2 = (m23 < t1 * t1) ? m23 : t1 * t1;
    return dx2 + dy2;
}

/* pair kernel shared by dm:617-702, 790-879, 973-1059: returns 0 if the
   pair is outside the cutoff, else writes abs_F*dx, abs_F*dy for particle
   i (species spi) against j (species spj); sig = sigma ratio, epsf =
   eps*sigma prefactor (1 for the 11 pair, whose code has no prefactor). */
static int wca_pair(const bdo_state *s, int spi, int i, int spj, int j, double sig,
                    double epsf, int scaled, double *fx, double *fy) {
    double B = s->B, L = s->p.L;
    double a2 = abstand2(s, spi, i, spj, j);
    if (scaled) {
        if (a2 * sig * sig > s->zh16 * s->zh16) return 0;
    } else {
        if (a2 > s->zh16 * s->zh16) return 0;
    }
    double x1 = s->git[spi][2 * i] * B + s->rel[spi][2 * i];
    double y1 = s->git[spi][2 * i + 1] * B + s->rel[spi][2 * i + 1];
    double x2 = s->git[spj][2 * j] * B + s->rel[spj][2 * j];
    double y2 = s->git[spj][2 * j + 1] * B + s->rel[spj][2 * j + 1];
    double dx = x1 - x2, dy = y1 - y2;
    if (dx * dx + dy * dy > a2) {
        if ((dx - L) * (dx - L) < dx * dx) x2 += L;
        else if ((dx + L) * (dx + L) < dx * dx) x2 -= L;
        dx = x1 - x2;
        if ((dy - L) * (dy - L) < dy * dy) y2 += L;
        else if ((dy + L) * (dy + L) < dy * dy) y2 -= L;
        dy = y1 - y2;
    }
    if (scaled) {
        a2 *= sig * sig;
        dx *= sig;
        dy *= sig;
    }
    double a_8 = 1.0 / (a2 * a2 * a2 * a2);
    double a_14 = 1.0 / (a2 * a2 * a2 * a2 * a2 * a2 * a2);
    double absF;
    if (scaled)
        absF = epsf * 4.0 * 6.0 * (2.0 * a_14 - a_8);
    else
        absF = 4.0 * 6.0 * (2.0 * a_14 - a_8);
    if (absF > s->p.max_pair_wca) absF = s->p.max_pair_wca;
    *fx = absF * dx;
    *fy = absF * dy;
    return 1;
}

/* berechne_WCA11 (dm:617-702) */
void bdo_berechne_wca11(bdo_state *s) {
    int n1 = s->p.n1, nc = s->nc;
    for (int i = 0; i < n1; ++i) { s->Fw[0][2 * i] = 0.0; s->Fw[0][2 * i + 1] = 0.0; }
    if (s->p.idgas_only) return;
    for (int i = 0; i < n1; ++i) {
        const ilist *l = &s->nb[0][s->git[0][2 * i] * nc + s->git[0][2 * i + 1]];
        for (int t = 0; t < l->n; ++t) {
            int j = l->a[t];
            if (j == i) continue;
            double fx, fy;
            if (!wca_pair(s, 0, i, 0, j, 1.0, 1.0, 0, &fx, &fy)) continue;
            s->Fw[0][2 * i] += fx;
            s->Fw[0][2 * i + 1] += fy;
        }
    }
}

/* berechne_WCA22 (dm:790-879) */
void bdo_berechne_wca22(bdo_state *s) {
    int n2 = s->p.n2, nc = s->nc;
    for (int i = 0; i < n2; ++i) { s->Fw[1][2 * i] = 0.0; s->Fw[1][2 * i + 1] = 0.0; }
    if (s->p.idgas_only) return;
    for (int i = 0; i < n2; ++i) {
        const ilist *l = &s->nb[1][s->git[1][2 * i] * nc + s->git[1][2 * i + 1]];
        for (int t = 0; t < l->n; ++t) {
            int j = l->a[t];
            if (j == i) continue;
            double fx, fy;
            if (!wca_pair(s, 1, i, 1, j, s->p.s11_22, s->p.e22_11 * s->p.s11_22, 1, &fx, &fy))
                continue;
            s->Fw[1][2 * i] += fx;
            s->Fw[1][2 * i + 1] += fy;
        }
    }
}

/* addiere_WCA12 (dm:973-1059) */
void bdo_addiere_wca12(bdo_state *s) {
    if (s->p.idgas_only) return;
    int n1 = s->p.n1, nc = s->nc;
    for (int i = 0; i < n1; ++i) {
        const ilist *l = &s->nb[1][s->git[0][2 * i] * nc + s->git[0][2 * i + 1]];
        for (int t = 0; t < l->n; ++t) {
            int j2 = l->a[t];
            double fx, fy;
            if (!wca_pair(s, 0, i, 1, j2, s->p.s11_12, s->p.e12_11 * s->p.s11_12, 1, &fx, &fy))
                continue;
            s->Fw[0][2 * i] += fx;
            s->Fw[0][2 * i + 1] += fy;
            s->Fw[1][2 * j2] -= fx;
            s->Fw[1][2 * j2 + 1] -= fy;
        }
    }
}

/* particle -> mesh index and the three weights per axis, per scheme:
   NGP gr:44-54; CIC gr:154-190; TSC gr:435-459 */
static void weights1d(const bdo_state *s, double x, int *k, double w[3]) {
    double h = s->h;
    int kk = (int)(x / h);
    *k = kk;
    if (s->p.scheme == 0) {
        w[0] = 0.0; w[1] = 1.0; w[2] = 0.0;
    } else if (s->p.scheme == 1) {
        double tmp = x / h - (kk + 0.5);
        w[1] = 1.0 - fabs(tmp);
        if (tmp < 0.0) { w[0] = -tmp; w[2] = 0.0; }
        else { w[0] = 0.0; w[2] = tmp; }
    } else {
        double d = x - (kk + 0.5) * h;
        w[1] = 0.75 - d * d / h / h;
        double tmp = fabs((d + h) / h);
        w[0] = 0.5 * (1.5 - tmp) * (1.5 - tmp);
        tmp = fabs((d - h) / h);
        w[2] = 0.5 * (1.5 - tmp) * (1.5 - tmp);
    }
}

/* gridDensity_{NGP,CIC,TSC} (gr:31-76, 131-280, 412-547) */
static void deposit(bdo_state *s) {
    int Z = s->p.Z;
    for (size_t c = 0; c < (size_t)Z * Z; ++c) s->rhox[2 * c] = 0.0;
    for (int sp = 0; sp < 2; ++sp)
        for (int i = 0; i < nsp(s, sp); ++i) {
            double x = s->git[sp][2 * i] * s->B + s->rel[sp][2 * i];
            double y = s->git[sp][2 * i + 1] * s->B + s->rel[sp][2 * i + 1];
            int k, l;
            double wx[3], wy[3];
            weights1d(s, x, &k, wx);
            weights1d(s, y, &l, wy);
            if (s->p.scheme == 0) {
                if (sp == 0) s->rhox[2 * iwz(s, k, l)] += s->drho;
                else s->rhox[2 * iwz(s, k, l)] += s->drho * s->p.f2_f1;
                continue;
            }
            int kx[3] = {(k - 1 + Z) % Z, k, (k + 1) % Z};
            int ly[3] = {(l - 1 + Z) % Z, l, (l + 1) % Z};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    if (sp == 0) s->rhox[2 * iwz(s, kx[a], ly[b])] += wx[a] * wy[b] * s->drho;
                    else s->rhox[2 * iwz(s, kx[a], ly[b])] += wx[a] * wy[b] * s->drho * s->p.f2_f1;
                }
        }
}

/* inv_gridDensity_{NGP,CIC,TSC} (gr:81-126, 284-407, 551-668) */
static void gather(bdo_state *s) {
    int Z = s->p.Z;
    for (int sp = 0; sp < 2; ++sp)
        for (int i = 0; i < nsp(s, sp); ++i) {
            double x = s->git[sp][2 * i] * s->B + s->rel[sp][2 * i];
            double y = s->git[sp][2 * i + 1] * s->B + s->rel[sp][2 * i + 1];
            int k, l;
            double wx[3], wy[3];
            weights1d(s, x, &k, wx);
            weights1d(s, y, &l, wy);
            double *F = s->Fk[sp] + 2 * i;
            if (s->p.scheme == 0) {
                size_t c = iwz(s, k, l);
                if (sp == 0) {
                    F[0] = s->Fx[2 * c];
                    F[1] = s->Fy[2 * c];
                } else {
                    F[0] = s->p.f2_f1 * s->Fx[2 * c];
                    /* gr:123 reads the imaginary part of Fx; ngp_fix reads Fy */
                    F[1] = s->p.ngp_fix ? s->p.f2_f1 * s->Fy[2 * c] : s->p.f2_f1 * s->Fx[2 * c + 1];
                }
                continue;
            }
            int kx[3] = {(k - 1 + Z) % Z, k, (k + 1) % Z};
            int ly[3] = {(l - 1 + Z) % Z, l, (l + 1) % Z};
            double fx = 0.0, fy = 0.0;
            for (int a = 0; a < 3; ++a) {
                double sx = wy[0] * s->Fx[2 * iwz(s, kx[a], ly[0])] + wy[1] * s->Fx[2 * iwz(s, kx[a], ly[1])] +
                            wy[2] * s->Fx[2 * iwz(s, kx[a], ly[2])];
                double sy = wy[0] * s->Fy[2 * iwz(s, kx[a], ly[0])] + wy[1] * s->Fy[2 * iwz(s, kx[a], ly[1])] +
                            wy[2] * s->Fy[2 * iwz(s, kx[a], ly[2])];
                if (a == 0) { fx = wx[0] * sx; fy = wx[0] * sy; }
                else { fx += wx[a] * sx; fy += wx[a] * sy; }
            }
            if (sp == 1) { fx *= s->p.f2_f1; fy *= s->p.f2_f1; }
            F[0] = fx;
            F[1] = fy;
        }
}

/* berechne_kapkraefte (dm:1180-1233) */
void bdo_berechne_kapkraefte(bdo_state *s) {
    if (s->p.idgas_only) {
        for (int sp = 0; sp < 2; ++sp)
            for (int i = 0; i < 2 * nsp(s, sp); ++i) s->Fk[sp][i] = 0.0;
        return;
    }
    int Z = s->p.Z;
    size_t Z2 = (size_t)Z * Z;
    deposit(s);
    for (size_t c = 0; c < Z2; ++c) s->rhox[2 * c + 1] = 0.0; /* dm:556-559 */
    /* rhok = FFT(rhox), in a copy (rhox is kept for bdo_get_mesh) */
    double *rhok = (double *)malloc(sizeof(double) * 2 * Z2);
    memcpy(rhok, s->rhox, sizeof(double) * 2 * Z2);
    bdo_fft2d(rhok, Z, -1);
    for (size_t j = 0; j < Z2; ++j) { /* dm:1210-1216 */
        s->Fx[2 * j] = -s->sinxG[j] * rhok[2 * j + 1];
        s->Fx[2 * j + 1] = s->sinxG[j] * rhok[2 * j];
        s->Fy[2 * j] = -s->sinyG[j] * rhok[2 * j + 1];
        s->Fy[2 * j + 1] = s->sinyG[j] * rhok[2 * j];
    }
    free(rhok);
    bdo_fft2d(s->Fx, Z, +1); /* dm:1221-1222 */
    bdo_fft2d(s->Fy, Z, +1);
    gather(s);
}

/* berechne_zufallskraefte (dm:1317-1339) and the counter-based modes */
void bdo_berechne_zufallskraefte(bdo_state *s) {
    const double wurzelfuenf = sqrt(5.0);
    for (int sp = 0; sp < 2; ++sp)
        for (int i = 0; i < nsp(s, sp); ++i) {
            double *F = s->Fn[sp] + 2 * i;
            if (s->p.noise_mode == 0) {
                double x, y;
                do {
                    x = wurzelfuenf * mt_vonbis(&s->rng, -1.0, 1.0);
                    y = wurzelfuenf * mt_vonbis(&s->rng, -1.0, 1.0);
                } while (x * x + y * y > 5.0);
                F[0] = x;
                F[1] = y;
            } else if (s->p.noise_mode == 3) {
                F[0] = s->ext_noise[sp][2 * i];
                F[1] = s->ext_noise[sp][2 * i + 1];
            } else {
                uint32_t pid = (uint32_t)(sp == 0 ? i : s->p.n1 + i);
                bdo_philox_noise(s->p.seed, pid, (uint64_t)s->step, s->p.noise_mode == 2,
                                 s->p.noise_var, F);
            }
        }
}

/* optimaler_zeitschritt (dm:1420-1491); std::min(a,b) = (b<a)?b:a */
double bdo_optimaler_zeitschritt(const bdo_state *s) {
    double ret = s->p.dt_max;
    const double mrd = s->p.mrd, T = s->p.T;
    for (int sp = 0; sp < 2; ++sp)
        for (int i = 0; i < nsp(s, sp); ++i)
            for (int c = 0; c < 2; ++c) {
                double F = s->Fw[sp][2 * i + c] + s->Fk[sp][2 * i + c];
                double n = s->Fn[sp][2 * i + c];
                double cand;
                if (fabs(F) < 1.0e-5)
                    cand = mrd * mrd / (2.0 * T * n * n);
                else {
                    double u = n * sqrt(0.5 * T) / F;
                    double w = sqrt(u * u + mrd / fabs(F));
                    cand = (w - fabs(u)) * (w - fabs(u));
                }
                ret = (cand < ret) ? cand : ret;
            }
    return ret;
}

/* refresh_erwNachbar (dm:216-325) */
void bdo_refresh_erwnachbar(bdo_state *s) {
    double B = s->B;
    int nc = s->nc;
    for (int sp = 0; sp < 2; ++sp)
        for (int i = 0; i < nsp(s, sp); ++i) {
            int32_t *g = s->git[sp] + 2 * i;
            double *r = s->rel[sp] + 2 * i;
            if (r[0] < 0.0 || r[0] > B || r[1] < 0.0 || r[1] > B) {
                lists_rem(s, sp, i, g[0], g[1]);
                int ic = (int)floor((g[0] * B + r[0]) / B);
                int jc = (int)floor((g[1] * B + r[1]) / B);
                ic = (ic + nc) % nc;
                jc = (jc + nc) % nc;
                if (ic < 0) ic += nc; /* reference UB for moves beyond -L */
                if (jc < 0) jc += nc;
                r[0] = fmod(r[0] + B, B);
                r[1] = fmod(r[1] + B, B);
                g[0] = ic;
                g[1] = jc;
                lists_add(s, sp, i, ic, jc);
            }
        }
}

/* zeitschritt (dm:74-163) */
double bdo_zeitschritt(bdo_state *s, double tmax) {
    if (tmax <= 0.0) return 1.0e10; /* dm:80-83 */
    bdo_berechne_wca11(s);
    bdo_berechne_wca22(s);
    bdo_addiere_wca12(s);
    bdo_berechne_kapkraefte(s);
    if (!s->p.no_wca)
        bdo_berechne_zufallskraefte(s);
    else
        for (int sp = 0; sp < 2; ++sp)
            for (int i = 0; i < 2 * nsp(s, sp); ++i) s->Fn[sp][i] = 0.0;
    double deltat = bdo_optimaler_zeitschritt(s);
    double T = s->p.T, B = s->B;
    for (int sp = 0; sp < 2; ++sp)
        for (int i = 0; i < nsp(s, sp); ++i) {
            double *r = s->rel[sp] + 2 * i;
            double dx = deltat * (s->Fw[sp][2 * i] + s->Fk[sp][2 * i]) + sqrt(2.0 * T * deltat) * s->Fn[sp][2 * i];
            double dy = deltat * (s->Fw[sp][2 * i + 1] + s->Fk[sp][2 * i + 1]) +
                        sqrt(2.0 * T * deltat) * s->Fn[sp][2 * i + 1];
            r[0] += dx;
            r[1] += dy;
            if (sp == 1 && s->p.restrict_radial) { /* dm:138-155 */
                const double mitte = 0.5 * s->p.L;
                double x = s->git[1][2 * i] * B + r[0];
                double y = s->git[1][2 * i + 1] * B + r[1];
                double dxm = mitte - x, dym = mitte - y;
                double abst = sqrt(dxm * dxm + dym * dym);
                double xs = mitte + abst * cos(2 * M_PI * (int)i / s->p.n2);
                double ys = mitte + abst * sin(2 * M_PI * (int)i / s->p.n2);
                r[0] += (xs - x);
                r[1] += (ys - y);
            }
        }
    bdo_refresh_erwnachbar(s);
    s->step++;
    return deltat;
}

/* zeitschritte_bis_obs (dm:444-455) with t reset by obs_point (au:90) */
int64_t bdo_zeitschritte_bis_obs(bdo_state *s, double obs_dt) {
    long double t = 0.0L;
    int64_t n = 0;
    while (t < obs_dt) {
        t += (long double)bdo_zeitschritt(s, (double)((long double)obs_dt - t));
        n++;
    }
    return n;
}

int64_t bdo_steps_for_interval(double obs_dt, double dt) {
    long double t = 0.0L;
    int64_t n = 0;
    while (t < obs_dt) {
        t += (long double)dt;
        n++;
    }
    return n;
}

int bdo_get_forces(const bdo_state *s, int which, int species, double *out) {
    int sp = species - 1;
    if (sp < 0 || sp > 1) return -1;
    const double *src = which == 0 ? s->Fw[sp] : which == 1 ? s->Fk[sp] : s->Fn[sp];
    memcpy(out, src, sizeof(double) * 2 * (size_t)nsp(s, sp));
    return 0;
}

void bdo_set_noise(bdo_state *s, const double *n1, const double *n2) {
    if (n1) memcpy(s->ext_noise[0], n1, sizeof(double) * 2 * (size_t)s->p.n1);
    if (n2) memcpy(s->ext_noise[1], n2, sizeof(double) * 2 * (size_t)s->p.n2);
}

void bdo_get_mesh(const bdo_state *s, int which, double *out) {
    const double *src = which == 0 ? s->rhox : which == 1 ? s->Fx : s->Fy;
    for (size_t c = 0; c < (size_t)s->p.Z * s->p.Z; ++c) out[c] = src[2 * c];
}

int bdo_get_list(const bdo_state *s, int species, int cx, int cy, int32_t *out, int cap) {
    if (species < 1 || species > 2 || cx < 0 || cx >= s->nc || cy < 0 || cy >= s->nc) return -1;
    const ilist *l = &s->nb[species - 1][cx * s->nc + cy];
    for (int i = 0; i < l->n && i < cap; ++i) out[i] = l->a[i];
    return l->n;
}

int64_t bdo_step_counter(const bdo_state *s) { return s->step; }

int bdo_write_pos_header(const char *path) {
    FILE *f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "# Format: t TAB x TAB y \n\n"); /* au:60 */
    fclose(f);
    return 0;
}

/* schreibe_pos (au:631-648) for one species */
int bdo_append_pos(const bdo_state *s, int species, const char *path, int nr, double obs_dt) {
    int sp = species - 1;
    if (sp < 0 || sp > 1) return -1;
    FILE *f = fopen(path, "a");
    if (!f) return -1;
    for (int i = 0; i < nsp(s, sp); ++i) {
        double x = s->rel[sp][2 * i] + s->B * s->git[sp][2 * i];
        double y = s->rel[sp][2 * i + 1] + s->B * s->git[sp][2 * i + 1];
        fprintf(f, "%g \t %g \t %g \n", nr * obs_dt, x, y);
    }
    fprintf(f, "\n");
    fclose(f);
    return 0;
}

void bdo_mt_seed(bdo_state *s, uint64_t seed) { mt_seed(&s->rng, seed); }
uint64_t bdo_mt_next(bdo_state *s) { return mt_next(&s->rng); }
double bdo_mt_canonical(bdo_state *s) { return mt_canonical(&s->rng); }
