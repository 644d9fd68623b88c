/*
 * bd_oracle.h -- FP64 CPU restatement of the Brownian-dynamics timestep of
 * Gully85/BD-Simulations (RunZustand::RunDynamik, dynamik_methoden.cpp,
 * gridRoutinen.cpp, allgemeine_methoden.cpp, zufall.cpp, auswertung.cpp).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity reference for the HIP
 * path: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  The product library never links or calls it.
 *
 * Everything is plain C99, FP64, single-threaded, and compiled without
 * floating-point contraction, so that the arithmetic follows the reference's
 * x86-64 build (no FMA) operation for operation.
 *
 * Pinning (what fixes each function independently of this restatement):
 *   mt19937_64 + canonical      : C++ <random> (libstdc++) golden stream, and the
 *                                 standard's KAT (10000th output of the
 *                                 default-seeded engine = 9981545732273789042)
 *   philox4x32-10               : Random123 published known-answer vectors
 *   cell lists                  : brute-force membership from the git formula
 *   WCA 11/22/12                : closed form 24(2r^-13 - r^-7), brute-force
 *                                 O(N^2) all-pair sum, Newton-3 sum = 0, cap
 *   deposit NGP/CIC/TSC         : charge conservation, partition of unity,
 *                                 textbook (Hockney-Eastwood) weights
 *   2D FFT                      : brute-force DFT and numpy.fft
 *   Green's function / gradient : closed-form symbols, two-particle force vs
 *                                 v K1(r/lambda)/(2 pi lambda) (scipy)
 *   gather                      : adjoint identity with the deposit
 *   dt, update, refresh         : closed forms, integer floor formula
 *   checkpoint step count       : long double accumulation (1001 for 0.1/1e-4)
 *   pos writer                  : printf("%g") format
 */
#ifndef BD_ORACLE_H
#define BD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters as in parameter.h:14-99 (user constants). */
typedef struct bdo_params {
    int32_t n1, n2;          /* N1, N2                         ph:15-16 */
    int32_t scheme;          /* densGrid_Schema 0 NGP 1 CIC 2 TSC ph:22 */
    int32_t Z;               /* FFT_samples (power of two)      ph:25 */
    double L;                /* box side                        ph:19 */
    double lambda;           /* lambda_kapillar                 ph:29 */
    double kap;              /* kapillar_vorfaktor              ph:32 */
    double T;                /* temperature                     ph:37 */
    double dt_max;           /* dt_max                          ph:41 */
    double mrd;              /* max_reisedistanz (inf = fixed dt) ph:44 */
    double s11_22;           /* sigma11_22                      ph:47 */
    double s11_12;           /* sigma11_12                      ph:50 */
    double e22_11;           /* eps22_11                        ph:56 */
    double e12_11;           /* eps12_11                        ph:59 */
    double f2_f1;            /* f2_f1                           ph:65 */
    double max_pair_wca;     /* maxPairWCA                      ph:81 */
    int32_t idgas_only;      /* ph:68 */
    int32_t no_wca;          /* ph:70 (also switches the noise off, dm:90-101) */
    int32_t restrict_radial; /* ph:76 */
    int32_t ngp_fix;         /* 0: reference NGP species-2 y gather (gr:123) */
    int32_t noise_mode;      /* 0 MT disc (reference), 1 philox disc,
                                2 philox gauss, 3 external arrays */
    int32_t pad0;
    double noise_var;        /* variance per component for mode 2 */
    uint64_t seed;           /* MT seed (mode 0) / philox key (modes 1,2) */
} bdo_params;

typedef struct bdo_state bdo_state;

/* lifecycle */
bdo_state *bdo_create(const bdo_params *p);
void bdo_destroy(bdo_state *s);
/* derived constants: [0]=zweihoch1_6 [1]=nc [2]=B [3]=h [4]=dq [5]=drho */
void bdo_derived(const bdo_state *s, double *out6);

/* positions in the reference's (git, rel) form, species 1 or 2, [n][2] */
int bdo_set_positions(bdo_state *s, int species, int n, const int32_t *git,
                      const double *rel);
int bdo_get_positions(const bdo_state *s, int species, int n, int32_t *git,
                      double *rel);
/* (re)build the extended neighbour lists from git (WCA_init, dm:459-526) */
void bdo_wca_init(bdo_state *s);
/* init_zufallspos (dm:1647-1678) from the MT stream, then WCA_init */
void bdo_init_zufallspos(bdo_state *s);

/* force stages (fill the internal force arrays) */
void bdo_berechne_wca11(bdo_state *s);   /* dm:617-702 */
void bdo_berechne_wca22(bdo_state *s);   /* dm:790-879 */
void bdo_addiere_wca12(bdo_state *s);    /* dm:973-1059 */
void bdo_berechne_kapkraefte(bdo_state *s); /* dm:1180-1233 */
void bdo_berechne_zufallskraefte(bdo_state *s); /* dm:1317-1339 */
double bdo_optimaler_zeitschritt(const bdo_state *s); /* dm:1420-1491 */
void bdo_refresh_erwnachbar(bdo_state *s); /* dm:216-325 */
double bdo_zeitschritt(bdo_state *s, double tmax); /* dm:74-163 */
/* zeitschritte_bis_obs (dm:444-455): returns number of steps taken */
int64_t bdo_zeitschritte_bis_obs(bdo_state *s, double obs_dt);
/* number of steps the long-double checkpoint loop takes for a constant dt */
int64_t bdo_steps_for_interval(double obs_dt, double dt);

/* force arrays: which 0=WCA 1=kap 2=noise; species 1/2; out [n][2] */
int bdo_get_forces(const bdo_state *s, int which, int species, double *out);
/* external noise (mode 3) for the next step: [n1][2] and [n2][2] */
void bdo_set_noise(bdo_state *s, const double *n1, const double *n2);
/* mesh arrays, Z*Z doubles: which 0 = rho (real part of rhox), 1 = Fx, 2 = Fy */
void bdo_get_mesh(const bdo_state *s, int which, double *out);
/* extended list of a cell: returns length, copies up to cap entries */
int bdo_get_list(const bdo_state *s, int species, int cx, int cy, int32_t *out,
                 int cap);
int64_t bdo_step_counter(const bdo_state *s);

/* pos writer (au:57-69 header, au:631-648 rows): append one checkpoint */
int bdo_write_pos_header(const char *path);
int bdo_append_pos(const bdo_state *s, int species, const char *path, int nr,
                   double obs_dt);

/* building blocks exposed for the known-answer tests */
void bdo_mt_seed(bdo_state *s, uint64_t seed);
uint64_t bdo_mt_next(bdo_state *s);
double bdo_mt_canonical(bdo_state *s);
void bdo_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* noise sample of the counter-based modes for one particle id and step */
void bdo_philox_noise(uint64_t seed, uint32_t pid, uint64_t step, int gauss,
                      double var, double out[2]);
/* in-place unnormalised 2D complex DFT, sign -1 forward / +1 backward,
   data = n*n interleaved (re,im), row-major; n power of two */
void bdo_fft2d(double *data, int n, int sign);

#ifdef __cplusplus
}
#endif
#endif
