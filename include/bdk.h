/*
 * bdk.h -- C-ABI of the MI355X Brownian-dynamics timestep (libbdk_hip.so).
 *
 * Drop-in boundary for RunZustand::RunDynamik of Gully85/BD-Simulations
 * (signaturen2.h:33-252).  Plain pointers and sizes only.  Every call returns
 * an int status (BDK_OK = 0); the message of the last failure on a handle is
 * bdk_last_error(h) (bdk_last_error(NULL) for bdk_create failures).
 *
 * Threading: one handle per host thread.  Each handle owns its HIP stream
 * and its device buffers.  Host arrays belong to the caller.
 *
 * Layout at the boundary (the reference's, signaturen2.h:197-233):
 *   positions  git [n][2] int32 (cell index x, y) and rel [n][2] double
 *              (offset inside the cell), x = git*B + rel, per species;
 *   forces     [n][2] double per species;
 *   meshes     Z*Z double, row-major [x][y] (iw(i,j) = i*Z + j, am:242-250).
 */
#ifndef BDK_H
#define BDK_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum bdk_status {
    BDK_OK = 0,
    BDK_ERR_ARG = 1,       /* invalid argument / parameter */
    BDK_ERR_HIP = 2,       /* HIP runtime failure */
    BDK_ERR_STATE = 3,     /* call out of order (e.g. positions not set) */
    BDK_ERR_NOMEM = 4,     /* device footprint does not fit */
    BDK_ERR_NONFINITE = 5  /* non-finite particle position detected */
};

enum bdk_scheme { BDK_NGP = 0, BDK_CIC = 1, BDK_TSC = 2 };     /* densGrid_Schema, ph:22 */
/* Precision ladder.  FP64: forces, mesh and FFT in FP64 (the reference's
   arithmetic).  FP32: forces, mesh and FFT in FP32, positions FP64.  The two
   further rungs keep FP32 forces and store the in-cell positions (rel) in
   FP32: POS32 rounds every new position to FP32; POS32K adds each step's
   increment with Kahan compensation (a per-particle FP32 compensation term
   that follows the particle). */
enum bdk_precision { BDK_FP64 = 0, BDK_FP32 = 1, BDK_FP32_POS32 = 2, BDK_FP32_POS32K = 3 };
enum bdk_noise {
    BDK_NOISE_DISC = 0,    /* uniform on the disc of radius sqrt 5 (dm:1317-1339), Philox */
    BDK_NOISE_GAUSS = 1,   /* Gaussian of variance noise_var, Philox */
    BDK_NOISE_UPLOAD = 2   /* caller-supplied per step (bdk_set_noise), e.g. the MT stream */
};
enum bdk_green {
    BDK_GREEN_DISCRETE = 0,   /* reference symbols: discrete Laplacian + central difference (dm:585-593, 1384-1394) */
    BDK_GREEN_CONTINUOUS = 1  /* i k / (k^2 + lambda^-2) */
};
enum bdk_mesh { BDK_MESH_RHO = 0, BDK_MESH_FX = 1, BDK_MESH_FY = 2 };

/* parameter.h:14-99 user constants, plus the build's own switches. */
typedef struct bd_params {
    int32_t n1, n2;          /* N1, N2 (ph:15-16) */
    int32_t scheme;          /* densGrid_Schema (ph:22) */
    int32_t Z;               /* FFT_samples: even, 2^a 3^b 5^c 7^d, 8..4096 (ph:24) */
    double L;                /* box side (ph:19) */
    double lambda;           /* lambda_kapillar (ph:29) */
    double kap;              /* kapillar_vorfaktor (ph:32) */
    double T;                /* temperature (ph:37) */
    double dt_max;           /* dt_max (ph:41) */
    double max_travel;       /* max_reisedistanz (ph:44); +inf gives dt = dt_max every step */
    double sigma11_22, sigma11_12, eps22_11, eps12_11; /* ph:47-59 */
    double f2_f1;            /* species-2 charge ratio (ph:65) */
    double max_pair_wca;     /* maxPairWCA, cap on |F|/r (ph:81) */
    int32_t idgas_only;      /* ph:68 */
    int32_t no_wca;          /* noWCA (ph:70): switches the noise off (dm:90-101) */
    int32_t restrict_radial; /* restrictRadial (ph:76) */
    int32_t ngp_fix;         /* 0: reference NGP species-2 y force (gr:123, ~0); 1: use Fy */
    int32_t precision;       /* bdk_precision (forces, mesh, FFT; positions FP64 unless a POS32 rung) */
    int32_t noise_mode;      /* bdk_noise */
    int32_t green_mode;      /* bdk_green */
    int32_t device;          /* HIP device ordinal */
    double noise_var;        /* variance per component for BDK_NOISE_GAUSS */
    uint64_t seed;           /* Philox key */
    int32_t ftrho_fix;       /* 0: bdk_ftrho_sums takes the reference's species-1 y = git_y B + rel_x
                                (au:333); 1: rel_y, the species' actual y */
} bd_params;

typedef struct bdk bdk_t;

/* Defaults: parameter.h's switches off, sigma/eps ratios 1, maxPairWCA 1e5,
   fixed dt (max_travel = +inf), FP32, disc noise, discrete symbols. */
int bdk_params_default(bd_params *p);

/* Derived constants of parameter.h:170-191: out[0]=2^(1/6), [1]=nachList_Zellen,
   [2]=nachList_Breite, [3]=densGrid_Breite, [4]=dq, [5]=drho (gr:28). */
int bdk_params_derived(const bd_params *p, double *out6);

/* RunZustand::RunDynamik constructor + init() (am:361-364, dm:459-613):
   validates p, allocates device state, builds the spectral tables. */
int bdk_create(const bd_params *p, bdk_t **out);
void bdk_destroy(bdk_t *h);
const char *bdk_last_error(const bdk_t *h);

/* Positions of one species (the aliases set at am:305-308).  Setting both
   species (re)builds the cell lists, as WCA_init does (dm:459-526). */
int bdk_set_positions(bdk_t *h, int species, int32_t n, const int32_t *git, const double *rel);
int bdk_get_positions(bdk_t *h, int species, int32_t n, int32_t *git, double *rel);

/* berechne_WCA11 + berechne_WCA22 + addiere_WCA12 (dm:617-702, 790-879,
   973-1059): WCA forces at the current positions; f1 [n1][2], f2 [n2][2]. */
int bdk_compute_wca(bdk_t *h, double *f1, double *f2);

/* berechne_kapkraefte (dm:1180-1233): capillary forces (deposit, FFT,
   Green's function and gradient, inverse FFT, gather). */
int bdk_compute_capillary(bdk_t *h, double *f1, double *f2);

/* Noise for the next step in BDK_NOISE_UPLOAD mode, [n1][2] and [n2][2]
   (the arrays berechne_zufallskraefte fills, dm:1317-1339). */
int bdk_set_noise(bdk_t *h, const double *n1, const double *n2);

/* Noise that the next step will use (Philox modes: a pure function of
   seed, particle and step counter). */
int bdk_get_noise(bdk_t *h, double *n1, double *n2);

/* optimaler_zeitschritt (dm:1420-1491) on the forces and noise of the
   current state, without stepping. */
int bdk_optimal_dt(bdk_t *h, double *dt);

/* zeitschritt(tmax) (dm:74-163): one step; *dt_out = the step taken.
   tmax <= 0 reproduces dm:80-83: no step, *dt_out = 1e10, BDK_ERR_ARG. */
int bdk_step(bdk_t *h, double tmax, double *dt_out);

/* k steps (hipGraph replay when dt is fixed); *t_out = sum of the dt's,
   accumulated in long double as zeitschritte_bis_obs does (dm:444-455). */
int bdk_run_steps(bdk_t *h, int64_t k, double *t_out);

/* zeitschritte_bis_obs (dm:444-455): steps until the long-double clock,
   starting at 0 (obs_point, au:90), reaches obs_dt. */
int bdk_run_interval(bdk_t *h, double obs_dt, int64_t *steps_out);
/* The same interval, also returning the long-double clock it ended with
   (RunZustand::t, s2:375) exactly as t_hi + t_lo (the 64-bit significand of
   the x87 long double fits in two doubles); either pointer may be NULL. */
int bdk_run_interval_clock(bdk_t *h, double obs_dt, int64_t *steps_out, double *t_hi, double *t_lo);

/* The travel cap of the following steps (max_reisedistanz; +inf gives
   dt = dt_max every step).  zeitschritt_debug(mrd, ...) (dm:168-211) takes
   the cap per call, as the mrd sweep of main.cpp:156-216 does.  Not while the
   handle belongs to a batch. */
int bdk_set_max_travel(bdk_t *h, double max_travel);

/* Number of steps zeitschritte_bis_obs takes for a constant dt. */
int bdk_steps_for_interval(double obs_dt, double dt, int64_t *steps_out);

/* refresh_erwNachbar (dm:216-325) on the current positions. */
int bdk_refresh(bdk_t *h);

/* Upload the positions set with bdk_set_positions, rebuild the cell lists
   and enqueue the forces of that state on the handle's stream, without
   waiting (otherwise the first step or force call does this).  The state and
   every later result are the same either way; a job server uses it to bring
   a new job up while other simulations run. */
int bdk_prepare(bdk_t *h);
int bdk_sync(bdk_t *h);
int64_t bdk_step_counter(const bdk_t *h);

/* Cell membership (the per-cell content behind erwNachbarn1/2, s2:215-216):
   start [nc*nc+1], ids [n1+n2] = particle ids (species 1: 0..n1-1, species 2:
   n1..n1+n2-1) of cell c in ids[start[c]..start[c+1]), ascending. */
int bdk_get_cells(bdk_t *h, int32_t *start, int32_t *ids);

/* Mesh arrays of the last capillary evaluation, Z*Z double, [x][y]:
   BDK_MESH_RHO = rhox[..][0] (gr:31-547), FX/FY = Fx/Fy[..][0] (dm:1221-1222). */
int bdk_get_mesh(bdk_t *h, int which, double *out);

/* Device memory this parameter set needs, bytes (admission by footprint). */
int bdk_footprint(const bd_params *p, uint64_t *bytes);

/* Per-kernel timing of k steps with HIP events on the handle's stream.
   names: buffer for up to nk '\n'-separated kernel names; ms [nk] average
   milliseconds per launch; *n_out = number of kernels in one step. */
int bdk_profile_steps(bdk_t *h, int64_t k, int32_t nk, double *ms, char *names, int32_t names_cap,
                      int32_t *n_out);

/* Average duration (ms) of one launch of a single kernel of the step,
   timed as `reps` back-to-back launches between two HIP events on the
   handle's stream.  kernel: BDK_K_* below.  The kernels timed this way are
   idempotent (they read the state and overwrite their outputs), except
   BDK_K_STEP, which advances the simulation by `reps` full steps. */
enum bdk_kernel_id {
    BDK_K_WCA = 0,        /* WCA forces */
    BDK_K_DEPOSIT = 1,    /* charge deposit */
    BDK_K_R2C = 2,        /* row r2c FFT pass */
    BDK_K_COL = 3,        /* column FFT + Green's function + inverse column FFTs */
    BDK_K_C2R = 4,        /* row c2r FFT pass */
    BDK_K_GATHER = 5,     /* capillary gather (no update) */
    BDK_K_STEP = 6,       /* one full step (graph replay when dt is fixed) */
    BDK_K_UPDATE = 7      /* gather + noise + Euler step + refresh + counting; the state it
                             changes is saved before and restored after the timing */
};
int bdk_profile_kernel(bdk_t *h, int kernel, int32_t reps, double *ms_per_launch);

/* ---- initial states (RunZustand::init, am:256-320) ---- */

/* Host mt19937_64 stream of zufall.cpp:18-39 (uniform_real_distribution over
   a seeded engine; the reference seeds with time(NULL)). */
typedef struct bdk_mt bdk_mt_t;
int bdk_mt_create(uint64_t seed, bdk_mt_t **out);
void bdk_mt_destroy(bdk_mt_t *g);
double bdk_mt_uniform(bdk_mt_t *g, double lo, double hi);
/* init_zufallspos (dm:1647-1678), incl. its ((int)x)/B cell index. */
int bdk_mt_init_positions(bdk_mt_t *g, const bd_params *p, int32_t *git1, double *rel1,
                          int32_t *git2, double *rel2);
/* berechne_zufallskraefte (dm:1317-1339): disc noise by rejection, S1 then S2. */
int bdk_mt_disc_noise(bdk_mt_t *g, int32_t n1, int32_t n2, double *out1, double *out2);

/* Uniform random sequential insertion with exclusion distance dmin
   (periodic), seeded; all particles, species 1 first. */
int bdk_init_uniform(const bd_params *p, uint64_t seed, double dmin, int32_t *git, double *rel);
/* Gaussian blob (std dev `width`, centred) plus a uniform background fraction
   `background`, by seeded thinning of a hexagonal lattice of spacing dmin:
   exclusion dmin holds exactly; target densities above the lattice density
   2/(sqrt3 dmin^2) are clipped to it.  *clipped_out = fraction of the target
   profile's particles lost to the clip (before renormalisation). */
int bdk_init_blob(const bd_params *p, uint64_t seed, double width, double background, double dmin,
                  int32_t *git, double *rel, double *clipped_out);

/* pos1_<run+1>.txt / pos2_<run+1>.txt writer (au:57-69, au:631-648):
   header on nr == 0 with truncate, then one block per checkpoint. */
int bdk_write_pos(bdk_t *h, const char *path1, const char *path2, int32_t nr, double obs_dt);
/* The same checkpoint written in the background (schreibe_pos, au:631-648,
   which blocks the reference's run loop): the current state is copied to
   pinned host memory on the handle's stream, ordered before any later step,
   and a host thread formats and appends the files while the device keeps
   stepping.  The files' bytes equal bdk_write_pos's.  At most one write is in
   flight per handle: a new call (or bdk_write_pos) first waits for the
   previous one and returns its error, if any. */
int bdk_write_pos_async(bdk_t *h, const char *path1, const char *path2, int32_t nr, double obs_dt);
/* Wait for the handle's asynchronous write; returns its status (BDK_OK when
   none is in flight).  bdk_destroy also waits. */
int bdk_write_wait(bdk_t *h);

/* One step like bdk_step that also fills the counters of TimestepInfo
   (signaturen2.h:71-145, zeitschritt_debug dm:168-211): the largest WCA,
   capillary and random force of the step by |F|^2 (vector, species 1/2,
   index within the species), the species counts of the 3x3 cells around
   the WCA maximum, the number of candidate pairs i != j the cell lists
   present (as dm:704-1078 counts them: species-1 particles against both
   lists, species-2 particles against list 2), the step's dt, and the
   particles that changed cell per species.
   Ties in |F|^2 at FP32 resolution go to the later list position. */
typedef struct {
    int64_t num_wca;
    double max_wca[2];
    int32_t max_wca_type, max_wca_index;
    int32_t max_wca_nlist[2];
    double max_kap[2];
    int32_t max_kap_type, max_kap_index;
    double max_rnd[2];
    int32_t max_rnd_type, max_rnd_index;
    double dt;
    int32_t cell_changes[2];
} bdk_step_info_t;
int bdk_step_info(bdk_t *h, double tmax, bdk_step_info_t *out);

/* ---------------------------------------------------------------- restart
   init_pos_aus_datei (dm:1576-1643) with its P8 defect fixed (rel = x - git*B):
   reads startpos.txt-format files (two header lines, then "nr typ x y" for
   species 1 then 2, nr from 0 per species, coordinates parsed as float like
   the reference, 0 <= x, y <= L) into git/rel, species 1 first. */
int bdk_read_startpos(const bd_params *p, const char *path, int32_t *git, double *rel);
/* Device snapshot (no counterpart in the reference, whose restart path is
   broken): parameters, step counters and the particle state in cell-list
   order.  bdk_load_state on a handle created with the same parameters
   continues the run bit for bit. */
int bdk_save_state(bdk_t *h, const char *path);
int bdk_load_state(bdk_t *h, const char *path);

/* ---------------------------------------------------------------- observables
   Radial density profiles of a blob (berechne_dichteprofile,
   observablen_auswertung.cpp:227-330) on the current positions.  Centres
   (berechne_mittelpunkte, observablen_auswertung.cpp:140-223): per species and
   axis the circular mean S = (atan2(-sum sin, -sum cos) + pi) L / 2pi;
   centers[6] = S1x, S1y, S2x, S2y, Sgx, Sgy with Sg = (S1 + S2)/2.
   profiles[6][nbins] = rho1ring (species 1 around Sg), rho2ring (species 2
   around S2, as the reference computes it), rho1zen1, rho1zen2, rho2zen1,
   rho2zen2 (species a around centre b); bin = (int)(r/dr), r without
   periodic images, overflow into the last bin; value = count * L^2 /
   (N_a dr^2 pi) / (2 bin + 1). */
int bdk_radial_profile(bdk_t *h, int32_t nbins, double dr, double *centers, double *profiles);

/* Pair correlation functions g11, g12, g22 (record_korrelationsfunktion11/12/22,
   auswertung.cpp:110-188): minimum-image pair distances binned by (int)(r/dr)
   for r < nbins*dr; each pair adds vorfaktor/(2 bin + 1) with vorfaktor =
   2/((N1-1) pi rho1 dr^2) (11), 1/((N1-1) pi rho2 dr^2) (12, as the reference
   writes it), 2/((N2-1) pi rho2 dr^2) (22), rho_a = N_a/L^2.  Any of the
   three outputs may be null. */
int bdk_pair_correlation(bdk_t *h, int32_t nbins, double dr, double *g11, double *g12, double *g22);
/* The same histograms without allocating anything: no counterpart in the
   reference; for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
   counts go through a fixed scratch array in device memory, one per GPU
   (96 KiB, kept until the process ends), and are written straight into the
   caller's arrays.  Limits: 1 <= nbins <= 4096; all three outputs are
   required; the particle state must already be on the device (any force,
   step or refresh call after bdk_set_positions), else BDK_ERR_STATE.  Calls
   that overlap on one GPU (several host threads) race on the array. */
int bdk_pair_correlation_noalloc(bdk_t *h, int32_t nbins, double dr, double *g11, double *g12, double *g22);

/* rho(k) of one species (record_rhoFFTW, auswertung.cpp:529-626): the number
   density of `species` on the mesh (the force deposit's scheme and weights
   drho, gridRoutinen.cpp:672-1040), its 2D FFT times (-1)^(i+j), averaged
   over shells bin = (int)(dq_ratio * |(i',j')| + 1/2) of the frequency
   indices, times L^2/Z^2.  The reference's dq_dqFT = dq/dq_rhoFFTW = 1/2 and
   nbins = Z * dq_ratio / 2 (parameter.h:194-200).  Shells at or beyond nbins
   are left out (the reference's bound check lets bin == nbins through).
   im[] receives the shell sums of the imaginary parts, which cancel over
   the full plane (zero here; rounding-size in the reference). */
int bdk_rho_k(bdk_t *h, int32_t species, double dq_ratio, int32_t nbins, double *re, double *im);

/* Local densities of the snapshot grid (main_auswertung.cpp:214-320) on the
   current positions: an m x m grid of width w = L/m (the reference uses
   m = 2 (int)(L/(2 snapgrid_crude) + 0.5), parameter.h:161-176); each particle
   adds its TSC weights (centre 3/4 - d^2/w^2, sides (3/2 - |d +- w|/w)^2 / 2,
   d from the centre of cell (int)(x/w)) times 1/w^2 to the nine cells around
   it on its species' grid.  rho1/rho2 [m][m], x index first; either may be
   NULL.  Weight products are summed in fixed point (2^-40), so the grid is
   independent of summation order; each contribution differs from its FP64
   value by at most 2^-41 / w^2. */
int bdk_density_grid(bdk_t *h, int32_t m, double *rho1, double *rho2);
/* Statistics of that grid (main_auswertung.cpp:324-385): hist[4][nbins] =
   fractions of the m*m cells per bin of rho1, rho2, rho = rho1 + rho2 (bin =
   (int)(x/dbin - 0.5), clamped to [0, nbins-1]) and rho1 - rho2 (bin shifted
   by nbins/2, integer division); over[7] = fractions with rho >
   rho_thresh[0..2] and |rho1 - rho2|/rho > drho_thresh[0..3]. */
int bdk_density_stats(bdk_t *h, int32_t m, int32_t nbins, double dbin, const double *rho_thresh,
                      const double *drho_thresh, double *hist, double *over);

/* record_ftrho_unkorrigiert (au:307-360): for the wave-vector index pairs
   (i, j), i, j < qbins, of the reference's loop (in loop order; the inner
   loop stops at the first j with i^2 + j^2 > qbins^2), the sums over one
   species of cos and sin(dq_ft (x i + y j)), x = git_x B + rel_x.  For
   species 1 the reference takes y = git_y B + rel_x (au:333), kept here
   unless bd_params.ftrho_fix is set (then rel_y, as species 2 uses).
   re/im receive *npairs (<= qbins^2) entries; the shells by i^2 + j^2 are
   the caller's (init_ftrho, au:262-300).  Device scratch: npairs x at most
   64 partial sums (16 B each), allocated for the call. */
int bdk_ftrho_sums(bdk_t *h, int32_t species, int32_t qbins, double dq_ft, double *re, double *im, int32_t *npairs);

/* ---------------------------------------------------------------- batches
   Many independent simulations (the reference's `runs` and parameter sweeps
   over N, L, lambda and scheme) advanced together with one kernel launch per
   stage of the step.  No counterpart in the reference, which runs one
   simulation at a time (main.cpp:129); a batch replaces a loop of
   zeitschritt calls over independent RunZustand objects.

   A batch is a view over member handles made with bdk_create and given
   positions; it does not own them.  Members share the device and the
   precision, have capillary forces on (idgas_only = 0) and do not use
   BDK_NOISE_UPLOAD.  Members may have a fixed dt (max_travel = +inf) or
   the adaptive dt of optimaler_zeitschritt (dm:1420-1491), each member its
   own dt per step; bdk_run_interval on a member afterwards keeps the
   reference's long-double clock from its own logged dt's.  Each member's state
   stays in its own buffers: after bdk_batch_sync the ordinary per-handle
   calls (bdk_get_positions, bdk_write_pos, ...) read it, and each member's
   trajectory is bit-identical to running that handle alone.  A member must
   not be stepped through its own handle while it belongs to a live batch. */
typedef struct bdk_batch bdk_batch_t;

int bdk_batch_create(bdk_t *const *members, int32_t n, bdk_batch_t **out);
void bdk_batch_destroy(bdk_batch_t *b);
const char *bdk_batch_last_error(const bdk_batch_t *b);
/* Enqueue k steps of every member (graph replay); returns before they end. */
int bdk_batch_run_steps(bdk_batch_t *b, int64_t k);
/* Wait for the batch; BDK_ERR_NONFINITE names the first member that ran away. */
int bdk_batch_sync(bdk_batch_t *b);
/* Wait for the batch and report each member's state: status[i] = BDK_OK, or
   BDK_ERR_NONFINITE for a member whose particle ran away (held in place; the
   flag stays until the member's positions are set again), or BDK_ERR_HIP for
   an internal inconsistency.  n = the number of members. */
int bdk_batch_status(bdk_batch_t *b, int32_t n, int32_t *status);
/* Project autotuner: times each knob on the batch's own members (reps
   launches of the stages that leave the particles unchanged) and keeps the
   fastest value: the FFT elements per thread of each pass (8, 16, 32), the
   WCA block size (128, 256, 512 threads), the deposit tile (list cells per
   side such that a tile owns at most 24^2, 32^2, 38^2, 48^2, 64^2 or 80^2 mesh points),
   and the persistent grid (one block per work item, or 4 or 8 blocks per
   compute unit that loop over the work items).  No choice changes a
   result.  choice_out[BDK_TUNE_COUNT] (may be NULL) receives the choice. */
enum bdk_tune_index {
    BDK_TUNE_EPT_R2C = 0, BDK_TUNE_EPT_COL = 1, BDK_TUNE_EPT_C2R = 2,
    BDK_TUNE_WCA_BLOCK = 3, BDK_TUNE_DEPOSIT_TILE = 4, BDK_TUNE_PERSIST = 5,
    BDK_TUNE_COUNT = 6
};
int bdk_batch_tune(bdk_batch_t *b, int32_t reps, int32_t *choice_out);
/* Apply a choice (e.g. one bdk_batch_tune made for the same mix): FFT
   elements per thread 8, 16 or 32 (raised to the members' minimum), WCA
   block size 128, 256 or 512, deposit tile 8..96 mesh points per side,
   persistent grid >= 0 (0: one block per work item).  No choice changes a
   result. */
int bdk_batch_set_tuning(bdk_batch_t *b, const int32_t *choice);
/* ms per launch of one stage (0 WCA, 1 deposit, 2 r2c, 3 col, 4 c2r, 7 whole
   step) over reps launches after one warm-up launch.  Stages 0-4 leave the
   particles unchanged; stage 7 advances every member by reps + 1 steps. */
int bdk_batch_profile(bdk_batch_t *b, int32_t stage, int32_t reps, double *ms_per_launch);

/* ---------------------------------------------------------------- serving
   A job server over batches (continuous batching).  No counterpart in the
   reference.  Jobs (parameters, start positions species 1 first, step count)
   are submitted at any time; bdk_serve_step runs one tick: it admits queued
   jobs in arrival order while their HBM footprints fit `hbm_budget` and at
   most `max_jobs` run, advances every running job by
   min(tick_steps, smallest remaining step count) steps in one batch, and
   retires the jobs that are done (positions copied back, device state
   freed).  Served jobs use capillary forces and device noise, with a fixed
   or an adaptive dt (max_travel); jobs of different precision advance in
   one batch per precision, side by side on the device.
   Each job's result equals running it alone with bdk_run_steps.  A job
   that cannot be admitted (bdk_create or the upload fails) or whose
   particles run away is finished as failed and the others go on;
   bdk_serve_step itself fails only when the device fails. */
typedef struct bdk_server bdk_server_t;

int bdk_serve_create(int32_t device, uint64_t hbm_budget, int32_t max_jobs, int64_t tick_steps,
                     bdk_server_t **out);
void bdk_serve_destroy(bdk_server_t *s);
const char *bdk_serve_last_error(const bdk_server_t *s);
int bdk_serve_submit(bdk_server_t *s, const bd_params *p, const int32_t *git, const double *rel, int64_t steps,
                     int64_t *job_id);
int bdk_serve_step(bdk_server_t *s, int32_t *n_active, int32_t *n_queued, int32_t *n_done);
/* Positions of a finished job ([n1+n2][2], species 1 first) and its admission
   and completion times in seconds after submission.  For a failed job the
   call returns the job's error code (bdk_serve_last_error says why) and
   writes only the times. */
int bdk_serve_result(bdk_server_t *s, int64_t job_id, int32_t *git, double *rel, double *t_admit, double *t_done);
int bdk_serve_stats(const bdk_server_t *s, int64_t *particle_steps, double *busy_s, uint64_t *hbm_used);
/* Host seconds the ticks spent so far in admission (bdk_create and the
   uploads), batch rebuilds, stepping (enqueue and wait) and retirement
   (positions read back, device state freed); any pointer may be NULL. */
int bdk_serve_timing(const bdk_server_t *s, double *admit_s, double *batch_s, double *run_s, double *retire_s);

#ifdef __cplusplus
}
#endif
#endif
