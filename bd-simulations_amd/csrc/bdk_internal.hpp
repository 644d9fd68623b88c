// Internal declarations shared by the HIP kernels (bdk_kernels.hip) and the
// C-ABI layer (bdk_api.cpp).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bdk.h"

namespace bdkx {

// Geometry and physics constants passed by value to every kernel.
// Derived values follow parameter.h:170-191 and gr:28.
struct Geo {
    int n1, n2, N;
    int nc;          // nachList_Zellen
    int ncell;       // nc*nc
    int Z, logZ;
    int scheme;
    int idgas, nowca, restrict_radial, ngp_fix, noise_mode, green_mode;
    int adaptive;    // max_travel finite: dt from the device min-reduction
    double B;        // nachList_Breite
    double L;
    double h;        // densGrid_Breite
    double drho;     // 1/h^2
    double cut2;     // zweihoch1_6*zweihoch1_6
    double s22, s12; // sigma ratios
    double e22f, e12f; // eps22_11*sigma11_22, eps12_11*sigma11_12 (dm:851, dm:1031)
    double maxpair;
    double f2_f1;
    double T, dt_max, mrd, noise_var;
    double sq_fixed;  // sqrt(2 T dt_max): the noise amplitude at fixed dt
    double lam2inv;  // 1/(lambda*lambda)
    double kap;
    int zero_mode_g; // G(0,0) = 0 when lambda > 0.2 L (dm:1389)
    int xcd;         // renumber workgroups per XCD (launch order only)
    float ncinv;     // 1/nc (fast exact division of cell indices)
    int fast_wca;    // FP32 forces and nc >= 6: k_wca_f32 (reads lrelf)
    int f32noise;    // FP32 forces: device noise evaluated in FP32
    int pos_mode;    // in-cell positions: 0 FP64, 1 FP32, 2 FP32 + Kahan-compensated increments
    double hinv;     // 1/h
    unsigned long long seed;
};

// Device-resident scalars of one simulation (graph replays read them).
// entries of the adaptive-dt log: the most steps replayed without reading it
constexpr int kDtLog = 4096;

struct DevScalars {
    unsigned long long step;      // step counter (noise counter)
    unsigned long long dt_bits;   // adaptive dt: running min, double bits
    unsigned long long exc_bits[2]; // max excursion of rel outside [0,B], double bits, by step parity
    unsigned long long dt_last;   // dt of the last update, double bits
    int nonfinite;                // a non-finite position appeared
    int fallback;                 // tiles that ran the unstaged WCA/deposit path
    int pad[4];
};

// Asynchronous pos writer (bdk_write_pos_async): pinned host copies of the
// slot arrays and scalars, the event recorded after the copies, and the host
// thread that formats and appends the files.  At most one write is in flight.
struct PosWriter {
    int *cell = nullptr;
    int *pid = nullptr;
    double *rel = nullptr;        // [cap][2]
    DevScalars *sc = nullptr;
    int cap = 0;
    hipStream_t st = nullptr;     // copies run here (the thread never waits on the handle's stream,
                                  // which may be under graph capture meanwhile)
    hipEvent_t ready = nullptr;   // on the handle's stream: the state to write
    hipEvent_t ev = nullptr;      // on st: copies done
    std::thread th;
    bool pending = false;         // copies enqueued, thread not yet joined
    int status = 0;
    std::string err;
};

struct DevBufs {
    // particles, storage order ("slot"); species from pid (< n1: species 1)
    int *cell = nullptr;      // packed cell gx*nc+gy (k_update writes slot q = list position q)
    double2 *rel = nullptr;   // offset inside the cell
    int *pid = nullptr;       // slot -> particle id
    void *fwca = nullptr;     // [N] real2, list order
    void *fkap = nullptr;     // [N] real2, list order
    double2 *noise_up = nullptr; // [N] by particle id (upload mode)
    double2 *noise_out = nullptr; // [N] by particle id (bdk_get_noise)
    int *rank = nullptr;      // [N] rank of the slot inside its cell
    // cell list
    int *count = nullptr;     // [ncell] particles per cell (kept current across steps)
    int *arr = nullptr;       // [ncell] arrivals of the current step (zero between steps)
    int *start = nullptr;     // [ncell+1]
    int *start_old = nullptr; // [ncell+1] previous starts: the slot arrays' grouping during a rebuild
    float2 *comp = nullptr;   // pos_mode 2: Kahan compensation by slot (written by k_update)
    float2 *lcomp = nullptr;  // pos_mode 2: the same in list order (written by k_fill)
    unsigned long long *exo = nullptr; // [64 x 16] max distance of a new position outside its previous cell (double bits)
    unsigned long long *dts = nullptr; // [64 x 16] adaptive dt: minima of the candidates by block (double bits;
                                       // all ones = none), reduced by the step's update (by k_dt_fold for
                                       // bdk_optimal_dt), re-armed at the end of the step
    double *dtlog = nullptr;  // [kDtLog] adaptive dt of step s at s mod kDtLog (written by the update)
    int *list = nullptr;      // [N] slots grouped by cell (slot | species << 31)
    // the state in list order (k_fill): cell, rel and particle id of list entry q
    int *lcell = nullptr;
    double2 *lrel = nullptr;
    int *lpid = nullptr;
    float2 *lrelf = nullptr;  // FP32 rel - B/2 of list entry q (FP32 WCA)
    int *bcount = nullptr;    // particles per block of 4096 cells, kept current by the steps
    int *own = nullptr;       // [nc+1] first mesh index owned by list cell c (deposit tiles)
    // mesh
    void *rho = nullptr;      // Z*Z real
    void *ht = nullptr;       // (Z/2)*Z complex, transposed half spectrum
    void *gt = nullptr;       // (Z/2)*Z complex pairs (Fx, Fy), transposed
    void *fmesh = nullptr;    // Z*Z real2 (Fx, Fy)
    // tables: tw[k] = exp(-2 pi i k/Z), sq[j] = kap sin(q_j h)/h, lap[j] = 4 sin^2(q_j h/2)/h^2, qv[j]
    void *tw = nullptr;
    double *sq = nullptr, *lap = nullptr, *qv = nullptr;  // FP64 1D tables
    float *sqf = nullptr, *lapf = nullptr, *qvf = nullptr;  // FP32 copies
    DevScalars *sc = nullptr;
};

// Deposit launch arguments (one tile per block).
struct DepArgs {
    int tc;        // list cells per tile side
    int seg_cap;   // staged list segments (rows, two per row when y wraps)
    int win_cap;   // owned mesh points per tile (capacity)
    int win_fast;  // FP32: points of the stencil window of a tile (0: no windowed path)
    float scale_f; // FP32 windowed path: 2^S with |charge| 2^S < 2^31
    double iscale_f;
    double scale_hi, scale_lo;      // 2^S1, 2^S2
    double iscale_hi, iscale_lo;    // 2^-S1, 2^-S2
    int reset;     // re-arm the per-step scalars (block 0)
    double q1, q2; // species charges (1 and f2_f1 for the capillary force)
    int slot;      // positions from the slot arrays grouped by start_old (during a rebuild), bound from exo
};

// One member of a batch, as the batched kernels read it (device memory).
struct RepDev {
    Geo g;
    DevBufs d;
    DepArgs dep;
    int dep_tx;                      // deposit tiles per row
    int rb, rc, rb2, sR1, sC, sR2;   // FFT rows per block and LDS row strides
    int scan_nb;                     // scan blocks
    int unit;                        // all sigma/epsilon ratios 1
};

// Block maps of the batched stages: blockIdx.x -> (member, block of member).
// BS_GATHER: particle blocks of the adaptive-dt members;
// BS_WCA: particle blocks of the WCA stage (its own block size).
enum BatchStage { BS_PART = 0, BS_CELL, BS_SCAN, BS_DEP, BS_R2C, BS_COL, BS_C2R, BS_GATHER, BS_WCA, BS_COUNT };

// The choices of the batch autotuner (bdk_batch_tune); none changes a result.
struct BatchTuning {
    int ept[3] = {8, 8, 8};  // FFT elements per thread (r2c, col, c2r)
    int wca_block = 256;     // threads per block of the WCA stage (128, 256 or 512)
    int dep_tile = 38;       // deposit tile: list cells per side such that a tile owns <= dep_tile^2 mesh points
    int persist = 0;         // blocks per launch of every stage (0: one block per map entry)
};

struct BatchPlan {
    int nmem = 0;
    int f32 = 1;
    RepDev *reps = nullptr;          // device [nmem]
    int2 *map[BS_COUNT] = {};        // device maps
    int nblk[BS_COUNT] = {};
    int nblk2[BS_COUNT] = {};        // leading entries of members with a power-of-two mesh
    int nblk16[BS_COUNT] = {};       // of those, the leading entries whose column and c2r passes are radix 16
    BatchTuning tune;
    int ept[3] = {8, 8, 8};          // the FFT elements per thread in use (at least the members' minimum)
    size_t dep_lds = 0, l1 = 0, lc = 0, l2 = 0;
};

}  // namespace bdkx

namespace bdkx {
struct StepInfoDev;
}

struct bdk {
    bd_params p;
    bdkx::Geo g;
    bdkx::DevBufs d;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;            // second stream of a step (mesh chain beside the list rebuild)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int dev = 0;
    bool pos_set[2] = {false, false};
    bool dev_valid = false;       // device particle state uploaded
    bool forces_ready = false;    // fwca / fmesh hold the forces of the current state
    std::vector<int32_t> hgit[2];
    std::vector<double> hrel[2];
    int scan_blocks = 0;
    hipGraph_t graph = nullptr;           // one step
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph_multi = nullptr;     // kGraphSteps steps
    hipGraphExec_t gexec_multi = nullptr;
    std::string err;
    int64_t host_step = 0;
    bdkx::StepInfoDev *info_dev = nullptr;  // step counters (bdk_step_info), allocated on first use
    bdkx::PosWriter writer;
};

namespace bdkx {

// Launchers implemented in bdk_kernels.hip.  All enqueue on h->stream.
hipError_t launch_wca(bdk_t *h, hipStream_t st = nullptr);  // WCA forces -> fwca (default: h->stream)
hipError_t launch_deposit(bdk_t *h, int reset, int slot = 0, hipStream_t st = nullptr);  // charges -> rho
                                                  // (reset: re-arm per-step scalars; slot: during a rebuild)
hipError_t launch_rearm(bdk_t *h);                // per-step scalars when no deposit runs
size_t deposit_lds_bytes(const bdk_t *h);          // LDS of one deposit block
hipError_t launch_fft(bdk_t *h);                  // rho -> fmesh (r2c, multiply, c2r)
// 0 r2c rows, 1 columns, 2 c2r rows (end_step: the c2r pass also ends the step)
hipError_t launch_fft_pass(bdk_t *h, int pass, hipStream_t st = nullptr, int end_step = 0);
hipError_t launch_gather(bdk_t *h, int with_dt, int fold = 1);  // fkap (+ dt candidates; + their fold)
hipError_t launch_update(bdk_t *h, int fused);    // noise, Euler step, refresh, counting
hipError_t launch_bin(bdk_t *h);                  // counting after an upload/refresh
hipError_t launch_own(bdk_t *h);                  // owned-mesh-index table of the deposit tiles
hipError_t launch_scan_fill(bdk_t *h);            // exclusive scan + fill of the cell list
hipError_t launch_refresh(bdk_t *h);              // refresh only (no motion)
hipError_t launch_noise_out(bdk_t *h);            // noise of the next step, by pid
// ordered-pair distance histograms [3][nbins] (11, 12, 22), zeroed by the caller
hipError_t launch_pair_hist(bdk_t *h, int nbins, double dr, unsigned long long *hist);
// rho(k) of one species: part [(Z/2+1)*nbins] scratch, out[nbins] shell sums
hipError_t launch_rhok(bdk_t *h, int species, double dqr, int nbins, double *part, double *out);
// radial density profiles: part [ceil(N/256)*8] scratch, centers[6], hist[6*nbins]
hipError_t launch_radial(bdk_t *h, int nbins, double dr, double *part, double *centers, int *hist);
// snapshot density grid (main_auswertung.cpp:214-375): acc [2][m*m] zeroed by
// the caller, rho [2][m*m]; stats: cnt [4*nbins + 7] zeroed by the caller
struct SnapThresh {
    double rho[3];
    double drho[4];
};
hipError_t launch_snap(bdk_t *h, int m, unsigned long long *acc, double *rho, int stats, int nbins, double dbin,
                       SnapThresh th, unsigned int *cnt);
// direct rho(k) sums (au:307-360) of one species for npairs index pairs ij;
// partial [npairs][ftrho_chunks(h)]
hipError_t launch_ftrho(bdk_t *h, const int2 *ij, int npairs, int species, double dq, double2 *partial, double *re,
                        double *im);
int ftrho_chunks(const bdk_t *h);
hipError_t set_kernel_attributes();
// device address of the allocation-free entry's scratch array on device
// `dev` (current), looked up once per device (bdk_noalloc.hip)
hipError_t noalloc_setup(int dev);
hipError_t build_tables(bdk_t *h);                // spectral tables in device memory

// batches (a batch is a view over member handles, each with its own buffers)
int batch_min_ept(int stage, int zmax);
hipError_t batch_build(const std::vector<bdk_t *> &mem, const BatchTuning &tune, BatchPlan &bp,
                       std::vector<RepDev> &reps, std::vector<int2> maps[BS_COUNT]);
hipError_t batch_stage(const BatchPlan &bp, int stage, hipStream_t st);
// one batched step of members whose forces are ready; with aux and the two
// events the mesh stages run beside the list rebuild
hipError_t batch_enqueue_step(const BatchPlan &bp, hipStream_t st, hipStream_t aux = nullptr,
                              hipEvent_t fork = nullptr, hipEvent_t join = nullptr);
// the forces of the members' current states (after an upload)
hipError_t batch_enqueue_forces(const BatchPlan &bp, hipStream_t st, hipStream_t aux = nullptr,
                                hipEvent_t fork = nullptr, hipEvent_t join = nullptr);

// Counters of one step (TimestepInfo, signaturen2.h:71-145), device side.
struct StepInfoDev {
    unsigned long long key[3];   // max |F|^2 (float bits) << 32 | list position: WCA, capillary, noise
    unsigned long long num_wca;  // candidate pairs i != j in the 3x3 cells
    int moves[2];                // particles that changed cell, per species
    int pid[3];                  // particle ids of the three maxima
    int nlist[2];                // species counts in the 3x3 cells of the WCA maximum
    double wca[2], kap[2], rnd[2];
};

// one full step as a sequence of launches; names (if non-null) get the
// kernel labels in launch order; info (if non-null) receives the counters
// Forces of the current state (WCA beside the deposit and the FFT passes),
// needed once after the state was set from the host; every step leaves the
// forces of its new state behind (h->forces_ready).
hipError_t enqueue_forces(bdk_t *h, std::vector<std::string> *names = nullptr, std::vector<hipEvent_t> *ev = nullptr);
// One step from a state whose forces are ready.
hipError_t enqueue_step(bdk_t *h, std::vector<std::string> *names, std::vector<hipEvent_t> *ev,
                        StepInfoDev *info = nullptr);

}  // namespace bdkx
