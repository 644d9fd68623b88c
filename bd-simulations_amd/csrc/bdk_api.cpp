// C-ABI of the MI355X Brownian-dynamics timestep (include/bdk.h).
//
// Host side of RunZustand::RunDynamik (signaturen2.h:33-252): parameter
// validation, device state, the particle storage permutation (particles are
// stored grouped by cell so that every kernel reads contiguous memory; the
// boundary always speaks the reference's per-species order), the step
// launch sequence, hipGraph replay of fixed-dt steps, the long-double
// checkpoint clock of zeitschritte_bis_obs (dm:444-455), initial states and
// the pos1_/pos2_ writer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "bdk_internal.hpp"

using bdkx::DevScalars;

namespace {

thread_local std::string g_create_error;

int fail(bdk_t *h, int code, const std::string &msg) {
    if (h)
        h->err = msg;
    else
        g_create_error = msg;
    return code;
}

int hipfail(bdk_t *h, hipError_t e, const char *where) {
    return fail(h, BDK_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIPCHK(h, call)                                      \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hipfail(h, e_, #call);  \
    } while (0)

bool is_pow2(int z) { return z > 0 && (z & (z - 1)) == 0; }

// Z = 2^a 3^b 5^c 7^d with a >= 1 (the FFT passes have radices 2, 4, 8, 3,
// 5 and 7); the reference takes any even Z through FFTW (parameter.h:24, its
// default 384 = 2^7 3)
bool mesh_size_ok(int z) {
    if (z < 8 || z > 4096 || (z & 1)) return false;
    for (int f : {3, 5, 7})
        while (z % f == 0) z /= f;
    return is_pow2(z);
}

int ilog2(int z) {
    int l = 0;
    while ((1 << l) < z) ++l;
    return l;
}

// parameter.h:170-191
struct Derived {
    double zh16, B, h, dq, drho;
    int nc;
};

Derived derive(const bd_params &p) {
    Derived d;
    d.zh16 = std::pow(2.0, 1.0 / 6.0);
    d.nc = (int)std::floor(p.L / d.zh16);
    d.B = p.L / ((double)d.nc);
    d.h = p.L / p.Z;
    d.dq = 2.0 * M_PI / p.L;
    d.drho = 1.0 / (d.h * d.h);
    return d;
}

std::string validate(const bd_params &p) {
    if (p.n1 < 0 || p.n2 < 0 || (int64_t)p.n1 + p.n2 < 1) return "n1, n2 must be >= 0 with n1+n2 >= 1";
    if ((int64_t)p.n1 + p.n2 > (1LL << 30)) return "n1+n2 too large";
    if (!(p.L > 0.0) || !std::isfinite(p.L)) return "L must be positive";
    if (p.scheme < 0 || p.scheme > 2) return "densGrid_Schema must be 0, 1 or 2 (NGP/CIC/TSC)";
    if (!mesh_size_ok(p.Z)) return "Z (FFT_samples) must be even, 2^a 3^b 5^c 7^d, in [8, 4096]";
    Derived d = derive(p);
    if (d.nc < 3) return "the box must hold at least 3 list cells per side (L >= 3*2^(1/6))";
    if ((int64_t)d.nc * d.nc > 4096LL * 4096LL) return "too many list cells (nc^2 > 4096^2)";
    if (!(p.lambda > 0.0)) return "lambda must be positive";
    if (!(p.T >= 0.0)) return "T must be >= 0";
    if (!(p.dt_max > 0.0)) return "dt_max must be positive";
    if (!(p.max_travel > 0.0)) return "max_travel must be positive (+inf for fixed dt)";
    if (p.precision < BDK_FP64 || p.precision > BDK_FP32_POS32K)
        return "precision must be BDK_FP64, BDK_FP32, BDK_FP32_POS32 or BDK_FP32_POS32K";
    if (p.noise_mode < 0 || p.noise_mode > 2) return "noise_mode must be DISC, GAUSS or UPLOAD";
    if (p.green_mode < 0 || p.green_mode > 1) return "green_mode must be DISCRETE or CONTINUOUS";
    if (p.noise_mode == BDK_NOISE_GAUSS && !(p.noise_var >= 0.0)) return "noise_var must be >= 0";
    {
        // the deposit owns at most (floor(B/h)+2)^2 mesh points per list cell tile in LDS
        const double side = std::floor(d.B / d.h) + 2.0;
        if (side * side * (p.precision == BDK_FP64 ? 16.0 : 8.0) > 120.0 * 1024.0)
            return "mesh spacing L/Z is too fine relative to the list cell (B/h > 84)";
    }
    return "";
}

size_t real_size(const bd_params &p) { return p.precision != BDK_FP64 ? 4 : 8; }

struct Footprint {
    size_t particles, cells, mesh, total;
};

Footprint footprint_of(const bd_params &p) {
    Derived d = derive(p);
    const size_t N = (size_t)p.n1 + p.n2, R = real_size(p);
    const size_t ncell = (size_t)d.nc * d.nc, Z = (size_t)p.Z;
    Footprint f;
    // cell, rel, pid, fwca, fkap, noise_up, noise_out, rank, list
    f.particles = N * (4 + 16 + 4 + 2 * R + 2 * R + 16 + 16 + 4 + 4 + 4 + 16 + 4 + 8) +
                  (p.precision == BDK_FP32_POS32K ? 2 * (N + 1) * 8 : 0);  // Kahan compensation
    f.cells = ncell * 16 + 8 + 4096 * 32 * 4 + 64 * 16 * 8;
    // rho, ht, gt, fmesh, tables
    f.mesh = Z * Z * R + (Z / 2) * Z * 2 * R + (Z / 2) * Z * 4 * R + Z * Z * 2 * R + Z * (2 * R + 3 * 8 + 3 * 4);
    f.total = f.particles + f.cells + f.mesh + sizeof(DevScalars);
    return f;
}

void release_streams(bdk_t *h) {
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_join) (void)hipEventDestroy(h->ev_join);
    if (h->aux) (void)hipStreamDestroy(h->aux);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    h->ev_fork = h->ev_join = nullptr;
    h->aux = h->stream = nullptr;
}

void free_dev(bdk_t *h) {
    void *ptrs[] = {h->d.cell, h->d.rel, h->d.pid, h->d.fwca, h->d.fkap, h->d.noise_up, h->d.noise_out,
                    h->d.rank, h->d.count, h->d.arr, h->d.start, h->d.list, h->d.lcell, h->d.lrel, h->d.lpid, h->d.lrelf, h->d.bcount, h->d.own, h->d.start_old, h->d.exo, h->d.dts, h->d.dtlog, h->d.comp, h->d.lcomp, h->d.rho, h->d.ht,
                    h->d.gt, h->d.fmesh, h->d.tw, h->d.sq, h->d.lap, h->d.qv, h->d.sqf, h->d.lapf,
                    h->d.qvf, h->d.sc};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    if (h->info_dev) (void)hipFree(h->info_dev);
    h->info_dev = nullptr;
    h->d = bdkx::DevBufs();
}

template <typename T> hipError_t dalloc(T *&p, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T) > 0 ? n * sizeof(T) : 16);
    p = (T *)q;
    return e;
}

hipError_t dalloc_v(void *&p, size_t bytes) { return hipMalloc(&p, bytes > 0 ? bytes : 16); }

// spectral tables (dm:578-594, dm:1384-1394), built in FP64 on the host
int build_tables_host(bdk_t *h) {
    const bd_params &p = h->p;
    const int Z = p.Z;
    Derived dv = derive(p);
    const double dx = dv.h;
    std::vector<double> sq(Z), lap(Z), qv(Z);
    for (int j = 0; j < Z; ++j) {
        double q = dv.dq * ((j + (int)(Z * 0.5)) % Z - 0.5 * Z);
        qv[j] = q;
        if (p.green_mode == BDK_GREEN_DISCRETE) {
            sq[j] = p.kap * std::sin(q * dx) / dx;
            lap[j] = 4 * std::sin(0.5 * q * dx) * std::sin(0.5 * q * dx) / dx / dx;
        } else {
            sq[j] = (j == Z / 2) ? 0.0 : p.kap * q;
            lap[j] = q * q;
        }
    }
    std::vector<float> sqf(sq.begin(), sq.end()), lapf(lap.begin(), lap.end()), qvf(qv.begin(), qv.end());
    HIPCHK(h, hipMemcpy(h->d.sq, sq.data(), Z * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d.lap, lap.data(), Z * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d.qv, qv.data(), Z * 8, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d.sqf, sqf.data(), Z * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d.lapf, lapf.data(), Z * 4, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->d.qvf, qvf.data(), Z * 4, hipMemcpyHostToDevice));
    // twiddles exp(-2 pi i k / Z)
    if (p.precision != BDK_FP64) {
        std::vector<float> tw(2 * Z);
        for (int k = 0; k < Z; ++k) {
            double a = 2.0 * M_PI * k / Z;
            tw[2 * k] = (float)std::cos(a);
            tw[2 * k + 1] = (float)-std::sin(a);
        }
        HIPCHK(h, hipMemcpy(h->d.tw, tw.data(), tw.size() * 4, hipMemcpyHostToDevice));
    } else {
        std::vector<double> tw(2 * Z);
        for (int k = 0; k < Z; ++k) {
            double a = 2.0 * M_PI * k / Z;
            tw[2 * k] = std::cos(a);
            tw[2 * k + 1] = -std::sin(a);
        }
        HIPCHK(h, hipMemcpy(h->d.tw, tw.data(), tw.size() * 8, hipMemcpyHostToDevice));
    }
    return BDK_OK;
}

// status code and message of the device flags (DevScalars::nonfinite)
int flag_status(int flags, std::string *msg) {
    if (flags & 2) {
        if (msg) *msg = "internal error: inconsistent cell lists (cell counts do not sum to N)";
        return BDK_ERR_HIP;
    }
    if (flags & 4) {
        if (msg)
            *msg = "internal error: a particle's charge stencil lay outside its deposit tile's window "
                   "(displacement bound too small)";
        return BDK_ERR_HIP;
    }
    if (flags) {
        if (msg) *msg = "a particle position became non-finite or ran away (|rel| >= 1e9); it was held in place";
        return BDK_ERR_NONFINITE;
    }
    return BDK_OK;
}

int check_nonfinite(bdk_t *h) {
    DevScalars sc;
    HIPCHK(h, hipMemcpyAsync(&sc, h->d.sc, sizeof(sc), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::string m;
    const int rc = flag_status(sc.nonfinite, &m);
    return rc ? fail(h, rc, m) : BDK_OK;
}

// a new particle state starts with no flags raised (they describe a run
// from the state before)
hipError_t clear_flags(bdk_t *h) {
    return hipMemsetAsync(&h->d.sc->nonfinite, 0, sizeof(int), h->stream);
}

// upload both species in storage order (grouped by cell, ascending id) and
// rebuild the cell lists (WCA_init, dm:459-526)
int upload(bdk_t *h) {
    const int n1 = h->p.n1, n2 = h->p.n2, N = n1 + n2, nc = h->g.nc;
    std::vector<int> cell(N);
    std::vector<double> rel(2 * (size_t)N);
    for (int sp = 0; sp < 2; ++sp) {
        const int n = sp == 0 ? n1 : n2, off = sp == 0 ? 0 : n1;
        for (int i = 0; i < n; ++i) {
            int gx = h->hgit[sp][2 * i], gy = h->hgit[sp][2 * i + 1];
            if (gx < 0 || gx >= nc || gy < 0 || gy >= nc)
                return fail(h, BDK_ERR_ARG, "git out of range [0, nachList_Zellen) for species " +
                                                std::to_string(sp + 1) + " particle " + std::to_string(i));
            double rx = h->hrel[sp][2 * i], ry = h->hrel[sp][2 * i + 1];
            if (!(std::fabs(rx) < 1e9 && std::fabs(ry) < 1e9))
                return fail(h, BDK_ERR_NONFINITE, "non-finite rel for species " + std::to_string(sp + 1) +
                                                      " particle " + std::to_string(i));
            cell[off + i] = gx * nc + gy;
            if (h->g.pos_mode) {  // FP32 storage of the in-cell positions
                rx = (double)(float)rx;
                ry = (double)(float)ry;
            }
            rel[2 * (size_t)(off + i)] = rx;
            rel[2 * (size_t)(off + i) + 1] = ry;
        }
    }
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cell[a] < cell[b]; });
    std::vector<int> scell(N), spid(N);
    std::vector<double> srel(2 * (size_t)N);
    for (int s = 0; s < N; ++s) {
        int p = order[s];
        scell[s] = cell[p];
        spid[s] = p;
        srel[2 * (size_t)s] = rel[2 * (size_t)p];
        srel[2 * (size_t)s + 1] = rel[2 * (size_t)p + 1];
    }
    HIPCHK(h, hipMemcpyAsync(h->d.cell, scell.data(), (size_t)N * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.pid, spid.data(), (size_t)N * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.rel, srel.data(), (size_t)N * 16, hipMemcpyHostToDevice, h->stream));
    if (h->d.comp) HIPCHK(h, hipMemsetAsync(h->d.comp, 0, (size_t)N * 8, h->stream));
    HIPCHK(h, hipMemsetAsync(h->d.count, 0, (size_t)h->g.ncell * 4, h->stream));
    HIPCHK(h, clear_flags(h));
    HIPCHK(h, bdkx::launch_bin(h));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->dev_valid = true;
    h->forces_ready = false;
    return BDK_OK;
}

int ensure(bdk_t *h) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    if (h->dev_valid) return BDK_OK;
    if (!h->pos_set[0] || (h->p.n2 > 0 && !h->pos_set[1]))
        return fail(h, BDK_ERR_STATE, "positions of both species must be set first (bdk_set_positions)");
    return upload(h);
}

// uploaded state whose forces are on the device (steps start from them)
int ensure_forces(bdk_t *h) {
    int rc = ensure(h);
    if (rc) return rc;
    if (h->forces_ready) return BDK_OK;
    HIPCHK(h, bdkx::enqueue_forces(h));
    h->forces_ready = true;
    return BDK_OK;
}

// list-ordered device array of real2 -> per-species [n][2] doubles
template <typename R>
int download_r2(bdk_t *h, const void *src, double *f1, double *f2) {
    const int N = h->g.N, n1 = h->p.n1;
    std::vector<R> buf(2 * (size_t)N);
    std::vector<int> pid(N);
    HIPCHK(h, hipMemcpyAsync(buf.data(), src, buf.size() * sizeof(R), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(pid.data(), h->d.lpid, (size_t)N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int s = 0; s < N; ++s) {
        int p = pid[s];
        double *dst = p < n1 ? (f1 ? f1 + 2 * (size_t)p : nullptr) : (f2 ? f2 + 2 * (size_t)(p - n1) : nullptr);
        if (!dst) continue;
        dst[0] = (double)buf[2 * (size_t)s];
        dst[1] = (double)buf[2 * (size_t)s + 1];
    }
    return BDK_OK;
}

int download_forces(bdk_t *h, const void *src, double *f1, double *f2) {
    return h->p.precision != BDK_FP64 ? download_r2<float>(h, src, f1, f2) : download_r2<double>(h, src, f1, f2);
}

// steps per replay of the multi-step graph
constexpr int kGraphSteps = 8;

void drop_graph(bdk_t *h) {
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    if (h->gexec_multi) (void)hipGraphExecDestroy(h->gexec_multi);
    if (h->graph_multi) (void)hipGraphDestroy(h->graph_multi);
    h->gexec = nullptr;
    h->graph = nullptr;
    h->gexec_multi = nullptr;
    h->graph_multi = nullptr;
}

int capture_steps(bdk_t *h, int nsteps, hipGraph_t *graph, hipGraphExec_t *exec) {
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int i = 0; i < nsteps && e == hipSuccess; ++i) e = bdkx::enqueue_step(h, nullptr, nullptr);  // forces ready
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess) return hipfail(h, e, "enqueue_step (capture)");
    if (e2 != hipSuccess) return hipfail(h, e2, "hipStreamEndCapture");
    *graph = g;
    HIPCHK(h, hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return BDK_OK;
}

// fixed-dt steps are replayed from hipGraphs: one of a single step and one
// of kGraphSteps steps (fewer launches per simulated time)
int ensure_graph(bdk_t *h) {
    if (h->gexec && h->gexec_multi) return BDK_OK;
    int rc = capture_steps(h, 1, &h->graph, &h->gexec);
    if (rc) return rc;
    return capture_steps(h, kGraphSteps, &h->graph_multi, &h->gexec_multi);
}

double dt_from_bits(unsigned long long b) {
    double v;
    std::memcpy(&v, &b, 8);
    return v;
}

}  // namespace

struct bdk_batch {
    std::vector<bdk_t *> mem;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;  // mesh stages beside the list rebuild
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int dev = 0;
    bdkx::BatchPlan bp;
    int ept_min[3] = {8, 8, 8};
    bdkx::BatchTuning tuning;  // the autotuner's current choice
    hipGraph_t graph = nullptr, graph_multi = nullptr;
    hipGraphExec_t gexec = nullptr, gexec_multi = nullptr;
    std::string err;

    int fail(int code, const std::string &msg) {
        err = msg;
        return code;
    }
    void drop_graphs() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (gexec_multi) (void)hipGraphExecDestroy(gexec_multi);
        if (graph_multi) (void)hipGraphDestroy(graph_multi);
        gexec = gexec_multi = nullptr;
        graph = graph_multi = nullptr;
    }
    static void free_plan(bdkx::BatchPlan &pl) {
        if (pl.reps) (void)hipFree(pl.reps);
        for (int k = 0; k < bdkx::BS_COUNT; ++k)
            if (pl.map[k]) (void)hipFree(pl.map[k]);
        pl = bdkx::BatchPlan();
    }
    void free_tables() { free_plan(bp); }
    // device copies of a plan's member table and block maps
    static hipError_t upload_plan(bdkx::BatchPlan &pl, const std::vector<bdkx::RepDev> &reps,
                                  const std::vector<int2> maps[bdkx::BS_COUNT]) {
        hipError_t e = hipMalloc((void **)&pl.reps, reps.size() * sizeof(bdkx::RepDev));
        if (e == hipSuccess)
            e = hipMemcpy(pl.reps, reps.data(), reps.size() * sizeof(bdkx::RepDev), hipMemcpyHostToDevice);
        for (int k = 0; k < bdkx::BS_COUNT && e == hipSuccess; ++k) {
            e = hipMalloc((void **)&pl.map[k], std::max<size_t>(1, maps[k].size()) * sizeof(int2));
            if (e == hipSuccess && !maps[k].empty())
                e = hipMemcpy(pl.map[k], maps[k].data(), maps[k].size() * sizeof(int2), hipMemcpyHostToDevice);
        }
        return e;
    }
    // device tables of the members for the tuning choice t
    hipError_t rebuild(const bdkx::BatchTuning &t) {
        drop_graphs();
        free_tables();
        int zmax = 8;
        for (bdk_t *h : mem) zmax = std::max(zmax, h->g.Z);
        for (int k = 0; k < 3; ++k) ept_min[k] = bdkx::batch_min_ept(k, zmax);
        tuning = t;
        std::vector<bdkx::RepDev> reps;
        std::vector<int2> maps[bdkx::BS_COUNT];
        hipError_t e = bdkx::batch_build(mem, t, bp, reps, maps);
        if (e == hipSuccess) e = upload_plan(bp, reps, maps);
        return e;
    }
    hipError_t capture(int nsteps, hipGraph_t *g, hipGraphExec_t *x) {
        hipError_t e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        for (int i = 0; i < nsteps && e == hipSuccess; ++i)
            e = bdkx::batch_enqueue_step(bp, stream, aux, ev_fork, ev_join);
        hipGraph_t gg = nullptr;
        hipError_t e2 = hipStreamEndCapture(stream, &gg);
        if (e == hipSuccess) e = e2;
        if (e != hipSuccess) return e;
        *g = gg;
        return hipGraphInstantiate(x, gg, nullptr, nullptr, 0);
    }
    // a member's asynchronous pos write copies on the member's stream: the
    // batch's steps wait for those copies
    hipError_t wait_writers() {
        for (bdk_t *h : mem)
            if (h->writer.pending) {
                hipError_t e = hipStreamWaitEvent(stream, h->writer.ev, 0);
                if (e != hipSuccess) return e;
            }
        return hipSuccess;
    }
    // members uploaded since their last step get their forces first.  When
    // only some need them (a server admitting jobs beside running ones), the
    // batched force stages run over those members alone, with the batch's
    // FFT tiling: a member's forces depend only on its own state and the
    // tiling, so they equal what the whole batch would compute.
    hipError_t ensure_forces() {
        std::vector<bdk_t *> need;
        for (bdk_t *h : mem)
            if (!h->forces_ready) need.push_back(h);
        if (need.empty()) return hipSuccess;
        hipError_t e = hipSuccess;
        if (need.size() == mem.size()) {
            e = bdkx::batch_enqueue_forces(bp, stream, aux, ev_fork, ev_join);
        } else {
            bdkx::BatchPlan sub;
            std::vector<bdkx::RepDev> reps;
            std::vector<int2> maps[bdkx::BS_COUNT];
            e = bdkx::batch_build(need, bp.tune, sub, reps, maps);
            if (e == hipSuccess) e = upload_plan(sub, reps, maps);
            if (e == hipSuccess) e = bdkx::batch_enqueue_forces(sub, stream, aux, ev_fork, ev_join);
            // the tables are freed once the force stages that read them are done
            const hipError_t es = hipStreamSynchronize(stream);
            if (e == hipSuccess) e = es;
            free_plan(sub);
        }
        if (e == hipSuccess)
            for (bdk_t *h : need) h->forces_ready = true;
        return e;
    }
    hipError_t ensure_graphs() {
        if (gexec && gexec_multi) return hipSuccess;
        hipError_t e = capture(1, &graph, &gexec);
        if (e == hipSuccess) e = capture(kGraphSteps, &graph_multi, &gexec_multi);
        return e;
    }
};

extern "C" {

int bdk_params_default(bd_params *p) {
    if (!p) return BDK_ERR_ARG;
    std::memset(p, 0, sizeof(*p));
    p->n1 = 1000;
    p->n2 = 1000;
    p->scheme = BDK_CIC;
    p->Z = 128;
    p->L = 64.0;
    p->lambda = 8.0;
    p->kap = 1.0;
    p->T = 1.0;
    p->dt_max = 1e-4;
    p->max_travel = INFINITY;
    p->sigma11_22 = 1.0;
    p->sigma11_12 = 1.0;
    p->eps22_11 = 1.0;
    p->eps12_11 = 1.0;
    p->f2_f1 = 0.5;
    p->max_pair_wca = 1.0e5;
    p->precision = BDK_FP32;
    p->noise_mode = BDK_NOISE_DISC;
    p->green_mode = BDK_GREEN_DISCRETE;
    p->noise_var = 1.0;
    p->seed = 1;
    return BDK_OK;
}

int bdk_params_derived(const bd_params *p, double *o) {
    if (!p || !o) return BDK_ERR_ARG;
    Derived d = derive(*p);
    o[0] = d.zh16;
    o[1] = d.nc;
    o[2] = d.B;
    o[3] = d.h;
    o[4] = d.dq;
    o[5] = d.drho;
    return BDK_OK;
}

int bdk_footprint(const bd_params *p, uint64_t *bytes) {
    if (!p || !bytes) return fail(nullptr, BDK_ERR_ARG, "null argument");
    std::string v = validate(*p);
    if (!v.empty()) return fail(nullptr, BDK_ERR_ARG, v);
    *bytes = footprint_of(*p).total;
    return BDK_OK;
}

int bdk_create(const bd_params *p, bdk_t **out) {
    if (!p || !out) return fail(nullptr, BDK_ERR_ARG, "null argument");
    *out = nullptr;
    std::string v = validate(*p);
    if (!v.empty()) return fail(nullptr, BDK_ERR_ARG, v);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, BDK_ERR_HIP, "no HIP device available");
    if (p->device < 0 || p->device >= ndev) return fail(nullptr, BDK_ERR_ARG, "device ordinal out of range");
    bdk_t *h = new bdk();
    h->p = *p;
    h->dev = p->device;
    hipError_t e = hipSetDevice(h->dev);
    if (e == hipSuccess) e = bdkx::noalloc_setup(h->dev);  // once per device: the no-allocation entry's scratch
    if (e != hipSuccess) {
        int rc = hipfail(nullptr, e, "hipSetDevice / scratch address");
        delete h;
        return rc;
    }
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) == hipSuccess && footprint_of(*p).total > freeb) {
        delete h;
        return fail(nullptr, BDK_ERR_NOMEM, "device footprint exceeds free device memory");
    }
    Derived dv = derive(*p);
    bdkx::Geo &g = h->g;
    std::memset(&g, 0, sizeof(g));
    g.n1 = p->n1;
    g.n2 = p->n2;
    g.N = p->n1 + p->n2;
    g.nc = dv.nc;
    g.ncell = dv.nc * dv.nc;
    g.Z = p->Z;
    g.logZ = ilog2(p->Z);
    g.scheme = p->scheme;
    g.idgas = p->idgas_only != 0;
    g.nowca = p->no_wca != 0;
    g.restrict_radial = p->restrict_radial != 0;
    g.ngp_fix = p->ngp_fix != 0;
    g.noise_mode = p->noise_mode;
    g.green_mode = p->green_mode;
    g.adaptive = std::isfinite(p->max_travel) ? 1 : 0;
    g.B = dv.B;
    g.L = p->L;
    g.h = dv.h;
    g.drho = dv.drho;
    g.cut2 = dv.zh16 * dv.zh16;
    g.s22 = p->sigma11_22;
    g.s12 = p->sigma11_12;
    g.e22f = p->eps22_11 * p->sigma11_22;
    g.e12f = p->eps12_11 * p->sigma11_12;
    g.maxpair = p->max_pair_wca;
    g.f2_f1 = p->f2_f1;
    g.T = p->T;
    g.dt_max = p->dt_max;
    g.sq_fixed = std::sqrt(2.0 * p->T * p->dt_max);
    g.mrd = p->max_travel;
    g.noise_var = p->noise_var;
    g.lam2inv = 1.0 / (p->lambda * p->lambda);
    g.kap = p->kap;
    g.zero_mode_g = p->lambda > 0.2 * p->L;
    g.ncinv = 1.0f / (float)g.nc;
    g.fast_wca = (p->precision != BDK_FP64 && g.nc >= 6) ? 1 : 0;
    g.f32noise = p->precision != BDK_FP64 ? 1 : 0;
    g.pos_mode = p->precision == BDK_FP32_POS32 ? 1 : (p->precision == BDK_FP32_POS32K ? 2 : 0);
    g.hinv = 1.0 / g.h;
    {
        // BDK_XCD_REMAP=0 keeps the hardware workgroup order (for comparison)
        const char *e = getenv("BDK_XCD_REMAP");
        g.xcd = (e && e[0] == '0') ? 0 : 1;
    }
    g.seed = p->seed;

    const size_t N = g.N, R = real_size(*p), ncell = g.ncell, Z = g.Z;
    bdkx::DevBufs &d = h->d;
    int rc = BDK_OK;
#define ALLOC(expr)                                                \
    if (rc == BDK_OK) {                                            \
        hipError_t ee = (expr);                                    \
        if (ee != hipSuccess) rc = hipfail(nullptr, ee, #expr);    \
    }
    ALLOC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    ALLOC(hipStreamCreateWithFlags(&h->aux, hipStreamNonBlocking));
    ALLOC(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    ALLOC(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    ALLOC(dalloc(d.cell, N));
    ALLOC(dalloc(d.rel, N));
    ALLOC(dalloc(d.pid, N));
    ALLOC(dalloc_v(d.fwca, N * 2 * R));
    ALLOC(dalloc_v(d.fkap, N * 2 * R));
    ALLOC(dalloc(d.noise_up, N));
    ALLOC(dalloc(d.noise_out, N));
    ALLOC(dalloc(d.rank, N));
    ALLOC(dalloc(d.count, ncell));
    ALLOC(dalloc(d.arr, ncell));
    ALLOC(dalloc(d.start, ncell + 1));
    ALLOC(dalloc(d.start_old, ncell + 1));
    if (h->g.pos_mode == 2) {
        ALLOC(dalloc(d.comp, N + 1));
        ALLOC(dalloc(d.lcomp, N + 1));
    }
    ALLOC(dalloc(d.exo, 64 * 16));
    ALLOC(dalloc(d.dts, 64 * 16));
    ALLOC(dalloc(d.dtlog, bdkx::kDtLog));
    // list-ordered arrays carry one padding entry (read ahead by k_wca_f32)
    ALLOC(dalloc(d.list, N + 1));
    ALLOC(dalloc(d.lcell, N + 1));
    ALLOC(dalloc(d.lrel, N + 1));
    ALLOC(dalloc(d.lpid, N + 1));
    ALLOC(dalloc(d.lrelf, N + 1));
    ALLOC(dalloc(d.bcount, 4096 * 32));  // one entry per 4096 cells, 128 bytes apart
    ALLOC(dalloc(d.own, h->g.nc + 1));
    ALLOC(dalloc_v(d.rho, Z * Z * R));
    ALLOC(dalloc_v(d.ht, (Z / 2) * Z * 2 * R));
    ALLOC(dalloc_v(d.gt, (Z / 2) * Z * 4 * R));
    ALLOC(dalloc_v(d.fmesh, Z * Z * 2 * R));
    ALLOC(dalloc_v(d.tw, Z * 2 * R));
    ALLOC(dalloc(d.sq, Z));
    ALLOC(dalloc(d.lap, Z));
    ALLOC(dalloc(d.qv, Z));
    ALLOC(dalloc(d.sqf, Z));
    ALLOC(dalloc(d.lapf, Z));
    ALLOC(dalloc(d.qvf, Z));
    ALLOC(dalloc(d.sc, 1));
    ALLOC(bdkx::set_kernel_attributes());
    ALLOC(bdkx::launch_own(h));
#undef ALLOC
    if (rc != BDK_OK) {
        free_dev(h);
        release_streams(h);
        delete h;
        return rc;
    }
    DevScalars sc;
    std::memset(&sc, 0, sizeof(sc));
    e = hipMemcpy(d.sc, &sc, sizeof(sc), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d.count, 0, ncell * 4);
    if (e == hipSuccess) e = hipMemset(d.exo, 0, 64 * 16 * 8);
    if (e == hipSuccess) e = hipMemset(d.dts, 0xff, 64 * 16 * 8);
    if (e == hipSuccess) e = hipMemset(d.dtlog, 0, bdkx::kDtLog * 8);
    if (e == hipSuccess && d.comp) e = hipMemset(d.comp, 0, (N + 1) * 8);
    if (e == hipSuccess && d.lcomp) e = hipMemset(d.lcomp, 0, (N + 1) * 8);
    if (e == hipSuccess) e = hipMemset(d.arr, 0, ncell * 4);
    if (e == hipSuccess) e = hipMemset(d.list, 0, (N + 1) * 4);
    if (e == hipSuccess) e = hipMemset(d.lrelf, 0, (N + 1) * 8);
    if (e == hipSuccess) e = hipMemset(d.fwca, 0, N * 2 * R);
    if (e == hipSuccess) e = hipMemset(d.fkap, 0, N * 2 * R);
    if (e == hipSuccess) e = hipMemset(d.noise_up, 0, N * 16);
    if (e == hipSuccess) e = hipMemset(d.rho, 0, Z * Z * R);
    if (e == hipSuccess) e = hipMemset(d.fmesh, 0, Z * Z * 2 * R);
    if (e != hipSuccess) {
        rc = hipfail(nullptr, e, "initialising device state");
        free_dev(h);
        release_streams(h);
        delete h;
        return rc;
    }
    rc = build_tables_host(h);
    if (rc != BDK_OK) {
        g_create_error = h->err;
        free_dev(h);
        release_streams(h);
        delete h;
        return rc;
    }
    h->hgit[0].assign(2 * (size_t)p->n1, 0);
    h->hgit[1].assign(2 * (size_t)p->n2, 0);
    h->hrel[0].assign(2 * (size_t)p->n1, 0.0);
    h->hrel[1].assign(2 * (size_t)p->n2, 0.0);
    h->pos_set[1] = p->n2 == 0;
    *out = h;
    return BDK_OK;
}

static void free_writer(bdk_t *h);

void bdk_destroy(bdk_t *h) {
    if (!h) return;
    (void)hipSetDevice(h->dev);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->aux) (void)hipStreamSynchronize(h->aux);
    free_writer(h);
    drop_graph(h);
    free_dev(h);
    release_streams(h);
    delete h;
}

const char *bdk_last_error(const bdk_t *h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int bdk_set_positions(bdk_t *h, int species, int32_t n, const int32_t *git, const double *rel) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    if (species != 1 && species != 2) return fail(h, BDK_ERR_ARG, "species must be 1 or 2");
    const int sp = species - 1;
    const int want = sp == 0 ? h->p.n1 : h->p.n2;
    if (n != want) return fail(h, BDK_ERR_ARG, "n does not match N" + std::to_string(species));
    if (n > 0 && (!git || !rel)) return fail(h, BDK_ERR_ARG, "null position array");
    if (n > 0) {
        std::memcpy(h->hgit[sp].data(), git, sizeof(int32_t) * 2 * (size_t)n);
        std::memcpy(h->hrel[sp].data(), rel, sizeof(double) * 2 * (size_t)n);
    }
    h->pos_set[sp] = true;
    h->dev_valid = false;
    return BDK_OK;
}

int bdk_get_positions(bdk_t *h, int species, int32_t n, int32_t *git, double *rel) {
    int rc = ensure(h);
    if (rc) return rc;
    if (species != 1 && species != 2) return fail(h, BDK_ERR_ARG, "species must be 1 or 2");
    const int n1 = h->p.n1, N = h->g.N, nc = h->g.nc;
    if (n != (species == 1 ? n1 : h->p.n2)) return fail(h, BDK_ERR_ARG, "n does not match");
    std::vector<int> cell(N), pid(N);
    std::vector<double> r(2 * (size_t)N);
    HIPCHK(h, hipMemcpyAsync(cell.data(), h->d.cell, (size_t)N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(pid.data(), h->d.pid, (size_t)N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(r.data(), h->d.rel, (size_t)N * 16, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int s = 0; s < N; ++s) {
        int p = pid[s];
        int i = species == 1 ? p : p - n1;
        if ((species == 1) != (p < n1)) continue;
        if (git) {
            git[2 * (size_t)i] = cell[s] / nc;
            git[2 * (size_t)i + 1] = cell[s] % nc;
        }
        if (rel) {
            rel[2 * (size_t)i] = r[2 * (size_t)s];
            rel[2 * (size_t)i + 1] = r[2 * (size_t)s + 1];
        }
    }
    return check_nonfinite(h);
}

int bdk_compute_wca(bdk_t *h, double *f1, double *f2) {
    int rc = ensure(h);
    if (rc) return rc;
    if (h->g.idgas) {
        HIPCHK(h, hipMemsetAsync(h->d.fwca, 0, (size_t)h->g.N * 2 * real_size(h->p), h->stream));
    } else {
        HIPCHK(h, bdkx::launch_wca(h));
    }
    return download_forces(h, h->d.fwca, f1, f2);
}

int bdk_compute_capillary(bdk_t *h, double *f1, double *f2) {
    int rc = ensure(h);
    if (rc) return rc;
    if (h->g.idgas) {
        HIPCHK(h, hipMemsetAsync(h->d.fkap, 0, (size_t)h->g.N * 2 * real_size(h->p), h->stream));
    } else {
        HIPCHK(h, bdkx::launch_deposit(h, 0));
        HIPCHK(h, bdkx::launch_fft(h));
        HIPCHK(h, bdkx::launch_gather(h, 0));
    }
    return download_forces(h, h->d.fkap, f1, f2);
}

int bdk_set_noise(bdk_t *h, const double *n1, const double *n2) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    if (h->p.noise_mode != BDK_NOISE_UPLOAD) return fail(h, BDK_ERR_STATE, "noise_mode is not BDK_NOISE_UPLOAD");
    if ((h->p.n1 && !n1) || (h->p.n2 && !n2)) return fail(h, BDK_ERR_ARG, "null noise array");
    if (h->p.n1)
        HIPCHK(h, hipMemcpyAsync(h->d.noise_up, n1, (size_t)h->p.n1 * 16, hipMemcpyHostToDevice, h->stream));
    if (h->p.n2)
        HIPCHK(h, hipMemcpyAsync(h->d.noise_up + h->p.n1, n2, (size_t)h->p.n2 * 16, hipMemcpyHostToDevice,
                                 h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return BDK_OK;
}

int bdk_get_noise(bdk_t *h, double *n1, double *n2) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    HIPCHK(h, bdkx::launch_noise_out(h));
    std::vector<double> buf(2 * (size_t)h->g.N);
    HIPCHK(h, hipMemcpyAsync(buf.data(), h->d.noise_out, buf.size() * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (n1) std::memcpy(n1, buf.data(), (size_t)h->p.n1 * 16);
    if (n2) std::memcpy(n2, buf.data() + 2 * (size_t)h->p.n1, (size_t)h->p.n2 * 16);
    return BDK_OK;
}

int bdk_optimal_dt(bdk_t *h, double *dt) {
    int rc = ensure(h);
    if (rc) return rc;
    if (!dt) return fail(h, BDK_ERR_ARG, "null dt");
    if (!h->g.adaptive) {
        *dt = h->p.dt_max;
        return BDK_OK;
    }
    if (h->g.idgas) {
        HIPCHK(h, hipMemsetAsync(h->d.fwca, 0, (size_t)h->g.N * 2 * real_size(h->p), h->stream));
        HIPCHK(h, hipMemsetAsync(h->d.fmesh, 0, (size_t)h->g.Z * h->g.Z * 2 * real_size(h->p), h->stream));
        HIPCHK(h, bdkx::launch_rearm(h));
    } else {
        HIPCHK(h, bdkx::launch_wca(h));
        HIPCHK(h, bdkx::launch_deposit(h, 1));
        HIPCHK(h, bdkx::launch_fft(h));
    }
    HIPCHK(h, bdkx::launch_gather(h, 1));
    DevScalars sc;
    HIPCHK(h, hipMemcpyAsync(&sc, h->d.sc, sizeof(sc), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *dt = dt_from_bits(sc.dt_bits);
    return BDK_OK;
}

int bdk_step(bdk_t *h, double tmax, double *dt_out) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    if (!(tmax > 0.0)) {  // dm:80-83
        if (dt_out) *dt_out = 1.0e10;
        return fail(h, BDK_ERR_ARG, "Fehler: Negative Zeitschrittweite gefordert (tmax <= 0)");
    }
    int rc = ensure_forces(h);
    if (rc) return rc;
    if (h->g.idgas) {
        HIPCHK(h, hipMemsetAsync(h->d.fmesh, 0, (size_t)h->g.Z * h->g.Z * 2 * real_size(h->p), h->stream));
    }
    HIPCHK(h, bdkx::enqueue_step(h, nullptr, nullptr));
    h->host_step++;
    if (h->g.adaptive) {
        DevScalars sc;
        HIPCHK(h, hipMemcpyAsync(&sc, h->d.sc, sizeof(sc), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (dt_out) *dt_out = dt_from_bits(sc.dt_last);
    } else if (dt_out) {
        *dt_out = h->p.dt_max;
    }
    return BDK_OK;
}

int bdk_step_info(bdk_t *h, double tmax, bdk_step_info_t *out) {
    if (!h || !out) return fail(h, BDK_ERR_ARG, "null argument");
    if (!(tmax > 0.0)) return fail(h, BDK_ERR_ARG, "Fehler: Negative Zeitschrittweite gefordert (tmax <= 0)");
    int rc = ensure_forces(h);
    if (rc) return rc;
    if (!h->info_dev) HIPCHK(h, hipMalloc((void **)&h->info_dev, sizeof(bdkx::StepInfoDev)));
    bdkx::StepInfoDev *dinfo = h->info_dev;
    bdkx::StepInfoDev hi;
    hipError_t e = hipMemsetAsync(dinfo, 0, sizeof hi, h->stream);
    if (e == hipSuccess && h->g.idgas)
        e = hipMemsetAsync(h->d.fmesh, 0, (size_t)h->g.Z * h->g.Z * 2 * real_size(h->p), h->stream);
    if (e == hipSuccess) e = bdkx::enqueue_step(h, nullptr, nullptr, dinfo);
    DevScalars sc;
    if (e == hipSuccess) e = hipMemcpyAsync(&hi, dinfo, sizeof hi, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&sc, h->d.sc, sizeof sc, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hipfail(h, e, "bdk_step_info");
    h->host_step++;
    const int n1 = h->p.n1;
    auto who = [&](int pid, int32_t *type, int32_t *index) {
        *type = pid < n1 ? 1 : 2;
        *index = pid < n1 ? pid : pid - n1;
    };
    out->num_wca = (int64_t)hi.num_wca;
    out->max_wca[0] = hi.wca[0];
    out->max_wca[1] = hi.wca[1];
    who(hi.pid[0], &out->max_wca_type, &out->max_wca_index);
    out->max_wca_nlist[0] = hi.nlist[0];
    out->max_wca_nlist[1] = hi.nlist[1];
    out->max_kap[0] = hi.kap[0];
    out->max_kap[1] = hi.kap[1];
    who(hi.pid[1], &out->max_kap_type, &out->max_kap_index);
    out->max_rnd[0] = hi.rnd[0];
    out->max_rnd[1] = hi.rnd[1];
    who(hi.pid[2], &out->max_rnd_type, &out->max_rnd_index);
    out->dt = h->g.adaptive ? dt_from_bits(sc.dt_last) : h->p.dt_max;
    out->cell_changes[0] = hi.moves[0];
    out->cell_changes[1] = hi.moves[1];
    return check_nonfinite(h);
}

// adaptive dt whose steps can be replayed from the step graphs (the dt
// stays on the device; the host reads the log of the dt's afterwards)
static bool adaptive_graphable(const bdk_t *h) {
    return h->g.adaptive && !h->g.idgas && h->p.noise_mode != BDK_NOISE_UPLOAD;
}

// k <= kDtLog adaptive steps replayed from the step graphs, then their dt's
// read from the device log and added to *t in step order, in long double
// as zeitschritte_bis_obs adds them (dm:444-455).  Forces must be ready.
static int run_logged(bdk_t *h, int64_t k, long double *t) {
    int rc = ensure_graph(h);
    if (rc) return rc;
    int64_t i = 0;
    for (; i + kGraphSteps <= k; i += kGraphSteps) HIPCHK(h, hipGraphLaunch(h->gexec_multi, h->stream));
    for (; i < k; ++i) HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
    std::vector<double> log(bdkx::kDtLog);
    DevScalars sc;
    HIPCHK(h, hipMemcpyAsync(log.data(), h->d.dtlog, bdkx::kDtLog * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(&sc, h->d.sc, sizeof(sc), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->host_step += k;
    const unsigned long long s0 = sc.step - (unsigned long long)k;
    for (int64_t j = 0; j < k; ++j) *t += (long double)log[(s0 + (unsigned long long)j) & (bdkx::kDtLog - 1)];
    return BDK_OK;
}

int bdk_run_steps(bdk_t *h, int64_t k, double *t_out) {
    int rc = ensure_forces(h);
    if (rc) return rc;
    if (k < 0) return fail(h, BDK_ERR_ARG, "k must be >= 0");
    long double t = 0.0L;
    if (adaptive_graphable(h)) {
        for (int64_t i = 0; i < k;) {
            const int64_t b = std::min<int64_t>(k - i, bdkx::kDtLog);
            rc = run_logged(h, b, &t);
            if (rc) return rc;
            i += b;
        }
    } else if (!h->g.adaptive && h->p.noise_mode != BDK_NOISE_UPLOAD) {
        rc = ensure_graph(h);
        if (rc) return rc;
        int64_t i = 0;
        for (; i + kGraphSteps <= k; i += kGraphSteps) HIPCHK(h, hipGraphLaunch(h->gexec_multi, h->stream));
        for (; i < k; ++i) HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
        for (int64_t j = 0; j < k; ++j) t += (long double)h->p.dt_max;
        h->host_step += k;
    } else {
        for (int64_t i = 0; i < k; ++i) {
            double dt = 0.0;
            rc = bdk_step(h, 1.0, &dt);
            if (rc) return rc;
            t += (long double)dt;
        }
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (t_out) *t_out = (double)t;
    return check_nonfinite(h);
}

int bdk_set_max_travel(bdk_t *h, double max_travel) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    if (!(max_travel > 0.0)) return fail(h, BDK_ERR_ARG, "max_travel must be > 0 (+inf: fixed dt)");
    if (h->p.max_travel == max_travel) return BDK_OK;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->p.max_travel = max_travel;
    h->g.adaptive = std::isfinite(max_travel) ? 1 : 0;
    h->g.mrd = max_travel;
    drop_graph(h);  // the captured steps carry the old Geo
    return BDK_OK;
}

int bdk_steps_for_interval(double obs_dt, double dt, int64_t *steps_out) {
    if (!steps_out || !(dt > 0.0)) return BDK_ERR_ARG;
    long double t = 0.0L;
    int64_t n = 0;
    while (t < obs_dt) {
        t += (long double)dt;
        ++n;
    }
    *steps_out = n;
    return BDK_OK;
}

int bdk_run_interval(bdk_t *h, double obs_dt, int64_t *steps_out) {
    return bdk_run_interval_clock(h, obs_dt, steps_out, nullptr, nullptr);
}

int bdk_run_interval_clock(bdk_t *h, double obs_dt, int64_t *steps_out, double *t_hi, double *t_lo) {
    int rc = ensure(h);
    if (rc) return rc;
    int64_t n = 0;
    long double t = 0.0L;
    if (!h->g.adaptive) {
        bdk_steps_for_interval(obs_dt, h->p.dt_max, &n);
        rc = bdk_run_steps(h, n, nullptr);
        if (rc) return rc;
        for (int64_t j = 0; j < n; ++j) t += (long double)h->p.dt_max;
    } else {
        if (adaptive_graphable(h)) {
            // dt <= dt_max, so while the clock is more than (k + 2) dt_max
            // short of obs_dt the next k steps are all taken: replay them
            // from the graphs and add their logged dt's in order
            rc = ensure_forces(h);
            if (rc) return rc;
            for (;;) {
                const long double room = ((long double)obs_dt - t) / (long double)h->p.dt_max;
                if (!(room > 18.0L)) break;
                const int64_t k = room > (long double)(bdkx::kDtLog + 2) ? bdkx::kDtLog : (int64_t)room - 2;
                rc = run_logged(h, k, &t);
                if (rc) return rc;
                n += k;
            }
        }
        while (t < obs_dt) {
            double dt = 0.0;
            rc = bdk_step(h, (double)((long double)obs_dt - t), &dt);
            if (rc) return rc;
            t += (long double)dt;
            ++n;
        }
        rc = check_nonfinite(h);
        if (rc) return rc;
    }
    if (steps_out) *steps_out = n;
    // the clock as an unevaluated sum of two doubles: the 64-bit significand
    // of the long double fits in the two 53-bit ones, so hi + lo is exact
    const double hi = (double)t;
    if (t_hi) *t_hi = hi;
    if (t_lo) *t_lo = (double)(t - (long double)hi);
    return BDK_OK;
}

int bdk_refresh(bdk_t *h) {
    int rc = ensure(h);
    if (rc) return rc;
    HIPCHK(h, bdkx::launch_refresh(h));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->forces_ready = false;
    return check_nonfinite(h);
}

int bdk_prepare(bdk_t *h) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    HIPCHK(h, hipSetDevice(h->dev));
    return ensure_forces(h);
}

int bdk_sync(bdk_t *h) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return BDK_OK;
}

int64_t bdk_step_counter(const bdk_t *h) { return h ? h->host_step : -1; }

int bdk_get_cells(bdk_t *h, int32_t *start, int32_t *ids) {
    int rc = ensure(h);
    if (rc) return rc;
    const int N = h->g.N, ncell = h->g.ncell;
    std::vector<int> st(ncell + 1), list(N), pid(N);
    HIPCHK(h, hipMemcpyAsync(st.data(), h->d.start, (size_t)(ncell + 1) * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(list.data(), h->d.list, (size_t)N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(pid.data(), h->d.pid, (size_t)N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (start) std::memcpy(start, st.data(), (size_t)(ncell + 1) * 4);
    if (ids) {
        for (int q = 0; q < N; ++q) ids[q] = pid[list[q] & 0x7fffffff];
        for (int c = 0; c < ncell; ++c) std::sort(ids + st[c], ids + st[c + 1]);
    }
    return BDK_OK;
}

int bdk_get_mesh(bdk_t *h, int which, double *out) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    if (!out || which < 0 || which > 2) return fail(h, BDK_ERR_ARG, "bad mesh request");
    const size_t Z2 = (size_t)h->g.Z * h->g.Z;
    const bool f32 = h->p.precision != BDK_FP64;
    if (which == BDK_MESH_RHO) {
        if (f32) {
            std::vector<float> b(Z2);
            HIPCHK(h, hipMemcpyAsync(b.data(), h->d.rho, Z2 * 4, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            for (size_t i = 0; i < Z2; ++i) out[i] = b[i];
        } else {
            HIPCHK(h, hipMemcpyAsync(out, h->d.rho, Z2 * 8, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
        }
    } else {
        const int c = which == BDK_MESH_FX ? 0 : 1;
        if (f32) {
            std::vector<float> b(2 * Z2);
            HIPCHK(h, hipMemcpyAsync(b.data(), h->d.fmesh, Z2 * 8, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            for (size_t i = 0; i < Z2; ++i) out[i] = b[2 * i + c];
        } else {
            std::vector<double> b(2 * Z2);
            HIPCHK(h, hipMemcpyAsync(b.data(), h->d.fmesh, Z2 * 16, hipMemcpyDeviceToHost, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            for (size_t i = 0; i < Z2; ++i) out[i] = b[2 * i + c];
        }
    }
    return BDK_OK;
}

int bdk_profile_steps(bdk_t *h, int64_t k, int32_t nk, double *ms, char *names, int32_t names_cap,
                      int32_t *n_out) {
    int rc = ensure_forces(h);
    if (rc) return rc;
    if (h->p.noise_mode == BDK_NOISE_UPLOAD) return fail(h, BDK_ERR_STATE, "profiling needs a Philox noise mode");
    std::vector<double> acc;
    std::vector<std::string> labels;
    for (int64_t it = 0; it < k; ++it) {
        std::vector<std::string> nm;
        std::vector<hipEvent_t> ev;
        HIPCHK(h, bdkx::enqueue_step(h, &nm, &ev));
        h->host_step++;
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (acc.empty()) {
            acc.assign(ev.size() - 1, 0.0);
            labels.assign(nm.begin() + 1, nm.end());
        }
        for (size_t i = 1; i < ev.size(); ++i) {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, ev[i - 1], ev[i]);
            acc[i - 1] += t;
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    if (n_out) *n_out = (int32_t)acc.size();
    for (size_t i = 0; i < acc.size() && (int32_t)i < nk; ++i) ms[i] = k ? acc[i] / (double)k : 0.0;
    if (names && names_cap > 0) {
        std::string all;
        for (size_t i = 0; i < labels.size(); ++i) all += labels[i] + (i + 1 < labels.size() ? "\n" : "");
        std::strncpy(names, all.c_str(), (size_t)names_cap - 1);
        names[names_cap - 1] = 0;
    }
    return check_nonfinite(h);
}

int bdk_profile_kernel(bdk_t *h, int kernel, int32_t reps, double *ms_per_launch) {
    int rc = ensure_forces(h);
    if (rc) return rc;
    if (reps < 1 || kernel < 0 || kernel > 7 || !ms_per_launch) return fail(h, BDK_ERR_ARG, "bad profile request");
    // k_update changes positions, cells, ranks, counts and scalars: snapshot
    // them (the copies are freed on every return path)
    struct Snap {
        void *p;
        void *src;
        size_t n;
    };
    struct Snaps {
        std::vector<Snap> v;
        ~Snaps() {
            for (Snap &sn : v) (void)hipFree(sn.p);
        }
    } snaps;
    if (kernel == BDK_K_UPDATE) {
        const size_t N = (size_t)h->g.N, nc2 = (size_t)h->g.ncell;
        Snap list[] = {{nullptr, h->d.cell, N * 4}, {nullptr, h->d.rel, N * 16}, {nullptr, h->d.pid, N * 4},
                       {nullptr, h->d.rank, N * 4},
                       {nullptr, h->d.count, nc2 * 4}, {nullptr, h->d.arr, nc2 * 4},
                       {nullptr, h->d.bcount, 4096 * 32 * 4}, {nullptr, h->d.exo, 64 * 16 * 8},
                       {nullptr, h->d.sc, sizeof(DevScalars)}};
        for (Snap sn : list) {
            HIPCHK(h, hipMalloc(&sn.p, sn.n));
            snaps.v.push_back(sn);
            HIPCHK(h, hipMemcpyAsync(sn.p, sn.src, sn.n, hipMemcpyDeviceToDevice, h->stream));
        }
        // forces of the current state for the update to consume
        HIPCHK(h, bdkx::launch_wca(h));
        HIPCHK(h, bdkx::launch_deposit(h, 1));
        HIPCHK(h, bdkx::launch_fft(h));
    }
    if (kernel == BDK_K_STEP && !h->g.adaptive && h->p.noise_mode != BDK_NOISE_UPLOAD) {
        rc = ensure_graph(h);
        if (rc) return rc;
    }
    auto launch = [&]() -> hipError_t {
        switch (kernel) {
        case BDK_K_WCA: return bdkx::launch_wca(h);
        case BDK_K_DEPOSIT: return bdkx::launch_deposit(h, 0);
        case BDK_K_R2C: return bdkx::launch_fft_pass(h, 0);
        case BDK_K_COL: return bdkx::launch_fft_pass(h, 1);
        case BDK_K_C2R: return bdkx::launch_fft_pass(h, 2);
        case BDK_K_GATHER: return bdkx::launch_gather(h, 0);
        case BDK_K_UPDATE: return bdkx::launch_update(h, h->g.adaptive ? 0 : 1);
        default:
            if (h->gexec) return hipGraphLaunch(h->gexec, h->stream);
            return bdkx::enqueue_step(h, nullptr, nullptr);
        }
    };
    struct Events {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    HIPCHK(h, launch());  // warm-up
    HIPCHK(h, hipEventCreate(&ev.e0));
    HIPCHK(h, hipEventCreate(&ev.e1));
    HIPCHK(h, hipEventRecord(ev.e0, h->stream));
    for (int i = 0; i < reps; ++i) HIPCHK(h, launch());
    HIPCHK(h, hipEventRecord(ev.e1, h->stream));
    HIPCHK(h, hipEventSynchronize(ev.e1));
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, ev.e0, ev.e1));
    if (kernel == BDK_K_STEP) h->host_step += reps + 1;
    for (Snap &sn : snaps.v) HIPCHK(h, hipMemcpyAsync(sn.src, sn.p, sn.n, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *ms_per_launch = ms / reps;
    return check_nonfinite(h);
}

// ------------------------------------------------------------------ initial states
struct bdk_mt {
    std::mt19937_64 engine;
};

int bdk_mt_create(uint64_t seed, bdk_mt_t **out) {
    if (!out) return BDK_ERR_ARG;
    *out = new bdk_mt{std::mt19937_64(seed)};
    return BDK_OK;
}

void bdk_mt_destroy(bdk_mt_t *g) { delete g; }

// zufall_gleichverteilt_vonbis (zufall.cpp:35-39)
double bdk_mt_uniform(bdk_mt_t *g, double lo, double hi) {
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    return lo + (hi - lo) * dist(g->engine);
}

int bdk_mt_init_positions(bdk_mt_t *g, const bd_params *p, int32_t *git1, double *rel1, int32_t *git2,
                          double *rel2) {
    if (!g || !p) return BDK_ERR_ARG;
    Derived d = derive(*p);
    for (int sp = 0; sp < 2; ++sp) {
        const int n = sp == 0 ? p->n1 : p->n2;
        int32_t *git = sp == 0 ? git1 : git2;
        double *rel = sp == 0 ? rel1 : rel2;
        for (int i = 0; i < n; ++i) {
            double x = bdk_mt_uniform(g, 0.0, p->L);
            double y = bdk_mt_uniform(g, 0.0, p->L);
            int jc = (int)((int)x / d.B);  // dm:1655: (int) binds to x only
            int kc = (int)((int)y / d.B);
            git[2 * i] = jc;
            git[2 * i + 1] = kc;
            rel[2 * i] = x - jc * d.B;
            rel[2 * i + 1] = y - kc * d.B;
        }
    }
    return BDK_OK;
}

int bdk_mt_disc_noise(bdk_mt_t *g, int32_t n1, int32_t n2, double *out1, double *out2) {
    if (!g) return BDK_ERR_ARG;
    const double w5 = std::sqrt(5.0);
    for (int sp = 0; sp < 2; ++sp) {
        const int n = sp == 0 ? n1 : n2;
        double *o = sp == 0 ? out1 : out2;
        for (int i = 0; i < n; ++i) {
            double x, y;
            do {
                x = w5 * bdk_mt_uniform(g, -1.0, 1.0);
                y = w5 * bdk_mt_uniform(g, -1.0, 1.0);
            } while (x * x + y * y > 5.0);
            o[2 * i] = x;
            o[2 * i + 1] = y;
        }
    }
    return BDK_OK;
}

static void to_git_rel(double x, double y, const Derived &d, int32_t *git, double *rel) {
    int gx = (int)std::floor(x / d.B), gy = (int)std::floor(y / d.B);
    gx = std::min(std::max(gx, 0), d.nc - 1);
    gy = std::min(std::max(gy, 0), d.nc - 1);
    git[0] = gx;
    git[1] = gy;
    rel[0] = x - gx * d.B;
    rel[1] = y - gy * d.B;
}

int bdk_init_uniform(const bd_params *p, uint64_t seed, double dmin, int32_t *git, double *rel) {
    if (!p || !git || !rel) return fail(nullptr, BDK_ERR_ARG, "null argument");
    const int N = p->n1 + p->n2;
    const double L = p->L;
    Derived d = derive(*p);
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> U(0.0, L);
    const int ng = std::max(1, (int)(L / std::max(dmin, 1e-9)));
    const double cw = L / ng;
    std::vector<std::vector<int>> grid(dmin > 0 ? (size_t)ng * ng : 0);
    std::vector<double> xs, ys;
    xs.reserve(N);
    ys.reserve(N);
    const int64_t max_tries = 1000LL * N + 100000;
    int64_t tries = 0;
    while ((int)xs.size() < N) {
        if (++tries > max_tries) return fail(nullptr, BDK_ERR_ARG, "random insertion jammed (density too high)");
        double x = U(eng), y = U(eng);
        if (x >= L) x = 0.0;
        if (y >= L) y = 0.0;
        if (dmin > 0) {
            int cx = std::min((int)(x / cw), ng - 1), cy = std::min((int)(y / cw), ng - 1);
            bool ok = true;
            for (int a = -1; a <= 1 && ok; ++a)
                for (int b = -1; b <= 1 && ok; ++b) {
                    int u = ((cx + a) % ng + ng) % ng, v = ((cy + b) % ng + ng) % ng;
                    for (int q : grid[(size_t)u * ng + v]) {
                        double dx = x - xs[q], dy = y - ys[q];
                        dx -= L * std::round(dx / L);
                        dy -= L * std::round(dy / L);
                        if (dx * dx + dy * dy < dmin * dmin) {
                            ok = false;
                            break;
                        }
                    }
                }
            if (!ok) continue;
            grid[(size_t)cx * ng + cy].push_back((int)xs.size());
        }
        xs.push_back(x);
        ys.push_back(y);
    }
    for (int i = 0; i < N; ++i) to_git_rel(xs[i], ys[i], d, git + 2 * i, rel + 2 * i);
    return BDK_OK;
}

int bdk_init_blob(const bd_params *p, uint64_t seed, double width, double background, double dmin, int32_t *git,
                  double *rel, double *clipped_out) {
    if (!p || !git || !rel || !(dmin > 0) || !(width > 0)) return fail(nullptr, BDK_ERR_ARG, "bad argument");
    const int N = p->n1 + p->n2;
    const double L = p->L, c = 0.5 * L;
    Derived d = derive(*p);
    // hexagonal lattice of spacing dmin with whole rows and columns that fit
    // the periodic box (the row spacing is stretched slightly to close it)
    const int nxl = std::max(1, (int)std::floor(L / dmin));
    int nyl = std::max(2, (int)std::floor(L / (dmin * std::sqrt(3.0) / 2.0)));
    if (nyl % 2) --nyl;
    const double ax = L / nxl, ay = L / nyl;
    const double rho_lat = 1.0 / (ax * ay);
    // target density profile: (1-bg) N Gaussian(width) + bg N / L^2
    const double amp = (1.0 - background) * N / (2.0 * M_PI * width * width);
    const double flat = background * N / (L * L);
    std::vector<double> sx, sy, w;
    sx.reserve((size_t)nxl * nyl);
    double wsum = 0.0, clip_lost = 0.0;
    for (int j = 0; j < nyl; ++j)
        for (int i = 0; i < nxl; ++i) {
            double x = (i + 0.5 * (j & 1) + 0.25) * ax, y = (j + 0.5) * ay;
            double dx = x - c, dy = y - c;
            double target = amp * std::exp(-(dx * dx + dy * dy) / (2.0 * width * width)) + flat;
            double prob = target / rho_lat;
            if (prob > 1.0) {
                clip_lost += (prob - 1.0);
                prob = 1.0;
            }
            sx.push_back(x);
            sy.push_back(y);
            w.push_back(prob);
            wsum += prob;
        }
    if ((double)sx.size() < N) return fail(nullptr, BDK_ERR_ARG, "lattice has fewer sites than N");
    // renormalise after the clip: scale the profile by alpha so that the
    // clipped keep-probabilities min(1, alpha w) sum to N (bisection)
    {
        std::vector<double> raw(w.size());
        for (size_t s = 0; s < w.size(); ++s) {
            double x = sx[s] - c, y = sy[s] - c;
            raw[s] = (amp * std::exp(-(x * x + y * y) / (2.0 * width * width)) + flat) / rho_lat;
        }
        auto kept = [&](double a) {
            double t = 0.0;
            for (double r : raw) t += std::min(1.0, a * r);
            return t;
        };
        double lo = 0.0, hi = 1.0;
        while (kept(hi) < N) hi *= 2.0;
        for (int it = 0; it < 100; ++it) {
            double mid = 0.5 * (lo + hi);
            (kept(mid) < N ? lo : hi) = mid;
        }
        for (size_t s = 0; s < w.size(); ++s) w[s] = std::min(1.0, hi * raw[s]);
    }
    // seeded thinning: keep each site independently with probability w
    // (expected count N), then correct the count to exactly N by dropping
    // random uncertain sites or adding sites accepted with probability w
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::vector<char> keep(sx.size(), 0);
    std::vector<int> chosen;
    chosen.reserve((size_t)N + 1024);
    for (size_t s = 0; s < sx.size(); ++s)
        if (U(eng) < w[s]) {
            keep[s] = 1;
            chosen.push_back((int)s);
        }
    std::uniform_int_distribution<size_t> pick(0, sx.size() - 1);
    while ((int)chosen.size() < N) {
        size_t s = pick(eng);
        if (!keep[s] && U(eng) < w[s]) {
            keep[s] = 1;
            chosen.push_back((int)s);
        }
    }
    while ((int)chosen.size() > N) {
        std::uniform_int_distribution<size_t> pc(0, chosen.size() - 1);
        size_t q = pc(eng);
        if (w[chosen[q]] >= 1.0) continue;
        keep[chosen[q]] = 0;
        chosen[q] = chosen.back();
        chosen.pop_back();
    }
    std::sort(chosen.begin(), chosen.end());
    // species labels: shuffle so both species share the profile
    std::shuffle(chosen.begin(), chosen.end(), eng);
    for (int i = 0; i < N; ++i) to_git_rel(sx[chosen[i]], sy[chosen[i]], d, git + 2 * i, rel + 2 * i);
    if (clipped_out) *clipped_out = clip_lost / (wsum + clip_lost);
    return BDK_OK;
}

// The lines "t x y" of particles [lo, hi) of one species (au:633-637) into out.
static void format_pos_lines(std::string &out, int lo, int hi, double B, double t, const int32_t *g,
                             const double *r) {
    out.clear();
    out.reserve((size_t)(hi - lo) * 40);
    char line[128];
    for (int i = lo; i < hi; ++i) {
        const double x = r[2 * (size_t)i] + B * g[2 * (size_t)i];
        const double y = r[2 * (size_t)i + 1] + B * g[2 * (size_t)i + 1];
        const int k = std::snprintf(line, sizeof line, "%g \t %g \t %g \n", t, x, y);
        if (k > 0) out.append(line, (size_t)std::min(k, (int)sizeof line - 1));
    }
}

// One checkpoint of pos1/pos2 (au:57-69, au:631-648): on nr == 0 the file is
// truncated and gets the header, then one line "t x y" per particle of the
// species in id order and a blank line.  g/r: [n][2] of each species.  The
// lines are formatted by up to 16 host threads (BDK_WRITER_THREADS sets the
// limit, 1-64) in contiguous chunks and written in order, so the bytes equal
// one thread's.  Returns an empty string or the error.
static std::string append_pos_files(const char *path1, const char *path2, int n1, int n2, double B, int32_t nr,
                                    double obs_dt, const int32_t *g1, const double *r1, const int32_t *g2,
                                    const double *r2) {
    const double t = nr * obs_dt;
    int hw = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    if (const char *e = std::getenv("BDK_WRITER_THREADS")) hw = std::max(1, std::min(64, std::atoi(e)));
    for (int sp = 0; sp < (n2 ? 2 : 1); ++sp) {
        const char *path = sp == 0 ? path1 : path2;
        if (!path) continue;
        const int n = sp == 0 ? n1 : n2;
        const int32_t *g = sp == 0 ? g1 : g2;
        const double *r = sp == 0 ? r1 : r2;
        const int nt = std::min(hw, 1 + n / 32768);
        std::vector<std::string> part((size_t)nt);
        std::vector<std::thread> th;
        for (int k = 1; k < nt; ++k)
            th.emplace_back(format_pos_lines, std::ref(part[k]), (int)((long)n * k / nt),
                            (int)((long)n * (k + 1) / nt), B, t, g, r);
        format_pos_lines(part[0], 0, (int)((long)n / nt), B, t, g, r);
        for (std::thread &x : th) x.join();
        FILE *f = std::fopen(path, nr == 0 ? "w" : "a");
        if (!f) return std::string("cannot open ") + path;
        bool ok = true;
        if (nr == 0) ok = std::fputs("# Format: t TAB x TAB y \n\n", f) >= 0;  // au:60
        for (const std::string &q : part) ok = ok && std::fwrite(q.data(), 1, q.size(), f) == q.size();
        ok = std::fputs("\n", f) >= 0 && ok;
        if (std::fclose(f) != 0 || !ok) return std::string("cannot write ") + path;
    }
    return std::string();
}

int bdk_write_pos(bdk_t *h, const char *path1, const char *path2, int32_t nr, double obs_dt) {
    if (!h || !path1) return fail(h, BDK_ERR_ARG, "null argument");
    int rc = bdk_write_wait(h);  // an earlier asynchronous write of the same files comes first
    if (rc) return rc;
    const int n1 = h->p.n1, n2 = h->p.n2;
    std::vector<int32_t> g1(2 * (size_t)n1), g2(2 * (size_t)n2);
    std::vector<double> r1(2 * (size_t)n1), r2(2 * (size_t)n2);
    rc = bdk_get_positions(h, 1, n1, g1.data(), r1.data());
    if (rc) return rc;
    if (n2) {
        rc = bdk_get_positions(h, 2, n2, g2.data(), r2.data());
        if (rc) return rc;
    }
    std::string e = append_pos_files(path1, path2, n1, n2, h->g.B, nr, obs_dt, g1.data(), r1.data(), g2.data(),
                                     r2.data());
    return e.empty() ? BDK_OK : fail(h, BDK_ERR_ARG, e);
}

static void free_writer(bdk_t *h) {
    bdkx::PosWriter &w = h->writer;
    if (w.th.joinable()) w.th.join();
    w.pending = false;
    if (w.cell) (void)hipHostFree(w.cell);
    if (w.pid) (void)hipHostFree(w.pid);
    if (w.rel) (void)hipHostFree(w.rel);
    if (w.sc) (void)hipHostFree(w.sc);
    if (w.ev) (void)hipEventDestroy(w.ev);
    if (w.ready) (void)hipEventDestroy(w.ready);
    if (w.st) (void)hipStreamDestroy(w.st);
    w.st = nullptr;
    w.ready = nullptr;
    w.cell = w.pid = nullptr;
    w.rel = nullptr;
    w.sc = nullptr;
    w.ev = nullptr;
    w.cap = 0;
}

int bdk_write_pos_async(bdk_t *h, const char *path1, const char *path2, int32_t nr, double obs_dt) {
    if (!h || !path1) return fail(h, BDK_ERR_ARG, "null argument");
    int rc = bdk_write_wait(h);
    if (rc) return rc;
    rc = ensure(h);
    if (rc) return rc;
    bdkx::PosWriter &w = h->writer;
    const int N = h->g.N;
    if (w.cap < N || !w.ev) {
        free_writer(h);
        HIPCHK(h, hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking));
        HIPCHK(h, hipEventCreateWithFlags(&w.ready, hipEventDisableTiming));
        HIPCHK(h, hipEventCreateWithFlags(&w.ev, hipEventDisableTiming));
        HIPCHK(h, hipHostMalloc((void **)&w.cell, (size_t)std::max(N, 1) * 4, hipHostMallocDefault));
        HIPCHK(h, hipHostMalloc((void **)&w.pid, (size_t)std::max(N, 1) * 4, hipHostMallocDefault));
        HIPCHK(h, hipHostMalloc((void **)&w.rel, (size_t)std::max(N, 1) * 16, hipHostMallocDefault));
        HIPCHK(h, hipHostMalloc((void **)&w.sc, sizeof(DevScalars), hipHostMallocDefault));
        w.cap = N;
    }
    // the copies start when the state is there, and later work on the
    // handle's stream waits for them, so no step overwrites the state first
    HIPCHK(h, hipEventRecord(w.ready, h->stream));
    HIPCHK(h, hipStreamWaitEvent(w.st, w.ready, 0));
    HIPCHK(h, hipMemcpyAsync(w.cell, h->d.cell, (size_t)N * 4, hipMemcpyDeviceToHost, w.st));
    HIPCHK(h, hipMemcpyAsync(w.pid, h->d.pid, (size_t)N * 4, hipMemcpyDeviceToHost, w.st));
    HIPCHK(h, hipMemcpyAsync(w.rel, h->d.rel, (size_t)N * 16, hipMemcpyDeviceToHost, w.st));
    HIPCHK(h, hipMemcpyAsync(w.sc, h->d.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, w.st));
    HIPCHK(h, hipEventRecord(w.ev, w.st));
    HIPCHK(h, hipStreamWaitEvent(h->stream, w.ev, 0));
    w.pending = true;
    w.status = BDK_OK;
    w.err.clear();
    const std::string p1 = path1, p2 = path2 ? path2 : "";
    const bool has2 = path2 != nullptr;
    const int n1 = h->p.n1, n2 = h->p.n2, nc = h->g.nc;
    const double B = h->g.B;
    bdkx::PosWriter *wp = &w;
    w.th = std::thread([=]() {
        hipError_t e = hipEventSynchronize(wp->ev);
        if (e != hipSuccess) {
            wp->status = BDK_ERR_HIP;
            wp->err = std::string("bdk_write_pos_async: ") + hipGetErrorString(e);
            return;
        }
        if (wp->sc->nonfinite & 2) {
            wp->status = BDK_ERR_HIP;
            wp->err = "internal error: inconsistent cell lists (cell counts do not sum to N)";
            return;
        }
        if (wp->sc->nonfinite) {
            wp->status = BDK_ERR_NONFINITE;
            wp->err = "a particle position became non-finite or ran away (|rel| >= 1e9); it was held in place";
            return;
        }
        std::vector<int32_t> g(2 * (size_t)N);
        std::vector<double> r(2 * (size_t)N);
        for (int s = 0; s < N; ++s) {  // slot -> particle id order, species 1 first
            const int q = wp->pid[s], c = wp->cell[s];
            g[2 * (size_t)q] = c / nc;
            g[2 * (size_t)q + 1] = c % nc;
            r[2 * (size_t)q] = wp->rel[2 * (size_t)s];
            r[2 * (size_t)q + 1] = wp->rel[2 * (size_t)s + 1];
        }
        std::string msg = append_pos_files(p1.c_str(), has2 ? p2.c_str() : nullptr, n1, n2, B, nr, obs_dt,
                                           g.data(), r.data(), g.data() + 2 * (size_t)n1,
                                           r.data() + 2 * (size_t)n1);
        if (!msg.empty()) {
            wp->status = BDK_ERR_ARG;
            wp->err = msg;
        }
    });
    return BDK_OK;
}

int bdk_write_wait(bdk_t *h) {
    if (!h) return fail(nullptr, BDK_ERR_ARG, "null handle");
    bdkx::PosWriter &w = h->writer;
    if (w.th.joinable()) w.th.join();
    w.pending = false;
    const int st = w.status;
    w.status = BDK_OK;
    if (st) return fail(h, st, w.err);
    return BDK_OK;
}

// ---------------------------------------------------------------- restart

// init_pos_aus_datei (dm:1576-1643) with P8 fixed: two header lines, then N1
// lines "nr 1 x y" and N2 lines "nr 2 x y" (nr counting from 0 per species),
// x and y read as float like the reference, 0 <= x, y <= L;
// git = (int)(x/B), rel = x - git*B.  Particles species 1 first.
int bdk_read_startpos(const bd_params *p, const char *path, int32_t *git, double *rel) {
    if (!p || !path || !git || !rel) return fail(nullptr, BDK_ERR_ARG, "null argument");
    std::string v = validate(*p);
    if (!v.empty()) return fail(nullptr, BDK_ERR_ARG, v);
    FILE *in = std::fopen(path, "r");
    if (!in) return fail(nullptr, BDK_ERR_ARG, std::string("cannot open ") + path);
    char buf[256];
    if (!std::fgets(buf, sizeof buf, in) || !std::fgets(buf, sizeof buf, in)) {
        std::fclose(in);
        return fail(nullptr, BDK_ERR_ARG, "startpos: cannot read header lines 1-2");
    }
    const double B = derive(*p).B, L = p->L;
    int line = 0;
    for (int sp = 1; sp <= 2; ++sp) {
        const int n = sp == 1 ? p->n1 : p->n2, off = sp == 1 ? 0 : p->n1;
        for (int i = 0; i < n; ++i, ++line) {
            int nr = -1, typ = -1;
            float xf = 0.f, yf = 0.f;
            if (std::fscanf(in, "%d %d %g %g", &nr, &typ, &xf, &yf) != 4 || nr != i || typ != sp) {
                std::fclose(in);
                return fail(nullptr, BDK_ERR_ARG, "startpos: bad particle line " + std::to_string(line + 3));
            }
            const double x = xf, y = yf;
            if (x < 0.0 || x > L || y < 0.0 || y > L) {
                std::fclose(in);
                return fail(nullptr, BDK_ERR_ARG, "startpos: coordinates out of range on line " + std::to_string(line + 3));
            }
            const int gx = (int)(x / B), gy = (int)(y / B);
            git[2 * (size_t)(off + i)] = gx;
            git[2 * (size_t)(off + i) + 1] = gy;
            rel[2 * (size_t)(off + i)] = x - gx * B;
            rel[2 * (size_t)(off + i) + 1] = y - gy * B;
        }
    }
    std::fclose(in);
    return BDK_OK;
}

namespace {
const char kStateMagic[8] = {'B', 'D', 'K', 'S', 'T', 'A', 'T', '1'};
}

// Device snapshot: parameters, step counters and the particle state in list
// order (lcell, lrel, lpid).  Loading it reproduces the list order, so a
// resumed run continues bit for bit.
int bdk_save_state(bdk_t *h, const char *path) {
    int rc = ensure(h);
    if (rc) return rc;
    if (!path) return fail(h, BDK_ERR_ARG, "null path");
    const int N = h->g.N;
    std::vector<int32_t> cell(N), pid(N);
    std::vector<double> rel(2 * (size_t)N);
    DevScalars sc;
    HIPCHK(h, hipMemcpyAsync(cell.data(), h->d.lcell, (size_t)N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(pid.data(), h->d.lpid, (size_t)N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(rel.data(), h->d.lrel, (size_t)N * 16, hipMemcpyDeviceToHost, h->stream));
    std::vector<float> comp(h->d.lcomp ? 2 * (size_t)N : 0);  // Kahan compensation (POS32K rung)
    if (h->d.lcomp)
        HIPCHK(h, hipMemcpyAsync(comp.data(), h->d.lcomp, (size_t)N * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(&sc, h->d.sc, sizeof(sc), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(h, BDK_ERR_ARG, std::string("cannot open ") + path);
    bd_params p = h->p;
    p.device = 0;
    const int64_t hs = h->host_step;
    const uint64_t step = sc.step;
    bool ok = std::fwrite(kStateMagic, 1, 8, f) == 8 && std::fwrite(&p, sizeof p, 1, f) == 1 &&
              std::fwrite(&step, 8, 1, f) == 1 && std::fwrite(&hs, 8, 1, f) == 1 &&
              std::fwrite(cell.data(), 4, N, f) == (size_t)N && std::fwrite(rel.data(), 16, N, f) == (size_t)N &&
              std::fwrite(pid.data(), 4, N, f) == (size_t)N &&
              (comp.empty() || std::fwrite(comp.data(), 8, N, f) == (size_t)N);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? BDK_OK : fail(h, BDK_ERR_ARG, std::string("write failed: ") + path);
}

int bdk_load_state(bdk_t *h, const char *path) {
    if (!h || !path) return fail(h, BDK_ERR_ARG, "null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(h, BDK_ERR_ARG, std::string("cannot open ") + path);
    char magic[8];
    bd_params p;
    uint64_t step = 0;
    int64_t hs = 0;
    const int N = h->g.N;
    std::vector<int32_t> cell(N), pid(N);
    std::vector<double> rel(2 * (size_t)N);
    std::vector<float> comp;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kStateMagic, 8) == 0 &&
              std::fread(&p, sizeof p, 1, f) == 1 && std::fread(&step, 8, 1, f) == 1 && std::fread(&hs, 8, 1, f) == 1;
    if (ok) {
        bd_params mine = h->p;
        mine.device = 0;
        ok = std::memcmp(&p, &mine, sizeof p) == 0;
        if (!ok) {
            std::fclose(f);
            return fail(h, BDK_ERR_ARG, "state file was written with different parameters");
        }
        ok = std::fread(cell.data(), 4, N, f) == (size_t)N && std::fread(rel.data(), 16, N, f) == (size_t)N &&
             std::fread(pid.data(), 4, N, f) == (size_t)N;
        if (ok && h->d.comp) {
            comp.resize(2 * (size_t)N);
            ok = std::fread(comp.data(), 8, N, f) == (size_t)N;
        }
    }
    std::fclose(f);
    if (!ok) return fail(h, BDK_ERR_ARG, std::string("not a readable state file: ") + path);
    // the slots take the saved list order; host copies per species for later uploads
    for (int sp = 0; sp < 2; ++sp) {
        const int n = sp == 0 ? h->p.n1 : h->p.n2;
        h->hgit[sp].assign(2 * (size_t)n, 0);
        h->hrel[sp].assign(2 * (size_t)n, 0.0);
    }
    std::vector<char> seen(N, 0);
    for (int q = 0; q < N; ++q) {
        const int id = pid[q];
        if (id < 0 || id >= N || seen[id] || cell[q] < 0 || cell[q] >= h->g.ncell)
            return fail(h, BDK_ERR_ARG, "corrupt state file");
        seen[id] = 1;
        const int sp = id < h->p.n1 ? 0 : 1, i = sp ? id - h->p.n1 : id;
        h->hgit[sp][2 * (size_t)i] = cell[q] / h->g.nc;
        h->hgit[sp][2 * (size_t)i + 1] = cell[q] % h->g.nc;
        h->hrel[sp][2 * (size_t)i] = rel[2 * (size_t)q];
        h->hrel[sp][2 * (size_t)i + 1] = rel[2 * (size_t)q + 1];
    }
    drop_graph(h);
    HIPCHK(h, hipMemcpyAsync(h->d.cell, cell.data(), (size_t)N * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.pid, pid.data(), (size_t)N * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.rel, rel.data(), (size_t)N * 16, hipMemcpyHostToDevice, h->stream));
    if (h->d.comp)  // slot order = the saved list order, like the state
        HIPCHK(h, hipMemcpyAsync(h->d.comp, comp.data(), (size_t)N * 8, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->d.sc, &step, 8, hipMemcpyHostToDevice, h->stream));  // DevScalars::step
    HIPCHK(h, hipMemsetAsync(h->d.count, 0, (size_t)h->g.ncell * 4, h->stream));
    HIPCHK(h, clear_flags(h));
    HIPCHK(h, bdkx::launch_bin(h));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->pos_set[0] = h->pos_set[1] = true;
    h->dev_valid = true;
    h->forces_ready = false;
    h->host_step = hs;
    return BDK_OK;
}

// ---------------------------------------------------------------- observables

int bdk_radial_profile(bdk_t *h, int32_t nbins, double dr, double *centers, double *profiles) {
    int rc = ensure(h);
    if (rc) return rc;
    if (nbins < 1 || !(dr > 0.0) || !profiles) return fail(h, BDK_ERR_ARG, "need nbins >= 1, dr > 0, profiles");
    const int nb = (h->g.N + 255) / 256;
    double *part = nullptr, *cen = nullptr;
    int *hist = nullptr;
    hipError_t e = hipMalloc((void **)&part, (size_t)nb * 8 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&cen, 6 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&hist, (size_t)6 * nbins * sizeof(int));
    if (e == hipSuccess) e = bdkx::launch_radial(h, nbins, dr, part, cen, hist);
    std::vector<int> cnt((size_t)6 * nbins);
    double c6[6] = {0, 0, 0, 0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), hist, cnt.size() * 4, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c6, cen, sizeof(c6), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (part) (void)hipFree(part);
    if (cen) (void)hipFree(cen);
    if (hist) (void)hipFree(hist);
    if (e != hipSuccess) return hipfail(h, e, "bdk_radial_profile");
    const double L = h->p.L;
    const double v1 = h->p.n1 ? L * L / (h->p.n1 * dr * dr * M_PI) : 0.0;  // observablen_auswertung.cpp:268
    const double v2 = h->p.n2 ? L * L / (h->p.n2 * dr * dr * M_PI) : 0.0;
    const double vf[6] = {v1, v2, v1, v1, v2, v2};
    for (int k = 0; k < 6; ++k)
        for (int b = 0; b < nbins; ++b)
            profiles[(size_t)k * nbins + b] = cnt[(size_t)k * nbins + b] * (vf[k] / (2 * b + 1));
    if (centers) std::memcpy(centers, c6, sizeof(c6));
    return BDK_OK;
}

int bdk_pair_correlation(bdk_t *h, int32_t nbins, double dr, double *g11, double *g12, double *g22) {
    int rc = ensure(h);
    if (rc) return rc;
    if (nbins < 1 || nbins > 4096 || !(dr > 0.0)) return fail(h, BDK_ERR_ARG, "need 1 <= nbins <= 4096, dr > 0");
    unsigned long long *hist = nullptr;
    const size_t hb = (size_t)3 * nbins * sizeof(unsigned long long);
    hipError_t e = hipMalloc((void **)&hist, hb);
    if (e == hipSuccess) e = hipMemsetAsync(hist, 0, hb, h->stream);
    if (e == hipSuccess) e = bdkx::launch_pair_hist(h, nbins, dr, hist);
    std::vector<unsigned long long> cnt((size_t)3 * nbins);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), hist, hb, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (hist) (void)hipFree(hist);
    if (e != hipSuccess) return hipfail(h, e, "bdk_pair_correlation");
    // auswertung.cpp:120-121, 149-150, 173-174 (g12 keeps the reference's N1-1)
    const double L = h->p.L, n1 = h->p.n1, n2 = h->p.n2;
    const double v11 = n1 > 1 ? 2.0 / ((n1 - 1) * M_PI * (n1 / L / L) * dr * dr) : 0.0;
    const double v12 = (n1 > 1 && n2 > 0) ? 1.0 / ((n1 - 1) * M_PI * (n2 / L / L) * dr * dr) : 0.0;
    const double v22 = n2 > 1 ? 2.0 / ((n2 - 1) * M_PI * (n2 / L / L) * dr * dr) : 0.0;
    for (int b = 0; b < nbins; ++b) {
        // same-species pairs were counted in both orders
        if (g11) g11[b] = (double)(cnt[b] / 2) * (v11 / (2 * b + 1));
        if (g12) g12[b] = (double)cnt[(size_t)nbins + b] * (v12 / (2 * b + 1));
        if (g22) g22[b] = (double)(cnt[2 * (size_t)nbins + b] / 2) * (v22 / (2 * b + 1));
    }
    return BDK_OK;
}

// Snapshot density grid and its statistics (main_auswertung.cpp:214-385).
static int density_snapshot(bdk_t *h, int32_t m, double *rho1, double *rho2, int32_t nbins, double dbin,
                            const double *rho_thresh, const double *drho_thresh, double *hist, double *over,
                            const char *what) {
    int rc = ensure(h);
    if (rc) return rc;
    if (m < 1 || m > 16384) return fail(h, BDK_ERR_ARG, std::string(what) + ": need 1 <= m <= 16384");
    const int stats = hist || over ? 1 : 0;
    if (stats && (nbins < 1 || nbins > 4096 || !(dbin > 0.0) || !rho_thresh || !drho_thresh))
        return fail(h, BDK_ERR_ARG, std::string(what) + ": need 1 <= nbins <= 4096, dbin > 0, thresholds");
    const size_t mm = (size_t)m * m;
    bdkx::SnapThresh th{};
    if (stats) {
        for (int t = 0; t < 3; ++t) th.rho[t] = rho_thresh[t];
        for (int t = 0; t < 4; ++t) th.drho[t] = drho_thresh[t];
    }
    unsigned long long *acc = nullptr;
    double *rho = nullptr;
    unsigned int *cnt = nullptr;
    const size_t nc = stats ? (size_t)4 * nbins + 7 : 1;
    hipError_t e = hipMalloc((void **)&acc, 2 * mm * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void **)&rho, 2 * mm * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&cnt, nc * sizeof(unsigned int));
    if (e == hipSuccess) e = hipMemsetAsync(acc, 0, 2 * mm * sizeof(unsigned long long), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(cnt, 0, nc * sizeof(unsigned int), h->stream);
    if (e == hipSuccess) e = bdkx::launch_snap(h, m, acc, rho, stats, nbins, dbin, th, cnt);
    std::vector<unsigned int> c(nc);
    if (e == hipSuccess && rho1) e = hipMemcpyAsync(rho1, rho, mm * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && rho2) e = hipMemcpyAsync(rho2, rho + mm, mm * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && stats)
        e = hipMemcpyAsync(c.data(), cnt, nc * sizeof(unsigned int), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (acc) (void)hipFree(acc);
    if (rho) (void)hipFree(rho);
    if (cnt) (void)hipFree(cnt);
    if (e != hipSuccess) return hipfail(h, e, what);
    if (stats) {
        const double pre = 1.0 / (double)mm;  // main_auswertung.cpp:328
        if (hist)
            for (size_t i = 0; i < (size_t)4 * nbins; ++i) hist[i] = pre * c[i];
        if (over)
            for (int t = 0; t < 7; ++t) over[t] = pre * c[(size_t)4 * nbins + t];
    }
    return BDK_OK;
}

int bdk_density_grid(bdk_t *h, int32_t m, double *rho1, double *rho2) {
    if (!h || (!rho1 && !rho2)) return fail(h, BDK_ERR_ARG, "null argument");
    return density_snapshot(h, m, rho1, rho2, 0, 0.0, nullptr, nullptr, nullptr, nullptr, "bdk_density_grid");
}

int bdk_density_stats(bdk_t *h, int32_t m, int32_t nbins, double dbin, const double *rho_thresh,
                      const double *drho_thresh, double *hist, double *over) {
    if (!h || !hist || !over) return fail(h, BDK_ERR_ARG, "null argument");
    return density_snapshot(h, m, nullptr, nullptr, nbins, dbin, rho_thresh, drho_thresh, hist, over,
                            "bdk_density_stats");
}

int bdk_ftrho_sums(bdk_t *h, int32_t species, int32_t qbins, double dq_ft, double *re, double *im, int32_t *npairs) {
    int rc = ensure(h);
    if (rc) return rc;
    if ((species != 1 && species != 2) || qbins < 1 || qbins > 256 || !re || !im || !npairs)
        return fail(h, BDK_ERR_ARG, "bdk_ftrho_sums: need species 1 or 2, 1 <= qbins <= 256, outputs");
    // the reference's loop order: the inner loop stops at the first j beyond the quarter circle
    std::vector<int2> ij;
    for (int i = 0; i < qbins; ++i)
        for (int j = 0; j < qbins; ++j) {
            if (i * i + j * j > qbins * qbins) break;
            ij.push_back(make_int2(i, j));
        }
    const int np = (int)ij.size();
    const size_t nch = (size_t)bdkx::ftrho_chunks(h);
    int2 *dij = nullptr;
    double2 *part = nullptr;
    double *out = nullptr;
    hipError_t e = hipMalloc((void **)&dij, (size_t)np * sizeof(int2));
    if (e == hipSuccess) e = hipMalloc((void **)&part, (size_t)np * nch * sizeof(double2));
    if (e == hipSuccess) e = hipMalloc((void **)&out, 2 * (size_t)np * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(dij, ij.data(), (size_t)np * sizeof(int2), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = bdkx::launch_ftrho(h, dij, np, species, dq_ft, part, out, out + np);
    if (e == hipSuccess) e = hipMemcpyAsync(re, out, (size_t)np * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(im, out + np, (size_t)np * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (dij) (void)hipFree(dij);
    if (part) (void)hipFree(part);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) return hipfail(h, e, "bdk_ftrho_sums");
    *npairs = np;
    return BDK_OK;
}

int bdk_rho_k(bdk_t *h, int32_t species, double dq_ratio, int32_t nbins, double *re, double *im) {
    int rc = ensure(h);
    if (rc) return rc;
    if ((species != 1 && species != 2) || !(dq_ratio > 0.0) || nbins < 1 || !re)
        return fail(h, BDK_ERR_ARG, "need species 1 or 2, dq_ratio > 0, nbins >= 1, re");
    const int Z = h->g.Z;
    double *part = nullptr, *out = nullptr;
    hipError_t e = hipMalloc((void **)&part, (size_t)(Z / 2 + 1) * nbins * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void **)&out, (size_t)nbins * sizeof(double));
    if (e == hipSuccess) e = bdkx::launch_rhok(h, species, dq_ratio, nbins, part, out);
    std::vector<double> sum(nbins);
    if (e == hipSuccess) e = hipMemcpyAsync(sum.data(), out, nbins * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (part) (void)hipFree(part);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) return hipfail(h, e, "bdk_rho_k");
    // lattice points per shell over the full plane (auswertung.cpp:497-510)
    std::vector<int64_t> cnt(nbins, 0);
    for (int i = 0; i < Z; ++i) {
        const double xq = ((i + Z / 2) % Z) - Z / 2;
        for (int j = 0; j < Z; ++j) {
            const double yq = ((j + Z / 2) % Z) - Z / 2;
            const int bin = (int)(dq_ratio * std::sqrt(xq * xq + yq * yq) + 0.5);
            if (bin >= 0 && bin < nbins) ++cnt[bin];
        }
    }
    const double L = h->p.L;
    for (int b = 0; b < nbins; ++b) {  // auswertung.cpp:572-576
        re[b] = cnt[b] ? sum[b] * (L * L / ((double)Z * Z)) / cnt[b] : 0.0;
        if (im) im[b] = 0.0;
    }
    return BDK_OK;
}

// ---------------------------------------------------------------- batches

int bdk_batch_create(bdk_t *const *members, int32_t n, bdk_batch_t **out) {
    if (!out || !members || n < 1) return fail(nullptr, BDK_ERR_ARG, "bdk_batch_create: need n >= 1 members");
    *out = nullptr;
    std::vector<bdk_t *> mem(members, members + n);
    int zmax = 8;
    for (int i = 0; i < n; ++i) {
        bdk_t *h = mem[i];
        if (!h) return fail(nullptr, BDK_ERR_ARG, "bdk_batch_create: null member");
        const std::string who = "member " + std::to_string(i) + ": ";
        if (h->dev != mem[0]->dev) return fail(nullptr, BDK_ERR_ARG, who + "members must share the device");
        if (h->p.precision != mem[0]->p.precision)
            return fail(nullptr, BDK_ERR_ARG, who + "members must share the precision");
        if (h->g.idgas) return fail(nullptr, BDK_ERR_ARG, who + "batches need capillary forces (idgas_only = 0)");
        if (h->p.noise_mode == BDK_NOISE_UPLOAD)
            return fail(nullptr, BDK_ERR_ARG, who + "batches do not take uploaded noise");
        for (int j = 0; j < i; ++j)
            if (mem[j] == h) return fail(nullptr, BDK_ERR_ARG, who + "a handle appears twice");
        int rc = ensure(h);
        if (rc) return fail(nullptr, rc, who + h->err);
        if (hipStreamSynchronize(h->stream) != hipSuccess)
            return fail(nullptr, BDK_ERR_HIP, who + "stream synchronisation failed");
        zmax = std::max(zmax, h->g.Z);
    }
    bdk_batch_t *b = new bdk_batch();
    b->mem = mem;
    b->dev = mem[0]->dev;
    int rc = BDK_OK;
    hipError_t e = hipSetDevice(b->dev);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->aux, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming);
    bdkx::BatchTuning t;
    for (int k = 0; k < 3; ++k) t.ept[k] = bdkx::batch_min_ept(k, zmax);
    if (e == hipSuccess) e = b->rebuild(t);
    if (e != hipSuccess) {
        rc = fail(nullptr, BDK_ERR_HIP, std::string("bdk_batch_create: ") + hipGetErrorString(e));
        bdk_batch_destroy(b);
        return rc;
    }
    *out = b;
    return BDK_OK;
}

void bdk_batch_destroy(bdk_batch_t *b) {
    if (!b) return;
    (void)hipSetDevice(b->dev);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->aux) (void)hipStreamSynchronize(b->aux);
    b->drop_graphs();
    b->free_tables();
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->aux) (void)hipStreamDestroy(b->aux);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *bdk_batch_last_error(const bdk_batch_t *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int bdk_batch_run_steps(bdk_batch_t *b, int64_t k) {
    if (!b) return fail(nullptr, BDK_ERR_ARG, "null batch");
    if (k < 0) return b->fail(BDK_ERR_ARG, "k must be >= 0");
    hipError_t e = hipSetDevice(b->dev);
    if (e == hipSuccess) e = b->ensure_forces();
    if (e == hipSuccess) e = b->ensure_graphs();
    if (e == hipSuccess) e = b->wait_writers();
    int64_t i = 0;
    for (; e == hipSuccess && i + kGraphSteps <= k; i += kGraphSteps) e = hipGraphLaunch(b->gexec_multi, b->stream);
    for (; e == hipSuccess && i < k; ++i) e = hipGraphLaunch(b->gexec, b->stream);
    if (e != hipSuccess) return b->fail(BDK_ERR_HIP, std::string("bdk_batch_run_steps: ") + hipGetErrorString(e));
    for (bdk_t *h : b->mem) h->host_step += k;
    return BDK_OK;
}

int bdk_batch_sync(bdk_batch_t *b) {
    if (!b) return fail(nullptr, BDK_ERR_ARG, "null batch");
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return b->fail(BDK_ERR_HIP, std::string("bdk_batch_sync: ") + hipGetErrorString(e));
    for (size_t i = 0; i < b->mem.size(); ++i) {
        DevScalars sc;
        e = hipMemcpy(&sc, b->mem[i]->d.sc, sizeof(sc), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return b->fail(BDK_ERR_HIP, std::string("bdk_batch_sync: ") + hipGetErrorString(e));
        std::string m;
        const int rc = flag_status(sc.nonfinite, &m);
        if (rc) return b->fail(rc, "member " + std::to_string(i) + ": " + m);
    }
    return BDK_OK;
}

int bdk_batch_status(bdk_batch_t *b, int32_t n, int32_t *status) {
    if (!b) return fail(nullptr, BDK_ERR_ARG, "null batch");
    if (!status || n != (int32_t)b->mem.size()) return b->fail(BDK_ERR_ARG, "status needs one entry per member");
    hipError_t e = hipStreamSynchronize(b->stream);
    for (int32_t i = 0; i < n && e == hipSuccess; ++i) {
        DevScalars sc;
        e = hipMemcpy(&sc, b->mem[i]->d.sc, sizeof(sc), hipMemcpyDeviceToHost);
        if (e == hipSuccess) status[i] = flag_status(sc.nonfinite, nullptr);
    }
    if (e != hipSuccess) return b->fail(BDK_ERR_HIP, std::string("bdk_batch_status: ") + hipGetErrorString(e));
    return BDK_OK;
}

int bdk_batch_profile(bdk_batch_t *b, int32_t stage, int32_t reps, double *ms_per_launch) {
    if (!b) return fail(nullptr, BDK_ERR_ARG, "null batch");
    // the update and the list rebuild only keep the state consistent as parts
    // of a whole step, so they are not timed alone
    if (stage < 0 || stage > 7 || stage == 5 || stage == 6 || reps < 1 || !ms_per_launch)
        return b->fail(BDK_ERR_ARG, "bad profile request (stages 0-4 and 7)");
    hipError_t e = hipSetDevice(b->dev);
    if (e == hipSuccess) e = b->ensure_forces();
    if (e == hipSuccess) e = b->ensure_graphs();
    auto launch = [&]() -> hipError_t {
        if (stage == 7) {
            hipError_t w = b->wait_writers();
            return w != hipSuccess ? w : hipGraphLaunch(b->gexec, b->stream);
        }
        return bdkx::batch_stage(b->bp, stage, b->stream);
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
    for (int i = 0; e == hipSuccess && i < reps; ++i) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return b->fail(BDK_ERR_HIP, std::string("bdk_batch_profile: ") + hipGetErrorString(e));
    // whole steps advance the members (the warm-up launch included)
    if (stage == 7)
        for (bdk_t *h : b->mem) h->host_step += reps + 1;
    *ms_per_launch = ms / reps;
    return BDK_OK;
}

int bdk_batch_set_tuning(bdk_batch_t *b, const int32_t *choice) {
    if (!b) return fail(nullptr, BDK_ERR_ARG, "null batch");
    if (!choice) return b->fail(BDK_ERR_ARG, "null choice");
    bdkx::BatchTuning t;
    for (int k = 0; k < 3; ++k) {
        const int v = choice[BDK_TUNE_EPT_R2C + k];
        if (v != 8 && v != 16 && v != 32) return b->fail(BDK_ERR_ARG, "FFT elements per thread must be 8, 16 or 32");
        t.ept[k] = v;
    }
    t.wca_block = choice[BDK_TUNE_WCA_BLOCK];
    if (t.wca_block != 128 && t.wca_block != 256 && t.wca_block != 512)
        return b->fail(BDK_ERR_ARG, "WCA block size must be 128, 256 or 512");
    t.dep_tile = choice[BDK_TUNE_DEPOSIT_TILE];
    if (t.dep_tile < 8 || t.dep_tile > 96) return b->fail(BDK_ERR_ARG, "deposit tile must be 8..96 mesh points");
    t.persist = choice[BDK_TUNE_PERSIST];
    if (t.persist < 0) return b->fail(BDK_ERR_ARG, "persistent grid must be >= 0");
    hipError_t e = hipSetDevice(b->dev);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess) e = b->rebuild(t);
    if (e != hipSuccess) return b->fail(BDK_ERR_HIP, std::string("bdk_batch_set_tuning: ") + hipGetErrorString(e));
    return BDK_OK;
}

// The project's autotuner.  Each knob is timed on the batch's own members
// with the stages that leave the particles unchanged, one knob at a time in
// this order, keeping the fastest value: the FFT elements per thread of each
// pass (8, 16, 32; stages 2-4), the WCA block size (128, 256, 512; stage 0),
// the deposit tile (mesh points per tile side 24, 32, 38, 48, 64, 80; stage 1:
// fine meshes put few particles in a small tile), and
// the persistent grid (one block per map entry, or 4 or 8 blocks per compute
// unit for every stage; stages 0-4 together).  No choice changes a result.
int bdk_batch_tune(bdk_batch_t *b, int32_t reps, int32_t *out) {
    if (!b) return fail(nullptr, BDK_ERR_ARG, "null batch");
    if (reps < 1) return b->fail(BDK_ERR_ARG, "reps must be >= 1");
    hipError_t e = hipSetDevice(b->dev);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->dev);
    bdkx::BatchTuning best = b->tuning;
    auto timed = [&](const bdkx::BatchTuning &t, std::initializer_list<int> stages, double *ms) -> int {
        e = b->rebuild(t);
        if (e != hipSuccess) return BDK_ERR_HIP;
        *ms = 0.0;
        for (int st : stages) {
            double one = 0.0;
            const int rc = bdk_batch_profile(b, st, reps, &one);
            if (rc) return rc;
            *ms += one;
        }
        return BDK_OK;
    };
    auto pick = [&](auto set, const std::vector<int> &values, std::initializer_list<int> stages) -> int {
        double tbest = 1e30;
        bdkx::BatchTuning keep = best;
        for (int v : values) {
            bdkx::BatchTuning t = best;
            if (!set(t, v)) continue;
            double ms = 0.0;
            const int rc = timed(t, stages, &ms);
            if (rc) return rc;
            if (ms < tbest) {
                tbest = ms;
                keep = t;
            }
        }
        best = keep;
        return BDK_OK;
    };
    int rc = e == hipSuccess ? BDK_OK : BDK_ERR_HIP;
    for (int stage = 0; stage < 3 && !rc; ++stage)
        rc = pick([&](bdkx::BatchTuning &t, int v) { return v >= b->ept_min[stage] ? (t.ept[stage] = v, true) : false; },
                  {8, 16, 32}, {2 + stage});
    if (!rc) rc = pick([](bdkx::BatchTuning &t, int v) { return (t.wca_block = v, true); }, {128, 256, 512}, {0});
    if (!rc) rc = pick([](bdkx::BatchTuning &t, int v) { return (t.dep_tile = v, true); }, {24, 32, 38, 48, 64, 80}, {1});
    if (!rc)
        rc = pick([](bdkx::BatchTuning &t, int v) { return (t.persist = v, true); }, {0, 4 * cus, 8 * cus},
                  {0, 1, 2, 3, 4});
    if (rc == BDK_ERR_HIP || e != hipSuccess)
        return b->fail(BDK_ERR_HIP, std::string("bdk_batch_tune: ") + hipGetErrorString(e));
    if (rc) return rc;
    e = b->rebuild(best);
    if (e != hipSuccess) return b->fail(BDK_ERR_HIP, std::string("bdk_batch_tune: ") + hipGetErrorString(e));
    // the timed stages overwrote the members' forces (with equal values):
    // recompute them with the chosen plan before the next step
    for (bdk_t *h : b->mem) h->forces_ready = false;
    if (out) {
        for (int k = 0; k < 3; ++k) out[BDK_TUNE_EPT_R2C + k] = b->bp.ept[k];
        out[BDK_TUNE_WCA_BLOCK] = best.wca_block;
        out[BDK_TUNE_DEPOSIT_TILE] = best.dep_tile;
        out[BDK_TUNE_PERSIST] = best.persist;
    }
    return BDK_OK;
}

}  // extern "C"
