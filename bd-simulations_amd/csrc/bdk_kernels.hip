// HIP kernels of the MI355X Brownian-dynamics timestep (gfx950).
//
// One step (dynamik_methoden.cpp:74-163) starts from a state whose forces
// are on the device and leaves the next state with its forces:
//   k_update    gather (gr:81-668), noise (dm:1317-1339), Euler step
//               (dm:116-157), refresh (dm:216-325); the cell counts are
//               updated by the particles that change cell;
// then on two streams side by side:
//   k_rank_scan rank of each particle that stayed in its cell (no atomics)
//               beside the exclusive scan of the cell counts;
//   k_fill      cell list = slots grouped by cell, species in bit 31, and
//               the particle state in list order (lcell, lrel, lpid, lrelf);
//   k_sort_cells  the arrivals of each cell in slot order;
//   k_wca       WCA forces (dm:617-1059) of the new lists;
// and
//   k_deposit   charge deposit (gr:31-547) per tile of cells into the tile's
//               own mesh points, integer LDS accumulation, here from the
//               slot arrays (grouped by the previous cells);
//   k_r2c_rows  row FFTs of the real mesh, transposed half-spectrum out;
//   k_col       column FFT, Green's function x gradient (dm:1210-1216), two
//               inverse column FFTs (Fx, Fy);
//   k_c2r_rows  inverse row FFT of (Fx + i Fy) -> real Fx, Fy mesh
//               (Stockham passes in LDS: radix 8, radix 16 for the column
//               and c2r passes of FP32 meshes of 1024 points and more, whose
//               r2c rows end in one radix-16 pass instead of radix 8 and 2).
// Every force, gather and update kernel runs over list positions q and
// reads the list-ordered state; k_update writes the new state to slot q.
// The slots are therefore re-sorted by cell every step: consecutive threads
// handle particles of neighbouring cells for the whole run.
// With a finite max_travel, k_gather (forces and the dt candidates,
// dm:1420-1491) runs before a non-fused k_update, which reduces them.
//
// Position arithmetic (x = git*B + rel, mesh indices, weights, refresh) is
// compiled without FP contraction so it rounds like the reference's x86-64
// build; cell indices are therefore bit-exact.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "bdk_internal.hpp"

namespace bdkx {

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };
template <typename R> struct V4;
template <> struct V4<float> { using T = float4; };
template <> struct V4<double> { using T = double4; };

// Complex helpers of the FFT and spectral arithmetic.  The complex product
// uses explicit fused multiply-adds (the file compiles without automatic
// contraction, so every kernel that inlines it rounds the same way).
template <typename C> __device__ __forceinline__ C cm(decltype(C::x) a, decltype(C::x) b) {
    C r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cm<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cm<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cconj(C a) { return cm<C>(a.x, -a.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
    return cm<C>(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
template <typename C> __device__ __forceinline__ C cmul_i(C a) { return cm<C>(-a.y, a.x); }    // i*a
template <typename C> __device__ __forceinline__ C cmul_mi(C a) { return cm<C>(a.y, -a.x); }   // -i*a
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cm<C>(a.x * s, a.y * s); }

__device__ __forceinline__ double dbits(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long bitsd(double v) { return (unsigned long long)__double_as_longlong(v); }

// ------------------------------------------------------------------ Philox
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32x32 -> 64-bit product each (a single multiply instruction
        // instead of separate low and high halves)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
}

// noise of particle `pid` at step `step`: disc of radius sqrt 5 (variance
// 5/4 per component, the distribution of dm:1321-1325) or Gaussian.
__device__ __forceinline__ double2 philox_noise(unsigned long long seed, uint32_t pid, unsigned long long step,
                                                int gauss, double var) {
    uint32_t c[4] = {pid, (uint32_t)step, (uint32_t)(step >> 32), 0x4244u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double inv53 = 1.0 / 9007199254740992.0;
    double u1 = (double)(((((unsigned long long)c[1]) << 32) | c[0]) >> 11) * inv53;
    double u2 = (double)(((((unsigned long long)c[3]) << 32) | c[2]) >> 11) * inv53;
    double r = gauss ? sqrt(-2.0 * var * log(1.0 - u1)) : sqrt(5.0 * u1);
    double a = 2.0 * M_PI * u2;
    double sa, ca;
    sincos(a, &sa, &ca);
    double2 o;
    o.x = r * ca;
    o.y = r * sa;
    return o;
}

// The same draws in FP32 (FP32 force mode: the random force is a force):
// u1, u2 are the high words of the FP64 path's 64-bit draws rounded to FP32
// (relative error <= 2^-24 + 2^-32/u), r and the angle in FP32.
__device__ __forceinline__ double2 philox_noise_f32(unsigned long long seed, uint32_t pid, unsigned long long step,
                                                    int gauss, double var) {
    uint32_t c[4] = {pid, (uint32_t)step, (uint32_t)(step >> 32), 0x4244u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float u1 = fminf((float)c[1] * 2.3283064365386963e-10f, 0.99999994f);  // 2^-32, below 1
    const float u2 = (float)c[3] * 2.3283064365386963e-10f;
    const float r = gauss ? sqrtf(-2.0f * (float)var * logf(1.0f - u1)) : sqrtf(5.0f * u1);
    float sa, ca;
    sincospif(2.0f * u2, &sa, &ca);
    double2 o;
    o.x = (double)(r * ca);
    o.y = (double)(r * sa);
    return o;
}

// ------------------------------------------------------------------ launch order
// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Renumbered so that every XCD runs one contiguous range of the logical
// grid, neighbouring tiles (whose inputs overlap) share an L2.
__device__ __forceinline__ int xcd_block(int on) {
    const int b = blockIdx.x + blockIdx.y * gridDim.x;
    if (!on) return b;
    const int nb = gridDim.x * gridDim.y;
    const int x = b & 7, i = b >> 3, base = nb >> 3, rem = nb & 7;
    return x < rem ? x * (base + 1) + i : rem * (base + 1) + (x - rem) * base + i;
}

// ------------------------------------------------------------------ geometry
__device__ __forceinline__ int wrapc(int u, int n) {
    int v = u % n;
    return v < 0 ? v + n : v;
}
// u in [-n, 2n)
__device__ __forceinline__ int wrap1(int u, int n) { return u < 0 ? u + n : (u >= n ? u - n : u); }
// Mesh code is compiled for two kinds of mesh size: P2 (Z a power of two:
// masks and shifts) and the other even sizes.  The kind is a template
// parameter of every mesh kernel, chosen by its launcher, so the
// power-of-two instruction stream carries no test of the other.
__device__ __forceinline__ bool zpow2(int Z) { return (Z & (Z - 1)) == 0; }
// mesh index u in [-Z, 2Z) into [0, Z)
template <bool P2> __device__ __forceinline__ int wrapz(int u, int Z) { return P2 ? (u & (Z - 1)) : wrap1(u, Z); }
// x / d and x % d for 0 <= x < 2^22: a shift when d is a power of two, else
// the float reciprocal corrected to the exact quotient
struct Dv {
    int d, sh;
    float inv;
    __device__ explicit Dv(int d_)
        : d(d_), sh((d_ & (d_ - 1)) == 0 ? __ffs(d_) - 1 : -1), inv(__builtin_amdgcn_rcpf((float)d_)) {}
    __device__ int q(int x) const {
        if (sh >= 0) return x >> sh;
        int v = (int)((float)x * inv);
        const int r = x - v * d;
        v += (r >= d) - (r < 0);
        return v;
    }
    __device__ int r(int x) const { return sh >= 0 ? (x & (d - 1)) : x - q(x) * d; }
};
// the same with the kind of d known at compile time (P2: a power of two)
template <bool P2> struct DvT : Dv {
    __device__ explicit DvT(int d_) : Dv(d_) {}
};
template <> struct DvT<true> {
    int d, sh;
    __device__ explicit DvT(int d_) : d(d_), sh(__ffs(d_) - 1) {}
    __device__ int q(int x) const { return x >> sh; }
    __device__ int r(int x) const { return x & (d - 1); }
};
// c / nc for 0 <= c < 2^24 (float reciprocal, corrected to the exact quotient)
__device__ __forceinline__ int div_nc(int c, const Geo &g) {
    int q = (int)((float)c * g.ncinv);
    const int r = c - q * g.nc;
    q += (r >= g.nc) - (r < 0);
    return q;
}

// mesh index and weights along one axis (NGP gr:44-54, CIC gr:154-190,
// TSC gr:435-459), evaluated in FP64 in the reference's operation order
template <int SCHEME>
__device__ __forceinline__ void weights1d(double x, double h, int &k, double w[3]) {
    int kk = (int)(x / h);
    k = kk;
    if (SCHEME == 0) {
        w[0] = 0.0;
        w[1] = 1.0;
        w[2] = 0.0;
    } else if (SCHEME == 1) {
        double tmp = x / h - (kk + 0.5);
        w[1] = 1.0 - fabs(tmp);
        if (tmp < 0.0) {
            w[0] = -tmp;
            w[2] = 0.0;
        } else {
            w[0] = 0.0;
            w[2] = tmp;
        }
    } else {
        double d = x - (kk + 0.5) * h;
        w[1] = 0.75 - d * d / h / h;
        double tmp = fabs((d + h) / h);
        w[0] = 0.5 * (1.5 - tmp) * (1.5 - tmp);
        tmp = fabs((d - h) / h);
        w[2] = 0.5 * (1.5 - tmp) * (1.5 - tmp);
    }
}

// Weights in the precision of the forces: FP64 follows the reference's
// operations exactly; FP32 takes the index from x/h in FP64 and evaluates
// the offset t = x/h - k - 1/2 and the textbook weights (CIC: 1-|t|, -t, t;
// TSC: 3/4-t^2, (1/2-t)^2/2, (1/2+t)^2/2) in FP32.
template <typename R, int SCHEME>
__device__ __forceinline__ void weights_r(double x, const Geo &g, int &k, R w[3]) {
    if (std::is_same<R, double>::value) {
        double wd[3];
        weights1d<SCHEME>(x, g.h, k, wd);
        w[0] = (R)wd[0];
        w[1] = (R)wd[1];
        w[2] = (R)wd[2];
        return;
    }
    if (SCHEME == 0) {
        k = (int)(x / g.h);
        w[0] = 0;
        w[1] = 1;
        w[2] = 0;
        return;
    }
    // CIC/TSC weights are continuous across a cell edge, so x * (1/h) may
    // pick the neighbouring index at an exact edge without changing the sum
    const double u = x * g.hinv;
    const int kk = (int)u;
    k = kk;
    const float t = (float)(u - (kk + 0.5));
    if (SCHEME == 1) {
        w[1] = 1.0f - fabsf(t);
        w[0] = t < 0.0f ? -t : 0.0f;
        w[2] = t < 0.0f ? 0.0f : t;
    } else {
        const float a = 0.5f - t, b = 0.5f + t;
        w[1] = 0.75f - t * t;
        w[0] = 0.5f * a * a;
        w[2] = 0.5f * b * b;
    }
}

// One WCA pair, the arithmetic of dm:640-698 (type 0, pair 11), dm:815-868
// (type 1, pair 22) and dm:987-1040 (type 2, pair 12): force on particle a
// at (xa,ya) from b at (xb,yb).  Minimum image as abstand2 (am:100-166) and
// the shift logic of dm:663-676.  Returns false outside the cutoff.
template <typename R>
__device__ __forceinline__ bool wca_pair(double xa, double ya, double xb, double yb, int type, const Geo &g,
                                         R &fx, R &fy) {
    const double L = g.L;
    double t1 = xa - xb, t2 = xa - xb + L, t3 = xa - xb - L;
    double m23 = (t3 * t3 < t2 * t2) ? t3 * t3 : t2 * t2;
    double dx2 = (m23 < t1 * t1) ? m23 : t1 * t1;
    t1 = ya - yb;
    t2 = ya - yb + L;
    t3 = ya - yb - L;
    m23 = (t3 * t3 < t2 * t2) ? t3 * t3 : t2 * t2;
    double dy2 = (m23 < t1 * t1) ? m23 : t1 * t1;
    double a2 = dx2 + dy2;
    double sig = type == 1 ? g.s22 : g.s12;
    if (type == 0) {
        if (a2 > g.cut2) return false;
    } else {
        if (a2 * sig * sig > g.cut2) return false;
    }
    double x2 = xb, y2 = yb;
    double dx = xa - x2, dy = ya - y2;
    if (dx * dx + dy * dy > a2) {
        if ((dx - L) * (dx - L) < dx * dx)
            x2 += L;
        else if ((dx + L) * (dx + L) < dx * dx)
            x2 -= L;
        dx = xa - x2;
        if ((dy - L) * (dy - L) < dy * dy)
            y2 += L;
        else if ((dy + L) * (dy + L) < dy * dy)
            y2 -= L;
        dy = ya - y2;
    }
    if (type != 0) {
        a2 *= sig * sig;
        dx *= sig;
        dy *= sig;
    }
    const R ra2 = (R)a2;
    const R a_8 = (R)1.0 / (ra2 * ra2 * ra2 * ra2);
    const R a_14 = (R)1.0 / (ra2 * ra2 * ra2 * ra2 * ra2 * ra2 * ra2);
    R absF;
    if (type == 0)
        absF = (R)4.0 * (R)6.0 * ((R)2.0 * a_14 - a_8);
    else
        absF = (R)(type == 1 ? g.e22f : g.e12f) * (R)4.0 * (R)6.0 * ((R)2.0 * a_14 - a_8);
    if (absF > (R)g.maxpair) absF = (R)g.maxpair;
    fx = absF * (R)dx;
    fy = absF * (R)dy;
    return true;
}

// ------------------------------------------------------------------ block scan
// exclusive scan of a[0..n) in LDS (n <= 4*blockDim.x, blockDim.x <= 1024),
// returns the total; all threads must call it
__device__ int block_exscan4(int *a, int n, int *wsum) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = (blockDim.x + 63) >> 6;
    int v[4];
    int s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int idx = 4 * tid + q;
        v[q] = idx < n ? a[idx] : 0;
        s += v[q];
    }
    int x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int woff = 0, total = 0;
    for (int q = 0; q < nw; ++q) {
        if (q < w) woff += wsum[q];
        total += wsum[q];
    }
    int e = woff + x - s;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        int idx = 4 * tid + q;
        if (idx < n) a[idx] = e;
        e += v[q];
    }
    __syncthreads();
    return total;
}

// ------------------------------------------------------------------ WCA
// One thread per list position.  Neighbours come from the cell list of
// the 3x3 cells around the particle's own cell (the extended-list semantics
// of dm:459-526): entries are slot | species << 31, each cell's entries
// sorted by slot, so every sum has a fixed order.  Same-species and
// cross-species contributions are summed separately and then added, the
// grouping of dm:617-702 + dm:973-1059.
constexpr unsigned LIST_SLOT = 0x7fffffffu;
// rank of a particle that changed cell during the step: its arrival index
// in the new cell, marked (ranks stay below 2^30)
constexpr int RANK_MOVED = 1 << 30;
// Scan of the cell counts: blocks of SCAN_BLOCK cells, SCAN_PER_THREAD
// consecutive cells per thread; bcount[c >> SCAN_SHIFT] is the particle count
// of the block holding cell c
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_SHIFT = 12;
constexpr int SCAN_BLOCK = 1 << SCAN_SHIFT;  // 4096 cells per block
constexpr int SCAN_PER_THREAD = SCAN_BLOCK / SCAN_THREADS;
static_assert(SCAN_PER_THREAD % 4 == 0, "int4 loads");
// bcount entries sit 128 bytes apart: the movers' atomics on the blocks of
// a dense region then spread over many L2 lines instead of a few
constexpr int BCOUNT_STRIDE = 32;

template <typename R>
__device__ __forceinline__ void wca_ref_body(const Geo &g, const DevBufs &d, int blk) {
    using R2 = typename V2<R>::T;
    const int qi = blk * blockDim.x + threadIdx.x;
    if (qi >= g.N) return;
    const int nc = g.nc;
    const unsigned *list = (const unsigned *)d.list;
    const int s = (int)(list[qi] & LIST_SLOT);
    const int c = d.lcell[qi];
    const int cx = div_nc(c, g), cy = c - cx * nc;
    const double2 ri = d.lrel[qi];
    const bool s2i = d.lpid[qi] >= g.n1;
    R fsx = 0, fsy = 0, fcx = 0, fcy = 0;
    {
        // the reference's FP64 arithmetic on absolute positions (am:100-166,
        // dm:640-698); the S2 side of a 12 pair is the negated S1 side
        const double xi = cx * g.B + ri.x, yi = cy * g.B + ri.y;
        for (int da = -1; da <= 1; ++da)
            for (int db = -1; db <= 1; ++db) {
                const int ux = wrap1(cx + da, nc), uy = wrap1(cy + db, nc);
                const int cc = ux * nc + uy;
                const int qe = d.start[cc + 1];
                for (int q = d.start[cc]; q < qe; ++q) {
                    const unsigned ent = list[q];
                    const int sj = (int)(ent & LIST_SLOT);
                    if (sj == s) continue;
                    const bool s2j = (ent >> 31) != 0;
                    const double2 rj = d.lrel[q];
                    const double xj = ux * g.B + rj.x, yj = uy * g.B + rj.y;
                    R fx, fy;
                    if (!s2i && !s2j) {
                        if (wca_pair<R>(xi, yi, xj, yj, 0, g, fx, fy)) { fsx += fx; fsy += fy; }
                    } else if (s2i && s2j) {
                        if (wca_pair<R>(xi, yi, xj, yj, 1, g, fx, fy)) { fsx += fx; fsy += fy; }
                    } else if (!s2i) {
                        if (wca_pair<R>(xi, yi, xj, yj, 2, g, fx, fy)) { fcx += fx; fcy += fy; }
                    } else {
                        if (wca_pair<R>(xj, yj, xi, yi, 2, g, fx, fy)) { fcx -= fx; fcy -= fy; }
                    }
                }
            }
    }
    R2 o;
    o.x = fsx + fcx;
    o.y = fsy + fcy;
    ((R2 *)d.fwca)[qi] = o;
}

template <typename R>
__global__ __launch_bounds__(256) void k_wca(Geo g, DevBufs d) {
    wca_ref_body<R>(g, d, xcd_block(g.xcd));
}

// FP32 forces for boxes of at least 6 cells per side: the displacement is
// the difference of the in-cell offsets (FP32 copies centred on the cell,
// lrelf) plus the cell offset, the minimum image by construction.  UNIT: all
// sigma and epsilon ratios are 1, so the per-type factors drop out and one
// sum collects all partners; otherwise same-species and cross-species
// partners are summed apart and added at the end (the grouping of
// dm:617-702 + dm:973-1059).  Partners are visited in list order, cell by
// cell, row by row.
//
// A block handles consecutive list positions.  Their neighbour cells in row
// da (da = -1, 0, 1: the cells (cx + da, cy - 1 .. cy + 1)) form one
// contiguous range of the list when no particle's row wraps around the box
// in x; the block copies the three ranges' offsets (and entries) to LDS once
// and every particle away from the y edge reads its candidates there.  The
// rest (the first and last cell columns, rows wrapping in y, ranges beyond
// the LDS capacity) read the list directly.  Both loops evaluate a candidate
// with the same code, so a particle's force has the same bits either way
// (and therefore for every block size).
// staged candidates per block: 8 KB of offsets, small enough to share a CU
// with the FFT passes of the other stream (measured in the two-stream step:
// 2048 entries 185.5 us, 1024 180.7 us, 768 183.4 us per step; a dense
// 256-particle block stages about 800)
constexpr int WCA_STAGE = 1024;

// One candidate (entry e, centred offset r) at list position q.  The cell
// offset is folded into this particle's coordinates: y + B below m1 (yb), y
// below m2 (ym), y - B from m2 on (ya), x + ox (xo).  SELF: the row holds
// this particle (position qi, excluded).  Outside the cutoff the force is
// selected away rather than branched around (every wave holds some partner
// inside, so the branch would not skip the arithmetic).  Contraction into
// FMAs is allowed in the FP32 force arithmetic, unlike the FP64 position
// arithmetic.
template <bool UNIT, bool SELF>
__device__ __forceinline__ void wca_f32_cand(unsigned e, float2 r, int q, int qi, bool s2i, int m1, int m2, float xo,
                                             float yb, float ym, float ya, float cut2, float cap, float sigA,
                                             float prefA, float sigB, float prefB, float &fsx, float &fsy,
                                             float &fcx, float &fcy) {
#pragma clang fp contract(fast)
    const float yo = q < m1 ? yb : (q < m2 ? ym : ya);
    const float dx = xo - r.x, dy = yo - r.y;
    const float r2 = dx * dx + dy * dy;
    if (UNIT) {
        const float inv = __builtin_amdgcn_rcpf(r2);
        const float i2 = inv * inv, i4 = i2 * i2;
        float absF = fminf(24.0f * (2.0f * (i4 * i2 * inv) - i4), cap);
        absF = (r2 <= cut2 && (!SELF || q != qi)) ? absF : 0.0f;
        fsx += absF * dx;
        fsy += absF * dy;
    } else {
        const bool s2j = (e >> 31) != 0;
        const float sig = s2j ? sigB : sigA;
        const float r2s = r2 * sig * sig;
        const float inv = __builtin_amdgcn_rcpf(r2s);
        const float i2 = inv * inv, i4 = i2 * i2;
        float absF = fminf((s2j ? prefB : prefA) * 24.0f * (2.0f * (i4 * i2 * inv) - i4), cap);
        absF = (r2s <= cut2 && (!SELF || q != qi)) ? absF : 0.0f;
        const float fx = absF * dx * sig, fy = absF * dy * sig;
        if (s2i == s2j) {
            fsx += fx;
            fsy += fy;
        } else {
            fcx += fx;
            fcy += fy;
        }
    }
}

// candidates q0..qe-1 read from the list, loaded one candidate ahead of the
// arithmetic (the list arrays carry one padding entry).  (Scalar arguments:
// a parameter struct here was placed in LDS by the compiler.)
template <bool UNIT, bool SELF>
__device__ __forceinline__ void wca_f32_row(const unsigned *__restrict__ list, const float2 *__restrict__ rf, int qi,
                                            bool s2i, int q0, int qe, int m1, int m2, float xo, float yb, float ym,
                                            float ya, float cut2, float cap, float sigA, float prefA, float sigB,
                                            float prefB, float &fsx, float &fsy, float &fcx, float &fcy) {
    unsigned ent = list[q0];
    float2 rj = rf[q0];
    for (int q = q0; q < qe; ++q) {
        const unsigned e = ent;
        const float2 r = rj;
        ent = list[q + 1];  // q + 1 <= N: padding entry
        rj = rf[q + 1];
        wca_f32_cand<UNIT, SELF>(e, r, q, qi, s2i, m1, m2, xo, yb, ym, ya, cut2, cap, sigA, prefA, sigB, prefB, fsx,
                                 fsy, fcx, fcy);
    }
}

// the same candidates read from the staged copy (list position q at q + delta)
template <bool UNIT, bool SELF>
__device__ __forceinline__ void wca_f32_row_lds(const float2 *sp, const unsigned *se, int delta, int qi, bool s2i,
                                                int q0, int qe, int m1, int m2, float xo, float yb, float ym,
                                                float ya, float cut2, float cap, float sigA, float prefA, float sigB,
                                                float prefB, float &fsx, float &fsy, float &fcx, float &fcy) {
    float2 rn = sp[q0 + delta];
    unsigned en = UNIT ? 0u : se[q0 + delta];
    for (int q = q0; q < qe; ++q) {
        const float2 r = rn;
        const unsigned e = en;
        rn = sp[q + 1 + delta];  // one ahead: the arrays hold WCA_STAGE + 1 entries
        if (!UNIT) en = se[q + 1 + delta];
        wca_f32_cand<UNIT, SELF>(e, r, q, qi, s2i, m1, m2, xo, yb, ym, ya, cut2, cap, sigA, prefA, sigB, prefB, fsx,
                                 fsy, fcx, fcy);
    }
}

// LDS of the staged rows: offsets sp[WCA_STAGE + 1], entries
// se[WCA_STAGE + 1] (not UNIT; the extra entry is read ahead, never used),
// and hdr[8] (each row's first list position and length, and whether the
// block stages).  Every thread of the block calls this.
template <bool UNIT>
__device__ __forceinline__ void wca_f32_body(const Geo &g, const DevBufs &d, int blk, float2 *sp, unsigned *se,
                                             int *hdr) {
    const int nb = blockDim.x, tid = threadIdx.x;
    const int q_first = blk * nb;
    const int q_last = min(q_first + nb, g.N) - 1;
    const int nc = g.nc;
    const unsigned *list = (const unsigned *)d.list;
    const float2 *rf = d.lrelf;
    if (tid < 3) {
        // lane da + 1: the list range of the cells (cx + da, cy - 1 .. cy + 1)
        // over the block's cells c_first..c_last (list order is cell order)
        const int cf = d.lcell[q_first], cl = d.lcell[q_last];
        const int ok = div_nc(cf, g) >= 1 && div_nc(cl, g) <= nc - 2;
        int g0 = 0, len = 0;
        if (ok) {
            const int da = tid - 1;
            const int lo = max(cf + da * nc - 1, 0), hi = min(cl + da * nc + 1, g.ncell - 1);
            g0 = d.start[lo];
            len = d.start[hi + 1] - g0;
        }
        hdr[tid] = g0;
        hdr[3 + tid] = len;
        if (tid == 0) hdr[6] = ok;
    }
    __syncthreads();
    const int l0 = hdr[3], l1 = hdr[4], total = l0 + l1 + hdr[5];
    const bool staged = hdr[6] && total <= WCA_STAGE;
    if (staged) {
        for (int t = tid; t < total; t += nb) {
            const int r = (t >= l0) + (t >= l0 + l1);
            const int q = hdr[r] + t - (r == 0 ? 0 : (r == 1 ? l0 : l0 + l1));
            sp[t] = rf[q];
            if (!UNIT) se[t] = list[q];
        }
    }
    __syncthreads();
    const int qi = q_first + tid;
    if (qi > q_last) return;
    const int c = d.lcell[qi];
    const int cx = div_nc(c, g), cy = c - cx * nc;
    const float2 ri = rf[qi];
    const bool s2i = d.lpid[qi] >= g.n1;
    const float B = (float)g.B, cut2 = (float)g.cut2, cap = (float)g.maxpair;
    // factors for a partner of species 1 (A) or species 2 (B)
    const float sigA = s2i ? (float)g.s12 : 1.0f, prefA = s2i ? (float)g.e12f : 1.0f;
    const float sigB = s2i ? (float)g.s22 : (float)g.s12, prefB = s2i ? (float)g.e22f : (float)g.e12f;
    const float yb = ri.y + B, ym = ri.y, ya = ri.y - B;
    float fsx = 0, fsy = 0, fcx = 0, fcy = 0;
    // cells cy-1, cy, cy+1 of a row are adjacent in the list
    const bool inner = cy >= 1 && cy <= nc - 2;
    if (staged && inner) {
#define BDK_WCA_ROW(SELF)                                                                                       \
    wca_f32_row_lds<UNIT, SELF>(sp, se, delta, qi, s2i, s0, s3, s1, s2, xo, yb, ym, ya, cut2, cap, sigA, prefA, \
                                sigB, prefB, fsx, fsy, fcx, fcy)
        for (int da = -1; da <= 1; ++da) {
            const float xo = ri.x - da * B;
            const int *st = d.start + (cx + da) * nc + cy - 1;
            const int s0 = st[0], s1 = st[1], s2 = st[2], s3 = st[3];
            const int delta = (da < 0 ? 0 : (da == 0 ? l0 : l0 + l1)) - hdr[da + 1];
            if (da == 0)
                BDK_WCA_ROW(true);
            else
                BDK_WCA_ROW(false);
        }
#undef BDK_WCA_ROW
    } else {
#define BDK_WCA_ROW(SELF, Q0, QE, M1, M2)                                                                    \
    wca_f32_row<UNIT, SELF>(list, rf, qi, s2i, Q0, QE, M1, M2, xo, yb, ym, ya, cut2, cap, sigA, prefA, sigB, \
                            prefB, fsx, fsy, fcx, fcy)
        for (int da = -1; da <= 1; ++da) {
            const int row = wrap1(cx + da, nc) * nc;
            const float xo = ri.x - da * B;
            if (inner) {
                const int *st = d.start + row + cy - 1;
                const int s0 = st[0], s1 = st[1], s2 = st[2], s3 = st[3];
                if (da == 0)
                    BDK_WCA_ROW(true, s0, s3, s1, s2);
                else
                    BDK_WCA_ROW(false, s0, s3, s1, s2);
            } else {
                for (int db = -1; db <= 1; ++db) {
                    const int cc = row + wrap1(cy + db, nc);
                    const int q0 = d.start[cc], qe = d.start[cc + 1];
                    const int m1 = db < 0 ? qe : q0, m2 = db > 0 ? q0 : qe;
                    if (da == 0 && db == 0)
                        BDK_WCA_ROW(true, q0, qe, m1, m2);
                    else
                        BDK_WCA_ROW(false, q0, qe, m1, m2);
                }
            }
        }
#undef BDK_WCA_ROW
    }
    float2 o;
    o.x = fsx + fcx;
    o.y = fsy + fcy;
    ((float2 *)d.fwca)[qi] = o;
}

template <bool UNIT>
__global__ __launch_bounds__(256) void k_wca_f32(Geo g, DevBufs d) {
    __shared__ float2 sp[WCA_STAGE + 1];
    __shared__ unsigned se[UNIT ? 1 : WCA_STAGE + 1];
    __shared__ int hdr[8];
    wca_f32_body<UNIT>(g, d, xcd_block(g.xcd), sp, se, hdr);
}

// ------------------------------------------------------------------ deposit
// Charge assignment (gr:31-547) per tile of tc x tc list cells.  A mesh
// point belongs to the tile whose cells contain its centre,
// owner(k) = floor((k+1/2)h/B).  Every particle of the tile's cells and of
// a halo of hx cells adds its (up to 9) weighted charges to the tile's own
// mesh points in LDS.  The additions are integer (fixed point, scale 2^S),
// so the sums do not depend on the order of the atomics: the deposit is
// deterministic without floating-point atomics.  FP64 mode splits every
// charge into a coarse and a fine integer part (two accumulators), which
// keeps the sum exact to ~2^-(S+40).
constexpr int DP_THREADS = 256;
constexpr int DP_SEG_TAB = 1536;  // staged particles with a segment table entry

__device__ int first_owned(int c, const Geo &g) {
    if (c <= 0) return 0;
    if (c >= g.nc) return g.Z;
    int k = (int)ceil(c * g.B / g.h - 0.5);
    if (k < 0) k = 0;
    if (k > g.Z) k = g.Z;
    auto owner = [&](int kk) { return (int)floor(((kk + 0.5) * g.h) / g.B); };
    while (k > 0 && owner(k - 1) >= c) --k;
    while (k < g.Z && owner(k) < c) ++k;
    return k;
}

__global__ void k_own(Geo g, int *own) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c <= g.nc) own[c] = first_owned(c, g);
}

// FP32 deposit into a window that holds the whole stencil of every staged
// particle (mesh rows kw0 .. kw0 + nwx - 1, columns lw0 .., unwrapped): one
// base index per particle and nine fixed offsets, no per-point bounds.  The
// charges are truncated to 32-bit fixed point (scale 2^S, |charge| 2^S <
// 2^31: a truncation error of 2^-31 of the largest charge, far below FP32
// resolution) and summed in 64 bits, so the sums stay exact and independent
// of the order of the atomics.  dqs: the particle's charge drho q times 2^S
// (exact: a power of two), so a point's fixed-point value is
// (int)((wx wy) dqs).
template <int SCHEME, bool P2>
__device__ __forceinline__ void deposit_particle_win(const Geo &g, const DevBufs &d, float dqs, int gx, int gy,
                                                     double2 rr, int kw0, int lw0, int nwx, int nwy,
                                                     unsigned long long *acc) {
    const double x = gx * g.B + rr.x, y = gy * g.B + rr.y;
    int kp, lp;
    float wx[3], wy[3];
    weights_r<float, SCHEME>(x, g, kp, wx);
    weights_r<float, SCHEME>(y, g, lp, wy);
    const int kb = wrapz<P2>(kp - 1 - kw0, g.Z), lb = wrapz<P2>(lp - 1 - lw0, g.Z);
    if (kb > nwx - 3 || lb > nwy - 3) {  // outside the window: not reachable, flagged
        atomicOr(&d.sc->nonfinite, 4);
        return;
    }
    unsigned long long *p = acc + kb * nwy + lb;
    if (SCHEME == 0) {
        atomicAdd(&p[nwy + 1], (unsigned long long)(long long)(int)dqs);
        return;
    }
#pragma unroll
    for (int ia = 0; ia < 3; ++ia)
#pragma unroll
        for (int ib = 0; ib < 3; ++ib) {
            const int X = (int)((wx[ia] * wy[ib]) * dqs);
            atomicAdd(&p[ia * nwy + ib], (unsigned long long)(long long)X);
        }
}

template <typename R, int SCHEME, bool P2>
__device__ __forceinline__ void deposit_particle(const Geo &g, const DevBufs &d, const DepArgs &a, float dqs, int gx,
                                                 int gy, double2 rr, bool s2, int klo, int khi, int llo, int lhi,
                                                 unsigned long long *acc_hi, unsigned long long *acc_lo) {
    const double x = gx * g.B + rr.x, y = gy * g.B + rr.y;
    int kp, lp;
    R wx[3], wy[3];
    weights_r<R, SCHEME>(x, g, kp, wx);
    weights_r<R, SCHEME>(y, g, lp, wy);
    const int Z = g.Z, nl = lhi - llo;
#pragma unroll
    for (int ia = 0; ia < 3; ++ia) {
        if (SCHEME == 0 && ia != 1) continue;
        if (wx[ia] == 0.0) continue;
        const int k = wrapz<P2>(kp + ia - 1, Z);
        if (k < klo || k >= khi) continue;
#pragma unroll
        for (int ib = 0; ib < 3; ++ib) {
            if (SCHEME == 0 && ib != 1) continue;
            if (wy[ib] == 0.0) continue;
            const int l = wrapz<P2>(lp + ib - 1, Z);
            if (l < llo || l >= lhi) continue;
            const int w = (k - klo) * nl + (l - llo);
            if (std::is_same<R, float>::value) {
                // the charge and its fixed point exactly as in the window path,
                // so the mesh does not depend on which path a tile takes
                const float cf = SCHEME == 0 ? dqs : ((float)wx[ia] * (float)wy[ib]) * dqs;
                atomicAdd(&acc_hi[w], (unsigned long long)(long long)(int)cf);
            } else {
                double cd;
                if (SCHEME == 0)
                    cd = g.drho * (s2 ? a.q2 : a.q1);
                else {
                    cd = wx[ia] * wy[ib] * g.drho;
                    cd = cd * (s2 ? a.q2 : a.q1);
                }
                const long long Xh = __double2ll_rn(cd * a.scale_hi);
                const double rem = cd - (double)Xh * a.iscale_hi;
                const long long Xl = __double2ll_rn(rem * a.scale_lo);
                atomicAdd(&acc_hi[w], (unsigned long long)Xh);
                atomicAdd(&acc_lo[w], (unsigned long long)Xl);
            }
        }
    }
}

// lb: tile index, tiles numbered x fastest over tx tiles per row
template <typename R, int SCHEME, bool P2>
__device__ __forceinline__ void deposit_body(const Geo &g, const DevBufs &d, const DepArgs &a, int lb, int tx) {
    extern __shared__ unsigned char smem_raw[];
    unsigned long long *acc_hi = (unsigned long long *)smem_raw;
    unsigned long long *acc_lo = acc_hi + a.win_cap;
    const int acc_n = max(a.win_cap * (std::is_same<R, double>::value ? 2 : 1), a.win_fast);
    int *seg_q = (int *)(acc_hi + acc_n);
    int *seg_pre = seg_q + a.seg_cap;       // [seg_cap + 1]
    int *seg_row = seg_pre + a.seg_cap + 1;
    unsigned short *seg_of = (unsigned short *)(seg_row + a.seg_cap);  // [DP_SEG_TAB]
    __shared__ int info[4];
    __shared__ int wsum[16];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nc = g.nc, tc = a.tc;
    const int cx0 = (lb % tx) * tc, cy0 = (lb / tx) * tc;
    const int nx = min(tc, nc - cx0), ny = min(tc, nc - cy0);
    const unsigned long long step = d.sc->step;
    // bound on how far a staged particle lies outside its staging cell: the
    // excursion of rel outside [0, B] (list order), or during a rebuild the
    // distance of the new positions outside their previous cells
    __shared__ double e_sh;
    if (a.slot) {
        if (tid < 64) {
            double m = dbits(d.exo[tid * 16]);
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
            if (tid == 0) e_sh = m;
        }
    } else if (tid == 0) {
        e_sh = dbits(d.sc->exc_bits[step & 1]);
    }
    const int *st = a.slot ? d.start_old : d.start;
    if (a.reset && lb == 0 && tid == 0) {
        // per-step scalars for the kernels that follow this one
        d.sc->dt_bits = bitsd(g.dt_max);
        d.sc->exc_bits[(step + 1) & 1] = 0ull;
    }
    if (tid < 4) info[tid] = d.own[tid == 0 ? cx0 : tid == 1 ? cx0 + nx : tid == 2 ? cy0 : cy0 + ny];
    __syncthreads();
    const int klo = info[0], khi = info[1], llo = info[2], lhi = info[3];
    // FP32: the species charges in fixed point, drho q 2^S (2^S exact)
    const float dqs1 = ((float)g.drho * (float)a.q1) * a.scale_f, dqs2 = ((float)g.drho * (float)a.q2) * a.scale_f;
    const double e = e_sh + 1e-9 * g.L;
    int hx = (int)floor((1.5 * g.h + e) / g.B) + 1;
    if (hx < 1) hx = 1;
    // staged cells: the tile plus hx halo cells, each distinct cell once (a
    // halo that would wrap onto the tile itself stages the whole axis)
    const bool allx = nx + 2 * hx >= nc, ally = ny + 2 * hx >= nc;
    const int SX = allx ? nc : nx + 2 * hx, SY = ally ? nc : ny + 2 * hx;
    const int sx0 = allx ? 0 : cx0 - hx, sy0 = ally ? 0 : cy0 - hx;
    // FP32: the window of mesh points reachable from the staged cells (their
    // positions lie within e of them; margins cover the index rounding)
    const int kw0 = (int)floor((sx0 * g.B - e) * g.hinv) - 3, lw0 = (int)floor((sy0 * g.B - e) * g.hinv) - 3;
    const int nwx = (int)floor(((sx0 + SX) * g.B + e) * g.hinv) + 4 - kw0;
    const int nwy = (int)floor(((sy0 + SY) * g.B + e) * g.hinv) + 4 - lw0;
    const bool win = std::is_same<R, float>::value && !allx && !ally && nwx <= g.Z && nwy <= g.Z &&
                     nwx * nwy <= a.win_fast;
    const int nwin = win ? nwx * nwy : (khi - klo) * (lhi - llo);
    for (int w = tid; w < nwin; w += nt) {
        acc_hi[w] = 0ull;
        if (std::is_same<R, double>::value) acc_lo[w] = 0ull;
    }
    // The list holds the particles in cell order, so the staged cells of one
    // row (fixed x) are one contiguous list range, or two when the y range
    // wraps: segments of consecutive list positions.
    const int ywrap = (sy0 < 0 || sy0 + SY > nc) ? 1 : 0;
    const int nsr = 1 + ywrap, nseg = SX * nsr;
    __syncthreads();  // accumulators cleared before the first atomic
    if (nseg <= a.seg_cap) {
        for (int i = tid; i < nseg; i += nt) {
            const int r = i / nsr, part = i - r * nsr;
            const int gx = wrap1(sx0 + r, nc), row = gx * nc;
            int ya, yb;
            if (!ywrap) {
                ya = sy0;
                yb = sy0 + SY;
            } else if (sy0 < 0) {
                ya = part == 0 ? sy0 + nc : 0;
                yb = part == 0 ? nc : sy0 + SY;
            } else {
                ya = part == 0 ? sy0 : 0;
                yb = part == 0 ? nc : sy0 + SY - nc;
            }
            const int qa = st[row + ya];
            seg_q[i] = qa;
            seg_pre[i] = max(st[row + yb] - qa, 0);
            seg_row[i] = gx;
        }
        __syncthreads();
        const int total = block_exscan4(seg_pre, nseg, wsum);
        if (tid == 0) seg_pre[nseg] = total;
        __syncthreads();
        // the segment of every staged particle, written once per segment
        // (a table lookup per particle instead of a binary search)
        const bool tab = total <= DP_SEG_TAB;
        if (tab) {
            for (int i = tid; i < nseg; i += nt)
                for (int t = seg_pre[i], te = seg_pre[i + 1]; t < te; ++t) seg_of[t] = (unsigned short)i;
            __syncthreads();
        }
        for (int t = tid; t < total; t += nt) {
            int lo = 0, hi = nseg - 1;  // last segment with seg_pre <= t
            if (tab)
                lo = seg_of[t];
            else
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (seg_pre[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
            const int q = seg_q[lo] + (t - seg_pre[lo]);
            int gx, gy;
            double2 rr;
            bool s2;
            if (a.slot) {  // the particle's new cell, not its staging row
                const int c = d.cell[q];
                gx = div_nc(c, g);
                gy = c - gx * nc;
                rr = d.rel[q];
                s2 = d.pid[q] >= g.n1;
            } else {
                gx = seg_row[lo];
                gy = d.lcell[q] - gx * nc;
                rr = d.lrel[q];
                s2 = (((const unsigned *)d.list)[q] >> 31) != 0;
            }
            const float dqs = s2 ? dqs2 : dqs1;
            if (win)
                deposit_particle_win<SCHEME, P2>(g, d, dqs, gx, gy, rr, kw0, lw0, nwx, nwy, acc_hi);
            else
                deposit_particle<R, SCHEME, P2>(g, d, a, dqs, gx, gy, rr, s2, klo, khi, llo, lhi, acc_hi, acc_lo);
        }
    } else {
        for (int idx = tid; idx < SX * SY; idx += nt) {
            const int ia = idx / SY, ib = idx - ia * SY;
            const int gx = wrap1(sx0 + ia, nc), gy = wrap1(sy0 + ib, nc);
            const int c = gx * nc + gy;
            for (int q = st[c]; q < st[c + 1]; ++q) {
                int px = gx, py = gy;
                double2 rr;
                bool s2;
                if (a.slot) {
                    const int cq = d.cell[q];
                    px = div_nc(cq, g);
                    py = cq - px * nc;
                    rr = d.rel[q];
                    s2 = d.pid[q] >= g.n1;
                } else {
                    rr = d.lrel[q];
                    s2 = (((const unsigned *)d.list)[q] >> 31) != 0;
                }
                const float dqs = s2 ? dqs2 : dqs1;
                if (win)
                    deposit_particle_win<SCHEME, P2>(g, d, dqs, px, py, rr, kw0, lw0, nwx, nwy, acc_hi);
                else
                    deposit_particle<R, SCHEME, P2>(g, d, a, dqs, px, py, rr, s2, klo, khi, llo, lhi, acc_hi,
                                                    acc_lo);
            }
        }
    }
    __syncthreads();
    R *rho = (R *)d.rho;
    const int nl = lhi - llo, nown = (khi - klo) * nl;
    // w / nl from a float reciprocal, corrected to the exact quotient
    // (w < 2^24): an integer division per point cost more than the point
    const float rnl = nl > 0 ? 1.0f / (float)nl : 0.0f;
    for (int w = tid; w < nown; w += nt) {
        int qk = (int)((float)w * rnl);
        int rl = w - qk * nl;
        if (rl >= nl) {
            ++qk;
            rl -= nl;
        } else if (rl < 0) {
            --qk;
            rl += nl;
        }
        const int k = klo + qk, l = llo + rl;
        double v;
        if (win) {
            v = (double)(long long)acc_hi[wrapz<P2>(k - kw0, g.Z) * nwy + wrapz<P2>(l - lw0, g.Z)] * a.iscale_f;
        } else if (std::is_same<R, float>::value) {
            v = (double)(long long)acc_hi[w] * a.iscale_f;
        } else {
            v = (double)(long long)acc_hi[w] * a.iscale_hi + (double)(long long)acc_lo[w] * a.iscale_lo;
        }
        rho[(size_t)k * g.Z + l] = (R)v;
    }
}

// hardware block order: measured faster than per-XCD ranges for this kernel
// P2: the mesh kind (chosen by the launcher; one body per kernel keeps the
// scalar registers of the power-of-two path as they were)
template <typename R, int SCHEME, bool P2>
__global__ __launch_bounds__(DP_THREADS) void k_deposit(Geo g, DevBufs d, DepArgs a) {
    deposit_body<R, SCHEME, P2>(g, d, a, blockIdx.x + blockIdx.y * gridDim.x, gridDim.x);
}

// End of a step: step counter + 1, and the per-step scalars of the next
// update re-armed (dt minimum and the 64 slots of dt candidates, excursion
// slot, displacement bound).  Run by threads 0..63 of one block of the
// step's last mesh pass (k_c2r_rows), which follows every reader of these
// scalars in the step.
__device__ __forceinline__ void step_end_body(const Geo &g, const DevBufs &d) {
    const int t = threadIdx.x;
    if (t < 64) {
        d.exo[t * 16] = 0ull;
        d.dts[t * 16] = ~0ull;
    }
    if (t == 0) {
        const unsigned long long s = d.sc->step;
        d.sc->dt_bits = bitsd(g.dt_max);
        d.sc->exc_bits[s & 1] = 0ull;  // filled by the update of step s + 1
        d.sc->step = s + 1;
    }
}

// per-step scalar re-arm when no deposit runs (ideal-gas runs), at the start
// of the step; one wave
__global__ void k_rearm(Geo g, DevBufs d) {
    d.dts[threadIdx.x * 16] = ~0ull;
    if (threadIdx.x == 0) {
        const unsigned long long step = d.sc->step;
        d.sc->dt_bits = bitsd(g.dt_max);
        d.sc->exc_bits[(step + 1) & 1] = 0ull;
    }
}
// ------------------------------------------------------------------ FFT in LDS
// tw[k] = exp(-2 pi i k / Z), k < Z; the inverse transforms use the conjugates.
// radix-4 butterfly in place, y = DFT4(v) (inverse: conjugate kernel)
template <typename C>
__device__ __forceinline__ void dft4(C *v, bool inv) {
    C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    C t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    C r3 = inv ? cmul_i(t3) : cmul_mi(t3);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, r3);
    v[3] = csub(t1, r3);
}

// radix-8 butterfly: y[2m] = DFT4(v[t] + v[t+4])[m],
// y[2m+1] = DFT4((v[t] - v[t+4]) W^t)[m], W = exp(-+2 pi i / 8)
template <typename C>
__device__ __forceinline__ void dft8(const C *v, C *y, bool inv) {
    using R = decltype(C::x);
    const R c = (R)0.70710678118654752440;
    C a[4], b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        a[t] = cadd(v[t], v[t + 4]);
        b[t] = csub(v[t], v[t + 4]);
    }
    // b[1] *= W, b[2] *= W^2, b[3] *= W^3
    if (!inv) {
        b[1] = cm<C>(c * (b[1].x + b[1].y), c * (b[1].y - b[1].x));
        b[2] = cmul_mi(b[2]);
        b[3] = cm<C>(c * (b[3].y - b[3].x), -c * (b[3].x + b[3].y));
    } else {
        b[1] = cm<C>(c * (b[1].x - b[1].y), c * (b[1].x + b[1].y));
        b[2] = cmul_i(b[2]);
        b[3] = cm<C>(-c * (b[3].x + b[3].y), c * (b[3].x - b[3].y));
    }
    dft4(a, inv);
    dft4(b, inv);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        y[2 * m] = a[m];
        y[2 * m + 1] = b[m];
    }
}

// radix-16 butterfly: y[2m] = DFT8(v[t] + v[t+8])[m],
// y[2m+1] = DFT8((v[t] - v[t+8]) W^t)[m], W = exp(-+2 pi i / 16)
template <typename C>
__device__ __forceinline__ void dft16(const C *v, C *y, bool inv) {
    using R = decltype(C::x);
    const R c1 = (R)0.92387953251128675613, s1 = (R)0.38268343236508977173;  // cos, sin of pi/8
    const R h = (R)0.70710678118654752440;
    // W^t = cos(t pi/8) -+ i sin(t pi/8), t = 1..7
    const R wr[8] = {(R)1, c1, h, s1, (R)0, -s1, -h, -c1};
    const R wi[8] = {(R)0, s1, h, c1, (R)1, c1, h, s1};
    C a[8], b[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        a[t] = cadd(v[t], v[t + 8]);
        b[t] = csub(v[t], v[t + 8]);
    }
    b[4] = inv ? cmul_i(b[4]) : cmul_mi(b[4]);
#pragma unroll
    for (int t = 1; t < 8; ++t) {
        if (t == 4) continue;
        const R si = inv ? wi[t] : -wi[t];
        b[t] = cm<C>(b[t].x * wr[t] - b[t].y * si, b[t].x * si + b[t].y * wr[t]);
    }
    C ya[8], yb[8];
    dft8(a, ya, inv);
    dft8(b, yb, inv);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        y[2 * m] = ya[m];
        y[2 * m + 1] = yb[m];
    }
}

// LDS layout of the FFT rows: element i of a row sits at lp(i), one padding
// element after every 8.  The output strides of the first Stockham passes
// (8 and 64 elements) then spread over the banks instead of piling onto a
// few (measured: half the LDS cycles of the unpadded passes were bank
// conflicts).  A row of n elements takes n + n/8 + 1 (odd) elements.
__device__ __forceinline__ int lp(int i) { return i + (i >> 3); }
__host__ __device__ __forceinline__ int lds_row(int n) { return n + n / 8 + 1; }

// One Stockham pass of radix RAD (2, 4 or 8) over nrows rows of length
// n = 2^logn held in LDS (row stride `stride`, layout lp), in place: every thread reads
// its butterflies' inputs (and, for small register footprints, their
// twiddles: powers 1, 2, 4 from the FP64-built table, 3, 5, 6, 7 as
// products) before the barrier and writes the outputs after it.  Twiddle of input t of butterfly j: tw[(t k) << twsh],
// k = j mod Ns (conjugated for the inverse transform).
// Output index ((j - k) << LR) + k + t Ns with Ns a power of 8 maps to
// lp(base) + t (Ns + Ns/8); input index j + t nb to lp(j) + t (nb + nb/8)
// when nb is a multiple of 8 (SMALL: nb < 8, mapped per element).
template <typename R, int EPT, int RAD, bool SMALL>
__device__ __forceinline__ void fft_pass(typename V2<R>::T *buf, int logn, int nrows, int stride, int logNs,
                                         const typename V2<R>::T *__restrict__ tw, int logZ, bool inv) {
    using C = typename V2<R>::T;
    constexpr int LR = RAD == 16 ? 4 : (RAD == 8 ? 3 : (RAD == 4 ? 2 : 1));
    // butterflies per thread (radix 16 at 8 elements per thread: one, on half
    // of the threads)
    constexpr int NB = EPT >= RAD ? EPT / RAD : 1;
    constexpr bool PRE = sizeof(R) == 4 && NB * (RAD - 1) <= 16;  // twiddles fetched before the barrier
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lognb = logn - LR, nb = 1 << lognb;
    const int tot = nb * nrows;
    const int Ns = 1 << logNs;
    const int twsh = logZ - (logNs + LR);
    const int nbp = nb + (nb >> 3), Nsp = Ns + (Ns >> 3);
    C v[NB][RAD];
    C w[PRE ? NB : 1][RAD - 1];
    // twiddles of butterfly k: tw^1, tw^2, tw^4 (and tw^8) from the table,
    // the other powers as products (one to three roundings deeper; 3 loads
    // instead of 7 for radix 8, 4 instead of 15 for radix 16).  Every tiling
    // forms them the same way -- before the barrier when the registers allow
    // (PRE), else after it -- so the transform's rounding does not depend on
    // the elements per thread.
    auto twiddles = [&](int k, C *ww) {
        ww[0] = tw[k << twsh];
        if (RAD >= 4) {
            ww[1] = tw[(2 * k) << twsh];
            ww[2] = cmul(ww[0], ww[1]);
        }
        if (RAD >= 8) {
            ww[3] = tw[(4 * k) << twsh];
            ww[4] = cmul(ww[3], ww[0]);
            ww[5] = cmul(ww[3], ww[1]);
            ww[6] = cmul(ww[3], ww[2]);
        }
        if (RAD == 16) {
            ww[7] = tw[(8 * k) << twsh];
#pragma unroll
            for (int t = 8; t < 15; ++t) ww[t] = cmul(ww[7], ww[t - 8]);
        }
    };
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        const int b = tid + e * nt;
        if (b < tot) {
            const int row = b >> lognb, j = b & (nb - 1);
            const C *p = buf + row * stride;
            if (SMALL) {
#pragma unroll
                for (int t = 0; t < RAD; ++t) v[e][t] = p[lp(j + t * nb)];
            } else {
                p += lp(j);
#pragma unroll
                for (int t = 0; t < RAD; ++t) v[e][t] = p[t * nbp];
            }
            if (PRE) twiddles(j & (Ns - 1), w[PRE ? e : 0]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        const int b = tid + e * nt;
        if (b < tot) {
            const int row = b >> lognb, j = b & (nb - 1);
            const int k = j & (Ns - 1);
            C wl[RAD - 1];
            if (!PRE) twiddles(k, wl);
#pragma unroll
            for (int t = 1; t < RAD; ++t) {
                C ww = PRE ? w[PRE ? e : 0][t - 1] : wl[t - 1];
                if (inv) ww.y = -ww.y;
                v[e][t] = cmul(v[e][t], ww);
            }
            C y[RAD];
            if (RAD == 2) {
                y[0] = cadd(v[e][0], v[e][1]);
                y[1] = csub(v[e][0], v[e][1]);
            } else if (RAD == 4) {
#pragma unroll
                for (int t = 0; t < 4; ++t) y[t] = v[e][t];
                dft4(y, inv);
            } else if (RAD == 8) {
                dft8(v[e], y, inv);
            } else {
                dft16(v[e], y, inv);
            }
            if (RAD == 16 && Ns < 8) {
                // the first radix-16 pass (Ns = 1): outputs 16 j + t straddle a
                // padding element, so each takes its own padded index
                C *q = buf + row * stride;
                const int base = (j - k) << LR;
#pragma unroll
                for (int t = 0; t < RAD; ++t) q[lp(base + k + t * Ns)] = y[t];
            } else {
                C *q = buf + row * stride + lp(((j - k) << LR) + k);
#pragma unroll
                for (int t = 0; t < RAD; ++t) q[t * Nsp] = y[t];
            }
        }
    }
    __syncthreads();
}

// radix-3 butterfly, y = DFT3(v): W = exp(-+2 pi i / 3) = -1/2 -+ i sqrt(3)/2
template <typename C>
__device__ __forceinline__ void dft3(const C *v, C *y, bool inv) {
    using R = decltype(C::x);
    const R h = (R)0.86602540378443864676;  // sqrt(3)/2
    const C s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    const C m = cm<C>(v[0].x - (R)0.5 * s.x, v[0].y - (R)0.5 * s.y);
    const C r = inv ? cm<C>(-h * d.y, h * d.x) : cm<C>(h * d.y, -h * d.x);  // (+-i) sqrt(3)/2 (v1 - v2)
    y[0] = cadd(v[0], s);
    y[1] = cadd(m, r);
    y[2] = csub(m, r);
}

// radix-P butterfly for P = 5, 7 as the P-point DFT, y[u] = sum_t v[t] W^(tu
// mod P), W = exp(-+2 pi i / P) = tw[Z / P] (conjugated for the inverse);
// P divides Z.
template <typename C, int P>
__device__ __forceinline__ void dftp(const C *v, C *y, const C *__restrict__ tw, int Z, bool inv) {
    const int zp = Z / P;
    C w[P];
#pragma unroll
    for (int k = 0; k < P; ++k) {
        w[k] = tw[k * zp];
        if (inv) w[k].y = -w[k].y;
    }
#pragma unroll
    for (int u = 0; u < P; ++u) {
        C acc = v[0];
#pragma unroll
        for (int t = 1; t < P; ++t) acc = cadd(acc, cmul(v[t], w[(t * u) % P]));
        y[u] = acc;
    }
}

// One Stockham pass of radix RAD (2, 3, 4, 5, 7 or 8) over rows of any
// length n divisible by RAD (the meshes whose size is not a power of two).
// The same butterflies as fft_pass, with per-element LDS addresses (lp of
// every index), exact integer division for the butterfly and row indices,
// and every twiddle power loaded from the table: tw[t k Z / (Ns RAD)].
// Inputs are read before the barrier, outputs written after it (in place).
template <typename R, int EPT, int RAD>
__device__ __forceinline__ void fft_pass_g(typename V2<R>::T *buf, int n, int nrows, int stride, int Ns,
                                           const typename V2<R>::T *__restrict__ tw, int Z, bool inv) {
    using C = typename V2<R>::T;
    constexpr int NB = (EPT + RAD - 1) / RAD;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nb = n / RAD, tot = nb * nrows;
    const int twm = Z / (Ns * RAD);
    const Dv dnb(nb), dns(Ns);
    C v[NB][RAD];
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        const int b = tid + e * nt;
        if (b < tot) {
            const int row = dnb.q(b), j = b - row * nb;
            const C *p = buf + row * stride;
#pragma unroll
            for (int t = 0; t < RAD; ++t) v[e][t] = p[lp(j + t * nb)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        const int b = tid + e * nt;
        if (b < tot) {
            const int row = dnb.q(b), j = b - row * nb;
            const int k = dns.r(j);
#pragma unroll
            for (int t = 1; t < RAD; ++t) {
                C ww = tw[t * k * twm];
                if (inv) ww.y = -ww.y;
                v[e][t] = cmul(v[e][t], ww);
            }
            C y[RAD];
            if constexpr (RAD == 2) {
                y[0] = cadd(v[e][0], v[e][1]);
                y[1] = csub(v[e][0], v[e][1]);
            } else if constexpr (RAD == 3) {
                dft3(v[e], y, inv);
            } else if constexpr (RAD == 4) {
#pragma unroll
                for (int t = 0; t < 4; ++t) y[t] = v[e][t];
                dft4(y, inv);
            } else if constexpr (RAD == 8) {
                dft8(v[e], y, inv);
            } else {
                dftp<C, RAD>(v[e], y, tw, Z, inv);
            }
            C *q = buf + row * stride;
            const int base = (j - k) * RAD + k;
#pragma unroll
            for (int t = 0; t < RAD; ++t) q[lp(base + t * Ns)] = y[t];
        }
    }
    __syncthreads();
}

// Transform of rows of length n = 2^p 3^c 5^d 7^f: radix-8 passes over the
// factor 2^p, one radix-4 or radix-2 pass for its remaining bits, then one
// radix-3, -5 or -7 pass per odd factor.  A power-of-two n takes fft_pass
// (below).
template <typename R, int EPT>
__device__ __forceinline__ void lds_fft_g(typename V2<R>::T *buf, int n, int nrows, int stride,
                                          const typename V2<R>::T *__restrict__ tw, int Z, bool inv) {
    int m = n, c3 = 0, c5 = 0, c7 = 0;
    while (m % 3 == 0) {
        m /= 3;
        ++c3;
    }
    while (m % 5 == 0) {
        m /= 5;
        ++c5;
    }
    while (m % 7 == 0) {
        m /= 7;
        ++c7;
    }
    int Ns = 1;
    while (m >= 8) {
        fft_pass_g<R, EPT, 8>(buf, n, nrows, stride, Ns, tw, Z, inv);
        Ns *= 8;
        m /= 8;
    }
    if (m == 4) {
        fft_pass_g<R, EPT, 4>(buf, n, nrows, stride, Ns, tw, Z, inv);
        Ns *= 4;
    } else if (m == 2) {
        fft_pass_g<R, EPT, 2>(buf, n, nrows, stride, Ns, tw, Z, inv);
        Ns *= 2;
    }
    for (; c3 > 0; --c3) {
        fft_pass_g<R, EPT, 3>(buf, n, nrows, stride, Ns, tw, Z, inv);
        Ns *= 3;
    }
    for (; c5 > 0; --c5) {
        fft_pass_g<R, EPT, 5>(buf, n, nrows, stride, Ns, tw, Z, inv);
        Ns *= 5;
    }
    for (; c7 > 0; --c7) {
        fft_pass_g<R, EPT, 7>(buf, n, nrows, stride, Ns, tw, Z, inv);
        Ns *= 7;
    }
}

// rows of length n: the power-of-two passes below, or lds_fft_g when n has an
// odd factor (tw holds exp(-2 pi i k / Z), k < Z; n divides Z).  R16: the
// kernel's power-of-two rows take radix-16 passes (FP32); a kernel's choice
// does not depend on its elements per thread, so neither does its rounding.
template <typename R, int EPT, bool R16 = false, bool R16T = false>
__device__ __forceinline__ void lds_fft_pow2(typename V2<R>::T *buf, int logn, int nrows, int stride,
                                             const typename V2<R>::T *__restrict__ tw, int logZ, bool inv);
template <typename R, int EPT>
__device__ __forceinline__ void lds_fft(typename V2<R>::T *buf, int n, int nrows, int stride,
                                        const typename V2<R>::T *__restrict__ tw, int Z, bool inv) {
    if ((n & (n - 1)) == 0 && (Z & (Z - 1)) == 0)
        lds_fft_pow2<R, EPT>(buf, __ffs(n) - 1, nrows, stride, tw, __ffs(Z) - 1, inv);
    else
        lds_fft_g<R, EPT>(buf, n, nrows, stride, tw, Z, inv);
}
// the same with the mesh kind known at compile time (P2: n and Z powers of two)
template <typename R, int EPT, bool P2, bool R16 = false, bool R16T = false>
__device__ __forceinline__ void lds_fft_t(typename V2<R>::T *buf, int n, int nrows, int stride,
                                          const typename V2<R>::T *__restrict__ tw, int Z, bool inv) {
    if (P2)
        lds_fft_pow2<R, EPT, R16, R16T>(buf, __ffs(n) - 1, nrows, stride, tw, __ffs(Z) - 1, inv);
    else
        lds_fft_g<R, EPT>(buf, n, nrows, stride, tw, Z, inv);
}

// radix-8 passes (radix-16 passes with R16), then one pass of a smaller radix
// for the remaining bits
template <typename R, int EPT, bool R16, bool R16T>
__device__ __forceinline__ void lds_fft_pow2(typename V2<R>::T *buf, int logn, int nrows, int stride,
                                             const typename V2<R>::T *__restrict__ tw, int logZ, bool inv) {
    int logNs = 0;
    if constexpr (R16T && sizeof(R) == 4) {
        // radix-8 passes, with the last 4 bits (when radix 8 would leave one
        // bit for a radix-2 pass) in one radix-16 pass: 1024 points in passes
        // of 8, 8 and 16
        if (logn >= 7 && logn % 3 == 1) {
            while (logn - logNs > 4) {
                fft_pass<R, EPT, 8, false>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
                logNs += 3;
            }
            fft_pass<R, EPT, 16, false>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
            return;
        }
    }
    if constexpr (R16 && sizeof(R) == 4) {
        // FP32 rows of at least 128: radix-16 passes, then one pass of radix
        // 8, 4 or 2 for the remaining bits (2048 points in 3 passes instead
        // of 4)
        if (logn >= 7) {
            while (logn - logNs >= 4) {
                fft_pass<R, EPT, 16, false>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
                logNs += 4;
            }
            const int rem = logn - logNs;
            if (rem == 3)
                fft_pass<R, EPT, 8, false>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
            else if (rem == 2)
                fft_pass<R, EPT, 4, false>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
            else if (rem == 1)
                fft_pass<R, EPT, 2, false>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
            return;
        }
    }
    const bool small8 = logn - 3 < 3;  // nb = n/8 below 8
    while (logn - logNs >= 3) {
        if (small8)
            fft_pass<R, EPT, 8, true>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
        else
            fft_pass<R, EPT, 8, false>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
        logNs += 3;
    }
    if (logn - logNs == 2) {
        if (logn - 2 < 3)
            fft_pass<R, EPT, 4, true>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
        else
            fft_pass<R, EPT, 4, false>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
    } else if (logn - logNs == 1) {
        if (logn - 1 < 3)
            fft_pass<R, EPT, 2, true>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
        else
            fft_pass<R, EPT, 2, false>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
    }
}

constexpr int FFT_THREADS = 256;

// Row r2c: rows k0..k0+RB-1 of the real mesh -> half spectrum X[m], m<Z/2,
// with X[0] packing (DC, Nyquist) (both real), written transposed:
// ht[m*Z + k].  Length-Z/2 complex FFT of (x[2n], x[2n+1]) + split; with
// R16 (FP32) a row length of 2^(3m+1) ends in one radix-16 pass.
template <typename R, int EPT, bool P2, bool R16 = false>
__device__ __forceinline__ void r2c_rows_body(const R *__restrict__ rho, typename V2<R>::T *__restrict__ ht,
                                              const typename V2<R>::T *__restrict__ tw, int Z, int RB,
                                              int stride, int blk) {
    using C = typename V2<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C *s = (C *)smem_raw;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = Z >> 1;
    const DvT<P2> dn(n), drb(RB);  // RB divides Z: a power of two with Z
    const int k0 = blk * RB;
    for (int idx = tid; idx < RB * n; idx += nt) {
        const int row = dn.q(idx), m = dn.r(idx);
        s[row * stride + lp(m)] = ((const C *)(rho + (size_t)(k0 + row) * Z))[m];
    }
    __syncthreads();
    lds_fft_t<R, EPT, P2, false, R16>(s, n, RB, stride, tw, Z, false);
    const int half = n >> 1;
    for (int idx = tid; idx < RB * (half + 1); idx += nt) {
        const int row = idx / (half + 1), m = idx - row * (half + 1);
        C *p = s + row * stride;
        if (m == 0) {
            C z = p[0];
            p[0] = cm<C>(z.x + z.y, z.x - z.y);
        } else {
            C a = p[lp(m)], bc = cconj(p[lp(n - m)]);
            C E = cscale(cadd(a, bc), (R)0.5);
            C O = cmul_mi(cscale(csub(a, bc), (R)0.5));
            C WO = cmul(tw[m], O);
            p[lp(m)] = cadd(E, WO);
            if (m != n - m) p[lp(n - m)] = cconj(csub(E, WO));
        }
    }
    __syncthreads();
    for (int idx = tid; idx < RB * n; idx += nt) {
        const int m = drb.q(idx), row = drb.r(idx);
        ht[(size_t)m * Z + k0 + row] = s[row * stride + lp(m)];
    }
}

template <typename R, int EPT, bool P2, bool R16 = false>
__global__ __launch_bounds__(FFT_THREADS) void k_r2c_rows(const R *__restrict__ rho, typename V2<R>::T *__restrict__ ht,
                                                          const typename V2<R>::T *__restrict__ tw, int Z,
                                                          int RB, int stride, int remap) {
    r2c_rows_body<R, EPT, P2, R16>(rho, ht, tw, Z, RB, stride, xcd_block(remap));
}

// spectral factors of one (j, m') pair: gx = sq[j] G / Z / Z, gy = sq[m'] G / Z / Z,
// G = 1/((lap[j] + lap[m']) + 1/lambda^2) (dm:585-593, dm:1384-1394)
template <typename R>
__device__ __forceinline__ void symbols(int j, int mm, const Geo &g, const R *sq, const R *lap, R &gx, R &gy) {
    R G;
    if (j == 0 && mm == 0 && g.zero_mode_g)
        G = (R)0.0;
    else if (sizeof(R) == 4)  // FP32: the hardware reciprocal (1 ulp) instead of a correctly rounded division
        G = (R)__builtin_amdgcn_rcpf((float)((lap[j] + lap[mm]) + (R)g.lam2inv));
    else
        G = (R)1.0 / ((lap[j] + lap[mm]) + (R)g.lam2inv);
    // x / Z / Z == x * 2^(-2 log2 Z) exactly (Z is a power of two)
    const R izz = (R)1.0 / ((R)g.Z * (R)g.Z);
    gx = sq[j] * G * izz;
    gy = sq[mm] * G * izz;
}

// Column pass: rows m0..m0+RC-1 of ht (length Z along x), forward FFT,
// multiply by i*gx and i*gy, inverse FFT of both, out gt[m*Z+k] = (Fx, Fy).
template <typename R, int EPT, bool P2, bool R16 = false>
__device__ __forceinline__ void col_body(const typename V2<R>::T *__restrict__ ht, typename V4<R>::T *__restrict__ gt,
                                         const typename V2<R>::T *__restrict__ tw, const R *__restrict__ sq,
                                         const R *__restrict__ lap, const Geo &g, int RC, int stride, int blk) {
    using C = typename V2<R>::T;
    using C4 = typename V4<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C *s = (C *)smem_raw;
    C *sB = s + RC * stride;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int Z = g.Z, Zh = Z >> 1;
    const DvT<P2> dz(Z);
    const int m0 = blk * RC;
    for (int idx = tid; idx < RC * Z; idx += nt) {
        const int row = dz.q(idx), k = dz.r(idx);
        s[row * stride + lp(k)] = ht[(size_t)(m0 + row) * Z + k];
    }
    __syncthreads();
    lds_fft_t<R, EPT, P2, R16>(s, Z, RC, stride, tw, Z, false);
    for (int idx = tid; idx < RC * Z; idx += nt) {
        const int row = dz.q(idx), j = dz.r(idx);
        const int m = m0 + row;
        if (m == 0) {
            // packed (DC, Nyquist) column: pairs (j, Z-j) by one thread
            if (j > Zh) continue;
            const int jn = P2 ? ((Z - j) & (Z - 1)) : (j == 0 ? 0 : Z - j);
            C P = s[lp(j)], Q = s[lp(jn)];
            C Sdc = cscale(cadd(P, cconj(Q)), (R)0.5);
            C Sny = cmul_mi(cscale(csub(P, cconj(Q)), (R)0.5));
            R gx0, gy0, gxN, gyN, hx0, hy0, hxN, hyN;
            symbols<R>(j, 0, g, sq, lap, gx0, gy0);
            symbols<R>(j, Zh, g, sq, lap, gxN, gyN);
            symbols<R>(jn, 0, g, sq, lap, hx0, hy0);
            symbols<R>(jn, Zh, g, sq, lap, hxN, hyN);
            // packed = i g(.,0) S_dc + i (i g(.,Z/2) S_ny)
            C fxj = cadd(cmul_i(cscale(Sdc, gx0)), cmul_i(cmul_i(cscale(Sny, gxN))));
            C fyj = cadd(cmul_i(cscale(Sdc, gy0)), cmul_i(cmul_i(cscale(Sny, gyN))));
            C fxn = cadd(cmul_i(cscale(cconj(Sdc), hx0)), cmul_i(cmul_i(cscale(cconj(Sny), hxN))));
            C fyn = cadd(cmul_i(cscale(cconj(Sdc), hy0)), cmul_i(cmul_i(cscale(cconj(Sny), hyN))));
            s[lp(j)] = fxj;
            sB[lp(j)] = fyj;
            if (jn != j) {
                s[lp(jn)] = fxn;
                sB[lp(jn)] = fyn;
            }
        } else {
            R gx, gy;
            symbols<R>(j, m, g, sq, lap, gx, gy);
            const int o = row * stride + lp(j);
            C S = s[o];
            s[o] = cm<C>(-gx * S.y, gx * S.x);
            sB[o] = cm<C>(-gy * S.y, gy * S.x);
        }
    }
    __syncthreads();
    lds_fft_t<R, EPT, P2, R16>(s, Z, 2 * RC, stride, tw, Z, true);
    for (int idx = tid; idx < RC * Z; idx += nt) {
        const int row = dz.q(idx), k = dz.r(idx);
        const int so = row * stride + lp(k);
        C a = s[so], b = sB[so];
        C4 o;
        o.x = a.x;
        o.y = a.y;
        o.z = b.x;
        o.w = b.y;
        gt[(size_t)(m0 + row) * Z + k] = o;
    }
}

template <typename R, int EPT, bool P2, bool R16 = false>
__global__ __launch_bounds__(FFT_THREADS) void k_col(const typename V2<R>::T *__restrict__ ht,
                                                     typename V4<R>::T *__restrict__ gt,
                                                     const typename V2<R>::T *__restrict__ tw, const R *__restrict__ sq,
                                                     const R *__restrict__ lap, Geo g, int RC, int stride) {
    col_body<R, EPT, P2, R16>(ht, gt, tw, sq, lap, g, RC, stride, blockIdx.x);
}

// Row c2r of the two real signals Fx, Fy: rows k0..k0+RB2-1, spectrum
// C[m] = X[m] + i Y[m] completed by Hermitian symmetry, inverse FFT of
// length Z, out fmesh[k*Z + l] = (Re, Im) = (Fx, Fy).
template <typename R, int EPT, bool P2, bool R16 = false>
__device__ __forceinline__ void c2r_rows_body(const typename V4<R>::T *__restrict__ gt,
                                              typename V2<R>::T *__restrict__ fm,
                                              const typename V2<R>::T *__restrict__ tw, int Z, int RB2,
                                              int stride, int blk) {
    using C = typename V2<R>::T;
    using C4 = typename V4<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C *s = (C *)smem_raw;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int Zh = Z >> 1;
    const DvT<P2> dz(Z), drb(RB2);  // RB2 divides Z: a power of two with Z
    const int k0 = blk * RB2;
    for (int idx = tid; idx < RB2 * Zh; idx += nt) {
        const int m = drb.q(idx), row = drb.r(idx);
        C4 v = gt[(size_t)m * Z + k0 + row];
        C *p = s + row * stride;
        if (m == 0) {
            p[0] = cm<C>(v.x, v.z);
            p[lp(Zh)] = cm<C>(v.y, v.w);
        } else {
            p[lp(m)] = cm<C>(v.x - v.w, v.y + v.z);
            p[lp(Z - m)] = cm<C>(v.x + v.w, -v.y + v.z);
        }
    }
    __syncthreads();
    lds_fft_t<R, EPT, P2, R16>(s, Z, RB2, stride, tw, Z, true);
    for (int idx = tid; idx < RB2 * Z; idx += nt) {
        const int row = dz.q(idx), l = dz.r(idx);
        fm[(size_t)(k0 + row) * Z + l] = s[row * stride + lp(l)];
    }
}

template <typename R, int EPT, bool P2, bool R16 = false>
__global__ __launch_bounds__(FFT_THREADS) void k_c2r_rows(const typename V4<R>::T *__restrict__ gt,
                                                          typename V2<R>::T *__restrict__ fm,
                                                          const typename V2<R>::T *__restrict__ tw, int Z,
                                                          int RB2, int stride, int remap, Geo g, DevBufs d,
                                                          int end_step) {
    if (end_step && blockIdx.x == 0) step_end_body(g, d);
    c2r_rows_body<R, EPT, P2, R16>(gt, fm, tw, Z, RB2, stride, xcd_block(remap));
}

// ------------------------------------------------------------------ gather, update
template <typename R, int SCHEME, bool P2>
__device__ __forceinline__ void gather_kap(double x, double y, bool s2, const typename V2<R>::T *__restrict__ fm,
                                           const Geo &g, double &fx, double &fy) {
    using R2 = typename V2<R>::T;
    const int Z = g.Z;
    int k, l;
    R wx[3], wy[3];
    weights_r<R, SCHEME>(x, g, k, wx);
    weights_r<R, SCHEME>(y, g, l, wy);
    k = wrapz<P2>(k, Z);
    l = wrapz<P2>(l, Z);
    if (SCHEME == 0) {
        R2 f = fm[(size_t)k * Z + l];
        if (!s2) {
            fx = (double)f.x;
            fy = (double)f.y;
        } else {
            fx = (double)((R)g.f2_f1 * f.x);
            // gr:123 reads Fx[..][1], the imaginary part of the x force (zero
            // up to rounding); ngp_fix reads Fy
            fy = g.ngp_fix ? (double)((R)g.f2_f1 * f.y) : 0.0;
        }
        return;
    }
    const int kx[3] = {wrapz<P2>(k - 1, Z), k, wrapz<P2>(k + 1, Z)};
    const int ly[3] = {wrapz<P2>(l - 1, Z), l, wrapz<P2>(l + 1, Z)};
    R ax = 0, ay = 0;
    // weighted sums of the mesh forces with explicit fused multiply-adds
    // (the weights and indices above follow the reference's rounding)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const R2 *row = fm + (size_t)kx[a] * Z;
        R2 f0 = row[ly[0]], f1 = row[ly[1]], f2 = row[ly[2]];
        const R sx = fma(wy[2], f2.x, fma(wy[1], f1.x, wy[0] * f0.x));
        const R sy = fma(wy[2], f2.y, fma(wy[1], f1.y, wy[0] * f0.y));
        if (a == 0) {
            ax = wx[0] * sx;
            ay = wx[0] * sy;
        } else {
            ax = fma(wx[a], sx, ax);
            ay = fma(wx[a], sy, ay);
        }
    }
    if (s2) {
        ax *= (R)g.f2_f1;
        ay *= (R)g.f2_f1;
    }
    fx = (double)ax;
    fy = (double)ay;
}

__device__ __forceinline__ double dt_candidate(double F, double n, const Geo &g) {
    if (fabs(F) < 1.0e-5) return g.mrd * g.mrd / (2.0 * g.T * n * n);
    double u = n * sqrt(0.5 * g.T) / F;
    double w = sqrt(u * u + g.mrd / fabs(F));
    return (w - fabs(u)) * (w - fabs(u));
}

__device__ __forceinline__ double2 particle_noise(const Geo &g, const DevBufs &d, int pid, unsigned long long step) {
    double2 n;
    if (g.nowca) {
        n.x = 0.0;
        n.y = 0.0;
    } else if (g.noise_mode == BDK_NOISE_UPLOAD) {
        n = d.noise_up[pid];
    } else if (g.f32noise) {
        n = philox_noise_f32(g.seed, (uint32_t)pid, step, g.noise_mode == BDK_NOISE_GAUSS, g.noise_var);
    } else {
        n = philox_noise(g.seed, (uint32_t)pid, step, g.noise_mode == BDK_NOISE_GAUSS, g.noise_var);
    }
    return n;
}

// forces only: fkap (and dt candidates when with_dt) of block blk of list
// positions; the block minimum of the candidates goes to dts slot `slot`
template <typename R, int SCHEME, bool P2>
__device__ __forceinline__ void gather_body(const Geo &g, const DevBufs &d, int with_dt, int blk, int slot) {
    using R2 = typename V2<R>::T;
    const int s = blk * blockDim.x + threadIdx.x;  // list position
    double cand = INFINITY;
    if (s < g.N) {
        const int c = d.lcell[s];
        const int gx = div_nc(c, g), gy = c - gx * g.nc;
        const double2 r = d.lrel[s];
        const int pid = d.lpid[s];
        // the WCA force and the noise before the mesh gather: their loads
        // overlap the gather's instead of waiting behind the fkap store
        R2 fw;
        double2 n;
        if (with_dt) {
            fw = ((const R2 *)d.fwca)[s];
            n = particle_noise(g, d, pid, d.sc->step);
        }
        double fx = 0.0, fy = 0.0;
        if (!g.idgas) gather_kap<R, SCHEME, P2>(gx * g.B + r.x, gy * g.B + r.y, pid >= g.n1, (const R2 *)d.fmesh, g, fx, fy);
        R2 o;
        o.x = (R)fx;
        o.y = (R)fy;
        ((R2 *)d.fkap)[s] = o;
        if (with_dt) {
            double c0 = dt_candidate((double)fw.x + (double)o.x, n.x, g);
            double c1 = dt_candidate((double)fw.y + (double)o.y, n.y, g);
            if (c0 < cand) cand = c0;
            if (c1 < cand) cand = c1;
        }
    }
    if (with_dt) {
        // NaN candidates never win (std::min semantics of dm:1420-1491).
        // Minimum over the block, then one atomicMin into one of 64 slots
        // (128 bytes apart), skipped when the slot already holds less.
        __shared__ double wmin[4];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            double v = __shfl_xor(cand, o, 64);
            if (v < cand) cand = v;
        }
        if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = cand;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
                if (wmin[w] < cand) cand = wmin[w];
            if (cand < INFINITY && cand >= 0.0) {
                unsigned long long *sp = &d.dts[(slot & 63) * 16];
                const unsigned long long b = bitsd(cand);
                if (b < *(volatile unsigned long long *)sp) atomicMin(sp, b);
            }
        }
    }
}

template <typename R, int SCHEME, bool P2>
__global__ __launch_bounds__(256) void k_gather(Geo g, DevBufs d, int with_dt) {
    gather_body<R, SCHEME, P2>(g, d, with_dt, xcd_block(g.xcd), blockIdx.x);
}

// dt for bdk_optimal_dt (the step's update reduces the slots itself):
// dt_max and the 64 slot minima of k_gather folded into
// sc->dt_bits and logged by step; the slots are re-armed (all ones) for the
// next reduction.  One wave.
__device__ __forceinline__ void dt_fold_body(const Geo &g, const DevBufs &d) {
    const int t = threadIdx.x;
    unsigned long long m = d.dts[t * 16];
    d.dts[t * 16] = ~0ull;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const unsigned long long v = __shfl_xor(m, o, 64);
        if (v < m) m = v;
    }
    if (t == 0) {
        const unsigned long long dm = bitsd(g.dt_max);
        const unsigned long long dt = m < dm ? m : dm;
        d.sc->dt_bits = dt;
        d.dtlog[d.sc->step & (kDtLog - 1)] = dbits(dt);
    }
}

__global__ __launch_bounds__(64) void k_dt_fold(Geo g, DevBufs d) { dt_fold_body(g, d); }

// fmod(y, B) for the refresh's y = rel + B (dm:249-250): for |y| < 2B the
// quotient is 0 or 1 and the remainder y or y - B, which is exact (y and B
// within a factor of two); larger moves take fmod itself.  Equal to fmod
// bit for bit everywhere (IEEE fmod is exact too).
__device__ __forceinline__ double fmod_cell(double y, double B) {
    if (y >= 0.0 && y < B) return y;
    if (y >= B && y < 2.0 * B) return y - B;
    if (y > -B && y < 0.0) return y;
    return fmod(y, B);
}

// Euler step (dm:116-157) + refresh (dm:216-325) + cell counting of list
// position s (written to slot s).  Returns how far the stored new position
// (cell and offset) lies outside the particle's previous cell (<= 0 inside).
// FUSED: capillary force gathered here (fixed dt); else read from fkap.
template <typename R, int SCHEME, bool FUSED, bool P2>
__device__ __forceinline__ double update_one(const Geo &g, const DevBufs &d, int s, double dt) {
    using R2 = typename V2<R>::T;
    const unsigned long long step = d.sc->step;
    const int c = d.lcell[s];
    int gx = div_nc(c, g), gy = c - gx * g.nc;
    double2 r = d.lrel[s];
    const int pid = d.lpid[s];
    const bool s2 = pid >= g.n1;
    double fwx = 0.0, fwy = 0.0, fkx = 0.0, fky = 0.0;
    if (!g.idgas) {
        const R2 fw = ((const R2 *)d.fwca)[s];
        fwx = (double)fw.x;
        fwy = (double)fw.y;
        if (FUSED) {
            gather_kap<R, SCHEME, P2>(gx * g.B + r.x, gy * g.B + r.y, s2, (const R2 *)d.fmesh, g, fkx, fky);
        } else {
            const R2 fk = ((const R2 *)d.fkap)[s];
            fkx = (double)fk.x;
            fky = (double)fk.y;
        }
    }
    const double2 n = particle_noise(g, d, pid, step);
    const double sq = g.adaptive ? sqrt(2.0 * g.T * dt) : g.sq_fixed;  // same expression, correctly rounded
    const double dx = dt * (fwx + fkx) + sq * n.x;
    const double dy = dt * (fwy + fky) + sq * n.y;
    double2 rn;
    float2 kc = make_float2(0.f, 0.f);  // Kahan compensation (pos_mode 2)
    if (g.pos_mode == 0) {
        rn.x = r.x + dx;
        rn.y = r.y + dy;
    } else {
        // FP32 in-cell positions: the increment rounded to FP32 and added in
        // FP32 (pos_mode 2: Kahan summation, c = (t - r) - y carries the
        // low-order part lost by the addition into the next step)
        const float rx = (float)r.x, ry = (float)r.y;
        if (g.pos_mode == 1) {
            rn.x = (double)(rx + (float)dx);
            rn.y = (double)(ry + (float)dy);
        } else {
            const float2 c = d.lcomp[s];
            const float yx = (float)dx - c.x, yy = (float)dy - c.y;
            const float tx = rx + yx, ty = ry + yy;
            kc.x = (tx - rx) - yx;
            kc.y = (ty - ry) - yy;
            rn.x = (double)tx;
            rn.y = (double)ty;
        }
    }
    if (s2 && g.restrict_radial) {  // dm:138-155
        const double mitte = 0.5 * g.L;
        const int i2 = pid - g.n1;
        double x = gx * g.B + rn.x, y = gy * g.B + rn.y;
        double dxm = mitte - x, dym = mitte - y;
        double abst = sqrt(dxm * dxm + dym * dym);
        double xs = mitte + abst * cos(2 * M_PI * (int)i2 / g.n2);
        double ys = mitte + abst * sin(2 * M_PI * (int)i2 / g.n2);
        rn.x += (xs - x);
        rn.y += (ys - y);
    }
    const double B = g.B;
    int nc_new = c;
    const bool finite = fabs(rn.x) < 1e9 && fabs(rn.y) < 1e9;
    if (!finite) {
        // non-finite or runaway position: keep the old state, raise the flag
        atomicOr(&d.sc->nonfinite, 1);
        rn = r;
    }
    // the stored new position (cell, offset) relative to the previous cell's
    // corner: the bound the deposit beside the rebuild stages by
    double ox = rn.x, oy = rn.y;
    if (finite && (rn.x < 0.0 || rn.x > B || rn.y < 0.0 || rn.y > B)) {
        const int ic0 = (int)floor((gx * B + rn.x) / B);
        const int jc0 = (int)floor((gy * B + rn.y) / B);
        const int ic = wrapc(ic0 + g.nc, g.nc);
        const int jc = wrapc(jc0 + g.nc, g.nc);
        rn.x = fmod_cell(rn.x + B, B);
        rn.y = fmod_cell(rn.y + B, B);
        // a move of more than one cell downwards keeps a negative offset
        // (fmod takes the sign of rel + B, dm:216-325): the stored position
        // then lies one cell below the moved one, so the bound follows the
        // stored cell and offset
        ox = (ic0 - gx) * B + rn.x;
        oy = (jc0 - gy) * B + rn.y;
        gx = ic;
        gy = jc;
        nc_new = gx * g.nc + gy;
        if (g.pos_mode) {
            // back to FP32 storage; the rounding of the shifted value joins
            // the compensation (the position is t - c)
            const float wx = (float)rn.x, wy = (float)rn.y;
            kc.x += (float)((double)wx - rn.x);
            kc.y += (float)((double)wy - rn.y);
            ox += (double)wx - rn.x;
            oy += (double)wy - rn.y;
            rn.x = (double)wx;
            rn.y = (double)wy;
        }
    }
    const double ex_prev = fmax(fmax(-ox, ox - B), fmax(-oy, oy - B));
    d.cell[s] = nc_new;
    d.rel[s] = rn;
    d.pid[s] = pid;
    if (g.pos_mode == 2) d.comp[s] = kc;
    if (nc_new != c) {
        // a cell change (a few percent of the particles per step): the
        // counts follow, and the rank among the new cell's arrivals comes
        // from arr; particles that stay are ranked by k_rank without atomics
        atomicSub(&d.count[c], 1);
        atomicAdd(&d.count[nc_new], 1);
        if ((c >> SCAN_SHIFT) != (nc_new >> SCAN_SHIFT)) {
            atomicSub(&d.bcount[(c >> SCAN_SHIFT) * BCOUNT_STRIDE], 1);
            atomicAdd(&d.bcount[(nc_new >> SCAN_SHIFT) * BCOUNT_STRIDE], 1);
        }
        d.rank[s] = RANK_MOVED | atomicAdd(&d.arr[nc_new], 1);
    }
    double ex = fmax(fmax(-rn.x, rn.x - B), fmax(-rn.y, rn.y - B));
    if (ex > 0.0) atomicMax(&d.sc->exc_bits[(step + 1) & 1], bitsd(ex));
    if (s == 0) d.sc->dt_last = bitsd(dt);
    return ex_prev;
}

// The update of one block of list positions.  The largest distance of a new
// position outside its previous cell bounds where the next deposit finds the
// particles when it stages them by their previous cells (during the list
// rebuild): reduced per wave, then one atomicMax per wave into one of 64
// slots (128 bytes apart), skipped when the slot already holds as much.
template <typename R, int SCHEME, bool FUSED, bool P2>
__device__ __forceinline__ void update_body(const Geo &g, const DevBufs &d, int blk) {
    const int s = blk * blockDim.x + threadIdx.x;  // list position in, slot out
    // the starts that group the slot arrays after this update (the previous
    // list order), kept for the deposit beside the rebuild
    const int stride = ((g.N + blockDim.x - 1) / blockDim.x) * blockDim.x;
    for (int c = s; c <= g.ncell; c += stride) d.start_old[c] = d.start[c];
    double dt = g.dt_max;
    if (!FUSED && g.adaptive) {
        // the step's dt (dm:1420-1491): dt_max and the 64 slot minima of the
        // candidates from k_gather (all ones = none), reduced by every wave;
        // the step's last mesh pass re-arms the slots (k_rearm at the start
        // of the step when no mesh pass runs)
        unsigned long long mm = d.dts[(threadIdx.x & 63) * 16];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const unsigned long long v = __shfl_xor(mm, o, 64);
            if (v < mm) mm = v;
        }
        const unsigned long long dm = bitsd(g.dt_max);
        dt = dbits(mm < dm ? mm : dm);
        if (blk == 0 && threadIdx.x == 0) {
            d.sc->dt_bits = bitsd(dt);
            d.dtlog[d.sc->step & (kDtLog - 1)] = dt;
        }
    }
    double m = 0.0;
    if (s < g.N) m = update_one<R, SCHEME, FUSED, P2>(g, d, s, dt);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.0) {
        unsigned long long *slot = &d.exo[((s >> 6) & 63) * 16];
        const unsigned long long b = bitsd(m);
        if (b > *(volatile unsigned long long *)slot) atomicMax(slot, b);
    }
}

// (without the fused gather the mesh kind plays no part: P2 = true).
// 128-thread blocks: measured a little faster than 256 (and 512 slower).
template <typename R, int SCHEME, bool FUSED, bool P2>
__global__ __launch_bounds__(128) void k_update(Geo g, DevBufs d) {
    update_body<R, SCHEME, FUSED, P2>(g, d, xcd_block(g.xcd));
}

// refresh (dm:216-325) without motion + counting; used by bdk_refresh
__global__ void k_refresh(Geo g, DevBufs d) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    const int c = d.cell[s];
    int gx = c / g.nc, gy = c - gx * g.nc;
    double2 rn = d.rel[s];
    const double B = g.B;
    int nc_new = c;
    if (!(fabs(rn.x) < 1e9 && fabs(rn.y) < 1e9)) {
        atomicOr(&d.sc->nonfinite, 1);
    } else if (rn.x < 0.0 || rn.x > B || rn.y < 0.0 || rn.y > B) {
        int ic = (int)floor((gx * B + rn.x) / B);
        int jc = (int)floor((gy * B + rn.y) / B);
        ic = wrapc(ic + g.nc, g.nc);
        jc = wrapc(jc + g.nc, g.nc);
        rn.x = fmod(rn.x + B, B);
        rn.y = fmod(rn.y + B, B);
        nc_new = ic * g.nc + jc;
    }
    d.cell[s] = nc_new;
    d.rel[s] = rn;
    d.rank[s] = atomicAdd(&d.count[nc_new], 1);
    double ex = fmax(fmax(-rn.x, rn.x - B), fmax(-rn.y, rn.y - B));
    if (ex > 0.0) atomicMax(&d.sc->exc_bits[d.sc->step & 1], bitsd(ex));
}

// counting only (after an upload): rank and excursion
__global__ void k_bin(Geo g, DevBufs d) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    const int c = d.cell[s];
    d.rank[s] = atomicAdd(&d.count[c], 1);
    const double2 rn = d.rel[s];
    const double B = g.B;
    double ex = fmax(fmax(-rn.x, rn.x - B), fmax(-rn.y, rn.y - B));
    if (ex > 0.0) atomicMax(&d.sc->exc_bits[d.sc->step & 1], bitsd(ex));
}

// clears the excursion slot that k_bin / k_refresh fill
__global__ void k_reset_exc(DevScalars *sc) { sc->exc_bits[sc->step & 1] = 0ull; }

// noise of the next step, by particle id (bdk_get_noise)
__global__ void k_noise_out(Geo g, DevBufs d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.N) return;
    d.noise_out[p] = particle_noise(g, d, p, d.sc->step);
}

// ------------------------------------------------------------------ scan + fill
// Full recount of bcount from the cell counts (after an upload or a
// refresh; during steps the update keeps bcount current).
__device__ __forceinline__ void scan_reduce_body(const int *__restrict__ count, int ncell, int *__restrict__ bcount,
                                                 int blk) {
    __shared__ int ws[SCAN_THREADS / 64];
    const int base = blk * SCAN_BLOCK;
    int s = 0;
    for (int q = threadIdx.x; q < SCAN_BLOCK; q += SCAN_THREADS) {
        int i = base + q;
        if (i < ncell) s += count[i];
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int w = 0; w < SCAN_THREADS / 64; ++w) t += ws[w];
        bcount[blk * BCOUNT_STRIDE] = t;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const int *__restrict__ count, int ncell,
                                                              int *__restrict__ bcount) {
    scan_reduce_body(count, ncell, bcount, blockIdx.x);
}

// Exclusive scan of the cell counts into start[] in one pass: a block scans
// its SCAN_BLOCK cells (int4 loads, wave scans, one LDS exchange) and adds
// the particles of the blocks before it, summed from bcount.  The last
// thread of the last block closes the step: start[ncell] = total, step
// counter += inc_step.
__device__ __forceinline__ void scan_body(const int *__restrict__ count, const int *__restrict__ bcount,
                                          int *__restrict__ start, int ncell, int nb, int nmax, DevScalars *sc,
                                          int inc_step, int blk) {
    constexpr int NW = SCAN_THREADS / 64;
    __shared__ int ws[2 * NW];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int p = 0;
    for (int i = tid; i < blk; i += SCAN_THREADS) p += bcount[i * BCOUNT_STRIDE];
    const int base = blk * SCAN_BLOCK + tid * SCAN_PER_THREAD;
    const bool full = base + SCAN_PER_THREAD <= ncell;
    int loc[SCAN_PER_THREAD];
    if (full) {
        const int4 *c4 = (const int4 *)(count + base);
#pragma unroll
        for (int k = 0; k < SCAN_PER_THREAD / 4; ++k) {
            const int4 v = c4[k];
            loc[4 * k] = v.x;
            loc[4 * k + 1] = v.y;
            loc[4 * k + 2] = v.z;
            loc[4 * k + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < SCAN_PER_THREAD; ++q) loc[q] = base + q < ncell ? count[base + q] : 0;
    }
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER_THREAD; ++q) s += loc[q];
    int x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) p += __shfl_xor(p, o, 64);
    if (lane == 63) ws[w] = x;
    if (lane == 0) ws[NW + w] = p;
    __syncthreads();
    int e = x - s;
#pragma unroll
    for (int k = 0; k < NW; ++k) e += ws[NW + k] + (k < w ? ws[k] : 0);
    // starts are clamped into [0, total of the recount]: list ranges stay
    // inside the arrays even if the counts were inconsistent (then flagged)
    int out[SCAN_PER_THREAD];
#pragma unroll
    for (int q = 0; q < SCAN_PER_THREAD; ++q) {
        out[q] = min(max(e, 0), nmax);
        e += loc[q];
    }
    if (full) {
        int4 *s4 = (int4 *)(start + base);
#pragma unroll
        for (int k = 0; k < SCAN_PER_THREAD / 4; ++k)
            s4[k] = make_int4(out[4 * k], out[4 * k + 1], out[4 * k + 2], out[4 * k + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < SCAN_PER_THREAD; ++q)
            if (base + q < ncell) start[base + q] = out[q];
    }
    if (blk == nb - 1 && tid == SCAN_THREADS - 1) {
        start[ncell] = min(max(e, 0), nmax);
        if (e != nmax) atomicOr(&sc->nonfinite, 2);
        if (inc_step) sc->step += 1;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan(const int *__restrict__ count, const int *__restrict__ bcount,
                                                       int *__restrict__ start, int ncell, int nb, int nmax,
                                                       DevScalars *sc, int inc_step) {
    scan_body(count, bcount, start, ncell, nb, nmax, sc, inc_step, blockIdx.x);
}

// Rank of a particle that stayed in its cell during the step: its position
// among the cell's previous entries, minus the entries before it that left.
// Reads the previous starts from the update's copy, so it runs beside the
// scan that overwrites start.
__device__ __forceinline__ void rank_body(const Geo &g, const DevBufs &d, int blk) {
    const int q = blk * blockDim.x + threadIdx.x;
    if (q >= g.N) return;
    const int c = d.lcell[q];
    if (d.cell[q] != c) return;  // moved: ranked among the arrivals
    const int b = d.start_old[c];
    int r = q - b;
    for (int p = b; p < q; ++p) r -= d.cell[p] != d.lcell[p];
    d.rank[q] = r;
}

// the ranks and the scan of the new counts as one launch: blocks [0, nrank)
// rank, the others scan
__global__ __launch_bounds__(256) void k_rank_scan(Geo g, DevBufs d, int nrank, int nb, int inc_step) {
    if ((int)blockIdx.x < nrank)
        rank_body(g, d, blockIdx.x);
    else
        scan_body(d.count, d.bcount, d.start, g.ncell, nb, g.N, d.sc, inc_step, blockIdx.x - nrank);
}
// List entries and the list-ordered state from the ranks.  After a step
// (moved_ranks) a cell holds its stayers first (ascending slot, from
// k_rank), then its arrivals (marked ranks, ordered by k_sort_cells); after
// a full binning every rank is the position in the cell (ordered by
// k_sort_cells).
__device__ __forceinline__ void fill_body(const Geo &g, const DevBufs &d, int moved_ranks, int blk) {
    const int s = blk * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    const int pid = d.pid[s];
    const int c = d.cell[s];
    const int rk = d.rank[s];
    int pos = d.start[c] + (rk & (RANK_MOVED - 1));
    if (moved_ranks && (rk & RANK_MOVED)) pos += d.start[c + 1] - d.start[c] - d.arr[c];
    if ((unsigned)pos >= (unsigned)g.N) {  // inconsistent counts: flagged, nothing written
        atomicOr(&d.sc->nonfinite, 2);
        return;
    }
    const double2 r = d.rel[s];
    ((unsigned *)d.list)[pos] = (unsigned)s | ((pid >= g.n1 ? 1u : 0u) << 31);
    d.lcell[pos] = c;
    d.lrel[pos] = r;
    d.lpid[pos] = pid;
    if (g.fast_wca) d.lrelf[pos] = make_float2((float)(r.x - 0.5 * g.B), (float)(r.y - 0.5 * g.B));
    if (g.pos_mode == 2) d.lcomp[pos] = d.comp[s];
}

__global__ void k_fill(Geo g, DevBufs d, int moved_ranks) { fill_body(g, d, moved_ranks, blockIdx.x); }

__global__ void k_zero_int(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Fixed order inside each cell: after a step only the arrivals (the last
// arr[c] entries) need sorting by slot, after a full binning the whole cell.
// Entries move together with their list-ordered state.
__device__ __forceinline__ void sort_cells_body(const Geo &g, const DevBufs &d, int moved_ranks, int blk) {
    const int c = blk * blockDim.x + threadIdx.x;
    if (c >= g.ncell) return;
    int lo = 0, n, b;
    if (moved_ranks) {
        // after a step most cells had no arrival and one or two had one: only
        // cells with two or more arrivals read their range and sort
        const int a = d.arr[c];
        if (a == 0) return;
        d.arr[c] = 0;  // arrivals of the step, consumed by k_fill
        if (a < 2) return;
        b = d.start[c];
        n = d.start[c + 1] - b;
        lo = n - a;
    } else {
        b = d.start[c];
        n = d.start[c + 1] - b;
        d.arr[c] = 0;
    }
    if (lo < 0 || n - lo < 2) return;  // (lo < 0: inconsistent counts, flagged by the scan)
    unsigned *l = (unsigned *)d.list + b;
    double2 *rl = d.lrel + b;
    int *pl = d.lpid + b;
    float2 *fl = d.lrelf + b;
    const bool f = g.fast_wca, kah = g.pos_mode == 2;
    float2 *cl = kah ? d.lcomp + b : nullptr;
    for (int i = lo + 1; i < n; ++i) {
        const unsigned v = l[i];
        const double2 vr = rl[i];
        const int vp = pl[i];
        const float2 vf = f ? fl[i] : make_float2(0.f, 0.f);
        const float2 vc = kah ? cl[i] : make_float2(0.f, 0.f);
        int j = i - 1;
        while (j >= lo && (l[j] & LIST_SLOT) > (v & LIST_SLOT)) {
            l[j + 1] = l[j];
            rl[j + 1] = rl[j];
            pl[j + 1] = pl[j];
            if (f) fl[j + 1] = fl[j];
            if (kah) cl[j + 1] = cl[j];
            --j;
        }
        l[j + 1] = v;
        rl[j + 1] = vr;
        pl[j + 1] = vp;
        if (f) fl[j + 1] = vf;
        if (kah) cl[j + 1] = vc;
    }
}

__global__ void k_sort_cells(Geo g, DevBufs d, int moved_ranks) { sort_cells_body(g, d, moved_ranks, blockIdx.x); }


// ------------------------------------------------------------------ batches
// A batch advances many independent simulations (members) with one launch
// per stage.  A batched kernel maps its block through the stage's table to
// (member, block of that member) and runs the same device body as the
// single-simulation kernel on that member's geometry and buffers, so every
// member's results are bit-identical alone and in a batch.

// Every batched kernel walks its stage's block map: one map entry per block,
// or in the persistent form (a grid smaller than the map) each block takes
// the entries blockIdx.x, blockIdx.x + gridDim.x, ...  Kernels whose bodies
// use LDS end every entry with a barrier, so the next entry may refill it.
// A member's results do not depend on the grid or on the WCA block size
// (per-particle sums, integer deposit sums, order-free min/max reductions).
// P (persistent) selects the form at compile time: without it a kernel
// handles exactly its own entry (no loop, so the code and its register use
// are those of a plain one-entry kernel).
template <bool P, bool SYNC, typename F> __device__ __forceinline__ void for_entries(int nmap, F &&f) {
    if (!P) {
        f((int)blockIdx.x);
        return;
    }
    for (int e = blockIdx.x; e < nmap; e += gridDim.x) {
        f(e);
        if (SYNC) __syncthreads();
    }
}

template <typename R, bool P>
__global__ __launch_bounds__(512) void kb_wca(const RepDev *m, const int2 *map, int nmap) {
    constexpr int S = std::is_same<R, float>::value ? WCA_STAGE + 1 : 1;  // FP64 members read the list directly
    __shared__ float2 sp[S];
    __shared__ unsigned se[S];
    __shared__ int hdr[8];
    for_entries<P, true>(nmap, [&](int e_) {
        const int2 t = map[e_];
        const RepDev &r = m[t.x];
        if (std::is_same<R, float>::value && r.g.fast_wca) {
            if (r.unit)
                wca_f32_body<true>(r.g, r.d, t.y, sp, se, hdr);
            else
                wca_f32_body<false>(r.g, r.d, t.y, sp, se, hdr);
        } else {
            wca_ref_body<R>(r.g, r.d, t.y);
        }
    });
}

template <typename R, bool P, bool P2>
__global__ __launch_bounds__(DP_THREADS) void kb_deposit(const RepDev *m, const int2 *map, int nmap, int slot) {
    for_entries<P, true>(nmap, [&](int e_) {
        const int2 t = map[e_];
        const RepDev &r = m[t.x];
        DepArgs a = r.dep;
        if (slot) {  // beside the rebuild: the step's end re-arms the scalars
            a.slot = 1;
            a.reset = 0;
        }
        switch (r.g.scheme) {
        case 0: deposit_body<R, 0, P2>(r.g, r.d, a, t.y, r.dep_tx); break;
        case 1: deposit_body<R, 1, P2>(r.g, r.d, a, t.y, r.dep_tx); break;
        default: deposit_body<R, 2, P2>(r.g, r.d, a, t.y, r.dep_tx); break;
        }
    });
}

template <typename R, int EPT, bool P, bool P2, bool R16 = false>
__global__ __launch_bounds__(FFT_THREADS) void kb_r2c(const RepDev *m, const int2 *map, int nmap) {
    using C = typename V2<R>::T;
    for_entries<P, true>(nmap, [&](int e_) {
        const int2 t = map[e_];
        const RepDev &r = m[t.x];
        r2c_rows_body<R, EPT, P2, R16>((const R *)r.d.rho, (C *)r.d.ht, (const C *)r.d.tw, r.g.Z, r.rb, r.sR1, t.y);
    });
}

template <typename R, int EPT, bool P, bool P2, bool R16 = false>
__global__ __launch_bounds__(FFT_THREADS) void kb_col(const RepDev *m, const int2 *map, int nmap) {
    using C = typename V2<R>::T;
    using C4 = typename V4<R>::T;
    for_entries<P, true>(nmap, [&](int e_) {
        const int2 t = map[e_];
        const RepDev &r = m[t.x];
        const R *sq = std::is_same<R, float>::value ? (const R *)r.d.sqf : (const R *)r.d.sq;
        const R *lap = std::is_same<R, float>::value ? (const R *)r.d.lapf : (const R *)r.d.lap;
        col_body<R, EPT, P2, R16>((const C *)r.d.ht, (C4 *)r.d.gt, (const C *)r.d.tw, sq, lap, r.g, r.rc, r.sC, t.y);
    });
}

template <typename R, int EPT, bool P, bool P2, bool R16 = false>
__global__ __launch_bounds__(FFT_THREADS) void kb_c2r(const RepDev *m, const int2 *map, int nmap, int end_step) {
    using C = typename V2<R>::T;
    using C4 = typename V4<R>::T;
    for_entries<P, true>(nmap, [&](int e_) {
        const int2 t = map[e_];
        const RepDev &r = m[t.x];
        if (end_step && t.y == 0) step_end_body(r.g, r.d);
        c2r_rows_body<R, EPT, P2, R16>((const C4 *)r.d.gt, (C *)r.d.fmesh, (const C *)r.d.tw, r.g.Z, r.rb2,
                                       r.sR2, t.y);
    });
}

// Members with a fixed dt gather the mesh force inside the update; members
// with adaptive dt read the forces of kb_gather and reduce its dt
// candidates, as their single-simulation steps do.
template <typename R, bool P, bool P2>
__global__ __launch_bounds__(256) void kb_update(const RepDev *m, const int2 *map, int nmap) {
    for_entries<P, false>(nmap, [&](int e_) {
        const int2 t = map[e_];
        const RepDev &r = m[t.x];
        if (r.g.adaptive) {
            switch (r.g.scheme) {
            case 0: update_body<R, 0, false, true>(r.g, r.d, t.y); break;
            case 1: update_body<R, 1, false, true>(r.g, r.d, t.y); break;
            default: update_body<R, 2, false, true>(r.g, r.d, t.y); break;
            }
        } else {
            switch (r.g.scheme) {
            case 0: update_body<R, 0, true, P2>(r.g, r.d, t.y); break;
            case 1: update_body<R, 1, true, P2>(r.g, r.d, t.y); break;
            default: update_body<R, 2, true, P2>(r.g, r.d, t.y); break;
            }
        }
    });
}

// capillary forces and dt candidates of the adaptive members (their
// particle blocks only)
template <typename R, bool P, bool P2>
__global__ __launch_bounds__(256) void kb_gather(const RepDev *m, const int2 *map, int nmap) {
    for_entries<P, true>(nmap, [&](int e_) {
        const int2 t = map[e_];
        const RepDev &r = m[t.x];
        switch (r.g.scheme) {
        case 0: gather_body<R, 0, P2>(r.g, r.d, 1, t.y, e_); break;
        case 1: gather_body<R, 1, P2>(r.g, r.d, 1, t.y, e_); break;
        default: gather_body<R, 2, P2>(r.g, r.d, 1, t.y, e_); break;
        }
    });
}

// entries [0, nrank): ranks (map mp); entries [nrank, nrank + nscan): scans (map ms)
template <bool P>
__global__ __launch_bounds__(256) void kb_rank_scan(const RepDev *m, const int2 *mp, int nrank, const int2 *ms,
                                                    int nscan) {
    for_entries<P, true>(nrank + nscan, [&](int e_) {
        if (e_ < nrank) {
            const int2 t = mp[e_];
            rank_body(m[t.x].g, m[t.x].d, t.y);
        } else {
            const int2 t = ms[e_ - nrank];
            const RepDev &r = m[t.x];
            scan_body(r.d.count, r.d.bcount, r.d.start, r.g.ncell, r.scan_nb, r.g.N, r.d.sc, 0, t.y);
        }
    });
}

template <bool P>
__global__ __launch_bounds__(256) void kb_fill(const RepDev *m, const int2 *map, int nmap) {
    for_entries<P, false>(nmap, [&](int e_) {
        const int2 t = map[e_];
        fill_body(m[t.x].g, m[t.x].d, 1, t.y);
    });
}

template <bool P>
__global__ __launch_bounds__(256) void kb_sort_cells(const RepDev *m, const int2 *map, int nmap) {
    for_entries<P, false>(nmap, [&](int e_) {
        const int2 t = map[e_];
        sort_cells_body(m[t.x].g, m[t.x].d, 1, t.y);
    });
}

// ------------------------------------------------------------------ launchers
static inline unsigned nblk(int n, int b) { return (unsigned)((n + b - 1) / b); }

// Elements per thread (EPT) of the three FFT kernels; a block of
// FFT_THREADS threads holds EPT * FFT_THREADS complex values.  The defaults
// (8, 16, 16) were the fastest of {8, 16, 32} per kernel at Z = 2048 (FP32):
// 2 rows per r2c block (1024 blocks), one column per k_col block (1024),
// 2 rows per c2r block (1024).  Each EPT is raised to the smallest value
// whose block still holds one transform.  BDK_FFT_EPT="a,b,c" (8, 16 or 32)
// overrides them.
struct FftCfg {
    int e1, e2, e3;      // EPT of r2c rows, columns, c2r rows
    int RB, RC, RB2;     // rows per block
    int sR1, sC, sR2;    // LDS row strides (padded)
    size_t l1, lc, l2;   // LDS bytes
};

static const int *fft_ept_default() {
    static int v[3] = {0, 0, 0};
    if (!v[0]) {
        int a = 8, b = 16, c = 16;
        if (const char *e = getenv("BDK_FFT_EPT")) {
            int x, y, z;
            if (sscanf(e, "%d,%d,%d", &x, &y, &z) == 3) {
                auto ok = [](int t) { return t == 8 || t == 16 || t == 32; };
                if (ok(x) && ok(y) && ok(z)) a = x, b = y, c = z;
            }
        }
        v[1] = b;
        v[2] = c;
        v[0] = a;
    }
    return v;
}

static int ept_for(int want, int need) {
    int e = want;
    while (e * FFT_THREADS < need && e < 32) e *= 2;
    return e;
}

// the largest divisor of m not above cap (at least 1)
static int div_at_most(int m, int cap) {
    cap = std::max(1, std::min(m, cap));
    while (m % cap) --cap;
    return cap;
}

// FFT rows per block for elements-per-thread (e1, e2, e3) at mesh size Z:
// as many rows as the block's elements hold, a divisor of the row count
// (for a power-of-two Z the element count itself)
static void fft_rows(int Z, const int e[3], int &rb, int &rc, int &rb2) {
    rb = div_at_most(Z, e[0] * FFT_THREADS / (Z / 2));
    rc = div_at_most(Z / 2, e[1] * FFT_THREADS / (2 * Z));
    rb2 = div_at_most(Z, e[2] * FFT_THREADS / Z);
}

template <typename R> static FftCfg fft_cfg(int Z) {
    using C = typename V2<R>::T;
    const int *w = fft_ept_default();
    FftCfg c;
    c.e1 = ept_for(w[0], Z / 2);
    c.e2 = ept_for(w[1], 2 * Z);
    c.e3 = ept_for(w[2], Z);
    const int e[3] = {c.e1, c.e2, c.e3};
    fft_rows(Z, e, c.RB, c.RC, c.RB2);
    c.sR1 = lds_row(Z / 2);
    c.sC = lds_row(Z);
    c.sR2 = lds_row(Z);
    c.l1 = (size_t)c.RB * c.sR1 * sizeof(C);
    c.lc = (size_t)2 * c.RC * c.sC * sizeof(C);
    c.l2 = (size_t)c.RB2 * c.sR2 * sizeof(C);
    return c;
}

hipError_t set_kernel_attributes() {
    const int big = 160 * 1024;
    hipError_t e = hipSuccess;
#define BDK_ATTR(K)                                                                            \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)K, hipFuncAttributeMaxDynamicSharedMemorySize, big);
#define BDK_ATTR_E(E, Q)                \
    BDK_ATTR((k_r2c_rows<float, E, Q>))  \
    BDK_ATTR((k_r2c_rows<double, E, Q>)) \
    BDK_ATTR((k_col<float, E, Q>))       \
    BDK_ATTR((k_col<double, E, Q>))      \
    BDK_ATTR((k_c2r_rows<float, E, Q>))  \
    BDK_ATTR((k_c2r_rows<double, E, Q>))
    BDK_ATTR_E(8, true)
    BDK_ATTR_E(16, true)
    BDK_ATTR_E(32, true)
    BDK_ATTR_E(8, false)
    BDK_ATTR_E(16, false)
    BDK_ATTR_E(32, false)
#undef BDK_ATTR_E
    // the radix-16 column and c2r passes (FP32, power-of-two meshes)
#define BDK_ATTR_16(E)                         \
    BDK_ATTR((k_r2c_rows<float, E, true, true>)) \
    BDK_ATTR((kb_r2c<float, E, false, true, true>)) \
    BDK_ATTR((kb_r2c<float, E, true, true, true>))  \
    BDK_ATTR((k_col<float, E, true, true>))      \
    BDK_ATTR((k_c2r_rows<float, E, true, true>)) \
    BDK_ATTR((kb_col<float, E, false, true, true>)) \
    BDK_ATTR((kb_col<float, E, true, true, true>))  \
    BDK_ATTR((kb_c2r<float, E, false, true, true>)) \
    BDK_ATTR((kb_c2r<float, E, true, true, true>))
    BDK_ATTR_16(8)
    BDK_ATTR_16(16)
    BDK_ATTR_16(32)
#undef BDK_ATTR_16
#define BDK_ATTR_B(E, P, Q)            \
    BDK_ATTR((kb_r2c<float, E, P, Q>))  \
    BDK_ATTR((kb_r2c<double, E, P, Q>)) \
    BDK_ATTR((kb_col<float, E, P, Q>))  \
    BDK_ATTR((kb_col<double, E, P, Q>)) \
    BDK_ATTR((kb_c2r<float, E, P, Q>))  \
    BDK_ATTR((kb_c2r<double, E, P, Q>))
#define BDK_ATTR_BQ(Q)           \
    BDK_ATTR_B(8, false, Q)  \
    BDK_ATTR_B(16, false, Q) \
    BDK_ATTR_B(32, false, Q) \
    BDK_ATTR_B(8, true, Q)   \
    BDK_ATTR_B(16, true, Q)  \
    BDK_ATTR_B(32, true, Q)
    BDK_ATTR_BQ(true)
    BDK_ATTR_BQ(false)
#undef BDK_ATTR_BQ
#undef BDK_ATTR_B
#undef BDK_ATTR
    // the deposit also has a little static LDS: ask for 128 KB of dynamic LDS
#define BDK_ATTR_D(K)                                                                          \
    if (e == hipSuccess) e = hipFuncSetAttribute((const void *)K, hipFuncAttributeMaxDynamicSharedMemorySize, 128 * 1024);
#define BDK_ATTR_DQ(Q)                      \
    BDK_ATTR_D((k_deposit<float, 0, Q>))    \
    BDK_ATTR_D((k_deposit<float, 1, Q>))    \
    BDK_ATTR_D((k_deposit<float, 2, Q>))    \
    BDK_ATTR_D((k_deposit<double, 0, Q>))   \
    BDK_ATTR_D((k_deposit<double, 1, Q>))   \
    BDK_ATTR_D((k_deposit<double, 2, Q>))   \
    BDK_ATTR_D((kb_deposit<float, false, Q>))  \
    BDK_ATTR_D((kb_deposit<double, false, Q>)) \
    BDK_ATTR_D((kb_deposit<float, true, Q>))   \
    BDK_ATTR_D((kb_deposit<double, true, Q>))
    BDK_ATTR_DQ(true)
    BDK_ATTR_DQ(false)
#undef BDK_ATTR_DQ
#undef BDK_ATTR_D
    return e;
}

hipError_t launch_wca(bdk_t *h, hipStream_t st) {
    // FP32 forces use the cell-relative displacement when the box holds at
    // least 6 cells per side (the neighbour offset is then the minimum image)
    const dim3 grid(nblk(h->g.N, 256)), blk(256);
    if (!st) st = h->stream;
    if (h->g.fast_wca) {
        const bool unit = h->p.sigma11_22 == 1.0 && h->p.sigma11_12 == 1.0 && h->p.eps22_11 == 1.0 &&
                          h->p.eps12_11 == 1.0;
        if (unit)
            hipLaunchKernelGGL(k_wca_f32<true>, grid, blk, 0, st, h->g, h->d);
        else
            hipLaunchKernelGGL(k_wca_f32<false>, grid, blk, 0, st, h->g, h->d);
    } else if (h->p.precision != BDK_FP64) {
        hipLaunchKernelGGL(k_wca<float>, grid, blk, 0, st, h->g, h->d);
    } else {
        hipLaunchKernelGGL(k_wca<double>, grid, blk, 0, st, h->g, h->d);
    }
    return hipGetLastError();
}

// Deposit tiling: tile side tc (list cells) such that a tile owns at most
// about tile_pts x tile_pts mesh points (38 unless a batch tuned it); fixed-point scales chosen so that no sum of
// N charges can overflow 63 bits.
struct DepPlan {
    DepArgs a;
    dim3 grid;
    size_t lds;
};

static DepPlan dep_plan(const bdk_t *h, int reset, int tile_pts = 38) {
    const Geo &g = h->g;
    const bool f64 = h->p.precision == BDK_FP64;
    DepPlan pl;
    int tc = (int)std::floor(tile_pts * g.h / g.B);
    if (tc < 1) tc = 1;
    if (tc > g.nc) tc = g.nc;
    const int side = (int)std::floor(tc * g.B / g.h) + 2;
    pl.a.tc = tc;
    pl.a.win_cap = side * side;
    pl.a.seg_cap = std::min(2 * (tc + 6), 4 * DP_THREADS);  // block_exscan4 limit
    const double cmax = g.drho * std::max(1.0, std::fabs(g.f2_f1));
    const double nlog = std::log2((double)g.N * cmax + 1.0);
    int s1 = (int)std::floor(62.0 - nlog);
    if (!f64) s1 = std::min(s1, 40);
    const int gap = std::min(40, 61 - (int)std::ceil(std::log2((double)g.N + 1.0)));
    const int s2 = s1 + gap;
    pl.a.scale_hi = std::ldexp(1.0, s1);
    pl.a.iscale_hi = std::ldexp(1.0, -s1);
    pl.a.scale_lo = std::ldexp(1.0, s2);
    pl.a.iscale_lo = std::ldexp(1.0, -s2);
    pl.a.reset = reset;
    pl.a.q1 = 1.0;  // species charges of the capillary deposit (gr:31-547)
    pl.a.q2 = g.f2_f1;
    pl.a.slot = 0;
    // FP32 stencil window, sized for halos of up to 2 cells and excursions
    // up to B/4 (larger ones take the bounds-checked path): the staged span
    // plus the excursion on both sides and the index margins
    pl.a.win_fast = 0;
    pl.a.scale_f = 1.0f;
    pl.a.iscale_f = 1.0;
    if (!f64) {
        const int wside = (int)std::ceil((tc + 4.5) * g.B / g.h) + 8;
        pl.a.win_fast = wside * wside;
        const int sf = (int)std::floor(30.0 - std::log2(cmax));
        pl.a.scale_f = std::ldexp(1.0f, sf);
        pl.a.iscale_f = std::ldexp(1.0, -sf);
    }
    pl.grid = dim3(nblk(g.nc, tc), nblk(g.nc, tc));
    pl.lds = (size_t)std::max(pl.a.win_cap * (f64 ? 2 : 1), pl.a.win_fast) * 8 + (size_t)(3 * pl.a.seg_cap + 1) * 4 +
             (size_t)DP_SEG_TAB * 2;
    return pl;
}

size_t deposit_lds_bytes(const bdk_t *h) { return dep_plan(h, 0).lds; }

static bool zp2(int Z) { return (Z & (Z - 1)) == 0; }

// FP32 power-of-two meshes of at least FFT_R16_MIN points per side run the
// column and c2r passes in radix-16 kernels and the r2c rows with a radix-16
// last pass where radix 8 would leave one bit (fewer LDS passes); smaller
// meshes keep the radix-8 kernels, whose registers the radix-16 code would
// raise.  The choice depends on the mesh only, so a simulation takes the same
// kernels alone and in a batch.
constexpr int FFT_R16_MIN = 1024;
static bool fft_r16(int Z) { return zp2(Z) && Z >= FFT_R16_MIN; }

template <typename R, int SCHEME> static void launch_deposit_t(bdk_t *h, const DepPlan &pl, hipStream_t st) {
    if (zp2(h->g.Z))
        hipLaunchKernelGGL((k_deposit<R, SCHEME, true>), pl.grid, dim3(DP_THREADS), pl.lds, st, h->g, h->d, pl.a);
    else
        hipLaunchKernelGGL((k_deposit<R, SCHEME, false>), pl.grid, dim3(DP_THREADS), pl.lds, st, h->g, h->d, pl.a);
}

hipError_t launch_deposit(bdk_t *h, int reset, int slot, hipStream_t st) {
    DepPlan pl = dep_plan(h, reset);
    pl.a.slot = slot;
    if (!st) st = h->stream;
    const bool f32 = h->p.precision != BDK_FP64;
    switch (h->g.scheme) {
    case 0: f32 ? launch_deposit_t<float, 0>(h, pl, st) : launch_deposit_t<double, 0>(h, pl, st); break;
    case 1: f32 ? launch_deposit_t<float, 1>(h, pl, st) : launch_deposit_t<double, 1>(h, pl, st); break;
    default: f32 ? launch_deposit_t<float, 2>(h, pl, st) : launch_deposit_t<double, 2>(h, pl, st); break;
    }
    return hipGetLastError();
}

hipError_t launch_rearm(bdk_t *h) {
    hipLaunchKernelGGL(k_rearm, dim3(1), dim3(64), 0, h->stream, h->g, h->d);
    return hipGetLastError();
}

#define BDK_EPT_SWITCH(E, CALL) \
    switch (E) {                \
    case 8: CALL(8); break;     \
    case 16: CALL(16); break;   \
    default: CALL(32); break;   \
    }

template <typename R, bool Q, bool R16> static void launch_fft_tq(bdk_t *h, int only, hipStream_t st, int end_step) {
    using C = typename V2<R>::T;
    using C4 = typename V4<R>::T;
    const int Z = h->g.Z;
    const FftCfg c = fft_cfg<R>(Z);
    const R *sq = std::is_same<R, float>::value ? (const R *)h->d.sqf : (const R *)h->d.sq;
    const R *lap = std::is_same<R, float>::value ? (const R *)h->d.lapf : (const R *)h->d.lap;
#define R2C(E)                                                                                              \
    hipLaunchKernelGGL((k_r2c_rows<R, E, Q, R16>), dim3(Z / c.RB), dim3(FFT_THREADS), c.l1, st,     \
                       (const R *)h->d.rho, (C *)h->d.ht, (const C *)h->d.tw, Z, c.RB, c.sR1, h->g.xcd)
#define COL(E)                                                                                              \
    hipLaunchKernelGGL((k_col<R, E, Q, R16>), dim3((Z / 2) / c.RC), dim3(FFT_THREADS), c.lc, st,    \
                       (const C *)h->d.ht, (C4 *)h->d.gt, (const C *)h->d.tw, sq, lap, h->g, c.RC, c.sC)
#define C2R(E)                                                                                              \
    hipLaunchKernelGGL((k_c2r_rows<R, E, Q, R16>), dim3(Z / c.RB2), dim3(FFT_THREADS), c.l2, st,    \
                       (const C4 *)h->d.gt, (C *)h->d.fmesh, (const C *)h->d.tw, Z, c.RB2, c.sR2, h->g.xcd, h->g, h->d,  \
                       end_step)
    if (only < 0 || only == 0) BDK_EPT_SWITCH(c.e1, R2C)
    if (only < 0 || only == 1) BDK_EPT_SWITCH(c.e2, COL)
    if (only < 0 || only == 2) BDK_EPT_SWITCH(c.e3, C2R)
#undef R2C
#undef COL
#undef C2R
}

template <typename R> static void launch_fft_t(bdk_t *h, int only = -1, hipStream_t st = nullptr, int end_step = 0) {
    if (!st) st = h->stream;
    if constexpr (std::is_same<R, float>::value) {
        if (fft_r16(h->g.Z)) {
            launch_fft_tq<R, true, true>(h, only, st, end_step);
            return;
        }
    }
    if (zp2(h->g.Z))
        launch_fft_tq<R, true, false>(h, only, st, end_step);
    else
        launch_fft_tq<R, false, false>(h, only, st, end_step);
}

hipError_t launch_fft_pass(bdk_t *h, int pass, hipStream_t st, int end_step) {
    if (h->p.precision != BDK_FP64)
        launch_fft_t<float>(h, pass, st, end_step);
    else
        launch_fft_t<double>(h, pass, st, end_step);
    return hipGetLastError();
}

hipError_t launch_fft(bdk_t *h) {
    if (h->p.precision != BDK_FP64)
        launch_fft_t<float>(h);
    else
        launch_fft_t<double>(h);
    return hipGetLastError();
}

template <typename R, int SCHEME> static void launch_gather_t(bdk_t *h, int with_dt, int fold) {
    if (zp2(h->g.Z))
        hipLaunchKernelGGL((k_gather<R, SCHEME, true>), dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d,
                           with_dt);
    else
        hipLaunchKernelGGL((k_gather<R, SCHEME, false>), dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g,
                           h->d, with_dt);
    if (with_dt && fold) hipLaunchKernelGGL(k_dt_fold, dim3(1), dim3(64), 0, h->stream, h->g, h->d);
}

hipError_t launch_gather(bdk_t *h, int with_dt, int fold) {
    const bool f32 = h->p.precision != BDK_FP64;
    switch (h->g.scheme) {
    case 0: f32 ? launch_gather_t<float, 0>(h, with_dt, fold) : launch_gather_t<double, 0>(h, with_dt, fold); break;
    case 1: f32 ? launch_gather_t<float, 1>(h, with_dt, fold) : launch_gather_t<double, 1>(h, with_dt, fold); break;
    default: f32 ? launch_gather_t<float, 2>(h, with_dt, fold) : launch_gather_t<double, 2>(h, with_dt, fold); break;
    }
    return hipGetLastError();
}

template <typename R, int SCHEME> static void launch_update_t(bdk_t *h, int fused) {
    if (fused && zp2(h->g.Z))
        hipLaunchKernelGGL((k_update<R, SCHEME, true, true>), dim3(nblk(h->g.N, 128)), dim3(128), 0, h->stream,
                           h->g, h->d);
    else if (fused)
        hipLaunchKernelGGL((k_update<R, SCHEME, true, false>), dim3(nblk(h->g.N, 128)), dim3(128), 0, h->stream,
                           h->g, h->d);
    else
        hipLaunchKernelGGL((k_update<R, SCHEME, false, true>), dim3(nblk(h->g.N, 128)), dim3(128), 0, h->stream,
                           h->g, h->d);
}

hipError_t launch_update(bdk_t *h, int fused) {
    const bool f32 = h->p.precision != BDK_FP64;
    switch (h->g.scheme) {
    case 0: f32 ? launch_update_t<float, 0>(h, fused) : launch_update_t<double, 0>(h, fused); break;
    case 1: f32 ? launch_update_t<float, 1>(h, fused) : launch_update_t<double, 1>(h, fused); break;
    default: f32 ? launch_update_t<float, 2>(h, fused) : launch_update_t<double, 2>(h, fused); break;
    }
    return hipGetLastError();
}

static void launch_scan_fill_inner(bdk_t *h, int inc_step) {
    const int ncell = h->g.ncell;
    const int nb = (int)nblk(ncell, SCAN_BLOCK);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, h->stream, h->d.count, ncell, h->d.bcount);
    hipLaunchKernelGGL(k_scan, dim3(nb), dim3(SCAN_THREADS), 0, h->stream, h->d.count, h->d.bcount, h->d.start, ncell,
                       nb, h->g.N, h->d.sc, inc_step);
    hipLaunchKernelGGL(k_fill, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d, 0);
    hipLaunchKernelGGL(k_sort_cells, dim3(nblk(ncell, 256)), dim3(256), 0, h->stream, h->g, h->d, 0);
}

hipError_t launch_scan_fill(bdk_t *h) {
    launch_scan_fill_inner(h, 1);
    return hipGetLastError();
}

hipError_t launch_own(bdk_t *h) {
    hipLaunchKernelGGL(k_own, dim3(nblk(h->g.nc + 1, 256)), dim3(256), 0, h->stream, h->g, h->d.own);
    return hipGetLastError();
}

hipError_t launch_bin(bdk_t *h) {
    hipLaunchKernelGGL(k_reset_exc, dim3(1), dim3(1), 0, h->stream, h->d.sc);
    hipLaunchKernelGGL(k_bin, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d);
    launch_scan_fill_inner(h, 0);
    return hipGetLastError();
}

hipError_t launch_refresh(bdk_t *h) {
    hipLaunchKernelGGL(k_reset_exc, dim3(1), dim3(1), 0, h->stream, h->d.sc);
    hipLaunchKernelGGL(k_zero_int, dim3(nblk(h->g.ncell, 256)), dim3(256), 0, h->stream, h->d.count, h->g.ncell);
    hipLaunchKernelGGL(k_refresh, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d);
    launch_scan_fill_inner(h, 0);
    return hipGetLastError();
}

hipError_t launch_noise_out(bdk_t *h) {
    hipLaunchKernelGGL(k_noise_out, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d);
    return hipGetLastError();
}

// rho(k) of one species (record_rhoFFTW, auswertung.cpp:529-626): the
// species' number density (weights drho, gridRoutinen.cpp:672-1040) is
// transformed (r2c rows + this column pass), multiplied by (-1)^(i+j)
// (the origin moved to the box centre) and the real parts are summed over
// shells bin = (int)(dq_dqFT |k|/dq + 1/2) of the full Z x Z plane: a
// column m of the half spectrum stands for itself and its mirror
// (Z-j, Z-m), whose conjugate adds the same real part (the imaginary parts
// cancel).  One thread per column sums in a fixed order into that column's
// partial row; k_rhok_final adds the partial rows in order.
template <typename R, int EPT>
__global__ __launch_bounds__(FFT_THREADS) void k_col_rhok(const typename V2<R>::T *__restrict__ ht,
                                                          const typename V2<R>::T *__restrict__ tw, int Z,
                                                          double dqr, int nbins, double *part) {
    using C = typename V2<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C *s = (C *)smem_raw;
    const int m = blockIdx.x, tid = threadIdx.x, Zh = Z >> 1;
    for (int k = tid; k < Z; k += blockDim.x) s[lp(k)] = ht[(size_t)m * Z + k];
    double *p0 = part + (size_t)m * nbins, *pn = part + (size_t)Zh * nbins;
    for (int b = tid; b < nbins; b += blockDim.x) {
        p0[b] = 0.0;
        if (m == 0) pn[b] = 0.0;
    }
    __syncthreads();
    lds_fft<R, EPT>(s, Z, 1, lds_row(Z), tw, Z, false);
    if (tid != 0) return;
    for (int j = 0; j < Z; ++j) {
        const double xq = (double)(wrap1(j + Zh, Z) - Zh);
        const double sg = ((j + m) & 1) ? -1.0 : 1.0;
        if (m > 0) {
            const int bin = (int)(dqr * sqrt(xq * xq + (double)m * m) + 0.5);
            if (bin < nbins) p0[bin] += 2.0 * sg * (double)s[lp(j)].x;
        } else {
            const C P = s[lp(j)], Q = s[lp(j == 0 ? 0 : Z - j)];
            const double dc = 0.5 * ((double)P.x + (double)Q.x);  // Re of the DC column
            const double ny = 0.5 * ((double)P.y + (double)Q.y);  // Re of the Nyquist column
            int bin = (int)(dqr * sqrt(xq * xq) + 0.5);
            if (bin < nbins) p0[bin] += sg * dc;
            bin = (int)(dqr * sqrt(xq * xq + (double)Zh * Zh) + 0.5);
            if (bin < nbins) pn[bin] += sg * ny;  // (-1)^(j + Z/2) = (-1)^j
        }
    }
}

__global__ void k_rhok_final(const double *part, int ncol, int nbins, double *out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins) return;
    double v = 0.0;
    for (int m = 0; m < ncol; ++m) v += part[(size_t)m * nbins + b];
    out[b] = v;
}

template <typename R> static void launch_rhok_t(bdk_t *h, double dqr, int nbins, double *part, double *out) {
    using C = typename V2<R>::T;
    const int Z = h->g.Z;
    const int ept = ept_for(8, Z);
    const size_t lds = (size_t)lds_row(Z) * sizeof(C);
#define RHOK(E)                                                                                              \
    (void)hipFuncSetAttribute((const void *)k_col_rhok<R, E>, hipFuncAttributeMaxDynamicSharedMemorySize,      \
                              160 * 1024);                                                                     \
    hipLaunchKernelGGL((k_col_rhok<R, E>), dim3(Z / 2), dim3(FFT_THREADS), lds, h->stream, (const C *)h->d.ht, \
                       (const C *)h->d.tw, Z, dqr, nbins, part)
    BDK_EPT_SWITCH(ept, RHOK)
#undef RHOK
    hipLaunchKernelGGL(k_rhok_final, dim3(nblk(nbins, 256)), dim3(256), 0, h->stream, part, Z / 2 + 1, nbins, out);
}

// species density (weights q1, q2 per species) -> rho, r2c rows, column pass + shells
hipError_t launch_rhok(bdk_t *h, int species, double dqr, int nbins, double *part, double *out) {
    DepPlan pl = dep_plan(h, 0);
    pl.a.q1 = species == 1 ? 1.0 : 0.0;
    pl.a.q2 = species == 2 ? 1.0 : 0.0;
    hipStream_t st = h->stream;
    const bool f32 = h->p.precision != BDK_FP64;
    switch (h->g.scheme) {
    case 0: f32 ? launch_deposit_t<float, 0>(h, pl, st) : launch_deposit_t<double, 0>(h, pl, st); break;
    case 1: f32 ? launch_deposit_t<float, 1>(h, pl, st) : launch_deposit_t<double, 1>(h, pl, st); break;
    default: f32 ? launch_deposit_t<float, 2>(h, pl, st) : launch_deposit_t<double, 2>(h, pl, st); break;
    }
    hipError_t e = launch_fft_pass(h, 0);
    if (e != hipSuccess) return e;
    if (f32)
        launch_rhok_t<float>(h, dqr, nbins, part, out);
    else
        launch_rhok_t<double>(h, dqr, nbins, part, out);
    return hipGetLastError();
}

// Pair correlation g11, g12, g22 (record_korrelationsfunktion11/12/22,
// auswertung.cpp:110-188): minimum-image distances (abstand2, am:100-166)
// of all pairs closer than nbins*dr, found through the cells within R cells,
// binned by (int)(r/dr).  Ordered pairs (12 pairs from the species-1 side)
// are counted with integer atomics in LDS per block, then added to the
// global histograms: the counts are exact and deterministic.
__global__ __launch_bounds__(256) void k_pair_hist(Geo g, DevBufs d, int nbins, double dr, int R,
                                                   unsigned long long *hist) {
    extern __shared__ unsigned int shist[];  // [3][nbins]
    for (int b = threadIdx.x; b < 3 * nbins; b += blockDim.x) shist[b] = 0u;
    __syncthreads();
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < g.N) {
        const int nc = g.nc, c = d.lcell[q];
        const int cx = div_nc(c, g), cy = c - cx * nc;
        const double2 ri = d.lrel[q];
        const double xi = cx * g.B + ri.x, yi = cy * g.B + ri.y, L = g.L;
        const bool s2i = d.lpid[q] >= g.n1;
        const bool all = 2 * R + 1 >= nc;
        const int a0 = all ? 0 : cx - R, a1 = all ? nc - 1 : cx + R;
        const int b0 = all ? 0 : cy - R, b1 = all ? nc - 1 : cy + R;
        for (int a = a0; a <= a1; ++a) {
            const int ux = wrapc(a, nc);
            for (int b = b0; b <= b1; ++b) {
                const int uy = wrapc(b, nc), cc = ux * nc + uy;
                const int qe = d.start[cc + 1];
                for (int j = d.start[cc]; j < qe; ++j) {
                    if (j == q) continue;
                    const bool s2j = d.lpid[j] >= g.n1;
                    if (s2i && !s2j) continue;
                    const double2 rj = d.lrel[j];
                    double t1 = xi - (ux * g.B + rj.x), t2 = t1 + L, t3 = t1 - L;
                    double m = t3 * t3 < t2 * t2 ? t3 * t3 : t2 * t2;
                    const double dx2 = m < t1 * t1 ? m : t1 * t1;
                    t1 = yi - (uy * g.B + rj.y);
                    t2 = t1 + L;
                    t3 = t1 - L;
                    m = t3 * t3 < t2 * t2 ? t3 * t3 : t2 * t2;
                    const double dy2 = m < t1 * t1 ? m : t1 * t1;
                    const int idx = (int)(sqrt(dx2 + dy2) / dr);
                    if (idx < 0 || idx >= nbins) continue;
                    const int type = s2i ? 2 : (s2j ? 1 : 0);
                    atomicAdd(&shist[type * nbins + idx], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 3 * nbins; b += blockDim.x)
        if (shist[b]) atomicAdd(&hist[b], (unsigned long long)shist[b]);
}

hipError_t launch_pair_hist(bdk_t *h, int nbins, double dr, unsigned long long *hist) {
    int R = (int)ceil(nbins * dr / h->g.B);
    if (R < 1) R = 1;
    const size_t lds = (size_t)3 * nbins * sizeof(unsigned int);
    (void)hipFuncSetAttribute((const void *)k_pair_hist, hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024);
    hipLaunchKernelGGL(k_pair_hist, dim3(nblk(h->g.N, 256)), dim3(256), lds, h->stream, h->g, h->d, nbins, dr, R,
                       hist);
    return hipGetLastError();
}

// ------------------------------------------------------------------ observables
// Radial density profiles of the blob (berechne_dichteprofile,
// observablen_auswertung.cpp:227-330) around the circular-mean centres of
// the species (berechne_mittelpunkte, observablen_auswertung.cpp:140-223).
// Sums use a fixed-order reduction (per-block tree, then one block over the
// block partials in order) and the histograms integer atomics, so the
// result is deterministic.
constexpr int OBS_THREADS = 256;

// block partial sums of cos/sin(2 pi x / L), cos/sin(2 pi y / L) per species
__global__ __launch_bounds__(OBS_THREADS) void k_circ_partial(Geo g, DevBufs d, double *part) {
    __shared__ double sm[OBS_THREADS][8];
    const int q = blockIdx.x * OBS_THREADS + threadIdx.x;
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (q < g.N) {
        const int c = d.lcell[q];
        const int gx = div_nc(c, g), gy = c - gx * g.nc;
        const double2 r = d.lrel[q];
        const double k = 2.0 * M_PI / g.L;
        const double x = gx * g.B + r.x, y = gy * g.B + r.y;
        const int o = d.lpid[q] >= g.n1 ? 4 : 0;
        v[o + 0] = cos(x * k);
        v[o + 1] = sin(x * k);
        v[o + 2] = cos(y * k);
        v[o + 3] = sin(y * k);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) sm[threadIdx.x][i] = v[i];
    __syncthreads();
    for (int w = OBS_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int i = 0; i < 8; ++i) sm[threadIdx.x][i] += sm[threadIdx.x + w][i];
        __syncthreads();
    }
    if (threadIdx.x < 8) part[blockIdx.x * 8 + threadIdx.x] = sm[0][threadIdx.x];
}

// centres S1, S2, Sgamma = (S1+S2)/2 from the block partials (one block)
__global__ __launch_bounds__(OBS_THREADS) void k_circ_final(Geo g, const double *part, int nb, double *centers) {
    __shared__ double sm[OBS_THREADS][8];
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nb; b += OBS_THREADS)
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] += part[b * 8 + i];
#pragma unroll
    for (int i = 0; i < 8; ++i) sm[threadIdx.x][i] = v[i];
    __syncthreads();
    for (int w = OBS_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int i = 0; i < 8; ++i) sm[threadIdx.x][i] += sm[threadIdx.x + w][i];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double k = 2.0 * M_PI / g.L;
        double S[4];
        for (int a = 0; a < 4; ++a)  // (S1x, S1y, S2x, S2y)
            S[a] = (atan2(-sm[0][2 * a + 1], -sm[0][2 * a]) + M_PI) / k;
        centers[0] = S[0];
        centers[1] = S[1];
        centers[2] = S[2];
        centers[3] = S[3];
        centers[4] = 0.5 * (S[0] + S[2]);
        centers[5] = 0.5 * (S[1] + S[3]);
    }
}

// counts per bin: rho1ring (Sgamma), rho2ring (S2), rho1zen1, rho1zen2,
// rho2zen1, rho2zen2; distances without periodic images, beyond the last
// bin into the last bin (as the reference)
__global__ __launch_bounds__(OBS_THREADS) void k_radial_hist(Geo g, DevBufs d, const double *centers, int nbins,
                                                             double dr, int *hist) {
    const int q = blockIdx.x * OBS_THREADS + threadIdx.x;
    if (q >= g.N) return;
    const int c = d.lcell[q];
    const int gx = div_nc(c, g), gy = c - gx * g.nc;
    const double2 r = d.lrel[q];
    const double x = gx * g.B + r.x, y = gy * g.B + r.y;
    const bool s2 = d.lpid[q] >= g.n1;
    // (histogram, centre index) pairs of this species
    const int hs[3] = {s2 ? 1 : 0, s2 ? 4 : 2, s2 ? 5 : 3};
    const int cs[3] = {s2 ? 1 : 2, 0, 1};  // 0: S1, 1: S2, 2: Sgamma
    for (int i = 0; i < 3; ++i) {
        const double dx = x - centers[2 * cs[i]], dy = y - centers[2 * cs[i] + 1];
        const double rr = sqrt(dx * dx + dy * dy);
        int bin = (int)(rr / dr);
        if (bin < 0 || bin >= nbins) bin = nbins - 1;
        atomicAdd(&hist[hs[i] * nbins + bin], 1);
    }
}

hipError_t launch_radial(bdk_t *h, int nbins, double dr, double *part, double *centers, int *hist) {
    const int nb = (int)nblk(h->g.N, OBS_THREADS);
    hipLaunchKernelGGL(k_circ_partial, dim3(nb), dim3(OBS_THREADS), 0, h->stream, h->g, h->d, part);
    hipLaunchKernelGGL(k_circ_final, dim3(1), dim3(OBS_THREADS), 0, h->stream, h->g, part, nb, centers);
    hipLaunchKernelGGL(k_zero_int, dim3(nblk(6 * nbins, 256)), dim3(256), 0, h->stream, hist, 6 * nbins);
    hipLaunchKernelGGL(k_radial_hist, dim3(nb), dim3(OBS_THREADS), 0, h->stream, h->g, h->d, centers, nbins, dr,
                       hist);
    return hipGetLastError();
}

// ------------------------------------------------------------------ snapshot grid
// Local densities on an m x m grid of width w = L/m (main_auswertung.cpp:
// 214-320): each particle adds its TSC weights times 1/w^2 to the nine cells
// around cell (int)(x/w) mod m of its species' grid.  The weight products are
// summed as 2^-40 fixed point in 64-bit integers, so the grid does not depend
// on the order of the additions.  The offset d from the cell centre uses the
// unwrapped cell index (the reference wraps first, which only differs when
// x/w rounds up to m).
constexpr double SNAP_SCALE = 1099511627776.0;  // 2^40

__device__ __forceinline__ void snap_tsc(double x, double w, int m, int &k, double wt[3]) {
    const int k0 = (int)(x / w);
    k = k0 % m;
    const double dd = x - (k0 + 0.5) * w;
    wt[1] = 0.75 - dd * dd / (w * w);
    double t = fabs((dd + w) / w);
    wt[0] = 0.5 * (1.5 - t) * (1.5 - t);
    t = fabs((dd - w) / w);
    wt[2] = 0.5 * (1.5 - t) * (1.5 - t);
}

__global__ __launch_bounds__(256) void k_snap_grid(Geo g, DevBufs d, int m, double w, unsigned long long *acc) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.N) return;
    const int c = d.lcell[q];
    const int gx = div_nc(c, g), gy = c - gx * g.nc;
    const double2 r = d.lrel[q];
    int k, l;
    double wx[3], wy[3];
    snap_tsc(gx * g.B + r.x, w, m, k, wx);
    snap_tsc(gy * g.B + r.y, w, m, l, wy);
    unsigned long long *a = acc + (d.lpid[q] >= g.n1 ? (size_t)m * m : 0);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const size_t row = (size_t)((k - 1 + i + m) % m) * m;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const long long v = __double2ll_rn(wx[i] * wy[j] * SNAP_SCALE);
            atomicAdd(&a[row + (l - 1 + j + m) % m], (unsigned long long)v);
        }
    }
}

// fixed point -> densities; with stats: the histograms of rho1, rho2,
// rho = rho1 + rho2 and rho1 - rho2 and the cells over the thresholds
// (main_auswertung.cpp:324-375), counted per block in LDS.
// cnt: [4][nbins] histograms, then 7 threshold counts.
__device__ __forceinline__ int snap_bin(double t, int nbins) {
    if (!(t > -1.0)) return 0;  // (int) truncates toward zero; below 0 clamps to 0
    if (t >= (double)nbins) return nbins - 1;
    const int b = (int)t;
    return b < 0 ? 0 : (b >= nbins ? nbins - 1 : b);
}

__global__ __launch_bounds__(256) void k_snap_final(int m, double drho, const unsigned long long *acc, double *rho,
                                                    int stats, int nbins, double dbin, SnapThresh th,
                                                    unsigned int *cnt) {
    extern __shared__ unsigned int sh[];  // [4 * nbins + 7]
    const int nsh = 4 * nbins + 7;
    if (stats)
        for (int i = threadIdx.x; i < nsh; i += blockDim.x) sh[i] = 0u;
    __syncthreads();
    const size_t mm = (size_t)m * m;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < mm; i += stride) {
        const double r1 = ((double)acc[i] / SNAP_SCALE) * drho;
        const double r2 = ((double)acc[mm + i] / SNAP_SCALE) * drho;
        rho[i] = r1;
        rho[mm + i] = r2;
        if (!stats) continue;
        const double rs = r1 + r2, dr = r1 - r2;
        for (int t = 0; t < 3; ++t)
            if (rs > th.rho[t]) atomicAdd(&sh[4 * nbins + t], 1u);
        const double f = rs > 0.0 ? fabs(dr) / rs : 0.0;  // 0/0 passes no threshold
        for (int t = 0; t < 4; ++t)
            if (rs > 0.0 && f > th.drho[t]) atomicAdd(&sh[4 * nbins + 3 + t], 1u);
        atomicAdd(&sh[snap_bin(r1 / dbin - 0.5, nbins)], 1u);
        atomicAdd(&sh[nbins + snap_bin(r2 / dbin - 0.5, nbins)], 1u);
        atomicAdd(&sh[2 * nbins + snap_bin(rs / dbin - 0.5, nbins)], 1u);
        atomicAdd(&sh[3 * nbins + snap_bin(dr / dbin - 0.5 + (double)(nbins / 2), nbins)], 1u);
    }
    if (!stats) return;
    __syncthreads();
    for (int i = threadIdx.x; i < nsh; i += blockDim.x)
        if (sh[i]) atomicAdd(&cnt[i], sh[i]);
}

hipError_t launch_snap(bdk_t *h, int m, unsigned long long *acc, double *rho, int stats, int nbins, double dbin,
                       SnapThresh th, unsigned int *cnt) {
    const double w = h->p.L / m;
    hipLaunchKernelGGL(k_snap_grid, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d, m, w, acc);
    const size_t mm = (size_t)m * m;
    const int nb = (int)std::min<size_t>(nblk(mm, 256), 4096);
    const size_t lds = stats ? (size_t)(4 * nbins + 7) * sizeof(unsigned int) : 0;
    hipLaunchKernelGGL(k_snap_final, dim3(nb), dim3(256), lds, h->stream, m, 1.0 / (w * w), acc, rho, stats, nbins,
                       dbin, th, cnt);
    return hipGetLastError();
}

// ------------------------------------------------------------------ direct rho(k)
// record_ftrho_unkorrigiert (au:307-360): for each wave-vector index pair
// (i, j) the sums over one species of cos and sin(dq (x i + y j)).  For
// species 1 the reference takes y = git_y B + rel_x (au:333), unless the
// parameters ask for rel_y (ftrho_fix); species 2 uses rel_y.  Block (chunk,
// pair): a fixed-order tree sum over the chunk's particles into
// partial[pair][chunk]; k_ftrho_final adds the chunks in order.
constexpr int FT_THREADS = 256, FT_PER_THREAD = 16, FT_MAX_CHUNKS = 64;

// chunk `chunk` of nchunk covers list positions [chunk per, (chunk + 1) per);
// per is a multiple of FT_THREADS, and nchunk is at most FT_MAX_CHUNKS, so
// the partial sums stay small at any N.  fix: species 1 takes rel_y.
__global__ __launch_bounds__(FT_THREADS) void k_ftrho(Geo g, DevBufs d, const int2 *ij, int species, int fix,
                                                      double dq, double2 *partial, int nchunk, int per) {
    __shared__ double sc[FT_THREADS], ss[FT_THREADS];
    const int p = blockIdx.y, chunk = blockIdx.x, tid = threadIdx.x;
    const int2 q = ij[p];
    double c = 0.0, s = 0.0;
    const int q0 = chunk * per, q1 = min(q0 + per, g.N);
    for (int t = q0 + tid; t < q1; t += FT_THREADS) {
        const bool s2 = d.lpid[t] >= g.n1;
        if (s2 != (species == 2)) continue;
        const int cell = d.lcell[t];
        const int gx = div_nc(cell, g), gy = cell - gx * g.nc;
        const double2 r = d.lrel[t];
        const double x = gx * g.B + r.x;
        const double y = gy * g.B + ((species == 1 && !fix) ? r.x : r.y);
        double sn, cs;
        sincos(dq * (x * q.x + y * q.y), &sn, &cs);
        c += cs;
        s += sn;
    }
    sc[tid] = c;
    ss[tid] = s;
    __syncthreads();
    for (int w = FT_THREADS / 2; w > 0; w >>= 1) {
        if (tid < w) {
            sc[tid] += sc[tid + w];
            ss[tid] += ss[tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) partial[(size_t)p * nchunk + chunk] = make_double2(sc[0], ss[0]);
}

__global__ void k_ftrho_final(const double2 *partial, int nchunk, int npairs, double *re, double *im) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    double c = 0.0, s = 0.0;
    for (int k = 0; k < nchunk; ++k) {
        const double2 v = partial[(size_t)p * nchunk + k];
        c += v.x;
        s += v.y;
    }
    re[p] = c;
    im[p] = s;
}

// chunks of at least FT_THREADS * FT_PER_THREAD positions, at most FT_MAX_CHUNKS
static void ftrho_split(const bdk_t *h, int &nchunk, int &per) {
    nchunk = std::min((int)nblk(h->g.N, FT_THREADS * FT_PER_THREAD), FT_MAX_CHUNKS);
    per = (int)nblk(h->g.N, nchunk * FT_THREADS) * FT_THREADS;
    nchunk = std::max(1, (int)nblk(h->g.N, per));
}

hipError_t launch_ftrho(bdk_t *h, const int2 *ij, int npairs, int species, double dq, double2 *partial, double *re,
                        double *im) {
    int nchunk, per;
    ftrho_split(h, nchunk, per);
    hipLaunchKernelGGL(k_ftrho, dim3(nchunk, npairs), dim3(FT_THREADS), 0, h->stream, h->g, h->d, ij, species,
                       h->p.ftrho_fix ? 1 : 0, dq, partial, nchunk, per);
    hipLaunchKernelGGL(k_ftrho_final, dim3(nblk(npairs, 256)), dim3(256), 0, h->stream, partial, nchunk, npairs, re,
                       im);
    return hipGetLastError();
}

int ftrho_chunks(const bdk_t *h) {
    int nchunk, per;
    ftrho_split(h, nchunk, per);
    return nchunk;
}

// ------------------------------------------------------------------ step counters
// TimestepInfo (signaturen2.h:71-145, filled by zeitschritt_debug,
// dm:168-211): the largest force of each kind by |F|^2 with its particle,
// the number of candidate pairs the cell lists present, and the particles
// that changed cell.  Maxima are 64-bit keys (|F|^2 as FP32 bits, list
// position) under atomicMax, so the result does not depend on the order.
template <typename R>
__global__ __launch_bounds__(256) void k_info_forces(Geo g, DevBufs d, StepInfoDev *info) {
    using R2 = typename V2<R>::T;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long cand = 0;
    if (q < g.N) {
        const R2 fw = ((const R2 *)d.fwca)[q], fk = ((const R2 *)d.fkap)[q];
        const double2 n = particle_noise(g, d, d.lpid[q], d.sc->step);
        const double v[3] = {(double)fw.x * fw.x + (double)fw.y * fw.y, (double)fk.x * fk.x + (double)fk.y * fk.y,
                             n.x * n.x + n.y * n.y};
        for (int k = 0; k < 3; ++k) {
            const unsigned long long key = ((unsigned long long)__float_as_uint((float)v[k]) << 32) | (unsigned)q;
            atomicMax(&info->key[k], key);
        }
        // as the reference counts them: a species-1 particle against both
        // lists, a species-2 particle against list 2 (the 12 pairs once)
        const bool s2 = d.lpid[q] >= g.n1;
        const int c = d.lcell[q], cx = div_nc(c, g), cy = c - cx * g.nc;
        for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b) {
                const int cc = wrap1(cx + a, g.nc) * g.nc + wrap1(cy + b, g.nc);
                if (!s2) {
                    cand += d.start[cc + 1] - d.start[cc];
                } else {
                    for (int j = d.start[cc]; j < d.start[cc + 1]; ++j)
                        cand += ((const unsigned *)d.list)[j] >> 31;
                }
            }
        cand -= 1;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) cand += __shfl_xor(cand, o, 64);
    if ((threadIdx.x & 63) == 0 && cand) atomicAdd(&info->num_wca, cand);
}

// the vectors and particles of the maxima (one thread), before the update
template <typename R>
__global__ void k_info_fetch(Geo g, DevBufs d, StepInfoDev *info) {
    using R2 = typename V2<R>::T;
    for (int k = 0; k < 3; ++k) {
        const int q = (int)(info->key[k] & 0xffffffffull);
        info->pid[k] = d.lpid[q];
        if (k == 0) {
            const R2 f = ((const R2 *)d.fwca)[q];
            info->wca[0] = f.x;
            info->wca[1] = f.y;
            const int c = d.lcell[q], cx = div_nc(c, g), cy = c - cx * g.nc;
            int n1 = 0, n2 = 0;
            for (int a = -1; a <= 1; ++a)
                for (int b = -1; b <= 1; ++b) {
                    const int cc = wrap1(cx + a, g.nc) * g.nc + wrap1(cy + b, g.nc);
                    for (int j = d.start[cc]; j < d.start[cc + 1]; ++j) {
                        if (((const unsigned *)d.list)[j] >> 31) ++n2;
                        else ++n1;
                    }
                }
            info->nlist[0] = n1;
            info->nlist[1] = n2;
        } else if (k == 1) {
            const R2 f = ((const R2 *)d.fkap)[q];
            info->kap[0] = f.x;
            info->kap[1] = f.y;
        } else {
            const double2 n = particle_noise(g, d, d.lpid[q], d.sc->step);
            info->rnd[0] = n.x;
            info->rnd[1] = n.y;
        }
    }
}

// after the update, before the list rebuild: cell changes per species
__global__ void k_info_moves(Geo g, DevBufs d, StepInfoDev *info) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.N) return;
    if (d.cell[q] != d.lcell[q]) atomicAdd(&info->moves[d.lpid[q] >= g.n1 ? 1 : 0], 1);
}

// Event marks for per-kernel timing: with ev != nullptr an event is
// recorded on the main stream after every launch and everything runs on the
// main stream in order; names receives the labels.
struct Marks {
    bdk_t *h;
    std::vector<std::string> *names;
    std::vector<hipEvent_t> *ev;
    hipError_t e = hipSuccess;
    void chk(hipError_t r) {
        if (e == hipSuccess && r != hipSuccess) e = r;
    }
    void mark(const char *nm) {
        if (names) names->push_back(nm);
        if (ev) {
            hipEvent_t x;
            if (hipEventCreate(&x) == hipSuccess) {
                chk(hipEventRecord(x, h->stream));
                ev->push_back(x);
            }
        }
    }
};

// Forces of the current state.  The WCA forces and the mesh chain (deposit
// from the list order, FFTs) only read the particle state: the WCA runs on
// the second stream beside the chain (fork and join by events).
hipError_t enqueue_forces(bdk_t *h, std::vector<std::string> *names, std::vector<hipEvent_t> *ev) {
    Marks m{h, names, ev};
    if (ev) m.mark("begin");
    if (h->g.idgas) return hipSuccess;  // no forces
    const bool fork = !ev && h->aux;
    if (fork) {
        m.chk(hipEventRecord(h->ev_fork, h->stream));
        m.chk(hipStreamWaitEvent(h->aux, h->ev_fork, 0));
        m.chk(launch_wca(h, h->aux));
        m.chk(hipEventRecord(h->ev_join, h->aux));
    } else {
        m.chk(launch_wca(h));
        m.mark("k_wca");
    }
    m.chk(launch_deposit(h, 1));
    m.mark("k_deposit");
    for (int p = 0; p < 3; ++p) {
        m.chk(launch_fft_pass(h, p));
        m.mark(p == 0 ? "k_r2c_rows" : p == 1 ? "k_col" : "k_c2r_rows");
    }
    if (fork) m.chk(hipStreamWaitEvent(h->stream, h->ev_join, 0));
    m.chk(hipGetLastError());
    return m.e;
}

// One step from a state whose forces are ready: the update (gather, noise,
// Euler step, refresh), then the two halves that only need its output side
// by side:
//   main stream: the list rebuild (rank, scan, fill, sort), then the WCA
//                forces of the new lists;
//   second stream: the deposit from the slot arrays (grouped by the previous
//                starts, which the update kept; halo widened by the
//                update's displacement bound), the FFT passes, the last of
//                which ends the step (counter, per-step scalars);
// joined.  The new state leaves with its forces ready.  Without forces (an
// ideal gas) the scan ends the step.
hipError_t enqueue_step(bdk_t *h, std::vector<std::string> *names, std::vector<hipEvent_t> *ev,
                        StepInfoDev *info) {
    Marks m{h, names, ev};
    if (ev) m.mark("begin");
    const bool kap = !h->g.idgas;
    const int ncell = h->g.ncell, nb = (int)nblk(ncell, SCAN_BLOCK);
    if (!kap) {
        m.chk(launch_rearm(h));
        m.mark("k_rearm");
    }
    if (info) {
        // capillary forces on their own (the step fuses them into the update)
        if (kap && !h->g.adaptive) m.chk(launch_gather(h, 0, 0));
        const bool f32 = h->p.precision != BDK_FP64;
        if (f32) {
            hipLaunchKernelGGL(k_info_forces<float>, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d, info);
            hipLaunchKernelGGL(k_info_fetch<float>, dim3(1), dim3(1), 0, h->stream, h->g, h->d, info);
        } else {
            hipLaunchKernelGGL(k_info_forces<double>, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d, info);
            hipLaunchKernelGGL(k_info_fetch<double>, dim3(1), dim3(1), 0, h->stream, h->g, h->d, info);
        }
    }
    if (h->g.adaptive) {
        m.chk(launch_gather(h, 1, 0));  // the update reduces the candidates itself
        m.mark("k_gather");
        m.chk(launch_update(h, 0));
    } else {
        m.chk(launch_update(h, kap ? 1 : 0));
    }
    m.mark("k_update");
    if (info) hipLaunchKernelGGL(k_info_moves, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d, info);
    const bool fork = kap && !ev && h->aux;
    hipStream_t side = fork ? h->aux : h->stream;
    if (kap) {
        if (fork) {
            m.chk(hipEventRecord(h->ev_fork, h->stream));
            m.chk(hipStreamWaitEvent(h->aux, h->ev_fork, 0));
        }
        m.chk(launch_deposit(h, 0, 1, side));
        m.mark("k_deposit");
        for (int p = 0; p < 3; ++p) {
            m.chk(launch_fft_pass(h, p, side, p == 2));
            m.mark(p == 0 ? "k_r2c_rows" : p == 1 ? "k_col" : "k_c2r_rows");
        }
        if (fork) m.chk(hipEventRecord(h->ev_join, side));
    }
    // bcount is current (the update adjusted it with the counts); the ranks
    // read the previous starts that the update kept
    const int nrank = (int)nblk(h->g.N, 256);
    hipLaunchKernelGGL(k_rank_scan, dim3(nrank + nb), dim3(256), 0, h->stream, h->g, h->d, nrank, nb, !kap);
    m.mark("k_rank_scan");
    hipLaunchKernelGGL(k_fill, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d, 1);
    m.mark("k_fill");
    hipLaunchKernelGGL(k_sort_cells, dim3(nblk(ncell, 256)), dim3(256), 0, h->stream, h->g, h->d, 1);
    m.mark("k_sort_cells");
    if (kap) {
        m.chk(launch_wca(h));
        m.mark("k_wca");
    }
    if (fork) m.chk(hipStreamWaitEvent(h->stream, h->ev_join, 0));
    m.chk(hipGetLastError());
    return m.e;
}

// ------------------------------------------------------------------ batch plans

int batch_min_ept(int stage, int zmax) {
    const int need = stage == 0 ? zmax / 2 : (stage == 1 ? 2 * zmax : zmax);
    return ept_for(8, need);
}

hipError_t batch_build(const std::vector<bdk_t *> &mem, const BatchTuning &tune, BatchPlan &bp,
                       std::vector<RepDev> &reps, std::vector<int2> maps[BS_COUNT]) {
    const int n = (int)mem.size();
    reps.assign(n, RepDev());
    for (int k = 0; k < BS_COUNT; ++k) {
        maps[k].clear();
        bp.nblk2[k] = 0;
        bp.nblk16[k] = 0;
    }
    bp.nmem = n;
    bp.f32 = mem.empty() ? 1 : (mem[0]->p.precision != BDK_FP64);
    bp.tune = tune;
    int zmax = 8;
    for (const bdk_t *h : mem) zmax = std::max(zmax, h->g.Z);
    for (int k = 0; k < 3; ++k) bp.ept[k] = std::max(tune.ept[k], batch_min_ept(k, zmax));
    const int *ept = bp.ept;
    bp.dep_lds = bp.l1 = bp.lc = bp.l2 = 0;
    const size_t csz = bp.f32 ? 8 : 16;
    // entries of the members with a power-of-two mesh first, those with
    // radix-16 column and c2r passes (FP32) before the others: the mesh
    // stages launch once per mesh kind present (the kind is a template
    // parameter)
    auto cls = [&](const bdk_t *h) { return bp.f32 && fft_r16(h->g.Z) ? 0 : (zp2(h->g.Z) ? 1 : 2); };
    std::vector<int> order;
    for (int pass = 0; pass < 3; ++pass)
        for (int i = 0; i < n; ++i)
            if (cls(mem[i]) == pass) order.push_back(i);
    for (int oi = 0; oi < n; ++oi) {
        const int i = order[oi];
        const bdk_t *h = mem[i];
        RepDev &r = reps[i];
        r.g = h->g;
        r.d = h->d;
        const DepPlan pl = dep_plan(h, 1, tune.dep_tile);
        r.dep = pl.a;
        r.dep_tx = (int)pl.grid.x;
        bp.dep_lds = std::max(bp.dep_lds, pl.lds);
        const int Z = h->g.Z;
        fft_rows(Z, ept, r.rb, r.rc, r.rb2);
        r.sR1 = lds_row(Z / 2);
        r.sC = lds_row(Z);
        r.sR2 = lds_row(Z);
        bp.l1 = std::max(bp.l1, (size_t)r.rb * r.sR1 * csz);
        bp.lc = std::max(bp.lc, (size_t)2 * r.rc * r.sC * csz);
        bp.l2 = std::max(bp.l2, (size_t)r.rb2 * r.sR2 * csz);
        r.scan_nb = (int)nblk(h->g.ncell, SCAN_BLOCK);
        r.unit = h->p.sigma11_22 == 1.0 && h->p.sigma11_12 == 1.0 && h->p.eps22_11 == 1.0 && h->p.eps12_11 == 1.0;
        const int ad = h->g.adaptive ? 1 : 0;
        const int nb[BS_COUNT] = {(int)nblk(h->g.N, 256), (int)nblk(h->g.ncell, 256), r.scan_nb,
                                  (int)(pl.grid.x * pl.grid.y), Z / r.rb, (Z / 2) / r.rc, Z / r.rb2,
                                  ad * (int)nblk(h->g.N, 256), (int)nblk(h->g.N, tune.wca_block)};
        for (int k = 0; k < BS_COUNT; ++k)
            for (int b = 0; b < nb[k]; ++b) maps[k].push_back(make_int2(i, b));
        if (zp2(h->g.Z))
            for (int k = 0; k < BS_COUNT; ++k) bp.nblk2[k] = (int)maps[k].size();
        if (cls(h) == 0)
            for (int k = 0; k < BS_COUNT; ++k) bp.nblk16[k] = (int)maps[k].size();
    }
    for (int k = 0; k < BS_COUNT; ++k) bp.nblk[k] = (int)maps[k].size();
    return hipSuccess;
}

#define BDK_EPT_SWITCH2(E, CALL) \
    switch (E) {                 \
    case 8: CALL(8); break;      \
    case 16: CALL(16); break;    \
    default: CALL(32); break;    \
    }

// blocks of a launch over a map of n entries (the persistent form caps them)
static unsigned stage_grid(const BatchPlan &bp, int n) {
    const int p = bp.tune.persist;
    return (unsigned)std::max(1, p > 0 ? std::min(n, p) : n);
}

template <typename R, bool P> static void batch_stage_tp(const BatchPlan &bp, int stage, hipStream_t st) {
    const RepDev *m = bp.reps;
    // f(kind, map, entries) once per mesh kind present in stage S's map
    auto kinds = [&](int S, auto &&f) {
        const int n2 = bp.nblk2[S], nall = bp.nblk[S];
        if (n2 > 0) f(std::true_type(), (const int2 *)bp.map[S], n2);
        if (nall > n2) f(std::false_type(), (const int2 *)bp.map[S] + n2, nall - n2);
    };
    // the same for the column and c2r passes, with the radix-16 members
    // (leading) apart: f(kind, radix 16, map, entries)
    auto kinds16 = [&](int S, auto &&f) {
        const int n16 = std::is_same<R, float>::value ? bp.nblk16[S] : 0, n2 = bp.nblk2[S], nall = bp.nblk[S];
        if (n16 > 0) f(std::true_type(), std::true_type(), (const int2 *)bp.map[S], n16);
        if (n2 > n16) f(std::true_type(), std::false_type(), (const int2 *)bp.map[S] + n16, n2 - n16);
        if (nall > n2) f(std::false_type(), std::false_type(), (const int2 *)bp.map[S] + n2, nall - n2);
    };
#define BDK_G(S) dim3(stage_grid(bp, bp.nblk[S]))
    switch (stage) {
    case 0:
        hipLaunchKernelGGL((kb_wca<R, P>), BDK_G(BS_WCA), dim3(bp.tune.wca_block), 0, st, m, bp.map[BS_WCA],
                           bp.nblk[BS_WCA]);
        break;
    case 1:  // from the list order
    case 8:  // from the slot arrays, beside the rebuild
        kinds(BS_DEP, [&](auto q, const int2 *mp, int nm) {
            constexpr bool Q = decltype(q)::value;
            hipLaunchKernelGGL((kb_deposit<R, P, Q>), dim3(stage_grid(bp, nm)), dim3(DP_THREADS), bp.dep_lds, st, m,
                               mp, nm, stage == 8 ? 1 : 0);
        });
        break;
    case 2:
        kinds16(BS_R2C, [&](auto q, auto r16, const int2 *mp, int nm) {
            constexpr bool Q = decltype(q)::value;
            constexpr bool R16 = decltype(r16)::value && std::is_same<R, float>::value;
#define R2C(E)                                                                                                 \
    hipLaunchKernelGGL((kb_r2c<R, E, P, Q, R16>), dim3(stage_grid(bp, nm)), dim3(FFT_THREADS), bp.l1, st, m, mp, \
                       nm)
            BDK_EPT_SWITCH2(bp.ept[0], R2C)
#undef R2C
        });
        break;
    case 3:
        kinds16(BS_COL, [&](auto q, auto r16, const int2 *mp, int nm) {
            constexpr bool Q = decltype(q)::value;
            constexpr bool R16 = decltype(r16)::value && std::is_same<R, float>::value;
#define COL(E)                                                                                                 \
    hipLaunchKernelGGL((kb_col<R, E, P, Q, R16>), dim3(stage_grid(bp, nm)), dim3(FFT_THREADS), bp.lc, st, m, mp, \
                       nm)
            BDK_EPT_SWITCH2(bp.ept[1], COL)
#undef COL
        });
        break;
    case 4:
    case 12:  // 12: also the end of the step
        kinds16(BS_C2R, [&](auto q, auto r16, const int2 *mp, int nm) {
            constexpr bool Q = decltype(q)::value;
            constexpr bool R16 = decltype(r16)::value && std::is_same<R, float>::value;
#define C2R(E)                                                                                                  \
    hipLaunchKernelGGL((kb_c2r<R, E, P, Q, R16>), dim3(stage_grid(bp, nm)), dim3(FFT_THREADS), bp.l2, st, m, mp, \
                       nm, stage == 12 ? 1 : 0)
            BDK_EPT_SWITCH2(bp.ept[2], C2R)
#undef C2R
        });
        break;
    case 5:
        kinds(BS_PART, [&](auto q, const int2 *mp, int nm) {
            constexpr bool Q = decltype(q)::value;
            hipLaunchKernelGGL((kb_update<R, P, Q>), dim3(stage_grid(bp, nm)), dim3(256), 0, st, m, mp, nm);
        });
        break;
    case 13:  // adaptive members: forces and dt candidates (their updates reduce them)
        if (bp.nblk[BS_GATHER] > 0) {
            kinds(BS_GATHER, [&](auto q, const int2 *mp, int nm) {
                constexpr bool Q = decltype(q)::value;
                hipLaunchKernelGGL((kb_gather<R, P, Q>), dim3(stage_grid(bp, nm)), dim3(256), 0, st, m, mp, nm);
            });
        }
        break;
    case 6:
        hipLaunchKernelGGL(kb_rank_scan<P>, dim3(stage_grid(bp, bp.nblk[BS_PART] + bp.nblk[BS_SCAN])), dim3(256), 0, st,
                           m, bp.map[BS_PART], bp.nblk[BS_PART], bp.map[BS_SCAN], bp.nblk[BS_SCAN]);
        break;
    default: {  // 11: fill, sort
        hipLaunchKernelGGL(kb_fill<P>, BDK_G(BS_PART), dim3(256), 0, st, m, bp.map[BS_PART], bp.nblk[BS_PART]);
        hipLaunchKernelGGL(kb_sort_cells<P>, BDK_G(BS_CELL), dim3(256), 0, st, m, bp.map[BS_CELL], bp.nblk[BS_CELL]);
        break;
    }
    }
#undef BDK_G
}

template <typename R> static void batch_stage_t(const BatchPlan &bp, int stage, hipStream_t st) {
    if (bp.tune.persist > 0)
        batch_stage_tp<R, true>(bp, stage, st);
    else
        batch_stage_tp<R, false>(bp, stage, st);
}

// stage 0 WCA, 1 deposit (list order), 2 r2c, 3 col, 4 c2r, 5 update,
// 6 rank + scan, 8 deposit (slot arrays), 11 fill + sort, 12 c2r + step end,
// 13 gather (forces and dt candidates) of the adaptive members
hipError_t batch_stage(const BatchPlan &bp, int stage, hipStream_t st) {
    if (bp.f32)
        batch_stage_t<float>(bp, stage, st);
    else
        batch_stage_t<double>(bp, stage, st);
    return hipGetLastError();
}

// The members' forces: the WCA stage on aux beside the mesh stages.
hipError_t batch_enqueue_forces(const BatchPlan &bp, hipStream_t st, hipStream_t aux, hipEvent_t fork,
                                hipEvent_t join) {
    hipError_t e = hipSuccess;
    if (aux) {
        e = hipEventRecord(fork, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(aux, fork, 0);
        if (e == hipSuccess) e = batch_stage(bp, 0, aux);
        if (e == hipSuccess) e = hipEventRecord(join, aux);
    } else {
        e = batch_stage(bp, 0, st);
    }
    for (int k = 1; k <= 4 && e == hipSuccess; ++k) e = batch_stage(bp, k, st);
    if (aux && e == hipSuccess) e = hipStreamWaitEvent(st, join, 0);
    return e;
}

// One batched step, laid out like the single-simulation step: (the adaptive
// members' gather,) update, then
// the rebuild and the WCA on st beside the slot-order deposit and the FFT
// stages on aux (the last of which ends the step), joined.
hipError_t batch_enqueue_step(const BatchPlan &bp, hipStream_t st, hipStream_t aux, hipEvent_t fork,
                              hipEvent_t join) {
    hipError_t e = bp.nblk[BS_GATHER] > 0 ? batch_stage(bp, 13, st) : hipSuccess;
    if (e == hipSuccess) e = batch_stage(bp, 5, st);
    hipStream_t side = aux ? aux : st;
    if (aux && e == hipSuccess) e = hipEventRecord(fork, st);
    if (aux && e == hipSuccess) e = hipStreamWaitEvent(aux, fork, 0);
    if (e == hipSuccess) e = batch_stage(bp, 8, side);
    for (int k = 2; k <= 3 && e == hipSuccess; ++k) e = batch_stage(bp, k, side);
    if (e == hipSuccess) e = batch_stage(bp, 12, side);
    if (aux && e == hipSuccess) e = hipEventRecord(join, side);
    if (e == hipSuccess) e = batch_stage(bp, 6, st);
    if (e == hipSuccess) e = batch_stage(bp, 11, st);
    if (e == hipSuccess) e = batch_stage(bp, 0, st);
    if (aux && e == hipSuccess) e = hipStreamWaitEvent(st, join, 0);
    return e;
}

}  // namespace bdkx
