// HIP kernels of the MI355X Brownian-dynamics timestep (gfx950).
//
// One step (dynamik_methoden.cpp:74-163) is this launch sequence:
//   k_force     WCA forces (dm:617-1059) and the charge deposit (gr:31-547),
//               per tile of 16x16 list cells, particles staged in LDS;
//   k_r2c_rows  row FFTs of the real mesh, transposed half-spectrum out;
//   k_col       column FFT, Green's function x gradient (dm:1210-1216), two
//               inverse column FFTs (Fx, Fy);
//   k_c2r_rows  inverse row FFT of (Fx + i Fy) -> real Fx, Fy mesh;
//   k_update    gather (gr:81-668), noise (dm:1317-1339), Euler step
//               (dm:116-157), refresh (dm:216-325), cell counting;
//   k_scan_*    exclusive scan of the cell counts;
//   k_fill      cell list = slots grouped by cell.
// With a finite max_travel, k_gather (forces + dt min-reduction, dm:1420-1491)
// runs before a non-fused k_update.
//
// Position arithmetic (x = git*B + rel, mesh indices, weights, refresh) is
// compiled without FP contraction so it rounds like the reference's x86-64
// build; cell indices are therefore bit-exact.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "bdk_internal.hpp"

namespace bdkx {

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };
template <typename R> struct V4;
template <> struct V4<float> { using T = float4; };
template <> struct V4<double> { using T = double4; };

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
    return cm<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
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
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]), lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]), lo1 = 0xCD9E8D57u * c[2];
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
    double2 o;
    o.x = r * cos(a);
    o.y = r * sin(a);
    return o;
}

// ------------------------------------------------------------------ geometry
__device__ __forceinline__ int wrapc(int u, int n) {
    int v = u % n;
    return v < 0 ? v + n : v;
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

// ------------------------------------------------------------------ k_force
// WCA forces and charge deposit for one tile of FT_TC x FT_TC list cells.
// The tile's cells plus a halo of hx cells are staged in LDS, particles of
// each cell in ascending particle id, so every sum has a fixed order and
// the result is deterministic (no floating-point atomics).  Mesh points are
// owned by the tile whose cells contain their centres; each owned point
// sums the contributions of the staged particles whose 3x3 stencil covers
// it (gather form of the deposit).
constexpr int FT_TC = 16;
constexpr int FT_HX = 2;
constexpr int FT_S = FT_TC + 2 * FT_HX;
constexpr int FT_NCS = FT_S * FT_S;
constexpr int FT_MAXP = 1024;
constexpr int FT_THREADS = 256;

template <typename R> struct ForceLDS {
    int cstart[FT_NCS + 1];
    int cwrap[FT_NCS];
    int wsum[16];
    int info[8];
    double x[FT_MAXP], y[FT_MAXP];
    int slot[FT_MAXP];
    int pid[FT_MAXP];
    int kp[FT_MAXP], lp[FT_MAXP];
    R wx[FT_MAXP][3], wy[FT_MAXP][3];
    unsigned short pcell[FT_MAXP];
};

// first mesh index whose centre lies in list cell >= c: owner(k) = floor((k+1/2)h/B)
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

// ascending-pid walk over the particles of global cell c (fallback path)
__device__ __forceinline__ int next_in_cell(const int *list, const int *pid, int beg, int end, int last_pid,
                                            int &slot_out) {
    int best = 0x7fffffff, bs = -1;
    for (int q = beg; q < end; ++q) {
        int s = list[q];
        int p = pid[s];
        if (p > last_pid && p < best) {
            best = p;
            bs = s;
        }
    }
    slot_out = bs;
    return best;
}

template <typename R, int SCHEME>
__device__ void force_tile_direct(const Geo &g, const DevBufs &d, int cx0, int cy0, int nx, int ny, int klo,
                                  int khi, int llo, int lhi, double e, int what) {
    using R2 = typename V2<R>::T;
    const int nc = g.nc, tid = threadIdx.x, nt = blockDim.x;
    R2 *fw = (R2 *)d.fwca;
    R *rho = (R *)d.rho;
    if (what & 1) {
        for (int ci = tid; ci < nx * ny; ci += nt) {
            int cx = cx0 + ci / ny, cy = cy0 + ci % ny;
            int c = cx * nc + cy;
            for (int q = d.start[c]; q < d.start[c + 1]; ++q) {
                int si = d.list[q];
                int pi = d.pid[si];
                double2 ri = d.rel[si];
                double xi = cx * g.B + ri.x, yi = cy * g.B + ri.y;
                bool s2i = pi >= g.n1;
                R fxs = 0, fys = 0;
                for (int pass = 0; pass < 2; ++pass)
                    for (int da = -1; da <= 1; ++da)
                        for (int db = -1; db <= 1; ++db) {
                            int ux = wrapc(cx + da, nc), uy = wrapc(cy + db, nc);
                            int cc = ux * nc + uy;
                            int beg = d.start[cc], end = d.start[cc + 1];
                            int last = -1, sj;
                            while (true) {
                                int pj = next_in_cell(d.list, d.pid, beg, end, last, sj);
                                if (sj < 0) break;
                                last = pj;
                                if (sj == si) continue;
                                bool s2j = pj >= g.n1;
                                if ((pass == 0) != (s2i == s2j)) continue;
                                double2 rj = d.rel[sj];
                                double xj = ux * g.B + rj.x, yj = uy * g.B + rj.y;
                                R fx, fy;
                                if (!s2i && !s2j) {
                                    if (wca_pair<R>(xi, yi, xj, yj, 0, g, fx, fy)) { fxs += fx; fys += fy; }
                                } else if (s2i && s2j) {
                                    if (wca_pair<R>(xi, yi, xj, yj, 1, g, fx, fy)) { fxs += fx; fys += fy; }
                                } else if (!s2i) {
                                    if (wca_pair<R>(xi, yi, xj, yj, 2, g, fx, fy)) { fxs += fx; fys += fy; }
                                } else {
                                    if (wca_pair<R>(xj, yj, xi, yi, 2, g, fx, fy)) { fxs -= fx; fys -= fy; }
                                }
                            }
                        }
                R2 o;
                o.x = fxs;
                o.y = fys;
                fw[si] = o;
            }
        }
    }
    if (what & 2) {
        const int Z = g.Z, nlx = lhi - llo;
        for (int pi = tid; pi < (khi - klo) * nlx; pi += nt) {
            int k = klo + pi / nlx, l = llo + pi % nlx;
            int culo = (int)ceil(((k - 1) * g.h - e) / g.B) - 1, cuhi = (int)ceil(((k + 2) * g.h + e) / g.B) - 1;
            int cvlo = (int)ceil(((l - 1) * g.h - e) / g.B) - 1, cvhi = (int)ceil(((l + 2) * g.h + e) / g.B) - 1;
            if (cuhi - culo + 1 > nc) cuhi = culo + nc - 1;
            if (cvhi - cvlo + 1 > nc) cvhi = cvlo + nc - 1;
            R acc = 0;
            for (int cu = culo; cu <= cuhi; ++cu)
                for (int cv = cvlo; cv <= cvhi; ++cv) {
                    int ux = wrapc(cu, nc), uy = wrapc(cv, nc);
                    int cc = ux * nc + uy;
                    int beg = d.start[cc], end = d.start[cc + 1];
                    int last = -1, sj;
                    while (true) {
                        int pj = next_in_cell(d.list, d.pid, beg, end, last, sj);
                        if (sj < 0) break;
                        last = pj;
                        double2 rj = d.rel[sj];
                        double x = ux * g.B + rj.x, y = uy * g.B + rj.y;
                        int kp, lp;
                        double wx[3], wy[3];
                        weights1d<SCHEME>(x, g.h, kp, wx);
                        weights1d<SCHEME>(y, g.h, lp, wy);
                        int dk = wrapc(k - kp + Z / 2, Z) - Z / 2, dl = wrapc(l - lp + Z / 2, Z) - Z / 2;
                        if (dk < -1 || dk > 1 || dl < -1 || dl > 1) continue;
                        if (SCHEME == 0) {
                            if (dk != 0 || dl != 0) continue;  // NGP: the particle's own mesh cell only
                            if (pj < g.n1) acc += (R)g.drho;
                            else acc += (R)(g.drho * g.f2_f1);
                        } else {
                            R c = (R)wx[dk + 1] * (R)wy[dl + 1] * (R)g.drho;
                            if (pj >= g.n1) c = c * (R)g.f2_f1;
                            acc += c;
                        }
                    }
                }
            rho[(size_t)k * Z + l] = acc;
        }
    }
}

template <typename R, int SCHEME>
__global__ __launch_bounds__(FT_THREADS) void k_force(Geo g, DevBufs d, int what) {
    using R2 = typename V2<R>::T;
    __shared__ ForceLDS<R> S;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nc = g.nc, Z = g.Z;
    const int cx0 = blockIdx.x * FT_TC, cy0 = blockIdx.y * FT_TC;
    const int nx = min(FT_TC, nc - cx0), ny = min(FT_TC, nc - cy0);
    const unsigned long long step = d.sc->step;
    // Excursion bound of this step's positions (written by the previous
    // update); the other parity slot is cleared for the coming update, and
    // the dt minimum is re-armed.  Both are read only by later kernels.
    const double e = dbits(d.sc->exc_bits[step & 1]) + 1e-9 * g.L;
    if (blockIdx.x == 0 && blockIdx.y == 0 && tid == 0 && (what & 4)) {
        d.sc->dt_bits = bitsd(g.dt_max);
        d.sc->exc_bits[(step + 1) & 1] = 0ull;
    }
    int hx = (int)floor((1.5 * g.h + e) / g.B) + 1;
    if (hx < 1) hx = 1;
    if (tid == 0) {
        S.info[0] = first_owned(cx0, g);
        S.info[1] = first_owned(cx0 + nx, g);
        S.info[2] = first_owned(cy0, g);
        S.info[3] = first_owned(cy0 + ny, g);
    }
    const int SX = nx + 2 * hx, SY = ny + 2 * hx;
    bool direct = hx > FT_HX;
    if (!direct) {
        for (int idx = tid; idx < SX * SY; idx += nt) {
            int a = idx / SY, b = idx - a * SY;
            int c = wrapc(cx0 - hx + a, nc) * nc + wrapc(cy0 - hx + b, nc);
            S.cwrap[idx] = c;
            S.cstart[idx] = d.start[c + 1] - d.start[c];
        }
        __syncthreads();
        int total = block_exscan4(S.cstart, SX * SY, S.wsum);
        if (tid == 0) S.cstart[SX * SY] = total;
        direct = total > FT_MAXP;
    }
    __syncthreads();
    const int klo = S.info[0], khi = S.info[1], llo = S.info[2], lhi = S.info[3];
    if (direct) {
        if (tid == 0) atomicAdd(&d.sc->fallback, 1);
        force_tile_direct<R, SCHEME>(g, d, cx0, cy0, nx, ny, klo, khi, llo, lhi, e, what);
        return;
    }
    // stage: one thread per staged cell, particles sorted by pid
    for (int idx = tid; idx < SX * SY; idx += nt) {
        const int c = S.cwrap[idx];
        const int gbeg = d.start[c], n = d.start[c + 1] - gbeg;
        const int base = S.cstart[idx];
        for (int q = 0; q < n; ++q) {
            int s = d.list[gbeg + q];
            int p = d.pid[s];
            int r = q - 1;
            while (r >= 0 && S.pid[base + r] > p) {
                S.pid[base + r + 1] = S.pid[base + r];
                S.slot[base + r + 1] = S.slot[base + r];
                --r;
            }
            S.pid[base + r + 1] = p;
            S.slot[base + r + 1] = s;
        }
        const int gx = c / nc, gy = c - gx * nc;
        for (int q = 0; q < n; ++q) {
            const int t = base + q;
            double2 rr = d.rel[S.slot[t]];
            double x = gx * g.B + rr.x, y = gy * g.B + rr.y;
            S.x[t] = x;
            S.y[t] = y;
            S.pcell[t] = (unsigned short)idx;
            if (what & 2) {
                int kp, lp;
                double wx[3], wy[3];
                weights1d<SCHEME>(x, g.h, kp, wx);
                weights1d<SCHEME>(y, g.h, lp, wy);
                S.kp[t] = wrapc(kp, Z);
                S.lp[t] = wrapc(lp, Z);
#pragma unroll
                for (int u = 0; u < 3; ++u) {
                    S.wx[t][u] = (R)wx[u];
                    S.wy[t][u] = (R)wy[u];
                }
            }
        }
    }
    __syncthreads();
    const int total = S.cstart[SX * SY];
    if (what & 1) {
        R2 *fw = (R2 *)d.fwca;
        for (int t = tid; t < total; t += nt) {
            const int idx = S.pcell[t];
            const int a = idx / SY, b = idx - a * SY;
            if (a < hx || a >= hx + nx || b < hx || b >= hx + ny) continue;
            const double xi = S.x[t], yi = S.y[t];
            const bool s2i = S.pid[t] >= g.n1;
            R fxs = 0, fys = 0;
            for (int pass = 0; pass < 2; ++pass)
                for (int da = -1; da <= 1; ++da)
                    for (int db = -1; db <= 1; ++db) {
                        const int cj = (a + da) * SY + (b + db);
                        const int ue = S.cstart[cj + 1];
                        for (int u = S.cstart[cj]; u < ue; ++u) {
                            if (u == t) continue;
                            const bool s2j = S.pid[u] >= g.n1;
                            if ((pass == 0) != (s2i == s2j)) continue;
                            R fx, fy;
                            if (!s2i && !s2j) {
                                if (wca_pair<R>(xi, yi, S.x[u], S.y[u], 0, g, fx, fy)) { fxs += fx; fys += fy; }
                            } else if (s2i && s2j) {
                                if (wca_pair<R>(xi, yi, S.x[u], S.y[u], 1, g, fx, fy)) { fxs += fx; fys += fy; }
                            } else if (!s2i) {
                                if (wca_pair<R>(xi, yi, S.x[u], S.y[u], 2, g, fx, fy)) { fxs += fx; fys += fy; }
                            } else {
                                if (wca_pair<R>(S.x[u], S.y[u], xi, yi, 2, g, fx, fy)) { fxs -= fx; fys -= fy; }
                            }
                        }
                    }
            R2 o;
            o.x = fxs;
            o.y = fys;
            fw[S.slot[t]] = o;
        }
    }
    if (what & 2) {
        R *rho = (R *)d.rho;
        const int nlx = lhi - llo;
        const int base_u = cx0 - hx, base_v = cy0 - hx;
        const R drho = (R)g.drho, q2 = (R)g.f2_f1;
        const R drho2 = (R)(g.drho * g.f2_f1);
        for (int pi = tid; pi < (khi - klo) * nlx; pi += nt) {
            const int k = klo + pi / nlx, l = llo + pi - (pi / nlx) * nlx;
            int culo = (int)ceil(((k - 1) * g.h - e) / g.B) - 1, cuhi = (int)ceil(((k + 2) * g.h + e) / g.B) - 1;
            int cvlo = (int)ceil(((l - 1) * g.h - e) / g.B) - 1, cvhi = (int)ceil(((l + 2) * g.h + e) / g.B) - 1;
            if (cuhi - culo + 1 > nc) cuhi = culo + nc - 1;
            if (cvhi - cvlo + 1 > nc) cvhi = cvlo + nc - 1;
            int alo = max(culo - base_u, 0), ahi = min(cuhi - base_u, SX - 1);
            int blo = max(cvlo - base_v, 0), bhi = min(cvhi - base_v, SY - 1);
            R acc = 0;
            for (int a = alo; a <= ahi; ++a)
                for (int b = blo; b <= bhi; ++b) {
                    const int cj = a * SY + b;
                    const int ue = S.cstart[cj + 1];
                    for (int u = S.cstart[cj]; u < ue; ++u) {
                        int dk = k - S.kp[u], dl = l - S.lp[u];
                        if (dk > Z / 2) dk -= Z;
                        if (dk < -Z / 2) dk += Z;
                        if (dl > Z / 2) dl -= Z;
                        if (dl < -Z / 2) dl += Z;
                        if (dk < -1 || dk > 1 || dl < -1 || dl > 1) continue;
                        const bool s2 = S.pid[u] >= g.n1;
                        if (SCHEME == 0) {
                            if (dk != 0 || dl != 0) continue;  // NGP: the particle's own mesh cell only
                            acc += s2 ? drho2 : drho;
                        } else {
                            R c = S.wx[u][dk + 1] * S.wy[u][dl + 1] * drho;
                            if (s2) c = c * q2;
                            acc += c;
                        }
                    }
                }
            rho[(size_t)k * Z + l] = acc;
        }
    }
}

// ------------------------------------------------------------------ FFT in LDS
// One Stockham pass of radix RAD over nrows rows of length n = 2^logn held
// in LDS (row stride `stride`), in place: every thread loads its butterflies'
// inputs into registers, the block synchronises, then it writes the outputs.
// tw[k] = exp(-2 pi i k / 2^logZ); the inverse uses the conjugates.
template <typename R, int EPT, int RAD>
__device__ __forceinline__ void fft_pass(typename V2<R>::T *buf, int logn, int nrows, int stride, int logNs,
                                         const typename V2<R>::T *__restrict__ tw, int logZ, bool inv) {
    using C = typename V2<R>::T;
    constexpr int LR = RAD == 4 ? 2 : 1;
    constexpr int NB = EPT / RAD;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lognb = logn - LR, nb = 1 << lognb;
    const int total = nb << 0;
    const int tot = total * nrows;
    const int Ns = 1 << logNs;
    const int twsh = logZ - (logNs + LR);
    C v[NB][RAD];
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        const int b = tid + e * nt;
        if (b < tot) {
            const int row = b >> lognb, j = b & (nb - 1);
            const C *p = buf + row * stride + j;
#pragma unroll
            for (int t = 0; t < RAD; ++t) v[e][t] = p[t * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        const int b = tid + e * nt;
        if (b < tot) {
            const int row = b >> lognb, j = b & (nb - 1);
            const int k = j & (Ns - 1);
#pragma unroll
            for (int t = 1; t < RAD; ++t) {
                C w = tw[(t * k) << twsh];
                if (inv) w.y = -w.y;
                v[e][t] = cmul(v[e][t], w);
            }
            C y[RAD];
            if (RAD == 2) {
                y[0] = cadd(v[e][0], v[e][1]);
                y[1] = csub(v[e][0], v[e][1]);
            } else {
                C t0 = cadd(v[e][0], v[e][2]), t1 = csub(v[e][0], v[e][2]);
                C t2 = cadd(v[e][1], v[e][3]), t3 = csub(v[e][1], v[e][3]);
                C r3 = inv ? cmul_i(t3) : cmul_mi(t3);
                y[0] = cadd(t0, t2);
                y[2] = csub(t0, t2);
                y[1] = cadd(t1, r3);
                y[3] = csub(t1, r3);
            }
            C *q = buf + row * stride + ((j - k) << LR) + k;
#pragma unroll
            for (int t = 0; t < RAD; ++t) q[t * Ns] = y[t];
        }
    }
    __syncthreads();
}

template <typename R, int EPT>
__device__ void lds_fft(typename V2<R>::T *buf, int logn, int nrows, int stride,
                        const typename V2<R>::T *__restrict__ tw, int logZ, bool inv) {
    int logNs = 0;
    while (logn - logNs >= 2) {
        fft_pass<R, EPT, 4>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
        logNs += 2;
    }
    if (logn - logNs == 1) fft_pass<R, EPT, 2>(buf, logn, nrows, stride, logNs, tw, logZ, inv);
}

constexpr int FFT_THREADS = 256;
constexpr int FFT_EPT = 32;

// Row r2c: rows k0..k0+RB-1 of the real mesh -> half spectrum X[m], m<Z/2,
// with X[0] packing (DC, Nyquist) (both real), written transposed:
// ht[m*Z + k].  Length-Z/2 complex FFT of (x[2n], x[2n+1]) + split.
template <typename R>
__global__ __launch_bounds__(FFT_THREADS) void k_r2c_rows(const R *__restrict__ rho, typename V2<R>::T *__restrict__ ht,
                                                          const typename V2<R>::T *__restrict__ tw, int Z, int logZ,
                                                          int RB, int stride) {
    using C = typename V2<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C *s = (C *)smem_raw;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int n = Z >> 1, logn = logZ - 1;
    const int k0 = blockIdx.x * RB;
    for (int idx = tid; idx < RB * n; idx += nt) {
        const int row = idx >> logn, m = idx & (n - 1);
        s[row * stride + m] = ((const C *)(rho + (size_t)(k0 + row) * Z))[m];
    }
    __syncthreads();
    lds_fft<R, FFT_EPT>(s, logn, RB, stride, tw, logZ, false);
    const int half = n >> 1;
    for (int idx = tid; idx < RB * (half + 1); idx += nt) {
        const int row = idx / (half + 1), m = idx - row * (half + 1);
        C *p = s + row * stride;
        if (m == 0) {
            C z = p[0];
            p[0] = cm<C>(z.x + z.y, z.x - z.y);
        } else {
            C a = p[m], bc = cconj(p[n - m]);
            C E = cscale(cadd(a, bc), (R)0.5);
            C O = cmul_mi(cscale(csub(a, bc), (R)0.5));
            C WO = cmul(tw[m], O);
            p[m] = cadd(E, WO);
            if (m != n - m) p[n - m] = cconj(csub(E, WO));
        }
    }
    __syncthreads();
    for (int idx = tid; idx < RB * n; idx += nt) {
        const int row = idx % RB, m = idx / RB;
        ht[(size_t)m * Z + k0 + row] = s[row * stride + m];
    }
}

// spectral factors of one (j, m') pair: gx = sq[j] G / Z / Z, gy = sq[m'] G / Z / Z,
// G = 1/((lap[j] + lap[m']) + 1/lambda^2) (dm:585-593, dm:1384-1394)
template <typename R>
__device__ __forceinline__ void symbols(int j, int mm, const Geo &g, const R *sq, const R *lap, R &gx, R &gy) {
    R G;
    if (j == 0 && mm == 0 && g.zero_mode_g)
        G = (R)0.0;
    else
        G = (R)1.0 / ((lap[j] + lap[mm]) + (R)g.lam2inv);
    const R Zr = (R)g.Z;
    gx = sq[j] * G / Zr / Zr;
    gy = sq[mm] * G / Zr / Zr;
}

// Column pass: rows m0..m0+RC-1 of ht (length Z along x), forward FFT,
// multiply by i*gx and i*gy, inverse FFT of both, out gt[m*Z+k] = (Fx, Fy).
template <typename R>
__global__ __launch_bounds__(FFT_THREADS) void k_col(const typename V2<R>::T *__restrict__ ht,
                                                     typename V4<R>::T *__restrict__ gt,
                                                     const typename V2<R>::T *__restrict__ tw, const R *__restrict__ sq,
                                                     const R *__restrict__ lap, Geo g, int RC, int stride) {
    using C = typename V2<R>::T;
    using C4 = typename V4<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C *s = (C *)smem_raw;
    C *sB = s + RC * stride;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int Z = g.Z, logZ = g.logZ, Zh = Z >> 1;
    const int m0 = blockIdx.x * RC;
    for (int idx = tid; idx < RC * Z; idx += nt) {
        const int row = idx >> logZ, k = idx & (Z - 1);
        s[row * stride + k] = ht[(size_t)(m0 + row) * Z + k];
    }
    __syncthreads();
    lds_fft<R, FFT_EPT>(s, logZ, RC, stride, tw, logZ, false);
    for (int idx = tid; idx < RC * Z; idx += nt) {
        const int row = idx >> logZ, j = idx & (Z - 1);
        const int m = m0 + row;
        if (m == 0) {
            // packed (DC, Nyquist) column: pairs (j, Z-j) by one thread
            if (j > Zh) continue;
            const int jn = (Z - j) & (Z - 1);
            C P = s[j], Q = s[jn];
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
            s[j] = fxj;
            sB[j] = fyj;
            if (jn != j) {
                s[jn] = fxn;
                sB[jn] = fyn;
            }
        } else {
            R gx, gy;
            symbols<R>(j, m, g, sq, lap, gx, gy);
            C S = s[row * stride + j];
            s[row * stride + j] = cm<C>(-gx * S.y, gx * S.x);
            sB[row * stride + j] = cm<C>(-gy * S.y, gy * S.x);
        }
    }
    __syncthreads();
    lds_fft<R, FFT_EPT>(s, logZ, 2 * RC, stride, tw, logZ, true);
    for (int idx = tid; idx < RC * Z; idx += nt) {
        const int row = idx >> logZ, k = idx & (Z - 1);
        C a = s[row * stride + k], b = sB[row * stride + k];
        C4 o;
        o.x = a.x;
        o.y = a.y;
        o.z = b.x;
        o.w = b.y;
        gt[(size_t)(m0 + row) * Z + k] = o;
    }
}

// Row c2r of the two real signals Fx, Fy: rows k0..k0+RB2-1, spectrum
// C[m] = X[m] + i Y[m] completed by Hermitian symmetry, inverse FFT of
// length Z, out fmesh[k*Z + l] = (Re, Im) = (Fx, Fy).
template <typename R>
__global__ __launch_bounds__(FFT_THREADS) void k_c2r_rows(const typename V4<R>::T *__restrict__ gt,
                                                          typename V2<R>::T *__restrict__ fm,
                                                          const typename V2<R>::T *__restrict__ tw, int Z, int logZ,
                                                          int RB2, int stride) {
    using C = typename V2<R>::T;
    using C4 = typename V4<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C *s = (C *)smem_raw;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int Zh = Z >> 1;
    const int k0 = blockIdx.x * RB2;
    for (int idx = tid; idx < RB2 * Zh; idx += nt) {
        const int row = idx % RB2, m = idx / RB2;
        C4 v = gt[(size_t)m * Z + k0 + row];
        C *p = s + row * stride;
        if (m == 0) {
            p[0] = cm<C>(v.x, v.z);
            p[Zh] = cm<C>(v.y, v.w);
        } else {
            p[m] = cm<C>(v.x - v.w, v.y + v.z);
            p[Z - m] = cm<C>(v.x + v.w, -v.y + v.z);
        }
    }
    __syncthreads();
    lds_fft<R, FFT_EPT>(s, logZ, RB2, stride, tw, logZ, true);
    for (int idx = tid; idx < RB2 * Z; idx += nt) {
        const int row = idx >> logZ, l = idx & (Z - 1);
        fm[(size_t)(k0 + row) * Z + l] = s[row * stride + l];
    }
}

// ------------------------------------------------------------------ gather, update
template <typename R, int SCHEME>
__device__ __forceinline__ void gather_kap(double x, double y, bool s2, const typename V2<R>::T *__restrict__ fm,
                                           const Geo &g, double &fx, double &fy) {
    using R2 = typename V2<R>::T;
    const int Z = g.Z;
    int k, l;
    double wxd[3], wyd[3];
    weights1d<SCHEME>(x, g.h, k, wxd);
    weights1d<SCHEME>(y, g.h, l, wyd);
    k = wrapc(k, Z);
    l = wrapc(l, Z);
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
    const int kx[3] = {(k - 1 + Z) & (Z - 1), k, (k + 1) & (Z - 1)};
    const int ly[3] = {(l - 1 + Z) & (Z - 1), l, (l + 1) & (Z - 1)};
    R wx[3], wy[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) {
        wx[u] = (R)wxd[u];
        wy[u] = (R)wyd[u];
    }
    R ax = 0, ay = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const R2 *row = fm + (size_t)kx[a] * Z;
        R2 f0 = row[ly[0]], f1 = row[ly[1]], f2 = row[ly[2]];
        R sx = wy[0] * f0.x + wy[1] * f1.x + wy[2] * f2.x;
        R sy = wy[0] * f0.y + wy[1] * f1.y + wy[2] * f2.y;
        if (a == 0) {
            ax = wx[0] * sx;
            ay = wx[0] * sy;
        } else {
            ax += wx[a] * sx;
            ay += wx[a] * sy;
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
    } else {
        n = philox_noise(g.seed, (uint32_t)pid, step, g.noise_mode == BDK_NOISE_GAUSS, g.noise_var);
    }
    return n;
}

// forces only: fkap (and dt candidates when with_dt)
template <typename R, int SCHEME>
__global__ __launch_bounds__(256) void k_gather(Geo g, DevBufs d, int with_dt) {
    using R2 = typename V2<R>::T;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    double cand = INFINITY;
    if (s < g.N) {
        const int c = d.cell[s];
        const int gx = c / g.nc, gy = c - gx * g.nc;
        const double2 r = d.rel[s];
        const int pid = d.pid[s];
        double fx = 0.0, fy = 0.0;
        if (!g.idgas) gather_kap<R, SCHEME>(gx * g.B + r.x, gy * g.B + r.y, pid >= g.n1, (const R2 *)d.fmesh, g, fx, fy);
        R2 o;
        o.x = (R)fx;
        o.y = (R)fy;
        ((R2 *)d.fkap)[s] = o;
        if (with_dt) {
            const R2 fw = ((const R2 *)d.fwca)[s];
            const double2 n = particle_noise(g, d, pid, d.sc->step);
            double c0 = dt_candidate((double)fw.x + (double)o.x, n.x, g);
            double c1 = dt_candidate((double)fw.y + (double)o.y, n.y, g);
            if (c0 < cand) cand = c0;
            if (c1 < cand) cand = c1;
        }
    }
    if (with_dt) {
        // NaN candidates never win (std::min semantics of dm:1420-1491)
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            double v = __shfl_xor(cand, o, 64);
            if (v < cand) cand = v;
        }
        if ((threadIdx.x & 63) == 0 && cand < INFINITY && cand >= 0.0) atomicMin(&d.sc->dt_bits, bitsd(cand));
    }
}

// Euler step (dm:116-157) + refresh (dm:216-325) + cell counting.
// FUSED: capillary force gathered here (fixed dt); else read from fkap.
template <typename R, int SCHEME, bool FUSED>
__global__ __launch_bounds__(256) void k_update(Geo g, DevBufs d) {
    using R2 = typename V2<R>::T;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    const unsigned long long step = d.sc->step;
    const int c = d.cell[s];
    int gx = c / g.nc, gy = c - gx * g.nc;
    double2 r = d.rel[s];
    const int pid = d.pid[s];
    const bool s2 = pid >= g.n1;
    double fwx = 0.0, fwy = 0.0, fkx = 0.0, fky = 0.0;
    if (!g.idgas) {
        const R2 fw = ((const R2 *)d.fwca)[s];
        fwx = (double)fw.x;
        fwy = (double)fw.y;
        if (FUSED) {
            gather_kap<R, SCHEME>(gx * g.B + r.x, gy * g.B + r.y, s2, (const R2 *)d.fmesh, g, fkx, fky);
        } else {
            const R2 fk = ((const R2 *)d.fkap)[s];
            fkx = (double)fk.x;
            fky = (double)fk.y;
        }
    }
    const double2 n = particle_noise(g, d, pid, step);
    const double dt = g.adaptive ? dbits(d.sc->dt_bits) : g.dt_max;
    const double sq = sqrt(2.0 * g.T * dt);
    const double dx = dt * (fwx + fkx) + sq * n.x;
    const double dy = dt * (fwy + fky) + sq * n.y;
    double2 rn;
    rn.x = r.x + dx;
    rn.y = r.y + dy;
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
    if (!(fabs(rn.x) < 1e9 && fabs(rn.y) < 1e9)) {
        // non-finite or runaway position: keep the old state, raise the flag
        atomicOr(&d.sc->nonfinite, 1);
        rn = r;
    } else if (rn.x < 0.0 || rn.x > B || rn.y < 0.0 || rn.y > B) {
        int ic = (int)floor((gx * B + rn.x) / B);
        int jc = (int)floor((gy * B + rn.y) / B);
        ic = wrapc(ic + g.nc, g.nc);
        jc = wrapc(jc + g.nc, g.nc);
        rn.x = fmod(rn.x + B, B);
        rn.y = fmod(rn.y + B, B);
        gx = ic;
        gy = jc;
        nc_new = gx * g.nc + gy;
    }
    d.cell[s] = nc_new;
    d.rel[s] = rn;
    d.rank[s] = atomicAdd(&d.count[nc_new], 1);
    double ex = fmax(fmax(-rn.x, rn.x - B), fmax(-rn.y, rn.y - B));
    if (ex > 0.0) atomicMax(&d.sc->exc_bits[(step + 1) & 1], bitsd(ex));
    if (s == 0) d.sc->dt_last = bitsd(dt);
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
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER_THREAD = 16;
constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_PER_THREAD;  // 4096 cells per block

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const int *__restrict__ count, int ncell,
                                                              int *__restrict__ bsum) {
    __shared__ int ws[16];
    const int base = blockIdx.x * SCAN_BLOCK;
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
        bsum[blockIdx.x] = t;
    }
}

// exclusive scan of the block sums (one block of 1024 threads, nb <= 4096);
// also closes the step: start[ncell] = total, step counter += inc_step
__global__ __launch_bounds__(1024) void k_scan_top(int *bsum, int nb, int *start, int ncell, DevScalars *sc,
                                                   int inc_step) {
    __shared__ int a[4096];
    __shared__ int ws[16];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) a[i] = bsum[i];
    __syncthreads();
    int total = block_exscan4(a, nb, ws);
    for (int i = threadIdx.x; i < nb; i += blockDim.x) bsum[i] = a[i];
    if (threadIdx.x == 0) {
        start[ncell] = total;
        if (inc_step) sc->step += 1;
    }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(int *__restrict__ count, const int *__restrict__ bsum,
                                                            int *__restrict__ start, int ncell) {
    __shared__ int a[SCAN_BLOCK];
    __shared__ int ws[16];
    const int base = blockIdx.x * SCAN_BLOCK;
    for (int q = threadIdx.x; q < SCAN_BLOCK; q += SCAN_THREADS) {
        int i = base + q;
        a[q] = i < ncell ? count[i] : 0;
    }
    __syncthreads();
    // 16 consecutive elements per thread
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int loc[SCAN_PER_THREAD];
    int s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_PER_THREAD; ++q) {
        loc[q] = a[tid * SCAN_PER_THREAD + q];
        s += loc[q];
    }
    int x = s;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    int woff = bsum[blockIdx.x];
    for (int q = 0; q < w; ++q) woff += ws[q];
    int e = woff + x - s;
#pragma unroll
    for (int q = 0; q < SCAN_PER_THREAD; ++q) {
        a[tid * SCAN_PER_THREAD + q] = e;
        e += loc[q];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < SCAN_BLOCK; q += SCAN_THREADS) {
        int i = base + q;
        if (i < ncell) {
            start[i] = a[q];
            count[i] = 0;
        }
    }
}

__global__ void k_fill(Geo g, DevBufs d) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= g.N) return;
    d.list[d.start[d.cell[s]] + d.rank[s]] = s;
}

// ------------------------------------------------------------------ launchers
static inline unsigned nblk(int n, int b) { return (unsigned)((n + b - 1) / b); }

template <typename R> struct FftCfg {
    int RB, RC, RB2, sR1, sC, sR2;
    size_t l1, lc, l2;
};

template <typename R> static FftCfg<R> fft_cfg(int Z) {
    using C = typename V2<R>::T;
    const int E = FFT_EPT * FFT_THREADS;  // complex elements per block
    FftCfg<R> c;
    c.RB = std::min(Z, E / (Z / 2));
    c.RC = std::max(1, std::min(Z / 2, E / (2 * Z)));
    c.RB2 = std::min(Z, E / Z);
    c.sR1 = Z / 2 + 1;
    c.sC = Z + 1;
    c.sR2 = Z + 1;
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
    BDK_ATTR(k_r2c_rows<float>)
    BDK_ATTR(k_r2c_rows<double>)
    BDK_ATTR(k_col<float>)
    BDK_ATTR(k_col<double>)
    BDK_ATTR(k_c2r_rows<float>)
    BDK_ATTR(k_c2r_rows<double>)
#undef BDK_ATTR
    return e;
}

template <typename R, int SCHEME> static void launch_force_t(bdk_t *h, int what) {
    dim3 grid(nblk(h->g.nc, FT_TC), nblk(h->g.nc, FT_TC));
    hipLaunchKernelGGL((k_force<R, SCHEME>), grid, dim3(FT_THREADS), 0, h->stream, h->g, h->d, what);
}

hipError_t launch_force(bdk_t *h, int what) {
    const bool f32 = h->p.precision == BDK_FP32;
    switch (h->g.scheme) {
    case 0: f32 ? launch_force_t<float, 0>(h, what) : launch_force_t<double, 0>(h, what); break;
    case 1: f32 ? launch_force_t<float, 1>(h, what) : launch_force_t<double, 1>(h, what); break;
    default: f32 ? launch_force_t<float, 2>(h, what) : launch_force_t<double, 2>(h, what); break;
    }
    return hipGetLastError();
}

template <typename R> static void launch_fft_t(bdk_t *h, int only = -1) {
    using C = typename V2<R>::T;
    using C4 = typename V4<R>::T;
    const int Z = h->g.Z, logZ = h->g.logZ;
    FftCfg<R> c = fft_cfg<R>(Z);
    const R *sq = std::is_same<R, float>::value ? (const R *)h->d.sqf : (const R *)h->d.sq;
    const R *lap = std::is_same<R, float>::value ? (const R *)h->d.lapf : (const R *)h->d.lap;
    if (only < 0 || only == 0)
        hipLaunchKernelGGL((k_r2c_rows<R>), dim3(Z / c.RB), dim3(FFT_THREADS), c.l1, h->stream,
                           (const R *)h->d.rho, (C *)h->d.ht, (const C *)h->d.tw, Z, logZ, c.RB, c.sR1);
    if (only < 0 || only == 1)
        hipLaunchKernelGGL((k_col<R>), dim3((Z / 2) / c.RC), dim3(FFT_THREADS), c.lc, h->stream,
                           (const C *)h->d.ht, (C4 *)h->d.gt, (const C *)h->d.tw, sq, lap, h->g, c.RC, c.sC);
    if (only < 0 || only == 2)
        hipLaunchKernelGGL((k_c2r_rows<R>), dim3(Z / c.RB2), dim3(FFT_THREADS), c.l2, h->stream,
                           (const C4 *)h->d.gt, (C *)h->d.fmesh, (const C *)h->d.tw, Z, logZ, c.RB2, c.sR2);
}

hipError_t launch_fft_pass(bdk_t *h, int pass) {
    if (h->p.precision == BDK_FP32)
        launch_fft_t<float>(h, pass);
    else
        launch_fft_t<double>(h, pass);
    return hipGetLastError();
}

hipError_t launch_fft(bdk_t *h) {
    if (h->p.precision == BDK_FP32)
        launch_fft_t<float>(h);
    else
        launch_fft_t<double>(h);
    return hipGetLastError();
}

template <typename R, int SCHEME> static void launch_gather_t(bdk_t *h, int with_dt) {
    hipLaunchKernelGGL((k_gather<R, SCHEME>), dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d,
                       with_dt);
}

hipError_t launch_gather(bdk_t *h, int with_dt) {
    const bool f32 = h->p.precision == BDK_FP32;
    switch (h->g.scheme) {
    case 0: f32 ? launch_gather_t<float, 0>(h, with_dt) : launch_gather_t<double, 0>(h, with_dt); break;
    case 1: f32 ? launch_gather_t<float, 1>(h, with_dt) : launch_gather_t<double, 1>(h, with_dt); break;
    default: f32 ? launch_gather_t<float, 2>(h, with_dt) : launch_gather_t<double, 2>(h, with_dt); break;
    }
    return hipGetLastError();
}

template <typename R, int SCHEME> static void launch_update_t(bdk_t *h, int fused) {
    if (fused)
        hipLaunchKernelGGL((k_update<R, SCHEME, true>), dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g,
                           h->d);
    else
        hipLaunchKernelGGL((k_update<R, SCHEME, false>), dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g,
                           h->d);
}

hipError_t launch_update(bdk_t *h, int fused) {
    const bool f32 = h->p.precision == BDK_FP32;
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
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, h->stream, h->d.count, ncell, h->d.bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, h->stream, h->d.bsum, nb, h->d.start, ncell, h->d.sc,
                       inc_step);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_THREADS), 0, h->stream, h->d.count, h->d.bsum, h->d.start,
                       ncell);
    hipLaunchKernelGGL(k_fill, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d);
}

hipError_t launch_scan_fill(bdk_t *h) {
    launch_scan_fill_inner(h, 1);
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
    hipLaunchKernelGGL(k_refresh, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d);
    launch_scan_fill_inner(h, 0);
    return hipGetLastError();
}

hipError_t launch_noise_out(bdk_t *h) {
    hipLaunchKernelGGL(k_noise_out, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d);
    return hipGetLastError();
}

// One step.  With ev != nullptr an event is recorded after every launch
// (profiling); names receives the labels.
hipError_t enqueue_step(bdk_t *h, std::vector<std::string> *names, std::vector<hipEvent_t> *ev) {
    auto mark = [&](const char *nm) {
        if (names) names->push_back(nm);
        if (ev) {
            hipEvent_t e;
            hipEventCreate(&e);
            hipEventRecord(e, h->stream);
            ev->push_back(e);
        }
    };
    if (ev) mark("begin");
    const bool kap = !h->g.idgas;
    const bool adaptive = h->g.adaptive;
    // what: 1 WCA, 2 deposit, 4 reset of the dt min
    int what = (h->g.idgas ? 0 : 1 | 2) | 4;
    hipError_t e = launch_force(h, what);
    mark("k_force");
    if (kap) {
        if (h->p.precision == BDK_FP32) {
            using C = float2;
            using C4 = float4;
            const int Z = h->g.Z, logZ = h->g.logZ;
            FftCfg<float> c = fft_cfg<float>(Z);
            hipLaunchKernelGGL((k_r2c_rows<float>), dim3(Z / c.RB), dim3(FFT_THREADS), c.l1, h->stream,
                               (const float *)h->d.rho, (C *)h->d.ht, (const C *)h->d.tw, Z, logZ, c.RB, c.sR1);
            mark("k_r2c_rows");
            hipLaunchKernelGGL((k_col<float>), dim3((Z / 2) / c.RC), dim3(FFT_THREADS), c.lc, h->stream,
                               (const C *)h->d.ht, (C4 *)h->d.gt, (const C *)h->d.tw, (const float *)h->d.sqf,
                               (const float *)h->d.lapf, h->g, c.RC, c.sC);
            mark("k_col");
            hipLaunchKernelGGL((k_c2r_rows<float>), dim3(Z / c.RB2), dim3(FFT_THREADS), c.l2, h->stream,
                               (const C4 *)h->d.gt, (C *)h->d.fmesh, (const C *)h->d.tw, Z, logZ, c.RB2, c.sR2);
            mark("k_c2r_rows");
        } else {
            using C = double2;
            using C4 = double4;
            const int Z = h->g.Z, logZ = h->g.logZ;
            FftCfg<double> c = fft_cfg<double>(Z);
            hipLaunchKernelGGL((k_r2c_rows<double>), dim3(Z / c.RB), dim3(FFT_THREADS), c.l1, h->stream,
                               (const double *)h->d.rho, (C *)h->d.ht, (const C *)h->d.tw, Z, logZ, c.RB, c.sR1);
            mark("k_r2c_rows");
            hipLaunchKernelGGL((k_col<double>), dim3((Z / 2) / c.RC), dim3(FFT_THREADS), c.lc, h->stream,
                               (const C *)h->d.ht, (C4 *)h->d.gt, (const C *)h->d.tw, (const double *)h->d.sq,
                               (const double *)h->d.lap, h->g, c.RC, c.sC);
            mark("k_col");
            hipLaunchKernelGGL((k_c2r_rows<double>), dim3(Z / c.RB2), dim3(FFT_THREADS), c.l2, h->stream,
                               (const C4 *)h->d.gt, (C *)h->d.fmesh, (const C *)h->d.tw, Z, logZ, c.RB2, c.sR2);
            mark("k_c2r_rows");
        }
    }
    if (adaptive) {
        launch_gather(h, 1);
        mark("k_gather");
        launch_update(h, 0);
    } else {
        launch_update(h, kap ? 1 : 0);
    }
    mark("k_update");
    const int ncell = h->g.ncell;
    const int nb = (int)nblk(ncell, SCAN_BLOCK);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, h->stream, h->d.count, ncell, h->d.bsum);
    mark("k_scan_reduce");
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, h->stream, h->d.bsum, nb, h->d.start, ncell, h->d.sc, 1);
    mark("k_scan_top");
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SCAN_THREADS), 0, h->stream, h->d.count, h->d.bsum, h->d.start,
                       ncell);
    mark("k_scan_down");
    hipLaunchKernelGGL(k_fill, dim3(nblk(h->g.N, 256)), dim3(256), 0, h->stream, h->g, h->d);
    mark("k_fill");
    if (e == hipSuccess) e = hipGetLastError();
    return e;
}

}  // namespace bdkx
