// Allocation-free pair correlation (bdk_pair_correlation_noalloc): the same
// histograms as bdk_pair_correlation (record_korrelationsfunktion11/12/22,
// auswertung.cpp:110-188), for callers that must not allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The counts go into `scratch`, a file-scope device array of a
// fixed size: one copy per GPU, alive until the process ends.  Host code
// never passes the array's name to a kernel (on the host it is a stand-in,
// not a device address): bdk_create looks up each device's address of the
// array once, with that device current, and every use goes through it.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "bdk_internal.hpp"

namespace {

constexpr int kMaxBins = 4096;   // bins per histogram (the LDS histogram of k_pair_hist holds 3 x 4096)
constexpr int kMaxDevices = 64;

// [3][kMaxBins] ordered-pair counts (11, 12, 22): 96 KiB per GPU
__device__ __align__(256) unsigned long long scratch[3 * kMaxBins];

std::mutex g_mu;
unsigned long long *g_scratch[kMaxDevices] = {};

__global__ void k_zero_counts(unsigned long long *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace

namespace bdkx {

// Called by bdk_create with device `dev` current: the device address of this
// GPU's scratch array, looked up once per device.
hipError_t noalloc_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[dev]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = (unsigned long long *)ptr;
    return hipSuccess;
}

}  // namespace bdkx

extern "C" int bdk_pair_correlation_noalloc(bdk_t *h, int32_t nbins, double dr, double *g11, double *g12,
                                            double *g22) {
    if (!h) return BDK_ERR_ARG;
    if (nbins < 1 || nbins > kMaxBins || !(dr > 0.0) || !g11 || !g12 || !g22) {
        h->err = "bdk_pair_correlation_noalloc: need 1 <= nbins <= " + std::to_string(kMaxBins) +
                 ", dr > 0 and all three outputs";
        return BDK_ERR_ARG;
    }
    if (!h->dev_valid) {
        h->err = "bdk_pair_correlation_noalloc: the particle state is not on the device yet (any force, step "
                 "or refresh call after bdk_set_positions puts it there)";
        return BDK_ERR_STATE;
    }
    unsigned long long *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (h->dev >= 0 && h->dev < kMaxDevices) ptr = g_scratch[h->dev];
    }
    if (!ptr) {
        h->err = "bdk_pair_correlation_noalloc: no scratch address for this device";
        return BDK_ERR_STATE;
    }
    hipError_t e = hipSetDevice(h->dev);
    const int n = 3 * nbins;
    // the array keeps the previous call's counts: clear the part this call uses
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_zero_counts, dim3((n + 255) / 256), dim3(256), 0, h->stream, ptr, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = bdkx::launch_pair_hist(h, nbins, dr, ptr);
    // the counts land in the caller's arrays as 64-bit integers, then are
    // scaled in place (both types are 8 bytes): no host buffer either
    double *out[3] = {g11, g12, g22};
    for (int k = 0; k < 3 && e == hipSuccess; ++k)
        e = hipMemcpyAsync(out[k], ptr + (size_t)k * nbins, (size_t)nbins * 8, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        h->err = std::string("bdk_pair_correlation_noalloc: ") + hipGetErrorString(e);
        return BDK_ERR_HIP;
    }
    // the prefactors of bdk_pair_correlation (auswertung.cpp:120-121, 149-150, 173-174)
    const double L = h->p.L, n1 = h->p.n1, n2 = h->p.n2;
    const double v11 = n1 > 1 ? 2.0 / ((n1 - 1) * M_PI * (n1 / L / L) * dr * dr) : 0.0;
    const double v12 = (n1 > 1 && n2 > 0) ? 1.0 / ((n1 - 1) * M_PI * (n2 / L / L) * dr * dr) : 0.0;
    const double v22 = n2 > 1 ? 2.0 / ((n2 - 1) * M_PI * (n2 / L / L) * dr * dr) : 0.0;
    for (int b = 0; b < nbins; ++b) {
        unsigned long long c[3];
        for (int k = 0; k < 3; ++k) std::memcpy(&c[k], &out[k][b], 8);
        g11[b] = (double)(c[0] / 2) * (v11 / (2 * b + 1));  // same-species pairs counted in both orders
        g12[b] = (double)c[1] * (v12 / (2 * b + 1));
        g22[b] = (double)(c[2] / 2) * (v22 / (2 * b + 1));
    }
    return BDK_OK;
}
