// Effective conductivity reduced on the device from a caller's flux field:
// k_eff = sum_i q_i V_i / (A_perp dT), the general form of the check in
// bin/examples/cross-plane/verify_kappa_crossplane.m:8 (mean(q) L / dT on
// full equal-area domains).  Two entry points share the kernels:
//   gbte_flux_keff          partial sums in a buffer it allocates per call
//   gbte_flux_keff_noalloc  partial sums in a fixed static device array, for
//                           callers that cannot allocate (graph capture,
//                           serving loops); stream-ordered, no host sync
// Both reduce in a fixed order (4096-cell blocks, then the block sums), so
// they give bit-identical results.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "giftbte.h"

namespace gbte {
std::string &thread_error();  // capi_problem.cpp
}

namespace {

constexpr int kKeffBlock = 256;
constexpr int kKeffPerThread = 16;
constexpr int kKeffCells = kKeffBlock * kKeffPerThread;  // cells per first-stage block
constexpr int kKeffMaxBlocks = 4096;                      // partial sums the static array holds
constexpr long kKeffMaxCells = (long)kKeffCells * kKeffMaxBlocks;
constexpr int kMaxDevices = 64;

// Partial sums of gbte_flux_keff_noalloc, one array per GPU (32 KiB), kept
// until the process ends.  Calls that overlap on one GPU share it and race.
static __device__ __align__(256) double scratch[kKeffMaxBlocks];

__device__ double block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kKeffBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

// part[b] = sum of q V over cells [b kKeffCells, (b + 1) kKeffCells)
__global__ void __launch_bounds__(kKeffBlock) k_flux_partial(const double *__restrict__ q,
                                                             const double *__restrict__ vol, int n,
                                                             double *__restrict__ part) {
    __shared__ double red[kKeffBlock];
    long base = (long)blockIdx.x * kKeffCells + threadIdx.x;
    double s = 0;
#pragma unroll
    for (int k = 0; k < kKeffPerThread; ++k) {
        long i = base + (long)k * kKeffBlock;
        if (i < n) s += q[i] * vol[i];
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[0] = scale * sum of part[0 .. nb)
__global__ void __launch_bounds__(kKeffBlock) k_flux_final(const double *__restrict__ part, int nb, double scale,
                                                           double *__restrict__ out) {
    __shared__ double red[kKeffBlock];
    double s = 0;
    for (int i = threadIdx.x; i < nb; i += kKeffBlock) s += part[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = s * scale;
}

int fail(int code, const std::string &msg) {
    gbte::thread_error() = msg;
    return code;
}

int check_args(const double *q, const double *vol, int n, double area, double dT, const double *out) {
    if (!q || !vol || !out) return fail(GBTE_BAD_INPUT, "gbte_flux_keff: null pointer");
    if (n <= 0) return fail(GBTE_BAD_INPUT, "gbte_flux_keff: n_cells must be positive");
    if ((long)n > kKeffMaxCells)
        return fail(GBTE_BAD_INPUT, "gbte_flux_keff: n_cells above the limit of " + std::to_string(kKeffMaxCells));
    if (!(area * dT != 0) || area != area || dT != dT)
        return fail(GBTE_BAD_INPUT, "gbte_flux_keff: zero or NaN area * delta_T");
    return GBTE_OK;
}

int launch(const double *q, const double *vol, int n, double area, double dT, double *out, double *part,
           hipStream_t s) {
    int nb = (int)((n + (long)kKeffCells - 1) / kKeffCells);
    hipLaunchKernelGGL(k_flux_partial, dim3(nb), dim3(kKeffBlock), 0, s, q, vol, n, part);
    hipLaunchKernelGGL(k_flux_final, dim3(1), dim3(kKeffBlock), 0, s, part, nb, 1.0 / (area * dT), out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(GBTE_HIP_ERROR, std::string("gbte_flux_keff: launch: ") + hipGetErrorString(e));
    return GBTE_OK;
}

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (the host-side name is not a device address).
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

int scratch_of_current(double **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(GBTE_HIP_ERROR, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(GBTE_HIP_ERROR, "gbte_flux_keff_noalloc: device index too large");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GBTE_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return GBTE_OK;
}

}  // namespace

extern "C" {

int gbte_flux_keff(const double *q_dev, const double *vol_dev, int n_cells, double area, double delta_T,
                   double *keff_dev, void *stream) {
    int rc = check_args(q_dev, vol_dev, n_cells, area, delta_T, keff_dev);
    if (rc != GBTE_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int nb = (int)((n_cells + (long)kKeffCells - 1) / kKeffCells);
    double *part = nullptr;
    hipError_t e = hipMalloc(&part, (size_t)nb * sizeof(double));
    if (e != hipSuccess) return fail(GBTE_HIP_ERROR, std::string("hipMalloc: ") + hipGetErrorString(e));
    rc = launch(q_dev, vol_dev, n_cells, area, delta_T, keff_dev, part, s);
    e = hipStreamSynchronize(s);  // the partials are freed below
    (void)hipFree(part);
    if (rc != GBTE_OK) return rc;
    if (e != hipSuccess) return fail(GBTE_HIP_ERROR, std::string("gbte_flux_keff: ") + hipGetErrorString(e));
    gbte::thread_error().clear();
    return GBTE_OK;
}

int gbte_flux_keff_noalloc(const double *q_dev, const double *vol_dev, int n_cells, double area, double delta_T,
                           double *keff_dev, void *stream) {
    int rc = check_args(q_dev, vol_dev, n_cells, area, delta_T, keff_dev);
    if (rc != GBTE_OK) return rc;
    double *part = nullptr;
    rc = scratch_of_current(&part);
    if (rc != GBTE_OK) return rc;
    // every partial that k_flux_final reads is written first by k_flux_partial
    // in the same stream, so nothing is cleared
    rc = launch(q_dev, vol_dev, n_cells, area, delta_T, keff_dev, part, static_cast<hipStream_t>(stream));
    if (rc == GBTE_OK) gbte::thread_error().clear();
    return rc;
}

}  // extern "C"
