// Device helpers shared by the 2D (solver.hip) and 3D (solver3d.hip)
// kernels: launch constants, wave / group reductions in a fixed order, the
// XCD-aware block mapping and the per-mode constants.
#pragma once

#include <hip/hip_runtime.h>

namespace gbte {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kSweepBlock = 1024;
constexpr int kMaxIn = 3;  // in-faces per cell of one direction (triangles, quads, tetrahedra)
constexpr int kNT = 8;     // row entries per thread in the row kernels
constexpr int kNTf = 4;    // the same for the register-heavy face / source kernels
constexpr int kRowBlockMax = 512;
constexpr int kStencilBatch = 4;  // LSQ stencil entries loaded together
constexpr int kNJ = 4;            // row entries per thread loaded together

__device__ __forceinline__ int wave_global() { return (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); }
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

// Sum over the nB lanes of one cell group; valid in the group's first lane.
__device__ __forceinline__ double group_sum(double v, int b, int nB) {
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_down(v, off, 64);
        if (off < nB && b + off < nB) v += o;
    }
    return v;
}

// Rows of n_dirs*n_bands values are processed by tpr threads (a power of two,
// tpr * rpb = kBlock); row_sum reduces over the tpr threads of a row in a
// fixed order and returns the sum to every thread of the row.
__device__ __forceinline__ double row_sum(double v, int tpr, double *sh) {
    int w = threadIdx.x >> 6;
    int lim = tpr < 64 ? tpr : 64;
    for (int off = lim >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (tpr <= 64) return v;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    int wpr = tpr >> 6, w0 = (w / wpr) * wpr;
    double s = 0;
    for (int k = 0; k < wpr; ++k) s += sh[w0 + k];
    return s;
}

// Sum over the 64 lanes of a wave (butterfly, fixed order), in every lane.
__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Rows are Morton-ordered, so giving every XCD one contiguous range of
// rows keeps neighbouring rows - which every stencil re-reads - in one L2.
constexpr int kXcd = 8;
// The grid must be a multiple of kXcd for this to be a bijection.
__device__ __forceinline__ int xcd_block() {
    int nb = gridDim.x, per = nb / kXcd;
    return (int)(blockIdx.x % kXcd) * per + (int)(blockIdx.x / kXcd);
}




inline int blocks_for(long n, int per_block) { return (int)((n + per_block - 1) / per_block); }

}  // namespace
}  // namespace gbte
