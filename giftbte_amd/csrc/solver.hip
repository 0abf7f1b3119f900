// MI355X (gfx950) solver for the steady StaticBTESynthetic path.
//
// One outer iteration of StaticBTESynthetic::solve
// (src/StaticBTESynthetic/StaticBTESyntheticSolver.cpp:63-221) becomes:
//   k_copy_old       copy()                                  StaticBTESynthetic.cpp:2655-2690
//   k_source         the parts of _get_Re that do not depend on the new
//                    upwind values, for all systems             :1427-1500
//   k_sweep_dir      the per-system SparseLU solve: one workgroup walks every
//                    upwind level of one direction            Solver.cpp:96-100
//   k_bound_out      _get_bound_ee                              :1502-1539
//   k_vertex_e       _set_vertex_energydensity                  :1541-1616
//   k_face_moments   _get_face_gradient + _get_non_fourier + _recover_temperature
//                    + _get_heat_flux, and the next step's _get_gradient_larger
//                                                              :1651-1937, :2485-2520, :1243-1340
//   k_boundary       _set_bound_ee_1 + _set_macro_boundary_condition  :2083-2274
//   macro loop       _marco_solution: k_vertex_T, k_rhs, PCG    :2276-2407, :2692-2734
//   k_err            _get_magin_check_error                     :2565-2611
//
// Layout in HBM.  Every per-system field is stored cell-major with one
// contiguous row of n_dirs*n_bands values per cell (row index t = dir*B + band):
//   e     [cell][dir][band]           energy density, updated in place
//   grad  [cell][dir][band][dim]      LSQ gradient of the previous sweep's e
//   eb    [entry][dir][band]          boundary values, entry = 2*bound + side
//   eV    [node][dir][band]           inverse-distance vertex values
//   S     [dir][position][band]       sweep sources in each direction's upwind order
// The spatial gathers (LSQ stencils, face neighbours, vertex stencils) use the
// same neighbours for every band and direction, so a workgroup owns one row and
// reads whole neighbour rows with coalesced loads.  The sweep reads and writes
// the 8*n_bands-byte (cell, dir) segments of a row.
#include "devcommon.hpp"
#include "solver.hpp"
#include "tiles.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>

namespace gbte {

namespace {

// Storage of the per-system fields (e, its gradient, boundary values, sweep
// sources): ST = double, or float for GBTE_FP32_SWEEP; arithmetic is FP64.
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld2(const double *p) { return *reinterpret_cast<const double2 *>(p); }
__device__ __forceinline__ double2 ld2(const float *p) {
    float2 v = *reinterpret_cast<const float2 *>(p);
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st2_nt(double *p, double a, double b) {
    __builtin_nontemporal_store(v2d{a, b}, reinterpret_cast<v2d *>(p));
}
__device__ __forceinline__ void st2_nt(float *p, double a, double b) {
    __builtin_nontemporal_store(v2f{(float)a, (float)b}, reinterpret_cast<v2f *>(p));
}

struct SlotGeo {
    const int *fsize, *nb, *ent;
    const double *nx, *ny, *nz, *area, *vol, *rx, *ry, *rz, *qx, *qy, *qz;
};

struct Rows {
    int DB, B, tpr, rpb;  // row length, bands, threads per row, rows per block
    int dstep, bstep;     // tpr = dstep * B + bstep: (dir, band) advance per stride
};

// (direction, band) of row entry t0 and its advance by tpr entries, without
// an integer division per entry.
struct DirBand {
    int d, b;
    __device__ __forceinline__ DirBand(int t0, int nB) : d(t0 / nB), b(t0 - (t0 / nB) * nB) {}
    __device__ __forceinline__ void next(const Rows &R) {
        d += R.dstep;
        b += R.bstep;
        if (b >= R.B) { b -= R.B; ++d; }
    }
};

// Row and lane of this thread: one row per block (uniform row index, so
// per-row data lives in scalar registers) or rpb rows of tpr lanes.
template <bool OneRow>
__device__ __forceinline__ void row_lane(const Rows &R, int &row, int &t0) {
    if (OneRow) {
        row = xcd_block();
        t0 = threadIdx.x;
    } else {
        row = xcd_block() * R.rpb + threadIdx.x / R.tpr;
        t0 = threadIdx.x % R.tpr;
    }
}

// ---------------------------------------------------------------- kernels

// _set_initial (StaticBTESynthetic.cpp:765-782): e = T C_b for all systems
// (held as (T - eref) C_b).
template <typename ST>
__global__ void k_init_e(ST *__restrict__ e, const double *__restrict__ T, const double *__restrict__ cap, int DB,
                         int nB, size_t total, double eref) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    e[i] = (T[i / DB] - eref) * cap[(i % DB) % nB];
}

// _get_gradient_larger, Order 2 (StaticBTESynthetic.cpp:1243-1340), for the
// whole row of one cell: grad = P [(e_j - e_i) / L_x].  Each thread owns kNT
// row entries (stride tpr), so every stencil neighbour costs kNT independent
// loads in flight.
template <typename ST>
__device__ __forceinline__ void row_gradient(const ST *__restrict__ e, ST *__restrict__ grad, int c,
                                             const int *__restrict__ gptr, const int *__restrict__ gidx,
                                             const double *__restrict__ gP, const int *__restrict__ gok, int dim,
                                             int nnz, double invLx, int t0, int tpr, int DB) {
    int a = gptr[c], z = gptr[c + 1];
    bool ok = gok[c] != 0;
    const ST *ec = e + (size_t)c * DB;
    ST *gc = grad + (size_t)c * dim * DB;
    for (int tb = t0; tb < DB; tb += kNT * tpr) {
        double e0[kNT], gx[kNT], gy[kNT], gz[kNT];
#pragma unroll
        for (int j = 0; j < kNT; ++j) {
            int t = tb + j * tpr;
            e0[j] = t < DB ? ec[t] : 0.0;
            gx[j] = 0; gy[j] = 0; gz[j] = 0;
        }
        if (ok) {
            for (int k = a; k < z; ++k) {
                const ST *en = e + (size_t)gidx[k] * DB;
                double p0 = gP[k], p1 = gP[(size_t)nnz + k], p2 = dim == 3 ? gP[2 * (size_t)nnz + k] : 0.0;
#pragma unroll
                for (int j = 0; j < kNT; ++j) {
                    int t = tb + j * tpr;
                    if (t < DB) {
                        double d1 = invLx * (en[t] - e0[j]);
                        gx[j] += p0 * d1;
                        gy[j] += p1 * d1;
                        gz[j] += p2 * d1;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kNT; ++j) {
            int t = tb + j * tpr;
            if (t >= DB) continue;
            double x = gx[j], y = gy[j], w = dim == 3 ? gz[j] : 0.0;
            if (isnan(x) || isnan(y) || isnan(w)) { x = 0; y = 0; w = 0; }
            gc[(size_t)t * dim] = x;
            gc[(size_t)t * dim + 1] = y;
            if (dim == 3) gc[(size_t)t * dim + 2] = w;
        }
    }
}

template <typename ST>
__global__ void __launch_bounds__(kRowBlockMax) k_gradient(const ST *__restrict__ e, ST *__restrict__ grad,
                                                     const int *__restrict__ gptr, const int *__restrict__ gidx,
                                                     const double *__restrict__ gP, const int *__restrict__ gok,
                                                     int dim, int nnz, double invLx, int nc, Rows R) {
    int c = xcd_block() * R.rpb + threadIdx.x / R.tpr;
    if (c >= nc) return;
    row_gradient(e, grad, c, gptr, gidx, gP, gok, dim, nnz, invLx, threadIdx.x % R.tpr, R.tpr, R.DB);
}

// _get_bound_ee (StaticBTESynthetic.cpp:1502-1539): outgoing boundary values
// e + r . grad e for every system, one row per boundary entry.
template <typename ST>
__global__ void __launch_bounds__(kRowBlockMax) k_bound_out(const ST *__restrict__ e, const ST *__restrict__ grad,
                                                      ST *__restrict__ eb, SlotGeo g, Modes m,
                                                      const int *__restrict__ ent_list, const int *__restrict__ be_cell,
                                                      const int *__restrict__ be_slot, int n_ent, int dim, Rows R) {
    int k = xcd_block() * R.rpb + threadIdx.x / R.tpr;
    if (k >= n_ent) return;
    int en = ent_list[k], c = be_cell[en], s = be_slot[en];
    double nx = g.nx[s], ny = g.ny[s], nz = g.nz[s], rx = g.rx[s], ry = g.ry[s], rz = g.rz[s];
    const ST *ec = e + (size_t)c * R.DB;
    const ST *gc = grad + (size_t)c * dim * R.DB;
    ST *eo = eb + (size_t)en * R.DB;
    for (int t = threadIdx.x % R.tpr; t < R.DB; t += R.tpr) {
        int d = t / R.B, b = t - d * R.B;
        double vb = m.vg[b];
        double dot = vb * m.sx[d] * nx + vb * m.sy[d] * ny + vb * m.sz[d] * nz;
        if (dot >= 0) {
            double gz = dim == 3 ? gc[(size_t)t * dim + 2] : 0.0;
            eo[t] = ec[t] + (rx * gc[(size_t)t * dim] + ry * gc[(size_t)t * dim + 1] + rz * gz);
        }
    }
}

// Parallel part of one transport step for every system:
//   S = C T_old + Q hr tau - sum_out a (r.grad e_c) - sum_in,interior a (r'.grad e_nb)
//       - sum_in,boundary a eb,      a = tau (A/V) (v.n),
// i.e. everything of _get_Re (StaticBTESynthetic.cpp:1427-1500) that does not
// depend on the new upwind values.  One row per cell; written to each
// direction's upwind order for the sweep.  2D cells (at most 4 faces).
constexpr int kMaxFace2D = 4;

// Per-cell face record of the source kernel, one contiguous block so that a
// few scalar loads fetch it: doubles [q * 4 + f] for q = nx, ny, A/V, r_x,
// r_y, r'_x, r'_y (kSrcD per cell), ints [f] = neighbour cell, or -(entry+1)
// for a boundary face, or the cell itself past its last face (kSrcI).
constexpr int kSrcD = 32, kSrcI = 4;
constexpr int kSrcBatch = 2;  // entries per lane whose loads are issued before their stores
constexpr int kSrcWaves = 4;  // waves per SIMD the register budget allows

// One wave per cell row (kSrcRows rows per workgroup); the row index is
// wave-uniform, so the cell's face record sits in scalar registers.  The
// per-band and per-direction mode constants and the rows' upwind positions
// are copied to LDS first (dynamic LDS: 4 nB + 2 nD doubles, kSrcRows nD ints).
constexpr int kSrcRows = 4;

// What the source kernels share: the mode constants and the rows' upwind
// positions in LDS, and one cell's face record (normal, A/V, r from this
// centre, r' from the neighbour's centre, the neighbour's gradient row or the
// boundary-value row).  Faces past the last have zero records, so their term
// is exactly 0.
struct SrcLds {
    double *vg, *tau, *cap, *hr, *sx, *sy;
    int *pos;
};

__device__ __forceinline__ SrcLds src_stage(double *lds, const int *__restrict__ posc, const Modes &m, int c0,
                                            int nrows, const int *__restrict__ cells, int nB, int nD) {
    SrcLds L;
    L.vg = lds;
    L.tau = L.vg + nB;
    L.cap = L.tau + nB;
    L.hr = L.cap + nB;
    L.sx = L.hr + nB;
    L.sy = L.sx + nD;
    L.pos = reinterpret_cast<int *>(L.sy + nD);
    for (int i = threadIdx.x; i < nB; i += blockDim.x) {
        L.vg[i] = m.vg[i];
        L.tau[i] = m.tau[i];
        L.cap[i] = m.cap[i];
        L.hr[i] = m.hr[i];
    }
    for (int i = threadIdx.x; i < nD; i += blockDim.x) {
        L.sx[i] = m.sx[i];
        L.sy[i] = m.sy[i];
    }
    for (int i = threadIdx.x; i < kSrcRows * nD; i += blockDim.x) {
        int r = c0 + i / nD;
        int c = r < nrows ? (cells ? cells[r] : r) : -1;
        L.pos[i] = c >= 0 ? posc[(size_t)c * nD + i % nD] : 0;
    }
    __syncthreads();
    return L;
}

template <typename ST>
struct SrcFaces {
    double nx[kMaxFace2D], ny[kMaxFace2D], av[kMaxFace2D], rx[kMaxFace2D], ry[kMaxFace2D], qx[kMaxFace2D],
        qy[kMaxFace2D];
    const ST *in[kMaxFace2D];  // neighbour gradient row, or boundary-value row
    bool interior[kMaxFace2D];
    __device__ __forceinline__ SrcFaces(const double *rd, const int *ri, const ST *grad, const ST *eb, int DB) {
#pragma unroll
        for (int f = 0; f < kMaxFace2D; ++f) {
            nx[f] = rd[f];
            ny[f] = rd[4 + f];
            av[f] = rd[8 + f];
            rx[f] = rd[12 + f];
            ry[f] = rd[16 + f];
            qx[f] = rd[20 + f];
            qy[f] = rd[24 + f];
            int nb = ri[f];
            interior[f] = nb >= 0;
            in[f] = nb >= 0 ? grad + (size_t)nb * 2 * DB : eb + (size_t)(-nb - 1) * DB;
        }
    }
};

template <typename ST>
__global__ void __launch_bounds__(64 * kSrcRows, kSrcWaves) k_source(const int *__restrict__ posc, ST *__restrict__ S,
                                                         const ST *__restrict__ grad, const ST *__restrict__ eb,
                                                         const double *__restrict__ Told, const double *__restrict__ Q,
                                                         const int *__restrict__ fsize, const double *__restrict__ srec,
                                                         const int *__restrict__ srei, Modes m, int nc, Rows R,
                                                         const int *__restrict__ cells, int nrows, int with_told) {
    extern __shared__ double lds[];
    int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int c0 = xcd_block() * kSrcRows, t0 = threadIdx.x & 63;
    int DB = R.DB, nB = R.B, nD = DB / nB;
    SrcLds L = src_stage(lds, posc, m, c0, nrows, cells, nB, nD);
    const double *mvg = L.vg, *mtau = L.tau, *mcap = L.cap, *mhr = L.hr, *msx = L.sx, *msy = L.sy;
    int row = c0 + w;
    if (row >= nrows) return;
    int c = cells ? cells[row] : row;
    (void)fsize;
    // the C T_old term of a deviation, or none (added by the sweep instead)
    double To = with_told ? Told[c] - m.eref : 0.0, Qc = Q[c];
    const ST *gc = grad + (size_t)c * 2 * DB;
    const int *pc = L.pos + w * nD;
    SrcFaces<ST> F(srec + (size_t)c * kSrcD, srei + (size_t)c * kSrcI, grad, eb, DB);
    const double *fnx = F.nx, *fny = F.ny, *fav = F.av, *frx = F.rx, *fry = F.ry, *fqx = F.qx, *fqy = F.qy;
    const ST *const *fin = F.in;
    const bool *fint = F.interior;
    DirBand db(t0, nB);
    double vb = mvg[db.b], tau = mtau[db.b], s0 = To * mcap[db.b] + Qc * mhr[db.b] * mtau[db.b];
    // entries in batches of kSrcBatch: every load of a batch is issued before
    // its stores (the stores are counted by the same wait counter)
    for (int tb = t0; tb < DB; tb += kSrcBatch * R.tpr) {
        double so[kSrcBatch];
        size_t sa[kSrcBatch];
#pragma unroll
        for (int j = 0; j < kSrcBatch; ++j) {
            int t = tb + j * R.tpr;
            bool in = t < DB;
            int tt = in ? t : t0, d = in ? db.d : 0, bb = in ? db.b : 0;
            if (R.bstep != 0) {
                vb = mvg[bb];
                tau = mtau[bb];
                s0 = To * mcap[bb] + Qc * mhr[bb] * mtau[bb];
            }
            double sx = msx[d], sy = msy[d];
            // straight-line: every face issues one load before any is used
            // (an incoming face's upwind row: the neighbour's gradient pair,
            // or the aligned pair of boundary values holding entry tt; an
            // outgoing face re-reads the cell's own pair), then the term is
            // selected
            double2 gv = ld2(gc + 2 * tt);
            double dot[kMaxFace2D];
            double2 nv[kMaxFace2D];
#pragma unroll
            for (int f = 0; f < kMaxFace2D; ++f) {
                dot[f] = vb * sx * fnx[f] + vb * sy * fny[f];
                size_t off = dot[f] >= 0 ? (size_t)(gc - fin[f]) + 2 * tt : (fint[f] ? 2 * tt : (tt & ~1));
                // an opaque offset: the compiler must not turn the
                // outgoing lanes' re-read into a branch around the load
                asm volatile("" : "+v"(off));
                nv[f] = ld2(fin[f] + off);
            }
            double gcx = gv.x, gcy = gv.y;
            double s = s0;
#pragma unroll
            for (int f = 0; f < kMaxFace2D; ++f) {
                double a = tau * fav[f] * dot[f];
                double uo = frx[f] * gcx + fry[f] * gcy;
                double ui = fint[f] ? fqx[f] * nv[f].x + fqy[f] * nv[f].y : ((tt & 1) ? nv[f].y : nv[f].x);
                s -= a * (dot[f] >= 0 ? uo : ui);
            }
            so[j] = s;
            sa[j] = in ? ((size_t)d * nc + pc[d]) * nB + bb : ~(size_t)0;
            db.next(R);
        }
#pragma unroll
        for (int j = 0; j < kSrcBatch; ++j)
            if (sa[j] != ~(size_t)0) __builtin_nontemporal_store((ST)so[j], S + sa[j]);  // read by the next launch
    }
}

// k_source for FP32 storage with an even band count and a row length that is
// a multiple of 4: each lane takes two neighbouring bands b, b + 1 (b even)
// of one direction, so that one 16-byte load brings both entries' gradient
// pairs (or the aligned quad of boundary values holding both) and one 8-byte
// store writes both sources.  With one entry per lane the float kernel issues
// as many memory instructions as the FP64 one for half the bytes.  A face
// loads its upwind row unless both entries are outgoing; each entry's
// arithmetic is k_source's.  R.tpr = 128 (64 lanes x 2 entries).
__global__ void __launch_bounds__(64 * kSrcRows, kSrcWaves) k_source_f32x2(
    const int *__restrict__ posc, float *__restrict__ S, const float *__restrict__ grad,
    const float *__restrict__ eb, const double *__restrict__ Told, const double *__restrict__ Q,
    const int *__restrict__ fsize, const double *__restrict__ srec, const int *__restrict__ srei, Modes m, int nc,
    Rows R, const int *__restrict__ cells, int nrows, int with_told) {
    extern __shared__ double lds[];
    int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int c0 = xcd_block() * kSrcRows, t0 = 2 * (int)(threadIdx.x & 63);
    int DB = R.DB, nB = R.B, nD = DB / nB;
    SrcLds L = src_stage(lds, posc, m, c0, nrows, cells, nB, nD);
    const double *mvg = L.vg, *mtau = L.tau, *mcap = L.cap, *mhr = L.hr, *msx = L.sx, *msy = L.sy;
    int row = c0 + w;
    if (row >= nrows) return;
    int c = cells ? cells[row] : row;
    (void)fsize;
    // the C T_old term of a deviation, or none (added by the sweep instead)
    double To = with_told ? Told[c] - m.eref : 0.0, Qc = Q[c];
    const float *gc = grad + (size_t)c * 2 * DB;
    const int *pc = L.pos + w * nD;
    SrcFaces<float> F(srec + (size_t)c * kSrcD, srei + (size_t)c * kSrcI, grad, eb, DB);
    const double *fnx = F.nx, *fny = F.ny, *fav = F.av, *frx = F.rx, *fry = F.ry, *fqx = F.qx, *fqy = F.qy;
    const float *const *fin = F.in;
    const bool *fint = F.interior;
    DirBand db(t0, nB);
    for (int tb = t0; tb < DB; tb += kSrcBatch * R.tpr) {
        double so0[kSrcBatch], so1[kSrcBatch];
        size_t sa[kSrcBatch];
#pragma unroll
        for (int j = 0; j < kSrcBatch; ++j) {
            int t = tb + j * R.tpr;
            bool in = t < DB;
            int tt = in ? t : t0, d = in ? db.d : 0, bb = in ? db.b : 0;
            double vb0 = mvg[bb], vb1 = mvg[bb + 1], tau0 = mtau[bb], tau1 = mtau[bb + 1];
            double s0 = To * mcap[bb] + Qc * mhr[bb] * mtau[bb];
            double s1 = To * mcap[bb + 1] + Qc * mhr[bb + 1] * mtau[bb + 1];
            double sx = msx[d], sy = msy[d];
            float4 gv = *reinterpret_cast<const float4 *>(gc + 2 * tt);
            double dot0[kMaxFace2D], dot1[kMaxFace2D];
            float4 nv[kMaxFace2D];
#pragma unroll
            for (int f = 0; f < kMaxFace2D; ++f) {
                dot0[f] = vb0 * sx * fnx[f] + vb0 * sy * fny[f];
                dot1[f] = vb1 * sx * fnx[f] + vb1 * sy * fny[f];
                bool out = dot0[f] >= 0 && dot1[f] >= 0;
                size_t off = out ? (size_t)(gc - fin[f]) + 2 * tt : (fint[f] ? 2 * tt : (tt & ~3));
                asm volatile("" : "+v"(off));  // keep it a plain load
                nv[f] = *reinterpret_cast<const float4 *>(fin[f] + off);
            }
            double g0x = gv.x, g0y = gv.y, g1x = gv.z, g1y = gv.w;
            bool hi = (tt & 2) != 0;
#pragma unroll
            for (int f = 0; f < kMaxFace2D; ++f) {
                double a0 = tau0 * fav[f] * dot0[f], a1 = tau1 * fav[f] * dot1[f];
                double uo0 = frx[f] * g0x + fry[f] * g0y, uo1 = frx[f] * g1x + fry[f] * g1y;
                double nx0 = nv[f].x, ny0 = nv[f].y, nx1 = nv[f].z, ny1 = nv[f].w;
                double ui0 = fint[f] ? fqx[f] * nx0 + fqy[f] * ny0 : (hi ? nx1 : nx0);
                double ui1 = fint[f] ? fqx[f] * nx1 + fqy[f] * ny1 : (hi ? ny1 : ny0);
                s0 -= a0 * (dot0[f] >= 0 ? uo0 : ui0);
                s1 -= a1 * (dot1[f] >= 0 ? uo1 : ui1);
            }
            so0[j] = s0;
            so1[j] = s1;
            sa[j] = in ? ((size_t)d * nc + pc[d]) * nB + bb : ~(size_t)0;
            db.next(R);
        }
#pragma unroll
        for (int j = 0; j < kSrcBatch; ++j)
            if (sa[j] != ~(size_t)0) st2_nt(S + sa[j], so0[j], so1[j]);
    }
}

// Order 2limit: the limiter of _get_gradient_larger (StaticBTESynthetic.cpp:
// 1372-1422) applied to the stored LSQ gradients.  The range is the cell's
// value, its face neighbours' and the incoming non-specular boundary values
// of side 0 of its boundary faces; the faces are then visited in order, and
// each face whose extrapolated value leaves the range sets the factor anew.
// The stored gradient becomes the limited one (g * limit; the reference keeps
// the factor apart and applies it to r.g).  One wave per cell row.
constexpr int kLimNone = -2147483647 - 1;  // face without a contribution to the range

template <typename ST>
__global__ void __launch_bounds__(64 * kSrcRows) k_limit(const ST *__restrict__ e, ST *__restrict__ grad,
                                                        const ST *__restrict__ eb, const int *__restrict__ fsize,
                                                        const double *__restrict__ srec, const int *__restrict__ slim,
                                                        Modes m, int nc, int DB, int nB) {
    int c = __builtin_amdgcn_readfirstlane(xcd_block() * kSrcRows + (int)(threadIdx.x >> 6));
    int lane = threadIdx.x & 63;
    if (c >= nc) return;  // whole wave
    int nf = fsize[c];
    const double *rd = srec + (size_t)c * kSrcD;
    const int *li = slim + (size_t)c * kSrcI;
    double fnx[kMaxFace2D], fny[kMaxFace2D], fax[kMaxFace2D], fay[kMaxFace2D];
    int fk[kMaxFace2D];
#pragma unroll
    for (int f = 0; f < kMaxFace2D; ++f) {
        fnx[f] = rd[f];
        fny[f] = rd[4 + f];
        fax[f] = rd[12 + f];
        fay[f] = rd[16 + f];
        fk[f] = li[f];
    }
    const ST *ec = e + (size_t)c * DB;
    ST *gc = grad + (size_t)c * 2 * DB;
    for (int t = lane; t < DB; t += 64) {
        int d = t / nB, b = t - d * nB;
        double vb = m.vg[b], sx = m.sx[d], sy = m.sy[d];
        double e0 = ec[t], mx = e0, mn = e0;
#pragma unroll
        for (int f = 0; f < kMaxFace2D; ++f) {
            if (f >= nf) break;
            int k = fk[f];
            double v;
            if (k >= 0) v = e[(size_t)k * DB + t];
            else if (k != kLimNone && vb * sx * fnx[f] + vb * sy * fny[f] < 0) v = eb[(size_t)(-k - 1) * DB + t];
            else continue;
            if (v > mx) mx = v;
            if (v < mn) mn = v;
        }
        double2 g = ld2(gc + 2 * t);
        double L = 1.0;
#pragma unroll
        for (int f = 0; f < kMaxFace2D; ++f) {
            if (f >= nf) break;
            double p = g.x * fax[f] + g.y * fay[f];
            if (p * L + e0 < mn) {
                double y = (mn - e0) / p;
                L = (y * y + 2 * y) / (y * y + y + 2);
            }
            if (p * L + e0 > mx) {
                double y = (mx - e0) / p;
                L = (y * y + 2 * y) / (y * y + y + 2);
            }
        }
        if (L != 1.0) {
            gc[2 * t] = (ST)(g.x * L);
            gc[2 * t + 1] = (ST)(g.y * L);
        }
    }
}

// Boundary values before their first write: the reference allocates them
// uninitialised (StaticBTESynthetic.cpp:310) and only the 2limit limiter of
// the set-up gradients reads such entries; they start as an absolute zero,
// which is -C_b Tref as a deviation.
template <typename ST>
__global__ void k_eb_zero(ST *__restrict__ eb, const double *__restrict__ cap, double eref, size_t n, int nB) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) eb[i] = (ST)(-cap[i % nB] * eref);
}

// The upwind solve of every direction (the Eigen SparseLU solve of
// StaticBTESyntheticSolver.cpp:96-100 on the matrix of _get_coefficient,
// StaticBTESynthetic.cpp:784-840): one workgroup walks all upwind levels of
// one direction for a chunk of Bc bands,
//   e_new = (S - sum_in,interior tau v g_f e_new_nb) / (1 + tau v G_out),
// with g_f = (A/V)(s.n) of the in-faces and G_out the sum over out-faces.
// Levels are separated by workgroup barriers.  The most recent R positions
// stay in an LDS ring.  Each slot's items are loaded two levels ahead (source,
// records, 1/(1 + tau v G)); upwind values that are already final at load
// time are folded into the source then, so that after a barrier only values
// of the two previous levels are read, from LDS.
constexpr int kSweepPF = 2;  // items per slot and level held in registers

struct SweepItem {
    double s, inv;
    int cell;
    int q[kMaxIn], c[kMaxIn];  // pending upwind positions / cells (q < 0: none)
    double a[kMaxIn];          // their coefficients tau v g_f
};

__global__ void __launch_bounds__(kSweepBlock) k_sweep_dir(
    const int *__restrict__ item_cell, const int *__restrict__ levptr, const int *__restrict__ dlev,
    const int *__restrict__ rec_q, const int *__restrict__ rec_c, const double *__restrict__ rec_g,
    const double *__restrict__ rec_G, const double *__restrict__ S, double *__restrict__ e, Modes m, int nc, int nB,
    int nD, int Bc, int nchunk, int R) {
    extern __shared__ double ring[];
    int d = blockIdx.x / nchunk, chunk = blockIdx.x - d * nchunk;
    int tid = threadIdx.x, bl = tid % Bc, slot = tid / Bc, cpp = blockDim.x / Bc;
    int b = chunk * Bc + bl;
    bool act = slot < cpp && b < nB;
    double tv = act ? m.tau[b] * m.vg[b] : 0.0;
    const int *lp = levptr + dlev[d];
    int L = dlev[d + 1] - dlev[d] - 1;
    size_t base = (size_t)d * nc;
    size_t DB = (size_t)nD * nB;
    double *ed = e + (size_t)d * nB + b;  // + cell * DB
    // upwind value at position q (cell c) that is final and visible:
    // from the ring while it holds q (ring positions >= top - R, top = end of
    // the latest level this value may be overwritten by), else from HBM
    auto upwind = [&](int q, int c, int top) -> double {
        return q >= top - R ? ring[(q % R) * Bc + bl] : ed[(size_t)c * DB];
    };
    // item k of the level ending at b1; upwind values at positions < done
    // are final now
    auto load = [&](SweepItem &it, int k, int b1, int done) {
        size_t P = base + k;
        double s = S[P * nB + b];
        it.inv = 1.0 / (1.0 + tv * rec_G[P]);
        it.cell = item_cell[P];
#pragma unroll
        for (int j = 0; j < kMaxIn; ++j) {
            int q = rec_q[P * kMaxIn + j], c = rec_c[P * kMaxIn + j];
            double a = tv * rec_g[P * kMaxIn + j];
            if (q >= 0 && q < done) {
                s -= a * upwind(q, c, b1);
                q = -1;
            }
            it.q[j] = q;
            it.c[j] = c;
            it.a[j] = a;
        }
        it.s = s;
    };
    auto solve = [&](const SweepItem &it, int k, int a1) {
        double s = it.s;
#pragma unroll
        for (int j = 0; j < kMaxIn; ++j)
            if (it.q[j] >= 0) s -= it.a[j] * upwind(it.q[j], it.c[j], a1);
        double en = s * it.inv;
        ed[(size_t)it.cell * DB] = en;
        ring[(k % R) * Bc + bl] = en;
    };
    // one level: solve its items, refill the buffer with level l + 2
    auto step = [&](int l, SweepItem(&pf)[kSweepPF]) {
        int a0 = lp[l], a1 = lp[l + 1];
        if (act) {
#pragma unroll
            for (int i = 0; i < kSweepPF; ++i) {
                int k = a0 + slot + i * cpp;
                if (k < a1) solve(pf[i], k, a1);
            }
            for (int k = a0 + slot + kSweepPF * cpp; k < a1; k += cpp) {
                SweepItem it;
                load(it, k, a1, a0);
                solve(it, k, a1);
            }
            if (l + 2 < L) {
                int c0 = lp[l + 2], c1 = lp[l + 3];
#pragma unroll
                for (int i = 0; i < kSweepPF; ++i) {
                    int k = c0 + slot + i * cpp;
                    if (k < c1) load(pf[i], k, c1, a0);
                }
            }
        }
        __syncthreads();
    };
    SweepItem pa[kSweepPF], pb[kSweepPF];
    if (act)
#pragma unroll
        for (int i = 0; i < kSweepPF; ++i) {
            if (L > 0 && lp[0] + slot + i * cpp < lp[1]) load(pa[i], lp[0] + slot + i * cpp, lp[1], 0);
            if (L > 1 && lp[1] + slot + i * cpp < lp[2]) load(pb[i], lp[1] + slot + i * cpp, lp[2], 0);
        }
    for (int l = 0; l < L; l += 2) {
        step(l, pa);
        if (l + 1 < L) step(l + 1, pb);
    }
}

// Wave-specialised variant of the upwind sweep (same equation as
// k_sweep_dir).  The host splits every level into sub-levels of at most
// kPcCap positions and at most kPcFar "far" upwind reads: upwind values that
// have left the LDS ring of the last R positions by the time the sub-level
// is solved.  Four producer waves stream each sub-level's operands from HBM
// kPcDepth sub-levels ahead (sources, position records, far upwind values)
// and stage them in LDS one sub-level ahead: the count of their loads per
// sub-level is fixed, so each wait covers only loads issued sub-levels ago.
// Twelve consumer waves read only LDS (staged operands and the ring of
// recent values) and store e.  Far values are final when the producers read
// them: the host guarantees (kPcDepth + 1) * kPcCap <= R, so a far position
// lies in a sub-level finished before the producers' read.  The band chunks of one
// direction run the same schedule at about the same pace and write the same
// cells' e rows, so they are placed on one XCD, where those partial-line
// writes merge in its L2.
constexpr int kPcBlock = 1024;
constexpr int kPcProd = 256;    // producer lanes (4 waves)
constexpr int kPcCap = 128;     // positions per sub-level (2 producer lanes each)
constexpr int kPcFar = 16;      // far upwind reads per sub-level on 2D meshes (16 producer lanes each)
constexpr int kPcFar3 = 64;     // the same on tetrahedra (four per producer lane), whose upwind
                                // neighbours more often lie behind the ring
constexpr int kPcBc = 10;       // bands per workgroup at most
constexpr int kPcBpl = 5;       // bands per producer lane (kPcBc / 2)
constexpr int kPcDepth = 4;     // sub-levels in flight
// staging buffer, in doubles (16-byte aligned sections, read as 16-byte
// vectors): S[kPcCap][Bc] | g0,g1,g2,G [kPcCap] | far[kPcFar][Bc] | ints
// cell,q0,q1,q2 [kPcCap] | T_old - Tref [kPcCap] (when the sources leave out
// their C T_old term)
static_assert(kPcCap <= 256 && kPcFar3 <= 255, "sub-level sizes pack into 8 + 8 bits");
constexpr int pc_stage(int far) { return kPcCap * kPcBc + kPcCap * 4 + far * kPcBc + kPcCap * 2 + kPcCap; }

template <typename ST, int FPL>
struct PcProdBuf {
    ST s[kPcBpl];        // loads in the storage type, converted when staged
    int4 r0;             // cell,q0,q1,q2 (even lane) or g2,G (odd lane, as bits)
    int4 r1;             // g0,g1 (both lanes)
    int fc[FPL];         // far upwind cells (far entries fe + 16 i)
    ST fv[FPL];          // far upwind values
    double tp;           // T_old - Tref of the position's cell (with tpos)
};

__device__ __forceinline__ double pc_dbl(int lo, int hi) { return __hiloint2double(hi, lo); }

// Position and far-entry offsets of one sub-level, advanced one sub-level at
// a time from the packed sizes (and held at the last sub-level past the end).
struct PcCursor {
    const unsigned short *sz;
    int l = 0, a0 = 0, f0, np, nf;
    __device__ PcCursor(const unsigned short *s, int fbase) : sz(s), f0(fbase) { unpack(); }
    __device__ void unpack() {
        int v = __builtin_amdgcn_readfirstlane(sz[l]);
        np = (v & 255) + 1;
        nf = v >> 8;
    }
    __device__ void next(int Lsub) {
        if (l + 1 >= Lsub) return;
        a0 += np;
        f0 += nf;
        ++l;
        unpack();
    }
};

// tpos: [dir][upwind position] T_old - Tref, or null when S holds the whole
// source; with it, the source is S + C_b tpos.
template <typename ST, int FAR>
__global__ void __launch_bounds__(kPcBlock) k_sweep_pc(PcSched P, const ST *__restrict__ S,
                                                       ST *__restrict__ e, Modes m, int nc, int nB, int nD, int Bc,
                                                       int nchunk, int R, const double *__restrict__ tpos) {
    extern __shared__ double lds[];
    int bid = gridDim.x % kXcd == 0 ? xcd_block() : (int)blockIdx.x;
    int d = bid / nchunk, chunk = bid - d * nchunk;
    int b0 = chunk * Bc;
    int sl0 = P.dsl[d], Lsub = P.dsl[d + 1] - sl0 - 1;
    double *ring = lds;
    double *stage = lds + (size_t)R * Bc;
    constexpr int kStage = pc_stage(FAR), FPL = FAR / 16;
    unsigned short *sz = reinterpret_cast<unsigned short *>(stage + 2 * kStage);
    for (int i = threadIdx.x; i <= Lsub; i += blockDim.x) sz[i] = P.size[sl0 + i];
    int fbase = P.f0[d];
    __syncthreads();
    size_t base = (size_t)d * nc;
    size_t DB = (size_t)nD * nB;
    int nstep = ((Lsub + kPcDepth - 1) / kPcDepth) * kPcDepth;
    int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (wave < kPcProd / 64) {
        // ---------------- producers
        int pl = threadIdx.x, p = pl >> 1, h = pl & 1;
        int fe = pl >> 4, fb = pl & 15;
        PcProdBuf<ST, FPL> buf[kPcDepth];
        // each of the three producer streams (loads, far cells, staging)
        // walks the sub-levels in order with a cursor of its own
        PcCursor cI(sz, fbase), cF(sz, fbase), cW(sz, fbase);
        // loads of the next sub-level (the last one again past the end:
        // addresses stay valid) into B
        auto issue = [&](PcProdBuf<ST, FPL> &B) {
            int a0 = cI.a0, np = cI.np;
            cI.next(Lsub);
            int pp = p < np ? p : 0;
            size_t Pp = base + a0 + pp;
            const ST *srow = S + Pp * nB + b0;
#pragma unroll
            for (int i = 0; i < kPcBpl; ++i) {
                int jb = h * kPcBpl + i;
                B.s[i] = srow[jb < Bc ? jb : 0];
            }
            const int4 *rr = reinterpret_cast<const int4 *>(P.rec + Pp);
            B.r0 = rr[h ? 2 : 0];
            B.r1 = rr[1];
            if (tpos) B.tp = tpos[Pp];
        };
        // far upwind cells of sub-level lv, loaded kPcDepth - 1 sub-levels
        // before the far values that they address (so that the dependent
        // far-value load waits only for loads that old)
        auto issue_fc = [&](PcProdBuf<ST, FPL> &B) {
            int f0 = cF.f0, nf = cF.nf;
            cF.next(Lsub);
#pragma unroll
            for (int i = 0; i < FPL; ++i) {
                int fi = fe + 16 * i;
                B.fc[i] = P.farcell[f0 + (fi < nf ? fi : 0)];
            }
        };
        auto issue_far = [&](PcProdBuf<ST, FPL> &B) {
#pragma unroll
            for (int i = 0; i < FPL; ++i) B.fv[i] = e[(size_t)B.fc[i] * DB + (size_t)d * nB + b0 + (fb < Bc ? fb : 0)];
        };
        // stage sub-level lv (its loads have arrived) into buffer lv & 1
        auto write = [&](const PcProdBuf<ST, FPL> &B, int lv) {
            if (lv >= Lsub) return;
            int np = cW.np, nf = cW.nf;
            cW.next(Lsub);
            double *st = stage + (size_t)(lv & 1) * kStage;
            double *sS = st, *sG = st + kPcCap * kPcBc, *sF = sG + kPcCap * 4;
            int *sN = reinterpret_cast<int *>(sF + FAR * kPcBc);
            double *sT = sF + FAR * kPcBc + kPcCap * 2;
            if (p < np) {
                if (tpos && h == 0) sT[p] = B.tp;
#pragma unroll
                for (int i = 0; i < kPcBpl; ++i) {
                    int jb = h * kPcBpl + i;
                    if (jb < Bc) sS[p * Bc + jb] = B.s[i];
                }
                if (h == 0) {
                    sN[p * 4 + 0] = B.r0.x;
                    sN[p * 4 + 1] = B.r0.y;
                    sN[p * 4 + 2] = B.r0.z;
                    sN[p * 4 + 3] = B.r0.w;
                    sG[p * 4 + 0] = pc_dbl(B.r1.x, B.r1.y);
                    sG[p * 4 + 1] = pc_dbl(B.r1.z, B.r1.w);
                } else {
                    sG[p * 4 + 2] = pc_dbl(B.r0.x, B.r0.y);
                    sG[p * 4 + 3] = pc_dbl(B.r0.z, B.r0.w);
                }
            }
#pragma unroll
            for (int i = 0; i < FPL; ++i)
                if (fe + 16 * i < nf && fb < Bc) sF[(fe + 16 * i) * Bc + fb] = B.fv[i];
        };
        // prologue: sub-levels 0 .. kPcDepth-1 in flight with their far
        // values, stage 0, far cells of kPcDepth .. 2 kPcDepth-2
#pragma unroll
        for (int q = 0; q < kPcDepth; ++q) {
            issue_fc(buf[q]);
            issue(buf[q]);
        }
#pragma unroll
        for (int q = 0; q < kPcDepth; ++q) issue_far(buf[q]);
        write(buf[0], 0);
#pragma unroll
        for (int q = 0; q + 1 < kPcDepth; ++q) issue_fc(buf[q]);  // sub-levels kPcDepth + q
        __builtin_amdgcn_s_waitcnt(0);  // the loop starts from a known state of outstanding loads
        __syncthreads();
        // iteration j (consumers solve j): stage j+1, far cells of j+2D-1
        // (into the buffer whose far cell was used last iteration), load j+D
        // and its far values.  Every load is consumed kPcDepth - 1 sub-levels
        // after it is issued; the far values of j+D lie in sub-levels up to
        // j-1 (the host checks this), which are finished.
        for (int j0 = 0; j0 < nstep; j0 += kPcDepth) {
#pragma unroll
            for (int q = 0; q < kPcDepth; ++q) {
                int j = j0 + q;
                write(buf[(q + 1) % kPcDepth], j + 1);
                issue_fc(buf[(q + kPcDepth - 1) % kPcDepth]);  // sub-level j + 2 kPcDepth - 1
                issue(buf[q]);                                 // sub-level j + kPcDepth
                issue_far(buf[q]);
                __syncthreads();
            }
        }
    } else {
        // ---------------- consumers
        int cl = threadIdx.x - kPcProd;
        int bl = cl % Bc, slot = cl / Bc, nslot = (kPcBlock - kPcProd) / Bc;
        int b = b0 + bl;
        bool act = slot < nslot && b < nB;
        double tvb = act ? m.tau[b] * m.vg[b] : 0.0;
        double cb = act ? m.cap[b] : 0.0;
        ST *ed = e + (size_t)d * nB + b;
        PcCursor cC(sz, fbase);
        __syncthreads();  // prologue
        for (int j = 0; j < nstep; ++j) {
            int a0 = cC.a0, np = cC.np;
            cC.next(Lsub);
            if (act && j < Lsub) {
                const double *st = stage + (size_t)(j & 1) * kStage;
                const double *sS = st, *sG = st + kPcCap * kPcBc, *sF = sG + kPcCap * 4;
                const int *sN = reinterpret_cast<const int *>(sF + FAR * kPcBc);
                const double *sT = sF + FAR * kPcBc + kPcCap * 2;
                for (int p = slot; p < np; p += nslot) {
                    int4 nq = *reinterpret_cast<const int4 *>(sN + p * 4);
                    double2 g23 = *reinterpret_cast<const double2 *>(sG + p * 4 + 2);
                    double den = 1.0 + tvb * g23.y;
                    int cell = nq.x;
                    double s = sS[p * Bc + bl];
                    if (tpos) s += cb * sT[p];
                    auto up = [&](int q, double gq) {
                        if (q == -1) return;
                        double u = q >= 0 ? ring[(q & (R - 1)) * Bc + bl] : sF[(-2 - q) * Bc + bl];
                        s -= tvb * gq * u;
                    };
                    up(nq.y, sG[p * 4]);
                    up(nq.z, sG[p * 4 + 1]);
                    up(nq.w, g23.x);
                    double en = s / den;
                    ed[(size_t)cell * DB] = en;
                    ring[((a0 + p) & (R - 1)) * Bc + bl] = en;
                }
            }
            __syncthreads();
        }
    }
}

// Inverse-distance vertex values (_set_vertex_energydensity,
// StaticBTESynthetic.cpp:1541-1616): interior nodes from cells, nodes on a
// boundary face from the boundary values of those faces.  One row per node;
// vwt holds the weights (1/d) / sum(1/d).
template <bool OneRow>
__global__ void __launch_bounds__(kRowBlockMax) k_vertex_e(const double *__restrict__ e,
                                                           const double *__restrict__ eb, double *__restrict__ eV,
                                                           const int *__restrict__ vptr, const int *__restrict__ vsrc,
                                                           const double *__restrict__ vwt, int nn, Rows R) {
    int v, t0;
    row_lane<OneRow>(R, v, t0);
    if (v >= nn) return;
    int a = vptr[v], z = vptr[v + 1];
    int DB = R.DB;
    for (int tb = t0; tb < DB; tb += kNT * R.tpr) {
        double sm[kNT];
#pragma unroll
        for (int j = 0; j < kNT; ++j) sm[j] = 0;
        int k = a;
        for (; k + 1 < z; k += 2) {  // two sources: 2 kNT loads in flight
            int s0 = vsrc[k], s1 = vsrc[k + 1];
            const double *r0 = s0 >= 0 ? e + (size_t)s0 * DB : eb + (size_t)(-s0 - 1) * DB;
            const double *r1 = s1 >= 0 ? e + (size_t)s1 * DB : eb + (size_t)(-s1 - 1) * DB;
            double w0 = vwt[k], w1 = vwt[k + 1];
            double x0[kNT], x1[kNT];
#pragma unroll
            for (int j = 0; j < kNT; ++j) {
                int t = tb + j * R.tpr;
                x0[j] = t < DB ? r0[t] : 0.0;
                x1[j] = t < DB ? r1[t] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < kNT; ++j) {
                sm[j] += x0[j] * w0;
                sm[j] += x1[j] * w1;
            }
        }
        if (k < z) {
            int s0 = vsrc[k];
            const double *r0 = s0 >= 0 ? e + (size_t)s0 * DB : eb + (size_t)(-s0 - 1) * DB;
            double w0 = vwt[k];
#pragma unroll
            for (int j = 0; j < kNT; ++j) {
                int t = tb + j * R.tpr;
                if (t < DB) sm[j] += r0[t] * w0;
            }
        }
#pragma unroll
        for (int j = 0; j < kNT; ++j) {
            int t = tb + j * R.tpr;
            if (t >= DB) continue;
            double x = sm[j];
            if (isnan(x)) x = 0;
            eV[(size_t)v * DB + t] = x;
        }
    }
}

struct FaceK {
    const int *kind, *other, *v0, *v1;
    const double *ga, *n0x, *n0y, *mxm, *mym;  // an/|a|^2, face normal n0, m/|m|^2
};

// One row per cell, over every (direction, band):
//   face gradients (_get_face_gradient, StaticBTESynthetic.cpp:1651-1783),
//   the non-Fourier residual (_get_non_fourier, :1883-1937),
//   the lattice-temperature and heat-flux moments (_recover_temperature,
//   _get_heat_flux, :2485-2520), reduced over the row,
//   and the LSQ gradient of the new e for the next transport step
//   (_get_gradient_larger, Order 2, :1243-1340).
// kind: 0 none (specular), 1 interior face owned first by this cell,
//       3 interior face owned first by the neighbour, 2 boundary face.
template <bool OneRow>
__global__ void __launch_bounds__(kRowBlockMax) k_face_moments(
    const double *__restrict__ e, const double *__restrict__ eb, const double *__restrict__ eV,
    double *__restrict__ grad, SlotGeo g, FaceK f, Modes m, const int *__restrict__ gptr, const int *__restrict__ gidx,
    const double *__restrict__ gP, const int *__restrict__ gok, int nnz, double invLx, double *__restrict__ T,
    double *__restrict__ qx, double *__restrict__ qy, double *__restrict__ qz, double *__restrict__ RM, int nc,
    Rows R, int want_grad) {
    __shared__ double sh[kRowBlockMax / 64];
    int c, t0;
    row_lane<OneRow>(R, c, t0);
    bool act = c < nc;
    int cc = act ? c : nc - 1;
    int DB = R.DB, nB = R.B;
    const double *ec = e + (size_t)cc * DB;
    int nf = g.fsize[cc];
    int fk[kMaxFace2D];
    const double *fo[kMaxFace2D], *f0[kMaxFace2D], *f1[kMaxFace2D];
    double fga[kMaxFace2D], fn0x[kMaxFace2D], fn0y[kMaxFace2D], fmx[kMaxFace2D], fmy[kMaxFace2D], fnx[kMaxFace2D],
        fny[kMaxFace2D], fA[kMaxFace2D];
#pragma unroll
    for (int i = 0; i < kMaxFace2D; ++i) {
        int s = cc * 6 + i;
        int kd = i < nf ? f.kind[s] : 0;
        fk[i] = kd;
        bool on = kd != 0;
        fo[i] = on ? (kd == 2 ? eb : e) + (size_t)f.other[s] * DB : ec;
        f0[i] = on ? eV + (size_t)f.v0[s] * DB : ec;
        f1[i] = on ? eV + (size_t)f.v1[s] * DB : ec;
        fga[i] = on ? f.ga[s] : 0.0;
        fn0x[i] = on ? f.n0x[s] : 0.0;
        fn0y[i] = on ? f.n0y[s] : 0.0;
        fmx[i] = on ? f.mxm[s] : 0.0;
        fmy[i] = on ? f.mym[s] : 0.0;
        fnx[i] = on ? g.nx[s] : 0.0;
        fny[i] = on ? g.ny[s] : 0.0;
        fA[i] = on ? g.area[s] : 0.0;
    }
    int ka = gptr[cc], kz = gptr[cc + 1];
    bool ok = gok[cc] != 0;
    double *gcx = grad + (size_t)cc * 2 * DB, *gcy = gcx + 1;  // [t][x, y]
    DirBand db(t0, nB);
    double aT = 0, ax = 0, ay = 0, az = 0, aR = 0;
    for (int tb = t0; tb < DB; tb += kNJ * R.tpr) {
        int tt[kNJ];
        double x[kNJ], wd[kNJ], vx[kNJ], vy[kNJ], tau[kNJ], kl[kNJ];
        bool hr[kNJ], in[kNJ];
#pragma unroll
        for (int j = 0; j < kNJ; ++j) {
            int t = tb + j * R.tpr;
            in[j] = t < DB;
            tt[j] = in[j] ? t : t0;
            int d = in[j] ? db.d : 0, b = in[j] ? db.b : 0;
            db.next(R);
            x[j] = ec[tt[j]];
            wd[j] = in[j] ? m.w[d] : 0.0;
            double vb = m.vg[b];
            vx[j] = vb * m.sx[d];
            vy[j] = vb * m.sy[d];
            double vz = vb * m.sz[d];
            tau[j] = m.tau[b];
            double lrc = m.lrc[b];
            kl[j] = m.kappa * lrc;
            hr[j] = m.hr[b] > 0;
            aT += lrc * x[j] * wd[j];
            ax += vx[j] * wd[j] * x[j];
            ay += vy[j] * wd[j] * x[j];
            az += vz * wd[j] * x[j];
        }
        // face gradients and the non-Fourier residual, one face at a time
        // (3 kNJ loads in flight)
        double acc[kNJ];
#pragma unroll
        for (int j = 0; j < kNJ; ++j) acc[j] = 0;
#pragma unroll
        for (int i = 0; i < kMaxFace2D; ++i) {
            if (fk[i] == 0) continue;
            double o[kNJ], w0[kNJ], w1[kNJ];
#pragma unroll
            for (int j = 0; j < kNJ; ++j) {
                o[j] = fo[i][tt[j]];
                w0[j] = f0[i][tt[j]];
                w1[j] = f1[i][tt[j]];
            }
#pragma unroll
            for (int j = 0; j < kNJ; ++j) {
                double de = fk[i] == 3 ? x[j] - o[j] : o[j] - x[j];
                double gn = de * fga[i];
                double dv = w1[j] - w0[j];
                double gx = gn * fn0x[i] + dv * fmx[i];
                double gy = gn * fn0y[i] + dv * fmy[i];
                double vn = vx[j] * fnx[i] + vy[j] * fny[i];
                double r = fA[i] * (kl[j] * (fnx[i] * gx + fny[i] * gy) - tau[j] * (vx[j] * gx + vy[j] * gy) * vn);
                acc[j] += hr[j] ? r : 0.0;
            }
        }
#pragma unroll
        for (int j = 0; j < kNJ; ++j) aR += wd[j] * acc[j];
        // LSQ gradient: kStencilBatch neighbours x kNJ entries in flight
        double gx[kNJ], gy[kNJ];
#pragma unroll
        for (int j = 0; j < kNJ; ++j) { gx[j] = 0; gy[j] = 0; }
        if (ok) {
            for (int k = ka; k < kz; k += kStencilBatch) {
                double nv[kStencilBatch][kNJ];
#pragma unroll
                for (int q = 0; q < kStencilBatch; ++q) {
                    const double *rn = e + (size_t)gidx[k + q] * DB;
#pragma unroll
                    for (int j = 0; j < kNJ; ++j) nv[q][j] = rn[tt[j]];
                }
#pragma unroll
                for (int q = 0; q < kStencilBatch; ++q) {
                    double p0 = gP[k + q], p1 = gP[(size_t)nnz + k + q];
#pragma unroll
                    for (int j = 0; j < kNJ; ++j) {
                        double d1 = invLx * (nv[q][j] - x[j]);
                        gx[j] += p0 * d1;
                        gy[j] += p1 * d1;
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kNJ; ++j) {
            double a = gx[j], b = gy[j];
            if (isnan(a) || isnan(b)) { a = 0; b = 0; }
            if (act && in[j] && want_grad) {  // Order 1 keeps the gradients zero
                gcx[2 * (size_t)tt[j]] = a;
                gcy[2 * (size_t)tt[j]] = b;
            }
        }
    }
    aT = row_sum(aT, R.tpr, sh);
    ax = row_sum(ax, R.tpr, sh);
    ay = row_sum(ay, R.tpr, sh);
    az = row_sum(az, R.tpr, sh);
    aR = row_sum(aR, R.tpr, sh);
    if (act && t0 == 0) {
        T[c] = aT + m.offT;
        qx[c] = ax + m.offQx;
        qy[c] = ay + m.offQy;
        qz[c] = az + m.offQz;
        RM[c] = aR;
    }
}

// LDS-staged variant of k_vertex_e + k_face_moments over cell tiles
// (tiles.cpp).  One workgroup per tile walks the row in 64-entry slices.  For
// each slice the tile's neighbour rows [cells | boundary entries] are staged
// in LDS and two linear maps of the slice run on the matrix cores:
//   LSQ gradient  G[cell + 8 comp][t] = sum_r C[cell + 8 comp][r] E[r][t]
//                 over the tile's cell rows, one 16-entry block per wave
//                 (v_mfma_f64_16x16x4f64);
//   moments       Y[j][r] += sum_t K_j(t) E[r][t] over the staged rows,
//                 accumulated in the matrix registers over all slices, with
//                 K_j(t) the per-(direction, band) factors of T, q and of the
//                 residual's face terms (v_mfma_f64_4x4x4f64: four j per
//                 product and no idle rows).
// The face vertices' inverse-distance values enter only through their
// moments, so after the last slice a vertex's Y is the weighted sum of its
// sources' Y, and each cell's T, q and non-Fourier residual are short
// combinations of Y entries (tiles.hpp).
constexpr int kTileW = 64;
constexpr int kTileStride = kTileW + 1;  // LDS row stride: the moment product reads 16 rows of one column
constexpr int kTileBlock = 256;
constexpr int kTileWaves = kTileBlock / 64;
constexpr int kTileCells = kTileCellsMax;  // cells per tile
constexpr int kTileRows = 60;              // staged rows per tile: 30.5 KiB of LDS in FP64, five workgroups per CU
constexpr int kStageBatch = 16;            // staged rows in flight per lane
constexpr int kTileJ = 8;                  // moment factors K_0..K_7, two groups of four
static_assert(kTileRows <= kYRows && kTileRows <= kTileRowsMax && kTileRows > 48, "four 16-row blocks of moments");
constexpr int kTileMomentLds = 2 * kYRows * kTileJ * 2;  // doubles of the moment partials and totals
static_assert(kTileMomentLds <= kTileRows * kTileStride, "moment partials fit in the FP64 row buffer");

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// k_tile_face work: the moments (T, q, residual) and / or the LSQ gradient
constexpr int kTileMoments = 1, kTileGradient = 2;

struct TileDev {
    const int *cell0, *rptr, *rsrc, *nh, *vptr, *vs_ptr, *vs_row, *fin_i;
    const double *vs_w, *cmf, *kwf, *fin_d;
};

__device__ __forceinline__ int lane_int(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// workgroups per CU: five in FP64 (30.5 KiB of LDS, 96 VGPRs); FP32 storage
// stages float rows (16 KiB) and runs the LSQ product in float, which fits
// seven (72 VGPRs)
template <typename ST>
__global__ void __launch_bounds__(kTileBlock, sizeof(ST) == 4 ? 7 : 5) k_tile_face(const ST *__restrict__ e, const ST *__restrict__ eb,
                                                          ST *__restrict__ grad, TileDev T,
                                                          double *__restrict__ To, double *__restrict__ qxo,
                                                          double *__restrict__ qyo, double *__restrict__ qzo,
                                                          double *__restrict__ RMo, double4 off, int DB, int ntiles,
                                                          int mode) {
    extern __shared__ double lds[];
    ST *rows = reinterpret_cast<ST *>(lds);  // staged rows in the storage type: no rounding
    bool moments = mode & kTileMoments, gradient = mode & kTileGradient;
    constexpr bool kKEarly = false;
    int tile = xcd_block();
    if (tile >= ntiles) return;  // whole workgroup
    int lane = threadIdx.x & 63, l16 = lane & 15, lq = lane >> 4;
    int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int r0 = T.rptr[tile], nr = T.rptr[tile + 1] - r0;
    int v0 = T.vptr[tile], nv = T.vptr[tile + 1] - v0;
    int c0 = T.cell0[tile], ncell = T.cell0[tile + 1] - c0;
    int nks = (T.nh[tile] + 3) >> 2;  // LSQ k-steps: rows [0, 4 nks)
    // moment work: (16-row block, j group) units, wave w takes group w & 1 of
    // blocks w >> 1 and (w >> 1) + 2; a single block is split over the
    // k-steps instead (halves ks0 = 0, 8)
    int nyb = (nr + 15) >> 4;
    int yg = w & 1;
    int yb = nyb == 1 ? 0 : w >> 1;
    int ks0 = nyb == 1 ? 8 * (w >> 1) : 0, nys = nyb == 1 ? 8 : kTileKs;
    bool y0 = yb < nyb, y1 = nyb > 2 && yb + 2 < nyb;
    // rows [nr, kTileRows) are read by the products with zero weight (LSQ) or
    // into unused columns (moments) and are written by nothing else: clear
    // them once
    for (int r = nr + w; r < kTileRows; r += kTileWaves) rows[r * kTileStride + lane] = 0;
    // LSQ coefficients (A operands), the same for every slice; in the storage
    // type, as the gradients they make are stored in it
    // FP32 storage runs the product on v_mfma_f32_16x16x4f32, whose lane l
    // holds D rows 4 (l / 16) + i (element i): the A rows are taken in the
    // order m -> C row 2 (m / 4) + (m % 4) / 2 + 8 (m % 2), so that a lane
    // holds (x, y) of cells 2 (l / 16) and 2 (l / 16) + 1
    constexpr bool kF32Lsq = sizeof(ST) == 4;
    int arow = lane;
    if constexpr (kF32Lsq) {
        int m = lane & 15;
        arow = (lane & ~15) | (2 * (m >> 2) + ((m & 3) >> 1) + 8 * (m & 1));
    }
    ST cm[kTileKs];
    const double *cmt = T.cmf + (size_t)tile * kTileKs * 64 + arow;
#pragma unroll
    for (int k = 0; k < kTileKs; ++k) cm[k] = k < nks ? (ST)cmt[k * 64] : (ST)0;
    // B operands: Y reads 16 rows of one column, G 4 rows of 16 columns
    // (rows past kTileRows feed only columns of rows past nr: clamp them)
    const ST *yrd = rows + min(16 * yb + l16, kTileRows - 1) * kTileStride + 4 * ks0 + lq;
    const ST *grd = rows + lq * kTileStride + 16 * w + l16;
    const ST *yrd1 = rows + min(16 * yb + 32 + l16, kTileRows - 1) * kTileStride + 4 * ks0 + lq;
    double acc0 = 0.0, acc1 = 0.0;
    int src0 = lane < nr ? T.rsrc[r0 + lane] : 0;
    int nslice = (DB + kTileW - 1) / kTileW;
    for (int slice = 0; slice < nslice; ++slice) {
        int t = slice * kTileW + lane;
        int tt = t < DB ? t : 0;
        // this wave's K fragments of the slice
        double kw[kTileKs];
        auto load_k = [&]() {
            const double *kws = T.kwf + (((size_t)slice * 2 + yg) * kTileKs + ks0) * 64 + lane;
#pragma unroll
            for (int s = 0; s < kTileKs; ++s) kw[s] = s < nys ? kws[s * 64] : 0.0;
        };
        if constexpr (kKEarly) load_k();  // in flight over the staging
        __syncthreads();  // the previous slice is done with the LDS rows
        for (int r = w; r < nr; r += kStageBatch * kTileWaves) {
            ST x[kStageBatch];
#pragma unroll
            for (int j = 0; j < kStageBatch; ++j) {
                int rr = r + j * kTileWaves;
                x[j] = 0;
                if (rr < nr) {
                    int src = lane_int(src0, rr);
                    const ST *row = src >= 0 ? e + (size_t)src * DB : eb + (size_t)(-src - 1) * DB;
                    x[j] = row[tt];
                }
            }
#pragma unroll
            for (int j = 0; j < kStageBatch; ++j) {
                int rr = r + j * kTileWaves;
                if (rr < nr) rows[rr * kTileStride + lane] = x[j];
            }
        }
        // float staging needs the registers: its K loads wait for the rows
        if constexpr (!kKEarly) load_k();
        __syncthreads();
        // moments: k-steps [ks0, ks0 + nys) of group yg, blocks yb and yb + 2
#pragma unroll
        for (int s = 0; s < kTileKs; ++s)
            if (moments && s < nys) {
                if (y0) acc0 = __builtin_amdgcn_mfma_f64_4x4x4f64(kw[s], (double)yrd[4 * s], acc0, 0, 0, 0);
                if (y1)
                    acc1 = __builtin_amdgcn_mfma_f64_4x4x4f64(kw[s], (double)yrd1[4 * s], acc1, 0, 0, 0);
            }
        // LSQ gradient of entries [16 w, 16 w + 16) of the slice
        v4d g = {0.0, 0.0, 0.0, 0.0};
        if constexpr (kF32Lsq) {
            v4f gf = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < kTileKs; ++k)
                if (gradient && k < nks)
                    gf = __builtin_amdgcn_mfma_f32_16x16x4f32((float)cm[k], (float)grd[4 * k * kTileStride], gf, 0, 0,
                                                              0);
            // lane holds (cell 2 lq, x), (2 lq, y), (2 lq + 1, x), (2 lq + 1, y)
            g = v4d{gf[0], gf[2], gf[1], gf[3]};
        } else {
#pragma unroll
            for (int k = 0; k < kTileKs; ++k)
                if (gradient && k < nks)
                    g = __builtin_amdgcn_mfma_f64_16x16x4f64((double)cm[k], (double)grd[4 * k * kTileStride], g, 0, 0,
                                                             0);
        }
        // g[h], g[2 + h] are (x, y) of the lane's cell h: FP64 lq + 4 h (its
        // rows lq + 4 i), FP32 2 lq + h
        int tg = slice * kTileW + 16 * w + l16;
        if (gradient && tg < DB) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                int cl = kF32Lsq ? 2 * lq + h : lq + 4 * h;
                double gx = g[h], gy = g[2 + h];
                if (isnan(gx) || isnan(gy)) {
                    gx = 0;
                    gy = 0;
                }
                // streaming stores: the gradients are read back by the next
                // launch, from HBM, and should not evict staged rows from L2
                if (cl < ncell) st2_nt(grad + (size_t)(c0 + cl) * 2 * DB + 2 * tg, gx, gy);
            }
        }
    }
    if (!moments) return;  // the set-up gradient (whole workgroup)
    // moment partials (lane holds Y[4 yg + lq][16 block + l16]) -> LDS
    // Yp[part][row][j], part 1 only for a single block split over k-steps
    __syncthreads();
    double *Yp = lds, *Y = lds + 2 * kYRows * kTileJ;
    int part = nyb == 1 ? w >> 1 : 0;
    if (y0) Yp[(part * kYRows + 16 * yb + l16) * kTileJ + 4 * yg + lq] = acc0;
    if (y1) Yp[(part * kYRows + 16 * (yb + 2) + l16) * kTileJ + 4 * yg + lq] = acc1;
    __syncthreads();
    for (int idx = threadIdx.x; idx < 16 * nyb * kTileJ; idx += kTileBlock)
        Y[idx] = nyb == 1 ? Yp[idx] + Yp[kYRows * kTileJ + idx] : Yp[idx];
    __syncthreads();
    // face vertices: Y_j = sum_k w_k Y_j[source k] for the residual's j < 4
    // (_set_vertex_energydensity, StaticBTESynthetic.cpp:1541-1616; its NaN -> 0
    // rule applies to the combined value)
    for (int idx = threadIdx.x; idx < 4 * nv; idx += kTileBlock) {
        int i = idx >> 2, j = idx & 3;
        int a = T.vs_ptr[v0 + i], z = T.vs_ptr[v0 + i + 1];
        double sm = 0;
        for (int k = a; k < z; ++k) sm += T.vs_w[k] * Y[T.vs_row[k] * kTileJ + j];
        if (isnan(sm)) sm = 0;
        Y[(kYRows + i) * kTileJ + j] = sm;
    }
    __syncthreads();
    if ((int)threadIdx.x < ncell) {
        int c = c0 + threadIdx.x;
        const int *fi = T.fin_i + (size_t)c * kFinI;
        const double *fd = T.fin_d + (size_t)c * kFinD;
        const double *yo = Y + fi[0] * kTileJ;
        To[c] = yo[4] + off.x;
        qxo[c] = yo[5] + off.y;
        qyo[c] = yo[6] + off.z;
        qzo[c] = yo[7] + off.w;
        double r = 0;
        int nf = fi[1];
        for (int f = 0; f < nf; ++f) {
            const double *ya = Y + fi[2 + 3 * f] * kTileJ, *ya0 = Y + fi[3 + 3 * f] * kTileJ,
                         *ya1 = Y + fi[4 + 3 * f] * kTileJ;
#pragma unroll
            for (int j = 0; j < 4; ++j) r += fd[8 * f + j] * (ya[j] - yo[j]) + fd[8 * f + 4 + j] * (ya1[j] - ya0[j]);
        }
        RMo[c] = r;
    }
}

// Output-only moments: energy-equivalent temperature (_get_total_energy,
// StaticBTESynthetic.cpp:2498-2507) and band temperatures
// (_recover_temperature, :2487-2496).  One wave per cell row, lanes over the
// bands: each lane sums its band over the directions in order (one coalesced
// load of the row's band values per direction) and E is the sum of the band
// sums.
constexpr int kOutRows = 4;

template <typename ST>
__global__ void __launch_bounds__(64 * kOutRows) k_output_moments(const ST *__restrict__ e, Modes m,
                                                                 double *__restrict__ E, double *__restrict__ Tb,
                                                                 int nc, int nB, int nD) {
    int c = xcd_block() * kOutRows + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (c >= nc) return;  // whole wave
    const ST *ec = e + (size_t)c * nD * nB;
    double aE = 0;
    for (int b = lane; b < nB; b += 64) {
        double s = 0;
        for (int d = 0; d < nD; ++d) s += ec[d * nB + b] * m.w[d];
        aE += s;
        Tb[(size_t)b * nc + c] = s / m.cap[b] + m.offW;
    }
    for (int o = 32; o > 0; o >>= 1) aE += __shfl_xor(aE, o);
    if (lane == 0) E[c] = aE / m.cbulk + m.offE;
}

// _set_bound_ee_1 (StaticBTESynthetic.cpp:2083-2232) at np=1 and
// _set_macro_boundary_condition (:2234-2274).  One group of band lanes per
// boundary entry; directions are visited in increasing order, as the
// reference's per-system Allgather makes later directions see earlier
// updates (this matters only for specular partners that are incoming).
template <typename ST>
__global__ void k_boundary(ST *__restrict__ eb, double *__restrict__ tmac, Modes m,
                           const int *__restrict__ ent_list, const int *__restrict__ be_type,
                           const int *__restrict__ be_conn, const int *__restrict__ be_spec,
                           const double *__restrict__ be_nx, const double *__restrict__ be_ny,
                           const double *__restrict__ be_nz, const double *__restrict__ be_tw,
                           const double *__restrict__ be_dT, const double *__restrict__ be_tov, int n_ent, int nB,
                           int nD, int cpw) {
    int lane = lane_id(), sub = lane / nB, b = lane - sub * nB;
    int k = wave_global() * cpw + sub;
    bool act = sub < cpw && k < n_ent;
    int en = ent_list[act ? k : 0];
    double vb = m.vg[b], Cb = m.cap[b], lb = m.lr[b], hb = m.hr[b];
    double nx = be_nx[en], ny = be_ny[en], nz = be_nz[en];
    int ty = be_type[en];
    size_t DB = (size_t)nD * nB;
    ST *row = eb + (size_t)en * DB + b;
    double tm = 0;
    if (act) {
        double diff = 0;
        if (ty == 2) {
            double num = 0, den = 0;
            for (int q = 0; q < nD; ++q) {
                double d1 = vb * m.sx[q] * nx + vb * m.sy[q] * ny + vb * m.sz[q] * nz;
                if (d1 >= 0) {
                    num += row[(size_t)q * nB] * (d1 * m.w[q]);
                    den += (d1 * m.w[q]);
                }
            }
            diff = num / den;
        }
        for (int d = 0; d < nD; ++d) {
            double dot = vb * m.sx[d] * nx + vb * m.sy[d] * ny + vb * m.sz[d] * nz;
            if (hb != 0 && dot < 0) {
                double v = row[(size_t)d * nB];
                if (ty == 1) v = Cb * be_tw[en];
                else if (ty == 2) v = diff;
                else if (ty == 3) v = row[(size_t)be_spec[(size_t)en * nD + d] * nB];
                else if (ty < 0) v = eb[(size_t)be_conn[en] * DB + (size_t)d * nB + b] + Cb * be_dT[en];
                row[(size_t)d * nB] = v;
            }
            tm += lb * row[(size_t)d * nB] / Cb * m.w[d];
        }
    }
    tm = group_sum(tm, b, nB);
    if (act && b == 0) {
        double ov = be_tov[en];
        tmac[en] = isnan(ov) ? tm + m.offMac : ov;
    }
}

// T_old - Tref in each direction's upwind order, for a sweep whose sources
// leave out their C T_old term (item_cell: [dir][position] -> cell).
__global__ void k_tpos(const int *__restrict__ item_cell, const double *__restrict__ Told, double *__restrict__ tpos,
                       double eref, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) tpos[i] = Told[item_cell[i]] - eref;
}

// copy() (StaticBTESynthetic.cpp:2655-2690): old fields.
__global__ void k_copy_old(const double *__restrict__ T, const double *__restrict__ qx, const double *__restrict__ qy,
                           const double *__restrict__ qz, double *__restrict__ To, double *__restrict__ qxo,
                           double *__restrict__ qyo, double *__restrict__ qzo, int nc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    To[i] = T[i];
    qxo[i] = qx[i];
    qyo[i] = qy[i];
    qzo[i] = qz[i];
}

// Cell -> node inverse-distance temperature (StaticBTESynthetic.cpp:2276-2305,
// StaticFourier.cpp:280-309).
__global__ void k_vertex_T(const double *__restrict__ T, double *__restrict__ TV, const int *__restrict__ ptr,
                           const int *__restrict__ src, const double *__restrict__ dist, int nn) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    double t = 0, r = 0;
    for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
        t += T[src[k]] / dist[k];
        r += 1.0 / dist[k];
    }
    TV[v] = t / r;
}

struct RhsK {
    const int *bent, *nonorth, *v0, *v1;
    const double *bcoef, *kA, *a2x, *a2y, *mx, *my, *m2;
};

// Right-hand side of the synthetic solve (_get_Re_macro,
// StaticBTESynthetic.cpp:2307-2407) or of the Fourier solve (_get_Re,
// StaticFourier.cpp:520-613), returned negated for the SPD form.
__global__ void k_rhs(const double *__restrict__ RM, const double *__restrict__ Q, const double *__restrict__ vol,
                      const int *__restrict__ fsize, const double *__restrict__ bval, const double *__restrict__ TV,
                      RhsK k, double *__restrict__ rhs, int nc) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double r = RM ? RM[c] : 0.0;
    r -= Q[c] * vol[c];
    int nf = fsize[c];
    for (int j = 0; j < nf; ++j) {
        int s = c * 6 + j;
        int be = k.bent[s];
        if (be >= 0) r += -k.bcoef[s] * bval[be];
        if (k.nonorth[s]) {
            double dv = TV[k.v1[s]] - TV[k.v0[s]];
            double tx = dv / k.m2[s] * k.mx[s], ty = dv / k.m2[s] * k.my[s];
            r -= k.kA[s] * (k.a2x[s] * tx + k.a2y[s] * ty);
        }
    }
    rhs[c] = -r;
}

// ---- preconditioned conjugate gradients on the SPD matrix -K ------------
__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sh[i];
    __syncthreads();
    return s;  // valid in thread 0
}

// every block sums all partials in the same order (deterministic)
__device__ __forceinline__ double sum_partials(const double *__restrict__ part, int np, double *sh) {
    double v = 0;
    for (int i = threadIdx.x; i < np; i += blockDim.x) v += part[i];
    double s = block_sum(v, sh);
    if (threadIdx.x == 0) sh[8] = s;
    __syncthreads();
    double r = sh[8];
    __syncthreads();
    return r;
}

__global__ void k_cg_init(const int *__restrict__ ptr, const int *__restrict__ col, const double *__restrict__ val,
                          const double *__restrict__ dinv, const double *__restrict__ bvec, const double *__restrict__ x,
                          double *__restrict__ r, double *__restrict__ z, double *__restrict__ p,
                          double *__restrict__ part, int n, int np) {
    __shared__ double sh[16];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0, rr = 0, bb = 0;
    if (i < n) {
        double ax = 0;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) ax += val[k] * x[col[k]];
        double ri = bvec[i] - ax;
        double zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        rz = ri * zi;
        rr = ri * ri;
        bb = bvec[i] * bvec[i];
    }
    double s0 = block_sum(rz, sh), s1 = block_sum(rr, sh), s2 = block_sum(bb, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s0;
        part[np + blockIdx.x] = s1;
        part[2 * np + blockIdx.x] = s2;
    }
}

// scal: [0],[1] rz ping-pong, [2] rr, [3] bb
__global__ void k_cg_init2(const double *__restrict__ part, double *__restrict__ scal, int np) {
    __shared__ double sh[16];
    double rz = sum_partials(part, np, sh);
    double rr = sum_partials(part + np, np, sh);
    double bb = sum_partials(part + 2 * np, np, sh);
    if (threadIdx.x == 0) {
        scal[0] = rz;
        scal[2] = rr;
        scal[3] = bb;
    }
}

__global__ void k_cg_spmv(const int *__restrict__ ptr, const int *__restrict__ col, const double *__restrict__ val,
                          const double *__restrict__ p, double *__restrict__ Ap, double *__restrict__ part, int n) {
    __shared__ double sh[16];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double pap = 0;
    if (i < n) {
        double s = 0;
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += val[k] * p[col[k]];
        Ap[i] = s;
        pap = p[i] * s;
    }
    double t = block_sum(pap, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_cg_update1(const double *__restrict__ part, const double *__restrict__ scal, int slot,
                             double *__restrict__ x, const double *__restrict__ p, double *__restrict__ r,
                             const double *__restrict__ Ap, const double *__restrict__ dinv, double *__restrict__ z,
                             double *__restrict__ part2, int n, int np) {
    __shared__ double sh[16];
    double pap = sum_partials(part, np, sh);
    double alpha = scal[slot] / pap;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0, rr = 0;
    if (i < n) {
        x[i] += alpha * p[i];
        double ri = r[i] - alpha * Ap[i];
        r[i] = ri;
        double zi = dinv[i] * ri;
        z[i] = zi;
        rz = ri * zi;
        rr = ri * ri;
    }
    double s0 = block_sum(rz, sh), s1 = block_sum(rr, sh);
    if (threadIdx.x == 0) {
        part2[blockIdx.x] = s0;
        part2[np + blockIdx.x] = s1;
    }
}

__global__ void k_cg_update2(const double *__restrict__ part2, double *__restrict__ scal, int slot,
                             const double *__restrict__ z, double *__restrict__ p, int n, int np) {
    __shared__ double sh[16];
    double rz = sum_partials(part2, np, sh);
    double beta = rz / scal[slot];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[slot ^ 1] = rz;
    }
    double rr = sum_partials(part2 + np, np, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[2] = rr;
}

// ---- direct solve with the nested-dissection Cholesky factor (ndchol.cpp)
// Per supernode k, F_k = [inv(L_kk); -L_Ak inv(L_kk)] ((ncol + nrow) x ncol).
// Forward, one launch per elimination-tree height: the front
// f = [b(cols) + children; children(rows)] (extend-add of the children's
// update vectors) and [y; u - f_A] = F_k f_cols.  Backward, one launch per
// depth: x = F_k^T [y; x_A].  Each workgroup takes 64 output rows of one
// supernode (lanes = rows, coalesced along the stored layout) and splits the
// reduction over its 4 waves; partial sums are added in a fixed order.
struct NDDev {
    const int *perm, *col0, *ncol, *nrow, *child_ptr, *child_idx, *map, *map_off, *rows, *rows_off, *u_off;
    const long *f_off;
    const double *frm, *fcm;  // F row-major / column-major
};

constexpr int kNdRows = 64;
constexpr int kNdWide = 512;      // levels with fewer row blocks ...
constexpr int kNdWideLen = 512;   // ... and reductions this long take 16 waves per block

template <int W>
__global__ void __launch_bounds__(64 * W) k_nd_forward(const int2 *__restrict__ items, NDDev F,
                                                       const double *__restrict__ b, double *__restrict__ y,
                                                       double *__restrict__ u) {
    extern __shared__ double f[];
    __shared__ double red[W][kNdRows];
    int2 it = items[blockIdx.x];
    int k = it.x, r0 = it.y;
    int nc = F.ncol[k], nr = F.nrow[k], m = nc + nr, c0 = F.col0[k];
    for (int i = threadIdx.x; i < m; i += blockDim.x) f[i] = i < nc ? b[F.perm[c0 + i]] : 0.0;
    __syncthreads();
    for (int q = F.child_ptr[k]; q < F.child_ptr[k + 1]; ++q) {
        int c = F.child_idx[q];
        const int *mp = F.map + F.map_off[c];
        const double *uc = u + F.u_off[c];
        for (int i = threadIdx.x; i < F.nrow[c]; i += blockDim.x) f[mp[i]] += uc[i];
        __syncthreads();
    }
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int i = r0 + lane;
    bool act = i < m;
    const double *Fc = F.fcm + F.f_off[k];  // element (i, a) at a * m + i
    int per = (nc + W - 1) / W, a0 = w * per, a1 = min(nc, a0 + per);
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    if (act) {
        int a = a0;
        // 16 columns' loads in flight before their products; each
        // accumulator still takes its columns in order (a, a + 4, ...)
        for (; a + 15 < a1; a += 16) {
            double x[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) x[q] = Fc[(size_t)(a + q) * m + i];
#pragma unroll
            for (int q = 0; q < 16; q += 4) {
                s0 += x[q] * f[a + q];
                s1 += x[q + 1] * f[a + q + 1];
                s2 += x[q + 2] * f[a + q + 2];
                s3 += x[q + 3] * f[a + q + 3];
            }
        }
        for (; a + 3 < a1; a += 4) {
            s0 += Fc[(size_t)a * m + i] * f[a];
            s1 += Fc[(size_t)(a + 1) * m + i] * f[a + 1];
            s2 += Fc[(size_t)(a + 2) * m + i] * f[a + 2];
            s3 += Fc[(size_t)(a + 3) * m + i] * f[a + 3];
        }
        for (; a < a1; ++a) s0 += Fc[(size_t)a * m + i] * f[a];
    }
    red[w][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (w == 0 && act) {
        double t = 0;
        for (int ww = 0; ww < W; ++ww) t += red[ww][lane];
        if (i < nc) y[c0 + i] = t;
        else u[F.u_off[k] + i - nc] = f[i] + t;
    }
}

template <int W>
__global__ void __launch_bounds__(64 * W) k_nd_backward(const int2 *__restrict__ items, NDDev F,
                                                        const double *__restrict__ y, double *__restrict__ xp,
                                                        double *__restrict__ out) {
    extern __shared__ double v[];
    __shared__ double red[W][kNdRows];
    int2 it = items[blockIdx.x];
    int k = it.x, r0 = it.y;
    int nc = F.ncol[k], nr = F.nrow[k], m = nc + nr, c0 = F.col0[k];
    const int *rw = F.rows + F.rows_off[k];
    for (int j = threadIdx.x; j < m; j += blockDim.x) v[j] = j < nc ? y[c0 + j] : xp[rw[j - nc]];
    __syncthreads();
    int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int a = r0 + lane;
    bool act = a < nc;
    const double *Fr = F.frm + F.f_off[k];  // element (j, a) at j * nc + a
    int per = (m + W - 1) / W, j0 = w * per, j1 = min(m, j0 + per);
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    if (act) {
        int j = j0;
        for (; j + 15 < j1; j += 16) {  // as in k_nd_forward
            double x[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) x[q] = Fr[(size_t)(j + q) * nc + a];
#pragma unroll
            for (int q = 0; q < 16; q += 4) {
                s0 += x[q] * v[j + q];
                s1 += x[q + 1] * v[j + q + 1];
                s2 += x[q + 2] * v[j + q + 2];
                s3 += x[q + 3] * v[j + q + 3];
            }
        }
        for (; j + 3 < j1; j += 4) {
            s0 += Fr[(size_t)j * nc + a] * v[j];
            s1 += Fr[(size_t)(j + 1) * nc + a] * v[j + 1];
            s2 += Fr[(size_t)(j + 2) * nc + a] * v[j + 2];
            s3 += Fr[(size_t)(j + 3) * nc + a] * v[j + 3];
        }
        for (; j < j1; ++j) s0 += Fr[(size_t)j * nc + a] * v[j];
    }
    red[w][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (w == 0 && act) {
        double t = 0;
        for (int ww = 0; ww < W; ++ww) t += red[ww][lane];
        xp[c0 + a] = t;
        out[F.perm[c0 + a]] = t;
    }
}

// partial sums for the outer error check and the macro inner error
__global__ void k_err(const double *__restrict__ T, const double *__restrict__ To, const double *__restrict__ qx,
                      const double *__restrict__ qy, const double *__restrict__ qz, const double *__restrict__ qxo,
                      const double *__restrict__ qyo, const double *__restrict__ qzo, double *__restrict__ part,
                      int nc, int np) {
    __shared__ double sh[16];
    __shared__ double mx[2][kWaves];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sT = 0, mT = 0, sQ = 0, mQ = 0;
    if (i < nc) {
        sT = sqrt(pow(T[i] - To[i], 2));
        mT = fabs(T[i]);
        double h = sqrt(pow(qx[i], 2) + pow(qy[i], 2) + pow(qz[i], 2));
        double h1 = sqrt(pow(qxo[i], 2) + pow(qyo[i], 2) + pow(qzo[i], 2));
        mQ = h;
        sQ = fabs(h - h1);
    }
    double a = block_sum(sT, sh), c2 = block_sum(sQ, sh);
    for (int off = 32; off >= 1; off >>= 1) {
        mT = fmax(mT, __shfl_down(mT, off, 64));
        mQ = fmax(mQ, __shfl_down(mQ, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        mx[0][threadIdx.x >> 6] = mT;
        mx[1][threadIdx.x >> 6] = mQ;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m0 = 0, m1 = 0;
        for (int w = 0; w < kWaves; ++w) {
            m0 = fmax(m0, mx[0][w]);
            m1 = fmax(m1, mx[1][w]);
        }
        part[blockIdx.x] = a;
        part[np + blockIdx.x] = m0;
        part[2 * np + blockIdx.x] = c2;
        part[3 * np + blockIdx.x] = m1;
    }
}

__global__ void k_sqdiff(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ part, int n) {
    __shared__ double sh[16];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0;
    if (i < n) v = pow(a[i] - b[i], 2);
    double s = block_sum(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// StaticFourier::_get_heat_flux, 2D (StaticFourier.cpp:615-710): q = -kappa grad T.
__global__ void k_fourier_flux(const double *__restrict__ T, const int *__restrict__ gptr, const int *__restrict__ gidx,
                               const double *__restrict__ gP, const int *__restrict__ gok,
                               const int *__restrict__ nint, int nnz, double invLx, double kappa,
                               double *__restrict__ qx, double *__restrict__ qy, double *__restrict__ qz, int nc,
                               int dim) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    if (dim == 2 && nint[c] < 2) {
        qx[c] = 0;
        qy[c] = 0;
        return;
    }
    double gx = 0, gy = 0, gz = 0;
    if (gok[c]) {
        for (int k = gptr[c]; k < gptr[c + 1]; ++k) {
            double d1 = invLx * (T[gidx[k]] - T[c]);
            gx += gP[k] * d1;
            gy += gP[(size_t)nnz + k] * d1;
            if (dim == 3) gz += gP[2 * (size_t)nnz + k] * d1;
        }
    }
    qx[c] = -kappa * gx;
    qy[c] = -kappa * gy;
    if (dim == 3) qz[c] = -kappa * gz;
}


// Independent host work items 0 .. n-1 on up to 16 host threads (the set-up
// of per-direction sweep schedules and of 3D tables); fn must not throw.
template <typename F>
void host_parallel(int n, F fn) {
    unsigned hw = std::thread::hardware_concurrency();
    int nt = std::max(1, std::min(n, (int)std::min(16u, hw ? hw : 1u)));
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (int i; (i = next.fetch_add(1)) < n;) fn(i);
        });
    for (auto &x : th) x.join();
}

}  // namespace

// ======================================================================
// host side
// ======================================================================

bool Solver::check(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    if (err_.empty()) err_ = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
    return false;
}


Solver::Solver(const Problem &p, const gbte_options &o) : Pp(renumber(p, cell_new_, node_new_)), P(Pp), opt(o) {
    fp32_ = opt.precision == GBTE_FP32_SWEEP;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        err_ = "no HIP device available";
        return;
    }
    if (!check(hipSetDevice(opt.device), "hipSetDevice")) return;
    // non-blocking: no implicit ordering with the legacy null stream, so a
    // synchronous copy in another thread (another handle's set-up) cannot
    // break this stream's graph capture; every public call ends with a sync
    // of this stream, so the synchronous host copies at call entry and exit
    // never overlap its work.  Default priority.
    if (!check(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking), "hipStreamCreateWithFlags")) return;
    for (hipEvent_t &ev : ev_)
        if (!check(hipEventCreate(&ev), "hipEventCreate")) return;
    if (!check(hipHostMalloc((void **)&h_pin, 64 * sizeof(double)), "hipHostMalloc")) return;
    build();
}

Solver::~Solver() {
    for (int q = 0; q < 2; ++q) {
        if (macro_exec_[q]) (void)hipGraphExecDestroy(macro_exec_[q]);
        if (macro_graph_[q]) (void)hipGraphDestroy(macro_graph_[q]);
    }
    if (h_part) (void)hipHostFree(h_part);
    if (h_err) (void)hipHostFree(h_err);
    if (h_pin) (void)hipHostFree(h_pin);
    if (s2_) (void)hipStreamSynchronize(s2_);
    for (hipEvent_t ev : ev_)
        if (ev) (void)hipEventDestroy(ev);
    for (auto &set : ev2_)
        for (hipEvent_t ev : set)
            if (ev) (void)hipEventDestroy(ev);
    if (s2_) (void)hipStreamDestroy(s2_);
    if (s_) (void)hipStreamDestroy(s_);
}

// The rows of -K with duplicate entries merged (periodic or repeated faces),
// in column order: the matrix the factorization and the PCG take.
static std::vector<std::vector<std::pair<int, double>>> negated_rows(
    const std::vector<std::vector<std::pair<int, double>>> &rows) {
    std::vector<std::vector<std::pair<int, double>>> merged(rows.size());
    for (size_t i = 0; i < rows.size(); ++i) {
        std::vector<std::pair<int, double>> r = rows[i];
        std::sort(r.begin(), r.end(), [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
        for (auto &e : r) {
            if (!merged[i].empty() && merged[i].back().first == e.first) merged[i].back().second += -e.second;
            else merged[i].push_back({e.first, -e.second});
        }
    }
    return merged;
}

bool Solver::build_matrix(const std::vector<std::vector<std::pair<int, double>>> &rows, DevCSR &m, DevND &f,
                          NDFactor *pre, const std::string *pre_err) {
    int n = (int)rows.size();
    std::vector<int> ptr(n + 1, 0), col;
    std::vector<double> val, dinv(n, 0.0);
    std::vector<std::vector<std::pair<int, double>>> merged(n);
    for (int i = 0; i < n; ++i) {
        std::vector<std::pair<int, double>> r = rows[i];
        std::sort(r.begin(), r.end(), [](const std::pair<int, double> &a, const std::pair<int, double> &b) { return a.first < b.first; });
        // merge duplicates (periodic or repeated faces)
        std::vector<std::pair<int, double>> mr;
        for (auto &e : r) {
            if (!mr.empty() && mr.back().first == e.first) mr.back().second += e.second;
            else mr.push_back(e);
        }
        for (auto &e : mr) {
            col.push_back(e.first);
            val.push_back(-e.second);
            merged[i].push_back({e.first, -e.second});
            if (e.first == i) dinv[i] = 1.0 / (-e.second);
        }
        ptr[i + 1] = (int)col.size();
        if (!(dinv[i] > 0) || !std::isfinite(dinv[i])) {
            err_ = "synthetic/Fourier matrix has a non-positive diagonal";
            return false;
        }
    }
    // symmetry check: the PCG needs K symmetric
    for (int i = 0; i < n; ++i)
        for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
            int j = col[k];
            bool found = false;
            for (int q = ptr[j]; q < ptr[j + 1]; ++q)
                if (col[q] == i) {
                    double a = val[k], b = val[q];
                    if (std::fabs(a - b) > 1e-12 * std::max(std::fabs(a), std::fabs(b))) {
                        err_ = "synthetic/Fourier matrix is not symmetric (multi-material meshes are not supported)";
                        return false;
                    }
                    found = true;
                }
            if (!found) {
                err_ = "synthetic/Fourier matrix pattern is not symmetric";
                return false;
            }
        }
    m.n = n;
    if (!upload(m.ptr, ptr) || !upload(m.col, col) || !upload(m.val, val) || !upload(m.dinv, dinv)) return false;
    if (opt.linear_solver != GBTE_LINSOLVE_DIRECT) return true;
    NDFactor F0;
    std::string e = pre ? *pre_err : nd_factor(merged, P.cx, P.cy, P.cz, dim, F0);
    if (!e.empty()) { err_ = e; return false; }
    NDFactor &F = pre ? *pre : F0;
    // LDS of the solve kernels: one vector of the largest front
    if ((size_t)F.max_front * sizeof(double) > 60 * 1024) {
        err_ = "nested-dissection front too large for the direct solve (use the PCG linear solver)";
        return false;
    }
    int nt = (int)F.ncol.size();
    std::vector<int> rows_off(nt + 1), u_off(nt + 1), map_off(F.map_off.begin(), F.map_off.end());
    for (int k = 0; k <= nt; ++k) { rows_off[k] = (int)F.rows_off[k]; u_off[k] = (int)F.u_off[k]; }
    int hmax = 0, dmax = 0;
    for (int k = 0; k < nt; ++k) { hmax = std::max(hmax, F.height[k]); dmax = std::max(dmax, F.depth[k]); }
    // work items (supernode, first output row) per level, 64 rows each
    std::vector<int2> fwd, bwd;
    f.fwd_ptr.assign(1, 0);
    f.fwd_len.clear();
    for (int h = 0; h <= hmax; ++h) {
        int len = 0;
        for (int k = 0; k < nt; ++k)
            if (F.height[k] == h) {
                for (int r = 0; r < F.ncol[k] + F.nrow[k]; r += kNdRows) fwd.push_back(make_int2(k, r));
                len = std::max(len, F.ncol[k]);
            }
        f.fwd_ptr.push_back((int)fwd.size());
        f.fwd_len.push_back(len);
    }
    f.bwd_ptr.assign(1, 0);
    f.bwd_len.clear();
    for (int d = 0; d <= dmax; ++d) {
        int len = 0;
        for (int k = 0; k < nt; ++k)
            if (F.depth[k] == d) {
                for (int r = 0; r < F.ncol[k]; r += kNdRows) bwd.push_back(make_int2(k, r));
                len = std::max(len, F.ncol[k] + F.nrow[k]);
            }
        f.bwd_ptr.push_back((int)bwd.size());
        f.bwd_len.push_back(len);
    }
    f.n = n;
    f.max_front = std::max(F.max_front, 1);
    f.nodes = nt;
    f.u_total = F.u_total;
    return upload(f.perm, F.perm) && upload(f.col0, F.col0) && upload(f.ncol, F.ncol) && upload(f.nrow, F.nrow) &&
           upload(f.child_ptr, F.child_ptr) && upload(f.child_idx, F.child_idx) && upload(f.map, F.map) &&
           upload(f.map_off, map_off) && upload(f.rows, F.rows) && upload(f.rows_off, rows_off) &&
           upload(f.u_off, u_off) && upload(f.f_off, F.f_off) && upload(f.frm, F.frm) && upload(f.fcm, F.fcm) &&
           upload(f.fwd_items, fwd) && upload(f.bwd_items, bwd);
}

bool Solver::build() {
    const Problem &p = P;
    // host time of each set-up phase, reported with the tuner's record
    auto tp = std::chrono::steady_clock::now();
    std::string phases;
    auto mark = [&](const char *what) {
        auto t = std::chrono::steady_clock::now();
        char b[256];
        snprintf(b, sizeof b, "%s%s %.2f s", phases.empty() ? "" : ", ", what,
                 std::chrono::duration<double>(t - tp).count());
        phases += b;
        tp = t;
    };
    nc = p.n_cells; nb = p.n_bounds; ne = 2 * nb; nn = p.n_nodes; nB = p.n_bands; nD = p.n_dirs; nf = p.n_faces;
    dim = p.dim;
    if (p.n_matters != 1) { err_ = "only single-material problems are supported"; return false; }
    if (dim != 2 && dim != 3) { err_ = "2D and 3D meshes only"; return false; }
    if (dim == 3 && p.deck.order == "2limit") {
        err_ = "Order 2limit is not supported on 3D meshes (the limiter runs on the 2D kernels only)";
        return false;
    }
    if (nB < 1 || nB > 64) { err_ = "the number of bands must be between 1 and 64"; return false; }
    for (int ib = 0; ib < nb; ++ib)
        if (p.bound_type[ib] == 4 || p.bound_type[ib] == 5) { err_ = "interface boundaries (type 4/5) are not supported"; return false; }
    cpw = 64 / nB;
    {
        // about kNT row entries per thread: long rows get one block of a
        // multiple of 64 threads, short rows share a 256-thread block
        // when a multiple of 64 lanes is also a multiple of n_bands, every
        // thread keeps one band for the whole row
        int DB = nD * nB, need = (DB + kNT - 1) / kNT;
        if (need > 64) {
            row_tpr = std::min(kRowBlockMax, ((need + 63) / 64) * 64);
            int l = 64 * nB / std::gcd(64, nB);
            if (l <= kRowBlockMax) row_tpr = std::min(kRowBlockMax / l, std::max(1, (row_tpr + l / 2) / l)) * l;
            row_rpb = 1;
        } else {
            row_tpr = 1;
            while (row_tpr < need) row_tpr <<= 1;
            row_rpb = kBlock / row_tpr;
        }
    }
    // ---- modes
    std::vector<double> sx(p.dir_x), sy(p.dir_y), sz(p.dir_z), w(p.weight);
    std::vector<double> vg(p.vg.begin(), p.vg.begin() + nB), tau(p.tau.begin(), p.tau.begin() + nB),
        cap(p.cap.begin(), p.cap.begin() + nB), lr(p.lr.begin(), p.lr.begin() + nB), hr(p.hr.begin(), p.hr.begin() + nB);
    for (int b = 0; b < nB; ++b)
        if (!(vg[b] > 0)) { err_ = "group velocities must be positive"; return false; }
    std::vector<double> lrc(nB);
    for (int b = 0; b < nB; ++b) lrc[b] = lr[b] / cap[b];
    // moments of the equilibrium C_b eref that the deviations leave out
    {
        eref_ = p.deck.tref;
        double sT = 0, sE = 0, sW = 0, sQx = 0, sQy = 0, sQz = 0, sM = 0;
        for (int d = 0; d < nD; ++d) {
            sW += w[d];
            for (int b = 0; b < nB; ++b) {
                sT += w[d] * lrc[b] * cap[b];
                sE += cap[b] * w[d] / P.capacity_bulk[0];
                sQx += w[d] * vg[b] * sx[d] * cap[b];
                sQy += w[d] * vg[b] * sy[d] * cap[b];
                sQz += w[d] * vg[b] * sz[d] * cap[b];
                sM += lr[b] * cap[b] / cap[b] * w[d];
            }
        }
        offT_ = eref_ * sT;
        offE_ = eref_ * sE;
        offW_ = eref_ * sW;
        offQ_[0] = eref_ * sQx;
        offQ_[1] = eref_ * sQy;
        offQ_[2] = eref_ * sQz;
        offMac_ = eref_ * sM;
    }
    if (!upload(d_vg, vg) || !upload(d_tau, tau) || !upload(d_cap, cap) || !upload(d_lr, lr) || !upload(d_hr, hr) ||
        !upload(d_sx, sx) || !upload(d_sy, sy) || !upload(d_sz, sz) || !upload(d_w, w) || !upload(d_lrc, lrc))
        return false;
    // ---- slots
    size_t ns = (size_t)nc * 6;
    std::vector<int> snb(ns, -1), sent(ns, -1);
    std::vector<double> srx(ns, 0), sry(ns, 0), srz(ns, 0), sqx(ns, 0), sqy(ns, 0), sqz(ns, 0);
    std::vector<int> be_cell(ne, -1), be_slot(ne, -1);
    for (int ib = 0; ib < nb; ++ib)
        for (int ic = 0; ic < 2; ++ic) {
            int c = p.bound_cell[ib * 2 + ic];
            if (c < 0) continue;
            int s = c * 6 + p.bound_face[ib * 2 + ic];
            be_cell[ib * 2 + ic] = c;
            be_slot[ib * 2 + ic] = s;
            sent[s] = ib * 2 + ic;
        }
    for (int c = 0; c < nc; ++c)
        for (int j = 0; j < p.face_size[c]; ++j) {
            size_t s = (size_t)c * 6 + j;
            srx[s] = p.face_cx[s] - p.cx[c];
            sry[s] = p.face_cy[s] - p.cy[c];
            srz[s] = p.face_cz[s] - p.cz[c];
            if (p.face_bound[s] == -1) {
                int n = p.face_neighbor[s];
                snb[s] = n;
                sent[s] = -1;
                sqx[s] = p.face_cx[s] - p.cx[n];
                sqy[s] = p.face_cy[s] - p.cy[n];
                sqz[s] = p.face_cz[s] - p.cz[n];
            } else if (sent[s] < 0) {
                err_ = "boundary face without a boundary entry";
                return false;
            }
        }
    if (dim == 2) {
        std::vector<double> srec((size_t)nc * kSrcD, 0.0);
        std::vector<int> srei((size_t)nc * kSrcI, 0), slim((size_t)nc * kSrcI, kLimNone);
        for (int c = 0; c < nc; ++c) {
            if (p.face_size[c] > kMaxFace2D) { err_ = "more than 4 faces per 2D cell"; return false; }
            for (int f = 0; f < kSrcI; ++f) {
                double *r = &srec[(size_t)c * kSrcD];
                srei[(size_t)c * kSrcI + f] = c;
                if (f >= p.face_size[c]) continue;
                size_t s = (size_t)c * 6 + f;
                r[f] = p.face_nx[s];
                r[4 + f] = p.face_ny[s];
                r[8 + f] = p.face_area[s] / p.volume[c];
                r[12 + f] = srx[s];
                r[16 + f] = sry[s];
                r[20 + f] = sqx[s];
                r[24 + f] = sqy[s];
                srei[(size_t)c * kSrcI + f] = snb[s] >= 0 ? snb[s] : -sent[s] - 1;
                // the limiter's range: the neighbour, or side 0 of a
                // non-specular boundary (StaticBTESynthetic.cpp:1380-1399)
                int fb = p.face_bound[s];
                if (fb == -1) slim[(size_t)c * kSrcI + f] = p.face_neighbor[s];
                else if (p.bound_type[fb] != 3) slim[(size_t)c * kSrcI + f] = -(2 * fb) - 1;
            }
        }
        if (!upload(d_srec, srec) || !upload(d_srei, srei) || !upload(d_slim, slim)) return false;
        // cells without / with a boundary face: the first sources read only
        // gradients, so they can be formed before the boundary update
        std::vector<int> ci, cbd;
        for (int c = 0; c < nc; ++c) {
            bool bnd = false;
            for (int f = 0; f < p.face_size[c]; ++f) bnd = bnd || srei[(size_t)c * kSrcI + f] < 0;
            (bnd ? cbd : ci).push_back(c);
        }
        n_src_int = (int)ci.size();
        n_src_bnd = (int)cbd.size();
        if ((!ci.empty() && !upload(d_src_int, ci)) || (!cbd.empty() && !upload(d_src_bnd, cbd))) return false;
    }
    if (!upload(d_fsize, p.face_size) || !upload(d_snb, snb) || !upload(d_sent, sent) || !upload(d_vol, p.volume) ||
        !upload(d_Q, p.heat_source) || !upload(d_snx, p.face_nx) || !upload(d_sny, p.face_ny) || !upload(d_snz, p.face_nz) ||
        !upload(d_sarea, p.face_area) || !upload(d_srx, srx) || !upload(d_sry, sry) || !upload(d_srz, srz) ||
        !upload(d_sqx, sqx) || !upload(d_sqy, sqy) || !upload(d_sqz, sqz))
        return false;
    mark("mode, geometry and boundary tables");
    // ---- per-direction upwind order.  Band 0 decides the upwind graph; the
    // band signs agree except for faces parallel to a direction to rounding,
    // whose coupling coefficient vanishes.
    {
        size_t NI = (size_t)nD * nc;
        std::vector<int> item_cell(NI), posc(NI), levptr, dlev(nD + 1, 0), rec_q(NI * kMaxIn, -1), rec_c(NI * kMaxIn, -1);
        std::vector<double> rec_g(NI * kMaxIn, 0.0), rec_G(NI, 0.0);
        double v0 = vg[0];
        // every direction's schedule is independent host work: one host
        // thread per direction at a time; each writes its own slices, and
        // the level pointers are concatenated afterwards
        std::vector<std::vector<int>> lpd(nD);
        std::vector<std::string> errd(nD);
        std::vector<int> posT(NI);  // [dir][cell] position, transposed into posc below
        host_parallel(nD, [&](int d) {
            auto dotf = [&](size_t s) { return v0 * sx[d] * p.face_nx[s] + v0 * sy[d] * p.face_ny[s] + v0 * sz[d] * p.face_nz[s]; };
            std::vector<int> level(nc, 0), indeg(nc, 0), q, cnt;
            int *pos = posT.data() + (size_t)d * nc;
            q.reserve(nc);
            for (int c = 0; c < nc; ++c)
                for (int j = 0; j < p.face_size[c]; ++j) {
                    size_t s = (size_t)c * 6 + j;
                    if (snb[s] >= 0 && dotf(s) < 0) indeg[c]++;
                }
            for (int c = 0; c < nc; ++c)
                if (indeg[c] == 0) { q.push_back(c); level[c] = 0; }
            size_t head = 0;
            while (head < q.size()) {
                int c = q[head++];
                for (int j = 0; j < p.face_size[c]; ++j) {
                    size_t s = (size_t)c * 6 + j;
                    int n = snb[s];
                    if (n < 0 || !(dotf(s) > 0)) continue;
                    // n's face towards c has the exactly negated normal
                    level[n] = std::max(level[n], level[c] + 1);
                    if (--indeg[n] == 0) q.push_back(n);
                }
            }
            if ((int)q.size() != nc) {
                errd[d] = "the upwind graph of direction " + std::to_string(d) + " has a cycle; an acyclic (e.g. Delaunay) mesh is required";
                return;
            }
            int L = 0;
            for (int c = 0; c < nc; ++c) L = std::max(L, level[c] + 1);
            cnt.assign(L + 1, 0);
            for (int c = 0; c < nc; ++c) cnt[level[c] + 1]++;
            for (int l = 0; l < L; ++l) cnt[l + 1] += cnt[l];
            lpd[d].assign(cnt.begin(), cnt.end());
            for (int c = 0; c < nc; ++c) pos[c] = cnt[level[c]]++;
            size_t base = (size_t)d * nc;
            for (int c = 0; c < nc; ++c) {
                size_t P0 = base + pos[c];
                item_cell[P0] = c;
                int nin = 0;
                double G = 0;
                for (int j = 0; j < p.face_size[c]; ++j) {
                    size_t s = (size_t)c * 6 + j;
                    double sn = sx[d] * p.face_nx[s] + sy[d] * p.face_ny[s] + sz[d] * p.face_nz[s];
                    double gf = p.face_area[s] / p.volume[c] * sn;
                    if (dotf(s) >= 0) {
                        G += gf;
                    } else if (snb[s] >= 0) {
                        if (nin >= kMaxIn) { errd[d] = "a cell has more in-faces than the sweep supports"; return; }
                        rec_q[P0 * kMaxIn + nin] = pos[snb[s]];
                        rec_c[P0 * kMaxIn + nin] = snb[s];
                        rec_g[P0 * kMaxIn + nin] = gf;
                        nin++;
                    }
                }
                rec_G[P0] = G;
            }
        });
        for (int d = 0; d < nD; ++d)
            if (!errd[d].empty()) { err_ = errd[d]; return false; }
        nlev = 0;
        for (int d = 0; d < nD; ++d) {
            int L = (int)lpd[d].size() - 1;
            nlev = std::max(nlev, L);
            dlev[d + 1] = dlev[d] + L + 1;
            levptr.insert(levptr.end(), lpd[d].begin(), lpd[d].end());
        }
        host_parallel((nc + 4095) / 4096, [&](int blk) {
            for (int c = blk * 4096; c < std::min(nc, (blk + 1) * 4096); ++c)
                for (int d = 0; d < nD; ++d) posc[(size_t)c * nD + d] = posT[(size_t)d * nc + c];
        });
        std::vector<int>().swap(posT);
        mark("upwind orders");
        if (dim == 3 && !opt.fallback_kernels) {
            // records of the level-synchronous chunked sweep: every level is
            // non-empty (the kernel's lanes past a level's end re-read its
            // last position) and every upwind cell lies on an earlier level
            // of its direction (both checked)
            std::vector<LvRec> lrec(NI);
            std::vector<char> badd(nD, 0);
            host_parallel(nD, [&](int d) {
                const int *lp = levptr.data() + dlev[d];
                int L = dlev[d + 1] - dlev[d] - 1;
                for (int l = 0; l < L; ++l) {
                    if (lp[l + 1] <= lp[l]) badd[d] = 1;
                    for (int k = lp[l]; k < lp[l + 1]; ++k) {
                        size_t P0 = (size_t)d * nc + k;
                        LvRec &r = lrec[P0];
                        r.cell = item_cell[P0];
                        r.G = rec_G[P0];
                        for (int j = 0; j < kMaxIn; ++j) {
                            int qq = rec_q[P0 * kMaxIn + j];
                            r.c[j] = qq >= 0 ? rec_c[P0 * kMaxIn + j] : -1;
                            r.g[j] = qq >= 0 ? rec_g[P0 * kMaxIn + j] : 0.0;
                            if (qq >= lp[l]) badd[d] = 1;
                        }
                    }
                }
            });
            for (int d = 0; d < nD; ++d)
                if (badd[d]) { err_ = "internal error: empty level, or upwind cell on the same or a later level"; return false; }
            if (!upload(d_lv_rec, lrec)) return false;
        }
        // about one workgroup per CU: split the bands of a direction into
        // chunks; the ring takes 64 KiB of LDS
        sweep_nchunk = std::max(1, std::min(nB, 256 / std::max(1, nD)));
        sweep_bc = (nB + sweep_nchunk - 1) / sweep_nchunk;
        sweep_nchunk = (nB + sweep_bc - 1) / sweep_bc;
        sweep_ring = 8192 / sweep_bc;
        if (!upload(d_item_cell, item_cell) || !upload(d_posc, posc) || !upload(d_levptr, levptr) || !upload(d_dlev, dlev) ||
            !upload(d_rec_q, rec_q) || !upload(d_rec_c, rec_c) || !upload(d_rec_g, rec_g) || !upload(d_rec_G, rec_G))
            return false;
        mark("level records");
        // ---- wave-specialised sweep schedule: sub-levels of at most kPcCap
        // positions and kPcFar far upwind reads (upwind positions older than
        // the ring of the last pc_ring positions when the sub-level runs)
        use_pc = false;
        pc_far = dim == 3 ? kPcFar3 : kPcFar;
        pc_nchunk = std::max((nB + kPcBc - 1) / kPcBc, std::min(nB, (256 + nD - 1) / nD));
        pc_bc = (nB + pc_nchunk - 1) / pc_nchunk;
        pc_nchunk = (nB + pc_bc - 1) / pc_bc;
        pc_ring = 1;
        while ((size_t)pc_ring * 2 * pc_bc * sizeof(double) <= 80 * 1024) pc_ring *= 2;
        // on tetrahedra the wave-specialised kernel is a candidate only for
        // levels about as narrow as its sub-levels: a wider level is cut into
        // several sub-levels that run one after another (the 295k-tet deck's
        // levels hold 590 cells on average, where the chunked kernel is
        // faster by 1.5 to 2.5x), so the schedule is not built there
        bool pc_fits = dim == 2 || (double)NI / std::max(1, dlev[nD] - nD) <= 2.0 * kPcCap;
        if (!opt.fallback_kernels && pc_fits && pc_bc <= kPcBc && (kPcDepth + 1) * kPcCap <= pc_ring) {
            std::vector<PcRec> prec(NI);
            std::vector<int> psl, pfp, pdsl(1, 0), pfar;
            bool fits = true;
            int R = pc_ring;
            // per direction (host threads): sub-level starts, far-entry
            // offsets local to the direction, far cells; concatenated below
            std::vector<std::vector<int>> psl_d(nD), pfp_d(nD), pfar_d(nD);
            std::vector<std::string> errd2(nD);
            host_parallel(nD, [&](int d) {
                std::vector<int> freq(nc, 0), touched;
                const int *lp = levptr.data() + dlev[d];
                int L = dlev[d + 1] - dlev[d] - 1;
                size_t base = (size_t)d * nc;
                std::vector<int> sub(1, 0);  // sub-level starts (direction-local positions)
                for (int l = 0; l < L; ++l) {
                    int s0 = lp[l], far = 0;
                    auto reset = [&]() { for (int t : touched) freq[t] = 0; touched.clear(); far = 0; };
                    reset();
                    for (int k = lp[l]; k < lp[l + 1]; ++k) {
                        // tentative end k+1: ring holds positions >= k+1-R
                        int lo_old = k - R, lo_new = k + 1 - R;
                        int add = (lo_old >= 0) ? freq[lo_old] : 0, own = 0;
                        for (int j = 0; j < kMaxIn; ++j) {
                            int qq = rec_q[(base + k) * kMaxIn + j];
                            if (qq >= 0 && qq < lo_new) own++;
                        }
                        if (k > s0 && (k - s0 + 1 > kPcCap || far + add + own > pc_far)) {
                            sub.push_back(k);
                            s0 = k;
                            reset();
                            add = 0;
                        }
                        far += add;
                        for (int j = 0; j < kMaxIn; ++j) {
                            int qq = rec_q[(base + k) * kMaxIn + j];
                            if (qq < 0) continue;
                            if (qq < lo_new) far++;
                            else { if (!freq[qq]) touched.push_back(qq); freq[qq]++; }
                        }
                    }
                    sub.push_back(lp[l + 1]);
                    for (int t : touched) freq[t] = 0;
                    touched.clear();
                }
                int Ls = (int)sub.size() - 1;
                std::vector<int> &sl = psl_d[d], &fp = pfp_d[d], &fc = pfar_d[d];
                for (int m = 0; m < Ls; ++m) {
                    sl.push_back(sub[m]);
                    fp.push_back((int)fc.size());
                    int a1 = sub[m + 1], nf = 0;
                    for (int k = sub[m]; k < a1; ++k) {
                        size_t P0 = base + k;
                        PcRec &r = prec[P0];
                        r.cell = item_cell[P0];
                        r.G = rec_G[P0];
                        for (int j = 0; j < kMaxIn; ++j) {
                            int qq = rec_q[P0 * kMaxIn + j];
                            r.g[j] = rec_g[P0 * kMaxIn + j];
                            if (qq < 0) r.q[j] = -1;
                            else if (qq >= a1 - R) r.q[j] = qq;
                            else {
                                r.q[j] = -2 - nf++;
                                fc.push_back(rec_c[P0 * kMaxIn + j]);
                            }
                        }
                    }
                    if (nf > pc_far) { errd2[d] = "internal error: sweep sub-level far budget"; return; }
                }
                sl.push_back(sub[Ls]);
                fp.push_back((int)fc.size());
            });
            for (int d = 0; d < nD; ++d)
                if (!errd2[d].empty()) { err_ = errd2[d]; return false; }
            for (int d = 0; d < nD; ++d) {
                int off = (int)pfar.size();
                psl.insert(psl.end(), psl_d[d].begin(), psl_d[d].end());
                for (int f : pfp_d[d]) pfp.push_back(off + f);
                pfar.insert(pfar.end(), pfar_d[d].begin(), pfar_d[d].end());
                pdsl.push_back((int)psl.size());
                std::vector<int>().swap(psl_d[d]);
                std::vector<int>().swap(pfp_d[d]);
                std::vector<int>().swap(pfar_d[d]);
            }
            // validate what the kernel assumes: sub-levels cover the
            // direction in order within kPcCap / kPcFar; a ring read is still
            // in the ring and finished; a far read is finished before the
            // producers read it, kPcDepth sub-levels ahead
            host_parallel(nD, [&](int d) {
                const int *sl = psl.data() + pdsl[d];
                const int *fp = pfp.data() + pdsl[d];
                int Ls = pdsl[d + 1] - pdsl[d] - 1;
                if (sl[0] != 0 || sl[Ls] != nc) { errd2[d] = "internal error: sweep sub-levels"; return; }
                for (int m2 = 0; m2 < Ls; ++m2) {
                    int a0 = sl[m2], a1 = sl[m2 + 1], nf = fp[m2 + 1] - fp[m2];
                    if (a1 <= a0 || a1 - a0 > kPcCap || nf < 0 || nf > pc_far) { errd2[d] = "internal error: sweep sub-level size"; return; }
                    int ready = sl[std::max(0, m2 - kPcDepth)];
                    for (int k = a0; k < a1; ++k) {
                        const PcRec &r = prec[(size_t)d * nc + k];
                        for (int j = 0; j < kMaxIn; ++j) {
                            int qq = r.q[j];
                            if (qq >= 0 && !(qq < a0 && qq >= a1 - pc_ring)) { errd2[d] = "internal error: sweep ring read"; return; }
                            if (qq <= -2) {
                                int f = -2 - qq;
                                int up = rec_q[((size_t)d * nc + k) * kMaxIn + j];
                                if (f >= nf || pfar[fp[m2] + f] != rec_c[((size_t)d * nc + k) * kMaxIn + j] || !(up < ready)) {
                                    errd2[d] = "internal error: sweep far read";
                                    return;
                                }
                            }
                        }
                    }
                }
            });
            for (int d = 0; d < nD; ++d)
                if (!errd2[d].empty()) { err_ = errd2[d]; return false; }
            pfar.push_back(0);  // padding read by clamped loads
            int maxsub0 = 0;
            for (int d = 0; d < nD; ++d) maxsub0 = std::max(maxsub0, pdsl[d + 1] - pdsl[d] - 1);
            // the kernel copies one direction's sub-level tables to LDS
            size_t lds = ((size_t)pc_ring * pc_bc + 2 * pc_stage(pc_far)) * sizeof(double) +
                     (size_t)(maxsub0 + 1) * sizeof(unsigned short);
            int dev_lds = 0;
            (void)hipDeviceGetAttribute(&dev_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, opt.device);
            if (opt.verbose)
                printf("wave-specialised sweep: ring %d, schedule %s, %d sub-levels at most, %zu of %d bytes of LDS\n",
                       pc_ring, fits ? "built" : "rejected", maxsub0, lds, dev_lds);
            if (fits && lds <= (size_t)dev_lds) {
                pc_lds = lds;
                std::vector<unsigned short> psize(psl.size(), 0);
                std::vector<int> pf0(nD);
                for (int d = 0; d < nD; ++d) {
                    pf0[d] = pfp[pdsl[d]];
                    for (int k = pdsl[d]; k + 1 < pdsl[d + 1]; ++k)
                        psize[k] = (unsigned short)((psl[k + 1] - psl[k] - 1) | (pfp[k + 1] - pfp[k]) << 8);
                }
                use_pc = upload(d_pc_rec, prec) && upload(d_pc_size, psize) && upload(d_pc_f0, pf0) &&
                         upload(d_pc_dsl, pdsl) && upload(d_pc_far, pfar);
                if (!use_pc) return false;
                int maxsub = 0;
                for (int d = 0; d < nD; ++d) maxsub = std::max(maxsub, pdsl[d + 1] - pdsl[d] - 1);
                nsub = maxsub;
            }
        }
    }
    mark("wave-specialised schedule");
    // ---- LSQ gradient stencil (_get_CellMatrix_larger + Get_inverse_2D,
    // StaticBTESynthetic.cpp:1000-1064, utility.cpp:401-466); 3D: build3d
    if (dim == 2) {
        int nnz = p.nbr_ptr[nc];
        std::vector<double> gP((size_t)nnz * dim, 0.0);
        std::vector<int> gok(nc, 0);
        double Lx = p.deck.Lx;
        for (int c = 0; c < nc; ++c) {
            int a = p.nbr_ptr[c], z = p.nbr_ptr[c + 1], m = z - a;
            if (m <= 2) continue;
            std::vector<double> jx(m), jy(m);
            for (int k = 0; k < m; ++k) {
                int n = p.nbr_idx[a + k];
                jx[k] = (1.0 / Lx * (p.cx[n] - p.cx[c]));
                jy[k] = (1.0 / Lx * (p.cy[n] - p.cy[c]));
            }
            double a00 = 0, a01 = 0, a10 = 0, a11 = 0;
            for (int k = 0; k < m; ++k) a00 += jx[k] * jx[k];
            for (int k = 0; k < m; ++k) a01 += jx[k] * jy[k];
            for (int k = 0; k < m; ++k) a10 += jy[k] * jx[k];
            for (int k = 0; k < m; ++k) a11 += jy[k] * jy[k];
            double det = a00 * a11 - a01 * a10;
            double i00 = a11 / det, i01 = -a01 / det, i11 = a00 / det, i10 = -a10 / det;
            for (int k = 0; k < m; ++k) {
                double p0 = 0, p1 = 0;
                p0 += i00 * jx[k];
                p0 += i01 * jy[k];
                p1 += i10 * jx[k];
                p1 += i11 * jy[k];
                gP[a + k] = p0;
                gP[(size_t)nnz + a + k] = p1;
            }
            gok[c] = 1;
        }
        // stencils padded to a multiple of kStencilBatch with the cell itself
        // at weight 0 (adds exact zeros), so the device loops run in batches
        std::vector<int> qptr(nc + 1, 0), qidx;
        std::vector<double> q0, q1;
        for (int c = 0; c < nc; ++c) {
            int a = p.nbr_ptr[c], z = p.nbr_ptr[c + 1];
            for (int k = a; k < z; ++k) { qidx.push_back(p.nbr_idx[k]); q0.push_back(gP[k]); q1.push_back(gP[(size_t)nnz + k]); }
            while ((qidx.size() - qptr[c]) % kStencilBatch) { qidx.push_back(c); q0.push_back(0.0); q1.push_back(0.0); }
            qptr[c + 1] = (int)qidx.size();
        }
        gnnz = qptr[nc];
        h_gptr = qptr;
        h_gidx = qidx;
        h_gok = gok;
        q0.insert(q0.end(), q1.begin(), q1.end());
        h_gP = q0;
        if (!upload(d_gptr, qptr) || !upload(d_gidx, qidx) || !upload(d_gP, q0) || !upload(d_gok, gok)) return false;
    }
    // ---- inverse-distance stencils
    {
        std::vector<std::vector<std::pair<int, double>>> ve(nn), vt(nn);
        std::vector<char> bnode(nn, 0);
        for (int c = 0; c < nc; ++c)
            for (int j = 0; j < p.cell_nvert[c]; ++j) {
                int v = p.cell_vert[(size_t)c * 8 + j];
                double dd = std::sqrt(std::pow(p.node_x[v] - p.cx[c], 2) + std::pow(p.node_y[v] - p.cy[c], 2) +
                                      std::pow(p.node_z[v] - p.cz[c], 2));
                ve[v].push_back({c, dd});
                vt[v].push_back({c, dd});
            }
        for (int en = 0; en < ne; ++en) {
            int s = be_slot[en];
            if (s < 0) continue;
            for (int k = 0; k < p.face_nvert[s]; ++k) {
                int v = p.face_vert[(size_t)s * 4 + k];
                if (!bnode[v]) { bnode[v] = 1; ve[v].clear(); }
            }
        }
        for (int en = 0; en < ne; ++en) {
            int s = be_slot[en];
            if (s < 0) continue;
            for (int k = 0; k < p.face_nvert[s]; ++k) {
                int v = p.face_vert[(size_t)s * 4 + k];
                double dd = std::sqrt(std::pow(p.node_x[v] - p.face_cx[s], 2) + std::pow(p.node_y[v] - p.face_cy[s], 2) +
                                      std::pow(p.node_z[v] - p.face_cz[s], 2));
                ve[v].push_back({-(en + 1), dd});
            }
        }
        std::vector<int> vptr(1, 0), vsrc, tptr(1, 0), tsrc;
        std::vector<double> vwt, tdist;
        for (int v = 0; v < nn; ++v) {
            double r = 0;
            for (auto &e : ve[v]) r += 1.0 / e.second;
            for (auto &e : ve[v]) { vsrc.push_back(e.first); vwt.push_back(1.0 / e.second / r); }
            vptr.push_back((int)vsrc.size());
            for (auto &e : vt[v]) { tsrc.push_back(e.first); tdist.push_back(e.second); }
            tptr.push_back((int)tsrc.size());
        }
        h_vptr = vptr;
        h_vsrc = vsrc;
        h_vwt = vwt;
        if (!upload(d_vptr, vptr) || !upload(d_vsrc, vsrc) || !upload(d_vwt, vwt) || !upload(d_tptr, tptr) ||
            !upload(d_tsrc, tsrc) || !upload(d_tdist, tdist))
            return false;
    }
    // ---- face-gradient slot data (2D; build3d holds the 3D face terms)
    if (dim == 2) {
        std::vector<int> kind(ns, 0), other(ns, 0), fv0(ns, 0), fv1(ns, 0);
        std::vector<double> an(ns, 1), aa(ns, 1), n0x(ns, 0), n0y(ns, 0), n0z(ns, 0), mx(ns, 0), my(ns, 0), m2(ns, 1);
        for (int c = 0; c < nc; ++c)
            for (int j = 0; j < p.face_size[c]; ++j) {
                size_t s = (size_t)c * 6 + j;
                int fb = p.face_bound[s];
                size_t s0 = s;
                double ax, ay, az;
                if (fb == -1) {
                    int f = p.face_index[s];
                    int i0 = p.fface_cell[(size_t)f * 2], j0 = p.fface_face[(size_t)f * 2], i1 = p.fface_cell[(size_t)f * 2 + 1];
                    s0 = (size_t)i0 * 6 + j0;
                    kind[s] = (c == i0) ? 1 : 3;
                    other[s] = (c == i0) ? i1 : i0;
                    ax = p.cx[i1] - p.cx[i0]; ay = p.cy[i1] - p.cy[i0]; az = p.cz[i1] - p.cz[i0];
                } else if (p.bound_type[fb] != 3) {
                    int ibc = 0;
                    for (int k = 0; k < 2; ++k)
                        if (p.bound_cell[fb * 2 + k] == c) ibc = k;
                    kind[s] = 2;
                    other[s] = fb * 2 + ibc;
                    ax = p.face_cx[s] - p.cx[c]; ay = p.face_cy[s] - p.cy[c]; az = p.face_cz[s] - p.cz[c];
                } else {
                    continue;
                }
                an[s] = ax * p.face_nx[s0] + ay * p.face_ny[s0] + az * p.face_nz[s0];
                aa[s] = ax * ax + ay * ay + az * az;
                n0x[s] = p.face_nx[s0]; n0y[s] = p.face_ny[s0]; n0z[s] = p.face_nz[s0];
                int v0 = p.face_vert[s0 * 4], v1 = p.face_vert[s0 * 4 + 1];
                fv0[s] = v0; fv1[s] = v1;
                mx[s] = p.node_x[v1] - p.node_x[v0];
                my[s] = p.node_y[v1] - p.node_y[v0];
                m2[s] = mx[s] * mx[s] + my[s] * my[s];
            }
        std::vector<double> ga(ns), mxm(ns), mym(ns);
        for (size_t s = 0; s < ns; ++s) {
            ga[s] = an[s] / aa[s];
            mxm[s] = mx[s] / m2[s];
            mym[s] = my[s] / m2[s];
        }
        h_fkind = kind;
        h_fother = other;
        h_fga = ga;
        h_fn0x = n0x;
        h_fn0y = n0y;
        h_fmxm = mxm;
        h_fmym = mym;
        h_fv0 = fv0;
        h_fv1 = fv1;
        if (!upload(d_fkind, kind) || !upload(d_fother, other) || !upload(d_fv0, fv0) || !upload(d_fv1, fv1) ||
            !upload(d_fga, ga) || !upload(d_fn0x, n0x) || !upload(d_fn0y, n0y) || !upload(d_fmx, mxm) ||
            !upload(d_fmy, mym))
            return false;
    }
    // ---- boundary entries
    {
        std::vector<int> btype(ne, 0), bconn(ne, 0), spec((size_t)ne * nD, 0);
        std::vector<double> bnx(ne, 0), bny(ne, 0), bnz(ne, 0), btw(ne, 0), bdT(ne, 0),
            btov(ne, std::numeric_limits<double>::quiet_NaN());
        h_ent_list.clear();
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < 2; ++ic) {
                int en = ib * 2 + ic, s = be_slot[en];
                if (s < 0) continue;
                h_ent_list.push_back(en);
                int ty = p.bound_type[ib];
                btype[en] = ty;
                bnx[en] = p.face_nx[s]; bny[en] = p.face_ny[s]; bnz[en] = p.face_nz[s];
                if (ty == 1) btw[en] = p.bound_temp[ib] - p.deck.tref;
                if (ty < 0) {
                    int cb = p.bound_connect[ib];
                    bconn[en] = cb * 2 + ic;
                    if (be_slot[cb * 2 + ic] < 0) { err_ = "periodic partner has no cell on this side"; return false; }
                    bdT[en] = (p.bound_temp[ib] - p.bound_temp[cb]);
                }
                if (ty == 3)
                    for (int d = 0; d < nD; ++d) {
                        double d1 = (sx[d] * p.face_nx[s] + sy[d] * p.face_ny[s] + sz[d] * p.face_nz[s]);
                        double rx = sx[d] - p.face_nx[s] * d1 * 2, ry = sy[d] - p.face_ny[s] * d1 * 2, rz = sz[d] - p.face_nz[s] * d1 * 2;
                        double close = 1;
                        int nfd = -1;
                        for (int k = 0; k < nD; ++k) {
                            double len = std::sqrt(std::pow(rx - sx[k], 2) + std::pow(ry - sy[k], 2) + std::pow(rz - sz[k], 2));
                            if (len < close) { nfd = k; close = len; }
                        }
                        if (nfd < 0) { err_ = "specular reflection has no direction within distance 1"; return false; }
                        spec[(size_t)en * nD + d] = nfd;
                    }
            }
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < 2; ++ic) {
                if (be_slot[ib * 2 + ic] < 0) continue;
                if (p.bound_type[ib] < 0 && p.bound_temp[ib] > 0) {
                    btov[ib * 2 + ic] = p.bound_temp[ib] / 2;
                    btov[p.bound_connect[ib] * 2 + ic] = -p.bound_temp[ib] / 2;
                }
            }
        std::vector<int> bcell(be_cell);
        if (!upload(d_ent_list, h_ent_list) || !upload(d_be_cell, bcell) || !upload(d_be_type, btype) ||
            !upload(d_be_conn, bconn) || !upload(d_be_spec, spec) || !upload(d_be_nx, bnx) || !upload(d_be_ny, bny) ||
            !upload(d_be_nz, bnz) || !upload(d_be_tw, btw) || !upload(d_be_dT, bdT) || !upload(d_be_tov, btov) ||
            !upload(d_be_slot, be_slot))  // face slot per entry, for k_bound_out
            return false;
    }
    mark("gradient, vertex, face and boundary tables");
    // ---- macro (set 0) and Fourier (set 1) right-hand sides and matrices
    {
        double kap = p.kappa_bulk[0];
        std::vector<std::vector<std::pair<int, double>>> rowsM(nc), rowsF(nc);
        std::vector<double> fbval(ne, 0.0);
        std::vector<double> btemp(nb, 0.0);
        for (int ib = 0; ib < nb; ++ib) {
            if (p.bound_type[ib] == 1) btemp[ib] = p.bound_temp[ib];
            else if (p.bound_type[ib] < 0 && p.bound_temp[ib] != 0) {
                btemp[ib] = p.bound_temp[ib] / 2;
                btemp[p.bound_connect[ib]] = -p.bound_temp[ib] / 2;
            }
        }
        for (int en = 0; en < ne; ++en) fbval[en] = btemp[en / 2];
        for (int set = 0; set < 2; ++set) {
            std::vector<int> bent(ns, -1), nonorth(ns, 0), v0(ns, 0), v1(ns, 0);
            std::vector<double> bcoef(ns, 0), kA(ns, 0), a2x(ns, 0), a2y(ns, 0), mx(ns, 0), my(ns, 0), m2(ns, 1);
            for (int c = 0; c < nc; ++c) {
                double diag = 0;
                for (int j = 0; j < p.face_size[c]; ++j) {
                    size_t s = (size_t)c * 6 + j;
                    int fb = p.face_bound[s];
                    double ax = 0, ay = 0, az = 0;
                    bool part = false;
                    if (fb == -1) {
                        int n = p.face_neighbor[s];
                        ax = p.cx[n] - p.cx[c]; ay = p.cy[n] - p.cy[c]; az = p.cz[n] - p.cz[c];
                        double tmp = (set == 0) ? kap * p.face_area[s] : 2 * (kap * kap) / (kap + kap) * p.face_area[s];
                        tmp = tmp / (p.face_nx[s] * ax + p.face_ny[s] * ay + p.face_nz[s] * az);
                        diag += -tmp;
                        (set == 0 ? rowsM : rowsF)[c].push_back({n, tmp});
                        kA[s] = (set == 0) ? kap * p.face_area[s] : 2 * (kap * kap) / (kap + kap) * p.face_area[s];
                        part = true;
                    } else {
                        int ty = p.bound_type[fb];
                        bool dir;
                        if (set == 0) dir = ty != 3;
                        else dir = ty == 1 || (ty < 0 && (p.bound_temp[fb] != 0 || p.bound_temp[p.bound_connect[fb]] != 0));
                        if (dir) {
                            int ibc = -1;
                            for (int k = 0; k < 2; ++k)
                                if (p.bound_cell[fb * 2 + k] == c) ibc = k;
                            ax = p.face_cx[s] - p.cx[c]; ay = p.face_cy[s] - p.cy[c]; az = p.face_cz[s] - p.cz[c];
                            double tmp = kap * p.face_area[s];
                            tmp = tmp / (p.face_nx[s] * ax + p.face_ny[s] * ay + p.face_nz[s] * az);
                            diag += -tmp;
                            bcoef[s] = tmp;
                            bent[s] = (set == 0) ? fb * 2 + ibc : fb * 2;
                            kA[s] = kap * p.face_area[s];
                            part = true;
                        }
                    }
                    if (!part) continue;
                    nonorth[s] = 1;
                    double dn = ax * p.face_nx[s] + ay * p.face_ny[s] + az * p.face_nz[s];
                    a2x[s] = p.face_nx[s] - ax / dn;
                    a2y[s] = p.face_ny[s] - ay / dn;
                    int w0 = p.face_vert[s * 4], w1 = p.face_vert[s * 4 + 1];
                    v0[s] = w0; v1[s] = w1;
                    mx[s] = p.node_x[w1] - p.node_x[w0];
                    my[s] = p.node_y[w1] - p.node_y[w0];
                    m2[s] = mx[s] * mx[s] + my[s] * my[s];
                }
                (set == 0 ? rowsM : rowsF)[c].push_back({c, diag});
            }
            if (!upload(d_mbent[set], bent) || !upload(d_mnonorth[set], nonorth) || !upload(d_mv0[set], v0) ||
                !upload(d_mv1[set], v1) || !upload(d_mbcoef[set], bcoef) || !upload(d_mkA[set], kA) ||
                !upload(d_ma2x[set], a2x) || !upload(d_ma2y[set], a2y) || !upload(d_mmx[set], mx) ||
                !upload(d_mmy[set], my) || !upload(d_mm2[set], m2))
                return false;
        }
        if (!upload(d_fourier_bval, fbval)) return false;
        // the two factorizations are independent host work: the Fourier one
        // runs on a second host thread
        NDFactor Ff;
        std::string ef;
        std::thread th;
        bool direct = opt.linear_solver == GBTE_LINSOLVE_DIRECT;
        if (direct) th = std::thread([&]() { ef = nd_factor(negated_rows(rowsF), P.cx, P.cy, P.cz, dim, Ff); });
        bool okM = build_matrix(rowsM, Kmacro, NDmacro, nullptr, nullptr);
        if (direct) th.join();
        if (!okM || !build_matrix(rowsF, Kfour, NDfour, direct ? &Ff : nullptr, direct ? &ef : nullptr)) return false;
    }
    mark("macro and Fourier matrices and factors");
    // ---- 3D tables (solver3d.hip)
    if (dim == 3 && !build3d(be_slot)) return false;
    if (dim == 3) {
        std::string what = "3D tables (" + phase3d_ + ")";
        mark(what.c_str());
    }
    // ---- LDS tiles of the face / vertex / moment kernel (2D)
    if (dim == 2) {
        Tiles T;
        std::string e = opt.fallback_kernels ? std::string("fallback kernels requested")
                                             : build_tiles(p, h_gptr, h_gidx, h_vptr, h_vsrc, h_vwt, h_fkind, h_fother,
                                                           h_fv0, h_fv1, kTileRows, kTileCells, T);
        if (e.find("mismatch") != std::string::npos) { err_ = "internal error: " + e; return false; }
        use_tiles = e.empty();  // otherwise (a stencil too large for LDS) the row kernels run
        if (use_tiles) {
            std::vector<const std::vector<double> *> face8 = {&h_fga, &h_fn0x, &h_fn0y, &h_fmxm,
                                                              &h_fmym, &p.face_nx, &p.face_ny, &p.face_area};
            std::string e2 = pack_tile_mfma(p, h_gptr, h_gP, h_gok, h_fkind, face8, 1.0 / P.deck.Lx, T);
            use_tiles = e2.empty();
        }
        if (use_tiles) {
            n_tiles = T.n_tiles;
            tile_rows = T.max_rows;
            // LDS: row slices, then the vertex stencils (ptr | rows | weights)
            tile_lds = std::max((size_t)kTileRows * kTileStride * (fp32_ ? sizeof(float) : sizeof(double)),
                                (size_t)kTileMomentLds * sizeof(double));
        }
        if (use_tiles) {
            // K_j(t) of the moment product, as A fragments of
            // v_mfma_f64_4x4x4f64 per slice, j group g and k-step: lane l
            // holds K_{4 g + l % 4}(64 slice + 4 k + l / 16) (the same for the
            // four blocks l / 4 % 4), zero past the row
            int DB = nD * nB, nslice = (DB + kTileW - 1) / kTileW;
            std::vector<double> kwf((size_t)nslice * 2 * kTileKs * 64, 0.0);
            double kap = P.kappa_bulk[0];
            for (int sl = 0; sl < nslice; ++sl)
              for (int gq = 0; gq < 2; ++gq)
                for (int k = 0; k < kTileKs; ++k)
                    for (int l = 0; l < 64; ++l) {
                        int j = 4 * gq + l % 4, t = sl * kTileW + 4 * k + l / 16;
                        if (t >= DB) continue;
                        int d = t / nB, b = t - d * nB;
                        double wd = w[d], vb = vg[b], h = hr[b] > 0 ? 1.0 : 0.0, tv2 = tau[b] * vb * vb;
                        double K[kTileJ] = {h * wd * kap * lrc[b],        h * wd * tv2 * sx[d] * sx[d],
                                            h * wd * tv2 * sx[d] * sy[d], h * wd * tv2 * sy[d] * sy[d],
                                            wd * lrc[b],                  wd * vb * sx[d],
                                            wd * vb * sy[d],              wd * vb * sz[d]};
                        kwf[(((size_t)sl * 2 + gq) * kTileKs + k) * 64 + l] = K[j];
                    }
            if (!upload(d_t_cell0, T.cell0) || !upload(d_t_rptr, T.rptr) || !upload(d_t_rsrc, T.rsrc) ||
                !upload(d_t_nh, T.nh) || !upload(d_t_vptr, T.vptr) || !upload(d_t_vs_ptr, T.vs_ptr) ||
                !upload(d_t_vs_row, T.vs_row) || !upload(d_t_vs_w, T.vs_w) ||
                !upload(d_t_fin_i, T.fin_i) || !upload(d_t_cmf, T.cmf) ||
                !upload(d_t_kwf, kwf) || !upload(d_t_fin_d, T.fin_d))
                return false;
            if (opt.verbose) {
                int rmax = 0;
                long rsum = 0, hsum = 0;
                for (int t = 0; t < n_tiles; ++t) {
                    int nr = T.rptr[t + 1] - T.rptr[t];
                    rmax = std::max(rmax, nr);
                    rsum += nr;
                    hsum += T.nh[t];
                }
                printf("[gbte] %d tiles: mean %.1f cell rows, mean %.1f / max %d staged rows\n", n_tiles,
                       (double)hsum / n_tiles, (double)rsum / n_tiles, rmax);
            }
        }
        // h_gptr / h_gidx / h_gP / h_gok and h_vptr / h_vsrc / h_vwt stay:
        // the adjoint transposes the same LSQ and inverse-distance stencils
        std::vector<int>().swap(h_fkind);
        std::vector<int>().swap(h_fother);
        std::vector<int>().swap(h_fv0);
        std::vector<int>().swap(h_fv1);
        for (auto *v : {&h_fga, &h_fn0x, &h_fn0y, &h_fmxm, &h_fmym}) std::vector<double>().swap(*v);
    }
    // ---- state
    size_t nE = (size_t)nD * nc * nB;
    // FP32 sweeps: e, its gradient, the boundary values and S in float
    // (buffers sized in doubles)
    if (fp32_ && (opt.fallback_kernels || (!use_pc && !(dim == 3 && d_lv_rec.p)) || (dim == 2 && !use_tiles))) {
        err_ = "FP32 sweeps need the tiled face kernel and the wave-specialised sweep (2D) or a sweep kernel that "
               "takes float storage (unsupported for this mesh or with fallback kernels)";
        return false;
    }
    auto half = [&](size_t n) { return fp32_ ? (n + 1) / 2 : n; };
    size_t nEB = (size_t)nD * ne * nB;
    // gradients: per component in 2D; on tetrahedra the projections on the
    // four face offsets
    if (!alloc(d_e, half(nE)) || !alloc(d_grad, half(nE * (dim == 3 ? 4 : dim))) || !alloc(d_eb, half(nEB + 2)) ||
        !alloc(d_eV, dim == 2 ? (size_t)nD * nn * nB : 0) || !alloc(d_T, nc) || !alloc(d_Told, nc) || !alloc(d_T1, nc) ||
        !alloc(d_TV, nn) || !alloc(d_E, nc) || !alloc(d_qx, nc) || !alloc(d_qy, nc) || !alloc(d_qz, nc) ||
        !alloc(d_qxo, nc) || !alloc(d_qyo, nc) || !alloc(d_qzo, nc) || !alloc(d_RM, nc) || !alloc(d_Tb, (size_t)nB * nc) ||
        !alloc(d_tmac, ne) || !alloc(d_rhs, nc) || !alloc(d_r, nc) || !alloc(d_z, nc) || !alloc(d_p, nc) ||
        !alloc(d_Ap, nc) || !alloc(d_S, half(nE)))
        return false;
    // tetrahedra: the gradient kernels write a face projection only where it
    // is read (the out-faces per direction); the other entries stay zero
    if (dim == 3 && !check(hipMemsetAsync(d_grad.p, 0, d_grad.n * sizeof(double), s_), "zero face projections"))
        return false;
    if (use_pc && dim == 2 && !alloc(d_tpos, (size_t)nD * nc)) return false;
    if (opt.linear_solver == GBTE_LINSOLVE_DIRECT &&
        (!alloc(d_ndy, nc) || !alloc(d_ndx, nc) || !alloc(d_ndu, (size_t)std::max<long>(1, std::max(NDmacro.u_total, NDfour.u_total)))))
        return false;
    red_blocks = blocks_for(nc, kBlock);
    if (!check(hipHostMalloc((void **)&h_part, (size_t)2 * red_blocks * sizeof(double)), "hipHostMalloc") ||
        !check(hipHostMalloc((void **)&h_err, (size_t)4 * red_blocks * sizeof(double)), "hipHostMalloc") ||
        !alloc(d_mpart, (size_t)2 * red_blocks))
        return false;
    if (!alloc(d_part, (size_t)red_blocks * 4 + 8) || !alloc(d_part2, (size_t)red_blocks * 4 + 8)) return false;
    if (!check(hipMemsetAsync(d_eb.p, 0, d_eb.n * sizeof(double), s_), "hipMemsetAsync") ||
        !check(hipMemsetAsync(d_qz.p, 0, nc * sizeof(double), s_), "hipMemsetAsync"))
        return false;
    if (P.deck.order == "2limit") {
        size_t n = nEB;
        dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
        if (fp32_)
            hipLaunchKernelGGL(k_eb_zero<float>, grid, dim3(kBlock), 0, s_, reinterpret_cast<float *>(d_eb.p),
                               d_cap.p, eref_, n, nB);
        else
            hipLaunchKernelGGL(k_eb_zero<double>, grid, dim3(kBlock), 0, s_, d_eb.p, d_cap.p, eref_, n, nB);
        if (!check(hipGetLastError(), "boundary value set-up")) return false;
    }
    // initial temperature and flux (Solution.cpp:78-90)
    if (!check(hipMemcpy(d_T.p, p.initial_temp.data(), nc * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        !check(hipMemsetAsync(d_qx.p, 0, nc * sizeof(double), s_), "hipMemsetAsync") ||
        !check(hipMemsetAsync(d_qy.p, 0, nc * sizeof(double), s_), "hipMemsetAsync"))
        return false;
    // interior face count per cell (the Fourier flux's LSQ switch)
    {
        std::vector<int> nint(nc, 0);
        for (int c = 0; c < nc; ++c)
            for (int j = 0; j < p.face_size[c]; ++j)
                if (p.face_bound[(size_t)c * 6 + j] == -1) nint[c]++;
        if (!upload(d_nint, nint)) return false;
    }
    mark("tiles and state");
    if (dim == 3 && !tune3d()) return false;
    if (dim == 3) mark("tuner");
    tune_log_ = "set-up (host): " + phases + "\n" + tune_log_;
    return check(hipStreamSynchronize(s_), "build sync");
}

bool Solver::reduce_host(const double *d_partials, int n, std::vector<double> &out, int stride) {
    std::vector<double> h((size_t)n * stride);
    if (!check(hipMemcpyAsync(h.data(), d_partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, s_), "D2H") ||
        !check(hipStreamSynchronize(s_), "sync"))
        return false;
    out = h;
    return true;
}

void Solver::launch_vertex_T(const double *T, double *TV) {
    hipLaunchKernelGGL(k_vertex_T, dim3(blocks_for(nn, kBlock)), dim3(kBlock), 0, s_, T, TV, d_tptr.p, d_tsrc.p,
                       d_tdist.p, nn);
}

void Solver::launch_rhs(bool macro, const double *bval, double *rhs) {
    if (dim == 3) {
        launch_rhs3(macro, bval, rhs);
        return;
    }
    int set = macro ? 0 : 1;
    RhsK k{d_mbent[set].p, d_mnonorth[set].p, d_mv0[set].p, d_mv1[set].p, d_mbcoef[set].p, d_mkA[set].p,
           d_ma2x[set].p, d_ma2y[set].p, d_mmx[set].p, d_mmy[set].p, d_mm2[set].p};
    hipLaunchKernelGGL(k_rhs, dim3(blocks_for(nc, kBlock)), dim3(kBlock), 0, s_, macro ? d_RM.p : nullptr, d_Q.p,
                       d_vol.p, d_fsize.p, bval, d_TV.p, k, rhs, nc);
}

bool Solver::inner_error(const double *a, const double *b, double *out) {
    hipLaunchKernelGGL(k_sqdiff, dim3(red_blocks), dim3(kBlock), 0, s_, a, b, d_part.p, nc);
    std::vector<double> h;
    if (!reduce_host(d_part.p, red_blocks, h, 1)) return false;
    double s = 0;
    for (double v : h) s += v;
    *out = s / nc;
    return true;
}

// One pass of the synthetic loop with the direct solver, T1 = T, T = K^-1
// rhs(T1), as one HIP graph (about two dozen small kernels): enqueued without
// a sync; graph q leaves the pass's squared-change partials in h_part + q
// red_blocks once the stream has reached it (macro_inner reads them).
bool Solver::macro_launch(int q) {
    return macro_capture(q) && check(hipGraphLaunch(macro_exec_[q], s_), "hipGraphLaunch");
}

// Record synthetic pass q as a graph (once per handle).  Nothing runs.
bool Solver::macro_capture(int q) {
    if (!macro_exec_[q]) {
        if (!check(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) return false;
        hipLaunchKernelGGL(k_copy, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_T1.p, nc);
        launch_vertex_T(d_T.p, d_TV.p);
        launch_rhs(true, d_tmac.p, d_rhs.p);
        bool ok = nd_solve(NDmacro, d_rhs.p, d_T.p);
        double *dp = d_mpart.p + (size_t)q * red_blocks;
        hipLaunchKernelGGL(k_sqdiff, dim3(red_blocks), dim3(kBlock), 0, s_, d_T1.p, d_T.p, dp, nc);
        ok = check(hipMemcpyAsync(h_part + (size_t)q * red_blocks, dp, (size_t)red_blocks * sizeof(double),
                                  hipMemcpyDeviceToHost, s_),
                   "D2H") &&
             ok;
        hipGraph_t g = nullptr;
        if (!check(hipStreamEndCapture(s_, &g), "hipStreamEndCapture") || !ok) return false;
        macro_graph_[q] = g;
        if (!check(hipGraphInstantiate(&macro_exec_[q], g, nullptr, nullptr, 0), "hipGraphInstantiate")) return false;
    }
    return true;
}

int Solver::prepare() {
    if (!ok()) return GBTE_HIP_ERROR;
    if (!check(hipSetDevice(opt.device), "hipSetDevice")) return GBTE_HIP_ERROR;
    if (opt.linear_solver == GBTE_LINSOLVE_DIRECT && P.deck.method != "Sequential")
        if (!macro_capture(0) || !macro_capture(1)) return GBTE_HIP_ERROR;
    return GBTE_OK;
}

// mean squared change of the pass run by graph q (after a sync), summed like
// inner_error
double Solver::macro_inner(int q) const {
    double sum = 0;
    for (int i = 0; i < red_blocks; ++i) sum += h_part[(size_t)q * red_blocks + i];
    return sum / nc;
}

bool Solver::linsolve(const DevCSR &K, const DevND &f, double *d_b, double *d_x, int *iters) {
    if (opt.linear_solver == GBTE_LINSOLVE_DIRECT) {
        if (iters) *iters = 0;
        return nd_solve(f, d_b, d_x);
    }
    return cg(K, d_b, d_x, iters);
}

bool Solver::nd_solve(const DevND &f, const double *d_b, double *d_x) {
    NDDev F{f.perm.p, f.col0.p, f.ncol.p, f.nrow.p, f.child_ptr.p, f.child_idx.p, f.map.p, f.map_off.p,
            f.rows.p, f.rows_off.p, f.u_off.p, f.f_off.p, f.frm.p, f.fcm.p};
    size_t lds = (size_t)f.max_front * sizeof(double);
    // levels of fewer row blocks than the chip has CUs whose reductions are
    // long (the large separators near the root) split each block's reduction
    // over 16 waves instead of 4, for more loads in flight, where the front
    // and the 16 waves' partial sums fit 64 KB of LDS
    bool wide = lds + 16 * kNdRows * sizeof(double) <= 65536;
    for (size_t h = 0; h + 1 < f.fwd_ptr.size(); ++h) {
        int cnt = f.fwd_ptr[h + 1] - f.fwd_ptr[h];
        if (cnt == 0) continue;
        if (wide && cnt < kNdWide && f.fwd_len[h] >= kNdWideLen)
            hipLaunchKernelGGL(k_nd_forward<16>, dim3(cnt), dim3(1024), lds, s_, f.fwd_items.p + f.fwd_ptr[h], F, d_b,
                               d_ndy.p, d_ndu.p);
        else
            hipLaunchKernelGGL(k_nd_forward<kWaves>, dim3(cnt), dim3(kBlock), lds, s_, f.fwd_items.p + f.fwd_ptr[h], F,
                               d_b, d_ndy.p, d_ndu.p);
    }
    for (size_t d = 0; d + 1 < f.bwd_ptr.size(); ++d) {
        int cnt = f.bwd_ptr[d + 1] - f.bwd_ptr[d];
        if (cnt == 0) continue;
        if (wide && cnt < kNdWide && f.bwd_len[d] >= kNdWideLen)
            hipLaunchKernelGGL(k_nd_backward<16>, dim3(cnt), dim3(1024), lds, s_, f.bwd_items.p + f.bwd_ptr[d], F,
                               d_ndy.p, d_ndx.p, d_x);
        else
            hipLaunchKernelGGL(k_nd_backward<kWaves>, dim3(cnt), dim3(kBlock), lds, s_, f.bwd_items.p + f.bwd_ptr[d],
                               F, d_ndy.p, d_ndx.p, d_x);
    }
    return check(hipGetLastError(), "nested-dissection solve");
}

bool Solver::cg(const DevCSR &K, double *d_b, double *d_x, int *iters) {
    int n = K.n, np = red_blocks;
    double rtol = opt.macro_rtol > 0 ? opt.macro_rtol : 1e-14;
    double *scal = d_part2.p + (size_t)np * 2;  // 4 scalars after the partials
    hipLaunchKernelGGL(k_cg_init, dim3(np), dim3(kBlock), 0, s_, K.ptr.p, K.col.p, K.val.p, K.dinv.p, d_b, d_x, d_r.p,
                       d_z.p, d_p.p, d_part.p, n, np);
    hipLaunchKernelGGL(k_cg_init2, dim3(1), dim3(kBlock), 0, s_, d_part.p, scal, np);
    if (!check(hipMemcpyAsync(h_pin, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, s_), "D2H") ||
        !check(hipStreamSynchronize(s_), "sync"))
        return false;
    double bb = h_pin[3], rr = h_pin[2];
    double tol2 = rtol * rtol * bb;
    int it = 0, max_it = std::max(1000, 20 * n);
    int slot = 0;
    double prev = rr;
    int stall = 0;
    while (rr > tol2 && rr > 0 && it < max_it) {
        for (int k = 0; k < 16 && it < max_it; ++k, ++it) {
            hipLaunchKernelGGL(k_cg_spmv, dim3(np), dim3(kBlock), 0, s_, K.ptr.p, K.col.p, K.val.p, d_p.p, d_Ap.p,
                               d_part.p, n);
            hipLaunchKernelGGL(k_cg_update1, dim3(np), dim3(kBlock), 0, s_, d_part.p, scal, slot, d_x, d_p.p, d_r.p,
                               d_Ap.p, K.dinv.p, d_z.p, d_part2.p, n, np);
            hipLaunchKernelGGL(k_cg_update2, dim3(np), dim3(kBlock), 0, s_, d_part2.p, scal, slot, d_z.p, d_p.p, n, np);
            slot ^= 1;
        }
        if (!check(hipMemcpyAsync(h_pin, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, s_), "D2H") ||
            !check(hipStreamSynchronize(s_), "sync"))
            return false;
        rr = h_pin[2];
        if (!std::isfinite(rr)) { err_ = "non-finite residual in the synthetic solve"; return false; }
        // stagnation at round-off: stop when 8 checks in a row gain nothing
        if (rr >= prev * 0.999) { if (++stall >= 8) break; } else stall = 0;
        prev = std::min(prev, rr);
    }
    if (iters) *iters = it;
    return true;
}

int Solver::fourier(gbte_stats *st) {
    if (!ok()) return GBTE_HIP_ERROR;
    if (!check(hipSetDevice(opt.device), "hipSetDevice")) return GBTE_HIP_ERROR;  // the calling thread's device
    // StaticFourier::_solve (StaticFourier.cpp:934-976) from T = initial
    double err = 1000;
    int passes = 0;
    if (!check(hipMemcpy(d_T.p, P.initial_temp.data(), nc * sizeof(double), hipMemcpyHostToDevice), "H2D"))
        return GBTE_HIP_ERROR;
    while (err > 1e-5) {
        hipLaunchKernelGGL(k_copy, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_T1.p, nc);
        launch_vertex_T(d_T.p, d_TV.p);
        launch_rhs(false, d_fourier_bval.p, d_rhs.p);
        if (!linsolve(Kfour, NDfour, d_rhs.p, d_T.p, nullptr)) return GBTE_HIP_ERROR;
        hipLaunchKernelGGL(k_fourier_flux, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_gptr.p, d_gidx.p, d_gP.p,
                           d_gok.p, d_nint.p, gnnz, 1.0 / P.deck.Lx, P.kappa_bulk[0], d_qx.p, d_qy.p, d_qz.p,
                           nc, dim);
        if (!inner_error(d_T1.p, d_T.p, &err)) return GBTE_HIP_ERROR;
        if (opt.verbose) printf("iterations residual: %g\n", err);
        if (++passes > 100000) { err_ = "Fourier loop does not converge"; return GBTE_DIVERGED; }
    }
    if (!check(hipMemsetAsync(d_qz.p, 0, nc * sizeof(double), s_), "memset")) return GBTE_HIP_ERROR;
    fourier_passes_ = passes;
    if (st) st->fourier_passes = passes;
    return check(hipStreamSynchronize(s_), "sync") ? GBTE_OK : GBTE_HIP_ERROR;
}

int Solver::set_initial(const double *T, const double *qx, const double *qy, const double *qz) {
    if (!ok()) return GBTE_HIP_ERROR;
    if (!check(hipSetDevice(opt.device), "hipSetDevice")) return GBTE_HIP_ERROR;  // the calling thread's device
    const double *src[4] = {T, qx, qy, qz};
    double *dst[4] = {d_T.p, d_qx.p, d_qy.p, d_qz.p};
    std::vector<double> tmp(nc);
    for (int k = 0; k < 4; ++k) {
        if (!src[k]) continue;
        for (int c = 0; c < nc; ++c) tmp[cell_new_[c]] = src[k][c];  // deck order -> device order
        if (!check(hipMemcpy(dst[k], tmp.data(), nc * sizeof(double), hipMemcpyHostToDevice), "H2D"))
            return GBTE_HIP_ERROR;
    }
    return GBTE_OK;
}

// grid sizes of the kernels that remap blocks per XCD (xcd_block) must be
// multiples of kXcd
static int xgrid_of(int n) { return ((n + kXcd - 1) / kXcd) * kXcd; }

Modes Solver::modes() const {
    return Modes{d_vg.p, d_tau.p, d_cap.p, d_lr.p, d_hr.p, d_sx.p, d_sy.p, d_sz.p, d_w.p, d_lrc.p, P.kappa_bulk[0],
                 P.capacity_bulk[0], eref_, offT_, offE_, offW_, offQ_[0], offQ_[1], offQ_[2], offMac_};
}

// One whole-source sweep of every direction (tetrahedral meshes): the
// wave-specialised kernel or the level-synchronous one.
void Solver::launch_sweep3(const Modes &m) {
    if (lv_bc > 0) {
        launch_sweep_lv(m, lv_bc);
    } else if (use_pc) {
        PcSched ps{d_pc_rec.p, d_pc_size.p, d_pc_f0.p, d_pc_dsl.p, d_pc_far.p};
        launch_sweep_pc(ps, m, nullptr);
    } else {
        hipLaunchKernelGGL(k_sweep_dir, dim3(nD * sweep_nchunk), dim3(kSweepBlock),
                           (size_t)sweep_ring * sweep_bc * sizeof(double), s_, d_item_cell.p, d_levptr.p, d_dlev.p,
                           d_rec_q.p, d_rec_c.p, d_rec_g.p, d_rec_G.p, d_S.p, d_e.p, m, nc, nB, nD, sweep_bc,
                           sweep_nchunk, sweep_ring);
    }
}

// The wave-specialised sweep at the handle's storage type and far budget.
void Solver::launch_sweep_pc(const PcSched &ps, const Modes &m, const double *tpos) {
    dim3 g(nD * pc_nchunk), b(kPcBlock);
#define GBTE_PC(ST_, F_)                                                                                          \
    hipLaunchKernelGGL((k_sweep_pc<ST_, F_>), g, b, pc_lds, s_, ps, reinterpret_cast<const ST_ *>(d_S.p),            \
                       reinterpret_cast<ST_ *>(d_e.p), m, nc, nB, nD, pc_bc, pc_nchunk, pc_ring, tpos)
    if (fp32_ && pc_far == kPcFar3) GBTE_PC(float, kPcFar3);
    else if (fp32_) GBTE_PC(float, kPcFar);
    else if (pc_far == kPcFar3) GBTE_PC(double, kPcFar3);
    else GBTE_PC(double, kPcFar);
#undef GBTE_PC
}

// The per-system fields in their storage type, and a launch of a kernel
// templated on it (FP64, or FP32 with GBTE_FP32_SWEEP).
template <typename ST>
static ST *rp(DBuf<double> &b) {
    return reinterpret_cast<ST *>(b.p);
}
#define GBTE_LAUNCH_ST(kernel, grid, block, lds, ...)                             \
    do {                                                                         \
        if (fp32_) {                                                             \
            using ST = float;                                                    \
            hipLaunchKernelGGL(kernel<float>, grid, block, lds, s_, __VA_ARGS__); \
        } else {                                                                 \
            using ST = double;                                                   \
            hipLaunchKernelGGL(kernel<double>, grid, block, lds, s_, __VA_ARGS__); \
        }                                                                        \
    } while (0)

int Solver::solve(int max_iter, double tolT, double tolQ, gbte_stats *st) {
    if (!ok()) return GBTE_HIP_ERROR;
    if (!check(hipSetDevice(opt.device), "hipSetDevice")) return GBTE_HIP_ERROR;  // the calling thread's device
    if (dim == 3) return solve3d(max_iter, tolT, tolQ, st);
    // an earlier call that failed mid-iteration may have left side-stream work
    if (s2_ && !check(hipStreamSynchronize(s2_), "sync")) return GBTE_HIP_ERROR;
    auto t0 = std::chrono::high_resolution_clock::now();
    bool sequential = P.deck.method == "Sequential", first_order = P.deck.order == "1",
         limiter = P.deck.order == "2limit";
    Modes m = modes();
    SlotGeo g{d_fsize.p, d_snb.p, d_sent.p, d_snx.p, d_sny.p, d_snz.p, d_sarea.p, d_vol.p,
              d_srx.p, d_sry.p, d_srz.p, d_sqx.p, d_sqy.p, d_sqz.p};
    FaceK fk{d_fkind.p, d_fother.p, d_fv0.p, d_fv1.p, d_fga.p, d_fn0x.p, d_fn0y.p, d_fmx.p, d_fmy.p};
    int nnz = gnnz;
    double invLx = 1.0 / P.deck.Lx;
    int n_ent = (int)h_ent_list.size();
    size_t nE = (size_t)nD * nc * nB;
    Rows R{nD * nB, nB, row_tpr, row_rpb, row_tpr / nB, row_tpr % nB};
    Rows Rsrc{nD * nB, nB, 64, kSrcRows, 64 / nB, 64 % nB};  // k_source: one wave per row
    // FP32 storage: two bands per lane where the bands pair up and the rows
    // hold whole 16-byte quads
    bool src_pairs = fp32_ && nB % 2 == 0 && (nD * nB) % 4 == 0;
    Rows Rsrc2{nD * nB, nB, 128, kSrcRows, 128 / nB, 128 % nB};
    size_t src_lds = (size_t)(4 * nB + 2 * nD) * sizeof(double) + (size_t)kSrcRows * nD * sizeof(int);
    bool one_row = row_rpb == 1;
    int rowblk = row_tpr * row_rpb;
    // row kernels remap blocks per XCD (xcd_block): grids are multiples of 8
    auto xgrid = [](int n) { return ((n + kXcd - 1) / kXcd) * kXcd; };
    int gw = xgrid(blocks_for(nc, row_rpb));
    int bw = xgrid(blocks_for(n_ent, row_rpb));
    int vw = xgrid(blocks_for(nn, row_rpb));
    int bbw = blocks_for(blocks_for(n_ent, cpw), kWaves);
    // _set_initial, gradients, outgoing and incoming boundary values
    // (StaticBTESyntheticSolver.cpp:19-62)
    GBTE_LAUNCH_ST(k_init_e, dim3(blocks_for((long)nE, kBlock)), dim3(kBlock), 0, rp<ST>(d_e), d_T.p, d_cap.p,
                   nD * nB, nB, nE, m.eref);
    TileDev td{d_t_cell0.p, d_t_rptr.p, d_t_rsrc.p, d_t_nh.p, d_t_vptr.p, d_t_vs_ptr.p,
               d_t_vs_row.p, d_t_fin_i.p, d_t_vs_w.p, d_t_cmf.p, d_t_kwf.p, d_t_fin_d.p};
    // _get_gradient_larger of e: the tile kernel's LSQ product alone (the
    // tiles hold every cell's full stencil), or the row kernel
    auto launch_gradient = [&]() {
        if (use_tiles) {
            GBTE_LAUNCH_ST(k_tile_face, dim3(xgrid(n_tiles)), dim3(kTileBlock), tile_lds, rp<ST>(d_e), rp<ST>(d_eb),
                           rp<ST>(d_grad), td, d_T.p, d_qx.p, d_qy.p, d_qz.p, d_RM.p, make_double4(0, 0, 0, 0),
                           nD * nB, n_tiles, kTileGradient);
        } else {
            GBTE_LAUNCH_ST(k_gradient, dim3(gw), dim3(rowblk), 0, rp<ST>(d_e), rp<ST>(d_grad), d_gptr.p, d_gidx.p,
                           d_gP.p, d_gok.p, dim, nnz, invLx, nc, R);
        }
    };
    if (first_order) {
        // Order 1 (solve_firstorder, StaticBTESyntheticSolver.cpp:487-966):
        // the gradients are zero throughout (:537-542), so the outgoing
        // boundary values are the cell values (_get_bound_ee_firstorder,
        // StaticBTESynthetic.cpp:2883-2906) and the source has no gradient
        // terms (_get_Re_firstorder, :2841-2881)
        if (!check(hipMemsetAsync(d_grad.p, 0, d_grad.n * sizeof(double), s_), "zero gradients"))
            return GBTE_HIP_ERROR;
    } else {
        launch_gradient();
    }
    if (limiter)
        GBTE_LAUNCH_ST(k_limit, dim3(xgrid(blocks_for(nc, kSrcRows))), dim3(64 * kSrcRows), 0, rp<ST>(d_e),
                       rp<ST>(d_grad), rp<ST>(d_eb), d_fsize.p, d_srec.p, d_slim.p, m, nc, nD * nB, nB);
    GBTE_LAUNCH_ST(k_bound_out, dim3(std::max(bw, 1)), dim3(rowblk), 0, rp<ST>(d_e), rp<ST>(d_grad), rp<ST>(d_eb), g, m,
                       d_ent_list.p, d_be_cell.p, d_be_slot.p, n_ent, dim, R);
    GBTE_LAUNCH_ST(k_boundary, dim3(std::max(bbw, 1)), dim3(kBlock), 0, rp<ST>(d_eb), d_tmac.p, m, d_ent_list.p,
                       d_be_type.p, d_be_conn.p, d_be_spec.p, d_be_nx.p, d_be_ny.p, d_be_nz.p, d_be_tw.p, d_be_dT.p,
                       d_be_tov.p, n_ent, nB, nD, cpw);
    if (!check(hipGetLastError(), "setup kernels")) return GBTE_HIP_ERROR;
    errInc_ = 0;
    errTold_ = 0;
    errQold_ = 0;
    errInner_ = 0;
    int nt = 0, status = GBTE_OK, macro_total = 0;
    bool conv = false;
    double errT = 0, errQ = 0;
    double source_ms = 0, sweep_ms = 0, face_ms = 0, transport_ms = 0, macro_ms = 0;
    long cg_total = 0;
    // Sources in two parts (not with 2limit, whose limiter rewrites the
    // gradients after the boundary update, nor with the level-synchronous
    // sweep): the cells without a boundary face read only gradients, so
    // their sources (without the C T_old term) are formed on the side stream
    // right after the face kernel, beside the boundary update, the synthetic
    // passes and the host's error check; the cells with one follow there once
    // the boundary update has run, and the sweep adds C_b (T_old - Tref).
    // Not in batched solves: their many small instances already run side by
    // side, and the extra launches cost more than the overlap gains there.
    bool split = use_pc && !limiter && !single_stream_, overlap = split;
    if (overlap && !s2_) {
        // created at first use, so that handles that never use it do not
        // take a share of the hardware queues (default priority as well)
        if (!check(hipStreamCreateWithFlags(&s2_, hipStreamNonBlocking), "hipStreamCreateWithFlags"))
            return GBTE_HIP_ERROR;
        for (auto &set : ev2_)
            for (hipEvent_t &ev : set)
                if (!check(hipEventCreate(&ev), "hipEventCreate")) return GBTE_HIP_ERROR;
    }
    auto launch_src = [&](hipStream_t st, const int *cells, int nrows, int with_told) {
        if (nrows <= 0) return;
        dim3 grid(xgrid(blocks_for(nrows, kSrcRows))), blk(64 * kSrcRows);
        if (src_pairs)
            hipLaunchKernelGGL(k_source_f32x2, grid, blk, src_lds, st, d_posc.p, rp<float>(d_S), rp<float>(d_grad),
                               rp<float>(d_eb), d_Told.p, d_Q.p, d_fsize.p, d_srec.p, d_srei.p, m, nc, Rsrc2, cells,
                               nrows, with_told);
        else if (fp32_)
            hipLaunchKernelGGL(k_source<float>, grid, blk, src_lds, st, d_posc.p, rp<float>(d_S), rp<float>(d_grad),
                               rp<float>(d_eb), d_Told.p, d_Q.p, d_fsize.p, d_srec.p, d_srei.p, m, nc, Rsrc, cells, nrows,
                               with_told);
        else
            hipLaunchKernelGGL(k_source<double>, grid, blk, src_lds, st, d_posc.p, d_S.p, d_grad.p, d_eb.p, d_Told.p,
                               d_Q.p, d_fsize.p, d_srec.p, d_srei.p, m, nc, Rsrc, cells, nrows, with_told);
    };
    bool side_pending = false;  // this iteration's interior sources are on the side stream
    for (nt = 0; nt < max_iter; ++nt) {
        bool side_src = side_pending;
        side_pending = false;
        hipLaunchKernelGGL(k_copy_old, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_qx.p, d_qy.p, d_qz.p, d_Told.p,
                           d_qxo.p, d_qyo.p, d_qzo.p, nc);
        // this iteration's gradients were formed by the previous iteration's
        // face kernel (the first iteration's by a fresh pass: the set-up ones
        // were limited in place); the limiter needs the boundary values
        // written since
        if (limiter && nt == 0) launch_gradient();
        if (limiter)
            GBTE_LAUNCH_ST(k_limit, dim3(xgrid(blocks_for(nc, kSrcRows))), dim3(64 * kSrcRows), 0, rp<ST>(d_e),
                           rp<ST>(d_grad), rp<ST>(d_eb), d_fsize.p, d_srec.p, d_slim.p, m, nc, nD * nB, nB);
        (void)hipEventRecord(ev_[0], s_);
        if (!split) {
            launch_src(s_, nullptr, nc, 1);
        } else {
            if (!side_src) {
                launch_src(s_, d_src_int.p, n_src_int, 0);
                launch_src(s_, d_src_bnd.p, n_src_bnd, 0);
            }
            size_t npos = (size_t)nD * nc;
            hipLaunchKernelGGL(k_tpos, dim3((unsigned)((npos + kBlock - 1) / kBlock)), dim3(kBlock), 0, s_,
                               d_item_cell.p, d_Told.p, d_tpos.p, m.eref, npos);
            if (side_src && !check(hipStreamWaitEvent(s_, ev2_[nt & 1][2], 0), "hipStreamWaitEvent"))
                return GBTE_HIP_ERROR;
        }
        (void)hipEventRecord(ev_[1], s_);
        if (use_pc) {
            PcSched ps{d_pc_rec.p, d_pc_size.p, d_pc_f0.p, d_pc_dsl.p, d_pc_far.p};
            launch_sweep_pc(ps, m, split ? d_tpos.p : nullptr);
        } else {
            hipLaunchKernelGGL(k_sweep_dir, dim3(nD * sweep_nchunk), dim3(kSweepBlock),
                               (size_t)sweep_ring * sweep_bc * sizeof(double), s_, d_item_cell.p, d_levptr.p, d_dlev.p,
                               d_rec_q.p, d_rec_c.p, d_rec_g.p, d_rec_G.p, d_S.p, d_e.p, m, nc, nB, nD, sweep_bc,
                               sweep_nchunk, sweep_ring);
        }
        (void)hipEventRecord(ev_[2], s_);
        GBTE_LAUNCH_ST(k_bound_out, dim3(std::max(bw, 1)), dim3(rowblk), 0, rp<ST>(d_e), rp<ST>(d_grad), rp<ST>(d_eb), g,
                       m,
                           d_ent_list.p, d_be_cell.p, d_be_slot.p, n_ent, dim, R);
        (void)hipEventRecord(ev_[3], s_);
        if (use_tiles) {
            GBTE_LAUNCH_ST(k_tile_face, dim3(xgrid(n_tiles)), dim3(kTileBlock), tile_lds, rp<ST>(d_e), rp<ST>(d_eb),
                           rp<ST>(d_grad), td, d_T.p, d_qx.p, d_qy.p, d_qz.p, d_RM.p,
                               make_double4(m.offT, m.offQx, m.offQy, m.offQz), nD * nB, n_tiles,
                               first_order ? kTileMoments : kTileMoments | kTileGradient);
        } else {
            hipLaunchKernelGGL(one_row ? k_vertex_e<true> : k_vertex_e<false>, dim3(vw), dim3(rowblk), 0, s_, d_e.p, d_eb.p,
                               d_eV.p, d_vptr.p, d_vsrc.p, d_vwt.p, nn, R);
            hipLaunchKernelGGL(one_row ? k_face_moments<true> : k_face_moments<false>, dim3(gw), dim3(rowblk), 0, s_, d_e.p,
                               d_eb.p, d_eV.p, d_grad.p, g, fk, m, d_gptr.p, d_gidx.p, d_gP.p, d_gok.p, nnz, invLx, d_T.p,
                               d_qx.p, d_qy.p, d_qz.p, d_RM.p, nc, R, first_order ? 0 : 1);
        }
        (void)hipEventRecord(ev_[4], s_);
        if (overlap && nt + 1 < max_iter) {
            hipEvent_t *e2 = ev2_[(nt + 1) & 1];
            (void)hipEventRecord(e2[0], s_);
            if (!check(hipStreamWaitEvent(s2_, e2[0], 0), "hipStreamWaitEvent")) return GBTE_HIP_ERROR;
            (void)hipEventRecord(e2[1], s2_);
            launch_src(s2_, d_src_int.p, n_src_int, 0);
            (void)hipEventRecord(e2[4], s2_);
            side_pending = true;
        }
        GBTE_LAUNCH_ST(k_boundary, dim3(std::max(bbw, 1)), dim3(kBlock), 0, rp<ST>(d_eb), d_tmac.p, m,
                           d_ent_list.p, d_be_type.p, d_be_conn.p, d_be_spec.p, d_be_nx.p, d_be_ny.p, d_be_nz.p,
                           d_be_tw.p, d_be_dT.p, d_be_tov.p, n_ent, nB, nD, cpw);
        if (side_pending) {
            // the cells with a boundary face read the values just updated
            hipEvent_t *e2 = ev2_[(nt + 1) & 1];
            (void)hipEventRecord(e2[3], s_);
            if (!check(hipStreamWaitEvent(s2_, e2[3], 0), "hipStreamWaitEvent")) return GBTE_HIP_ERROR;
            (void)hipEventRecord(e2[5], s2_);
            launch_src(s2_, d_src_bnd.p, n_src_bnd, 0);
            (void)hipEventRecord(e2[2], s2_);
        }
        (void)hipEventRecord(ev_[5], s_);
        if (!check(hipGetLastError(), "iteration kernels")) return GBTE_HIP_ERROR;
        // _marco_solution (StaticBTESynthetic.cpp:2692-2734)
        double inner = 1000;
        int inc = 0, passes = 0;
        bool have_err = false;  // h_err already holds this iteration's outer error partials
        auto account = [&](double in) {  // the loop's bookkeeping after one pass; false: stop
            inner = in;
            if (inner > errInner_) inc++;
            errInner_ = inner;
            passes++;
            if (opt.verbose) printf("internal residual: %g\n", inner);
            if (inc >= 10) { err_ = "cannot congerve, try to set Order=1 or 2limit"; status = GBTE_DIVERGED; return false; }
            return inner > 1e-6;
        };
        if (sequential) {
            // solve_DOM (StaticBTESyntheticSolver.cpp:968-1178): no synthetic
            // correction, T stays the recovered lattice temperature
            (void)hipEventRecord(ev_[6], s_);
        } else if (opt.linear_solver == GBTE_LINSOLVE_DIRECT) {
            // two passes and the outer error of their result are enqueued
            // behind one sync (most iterations take exactly two passes); a
            // loop that stops after the first pass undoes the second (T = T1)
            if (!macro_launch(0) || !macro_launch(1)) return GBTE_HIP_ERROR;
            (void)hipEventRecord(ev_[6], s_);
            hipLaunchKernelGGL(k_err, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_Told.p, d_qx.p, d_qy.p, d_qz.p,
                               d_qxo.p, d_qyo.p, d_qzo.p, d_part.p, nc, red_blocks);
            if (!check(hipMemcpyAsync(h_err, d_part.p, (size_t)4 * red_blocks * sizeof(double), hipMemcpyDeviceToHost,
                                      s_),
                       "D2H") ||
                !check(hipStreamSynchronize(s_), "sync"))
                return GBTE_HIP_ERROR;
            if (!account(macro_inner(0))) {
                hipLaunchKernelGGL(k_copy, dim3(red_blocks), dim3(kBlock), 0, s_, d_T1.p, d_T.p, nc);
            } else if (!account(macro_inner(1))) {
                have_err = status == GBTE_OK;
            } else {
                for (bool more = true; more;) {
                    if (!macro_launch(passes & 1)) return GBTE_HIP_ERROR;
                    (void)hipEventRecord(ev_[6], s_);
                    if (!check(hipStreamSynchronize(s_), "sync")) return GBTE_HIP_ERROR;
                    more = account(macro_inner(passes & 1));
                }
            }
        } else {
            while (inner > 1e-6) {
                hipLaunchKernelGGL(k_copy, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_T1.p, nc);
                launch_vertex_T(d_T.p, d_TV.p);
                launch_rhs(true, d_tmac.p, d_rhs.p);
                int cgit = 0;
                if (!linsolve(Kmacro, NDmacro, d_rhs.p, d_T.p, &cgit)) return GBTE_HIP_ERROR;
                cg_total += cgit;
                double in = 0;
                if (!inner_error(d_T1.p, d_T.p, &in)) return GBTE_HIP_ERROR;
                if (!account(in)) break;
            }
            (void)hipEventRecord(ev_[6], s_);
        }
        macro_total += passes;
        last_passes_ = passes;
        auto add_times = [&]() {  // after a host sync on s_
            float ms = 0;
            if (side_src) {
                // the two side-stream source launches, each between its own
                // events (completed: this stream waited for them and has been
                // synced); the main stream's interval holds k_tpos and the wait
                (void)hipEventElapsedTime(&ms, ev2_[nt & 1][1], ev2_[nt & 1][4]);
                source_ms += ms;
                (void)hipEventElapsedTime(&ms, ev2_[nt & 1][5], ev2_[nt & 1][2]);
                source_ms += ms;
            } else {
                (void)hipEventElapsedTime(&ms, ev_[0], ev_[1]);
                source_ms += ms;
            }
            (void)hipEventElapsedTime(&ms, ev_[1], ev_[2]);
            sweep_ms += ms;
            (void)hipEventElapsedTime(&ms, ev_[3], ev_[4]);
            face_ms += ms;
            (void)hipEventElapsedTime(&ms, ev_[0], ev_[4]);
            transport_ms += ms;
            (void)hipEventElapsedTime(&ms, ev_[5], ev_[6]);
            macro_ms += ms;
        };
        if (status != GBTE_OK) {
            if (!check(hipEventSynchronize(ev_[6]), "event sync")) return GBTE_HIP_ERROR;
            add_times();
            ++nt;
            break;
        }
        // _get_magin_check_error
        std::vector<double> h;
        if (have_err) {
            h.assign(h_err, h_err + (size_t)4 * red_blocks);
        } else {
            hipLaunchKernelGGL(k_err, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_Told.p, d_qx.p, d_qy.p, d_qz.p,
                               d_qxo.p, d_qyo.p, d_qzo.p, d_part.p, nc, red_blocks);
            if (!reduce_host(d_part.p, red_blocks, h, 4)) return GBTE_HIP_ERROR;
        }
        add_times();
        double sT = 0, mT = 0, sQ = 0, mQ = 0;
        for (int k = 0; k < red_blocks; ++k) {
            sT += h[k];
            mT = std::max(mT, h[red_blocks + k]);
            sQ += h[2 * red_blocks + k];
            mQ = std::max(mQ, h[3 * red_blocks + k]);
        }
        errT = sT / nc / mT;
        errQ = sQ / nc / mQ;
        if (opt.verbose)
            printf("Iteration #%d\t Margin per band per cell: Temperature:%g Heat Flux:%g\n", nt, errT, errQ);
        if ((errT > errTold_ && errQ > errQold_) || std::isnan(errT) || std::isnan(errQ) || std::isinf(errT) || std::isinf(errQ))
            errInc_++;
        else
            errInc_ = 0;
        errTold_ = errT;
        errQold_ = errQ;
        if (std::isnan(errT) || std::isnan(errQ)) { status = GBTE_NAN; err_ = "NaN in the iteration errors"; ++nt; break; }
        if (errT <= tolT && errQ <= tolQ) { conv = true; ++nt; break; }
        if (errInc_ >= 10) { status = GBTE_DIVERGED; err_ = "Error increases for 10 times"; ++nt; break; }
    }
    GBTE_LAUNCH_ST(k_output_moments, dim3(xgrid(blocks_for(nc, kOutRows))), dim3(64 * kOutRows), 0, rp<ST>(d_e), m,
                   d_E.p, d_Tb.p, nc, nB, nD);
    // the side stream may hold the sources of an iteration that did not run
    if ((s2_ && !check(hipStreamSynchronize(s2_), "sync")) || !check(hipStreamSynchronize(s_), "sync"))
        return GBTE_HIP_ERROR;
    auto t1 = std::chrono::high_resolution_clock::now();
    have_state = true;
    if (st) {
        st->iterations = nt;
        st->macro_passes = macro_total;
        st->fourier_passes = fourier_passes_;
        st->converged = conv ? 1 : 0;
        st->err_temp = errT;
        st->err_flux = errQ;
        st->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->sweep_ms = sweep_ms;
        st->transport_ms = transport_ms;
        st->macro_ms = macro_ms;
        st->sweep_levels = use_pc ? nsub : nlev;
        st->macro_cg_iters = (int)cg_total;
        st->source_ms = source_ms;
        st->face_ms = face_ms;
        st->grad_ms = 0;  // formed inside the face kernel
    }
    return status;
}

// StaticBTESynthetic::solve on tetrahedra (StaticBTESyntheticSolver.cpp:12-247,
// solve_DOM :968-1178 for the Sequential scheme, the first-order variants
// :487-966, :1394-1597 with zero gradients): the 3D kernels of solver3d.hip,
// the sweep, boundary update, synthetic passes and error check of the 2D loop.
int Solver::solve3d(int max_iter, double tolT, double tolQ, gbte_stats *st) {
    auto t0 = std::chrono::high_resolution_clock::now();
    bool sequential = P.deck.method == "Sequential", first_order = P.deck.order == "1";
    Modes m = modes();
    size_t nE = (size_t)nD * nc * nB;
    int n_ent = (int)h_ent_list.size();
    int bbw = blocks_for(blocks_for(n_ent, cpw), kWaves);
    auto launch_boundary = [&]() {
        GBTE_LAUNCH_ST(k_boundary, dim3(std::max(bbw, 1)), dim3(kBlock), 0, rp<ST>(d_eb), d_tmac.p, m, d_ent_list.p,
                       d_be_type.p, d_be_conn.p, d_be_spec.p, d_be_nx.p, d_be_ny.p, d_be_nz.p, d_be_tw.p, d_be_dT.p,
                       d_be_tov.p, n_ent, nB, nD, cpw);
    };
    auto launch_gradient = [&]() {
        if (first_order) {
            // Order 1: the gradients stay zero (StaticBTESyntheticSolver.cpp:537-542)
            (void)check(hipMemsetAsync(d_grad.p, 0, d_grad.n * sizeof(double), s_), "zero gradients");
        } else {
            launch_grad3();
        }
    };
    // _set_initial, gradients, outgoing and incoming boundary values
    // (StaticBTESyntheticSolver.cpp:19-62)
    GBTE_LAUNCH_ST(k_init_e, dim3(blocks_for((long)nE, kBlock)), dim3(kBlock), 0, rp<ST>(d_e), d_T.p, d_cap.p,
                   nD * nB, nB, nE, m.eref);
    launch_gradient();
    launch_bound_out3(m);
    launch_boundary();
    if (!check(hipGetLastError(), "setup kernels")) return GBTE_HIP_ERROR;
    errInc_ = 0;
    errTold_ = 0;
    errQold_ = 0;
    errInner_ = 0;
    int nt = 0, status = GBTE_OK, macro_total = 0;
    bool conv = false;
    double errT = 0, errQ = 0;
    double source_ms = 0, sweep_ms = 0, face_ms = 0, grad_ms = 0, transport_ms = 0, macro_ms = 0;
    long cg_total = 0;
    for (nt = 0; nt < max_iter; ++nt) {
        hipLaunchKernelGGL(k_copy_old, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_qx.p, d_qy.p, d_qz.p, d_Told.p,
                           d_qxo.p, d_qyo.p, d_qzo.p, nc);
        (void)hipEventRecord(ev_[0], s_);
        launch_source3(m);
        (void)hipEventRecord(ev_[1], s_);
        launch_sweep3(m);
        (void)hipEventRecord(ev_[2], s_);
        launch_bound_out3(m);
        (void)hipEventRecord(ev_[3], s_);
        if (!first_order) launch_grad3();
        (void)hipEventRecord(ev_[7], s_);
        launch_moments3(m);
        (void)hipEventRecord(ev_[4], s_);
        launch_boundary();
        (void)hipEventRecord(ev_[5], s_);
        if (!check(hipGetLastError(), "iteration kernels")) return GBTE_HIP_ERROR;
        // _marco_solution (StaticBTESynthetic.cpp:2692-2734)
        double inner = 1000;
        int inc = 0, passes = 0;
        auto account = [&](double in) {
            inner = in;
            if (inner > errInner_) inc++;
            errInner_ = inner;
            passes++;
            if (opt.verbose) printf("internal residual: %g\n", inner);
            if (inc >= 10) { err_ = "cannot congerve, try to set Order=1 or 2limit"; status = GBTE_DIVERGED; return false; }
            return inner > 1e-6;
        };
        if (sequential) {
            (void)hipEventRecord(ev_[6], s_);
        } else if (opt.linear_solver == GBTE_LINSOLVE_DIRECT) {
            for (bool more = true; more;) {
                if (!macro_launch(passes & 1)) return GBTE_HIP_ERROR;
                (void)hipEventRecord(ev_[6], s_);
                if (!check(hipStreamSynchronize(s_), "sync")) return GBTE_HIP_ERROR;
                more = account(macro_inner((passes) & 1));
            }
        } else {
            while (inner > 1e-6) {
                hipLaunchKernelGGL(k_copy, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_T1.p, nc);
                launch_vertex_T(d_T.p, d_TV.p);
                launch_rhs(true, d_tmac.p, d_rhs.p);
                int cgit = 0;
                if (!linsolve(Kmacro, NDmacro, d_rhs.p, d_T.p, &cgit)) return GBTE_HIP_ERROR;
                cg_total += cgit;
                double in = 0;
                if (!inner_error(d_T1.p, d_T.p, &in)) return GBTE_HIP_ERROR;
                if (!account(in)) break;
            }
            (void)hipEventRecord(ev_[6], s_);
        }
        macro_total += passes;
        last_passes_ = passes;
        auto add_times = [&]() {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev_[0], ev_[1]);
            source_ms += ms;
            (void)hipEventElapsedTime(&ms, ev_[1], ev_[2]);
            sweep_ms += ms;
            (void)hipEventElapsedTime(&ms, ev_[3], ev_[4]);
            face_ms += ms;
            (void)hipEventElapsedTime(&ms, ev_[3], ev_[7]);
            grad_ms += ms;
            (void)hipEventElapsedTime(&ms, ev_[0], ev_[4]);
            transport_ms += ms;
            (void)hipEventElapsedTime(&ms, ev_[5], ev_[6]);
            macro_ms += ms;
        };
        if (status != GBTE_OK) {
            if (!check(hipEventSynchronize(ev_[6]), "event sync")) return GBTE_HIP_ERROR;
            add_times();
            ++nt;
            break;
        }
        // _get_magin_check_error
        std::vector<double> h;
        hipLaunchKernelGGL(k_err, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_Told.p, d_qx.p, d_qy.p, d_qz.p,
                           d_qxo.p, d_qyo.p, d_qzo.p, d_part.p, nc, red_blocks);
        if (!reduce_host(d_part.p, red_blocks, h, 4)) return GBTE_HIP_ERROR;
        add_times();
        double sT = 0, mT = 0, sQ = 0, mQ = 0;
        for (int k = 0; k < red_blocks; ++k) {
            sT += h[k];
            mT = std::max(mT, h[red_blocks + k]);
            sQ += h[2 * red_blocks + k];
            mQ = std::max(mQ, h[3 * red_blocks + k]);
        }
        errT = sT / nc / mT;
        errQ = sQ / nc / mQ;
        if (opt.verbose)
            printf("Iteration #%d\t Margin per band per cell: Temperature:%g Heat Flux:%g\n", nt, errT, errQ);
        if ((errT > errTold_ && errQ > errQold_) || std::isnan(errT) || std::isnan(errQ) || std::isinf(errT) ||
            std::isinf(errQ))
            errInc_++;
        else
            errInc_ = 0;
        errTold_ = errT;
        errQold_ = errQ;
        if (std::isnan(errT) || std::isnan(errQ)) { status = GBTE_NAN; err_ = "NaN in the iteration errors"; ++nt; break; }
        if (errT <= tolT && errQ <= tolQ) { conv = true; ++nt; break; }
        if (errInc_ >= 10) { status = GBTE_DIVERGED; err_ = "Error increases for 10 times"; ++nt; break; }
    }
    GBTE_LAUNCH_ST(k_output_moments, dim3(xgrid_of(blocks_for(nc, kOutRows))), dim3(64 * kOutRows), 0, rp<ST>(d_e), m,
                   d_E.p, d_Tb.p, nc, nB, nD);
    if (!check(hipStreamSynchronize(s_), "sync")) return GBTE_HIP_ERROR;
    auto t1 = std::chrono::high_resolution_clock::now();
    have_state = true;
    if (st) {
        st->iterations = nt;
        st->macro_passes = macro_total;
        st->fourier_passes = fourier_passes_;
        st->converged = conv ? 1 : 0;
        st->err_temp = errT;
        st->err_flux = errQ;
        st->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->sweep_ms = sweep_ms;
        st->transport_ms = transport_ms;
        st->macro_ms = macro_ms;
        st->sweep_levels = use_pc ? nsub : nlev;
        st->macro_cg_iters = (int)cg_total;
        st->source_ms = source_ms;
        st->face_ms = face_ms;
        st->grad_ms = grad_ms;
    }
    return status;
}

int Solver::get_fields(double *TL, double *TE, double *qx, double *qy, double *qz, double *Tb) {
    if (!ok()) return GBTE_HIP_ERROR;
    if (!check(hipSetDevice(opt.device), "hipSetDevice")) return GBTE_HIP_ERROR;  // the calling thread's device
    struct {
        double *h;
        const double *d;
        int rows;
    } f[6] = {{TL, d_T.p, 1}, {TE, d_E.p, 1}, {qx, d_qx.p, 1}, {qy, d_qy.p, 1}, {qz, d_qz.p, 1}, {Tb, d_Tb.p, nB}};
    std::vector<double> tmp;
    for (auto &x : f) {
        if (!x.h) continue;
        tmp.resize((size_t)x.rows * nc);
        if (!check(hipMemcpy(tmp.data(), x.d, tmp.size() * sizeof(double), hipMemcpyDeviceToHost), "D2H"))
            return GBTE_HIP_ERROR;
        // device order -> deck order
        for (int r = 0; r < x.rows; ++r)
            for (int c = 0; c < nc; ++c) x.h[(size_t)r * nc + c] = tmp[(size_t)r * nc + cell_new_[c]];
    }
    return GBTE_OK;
}

int Solver::keff(int axis, double dT, double *out) {
    if (!ok()) return GBTE_HIP_ERROR;
    if (!check(hipSetDevice(opt.device), "hipSetDevice")) return GBTE_HIP_ERROR;  // the calling thread's device
    if (axis < 0 || axis >= dim || dT == 0) { err_ = "keff: bad axis or zero temperature difference"; return GBTE_BAD_INPUT; }
    std::vector<double> q(nc);
    const double *src = axis == 0 ? d_qx.p : (axis == 1 ? d_qy.p : d_qz.p);
    if (!check(hipMemcpy(q.data(), src, nc * sizeof(double), hipMemcpyDeviceToHost), "D2H")) return GBTE_HIP_ERROR;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v = 0; v < nn; ++v) {
        double x[3] = {P.node_x[v], P.node_y[v], P.node_z[v]};
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], x[k]); hi[k] = std::max(hi[k], x[k]); }
    }
    double ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    double s = 0;
    for (int c = 0; c < nc; ++c) s += q[c] * P.volume[c];
    // sum q V / (A_perp dT) with A_perp the cross-section normal to the axis
    double area = 1;
    for (int k = 0; k < dim; ++k)
        if (k != axis) area *= ext[k];
    *out = s / (area * dT);
    return GBTE_OK;
}

}  // namespace gbte
