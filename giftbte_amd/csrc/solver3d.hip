// MI355X (gfx950) kernels and host set-up for tetrahedral (3D) meshes.
//
// One outer iteration of StaticBTESynthetic::solve
// (src/StaticBTESynthetic/StaticBTESyntheticSolver.cpp:63-221) on tetrahedra:
//   k_source3     the part of _get_Re that does not depend on the new upwind
//                 values, for all systems              StaticBTESynthetic.cpp:1427-1500
//   k_sweep_lv    the per-system SparseLU solve as a level-synchronous
//                 upwind sweep (or k_sweep_pc / k_sweep_dir of solver.hip;
//                 the set-up tuner picks the fastest)  StaticBTESyntheticSolver.cpp:96-100
//   k_bound_out3  _get_bound_ee                                      :1502-1539
//   k_vsum3 +     the next step's _get_gradient_larger (3D branch)   :1329-1371
//   k_grad3v      through sums over the mesh vertices (or k_grad3g, the
//                 direct stencil gather, or k_grad3m on the matrix cores;
//                 the set-up tuner picks the fastest)
//   k_mom3        per-row moments of e and of the boundary values: T, q and
//                 the seven direction factors of the non-Fourier face terms
//   k_vert3       _set_vertex_energydensity on those moments         :1541-1616
//   k_fin3        _recover_temperature, _get_heat_flux and
//                 _get_non_fourier with _get_face_gradient's 3D branch
//                 (FaceMatrix, _get_face_matrix :1617-1649)        :1651-1937, :2485-2520
//   k_boundary    _set_bound_ee_1 + _set_macro_boundary_condition (solver.hip)
//   synthetic     _marco_solution with the 3D _get_Re_macro (k_rhs3) :2307-2407, :2692-2734
//
// Layout: e [cell][dir][band] as in 2D.  The LSQ gradient is used only
// through its projections on the cell's four face offsets (r_f . grad e, in
// the sources and the outgoing boundary values), so it is stored as those:
// phi [cell][face][dir][band], each face's row one coalesced load.
//
// The LSQ gradient of a tetrahedron reads its whole vertex-neighbour stencil
// (about 80 cells on a Delaunay mesh), so every row is re-read by about 80
// gradients.  k_grad3g gathers them from L2, where the waves of
// Morton-consecutive cells find largely the same rows; k_vsum3 + k_grad3v
// first sum each vertex's cells once and then read four vertex sums and
// about 20 rows per cell.  (Staging the rows of a tile of cells in LDS was
// measured slower: the tile's stencil union is about 190 rows for 8 cells,
// and the per-entry weights then have to reach every lane from LDS or scalar
// loads.)
//
// The non-Fourier residual is linear in the face values, so the per-system
// sums over faces factor through per-row moments Y_j = sum_sys K_j(sys) e:
//   Y_0            = sum [hr>0] w kappa lr/C e
//   Y_1..6 (ab)    = sum [hr>0] w tau v^2 s_a s_b e   (xx, xy, xz, yy, yz, zz)
//   Y_7            = sum w lr/C e                     (lattice temperature)
//   Y_8..10 (a)    = sum w v s_a e                    (heat flux)
// with the vertex moments the inverse-distance combinations of the row
// moments.  Each face then contributes fixed geometric multiples of
// differences of Y (host-computed, 21 per face slot), which is the
// reference's sum over systems reordered.
#include "devcommon.hpp"
#include "solver.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <thread>
#include <type_traits>

namespace gbte {

namespace {

constexpr int kY = 12;            // row moments (11 used)
constexpr int kYV = 8;            // vertex moments (the seven residual factors)
constexpr int kS3D = 40;          // source record doubles per cell: 4 faces x (n, A/V, r, r')
constexpr int kS3I = 4;           // source record ints per cell
constexpr int kSrc3Rows = 4;      // k_source3: cell rows (waves) per workgroup
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kG3W = 64;          // entries per slice
constexpr int kG3Ent = 8;         // stencil entries per batch
constexpr int kG3Pad = 4;         // stencils padded to a multiple (a batch of 8, then one of 4)
constexpr int kFin = 21;          // k_fin3 coefficients per face slot
constexpr int kMomRows = 4;       // k_mom3: rows (waves) per workgroup
// S layouts between the source and the chunked sweep (s_sh): >= 0
// [dir][band chunk][position][2^s_sh], -1 [dir][position][band], kSCell
// [cell][dir][band] (e's layout: the source's stores are whole rows, the
// sweep reads S at the position's cell like its upwind values)
constexpr int kSCell = -2;

// One LSQ stencil entry of k_grad3g: the neighbour row and the weights / L_x.
template <typename AT>
struct G3Ent;
template <>
struct alignas(32) G3Ent<double> {
    int loc, pad;
    double px, py, pz;
};
template <>
struct alignas(16) G3Ent<float> {
    int loc;
    float px, py, pz;
};

// ----------------------------------------------------------------- kernels

// Row loads of the gradient kernels: a buffer resource per row (built from
// the wave-uniform row address by scalar instructions) and a per-lane 32-bit
// byte offset, so that a load takes no per-lane 64-bit address arithmetic.
// Word 3 is the gfx9 default (32-bit data format, no swizzle).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void *row, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(row), (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T ld_row(__amdgpu_buffer_rsrc_t r, unsigned off);
template <>
__device__ __forceinline__ float ld_row<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ double ld_row<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ v4f ld_row<v4f>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(v4f, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}
template <>
__device__ __forceinline__ v4d ld_row<v4d>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    v2d lo = __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
    v2d hi = __builtin_bit_cast(v2d, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off + 16, 0, 0));
    return v4d{lo[0], lo[1], hi[0], hi[1]};
}

// One batch of N stencil-table entries: the table entries by scalar loads
// (they serve all GS slices of the wave), then N * GS row loads before the
// products.
template <typename ST, int GS, bool S0, int N>
__device__ __forceinline__ void g3_batch(const ST *__restrict__ X, const int *__restrict__ gidx,
                                         const G3Ent<ST> *__restrict__ ent, int k, const unsigned *to, size_t DB,
                                         const ST *self, ST *gx, ST *gy, ST *gz, ST *s0) {
    G3Ent<ST> w[N];
    __amdgpu_buffer_rsrc_t row[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        w[j] = ent[k + j];
        row[j] = row_rsrc(X + (size_t)gidx[k + j] * DB, (unsigned)(DB * sizeof(ST)));
    }
    ST v[N][GS];
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int q = 0; q < GS; ++q) v[j][q] = ld_row<ST>(row[j], to[q]);
#pragma unroll
    for (int j = 0; j < N; ++j)
#pragma unroll
        for (int q = 0; q < GS; ++q) {
            ST d = v[j][q] - self[q];
            gx[q] += w[j].px * d;
            gy[q] += w[j].py * d;
            gz[q] += w[j].pz * d;
            if constexpr (S0) s0[q] += d;
        }
}

// The entry loop of the gathering LSQ kernels: g += w_j (X[row j] - self)
// over the stencil-table entries a..z (padded to a multiple of kG3Pad), with
// S0 also the plain sum of the differences; to[q] is slice q's byte offset
// in a row.
template <typename ST, int GS, bool S0>
__device__ __forceinline__ void g3_entries(const ST *__restrict__ X, const int *__restrict__ gidx,
                                           const G3Ent<ST> *__restrict__ ent, int a, int z, const unsigned *to, int DB,
                                           const ST *self, ST *gx, ST *gy, ST *gz, ST *s0) {
    int k = a;
    for (; k + kG3Ent <= z; k += kG3Ent) g3_batch<ST, GS, S0, kG3Ent>(X, gidx, ent, k, to, DB, self, gx, gy, gz, s0);
    if (k < z) g3_batch<ST, GS, S0, kG3Pad>(X, gidx, ent, k, to, DB, self, gx, gy, gz, s0);
}

// A face projection phi_f is read only where the face is an out-face of
// the cell for the entry's direction (s . n_f >= 0: by the cell's own source
// and outgoing boundary value, and by the downwind neighbour's source), so
// it is written only there: about half of the four rows.  The write is
// skipped only where s . n_f < -kInFace, where every reader's rounded
// v s . n is negative too; the rows are zeroed once at set-up.
constexpr double kInFace = 1e-12;

// The direction of entry t = d nB + b: floor((t + 0.5) / nB) in FP32 with
// rnB = 1 / nB, exact while t < 2^21 (the relative error, below 2^-22, then
// moves the quotient by less than its distance 0.5 / nB to an integer); the
// 3D set-up rejects larger dir x band counts.
__device__ __forceinline__ int dir_of(int t, float rnB) { return (int)(((float)t + 0.5f) * rnB); }

// Out-face flags per (cell, direction), bit f where s . n_f >= -kInFace:
// the faces whose projections the gradient kernels write (set-up; one byte
// load per lane then replaces three direction loads and the dot products).
__global__ void k_oflags3(const double *__restrict__ rec, Modes m, int nc, int nD, unsigned char *__restrict__ fl) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nc * nD) return;
    int c = (int)(i / nD), d = (int)(i - (size_t)c * nD);
    const double *rc = rec + (size_t)c * kS3D;
    double sx = m.sx[d], sy = m.sy[d], sz = m.sz[d];
    unsigned v = 0;
#pragma unroll
    for (int f = 0; f < 4; ++f)
        if (sx * rc[f * 10] + sy * rc[f * 10 + 1] + sz * rc[f * 10 + 2] >= -kInFace) v |= 1u << f;
    fl[i] = (unsigned char)v;
}

// phi_f = r_f . g for the out-faces of cell c in the entries tb + q kG3W
// (q < GS, those below DB), NaN -> 0 (the reference's zero gradient), the
// out-faces from the set-up flags.  FP64 storage takes the projections in
// FP64, FP32 storage in FP32 with the face offsets converted once per wave.
template <typename ST, int GS>
__device__ __forceinline__ void g3_store_phi(ST *__restrict__ phi, const double *__restrict__ rc,
                                             const unsigned char *__restrict__ oflag, int c, int tb, int DB, int nB,
                                             ST *gx, ST *gy, ST *gz) {
    float rnB = 1.0f / (float)nB;
    int nD = DB / nB;
    __amdgpu_buffer_rsrc_t fr_c = row_rsrc(oflag + (size_t)c * nD, (unsigned)nD);
    ST *pc = phi + (size_t)c * 4 * DB;
    using AT = typename std::conditional<sizeof(ST) == 4, float, double>::type;
    AT fr[4][3];
#pragma unroll
    for (int f = 0; f < 4; ++f)
#pragma unroll
        for (int a = 0; a < 3; ++a) fr[f][a] = (AT)rc[f * 10 + 4 + a];
#pragma unroll
    for (int q = 0; q < GS; ++q) {
        int t = tb + q * kG3W;
        if (t >= DB) continue;
        AT x = gx[q], y = gy[q], w = gz[q];
        if (isnan(x) || isnan(y) || isnan(w)) { x = 0; y = 0; w = 0; }
        unsigned fl = __builtin_amdgcn_raw_buffer_load_b8(fr_c, dir_of(t, rnB), 0, 0);
#pragma unroll
        for (int f = 0; f < 4; ++f)
            if (fl & (1u << f)) __builtin_nontemporal_store((ST)(fr[f][0] * x + fr[f][1] * y + fr[f][2] * w), pc + f * DB + t);
    }
}

// The per-lane part of a gradient wave: slice q's entry t (clamped to 0
// past the row's end) as a byte offset in a row.
template <typename ST, int GS>
__device__ __forceinline__ void g3_offsets(int tb, int DB, unsigned *to) {
#pragma unroll
    for (int q = 0; q < GS; ++q) {
        int t = tb + q * kG3W;
        to[q] = (unsigned)(t < DB ? t : 0) * (unsigned)sizeof(ST);
    }
}

// _get_gradient_larger, 3D (StaticBTESynthetic.cpp:1329-1371):
// grad = sum_j (P_j / L_x) (e_j - e_i), zero for a stencil of 3 or fewer
// cells (zero weights) or a NaN, stored as its projections on the cell's four
// face offsets, phi_f = r_f . grad with r_f the face centre minus the cell
// centre: the sources and the outgoing boundary values use the gradient only
// through these (_get_Re :1427-1500, _get_bound_ee :1502-1539), so one load
// per face replaces three.  One wave per (cell, group of GS 64-entry
// slices), the lanes over the entries, the neighbour rows read straight from
// memory (L2).  The WPB waves of a workgroup take Morton-consecutive cells,
// whose stencils (in increasing row order) overlap, so they read largely the
// same neighbour rows at about the same time.
template <typename ST, int GS, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_grad3g(const ST *__restrict__ e, ST *__restrict__ phi,
                                                     const int *__restrict__ sptr, const int *__restrict__ gidx,
                                                     const G3Ent<ST> *__restrict__ ent, const double *__restrict__ rec,
                                                     const unsigned char *__restrict__ oflag, int nc, int DB, int nB,
                                                     int ngroups) {
    int wv = __builtin_amdgcn_readfirstlane(xcd_block() * WPB + (int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    int sg = wv / nc, c = wv - sg * nc;
    if (sg >= ngroups) return;  // whole wave
    int tb = sg * GS * kG3W + lane;
    ST self[GS], gx[GS], gy[GS], gz[GS];
    unsigned to[GS];
    g3_offsets<ST, GS>(tb, DB, to);
    __amdgpu_buffer_rsrc_t ec = row_rsrc(e + (size_t)c * DB, (unsigned)(DB * sizeof(ST)));
#pragma unroll
    for (int q = 0; q < GS; ++q) {
        self[q] = ld_row<ST>(ec, to[q]);
        gx[q] = 0; gy[q] = 0; gz[q] = 0;
    }
    g3_entries<ST, GS, false>(e, gidx, ent, sptr[c], sptr[c + 1], to, DB, self, gx, gy, gz, nullptr);
    g3_store_phi<ST, GS>(phi, rec + (size_t)c * kS3D, oflag, c, tb, DB, nB, gx, gy, gz);
}

// The same gradient through sums over the mesh vertices, which halves the
// rows a cell reads.  The stencil of cell i is the union of the cells C(v)
// around its four vertices; a cell k of it is found m_ik times (the vertices
// it shares with i: 1, 2 or 3), so with j_k = (x_k - x_i) / L_x, M the 3 x 3
// inverse of the LSQ normal matrix and r_v one chosen cell of C(v),
//   sum_k j_k (e_k - e_i) = sum_v (S1_v + d_v S0_v)
//                           + sum_k c_k (e_k - e_i),  d_v = (x_v - x_i) / L_x,
//   S0_v = sum_{k in C(v)} (e_k - e_{r_v}),
//   S1_v = sum_{k in C(v)} ((x_k - x_v) / L_x) (e_k - e_{r_v}),
//   c_k  = -(m_ik - 1) j_k + sum_{v: r_v = k} sum_{l in C(v)} j_l,
// and grad = (M / L_x) times that: exact algebra, the weights c_k summing to
// zero.  Only the cells sharing an edge with i and the chosen r_v keep an own
// entry (about 20 against about 78), plus the four vertex sums.  Subtracting
// e_{r_v} keeps the vertex sums at the size of differences (no cancellation
// in FP32).
// k_vsum3: one wave per (vertex, slice group): the quad (S1x, S1y, S1z, S0)
// per entry, [vertex][entry][4].  Vertices are numbered in the order of
// their first cell, so consecutive waves share cells.
template <typename ST, int GS, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_vsum3(const ST *__restrict__ e, ST *__restrict__ vs,
                                                    const int *__restrict__ vptr, const int *__restrict__ vidx,
                                                    const G3Ent<ST> *__restrict__ vent, const int *__restrict__ vref,
                                                    int nv, int DB, int ngroups) {
    using Q4 = typename std::conditional<sizeof(ST) == 4, v4f, v4d>::type;
    int wv = __builtin_amdgcn_readfirstlane(xcd_block() * WPB + (int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    int sg = wv / nv, v = wv - sg * nv;
    if (sg >= ngroups) return;  // whole wave
    int tb = sg * GS * kG3W + lane;
    ST self[GS], gx[GS], gy[GS], gz[GS], s0[GS];
    unsigned to[GS];
    g3_offsets<ST, GS>(tb, DB, to);
    __amdgpu_buffer_rsrc_t er = row_rsrc(e + (size_t)vref[v] * DB, (unsigned)(DB * sizeof(ST)));
#pragma unroll
    for (int q = 0; q < GS; ++q) {
        self[q] = ld_row<ST>(er, to[q]);
        gx[q] = 0; gy[q] = 0; gz[q] = 0; s0[q] = 0;
    }
    g3_entries<ST, GS, true>(e, vidx, vent, vptr[v], vptr[v + 1], to, DB, self, gx, gy, gz, s0);
    Q4 *vv = reinterpret_cast<Q4 *>(vs + (size_t)v * DB * 4);
#pragma unroll
    for (int q = 0; q < GS; ++q) {
        int t = tb + q * kG3W;
        if (t < DB) vv[t] = Q4{gx[q], gy[q], gz[q], s0[q]};
    }
}

// k_grad3v's per-cell record: the four vertices (k_vsum3 numbering),
// a = M / L_x (row-major) and b_v = (M / L_x) d_v
template <typename ST>
struct alignas(16) G3VRec {
    int v[4];
    ST a[9];
    ST b[12];
};

// k_grad3v: the face projections from the vertex sums and the remaining
// entries (one wave per (cell, slice group), as k_grad3g)
template <typename ST, int GS, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_grad3v(const ST *__restrict__ e, const ST *__restrict__ vs,
                                                     ST *__restrict__ phi, const G3VRec<ST> *__restrict__ vrec,
                                                     const int *__restrict__ sptr, const int *__restrict__ gidx,
                                                     const G3Ent<ST> *__restrict__ ent, const double *__restrict__ rec,
                                                     const unsigned char *__restrict__ oflag, int nc, int DB, int nB,
                                                     int ngroups) {
    using Q4 = typename std::conditional<sizeof(ST) == 4, v4f, v4d>::type;
    int wv = __builtin_amdgcn_readfirstlane(xcd_block() * WPB + (int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    int sg = wv / nc, c = wv - sg * nc;
    if (sg >= ngroups) return;  // whole wave
    int tb = sg * GS * kG3W + lane;
    const G3VRec<ST> &R = vrec[c];
    ST self[GS], gx[GS], gy[GS], gz[GS], sx[GS], sy[GS], sz[GS];
    unsigned to[GS];
    g3_offsets<ST, GS>(tb, DB, to);
    __amdgpu_buffer_rsrc_t ec = row_rsrc(e + (size_t)c * DB, (unsigned)(DB * sizeof(ST)));
#pragma unroll
    for (int q = 0; q < GS; ++q) {
        self[q] = ld_row<ST>(ec, to[q]);
        gx[q] = 0; gy[q] = 0; gz[q] = 0; sx[q] = 0; sy[q] = 0; sz[q] = 0;
    }
    // the vertex sums: S1 summed over the four vertices, S0 weighted per vertex
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        __amdgpu_buffer_rsrc_t vb = row_rsrc(vs + (size_t)R.v[i] * DB * 4, (unsigned)(DB * 4 * sizeof(ST)));
        Q4 qv[GS];
#pragma unroll
        for (int q = 0; q < GS; ++q) qv[q] = ld_row<Q4>(vb, to[q] * 4);
#pragma unroll
        for (int q = 0; q < GS; ++q) {
            sx[q] += qv[q][0]; sy[q] += qv[q][1]; sz[q] += qv[q][2];
            gx[q] += R.b[3 * i] * qv[q][3]; gy[q] += R.b[3 * i + 1] * qv[q][3]; gz[q] += R.b[3 * i + 2] * qv[q][3];
        }
    }
#pragma unroll
    for (int q = 0; q < GS; ++q) {
        gx[q] += R.a[0] * sx[q] + R.a[1] * sy[q] + R.a[2] * sz[q];
        gy[q] += R.a[3] * sx[q] + R.a[4] * sy[q] + R.a[5] * sz[q];
        gz[q] += R.a[6] * sx[q] + R.a[7] * sy[q] + R.a[8] * sz[q];
    }
    g3_entries<ST, GS, false>(e, gidx, ent, sptr[c], sptr[c + 1], to, DB, self, gx, gy, gz, nullptr);
    g3_store_phi<ST, GS>(phi, rec + (size_t)c * kS3D, oflag, c, tb, DB, nB, gx, gy, gz);
}

// The same face projections on the matrix cores.  Tiles of up to four
// Morton-consecutive cells; the rows a tile needs (its cells and their
// stencils, at most 4 kGmKS) are the K dimension of one 16x16x4 product per
// four rows:
//   phi[cell q][face f][t] = sum_k A'[f, q][k] (E[row k][t] - E[ref][t]),
//   A'[f, q][k] = r_f(q) . W_q[k],  W_q[k] = P_k / L_x on q's stencil rows,
//   -sum P / L_x on q's own row, 0 elsewhere,
// so the 16 matrix rows are the 4 faces x 4 cells and the output needs no
// further arithmetic.  Every row of A' sums to zero, so subtracting one row
// of the tile (ref) from every staged value leaves the product unchanged
// while keeping the operands at the size of the differences.  One wave per
// tile, over slices of 16 NB entries: in K-step s lane l loads row
// 4 s + l / 16 at entries NB (l % 16) .. NB (l % 16) + NB - 1 of the slice,
// which is exactly its B operand of K-step s for each of the NB products
// (product j takes the entries NB n + j), so the rows go from memory (L2,
// shared by neighbouring tiles) straight into the products without LDS.
// The A' operands stay in registers for the whole tile; in FP32 each B
// register is reloaded with the next slice's value as soon as its product has
// issued, so the next slice's loads are in flight behind this slice's
// products (FP64 keeps two waves per SIMD instead).
// Output: lane l holds cell l / 16, faces 0..3 (elements), entries
// NB (l % 16) + j (FP64: D[4 i + l / 16][l % 16] in element i, so A' row
// m = 4 f + q; FP32: D[4 (l / 16) + i][l % 16], so m = 4 q + f).  A NaN
// zeroes the four projections (the reference's NaN -> 0 gradient); as in any
// dense product, a non-finite staged value reaches every cell of the tile at
// that entry.
constexpr int kGmKS = 40;  // K-steps per tile (at most 160 rows)
template <typename ST, int NB>
__global__ void __launch_bounds__(64) k_grad3m(const ST *__restrict__ e, ST *__restrict__ phi,
                                               const int *__restrict__ trows, const int *__restrict__ tcell,
                                               const ST *__restrict__ tA, const double *__restrict__ rec, Modes m,
                                               int n_tiles, int DB, int nB, int spg) {
    using BV = typename std::conditional<NB == 2, typename std::conditional<sizeof(ST) == 4, v2f, v2d>::type, ST>::type;
    // wave = (slice group, tile), tiles fastest: the waves that run together
    // work on the same slice group, so the rows that neighbouring tiles share
    // are read from L2
    int id = xcd_block(), sg = id / n_tiles, tile = id - sg * n_tiles;
    constexpr int W = 16 * NB;  // entries per slice
    constexpr bool kRoll = sizeof(ST) == 4;
    int s_begin = sg * spg * W, s_end = min(DB, s_begin + spg * W);
    if (s_begin >= DB) return;  // whole wave
    int lane = threadIdx.x, q = lane >> 4, n = lane & 15;
    ST a[kGmKS];
#pragma unroll
    for (int s = 0; s < kGmKS; ++s) a[s] = tA[((size_t)tile * kGmKS + s) * 64 + lane];
    // the tile's rows in LDS (registers go to the operands): K-step s, lane
    // group q reads srow[4 s + q]
    __shared__ int srow[kGmKS * 4];
    for (int i = lane; i < kGmKS * 4; i += 64) srow[i] = trows[(size_t)tile * kGmKS * 4 + i];
    __syncthreads();
    auto row = [&](int s) { return srow[4 * s + q]; };
    int cell = tcell[tile * 4 + q];
    int ref = trows[(size_t)tile * kGmKS * 4];
    // the lane's cell's face normals (the out-faces per direction)
    double fnx[4], fny[4], fnz[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const double *r = rec + (size_t)(cell >= 0 ? cell : 0) * kS3D + f * 10;
        fnx[f] = r[0]; fny[f] = r[1]; fnz[f] = r[2];
    }
    // lane's first entry of a slice, clamped to a whole vector inside the row
    auto ent0 = [&](int s0) {
        int t = s0 + NB * n;
        return t + NB <= DB ? t : DB - NB;
    };
    auto ld = [&](int r, int t) { return *reinterpret_cast<const BV *>(e + (size_t)r * DB + t); };
    BV b[kGmKS];
    BV rn;
    {
        int t0 = ent0(s_begin);
        rn = ld(ref, t0);
#pragma unroll
        for (int s = 0; s < kGmKS; ++s) b[s] = ld(row(s), t0);
    }
    for (int s0 = s_begin; s0 < s_end; s0 += W) {
        // the row indices are read from LDS again for every slice: kept in
        // registers (or as 64-bit addresses) they would halve the waves per SIMD
        asm volatile("" ::: "memory");
        int tn = ent0(s0 + W < s_end ? s0 + W : s0);  // the next slice (the last one again at the end)
        if constexpr (!kRoll) {
            // FP64: the registers of a second set of B operands would halve
            // the waves per SIMD; this slice's loads are issued here instead
            if (s0 != s_begin) {
                int tc = ent0(s0);
                rn = ld(ref, tc);
#pragma unroll
                for (int s = 0; s < kGmKS; ++s) {
                    b[s] = ld(row(s), tc);
                }
            }
        }
        BV rc = rn;
        if constexpr (kRoll) rn = ld(ref, tn);
        double v[NB][4];
        if constexpr (sizeof(ST) == 4) {
            v4f acc[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) acc[j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < kGmKS; ++s) {
                BV d = b[s] - rc;
                if constexpr (kRoll) b[s] = ld(row(s), tn);
                if constexpr (NB == 2) {
                    acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], d[0], acc[0], 0, 0, 0);
                    acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], d[1], acc[1], 0, 0, 0);
                } else {
                    acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], d, acc[0], 0, 0, 0);
                }
            }
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) v[j][i] = acc[j][i];
        } else {
            v4d acc[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < kGmKS; ++s) {
                BV d = b[s] - rc;
                if constexpr (kRoll) b[s] = ld(row(s), tn);
                if constexpr (NB == 2) {
                    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], d[0], acc[0], 0, 0, 0);
                    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], d[1], acc[1], 0, 0, 0);
                } else {
                    acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], d, acc[0], 0, 0, 0);
                }
            }
#pragma unroll
            for (int j = 0; j < NB; ++j)
#pragma unroll
                for (int i = 0; i < 4; ++i) v[j][i] = acc[j][i];
        }
        int t = s0 + NB * n;
        if (cell >= 0) {
            bool bad[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) bad[j] = isnan(v[j][0]) || isnan(v[j][1]) || isnan(v[j][2]) || isnan(v[j][3]);
            ST *pc = phi + (size_t)cell * 4 * DB + t;
            // out-faces only (as g3_store_phi); a pair shares its direction
            // when the bands per direction are even, else both entries'
            bool out[NB][4];
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                int d = min(t + j, DB - 1) / nB;
                double sx = m.sx[d], sy = m.sy[d], sz = m.sz[d];
#pragma unroll
                for (int i = 0; i < 4; ++i) out[j][i] = sx * fnx[i] + sy * fny[i] + sz * fnz[i] >= -kInFace;
            }
            bool pair = false;
            if constexpr (NB == 2) {
                // both entries of the pair in one store per face
                pair = t + 1 < DB;
                if (pair)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if (!out[0][i] && !out[1][i]) continue;
                        BV o;
                        o[0] = (ST)(bad[0] ? 0.0 : v[0][i]);
                        o[1] = (ST)(bad[1] ? 0.0 : v[1][i]);
                        __builtin_nontemporal_store(o, reinterpret_cast<BV *>(pc + i * DB));
                    }
            }
            if (!pair)
#pragma unroll
                for (int j = 0; j < NB; ++j) {
                    if (t + j >= DB) continue;
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (out[j][i]) __builtin_nontemporal_store((ST)(bad[j] ? 0.0 : v[j][i]), pc + j + i * DB);
                }
        }
    }
}

// The source of one transport step for every system (_get_Re,
// StaticBTESynthetic.cpp:1427-1500, less the new upwind values):
//   S = C (T_old - Tref) + Q hr tau - sum_out a (r . g_c) - sum_in,interior a (r' . g_nb)
//       - sum_in,boundary a eb,      a = tau (A/V) (v . n),
// written to each direction's upwind order for the sweep: S[dir][position][band]
// (sh < 0), or S[dir][band chunk][position][band in chunk] with chunks of
// 2^sh bands (the layout of k_sweep_lv).  r . g_c is the cell's own face
// projection phi_f and r' . g_nb the neighbour's projection on the shared
// face (k_grad3g), so each face is one load, from the row its sign of
// v . n selects.  One wave per cell row; the cell's face record (normal,
// A/V) sits in scalar registers, and every face issues its load before any
// term is formed.
template <typename ST>
__global__ void __launch_bounds__(64 * kSrc3Rows) k_source3(const int *__restrict__ posc, ST *__restrict__ S,
                                                            const ST *__restrict__ phi, const ST *__restrict__ eb,
                                                            const double *__restrict__ Told,
                                                            const double *__restrict__ Q,
                                                            const double *__restrict__ rec,
                                                            const int *__restrict__ reci, Modes m, int nc, int nB,
                                                            int nD, int sh) {
    extern __shared__ double lds[];
    double *mvg = lds, *mtau = mvg + nB, *mcap = mtau + nB, *mhr = mcap + nB, *msx = mhr + nB, *msy = msx + nD,
           *msz = msy + nD;
    int *pos = reinterpret_cast<int *>(msz + nD);
    int row0 = xcd_block() * kSrc3Rows;
    for (int i = threadIdx.x; i < nB; i += blockDim.x) {
        mvg[i] = m.vg[i];
        mtau[i] = m.tau[i];
        mcap[i] = m.cap[i];
        mhr[i] = m.hr[i];
    }
    for (int i = threadIdx.x; i < nD; i += blockDim.x) {
        msx[i] = m.sx[i];
        msy[i] = m.sy[i];
        msz[i] = m.sz[i];
    }
    for (int i = threadIdx.x; i < kSrc3Rows * nD; i += blockDim.x) {
        int r = row0 + i / nD;
        pos[i] = r < nc ? posc[(size_t)r * nD + i % nD] : 0;
    }
    __syncthreads();
    int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int c = row0 + w, lane = threadIdx.x & 63;
    if (c >= nc) return;
    int DB = nD * nB;
    const double *rd = rec + (size_t)c * kS3D;
    double fnx[4], fny[4], fnz[4], fav[4];
    const ST *fo[4], *fin[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const double *r = rd + f * 10;
        fnx[f] = r[0]; fny[f] = r[1]; fnz[f] = r[2]; fav[f] = r[3];
        // interior: 4 nb + the neighbour's slot of the shared face;
        // boundary: -(entry) - 1
        int q = reci[(size_t)c * kS3I + f];
        fo[f] = phi + ((size_t)c * 4 + f) * DB;
        fin[f] = q >= 0 ? phi + (size_t)q * DB : eb + (size_t)(-q - 1) * DB;
    }
    double To = Told[c] - m.eref, Qc = Q[c];
    const int *pc = pos + w * nD;
    for (int t = lane; t < DB; t += 64) {
        int d = t / nB, b = t - d * nB;
        double vb = mvg[b], tau = mtau[b], sx = msx[d], sy = msy[d], sz = msz[d];
        double dot[4], u[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            dot[f] = vb * sx * fnx[f] + vb * sy * fny[f] + vb * sz * fnz[f];
            u[f] = (dot[f] >= 0 ? fo[f] : fin[f])[t];
        }
        double s = To * mcap[b] + Qc * mhr[b] * mtau[b];
#pragma unroll
        for (int f = 0; f < 4; ++f) s -= tau * fav[f] * dot[f] * u[f];
        size_t si = sh == kSCell ? (size_t)c * DB + t
                    : sh < 0     ? ((size_t)d * nc + pc[d]) * nB + b
                                 : ((((size_t)d * (nB >> sh) + (b >> sh)) * nc + pc[d]) << sh) + (b & ((1 << sh) - 1));
        __builtin_nontemporal_store((ST)s, S + si);
    }
}

// FP32 storage, a band count divisible by V (2 or 4): the same source with V
// bands of one direction per lane, so every face value is one 8- or 16-byte
// load and the V sources go out in one store.  The vector memory pipeline
// bounds this kernel, and it issues a 16-byte load per lane at about the cost
// of an 8-byte one, so V = 4 moves twice the bytes per instruction.  Each
// entry's arithmetic is k_source3's.  A face's row (own projection or the
// upwind side's) is chosen once per direction by the sign of s . n, which the
// V bands share: their v s . n can differ in sign only where s . n vanishes
// to rounding, and there the term itself is that small; the projections are
// written where s . n >= -1e-12 on either side, so both rows hold values
// there.  Two iterations' loads are in flight.
template <int V>
__global__ void __launch_bounds__(64 * kSrc3Rows) k_source3_f32v(
    const int *__restrict__ posc, float *__restrict__ S, const float *__restrict__ phi, const float *__restrict__ eb,
    const double *__restrict__ Told, const double *__restrict__ Q, const double *__restrict__ rec,
    const int *__restrict__ reci, Modes m, int nc, int nB, int nD, int sh) {
    using VT = typename std::conditional<V == 4, v4f, v2f>::type;
    extern __shared__ double lds[];
    double *mvg = lds, *mtau = mvg + nB, *mcap = mtau + nB, *mhr = mcap + nB, *msx = mhr + nB, *msy = msx + nD,
           *msz = msy + nD;
    int *pos = reinterpret_cast<int *>(msz + nD);
    int row0 = xcd_block() * kSrc3Rows;
    for (int i = threadIdx.x; i < nB; i += blockDim.x) {
        mvg[i] = m.vg[i];
        mtau[i] = m.tau[i];
        mcap[i] = m.cap[i];
        mhr[i] = m.hr[i];
    }
    for (int i = threadIdx.x; i < nD; i += blockDim.x) {
        msx[i] = m.sx[i];
        msy[i] = m.sy[i];
        msz[i] = m.sz[i];
    }
    for (int i = threadIdx.x; i < kSrc3Rows * nD; i += blockDim.x) {
        int r = row0 + i / nD;
        pos[i] = r < nc ? posc[(size_t)r * nD + i % nD] : 0;
    }
    __syncthreads();
    int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    int c = row0 + w, lane = threadIdx.x & 63;
    if (c >= nc) return;
    int DB = nD * nB;
    const double *rd = rec + (size_t)c * kS3D;
    double fnx[4], fny[4], fnz[4], fav[4];
    const float *fo[4], *fin[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const double *r = rd + f * 10;
        fnx[f] = r[0]; fny[f] = r[1]; fnz[f] = r[2]; fav[f] = r[3];
        int q = reci[(size_t)c * kS3I + f];
        fo[f] = phi + ((size_t)c * 4 + f) * DB;
        fin[f] = q >= 0 ? phi + (size_t)q * DB : eb + (size_t)(-q - 1) * DB;
    }
    double To = Told[c] - m.eref, Qc = Q[c];
    const int *pc = pos + w * nD;
#pragma unroll 2
    for (int t = V * lane; t < DB; t += 64 * V) {
        int d = t / nB, b = t - d * nB;
        double sx = msx[d], sy = msy[d], sz = msz[d];
        VT u[4];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            double sn = sx * fnx[f] + sy * fny[f] + sz * fnz[f];
            u[f] = *reinterpret_cast<const VT *>((sn >= 0 ? fo[f] : fin[f]) + t);
        }
        VT o;
#pragma unroll
        for (int j = 0; j < V; ++j) {
            double vb = mvg[b + j], tb = mtau[b + j];
            double sj = To * mcap[b + j] + Qc * mhr[b + j] * tb;
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                double dot = vb * sx * fnx[f] + vb * sy * fny[f] + vb * sz * fnz[f];
                sj -= tb * fav[f] * dot * (double)u[f][j];
            }
            o[j] = (float)sj;
        }
        size_t si = sh == kSCell ? (size_t)c * DB + t
                    : sh < 0     ? ((size_t)d * nc + pc[d]) * nB + b
                                 : ((((size_t)d * (nB >> sh) + (b >> sh)) * nc + pc[d]) << sh) + (b & ((1 << sh) - 1));
        __builtin_nontemporal_store(o, reinterpret_cast<VT *>(S + si));
    }
}

// 16-byte vector moves of N values of type T (N * sizeof(T) a multiple of 16,
// the address 16-byte aligned)
template <typename T, int N>
__device__ __forceinline__ void vld(T (&v)[N], const T *p) {
    constexpr int NV = N * (int)sizeof(T) / 16;
    static_assert(NV * 16 == N * (int)sizeof(T), "whole 16-byte vectors");
    int4 t[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) t[k] = reinterpret_cast<const int4 *>(p)[k];
    __builtin_memcpy(v, t, sizeof(v));
}
template <typename T, int N>
__device__ __forceinline__ void vst(T *p, const T (&v)[N]) {
    constexpr int NV = N * (int)sizeof(T) / 16;
    int4 t[NV];
    __builtin_memcpy(t, v, sizeof(v));
#pragma unroll
    for (int k = 0; k < NV; ++k) reinterpret_cast<int4 *>(p)[k] = t[k];
}

// The per-system SparseLU solve (StaticBTESyntheticSolver.cpp:96-100 on
// _get_coefficient :784-840) as a level-synchronous upwind sweep for meshes
// whose levels are wide (tetrahedra: about 500 levels of about 600 cells per
// direction):
//   e_new = (S - sum_in,interior tau v g_f e_new_nb) / (1 + tau v G_out).
// One TPB-thread workgroup per (direction, chunk of BC bands), so that every
// CU holds several.  V lanes take one position, each lane BC / V bands of
// it as whole 16-byte vectors: with V = 1 a lane moves the position's whole
// chunk (its record read once), with V lanes per position the chunk's
// upwind reads and stores are one 16-byte vector per lane, and the V lanes
// of a position touch one cache line together instead of one lane touching
// it V times (the L1 processes scattered reads line by line).  The upwind
// values are read from e, where this workgroup wrote them on earlier levels;
// levels are separated by workgroup barriers.
//
// Software pipeline: the positions are taken in chunks of TPB NP / V (a wide
// level is several chunks, with no barrier between them), and a chunk's
// records and sources are loaded during the previous chunk, issued after its
// upwind loads, so that only the upwind loads (L2) and the arithmetic lie on
// the chain of dependent levels.  The level pointers of the direction sit in LDS.  The
// chunks of one direction take consecutive workgroup ids, which xcd_block
// places on one XCD, so they share its L2 for the records and for the
// partial lines of the e rows they write.
//
// V = 0: chunks of up to BC bands with one 16-byte vector per lane, the lanes
// per position set per workgroup (the last chunk of a direction may be
// narrower: 40 FP32 bands are chunks of 16, 16 and 8, four or two lanes per
// position), S per position.
template <typename ST, int BC, int NP, int TPB, int V, bool CM>
__global__ void __launch_bounds__(TPB) k_sweep_lv(const LvRec *__restrict__ rec, const int *__restrict__ levptr,
                                                  const int *__restrict__ dlev, const ST *S, ST *e, Modes m, int nc,
                                                  int nB, int nD, int smajor) {
    constexpr int BL = V > 0 ? BC / V : 16 / (int)sizeof(ST);  // bands per lane
    static_assert((V == 0 || (BC % V == 0 && TPB % V == 0)) && BC % BL == 0 && (BL * sizeof(ST)) % 16 == 0,
                  "whole 16-byte vectors per lane");
    extern __shared__ int slp[];  // level pointers of the direction
    int nchunk = V > 0 ? nB / BC : (nB + BC - 1) / BC;
    int id = xcd_block();
    if (id >= nD * nchunk) return;  // whole workgroup
    int d = id / nchunk, chunk = id - d * nchunk;
    int vl = V > 0 ? V : min(BC, nB - chunk * BC) / BL;  // lanes per position
    const int PP = TPB / vl;                                // positions per pass
    int tid = (int)threadIdx.x / vl, sub = (int)threadIdx.x - tid * vl;  // position slot, vector of the chunk
    int wpos = __builtin_amdgcn_readfirstlane(tid);                      // the wave's first position slot
    int b0 = chunk * BC + sub * BL;
    double tv[BL];
#pragma unroll
    for (int i = 0; i < BL; ++i) tv[i] = m.tau[b0 + i] * m.vg[b0 + i];
    int L = dlev[d + 1] - dlev[d] - 1;
    for (int i = threadIdx.x; i <= L; i += TPB) slp[i] = levptr[dlev[d] + i];
    __syncthreads();
    const LvRec *rd = rec + (size_t)d * nc;
    // S[dir][chunk][position][BC], S[dir][position][band] (smajor), or
    // S[cell][dir][band] (CM: read at the position's cell with the upwind
    // values, once its record has arrived)
    const ST *Sd = smajor ? S + (size_t)d * nc * nB + b0 : S + ((size_t)d * nchunk + chunk) * nc * BC + sub * BL;
    size_t sst = smajor ? (size_t)nB : (size_t)BC;
    size_t DB = (size_t)nD * nB;
    const ST *Sc = S + (size_t)d * nB + b0;
    ST *ed = e + (size_t)d * nB + b0;
    // Every load of the pipelined part is unconditional (lanes past the end
    // of a level re-read its last position, absent upwind cells read row 0,
    // both masked afterwards): with no branch around a load, the wait before
    // the products counts exactly this level's upwind loads and leaves the
    // next level's records and sources in flight.
    auto upwind = [&](const LvRec &r, ST(&u)[3][BL]) {
#pragma unroll
        for (int j = 0; j < 3; ++j) vld(u[j], ed + (size_t)max(r.c[j], 0) * DB);
    };
    auto solve = [&](const LvRec &r, const ST(&s)[BL], const ST(&u)[3][BL], bool store) {
        ST en[BL];
#pragma unroll
        for (int b = 0; b < BL; ++b) {
            double v = s[b];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double t = tv[b] * r.g[j] * (double)u[j][b];
                v -= r.c[j] >= 0 ? t : 0.0;  // an absent upwind cell: exactly v
            }
            en[b] = (ST)(v / (1.0 + tv[b] * r.G));
        }
        if (store) vst(ed + (size_t)r.cell * DB, en);
    };
    // The positions of a direction are taken in chunks of up to PP NP: a
    // level is one chunk or several, and a barrier separates the levels, not
    // the chunks of one level (whose upwind cells all lie on earlier levels).
    // Chunk (l, p0) from the buffers (rc, sc); the next chunk's positions
    // into (rn, sn); the last chunk re-reads its own.
    auto next = [&](int l, int p0, int &nl, int &np0) {
        if (p0 + PP * NP < slp[l + 1]) { nl = l; np0 = p0 + PP * NP; }
        else { nl = l + 1; np0 = slp[l + 1]; }
    };
    auto fetch = [&](int l, int p0, LvRec(&rr)[NP], ST(&ss)[NP][BL]) {
        int a1 = slp[l + 1];
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            int p = min(p0 + tid + i * PP, a1 - 1);
            rr[i] = rd[p];
            if constexpr (!CM) vld(ss[i], Sd + (size_t)p * sst);
        }
    };
    auto chunk_step = [&](int l, int p0, LvRec(&rc)[NP], ST(&sc)[NP][BL], LvRec(&rn)[NP], ST(&sn)[NP][BL],
                          int &nl, int &np0) {
        int a1 = slp[l + 1];
        next(l, p0, nl, np0);
        ST u[NP][3][BL];
#pragma unroll
        for (int i = 0; i < NP; ++i) upwind(rc[i], u[i]);
        if constexpr (CM) {
#pragma unroll
            for (int i = 0; i < NP; ++i) vld(sc[i], Sc + (size_t)rc[i].cell * DB);
        }
        fetch(nl < L ? nl : l, nl < L ? np0 : p0, rn, sn);
        // a wave none of whose positions lies in the chunk skips the
        // arithmetic (a uniform branch with no load in it)
#pragma unroll
        for (int i = 0; i < NP; ++i)
            if (p0 + wpos + i * PP < a1) solve(rc[i], sc[i], u[i], p0 + tid + i * PP < a1);
        if (nl != l) __syncthreads();
    };
    LvRec rA[NP], rB[NP];
    ST sA[NP][BL], sB[NP][BL];
    if (L > 0) {
        int l = 0, p0 = slp[0], nl, np0;
        fetch(0, p0, rA, sA);
        for (;;) {
            chunk_step(l, p0, rA, sA, rB, sB, nl, np0);
            if (nl >= L) break;
            l = nl; p0 = np0;
            chunk_step(l, p0, rB, sB, rA, sA, nl, np0);
            if (nl >= L) break;
            l = nl; p0 = np0;
        }
    }
}

// _get_bound_ee (StaticBTESynthetic.cpp:1502-1539): outgoing boundary values
// e + r . grad e, the cell's face projection phi_f of k_grad3g; one wave per
// boundary entry.
template <typename ST>
__global__ void __launch_bounds__(256) k_bound_out3(const ST *__restrict__ e, const ST *__restrict__ phi,
                                                    ST *__restrict__ eb, Modes m, const int *__restrict__ ent_list,
                                                    const int *__restrict__ be_cell, const int *__restrict__ be_slot,
                                                    const double *__restrict__ snx, const double *__restrict__ sny,
                                                    const double *__restrict__ snz, int n_ent, int nB, int nD) {
    // the band and direction constants in LDS (read per lane there instead
    // of from memory)
    extern __shared__ double bo_lds[];
    double *mvg = bo_lds, *msx = mvg + nB, *msy = msx + nD, *msz = msy + nD;
    for (int i = threadIdx.x; i < nB; i += blockDim.x) mvg[i] = m.vg[i];
    for (int i = threadIdx.x; i < nD; i += blockDim.x) {
        msx[i] = m.sx[i];
        msy[i] = m.sy[i];
        msz[i] = m.sz[i];
    }
    __syncthreads();
    int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    if (k >= n_ent) return;
    int en = ent_list[k], c = be_cell[en], s = be_slot[en];
    double nx = snx[s], ny = sny[s], nz = snz[s];
    int DB = nD * nB;
    float rnB = 1.0f / (float)nB;
    const ST *ec = e + (size_t)c * DB, *pf = phi + ((size_t)c * 4 + (s - c * 6)) * DB;
    ST *eo = eb + (size_t)en * DB;
    for (int t = lane; t < DB; t += 64) {
        int d = dir_of(t, rnB), b = t - d * nB;
        double vb = mvg[b];
        double dot = vb * msx[d] * nx + vb * msy[d] * ny + vb * msz[d] * nz;
        if (dot >= 0) eo[t] = (ST)((double)ec[t] + (double)pf[t]);
    }
}

// Row moments Y_0..10 (see the top of this file) of the cell rows of e and
// of the boundary-entry rows of eb.  One wave per row with the lanes over the
// bands: each lane sums its band over the directions with the direction
// factors (w, w s_a s_b, w s_a; a set-up table, kDirF per direction, read by
// scalar loads), then multiplies by the band factors and the wave adds the
// bands in a fixed order.
constexpr int kDirF = 10;
template <typename ST>
__global__ void __launch_bounds__(64 * kMomRows) k_mom3(const ST *__restrict__ e, const ST *__restrict__ eb,
                                                        const int *__restrict__ ent_list, const double *__restrict__ dirf,
                                                        Modes m, double *__restrict__ Y, int nc, int n_ent, int nB,
                                                        int nD) {
    int r = xcd_block() * kMomRows + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= nc + n_ent) return;  // whole wave
    const ST *row = r < nc ? e + (size_t)r * nD * nB : eb + (size_t)ent_list[r - nc] * nD * nB;
    int b = lane < nB ? lane : 0;  // nB <= 64; lanes past nB add nothing (their band factors are 0)
    double f[kDirF];
#pragma unroll
    for (int j = 0; j < kDirF; ++j) f[j] = 0;
#pragma unroll 8
    for (int d = 0; d < nD; ++d) {
        double v = row[d * nB + b];
        const double *df = dirf + (size_t)d * kDirF;
#pragma unroll
        for (int j = 0; j < kDirF; ++j) f[j] += df[j] * v;
    }
    bool act = lane < nB;
    double vb = act ? m.vg[b] : 0.0, lrc = act ? m.lrc[b] : 0.0;
    double h = act && m.hr[b] > 0 ? 1.0 : 0.0;
    double kk = h * m.kappa * lrc, tv = h * (act ? m.tau[b] : 0.0) * vb * vb;
    double y[11];
    y[0] = kk * f[0];
#pragma unroll
    for (int j = 1; j < 7; ++j) y[j] = tv * f[j];
    y[7] = lrc * f[0];
    y[8] = vb * f[7];
    y[9] = vb * f[8];
    y[10] = vb * f[9];
#pragma unroll
    for (int j = 0; j < 11; ++j) y[j] = wave_sum(y[j]);
    // boundary entry en's moments are row nc + en
    size_t yr = r < nc ? (size_t)r : (size_t)nc + ent_list[r - nc];
    if (lane < 11) {
        double v = y[0];
#pragma unroll
        for (int j = 1; j < 11; ++j) v = lane == j ? y[j] : v;
        Y[yr * kY + lane] = v;
    }
}

// Vertex moments: the inverse-distance combination of the sources' row
// moments (cells, or the boundary entries of a boundary node), as
// _set_vertex_energydensity combines e (StaticBTESynthetic.cpp:1541-1616);
// a node without sources gets 0 (the reference's NaN -> 0).
__global__ void k_vert3(const double *__restrict__ Y, const int *__restrict__ vptr, const int *__restrict__ vsrc,
                        const double *__restrict__ vwt, double *__restrict__ YV, int nn, int nc) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nn) return;
    double a[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = vptr[v]; k < vptr[v + 1]; ++k) {
        int s = vsrc[k];
        const double *y = Y + (size_t)(s >= 0 ? s : nc + (-s - 1)) * kY;
        double wt = vwt[k];
#pragma unroll
        for (int j = 0; j < 7; ++j) a[j] += wt * y[j];
    }
#pragma unroll
    for (int j = 0; j < 7; ++j) YV[(size_t)v * kYV + j] = a[j];
}

// Per cell: T = Y_7, q = Y_8..10 (plus the moments of C_b Tref) and the
// non-Fourier residual, sum over the face slots of
//   sum_j ce_j (Y_j[other] - Y_j[cell]) + cv1_j (YV_j[u1] - YV_j[u0]) + cv2_j (YV_j[u2] - YV_j[u1]).
__global__ void k_fin3(const double *__restrict__ Y, const double *__restrict__ YV, const int *__restrict__ fi,
                       const double *__restrict__ fd, Modes m, double *__restrict__ T, double *__restrict__ qx,
                       double *__restrict__ qy, double *__restrict__ qz, double *__restrict__ RM, int nc) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const double *yc = Y + (size_t)c * kY;
    T[c] = yc[7] + m.offT;
    qx[c] = yc[8] + m.offQx;
    qy[c] = yc[9] + m.offQy;
    qz[c] = yc[10] + m.offQz;
    double r = 0;
    for (int f = 0; f < 4; ++f) {
        const int *q = fi + ((size_t)c * 4 + f) * 4;
        if (q[0] < 0) continue;  // no term (specular wall)
        const double *yo = Y + (size_t)q[0] * kY;
        const double *v0 = YV + (size_t)q[1] * kYV, *v1 = YV + (size_t)q[2] * kYV, *v2 = YV + (size_t)q[3] * kYV;
        const double *k = fd + ((size_t)c * 4 + f) * kFin;
#pragma unroll
        for (int j = 0; j < 7; ++j)
            r += k[j] * (yo[j] - yc[j]) + k[7 + j] * (v1[j] - v0[j]) + k[14 + j] * (v2[j] - v1[j]);
    }
    RM[c] = r;
}

// Right-hand side of the synthetic solve (_get_Re_macro, 3D branch,
// StaticBTESynthetic.cpp:2307-2407) or of the Fourier solve (_get_Re,
// StaticFourier.cpp:520-613), negated for the SPD form: the boundary
// temperatures' terms and the non-orthogonal terms kappa A a2 . g_t, with the
// tangential gradient g_t of the vertex temperatures a fixed combination of
// the face's two vertex differences (w1, w2 per slot).
__global__ void k_rhs3(const double *__restrict__ RM, const double *__restrict__ Q, const double *__restrict__ vol,
                       const double *__restrict__ bval, const double *__restrict__ TV, const int *__restrict__ ri,
                       const double *__restrict__ rd, double *__restrict__ rhs, int nc) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double r = RM ? RM[c] : 0.0;
    r -= Q[c] * vol[c];
    for (int f = 0; f < 4; ++f) {
        const int *q = ri + ((size_t)c * 4 + f) * 4;
        const double *k = rd + ((size_t)c * 4 + f) * 3;
        if (q[0] >= 0) r += -k[0] * bval[q[0]];
        if (q[1] >= 0) r -= k[1] * (TV[q[2]] - TV[q[1]]) + k[2] * (TV[q[3]] - TV[q[2]]);
    }
    rhs[c] = -r;
}

// fn(lo, hi) over [0, n) in contiguous blocks on up to 16 host threads
template <typename F>
void host_blocks(int n, F fn) {
    int nth = std::max(1, std::min({16, (int)std::thread::hardware_concurrency(), (n + 1023) / 1024}));
    if (nth <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    for (int w = 0; w < nth; ++w) {
        int lo = (int)((long)n * w / nth), hi = (int)((long)n * (w + 1) / nth);
        th.emplace_back([&fn, lo, hi]() { fn(lo, hi); });
    }
    for (auto &x : th) x.join();
}

// (J^T J)^{-1} of a J with three columns (Get_inverse_3D,
// src/utility/utility.cpp:488-548: the adjugate over the determinant)
void jtj_inv(const std::vector<double> &J, int m, double I[3][3]) {
    double A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            A[i][j] = 0;
            for (int k = 0; k < m; ++k) A[i][j] += J[k * 3 + i] * J[k * 3 + j];
        }
    double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) + A[0][1] * (A[1][2] * A[2][0] - A[1][0] * A[2][2]) +
                 A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    I[0][0] = 1.0 / det * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
    I[1][0] = 1.0 / det * (A[0][2] * A[2][1] - A[0][1] * A[2][2]);
    I[2][0] = 1.0 / det * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
    I[0][1] = 1.0 / det * (A[1][2] * A[2][0] - A[1][0] * A[2][2]);
    I[1][1] = 1.0 / det * (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
    I[2][1] = 1.0 / det * (A[1][0] * A[0][2] - A[0][0] * A[1][2]);
    I[0][2] = 1.0 / det * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    I[1][2] = 1.0 / det * (A[0][1] * A[2][0] - A[0][0] * A[2][1]);
    I[2][2] = 1.0 / det * (A[0][0] * A[1][1] - A[1][0] * A[0][1]);
}

// (J^T J)^{-1} J^T of a J with three columns (Get_inverse_3D,
// src/utility/utility.cpp:488-548), row-major 3 x m.
void lsq3(const std::vector<double> &J, int m, std::vector<double> &P) {
    double I[3][3];
    jtj_inv(J, m, I);
    P.assign((size_t)3 * m, 0.0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < m; ++j)
            for (int k = 0; k < 3; ++k) P[(size_t)i * m + j] += I[i][k] * J[j * 3 + k];
}

// the face's vertex-gradient matrix: J rows [n; (x_{k+1} - x_k) / L_x] of the
// slot's vertices in their order (_get_face_matrix, StaticBTESynthetic.cpp:
// 1617-1649, and Get_gradient_3D's J in StaticFourier.cpp:571-593)
void face_lsq(const Problem &p, size_t s, double invLx, std::vector<double> &P) {
    int nv = p.face_nvert[s];
    std::vector<double> J((size_t)nv * 3);
    J[0] = p.face_nx[s]; J[1] = p.face_ny[s]; J[2] = p.face_nz[s];
    for (int k = 0; k + 1 < nv; ++k) {
        int a = p.face_vert[s * 4 + k], b = p.face_vert[s * 4 + k + 1];
        J[(k + 1) * 3 + 0] = (invLx * (p.node_x[b] - p.node_x[a]));
        J[(k + 1) * 3 + 1] = (invLx * (p.node_y[b] - p.node_y[a]));
        J[(k + 1) * 3 + 2] = (invLx * (p.node_z[b] - p.node_z[a]));
    }
    lsq3(J, nv, P);
}

}  // namespace

// ======================================================================
// host side
// ======================================================================

// Device tables of the tetrahedral path (called from build() for dim 3).
// be_slot: boundary entry -> face slot.
bool Solver::build3d(const std::vector<int> &be_slot) {
    const Problem &p = P;
    auto t0 = std::chrono::steady_clock::now();
    phase3d_.clear();
    auto mark = [&](const char *what) {
        auto t = std::chrono::steady_clock::now();
        char b[96];
        snprintf(b, sizeof b, "%s%s %.2f s", phase3d_.empty() ? "" : ", ", what,
                 std::chrono::duration<double>(t - t0).count());
        phase3d_ += b;
        t0 = t;
    };
    size_t ns = (size_t)nc * 6;
    double invLx = 1.0 / p.deck.Lx;
    for (int c = 0; c < nc; ++c)
        if (p.face_size[c] != 4) { err_ = "3D meshes of tetrahedra only"; return false; }
    // the gradient kernels find an entry's direction by an FP32 quotient
    if ((long long)nD * nB >= (1LL << 21)) {
        err_ = "3D meshes take fewer than 2^21 direction x band systems";
        return false;
    }
    // ---- source records (kept: the face offsets also weight the gradient's
    // face projections)
    std::vector<double> rec_h((size_t)nc * kS3D, 0.0);
    {
        std::vector<double> &rec = rec_h;
        std::vector<int> reci((size_t)nc * kS3I, 0);
        std::vector<char> bad(nc, 0);
        host_blocks(nc, [&](int lo, int hi) {
        for (int c = lo; c < hi; ++c)
            for (int f = 0; f < 4; ++f) {
                size_t s = (size_t)c * 6 + f;
                double *r = &rec[(size_t)c * kS3D + f * 10];
                r[0] = p.face_nx[s]; r[1] = p.face_ny[s]; r[2] = p.face_nz[s];
                r[3] = p.face_area[s] / p.volume[c];
                r[4] = p.face_cx[s] - p.cx[c]; r[5] = p.face_cy[s] - p.cy[c]; r[6] = p.face_cz[s] - p.cz[c];
                int fb = p.face_bound[s];
                if (fb == -1) {
                    int n = p.face_neighbor[s];
                    r[7] = p.face_cx[s] - p.cx[n]; r[8] = p.face_cy[s] - p.cy[n]; r[9] = p.face_cz[s] - p.cz[n];
                    // the neighbour's slot of the shared face: its phi row
                    int fn = -1;
                    for (int k = 0; k < 4; ++k)
                        if (p.face_neighbor[(size_t)n * 6 + k] == c && p.face_bound[(size_t)n * 6 + k] == -1) fn = k;
                    if (fn < 0) { bad[c] = 1; continue; }
                    reci[(size_t)c * kS3I + f] = n * 4 + fn;
                } else {
                    int ibc = -1;
                    for (int k = 0; k < 2; ++k)
                        if (p.bound_cell[fb * 2 + k] == c && p.bound_face[fb * 2 + k] == f) ibc = k;
                    if (ibc < 0) { bad[c] = 2; continue; }
                    reci[(size_t)c * kS3I + f] = -(fb * 2 + ibc) - 1;
                }
            }
        });
        for (int c = 0; c < nc; ++c)
            if (bad[c]) {
                err_ = bad[c] == 1 ? "internal error: interior face without its neighbour's side"
                                   : "boundary face without a boundary entry";
                return false;
            }
        if (!upload(d_s3d, rec) || !upload(d_s3i, reci)) return false;
        // the moments' direction factors: w, w s_a s_b (xx xy xz yy yz zz), w s_a
        std::vector<double> dirf((size_t)nD * kDirF);
        for (int d = 0; d < nD; ++d) {
            double w = p.weight[d], x = p.dir_x[d], y = p.dir_y[d], z = p.dir_z[d];
            double *f = &dirf[(size_t)d * kDirF];
            f[0] = w;
            f[1] = w * (x * x); f[2] = w * (x * y); f[3] = w * (x * z);
            f[4] = w * (y * y); f[5] = w * (y * z); f[6] = w * (z * z);
            f[7] = w * x; f[8] = w * y; f[9] = w * z;
        }
        if (!upload(d_dirf, dirf)) return false;
        size_t nfl = (size_t)nc * nD;
        if (!alloc(d_oflag, nfl)) return false;
        hipLaunchKernelGGL(k_oflags3, dim3((unsigned)((nfl + 255) / 256)), dim3(256), 0, s_, d_s3d.p, modes(), nc, nD,
                           d_oflag.p);
        if (!check(hipGetLastError(), "k_oflags3")) return false;
    }
    // ---- LSQ weights (_get_CellMatrix_larger + Get_inverse_3D): padded row
    // tables for the Fourier flux kernel, and the k_grad3 tiles
    std::vector<double> P3;  // [comp][nbr entry], weights (not yet / L_x)
    std::vector<int> gok(nc, 0);
    {
        int nnz = p.nbr_ptr[nc];
        P3.assign((size_t)3 * nnz, 0.0);
        host_blocks(nc, [&](int lo, int hi) {
            std::vector<double> J, Pc;
            for (int c = lo; c < hi; ++c) {
                int a = p.nbr_ptr[c], m = p.nbr_ptr[c + 1] - a;
                if (m <= 3) continue;
                J.assign((size_t)m * 3, 0.0);
                for (int k = 0; k < m; ++k) {
                    int n = p.nbr_idx[a + k];
                    J[k * 3 + 0] = (invLx * (p.cx[n] - p.cx[c]));
                    J[k * 3 + 1] = (invLx * (p.cy[n] - p.cy[c]));
                    J[k * 3 + 2] = (invLx * (p.cz[n] - p.cz[c]));
                }
                lsq3(J, m, Pc);
                for (int i = 0; i < 3; ++i)
                    for (int k = 0; k < m; ++k) P3[(size_t)i * nnz + a + k] = Pc[(size_t)i * m + k];
                gok[c] = 1;
            }
        });
        // the Fourier flux kernel's table: each cell's stencil padded to
        // whole batches with its own row at weight 0
        std::vector<int> qptr(nc + 1, 0);
        for (int c = 0; c < nc; ++c) {
            int m = p.nbr_ptr[c + 1] - p.nbr_ptr[c];
            qptr[c + 1] = qptr[c] + (m + kStencilBatch - 1) / kStencilBatch * kStencilBatch;
        }
        gnnz = qptr[nc];
        std::vector<int> qidx(gnnz);
        std::vector<double> qa((size_t)3 * gnnz, 0.0);
        host_blocks(nc, [&](int lo, int hi) {
            for (int c = lo; c < hi; ++c) {
                int o = qptr[c];
                for (int k = p.nbr_ptr[c]; k < p.nbr_ptr[c + 1]; ++k, ++o) {
                    qidx[o] = p.nbr_idx[k];
                    for (int i = 0; i < 3; ++i) qa[(size_t)i * gnnz + o] = P3[(size_t)i * nnz + k];
                }
                for (; o < qptr[c + 1]; ++o) qidx[o] = c;
            }
        });
        if (!upload(d_gptr, qptr) || !upload(d_gidx, qidx) || !upload(d_gP, qa) || !upload(d_gok, gok)) return false;
    }
    mark("source records and LSQ weights");
    // ---- vertex-sum gradient (k_vsum3 + k_grad3v): the cells around each
    // vertex, one chosen cell r_v per vertex, per-vertex and per-cell tables
    gv_ok = build_gv(invLx);
    if (!err_.empty()) return false;
    mark("vertex sums");
    // the direct gather and the matrix-core tiles read about twice the rows
    // of the vertex sums per cell: their tables are built only when the
    // vertex sums do not take the mesh or the kernel is asked for
    bool need_g = !gv_ok || opt.gradient_kernel == GBTE_GRAD_GATHER;
    bool need_m = !gv_ok || opt.gradient_kernel == GBTE_GRAD_MATRIX;
    // ---- gradient stencil tables: global neighbour rows, padded per cell to
    // a multiple of kG3Pad with the cell's own row at weight 0 (each adds an
    // exact zero)
    {
        int nnz = p.nbr_ptr[nc];
        if (need_g) {
        std::vector<int> gptr2(nc + 1, 0);
        for (int cc = 0; cc < nc; ++cc) {
            int m = p.nbr_ptr[cc + 1] - p.nbr_ptr[cc];
            gptr2[cc + 1] = gptr2[cc] + (m + kG3Pad - 1) / kG3Pad * kG3Pad;
        }
        size_t n2 = (size_t)gptr2[nc];
        std::vector<int> gidx2(n2);
        std::vector<int> gpos(n2);  // stencil position per entry (-1: padding)
        host_blocks(nc, [&](int lo, int hi) {
            std::vector<std::pair<int, int>> st;
            for (int cc = lo; cc < hi; ++cc) {
                // in increasing neighbour order: the waves of a workgroup take
                // consecutive cells, whose stencils overlap, and then read the
                // shared rows at about the same time (L1 hits)
                st.clear();
                for (int k = p.nbr_ptr[cc]; k < p.nbr_ptr[cc + 1]; ++k) st.push_back({p.nbr_idx[k], k});
                std::sort(st.begin(), st.end());
                int o = gptr2[cc];
                for (auto &q : st) { gidx2[o] = q.first; gpos[o] = q.second; ++o; }
                for (; o < gptr2[cc + 1]; ++o) { gidx2[o] = cc; gpos[o] = -1; }
            }
        });
        if (fp32_) {
            std::vector<G3Ent<float>> e2(n2);
            host_blocks((int)n2, [&](int lo, int hi) {
                for (int k = lo; k < hi; ++k) {
                    int j = gpos[k];
                    e2[k].loc = gidx2[k];
                    e2[k].px = j < 0 ? 0.0f : (float)(P3[j] * invLx);
                    e2[k].py = j < 0 ? 0.0f : (float)(P3[(size_t)nnz + j] * invLx);
                    e2[k].pz = j < 0 ? 0.0f : (float)(P3[2 * (size_t)nnz + j] * invLx);
                }
            });
            if (!alloc(d_g3g_ent, n2 * sizeof(e2[0])) ||
                !check(hipMemcpy(d_g3g_ent.p, e2.data(), n2 * sizeof(e2[0]), hipMemcpyHostToDevice), "H2D"))
                return false;
        } else {
            std::vector<G3Ent<double>> e2(n2);
            host_blocks((int)n2, [&](int lo, int hi) {
                for (int k = lo; k < hi; ++k) {
                    int j = gpos[k];
                    e2[k].loc = gidx2[k];
                    e2[k].pad = 0;
                    e2[k].px = j < 0 ? 0.0 : P3[j] * invLx;
                    e2[k].py = j < 0 ? 0.0 : P3[(size_t)nnz + j] * invLx;
                    e2[k].pz = j < 0 ? 0.0 : P3[2 * (size_t)nnz + j] * invLx;
                }
            });
            if (!alloc(d_g3g_ent, n2 * sizeof(e2[0])) ||
                !check(hipMemcpy(d_g3g_ent.p, e2.data(), n2 * sizeof(e2[0]), hipMemcpyHostToDevice), "H2D"))
                return false;
        }
        if (!upload(d_g3g_ptr, gptr2) || !upload(d_g3g_idx, gidx2)) return false;
        }
        mark("gather tables");
        // ---- matrix-core tiles (k_grad3m): up to four consecutive cells while
        // the rows they need fit kGmKS K-steps
        if (need_m) {
            // tile boundaries and rows (sequential, greedy), then the A'
            // operands of each tile (host threads)
            std::vector<int> trows, tcell, tnrow;
            std::vector<int> loc_of(nc, -1), cur;
            auto rows_with = [&](int cc) {
                int add = loc_of[cc] < 0 ? 1 : 0;
                for (int k = p.nbr_ptr[cc]; k < p.nbr_ptr[cc + 1]; ++k) add += loc_of[p.nbr_idx[k]] < 0 ? 1 : 0;
                return (int)cur.size() + add;
            };
            auto add_row = [&](int r) {
                if (loc_of[r] < 0) { loc_of[r] = (int)cur.size(); cur.push_back(r); }
            };
            const int cap = 4 * kGmKS;
            bool ok = true;
            int c = 0;
            while (c < nc && ok) {
                cur.clear();
                int c0 = c;
                while (c < nc && c - c0 < 4 && (c == c0 || rows_with(c) <= cap)) {
                    if (rows_with(c) > cap) { ok = false; break; }
                    add_row(c);
                    for (int k = p.nbr_ptr[c]; k < p.nbr_ptr[c + 1]; ++k) add_row(p.nbr_idx[k]);
                    ++c;
                }
                if (!ok) break;
                for (int qq = 0; qq < 4; ++qq) tcell.push_back(c0 + qq < c ? c0 + qq : -1);
                for (int k = 0; k < cap; ++k) trows.push_back(k < (int)cur.size() ? cur[k] : cur[0]);
                tnrow.push_back((int)cur.size());
                for (int r : cur) loc_of[r] = -1;
            }
            gm_ok = ok;
            if (ok) {
                // validate: every tile row names a cell, every cell is in one tile
                for (int r : trows)
                    if (r < 0 || r >= nc) { err_ = "internal error: matrix-core tile row"; return false; }
                gm_ntiles = (int)tcell.size() / 4;
                int seen = 0;
                for (int t2 = 0; t2 < gm_ntiles; ++t2)
                    for (int qq = 0; qq < 4; ++qq) seen += tcell[t2 * 4 + qq] >= 0;
                if (seen != nc) { err_ = "internal error: matrix-core tiles do not cover the cells"; return false; }
                // A'[m][k] in double, m = 4 f + q (FP64) or 4 q + f (FP32),
                // stored [tile][K-step][lane] in the storage type
                size_t per = (size_t)kGmKS * 64;
                std::vector<double> tA64(fp32_ ? 0 : (size_t)gm_ntiles * per);
                std::vector<float> tA32(fp32_ ? (size_t)gm_ntiles * per : 0);
                int nth = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
                std::vector<std::thread> th;
                for (int w = 0; w < nth; ++w)
                    th.emplace_back([&, w]() {
                        std::vector<int> loc(nc, -1);
                        std::vector<double> W((size_t)16 * cap);
                        for (int t2 = w; t2 < gm_ntiles; t2 += nth) {
                            const int *rw = trows.data() + (size_t)t2 * cap;
                            for (int k = 0; k < tnrow[t2]; ++k) loc[rw[k]] = k;
                            std::fill(W.begin(), W.end(), 0.0);
                            for (int qq = 0; qq < 4; ++qq) {
                                int cc = tcell[t2 * 4 + qq];
                                if (cc < 0 || !gok[cc]) continue;
                                for (int f = 0; f < 4; ++f) {
                                    int m = fp32_ ? 4 * qq + f : 4 * f + qq;
                                    const double *r = rec_h.data() + (size_t)cc * kS3D + f * 10 + 4;
                                    double sum = 0;
                                    for (int k = p.nbr_ptr[cc]; k < p.nbr_ptr[cc + 1]; ++k) {
                                        double wk = (r[0] * P3[k] + r[1] * P3[(size_t)nnz + k] +
                                                     r[2] * P3[2 * (size_t)nnz + k]) * invLx;
                                        W[(size_t)m * cap + loc[p.nbr_idx[k]]] += wk;
                                        sum += wk;
                                    }
                                    W[(size_t)m * cap + loc[cc]] -= sum;
                                }
                            }
                            for (int k = 0; k < tnrow[t2]; ++k) loc[rw[k]] = -1;
                            size_t o = (size_t)t2 * per;
                            for (int st = 0; st < kGmKS; ++st)
                                for (int l = 0; l < 64; ++l) {
                                    double v = W[(size_t)(l & 15) * cap + 4 * st + (l >> 4)];
                                    if (fp32_) tA32[o + st * 64 + l] = (float)v;
                                    else tA64[o + st * 64 + l] = v;
                                }
                        }
                    });
                for (auto &x : th) x.join();
                if (!upload(d_gm_rows, trows) || !upload(d_gm_cell, tcell)) return false;
                size_t bytes = fp32_ ? tA32.size() * sizeof(float) : tA64.size() * sizeof(double);
                const void *src = fp32_ ? (const void *)tA32.data() : (const void *)tA64.data();
                if (!alloc(d_gm_A, bytes) || !check(hipMemcpy(d_gm_A.p, src, bytes, hipMemcpyHostToDevice), "H2D"))
                    return false;
            }
        }
    }
    mark("matrix-core tiles");
    // ---- moments and the residual's face coefficients
    std::vector<std::vector<double>> FM(nf);  // per unique face, from its first side: 3 x nv
    for (int f = 0; f < nf; ++f) {
        int i0 = p.fface_cell[(size_t)f * 2], j0 = p.fface_face[(size_t)f * 2];
        face_lsq(p, (size_t)i0 * 6 + j0, invLx, FM[f]);
    }
    {
        std::vector<int> fi(ns / 6 * 4 * 4, -1);
        std::vector<double> fd(ns / 6 * 4 * kFin, 0.0);
        const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {0, 1, 2, 1, 2, 2};
        for (int c = 0; c < nc; ++c)
            for (int f = 0; f < 4; ++f) {
                size_t s = (size_t)c * 6 + f;
                int *q = &fi[((size_t)c * 4 + f) * 4];
                double *k = &fd[((size_t)c * 4 + f) * kFin];
                int fb = p.face_bound[s], fidx = p.face_index[s];
                size_t s0;        // the slot whose normal and vertex order define the face gradient
                double sign, ax, ay, az, At;
                if (fb == -1) {
                    int i0 = p.fface_cell[(size_t)fidx * 2], j0 = p.fface_face[(size_t)fidx * 2];
                    int i1 = p.fface_cell[(size_t)fidx * 2 + 1];
                    s0 = (size_t)i0 * 6 + j0;
                    sign = c == i0 ? 1.0 : -1.0;
                    q[0] = c == i0 ? i1 : i0;
                    ax = p.cx[i1] - p.cx[i0]; ay = p.cy[i1] - p.cy[i0]; az = p.cz[i1] - p.cz[i0];
                    At = p.face_area[s0];  // the reference's second side takes the first side's area here
                } else if (p.bound_type[fb] != 3) {
                    int ibc = 0;
                    for (int kk = 0; kk < 2; ++kk)
                        if (p.bound_cell[fb * 2 + kk] == c) ibc = kk;
                    s0 = s;
                    sign = 1.0;
                    q[0] = nc + fb * 2 + ibc;
                    ax = p.face_cx[s] - p.cx[c]; ay = p.face_cy[s] - p.cy[c]; az = p.face_cz[s] - p.cz[c];
                    At = p.face_area[s];
                } else {
                    continue;  // specular: no non-Fourier term (StaticBTESynthetic.cpp:1911)
                }
                const std::vector<double> &F = FM[fidx];
                int nv = p.face_nvert[s0];
                if (nv != 3 || (int)F.size() != 9) { err_ = "3D faces must be triangles"; return false; }
                for (int kk = 0; kk < 3; ++kk) q[1 + kk] = p.face_vert[s0 * 4 + kk];
                double n0[3] = {p.face_nx[s0], p.face_ny[s0], p.face_nz[s0]};
                double n[3] = {p.face_nx[s], p.face_ny[s], p.face_nz[s]};
                double Ak = p.face_area[s];
                double ga = (ax * n0[0] + ay * n0[1] + az * n0[2]) / (ax * ax + ay * ay + az * az);
                // F columns: F[i * 3 + m] is row i (x, y, z), column m (m = 1, 2 the vertex differences)
                double nn0 = n[0] * n0[0] + n[1] * n0[1] + n[2] * n0[2];
                k[0] = Ak * sign * ga * nn0;
                for (int j = 0; j < 6; ++j) {
                    int a = pa[j], b = pb[j];
                    double cp = a == b ? n[a] * n0[a] : n[b] * n0[a] + n[a] * n0[b];
                    k[1 + j] = -At * sign * ga * cp;
                }
                for (int mcol = 1; mcol <= 2; ++mcol) {
                    double fx = F[0 * 3 + mcol], fy = F[1 * 3 + mcol], fz = F[2 * 3 + mcol];
                    double fv[3] = {fx, fy, fz};
                    double *kv = k + 7 * mcol;
                    kv[0] = Ak * (n[0] * fx + n[1] * fy + n[2] * fz) * invLx;
                    for (int j = 0; j < 6; ++j) {
                        int a = pa[j], b = pb[j];
                        double cp = a == b ? n[a] * fv[a] : n[b] * fv[a] + n[a] * fv[b];
                        kv[1 + j] = -At * cp * invLx;
                    }
                }
            }
        if (!upload(d_f3i, fi) || !upload(d_f3d, fd)) return false;
        if (!alloc(d_Y, (size_t)(nc + ne) * kY) || !alloc(d_YV, (size_t)nn * kYV)) return false;
        if (!check(hipMemsetAsync(d_Y.p, 0, d_Y.n * sizeof(double), s_), "hipMemsetAsync")) return false;
    }
    // ---- synthetic (set 0) and Fourier (set 1) right-hand sides
    {
        double kap = p.kappa_bulk[0];
        for (int set = 0; set < 2; ++set) {
            std::vector<int> ri((size_t)nc * 16, -1);
            std::vector<double> rd((size_t)nc * 12, 0.0);
            std::vector<double> Pf;
            for (int c = 0; c < nc; ++c)
                for (int f = 0; f < 4; ++f) {
                    size_t s = (size_t)c * 6 + f;
                    int *q = &ri[((size_t)c * 4 + f) * 4];
                    double *k = &rd[((size_t)c * 4 + f) * 3];
                    int fb = p.face_bound[s];
                    double ax, ay, az, kA;
                    if (fb == -1) {
                        int n = p.face_neighbor[s];
                        ax = p.cx[n] - p.cx[c]; ay = p.cy[n] - p.cy[c]; az = p.cz[n] - p.cz[c];
                        kA = (set == 0) ? kap * p.face_area[s] : 2 * (kap * kap) / (kap + kap) * p.face_area[s];
                    } else {
                        int ty = p.bound_type[fb];
                        bool dir = set == 0 ? ty != 3
                                            : (ty == 1 || (ty < 0 && (p.bound_temp[fb] != 0 ||
                                                                       p.bound_temp[p.bound_connect[fb]] != 0)));
                        if (!dir) continue;
                        int ibc = 0;
                        for (int kk = 0; kk < 2; ++kk)
                            if (p.bound_cell[fb * 2 + kk] == c) ibc = kk;
                        ax = p.face_cx[s] - p.cx[c]; ay = p.face_cy[s] - p.cy[c]; az = p.face_cz[s] - p.cz[c];
                        double tmp = kap * p.face_area[s];
                        tmp = tmp / (p.face_nx[s] * ax + p.face_ny[s] * ay + p.face_nz[s] * az);
                        k[0] = tmp;
                        q[0] = set == 0 ? fb * 2 + ibc : fb * 2;
                        kA = kap * p.face_area[s];
                    }
                    double dn = ax * p.face_nx[s] + ay * p.face_ny[s] + az * p.face_nz[s];
                    double a2[3] = {p.face_nx[s] - ax / dn, p.face_ny[s] - ay / dn, p.face_nz[s] - az / dn};
                    // tangential gradient: the synthetic solve takes the face's
                    // first side (FaceMatrix), the Fourier solve the cell's own
                    // face (Get_gradient_3D)
                    size_t sv;
                    if (set == 0) {
                        int fidx = p.face_index[s];
                        sv = (size_t)p.fface_cell[(size_t)fidx * 2] * 6 + p.fface_face[(size_t)fidx * 2];
                        Pf = FM[fidx];
                    } else {
                        sv = s;
                        face_lsq(p, s, invLx, Pf);
                    }
                    for (int kk = 0; kk < 3; ++kk) q[1 + kk] = p.face_vert[sv * 4 + kk];
                    for (int mcol = 1; mcol <= 2; ++mcol)
                        k[mcol] = kA * (a2[0] * Pf[0 * 3 + mcol] + a2[1] * Pf[1 * 3 + mcol] + a2[2] * Pf[2 * 3 + mcol]) * invLx;
                }
            if (!upload(d_r3i[set], ri) || !upload(d_r3d[set], rd)) return false;
        }
    }
    (void)be_slot;
    mark("face and right-hand-side tables");
    return true;
}

// Tables of the vertex-sum gradient (k_vsum3 + k_grad3v; the algebra is at
// k_vsum3).  Returns false, with err_ empty, when the mesh does not take it:
// a cell that is not a tetrahedron, or a stencil that is not exactly the
// cells around the cell's vertices (cells of another material).
bool Solver::build_gv(double invLx) {
    const Problem &p = P;
    const int nth = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
    auto par = [&](int n, auto fn) {
        std::vector<std::thread> th;
        for (int w = 0; w < nth; ++w)
            th.emplace_back([&, w]() {
                for (int i = w; i < n; i += nth) fn(i);
            });
        for (auto &x : th) x.join();
    };
    for (int c = 0; c < nc; ++c)
        if (p.cell_nvert[c] != 4) return false;
    // cells around each vertex, in increasing cell order
    std::vector<int> cnt(nn + 1, 0);
    for (int c = 0; c < nc; ++c)
        for (int i = 0; i < 4; ++i) {
            int v = p.cell_vert[(size_t)c * 8 + i];
            if (v < 0 || v >= nn) { err_ = "internal error: tetrahedron vertex out of range"; return false; }
            ++cnt[v + 1];
        }
    for (int v = 0; v < nn; ++v) cnt[v + 1] += cnt[v];
    std::vector<int> v2c(cnt[nn]);
    {
        std::vector<int> fill(cnt.begin(), cnt.end() - 1);
        for (int c = 0; c < nc; ++c)
            for (int i = 0; i < 4; ++i) v2c[fill[p.cell_vert[(size_t)c * 8 + i]]++] = c;
    }
    // the used vertices, numbered in the order of their first cell
    std::vector<int> order, newid(nn, -1);
    for (int v = 0; v < nn; ++v)
        if (cnt[v + 1] > cnt[v]) order.push_back(v);
    std::sort(order.begin(), order.end(), [&](int x, int y) {
        return v2c[cnt[x]] != v2c[cnt[y]] ? v2c[cnt[x]] < v2c[cnt[y]] : x < y;
    });
    int nvu = (int)order.size();
    for (int i = 0; i < nvu; ++i) newid[order[i]] = i;
    // r_v: the cell of C(v) that shares an edge with the most cells of C(v)
    // (those then need no own entry for it)
    std::vector<int> rv(nn, -1);
    auto shares_edge = [&](int x, int y) {
        int n = 0;
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) n += p.cell_vert[(size_t)x * 8 + i] == p.cell_vert[(size_t)y * 8 + j];
        return n >= 2;
    };
    par(nvu, [&](int i) {
        int v = order[i], best = -1;
        for (int a = cnt[v]; a < cnt[v + 1]; ++a) {
            int s = 0;
            for (int b = cnt[v]; b < cnt[v + 1]; ++b) s += shares_edge(v2c[a], v2c[b]);
            if (s > best) { best = s; rv[v] = v2c[a]; }
        }
    });
    // per-vertex entries: (cell, (x_k - x_v) / L_x), padded with r_v at 0
    std::vector<int> vptr(nvu + 1, 0), vidx, vref(nvu);
    std::vector<double> vw;
    for (int i = 0; i < nvu; ++i) {
        int v = order[i];
        vref[i] = rv[v];
        for (int a = cnt[v]; a < cnt[v + 1]; ++a) {
            int k = v2c[a];
            vidx.push_back(k);
            vw.push_back(invLx * (p.cx[k] - p.node_x[v]));
            vw.push_back(invLx * (p.cy[k] - p.node_y[v]));
            vw.push_back(invLx * (p.cz[k] - p.node_z[v]));
        }
        while ((vidx.size() - vptr[i]) % kG3Pad) {
            vidx.push_back(rv[v]);
            vw.insert(vw.end(), {0.0, 0.0, 0.0});
        }
        vptr[i + 1] = (int)vidx.size();
    }
    // per-cell entries and records, built per cell on host threads
    struct CellOut {
        std::vector<int> idx;
        std::vector<double> w;  // 3 per entry
        double a[9], b[12];
        int v[4];
        bool ok = true;
    };
    std::vector<CellOut> out(nc);
    par(nc, [&](int c) {
        CellOut &o = out[c];
        std::vector<int> L;
        for (int i = 0; i < 4; ++i) {
            int v = p.cell_vert[(size_t)c * 8 + i];
            L.insert(L.end(), v2c.begin() + cnt[v], v2c.begin() + cnt[v + 1]);
        }
        std::sort(L.begin(), L.end());
        // the distinct cells other than c must be exactly the stencil
        std::vector<int> nb(p.nbr_idx.begin() + p.nbr_ptr[c], p.nbr_idx.begin() + p.nbr_ptr[c + 1]);
        std::sort(nb.begin(), nb.end());
        std::vector<std::pair<int, int>> mult;  // (cell, shared vertices)
        for (size_t a = 0; a < L.size();) {
            size_t z = a;
            while (z < L.size() && L[z] == L[a]) ++z;
            if (L[a] != c) mult.push_back({L[a], (int)(z - a)});
            a = z;
        }
        if (mult.size() != nb.size()) { o.ok = false; return; }
        for (size_t k = 0; k < nb.size(); ++k)
            if (mult[k].first != nb[k]) { o.ok = false; return; }
        int m = (int)nb.size();
        double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (m > 3) {
            std::vector<double> J((size_t)m * 3);
            for (int k = 0; k < m; ++k) {
                int n = p.nbr_idx[p.nbr_ptr[c] + k];
                J[k * 3 + 0] = invLx * (p.cx[n] - p.cx[c]);
                J[k * 3 + 1] = invLx * (p.cy[n] - p.cy[c]);
                J[k * 3 + 2] = invLx * (p.cz[n] - p.cz[c]);
            }
            jtj_inv(J, m, M);
        }
        auto jv = [&](int k, double *d) {
            d[0] = invLx * (p.cx[k] - p.cx[c]); d[1] = invLx * (p.cy[k] - p.cy[c]); d[2] = invLx * (p.cz[k] - p.cz[c]);
        };
        // c_k: -(m_ik - 1) j_k for the cells sharing an edge, + Q_v at r_v
        std::vector<std::pair<int, std::array<double, 3>>> cw;
        auto add = [&](int k, const double *d, double f) {
            for (auto &x : cw)
                if (x.first == k) { for (int i = 0; i < 3; ++i) x.second[i] += f * d[i]; return; }
            cw.push_back({k, {f * d[0], f * d[1], f * d[2]}});
        };
        for (auto &km : mult)
            if (km.second >= 2) {
                double d[3];
                jv(km.first, d);
                add(km.first, d, -(double)(km.second - 1));
            }
        for (int i = 0; i < 4; ++i) {
            int v = p.cell_vert[(size_t)c * 8 + i];
            double Q[3] = {0, 0, 0}, d[3];
            for (int a = cnt[v]; a < cnt[v + 1]; ++a) {
                jv(v2c[a], d);
                for (int j = 0; j < 3; ++j) Q[j] += d[j];
            }
            if (rv[v] != c) add(rv[v], Q, 1.0);
            o.v[i] = newid[v];
            double dv[3] = {invLx * (p.node_x[v] - p.cx[c]), invLx * (p.node_y[v] - p.cy[c]),
                            invLx * (p.node_z[v] - p.cz[c])};
            for (int r = 0; r < 3; ++r)
                o.b[3 * i + r] = invLx * (M[r][0] * dv[0] + M[r][1] * dv[1] + M[r][2] * dv[2]);
        }
        for (int r = 0; r < 3; ++r)
            for (int j = 0; j < 3; ++j) o.a[3 * r + j] = invLx * M[r][j];
        std::sort(cw.begin(), cw.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        for (auto &x : cw) {
            o.idx.push_back(x.first);
            for (int r = 0; r < 3; ++r)
                o.w.push_back(invLx * (M[r][0] * x.second[0] + M[r][1] * x.second[1] + M[r][2] * x.second[2]));
        }
        while (o.idx.size() % kG3Pad) {
            o.idx.push_back(c);
            o.w.insert(o.w.end(), {0.0, 0.0, 0.0});
        }
    });
    for (int c = 0; c < nc; ++c)
        if (!out[c].ok) return false;
    std::vector<int> gptr(nc + 1, 0), gidx;
    std::vector<double> gw;
    for (int c = 0; c < nc; ++c) {
        gidx.insert(gidx.end(), out[c].idx.begin(), out[c].idx.end());
        gw.insert(gw.end(), out[c].w.begin(), out[c].w.end());
        gptr[c + 1] = (int)gidx.size();
    }
    for (int k : gidx)
        if (k < 0 || k >= nc) { err_ = "internal error: vertex-gradient entry"; return false; }
    for (int k : vidx)
        if (k < 0 || k >= nc) { err_ = "internal error: vertex-sum entry"; return false; }
    // upload in the storage type
    auto put = [&](DBuf<char> &dst, auto tag, const std::vector<int> &idx, const std::vector<double> &w) -> bool {
        using ST = decltype(tag);
        std::vector<G3Ent<ST>> h(idx.size());
        for (size_t k = 0; k < idx.size(); ++k) {
            h[k] = G3Ent<ST>{};
            h[k].loc = idx[k];
            h[k].px = (ST)w[3 * k]; h[k].py = (ST)w[3 * k + 1]; h[k].pz = (ST)w[3 * k + 2];
        }
        return alloc(dst, h.size() * sizeof(h[0])) &&
               check(hipMemcpy(dst.p, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice), "H2D");
    };
    auto put_rec = [&](auto tag) -> bool {
        using ST = decltype(tag);
        std::vector<G3VRec<ST>> h(nc);
        for (int c = 0; c < nc; ++c) {
            h[c] = G3VRec<ST>{};
            for (int i = 0; i < 4; ++i) h[c].v[i] = out[c].v[i];
            for (int i = 0; i < 9; ++i) h[c].a[i] = (ST)out[c].a[i];
            for (int i = 0; i < 12; ++i) h[c].b[i] = (ST)out[c].b[i];
        }
        return alloc(d_gv_rec, h.size() * sizeof(h[0])) &&
               check(hipMemcpy(d_gv_rec.p, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice), "H2D");
    };
    bool ok = fp32_ ? put(d_gv_vent, 0.0f, vidx, vw) && put(d_gv_ent, 0.0f, gidx, gw) && put_rec(0.0f)
                    : put(d_gv_vent, 0.0, vidx, vw) && put(d_gv_ent, 0.0, gidx, gw) && put_rec(0.0);
    if (!ok || !upload(d_gv_vptr, vptr) || !upload(d_gv_vidx, vidx) || !upload(d_gv_vref, vref) ||
        !upload(d_gv_ptr, gptr) || !upload(d_gv_idx, gidx))
        return false;
    gv_nv = nvu;
    size_t sums = (size_t)nvu * nD * nB * 4 * (fp32_ ? sizeof(float) : sizeof(double));
    return alloc(d_gv_sums, sums);
}

void Solver::launch_rhs3(bool macro, const double *bval, double *rhs) {
    int set = macro ? 0 : 1;
    hipLaunchKernelGGL(k_rhs3, dim3(blocks_for(nc, kBlock)), dim3(kBlock), 0, s_, macro ? d_RM.p : nullptr, d_Q.p,
                       d_vol.p, bval, d_TV.p, d_r3i[set].p, d_r3d[set].p, rhs, nc);
}

// k_vsum3 at its own shape (gv_gs slices per wave, gv_wpb waves per
// workgroup)
void Solver::launch_vsum3() {
    int DB = nD * nB;
    int ngroups = (DB + gv_gs * kG3W - 1) / (gv_gs * kG3W);
    int gridv = ((blocks_for((long)ngroups * gv_nv, gv_wpb) + kXcd - 1) / kXcd) * kXcd;
#define GBTE_VS(ST_, GS_, W_)                                                                                    \
    hipLaunchKernelGGL((k_vsum3<ST_, GS_, W_>), dim3(gridv), dim3(64 * W_), 0, s_,                               \
                       reinterpret_cast<const ST_ *>(d_e.p), reinterpret_cast<ST_ *>(d_gv_sums.p), d_gv_vptr.p,  \
                       d_gv_vidx.p, reinterpret_cast<const G3Ent<ST_> *>(d_gv_vent.p), d_gv_vref.p, gv_nv, DB,   \
                       ngroups)
#define GBTE_VS_W(ST_, GS_)                      \
    do {                                         \
        if (gv_wpb == 16) GBTE_VS(ST_, GS_, 16); \
        else GBTE_VS(ST_, GS_, 4);               \
    } while (0)
#define GBTE_VS_GS(ST_)                        \
    do {                                       \
        if (gv_gs == 2) GBTE_VS_W(ST_, 2);     \
        else if (gv_gs == 4) GBTE_VS_W(ST_, 4); \
        else GBTE_VS_W(ST_, 8);                \
    } while (0)
    if (fp32_) GBTE_VS_GS(float);
    else GBTE_VS_GS(double);
#undef GBTE_VS_GS
#undef GBTE_VS_W
#undef GBTE_VS
}

// The 3D transport kernels of one outer iteration (and of the set-up), on
// the solver's stream.
void Solver::launch_grad3() {
    int DB = nD * nB;
    Modes md = modes();
    if (use_gm) {
        // slice groups of 256 entries
        int nb = fp32_ && DB % 2 == 0 ? 2 : 1, spg = 256 / (16 * nb);
        int ngroups = (DB + 255) / 256;
        int grid = (int)((((long)gm_ntiles * ngroups) + kXcd - 1) / kXcd) * kXcd;
        // FP32 reads pairs of entries (8-byte loads, two products per K-step)
        // where the row length allows whole pairs
        if (fp32_ && DB % 2 == 0)
            hipLaunchKernelGGL((k_grad3m<float, 2>), dim3(grid), dim3(64), 0, s_, reinterpret_cast<const float *>(d_e.p),
                               reinterpret_cast<float *>(d_grad.p), d_gm_rows.p, d_gm_cell.p,
                               reinterpret_cast<const float *>(d_gm_A.p), d_s3d.p, md, gm_ntiles, DB, nB, spg);
        else if (fp32_)
            hipLaunchKernelGGL((k_grad3m<float, 1>), dim3(grid), dim3(64), 0, s_, reinterpret_cast<const float *>(d_e.p),
                               reinterpret_cast<float *>(d_grad.p), d_gm_rows.p, d_gm_cell.p,
                               reinterpret_cast<const float *>(d_gm_A.p), d_s3d.p, md, gm_ntiles, DB, nB, spg);
        else
            hipLaunchKernelGGL((k_grad3m<double, 1>), dim3(grid), dim3(64), 0, s_, d_e.p, d_grad.p, d_gm_rows.p,
                               d_gm_cell.p, reinterpret_cast<const double *>(d_gm_A.p), d_s3d.p, md, gm_ntiles, DB,
                               nB, spg);
        return;
    }
    if (use_gv) launch_vsum3();
    int ngroups = (DB + g3_gs * kG3W - 1) / (g3_gs * kG3W);
    int grid = ((blocks_for((long)ngroups * nc, g3_wpb) + kXcd - 1) / kXcd) * kXcd;
    // the projections from the vertex sums (launched above)
#define GBTE_G3V(ST_, GS_, W_)                                                                                   \
    do {                                                                                                         \
        hipLaunchKernelGGL((k_grad3v<ST_, GS_, W_>), dim3(grid), dim3(64 * W_), 0, s_,                           \
                           reinterpret_cast<const ST_ *>(d_e.p), reinterpret_cast<const ST_ *>(d_gv_sums.p),     \
                           reinterpret_cast<ST_ *>(d_grad.p), reinterpret_cast<const G3VRec<ST_> *>(d_gv_rec.p), \
                           d_gv_ptr.p, d_gv_idx.p, reinterpret_cast<const G3Ent<ST_> *>(d_gv_ent.p), d_s3d.p,     \
                           d_oflag.p, nc, DB, nB, ngroups);                                                      \
    } while (0)
#define GBTE_G3(ST_, GS_, W_)                                                                                    \
    do {                                                                                                         \
        if (use_gv)                                                                                              \
            GBTE_G3V(ST_, GS_, W_);                                                                              \
        else                                                                                                     \
            hipLaunchKernelGGL((k_grad3g<ST_, GS_, W_>), dim3(grid), dim3(64 * W_), 0, s_,                       \
                               reinterpret_cast<const ST_ *>(d_e.p), reinterpret_cast<ST_ *>(d_grad.p),          \
                               d_g3g_ptr.p, d_g3g_idx.p, reinterpret_cast<const G3Ent<ST_> *>(d_g3g_ent.p),      \
                               d_s3d.p, d_oflag.p, nc, DB, nB, ngroups);                                         \
    } while (0)
#define GBTE_G3_W(ST_, GS_)                         \
    do {                                            \
        if (g3_wpb == 16) GBTE_G3(ST_, GS_, 16);    \
        else GBTE_G3(ST_, GS_, 4);                  \
    } while (0)
#define GBTE_G3_GS(ST_)                        \
    do {                                       \
        if (g3_gs == 1) GBTE_G3_W(ST_, 1);     \
        else if (g3_gs == 2) GBTE_G3_W(ST_, 2); \
        else if (g3_gs == 4) GBTE_G3_W(ST_, 4); \
        else GBTE_G3_W(ST_, 8);                \
    } while (0)
    if (fp32_) GBTE_G3_GS(float);
    else GBTE_G3_GS(double);
#undef GBTE_G3_GS
#undef GBTE_G3_W
#undef GBTE_G3
#undef GBTE_G3V
}

// FP32 3D source: V bands per lane take a band count divisible by V and, for
// S per band chunk, chunks of a multiple of V bands
bool Solver::src_width_ok(int v) const {
    return v == 1 || (fp32_ && nB % v == 0 && (s_sh < 0 || (1 << s_sh) % v == 0));
}

// The sweep source: in e's layout it is written into e itself, each entry
// then read by the sweep just before the sweep overwrites it with the new e
// (the source reads no e, and nothing reads the old e after it), so S and e
// share their cache lines and pages
double *Solver::sbuf() { return s_sh == kSCell ? d_e.p : d_S.p; }

void Solver::launch_source3(const Modes &m) {
    int grid = ((blocks_for(nc, kSrc3Rows) + kXcd - 1) / kXcd) * kXcd;
    size_t lds = (size_t)(4 * nB + 3 * nD) * sizeof(double) + (size_t)kSrc3Rows * nD * sizeof(int);
    int v = fp32_ ? (src_v > 0 && src_width_ok(src_v) ? src_v : src_width_ok(2) ? 2 : 1) : 1;
    double *S = sbuf();
    if (v == 4)
        hipLaunchKernelGGL(k_source3_f32v<4>, dim3(grid), dim3(64 * kSrc3Rows), lds, s_, d_posc.p,
                           reinterpret_cast<float *>(S), reinterpret_cast<const float *>(d_grad.p),
                           reinterpret_cast<const float *>(d_eb.p), d_Told.p, d_Q.p, d_s3d.p, d_s3i.p, m, nc, nB, nD, s_sh);
    else if (v == 2)
        hipLaunchKernelGGL(k_source3_f32v<2>, dim3(grid), dim3(64 * kSrc3Rows), lds, s_, d_posc.p,
                           reinterpret_cast<float *>(S), reinterpret_cast<const float *>(d_grad.p),
                           reinterpret_cast<const float *>(d_eb.p), d_Told.p, d_Q.p, d_s3d.p, d_s3i.p, m, nc, nB, nD, s_sh);
    else if (fp32_)
        hipLaunchKernelGGL(k_source3<float>, dim3(grid), dim3(64 * kSrc3Rows), lds, s_, d_posc.p,
                           reinterpret_cast<float *>(S), reinterpret_cast<const float *>(d_grad.p),
                           reinterpret_cast<const float *>(d_eb.p), d_Told.p, d_Q.p, d_s3d.p, d_s3i.p, m, nc, nB, nD, s_sh);
    else
        hipLaunchKernelGGL(k_source3<double>, dim3(grid), dim3(64 * kSrc3Rows), lds, s_, d_posc.p, S, d_grad.p,
                           d_eb.p, d_Told.p, d_Q.p, d_s3d.p, d_s3i.p, m, nc, nB, nD, s_sh);
}

// The level-synchronous chunked sweep with bc bands per workgroup and lv_np
// positions per thread in flight (the tuner's candidates).
void Solver::launch_sweep_lv(const Modes &m, int bc) {
    int grid = ((nD * (nB / bc) + kXcd - 1) / kXcd) * kXcd;
    size_t lds = (size_t)(nlev + 1) * sizeof(int);
#define GBTE_LV1(ST_, BC_, NP_, TPB_, V_, CM_)                                                                       \
    hipLaunchKernelGGL((k_sweep_lv<ST_, BC_, NP_, TPB_, V_, CM_>), dim3(grid), dim3(TPB_), lds, s_, d_lv_rec.p,      \
                       d_levptr.p, d_dlev.p, reinterpret_cast<const ST_ *>(sbuf()), reinterpret_cast<ST_ *>(d_e.p), m, \
                       nc, nB, nD, s_sh < 0 ? 1 : 0)
#define GBTE_LV(ST_, BC_, NP_, TPB_, V_)                           \
    do {                                                           \
        if (s_sh == kSCell) GBTE_LV1(ST_, BC_, NP_, TPB_, V_, true); \
        else GBTE_LV1(ST_, BC_, NP_, TPB_, V_, false);             \
    } while (0)
#define GBTE_LV_NP(ST_, BC_)                            \
    do {                                                \
        if (lv_np == 4) GBTE_LV(ST_, BC_, 4, 256, 1);   \
        else if (lv_np == 2) GBTE_LV(ST_, BC_, 2, 256, 1); \
        else GBTE_LV(ST_, BC_, 1, 256, 1);              \
    } while (0)
#define GBTE_LV_NP2(ST_, BC_)                           \
    do {                                                \
        if (lv_np == 2) GBTE_LV(ST_, BC_, 2, 256, 1);   \
        else GBTE_LV(ST_, BC_, 1, 256, 1);              \
    } while (0)
    // one 16-byte vector per lane: V = BC * word / 16 lanes per position
#define GBTE_LVV(ST_, BC_)                                                     \
    do {                                                                       \
        constexpr int V_ = BC_ * (int)sizeof(ST_) / 16;                        \
        if (lv_tpb == 1024) {                                                  \
            GBTE_LV(ST_, BC_, 2, 1024, V_);                                    \
        } else if (lv_tpb == 512) {                                            \
            if (lv_np == 4) GBTE_LV(ST_, BC_, 4, 512, V_);                     \
            else GBTE_LV(ST_, BC_, 2, 512, V_);                                \
        } else {                                                               \
            if (lv_np == 4) GBTE_LV(ST_, BC_, 4, 256, V_);                     \
            else GBTE_LV(ST_, BC_, 2, 256, V_);                                \
        }                                                                      \
    } while (0)
    if (lv_v == 0) {
        // chunks of up to 16 bands (FP32) or 16 / 8 bands (FP64), lanes per
        // position set per chunk; the grid counts the partial last chunk
        grid = ((nD * ((nB + bc - 1) / bc) + kXcd - 1) / kXcd) * kXcd;
        if (fp32_) {
            if (lv_np == 4) GBTE_LV(float, 16, 4, 512, 0);
            else GBTE_LV(float, 16, 2, 512, 0);
        } else if (bc == 16) {
            if (lv_np == 4) GBTE_LV(double, 16, 4, 512, 0);
            else GBTE_LV(double, 16, 2, 512, 0);
        } else {
            if (lv_np == 4) GBTE_LV(double, 8, 4, 512, 0);
            else GBTE_LV(double, 8, 2, 512, 0);
        }
    } else if (lv_v > 1) {
        if (fp32_) GBTE_LVV(float, 8);
        else if (bc == 4) GBTE_LVV(double, 4);
        else GBTE_LVV(double, 8);
    } else if (fp32_) {
        if (bc == 4) GBTE_LV_NP(float, 4);
        else if (bc == 8) GBTE_LV_NP(float, 8);
        else GBTE_LV_NP2(float, 20);
    } else {
        if (bc == 2) GBTE_LV_NP(double, 2);
        else if (bc == 4) GBTE_LV_NP(double, 4);
        else if (bc == 8) GBTE_LV_NP(double, 8);
        else if (bc == 10) GBTE_LV_NP2(double, 10);
        else GBTE_LV_NP2(double, 20);
    }
#undef GBTE_LVV
#undef GBTE_LV_NP2
#undef GBTE_LV_NP
#undef GBTE_LV
#undef GBTE_LV1
}

void Solver::launch_bound_out3(const Modes &m) {
    int n_ent = (int)h_ent_list.size();
    if (n_ent == 0) return;
    dim3 grid(blocks_for((long)n_ent * 64, 256));
    size_t bo_lds = (size_t)(nB + 3 * nD) * sizeof(double);
    if (fp32_)
        hipLaunchKernelGGL(k_bound_out3<float>, grid, dim3(256), bo_lds, s_, reinterpret_cast<const float *>(d_e.p),
                           reinterpret_cast<const float *>(d_grad.p), reinterpret_cast<float *>(d_eb.p), m,
                           d_ent_list.p, d_be_cell.p, d_be_slot.p, d_snx.p, d_sny.p, d_snz.p, n_ent, nB, nD);
    else
        hipLaunchKernelGGL(k_bound_out3<double>, grid, dim3(256), bo_lds, s_, d_e.p, d_grad.p, d_eb.p, m, d_ent_list.p,
                           d_be_cell.p, d_be_slot.p, d_snx.p, d_sny.p, d_snz.p, n_ent, nB, nD);
}

void Solver::launch_moments3(const Modes &m) {
    int n_ent = (int)h_ent_list.size();
    int grid = ((blocks_for(nc + n_ent, kMomRows) + kXcd - 1) / kXcd) * kXcd;
    if (fp32_)
        hipLaunchKernelGGL(k_mom3<float>, dim3(grid), dim3(64 * kMomRows), 0, s_, reinterpret_cast<const float *>(d_e.p),
                           reinterpret_cast<const float *>(d_eb.p), d_ent_list.p, d_dirf.p, m, d_Y.p, nc, n_ent, nB,
                           nD);
    else
        hipLaunchKernelGGL(k_mom3<double>, dim3(grid), dim3(64 * kMomRows), 0, s_, d_e.p, d_eb.p, d_ent_list.p,
                           d_dirf.p, m, d_Y.p, nc, n_ent, nB, nD);
    hipLaunchKernelGGL(k_vert3, dim3(blocks_for(nn, kBlock)), dim3(kBlock), 0, s_, d_Y.p, d_vptr.p, d_vsrc.p,
                       d_vwt.p, d_YV.p, nn, nc);
    hipLaunchKernelGGL(k_fin3, dim3(blocks_for(nc, kBlock)), dim3(kBlock), 0, s_, d_Y.p, d_YV.p, d_f3i.p, d_f3d.p,
                       m, d_T.p, d_qx.p, d_qy.p, d_qz.p, d_RM.p, nc);
}

// Project autotuner for tetrahedral meshes: times each candidate shape of
// the gradient kernels (k_grad3g slices per wave and waves per workgroup,
// k_grad3m), of the sweeps (wave-specialised, level-synchronous, chunked at
// each band chunk and positions per thread) and the layout of S between the
// source and the chunked sweep, on this mesh, and keeps the fastest.  The
// candidates, timings and choices are kept in tune_log_ (gbte_tuning_log).
bool Solver::tune3d() {
    if (!check(hipMemsetAsync(d_e.p, 0, d_e.n * sizeof(double), s_), "hipMemsetAsync")) return false;
    hipEvent_t a, b;
    if (!check(hipEventCreate(&a), "hipEventCreate") || !check(hipEventCreate(&b), "hipEventCreate")) return false;
    char line[160];
    // one candidate: a warm launch and a timed one, then two more timed
    // launches only when the first is within 15% of the best so far (a
    // clearly slower candidate costs two launches instead of four)
    auto time_it = [&](auto launch, double best, double &per) -> bool {
        launch();
        (void)hipEventRecord(a, s_);
        launch();
        (void)hipEventRecord(b, s_);
        if (!check(hipEventSynchronize(b), "tune sync")) return false;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        per = ms;
        if (per <= 1.15 * best) {
            (void)hipEventRecord(a, s_);
            launch();
            launch();
            (void)hipEventRecord(b, s_);
            if (!check(hipEventSynchronize(b), "tune sync")) return false;
            (void)hipEventElapsedTime(&ms, a, b);
            per = (per + ms) / 3.0;
        }
        return true;
    };
    {
        // the gradient: 64-entry slices per wave (one stencil-table batch
        // serves them all; more slices, more loads in flight per wave)
        double best = 1e300;
        int pick = 4, pickw = 4, kind = 0;  // kind 0: k_grad3g, 1: k_vsum3 + k_grad3v, 2: k_grad3m
        int want = opt.gradient_kernel;
        if (want < GBTE_GRAD_AUTO || want > GBTE_GRAD_VERTEX) { err_ = "unknown gradient_kernel"; return false; }
        bool want_g = want == GBTE_GRAD_AUTO || want == GBTE_GRAD_GATHER;
        bool want_v = want == GBTE_GRAD_AUTO || want == GBTE_GRAD_VERTEX;
        bool want_m = want == GBTE_GRAD_AUTO || want == GBTE_GRAD_MATRIX;
        if (want == GBTE_GRAD_MATRIX && !gm_ok) {
            err_ = "the matrix-core gradient does not take this mesh's stencils";
            return false;
        }
        if (want == GBTE_GRAD_VERTEX && !gv_ok) {
            err_ = "the vertex-sum gradient does not take this mesh (tetrahedra of one material only)";
            return false;
        }
        // the vertex sums read about half the rows of the direct gather, so
        // the gather shapes are timed only when the vertex sums do not take
        // the mesh or the gather is asked for
        if (want == GBTE_GRAD_AUTO && gv_ok) want_g = false;
        // the vertex sums first, at their own shape (the projections read
        // them in any slice order); then the projections' shapes, each
        // timed as the pair
        if (want_v && gv_ok) {
            double bv = 1e300;
            int pgs = 4, pw = 4;
            for (int wpb : {4, 16})
                for (int gs : {2, 4, 8}) {
                    if (gs == 8 && !fp32_) continue;
                    gv_gs = gs;
                    gv_wpb = wpb;
                    double per = 0;
                    if (!time_it([&]() { launch_vsum3(); }, bv, per)) return false;
                    snprintf(line, sizeof line, "k_vsum3 %d slices per wave, %d waves per workgroup: %.3f ms per launch\n",
                             gs, wpb, per);
                    tune_log_ += line;
                    if (per < bv) { bv = per; pgs = gs; pw = wpb; }
                }
            gv_gs = pgs;
            gv_wpb = pw;
        }
        for (int v : {1, 0}) {
            if (v == 0 ? !want_g : !(want_v && gv_ok)) continue;
            use_gv = v == 1;
            for (int wpb : {4, 16})
                for (int gs : {1, 2, 4, 8}) {
                    if (gs == 8 && !fp32_) continue;  // FP64 is bound by the gather at 2 slices
                    if (v == 1 && gs == 1) continue;
                    g3_gs = gs;
                    g3_wpb = wpb;
                    double per = 0;
                    if (!time_it([&]() { launch_grad3(); }, best, per)) return false;
                    snprintf(line, sizeof line, "%s %d slices per wave, %d waves per workgroup: %.3f ms per launch\n",
                             v ? "k_vsum3 + k_grad3v" : "k_grad3g", gs, wpb, per);
                    tune_log_ += line;
                    if (per < best) { best = per; pick = gs; pickw = wpb; kind = v; }
                }
        }
        use_gv = false;
        if (gm_ok && want_m) {
            use_gm = true;
            double per = 0;
            if (!time_it([&]() { launch_grad3(); }, best, per)) return false;
            snprintf(line, sizeof line, "k_grad3m (matrix cores, %d tiles): %.3f ms per launch\n", gm_ntiles, per);
            tune_log_ += line;
            use_gm = per < best;
            if (use_gm) { best = per; kind = 2; }
        }
        g3_gs = pick;
        g3_wpb = pickw;
        use_gv = kind == 1;
        if (use_gm)
            snprintf(line, sizeof line, "gradient chosen: k_grad3m (matrix cores)\n");
        else
            snprintf(line, sizeof line,
                     use_gv ? "gradient chosen: k_vsum3 + k_grad3v, %d slices per wave, %d waves per workgroup "
                              "(k_vsum3 %d and %d)\n"
                            : "gradient chosen: k_grad3g, %d slices per wave, %d waves per workgroup\n",
                     pick, pickw, gv_gs, gv_wpb);
        tune_log_ += line;
    }
    // sweeps: the wave-specialised kernel (when its schedule was built), the
    // one-workgroup-per-direction level-synchronous kernel (FP64) and the
    // chunked level-synchronous kernel at each band chunk and positions per
    // thread in flight that the storage type allows
    {
        if (!check(hipMemsetAsync(d_S.p, 0, d_S.n * sizeof(double), s_), "hipMemsetAsync")) return false;
        if (!check(hipMemsetAsync(d_e.p, 0, d_e.n * sizeof(double), s_), "hipMemsetAsync")) return false;
        Modes m = modes();
        struct Cand { int kind, bc, np, v = 1, tpb = 256; };  // kind 0: wave-specialised, 1: level-synchronous, 2: chunked
        std::vector<Cand> cands;
        bool pc_built = use_pc;
        int want = opt.sweep_kernel;  // GBTE_SWEEP_AUTO: every family
        if (want < GBTE_SWEEP_AUTO || want > GBTE_SWEEP_LANES) { err_ = "unknown sweep_kernel"; return false; }
        if (pc_built && (want == GBTE_SWEEP_AUTO || want == GBTE_SWEEP_WAVESPEC)) cands.push_back({0, 0, 0});
        if (!fp32_ && (want == GBTE_SWEEP_AUTO || want == GBTE_SWEEP_LEVEL)) cands.push_back({1, 0, 0});
        int wsz = fp32_ ? 4 : 8;
        // band chunks of whole 16-byte vectors (the non-power-of-two ones with
        // S stored per position), up to 160 bytes of each operand per thread
        for (int bc : {2, 4, 8, 10, 20})
            if (d_lv_rec.p && nB % bc == 0 && (bc * wsz) % 16 == 0 && (nB * wsz) % 16 == 0 &&
                (size_t)(nlev + 1) * sizeof(int) <= 48 * 1024 &&
                (want == GBTE_SWEEP_AUTO || want == GBTE_SWEEP_CHUNKED))
                for (int np : {1, 2, 4})
                    if (bc * wsz * np <= 160 && (np <= 2 || (bc & (bc - 1)) == 0)) cands.push_back({2, bc, np});
        // several lanes per position, one 16-byte vector each (8 bands FP32;
        // 4 or 8 bands FP64), 256 or 512 threads per workgroup
        for (int bc : {4, 8})
            if (d_lv_rec.p && nB % bc == 0 && (fp32_ ? bc == 8 : true) && (nB * wsz) % 16 == 0 &&
                (size_t)(nlev + 1) * sizeof(int) <= 48 * 1024 &&
                (want == GBTE_SWEEP_AUTO || want == GBTE_SWEEP_LANES))
                for (int tpb : {256, 512, 1024})
                    for (int np : {2, 4})
                        if (tpb < 1024 || np == 2) cands.push_back({2, bc, np, bc * wsz / 16, tpb});
        // chunks of up to 16 bands (FP32 and FP64) or 8 (FP64) where the bands
        // split into such chunks with 1, 2, 4 or 8 lanes per position, S per
        // position (v = 0 marks the per-chunk lane count)
        for (int bc : {16, 8}) {
            int bl = 16 / wsz, last = nB % bc == 0 ? bc : nB % bc;
            bool pow2 = last % bl == 0 && ((last / bl) & (last / bl - 1)) == 0;
            if ((fp32_ && bc != 16) || nB % bc == 0 || !pow2 || !d_lv_rec.p ||
                (size_t)(nlev + 1) * sizeof(int) > 48 * 1024 || !(want == GBTE_SWEEP_AUTO || want == GBTE_SWEEP_LANES))
                continue;
            for (int np : {2, 4}) cands.push_back({2, bc, np, 0, 512});
        }
        // the lane forms first: usually the fastest, they set the bar early so
        // that slower candidates get only their warm and one timed launch
        std::stable_partition(cands.begin(), cands.end(), [](const Cand &c) { return c.kind == 2 && c.v != 1; });
        double best = 1e300;
        Cand pick{-1, 0, 0};
        auto chunk_sh = [](int bc) {  // S per band chunk where the chunk is a power of two
            for (int k = 1; k <= 3; ++k)
                if ((1 << k) == bc) return k;
            return -1;
        };
        for (const Cand &c : cands) {
            use_pc = c.kind == 0;
            lv_bc = c.kind == 2 ? c.bc : 0;
            lv_np = c.kind == 2 ? c.np : 2;
            lv_v = c.v;
            lv_tpb = c.tpb;
            s_sh = c.kind == 2 && c.v != 0 ? chunk_sh(c.bc) : -1;
            double per = 0;
            if (!time_it([&]() { launch_sweep3(m); }, best, per)) return false;
            if (c.kind == 2 && c.v == 0)
                snprintf(line, sizeof line,
                         "sweep chunked up to %d bands, one vector per lane, %d positions per lane, %d threads: %.3f ms per launch\n",
                         c.bc, c.np, c.tpb, per);
            else if (c.kind == 2 && c.v > 1)
                snprintf(line, sizeof line,
                         "sweep chunked %d bands, %d lanes per position, %d positions per lane, %d threads: %.3f ms per launch\n",
                         c.bc, c.v, c.np, c.tpb, per);
            else if (c.kind == 2)
                snprintf(line, sizeof line, "sweep chunked %d bands, %d positions per thread: %.3f ms per launch\n",
                         c.bc, c.np, per);
            else
                snprintf(line, sizeof line, "sweep %s: %.3f ms per launch\n",
                         c.kind == 0 ? "wave-specialised" : "level-synchronous", per);
            tune_log_ += line;
            if (per < best) { best = per; pick = c; }
        }
        if (pick.kind < 0) { err_ = "no sweep kernel takes this mesh at this storage precision"; return false; }
        use_pc = pick.kind == 0;
        lv_bc = pick.kind == 2 ? pick.bc : 0;
        lv_np = pick.kind == 2 ? pick.np : 2;
        lv_v = pick.v;
        lv_tpb = pick.tpb;
        s_sh = pick.kind == 2 && pick.v != 0 ? chunk_sh(pick.bc) : -1;
        if (pick.kind == 2 && pick.v == 0)
            snprintf(line, sizeof line,
                     "sweep chosen: chunked up to %d bands, one vector per lane, %d positions per lane, %d threads\n",
                     pick.bc, pick.np, pick.tpb);
        else if (pick.kind == 2 && pick.v > 1)
            snprintf(line, sizeof line,
                     "sweep chosen: chunked %d bands, %d lanes per position, %d positions per lane, %d threads\n",
                     pick.bc, pick.v, pick.np, pick.tpb);
        else if (pick.kind == 2)
            snprintf(line, sizeof line, "sweep chosen: chunked %d bands, %d positions per thread\n", pick.bc, pick.np);
        else
            snprintf(line, sizeof line, "sweep chosen: %s\n", use_pc ? "wave-specialised" : "level-synchronous");
        tune_log_ += line;
        int wl = opt.source_layout;
        if (wl < GBTE_SLAYOUT_AUTO || wl > GBTE_SLAYOUT_CELL) { err_ = "unknown source_layout"; return false; }
        if (wl != GBTE_SLAYOUT_AUTO && pick.kind != 2) {
            err_ = "source_layout applies to the chunked sweep only";
            return false;
        }
        if (wl == GBTE_SLAYOUT_CHUNK && s_sh < 0) {
            err_ = "S per band chunk needs a sweep chunk of a power-of-two band count";
            return false;
        }
        if (pick.kind == 2) {
            // the layout of S between the source kernel and the chunked sweep:
            // per band chunk (the sweep's loads contiguous), per position with
            // all bands, or per cell in e's layout (the source's stores whole
            // rows, the sweep's loads at the positions' cells); the pair is
            // timed
            int shc = s_sh;
            double tl[3] = {0, 0, 0};
            const int lay[3] = {shc, -1, kSCell};
            const char *lname[3] = {"per band chunk", "per position", "per cell (e's layout)"};
            for (int v = 0; v < 3; ++v) {
                if ((v == 0 && shc < 0) || (wl != GBTE_SLAYOUT_AUTO && wl != v + 1)) {
                    tl[v] = 1e30;  // per position is then the sweep's own layout, or another one is asked for
                    continue;
                }
                s_sh = lay[v];
                launch_source3(m);
                launch_sweep3(m);  // warm
                (void)hipEventRecord(a, s_);
                for (int r = 0; r < 3; ++r) {
                    launch_source3(m);
                    launch_sweep3(m);
                }
                (void)hipEventRecord(b, s_);
                if (!check(hipEventSynchronize(b), "tune sync")) return false;
                float ms = 0;
                (void)hipEventElapsedTime(&ms, a, b);
                tl[v] = ms / 3.0;
                snprintf(line, sizeof line, "source + sweep, S %s: %.3f ms per pair\n", lname[v], tl[v]);
                tune_log_ += line;
            }
            int lb = 0;
            for (int v = 1; v < 3; ++v)
                if (tl[v] < tl[lb]) lb = v;
            s_sh = lay[lb];
            snprintf(line, sizeof line, "S layout chosen: %s\n", lname[lb]);
            tune_log_ += line;
        }
    }
    if (fp32_) {
        // FP32 source: bands per lane (8- or 16-byte face loads)
        Modes m = modes();
        double best = 1e30;
        int pick = 0;
        for (int v : {2, 4}) {
            if (!src_width_ok(v)) continue;
            src_v = v;
            double per = 0;
            if (!time_it([&]() { launch_source3(m); }, best, per)) return false;
            snprintf(line, sizeof line, "source %d bands per lane: %.3f ms per launch\n", v, per);
            tune_log_ += line;
            if (per < best) { best = per; pick = v; }
        }
        src_v = pick;
        if (pick) {
            snprintf(line, sizeof line, "source chosen: %d bands per lane\n", pick);
            tune_log_ += line;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (opt.verbose) printf("%s", tune_log_.c_str());
    if (!check(hipGetLastError(), "tune launches")) return false;
    // S lives in e from here on (the chunked sweep is the only one that
    // takes this layout): its own buffer is freed
    if (s_sh == kSCell) d_S.release();
    return true;
}

}  // namespace gbte
