// Discrete adjoint of the steady StaticBTESynthetic fixed point (2D, FP64):
// dk_eff / d(relaxation-time scale of each band) and dk_eff / d(wall
// temperature of each isothermal boundary face).  No reference counterpart.
//
// One outer iteration of the solver (StaticBTESyntheticSolver.cpp:63-221) is
// an affine map x' = F(x) of the state x = (e, incoming boundary values,
// macro T):
//   g   = LSQ(e)                                   _get_gradient_larger
//   S   = C T + Q hr tau - sum a (r.g) - sum_in,bnd a eb_in      _get_Re
//   e'  = M^-1 S   (M the upwind matrix; e_old cancels)          sweep
//   eb_out' = e' + r.g                             _get_bound_ee
//   R   = sum_sys NF(face gradients of e', IDW(e', eb))   _get_non_fourier
//   T_rec = sum lr w e' / C
//   eb_in' = BC(eb_out', T_wall)                   _set_bound_ee_1
//   tmac = sum lr w eb' / C                        _set_macro_boundary_condition
//   T' = m synthetic passes T^k = K^-1 (R - QV - B tmac - Corr(IDW(T^{k-1})))
//        from T^0 = T_rec                          _marco_solution
// (m is the pass count of the forward's last iteration).  At the converged
// state x*, k_eff = c . e* with c = v s_axis w V / (A dT), and
// dk/dtheta = lambda . dF/dtheta with lambda = J^T lambda + c, J = dF/dx,
// solved by the same fixed-point iteration run backwards: each adjoint step
// applies J^T (the macro passes in reverse order, the transposed boundary
// conditions, the transposed non-Fourier / IDW / face-gradient chain, a sweep
// of M^T in reversed upwind order, the transposed source and LSQ gradient).
//
// The transposed non-Fourier chain factors through per-system constants:
// its contribution to the adjoint of e'(sys) is sum_j K_j(sys) W_j(cell)
// with K = ([hr>0] w tau v^2 s_a s_b for ab = xx, xy, yy; [hr>0] w lr / C),
// the recovery adds (w lr / C) lambda_Trec, and W_j are geometric transposes
// of the cell vector mu = sum_k K^-1 lambda_T^k formed on the host.
//
// Parameter terms (tau_b enters a = tau (A/V)(v.n), the tau part of R, the
// lattice ratios lr = (C/tau) / sum (C/tau) and kappa = sum C v^2 tau / dimMat):
//   sweep:  lambda_S . (e* - C T*) / tau_b    (dM/dtau e' = (S - e') / tau,
//                                              dS/dtau = (S - C T) / tau)
//   R:      mu . R_tau,b / tau_b
//   lr_k:   w / C_k (lambda_Trec . e*_k + lambda_tmac . eb*_k) + mu . R_kappa,k / lr_k
//   kappa:  mu . R_kappa / kappa - sum_k nu_k . (R* - QV) / kappa
// and d/dT_wall sums C_b over the adjoint of the wall's incoming values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "devcommon.hpp"
#include "solver.hpp"

namespace gbte {
namespace {

constexpr int kAdjRow = 256;    // threads per cell / boundary-entry row
constexpr int kAdjSweep = 256;  // threads of one reversed-sweep workgroup
constexpr int kAdjFaces = 4;    // face slots per 2D cell
constexpr int kAdjW = 5;        // geometric weights per cell / entry: xx, xy, yy, kappa, T
constexpr int kColChunk = 2048; // cells (or entries) per partial of k_adj_colsum
constexpr int kCols = 11;       // k_adj_colsum outputs per column

// Face slot of a cell as the adjoint kernels need it.
struct AdjFace {
    int kind;          // 0 none, 1 interior, 2 boundary
    int other;         // interior: neighbour cell; boundary: entry
    double av;         // A / V of this slot
    double nx, ny;     // slot normal
    double rx, ry;     // face centre - cell centre
    double avn;        // interior: A / V of the neighbour's slot
    double nnx, nny;   // interior: the neighbour's slot normal
    double rnx, rny;   // interior: the neighbour's face centre - this cell's centre
};

struct AdjSys {
    double tv, vb, cb, hr, tau;
    double kxx, kxy, kyy, kk, kt;
};

__device__ __forceinline__ AdjSys adj_sys(const Modes &m, int d, int b) {
    AdjSys s;
    s.vb = m.vg[b];
    s.tau = m.tau[b];
    s.tv = s.tau * s.vb;
    s.cb = m.cap[b];
    s.hr = m.hr[b];
    double w = m.w[d], sx = m.sx[d], sy = m.sy[d];
    double h = s.hr > 0 ? 1.0 : 0.0;
    double tvv = h * w * s.tau * s.vb * s.vb;
    s.kxx = tvv * sx * sx;
    s.kxy = tvv * sx * sy;
    s.kyy = tvv * sy * sy;
    s.kk = h * w * m.lr[b] / s.cb;
    s.kt = w * m.lr[b] / s.cb;
    return s;
}

__device__ __forceinline__ double adj_nf(const AdjSys &s, const double *W) {
    return s.kxx * W[0] + s.kxy * W[1] + s.kyy * W[2] + s.kk * W[3];
}

// Incoming boundary values: their total adjoint Lin (state adjoint plus the
// macro boundary temperatures' part), the new state adjoint (identity where
// the boundary condition keeps the old value, plus the non-Fourier chain's
// part) and the wall-temperature sums.  One lane per band.
__global__ void k_adj_bc_in(double *__restrict__ Ab, double *__restrict__ Lin, double *__restrict__ twp,
                            const double *__restrict__ Wb, Modes m, const int *__restrict__ ent_list,
                            const int *__restrict__ be_type, const double *__restrict__ be_nx,
                            const double *__restrict__ be_ny, int n_ent, int nB, int nD, int cpw) {
    int lane = lane_id(), sub = lane / nB, b = lane - sub * nB;
    int k = wave_global() * cpw + sub;
    bool act = sub < cpw && k < n_ent;
    int en = ent_list[act ? k : 0];
    size_t DB = (size_t)nD * nB;
    double tw = 0;
    if (act) {
        int ty = be_type[en];
        bool handled = ty == 1 || ty == 2 || ty == 3 || ty < 0;
        double nx = be_nx[en], ny = be_ny[en];
        const double *W = Wb + (size_t)en * kAdjW;
        for (int d = 0; d < nD; ++d) {
            AdjSys s = adj_sys(m, d, b);
            size_t i = (size_t)en * DB + (size_t)d * nB + b;
            double dot = s.vb * (m.sx[d] * nx + m.sy[d] * ny);
            if (dot < 0) {
                double lin = Ab[i] + s.kt * W[4];
                Lin[i] = lin;
                bool keep = !(s.hr != 0) || !handled;
                Ab[i] = (keep ? lin : 0.0) + adj_nf(s, W);
                if (ty == 1 && s.hr != 0) tw += s.cb * lin;
            } else {
                Lin[i] = 0;
                Ab[i] = 0;
            }
        }
    }
    tw = group_sum(tw, b, nB);
    if (act && b == 0) twp[en] = tw;
}

// Adjoint of the outgoing boundary values: the macro boundary temperatures,
// the non-Fourier chain and the transposed boundary conditions (diffuse:
// the outgoing values' weights; periodic: the partner's incoming values;
// specular: the reflected directions' incoming values).
__global__ void k_adj_bc_out(double *__restrict__ Lout, const double *__restrict__ Lin,
                             const double *__restrict__ Wb, Modes m, const int *__restrict__ ent_list,
                             const int *__restrict__ be_type, const int *__restrict__ inv_conn,
                             const int *__restrict__ be_spec, const double *__restrict__ be_nx,
                             const double *__restrict__ be_ny, int n_ent, int nB, int nD, int cpw) {
    int lane = lane_id(), sub = lane / nB, b = lane - sub * nB;
    int k = wave_global() * cpw + sub;
    bool act = sub < cpw && k < n_ent;
    if (!act) return;
    int en = ent_list[k];
    size_t DB = (size_t)nD * nB;
    int ty = be_type[en];
    double nx = be_nx[en], ny = be_ny[en];
    const double *W = Wb + (size_t)en * kAdjW;
    const double *lin = Lin + (size_t)en * DB + b;
    double vb = m.vg[b], hb = m.hr[b];
    double den = 0, sin_ = 0;
    if (ty == 2) {
        for (int q = 0; q < nD; ++q) {
            double d1 = vb * (m.sx[q] * nx + m.sy[q] * ny);
            if (d1 >= 0) den += d1 * m.w[q];
            else sin_ += lin[(size_t)q * nB];
        }
    }
    int pin = ty < 0 ? inv_conn[en] : -1;
    for (int d = 0; d < nD; ++d) {
        AdjSys s = adj_sys(m, d, b);
        size_t i = (size_t)en * DB + (size_t)d * nB + b;
        double dot = vb * (m.sx[d] * nx + m.sy[d] * ny);
        if (dot < 0) {
            Lout[i] = 0;
            continue;
        }
        double acc = s.kt * W[4] + adj_nf(s, W);
        if (hb != 0) {
            if (ty == 2) acc += dot * m.w[d] / den * sin_;
            else if (ty < 0 && pin >= 0) acc += Lin[(size_t)pin * DB + (size_t)d * nB + b];
            else if (ty == 3)
                for (int q = 0; q < nD; ++q)
                    if (be_spec[(size_t)en * nD + q] == d) acc += lin[(size_t)q * nB];
        }
        Lout[i] = acc;
    }
}

// Adjoint of e' before the sweep: the carried state adjoint, the non-Fourier
// chain, the recovery of T and the outgoing boundary values (which also start
// the adjoint of the gradient).
__global__ void __launch_bounds__(kAdjRow) k_adj_pre(double *__restrict__ Ae, double *__restrict__ LGx,
                                                     double *__restrict__ LGy, const double *__restrict__ Lout,
                                                     const double *__restrict__ Wc, const AdjFace *__restrict__ F,
                                                     Modes m, int nc, int nB, int nD) {
    int c = blockIdx.x;
    if (c >= nc) return;
    int DB = nD * nB;
    const double *W = Wc + (size_t)c * kAdjW;
    const AdjFace *fc = F + (size_t)c * kAdjFaces;
    for (int t = threadIdx.x; t < DB; t += blockDim.x) {
        int d = t / nB, b = t - d * nB;
        AdjSys s = adj_sys(m, d, b);
        size_t i = (size_t)c * DB + t;
        double le = Ae[i] + adj_nf(s, W) + s.kt * W[4];
        double gx = 0, gy = 0;
#pragma unroll
        for (int f = 0; f < kAdjFaces; ++f) {
            if (fc[f].kind != 2) continue;
            double dot = s.vb * (m.sx[d] * fc[f].nx + m.sy[d] * fc[f].ny);
            if (dot < 0) continue;
            double l = Lout[(size_t)fc[f].other * DB + t];
            le += l;
            gx += fc[f].rx * l;
            gy += fc[f].ry * l;
        }
        Ae[i] = le;
        LGx[i] = gx;
        LGy[i] = gy;
    }
}

// M^T lambda_S = lambda_e' for every system: the upwind levels in reverse;
// a cell's downwind cells lie in later levels, finished before it.
// dn_c / dn_g: per position up to kMaxIn downwind cells and their in-face
// (A/V)(s.n); rec_G the position's out-face sum.
__global__ void __launch_bounds__(kAdjSweep) k_adj_sweep(double *__restrict__ LS, const double *__restrict__ Ae,
                                                         const int *__restrict__ item_cell,
                                                         const int *__restrict__ levptr, const int *__restrict__ dlev,
                                                         const int *__restrict__ dn_c, const double *__restrict__ dn_g,
                                                         const double *__restrict__ rec_G, Modes m, int nc, int nB,
                                                         int nD, int Bc, int nchunk) {
    int d = blockIdx.x / nchunk, chunk = blockIdx.x - d * nchunk;
    const int *lp = levptr + dlev[d];
    int L = dlev[d + 1] - dlev[d] - 1;
    size_t base = (size_t)d * nc;
    size_t DB = (size_t)nD * nB;
    int b0 = chunk * Bc, nb = min(Bc, nB - b0);
    for (int l = L - 1; l >= 0; --l) {
        int a = lp[l], z = lp[l + 1];
        for (int idx = threadIdx.x; idx < (z - a) * nb; idx += blockDim.x) {
            int k = a + idx / nb, b = b0 + idx % nb;
            size_t P = base + k;
            double tv = m.tau[b] * m.vg[b];
            int c = item_cell[P];
            size_t col = (size_t)d * nB + b;
            double s = Ae[(size_t)c * DB + col];
#pragma unroll
            for (int j = 0; j < kMaxIn; ++j) {
                int cj = dn_c[P * kMaxIn + j];
                if (cj >= 0) s -= tv * dn_g[P * kMaxIn + j] * LS[(size_t)cj * DB + col];
            }
            LS[(size_t)c * DB + col] = s / (1.0 + tv * rec_G[P]);
        }
        __syncthreads();
    }
}

// Transposed source: the gradient adjoint (own out-faces and the downwind
// neighbours' in-faces), the incoming boundary values' adjoint and the
// macro T's adjoint sum_sys C lambda_S.
__global__ void __launch_bounds__(kAdjRow) k_adj_src(double *__restrict__ LGx, double *__restrict__ LGy,
                                                     double *__restrict__ Ab, double *__restrict__ ATn,
                                                     const double *__restrict__ LS, const AdjFace *__restrict__ F,
                                                     Modes m, int nc, int nB, int nD, int first_order) {
    __shared__ double red[kAdjRow];
    int c = blockIdx.x;
    int DB = nD * nB;
    const AdjFace *fc = F + (size_t)c * kAdjFaces;
    double at = 0;
    for (int t = threadIdx.x; t < DB; t += blockDim.x) {
        int d = t / nB, b = t - d * nB;
        double vb = m.vg[b], tau = m.tau[b], sx = m.sx[d], sy = m.sy[d];
        size_t i = (size_t)c * DB + t;
        double ls = LS[i];
        at += m.cap[b] * ls;
        double gx = 0, gy = 0;
#pragma unroll
        for (int f = 0; f < kAdjFaces; ++f) {
            int kd = fc[f].kind;
            if (kd == 0) continue;
            double sn = sx * fc[f].nx + sy * fc[f].ny;
            double ac = tau * fc[f].av * vb * sn;
            if (vb * sn >= 0) {
                gx -= fc[f].rx * ac * ls;
                gy -= fc[f].ry * ac * ls;
            } else if (kd == 2) {
                Ab[(size_t)fc[f].other * DB + t] -= ac * ls;
            }
            if (kd == 1) {
                double snj = sx * fc[f].nnx + sy * fc[f].nny;
                if (vb * snj < 0) {
                    double aj = tau * fc[f].avn * vb * snj;
                    double lj = LS[(size_t)fc[f].other * DB + t];
                    gx -= fc[f].rnx * aj * lj;
                    gy -= fc[f].rny * aj * lj;
                }
            }
        }
        if (!first_order) {
            LGx[i] += gx;
            LGy[i] += gy;
        }
    }
    red[threadIdx.x] = at;
    __syncthreads();
    for (int s = kAdjRow / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) ATn[c] = red[0];
}

// Transposed LSQ gradient plus the objective: lambda_e = L^T lambda_g + c,
// c = v s_axis w V / (A dT).  tptr / tsrc / twx / twy: per cell the cells
// whose gradient reads it and the weights (its own row with minus the sums).
__global__ void __launch_bounds__(kAdjRow) k_adj_lsqT(double *__restrict__ Ae, const double *__restrict__ LGx,
                                                      const double *__restrict__ LGy, const int *__restrict__ tptr,
                                                      const int *__restrict__ tsrc, const double *__restrict__ twx,
                                                      const double *__restrict__ twy, const double *__restrict__ cfac,
                                                      Modes m, int axis, int nc, int nB, int nD, int first_order) {
    int c = blockIdx.x;
    int DB = nD * nB;
    int a = tptr[c], z = tptr[c + 1];
    double cf = cfac[c];
    for (int t = threadIdx.x; t < DB; t += blockDim.x) {
        int d = t / nB, b = t - d * nB;
        double s = m.vg[b] * (axis == 0 ? m.sx[d] : m.sy[d]) * m.w[d] * cf;
        if (!first_order)
            for (int k = a; k < z; ++k) {
                size_t j = (size_t)tsrc[k] * DB + t;
                s += twx[k] * LGx[j] + twy[k] * LGy[j];
            }
        Ae[(size_t)c * DB + t] = s;
    }
}

// Column sums for the parameter terms, per (direction, band) column t:
//   0     sum_c lambda_S (e* - C (T* - Tref))      (e* held as e - C Tref)
//   1..5  sum_c Wc_j e*                            (j = xx, xy, yy, kappa, T)
//   6..10 sum_entries Wb_j eb*
// grid.y: cell chunks, then entry chunks; out[chunk][kCols][DB].
__global__ void __launch_bounds__(kAdjRow) k_adj_colsum(double *__restrict__ out, const double *__restrict__ LS,
                                                        const double *__restrict__ e, const double *__restrict__ T,
                                                        const double *__restrict__ eb, const double *__restrict__ Wc,
                                                        const double *__restrict__ Wb, const int *__restrict__ ent_list,
                                                        Modes m, int nc, int n_ent, int nB, int nD, int ncc) {
    int DB = nD * nB;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= DB) return;
    int ch = blockIdx.y;
    double acc[kCols];
#pragma unroll
    for (int j = 0; j < kCols; ++j) acc[j] = 0;
    int b = t % nB;
    double cb = m.cap[b];
    if (ch < ncc) {
        int c0 = ch * kColChunk, c1 = min(nc, c0 + kColChunk);
        for (int c = c0; c < c1; ++c) {
            size_t i = (size_t)c * DB + t;
            double ev = e[i];
            acc[0] += LS[i] * (ev - cb * (T[c] - m.eref));
#pragma unroll
            for (int j = 0; j < kAdjW; ++j) acc[1 + j] += Wc[(size_t)c * kAdjW + j] * ev;
        }
    } else {
        int k0 = (ch - ncc) * kColChunk, k1 = min(n_ent, k0 + kColChunk);
        for (int k = k0; k < k1; ++k) {
            int en = ent_list[k];
            double ev = eb[(size_t)en * DB + t];
#pragma unroll
            for (int j = 0; j < kAdjW; ++j) acc[6 + j] += Wb[(size_t)en * kAdjW + j] * ev;
        }
    }
#pragma unroll
    for (int j = 0; j < kCols; ++j) out[((size_t)ch * kCols + j) * DB + t] = acc[j];
}

}  // namespace

int Solver::adjoint(int axis, double dT, int max_iter, double tol, double *dk_ds, double *dk_dtb,
                    gbte_adjoint_stats *st) {
    auto t0 = std::chrono::high_resolution_clock::now();
    if (!ok()) return GBTE_HIP_ERROR;
    if (!check(hipSetDevice(opt.device), "hipSetDevice")) return GBTE_HIP_ERROR;
    const Problem &p = P;
    bool first_order = p.deck.order == "1", sequential = p.deck.method == "Sequential";
    if (dim != 2 || fp32_ || p.deck.order == "2limit" || p.n_matters != 1) {
        err_ = "the adjoint covers 2D meshes of one material, FP64 storage, Order 2 or 1";
        return GBTE_UNSUPPORTED;
    }
    for (int ib = 0; ib < nb; ++ib)
        if (p.bound_type[ib] == 4 || p.bound_type[ib] == 5) {
            err_ = "the adjoint does not cover interface boundaries";
            return GBTE_UNSUPPORTED;
        }
    if (!have_state) {
        err_ = "gbte_adjoint needs a converged gbte_solve on this handle first";
        return GBTE_BAD_INPUT;
    }
    if (axis < 0 || axis >= dim || dT == 0 || max_iter <= 0) {
        err_ = "adjoint: bad axis, zero temperature difference or no iterations";
        return GBTE_BAD_INPUT;
    }
    int m_pass = sequential ? 0 : std::max(1, last_passes_);
    size_t DB = (size_t)nD * nB, nE = (size_t)nc * DB, nEB = (size_t)ne * DB;
    int n_ent = (int)h_ent_list.size();
    double kap = p.kappa_bulk[0], Lx = p.deck.Lx;
    auto slot = [](int c, int j) { return (size_t)c * 6 + j; };
    // ---- host tables (device order)
    std::vector<int> be_slot(ne, -1), be_cell(ne, -1);
    for (int ib = 0; ib < nb; ++ib)
        for (int ic = 0; ic < 2; ++ic) {
            int c = p.bound_cell[ib * 2 + ic];
            if (c < 0) continue;
            be_slot[ib * 2 + ic] = (int)slot(c, p.bound_face[ib * 2 + ic]);
            be_cell[ib * 2 + ic] = c;
        }
    std::vector<int> slot_entry((size_t)nc * 6, -1);
    for (int en = 0; en < ne; ++en)
        if (be_slot[en] >= 0) slot_entry[be_slot[en]] = en;
    std::vector<AdjFace> faces((size_t)nc * kAdjFaces);
    for (int c = 0; c < nc; ++c) {
        if (p.face_size[c] > kAdjFaces) {
            err_ = "the adjoint covers cells of at most 4 faces";
            return GBTE_UNSUPPORTED;
        }
        for (int j = 0; j < kAdjFaces; ++j) {
            AdjFace &f = faces[(size_t)c * kAdjFaces + j];
            f = AdjFace{};
            if (j >= p.face_size[c]) continue;
            size_t s = slot(c, j);
            f.av = p.face_area[s] / p.volume[c];
            f.nx = p.face_nx[s];
            f.ny = p.face_ny[s];
            f.rx = p.face_cx[s] - p.cx[c];
            f.ry = p.face_cy[s] - p.cy[c];
            if (p.face_bound[s] == -1) {
                int n = p.face_neighbor[s];
                f.kind = 1;
                f.other = n;
                for (int k = 0; k < p.face_size[n]; ++k) {
                    size_t s2 = slot(n, k);
                    if (p.face_bound[s2] == -1 && p.face_neighbor[s2] == c) {
                        f.avn = p.face_area[s2] / p.volume[n];
                        f.nnx = p.face_nx[s2];
                        f.nny = p.face_ny[s2];
                        f.rnx = p.face_cx[s2] - p.cx[c];
                        f.rny = p.face_cy[s2] - p.cy[c];
                    }
                }
            } else {
                f.kind = slot_entry[s] >= 0 ? 2 : 0;
                f.other = slot_entry[s];
            }
        }
    }
    // transposed LSQ stencils (the forward's padded stencils, zero weights dropped)
    std::vector<std::vector<std::pair<int, std::pair<double, double>>>> lt(nc);
    if (!first_order) {
        int nnz = gnnz;
        for (int i = 0; i < nc; ++i) {
            if (!h_gok[i]) continue;
            double sx = 0, sy = 0;
            for (int k = h_gptr[i]; k < h_gptr[i + 1]; ++k) {
                double wx = h_gP[k] / Lx, wy = h_gP[(size_t)nnz + k] / Lx;
                if (wx == 0 && wy == 0) continue;
                lt[h_gidx[k]].push_back({i, {wx, wy}});
                sx += wx;
                sy += wy;
            }
            lt[i].push_back({i, {-sx, -sy}});
        }
    }
    std::vector<int> tptr(1, 0), tsrc;
    std::vector<double> twx, twy;
    for (int c = 0; c < nc; ++c) {
        for (auto &q : lt[c]) { tsrc.push_back(q.first); twx.push_back(q.second.first); twy.push_back(q.second.second); }
        tptr.push_back((int)tsrc.size());
    }
    // objective weights V / (A dT), A the node extent normal to the axis
    double lo = 1e300, hi = -1e300;
    for (int v = 0; v < nn; ++v) {
        double y = axis == 0 ? p.node_y[v] : p.node_x[v];
        lo = std::min(lo, y);
        hi = std::max(hi, y);
    }
    std::vector<double> cfac(nc);
    for (int c = 0; c < nc; ++c) cfac[c] = p.volume[c] / ((hi - lo) * dT);
    // downwind records from the forward's upwind schedule
    std::vector<int> item_cell((size_t)nD * nc), rq((size_t)nD * nc * kMaxIn), rc((size_t)nD * nc * kMaxIn);
    std::vector<double> rg((size_t)nD * nc * kMaxIn);
    if (!check(hipMemcpy(item_cell.data(), d_item_cell.p, item_cell.size() * sizeof(int), hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(rq.data(), d_rec_q.p, rq.size() * sizeof(int), hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(rc.data(), d_rec_c.p, rc.size() * sizeof(int), hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(rg.data(), d_rec_g.p, rg.size() * sizeof(double), hipMemcpyDeviceToHost), "D2H"))
        return GBTE_HIP_ERROR;
    std::vector<int> dn_c(rq.size(), -1), fill((size_t)nD * nc, 0);
    std::vector<double> dn_g(rq.size(), 0.0);
    for (int d = 0; d < nD; ++d)
        for (int k = 0; k < nc; ++k) {
            size_t P0 = (size_t)d * nc + k;
            for (int j = 0; j < kMaxIn; ++j) {
                int q = rq[P0 * kMaxIn + j];
                if (q < 0) continue;
                size_t Pu = (size_t)d * nc + q;
                int &n = fill[Pu];
                if (n >= kMaxIn) {
                    err_ = "adjoint: more than three downwind cells";
                    return GBTE_UNSUPPORTED;
                }
                dn_c[Pu * kMaxIn + n] = item_cell[P0];
                dn_g[Pu * kMaxIn + n] = rg[P0 * kMaxIn + j];
                ++n;
            }
        }
    // periodic partners: the entry whose incoming values read this one
    std::vector<int> btype(ne, 0), bconn(ne, 0), inv_conn(ne, -1);
    std::vector<double> btov(ne);
    if (!check(hipMemcpy(btype.data(), d_be_type.p, ne * sizeof(int), hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(bconn.data(), d_be_conn.p, ne * sizeof(int), hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(btov.data(), d_be_tov.p, ne * sizeof(double), hipMemcpyDeviceToHost), "D2H"))
        return GBTE_HIP_ERROR;
    for (int en : h_ent_list)
        if (btype[en] < 0) {
            int q = bconn[en];
            if (inv_conn[q] >= 0) {
                err_ = "adjoint: a periodic entry read by two partners";
                return GBTE_UNSUPPORTED;
            }
            inv_conn[q] = en;
        }
    // macro operators: boundary coefficients B (per entry), the
    // non-orthogonal correction weights (per slot, on its two vertices) and
    // the cell-to-node IDW of temperatures
    std::vector<double> Bent(ne, 0.0), corr((size_t)nc * 6, 0.0);
    std::vector<int> cv0((size_t)nc * 6, 0), cv1((size_t)nc * 6, 0);
    for (int c = 0; c < nc; ++c)
        for (int j = 0; j < p.face_size[c]; ++j) {
            size_t s = slot(c, j);
            int fb = p.face_bound[s];
            double ax, ay;
            if (fb >= 0) {
                if (p.bound_type[fb] == 3) continue;
                ax = p.face_cx[s] - p.cx[c];
                ay = p.face_cy[s] - p.cy[c];
                for (int ic = 0; ic < 2; ++ic)
                    if (p.bound_cell[fb * 2 + ic] == c)
                        Bent[fb * 2 + ic] = kap * p.face_area[s] / (p.face_nx[s] * ax + p.face_ny[s] * ay);
            } else {
                int n = p.face_neighbor[s];
                ax = p.cx[n] - p.cx[c];
                ay = p.cy[n] - p.cy[c];
            }
            double dn = ax * p.face_nx[s] + ay * p.face_ny[s];
            double a2x = p.face_nx[s] - ax / dn, a2y = p.face_ny[s] - ay / dn;
            int v0 = p.face_vert[s * 4], v1 = p.face_vert[s * 4 + 1];
            double gx = p.node_x[v1] - p.node_x[v0], gy = p.node_y[v1] - p.node_y[v0];
            corr[s] = kap * p.face_area[s] * (a2x * gx + a2y * gy) / (gx * gx + gy * gy);
            cv0[s] = v0;
            cv1[s] = v1;
        }
    std::vector<double> tR(nn, 0.0);
    std::vector<double> tdist((size_t)nc * 8, 0.0);
    for (int c = 0; c < nc; ++c)
        for (int j = 0; j < p.cell_nvert[c]; ++j) {
            int v = p.cell_vert[(size_t)c * 8 + j];
            double dd = std::sqrt(std::pow(p.node_x[v] - p.cx[c], 2) + std::pow(p.node_y[v] - p.cy[c], 2) +
                                  std::pow(p.node_z[v] - p.cz[c], 2));
            tdist[(size_t)c * 8 + j] = dd;
            tR[v] += 1.0 / dd;
        }
    // ---- device buffers
    DBuf<double> Ae, Ab, Lin, Lout, LS, LGx, LGy, Wc, Wb, ATn, twp, xrhs, xsol, cf, twxd, twyd, dng, part;
    DBuf<int> tptrd, tsrcd, dnc, invd;
    DBuf<AdjFace> fd;
    int ncc = (nc + kColChunk - 1) / kColChunk, nec = (std::max(n_ent, 1) + kColChunk - 1) / kColChunk;
    if (!alloc(Ae, nE) || !alloc(LS, nE) || !alloc(LGx, nE) || !alloc(LGy, nE) || !alloc(Ab, nEB) ||
        !alloc(Lin, nEB) || !alloc(Lout, nEB) || !alloc(Wc, (size_t)nc * kAdjW) || !alloc(Wb, (size_t)ne * kAdjW) ||
        !alloc(ATn, nc) || !alloc(twp, ne) || !alloc(xrhs, nc) || !alloc(xsol, nc) ||
        !alloc(part, (size_t)(ncc + nec) * kCols * DB) || !upload(cf, cfac) || !upload(tptrd, tptr) ||
        !upload(tsrcd, tsrc) || !upload(twxd, twx) || !upload(twyd, twy) || !upload(dnc, dn_c) ||
        !upload(dng, dn_g) || !upload(invd, inv_conn) || !upload(fd, faces))
        return GBTE_HIP_ERROR;
    Modes m = modes();
    // start: lambda = c (the objective), zero boundary and T adjoints
    if (!check(hipMemsetAsync(Ab.p, 0, nEB * sizeof(double), s_), "memset") ||
        !check(hipMemsetAsync(twp.p, 0, ne * sizeof(double), s_), "memset"))
        return GBTE_HIP_ERROR;
    hipLaunchKernelGGL(k_adj_lsqT, dim3(nc), dim3(kAdjRow), 0, s_, Ae.p, LGx.p, LGy.p, tptrd.p, tsrcd.p, twxd.p,
                       twyd.p, cf.p, m, axis, nc, nB, nD, 1);
    std::vector<double> AT(nc, 0.0), ATnew(nc), mu(nc), lamT(nc), nu(nc), lamTV(nn), tm(ne), Wch((size_t)nc * kAdjW),
        Wbh((size_t)ne * kAdjW), eV(nn);
    std::vector<std::vector<double>> nus;
    int Bc = std::min(nB, 16), nchunk = (nB + Bc - 1) / Bc;
    int cpw = std::max(1, 64 / nB), bw = blocks_for(blocks_for(n_ent, cpw), kWaves);
    int it = 0;
    double change = 0;
    bool conv = false;
    for (it = 0; it < max_iter; ++it) {
        // macro passes in reverse: nu_k = K^-1 lambda_T^k, mu = sum nu_k,
        // lambda_tmac = -B nu, lambda_T^{k-1} = -IDW_T^T Corr^T nu_k
        std::fill(mu.begin(), mu.end(), 0.0);
        std::fill(tm.begin(), tm.end(), 0.0);
        lamT = AT;
        nus.clear();
        for (int k = m_pass; k >= 1; --k) {
            // nd_solve / linsolve solve (-K) x = b: b = -lambda gives K^-1 lambda
            for (int c = 0; c < nc; ++c) nu[c] = -lamT[c];
            if (!check(hipMemcpyAsync(xrhs.p, nu.data(), nc * sizeof(double), hipMemcpyHostToDevice, s_), "H2D"))
                return GBTE_HIP_ERROR;
            int cgit = 0;
            if (!check(hipMemsetAsync(xsol.p, 0, nc * sizeof(double), s_), "memset") ||
                !linsolve(Kmacro, NDmacro, xrhs.p, xsol.p, &cgit))
                return GBTE_HIP_ERROR;
            if (!check(hipMemcpyAsync(nu.data(), xsol.p, nc * sizeof(double), hipMemcpyDeviceToHost, s_), "D2H") ||
                !check(hipStreamSynchronize(s_), "sync"))
                return GBTE_HIP_ERROR;
            nus.push_back(nu);
            for (int c = 0; c < nc; ++c) mu[c] += nu[c];
            for (int en = 0; en < ne; ++en)
                if (be_cell[en] >= 0 && std::isnan(btov[en])) tm[en] -= Bent[en] * nu[be_cell[en]];
            std::fill(lamTV.begin(), lamTV.end(), 0.0);
            for (int c = 0; c < nc; ++c)
                for (int j = 0; j < p.face_size[c]; ++j) {
                    size_t s = slot(c, j);
                    if (corr[s] == 0) continue;
                    lamTV[cv1[s]] -= corr[s] * nu[c];
                    lamTV[cv0[s]] += corr[s] * nu[c];
                }
            for (int c = 0; c < nc; ++c) {
                double a = 0;
                for (int j = 0; j < p.cell_nvert[c]; ++j) {
                    int v = p.cell_vert[(size_t)c * 8 + j];
                    a += (1.0 / tdist[(size_t)c * 8 + j]) / tR[v] * lamTV[v];
                }
                lamT[c] = a;
            }
        }
        // non-Fourier chain transposed for the four system factors:
        // Wc / Wb[j] = geometric adjoint of the cell / entry values
        std::fill(Wch.begin(), Wch.end(), 0.0);
        std::fill(Wbh.begin(), Wbh.end(), 0.0);
        if (!sequential)
            for (int j = 0; j < 4; ++j) {
                std::fill(eV.begin(), eV.end(), 0.0);
                auto lam_side = [&](int c, size_t s, double area, double &lx, double &ly) {
                    double nx = p.face_nx[s], ny = p.face_ny[s], u = mu[c];
                    if (j == 0) { lx -= u * area * nx; }
                    else if (j == 1) { lx -= u * area * ny; ly -= u * area * nx; }
                    else if (j == 2) { ly -= u * area * ny; }
                    else { lx += u * kap * p.face_area[s] * nx; ly += u * kap * p.face_area[s] * ny; }
                };
                for (int f = 0; f < nf; ++f) {
                    int fb = p.fface_bound[f];
                    if (fb == -1) {
                        int i = p.fface_cell[f * 2], i2 = p.fface_cell[f * 2 + 1];
                        size_t s0 = slot(i, p.fface_face[f * 2]), s1 = slot(i2, p.fface_face[f * 2 + 1]);
                        double lx = 0, ly = 0;
                        lam_side(i, s0, p.face_area[s0], lx, ly);
                        lam_side(i2, s1, p.face_area[s0], lx, ly);
                        double ax = p.cx[i2] - p.cx[i], ay = p.cy[i2] - p.cy[i], az = p.cz[i2] - p.cz[i];
                        double al = (ax * p.face_nx[s0] + ay * p.face_ny[s0] + az * p.face_nz[s0]) /
                                    (ax * ax + ay * ay + az * az);
                        double lgn = p.face_nx[s0] * lx + p.face_ny[s0] * ly;
                        Wch[(size_t)i2 * kAdjW + j] += al * lgn;
                        Wch[(size_t)i * kAdjW + j] -= al * lgn;
                        int v0 = p.face_vert[s0 * 4], v1 = p.face_vert[s0 * 4 + 1];
                        double mx = p.node_x[v1] - p.node_x[v0], my = p.node_y[v1] - p.node_y[v0];
                        double lt2 = (mx * lx + my * ly) / (mx * mx + my * my);
                        eV[v1] += lt2;
                        eV[v0] -= lt2;
                    } else if (p.bound_type[fb] != 3) {
                        for (int ic = 0; ic < 2; ++ic) {
                            int ie = p.fface_cell[f * 2 + ic];
                            if (ie < 0) continue;
                            int ibc = 0;
                            for (int k = 0; k < 2; ++k)
                                if (p.bound_cell[fb * 2 + k] == ie) ibc = k;
                            size_t s = slot(ie, p.fface_face[f * 2 + ic]);
                            double lx = 0, ly = 0;
                            lam_side(ie, s, p.face_area[s], lx, ly);
                            double ax = p.face_cx[s] - p.cx[ie], ay = p.face_cy[s] - p.cy[ie], az = p.face_cz[s] - p.cz[ie];
                            double al = (ax * p.face_nx[s] + ay * p.face_ny[s] + az * p.face_nz[s]) /
                                        (ax * ax + ay * ay + az * az);
                            double lgn = p.face_nx[s] * lx + p.face_ny[s] * ly;
                            Wbh[(size_t)(fb * 2 + ibc) * kAdjW + j] += al * lgn;
                            Wch[(size_t)ie * kAdjW + j] -= al * lgn;
                            int v0 = p.face_vert[s * 4], v1 = p.face_vert[s * 4 + 1];
                            double mx = p.node_x[v1] - p.node_x[v0], my = p.node_y[v1] - p.node_y[v0];
                            double lt2 = (mx * lx + my * ly) / (mx * mx + my * my);
                            eV[v1] += lt2;
                            eV[v0] -= lt2;
                        }
                    }
                }
                // energy-density IDW transposed (cells or boundary entries)
                for (int v = 0; v < nn; ++v)
                    for (int k = h_vptr[v]; k < h_vptr[v + 1]; ++k) {
                        int sidx = h_vsrc[k];
                        double w = h_vwt[k] * eV[v];
                        if (sidx >= 0) Wch[(size_t)sidx * kAdjW + j] += w;
                        else Wbh[(size_t)(-sidx - 1) * kAdjW + j] += w;
                    }
            }
        for (int c = 0; c < nc; ++c) Wch[(size_t)c * kAdjW + 4] = sequential ? AT[c] : lamT[c];
        for (int en = 0; en < ne; ++en) Wbh[(size_t)en * kAdjW + 4] = tm[en];
        if (!check(hipMemcpyAsync(Wc.p, Wch.data(), Wch.size() * sizeof(double), hipMemcpyHostToDevice, s_), "H2D") ||
            !check(hipMemcpyAsync(Wb.p, Wbh.data(), Wbh.size() * sizeof(double), hipMemcpyHostToDevice, s_), "H2D"))
            return GBTE_HIP_ERROR;
        // one application of J^T (+ c)
        if (n_ent > 0) {
            hipLaunchKernelGGL(k_adj_bc_in, dim3(std::max(bw, 1)), dim3(kBlock), 0, s_, Ab.p, Lin.p, twp.p, Wb.p, m,
                               d_ent_list.p, d_be_type.p, d_be_nx.p, d_be_ny.p, n_ent, nB, nD, cpw);
            hipLaunchKernelGGL(k_adj_bc_out, dim3(std::max(bw, 1)), dim3(kBlock), 0, s_, Lout.p, Lin.p, Wb.p, m,
                               d_ent_list.p, d_be_type.p, invd.p, d_be_spec.p, d_be_nx.p, d_be_ny.p, n_ent, nB, nD,
                               cpw);
        }
        hipLaunchKernelGGL(k_adj_pre, dim3(nc), dim3(kAdjRow), 0, s_, Ae.p, LGx.p, LGy.p, Lout.p, Wc.p, fd.p, m, nc,
                           nB, nD);
        hipLaunchKernelGGL(k_adj_sweep, dim3(nD * nchunk), dim3(kAdjSweep), 0, s_, LS.p, Ae.p, d_item_cell.p,
                           d_levptr.p, d_dlev.p, dnc.p, dng.p, d_rec_G.p, m, nc, nB, nD, Bc, nchunk);
        hipLaunchKernelGGL(k_adj_src, dim3(nc), dim3(kAdjRow), 0, s_, LGx.p, LGy.p, Ab.p, ATn.p, LS.p, fd.p, m, nc,
                           nB, nD, first_order ? 1 : 0);
        hipLaunchKernelGGL(k_adj_lsqT, dim3(nc), dim3(kAdjRow), 0, s_, Ae.p, LGx.p, LGy.p, tptrd.p, tsrcd.p, twxd.p,
                           twyd.p, cf.p, m, axis, nc, nB, nD, first_order ? 1 : 0);
        if (!check(hipGetLastError(), "adjoint kernels") ||
            !check(hipMemcpyAsync(ATnew.data(), ATn.p, nc * sizeof(double), hipMemcpyDeviceToHost, s_), "D2H") ||
            !check(hipStreamSynchronize(s_), "sync"))
            return GBTE_HIP_ERROR;
        double dmax = 0, amax = 0;
        for (int c = 0; c < nc; ++c) {
            dmax = std::max(dmax, std::fabs(ATnew[c] - AT[c]));
            amax = std::max(amax, std::fabs(ATnew[c]));
        }
        AT.swap(ATnew);
        change = amax > 0 ? dmax / amax : dmax;
        if (std::isnan(change)) {
            err_ = "adjoint iteration produced NaN";
            return GBTE_NAN;
        }
        if (it > 0 && change <= tol) {
            conv = true;
            ++it;
            break;
        }
    }
    // ---- parameter terms from the last application (its lambda_S, mu, nu,
    // lambda_Trec / lambda_tmac and incoming-value adjoints)
    int ny = ncc + nec;
    hipLaunchKernelGGL(k_adj_colsum, dim3(blocks_for((long)DB, kAdjRow), ny), dim3(kAdjRow), 0, s_, part.p, LS.p,
                       d_e.p, d_T.p, d_eb.p, Wc.p, Wb.p, d_ent_list.p, m, nc, n_ent, nB, nD, ncc);
    std::vector<double> ph((size_t)ny * kCols * DB), tw(ne), RM(nc), col((size_t)kCols * DB, 0.0);
    // the handle's stream is non-blocking: wait for it before the copies
    if (!check(hipGetLastError(), "colsum") || !check(hipStreamSynchronize(s_), "sync") ||
        !check(hipMemcpy(ph.data(), part.p, ph.size() * sizeof(double), hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(tw.data(), twp.p, ne * sizeof(double), hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(RM.data(), d_RM.p, nc * sizeof(double), hipMemcpyDeviceToHost), "D2H"))
        return GBTE_HIP_ERROR;
    for (int y = 0; y < ny; ++y)
        for (size_t k = 0; k < col.size(); ++k) col[k] += ph[(size_t)y * kCols * DB + k];
    std::vector<double> sw(nB, 0.0), nft(nB, 0.0), nfk(nB, 0.0), slr(nB, 0.0);
    for (int d = 0; d < nD; ++d)
        for (int b = 0; b < nB; ++b) {
            size_t t = (size_t)d * nB + b;
            double w = p.weight[d], sx = p.dir_x[d], sy = p.dir_y[d], vb = p.vg[b], tau = p.tau[b], cb = p.cap[b];
            double h = p.hr[b] > 0 ? 1.0 : 0.0;
            double Y[kCols];
            for (int j = 0; j < kCols; ++j) Y[j] = col[(size_t)j * DB + t];
            sw[b] += Y[0];
            double tvv = h * w * tau * vb * vb;
            nft[b] += tvv * (sx * sx * (Y[1] + Y[6]) + sx * sy * (Y[2] + Y[7]) + sy * sy * (Y[3] + Y[8]));
            nfk[b] += h * w * p.lr[b] / cb * (Y[4] + Y[9]);
            slr[b] += w / cb * (Y[5] + Y[10]) + h * w / cb * (Y[4] + Y[9]);
        }
    double nfk_all = 0;
    for (int b = 0; b < nB; ++b) nfk_all += nfk[b];
    double s_kap = nfk_all / kap;
    for (auto &v : nus)
        for (int c = 0; c < nc; ++c) s_kap -= v[c] * (RM[c] - p.heat_source[c] * p.volume[c]) / kap;
    double lrsum = 0;
    for (int k = 0; k < nB; ++k) lrsum += slr[k] * p.lr[k];
    int dmat = p.deck.material_dim;
    for (int b = 0; b < nB; ++b) {
        double tau = p.tau[b];
        // sum_k s_k d lr_k / d tau_b = (lr_b sum_k s_k lr_k - s_b lr_b) / tau_b
        double g = (sw[b] + nft[b]) / tau + (p.lr[b] * lrsum - slr[b] * p.lr[b]) / tau +
                   s_kap * p.cap[b] * p.vg[b] * p.vg[b] / dmat;
        dk_ds[b] = tau * g;
    }
    if (dk_dtb) {
        for (int ib = 0; ib < nb; ++ib) dk_dtb[ib] = 0;
        for (int en = 0; en < ne; ++en)
            if (be_cell[en] >= 0 && p.bound_type[en / 2] == 1) dk_dtb[en / 2] += tw[en];
    }
    if (st) {
        st->iterations = it;
        st->converged = conv ? 1 : 0;
        st->change = change;
        st->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count();
    }
    return GBTE_OK;
}

}  // namespace gbte
