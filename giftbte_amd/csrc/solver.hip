// MI355X (gfx950) solver for the steady StaticBTESynthetic path.
//
// One outer iteration of StaticBTESynthetic::solve
// (src/StaticBTESynthetic/StaticBTESyntheticSolver.cpp:63-221) becomes:
//   k_copy_old         copy()                       StaticBTESynthetic.cpp:2655-2690
//   k_sweep  (x L)     _get_Re + SparseLU solve + _get_bound_ee, all systems,
//                      one launch per upwind level   :1427-1539, Solver.cpp:84-108
//   k_vertex_e         _set_vertex_energydensity     :1541-1616
//   k_face_moments     _get_face_gradient + _get_non_fourier + moments
//                      + next iteration's _get_gradient_larger
//                                                    :1651-1937, :2485-2520, :1243-1340
//   k_boundary         _set_bound_ee_1 + _set_macro_boundary_condition
//                                                    :2083-2274
//   macro loop         _marco_solution: k_vertex_T, k_rhs, PCG   :2276-2407, :2692-2734
//   k_err              _get_magin_check_error        :2565-2611
//
// Layout in HBM (lanes of a wavefront run over bands; all bands of one
// direction share geometry, upwind order and coefficients):
//   e     [dir][cell][band]           energy density, in place
//   grad  [dir][cell][dim][band]      LSQ gradient of the previous sweep's e
//   eb    [dir][entry][band]          boundary values, entry = 2*bound + side
//   eV    [dir][node][band]           vertex (IDW) values
// A wavefront holds floor(64 / n_bands) cells of one (direction, level) list.
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

namespace gbte {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

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

struct Modes {
    const double *vg, *tau, *cap, *lr, *hr, *sx, *sy, *sz, *w;
    double kappa, cbulk;
};

struct SlotGeo {
    const int *fsize, *nb, *ent;
    const double *nx, *ny, *nz, *area, *vol, *rx, *ry, *rz, *qx, *qy, *qz;
};

// ---------------------------------------------------------------- kernels

// _set_initial (StaticBTESynthetic.cpp:765-782): e = T C_b for all systems.
__global__ void k_init_e(double *__restrict__ e, const double *__restrict__ T, const double *__restrict__ cap, int nc,
                         int nB, size_t total) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    int b = (int)(i % nB);
    size_t dc = i / nB;
    int c = (int)(dc % nc);
    e[i] = T[c] * cap[b];
}

// _get_gradient_larger, Order 2 (StaticBTESynthetic.cpp:1243-1340).
__device__ __forceinline__ void lsq_gradient(const double *__restrict__ ed, int c, int b, int nB, double ec,
                                             const int *__restrict__ gptr, const int *__restrict__ gidx,
                                             const double *__restrict__ gP, const int *__restrict__ gok, int dim,
                                             int nnz, double invLx, double &gx, double &gy, double &gz) {
    gx = 0; gy = 0; gz = 0;
    if (!gok[c]) return;
    int a = gptr[c], z = gptr[c + 1];
    for (int k = a; k < z; ++k) {
        double d1 = invLx * (ed[(size_t)gidx[k] * nB + b] - ec);
        gx += gP[k] * d1;
        gy += gP[(size_t)nnz + k] * d1;
        if (dim == 3) gz += gP[2 * (size_t)nnz + k] * d1;
    }
    if (isnan(gx) || isnan(gy) || isnan(gz)) { gx = 0; gy = 0; gz = 0; }
}

__global__ void k_gradient(const double *__restrict__ e, double *__restrict__ grad, const int *__restrict__ gptr,
                           const int *__restrict__ gidx, const double *__restrict__ gP, const int *__restrict__ gok,
                           int dim, int nnz, double invLx, int nc, int nB, int nD, int cpw) {
    int lane = lane_id(), sub = lane / nB, b = lane - sub * nB;
    long item = (long)wave_global() * cpw + sub;
    if (sub >= cpw || item >= (long)nc * nD) return;
    int d = (int)(item / nc), c = (int)(item % nc);
    const double *ed = e + (size_t)d * nc * nB;
    double gx, gy, gz;
    lsq_gradient(ed, c, b, nB, ed[(size_t)c * nB + b], gptr, gidx, gP, gok, dim, nnz, invLx, gx, gy, gz);
    double *g = grad + (((size_t)d * nc + c) * dim) * nB + b;
    g[0] = gx;
    g[nB] = gy;
    if (dim == 3) g[2 * nB] = gz;
}

// _get_bound_ee (StaticBTESynthetic.cpp:1502-1539) for all systems; used once
// before the loop (inside the loop the sweep writes outgoing values).
__global__ void k_bound_out(const double *__restrict__ e, const double *__restrict__ grad, double *__restrict__ eb,
                            SlotGeo g, Modes m, const int *__restrict__ ent_list, const int *__restrict__ be_cell,
                            const int *__restrict__ be_slot, int n_ent, int nc, int ne, int nB, int nD, int dim, int cpw) {
    int lane = lane_id(), sub = lane / nB, b = lane - sub * nB;
    long item = (long)wave_global() * cpw + sub;
    if (sub >= cpw || item >= (long)n_ent * nD) return;
    int d = (int)(item / n_ent), k = (int)(item % n_ent);
    int en = ent_list[k], c = be_cell[en], s = be_slot[en];
    double vb = m.vg[b];
    double dot = vb * m.sx[d] * g.nx[s] + vb * m.sy[d] * g.ny[s] + vb * m.sz[d] * g.nz[s];
    if (dot >= 0) {
        const double *gr = grad + (((size_t)d * nc + c) * dim) * nB + b;
        double gz = dim == 3 ? gr[2 * nB] : 0.0;
        double ev = e[((size_t)d * nc + c) * nB + b] + (g.rx[s] * gr[0] + g.ry[s] * gr[nB] + g.rz[s] * gz);
        eb[((size_t)d * ne + en) * nB + b] = ev;
    }
}

// One upwind level of every direction.  Per (direction, cell) and band lane:
//   e_new = [C T_old + Q hr tau - sum_out a (r.grad e_c)
//            - sum_in,interior a (e_new_nb + r'.grad e_nb) - sum_in,boundary a eb]
//           / (1 + sum_out a),     a = tau (A/V) (v.n)
// which is e + delta with M delta = R of _get_Re / the SparseLU solve
// (StaticBTESynthetic.cpp:784-840, 1427-1500; Solver.cpp:96-100) written
// for e_new directly.  Outgoing boundary values of _get_bound_ee follow.
__global__ void __launch_bounds__(kBlock) k_sweep(const int *__restrict__ item_c, const int *__restrict__ item_d,
                                                  int n_items, double *__restrict__ e, const double *__restrict__ grad,
                                                  double *__restrict__ eb, const double *__restrict__ Told,
                                                  const double *__restrict__ Q, SlotGeo g, Modes m, int nc, int ne,
                                                  int nB, int dim, int cpw) {
    int lane = lane_id(), sub = lane / nB, b = lane - sub * nB;
    int item = wave_global() * cpw + sub;
    if (sub >= cpw || item >= n_items) return;
    int c = item_c[item], d = item_d[item];
    double vb = m.vg[b], tb = m.tau[b], Cb = m.cap[b], hb = m.hr[b];
    double vx = vb * m.sx[d], vy = vb * m.sy[d], vz = vb * m.sz[d];
    double *ed = e + (size_t)d * nc * nB;
    const double *gd = grad + (size_t)d * nc * dim * nB;
    double *ebd = eb + (size_t)d * ne * nB;
    const double *gc = gd + (size_t)c * dim * nB + b;
    double gcx = gc[0], gcy = gc[nB], gcz = dim == 3 ? gc[2 * nB] : 0.0;
    double S = Told[c] * Cb + Q[c] * hb * tb;
    double M = 0;
    int nf = g.fsize[c];
    double vol = g.vol[c];
    for (int j = 0; j < nf; ++j) {
        int s = c * 6 + j;
        double dot = vx * g.nx[s] + vy * g.ny[s] + vz * g.nz[s];
        double a = tb * g.area[s] / vol * dot;
        if (dot >= 0) {
            M += a;
            S -= a * (g.rx[s] * gcx + g.ry[s] * gcy + g.rz[s] * gcz);
        } else {
            int nb = g.nb[s];
            if (nb >= 0) {
                const double *gn = gd + (size_t)nb * dim * nB + b;
                double gnz = dim == 3 ? gn[2 * nB] : 0.0;
                double up = ed[(size_t)nb * nB + b] + (g.qx[s] * gn[0] + g.qy[s] * gn[nB] + g.qz[s] * gnz);
                S -= a * up;
            } else {
                S -= a * ebd[(size_t)g.ent[s] * nB + b];
            }
        }
    }
    M += 1.0;
    double en = S / M;
    ed[(size_t)c * nB + b] = en;
    for (int j = 0; j < nf; ++j) {
        int s = c * 6 + j;
        int et = g.ent[s];
        if (et < 0) continue;
        double dot = vx * g.nx[s] + vy * g.ny[s] + vz * g.nz[s];
        if (dot >= 0) ebd[(size_t)et * nB + b] = en + (g.rx[s] * gcx + g.ry[s] * gcy + g.rz[s] * gcz);
    }
}

// Inverse-distance vertex values (_set_vertex_energydensity,
// StaticBTESynthetic.cpp:1541-1616): interior nodes from cells, nodes on a
// boundary face from the boundary values of those faces.
__global__ void k_vertex_e(const double *__restrict__ e, const double *__restrict__ eb, double *__restrict__ eV,
                           const int *__restrict__ vptr, const int *__restrict__ vsrc, const double *__restrict__ vdist,
                           int nn, int nc, int ne, int nB, int nD, int cpw) {
    int lane = lane_id(), sub = lane / nB, b = lane - sub * nB;
    long item = (long)wave_global() * cpw + sub;
    if (sub >= cpw || item >= (long)nn * nD) return;
    int d = (int)(item / nn), v = (int)(item % nn);
    const double *ed = e + (size_t)d * nc * nB;
    const double *ebd = eb + (size_t)d * ne * nB;
    double t = 0, r = 0;
    for (int k = vptr[v]; k < vptr[v + 1]; ++k) {
        int src = vsrc[k];
        double val = src >= 0 ? ed[(size_t)src * nB + b] : ebd[(size_t)(-src - 1) * nB + b];
        t += val / vdist[k];
        r += 1.0 / vdist[k];
    }
    double x = t / r;
    if (isnan(x)) x = 0;
    eV[((size_t)d * nn + v) * nB + b] = x;
}

struct FaceK {
    const int *kind, *other, *v0, *v1;
    const double *an, *aa, *n0x, *n0y, *n0z, *mx, *my, *m2;
};

// Per cell, looping over directions with band lanes:
//   face gradients (_get_face_gradient, StaticBTESynthetic.cpp:1651-1783),
//   the non-Fourier residual (_get_non_fourier, :1883-1937),
//   the moments (_recover_temperature, _get_total_energy, _get_heat_flux,
//   :2485-2520), and the LSQ gradient of the new e for the next sweep.
// kind: 0 none (specular), 1 interior face owned first by this cell,
//       3 interior face owned first by the neighbour, 2 boundary face.
__global__ void __launch_bounds__(kBlock) k_face_moments(
    const double *__restrict__ e, const double *__restrict__ eb, const double *__restrict__ eV,
    double *__restrict__ grad, SlotGeo g, FaceK f, Modes m, const int *__restrict__ gptr, const int *__restrict__ gidx,
    const double *__restrict__ gP, const int *__restrict__ gok, int nnz, double invLx, double *__restrict__ T,
    double *__restrict__ E, double *__restrict__ qx, double *__restrict__ qy, double *__restrict__ qz,
    double *__restrict__ RM, double *__restrict__ Tb, int nc, int ne, int nn, int nB, int nD, int dim, int cpw) {
    int lane = lane_id(), sub = lane / nB, b = lane - sub * nB;
    int c = wave_global() * cpw + sub;
    bool act = sub < cpw && c < nc;
    int cc = act ? c : 0;
    double vb = m.vg[b], tb = m.tau[b], Cb = m.cap[b], lb = m.lr[b], hb = m.hr[b];
    double aT = 0, aTb = 0, aE = 0, ax = 0, ay = 0, az = 0, aR = 0;
    int nf = g.fsize[cc];
    for (int d = 0; d < nD; ++d) {
        const double *ed = e + (size_t)d * nc * nB;
        const double *ebd = eb + (size_t)d * ne * nB;
        const double *eVd = eV + (size_t)d * nn * nB;
        double ec = ed[(size_t)cc * nB + b];
        double wd = m.w[d];
        double vx = vb * m.sx[d], vy = vb * m.sy[d], vz = vb * m.sz[d];
        aT += lb * ec * wd / Cb;
        aTb += ec * wd / Cb;
        aE += ec * wd / m.cbulk;
        ax += vx * wd * ec;
        ay += vy * wd * ec;
        az += vz * wd * ec;
        if (hb > 0) {
            for (int j = 0; j < nf; ++j) {
                int s = cc * 6 + j;
                int kd = f.kind[s];
                if (kd == 0) continue;
                double de;
                int oth = f.other[s];
                if (kd == 1) de = ed[(size_t)oth * nB + b] - ec;
                else if (kd == 3) de = ec - ed[(size_t)oth * nB + b];
                else de = ebd[(size_t)oth * nB + b] - ec;
                double gn = de * f.an[s] / f.aa[s];
                double dv = eVd[(size_t)f.v1[s] * nB + b] - eVd[(size_t)f.v0[s] * nB + b];
                double gx = dv / f.m2[s] * f.mx[s], gy = dv / f.m2[s] * f.my[s], gz = 0;
                gx = gn * f.n0x[s] + gx;
                gy = gn * f.n0y[s] + gy;
                gz = gn * f.n0z[s] + gz;
                double nx = g.nx[s], ny = g.ny[s], nz = g.nz[s], A = g.area[s];
                aR -= wd * A * tb * (vx * gx + vy * gy + vz * gz) * (vx * nx + vy * ny + vz * nz);
                aR += m.kappa * wd * A * lb / Cb * (nx * gx + ny * gy + nz * gz);
            }
        }
        double gx, gy, gz;
        lsq_gradient(ed, cc, b, nB, ec, gptr, gidx, gP, gok, dim, nnz, invLx, gx, gy, gz);
        if (act) {
            double *gr = grad + (((size_t)d * nc + cc) * dim) * nB + b;
            gr[0] = gx;
            gr[nB] = gy;
            if (dim == 3) gr[2 * nB] = gz;
        }
    }
    if (act) Tb[(size_t)b * nc + c] = aTb;
    aT = group_sum(aT, b, nB);
    aE = group_sum(aE, b, nB);
    ax = group_sum(ax, b, nB);
    ay = group_sum(ay, b, nB);
    az = group_sum(az, b, nB);
    aR = group_sum(aR, b, nB);
    if (act && b == 0) {
        T[c] = aT;
        E[c] = aE;
        qx[c] = ax;
        qy[c] = ay;
        qz[c] = az;
        RM[c] = aR;
    }
}

// _set_bound_ee_1 (StaticBTESynthetic.cpp:2083-2232) at np=1 and
// _set_macro_boundary_condition (:2234-2274).  One group of band lanes per
// boundary entry; directions are visited in increasing order, as the
// reference's per-system Allgather makes later directions see earlier
// updates (this matters only for specular partners that are incoming).
__global__ void k_boundary(double *__restrict__ eb, double *__restrict__ tmac, Modes m,
                           const int *__restrict__ ent_list, const int *__restrict__ be_type,
                           const int *__restrict__ be_conn, const int *__restrict__ be_spec,
                           const double *__restrict__ be_nx, const double *__restrict__ be_ny,
                           const double *__restrict__ be_nz, const double *__restrict__ be_tw,
                           const double *__restrict__ be_dT, const double *__restrict__ be_tov, int n_ent, int ne,
                           int nB, int nD, int cpw) {
    int lane = lane_id(), sub = lane / nB, b = lane - sub * nB;
    int k = wave_global() * cpw + sub;
    bool act = sub < cpw && k < n_ent;
    int en = ent_list[act ? k : 0];
    double vb = m.vg[b], Cb = m.cap[b], lb = m.lr[b], hb = m.hr[b];
    double nx = be_nx[en], ny = be_ny[en], nz = be_nz[en];
    int ty = be_type[en];
    double tm = 0;
    if (act) {
        double diff = 0;
        if (ty == 2) {
            double num = 0, den = 0;
            for (int q = 0; q < nD; ++q) {
                double d1 = vb * m.sx[q] * nx + vb * m.sy[q] * ny + vb * m.sz[q] * nz;
                if (d1 >= 0) {
                    num += eb[((size_t)q * ne + en) * nB + b] * (d1 * m.w[q]);
                    den += (d1 * m.w[q]);
                }
            }
            diff = num / den;
        }
        for (int d = 0; d < nD; ++d) {
            size_t at = ((size_t)d * ne + en) * nB + b;
            double dot = vb * m.sx[d] * nx + vb * m.sy[d] * ny + vb * m.sz[d] * nz;
            if (hb != 0 && dot < 0) {
                if (ty == 1) eb[at] = Cb * be_tw[en];
                else if (ty == 2) eb[at] = diff;
                else if (ty == 3) eb[at] = eb[((size_t)be_spec[(size_t)en * nD + d] * ne + en) * nB + b];
                else if (ty < 0) eb[at] = eb[((size_t)d * ne + be_conn[en]) * nB + b] + Cb * be_dT[en];
            }
            tm += lb * eb[at] / Cb * m.w[d];
        }
    }
    tm = group_sum(tm, b, nB);
    if (act && b == 0) {
        double ov = be_tov[en];
        tmac[en] = isnan(ov) ? tm : ov;
    }
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

inline int blocks_for(long n, int per_block) { return (int)((n + per_block - 1) / per_block); }

}  // namespace

// ======================================================================
// host side
// ======================================================================

bool Solver::check(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    if (err_.empty()) err_ = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
    return false;
}

template <typename T>
bool Solver::alloc(DBuf<T> &b, size_t n) {
    b.release();
    b.n = n;
    if (n == 0) return true;
    return check(hipMalloc((void **)&b.p, n * sizeof(T)), "hipMalloc");
}

template <typename T>
bool Solver::upload(DBuf<T> &b, const std::vector<T> &h) {
    if (!alloc(b, h.size())) return false;
    if (h.empty()) return true;
    return check(hipMemcpy(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

Solver::Solver(const Problem &p, const gbte_options &o) : P(p), opt(o) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        err_ = "no HIP device available";
        return;
    }
    if (!check(hipSetDevice(opt.device), "hipSetDevice")) return;
    if (!check(hipStreamCreate(&s_), "hipStreamCreate")) return;
    if (!check(hipEventCreate(&ev0_), "hipEventCreate") || !check(hipEventCreate(&ev1_), "hipEventCreate") ||
        !check(hipEventCreate(&ev2_), "hipEventCreate"))
        return;
    if (!check(hipHostMalloc((void **)&h_pin, 64 * sizeof(double)), "hipHostMalloc")) return;
    build();
}

Solver::~Solver() {
    if (h_pin) (void)hipHostFree(h_pin);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (ev2_) (void)hipEventDestroy(ev2_);
    if (s_) (void)hipStreamDestroy(s_);
}

bool Solver::build_matrix(const std::vector<std::vector<std::pair<int, double>>> &rows, DevCSR &m) {
    int n = (int)rows.size();
    std::vector<int> ptr(n + 1, 0), col;
    std::vector<double> val, dinv(n, 0.0);
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
    return upload(m.ptr, ptr) && upload(m.col, col) && upload(m.val, val) && upload(m.dinv, dinv);
}

bool Solver::build() {
    const Problem &p = P;
    nc = p.n_cells; nb = p.n_bounds; ne = 2 * nb; nn = p.n_nodes; nB = p.n_bands; nD = p.n_dirs; nf = p.n_faces;
    dim = p.dim;
    if (p.n_matters != 1) { err_ = "only single-material problems are supported"; return false; }
    if (dim != 2) { err_ = "only 2D meshes are supported by this solver build"; return false; }
    if (nB < 1 || nB > 64) { err_ = "the number of bands must be between 1 and 64"; return false; }
    for (int ib = 0; ib < nb; ++ib)
        if (p.bound_type[ib] == 4 || p.bound_type[ib] == 5) { err_ = "interface boundaries (type 4/5) are not supported"; return false; }
    cpw = 64 / nB;
    // ---- modes
    std::vector<double> sx(p.dir_x), sy(p.dir_y), sz(p.dir_z), w(p.weight);
    std::vector<double> vg(p.vg.begin(), p.vg.begin() + nB), tau(p.tau.begin(), p.tau.begin() + nB),
        cap(p.cap.begin(), p.cap.begin() + nB), lr(p.lr.begin(), p.lr.begin() + nB), hr(p.hr.begin(), p.hr.begin() + nB);
    for (int b = 0; b < nB; ++b)
        if (!(vg[b] > 0)) { err_ = "group velocities must be positive"; return false; }
    if (!upload(d_vg, vg) || !upload(d_tau, tau) || !upload(d_cap, cap) || !upload(d_lr, lr) || !upload(d_hr, hr) ||
        !upload(d_sx, sx) || !upload(d_sy, sy) || !upload(d_sz, sz) || !upload(d_w, w))
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
    if (!upload(d_fsize, p.face_size) || !upload(d_snb, snb) || !upload(d_sent, sent) || !upload(d_vol, p.volume) ||
        !upload(d_Q, p.heat_source) || !upload(d_snx, p.face_nx) || !upload(d_sny, p.face_ny) || !upload(d_snz, p.face_nz) ||
        !upload(d_sarea, p.face_area) || !upload(d_srx, srx) || !upload(d_sry, sry) || !upload(d_srz, srz) ||
        !upload(d_sqx, sqx) || !upload(d_sqy, sqy) || !upload(d_sqz, sqz))
        return false;
    // ---- per-direction upwind levels (band 0 decides; band signs agree
    // except for faces parallel to a direction to rounding, whose coupling
    // coefficient vanishes)
    {
        std::vector<std::vector<int>> lev_c, lev_d;
        std::vector<int> level(nc), indeg(nc), q;
        double v0 = vg[0];
        for (int d = 0; d < nD; ++d) {
            auto dotf = [&](size_t s) { return v0 * sx[d] * p.face_nx[s] + v0 * sy[d] * p.face_ny[s] + v0 * sz[d] * p.face_nz[s]; };
            std::fill(indeg.begin(), indeg.end(), 0);
            for (int c = 0; c < nc; ++c)
                for (int j = 0; j < p.face_size[c]; ++j) {
                    size_t s = (size_t)c * 6 + j;
                    if (snb[s] >= 0 && dotf(s) < 0) indeg[c]++;
                }
            q.clear();
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
                err_ = "the upwind graph of direction " + std::to_string(d) + " has a cycle; an acyclic (e.g. Delaunay) mesh is required";
                return false;
            }
            for (int c = 0; c < nc; ++c) {
                if ((int)lev_c.size() <= level[c]) { lev_c.resize(level[c] + 1); lev_d.resize(level[c] + 1); }
                lev_c[level[c]].push_back(c);
                lev_d[level[c]].push_back(d);
            }
        }
        nlev = (int)lev_c.size();
        std::vector<int> ic, id;
        lev_ptr.assign(1, 0);
        for (int l = 0; l < nlev; ++l) {
            ic.insert(ic.end(), lev_c[l].begin(), lev_c[l].end());
            id.insert(id.end(), lev_d[l].begin(), lev_d[l].end());
            lev_ptr.push_back((int)ic.size());
        }
        if (!upload(d_item_c, ic) || !upload(d_item_d, id)) return false;
    }
    // ---- LSQ gradient stencil (_get_CellMatrix_larger + Get_inverse_2D,
    // StaticBTESynthetic.cpp:1000-1064, utility.cpp:401-466)
    {
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
        if (!upload(d_gptr, p.nbr_ptr) || !upload(d_gidx, p.nbr_idx) || !upload(d_gP, gP) || !upload(d_gok, gok)) return false;
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
        std::vector<double> vdist, tdist;
        for (int v = 0; v < nn; ++v) {
            for (auto &e : ve[v]) { vsrc.push_back(e.first); vdist.push_back(e.second); }
            vptr.push_back((int)vsrc.size());
            for (auto &e : vt[v]) { tsrc.push_back(e.first); tdist.push_back(e.second); }
            tptr.push_back((int)tsrc.size());
        }
        if (!upload(d_vptr, vptr) || !upload(d_vsrc, vsrc) || !upload(d_vdist, vdist) || !upload(d_tptr, tptr) ||
            !upload(d_tsrc, tsrc) || !upload(d_tdist, tdist))
            return false;
    }
    // ---- face-gradient slot data
    {
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
        if (!upload(d_fkind, kind) || !upload(d_fother, other) || !upload(d_fv0, fv0) || !upload(d_fv1, fv1) ||
            !upload(d_fan, an) || !upload(d_faa, aa) || !upload(d_fn0x, n0x) || !upload(d_fn0y, n0y) || !upload(d_fn0z, n0z) ||
            !upload(d_fmx, mx) || !upload(d_fmy, my) || !upload(d_fm2, m2))
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
                if (ty == 1) btw[en] = p.bound_temp[ib];
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
            !upload(d_be_nz, bnz) || !upload(d_be_tw, btw) || !upload(d_be_dT, bdT) || !upload(d_be_tov, btov))
            return false;
        // slot index per entry for k_bound_out
        DBuf<int> tmp;
        (void)tmp;
    }
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
        if (!build_matrix(rowsM, Kmacro) || !build_matrix(rowsF, Kfour)) return false;
    }
    // ---- state
    size_t nE = (size_t)nD * nc * nB;
    if (!alloc(d_e, nE) || !alloc(d_grad, nE * dim) || !alloc(d_eb, (size_t)nD * ne * nB) ||
        !alloc(d_eV, (size_t)nD * nn * nB) || !alloc(d_T, nc) || !alloc(d_Told, nc) || !alloc(d_T1, nc) ||
        !alloc(d_TV, nn) || !alloc(d_E, nc) || !alloc(d_qx, nc) || !alloc(d_qy, nc) || !alloc(d_qz, nc) ||
        !alloc(d_qxo, nc) || !alloc(d_qyo, nc) || !alloc(d_qzo, nc) || !alloc(d_RM, nc) || !alloc(d_Tb, (size_t)nB * nc) ||
        !alloc(d_tmac, ne) || !alloc(d_rhs, nc) || !alloc(d_r, nc) || !alloc(d_z, nc) || !alloc(d_p, nc) ||
        !alloc(d_Ap, nc))
        return false;
    red_blocks = blocks_for(nc, kBlock);
    if (!alloc(d_part, (size_t)red_blocks * 4 + 8) || !alloc(d_part2, (size_t)red_blocks * 4 + 8)) return false;
    if (!check(hipMemsetAsync(d_eb.p, 0, d_eb.n * sizeof(double), s_), "hipMemsetAsync") ||
        !check(hipMemsetAsync(d_qz.p, 0, nc * sizeof(double), s_), "hipMemsetAsync"))
        return false;
    // initial temperature and flux (Solution.cpp:78-90)
    if (!check(hipMemcpy(d_T.p, p.initial_temp.data(), nc * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy") ||
        !check(hipMemsetAsync(d_qx.p, 0, nc * sizeof(double), s_), "hipMemsetAsync") ||
        !check(hipMemsetAsync(d_qy.p, 0, nc * sizeof(double), s_), "hipMemsetAsync"))
        return false;
    // entry -> slot table for k_bound_out lives in d_be_slot (reuse d_fother layout is not possible)
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
    // StaticFourier::_solve (StaticFourier.cpp:934-976) from T = initial
    double err = 1000;
    int passes = 0;
    std::vector<int> nint(nc, 0);
    for (int c = 0; c < nc; ++c)
        for (int j = 0; j < P.face_size[c]; ++j)
            if (P.face_bound[(size_t)c * 6 + j] == -1) nint[c]++;
    DBuf<int> d_nint;
    if (!upload(d_nint, nint)) return GBTE_HIP_ERROR;
    if (!check(hipMemcpy(d_T.p, P.initial_temp.data(), nc * sizeof(double), hipMemcpyHostToDevice), "H2D"))
        return GBTE_HIP_ERROR;
    while (err > 1e-5) {
        hipLaunchKernelGGL(k_copy, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_T1.p, nc);
        launch_vertex_T(d_T.p, d_TV.p);
        launch_rhs(false, d_fourier_bval.p, d_rhs.p);
        if (!cg(Kfour, d_rhs.p, d_T.p, nullptr)) return GBTE_HIP_ERROR;
        hipLaunchKernelGGL(k_fourier_flux, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_gptr.p, d_gidx.p, d_gP.p,
                           d_gok.p, d_nint.p, P.nbr_ptr[nc], 1.0 / P.deck.Lx, P.kappa_bulk[0], d_qx.p, d_qy.p, d_qz.p,
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
    const double *src[4] = {T, qx, qy, qz};
    double *dst[4] = {d_T.p, d_qx.p, d_qy.p, d_qz.p};
    for (int k = 0; k < 4; ++k)
        if (src[k] && !check(hipMemcpy(dst[k], src[k], nc * sizeof(double), hipMemcpyHostToDevice), "H2D"))
            return GBTE_HIP_ERROR;
    return GBTE_OK;
}

int Solver::solve(int max_iter, double tolT, double tolQ, gbte_stats *st) {
    if (!ok()) return GBTE_HIP_ERROR;
    auto t0 = std::chrono::high_resolution_clock::now();
    Modes m{d_vg.p, d_tau.p, d_cap.p, d_lr.p, d_hr.p, d_sx.p, d_sy.p, d_sz.p, d_w.p, P.kappa_bulk[0], P.capacity_bulk[0]};
    SlotGeo g{d_fsize.p, d_snb.p, d_sent.p, d_snx.p, d_sny.p, d_snz.p, d_sarea.p, d_vol.p,
              d_srx.p, d_sry.p, d_srz.p, d_sqx.p, d_sqy.p, d_sqz.p};
    FaceK fk{d_fkind.p, d_fother.p, d_fv0.p, d_fv1.p, d_fan.p, d_faa.p, d_fn0x.p, d_fn0y.p, d_fn0z.p,
             d_fmx.p, d_fmy.p, d_fm2.p};
    int nnz = P.nbr_ptr[nc];
    double invLx = 1.0 / P.deck.Lx;
    int n_ent = (int)h_ent_list.size();
    std::vector<int> be_slot(ne, -1);
    for (int ib = 0; ib < nb; ++ib)
        for (int ic = 0; ic < 2; ++ic) {
            int c = P.bound_cell[ib * 2 + ic];
            if (c >= 0) be_slot[ib * 2 + ic] = c * 6 + P.bound_face[ib * 2 + ic];
        }
    DBuf<int> d_be_slot;
    if (!upload(d_be_slot, be_slot)) return GBTE_HIP_ERROR;
    size_t nE = (size_t)nD * nc * nB;
    // _set_initial, gradients, outgoing and incoming boundary values
    // (StaticBTESyntheticSolver.cpp:19-62)
    hipLaunchKernelGGL(k_init_e, dim3(blocks_for((long)nE, kBlock)), dim3(kBlock), 0, s_, d_e.p, d_T.p, d_cap.p, nc, nB, nE);
    int gw = blocks_for(blocks_for((long)nc * nD, cpw), kWaves);
    hipLaunchKernelGGL(k_gradient, dim3(gw), dim3(kBlock), 0, s_, d_e.p, d_grad.p, d_gptr.p, d_gidx.p, d_gP.p, d_gok.p,
                       dim, nnz, invLx, nc, nB, nD, cpw);
    int bw = blocks_for(blocks_for((long)n_ent * nD, cpw), kWaves);
    hipLaunchKernelGGL(k_bound_out, dim3(std::max(bw, 1)), dim3(kBlock), 0, s_, d_e.p, d_grad.p, d_eb.p, g, m,
                       d_ent_list.p, d_be_cell.p, d_be_slot.p, n_ent, nc, ne, nB, nD, dim, cpw);
    int bbw = blocks_for(blocks_for(n_ent, cpw), kWaves);
    hipLaunchKernelGGL(k_boundary, dim3(std::max(bbw, 1)), dim3(kBlock), 0, s_, d_eb.p, d_tmac.p, m, d_ent_list.p,
                       d_be_type.p, d_be_conn.p, d_be_spec.p, d_be_nx.p, d_be_ny.p, d_be_nz.p, d_be_tw.p, d_be_dT.p,
                       d_be_tov.p, n_ent, ne, nB, nD, cpw);
    if (!check(hipGetLastError(), "setup kernels")) return GBTE_HIP_ERROR;
    errInc_ = 0;
    errTold_ = 0;
    errQold_ = 0;
    errInner_ = 0;
    int nt = 0, status = GBTE_OK, macro_total = 0;
    bool conv = false;
    double errT = 0, errQ = 0;
    double sweep_ms = 0, transport_ms = 0, macro_ms = 0;
    long cg_total = 0;
    int fmw = blocks_for(blocks_for(nc, cpw), kWaves);
    int vw = blocks_for(blocks_for((long)nn * nD, cpw), kWaves);
    for (nt = 0; nt < max_iter; ++nt) {
        hipLaunchKernelGGL(k_copy_old, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_qx.p, d_qy.p, d_qz.p, d_Told.p,
                           d_qxo.p, d_qyo.p, d_qzo.p, nc);
        (void)hipEventRecord(ev0_, s_);
        for (int l = 0; l < nlev; ++l) {
            int a = lev_ptr[l], n_items = lev_ptr[l + 1] - a;
            int blocks = blocks_for(blocks_for(n_items, cpw), kWaves);
            hipLaunchKernelGGL(k_sweep, dim3(blocks), dim3(kBlock), 0, s_, d_item_c.p + a, d_item_d.p + a, n_items,
                               d_e.p, d_grad.p, d_eb.p, d_Told.p, d_Q.p, g, m, nc, ne, nB, dim, cpw);
        }
        (void)hipEventRecord(ev1_, s_);
        hipLaunchKernelGGL(k_vertex_e, dim3(vw), dim3(kBlock), 0, s_, d_e.p, d_eb.p, d_eV.p, d_vptr.p, d_vsrc.p,
                           d_vdist.p, nn, nc, ne, nB, nD, cpw);
        hipLaunchKernelGGL(k_face_moments, dim3(fmw), dim3(kBlock), 0, s_, d_e.p, d_eb.p, d_eV.p, d_grad.p, g, fk, m,
                           d_gptr.p, d_gidx.p, d_gP.p, d_gok.p, nnz, invLx, d_T.p, d_E.p, d_qx.p, d_qy.p, d_qz.p,
                           d_RM.p, d_Tb.p, nc, ne, nn, nB, nD, dim, cpw);
        (void)hipEventRecord(ev2_, s_);
        hipLaunchKernelGGL(k_boundary, dim3(std::max(bbw, 1)), dim3(kBlock), 0, s_, d_eb.p, d_tmac.p, m,
                           d_ent_list.p, d_be_type.p, d_be_conn.p, d_be_spec.p, d_be_nx.p, d_be_ny.p, d_be_nz.p,
                           d_be_tw.p, d_be_dT.p, d_be_tov.p, n_ent, ne, nB, nD, cpw);
        if (!check(hipGetLastError(), "iteration kernels")) return GBTE_HIP_ERROR;
        // _marco_solution (StaticBTESynthetic.cpp:2692-2734)
        auto tm0 = std::chrono::high_resolution_clock::now();
        double inner = 1000;
        int inc = 0, passes = 0;
        while (inner > 1e-6) {
            hipLaunchKernelGGL(k_copy, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_T1.p, nc);
            launch_vertex_T(d_T.p, d_TV.p);
            launch_rhs(true, d_tmac.p, d_rhs.p);
            int cgit = 0;
            if (!cg(Kmacro, d_rhs.p, d_T.p, &cgit)) return GBTE_HIP_ERROR;
            cg_total += cgit;
            if (!inner_error(d_T1.p, d_T.p, &inner)) return GBTE_HIP_ERROR;
            if (inner > errInner_) inc++;
            errInner_ = inner;
            passes++;
            if (opt.verbose) printf("internal residual: %g\n", inner);
            if (inc >= 10) { err_ = "cannot congerve, try to set Order=1 or 2limit"; status = GBTE_DIVERGED; break; }
        }
        macro_total += passes;
        macro_ms += std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - tm0).count();
        float ms = 0;
        (void)hipEventElapsedTime(&ms, ev0_, ev1_);
        sweep_ms += ms;
        (void)hipEventElapsedTime(&ms, ev0_, ev2_);
        transport_ms += ms;
        if (status != GBTE_OK) { ++nt; break; }
        // _get_magin_check_error
        hipLaunchKernelGGL(k_err, dim3(red_blocks), dim3(kBlock), 0, s_, d_T.p, d_Told.p, d_qx.p, d_qy.p, d_qz.p,
                           d_qxo.p, d_qyo.p, d_qzo.p, d_part.p, nc, red_blocks);
        std::vector<double> h;
        if (!reduce_host(d_part.p, red_blocks, h, 4)) return GBTE_HIP_ERROR;
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
        st->sweep_levels = nlev;
        st->macro_cg_iters = (int)cg_total;
    }
    return status;
}

int Solver::get_fields(double *TL, double *TE, double *qx, double *qy, double *qz, double *Tb) {
    if (!ok()) return GBTE_HIP_ERROR;
    struct {
        double *h;
        const double *d;
        size_t n;
    } f[6] = {{TL, d_T.p, (size_t)nc}, {TE, d_E.p, (size_t)nc}, {qx, d_qx.p, (size_t)nc},
              {qy, d_qy.p, (size_t)nc}, {qz, d_qz.p, (size_t)nc}, {Tb, d_Tb.p, (size_t)nB * nc}};
    for (auto &x : f)
        if (x.h && !check(hipMemcpy(x.h, x.d, x.n * sizeof(double), hipMemcpyDeviceToHost), "D2H")) return GBTE_HIP_ERROR;
    return GBTE_OK;
}

int Solver::keff(int axis, double dT, double *out) {
    if (!ok()) return GBTE_HIP_ERROR;
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
