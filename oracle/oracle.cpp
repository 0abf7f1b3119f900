// ORACLE - test infrastructure only.  Not part of the product: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load this file's
// library, as the checker.  The product (libgiftbte_hip.so) never links it.
//
// Plain FP64 CPU restatement, at np=1, of the reference's steady path:
//   StaticFourier::_solve                 src/StaticFourier/StaticFourier.cpp:934-976
//   StaticBTESynthetic::solve             src/StaticBTESynthetic/StaticBTESyntheticSolver.cpp:12-247
//   SolutionAll glue                      src/Solution/Solution.cpp:78-124, 236-414
// It consumes the flattened problem tables (include/giftbte.h gbte_tables),
// which the test suite pins entry by entry against the compiled reference.
// Each function cites the reference routine it restates and keeps its loop
// order and operation order, with two deliberate replacements:
//   * the per-system Eigen SparseLU solve (StaticBTESyntheticSolver.cpp:96)
//     becomes an exact sweep in upwind topological order (the upwind matrix
//     is a permuted triangular matrix when the direction graph is acyclic;
//     a cycle is reported as an error);
//   * the macro / Fourier SparseLU (StaticBTESynthetic.cpp:881-883,2705;
//     StaticFourier.cpp:515-517,954) becomes a banded LU without pivoting on
//     a reverse Cuthill-McKee ordering (the matrices are diagonally dominant
//     M-matrices).
// Both are exact up to rounding.
//
// Parity pins (see tests/): the compiled reference's full-precision dumps
// (oracle/_ref/ref_dump) for whole runs and for every per-system stage of
// iteration 0; closed forms (LSQ gradients of linear fields, quadrature
// identities, equilibrium T_hot = T_cold, linearity in dT, mirror symmetry).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "giftbte.h"

namespace {

using std::vector;

thread_local std::string g_err;

// ------------------------------------------------------------ banded LU
struct Banded {
    int n = 0, bw = 0;                 // half bandwidth in the permuted order
    vector<int> perm, iperm;           // perm[new] = old
    vector<double> a;                  // row-major, (2bw+1) per row, column offset j-i+bw
    double &at(int i, int j) { return a[(size_t)i * (2 * bw + 1) + (j - i + bw)]; }
};

// Reverse Cuthill-McKee on the symmetric pattern of rows.
vector<int> rcm(int n, const vector<vector<std::pair<int, double>>> &rows) {
    vector<int> deg(n), order, mark(n, 0);
    for (int i = 0; i < n; ++i) deg[i] = (int)rows[i].size();
    for (int s = 0; s < n; ++s) {
        if (mark[s]) continue;
        int start = s;
        for (int i = s; i < n; ++i)
            if (!mark[i] && deg[i] < deg[start]) start = i;
        std::deque<int> q;
        q.push_back(start);
        mark[start] = 1;
        while (!q.empty()) {
            int u = q.front();
            q.pop_front();
            order.push_back(u);
            vector<int> nb;
            for (auto &e : rows[u])
                if (!mark[e.first]) { nb.push_back(e.first); mark[e.first] = 1; }
            std::sort(nb.begin(), nb.end(), [&](int x, int y) { return deg[x] < deg[y] || (deg[x] == deg[y] && x < y); });
            for (int v : nb) q.push_back(v);
        }
    }
    std::reverse(order.begin(), order.end());
    return order;
}

bool banded_factor(const vector<vector<std::pair<int, double>>> &rows, Banded &B) {
    int n = (int)rows.size();
    B.n = n;
    B.perm = rcm(n, rows);
    B.iperm.assign(n, 0);
    for (int i = 0; i < n; ++i) B.iperm[B.perm[i]] = i;
    int bw = 0;
    for (int i = 0; i < n; ++i)
        for (auto &e : rows[i]) bw = std::max(bw, std::abs(B.iperm[i] - B.iperm[e.first]));
    B.bw = bw;
    B.a.assign((size_t)n * (2 * bw + 1), 0.0);
    for (int i = 0; i < n; ++i)
        for (auto &e : rows[i]) B.at(B.iperm[i], B.iperm[e.first]) += e.second;
    for (int k = 0; k < n; ++k) {
        double piv = B.at(k, k);
        if (piv == 0 || !std::isfinite(piv)) return false;
        int hi = std::min(n - 1, k + bw);
        for (int i = k + 1; i <= hi; ++i) {
            double l = B.at(i, k) / piv;
            if (l == 0) continue;
            B.at(i, k) = l;
            for (int j = k + 1; j <= hi; ++j) B.at(i, j) -= l * B.at(k, j);
        }
    }
    return true;
}

void banded_solve(Banded &B, const vector<double> &rhs, vector<double> &x) {
    int n = B.n, bw = B.bw;
    vector<double> y(n);
    for (int i = 0; i < n; ++i) y[i] = rhs[B.perm[i]];
    for (int i = 0; i < n; ++i) {
        double s = y[i];
        for (int j = std::max(0, i - bw); j < i; ++j) s -= B.at(i, j) * y[j];
        y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j <= std::min(n - 1, i + bw); ++j) s -= B.at(i, j) * y[j];
        y[i] = s / B.at(i, i);
    }
    x.assign(n, 0.0);
    for (int i = 0; i < n; ++i) x[B.perm[i]] = y[i];
}

// ------------------------------------------------------------ LSQ helpers
// (J^T J)^{-1} J^T for the vertex-neighbour stencil: Get_inverse_2D /
// Get_inverse_3D, src/utility/utility.cpp:401-548.
void lsq_2d(const vector<vector<double>> &J, vector<vector<double>> &P) {
    size_t m = J.size();
    double a00 = 0, a01 = 0, a10 = 0, a11 = 0;
    for (size_t k = 0; k < m; ++k) a00 += J[k][0] * J[k][0];
    for (size_t k = 0; k < m; ++k) a01 += J[k][0] * J[k][1];
    for (size_t k = 0; k < m; ++k) a10 += J[k][1] * J[k][0];
    for (size_t k = 0; k < m; ++k) a11 += J[k][1] * J[k][1];
    double det = a00 * a11 - a01 * a10;
    double i00 = a11 / det, i01 = -a01 / det, i11 = a00 / det, i10 = -a10 / det;
    P.assign(2, vector<double>(m, 0.0));
    for (size_t j = 0; j < m; ++j) {
        P[0][j] += i00 * J[j][0];
        P[0][j] += i01 * J[j][1];
        P[1][j] += i10 * J[j][0];
        P[1][j] += i11 * J[j][1];
    }
}

void lsq_3d(const vector<vector<double>> &J, vector<vector<double>> &P) {
    size_t m = J.size();
    double A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            A[i][j] = 0;
            for (size_t k = 0; k < m; ++k) A[i][j] += J[k][i] * J[k][j];
        }
    double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) + A[0][1] * (A[1][2] * A[2][0] - A[1][0] * A[2][2]) +
                 A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    double I[3][3];
    I[0][0] = 1.0 / det * (A[1][1] * A[2][2] - A[1][2] * A[2][1]);
    I[1][0] = 1.0 / det * (A[0][2] * A[2][1] - A[0][1] * A[2][2]);
    I[2][0] = 1.0 / det * (A[0][1] * A[1][2] - A[0][2] * A[1][1]);
    I[0][1] = 1.0 / det * (A[1][2] * A[2][0] - A[1][0] * A[2][2]);
    I[1][1] = 1.0 / det * (A[0][0] * A[2][2] - A[0][2] * A[2][0]);
    I[2][1] = 1.0 / det * (A[1][0] * A[0][2] - A[0][0] * A[1][2]);
    I[0][2] = 1.0 / det * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
    I[1][2] = 1.0 / det * (A[0][1] * A[2][0] - A[0][0] * A[2][1]);
    I[2][2] = 1.0 / det * (A[0][0] * A[1][1] - A[1][0] * A[0][1]);
    P.assign(3, vector<double>(m, 0.0));
    for (int i = 0; i < 3; ++i)
        for (size_t j = 0; j < m; ++j)
            for (int k = 0; k < 3; ++k) P[i][j] += I[i][k] * J[j][k];
}

// ------------------------------------------------------------ the solver
struct Oracle {
    const gbte_tables &t;
    int nc, nb, nf, nn, nB, nD, nm, dim;
    double Lx;
    vector<vector<vector<double>>> cellMat;  // [cell][dim][nbr]
    vector<int> cmSet;                       // CellMatrix assigned
    vector<vector<vector<double>>> faceMat;  // 3D: [face][3][face vertices]
    // BTE state
    vector<double> e;                        // [band][dir][cell]
    vector<double> ebound;                   // [band][dir][bound][2]
    vector<double> eboundLocal;              // [band][bound][2] (np=1: numBandLocal = nB)
    vector<double> T, Told, Tband, E, qx, qy, qz, qxo, qyo, qzo, ReMacro, gradX, gradY, gradZ, eV, TV, T1;
    vector<double> gfx, gfy, gfz;            // [face][2]
    vector<vector<double>> tmac;             // [bound][2]
    double errT = 0, errQ = 0, errTold = 0, errQold = 0, errInner = 0;
    int errInc = 0;
    Banded Kmacro;
    int macroPasses = 0;

    bool first = false;    // Order 1
    bool limiter = false;  // Order 2limit
    vector<double> lim;    // limiter factor per cell of the current system
    explicit Oracle(const gbte_tables &tt) : t(tt) {
        nc = t.n_cells; nb = t.n_bounds; nf = t.n_faces; nn = t.n_nodes; nB = t.n_bands; nD = t.n_dirs;
        nm = t.n_matters; dim = t.dim; Lx = t.Lx;
        lim.assign(nc, 1.0);
    }
    size_t slot(int i, int j) const { return (size_t)i * 6 + j; }
    double vgx(int m, int b, int d) const { return t.vg[(size_t)m * nB + b] * t.dir_x[d]; }
    double vgy(int m, int b, int d) const { return t.vg[(size_t)m * nB + b] * t.dir_y[d]; }
    double vgz(int m, int b, int d) const { return t.vg[(size_t)m * nB + b] * t.dir_z[d]; }
    double tau(int m, int b) const { return t.tau[(size_t)m * nB + b]; }
    double cap(int m, int b) const { return t.cap[(size_t)m * nB + b]; }
    double lr(int m, int b) const { return t.lr[(size_t)m * nB + b]; }
    double hr(int m, int b) const { return t.hr[(size_t)m * nB + b]; }
    double w(int d) const { return t.weight[d]; }
    double *E_(int b, int d) { return &e[((size_t)b * nD + d) * nc]; }
    double &EB(int b, int d, int ib, int ic) { return ebound[(((size_t)b * nD + d) * nb + ib) * 2 + ic]; }
    double dotvn(int m, int b, int d, size_t s) const {
        return vgx(m, b, d) * t.face_nx[s] + vgy(m, b, d) * t.face_ny[s] + vgz(m, b, d) * t.face_nz[s];
    }

    // _get_CellMatrix_larger (StaticBTESynthetic.cpp:1000-1064) /
    // StaticFourier::_get_CellMatrix: same stencil, same matrices.
    void cell_matrix() {
        cellMat.assign(nc, {});
        cmSet.assign(nc, 0);
        for (int i = 0; i < nc; ++i) {
            int a = t.nbr_ptr[i], b = t.nbr_ptr[i + 1];
            vector<vector<double>> J(b - a, vector<double>(dim));
            for (int k = a; k < b; ++k) {
                int j = t.nbr_idx[k];
                J[k - a][0] = (1.0 / Lx * (t.cx[j] - t.cx[i]));
                J[k - a][1] = (1.0 / Lx * (t.cy[j] - t.cy[i]));
                if (dim == 3) J[k - a][2] = (1.0 / Lx * (t.cz[j] - t.cz[i]));
            }
            if (dim == 2 && b - a > 2) { lsq_2d(J, cellMat[i]); cmSet[i] = 1; }
            if (dim == 3 && b - a > 3) { lsq_3d(J, cellMat[i]); cmSet[i] = 1; }
        }
    }

    // IDW of a cell field to the nodes (StaticFourier.cpp:280-309 and
    // StaticBTESynthetic.cpp:2276-2305; one material).
    void vertex_from_cells(const vector<double> &f, vector<double> &out) {
        vector<double> tmp((size_t)nn * nm, 0.0), R((size_t)nn * nm, 0.0);
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < t.cell_nvert[i]; ++j) {
                int v = t.cell_vert[(size_t)i * 8 + j];
                double d = std::sqrt(std::pow(t.node_x[v] - t.cx[i], 2) + std::pow(t.node_y[v] - t.cy[i], 2) +
                                     std::pow(t.node_z[v] - t.cz[i], 2));
                tmp[(size_t)v * nm + t.matter[i]] += f[i] / d;
                R[(size_t)v * nm + t.matter[i]] += 1.0 / d;
            }
        out.assign((size_t)nn * nm, 0.0);
        for (size_t k = 0; k < out.size(); ++k) out[k] = tmp[k] / R[k];
    }

    // _get_face_matrix (StaticBTESynthetic.cpp:1617-1649), 3D: per unique
    // face, from its first side, (J^T J)^{-1} J^T of the rows [n; (x_{m+1} -
    // x_m) / L_x] (Get_inverse_3D), so that FaceMatrix . [0; dv / L_x] is the
    // tangential gradient of a vertex field.
    void face_matrix() {
        faceMat.assign(nf, {});
        if (dim != 3) return;
        for (int f = 0; f < nf; ++f) {
            int i = t.fface_cell[f * 2], j = t.fface_face[f * 2];
            size_t s = slot(i, j);
            int nv = t.face_nvert[s];
            vector<vector<double>> J(nv, vector<double>(3));
            J[0][0] = t.face_nx[s]; J[0][1] = t.face_ny[s]; J[0][2] = t.face_nz[s];
            for (int m = 0; m < nv - 1; ++m) {
                int a = t.face_vert[s * 4 + m], b = t.face_vert[s * 4 + m + 1];
                J[m + 1][0] = (1.0 / Lx * (t.node_x[b] - t.node_x[a]));
                J[m + 1][1] = (1.0 / Lx * (t.node_y[b] - t.node_y[a]));
                J[m + 1][2] = (1.0 / Lx * (t.node_z[b] - t.node_z[a]));
            }
            lsq_3d(J, faceMat[f]);
        }
    }
    // FaceMatrix[f] . d1 with d1 = [0, (v[m+1] - v[m]) / L_x], the vertices of
    // slot s in their order, vertex values vf[vertex * stride + off]
    void tangential(int f, size_t s, const vector<double> &vf, int stride, int off, double g[3]) const {
        int nv = t.face_nvert[s];
        vector<double> d1(nv, 0.0);
        for (int m = 0; m < nv - 1; ++m) {
            int a = t.face_vert[s * 4 + m], b = t.face_vert[s * 4 + m + 1];
            d1[m + 1] = 1.0 / Lx * (vf[(size_t)b * stride + off] - vf[(size_t)a * stride + off]);
        }
        g[0] = 0; g[1] = 0; g[2] = 0;
        for (int m = 0; m < nv; ++m) {
            g[0] += faceMat[f][0][m] * d1[m];
            g[1] += faceMat[f][1][m] * d1[m];
            g[2] += faceMat[f][2][m] * d1[m];
        }
    }

    // ------------------------------------------------ Fourier warm start
    int fourier(vector<double> &TF, vector<double> &fqx, vector<double> &fqy, vector<double> &fqz) {
        // boundaryTemp (Solution.cpp:96-118) as handed to StaticFourier
        vector<double> btemp(nb, 0.0);
        for (int ib = 0; ib < nb; ++ib) {
            if (t.bound_type[ib] == 1) btemp[ib] = t.bound_temp[ib];
            else if (t.bound_type[ib] < 0 && t.bound_temp[ib] != 0) {
                btemp[ib] = t.bound_temp[ib] / 2;
                btemp[t.bound_connect[ib]] = -t.bound_temp[ib] / 2;
            }
        }
        auto kap = [&](int i) { return t.kappa_bulk[t.matter[i]]; };
        vector<double> resis(nb, 0.0);  // StaticFourier.cpp:215-226
        for (int ib = 0; ib < nb; ++ib)
            if (t.bound_type[ib] == 5) { g_err = "interface (type 5) boundaries are not in the oracle"; return -1; }
        // _get_coefficient_macro (StaticFourier.cpp:449-518)
        vector<vector<std::pair<int, double>>> rows(nc);
        for (int ie = 0; ie < nc; ++ie) {
            double diag = 0;
            for (int j = 0; j < t.face_size[ie]; ++j) {
                size_t s = slot(ie, j);
                int fb = t.face_bound[s];
                if (fb == -1) {
                    int nbr = t.face_neighbor[s];
                    double tmp = 2 * (kap(ie) * kap(nbr)) / (kap(ie) + kap(nbr)) * t.face_area[s];
                    double ax = t.cx[nbr] - t.cx[ie], ay = t.cy[nbr] - t.cy[ie], az = t.cz[nbr] - t.cz[ie];
                    tmp = tmp / (t.face_nx[s] * ax + t.face_ny[s] * ay + t.face_nz[s] * az);
                    diag += -tmp;
                    rows[ie].push_back({nbr, tmp});
                } else if (t.bound_type[fb] == 1) {
                    double tmp = kap(ie) * t.face_area[s];
                    double ax = t.face_cx[s] - t.cx[ie], ay = t.face_cy[s] - t.cy[ie], az = t.face_cz[s] - t.cz[ie];
                    tmp = tmp / (t.face_nx[s] * ax + t.face_ny[s] * ay + t.face_nz[s] * az);
                    diag += -tmp;
                } else if (t.bound_type[fb] < 0) {
                    if (t.bound_temp[fb] != 0 || t.bound_temp[t.bound_connect[fb]] != 0) {
                        double tmp = kap(ie) * t.face_area[s];
                        double ax = t.face_cx[s] - t.cx[ie], ay = t.face_cy[s] - t.cy[ie], az = t.face_cz[s] - t.cz[ie];
                        tmp = tmp / (t.face_nx[s] * ax + t.face_ny[s] * ay + t.face_nz[s] * az);
                        diag += -tmp;
                    }
                }
            }
            rows[ie].push_back({ie, diag});
        }
        Banded K;
        if (!banded_factor(rows, K)) { g_err = "Fourier matrix is singular"; return -1; }
        cell_matrix();
        TF.assign(t.initial_temp, t.initial_temp + nc);
        fqx.assign(nc, 0.0); fqy.assign(nc, 0.0); fqz.assign(nc, 0.0);
        vector<double> Tv, Re(nc), Tn;
        double err = 1000;
        int passes = 0;
        while (err > 1e-5) {
            vector<double> Tprev = TF;
            vertex_from_cells(TF, Tv);
            // _get_Re (StaticFourier.cpp:520-613)
            for (int i = 0; i < nc; ++i) Re[i] = 0;
            for (int i = 0; i < nc; ++i) Re[i] -= t.heat_source[i] * t.volume[i];
            for (int i = 0; i < nc; ++i)
                for (int j = 0; j < t.face_size[i]; ++j) {
                    size_t s = slot(i, j);
                    int fb = t.face_bound[s];
                    double ax = 0, ay = 0, az = 0;
                    if (fb >= 0) {
                        if (t.bound_type[fb] == 1 ||
                            (t.bound_type[fb] < 0 && (t.bound_temp[fb] != 0 || t.bound_temp[t.bound_connect[fb]] != 0))) {
                            double tmp = kap(i) * t.face_area[s];
                            ax = t.face_cx[s] - t.cx[i]; ay = t.face_cy[s] - t.cy[i]; az = t.face_cz[s] - t.cz[i];
                            tmp = tmp / (t.face_nx[s] * ax + t.face_ny[s] * ay + t.face_nz[s] * az);
                            Re[i] += -tmp * btemp[fb];
                        } else {
                            continue;  // no non-orthogonal term on other boundary types
                        }
                    } else {
                        int nbr = t.face_neighbor[s];
                        ax = t.cx[nbr] - t.cx[i]; ay = t.cy[nbr] - t.cy[i]; az = t.cz[nbr] - t.cz[i];
                    }
                    double dn = ax * t.face_nx[s] + ay * t.face_ny[s] + az * t.face_nz[s];
                    double a2x = t.face_nx[s] - ax / dn, a2y = t.face_ny[s] - ay / dn, a2z = t.face_nz[s] - az / dn;
                    double gx = 0, gy = 0, gz = az;
                    if (dim == 2) {
                        int v0 = t.face_vert[s * 4], v1 = t.face_vert[s * 4 + 1];
                        double mx = t.node_x[v1] - t.node_x[v0], my = t.node_y[v1] - t.node_y[v0];
                        gx = (Tv[v1] - Tv[v0]) / (mx * mx + my * my) * mx;
                        gy = (Tv[v1] - Tv[v0]) / (mx * mx + my * my) * my;
                    } else {
                        // Get_gradient_3D on the cell's own face (StaticFourier.cpp:571-593)
                        int nv = t.face_nvert[s];
                        vector<vector<double>> J(nv, vector<double>(3)), Pf;
                        vector<double> d1(nv, 0.0);
                        J[0][0] = t.face_nx[s]; J[0][1] = t.face_ny[s]; J[0][2] = t.face_nz[s];
                        for (int m = 0; m < nv - 1; ++m) {
                            int a = t.face_vert[s * 4 + m], c = t.face_vert[s * 4 + m + 1];
                            J[m + 1][0] = (1.0 / Lx * (t.node_x[c] - t.node_x[a]));
                            J[m + 1][1] = (1.0 / Lx * (t.node_y[c] - t.node_y[a]));
                            J[m + 1][2] = (1.0 / Lx * (t.node_z[c] - t.node_z[a]));
                            d1[m + 1] = 1.0 / Lx * (Tv[c] - Tv[a]);
                        }
                        lsq_3d(J, Pf);
                        gx = 0; gy = 0; gz = 0;
                        for (int m = 0; m < nv; ++m) {
                            gx += Pf[0][m] * d1[m];
                            gy += Pf[1][m] * d1[m];
                            gz += Pf[2][m] * d1[m];
                        }
                    }
                    if (fb == -1) {
                        int nbr = t.face_neighbor[s];
                        Re[i] -= 2 * (kap(i) * kap(nbr)) / (kap(i) + kap(nbr)) * t.face_area[s] * (a2x * gx + a2y * gy + a2z * gz);
                    } else {
                        Re[i] -= kap(i) * t.face_area[s] * (a2x * gx + a2y * gy + a2z * gz);
                    }
                }
            banded_solve(K, Re, Tn);
            TF = Tn;
            // _get_heat_flux (StaticFourier.cpp:615-710), 2D and 3D
            for (int i = 0; i < nc; ++i) {
                int nump = 0;
                for (int j = 0; j < t.face_size[i]; ++j)
                    if (t.face_bound[slot(i, j)] == -1) nump++;
                if (dim == 2 && nump < 2) { fqx[i] = 0; fqy[i] = 0; continue; }
                double gx = 0, gy = 0, gz = 0;
                for (int k = t.nbr_ptr[i]; k < t.nbr_ptr[i + 1]; ++k) {
                    double d1 = 1.0 / Lx * (TF[t.nbr_idx[k]] - TF[i]);
                    int m = k - t.nbr_ptr[i];
                    double p0 = cmSet[i] ? cellMat[i][0][m] : 0.0, p1 = cmSet[i] ? cellMat[i][1][m] : 0.0;
                    gx += p0 * d1;
                    gy += p1 * d1;
                    if (dim == 3) gz += (cmSet[i] ? cellMat[i][2][m] : 0.0) * d1;
                }
                fqx[i] = -kap(i) * gx;
                fqy[i] = -kap(i) * gy;
                if (dim == 3) fqz[i] = -kap(i) * gz;
            }
            err = 0;
            for (int i = 0; i < nc; ++i) err += std::pow(Tprev[i] - TF[i], 2);
            err = err / nc;
            passes++;
            if (passes > 100000) { g_err = "Fourier loop does not converge"; return -1; }
        }
        return passes;
    }

    // ------------------------------------------------ BTE per-system kernels
    // _get_gradient_larger, Order 2 (StaticBTESynthetic.cpp:1242-1340)
    void gradient(int b, int d) {
        const double *ed = E_(b, d);
        for (int i = 0; i < nc; ++i) { gradX[i] = 0; gradY[i] = 0; gradZ[i] = 0; }
        for (int i = 0; i < nc; ++i) {
            int a = t.nbr_ptr[i], n = t.nbr_ptr[i + 1] - a;
            if (n < 3) continue;
            double gx = 0, gy = 0, gz = 0;
            for (int m = 0; m < n; ++m) {
                double d1 = 1.0 / Lx * (ed[t.nbr_idx[a + m]] - ed[i]);
                gx += cellMat[i][0][m] * d1;
                gy += cellMat[i][1][m] * d1;
                if (dim == 3) gz += cellMat[i][2][m] * d1;
            }
            if (std::isnan(gx) || std::isnan(gy) || std::isnan(gz)) continue;
            gradX[i] = gx; gradY[i] = gy; gradZ[i] = gz;
        }
        for (int i = 0; i < nc; ++i) lim[i] = 1.0;
        if (!limiter) return;
        // Order 2limit (StaticBTESynthetic.cpp:1372-1422): the range of the
        // cell, its face neighbours and the incoming non-specular boundary
        // values (side 0), then the faces in order; each face that
        // overshoots the range sets the factor anew
        for (int i = 0; i < nc; ++i) {
            int m = t.matter[i];
            double mx = ed[i], mn = ed[i];
            for (int j = 0; j < t.face_size[i]; ++j) {
                size_t s = slot(i, j);
                int fb = t.face_bound[s];
                double v;
                if (fb == -1) v = ed[t.face_neighbor[s]];
                else if (t.bound_type[fb] != 3 && dotvn(m, b, d, s) < 0) v = EB(b, d, fb, 0);
                else continue;
                if (v > mx) mx = v;
                if (v < mn) mn = v;
            }
            for (int j = 0; j < t.face_size[i]; ++j) {
                size_t s = slot(i, j);
                double ax = t.face_cx[s] - t.cx[i], ay = t.face_cy[s] - t.cy[i], az = t.face_cz[s] - t.cz[i];
                double p = gradX[i] * ax + gradY[i] * ay + gradZ[i] * az;
                if (p * lim[i] + ed[i] < mn) {
                    double y = (mn - ed[i]) / p;
                    lim[i] = (y * y + 2 * y) / (y * y + y + 2);
                }
                if (p * lim[i] + ed[i] > mx) {
                    double y = (mx - ed[i]) / p;
                    lim[i] = (y * y + 2 * y) / (y * y + y + 2);
                }
            }
        }
    }

    // _get_Re (StaticBTESynthetic.cpp:1427-1500) followed by the exact solve
    // of the upwind matrix of _get_coefficient (:784-840) and e += delta
    // (StaticBTESyntheticSolver.cpp:96-100).
    // first: Order 1, _get_Re_firstorder (StaticBTESynthetic.cpp:2841-2881):
    // Re holds only C T_old, the heat source and the incoming boundary values,
    // and the solve gives e itself (StaticBTESyntheticSolver.cpp:575-579)
    bool sweep(int b, int d, vector<int> &order, vector<int> &indeg, vector<double> &Re, vector<double> &diag) {
        double *ed = E_(b, d);
        for (int i = 0; i < nc; ++i) Re[i] = 0;
        for (int ie = 0; ie < nc; ++ie) {
            int m = t.matter[ie];
            double dg = 0;
            for (int j = 0; j < t.face_size[ie]; ++j) {
                size_t s = slot(ie, j);
                double dp = dotvn(m, b, d, s);
                double tmp = tau(m, b) * t.face_area[s] / t.volume[ie] * dp;
                if (first) {
                    if (dp >= 0) dg += tmp;
                } else if (dp >= 0) {
                    double ax = t.face_cx[s] - t.cx[ie], ay = t.face_cy[s] - t.cy[ie], az = t.face_cz[s] - t.cz[ie];
                    double ev = (ed[ie] + (ax * gradX[ie] + ay * gradY[ie] + az * gradZ[ie]) * lim[ie]);
                    Re[ie] -= tmp * ev;
                    dg += tmp;
                } else if (t.face_bound[s] == -1) {
                    int nbr = t.face_neighbor[s];
                    double ax = t.face_cx[s] - t.cx[nbr], ay = t.face_cy[s] - t.cy[nbr], az = t.face_cz[s] - t.cz[nbr];
                    double ev = (ed[nbr] + (ax * gradX[nbr] + ay * gradY[nbr] + az * gradZ[nbr]) * lim[nbr]);
                    Re[ie] -= tmp * ev;
                }
            }
            dg += 1.0;
            diag[ie] = dg;
            Re[ie] += Told[ie] * cap(m, b);
            Re[ie] += t.heat_source[ie] * hr(m, b) * tau(m, b);
            if (!first) Re[ie] -= ed[ie];
        }
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < 2; ++ic) {
                int ie = t.bound_cell[ib * 2 + ic], jf = t.bound_face[ib * 2 + ic];
                if (ie < 0) continue;
                int m = t.matter[ie];
                size_t s = slot(ie, jf);
                double dp = dotvn(m, b, d, s);
                if (dp < 0) {
                    double tmp = tau(m, b) * t.face_area[s] / t.volume[ie] * dp;
                    Re[ie] -= tmp * EB(b, d, ib, ic);
                }
            }
        // topological order of the upwind graph: i depends on nbr through an
        // interior face with v.n < 0
        order.clear();
        indeg.assign(nc, 0);
        for (int ie = 0; ie < nc; ++ie)
            for (int j = 0; j < t.face_size[ie]; ++j) {
                size_t s = slot(ie, j);
                if (t.face_bound[s] == -1 && dotvn(t.matter[ie], b, d, s) < 0) indeg[ie]++;
            }
        vector<int> q;
        for (int i = 0; i < nc; ++i)
            if (indeg[i] == 0) q.push_back(i);
        vector<double> delta(nc, 0.0);
        size_t head = 0;
        while (head < q.size()) {
            int ie = q[head++];
            int m = t.matter[ie];
            double r = Re[ie];
            for (int j = 0; j < t.face_size[ie]; ++j) {
                size_t s = slot(ie, j);
                if (t.face_bound[s] != -1) continue;
                double dp = dotvn(m, b, d, s);
                if (dp < 0) r -= tau(m, b) * t.face_area[s] / t.volume[ie] * dp * delta[t.face_neighbor[s]];
            }
            delta[ie] = r / diag[ie];
            // release downwind neighbours
            for (int j = 0; j < t.face_size[ie]; ++j) {
                size_t s = slot(ie, j);
                if (t.face_bound[s] != -1) continue;
                int nbr = t.face_neighbor[s];
                for (int k = 0; k < t.face_size[nbr]; ++k) {
                    size_t s2 = slot(nbr, k);
                    if (t.face_neighbor[s2] == ie && t.face_bound[s2] == -1 && dotvn(t.matter[nbr], b, d, s2) < 0)
                        if (--indeg[nbr] == 0) q.push_back(nbr);
                }
            }
        }
        if ((int)q.size() != nc) {
            g_err = "upwind graph of a direction has a cycle; the exact sweep needs an acyclic mesh";
            return false;
        }
        for (int i = 0; i < nc; ++i) ed[i] = first ? delta[i] : delta[i] + ed[i];
        return true;
    }

    // _get_bound_ee (StaticBTESynthetic.cpp:1502-1539)
    void bound_ee(int b, int d) {
        double *ed = E_(b, d);
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < 2; ++ic) {
                int ie = t.bound_cell[ib * 2 + ic], jf = t.bound_face[ib * 2 + ic];
                if (ie < 0) continue;
                size_t s = slot(ie, jf);
                if (dotvn(t.matter[ie], b, d, s) >= 0) {
                    double ax = t.face_cx[s] - t.cx[ie], ay = t.face_cy[s] - t.cy[ie], az = t.face_cz[s] - t.cz[ie];
                    double ev = (ed[ie] + (ax * gradX[ie] + ay * gradY[ie] + az * gradZ[ie]) * lim[ie]);
                    eboundLocal[((size_t)b * nb + ib) * 2 + ic] = ev;
                    EB(b, d, ib, ic) = ev;
                }
            }
    }

    // _set_vertex_energydensity (StaticBTESynthetic.cpp:1541-1616)
    void vertex_e(int b, int d) {
        const double *ed = E_(b, d);
        vector<double> tmp((size_t)nn * nm, 0.0), R((size_t)nn * nm, 0.0);
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < t.cell_nvert[i]; ++j) {
                int v = t.cell_vert[(size_t)i * 8 + j];
                double dd = std::sqrt(std::pow(t.node_x[v] - t.cx[i], 2) + std::pow(t.node_y[v] - t.cy[i], 2) +
                                      std::pow(t.node_z[v] - t.cz[i], 2));
                tmp[(size_t)v * nm + t.matter[i]] += ed[i] / dd;
                R[(size_t)v * nm + t.matter[i]] += 1.0 / dd;
            }
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < 2; ++ic) {
                int i = t.bound_cell[ib * 2 + ic], j = t.bound_face[ib * 2 + ic];
                if (i < 0) continue;
                size_t s = slot(i, j);
                for (int k = 0; k < t.face_nvert[s]; ++k) {
                    int v = t.face_vert[s * 4 + k];
                    tmp[(size_t)v * nm + t.matter[i]] = 0;
                    R[(size_t)v * nm + t.matter[i]] = 0;
                }
            }
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < 2; ++ic) {
                int i = t.bound_cell[ib * 2 + ic], j = t.bound_face[ib * 2 + ic];
                if (i < 0) continue;
                size_t s = slot(i, j);
                for (int k = 0; k < t.face_nvert[s]; ++k) {
                    int v = t.face_vert[s * 4 + k];
                    double dd = std::sqrt(std::pow(t.node_x[v] - t.face_cx[s], 2) + std::pow(t.node_y[v] - t.face_cy[s], 2) +
                                          std::pow(t.node_z[v] - t.face_cz[s], 2));
                    tmp[(size_t)v * nm + t.matter[i]] += EB(b, d, ib, ic) / dd;
                    R[(size_t)v * nm + t.matter[i]] += 1.0 / dd;
                }
            }
        for (size_t k = 0; k < eV.size(); ++k) {
            eV[k] = tmp[k] / R[k];
            if (std::isnan(eV[k])) eV[k] = 0;
        }
    }

    // _get_face_gradient (StaticBTESynthetic.cpp:1651-1783); in 3D the
    // tangential part is FaceMatrix . d1 (:1680-1699, :1751-1770)
    bool face_gradient(int b, int d) {
        const double *ed = E_(b, d);
        for (int f = 0; f < nf; ++f) {
            int fb = t.fface_bound[f];
            if (fb == -1) {
                int i = t.fface_cell[f * 2], j = t.fface_face[f * 2], i2 = t.fface_cell[f * 2 + 1];
                size_t s = slot(i, j);
                double ax = t.cx[i2] - t.cx[i], ay = t.cy[i2] - t.cy[i], az = t.cz[i2] - t.cz[i];
                double gn = (ed[i2] - ed[i]) * (ax * t.face_nx[s] + ay * t.face_ny[s] + az * t.face_nz[s]) / (ax * ax + ay * ay + az * az);
                double gx, gy, gz;
                if (dim == 2) {
                    int v0 = t.face_vert[s * 4], v1 = t.face_vert[s * 4 + 1];
                    double mx = t.node_x[v1] - t.node_x[v0], my = t.node_y[v1] - t.node_y[v0];
                    double dv = eV[(size_t)v1 * nm + t.matter[i]] - eV[(size_t)v0 * nm + t.matter[i]];
                    gx = dv / (mx * mx + my * my) * mx; gy = dv / (mx * mx + my * my) * my; gz = 0;
                } else {
                    double g3[3];
                    tangential(f, s, eV, nm, t.matter[i], g3);
                    gx = g3[0]; gy = g3[1]; gz = g3[2];
                }
                gx = gn * t.face_nx[s] + gx;
                gy = gn * t.face_ny[s] + gy;
                gz = gn * t.face_nz[s] + gz;
                gfx[f * 2] = gx; gfy[f * 2] = gy; gfz[f * 2] = gz;
                gfx[f * 2 + 1] = gx; gfy[f * 2 + 1] = gy; gfz[f * 2 + 1] = gz;
            } else if (t.bound_type[fb] != 3) {
                for (int ic = 0; ic < 2; ++ic) {
                    int ie = t.fface_cell[f * 2 + ic], jf = t.fface_face[f * 2 + ic];
                    if (ie < 0) continue;
                    int ibc = 0;
                    for (int k = 0; k < 2; ++k)
                        if (t.bound_cell[fb * 2 + k] == ie) ibc = k;
                    size_t s = slot(ie, jf);
                    double ax = t.face_cx[s] - t.cx[ie], ay = t.face_cy[s] - t.cy[ie], az = t.face_cz[s] - t.cz[ie];
                    double ef = EB(b, d, fb, ibc);
                    double gn = (ef - ed[ie]) * (ax * t.face_nx[s] + ay * t.face_ny[s] + az * t.face_nz[s]) / (ax * ax + ay * ay + az * az);
                    // the reference indexes the vertex field without the
                    // matter offset here (StaticBTESynthetic.cpp:1741-1742, 1751-1758)
                    double gx, gy, gz;
                    if (dim == 2) {
                        int v0 = t.face_vert[s * 4], v1 = t.face_vert[s * 4 + 1];
                        double mx = t.node_x[v1] - t.node_x[v0], my = t.node_y[v1] - t.node_y[v0];
                        double dv = eV[v1] - eV[v0];
                        gx = dv / (mx * mx + my * my) * mx; gy = dv / (mx * mx + my * my) * my; gz = 0;
                    } else {
                        double g3[3];
                        tangential(f, s, eV, 1, 0, g3);
                        gx = g3[0]; gy = g3[1]; gz = g3[2];
                    }
                    gx = gn * t.face_nx[s] + gx;
                    gy = gn * t.face_ny[s] + gy;
                    gz = gn * t.face_nz[s] + gz;
                    gfx[f * 2 + ic] = gx; gfy[f * 2 + ic] = gy; gfz[f * 2 + ic] = gz;
                }
            }
        }
        return true;
    }

    // _get_non_fourier (StaticBTESynthetic.cpp:1883-1937)
    void non_fourier(int b, int d) {
        auto add = [&](int ie, size_t s, double area, size_t g) {
            int m = t.matter[ie];
            if (!(hr(m, b) > 0)) return;
            ReMacro[ie] -= w(d) * area * tau(m, b) *
                           (vgx(m, b, d) * gfx[g] + vgy(m, b, d) * gfy[g] + vgz(m, b, d) * gfz[g]) *
                           (vgx(m, b, d) * t.face_nx[s] + vgy(m, b, d) * t.face_ny[s] + vgz(m, b, d) * t.face_nz[s]);
            ReMacro[ie] += t.kappa_bulk[m] * w(d) * t.face_area[s] * lr(m, b) / cap(m, b) *
                           (t.face_nx[s] * gfx[g] + t.face_ny[s] * gfy[g] + t.face_nz[s] * gfz[g]);
        };
        for (int f = 0; f < nf; ++f) {
            int fb = t.fface_bound[f];
            if (fb == -1) {
                int ie = t.fface_cell[f * 2], jf = t.fface_face[f * 2];
                int ie2 = t.fface_cell[f * 2 + 1], jf2 = t.fface_face[f * 2 + 1];
                add(ie, slot(ie, jf), t.face_area[slot(ie, jf)], (size_t)f * 2);
                add(ie2, slot(ie2, jf2), t.face_area[slot(ie, jf)], (size_t)f * 2 + 1);
            } else if (t.bound_type[fb] != 3) {
                for (int ic = 0; ic < 2; ++ic) {
                    int ie = t.fface_cell[f * 2 + ic], jf = t.fface_face[f * 2 + ic];
                    if (ie < 0) continue;
                    add(ie, slot(ie, jf), t.face_area[slot(ie, jf)], (size_t)f * 2 + ic);
                }
            }
        }
    }

    // _recover_temperature / _get_total_energy / _get_heat_flux
    // (StaticBTESynthetic.cpp:2485-2520)
    void moments(int b, int d) {
        const double *ed = E_(b, d);
        for (int i = 0; i < nc; ++i) {
            int m = t.matter[i];
            T[i] += lr(m, b) * ed[i] * w(d) / cap(m, b);
            Tband[(size_t)b * nc + i] += ed[i] * w(d) / cap(m, b);
        }
        for (int i = 0; i < nc; ++i) E[i] += ed[i] * w(d) / t.capacity_bulk[t.matter[i]];
        for (int i = 0; i < nc; ++i) {
            int m = t.matter[i];
            qx[i] += vgx(m, b, d) * w(d) * ed[i];
            qy[i] += vgy(m, b, d) * w(d) * ed[i];
            qz[i] += vgz(m, b, d) * w(d) * ed[i];
        }
    }

    // _set_bound_ee_1 (StaticBTESynthetic.cpp:2083-2232) at np=1: the
    // Allgather of each system copies eboundLocal back into ebound before the
    // next system is visited, so later systems see earlier updates.
    bool set_bound_ee_1() {
        for (int d = 0; d < nD; d++)
            for (int b = 0; b < nB; ++b) {
                for (int ib = 0; ib < nb; ++ib)
                    for (int ic = 0; ic < 2; ++ic) {
                        int ie = t.bound_cell[ib * 2 + ic], jf = t.bound_face[ib * 2 + ic];
                        if (ie < 0) continue;
                        int m = t.matter[ie];
                        size_t s = slot(ie, jf);
                        double &loc = eboundLocal[((size_t)b * nb + ib) * 2 + ic];
                        double dp = dotvn(m, b, d, s);
                        if (hr(m, b) != 0 && dp < 0) {
                            int ty = t.bound_type[ib];
                            if (ty == 1) {
                                loc = cap(m, b) * t.bound_temp[ib];
                            } else if (ty == 2) {
                                double num = 0, den = 0;
                                for (int k = 0; k < nD; ++k) {
                                    double d1 = dotvn(m, b, k, s);
                                    if (d1 >= 0) {
                                        num += EB(b, k, ib, ic) * (d1 * w(k));
                                        den += (d1 * w(k));
                                    }
                                }
                                loc = num / den;
                            } else if (ty == 3) {
                                double d1 = (t.dir_x[d] * t.face_nx[s] + t.dir_y[d] * t.face_ny[s] + t.dir_z[d] * t.face_nz[s]);
                                double rx = t.dir_x[d] - t.face_nx[s] * d1 * 2, ry = t.dir_y[d] - t.face_ny[s] * d1 * 2,
                                       rz = t.dir_z[d] - t.face_nz[s] * d1 * 2;
                                double close = 1;
                                int nfd = -1;
                                for (int k = 0; k < nD; ++k) {
                                    double len = std::sqrt(std::pow(rx - t.dir_x[k], 2) + std::pow(ry - t.dir_y[k], 2) + std::pow(rz - t.dir_z[k], 2));
                                    if (len < close) { nfd = k; close = len; }
                                }
                                if (nfd < 0) { g_err = "specular reflection has no direction within distance 1"; return false; }
                                loc = EB(b, nfd, ib, ic);
                            } else if (ty < 0) {
                                double ev = EB(b, d, t.bound_connect[ib], ic);
                                ev = ev + cap(m, b) * (t.bound_temp[ib] - t.bound_temp[t.bound_connect[ib]]);
                                loc = ev;
                            } else if (ty == 5 || ty == 4) {
                                g_err = "interface (type 4/5) boundaries are not in the oracle";
                                return false;
                            } else {
                                loc = EB(b, d, ib, ic);
                            }
                        } else {
                            loc = EB(b, d, ib, ic);
                        }
                    }
                // Allgather at np=1
                for (int ib = 0; ib < nb; ++ib)
                    for (int ic = 0; ic < 2; ++ic) EB(b, d, ib, ic) = eboundLocal[((size_t)b * nb + ib) * 2 + ic];
            }
        return true;
    }

    // _set_macro_boundary_condition (StaticBTESynthetic.cpp:2234-2274)
    void macro_bc() {
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < 2; ++ic) {
                tmac[ib][ic] = 0;
                int ie = t.bound_cell[ib * 2 + ic];
                if (ie < 0) continue;
                int m = t.matter[ie];
                for (int b = 0; b < nB; ++b)
                    for (int d = 0; d < nD; d++) tmac[ib][ic] += lr(m, b) * EB(b, d, ib, ic) / cap(m, b) * w(d);
            }
        for (int ib = 0; ib < nb; ++ib)
            for (int ic = 0; ic < 2; ++ic) {
                if (t.bound_cell[ib * 2 + ic] < 0) continue;
                if (t.bound_type[ib] < 0 && t.bound_temp[ib] > 0) {
                    tmac[ib][ic] = t.bound_temp[ib] / 2;
                    tmac[t.bound_connect[ib]][ic] = -t.bound_temp[ib] / 2;
                }
            }
    }

    // _get_coefficient_macro (StaticBTESynthetic.cpp:842-884)
    bool macro_matrix() {
        vector<vector<std::pair<int, double>>> rows(nc);
        for (int ie = 0; ie < nc; ++ie) {
            double diag = 0;
            double kap = t.kappa_bulk[t.matter[ie]];
            for (int j = 0; j < t.face_size[ie]; ++j) {
                size_t s = slot(ie, j);
                double tmp = kap * t.face_area[s];
                int fb = t.face_bound[s];
                if (fb == -1) {
                    int nbr = t.face_neighbor[s];
                    double ax = t.cx[nbr] - t.cx[ie], ay = t.cy[nbr] - t.cy[ie], az = t.cz[nbr] - t.cz[ie];
                    tmp = tmp / (t.face_nx[s] * ax + t.face_ny[s] * ay + t.face_nz[s] * az);
                    diag += -tmp;
                    rows[ie].push_back({nbr, tmp});
                } else if (t.bound_type[fb] != 3) {
                    double ax = t.face_cx[s] - t.cx[ie], ay = t.face_cy[s] - t.cy[ie], az = t.face_cz[s] - t.cz[ie];
                    tmp = tmp / (t.face_nx[s] * ax + t.face_ny[s] * ay + t.face_nz[s] * az);
                    diag += -tmp;
                }
            }
            rows[ie].push_back({ie, diag});
        }
        if (!banded_factor(rows, Kmacro)) { g_err = "macro matrix is singular"; return false; }
        return true;
    }

    // _get_Re_macro (StaticBTESynthetic.cpp:2307-2407); the 3D tangential
    // gradient comes from the face's first side (:2351-2381)
    bool macro_rhs(vector<double> &Re) {
        for (int ie = 0; ie < nc; ++ie) {
            double r = ReMacro[ie];
            double kap = t.kappa_bulk[t.matter[ie]];
            r -= t.heat_source[ie] * t.volume[ie];
            for (int j = 0; j < t.face_size[ie]; ++j) {
                size_t s = slot(ie, j);
                int fb = t.face_bound[s];
                double ax = 0, ay = 0, az = 0;
                if (fb >= 0) {
                    if (t.bound_type[fb] == 3) continue;
                    for (int ic = 0; ic < 2; ++ic)
                        if (t.bound_cell[fb * 2 + ic] == ie) {
                            double tmp = kap * t.face_area[s];
                            ax = t.face_cx[s] - t.cx[ie]; ay = t.face_cy[s] - t.cy[ie]; az = t.face_cz[s] - t.cz[ie];
                            tmp = tmp / (t.face_nx[s] * ax + t.face_ny[s] * ay + t.face_nz[s] * az);
                            r += -tmp * tmac[fb][ic];
                        }
                } else {
                    int nbr = t.face_neighbor[s];
                    ax = t.cx[nbr] - t.cx[ie]; ay = t.cy[nbr] - t.cy[ie]; az = t.cz[nbr] - t.cz[ie];
                }
                double dn = ax * t.face_nx[s] + ay * t.face_ny[s] + az * t.face_nz[s];
                double a2x = t.face_nx[s] - ax / dn, a2y = t.face_ny[s] - ay / dn, a2z = t.face_nz[s] - az / dn;
                if (dim == 2) {
                    int v0 = t.face_vert[s * 4], v1 = t.face_vert[s * 4 + 1];
                    double gx = t.node_x[v1] - t.node_x[v0], gy = t.node_y[v1] - t.node_y[v0];
                    double dv = TV[(size_t)v1 * nm + t.matter[ie]] - TV[(size_t)v0 * nm + t.matter[ie]];
                    double tx = dv / (gx * gx + gy * gy) * gx, ty = dv / (gx * gx + gy * gy) * gy;
                    r -= kap * t.face_area[s] * (a2x * tx + a2y * ty + a2z * az);
                } else {
                    int f = t.face_index[s];
                    int i1 = t.fface_cell[f * 2], j1 = t.fface_face[f * 2];
                    double g3[3];
                    tangential(f, slot(i1, j1), TV, nm, t.matter[i1], g3);
                    r -= kap * t.face_area[s] * (a2x * g3[0] + a2y * g3[1] + a2z * g3[2]);
                }
            }
            Re[ie] = r;
        }
        return true;
    }

    // _marco_solution (StaticBTESynthetic.cpp:2692-2734)
    int macro_solution() {
        double inner = 1000;
        int inc = 0, passes = 0;
        vector<double> Re(nc), sol;
        while (inner > 1e-6) {
            T1 = T;
            vertex_from_cells(T, TV);
            if (!macro_rhs(Re)) return -1;
            banded_solve(Kmacro, Re, sol);
            inner = 0;
            for (int i = 0; i < nc; ++i) {
                T[i] = sol[i];
                inner += std::pow(T1[i] - T[i], 2);
            }
            inner = inner / nc;
            if (inner > errInner) inc++;
            errInner = inner;
            passes++;
            if (inc >= 10) { g_err = "cannot congerve, try to set Order=1 or 2limit"; return -2; }
        }
        return passes;
    }

    // _get_magin_check_error (StaticBTESynthetic.cpp:2565-2611)
    bool check_error(double tolT, double tolQ) {
        errT = 0;
        double tmax = 0;
        for (int i = 0; i < nc; ++i)
            if (tmax < std::fabs(T[i])) tmax = std::fabs(T[i]);
        for (int i = 0; i < nc; ++i) errT += std::sqrt(std::pow(T[i] - Told[i], 2));
        errT = errT / nc / tmax;
        errQ = 0;
        double hmax = 0;
        for (int i = 0; i < nc; ++i) {
            double h = std::sqrt(std::pow(qx[i], 2) + std::pow(qy[i], 2) + std::pow(qz[i], 2));
            double h1 = std::sqrt(std::pow(qxo[i], 2) + std::pow(qyo[i], 2) + std::pow(qzo[i], 2));
            if (hmax < h) hmax = h;
            errQ += std::fabs(h - h1);
        }
        errQ = errQ / nc / hmax;
        if ((errT > errTold && errQ > errQold) || std::isnan(errT) || std::isnan(errQ) || std::isinf(errT) || std::isinf(errQ))
            errInc++;
        else
            errInc = 0;
        errTold = errT;
        errQold = errQ;
        return errT <= tolT && errQ <= tolQ;
    }

    // StaticBTESynthetic::solve (StaticBTESyntheticSolver.cpp:12-247)
    int solve(const double *T0, const double *q0x, const double *q0y, const double *q0z, int maxIter, double tolT,
              double tolQ, int *iters, int *status) {
        limiter = t.order == GBTE_ORDER_2LIMIT;
        // Order 1 (solve_firstorder, StaticBTESyntheticSolver.cpp:487-966):
        // the gradients stay zero, so the boundary values are the cell values
        // (_get_bound_ee_firstorder, StaticBTESynthetic.cpp:2883-2906)
        first = t.order == GBTE_ORDER_1;
        // Sequential: StaticBTESynthetic::solve_DOM (StaticBTESyntheticSolver.cpp:968-1178),
        // the same per-system steps without the vertex / face-gradient /
        // non-Fourier terms and without the synthetic correction
        bool seq = t.scheme == GBTE_SCHEME_SEQUENTIAL;
        for (int ib = 0; ib < nb; ++ib)
            if (t.bound_type[ib] == 4 || t.bound_type[ib] == 5) { g_err = "interface boundaries are not in the oracle"; return -1; }
        size_t ne = (size_t)nB * nD * nc;
        e.assign(ne, 0.0);
        ebound.assign((size_t)nB * nD * nb * 2, 0.0);
        eboundLocal.assign((size_t)nB * nb * 2, 0.0);
        T.assign(T0, T0 + nc); Told.assign(nc, 0.0); T1.assign(nc, 0.0);
        qx.assign(q0x, q0x + nc); qy.assign(q0y, q0y + nc); qz.assign(q0z, q0z + nc);
        qxo.assign(nc, 0.0); qyo.assign(nc, 0.0); qzo.assign(nc, 0.0);
        Tband.assign((size_t)nB * nc, 0.0); E.assign(nc, 0.0); ReMacro.assign(nc, 0.0);
        gradX.assign(nc, 0.0); gradY.assign(nc, 0.0); gradZ.assign(nc, 0.0); lim.assign(nc, 1.0);
        eV.assign((size_t)nn * nm, 0.0); TV.assign((size_t)nn * nm, 0.0);
        gfx.assign((size_t)nf * 2, 0.0); gfy.assign((size_t)nf * 2, 0.0); gfz.assign((size_t)nf * 2, 0.0);
        tmac.assign(nb, vector<double>(2, 0.0));
        errInc = 0;
        // _set_initial (StaticBTESynthetic.cpp:765-782)
        for (int b = 0; b < nB; ++b)
            for (int d = 0; d < nD; ++d) {
                double *ed = E_(b, d);
                for (int i = 0; i < nc; ++i) ed[i] = T[i] * cap(t.matter[i], b);
            }
        if (!macro_matrix()) return -1;
        cell_matrix();
        face_matrix();
        for (int d = 0; d < nD; d++)
            for (int b = 0; b < nB; ++b) {
                if (!first) gradient(b, d);
                bound_ee(b, d);
            }
        if (!set_bound_ee_1()) return -1;
        vector<int> order, indeg;
        vector<double> Re(nc), diag(nc);
        int nt = 0;
        *status = GBTE_OK;
        for (nt = 0; nt < maxIter; ++nt) {
            // copy() (StaticBTESynthetic.cpp:2655-2690)
            for (int i = 0; i < nc; ++i) { Told[i] = T[i]; qxo[i] = qx[i]; qyo[i] = qy[i]; qzo[i] = qz[i]; }
            std::fill(T.begin(), T.end(), 0.0); std::fill(E.begin(), E.end(), 0.0);
            std::fill(qx.begin(), qx.end(), 0.0); std::fill(qy.begin(), qy.end(), 0.0); std::fill(qz.begin(), qz.end(), 0.0);
            std::fill(ReMacro.begin(), ReMacro.end(), 0.0); std::fill(Tband.begin(), Tband.end(), 0.0);
            std::fill(eboundLocal.begin(), eboundLocal.end(), 0.0);
            for (int d = 0; d < nD; d++)
                for (int b = 0; b < nB; ++b) {
                    if (!first) gradient(b, d);
                    if (!sweep(b, d, order, indeg, Re, diag)) return -1;
                    bound_ee(b, d);
                    if (!seq) {
                        vertex_e(b, d);
                        if (!face_gradient(b, d)) return -1;
                        non_fourier(b, d);
                    }
                    moments(b, d);
                }
            if (!set_bound_ee_1()) return -1;
            if (!seq) {
                macro_bc();
                int p = macro_solution();
                if (p < 0) { *status = GBTE_DIVERGED; ++nt; break; }
                macroPasses += p;
            }
            bool conv = check_error(tolT, tolQ);
            if (std::isnan(errT) || std::isnan(errQ)) { *status = GBTE_NAN; ++nt; break; }
            if (conv) { ++nt; break; }
            if (errInc >= 10) { *status = GBTE_DIVERGED; ++nt; break; }
        }
        *iters = nt;
        return 0;
    }
};

}  // namespace

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

// StaticFourier::_solve from the initial temperature, then T/q as handed to
// the BTE solver (Solution.cpp:236-273).  Returns passes, or -1.
int oracle_fourier(const gbte_tables *t, double *T, double *qx, double *qy, double *qz) {
    Oracle o(*t);
    std::vector<double> a, b, c, d;
    int p = o.fourier(a, b, c, d);
    if (p < 0) return -1;
    std::copy(a.begin(), a.end(), T);
    std::copy(b.begin(), b.end(), qx);
    std::copy(c.begin(), c.end(), qy);
    std::copy(d.begin(), d.end(), qz);
    return p;
}

// Full steady run: Fourier warm start, then StaticBTESynthetic::solve.
// info[0..3] = outer iterations, macro passes, Fourier passes, status.
// err[0..1] = last temperature / flux errors.  Any output may be NULL.
int oracle_run(const gbte_tables *t, int max_iter, double tolT, double tolQ, double *T_lattice, double *T_energy,
               double *qx, double *qy, double *qz, double *T_band, int *info, double *err) {
    Oracle o(*t);
    std::vector<double> TF, fx, fy, fz;
    int fp = o.fourier(TF, fx, fy, fz);
    if (fp < 0) return -1;
    if (t->dim == 2) std::fill(fz.begin(), fz.end(), 0.0);
    int iters = 0, status = 0;
    if (o.solve(TF.data(), fx.data(), fy.data(), fz.data(), max_iter, tolT, tolQ, &iters, &status) != 0) return -1;
    size_t nc = t->n_cells;
    if (T_lattice) std::copy(o.T.begin(), o.T.end(), T_lattice);
    if (T_energy) std::copy(o.E.begin(), o.E.end(), T_energy);
    if (qx) std::copy(o.qx.begin(), o.qx.end(), qx);
    if (qy) std::copy(o.qy.begin(), o.qy.end(), qy);
    if (qz) std::copy(o.qz.begin(), o.qz.end(), qz);
    if (T_band) std::copy(o.Tband.begin(), o.Tband.end(), T_band);
    if (info) { info[0] = iters; info[1] = o.macroPasses; info[2] = fp; info[3] = status; }
    if (err) { err[0] = o.errT; err[1] = o.errQ; }
    (void)nc;
    return 0;
}

// One outer iteration's worth of per-system stages at iteration 0 for system
// (band, dir), for comparison with the reference stage dumps.
// out arrays: grad[3*nc], e_new[nc], ebound_sys[nb*2], e_vertex[nn*nm],
// gface[3*nf*2], remacro[nc].
int oracle_stages(const gbte_tables *t, int band, int dir, double *grad, double *e_new, double *ebound_sys,
                  double *e_vertex, double *gface, double *remacro, double *ebound_init) {
    Oracle o(*t);
    std::vector<double> TF, fx, fy, fz;
    if (o.fourier(TF, fx, fy, fz) < 0) return -1;
    int nc = t->n_cells, nb = t->n_bounds, nB = t->n_bands, nD = t->n_dirs, nf = t->n_faces;
    // set-up of solve() and the copy() of iteration 0
    o.e.assign((size_t)nB * nD * nc, 0.0);
    o.ebound.assign((size_t)nB * nD * nb * 2, 0.0);
    o.eboundLocal.assign((size_t)nB * nb * 2, 0.0);
    o.T = TF; o.Told = TF;
    o.gradX.assign(nc, 0); o.gradY.assign(nc, 0); o.gradZ.assign(nc, 0);
    o.eV.assign((size_t)t->n_nodes * t->n_matters, 0.0);
    o.gfx.assign((size_t)nf * 2, 0.0); o.gfy.assign((size_t)nf * 2, 0.0); o.gfz.assign((size_t)nf * 2, 0.0);
    o.ReMacro.assign(nc, 0.0);
    for (int b = 0; b < nB; ++b)
        for (int d = 0; d < nD; ++d)
            for (int i = 0; i < nc; ++i) o.E_(b, d)[i] = TF[i] * o.cap(t->matter[i], b);
    o.cell_matrix();
    o.face_matrix();
    for (int d = 0; d < nD; d++)
        for (int b = 0; b < nB; ++b) { o.gradient(b, d); o.bound_ee(b, d); }
    if (!o.set_bound_ee_1()) return -1;
    if (ebound_init) std::copy(o.ebound.begin(), o.ebound.end(), ebound_init);
    std::fill(o.eboundLocal.begin(), o.eboundLocal.end(), 0.0);
    std::vector<int> order, indeg;
    std::vector<double> Re(nc), diag(nc);
    o.gradient(band, dir);
    for (int i = 0; i < nc; ++i) { grad[i] = o.gradX[i]; grad[nc + i] = o.gradY[i]; grad[2 * nc + i] = o.gradZ[i]; }
    if (!o.sweep(band, dir, order, indeg, Re, diag)) return -1;
    std::copy(o.E_(band, dir), o.E_(band, dir) + nc, e_new);
    o.bound_ee(band, dir);
    for (int ib = 0; ib < nb; ++ib)
        for (int ic = 0; ic < 2; ++ic) ebound_sys[ib * 2 + ic] = o.EB(band, dir, ib, ic);
    o.vertex_e(band, dir);
    std::copy(o.eV.begin(), o.eV.end(), e_vertex);
    if (!o.face_gradient(band, dir)) return -1;
    for (int k = 0; k < nf * 2; ++k) { gface[k] = o.gfx[k]; gface[nf * 2 + k] = o.gfy[k]; gface[4 * nf + k] = o.gfz[k]; }
    o.non_fourier(band, dir);
    std::copy(o.ReMacro.begin(), o.ReMacro.end(), remacro);
    return 0;
}

// CPU-baseline sample: wall time of the per-system transport work of one
// outer iteration (gradient, _get_Re + exact sweep, _get_bound_ee,
// _set_vertex_energydensity, _get_face_gradient, _get_non_fourier, moments;
// StaticBTESyntheticSolver.cpp:79-127) over the first n_sys systems in the
// reference's (direction, band) order, starting from e = C T.  Set-up and the
// synthetic solve are excluded; boundary values start at equilibrium with
// the cells, which changes no operation count.  Returns seconds, or -1.
double oracle_transport_sample(const gbte_tables *t, const double *T0, int n_sys) {
    Oracle o(*t);
    int nc = t->n_cells, nb = t->n_bounds, nB = t->n_bands, nD = t->n_dirs, nf = t->n_faces;
    o.e.assign((size_t)nB * nD * nc, 0.0);
    o.ebound.assign((size_t)nB * nD * nb * 2, 0.0);
    o.eboundLocal.assign((size_t)nB * nb * 2, 0.0);
    o.T.assign(T0, T0 + nc); o.Told = o.T;
    o.E.assign(nc, 0.0); o.qx.assign(nc, 0.0); o.qy.assign(nc, 0.0); o.qz.assign(nc, 0.0);
    o.Tband.assign((size_t)nB * nc, 0.0); o.ReMacro.assign(nc, 0.0);
    o.gradX.assign(nc, 0); o.gradY.assign(nc, 0); o.gradZ.assign(nc, 0);
    o.eV.assign((size_t)t->n_nodes * t->n_matters, 0.0);
    o.gfx.assign((size_t)nf * 2, 0.0); o.gfy.assign((size_t)nf * 2, 0.0); o.gfz.assign((size_t)nf * 2, 0.0);
    for (int b = 0; b < nB; ++b)
        for (int d = 0; d < nD; ++d)
            for (int i = 0; i < nc; ++i) o.E_(b, d)[i] = o.T[i] * o.cap(t->matter[i], b);
    o.cell_matrix();
    o.face_matrix();
    // boundary values at equilibrium with the cells (timing only)
    for (int b = 0; b < nB; ++b)
        for (int d = 0; d < nD; ++d)
            for (int ib = 0; ib < nb; ++ib)
                for (int ic = 0; ic < 2; ++ic) {
                    int c = t->bound_cell[ib * 2 + ic];
                    if (c >= 0) o.EB(b, d, ib, ic) = o.T[c] * o.cap(t->matter[c], b);
                }
    std::vector<int> order, indeg;
    std::vector<double> Re(nc), diag(nc);
    auto t0 = std::chrono::steady_clock::now();
    int k = 0;
    for (int d = 0; d < nD && k < n_sys; d++)
        for (int b = 0; b < nB && k < n_sys; ++b, ++k) {
            o.gradient(b, d);
            if (!o.sweep(b, d, order, indeg, Re, diag)) return -1;
            o.bound_ee(b, d);
            o.vertex_e(b, d);
            if (!o.face_gradient(b, d)) return -1;
            o.non_fourier(b, d);
            o.moments(b, d);
        }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// LSQ cell gradient of an arbitrary cell field with the stencil and matrices
// of _get_CellMatrix_larger / _get_gradient_larger (StaticBTESynthetic.cpp:
// 1000-1064, 1243-1340): g = [gx(nc), gy(nc), gz(nc)].
int oracle_lsq_gradient(const gbte_tables *t, const double *field, double *g) {
    Oracle o(*t);
    int nc = t->n_cells;
    o.cell_matrix();
    o.e.assign(field, field + nc);  // system (0, 0) storage
    o.gradX.assign(nc, 0); o.gradY.assign(nc, 0); o.gradZ.assign(nc, 0);
    o.gradient(0, 0);
    for (int i = 0; i < nc; ++i) { g[i] = o.gradX[i]; g[nc + i] = o.gradY[i]; g[2 * nc + i] = o.gradZ[i]; }
    return 0;
}

}  // extern "C"
