// Nested-dissection ordering and multifrontal Cholesky factorization (host).
//
// The synthetic (macro) matrix of _get_coefficient_macro
// (src/StaticBTESynthetic/StaticBTESynthetic.cpp:842-884) and the Fourier
// matrix of StaticFourier::_get_coefficient_macro (StaticFourier.cpp:449-518)
// are constant for a run and are factorized once by the reference (Eigen
// SparseLU).  Negated, both are symmetric positive definite M-matrices; this
// file factorizes them exactly as L L^T in a nested-dissection order:
//   * ordering: recursive coordinate bisection of the cell centres; each
//     bisection's separator is the set of cells on one side that touch the
//     other side in the matrix graph;
//   * factorization: multifrontal, with a dense frontal matrix per separator
//     (supernode) and extend-add of the children's update matrices;
//   * per supernode k the solve needs inv(L_kk) and L_{A(k),k}, where A(k)
//     are the ancestor positions coupled to k after fill; both are folded
//     into F_k = [inv(L_kk); -L_{A(k),k} inv(L_kk)], so that each supernode's
//     forward and backward step is one dense matrix-vector product.
// The GPU solve (solver.hip, k_nd_forward / k_nd_backward) processes all
// supernodes of one tree level per launch, in blocks of 64 rows.
#include "ndchol.hpp"

#include <algorithm>
#include <cmath>
#include <functional>

namespace gbte {
namespace {

constexpr int kLeaf = 96;

struct Dissect {
    const std::vector<std::vector<std::pair<int, double>>> &A;
    const std::vector<double> *X[3];
    int dim;
    std::vector<std::vector<int>> cells;   // per tree node
    std::vector<std::vector<int>> kids;
    std::vector<char> mark;

    int build(std::vector<int> S) {
        int id = (int)cells.size();
        cells.emplace_back();
        kids.emplace_back();
        if ((int)S.size() <= kLeaf) {
            cells[id] = std::move(S);
            return id;
        }
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int c : S)
            for (int k = 0; k < dim; ++k) {
                lo[k] = std::min(lo[k], (*X[k])[c]);
                hi[k] = std::max(hi[k], (*X[k])[c]);
            }
        int ax = 0;
        for (int k = 1; k < dim; ++k)
            if (hi[k] - lo[k] > hi[ax] - lo[ax]) ax = k;
        const std::vector<double> &coord = *X[ax];
        size_t half = S.size() / 2;
        std::nth_element(S.begin(), S.begin() + half, S.end(), [&](int a, int b) {
            return coord[a] < coord[b] || (coord[a] == coord[b] && a < b);
        });
        std::vector<int> L(S.begin(), S.begin() + half), R(S.begin() + half, S.end());
        for (int c : R) mark[c] = 1;
        std::vector<int> sep, Lr;
        for (int c : L) {
            bool touch = false;
            for (auto &e : A[c])
                if (e.first != c && mark[e.first]) { touch = true; break; }
            (touch ? sep : Lr).push_back(c);
        }
        for (int c : R) mark[c] = 0;
        std::sort(sep.begin(), sep.end());
        std::vector<int> ch;
        if (!Lr.empty()) ch.push_back(build(std::move(Lr)));
        if (!R.empty()) ch.push_back(build(std::move(R)));
        cells[id] = std::move(sep);
        kids[id] = ch;
        return id;
    }
};

}  // namespace

std::string nd_factor(const std::vector<std::vector<std::pair<int, double>>> &A, const std::vector<double> &x,
                      const std::vector<double> &y, const std::vector<double> &z, int dim, NDFactor &F) {
    int n = (int)A.size();
    F = NDFactor();
    F.n = n;
    Dissect D{A, {&x, &y, &z}, dim, {}, {}, std::vector<char>(n, 0)};
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) all[i] = i;
    int root = D.build(all);
    // postorder
    int nt = (int)D.cells.size();
    std::vector<int> order, post_of(nt, -1);
    std::function<void(int)> visit = [&](int k) {
        for (int c : D.kids[k]) visit(c);
        post_of[k] = (int)order.size();
        order.push_back(k);
    };
    visit(root);
    F.perm.resize(n);
    std::vector<int> pos(n);
    F.col0.resize(nt); F.ncol.resize(nt); F.nrow.resize(nt); F.parent.assign(nt, -1);
    F.height.assign(nt, 0); F.depth.assign(nt, 0);
    int p = 0;
    for (int k = 0; k < nt; ++k) {
        const auto &cs = D.cells[order[k]];
        F.col0[k] = p;
        F.ncol[k] = (int)cs.size();
        for (int c : cs) { F.perm[p] = c; pos[c] = p; ++p; }
    }
    F.child_ptr.assign(nt + 1, 0);
    for (int k = 0; k < nt; ++k) {
        for (int c : D.kids[order[k]]) {
            F.child_idx.push_back(post_of[c]);
            F.parent[post_of[c]] = k;
        }
        F.child_ptr[k + 1] = (int)F.child_idx.size();
    }
    for (int k = 0; k < nt; ++k)
        for (int q = F.child_ptr[k]; q < F.child_ptr[k + 1]; ++q)
            F.height[k] = std::max(F.height[k], F.height[F.child_idx[q]] + 1);
    for (int k = nt - 1; k >= 0; --k)
        if (F.parent[k] >= 0) F.depth[k] = F.depth[F.parent[k]] + 1;
    // symbolic: rows(k)
    std::vector<std::vector<int>> rows(nt);
    for (int k = 0; k < nt; ++k) {
        int hi = F.col0[k] + F.ncol[k];
        std::vector<int> r;
        for (int a = 0; a < F.ncol[k]; ++a)
            for (auto &e : A[F.perm[F.col0[k] + a]])
                if (pos[e.first] >= hi) r.push_back(pos[e.first]);
        for (int q = F.child_ptr[k]; q < F.child_ptr[k + 1]; ++q)
            for (int v : rows[F.child_idx[q]])
                if (v >= hi) r.push_back(v);
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        rows[k] = std::move(r);
        F.nrow[k] = (int)rows[k].size();
    }
    // numeric, multifrontal
    std::vector<std::vector<double>> U(nt);
    F.rows_off.assign(nt + 1, 0); F.f_off.assign(nt + 1, 0);
    F.u_off.assign(nt + 1, 0); F.map_off.assign(nt + 1, 0);
    for (int k = 0; k < nt; ++k) {
        F.rows_off[k + 1] = F.rows_off[k] + F.nrow[k];
        F.f_off[k + 1] = F.f_off[k] + (long)(F.ncol[k] + F.nrow[k]) * F.ncol[k];
        F.u_off[k + 1] = F.u_off[k] + F.nrow[k];
    }
    F.u_total = F.u_off[nt];
    F.rows.resize(F.rows_off[nt]);
    F.frm.assign(F.f_off[nt], 0.0);
    F.fcm.assign(F.f_off[nt], 0.0);
    std::vector<double> Li;
    std::vector<int> map_of_node(nt, -1);
    for (int k = 0; k < nt; ++k) {
        int nc_ = F.ncol[k], nr = F.nrow[k], m = nc_ + nr, c0 = F.col0[k], hi = c0 + nc_;
        F.max_front = std::max(F.max_front, m);
        std::copy(rows[k].begin(), rows[k].end(), F.rows.begin() + F.rows_off[k]);
        auto loc = [&](int pp) -> int {
            if (pp < hi) return pp - c0;
            return nc_ + (int)(std::lower_bound(rows[k].begin(), rows[k].end(), pp) - rows[k].begin());
        };
        std::vector<double> Fm((size_t)m * m, 0.0);
        for (int a = 0; a < nc_; ++a)
            for (auto &e : A[F.perm[c0 + a]]) {
                int pp = pos[e.first];
                if (pp >= c0 + a) Fm[(size_t)loc(pp) * m + a] += e.second;
            }
        for (int q = F.child_ptr[k]; q < F.child_ptr[k + 1]; ++q) {
            int c = F.child_idx[q];
            int nrc = F.nrow[c];
            std::vector<int> mp(nrc);
            for (int i = 0; i < nrc; ++i) mp[i] = loc(rows[c][i]);
            map_of_node[c] = (int)F.map.size();
            F.map.insert(F.map.end(), mp.begin(), mp.end());
            const std::vector<double> &Uc = U[c];
            for (int i = 0; i < nrc; ++i)
                for (int j = 0; j <= i; ++j) {
                    int li = mp[i], lj = mp[j];
                    if (li >= lj) Fm[(size_t)li * m + lj] += Uc[(size_t)i * nrc + j];
                    else Fm[(size_t)lj * m + li] += Uc[(size_t)i * nrc + j];
                }
            std::vector<double>().swap(U[c]);
        }
        // partial dense Cholesky of the first nc_ columns
        for (int a = 0; a < nc_; ++a) {
            double d = Fm[(size_t)a * m + a];
            if (!(d > 0) || !std::isfinite(d)) return "synthetic/Fourier matrix is not positive definite";
            double l = std::sqrt(d);
            Fm[(size_t)a * m + a] = l;
            for (int i = a + 1; i < m; ++i) Fm[(size_t)i * m + a] /= l;
            for (int j = a + 1; j < m; ++j) {
                double lj = Fm[(size_t)j * m + a];
                if (lj == 0) continue;
                for (int i = j; i < m; ++i) Fm[(size_t)i * m + j] -= Fm[(size_t)i * m + a] * lj;
            }
        }
        // inverse of L_kk (lower triangular)
        Li.assign((size_t)nc_ * nc_, 0.0);
        for (int col = 0; col < nc_; ++col) {
            Li[(size_t)col * nc_ + col] = 1.0 / Fm[(size_t)col * m + col];
            for (int i = col + 1; i < nc_; ++i) {
                double s = 0;
                for (int j = col; j < i; ++j) s += Fm[(size_t)i * m + j] * Li[(size_t)j * nc_ + col];
                Li[(size_t)i * nc_ + col] = -s / Fm[(size_t)i * m + i];
            }
        }
        // F = [Li; -L_A Li]
        double *Fr = F.frm.data() + F.f_off[k];
        for (int i = 0; i < nc_; ++i)
            for (int a = 0; a <= i; ++a) Fr[(size_t)i * nc_ + a] = Li[(size_t)i * nc_ + a];
        for (int r = 0; r < nr; ++r) {
            const double *la = &Fm[(size_t)(nc_ + r) * m];
            for (int a = 0; a < nc_; ++a) {
                double s = 0;
                for (int j = a; j < nc_; ++j) s += la[j] * Li[(size_t)j * nc_ + a];
                Fr[(size_t)(nc_ + r) * nc_ + a] = -s;
            }
        }
        double *Fc = F.fcm.data() + F.f_off[k];
        for (int i = 0; i < m; ++i)
            for (int a = 0; a < nc_; ++a) Fc[(size_t)a * m + i] = Fr[(size_t)i * nc_ + a];
        U[k].assign((size_t)nr * nr, 0.0);
        for (int i = 0; i < nr; ++i)
            for (int j = 0; j <= i; ++j) U[k][(size_t)i * nr + j] = Fm[(size_t)(nc_ + i) * m + nc_ + j];
    }
    // map offsets per node (as child of its parent)
    for (int k = 0; k < nt; ++k) F.map_off[k] = map_of_node[k] < 0 ? 0 : map_of_node[k];
    if (F.nrow[nt - 1] != 0) return "nested dissection: root supernode has ancestor rows";
    // self-check: A x = A 1 must give x = 1 to round-off
    std::vector<double> b(n, 0.0), xs;
    for (int i = 0; i < n; ++i)
        for (auto &e : A[i]) b[i] += e.second;
    nd_solve_host(F, b, xs);
    double err = 0;
    for (int i = 0; i < n; ++i) err = std::max(err, std::fabs(xs[i] - 1.0));
    if (!(err < 1e-8)) return "nested-dissection factorization failed its self-check";
    return "";
}

void nd_solve_host(const NDFactor &F, const std::vector<double> &b, std::vector<double> &x) {
    int nt = (int)F.ncol.size();
    std::vector<double> y(F.n), u(F.u_total), xp(F.n);
    std::vector<double> f, v;
    for (int k = 0; k < nt; ++k) {  // postorder = a valid forward order
        int nc_ = F.ncol[k], nr = F.nrow[k], m = nc_ + nr;
        f.assign(m, 0.0);
        for (int a = 0; a < nc_; ++a) f[a] = b[F.perm[F.col0[k] + a]];
        for (int q = F.child_ptr[k]; q < F.child_ptr[k + 1]; ++q) {
            int c = F.child_idx[q];
            for (int i = 0; i < F.nrow[c]; ++i) f[F.map[F.map_off[c] + i]] += u[F.u_off[c] + i];
        }
        const double *Fr = F.frm.data() + F.f_off[k];
        for (int i = 0; i < m; ++i) {
            double s = 0;
            for (int a = 0; a < nc_; ++a) s += Fr[(size_t)i * nc_ + a] * f[a];
            if (i < nc_) y[F.col0[k] + i] = s;
            else u[F.u_off[k] + i - nc_] = f[i] + s;
        }
    }
    for (int k = nt - 1; k >= 0; --k) {  // reverse postorder = a valid backward order
        int nc_ = F.ncol[k], nr = F.nrow[k], m = nc_ + nr;
        const double *Fr = F.frm.data() + F.f_off[k];
        v.assign(m, 0.0);
        for (int a = 0; a < nc_; ++a) v[a] = y[F.col0[k] + a];
        for (int r = 0; r < nr; ++r) v[nc_ + r] = xp[F.rows[F.rows_off[k] + r]];
        for (int a = 0; a < nc_; ++a) {
            double s = 0;
            for (int j = 0; j < m; ++j) s += Fr[(size_t)j * nc_ + a] * v[j];
            xp[F.col0[k] + a] = s;
        }
    }
    x.assign(F.n, 0.0);
    for (int p = 0; p < F.n; ++p) x[F.perm[p]] = xp[p];
}

}  // namespace gbte
