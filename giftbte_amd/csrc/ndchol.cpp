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
#include <thread>

namespace gbte {
namespace {

constexpr int kLeaf = 96;
constexpr int kPanel = 32;  // pivot columns per trailing update of a front

// body(j) for j in [lo, hi) on the host's cores when the work is large
// (columns are independent); the split balances sum (len - j).
// set on the host threads that factorize whole subtrees, which then run
// their fronts' columns serially
thread_local bool t_serial_cols = false;

template <typename Body>
void parallel_cols(int lo, int hi, int len, Body body) {
    long work = 0;
    for (int j = lo; j < hi; ++j) work += len - j;
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int nth = t_serial_cols ? 1 : (int)std::min<long>(hw, std::max<long>(1, work / 200000));
    if (nth <= 1 || hi - lo < 2 * nth) {
        for (int j = lo; j < hi; ++j) body(j);
        return;
    }
    std::vector<int> cut(nth + 1, hi);
    cut[0] = lo;
    long acc = 0, per = (work + nth - 1) / nth;
    int t = 1;
    for (int j = lo; j < hi && t < nth; ++j) {
        acc += len - j;
        if (acc >= per * t) cut[t++] = j + 1;
    }
    std::vector<std::thread> th;
    for (int q = 0; q < nth; ++q)
        if (cut[q] < cut[q + 1])
            th.emplace_back([&, q]() {
                for (int j = cut[q]; j < cut[q + 1]; ++j) body(j);
            });
    for (auto &x : th) x.join();
}

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
    // where each child's map of its update rows into its parent's front
    // goes: the children of every node in postorder, in child order
    std::vector<int> map_of_node(nt, -1);
    {
        int run = 0;
        for (int k = 0; k < nt; ++k)
            for (int q = F.child_ptr[k]; q < F.child_ptr[k + 1]; ++q) {
                map_of_node[F.child_idx[q]] = run;
                run += F.nrow[F.child_idx[q]];
            }
        F.map.assign(run, 0);
    }
    for (int k = 0; k < nt; ++k) F.max_front = std::max(F.max_front, F.ncol[k] + F.nrow[k]);
    // one supernode: its front (the rows of A and the children's update
    // matrices), partial Cholesky, F_k and its own update matrix.  A node
    // needs only its children, so disjoint subtrees are independent.
    auto factor_node = [&](int k) -> std::string {
        std::vector<double> Li;
        int nc_ = F.ncol[k], nr = F.nrow[k], m = nc_ + nr, c0 = F.col0[k], hi = c0 + nc_;
        std::copy(rows[k].begin(), rows[k].end(), F.rows.begin() + F.rows_off[k]);
        auto loc = [&](int pp) -> int {
            if (pp < hi) return pp - c0;
            return nc_ + (int)(std::lower_bound(rows[k].begin(), rows[k].end(), pp) - rows[k].begin());
        };
        // the front, lower triangle column-major: element (i, j), i >= j, at C[j * m + i]
        std::vector<double> C((size_t)m * m, 0.0);
        for (int a = 0; a < nc_; ++a)
            for (auto &e : A[F.perm[c0 + a]]) {
                int pp = pos[e.first];
                if (pp >= c0 + a) C[(size_t)a * m + loc(pp)] += e.second;
            }
        for (int q = F.child_ptr[k]; q < F.child_ptr[k + 1]; ++q) {
            int c = F.child_idx[q];
            int nrc = F.nrow[c];
            std::vector<int> mp(nrc);
            for (int i = 0; i < nrc; ++i) mp[i] = loc(rows[c][i]);
            std::copy(mp.begin(), mp.end(), F.map.begin() + map_of_node[c]);
            const std::vector<double> &Uc = U[c];
            for (int i = 0; i < nrc; ++i)
                for (int j = 0; j <= i; ++j) {
                    int li = mp[i], lj = mp[j];
                    if (li >= lj) C[(size_t)lj * m + li] += Uc[(size_t)i * nrc + j];
                    else C[(size_t)li * m + lj] += Uc[(size_t)i * nrc + j];
                }
            std::vector<double>().swap(U[c]);
        }
        // partial dense Cholesky of the first nc_ columns, right-looking in
        // panels of kPanel columns; every entry is updated by the pivot
        // columns in increasing order, as in the unblocked form
        for (int a0 = 0; a0 < nc_; a0 += kPanel) {
            int a1 = std::min(nc_, a0 + kPanel);
            for (int a = a0; a < a1; ++a) {
                double *ca = &C[(size_t)a * m];
                double d = ca[a];
                if (!(d > 0) || !std::isfinite(d)) return "synthetic/Fourier matrix is not positive definite";
                double l = std::sqrt(d);
                ca[a] = l;
                for (int i = a + 1; i < m; ++i) ca[i] /= l;
                for (int j = a + 1; j < a1; ++j) {
                    double lj = ca[j];
                    if (lj == 0) continue;
                    double *cj = &C[(size_t)j * m];
                    for (int i = j; i < m; ++i) cj[i] -= ca[i] * lj;
                }
            }
            parallel_cols(a1, m, m, [&](int j) {
                double *cj = &C[(size_t)j * m];
                for (int a = a0; a < a1; ++a) {
                    const double *ca = &C[(size_t)a * m];
                    double lj = ca[j];
                    if (lj == 0) continue;
                    for (int i = j; i < m; ++i) cj[i] -= ca[i] * lj;
                }
            });
        }
        // inverse of L_kk (lower triangular), column-major: Li[a * nc_ + i];
        // column a by forward substitution, the terms of each entry in
        // increasing order
        Li.assign((size_t)nc_ * nc_, 0.0);
        parallel_cols(0, nc_, nc_, [&](int col) {
            double *x = &Li[(size_t)col * nc_];
            x[col] = 1.0;
            for (int j = col; j < nc_; ++j) {
                const double *cj = &C[(size_t)j * m];
                if (j == col) {
                    x[j] = 1.0 / cj[j];
                } else {
                    x[j] = x[j] / cj[j];
                }
                double xj = x[j];
                for (int i = j + 1; i < nc_; ++i) x[i] -= cj[i] * xj;
            }
        });
        // F = [Li; -L_A Li]
        double *Fr = F.frm.data() + F.f_off[k];
        for (int i = 0; i < nc_; ++i)
            for (int a = 0; a <= i; ++a) Fr[(size_t)i * nc_ + a] = Li[(size_t)a * nc_ + i];
        parallel_cols(0, nc_, nc_, [&](int a) {
            std::vector<double> acc(nr, 0.0);
            const double *la = &Li[(size_t)a * nc_];
            for (int j = a; j < nc_; ++j) {
                double lv = la[j];
                const double *cj = &C[(size_t)j * m + nc_];
                for (int r = 0; r < nr; ++r) acc[r] += cj[r] * lv;
            }
            for (int r = 0; r < nr; ++r) Fr[(size_t)(nc_ + r) * nc_ + a] = -acc[r];
        });
        double *Fc = F.fcm.data() + F.f_off[k];
        for (int i = 0; i < m; ++i)
            for (int a = 0; a < nc_; ++a) Fc[(size_t)a * m + i] = Fr[(size_t)i * nc_ + a];
        U[k].assign((size_t)nr * nr, 0.0);
        for (int j = 0; j < nr; ++j) {
            const double *cj = &C[(size_t)(nc_ + j) * m + nc_];
            for (int i = j; i < nr; ++i) U[k][(size_t)i * nr + j] = cj[i];
        }
        return std::string();
    };
    // the subtrees below the first depth with at least 64 nodes on host
    // threads (each subtree in postorder, columns serially), then the nodes
    // above them in postorder with the columns of each front in parallel;
    // the arithmetic of every node is the same either way
    {
        std::vector<int> size(nt, 1);
        for (int k = 0; k < nt; ++k)
            for (int q = F.child_ptr[k]; q < F.child_ptr[k + 1]; ++q) size[k] += size[F.child_idx[q]];
        int maxd = 0;
        for (int k = 0; k < nt; ++k) maxd = std::max(maxd, F.depth[k]);
        int D = maxd + 1;
        for (int dd = 0; dd <= maxd; ++dd) {
            int cntd = 0;
            for (int k = 0; k < nt; ++k) cntd += F.depth[k] == dd;
            if (cntd >= 64) { D = dd; break; }
        }
        std::vector<int> roots;
        for (int k = 0; k < nt; ++k)
            if (F.depth[k] == D) roots.push_back(k);
        std::vector<std::string> rerr(roots.size());
        unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        int nth = (int)std::min<size_t>(std::min(16u, hw), roots.size());
        if (nth > 0) {
            std::vector<std::thread> th;
            for (int w = 0; w < nth; ++w)
                th.emplace_back([&, w]() {
                    t_serial_cols = true;
                    for (size_t i = w; i < roots.size(); i += nth) {
                        int r = roots[i];
                        for (int k = r - size[r] + 1; k <= r && rerr[i].empty(); ++k) rerr[i] = factor_node(k);
                    }
                });
            for (auto &x : th) x.join();
        }
        for (auto &e : rerr)
            if (!e.empty()) return e;
        for (int k = 0; k < nt; ++k)
            if (F.depth[k] < D) {
                std::string e = factor_node(k);
                if (!e.empty()) return e;
            }
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
