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
constexpr int kPanel = 32;    // pivot columns per trailing update of a front
constexpr int kGroup = 8;     // columns updated together (one pass over a pivot column); the
                              // trailing update's 8-way kernel assumes 8
constexpr int kRowTile = 256; // rows of -L_A Li per cache tile

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

// body(g) for g in [0, n) on the host's cores when the total work (about
// `work` multiply-adds) is large; contiguous ranges of g per thread
template <typename Body>
void parallel_groups(int n, double work, Body body) {
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    int nth = t_serial_cols ? 1 : (int)std::min<double>(std::min<double>(hw, n), std::max(1.0, work / 200000.0));
    if (nth <= 1) {
        for (int g = 0; g < n; ++g) body(g);
        return;
    }
    std::vector<std::thread> th;
    for (int q = 0; q < nth; ++q)
        th.emplace_back([&, q]() {
            for (int g = q; g < n; g += nth) body(g);
        });
    for (auto &x : th) x.join();
}

// One group of kGroup trailing columns j0.. of a front (column-major, m
// rows) updated by the pivot columns a0..a1-1: C[i][j] -= C[i][a] C[j][a],
// in register blocks of 4 rows x kGroup columns that take all pivot columns
// before they are stored (each entry still takes the pivots in increasing
// order, so the result does not depend on the blocking).  Built for the
// host's vector units where they have AVX2 / FMA-class instructions, with a
// baseline copy chosen at load time otherwise.
__attribute__((target_clones("arch=x86-64-v3", "default"))) void trailing_group(double *Cp, int m, int a0, int a1,
                                                                               int j0) {
    int nj = std::min(kGroup, m - j0), np = a1 - a0;
    double *cj[kGroup];
    for (int q = 0; q < kGroup; ++q) cj[q] = Cp + (size_t)(j0 + std::min(q, nj - 1)) * m;
    if (nj < kGroup) {
        for (int a = a0; a < a1; ++a) {
            const double *ca = Cp + (size_t)a * m;
            for (int q = 0; q < nj; ++q) {
                double lj = ca[j0 + q];
                for (int i = j0; i < m; ++i) cj[q][i] -= ca[i] * lj;
            }
        }
        return;
    }
    // the pivots' entries on the group's rows, [a][q]
    double lt[kPanel][kGroup];
    for (int a = 0; a < np; ++a)
        for (int q = 0; q < kGroup; ++q) lt[a][q] = Cp[(size_t)(a0 + a) * m + j0 + q];
    int i = j0;
    for (; i + 4 <= m; i += 4) {
        double acc[4][kGroup];
        for (int q = 0; q < kGroup; ++q)
            for (int r = 0; r < 4; ++r) acc[r][q] = cj[q][i + r];
        for (int a = 0; a < np; ++a) {
            const double *ca = Cp + (size_t)(a0 + a) * m + i;
            double v0 = ca[0], v1 = ca[1], v2 = ca[2], v3 = ca[3];
            for (int q = 0; q < kGroup; ++q) {
                double l = lt[a][q];
                acc[0][q] -= v0 * l;
                acc[1][q] -= v1 * l;
                acc[2][q] -= v2 * l;
                acc[3][q] -= v3 * l;
            }
        }
        for (int q = 0; q < kGroup; ++q)
            for (int r = 0; r < 4; ++r) cj[q][i + r] = acc[r][q];
    }
    for (; i < m; ++i)
        for (int q = 0; q < kGroup; ++q) {
            double c = cj[q][i];
            for (int a = 0; a < np; ++a) c -= Cp[(size_t)(a0 + a) * m + i] * lt[a][q];
            cj[q][i] = c;
        }
}

// Columns col0.. (kGroup of them) of inv(L_kk) by forward substitution,
// L_kk the first nc columns of a front (column-major, m rows); each column
// is computed exactly as alone (entries in increasing j), the group only
// sharing the reads of L.
__attribute__((target_clones("arch=x86-64-v3", "default"))) void inverse_group(const double *Cp, int m, int nc,
                                                                              double *Li, int col0) {
    int ng = std::min(kGroup, nc - col0);
    double *x[kGroup];
    for (int q = 0; q < kGroup; ++q) x[q] = Li + (size_t)(col0 + std::min(q, ng - 1)) * nc;
    for (int q = 0; q < ng; ++q) x[q][col0 + q] = 1.0;
    for (int j = col0; j < nc; ++j) {
        const double *cj = Cp + (size_t)j * m;
        int act = std::min(ng, j - col0 + 1);  // columns col0 .. col0 + act - 1 have reached row j
        double xj[kGroup];
        for (int q = 0; q < act; ++q) {
            double *xq = x[q];
            xq[j] = j == col0 + q ? 1.0 / cj[j] : xq[j] / cj[j];
            xj[q] = xq[j];
        }
        if (act == kGroup) {
            double *x0 = x[0], *x1 = x[1], *x2 = x[2], *x3 = x[3], *x4 = x[4], *x5 = x[5], *x6 = x[6], *x7 = x[7];
            for (int i = j + 1; i < nc; ++i) {
                double c = cj[i];
                x0[i] -= c * xj[0]; x1[i] -= c * xj[1]; x2[i] -= c * xj[2]; x3[i] -= c * xj[3];
                x4[i] -= c * xj[4]; x5[i] -= c * xj[5]; x6[i] -= c * xj[6]; x7[i] -= c * xj[7];
            }
        } else {
            for (int q = 0; q < act; ++q) {
                double *xq = x[q];
                for (int i = j + 1; i < nc; ++i) xq[i] -= cj[i] * xj[q];
            }
        }
    }
}

// Rows nc.. of F = -L_A inv(L_kk) for the kGroup columns a0g.., in tiles of
// kRowTile rows: each L_A column tile is read once per group, and each
// entry sums its terms in increasing j (the zeros of inv(L_kk) above its
// diagonal are skipped, as alone).
__attribute__((target_clones("arch=x86-64-v3", "default"))) void lali_group(const double *Cp, int m, int nc, int nr,
                                                                           const double *Li, double *Fr, int a0g) {
    int na = std::min(kGroup, nc - a0g);
    double acc[kGroup][kRowTile];
    for (int r0 = 0; r0 < nr; r0 += kRowTile) {
        int rn = std::min(kRowTile, nr - r0);
        for (int q = 0; q < kGroup; ++q)
            for (int r = 0; r < rn; ++r) acc[q][r] = 0.0;
        for (int j = a0g; j < nc; ++j) {
            const double *cj = Cp + (size_t)j * m + nc + r0;
            if (na == kGroup && j >= a0g + kGroup - 1) {
                double l[kGroup];
                for (int q = 0; q < kGroup; ++q) l[q] = Li[(size_t)(a0g + q) * nc + j];
                for (int r = 0; r < rn; ++r) {
                    double c = cj[r];
                    for (int q = 0; q < kGroup; ++q) acc[q][r] += c * l[q];
                }
            } else {
                for (int q = 0; q < na; ++q) {
                    if (j < a0g + q) continue;
                    double l = Li[(size_t)(a0g + q) * nc + j];
                    for (int r = 0; r < rn; ++r) acc[q][r] += cj[r] * l;
                }
            }
        }
        for (int q = 0; q < na; ++q)
            for (int r = 0; r < rn; ++r) Fr[(size_t)(nc + r0 + r) * nc + a0g + q] = -acc[q][r];
    }
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
            // the child's update matrix, lower triangle column-major; each
            // front entry receives at most one term from a child, so the
            // order of the loops does not change the sums
            const std::vector<double> &Uc = U[c];
            for (int j = 0; j < nrc; ++j) {
                int lj = mp[j];
                const double *uj = &Uc[(size_t)j * nrc];
                for (int i = j; i < nrc; ++i) {
                    int li = mp[i];
                    if (li >= lj) C[(size_t)lj * m + li] += uj[i];
                    else C[(size_t)li * m + lj] += uj[i];
                }
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
            // trailing columns in groups of kGroup (rows above a column's
            // diagonal lie in the unused upper triangle)
            parallel_groups((m - a1 + kGroup - 1) / kGroup, 0.5 * (double)(m - a1) * (m - a1) * (a1 - a0), [&](int g) {
                trailing_group(C.data(), m, a0, a1, a1 + g * kGroup);
            });
        }
        // inverse of L_kk (lower triangular), column-major: Li[a * nc_ + i];
        // column a by forward substitution, the terms of each entry in
        // increasing order
        Li.assign((size_t)nc_ * nc_, 0.0);
        parallel_groups((nc_ + kGroup - 1) / kGroup, (double)nc_ * nc_ * nc_ / 6.0,
                        [&](int g) { inverse_group(C.data(), m, nc_, Li.data(), g * kGroup); });
        // F = [Li; -L_A Li]
        double *Fr = F.frm.data() + F.f_off[k];
        for (int i = 0; i < nc_; ++i)
            for (int a = 0; a <= i; ++a) Fr[(size_t)i * nc_ + a] = Li[(size_t)a * nc_ + i];
        // -L_A Li in groups of kGroup columns
        parallel_groups((nc_ + kGroup - 1) / kGroup, 0.5 * (double)nc_ * nc_ * nr,
                        [&](int g) { lali_group(C.data(), m, nc_, nr, Li.data(), Fr, g * kGroup); });
        // the same F column-major, transposed in 32 x 32 tiles
        double *Fc = F.fcm.data() + F.f_off[k];
        parallel_groups((m + 31) / 32, 4.0 * (double)m * nc_, [&](int ti) {
            int i0 = ti * 32, i1 = std::min(m, i0 + 32);
            for (int a0t = 0; a0t < nc_; a0t += 32) {
                int a1t = std::min(nc_, a0t + 32);
                for (int i = i0; i < i1; ++i)
                    for (int a = a0t; a < a1t; ++a) Fc[(size_t)a * m + i] = Fr[(size_t)i * nc_ + a];
            }
        });
        U[k].assign((size_t)nr * nr, 0.0);
        for (int j = 0; j < nr; ++j) {
            const double *cj = &C[(size_t)(nc_ + j) * m + nc_];
            std::copy(cj + j, cj + nr, U[k].begin() + (size_t)j * nr + j);
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
