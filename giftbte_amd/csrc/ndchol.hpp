// Nested-dissection supernodal Cholesky of the (negated) synthetic / Fourier
// matrices, factored once on the host; the triangular solves run on the GPU,
// one launch per elimination-tree level.
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace gbte {

struct NDFactor {
    int n = 0;
    std::vector<int> perm;  // perm[position] = row of the original matrix
    // elimination-tree nodes in postorder (children before parents)
    std::vector<int> col0, ncol, nrow, parent, height, depth;
    std::vector<long> rows_off, f_off, u_off, map_off;
    std::vector<int> rows;        // A(k): positions (> last column of k), sorted
    // per node, F = [inv(L_kk); -L_{A(k),k} inv(L_kk)], (ncol + nrow) x ncol,
    // stored row-major (frm) and column-major (fcm):
    //   forward   y = inv(L_kk) f_cols,  u = f_A - L_Ak y  =  F f_cols (+ [0; f_A])
    //   backward  x = inv(L_kk)^T (y - L_Ak^T x_A)        =  F^T [y; x_A]
    std::vector<double> frm, fcm;
    std::vector<int> child_ptr, child_idx;  // children of each node
    std::vector<int> map;         // per child c: local index of rows(c) in its parent's front
    long u_total = 0;
    int max_front = 0;
};

// rows[i] = (column, value) entries of a symmetric positive definite matrix;
// x, y, z are the coordinates used for the geometric bisection.
std::string nd_factor(const std::vector<std::vector<std::pair<int, double>>> &rows, const std::vector<double> &x,
                      const std::vector<double> &y, const std::vector<double> &z, int dim, NDFactor &F);

// Host restatement of the device solve (forward by postorder, backward by
// reverse postorder); used as the factorization's self-check.
void nd_solve_host(const NDFactor &F, const std::vector<double> &b, std::vector<double> &x);

}  // namespace gbte
