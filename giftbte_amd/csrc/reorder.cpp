// Locality-preserving renumbering of cells and nodes for the device layout.
//
// The deck's cell order (a Delaunay generator's simplex order, a COMSOL
// export order) scatters neighbouring cells across memory; every gather in
// the solver (upwind neighbours, LSQ stencils, inverse-distance vertex
// values) then touches unrelated cache lines.  Cells and nodes are sorted
// along a Morton (Z-order) curve of their coordinates; all tables are
// relabelled consistently and the per-cell lists keep their order, so every
// per-cell sum is evaluated in the same order as before.  Results are mapped
// back to deck order at the C ABI.
#include <algorithm>
#include <cstdint>

#include "problem.hpp"

namespace gbte {
namespace {

uint64_t spread2(uint64_t v) {
    v &= 0xffffffffull;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

uint64_t spread3(uint64_t v) {
    v &= 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// new_of_old for points sorted along the Morton curve
std::vector<int> morton_order(const std::vector<double> &x, const std::vector<double> &y, const std::vector<double> &z,
                              int dim) {
    size_t n = x.size();
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    const std::vector<double> *c[3] = {&x, &y, &z};
    for (int k = 0; k < 3; ++k)
        for (size_t i = 0; i < n; ++i) {
            lo[k] = std::min(lo[k], (*c[k])[i]);
            hi[k] = std::max(hi[k], (*c[k])[i]);
        }
    double scale = 0;
    for (int k = 0; k < dim; ++k) scale = std::max(scale, hi[k] - lo[k]);
    if (scale <= 0) scale = 1;
    const double bits = dim == 2 ? 4294967295.0 : 2097151.0;
    std::vector<std::pair<uint64_t, int>> key(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t q[3];
        for (int k = 0; k < 3; ++k) q[k] = (uint64_t)(((*c[k])[i] - lo[k]) / scale * bits);
        key[i].first = dim == 2 ? (spread2(q[0]) | (spread2(q[1]) << 1))
                                : (spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2));
        key[i].second = (int)i;
    }
    std::stable_sort(key.begin(), key.end(), [](const std::pair<uint64_t, int> &a, const std::pair<uint64_t, int> &b) {
        return a.first < b.first;
    });
    std::vector<int> new_of_old(n);
    for (size_t k = 0; k < n; ++k) new_of_old[key[k].second] = (int)k;
    return new_of_old;
}

template <typename T>
void gather_rows(const std::vector<T> &src, std::vector<T> &dst, const std::vector<int> &old_of_new, size_t width) {
    dst.resize(src.size());
    for (size_t i = 0; i < old_of_new.size(); ++i)
        std::copy(src.begin() + (size_t)old_of_new[i] * width, src.begin() + (size_t)(old_of_new[i] + 1) * width,
                  dst.begin() + i * width);
}

}  // namespace

Problem renumber(const Problem &p, std::vector<int> &cell_new, std::vector<int> &node_new) {
    cell_new = morton_order(p.cx, p.cy, p.cz, p.dim);
    node_new = morton_order(p.node_x, p.node_y, p.node_z, p.dim);
    int nc = p.n_cells, nn = p.n_nodes;
    std::vector<int> cell_old(nc), node_old(nn);
    for (int i = 0; i < nc; ++i) cell_old[cell_new[i]] = i;
    for (int i = 0; i < nn; ++i) node_old[node_new[i]] = i;
    Problem q = p;
    auto cmap = [&](int c) { return c >= 0 ? cell_new[c] : c; };
    auto nmap = [&](int v) { return v >= 0 ? node_new[v] : v; };
    gather_rows(p.face_size, q.face_size, cell_old, 1);
    gather_rows(p.face_neighbor, q.face_neighbor, cell_old, 6);
    for (auto &v : q.face_neighbor) v = cmap(v);
    gather_rows(p.face_bound, q.face_bound, cell_old, 6);
    gather_rows(p.face_index, q.face_index, cell_old, 6);
    gather_rows(p.face_nvert, q.face_nvert, cell_old, 6);
    gather_rows(p.face_vert, q.face_vert, cell_old, 24);
    for (auto &v : q.face_vert) v = nmap(v);
    gather_rows(p.cell_nvert, q.cell_nvert, cell_old, 1);
    gather_rows(p.cell_vert, q.cell_vert, cell_old, 8);
    for (auto &v : q.cell_vert) v = nmap(v);
    gather_rows(p.matter, q.matter, cell_old, 1);
    gather_rows(p.cell_region, q.cell_region, cell_old, 1);
    for (auto pr : {std::make_pair(&p.face_area, &q.face_area), std::make_pair(&p.face_nx, &q.face_nx),
                    std::make_pair(&p.face_ny, &q.face_ny), std::make_pair(&p.face_nz, &q.face_nz),
                    std::make_pair(&p.face_cx, &q.face_cx), std::make_pair(&p.face_cy, &q.face_cy),
                    std::make_pair(&p.face_cz, &q.face_cz)})
        gather_rows(*pr.first, *pr.second, cell_old, 6);
    for (auto pr : {std::make_pair(&p.volume, &q.volume), std::make_pair(&p.cx, &q.cx), std::make_pair(&p.cy, &q.cy),
                    std::make_pair(&p.cz, &q.cz), std::make_pair(&p.heat_source, &q.heat_source),
                    std::make_pair(&p.initial_temp, &q.initial_temp)})
        gather_rows(*pr.first, *pr.second, cell_old, 1);
    q.nbr_ptr.assign(nc + 1, 0);
    q.nbr_idx.clear();
    for (int i = 0; i < nc; ++i) {
        int o = cell_old[i];
        for (int k = p.nbr_ptr[o]; k < p.nbr_ptr[o + 1]; ++k) q.nbr_idx.push_back(cell_new[p.nbr_idx[k]]);
        q.nbr_ptr[i + 1] = (int)q.nbr_idx.size();
    }
    gather_rows(p.node_x, q.node_x, node_old, 1);
    gather_rows(p.node_y, q.node_y, node_old, 1);
    gather_rows(p.node_z, q.node_z, node_old, 1);
    for (auto &v : q.bound_cell) v = cmap(v);
    for (auto &v : q.fface_cell) v = cmap(v);
    return q;
}

}  // namespace gbte
