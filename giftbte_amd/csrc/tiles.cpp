// Cell tiles for the LDS-staged face / vertex / moment kernel.
//
// Cells are in Morton order, so a run of consecutive cells is a compact
// patch.  Each tile is such a run plus the rows its computation touches:
//   H  every cell of the tile and every LSQ-stencil neighbour (cells sharing a
//      vertex, which includes the face neighbours and every cell around each
//      vertex of the tile),
//   B  the boundary entries of the tile's (non-specular) boundary faces and
//      the entries that feed the inverse-distance values of the tile's
//      boundary vertices,
//   V  the two vertices of every face-gradient face of the tile (in the order
//      of the face's first owner), inverse-distance combinations of H and B
//      rows.
// The device stages [H | B] row slices in LDS and reads every neighbour value
// from there; a vertex enters only through its moments, which are the same
// combination of its sources' moments.  Tiles grow greedily along the Morton
// order until the staged rows would exceed max_rows or the vertices kYRows.
#include "tiles.hpp"

#include <algorithm>
#include <unordered_map>

namespace gbte {

std::string build_tiles(const Problem &p, const std::vector<int> &gptr, const std::vector<int> &gidx,
                        const std::vector<int> &vptr, const std::vector<int> &vsrc, const std::vector<double> &vwt,
                        const std::vector<int> &fkind, const std::vector<int> &fother, const std::vector<int> &fv0,
                        const std::vector<int> &fv1, int max_rows, int max_cells, Tiles &T) {
    int nc = p.n_cells;
    T = Tiles();
    T.cell0.push_back(0);
    T.rptr.push_back(0);
    T.vptr.push_back(0);
    T.vs_ptr.push_back(0);
    T.crow.assign(nc, 0);
    T.st_row.assign(gidx.size(), 0);
    T.f_orow.assign((size_t)nc * 6, 0);
    T.f_v0row.assign((size_t)nc * 6, 0);
    T.f_v1row.assign((size_t)nc * 6, 0);

    struct Sets {
        std::vector<int> H, B, V;
        std::unordered_map<int, int> hpos, bpos, vpos;
        void clear() { H.clear(); B.clear(); V.clear(); hpos.clear(); bpos.clear(); vpos.clear(); }
    } cur, trial;

    auto add_cell = [&](Sets &S, int c) {
        auto addH = [&](int x) { if (!S.hpos.count(x)) { S.hpos[x] = (int)S.H.size(); S.H.push_back(x); } };
        auto addB = [&](int x) { if (!S.bpos.count(x)) { S.bpos[x] = (int)S.B.size(); S.B.push_back(x); } };
        auto addV = [&](int x) { if (!S.vpos.count(x)) { S.vpos[x] = (int)S.V.size(); S.V.push_back(x); } };
        addH(c);
        for (int k = gptr[c]; k < gptr[c + 1]; ++k) addH(gidx[k]);
        for (int j = 0; j < p.face_size[c]; ++j) {
            size_t s = (size_t)c * 6 + j;
            int kd = fkind[s];
            if (kd == 0) continue;
            if (kd == 2) addB(fother[s]);
            else addH(fother[s]);
            for (int q = 0; q < 2; ++q) {
                int v = q == 0 ? fv0[s] : fv1[s];
                if (S.vpos.count(v)) continue;
                addV(v);
                for (int k = vptr[v]; k < vptr[v + 1]; ++k) {
                    if (vsrc[k] >= 0) addH(vsrc[k]);
                    else addB(-vsrc[k] - 1);
                }
            }
        }
    };
    auto rows = [](const Sets &S) { return (int)(S.H.size() + S.B.size()); };
    auto fits = [&](const Sets &S) { return rows(S) <= max_rows && (int)S.V.size() <= kYRows; };

    auto close_tile = [&](int c_end) {
        const Sets &S = cur;
        int nh = (int)S.H.size(), nbr = (int)S.B.size();
        for (int x : S.H) T.rsrc.push_back(x);
        for (int x : S.B) T.rsrc.push_back(-x - 1);
        T.rptr.push_back((int)T.rsrc.size());
        T.nh.push_back(nh);
        int base_v = nh + nbr;
        for (int v : S.V) {
            T.vnode.push_back(v);
            int first = 0;
            for (int k = vptr[v]; k < vptr[v + 1]; ++k) {
                int r = vsrc[k] >= 0 ? S.hpos.at(vsrc[k]) : nh + S.bpos.at(-vsrc[k] - 1);
                if (k == vptr[v]) first = r;
                T.vs_row.push_back(r);
                T.vs_w.push_back(vwt[k]);
            }
            // pad to a multiple of 4 with the first source at weight 0, so
            // the device sums in batches of 4 (x * 0 adds an exact zero
            // unless x is not finite, which already makes the sum so)
            while ((T.vs_row.size() - T.vs_ptr.back()) % 4) {
                T.vs_row.push_back(first);
                T.vs_w.push_back(0.0);
            }
            T.vs_ptr.push_back((int)T.vs_row.size());
        }
        T.vptr.push_back((int)T.vnode.size());
        for (int c = T.cell0.back(); c < c_end; ++c) {
            T.crow[c] = S.hpos.at(c);
            for (int k = gptr[c]; k < gptr[c + 1]; ++k) T.st_row[k] = S.hpos.at(gidx[k]);
            for (int j = 0; j < p.face_size[c]; ++j) {
                size_t s = (size_t)c * 6 + j;
                int kd = fkind[s];
                if (kd == 0) continue;
                T.f_orow[s] = kd == 2 ? nh + S.bpos.at(fother[s]) : S.hpos.at(fother[s]);
                T.f_v0row[s] = base_v + S.vpos.at(fv0[s]);
                T.f_v1row[s] = base_v + S.vpos.at(fv1[s]);
            }
        }
        T.max_rows = std::max(T.max_rows, rows(S));
        T.cell0.push_back(c_end);
    };

    int count = 0;
    for (int c = 0; c < nc; ++c) {
        trial = cur;
        add_cell(trial, c);
        if (count > 0 && (!fits(trial) || count >= max_cells)) {
            close_tile(c);
            cur.clear();
            count = 0;
            trial = cur;
            add_cell(trial, c);
        }
        if (!fits(trial)) return "a cell's stencil exceeds the LDS tile";
        cur = std::move(trial);
        ++count;
    }
    if (count > 0) close_tile(nc);
    T.n_tiles = (int)T.cell0.size() - 1;
    // every tile-local row the device reads must lie inside the tile and
    // name the intended cell, entry or vertex
    for (int t = 0; t < T.n_tiles; ++t) {
        int r0 = T.rptr[t], nr = T.rptr[t + 1] - r0, v0 = T.vptr[t], nv = T.vptr[t + 1] - v0;
        if (nr > max_rows || nv > kYRows) return "tile exceeds its row budget";
        auto staged = [&](int r, int want) { return r >= 0 && r < nr && T.rsrc[r0 + r] == want; };
        auto vrow = [&](int r, int node) { return r >= nr && r < nr + nv && T.vnode[v0 + r - nr] == node; };
        for (int i = 0; i < nv; ++i) {
            int v = T.vnode[v0 + i];
            int padded = (vptr[v + 1] - vptr[v] + 3) / 4 * 4;
            if (T.vs_ptr[v0 + i + 1] - T.vs_ptr[v0 + i] != padded) return "tile vertex stencil mismatch";
            for (int k = vptr[v + 1] - vptr[v]; k < padded; ++k)
                if (T.vs_w[T.vs_ptr[v0 + i] + k] != 0.0 || T.vs_row[T.vs_ptr[v0 + i] + k] < 0 ||
                    T.vs_row[T.vs_ptr[v0 + i] + k] >= nr)
                    return "tile vertex padding mismatch";
            for (int k = 0; k < vptr[v + 1] - vptr[v]; ++k) {
                int src = vsrc[vptr[v] + k];
                if (!staged(T.vs_row[T.vs_ptr[v0 + i] + k], src)) return "tile vertex source mismatch";
            }
        }
        for (int c = T.cell0[t]; c < T.cell0[t + 1]; ++c) {
            if (!staged(T.crow[c], c)) return "tile cell row mismatch";
            for (int k = gptr[c]; k < gptr[c + 1]; ++k)
                if (!staged(T.st_row[k], gidx[k])) return "tile stencil row mismatch";
            for (int j = 0; j < p.face_size[c]; ++j) {
                size_t s = (size_t)c * 6 + j;
                int kd = fkind[s];
                if (kd == 0) continue;
                if (!staged(T.f_orow[s], kd == 2 ? -fother[s] - 1 : fother[s])) return "tile face row mismatch";
                if (!vrow(T.f_v0row[s], fv0[s]) || !vrow(T.f_v1row[s], fv1[s])) return "tile face vertex mismatch";
            }
        }
    }
    return "";
}

std::string pack_tile_mfma(const Problem &p, const std::vector<int> &gptr, const std::vector<double> &gP,
                           const std::vector<int> &gok, const std::vector<int> &fkind,
                           const std::vector<const std::vector<double> *> &face8, double invL, Tiles &T) {
    int nc = p.n_cells;
    size_t nnz = gptr[nc];
    T.cmf.assign((size_t)T.n_tiles * kTileKs * 64, 0.0);
    T.fin_i.assign((size_t)nc * kFinI, 0);
    T.fin_d.assign((size_t)nc * kFinD, 0.0);
    const std::vector<double> &ga = *face8[0], &n0x = *face8[1], &n0y = *face8[2], &mxm = *face8[3],
                              &mym = *face8[4], &fnx = *face8[5], &fny = *face8[6], &area = *face8[7];
    for (int t = 0; t < T.n_tiles; ++t) {
        int c0 = T.cell0[t], ncell = T.cell0[t + 1] - c0, nh = T.nh[t];
        if (ncell > kTileCellsMax || nh > kTileRowsMax) return "tile exceeds the matrix-core shape";
        // C[m][r], m = cell + 8 * component
        std::vector<double> C((size_t)16 * kTileRowsMax, 0.0);
        for (int i = 0; i < ncell; ++i) {
            int c = c0 + i;
            if (!gok[c]) continue;
            int own = T.crow[c];
            for (int q = 0; q < 2; ++q) {
                double *Cm = &C[(size_t)(i + 8 * q) * kTileRowsMax];
                double sum = 0;
                for (int k = gptr[c]; k < gptr[c + 1]; ++k) {
                    double wgt = invL * gP[q * nnz + k];
                    Cm[T.st_row[k]] += wgt;
                    sum += wgt;
                }
                Cm[own] -= sum;
            }
        }
        double *F = &T.cmf[(size_t)t * kTileKs * 64];
        for (int k = 0; k < kTileKs; ++k)
            for (int l = 0; l < 64; ++l) {
                int r = 4 * k + l / 16;
                F[k * 64 + l] = r < nh ? C[(size_t)(l % 16) * kTileRowsMax + r] : 0.0;
            }
        // moment rows are the staged rows; a face vertex's moments are the
        // inverse-distance combination of its sources' (kYRows + its index)
        int nr = T.rptr[t + 1] - T.rptr[t];
        for (int i = 0; i < ncell; ++i) {
            int c = c0 + i;
            int nf = p.face_size[c];
            if (nf > 4) return "more than 4 faces per cell in a tile";
            int *fi = &T.fin_i[(size_t)c * kFinI];
            double *fd = &T.fin_d[(size_t)c * kFinD];
            int yo = T.crow[c];
            fi[0] = yo;
            fi[1] = nf;
            for (int j = 0; j < nf; ++j) {
                size_t s = (size_t)c * 6 + j;
                int kd = fkind[s];
                fi[2 + 3 * j] = fi[3 + 3 * j] = fi[4 + 3 * j] = yo;
                if (kd == 0) continue;
                int v0 = T.f_v0row[s] - nr, v1 = T.f_v1row[s] - nr;
                if (T.f_orow[s] < 0 || T.f_orow[s] >= nr || v0 < 0 || v1 < 0 || v0 >= kYRows || v1 >= kYRows)
                    return "tile moment row mismatch";
                fi[2 + 3 * j] = T.f_orow[s];
                fi[3 + 3 * j] = kYRows + v0;
                fi[4 + 3 * j] = kYRows + v1;
                // face term area (kl n.g - tau (v.g)(v.n)), g = de ga n0 + dv m,
                // de = sg (other - own), dv = v1 - v0, v = vg s; its
                // direction part expands over sx^2, sx sy, sy^2
                double sg = kd == 3 ? -1.0 : 1.0, A = area[s], nx = fnx[s], ny = fny[s];
                double ax = n0x[s], ay = n0y[s], bx = mxm[s], by = mym[s], Ag = sg * A * ga[s];
                fd[8 * j + 0] = Ag * (nx * ax + ny * ay);
                fd[8 * j + 1] = -Ag * ax * nx;
                fd[8 * j + 2] = -Ag * (ax * ny + ay * nx);
                fd[8 * j + 3] = -Ag * ay * ny;
                fd[8 * j + 4] = A * (nx * bx + ny * by);
                fd[8 * j + 5] = -A * bx * nx;
                fd[8 * j + 6] = -A * (bx * ny + by * nx);
                fd[8 * j + 7] = -A * by * ny;
            }
        }
        if (nr > kYRows || T.vptr[t + 1] - T.vptr[t] > kYRows) return "tile has more moment rows than the matrix-core shape";
    }
    return "";
}

}  // namespace gbte
