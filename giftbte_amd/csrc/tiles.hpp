// Cell tiles for the LDS-staged face / vertex / moment kernel (tiles.cpp).
#pragma once

#include <string>
#include <vector>

#include "problem.hpp"

namespace gbte {

struct Tiles {
    int n_tiles = 0, max_rows = 0;
    std::vector<int> cell0;            // tile t owns cells [cell0[t], cell0[t+1])
    std::vector<int> rptr, rsrc;       // staged rows of tile t: rsrc >= 0 cell row of e, < 0 -(entry+1) row of eb
    std::vector<int> nh;               // per tile: its H rows (cells) come first among the staged rows
    std::vector<int> vptr, vnode;      // vertex rows of tile t (placed after its staged rows)
    std::vector<int> vs_ptr, vs_row;   // per vertex row: sources (tile-local rows) ...
    std::vector<double> vs_w;          // ... and inverse-distance weights
    std::vector<int> crow;             // per cell: its tile-local row
    std::vector<int> st_row;           // per LSQ stencil entry: tile-local row
    std::vector<int> f_orow, f_v0row, f_v1row;  // per (cell, face) slot: tile-local rows
    // matrix-core tables (pack_tile_mfma)
    std::vector<double> cmf;           // per tile: LSQ coefficients as A fragments [kTileKs][64]
    std::vector<int> fin_i;            // per cell [kFinI]: own Y index, face count, per face Y(other, v0, v1)
    std::vector<double> fin_d;         // per cell [kFinD]: per face the factors of Y_0..3 (face value, vertex)
};

// The LSQ gradient of a tile's 8 cells is one product per slice, G = C E:
// C[m][r], m = cell + 8 * component, r < nh a tile row, holds the stencil
// weights / Lx and minus their sum on the cell's own row.  cmf stores C in
// the A-operand order of v_mfma_f64_16x16x4f64: k-step k, lane l holds
// C[l % 16][4 k + l / 16].
//
// The moments are Y[j][r] = sum over the row entries t of K_j(t) E[r][t] for
// the staged rows r, and for a face vertex the inverse-distance combination
// of its sources' Y (Y is linear in the rows; Y index kYRows + the vertex's
// tile index).  T, q_x, q_y, q_z are Y_4..7 of the cell's own row and the
// non-Fourier residual is sum_faces sum_{j<4} cde_j (Y_j[other] - Y_j[own]) +
// cdv_j (Y_j[v1] - Y_j[v0]), with K_0..3 the per-(direction, band) factors
// kl, tau vg^2 (sx^2, sx sy, sy^2) of its face terms.
constexpr int kTileCellsMax = 8, kTileRowsMax = 64, kTileKs = kTileRowsMax / 4, kYRows = 64;
constexpr int kFinI = 16, kFinD = 32;

// gptr/gidx: LSQ stencil; vptr/vsrc/vwt: inverse-distance vertex stencil
// (vsrc >= 0 cell, < 0 -(entry+1)); fkind/fother/fv0/fv1 per (cell, face)
// slot: face-gradient kind (0 none, 1/3 interior, 2 boundary), the other
// cell or boundary entry, and the face's two vertices.
std::string build_tiles(const Problem &p, const std::vector<int> &gptr, const std::vector<int> &gidx,
                        const std::vector<int> &vptr, const std::vector<int> &vsrc, const std::vector<double> &vwt,
                        const std::vector<int> &fkind, const std::vector<int> &fother, const std::vector<int> &fv0,
                        const std::vector<int> &fv1, int max_rows, int max_cells, Tiles &T);

// Fill the matrix-core tables from the tile tables, the LSQ weights gP (x
// then y, per stencil entry), the gradient flags and the per-slot face data
// face8 = (ga, n0x, n0y, mxm, mym, nx, ny, area); invL scales the gradient.
std::string pack_tile_mfma(const Problem &p, const std::vector<int> &gptr, const std::vector<double> &gP,
                           const std::vector<int> &gok, const std::vector<int> &fkind,
                           const std::vector<const std::vector<double> *> &face8, double invL, Tiles &T);

}  // namespace gbte
