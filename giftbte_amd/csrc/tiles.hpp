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
    std::vector<int> vptr, vnode;      // vertex rows of tile t (placed after its staged rows)
    std::vector<int> vs_ptr, vs_row;   // per vertex row: sources (tile-local rows) ...
    std::vector<double> vs_w;          // ... and inverse-distance weights
    std::vector<int> crow;             // per cell: its tile-local row
    std::vector<int> st_row;           // per LSQ stencil entry: tile-local row
    std::vector<int> f_orow, f_v0row, f_v1row;  // per (cell, face) slot: tile-local rows
    // lane-indexed records, one 64-lane vector per cell (pack_tile_records)
    std::vector<int> crec_i;
    std::vector<double> crec_a, crec_b;
};

// Lane layout of a cell record (kRec = 64 lanes):
//   ints   [0, kRecSt) stencil rows, [24,28) face kind, [28,32) other row,
//          [32,36) vertex-0 row, [36,40) vertex-1 row, 40 own row, 41 gradient
//          flag, 42 stencil count, 43 face count
//   dbl a  [0, kRecSt) stencil weight x, [24..56) per face j: ga, n0x, n0y,
//          m_x/|m|^2, m_y/|m|^2, nx, ny, area at 24 + 4*q + j
//   dbl b  [0, kRecSt) stencil weight y
constexpr int kRec = 64, kRecSt = 24;

// gptr/gidx: LSQ stencil; vptr/vsrc/vwt: inverse-distance vertex stencil
// (vsrc >= 0 cell, < 0 -(entry+1)); fkind/fother/fv0/fv1 per (cell, face)
// slot: face-gradient kind (0 none, 1/3 interior, 2 boundary), the other
// cell or boundary entry, and the face's two vertices.
std::string build_tiles(const Problem &p, const std::vector<int> &gptr, const std::vector<int> &gidx,
                        const std::vector<int> &vptr, const std::vector<int> &vsrc, const std::vector<double> &vwt,
                        const std::vector<int> &fkind, const std::vector<int> &fother, const std::vector<int> &fv0,
                        const std::vector<int> &fv1, int max_rows, int max_cells, Tiles &T);

// Fill the lane-indexed records from the tile tables and the per-slot face
// data (ga, n0x, n0y, mxm, mym, nx, ny, area) and LSQ weights gP (x then y).
std::string pack_tile_records(const Problem &p, const std::vector<int> &gptr, const std::vector<double> &gP,
                              const std::vector<int> &gok, const std::vector<int> &fkind,
                              const std::vector<const std::vector<double> *> &face8, Tiles &T);

}  // namespace gbte
