// Host-side problem: deck parsing, mesh ingestion, geometry and the
// flattened tables the solver consumes.  Mirrors the reference's set-up
// chain src/BTEMain.cpp:87-362 -> BTEBand / BTEBoundaryCondition /
// DistributeMesh(ReadIn, BTEMesh::setMeshParams, setMeshParams1) / BTEAngle
// -> StaticBTESynthetic constructor tables.
#pragma once

#include <string>
#include <vector>

#include "giftbte.h"

namespace gbte {

struct Deck {
    // CONTROL (src/BTEMain.cpp:63-78 defaults)
    std::string state = "Steady", order = "2", method = "Sequential", matrix = "LU";
    std::string multiscale;
    int max_iter = 10000;
    double tol_temp = 1e-6, tol_flux = 1e-4, tref = 300.0;
    // GEOMETRY
    int dim = -1;
    double Lx = 1, Ly = 1, Lz = 1, uniform_heat = 0;
    std::string bc_file, mesh_type, mesh_file, heat_type, heat_file;
    // PHONON
    int material_dim = 3, angle_method = 2, ntheta = 4, nphi = 4;
};

struct BoundaryCond {
    int index = 0, type = 0;
    double temperature = 0;
    std::string interface_file;
    std::vector<std::pair<int, std::vector<double>>> matter_trans;
};

struct Problem {
    Deck deck;
    int dim = 2, n_cells = 0, n_nodes = 0, n_bounds = 0, n_faces = 0, n_bands = 0, n_dirs = 0, n_matters = 0;
    // cells (6 face slots, 4 vertices per slot, 8 vertices per cell)
    std::vector<int> face_size, face_neighbor, face_bound, face_index, face_nvert, face_vert, cell_nvert, cell_vert,
        matter, nbr_ptr, nbr_idx, cell_region;
    std::vector<double> face_area, face_nx, face_ny, face_nz, face_cx, face_cy, face_cz, volume, cx, cy, cz,
        heat_source, initial_temp;
    std::vector<double> node_x, node_y, node_z;
    // boundaries
    std::vector<int> bound_cell, bound_face, bound_type, bound_connect;
    std::vector<double> bound_temp;
    std::vector<double> bound_trans;  // [n_bounds][n_bands][2]
    // unique faces
    std::vector<int> fface_cell, fface_face, fface_bound;
    // bands / directions
    std::vector<double> kappa_bulk, capacity_bulk, vg, tau, cap, lr, hr;
    std::vector<double> dir_x, dir_y, dir_z, weight;

    gbte_tables tables() const;
};

// Returns "" on success, else an error message.
std::string load_problem(const std::string &deck_dir, Problem &p);

// Gauss-Legendre nodes/weights on [a, b], N points, interleaved
// (x0, w0, x1, w1, ...), the layout src/utility/utility.cpp:116-172 returns.
std::vector<double> gauss_legendre(double a, double b, int n);

}  // namespace gbte

// Opaque C handle of include/giftbte.h.
struct gbte_problem {
    gbte::Problem p;
};
