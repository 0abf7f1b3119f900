// C ABI: the reference's output files (declared in include/giftbte.h).
//
// After a steady BTE run the reference leaves, in its working directory:
//   TempLattice.dat  x y z T (the lattice / macro temperature)
//   Tempband.dat     band x y z T_band
//                    (StaticBTESynthetic::_print_out, StaticBTESynthetic.cpp:2613-2653)
//   Temperature.dat  x y z T_energy
//   HeatFlux.dat     x y z qx qy qz in width-15 columns
//   Result.vtk       legacy VTK unstructured grid with the energy temperature
//                    and the heat flux per cell
//                    (SolutionAll::_print_out, Solution.cpp:128-233, after the
//                    hand-back of Solution.cpp:408-413 and BTEMain.cpp:393)
// in default ostream formatting (6 significant digits).  The files are
// written with the same formatting so that they read the same.
#include <fstream>
#include <iomanip>
#include <string>

#include "giftbte.h"
#include "problem.hpp"

namespace {

std::string in_dir(const char *dir, const char *name) {
    std::string d = dir ? dir : ".";
    if (d.empty()) d = ".";
    return d + "/" + name;
}

}  // namespace

extern "C" {

int gbte_write_fields(const gbte_problem *pp, const double *T_lattice, const double *T_energy, const double *qx,
                      const double *qy, const double *qz, const double *T_band, const char *dir) {
    if (!pp || !T_lattice || !T_energy || !qx || !qy || !qz || !T_band) return GBTE_BAD_INPUT;
    const gbte::Problem &P = pp->p;
    int nc = P.n_cells, nB = P.n_bands;
    std::ofstream a(in_dir(dir, "TempLattice.dat"));
    std::ofstream b(in_dir(dir, "Tempband.dat"));
    std::ofstream c(in_dir(dir, "Temperature.dat"));
    std::ofstream h(in_dir(dir, "HeatFlux.dat"));
    std::ofstream v(in_dir(dir, "Result.vtk"));
    if (!a || !b || !c || !h || !v) return GBTE_BAD_INPUT;
    for (int i = 0; i < nc; ++i) a << P.cx[i] << " " << P.cy[i] << " " << P.cz[i] << " " << T_lattice[i] << std::endl;
    for (int j = 0; j < nB; ++j)
        for (int i = 0; i < nc; ++i)
            b << j + 1 << " " << P.cx[i] << " " << P.cy[i] << " " << P.cz[i] << " " << T_band[(size_t)j * nc + i]
              << std::endl;
    for (int i = 0; i < nc; ++i) c << P.cx[i] << " " << P.cy[i] << " " << P.cz[i] << " " << T_energy[i] << std::endl;
    for (int i = 0; i < nc; ++i)
        h << std::setw(15) << P.cx[i] << std::setw(15) << P.cy[i] << std::setw(15) << P.cz[i] << std::setw(15)
          << qx[i] << std::setw(15) << qy[i] << std::setw(15) << qz[i] << std::endl;
    v << "# vtk DataFile Version 2.0" << std::endl << "RESULT" << std::endl << "ASCII" << std::endl
      << "DATASET UNSTRUCTURED_GRID" << std::endl;
    v << "POINTS " << P.n_nodes << " float" << std::endl;
    for (int k = 0; k < P.n_nodes; ++k)
        v << P.node_x[k] / P.deck.Lx << " " << P.node_y[k] / P.deck.Ly << " " << P.node_z[k] / P.deck.Lz << std::endl;
    long total = 0;
    for (int i = 0; i < nc; ++i) total += P.cell_nvert[i] + 1;
    v << std::endl << "CELLS " << nc << " " << total << std::endl;
    for (int i = 0; i < nc; ++i) {
        v << P.cell_nvert[i];
        for (int j = 0; j < P.cell_nvert[i]; ++j) v << " " << P.cell_vert[(size_t)i * 8 + j];
        v << std::endl;
    }
    v << std::endl << "CELL_TYPES " << nc << std::endl;
    for (int i = 0; i < nc; ++i) {
        // the reference's codes: 2D quads 8 (VTK_PIXEL), triangles 5; 3D
        // tetrahedra 10, hexahedra 11, wedges 13
        int nv = P.cell_nvert[i], code = -1;
        if (P.dim == 2 && nv == 4) code = 8;
        else if (P.dim == 2 && nv == 3) code = 5;
        else if (P.dim == 3 && nv == 4) code = 10;
        else if (P.dim == 3 && nv == 8) code = 11;
        else if (P.dim == 3 && nv == 6) code = 13;
        if (code >= 0) v << code << std::endl;
    }
    v << std::endl << "CELL_DATA" << " " << nc << std::endl << "SCALARS" << " Temperature" << " float 1" << std::endl
      << "LOOKUP_TABLE default" << std::endl;
    for (int i = 0; i < nc; ++i) v << T_energy[i] << std::endl;
    v << "VECTORS" << " HeatFLux" << " float" << std::endl;
    for (int i = 0; i < nc; ++i) v << qx[i] << " " << qy[i] << " " << qz[i] << std::endl;
    return (a && b && c && h && v) ? GBTE_OK : GBTE_BAD_INPUT;
}

}  // extern "C"
