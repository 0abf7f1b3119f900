// Dump driver for the compiled reference (test infrastructure only).
//
// Links against the reference GiftBTE objects built by oracle/ref.mk and
// replays the steady flow of src/BTEMain.cpp:15-421 at np=1 for the
// Order=1, 2 or 2limit / IterativeScheme=Synthetic or Sequential / MatrixSolver=LU or BICGSTAB branches
// (src/Solution/Solution.cpp:262-414), then writes every field the GPU
// solver is compared against as .npy files at full FP64 precision.  The
// reference's own .dat outputs carry 6 significant digits only
// (src/StaticBTESynthetic/StaticBTESynthetic.cpp:2613-2653).
//
// usage: ref_dump <deck_dir> <out_dir> <max_iter> <tolT> <tolQ> [stage_band stage_dir] [--tables DIR [--time]]
//   deck_dir  a COPY of a deck (the reference writes its .dat files into the
//             working directory on every iteration)
//   out_dir   where the .npy files go
//   max_iter  outer-iteration cap handed to StaticBTESynthetic::solve
//   tolT/tolQ stop thresholds (0 0 runs exactly max_iter iterations)
//   stage_*   optional: also dump every per-system stage of iteration 0 for
//             system (band, direction)
//   --tables  build the mesh from flattened tables in DIR instead of the
//             deck's mesh file (3D decks, whose reference reader was removed)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <string>
#include <vector>
#include <fstream>
#include <sstream>
#include <iostream>
#include <unistd.h>
#include <sys/stat.h>

#include "StaticBTESynthetic/StaticBTESynthetic.h"
#include "StaticFourier/StaticFourier.h"
#include "BTEMesh/Distribute.h"
#include "Solution/Solution.h"

namespace {

std::string g_out;

void write_npy(const std::string &name, const char *descr, size_t elem,
               const void *data, const std::vector<size_t> &shape) {
    std::string sh = "(";
    size_t n = 1;
    for (size_t k = 0; k < shape.size(); ++k) {
        sh += std::to_string(shape[k]);
        sh += (shape.size() == 1) ? "," : (k + 1 < shape.size() ? ", " : "");
        n *= shape[k];
    }
    sh += ")";
    std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
    size_t total = 10 + hdr.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    hdr.append(pad, ' ');
    hdr.push_back('\n');
    std::string path = g_out + "/" + name + ".npy";
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) { perror(path.c_str()); exit(2); }
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    fwrite(magic, 1, 8, f);
    uint16_t hl = (uint16_t)hdr.size();
    fwrite(&hl, 2, 1, f);
    fwrite(hdr.data(), 1, hdr.size(), f);
    if (n) fwrite(data, elem, n, f);
    fclose(f);
}
void dump_d(const std::string &name, const double *p, std::vector<size_t> shape) { write_npy(name, "<f8", 8, p, shape); }
void dump_i(const std::string &name, const int *p, std::vector<size_t> shape) { write_npy(name, "<i4", 4, p, shape); }
void dump_d(const std::string &name, const std::vector<double> &v) { dump_d(name, v.data(), {v.size()}); }
void dump_i(const std::string &name, const std::vector<int> &v) { dump_i(name, v.data(), {v.size()}); }

// Keyword scan with the reference's semantics (src/BTEMain.cpp:97-193):
// a line that contains the keyword and no '!' makes the next stream token the
// value.
struct Deck {
    std::string order = "2", method = "Sequential", matrix = "LU", state = "Steady";
    double tref = 300, errT = 1e-6, errQ = 1e-4;
    int maxIter = 10000;
    int dimGeo = -1;
    double Lx = 1, Ly = 1, Lz = 1, uniformHeat = 0;
    std::string bcFile, meshType, meshFile, heatType, heatFile, multiscale;
    int dimMat = 3, angleMethod = 2, ntheta = 4, nphi = 4;
};

bool has_kw(const std::string &s, const char *kw) {
    return s.find(kw) != std::string::npos && s.find('!') == std::string::npos;
}

Deck parse_deck() {
    Deck d;
    {
        std::ifstream f("input/CONTROL");
        std::string s;
        while (std::getline(f, s)) {
            if (has_kw(s, "State")) f >> d.state;
            if (has_kw(s, "Order")) f >> d.order;
            if (has_kw(s, "IterativeScheme")) f >> d.method;
            if (has_kw(s, "MatrixSolver")) f >> d.matrix;
            if (has_kw(s, "Tref")) f >> d.tref;
            if (has_kw(s, "MaxNumIter")) f >> d.maxIter;
            if (has_kw(s, "ResidualTemp")) f >> d.errT;
            if (has_kw(s, "ResidualFlux")) f >> d.errQ;
            if (has_kw(s, "Multiscale")) f >> d.multiscale;
        }
    }
    {
        std::ifstream f("input/GEOMETRY");
        std::string s;
        while (std::getline(f, s)) {
            if (has_kw(s, "GeometryDimension")) f >> d.dimGeo;
            if (has_kw(s, "ScaleX")) f >> d.Lx;
            if (has_kw(s, "ScaleY")) f >> d.Ly;
            if (has_kw(s, "ScaleZ")) f >> d.Lz;
            if (has_kw(s, "UniformHeat")) f >> d.uniformHeat;
            if (has_kw(s, "BCFile")) f >> d.bcFile;
            if (has_kw(s, "MeshType")) f >> d.meshType;
            if (has_kw(s, "MeshFile")) f >> d.meshFile;
            if (has_kw(s, "HeatType")) f >> d.heatType;
            if (has_kw(s, "HeatFile")) f >> d.heatFile;
        }
    }
    {
        std::ifstream f("input/PHONON");
        std::string s;
        while (std::getline(f, s)) {
            if (has_kw(s, "MaterialDimension")) f >> d.dimMat;
            if (has_kw(s, "Angle_method")) f >> d.angleMethod;
            if (has_kw(s, "Ntheta")) f >> d.ntheta;
            if (has_kw(s, "Nphi")) f >> d.nphi;
        }
    }
    return d;
}

// Flattened solver tables built by the reference constructor
// (src/StaticBTESynthetic/StaticBTESynthetic.cpp:224-626).
void dump_tables(StaticBTESynthetic &s, const std::string &pfx) {
    size_t nc = s.numCell, nb = s.numBound, nf = s.numFace, nn = s.numNode;
    dump_i(pfx + "elementFaceSize", s.elementFaceSize, {nc});
    dump_i(pfx + "elementFaceNeighbor", s.elementFaceNeighobr, {nc, 6});
    dump_i(pfx + "elementFaceBound", s.elementFaceBound, {nc, 6});
    dump_i(pfx + "elementFaceIndex", s.elementFaceIndex, {nc, 6});
    dump_d(pfx + "elementFaceArea", s.elementFaceArea, {nc, 6});
    dump_d(pfx + "elementFaceNormX", s.elementFaceNormX, {nc, 6});
    dump_d(pfx + "elementFaceNormY", s.elementFaceNormY, {nc, 6});
    dump_d(pfx + "elementFaceNormZ", s.elementFaceNormZ, {nc, 6});
    dump_d(pfx + "elementFaceCenterX", s.elementFaceCenterX, {nc, 6});
    dump_d(pfx + "elementFaceCenterY", s.elementFaceCenterY, {nc, 6});
    dump_d(pfx + "elementFaceCenterZ", s.elementFaceCenterZ, {nc, 6});
    dump_i(pfx + "elementVertexesSize", s.elementVertexesSize, {nc});
    dump_d(pfx + "elementVolume", s.elementVolume, {nc});
    dump_d(pfx + "elementCenterX", s.elementCenterX, {nc});
    dump_d(pfx + "elementCenterY", s.elementCenterY, {nc});
    dump_d(pfx + "elementCenterZ", s.elementCenterZ, {nc});
    dump_d(pfx + "elementHeatSource", s.elementHeatSource, {nc});
    dump_i(pfx + "matter", s.matter, {nc});
    std::vector<int> fv(nc * 6 * 4, -1), fvs(nc * 6, 0), cv(nc * 8, -1);
    for (size_t i = 0; i < nc; ++i) {
        for (int j = 0; j < s.elementVertexesSize[i]; ++j) cv[i * 8 + j] = s.elementVertexes[i * 8 + j];
        for (int j = 0; j < s.elementFaceSize[i]; ++j) {
            fvs[i * 6 + j] = s.elementFaceVertexesSize[i * 6 + j];
            for (int k = 0; k < fvs[i * 6 + j]; ++k) fv[(i * 6 + j) * 4 + k] = s.elementFaceVertexes[(i * 6 + j) * 4 + k];
        }
    }
    dump_i(pfx + "elementVertexes", cv.data(), {nc, 8});
    dump_i(pfx + "elementFaceVertexesSize", fvs.data(), {nc, 6});
    dump_i(pfx + "elementFaceVertexes", fv.data(), {nc, 6, 4});
    // vertex-neighbour list (same matter), CSR
    std::vector<int> nptr(nc + 1, 0), nidx;
    for (size_t i = 0; i < nc; ++i) {
        for (int c : s.elementNeighborList[i]) nidx.push_back(c);
        nptr[i + 1] = (int)nidx.size();
    }
    dump_i(pfx + "neighborPtr", nptr);
    dump_i(pfx + "neighborIdx", nidx);
    dump_d(pfx + "nodeX", s.nodeX, {nn});
    dump_d(pfx + "nodeY", s.nodeY, {nn});
    dump_d(pfx + "nodeZ", s.nodeZ, {nn});
    std::vector<int> bc(nb * 2), bf(nb * 2);
    for (size_t i = 0; i < nb; ++i)
        for (int k = 0; k < 2; ++k) { bc[i * 2 + k] = s.boundaryCell[i][k]; bf[i * 2 + k] = s.boundaryFace[i][k]; }
    dump_i(pfx + "boundaryCell", bc.data(), {nb, 2});
    dump_i(pfx + "boundaryFace", bf.data(), {nb, 2});
    dump_i(pfx + "boundaryType", s.boundaryType, {nb});
    dump_d(pfx + "boundaryThermal", s.boundaryThermal, {nb});
    dump_i(pfx + "boundaryConnect", s.boundaryConnect, {nb});
    std::vector<int> fc(nf * 2), ff(nf * 2);
    for (size_t i = 0; i < nf; ++i)
        for (int k = 0; k < 2; ++k) { fc[i * 2 + k] = s.faceCell[i][k]; ff[i * 2 + k] = s.faceFace[i][k]; }
    dump_i(pfx + "faceCell", fc.data(), {nf, 2});
    dump_i(pfx + "faceFace", ff.data(), {nf, 2});
    dump_i(pfx + "faceBound", s.faceBound, {nf});
    size_t nm = s.numofMatter, nB = s.numBand, nD = s.numDirection;
    dump_d(pfx + "kappaBulk", s.kappaBulk, {nm});
    dump_d(pfx + "capacityBulk", s.capacityBulk, {nm});
    dump_d(pfx + "directionX", s.directionX, {nD});
    dump_d(pfx + "directionY", s.directionY, {nD});
    dump_d(pfx + "directionZ", s.directionZ, {nD});
    std::vector<double> w(nD), vg(nm * nB), tau(nm * nB), cap(nm * nB), lr(nm * nB), hr(nm * nB);
    for (size_t d = 0; d < nD; ++d) w[d] = s.modeWeight[0][0][d];
    for (size_t m = 0; m < nm; ++m)
        for (size_t b = 0; b < nB; ++b) {
            // |v| recovered from the stored v*s products of the first direction
            double vx = s.groupVelocityX[m][b][0], vy = s.groupVelocityY[m][b][0], vz = s.groupVelocityZ[m][b][0];
            double sx = s.directionX[0], sy = s.directionY[0], sz = s.directionZ[0];
            double ss = sx * sx + sy * sy + sz * sz;
            vg[m * nB + b] = (vx * sx + vy * sy + vz * sz) / ss;
            tau[m * nB + b] = s.relaxationTime[m][b][0];
            cap[m * nB + b] = s.heatCapacity[m][b][0];
            lr[m * nB + b] = s.latticeRatio[m][b][0];
            hr[m * nB + b] = s.heatRatio[m][b][0];
        }
    dump_d(pfx + "modeWeight", w);
    dump_d(pfx + "groupVelocity", vg.data(), {nm, nB});
    dump_d(pfx + "relaxationTime", tau.data(), {nm, nB});
    dump_d(pfx + "heatCapacity", cap.data(), {nm, nB});
    dump_d(pfx + "latticeRatio", lr.data(), {nm, nB});
    dump_d(pfx + "heatRatio", hr.data(), {nm, nB});
    int dims[8] = {s.numCell, s.numBound, s.numFace, s.numNode, s.numBand, s.numDirection, s.numofMatter, s.dimension};
    dump_i(pfx + "sizes", dims, {8});
    double scale[3] = {s.L_x, s.L_y, s.L_z};
    dump_d(pfx + "scale", scale, {3});
}

void dump_cellmatrix(StaticBTESynthetic &s, const std::string &name) {
    int dim = s.dimension;
    std::vector<double> cm;
    for (int i = 0; i < s.numCell; ++i) {
        size_t n = s.elementNeighborList[i].size();
        bool set = (dim == 2) ? n > 2 : n > 3;
        for (int r = 0; r < dim; ++r)
            for (size_t m = 0; m < n; ++m) cm.push_back(set ? s.CellMatrix[i][r][m] : 0.0);
    }
    dump_d(name, cm);
}

void init_solver_state(StaticBTESynthetic &solver, DistributeMesh &dm, SolutionAll &sa) {
    // Solution.cpp:268-285 (hand-over from the Fourier warm start)
    for (size_t i = 0; i < dm.volumeElementsBTE.size(); ++i) {
        solver.temperature[i] = sa.solution.Temperature[dm.elementProjection[i]];
        solver.heatFluxXGlobal[i] = sa.solution.heatFluxX[dm.elementProjection[i]];
        solver.heatFluxYGlobal[i] = sa.solution.heatFluxY[dm.elementProjection[i]];
        solver.heatFluxZGlobal[i] = sa.solution.heatFluxZ[dm.elementProjection[i]];
    }
    for (size_t ib = 0; ib < dm.BTEMeshes.Boundaries.size(); ++ib) {
        int ie = dm.BTEMeshes.Boundaries[ib].cellindex[0];
        int jface = dm.BTEMeshes.Boundaries[ib].faceindex[0];
        Cell &fc = dm.FourierMeshes.Elements[dm.elementProjection[ie]];
        if (fc.faces[jface].bound == -1) {
            double d1 = get_distance(fc.center, fc.faces[jface].center);
            double d2 = get_distance(dm.FourierMeshes.Elements[fc.faces[jface].neighbor].center, fc.faces[jface].center);
            double t = sa.solution.Temperature[dm.elementProjection[ie]] / d1 +
                       sa.solution.Temperature[fc.faces[jface].neighbor] / d2;
            solver.boundaryThermal[ib] = t / (1.0 / d1 + 1.0 / d2);
        }
    }
}

void stage_dump(StaticBTESynthetic &s, int sb, int sd) {
    // Setup part of StaticBTESyntheticSolver.cpp:15-62, then iteration 0 of
    // the per-system loop (:78-127) for system (sb, sd), one stage at a time.
    s._set_face_matrix();
    s._set_cell_matrix_larger();
    s._set_matrix("LU");
    s.errorIncreaseTime = 0;
    s._set_initial(0);
    for (int d = 0; d < s.numDirectionLocal; d++)
        for (int b = 0; b < s.numBandLocal; ++b) s._get_coefficient(b, d);
    s._get_coefficient_macro();
    s._get_face_matrix();
    s._get_CellMatrix_larger();
    dump_cellmatrix(s, "cellMatrix");
    size_t nc = s.numCell, nb = s.numBound, nB = s.numBand, nD = s.numDirection;
    for (int d = 0; d < s.numDirectionLocal; d++)
        for (int b = 0; b < s.numBandLocal; ++b) {
            s._get_gradient_larger(0, b, d);
            s._get_bound_ee(b, d);
            MPI_Allgather(s.eboundLocal + b * nb * 2, nb * 2, MPI_DOUBLE,
                          s.ebound + nb * 2 * d + nD * nb * 2 * b, nb * 2, MPI_DOUBLE, MPI_COMM_WORLD);
        }
    s._set_bound_ee_1();
    dump_d("stage_ebound_init", s.ebound, {nB, nD, nb, 2});
    s.copy();
    dump_d("stage_e_old", s.energyDensity[sb][sd], {nc});
    dump_d("stage_T_old", s.temperatureOld, {nc});
    s._get_gradient_larger(0, sb, sd);
    dump_d("stage_gradX", s.gradientX, {nc});
    dump_d("stage_gradY", s.gradientY, {nc});
    dump_d("stage_gradZ", s.gradientZ, {nc});
    s._get_Re(sb, sd);
    dump_d("stage_Re", s.Re, {nc});
    s.sol1 = s.cg[sb][sd].solve(s.Re1);
    for (size_t i = 0; i < nc; ++i) s.energyDensity[sb][sd][i] += s.sol1(i);
    dump_d("stage_e_new", s.energyDensity[sb][sd], {nc});
    s._get_bound_ee(sb, sd);
    dump_d("stage_ebound_sys", s.ebound + (sb * nD + sd) * nb * 2, {nb, 2});
    s._set_vertex_energydensity(sb, sd);
    dump_d("stage_e_vertex", s.energyDensityVertex, {(size_t)s.numNode * s.numofMatter});
    s._get_face_gradient(sb, sd);
    size_t nf = s.numFace;
    dump_d("stage_gradFaceX", s.gradientFaceX, {nf, 2});
    dump_d("stage_gradFaceY", s.gradientFaceY, {nf, 2});
    dump_d("stage_gradFaceZ", s.gradientFaceZ, {nf, 2});
    s._get_non_fourier(sb, sd);
    dump_d("stage_ReMacro", s.ReMacroLocal, {nc});
    s._recover_temperature(sb, sd);
    s._get_total_energy(sb, sd);
    s._get_heat_flux(sb, sd);
    dump_d("stage_T_sys", s.temperatureLocal, {nc});
    dump_d("stage_E_sys", s.totalEnergyLocal, {nc});
    dump_d("stage_qx_sys", s.heatFluxXLocal, {nc});
    dump_d("stage_qy_sys", s.heatFluxYLocal, {nc});
    dump_d("stage_qz_sys", s.heatFluxZLocal, {nc});
    int sys[2] = {sb, sd};
    dump_i("stage_system", sys, {2});
}

// ---- 3D: the mesh from flattened tables.  The reference's 3D readers and
// 3D geometry builder were removed (src/BTEMesh/ReadinMesh.cpp:404-406,
// 531-533, src/BTEMesh/BTEMesh.cpp:841-843), but its solvers keep their 3D
// branches; this fills the public members of a BTEMesh
// (includes/BTEMesh/BTEMesh.h:21-37) from raw little-endian arrays written
// by the test suite (one <name>.bin per table of include/giftbte.h) and runs
// the reference's StaticFourier and StaticBTESynthetic on it.
template <typename T>
std::vector<T> read_bin(const std::string &dir, const std::string &name, size_t n) {
    std::vector<T> v(n);
    std::string path = dir + "/" + name + ".bin";
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) { perror(path.c_str()); exit(2); }
    if (n && fread(v.data(), sizeof(T), n, f) != n) { fprintf(stderr, "short read: %s\n", path.c_str()); exit(2); }
    fclose(f);
    return v;
}

void mesh_from_tables(const std::string &dir, BTEMesh &m) {
    std::vector<int> sz = read_bin<int>(dir, "sizes", 5);
    std::vector<double> sc = read_bin<double>(dir, "scale", 3);
    size_t nc = sz[0], nn = sz[1], nb = sz[2], nf = sz[3];
    m.Dimension = sz[4];
    m.L_x = sc[0]; m.L_y = sc[1]; m.L_z = sc[2];
    auto fsize = read_bin<int>(dir, "face_size", nc), fnb = read_bin<int>(dir, "face_neighbor", nc * 6),
         fbd = read_bin<int>(dir, "face_bound", nc * 6), fidx = read_bin<int>(dir, "face_index", nc * 6),
         fnv = read_bin<int>(dir, "face_nvert", nc * 6), fv = read_bin<int>(dir, "face_vert", nc * 24),
         cnv = read_bin<int>(dir, "cell_nvert", nc), cv = read_bin<int>(dir, "cell_vert", nc * 8),
         mat = read_bin<int>(dir, "matter", nc);
    auto fa = read_bin<double>(dir, "face_area", nc * 6), fnx = read_bin<double>(dir, "face_nx", nc * 6),
         fny = read_bin<double>(dir, "face_ny", nc * 6), fnz = read_bin<double>(dir, "face_nz", nc * 6),
         fcx = read_bin<double>(dir, "face_cx", nc * 6), fcy = read_bin<double>(dir, "face_cy", nc * 6),
         fcz = read_bin<double>(dir, "face_cz", nc * 6), vol = read_bin<double>(dir, "volume", nc),
         cx = read_bin<double>(dir, "cx", nc), cy = read_bin<double>(dir, "cy", nc), cz = read_bin<double>(dir, "cz", nc),
         heat = read_bin<double>(dir, "heat_source", nc), t0 = read_bin<double>(dir, "initial_temp", nc);
    auto nx = read_bin<double>(dir, "node_x", nn), ny = read_bin<double>(dir, "node_y", nn),
         nz = read_bin<double>(dir, "node_z", nn);
    auto bc = read_bin<int>(dir, "bound_cell", nb * 2), bf = read_bin<int>(dir, "bound_face", nb * 2),
         bt = read_bin<int>(dir, "bound_type", nb), bconn = read_bin<int>(dir, "bound_connect", nb);
    auto btemp = read_bin<double>(dir, "bound_temp", nb);
    auto ffc = read_bin<int>(dir, "fface_cell", nf * 2), fff = read_bin<int>(dir, "fface_face", nf * 2),
         ffb = read_bin<int>(dir, "fface_bound", nf);
    m.Nodes.resize(nn);
    for (size_t i = 0; i < nn; ++i) { m.Nodes[i].x = nx[i]; m.Nodes[i].y = ny[i]; m.Nodes[i].z = nz[i]; }
    m.Elements.resize(nc);
    for (size_t i = 0; i < nc; ++i) {
        Cell &c = m.Elements[i];
        c.volume = vol[i];
        c.index = 0;
        c.matter = mat[i];
        c.heat_source = heat[i];
        c.initial_temperature = t0[i];
        c.center = Point(cx[i], cy[i], cz[i]);
        c.vertexes.assign(cv.begin() + i * 8, cv.begin() + i * 8 + cnv[i]);
        c.faces.resize(fsize[i]);
        for (int j = 0; j < fsize[i]; ++j) {
            size_t k = i * 6 + j;
            Face &f = c.faces[j];
            f.area = fa[k];
            f.neighbor = fnb[k];
            f.bound = fbd[k];
            f.index = fidx[k];
            f.norm = vec(fnx[k], fny[k], fnz[k]);
            f.center = Point(fcx[k], fcy[k], fcz[k]);
            f.vertexes.assign(fv.begin() + k * 4, fv.begin() + k * 4 + fnv[k]);
        }
        for (int v : c.vertexes) m.Nodes[v].cells.push_back((int)i);
    }
    m.Boundaries.resize(nb);
    for (size_t i = 0; i < nb; ++i) {
        Boundary &b = m.Boundaries[i];
        b.index = (int)i;
        b.connection = bconn[i];
        b.type = bt[i];
        b.Temperature = btemp[i];
        for (int k = 0; k < 2; ++k)
            if (bc[i * 2 + k] >= 0) { b.cellindex.push_back(bc[i * 2 + k]); b.faceindex.push_back(bf[i * 2 + k]); }
    }
    m.Element_Faces.resize(nf);
    for (size_t i = 0; i < nf; ++i) {
        faces &f = m.Element_Faces[i];
        f.Boundindex = ffb[i];
        for (int k = 0; k < 2; ++k)
            if (ffc[i * 2 + k] >= 0) { f.cellindex.push_back(ffc[i * 2 + k]); f.faceindex.push_back(fff[i * 2 + k]); }
    }
}

// Timing run on a mesh built from tables: the reference's StaticBTESynthetic
// with its systems distributed over the MPI ranks (as BTEMain.cpp runs it with
// mpiexec), from the initial temperature, for max_iter outer iterations of
// solve() (StaticBTESyntheticSolver.cpp:12-247; rank 0 prints its
// per-iteration "Time taken by inner loop" line).  Nothing is dumped.
int time_from_tables(const Deck &deck, BTEBand *bands, BTEBoundaryCondition *bcs, const std::string &tables,
                     int maxIter) {
    int size = 1, rank = 0;
    MPI_Comm_size(MPI_COMM_WORLD, &size);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    BTEMesh mesh;
    mesh_from_tables(tables, mesh);
    BTEAngle *angles = new BTEAngle(deck.ntheta, deck.nphi, deck.dimMat, deck.angleMethod);
    StaticBTESynthetic solver(&mesh, bcs, bands, angles, size, rank);
    for (size_t i = 0; i < mesh.Elements.size(); ++i) solver.temperature[i] = mesh.Elements[i].initial_temperature;
    std::fill(solver.ebound, solver.ebound + (size_t)solver.numBand * solver.numDirection * solver.numBound * 2, 0.0);
    solver.error_old = 0;
    solver.errorOld = 0;
    solver.errorOld_heat = 0;
    solver.solve(0, maxIter, 0, 0.0, 0.0);
    MPI_Finalize();
    return 0;
}

// The steady flow of main on a mesh built from tables: _set_initial
// (Solution.cpp:78-124), the Fourier warm start (:236-260) and the BTE solve
// with its hand-over (:262-414), at np=1.
int run_from_tables(const Deck &deck, BTEBand *bands, BTEBoundaryCondition *bcs, const std::string &tables,
                    int maxIter, double tolT, double tolQ, int sb, int sd) {
    BTEMesh mesh;
    mesh_from_tables(tables, mesh);
    BTEAngle *angles = new BTEAngle(deck.ntheta, deck.nphi, deck.dimMat, deck.angleMethod);
    size_t nc = mesh.Elements.size(), nbd = mesh.Boundaries.size();
    std::vector<double> T(nc), qx(nc, 0.0), qy(nc, 0.0), qz(nc, 0.0);
    {
        StaticFourier fs(&mesh, bcs, bands, 1, 0);
        for (size_t i = 0; i < nc; ++i) {
            fs.temperature[i] = mesh.Elements[i].initial_temperature;
            fs.heatFluxXGlobal[i] = 0; fs.heatFluxYGlobal[i] = 0; fs.heatFluxZGlobal[i] = 0;
        }
        for (size_t ib = 0; ib < nbd; ++ib) {
            const Boundary &b = mesh.Boundaries[ib];
            fs.boundaryTemp[ib] = 0;
            fs.boundaryFlux[ib] = 0;
            if (b.type == 1) fs.boundaryTemp[ib] = b.Temperature;
        }
        for (size_t ib = 0; ib < nbd; ++ib) {
            const Boundary &b = mesh.Boundaries[ib];
            if (b.type < 0 && b.Temperature != 0) {
                fs.boundaryTemp[ib] = b.Temperature / 2;
                fs.boundaryTemp[b.connection] = -b.Temperature / 2;
            }
        }
        fs._solve();
        for (size_t i = 0; i < nc; ++i) {
            T[i] = fs.temperature[i];
            qx[i] = fs.heatFluxXGlobal[i]; qy[i] = fs.heatFluxYGlobal[i]; qz[i] = fs.heatFluxZGlobal[i];
        }
    }
    if (mesh.Dimension != 3) std::fill(qz.begin(), qz.end(), 0.0);
    dump_d("fourier_T", T);
    dump_d("fourier_qx", qx);
    dump_d("fourier_qy", qy);
    dump_d("fourier_qz", qz);
    auto hand_over = [&](StaticBTESynthetic &s) {
        for (size_t i = 0; i < nc; ++i) {
            s.temperature[i] = T[i];
            s.heatFluxXGlobal[i] = qx[i]; s.heatFluxYGlobal[i] = qy[i]; s.heatFluxZGlobal[i] = qz[i];
        }
    };
    if (sb >= 0) {
        StaticBTESynthetic st(&mesh, bcs, bands, angles, 1, 0);
        hand_over(st);
        st.error_old = 0;
        stage_dump(st, sb, sd);
    }
    StaticBTESynthetic solver(&mesh, bcs, bands, angles, 1, 0);
    hand_over(solver);
    std::fill(solver.ebound, solver.ebound + (size_t)solver.numBand * solver.numDirection * solver.numBound * 2, 0.0);
    solver.error_old = 0;
    solver.errorOld = 0;
    solver.errorOld_heat = 0;
    dump_tables(solver, "");
    bool seq = deck.method == "Sequential";
    if (deck.order == "1") {
        if (seq) solver.solve_DOM_firstorder(0, maxIter, 0, tolT, tolQ);
        else solver.solve_firstorder(0, maxIter, 0, tolT, tolQ);
    } else {
        int lim = deck.order == "2limit" ? 1 : 0;
        if (seq) solver.solve_DOM(0, maxIter, lim, tolT, tolQ);
        else solver.solve(0, maxIter, lim, tolT, tolQ);
    }
    size_t nB = solver.numBand, nD = solver.numDirection, nb = solver.numBound;
    dump_d("T_lattice", solver.temperature, {nc});
    dump_d("T_energy", solver.totalEnergy, {nc});
    dump_d("qx", solver.heatFluxXGlobal, {nc});
    dump_d("qy", solver.heatFluxYGlobal, {nc});
    dump_d("qz", solver.heatFluxZGlobal, {nc});
    dump_d("T_band", solver.temperatureband, {nB, nc});
    dump_d("ebound", solver.ebound, {nB, nD, nb, 2});
    dump_d("T_old", solver.temperatureOld, {nc});
    double errs[2] = {solver.error, solver.error_heat};
    dump_d("final_errors", errs, {2});
    MPI_Finalize();
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    if (argc < 6) {
        fprintf(stderr, "usage: ref_dump <deck_dir> <out_dir> <max_iter> <tolT> <tolQ> [stage_band stage_dir]\n");
        return 2;
    }
    char cwd0[4096];
    if (!getcwd(cwd0, sizeof cwd0)) return 2;
    std::string out = argv[2];
    if (out[0] != '/') out = std::string(cwd0) + "/" + out;
    g_out = out;
    mkdir(g_out.c_str(), 0755);
    if (chdir(argv[1]) != 0) { perror(argv[1]); return 2; }
    int maxIter = atoi(argv[3]);
    double tolT = atof(argv[4]), tolQ = atof(argv[5]);
    int sb = -1, sd = -1;
    std::string tables;  // --tables DIR: the mesh from flattened tables (3D decks)
    bool timing = false; // --time (with --tables): a timing run over the MPI ranks, nothing dumped
    std::vector<std::string> pos;
    for (int k = 6; k < argc; ++k) {
        if (std::string(argv[k]) == "--tables" && k + 1 < argc) {
            tables = argv[++k];
            if (tables[0] != '/') tables = std::string(cwd0) + "/" + tables;
        } else if (std::string(argv[k]) == "--time") {
            timing = true;
        } else {
            pos.push_back(argv[k]);
        }
    }
    if (pos.size() >= 2) { sb = atoi(pos[0].c_str()); sd = atoi(pos[1].c_str()); }

    Deck deck = parse_deck();
    if ((deck.order != "2" && deck.order != "1" && deck.order != "2limit") ||
        (deck.method != "Synthetic" && deck.method != "Sequential") ||
        (deck.matrix != "LU" && deck.matrix != "BICGSTAB")) {
        fprintf(stderr, "ref_dump covers Order=1/2/2limit Synthetic / Sequential LU / BICGSTAB only\n");
        return 3;
    }
    std::ifstream bandFile("input/PHONON");
    BTEBand *bands = new BTEBand(bandFile, deck.dimMat);
    bandFile.close();
    std::ifstream bcFile(deck.bcFile);
    BTEBoundaryCondition *bcs = new BTEBoundaryCondition(bcFile, bands);
    bcFile.close();
    if (!tables.empty() && timing) return time_from_tables(deck, bands, bcs, tables, maxIter);
    if (!tables.empty()) return run_from_tables(deck, bands, bcs, tables, maxIter, tolT, tolQ, sb, sd);
    std::ifstream heatfile(deck.heatFile), inittemp(""), geofile(deck.meshFile);
    DistributeMesh *dm = new DistributeMesh(deck.dimGeo, geofile, deck.Lx, deck.Ly, deck.Lz, deck.meshType, bands, bcs,
                                            heatfile, deck.uniformHeat, deck.heatType, inittemp, deck.tref);
    BTEAngle *angles = new BTEAngle(deck.ntheta, deck.nphi, deck.dimMat, deck.angleMethod);

    SolutionAll sa(dm, bcs, bands, angles, 1, 0);
    sa._set_initial(dm, bands, angles);
    size_t ncF = dm->volumeElementsFourier.size();
    // 2D Fourier runs never write heatFluxZ (StaticFourier.cpp:615-710); pin
    // it to zero so the hand-over below is defined.
    for (size_t i = 0; i < ncF; ++i) sa.solution.heatFluxZ[i] = 0.0;
    sa._Fourier_Solver(dm, bcs, bands, 1, 0);
    if (dm->FourierMeshes.Dimension != 3)
        for (size_t i = 0; i < ncF; ++i) sa.solution.heatFluxZ[i] = 0.0;
    dump_d("fourier_T", sa.solution.Temperature, {ncF});
    dump_d("fourier_qx", sa.solution.heatFluxX, {ncF});
    dump_d("fourier_qy", sa.solution.heatFluxY, {ncF});
    dump_d("fourier_qz", sa.solution.heatFluxZ, {ncF});

    std::ifstream heatfile2(deck.heatFile);
    dm->_build_BTEMesh(deck.dimGeo, deck.Lx, deck.Ly, deck.Lz, bands, bcs, heatfile2, deck.uniformHeat, deck.multiscale);

    if (sb >= 0) {
        StaticBTESynthetic st(&dm->BTEMeshes, bcs, bands, angles, 1, 0);
        init_solver_state(st, *dm, sa);
        st.error_old = 0;
        stage_dump(st, sb, sd);
    }

    StaticBTESynthetic solver(&dm->BTEMeshes, bcs, bands, angles, 1, 0);
    init_solver_state(solver, *dm, sa);
    // ebound is allocated without initialisation (StaticBTESynthetic.cpp:310).
    // Only the 2limit limiter reads entries before the first _set_bound_ee_1
    // writes them (the incoming side-0 values of the set-up gradients,
    // :1389-1399), so that run depends on what the allocation held: zero it,
    // which is what fresh pages give, so the fixture is defined.
    std::fill(solver.ebound, solver.ebound + (size_t)solver.numBand * solver.numDirection * solver.numBound * 2, 0.0);
    solver.error_old = 0;
    solver.errorOld = 0;
    solver.errorOld_heat = 0;
    dump_tables(solver, "");
    // the dispatch of Solution.cpp:287-340
    bool seq = deck.method == "Sequential", bicg = deck.matrix == "BICGSTAB";
    if (deck.order == "1") {
        if (seq && bicg) solver.solve_DOM_firstorder_Iterative(0, maxIter, 0, tolT, tolQ);  // :1598
        else if (seq) solver.solve_DOM_firstorder(0, maxIter, 0, tolT, tolQ);  // StaticBTESyntheticSolver.cpp:1394
        else if (bicg) solver.solve_firstorder_Iterative(0, maxIter, 0, tolT, tolQ);  // :727
        else solver.solve_firstorder(0, maxIter, 0, tolT, tolQ);         // :487
    } else {
        int lim = deck.order == "2limit" ? 1 : 0;  // Use_Limiter (Solution.cpp:363-400)
        if (seq && bicg) solver.solve_DOM_Iterative(0, maxIter, lim, tolT, tolQ);  // :1180
        else if (seq) solver.solve_DOM(0, maxIter, lim, tolT, tolQ);  // :968
        else if (bicg) solver.solve_Iterative(0, maxIter, lim, tolT, tolQ);  // :249
        else solver.solve(0, maxIter, lim, tolT, tolQ);
    }
    size_t nc = solver.numCell, nB = solver.numBand, nD = solver.numDirection, nb = solver.numBound;
    dump_d("T_lattice", solver.temperature, {nc});
    dump_d("T_energy", solver.totalEnergy, {nc});
    dump_d("qx", solver.heatFluxXGlobal, {nc});
    dump_d("qy", solver.heatFluxYGlobal, {nc});
    dump_d("qz", solver.heatFluxZGlobal, {nc});
    dump_d("T_band", solver.temperatureband, {nB, nc});
    dump_d("ebound", solver.ebound, {nB, nD, nb, 2});
    dump_d("T_old", solver.temperatureOld, {nc});
    double errs[2] = {solver.error, solver.error_heat};
    dump_d("final_errors", errs, {2});
    // the hand-back and the final outputs of main (Solution.cpp:408-413,
    // BTEMain.cpp:393): Temperature.dat, HeatFlux.dat and Result.vtk; the
    // solver wrote TempLattice.dat and Tempband.dat itself
    for (size_t i = 0; i < dm->volumeElementsBTE.size(); ++i) {
        sa.solution.Temperature[dm->elementProjection[i]] = solver.totalEnergy[i];
        sa.solution.heatFluxX[dm->elementProjection[i]] = solver.heatFluxXGlobal[i];
        sa.solution.heatFluxY[dm->elementProjection[i]] = solver.heatFluxYGlobal[i];
        sa.solution.heatFluxZ[dm->elementProjection[i]] = solver.heatFluxZGlobal[i];
    }
    sa._print_out(dm);
    MPI_Finalize();
    return 0;
}
