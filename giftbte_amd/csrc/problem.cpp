// Deck parsing, mesh ingestion and geometry for the steady BTE path.
//
// Every routine states the reference routine whose observable behaviour it
// reproduces; the tables it fills are compared entry by entry against the
// compiled reference in the test suite.  The reference keeps AoS objects
// (Cell/Face/Boundary, includes/utility/utility.h:27-137) and flattens them in
// the StaticBTESynthetic constructor; this file builds the flat tables
// directly.
#include "problem.hpp"

#include "bands.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

namespace gbte {
namespace {

struct V3 {
    double x = 0, y = 0, z = 0;
};

inline double dist(const V3 &a, const V3 &b) {
    return std::sqrt(std::pow(a.x - b.x, 2) + std::pow(a.y - b.y, 2) + std::pow(a.z - b.z, 2));
}

// Heron's formula, the reference's get_tri_area (src/utility/utility.cpp:87-95).
inline double tri_area(const V3 &p0, const V3 &p1, const V3 &p2) {
    double a = dist(p0, p1), b = dist(p1, p2), c = dist(p0, p2);
    double s = (a + b + c) / 2;
    return std::sqrt(s * (s - a) * (s - b) * (s - c));
}

bool has_kw(const std::string &s, const char *kw) {
    return s.find(kw) != std::string::npos && s.find('!') == std::string::npos;
}

std::string join(const std::string &dir, const std::string &f) {
    if (f.empty() || f[0] == '/') return f;
    return dir + "/" + f;
}

// src/BTEMain.cpp:87-315: a line that contains a keyword (and no '!') makes
// the next whitespace-separated token of the stream its value.
std::string parse_deck(const std::string &dir, Deck &d) {
    std::ifstream c(join(dir, "input/CONTROL"));
    if (!c.is_open()) return "Error: need CONTROL";
    std::string s;
    while (std::getline(c, s)) {
        if (has_kw(s, "State")) c >> d.state;
        if (has_kw(s, "Order")) c >> d.order;
        if (has_kw(s, "IterativeScheme")) c >> d.method;
        if (has_kw(s, "MatrixSolver")) c >> d.matrix;
        if (has_kw(s, "Tref")) c >> d.tref;
        if (has_kw(s, "MaxNumIter")) c >> d.max_iter;
        if (has_kw(s, "ResidualTemp")) c >> d.tol_temp;
        if (has_kw(s, "ResidualFlux")) c >> d.tol_flux;
        if (has_kw(s, "Multiscale")) c >> d.multiscale;
        if (has_kw(s, "InitialTemperatureFile")) return "InitialTemperatureFile is not supported by the reference either";
    }
    std::ifstream g(join(dir, "input/GEOMETRY"));
    if (!g.is_open()) return "Error: need GEOMETRY";
    while (std::getline(g, s)) {
        if (has_kw(s, "GeometryDimension")) g >> d.dim;
        if (has_kw(s, "ScaleX")) g >> d.Lx;
        if (has_kw(s, "ScaleY")) g >> d.Ly;
        if (has_kw(s, "ScaleZ")) g >> d.Lz;
        if (has_kw(s, "UniformHeat")) g >> d.uniform_heat;
        if (has_kw(s, "BCFile")) g >> d.bc_file;
        if (has_kw(s, "MeshType")) g >> d.mesh_type;
        if (has_kw(s, "MeshFile")) g >> d.mesh_file;
        if (has_kw(s, "HeatType")) g >> d.heat_type;
        if (has_kw(s, "HeatFile")) g >> d.heat_file;
    }
    if (d.dim == -1) return "Error: Please provide DimensionGeometry";
    std::ifstream ph(join(dir, "input/PHONON"));
    if (!ph.is_open()) return "Error: need PHONON";
    while (std::getline(ph, s)) {
        if (has_kw(s, "MaterialDimension")) ph >> d.material_dim;
        if (has_kw(s, "Angle_method")) ph >> d.angle_method;
        if (has_kw(s, "Ntheta")) ph >> d.ntheta;
        if (has_kw(s, "Nphi")) ph >> d.nphi;
    }
    return "";
}

struct Bands {
    int nband = 0, nmat = 0;
    std::vector<double> v, tau, C, lr, hr;  // [mat][band]
    std::vector<double> kappa, cbulk;
    std::vector<int> geo_matter;
};

// src/BTEBand/BTEBand.cpp:14-80 (DATABASE and inline rows), :81-740
// (ShengBTE / ALAMODE binning, bands.cpp) and :791-844.
std::string parse_bands(const std::string &dir, int dim_mat, Bands &B) {
    std::ifstream f(join(dir, "input/PHONON"));
    std::string s;
    while (std::getline(f, s)) {
        if (s.find("MaterialNumber") != std::string::npos && s.find('!') == std::string::npos) f >> B.nmat;
        if (s.find("BandNumber") != std::string::npos && s.find('!') == std::string::npos) f >> B.nband;
    }
    if (B.nmat <= 0 || B.nband <= 0) return "Error: please provide Number of matter and Number of bands";
    size_t n = (size_t)B.nmat * B.nband;
    B.v.assign(n, 0); B.tau.assign(n, 0); B.C.assign(n, 0); B.lr.assign(n, 0); B.hr.assign(n, 0);
    std::ifstream g(join(dir, "input/PHONON"));
    for (int m = 0; m < B.nmat; ++m) {
        while (std::getline(g, s)) {
            if (s.find("MaterialFile") == std::string::npos) continue;
            if (s.find("DATABASE") != std::string::npos && s.find("DATABASE") > 0) {
                std::string file;
                g >> file;
                std::ifstream db(join(dir, file));
                if (!db.is_open()) return "Error: DATABASE band file not open";
                for (int b = 0; b < B.nband; ++b) {
                    size_t k = (size_t)m * B.nband + b;
                    db >> B.v[k] >> B.tau[k] >> B.C[k] >> B.lr[k] >> B.hr[k];
                }
            } else if ((s.find("ShengBTE") != std::string::npos && s.find("ShengBTE") > 0) ||
                       (s.find("ALAMODE") != std::string::npos && s.find("ALAMODE") > 0)) {
                // spectral binning of ShengBTE / ALAMODE mode data (bands.cpp)
                bool sheng = s.find("ShengBTE") != std::string::npos && s.find("ShengBTE") > 0;
                std::string path;
                g >> path;
                BandColumns bc;
                std::string err = sheng ? bin_shengbte(join(dir, path), B.nband, bc)
                                        : bin_alamode(join(dir, path), s.length(), B.nband, bc);
                if (!err.empty()) return err;
                for (int b = 0; b < B.nband; ++b) {
                    size_t k = (size_t)m * B.nband + b;
                    B.v[k] = bc.v[b]; B.tau[k] = bc.tau[b]; B.C[k] = bc.C[b]; B.lr[k] = bc.lr[b]; B.hr[k] = bc.hr[b];
                }
            } else {
                for (int b = 0; b < B.nband; ++b) {
                    std::getline(g, s);
                    std::stringstream ss(s);
                    size_t k = (size_t)m * B.nband + b;
                    ss >> B.v[k] >> B.tau[k] >> B.C[k] >> B.lr[k] >> B.hr[k];
                }
            }
            break;
        }
    }
    B.kappa.assign(B.nmat, 0);
    B.cbulk.assign(B.nmat, 0);
    for (int m = 0; m < B.nmat; ++m) {
        double sumL = 0;
        for (int b = 0; b < B.nband; ++b) {
            size_t k = (size_t)m * B.nband + b;
            if (dim_mat == 2) B.kappa[m] += B.C[k] * B.v[k] * B.v[k] * B.tau[k] / 2;
            else if (dim_mat == 3) B.kappa[m] += B.C[k] * B.v[k] * B.v[k] * B.tau[k] / 3;
            B.cbulk[m] += B.C[k];
        }
        if (std::isnan(B.kappa[m])) return "Error: mistake in phonon properties, nan";
        for (int b = 0; b < B.nband; ++b) sumL += B.lr[(size_t)m * B.nband + b];
        if (std::fabs(sumL - 1.0) > 1e-6) return "Error: mistake in phonon properties, enegy conservation does not obey";
    }
    while (std::getline(g, s)) {
        if (s.find("GeometryMaterialType") == std::string::npos) continue;
        while (std::getline(g, s)) {
            if (s.length() <= 1) break;
            std::stringstream ss(s);
            int a = 0, b = 0;
            ss >> a >> b;
            if (std::abs(b) > 100) break;
            B.geo_matter.push_back(b - 1);
        }
    }
    return "";
}

// src/BTEBoundaryCondition/BTEBoundaryCondition.cpp:9-70
std::string parse_bcs(const std::string &dir, const std::string &file, int nband, std::vector<BoundaryCond> &out) {
    std::ifstream f(join(dir, file));
    if (!f.is_open()) return "Error: BC file not open";
    int num = 0;
    std::string line;
    f >> num;
    for (int i = 0; i < 3; ++i) std::getline(f, line);
    for (int i = 0; i < num; ++i) {
        BoundaryCond bc;
        f >> bc.index >> bc.type;
        if (bc.type == 4 || bc.type == 5) f >> bc.temperature >> bc.interface_file;
        else f >> bc.temperature;
        std::getline(f, line);
        out.push_back(bc);
    }
    for (auto &bc : out) {
        if (bc.type != 4 && bc.type != 5) continue;
        std::ifstream in(join(dir, bc.interface_file));
        if (!in.is_open()) return "Error: interface file not open";
        while (std::getline(in, line)) {
            size_t pos = line.find("matter");
            if (pos == std::string::npos || pos == 0) continue;
            std::pair<int, std::vector<double>> t;
            t.first = std::stoi(line.substr(0, line.find(' ')));
            while (std::getline(in, line)) {
                if (line.length() <= 1) break;
                std::stringstream ss(line);
                double v;
                ss >> v;
                t.second.push_back(v);
            }
            if ((int)t.second.size() != nband) return "Error: mismatch band vs interface";
            bc.matter_trans.push_back(t);
        }
    }
    return "";
}

struct RawMesh {
    std::vector<double> x, y, z;
    std::vector<std::vector<int>> cells;
    std::vector<int> cell_index;
    std::vector<std::vector<int>> bounds;
    std::vector<int> bound_index;
};

bool getline_find(std::istream &in, const char *key, std::string &s) {
    while (std::getline(in, s)) {
        size_t p = s.find(key);
        if (p != std::string::npos && p > 0) return true;
    }
    return false;
}

// src/BTEMesh/ReadinMesh.cpp:245-384 (COMSOL, 2D)
std::string read_comsol_2d(std::istream &in, double Lx, double Ly, RawMesh &m) {
    std::string s;
    std::getline(in, s);
    size_t p = s.find("COMSOL");
    if (p == std::string::npos || p == 0) return "Error: mesh file format error";
    if (!getline_find(in, "# sdim", s)) return "Error: mesh file format error";
    int dim = std::stoi(s.substr(0, s.find(' ')));
    if (dim != 2) return "Error: please check the dimension is correct or not";
    if (!getline_find(in, "# number of mesh", s)) return "Error: mesh file format error";
    int nn = std::stoi(s.substr(0, s.find(' ')));
    m.x.resize(nn); m.y.resize(nn); m.z.assign(nn, 0.0);
    for (int i = 0; i < 3; ++i) std::getline(in, s);
    for (int i = 0; i < nn; ++i) {
        in >> m.x[i] >> m.y[i];
        m.x[i] = m.x[i] * Lx;
        m.y[i] = m.y[i] * Ly;
    }
    if (!getline_find(in, "# number of element types", s)) return "Error: mesh file format error";
    int ntypes = std::stoi(s.substr(0, s.find(' ')));
    while (ntypes-- > 0) {
        int nv = 0;
        bool found = false;
        while (std::getline(in, s)) {
            size_t a = s.find("# number of vertices per element"), b = s.find("# number of nodes per element");
            if ((a != std::string::npos && a > 0) || (b != std::string::npos && b > 0)) {
                nv = std::stoi(s.substr(0, s.find(' ')));
                found = true;
                break;
            }
        }
        if (!found) break;
        if (nv < 2 || nv > 4) continue;
        if (!getline_find(in, "# number of elements", s)) return "Error: mesh file format error";
        int ne = std::stoi(s.substr(0, s.find(' ')));
        std::getline(in, s);
        size_t first = (nv == 2) ? m.bounds.size() : m.cells.size();
        for (int e = 0; e < ne; ++e) {
            std::vector<int> c(nv);
            for (int k = 0; k < nv; ++k) in >> c[k];
            if (nv == 2) m.bounds.push_back(c);
            else m.cells.push_back(c);
        }
        for (int i = 0; i < 4; ++i) std::getline(in, s);
        for (int e = 0; e < ne; ++e) {
            int idx;
            in >> idx;
            if (nv == 2) m.bound_index.push_back(idx + 1);
            else m.cell_index.push_back(idx);
        }
        (void)first;
    }
    for (auto &i : m.cell_index) i = i - 1;
    if (!in.good() && !in.eof()) return "Error: mesh file format error";
    return "";
}

// src/BTEMesh/ReadinMesh.cpp:387-532 (Gmsh MSH 4.1, 2D)
std::string read_msh_2d(std::istream &in, double Lx, double Ly, double Lz, RawMesh &m) {
    std::string s;
    std::getline(in, s);
    size_t p = s.find("MeshFormat");
    if (p == std::string::npos || p == 0) return "Error: mesh file format error";
    if (!getline_find(in, "Nodes", s)) return "Error: mesh file format error";
    int nblock, nn;
    in >> nblock >> nn;
    std::getline(in, s);
    m.x.assign(nn, 0); m.y.assign(nn, 0); m.z.assign(nn, 0);
    for (int b = 0; b < nblock; ++b) {
        int dim, tag, para, nin;
        in >> dim >> tag >> para >> nin;
        std::vector<int> tags(nin);
        for (int j = 0; j < nin; ++j) in >> tags[j];
        for (int j = 0; j < nin; ++j) {
            double x, y, z;
            in >> x >> y >> z;
            m.x[tags[j] - 1] = x * Lx;
            m.y[tags[j] - 1] = y * Ly;
            m.z[tags[j] - 1] = z * Lz;
        }
    }
    if (!getline_find(in, "Elements", s)) return "Error: mesh file format error";
    int ncell;
    in >> nblock >> ncell;
    std::getline(in, s);
    for (int b = 0; b < nblock; ++b) {
        int dim, tag, type, nin;
        in >> dim >> tag >> type >> nin;
        std::getline(in, s);
        for (int j = 0; j < nin; ++j) {
            int t;
            if (type == 1) {
                std::vector<int> a(2);
                in >> t >> a[0] >> a[1];
                a[0] -= 1; a[1] -= 1;
                m.bound_index.push_back(tag);
                m.bounds.push_back(a);
            } else if (type == 2) {
                std::vector<int> a(3);
                in >> t >> a[0] >> a[1] >> a[2];
                for (auto &v : a) v -= 1;
                m.cells.push_back(a);
                m.cell_index.push_back(tag - 1);
            } else if (type == 3) {
                std::vector<int> a(4);
                in >> t >> a[0] >> a[1] >> a[3] >> a[2];
                for (auto &v : a) v -= 1;
                m.cells.push_back(a);
                m.cell_index.push_back(tag - 1);
            } else {
                std::getline(in, s);
            }
            if (dim == 3) return "Error: find three dimensional element,please check the dimension is correct or not";
        }
    }
    return "";
}

// Gmsh MSH 4.1, 3D.  The reference's 3D readers were removed
// (src/BTEMesh/ReadinMesh.cpp:404-406, 531-533); this one reads the same
// subset as its 2D reader above: one node section, boundary triangles
// (element type 2, entity tag = BC index) and tetrahedra (type 4, entity
// tag - 1 = region).  Points and lines are skipped.
std::string read_msh_3d(std::istream &in, double Lx, double Ly, double Lz, RawMesh &m) {
    std::string s;
    std::getline(in, s);
    size_t p = s.find("MeshFormat");
    if (p == std::string::npos || p == 0) return "Error: mesh file format error";
    if (!getline_find(in, "Nodes", s)) return "Error: mesh file format error";
    long nblock, nn;
    in >> nblock >> nn;
    std::getline(in, s);
    if (!in || nn <= 0) return "Error: mesh file format error";
    m.x.assign(nn, 0); m.y.assign(nn, 0); m.z.assign(nn, 0);
    for (long b = 0; b < nblock; ++b) {
        long dim, tag, para, nin;
        in >> dim >> tag >> para >> nin;
        std::vector<long> tags(nin);
        for (long j = 0; j < nin; ++j) in >> tags[j];
        for (long j = 0; j < nin; ++j) {
            double x, y, z;
            in >> x >> y >> z;
            if (tags[j] < 1 || tags[j] > nn) return "Error: mesh node tag out of range";
            m.x[tags[j] - 1] = x * Lx;
            m.y[tags[j] - 1] = y * Ly;
            m.z[tags[j] - 1] = z * Lz;
        }
    }
    if (!in) return "Error: mesh file format error";
    if (!getline_find(in, "Elements", s)) return "Error: mesh file format error";
    long ncell;
    in >> nblock >> ncell;
    std::getline(in, s);
    for (long b = 0; b < nblock; ++b) {
        long dim, tag, type, nin;
        in >> dim >> tag >> type >> nin;
        std::getline(in, s);
        for (long j = 0; j < nin; ++j) {
            long t;
            if (type == 2) {
                std::vector<int> a(3);
                in >> t >> a[0] >> a[1] >> a[2];
                for (auto &v : a) v -= 1;
                m.bound_index.push_back((int)tag);
                m.bounds.push_back(a);
            } else if (type == 4) {
                std::vector<int> a(4);
                in >> t >> a[0] >> a[1] >> a[2] >> a[3];
                for (auto &v : a) v -= 1;
                m.cells.push_back(a);
                m.cell_index.push_back((int)tag - 1);
            } else if (type == 1 || type == 15) {
                std::getline(in, s);
            } else {
                return "Error: 3D meshes of tetrahedra only (Gmsh element type " + std::to_string(type) + ")";
            }
        }
    }
    if (!in && !in.eof()) return "Error: mesh file format error";
    return "";
}

inline bool contains(const std::vector<int> &v, int a) { return std::find(v.begin(), v.end(), a) != v.end(); }

struct CellG {
    double volume = 0;
    int index = 0, matter = 0;
    V3 center;
    std::vector<int> vert;
    struct F {
        double area = 0;
        int neighbor = -1, bound = -1, index = -1;
        std::vector<int> vert;
        V3 n, c;
    };
    std::vector<F> faces;
};

struct BoundG {
    int index = 0, connection = -1, type = 0;
    double temperature = 0;
    std::vector<int> cell, face;
    std::vector<std::vector<double>> trans;
};

// BTEMesh::setMeshParams, 2D branch (src/BTEMesh/BTEMesh.cpp:646-839).
std::string build_2d(const RawMesh &rm, const Bands &B, std::vector<V3> &nodes, std::vector<CellG> &cells,
                     std::vector<BoundG> &bounds, std::vector<std::vector<int>> &node_cells) {
    size_t nn = rm.x.size();
    nodes.resize(nn);
    for (size_t i = 0; i < nn; ++i) nodes[i] = {rm.x[i], rm.y[i], 0.0};
    cells.resize(rm.cells.size());
    for (size_t i = 0; i < cells.size(); ++i) {
        cells[i].vert = rm.cells[i];
        cells[i].index = rm.cell_index[i];
        if (cells[i].index < 0 || cells[i].index >= (int)B.geo_matter.size())
            return "Error: some region does not have material, check PHONON and GEOMETRY";
        cells[i].matter = B.geo_matter[cells[i].index];
        for (int v : cells[i].vert)
            if (v < 0 || v >= (int)nn) return "Error: mesh element references a missing node";
    }
    for (auto &c : cells) {
        const auto &v = c.vert;
        if (v.size() == 3) {
            c.volume = tri_area(nodes[v[0]], nodes[v[1]], nodes[v[2]]);
            c.center.x = ((nodes[v[0]].x + nodes[v[1]].x) + nodes[v[2]].x) / 3;
            c.center.y = ((nodes[v[0]].y + nodes[v[1]].y) + nodes[v[2]].y) / 3;
            c.center.z = ((nodes[v[0]].z + nodes[v[1]].z) + nodes[v[2]].z) / 3;
        } else if (v.size() == 4) {
            c.volume = tri_area(nodes[v[0]], nodes[v[1]], nodes[v[2]]);
            c.volume += tri_area(nodes[v[1]], nodes[v[2]], nodes[v[3]]);
            c.center.x = (((nodes[v[0]].x + nodes[v[1]].x) + nodes[v[2]].x) + nodes[v[3]].x) / 4;
            c.center.y = (((nodes[v[0]].y + nodes[v[1]].y) + nodes[v[2]].y) + nodes[v[3]].y) / 4;
            c.center.z = (((nodes[v[0]].z + nodes[v[1]].z) + nodes[v[2]].z) + nodes[v[3]].z) / 4;
        } else {
            return "Error: 2D cells must have 3 or 4 vertices";
        }
        int nf = (int)v.size();
        c.faces.resize(nf);
        for (int j = 0; j < nf; ++j) {
            auto &f = c.faces[j];
            if (nf == 4) {
                static const int q[4][2] = {{0, 1}, {1, 3}, {3, 2}, {2, 0}};
                f.vert = {v[q[j][0]], v[q[j][1]]};
            } else {
                f.vert = {v[(j + 1) % 3], v[(j + 2) % 3]};
            }
            const V3 &a0 = nodes[f.vert[0]], &a1 = nodes[f.vert[1]];
            f.area = dist(a0, a1);
            f.c.x = (a0.x + a1.x) / 2;
            f.c.y = (a0.y + a1.y) / 2;
            f.c.z = (a0.z + a1.z) / 2;
            double ax = (a0.x - a1.x) / f.area, ay = (a0.y - a1.y) / f.area;
            f.n = {ay, -ax, 0.0};
            double bx = c.center.x - f.c.x, by = c.center.y - f.c.y, bz = c.center.z - f.c.z;
            if (bx * f.n.x + by * f.n.y + bz * f.n.z > 0) {
                f.n.x = -f.n.x;
                f.n.y = -f.n.y;
            }
        }
    }
    node_cells.assign(nn, {});
    for (size_t i = 0; i < cells.size(); ++i)
        for (int v : cells[i].vert) node_cells[v].push_back((int)i);
    for (size_t i = 0; i < cells.size(); ++i)
        for (auto &f : cells[i].faces) {
            f.neighbor = -1;
            for (int k : node_cells[f.vert[0]])
                if (k != (int)i && contains(cells[k].vert, f.vert[0]) && contains(cells[k].vert, f.vert[1]))
                    f.neighbor = k;
            f.bound = -1;
        }
    bounds.resize(rm.bounds.size());
    for (size_t i = 0; i < rm.bounds.size(); ++i) {
        bounds[i].index = rm.bound_index[i];
        const auto &be = rm.bounds[i];
        if (be.size() < 2 || be[0] < 0 || be[0] >= (int)nn || be[1] < 0 || be[1] >= (int)nn)
            return "Error: boundary element references a missing node";
        for (int k : node_cells[be[0]])
            for (size_t j = 0; j < cells[k].faces.size(); ++j) {
                const auto &fv = cells[k].faces[j].vert;
                if ((be[1] == fv[0] || be[1] == fv[1]) && (be[0] == fv[0] || be[0] == fv[1])) {
                    bounds[i].cell.push_back(k);
                    bounds[i].face.push_back((int)j);
                }
            }
    }
    return "";
}

// Tetrahedral geometry, the counterpart of build_2d for the removed 3D
// branch of BTEMesh::setMeshParams (src/BTEMesh/BTEMesh.cpp:841-843):
//   volume   |det(v0 - v3, v1 - v3, v2 - v3)| / 6, the closed form of the
//            reference's get_tet_volume (src/utility/utility.cpp:71-86);
//   centre   the vertex mean; face j the triangle opposite vertex j,
//            (v[j+1], v[j+2], v[j+3]) mod 4, area by Heron's formula
//            (get_tri_area), centre the vertex mean, unit normal the
//            normalised cross product turned away from the cell centre;
//   neighbour the other cell holding the face's three vertices, and each
//            boundary triangle the (cell, face) holding its vertices.
std::string build_3d(const RawMesh &rm, const Bands &B, std::vector<V3> &nodes, std::vector<CellG> &cells,
                     std::vector<BoundG> &bounds, std::vector<std::vector<int>> &node_cells) {
    size_t nn = rm.x.size();
    nodes.resize(nn);
    for (size_t i = 0; i < nn; ++i) nodes[i] = {rm.x[i], rm.y[i], rm.z[i]};
    cells.resize(rm.cells.size());
    for (size_t i = 0; i < cells.size(); ++i) {
        cells[i].vert = rm.cells[i];
        cells[i].index = rm.cell_index[i];
        if (cells[i].index < 0 || cells[i].index >= (int)B.geo_matter.size())
            return "Error: some region does not have material, check PHONON and GEOMETRY";
        cells[i].matter = B.geo_matter[cells[i].index];
        if (cells[i].vert.size() != 4) return "Error: 3D cells must be tetrahedra";
        for (int v : cells[i].vert)
            if (v < 0 || v >= (int)nn) return "Error: mesh element references a missing node";
    }
    for (auto &c : cells) {
        const auto &v = c.vert;
        const V3 &p0 = nodes[v[0]], &p1 = nodes[v[1]], &p2 = nodes[v[2]], &p3 = nodes[v[3]];
        double a[3][3] = {{p0.x - p3.x, p1.x - p3.x, p2.x - p3.x},
                          {p0.y - p3.y, p1.y - p3.y, p2.y - p3.y},
                          {p0.z - p3.z, p1.z - p3.z, p2.z - p3.z}};
        double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                     a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
        c.volume = std::fabs(det) / 6;
        if (!(c.volume > 0)) return "Error: degenerate tetrahedron";
        c.center.x = (((p0.x + p1.x) + p2.x) + p3.x) / 4;
        c.center.y = (((p0.y + p1.y) + p2.y) + p3.y) / 4;
        c.center.z = (((p0.z + p1.z) + p2.z) + p3.z) / 4;
        c.faces.resize(4);
        for (int j = 0; j < 4; ++j) {
            auto &f = c.faces[j];
            f.vert = {v[(j + 1) % 4], v[(j + 2) % 4], v[(j + 3) % 4]};
            const V3 &a0 = nodes[f.vert[0]], &a1 = nodes[f.vert[1]], &a2 = nodes[f.vert[2]];
            f.area = tri_area(a0, a1, a2);
            f.c.x = ((a0.x + a1.x) + a2.x) / 3;
            f.c.y = ((a0.y + a1.y) + a2.y) / 3;
            f.c.z = ((a0.z + a1.z) + a2.z) / 3;
            double ux = a1.x - a0.x, uy = a1.y - a0.y, uz = a1.z - a0.z;
            double wx = a2.x - a0.x, wy = a2.y - a0.y, wz = a2.z - a0.z;
            double nx = uy * wz - uz * wy, ny = uz * wx - ux * wz, nz = ux * wy - uy * wx;
            double len = std::sqrt(nx * nx + ny * ny + nz * nz);
            f.n = {nx / len, ny / len, nz / len};
            double bx = c.center.x - f.c.x, by = c.center.y - f.c.y, bz = c.center.z - f.c.z;
            if (bx * f.n.x + by * f.n.y + bz * f.n.z > 0) f.n = {-f.n.x, -f.n.y, -f.n.z};
        }
    }
    node_cells.assign(nn, {});
    for (size_t i = 0; i < cells.size(); ++i)
        for (int v : cells[i].vert) node_cells[v].push_back((int)i);
    auto holds = [&](int k, const std::vector<int> &fv) {
        const auto &cv = cells[k].vert;
        return contains(cv, fv[0]) && contains(cv, fv[1]) && contains(cv, fv[2]);
    };
    for (size_t i = 0; i < cells.size(); ++i)
        for (auto &f : cells[i].faces) {
            f.neighbor = -1;
            for (int k : node_cells[f.vert[0]])
                if (k != (int)i && holds(k, f.vert)) f.neighbor = k;
            f.bound = -1;
        }
    bounds.resize(rm.bounds.size());
    for (size_t i = 0; i < rm.bounds.size(); ++i) {
        bounds[i].index = rm.bound_index[i];
        const auto &be = rm.bounds[i];
        for (int v : be)
            if (v < 0 || v >= (int)nn) return "Error: boundary element references a missing node";
        for (int k : node_cells[be[0]])
            for (size_t j = 0; j < cells[k].faces.size(); ++j) {
                const auto &fv = cells[k].faces[j].vert;
                if (contains(fv, be[0]) && contains(fv, be[1]) && contains(fv, be[2])) {
                    bounds[i].cell.push_back(k);
                    bounds[i].face.push_back((int)j);
                }
            }
    }
    return "";
}

// BTEMesh::setMeshParams1 (src/BTEMesh/BTEMesh.cpp:845-1023) plus the
// boundary-condition extension of DistributeMesh::_build_BTEMesh
// (src/BTEMesh/Distribute.cpp:158-196), which is a no-op on a valid deck
// without a multiscale file.
std::string apply_bcs(const std::vector<BoundaryCond> &bcs, const std::vector<V3> &nodes, std::vector<CellG> &cells,
                      std::vector<BoundG> &bounds, int nband) {
    (void)nodes;
    std::vector<BoundG> kept;
    for (auto &b : bounds)
        for (size_t j = 0; j < bcs.size(); ++j)
            if (b.index == bcs[j].index) {
                BoundG c = b;
                c.index = (int)j;
                kept.push_back(c);
            }
    bounds.swap(kept);
    for (size_t i = 0; i < bounds.size(); ++i)
        for (size_t j = 0; j < bounds[i].cell.size(); ++j) cells[bounds[i].cell[j]].faces[bounds[i].face[j]].bound = (int)i;
    for (size_t i = 0; i < cells.size(); ++i)
        for (size_t j = 0; j < cells[i].faces.size(); ++j)
            if (cells[i].faces[j].bound == -1 && cells[i].faces[j].neighbor == -1) {
                char msg[160];
                snprintf(msg, sizeof msg,
                         "cell %zu face %zu is neither boundary nor inner face, please check whether your boundary condition set is correct or not",
                         i, j);
                return msg;
            }
    for (auto &b : bounds)
        if (b.cell.empty()) return "Error: boundary element without a cell";
    for (auto &b : bounds) {
        b.connection = -1;
        if (bcs[b.index].type < 0) {
            double best = 1000;
            const V3 &c0 = cells[b.cell[0]].faces[b.face[0]].c;
            for (size_t j = 0; j < bounds.size(); ++j)
                if (bcs[b.index].index == -bcs[bounds[j].index].type) {
                    double d1 = dist(c0, cells[bounds[j].cell[0]].faces[bounds[j].face[0]].c);
                    if (d1 < best) {
                        best = d1;
                        b.connection = (int)j;
                    }
                }
            if (b.connection < 0) return "Error: periodic boundary without a partner";
        }
    }
    for (auto &b : bounds) {
        b.type = bcs[b.index].type;
        b.temperature = bcs[b.index].temperature;
        b.trans.assign(b.cell.size(), std::vector<double>());
        for (size_t i = 0; i < b.cell.size(); ++i)
            for (const auto &mt : bcs[b.index].matter_trans)
                if (cells[b.cell[i]].matter == mt.first) b.trans[i] = mt.second;
        if (b.type == 4 || b.type == 5)
            for (auto &t : b.trans)
                if ((int)t.size() != nband) return "Error: interface boundary without transmissivities for its matter";
    }
    // unique faces in first-encounter order
    std::vector<std::vector<int>> pending(cells.size());
    for (size_t i = 0; i < cells.size(); ++i) pending[i].assign(cells[i].faces.size(), 1);
    int nf = 0;
    for (size_t i = 0; i < cells.size(); ++i)
        for (size_t j = 0; j < cells[i].faces.size(); ++j) {
            if (!pending[i][j]) continue;
            auto &f = cells[i].faces[j];
            f.index = nf;
            if (f.bound == -1) {
                auto &nb = cells[f.neighbor];
                for (size_t k = 0; k < nb.faces.size(); ++k)
                    if (nb.faces[k].neighbor == (int)i) {
                        pending[f.neighbor][k] = 0;
                        nb.faces[k].index = nf;
                    }
            }
            ++nf;
        }
    return "";
}

// BTEAngle (src/BTEAngle/BTEAngle.cpp:11-190).
std::string make_angles(int ntheta, int nphi, int dim_mat, int method, std::vector<double> &dx, std::vector<double> &dy,
                        std::vector<double> &dz, std::vector<double> &w, double &solid) {
    const double PI = M_PI;
    if (method == 1) {
        if (dim_mat == 2) {
            solid = 2 * PI;
            int n = 4 * nphi;
            double dphi = 0.5 * PI / nphi;
            std::vector<double> phi(n);
            phi[0] = 0.5 * dphi;
            for (int k = 1; k < n; ++k) phi[k] = phi[k - 1] + dphi;
            for (int k = 0; k < n; ++k) {
                w.push_back(solid / n);
                dx.push_back(std::cos(phi[k]));
                dy.push_back(std::sin(phi[k]));
                dz.push_back(0);
            }
        } else if (dim_mat == 3) {
            solid = 4 * PI;
            double dth = 0.5 * PI / ntheta, dph = 0.5 * PI / nphi;
            std::vector<double> th(2 * ntheta), ph(4 * nphi);
            th[0] = 0.5 * dth;
            ph[0] = 0.5 * dph;
            for (int k = 1; k < 2 * ntheta; ++k) th[k] = th[k - 1] + dth;
            for (int k = 1; k < 4 * nphi; ++k) ph[k] = ph[k - 1] + dph;
            for (int t = 0; t < 2 * ntheta; ++t)
                for (int p = 0; p < 4 * nphi; ++p) {
                    w.push_back(2. * std::fabs(std::sin(th[t])) * std::sin(0.5 * dth) * dph);
                    dx.push_back(std::sin(th[t]) * std::cos(ph[p]));
                    dy.push_back(std::sin(th[t]) * std::sin(ph[p]));
                    dz.push_back(std::cos(th[t]));
                }
        } else {
            return "Error: MaterialDimension must be 2 or 3";
        }
    } else {
        if (dim_mat == 2) {
            solid = 2 * PI;
            std::vector<double> g = gauss_legendre(0, PI / 2, nphi);
            std::vector<double> ct, st, W;
            for (int k = 0; k < nphi; ++k) {
                ct.push_back(std::cos(g[2 * k])); W.push_back(g[2 * k + 1]); st.push_back(std::pow(1 - ct[k] * ct[k], 0.5));
            }
            for (int k = nphi - 1; k >= 0; --k) {
                ct.push_back(-std::cos(g[2 * k])); W.push_back(g[2 * k + 1]); st.push_back(std::pow(1 - ct[k] * ct[k], 0.5));
            }
            for (int k = 0; k < nphi; ++k) {
                ct.push_back(-std::cos(g[2 * k])); W.push_back(g[2 * k + 1]); st.push_back(-std::pow(1 - ct[k] * ct[k], 0.5));
            }
            for (int k = nphi - 1; k >= 0; --k) {
                ct.push_back(std::cos(g[2 * k])); W.push_back(g[2 * k + 1]); st.push_back(-std::pow(1 - ct[k] * ct[k], 0.5));
            }
            for (int k = 0; k < 4 * nphi; ++k) {
                w.push_back(W[k]); dx.push_back(ct[k]); dy.push_back(st[k]); dz.push_back(0);
            }
        } else if (dim_mat == 3) {
            solid = 4 * PI;
            std::vector<double> g = gauss_legendre(-1, 0, ntheta), gp = gauss_legendre(0, PI / 2, nphi);
            std::vector<double> ct, st, W, cp, sp, Wp;
            for (int k = 0; k < ntheta; ++k) {
                ct.push_back(-g[2 * k]); W.push_back(g[2 * k + 1]); st.push_back(std::pow(1 - ct[k] * ct[k], 0.5));
            }
            for (int k = ntheta - 1; k >= 0; --k) {
                ct.push_back(g[2 * k]); W.push_back(g[2 * k + 1]); st.push_back(std::pow(1 - ct[k] * ct[k], 0.5));
            }
            for (int k = 0; k < nphi; ++k) {
                cp.push_back(std::cos(gp[2 * k])); Wp.push_back(gp[2 * k + 1]); sp.push_back(std::pow(1 - cp[k] * cp[k], 0.5));
            }
            for (int k = nphi - 1; k >= 0; --k) {
                cp.push_back(-std::cos(gp[2 * k])); Wp.push_back(gp[2 * k + 1]); sp.push_back(std::pow(1 - cp[k] * cp[k], 0.5));
            }
            for (int k = 0; k < nphi; ++k) {
                cp.push_back(-std::cos(gp[2 * k])); Wp.push_back(gp[2 * k + 1]); sp.push_back(-std::pow(1 - cp[k] * cp[k], 0.5));
            }
            for (int k = nphi - 1; k >= 0; --k) {
                cp.push_back(std::cos(gp[2 * k])); Wp.push_back(gp[2 * k + 1]); sp.push_back(-std::pow(1 - cp[k] * cp[k], 0.5));
            }
            for (int t = 0; t < 2 * ntheta; ++t)
                for (int p = 0; p < 4 * nphi; ++p) {
                    w.push_back(W[t] * Wp[p]);
                    dx.push_back(st[t] * cp[p]);
                    dy.push_back(st[t] * sp[p]);
                    dz.push_back(ct[t]);
                }
        } else {
            return "Error: MaterialDimension must be 2 or 3";
        }
    }
    return "";
}

}  // namespace

// Newton iteration on the Legendre three-term recurrence from Chebyshev
// starting points, in the form of the published "lgwt" routine
// (G. von Winckel, 2004), which the reference's GaussIntegrationPoints
// (src/utility/utility.cpp:116-172) follows: the derivative carries the
// factor (N+2) and the weights the correction ((N+2)/(N+1))^2.
std::vector<double> gauss_legendre(double a, double b, int n) {
    int N = n - 1, N1 = N + 1, N2 = N + 2;
    std::vector<double> y(N1), y0(N1, 2.0), Lp(N1);
    std::vector<double> L((size_t)N1 * N2);
    for (int i = 0; i < N1; ++i) y[i] = std::cos((2 * (N - i) + 1) * M_PI / (2 * N + 2));
    auto margin = [&]() {
        double r = 0;
        for (int i = 0; i < N1; ++i) r = std::max(std::fabs(y[i] - y0[i]), r);
        return r;
    };
    const double eps = std::pow(2, -52);
    for (int i = 0; i < N1; ++i) L[(size_t)i * N2] = 1;
    while (margin() > eps) {
        for (int i = 0; i < N1; ++i) L[(size_t)i * N2 + 1] = y[i];
        for (int j = 2; j < N1 + 1; ++j)
            for (int i = 0; i < N1; ++i)
                L[(size_t)i * N2 + j] = ((2 * j - 1) * y[i] * L[(size_t)i * N2 + j - 1] - (j - 1) * L[(size_t)i * N2 + j - 2]) / j;
        for (int i = 0; i < N1; ++i) {
            Lp[i] = (N2) * (L[(size_t)i * N2 + N1 - 1] - y[i] * L[(size_t)i * N2 + N2 - 1]) / (1 - y[i] * y[i]);
            y0[i] = y[i];
            y[i] = y0[i] - L[(size_t)i * N2 + N2 - 1] / Lp[i];
        }
    }
    std::vector<double> out;
    for (int i = 0; i < N1; ++i) {
        out.push_back((a * (1 - y[i]) + b * (1 + y[i])) / 2);
        out.push_back((b - a) / ((1 - y[i] * y[i]) * Lp[i] * Lp[i]) * N2 / N1 * N2 / N1);
    }
    return out;
}

std::string load_problem(const std::string &dir, Problem &P) {
    Deck &d = P.deck;
    std::string err = parse_deck(dir, d);
    if (!err.empty()) return err;
    if (d.state != "Steady") return "State=" + d.state + " is not supported: this build is the steady solver";
    // the dispatch of SolutionAll::_BTE_Solver (Solution.cpp:287-406), which
    // exits on any other value
    if (d.order != "1" && d.order != "2" && d.order != "2limit") return "no this Order: " + d.order;
    if (d.method != "Synthetic" && d.method != "Sequential") return "no this IterativeScheme: " + d.method;
    if (d.matrix != "LU" && d.matrix != "BICGSTAB") return "no this MatrixSolver: " + d.matrix;
    if (d.dim != 2 && d.dim != 3)
        return "GeometryDimension " + std::to_string(d.dim) + " decks are not supported (2D and 3D only)";
    if (d.dim == 3 && d.mesh_type != "MSH") return "3D decks are read from Gmsh 4.1 meshes (MeshType MSH)";
    if (!d.multiscale.empty()) return "Multiscale Fourier/BTE region files are not supported";
    Bands B;
    err = parse_bands(dir, d.material_dim, B);
    if (!err.empty()) return err;
    std::vector<BoundaryCond> bcs;
    err = parse_bcs(dir, d.bc_file, B.nband, bcs);
    if (!err.empty()) return err;
    RawMesh rm;
    {
        std::ifstream mf(join(dir, d.mesh_file));
        if (!mf.is_open()) return "Error: Mesh file not open";
        if (d.mesh_type == "COMSOL") err = read_comsol_2d(mf, d.Lx, d.Ly, rm);
        else if (d.mesh_type == "MSH" && d.dim == 3) err = read_msh_3d(mf, d.Lx, d.Ly, d.Lz, rm);
        else if (d.mesh_type == "MSH") err = read_msh_2d(mf, d.Lx, d.Ly, d.Lz, rm);
        else err = "Error: does not support this mesh type";
        if (!err.empty()) return err;
    }
    std::vector<V3> nodes;
    std::vector<CellG> cells;
    std::vector<BoundG> bounds;
    std::vector<std::vector<int>> node_cells;
    err = d.dim == 3 ? build_3d(rm, B, nodes, cells, bounds, node_cells) : build_2d(rm, B, nodes, cells, bounds, node_cells);
    if (!err.empty()) return err;
    err = apply_bcs(bcs, nodes, cells, bounds, B.nband);
    if (!err.empty()) return err;

    // heat source: BTEMesh_heatin (src/BTEMesh/BTEMesh.cpp:1056-1107)
    std::vector<double> heat(cells.size(), d.uniform_heat);
    if (!d.heat_file.empty()) {
        std::ifstream hf(join(dir, d.heat_file));
        if (hf.is_open()) {
            if (d.heat_type == "REGION") {
                int nm = 0;
                hf >> nm;
                std::string s;
                std::getline(hf, s);
                std::vector<int> idx(nm);
                std::vector<double> hs(nm);
                for (int i = 0; i < nm; ++i) hf >> idx[i] >> hs[i];
                for (size_t i = 0; i < cells.size(); ++i) {
                    if (cells[i].index > nm || cells[i].index >= nm) return "Wrong heat file !";
                    heat[i] = hs[cells[i].index];
                }
            } else if (d.heat_type == "COORDINATE") {
                return "HeatType COORDINATE is not supported by the reference either";
            }
        }
    }

    // ---- flatten (StaticBTESynthetic.cpp:224-626)
    int nc = (int)cells.size(), nb = (int)bounds.size(), nn = (int)nodes.size();
    P.dim = d.dim;
    P.n_cells = nc;
    P.n_nodes = nn;
    P.n_bounds = nb;
    P.n_bands = B.nband;
    P.n_matters = B.nmat;
    P.face_size.assign(nc, 0);
    P.face_neighbor.assign((size_t)nc * 6, 0);
    P.face_bound.assign((size_t)nc * 6, 0);
    P.face_index.assign((size_t)nc * 6, 0);
    P.face_nvert.assign((size_t)nc * 6, 0);
    P.face_vert.assign((size_t)nc * 24, -1);
    P.cell_nvert.assign(nc, 0);
    P.cell_vert.assign((size_t)nc * 8, -1);
    P.face_area.assign((size_t)nc * 6, 0);
    for (auto *v : {&P.face_nx, &P.face_ny, &P.face_nz, &P.face_cx, &P.face_cy, &P.face_cz}) v->assign((size_t)nc * 6, 0);
    P.volume.resize(nc); P.cx.resize(nc); P.cy.resize(nc); P.cz.resize(nc);
    P.heat_source = heat;
    P.initial_temp.assign(nc, d.tref);
    P.matter.resize(nc);
    P.cell_region.resize(nc);
    for (int i = 0; i < nc; ++i) {
        const auto &c = cells[i];
        P.face_size[i] = (int)c.faces.size();
        for (size_t j = 0; j < c.faces.size(); ++j) {
            size_t k = (size_t)i * 6 + j;
            const auto &f = c.faces[j];
            P.face_area[k] = f.area;
            P.face_neighbor[k] = f.neighbor;
            P.face_index[k] = f.index;
            P.face_bound[k] = f.bound;
            P.face_nx[k] = f.n.x; P.face_ny[k] = f.n.y; P.face_nz[k] = f.n.z;
            P.face_cx[k] = f.c.x; P.face_cy[k] = f.c.y; P.face_cz[k] = f.c.z;
            P.face_nvert[k] = (int)f.vert.size();
            for (size_t v = 0; v < f.vert.size(); ++v) P.face_vert[k * 4 + v] = f.vert[v];
        }
        P.cell_nvert[i] = (int)c.vert.size();
        for (size_t v = 0; v < c.vert.size(); ++v) P.cell_vert[(size_t)i * 8 + v] = c.vert[v];
        P.volume[i] = c.volume;
        P.cx[i] = c.center.x; P.cy[i] = c.center.y; P.cz[i] = c.center.z;
        P.matter[i] = c.matter;
        P.cell_region[i] = c.index;
    }
    // same-matter vertex neighbours (StaticBTESynthetic.cpp:453-465)
    P.nbr_ptr.assign(nc + 1, 0);
    P.nbr_idx.clear();
    {
        std::vector<int> stamp(nc, -1);  // stamp[k] == i: k is already in cell i's list
        for (int i = 0; i < nc; ++i) {
            for (int v : cells[i].vert)
                for (int k : node_cells[v])
                    if (stamp[k] != i && cells[k].matter == cells[i].matter && k != i) {
                        stamp[k] = i;
                        P.nbr_idx.push_back(k);
                    }
            P.nbr_ptr[i + 1] = (int)P.nbr_idx.size();
        }
    }
    P.node_x.resize(nn); P.node_y.resize(nn); P.node_z.resize(nn);
    for (int i = 0; i < nn; ++i) { P.node_x[i] = nodes[i].x; P.node_y[i] = nodes[i].y; P.node_z[i] = nodes[i].z; }
    P.bound_cell.assign((size_t)nb * 2, -1);
    P.bound_face.assign((size_t)nb * 2, -1);
    P.bound_type.resize(nb);
    P.bound_temp.resize(nb);
    P.bound_connect.resize(nb);
    P.bound_trans.assign((size_t)nb * B.nband * 2, 0.0);
    for (int i = 0; i < nb; ++i) {
        const auto &b = bounds[i];
        for (size_t j = 0; j < b.cell.size() && j < 2; ++j) {
            P.bound_cell[(size_t)i * 2 + j] = b.cell[j];
            P.bound_face[(size_t)i * 2 + j] = b.face[j];
        }
        P.bound_type[i] = b.type;
        P.bound_temp[i] = b.temperature;
        P.bound_connect[i] = b.connection;
        if (b.type == 4 || b.type == 5)
            for (int band = 0; band < B.nband; ++band)
                for (size_t k = 0; k < 2 && k < b.trans.size(); ++k) P.bound_trans[((size_t)i * B.nband + band) * 2 + k] = b.trans[k][band];
    }
    // unique faces
    int nf = 0;
    for (int i = 0; i < nc; ++i)
        for (int j = 0; j < P.face_size[i]; ++j) nf = std::max(nf, P.face_index[(size_t)i * 6 + j] + 1);
    P.n_faces = nf;
    P.fface_cell.assign((size_t)nf * 2, -1);
    P.fface_face.assign((size_t)nf * 2, -1);
    P.fface_bound.assign(nf, -1);
    {
        std::vector<int> seen(nf, 0);
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < P.face_size[i]; ++j) {
                size_t k = (size_t)i * 6 + j;
                int f = P.face_index[k];
                int s = seen[f]++;
                if (s == 0) {
                    P.fface_cell[(size_t)f * 2] = i;
                    P.fface_face[(size_t)f * 2] = j;
                    P.fface_bound[f] = P.face_bound[k];
                } else if (P.face_bound[(size_t)P.fface_cell[(size_t)f * 2] * 6 + P.fface_face[(size_t)f * 2]] == -1) {
                    P.fface_cell[(size_t)f * 2 + 1] = i;
                    P.fface_face[(size_t)f * 2 + 1] = j;
                }
            }
    }
    // bands and directions
    size_t nmb = (size_t)B.nmat * B.nband;
    P.kappa_bulk = B.kappa;
    P.capacity_bulk = B.cbulk;
    P.vg.assign(B.v.begin(), B.v.begin() + nmb);
    P.tau = B.tau; P.cap = B.C; P.lr = B.lr; P.hr = B.hr;
    double solid = 0;
    std::vector<double> w;
    err = make_angles(d.ntheta, d.nphi, d.material_dim, d.angle_method, P.dir_x, P.dir_y, P.dir_z, w, solid);
    if (!err.empty()) return err;
    P.n_dirs = (int)w.size();
    P.weight.resize(w.size());
    for (size_t k = 0; k < w.size(); ++k) P.weight[k] = w[k] / solid;
    return "";
}

gbte_tables Problem::tables() const {
    gbte_tables t{};
    t.dim = dim;
    t.n_cells = n_cells; t.n_nodes = n_nodes; t.n_bounds = n_bounds; t.n_faces = n_faces;
    t.n_bands = n_bands; t.n_dirs = n_dirs; t.n_matters = n_matters;
    t.material_dim = deck.material_dim;
    t.Lx = deck.Lx; t.Ly = deck.Ly; t.Lz = deck.Lz;
    t.face_size = face_size.data(); t.face_neighbor = face_neighbor.data(); t.face_bound = face_bound.data();
    t.face_index = face_index.data(); t.face_area = face_area.data();
    t.face_nx = face_nx.data(); t.face_ny = face_ny.data(); t.face_nz = face_nz.data();
    t.face_cx = face_cx.data(); t.face_cy = face_cy.data(); t.face_cz = face_cz.data();
    t.face_nvert = face_nvert.data(); t.face_vert = face_vert.data();
    t.cell_nvert = cell_nvert.data(); t.cell_vert = cell_vert.data();
    t.volume = volume.data(); t.cx = cx.data(); t.cy = cy.data(); t.cz = cz.data();
    t.heat_source = heat_source.data(); t.matter = matter.data();
    t.nbr_ptr = nbr_ptr.data(); t.nbr_idx = nbr_idx.data();
    t.node_x = node_x.data(); t.node_y = node_y.data(); t.node_z = node_z.data();
    t.bound_cell = bound_cell.data(); t.bound_face = bound_face.data(); t.bound_type = bound_type.data();
    t.bound_temp = bound_temp.data(); t.bound_connect = bound_connect.data();
    t.fface_cell = fface_cell.data(); t.fface_face = fface_face.data(); t.fface_bound = fface_bound.data();
    t.kappa_bulk = kappa_bulk.data(); t.capacity_bulk = capacity_bulk.data();
    t.vg = vg.data(); t.tau = tau.data(); t.cap = cap.data(); t.lr = lr.data(); t.hr = hr.data();
    t.dir_x = dir_x.data(); t.dir_y = dir_y.data(); t.dir_z = dir_z.data(); t.weight = weight.data();
    t.initial_temp = initial_temp.data();
    t.max_iter = deck.max_iter; t.tol_temp = deck.tol_temp; t.tol_flux = deck.tol_flux; t.tref = deck.tref;
    t.order = deck.order == "1" ? GBTE_ORDER_1 : deck.order == "2limit" ? GBTE_ORDER_2LIMIT : GBTE_ORDER_2;
    t.scheme = deck.method == "Sequential" ? GBTE_SCHEME_SEQUENTIAL : GBTE_SCHEME_SYNTHETIC;
    t.matrix_solver = deck.matrix == "BICGSTAB" ? GBTE_MATRIX_BICGSTAB : GBTE_MATRIX_LU;
    return t;
}

}  // namespace gbte
