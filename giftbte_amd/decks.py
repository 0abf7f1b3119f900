"""Seeded synthetic input decks in the reference's own file formats.

Writes a deck directory laid out like the reference examples
(bin/examples/*/input): ``input/CONTROL``, ``input/GEOMETRY``,
``input/PHONON``, a boundary-condition file, a DATABASE band file and a 2D
COMSOL ``.mphtxt`` mesh.  The formats follow the readers they feed:

* CONTROL / GEOMETRY / PHONON keyword scan: src/BTEMain.cpp:97-315
* band DATABASE rows ``v tau C lattice_ratio heat_ratio``: src/BTEBand/BTEBand.cpp:64-80
* boundary-condition file: src/BTEBoundaryCondition/BTEBoundaryCondition.cpp:9-33
* COMSOL 2D mesh: src/BTEMesh/ReadinMesh.cpp:245-384 (edge entity indices are
  0-based in the file and become 1-based BC indices; domain indices are 1-based)

Mesh coordinates are written in units of the domain edge (0..1) and scaled by
ScaleX/ScaleY, as the reader does (ReadinMesh.cpp:287-294).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import dataclasses

import numpy as np

# edge entity ids used by the generators (BC index = entity + 1)
BOTTOM, RIGHT, TOP, LEFT, PORE = 0, 1, 2, 3, 4


@dataclass
class Mesh2D:
    nodes: np.ndarray          # (N, 2) float64, unit-square coordinates
    tris: np.ndarray           # (M, 3) int
    edges: np.ndarray          # (E, 2) int, boundary edges
    edge_entity: np.ndarray    # (E,) int, 0-based entity id


@dataclass
class Mesh3D:
    nodes: np.ndarray          # (N, 3) float64, unit-cube coordinates
    tets: np.ndarray           # (M, 4) int, positively oriented
    faces: np.ndarray          # (F, 3) int, boundary triangles
    face_entity: np.ndarray    # (F,) int, 0-based entity id (BC index = entity + 1)


# boundary entity ids of the tetrahedral cube generator (BC index = entity + 1)
X0, X1, Y0, Y1, Z0, Z1, VOID = 0, 1, 2, 3, 4, 5, 6


@dataclass
class Bands:
    v: np.ndarray
    tau: np.ndarray
    C: np.ndarray
    lr: np.ndarray
    hr: np.ndarray


@dataclass
class DeckSpec:
    mesh: "Mesh2D | Mesh3D"
    bands: Bands
    bcs: list                      # [(index, type, value), ...]
    scale: float = 1e-7
    material_dim: int = 2
    angle_method: int = 1
    ntheta: int = 1
    nphi: int = 8
    order: str = "2"
    method: str = "Synthetic"
    matrix: str = "LU"
    max_iter: int = 10000
    tol_temp: float = 1e-8
    tol_flux: float = 1e-4
    tref: float = 300.0
    uniform_heat: float = 0.0
    extra_geometry: dict = field(default_factory=dict)
    # (kind, writer): band data binned by the reader instead of the DATABASE
    # rows; kind "ShengBTE" or "ALAMODE", writer(input_dir) writes the files
    # and returns the path named in PHONON (relative to the deck root)
    band_source: tuple | None = None


# ----------------------------------------------------------------- meshes

def square_split_mesh(n: int) -> Mesh2D:
    """n x n squares on the unit square, each cut along the same diagonal."""
    xs = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)

    def nid(i, j):
        return i * (n + 1) + j

    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            tris.append((a, b, c))
            tris.append((a, c, d))
    edges, ent = [], []
    for k in range(n):
        edges.append((nid(k, 0), nid(k + 1, 0))); ent.append(BOTTOM)
        edges.append((nid(n, k), nid(n, k + 1))); ent.append(RIGHT)
        edges.append((nid(k + 1, n), nid(k, n))); ent.append(TOP)
        edges.append((nid(0, k + 1), nid(0, k))); ent.append(LEFT)
    return Mesh2D(nodes, np.asarray(tris, np.int64), np.asarray(edges, np.int64), np.asarray(ent, np.int64))


def square_quad_mesh(n: int) -> Mesh2D:
    """n x n quadrilaterals on the unit square, COMSOL vertex order
    (v0, v1, v2, v3) = (i,j), (i+1,j), (i,j+1), (i+1,j+1)."""
    xs = np.linspace(0.0, 1.0, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    nodes = np.stack([X.ravel(), Y.ravel()], axis=1)

    def nid(i, j):
        return i * (n + 1) + j

    quads = [(nid(i, j), nid(i + 1, j), nid(i, j + 1), nid(i + 1, j + 1)) for i in range(n) for j in range(n)]
    edges, ent = [], []
    for k in range(n):
        edges.append((nid(k, 0), nid(k + 1, 0))); ent.append(BOTTOM)
        edges.append((nid(n, k), nid(n, k + 1))); ent.append(RIGHT)
        edges.append((nid(k + 1, n), nid(k, n))); ent.append(TOP)
        edges.append((nid(0, k + 1), nid(0, k))); ent.append(LEFT)
    return Mesh2D(nodes, np.asarray(quads, np.int64), np.asarray(edges, np.int64), np.asarray(ent, np.int64))


def porous_delaunay_mesh(n_target: int, pore_radius: float, seed: int) -> Mesh2D:
    """Seeded Delaunay mesh of the unit square minus a centred circular pore.

    All points (square boundary, pore circle, seeded interior points) are
    triangulated together and triangles whose centroid lies in the pore are
    dropped, which keeps the triangulation Delaunay and therefore every
    direction's upwind graph acyclic.  Top and bottom edges carry the same x
    positions so that they can be paired periodically.
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    r = pore_radius
    area = 1.0 - math.pi * r * r
    # a Delaunay triangulation of N points has about 2N triangles
    n_pts = max(int(n_target / 2), 16)
    h = math.sqrt(area / n_pts)
    nb = max(int(round(1.0 / h)), 4)
    s = np.linspace(0.0, 1.0, nb + 1)
    pts = [np.stack([s, np.zeros_like(s)], 1),
           np.stack([s, np.ones_like(s)], 1),
           np.stack([np.zeros(nb - 1), s[1:-1]], 1),
           np.stack([np.ones(nb - 1), s[1:-1]], 1)]
    n_c = 0
    if r > 0:
        n_c = max(int(round(2 * math.pi * r / h)), 8)
        th = 2 * math.pi * np.arange(n_c) / n_c
        pts.append(np.stack([0.5 + r * np.cos(th), 0.5 + r * np.sin(th)], 1))
    # seeded interior points, rejected near walls and the pore
    need = n_pts - sum(len(p) for p in pts)
    inner = []
    margin = 0.45 * h
    while need > 0:
        cand = rng.random((2 * need + 16, 2))
        ok = (cand[:, 0] > margin) & (cand[:, 0] < 1 - margin) & (cand[:, 1] > margin) & (cand[:, 1] < 1 - margin)
        if r > 0:
            ok &= np.hypot(cand[:, 0] - 0.5, cand[:, 1] - 0.5) > r + margin
        cand = cand[ok][:need]
        inner.append(cand)
        need -= len(cand)
    pts.append(np.concatenate(inner, 0) if inner else np.zeros((0, 2)))
    P = np.concatenate(pts, 0)
    tri = Delaunay(P).simplices.astype(np.int64)
    cen = P[tri].mean(axis=1)
    if r > 0:
        tri = tri[np.hypot(cen[:, 0] - 0.5, cen[:, 1] - 0.5) > r]
    # drop unused points, renumber
    used = np.unique(tri)
    remap = -np.ones(len(P), np.int64)
    remap[used] = np.arange(len(used))
    P = P[used]
    tri = remap[tri]
    # boundary edges = edges used by exactly one triangle
    e = np.concatenate([tri[:, [1, 2]], tri[:, [2, 0]], tri[:, [0, 1]]], 0)
    es = np.sort(e, 1)
    _, inv, cnt = np.unique(es, axis=0, return_inverse=True, return_counts=True)
    bmask = cnt[inv.ravel()] == 1
    bedges = e[bmask]
    mid = P[bedges].mean(axis=1)
    tol = 1e-9
    ent = np.full(len(bedges), -1, np.int64)
    ent[np.abs(mid[:, 1]) < tol] = BOTTOM
    ent[np.abs(mid[:, 0] - 1) < tol] = RIGHT
    ent[np.abs(mid[:, 1] - 1) < tol] = TOP
    ent[np.abs(mid[:, 0]) < tol] = LEFT
    ent[(ent < 0) & (np.hypot(mid[:, 0] - 0.5, mid[:, 1] - 0.5) < r + h)] = PORE
    if (ent < 0).any():
        raise RuntimeError("porous_delaunay_mesh: unclassified boundary edge")
    return Mesh2D(P, tri, bedges, ent)


def cube_void_tet_mesh(m: int, void: float = 0.4, seed: int = 1, jitter: float = 0.3) -> Mesh3D:
    """Seeded Delaunay tetrahedra of the unit cube minus a centred cubic void.

    The points are the nodes of an (m+1)^3 lattice outside the void (the void
    faces sit on the lattice planes nearest to (1 -+ void) / 2), each moved by
    a seeded jitter of up to ``jitter`` lattice spacings within the planes it
    lies on: interior nodes in 3D, face nodes within their face, edge nodes
    along their edge, corners not at all.  All points are tetrahedralised
    together (scipy's Qhull) and tetrahedra whose centroid lies in the void
    are dropped.  A Delaunay tetrahedralisation has an acyclic upwind graph
    for every direction, and dropping cells keeps it so.  Boundary triangles
    are classified by the cube face they lie on; the rest face the void.
    """
    from scipy.spatial import Delaunay

    rng = np.random.default_rng(seed)
    h = 1.0 / m
    lo = int(round((1.0 - void) / 2 * m)) if void > 0 else m
    hi = m - lo if void > 0 else m
    g = np.arange(m + 1)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    ijk = np.stack([I.ravel(), J.ravel(), K.ravel()], 1)
    inside = np.all((ijk > lo) & (ijk < hi), axis=1) if hi > lo else np.zeros(len(ijk), bool)
    ijk = ijk[~inside]
    # a coordinate may move when it is not on an outer or a void plane
    on_outer = (ijk == 0) | (ijk == m)
    in_void_box = np.all((ijk >= lo) & (ijk <= hi), axis=1) if hi > lo else np.zeros(len(ijk), bool)
    on_void = in_void_box[:, None] & ((ijk == lo) | (ijk == hi))
    free = ~(on_outer | on_void)
    P = ijk * h + free * rng.uniform(-jitter, jitter, ijk.shape) * h
    tet = Delaunay(P).simplices.astype(np.int64)
    cen = P[tet].mean(axis=1)
    a, b = lo * h, hi * h
    if hi > lo:
        tet = tet[~np.all((cen > a) & (cen < b), axis=1)]
    used = np.unique(tet)
    remap = -np.ones(len(P), np.int64)
    remap[used] = np.arange(len(used))
    P = P[used]
    tet = remap[tet]
    # positive orientation
    d = np.einsum("ij,ij->i", np.cross(P[tet[:, 1]] - P[tet[:, 0]], P[tet[:, 2]] - P[tet[:, 0]]),
                  P[tet[:, 3]] - P[tet[:, 0]])
    neg = d < 0
    tet[neg] = tet[neg][:, [0, 2, 1, 3]]
    # Qhull can close a planar patch of the hull (four cocircular points on a
    # cube face) with a flat tetrahedron; dropping it leaves that patch's
    # triangles on the face, which the classification below checks
    tet = tet[np.abs(d) > 1e-9 * h ** 3]
    used = np.unique(tet)
    remap = -np.ones(len(P), np.int64)
    remap[used] = np.arange(len(used))
    P = P[used]
    tet = remap[tet]
    # boundary triangles = faces of exactly one tetrahedron
    f = np.concatenate([tet[:, [1, 2, 3]], tet[:, [0, 3, 2]], tet[:, [0, 1, 3]], tet[:, [0, 2, 1]]], 0)
    fs = np.sort(f, 1)
    _, inv, cnt = np.unique(fs, axis=0, return_inverse=True, return_counts=True)
    bf = f[cnt[inv.ravel()] == 1]
    X = P[bf]  # (F, 3 vertices, 3)
    tol = 1e-12
    ent = np.full(len(bf), -1, np.int64)
    for e, (ax, val) in enumerate([(0, 0.0), (0, 1.0), (1, 0.0), (1, 1.0), (2, 0.0), (2, 1.0)]):
        ent[np.all(np.abs(X[:, :, ax] - val) < tol, axis=1)] = e
    if hi > lo:
        in_box = np.all((X > a - tol) & (X < b + tol), axis=(1, 2))
        ent[(ent < 0) & in_box] = VOID
    if (ent < 0).any():
        raise RuntimeError("cube_void_tet_mesh: a boundary triangle lies on neither the cube nor the void")
    return Mesh3D(P, tet, bf, ent)


# ------------------------------------------------------------------ bands

def gray_band(v=6400.0, tau=15.6e-12, C=1.66e6) -> Bands:
    one = np.ones(1)
    return Bands(v * one, tau * one, C * one, one.copy(), one.copy())


def debye_bands(nb: int, seed: int, omega_max=1e14, vmax=6400.0, mfp_min=1e-9, mfp_max=1e-5,
                T=300.0, c_total=1.66e6, sigma=0.1) -> Bands:
    """Isotropic Debye-like dispersion on nb bands.

    omega_b = (b - 1/2) omega_max / nb, v_b = vmax cos(pi omega_b / (2 omega_max));
    mean free paths log-uniform from mfp_max (lowest band) down to mfp_min
    (highest band), tau_b = MFP_b / v_b times a seeded LogNormal(0, sigma)
    factor; C_b follows the Debye weight omega^2 x^2 e^x / (e^x - 1)^2 at T and
    sums to c_total.  lattice_ratio_b = (C/tau)_b / sum(C/tau) and
    heat_ratio_b = C_b / sum(C).
    """
    rng = np.random.default_rng(seed)
    b = np.arange(nb) + 0.5
    w = b * omega_max / nb
    v = vmax * np.cos(np.pi * w / (2 * omega_max))
    mfp = mfp_max * (mfp_min / mfp_max) ** (b / nb)
    tau = mfp / v * rng.lognormal(0.0, sigma, nb)
    hbar, kb = 1.054571817e-34, 1.380649e-23
    x = hbar * w / (kb * T)
    wt = w ** 2 * x ** 2 * np.exp(x) / np.expm1(x) ** 2
    C = c_total * wt / wt.sum()
    lr = (C / tau) / (C / tau).sum()
    hr = C / C.sum()
    return Bands(v, tau, C, lr, hr)


def scale_tau(bands: Bands, s) -> Bands:
    """Relaxation times tau_b -> s_b tau_b with the lattice ratios following,
    lr = (C/tau) / sum(C/tau) (the relation debye_bands uses); group
    velocities, capacities and heat ratios unchanged."""
    tau = bands.tau * np.asarray(s, dtype=np.float64)
    lr = (bands.C / tau) / (bands.C / tau).sum()
    return Bands(bands.v.copy(), tau, bands.C.copy(), lr, bands.hr.copy())


def set_wall_temperature(spec: DeckSpec, index: int, value: float) -> DeckSpec:
    """The spec with boundary condition `index` (type 1) at temperature value."""
    bcs = [(i, t, value if i == index else v) for (i, t, v) in spec.bcs]
    return dataclasses.replace(spec, bcs=bcs)


# ------------------------------------------------------------------ writers

def _f(x) -> str:
    """Shortest round-trip decimal text of a float."""
    return repr(float(x))


def write_mphtxt(path: str, m: Mesh2D) -> None:
    """COMSOL text mesh, 2D: vtx/edg/tri element types."""
    out = []
    out += ["# Created by COMSOL Multiphysics.", "", "# Major & minor version", "0 1 ",
            "1 # number of tags", "# Tags", "5 mesh1 ", "1 # number of types", "# Types", "3 obj ", "",
            "# --------- Object 0 ----------", "", "0 0 1 ", "4 Mesh # class", "4 # version",
            "2 # sdim", f"{len(m.nodes)} # number of mesh vertices", "0 # lowest mesh vertex index", "",
            "# Mesh vertex coordinates"]
    out += [f"{_f(x)} {_f(y)} " for x, y in m.nodes.tolist()]
    out += ["", "2 # number of element types", "", "# Type #0", "", "3 edg # type name", "", "",
            "2 # number of vertices per element", f"{len(m.edges)} # number of elements", "# Elements"]
    out += [f"{a} {b} " for a, b in m.edges.tolist()]
    out += ["", f"{len(m.edges)} # number of geometric entity indices", "# Geometric entity indices"]
    out += [f"{e} " for e in m.edge_entity.tolist()]
    nv = m.tris.shape[1]
    tname = "tri" if nv == 3 else "quad"
    out += ["", "# Type #1", "", f"{len(tname)} {tname} # type name", "", "",
            f"{nv} # number of vertices per element", f"{len(m.tris)} # number of elements", "# Elements"]
    out += [" ".join(str(v) for v in row) + " " for row in m.tris.tolist()]
    out += ["", f"{len(m.tris)} # number of geometric entity indices", "# Geometric entity indices"]
    out += ["1 "] * len(m.tris)
    out.append("")
    with open(path, "w") as f:
        f.write("\n".join(out))


def write_msh(path: str, m: Mesh2D) -> None:
    """Gmsh 4.1 text mesh, 2D, in the subset src/BTEMesh/ReadinMesh.cpp:387-532
    reads: one node block, one line block per boundary entity (entity tag =
    BC index), one triangle block (entity tag 1 -> region 0)."""
    n = len(m.nodes)
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$Nodes", f"1 {n} 1 {n}", f"2 1 0 {n}"]
    lines += [str(i + 1) for i in range(n)]
    lines += [f"{_f(x)} {_f(y)} 0" for x, y in m.nodes.tolist()]
    lines.append("$EndNodes")
    ents = sorted(set(m.edge_entity.tolist()))
    nel = len(m.edges) + len(m.tris)
    lines += ["$Elements", f"{len(ents) + 1} {nel} 1 {nel}"]
    tag = 1
    for e in ents:
        sel = m.edges[m.edge_entity == e]
        lines.append(f"1 {e + 1} 1 {len(sel)}")
        for a, b in sel.tolist():
            lines.append(f"{tag} {a + 1} {b + 1}")
            tag += 1
    if m.tris.shape[1] != 3:
        raise ValueError("write_msh writes triangle meshes")
    lines.append(f"2 1 2 {len(m.tris)}")
    for a, b, c in m.tris.tolist():
        lines.append(f"{tag} {a + 1} {b + 1} {c + 1}")
        tag += 1
    lines.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_msh3d(path: str, m: Mesh3D) -> None:
    """Gmsh 4.1 text mesh, 3D, in the same subset as the 2D writer: one node
    block, one triangle block per boundary entity (entity tag = BC index),
    one tetrahedron block (entity tag 1 -> region 0)."""
    n = len(m.nodes)
    out = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$Nodes", f"1 {n} 1 {n}", f"3 1 0 {n}"]
    out.append("\n".join(str(i + 1) for i in range(n)))
    out.append("\n".join(f"{_f(x)} {_f(y)} {_f(z)}" for x, y, z in m.nodes.tolist()))
    out.append("$EndNodes")
    ents = sorted(set(m.face_entity.tolist()))
    nel = len(m.faces) + len(m.tets)
    out += ["$Elements", f"{len(ents) + 1} {nel} 1 {nel}"]
    tag = 1
    for e in ents:
        sel = m.faces[m.face_entity == e] + 1
        out.append(f"2 {e + 1} 2 {len(sel)}")
        out.append("\n".join(f"{tag + i} {a} {b} {c}" for i, (a, b, c) in enumerate(sel.tolist())))
        tag += len(sel)
    t = m.tets + 1
    out.append(f"3 1 4 {len(t)}")
    out.append("\n".join(f"{tag + i} {a} {b} {c} {d}" for i, (a, b, c, d) in enumerate(t.tolist())))
    out.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_deck(root: str, spec: DeckSpec, mesh_format: str = "COMSOL") -> str:
    """Write a complete deck under ``root``; returns ``root``.  A Mesh3D is
    written as a Gmsh 4.1 3D mesh (GeometryDimension 3)."""
    inp = os.path.join(root, "input")
    os.makedirs(inp, exist_ok=True)
    is3d = isinstance(spec.mesh, Mesh3D)
    if is3d:
        mesh_format = "MSH"
        write_msh3d(os.path.join(inp, "mesh.msh"), spec.mesh)
    elif mesh_format == "MSH":
        write_msh(os.path.join(inp, "mesh.msh"), spec.mesh)
    else:
        write_mphtxt(os.path.join(inp, "mesh.mphtxt"), spec.mesh)
    bd = spec.bands
    with open(os.path.join(inp, "bands.dat"), "w") as f:
        for k in range(len(bd.v)):
            f.write(" ".join(_f(a[k]) for a in (bd.v, bd.tau, bd.C, bd.lr, bd.hr)) + "\n")
    with open(os.path.join(inp, "bc.dat"), "w") as f:
        f.write(f"{len(spec.bcs)} #Number of boundaries\n\n")
        f.write("BoundaryIndex BoundaryType BoundaryValue\n")
        for idx, typ, val in spec.bcs:
            f.write(f"{int(idx)} {int(typ)} {_f(val)}\n")
    with open(os.path.join(inp, "CONTROL"), "w") as f:
        f.write(f"State=\nSteady\nOrder=\n{spec.order}\nIterativeScheme=\n{spec.method}\n"
                f"MatrixSolver=\n{spec.matrix}\nMaxNumIter=\n{spec.max_iter}\n"
                f"ResidualTemp=\n{_f(spec.tol_temp)}\nResidualFlux=\n{_f(spec.tol_flux)}\n"
                f"Tref=\n{_f(spec.tref)}\n")
    with open(os.path.join(inp, "GEOMETRY"), "w") as f:
        f.write(f"GeometryDimension=\n{3 if is3d else 2}\nScaleX=\n{_f(spec.scale)}\nScaleY=\n{_f(spec.scale)}\n"
                + (f"ScaleZ=\n{_f(spec.scale)}\n" if is3d else "") +
                f"BCFile=\ninput/bc.dat\nMeshType=\n{mesh_format}\nMeshFile=\n"
                f"input/mesh.{'msh' if mesh_format == 'MSH' else 'mphtxt'}\n")
        if spec.uniform_heat:
            f.write(f"UniformHeat=\n{_f(spec.uniform_heat)}\n")
    material = "DATABASE\ninput/bands.dat"
    if spec.band_source is not None:
        kind, writer = spec.band_source
        material = f"{kind}\n{writer(inp)}"
    with open(os.path.join(inp, "PHONON"), "w") as f:
        f.write(f"MaterialNumber=\n1\nBandNumber=\n{len(bd.v)}\n\n"
                f"MaterialFile 1= {material}\n\n\n"
                f"GeometryMaterialType=\n1 1\n\n"
                f"MaterialDimension=\n{spec.material_dim}\nAngle_method=\n{spec.angle_method}\n"
                f"Ntheta=\n{spec.ntheta}\nNphi=\n{spec.nphi}\n")
    return root


# ------------------------------------------------------- named workloads

def gray_square_spec(n=40, **kw) -> DeckSpec:
    """2D gray film: n x n split squares, left 301 K / right 299 K, top and
    bottom specular, 1 band (v 6400 m/s, tau 15.6 ps, C 1.66e6), CADOM
    in-plane directions (4 * nphi)."""
    bcs = [(BOTTOM + 1, 3, 0.0), (RIGHT + 1, 1, 299.0), (TOP + 1, 3, 0.0), (LEFT + 1, 1, 301.0)]
    base = dict(mesh=square_split_mesh(n), bands=gray_band(), bcs=bcs, scale=1e-7,
                material_dim=2, angle_method=1, nphi=8, tol_temp=1e-8)
    base.update(kw)
    return DeckSpec(**base)


def porous_spec(n_target=50000, nbands=40, nphi=16, pore_radius=0.3, seed=1, scale=5e-7, **kw) -> DeckSpec:
    """2D nanoporous film: unit square (scaled) minus a centred circular pore
    (diffuse), left/right isothermal at 301/299 K, top/bottom periodic,
    Debye-like bands, CADOM in-plane directions (4 * nphi)."""
    bcs = [(BOTTOM + 1, -(TOP + 1), 0.0), (RIGHT + 1, 1, 299.0), (TOP + 1, -(BOTTOM + 1), 0.0),
           (LEFT + 1, 1, 301.0), (PORE + 1, 2, 0.0)]
    base = dict(mesh=porous_delaunay_mesh(n_target, pore_radius, seed), bands=debye_bands(nbands, seed),
                bcs=bcs, scale=scale, material_dim=2, angle_method=1, nphi=nphi, tol_temp=1e-8)
    base.update(kw)
    return DeckSpec(**base)


def cube_void_spec(m=38, nbands=40, ntheta=2, nphi=6, void=0.4, seed=1, scale=2e-7, **kw) -> DeckSpec:
    """3D film: unit cube (scaled; 200 nm by default) minus a centred cubic
    void (0.4 of the edge), seeded Delaunay tetrahedra from an (m+1)^3
    jittered lattice; x = 0 at 301 K and x = 1 at 299 K (isothermal), the
    other four faces and the void diffuse; Debye-like bands; Gauss-Legendre
    3D directions (8 * ntheta * nphi)."""
    bcs = [(X0 + 1, 1, 301.0), (X1 + 1, 1, 299.0), (Y0 + 1, 2, 0.0), (Y1 + 1, 2, 0.0), (Z0 + 1, 2, 0.0),
           (Z1 + 1, 2, 0.0)]
    if void > 0:
        bcs.append((VOID + 1, 2, 0.0))
    base = dict(mesh=cube_void_tet_mesh(m, void, seed), bands=debye_bands(nbands, seed), bcs=bcs, scale=scale,
                material_dim=3, angle_method=2, ntheta=ntheta, nphi=nphi, tol_temp=1e-8)
    base.update(kw)
    return DeckSpec(**base)


def _fortran_row(vals) -> str:
    """One data row as ShengBTE writes them: right-aligned E-format fields,
    no trailing blanks (the reader's token scan needs that)."""
    return "".join(("%20.10E" % v) if np.isfinite(v) else "%20s" % "NaN" for v in vals)


def shengbte_writer(seed: int, nirr: int = 6, nbranch: int = 6, nfull: int = 16):
    """Seeded ShengBTE-format mode data (CONTROL, BTE.omega,
    BTE.qpoints_full, BTE.w_final, BTE.v) for the band binning of
    src/BTEBand/BTEBand.cpp:81-628: nirr irreducible q-points with nbranch
    branches, expanded to nfull q-points.  The first q-point is Gamma (zero
    acoustic frequencies and NaN scattering rates, as ShengBTE writes them)."""
    def write(inp):
        rng = np.random.default_rng(seed)
        d = os.path.join(inp, "shengbte")
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, "CONTROL"), "w") as f:
            f.write("&allocations\n        nelements=1,\n        natoms=2,\n        ngrid(:)=4 2 2\n&end\n"
                    "&crystal\n        lfactor=0.0529177210818,\n"
                    "        lattvec(:,1)= -5.372296320673052 0.000000000000000 5.372296320673052\n"
                    "        lattvec(:,2)=  0.000000000000000 5.372296320673052 5.372296320673052\n"
                    "        lattvec(:,3)= -5.372296320673052 5.372296320673052 0.000000000000000\n"
                    "        elements=\"Ge\"\n&end\n&parameters\n        T=300\n        scalebroad=1\n&end\n"
                    "&flags\n        convergence = .true.\n        isotopes=.true.\n&end\n")
        om = np.sort(rng.uniform(2.0, 55.0, (nirr, nbranch)), axis=1)
        om[0, :3] = 0.0
        rate = rng.uniform(1e-3, 2.0, (nbranch, nirr))
        vel = rng.uniform(-5.0, 5.0, (nbranch, nirr, 3))
        with open(os.path.join(d, "BTE.omega"), "w") as f:
            for q in range(nirr):
                f.write(_fortran_row(om[q]) + "\n")
        irr = np.concatenate([np.arange(1, nirr + 1), rng.integers(1, nirr + 1, nfull - nirr)])
        with open(os.path.join(d, "BTE.qpoints_full"), "w") as f:
            for i, q in enumerate(irr):
                f.write("%9d %9d" % (i + 1, q) + _fortran_row(rng.uniform(0, 0.5, 3)) + "\n")
        with open(os.path.join(d, "BTE.w_final"), "w") as f:
            for b in range(nbranch):
                for q in range(nirr):
                    f.write(_fortran_row([om[q, b], np.nan if om[q, b] == 0 else rate[b, q]]) + "\n")
        with open(os.path.join(d, "BTE.v"), "w") as f:
            for b in range(nbranch):
                for q in range(nirr):
                    f.write(_fortran_row(vel[b, q] * (0.0 if om[q, b] == 0 else 1.0)) + "\n")
        return "input/shengbte"
    return write


def alamode_writer(seed: int, nk: int = 12, nmode: int = 6):
    """Seeded ALAMODE lifetime file (the analyzer's "Phonon lifetime" table:
    ik, is, frequency, lifetime [ps], |v|, MFP, multiplicity and the 9
    conductivity tensor components) for src/BTEBand/BTEBand.cpp:629-740.
    The first k-point has zero lifetimes and velocities (Gamma)."""
    def write(inp):
        rng = np.random.default_rng(seed)
        path = os.path.join(inp, "tau_alamode.dat")
        with open(path, "w") as f:
            f.write("# Result analyzer ver. 1.0.5\n# Input file : synthetic.result\n"
                    "# Phonon lifetime at temperature 300 K.\n"
                    f"# kpoint range 1 {nk}\n# mode   range 1 {nmode}\n"
                    "#  ik,  is, Frequency [cm^{-1}], Lifetime [ps], |Velocity| [m/s], MFP [nm], Multiplicity, "
                    "Thermal conductivity par mode (xx, xy, ...) [W/mK]\n")
            for k in range(nk):
                for m in range(nmode):
                    freq = rng.uniform(5.0, 500.0)
                    tau = 0.0 if k == 0 else rng.uniform(1.0, 5000.0)
                    v = 0.0 if k == 0 else rng.uniform(100.0, 9000.0)
                    mult = int(rng.integers(1, 9))
                    kap = rng.uniform(0, 0.3, 9) if k else np.zeros(9)
                    f.write("%5d%5d" % (k + 1, m + 1) + "".join("%15.6g" % x for x in (freq, tau, v, tau * v * 1e-3))
                            + "%5d" % mult + "".join("%15.6g" % x for x in kap) + "\n")
        return "input/tau_alamode.dat"
    return write


def read_band_table(path: str) -> Bands:
    """DATABASE band rows: v tau C lattice_ratio heat_ratio."""
    a = np.loadtxt(path, ndmin=2)
    return Bands(a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy(), a[:, 3].copy(), a[:, 4].copy())


def crossplane_spec(band_file: str, length: float = 1e-7, n: int = 100, **kw) -> DeckSpec:
    """Cross-plane film of thickness ``length`` (the reference's
    bin/examples/cross-plane decks): n x n quads, bottom 301 K and top 300 K
    isothermal, left/right specular, 3D Gauss directions (Ntheta 4, Nphi 4)."""
    bcs = [(BOTTOM + 1, 1, 301.0), (RIGHT + 1, 3, 0.0), (TOP + 1, 1, 300.0), (LEFT + 1, 3, 0.0)]
    base = dict(mesh=square_quad_mesh(n), bands=read_band_table(band_file), bcs=bcs, scale=length,
                material_dim=3, angle_method=2, ntheta=4, nphi=4, tol_temp=1e-8, tol_flux=1e-2, max_iter=100000)
    base.update(kw)
    return DeckSpec(**base)
