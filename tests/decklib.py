"""Named test decks (seeded) shared by the CPU and GPU tests."""
from __future__ import annotations

import dataclasses
import os

from giftbte_amd import decks

# name -> (spec factory, run settings)
DECKS = {
    # 2D gray film, 40x40 split squares, 32 in-plane directions, specular
    # top/bottom, left 301 K / right 299 K, converged to the deck tolerance.
    "gray40": (lambda: decks.gray_square_spec(40), dict(max_iter=500, tol_temp=1e-8, tol_flux=1e-4)),
    # small porous film: periodic top/bottom, diffuse pore, 4 bands x 16 dirs
    "porous_b4": (lambda: decks.porous_spec(n_target=2000, nbands=4, nphi=4, seed=3),
                  dict(max_iter=25, tol_temp=0.0, tol_flux=0.0)),
    # 40 bands x 64 directions on a coarse porous mesh (one cell per wavefront)
    "porous_b40": (lambda: decks.porous_spec(n_target=400, nbands=40, nphi=16, seed=5),
                   dict(max_iter=4, tol_temp=0.0, tol_flux=0.0)),
    # gray film with a uniform heat source and cold walls on all sides
    "gray_heat": (lambda: decks.gray_square_spec(
        12, uniform_heat=1e17, bcs=[(1, 1, 300.0), (2, 1, 300.0), (3, 1, 300.0), (4, 1, 300.0)]),
                  dict(max_iter=6, tol_temp=0.0, tol_flux=0.0)),
    # 20 bands (three cells per wavefront), diffuse walls on top/bottom
    "square_b20_diffuse": (lambda: decks.DeckSpec(
        mesh=decks.square_split_mesh(10), bands=decks.debye_bands(20, 7),
        bcs=[(1, 2, 0.0), (2, 1, 299.0), (3, 2, 0.0), (4, 1, 301.0)], scale=3e-7, nphi=4),
                           dict(max_iter=5, tol_temp=0.0, tol_flux=0.0)),
    # IterativeScheme Sequential (StaticBTESynthetic::solve_DOM: no synthetic
    # correction): a fixed run and one stopped by the tolerances
    "porous_b4_seq": (lambda: dataclasses.replace(decks.porous_spec(n_target=2000, nbands=4, nphi=4, seed=3),
                                                  method="Sequential"),
                      dict(max_iter=25, tol_temp=0.0, tol_flux=0.0)),
    "gray12_seq": (lambda: dataclasses.replace(decks.gray_square_spec(12), method="Sequential"),
                   dict(max_iter=400, tol_temp=1e-6, tol_flux=1e-4)),
    # Order 1 (solve_firstorder / solve_DOM_firstorder: no gradient terms)
    "porous_b4_o1": (lambda: dataclasses.replace(decks.porous_spec(n_target=2000, nbands=4, nphi=4, seed=3), order="1"),
                     dict(max_iter=25, tol_temp=0.0, tol_flux=0.0)),
    "gray12_o1_seq": (lambda: dataclasses.replace(decks.gray_square_spec(12), method="Sequential", order="1"),
                      dict(max_iter=400, tol_temp=1e-6, tol_flux=1e-4)),
    # Order 2limit (the gradient limiter)
    "porous_b4_2limit": (lambda: dataclasses.replace(decks.porous_spec(n_target=2000, nbands=4, nphi=4, seed=3),
                                                     order="2limit"),
                         dict(max_iter=25, tol_temp=0.0, tol_flux=0.0)),
    "square_b20_2limit_seq": (lambda: dataclasses.replace(DECKS["square_b20_diffuse"][0](), order="2limit",
                                                          method="Sequential"),
                              dict(max_iter=5, tol_temp=0.0, tol_flux=0.0)),
    # MatrixSolver BICGSTAB (solve_Iterative / solve_DOM_Iterative: Eigen
    # BiCGSTAB per system instead of the LU factor)
    "porous_b4_bicg": (lambda: dataclasses.replace(decks.porous_spec(n_target=2000, nbands=4, nphi=4, seed=3),
                                                   matrix="BICGSTAB"),
                       dict(max_iter=25, tol_temp=0.0, tol_flux=0.0)),
    "gray12_bicg_seq": (lambda: dataclasses.replace(decks.gray_square_spec(12), method="Sequential",
                                                    matrix="BICGSTAB"),
                        dict(max_iter=400, tol_temp=1e-6, tol_flux=1e-4)),
    # 3D: tetrahedra of a cube with a cubic void (diffuse walls and void,
    # isothermal x = 0 / x = 1), 8 bands, Gauss 3D directions (32)
    "cube3d_b8": (lambda: decks.cube_void_spec(m=6, nbands=8, ntheta=2, nphi=2, seed=6),
                  dict(max_iter=8, tol_temp=0.0, tol_flux=0.0)),
    # 3D, 20 bands: a band count that 8-band chunks do not divide (the sweep
    # then runs chunks of 16 and 4 bands with per-chunk lanes per position)
    "cube3d_b20": (lambda: decks.cube_void_spec(m=5, nbands=20, ntheta=2, nphi=2, seed=9),
                   dict(max_iter=6, tol_temp=0.0, tol_flux=0.0)),
    # 3D gray solid cube (no void), Sequential scheme, stopped by the tolerances
    "cube3d_gray_seq": (lambda: dataclasses.replace(
        decks.cube_void_spec(m=5, nbands=1, ntheta=1, nphi=3, void=0.0, seed=8), bands=decks.gray_band(),
        method="Sequential"), dict(max_iter=300, tol_temp=1e-6, tol_flux=1e-4)),
}


def make(name: str, root) -> str:
    d = os.path.join(str(root), name)
    if not os.path.exists(os.path.join(d, "input", "CONTROL")):
        spec, _ = DECKS[name]
        decks.write_deck(d, spec())
    return d


def settings(name: str) -> dict:
    return dict(DECKS[name][1])
