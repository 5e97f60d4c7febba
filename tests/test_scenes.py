"""The full-size DEM workloads keep every touching bond.

Bonds come from the initial pair list (addInitialCondition in
bondedParticleBeam.cpp:45-52 and iceCylinderInteraction.cpp:63-70, through
bondedInteraction::add, DEMSolver.h:88-135).  The pair walk only sees the
27 cells around a ball, so a lattice whose spacing equals the cell size can
lose touching neighbours to the rounding of (x - min) / cellSize.  The
builders give every axis half a cell to spare (scenes.half_cell_bounds); the
oracle's own neighbour search must then bond every nearest-neighbour pair.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle
from this_studio_amd import scenes


def _lattice_bonds(nx, ny, nz):
    """Nearest-neighbour pairs of an nx x ny x nz simple-cubic lattice whose
    x = 0 layer is fixed: pairs of two fixed balls are not listed
    (ballNeighborSearchKernel.cu:49)."""
    total = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    fixed_fixed = (ny - 1) * nz + ny * (nz - 1)
    return total - fixed_fixed


def test_half_cell_bounds_cell_size_exceeds_spacing():
    mn, mx = scenes.half_cell_bounds([-0.008, -0.008, -0.5], [0.52, 0.072, 2.04], 0.004)
    g = oracle.grid_set(mn, mx, 0.004)
    for L, n, c in ((mx[0] - mn[0], g.nx, g.cx), (mx[1] - mn[1], g.ny, g.cy), (mx[2] - mn[2], g.nz, g.cz)):
        assert n == int(L / 0.004)
        assert c > 0.004 * (1 + 1e-6)


def test_full_size_cantilever_keeps_every_lattice_bond():
    """128 x 16 x 16 beam (32,768 balls) with its 20,000 clumps above."""
    s = scenes.cantilever_with_clumps()
    o = oracle.OracleDEM(s)
    assert _lattice_bonds(128, 16, 16) == 93472
    a, b = o.get("bond_pointed"), o.get("bond_pointing")
    assert a.size == 93472
    # every bond joins two beam balls one spacing apart
    pos = o.get("pos")
    d = np.linalg.norm(pos[a] - pos[b], axis=1)
    assert np.abs(d - 0.004).max() < 1e-12
    o.close()


def test_full_size_ice_sheet_keeps_every_touching_bond():
    """The coupled workload's two hex layers (50,665 balls): every pair
    within the 1.1 (rA + rB) cut is bonded, as brute force finds them."""
    spec, d = scenes.ice_cylinder_coupled()
    pos = np.concatenate([o["pos"] for k, o in d.objects if k == "balls"])
    pairs = cKDTree(pos).query_pairs(1.1 * 2 * 0.004, output_type="ndarray")
    o = oracle.OracleDEM(d)
    a, b = o.get("bond_pointed"), o.get("bond_pointing")
    assert a.size == len(pairs) == 225861
    got = set(zip(np.minimum(a, b).tolist(), np.maximum(a, b).tolist()))
    assert got == set(map(tuple, np.sort(pairs, axis=1).tolist()))
    o.close()
