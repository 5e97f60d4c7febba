"""Periodic boundaries in the FP64 oracle (setPeriodicBoundary,
DEMSolver.h:1944-1959), pinned without the reference:

* the pair list against a brute-force minimum-image search;
* translation invariance: a y-periodic sheet and the same sheet rolled by
  some rows evolve identically (up to summation-order rounding);
* a bonded sheet translating rigidly through the periodic face feels no
  force and keeps every bond;
* what the reference's ghost passes would list on the same sheet (double
  counted contacts, missing cross-face contacts), the reason the intended
  rule is used.
"""
import itertools
import math

import numpy as np
import pytest

import oracle
from this_studio_amd import scenes


def _box_scene(n=400, seed=3, per=(True, True, True)):
    rng = np.random.default_rng(seed)
    s = scenes.DEMScene("periodic_box")
    box = 0.2
    pos = rng.uniform(0.0, box, (n, 3))
    # a few balls just outside the faces: wrapped at the first search
    pos[:8, 0] = box + rng.uniform(0.0, 0.004, 8)
    pos[8:16, 1] = -rng.uniform(1e-6, 0.004, 8)
    rad = rng.uniform(0.006, 0.01, n)
    s.add_balls(pos, np.zeros(3), np.zeros(3), rad, 1000.0, 0)
    s.set_linear_pair(0, 0, 1e3, 400.0, 0, 0, 0.1, 0.1, 0, 0, 0.3, 0, 0)
    s.set_periodic_boundary(*per)
    s.min_bound = np.zeros(3)
    s.max_bound = box * np.ones(3)
    s.gravity = np.zeros(3)
    s.dt = 1e-5
    return s


@pytest.mark.parametrize("per", [(True, True, True), (False, True, False), (True, False, True)])
def test_periodic_pairs_match_minimum_image_brute_force(per):
    s = _box_scene(per=per)
    o = oracle.OracleDEM(s, solve_begin=False)
    o.neighbor_search()
    pos = o.get("pos")
    rad = o.get("radius")
    L = s.max_bound - s.min_bound
    assert (pos[:, [a for a in range(3) if per[a]]] >= 0).all() and (pos[:, [a for a in range(3) if per[a]]] < 0.2).all()
    got = list(zip(o.get("pointed").tolist(), o.get("pointing").tolist(), o.get("pairShift").tolist()))
    assert len({(a, b) for a, b, _ in got}) == len(got)  # each pair once
    codes = [c for c in itertools.product((-1, 0, 1), repeat=3) if all(per[a] or c[a] == 0 for a in range(3))]
    within, touching = set(), set()
    n = pos.shape[0]
    for a in range(n):
        for b in range(a + 1, n):
            for c in codes:
                d = pos[a] - (pos[b] + np.array(c) * L)
                r2 = d @ d
                code = (c[0] + 1) + 3 * (c[1] + 1) + 9 * (c[2] + 1)
                if r2 <= (1.1 * (rad[a] + rad[b])) ** 2:
                    within.add((a, b, code))
                if r2 < (rad[a] + rad[b]) ** 2:
                    touching.add((a, b, code))
    assert set(got) <= within
    assert touching <= set(got)
    assert any(c != 13 for _, _, c in got)  # contacts through the faces exist


def _wrap_diff(a, b, L):
    d = a - b
    return d - L * np.round(d / L)


def test_periodic_sheet_is_translation_invariant():
    """The sheet and the sheet rolled by 6 rows: positions differ by the roll
    (modulo the period) and the bond sets are equal, after 400 steps of
    shear and compression waves that break bonds.  The roll puts a row just
    below the maximum face, where (y - min) / cellSize rounds to the cell
    past the last one (wrapped to the image below the minimum face).  The
    lattice has a 0.1 % gap: at zero overlap the rounding of the overlap
    decides whether a damped contact acts, which a touching lattice follows
    chaotically."""
    a = scenes.periodic_ice_sheet(gap=1e-3)
    b = scenes.periodic_ice_sheet(roll_rows=6, gap=1e-3)
    oa, ob = oracle.OracleDEM(a), oracle.OracleDEM(b)
    bonds_a = set(zip(oa.get("bond_pointed").tolist(), oa.get("bond_pointing").tolist()))
    bonds_b = set(zip(ob.get("bond_pointed").tolist(), ob.get("bond_pointing").tolist()))
    assert bonds_a == bonds_b
    ly = a.max_bound[1] - a.min_bound[1]
    # every ball of the periodic hex sheet has six bonded neighbours
    deg = np.bincount(oa.get("bond_pointed"), minlength=480) + np.bincount(oa.get("bond_pointing"), minlength=480)
    interior = (a.objects[0][1]["pos"][:, 0] > 0.15) & (a.objects[0][1]["pos"][:, 0] < 4.6)
    assert (deg[interior] == 6).all()
    oa.step(400)
    ob.step(400)
    dy = 0.2 * 1.001 * math.sqrt(3.0) * 0.5
    pa, pb = oa.get("pos"), ob.get("pos")
    err = np.abs(np.concatenate([pb[:, [0, 2]] - pa[:, [0, 2]],
                                 _wrap_diff(pb[:, [1]], pa[:, [1]] + 6 * dy, ly)], axis=1))
    assert err.max() < 1e-12, err.max()
    def broken(o):  # bond order follows the pair list, which differs
        return {(a, b) for a, b, k in zip(o.get("bond_pointed"), o.get("bond_pointing"), o.get("bond_isBonded"))
                if k == 0}
    assert broken(oa) == broken(ob)
    assert len(broken(oa)) > 10  # the waves break bonds
    assert np.abs(ob.get("vel") - oa.get("vel")).max() < 1e-10


def test_rigid_translation_through_the_face():
    """Every ball moves at 50 m/s along y: the sheet crosses the periodic
    face (positions are wrapped into the cell), no bond breaks and no force
    acts beyond rounding."""
    s = scenes.periodic_ice_sheet(uniform_velocity=[0.0, 50.0, 0.0])
    o = oracle.OracleDEM(s)
    p0 = o.get("pos").copy()
    nb = o.get("bond_isBonded").size
    o.step(1500)
    ly = s.max_bound[1] - s.min_bound[1]
    p = o.get("pos")
    assert (p[:, 1] >= 0).all() and (p[:, 1] < ly + 5 * 50.0 * s.dt).all()
    moved = 1500 * s.dt * 50.0
    assert moved > 0.25  # more than one cell through the face
    d = _wrap_diff(p[:, 1], p0[:, 1] + moved, ly)
    assert np.abs(d).max() < 1e-9
    assert np.abs(p[:, [0, 2]] - p0[:, [0, 2]]).max() < 1e-9
    assert o.get("bond_isBonded").sum() == nb
    m = 4 / 3 * math.pi * 0.1 ** 3 * 900.0
    assert np.abs(o.get("force")).max() < 1e-6 * m * 9.81
    assert np.abs(o.get("vel") - [0.0, 50.0, 0.0]).max() < 1e-9


def test_reference_ghost_passes_double_count_and_miss():
    """What the reference's ghost pass would list on the y-periodic sheet:
    contacts between the balls of the first cell layer appear again (both
    are shifted), and the contacts across the face are missed whenever the
    ball at the maximum face has the larger index, which holds for every
    such contact of this sheet (rows are inserted in y order)."""
    s = scenes.periodic_ice_sheet()
    o = oracle.OracleDEM(s, solve_begin=False)
    o.neighbor_search()
    g, _ = o.grids()
    pos = o.get("pos")
    n = pos.shape[0]
    ref = oracle.reference_periodic_pairs(pos, o.get("radius"), o.get("invMass"), np.full(n, -1), g, (0, 1, 0))
    pairs = list(zip(o.get("pointed").tolist(), o.get("pointing").tolist(), o.get("pairShift").tolist()))
    real = {(a, b) for a, b, c in pairs if c == 13}
    images = {(a, b) for a, b, c in pairs if c != 13}
    assert len(images) > 0
    dup = [p for p in ref if p in real]
    missed = [p for p in images if p not in set(ref)]
    assert len(dup) == len(ref) > 0  # every ghost pair repeats a contact already listed
    assert len(missed) == len(images)  # no contact through the face is found
    # the intended list: every touching pair of the periodic lattice, each
    # once: 23 in each of the 20 rows, 47 between each of the 20 row pairs,
    # of which the 47 of the last and first row meet through the face
    assert len(real) + len(images) == 20 * 23 + 20 * 47
    assert len(images) == 47
