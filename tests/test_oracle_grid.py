"""Pins of the oracle's grid, hash, sort, cell-range and SPH pair-list
functions (oracle/oracle_sph.c) against closed forms and brute force."""
import numpy as np
import pytest

import oracle


def test_grid_set_closed_form():
    # spatialGrid::set (boundary.h:70-94): gridSize = int(L/c) if L > c else 1
    g = oracle.grid_set([0, 0, 0], [1.0, 0.3, 0.05], 0.1)
    # 0.3 / 0.1 = 2.9999999999999996 in binary64, so truncation gives 2 cells
    assert (g.nx, g.ny, g.nz) == (10, 2, 1)
    assert g.cx == pytest.approx(0.1) and g.cy == pytest.approx(0.15) and g.cz == pytest.approx(0.05)
    g = oracle.grid_set([-1, -2, -3], [1, 2, 3], 0.7)
    assert (g.nx, g.ny, g.nz) == (2, 5, 8)
    assert g.cx == 2.0 / 2 and g.cy == 4.0 / 5 and g.cz == 6.0 / 8
    with pytest.raises(ValueError):
        oracle.grid_set([0, 0, 0], [0, 1, 1], 0.1)
    with pytest.raises(ValueError):
        oracle.grid_set([0, 0, 0], [1, 1, 1], 0.0)


def test_hash_known_cells_and_clamps():
    g = oracle.grid_set([0, 0, 0], [1.0, 0.5, 0.375], 0.125)
    assert (g.nx, g.ny, g.nz) == (8, 4, 3)
    # cell centres hash to z*ny*nx + y*nx + x
    rng = np.random.default_rng(1)
    ijk = np.stack([rng.integers(0, 8, 500), rng.integers(0, 4, 500), rng.integers(0, 3, 500)], axis=1)
    pts = (ijk + 0.5) * np.array([g.cx, g.cy, g.cz])
    h = oracle.hash_points(g, pts)
    assert np.array_equal(h, ijk[:, 2] * 32 + ijk[:, 1] * 8 + ijk[:, 0])
    # below the minimum -> cell 0 on that axis; at or above the maximum -> last cell
    out = np.array([[-5.0, 0.05, 0.05], [1.0, 0.05, 0.05], [7.0, 9.0, -1.0], [0.0, 0.5, 0.375]])
    h = oracle.hash_points(g, out)
    assert h.tolist() == [0, 7, 3 * 8 + 7, 2 * 32 + 3 * 8 + 0]


def test_sort_is_stable_lexicographic():
    rng = np.random.default_rng(2)
    for n, ng in [(1, 4), (17, 3), (5000, 97), (20000, 20000)]:
        h = rng.integers(0, ng, n).astype(np.int32)
        sh, si = oracle.sort_by_hash(h, ng)
        ref = np.argsort(h, kind="stable")
        assert np.array_equal(si, ref)
        assert np.array_equal(sh, h[ref])


def test_cell_ranges_brute_force():
    rng = np.random.default_rng(3)
    ng = 64
    h = rng.integers(0, ng, 3000).astype(np.int32)
    h[:10] = ng + 5  # out of range keys are ignored by findStartAndEnd
    sh, si = oracle.sort_by_hash(h, ng)
    s, e = oracle.cell_ranges(sh, ng)
    for c in range(ng):
        idx = np.nonzero(sh == c)[0]
        if idx.size == 0:
            assert s[c] == -1 and e[c] == -1
        else:
            assert s[c] == idx[0] and e[c] == idx[-1] + 1


def _pairs_oracle(g, pos, h):
    hv = oracle.hash_points(g, pos)
    sh, si = oracle.sort_by_hash(hv, g.num_grids)
    cs, ce = oracle.cell_ranges(sh, g.num_grids)
    n = pos.shape[0]
    cnt = np.zeros(n, dtype=np.int32)
    L = oracle.lib()
    tot = L.o_sph_count(oracle.Grid.from_buffer_copy(g), oracle.d(pos), oracle.d(h), n, oracle.d(pos), oracle.d(h),
                        oracle.i(si), oracle.i(cs), oracle.i(ce), 1, oracle.i(cnt))
    pref = np.cumsum(cnt).astype(np.int32)
    pa = np.empty(tot, dtype=np.int32)
    pb = np.empty(tot, dtype=np.int32)
    L.o_sph_write(oracle.Grid.from_buffer_copy(g), oracle.d(pos), oracle.d(h), n, oracle.d(pos), oracle.d(h),
                  oracle.i(si), oracle.i(cs), oracle.i(ce), 1, oracle.i(pref), oracle.i(pa), oracle.i(pb))
    return cnt, pa, pb


def test_sph_pairs_match_brute_force():
    rng = np.random.default_rng(4)
    n = 1500
    pos = rng.uniform(0.0, 1.0, (n, 3))
    h = rng.uniform(0.02, 0.05, n)
    g = oracle.grid_set([0, 0, 0], [1, 1, 1], 2.0 * h.max())
    cnt, pa, pb = _pairs_oracle(g, pos, h)
    # brute force O(N^2): every j != i with |r|^2 <= (2 max(hi, hj))^2
    d2 = ((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
    cut = 2.0 * np.maximum(h[:, None], h[None, :])
    mask = (cut * cut - d2) >= 0
    np.fill_diagonal(mask, False)
    assert np.array_equal(cnt, mask.sum(1))
    got = set(zip(pa.tolist(), pb.tolist()))
    want = set(zip(*np.nonzero(mask)))
    assert got == {(int(a), int(b)) for a, b in want}
    # pointed array is grouped by ascending particle index
    assert np.all(np.diff(pa) >= 0)


def test_sph_pair_set_independent_of_cell_size():
    rng = np.random.default_rng(5)
    pos = rng.uniform(0.0, 1.0, (2000, 3))
    h = np.full(2000, 0.03)
    sets = []
    for c in (0.06, 0.09, 0.2):
        g = oracle.grid_set([0, 0, 0], [1, 1, 1], c)
        cnt, pa, pb = _pairs_oracle(g, pos, h)
        sets.append(set(zip(pa.tolist(), pb.tolist())))
    assert sets[0] == sets[1] == sets[2]
