"""demsph_build_cells_noalloc (sort scratch in a fixed array inside the
library, no device allocation) against demsph_build_cells, through the C ABI.

Bit-exact: sorted hashes, the sort permutation and both cell-range arrays,
on two calls in a row, the second with fewer objects and another grid (the
fixed scratch keeps the first call's contents).
"""
import numpy as np
import pytest

import this_studio_amd as ts

pytestmark = pytest.mark.gpu


def _grid(mn, mx, c):
    cs, gs = ts.grid_geometry(mn, mx, c)
    return list(mn), list(mx), list(cs), [int(v) for v in gs], int(np.prod(gs))


def _run(fn, dpos, n, grid):
    import torch

    mn, mx, cs, gs, ng = grid
    dev = dpos.device
    hidx = torch.full((n,), -7, dtype=torch.int32, device=dev)
    hval = torch.full((n,), -7, dtype=torch.int32, device=dev)
    dcs = torch.full((ng,), -7, dtype=torch.int32, device=dev)
    dce = torch.full((ng,), -7, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    fn(dpos.data_ptr(), hidx.data_ptr(), hval.data_ptr(), dcs.data_ptr(), dce.data_ptr(), mn, mx, cs, gs, ng, n)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (hval, hidx, dcs, dce)]


def test_noalloc_build_cells_equals_build_cells_on_two_calls():
    import torch

    rng = np.random.default_rng(20261203)
    dev = torch.device("cuda:0")
    # first call: 300,000 objects (device scan of the digit table, 9-bit digits);
    # second: 40,000 objects on a coarser grid (bases fused into the scatter)
    cases = [(300_000, _grid([0.0, 0.0, 0.0], [1.0, 0.7, 0.4], 0.021)),
             (40_000, _grid([-0.5, -0.5, 0.0], [0.5, 0.5, 0.3], 0.05))]
    for n, grid in cases:
        mn, mx = np.array(grid[0]), np.array(grid[1])
        pos = rng.uniform(mn - 0.1, mx + 0.1, (n, 3))  # some points outside the domain hit the clamps
        dpos = torch.from_numpy(pos.reshape(-1)).to(dev)
        print(f"[no-alloc] n={n} cells={grid[4]}", flush=True)
        want = _run(ts.build_cells, dpos, n, grid)
        got = _run(ts.build_cells_noalloc, dpos, n, grid)
        for name, a, b in zip(("hashValue", "hashIndex", "cellHashStart", "cellHashEnd"), got, want):
            assert np.array_equal(a, b), f"{name} differs at n={n}"
        assert np.all(np.diff(got[0]) >= 0)
        assert np.array_equal(np.sort(got[1]), np.arange(n))


def test_noalloc_build_cells_rejects_oversized_input():
    import torch

    cap = ts.build_cells_noalloc_max_objects()
    assert cap == 1 << 21
    d = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    p = d.data_ptr()
    with pytest.raises(ts.DemsphError, match="scratch capacity"):
        ts.build_cells_noalloc(p, p, p, p, p, [0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [2, 2, 2], 8, cap + 1)
