"""The C-ABI library loads and exports every entry point its headers declare.

Runs without a GPU: only host-side logic (argument validation, grid
geometry, error reporting) is exercised; nothing launches a kernel.
"""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import this_studio_amd as ts

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_symbols():
    names = set()
    for h in HEADERS:
        txt = h.read_text()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"^\s*(?:int|size_t|const char\*|void)\s+(demsph_\w+|launch\w+)\s*\(", txt, flags=re.M))
    return sorted(names)


def test_headers_declare_entry_points():
    assert len(declared_symbols()) >= 20


@pytest.mark.parametrize("name", declared_symbols())
def test_symbol_exported(name):
    lib = C.CDLL(str(ts.LIB_PATH))
    assert hasattr(lib, name), f"{name} declared in include/ but not exported by libdemsph.so"


def test_version_and_error_reporting():
    L = ts.lib()
    assert b"gfx950" in L.demsph_version()
    mn = np.zeros(3)
    mx = np.array([1.0, 0.0, 1.0])  # degenerate y extent: the reference ignores such a grid
    cs = np.zeros(3)
    gs = np.zeros(3, dtype=np.int32)
    st = L.demsph_grid_geometry(mn.ctypes.data_as(C.POINTER(C.c_double)), mx.ctypes.data_as(C.POINTER(C.c_double)),
                                0.1, cs.ctypes.data_as(C.POINTER(C.c_double)), gs.ctypes.data_as(C.POINTER(C.c_int)))
    assert st == 1
    assert b"degenerate" in L.demsph_last_error()


def test_grid_geometry_matches_oracle():
    import oracle

    for mn, mx, c in [([0, 0, 0], [1.0, 0.3, 0.05], 0.1), ([-0.015] * 3, [1.615, 0.415, 0.615], 0.013),
                      ([-1, -2, -3], [1, 2, 3], 0.7)]:
        cs, gs = ts.grid_geometry(mn, mx, c)
        g = oracle.grid_set(mn, mx, c)
        assert tuple(gs) == (g.nx, g.ny, g.nz)
        assert tuple(cs) == (g.cx, g.cy, g.cz)


def test_scene_host_validation_without_gpu():
    s = ts.SPHScene()
    with pytest.raises(ts.DemsphError):
        s.set_eos(3)
    with pytest.raises(ts.DemsphError):
        s.neighbor_search()  # not initialised
    s.add_fluid(np.zeros((2, 3)), None, [10.0, -1.0], 0.01, 1000.0, 1e-3)  # c <= 0 is skipped (WCSPHSolver.h:526)
    assert s.sizes()[0] == 1
    s.close()


def test_noalloc_build_cells_refuses_oversized_input_on_the_host():
    """The capacity check runs before any device call, so it needs no GPU."""
    cap = ts.build_cells_noalloc_max_objects()
    assert cap == 1 << 21
    fake = 0x1000  # never dereferenced: the call is refused first
    with pytest.raises(ts.DemsphError, match="scratch capacity"):
        ts.build_cells_noalloc(fake, fake, fake, fake, fake, [0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [2, 2, 2], 8,
                               cap + 1)
