"""MI355X-native DEMSPH particle path: Python host bindings over libdemsph.so.

The product is the C ABI in include/demsph.h, implemented by hand-written
HIP kernels for gfx950 in this_studio_amd/csrc.  This module loads that
library with ctypes and mirrors the reference solver classes' setup calls
(DEMSPH/kernel/WCSPHSolver.h, DEMSolver.h) for Python callers, tests and the
benchmark.  There is no CPU fallback: without the built library every entry
point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libdemsph.so"

_lib = None


class DemsphError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Loads libdemsph.so (once).  Raises when the HIP library is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise DemsphError(
                f"{LIB_PATH} is missing: build it with `python -m this_studio_amd.build` "
                "(there is no CPU fallback)")
        _lib = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_sz = C.c_size_t


def _declare(L: C.CDLL) -> None:
    L.demsph_last_error.restype = C.c_char_p
    L.demsph_version.restype = C.c_char_p
    sig = {
        "demsph_build_cells": [_vp, _vp, _vp, _vp, _vp, _dp, _dp, _dp, _ip, _sz, _sz, _vp],
        "demsph_grid_geometry": [_dp, _dp, C.c_double, _dp, _ip],
        "demsph_sph_create": [C.POINTER(_vp), _vp],
        "demsph_sph_destroy": [_vp],
        "demsph_sph_add_fluid": [_vp, _sz, _dp, _dp, _dp, _dp, _dp, _dp],
        "demsph_sph_add_dummy": [_vp, _sz, _dp, _dp, _dp, _dp, _dp, _dp],
        "demsph_sph_set_eos": [_vp, C.c_int, C.c_double],
        "demsph_sph_set_capacity": [_vp, C.c_int, C.c_int],
        "demsph_sph_initialize": [_vp, _dp, _dp, _dp],
        "demsph_sph_neighbor_search": [_vp],
        "demsph_sph_first_half": [_vp, C.c_double],
        "demsph_sph_second_half": [_vp, C.c_double],
        "demsph_sph_step": [_vp, C.c_double, C.c_int],
        "demsph_sph_synchronize": [_vp],
        "demsph_sph_sizes": [_vp, C.POINTER(_sz), C.POINTER(_sz), C.POINTER(_sz)],
        "demsph_sph_get": [_vp, C.c_char_p, _vp, _sz, C.POINTER(_sz)],
        "demsph_sph_set": [_vp, C.c_char_p, _vp, _sz],
        "demsph_sph_enable_timing": [_vp, C.c_int],
        "demsph_sph_timing": [_vp, C.c_int, C.c_char_p, _sz, C.POINTER(C.c_long), C.POINTER(C.c_double)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int


def check(status: int) -> None:
    if status != 0:
        msg = lib().demsph_last_error().decode()
        raise DemsphError(f"demsph status {status}: {msg}")


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _f64(a, n: int, comps: int = 1) -> np.ndarray:
    a = np.asarray(a, dtype=np.float64)
    if comps == 1 and a.ndim == 0:
        a = np.full(n, float(a))
    a = np.ascontiguousarray(a).reshape(-1)
    if a.size != n * comps:
        raise ValueError(f"expected {n * comps} values, got {a.size}")
    return a


INT_FIELDS = {"hashValue", "hashIndex", "cellHashStart", "cellHashEnd", "neighborCount",
              "dummyNeighborCount", "objectPointing", "dummyObjectPointing"}


class SPHScene:
    """One WCSPH scene on the GPU (the reference's WCSPHSolver,
    DEMSPH/kernel/WCSPHSolver.h:18-636), driven through the C ABI."""

    def __init__(self, stream: int | None = None):
        self._h = _vp()
        check(lib().demsph_sph_create(C.byref(self._h), _vp(stream) if stream else None))

    def close(self) -> None:
        if self._h:
            check(lib().demsph_sph_destroy(self._h))
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _add(self, fn, pos, vel, c, spacing, rho0, nu):
        pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
        n = pos.shape[0]
        vel = _f64(np.zeros((n, 3)) if vel is None else vel, n, 3)
        args = [_f64(x, n) for x in (c, spacing, rho0, nu)]
        check(fn(self._h, n, _d(pos.reshape(-1)), _d(vel), *[_d(x) for x in args]))

    def add_fluid(self, pos, vel, sound_speed, spacing, density, viscosity):
        """addSPH (WCSPHSolver.h:523-531), batched."""
        self._add(lib().demsph_sph_add_fluid, pos, vel, sound_speed, spacing, density, viscosity)

    def add_dummy(self, pos, vel, sound_speed, spacing, density, viscosity):
        """addDummy (WCSPHSolver.h:533-541), batched."""
        self._add(lib().demsph_sph_add_dummy, pos, vel, sound_speed, spacing, density, viscosity)

    def set_eos(self, eos: int, gamma: float = 7.0):
        check(lib().demsph_sph_set_eos(self._h, int(eos), float(gamma)))

    def set_capacity(self, fluid: int, dummy: int):
        check(lib().demsph_sph_set_capacity(self._h, int(fluid), int(dummy)))

    def initialize(self, min_bound, max_bound, gravity):
        mn, mx, g = (_f64(x, 3) for x in (min_bound, max_bound, gravity))
        check(lib().demsph_sph_initialize(self._h, _d(mn), _d(mx), _d(g)))

    def neighbor_search(self):
        check(lib().demsph_sph_neighbor_search(self._h))

    def first_half(self, dt: float):
        check(lib().demsph_sph_first_half(self._h, float(dt)))

    def second_half(self, dt: float):
        check(lib().demsph_sph_second_half(self._h, float(dt)))

    def step(self, dt: float, nsteps: int = 1):
        check(lib().demsph_sph_step(self._h, float(dt), int(nsteps)))

    def synchronize(self):
        check(lib().demsph_sph_synchronize(self._h))

    def sizes(self):
        a, b, c = _sz(), _sz(), _sz()
        check(lib().demsph_sph_sizes(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def get(self, field: str) -> np.ndarray:
        cnt = _sz()
        check(lib().demsph_sph_get(self._h, field.encode(), None, 0, C.byref(cnt)))
        dt = np.int32 if field in INT_FIELDS else np.float64
        out = np.empty(cnt.value, dtype=dt)
        check(lib().demsph_sph_get(self._h, field.encode(), out.ctypes.data_as(_vp), cnt.value, C.byref(cnt)))
        if field in ("position", "velocity", "dummyNormal"):
            out = out.reshape(-1, 3)
        return out

    def set(self, field: str, values) -> None:
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
        check(lib().demsph_sph_set(self._h, field.encode(), v.ctypes.data_as(_vp), v.size))

    def enable_timing(self, on: bool = True):
        check(lib().demsph_sph_enable_timing(self._h, 1 if on else 0))

    def timing(self) -> dict:
        out = {}
        i = 0
        buf = C.create_string_buffer(128)
        while True:
            n, ms = C.c_long(), C.c_double()
            st = lib().demsph_sph_timing(self._h, i, buf, 128, C.byref(n), C.byref(ms))
            if st != 0:
                break
            out[buf.value.decode()] = (n.value, ms.value)
            i += 1
        return out


def build_cells(pos_dev_ptr: int, hash_index_ptr: int, hash_value_ptr: int, cell_start_ptr: int,
                cell_end_ptr: int, min_bound, max_bound, cell_size, grid_size, num_grids: int, n: int,
                stream: int | None = None) -> None:
    """demsph_build_cells: the drop-in for launchUpdateSpatialGridCellHashStartEnd
    (DEMSPH/kernel/neighborSearchKernel.cu:137-174) over raw device pointers."""
    mn, mx, cs = (_f64(x, 3) for x in (min_bound, max_bound, cell_size))
    gs = np.ascontiguousarray(np.asarray(grid_size, dtype=np.int32))
    check(lib().demsph_build_cells(_vp(pos_dev_ptr), _vp(hash_index_ptr), _vp(hash_value_ptr), _vp(cell_start_ptr),
                                   _vp(cell_end_ptr), _d(mn), _d(mx), _d(cs), gs.ctypes.data_as(_ip), num_grids, n,
                                   _vp(stream) if stream else None))


def grid_geometry(min_bound, max_bound, cell_one_dim: float):
    """spatialGrid::set (DEMSPH/kernel/boundary.h:70-94) -> (cellSize[3], gridSize[3])."""
    mn, mx = _f64(min_bound, 3), _f64(max_bound, 3)
    cs = np.zeros(3)
    gs = np.zeros(3, dtype=np.int32)
    check(lib().demsph_grid_geometry(_d(mn), _d(mx), float(cell_one_dim), _d(cs), gs.ctypes.data_as(_ip)))
    return cs, gs
