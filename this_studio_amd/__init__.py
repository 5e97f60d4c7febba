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


class TuneReport(C.Structure):
    """demsph_sph_tune_report: what the autotuner chose and measured."""
    _fields_ = [("listLanes", C.c_int), ("listPrefilter", C.c_int), ("pairLanes", C.c_int),
                ("rulesListLanes", C.c_int), ("rulesPairLanes", C.c_int), ("configs", C.c_int),
                ("cached", C.c_int), ("listUs", C.c_double), ("pairUs", C.c_double),
                ("rulesListUs", C.c_double), ("rulesPairUs", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


def _declare(L: C.CDLL) -> None:
    L.demsph_last_error.restype = C.c_char_p
    L.demsph_version.restype = C.c_char_p
    sig = {
        "demsph_build_cells": [_vp, _vp, _vp, _vp, _vp, _dp, _dp, _dp, _ip, _sz, _sz, _vp],
        "demsph_build_cells_noalloc": [_vp, _vp, _vp, _vp, _vp, _dp, _dp, _dp, _ip, _sz, _sz, _vp],
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
        "demsph_sph_set_graphs": [_vp, C.c_int],
        "demsph_sph_graph_captures": [_vp, C.POINTER(C.c_long)],
        "demsph_sph_tune": [_vp, C.c_char_p, C.c_int],
        "demsph_sph_tune_get": [_vp, C.c_char_p, C.POINTER(C.c_int)],
        "demsph_sph_autotune": [_vp, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(TuneReport)],
        "demsph_sph_autotune_clear_cache": [],
        "demsph_sph_write_vtu": [_vp, C.c_char_p, C.c_int, C.c_long, C.c_long, C.c_double],
        "demsph_sph_write_vtu_async": [_vp, C.c_char_p, C.c_int, C.c_long, C.c_long, C.c_double],
        "demsph_frames_wait": [],
        "demsph_sph_export_state": [_vp, _vp, _sz, C.POINTER(_sz)],
        "demsph_sph_import_state": [_vp, _vp, _sz],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.demsph_build_cells_noalloc_max_objects.argtypes = []
    L.demsph_build_cells_noalloc_max_objects.restype = C.c_size_t


def frames_wait() -> None:
    """Waits for every frame written in the background; raises on the first
    failure since the last call."""
    check(lib().demsph_frames_wait())


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
        self.L = lib()  # kept so that close() works during interpreter teardown
        check(self.L.demsph_sph_create(C.byref(self._h), _vp(stream) if stream else None))

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, _vp()
            st = self.L.demsph_sph_destroy(h)
            if st != 0:
                raise DemsphError(f"demsph status {st}: {self.L.demsph_last_error().decode()}")

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

    def export_state(self) -> bytes:
        """The fluid's dynamic state (current cell-sorted order) as bytes."""
        n = _sz()
        check(lib().demsph_sph_export_state(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().demsph_sph_export_state(self._h, buf, n.value, C.byref(n)))
        return buf.raw

    def import_state(self, state: bytes) -> None:
        """Continues from a state exported by a scene of the same particles."""
        check(lib().demsph_sph_import_state(self._h, state, len(state)))

    def write_vtu(self, directory, frame: int, step: int, time: float, fluid: bool = True,
                  dummies: bool = True, background: bool = False) -> None:
        """The reference's SPH_FFFF.vtu / dummy_FFFF.vtu frame files.  With
        background=True the fluid frame is snapshotted in stream order and
        written by the frame thread (frames_wait() waits for it)."""
        what = (1 if fluid else 0) | (2 if dummies else 0)
        fn = lib().demsph_sph_write_vtu_async if background else lib().demsph_sph_write_vtu
        check(fn(self._h, str(directory).encode(), what, int(frame), int(step), float(time)))

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

    def set_graphs(self, on: bool = True):
        """Replay captured two-step graphs in step() (on by default)."""
        check(lib().demsph_sph_set_graphs(self._h, 1 if on else 0))

    def tune(self, knob: str, value: int) -> None:
        """Kernel tuning knob ("pairLanes": 0 automatic, 1, 2, 4, 8)."""
        check(lib().demsph_sph_tune(self._h, knob.encode(), int(value)))

    def tuned(self, knob: str) -> int:
        v = C.c_int()
        check(lib().demsph_sph_tune_get(self._h, knob.encode(), C.byref(v)))
        return v.value

    def autotune(self, dt: float, trial_steps: int = 4, mode: str = "exact", use_cache: bool = True) -> dict:
        """Time the list-build and pair-pass variants on this scene's own steps
        (state restored afterwards) and keep the fastest.  mode "exact" tries
        only knobs that leave every result bit-identical; "pairs" also tries
        pair lanes (FP32 summation order).  Returns the report as a dict."""
        modes = {"exact": 0, "pairs": 1}
        if mode not in modes:
            raise ValueError(f"mode must be one of {sorted(modes)}")
        r = TuneReport()
        check(lib().demsph_sph_autotune(self._h, float(dt), int(trial_steps), modes[mode],
                                        1 if use_cache else 0, C.byref(r)))
        return r.as_dict()

    @staticmethod
    def clear_autotune_cache() -> None:
        check(lib().demsph_sph_autotune_clear_cache())

    def graph_captures(self) -> int:
        n = C.c_long()
        check(lib().demsph_sph_graph_captures(self._h, C.byref(n)))
        return n.value

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


def build_cells_noalloc(pos_dev_ptr: int, hash_index_ptr: int, hash_value_ptr: int, cell_start_ptr: int,
                        cell_end_ptr: int, min_bound, max_bound, cell_size, grid_size, num_grids: int, n: int,
                        stream: int | None = None) -> None:
    """demsph_build_cells_noalloc: build_cells with its sort scratch in a fixed
    array inside the library (no device allocation; at most
    build_cells_noalloc_max_objects() objects; overlapping calls on one GPU race)."""
    mn, mx, cs = (_f64(x, 3) for x in (min_bound, max_bound, cell_size))
    gs = np.ascontiguousarray(np.asarray(grid_size, dtype=np.int32))
    check(lib().demsph_build_cells_noalloc(_vp(pos_dev_ptr), _vp(hash_index_ptr), _vp(hash_value_ptr),
                                           _vp(cell_start_ptr), _vp(cell_end_ptr), _d(mn), _d(mx), _d(cs),
                                           gs.ctypes.data_as(_ip), num_grids, n, _vp(stream) if stream else None))


def build_cells_noalloc_max_objects() -> int:
    return int(lib().demsph_build_cells_noalloc_max_objects())


def grid_geometry(min_bound, max_bound, cell_one_dim: float):
    """spatialGrid::set (DEMSPH/kernel/boundary.h:70-94) -> (cellSize[3], gridSize[3])."""
    mn, mx = _f64(min_bound, 3), _f64(max_bound, 3)
    cs = np.zeros(3)
    gs = np.zeros(3, dtype=np.int32)
    check(lib().demsph_grid_geometry(_d(mn), _d(mx), float(cell_one_dim), _d(cs), gs.ctypes.data_as(_ip)))
    return cs, gs


# --------------------------------------------------------------------- DEM
_DEM_DECLARED = False
HOOK_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t, C.c_double)


class BallViews(C.Structure):
    _fields_ = [("force", C.c_void_p), ("torque", C.c_void_p), ("position", C.c_void_p), ("velocity", C.c_void_p),
                ("angularVelocity", C.c_void_p), ("radius", C.c_void_p), ("inverseMass", C.c_void_p),
                ("numBalls", C.c_size_t)]


def _declare_dem(L):
    global _DEM_DECLARED
    if _DEM_DECLARED:
        return
    d = C.c_double
    sig = {
        "demsph_dem_create": [C.POINTER(_vp), _vp],
        "demsph_dem_destroy": [_vp],
        "demsph_dem_set_linear_pair": [_vp, C.c_int, C.c_int] + [d] * 11,
        "demsph_dem_set_hertzian_pair": [_vp, C.c_int, C.c_int] + [d] * 8,
        "demsph_dem_set_bonded_pair": [_vp, C.c_int, C.c_int] + [d] * 6,
        "demsph_dem_add_balls": [_vp, _sz, _dp, _dp, _dp, _dp, _dp, _ip, _ip],
        "demsph_dem_add_fixed_balls": [_vp, _sz, _dp, _dp, _ip, _ip],
        "demsph_dem_add_clump": [_vp, _sz, _dp, _dp, _dp, _dp, _dp, d, _dp, C.c_int],
        "demsph_dem_add_fixed_clump": [_vp, _sz, _dp, _dp, _dp, C.c_int],
        "demsph_dem_add_mesh_wall": [_vp, _sz, _dp, _sz, _ip, _ip, _ip, _dp, _dp, _dp, C.c_int],
        "demsph_dem_add_bonded_interactions": [_vp, _sz, _ip, _ip, _sz, _dp],
        "demsph_dem_set_precision": [_vp, C.c_int],
        "demsph_dem_set_periodic_boundary": [_vp, C.c_int, C.c_int, C.c_int],
        "demsph_dem_initialize": [_vp, _dp, _dp],
        "demsph_dem_neighbor_search": [_vp],
        "demsph_dem_contact": [_vp, d],
        "demsph_dem_triangle_contact": [_vp, d],
        "demsph_dem_first_half": [_vp, _dp, d],
        "demsph_dem_second_half": [_vp, _dp, d],
        "demsph_dem_solve_begin": [_vp],
        "demsph_dem_step": [_vp, _dp, d, C.c_int],
        "demsph_dem_synchronize": [_vp],
        "demsph_dem_clock": [_vp, C.POINTER(_sz), C.POINTER(d)],
        "demsph_dem_set_hook": [_vp, HOOK_FN, _vp],
        "demsph_dem_views": [_vp, C.POINTER(BallViews)],
        "demsph_dem_hook_clear": [_vp],
        "demsph_dem_hook_constant_force": [_vp, _dp, d],
        "demsph_dem_hook_global_damping": [_vp, d],
        "demsph_dem_hook_buoyancy_drag": [_vp, d, d, _dp, _dp, d],
        "demsph_dem_sizes": [_vp] + [C.POINTER(_sz)] * 6,
        "demsph_dem_get": [_vp, C.c_char_p, _vp, _sz, C.POINTER(_sz)],
        "demsph_dem_write_vtu": [_vp, C.c_char_p, C.c_int, C.c_long, C.c_long, C.c_double],
        "demsph_dem_write_vtu_async": [_vp, C.c_char_p, C.c_int, C.c_long, C.c_long, C.c_double],
        "demsph_dem_record_contacts": [_vp, C.c_int],
        "demsph_dem_export_state": [_vp, _vp, _sz, C.POINTER(_sz)],
        "demsph_dem_import_state": [_vp, _vp, _sz],
        "demsph_dem_set": [_vp, C.c_char_p, _vp, _sz],
        "demsph_dem_set_graphs": [_vp, C.c_int],
        "demsph_dem_tune": [_vp, C.c_char_p, C.c_int],
        "demsph_dem_tune_get": [_vp, C.c_char_p, C.POINTER(C.c_int)],
        "demsph_dem_autotune": [_vp, _dp, C.c_double, C.c_int, C.POINTER(C.c_int)],
        "demsph_dem_graph_captures": [_vp, C.POINTER(C.c_long)],
        "demsph_dem_enable_timing": [_vp, C.c_int],
        "demsph_dem_timing": [_vp, C.c_int, C.c_char_p, _sz, C.POINTER(C.c_long), C.POINTER(C.c_double)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.demsph_build_cells_noalloc_max_objects.argtypes = []
    L.demsph_build_cells_noalloc_max_objects.restype = C.c_size_t
    _DEM_DECLARED = True


DEM_INT_FIELDS = {"objectPointed", "objectPointing", "pairShift", "triPointed", "triPointing", "triCancel", "bondPointed",
                  "bondPointing", "bondIsBonded", "hashValue", "hashIndex", "cellHashStart", "cellHashEnd"}
DEM_VEC_FIELDS = {"position", "velocity", "angularVelocity", "force", "torque", "slidingSpring", "rollingSpring",
                  "torsionSpring", "triSlidingSpring", "triRollingSpring", "triTorsionSpring", "bondShearForce", "bondBendingTorque", "bondPoint",
                  "bondNormal", "clumpPosition", "clumpVelocity", "clumpAngularVelocity", "wallPosition",
                  "vertexGlobalPosition", "triangleCircumcenter"}


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))


class DEMSceneGPU:
    """One DEM scene on the GPU (the reference's DEMSolver,
    DEMSPH/kernel/DEMSolver.h:144-2079) through the C ABI, built from a
    this_studio_amd.scenes.DEMScene recording."""

    def __init__(self, scene=None, fp64: bool = False, solve_begin: bool = True, stream: int | None = None,
                 initialize: bool = True):
        """initialize=False stages the scene's materials, balls and walls on
        the host only (a job for Server.submit, which initialises it on
        admission); hooks and bonds need an initialised scene."""
        L = lib()
        _declare_dem(L)
        self.L = L
        self._h = _vp()
        self._hook_ref = None
        check(L.demsph_dem_create(C.byref(self._h), _vp(stream) if stream else None))
        check(L.demsph_dem_set_precision(self._h, 1 if fp64 else 0))
        self.scene = scene
        if scene is not None:
            if not initialize and (scene.hooks or scene.bonds_from_initial_pairs or scene.explicit_bonds is not None):
                raise ValueError("hooks and bonds need an initialised scene")
            self._replay(scene, initialize)
            if solve_begin and initialize:
                self.solve_begin()

    def _replay(self, s, initialize: bool = True):
        L, h = self.L, self._h
        per = getattr(s, "periodic", (False, False, False))
        if any(per):
            check(L.demsph_dem_set_periodic_boundary(h, int(per[0]), int(per[1]), int(per[2])))
        for kind, a, b, p in s.materials:
            fn = {"linear": L.demsph_dem_set_linear_pair, "hertz": L.demsph_dem_set_hertzian_pair,
                  "bonded": L.demsph_dem_set_bonded_pair}[kind]
            check(fn(h, int(a), int(b), *[float(x) for x in p]))
        for kind, o in s.objects:
            if kind == "balls":
                n = o["pos"].shape[0]
                check(L.demsph_dem_add_balls(h, n, _d(_f64(o["pos"], n, 3)), _d(_f64(o["vel"], n, 3)),
                                             _d(_f64(o["w"], n, 3)), _d(_f64(o["r"], n)), _d(_f64(o["density"], n)),
                                             _i32(o["mat"]).ctypes.data_as(_ip), _i32(o["clump"]).ctypes.data_as(_ip)))
            elif kind == "fixed":
                n = o["pos"].shape[0]
                check(L.demsph_dem_add_fixed_balls(h, n, _d(_f64(o["pos"], n, 3)), _d(_f64(o["r"], n)),
                                                   _i32(o["mat"]).ctypes.data_as(_ip),
                                                   _i32(o["clump"]).ctypes.data_as(_ip)))
            elif kind == "clump":
                n = o["points"].shape[0]
                check(L.demsph_dem_add_clump(h, n, _d(_f64(o["points"], n, 3)), _d(_f64(o["radii"], n)),
                                             _d(_f64(o["centroid"], 3)), _d(_f64(o["vel"], 3)), _d(_f64(o["w"], 3)),
                                             float(o["mass"]), _d(_f64(o["inertia"], 6)), int(o["mat"])))
            elif kind == "fixed_clump":
                n = o["points"].shape[0]
                check(L.demsph_dem_add_fixed_clump(h, n, _d(_f64(o["points"], n, 3)), _d(_f64(o["radii"], n)),
                                                   _d(_f64(o["centroid"], 3)), int(o["mat"])))
            elif kind == "wall":
                nv = o["verts"].shape[0]
                t0, t1, t2 = _i32(o["t0"]), _i32(o["t1"]), _i32(o["t2"])
                check(L.demsph_dem_add_mesh_wall(h, nv, _d(_f64(o["verts"], nv, 3)), t0.size,
                                                 t0.ctypes.data_as(_ip), t1.ctypes.data_as(_ip),
                                                 t2.ctypes.data_as(_ip), _d(_f64(o["pos"], 3)), _d(_f64(o["vel"], 3)),
                                                 _d(_f64(o["w"], 3)), int(o["mat"])))
        if not initialize:
            return
        check(L.demsph_dem_initialize(h, _d(_f64(s.min_bound, 3)), _d(_f64(s.max_bound, 3))))
        n = self.sizes()[0]
        for hk in s.hooks:
            if hk[0] == "const":
                check(L.demsph_dem_hook_constant_force(h, _d(_f64(hk[1], n, 3)), float(hk[2])))
            elif hk[0] == "damping":
                check(L.demsph_dem_hook_global_damping(h, float(hk[1])))
            elif hk[0] == "buoyancy":
                check(L.demsph_dem_hook_buoyancy_drag(h, float(hk[1]), float(hk[2]), _d(_f64(hk[3], 3)),
                                                      _d(_f64(hk[4], 3)), float(hk[5])))

    def solve_begin(self):
        """solve() up to the loop, with addInitialCondition's bonds from the
        initial pair list when the scene asks for them (DEMSolver.h:1994-1998)."""
        check(self.L.demsph_dem_solve_begin(self._h))
        s = self.scene
        bonds = None
        if s is not None and s.bonds_from_initial_pairs:
            bonds = (self.get("objectPointed"), self.get("objectPointing"))
        if s is not None and s.explicit_bonds is not None:
            bonds = s.explicit_bonds
        if bonds is not None:
            pos = self.get("position")
            a, b = _i32(bonds[0]), _i32(bonds[1])
            check(self.L.demsph_dem_add_bonded_interactions(self._h, a.size, a.ctypes.data_as(_ip),
                                                            b.ctypes.data_as(_ip), pos.shape[0],
                                                            _d(_f64(pos, pos.shape[0], 3))))
            check(self.L.demsph_dem_initialize(self._h, _d(_f64(s.min_bound, 3)), _d(_f64(s.max_bound, 3))))

    def close(self):
        if self._h:
            h, self._h = self._h, _vp()
            st = self.L.demsph_dem_destroy(h)
            if st != 0:
                raise DemsphError(f"demsph status {st}: {self.L.demsph_last_error().decode()}")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1, dt: float | None = None, gravity=None):
        g = _f64(self.scene.gravity if gravity is None else gravity, 3)
        check(self.L.demsph_dem_step(self._h, _d(g), float(dt or self.scene.dt), int(n)))

    def neighbor_search(self):
        check(self.L.demsph_dem_neighbor_search(self._h))

    def contact(self, dt):
        check(self.L.demsph_dem_contact(self._h, float(dt)))

    def triangle_contact(self, dt):
        check(self.L.demsph_dem_triangle_contact(self._h, float(dt)))

    def synchronize(self):
        check(self.L.demsph_dem_synchronize(self._h))

    def clock(self):
        i, t = _sz(), C.c_double()
        check(self.L.demsph_dem_clock(self._h, C.byref(i), C.byref(t)))
        return i.value, t.value

    def sizes(self):
        v = [_sz() for _ in range(6)]
        check(self.L.demsph_dem_sizes(self._h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_hook(self, fn):
        """fn(iStep, time) runs on the host at each addExternalForceTorque
        point with views() valid; None removes it."""
        if fn is None:
            self._hook_ref = None
            check(self.L.demsph_dem_set_hook(self._h, HOOK_FN(), None))
            return
        self._hook_ref = HOOK_FN(lambda user, i, t: fn(i, t))
        check(self.L.demsph_dem_set_hook(self._h, self._hook_ref, None))

    def views(self) -> BallViews:
        v = BallViews()
        check(self.L.demsph_dem_views(self._h, C.byref(v)))
        return v

    def record_contacts(self, on: bool = True) -> None:
        """Keep each ball-triangle contact of every evaluation (for the
        ballTriangleInteraction frame file)."""
        check(self.L.demsph_dem_record_contacts(self._h, 1 if on else 0))

    def export_state(self) -> bytes:
        """Ball, contact-history, clump, wall and bond state, step and time."""
        n = _sz()
        check(self.L.demsph_dem_export_state(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(self.L.demsph_dem_export_state(self._h, buf, n.value, C.byref(n)))
        return buf.raw

    def import_state(self, state: bytes) -> None:
        """Continues from a state exported by a scene staged the same way."""
        check(self.L.demsph_dem_import_state(self._h, state, len(state)))

    def write_vtu(self, directory, frame: int, step: int, time: float, balls: bool = True,
                  walls: bool = True, wall_contacts: bool = False, pairs: bool = False,
                  bonds: bool = False, background: bool = False) -> None:
        """The reference's ball_FFFF.vtu / triangleWall_FFFF.vtu /
        ballTriangleInteraction_FFFF.vtu / ballInteraction_FFFF.vtu frame
        files (the contact files need record_contacts before the
        evaluation they report)."""
        what = ((1 if balls else 0) | (2 if walls else 0) | (4 if wall_contacts else 0) | (8 if pairs else 0)
                | (16 if bonds else 0))
        fn = self.L.demsph_dem_write_vtu_async if background else self.L.demsph_dem_write_vtu
        check(fn(self._h, str(directory).encode(), what, int(frame), int(step), float(time)))

    def get(self, field: str) -> np.ndarray:
        cnt = _sz()
        check(self.L.demsph_dem_get(self._h, field.encode(), None, 0, C.byref(cnt)))
        dt = np.int32 if field in DEM_INT_FIELDS else np.float64
        out = np.empty(cnt.value, dtype=dt)
        check(self.L.demsph_dem_get(self._h, field.encode(), out.ctypes.data_as(_vp), cnt.value, C.byref(cnt)))
        if field in DEM_VEC_FIELDS:
            out = out.reshape(-1, 3)
        if field in ("clumpOrientation", "wallOrientation"):
            out = out.reshape(-1, 4)
        return out

    def set(self, field: str, values):
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
        check(self.L.demsph_dem_set(self._h, field.encode(), v.ctypes.data_as(_vp), v.size))

    def set_graphs(self, on: bool = True):
        """Replay captured four-step graphs between pair rebuilds (on by
        default; needs a scene stream other than the null stream)."""
        check(self.L.demsph_dem_set_graphs(self._h, 1 if on else 0))

    def tune(self, knob: str, value: int) -> None:
        """Kernel tuning knob ("pairLanes": 0 automatic, 1, 2, 4, 8)."""
        check(self.L.demsph_dem_tune(self._h, knob.encode(), int(value)))

    def tuned(self, knob: str) -> int:
        v = C.c_int()
        check(self.L.demsph_dem_tune_get(self._h, knob.encode(), C.byref(v)))
        return v.value

    def autotune(self, trial_steps: int = 20, dt: float | None = None, gravity=None) -> int:
        """Times each "pairLanes" setting from the current state, keeps the
        fastest and restores the state; returns the lanes chosen."""
        s = self.scene
        dt = s.dt if dt is None else dt
        g = _f64(s.gravity if gravity is None else gravity, 3)
        v = C.c_int()
        check(self.L.demsph_dem_autotune(self._h, _d(g), float(dt), int(trial_steps), C.byref(v)))
        return v.value

    def graph_captures(self) -> int:
        n = C.c_long()
        check(self.L.demsph_dem_graph_captures(self._h, C.byref(n)))
        return n.value

    def enable_timing(self, on: bool = True):
        check(self.L.demsph_dem_enable_timing(self._h, 1 if on else 0))

    def timing(self) -> dict:
        out, i = {}, 0
        buf = C.create_string_buffer(128)
        while True:
            n, ms = C.c_long(), C.c_double()
            if self.L.demsph_dem_timing(self._h, i, buf, 128, C.byref(n), C.byref(ms)) != 0:
                break
            out[buf.value.decode()] = (n.value, ms.value)
            i += 1
        return out


class CoupledScene:
    """Two-way DEM-SPH coupling of one fluid scene and one DEM scene on the
    GPU (demsph_couple_*): both advance together with one time step."""

    def __init__(self, fluid: "SPHScene", dem: "DEMSceneGPU", capacity: int | None = None):
        L = lib()
        L.demsph_couple_create.argtypes = [C.POINTER(_vp), _vp, _vp]
        L.demsph_couple_destroy.argtypes = [_vp]
        L.demsph_couple_set_capacity.argtypes = [_vp, C.c_int]
        L.demsph_couple_step.argtypes = [_vp, _dp, C.c_double, C.c_int]
        L.demsph_couple_get.argtypes = [_vp, C.c_char_p, _vp, _sz, C.POINTER(_sz)]
        self.L, self.fluid, self.dem = L, fluid, dem
        self._h = _vp()
        check(L.demsph_couple_create(C.byref(self._h), fluid._h, dem._h))
        if capacity:
            check(L.demsph_couple_set_capacity(self._h, int(capacity)))

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, _vp()
            st = self.L.demsph_couple_destroy(h)
            if st != 0:
                raise DemsphError(f"demsph status {st}: {self.L.demsph_last_error().decode()}")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, dt: float, n: int = 1, gravity=(0.0, 0.0, -9.81)):
        g = _f64(gravity, 3)
        check(self.L.demsph_couple_step(self._h, _d(g), float(dt), int(n)))

    def get(self, field: str) -> np.ndarray:
        cnt = _sz()
        check(self.L.demsph_couple_get(self._h, field.encode(), None, 0, C.byref(cnt)))
        out = np.empty(cnt.value, dtype=np.float64)
        check(self.L.demsph_couple_get(self._h, field.encode(), out.ctypes.data_as(_vp), cnt.value, C.byref(cnt)))
        if field not in ("ballPressure", "blockReplays"):
            out = out.reshape(-1, 3)
        return out



class JobInfo(C.Structure):
    _fields_ = [("kind", C.c_int), ("status", C.c_int), ("particles", C.c_long), ("steps", C.c_long),
                ("stepsDone", C.c_long), ("bytes", C.c_size_t), ("arrival", C.c_double), ("admitted", C.c_double),
                ("started", C.c_double), ("stepped", C.c_double), ("finished", C.c_double)]


class ServerSummary(C.Structure):
    _fields_ = [("wall", C.c_double), ("particleSteps", C.c_double), ("particleStepsPerSecond", C.c_double),
                ("p50", C.c_double), ("p99", C.c_double), ("jobsDone", C.c_int), ("jobsFailed", C.c_int),
                ("peakBytes", C.c_size_t), ("peakActive", C.c_int)]


_SERVER_DECLARED = False


PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_double)


def _declare_server(L):
    global _SERVER_DECLARED
    if _SERVER_DECLARED:
        return
    sig = {
        "demsph_server_create": [C.POINTER(_vp), C.c_int, C.c_double, C.c_int],
        "demsph_server_destroy": [_vp],
        "demsph_server_set_block": [_vp, C.c_long],
        "demsph_server_set_autotune": [_vp, C.c_int, C.c_int],
        "demsph_server_set_autotune_min_steps": [_vp, C.c_long],
        "demsph_server_submit": [_vp, _vp, _vp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_long, C.c_double,
                                 C.POINTER(C.c_int)],
        "demsph_server_run": [_vp],
        "demsph_server_set_progress": [_vp, PROGRESS_FN, _vp],
        "demsph_server_job_info": [_vp, C.c_int, C.POINTER(JobInfo), C.c_char_p, _sz],
        "demsph_server_job_latencies": [_vp, C.c_int, _dp, _sz, C.POINTER(_sz)],
        "demsph_server_job_result": [_vp, C.c_int, C.c_char_p, _dp, _sz, C.POINTER(_sz)],
        "demsph_server_stats": [_vp, C.POINTER(ServerSummary)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.demsph_build_cells_noalloc_max_objects.argtypes = []
    L.demsph_build_cells_noalloc_max_objects.restype = C.c_size_t
    _SERVER_DECLARED = True


class Server:
    """Batched serving of independent scenes on one GPU (demsph_server_*):
    FIFO admission by arrival time, free slot and device-memory budget, each
    admitted job on its own stream, per-step device timestamps.  slots=1,
    graphs=False is the plain one-scene-at-a-time path."""

    KINDS = ("wcsph", "dem", "coupled")

    def __init__(self, slots: int = 4, memory_budget: float = 288e9, graphs: bool = True, block: int = 100,
                 autotune: str | None = None, tune_steps: int = 3, tune_min_steps: int = 2000):
        """autotune: None (off), "exact" (result-invariant knobs) or "pairs"
        (also pair lanes): the fluid and DEM scenes of every job of at least
        tune_min_steps steps are tuned on admission."""
        modes = {None: -1, "exact": 0, "pairs": 1}
        if autotune not in modes:
            raise ValueError("autotune must be None, 'exact' or 'pairs'")
        L = lib()
        _declare_server(L)
        self.L = L
        self._h = _vp()
        check(L.demsph_server_create(C.byref(self._h), int(slots), float(memory_budget), 1 if graphs else 0))
        check(L.demsph_server_set_block(self._h, int(block)))
        check(L.demsph_server_set_autotune(self._h, modes[autotune], int(tune_steps)))
        check(L.demsph_server_set_autotune_min_steps(self._h, int(tune_min_steps)))

    def close(self) -> None:
        if self._h:
            h, self._h = self._h, _vp()
            st = self.L.demsph_server_destroy(h)
            if st != 0:
                raise DemsphError(f"demsph status {st}: {self.L.demsph_last_error().decode()}")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, fluid: "SPHScene | None", dem: "DEMSceneGPU | None", gravity, dt: float, steps: int,
               arrival: float = 0.0, fluid_bounds=None, dem_bounds=None) -> int:
        """Hands the staged (not initialised) scenes to the server; bounds are
        (min, max) pairs.  Returns the job id."""
        fb = [_f64(x, 3) for x in fluid_bounds] if fluid is not None else [None, None]
        db = [_f64(x, 3) for x in dem_bounds] if dem is not None else [None, None]
        g = _f64(gravity, 3)
        jid = C.c_int()
        p = lambda a: _d(a) if a is not None else None
        check(self.L.demsph_server_submit(self._h, fluid._h if fluid is not None else None,
                                          dem._h if dem is not None else None, p(fb[0]), p(fb[1]), p(db[0]),
                                          p(db[1]), _d(g), float(dt), int(steps), float(arrival), C.byref(jid)))
        # the server owns the handles now
        if fluid is not None:
            fluid._h = _vp()
        if dem is not None:
            dem._h = _vp()
        return jid.value

    def set_progress(self, fn) -> None:
        """fn(jobs_finished, jobs_total, seconds) runs on a slot thread each
        time a job finishes during run(); None removes it."""
        if fn is None:
            self._progress_ref = None
            check(self.L.demsph_server_set_progress(self._h, PROGRESS_FN(), None))
            return
        self._progress_ref = PROGRESS_FN(lambda user, done, total, t: fn(done, total, t))
        check(self.L.demsph_server_set_progress(self._h, self._progress_ref, None))

    def run(self) -> None:
        check(self.L.demsph_server_run(self._h))

    def info(self, job: int) -> dict:
        i = JobInfo()
        err = C.create_string_buffer(512)
        check(self.L.demsph_server_job_info(self._h, int(job), C.byref(i), err, 512))
        return {"kind": self.KINDS[i.kind], "status": i.status, "error": err.value.decode(),
                "particles": i.particles, "steps": i.steps, "steps_done": i.stepsDone, "bytes": i.bytes,
                "arrival": i.arrival, "admitted": i.admitted, "started": i.started, "stepped": i.stepped,
                "finished": i.finished}

    def latencies(self, job: int) -> np.ndarray:
        n = _sz()
        check(self.L.demsph_server_job_latencies(self._h, int(job), None, 0, C.byref(n)))
        out = np.empty(n.value)
        check(self.L.demsph_server_job_latencies(self._h, int(job), _d(out), n.value, C.byref(n)))
        return out

    def result(self, job: int, field: str) -> np.ndarray:
        n = _sz()
        check(self.L.demsph_server_job_result(self._h, int(job), field.encode(), None, 0, C.byref(n)))
        out = np.empty(n.value)
        check(self.L.demsph_server_job_result(self._h, int(job), field.encode(), _d(out), n.value, C.byref(n)))
        return out.reshape(-1, 3)

    def stats(self) -> dict:
        s = ServerSummary()
        check(self.L.demsph_server_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerSummary._fields_}
