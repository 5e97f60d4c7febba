"""Scene construction recipes (host side).

Point-cloud generators follow DEMSPH/pointCloudGeneration.h (uniform lattice
:7-56); the scene recipes follow the reference apps (damBreak.cpp:14-54) and
the workload definitions of this project (WCSPH dam break at dx = 5 mm,
DEM collapse of 4,096 spheres).  Pure numpy; no device code.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def uniform_point_cloud(min_bound, max_bound, dx: float) -> np.ndarray:
    """generateUniformPointCloud (pointCloudGeneration.h:7-56): inclusive
    lattice with count floor(len/dx + 1e-12) + 1 per axis, x fastest."""
    if dx <= 0:
        return np.zeros((0, 3))
    eps = 1e-12
    lo = np.minimum(min_bound, max_bound).astype(np.float64)
    hi = np.maximum(min_bound, max_bound).astype(np.float64)
    counts = [max(0, int(math.floor((hi[a] - lo[a]) / dx + eps)) + 1) for a in range(3)]
    axes = []
    for a in range(3):
        v = lo[a] + np.arange(counts[a]) * dx
        axes.append(v[v <= hi[a] + eps])
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)


def splitmix64(seed: int):
    """SplitMix64 stream of uint64 (the project's seeded PRNG)."""
    s = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def uniform01(seed: int, n: int) -> np.ndarray:
    g = splitmix64(seed)
    return np.array([(next(g) >> 11) * (1.0 / 9007199254740992.0) for _ in range(n)])


@dataclass
class SPHSceneSpec:
    """Arrays of one WCSPH scene in the shape addSPH/addDummy take them."""
    fluid_pos: np.ndarray
    dummy_pos: np.ndarray
    dx: float
    rho0: float
    c0: float
    nu: float
    min_bound: np.ndarray
    max_bound: np.ndarray
    gravity: np.ndarray
    dt: float
    eos: int = 0
    gamma: float = 7.0
    fluid_vel: np.ndarray | None = None
    dummy_vel: np.ndarray | None = None
    name: str = ""

    @property
    def nf(self) -> int:
        return self.fluid_pos.shape[0]

    @property
    def nd(self) -> int:
        return self.dummy_pos.shape[0]


def tank_dummies(tank, dx: float, layers: int = 3) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Dummy shell of an open-top tank [0,Lx]x[0,Ly]x[0,..] built like
    damBreak.cpp:23-34: lattice over the tank grown by `layers` spacings,
    keeping the points outside the walls (below z=0, outside x or y)."""
    lx, ly, lz = tank
    thick = layers * dx
    mn = np.array([-thick, -thick, -thick])
    mx = np.array([lx + thick, ly + thick, lz + thick])
    p0 = uniform_point_cloud(mn, mx, dx)
    keep = (p0[:, 0] < 0) | (p0[:, 1] < 0) | (p0[:, 2] < 0) | (p0[:, 0] > lx) | (p0[:, 1] > ly)
    return p0[keep], mn, mx


def dam_break(dx: float = 0.005, column=(0.4, 0.4, 0.2), tank=(1.6, 0.4, 0.6), eos: int = 1,
              gamma: float = 7.0, cfl: float = 0.2, nu: float = 1.0e-3, rho0: float = 1000.0,
              lattice_counts=None) -> SPHSceneSpec:
    """WCSPH dam break: a water column of `column` metres at spacing dx in an
    open tank with a 3-layer dummy shell.  c0 = 10 sqrt(2 g H), h = 1.3 dx,
    dt = cfl * h / c0.  The column lattice holds round(L/dx) points per axis
    starting at the tank corner (80 x 80 x 40 = 256,000 at dx = 5 mm)."""
    g = 9.81
    H = column[2]
    c0 = 10.0 * math.sqrt(2.0 * g * H)
    if lattice_counts is None:
        lattice_counts = [int(round(column[a] / dx)) for a in range(3)]
    nx, ny, nz = lattice_counts
    xs, ys, zs = np.arange(nx) * dx, np.arange(ny) * dx, np.arange(nz) * dx
    z, y, x = np.meshgrid(zs, ys, xs, indexing="ij")
    fluid = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    dummies, mn, mx = tank_dummies(tank, dx)
    h = 1.3 * dx
    return SPHSceneSpec(fluid_pos=fluid, dummy_pos=dummies, dx=dx, rho0=rho0, c0=c0, nu=nu,
                        min_bound=mn, max_bound=mx, gravity=np.array([0.0, 0.0, -g]),
                        dt=cfl * h / c0, eos=eos, gamma=gamma, name="wcsph_dam_break")


def reference_dam_break(H: float = 0.1, spacing: float = 0.005) -> SPHSceneSpec:
    """The reference's damBreak.cpp scene (damBreak.cpp:14-54): fluid block
    H x 2H x H, tank 5H x 2H x 3H, c = 20 sqrt(g H), linear EOS, dt =
    0.25*1.3*dx/(c + 2 sqrt(g H))."""
    n = int(H / spacing)
    if n > 0:
        spacing = H / n
    fluid = uniform_point_cloud(np.zeros(3), np.array([H, 2 * H, H]), spacing)
    thick = np.array([3 * spacing] * 3)
    p0 = uniform_point_cloud(-thick, np.array([5 * H, 2 * H, 3 * H]) + thick, spacing)
    keep = (p0[:, 0] < 0) | (p0[:, 1] < 0) | (p0[:, 2] < 0) | (p0[:, 0] > 5 * H) | (p0[:, 1] > 2 * H)
    c = 20.0 * math.sqrt(9.81 * H)
    dt = 0.25 * 1.3 * spacing / (c + 2.0 * math.sqrt(9.81 * H))
    return SPHSceneSpec(fluid_pos=fluid, dummy_pos=p0[keep], dx=spacing, rho0=1000.0, c0=c, nu=1.0e-3,
                        min_bound=-thick, max_bound=np.array([5 * H, 2 * H, 3 * H]) + thick,
                        gravity=np.array([0.0, 0.0, -9.81]), dt=dt, eos=0, name="reference_dam_break")


def small_dam_break(nx: int = 12, ny: int = 6, nz: int = 10, dx: float = 0.01, eos: int = 0) -> SPHSceneSpec:
    """A small dam break (a few thousand particles) for parity tests."""
    tank = (nx * dx * 3.0, ny * dx + dx, nz * dx * 1.5)
    spec = dam_break(dx=dx, column=(nx * dx, ny * dx, nz * dx), tank=tank, eos=eos,
                     lattice_counts=(nx, ny, nz))
    spec.name = "small_dam_break"
    return spec
