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
    """The first n values of splitmix64(seed) as doubles in [0, 1) (top 53
    bits), vectorised: state k is seed + (k + 1) * golden gamma mod 2^64."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


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
    # per-fluid-particle properties (several phases); None: the scalars above
    fluid_c0: np.ndarray | None = None
    fluid_dx: np.ndarray | None = None
    fluid_rho0: np.ndarray | None = None
    fluid_nu: np.ndarray | None = None

    def fluid_props(self):
        """(sound speed, spacing, density, viscosity) of the fluid particles."""
        pick = lambda a, d: d if a is None else a
        return (pick(self.fluid_c0, self.c0), pick(self.fluid_dx, self.dx), pick(self.fluid_rho0, self.rho0),
                pick(self.fluid_nu, self.nu))

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


# ---------------------------------------------------------------------------
# DEM scenes: a recorder of the reference's DEMSolver setup calls
# (DEMSolver.h:1637-1942) that can be replayed into the GPU scene or the
# oracle, so both see identical inputs in identical order.
# ---------------------------------------------------------------------------
class DEMScene:
    def __init__(self, name: str = ""):
        self.name = name
        self.materials = []      # ("linear"|"hertz"|"bonded", a, b, params)
        self.objects = []        # ordered ball-producing calls and walls
        self.min_bound = np.zeros(3)
        self.max_bound = np.ones(3)
        self.gravity = np.array([0.0, 0.0, -9.81])
        self.dt = 1e-5
        self.hooks = []          # ("const", f[n,3], ramp) | ("damping", c) | ("buoyancy", rho, cd, g, vf, level)
        self.bonds_from_initial_pairs = False
        self.explicit_bonds = None
        self.periodic = (False, False, False)

    def set_periodic_boundary(self, x: bool, y: bool, z: bool):
        """setPeriodicBoundary (DEMSolver.h:1944-1959)."""
        self.periodic = (bool(x), bool(y), bool(z))

    # material rows ------------------------------------------------------
    def set_linear_pair(self, a, b, kn, ks, kr, kt, dn, ds, dr, dt, mus, mur, mut):
        self.materials.append(("linear", a, b, [kn, ks, kr, kt, dn, ds, dr, dt, mus, mur, mut]))

    def set_hertzian_pair(self, a, b, E, G, e, krks, ktks, mus, mur, mut):
        self.materials.append(("hertz", a, b, [E, G, e, krks, ktks, mus, mur, mut]))

    def set_bonded_pair(self, a, b, gamma, E, knks, sigt, C, mu):
        self.materials.append(("bonded", a, b, [gamma, E, knks, sigt, C, mu]))

    # objects --------------------------------------------------------------
    def add_balls(self, pos, vel, w, radius, density, mat, clump=-1):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        b = lambda x: np.broadcast_to(np.asarray(x, dtype=np.float64), (n, 3)).copy()
        s = lambda x, t=np.float64: np.broadcast_to(np.asarray(x, dtype=t), (n,)).copy()
        self.objects.append(("balls", dict(pos=pos, vel=b(vel), w=b(w), r=s(radius), density=s(density),
                                           mat=s(mat, np.int32), clump=s(clump, np.int32))))

    def add_fixed_balls(self, pos, radius, mat, clump=-1):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        s = lambda x, t=np.float64: np.broadcast_to(np.asarray(x, dtype=t), (n,)).copy()
        self.objects.append(("fixed", dict(pos=pos, r=s(radius), mat=s(mat, np.int32), clump=s(clump, np.int32))))

    def add_clump(self, points, radii, centroid, vel, w, mass, inertia, mat):
        self.objects.append(("clump", dict(points=np.asarray(points, dtype=np.float64).reshape(-1, 3),
                                           radii=np.asarray(radii, dtype=np.float64).reshape(-1),
                                           centroid=np.asarray(centroid, dtype=np.float64), vel=np.asarray(vel, float),
                                           w=np.asarray(w, float), mass=float(mass),
                                           inertia=np.asarray(inertia, float).reshape(6), mat=int(mat))))

    def add_fixed_clump(self, points, radii, centroid, mat):
        self.objects.append(("fixed_clump", dict(points=np.asarray(points, dtype=np.float64).reshape(-1, 3),
                                                 radii=np.asarray(radii, dtype=np.float64).reshape(-1),
                                                 centroid=np.asarray(centroid, dtype=np.float64), mat=int(mat))))

    def add_mesh_wall(self, verts, t0, t1, t2, pos, vel, w, mat):
        self.objects.append(("wall", dict(verts=np.asarray(verts, dtype=np.float64).reshape(-1, 3),
                                          t0=np.asarray(t0, dtype=np.int32), t1=np.asarray(t1, dtype=np.int32),
                                          t2=np.asarray(t2, dtype=np.int32), pos=np.asarray(pos, float),
                                          vel=np.asarray(vel, float), w=np.asarray(w, float), mat=int(mat))))

    @property
    def num_balls(self) -> int:
        n = 0
        for kind, o in self.objects:
            if kind in ("balls", "fixed"):
                n += o["pos"].shape[0]
            elif kind in ("clump", "fixed_clump"):
                n += o["points"].shape[0]
        return n

    def initial_positions(self) -> np.ndarray:
        out = []
        for kind, o in self.objects:
            if kind in ("balls", "fixed"):
                out.append(o["pos"])
            elif kind in ("clump", "fixed_clump"):
                out.append(o["points"])
        return np.concatenate(out) if out else np.zeros((0, 3))


def box_mesh(lo, hi):
    """A closed axis-aligned box as 8 vertices and 12 triangles (inward or
    outward orientation does not matter to the sphere-triangle test)."""
    lo = np.asarray(lo, float)
    hi = np.asarray(hi, float)
    v = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]], [lo[0], hi[1], lo[2]],
                  [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]])
    quads = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (1, 5, 6, 2), (2, 6, 7, 3), (3, 7, 4, 0)]
    t0, t1, t2 = [], [], []
    for a, b, c, d in quads:
        t0 += [a, a]
        t1 += [b, c]
        t2 += [c, d]
    return v, np.array(t0), np.array(t1), np.array(t2)


def cylinder_mesh(center, axis, radius, height, n_around, n_height, caps=True):
    """makeCylinderMesh (pointCloudGeneration.h:230-327): rings of n_around
    vertices, two triangles per side quad, optional cap fans."""
    w = np.asarray(axis, float)
    w = w / np.linalg.norm(w)
    tmp = np.array([0.0, 0.0, 1.0]) if abs(w[2]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(tmp, w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    v /= np.linalg.norm(v)
    center = np.asarray(center, float)
    verts = []
    for iz in range(n_height + 1):
        z = (iz / n_height - 0.5) * height
        c = center + z * w
        for ia in range(n_around):
            ang = 2.0 * math.pi * (ia / n_around)
            verts.append(c + radius * (math.cos(ang) * u + math.sin(ang) * v))
    vid = lambda iz, ia: iz * n_around + (ia % n_around)
    t0, t1, t2 = [], [], []
    for iz in range(n_height):
        for ia in range(n_around):
            i00, i01, i10, i11 = vid(iz, ia), vid(iz, ia + 1), vid(iz + 1, ia), vid(iz + 1, ia + 1)
            t0 += [i00, i00]
            t1 += [i10, i11]
            t2 += [i11, i01]
    if caps:
        bc = len(verts)
        verts.append(center - 0.5 * height * w)
        tc = len(verts)
        verts.append(center + 0.5 * height * w)
        for ia in range(n_around):
            t0.append(bc); t1.append(vid(0, ia + 1)); t2.append(vid(0, ia))
        for ia in range(n_around):
            t0.append(tc); t1.append(vid(n_height, ia)); t2.append(vid(n_height, ia + 1))
    return np.array(verts), np.array(t0), np.array(t1), np.array(t2)


def hex_sheet(min_bound, max_bound, spacing, periodic_x=False, periodic_y=False, with_cell=False):
    """generateHexPointCloud (pointCloudGeneration.h:92-195): rows
    dy = a sqrt(3)/2 apart, odd rows shifted by a/2.  A periodic axis is
    snapped to a whole number of steps (ceil), a non-periodic one floors and
    stops before max_bound.  with_cell=True also returns the periodic cell
    (min, max) of makeHexPBC2D."""
    dx = spacing
    dy = spacing * math.sqrt(3.0) * 0.5
    lx = max(0.0, max_bound[0] - min_bound[0])
    ly = max(0.0, max_bound[1] - min_bound[1])
    nx = max(1, int(math.ceil(lx / dx - 1e-12))) if periodic_x else max(1, int(math.floor(lx / dx + 1e-12)))
    ny = max(1, int(math.ceil(ly / dy - 1e-12))) if periodic_y else max(1, int(math.floor(ly / dy + 1e-12)))
    pts = []
    for j in range(ny):
        y = min_bound[1] + j * dy
        if not periodic_y and y >= max_bound[1]:
            break
        shift = 0.5 * dx if (j & 1) else 0.0
        for i in range(nx):
            x = min_bound[0] + i * dx + shift
            if not periodic_x and x >= max_bound[0]:
                break
            pts.append((x, y, min_bound[2]))
    pts = np.array(pts)
    if not with_cell:
        return pts
    cmax = np.array([min_bound[0] + nx * dx if periodic_x else max_bound[0],
                     min_bound[1] + ny * dy if periodic_y else max_bound[1], max_bound[2]], dtype=np.float64)
    return pts, np.asarray(min_bound, dtype=np.float64), cmax


def static_sphere_water_level(R: float, rho_ball: float, rho_water: float) -> float:
    """Height of the water plane, relative to the centre of a floating sphere,
    at which the submerged cap volume fraction (R+h)^2 (2R-h) / (4R^3)
    equals rho_ball/rho_water: 80 bisection steps on [-R, R]
    (iceCylinderInteraction.cpp:5-35)."""
    if not (R > 0.0) or not (rho_water > 0.0) or not (rho_ball >= 0.0):
        return R
    f = rho_ball / rho_water
    if f <= 0.0:
        return -R
    if f >= 1.0:
        return R
    lo, hi = -R, R
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        b = R + mid
        if (b * b) * (2.0 * R - mid) / (4.0 * R * R * R) < f:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def ice_cylinder() -> DEMScene:
    """iceCylinderInteraction.cpp: a 100 x 116 hex sheet of bonded ice balls
    (r = 0.1 m, 900 kg/m^3, periodic snap in y) drifting at 0.4 m/s onto a
    36 x 36 triangulated cylinder (R = 1 m, H = 6 m), buoyancy + drag hook,
    linear contacts with Coulomb friction, bonds breaking at 0.5 MPa,
    dt = t_c / 50 with t_c = 2e-4 s."""
    tc, r, den = 0.0002, 0.1, 900.0
    m = r ** 3 * math.pi * 4.0 / 3.0 * den
    kn = 0.5 * m * math.pi ** 2 / tc / tc
    E = kn * 2.0 * r / (r * r * math.pi)
    s = DEMScene("ice_cylinder")
    s.set_bonded_pair(0, 0, 1.0, E, 2.6, 0.5e6, 0.5e6, 0.1)
    s.set_linear_pair(0, 0, kn, kn / 2.6, 0.0, 0.0, 0.1, 0.1, 0.0, 0.0, 0.1, 0.0, 0.0)
    s.set_linear_pair(0, 1, kn, kn / 2.6, 0.0, 0.0, 0.1, 0.1, 0.0, 0.0, 0.1, 0.0, 0.0)
    pts, cmin, cmax = hex_sheet([0.0, 0.0, 0.0], [200 * r, 200 * r, 0.0], 2 * r, False, True, with_cell=True)
    s.add_balls(pts, [-0.4, 0.0, 0.0], [0.0, 0.0, 0.0], r, den, 0)
    v, t0, t1, t2 = cylinder_mesh([0.0, 0.0, 0.0], [0.0, 0.0, 1.0], 1.0, 6.0, 36, 36)
    s.add_mesh_wall(v, t0, t1, t2, [-1.1, 100 * r, 0.0], [0, 0, 0], [0, 0, 0], 1)
    level = static_sphere_water_level(r, den, 1000.0)
    s.hooks = [("buoyancy", 1000.0, 0.1, [0.0, 0.0, -9.81], [-0.4, 0.0, 0.0], level)]
    s.bonds_from_initial_pairs = True
    s.min_bound = np.array([-cmax[0], cmin[1], -3.0])
    s.max_bound = np.array([cmax[0], cmax[1], 3.0])
    s.gravity = np.array([0.0, 0.0, -9.81])
    s.dt = tc / 50.0
    return s


def periodic_ice_sheet(nx: int = 24, ny: int = 20, roll_rows: int = 0, speed: float = 0.3,
                       uniform_velocity=None, bonded: bool = True, gap: float = 0.0) -> DEMScene:
    """The ice sheet of iceCylinderInteraction.cpp (hex rows of r = 0.1 m
    balls at 900 kg/m^3, bonds from the initial pairs breaking at 0.5 MPa,
    linear contacts, dt = t_c / 50) on its own y-periodic cell: the rows of
    generateHexPointCloud with the y axis snapped to ny rows (ny even, so the
    row pattern continues across the face) and setPeriodicBoundary(false,
    true, false) on y in [0, ny dy).  No walls, no gravity.

    Velocities are periodic in y: v = speed (sin k y, 0.7 cos k y + 0.3,
    0.2 sin 2 k y), k = 2 pi / Ly, evaluated at each ball's lattice row, or
    uniform_velocity for every ball.  roll_rows (even) moves every ball by
    that many rows along y, wrapped into the cell, with the velocity it had:
    the same periodic system translated, balls in the same insertion order.
    gap > 0 spaces the lattice at 2 r (1 + gap): touching balls (gap 0, as
    the application) start at zero overlap, where the rounding of the
    overlap decides whether a damped contact acts, so such a sheet follows
    the rounding of its positions chaotically."""
    if ny % 2 or roll_rows % 2:
        raise ValueError("ny and roll_rows must be even (hex row parity)")
    tc, r, den = 0.0002, 0.1, 900.0
    m = r ** 3 * math.pi * 4.0 / 3.0 * den
    kn = 0.5 * m * math.pi ** 2 / tc / tc
    E = kn * 2.0 * r / (r * r * math.pi)
    sp = 2 * r * (1.0 + gap)
    dy = sp * math.sqrt(3.0) * 0.5
    s = DEMScene("periodic_ice_sheet")
    if bonded:
        s.set_bonded_pair(0, 0, 1.0, E, 2.6, 0.5e6, 0.5e6, 0.1)
        s.bonds_from_initial_pairs = True
    s.set_linear_pair(0, 0, kn, kn / 2.6, 0.0, 0.0, 0.1, 0.1, 0.0, 0.0, 0.1, 0.0, 0.0)
    pts, cmin, cmax = hex_sheet([0.0, 0.0, 0.0], [sp * nx, ny * dy, 0.0], sp, False, True, with_cell=True)
    ly = cmax[1] - cmin[1]
    k = 2.0 * math.pi / ly
    y0 = pts[:, 1].copy()
    if uniform_velocity is not None:
        vel = np.broadcast_to(np.asarray(uniform_velocity, dtype=np.float64), pts.shape).copy()
    else:
        vel = speed * np.stack([np.sin(k * y0), 0.7 * np.cos(k * y0) + 0.3, 0.2 * np.sin(2 * k * y0)], axis=1)
    if roll_rows:
        y = y0 + roll_rows * dy
        y = np.where(y >= cmax[1], y - ly, y)
        pts = pts.copy()
        pts[:, 1] = y
    s.add_balls(pts, vel, [0.0, 0.0, 0.0], r, den, 0)
    s.set_periodic_boundary(False, True, False)
    # x and z: cells strictly above 2r (the lattice never sits on cell faces);
    # y: the snapped period, whose cells are above 2r since dy < 2r
    lo, hi = half_cell_bounds([-1.0, cmin[1], -1.0], [sp * nx + 1.0, cmax[1], 1.0], sp)
    s.min_bound = lo
    s.max_bound = np.array([hi[0], cmax[1], hi[2]])
    s.gravity = np.zeros(3)
    s.dt = tc / 50.0
    return s


def elastic_bounce() -> DEMScene:
    """elasticBounce.cpp:3-58: one ball above a two-triangle plane, linear
    spring with contact time 5 ms, dt = t_c / 50."""
    s = DEMScene("elastic_bounce")
    r, tc = 0.01, 0.005
    s.add_balls([0.0, 0.0, 0.0], [0, 0, 0], [0, 0, 0], r, 2000, 0)
    verts = np.array([[-1.0, -1.0, -1.0], [1.0, -1.0, -1.0], [1.0, 1.0, -1.0], [-1.0, 1.0, -1.0]])
    s.add_mesh_wall(verts, [0, 1], [1, 2], [3, 3], [0, 0, 0], [0, 0, 0], [0, 0, 0], 0)
    mass = 4.0 / 3.0 * r ** 3 * math.pi * 2000
    s.set_linear_pair(0, 0, 0.5 * mass * math.pi ** 2 / tc / tc, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
    s.min_bound = np.array([-1.0, -1.0, -1.0])
    s.max_bound = np.array([1.0, 1.0, 1.0])
    s.gravity = np.array([0.0, 0.0, -9.81])
    s.dt = tc / 50.0
    return s


def bonded_beam() -> DEMScene:
    """bondedParticleBeam.cpp: 11-ball cantilever, fixed first ball, bonds from
    the initial pair list, tip load 1e5 N ramped over 1 s, global damping 0.1."""
    s = DEMScene("bonded_beam")
    L, nb = 4.0, 10
    r = L / nb / 2.0
    s.set_bonded_pair(0, 0, 1.0, 200e9, 2.6, 0.0, 0.0, 0.0)
    s.add_fixed_balls([0.0, 0.0, 0.0], r, 0)
    s.add_balls([[(i + 1) * 2.0 * r, 0.0, 0.0] for i in range(nb)], [0, 0, 0], [0, 0, 0], r, 7800, 0)
    f = np.zeros((nb + 1, 3))
    f[nb] = [0, 0, 100e3]
    s.hooks = [("const", f, 1.0), ("damping", 0.1)]
    s.bonds_from_initial_pairs = True
    s.min_bound = np.array([0.0, -r, -r])
    s.max_bound = np.array([L, r, r])
    s.gravity = np.zeros(3)
    s.dt = 1e-5
    return s


def granular_collapse(n_side: int = 16, seed: int = 20261202) -> DEMScene:
    """DEM granular collapse: n_side^3 spheres r = 5 mm, rho = 2500 on an
    11 mm lattice with seeded U(-1 mm, +1 mm) jitter per axis, in a 0.3 m box
    of 12 triangles, Hertz-Mindlin E = 1e7, nu = 0.3 (E* = E/(2(1-nu^2)),
    G* = E/(4(2-nu)(1+nu))), e = 0.5, mu_s = 0.3, dt = 1e-5.  A jitter that
    would overlap an earlier sphere is redrawn (up to 64 times, then zero)."""
    s = DEMScene("granular_collapse")
    r, sp, L = 0.005, 0.011, 0.3
    E, nu = 1e7, 0.3
    Es = E / (2 * (1 - nu * nu))
    Gs = E / (4 * (2 - nu) * (1 + nu))
    s.set_hertzian_pair(0, 0, Es, Gs, 0.5, 0.0, 0.0, 0.3, 0.0, 0.0)
    start = np.array([(L - (n_side - 1) * sp) / 2, (L - (n_side - 1) * sp) / 2, 0.01])
    g = splitmix64(seed)
    u = lambda: (next(g) >> 11) * (1.0 / 9007199254740992.0)
    base = np.array([[i, j, k] for k in range(n_side) for j in range(n_side) for i in range(n_side)], float) * sp + start
    pos = np.empty_like(base)
    grid = {}
    for idx, p0 in enumerate(base):
        cell = tuple(np.floor(p0 / sp).astype(int))
        for attempt in range(65):
            jit = np.array([u(), u(), u()]) * 0.002 - 0.001 if attempt < 64 else np.zeros(3)
            p = p0 + jit
            ok = True
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dxx in (-1, 0, 1):
                        q = grid.get((cell[0] + dxx, cell[1] + dy, cell[2] + dz))
                        if q is not None and np.sum((pos[q] - p) ** 2) < (2 * r) ** 2:
                            ok = False
            if ok:
                break
        pos[idx] = p
        grid[cell] = idx
    s.add_balls(pos, [0, 0, 0], [0, 0, 0], r, 2500, 0)
    v, t0, t1, t2 = box_mesh([0, 0, 0], [L, L, L])
    s.add_mesh_wall(v, t0, t1, t2, [0, 0, 0], [0, 0, 0], [0, 0, 0], 0)
    s.min_bound = np.zeros(3)
    s.max_bound = np.array([L, L, L])
    s.dt = 1e-5
    return s


# ---------------------------------------------------------------------------
# Coupled DEM-SPH scenes
# ---------------------------------------------------------------------------
def half_cell_bounds(min_bound, max_bound, cell: float) -> tuple[np.ndarray, np.ndarray]:
    """Grows max_bound so that every axis spans a whole number m of cells
    plus one half.  The grid then takes gridSize = m and the cell size
    length / m = cell (1 + 1/(2m)) (boundary.h:85-91), strictly above `cell`
    by far more than rounding: two balls at most `cell` apart along an axis
    land at most one cell apart there, wherever the lattice sits relative to
    the cell faces.  Without the margin, a lattice of spacing equal to the
    cell size can put touching neighbours two cells apart through the
    rounding of (x - min) / cellSize, out of each other's 27-cell stencil."""
    mn = np.asarray(min_bound, dtype=np.float64)
    mx = np.asarray(max_bound, dtype=np.float64)
    m = np.maximum(1.0, np.ceil((mx - mn) / cell - 0.5))
    return mn.copy(), mn + (m + 0.5) * cell


def lattice(nx: int, ny: int, nz: int, dx: float, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Simple cubic lattice, x fastest."""
    z, y, x = np.meshgrid(np.arange(nz) * dx, np.arange(ny) * dx, np.arange(nx) * dx, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1) + np.asarray(origin, dtype=np.float64)


def column_with_ball(velocity=None, gravity=(0.0, 0.0, -9.81)):
    """A 17 x 17 x 25 fluid lattice at dx = 1 cm (rho0 = 1000, c0 = 20,
    linear EOS) whose centre particle is replaced by a fixed ball of the same
    volume, r = dx (3/(4 pi))^(1/3).  Returns (spec, drho, ball scene, r):
    drho is the hydrostatic density perturbation rho0 |g| (H - z) / c0^2 of
    a column at rest."""
    dx, rho0, c0 = 0.01, 1000.0, 20.0
    pos = lattice(17, 17, 25, dx)
    centre = np.array([8, 8, 12]) * dx
    pos = pos[np.linalg.norm(pos - centre, axis=1) > 1e-9]
    g = np.asarray(gravity, dtype=np.float64)
    H = pos[:, 2].max() + 0.5 * dx
    vel = None if velocity is None else np.tile(np.asarray(velocity, float), (pos.shape[0], 1))
    spec = SPHSceneSpec(fluid_pos=pos, dummy_pos=np.zeros((0, 3)), dx=dx, rho0=rho0, c0=c0, nu=1e-6,
                        min_bound=np.array([-0.05, -0.05, -0.05]), max_bound=np.array([0.22, 0.22, 0.3]),
                        gravity=g, dt=1e-6, fluid_vel=vel)
    drho = rho0 * abs(g[2]) * (H - pos[:, 2]) / c0 ** 2
    r = dx * (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)
    d = DEMScene("ball in a column")
    d.add_fixed_balls(centre, r, 0)
    d.set_hertzian_pair(0, 0, 1e7, 1e6, 0.5, 0.0, 0.0, 0.3, 0.0, 0.0)
    d.min_bound, d.max_bound, d.gravity = spec.min_bound, spec.max_bound, g
    return spec, drho, d, r


def balls_in_dam_break():
    """small_dam_break() with twelve 6 mm balls (2500 kg/m^3) sinking at
    0.1 m/s inside the water column."""
    spec = small_dam_break()
    lo = spec.fluid_pos.min(axis=0)
    d = DEMScene("balls in water")
    centres = lattice(3, 2, 2, 0.025, origin=lo + np.array([0.02, 0.015, 0.03]))
    d.add_balls(centres, [0.0, 0.0, -0.1], [0.0, 0.0, 0.0], 0.006, 2500.0, 0)
    d.set_hertzian_pair(0, 0, 1e7, 1e6, 0.5, 0.0, 0.0, 0.3, 0.0, 0.0)
    d.min_bound, d.max_bound, d.gravity = spec.min_bound, spec.max_bound, spec.gravity
    return spec, d


# ---------------------------------------------------------------------------
# Batched serving mix: independent scenes of log-uniform size
# ---------------------------------------------------------------------------
def mix_wcsph(n: int, dx: float = 0.01, eos: int = 1) -> SPHSceneSpec:
    """A dam break of about n fluid particles: a 2m x m x m column (2 m^3
    ~ n) at spacing dx in a 4m x (m+1) x 1.5m open tank with its dummy
    shell, Tait EOS, dt from the CFL rule of dam_break()."""
    m = max(4, int(round((n / 2.0) ** (1.0 / 3.0))))
    nx, ny, nz = 2 * m, m, m
    spec = dam_break(dx=dx, column=(nx * dx, ny * dx, nz * dx), tank=(2 * nx * dx, (ny + 1) * dx, 1.5 * nz * dx),
                     eos=eos, lattice_counts=(nx, ny, nz))
    spec.name = f"mix_wcsph_{spec.nf}"
    return spec


def _hertz_box(name: str, L, H) -> DEMScene:
    s = DEMScene(name)
    E, nu = 1e7, 0.3
    s.set_hertzian_pair(0, 0, E / (2 * (1 - nu * nu)), E / (4 * (2 - nu) * (1 + nu)), 0.5, 0.0, 0.0, 0.3, 0.0, 0.0)
    v, t0, t1, t2 = box_mesh([0, 0, 0], [L[0], L[1], H])
    s.add_mesh_wall(v, t0, t1, t2, [0, 0, 0], [0, 0, 0], [0, 0, 0], 0)
    s.min_bound = np.zeros(3)
    s.max_bound = np.array([L[0], L[1], H])
    s.dt = 1e-5
    return s


def mix_dem(n: int, seed: int, r: float = 0.005) -> DEMScene:
    """n Hertz-Mindlin spheres (r = 5 mm, rho = 2500, the granular_collapse
    material) on a 2.2 r lattice, m x m wide and as tall as needed, with
    seeded jitter of at most 0.09 r per axis (never an overlap: neighbours
    stay at least 2.02 r apart along the lattice axis), settling in a closed
    box twice the pile's height, dt = 1e-5."""
    sp = 2.2 * r
    m = max(2, int(round(n ** (1.0 / 3.0))))
    mz = -(-n // (m * m))
    base = lattice(m, m, mz, sp, origin=(1.5 * sp, 1.5 * sp, 1.5 * sp))[:n]
    jit = (uniform01(seed, 3 * n).reshape(n, 3) * 2.0 - 1.0) * 0.09 * r
    L = (m + 2) * sp
    s = _hertz_box(f"mix_dem_{n}", (L, L), 2.0 * (mz + 2) * sp)
    s.add_balls(base + jit, [0, 0, 0], [0, 0, 0], r, 2500.0, 0)
    return s


def mix_coupled(n: int, seed: int, dx: float = 0.01):
    """A dam break of about 0.95 n fluid particles (mix_wcsph) with n/20
    balls (r = 0.6 dx, rho = 2500, Hertz-Mindlin) on a lattice of 2.7 dx
    inside the column, seeded jitter of 0.1 dx per axis, and the tank as a
    closed DEM box.  Returns (fluid spec, DEM scene); both scenes share the
    fluid's bounds and time step."""
    spec = mix_wcsph(int(0.95 * n), dx)
    nb = max(1, n // 20)
    lo, hi = spec.fluid_pos.min(axis=0) + 1.5 * dx, spec.fluid_pos.max(axis=0) - 1.5 * dx
    sp = 2.7 * dx
    cnt = np.maximum(1, np.floor((hi - lo) / sp).astype(int) + 1)
    pts = lattice(int(cnt[0]), int(cnt[1]), int(cnt[2]), sp, origin=lo)[:nb]
    pts = pts + (uniform01(seed, pts.size).reshape(-1, 3) * 2.0 - 1.0) * 0.1 * dx
    d = DEMScene(f"mix_coupled_{spec.nf}_{pts.shape[0]}")
    d.set_hertzian_pair(0, 0, 1e7, 1e6, 0.5, 0.0, 0.0, 0.3, 0.0, 0.0)
    d.add_balls(pts, [0, 0, 0], [0, 0, 0], 0.6 * dx, 2500.0, 0)
    # the tank of mix_wcsph as a closed DEM box (the fluid's dummies do not act on balls)
    top = spec.fluid_pos.max(axis=0) + dx
    v, t0, t1, t2 = box_mesh([0, 0, 0], [2.0 * top[0], top[1] + dx, 1.5 * top[2]])
    d.add_mesh_wall(v, t0, t1, t2, [0, 0, 0], [0, 0, 0], [0, 0, 0], 0)
    d.min_bound, d.max_bound, d.gravity = spec.min_bound, spec.max_bound, spec.gravity
    d.dt = spec.dt
    spec.name = d.name
    return spec, d


@dataclass
class MixJob:
    kind: str        # "wcsph" | "dem" | "coupled"
    n: int           # requested particles (fluid + balls)
    steps: int
    arrival: float   # seconds after the start of serving
    seed: int


def serving_mix(count: int = 256, seed: int = 20261202, n_range=(2.0e3, 2.0e5), steps_range=(1000, 20000),
                mean_interarrival: float = 0.05, fractions=(0.5, 0.3, 0.2)) -> list[MixJob]:
    """The batched-serving job stream: sizes exp(U(ln a, ln b)), kind drawn
    with the given DEM / WCSPH / coupled fractions, steps U{lo..hi},
    Poisson arrivals (exponential gaps of the given mean)."""
    u = uniform01(seed, 4 * count).reshape(count, 4)
    jobs, t = [], 0.0
    la, lb = math.log(n_range[0]), math.log(n_range[1])
    for k in range(count):
        n = int(round(math.exp(la + u[k, 0] * (lb - la))))
        c = u[k, 1]
        kind = "dem" if c < fractions[0] else ("wcsph" if c < fractions[0] + fractions[1] else "coupled")
        steps = steps_range[0] + int(u[k, 2] * (steps_range[1] - steps_range[0] + 1))
        steps = min(steps, steps_range[1])
        if k > 0:
            t += -mean_interarrival * math.log(max(1.0 - u[k, 3], 1e-300))
        jobs.append(MixJob(kind, n, steps, t, seed + 7919 * (k + 1)))
    return jobs


# ---------------------------------------------------------------------------
# Bonded cantilever with falling tetrahedral clumps
# ---------------------------------------------------------------------------
def quat_uniform(u1: float, u2: float, u3: float) -> np.ndarray:
    """Uniformly distributed unit quaternion (w, x, y, z) from three uniforms
    (Shoemake, Graphics Gems III, "Uniform random rotations")."""
    a, b = math.sqrt(1.0 - u1), math.sqrt(u1)
    t1, t2 = 2.0 * math.pi * u2, 2.0 * math.pi * u3
    x, y, z, w = a * math.sin(t1), a * math.cos(t1), b * math.sin(t2), b * math.cos(t2)
    return np.array([w, x, y, z])


def quat_rotate(q: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Rotates the rows of v by the unit quaternion q = (w, x, y, z)."""
    w, u = q[0], q[1:]
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def cantilever_with_clumps(nx: int = 128, ny: int = 16, nz: int = 16, r: float = 0.002, n_clumps: int = 20000,
                           E: float = 1e9, tip_force: float = 10.0, ramp_steps: int = 10000,
                           seed: int = 20261202) -> DEMScene:
    """Bonded DEM cantilever with falling clumps.

    * The beam is an nx x ny x nz simple-cubic lattice of touching spheres (spacing 2r, rho 2500). Its x = 0
      layer is fixed.
    * Parallel bonds come from the initial pair list: E = 1e9, gamma = 1, kn/ks = 2.6, sigma_t = C = 0 (never
      breaking).
    * The tip load is split evenly over the x = max layer and ramped linearly over `ramp_steps` steps by the hook.
    * Above the beam are n_clumps rigid tetrahedral 4-sphere clumps (pebble radius r, pebble centres r from the
      clump centre), with seeded Shoemake-uniform orientations on a 3r-clearance lattice over the beam's
      footprint.
    * Hertz-Mindlin contacts everywhere (E* and G* of E, nu = 0.3), e = 0.5, mu = 0.3.
    * Gravity -z.  dt = 5e-7 s, 1/65 of the bond oscillation period 2 pi sqrt(m / (E pi r / 2)).
    """
    s = DEMScene("cantilever_clumps")
    rho, nu = 2500.0, 0.3
    Es, Gs = E / (2 * (1 - nu * nu)), E / (4 * (2 - nu) * (1 + nu))
    s.set_bonded_pair(0, 0, 1.0, E, 2.6, 0.0, 0.0, 0.0)
    for a, b in ((0, 0), (0, 1), (1, 1)):
        s.set_hertzian_pair(a, b, Es, Gs, 0.5, 0.0, 0.0, 0.3, 0.0, 0.0)
    beam = lattice(nx, ny, nz, 2.0 * r)
    fixed = beam[:, 0] < 0.5 * r
    s.add_fixed_balls(beam[fixed], r, 0)
    s.add_balls(beam[~fixed], [0, 0, 0], [0, 0, 0], r, rho, 0)
    n_beam = beam.shape[0]
    tip = beam[~fixed][:, 0] > (nx - 1.5) * 2.0 * r
    # clumps
    d = r
    tet = np.array([[1.0, 1.0, 1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]]) / math.sqrt(3.0) * d
    mp = rho * 4.0 / 3.0 * math.pi * r ** 3
    I = 4.0 * (0.4 * mp * r * r + mp * (2.0 / 3.0) * d * d)
    pitch = 2.0 * (d + r) + 3.0 * r
    cx = max(1, int((nx * 2.0 * r) // pitch))
    cy = max(1, int((ny * 2.0 * r) // pitch))
    u = uniform01(seed, 3 * n_clumps).reshape(n_clumps, 3)
    z0 = nz * 2.0 * r + pitch
    for k in range(n_clumps):
        i, j, l = k % cx, (k // cx) % cy, k // (cx * cy)
        c = np.array([(i + 0.5) * pitch, (j + 0.5) * pitch, z0 + l * pitch])
        q = quat_uniform(*u[k])
        s.add_clump(quat_rotate(q, tet) + c, np.full(4, r), c, [0, 0, 0], [0, 0, 0], 4 * mp, [I, I, I, 0, 0, 0], 1)
    n_total = n_beam + 4 * n_clumps
    f = np.zeros((n_total, 3))
    # insertion order: fixed layer, free beam balls, then the clumps' pebbles
    nfix = int(fixed.sum())
    f[nfix + np.flatnonzero(tip), 2] = -tip_force / int(tip.sum())
    dt = 5e-7
    s.hooks = [("const", f, ramp_steps * dt)]
    s.bonds_from_initial_pairs = True
    top = z0 + (n_clumps // (cx * cy) + 1) * pitch
    s.min_bound, s.max_bound = half_cell_bounds([-4.0 * r, -4.0 * r, -0.5],
                                                [nx * 2.0 * r + 4.0 * r, ny * 2.0 * r + 4.0 * r, top], 2.0 * r)
    s.gravity = np.array([0.0, 0.0, -9.81])
    s.dt = dt
    return s


# ---------------------------------------------------------------------------
# Coupled ice sheet, current and cylinder
# ---------------------------------------------------------------------------
def ice_cylinder_coupled(nx: int = 500, ny: int = 200, nz: int = 10, dx: float = 0.004, R: float = 0.1,
                         n_around: int = 32, n_height: int = 32, current: float = 0.5, ice_cover: float = 0.93,
                         eos: int = 1):
    """Coupled DEM-SPH ice sheet pushed by a current onto a rigid cylinder.

    * Water: an nx x ny x nz lattice at spacing dx filling an open tank (3-layer dummy shell), moving at
      `current` m/s along +x, Tait EOS, c0 = 10 sqrt(2 g H).
    * Cylinder: vertical, radius R, at the tank centre; a DEM mesh wall of 2 n_around n_height triangles, and,
      for the fluid, dummy particles in the annulus R - 2.5 dx <= r <= R - 0.5 dx.  Fluid within R + 0.5 dx of
      its axis is removed.
    * Ice: two stacked hex layers of bonded balls (r = dx, rho = 917) on the water surface over the first
      `ice_cover` of the tank's length, drifting with the current.  Bonds come from the initial pairs
      (sigma_t = C = 0.5 MPa, mu = 0.1, kn/ks = 2.6, as iceCylinderInteraction.cpp:85).  Linear contacts use
      kn from a 50 dt contact time.
    * The cylinder is the only DEM wall: the triangle grid takes its cell from the largest triangle
      (DEMSolver.h:432-449), so a tank box of a few large triangles would make every ball test every facet.

    dt = 5e-6 s.  Returns (fluid spec, DEM scene); the DEM grid spans the fluid's bounds (see
    half_cell_bounds).
    """
    g, rho0, dt = 9.81, 1000.0, 5e-6
    Lx, Ly = nx * dx, ny * dx
    H = nz * dx
    tank = (Lx, Ly + dx, 2.5 * H + 4 * dx)
    fluid = lattice(nx, ny, nz, dx)
    cxy = np.array([0.5 * Lx, 0.5 * Ly])
    rf = np.linalg.norm(fluid[:, :2] - cxy, axis=1)
    fluid = fluid[rf >= R + 0.5 * dx]
    shell, mn, mx = tank_dummies(tank, dx)
    # cylinder annulus on the same lattice, from the floor shell to the tank top
    ca = int(math.ceil(R / dx)) + 1
    ix = np.arange(int(cxy[0] / dx) - ca, int(cxy[0] / dx) + ca + 1)
    iy = np.arange(int(cxy[1] / dx) - ca, int(cxy[1] / dx) + ca + 1)
    iz = np.arange(-3, int(tank[2] / dx) + 1)
    zz, yy, xx = np.meshgrid(iz * dx, iy * dx, ix * dx, indexing="ij")
    cand = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)
    rc = np.linalg.norm(cand[:, :2] - cxy, axis=1)
    ring = cand[(rc >= R - 2.5 * dx) & (rc <= R - 0.5 * dx)]
    dummies = np.concatenate([shell, ring])
    c0 = 10.0 * math.sqrt(2.0 * g * H)
    vel = np.tile([current, 0.0, 0.0], (fluid.shape[0], 1))
    spec = SPHSceneSpec(fluid_pos=fluid, dummy_pos=dummies, dx=dx, rho0=rho0, c0=c0, nu=1.0e-6, min_bound=mn,
                        max_bound=mx, gravity=np.array([0.0, 0.0, -g]), dt=dt, eos=eos, fluid_vel=vel,
                        name="ice_cylinder_coupled")
    # ice
    r, den = dx, 917.0
    tc = 50.0 * dt
    m = den * 4.0 / 3.0 * math.pi * r ** 3
    kn = 0.5 * m * math.pi ** 2 / tc / tc
    E = kn * 2.0 * r / (math.pi * r * r)
    d = DEMScene("ice_cylinder_coupled")
    d.set_bonded_pair(0, 0, 1.0, E, 2.6, 0.5e6, 0.5e6, 0.1)
    d.set_linear_pair(0, 0, kn, kn / 2.6, 0.0, 0.0, 0.1, 0.1, 0.0, 0.0, 0.1, 0.0, 0.0)
    d.set_linear_pair(0, 1, kn, kn / 2.6, 0.0, 0.0, 0.1, 0.1, 0.0, 0.0, 0.1, 0.0, 0.0)
    ztop = (nz - 1) * dx
    lo = [2.0 * r, 2.0 * r, ztop + r]
    hi = [ice_cover * Lx, Ly - 2.0 * r, ztop + r]
    l1 = hex_sheet(lo, hi, 2.0 * r)
    l2 = hex_sheet([lo[0] + r, lo[1] + r / math.sqrt(3.0), ztop + r + 2.0 * r * math.sqrt(2.0 / 3.0)],
                   [hi[0], hi[1], 0.0], 2.0 * r)
    l2[:, 2] = ztop + r + 2.0 * r * math.sqrt(2.0 / 3.0)
    ice = np.concatenate([l1, l2])
    ice = ice[np.linalg.norm(ice[:, :2] - cxy, axis=1) >= R + 2.0 * r]
    d.add_balls(ice, [current, 0.0, 0.0], [0.0, 0.0, 0.0], r, den, 0)
    v, t0, t1, t2 = cylinder_mesh([cxy[0], cxy[1], 0.5 * tank[2]], [0.0, 0.0, 1.0], R, tank[2], n_around, n_height,
                                  caps=False)
    d.add_mesh_wall(v, t0, t1, t2, [0, 0, 0], [0, 0, 0], [0, 0, 0], 1)
    d.bonds_from_initial_pairs = True
    # the ball grid spans the tank with half a cell to spare per axis, so the
    # hex layers' touching neighbours never round two cells apart
    d.min_bound, d.max_bound = half_cell_bounds(spec.min_bound, spec.max_bound, 2.0 * r)
    d.gravity = spec.gravity
    d.dt = dt
    return spec, d



def two_phase_tank(eos: int = 0) -> SPHSceneSpec:
    """Two fluid phases side by side in a small tank: 8 x 6 x 8 particles at
    dx = 1 cm (1000 kg/m^3, nu = 1e-6) and, 1 cm to their right, 6 x 5 x 6 at
    1.25 cm (800 kg/m^3, nu = 1e-3), with a 3-layer dummy shell at 1 cm.
    Per-pair cuts 2 max(h_A, h_B) and averaged smoothing lengths come into
    play across the interface.  dt = 0.2 h_min / c0."""
    dx1, dx2 = 0.01, 0.0125
    a = lattice(8, 6, 8, dx1)
    b = lattice(6, 5, 6, dx2, origin=(8 * dx1 + dx1, 0.0, 0.0))
    pos = np.concatenate([a, b])
    na, nb = a.shape[0], b.shape[0]
    tank = (0.3, 6 * dx1 + dx1, 0.15)
    dummies, mn, mx = tank_dummies(tank, dx1)
    c0 = 10.0 * math.sqrt(2.0 * 9.81 * 8 * dx1)
    spec = SPHSceneSpec(fluid_pos=pos, dummy_pos=dummies, dx=dx1, rho0=1000.0, c0=c0, nu=1e-6, min_bound=mn,
                        max_bound=mx, gravity=np.array([0.0, 0.0, -9.81]), dt=0.2 * 1.3 * dx1 / c0, eos=eos,
                        name="two_phase_tank")
    spec.fluid_dx = np.concatenate([np.full(na, dx1), np.full(nb, dx2)])
    spec.fluid_rho0 = np.concatenate([np.full(na, 1000.0), np.full(nb, 800.0)])
    spec.fluid_nu = np.concatenate([np.full(na, 1e-6), np.full(nb, 1e-3)])
    spec.fluid_c0 = np.full(na + nb, c0)
    return spec


# ---------------------------------------------------------------------------
# Closed-form scenes for bond breakage, moving walls and clump force sums
# ---------------------------------------------------------------------------
# Rates of the bonded-pair load cases: per force evaluation (spring time
# dt/2) the governing stress grows by inc = 5e4 Pa; the break evaluation is
# the first k with k * inc above the threshold.
BOND_LOAD_CASES = {
    # name: (v0 of ball 0, omega0 of ball 0, sigma_t / inc, C / inc, mu, governing stress)
    "tension": ((0.1, 0.0, 0.0), (0.0, 0.0, 0.0), 10.5, 0.0, 0.0, "sigma"),
    "cohesion": ((0.1, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0, 25.0, 2.0, "sigma"),
    "bending": ((0.0, 0.0, 0.0), (0.0, 0.0, 10.0), 8.5, 0.0, 0.0, "sigma"),
    "torsion": ((0.0, 0.0, 0.0), (26.0, 0.0, 0.0), 0.0, 10.5, 0.0, "tau"),
    "shear": ((0.0, 0.26, 0.0), (0.0, 0.0, 0.0), 0.0, 6.5, 0.0, "tau"),
}


def bonded_pair_load(case: str, dt: float = 1e-5):
    """Two touching bonded balls driven at constant relative rates until the
    bond breaks (ParallelBondedContact, contactKernel.cuh:4-62).

    * Balls of r = 1 cm at x = +-r, ball 0 at +r, so the bond normal is +x.
      Ball 0 moves with (v0, omega0), ball 1 with (-v0, -omega0).  A density of
      1e15 kg/m^3 keeps the rates fixed: the bond changes them by about 1e-11
      relative over the test.  No gravity, no contact rows (zero contact force,
      the reference's Hertz defaults), bonds from the initial pairs.
    * Bond: E = 1e9 Pa, gamma = 1 (R_b = r, A = pi r^2, I = pi r^4 / 4, J = 2 I),
      kn/ks = 2.6, so kn'' = E / 2r = 5e10 Pa/m and ks'' = kn'' / 2.6.
    * Each force evaluation integrates the bond over dt/2, so the governing
      stress after k evaluations is k * inc with inc = 5e4 Pa:
        tension   sigma = 2 v (dt/2) kn''           (normal separation)
        cohesion  the same motion with sigma_t = 0: tau = 0 exceeds C - mu sigma
                  once sigma > C / mu
        bending   sigma = 2 w (dt/2) kn'' R_b       (omega about z, across the bond)
        torsion   tau = 2 w (dt/2) ks'' R_b         (omega about the bond normal)
        shear     tau = 2 v (dt/2) ks''             (sliding along y; the normal turns
                  by about 1e-3 rad, a 1e-6 relative change)

    Returns (scene, inc, k_break, governing "sigma" or "tau")."""
    v0, w0, st, c, mu, gov = BOND_LOAD_CASES[case]
    r, E, knks = 0.01, 1e9, 2.6
    h = 0.5 * dt
    kn = E / (2 * r)
    ks = kn / knks
    inc = {"tension": 2 * v0[0] * h * kn, "cohesion": 2 * v0[0] * h * kn, "bending": 2 * w0[2] * h * kn * r,
           "torsion": 2 * w0[0] * h * ks * r, "shear": 2 * v0[1] * h * ks}[case]
    thresh = st if st > 0 else c / mu if mu > 0 else c
    k_break = int(math.floor(thresh)) + 1
    s = DEMScene(f"bond_{case}")
    s.set_bonded_pair(0, 0, 1.0, E, knks, st * inc, c * inc, mu)
    s.add_balls([[r, 0.0, 0.0], [-r, 0.0, 0.0]], [v0, [-x for x in v0]], [w0, [-x for x in w0]], r, 1e15, 0)
    s.bonds_from_initial_pairs = True
    s.min_bound, s.max_bound = -0.1 * np.ones(3), 0.1 * np.ones(3)
    s.gravity = np.zeros(3)
    s.dt = dt
    return s, inc, k_break, gov


def bonded_strip_pull(nx: int = 24, ny: int = 4, r: float = 0.005, force: float = 40.0, dt: float = 2e-6):
    """A bonded nx x ny strip of touching balls in the z = 0 plane, its x = 0
    column fixed, pulled at its free end by a constant force (split over the
    end column) in +x and -y, so that bonds break progressively near the
    root under combined tension, bending and shear.  E = 1e8 Pa, kn/ks = 2.6,
    sigma_t = C = 2e6 Pa, mu = 0.3; linear contact rows (kn = 1e5 N/m)
    between the balls.  No gravity.  The GPU and the oracle should break the
    same bonds at the same steps in FP64."""
    s = DEMScene("bonded_strip_pull")
    s.set_bonded_pair(0, 0, 1.0, 1e8, 2.6, 2e6, 2e6, 0.3)
    s.set_linear_pair(0, 0, 1e5, 1e5 / 2.6, 0.0, 0.0, 0.1, 0.1, 0.0, 0.0, 0.3, 0.0, 0.0)
    pts = lattice(nx, ny, 1, 2.0 * r)
    fixed = pts[:, 0] < 0.5 * r
    s.add_fixed_balls(pts[fixed], r, 0)
    s.add_balls(pts[~fixed], [0, 0, 0], [0, 0, 0], r, 2500.0, 0)
    n = pts.shape[0]
    f = np.zeros((n, 3))
    nfix = int(fixed.sum())
    end = np.flatnonzero(pts[~fixed][:, 0] > (nx - 1.5) * 2.0 * r)
    f[nfix + end] = np.array([force, -0.5 * force, 0.0]) / end.size
    s.hooks = [("const", f, 0.0)]
    s.bonds_from_initial_pairs = True
    s.min_bound, s.max_bound = half_cell_bounds([-4.0 * r, -0.5, -4.0 * r], [0.5, 0.5, 4.0 * r], 2.0 * r)
    s.gravity = np.zeros(3)
    s.dt = dt
    return s


def moving_plate(ball_on_plate: bool = True, half: float = 0.1, r: float = 0.01):
    """A square plate of two triangles (side 2 half, in the z = 0 plane of
    its wall frame, wall position at the origin), with one ball.  The wall's
    velocity and angular velocity are set after initialisation (addMeshWall
    stores zeros, DEMSolver.h:1901-1911).  With ball_on_plate the ball rests
    on the plate under gravity with Coulomb friction (linear contact), and
    the moving plate drags it; otherwise the ball floats 0.5 m above it in
    zero gravity and the wall moves freely."""
    s = DEMScene("moving_plate")
    v = np.array([[-half, -half, 0.0], [half, -half, 0.0], [half, half, 0.0], [-half, half, 0.0]])
    s.set_linear_pair(0, 0, 2e4, 2e4 / 2.6, 0.0, 0.0, 0.2, 0.2, 0.0, 0.0, 0.5, 0.0, 0.0)
    s.set_linear_pair(0, 1, 2e4, 2e4 / 2.6, 0.0, 0.0, 0.2, 0.2, 0.0, 0.0, 0.5, 0.0, 0.0)
    z = r - 2e-5 if ball_on_plate else 0.5
    s.add_balls([[0.03, 0.02, z]], [0, 0, 0], [0, 0, 0], r, 2500.0, 0)
    s.add_mesh_wall(v, [0, 0], [1, 2], [2, 3], [0, 0, 0], [0, 0, 0], [0, 0, 0], 1)
    s.min_bound, s.max_bound = np.array([-0.4, -0.4, -0.2]), np.array([0.4, 0.4, 0.8])
    s.gravity = np.array([0.0, 0.0, -9.81]) if ball_on_plate else np.zeros(3)
    s.dt = 1e-5
    return s


def pushed_clump(mode: str = "force", f: float = 2.0):
    """One rigid tetrahedral clump of four pebbles (r = 4 mm, centres 6 mm
    from the clump centre, mass 0.05 kg, isotropic inertia 2e-5 kg m^2) in
    zero gravity, pushed by constant hook forces on its pebbles
    (sumClumpForceTorqueKernel, ballIntegrationKernel.cu:3-28, sums them).

    * mode "force": f along +x on every pebble.  The pebbles sit symmetrically
      about the clump centre, so the torque is zero and the clump translates:
      after N steps v = N dt 4f / M and x = (4f / M) (N dt)^2 / 2 (KDK is exact
      for a constant force).
    * mode "torque": +f z on pebble 0 and -f z on pebble 1: zero net force and
      torque T = (p0 - p1) x f z, so omega = 2 (dt/2) T / I after one step, to
      O(|omega| dt)."""
    s = DEMScene(f"clump_{mode}")
    a = 0.006 / math.sqrt(3.0)
    pts = np.array([[a, a, a], [a, -a, -a], [-a, a, -a], [-a, -a, a]])
    s.set_linear_pair(0, 0, 1e3, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
    s.add_clump(pts, np.full(4, 0.004), [0.0, 0.0, 0.0], [0, 0, 0], [0, 0, 0], 0.05, [2e-5, 2e-5, 2e-5, 0, 0, 0], 0)
    fo = np.zeros((4, 3))
    if mode == "force":
        fo[:, 0] = f
    else:
        fo[0, 2], fo[1, 2] = f, -f
    s.hooks = [("const", fo, 0.0)]
    s.min_bound, s.max_bound = -np.ones(3), np.ones(3)
    s.gravity = np.zeros(3)
    s.dt = 1e-5
    return s
