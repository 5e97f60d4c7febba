"""DEM-SPH coupling on the GPU (demsph_couple_*).  The reference has no
coupling, so parity with it is unpinned; the model is pinned by what the
physics fixes: pairwise momentum exchange (Newton's third law), Archimedes'
force on a ball at rest in a hydrostatic column, and the direction of the
drag in a uniform stream."""
import math

import numpy as np
import pytest

import this_studio_amd as ts
from this_studio_amd import scenes

pytestmark = pytest.mark.gpu


def lattice(nx, ny, nz, dx, origin=(0.0, 0.0, 0.0)):
    z, y, x = np.meshgrid(np.arange(nz) * dx, np.arange(ny) * dx, np.arange(nx) * dx, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1) + np.asarray(origin)


def fluid_scene(spec, drho=None):
    s = ts.SPHScene()
    s.set_eos(spec.eos, spec.gamma)
    s.add_fluid(spec.fluid_pos, spec.fluid_vel, spec.c0, spec.dx, spec.rho0, spec.nu)
    if spec.nd:
        s.add_dummy(spec.dummy_pos, spec.dummy_vel, spec.c0, spec.dx, spec.rho0, spec.nu)
    s.initialize(spec.min_bound, spec.max_bound, spec.gravity)
    if drho is not None:
        s.set("densityChange", drho)
    return s


def column_with_ball(velocity=None, gravity=(0.0, 0.0, -9.81)):
    """A 17 x 17 x 25 lattice at dx = 1 cm with the centre particle replaced
    by a fixed ball of the same volume (r = dx (3/(4 pi))^(1/3)); at rest the
    column carries the hydrostatic pressure rho0 |g| (H - z)."""
    dx, rho0, c0 = 0.01, 1000.0, 20.0
    pos = lattice(17, 17, 25, dx)
    centre = np.array([8, 8, 12]) * dx
    keep = np.linalg.norm(pos - centre, axis=1) > 1e-9
    pos = pos[keep]
    g = np.asarray(gravity, float)
    H = pos[:, 2].max() + 0.5 * dx
    vel = None if velocity is None else np.tile(np.asarray(velocity, float), (pos.shape[0], 1))
    spec = scenes.SPHSceneSpec(fluid_pos=pos, dummy_pos=np.zeros((0, 3)), dx=dx, rho0=rho0, c0=c0, nu=1e-6,
                               min_bound=np.array([-0.05, -0.05, -0.05]), max_bound=np.array([0.22, 0.22, 0.3]),
                               gravity=g, dt=1e-6, fluid_vel=vel)
    p = rho0 * abs(g[2]) * (H - pos[:, 2])
    fluid = fluid_scene(spec, p / c0 ** 2)
    r = dx * (3.0 / (4.0 * math.pi)) ** (1.0 / 3.0)
    d = scenes.DEMScene("ball")
    d.add_fixed_balls(centre, r, 0)
    d.set_hertzian_pair(0, 0, 1e7, 1e6, 0.5, 0.0, 0.0, 0.3, 0.0, 0.0)
    d.min_bound, d.max_bound, d.gravity = spec.min_bound, spec.max_bound, g
    dem = ts.DEMSceneGPU(d)
    return spec, fluid, dem, r


def test_ball_at_rest_in_hydrostatic_column_feels_archimedes():
    spec, fluid, dem, r = column_with_ball()
    c = ts.CoupledScene(fluid, dem)
    c.step(spec.dt, 1, spec.gravity)
    F = c.get("couplingForce")[0]
    V = 4.0 / 3.0 * math.pi * r ** 3
    archimedes = spec.rho0 * 9.81 * V
    assert F[2] == pytest.approx(archimedes, rel=0.05)
    assert abs(F[0]) < 1e-3 * archimedes and abs(F[1]) < 1e-3 * archimedes
    # the interpolated ball pressure is the hydrostatic pressure at the centre
    H = spec.fluid_pos[:, 2].max() + 0.5 * spec.dx
    assert c.get("ballPressure")[0] == pytest.approx(spec.rho0 * 9.81 * (H - 0.12), rel=0.02)


def test_fixed_ball_in_uniform_stream_is_pushed_downstream():
    spec, fluid, dem, r = column_with_ball(velocity=(0.2, 0.0, 0.0), gravity=(0.0, 0.0, 0.0))
    c = ts.CoupledScene(fluid, dem)
    c.step(spec.dt, 1, spec.gravity)
    F = c.get("couplingForce")[0]
    assert F[0] > 0.0
    assert abs(F[1]) < 1e-3 * F[0] and abs(F[2]) < 1e-3 * F[0]


def falling_balls():
    spec = scenes.small_dam_break()
    fluid = fluid_scene(spec)
    d = scenes.DEMScene("balls in water")
    lo = spec.fluid_pos.min(axis=0)
    hi = spec.fluid_pos.max(axis=0)
    centres = lattice(3, 2, 2, 0.025, origin=lo + np.array([0.02, 0.015, 0.03]))
    assert np.all(centres < hi)
    d.add_balls(centres, [0.0, 0.0, -0.1], [0.0, 0.0, 0.0], 0.006, 2500.0, 0)
    d.set_hertzian_pair(0, 0, 1e7, 1e6, 0.5, 0.0, 0.0, 0.3, 0.0, 0.0)
    d.min_bound, d.max_bound, d.gravity = spec.min_bound, spec.max_bound, spec.gravity
    dem = ts.DEMSceneGPU(d)
    return spec, fluid, dem


def test_momentum_exchange_is_pairwise():
    """Sum of the coupling forces on the balls plus the fluid's coupling
    momentum rate (same evaluation) vanishes to FP32 summation rounding."""
    spec, fluid, dem = falling_balls()
    c = ts.CoupledScene(fluid, dem)
    c.step(spec.dt, 3, spec.gravity)
    F = c.get("couplingForce")
    a = c.get("fluidCouplingAcceleration")
    m = spec.dx ** 3 * spec.rho0
    total = F.sum(axis=0) + m * a.sum(axis=0)
    scale = np.abs(F).sum() + m * np.abs(a).sum()
    assert scale > 0
    assert np.abs(total).max() < 1e-5 * scale, (total, scale)


def test_coupled_dam_break_with_balls_stays_finite_and_drags_balls():
    spec, fluid, dem = falling_balls()
    v0 = dem.get("velocity")[:, 2].copy()
    c = ts.CoupledScene(fluid, dem)
    c.step(spec.dt, 50, spec.gravity)
    pos, vel = dem.get("position"), dem.get("velocity")
    assert np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
    assert np.all(np.isfinite(fluid.get("position"))) and np.all(np.isfinite(fluid.get("density")))
    # buoyancy and drag slow the sinking relative to free fall over the same time
    t = 50 * spec.dt
    free = v0 - 9.81 * t
    assert np.all(vel[:, 2] > free)
