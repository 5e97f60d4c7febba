"""GPU parity of the DEM path (libdemsph.so on the MI355X) against the FP64
oracle, closed forms and the reference's own output frames.

Bit-exact: hashes, sort permutation, cell ranges, ball-pair lists (as the
reference's objectPointed/objectPointing arrays, in its order) and
ball-triangle candidate lists.  Forces run in FP32 with FP64 geometry and
spring history (default) or fully FP64; tolerances are stated per test.
"""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
import this_studio_amd as ts
from this_studio_amd import scenes

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def random_scene(n=3000, seed=7, box=0.4):
    rng = np.random.default_rng(seed)
    s = scenes.DEMScene("random")
    pos = rng.uniform(0.02, box - 0.02, (n, 3))
    rad = rng.uniform(0.005, 0.01, n)
    clump = np.where(rng.uniform(size=n) < 0.1, rng.integers(0, 50, n), -1).astype(np.int32)
    fixed = rng.uniform(size=n) < 0.1
    vel = rng.normal(scale=0.1, size=(n, 3))
    for k in range(n):
        if fixed[k]:
            s.add_fixed_balls(pos[k], rad[k], 0, clump[k])
        else:
            s.add_balls(pos[k], vel[k], [0, 0, 0], rad[k], 1000.0, 0, clump[k])
    s.set_hertzian_pair(0, 0, 5e6, 1e6, 0.5, 0.1, 0.1, 0.3, 0.05, 0.05)
    s.min_bound = np.zeros(3)
    s.max_bound = box * np.ones(3)
    s.dt = 1e-5
    return s


@pytest.mark.parametrize("name", ["random", "collapse", "random_40k"])
def test_neighbor_search_bit_exact(name):
    """Hash, cells and pair lists against the oracle, bit for bit.  Up to
    32,768 balls the pair walk runs 8 lanes per ball, above it one (the
    40k-ball scene)."""
    if name == "random":
        s = random_scene()
    elif name == "random_40k":
        s = random_scene(n=40000, seed=11, box=0.8)
    else:
        s = scenes.granular_collapse()
    o = oracle.OracleDEM(s, solve_begin=False)
    g = ts.DEMSceneGPU(s, solve_begin=False)
    o.neighbor_search()
    g.neighbor_search()
    assert np.array_equal(g.get("hashIndex"), o.get("index"))
    assert np.array_equal(g.get("cellHashStart"), o.get("cellStart"))
    assert np.array_equal(g.get("cellHashEnd"), o.get("cellEnd"))
    assert np.array_equal(g.get("objectPointed"), o.get("pointed"))
    assert np.array_equal(g.get("objectPointing"), o.get("pointing"))


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("fp64", [False, True])
def test_forces_after_solve_begin(fp64):
    s = random_scene()
    o = oracle.OracleDEM(s)
    g = ts.DEMSceneGPU(s, fp64=fp64)
    # FP32 forces: 1e-5 of the largest force; FP64: 1e-10
    tol = 1e-10 if fp64 else 1e-5
    assert _rel(g.get("force"), o.get("force")) < tol
    assert _rel(g.get("torque"), o.get("torque")) < tol
    np.testing.assert_allclose(g.get("overlap"), o.get("overlap"), atol=1e-15)


@pytest.mark.parametrize("fp64", [False, True])
def test_random_scene_100_steps(fp64):
    s = random_scene()
    o = oracle.OracleDEM(s)
    g = ts.DEMSceneGPU(s, fp64=fp64)
    o.step(100)
    g.step(100)
    # positions within 1e-5 of the smallest diameter (FP32 forces) or 1e-10 (FP64)
    d = 2 * 0.005
    err = np.abs(g.get("position") - o.get("pos")).max() / d
    assert err < (1e-10 if fp64 else 1e-5), err
    assert np.array_equal(g.get("objectPointing"), o.get("pointing"))
    # spring history carried through 20 rebuilds
    assert _rel(g.get("slidingSpring"), o.get("sliding")) < (1e-8 if fp64 else 1e-3)


def test_elastic_bounce_matches_reference_frames():
    s = scenes.elastic_bounce()
    g = ts.DEMSceneGPU(s)
    fx = np.load(GOLDEN / "elastic_bounce_frames.npz")["frames"]
    z, vz = [], []
    for _ in range(200):
        g.step(100)
        z.append(g.get("position")[0, 2])
        vz.append(g.get("velocity")[0, 2])
    # FP32 contact force arithmetic: positions within 1e-6 m, velocities 1e-5 m/s
    assert np.abs(np.array(z) - fx[1:, 4]).max() < 1e-6
    assert np.abs(np.array(vz) - fx[1:, 7]).max() < 1e-5


def test_bonded_beam_matches_oracle_and_closed_form():
    s = scenes.bonded_beam()
    o = oracle.OracleDEM(s)
    g = ts.DEMSceneGPU(s)
    assert g.get("bondPointed").tolist() == list(range(10))
    o.step(20000)
    g.step(20000)
    assert _rel(g.get("position")[:, 2], o.get("pos")[:, 2]) < 1e-5
    # the axial force of a transversely loaded beam is a small difference of
    # bond elongations; compare it on the scale of the bond shear load
    shear = np.abs(o.get("bond_shearForce")).max()
    assert np.abs(g.get("bondNormalForce") - o.get("bond_normalForce")).max() / shear < 1e-5
    assert _rel(g.get("bondShearForce"), o.get("bond_shearForce")) < 1e-4
    g.step(100000)
    tip = g.get("position")[-1, 2]
    eb = 1e5 * 4.0 ** 3 / (3 * 200e9 * math.pi * 0.2 ** 4 / 4)
    assert abs(tip - eb) / eb < 0.01


def test_hertz_collision_closed_form():
    E, r, rho, v0 = 1e8, 0.01, 2500.0, 0.5
    s = scenes.DEMScene("two_balls")
    s.set_hertzian_pair(0, 0, E, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    s.add_balls([[-r - 1e-5, 0.0, 0.0], [r + 1e-5, 0.0, 0.0]], [[v0, 0, 0], [-v0, 0, 0]], [0, 0, 0], r, rho, 0)
    s.min_bound, s.max_bound, s.gravity, s.dt = -0.1 * np.ones(3), 0.1 * np.ones(3), np.zeros(3), 1e-6
    g = ts.DEMSceneGPU(s)
    steps, dmax = 0, 0.0
    for k in range(1000):
        g.step(1)
        ov = g.get("overlap")
        if ov.size and ov[0] > 0:
            steps += 1
            dmax = max(dmax, ov[0])
        elif steps:
            break
    m = 4 / 3 * math.pi * r ** 3 * rho
    ms, Rs, v = m / 2, r / 2, 2 * v0
    tc = 2.868 * (ms ** 2 / (Rs * E ** 2 * v)) ** 0.2
    assert abs(steps * s.dt - tc) / tc < 0.01
    assert abs(dmax - (15 * ms * v ** 2 / (16 * E * math.sqrt(Rs))) ** 0.4) / dmax < 0.005
    assert abs(g.get("velocity")[1, 0] - v0) / v0 < 2e-3


def test_granular_collapse_matches_oracle():
    s = scenes.granular_collapse()
    o = oracle.OracleDEM(s)
    g = ts.DEMSceneGPU(s)
    o.step(300)
    g.step(300)
    err = np.abs(g.get("position") - o.get("pos")).max() / 0.01
    assert err < 1e-5, err
    assert np.array_equal(g.get("objectPointing"), o.get("pointing"))
    assert np.array_equal(g.get("triPointing"), o.get("tri_pointing"))
    assert np.array_equal(g.get("triCancel"), o.get("tri_cancel"))


def test_clumps_on_plane_match_oracle():
    s = scenes.DEMScene("clumps")
    rng = np.random.default_rng(5)
    a = 0.02 / math.sqrt(8.0)
    tet = np.array([[a, a, a], [a, -a, -a], [-a, a, -a], [-a, -a, a]])
    for k in range(24):
        c = np.array([0.05 + 0.06 * (k % 6), 0.05 + 0.06 * (k // 6), 0.05 + 0.01 * (k % 3)])
        s.add_clump(tet + c, np.full(4, 0.008), c, rng.normal(scale=0.05, size=3), rng.normal(size=3), 0.02,
                    [2e-6, 2.5e-6, 3e-6, 0, 0, 0], 0)
    s.set_linear_pair(0, 0, 2e3, 2e3 / 2.6, 0, 0, 0.2, 0.2, 0, 0, 0.4, 0, 0)
    v = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [1.0, 1.0, 0.0], [-1.0, 1.0, 0.0]])
    s.add_mesh_wall(v, [0, 1], [1, 2], [3, 3], [0, 0, 0], [0, 0, 0], [0, 0, 0], 0)
    s.min_bound = np.array([-0.05, -0.05, -0.05])
    s.max_bound = np.array([0.45, 0.25, 0.2])
    s.dt = 2e-5
    o = oracle.OracleDEM(s)
    g = ts.DEMSceneGPU(s)
    o.step(1500)
    g.step(1500)
    assert np.abs(g.get("clumpPosition") - o.get("clump_pos")).max() < 1e-7
    assert np.abs(g.get("position") - o.get("pos")).max() < 1e-7


def test_cantilever_with_clumps_matches_oracle():
    """A reduced bonded cantilever with falling tetrahedral clumps (16 x 4 x 4
    beam, 40 clumps; the bench's scene at full size): bonds from the initial
    pairs, ramped tip load through the hook, clumps under gravity; 3,000 steps
    against the FP64 oracle (FP32 contact and bond forces: 1e-5 of the
    diameter)."""
    s = scenes.cantilever_with_clumps(nx=16, ny=4, nz=4, n_clumps=40, ramp_steps=1000)
    o = oracle.OracleDEM(s)
    g = ts.DEMSceneGPU(s)
    assert np.array_equal(g.get("bondPointed"), o.get("bond_pointed"))
    assert np.array_equal(g.get("bondPointing"), o.get("bond_pointing"))
    o.step(3000)
    g.step(3000)
    d = 0.004
    assert np.abs(g.get("position") - o.get("pos")).max() / d < 1e-5
    assert np.abs(g.get("clumpPosition") - o.get("clump_pos")).max() / d < 1e-5
    # the tip layer sags under the ramped load
    x0 = np.asarray(s.initial_positions())
    tip = np.isclose(x0[:, 0], 15 * d) & (x0[:, 2] < 4 * d)
    assert g.get("position")[tip, 2].mean() < x0[tip, 2].mean() - 1e-7


def test_graph_replays_equal_eager_steps():
    """Four-step graphs between pair rebuilds against eager launches, bit for
    bit, on the collapse (static box walls) run on a stream of its own, over
    blocks that start on and off the rebuild cycle."""
    import torch

    st = torch.cuda.Stream()
    s = scenes.granular_collapse()
    a = ts.DEMSceneGPU(s, stream=st.cuda_stream)
    b = ts.DEMSceneGPU(s, stream=st.cuda_stream)
    b.set_graphs(False)
    for n in (23, 40, 7, 30):
        a.step(n)
        b.step(n)
        for f in ("position", "velocity", "angularVelocity", "slidingSpring"):
            assert np.array_equal(a.get(f), b.get(f)), (n, f)
    assert a.clock() == b.clock()
    assert a.graph_captures() >= 2
    assert b.graph_captures() == 0
    a.close()
    b.close()


def test_user_hook_through_views_equals_builtin_hook():
    """addExternalForceTorque through the C-ABI callback, calling the HIP
    launcher on the insertion-ordered FP64 views, against the built-in
    device hook program: bit-identical trajectories."""
    s = scenes.bonded_beam()
    builtin = ts.DEMSceneGPU(s)
    s2 = scenes.bonded_beam()
    s2.hooks = []
    user = ts.DEMSceneGPU(s2)
    L = ts.lib()
    L.launchAddConstantForce.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_void_p]
    L.launchAddGlobalDampingForceTorque.argtypes = [C.c_void_p] * 4 + [C.c_double, C.c_size_t, C.c_size_t,
                                                                         C.c_size_t, C.c_void_p]
    import torch

    f_e = torch.zeros((11, 3), dtype=torch.float64, device="cuda:0")

    def hook(i, t):
        v = user.views()
        c = min(t, 1.0)
        f_e[10, 2] = 1e5 * c
        torch.cuda.synchronize()
        L.launchAddConstantForce(v.force, f_e.data_ptr(), v.numBalls, 1, 64, None)
        L.launchAddGlobalDampingForceTorque(v.force, v.torque, v.velocity, v.angularVelocity, 0.1, v.numBalls, 1, 64,
                                            None)

    user.set_hook(hook)
    builtin.step(300)
    user.step(300)
    assert np.array_equal(builtin.get("position"), user.get("position"))
    assert np.array_equal(builtin.get("velocity"), user.get("velocity"))


def test_ice_cylinder_before_fracture_matches_oracle():
    """iceCylinderInteraction.cpp scene (11,600 bonded balls, buoyancy hook,
    2,664-triangle cylinder): 500 steps, before any bond breaks, against the
    FP64 oracle.  FP32 forces: positions within 1e-5 of a diameter."""
    s = scenes.ice_cylinder()
    o = oracle.OracleDEM(s)
    g = ts.DEMSceneGPU(s)
    assert g.get("bondPointed").size == o.get("bond_pointed").size == 33035
    o.step(500)
    g.step(500)
    assert np.all(g.get("bondIsBonded") == 1)
    err = np.abs(g.get("position") - o.get("pos")).max() / 0.2
    assert err < 1e-5, err
    assert _rel(g.get("velocity"), o.get("vel")) < 1e-4
    assert np.array_equal(g.get("triPointing"), o.get("tri_pointing"))


def test_ice_cylinder_reference_frame_observables():
    """The reference's own frame at step 12,500 (tests/golden/
    ice_cylinder_frames.npz).  Past the first bond failures (about step
    3,000) the crack path depends on threshold events and on the
    reference's nondeterministic FP64 atomic bond sums, so the comparison is
    at the level of observables: the sheet's median drift within 5 %, a
    fractured sheet, and contact with the cylinder."""
    s = scenes.ice_cylinder()
    ref = np.load(GOLDEN / "ice_cylinder_frames.npz")
    g = ts.DEMSceneGPU(s)
    g.step(12500)
    p = g.get("position")
    assert np.all(np.isfinite(p))
    drift = np.median(p[:, 0] - ref["ball_pos0"][:, 0])
    drift_ref = np.median(ref["ball_pos1"][:, 0].astype(np.float64) - ref["ball_pos0"][:, 0])
    assert abs(drift - drift_ref) < 0.05 * abs(drift_ref), (drift, drift_ref)
    assert np.abs(p[:, 2]).max() < 1e-9  # the sheet floats at its buoyant level
    assert (g.get("bondIsBonded") == 0).sum() > 100
    assert g.get("triPointing").size > 0


def _vtu_arrays(path):
    import re
    out = {}
    for attrs, body in re.findall(r'<DataArray([^>]*)>(.*?)</DataArray>', path.read_text(), re.S):
        m = re.search(r'Name="([^"]+)"', attrs)
        name = m.group(1) if m else "Points"
        v = np.array(body.split(), dtype=np.float64)
        if 'NumberOfComponents="3"' in attrs:
            v = v.reshape(-1, 3)
        out[name] = v
    return out


def test_ball_triangle_frame_reports_contact_closed_form(tmp_path):
    """A ball pressed into a flat two-triangle plane by a known overlap, with
    a linear spring and no damping, gravity or motion.  After the first
    force evaluation the ballTriangleInteraction frame holds one contact
    with the closed-form point, normal, overlap and force kn*delta along
    the normal and zero torque.  Its overlap and springs equal the
    triOverlap / triSlidingSpring fields."""
    r, delta, kn = 0.01, 1.0e-4, 2.0e4
    s = scenes.DEMScene("press")
    s.add_balls([0.3, -0.2, r - delta], [0, 0, 0], [0, 0, 0], r, 2000.0, 0)
    verts = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [1.0, 1.0, 0.0], [-1.0, 1.0, 0.0]])
    s.add_mesh_wall(verts, [0, 0], [1, 2], [2, 3], [0, 0, 0], [0, 0, 0], [0, 0, 0], 0)
    s.set_linear_pair(0, 0, kn, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
    s.min_bound, s.max_bound = np.array([-1.0, -1.0, -1.0]), np.array([1.0, 1.0, 1.0])
    s.gravity = np.zeros(3)
    s.dt = 1e-5
    g = ts.DEMSceneGPU(s, solve_begin=False)
    g.record_contacts(True)
    g.solve_begin()
    g.write_vtu(tmp_path, 0, 0, 0.0, walls=False, wall_contacts=True)
    a = _vtu_arrays(tmp_path / "ballTriangleInteraction_0000.vtu")
    # both triangles are candidates (bounding boxes); the ball touches the
    # interior of the first only
    assert a["Points"].shape == (2, 3) and g.get("triPointing").size == 2
    k = int(np.argmax(a["overlap"]))
    np.testing.assert_allclose(a["Points"][k], [0.3, -0.2, 0.0], atol=1e-10)
    np.testing.assert_allclose(a["contact normal"][k], [0.0, 0.0, 1.0], atol=1e-10)
    assert a["overlap"][k] == pytest.approx(delta, rel=1e-9)
    # FP32 force arithmetic: 1e-6 relative
    np.testing.assert_allclose(a["contact force"][k], [0.0, 0.0, kn * delta], rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(a["contact torque"], 0.0, atol=1e-12)
    assert a["overlap"][1 - k] < 0 and np.all(a["contact force"][1 - k] == 0)
    np.testing.assert_allclose(np.linalg.norm(a["contact normal"], axis=1), 1.0, atol=1e-12)
    assert np.allclose(np.sort(a["overlap"]), np.sort(g.get("triOverlap")), rtol=0, atol=5e-11)
    # the ball's total force is the contact's
    np.testing.assert_allclose(g.get("force")[0], a["contact force"][k], rtol=1e-6, atol=1e-12)


def test_ball_pair_frame_reports_contact_closed_form(tmp_path):
    """Two equal balls overlapping by delta on a linear spring, at rest, no
    gravity.  The ballInteraction frame after the first evaluation holds
    the pair: contact point at the midpoint, normal along the centre line,
    overlap delta, force kn*delta along the normal on the pointed ball, zero
    torque.  Its entries follow the pair getters' order."""
    r, delta, kn = 0.01, 2.0e-4, 5.0e3
    s = scenes.DEMScene("pair")
    s.add_balls([[0.1, 0.2, 0.3], [0.1 + 2 * r - delta, 0.2, 0.3]], [0, 0, 0], [0, 0, 0], r, 2000.0, 0)
    s.set_linear_pair(0, 0, kn, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
    s.min_bound, s.max_bound = np.array([-1.0, -1.0, -1.0]), np.array([1.0, 1.0, 1.0])
    s.gravity = np.zeros(3)
    s.dt = 1e-5
    g = ts.DEMSceneGPU(s, solve_begin=False)
    g.record_contacts(True)
    g.solve_begin()
    g.write_vtu(tmp_path, 0, 0, 0.0, walls=False, pairs=True)
    a = _vtu_arrays(tmp_path / "ballInteraction_0000.vtu")
    assert a["Points"].shape == (1, 3)
    assert g.get("objectPointed").size == 1
    np.testing.assert_allclose(a["Points"][0], [0.1 + r - 0.5 * delta, 0.2, 0.3], atol=1e-12)
    nrm = a["contact normal"][0]
    np.testing.assert_allclose(np.abs(nrm), [1.0, 0.0, 0.0], atol=1e-12)
    assert a["overlap"][0] == pytest.approx(delta, rel=1e-9)
    f = a["contact force"][0]
    assert float(f @ nrm) == pytest.approx(kn * delta, rel=1e-6)
    np.testing.assert_allclose(np.cross(f, nrm), 0.0, atol=1e-9)
    np.testing.assert_allclose(a["contact torque"], 0.0, atol=1e-12)
    # frame values carry 10 decimals
    np.testing.assert_allclose(a["overlap"], g.get("overlap"), rtol=0, atol=5e-11)
    # the pointed ball's force is the contact force
    pointed = int(g.get("objectPointed")[0])
    np.testing.assert_allclose(g.get("force")[pointed], f, rtol=1e-6, atol=1e-12)


def test_background_ball_frame_is_a_snapshot(tmp_path):
    """The ball frame written in the background equals the blocking write of
    the same state while the scene keeps stepping."""
    g = ts.DEMSceneGPU(scenes.granular_collapse())
    g.step(20)
    g.write_vtu(tmp_path / "now", 1, 20, 0.0002, walls=False)
    g.write_vtu(tmp_path / "bg", 1, 20, 0.0002, walls=False, background=True)
    g.step(50)
    ts.frames_wait()
    assert (tmp_path / "now" / "ball_0001.vtu").read_bytes() == (tmp_path / "bg" / "ball_0001.vtu").read_bytes()


_STATE_FIELDS = ("position", "velocity", "angularVelocity", "force", "torque", "slidingSpring", "rollingSpring",
                 "torsionSpring", "overlap", "objectPointed", "objectPointing")


@pytest.mark.parametrize("name", ["collapse", "collapse_fp64", "cantilever_clumps"])
def test_exported_state_resumes_bit_for_bit(name):
    """A DEM scene exported between pair rebuilds and imported into a freshly
    built scene of the same staging continues bit for bit as the scene that
    kept running: ball and wall contact history on the collapse (box walls,
    FP32 and FP64 contact arithmetic), bond state, clumps and the time-ramped
    hook load on the cantilever.  The importer has either stepped already or
    (collapse) never run solve_begin; a scene of the other precision refuses
    the state."""
    fp64 = name == "collapse_fp64"
    if name.startswith("collapse"):
        make, extra = scenes.granular_collapse, ("triPointed", "triSlidingSpring", "triOverlap")
    else:
        def make():
            return scenes.cantilever_with_clumps(nx=8, ny=3, nz=3, n_clumps=12, ramp_steps=50)
        extra = ("clumpPosition", "clumpOrientation", "clumpAngularVelocity", "bondIsBonded", "bondNormalForce",
                 "bondShearForce", "bondBendingTorque", "bondTorsionTorque")
    a = ts.DEMSceneGPU(make(), fp64=fp64)
    a.step(23)  # 23 % 5 != 0: mid-way between rebuilds
    state = a.export_state()
    for graphs in (True, False):
        # the cantilever's bonds come from solve_begin's initial pair list
        b = ts.DEMSceneGPU(make(), fp64=fp64, solve_begin=graphs or name == "cantilever_clumps")
        b.set_graphs(graphs)
        if graphs:
            b.step(7)  # the state replaces whatever b held
        b.import_state(state)
        assert b.clock() == a.clock()
        ref = ts.DEMSceneGPU(make(), fp64=fp64)
        ref.step(23)
        ref.step(41)
        b.step(41)
        for f in _STATE_FIELDS + extra:
            assert np.array_equal(ref.get(f), b.get(f)), (name, graphs, f)
        assert ref.clock() == b.clock()
    with pytest.raises(ts.DemsphError):
        b.import_state(state[:-8])
    with pytest.raises(ts.DemsphError):
        ts.DEMSceneGPU(scenes.bonded_beam()).import_state(state)
    with pytest.raises(ts.DemsphError):
        ts.DEMSceneGPU(make(), fp64=not fp64).import_state(state)


@pytest.mark.parametrize("name", ["collapse", "random_40k"])
def test_pair_lane_settings_build_identical_lists(name):
    """Every "pairLanes" setting builds the oracle's pair lists in the
    oracle's order; the default depends on the ball count."""
    s = scenes.granular_collapse() if name == "collapse" else random_scene(n=40000, seed=11, box=0.8)
    o = oracle.OracleDEM(s, solve_begin=False)
    o.neighbor_search()
    g = ts.DEMSceneGPU(s, solve_begin=False)
    assert g.tuned("pairLanes") == (8 if name == "collapse" else 1)
    for lanes in (1, 2, 4, 8):
        g.tune("pairLanes", lanes)
        assert g.tuned("pairLanes") == lanes
        g.neighbor_search()
        assert np.array_equal(g.get("objectPointed"), o.get("pointed")), lanes
        assert np.array_equal(g.get("objectPointing"), o.get("pointing")), lanes
    with pytest.raises(ts.DemsphError):
        g.tune("pairLanes", 3)


def test_autotune_restores_state_and_keeps_results():
    """The DEM autotuner times every "pairLanes" setting from a checkpoint
    taken with the state export, restores the state byte for byte, and the
    tuned scene continues bit for bit as an untuned one."""
    s = scenes.granular_collapse()
    a = ts.DEMSceneGPU(s)
    b = ts.DEMSceneGPU(s)
    a.enable_timing(True)
    a.step(12)
    b.step(12)
    timed_before = a.timing()
    before = a.export_state()
    lanes = a.autotune(trial_steps=10)
    assert lanes in (1, 2, 4, 8) and a.tuned("pairLanes") == lanes
    assert a.export_state() == before
    assert a.timing() == timed_before  # the trials' own timings are not kept
    a.enable_timing(False)
    a.step(30)
    b.step(30)
    for f in ("position", "velocity", "angularVelocity", "slidingSpring", "objectPointed"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert a.clock() == b.clock()
    with pytest.raises(ts.DemsphError):
        a.autotune(trial_steps=3)


@pytest.mark.parametrize("case", list(scenes.BOND_LOAD_CASES))
def test_bond_breaks_at_closed_form_evaluation(case):
    """The GPU bond kernel under constant relative rates (tension, tension
    against cohesion with friction, bending, torsion, shear): the governing
    stress grows by inc per evaluation and the bond breaks at the same
    evaluation as the closed form and the oracle (ParallelBondedContact,
    contactKernel.cuh:4-62).  Bond state and stresses are FP64 on the GPU."""
    s, inc, kb, gov = scenes.bonded_pair_load(case)
    g = ts.DEMSceneGPU(s)
    o = oracle.OracleDEM(s)
    assert g.get("bondPointed").tolist() == [0] and g.get("bondPointing").tolist() == [1]
    tol = 2e-5 if case == "shear" else 1e-9
    field = "bondMaxNormalStress" if gov == "sigma" else "bondMaxShearStress"
    for step in range(1, kb // 2 + 3):
        g.step(1)
        o.step(1)
        k = 2 * step
        val = g.get(field)[0]
        assert g.get("bondIsBonded")[0] == o.get("bond_isBonded")[0] == (1 if k < kb else 0), (case, step)
        assert val == pytest.approx(min(k, kb) * inc, rel=tol)
        assert val == pytest.approx(o.get("bond_maxNormalStress" if gov == "sigma" else "bond_maxShearStress")[0],
                                    rel=1e-9)
    g.close()
    o.close()


def test_bonded_strip_breaks_like_oracle():
    """Progressive fracture: a bonded 24 x 4 strip, clamped at one end and
    pulled at the other, loses bonds from step ~800 on.  In FP64 the GPU
    breaks the same bonds at the same steps as the oracle and the strip's
    pieces stay on the oracle's trajectories."""
    s = scenes.bonded_strip_pull(force=150.0)
    g = ts.DEMSceneGPU(s, fp64=True)
    o = oracle.OracleDEM(s)
    nb = g.get("bondIsBonded").size
    broken = 0
    for blk in range(12):
        g.step(100)
        o.step(100)
        gb, ob = g.get("bondIsBonded"), o.get("bond_isBonded")
        assert np.array_equal(gb, ob), (blk + 1) * 100
        broken = nb - int(gb.sum())
        assert np.abs(g.get("position") - o.get("pos")).max() < 1e-9
    assert broken >= 8, broken
    g.close()
    o.close()


def test_moving_wall_kinematics_closed_form():
    """Wall motion on the GPU (k_wall_move, k_vertex_update, k_circumcenters;
    wallIntegrationKernel.cu:3-53, a full dt at each of the two evaluations
    of a step, DEMSolver.h:556-564) against the closed-form pose."""
    from test_oracle_dem import moving_wall_closed_form

    s = scenes.moving_plate(ball_on_plate=False)
    g = ts.DEMSceneGPU(s)
    v, w, N = np.array([0.05, -0.02, 0.01]), np.array([10.0, -20.0, 40.0]), 500
    g.set("wallVelocity", v)
    g.set("wallAngularVelocity", w)
    g.step(N)
    local = s.objects[1][1]["verts"]
    p, q, verts = moving_wall_closed_form(v, w, s.dt, N, local)
    np.testing.assert_allclose(g.get("wallPosition").reshape(-1), p, rtol=0, atol=1e-15)
    np.testing.assert_allclose(g.get("wallOrientation").reshape(-1), q, rtol=0, atol=1e-12)
    np.testing.assert_allclose(g.get("vertexGlobalPosition").reshape(-1, 3), verts, rtol=0, atol=1e-13)
    np.testing.assert_allclose(g.get("triangleCircumcenter").reshape(-1, 3), [p, p], rtol=0, atol=1e-13)
    g.close()


@pytest.mark.parametrize("fp64", [False, True])
def test_ball_on_moving_plate_matches_oracle(fp64):
    """A ball resting on a plate that translates and spins about its normal:
    friction drags the ball (the wall's rigid velocity at the contact point
    enters the relative velocity, contactKernel.cu:420-533), and the plate's
    triangles are re-hashed every evaluation.  GPU against the FP64 oracle
    after 3,000 steps."""
    s = scenes.moving_plate(ball_on_plate=True)
    g = ts.DEMSceneGPU(s, fp64=fp64)
    o = oracle.OracleDEM(s)
    v, w = np.array([0.2, 0.0, 0.0]), np.array([0.0, 0.0, 3.0])
    for obj in (g, o):
        obj.set("wallVelocity" if obj is g else "wall_vel", v)
        obj.set("wallAngularVelocity" if obj is g else "wall_angVel", w)
    g.step(3000)
    o.step(3000)
    gp, op = g.get("position")[0], o.get("pos")[0]
    # the ball was dragged along: it moved about a millimetre in x and y
    assert np.linalg.norm(op[:2] - [0.03, 0.02]) > 5e-4
    d = 0.02
    assert np.abs(gp - op).max() / d < (1e-10 if fp64 else 1e-5)
    assert np.abs(g.get("velocity") - o.get("vel")).max() < (1e-8 if fp64 else 1e-4)
    np.testing.assert_allclose(g.get("wallPosition"), o.get("wall_pos").reshape(-1, 3), rtol=0, atol=1e-14)
    g.close()
    o.close()


def test_clump_sums_closed_form():
    """The clump force and torque sums (sumClumpForceTorqueKernel,
    ballIntegrationKernel.cu:3-28, intended form) on the GPU: four equal
    pebble forces translate the clump at 4f/M exactly; an opposed pair of
    pebble forces turns it at omega = dt T / I after one step."""
    s = scenes.pushed_clump("force", f=2.0)
    N, dt, M = 200, s.dt, 0.05
    g = ts.DEMSceneGPU(s)
    g.step(N)
    acc = 4 * 2.0 / M
    assert g.get("clumpVelocity")[0] == pytest.approx([acc * N * dt, 0, 0], rel=1e-12, abs=1e-15)
    assert g.get("clumpPosition")[0] == pytest.approx([acc * (N * dt) ** 2 / 2, 0, 0], rel=1e-12, abs=1e-15)
    assert np.abs(g.get("clumpAngularVelocity")).max() < 1e-12
    g.close()
    s = scenes.pushed_clump("torque", f=2.0)
    g = ts.DEMSceneGPU(s)
    o = oracle.OracleDEM(s)
    g.step(1)
    o.step(1)
    pts = s.objects[0][1]["points"]
    T = np.cross(pts[0] - pts[1], [0.0, 0.0, 2.0])
    np.testing.assert_allclose(g.get("clumpAngularVelocity")[0], dt * T / 2e-5, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(g.get("clumpAngularVelocity")[0], o.get("clump_angVel")[0], rtol=1e-9, atol=1e-15)
    assert not g.get("clumpVelocity").any()
    g.close()
    o.close()


def _periodic_box(per, seed):
    rng = np.random.default_rng(seed)
    s = scenes.DEMScene("periodic_box")
    box = 0.2
    pos = rng.uniform(0.0, box, (3000, 3))
    pos[:20, 0] = box + rng.uniform(0.0, 0.004, 20)  # outside the faces: wrapped at the first search
    pos[20:40, 1] = -rng.uniform(1e-6, 0.004, 20)
    rad = rng.uniform(0.004, 0.006, 3000)
    s.add_balls(pos, rng.normal(scale=0.05, size=(3000, 3)), np.zeros(3), rad, 1000.0, 0)
    s.set_hertzian_pair(0, 0, 5e6, 1e6, 0.5, 0.1, 0.1, 0.3, 0.05, 0.05)
    s.set_periodic_boundary(*per)
    s.min_bound, s.max_bound, s.gravity, s.dt = np.zeros(3), box * np.ones(3), np.zeros(3), 1e-5
    return s


@pytest.mark.parametrize("per,lanes", [((True, True, True), 8), ((True, True, True), 1), ((False, True, False), 4),
                                       ((True, False, True), 2)])
def test_periodic_lists_and_steps_match_oracle(per, lanes):
    """Periodic boundaries (setPeriodicBoundary, DEMSolver.h:1944-1959, with
    the wrapped stencil): wrapped positions, pair lists with their image
    codes bit for bit against the oracle, with the pair walk on 1 to 8
    lanes per ball; then 100 steps (20 rebuilds) of a dense random box."""
    s = _periodic_box(per, seed=17)
    o = oracle.OracleDEM(s, solve_begin=False)
    g = ts.DEMSceneGPU(s, solve_begin=False)
    g.tune("pairLanes", lanes)
    o.neighbor_search()
    g.neighbor_search()
    assert np.array_equal(g.get("position"), o.get("pos"))  # the wrap is exact
    for f, of in (("objectPointed", "pointed"), ("objectPointing", "pointing"), ("pairShift", "pairShift")):
        assert np.array_equal(g.get(f), o.get(of)), f
    assert (g.get("pairShift") != 13).sum() > 50
    # a dense random box starts with large overlaps: compare the trajectory
    # with FP64 forces, where rounding differences stay small
    o2 = oracle.OracleDEM(s)
    g2 = ts.DEMSceneGPU(s, fp64=True)
    g2.tune("pairLanes", lanes)
    o2.step(100)
    g2.step(100)
    assert np.array_equal(g2.get("objectPointing"), o2.get("pointing"))
    assert np.array_equal(g2.get("pairShift"), o2.get("pairShift"))
    err = np.abs(g2.get("position") - o2.get("pos")).max() / 0.008
    assert err < 1e-10, err


@pytest.mark.parametrize("fp64", [False, True])
def test_periodic_ice_sheet_matches_oracle(fp64):
    """The y-periodic bonded ice sheet (rows of iceCylinderInteraction.cpp on
    its snapped periodic cell), rolled by 6 rows so that a row starts just
    below the maximum face: bonds across the face from the initial pairs,
    400 steps of waves that break bonds, against the FP64 oracle."""
    s = scenes.periodic_ice_sheet(roll_rows=6, gap=1e-3)
    o = oracle.OracleDEM(s)
    g = ts.DEMSceneGPU(s, fp64=fp64)
    assert np.array_equal(g.get("bondPointed"), o.get("bond_pointed"))
    assert np.array_equal(g.get("bondPointing"), o.get("bond_pointing"))
    assert np.array_equal(g.get("pairShift"), o.get("pairShift"))
    o.step(400)
    g.step(400)
    err = np.abs(g.get("position") - o.get("pos")).max() / 0.2
    assert err < (1e-12 if fp64 else 1e-6), err
    assert np.array_equal(g.get("bondIsBonded"), o.get("bond_isBonded"))
    assert (o.get("bond_isBonded") == 0).sum() > 10


def test_periodic_sheet_translates_rigidly_through_the_face():
    """A bonded periodic sheet moving at 50 m/s along y crosses the face
    (positions wrapped into the cell at each rebuild) with every bond intact
    and no force beyond rounding."""
    s = scenes.periodic_ice_sheet(uniform_velocity=[0.0, 50.0, 0.0])
    g = ts.DEMSceneGPU(s)
    p0 = g.get("position").copy()
    nb = g.get("bondIsBonded").size
    g.step(1500)
    ly = s.max_bound[1] - s.min_bound[1]
    p = g.get("position")
    d = p[:, 1] - (p0[:, 1] + 1500 * s.dt * 50.0)
    d -= ly * np.round(d / ly)
    assert np.abs(d).max() < 1e-9
    assert (p[:, 1] < ly + 5 * 50.0 * s.dt).all() and (p[:, 1] > -5 * 50.0 * s.dt).all()
    assert g.get("bondIsBonded").sum() == nb
    assert np.abs(g.get("velocity") - [0.0, 50.0, 0.0]).max() < 1e-9


def test_collapse_history_matches_oracle():
    """The 4,096-ball granular collapse over its full 10,000 steps against the
    FP64 oracle's history (tests/golden/collapse_history.npz, written by
    tests/golden/make_collapse_history.py), every 100 steps: translational
    kinetic energy within 0.2 % of its peak, rotational within 3 % of its
    peak (rolling starts with frictional contacts, sensitive to FP32 force
    rounding), mean and top ball height within 10 um."""
    ref = np.load(GOLDEN / "collapse_history.npz")["history"]
    s = scenes.granular_collapse()
    g = ts.DEMSceneGPU(s)
    r, rho = 0.005, 2500.0
    m = 4.0 / 3.0 * math.pi * r ** 3 * rho
    rows = []
    for k in range(ref.shape[0]):
        if k:
            g.step(100)
        v = g.get("velocity")
        w = g.get("angularVelocity")
        z = g.get("position")[:, 2]
        rows.append([0.5 * m * (v * v).sum(), 0.5 * 0.4 * m * r * r * (w * w).sum(), z.mean(), z.max()])
    h = np.array(rows)
    assert np.abs(h[:, 0] - ref[:, 1]).max() < 2e-3 * ref[:, 1].max()
    assert np.abs(h[:, 1] - ref[:, 2]).max() < 3e-2 * ref[:, 2].max()
    assert np.abs(h[:, 2] - ref[:, 3]).max() < 1e-5
    assert np.abs(h[:, 3] - ref[:, 4]).max() < 1e-5
    assert ref[-1, 1] > 0.9  # the collapse is under way at the end of the run


def test_pair_lists_grown_on_device_replay_bit_for_bit():
    """Pair lists built inside step blocks are sized on the device: a list
    that outgrows its capacity is truncated and noted, and the block is
    replayed from its device checkpoint with grown lists (no host read per
    rebuild; whole rebuild cycles replay one graph).  Lists that start at a
    quarter of what the scene needs replay and then match a scene with ample
    lists bit for bit, with graphs and with eager launches, over blocks that
    start on and off the rebuild cycle."""
    import torch

    st = torch.cuda.Stream()
    s = scenes.granular_collapse()
    a = ts.DEMSceneGPU(s, stream=st.cuda_stream)
    need = a.get("objectPointed").size
    small = []
    for graphs in (True, False):
        b = ts.DEMSceneGPU(s, solve_begin=False, stream=st.cuda_stream)
        b.set_graphs(graphs)
        b.tune("pairCapacity", max(need // 4, 1))
        # solve_begin's own search reads the total back and grows the list
        # it writes; the other list keeps the small capacity, so the first
        # rebuild inside a step block overflows
        b.solve_begin()
        assert b.get("objectPointed").size == need
        small.append(b)
    for n in (23, 40, 7, 30):
        for g in [a] + small:
            g.step(n)
        for b in small:
            for f in ("position", "velocity", "angularVelocity", "slidingSpring", "objectPointed",
                      "objectPointing"):
                assert np.array_equal(a.get(f), b.get(f)), (n, f)
            assert a.clock() == b.clock()
    assert a.tuned("pairReplays") == 0
    for b in small:
        assert b.tuned("pairReplays") >= 1
        assert b.tuned("pairCapacity") >= a.get("objectPointed").size
    assert a.graph_captures() >= 2 and small[0].graph_captures() >= 2  # rebuild cycles replay as graphs
    assert small[1].graph_captures() == 0
    for g in [a] + small:
        g.close()


def test_wall_register_slots_equal_general_wall_path():
    """Small walled scenes take ball-wall contacts through register slots
    (cells with up to four candidate triangles); the general candidate path
    must give the same wall histories and trajectories bit for bit, in FP32
    and FP64."""
    for fp64 in (False, True):
        s = scenes.granular_collapse()
        a = ts.DEMSceneGPU(s, fp64=fp64)
        b = ts.DEMSceneGPU(s, fp64=fp64, solve_begin=False)
        b.tune("wallSlots", 0)
        b.solve_begin()
        assert a.tuned("wallSlots") == 1 and b.tuned("wallSlots") == 0
        for n in (3, 597, 3000):  # the lowest layer lands after ~3,200 steps
            a.step(n)
            b.step(n)
            for f in ("position", "velocity", "angularVelocity", "slidingSpring", "triPointing", "triCancel",
                      "triSlidingSpring", "triOverlap"):
                assert np.array_equal(a.get(f), b.get(f)), (fp64, n, f)
        assert (a.get("triCancel") == 0).sum() > 50  # the lowest layer reaches the floor
        a.close()
        b.close()


def _lattice_beam_static_tip(nx, ny, nz, r=0.002, E=1e9, F=10.0, ks_ratio=2.6, rho=2500.0, dt=5e-7):
    """Static tip deflection of the simple-cubic bonded beam under a tip load
    F, from beam theory with the lattice's own stiffnesses:
    * each chain of bonds along x is a spring k = E pi r / 2 per spacing 2r
      (parallel bond of radius r, normal stiffness E / 2r per area);
    * bending: EI = k (2r)^3 ny nz (nz^2 - 1) / 12 from the chains, plus each
      bond's own bending stiffness E pi r^4 / 4;
    * shear (Timoshenko, kappa = 5/6): G = (E pi / 4) / ks_ratio over the
      section (ny 2r) x (nz 2r);
    * span L = (nx - 1) 2r from the clamped layer to the loaded layer.
    Also the first bending mode's period in steps."""
    k = E * math.pi * r / 2
    L = (nx - 1) * 2 * r
    EI = k * (2 * r) ** 3 * ny * nz * (nz * nz - 1) / 12 + ny * nz * E * math.pi * r ** 4 / 4
    G = (E * math.pi / 4) / ks_ratio
    A = (ny * 2 * r) * (nz * 2 * r)
    static = F * L ** 3 / (3 * EI) + F * L / (5 / 6 * G * A)
    mu = rho * 4 / 3 * math.pi * r ** 3 * ny * nz / (2 * r)
    w1 = 1.875104 ** 2 * math.sqrt(EI / (mu * L ** 4))
    return static, int(round(2 * math.pi / w1 / dt))


def test_full_size_cantilever_tip_deflection_is_beam_theory():
    """The 128 x 16 x 16 bonded cantilever (32,768 balls, 93,472 bonds; no
    clumps and no gravity, so the tip load alone bends it): after the
    10,000-step ramp the undamped beam oscillates about its static shape,
    and the tip deflection averaged over two periods of the first bending
    mode is the beam-theory value within 2 %.  (The same average on the
    oracle: 32 x 4 x 4 within 0.3 %, 48 x 6 x 6 within 1.1 %.)"""
    s = scenes.cantilever_with_clumps(n_clumps=0)
    s.gravity = np.zeros(3)
    g = ts.DEMSceneGPU(s)
    assert g.get("bondPointed").size == 93472
    static, T1 = _lattice_beam_static_tip(128, 16, 16)
    p0 = g.get("position")
    tip = p0[:, 0] > (128 - 1.5) * 2 * 0.002
    assert tip.sum() == 256
    g.step(10000 + T1 // 2)
    every = T1 // 64
    defl = []
    for _ in range(128):
        g.step(every)
        defl.append(-(g.get("position")[tip, 2] - p0[tip, 2]).mean())
    mean = float(np.mean(defl))
    assert abs(mean - static) / static < 0.02, (mean, static, T1)
    assert g.get("bondIsBonded").sum() == 93472  # the bonds never break (zero strengths)
    g.close()


def test_full_size_cantilever_matches_oracle():
    """The same full-size beam with its clumps and gravity, GPU in FP64
    against the FP64 oracle over the first 100 steps (the clumps fall, the
    load ramps): bonds bit-exact, positions and velocities within 1e-8 of
    their largest magnitude."""
    s = scenes.cantilever_with_clumps()
    g = ts.DEMSceneGPU(s, fp64=True)
    o = oracle.OracleDEM(s)
    assert np.array_equal(g.get("bondPointed"), o.get("bond_pointed"))
    assert np.array_equal(g.get("bondPointing"), o.get("bond_pointing"))
    g.step(100)
    o.step(100)
    for f, of in (("position", "pos"), ("velocity", "vel")):
        a, b = g.get(f), o.get(of)
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 1e-8 * scale, (f, np.abs(a - b).max(), scale)
    v = o.get("vel")
    assert np.abs(v).max() > 1e-4  # the clumps and the beam are moving
    g.close()


def _ball_on_fine_plate(square=0.001, n=20, r=0.01, gap=1e-4, v0=-0.1):
    """A ball falling onto a plate meshed into n x n squares of side `square`
    (two triangles each): at the impact its contact patch covers more
    triangles boxes than the default 16 wall candidates per ball."""
    s = scenes.DEMScene("ball on a fine plate")
    xs = (np.arange(n + 1) - n / 2) * square
    vx, vy = np.meshgrid(xs, xs, indexing="ij")
    verts = np.stack([vx.ravel(), vy.ravel(), np.zeros(vx.size)], axis=1)
    t0, t1, t2 = [], [], []
    for i in range(n):
        for j in range(n):
            a, b, c, d = i * (n + 1) + j, (i + 1) * (n + 1) + j, (i + 1) * (n + 1) + j + 1, i * (n + 1) + j + 1
            t0 += [a, a]
            t1 += [b, c]
            t2 += [c, d]
    E, nu = 1e7, 0.3
    s.set_hertzian_pair(0, 0, E / (2 * (1 - nu * nu)), E / (4 * (2 - nu) * (1 + nu)), 0.5, 0.0, 0.0, 0.3, 0.0, 0.0)
    s.add_balls([[0.0001, 0.0002, r + gap]], [0.0, 0.0, v0], [0, 0, 0], r, 2500.0, 0)
    s.add_mesh_wall(verts, t0, t1, t2, [0, 0, 0], [0, 0, 0], [0, 0, 0], 0)
    s.min_bound, s.max_bound = np.array([-0.05, -0.05, -0.05]), np.array([0.05, 0.05, 0.05])
    s.gravity = np.array([0.0, 0.0, -9.81])
    s.dt = 1e-5
    return s


@pytest.mark.parametrize("fp64", [False, True])
def test_wall_capacity_grows_inside_a_step_block(fp64):
    """The ball's impact brings more than 16 triangle candidates within
    reach: the step call goes back to its device checkpoint, the wall
    history grows to 64 entries per ball and the block runs again; the
    bounce then matches the oracle (which has no per-ball limit)."""
    s = _ball_on_fine_plate()
    g = ts.DEMSceneGPU(s, fp64=fp64)
    o = oracle.OracleDEM(s)
    assert g.tuned("wallCapacity") == 16
    for n in (150, 150, 100):
        g.step(n)
        o.step(n)
    assert g.tuned("wallCapacity") == 64 and g.tuned("wallReplays") >= 1
    tol = 1e-9 if fp64 else 1e-6
    assert np.abs(g.get("position") - o.get("pos")).max() < tol * 0.01
    assert np.abs(g.get("velocity") - o.get("vel")).max() < tol * 0.1 * 10
    assert o.get("vel")[0, 2] > 0  # it bounced
    g.close()
