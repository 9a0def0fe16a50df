"""The oracle's mesh (convex hull) narrow phase against the reference's own mesh
tests and closed forms (CPU only).

Restated from crates/pippin/tests/mesh.rs:
  mesh_cube_matches_box_mass_and_rests   (:22-45, both solvers)
  mesh_cubes_stack                       (:47-64)
  mesh_from_obj_file_with_scale          (:85-116)
Closed forms: a mesh cube resting on a plane or touched by a sphere, a box, a
capsule and another mesh cube at known poses gives the depths, normals and
points the geometry fixes; a plane's hull contacts equal the primitive
plane-box ones; a sphere over a hull face equals sphere-box.
"""
import numpy as np
import pytest

from oracle import mjcf
from oracle.pyoracle import OracleSim

CUBE = ("-0.05 -0.05 -0.05  0.05 -0.05 -0.05  -0.05 0.05 -0.05  0.05 0.05 -0.05 "
        "-0.05 -0.05 0.05  0.05 -0.05 0.05  -0.05 0.05 0.05  0.05 0.05 0.05")
both = pytest.mark.parametrize("solver", [0, 1])


def sim(xml, solver=0, n=1, directory="."):
    arr = mjcf.load_string(xml, directory)
    arr["solver_kind"] = solver
    return OracleSim(arr, n), arr


def body_pos(s, arr, b):
    """World position of body b's origin for free bodies (qpos of its joint)."""
    j = int(np.asarray(arr["body_jntadr"])[b])
    a = int(np.asarray(arr["jnt_qposadr"])[j])
    return s.get("qpos")[0, a:a + 3]


@both
def test_mesh_cube_matches_box_mass_and_rests(solver):
    xml = f"""<mujoco><asset><mesh name="cube" vertex="{CUBE}"/></asset><worldbody>
        <geom type="plane" size="5 5 .1"/>
        <body name="m" pos="0 0 0.3" euler="4 7 0"><freejoint/><geom type="mesh" mesh="cube"/></body>
        <body name="b" pos="1 0 0.3" euler="4 7 0"><freejoint/><geom type="box" size=".05 .05 .05"/></body>
    </worldbody></mujoco>"""
    s, arr = sim(xml, solver)
    mass = np.asarray(arr["body_mass"])
    inertia = np.asarray(arr["body_inertia"]).reshape(-1, 9)
    assert abs(mass[1] - mass[2]) < 1e-12
    assert np.max(np.abs(inertia[1] - inertia[2])) < 1e-12
    s.step(0, 1, int(round(3.0 / 0.002)))
    pm, pb = body_pos(s, arr, 1), body_pos(s, arr, 2)
    assert abs(pm[2] - pb[2]) < 1e-3, (pm, pb)
    assert abs(pm[2] - 0.05) < 2e-3
    assert np.linalg.norm(s.get("qvel")[0]) < 1e-3


def test_mesh_cubes_stack():
    xml = f'<mujoco><asset><mesh name="cube" vertex="{CUBE}"/></asset><worldbody><geom type="plane" size="5 5 .1"/>'
    for i in range(4):
        xml += (f'<body name="c{i}" pos="0 0 {0.05 + 0.101 * i}" euler="0 0 {9 * i}"><freejoint/>'
                f'<geom type="mesh" mesh="cube"/></body>')
    xml += "</worldbody></mujoco>"
    s, arr = sim(xml)
    s.step(0, 1, int(round(5.0 / 0.002)))
    top = body_pos(s, arr, 4)
    assert abs(top[0]) < 5e-3 and abs(top[1]) < 5e-3, top
    assert abs(top[2] - 0.35) < 1e-2, top


@both
def test_mesh_from_obj_file_with_scale(tmp_path, solver):
    (tmp_path / "wedge.obj").write_text(
        "v 0 0 0\nv 0.1 0 0\nv 0 0.1 0\nv 0 0 0.1\nv 0.1 0.1 0\nf 1 3 2\nf 2 3 5\nf 1 2 4\nf 1 4 3\nf 2 5 4\nf 3 4 5\n")
    xml = """<mujoco><asset><mesh file="wedge.obj" scale="2 2 2"/></asset><worldbody>
        <geom type="plane" size="5 5 .1"/>
        <body name="w" pos="0 0 0.5"><freejoint/><geom type="mesh" mesh="wedge"/></body>
    </worldbody></mujoco>"""
    s, arr = sim(xml, solver, directory=tmp_path)
    hv = np.asarray(arr["mesh_vert"]).reshape(-1, 3)
    assert abs(hv[:, 0].max() - 0.2) < 1e-6  # OBJ coordinates are f32
    from oracle import mesh
    faces = [list(np.asarray(arr["face_vert"])[a:a + k]) for a, k in zip(arr["face_vertadr"], arr["face_vertnum"])]
    vol, _, _ = mesh.mass_properties(hv, mesh.triangles(faces))
    assert abs(vol - 0.2 * 0.2 * 0.2 / 3.0) < 1e-8
    s.step(0, 1, int(round(3.0 / 0.002)))
    assert np.linalg.norm(s.get("qvel")[0]) < 1e-3, "wedge did not settle"
    q = s.get("qpos")[0]
    w, x, y, z = q[3:7]
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    lowest = min((q[:3] + R @ v)[2] for v in hv)
    assert abs(lowest) < 2e-3, lowest


# ---------------------------------------------------------------- closed forms
def _contacts_at(xml, qpos):
    """Contacts of the first step from the given pose (zero gravity, at rest):
    the narrow phase runs on exactly that pose."""
    arr = mjcf.load_string(xml)
    arr["gravity"] = np.zeros(3)
    s = OracleSim(arr, 1)
    s.set("qpos", 0, 1, np.asarray(qpos, dtype=float)[None])
    s.step(0, 1, 1)
    return s.contacts(0)


def _scene(other, mesh_first=False):
    body_mesh = '<body><freejoint/><geom type="mesh" mesh="cube"/></body>'
    body_other = f"<body><freejoint/>{other}</body>"
    bodies = body_mesh + body_other if mesh_first else body_other + body_mesh
    return (f'<mujoco><asset><mesh name="cube" vertex="{CUBE}"/></asset><worldbody>'
            f'{bodies}</worldbody></mujoco>')


def test_plane_hull_contacts_equal_plane_box():
    q = [0.01, -0.02, 0.046, 1, 0, 0, 0]
    plane = '<geom type="plane" size="5 5 .1"/>'
    hull = _contacts_at(f'<mujoco><asset><mesh name="cube" vertex="{CUBE}"/></asset><worldbody>{plane}'
                        f'<body><freejoint/><geom type="mesh" mesh="cube"/></body></worldbody></mujoco>', q)
    box = _contacts_at(f'<mujoco><worldbody>{plane}<body><freejoint/><geom type="box" size=".05 .05 .05"/>'
                       f'</body></worldbody></mujoco>', q)
    assert len(hull) == len(box) == 4
    key = lambda c: tuple(np.round(c["pos"], 9))
    for h, b in zip(sorted(hull, key=key), sorted(box, key=key)):
        assert abs(h["depth"] - 0.004) < 1e-15 and abs(b["depth"] - 0.004) < 1e-15
        assert np.allclose(h["normal"], [0, 0, 1], atol=0) and np.allclose(h["pos"], b["pos"], atol=1e-15)
        assert abs(h["pos"][2] - (-0.002)) < 1e-15  # midway between the corner and the plane


@pytest.mark.parametrize("mesh_first", [False, True])
def test_sphere_over_a_hull_face(mesh_first):
    """A sphere (r 0.03) over the top face of the mesh cube at 0.05 + 0.03 - 0.004:
    depth 0.004, normal +z from the cube to the sphere (or -z in the other
    order), the point midway between the two surfaces."""
    r, pen = 0.03, 0.004
    zs = 0.05 + r - pen
    qm, qs = [0, 0, 0, 1, 0, 0, 0], [0.012, -0.02, zs, 1, 0, 0, 0]
    qpos = qm + qs if mesh_first else qs + qm
    cs = _contacts_at(_scene(f'<geom type="sphere" size="{r}"/>', mesh_first), qpos)
    assert len(cs) == 1
    c = cs[0]
    assert abs(c["depth"] - pen) < 1e-12
    assert np.allclose(c["normal"], [0, 0, 1 if mesh_first else -1], atol=1e-12)
    assert np.allclose(c["pos"], [0.012, -0.02, 0.05 - pen / 2], atol=1e-12)


def test_box_face_on_hull_face():
    """A box (half 0.03, 0.02, 0.01) lying on the mesh cube's top face with
    penetration 0.002: four hits at its bottom corners, depth 0.002, normal +z."""
    qm, qb = [0, 0, 0, 1, 0, 0, 0], [0.01, 0.005, 0.05 + 0.01 - 0.002, 1, 0, 0, 0]
    cs = _contacts_at(_scene('<geom type="box" size=".03 .02 .01"/>', mesh_first=True), qm + qb)
    assert len(cs) == 4
    for c in cs:
        assert abs(c["depth"] - 0.002) < 1e-12
        assert np.allclose(c["normal"], [0, 0, 1], atol=1e-12)
    xy = sorted((round(c["pos"][0], 9), round(c["pos"][1], 9)) for c in cs)
    assert xy == sorted((round(0.01 + sx * 0.03, 9), round(0.005 + sy * 0.02, 9)) for sx in (-1, 1) for sy in (-1, 1))


def test_hull_on_hull_offset():
    """Two mesh cubes, the upper one shifted by (0.03, 0.02) and sunk 0.001:
    the overlap of the two faces is a 0.07 x 0.08 rectangle, four hits of
    depth 0.001 at its corners."""
    xml = (f'<mujoco><asset><mesh name="cube" vertex="{CUBE}"/></asset><worldbody>'
           '<body><freejoint/><geom type="mesh" mesh="cube"/></body>'
           '<body><freejoint/><geom type="mesh" mesh="cube"/></body></worldbody></mujoco>')
    cs = _contacts_at(xml, [0, 0, 0, 1, 0, 0, 0, 0.03, 0.02, 0.099, 1, 0, 0, 0])
    assert len(cs) == 4
    for c in cs:
        assert abs(c["depth"] - 0.001) < 1e-12
        assert np.allclose(c["normal"], [0, 0, 1], atol=1e-12)
    xs = sorted({round(c["pos"][0], 9) for c in cs})
    ys = sorted({round(c["pos"][1], 9) for c in cs})
    assert xs == [-0.02, 0.05] and ys == [-0.03, 0.05]


def test_capsule_across_a_hull_face():
    """A horizontal capsule (r 0.01, half-length 0.02) lying across the top
    face, sunk 0.003: two hits at its segment's ends, depth 0.003."""
    qm = [0, 0, 0, 1, 0, 0, 0]
    s2 = np.sqrt(0.5)
    qc = [0.0, 0.01, 0.05 + 0.01 - 0.003, s2, 0, s2, 0]  # capsule axis along x
    cs = _contacts_at(_scene('<geom type="capsule" size=".01 .02"/>', mesh_first=True), qm + qc)
    assert len(cs) == 2
    for c in cs:
        assert abs(c["depth"] - 0.003) < 1e-12
        assert np.allclose(c["normal"], [0, 0, 1], atol=1e-12)
    assert sorted(round(c["pos"][0], 9) for c in cs) == [-0.02, 0.02]
