"""URDF import (crates/pippin/src/urdf.rs) through the library's loader
(pippin_hip_model_load_urdf) and the oracle's (oracle/urdf.py), CPU only.

Both loaders must give the same raw model array for array; the conversion's
rules are checked against what URDF fixes: rpy as R = Rz(y) Ry(p) Rx(r), box
sizes as full extents, cylinder and capsule lengths as full lengths, inertia
as given in the inertial frame, joint types and limits, visual geometry that
neither collides nor weighs, collision geometry hidden in group 3 under
visuals, the root welded to the world unless floating, and mesh files found
by package://, file:// and relative paths.
"""
import numpy as np
import pytest

from oracle import urdf as ourdf
from oracle.mjcf import MjcfError
from pippin_amd import physics

WEDGE_OBJ = "v 0 0 0\nv 0.1 0 0\nv 0 0.1 0\nv 0 0 0.1\nv 0.1 0.1 0\nf 1 3 2\nf 2 3 5\nf 1 2 4\nf 1 4 3\nf 2 5 4\nf 3 4 5\n"

ARM = """<robot name="arm3">
  <material name="blue"><color rgba="0.1 0.2 0.9 1"/></material>
  <link name="base">
    <collision><origin xyz="0 0 -0.05"/><geometry><box size="2 2 0.1"/></geometry></collision>
  </link>
  <link name="upper">
    <inertial><origin xyz="0 0 0.15" rpy="0.1 -0.2 0.3"/><mass value="1.5"/>
      <inertia ixx="0.02" iyy="0.03" izz="0.01" ixy="0.001" ixz="-0.002" iyz="0.0005"/></inertial>
    <collision><origin xyz="0 0 0.15"/><geometry><capsule radius="0.04" length="0.3"/></geometry></collision>
    <visual><origin xyz="0 0 0.15"/><geometry><cylinder radius="0.04" length="0.3"/></geometry><material name="blue"/></visual>
  </link>
  <link name="fore">
    <collision><origin xyz="0 0 0.1" rpy="0 1.5707963267948966 0"/><geometry><cylinder radius="0.03" length="0.2"/></geometry></collision>
    <visual><geometry><sphere radius="0.05"/></geometry><material name="x"><color rgba="1 0 0 0.5"/></material></visual>
  </link>
  <link name="slider"><collision><geometry><sphere radius="0.05"/></geometry></collision></link>
  <link name="tool"><collision><geometry><box size="0.05 0.06 0.07"/></geometry></collision></link>
  <link name="spin"><collision><origin xyz="0.02 0 0"/><geometry><sphere radius="0.02"/></geometry></collision></link>
  <joint name="shoulder" type="revolute"><parent link="base"/><child link="upper"/>
    <origin xyz="0 0 0.4" rpy="0.3 0.2 -0.7"/><axis xyz="0 1 0"/><limit lower="-1.5" upper="1.2"/><dynamics damping="0.7"/></joint>
  <joint name="elbow" type="continuous"><parent link="upper"/><child link="fore"/>
    <origin xyz="0 0 0.3"/><axis xyz="0 1 0"/></joint>
  <joint name="slide" type="prismatic"><parent link="fore"/><child link="slider"/>
    <origin xyz="0 0 0.2"/><axis xyz="0 0 1"/><limit lower="-0.05" upper="0.1"/></joint>
  <joint name="weld" type="fixed"><parent link="slider"/><child link="tool"/><origin xyz="0.1 0 0"/></joint>
  <joint name="free_spin" type="revolute"><parent link="tool"/><child link="spin"/><axis xyz="1 0 0"/></joint>
</robot>"""


def _write(tmp_path, text, name="robot.urdf"):
    p = tmp_path / name
    p.write_text(text)
    return p


def _both(path, **kw):
    a = physics.load_urdf(path, **kw)
    b = ourdf.load_file(path, **kw)
    return a, b


def _assert_equal(a, b):
    assert a.keys() == b.keys()
    for k in a:
        va, vb = np.asarray(a[k]).reshape(-1), np.asarray(b[k]).reshape(-1)
        assert va.size == vb.size, k
        assert np.array_equal(va.astype(np.float64), vb.astype(np.float64), equal_nan=True), k


def _quat_mat(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _rpy_mat(r, p, y):
    cx, sx, cy, sy, cz, sz = np.cos(r), np.sin(r), np.cos(p), np.sin(p), np.cos(y), np.sin(y)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


@pytest.mark.parametrize("floating", [False, True])
def test_arm_loaders_agree_and_follow_urdf(tmp_path, floating):
    a, b = _both(_write(tmp_path, ARM), floating_base=floating)
    _assert_equal(a, b)
    # bodies: world, base, upper, fore, slider, tool, spin
    assert int(a["nbody"]) == 7
    jt = list(np.asarray(a["jnt_type"]))
    hinge, slide, free = 3, 2, 0
    assert jt == ([free] if floating else []) + [hinge, hinge, slide, hinge]
    lim = list(np.asarray(a["jnt_limited"]).astype(bool))
    rng = np.asarray(a["jnt_range"]).reshape(-1, 2)
    off = 1 if floating else 0
    assert lim[off:] == [True, False, True, False]  # lower < upper gives a range; none for continuous or 0..0
    assert tuple(rng[off]) == (-1.5, 1.2) and tuple(rng[off + 2]) == (-0.05, 0.1)
    assert float(np.asarray(a["dof_damping"])[6 if floating else 0]) == 0.7  # the shoulder's dof
    # the shoulder's origin rpy: R = Rz Ry Rx
    bq = np.asarray(a["body_quat"]).reshape(-1, 4)
    assert np.allclose(_quat_mat(bq[2]), _rpy_mat(0.3, 0.2, -0.7), atol=1e-14)
    # sizes: box full extents, capsule and cylinder full length
    gt = np.asarray(a["geom_type"])
    gs = np.asarray(a["geom_size"]).reshape(-1, 3)
    assert tuple(gs[0]) == (1.0, 1.0, 0.05) and gt[0] == 3
    assert gt[1] == 2 and tuple(gs[1][:2]) == (0.04, 0.15)  # upper's collision capsule
    assert gt[2] == 4 and tuple(gs[2][:2]) == (0.04, 0.15)  # upper's visual cylinder
    # visual geoms neither collide nor weigh; collision geoms under visuals sit in group 3
    ct, ca, grp = np.asarray(a["geom_contype"]), np.asarray(a["geom_conaffinity"]), np.asarray(a["geom_group"])
    assert ct[2] == 0 and ca[2] == 0 and grp[2] == 1 and grp[1] == 3 and grp[0] == 0
    rgba = np.asarray(a["geom_rgba"]).reshape(-1, 4)
    assert np.allclose(rgba[2], [0.1, 0.2, 0.9, 1.0]) and np.allclose(rgba[4], [1, 0, 0, 0.5])
    # the upper link's inertial: mass as given, inertia rotated into the body frame
    mass = np.asarray(a["body_mass"])
    assert mass[2] == 1.5
    I = np.array([[0.02, 0.001, -0.002], [0.001, 0.03, 0.0005], [-0.002, 0.0005, 0.01]])
    R = _rpy_mat(0.1, -0.2, 0.3)
    inertia = np.asarray(a["body_inertia"]).reshape(-1, 3, 3)
    assert np.allclose(inertia[2], R @ I @ R.T, atol=1e-15)
    assert np.allclose(np.asarray(a["body_ipos"]).reshape(-1, 3)[2], [0, 0, 0.15])
    # the fore link has no inertial: its mass is its collision cylinder's, not its visual sphere's
    assert abs(mass[3] - 1000.0 * np.pi * 0.03 ** 2 * 0.2) < 1e-12


def test_meshes_by_package_file_and_relative_paths(tmp_path, monkeypatch):
    monkeypatch.delenv("ROS_PACKAGE_PATH", raising=False)
    pkg = tmp_path / "share" / "my_robot"
    (pkg / "meshes").mkdir(parents=True)
    (pkg / "meshes" / "wedge.obj").write_text(WEDGE_OBJ)
    (pkg / "meshes" / "ignored.dae").write_text("<COLLADA/>")
    udir = tmp_path / "robots"
    udir.mkdir()
    text = f"""<robot name="m">
      <link name="root"><collision><geometry><box size="1 1 0.1"/></geometry></collision></link>
      <link name="a">
        <collision><geometry><mesh filename="package://my_robot/meshes/wedge.obj"/></geometry></collision>
        <collision><origin xyz="0.2 0 0"/><geometry><mesh filename="package://my_robot/meshes/wedge.obj"/></geometry></collision>
        <collision><geometry><mesh filename="package://my_robot/meshes/wedge.obj" scale="2 2 2"/></geometry></collision>
        <collision><geometry><mesh filename="file://{pkg}/meshes/wedge.obj"/></geometry></collision>
        <collision><geometry><mesh filename="../share/my_robot/meshes/wedge.obj"/></geometry></collision>
        <collision><geometry><mesh filename="package://my_robot/meshes/ignored.dae"/></geometry></collision>
        <collision><geometry><mesh filename="package://other/missing.stl"/></geometry></collision>
      </link>
      <joint name="j" type="floating"><parent link="root"/><child link="a"/><origin xyz="0 0 0.5"/></joint>
    </robot>"""
    path = _write(udir, text)
    a, b = _both(path, package_dirs=[tmp_path / "share"])
    _assert_equal(a, b)
    # one asset per (file, scale): unscaled and scaled
    assert int(a["nmesh"]) == 2
    assert list(np.asarray(a["geom_type"])).count(5) == 5  # the .dae and the missing file are skipped
    _, warnings = ourdf.to_mjcf(path.read_text(), str(udir), package_dirs=[tmp_path / "share"])
    assert any(".dae not supported" in w for w in warnings) and any("not found" in w for w in warnings)
    # the conventional layout next to the URDF file needs no package_dirs
    c = physics.load_urdf(_write(tmp_path / "share", text.replace("../share/", "")))
    assert int(c["nmesh"]) == 2


def test_mjcf_texts_of_both_converters_load_alike(tmp_path):
    base = str(tmp_path)
    mine = physics.urdf_to_mjcf(ARM, base)
    theirs, _ = ourdf.to_mjcf(ARM, base)
    from oracle import mjcf
    _assert_equal(physics.load_model(mine, is_xml=True), mjcf.load_string(theirs, base))
    assert mine.startswith('<mujoco model="arm3"><compiler angle="radian" autolimits="true"/>')


@pytest.mark.parametrize("text,msg", [
    ("<mujoco/>", "root element must be <robot>"),
    ('<robot><link name="a"/><link name="b"/></robot>', "exactly one root link"),
    ('<robot><link name="a"/><link name="b"/><link name="c"/>'
     '<joint name="j1" type="fixed"><parent link="a"/><child link="c"/></joint>'
     '<joint name="j2" type="fixed"><parent link="b"/><child link="c"/></joint></robot>', "two parent joints"),
    ('<robot><link name="a"/><joint name="j" type="fixed"><parent link="a"/><child link="zz"/></joint></robot>',
     "references unknown link"),
    ('<robot><link name="a"/><joint name="j" type="fixed"><parent link="a"/></joint></robot>', "missing parent/child"),
])
def test_invalid_urdf_is_refused_by_both(tmp_path, text, msg):
    path = _write(tmp_path, text)
    with pytest.raises(physics.HipError, match=msg):
        physics.load_urdf(path)
    with pytest.raises(MjcfError, match=msg):
        ourdf.load_file(path)


def test_rpy_quaternion_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(50):
        r, p, y = rng.uniform(-3.1, 3.1, 3)
        q = ourdf.rpy_quat([r, p, y])
        assert abs(np.linalg.norm(q) - 1.0) < 1e-15
        assert np.allclose(_quat_mat(q), _rpy_mat(r, p, y), atol=1e-14)
