"""The oracle's general convex narrow phase (oracle/general.cpp): the pairs the
reference hands to parry 0.31.1 (crates/pippin/src/collision.rs:102-147), here a
cylinder against a sphere, capsule, box or cylinder.

parry is absent from the reference tree, so its arithmetic is parity unpinned.
These tests pin the restatement in two ways:
- the reference's own physical test of the path
  (crates/pippin/tests/mesh.rs:66-83), with its tolerances;
- closed forms: depths, normals and contact points of posed pairs whose answer
  follows from the geometry. Each would catch a flipped normal, a dropped radius,
  a wrong support feature or a lost clip edge.
"""
import numpy as np
import pytest

from oracle import mjcf
from oracle import pyoracle as po
from oracle.pyoracle import OracleSim
from tests.scenes import general_xml

SPH, CAP, BOX, CYL = 1, 2, 3, 4
I3 = np.eye(3).ravel()


def rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]]).ravel()


def ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]]).ravel()


def rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]]).ravel()


@pytest.mark.parametrize("solver", [0, 1])
def test_reference_cylinders_rest_upright_on_side_and_stacked(solver):
    """crates/pippin/tests/mesh.rs:66-83 (there with the default 0.002 s step)."""
    xml = general_xml("cylinders_on_box").replace('timestep="0.004"', 'timestep="0.002"')
    arr = mjcf.load_string(xml)
    arr["solver_kind"] = solver
    s = OracleSim(arr, 1)
    for _ in range(1500):
        s.step(0, 1, 1)
    q = s.get("qpos")[0]
    assert abs(q[2] - 0.06) < 2e-3, q[2]
    assert abs(q[9] - 0.04) < 2e-3, q[9]
    assert abs(q[16] - (0.12 + 0.03)) < 3e-3, q[16]


@pytest.mark.parametrize("gap", [5e-4, -3e-3, -1.5e-2])
def test_sphere_over_cap(gap):
    h = po.collide(CYL, [0, 0, 0], I3, [.04, .06, 0], SPH, [0.01, 0.005, .08 + gap], I3, [.02, 0, 0])
    assert len(h) == 1
    assert h[0]["depth"] == pytest.approx(-gap, abs=1e-12)
    assert np.allclose(h[0]["normal"], [0, 0, 1], atol=1e-12)
    # midpoint of the cap point and the sphere's lowest point
    assert np.allclose(h[0]["pos"], [0.01, 0.005, 0.06 + gap / 2], atol=1e-12)


@pytest.mark.parametrize("gap", [5e-4, -3e-3])
def test_sphere_beside_side_and_order_flips_normal(gap):
    a = po.collide(SPH, [.06 + gap, 0, .01], I3, [.02, 0, 0], CYL, [0, 0, 0], I3, [.04, .06, 0])
    b = po.collide(CYL, [0, 0, 0], I3, [.04, .06, 0], SPH, [.06 + gap, 0, .01], I3, [.02, 0, 0])
    assert len(a) == 1 and len(b) == 1
    assert a[0]["depth"] == pytest.approx(-gap, abs=1e-9)
    assert np.allclose(a[0]["normal"], [-1, 0, 0], atol=1e-9)
    assert np.allclose(b[0]["normal"], [1, 0, 0], atol=1e-9)
    assert np.allclose(a[0]["pos"], [0.04 + gap / 2, 0, 0.01], atol=1e-9)
    assert b[0]["depth"] == pytest.approx(a[0]["depth"], abs=1e-9)


def test_sphere_far_away_and_within_margin():
    assert len(po.collide(CYL, [0, 0, 0], I3, [.04, .06, 0], SPH, [0, 0, .0815], I3, [.02, 0, 0], margin=1e-3)) == 0
    h = po.collide(CYL, [0, 0, 0], I3, [.04, .06, 0], SPH, [0, 0, .0805], I3, [.02, 0, 0], margin=1e-3)
    assert len(h) == 1 and h[0]["depth"] == pytest.approx(-5e-4, abs=1e-12)


@pytest.mark.parametrize("yaw", [0.0, 0.3, 1.1])
def test_upright_cylinder_on_box_four_rim_points(yaw):
    pen = 2e-3
    h = po.collide(BOX, [0, 0, 0], I3, [.05, .05, .05], CYL, [0.01, 0, .05 + .06 - pen], rz(yaw), [.03, .06, 0])
    assert len(h) == 4
    assert np.allclose(h["depth"], pen, atol=1e-12)
    assert np.allclose(h["normal"], [0, 0, 1], atol=1e-12)
    # the cap's square: four points on the rim, a quarter turn apart
    rim = h["pos"][:, :2] - [0.01, 0]
    assert np.allclose(np.linalg.norm(rim, axis=1), 0.03, atol=1e-12)
    assert np.allclose(h["pos"][:, 2], 0.05 - pen / 2, atol=1e-12)
    assert np.allclose(rim.sum(axis=0), 0, atol=1e-12)


def test_cylinder_on_its_side_on_box_line_ends():
    pen = 1e-3
    h = po.collide(BOX, [0, 0, 0], I3, [.2, .2, .05], CYL, [0, 0, .05 + .04 - pen], rx(np.pi / 2), [.04, .06, 0])
    assert len(h) == 2
    assert np.allclose(h["depth"], pen, atol=1e-12)
    assert np.allclose(sorted(h["pos"][:, 1]), [-0.06, 0.06], atol=1e-12)
    # the side meets the box through EPA on a curved surface: 1e-9 tolerance
    assert np.allclose(h["pos"][:, [0, 2]], [0, 0.05 - pen / 2], atol=1e-9)


def test_capsule_across_cap_clipped_to_the_cap():
    pen = 2e-3
    h = po.collide(CYL, [0, 0, 0], I3, [.04, .06, 0], CAP, [0, 0, .06 + .01 - pen], rx(np.pi / 2), [.01, .05, 0])
    assert len(h) == 2
    assert np.allclose(h["depth"], pen, atol=1e-12)
    assert np.allclose(h["normal"], [0, 0, 1], atol=1e-12)
    # the capsule's segment (|y| <= .05) clipped by the cap's square (|y| <= .04)
    assert np.allclose(sorted(h["pos"][:, 1]), [-0.04, 0.04], atol=1e-12)


def test_parallel_cylinders_overlap_interval():
    pen = 1e-3
    h = po.collide(CYL, [0, 0, 0], I3, [.04, .06, 0], CYL, [.08 - pen, 0, .02], I3, [.04, .06, 0])
    assert len(h) == 2
    # EPA on curved sides converges to its 1e-9 tolerance
    assert np.allclose(h["depth"], pen, atol=2e-9)
    assert np.allclose(h["normal"], [[1, 0, 0]] * 2, atol=1e-3)
    assert np.allclose(sorted(h["pos"][:, 2]), [-0.04, 0.06], atol=1e-6)


def test_general_box_contact_matches_the_plane_primitive():
    """An upright cylinder sunk into a large box (general path) and into a
    plane at the box's top (primitive plane_cylinder, collision.rs:238) have
    the same depth."""
    pen = 1.5e-3
    cpos = [0.02, -0.01, .06 - pen]
    g = po.collide(BOX, [0, 0, -.1], I3, [1, 1, .1], CYL, cpos, I3, [.03, .06, 0])
    p = po.collide(0, [0, 0, 0], I3, [5, 5, .1], CYL, cpos, I3, [.03, .06, 0])
    assert len(g) == 4 and len(p) >= 1
    assert np.allclose(g["depth"], p["depth"].max(), atol=1e-12)


def test_crossed_cylinders_side_to_side():
    """Perpendicular axes: the sides touch at one point, depth = r1 + r2 - d."""
    r1, r2, d = 0.04, 0.03, 0.068
    h = po.collide(CYL, [0, 0, 0], rx(np.pi / 2), [r1, .1, 0], CYL, [0, 0, d], ry(np.pi / 2), [r2, .1, 0])
    assert len(h) >= 1
    assert np.allclose(h["depth"], r1 + r2 - d, atol=2e-9)
    # both sides curved: the normal within the EPA stop's angle, as above
    assert np.allclose(h["normal"], [[0, 0, 1]] * len(h), atol=3e-4)
    assert np.allclose(h["pos"][:, :2], 0, atol=1e-4)


@pytest.mark.parametrize("where", ["below_cap", "inside_side"])
def test_sphere_centre_inside_cylinder(where):
    """Deep penetration (the centre inside the cylinder): EPA finds the
    nearest face; depth = distance to it + radius."""
    r = 0.01
    if where == "below_cap":
        c, depth, n = [0.0, 0.0, 0.055], (0.06 - 0.055) + r, [0, 0, 1]
    else:
        c, depth, n = [0.035, 0.0, 0.0], (0.04 - 0.035) + r, [1, 0, 0]
    h = po.collide(CYL, [0, 0, 0], I3, [.04, .06, 0], SPH, c, I3, [r, 0, 0])
    assert len(h) == 1
    assert h[0]["depth"] == pytest.approx(depth, abs=2e-9)
    # EPA stops once the support gains <= 1e-9 m; on the curved side that
    # leaves the normal's angle within sqrt(2 * 1e-9 / 0.04) ~ 2.2e-4 rad
    assert np.allclose(h[0]["normal"], n, atol=3e-4)
