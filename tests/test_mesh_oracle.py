"""oracle/mesh.py against the reference's known answers for meshes and the
invariants of a convex hull. (The device does not load meshes yet: this is the
oracle side of that row, pinned before the kernels that will be checked
against it.)"""

import math

import numpy as np
import pytest

from oracle import mesh


def _cube(h):
    # crates/pippin/src/mesh.rs:111-116: corner i has x = +-h by bit 0, y by bit 1, z by bit 2
    return [[h if i & 1 else -h, h if i & 2 else -h, h if i & 4 else -h] for i in range(8)]


def test_offset_cube_mass_properties_are_exact():
    """mesh.rs:117-134."""
    h = 0.3
    offset = np.array([1.0, -2.0, 0.5])
    V, F = mesh.convex_hull(_cube(h))
    V = V + offset
    vol, com, inertia = mesh.mass_properties(V, mesh.triangles(F))
    side = 2.0 * h
    assert abs(vol - side ** 3) < 1e-12
    assert np.linalg.norm(com - offset) < 1e-12
    expected = vol * side * side / 6.0
    for k in range(3):
        assert abs(inertia[k, k] - expected) < 1e-12
    assert abs(inertia[0, 1]) < 1e-12 and abs(inertia[0, 2]) < 1e-12 and abs(inertia[1, 2]) < 1e-12


def test_mesh_cube_mass_and_inertia_equal_the_box():
    """tests/mesh.rs:24-45: a 0.1 m mesh cube and a box of half-size 0.05 at
    density 1000 (box volume and inertia from mjcf.rs:783-787)."""
    V, F = mesh.convex_hull(_cube(0.05))
    vol, com, icom = mesh.mass_properties(V, mesh.triangles(F))
    rho = 1000.0
    a = b = c = 0.05
    bv = 8.0 * a * b * c
    bi = [bv * (b * b + c * c) / 3.0, bv * (a * a + c * c) / 3.0, bv * (a * a + b * b) / 3.0]
    assert abs(rho * vol - rho * bv) < 1e-12
    assert np.linalg.norm(com) < 1e-15
    for k in range(3):
        assert abs(rho * icom[k, k] - rho * bi[k]) < 1e-12


def test_scaled_obj_wedge():
    """tests/mesh.rs:85-116: the OBJ wedge scaled by 2 spans x up to 0.2 (float32
    coordinates) and has the volume of a square pyramid, base 0.2 x 0.2, height 0.2."""
    obj = "v 0 0 0\nv 0.1 0 0\nv 0 0.1 0\nv 0 0 0.1\nv 0.1 0.1 0\nf 1 3 2\nf 2 3 5\nf 1 2 4\nf 1 4 3\nf 2 5 4\nf 3 4 5\n"
    pts = mesh.read_obj_vertices(obj, (2.0, 2.0, 2.0))
    V, F = mesh.convex_hull(pts)
    assert abs(V[:, 0].max() - 0.2) < 1e-6
    vol, _, _ = mesh.mass_properties(V, mesh.triangles(F))
    assert abs(vol - 0.2 * 0.2 * 0.2 / 3.0) < 1e-8


def _lumpy(n):
    # points on a lumpy ellipsoid, after fasthull.rs:221-240 (fewer points: brute-force hull)
    pts = []
    for i in range(n):
        t, p = i * 2.399963, math.acos(1.0 - 2.0 * (i + 0.5) / n)
        r = 1.0 + 0.2 * math.sin(3.0 * t) * math.cos(2.0 * p)
        pts.append([r * math.sin(p) * math.cos(t), r * math.sin(p) * math.sin(t), 0.7 * r * math.cos(p)])
    return np.array(pts)


@pytest.mark.parametrize("pts", [_cube(0.02), _lumpy(40), _lumpy(24) * 0.03])
def test_hull_invariants(pts):
    P = np.asarray(pts, dtype=np.float64)
    V, F = mesh.convex_hull(P)
    # every point on or inside every face plane, faces planar and convex, outward
    for f in F:
        nrm = np.cross(V[f[1]] - V[f[0]], V[f[2]] - V[f[0]])
        nrm /= np.linalg.norm(nrm)
        assert np.all((P - V[f[0]]) @ nrm <= 1e-9)
        assert np.all(np.abs((V[f] - V[f[0]]) @ nrm) <= 1e-9)
        for k in range(len(f)):
            a, b, c = V[f[k]], V[f[(k + 1) % len(f)]], V[f[(k + 2) % len(f)]]
            assert np.cross(b - a, c - b) @ nrm > 0
    # Euler: V - E + F = 2, each edge shared by two faces
    edges = [tuple(sorted((f[k], f[(k + 1) % len(f)]))) for f in F for k in range(len(f))]
    assert all(edges.count(e) == 2 for e in set(edges))
    assert len(V) - len(set(edges)) + len(F) == 2
    # support over the hull equals a scan over all points (fasthull.rs:221-240)
    for i in range(200):
        a = i * 0.731
        d = np.array([math.sin(a) * math.cos(1.3 * a), math.cos(0.7 * a), math.sin(2.1 * a)])
        assert abs((V @ d).max() - (P @ d).max()) < 1e-12


def test_hull_volume_matches_a_tetrahedra_sum_and_is_order_free():
    """The volume of the lumpy hull is the same for any order of the input
    points and equals a fan of tetrahedra from the centroid."""
    P = _lumpy(30)
    V, F = mesh.convex_hull(P)
    vol, com, _ = mesh.mass_properties(V, mesh.triangles(F))
    rng = np.random.default_rng(3)
    V2, F2 = mesh.convex_hull(P[rng.permutation(len(P))])
    assert np.array_equal(V, V2) and F == F2
    c0 = V.mean(axis=0)
    ref = sum(abs(np.dot(V[a] - c0, np.cross(V[b] - c0, V[d] - c0))) / 6.0 for a, b, d in mesh.triangles(F))
    assert abs(vol - ref) < 1e-12
    assert vol > 0


def test_flat_points_have_no_hull():
    assert mesh.convex_hull([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]) is None

