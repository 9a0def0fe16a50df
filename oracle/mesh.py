"""ORACLE — test infrastructure only. Nothing in the product may import it.

Meshes as the reference compiles them (crates/pippin/src/mesh.rs):
- ``read_obj_vertices``: the vertex positions of an OBJ file, read as float32
  the way tobj hands them over (mesh.rs:42-53), then widened to float64;
- ``convex_hull``: the hull the reference takes from parry3d-f64 0.31.1
  (``transformation::convex_hull``, mesh.rs:63-69). parry is not in the
  reference tree, so this restates the hull as a set: its extreme points
  (sorted by x, y, z) and its faces as convex polygons wound counter-clockwise
  seen from outside. The vertex and triangle order of parry's output is
  parity unpinned; the set it describes is fixed by the mathematics;
- ``triangles``: the faces fanned into outward-wound triangles;
- ``mass_properties``: volume, centre of mass and inertia about the centre of
  mass per unit density by signed tetrahedra, mesh.rs:72-105, operation for
  operation (no fused multiply-adds, sums in the reference's order).

Pinned by the reference's known answers: mesh.rs:117-134 (an offset cube's
mass properties exact to 1e-12), tests/mesh.rs:24-45 (a mesh cube's mass and
inertia equal a box's to 1e-12), tests/mesh.rs:85-116 (a scaled OBJ wedge's
extent and volume), and by the hull's own invariants (every point inside every
face plane, Euler's formula, support equal to a scan over all points as in
fasthull.rs:221-240).
"""
from __future__ import annotations

import struct

import numpy as np


def read_obj_vertices(text: str, scale=(1.0, 1.0, 1.0)):
    """'v x y z' lines, each coordinate rounded to float32, then scaled
    component-wise (mjcf.rs parse_mesh: ``v.mul_elem(scale)``)."""
    out = []
    for line in text.splitlines():
        f = line.split()
        if len(f) >= 4 and f[0] == "v":
            xyz = [struct.unpack("f", struct.pack("f", float(t)))[0] for t in f[1:4]]
            out.append([xyz[0] * scale[0], xyz[1] * scale[1], xyz[2] * scale[2]])
    return np.array(out, dtype=np.float64).reshape(-1, 3)


def read_stl_vertices(data: bytes, scale=(1.0, 1.0, 1.0)):
    """Triangle corners of a binary STL (50 bytes a triangle after an 84-byte
    header) or an ASCII one ('vertex x y z'), as float32 (stl_io's precision),
    widened and scaled per axis."""
    out = []
    if len(data) >= 84:
        n = struct.unpack_from("<I", data, 80)[0]
        if len(data) == 84 + 50 * n:
            for t in range(n):
                for v in range(3):
                    xyz = struct.unpack_from("<3f", data, 84 + 50 * t + 12 + 12 * v)
                    out.append([xyz[0] * scale[0], xyz[1] * scale[1], xyz[2] * scale[2]])
            return np.array(out, dtype=np.float64).reshape(-1, 3)
    toks = data.decode(errors="replace").split()
    for i, tk in enumerate(toks):
        if tk == "vertex":
            xyz = [struct.unpack("f", struct.pack("f", float(t)))[0] for t in toks[i + 1:i + 4]]
            out.append([xyz[0] * scale[0], xyz[1] * scale[1], xyz[2] * scale[2]])
    return np.array(out, dtype=np.float64).reshape(-1, 3)


def _hull2d(pts2, idx):
    """Monotone chain: indices of the strict convex hull of 2-D points,
    counter-clockwise, collinear points dropped."""
    order = sorted(range(len(idx)), key=lambda k: (pts2[k][0], pts2[k][1]))

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])

    lower, upper = [], []
    for k in order:
        while len(lower) >= 2 and cross(pts2[lower[-2]], pts2[lower[-1]], pts2[k]) <= 0:
            lower.pop()
        lower.append(k)
    for k in reversed(order):
        while len(upper) >= 2 and cross(pts2[upper[-2]], pts2[upper[-1]], pts2[k]) <= 0:
            upper.pop()
        upper.append(k)
    return [idx[k] for k in lower[:-1] + upper[:-1]]


def convex_hull(points, rel_tol=1e-10):
    """(vertices m x 3, faces: list of vertex-index lists) or None when the
    points span no volume (the reference then refuses the mesh as degenerate,
    mesh.rs MeshError::Degenerate). Brute force over point triples: for the small
    hulls of tabletop objects (at most 64 points)."""
    P = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    if len(P) > 64:
        raise ValueError("convex_hull restates small hulls only (at most 64 points)")
    uniq = []
    for p in P:
        if not any(np.array_equal(p, q) for q in uniq):
            uniq.append(p)
    P = np.array(uniq)
    n = len(P)
    scale = max(float(np.abs(P).max()) if n else 0.0, 1e-300)
    tol = rel_tol * scale
    planes = {}
    for i in range(n):
        for j in range(i + 1, n):
            for k in range(j + 1, n):
                nrm = np.cross(P[j] - P[i], P[k] - P[i])
                ln = float(np.linalg.norm(nrm))
                if ln <= tol * scale:
                    continue
                nrm = nrm / ln
                d = (P - P[i]) @ nrm
                if np.all(d <= tol):
                    pass
                elif np.all(d >= -tol):
                    nrm, d = -nrm, -d
                else:
                    continue
                on = tuple(int(t) for t in np.nonzero(np.abs(d) <= tol)[0])
                if on not in planes:
                    planes[on] = nrm
    # no volume: every point on one plane (or too few points for a plane)
    if not planes or any(len(on) == n for on in planes):
        return None
    polys = []
    for on, nrm in planes.items():
        a = np.array([1.0, 0.0, 0.0]) if abs(nrm[0]) < 0.57 else np.array([0.0, 1.0, 0.0])
        u = np.cross(nrm, a)
        u = u / np.linalg.norm(u)
        v = np.cross(nrm, u)
        pts2 = [(float(P[t] @ u), float(P[t] @ v)) for t in on]
        poly = _hull2d(pts2, list(on))
        if len(poly) >= 3:
            polys.append(poly)
    used = sorted({t for poly in polys for t in poly}, key=lambda t: tuple(P[t]))
    remap = {t: r for r, t in enumerate(used)}
    faces = []
    for poly in polys:
        f = [remap[t] for t in poly]
        m = f.index(min(f))
        faces.append(f[m:] + f[:m])
    faces.sort()
    return P[used].copy(), faces


def triangles(faces):
    """Fan triangulation (f0, fi, fi+1) of each polygon, winding kept."""
    return [(f[0], f[i], f[i + 1]) for f in faces for i in range(1, len(f) - 1)]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, o):
    return (a[1] * o[2] - a[2] * o[1], a[2] * o[0] - a[0] * o[2], a[0] * o[1] - a[1] * o[0])


def mass_properties(vertices, tris):
    """mesh.rs:72-105: (volume, centre of mass, 3x3 inertia about the centre of
    mass per unit density) of a closed, outward-wound triangle mesh."""
    V = [tuple(float(x) for x in v) for v in np.asarray(vertices, dtype=np.float64).reshape(-1, 3)]
    vol = 0.0
    first = [0.0, 0.0, 0.0]
    c = [[0.0] * 3 for _ in range(3)]
    for ia, ib, idd in tris:
        a, b, d = V[ia], V[ib], V[idd]
        v6 = _dot(a, _cross(b, d))
        vol += v6 / 6.0
        s3 = [(a[k] + b[k]) + d[k] for k in range(3)]
        f = v6 / 24.0
        first = [first[k] + s3[k] * f for k in range(3)]
        for i in range(3):
            for j in range(3):
                s = a[i] * a[j] + b[i] * b[j] + d[i] * d[j]
                p = (a[i] + b[i] + d[i]) * (a[j] + b[j] + d[j])
                c[i][j] += v6 / 120.0 * (s + p)
    if abs(vol) < 1e-18:
        return 0.0, np.zeros(3), np.zeros((3, 3))
    inv = 1.0 / vol
    com = [first[k] * inv for k in range(3)]
    tr = c[0][0] + c[1][1] + c[2][2]
    io = [(tr if i == j else 0.0) - c[i][j] for i in range(3) for j in range(3)]
    cx = [0.0, -com[2], com[1], com[2], 0.0, -com[0], -com[1], com[0], 0.0]
    cxt = [cx[0], cx[3], cx[6], cx[1], cx[4], cx[7], cx[2], cx[5], cx[8]]
    prod = []
    for i in range(3):
        for j in range(3):
            acc = -0.0  # math.rs:169-177: each entry an iterator sum from -0.0
            for k in range(3):
                acc = acc + cx[3 * i + k] * cxt[3 * k + j]
            prod.append(acc)
    icom = [io[t] + prod[t] * -vol for t in range(9)]
    return vol, np.array(com), np.array(icom).reshape(3, 3)
