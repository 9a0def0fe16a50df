"""Synthetic workloads shared by bench.py and the tests (no reference counterpart)."""
from __future__ import annotations


def random_pile_xml(seed: int, general: bool = False) -> str:
    """2 to 6 free primitives (box, sphere, capsule, cylinder) of random sizes
    and orientations dropped in a column onto a plane, after the reference's
    tabletop tumble scenes. With general=False cylinders collide with the plane
    only; with general=True they also meet every other body, through the
    general convex narrow phase."""
    import numpy as np

    rng = np.random.default_rng(seed)
    nb = int(rng.integers(2, 7))
    bodies = []
    for i in range(nb):
        t = rng.choice(["box", "sphere", "capsule", "cylinder"])
        x, y = rng.uniform(-0.08, 0.08, 2)
        z = 0.05 + 0.07 * i
        e = rng.uniform(0, 180, 3)
        if t == "box":
            g = '<geom type="box" size="%.4f %.4f %.4f"/>' % tuple(rng.uniform(0.015, 0.035, 3))
        elif t == "sphere":
            g = '<geom type="sphere" size="%.4f"/>' % rng.uniform(0.015, 0.035)
        elif t == "capsule":
            g = '<geom type="capsule" size="%.4f %.4f"/>' % (rng.uniform(0.015, 0.03), rng.uniform(0.02, 0.05))
        else:
            mask = "" if general else ' contype="2" conaffinity="0"'
            g = ('<geom type="cylinder" size="%.4f %.4f"%s/>'
                 % (rng.uniform(0.015, 0.03), rng.uniform(0.02, 0.05), mask))
        bodies.append('<body pos="%.4f %.4f %.4f" euler="%.1f %.1f %.1f"><freejoint/>%s</body>'
                      % (x, y, z, e[0], e[1], e[2], g))
    return ('<mujoco><option timestep="0.004"/><worldbody>'
            '<geom type="plane" size="5 5 .1" contype="3" conaffinity="3"/>' + "".join(bodies) +
            "</worldbody></mujoco>")

