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



def arm_gripper_xml(seed: int) -> str:
    """A 7-DoF serial arm (hinge joints alternating yaw and pitch, link masses
    U[0.5, 3] kg, position actuators kp = 100) with a two-finger gripper whose
    slide fingers are coupled by a fixed tendon and a joint equality, over a
    cube (half-extent U[0.02, 0.03] m, friction U[0.6, 1.0]) resting on a
    table; implicitfast. After the survey's arm configuration, which has no
    model file in the reference."""
    import numpy as np

    rng = np.random.default_rng(seed)
    mass = rng.uniform(0.5, 3.0, 7)
    half = rng.uniform(0.02, 0.03)
    fric = rng.uniform(0.6, 1.0)
    axes = ["0 0 1", "0 1 0", "0 0 1", "0 1 0", "0 0 1", "0 1 0", "0 0 1"]
    length = [0.10, 0.16, 0.06, 0.14, 0.06, 0.10, 0.04]
    # a folded default pose that puts the hand above the cube
    links, close = "", ""
    for i in range(7):
        d = length[i]
        links += (f'<body name="l{i}" pos="0 0 {0.0 if i == 0 else length[i - 1]:.3f}">'
                  f'<joint name="j{i}" axis="{axes[i]}" range="-2.5 2.5" limited="true"/>'
                  f'<geom type="capsule" fromto="0 0 0 0 0 {d:.3f}" size="0.025" mass="{mass[i]:.4f}" '
                  f'contype="0" conaffinity="0"/>')
        close += "</body>"
    hand = ('<body name="palm" pos="0 0 0.04"><geom type="box" size="0.035 0.045 0.01" mass="0.2"/>'
            '<body name="f1" pos="0 0.03 0.04"><joint name="finger1" type="slide" axis="0 1 0" range="-0.03 0.03" '
            'limited="true"/><geom type="box" size="0.01 0.005 0.03" mass="0.05"/></body>'
            '<body name="f2" pos="0 -0.03 0.04"><joint name="finger2" type="slide" axis="0 -1 0" range="-0.03 0.03" '
            'limited="true"/><geom type="box" size="0.01 0.005 0.03" mass="0.05"/></body></body>')
    acts = "".join(f'<position joint="j{i}" kp="100" kv="5"/>' for i in range(7))
    return ('<mujoco model="arm7_gripper"><option timestep="0.004" integrator="implicitfast"/>'
            '<default><joint damping="1" armature="0.05"/><geom friction="0.8"/></default><worldbody>'
            '<geom type="plane" size="2 2 .1"/>'
            '<geom name="table" type="box" size="0.4 0.4 0.1" pos="0.35 0 0.1"/>'
            f'<body name="base" pos="0 0 0.2" euler="0 110 0">{links}{hand}{close}</body>'
            f'<body name="cube" pos="0.42 0 {0.2 + half:.4f}"><freejoint/>'
            f'<geom type="box" size="{half:.4f} {half:.4f} {half:.4f}" friction="{fric:.3f}" density="800"/></body>'
            '</worldbody>'
            '<tendon><fixed name="grip"><joint joint="finger1" coef="0.5"/><joint joint="finger2" coef="0.5"/>'
            '</fixed></tendon>'
            '<equality><joint joint1="finger2" joint2="finger1" polycoef="0 1 0 0 0"/></equality>'
            f'<actuator>{acts}<position tendon="grip" kp="50" kv="1" ctrlrange="-0.03 0.03" ctrllimited="true"/>'
            '</actuator></mujoco>')


def arm_targets(seed: int, n: int, steps: int, nu: int, every: int = 50):
    """Joint-space targets along random smooth waypoints (cosine blends
    between waypoints drawn every `every` steps), per env; the last actuator
    (the gripper) opens and closes."""
    import numpy as np

    rng = np.random.default_rng(seed)
    nw = steps // every + 2
    way = rng.uniform(-0.6, 0.6, size=(nw, n, nu))
    way[:, :, -1] = rng.uniform(-0.03, 0.03, size=(nw, n))
    t = np.arange(steps)
    k, f = t // every, (t % every) / every
    w = 0.5 - 0.5 * np.cos(np.pi * f)
    return way[k] * (1 - w)[:, None, None] + way[k + 1] * w[:, None, None]
