"""Small scenes shared by the CPU and GPU parity tests."""

# one free body per scene (a single dof block), one scene per primitive pair
# of the device narrow phase (dcollide.hpp vs collision.rs)
PRIMITIVE_SCENES = {
    "plane_box": '<geom type="plane" size="5 5 .1"/><body pos="0 0 .12" euler="10 20 5"><freejoint/><geom type="box" size=".1 .07 .05"/></body>',
    "box_box": '<geom type="box" size="1 1 .1" pos="0 0 -.1"/><body pos=".1 0 .2" euler="8 -12 30"><freejoint/><geom type="box" size=".05 .06 .07"/></body>',
    "capsule_box": '<geom type="box" size="1 1 .1" pos="0 0 -.1"/><body pos="0 0 .2" euler="0 80 0"><freejoint/><geom type="capsule" size=".03 .1"/></body>',
    "sphere_box": '<geom type="box" size="1 1 .1" pos="0 0 -.1"/><body pos="0 0 .2"><freejoint/><geom type="sphere" size=".05"/></body>',
    "plane_cylinder": '<geom type="plane" size="5 5 .1"/><body pos="0 0 .15" euler="30 10 0"><freejoint/><geom type="cylinder" size=".05 .08"/></body>',
    "plane_sphere_capsule": '<geom type="plane" size="5 5 .1"/><body pos="0 0 .2" euler="40 0 0"><freejoint/><geom type="capsule" size=".04 .1"/><geom type="sphere" size=".05" pos=".1 0 0"/></body>',
    "capsule_capsule": '<geom type="capsule" fromto="-1 0 0 1 0 0" size=".05"/><body pos="0 0 .2" euler="0 0 40"><freejoint/><geom type="capsule" fromto="-.2 0 0 .2 0 0" size=".04"/></body>',
    "sphere_sphere": '<geom type="sphere" size=".2"/><body pos=".05 0 .4"><freejoint/><geom type="sphere" size=".1"/></body>',
    "sphere_capsule": '<geom type="capsule" fromto="-1 0 0 1 0 0" size=".05"/><body pos=".1 .02 .2"><freejoint/><geom type="sphere" size=".05"/></body>',
}


# pairs of the general convex narrow phase (a cylinder against a non-plane):
# oracle/general.cpp against pippin_amd/csrc/dgeneral.hpp
GENERAL_SCENES = {
    # crates/pippin/tests/mesh.rs:66-83: upright, on its side, and stacked
    "cylinders_on_box": '<geom type="box" size="1 1 .1" pos="0 0 -.1"/>'
                        '<body name="up" pos="0 0 0.3"><freejoint/><geom type="cylinder" size=".04 .06"/></body>'
                        '<body name="side" pos="0.5 0 0.3" euler="90 0 0"><freejoint/><geom type="cylinder" size=".04 .06"/></body>'
                        '<body name="onup" pos="0 0 0.5"><freejoint/><geom type="cylinder" size=".03 .03"/></body>',
    "sphere_on_cylinder": '<geom type="plane" size="5 5 .1"/><geom type="cylinder" size=".05 .05" pos="0 0 .05"/>'
                          '<body pos=".03 .01 .2"><freejoint/><geom type="sphere" size=".03"/></body>',
    "capsule_on_cylinder": '<geom type="plane" size="5 5 .1"/><geom type="cylinder" size=".04 .1" pos="0 0 .04" euler="90 0 0"/>'
                           '<body pos=".01 0 .2" euler="0 80 20"><freejoint/><geom type="capsule" size=".02 .06"/></body>',
    "cylinder_on_cylinder": '<geom type="plane" size="5 5 .1"/><geom type="cylinder" size=".04 .1" pos="0 0 .04" euler="90 0 0"/>'
                            '<body pos="0 .02 .2" euler="10 70 5"><freejoint/><geom type="cylinder" size=".03 .05"/></body>',
    "box_on_cylinder": '<geom type="plane" size="5 5 .1"/><geom type="cylinder" size=".06 .05" pos="0 0 .05"/>'
                       '<body pos=".01 0 .2" euler="5 8 30"><freejoint/><geom type="box" size=".04 .03 .02"/></body>',
}


def general_xml(name: str) -> str:
    return f'<mujoco><option timestep="0.004"/><worldbody>{GENERAL_SCENES[name]}</worldbody></mujoco>'


def primitive_xml(name: str) -> str:
    return f'<mujoco><option timestep="0.004"/><worldbody>{PRIMITIVE_SCENES[name]}</worldbody></mujoco>'


# several free bodies: one dof block each, islands formed by contacts
MULTIBODY_SCENES = {
    "box_drop": None,  # pippin_amd/assets/box_drop.xml
    "stack3": '<mujoco><option timestep="0.004"/><worldbody><geom type="plane" size="5 5 .1"/>'
              '<body pos="0 0 .06"><freejoint/><geom type="box" size=".05 .05 .05"/></body>'
              '<body pos=".01 0 .17" euler="0 0 20"><freejoint/><geom type="box" size=".05 .05 .05"/></body>'
              '<body pos="0.3 0 .1"><freejoint/><geom type="sphere" size=".05"/></body>'
              '</worldbody></mujoco>',
    "pile": '<mujoco><option timestep="0.004"/><worldbody><geom type="plane" size="5 5 .1"/>'
            '<body pos="0 0 .05"><freejoint/><geom type="capsule" size=".03 .1" euler="0 90 0"/></body>'
            '<body pos="0 0 .12" euler="0 90 30"><freejoint/><geom type="capsule" size=".03 .1"/></body>'
            '<body pos=".05 .05 .2"><freejoint/><geom type="sphere" size=".04"/></body>'
            '<body pos="-.4 0 .2"><freejoint/><geom type="box" size=".04 .03 .05"/></body>'
            '</worldbody></mujoco>',
}

# Arm with a two-finger gripper and a free cube on the floor, after the
# reference's arm + gripper configuration: a fixed tendon drives both fingers
# (position actuator), a joint equality couples the fingers, the integrator is
# implicitfast, and the arm and the cube are separate dof blocks joined into
# one Newton island only while they touch.
GRIPPER = """<mujoco model="gripper">
  <option timestep="0.004" integrator="implicitfast"/>
  <default>
    <joint damping="0.5" armature="0.01"/>
    <geom friction="0.8"/>
  </default>
  <worldbody>
    <geom type="plane" size="2 2 .1"/>
    <body name="base" pos="0 0 0.45">
      <joint name="yaw" axis="0 0 1"/>
      <geom type="capsule" fromto="0 0 0 0 0 0.05" size="0.04" contype="0" conaffinity="0"/>
      <body name="upper" pos="0 0 0.05">
        <joint name="shoulder" axis="0 1 0" range="-1.5 1.5" limited="true"/>
        <geom type="capsule" fromto="0 0 0 0.25 0 0" size="0.03" contype="0" conaffinity="0"/>
        <body name="fore" pos="0.25 0 0">
          <joint name="elbow" axis="0 1 0" range="-2 2" limited="true"/>
          <geom type="capsule" fromto="0 0 0 0 0 -0.25" size="0.025" contype="0" conaffinity="0"/>
          <body name="palm" pos="0 0 -0.27">
            <geom type="box" size="0.04 0.05 0.01"/>
            <body name="f1" pos="0 0.03 -0.04">
              <joint name="finger1" type="slide" axis="0 1 0" range="-0.03 0.03" limited="true"/>
              <geom type="box" size="0.01 0.005 0.03"/>
            </body>
            <body name="f2" pos="0 -0.03 -0.04">
              <joint name="finger2" type="slide" axis="0 -1 0" range="-0.03 0.03" limited="true"/>
              <geom type="box" size="0.01 0.005 0.03"/>
            </body>
          </body>
        </body>
      </body>
    </body>
    <body name="cube" pos="0.25 0 0.03">
      <freejoint/>
      <geom type="box" size="0.025 0.025 0.025"/>
    </body>
  </worldbody>
  <tendon>
    <fixed name="grip">
      <joint joint="finger1" coef="0.5"/>
      <joint joint="finger2" coef="0.5"/>
    </fixed>
  </tendon>
  <equality>
    <joint joint1="finger2" joint2="finger1" polycoef="0 1 0 0 0"/>
  </equality>
  <actuator>
    <position joint="yaw" kp="20" kv="2"/>
    <position joint="shoulder" kp="40" kv="4"/>
    <position joint="elbow" kp="40" kv="4"/>
    <position tendon="grip" kp="50" kv="1" ctrlrange="-0.03 0.03" ctrllimited="true"/>
  </actuator>
</mujoco>
"""


from pippin_amd.workloads import random_pile_xml  # noqa: E402,F401  (shared with bench.py)


# mesh geoms (convex hulls): the reference's mesh cube (crates/pippin/tests/mesh.rs)
# and the 12-vertex wedge of the tumble workload, against a plane, every
# primitive and each other; full documents (they need an <asset> section)
MESH_CUBE = ("-0.05 -0.05 -0.05  0.05 -0.05 -0.05  -0.05 0.05 -0.05  0.05 0.05 -0.05 "
             "-0.05 -0.05 0.05  0.05 -0.05 0.05  -0.05 0.05 0.05  0.05 0.05 0.05")


def _wedge_vertex():
    from pippin_amd.workloads import wedge_vertices
    return " ".join("%.6f %.6f %.6f" % v for v in wedge_vertices())


def mesh_xml(name: str) -> str:
    assets = f'<asset><mesh name="cube" vertex="{MESH_CUBE}"/><mesh name="wedge" vertex="{_wedge_vertex()}"/></asset>'
    return f'<mujoco><option timestep="0.004"/>{assets}<worldbody>{MESH_SCENES[name]}</worldbody></mujoco>'


MESH_SCENES = {
    "cube_mesh_and_box": '<geom type="plane" size="5 5 .1"/>'
                         '<body pos="0 0 0.2" euler="4 7 0"><freejoint/><geom type="mesh" mesh="cube"/></body>'
                         '<body pos="0.5 0 0.2" euler="4 7 0"><freejoint/><geom type="box" size=".05 .05 .05"/></body>',
    "mesh_cubes_stack": '<geom type="plane" size="5 5 .1"/>'
                        '<body pos="0 0 .06"><freejoint/><geom type="mesh" mesh="cube"/></body>'
                        '<body pos=".01 0 .17" euler="0 0 20"><freejoint/><geom type="mesh" mesh="cube"/></body>'
                        '<body pos="-.01 .01 .28" euler="0 0 40"><freejoint/><geom type="mesh" mesh="cube"/></body>',
    "wedge_on_plane": '<geom type="plane" size="5 5 .1"/>'
                      '<body pos="0 0 0.15" euler="20 35 10"><freejoint/><geom type="mesh" mesh="wedge"/></body>',
    "primitives_on_wedge": '<geom type="plane" size="5 5 .1"/>'
                           '<body pos="0 0 0.03" euler="0 0 15"><freejoint/><geom type="mesh" mesh="wedge" density="3000"/></body>'
                           '<body pos="-.01 .005 .1"><freejoint/><geom type="sphere" size=".015"/></body>'
                           '<body pos=".02 -.01 .16" euler="80 10 0"><freejoint/><geom type="capsule" size=".01 .02"/></body>'
                           '<body pos="0 0 .22" euler="5 85 0"><freejoint/><geom type="cylinder" size=".012 .02"/></body>'
                           '<body pos=".005 .01 .28" euler="30 20 10"><freejoint/><geom type="box" size=".015 .012 .01"/></body>',
    "wedge_on_box_and_wedge": '<geom type="box" size="1 1 .1" pos="0 0 -.1"/>'
                              '<body pos="0 0 0.03"><freejoint/><geom type="mesh" mesh="wedge"/></body>'
                              '<body pos=".01 .005 0.12" euler="180 10 30"><freejoint/><geom type="mesh" mesh="wedge"/></body>',
    # the plane listed after the mesh body: the pair is (mesh, plane)
    "mesh_before_plane": '<body pos="0 0 0.2" euler="10 -20 5"><freejoint/><geom type="mesh" mesh="cube"/></body>'
                         '<geom type="plane" size="5 5 .1"/>',
}
