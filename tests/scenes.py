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
