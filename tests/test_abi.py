"""The C-ABI library on the CPU: it loads, exports every entry point the header
declares, its MJCF loader agrees with the oracle's, and it fails loudly (no CPU
fallback) where a GPU would be needed."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import mjcf
from pippin_amd import physics

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "pippin_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(pippin_hip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_trait_surface():
    names = declared_functions()
    # Physics trait (crates/pippin-env/src/lib.rs:110-155) plus lifecycle and diagnostics
    for need in ["create", "destroy", "dims", "num_envs", "step", "reset", "get", "set", "poses",
                 "set_param", "get_param", "appearance", "ncon", "contacts", "solver_iterations",
                 "mass_bias", "collision_pairs", "model_load_file", "model_load_xml", "last_error"]:
        assert f"pippin_hip_{need}" in names, need


def test_library_exports_every_declared_symbol():
    L = C.CDLL(str(physics.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python front end binds exactly what the header declares
    physics.lib()
    assert physics.BOUND == declared_functions()


@pytest.mark.parametrize("name", ["ant.xml", "cartpole.xml", "pendulum.xml", "box_drop.xml", "tumble.xml",
                                  "arm3d.xml"])
def test_library_loader_matches_oracle_loader(assets, name):
    a = physics.load_model(assets / name)
    b = mjcf.load_file(assets / name)
    assert set(a) == set(b)
    for k in b:
        va, vb = np.asarray(a[k]), np.asarray(b[k])
        assert np.array_equal(va.reshape(-1), vb.reshape(-1).astype(va.dtype)), k


def _generated_models():
    from pippin_amd.workloads import arm_gripper_xml, random_pile_xml
    from tests.scenes import GENERAL_SCENES, MULTIBODY_SCENES, PRIMITIVE_SCENES, general_xml, primitive_xml

    out = [(f"primitive {k}", primitive_xml(k)) for k in sorted(PRIMITIVE_SCENES)]
    out += [(f"multibody {k}", x) for k, x in sorted(MULTIBODY_SCENES.items()) if x is not None]
    out += [(f"general {k}", general_xml(k)) for k in sorted(GENERAL_SCENES)]
    out += [(f"pile {s}", random_pile_xml(s)) for s in (11, 29)]
    out += [(f"general pile {s}", random_pile_xml(s, general=True)) for s in (13, 18)]
    out += [("arm gripper 5", arm_gripper_xml(5)), ("arm gripper 0", arm_gripper_xml(0))]
    from pippin_amd.workloads import tabletop_tumble
    from tests.scenes import GRIPPER, MESH_SCENES, mesh_xml

    out.append(("gripper", GRIPPER))
    out += [(f"mesh {k}", mesh_xml(k)) for k in sorted(MESH_SCENES)]
    out.append(("tabletop tumble", tabletop_tumble(2, 0).xml))
    return out


@pytest.mark.parametrize("name,xml", _generated_models(), ids=[n for n, _ in _generated_models()])
def test_library_loader_matches_oracle_loader_on_generated_models(name, xml):
    """Every generated model the GPU tests step (primitive pairs, multi-body
    scenes, cylinder scenes, random piles, the arm with its gripper) compiles
    to the same arrays through the library's loader as through the oracle's."""
    a = physics.load_model(xml, is_xml=True)
    b = mjcf.load_string(xml)
    assert set(a) == set(b)
    for k in b:
        va, vb = np.asarray(a[k]), np.asarray(b[k])
        assert np.array_equal(va.reshape(-1), vb.reshape(-1).astype(va.dtype)), (name, k)


def test_loader_reports_errors():
    with pytest.raises(physics.HipError, match="."):
        physics.load_model("<mujoco><worldbody><body><joint type='nope'/></body></worldbody></mujoco>",
                           is_xml=True)
    with pytest.raises(physics.HipError):
        physics.load_model("/nonexistent/model.xml")


def test_create_without_gpu_fails_loudly(assets):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    arr = physics.load_model(assets / "ant.xml")
    with pytest.raises(physics.HipError):
        physics.HipPhysics(arr, 4)


def test_library_loader_reads_obj_and_stl_meshes(tmp_path):
    """Mesh files: an OBJ (the reference's wedge test, tests/mesh.rs:85-116) and
    the same points as binary and ASCII STL give the same hull through both
    loaders."""
    import struct

    pts = [(0, 0, 0), (0.1, 0, 0), (0, 0.1, 0), (0, 0, 0.1), (0.1, 0.1, 0)]
    tris = [(0, 2, 1), (1, 2, 4), (0, 1, 3), (0, 3, 2), (1, 4, 3), (2, 3, 4)]
    (tmp_path / "wedge.obj").write_text("".join("v %g %g %g\n" % p for p in pts)
                                        + "".join("f %d %d %d\n" % tuple(i + 1 for i in t) for t in tris))
    binary = b"\0" * 80 + struct.pack("<I", len(tris))
    for t in tris:
        binary += struct.pack("<3f", 0, 0, 0) + b"".join(struct.pack("<3f", *pts[i]) for i in t) + b"\0\0"
    (tmp_path / "wedge_b.stl").write_bytes(binary)
    (tmp_path / "wedge_a.stl").write_text(
        "solid w\n" + "".join("facet normal 0 0 0\nouter loop\n" + "".join("vertex %g %g %g\n" % pts[i] for i in t)
                               + "endloop\nendfacet\n" for t in tris) + "endsolid w\n")
    hulls = []
    for f in ("wedge.obj", "wedge_b.stl", "wedge_a.stl"):
        xml = (f'<mujoco><asset><mesh name="w" file="{f}" scale="2 2 2"/></asset><worldbody>'
               '<body><freejoint/><geom type="mesh" mesh="w"/></body></worldbody></mujoco>')
        (tmp_path / "m.xml").write_text(xml)
        a = physics.load_model(tmp_path / "m.xml")
        b = mjcf.load_file(tmp_path / "m.xml")
        for k in b:
            assert np.array_equal(np.asarray(a[k]).reshape(-1), np.asarray(b[k]).reshape(-1).astype(np.asarray(a[k]).dtype)), (f, k)
        hulls.append((np.asarray(a["mesh_vert"]).copy(), np.asarray(a["face_vert"]).copy(), np.asarray(a["body_mass"])))
        assert abs(np.asarray(a["mesh_vert"]).reshape(-1, 3)[:, 0].max() - 0.2) < 1e-6
    for h in hulls[1:]:
        assert all(np.array_equal(x, y) for x, y in zip(h, hulls[0]))


def test_degenerate_mesh_is_refused():
    flat = "0 0 0  1 0 0  0 1 0  1 1 0"
    xml = (f'<mujoco><asset><mesh name="f" vertex="{flat}"/></asset><worldbody>'
           '<body><freejoint/><geom type="mesh" mesh="f"/></body></worldbody></mujoco>')
    with pytest.raises(physics.HipError, match="degenerate"):
        physics.load_model(xml, is_xml=True)
    with pytest.raises(mjcf.MjcfError, match="degenerate"):
        mjcf.load_string(xml)


def test_includes_splice_in_both_loaders(tmp_path):
    """<include file=.../> (mjcf.rs:178-209): replaced by the children of the
    included root, relative to the including file, recursively; the result is
    the same model as the flat document, through both loaders; a cycle is an
    error."""
    (tmp_path / "parts").mkdir()
    (tmp_path / "parts" / "legs.xml").write_text(
        '<mujoco><body name="leg" pos="0.2 0 0"><joint axis="0 1 0"/>'
        '<geom type="capsule" fromto="0 0 0 0.2 0 0" size="0.03"/></body></mujoco>')
    (tmp_path / "parts" / "torso.xml").write_text(
        '<mujoco><body name="torso" pos="0 0 0.5"><freejoint/><geom type="sphere" size="0.1"/>'
        '<include file="legs.xml"/></body></mujoco>')
    (tmp_path / "opts.xml").write_text('<mujoco><option timestep="0.005"/></mujoco>')
    main = ('<mujoco><include file="opts.xml"/><worldbody><geom type="plane" size="5 5 .1"/>'
            '<include file="parts/torso.xml"/></worldbody></mujoco>')
    (tmp_path / "main.xml").write_text(main)
    flat = ('<mujoco><option timestep="0.005"/><worldbody><geom type="plane" size="5 5 .1"/>'
            '<body name="torso" pos="0 0 0.5"><freejoint/><geom type="sphere" size="0.1"/>'
            '<body name="leg" pos="0.2 0 0"><joint axis="0 1 0"/>'
            '<geom type="capsule" fromto="0 0 0 0.2 0 0" size="0.03"/></body></body></worldbody></mujoco>')
    a = physics.load_model(tmp_path / "main.xml")
    b = mjcf.load_file(tmp_path / "main.xml")
    c = mjcf.load_string(flat)
    assert a["timestep"] == 0.005 and int(a["nbody"]) == 3
    for k in c:
        vc = np.asarray(c[k]).reshape(-1)
        for other in (a, b):
            assert np.array_equal(np.asarray(other[k]).reshape(-1).astype(vc.dtype), vc), k
    (tmp_path / "loop.xml").write_text('<mujoco><include file="loop.xml"/></mujoco>')
    with pytest.raises(physics.HipError, match="nested too deeply"):
        physics.load_model(tmp_path / "loop.xml")
    with pytest.raises(mjcf.MjcfError, match="nested too deeply"):
        mjcf.load_file(tmp_path / "loop.xml")


@pytest.mark.parametrize("levels,ok", [(16, True), (17, False)])
def test_include_depth_limit(tmp_path, levels, ok):
    """mjcf.rs:179-183: included text is expanded up to 16 levels below the
    top document; the 17th level is an error in both loaders."""
    for k in range(1, levels + 1):
        inner = f'<include file="l{k + 1}.xml"/>' if k < levels else '<geom type="plane" size="1 1 .1"/>'
        (tmp_path / f"l{k}.xml").write_text(f"<mujoco>{inner}</mujoco>")
    main = '<mujoco><worldbody><include file="l1.xml"/></worldbody></mujoco>'
    (tmp_path / "main.xml").write_text(main)
    if ok:
        assert int(physics.load_model(tmp_path / "main.xml")["ngeom"]) == 1
        assert int(mjcf.load_file(tmp_path / "main.xml")["ngeom"]) == 1
    else:
        with pytest.raises(physics.HipError, match="nested too deeply"):
            physics.load_model(tmp_path / "main.xml")
        with pytest.raises(mjcf.MjcfError, match="nested too deeply"):
            mjcf.load_file(tmp_path / "main.xml")
