"""The C-ABI boundary without a GPU: the library loads, exports every entry
point include/*.h declares, and the Python mirror binds the same set; record
layouts match the header; host-only entry points behave."""
import ctypes
import glob
import os
import re

import pytest

import tressfx_amd as tfx

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(tfx_[A-Za-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_the_reference_surface():
    names = declared_functions()
    for n in ("tfx_hair_create", "tfx_hair_update_bone_matrices", "tfx_hair_update_simulation_parameters",
              "tfx_hair_reset_positions", "tfx_simulation_create", "tfx_simulation_simulate", "tfx_sdf_create",
              "tfx_sdf_update", "tfx_sdf_collide_with_hair", "tfx_sdf_collision_system_create"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = tfx.load_library()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_mirror_binds_every_declared_symbol():
    assert declared_functions() == set(tfx.EXPORTS), (declared_functions() ^ set(tfx.EXPORTS))


RECORDS = {
    "tfx_simulation_settings": tfx.SimulationSettings,
    "tfx_simulation_params": tfx.SimulationParams,
    "tfx_asset_desc": tfx.AssetDesc,
    "tfx_hair_info": tfx.HairInfo,
    "tfx_dynamic_state": tfx.DynamicState,
    "tfx_sdf_mesh_desc": tfx.SdfMeshDesc,
    "tfx_sdf_collision_params": tfx.SdfCollisionParams,
    "tfx_sdf_grid_info": tfx.SdfGridInfo,
    "tfx_skinned_mesh_desc": tfx.SkinnedMeshDesc,
    "tfx_server_config": tfx.ServerConfig,
    "tfx_instance_desc": tfx.InstanceDesc,
    "tfx_server_stats": tfx.ServerStats,
    "tfx_rendering_lod_settings": tfx.RenderingLodSettings,
    "tfx_autotune_record": tfx.AutotuneRecord,
    "tfx_hair_to_triangle_mapping": tfx.HairToTriangleMapping,
    "tfx_strand_transform": tfx.StrandTransform,
}


def test_record_layouts_match_header(tmp_path):
    """sizeof/offsetof of every record as the C compiler lays out the header,
    against the ctypes mirror."""
    import subprocess
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tressfx_amd.h"', "int main(void) {"]
    for cname, py in RECORDS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                          check=True).stdout.split("\n") if l)
    for cname, py in RECORDS.items():
        assert int(out[cname]) == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"
    # the reference's own byte sizes: StrandLevelData 48 B, BoneSkinningData 32 B,
    # StandardVertex 32 B, TressFXSDFCollisionParams 64 B
    assert int(out["tfx_sdf_collision_params"]) == 64


def test_settings_defaults_match_reference():
    lib = tfx.load_library()
    s = tfx.SimulationSettings()
    lib.tfx_simulation_settings_default(ctypes.byref(s))
    d = tfx.SimulationSettings.defaults()
    for name, _ in tfx.SimulationSettings._fields_:
        a, b = getattr(s, name), getattr(d, name)
        if isinstance(a, ctypes.Array):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    # TressFXSettings.h:38-56
    assert (s.local_constraints_iterations, s.length_constraints_iterations) == (2, 2)
    assert abs(s.vsp_coeff - 0.8) < 1e-7 and abs(s.damping - 0.08) < 1e-7


def test_host_only_calls_and_errors():
    lib = tfx.load_library()
    ma, mi, pa = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.tfx_get_version(ctypes.byref(ma), ctypes.byref(mi), ctypes.byref(pa)) == 0
    assert (ma.value, mi.value, pa.value) == (4, 0, 0)
    # invalid arguments return a status and a message instead of aborting
    assert lib.tfx_hair_create(None, None) == 1
    assert b"null" in lib.tfx_last_error()
    assert lib.tfx_simulation_set_launch_shape(None, 32, 0, 0) == 1
    h = ctypes.c_void_p()
    assert lib.tfx_simulation_create(ctypes.byref(h)) == 0
    assert lib.tfx_simulation_set_launch_shape(h, 12, 0, 0) == 1
    assert lib.tfx_simulation_set_launch_shape(h, 32, 1, 2) == 0
    assert lib.tfx_simulation_destroy(h) == 0


def test_asset_validation_rejects_bad_shapes():
    import numpy as np
    lib = tfx.load_library()
    a = tfx.HairAsset(1, 0, 12, np.zeros((12, 4)), np.zeros((12, 4)), np.zeros(12), np.zeros((1, 4)),
                      np.zeros((1, 8)))
    d = a.desc()
    h = ctypes.c_void_p()
    assert lib.tfx_hair_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b"vertices per strand" in lib.tfx_last_error()
    skin = np.zeros((1, 8), np.float32)
    skin[0, 0] = 600
    skin[0, 4] = 1
    a = tfx.HairAsset(1, 0, 16, np.zeros((16, 4)), np.zeros((16, 4)), np.zeros(16), np.zeros((1, 4)), skin)
    d = a.desc()
    assert lib.tfx_hair_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b"bone index" in lib.tfx_last_error()


def test_cpp_shim_compiles_against_the_header(tmp_path):
    """include/tressfx_amd.hpp (the reference-named C++ classes) and the shim
    test program compile warning-free and link against the library."""
    import subprocess
    from tressfx_amd import build
    out = str(tmp_path / "shim_test")
    res = subprocess.run(build.cpp_test_command(out), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert res.returncode == 0, res.stdout
    assert os.path.exists(out)
