"""The C-ABI boundary: libphoton_hip.so loads, exports exactly what include/photon_hip.h declares, and its
error paths behave like createDevice / RenderDevice::render (engine/render_device.cpp:110-157,
test_render_session.cpp:78-94): an unknown or unusable device gives an error with a message, never a crash.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import pytest

from photon_hip import capi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols() -> list[str]:
    text = (ROOT / "include" / "photon_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pe_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """Every ctypes mirror has the size and field offsets gcc gives the header's struct."""
    import subprocess

    structs = {"pe_device_info": capi.DeviceInfo, "pe_device_features": capi.DeviceFeatures,
               "pe_material_desc": capi.MaterialDesc, "pe_mesh_desc": capi.MeshDesc, "pe_light_desc": capi.LightDesc,
               "pe_env_desc": capi.EnvDesc, "pe_scene_desc": capi.SceneDesc, "pe_camera": capi.Camera,
               "pe_settings": capi.Settings, "pe_batch_result": capi.BatchResult,
               "pe_kernel_stats": capi.KernelStats, "pe_server_config": capi.ServerConfig,
               "pe_job_desc": capi.JobDesc, "pe_job_status": capi.JobStatus,
               "pe_guide_config": capi.GuideConfig, "pe_server_stats": capi.ServerStats,
               "pe_targets": capi.Targets, "pe_pass_targets": capi.PassTargets}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "photon_hip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'printf("{cname}.{f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                              check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f in py._fields_:
            assert int(out[f"{cname}.{f[0]}"]) == getattr(py, f[0]).offset, (cname, f[0])


def test_abi_version_and_device_registry_without_gpu():
    lib = capi.load()
    assert lib.pe_abi_version() == 2
    n = C.c_int(-1)
    assert lib.pe_device_count(C.byref(n)) == capi.PE_OK
    assert n.value >= 0
    if n.value == 0:
        h = C.c_void_p()
        rc = lib.pe_device_create(0, C.byref(h))
        assert rc == capi.PE_ERR_NO_DEVICE
        assert b"no such device" in lib.pe_last_error()
        assert not h.value


def test_null_arguments_are_rejected():
    lib = capi.load()
    assert lib.pe_device_count(None) == capi.PE_ERR_INVALID
    assert lib.pe_scene_upload(None, None, None) == capi.PE_ERR_INVALID
    assert lib.pe_film_create(None, 4, 4, None) == capi.PE_ERR_INVALID
    assert lib.pe_scene_release(None) == capi.PE_OK
    assert lib.pe_film_destroy(None) == capi.PE_OK
    assert lib.pe_device_destroy(None) == capi.PE_OK
    res = capi.BatchResult()
    assert lib.pe_render(None, None, None, None, None, 0, 1, None, C.byref(res)) == capi.PE_ERR_INVALID
    assert res.completed == 0 and res.samples_traced == 0


def test_camera_matches_reference_construction():
    from oracle import bindings
    from photon_hip import camera_perspective

    cam = camera_perspective((0, 0, 0.95), (0, -0.2, 0), (0, 1, 0), 60.0, 1.0)
    got = [list(cam.position), list(cam.lower_left), list(cam.horizontal), list(cam.vertical)]
    ref = bindings.reference(strict=True)
    if ref is None:
        pytest.skip("oracle/_ref not built here")
    out = (C.c_float * 12)()
    f3 = lambda v: (C.c_float * 3)(*v)  # noqa: E731
    ref.ref_camera_describe(f3((0, 0, 0.95)), f3((0, -0.2, 0)), f3((0, 1, 0)), 60.0, 1.0, out)
    assert [list(out[3 * k:3 * k + 3]) for k in range(4)] == got
