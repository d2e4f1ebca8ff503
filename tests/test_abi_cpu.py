"""The C ABI on the CPU: libwavetool.so loads, exports every entry point include/wt.h
declares, its structs have the header's layout, and the host-only entry points behave
(no GPU compute here)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import wavetool_amd as wt
from wavetool_amd import _abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wt.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = wt.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # and the ctypes prototype table covers the same set
    assert sorted(p[0] for p in _abi.PROTOTYPES) == names


def test_abi_version_and_status_strings():
    lib = wt.lib()
    assert lib.wt_abi_version() == 1
    for s in range(7):
        assert lib.wt_status_string(s)


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "wt.h"
#define S(T) printf(#T " %zu\n", sizeof(T));
#define O(T, f) printf(#T "." #f " %zu\n", offsetof(T, f));
int main(void) {
  S(wt_gerstner_wave) O(wt_gerstner_wave, xzDirection_D) O(wt_gerstner_wave, phaseOffset_phi0)
  S(wt_camera) O(wt_camera, position) O(wt_camera, pitch_deg)
  S(wt_heightmap) O(wt_heightmap, wrap)
  S(wt_water_params) O(wt_water_params, cameraPosition) O(wt_water_params, gridLength)
  O(wt_water_params, heightmapScrollVelocity) O(wt_water_params, waveAnimationTimeInSeconds)
  S(wt_vertex_buffers) S(wt_launch_config)
  S(wt_patch) O(wt_patch, gridLength) O(wt_patch, vertex_offset) O(wt_patch, waveAnimationTimeInSeconds)
  O(wt_patch, heightmapScrollVelocity)
  S(wt_fit_problem) O(wt_fit_problem, spacing) O(wt_fit_problem, target) O(wt_fit_problem, params)
  S(wt_adam_config)
  S(wt_varying_buffers) O(wt_varying_buffers, ndc_xy_height0) O(wt_varying_buffers, heightmap_sample)
  S(wt_shade_params) O(wt_shade_params, sunPosition) O(wt_shade_params, viewportWidthHeight) O(wt_shade_params, zNear)
  S(wt_rgba8_image) O(wt_rgba8_image, height)
  S(wt_shade_textures) O(wt_shade_textures, local_reflections) O(wt_shade_textures, local_refractions)
  O(wt_shade_textures, skybox) O(wt_shade_textures, skybox_length)
  S(wt_water_fragment) O(wt_water_fragment, ndc_xy_height0) O(wt_water_fragment, frag_depth)
  return 0;
}
"""


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_C)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    m = {"wt_gerstner_wave": _abi.GerstnerWave, "wt_camera": _abi.Camera, "wt_heightmap": _abi.Heightmap,
         "wt_water_params": _abi.WaterParams, "wt_vertex_buffers": _abi.VertexBuffers,
         "wt_launch_config": _abi.LaunchConfig, "wt_patch": _abi.Patch, "wt_fit_problem": _abi.FitProblem,
         "wt_adam_config": _abi.AdamConfig, "wt_varying_buffers": _abi.VaryingBuffers,
         "wt_shade_params": _abi.ShadeParams, "wt_rgba8_image": _abi.Rgba8Image,
         "wt_shade_textures": _abi.ShadeTextures, "wt_water_fragment": None}
    # the fragment record: 12 floats, read by the kernel as three float4
    assert got["wt_water_fragment"] == str(_abi.WT_BYTES_PER_FRAGMENT)
    assert got["wt_water_fragment.ndc_xy_height0"] == "36" and got["wt_water_fragment.frag_depth"] == "44"
    for key, val in got.items():
        if key.startswith("wt_water_fragment"):
            continue
        if "." in key:
            t, f = key.split(".")
            assert getattr(m[t], f).offset == int(val), key
        else:
            assert ctypes.sizeof(m[key]) == int(val), key


def test_normalize_steepness_reference_rule():
    """render-engine.cpp:1073: Q_i = (w A != 0) ? q / (w A Count) : 0, in FP32."""
    waves = scenes.gen_waves(3, 6).copy()
    waves[2, 0] = 0.0  # A = 0 -> Q_i = 0
    q = np.linspace(0.1, 0.9, 6).astype(np.float32)
    wt.normalize_steepness(waves, q, 4)
    for i in range(6):
        wa = np.float32(waves[i, 1] * waves[i, 0])
        expect = np.float32(q[i] / np.float32(wa * np.float32(4))) if wa != 0 else np.float32(0)
        assert waves[i, 3] == expect


def test_host_entry_points_reject_bad_arguments():
    with pytest.raises(wt.WtError) as e:
        wt.projector_corners(scenes.camera_tuple(fov_deg=0.0), 2.0)
    assert e.value.status == wt.WT_ERR_INVALID_ARG
    with pytest.raises(wt.WtError) as e:
        wt.projector_corners(scenes.camera_tuple(height=100.0, pitch_deg=80.0, far=50.0), 2.0)
    assert e.value.status == wt.WT_ERR_NO_INTERSECTION
    lib = wt.lib()
    vb = _abi.VertexBuffers()
    assert lib.wt_water_displace(None, None, None, 0, None, ctypes.byref(vb), None) == wt.WT_ERR_INVALID_ARG
    assert lib.wt_fit_create(None, None, 0, None) == wt.WT_ERR_INVALID_ARG
    assert lib.wt_batch_create(None, None, 0, None) == wt.WT_ERR_INVALID_ARG
    # the device-side entry points reject a missing context before touching the GPU
    assert lib.wt_projector_corners_batched(None, None, None, 1, None, None, None) == wt.WT_ERR_INVALID_ARG
    assert lib.wt_batch_set_cameras(None, None, None, None, None, None) == wt.WT_ERR_INVALID_ARG
    assert lib.wt_grid_triangle_indices(None, 16, 0, None, None) == wt.WT_ERR_INVALID_ARG
    assert lib.wt_autotune_batch(None, 64, 8, 1, None, None) == wt.WT_ERR_INVALID_ARG
    assert lib.wt_vertex_varyings(None, None, None, None, None, None, None, 0, None, None) == wt.WT_ERR_INVALID_ARG
    assert lib.wt_water_shade(None, None, None, None, None, 4, 4, None, 0, None) == wt.WT_ERR_INVALID_ARG


def test_phase_plan():
    fp = _abi.FitProblem()
    fp.grid_side, fp.snapshots, fp.wave_count = 64, 4, 8
    b, rec = wt.phase_plan(fp)
    assert b == 64 * 64 * 4 * 8 * 8 and rec


def test_context_creation_without_a_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    h = ctypes.c_void_p()
    assert wt.lib().wt_create(ctypes.byref(h)) == wt._abi.WT_ERR_HIP
