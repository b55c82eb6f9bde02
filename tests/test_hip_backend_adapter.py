"""include/sing_hip_backend.hpp: HipBackend implements the reference's
RenderBackend (render/include/render_backend.hpp:107-132) over the C ABI.
build() compiles tests/hip_backend_check.cpp against the reference's own
header where the reference tree is present (singularity_amd/hip_backend_check).

CPU: the adapter compiles against the reference interface, reports its name,
and initialize() returns false without a device (CudaBackend's contract,
cuda_backend.cu:111-118).  GPU: two render_frame calls through the interface
equal the library's own FP32 RGBA8 frame at the second frame's virtual clock
(1/30 s, cuda_backend.cu:149-151), bit for bit."""

from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import singularity_amd as sa

ROOT = Path(__file__).resolve().parents[1]
CHECK = ROOT / "singularity_amd" / "hip_backend_check"
needs_check = pytest.mark.skipif(not CHECK.exists(), reason="adapter check not built (no reference tree at build)")


@needs_check
def test_adapter_reports_name_and_fails_cleanly_without_device(tmp_path):
    r = subprocess.run([str(CHECK), str(tmp_path / "f.rgba"), "32", "16", "1"], capture_output=True, text=True,
                       timeout=120)
    assert "name=HIP-gfx950" in r.stdout
    if r.returncode == 3:
        assert "initialize failed" in r.stderr
    else:
        assert r.returncode == 0 and (tmp_path / "f.rgba").stat().st_size == 32 * 16 * 4


@needs_check
def test_usage_error_exit_code(tmp_path):
    r = subprocess.run([str(CHECK)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@needs_check
@pytest.mark.gpu
def test_adapter_frames_equal_library_frames(tmp_path):
    W, H, spp = 96, 72, 2
    out = tmp_path / "f.rgba"
    print(f"hip_backend_check {W}x{H} spp {spp}", flush=True)
    r = subprocess.run([str(CHECK), str(out), str(W), str(H), str(spp)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    t = np.float32(0.0) + np.float32(1.0) / np.float32(30.0)
    assert f"t={float(t + np.float32(1.0) / np.float32(30.0)):.9g}" in r.stdout
    got = np.fromfile(out, np.uint8).reshape(H, W, 4)
    s = sa.scene_defaults()
    s.supersample = spp
    u = sa.pack_uniforms(s, sa.orbital_camera(30.0, 0.15, 0.0, 60.0), W, H, float(t))
    be = sa.HipBackend(W, H)
    be.render(u)
    want = be.capture_frame()
    be.shutdown()
    np.testing.assert_array_equal(got, want)
